// server_core.h -- request scheduler of the served form, independent of the
// device.  Pure C++17 (threads, mutexes): the HIP backend (server.hip)
// implements the Backend interface; a host-only test double drives the same
// scheduler in tests.
//
// Continuous batching: a request (one scene, several target models) is
// admitted when its device-memory estimate fits the free part of the budget.
// Its scene is uploaded and each target becomes a job.  The scheduler thread
// runs vote rounds: one launch over slices of reference rows of every voting
// job, sized so that a round takes about `round_ms`, each job getting an equal
// share of the round's work.  Requests that arrive meanwhile join at the next
// round.  A job whose rows are all voted moves to the lanes: host threads
// with their own streams that run threshold/order, poses and clustering while
// the next rounds vote; a lane takes every queued job it can finish together
// with the first (one launch per stage for all of them).  When all of a request's jobs are finished its device
// memory returns to the budget.
//
// The reference has no served form: it registers one (scene, model) pair at
// a time (src/cuda/ppf.cu:63-98).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ppf_amd.h"

namespace ppfs {

using Clock = std::chrono::steady_clock;

enum class State { Queued, Voting, Finishing, Done, Failed, Cancelled };

struct Job {
  int32_t target = 0;    // position in the request's target list
  int32_t model_id = 0;
  int64_t rows = 0;      // reference rows of the scene
  int64_t next_row = 0;  // first row not yet voted in the current pass
  double row_units = 1;  // work estimate of one row (n_s * n_m)
  int passes = 0;        // vote passes started (more than one after an overflow)
  int status = PPF_OK;
  std::string error;
  std::shared_ptr<void> dev;  // backend state of the job
};

struct Request {
  uint64_t ticket = 0;
  int64_t n_points = 0;
  uint32_t rpdf = 1;
  int64_t rows = 0;
  std::vector<float> points, normals;  // host copies, dropped after upload
  std::vector<int32_t> model_ids;
  ppf_lookup_params params{};
  std::vector<Job> jobs;
  State state = State::Queued;
  int status = PPF_OK;
  std::string error;
  uint64_t bytes = 0;  // device-memory estimate used for admission
  int jobs_left = 0;
  int rounds = 0;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<std::array<float, 16>> poses;
  std::shared_ptr<void> dev;                // backend state of the request (scene)
  std::vector<std::shared_ptr<void>> results;  // backend results per target
};

struct Slice {
  Request *req;
  Job *job;
  int64_t row0, rows;
};

class Backend {
 public:
  virtual ~Backend() = default;
  // Device bytes the request holds while in flight.
  virtual uint64_t request_bytes(const Request &r) = 0;
  // Scheduler thread.  Upload the scene; start a vote pass of job j.
  virtual int upload(Request &r, std::string &err) = 0;
  virtual int start_pass(Request &r, Job &j, std::string &err) = 0;
  // Scheduler thread: one vote launch over the slices; returns when done.
  virtual int vote_round(std::vector<Slice> &slices, double *ms, std::string &err) = 0;
  // Scheduler thread, after the last row of a pass: *rerun asks for another
  // pass (candidate overflow); start_pass is called again for it.
  virtual int pass_done(Request &r, Job &j, bool *rerun, std::string &err) = 0;
  // Lane thread `lane`: stages after voting; fills r.poses[j.target] and
  // r.results[j.target], and the job's outcome (PPF_OK or PPF_ERR_NO_VOTES)
  // in *job_status.
  virtual int finish(int lane, Request &r, Job &j, int *job_status, std::string &err) = 0;
  // Jobs whose keys are equal (and >= 0) may be finished together by one
  // finish_group call; -1: this job is finished alone.  Called under the
  // scheduler's lock: it reads only what is fixed at submission.
  virtual int64_t group_key(const Request &, const Job &) { return -1; }
  // Lane thread `lane`: finish group[k] = (request, job) for every k, as
  // finish() would one by one: rc[k], status[k] and err[k] per job.
  virtual void finish_group(int lane, std::vector<std::pair<Request *, Job *>> &group, std::vector<int> &rc,
                            std::vector<int> &status, std::vector<std::string> &err) {
    for (size_t k = 0; k < group.size(); ++k)
      rc[k] = finish(lane, *group[k].first, *group[k].second, &status[k], err[k]);
  }
  // Any thread: free the request's device state (results stay).
  virtual void release(Request &r) = 0;
  // Reference rows the device votes at once (blocks of the vote launch).
  virtual int parallel_slots() { return 1; }
};

struct Config {
  uint64_t budget = 0;          // device bytes for requests in flight
  int max_active = 64;          // requests in flight at once
  int lanes = 4;                // finishing lanes
  int max_group = 64;           // jobs a lane finishes together at most
  double round_ms = 50.0;       // target duration of one vote round
  double tail_rows = 8.0;       // a round lasts at least this many rows per slot
  double aging = 1.0;           // ms of priority a request gains per ms in flight
  int64_t min_round_rows = 1;   // at least this many rows per round (if available)
  int64_t max_round_rows = 1 << 20;
};

struct Stats {
  int64_t submitted = 0, completed = 0, failed = 0, cancelled = 0, rejected = 0;
  int64_t rounds = 0, rows_voted = 0, queued = 0, active = 0;
  uint64_t inflight_bytes = 0, peak_inflight_bytes = 0;
  int64_t max_requests_in_round = 0;  // most distinct requests voted in one round
  double vote_ms = 0;
  int64_t finish_groups = 0, max_finish_group = 0;  // lane finishing calls, most jobs in one
};

class Scheduler {
 public:
  Scheduler(Backend *backend, const Config &cfg) : be_(backend), cfg_(cfg) {
    if (cfg_.lanes < 1) cfg_.lanes = 1;
    if (cfg_.max_active < 1) cfg_.max_active = 1;
    sched_ = std::thread([this] { loop(); });
    for (int l = 0; l < cfg_.lanes; ++l) lanes_.emplace_back([this, l] { lane_loop(l); });
  }

  ~Scheduler() { shutdown(); }

  // Stop: cancel queued requests, let requests in flight complete, join.
  void shutdown() {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (stopping_) return;
      stopping_ = true;
      for (Request *r : queue_) finalize_locked(*r, State::Cancelled, PPF_ERR_CANCELLED, "server stopped");
      queue_.clear();
    }
    sched_cv_.notify_all();
    if (sched_.joinable()) sched_.join();
    {
      std::lock_guard<std::mutex> g(mu_);
      lanes_stop_ = true;
    }
    lane_cv_.notify_all();
    for (auto &t : lanes_)
      if (t.joinable()) t.join();
  }

  // Takes ownership; returns the ticket (> 0).
  uint64_t submit(std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> g(mu_);
    r->ticket = ++next_ticket_;
    r->t_submit = Clock::now();
    r->state = State::Queued;
    r->jobs_left = (int)r->model_ids.size();
    r->poses.assign(r->model_ids.size(), std::array<float, 16>{});
    r->results.assign(r->model_ids.size(), nullptr);
    r->bytes = be_->request_bytes(*r);
    Request *p = r.get();
    reqs_[p->ticket] = std::move(r);
    ++st_.submitted;
    if (stopping_) {
      finalize_locked(*p, State::Cancelled, PPF_ERR_CANCELLED, "server stopped");
    } else {
      queue_.push_back(p);
      sched_cv_.notify_all();
    }
    return p->ticket;
  }

  // Wait for the request; on return (any status but PPF_ERR_TIMEOUT) the
  // record is removed and handed to *out.
  int wait(uint64_t ticket, double timeout_s, std::unique_ptr<Request> *out) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return PPF_ERR_ARG;
    Request *r = it->second.get();
    auto done = [&] { return is_final(r->state); };
    if (timeout_s < 0) {
      done_cv_.wait(lk, done);
    } else {
      // a system_clock deadline waits in pthread_cond_timedwait, which thread
      // checkers understand (steady_clock waits go through clockwait)
      const auto until = std::chrono::system_clock::now() +
                         std::chrono::duration_cast<std::chrono::system_clock::duration>(
                             std::chrono::duration<double>(timeout_s));
      if (!done_cv_.wait_until(lk, until, done)) return PPF_ERR_TIMEOUT;
    }
    *out = std::move(it->second);
    reqs_.erase(it);
    return PPF_OK;
  }

  // 1 if the request has completed (any final state), 0 if not.
  int poll(uint64_t ticket, int *done) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return PPF_ERR_ARG;
    *done = is_final(it->second->state) ? 1 : 0;
    return PPF_OK;
  }

  // Cancel a request that has not been admitted yet.
  int cancel(uint64_t ticket) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return PPF_ERR_ARG;
    Request *r = it->second.get();
    if (r->state != State::Queued) return PPF_ERR_STATE;
    for (auto q = queue_.begin(); q != queue_.end(); ++q)
      if (*q == r) {
        queue_.erase(q);
        break;
      }
    finalize_locked(*r, State::Cancelled, PPF_ERR_CANCELLED, "cancelled");
    return PPF_OK;
  }

  Stats stats() {
    std::lock_guard<std::mutex> g(mu_);
    Stats s = st_;
    s.queued = (int64_t)queue_.size();
    s.active = (int64_t)active_.size();
    s.inflight_bytes = inflight_;
    return s;
  }

 private:
  static bool is_final(State s) {
    return s == State::Done || s == State::Failed || s == State::Cancelled;
  }

  void finalize_locked(Request &r, State s, int status, const std::string &err) {
    r.state = s;
    if (status != PPF_OK && r.status == PPF_OK) {
      r.status = status;
      r.error = err;
    }
    r.t_done = Clock::now();
    if (s == State::Done) ++st_.completed;
    if (s == State::Failed) ++st_.failed;
    if (s == State::Cancelled) ++st_.cancelled;
    r.points.clear();
    r.points.shrink_to_fit();
    r.normals.clear();
    r.normals.shrink_to_fit();
    done_cv_.notify_all();
  }

  // Return the request's memory and drop it from the active set.
  void retire_locked(Request &r) {
    inflight_ -= r.bytes;
    for (auto a = active_.begin(); a != active_.end(); ++a)
      if (*a == &r) {
        active_.erase(a);
        break;
      }
    sched_cv_.notify_all();
  }

  // A request failed before all of its jobs reached the lanes.
  void fail_inflight_locked(std::unique_lock<std::mutex> &lk, Request &r, int status,
                            const std::string &err) {
    for (Job &j : r.jobs)
      if (j.status == PPF_OK) {
        j.status = status;
        j.error = err;
      }
    r.status = status;
    r.error = err;
    // jobs already with the lanes finish there; the rest are dropped now
    int queued = 0;
    for (auto &f : finq_)
      if (f.first == &r) ++queued;
    const int with_lanes = queued + busy_jobs_of(r);
    r.jobs_left = with_lanes;
    if (with_lanes == 0) {
      lk.unlock();
      be_->release(r);
      lk.lock();
      retire_locked(r);
      finalize_locked(r, State::Failed, status, err);
    }
  }

  int busy_jobs_of(const Request &r) {
    int n = 0;
    for (const Request *b : busy_) n += (b == &r);
    return n;
  }

  void loop() {
    std::unique_lock<std::mutex> lk(mu_);
    double unit_ms = 0.0;  // measured device ms per work unit (row * n_s * n_m)
    for (;;) {
      // ---- admission (FIFO; a request never fitting the budget is rejected)
      while (!queue_.empty()) {
        Request *r = queue_.front();
        if (cfg_.budget && r->bytes > cfg_.budget) {
          queue_.pop_front();
          ++st_.rejected;
          finalize_locked(*r, State::Failed, PPF_ERR_NOMEM,
                          "request needs " + std::to_string(r->bytes) + " device bytes, budget is " +
                              std::to_string(cfg_.budget));
          continue;
        }
        if ((cfg_.budget && inflight_ + r->bytes > cfg_.budget) ||
            (int)active_.size() >= cfg_.max_active)
          break;
        queue_.pop_front();
        inflight_ += r->bytes;
        st_.peak_inflight_bytes = std::max(st_.peak_inflight_bytes, inflight_);
        active_.push_back(r);
        r->state = State::Voting;
        r->t_admit = Clock::now();
        lk.unlock();
        std::string err;
        int rc = be_->upload(*r, err);
        for (size_t k = 0; rc == PPF_OK && k < r->jobs.size(); ++k) {
          rc = be_->start_pass(*r, r->jobs[k], err);
          r->jobs[k].passes = 1;
        }
        lk.lock();
        if (rc != PPF_OK) fail_inflight_locked(lk, *r, rc, err);
      }
      // ---- the voting jobs
      std::vector<std::pair<Request *, Job *>> voting;
      for (Request *r : active_)
        if (r->state == State::Voting && r->status == PPF_OK)
          for (Job &j : r->jobs)
            if (j.status == PPF_OK && j.next_row < j.rows) voting.emplace_back(r, &j);
      if (voting.empty()) {
        if (stopping_ && queue_.empty() && active_.empty()) break;
        sched_cv_.wait(lk);
        continue;
      }
      // ---- plan the round.  Jobs in order of estimated remaining device
      // time minus their age (shortest remaining first, so short requests
      // finish early; age keeps long ones from starving).  The round's work
      // budget is round_ms, but at least tail_rows rows per slot of the
      // costliest row: a round ends with its slowest block, so rounds much
      // longer than one row keep the device busy.
      const auto now = Clock::now();
      const double slots = (double)std::max(1, be_->parallel_slots());
      double max_row_units = 0.0;
      for (auto &v : voting) max_row_units = std::max(max_row_units, v.second->row_units);
      std::vector<std::pair<double, size_t>> order;
      for (size_t k = 0; k < voting.size(); ++k) {
        const Job &j = *voting[k].second;
        const double rem = (double)(j.rows - j.next_row) * j.row_units;
        const double age = std::chrono::duration<double, std::milli>(now - voting[k].first->t_admit).count();
        order.emplace_back(unit_ms > 0.0 ? rem * unit_ms - cfg_.aging * age : rem, k);
      }
      std::stable_sort(order.begin(), order.end(),
                       [](const std::pair<double, size_t> &x, const std::pair<double, size_t> &y) {
                         return x.first < y.first;
                       });
      std::vector<Slice> slices;
      int64_t total_rows = 0;
      if (unit_ms > 0.0) {
        double budget = std::max(cfg_.round_ms / unit_ms, cfg_.tail_rows * slots * max_row_units);
        for (auto &o : order) {
          if (budget <= 0.0 && total_rows >= cfg_.min_round_rows) break;
          Request *r = voting[o.second].first;
          Job &j = *voting[o.second].second;
          const int64_t left = j.rows - j.next_row;
          int64_t rows = (int64_t)std::ceil(std::max(0.0, budget) / std::max(1.0, j.row_units));
          rows = std::min(left, std::max<int64_t>(rows, 1));
          slices.push_back({r, &j, j.next_row, rows});
          total_rows += rows;
          budget -= (double)rows * j.row_units;
        }
      } else {
        // first round: no cost estimate yet
        const int64_t each = std::max<int64_t>(1, cfg_.min_round_rows / (int64_t)voting.size());
        for (auto &o : order) {
          Job &j = *voting[o.second].second;
          const int64_t rows = std::min(j.rows - j.next_row, each);
          slices.push_back({voting[o.second].first, &j, j.next_row, rows});
          total_rows += rows;
        }
      }
      // costliest rows first in the launch, so the round's tail holds cheap rows
      std::stable_sort(slices.begin(), slices.end(), [](const Slice &x, const Slice &y) {
        return x.job->row_units > y.job->row_units;
      });
      double units = 0.0;
      int64_t distinct = 0;
      {
        std::vector<const Request *> seen;
        for (Slice &s : slices) {
          units += (double)s.rows * s.job->row_units;
          if (std::find(seen.begin(), seen.end(), s.req) == seen.end()) seen.push_back(s.req);
        }
        distinct = (int64_t)seen.size();
      }
      lk.unlock();
      double ms = 0.0;
      std::string err;
      int rc = be_->vote_round(slices, &ms, err);
      lk.lock();
      ++st_.rounds;
      st_.vote_ms += ms;
      st_.rows_voted += total_rows;
      st_.max_requests_in_round = std::max(st_.max_requests_in_round, distinct);
      if (rc != PPF_OK) {
        std::vector<Request *> hit;
        for (Slice &s : slices)
          if (std::find(hit.begin(), hit.end(), s.req) == hit.end()) hit.push_back(s.req);
        for (Request *r : hit)
          if (r->status == PPF_OK) fail_inflight_locked(lk, *r, rc, err);
        continue;
      }
      if (units > 0.0 && ms > 0.0) {
        const double u = ms / units;
        unit_ms = unit_ms > 0.0 ? 0.5 * unit_ms + 0.5 * u : u;
      }
      {
        std::vector<Request *> seen;
        for (Slice &s : slices) {
          s.job->next_row += s.rows;
          if (std::find(seen.begin(), seen.end(), s.req) == seen.end()) {
            seen.push_back(s.req);
            ++s.req->rounds;
          }
        }
      }
      // ---- jobs whose pass ended
      for (Slice &s : slices) {
        Request &r = *s.req;
        Job &j = *s.job;
        if (j.next_row < j.rows || r.status != PPF_OK) continue;
        bool rerun = false;
        lk.unlock();
        std::string e2;
        int rc2 = be_->pass_done(r, j, &rerun, e2);
        if (rc2 == PPF_OK && rerun) {
          j.next_row = 0;
          ++j.passes;
          rc2 = be_->start_pass(r, j, e2);
        }
        lk.lock();
        if (rc2 != PPF_OK) {
          fail_inflight_locked(lk, r, rc2, e2);
          continue;
        }
        if (!rerun) {
          finq_.emplace_back(&r, &j);
          lane_cv_.notify_one();
        }
      }
      // a request whose jobs have all left voting is finishing
      for (Request *r : active_)
        if (r->state == State::Voting && r->status == PPF_OK) {
          bool all = true;
          for (Job &j : r->jobs) all = all && j.next_row >= j.rows;
          if (all) r->state = State::Finishing;
        }
    }
  }

  void lane_loop(int lane) {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      lane_cv_.wait(lk, [&] { return !finq_.empty() || lanes_stop_; });
      if (finq_.empty()) {
        if (lanes_stop_) break;
        continue;
      }
      // the front job, and with it every queued job the backend can finish
      // together with it (same group key, not failed), in queue order
      std::vector<std::pair<Request *, Job *>> group{finq_.front()};
      finq_.pop_front();
      const int64_t key =
          group[0].second->status == PPF_OK ? be_->group_key(*group[0].first, *group[0].second) : -1;
      if (key >= 0) {
        for (auto q = finq_.begin(); q != finq_.end() && (int)group.size() < cfg_.max_group;) {
          if (q->second->status == PPF_OK && be_->group_key(*q->first, *q->second) == key) {
            group.push_back(*q);
            q = finq_.erase(q);
          } else {
            ++q;
          }
        }
      }
      for (auto &g : group) busy_.push_back(g.first);
      const size_t n = group.size();
      ++st_.finish_groups;
      st_.max_finish_group = std::max<int64_t>(st_.max_finish_group, (int64_t)n);
      std::vector<int> rc(n, PPF_OK), js(n, PPF_OK);
      std::vector<std::string> err(n);
      std::vector<std::pair<Request *, Job *>> todo;
      std::vector<size_t> where;
      for (size_t k = 0; k < n; ++k)
        if (group[k].second->status == PPF_OK) {
          todo.push_back(group[k]);
          where.push_back(k);
        }
      lk.unlock();
      if (todo.size() == 1) {
        rc[where[0]] = be_->finish(lane, *todo[0].first, *todo[0].second, &js[where[0]], err[where[0]]);
      } else if (!todo.empty()) {
        std::vector<int> rc2(todo.size(), PPF_OK), js2(todo.size(), PPF_OK);
        std::vector<std::string> e2(todo.size());
        be_->finish_group(lane, todo, rc2, js2, e2);
        for (size_t k = 0; k < todo.size(); ++k) {
          rc[where[k]] = rc2[k];
          js[where[k]] = js2[k];
          err[where[k]] = e2[k];
        }
      }
      lk.lock();
      for (size_t k = 0; k < n; ++k) {
        Request &r = *group[k].first;
        Job &j = *group[k].second;
        if (rc[k] == PPF_OK && j.status == PPF_OK) j.status = js[k];
        for (auto b = busy_.begin(); b != busy_.end(); ++b)
          if (*b == &r) {
            busy_.erase(b);
            break;
          }
        if (rc[k] != PPF_OK && j.status == PPF_OK) {
          j.status = rc[k];
          j.error = err[k];
        }
        if (--r.jobs_left == 0) {
          lk.unlock();
          be_->release(r);
          lk.lock();
          retire_locked(r);
          int worst = r.status;
          std::string e = r.error;
          for (Job &x : r.jobs)
            if (worst == PPF_OK && x.status != PPF_OK && x.status != PPF_ERR_NO_VOTES) {
              worst = x.status;
              e = x.error;
            }
          finalize_locked(r, worst == PPF_OK ? State::Done : State::Failed, worst, e);
        }
      }
    }
  }

  Backend *be_;
  Config cfg_;
  std::mutex mu_;
  std::condition_variable sched_cv_, lane_cv_, done_cv_;
  std::map<uint64_t, std::unique_ptr<Request>> reqs_;
  std::deque<Request *> queue_;
  std::vector<Request *> active_;
  std::deque<std::pair<Request *, Job *>> finq_;
  std::vector<Request *> busy_;
  uint64_t inflight_ = 0;
  uint64_t next_ticket_ = 0;
  bool stopping_ = false, lanes_stop_ = false;
  Stats st_;
  std::thread sched_;
  std::vector<std::thread> lanes_;
};

}  // namespace ppfs
