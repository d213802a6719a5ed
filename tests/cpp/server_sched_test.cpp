// Host-only tests of the served form's scheduler (csrc/server_core.h) with a
// test double in place of the device backend: admission by budget,
// continuous batching across rounds, row coverage of every vote pass,
// overflow reruns, per-request error isolation, cancellation, timeouts and
// shutdown.  Prints "ok <name>" per passing check; exits non-zero on failure.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <chrono>
#include <thread>
#include <vector>

#include "server_core.h"

using namespace ppfs;

static int g_fail = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);          \
      g_fail = 1;                                                       \
    }                                                                   \
  } while (0)

struct FakeJob {
  std::vector<int64_t> rows;  // rows voted in the current pass, in order
  int passes = 0;
};

class FakeBackend : public Backend {
 public:
  std::mutex mu;
  std::vector<std::string> log;
  std::set<uint64_t> fail_upload, fail_finish, overflow_once;
  std::vector<std::set<uint64_t>> rounds;  // tickets voted in each round
  int64_t max_rows_in_round = 0;
  double row_us = 20.0;  // simulated device time per row
  std::atomic<int> finishing{0}, max_finishing{0};
  // grouping: jobs of even / odd model ids are finished together
  bool grouping = false;
  std::atomic<int> max_group{0}, groups{0};

  void note(const std::string &s) {
    std::lock_guard<std::mutex> g(mu);
    log.push_back(s);
  }
  uint64_t request_bytes(const Request &r) override { return (uint64_t)r.n_points * 100; }
  int upload(Request &r, std::string &err) override {
    note("upload " + std::to_string(r.ticket));
    if (fail_upload.count(r.ticket)) {
      err = "injected upload failure";
      return PPF_ERR_NOMEM;
    }
    r.dev = std::make_shared<int>(1);
    return PPF_OK;
  }
  int start_pass(Request &, Job &j, std::string &) override {
    auto f = std::static_pointer_cast<FakeJob>(j.dev);
    if (!f) {
      f = std::make_shared<FakeJob>();
      j.dev = f;
    }
    f->rows.clear();
    ++f->passes;
    return PPF_OK;
  }
  int vote_round(std::vector<Slice> &sl, double *ms, std::string &) override {
    int64_t rows = 0;
    std::set<uint64_t> t;
    for (Slice &s : sl) {
      auto f = std::static_pointer_cast<FakeJob>(s.job->dev);
      for (int64_t k = 0; k < s.rows; ++k) f->rows.push_back(s.row0 + k);
      rows += s.rows;
      t.insert(s.req->ticket);
    }
    {
      std::lock_guard<std::mutex> g(mu);
      rounds.push_back(t);
      max_rows_in_round = std::max(max_rows_in_round, rows);
    }
    std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(row_us * (double)rows)));
    *ms = row_us * (double)rows / 1000.0;
    return PPF_OK;
  }
  int pass_done(Request &r, Job &j, bool *rerun, std::string &) override {
    auto f = std::static_pointer_cast<FakeJob>(j.dev);
    *rerun = overflow_once.count(r.ticket) && f->passes == 1;
    return PPF_OK;
  }
  int finish(int, Request &r, Job &j, int *js, std::string &err) override {
    int now = ++finishing;
    int m = max_finishing.load();
    while (now > m && !max_finishing.compare_exchange_weak(m, now)) {
    }
    std::this_thread::sleep_for(std::chrono::milliseconds(2));
    --finishing;
    if (fail_finish.count(r.ticket) && j.target == 0) {
      err = "injected finish failure";
      return PPF_ERR_HIP;
    }
    auto f = std::static_pointer_cast<FakeJob>(j.dev);
    // pose[0] = number of rows voted in the final pass, pose[1] = 1 if they
    // were exactly 0..rows-1 in order, pose[2] = passes
    bool ordered = (int64_t)f->rows.size() == j.rows;
    for (size_t k = 0; ordered && k < f->rows.size(); ++k) ordered = f->rows[k] == (int64_t)k;
    r.poses[(size_t)j.target][0] = (float)f->rows.size();
    r.poses[(size_t)j.target][1] = ordered ? 1.0f : 0.0f;
    r.poses[(size_t)j.target][2] = (float)f->passes;
    *js = j.model_id == 99 ? PPF_ERR_NO_VOTES : PPF_OK;
    return PPF_OK;
  }
  int64_t group_key(const Request &, const Job &j) override { return grouping ? (j.model_id & 1) : -1; }
  void finish_group(int lane, std::vector<std::pair<Request *, Job *>> &g, std::vector<int> &rc,
                    std::vector<int> &js, std::vector<std::string> &err) override {
    ++groups;
    int m = max_group.load();
    while ((int)g.size() > m && !max_group.compare_exchange_weak(m, (int)g.size())) {
    }
    for (auto &x : g)
      if ((x.second->model_id & 1) != (g[0].second->model_id & 1) || x.second->status != PPF_OK) {
        std::printf("FAIL: a group mixes keys or holds a failed job\n");
        ++g_fail;
      }
    Backend::finish_group(lane, g, rc, js, err);
  }
  void release(Request &r) override {
    note("release " + std::to_string(r.ticket));
    r.dev.reset();
  }
};

static std::unique_ptr<Request> make_req(int64_t n, uint32_t rpdf, std::vector<int32_t> models) {
  std::unique_ptr<Request> r(new Request());
  r->n_points = n;
  r->rpdf = rpdf;
  r->rows = (n + rpdf - 1) / rpdf;
  r->model_ids = models;
  for (size_t k = 0; k < models.size(); ++k) {
    Job j;
    j.target = (int32_t)k;
    j.model_id = models[k];
    j.rows = r->rows;
    j.row_units = (double)n * (double)(100 + models[k]);
    r->jobs.push_back(j);
  }
  return r;
}

static int index_of(const std::vector<std::string> &log, const std::string &s) {
  for (size_t i = 0; i < log.size(); ++i)
    if (log[i] == s) return (int)i;
  return -1;
}

static void test_coverage_and_rounds() {
  FakeBackend be;
  Config c;
  c.budget = 1ull << 40;
  c.round_ms = 2.0;
  c.min_round_rows = 8;
  Scheduler s(&be, c);
  std::vector<uint64_t> t;
  t.push_back(s.submit(make_req(5000, 5, {0, 1, 2})));
  t.push_back(s.submit(make_req(300, 1, {3})));
  t.push_back(s.submit(make_req(20000, 7, {1})));
  for (uint64_t k : t) {
    std::unique_ptr<Request> r;
    CHECK(s.wait(k, 60.0, &r) == PPF_OK);
    CHECK(r->state == State::Done && r->status == PPF_OK);
    for (size_t j = 0; j < r->jobs.size(); ++j) {
      CHECK(r->poses[j][0] == (float)r->rows);
      CHECK(r->poses[j][1] == 1.0f);
      CHECK(r->poses[j][2] == 1.0f);
      CHECK(r->jobs[j].status == PPF_OK);
    }
    CHECK(r->rounds >= 1);
  }
  Stats st = s.stats();
  CHECK(st.completed == 3 && st.failed == 0 && st.inflight_bytes == 0);
  CHECK(st.rounds > 3);  // the large scene is voted over several rounds
  CHECK(st.max_requests_in_round >= 2);  // requests share rounds
  std::printf("ok coverage_and_rounds (%lld rounds)\n", (long long)st.rounds);
}

static void test_budget_admission() {
  FakeBackend be;
  Config c;
  c.budget = 600000;  // one 5000-point request (500000 bytes) at a time
  c.round_ms = 1.0;
  Scheduler s(&be, c);
  uint64_t a = s.submit(make_req(5000, 5, {0}));
  uint64_t b = s.submit(make_req(5000, 5, {1}));
  uint64_t big = s.submit(make_req(7000, 5, {2}));  // 700000 > budget: rejected
  std::unique_ptr<Request> ra, rb, rbig;
  CHECK(s.wait(a, 60.0, &ra) == PPF_OK && ra->status == PPF_OK);
  CHECK(s.wait(b, 60.0, &rb) == PPF_OK && rb->status == PPF_OK);
  CHECK(s.wait(big, 60.0, &rbig) == PPF_OK);
  CHECK(rbig->state == State::Failed && rbig->status == PPF_ERR_NOMEM);
  std::lock_guard<std::mutex> g(be.mu);
  const int ra_rel = index_of(be.log, "release " + std::to_string(a));
  const int rb_up = index_of(be.log, "upload " + std::to_string(b));
  CHECK(ra_rel >= 0 && rb_up > ra_rel);  // b admitted only after a released its memory
  CHECK(index_of(be.log, "upload " + std::to_string(big)) < 0);
  Stats st = s.stats();
  CHECK(st.rejected == 1 && st.peak_inflight_bytes <= c.budget);
  std::printf("ok budget_admission\n");
}

static void test_continuous_batching() {
  FakeBackend be;
  be.row_us = 200.0;
  Config c;
  c.budget = 1ull << 40;
  c.round_ms = 5.0;
  c.min_round_rows = 4;
  Scheduler s(&be, c);
  uint64_t big = s.submit(make_req(50000, 5, {0}));  // ~10000 rows: many rounds
  std::this_thread::sleep_for(std::chrono::milliseconds(30));
  int done = 0;
  CHECK(s.poll(big, &done) == PPF_OK && done == 0);
  uint64_t small = s.submit(make_req(1000, 5, {1}));  // arrives while big is voting
  std::unique_ptr<Request> rs, rb;
  CHECK(s.wait(small, 60.0, &rs) == PPF_OK && rs->status == PPF_OK);
  CHECK(s.poll(big, &done) == PPF_OK && done == 0);  // small finished first
  CHECK(s.wait(big, 120.0, &rb) == PPF_OK && rb->status == PPF_OK);
  bool shared = false;
  {
    std::lock_guard<std::mutex> g(be.mu);
    for (auto &r : be.rounds) shared = shared || (r.count(big) && r.count(small));
  }
  CHECK(shared);
  CHECK(rs->poses[0][1] == 1.0f && rb->poses[0][1] == 1.0f);
  std::printf("ok continuous_batching\n");
}

static void test_overflow_rerun() {
  FakeBackend be;
  be.overflow_once.insert(1);  // tickets are issued 1, 2, ...
  Config c;
  c.budget = 1ull << 40;
  c.round_ms = 1.0;
  Scheduler s(&be, c);
  uint64_t t = s.submit(make_req(3000, 3, {0, 1}));
  CHECK(t == 1);
  std::unique_ptr<Request> r;
  CHECK(s.wait(t, 60.0, &r) == PPF_OK && r->status == PPF_OK);
  for (size_t j = 0; j < r->jobs.size(); ++j) {
    CHECK(r->poses[j][2] == 2.0f);  // two passes
    CHECK(r->poses[j][1] == 1.0f);  // the second pass covered every row in order
    CHECK(r->jobs[j].passes == 2);
  }
  std::printf("ok overflow_rerun\n");
}

static void test_error_isolation() {
  FakeBackend be;
  Config c;
  c.budget = 1ull << 40;
  c.round_ms = 1.0;
  be.fail_upload.insert(2);
  be.fail_finish.insert(3);
  Scheduler s(&be, c);
  uint64_t a = s.submit(make_req(2000, 5, {0}));
  uint64_t b = s.submit(make_req(2000, 5, {0, 1}));      // upload fails
  uint64_t d = s.submit(make_req(2000, 5, {0, 1}));      // target 0 fails in finish
  uint64_t e = s.submit(make_req(2000, 5, {99, 1}));     // target 0: no votes (not an error)
  std::unique_ptr<Request> ra, rb, rd, re;
  CHECK(s.wait(a, 60.0, &ra) == PPF_OK && ra->status == PPF_OK);
  CHECK(s.wait(b, 60.0, &rb) == PPF_OK && rb->state == State::Failed && rb->status == PPF_ERR_NOMEM);
  CHECK(s.wait(d, 60.0, &rd) == PPF_OK && rd->state == State::Failed && rd->status == PPF_ERR_HIP);
  CHECK(rd->jobs[0].status == PPF_ERR_HIP && rd->jobs[1].status == PPF_OK);
  CHECK(s.wait(e, 60.0, &re) == PPF_OK && re->state == State::Done && re->status == PPF_OK);
  CHECK(re->jobs[0].status == PPF_ERR_NO_VOTES && re->jobs[1].status == PPF_OK);
  Stats st = s.stats();
  CHECK(st.inflight_bytes == 0 && st.active == 0);
  std::printf("ok error_isolation\n");
}

static void test_cancel_timeout_shutdown() {
  FakeBackend be;
  be.row_us = 500.0;
  Config c;
  c.budget = 600000;
  c.round_ms = 5.0;
  auto s = std::make_unique<Scheduler>(&be, c);
  uint64_t a = s->submit(make_req(5000, 5, {0}));   // ~0.5 s of simulated voting
  uint64_t b = s->submit(make_req(5000, 5, {0}));   // waits for budget
  std::unique_ptr<Request> r;
  CHECK(s->wait(a, 0.001, &r) == PPF_ERR_TIMEOUT);
  CHECK(s->cancel(b) == PPF_OK);
  // a is admitted by the scheduler thread; on a loaded host that can take a
  // while, and only an admitted request refuses cancellation
  for (int i = 0; i < 30000 && s->stats().active < 1; ++i)
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  CHECK(s->cancel(a) == PPF_ERR_STATE);
  CHECK(s->wait(b, 30.0, &r) == PPF_OK && r->state == State::Cancelled && r->status == PPF_ERR_CANCELLED);
  uint64_t c2 = s->submit(make_req(5000, 5, {1}));  // queued behind a
  s->shutdown();                                     // a completes, c2 is cancelled
  CHECK(s->wait(a, 30.0, &r) == PPF_OK && r->state == State::Done);
  CHECK(s->wait(c2, 30.0, &r) == PPF_OK && r->state == State::Cancelled);
  uint64_t late = s->submit(make_req(100, 1, {0}));
  CHECK(s->wait(late, 30.0, &r) == PPF_OK && r->state == State::Cancelled);
  CHECK(s->wait(12345, 0.0, &r) == PPF_ERR_ARG);
  {
    std::lock_guard<std::mutex> g(be.mu);
    CHECK(index_of(be.log, "upload " + std::to_string(b)) < 0);
  }
  s.reset();
  std::printf("ok cancel_timeout_shutdown\n");
}

static void test_lanes_overlap() {
  FakeBackend be;
  Config c;
  c.budget = 1ull << 40;
  c.lanes = 4;
  c.round_ms = 1.0;
  Scheduler s(&be, c);
  std::vector<uint64_t> t;
  // one row per job: every job leaves voting in the first round
  for (int k = 0; k < 8; ++k) t.push_back(s.submit(make_req(8, 8, {0, 1, 2, 3})));
  for (uint64_t k : t) {
    std::unique_ptr<Request> r;
    CHECK(s.wait(k, 60.0, &r) == PPF_OK && r->status == PPF_OK);
  }
  CHECK(be.max_finishing.load() >= 2);  // several lanes finished jobs at once
  std::printf("ok lanes_overlap (max %d)\n", be.max_finishing.load());
}

static void test_group_finishing() {
  FakeBackend be;
  be.grouping = true;
  be.fail_finish.insert(5);
  Config c;
  c.budget = 1ull << 40;
  c.lanes = 1;
  c.round_ms = 1.0;
  Scheduler s(&be, c);
  std::vector<uint64_t> t;
  // one row per job: all jobs leave voting in the first rounds and queue up
  // behind the single lane, which then takes them in groups by key
  for (int k = 0; k < 10; ++k) t.push_back(s.submit(make_req(8, 8, {0, 1, 2, 3, 99})));
  for (uint64_t k : t) {
    std::unique_ptr<Request> r;
    CHECK(s.wait(k, 60.0, &r) == PPF_OK);
    if (k == 5) {
      // target 0 fails in finish; the request fails, its other jobs finish
      CHECK(r->state == State::Failed && r->status == PPF_ERR_HIP);
      continue;
    }
    CHECK(r->state == State::Done && r->status == PPF_OK);
    for (size_t j = 0; j < r->poses.size(); ++j) CHECK(r->poses[j][0] == 1.0f && r->poses[j][1] == 1.0f);
  }
  CHECK(be.max_group.load() >= 2);
  std::printf("ok group_finishing (%d groups, max %d jobs)\n", be.groups.load(), be.max_group.load());
}

int main() {
  test_coverage_and_rounds();
  test_budget_admission();
  test_continuous_batching();
  test_overflow_rerun();
  test_error_isolation();
  test_cancel_timeout_shutdown();
  test_lanes_overlap();
  test_group_finishing();
  if (g_fail) {
    std::printf("FAILED\n");
    return 1;
  }
  std::printf("ALL OK\n");
  return 0;
}
