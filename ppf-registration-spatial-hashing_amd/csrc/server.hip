// server.hip -- served form: the HIP backend of the request scheduler
// (server_core.h) and the ppf_server_* C ABI.
//
// Device work per request: the scene is uploaded on the scheduler's lane;
// each target model is a vote job whose candidate buffer lives until the job
// is finished; vote rounds are single persistent launches of the vote kernel
// over row slices of every voting job (VoteJob.r_begin selects the slice);
// finishing (threshold/order, poses, clustering) runs on the server's
// finishing lanes, concurrently with the next rounds; a lane finishes every
// queued job with the same clustering flags together (finish_together).
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ppf_batch.h"
#include "server_core.h"

using namespace ppfi;

namespace {

struct SceneState {
  ppf_scene scene = nullptr;
  ~SceneState() { ppf_scene_destroy(scene); }
};

struct JobState {
  uint32_t cap = 0;
  TBuf cand;  // (code u64 [cap], count u32 [cap])
  TBuf meta;  // [0] candidates, [1] running maximum, [2..3] votes
  uint32_t n_cand = 0, max_count = 0;
  int64_t votes = 0;
};

struct ResultHolder {
  ppf_result r = nullptr;
  ~ResultHolder() { ppf_result_destroy(r); }
};

class HipBackend : public ppfs::Backend {
 public:
  ppf_ctx root = nullptr;
  ppf_ctx sched = nullptr;          // uploads and vote rounds
  std::vector<ppf_ctx> lanes;       // finishing
  std::vector<ppf_model> models;
  int slots = 1;

  int parallel_slots() override { return slots; }

  uint64_t request_bytes(const ppfs::Request &r) override {
    // scene: SoA (7 floats per point) + frames of the reference rows; each
    // target: candidate buffer and its counters.  Results are allocated when
    // a target finishes and are accounted by the context, not here.
    uint64_t b = (uint64_t)r.n_points * 28 + (uint64_t)r.rows * 64 + 512;
    const uint32_t cap =
        r.params.candidate_capacity ? round_cap(r.params.candidate_capacity) : batch_cap(r.rows);
    for (size_t k = 0; k < r.model_ids.size(); ++k) b += (uint64_t)cap * 12 + 512;
    return b;
  }

  int upload(ppfs::Request &r, std::string &err) override {
    if (hipSetDevice(root->device) != hipSuccess) {
      err = "server: hipSetDevice failed";
      return PPF_ERR_HIP;
    }
    auto st = std::make_shared<SceneState>();
    int rc = scene_create(sched, r.points.data(), r.normals.data(), r.n_points, r.rpdf, &st->scene);
    if (rc) {
      err = ppf_last_error();
      return rc;
    }
    r.dev = st;
    r.points.clear();
    r.points.shrink_to_fit();
    r.normals.clear();
    r.normals.shrink_to_fit();
    return PPF_OK;
  }

  int start_pass(ppfs::Request &r, ppfs::Job &j, std::string &err) override {
    auto js = std::static_pointer_cast<JobState>(j.dev);
    if (!js) {
      js = std::make_shared<JobState>();
      js->cap = r.params.candidate_capacity ? round_cap(r.params.candidate_capacity) : batch_cap(r.rows);
      j.dev = js;
    } else {
      js->cap = std::max(js->cap, round_cap(js->n_cand));  // rerun after an overflow
    }
    int rc = 0;
    if (js->cand.bytes < (size_t)js->cap * 12) rc = js->cand.alloc(root, (size_t)js->cap * 12);
    if (!rc && !js->meta.ptr) rc = js->meta.alloc(root, 16);
    if (rc) {
      err = ppf_last_error();
      return rc;
    }
    // a rerun starts from the final maximum of the previous pass
    const uint32_t preset[4] = {0u, js->max_count, 0u, 0u};
    if (hipMemcpyAsync(js->meta.ptr, preset, 16, hipMemcpyHostToDevice, sched->stream) != hipSuccess ||
        hipStreamSynchronize(sched->stream) != hipSuccess) {
      err = "server: meta preset failed";
      return PPF_ERR_HIP;
    }
    return PPF_OK;
  }

  int vote_round(std::vector<ppfs::Slice> &slices, double *ms, std::string &err) override {
    TraceRange tr("ppf.server.vote_round");
    *ms = 0.0;
    for (int exact = 0; exact < 2; ++exact) {
      std::vector<ppfk::VoteJob> vj;
      uint64_t items = 0;
      int max_tm = 0;
      int64_t max_ns = 0;
      for (ppfs::Slice &s : slices) {
        if ((s.req->params.exact_votes != 0) != (exact != 0)) continue;
        auto sc = std::static_pointer_cast<SceneState>(s.req->dev);
        auto js = std::static_pointer_cast<JobState>(s.job->dev);
        ppf_model m = models[(size_t)s.job->model_id];
        auto *cc = js->cand.as<unsigned long long>();
        auto *cn = reinterpret_cast<uint32_t *>(cc + js->cap);
        ppfk::VoteJob J = vote_job(m, sc->scene, s.req->params.vote_count_threshold, cc, cn,
                                   js->meta.as<uint32_t>(), js->cap);
        J.per_ref = s.req->params.per_reference_argmax != 0;
        J.r_begin = (int32_t)s.row0;
        J.item_begin = items;
        items += (uint64_t)s.rows;
        max_tm = std::max(max_tm, (int)m->TM);
        max_ns = std::max(max_ns, sc->scene->cloud.n);
        vj.push_back(J);
      }
      if (vj.empty()) continue;
      float t = 0.0f;
      int rc = ppfk::run_votes(sched, vj.data(), (int)vj.size(), items, max_tm, max_ns, exact != 0, &t);
      if (rc) {
        err = ppf_last_error();
        return rc;
      }
      *ms += t;
    }
    return PPF_OK;
  }

  int pass_done(ppfs::Request &, ppfs::Job &j, bool *rerun, std::string &err) override {
    auto js = std::static_pointer_cast<JobState>(j.dev);
    uint32_t mh[4];
    if (hipMemcpy(mh, js->meta.ptr, 16, hipMemcpyDeviceToHost) != hipSuccess) {
      err = "server: reading vote counters failed";
      return PPF_ERR_HIP;
    }
    js->n_cand = mh[0];
    js->max_count = mh[1];
    js->votes = (int64_t)(((uint64_t)mh[3] << 32) | mh[2]);
    *rerun = js->n_cand > js->cap;
    if (*rerun && (j.passes >= 3 || js->n_cand > CAND_CAP_MAX)) {
      err = "server: kept-vote buffer overflow (" + std::to_string(js->n_cand) + " candidates)";
      return PPF_ERR_NOMEM;
    }
    return PPF_OK;
  }

  int finish(int lane, ppfs::Request &r, ppfs::Job &j, int *job_status, std::string &err) override {
    TraceRange tr("ppf.server.finish");
    ppf_ctx L = lanes[(size_t)lane];
    if (hipSetDevice(L->device) != hipSuccess) {
      err = "server: hipSetDevice failed";
      return PPF_ERR_HIP;
    }
    auto sc = std::static_pointer_cast<SceneState>(r.dev);
    auto js = std::static_pointer_cast<JobState>(j.dev);
    ppf_model m = models[(size_t)j.model_id];
    std::unique_ptr<ppf_result_s> res(new ppf_result_s());
    res->info.lookups = (int64_t)sc->scene->R * (sc->scene->cloud.n - 1);
    res->info.votes = js->votes;
    res->info.max_count = js->max_count;
    res->info.n_candidates = js->n_cand;
    auto *cc = js->cand.as<unsigned long long>();
    auto *cn = reinterpret_cast<uint32_t *>(cc + js->cap);
    int rc = finish_lookup(L, m, sc->scene, &r.params, cc, cn, js->n_cand, js->max_count, res.get());
    js->cand.release();
    if (rc) {
      err = ppf_last_error();
      return rc;
    }
    std::memcpy(r.poses[(size_t)j.target].data(), res->pose, 64);
    *job_status = res->info.status;
    auto h = std::make_shared<ResultHolder>();
    h->r = res.release();
    r.results[(size_t)j.target] = h;
    return PPF_OK;
  }

  // Jobs with GPU clustering and the same clustering flags are finished
  // together (finish_together: one launch per stage for the group).
  int64_t group_key(const ppfs::Request &r, const ppfs::Job &) override {
    if (r.params.cpu_clustering) return -1;
    return (r.params.use_l1_norm ? 1 : 0) | (r.params.use_averaged_clusters ? 2 : 0);
  }

  void finish_group(int lane, std::vector<std::pair<ppfs::Request *, ppfs::Job *>> &g, std::vector<int> &rc,
                    std::vector<int> &status, std::vector<std::string> &err) override {
    TraceRange tr("ppf.server.finish_group");
    ppf_ctx L = lanes[(size_t)lane];
    std::vector<FinishInput> in;
    for (auto &x : g) {
      auto sc = std::static_pointer_cast<SceneState>(x.first->dev);
      auto js = std::static_pointer_cast<JobState>(x.second->dev);
      auto *cc = js->cand.as<unsigned long long>();
      in.push_back(FinishInput{models[(size_t)x.second->model_id], sc->scene, x.first->params.vote_count_threshold,
                               cc, reinterpret_cast<const uint32_t *>(cc + js->cap), js->n_cand, js->max_count,
                               js->votes});
    }
    std::vector<std::unique_ptr<ppf_result_s>> res;
    int e = hipSetDevice(L->device) == hipSuccess ? PPF_OK : PPF_ERR_HIP;
    if (e == PPF_OK && !finish_together_ok(in, &g[0].first->params)) e = -1;
    if (e == PPF_OK) e = finish_together(L, in, &g[0].first->params, res, nullptr);
    if (e == -1) {  // keys too wide for one sort: one by one
      ppfs::Backend::finish_group(lane, g, rc, status, err);
      return;
    }
    const std::string msg = e != PPF_OK ? std::string(ppf_last_error()) : std::string();
    for (size_t k = 0; k < g.size(); ++k) {
      auto js = std::static_pointer_cast<JobState>(g[k].second->dev);
      js->cand.release();
      if (e != PPF_OK) {
        rc[k] = e;
        err[k] = msg;
        continue;
      }
      ppfs::Request &r = *g[k].first;
      const int32_t t = g[k].second->target;
      std::memcpy(r.poses[(size_t)t].data(), res[k]->pose, 64);
      status[k] = res[k]->info.status;
      auto h = std::make_shared<ResultHolder>();
      h->r = res[k].release();
      r.results[(size_t)t] = h;
      rc[k] = PPF_OK;
    }
  }

  void release(ppfs::Request &r) override {
    r.dev.reset();
    for (ppfs::Job &j : r.jobs) j.dev.reset();
  }
};

}  // namespace

struct ppf_server_s {
  ppf_ctx root = nullptr;
  std::vector<ppf_ctx> own;  // lanes created for this server
  std::unique_ptr<HipBackend> be;
  std::unique_ptr<ppfs::Scheduler> sched;
  ppfs::Config cfg;
};

namespace {

void free_lanes(std::vector<ppf_ctx> &v) {
  for (ppf_ctx l : v) {
    ctx_release(l);
    delete l;
  }
  v.clear();
}

#define S_ARG(cond, msg)  \
  do {                    \
    if (!(cond)) {        \
      set_error(msg);     \
      return PPF_ERR_ARG; \
    }                     \
  } while (0)

}  // namespace

extern "C" {

int ppf_server_create(ppf_ctx ctx, int32_t n_models, const ppf_model *models,
                      const ppf_server_config *cfg, ppf_server *out) {
  S_ARG(ctx && models && out, "ppf_server_create: NULL argument");
  *out = nullptr;
  S_ARG(n_models >= 1, "ppf_server_create: need at least one model");
  ppf_ctx root = ctx->root;
  for (int32_t k = 0; k < n_models; ++k)
    S_ARG(models[k] && models[k]->ctx == root, "ppf_server_create: NULL model or model of another context");
  PPF_HIP(hipSetDevice(root->device));
  std::unique_ptr<ppf_server_s> s(new ppf_server_s());
  s->root = root;
  ppfs::Config c;
  c.lanes = cfg && cfg->lanes > 0 ? cfg->lanes : 4;
  c.max_active = cfg && cfg->max_active > 0 ? cfg->max_active : 64;
  c.round_ms = cfg && cfg->round_ms > 0 ? cfg->round_ms : 50.0;
  c.tail_rows = 8.0;
  c.aging = 1.0;
  if (cfg && cfg->hbm_budget_bytes) {
    c.budget = cfg->hbm_budget_bytes;
  } else {
    size_t fr = 0, tot = 0;
    PPF_HIP(hipMemGetInfo(&fr, &tot));
    c.budget = (uint64_t)((double)fr * 0.8);
  }
  // a round should give every block of the vote launch a few rows
  c.min_round_rows = (int64_t)root->n_cu * 2 * 4;
  c.max_round_rows = (int64_t)1 << 22;
  for (int k = 0; k < c.lanes + 1; ++k) {
    auto *l = new ppf_ctx_s();
    l->root = root;
    int rc = ctx_init(l, root->device);
    s->own.push_back(l);
    if (rc) {
      free_lanes(s->own);
      return rc;
    }
  }
  s->be.reset(new HipBackend());
  s->be->root = root;
  s->be->sched = s->own[0];
  s->be->lanes.assign(s->own.begin() + 1, s->own.end());
  s->be->models.assign(models, models + n_models);
  int max_tm = 32;
  for (int32_t k = 0; k < n_models; ++k) max_tm = std::max(max_tm, (int)models[k]->TM);
  {
    int rc = ppfk::vote_slots(root, max_tm, &s->be->slots);
    if (rc) {
      s->be.reset();
      free_lanes(s->own);
      return rc;
    }
  }
  s->cfg = c;
  s->sched.reset(new ppfs::Scheduler(s->be.get(), c));
  *out = s.release();
  return PPF_OK;
}

int ppf_server_destroy(ppf_server s) {
  if (!s) return PPF_OK;
  (void)hipSetDevice(s->root->device);
  s->sched.reset();  // drains requests in flight, frees what nobody waited for
  s->be.reset();
  free_lanes(s->own);
  delete s;
  return PPF_OK;
}

int ppf_server_submit(ppf_server s, const float *points, const float *normals, int64_t n,
                      uint32_t rpdf, int32_t n_targets, const int32_t *model_ids,
                      const ppf_lookup_params *params, uint64_t *ticket) {
  S_ARG(s && points && normals && model_ids && params && ticket, "ppf_server_submit: NULL argument");
  *ticket = 0;
  S_ARG(n >= 2 && n < ((int64_t)1 << 31), "ppf_server_submit: need 2 <= n < 2^31 points");
  S_ARG(rpdf >= 1, "ppf_server_submit: ref_point_downsample_factor must be >= 1");
  S_ARG(n_targets >= 1 && n_targets <= 4096, "ppf_server_submit: need 1..4096 targets");
  const int32_t nm = (int32_t)s->be->models.size();
  for (int32_t k = 0; k < n_targets; ++k)
    S_ARG(model_ids[k] >= 0 && model_ids[k] < nm, "ppf_server_submit: model id out of range");
  std::unique_ptr<ppfs::Request> r(new ppfs::Request());
  r->n_points = n;
  r->rpdf = rpdf;
  r->rows = (n + rpdf - 1) / rpdf;
  r->points.assign(points, points + n * 3);
  r->normals.assign(normals, normals + n * 3);
  r->model_ids.assign(model_ids, model_ids + n_targets);
  r->params = *params;
  for (int32_t k = 0; k < n_targets; ++k) {
    ppfs::Job j;
    j.target = k;
    j.model_id = model_ids[k];
    j.rows = r->rows;
    j.row_units = (double)n * (double)s->be->models[(size_t)model_ids[k]]->cloud.n;
    r->jobs.push_back(std::move(j));
  }
  *ticket = s->sched->submit(std::move(r));
  return PPF_OK;
}

int ppf_server_wait(ppf_server s, uint64_t ticket, double timeout_s, float *poses_out,
                    int32_t *status_out, ppf_result *results_out, ppf_request_info_t *info) {
  S_ARG(s, "ppf_server_wait: NULL server");
  std::unique_ptr<ppfs::Request> r;
  int rc = s->sched->wait(ticket, timeout_s, &r);
  if (rc == PPF_ERR_TIMEOUT) {
    set_error("ppf_server_wait: request not done yet");
    return rc;
  }
  if (rc) {
    set_error("ppf_server_wait: unknown ticket");
    return rc;
  }
  const size_t nt = r->model_ids.size();
  for (size_t k = 0; k < nt; ++k) {
    if (poses_out) std::memcpy(poses_out + 16 * k, r->poses[k].data(), 64);
    if (status_out) status_out[k] = r->jobs[k].status;
    if (results_out) {
      results_out[k] = nullptr;
      if (r->results[k]) {
        auto h = std::static_pointer_cast<ResultHolder>(r->results[k]);
        results_out[k] = h->r;
        h->r = nullptr;
      }
    }
  }
  if (info) {
    info->status = r->status;
    info->n_targets = (int32_t)nt;
    info->vote_rounds = r->rounds;
    info->pad = 0;
    const bool admitted = r->t_admit.time_since_epoch().count() != 0;
    info->queue_ms =
        admitted ? std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count() : 0.0;
    info->latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    info->device_bytes = r->bytes;
  }
  if (r->status != PPF_OK) {
    set_error(r->error);
    return r->status;
  }
  return PPF_OK;
}

int ppf_server_poll(ppf_server s, uint64_t ticket, int32_t *done) {
  S_ARG(s && done, "ppf_server_poll: NULL argument");
  int d = 0;
  int rc = s->sched->poll(ticket, &d);
  if (rc) {
    set_error("ppf_server_poll: unknown ticket");
    return rc;
  }
  *done = d;
  return PPF_OK;
}

int ppf_server_cancel(ppf_server s, uint64_t ticket) {
  S_ARG(s, "ppf_server_cancel: NULL server");
  int rc = s->sched->cancel(ticket);
  if (rc == PPF_ERR_STATE) set_error("ppf_server_cancel: request already admitted");
  if (rc == PPF_ERR_ARG) set_error("ppf_server_cancel: unknown ticket");
  return rc;
}

int ppf_server_stats(ppf_server s, ppf_server_stats_t *st) {
  S_ARG(s && st, "ppf_server_stats: NULL argument");
  ppfs::Stats x = s->sched->stats();
  st->submitted = x.submitted;
  st->completed = x.completed;
  st->failed = x.failed;
  st->cancelled = x.cancelled;
  st->rejected = x.rejected;
  st->rounds = x.rounds;
  st->rows_voted = x.rows_voted;
  st->queued = x.queued;
  st->active = x.active;
  st->inflight_bytes = x.inflight_bytes;
  st->peak_inflight_bytes = x.peak_inflight_bytes;
  st->budget_bytes = s->cfg.budget;
  st->max_requests_in_round = x.max_requests_in_round;
  st->vote_ms = x.vote_ms;
  st->finish_groups = x.finish_groups;
  st->max_finish_group = x.max_finish_group;
  return PPF_OK;
}

}  // extern "C"
