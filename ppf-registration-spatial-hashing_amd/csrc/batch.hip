// batch.hip -- batched registration: many (model, scene) pairs of uneven
// sizes voted in one persistent launch over device-resident model hash
// arrays, then finished (threshold/order, poses, clustering) for all jobs
// together, one launch per stage (finish_batch), or job by job concurrently
// on the context's lanes.
//
// The reference registers one pair at a time and rebuilds the model for every
// scene (src/cuda/ppf.cu:63-98); each pair's result here is bit-identical to
// ppf_lookup on the same pair, because a job's votes depend only on its own
// (model, scene) and its candidates are pruned against its own running
// maximum.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ppf_batch.h"

using namespace ppfi;

namespace {

#define B_ARG(cond, msg)    \
  do {                      \
    if (!(cond)) {          \
      set_error(msg);       \
      return PPF_ERR_ARG;   \
    }                       \
  } while (0)

}  // namespace

namespace ppfi {

int batch_votes(ppf_ctx ctx, std::vector<BatchJob> &jobs, bool exact, float *vote_ms,
                int *launches, bool *graphs) {
  *vote_ms = 0.0f;
  *launches = 0;
  const bool want_graph = graphs && *graphs && ctx->root->use_graphs != 0;
  if (graphs) *graphs = want_graph;
  const size_t nj = jobs.size();
  if (nj == 0) return 0;
  TBuf meta;
  PPF_TRY(meta.alloc(ctx, nj * 16));
  uint32_t *md = meta.as<uint32_t>();
  std::vector<uint32_t> mh(nj * 4, 0u);
  std::vector<size_t> todo(nj);
  std::iota(todo.begin(), todo.end(), (size_t)0);
  // with progress reports (ppf_ctx_set_progress), the jobs' reference rows in
  // slices of progress_rows, voted in launches of about that many rows, with
  // a report after each launch; a job's slices share its candidates and
  // running maximum, as the plain lookup's row launches do
  const ppf_ctx_s *root = ctx->root;
  const int64_t step = root->progress && root->progress_rows > 0 ? root->progress_rows : 0;
  for (int attempt = 0; !todo.empty(); ++attempt) {
    std::vector<ppfk::VoteJob> vj;  // one per (job, row slice), in launch order
    std::vector<int64_t> rows;      // rows of each entry
    int max_tm = 0;
    int64_t max_ns = 0, total_rows = 0;
    for (size_t j : todo) {
      BatchJob &b = jobs[j];
      if (b.cand.bytes < (size_t)b.cap * 12) PPF_TRY(b.cand.alloc(ctx, (size_t)b.cap * 12));
      auto *cc = b.cand.as<unsigned long long>();
      auto *cn = reinterpret_cast<uint32_t *>(cc + b.cap);
      ppfk::VoteJob J = vote_job(b.m, b.s, b.thr, cc, cn, md + 4 * j, b.cap);
      J.per_ref = b.per_ref;
      const int64_t R = b.s->R, sl = step > 0 ? step : std::max<int64_t>(R, 1);
      for (int64_t r0 = 0; r0 < R; r0 += sl) {
        J.r_begin = (int32_t)r0;
        vj.push_back(J);
        rows.push_back(std::min(sl, R - r0));
      }
      total_rows += R;
      max_tm = std::max(max_tm, (int)b.m->TM);
      max_ns = std::max(max_ns, b.s->cloud.n);
      // preset: no candidates yet, running maximum from the previous pass
      mh[4 * j] = 0u;
      mh[4 * j + 2] = mh[4 * j + 3] = 0u;
    }
    PPF_HIP(hipMemcpyAsync(md, mh.data(), nj * 16, hipMemcpyHostToDevice, ctx->stream));
    int64_t done = 0;
    for (size_t a = 0; a < vj.size();) {
      // one launch: the whole list, or consecutive slices of about `step` rows
      size_t e = a;
      uint64_t items = 0;
      while (e < vj.size() && (step == 0 || e == a || (int64_t)items + rows[e] <= step)) {
        vj[e].item_begin = items;
        items += (uint64_t)rows[e];
        ++e;
      }
      float ms = 0.0f;
      bool g = want_graph && a == 0 && e == vj.size();
      PPF_TRY(ppfk::run_votes(ctx, vj.data() + a, (int)(e - a), items, max_tm, max_ns, exact, &ms, 0, &g));
      if (graphs) *graphs = *graphs && g;
      *vote_ms += ms;
      ++*launches;
      done += (int64_t)items;
      if (step > 0) root->progress(root->progress_user, "vote", done, total_rows);
      a = e;
    }
    PPF_HIP(hipMemcpy(mh.data(), md, nj * 16, hipMemcpyDeviceToHost));
    std::vector<size_t> again;
    for (size_t j : todo) {
      BatchJob &b = jobs[j];
      b.n_cand = mh[4 * j];
      b.max_count = mh[4 * j + 1];
      b.votes = (int64_t)(((uint64_t)mh[4 * j + 3] << 32) | mh[4 * j + 2]);
      if (b.n_cand <= b.cap) continue;
      if (attempt >= 2 || b.n_cand > CAND_CAP_MAX) {
        set_error("batch: kept-vote buffer overflow (" + std::to_string(b.n_cand) + " candidates)");
        return PPF_ERR_NOMEM;
      }
      // the rerun starts from the final maximum, so it emits exactly the
      // kept votes, of which there are at most n_cand
      b.cap = std::max(b.cap, round_cap(b.n_cand));
      again.push_back(j);
    }
    todo.swap(again);
  }
  return 0;
}

void make_batch_jobs(int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                     const ppf_lookup_params *params, std::vector<size_t> &order, std::vector<BatchJob> &q) {
  order.resize((size_t)n_jobs);
  std::iota(order.begin(), order.end(), (size_t)0);
  // longest jobs first in the item queue (work ~ R * n_s * n_m)
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    auto w = [&](size_t j) {
      return (double)scenes[j]->R * (double)scenes[j]->cloud.n * (double)models[j]->cloud.n;
    };
    return w(a) > w(b);
  });
  q.clear();
  q.reserve((size_t)n_jobs);
  for (size_t j : order) {
    BatchJob b;
    b.m = models[j];
    b.s = scenes[j];
    b.thr = params->vote_count_threshold;
    b.per_ref = params->per_reference_argmax != 0;
    b.cap = params->candidate_capacity ? round_cap(params->candidate_capacity) : batch_cap(scenes[j]->R);
    b.index = (int64_t)j;
    q.push_back(std::move(b));
  }
}

namespace {
// The kept votes and clustering arrays of a batch finished together; each
// job's result views its range and shares the ownership.
struct BatchStore {
  TBuf codes, counts, T, vc, trans, quats;
};
}  // namespace

bool finish_together_ok(const std::vector<FinishInput> &in, const ppf_lookup_params *p) {
  if (p->cpu_clustering || in.empty()) return false;
  uint64_t C = 0;
  uint32_t mx = 0;
  int64_t ns = 0, nm = 0;
  for (const FinishInput &b : in) {
    C += b.n_cand;
    mx = std::max(mx, b.max_count);
    ns = std::max(ns, b.s->cloud.n);
    nm = std::max(nm, b.m->cloud.n);
  }
  return ppfk::finish_batch_fits((int)in.size(), C, mx, ns, nm);
}

int finish_together(ppf_ctx ctx, const std::vector<FinishInput> &in, const ppf_lookup_params *p,
                    std::vector<std::unique_ptr<ppf_result_s>> &results, ppf_batch_info_t *stages) {
  ppf_ctx root = ctx->root;
  const size_t nj = in.size();
  std::vector<ppfk::FinJob> fj(nj);
  std::vector<uint32_t> n_cand(nj);
  uint32_t mx = 0;
  int64_t ns = 0, nm = 0;
  for (size_t j = 0; j < nj; ++j) {
    const FinishInput &b = in[j];
    fj[j].cand_code = b.cand_code;
    fj[j].cand_count = b.cand_count;
    fj[j].thr_f = b.thr * (float)b.max_count;  // model.cu:164, as select_and_sort
    fj[j].rpdf = b.s->rpdf;
    fj[j].mframe = b.m->frames.as<float>();
    fj[j].sframe = b.s->frames.as<float>();
    fj[j].weights = b.m->weights.as<float>();
    fj[j].d_dist = b.m->d_dist;
    n_cand[j] = b.n_cand;
    mx = std::max(mx, b.max_count);
    ns = std::max(ns, b.s->cloud.n);
    nm = std::max(nm, b.m->cloud.n);
  }
  auto store = std::make_shared<BatchStore>();
  auto alloc = [&](int64_t P, ppfk::FinishArrays &A) -> int {
    const size_t n = (size_t)std::max<int64_t>(P, 1);
    PPF_TRY(store->codes.alloc(root, n * 8));
    PPF_TRY(store->counts.alloc(root, n * 4));
    PPF_TRY(store->T.alloc(root, n * 64));
    PPF_TRY(store->vc.alloc(root, n * 4));
    PPF_TRY(store->trans.alloc(root, n * 12));
    PPF_TRY(store->quats.alloc(root, n * 16));
    A.codes = store->codes.as<unsigned long long>();
    A.counts = store->counts.as<uint32_t>();
    A.T = store->T.as<float>();
    A.vc = store->vc.as<float>();
    A.trans = store->trans.as<float>();
    A.quats = store->quats.as<float>();
    return 0;
  };
  ppfk::FinishOut out;
  PPF_TRY(ppfk::finish_batch(ctx, fj, n_cand, mx, ns, nm, p->use_l1_norm != 0, p->use_averaged_clusters != 0,
                             alloc, out));
  results.clear();
  results.resize(nj);
  for (size_t j = 0; j < nj; ++j) {
    const FinishInput &b = in[j];
    std::unique_ptr<ppf_result_s> r(new ppf_result_s());
    r->ctx = root;
    r->n_m = b.m->cloud.n;
    r->cpu = false;
    r->info.lookups = (int64_t)b.s->R * (b.s->cloud.n - 1);
    r->info.votes = b.votes;
    r->info.max_count = b.max_count;
    r->info.n_candidates = b.n_cand;
    const int64_t k0 = (int64_t)out.kbeg[j], P = (int64_t)out.kbeg[j + 1] - k0;
    r->P = P;
    r->info.n_kept = P;
    r->info.max_idx = out.max_idx[j];
    // the group's stage times: every job's finishing ran in the same launches
    r->info.select_ms = out.select_ms;
    r->info.pose_ms = out.pose_ms;
    r->info.cluster_ms = out.cluster_ms;
    r->info.cluster_kernel_ms = out.cluster_kernel_ms;
    r->info.cluster_tests = out.tests[j];
    r->info.cluster_mixed_chunks = out.mixed[j];
    if (P > 0) {
      r->codes.set_view(store->codes.as<char>() + (size_t)k0 * 8, (size_t)P * 8);
      r->counts.set_view(store->counts.as<char>() + (size_t)k0 * 4, (size_t)P * 4);
      r->T.set_view(store->T.as<char>() + (size_t)k0 * 64, (size_t)P * 64);
      r->vc.set_view(store->vc.as<char>() + (size_t)k0 * 4, (size_t)P * 4);
      r->trans.set_view(store->trans.as<char>() + (size_t)k0 * 12, (size_t)P * 12);
      r->quats.set_view(store->quats.as<char>() + (size_t)k0 * 16, (size_t)P * 16);
      r->store = store;
      std::memcpy(r->pose, out.pose.data() + 16 * j, 64);
      r->info.status = PPF_OK;
    } else {
      r->info.status = PPF_ERR_NO_VOTES;
    }
    results[j] = std::move(r);
  }
  if (stages) {
    stages->fused = 1;
    stages->select_ms = out.select_ms;
    stages->pose_ms = out.pose_ms;
    stages->cluster_ms = out.cluster_ms;
    stages->cluster_kernel_ms = out.cluster_kernel_ms;
  }
  return 0;
}

namespace {
std::vector<FinishInput> finish_inputs(const std::vector<BatchJob> &jobs) {
  std::vector<FinishInput> in(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) {
    const BatchJob &b = jobs[j];
    auto *cc = b.cand.as<unsigned long long>();
    in[j] = FinishInput{b.m, b.s, b.thr, cc, reinterpret_cast<const uint32_t *>(cc + b.cap), b.n_cand,
                        b.max_count, b.votes};
  }
  return in;
}
}  // namespace

bool batch_fused_ok(const std::vector<BatchJob> &jobs, const ppf_lookup_params *p) {
  return finish_together_ok(finish_inputs(jobs), p);
}

int batch_finish_fused(ppf_ctx ctx, std::vector<BatchJob> &jobs, const ppf_lookup_params *p,
                       std::vector<std::unique_ptr<ppf_result_s>> &results, ppf_batch_info_t *stages) {
  return finish_together(ctx, finish_inputs(jobs), p, results, stages);
}

int batch_finish(ppf_ctx ctx, std::vector<BatchJob> &jobs, const ppf_lookup_params *p, int n_lanes,
                 std::vector<std::unique_ptr<ppf_result_s>> &results) {
  const size_t nj = jobs.size();
  results.clear();
  results.resize(nj);
  std::vector<ppf_ctx> lanes;
  PPF_TRY(ctx_lanes(ctx->root, std::max(1, std::min<int>(n_lanes, (int)nj)), lanes));
  // largest candidate sets first
  std::vector<size_t> order(nj);
  std::iota(order.begin(), order.end(), (size_t)0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return jobs[a].n_cand > jobs[b].n_cand; });
  std::atomic<size_t> next{0};
  std::vector<int> rc(lanes.size(), 0);
  std::vector<std::string> err(lanes.size());
  auto work = [&](size_t li) {
    ppf_ctx lane = lanes[li];
    if (hipSetDevice(lane->device) != hipSuccess) {
      rc[li] = PPF_ERR_HIP;
      err[li] = "batch: hipSetDevice failed";
      return;
    }
    for (;;) {
      const size_t k = next.fetch_add(1);
      if (k >= nj || rc[li]) break;
      BatchJob &b = jobs[order[k]];
      std::unique_ptr<ppf_result_s> r(new ppf_result_s());
      r->info.lookups = (int64_t)b.s->R * (b.s->cloud.n - 1);
      r->info.votes = b.votes;
      r->info.max_count = b.max_count;
      r->info.n_candidates = b.n_cand;
      auto *cc = b.cand.as<unsigned long long>();
      auto *cn = reinterpret_cast<uint32_t *>(cc + b.cap);
      int e = finish_lookup(lane, b.m, b.s, p, cc, cn, b.n_cand, b.max_count, r.get());
      if (e) {
        rc[li] = e;
        err[li] = std::string("batch job ") + std::to_string(b.index) + ": " + ppf_last_error();
        break;
      }
      results[order[k]] = std::move(r);
    }
  };
  if (lanes.size() == 1) {
    work(0);
  } else {
    std::vector<std::thread> th;
    for (size_t li = 0; li < lanes.size(); ++li) th.emplace_back(work, li);
    for (auto &t : th) t.join();
  }
  for (size_t li = 0; li < lanes.size(); ++li)
    if (rc[li]) {
      set_error(err[li]);
      results.clear();
      return rc[li];
    }
  return 0;
}

}  // namespace ppfi

extern "C" {

int ppf_batch_lookup(ppf_ctx ctx, int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                     const ppf_lookup_params *params, ppf_result *results_out,
                     ppf_batch_info_t *info) {
  B_ARG(ctx && models && scenes && params && results_out, "ppf_batch_lookup: NULL argument");
  B_ARG(n_jobs >= 1 && n_jobs <= PPF_MAX_BATCH_JOBS, "ppf_batch_lookup: n_jobs out of range");
  ppf_ctx root = ctx->root;
  for (int32_t j = 0; j < n_jobs; ++j) {
    results_out[j] = nullptr;
    B_ARG(models[j] && scenes[j], "ppf_batch_lookup: NULL model or scene");
    B_ARG(models[j]->ctx == root && scenes[j]->ctx == root,
          "ppf_batch_lookup: objects belong to another context");
  }
  PPF_HIP(hipSetDevice(root->device));
  TraceRange tr("ppf.batch");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<size_t> order;
  std::vector<BatchJob> q;
  make_batch_jobs(n_jobs, models, scenes, params, order, q);
  float vote_ms = 0.0f;
  int launches = 0;
  bool graphs = true;
  {
    TraceRange trv("ppf.batch.vote");
    PPF_TRY(batch_votes(root, q, params->exact_votes != 0, &vote_ms, &launches, &graphs));
  }
  std::vector<std::unique_ptr<ppf_result_s>> res;
  const auto t1 = std::chrono::steady_clock::now();
  const int lanes = root->tune_lanes > 0 ? root->tune_lanes : PPF_BATCH_LANES;
  ppf_batch_info_t stages{};
  // one launch per finishing stage over all jobs, unless the context asks for
  // lanes (tuner or diagnostics) or the batch's sort keys do not fit
  const bool fused = root->finish_mode != PPF_FINISH_LANES && batch_fused_ok(q, params);
  if (fused)
    PPF_TRY(batch_finish_fused(root, q, params, res, &stages));
  else
    PPF_TRY(batch_finish(root, q, params, lanes, res));
  const auto t2 = std::chrono::steady_clock::now();
  int64_t L = 0, V = 0;
  for (size_t k = 0; k < order.size(); ++k) {
    res[k]->info.vote_ms = vote_ms;
    res[k]->info.total_ms = vote_ms + res[k]->info.select_ms + res[k]->info.pose_ms + res[k]->info.cluster_ms;
    L += res[k]->info.lookups;
    V += res[k]->info.votes;
    results_out[order[k]] = res[k].release();
  }
  if (info) {
    info->n_jobs = n_jobs;
    info->lookups = L;
    info->votes = V;
    info->vote_launches = launches;
    info->lanes = fused ? 0 : std::min<int>(lanes, n_jobs);
    info->fused = fused ? 1 : 0;
    info->vote_graph = graphs ? 1 : 0;
    info->select_ms = stages.select_ms;
    info->pose_ms = stages.pose_ms;
    info->cluster_ms = stages.cluster_ms;
    info->cluster_kernel_ms = stages.cluster_kernel_ms;
    info->vote_ms = vote_ms;
    info->finish_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    info->total_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  }
  return PPF_OK;
}

}  // extern "C"
