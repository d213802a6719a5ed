// tune.hip -- the batched form's autotuner (ppf_batch_tune).  No counterpart
// in the reference, whose launch shapes are constants (BLOCK_SIZE /
// MAX_NBLOCKS, include/kernel.h:9-10).
//
// Three launch parameters are chosen by timing candidates on the caller's
// own batch, and kept on the context for every later ppf_batch_lookup (and
// the server, which shares the vote path):
//   * vote blocks per CU (1 or the most the LDS allows, 2): two blocks hide
//     latency, one leaves the whole CU to a block;
//   * finishing: all jobs together (one launch per stage), or job by job on
//     1, 2 or 4 lanes (host threads with their own streams);
//   * one-sweep radix-sort tile (256 or 512 threads x 8 or 16 keys).
// None of them changes a result: vote counts are integer atomics, every job
// is finished by the same stage sequence on whatever lane, and the sort is
// stable for any tile.
#include <algorithm>
#include <chrono>
#include <random>

#include "ppf_batch.h"

using namespace ppfi;

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int ppf_batch_tune(ppf_ctx ctx, int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                   const ppf_lookup_params *params, ppf_tune_info_t *info) {
  if (!(ctx && models && scenes && params && info) || n_jobs < 1 || n_jobs > PPF_MAX_BATCH_JOBS) {
    set_error("ppf_batch_tune: bad arguments");
    return PPF_ERR_ARG;
  }
  ppf_ctx root = ctx->root;
  for (int32_t j = 0; j < n_jobs; ++j)
    if (!models[j] || !scenes[j] || models[j]->ctx != root || scenes[j]->ctx != root) {
      set_error("ppf_batch_tune: NULL object or one of another context");
      return PPF_ERR_ARG;
    }
  PPF_HIP(hipSetDevice(root->device));
  const auto t_start = std::chrono::steady_clock::now();
  *info = ppf_tune_info_t{};
  std::vector<size_t> order;
  std::vector<BatchJob> q;
  make_batch_jobs(n_jobs, models, scenes, params, order, q);

  // ---- vote blocks per CU: the fastest of two timed vote passes each
  int max_per_cu = 0;
  {
    int max_tm = 0;
    for (const BatchJob &b : q) max_tm = std::max(max_tm, (int)b.m->TM);
    int slots = 0;
    PPF_TRY(ppfk::vote_slots(root, max_tm, &slots));
    max_per_cu = std::max(1, slots / std::max(1, root->n_cu));
  }
  double best_vote = 1e300;
  int best_per_cu = max_per_cu;
  for (int per_cu = max_per_cu; per_cu >= 1; --per_cu) {
    root->tune_vote_per_cu = per_cu;
    double t = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      float ms = 0.0f;
      int launches = 0;
      PPF_TRY(batch_votes(root, q, params->exact_votes != 0, &ms, &launches));
      t = std::min(t, (double)ms);
    }
    if (t < best_vote) {
      best_vote = t;
      best_per_cu = per_cu;
    }
  }
  root->tune_vote_per_cu = best_per_cu;
  // candidates of the chosen shape for the finishing runs
  {
    float ms = 0.0f;
    int launches = 0;
    PPF_TRY(batch_votes(root, q, params->exact_votes != 0, &ms, &launches));
  }

  // ---- one-sweep tile: sorts shaped like the finishing stages' (candidate
  // keys alone, pose keys with values), sized by the batch's largest
  // candidate set
  uint32_t max_cand = 1u << 16;
  for (const BatchJob &b : q) max_cand = std::max(max_cand, b.n_cand);
  const int64_t n_sort = std::min<int64_t>((int64_t)max_cand, (int64_t)1 << 26);
  {
    std::vector<uint64_t> hk((size_t)n_sort);
    std::mt19937_64 rng(7);
    for (auto &x : hk) x = rng();
    DevBuf keys, keys0, vals;
    PPF_TRY(keys.ensure((size_t)n_sort * 8));
    PPF_TRY(keys0.ensure((size_t)n_sort * 8));
    PPF_TRY(vals.ensure((size_t)n_sort * 4));
    PPF_HIP(hipMemcpy(keys0.ptr, hk.data(), (size_t)n_sort * 8, hipMemcpyHostToDevice));
    double best_sort = 1e300;
    int best_items = 16, best_block = 256;
    for (const auto shape : {std::pair<int, int>{256, 16}, {512, 16}, {512, 8}, {256, 8}}) {
      const int block = shape.first, items = shape.second;
      root->sort.tile_block = block;
      root->sort.tile_items = items;
      double t = 1e300;
      for (int rep = 0; rep < 3; ++rep) {
        PPF_HIP(hipMemcpyAsync(keys.ptr, keys0.ptr, (size_t)n_sort * 8, hipMemcpyDeviceToDevice, root->stream));
        PPF_HIP(hipStreamSynchronize(root->stream));
        const auto t0 = std::chrono::steady_clock::now();
        PPF_TRY(radix_sort_u64(keys.as<uint64_t>(), nullptr, n_sort, 0, 40, root->sort, root->stream));
        PPF_TRY(radix_sort_u64(keys.as<uint64_t>(), vals.as<uint32_t>(), n_sort, 0, 64, root->sort,
                               root->stream));
        PPF_HIP(hipStreamSynchronize(root->stream));
        t = std::min(t, ms_since(t0));
      }
      if (t < best_sort) {
        best_sort = t;
        best_items = items;
        best_block = block;
      }
    }
    root->sort.tile_items = best_items;
    root->sort.tile_block = best_block;
    for (ppf_ctx_s *l : root->lanes) {
      l->sort.tile_items = best_items;
      l->sort.tile_block = best_block;
    }
    info->sort_tile_keys = best_block * best_items;
    info->sort_tile_threads = best_block;
    info->best_sort_ms = best_sort;
  }

  // ---- finishing: the same candidates finished together (one launch per
  // stage over all jobs; lanes = 0) and with 4, 2 and 1 lanes
  double best_finish = 1e300;
  int best_lanes = PPF_BATCH_LANES;
  const bool fused_ok = batch_fused_ok(q, params);
  for (int lanes : {0, PPF_BATCH_LANES, 2, 1}) {
    if (lanes == 0 && !fused_ok) continue;
    std::vector<std::unique_ptr<ppf_result_s>> res;
    const auto t0 = std::chrono::steady_clock::now();
    if (lanes == 0)
      PPF_TRY(batch_finish_fused(root, q, params, res, nullptr));
    else
      PPF_TRY(batch_finish(root, q, params, lanes, res));
    res.clear();
    const double t = ms_since(t0);
    if (t < best_finish) {
      best_finish = t;
      best_lanes = lanes;
    }
  }
  root->tune_lanes = best_lanes > 0 ? best_lanes : 0;
  root->finish_mode = best_lanes > 0 ? PPF_FINISH_LANES : PPF_FINISH_AUTO;

  info->vote_blocks_per_cu = best_per_cu;
  info->finish_lanes = best_lanes;
  info->best_vote_ms = best_vote;
  info->best_finish_ms = best_finish;
  info->tune_ms = ms_since(t_start);
  return PPF_OK;
}

int ppf_tune_reset(ppf_ctx ctx) {
  if (!ctx) {
    set_error("ppf_tune_reset: NULL context");
    return PPF_ERR_ARG;
  }
  ppf_ctx root = ctx->root;
  root->tune_vote_per_cu = 0;
  root->tune_lanes = 0;
  root->finish_mode = PPF_FINISH_AUTO;
  root->sort.tile_items = 16;
  root->sort.tile_block = 256;
  for (ppf_ctx_s *l : root->lanes) {
    l->sort.tile_items = 16;
    l->sort.tile_block = 256;
  }
  return PPF_OK;
}

}  // extern "C"
