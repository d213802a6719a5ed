// ppf_batch.h -- shared pieces of the batched and served forms.
#pragma once

#include <chrono>
#include <memory>
#include <vector>

#include "ppf_objects.h"

// Lanes (streams with their own scratch and host thread) that finish batched
// jobs concurrently.  Four matches the hardware queues HIP gives a process by
// default.
#define PPF_BATCH_LANES 4

namespace ppfi {

// One (model, scene) pair of a batch, with its candidate buffer and vote
// statistics.
struct BatchJob {
  ppf_model m = nullptr;
  ppf_scene s = nullptr;
  float thr = 0.0f;
  int32_t per_ref = 0;   // per-reference-point argmax
  uint32_t cap = 0;      // candidate capacity
  TBuf cand;             // (code u64 [cap], count u32 [cap])
  uint32_t n_cand = 0;   // candidates emitted by the last pass
  uint32_t max_count = 0;
  int64_t votes = 0;
  int64_t index = 0;     // position in the caller's list
};

// Same initial capacity as a plain lookup: a job whose candidates overflow
// is voted again (rerun launch), which costs far more than the 12 B per
// candidate of a larger buffer.
inline uint32_t batch_cap(int64_t R) {
  return round_cap((uint64_t)std::max<int64_t>(CAND_CAP_MIN, R * (int64_t)CAND_PER_REF));
}

// The jobs of a batch, longest first (order[k] = caller's index of q[k]).
void make_batch_jobs(int32_t n_jobs, const ppf_model *models, const ppf_scene *scenes,
                     const ppf_lookup_params *params, std::vector<size_t> &order, std::vector<BatchJob> &q);
// Vote every job in one persistent launch (more only for jobs whose candidate
// buffer overflowed, rerun with their final maximum preset).
// graphs (in/out, may be NULL): asked for, the vote launches run as the
// context's captured vote graphs; true on return when every one did.
int batch_votes(ppf_ctx ctx, std::vector<BatchJob> &jobs, bool exact, float *vote_ms, int *launches,
                bool *graphs = nullptr);
// Finish every job on up to n_lanes lanes; results[k] belongs to jobs[k].
int batch_finish(ppf_ctx ctx, std::vector<BatchJob> &jobs, const ppf_lookup_params *p, int n_lanes,
                 std::vector<std::unique_ptr<ppf_result_s>> &results);
// A job after voting: its model, scene, threshold and candidates.
struct FinishInput {
  ppf_model m;
  ppf_scene s;
  float thr;
  const unsigned long long *cand_code;
  const uint32_t *cand_count;
  uint32_t n_cand, max_count;
  int64_t votes;
};
// Whether finish_together can take the jobs (GPU clustering, sort keys of 64
// bits).
bool finish_together_ok(const std::vector<FinishInput> &in, const ppf_lookup_params *p);
// Threshold/order, poses, clustering and argmax of every job at once on ctx
// (a root or a lane): one launch per stage; results[k] belongs to in[k],
// identical to a plain lookup's, and the results share one allocation.
// stages (may be NULL) receives the stage times and fused = 1.
int finish_together(ppf_ctx ctx, const std::vector<FinishInput> &in, const ppf_lookup_params *p,
                    std::vector<std::unique_ptr<ppf_result_s>> &results, ppf_batch_info_t *stages);
// Whether batch_finish_fused can take the jobs (GPU clustering, sort keys of
// 64 bits).
bool batch_fused_ok(const std::vector<BatchJob> &jobs, const ppf_lookup_params *p);
// Finish every job together: one launch per stage (threshold/order, poses,
// clustering, argmax) over all jobs; results[k] belongs to jobs[k] and is
// identical to the lane path's.  stages (may be NULL) receives the stage
// times and fused = 1.
int batch_finish_fused(ppf_ctx ctx, std::vector<BatchJob> &jobs, const ppf_lookup_params *p,
                       std::vector<std::unique_ptr<ppf_result_s>> &results, ppf_batch_info_t *stages);

}  // namespace ppfi
