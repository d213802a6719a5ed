// ppf_objects.h -- definitions of the opaque handle types and of the device
// launchers shared by kernels.hip and api.hip.
#pragma once

#include <atomic>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/ppf_amd.h"
#include "ppf_internal.h"

// Vote records: bits 0..8 model point within its tile, bit 9 "evaluate with
// the full float path", bits 10..31 the angle phi_u = atan2(u.z, u.y) of the
// model pair in units of 2*pi/2^22.
constexpr uint32_t REC_M_MASK = 0x1ffu;
constexpr uint32_t REC_SLOW_BIT = 0x200u;
constexpr uint32_t REC_PHI_SHIFT = 10;
constexpr int MAX_TILE_POINTS = 512;
constexpr int DEFAULT_TILE_POINTS = 512;
constexpr int N_BINS = 31;
constexpr int VOTE_BLOCK = 1024;
// LDS of a vote block beside its accumulator tile (shared state, the waves'
// queues of undecided votes), in bytes; kernels.hip checks the bound
constexpr size_t VOTE_LDS_RESERVE = 12288;

// Finishing of a batch: together (one launch per stage over all jobs) when
// the batch allows it, or job by job on lanes.
constexpr int32_t PPF_FINISH_AUTO = 0;
constexpr int32_t PPF_FINISH_LANES = 1;

// A context is one HIP stream with its own scratch buffers on one device.
// The context a caller creates is a root; batched and served work also run
// on lanes, contexts that share the root's device and device-memory account
// (budget, bytes) but own their stream, events and scratch, so that several
// host threads can drive independent stage chains at once.
struct ppf_ctx_s {
  ppf_ctx_s *root = this;
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  std::atomic<uint64_t> bytes{0};
  int n_cu = 0;
  size_t max_lds = 0;
  ppfi::SortScratch sort;
  ppfi::DevBuf tmp_a, tmp_b, tmp_c, tmp_d, hits, queue;
  ppfi::DevBuf sel;  // kept votes of the lookup being finished (grows, never shrinks)
  // scratch of the batched finishing (finish_batch)
  ppfi::DevBuf fin_meta, fin_flags, fin_keys, fin_cl, fin_chunk, fin_order, fin_scan, fin_kd;
  ppfi::DevBuf tmp_e;  // k-d slot grouping of a plain lookup's clustering
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  hipEvent_t ev4 = nullptr, ev5 = nullptr, ev6 = nullptr;
  std::vector<ppf_ctx_s *> lanes;  // root only
  // root only: the lane of ppf_lookup_noalloc, whose scratch buffers are only
  // ever pieces of the device's static scratch array
  ppf_ctx_s *noalloc = nullptr;
  // root only, diagnostics (ppf_debug_set): distinct model buckets a
  // reference row may group before it is voted hit by hit (0 = the LDS
  // grouping table's size)
  uint32_t dbg_group_cap = 0;
  // root only: the batched form's tuned launch shape (ppf_batch_tune; 0 =
  // default): vote blocks per CU, finishing lanes
  int32_t tune_vote_per_cu = 0;
  int32_t tune_lanes = 0;
  // root only: how batches are finished (PPF_FINISH_*; tuner or ppf_debug_set)
  int32_t finish_mode = 0;
  // the batched vote stage as two instantiated graphs (prelude, vote launch)
  // and the launch signature they were captured for
  hipGraphExec_t vote_pre = nullptr, vote_run = nullptr;
  uint64_t vote_sig[17] = {};
  // root only: run the batched vote stage from captured graphs (ppf_debug_set)
  int32_t use_graphs = 1;
  // root only: progress reports of plain lookups (ppf_ctx_set_progress)
  ppf_progress_fn progress = nullptr;
  void *progress_user = nullptr;
  int64_t progress_rows = 0;

  // Account n bytes against the root's budget (false: over budget).
  bool reserve(uint64_t n) {
    ppf_ctx_s *r = root;
    uint64_t cur = r->bytes.load();
    do {
      if (r->budget && cur + n > r->budget) return false;
    } while (!r->bytes.compare_exchange_weak(cur, cur + n));
    return true;
  }
  void unreserve(uint64_t n) { root->bytes -= n; }
};

namespace ppfi {
// Create the stream, events and limits of a context on `device`.
int ctx_init(ppf_ctx_s *c, int device);
// Release everything ctx_init and the scratch buffers hold.
void ctx_release(ppf_ctx_s *c);
// The root's lanes, created on first use (at most max_lanes).
int ctx_lanes(ppf_ctx_s *root, int n, std::vector<ppf_ctx_s *> &out);
}  // namespace ppfi

// Tracked device allocation owned by a model/scene/result.
struct TBuf {
  ppf_ctx ctx = nullptr;
  void *ptr = nullptr;
  size_t bytes = 0;
  bool arena = false;  // a piece of the current scratch arena (not budgeted, never freed)
  bool view = false;   // a range of an allocation someone else owns (never freed here)
  int alloc(ppf_ctx c, size_t n);
  void release();
  // Point at n bytes of another owner's allocation.
  void set_view(void *p, size_t n) {
    release();
    ptr = p;
    bytes = n;
    view = true;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
  ~TBuf() { release(); }
  TBuf() = default;
  TBuf(const TBuf &) = delete;
  TBuf &operator=(const TBuf &) = delete;
  TBuf(TBuf &&o) noexcept : ctx(o.ctx), ptr(o.ptr), bytes(o.bytes), arena(o.arena), view(o.view) {
    o.ptr = nullptr;
    o.bytes = 0;
    o.arena = false;
    o.view = false;
  }
  TBuf &operator=(TBuf &&o) noexcept {
    if (this != &o) {
      release();
      ctx = o.ctx;
      ptr = o.ptr;
      bytes = o.bytes;
      arena = o.arena;
      view = o.view;
      o.ptr = nullptr;
      o.bytes = 0;
      o.arena = false;
      o.view = false;
    }
    return *this;
  }
};

// Structure-of-arrays cloud: x y z nx ny nz |n|, each n floats.
struct CloudDev {
  int64_t n = 0;
  TBuf soa;
  const float *x() const { return soa.as<float>(); }
  const float *y() const { return soa.as<float>() + n; }
  const float *z() const { return soa.as<float>() + 2 * n; }
  const float *nx() const { return soa.as<float>() + 3 * n; }
  const float *ny() const { return soa.as<float>() + 4 * n; }
  const float *nz() const { return soa.as<float>() + 5 * n; }
  const float *nn() const { return soa.as<float>() + 6 * n; }
};

struct ppf_model_s {
  ppf_ctx ctx = nullptr;
  CloudDev cloud;
  float d_dist = 0;
  int32_t TM = 0, T = 0;
  int64_t U = 0;
  uint32_t hcap = 0, hshift = 0;
  TBuf frames;   // [n][16] local frames T_{m->g}
  TBuf hkeys;    // [U] unique keys (ascending)
  TBuf hcount;   // [U]
  TBuf hfirst;   // [U]
  TBuf pair;     // [n*n] pair index m_r*n + m_i, sorted by (key, pair): hashkeyToDataMap
  int32_t acc_pitch = 0;  // accumulator row pitch in words: 2^k >= TM, >= 64
  TBuf vpair;    // [n*n] pair index in vote order (key, tile, rank in class, class)
  TBuf rec;      // [n*n] packed vote records, vote order
  TBuf toff;     // [U][T+1] start of each tile's sub-range
  TBuf htab;     // [hcap][2] (key, u) open addressing, key 0 = empty
  TBuf weights;  // [n]
};

struct ppf_scene_s {
  ppf_ctx ctx = nullptr;
  CloudDev cloud;
  uint32_t rpdf = 1;
  int64_t R = 0;
  TBuf frames;  // [R][16] frames of reference points
};

struct ppf_result_s {
  ppf_ctx ctx = nullptr;
  ppf_result_info_t info{};
  int64_t P = 0;
  int64_t n_m = 0;
  TBuf codes, counts, T, vc, trans, quats;
  // results of a batch finished together: the arrays above are views of the
  // batch's allocation, which lives until its last result is destroyed
  std::shared_ptr<void> store;
  float pose[16] = {0};
  bool cpu = false;
  // CPU clustering: the (up to) three best clusters' averaged poses and votes
  int32_t n_cpu = 0;
  float cpu_poses[3][16] = {};
  uint32_t cpu_votes[3] = {};
};

// ------------------------------------------------------------- launchers
namespace ppfk {

struct VoteJob {
  const float *sx, *sy, *sz, *snx, *sny, *snz, *snn;
  const float *sframe;
  int32_t n_s, rpdf, R, r_begin;  // items of this job: reference rows r_begin, r_begin+1, ...
  const uint32_t *htab;
  uint32_t hshift, hmask;
  const uint32_t *toff;
  const uint32_t *rec;
  const uint32_t *pair;
  const float *mframe;
  const float *mx, *my, *mz;
  int32_t n_m, TM, T, acc_pitch;
  float d_dist, thr;
  uint32_t *max_count;
  unsigned long long *cand_code;
  uint32_t *cand_count;
  uint32_t *cand_n;
  uint32_t cand_cap;
  int32_t per_ref;  // 1: emit only each reference row's first maximum
  unsigned long long *votes;  // total votes of the job (statistics)
  uint64_t item_begin;
  uint32_t *ungrouped;  // rows voted hit by hit (grouping table overflow); may be NULL
  uint32_t group_cap;   // distinct buckets a row may group before it is voted hit by hit; 0 = table size
};

int prep_cloud(ppf_ctx ctx, const float *d_aos_pts, const float *d_aos_nrm, int64_t n,
               float *soa, float *frames, uint32_t rpdf);
int cloud_pair_words(ppf_ctx ctx, const CloudDev &c, float d_dist, uint32_t rpdf, uint64_t *words);
int cloud_pair_keys(ppf_ctx ctx, const CloudDev &c, float d_dist, uint32_t rpdf, uint32_t *keys);
// sorted words -> number of unique keys U (host value); then unique keys
// (hkeys), first indices (hfirst), counts (hcount) and the low halves of the
// words (low, may be NULL).  unique_emit must follow unique_count on the same
// context with the same words.  The key of a word is word >> shift: hkeys
// gets its low 32 bits, hjob (may be NULL) the bits above, low the bits below.
int unique_count(ppf_ctx ctx, const uint64_t *words, int64_t n, int64_t *U_out, int shift = 32);
int unique_emit(ppf_ctx ctx, const uint64_t *words, int64_t n, int64_t U, uint32_t *hkeys,
                uint32_t *hfirst, uint32_t *hcount, uint32_t *low, int shift = 32,
                uint32_t *hjob = nullptr);
int build_htable(ppf_ctx ctx, const uint32_t *hkeys, int64_t U, uint32_t *htab, uint32_t hcap,
                 uint32_t hshift);
int build_tile_offsets(ppf_ctx ctx, const uint32_t *hfirst, const uint32_t *hcount,
                       const uint32_t *pair, int64_t U, int32_t n_m, int32_t TM, int32_t T,
                       uint32_t *toff);
// words (key << 32 | pair, sorted by key) -> vote order (key, tile, phi_u);
// vals receives the pair index of each entry in that order.
// Vote order of the model pairs (vals receives the pair indices, toff the
// [U][T+1] start of every bucket's tile sub-range); words are sorted by key.
int vote_order(ppf_ctx ctx, uint64_t *words, uint32_t *vals, int64_t N, const CloudDev &m,
               int32_t TM, int32_t T, const uint32_t *hfirst, const uint32_t *hcount, int64_t U,
               uint32_t *toff);
int build_records(ppf_ctx ctx, const uint32_t *pair, int64_t N, const float *mframe,
                  const CloudDev &m, int32_t TM, uint32_t *rec);
// max_blocks > 0 caps the persistent grid (each block holds max_ns hits of
// scratch); results do not depend on the grid size.  graph (in/out): asked
// for, the launches run as the context's captured vote graphs (*graph = true
// when they did).
int run_votes(ppf_ctx ctx, const VoteJob *h_jobs, int n_jobs, uint64_t n_items, int max_tm,
              int64_t max_ns, bool exact, float *vote_ms, int max_blocks = 0, bool *graph = nullptr);
// Destroy the context's captured vote graphs.
void release_vote_graphs(ppf_ctx ctx);
// Device scratch of one vote block for scenes of up to max_ns points.
size_t vote_hit_bytes_per_block(int64_t max_ns);
// Blocks of a vote launch that run at once (reference rows voted in parallel).
int vote_slots(ppf_ctx ctx, int max_tm, int *slots);
int select_and_sort(ppf_ctx ctx, const unsigned long long *cand_code, const uint32_t *cand_count,
                    uint32_t n_cand, uint32_t max_count, float thr, int64_t n_s, int64_t n_m,
                    unsigned long long *out_codes, uint32_t *out_counts, int64_t *P_out);
int compute_poses(ppf_ctx ctx, const unsigned long long *codes, int64_t P, const float *mframe,
                  const float *sframe, uint32_t rpdf, float *T);
int cluster_spatial(ppf_ctx ctx, const float *T, const unsigned long long *codes,
                    const uint32_t *counts, const float *weights, int64_t P, float d_dist,
                    bool use_l1, bool use_avg, float *vc, float *trans_out, float *quats,
                    int64_t *max_idx, int64_t *pair_tests, int64_t *mixed_chunks, double *kernel_ms);
// One job of a batch for the finishing stages: its candidates (code, count)
// and threshold, and what its poses and clustering read.
struct FinJob {
  const unsigned long long *cand_code;
  const uint32_t *cand_count;
  float thr_f;  // fl(vote_count_threshold * fl(max_count))
  uint32_t rpdf;
  const float *mframe, *sframe, *weights;
  float d_dist;
};
// Results of the batched finishing: job j owns the kept votes
// [kbeg[j], kbeg[j+1]) of the output arrays (codes, counts, T, vc, trans,
// quats), in the order a plain lookup of the job gives them.
struct FinishOut {
  std::vector<uint64_t> kbeg;
  std::vector<float> pose;        // [nj][16] final pose (translation from the clustering)
  std::vector<int64_t> max_idx;   // first maximum of vc within the job (-1: no kept vote)
  std::vector<int64_t> tests, mixed;
  float select_ms = 0, pose_ms = 0, cluster_ms = 0, cluster_kernel_ms = 0;
};
// Whether finish_batch can take these jobs: every sort key fits 64 bits.
bool finish_batch_fits(int nj, uint64_t total_cand, uint32_t max_count, int64_t max_ns, int64_t max_nm);
// Threshold/order, poses and GPU spatial-hash clustering of all jobs, one
// launch per stage.  n_cand[j] candidates of job j; max_count = the largest
// of the jobs' maxima.  alloc(P_total) is called once the kept votes are
// counted and returns the output arrays (codes u64, counts u32, T 16 f32,
// vc f32, trans 3 f32, quats 4 f32, each P_total long).
struct FinishArrays {
  unsigned long long *codes;
  uint32_t *counts;
  float *T, *vc, *trans, *quats;
};
int finish_batch(ppf_ctx ctx, const std::vector<FinJob> &jobs, const std::vector<uint32_t> &n_cand,
                 uint32_t max_count, int64_t max_ns, int64_t max_nm, bool use_l1, bool use_avg,
                 const std::function<int(int64_t, FinishArrays &)> &alloc, FinishOut &out);
int debug_numerics(ppf_ctx ctx, int op, const float *a, const float *b, float *out, int64_t n);
int debug_copy(ppf_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n);  // n multiple of 4
// blocks x 256 threads of independent FMA chains, iters x 16 FMAs per thread
int debug_fma_stream(ppf_ctx ctx, float *out, int64_t blocks, int iters);

}  // namespace ppfk

namespace ppfi {
// Candidate-buffer sizing shared by the plain, batched and served paths:
// (code, count) pairs per reference point before growth on overflow.  An
// overflow costs a second vote pass.  First-pass candidates per reference
// point depend on the model's bucket sizes: 1,205 for a 109,414-point scene
// against a 19,073-point model, 2,209 for 75,366 against 11,048 points (both
// one pass at 4,096), 9,394 for 45,069 against 16,863 points (two passes).
constexpr uint32_t CAND_PER_REF = 4096;
constexpr uint32_t CAND_CAP_MIN = 1u << 23;  // plain path: at least 8M candidates
constexpr uint32_t CAND_CAP_MAX = 1u << 28;
constexpr size_t CAND_HEADER = 256;  // counters ahead of the (code, count) arrays
// Candidate capacities are multiples of 256 (keeps every array aligned).
inline uint32_t round_cap(uint64_t c) {
  const uint64_t r = (c + 255u) / 256u * 256u;
  return (uint32_t)(r < CAND_CAP_MAX ? r : CAND_CAP_MAX);
}
// Vote job of one (model, scene) pair over all its reference points.  meta:
// [0] candidates emitted, [1] running maximum, [2..3] vote total (u64).
ppfk::VoteJob vote_job(ppf_model m, ppf_scene s, float thr, unsigned long long *cand_code,
                       uint32_t *cand_count, uint32_t *meta, uint32_t cap);
// Upload a scene through `lane` (its stream and scratch); owned by lane->root.
int scene_create(ppf_ctx lane, const float *points, const float *normals, int64_t n, uint32_t rpdf,
                 ppf_scene *out);
// Stages after voting (threshold/order, poses, clustering, final pose).
int finish_lookup(ppf_ctx ctx, ppf_model m, ppf_scene s, const ppf_lookup_params *p,
                  const unsigned long long *ccode, const uint32_t *ccount, uint32_t n_cand,
                  uint32_t max_count, ppf_result_s *r);
}  // namespace ppfi

