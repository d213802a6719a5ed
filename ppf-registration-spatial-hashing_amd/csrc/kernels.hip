// kernels.hip -- gfx950 kernels of the PPF registration path and their launchers.
//
//   prep_cloud          AoS -> SoA, |n|, local frames (Scene/Model ctors,
//                       src/cuda/scene.cu:24-70, src/cuda/model.cu:43-66)
//   cloud_pair_words    quantized PPF + FNV key of every ordered pair, packed with
//                       the pair index (ppf_kernel + ppf_hash_kernel,
//                       src/cuda/kernel.cu:404-477)
//   unique_words        unique keys / first index / counts of the sorted words
//                       (histogram + exclusive_scan, include/impl/util.hpp:30-52,
//                       include/impl/parallel_hash_array.hpp:67-76)
//   build_htable        key -> unique-key index table (replaces lower_bound,
//                       parallel_hash_array.hpp:80-92; same hit/miss set)
//   build_tile_offsets  per unique key, start of each model-point tile
//   build_records       packed per-pair vote records (model point, phi_u)
//   run_votes           persistent voting kernel (ppf_vote_count_kernel +
//                       ppf_vote_kernel + the vote sort/histogram of
//                       src/cuda/model.cu:95-151, kernel.cu:480-554)
//   select_and_sort     threshold + order of kept votes (model.cu:155-170)
//   compute_poses       trans_calc_kernel2 (kernel.cu:605-645)
//   cluster_spatial     mat2transquat, trans2idx, rot_clustering, max_element
//                       (kernel.cu:647-763, model.cu:202-244,293-295)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>

#include "ppf_cluster.h"
#include "ppf_geom.h"
#include "ppf_objects.h"

using namespace ppfi;

namespace ppfk {

namespace {

constexpr int B256 = 256;

// Load through a global-address-space pointer.  Pointers read from the job
// table are generic (flat) to the compiler; flat loads also wait on the LDS
// counter, which serialises the vote loop.
template <class T>
__device__ __forceinline__ T gld(const T *p) {
  return *(const __attribute__((address_space(1))) T *)p;
}
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 gld2(const uint32_t *p) {
  const u32x2 v = *(const __attribute__((address_space(1))) u32x2 *)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 gld4(const uint4 *p) {
  const u32x4 v = *(const __attribute__((address_space(1))) u32x4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ uint32_t htab_find(const uint32_t *__restrict__ htab, uint32_t hshift,
                                              uint32_t hmask, uint32_t key) {
  uint32_t slot = (key * 2654435761u) >> hshift;
  for (;;) {
    const uint2 e = gld2(htab + 2 * slot);
    if (e.x == key) return e.y;
    if (e.x == 0u) return 0xffffffffu;
    slot = (slot + 1u) & hmask;
  }
}

// ------------------------------------------------------------ point prep
__global__ void k_prep_cloud(const float *__restrict__ pts, const float *__restrict__ nrm, int64_t n,
                             float *__restrict__ soa, float *__restrict__ frames, uint32_t rpdf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
  const float nx = nrm[3 * i], ny = nrm[3 * i + 1], nz = nrm[3 * i + 2];
  soa[i] = px;
  soa[n + i] = py;
  soa[2 * n + i] = pz;
  soa[3 * n + i] = nx;
  soa[4 * n + i] = ny;
  soa[5 * n + i] = nz;
  soa[6 * n + i] = ppfg::vnorm(nx, ny, nz);
  if (frames && (i % rpdf) == 0) {
    float T[16];
    ppfg::frame_to_global(px, py, pz, nx, ny, nz, T);
    float *o = frames + (i / rpdf) * 16;
#pragma unroll
    for (int k = 0; k < 16; ++k) o[k] = T[k];
  }
}

// --------------------------------------------------------- pair keys
// blockIdx.y = reference point r, x-dimension over partners i.  Rows with
// r % rpdf != 0 and the diagonal get key 0 (reference: .x = NaN,
// kernel.cu:432-441).
template <bool WORDS>
__global__ __launch_bounds__(B256) void k_pair_keys(const float *__restrict__ soa, int64_t n,
                                                    float d_dist, uint32_t rpdf,
                                                    void *__restrict__ out) {
  const int64_t r = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * B256 + threadIdx.x;
  if (i >= n) return;
  const float *X = soa, *Y = soa + n, *Z = soa + 2 * n, *NX = soa + 3 * n, *NY = soa + 4 * n,
              *NZ = soa + 5 * n, *NN = soa + 6 * n;
  uint32_t key = 0;
  if ((r % rpdf) == 0 && i != r) {
    float f[4];
    ppfg::ppf_feature_q(X[r], Y[r], Z[r], NX[r], NY[r], NZ[r], NN[r], X[i], Y[i], Z[i], NX[i], NY[i],
                        NZ[i], NN[i], d_dist, f);
    key = ppfm::ppf_key(f[0], f[1], f[2], f[3]);
  }
  const int64_t g = r * n + i;
  if (WORDS)
    static_cast<uint64_t *>(out)[g] = ((uint64_t)key << 32) | (uint64_t)(uint32_t)g;
  else
    static_cast<uint32_t *>(out)[g] = key;
}

// ------------------------------------------------------------ unique
constexpr int UQ_MAX_BLOCKS = 1024;

// A run of equal keys starts at j; the key of a word is w >> shift (32: the
// high half; the batched clustering's words carry a (job, cell) key above a
// pose index of fewer bits).
__device__ __forceinline__ bool is_head(const uint64_t *__restrict__ w, int64_t j, int shift) {
  return j == 0 || (w[j] >> shift) != (w[j - 1] >> shift);
}

__global__ __launch_bounds__(B256) void k_unique_count(const uint64_t *__restrict__ w, int64_t n,
                                                       int64_t per, int shift,
                                                       uint32_t *__restrict__ counts) {
  __shared__ uint32_t s[B256 / WAVE];
  const int64_t beg = (int64_t)blockIdx.x * per, end = min(beg + per, n);
  uint32_t c = 0;
  for (int64_t j = beg + threadIdx.x; j < end; j += B256) c += is_head(w, j, shift) ? 1u : 0u;
  c = wave_sum_u32(c);
  if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int k = 0; k < B256 / WAVE; ++k) t += s[k];
    counts[blockIdx.x] = t;
  }
}

// hkeys[u]: the low 32 bits of run u's key; hjob[u] (may be NULL): the bits
// above them; low[j]: the bits of word j below the key.
__global__ __launch_bounds__(B256) void k_unique_emit(const uint64_t *__restrict__ w, int64_t n,
                                                      int64_t per, int shift,
                                                      const uint32_t *__restrict__ offsets,
                                                      uint32_t *__restrict__ hkeys,
                                                      uint32_t *__restrict__ hfirst,
                                                      uint32_t *__restrict__ low,
                                                      uint32_t *__restrict__ hjob) {
  __shared__ uint32_t s_w[B256 / WAVE];
  __shared__ uint32_t s_run;
  const int tid = threadIdx.x, wv = tid / WAVE, l = tid % WAVE;
  const int64_t beg = (int64_t)blockIdx.x * per, end = min(beg + per, n);
  if (tid == 0) s_run = offsets[blockIdx.x];
  __syncthreads();
  const uint64_t lt = lanemask_lt();
  for (int64_t c0 = beg; c0 < end; c0 += B256) {
    const int64_t j = c0 + tid;
    const bool valid = j < end;
    uint64_t wj = valid ? w[j] : 0ull;
    const bool h = valid && is_head(w, j, shift);
    if (valid && low) low[j] = (uint32_t)(wj & ((1ull << shift) - 1ull));
    const uint64_t m = __ballot(h);
    if (l == 0) s_w[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = s_run;
    for (int k = 0; k < wv; ++k) base += s_w[k];
    if (h) {
      const uint32_t u = base + (uint32_t)__popcll(m & lt);
      hkeys[u] = (uint32_t)(wj >> shift);
      hfirst[u] = (uint32_t)j;
      if (hjob) hjob[u] = (uint32_t)(wj >> shift >> 32);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t t = 0;
      for (int k = 0; k < B256 / WAVE; ++k) t += s_w[k];
      s_run += t;
    }
    __syncthreads();
  }
}

__global__ void k_unique_counts(const uint32_t *__restrict__ hfirst, int64_t U, int64_t n,
                                uint32_t *__restrict__ hcount) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  const uint64_t nxt = (u + 1 < U) ? hfirst[u + 1] : (uint64_t)n;
  hcount[u] = (uint32_t)(nxt - hfirst[u]);
}

// ---------------------------------------------------------- hash table
__global__ void k_htable_insert(const uint32_t *__restrict__ hkeys, int64_t U, uint32_t *htab,
                                uint32_t hshift, uint32_t hmask) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  const uint32_t key = hkeys[u];
  if (key == 0u) return;  // key 0 never matches (kernel.cu:491)
  uint32_t slot = (key * 2654435761u) >> hshift;
  for (;;) {
    const uint32_t prev = atomicCAS(htab + 2 * slot, 0u, key);
    if (prev == 0u) {
      htab[2 * slot + 1] = (uint32_t)u;
      return;
    }
    slot = (slot + 1u) & hmask;
  }
}

// -------------------------------------------------------- tile offsets
__global__ void k_tile_offsets(const uint32_t *__restrict__ hfirst, const uint32_t *__restrict__ hcount,
                               const uint32_t *__restrict__ pair, int64_t U, int32_t n_m, int32_t TM,
                               int32_t T, uint32_t *__restrict__ toff) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  const uint32_t f = hfirst[u], c = hcount[u];
  uint32_t *o = toff + u * (T + 1);
  o[0] = f;
  for (int t = 1; t < T; ++t) {
    const uint64_t bound = (uint64_t)t * TM * (uint64_t)n_m;
    uint32_t lo = 0, hi = c;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((uint64_t)pair[f + mid] < bound)
        lo = mid + 1;
      else
        hi = mid;
    }
    o[t] = f + lo;
  }
  o[T] = f + c;
}

// -------------------------------------------------------------- records
__global__ void k_records(const uint32_t *__restrict__ pair, int64_t N, const float *__restrict__ mframe,
                          const float *__restrict__ soa, int32_t n_m, int32_t TM,
                          uint32_t *__restrict__ rec) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint32_t p = pair[j];
  const uint32_t m_r = p / (uint32_t)n_m, m_i = p - m_r * (uint32_t)n_m;
  const float *F = mframe + (int64_t)m_r * 16;
  float uy, uz;
  ppfg::frame_apply_yz(F + 4, F + 8, soa[m_i], soa[n_m + m_i], soa[2 * n_m + m_i], uy, uz);
  const double r2 = (double)uy * uy + (double)uz * uz;
  uint32_t slow = 0u, q = 0u;
  if (!(r2 > 1e-30) || !(r2 < 1e30)) {
    slow = REC_SLOW_BIT;
  } else {
    const double phi = atan2((double)uz, (double)uy);  // (-pi, pi]
    const double turns = phi * (4194304.0 / 6.283185307179586);
    q = (uint32_t)(int32_t)llrint(turns) & 0x3fffffu;
  }
  rec[j] = (q << REC_PHI_SHIFT) | slow | (m_r % (uint32_t)TM);
}

// Vote-record order.  Records are sorted by (bucket, tile, m_r mod 64),
// stably (pair order within), then within each (bucket, tile) segment by
// (rank in its class, class): the first rows of 64 hold one record of every
// class, and a class only repeats in a row once fewer than 64 classes have
// records left.  With the accumulator at acc[bin * pitch + m], pitch a
// multiple of 64, an update's LDS bank is m mod 64 whatever its bin, so lanes
// with distinct classes never share a bank (the bins of a wave's lanes differ
// from record to record).  Any order gives the same votes; this one keeps the
// 64 atomic adds of a wave instruction in distinct banks as far as possible.
// The sort words carry the bucket index u (keys ascend with u) in as few bits
// as U, T and the largest rank need, so the radix sorts take few passes.
__device__ __forceinline__ uint32_t bucket_of(const uint32_t *__restrict__ hfirst, int64_t U, int64_t j) {
  int64_t lo = 0, hi = U - 1;  // last u with hfirst[u] <= j
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if ((int64_t)hfirst[mid] <= j)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (uint32_t)lo;
}

// words in (key, pair) order -> (u << tb+6 | tile << 6 | class), vals = pair
__global__ void k_vote_order_words(uint64_t *__restrict__ words, uint32_t *__restrict__ vals,
                                   int64_t N, int32_t n_m, int32_t TM, const uint32_t *__restrict__ hfirst,
                                   int64_t U, int tb) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint32_t p = (uint32_t)words[j];
  const uint32_t m_r = p / (uint32_t)n_m;
  const uint32_t tile = m_r / (uint32_t)TM;
  const uint32_t u = bucket_of(hfirst, U, j);
  words[j] = ((uint64_t)u << (tb + 6)) | ((uint64_t)tile << 6) | (uint64_t)(m_r & 63u);
  vals[j] = p;
}

// words sorted by (u, tile, class) -> rank of each record within its
// (u, tile, class) run
__global__ void k_vote_ranks(const uint64_t *__restrict__ words, const uint32_t *__restrict__ toff,
                             int32_t T, int tb, int64_t N, uint32_t *__restrict__ rank) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint32_t tcm = (1u << (tb + 6)) - 1u;
  const uint64_t w = words[j];
  const uint32_t tc = (uint32_t)w & tcm;  // tile << 6 | class
  const uint64_t u = w >> (tb + 6);
  // first record of this (tile, class) in the bucket's tile segment
  uint32_t a = toff[u * (uint64_t)(T + 1) + (tc >> 6)], b = (uint32_t)j;
  while (a < b) {
    const uint32_t mid = (a + b) >> 1;
    if (((uint32_t)words[mid] & tcm) < tc)
      a = mid + 1;
    else
      b = mid;
  }
  rank[j] = (uint32_t)j - a;
}

// the last record of each (u, tile, class) run writes the run's length
__global__ void k_vote_run_counts(const uint64_t *__restrict__ words, const uint32_t *__restrict__ rank,
                                  int64_t N, int32_t T, int tb, uint32_t *__restrict__ cnt) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint64_t w = words[j];
  if (j + 1 < N && words[j + 1] == w) return;
  const uint64_t u = w >> (tb + 6);
  const uint64_t tile = (w >> 6) & ((1u << tb) - 1u);
  cnt[((u * (uint64_t)T + tile) << 6) + (w & 63u)] = rank[j] + 1u;
}

// Within a (u, tile) segment with n_c records of class c, the record of rank
// r and class c goes after every record of lower rank (sum over classes of
// min(n_c', r)) and after the rank-r records of lower classes.
__global__ void k_vote_place(const uint64_t *__restrict__ words, const uint32_t *__restrict__ rank,
                             const uint32_t *__restrict__ vals_in, const uint32_t *__restrict__ toff,
                             const uint32_t *__restrict__ cnt, int64_t N, int32_t T, int tb,
                             uint32_t *__restrict__ vals_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const uint64_t w = words[j];
  const uint64_t u = w >> (tb + 6);
  const uint64_t tile = (w >> 6) & ((1u << tb) - 1u);
  const uint32_t c = (uint32_t)(w & 63u), r = rank[j];
  const uint4 *n4 = reinterpret_cast<const uint4 *>(cnt + ((u * (uint64_t)T + tile) << 6));
  uint32_t pos = toff[u * (uint64_t)(T + 1) + tile];
#pragma unroll 4
  for (int k = 0; k < 16; ++k) {
    const uint4 n = n4[k];
    const uint32_t nk[4] = {n.x, n.y, n.z, n.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) pos += min(nk[e], r) + ((uint32_t)(4 * k + e) < c && nk[e] > r ? 1u : 0u);
  }
  vals_out[pos] = vals_in[j];
}

// --------------------------------------------------------------- voting
// Fast alpha bin: with theta = phi_v - phi_u the exact angle between the y/z
// parts of u and v, the float pipeline of alpha_bin() yields
// floor((theta + pi_f)/D) except within 1.13e-6 rad of a bin edge (rounding
// of cross/dot 1.2e-7, atan2f error <= 3.2 ulp = 7.7e-7, rounding of
// alpha + pi 2.4e-7).  The fast path works in 24-bit turns: a = (v24 - phi24)
// mod 2^24 with v24 = phi_v + pi and phi24 = phi_u, then
// a * 30 * 2^8 = (bin << 32) | frac exactly (the 24 x 24 bit multiplies
// v_mul_u32_u24 / v_mul_hi_u32_u24).  phi_u carries <= 7.5e-7 rad of
// quantization (22 bits, rounded to nearest from a double atan2), v24 the
// rounding to 24 bits of a double atan2 (<= 1.9e-7 rad), and nominal edges
// k*2pi/30 differ from the float pipeline's k*D - pi_f by < 2.6e-7 rad:
// < 2.4e-6 rad = 1.13e-5 bin in all.  Whenever frac is farther than
// FAST_MARGIN_BINS (1.5e-5 bin = 3.1e-6 rad) from an edge the bin is decided;
// otherwise (and for degenerate u or v) the exact float pipeline runs.  The
// wrap point theta = +-pi (position 0 or 30) is always within the margin, so
// bin 30 only comes from the exact path.
constexpr float FAST_MARGIN_BINS = 1.5e-5f;
// frac (as a 32-bit fraction) is decided when it lies in [OFF, 2^32 - OFF]
// with OFF = floor(1.5e-5 * 2^32) + 1, i.e. when frac - OFF < RANGE (mod 2^32)
constexpr uint32_t FAST_OFF = 64425u;
constexpr uint32_t FAST_RANGE = (uint32_t)(4294967296ull - 2ull * (FAST_OFF - 1u) - 1ull);
static_assert(FAST_OFF - 1u == (uint32_t)(1.5e-5 * 4294967296.0), "fast-bin margin");
constexpr int VOTE_UNROLL = 4;
// Undecided votes of a wave wait in an LDS queue (record index, hit, the bin
// the fast loop used) and take the float pipeline 64 at a time.
constexpr uint32_t VQ_CAP = WAVE;
constexpr int GT_LOG2 = 12;
constexpr int GT_SIZE = 1 << GT_LOG2;  // hit-grouping table slots (2 x 16 KB of LDS)

// Per-block scalars of the vote kernel, kept in dynamic LDS after the
// accumulator region.
struct VoteShared {
  unsigned long long item;
  int job;
  uint32_t nh, next, ovf, ngroups;
  uint32_t ndist, pad0;
  uint32_t wmax[VOTE_BLOCK / WAVE];
  uint32_t wsum[VOTE_BLOCK / WAVE][2];
  float thr;
  uint32_t wbest[VOTE_BLOCK / WAVE];  // per-wave best (count, code) of a tile
  unsigned long long wcode[VOTE_BLOCK / WAVE];
  uint32_t best;                       // the reference row's best so far
  unsigned long long best_code;
  uint32_t cls[64];  // groups per work class, then the class cursors
};
static_assert(sizeof(VoteShared) % 8 == 0, "the vote queues follow VoteShared");
static_assert(sizeof(VoteShared) + (VOTE_BLOCK / WAVE) * VQ_CAP * 8 + 8 <= VOTE_LDS_RESERVE,
              "vote block LDS beside the accumulator");

// (count, code) order of the per-reference-row maximum: larger count, then
// smaller code (the first maximum in code order)
__device__ __forceinline__ bool better_vote(uint32_t c, unsigned long long k, uint32_t bc,
                                            unsigned long long bk) {
  return c > bc || (c == bc && c > 0u && k < bk);
}

__device__ __forceinline__ int job_of_item(const VoteJob *__restrict__ jobs, int n_jobs, uint64_t it) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].item_begin <= it)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Estimated cost of every item (reference row): the bucket sizes of the keys
// of every 16th partner, one wave per item.  Writes (~cost << 32 | item), so
// that an ascending sort orders the items heaviest first: persistent blocks
// then finish on the cheap rows instead of waiting for a late heavy one.
constexpr int COST_STRIDE = 16;
__global__ __launch_bounds__(256) void k_row_cost(const VoteJob *__restrict__ jobs, int n_jobs,
                                                  uint64_t n_items, uint64_t *__restrict__ keys) {
  const uint64_t item = (uint64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  const int lane = threadIdx.x % WAVE;
  if (item >= n_items) return;
  const VoteJob &J = jobs[job_of_item(jobs, n_jobs, item)];
  const uint32_t r = (uint32_t)(item - J.item_begin) + (uint32_t)J.r_begin;
  const int32_t s_r = (int32_t)r * J.rpdf, n_s = J.n_s, T = J.T;
  const float srx = gld(J.sx + s_r), sry = gld(J.sy + s_r), srz = gld(J.sz + s_r);
  const float snx = gld(J.snx + s_r), sny = gld(J.sny + s_r), snz = gld(J.snz + s_r);
  const float snn = gld(J.snn + s_r);
  uint64_t c = 0;
  for (int32_t s_i = lane * COST_STRIDE; s_i < n_s; s_i += WAVE * COST_STRIDE) {
    if (s_i == s_r) continue;
    float f[4];
    ppfg::ppf_feature_q(srx, sry, srz, snx, sny, snz, snn, gld(J.sx + s_i), gld(J.sy + s_i), gld(J.sz + s_i),
                        gld(J.snx + s_i), gld(J.sny + s_i), gld(J.snz + s_i), gld(J.snn + s_i), J.d_dist, f);
    const uint32_t key = ppfm::ppf_key(f[0], f[1], f[2], f[3]);
    if (key == 0u) continue;
    const uint32_t u = htab_find(J.htab, J.hshift, J.hmask, key);
    if (u == 0xffffffffu) continue;
    c += gld(J.toff + (size_t)u * (T + 1) + T) - gld(J.toff + (size_t)u * (T + 1));
  }
  c = wave_sum_u64(c);
  if (lane == 0) keys[item] = ((uint64_t)(0xffffffffu - (uint32_t)min(c, (uint64_t)0xffffffffu)) << 32) | item;
}

// The float pipeline for one vote: record j (in vote order) against a hit's
// v = (vy, vz).  moved_approx = 0x100 | bin when the fast loop already added
// the vote to `bin`: then it is moved only if the exact bin differs.  Out of
// line: it runs for a few votes in 10^5, and inlined its registers would
// count against the fast loop's.
__device__ __noinline__ void exact_vote_fn(const VoteJob *__restrict__ J, uint32_t *acc, uint32_t pitch,
                                           uint32_t tile_base, uint32_t j, float vy, float vz,
                                           uint32_t moved_approx) {
  const uint32_t n_m = (uint32_t)J->n_m;
  const uint32_t pr = gld(J->pair + j);
  const uint32_t m_r = pr / n_m, m_i = pr - m_r * n_m;
  const float *Fm = J->mframe + (size_t)m_r * 16;
  float r1[4], r2[4];
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    r1[kk] = gld(Fm + 4 + kk);
    r2[kk] = gld(Fm + 8 + kk);
  }
  float uy, uz;
  ppfg::frame_apply_yz(r1, r2, gld(J->mx + m_i), gld(J->my + m_i), gld(J->mz + m_i), uy, uz);
  const uint32_t bin = ppfg::alpha_bin(uy, uz, vy, vz);
  const uint32_t mloc = m_r - tile_base;
  const bool moved = (moved_approx & 0x100u) != 0u;
  const uint32_t approx = moved_approx & 0xffu;
  if (!moved || bin != approx) {
    if (moved) atomicSub(&acc[approx * pitch + mloc], 1u);
    atomicAdd(&acc[bin * pitch + mloc], 1u);
  }
}

template <bool EXACT>
__global__ __launch_bounds__(VOTE_BLOCK, 8) void k_vote(const VoteJob *__restrict__ jobs, int n_jobs,
                                                     uint64_t n_items, const uint64_t *__restrict__ order,
                                                     unsigned long long *__restrict__ queue,
                                                     uint4 *__restrict__ hit_scratch,
                                                     uint32_t hit_cap, uint32_t stride_u4,
                                                     uint32_t misc_words) {
  // dynamic LDS only, so that the accumulator starts at LDS address 0:
  // [accumulator tile | ...][VoteShared at misc_words]
  extern __shared__ uint32_t acc[];
  VoteShared &sh = *reinterpret_cast<VoteShared *>(acc + misc_words);
  const int tid = threadIdx.x, wv = tid / WAVE, lane = tid % WAVE;
  const uint64_t lt = lanemask_lt();
  // per-block scratch: hits, hits in bucket order, slot of each hit, groups
  uint4 *hits = hit_scratch + (size_t)blockIdx.x * stride_u4;
  uint4 *sorted = hits + hit_cap;
  uint32_t *slotid = reinterpret_cast<uint32_t *>(sorted + hit_cap);
  uint2 *groups = reinterpret_cast<uint2 *>(slotid + hit_cap + (hit_cap & 1u));

  for (;;) {
    if (tid == 0) {
      unsigned long long it = atomicAdd(queue, 1ull);
      if (it < n_items && order) it = (uint32_t)order[it];
      sh.item = it;
      if (it < n_items) sh.job = job_of_item(jobs, n_jobs, it);
      sh.nh = 0;
    }
    __syncthreads();
    const unsigned long long item = sh.item;
    if (item >= n_items) break;
    const VoteJob &J = jobs[sh.job];
    const int32_t n_s = J.n_s;
    const uint32_t r = (uint32_t)(item - J.item_begin) + (uint32_t)J.r_begin;
    const int32_t s_r = (int32_t)r * J.rpdf;
    const float *Fs = J.sframe + (size_t)r * 16;
    const float srx = gld(J.sx + s_r), sry = gld(J.sy + s_r), srz = gld(J.sz + s_r);
    const float snx = gld(J.snx + s_r), sny = gld(J.sny + s_r), snz = gld(J.snz + s_r);
    const float snn = gld(J.snn + s_r);
    float row1[4], row2[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      row1[k] = gld(Fs + 4 + k);
      row2[k] = gld(Fs + 8 + k);
    }
    const float d_dist = J.d_dist;
    const uint32_t *__restrict__ htab = J.htab;
    const uint32_t hshift = J.hshift, hmask = J.hmask;

    // ---- phase 1: key, lookup and v for every partner
    for (int32_t c0 = 0; c0 < n_s; c0 += VOTE_BLOCK) {
      const int32_t s_i = c0 + tid;
      bool hit = false;
      uint4 hr;
      if (s_i < n_s && s_i != s_r) {
        const float px = gld(J.sx + s_i), py = gld(J.sy + s_i), pz = gld(J.sz + s_i);
        float f[4];
        ppfg::ppf_feature_q(srx, sry, srz, snx, sny, snz, snn, px, py, pz, gld(J.snx + s_i),
                            gld(J.sny + s_i), gld(J.snz + s_i), gld(J.snn + s_i), d_dist, f);
        const uint32_t key = ppfm::ppf_key(f[0], f[1], f[2], f[3]);
        if (key != 0u) {
          const uint32_t u = htab_find(htab, hshift, hmask, key);
          if (u != 0xffffffffu) {
            hit = true;
            float vy, vz;
            ppfg::frame_apply_yz(row1, row2, px, py, pz, vy, vz);
            uint32_t flag = 0u, v32 = 0u;
            const float r2 = vy * vy + vz * vz;
            if (EXACT || !(r2 > 1e-30f) || !(r2 < 1e30f)) {
              flag = 0x80000000u;
            } else {
              // 24-bit turns of phi_v + pi, rounded to nearest from a double
              // atan2 of the float pair (error < 2^-25 turn)
              const double turns = atan2((double)vz, (double)vy) * (16777216.0 / 6.283185307179586);
              v32 = ((uint32_t)(long long)llrint(turns) + 0x800000u) & 0xffffffu;
            }
            // the float-path flag rides on both u and v24 (bit 31)
            hr = make_uint4(u | flag, v32 | flag, ppfm::fbits(vy), ppfm::fbits(vz));
          }
        }
      }
      const uint64_t m = __ballot(hit);
      uint32_t base = 0;
      if (lane == 0 && m) base = atomicAdd(&sh.nh, (uint32_t)__popcll(m));
      base = __shfl(base, 0, WAVE);
      if (hit) hits[base + (uint32_t)__popcll(m & lt)] = hr;
    }
    __syncthreads();
    const uint32_t nh = sh.nh;

    // ---- phase 1b: group hits by model bucket.  Partners of one reference
    // point hit few distinct buckets (tens of hits per bucket is common), so
    // phase 2 reads each bucket's records once per group.  LDS hash table over
    // u (in the accumulator space, not yet in use), then a counting sort.
    uint32_t *tkey = acc;
    uint32_t *tcnt = acc + GT_SIZE;
    for (int i = tid; i < GT_SIZE; i += VOTE_BLOCK) {
      tkey[i] = 0u;
      tcnt[i] = 0u;
    }
    if (tid < 64) sh.cls[tid] = 0u;
    if (tid == 0) {
      sh.ovf = 0u;
      sh.ndist = 0u;
    }
    __syncthreads();
    const uint32_t group_cap = J.group_cap;
    for (uint32_t h = tid; h < nh; h += VOTE_BLOCK) {
      const uint32_t u = gld(reinterpret_cast<const uint32_t *>(hits + h)) & 0x7fffffffu;
      const uint32_t key = u + 1u;
      uint32_t slot = (u * 2654435761u) >> (32 - GT_LOG2);
      for (uint32_t probes = 0;; ++probes) {
        if (probes == (uint32_t)GT_SIZE) {
          slot = 0xffffffffu;
          break;
        }
        const uint32_t prev = atomicCAS(&tkey[slot], 0u, key);
        if (prev == 0u) {
          // a new group; beyond group_cap groups the row is voted hit by hit
          if (group_cap != 0u && atomicAdd(&sh.ndist, 1u) >= group_cap) sh.ovf = 1u;
          break;
        }
        if (prev == key) break;
        slot = (slot + 1u) & (GT_SIZE - 1);
      }
      if (slot == 0xffffffffu)
        sh.ovf = 1u;
      else
        atomicAdd(&tcnt[slot], 1u);
      slotid[h] = slot;
    }
    __syncthreads();
    const bool identity = sh.ovf != 0u;  // table full: one group per hit
    if (identity && tid == 0 && J.ungrouped) atomicAdd(J.ungrouped, 1u);
    if (!identity) {
      // exclusive scans of the slot counts (group starts) and of the used
      // slots (group ids); each thread owns GT_SIZE / VOTE_BLOCK slots
      constexpr int PER = GT_SIZE / VOTE_BLOCK;
      uint32_t c[PER], cs = 0, us = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        c[q] = tcnt[tid * PER + q];
        cs += c[q];
        us += c[q] ? 1u : 0u;
      }
      const uint32_t ci = wave_incl_scan_u32(cs), ui = wave_incl_scan_u32(us);
      if (lane == WAVE - 1) {
        sh.wsum[wv][0] = ci;
        sh.wsum[wv][1] = ui;
      }
      __syncthreads();
      uint32_t cb = ci - cs, utot = 0;
      for (int k = 0; k < VOTE_BLOCK / WAVE; ++k) {
        if (k < wv) cb += sh.wsum[k][0];
        utot += sh.wsum[k][1];
      }
      // heaviest groups first, so that the waves of a tile finish together:
      // work class = floor(log2(hits x bucket records)), counted per class,
      // then placed in descending class order (the order of the groups does
      // not change any count: the accumulator adds are integer atomics)
      uint32_t cl[PER];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        cl[q] = 0u;
        if (c[q]) {
          const uint32_t u = tkey[tid * PER + q] - 1u;
          const uint32_t *tu = J.toff + (size_t)u * (J.T + 1);
          const uint64_t wk = (uint64_t)c[q] * (uint64_t)(gld(tu + J.T) - gld(tu));
          cl[q] = wk ? 63u - (uint32_t)__builtin_clzll(wk) : 0u;
          atomicAdd(&sh.cls[cl[q]], 1u);
        }
      }
      __syncthreads();
      if (tid == 0) {
        uint32_t run = 0;
        for (int k = 63; k >= 0; --k) {
          const uint32_t n = sh.cls[k];
          sh.cls[k] = run;
          run += n;
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        if (c[q]) groups[atomicAdd(&sh.cls[cl[q]], 1u)] = make_uint2(cb, c[q]);
        tcnt[tid * PER + q] = cb;  // cursor
        cb += c[q];
      }
      if (tid == 0) sh.ngroups = utot;
      __syncthreads();
      for (uint32_t h = tid; h < nh; h += VOTE_BLOCK) {
        const uint32_t pos = atomicAdd(&tcnt[slotid[h]], 1u);
        sorted[pos] = gld4(hits + h);
      }
    } else if (tid == 0) {
      sh.ngroups = nh;
    }
    __syncthreads();
    const uint32_t ngroups = sh.ngroups;
    const uint4 *ghits = identity ? hits : sorted;

    // ---- phase 2: per model-point tile, accumulate then emit
    unsigned long long wave_votes = 0;
    const int32_t T = J.T, TM = J.TM, n_m = J.n_m;
    // bin-major accumulator acc[bin * pitch + m], pitch = 2^k a multiple of
    // 64: the LDS bank of an update is m mod 64 (LDS byte addresses below)
    const uint32_t pitch = (uint32_t)J.acc_pitch, pitch4 = pitch * 4u;
    const uint32_t acc_base = (uint32_t)(uintptr_t)acc;
    const int32_t nb = N_BINS * (int32_t)pitch;
    const uint32_t *__restrict__ toff = J.toff;
    const uint32_t *__restrict__ rec = J.rec;
    uint2 *wq = reinterpret_cast<uint2 *>(&sh + 1) + (size_t)wv * VQ_CAP;
    for (int32_t t = 0; t < T; ++t) {
      const int32_t tile_base = t * TM;
      // the float pipeline for one vote: record j against a hit's v = (vy,
      // vz); a vote the fast loop added to bin `approx` (moved) is moved
      // when its exact bin differs
      auto exact_vote = [&](uint32_t j, float vy, float vz, bool moved, uint32_t approx) {
        exact_vote_fn(&J, acc, pitch, (uint32_t)tile_base, j, vy, vz, moved ? (0x100u | approx) : 0u);
      };
      for (int32_t i = tid; i < nb; i += VOTE_BLOCK) acc[i] = 0u;
      if (tid == 0) sh.next = 0;
      __syncthreads();
      for (;;) {
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(&sh.next, 1u);
        g = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(g, 0, WAVE));
        if (g >= ngroups) break;
        uint32_t gs = g, gc = 1;
        if (!identity) {
          const uint2 gr = gld2(reinterpret_cast<const uint32_t *>(groups + g));
          gs = gr.x;
          gc = gr.y;
        }
        const uint32_t u = gld(reinterpret_cast<const uint32_t *>(ghits + gs)) & 0x7fffffffu;
        const uint32_t b = gld(toff + (size_t)u * (T + 1) + t);
        const uint32_t e = gld(toff + (size_t)u * (T + 1) + t + 1);
        if (b >= e) continue;
        wave_votes += (unsigned long long)(e - b) * gc;
        for (uint32_t s0 = 0; s0 < gc; s0 += WAVE) {
          const uint32_t sc =
              (uint32_t)__builtin_amdgcn_readfirstlane((int)min((uint32_t)WAVE, gc - s0));
          uint4 my = make_uint4(0u, 0u, 0u, 0u);
          if ((uint32_t)lane < sc) my = gld4(ghits + gs + s0 + lane);
          // queued undecided votes: (record, hit k | approx bin << 6 | moved << 11)
          uint32_t qn = 0;
          auto flush = [&]() {
            // the wave's queue writes before its reads (and the reads before
            // later writes)
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            const bool act = (uint32_t)lane < qn;
            const uint2 en = act ? wq[lane] : make_uint2(0u, 0u);
            const int kk = (int)(en.y & 63u);
            const float vy = ppfm::bitsf((uint32_t)__shfl((int)my.z, kk, WAVE));
            const float vz = ppfm::bitsf((uint32_t)__shfl((int)my.w, kk, WAVE));
            if (act) exact_vote(en.x, vy, vz, ((en.y >> 11) & 1u) != 0u, (en.y >> 6) & 31u);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
          };
          for (uint32_t jb = b; jb < e; jb += VOTE_UNROLL * WAVE) {
            // per record: phi_u in 24-bit turns, byte address of its
            // accumulator column, and the fast-path range (0: never fast).
            // Slot q holds records jb + 64 q + lane; the first `left[q]`
            // lanes of it are records (a uniform count, so the lane masks
            // are scalar).  Lanes past the end re-read the last record.
            const uint32_t j0 = jb + (uint32_t)lane;
            uint32_t ph[VOTE_UNROLL], row4[VOTE_UNROLL], rg[VOTE_UNROLL], inc[VOTE_UNROLL];
            uint64_t vmask[VOTE_UNROLL];
            int32_t left[VOTE_UNROLL];
            uint32_t rr[VOTE_UNROLL];
#pragma unroll
            for (int q = 0; q < VOTE_UNROLL; ++q) rr[q] = gld(rec + min(j0 + q * WAVE, e - 1u));
#pragma unroll
            for (int q = 0; q < VOTE_UNROLL; ++q) {
              left[q] = (int32_t)(e - jb) - q * WAVE;
              vmask[q] = left[q] >= WAVE ? ~0ull : (left[q] <= 0 ? 0ull : (1ull << left[q]) - 1ull);
              const bool valid = lane < left[q];
              const uint32_t r = rr[q];
              ph[q] = (r >> (REC_PHI_SHIFT - 2)) & 0xfffffcu;  // phi_u in 24-bit turns
              const bool fastrec = valid && !(r & REC_SLOW_BIT);
              // lanes without a fast record add 0, in their own bank's column
              row4[q] = acc_base + ((fastrec ? (r & REC_M_MASK) : (uint32_t)lane) << 2);
              rg[q] = fastrec ? FAST_RANGE : 0u;
              inc[q] = fastrec ? 1u : 0u;
              asm volatile("" : "+v"(inc[q]));  // keep it a register, not a mask
            }
            for (uint32_t k = 0; k < sc; ++k) {
              // 24-bit turns of phi_v + pi; bit 31: every vote of this hit
              // takes the float path (degenerate v, or exact mode)
              const uint32_t v24 = (uint32_t)__builtin_amdgcn_readlane((int)my.y, (int)k);
              uint64_t bad = 0;
              if (EXACT || (v24 >> 31) != 0u) {
                // degenerate v (or exact mode): every vote of this hit exact
#pragma unroll
                for (int q = 0; q < VOTE_UNROLL; ++q) bad |= vmask[q];
              } else {
#pragma unroll
                for (int q = 0; q < VOTE_UNROLL; ++q) {
                  // a = 24-bit turns of theta + pi; a * 30 * 2^8 = (bin << 32) | frac.
                  // Every lane adds inc to the bin of a; the float path below
                  // moves the undecided votes whose bin differs.
                  const uint32_t a = v24 - ph[q];
                  const bool ok = (a & 0xffffffu) * (30u << 8) - FAST_OFF < rg[q];
                  bad |= ~__ballot(ok) & vmask[q];
                  uint32_t bin;
                  asm("v_mul_hi_u32_u24 %0, 0x1e00, %1" : "=v"(bin) : "v"(a));
                  const uint32_t addr = row4[q] + __umul24(bin, pitch4);
                  asm volatile("ds_add_u32 %0, %1" : : "v"(addr), "v"(inc[q]) : "memory");
                }
              }
              if (bad) {
                if (EXACT || (v24 >> 31) != 0u) {
                  // every vote of this hit takes the float pipeline, now
                  const float vy = ppfm::bitsf((uint32_t)__builtin_amdgcn_readlane((int)my.z, (int)k));
                  const float vz = ppfm::bitsf((uint32_t)__builtin_amdgcn_readlane((int)my.w, (int)k));
#pragma unroll
                  for (int q = 0; q < VOTE_UNROLL; ++q)
                    if (lane < left[q]) exact_vote(j0 + q * WAVE, vy, vz, false, 0u);
                } else {
                  // undecided votes (a few lanes of a few hits) join the
                  // queue with the bin the fast loop added them to
#pragma unroll
                  for (int q = 0; q < VOTE_UNROLL; ++q) {
                    const uint32_t a = (v24 - ph[q]) & 0xffffffu;
                    const bool need = lane < left[q] && !(a * (30u << 8) - FAST_OFF < rg[q]);
                    const uint64_t nm = __ballot(need);
                    if (nm != 0ull) {
                      const uint32_t cnt = (uint32_t)__popcll(nm);
                      if (qn + cnt > VQ_CAP) {
                        flush();
                        qn = 0;
                      }
                      if (need)
                        wq[qn + (uint32_t)__popcll(nm & lt)] =
                            make_uint2(j0 + q * WAVE, k | (((a * 30u) >> 24) << 6) | (rg[q] != 0u ? 1u << 11 : 0u));
                      qn += cnt;
                    }
                  }
                }
              }
            }
          }
          if (qn != 0u) {  // before the next chunk of hits replaces my
            flush();
            qn = 0;
          }
        }
      }
      __syncthreads();
      // local maximum -> running job maximum -> pruning threshold
      uint32_t lm = 0;
      for (int32_t i = tid; i < nb; i += VOTE_BLOCK) lm = max(lm, acc[i]);
      lm = wave_max_u32(lm);
      if (lane == 0) sh.wmax[wv] = lm;
      __syncthreads();
      if (tid == 0) {
        uint32_t m = 0;
        for (int k = 0; k < VOTE_BLOCK / WAVE; ++k) m = max(m, sh.wmax[k]);
        uint32_t g = atomicMax(J.max_count, m);
        g = max(g, m);
        sh.thr = J.thr * (float)g;
      }
      __syncthreads();
      const float thr_lb = sh.thr;
      if (J.per_ref) {
        // per-reference-point argmax: the tile's first maximum in code order,
        // kept if it beats the row's best of the earlier tiles
        uint32_t bc = 0;
        unsigned long long bk = ~0ull;
        for (int32_t i = tid; i < nb; i += VOTE_BLOCK) {
          const uint32_t c = acc[i];
          const uint32_t a = (uint32_t)i / pitch;
          const uint32_t mr = (uint32_t)tile_base + ((uint32_t)i - a * pitch);
          const unsigned long long k = ((unsigned long long)(uint32_t)s_r << 32) | (mr << 6) | a;
          if (better_vote(c, k, bc, bk)) {
            bc = c;
            bk = k;
          }
        }
        for (int o = WAVE / 2; o > 0; o >>= 1) {
          const uint32_t c2 = (uint32_t)__shfl_xor((int)bc, o, WAVE);
          const unsigned long long k2 = __shfl_xor(bk, o, WAVE);
          if (better_vote(c2, k2, bc, bk)) {
            bc = c2;
            bk = k2;
          }
        }
        if (lane == 0) {
          sh.wbest[wv] = bc;
          sh.wcode[wv] = bk;
        }
        __syncthreads();
        if (tid == 0) {
          if (t == 0) {
            sh.best = 0u;
            sh.best_code = ~0ull;
          }
          for (int k = 0; k < VOTE_BLOCK / WAVE; ++k)
            if (better_vote(sh.wbest[k], sh.wcode[k], sh.best, sh.best_code)) {
              sh.best = sh.wbest[k];
              sh.best_code = sh.wcode[k];
            }
          if (t == T - 1 && sh.best > 0u && (float)sh.best > thr_lb) {
            const uint32_t pos = atomicAdd(J.cand_n, 1u);
            if (pos < J.cand_cap) {
              J.cand_code[pos] = sh.best_code;
              J.cand_count[pos] = sh.best;
            }
          }
        }
        __syncthreads();
        continue;
      }
      const int32_t nb_r = (nb + VOTE_BLOCK - 1) / VOTE_BLOCK * VOTE_BLOCK;
      for (int32_t i = tid; i < nb_r; i += VOTE_BLOCK) {
        const uint32_t c = i < nb ? acc[i] : 0u;
        const bool emit = c > 0u && (float)c > thr_lb;
        const uint64_t m = __ballot(emit);
        uint32_t base = 0;
        if (lane == 0 && m) base = atomicAdd(J.cand_n, (uint32_t)__popcll(m));
        base = __shfl(base, 0, WAVE);
        if (emit) {
          const uint32_t pos = base + (uint32_t)__popcll(m & lt);
          if (pos < J.cand_cap) {
            const uint32_t a = (uint32_t)i / pitch;
            const uint32_t mr = (uint32_t)tile_base + ((uint32_t)i - a * pitch);
            J.cand_code[pos] = ((unsigned long long)(uint32_t)s_r << 32) | (unsigned long long)((mr << 6) | a);
            J.cand_count[pos] = c;
          }
        }
      }
      __syncthreads();
    }
    if (lane == 0 && wave_votes) atomicAdd(J.votes, wave_votes);
  }
}

// ------------------------------------------------------- vote selection
__global__ void k_select_flags(const uint32_t *__restrict__ cnt, uint32_t n, float thr_f,
                               uint32_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flags[i] = ((float)cnt[i] > thr_f) ? 1u : 0u;
}

// compact survivors into composite sort keys (count desc, code asc) when the
// fields fit in 64 bits, else into (code, count) pairs
__global__ void k_select_scatter(const unsigned long long *__restrict__ code,
                                 const uint32_t *__restrict__ cnt, uint32_t n,
                                 const uint32_t *__restrict__ flags, const uint32_t *__restrict__ pos,
                                 uint32_t max_count, int composite, int sh_m, int sh_c,
                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const unsigned long long c = code[i];
  const uint32_t p = pos[i];
  if (composite) {
    const uint64_t s_r = c >> 32, mc = c & 0xffffffffull;
    keys[p] = ((uint64_t)(max_count - cnt[i]) << sh_c) | (s_r << sh_m) | mc;
  } else {
    keys[p] = c;
    vals[p] = cnt[i];
  }
}

__global__ void k_select_decode(const uint64_t *__restrict__ keys, int64_t P, uint32_t max_count,
                                int sh_m, int sh_c, unsigned long long *__restrict__ out_code,
                                uint32_t *__restrict__ out_cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint64_t k = keys[i];
  const uint64_t mc = k & ((1ull << sh_m) - 1ull);
  const uint64_t s_r = (k & ((1ull << sh_c) - 1ull)) >> sh_m;
  out_code[i] = (s_r << 32) | mc;
  out_cnt[i] = max_count - (uint32_t)(k >> sh_c);
}

__global__ void k_iota_inv(const uint32_t *__restrict__ cnt, int64_t P, uint32_t max_count,
                           uint64_t *__restrict__ keys2, uint32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  keys2[i] = (uint64_t)(max_count - cnt[i]);
  idx[i] = (uint32_t)i;
}

__global__ void k_gather_pairs(const uint64_t *__restrict__ codes, const uint32_t *__restrict__ cnt,
                               const uint32_t *__restrict__ idx, int64_t P,
                               unsigned long long *__restrict__ out_code,
                               uint32_t *__restrict__ out_cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  out_code[i] = codes[idx[i]];
  out_cnt[i] = cnt[idx[i]];
}

// ----------------------------------------------------------------- poses
// The pose of kept vote code c into T[0..15].
__device__ __forceinline__ void pose_of_code(unsigned long long c, const float *__restrict__ mframe,
                                             const float *__restrict__ sframe, uint32_t rpdf,
                                             float *__restrict__ T) {
  const uint32_t s_r = (uint32_t)(c >> 32);
  const uint32_t mc = (uint32_t)c;
  const uint32_t m_r = mc >> 6, a = mc & 63u;
  float out[16];
  if (s_r == 0 && m_r == 0 && a == 0) {
    // code 0 is skipped by the reference and keeps a zero matrix (kernel.cu:628-630)
#pragma unroll
    for (int k = 0; k < 16; ++k) out[k] = 0.0f;
  } else {
    float Tm[16], Ts[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      Tm[k] = mframe[(size_t)m_r * 16 + k];
      Ts[k] = sframe[(size_t)(s_r / rpdf) * 16 + k];
    }
    ppfg::pose_from_frames(Tm, Ts, a, out);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) T[k] = out[k];
}

__global__ void k_poses(const unsigned long long *__restrict__ codes, int64_t P,
                        const float *__restrict__ mframe, const float *__restrict__ sframe,
                        uint32_t rpdf, float *__restrict__ T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  pose_of_code(codes[i], mframe, sframe, rpdf, T + i * 16);
}

// ------------------------------------------------------------ clustering
// Translation, quaternion, weighted count and the 26 neighbour-cell hashes of
// pose i (mat2transquat + trans2idx, kernel.cu:647-722); returns its cell hash.
__device__ __forceinline__ uint32_t cluster_prep_one(int64_t i, const float *__restrict__ T,
                                                     const unsigned long long *__restrict__ codes,
                                                     const uint32_t *__restrict__ cnt,
                                                     const float *__restrict__ weights, float d_dist,
                                                     float *__restrict__ trans, float *__restrict__ quats,
                                                     float *__restrict__ w, uint32_t *__restrict__ adj) {
  float M[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) M[k] = T[i * 16 + k];
  trans[3 * i] = M[3];
  trans[3 * i + 1] = M[7];
  trans[3 * i + 2] = M[11];
  float q[4];
  ppfg::rotmat_to_quat(M, q);
#pragma unroll
  for (int k = 0; k < 4; ++k) quats[4 * i + k] = q[k];
  const uint32_t m_r = ((uint32_t)codes[i]) >> 6;
  w[i] = (weights ? weights[m_r] : 1.0f) * (float)cnt[i];
  int32_t c[3];
  ppfg::cell_of(M[3], M[7], M[11], d_dist, c);
  int cc = 0;
  for (int a = -1; a < 2; ++a)
    for (int b = -1; b < 2; ++b)
      for (int k = -1; k < 2; ++k, ++cc)
        adj[27 * i + cc] = (a == 0 && b == 0 && k == 0) ? 0u
                                                         : ppfm::cell_hash(c[0] + a, c[1] + b, c[2] + k);
  return ppfm::cell_hash(c[0], c[1], c[2]);
}

__global__ void k_cluster_prep(const float *__restrict__ T, const unsigned long long *__restrict__ codes,
                               const uint32_t *__restrict__ cnt, const float *__restrict__ weights,
                               int64_t P, float d_dist, float *__restrict__ trans,
                               float *__restrict__ quats, float *__restrict__ w,
                               uint64_t *__restrict__ words, uint32_t *__restrict__ adj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint32_t h = cluster_prep_one(i, T, codes, cnt, weights, d_dist, trans, quats, w, adj);
  words[i] = ((uint64_t)h << 32) | (uint64_t)i;
}

// Clustering blocks: CL_PER poses per thread, CL_CHUNK poses per block.  One
// wave per block: the waves of a larger block wait for each other at every
// member tile, and their pose slots test different numbers of members.
constexpr int CL_BLOCK = 64;
constexpr int CL_PER = 2;  // 3 and 4 measured slower
constexpr int CL_CHUNK = CL_BLOCK * CL_PER;

// Work items of the tiled clustering: (unique cell hash u, CL_CHUNK-pose chunk).
__global__ void k_cluster_chunks(const uint32_t *__restrict__ ucount, int64_t U,
                                 uint32_t *__restrict__ nchunk, uint32_t *__restrict__ max_count) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  nchunk[u] = (ucount[u] + (uint32_t)CL_CHUNK - 1u) / (uint32_t)CL_CHUNK;
  atomicMax(max_count, ucount[u]);
}

// Order of the poses of a cell for forming clustering chunks: the position
// of the normalized quaternion (8 bits per component) along a 4-D Hilbert
// curve, so that the 64 poses of a wave's slot have close rotations (their
// rotation caps are smaller than along a Morton order, whose runs jump).  The
// order of the members walked in each neighbour cell is unchanged (pose
// order); only which poses share a wave depends on this key.
__device__ __forceinline__ uint32_t quat_order_key(const float *__restrict__ quats, uint32_t i) {
  float q[4], n2 = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    q[c] = quats[4 * (size_t)i + c];
    n2 += q[c] * q[c];
  }
  const float inv = n2 > 0.0f && n2 < 1e30f ? 1.0f / sqrtf(n2) : 0.0f;
  uint32_t b[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float x = q[c] * inv;
    x = x == x ? fminf(fmaxf(x, -1.0f), 1.0f) : 0.0f;
    b[c] = (uint32_t)((x + 1.0f) * 127.5f);
  }
  return ppfc::hilbert4<8>(b[0], b[1], b[2], b[3]);
}

__global__ void k_cluster_order_keys(const float *__restrict__ quats, const uint64_t *__restrict__ words,
                                     int64_t P, uint64_t *__restrict__ okeys,
                                     uint32_t *__restrict__ ovals) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const uint64_t wd = words[k];  // sorted (cell hash, pose)
  const uint32_t i = (uint32_t)wd;
  okeys[k] = (wd & 0xffffffff00000000ull) | (uint64_t)quat_order_key(quats, i);
  ovals[k] = i;
}

// Estimated cost of every clustering chunk: its poses times the members of
// its cell's 26 neighbour buckets.  Writes (~cost << 32 | chunk), so that an
// ascending sort launches the heaviest chunks first.
// Cells of a batch of jobs: ukeys holds each job's cell hashes, ascending,
// in the range [job_ub[j], job_ub[j+1]) of job j = ujob[u]; a pose's
// neighbour cells are looked up in its own job's range.  Without a batch
// (ujob NULL) the range is the whole array.
__device__ __forceinline__ void cell_range(const uint32_t *__restrict__ ujob,
                                           const uint32_t *__restrict__ job_ub, int64_t U, int64_t u,
                                           int64_t &lo, int64_t &hi, uint32_t &job) {
  if (ujob) {
    job = ujob[u];
    lo = job_ub[job];
    hi = job_ub[job + 1];
  } else {
    job = 0;
    lo = 0;
    hi = U;
  }
}

__global__ void k_cluster_chunk_cost(const uint32_t *__restrict__ chunk_start, int64_t U,
                                     uint32_t n_chunks, const uint32_t *__restrict__ ukeys,
                                     const uint32_t *__restrict__ ufirst, const uint32_t *__restrict__ ucount,
                                     const uint32_t *__restrict__ centers, const uint32_t *__restrict__ adj,
                                     const uint32_t *__restrict__ ujob, const uint32_t *__restrict__ job_ub,
                                     uint64_t *__restrict__ keys) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_chunks) return;
  int64_t lo = 0, hi = U - 1;  // cell: last u with chunk_start[u] <= b
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (chunk_start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint32_t c = b - chunk_start[lo], n = ucount[lo];
  const uint32_t poses = min((uint32_t)CL_CHUNK, n - c * (uint32_t)CL_CHUNK);
  const uint32_t i0 = centers[ufirst[lo]];
  int64_t r0, r1;
  uint32_t job;
  cell_range(ujob, job_ub, U, lo, r0, r1, job);
  uint64_t nb = 0;
  for (int q = 0; q < 27; ++q) {
    const uint32_t h = adj[27 * (size_t)i0 + q];
    if (h == 0u) continue;
    int64_t a = r0, e = r1;
    while (a < e) {
      const int64_t mid = (a + e) >> 1;
      if (ukeys[mid] < h)
        a = mid + 1;
      else
        e = mid;
    }
    if (a < r1 && ukeys[a] == h) nb += ucount[a];
  }
  const uint64_t cost = min((uint64_t)poses * nb, (uint64_t)0xffffffffu);
  keys[b] = ((uint64_t)(0xffffffffu - (uint32_t)cost) << 32) | b;
}

// Member data in cell order: quaternion, and translation + weight.
__global__ void k_cluster_gather(const uint32_t *__restrict__ members, int64_t P,
                                 const float *__restrict__ trans, const float *__restrict__ quats,
                                 const float *__restrict__ w, float4 *__restrict__ cq,
                                 float4 *__restrict__ ct) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const uint32_t j = members[k];
  cq[k] = make_float4(quats[4 * (size_t)j], quats[4 * (size_t)j + 1], quats[4 * (size_t)j + 2],
                      quats[4 * (size_t)j + 3]);
  ct[k] = make_float4(trans[3 * (size_t)j], trans[3 * (size_t)j + 1], trans[3 * (size_t)j + 2], w[j]);
}

// Poses that share a voxel cell share the 26 neighbour hashes, so one block
// takes up to CL_CHUNK poses of one cell (two per thread, in the rotation
// order above) and streams every neighbour bucket through LDS once.  Each
// thread walks the bucket in member order, which keeps the summation order
// of rot_clustering_kernel (kernel.cu:724-756).  Chunks
// whose poses do not all lie in one cell (an FNV collision between cells)
// walk their own neighbour lists per pose.
//
// Translation test: sqrtf is correctly rounded and monotone, so
// sqrtf(s) < d_dist holds exactly when s <= s_max, the largest float whose
// square root rounds below d_dist (computed on the host).  s is the same
// ((dx*dx + dy*dy) + dz*dz) sum as norm(); no square root is taken.
//
// Wave pre-test: a wave bounds its poses by a rotation cap (centre c, angle
// rho, largest quaternion norm) and a translation ball (centre, radius).
// The rotation test can pass only when the float dot q_i.q_j exceeds
// 1 - ROT_THRESH^2/8 > 0.978, i.e. when the angle between the quaternions is
// below acos(0.978 / (|q_i||q_j|)); by the triangle inequality on the unit
// 3-sphere a member whose angle to c exceeds that plus rho fails for every
// pose of the wave.  Likewise a member farther than d_dist + radius from the
// ball's centre fails the translation test.  Both bounds carry margins far
// above float rounding (5e-3 rad, 1e-5 relative), and a member is skipped
// only when a bound proves that it fails for all poses of the wave; skipped
// members would add nothing, so every sum keeps its order and bits.

using ppfc::ClPose;
using ppfc::ClWave;

// The cell of hash h among ukeys[r0, r1) (ascending), or -1.
__device__ __forceinline__ int64_t cell_lower_bound(const uint32_t *__restrict__ ukeys, int64_t r0,
                                                    int64_t r1, uint32_t h) {
  int64_t lo = r0, hi = r1;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ukeys[mid] < h)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < r1 && ukeys[lo] == h) ? lo : -1;
}

__device__ __forceinline__ void load_pose(ClPose &P, const float *__restrict__ trans,
                                          const float *__restrict__ quats, uint32_t i) {
  P.tx = trans[3 * (size_t)i];
  P.ty = trans[3 * (size_t)i + 1];
  P.tz = trans[3 * (size_t)i + 2];
  P.qx = quats[4 * (size_t)i];
  P.qy = quats[4 * (size_t)i + 1];
  P.qz = quats[4 * (size_t)i + 2];
  P.qw = quats[4 * (size_t)i + 3];
  P.vc = 1.0f;
  P.ox = P.tx;
  P.oy = P.ty;
  P.oz = P.tz;
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
  return v;
}

__device__ __forceinline__ float readfirst_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float wave_sum_f32(float v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

__device__ __forceinline__ float wave_min_f32(float v) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, WAVE));
  return v;
}

// Bounds of the 64 poses a wave holds in one slot, around their mean
// quaternion and the midpoint of their translations (lane 0's pose where a
// mean is unusable).  Poses with non-finite quaternions never pass the
// rotation test and are left out of the mean.
__device__ ClWave cluster_slot_bounds(const ClPose &p, bool have, float d_dist) {
  ClPose f = p;
  f.tx = readfirst_f32(f.tx);
  f.ty = readfirst_f32(f.ty);
  f.tz = readfirst_f32(f.tz);
  f.qx = readfirst_f32(f.qx);
  f.qy = readfirst_f32(f.qy);
  f.qz = readfirst_f32(f.qz);
  f.qw = readfirst_f32(f.qw);
  const bool qok = have && isfinite(p.qx) && isfinite(p.qy) && isfinite(p.qz) && isfinite(p.qw);
  const float sq[4] = {wave_sum_f32(qok ? p.qx : 0.0f), wave_sum_f32(qok ? p.qy : 0.0f),
                       wave_sum_f32(qok ? p.qz : 0.0f), wave_sum_f32(qok ? p.qw : 0.0f)};
  const float nq = wave_sum_f32(qok ? 1.0f : 0.0f);
  const float lo[3] = {wave_min_f32(have ? p.tx : INFINITY), wave_min_f32(have ? p.ty : INFINITY),
                       wave_min_f32(have ? p.tz : INFINITY)};
  const float hi[3] = {wave_max_f32(have ? p.tx : -INFINITY), wave_max_f32(have ? p.ty : -INFINITY),
                       wave_max_f32(have ? p.tz : -INFINITY)};
  const ClPose c = ppfc::cl_group_centre(sq, nq, lo, hi, f);
  const bool h0 = __ballot(have) != 0ull;
  ClWave W;
  ppfc::cl_wave_center(c, h0, W);
  float rho = 0.0f, nmax = 0.0f, rad = 0.0f;
  if (have) ppfc::cl_wave_add(W, p, rho, nmax, rad);
  ppfc::cl_wave_finish(W, wave_max_f32(rho), wave_max_f32(nmax), wave_max_f32(rad), d_dist);
  return W;
}

// The flags are template parameters: compiled in as constants, the test of
// the common case (no averaging) carries none of the averaging's code.
template <int USE_L1, int USE_AVG>
__device__ __forceinline__ void cl_test(ClPose &P, const float4 q, const float4 t, float s_max) {
  ppfc::cluster_test(P, q.x, q.y, q.z, q.w, t.x, t.y, t.z, t.w, s_max, USE_L1, USE_AVG);
}

template <int use_l1, int use_avg>
__global__ __launch_bounds__(CL_BLOCK) void k_cluster_tiled(
    const float *__restrict__ trans, const float *__restrict__ quats, const float4 *__restrict__ cq,
    const float4 *__restrict__ ct, const uint32_t *__restrict__ adj,
    const uint32_t *__restrict__ ukeys, const uint32_t *__restrict__ ufirst,
    const uint32_t *__restrict__ ucount, int64_t U, const uint32_t *__restrict__ chunk_start,
    const uint32_t *__restrict__ centers, const uint64_t *__restrict__ order, float s_max, float d_dist,
    float *__restrict__ vc_out, float *__restrict__ trans_out,
    unsigned long long *__restrict__ n_tests, unsigned long long *__restrict__ n_mixed,
    const uint32_t *__restrict__ ujob, const uint32_t *__restrict__ job_ub,
    const float2 *__restrict__ job_dist) {
  __shared__ float4 s_q[CL_BLOCK];
  __shared__ float4 s_t[CL_BLOCK];
  __shared__ int s_u, s_c, s_uniform;
  __shared__ uint32_t s_adj[27];
  __shared__ int64_t s_bucket;
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  if (tid == 0) {
    // which (cell, chunk) this block owns (heaviest chunks first)
    const uint32_t b = order ? (uint32_t)order[blockIdx.x] : blockIdx.x;
    int64_t lo = 0, hi = U - 1;
    while (lo < hi) {
      const int64_t mid = (lo + hi + 1) >> 1;
      if (chunk_start[mid] <= b)
        lo = mid;
      else
        hi = mid - 1;
    }
    s_u = (int)lo;
    s_c = (int)(b - chunk_start[lo]);
    s_uniform = 1;
  }
  __syncthreads();
  const int u = s_u;
  // a batch of jobs: this chunk's job -- its cells' range, d_dist and s_max
  // (job_dist[j] = (s_max, d_dist)), and its own statistics counters
  int64_t r0, r1;
  uint32_t job;
  cell_range(ujob, job_ub, U, u, r0, r1, job);
  if (ujob) {
    const float2 jd = job_dist[job];
    s_max = jd.x;
    d_dist = jd.y;
    n_tests += job;
    n_mixed += job;
  }
  const uint32_t f = ufirst[u], n = ucount[u];
  const uint32_t base = (uint32_t)s_c * CL_CHUNK;
  const uint32_t i0 = centers[f + base];
  bool have[CL_PER];
  uint32_t idx[CL_PER];
  ClPose P[CL_PER];
#pragma unroll
  for (int r = 0; r < CL_PER; ++r) {
    // wave wv holds chunk positions [128 wv, 128 wv + 128): consecutive in
    // the rotation order
    const uint32_t k = base + (uint32_t)(wv * (WAVE * CL_PER) + r * WAVE + lane);
    have[r] = k < n;
    idx[r] = have[r] ? centers[f + k] : i0;
    if (have[r]) {
      for (int q = 0; q < 27; ++q)
        if (adj[27 * (size_t)idx[r] + q] != adj[27 * (size_t)i0 + q]) s_uniform = 0;
    }
    load_pose(P[r], trans, quats, idx[r]);
  }
  if (tid < 27) s_adj[tid] = adj[27 * (size_t)i0 + tid];
  __syncthreads();
  unsigned long long tests = 0;
  if (s_uniform) {
    // one bound per slot: 64 consecutive poses of the rotation order each
    ClWave W[CL_PER];
    bool any[CL_PER];
#pragma unroll
    for (int r = 0; r < CL_PER; ++r) {
      W[r] = cluster_slot_bounds(P[r], have[r], d_dist);
      any[r] = __ballot(have[r]) != 0ull;
    }
    for (int blk = 0; blk < 27; ++blk) {
      const uint32_t h = s_adj[blk];
      if (h == 0u) continue;
      if (tid == 0) s_bucket = cell_lower_bound(ukeys, r0, r1, h);
      __syncthreads();
      const int64_t bu = s_bucket;
      __syncthreads();
      if (bu < 0) continue;
      const uint32_t bf = ufirst[bu], bn = ucount[bu];
#pragma unroll
      for (int r = 0; r < CL_PER; ++r) tests += have[r] ? bn : 0u;
      for (uint32_t c0 = 0; c0 < bn; c0 += CL_BLOCK) {
        const uint32_t cn = min((uint32_t)CL_BLOCK, bn - c0);
        if ((uint32_t)tid < cn) {
          s_q[tid] = cq[bf + c0 + tid];
          s_t[tid] = ct[bf + c0 + tid];
        }
        __syncthreads();
        for (uint32_t g0 = 0; g0 < cn; g0 += WAVE) {
          const uint32_t gj = g0 + (uint32_t)lane;
          bool live[CL_PER];
#pragma unroll
          for (int r = 0; r < CL_PER; ++r) live[r] = false;
          if (gj < cn) {
            const float4 mq = s_q[gj], mt = s_t[gj];
#pragma unroll
            for (int r = 0; r < CL_PER; ++r)
              live[r] = any[r] && !ppfc::cluster_skip(W[r], mq.x, mq.y, mq.z, mq.w, mt.x, mt.y, mt.z, use_l1);
          }
          uint64_t m[CL_PER], mask = 0;
#pragma unroll
          for (int r = 0; r < CL_PER; ++r) {
            m[r] = __ballot(live[r]);
            mask |= m[r];
          }
          while (mask) {
            const uint32_t b = (uint32_t)__builtin_ctzll(mask);
            mask &= mask - 1;
            const float4 q0 = s_q[g0 + b], t0 = s_t[g0 + b];
#pragma unroll
            for (int r = 0; r < CL_PER; ++r)
              if ((m[r] >> b) & 1ull) cl_test<use_l1, use_avg>(P[r], q0, t0, s_max);
          }
        }
        __syncthreads();
      }
    }
  } else {
    // the chunk's poses lie in several cells that share a hash (an FNV
    // collision): each pose walks its own neighbour list
    if (tid == 0) atomicAdd(n_mixed, 1ull);
#pragma unroll
    for (int r = 0; r < CL_PER; ++r) {
      if (!have[r]) continue;
      for (int blk = 0; blk < 27; ++blk) {
        const uint32_t h = adj[27 * (size_t)idx[r] + blk];
        if (h == 0u) continue;
        const int64_t bu = cell_lower_bound(ukeys, r0, r1, h);
        if (bu < 0) continue;
        const uint32_t bf = ufirst[bu], bn = ucount[bu];
        tests += bn;
        for (uint32_t jj = 0; jj < bn; ++jj)
          cl_test<use_l1, use_avg>(P[r], cq[bf + jj], ct[bf + jj], s_max);
      }
    }
  }
  tests = wave_sum_u64(tests);
  if ((tid % WAVE) == 0 && tests) atomicAdd(n_tests, tests);
#pragma unroll
  for (int r = 0; r < CL_PER; ++r) {
    if (have[r]) {
      vc_out[idx[r]] = P[r].vc;
      trans_out[3 * (size_t)idx[r]] = P[r].ox;
      trans_out[3 * (size_t)idx[r] + 1] = P[r].oy;
      trans_out[3 * (size_t)idx[r] + 2] = P[r].oz;
    }
  }
}

// First index of the maximum (thrust::max_element, model.cu:293-295) as one
// 64-bit atomicMax over the grid: key = order-preserving bits of the value
// (-0 counted as +0, NaN below every number) << 32 | ~index, so the largest
// value wins and, among equal values, the smallest index.
__device__ __forceinline__ unsigned long long argmax_key(float x, uint32_t i) {
  uint32_t b = ppfm::fbits(x);
  if ((b & 0x7fffffffu) == 0u) b = 0u;
  uint32_t k;
  if ((b & 0x7fffffffu) > 0x7f800000u)
    k = 0u;  // NaN
  else
    k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((unsigned long long)k << 32) | (unsigned long long)(0xffffffffu - i);
}

__global__ __launch_bounds__(256) void k_argmax(const float *__restrict__ v, int64_t P,
                                                unsigned long long *__restrict__ out) {
  __shared__ unsigned long long sk[256 / WAVE];
  unsigned long long best = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < P; i += (int64_t)gridDim.x * 256) {
    const unsigned long long k = argmax_key(v[i], (uint32_t)i);
    best = k > best ? k : best;
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    const unsigned long long w = __shfl_xor(best, o, WAVE);
    best = w > best ? w : best;
  }
  if (threadIdx.x % WAVE == 0) sk[threadIdx.x / WAVE] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 256 / WAVE; ++k) best = sk[k] > best ? sk[k] : best;
    atomicMax(out, best);
  }
}

// ------------------------------------------------- batched finishing
// Threshold/order, poses and clustering of every job of a batch, one launch
// per stage.  Job j's elements are a contiguous range [begin[j], begin[j+1])
// of each stage's arrays (candidates, then kept votes), jobs in order, so
// every per-job result is the sub-range the plain path computes alone.

// The job of element i: the last j with begin[j] <= i (begin ascending, nj
// jobs; an empty job shares its begin with the next one and is skipped).
__device__ __forceinline__ uint32_t job_of(const uint64_t *__restrict__ begin, int nj, uint64_t i) {
  int lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (begin[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (uint32_t)lo;
}

// select flags: candidate count above the job's fl(thr * fl(max)) (model.cu:155-170)
__global__ void k_bsel_flags(const FinJob *__restrict__ jobs, const uint64_t *__restrict__ cbeg, int nj,
                             uint64_t C, uint32_t *__restrict__ flags) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C) return;
  const uint32_t j = job_of(cbeg, nj, i);
  flags[i] = ((float)jobs[j].cand_count[i - cbeg[j]] > jobs[j].thr_f) ? 1u : 0u;
}

// kbeg[j] = the first kept vote of job j (j = 0..nj)
__global__ void k_bsel_begins(const uint64_t *__restrict__ cbeg, int nj, uint64_t C,
                              const uint32_t *__restrict__ pos, const uint32_t *__restrict__ total,
                              uint64_t *__restrict__ kbeg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > nj) return;
  kbeg[j] = cbeg[j] < C ? pos[cbeg[j]] : *total;
}

// kept candidates as one sort key: job, then count descending, then code
// ascending (the plain path's composite key below the job bits)
__global__ void k_bsel_scatter(const FinJob *__restrict__ jobs, const uint64_t *__restrict__ cbeg, int nj,
                               uint64_t C, const uint32_t *__restrict__ flags,
                               const uint32_t *__restrict__ pos, uint32_t max_count, int sh_m, int sh_c,
                               int sh_j, uint64_t *__restrict__ keys) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= C || !flags[i]) return;
  const uint32_t j = job_of(cbeg, nj, i);
  const uint64_t l = i - cbeg[j];
  const unsigned long long c = jobs[j].cand_code[l];
  const uint64_t s_r = c >> 32, mc = c & 0xffffffffull;
  keys[pos[i]] = ((uint64_t)j << sh_j) | ((uint64_t)(max_count - jobs[j].cand_count[l]) << sh_c) |
                 (s_r << sh_m) | mc;
}

__global__ void k_bsel_decode(const uint64_t *__restrict__ keys, int64_t P, uint32_t max_count, int sh_m,
                              int sh_c, int bc, unsigned long long *__restrict__ out_code,
                              uint32_t *__restrict__ out_cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint64_t k = keys[i];
  const uint64_t mc = k & ((1ull << sh_m) - 1ull);
  const uint64_t s_r = (k & ((1ull << sh_c) - 1ull)) >> sh_m;
  out_code[i] = (s_r << 32) | mc;
  out_cnt[i] = max_count - (uint32_t)((k >> sh_c) & ((1ull << bc) - 1ull));
}

__global__ void k_bposes(const unsigned long long *__restrict__ codes, int64_t P,
                         const uint64_t *__restrict__ kbeg, int nj, const FinJob *__restrict__ jobs,
                         float *__restrict__ T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const FinJob &J = jobs[job_of(kbeg, nj, (uint64_t)i)];
  pose_of_code(codes[i], J.mframe, J.sframe, J.rpdf, T + i * 16);
}

// words: (job, cell hash) above the pose index (bg bits)
__global__ void k_bcluster_prep(const float *__restrict__ T, const unsigned long long *__restrict__ codes,
                                const uint32_t *__restrict__ cnt, int64_t P, const uint64_t *__restrict__ kbeg,
                                int nj, const FinJob *__restrict__ jobs, int bg, float *__restrict__ trans,
                                float *__restrict__ quats, float *__restrict__ w,
                                uint64_t *__restrict__ words, uint32_t *__restrict__ adj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const uint32_t j = job_of(kbeg, nj, (uint64_t)i);
  const uint32_t h = cluster_prep_one(i, T, codes, cnt, jobs[j].weights, jobs[j].d_dist, trans, quats, w, adj);
  words[i] = (((uint64_t)j << 32 | (uint64_t)h) << bg) | (uint64_t)i;
}

// job_ub[j] = the first cell of job j among the unique cells (j = 0..nj)
__global__ void k_bcell_begins(const uint32_t *__restrict__ ujob, int64_t U, int nj,
                               uint32_t *__restrict__ job_ub) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > nj) return;
  int64_t lo = 0, hi = U;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (ujob[mid] < (uint32_t)j)
      lo = mid + 1;
    else
      hi = mid;
  }
  job_ub[j] = (uint32_t)lo;
}

// chunk order inside each cell: (cell index u, rotation Morton code)
__global__ void k_bcluster_order_keys(const float *__restrict__ quats, const uint32_t *__restrict__ members,
                                      int64_t P, const uint32_t *__restrict__ ufirst, int64_t U,
                                      uint64_t *__restrict__ okeys, uint32_t *__restrict__ ovals) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  int64_t lo = 0, hi = U - 1;  // last u with ufirst[u] <= k
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (ufirst[mid] <= (uint32_t)k)
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint32_t i = members[k];
  okeys[k] = ((uint64_t)lo << 32) | (uint64_t)quat_order_key(quats, i);
  ovals[k] = i;
}

// first maximum of each job's vote counts (index within the job)
__global__ __launch_bounds__(256) void k_bargmax(const float *__restrict__ v, int64_t P,
                                                 const uint64_t *__restrict__ kbeg, int nj,
                                                 unsigned long long *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool valid = i < P;
  const uint32_t j = valid ? job_of(kbeg, nj, (uint64_t)i) : 0xffffffffu;
  unsigned long long key = valid ? argmax_key(v[i], (uint32_t)((uint64_t)i - kbeg[j])) : 0ull;
  const uint32_t j0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)j);
  if (__ballot(valid && j != j0) == 0ull) {
    // the wave's elements belong to one job: one atomic
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      const unsigned long long x = __shfl_xor(key, o, WAVE);
      key = x > key ? x : key;
    }
    if (lane_id() == 0 && j0 != 0xffffffffu) atomicMax(out + j0, key);
  } else if (valid) {
    atomicMax(out + j, key);
  }
}

// per job: final pose (T of the first maximum with the clustering's
// translation, ppf.cu:78-93), its index and the clustering statistics
__global__ void k_bfinal(const uint64_t *__restrict__ kbeg, int nj, const unsigned long long *__restrict__ amax,
                         const float *__restrict__ T, const float *__restrict__ trans_out,
                         const unsigned long long *__restrict__ jtests,
                         const unsigned long long *__restrict__ jmixed, uint32_t *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nj) return;
  uint32_t *o = out + 24 * (size_t)j;
  const uint64_t b = kbeg[j], P = kbeg[j + 1] - b;
  uint32_t mi = 0xffffffffu;
  if (P > 0) {
    mi = 0xffffffffu - (uint32_t)(amax[j] & 0xffffffffull);
    const float *t = T + (b + mi) * 16;
    for (int k = 0; k < 16; ++k) o[k] = ppfm::fbits(t[k]);
    o[3] = ppfm::fbits(trans_out[(b + mi) * 3]);
    o[7] = ppfm::fbits(trans_out[(b + mi) * 3 + 1]);
    o[11] = ppfm::fbits(trans_out[(b + mi) * 3 + 2]);
  } else {
    for (int k = 0; k < 16; ++k) o[k] = 0u;
  }
  o[16] = mi;
  o[17] = 0u;
  o[18] = (uint32_t)jtests[j];
  o[19] = (uint32_t)(jtests[j] >> 32);
  o[20] = (uint32_t)jmixed[j];
  o[21] = (uint32_t)(jmixed[j] >> 32);
}

// ------------------------------------------------ k-d slots within cells
// The poses of each cell are split recursively at the median of the widest
// of seven coordinates -- the normalized quaternion and the translation
// times KD_TRANS_SCALE / d_dist -- into runs of WAVE poses: a run of s > 64
// poses splits after round_up(floor(s / 2), 64) of them, so every slot but a
// cell's last is full and starts at a multiple of 64 from the cell's first
// pose.  A slot of a k-d leaf is more compact than 64 poses consecutive in
// the Hilbert order, so its bounds skip more members.  Which poses share a
// slot changes no result.
constexpr float KD_TRANS_SCALE = 0.6f;  // 0.2: clustering kernel 2% slower on the batched headline
constexpr int64_t KD_MIN_POSES = 8192;  // below this the levels cost more than they save

// The run [b, e) of cell-relative position rel after `level` splits of [0, n).
__device__ __forceinline__ void kd_run(uint32_t n, uint32_t rel, int level, uint32_t &b, uint32_t &e) {
  b = 0u;
  e = n;
  for (int l = 0; l < level && e - b > (uint32_t)WAVE; ++l) {
    const uint32_t L = ((e - b) / 2u + (uint32_t)WAVE - 1u) / (uint32_t)WAVE * (uint32_t)WAVE;
    if (rel < b + L)
      e = b + L;
    else
      b += L;
  }
}

// order-preserving u32 of a float (for atomic min/max)
__device__ __forceinline__ uint32_t ord_f32(float x) {
  const uint32_t b = ppfm::fbits(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t o) {
  return ppfm::bitsf((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// seven k-d coordinates of every pose (job j's translation scale for the
// poses in [kbeg[j], kbeg[j+1]))
__global__ void k_kd_feat(const float *__restrict__ trans, const float *__restrict__ quats, int64_t P,
                          const uint64_t *__restrict__ kbeg, int nj, const float2 *__restrict__ job_dist,
                          float d_dist, float *__restrict__ feat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  // job_dist[j] = (s_max, d_dist) of job j when the poses are a batch's
  const float d = job_dist ? job_dist[job_of(kbeg, nj, (uint64_t)i)].y : d_dist;
  const float sc = d > 0.0f ? KD_TRANS_SCALE / d : 0.0f;
  float q[4], n2 = 0.0f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    q[c] = quats[4 * i + c];
    n2 += q[c] * q[c];
  }
  const float inv = n2 > 0.0f && n2 < 1e30f ? 1.0f / sqrtf(n2) : 0.0f;
  float *f = feat + 8 * i;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float x = q[c] * inv;
    f[c] = x == x ? x : 0.0f;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float x = trans[3 * i + c] * sc;
    f[4 + c] = (x == x && fabsf(x) < 1e30f) ? x : 0.0f;
  }
}

// cell[k] = the cell of position k: the last u with ufirst[u] <= k
__global__ void k_kd_cells(int64_t P, const uint32_t *__restrict__ ufirst, int64_t U, uint32_t *__restrict__ cell) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  int64_t lo = 0, hi = U - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (ufirst[mid] <= (uint32_t)k)
      lo = mid;
    else
      hi = mid - 1;
  }
  cell[k] = (uint32_t)lo;
}

// The run of position k (in `centers`) at this level: global first position
// gb and size; leaf = at most one slot.
__device__ __forceinline__ void kd_locate(int64_t k, const uint32_t *__restrict__ cell,
                                          const uint32_t *__restrict__ ufirst,
                                          const uint32_t *__restrict__ ucount, int level, uint32_t &gb,
                                          uint32_t &size) {
  const uint32_t u = cell[k];
  uint32_t b, e;
  kd_run(ucount[u], (uint32_t)k - ufirst[u], level, b, e);
  gb = ufirst[u] + b;
  size = e - b;
}

// per run with more than one slot: min and max of each coordinate.  Runs are
// contiguous and longer than a wave, so a wave spans at most two of them:
// each is reduced across the wave first (and across the block when the whole
// block lies in one run) before the atomics.
constexpr int KD_BLOCK = 1024;

__device__ __forceinline__ void kd_wave_reduce(uint32_t lo[7], uint32_t hi[7]) {
#pragma unroll
  for (int c = 0; c < 7; ++c)
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      lo[c] = min(lo[c], (uint32_t)__shfl_xor((int)lo[c], o, WAVE));
      hi[c] = max(hi[c], (uint32_t)__shfl_xor((int)hi[c], o, WAVE));
    }
}

__global__ __launch_bounds__(KD_BLOCK) void k_kd_bounds(const uint32_t *__restrict__ centers, int64_t P,
                                                        const uint32_t *__restrict__ cell,
                                                        const uint32_t *__restrict__ ufirst,
                                                        const uint32_t *__restrict__ ucount, int level,
                                                        const float *__restrict__ feat,
                                                        uint32_t *__restrict__ mm) {
  __shared__ uint32_t s_gb;
  __shared__ int s_mixed;
  __shared__ uint32_t s_lo[KD_BLOCK / WAVE][7], s_hi[KD_BLOCK / WAVE][7];
  const int tid = threadIdx.x, wv = tid / WAVE, lane = tid % WAVE;
  const int64_t k = (int64_t)blockIdx.x * KD_BLOCK + tid;
  const bool valid = k < P;
  uint32_t gb = 0xffffffffu, size = 0u;
  if (valid) kd_locate(k, cell, ufirst, ucount, level, gb, size);
  const bool act = valid && size > (uint32_t)WAVE;
  // the pose's 32-byte record as two 16-byte loads (one cache line)
  float fv[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (act) {
    const float4 *rp = reinterpret_cast<const float4 *>(feat + 8 * (size_t)centers[k]);
    const float4 a = rp[0], b = rp[1];
    fv[0] = a.x;
    fv[1] = a.y;
    fv[2] = a.z;
    fv[3] = a.w;
    fv[4] = b.x;
    fv[5] = b.y;
    fv[6] = b.z;
  }
  uint32_t flo[7], fhi[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    const uint32_t o = act ? ord_f32(fv[c]) : 0u;
    flo[c] = act ? o : 0xffffffffu;
    fhi[c] = o;
  }
  if (tid == 0) {
    s_gb = 0xffffffffu;
    s_mixed = 0;
  }
  __syncthreads();
  if (act) atomicMin(&s_gb, gb);
  __syncthreads();
  if (act && gb != s_gb) s_mixed = 1;
  __syncthreads();
  if (!s_mixed) {
    // every active position of the block is in one run (or none is active)
    if (s_gb == 0xffffffffu) return;
    kd_wave_reduce(flo, fhi);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < 7; ++c) {
        s_lo[wv][c] = flo[c];
        s_hi[wv][c] = fhi[c];
      }
    }
    __syncthreads();
    if (tid < 7) {
      uint32_t lo = 0xffffffffu, hi = 0u;
      for (int q = 0; q < KD_BLOCK / WAVE; ++q) {
        lo = min(lo, s_lo[q][tid]);
        hi = max(hi, s_hi[q][tid]);
      }
      atomicMin(&mm[16 * (size_t)s_gb + tid], lo);
      atomicMax(&mm[16 * (size_t)s_gb + 8 + tid], hi);
    }
    return;
  }
  // at most two runs per wave: the first active lane's, then the rest
  const uint64_t am = __ballot(act);
  if (am == 0ull) return;
  const uint32_t g0 = (uint32_t)__builtin_amdgcn_readlane((int)gb, (int)__builtin_ctzll(am));
  for (int pass = 0; pass < 2; ++pass) {
    const bool mine = act && (pass == 0 ? gb == g0 : gb != g0);
    const uint64_t mm_ = __ballot(mine);
    if (mm_ == 0ull) break;
    const uint32_t g = (uint32_t)__builtin_amdgcn_readlane((int)gb, (int)__builtin_ctzll(mm_));
    uint32_t lo[7], hi[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      lo[c] = mine ? flo[c] : 0xffffffffu;
      hi[c] = mine ? fhi[c] : 0u;
    }
    kd_wave_reduce(lo, hi);
    if (lane < 7) {
      uint32_t l = lo[0], h = hi[0];
#pragma unroll
      for (int c = 1; c < 7; ++c)
        if (lane == c) {
          l = lo[c];
          h = hi[c];
        }
      atomicMin(&mm[16 * (size_t)g + lane], l);
      atomicMax(&mm[16 * (size_t)g + 8 + lane], h);
    }
  }
}

// reset the bounds of the runs that start at k
__global__ void k_kd_reset(int64_t P, const uint32_t *__restrict__ cell, const uint32_t *__restrict__ ufirst,
                           const uint32_t *__restrict__ ucount, int level, uint32_t *__restrict__ mm) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  uint32_t gb, size;
  kd_locate(k, cell, ufirst, ucount, level, gb, size);
  if (gb != (uint32_t)k || size <= (uint32_t)WAVE) return;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    mm[16 * (size_t)k + c] = 0xffffffffu;
    mm[16 * (size_t)k + 8 + c] = 0u;
  }
}

// sort key of position k: its run's first position, then its coordinate
// along the run's widest axis (8 bits: equal ones keep their order, so the
// split lands within 1/256 of the run's extent of the median); single-slot
// runs keep their order
__global__ void k_kd_keys(const uint32_t *__restrict__ centers, int64_t P, const uint32_t *__restrict__ cell,
                          const uint32_t *__restrict__ ufirst, const uint32_t *__restrict__ ucount, int level,
                          const float *__restrict__ feat, const uint32_t *__restrict__ mm,
                          uint64_t *__restrict__ keys) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  uint32_t gb, size;
  kd_locate(k, cell, ufirst, ucount, level, gb, size);
  uint32_t q = 0u;
  if (size > (uint32_t)WAVE) {
    int best = 0;
    float wbest = -1.0f, lob = 0.0f;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      const float lo = unord_f32(mm[16 * (size_t)gb + c]), hi = unord_f32(mm[16 * (size_t)gb + 8 + c]);
      const float wdt = hi - lo;
      if (wdt > wbest) {
        wbest = wdt;
        best = c;
        lob = lo;
      }
    }
    if (wbest > 0.0f) {
      const float x = (feat[8 * (size_t)centers[k] + best] - lob) / wbest;
      q = (uint32_t)(fminf(fmaxf(x, 0.0f), 1.0f) * 255.0f);
    }
  }
  keys[k] = ((uint64_t)gb << 8) | (uint64_t)q;
}

// ---------------------------------------------------------- diagnostics
// A stream of independent full-wave FMAs (eight chains per lane, sixteen
// v_fma_f32 per iteration): the reference VALU issue rate for instruction
// counters, as the copy kernel is for traffic counters.
__global__ __launch_bounds__(256) void k_fma_stream(float *__restrict__ out, int iters, float a, float b) {
  float x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = (float)(threadIdx.x + k);
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
      for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], a, b);
  }
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += x[k];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void k_copy_u32(const uint4 *__restrict__ in, uint4 *__restrict__ out, int64_t n4) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n4) out[i] = in[i];
}

__global__ void k_debug_numerics(int op, const float *__restrict__ a, const float *__restrict__ b,
                                 float *__restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = 0.0f;
  switch (op) {
    case 0: r = ppfm::acosf_own(a[i]); break;
    case 1: r = ppfm::atan2f_own(a[i], b[i]); break;
    case 2: r = ppfm::sinf_own(a[i]); break;
    case 3: r = ppfm::cosf_own(a[i]); break;
    case 4: r = ppfm::fmodf_exact(a[i], b[i]); break;
    case 5: r = ppfm::quant_downf(a[i], b[i]); break;
    case 6: r = a[i] / b[i]; break;
    case 7: r = sqrtf(a[i]); break;
    case 8: r = (float)ppfg::alpha_bin(a[2 * i], a[2 * i + 1], b[2 * i], b[2 * i + 1]); break;
    default: break;
  }
  out[i] = r;
}

inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// Largest float s with sqrtf(s) < t (t > 0 finite); sqrtf is correctly
// rounded and monotone, so {s : sqrtf(s) < t} is a prefix of the floats.
float sqrt_below_max(float t) {
  float s = t * t;
  if (sqrtf(s) < t) {
    for (;;) {
      const float nx = nextafterf(s, INFINITY);
      if (!(sqrtf(nx) < t)) break;
      s = nx;
    }
  } else {
    do {
      s = nextafterf(s, 0.0f);
    } while (!(sqrtf(s) < t) && s > 0.0f);
  }
  return s;
}

int bitlen(uint64_t x) {
  int b = 0;
  while (x) {
    ++b;
    x >>= 1;
  }
  return b;
}

}  // namespace

// =============================================================== launchers
int prep_cloud(ppf_ctx ctx, const float *d_aos_pts, const float *d_aos_nrm, int64_t n, float *soa,
               float *frames, uint32_t rpdf) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_prep_cloud, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, d_aos_pts, d_aos_nrm,
                     n, soa, frames, rpdf);
  PPF_HIP(hipGetLastError());
  return 0;
}

int cloud_pair_words(ppf_ctx ctx, const CloudDev &c, float d_dist, uint32_t rpdf, uint64_t *words) {
  if (c.n <= 0) return 0;
  dim3 g(nblk(c.n, B256), (unsigned)c.n);
  hipLaunchKernelGGL(k_pair_keys<true>, g, dim3(B256), 0, ctx->stream, c.soa.as<float>(), c.n, d_dist,
                     rpdf, (void *)words);
  PPF_HIP(hipGetLastError());
  return 0;
}

int cloud_pair_keys(ppf_ctx ctx, const CloudDev &c, float d_dist, uint32_t rpdf, uint32_t *keys) {
  if (c.n <= 0) return 0;
  dim3 g(nblk(c.n, B256), (unsigned)c.n);
  hipLaunchKernelGGL(k_pair_keys<false>, g, dim3(B256), 0, ctx->stream, c.soa.as<float>(), c.n,
                     d_dist, rpdf, (void *)keys);
  PPF_HIP(hipGetLastError());
  return 0;
}

int unique_count(ppf_ctx ctx, const uint64_t *words, int64_t n, int64_t *U_out, int shift) {
  *U_out = 0;
  if (n <= 0) return 0;
  const int64_t chunks = (n + B256 - 1) / B256;
  const int G = (int)std::min<int64_t>(chunks, UQ_MAX_BLOCKS);
  const int64_t per = (chunks + G - 1) / G * B256;
  PPF_TRY(ctx->tmp_c.ensure((size_t)G * 8 + 16));
  uint32_t *counts = ctx->tmp_c.as<uint32_t>();
  uint32_t *offs = counts + G;
  uint32_t *total = offs + G;
  hipLaunchKernelGGL(k_unique_count, dim3(G), dim3(B256), 0, ctx->stream, words, n, per, shift, counts);
  PPF_TRY(exclusive_scan_u32(counts, offs, G, ctx->tmp_d, ctx->stream, total));
  PPF_HIP(hipGetLastError());
  uint32_t U = 0;
  PPF_HIP(hipMemcpyAsync(&U, total, 4, hipMemcpyDeviceToHost, ctx->stream));
  PPF_HIP(hipStreamSynchronize(ctx->stream));
  *U_out = U;
  return 0;
}

int unique_emit(ppf_ctx ctx, const uint64_t *words, int64_t n, int64_t U, uint32_t *hkeys,
                uint32_t *hfirst, uint32_t *hcount, uint32_t *low, int shift, uint32_t *hjob) {
  if (n <= 0) return 0;
  const int64_t chunks = (n + B256 - 1) / B256;
  const int G = (int)std::min<int64_t>(chunks, UQ_MAX_BLOCKS);
  const int64_t per = (chunks + G - 1) / G * B256;
  const uint32_t *offs = ctx->tmp_c.as<uint32_t>() + G;  // left by unique_count
  hipLaunchKernelGGL(k_unique_emit, dim3(G), dim3(B256), 0, ctx->stream, words, n, per, shift, offs,
                     hkeys, hfirst, low, hjob);
  if (U > 0)
    hipLaunchKernelGGL(k_unique_counts, dim3(nblk(U, 256)), dim3(256), 0, ctx->stream, hfirst, U, n,
                       hcount);
  PPF_HIP(hipGetLastError());
  return 0;
}

int build_htable(ppf_ctx ctx, const uint32_t *hkeys, int64_t U, uint32_t *htab, uint32_t hcap,
                 uint32_t hshift) {
  PPF_HIP(hipMemsetAsync(htab, 0, (size_t)hcap * 8, ctx->stream));
  if (U <= 0) return 0;
  hipLaunchKernelGGL(k_htable_insert, dim3(nblk(U, 256)), dim3(256), 0, ctx->stream, hkeys, U, htab,
                     hshift, hcap - 1);
  PPF_HIP(hipGetLastError());
  return 0;
}

int build_tile_offsets(ppf_ctx ctx, const uint32_t *hfirst, const uint32_t *hcount,
                       const uint32_t *pair, int64_t U, int32_t n_m, int32_t TM, int32_t T,
                       uint32_t *toff) {
  if (U <= 0) return 0;
  hipLaunchKernelGGL(k_tile_offsets, dim3(nblk(U, 256)), dim3(256), 0, ctx->stream, hfirst, hcount,
                     pair, U, n_m, TM, T, toff);
  PPF_HIP(hipGetLastError());
  return 0;
}

int vote_order(ppf_ctx ctx, uint64_t *words, uint32_t *vals, int64_t N, const CloudDev &m,
               int32_t TM, int32_t T, const uint32_t *hfirst, const uint32_t *hcount, int64_t U,
               uint32_t *toff) {
  if (N <= 0 || U <= 0) return 0;
  auto bits = [](uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; };
  const int ub = bits((uint64_t)U - 1), tb = bits((uint64_t)T - 1);
  hipLaunchKernelGGL(k_vote_order_words, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, words, vals, N,
                     (int32_t)m.n, TM, hfirst, U, tb);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(words, vals, N, 0, ub + tb + 6, ctx->sort, ctx->stream));
  // tile offsets (the same after the second sort: it reorders within segments)
  hipLaunchKernelGGL(k_tile_offsets, dim3(nblk(U, 256)), dim3(256), 0, ctx->stream, hfirst, hcount, vals,
                     U, (int32_t)m.n, TM, T, toff);
  PPF_HIP(hipGetLastError());
  // vote order within each (u, tile) segment: (rank in class, class), placed
  // directly from the run lengths.  It only spreads LDS banks; with more
  // (u, tile, class) runs than records (many tiny buckets) the run table would
  // outgrow the records, and the (u, tile, class, pair) order stays.
  const size_t n_cnt = (size_t)U * (size_t)T * 64;
  if (n_cnt > (size_t)N) return 0;
  PPF_TRY(ctx->tmp_a.ensure((size_t)N * 4));
  uint32_t *rank = ctx->tmp_a.as<uint32_t>();
  hipLaunchKernelGGL(k_vote_ranks, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, words, toff, T, tb, N,
                     rank);
  PPF_HIP(hipGetLastError());
  PPF_TRY(ctx->sort.tmp_keys.ensure(n_cnt * 4));
  PPF_TRY(ctx->sort.tmp_vals.ensure((size_t)N * 4));
  uint32_t *cnt = ctx->sort.tmp_keys.as<uint32_t>();
  uint32_t *vcopy = ctx->sort.tmp_vals.as<uint32_t>();
  PPF_HIP(hipMemsetAsync(cnt, 0, n_cnt * 4, ctx->stream));
  PPF_HIP(hipMemcpyAsync(vcopy, vals, (size_t)N * 4, hipMemcpyDeviceToDevice, ctx->stream));
  hipLaunchKernelGGL(k_vote_run_counts, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, words, rank, N, T,
                     tb, cnt);
  hipLaunchKernelGGL(k_vote_place, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, words, rank, vcopy, toff,
                     cnt, N, T, tb, vals);
  PPF_HIP(hipGetLastError());
  return 0;
}

int build_records(ppf_ctx ctx, const uint32_t *pair, int64_t N, const float *mframe,
                  const CloudDev &m, int32_t TM, uint32_t *rec) {
  if (N <= 0) return 0;
  hipLaunchKernelGGL(k_records, dim3(nblk(N, 256)), dim3(256), 0, ctx->stream, pair, N, mframe,
                     m.soa.as<float>(), (int32_t)m.n, TM, rec);
  PPF_HIP(hipGetLastError());
  return 0;
}

// Words of dynamic LDS before VoteShared: the largest accumulator (31 rows of
// pitch 2^k >= 64, or 2^k + 1 for model files written with that pitch), or
// the hit-grouping table, rounded to 8 bytes.
uint32_t vote_misc_words(int max_tm) {
  uint32_t pw = 64;
  while ((int)pw < max_tm) pw <<= 1;
  const uint32_t w = std::max((pw + 1u) * N_BINS, (uint32_t)2 * GT_SIZE);
  return (w + 1u) & ~1u;
}

// Dynamic LDS of k_vote: accumulator or grouping table, VoteShared, and the
// waves' queues of undecided votes.
size_t vote_lds_bytes(uint32_t misc_words) {
  return (size_t)misc_words * 4 + sizeof(VoteShared) + (size_t)(VOTE_BLOCK / WAVE) * VQ_CAP * sizeof(uint2);
}

// hits + sorted hits (16 B each), slot ids (4 B), groups (8 B x GT_SIZE), in uint4
uint32_t vote_hit_stride_u4(uint32_t hit_cap) {
  return 2 * hit_cap + (uint32_t)(((size_t)(hit_cap + (hit_cap & 1u)) * 4 + (size_t)GT_SIZE * 8 + 15) / 16);
}

size_t vote_hit_bytes_per_block(int64_t max_ns) { return (size_t)vote_hit_stride_u4((uint32_t)max_ns) * 16; }

// Capture what `body` launches on stream st into an instantiated graph.
static int capture_graph(hipStream_t st, const std::function<int()> &body, hipGraphExec_t *out) {
  *out = nullptr;
  PPF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = body();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != 0) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("vote graph capture: ") + hipGetErrorString(e));
    return PPF_ERR_HIP;
  }
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    *out = nullptr;
    set_error(std::string("vote graph instantiation: ") + hipGetErrorString(ei));
    return PPF_ERR_HIP;
  }
  return 0;
}

void release_vote_graphs(ppf_ctx ctx) {
  for (hipGraphExec_t *g : {&ctx->vote_pre, &ctx->vote_run})
    if (*g) {
      (void)hipGraphExecDestroy(*g);
      *g = nullptr;
    }
}

int run_votes(ppf_ctx ctx, const VoteJob *h_jobs, int n_jobs, uint64_t n_items, int max_tm,
              int64_t max_ns, bool exact, float *vote_ms, int max_blocks, bool *graph) {
  *vote_ms = 0.0f;
  const bool want_graph = graph && *graph;
  if (graph) *graph = false;
  if (n_items == 0 || n_jobs <= 0) return 0;
  const uint32_t misc_words = vote_misc_words(max_tm);
  const size_t lds = vote_lds_bytes(misc_words);
  if (lds > ctx->max_lds) {
    set_error("run_votes: accumulator tile exceeds the LDS limit");
    return PPF_ERR_ARG;
  }
  auto kern = exact ? k_vote<true> : k_vote<false>;
  PPF_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  PPF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, VOTE_BLOCK, lds));
  per_cu = std::max(1, per_cu);
  if (ctx->root->tune_vote_per_cu > 0) per_cu = std::min(per_cu, ctx->root->tune_vote_per_cu);
  int G = ctx->n_cu * per_cu;
  if (max_blocks > 0) G = std::min(G, max_blocks);
  PPF_TRY(ctx->tmp_a.ensure(sizeof(VoteJob) * n_jobs));
  PPF_TRY(ctx->queue.ensure(16));
  const uint32_t hit_cap = (uint32_t)max_ns;
  const uint32_t stride_u4 = vote_hit_stride_u4(hit_cap);
  PPF_TRY(ctx->hits.ensure((size_t)G * stride_u4 * 16));
  PPF_HIP(hipMemcpyAsync(ctx->tmp_a.ptr, h_jobs, sizeof(VoteJob) * n_jobs, hipMemcpyHostToDevice,
                         ctx->stream));
  PPF_HIP(hipMemsetAsync(ctx->queue.ptr, 0, 16, ctx->stream));
  // heaviest rows first, when there are more rows than blocks
  const bool use_order = n_items > (uint64_t)G && n_items < ((uint64_t)1 << 32);
  if (want_graph && n_items < ((uint64_t)1 << 30)) {
    // The same launches as below as two captured graphs -- the prelude (queue
    // reset, row costs, their sort) and the vote launch -- instantiated once
    // per launch signature and replayed while it stays the same (repeated
    // batches); the prelude clears with fill kernels, not memset nodes.
    if (use_order) {
      PPF_TRY(ctx->tmp_c.ensure((size_t)n_items * 8));
      PPF_TRY(radix_sort_async_prepare((int64_t)n_items, 32, 64, ctx->sort, false, ctx->stream));
    }
    const uint64_t sig[] = {(uint64_t)(uintptr_t)kern, (uint64_t)G, (uint64_t)lds, (uint64_t)n_jobs, n_items,
                            hit_cap, stride_u4, misc_words, (uint64_t)(uintptr_t)ctx->tmp_a.ptr,
                            (uint64_t)(uintptr_t)ctx->queue.ptr, (uint64_t)(uintptr_t)ctx->hits.ptr,
                            (uint64_t)(uintptr_t)ctx->tmp_c.ptr, (uint64_t)(uintptr_t)ctx->sort.counts.ptr,
                            (uint64_t)(uintptr_t)ctx->sort.tmp_keys.ptr, (uint64_t)ctx->sort.tile_block,
                            (uint64_t)ctx->sort.tile_items, (uint64_t)use_order};
    static_assert(sizeof sig == sizeof ctx->vote_sig, "vote graph signature size");
    if (!ctx->vote_pre || !ctx->vote_run || std::memcmp(sig, ctx->vote_sig, sizeof sig) != 0) {
      release_vote_graphs(ctx);
      uint64_t *keys = ctx->tmp_c.as<uint64_t>();
      int rc = capture_graph(ctx->stream, [&]() -> int {
        PPF_TRY(fill_u32(ctx->queue.as<uint32_t>(), 0u, 4, ctx->stream));
        if (use_order) {
          hipLaunchKernelGGL(k_row_cost, dim3((unsigned)((n_items + 3) / 4)), dim3(256), 0, ctx->stream,
                             ctx->tmp_a.as<const VoteJob>(), n_jobs, n_items, keys);
          PPF_HIP(hipGetLastError());
          PPF_TRY(radix_sort_u64_async(keys, nullptr, (int64_t)n_items, 32, 64, ctx->sort, ctx->stream));
        }
        return 0;
      }, &ctx->vote_pre);
      if (rc == 0)
        rc = capture_graph(ctx->stream, [&]() -> int {
          hipLaunchKernelGGL(kern, dim3(G), dim3(VOTE_BLOCK), lds, ctx->stream, ctx->tmp_a.as<const VoteJob>(),
                             n_jobs, n_items, use_order ? (const uint64_t *)keys : nullptr,
                             ctx->queue.as<unsigned long long>(), ctx->hits.as<uint4>(), hit_cap, stride_u4,
                             misc_words);
          PPF_HIP(hipGetLastError());
          return 0;
        }, &ctx->vote_run);
      if (rc == 0) {
        std::memcpy(ctx->vote_sig, sig, sizeof sig);
      } else {
        // a runtime that cannot capture these launches: launch them directly
        release_vote_graphs(ctx);
        (void)hipGetLastError();
      }
    }
  }
  if (want_graph && ctx->vote_pre && ctx->vote_run && n_items < ((uint64_t)1 << 30)) {
    PPF_HIP(hipGraphLaunch(ctx->vote_pre, ctx->stream));
    PPF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    PPF_HIP(hipGraphLaunch(ctx->vote_run, ctx->stream));
    PPF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    PPF_HIP(hipEventSynchronize(ctx->ev1));
    PPF_HIP(hipEventElapsedTime(vote_ms, ctx->ev0, ctx->ev1));
    if (use_order) PPF_TRY(radix_sort_async_check((int64_t)n_items, 32, 64, ctx->sort, ctx->stream));
    *graph = true;
    return 0;
  }
  const uint64_t *order = nullptr;
  if (use_order) {
    PPF_TRY(ctx->tmp_c.ensure((size_t)n_items * 8));
    uint64_t *keys = ctx->tmp_c.as<uint64_t>();
    hipLaunchKernelGGL(k_row_cost, dim3((unsigned)((n_items + 3) / 4)), dim3(256), 0, ctx->stream,
                       ctx->tmp_a.as<const VoteJob>(), n_jobs, n_items, keys);
    PPF_HIP(hipGetLastError());
    PPF_TRY(radix_sort_u64(keys, nullptr, (int64_t)n_items, 32, 64, ctx->sort, ctx->stream));
    order = keys;
  }
  // vote_ms times the vote kernel alone
  PPF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(kern, dim3(G), dim3(VOTE_BLOCK), lds, ctx->stream,
                     ctx->tmp_a.as<const VoteJob>(), n_jobs, n_items, order,
                     ctx->queue.as<unsigned long long>(), ctx->hits.as<uint4>(), hit_cap, stride_u4,
                     misc_words);
  PPF_HIP(hipGetLastError());
  PPF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  PPF_HIP(hipEventSynchronize(ctx->ev1));
  PPF_HIP(hipEventElapsedTime(vote_ms, ctx->ev0, ctx->ev1));
  return 0;
}

int vote_slots(ppf_ctx ctx, int max_tm, int *slots) {
  const uint32_t misc_words = vote_misc_words(max_tm);
  const size_t lds = vote_lds_bytes(misc_words);
  auto kern = k_vote<false>;
  PPF_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  PPF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, VOTE_BLOCK, lds));
  *slots = ctx->n_cu * std::max(1, per_cu);
  return 0;
}

int select_and_sort(ppf_ctx ctx, const unsigned long long *cand_code, const uint32_t *cand_count,
                    uint32_t n_cand, uint32_t max_count, float thr, int64_t n_s, int64_t n_m,
                    unsigned long long *out_codes, uint32_t *out_counts, int64_t *P_out) {
  *P_out = 0;
  if (n_cand == 0) return 0;
  const float thr_f = thr * (float)max_count;  // model.cu:164
  PPF_TRY(ctx->tmp_b.ensure((size_t)n_cand * 8 + 16));
  uint32_t *flags = ctx->tmp_b.as<uint32_t>();
  uint32_t *pos = flags + n_cand;
  uint32_t *total = pos + n_cand;
  hipLaunchKernelGGL(k_select_flags, dim3(nblk(n_cand, 256)), dim3(256), 0, ctx->stream, cand_count,
                     n_cand, thr_f, flags);
  PPF_TRY(exclusive_scan_u32(flags, pos, n_cand, ctx->tmp_d, ctx->stream, total));
  uint32_t P = 0;
  PPF_HIP(hipMemcpyAsync(&P, total, 4, hipMemcpyDeviceToHost, ctx->stream));
  PPF_HIP(hipStreamSynchronize(ctx->stream));
  *P_out = P;
  if (P == 0) return 0;
  const int bs = std::max(1, bitlen((uint64_t)(n_s > 0 ? n_s - 1 : 0)));
  const int bm = bitlen((uint64_t)(n_m > 0 ? n_m - 1 : 0));
  const int bc = bitlen(max_count);
  const int sh_m = 6 + bm, sh_c = 6 + bm + bs;
  const bool composite = sh_c + bc <= 64;
  PPF_TRY(ctx->tmp_c.ensure((size_t)P * 16));
  uint64_t *keys = ctx->tmp_c.as<uint64_t>();
  uint32_t *vals = reinterpret_cast<uint32_t *>(keys + P);
  hipLaunchKernelGGL(k_select_scatter, dim3(nblk(n_cand, 256)), dim3(256), 0, ctx->stream, cand_code,
                     cand_count, n_cand, flags, pos, max_count, composite ? 1 : 0, sh_m, sh_c, keys,
                     vals);
  PPF_HIP(hipGetLastError());
  if (composite) {
    PPF_TRY(radix_sort_u64(keys, nullptr, P, 0, sh_c + bc, ctx->sort, ctx->stream));
    hipLaunchKernelGGL(k_select_decode, dim3(nblk(P, 256)), dim3(256), 0, ctx->stream, keys,
                       (int64_t)P, max_count, sh_m, sh_c, out_codes, out_counts);
  } else {
    PPF_TRY(radix_sort_u64(keys, vals, P, 0, 64, ctx->sort, ctx->stream));
    PPF_TRY(ctx->tmp_d.ensure((size_t)P * 12));
    uint64_t *k2 = ctx->tmp_d.as<uint64_t>();
    uint32_t *ix = reinterpret_cast<uint32_t *>(k2 + P);
    hipLaunchKernelGGL(k_iota_inv, dim3(nblk(P, 256)), dim3(256), 0, ctx->stream, vals, (int64_t)P,
                       max_count, k2, ix);
    PPF_TRY(radix_sort_u64(k2, ix, P, 0, std::max(8, bc), ctx->sort, ctx->stream));
    hipLaunchKernelGGL(k_gather_pairs, dim3(nblk(P, 256)), dim3(256), 0, ctx->stream, keys, vals, ix,
                       (int64_t)P, out_codes, out_counts);
  }
  PPF_HIP(hipGetLastError());
  return 0;
}

int compute_poses(ppf_ctx ctx, const unsigned long long *codes, int64_t P, const float *mframe,
                  const float *sframe, uint32_t rpdf, float *T) {
  if (P <= 0) return 0;
  hipLaunchKernelGGL(k_poses, dim3(nblk(P, 128)), dim3(128), 0, ctx->stream, codes, P, mframe, sframe,
                     rpdf, T);
  PPF_HIP(hipGetLastError());
  return 0;
}

// Reorder each cell's poses in `centers` into k-d slots (the kernels above);
// n_max = the largest cell.  Cells of at most one slot, and few poses, are
// left in their order.
int kd_slots(ppf_ctx ctx, hipStream_t st, uint32_t *centers, int64_t P, const uint32_t *ufirst,
             const uint32_t *ucount, int64_t U, uint32_t n_max, const float *trans, const float *quats,
             const uint64_t *kbeg, int nj, const float2 *job_dist, float d_dist, DevBuf &buf) {
  if (P < KD_MIN_POSES || n_max <= (uint32_t)WAVE || U <= 0) return 0;
  TraceRange tr("ppf.cluster.kd_slots");
  const int levels = bitlen((uint64_t)((n_max - 1u) / (uint32_t)WAVE)) + 1;
  PPF_TRY(buf.ensure((size_t)P * (32 + 64 + 8 + 4)));
  float *feat = buf.as<float>();
  uint32_t *mm = reinterpret_cast<uint32_t *>(feat + (size_t)P * 8);
  uint64_t *keys = reinterpret_cast<uint64_t *>(mm + (size_t)P * 16);
  uint32_t *cell = reinterpret_cast<uint32_t *>(keys + P);
  hipLaunchKernelGGL(k_kd_feat, dim3(nblk(P, 256)), dim3(256), 0, st, trans, quats, P, kbeg, nj, job_dist, d_dist,
                     feat);
  hipLaunchKernelGGL(k_kd_cells, dim3(nblk(P, 256)), dim3(256), 0, st, P, ufirst, U, cell);
  const int kbits = 8 + std::max(1, bitlen((uint64_t)P));
  // the levels' sorts are queued without host round trips; their look-back
  // time-out flags are read once, after the last
  PPF_TRY(radix_sort_async_prepare(P, 0, kbits, ctx->sort, true, st));
  for (int level = 0; level < levels; ++level) {
    hipLaunchKernelGGL(k_kd_reset, dim3(nblk(P, 256)), dim3(256), 0, st, P, cell, ufirst, ucount, level, mm);
    hipLaunchKernelGGL(k_kd_bounds, dim3(nblk(P, KD_BLOCK)), dim3(KD_BLOCK), 0, st, centers, P, cell, ufirst,
                       ucount, level, feat, mm);
    hipLaunchKernelGGL(k_kd_keys, dim3(nblk(P, 256)), dim3(256), 0, st, centers, P, cell, ufirst, ucount, level,
                       feat, mm, keys);
    PPF_HIP(hipGetLastError());
    PPF_TRY(radix_sort_u64_async(keys, centers, P, 0, kbits, ctx->sort, st));
  }
  PPF_TRY(radix_sort_async_check(P, 0, kbits, ctx->sort, st));
  return 0;
}

int cluster_spatial(ppf_ctx ctx, const float *T, const unsigned long long *codes,
                    const uint32_t *counts, const float *weights, int64_t P, float d_dist,
                    bool use_l1, bool use_avg, float *vc, float *trans_out, float *quats,
                    int64_t *max_idx, int64_t *pair_tests, int64_t *mixed_chunks, double *kernel_ms) {
  *max_idx = -1;
  *pair_tests = 0;
  *mixed_chunks = 0;
  *kernel_ms = 0.0;
  if (P <= 0) return 0;
  // cq, ct (2 x float4 P) words, okeys (2 x u64 P) trans (3P) w (P) adj (27P)
  // ukeys/ufirst/ucount (3P) members, centers (2P) argmax (8) test counter (8)
  const size_t bytes = (size_t)P * (32 + 16 + 12 + 4 + 108 + 12 + 8) + 128;
  PPF_TRY(ctx->tmp_b.ensure(bytes));
  char *p = ctx->tmp_b.as<char>();
  float4 *cq = reinterpret_cast<float4 *>(p);
  p += (size_t)P * 16;
  float4 *ct = reinterpret_cast<float4 *>(p);
  p += (size_t)P * 16;
  uint64_t *okeys = reinterpret_cast<uint64_t *>(p);
  p += (size_t)P * 8;
  uint64_t *words = reinterpret_cast<uint64_t *>(p);
  p += (size_t)P * 8;
  float *trans = reinterpret_cast<float *>(p);
  p += (size_t)P * 12;
  float *w = reinterpret_cast<float *>(p);
  p += (size_t)P * 4;
  uint32_t *adj = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 108;
  uint32_t *ukeys = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *ufirst = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *ucount = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *members = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *centers = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  long long *am = reinterpret_cast<long long *>(((uintptr_t)p + 7) & ~(uintptr_t)7);
  unsigned long long *ntest = reinterpret_cast<unsigned long long *>(am + 1);
  unsigned long long *nmixed = ntest + 1;
  unsigned long long *amax = ntest + 2;
  PPF_HIP(hipMemsetAsync(ntest, 0, 24, ctx->stream));
  hipLaunchKernelGGL(k_cluster_prep, dim3(nblk(P, 128)), dim3(128), 0, ctx->stream, T, codes, counts,
                     weights, P, d_dist, trans, quats, w, words, adj);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(words, nullptr, P, 32, 64, ctx->sort, ctx->stream));
  int64_t U = 0;
  PPF_TRY(unique_count(ctx, words, P, &U));
  PPF_TRY(unique_emit(ctx, words, P, U, ukeys, ufirst, ucount, members));
  // chunk order of each cell's poses (rotation Morton code); member data in
  // cell order for contiguous tile loads
  hipLaunchKernelGGL(k_cluster_order_keys, dim3(nblk(P, 256)), dim3(256), 0, ctx->stream, quats, words,
                     P, okeys, centers);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(okeys, centers, P, 0, 64, ctx->sort, ctx->stream));
  hipLaunchKernelGGL(k_cluster_gather, dim3(nblk(P, 256)), dim3(256), 0, ctx->stream, members, P, trans,
                     quats, w, cq, ct);
  PPF_HIP(hipGetLastError());
  // chunk table: chunk_start[u] = exclusive scan of ceil(count/256)
  PPF_TRY(ctx->tmp_a.ensure((size_t)(U + 1) * 8 + 64));
  uint32_t *nch = ctx->tmp_a.as<uint32_t>();
  uint32_t *cstart = nch + U + 1;
  uint32_t *ctot = cstart + U;
  PPF_HIP(hipMemsetAsync(ctot + 1, 0, 4, ctx->stream));
  hipLaunchKernelGGL(k_cluster_chunks, dim3(nblk(U, 256)), dim3(256), 0, ctx->stream, ucount, U, nch, ctot + 1);
  PPF_TRY(exclusive_scan_u32(nch, cstart, U, ctx->tmp_d, ctx->stream, ctot));
  uint32_t nit[2] = {0, 0};
  PPF_HIP(hipMemcpyAsync(nit, ctot, 8, hipMemcpyDeviceToHost, ctx->stream));
  PPF_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t n_items = nit[0];
  PPF_TRY(kd_slots(ctx, ctx->stream, centers, P, ufirst, ucount, U, nit[1], trans, quats, nullptr, 1, nullptr,
                   d_dist, ctx->tmp_e));
  if (n_items > 0) {
    PPF_TRY(ctx->tmp_c.ensure((size_t)n_items * 8));
    uint64_t *order = ctx->tmp_c.as<uint64_t>();
    hipLaunchKernelGGL(k_cluster_chunk_cost, dim3(nblk(n_items, 256)), dim3(256), 0, ctx->stream, cstart, U,
                       n_items, ukeys, ufirst, ucount, centers, adj, nullptr, nullptr, order);
    PPF_HIP(hipGetLastError());
    PPF_TRY(radix_sort_u64(order, nullptr, n_items, 32, 64, ctx->sort, ctx->stream));
    auto kern = use_avg ? (use_l1 ? k_cluster_tiled<1, 1> : k_cluster_tiled<0, 1>)
                        : (use_l1 ? k_cluster_tiled<1, 0> : k_cluster_tiled<0, 0>);
    PPF_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(kern, dim3(n_items), dim3(CL_BLOCK), 0, ctx->stream, trans, quats, cq, ct, adj,
                       ukeys, ufirst, ucount, U, cstart, centers, order, sqrt_below_max(d_dist), d_dist,
                       vc, trans_out, ntest, nmixed, nullptr, nullptr, nullptr);
    PPF_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  }
  hipLaunchKernelGGL(k_argmax, dim3((unsigned)std::min<int64_t>(nblk(P, 256), 1024)), dim3(256), 0,
                     ctx->stream, vc, P, amax);
  PPF_HIP(hipGetLastError());
  unsigned long long mi_t[3] = {0, 0, 0};
  PPF_HIP(hipMemcpyAsync(mi_t, ntest, 24, hipMemcpyDeviceToHost, ctx->stream));
  PPF_HIP(hipStreamSynchronize(ctx->stream));
  *max_idx = (int64_t)(0xffffffffu - (uint32_t)(mi_t[2] & 0xffffffffull));
  *pair_tests = (int64_t)mi_t[0];
  *mixed_chunks = (int64_t)mi_t[1];
  if (n_items > 0) {
    float ms = 0.0f;
    PPF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *kernel_ms = ms;
  }
  return 0;
}

bool finish_batch_fits(int nj, uint64_t total_cand, uint32_t max_count, int64_t max_ns, int64_t max_nm) {
  if (nj < 1 || total_cand >= ((uint64_t)1 << 31)) return false;
  const int bj = std::max(1, bitlen((uint64_t)(nj - 1)));
  const int bs = std::max(1, bitlen((uint64_t)(max_ns > 0 ? max_ns - 1 : 0)));
  const int bm = bitlen((uint64_t)(max_nm > 0 ? max_nm - 1 : 0));
  const int bc = std::max(1, bitlen(max_count));
  const int bg = std::max(1, bitlen(total_cand));
  return bj + bc + bs + bm + 6 <= 64 && bj + 32 + bg <= 64;
}

int finish_batch(ppf_ctx ctx, const std::vector<FinJob> &jobs, const std::vector<uint32_t> &n_cand,
                 uint32_t max_count, int64_t max_ns, int64_t max_nm, bool use_l1, bool use_avg,
                 const std::function<int(int64_t, FinishArrays &)> &alloc, FinishOut &out) {
  hipStream_t st = ctx->stream;
  const int nj = (int)jobs.size();
  std::vector<uint64_t> cbeg((size_t)nj + 1, 0);
  for (int j = 0; j < nj; ++j) cbeg[j + 1] = cbeg[j] + n_cand[j];
  const uint64_t C = cbeg[nj];
  if (!finish_batch_fits(nj, C, max_count, max_ns, max_nm)) {
    set_error("finish_batch: the batch's sort keys exceed 64 bits");
    return PPF_ERR_ARG;
  }
  const int bj = std::max(1, bitlen((uint64_t)(nj - 1)));
  const int bs = std::max(1, bitlen((uint64_t)(max_ns > 0 ? max_ns - 1 : 0)));
  const int bm = bitlen((uint64_t)(max_nm > 0 ? max_nm - 1 : 0));
  const int bc = std::max(1, bitlen(max_count));
  const int bg = std::max(1, bitlen(C));
  const int sh_m = 6 + bm, sh_c = 6 + bm + bs, sh_j = 6 + bm + bs + bc;

  // job table, candidate begins and per-job constants, then device-written
  // kept begins, cell begins, argmax keys, statistics and final poses
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_cbeg = al(sizeof(FinJob) * nj), o_dist = o_cbeg + al(8 * ((size_t)nj + 1));
  const size_t o_kbeg = o_dist + al(8 * (size_t)nj), o_jub = o_kbeg + al(8 * ((size_t)nj + 1));
  const size_t o_amax = o_jub + al(4 * ((size_t)nj + 1)), o_jt = o_amax + al(8 * (size_t)nj);
  const size_t o_jm = o_jt + al(8 * (size_t)nj), o_out = o_jm + al(8 * (size_t)nj);
  const size_t meta_bytes = o_out + al(96 * (size_t)nj);
  PPF_TRY(ctx->fin_meta.ensure(meta_bytes));
  char *meta = ctx->fin_meta.as<char>();
  const FinJob *d_jobs = reinterpret_cast<const FinJob *>(meta);
  const uint64_t *d_cbeg = reinterpret_cast<const uint64_t *>(meta + o_cbeg);
  const float2 *d_dist = reinterpret_cast<const float2 *>(meta + o_dist);
  uint64_t *d_kbeg = reinterpret_cast<uint64_t *>(meta + o_kbeg);
  uint32_t *d_jub = reinterpret_cast<uint32_t *>(meta + o_jub);
  unsigned long long *d_amax = reinterpret_cast<unsigned long long *>(meta + o_amax);
  unsigned long long *d_jt = reinterpret_cast<unsigned long long *>(meta + o_jt);
  unsigned long long *d_jm = reinterpret_cast<unsigned long long *>(meta + o_jm);
  uint32_t *d_out = reinterpret_cast<uint32_t *>(meta + o_out);
  {
    std::vector<char> h(o_kbeg, 0);
    std::memcpy(h.data(), jobs.data(), sizeof(FinJob) * nj);
    std::memcpy(h.data() + o_cbeg, cbeg.data(), 8 * ((size_t)nj + 1));
    float2 *dd = reinterpret_cast<float2 *>(h.data() + o_dist);
    for (int j = 0; j < nj; ++j) dd[j] = make_float2(sqrt_below_max(jobs[j].d_dist), jobs[j].d_dist);
    PPF_HIP(hipMemcpyAsync(meta, h.data(), o_kbeg, hipMemcpyHostToDevice, st));
    PPF_HIP(hipMemsetAsync(meta + o_amax, 0, o_out - o_amax, st));
  }
  PPF_HIP(hipEventRecord(ctx->ev4, st));
  TraceRange tr("ppf.finish_together");

  // ---- threshold and order (model.cu:148-170)
  PPF_TRY(ctx->fin_flags.ensure((size_t)C * 8 + 16));
  uint32_t *flags = ctx->fin_flags.as<uint32_t>();
  uint32_t *pos = flags + C;
  uint32_t *total = pos + C;
  if (C > 0) {
    hipLaunchKernelGGL(k_bsel_flags, dim3(nblk((int64_t)C, 256)), dim3(256), 0, st, d_jobs, d_cbeg, nj, C,
                       flags);
    PPF_TRY(exclusive_scan_u32(flags, pos, (int64_t)C, ctx->fin_scan, st, total));
  } else {
    PPF_HIP(hipMemsetAsync(total, 0, 4, st));
  }
  hipLaunchKernelGGL(k_bsel_begins, dim3(nblk(nj + 1, 256)), dim3(256), 0, st, d_cbeg, nj, C, pos, total,
                     d_kbeg);
  PPF_HIP(hipGetLastError());
  out.kbeg.assign((size_t)nj + 1, 0);
  PPF_HIP(hipMemcpyAsync(out.kbeg.data(), d_kbeg, 8 * ((size_t)nj + 1), hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  const int64_t P = (int64_t)out.kbeg[nj];
  FinishArrays A{};
  PPF_TRY(alloc(P, A));
  out.pose.assign((size_t)nj * 16, 0.0f);
  out.max_idx.assign((size_t)nj, -1);
  out.tests.assign((size_t)nj, 0);
  out.mixed.assign((size_t)nj, 0);
  if (P == 0) {
    PPF_HIP(hipEventRecord(ctx->ev5, st));
    PPF_HIP(hipEventRecord(ctx->ev6, st));
    PPF_HIP(hipEventRecord(ctx->ev3, st));
    PPF_HIP(hipEventSynchronize(ctx->ev3));
    return 0;
  }
  PPF_TRY(ctx->fin_keys.ensure((size_t)P * 8));
  uint64_t *keys = ctx->fin_keys.as<uint64_t>();
  hipLaunchKernelGGL(k_bsel_scatter, dim3(nblk((int64_t)C, 256)), dim3(256), 0, st, d_jobs, d_cbeg, nj, C,
                     flags, pos, max_count, sh_m, sh_c, sh_j, keys);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(keys, nullptr, P, 0, sh_j + bj, ctx->sort, st));
  hipLaunchKernelGGL(k_bsel_decode, dim3(nblk(P, 256)), dim3(256), 0, st, keys, P, max_count, sh_m, sh_c, bc,
                     A.codes, A.counts);
  PPF_HIP(hipGetLastError());
  PPF_HIP(hipEventRecord(ctx->ev5, st));

  // ---- poses (model.cu:191-200)
  hipLaunchKernelGGL(k_bposes, dim3(nblk(P, 128)), dim3(128), 0, st, A.codes, P, d_kbeg, nj, d_jobs, A.T);
  PPF_HIP(hipGetLastError());
  PPF_HIP(hipEventRecord(ctx->ev6, st));

  // ---- spatial-hash clustering (model.cu:202-244): cells of (job, hash)
  const size_t bytes = (size_t)P * (32 + 16 + 12 + 4 + 108 + 16 + 8) + 256;
  PPF_TRY(ctx->fin_cl.ensure(bytes));
  char *p = ctx->fin_cl.as<char>();
  float4 *cq = reinterpret_cast<float4 *>(p);
  p += (size_t)P * 16;
  float4 *ct = reinterpret_cast<float4 *>(p);
  p += (size_t)P * 16;
  uint64_t *okeys = reinterpret_cast<uint64_t *>(p);
  p += (size_t)P * 8;
  uint64_t *words = reinterpret_cast<uint64_t *>(p);
  p += (size_t)P * 8;
  float *trans = reinterpret_cast<float *>(p);
  p += (size_t)P * 12;
  float *w = reinterpret_cast<float *>(p);
  p += (size_t)P * 4;
  uint32_t *adj = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 108;
  uint32_t *ukeys = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *ufirst = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *ucount = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *ujob = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *members = reinterpret_cast<uint32_t *>(p);
  p += (size_t)P * 4;
  uint32_t *centers = reinterpret_cast<uint32_t *>(p);
  hipLaunchKernelGGL(k_bcluster_prep, dim3(nblk(P, 128)), dim3(128), 0, st, A.T, A.codes, A.counts, P, d_kbeg,
                     nj, d_jobs, bg, trans, A.quats, w, words, adj);
  PPF_HIP(hipGetLastError());
  // stable on the key bits: the pose index below them is ascending already
  PPF_TRY(radix_sort_u64(words, nullptr, P, bg, bg + 32 + bj, ctx->sort, st));
  int64_t U = 0;
  PPF_TRY(unique_count(ctx, words, P, &U, bg));
  PPF_TRY(unique_emit(ctx, words, P, U, ukeys, ufirst, ucount, members, bg, ujob));
  hipLaunchKernelGGL(k_bcell_begins, dim3(nblk(nj + 1, 256)), dim3(256), 0, st, ujob, U, nj, d_jub);
  hipLaunchKernelGGL(k_bcluster_order_keys, dim3(nblk(P, 256)), dim3(256), 0, st, A.quats, members, P, ufirst,
                     U, okeys, centers);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(okeys, centers, P, 0, 32 + std::max(1, bitlen((uint64_t)U)), ctx->sort, st));
  hipLaunchKernelGGL(k_cluster_gather, dim3(nblk(P, 256)), dim3(256), 0, st, members, P, trans, A.quats, w,
                     cq, ct);
  PPF_HIP(hipGetLastError());
  PPF_TRY(ctx->fin_chunk.ensure((size_t)(U + 1) * 8 + 64));
  uint32_t *nch = ctx->fin_chunk.as<uint32_t>();
  uint32_t *cstart = nch + U + 1;
  uint32_t *ctot = cstart + U;
  PPF_HIP(hipMemsetAsync(ctot + 1, 0, 4, st));
  hipLaunchKernelGGL(k_cluster_chunks, dim3(nblk(U, 256)), dim3(256), 0, st, ucount, U, nch, ctot + 1);
  PPF_TRY(exclusive_scan_u32(nch, cstart, U, ctx->fin_scan, st, ctot));
  uint32_t nit[2] = {0, 0};
  PPF_HIP(hipMemcpyAsync(nit, ctot, 8, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  const uint32_t n_items = nit[0];
  PPF_TRY(kd_slots(ctx, st, centers, P, ufirst, ucount, U, nit[1], trans, A.quats, d_kbeg, nj, d_dist, 0.0f,
                   ctx->fin_kd));
  if (n_items > 0) {
    PPF_TRY(ctx->fin_order.ensure((size_t)n_items * 8));
    uint64_t *order = ctx->fin_order.as<uint64_t>();
    hipLaunchKernelGGL(k_cluster_chunk_cost, dim3(nblk(n_items, 256)), dim3(256), 0, st, cstart, U, n_items,
                       ukeys, ufirst, ucount, centers, adj, ujob, d_jub, order);
    PPF_HIP(hipGetLastError());
    PPF_TRY(radix_sort_u64(order, nullptr, n_items, 32, 64, ctx->sort, st));
    auto kern = use_avg ? (use_l1 ? k_cluster_tiled<1, 1> : k_cluster_tiled<0, 1>)
                        : (use_l1 ? k_cluster_tiled<1, 0> : k_cluster_tiled<0, 0>);
    PPF_HIP(hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(kern, dim3(n_items), dim3(CL_BLOCK), 0, st, trans, A.quats, cq, ct, adj, ukeys, ufirst,
                       ucount, U, cstart, centers, order, 0.0f, 0.0f, A.vc, A.trans, d_jt, d_jm, ujob, d_jub,
                       d_dist);
    PPF_HIP(hipGetLastError());
    PPF_HIP(hipEventRecord(ctx->ev1, st));
  }
  hipLaunchKernelGGL(k_bargmax, dim3(nblk(P, 256)), dim3(256), 0, st, A.vc, P, d_kbeg, nj, d_amax);
  hipLaunchKernelGGL(k_bfinal, dim3(nblk(nj, 64)), dim3(64), 0, st, d_kbeg, nj, d_amax, A.T, A.trans, d_jt, d_jm,
                     d_out);
  PPF_HIP(hipGetLastError());
  std::vector<uint32_t> h((size_t)nj * 24);
  PPF_HIP(hipMemcpyAsync(h.data(), d_out, h.size() * 4, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipEventRecord(ctx->ev3, st));
  PPF_HIP(hipEventSynchronize(ctx->ev3));
  for (int j = 0; j < nj; ++j) {
    const uint32_t *o = h.data() + 24 * (size_t)j;
    std::memcpy(out.pose.data() + 16 * (size_t)j, o, 64);
    out.max_idx[j] = o[16] == 0xffffffffu ? -1 : (int64_t)o[16];
    out.tests[j] = (int64_t)(((uint64_t)o[19] << 32) | o[18]);
    out.mixed[j] = (int64_t)(((uint64_t)o[21] << 32) | o[20]);
  }
  float ms = 0.0f;
  PPF_HIP(hipEventElapsedTime(&ms, ctx->ev4, ctx->ev5));
  out.select_ms = ms;
  PPF_HIP(hipEventElapsedTime(&ms, ctx->ev5, ctx->ev6));
  out.pose_ms = ms;
  PPF_HIP(hipEventElapsedTime(&ms, ctx->ev6, ctx->ev3));
  out.cluster_ms = ms;
  out.cluster_kernel_ms = 0.0f;
  if (n_items > 0) {
    PPF_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    out.cluster_kernel_ms = ms;
  }
  return 0;
}

int debug_fma_stream(ppf_ctx ctx, float *out, int64_t blocks, int iters) {
  if (blocks <= 0) return 0;
  hipLaunchKernelGGL(k_fma_stream, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, out, iters, 0.999f,
                     0.001f);
  PPF_HIP(hipGetLastError());
  return 0;
}

int debug_copy(ppf_ctx ctx, const uint32_t *in, uint32_t *out, int64_t n) {
  const int64_t n4 = n / 4;
  if (n4 > 0)
    hipLaunchKernelGGL(k_copy_u32, dim3(nblk(n4, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const uint4 *>(in), reinterpret_cast<uint4 *>(out), n4);
  PPF_HIP(hipGetLastError());
  return 0;
}

int debug_numerics(ppf_ctx ctx, int op, const float *a, const float *b, float *out, int64_t n) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_debug_numerics, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, op, a, b, out, n);
  PPF_HIP(hipGetLastError());
  return 0;
}

}  // namespace ppfk
