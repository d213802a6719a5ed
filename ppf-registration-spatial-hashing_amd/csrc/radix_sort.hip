// radix_sort.hip -- stable LSD radix sort of 64-bit keys (+ optional 32-bit
// values) and an exclusive scan, written for wave64 CDNA.
//
// Replaces the reference's thrust::sort_by_key / thrust::sort / exclusive_scan
// call sites (include/impl/parallel_hash_array.hpp:63-65,73-76;
// src/cuda/model.cu:111-113,148,155-158,222-223).
//
// Per 8-bit digit pass:
//   count   : each block owns a contiguous run of 2048-element tiles and writes
//             its 256 digit counts, digit-major, to counts[d*G + b];
//   scan    : exclusive scan of counts gives each (digit, block) its base;
//   scatter : each block re-reads its tiles in order; inside a tile every wave
//             ranks its 64-element slots with 8 ballots (lanes holding the same
//             digit find each other), which keeps the sort stable; per-wave
//             digit totals in LDS turn the ranks into global positions.
#include <algorithm>

#include "ppf_internal.h"

#include <vector>

namespace ppfi {

namespace {

constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / WAVE;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;  // 2048
constexpr int RS_MAX_BLOCKS = 4096;

__global__ __launch_bounds__(RS_BLOCK) void k_radix_count(const uint64_t *__restrict__ keys,
                                                          int64_t n, int shift,
                                                          int64_t tiles_per_block, int64_t n_tiles,
                                                          uint32_t *__restrict__ counts, int G) {
  __shared__ uint32_t h[RS_WAVES][256];
  const int tid = threadIdx.x, w = tid / WAVE;
  for (int i = tid; i < RS_WAVES * 256; i += RS_BLOCK) (&h[0][0])[i] = 0;
  __syncthreads();
  const int64_t t0 = (int64_t)blockIdx.x * tiles_per_block;
  const int64_t t1 = min(t0 + tiles_per_block, n_tiles);
  for (int64_t t = t0; t < t1; ++t) {
    const int64_t base = t * RS_TILE;
    // all loads first (lanes past the end re-read the last key), then count
    uint64_t kk[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) kk[i] = keys[min(base + (int64_t)i * RS_BLOCK + tid, n - 1)];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i)
      if (base + (int64_t)i * RS_BLOCK + tid < n) atomicAdd(&h[w][(uint32_t)(kk[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < RS_WAVES; ++k) s += h[k][tid];
  counts[(int64_t)tid * G + blockIdx.x] = s;
}

template <bool VALS>
__global__ __launch_bounds__(RS_BLOCK) void k_radix_scatter(
    const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, int64_t tiles_per_block, int64_t n_tiles,
    const uint32_t *__restrict__ offsets, int G) {
  __shared__ uint32_t s_off[256];
  __shared__ uint32_t s_cnt[RS_WAVES][256];
  const int tid = threadIdx.x, w = tid / WAVE, l = tid % WAVE;
  s_off[tid] = offsets[(int64_t)tid * G + blockIdx.x];
  const uint64_t lt = lanemask_lt();
  const int64_t t0 = (int64_t)blockIdx.x * tiles_per_block;
  const int64_t t1 = min(t0 + tiles_per_block, n_tiles);
  for (int64_t t = t0; t < t1; ++t) {
#pragma unroll
    for (int k = 0; k < 4; ++k) s_cnt[w][l + 64 * k] = 0;
    __syncthreads();
    // wave w owns elements [base_w, base_w + 64*ITEMS); slot i, lane l
    const int64_t base_w = t * RS_TILE + (int64_t)w * WAVE * RS_ITEMS;
    uint64_t k[RS_ITEMS];
    uint32_t v[RS_ITEMS];
    uint32_t rank[RS_ITEMS];
    // all loads first (lanes past the end re-read the last element)
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
      const int64_t idx = min(base_w + (int64_t)i * WAVE + l, n - 1);
      k[i] = kin[idx];
      if (VALS) v[i] = vin[idx];
    }
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
      const int64_t idx = base_w + (int64_t)i * WAVE + l;
      const bool valid = idx < n;
      const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      const uint32_t before = (uint32_t)__popcll(peers & lt);
      uint32_t prev = 0;
      if (valid) prev = s_cnt[w][d];
      rank[i] = prev + before;
      if (valid && before == 0) s_cnt[w][d] = prev + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
      const int d = tid;
      uint32_t c[RS_WAVES];
#pragma unroll
      for (int q = 0; q < RS_WAVES; ++q) c[q] = s_cnt[q][d];
      uint32_t run = s_off[d];
#pragma unroll
      for (int q = 0; q < RS_WAVES; ++q) {
        s_cnt[q][d] = run;
        run += c[q];
      }
      s_off[d] = run;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; ++i) {
      const int64_t idx = base_w + (int64_t)i * WAVE + l;
      if (idx < n) {
        const uint32_t d = (uint32_t)(k[i] >> shift) & 255u;
        const uint32_t pos = s_cnt[w][d] + rank[i];
        kout[pos] = k[i];
        if (VALS) vout[pos] = v[i];
      }
    }
    __syncthreads();
  }
}

// ---- exclusive scan, three kernels (chunk sums, scan of sums, down-sweep)
constexpr int SC_BLOCK = 1024;
constexpr int SC_ITEMS = 4;
constexpr int SC_CHUNK = SC_BLOCK * SC_ITEMS;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_wave,
                                                    uint32_t *total) {
  const int tid = threadIdx.x, w = tid / WAVE, l = tid % WAVE;
  const int nw = blockDim.x / WAVE;
  uint32_t inc = wave_incl_scan_u32(v);
  if (l == WAVE - 1) s_wave[w] = inc;
  __syncthreads();
  if (w == 0) {
    uint32_t x = l < nw ? s_wave[l] : 0u;
    uint32_t xi = wave_incl_scan_u32(x);
    if (l < nw) s_wave[l] = xi - x;
    if (l == nw - 1) s_wave[nw] = xi;
  }
  __syncthreads();
  uint32_t r = inc - v + s_wave[w];
  *total = s_wave[nw];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(SC_BLOCK) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                          int64_t n, uint32_t *__restrict__ sums) {
  __shared__ uint32_t s[SC_BLOCK / WAVE + 1];
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i)
    if (base + i < n) v += in[base + i];
  v = wave_sum_u32(v);
  if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < SC_BLOCK / WAVE; ++i) t += s[i];
    sums[blockIdx.x] = t;
  }
}

// one block; nb <= SC_CHUNK; writes exclusive scan in place and the total
__global__ __launch_bounds__(SC_BLOCK) void k_scan_top(uint32_t *__restrict__ sums, int64_t nb,
                                                       uint32_t *__restrict__ total_out) {
  __shared__ uint32_t s[SC_BLOCK / WAVE + 1];
  const int64_t base = (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t x[SC_ITEMS];
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    x[i] = base + i < nb ? sums[base + i] : 0u;
    v += x[i];
  }
  uint32_t tot;
  uint32_t r = block_excl_scan(v, s, &tot);
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    if (base + i < nb) sums[base + i] = r;
    r += x[i];
  }
  if (threadIdx.x == 0 && total_out) *total_out = tot;
}

__global__ __launch_bounds__(SC_BLOCK) void k_scan_down(const uint32_t *__restrict__ in, int64_t n,
                                                        const uint32_t *__restrict__ sums,
                                                        uint32_t *__restrict__ out) {
  __shared__ uint32_t s[SC_BLOCK / WAVE + 1];
  const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
  uint32_t x[SC_ITEMS];
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    x[i] = base + i < n ? in[base + i] : 0u;
    v += x[i];
  }
  uint32_t tot;
  uint32_t r = block_excl_scan(v, s, &tot) + sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SC_ITEMS; ++i) {
    if (base + i < n) out[base + i] = r;
    r += x[i];
  }
}

// ---- one-sweep LSD radix sort: one histogram pass over the keys for every
// digit, then one kernel per digit that ranks a tile, finds the tile's
// global offsets by decoupled look-back over the preceding tiles' digit
// counts, and writes the tile in digit order through LDS so that each digit's
// run leaves the block as contiguous stores.
constexpr int OS_BLOCK = 256;
constexpr int OS_WAVES = OS_BLOCK / WAVE;
constexpr int OS_MAX_PASSES = 8;
// look-back status word per (tile, digit): flag in bits 30-31, count below
constexpr uint32_t OS_AGG = 1u << 30;      // the tile's own count of the digit
constexpr uint32_t OS_INC = 2u << 30;      // inclusive prefix through this tile
constexpr uint32_t OS_COUNT = (1u << 30) - 1u;
// a wait longer than this many status reads (about a second) is reported as
// an error instead of waiting on; every other waiting tile then stops too
constexpr uint32_t OS_SPIN_LIMIT = 1u << 20;

__global__ __launch_bounds__(OS_BLOCK) void k_os_hist(const uint64_t *__restrict__ keys, int64_t n,
                                                      int lo, int npass, uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[OS_MAX_PASSES][256];
  const int tid = threadIdx.x;
  for (int i = tid; i < OS_MAX_PASSES * 256; i += OS_BLOCK) (&h[0][0])[i] = 0u;
  __syncthreads();
  // each thread counts OS_HSEG keys, OS_BLOCK apart (coalesced loads), and
  // adds one count per run of equal digits: keys with long runs of equal
  // digits (vote-order words) would otherwise send 64 lanes' atomics to one
  // LDS word at once
  constexpr int OS_HSEG = 16;
  for (int64_t base = (int64_t)blockIdx.x * OS_BLOCK * OS_HSEG + tid; base < n;
       base += (int64_t)gridDim.x * OS_BLOCK * OS_HSEG) {
    const int cnt_k = (int)min((int64_t)OS_HSEG, (n - base + OS_BLOCK - 1) / OS_BLOCK);
    uint64_t k[OS_HSEG];
#pragma unroll
    for (int i = 0; i < OS_HSEG; ++i) k[i] = keys[base + (int64_t)min(i, cnt_k - 1) * OS_BLOCK];
    for (int p = 0; p < npass; ++p) {
      const int sh = lo + 8 * p;
      uint32_t cur = (uint32_t)(k[0] >> sh) & 255u, run = 1u;
#pragma unroll
      for (int i = 1; i < OS_HSEG; ++i) {
        if (i >= cnt_k) break;
        const uint32_t d = (uint32_t)(k[i] >> sh) & 255u;
        if (d == cur) {
          ++run;
        } else {
          atomicAdd(&h[p][cur], run);
          cur = d;
          run = 1u;
        }
      }
      atomicAdd(&h[p][cur], run);
    }
  }
  __syncthreads();
  for (int i = tid; i < npass * 256; i += OS_BLOCK) {
    const uint32_t c = (&h[0][0])[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// hist [npass][256] -> exclusive digit starts, in place; one block
__global__ __launch_bounds__(256) void k_os_starts(uint32_t *__restrict__ hist, int npass) {
  __shared__ uint32_t s[256 / WAVE + 1];
  for (int p = 0; p < npass; ++p) {
    uint32_t tot;
    const uint32_t v = hist[p * 256 + threadIdx.x];
    const uint32_t r = block_excl_scan(v, s, &tot);
    hist[p * 256 + threadIdx.x] = r;
  }
}

// Rank table of a sort by few distinct keys (sort_key32_ranked): a perfect
// hash of the distinct high halves, built on the host once they are known --
// bucket b = h1(key), slot = h2(key ^ seed(disp[b])) -- so a lookup is two
// gathers from 16-bit tables small enough for L1 (every key looked up is in
// the set, so no key is stored).
constexpr uint32_t RK_EMPTY = 0xFFFFFFFFu;
__host__ __device__ __forceinline__ uint32_t rk_hash(uint32_t key, int lg) { return (key * 2654435761u) >> (32 - lg); }
struct RankTable {
  const uint16_t *disp;  // per bucket
  const uint16_t *rank;  // per slot
  uint32_t m1, m2;       // odd multipliers of h1 and h2
  int lgb, lgs;          // log2 of the buckets and of the slots
};
__host__ __device__ __forceinline__ uint32_t rk_bucket(uint32_t key, const RankTable &rt) {
  return (key * rt.m1) >> (32 - rt.lgb);
}
__host__ __device__ __forceinline__ uint32_t rk_slot(uint32_t key, uint32_t d, const RankTable &rt) {
  return ((key ^ (d * 0x9E3779B9u)) * rt.m2) >> (32 - rt.lgs);
}

// The digit of word k at `shift`: of the word itself, or (RANKED) of the rank
// of its high 32 bits.
template <bool RANKED>
__device__ __forceinline__ uint32_t os_digit(uint64_t k, int shift, const RankTable &rt) {
  if (!RANKED) return (uint32_t)(k >> shift) & 255u;
  const uint32_t key = (uint32_t)(k >> 32);
  const uint32_t slot = rk_slot(key, rt.disp[rk_bucket(key, rt)], rt);
  return ((uint32_t)rt.rank[slot] >> shift) & 255u;
}

// BLOCK threads x ITEMS keys per thread per tile (the tuner picks the shape);
// threads 0..255 own one digit each in the count and look-back steps.
template <bool VALS, int BLOCK, int ITEMS, bool RANKED = false>
__global__ __launch_bounds__(BLOCK) void k_os_pass(
    const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint64_t *__restrict__ kout,
    uint32_t *__restrict__ vout, int64_t n, int shift, const uint32_t *__restrict__ starts,
    uint32_t *__restrict__ status, uint32_t *__restrict__ counters, RankTable rt) {
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int WV = BLOCK / WAVE;
  __shared__ uint64_t s_key[TILE];
  __shared__ uint32_t s_val[VALS ? TILE : 1];
  __shared__ uint8_t s_dig[RANKED ? TILE : 1];  // looked-up digits, in tile order
  __shared__ uint32_t s_cnt[WV][256];
  __shared__ uint32_t s_off[256];   // tile-local start of each digit
  __shared__ uint32_t s_glob[256];  // global position of tile-local position 0 of each digit
  __shared__ uint32_t s_wsum[WV + 1];
  __shared__ uint32_t s_tile;
  const int tid = threadIdx.x, w = tid / WAVE, l = tid % WAVE;
  if (tid == 0) s_tile = atomicAdd(&counters[0], 1u);  // tiles in acquisition order
  for (int i = tid; i < WV * 256; i += BLOCK) (&s_cnt[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t t = s_tile;
  const int64_t base = (int64_t)t * TILE;
  const int64_t base_w = base + (int64_t)w * WAVE * ITEMS;
  const uint64_t lt = lanemask_lt();
  uint64_t k[ITEMS];
  uint32_t v[ITEMS];
  uint32_t rank[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const int64_t idx = min(base_w + (int64_t)i * WAVE + l, n - 1);
    k[i] = kin[idx];
    if (VALS) v[i] = vin[idx];
  }
  // stable rank inside the wave: slots in order, lanes in order
  uint32_t dg[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) dg[i] = os_digit<RANKED>(k[i], shift, rt);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = base_w + (int64_t)i * WAVE + l < n;
    const uint32_t d = dg[i];
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t before = (uint32_t)__popcll(peers & lt);
    uint32_t prev = 0;
    if (valid) prev = s_cnt[w][d];
    rank[i] = prev + before;
    if (valid && before == 0) s_cnt[w][d] = prev + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // thread d < 256: the digit's count in the tile, its waves' offsets
  const int d = tid;
  const bool owner = d < 256;
  uint32_t cnt = 0;
  if (owner) {
#pragma unroll
    for (int q = 0; q < WV; ++q) {
      const uint32_t c = s_cnt[q][d];
      s_cnt[q][d] = cnt;
      cnt += c;
    }
  }
  // publish the tile's count at once, so that later tiles can look back
  uint32_t *st_row = status + (size_t)t * 256;
  if (owner)
    __hip_atomic_store(&st_row[d], (t == 0 ? OS_INC : OS_AGG) | cnt, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  uint32_t tot;
  const uint32_t off = block_excl_scan(cnt, s_wsum, &tot);  // tile-local start of digit d
  if (owner) {
    s_off[d] = off;
    // decoupled look-back over the preceding tiles' counts of digit d
    uint32_t excl = 0;
    if (t > 0) {
      int64_t j = (int64_t)t - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        const uint32_t sv = __hip_atomic_load(&status[(size_t)j * 256 + d], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t flag = sv & ~OS_COUNT;
        if (flag == 0u) {
          ++spins;
          if (spins > OS_SPIN_LIMIT) {  // never expected: report instead of waiting forever
            atomicOr(&counters[1], 1u);
            break;
          }
          if ((spins & 1023u) == 0u &&
              __hip_atomic_load(&counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
            break;
          continue;
        }
        excl += sv & OS_COUNT;
        if (flag == OS_INC) break;
        --j;
      }
      __hip_atomic_store(&st_row[d], OS_INC | (excl + cnt), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_glob[d] = starts[d] + excl - off;
  }
  __syncthreads();
  // the tile in digit order, in LDS
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if (base_w + (int64_t)i * WAVE + l < n) {
      const uint32_t dd = dg[i];
      const uint32_t lp = s_off[dd] + s_cnt[w][dd] + rank[i];
      s_key[lp] = k[i];
      if (VALS) s_val[lp] = v[i];
      if (RANKED) s_dig[lp] = (uint8_t)dd;
    }
  }
  __syncthreads();
  const int nt = (int)min((int64_t)TILE, n - base);
  for (int i = tid; i < nt; i += BLOCK) {
    const uint64_t kk = s_key[i];
    const uint32_t pos = s_glob[RANKED ? (uint32_t)s_dig[i] : os_digit<false>(kk, shift, rt)] + (uint32_t)i;
    kout[pos] = kk;
    if (VALS) vout[pos] = s_val[i];
  }
}

// the tile counters of npass one-sweep passes (even words of ctr)
__global__ void k_os_clear_tiles(uint32_t *__restrict__ ctr, int npass) {
  if ((int)threadIdx.x < npass) ctr[2 * threadIdx.x] = 0u;
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------ sort by few distinct keys
// The words' high 32 bits (the PPF keys of a model: a few thousand distinct
// values) are collected with their counts in one read pass -- a hash table
// in each block's LDS, merged into a global one -- sorted, and replaced by
// their rank: two one-sweep passes over the rank's 8-bit digits then place
// the words (40 bytes moved per word instead of 72 for four passes over the
// key).  More distinct keys than RK_MAX: the caller sorts by the key itself.
constexpr int RK_MAX = 8192;  // distinct keys at most (13-bit ranks)
constexpr int RK_LG = 14;     // global and block tables: 16,384 probed slots (+1)
constexpr int RK_BLOCK = 1024;
constexpr int RK_UNROLL = 4;

__global__ void k_rk_init(uint2 *__restrict__ gent, uint32_t *__restrict__ meta) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= (1u << RK_LG)) gent[i] = make_uint2(RK_EMPTY, 0u);
  if (i < 4) meta[i] = 0u;
}

// meta[0]: distinct keys, meta[1]: too many (the caller falls back)
__global__ __launch_bounds__(RK_BLOCK) void k_rk_discover(const uint64_t *__restrict__ words, int64_t n,
                                                          uint2 *__restrict__ gent,
                                                          uint32_t *__restrict__ meta) {
  constexpr uint32_t NS = 1u << RK_LG;
  __shared__ uint32_t s_tag[NS];
  __shared__ uint32_t s_cnt[NS + 1];  // [NS]: the key RK_EMPTY
  __shared__ uint32_t s_n, s_ovf;
  const int tid = threadIdx.x;
  for (uint32_t i = tid; i < NS; i += RK_BLOCK) {
    s_tag[i] = RK_EMPTY;
    s_cnt[i] = 0u;
  }
  if (tid == 0) {
    s_cnt[NS] = 0u;
    s_n = 0u;
    s_ovf = 0u;
  }
  __syncthreads();
  const int64_t step = (int64_t)gridDim.x * RK_BLOCK * RK_UNROLL;
  for (int64_t i0 = (int64_t)blockIdx.x * RK_BLOCK * RK_UNROLL + tid; i0 < n; i0 += step) {
    uint32_t kk[RK_UNROLL];
#pragma unroll
    for (int u = 0; u < RK_UNROLL; ++u) {
      const int64_t idx = i0 + (int64_t)u * RK_BLOCK;
      kk[u] = idx < n ? (uint32_t)(words[idx] >> 32) : 0u;
    }
#pragma unroll
    for (int u = 0; u < RK_UNROLL; ++u) {
      if (i0 + (int64_t)u * RK_BLOCK >= n) break;
      const uint32_t key = kk[u];
      if (key == RK_EMPTY) {
        atomicAdd(&s_cnt[NS], 1u);
        continue;
      }
      uint32_t slot = rk_hash(key, RK_LG);
      uint32_t p = 0;
      for (; p < NS; ++p) {
        uint32_t cur = s_tag[slot];
        if (cur == RK_EMPTY) {
          cur = atomicCAS(&s_tag[slot], RK_EMPTY, key);
          if (cur == RK_EMPTY) {
            if (atomicAdd(&s_n, 1u) >= (uint32_t)RK_MAX) s_ovf = 1u;
            cur = key;
          }
        }
        if (cur == key) {
          atomicAdd(&s_cnt[slot], 1u);
          break;
        }
        slot = (slot + 1u) & (NS - 1u);
      }
      if (p == NS) s_ovf = 1u;
    }
    if (s_ovf) break;
  }
  __syncthreads();
  if (s_ovf) {
    if (tid == 0) atomicOr(&meta[1], 1u);
    return;
  }
  // merge into the global table
  for (uint32_t i = tid; i <= NS; i += RK_BLOCK) {
    const uint32_t c = s_cnt[i];
    if (c == 0u) continue;
    uint32_t slot = NS;
    if (i < NS) {
      const uint32_t key = s_tag[i];
      slot = rk_hash(key, RK_LG);
      uint32_t p = 0;
      for (; p < NS; ++p) {
        const uint32_t cur = atomicCAS(&gent[slot].x, RK_EMPTY, key);
        if (cur == RK_EMPTY) {
          if (atomicAdd(&meta[0], 1u) >= (uint32_t)RK_MAX) atomicOr(&meta[1], 1u);
          break;
        }
        if (cur == key) break;
        slot = (slot + 1u) & (NS - 1u);
      }
      if (p == NS) {
        atomicOr(&meta[1], 1u);
        continue;
      }
    } else if (atomicCAS(&gent[NS].x, RK_EMPTY, 0u) == RK_EMPTY) {  // first block with the key RK_EMPTY
      if (atomicAdd(&meta[0], 1u) >= (uint32_t)RK_MAX) atomicOr(&meta[1], 1u);
    }
    atomicAdd(&gent[slot].y, c);
  }
}

// the distinct keys, as (key << 32 | slot), in table order
__global__ void k_rk_compact(const uint2 *__restrict__ gent, uint64_t *__restrict__ ukeys,
                             uint32_t *__restrict__ pos) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr uint32_t NS = 1u << RK_LG;
  if (slot > NS) return;
  const uint32_t tag = gent[slot].x;
  if (tag == RK_EMPTY) return;
  const uint32_t key = slot == NS ? RK_EMPTY : tag;
  const uint32_t p = atomicAdd(pos, 1u);
  if (p < (uint32_t)RK_MAX) ukeys[p] = ((uint64_t)key << 32) | slot;
}

// the histograms of both digits of the ranks of the sorted distinct keys
__global__ void k_rk_hist(const uint64_t *__restrict__ ukeys, uint32_t U, const uint2 *__restrict__ gent,
                          uint32_t *__restrict__ hist) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= U) return;
  const uint32_t c = gent[(uint32_t)ukeys[r]].y;
  atomicAdd(&hist[r & 255u], c);
  atomicAdd(&hist[256 + (r >> 8)], c);
}

}  // namespace

// Scan of n <= SC_CHUNK^2 elements using `sums` (>= nb entries) as scratch.
static int scan_level(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *sums, hipStream_t st,
                      uint32_t *total_out_dev) {
  const int64_t nb = (n + SC_CHUNK - 1) / SC_CHUNK;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(SC_BLOCK), 0, st, in, n, sums);
  const uint32_t *scanned = sums;
  if (nb <= SC_CHUNK) {
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SC_BLOCK), 0, st, sums, nb, total_out_dev);
  } else {
    // one more level: scan the chunk sums into sums + nb
    PPF_TRY(scan_level(sums, sums + nb, nb, sums + 2 * nb, st, total_out_dev));
    scanned = sums + nb;
  }
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(SC_BLOCK), 0, st, in, n, scanned, out);
  PPF_HIP(hipGetLastError());
  return 0;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp, hipStream_t st,
                       uint32_t *total_out_dev) {
  if (n <= 0) {
    if (total_out_dev) PPF_HIP(hipMemsetAsync(total_out_dev, 0, 4, st));
    return 0;
  }
  if (n > ((int64_t)1 << 32) - 1) {
    set_error("exclusive_scan_u32: input too long");
    return 1;
  }
  // scratch for every level: 2*nb + 2*nb/SC_CHUNK + ... (< 3*nb + 8)
  const int64_t nb = (n + SC_CHUNK - 1) / SC_CHUNK;
  PPF_TRY(tmp.ensure((size_t)(3 * nb + 16) * 4));
  return scan_level(in, out, n, tmp.as<uint32_t>(), st, total_out_dev);
}

// Tile shape of a one-sweep sort of n keys: 512 x 16 keys for sorts of 2^24
// keys and more (the model build's, measured fastest there); the tuned shape
// below that.  Scratch: [hist/starts: 8 x 256][counters: 2 per pass]
// [status: n_tiles x 256] words.
struct OsShape {
  int block, items;
  int64_t n_tiles;
  size_t counts_bytes;
};
static OsShape onesweep_shape(int64_t n, const SortScratch &s) {
  const bool big = n >= ((int64_t)1 << 24);
  OsShape sh;
  sh.items = big ? 16 : (s.tile_items == 8 ? 8 : 16);
  sh.block = big ? 512 : (s.tile_block == 512 ? 512 : 256);
  const int64_t tile = (int64_t)sh.block * sh.items;
  sh.n_tiles = (n + tile - 1) / tile;
  sh.counts_bytes = ((size_t)OS_MAX_PASSES * 256 + 2 * OS_MAX_PASSES + (size_t)sh.n_tiles * 256) * 4;
  return sh;
}

// One-sweep form (n < 2^30): a histogram kernel, a start kernel, then one
// kernel per 8-bit digit.  Counters are cleared by a fill kernel (measured
// faster per call than the runtime's memset, and capturable in a HIP graph).
// async: no read-back of the look-back time-out flags (radix_sort_async_check
// reads them), so that sorts queue back to back.
static int radix_sort_onesweep(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi,
                               SortScratch &s, hipStream_t st, bool async = false) {
  const int npass = (bit_hi - bit_lo + 7) / 8;
  const OsShape shp = onesweep_shape(n, s);
  const int items = shp.items, block = shp.block;
  const int64_t n_tiles = shp.n_tiles;
  const size_t hist_words = (size_t)OS_MAX_PASSES * 256, ctr_words = 2 * OS_MAX_PASSES;
  PPF_TRY(s.counts.ensure(shp.counts_bytes));
  PPF_TRY(s.tmp_keys.ensure((size_t)n * 8));
  if (vals) PPF_TRY(s.tmp_vals.ensure((size_t)n * 4));
  uint32_t *hist = s.counts.as<uint32_t>();
  uint32_t *ctr = hist + hist_words;
  uint32_t *status = ctr + ctr_words;
  auto zero = [&](uint32_t *p, size_t words) -> int {
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, st, p, 0u,
                       (int64_t)words);
    PPF_HIP(hipGetLastError());
    return 0;
  };
  if (async) {
    // the time-out flags (odd counter words) stay set until
    // radix_sort_async_check reads them: only the tile counters are cleared
    PPF_TRY(zero(hist, hist_words));
    hipLaunchKernelGGL(k_os_clear_tiles, dim3(1), dim3(64), 0, st, ctr, npass);
    PPF_HIP(hipGetLastError());
  } else {
    PPF_TRY(zero(hist, hist_words + ctr_words));
  }
  const unsigned hist_blocks = (unsigned)std::min<int64_t>((n + OS_BLOCK * 64 - 1) / (OS_BLOCK * 64), 2048);
  hipLaunchKernelGGL(k_os_hist, dim3(hist_blocks), dim3(OS_BLOCK), 0, st, keys, n, bit_lo, npass, hist);
  hipLaunchKernelGGL(k_os_starts, dim3(1), dim3(256), 0, st, hist, npass);
  uint64_t *ka = keys, *kb = s.tmp_keys.as<uint64_t>();
  uint32_t *va = vals, *vb = vals ? s.tmp_vals.as<uint32_t>() : nullptr;
  for (int p = 0; p < npass; ++p) {
    PPF_TRY(zero(status, (size_t)n_tiles * 256));
    const int shift = bit_lo + 8 * p;
    decltype(&k_os_pass<true, 256, 16>) kern;
    if (block == 512)
      kern = vals ? (items == 8 ? k_os_pass<true, 512, 8> : k_os_pass<true, 512, 16>)
                  : (items == 8 ? k_os_pass<false, 512, 8> : k_os_pass<false, 512, 16>);
    else
      kern = vals ? (items == 8 ? k_os_pass<true, 256, 8> : k_os_pass<true, 256, 16>)
                  : (items == 8 ? k_os_pass<false, 256, 8> : k_os_pass<false, 256, 16>);
    hipLaunchKernelGGL(kern, dim3((unsigned)n_tiles), dim3(block), 0, st, ka, va, kb, vb, n, shift,
                       hist + 256 * p, status, ctr + 2 * p, RankTable{});
    PPF_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (npass & 1) {
    PPF_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    if (vals) PPF_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  }
  if (async) return 0;
  // look-back time-out flags (never expected)
  uint32_t h_ctr[2 * OS_MAX_PASSES];
  PPF_HIP(hipMemcpyAsync(h_ctr, ctr, sizeof h_ctr, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  for (int p = 0; p < npass; ++p)
    if (h_ctr[2 * p + 1]) {
      set_error("radix_sort_u64: look-back wait exceeded its spin limit");
      return 2;
    }
  return 0;
}

// The perfect hash of the sorted distinct keys (hash and displace): buckets
// in decreasing size each take the first 16-bit seed that sends all their
// keys to distinct free slots.  False if some bucket finds none.
static bool rk_build(const std::vector<uint32_t> &keys, RankTable &rt, std::vector<uint16_t> &disp,
                     std::vector<uint16_t> &rank) {
  const uint32_t U = (uint32_t)keys.size();
  const uint32_t B = 1u << rt.lgb, S = 1u << rt.lgs;
  std::vector<uint32_t> bstart(B + 1, 0u), members(U);
  for (uint32_t r = 0; r < U; ++r) ++bstart[rk_bucket(keys[r], rt) + 1];
  for (uint32_t b = 0; b < B; ++b) bstart[b + 1] += bstart[b];
  {
    std::vector<uint32_t> fill(bstart.begin(), bstart.end() - 1);
    for (uint32_t r = 0; r < U; ++r) members[fill[rk_bucket(keys[r], rt)]++] = r;
  }
  std::vector<uint32_t> order(B);
  for (uint32_t b = 0; b < B; ++b) order[b] = b;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return bstart[a + 1] - bstart[a] > bstart[b + 1] - bstart[b];
  });
  disp.assign(B, 0);
  rank.assign(S, 0);
  std::vector<uint8_t> used(S, 0);
  std::vector<uint32_t> slots;
  for (uint32_t b : order) {
    const uint32_t m0 = bstart[b], m1 = bstart[b + 1];
    if (m0 == m1) break;  // the rest are empty
    bool placed = false;
    for (uint32_t d = 0; d < 65536u && !placed; ++d) {
      slots.clear();
      bool ok = true;
      for (uint32_t m = m0; m < m1 && ok; ++m) {
        const uint32_t sl = rk_slot(keys[members[m]], d, rt);
        if (used[sl] || std::find(slots.begin(), slots.end(), sl) != slots.end()) ok = false;
        slots.push_back(sl);
      }
      if (!ok) continue;
      for (uint32_t m = m0; m < m1; ++m) {
        used[slots[m - m0]] = 1;
        rank[slots[m - m0]] = (uint16_t)members[m];
      }
      disp[b] = (uint16_t)d;
      placed = true;
    }
    if (!placed) return false;
  }
  return true;
}

int sort_key32_ranked(uint64_t *words, int64_t n, SortScratch &s, hipStream_t st, bool *done) {
  *done = false;
  if (n < ((int64_t)1 << 20) || n >= ((int64_t)1 << 30)) return 0;
  constexpr size_t NE = ((size_t)1 << RK_LG) + 1;
  // [entries: NE uint2][meta: 4 u32 (+ pad)][ukeys: RK_MAX u64]
  PPF_TRY(s.rank.ensure(NE * 8 + 32 + (size_t)RK_MAX * 8));
  uint2 *gent = s.rank.as<uint2>();
  uint32_t *meta = reinterpret_cast<uint32_t *>(gent + NE);
  uint64_t *ukeys = reinterpret_cast<uint64_t *>(meta + 8);
  hipLaunchKernelGGL(k_rk_init, dim3((unsigned)((NE + 255) / 256)), dim3(256), 0, st, gent, meta);
  const unsigned blocks =
      (unsigned)std::min<int64_t>((n + (int64_t)RK_BLOCK * RK_UNROLL * 16 - 1) / ((int64_t)RK_BLOCK * RK_UNROLL * 16), 256);
  hipLaunchKernelGGL(k_rk_discover, dim3(blocks), dim3(RK_BLOCK), 0, st, words, n, gent, meta);
  PPF_HIP(hipGetLastError());
  uint32_t mh[4] = {0, 0, 0, 0};
  PPF_HIP(hipMemcpyAsync(mh, meta, 16, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  if (mh[1] != 0u || mh[0] == 0u || mh[0] > (uint32_t)RK_MAX) return 0;  // the caller sorts by the key
  const uint32_t U = mh[0];
  // [hist/starts: 2 x 256][counters: 2 per pass][status: n_tiles x 256]; the
  // tile keeps its digits in LDS too, so 256 x 16 keys leave room for three
  // tiles per CU
  constexpr int block = 256, items = 16;
  const int64_t tile = (int64_t)block * items;
  const int64_t n_tiles = (n + tile - 1) / tile;
  const size_t hist_words = 2 * 256, ctr_words = 4;
  // sized before the sort of the distinct keys, which uses the same scratch
  PPF_TRY(s.counts.ensure((hist_words + ctr_words + (size_t)n_tiles * 256) * 4));
  PPF_TRY(s.tmp_keys.ensure((size_t)n * 8));
  hipLaunchKernelGGL(k_rk_compact, dim3((unsigned)((NE + 255) / 256)), dim3(256), 0, st, gent, ukeys, meta + 2);
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(ukeys, nullptr, U, 32, 64, s, st));
  std::vector<uint64_t> uk(U);
  PPF_HIP(hipMemcpyAsync(uk.data(), ukeys, (size_t)U * 8, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  std::vector<uint32_t> keys(U);
  for (uint32_t r = 0; r < U; ++r) keys[r] = (uint32_t)(uk[r] >> 32);
  // slots: a power of two >= 1.6 U; buckets: a quarter of that
  int lgs = 1;
  while ((1u << lgs) * 5u < U * 8u) ++lgs;
  RankTable rt{nullptr, nullptr, 0u, 0u, std::max(1, lgs - 2), lgs};
  std::vector<uint16_t> disp, rank;
  static const uint32_t mults[4][2] = {{0x9E3779B1u, 0x85EBCA6Bu}, {0xC2B2AE35u, 0x27D4EB2Fu},
                                       {0x165667B1u, 0xD3A2646Du}, {0xFD7046C5u, 0xB55A4F09u}};
  bool built = false;
  for (int t = 0; t < 4 && !built; ++t) {
    rt.m1 = mults[t][0];
    rt.m2 = mults[t][1];
    built = rk_build(keys, rt, disp, rank);
  }
  if (!built) return 0;  // the caller sorts by the key
  PPF_TRY(s.rank_tab.ensure((disp.size() + rank.size()) * 2));
  uint16_t *d_disp = s.rank_tab.as<uint16_t>();
  uint16_t *d_rank = d_disp + disp.size();
  PPF_HIP(hipMemcpy(d_disp, disp.data(), disp.size() * 2, hipMemcpyHostToDevice));
  PPF_HIP(hipMemcpy(d_rank, rank.data(), rank.size() * 2, hipMemcpyHostToDevice));
  rt.disp = d_disp;
  rt.rank = d_rank;
  uint32_t *hist = s.counts.as<uint32_t>();
  uint32_t *ctr = hist + hist_words;
  uint32_t *status = ctr + ctr_words;
  PPF_HIP(hipMemsetAsync(hist, 0, (hist_words + ctr_words) * 4, st));
  hipLaunchKernelGGL(k_rk_hist, dim3((U + 255) / 256), dim3(256), 0, st, ukeys, U, gent, hist);
  hipLaunchKernelGGL(k_os_starts, dim3(1), dim3(256), 0, st, hist, 2);
  PPF_HIP(hipGetLastError());
  uint64_t *ka = words, *kb = s.tmp_keys.as<uint64_t>();
  for (int p = 0; p < 2; ++p) {
    PPF_HIP(hipMemsetAsync(status, 0, (size_t)n_tiles * 256 * 4, st));
    hipLaunchKernelGGL((k_os_pass<false, block, items, true>), dim3((unsigned)n_tiles), dim3(block), 0, st, ka,
                       nullptr, kb, nullptr, n, 8 * p, hist + 256 * p, status, ctr + 2 * p, rt);
    PPF_HIP(hipGetLastError());
    std::swap(ka, kb);
  }
  uint32_t h_ctr[4];
  PPF_HIP(hipMemcpyAsync(h_ctr, ctr, sizeof h_ctr, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  if (h_ctr[1] || h_ctr[3]) {
    set_error("sort_key32_ranked: look-back wait exceeded its spin limit");
    return 2;
  }
  *done = true;
  return 0;
}

int fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
  PPF_HIP(hipGetLastError());
  return 0;
}

static bool async_sortable(int64_t n, int bit_lo, int bit_hi) {
  return n < ((int64_t)1 << 30) && bit_hi > bit_lo && bit_hi - bit_lo <= 8 * OS_MAX_PASSES;
}

int radix_sort_async_prepare(int64_t n, int bit_lo, int bit_hi, SortScratch &s, bool vals, hipStream_t st) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  if (!async_sortable(n, bit_lo, bit_hi)) {
    set_error("radix_sort_u64_async: more than 2^30 keys or 64 bits");
    return 1;
  }
  PPF_TRY(s.counts.ensure(onesweep_shape(n, s).counts_bytes));
  PPF_TRY(s.tmp_keys.ensure((size_t)n * 8));
  if (vals) PPF_TRY(s.tmp_vals.ensure((size_t)n * 4));
  // clear the time-out flags, which the sorts that follow only set
  PPF_HIP(hipMemsetAsync(s.counts.as<uint32_t>() + (size_t)OS_MAX_PASSES * 256, 0, 2 * OS_MAX_PASSES * 4, st));
  return 0;
}

int radix_sort_u64_async(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi, SortScratch &s,
                         hipStream_t st) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  if (!async_sortable(n, bit_lo, bit_hi)) {
    set_error("radix_sort_u64_async: more than 2^30 keys or 64 bits");
    return 1;
  }
  return radix_sort_onesweep(keys, vals, n, bit_lo, bit_hi, s, st, true);
}

int radix_sort_async_check(int64_t n, int bit_lo, int bit_hi, SortScratch &s, hipStream_t st) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  uint32_t h_ctr[2 * OS_MAX_PASSES];
  PPF_HIP(hipMemcpyAsync(h_ctr, s.counts.as<uint32_t>() + (size_t)OS_MAX_PASSES * 256, sizeof h_ctr,
                         hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  for (int p = 0; p < OS_MAX_PASSES; ++p)
    if (h_ctr[2 * p + 1]) {
      set_error("radix_sort_u64_async: look-back wait exceeded its spin limit");
      return 2;
    }
  return 0;
}

int radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi,
                   SortScratch &s, hipStream_t st) {
  if (n <= 1 || bit_hi <= bit_lo) return 0;
  if (n >= (int64_t)1 << 32) {
    set_error("radix_sort_u64: more than 2^32 elements");
    return 1;
  }
  if (n < ((int64_t)1 << 30) && bit_hi - bit_lo <= 8 * OS_MAX_PASSES) return radix_sort_onesweep(keys, vals, n, bit_lo, bit_hi, s, st);
  const int64_t n_tiles = (n + RS_TILE - 1) / RS_TILE;
  const int G = (int)std::min<int64_t>(n_tiles, RS_MAX_BLOCKS);
  const int64_t tpb = (n_tiles + G - 1) / G;
  PPF_TRY(s.counts.ensure((size_t)256 * G * 4 * 2));
  PPF_TRY(s.tmp_keys.ensure((size_t)n * 8));
  if (vals) PPF_TRY(s.tmp_vals.ensure((size_t)n * 4));
  uint32_t *counts = s.counts.as<uint32_t>();
  uint32_t *offs = counts + (size_t)256 * G;
  uint64_t *ka = keys, *kb = s.tmp_keys.as<uint64_t>();
  uint32_t *va = vals, *vb = vals ? s.tmp_vals.as<uint32_t>() : nullptr;
  int passes = 0;
  for (int shift = bit_lo; shift < bit_hi; shift += 8, ++passes) {
    hipLaunchKernelGGL(k_radix_count, dim3(G), dim3(RS_BLOCK), 0, st, ka, n, shift, tpb, n_tiles,
                       counts, G);
    PPF_TRY(exclusive_scan_u32(counts, offs, (int64_t)256 * G, s.scan, st, nullptr));
    if (vals)
      hipLaunchKernelGGL(k_radix_scatter<true>, dim3(G), dim3(RS_BLOCK), 0, st, ka, va, kb, vb, n,
                         shift, tpb, n_tiles, offs, G);
    else
      hipLaunchKernelGGL(k_radix_scatter<false>, dim3(G), dim3(RS_BLOCK), 0, st, ka, va, kb, vb, n,
                         shift, tpb, n_tiles, offs, G);
    PPF_HIP(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (passes & 1) {
    PPF_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
    if (vals) PPF_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  }
  return 0;
}

}  // namespace ppfi
