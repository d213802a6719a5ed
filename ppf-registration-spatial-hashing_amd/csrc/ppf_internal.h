// ppf_internal.h -- declarations shared by the HIP translation units of
// libppf_amd.so (not part of the public C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>

#include <string>

namespace ppfi {

// A roctx range over a host scope: the stages of the path appear by name
// in rocprofv3 --marker-trace (the reference's DEBUG stage timers,
// scene.cu:72-98, model.cu:271-305).  Costs a function call when no tool
// is attached.
struct TraceRange {
  explicit TraceRange(const char *name) { roctxRangePush(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;
};

constexpr int WAVE = 64;

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
#define PPF_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::ppfi::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + \
                        __FILE__ + ":" + std::to_string(__LINE__) + ")");          \
      return 2; /* PPF_ERR_HIP */                                                  \
    }                                                                              \
  } while (0)
#define PPF_TRY(expr)      \
  do {                     \
    int r_ = (expr);       \
    if (r_ != 0) return r_; \
  } while (0)

// -------------------------------------------------------- scratch arena
// While a thread has an arena set (ArenaScope), DevBuf::ensure and
// TBuf::alloc carve 256-byte aligned pieces of the arena's device memory in
// order instead of allocating; the pieces are never freed, the arena is reset
// as a whole.  Used by the allocation-free lookup entry point.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
};
Arena *&current_arena();  // thread-local; nullptr when no arena is active
void *arena_take(size_t n);  // nullptr (and an error message) when the arena is full

// -------------------------------------------------------- device buffers
// A growable device buffer owned by whoever holds it; never shrinks.  Inside
// an arena it is a piece of the arena (arena = true) and release() only
// forgets it.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool arena = false;
  int ensure(size_t n);  // grow to at least n bytes (contents not kept)
  void release();
  template <class T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// ------------------------------------------------------------ radix sort
// Stable LSD radix sort of 64-bit keys (optionally carrying 32-bit values) on
// bits [bit_lo, bit_hi), 8-bit digits.  Result ends in keys/vals (the input
// buffers); tmp_keys/tmp_vals are ping-pong space of the same length.
struct SortScratch {
  DevBuf counts, scan, tmp_keys, tmp_vals;
  DevBuf rank, rank_tab;  // sort_key32_ranked's key table and perfect hash
  int tile_items = 16;   // keys per thread of a one-sweep tile (8 or 16; tuned)
  int tile_block = 256;  // threads per one-sweep tile (256 or 512; tuned)
};
int radix_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi,
                   SortScratch &s, hipStream_t st);
// radix_sort_u64 (n < 2^30) that can be queued back to back and captured in
// a HIP graph: no allocation, no memset (fill kernels) and no host
// synchronization.  radix_sort_async_prepare (outside any capture) sizes the
// scratch for sorts of up to n keys (with values if `vals`) and clears the
// look-back time-out flags, which the sorts that follow only set;
// radix_sort_async_check reads them once those sorts have run.
int radix_sort_async_prepare(int64_t n, int bit_lo, int bit_hi, SortScratch &s, bool vals, hipStream_t st);
int radix_sort_u64_async(uint64_t *keys, uint32_t *vals, int64_t n, int bit_lo, int bit_hi, SortScratch &s,
                         hipStream_t st);
int radix_sort_async_check(int64_t n, int bit_lo, int bit_hi, SortScratch &s, hipStream_t st);
// p[0..n) = v by a kernel (a graph node, where a memset node is not wanted)
int fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t st);
// The same result as radix_sort_u64(words, NULL, n, 32, 64, ...) -- a stable
// sort by the high 32 bits -- in two passes over the ranks of the distinct
// high halves when there are at most 8,192 of them (*done = true); otherwise
// (or for fewer than 2^20 words) nothing is done and *done = false.
int sort_key32_ranked(uint64_t *words, int64_t n, SortScratch &s, hipStream_t st, bool *done);

// ------------------------------------------------------- scans / reduces
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp, hipStream_t st,
                       uint32_t *total_out_dev);

}  // namespace ppfi

// ----------------------------------------------------------- wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// inclusive prefix sum across the wave
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  int l = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t w = __shfl_up(v, o, 64);
    if (l >= o) v += w;
  }
  return v;
}
