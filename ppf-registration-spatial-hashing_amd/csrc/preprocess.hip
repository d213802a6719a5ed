// preprocess.hip -- the steps before the registration path in the
// reference's alignment driver: voxel-grid downsampling of scenes and models
// (pcl::VoxelGrid with leaf (l, l, l), src/alignment.cpp:78-87,259-288) and the
// bounding-box extent that defines d_dist (src/alignment.cpp:246-257).
//
// VoxelGrid (PCL 1.x applyFilter, all fields averaged): over the points with
// finite x, y, z, min_b = floor(min * inv), max_b = floor(max * inv) with
// inv = 1/leaf in float; if the number of voxels of that box exceeds
// INT32_MAX the cloud is returned unchanged; each point's voxel index is
// (i - min_b.x) + (j - min_b.y) * dx + (k - min_b.z) * dx * dy; points are
// sorted by that index and every voxel becomes one point whose fields are the
// float sum of its points divided by their count, voxels in ascending index.
// PCL sorts with std::sort, which leaves the order of points inside a voxel
// unspecified; here it is point order (a stable sort), so sums are
// deterministic and agree with PCL's to float rounding.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ppf_objects.h"

using namespace ppfi;

namespace {

constexpr int VX_BLOCK = 256;

__device__ __forceinline__ bool finite3(float x, float y, float z) {
  return isfinite(x) && isfinite(y) && isfinite(z);
}

// per-block min/max of floor(p * inv) over finite points -> atomics on [6]
__global__ void k_voxel_bounds(const float *__restrict__ pts, int64_t n, float inv,
                               int32_t *__restrict__ bounds) {
  __shared__ int32_t s[6][VX_BLOCK / WAVE];
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    if (!finite3(x, y, z)) continue;
    const float c[3] = {x, y, z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int32_t b = (int32_t)floorf(c[d] * inv);
      lo[d] = min(lo[d], b);
      hi[d] = max(hi[d], b);
    }
  }
  const int lane = threadIdx.x % WAVE, wv = threadIdx.x / WAVE;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    for (int o = WAVE / 2; o > 0; o >>= 1) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], o, WAVE));
      hi[d] = max(hi[d], __shfl_xor(hi[d], o, WAVE));
    }
    if (lane == 0) {
      s[d][wv] = lo[d];
      s[3 + d][wv] = hi[d];
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    int32_t a = INT32_MAX, b = INT32_MIN;
    for (int k = 0; k < VX_BLOCK / WAVE; ++k) {
      a = min(a, s[d][k]);
      b = max(b, s[3 + d][k]);
    }
    atomicMin(&bounds[d], a);
    atomicMax(&bounds[3 + d], b);
  }
}

// words = voxel index << 32 | point index (non-finite points: index ~0u)
__global__ void k_voxel_words(const float *__restrict__ pts, int64_t n, float inv, int32_t bx, int32_t by,
                              int32_t bz, int64_t dx, int64_t dxy, uint64_t *__restrict__ words) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
  uint32_t idx = 0xffffffffu;
  if (finite3(x, y, z)) {
    const int64_t a = (int64_t)(int32_t)floorf(x * inv) - bx;
    const int64_t b = (int64_t)(int32_t)floorf(y * inv) - by;
    const int64_t c = (int64_t)(int32_t)floorf(z * inv) - bz;
    idx = (uint32_t)(a + b * dx + c * dxy);
  }
  words[i] = ((uint64_t)idx << 32) | (uint64_t)i;
}

// one thread per voxel: sum of its points' fields in point order / count
__global__ void k_voxel_centroids(const float *__restrict__ pts, const float *__restrict__ nrm,
                                  const uint32_t *__restrict__ first, const uint32_t *__restrict__ count,
                                  const uint32_t *__restrict__ members, int64_t V,
                                  float *__restrict__ out_pts, float *__restrict__ out_nrm) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint32_t f = first[v], c = count[v];
  float s[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k = 0; k < c; ++k) {
    const uint32_t i = members[f + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      s[d] += pts[3 * (size_t)i + d];
      s[3 + d] += nrm[3 * (size_t)i + d];
    }
  }
  const float cf = (float)c;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    out_pts[3 * v + d] = s[d] / cf;
    out_nrm[3 * v + d] = s[3 + d] / cf;
  }
}

}  // namespace

extern "C" {

int ppf_voxel_downsample(ppf_ctx ctx, const float *points, const float *normals, int64_t n, float leaf,
                         float *points_out, float *normals_out, int64_t *n_out) {
  if (!ctx || !points || !normals || !n_out) {
    set_error("ppf_voxel_downsample: NULL argument");
    return PPF_ERR_ARG;
  }
  if (n < 0 || n >= ((int64_t)1 << 31) || !(leaf > 0.0f) || !std::isfinite(leaf)) {
    set_error("ppf_voxel_downsample: need 0 <= n < 2^31 and a finite leaf > 0");
    return PPF_ERR_ARG;
  }
  *n_out = 0;
  if (n == 0) return PPF_OK;
  PPF_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const float inv = 1.0f / leaf;
  DevBuf dp, dn, dw, meta;
  PPF_TRY(dp.ensure((size_t)n * 12));
  PPF_TRY(dn.ensure((size_t)n * 12));
  PPF_TRY(dw.ensure((size_t)n * 8));
  PPF_TRY(meta.ensure(64));
  PPF_HIP(hipMemcpyAsync(dp.ptr, points, (size_t)n * 12, hipMemcpyHostToDevice, st));
  PPF_HIP(hipMemcpyAsync(dn.ptr, normals, (size_t)n * 12, hipMemcpyHostToDevice, st));
  const int32_t init[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
  PPF_HIP(hipMemcpyAsync(meta.ptr, init, sizeof(init), hipMemcpyHostToDevice, st));
  const int G = (int)std::min<int64_t>((n + VX_BLOCK - 1) / VX_BLOCK, 1024);
  hipLaunchKernelGGL(k_voxel_bounds, dim3(G), dim3(VX_BLOCK), 0, st, dp.as<float>(), n, inv,
                     meta.as<int32_t>());
  PPF_HIP(hipGetLastError());
  int32_t b[6];
  PPF_HIP(hipMemcpyAsync(b, meta.ptr, sizeof(b), hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  if (b[0] > b[3]) return PPF_OK;  // no finite point
  const int64_t dx = (int64_t)b[3] - b[0] + 1, dy = (int64_t)b[4] - b[1] + 1, dz = (int64_t)b[5] - b[2] + 1;
  if (dx * dy * dz > (int64_t)INT32_MAX) {
    // PCL: "Leaf size is too small for the input dataset" -> the input as is
    if (points_out) std::copy(points, points + n * 3, points_out);
    if (normals_out) std::copy(normals, normals + n * 3, normals_out);
    *n_out = n;
    return PPF_OK;
  }
  hipLaunchKernelGGL(k_voxel_words, dim3((unsigned)((n + VX_BLOCK - 1) / VX_BLOCK)), dim3(VX_BLOCK), 0, st,
                     dp.as<float>(), n, inv, b[0], b[1], b[2], dx, dx * dy, dw.as<uint64_t>());
  PPF_HIP(hipGetLastError());
  PPF_TRY(radix_sort_u64(dw.as<uint64_t>(), nullptr, n, 32, 64, ctx->sort, st));
  int64_t U = 0;
  PPF_TRY(ppfk::unique_count(ctx, dw.as<uint64_t>(), n, &U));
  DevBuf keys, first, cnt, members, op, on;
  PPF_TRY(keys.ensure((size_t)U * 4));
  PPF_TRY(first.ensure((size_t)U * 4));
  PPF_TRY(cnt.ensure((size_t)U * 4));
  PPF_TRY(members.ensure((size_t)n * 4));
  PPF_TRY(ppfk::unique_emit(ctx, dw.as<uint64_t>(), n, U, keys.as<uint32_t>(), first.as<uint32_t>(),
                            cnt.as<uint32_t>(), members.as<uint32_t>()));
  // the non-finite points, if any, form the last run (index ~0u)
  uint32_t last_key = 0;
  PPF_HIP(hipMemcpyAsync(&last_key, keys.as<uint32_t>() + U - 1, 4, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  const int64_t V = last_key == 0xffffffffu ? U - 1 : U;
  *n_out = V;
  if (V == 0 || (!points_out && !normals_out)) return PPF_OK;
  PPF_TRY(op.ensure((size_t)V * 12));
  PPF_TRY(on.ensure((size_t)V * 12));
  hipLaunchKernelGGL(k_voxel_centroids, dim3((unsigned)((V + VX_BLOCK - 1) / VX_BLOCK)), dim3(VX_BLOCK), 0,
                     st, dp.as<float>(), dn.as<float>(), first.as<uint32_t>(), cnt.as<uint32_t>(),
                     members.as<uint32_t>(), V, op.as<float>(), on.as<float>());
  PPF_HIP(hipGetLastError());
  if (points_out) PPF_HIP(hipMemcpyAsync(points_out, op.ptr, (size_t)V * 12, hipMemcpyDeviceToHost, st));
  if (normals_out) PPF_HIP(hipMemcpyAsync(normals_out, on.ptr, (size_t)V * 12, hipMemcpyDeviceToHost, st));
  PPF_HIP(hipStreamSynchronize(st));
  return PPF_OK;
}

int ppf_bbox_extent(const float *points, int64_t n, float extent_out[3]) {
  if (!points || !extent_out || n < 0) {
    set_error("ppf_bbox_extent: NULL argument or n < 0");
    return PPF_ERR_ARG;
  }
  // pcl::getMinMax3D over the finite points (alignment.cpp:246-257)
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool any = false;
  for (int64_t i = 0; i < n; ++i) {
    const float *p = points + 3 * i;
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) continue;
    any = true;
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], p[d]);
      hi[d] = std::max(hi[d], p[d]);
    }
  }
  for (int d = 0; d < 3; ++d) extent_out[d] = any ? hi[d] - lo[d] : 0.0f;
  return PPF_OK;
}

}  // extern "C"
