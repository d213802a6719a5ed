// api.hip -- the C ABI of libppf_amd.so (include/ppf_amd.h): objects, stage
// orchestration of Model::ppf_lookup and ppf_registration, and the host greedy
// pose clustering of the reference's --cpu_clustering path.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <utility>

#include "ppf_geom.h"
#include "ppf_objects.h"

using namespace ppfi;

// ------------------------------------------------------------------ errors
namespace ppfi {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

Arena *&current_arena() {
  static thread_local Arena *a = nullptr;
  return a;
}

void *arena_take(size_t n) {
  Arena *a = current_arena();
  const size_t off = (a->used + 255) & ~(size_t)255;
  if (off > a->cap || n > a->cap - off) {
    set_error("device scratch exhausted: this call needs more than the " + std::to_string(a->cap >> 20) +
              " MiB of the allocation-free scratch (" + std::to_string(off + n) + " bytes)");
    return nullptr;
  }
  a->used = off + n;
  return a->base + off;
}

int DevBuf::ensure(size_t n) {
  if (n <= bytes && ptr) return 0;
  if (current_arena()) {
    if (ptr && !arena) {
      set_error("internal: an allocated buffer used inside the scratch arena");
      return PPF_ERR_STATE;
    }
    void *p = arena_take(std::max<size_t>(n, 256));
    if (!p) return PPF_ERR_NOMEM;
    ptr = p;
    bytes = std::max<size_t>(n, 256);
    arena = true;
    return 0;
  }
  release();
  size_t want = std::max<size_t>(n, 256);
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
    return PPF_ERR_NOMEM;
  }
  bytes = want;
  return 0;
}
void DevBuf::release() {
  if (ptr && !arena) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
  arena = false;
}
}  // namespace ppfi

int TBuf::alloc(ppf_ctx c, size_t n) {
  release();
  view = false;
  ctx = c->root;
  size_t want = std::max<size_t>(n, 256);
  if (current_arena()) {
    ptr = arena_take(want);
    if (!ptr) return PPF_ERR_NOMEM;
    bytes = want;
    arena = true;
    return 0;
  }
  if (!ctx->reserve(want)) {
    set_error("device memory budget exceeded (" + std::to_string(ctx->bytes.load() + want) + " > " +
              std::to_string(ctx->budget) + " bytes)");
    return PPF_ERR_NOMEM;
  }
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    ctx->unreserve(want);
    set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
    return PPF_ERR_NOMEM;
  }
  bytes = want;
  return 0;
}
void TBuf::release() {
  if (ptr && !arena && !view) {
    (void)hipFree(ptr);
    if (ctx) ctx->unreserve(bytes);
  }
  ptr = nullptr;
  bytes = 0;
  arena = false;
  view = false;
}

namespace ppfi {
int ctx_init(ppf_ctx_s *c, int device) {
  c->device = device;
  hipError_t e = hipSetDevice(c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, c->device);
  int lds = 0;
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device);
  c->max_lds = (size_t)lds;
  for (hipEvent_t *ev : {&c->ev0, &c->ev1, &c->ev2, &c->ev3, &c->ev4, &c->ev5, &c->ev6})
    if (e == hipSuccess) e = hipEventCreate(ev);
  if (e != hipSuccess) {
    set_error(std::string("context set-up: ") + hipGetErrorString(e));
    return PPF_ERR_HIP;
  }
  return 0;
}

void ctx_release(ppf_ctx_s *c) {
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (ppf_ctx_s *l : c->lanes) {
    ctx_release(l);
    delete l;
  }
  c->lanes.clear();
  if (c->noalloc) {
    ctx_release(c->noalloc);
    delete c->noalloc;
    c->noalloc = nullptr;
  }
  ppfk::release_vote_graphs(c);
  c->sort.counts.release();
  c->sort.scan.release();
  c->sort.tmp_keys.release();
  c->sort.tmp_vals.release();
  c->sort.rank.release();
  c->sort.rank_tab.release();
  for (DevBuf *b : {&c->tmp_a, &c->tmp_b, &c->tmp_c, &c->tmp_d, &c->hits, &c->queue, &c->sel, &c->fin_meta,
                    &c->fin_flags, &c->fin_keys, &c->fin_cl, &c->fin_chunk, &c->fin_order, &c->fin_scan,
                    &c->fin_kd, &c->tmp_e})
    b->release();
  for (hipEvent_t *ev : {&c->ev0, &c->ev1, &c->ev2, &c->ev3, &c->ev4, &c->ev5, &c->ev6})
    if (*ev) {
      (void)hipEventDestroy(*ev);
      *ev = nullptr;
    }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

int ctx_lanes(ppf_ctx_s *root, int n, std::vector<ppf_ctx_s *> &out) {
  while ((int)root->lanes.size() < n) {
    auto *l = new ppf_ctx_s();
    l->root = root;
    int rc = ctx_init(l, root->device);
    if (rc) {
      ctx_release(l);
      delete l;
      return rc;
    }
    l->sort.tile_items = root->sort.tile_items;
    l->sort.tile_block = root->sort.tile_block;
    root->lanes.push_back(l);
  }
  out.assign(root->lanes.begin(), root->lanes.begin() + n);
  return 0;
}
}  // namespace ppfi

#define ARG_CHECK(cond, msg)  \
  do {                        \
    if (!(cond)) {            \
      set_error(msg);         \
      return PPF_ERR_ARG;     \
    }                         \
  } while (0)

#define HIPC(expr)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) {                                                                \
      set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                        \
      return PPF_ERR_HIP;                                                                  \
    }                                                                                      \
  } while (0)
#define TRY(expr)                         \
  do {                                    \
    int r_ = (expr);                      \
    if (r_ != 0) return r_;               \
  } while (0)

namespace {

int set_device(ppf_ctx c) {
  HIPC(hipSetDevice(c->device));
  return 0;
}

// Copy an AoS host cloud to the device and build SoA + frames.
int upload_cloud(ppf_ctx ctx, const float *pts, const float *nrm, int64_t n, CloudDev &cloud,
                 TBuf *frames, uint32_t rpdf, int64_t n_frames) {
  cloud.n = n;
  TRY(cloud.soa.alloc(ctx, (size_t)n * 7 * 4));
  if (frames) TRY(frames->alloc(ctx, (size_t)std::max<int64_t>(n_frames, 1) * 64));
  TRY(ctx->tmp_a.ensure((size_t)n * 24));
  float *d = ctx->tmp_a.as<float>();
  HIPC(hipMemcpyAsync(d, pts, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(d + 3 * n, nrm, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
  TRY(ppfk::prep_cloud(ctx, d, d + 3 * n, n, cloud.soa.as<float>(),
                       frames ? frames->as<float>() : nullptr, rpdf));
  HIPC(hipStreamSynchronize(ctx->stream));
  return 0;
}

uint32_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return (uint32_t)p;
}

// ---------------------------------------------------------- CPU clustering
// Greedy clustering of transformation_clustering.cpp:62-123 with the pose test
// of :127-137.  Deviations (documented): stable sorts instead of std::sort,
// rotation() taken as the linear part, rotation angle 2*atan2(|q.vec|, |q.w|).
struct HostPose {
  float R[3][3];
  float t[3];
  uint32_t votes;
};

void quat_of(const float R[3][3], float &x, float &y, float &z, float &w) {
  // Shepperd's method: trace branch, else the largest diagonal element.
  float t = R[0][0] + R[1][1] + R[2][2];
  if (t > 0) {
    t = sqrtf(t + 1.0f);
    w = 0.5f * t;
    t = 0.5f / t;
    x = (R[2][1] - R[1][2]) * t;
    y = (R[0][2] - R[2][0]) * t;
    z = (R[1][0] - R[0][1]) * t;
    return;
  }
  int i = 0;
  if (R[1][1] > R[0][0]) i = 1;
  if (R[2][2] > R[i][i]) i = 2;
  int j = (i + 1) % 3, k = (j + 1) % 3;
  float c[3];
  t = sqrtf(R[i][i] - R[j][j] - R[k][k] + 1.0f);
  c[i] = 0.5f * t;
  t = 0.5f / t;
  w = (R[k][j] - R[j][k]) * t;
  c[j] = (R[j][i] + R[i][j]) * t;
  c[k] = (R[k][i] + R[i][k]) * t;
  x = c[0];
  y = c[1];
  z = c[2];
}

float angle_between(const float A[3][3], const float B[3][3]) {
  float R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = A[0][i] * B[0][j] + A[1][i] * B[1][j] + A[2][i] * B[2][j];
  float x, y, z, w;
  quat_of(R, x, y, z, w);
  const float n = sqrtf(x * x + y * y + z * z);
  return 2.0f * ppfm::atan2f_own(n, fabsf(w));
}

bool within_bounds(const HostPose &a, const HostPose &b, float tt, float rt) {
  const float dx = a.t[0] - b.t[0], dy = a.t[1] - b.t[1], dz = a.t[2] - b.t[2];
  const float pd = sqrtf(dx * dx + dy * dy + dz * dz);
  const float ang = fabsf(angle_between(a.R, b.R));
  return pd < tt && ang < rt;
}

// Up to three clusters, most votes first: cluster vote sums, and per cluster
// the translation mean and the normalised mean of the quaternion coefficients
// (transformation_clustering.cpp:96-122).
void greedy_cluster(const std::vector<float> &T, const std::vector<uint32_t> &counts, float tt,
                    float rt, float poses_out[3][16], uint32_t votes_out[3], int32_t *n_out,
                    int64_t *n_clusters) {
  const size_t P = counts.size();
  std::vector<HostPose> poses(P);
  for (size_t i = 0; i < P; ++i) {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) poses[i].R[r][c] = T[16 * i + 4 * r + c];
      poses[i].t[r] = T[16 * i + 4 * r + 3];
    }
    poses[i].votes = counts[i];
  }
  std::stable_sort(poses.begin(), poses.end(),
                   [](const HostPose &a, const HostPose &b) { return a.votes > b.votes; });
  std::vector<size_t> first;            // first pose of each cluster
  std::vector<std::vector<size_t>> members;
  std::vector<std::pair<size_t, uint32_t>> votes;
  for (size_t i = 0; i < P; ++i) {
    bool found = false;
    for (size_t c = 0; c < first.size(); ++c) {
      if (within_bounds(poses[i], poses[first[c]], tt, rt)) {
        members[c].push_back(i);
        votes[c].second += poses[i].votes;
        found = true;
        break;
      }
    }
    if (!found) {
      first.push_back(i);
      members.push_back({i});
      votes.push_back({first.size() - 1, poses[i].votes});
    }
  }
  std::stable_sort(votes.begin(), votes.end(),
                   [](const std::pair<size_t, uint32_t> &a, const std::pair<size_t, uint32_t> &b) {
                     return a.second > b.second;
                   });
  *n_clusters = (int64_t)first.size();
  const int32_t nres = (int32_t)std::min<size_t>(3, first.size());
  *n_out = nres;
  for (int32_t ci = 0; ci < nres; ++ci) {
    const auto &cl = members[votes[(size_t)ci].first];
    float ta[3] = {0, 0, 0}, qa[4] = {0, 0, 0, 0};
    for (size_t pi : cl) {
      for (int r = 0; r < 3; ++r) ta[r] += poses[pi].t[r];
      float x, y, z, w;
      quat_of(poses[pi].R, x, y, z, w);
      qa[0] += x;
      qa[1] += y;
      qa[2] += z;
      qa[3] += w;
    }
    const float sz = (float)cl.size();
    for (int r = 0; r < 3; ++r) ta[r] /= sz;
    for (int r = 0; r < 4; ++r) qa[r] /= sz;
    const float qn = sqrtf(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2] + qa[3] * qa[3]);
    const float x = qa[0] / qn, y = qa[1] / qn, z = qa[2] / qn, w = qa[3] / qn;
    const float tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const float twx = tx * w, twy = ty * w, twz = tz * w;
    const float txx = tx * x, txy = ty * x, txz = tz * x;
    const float tyy = ty * y, tyz = tz * y, tzz = tz * z;
    float R[3][3] = {{1 - (tyy + tzz), txy - twz, txz + twy},
                     {txy + twz, 1 - (txx + tzz), tyz - twx},
                     {txz - twy, tyz + twx, 1 - (txx + tyy)}};
    float *po = poses_out[ci];
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) po[4 * r + c] = R[r][c];
      po[4 * r + 3] = ta[r];
    }
    po[12] = po[13] = po[14] = 0.0f;
    po[15] = 1.0f;
    votes_out[ci] = votes[(size_t)ci].second;
  }
}


}  // namespace

namespace ppfi {
int scene_create(ppf_ctx lane, const float *points, const float *normals, int64_t n, uint32_t rpdf,
                 ppf_scene *out) {
  auto *s = new ppf_scene_s();
  s->ctx = lane->root;
  s->rpdf = rpdf;
  s->R = (n + rpdf - 1) / rpdf;
  int rc = upload_cloud(lane, points, normals, n, s->cloud, &s->frames, rpdf, s->R);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return PPF_OK;
}
}  // namespace ppfi

namespace ppfi {
ppfk::VoteJob vote_job(ppf_model m, ppf_scene s, float thr, unsigned long long *cand_code,
                       uint32_t *cand_count, uint32_t *meta, uint32_t cap) {
  ppfk::VoteJob J{};
  J.sx = s->cloud.x();
  J.sy = s->cloud.y();
  J.sz = s->cloud.z();
  J.snx = s->cloud.nx();
  J.sny = s->cloud.ny();
  J.snz = s->cloud.nz();
  J.snn = s->cloud.nn();
  J.sframe = s->frames.as<float>();
  J.n_s = (int32_t)s->cloud.n;
  J.rpdf = (int32_t)s->rpdf;
  J.R = (int32_t)s->R;
  J.r_begin = 0;
  J.htab = m->htab.as<uint32_t>();
  J.hshift = m->hshift;
  J.hmask = m->hcap - 1;
  J.toff = m->toff.as<uint32_t>();
  J.rec = m->rec.as<uint32_t>();
  J.pair = m->vpair.as<uint32_t>();
  J.mframe = m->frames.as<float>();
  J.mx = m->cloud.x();
  J.my = m->cloud.y();
  J.mz = m->cloud.z();
  J.n_m = (int32_t)m->cloud.n;
  J.TM = m->TM;
  J.T = m->T;
  J.acc_pitch = m->acc_pitch;
  J.d_dist = m->d_dist;
  J.thr = thr;
  J.max_count = meta + 1;
  J.cand_code = cand_code;
  J.cand_count = cand_count;
  J.cand_n = meta;
  J.cand_cap = cap;
  J.votes = reinterpret_cast<unsigned long long *>(meta + 2);
  J.item_begin = 0;
  J.per_ref = 0;
  J.ungrouped = nullptr;
  J.group_cap = m->ctx->root->dbg_group_cap;
  return J;
}
}  // namespace ppfi

namespace ppfi {
// Everything after voting for one (model, scene) pair, on one context or
// lane: the final threshold and order of the candidates (model.cu:148-170),
// one pose per kept vote (model.cu:191-200), then GPU spatial-hash or host
// greedy clustering and the final pose (model.cu:202-266, ppf.cu:74-93).
// Records ev4 (start), ev5, ev6 and ev3 (end) on the context's stream and
// fills the timings and outputs of r; r->info vote fields are the caller's.
int finish_lookup(ppf_ctx ctx, ppf_model m, ppf_scene s, const ppf_lookup_params *p,
                  const unsigned long long *ccode, const uint32_t *ccount, uint32_t n_cand,
                  uint32_t max_count, ppf_result_s *r) {
  hipStream_t st = ctx->stream;
  r->ctx = ctx->root;
  r->n_m = m->cloud.n;
  r->cpu = p->cpu_clustering != 0;
  r->info.max_idx = -1;
  // a buffer of the context: a freed device buffer would synchronize the
  // whole device, and lanes finish lookups concurrently
  const size_t sel_n = std::max<uint32_t>(n_cand, 1);
  TRY(ctx->sel.ensure(sel_n * 12));
  unsigned long long *scode = ctx->sel.as<unsigned long long>();
  uint32_t *scount = reinterpret_cast<uint32_t *>(scode + sel_n);
  int64_t P = 0;
  HIPC(hipEventRecord(ctx->ev4, st));
  {
    TraceRange tr("ppf.select");
    TRY(ppfk::select_and_sort(ctx, ccode, ccount, n_cand, max_count, p->vote_count_threshold,
                              s->cloud.n, m->cloud.n, scode, scount, &P));
  }
  HIPC(hipEventRecord(ctx->ev5, st));
  // progress reports of plain lookups only (lanes finish batched jobs concurrently)
  const ppf_progress_fn prog = ctx == ctx->root ? ctx->root->progress : nullptr;
  if (prog) prog(ctx->root->progress_user, "select", P, P);
  r->P = P;
  r->info.n_kept = P;
  if (P > 0) {
    TraceRange tr("ppf.poses_and_cluster");
    TRY(r->codes.alloc(ctx, (size_t)P * 8));
    TRY(r->counts.alloc(ctx, (size_t)P * 4));
    TRY(r->T.alloc(ctx, (size_t)P * 64));
    HIPC(hipMemcpyAsync(r->codes.ptr, scode, (size_t)P * 8, hipMemcpyDeviceToDevice, st));
    HIPC(hipMemcpyAsync(r->counts.ptr, scount, (size_t)P * 4, hipMemcpyDeviceToDevice, st));
    TRY(ppfk::compute_poses(ctx, r->codes.as<unsigned long long>(), P, m->frames.as<float>(),
                            s->frames.as<float>(), s->rpdf, r->T.as<float>()));
    HIPC(hipEventRecord(ctx->ev6, st));
    if (prog) {
      HIPC(hipEventSynchronize(ctx->ev6));
      prog(ctx->root->progress_user, "poses", P, P);
    }
    if (p->cpu_clustering) {
      std::vector<float> T((size_t)P * 16);
      std::vector<uint32_t> counts((size_t)P);
      HIPC(hipMemcpyAsync(T.data(), r->T.ptr, (size_t)P * 64, hipMemcpyDeviceToHost, st));
      HIPC(hipMemcpyAsync(counts.data(), r->counts.ptr, (size_t)P * 4, hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      greedy_cluster(T, counts, m->d_dist, ppfm::D_ANGLE0, r->cpu_poses, r->cpu_votes, &r->n_cpu,
                     &r->info.n_clusters);
      std::memcpy(r->pose, r->cpu_poses[0], 64);
    } else {
      TRY(r->vc.alloc(ctx, (size_t)P * 4));
      TRY(r->trans.alloc(ctx, (size_t)P * 12));
      TRY(r->quats.alloc(ctx, (size_t)P * 16));
      int64_t mi = -1;
      TRY(ppfk::cluster_spatial(ctx, r->T.as<float>(), r->codes.as<unsigned long long>(),
                                r->counts.as<uint32_t>(), m->weights.as<float>(), P, m->d_dist,
                                p->use_l1_norm != 0, p->use_averaged_clusters != 0, r->vc.as<float>(),
                                r->trans.as<float>(), r->quats.as<float>(), &mi,
                                &r->info.cluster_tests, &r->info.cluster_mixed_chunks,
                                &r->info.cluster_kernel_ms));
      r->info.max_idx = mi;
      float t3[3];
      HIPC(hipMemcpyAsync(r->pose, r->T.as<float>() + mi * 16, 64, hipMemcpyDeviceToHost, st));
      HIPC(hipMemcpyAsync(t3, r->trans.as<float>() + mi * 3, 12, hipMemcpyDeviceToHost, st));
      HIPC(hipStreamSynchronize(st));
      // src/cuda/ppf.cu:78-93: translation column from transformation_trans
      r->pose[3] = t3[0];
      r->pose[7] = t3[1];
      r->pose[11] = t3[2];
    }
    r->info.status = PPF_OK;
    if (prog) prog(ctx->root->progress_user, "cluster", P, P);
  } else {
    r->info.status = PPF_ERR_NO_VOTES;
  }
  HIPC(hipEventRecord(ctx->ev3, st));
  HIPC(hipEventSynchronize(ctx->ev3));
  float tm = 0.0f;
  HIPC(hipEventElapsedTime(&tm, ctx->ev4, ctx->ev5));
  r->info.select_ms = tm;
  if (P > 0) {
    HIPC(hipEventElapsedTime(&tm, ctx->ev5, ctx->ev6));
    r->info.pose_ms = tm;
    HIPC(hipEventElapsedTime(&tm, ctx->ev6, ctx->ev3));
    r->info.cluster_ms = tm;
  }
  return 0;
}
}  // namespace ppfi

// ---------------------------------------------------------- model files
// Little-endian file: a 128-byte header, then the model's arrays in a fixed
// order, with a 64-bit FNV-1a checksum in the header over the header itself
// (checksum field zero) and then the arrays.  The
// arrays are this build's device layout (version 1), so a loaded model is
// used as is, without rebuilding its hash array.
namespace {

constexpr char MODEL_MAGIC[8] = {'P', 'P', 'F', 'M', 'O', 'D', 'E', 'L'};
constexpr uint32_t MODEL_VERSION = 2;

// accumulator pitch of a tile of TM points: 2^k >= max(TM, 64); files written
// with 2^k + 1 (same records, more LDS bank conflicts) load as they are
bool pitch_ok(int32_t TM, int32_t pitch) {
  int32_t pw = 64;
  while (pw < TM) pw <<= 1;
  return pitch == pw || pitch == pw + 1;
}

struct ModelFileHeader {
  char magic[8];
  uint32_t version, header_bytes;
  int64_t n_points;
  float d_dist;
  int32_t n_angle, TM, T, acc_pitch;
  int32_t pad0;
  int64_t U, N;
  uint32_t hcap, hshift;
  uint64_t payload_bytes, checksum;
  uint8_t reserved[128 - 88];
};
static_assert(sizeof(ModelFileHeader) == 128, "model file header");

uint64_t fnv1a64(uint64_t h, const void *data, size_t n) {
  const unsigned char *p = static_cast<const unsigned char *>(data);
  for (size_t i = 0; i < n; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

// (buffer, bytes) of every saved array, in file order
std::vector<std::pair<TBuf *, size_t>> model_arrays(ppf_model m) {
  const size_t n = (size_t)m->cloud.n, N = n * n, U = (size_t)m->U;
  return {{&m->cloud.soa, n * 7 * 4}, {&m->frames, n * 64},           {&m->hkeys, U * 4},
          {&m->hcount, U * 4},        {&m->hfirst, U * 4},            {&m->pair, N * 4},
          {&m->vpair, N * 4},         {&m->rec, N * 4},               {&m->toff, U * (m->T + 1) * 4},
          {&m->htab, (size_t)m->hcap * 8}, {&m->weights, n * 4}};
}

}  // namespace

extern "C" {

int ppf_model_save(ppf_model m, const char *path) {
  ARG_CHECK(m && path, "ppf_model_save: NULL argument");
  TRY(set_device(m->ctx));
  HIPC(hipStreamSynchronize(m->ctx->stream));
  ModelFileHeader h{};
  std::memcpy(h.magic, MODEL_MAGIC, 8);
  h.version = MODEL_VERSION;
  h.header_bytes = sizeof(ModelFileHeader);
  h.n_points = m->cloud.n;
  h.d_dist = m->d_dist;
  h.n_angle = 30;
  h.TM = m->TM;
  h.T = m->T;
  h.acc_pitch = m->acc_pitch;
  h.U = m->U;
  h.N = m->cloud.n * m->cloud.n;
  h.hcap = m->hcap;
  h.hshift = m->hshift;
  std::vector<std::vector<char>> host;
  uint64_t bytes = 0;
  for (auto &a : model_arrays(m)) {
    host.emplace_back(a.second);
    if (a.second) HIPC(hipMemcpy(host.back().data(), a.first->ptr, a.second, hipMemcpyDeviceToHost));
    bytes += a.second;
  }
  h.payload_bytes = bytes;
  // the checksum covers the header (checksum field zero) and then the arrays
  h.checksum = 0;
  uint64_t ck = fnv1a64(14695981039346656037ull, &h, sizeof h);
  for (auto &b : host) ck = fnv1a64(ck, b.data(), b.size());
  h.checksum = ck;
  FILE *f = std::fopen(path, "wb");
  ARG_CHECK(f, std::string("ppf_model_save: cannot open ") + path);
  bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
  for (auto &b : host) ok = ok && (b.empty() || std::fwrite(b.data(), 1, b.size(), f) == b.size());
  ok = (std::fclose(f) == 0) && ok;
  ARG_CHECK(ok, std::string("ppf_model_save: write failed: ") + path);
  return PPF_OK;
}

int ppf_model_load(ppf_ctx ctx, const char *path, ppf_model *out) {
  ARG_CHECK(ctx && path && out, "ppf_model_load: NULL argument");
  *out = nullptr;
  FILE *f = std::fopen(path, "rb");
  ARG_CHECK(f, std::string("ppf_model_load: cannot open ") + path);
  std::unique_ptr<FILE, int (*)(FILE *)> fc(f, std::fclose);
  ModelFileHeader h{};
  ARG_CHECK(std::fread(&h, sizeof h, 1, f) == 1 && std::memcmp(h.magic, MODEL_MAGIC, 8) == 0,
            "ppf_model_load: not a model file");
  ARG_CHECK(h.version == MODEL_VERSION && h.header_bytes == sizeof(ModelFileHeader),
            "ppf_model_load: unsupported model file version " + std::to_string(h.version));
  ARG_CHECK(h.n_angle == 30, "ppf_model_load: model quantized with another angle count");
  ARG_CHECK(h.n_points >= 2 && h.n_points <= PPF_MAX_MODEL_POINTS && h.N == h.n_points * h.n_points &&
                h.U >= 1 && h.U <= h.N && h.T >= 1 && h.TM >= 1 && h.TM <= MAX_TILE_POINTS &&
                (int64_t)h.T == (h.n_points + h.TM - 1) / h.TM && h.hcap >= 2 * h.U && (h.hcap & (h.hcap - 1)) == 0 &&
                pitch_ok(h.TM, h.acc_pitch) && h.d_dist > 0.0f && h.hshift >= 1 && h.hshift <= 31 &&
                (1ull << (32 - h.hshift)) == h.hcap,
            "ppf_model_load: inconsistent model file header");
  const size_t lds_points = (ctx->max_lds - VOTE_LDS_RESERVE) / (N_BINS * 4) / 32 * 32;
  ARG_CHECK((size_t)h.TM <= lds_points, "ppf_model_load: model tiles exceed this device's LDS");
  TRY(set_device(ctx));
  std::unique_ptr<ppf_model_s> m(new ppf_model_s());
  m->ctx = ctx->root;
  m->cloud.n = h.n_points;
  m->d_dist = h.d_dist;
  m->TM = h.TM;
  m->T = h.T;
  m->acc_pitch = h.acc_pitch;
  m->U = h.U;
  m->hcap = h.hcap;
  m->hshift = h.hshift;
  ModelFileHeader h0 = h;
  h0.checksum = 0;
  uint64_t ck = fnv1a64(14695981039346656037ull, &h0, sizeof h0), bytes = 0;
  std::vector<char> buf;
  for (auto &a : model_arrays(m.get())) {
    buf.resize(a.second);
    ARG_CHECK(a.second == 0 || std::fread(buf.data(), 1, a.second, f) == a.second,
              "ppf_model_load: file ends early");
    ck = fnv1a64(ck, buf.data(), a.second);
    bytes += a.second;
    TRY(a.first->alloc(ctx, a.second));
    HIPC(hipMemcpy(a.first->ptr, buf.data(), a.second, hipMemcpyHostToDevice));
  }
  ARG_CHECK(bytes == h.payload_bytes && ck == h.checksum, "ppf_model_load: checksum mismatch");
  *out = m.release();
  return PPF_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free lookup
// Device scratch of ppf_lookup_noalloc: one static array per device (each GPU
// holds its own copy until the process ends).  In host code `scratch` names a
// host-side stand-in, never the device memory: host code uses only the
// address hipGetSymbolAddress returns for the current device, kept per device
// in g_scratch_dev.
static __device__ __align__(256) unsigned char scratch[PPF_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[MAX_DEVICES] = {};
// Share of the scratch the vote blocks' hit lists may take; the grid is
// capped so that they fit.
constexpr size_t NOALLOC_HIT_BYTES = PPF_NOALLOC_SCRATCH_BYTES * 3 / 8;
constexpr uint32_t NOALLOC_CAND_CAP = 1u << 20;

// Set-up at context creation, with the context's device current: the address
// of this device's scratch array (once per device) and the context's
// allocation-free lane (its own stream and events; its scratch buffers are
// only ever pieces of the array).
int noalloc_setup(ppf_ctx_s *c) {
  ARG_CHECK(c->device >= 0 && c->device < MAX_DEVICES, "ppf_ctx_create: device index too large");
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
      void *ptr = nullptr;
      HIPC(hipSetDevice(c->device));
      HIPC(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      ARG_CHECK(ptr != nullptr, "ppf_ctx_create: no address for the device scratch array");
      g_scratch_dev[c->device] = ptr;
    }
  }
  auto *l = new ppf_ctx_s();
  l->root = c;
  c->noalloc = l;
  return ctx_init(l, c->device);
}

// Makes the device's scratch array the calling thread's arena for one call
// and forgets the lane's pieces of it before and after.
struct ArenaScope {
  ppfi::Arena arena;
  ppf_ctx_s *lane;
  static void forget(ppf_ctx_s *l) {
    for (DevBuf *b : {&l->tmp_a, &l->tmp_b, &l->tmp_c, &l->tmp_d, &l->hits, &l->queue, &l->sel,
                      &l->sort.counts, &l->sort.scan, &l->sort.tmp_keys, &l->sort.tmp_vals})
      b->release();
  }
  ArenaScope(void *base, ppf_ctx_s *l) : lane(l) {
    arena.base = static_cast<char *>(base);
    arena.cap = PPF_NOALLOC_SCRATCH_BYTES;
    forget(lane);
    ppfi::current_arena() = &arena;
  }
  ~ArenaScope() {
    ppfi::current_arena() = nullptr;
    forget(lane);
  }
};
}  // namespace

extern "C" {

const char *ppf_last_error(void) { return g_err.c_str(); }

void ppf_abi_sizes(uint32_t sizes_out[PPF_ABI_N_STRUCTS]) {
  sizes_out[0] = (uint32_t)sizeof(ppf_model_info_t);
  sizes_out[1] = (uint32_t)sizeof(ppf_lookup_params);
  sizes_out[2] = (uint32_t)sizeof(ppf_result_info_t);
  sizes_out[3] = (uint32_t)sizeof(ppf_batch_info_t);
  sizes_out[4] = (uint32_t)sizeof(ppf_server_config);
  sizes_out[5] = (uint32_t)sizeof(ppf_request_info_t);
  sizes_out[6] = (uint32_t)sizeof(ppf_server_stats_t);
  sizes_out[7] = (uint32_t)sizeof(ppf_tune_info_t);
}

int ppf_ctx_create(int device, uint64_t hbm_budget_bytes, ppf_ctx *out) {
  ARG_CHECK(out, "ppf_ctx_create: out is NULL");
  *out = nullptr;
  int n = 0;
  HIPC(hipGetDeviceCount(&n));
  ARG_CHECK(n > 0, "ppf_ctx_create: no HIP device");
  auto *c = new ppf_ctx_s();
  c->budget = hbm_budget_bytes;
  int rc = ctx_init(c, std::min(n - 1, std::max(0, device)));
  if (rc == 0) rc = noalloc_setup(c);
  if (rc) {
    ctx_release(c);
    delete c;
    return rc;
  }
  *out = c;
  return PPF_OK;
}

int ppf_ctx_destroy(ppf_ctx ctx) {
  if (!ctx) return PPF_OK;
  ARG_CHECK(ctx->root == ctx, "ppf_ctx_destroy: not a context created by ppf_ctx_create");
  ctx_release(ctx);
  delete ctx;
  return PPF_OK;
}

uint64_t ppf_ctx_device_bytes(ppf_ctx ctx) { return ctx ? ctx->root->bytes.load() : 0; }

int ppf_ctx_set_progress(ppf_ctx ctx, ppf_progress_fn fn, void *user, int64_t rows_per_launch) {
  ARG_CHECK(ctx, "ppf_ctx_set_progress: NULL context");
  ARG_CHECK(rows_per_launch >= 0, "ppf_ctx_set_progress: negative rows_per_launch");
  ctx->root->progress = fn;
  ctx->root->progress_user = user;
  ctx->root->progress_rows = rows_per_launch;
  return PPF_OK;
}

// ------------------------------------------------------------------ model
int ppf_model_build(ppf_ctx ctx, const float *points, const float *normals, int64_t n,
                    float d_dist, ppf_model *out) {
  ARG_CHECK(ctx && out, "ppf_model_build: NULL handle");
  *out = nullptr;
  ARG_CHECK(points && normals, "ppf_model_build: NULL cloud");
  ARG_CHECK(n >= 2 && n <= PPF_MAX_MODEL_POINTS, "ppf_model_build: need 2 <= n <= 65535 points");
  ARG_CHECK(d_dist > 0.0f && std::isfinite(d_dist), "ppf_model_build: d_dist must be > 0");
  TRY(set_device(ctx));
  TraceRange tr("ppf.model_build");
  std::unique_ptr<ppf_model_s> m(new ppf_model_s());
  m->ctx = ctx;
  m->d_dist = d_dist;
  hipStream_t st = ctx->stream;
  TRY(upload_cloud(ctx, points, normals, n, m->cloud, &m->frames, 1, n));
  // accumulator tile: TM model points x 31 bins x 4 B must fit the LDS of one block
  const int64_t lds_points = (int64_t)((ctx->max_lds - VOTE_LDS_RESERVE) / (N_BINS * 4)) / 32 * 32;
  m->TM = (int32_t)std::min<int64_t>(n, std::min<int64_t>(DEFAULT_TILE_POINTS, lds_points));
  ARG_CHECK(m->TM >= 32, "ppf_model_build: device LDS too small for the vote accumulator");
  m->T = (int32_t)((n + m->TM - 1) / m->TM);
  const int64_t N = n * n;
  // words = key << 32 | pair index, in pair order; a stable sort on the key
  // bits gives (key, pair) order = the reference's stable sort_by_key
  ppfi::DevBuf words;
  TRY(words.ensure((size_t)N * 8));
  TRY(ppfk::cloud_pair_words(ctx, m->cloud, d_dist, 1, words.as<uint64_t>()));
  {
    // a model has a few thousand distinct keys: sort by their ranks
    bool ranked = false;
    TRY(ppfi::sort_key32_ranked(words.as<uint64_t>(), N, ctx->sort, st, &ranked));
    if (!ranked) TRY(ppfi::radix_sort_u64(words.as<uint64_t>(), nullptr, N, 32, 64, ctx->sort, st));
  }
  int64_t U = 0;
  TRY(ppfk::unique_count(ctx, words.as<uint64_t>(), N, &U));
  m->U = U;
  TRY(m->pair.alloc(ctx, (size_t)N * 4));
  TRY(m->hkeys.alloc(ctx, (size_t)U * 4));
  TRY(m->hfirst.alloc(ctx, (size_t)U * 4));
  TRY(m->hcount.alloc(ctx, (size_t)U * 4));
  TRY(ppfk::unique_emit(ctx, words.as<uint64_t>(), N, U, m->hkeys.as<uint32_t>(),
                        m->hfirst.as<uint32_t>(), m->hcount.as<uint32_t>(), m->pair.as<uint32_t>()));
  // vote order and per-tile sub-ranges
  TRY(m->vpair.alloc(ctx, (size_t)N * 4));
  TRY(m->toff.alloc(ctx, (size_t)U * (m->T + 1) * 4));
  TRY(ppfk::vote_order(ctx, words.as<uint64_t>(), m->vpair.as<uint32_t>(), N, m->cloud, m->TM, m->T,
                       m->hfirst.as<uint32_t>(), m->hcount.as<uint32_t>(), U, m->toff.as<uint32_t>()));
  {
    int32_t pw = 64;
    while (pw < m->TM) pw <<= 1;
    m->acc_pitch = pw;
  }
  HIPC(hipStreamSynchronize(st));
  words.release();
  // key -> unique index table
  m->hcap = std::max<uint32_t>(1024, next_pow2((uint64_t)U * 2));
  uint32_t lg = 0;
  while ((1u << lg) < m->hcap) ++lg;
  m->hshift = 32 - lg;
  TRY(m->htab.alloc(ctx, (size_t)m->hcap * 8));
  TRY(ppfk::build_htable(ctx, m->hkeys.as<uint32_t>(), U, m->htab.as<uint32_t>(), m->hcap, m->hshift));
  // packed vote records
  TRY(m->rec.alloc(ctx, (size_t)N * 4));
  TRY(ppfk::build_records(ctx, m->vpair.as<uint32_t>(), N, m->frames.as<float>(), m->cloud, m->TM,
                          m->rec.as<uint32_t>()));
  TRY(m->weights.alloc(ctx, (size_t)n * 4));
  std::vector<float> ones((size_t)n, 1.0f);
  HIPC(hipMemcpyAsync(m->weights.ptr, ones.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  HIPC(hipStreamSynchronize(st));
  *out = m.release();
  return PPF_OK;
}

int ppf_model_destroy(ppf_model model) {
  if (!model) return PPF_OK;
  (void)hipSetDevice(model->ctx->device);
  (void)hipStreamSynchronize(model->ctx->stream);
  delete model;
  return PPF_OK;
}

int ppf_model_info(ppf_model m, ppf_model_info_t *info) {
  ARG_CHECK(m && info, "ppf_model_info: NULL");
  info->n_points = m->cloud.n;
  info->d_dist = m->d_dist;
  info->n_pairs = m->cloud.n * m->cloud.n;
  info->n_unique_keys = m->U;
  info->tile_points = m->TM;
  info->n_tiles = m->T;
  info->device_bytes = m->cloud.soa.bytes + m->frames.bytes + m->hkeys.bytes + m->hcount.bytes +
                       m->hfirst.bytes + m->pair.bytes + m->vpair.bytes + m->rec.bytes + m->toff.bytes +
                       m->htab.bytes + m->weights.bytes;
  return PPF_OK;
}

int ppf_model_hash_array(ppf_model m, uint32_t *hashkeys, uint32_t *counts, uint32_t *first,
                         uint32_t *key2ppf) {
  ARG_CHECK(m, "ppf_model_hash_array: NULL model");
  TRY(set_device(m->ctx));
  HIPC(hipStreamSynchronize(m->ctx->stream));
  if (hashkeys) HIPC(hipMemcpy(hashkeys, m->hkeys.ptr, (size_t)m->U * 4, hipMemcpyDeviceToHost));
  if (counts) HIPC(hipMemcpy(counts, m->hcount.ptr, (size_t)m->U * 4, hipMemcpyDeviceToHost));
  if (first) HIPC(hipMemcpy(first, m->hfirst.ptr, (size_t)m->U * 4, hipMemcpyDeviceToHost));
  if (key2ppf)
    HIPC(hipMemcpy(key2ppf, m->pair.ptr, (size_t)m->cloud.n * m->cloud.n * 4, hipMemcpyDeviceToHost));
  return PPF_OK;
}

int ppf_model_set_weights(ppf_model m, const float *weights, int64_t n) {
  ARG_CHECK(m && weights, "ppf_model_set_weights: NULL");
  ARG_CHECK(n == m->cloud.n, "ppf_model_set_weights: need one weight per model point");
  TRY(set_device(m->ctx));
  HIPC(hipMemcpy(m->weights.ptr, weights, (size_t)n * 4, hipMemcpyHostToDevice));
  return PPF_OK;
}

// ------------------------------------------------------------------ scene
int ppf_scene_upload(ppf_ctx ctx, const float *points, const float *normals, int64_t n,
                     uint32_t rpdf, ppf_scene *out) {
  ARG_CHECK(ctx && out, "ppf_scene_upload: NULL handle");
  *out = nullptr;
  ARG_CHECK(points && normals, "ppf_scene_upload: NULL cloud");
  ARG_CHECK(n >= 2 && n < ((int64_t)1 << 31), "ppf_scene_upload: need 2 <= n < 2^31 points");
  TraceRange tr("ppf.scene_upload");
  ARG_CHECK(rpdf >= 1, "ppf_scene_upload: ref_point_downsample_factor must be >= 1");
  TRY(set_device(ctx));
  return ppfi::scene_create(ctx, points, normals, n, rpdf, out);
}

int ppf_scene_destroy(ppf_scene scene) {
  if (!scene) return PPF_OK;
  (void)hipSetDevice(scene->ctx->device);
  (void)hipStreamSynchronize(scene->ctx->stream);
  delete scene;
  return PPF_OK;
}

int ppf_scene_keys(ppf_scene s, float d_dist, uint32_t *keys_out) {
  ARG_CHECK(s && keys_out, "ppf_scene_keys: NULL");
  ARG_CHECK(s->cloud.n <= 8192, "ppf_scene_keys: inspection limited to 8192 points");
  ARG_CHECK(d_dist > 0.0f, "ppf_scene_keys: d_dist must be > 0");
  ppf_ctx ctx = s->ctx;
  TRY(set_device(ctx));
  const int64_t N = s->cloud.n * s->cloud.n;
  ppfi::DevBuf k;
  TRY(k.ensure((size_t)N * 4));
  TRY(ppfk::cloud_pair_keys(ctx, s->cloud, d_dist, s->rpdf, k.as<uint32_t>()));
  HIPC(hipMemcpyAsync(keys_out, k.ptr, (size_t)N * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return PPF_OK;
}


// ----------------------------------------------------------------- lookup
}  // extern "C"

namespace {
// Model::ppf_lookup on `ctx` (the model's context, or a lane of it: its
// stream, events and scratch).  fixed_cap > 0: the candidate buffer has
// exactly that capacity (no growth); max_blocks > 0 caps the vote grid.
int lookup_core(ppf_ctx ctx, ppf_model m, ppf_scene s, const ppf_lookup_params *p, uint32_t fixed_cap,
                int max_blocks, ppf_result_s *r) {
  hipStream_t st = ctx->stream;
  TraceRange tr("ppf.lookup");
  r->ctx = ctx->root;
  r->n_m = m->cloud.n;
  r->cpu = p->cpu_clustering != 0;
  HIPC(hipEventRecord(ctx->ev2, st));

  // ---- votes.  Candidates (code, count) are pruned against a running lower
  // bound of the global maximum; on overflow the vote pass is repeated with
  // the final maximum preset, which emits exactly the kept votes.
  ppfi::DevBuf cand;
  // candidates per reference point are bounded by the pruning; start at
  // CAND_PER_REF per reference point (at least 8M), grow on overflow
  uint32_t cap = fixed_cap ? fixed_cap
                 : p->candidate_capacity
                     ? round_cap(p->candidate_capacity)
                     : round_cap((uint64_t)std::max<int64_t>(CAND_CAP_MIN, s->R * (int64_t)CAND_PER_REF));
  uint32_t meta_h[2] = {0, 0};
  int64_t votes = 0;
  float vote_ms = 0.0f;
  for (int attempt = 0;; ++attempt) {
    TraceRange trv("ppf.vote");
    // [meta: 32 B, 256 B reserved][codes: cap x u64][counts: cap x u32].  The
    // counters come first so that the 64-bit vote total stays 8-byte aligned
    // for any cap (a misaligned 64-bit atomic faults).
    TRY(cand.ensure(CAND_HEADER + (size_t)cap * 12));
    // [0] candidates emitted, [1] running maximum, [2..3] votes, [4] rows
    // voted hit by hit
    uint32_t *meta = cand.as<uint32_t>();
    unsigned long long *cc = reinterpret_cast<unsigned long long *>(cand.as<char>() + CAND_HEADER);
    uint32_t *cn = reinterpret_cast<uint32_t *>(cc + cap);
    const uint32_t preset[8] = {0u, meta_h[1], 0u, 0u, 0u, 0u, 0u, 0u};
    HIPC(hipMemcpyAsync(meta, preset, 32, hipMemcpyHostToDevice, st));
    ppfk::VoteJob J = ppfi::vote_job(m, s, p->vote_count_threshold, cc, cn, meta, cap);
    J.per_ref = p->per_reference_argmax != 0;
    J.ungrouped = meta + 4;
    // with progress reports, the rows in launches of progress_rows (the
    // running maximum and the candidates carry over between launches)
    const ppf_ctx_s *root = ctx->root;
    const int64_t step = root->progress && root->progress_rows > 0 ? root->progress_rows : s->R;
    for (int64_t r0 = 0; r0 < s->R; r0 += step) {
      const int64_t nr = std::min<int64_t>(step, s->R - r0);
      J.r_begin = (int32_t)r0;
      float ms = 0.0f;
      TRY(ppfk::run_votes(ctx, &J, 1, (uint64_t)nr, m->TM, s->cloud.n, p->exact_votes != 0, &ms,
                          max_blocks));
      vote_ms += ms;
      if (root->progress) root->progress(root->progress_user, "vote", r0 + nr, s->R);
    }
    uint32_t mh[8];
    HIPC(hipMemcpyAsync(mh, meta, 32, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    meta_h[0] = mh[0];
    meta_h[1] = mh[1];
    votes = (int64_t)(((uint64_t)mh[3] << 32) | mh[2]);
    r->info.vote_rows_ungrouped = mh[4];
    if (meta_h[0] <= cap) break;
    if (attempt >= 2 || meta_h[0] > CAND_CAP_MAX || (fixed_cap && attempt >= 1)) {
      set_error("ppf_lookup: kept-vote buffer overflow (" + std::to_string(meta_h[0]) + " candidates" +
                (fixed_cap ? ", more than the fixed capacity of " + std::to_string(fixed_cap) : "") + ")");
      return PPF_ERR_NOMEM;
    }
    // the rerun starts from the final maximum, so it emits only the kept
    // votes, of which there are at most meta_h[0]: sized so, it needs no
    // third pass (a fixed capacity reruns once at that capacity)
    if (!fixed_cap) cap = round_cap(meta_h[0]);
  }
  const uint32_t n_cand = meta_h[0];
  r->info.lookups = (int64_t)s->R * (s->cloud.n - 1);
  r->info.votes = votes;
  r->info.max_count = meta_h[1];
  r->info.n_candidates = n_cand;
  r->info.vote_ms = vote_ms;
  const unsigned long long *ccode =
      reinterpret_cast<const unsigned long long *>(cand.as<char>() + CAND_HEADER);
  const uint32_t *ccount = reinterpret_cast<const uint32_t *>(ccode + cap);
  TRY(ppfi::finish_lookup(ctx, m, s, p, ccode, ccount, n_cand, meta_h[1], r));
  float tm = 0.0f;
  HIPC(hipEventElapsedTime(&tm, ctx->ev2, ctx->ev3));
  r->info.total_ms = tm;
  return PPF_OK;
}
}  // namespace

extern "C" {

int ppf_lookup(ppf_model m, ppf_scene s, const ppf_lookup_params *p, ppf_result *out) {
  ARG_CHECK(m && s && p && out, "ppf_lookup: NULL argument");
  *out = nullptr;
  ARG_CHECK(m->ctx == s->ctx, "ppf_lookup: model and scene belong to different contexts");
  TRY(set_device(m->ctx));
  std::unique_ptr<ppf_result_s> r(new ppf_result_s());
  TRY(lookup_core(m->ctx, m, s, p, 0, 0, r.get()));
  *out = r.release();
  return PPF_OK;
}

int ppf_lookup_noalloc(ppf_model m, ppf_scene s, const ppf_lookup_params *p, float *pose_out,
                       ppf_result_info_t *info_out) {
  ARG_CHECK(m && s && p && pose_out && info_out, "ppf_lookup_noalloc: NULL argument");
  ARG_CHECK(m->ctx == s->ctx, "ppf_lookup_noalloc: model and scene belong to different contexts");
  ARG_CHECK(s->cloud.n <= PPF_NOALLOC_MAX_SCENE_POINTS,
            "ppf_lookup_noalloc: scene of " + std::to_string(s->cloud.n) + " points exceeds the limit of " +
                std::to_string(PPF_NOALLOC_MAX_SCENE_POINTS) + " (use ppf_lookup)");
  ppf_ctx root = m->ctx->root;
  ARG_CHECK(root->noalloc, "ppf_lookup_noalloc: context has no allocation-free lane");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    base = g_scratch_dev[root->device];
  }
  ARG_CHECK(base, "ppf_lookup_noalloc: device scratch not set up");
  TRY(set_device(root));
  // the model's and scene's uploads (on the root stream) are complete: both
  // constructors synchronize before returning
  const size_t per_block = ppfk::vote_hit_bytes_per_block(s->cloud.n);
  const int max_blocks = (int)std::max<size_t>(1, NOALLOC_HIT_BYTES / per_block);
  int rc;
  {
    ArenaScope scope(base, root->noalloc);
    ppf_result_s r;
    rc = lookup_core(root->noalloc, m, s, p, NOALLOC_CAND_CAP, max_blocks, &r);
    if (rc == PPF_OK) {
      std::memcpy(pose_out, r.pose, 64);
      *info_out = r.info;
    }
    HIPC(hipStreamSynchronize(root->noalloc->stream));
  }
  if (rc) return rc;
  return info_out->status == PPF_OK ? PPF_OK : info_out->status;
}

int ppf_result_destroy(ppf_result r) {
  if (!r) return PPF_OK;
  (void)hipSetDevice(r->ctx->device);
  (void)hipStreamSynchronize(r->ctx->stream);
  delete r;
  return PPF_OK;
}

int ppf_result_info(ppf_result r, ppf_result_info_t *info) {
  ARG_CHECK(r && info, "ppf_result_info: NULL");
  *info = r->info;
  return PPF_OK;
}

int ppf_result_votes(ppf_result r, uint64_t *codes, uint32_t *counts) {
  ARG_CHECK(r, "ppf_result_votes: NULL");
  if (r->P == 0) return PPF_OK;
  TRY(set_device(r->ctx));
  if (codes) HIPC(hipMemcpy(codes, r->codes.ptr, (size_t)r->P * 8, hipMemcpyDeviceToHost));
  if (counts) HIPC(hipMemcpy(counts, r->counts.ptr, (size_t)r->P * 4, hipMemcpyDeviceToHost));
  return PPF_OK;
}

int ppf_result_transforms(ppf_result r, float *T) {
  ARG_CHECK(r && T, "ppf_result_transforms: NULL");
  if (r->P == 0) return PPF_OK;
  TRY(set_device(r->ctx));
  HIPC(hipMemcpy(T, r->T.ptr, (size_t)r->P * 64, hipMemcpyDeviceToHost));
  return PPF_OK;
}

int ppf_result_clusters(ppf_result r, float *vc, float *trans, float *quats) {
  ARG_CHECK(r, "ppf_result_clusters: NULL");
  if (r->P == 0) return PPF_OK;
  if (r->cpu) {
    set_error("ppf_result_clusters: lookup used CPU clustering");
    return PPF_ERR_STATE;
  }
  TRY(set_device(r->ctx));
  if (vc) HIPC(hipMemcpy(vc, r->vc.ptr, (size_t)r->P * 4, hipMemcpyDeviceToHost));
  if (trans) HIPC(hipMemcpy(trans, r->trans.ptr, (size_t)r->P * 12, hipMemcpyDeviceToHost));
  if (quats) HIPC(hipMemcpy(quats, r->quats.ptr, (size_t)r->P * 16, hipMemcpyDeviceToHost));
  return PPF_OK;
}

int ppf_result_cpu_clusters(ppf_result r, float *poses_out, uint32_t *votes_out, int32_t *n_out) {
  ARG_CHECK(r && n_out, "ppf_result_cpu_clusters: NULL");
  if (!r->cpu) {
    set_error("ppf_result_cpu_clusters: lookup used GPU clustering");
    return PPF_ERR_STATE;
  }
  *n_out = r->n_cpu;
  for (int32_t i = 0; i < r->n_cpu; ++i) {
    if (poses_out) std::memcpy(poses_out + 16 * i, r->cpu_poses[i], 64);
    if (votes_out) votes_out[i] = r->cpu_votes[i];
  }
  return PPF_OK;
}

int ppf_result_pose(ppf_result r, float *pose16) {
  ARG_CHECK(r && pose16, "ppf_result_pose: NULL");
  std::memcpy(pose16, r->pose, 64);
  return r->info.status == PPF_OK ? PPF_OK : r->info.status;
}

// ------------------------------------------------------------ registration
int ppf_registration(ppf_ctx ctx, int32_t n_scenes, const float *const *scene_points,
                     const float *const *scene_normals, const int64_t *scene_sizes,
                     int32_t n_models, const float *const *model_points,
                     const float *const *model_normals, const int64_t *model_sizes,
                     const float *model_d_dists, uint32_t rpdf, float vote_count_threshold,
                     int32_t cpu_clustering, int32_t use_l1_norm, int32_t use_averaged_clusters,
                     const float *const *model_weights, float *poses_out, int32_t *status_out) {
  ARG_CHECK(ctx && poses_out, "ppf_registration: NULL argument");
  ARG_CHECK(n_scenes >= 0 && n_models >= 0, "ppf_registration: negative counts");
  ARG_CHECK(n_models == 0 || (model_points && model_normals && model_sizes && model_d_dists),
            "ppf_registration: NULL model arrays");
  ARG_CHECK(n_scenes == 0 || (scene_points && scene_normals && scene_sizes),
            "ppf_registration: NULL scene arrays");
  std::vector<ppf_model> models((size_t)n_models, nullptr);
  auto cleanup = [&]() {
    for (auto mm : models) ppf_model_destroy(mm);
  };
  // Each model is built once (the reference rebuilds it per scene, ppf.cu:69;
  // the hash array does not depend on the scene).
  for (int32_t j = 0; j < n_models; ++j) {
    int rc = ppf_model_build(ctx, model_points[j], model_normals[j], model_sizes[j], model_d_dists[j],
                             &models[j]);
    if (rc) {
      cleanup();
      return rc;
    }
    if (model_weights && model_weights[j]) {
      rc = ppf_model_set_weights(models[j], model_weights[j], model_sizes[j]);
      if (rc) {
        cleanup();
        return rc;
      }
    }
  }
  ppf_lookup_params prm{vote_count_threshold, cpu_clustering, use_l1_norm, use_averaged_clusters, 0};
  // every model against one scene in one batched lookup: one vote launch,
  // the finishing stages overlapped on lanes; each pose is the one
  // ppf_lookup gives for the pair
  std::vector<ppf_scene> sv((size_t)n_models);
  std::vector<ppf_result> res((size_t)n_models, nullptr);
  for (int32_t i = 0; i < n_scenes && n_models > 0; ++i) {
    ppf_scene sc = nullptr;
    int rc = ppf_scene_upload(ctx, scene_points[i], scene_normals[i], scene_sizes[i], rpdf, &sc);
    if (rc) {
      cleanup();
      return rc;
    }
    std::fill(sv.begin(), sv.end(), sc);
    rc = ppf_batch_lookup(ctx, n_models, models.data(), sv.data(), &prm, res.data(), nullptr);
    for (int32_t j = 0; j < n_models && rc == PPF_OK; ++j) {
      float *dst = poses_out + ((size_t)i * n_models + j) * 16;
      const int st = ppf_result_pose(res[(size_t)j], dst);
      if (status_out) status_out[(size_t)i * n_models + j] = st;
    }
    for (auto &r : res) {
      ppf_result_destroy(r);
      r = nullptr;
    }
    ppf_scene_destroy(sc);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  cleanup();
  return PPF_OK;
}

int ppf_debug_numerics(ppf_ctx ctx, int32_t op, const float *a, const float *b, float *out,
                       int64_t n) {
  ARG_CHECK(ctx && a && out, "ppf_debug_numerics: NULL");
  ARG_CHECK(op >= 0 && op <= 8, "ppf_debug_numerics: unknown op");
  TRY(set_device(ctx));
  const int64_t w = op == 8 ? 2 : 1;
  ppfi::DevBuf da, db, dout;
  TRY(da.ensure((size_t)n * w * 4));
  TRY(db.ensure((size_t)n * w * 4));
  TRY(dout.ensure((size_t)n * 4));
  HIPC(hipMemcpy(da.ptr, a, (size_t)n * w * 4, hipMemcpyHostToDevice));
  if (b) HIPC(hipMemcpy(db.ptr, b, (size_t)n * w * 4, hipMemcpyHostToDevice));
  TRY(ppfk::debug_numerics(ctx, op, da.as<float>(), db.as<float>(), dout.as<float>(), n));
  HIPC(hipMemcpyAsync(out, dout.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return PPF_OK;
}

int ppf_debug_set(ppf_ctx ctx, int32_t knob, int64_t value) {
  ARG_CHECK(ctx, "ppf_debug_set: NULL context");
  ARG_CHECK(knob >= 0 && knob <= 4, "ppf_debug_set: unknown knob");
  if (knob == 4) {
    ARG_CHECK(value == 0 || value == 1, "ppf_debug_set: vote graphs must be 0 or 1");
    ctx->root->use_graphs = (int32_t)value;
    return PPF_OK;
  }
  if (knob == 3) {
    ARG_CHECK(value == PPF_FINISH_AUTO || value == PPF_FINISH_LANES, "ppf_debug_set: finish mode must be 0 or 1");
    ctx->root->finish_mode = (int32_t)value;
    return PPF_OK;
  }
  if (knob == 1) {
    ARG_CHECK(value == 8 || value == 16, "ppf_debug_set: sort tile keys per thread must be 8 or 16");
    ctx->root->sort.tile_items = (int)value;
    for (ppf_ctx_s *l : ctx->root->lanes) l->sort.tile_items = (int)value;
    return PPF_OK;
  }
  if (knob == 2) {
    ARG_CHECK(value == 256 || value == 512, "ppf_debug_set: sort tile threads must be 256 or 512");
    ctx->root->sort.tile_block = (int)value;
    for (ppf_ctx_s *l : ctx->root->lanes) l->sort.tile_block = (int)value;
    return PPF_OK;
  }
  ARG_CHECK(value >= 0 && value <= 0xffffffffll, "ppf_debug_set: value out of range");
  ctx->root->dbg_group_cap = (uint32_t)value;
  return PPF_OK;
}

int ppf_debug_cluster(ppf_ctx ctx, const float *T, const uint64_t *codes, const uint32_t *counts,
                      const float *weights, int64_t n_weights, int64_t P, float d_dist, int32_t use_l1,
                      int32_t use_avg, float *vc_out, float *trans_out, float *quats_out,
                      int64_t stats_out[3]) {
  ARG_CHECK(ctx && T && codes && counts && vc_out && trans_out && quats_out && stats_out,
            "ppf_debug_cluster: NULL argument");
  ARG_CHECK(P >= 1 && P < ((int64_t)1 << 31), "ppf_debug_cluster: need 1 <= P < 2^31 poses");
  ARG_CHECK(d_dist > 0.0f, "ppf_debug_cluster: d_dist must be > 0");
  for (int64_t i = 0; i < P; ++i)
    ARG_CHECK(!weights || (int64_t)(((uint32_t)codes[i]) >> 6) < n_weights,
              "ppf_debug_cluster: a code's model point has no weight");
  TRY(set_device(ctx));
  hipStream_t st = ctx->stream;
  ppfi::DevBuf dT, dc, dn, dw, dvc, dtr, dq;
  TRY(dT.ensure((size_t)P * 64));
  TRY(dc.ensure((size_t)P * 8));
  TRY(dn.ensure((size_t)P * 4));
  TRY(dvc.ensure((size_t)P * 4));
  TRY(dtr.ensure((size_t)P * 12));
  TRY(dq.ensure((size_t)P * 16));
  HIPC(hipMemcpy(dT.ptr, T, (size_t)P * 64, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dc.ptr, codes, (size_t)P * 8, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(dn.ptr, counts, (size_t)P * 4, hipMemcpyHostToDevice));
  if (weights) {
    TRY(dw.ensure((size_t)n_weights * 4));
    HIPC(hipMemcpy(dw.ptr, weights, (size_t)n_weights * 4, hipMemcpyHostToDevice));
  }
  int64_t mi = -1, tests = 0, mixed = 0;
  double kms = 0.0;
  TRY(ppfk::cluster_spatial(ctx, dT.as<float>(), dc.as<unsigned long long>(), dn.as<uint32_t>(),
                            weights ? dw.as<float>() : nullptr, P, d_dist, use_l1 != 0, use_avg != 0,
                            dvc.as<float>(), dtr.as<float>(), dq.as<float>(), &mi, &tests, &mixed, &kms));
  HIPC(hipMemcpyAsync(vc_out, dvc.ptr, (size_t)P * 4, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(trans_out, dtr.ptr, (size_t)P * 12, hipMemcpyDeviceToHost, st));
  HIPC(hipMemcpyAsync(quats_out, dq.ptr, (size_t)P * 16, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  stats_out[0] = mi;
  stats_out[1] = tests;
  stats_out[2] = mixed;
  return PPF_OK;
}

int ppf_debug_primitive(ppf_ctx ctx, int32_t op, void *keys_io, uint32_t *vals_io, int64_t n,
                        int32_t bit_lo, int32_t bit_hi) {
  ARG_CHECK(ctx && keys_io && n >= 0, "ppf_debug_primitive: bad arguments");
  ARG_CHECK(op >= 0 && op <= 4, "ppf_debug_primitive: unknown op");
  TRY(set_device(ctx));
  hipStream_t st = ctx->stream;
  if (op == 3) {
    ARG_CHECK(n > 0 && n <= (1 << 20) && bit_lo > 0, "ppf_debug_primitive: FMA stream needs blocks and iterations");
    ppfi::DevBuf o;
    TRY(o.ensure((size_t)n * 256 * 4));
    TRY(ppfk::debug_fma_stream(ctx, o.as<float>(), n, bit_lo));
    HIPC(hipMemcpyAsync(keys_io, o.ptr, (size_t)n * 256 * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return PPF_OK;
  }
  if (op == 2) {
    ARG_CHECK(vals_io && n % 4 == 0, "ppf_debug_primitive: copy needs an output and n % 4 == 0");
    ppfi::DevBuf a, b;
    TRY(a.ensure((size_t)n * 4 + 16));
    TRY(b.ensure((size_t)n * 4 + 16));
    HIPC(hipMemcpy(a.ptr, keys_io, (size_t)n * 4, hipMemcpyHostToDevice));
    TRY(ppfk::debug_copy(ctx, a.as<uint32_t>(), b.as<uint32_t>(), n));
    HIPC(hipMemcpyAsync(vals_io, b.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return PPF_OK;
  }
  if (op == 4) {
    ARG_CHECK(vals_io, "ppf_debug_primitive: key sort needs vals_io[0] for the path taken");
    ppfi::DevBuf k;
    TRY(k.ensure((size_t)n * 8 + 8));
    HIPC(hipMemcpy(k.ptr, keys_io, (size_t)n * 8, hipMemcpyHostToDevice));
    bool ranked = false;
    TRY(ppfi::sort_key32_ranked(k.as<uint64_t>(), n, ctx->sort, st, &ranked));
    if (!ranked) TRY(ppfi::radix_sort_u64(k.as<uint64_t>(), nullptr, n, 32, 64, ctx->sort, st));
    HIPC(hipMemcpyAsync(keys_io, k.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    vals_io[0] = ranked ? 1u : 0u;
    return PPF_OK;
  }
  if (op == 0) {
    ARG_CHECK(vals_io, "ppf_debug_primitive: scan needs vals_io[0] for the total");
    ppfi::DevBuf a, b, tmp;
    TRY(a.ensure((size_t)n * 4 + 8));
    TRY(b.ensure((size_t)n * 4 + 8));
    HIPC(hipMemcpy(a.ptr, keys_io, (size_t)n * 4, hipMemcpyHostToDevice));
    uint32_t *tot = b.as<uint32_t>() + n;
    TRY(ppfi::exclusive_scan_u32(a.as<uint32_t>(), b.as<uint32_t>(), n, tmp, st, tot));
    HIPC(hipMemcpyAsync(keys_io, b.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
    HIPC(hipMemcpyAsync(vals_io, tot, 4, hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    return PPF_OK;
  }
  ARG_CHECK(bit_lo >= 0 && bit_hi <= 64 && bit_lo <= bit_hi, "ppf_debug_primitive: bad bit range");
  ppfi::DevBuf k, v;
  TRY(k.ensure((size_t)n * 8 + 8));
  if (vals_io) TRY(v.ensure((size_t)n * 4 + 8));
  HIPC(hipMemcpy(k.ptr, keys_io, (size_t)n * 8, hipMemcpyHostToDevice));
  if (vals_io) HIPC(hipMemcpy(v.ptr, vals_io, (size_t)n * 4, hipMemcpyHostToDevice));
  TRY(ppfi::radix_sort_u64(k.as<uint64_t>(), vals_io ? v.as<uint32_t>() : nullptr, n, bit_lo, bit_hi,
                           ctx->sort, st));
  HIPC(hipMemcpyAsync(keys_io, k.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  if (vals_io) HIPC(hipMemcpyAsync(vals_io, v.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  HIPC(hipStreamSynchronize(st));
  return PPF_OK;
}

}  // extern "C"
