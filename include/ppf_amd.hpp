// ppf_amd.hpp -- header-only C++ facade over the C ABI of libppf_amd.so.
//
// Mirrors the host API of nicolasavru/ppf-registration-spatial-hashing so that
// a caller switches by changing includes and type names only:
//   Scene(cloud_ptr, d_dist, ref_point_downsample_factor=1)   include/scene.h:15-16
//   Model(cloud, d_dist, vote_count_threshold, cpu_clustering,
//         use_l1_norm, use_averaged_clusters)                 include/model.h:17-19
//   Model::ppf_lookup(Scene*), SetModelPointVoteWeights,
//   getTransformations(), and the public results max_idx, votes,
//   voteCounts, transformations, weightedVoteCounts,
//   transformation_trans, transformation_rots, vote_counts_out,
//   cpu_transformations                                      include/model.h:22-113
//   ppf_registration(scene_clouds, model_clouds, model_d_dists, rpdf, thr,
//                    cpu_clustering, use_l1_norm, use_averaged_clusters,
//                    devUse, model_weights)                  include/ppf.h:9-15
// pcl::PointCloud<pcl::PointNormal> becomes ppf::PointCloud<ppf::PointNormal>
// (same field names, ::Ptr); Eigen::Matrix4f becomes ppf::Matrix4f (row-major
// float[16], operator()(r, c)); Eigen::Affine3f (in PoseWithVotes) becomes
// ppf::Affine3f; thrust::device_vector<T> results become host std::vector<T>
// copies; float3/float4 are ppf::float3/ppf::float4.
//
// Objects made with the reference's constructors live on the default context
// of device 0 (ppf::default_context); every constructor also has a form that
// takes a ppf::Context explicitly.  Errors throw ppf::Error carrying the
// status code and ppf_last_error() (the reference exits the process).
#pragma once

#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "ppf_amd.h"

namespace ppf {

struct PointNormal {
  float x, y, z;
  float normal_x, normal_y, normal_z;
};

// The parts of pcl::PointCloud<PointT> the reference's host code uses.
template <typename PointT>
struct PointCloud {
  using Ptr = std::shared_ptr<PointCloud<PointT>>;
  using ConstPtr = std::shared_ptr<const PointCloud<PointT>>;
  std::vector<PointT> points;
  uint32_t width = 0, height = 1;
  size_t size() const { return points.size(); }
  bool empty() const { return points.empty(); }
  void resize(size_t n) {
    points.resize(n);
    width = (uint32_t)n;
    height = 1;
  }
  void push_back(const PointT &p) {
    points.push_back(p);
    width = (uint32_t)points.size();
    height = 1;
  }
  PointT &operator[](size_t i) { return points[i]; }
  const PointT &operator[](size_t i) const { return points[i]; }
  typename std::vector<PointT>::iterator begin() { return points.begin(); }
  typename std::vector<PointT>::iterator end() { return points.end(); }
  typename std::vector<PointT>::const_iterator begin() const { return points.begin(); }
  typename std::vector<PointT>::const_iterator end() const { return points.end(); }
};
using Cloud = PointCloud<PointNormal>;

struct float3 {
  float x, y, z;
};
struct float4 {
  float x, y, z, w;
};

struct Matrix4f {
  std::array<float, 16> m{};  // row-major
  float &operator()(int r, int c) { return m[4 * r + c]; }
  float operator()(int r, int c) const { return m[4 * r + c]; }
  static Matrix4f Zero() { return Matrix4f{}; }
  static Matrix4f Identity() {
    Matrix4f I;
    I(0, 0) = I(1, 1) = I(2, 2) = I(3, 3) = 1.0f;
    return I;
  }
};

// The rigid transform of a clustered pose (Eigen::Affine3f in the reference).
struct Affine3f {
  Matrix4f T = Matrix4f::Identity();
  Matrix4f matrix() const { return T; }
  float3 translation() const { return {T(0, 3), T(1, 3), T(2, 3)}; }
};

// include/transformation_clustering.h:4-23
struct PoseWithVotes {
  Affine3f pose;
  unsigned int votes = 0;
};
using PoseWithVotesList = std::vector<PoseWithVotes>;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != PPF_OK) throw Error(rc, std::string("ppf: ") + ppf_last_error());
}

namespace detail {
inline void split(const Cloud &c, std::vector<float> &p, std::vector<float> &n) {
  p.resize(c.size() * 3);
  n.resize(c.size() * 3);
  for (size_t i = 0; i < c.size(); ++i) {
    p[3 * i] = c[i].x;
    p[3 * i + 1] = c[i].y;
    p[3 * i + 2] = c[i].z;
    n[3 * i] = c[i].normal_x;
    n[3 * i + 1] = c[i].normal_y;
    n[3 * i + 2] = c[i].normal_z;
  }
}
}  // namespace detail

// One device and one HIP stream (src/cuda/ppf.cu:37-49 device selection).
class Context {
 public:
  explicit Context(int device = 0, uint64_t hbm_budget_bytes = 0) {
    check(ppf_ctx_create(device, hbm_budget_bytes, &h_));
  }
  ~Context() { ppf_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  ppf_ctx get() const { return h_; }

 private:
  ppf_ctx h_ = nullptr;
};

// The context of `device` that the reference-signature constructors use,
// created on first use and kept until the process ends (never destroyed, so
// that static destruction order cannot touch a torn-down HIP runtime).
inline Context &default_context(int device = 0) {
  static std::mutex mu;
  static std::map<int, Context *> ctxs;
  std::lock_guard<std::mutex> g(mu);
  Context *&c = ctxs[device];
  if (!c) c = new Context(device);
  return *c;
}

class Scene {
 public:
  // Scene(cloud_ptr, d_dist, ref_point_downsample_factor) (include/scene.h:15-16).
  // d_dist is kept for getHashKeys(): scene features are quantized with the
  // model's d_dist during the lookup.
  Scene(const Cloud *cloud_ptr, float d_dist, unsigned int ref_point_downsample_factor = 1)
      : Scene(default_context(), cloud_ptr, d_dist, ref_point_downsample_factor) {}
  Scene(Context &ctx, const Cloud *cloud_ptr, float d_dist,
        unsigned int ref_point_downsample_factor = 1)
      : cloud_ptr(cloud_ptr), d_dist_(d_dist), n_((int)cloud_ptr->size()) {
    std::vector<float> p, n;
    detail::split(*cloud_ptr, p, n);
    check(ppf_scene_upload(ctx.get(), p.data(), n.data(), (int64_t)cloud_ptr->size(),
                           ref_point_downsample_factor, &h_));
  }
  ~Scene() { ppf_scene_destroy(h_); }
  Scene(const Scene &) = delete;
  Scene &operator=(const Scene &) = delete;
  int numPoints() const { return n_; }
  std::vector<uint32_t> getHashKeys() const {
    std::vector<uint32_t> k((size_t)n_ * n_);
    check(ppf_scene_keys(h_, d_dist_, k.data()));
    return k;
  }
  ppf_scene get() const { return h_; }

  const Cloud *cloud_ptr;  // borrowed, as in the reference (scene.cu:27)

 private:
  ppf_scene h_ = nullptr;
  float d_dist_;
  int n_;
};

class Model {
 public:
  // Model(cloud, d_dist, vote_count_threshold, cpu_clustering, use_l1_norm,
  // use_averaged_clusters) (include/model.h:17-19).
  Model(const Cloud *cloud, float d_dist, float vote_count_threshold, bool cpu_clustering,
        bool use_l1_norm, bool use_averaged_clusters)
      : Model(default_context(), cloud, d_dist, vote_count_threshold, cpu_clustering, use_l1_norm,
              use_averaged_clusters) {}
  Model(Context &ctx, const Cloud *cloud, float d_dist, float vote_count_threshold,
        bool cpu_clustering, bool use_l1_norm, bool use_averaged_clusters)
      : cloud_ptr(cloud),
        vote_count_threshold(vote_count_threshold),
        cpu_clustering(cpu_clustering),
        use_l1_norm(use_l1_norm),
        use_averaged_clusters(use_averaged_clusters),
        modelPointVoteWeights(cloud->size(), 1.0f) {
    std::vector<float> p, n;
    detail::split(*cloud, p, n);
    check(ppf_model_build(ctx.get(), p.data(), n.data(), (int64_t)cloud->size(), d_dist, &h_));
  }
  ~Model() {
    ppf_result_destroy(res_);
    ppf_model_destroy(h_);
  }
  Model(const Model &) = delete;
  Model &operator=(const Model &) = delete;

  // Model::SetModelPointVoteWeights (src/cuda/model.cu:91-93)
  void SetModelPointVoteWeights(const std::vector<float> &w) {
    check(ppf_model_set_weights(h_, w.data(), (int64_t)w.size()));
    modelPointVoteWeights = w;
  }

  // Model::ppf_lookup (src/cuda/model.cu:269-306): fills the public results
  void ppf_lookup(Scene *scene) {
    ppf_result_destroy(res_);
    res_ = nullptr;
    const ppf_lookup_params prm{vote_count_threshold, cpu_clustering ? 1 : 0, use_l1_norm ? 1 : 0,
                                use_averaged_clusters ? 1 : 0, 0, 0, 0};
    check(::ppf_lookup(h_, scene->get(), &prm, &res_));
    check(ppf_result_info(res_, &info_));
    const size_t P = (size_t)info_.n_kept;
    votes_.assign(P, 0);
    voteCounts_.assign(P, 0);
    transformations.assign(P * 16, 0.0f);
    weightedVoteCounts.assign(P, 0.0f);
    if (P) {
      std::vector<uint64_t> codes(P);
      check(ppf_result_votes(res_, codes.data(), voteCounts_.data()));
      for (size_t i = 0; i < P; ++i) votes_[i] = (unsigned long)codes[i];
      check(ppf_result_transforms(res_, transformations.data()));
      // ComputeWeightedVoteCounts (model.cu:173-189, kernel.cu:766-782)
      for (size_t i = 0; i < P; ++i)
        weightedVoteCounts[i] =
            modelPointVoteWeights[(size_t)((codes[i] >> 6) & 0x3ffffffu)] * (float)voteCounts_[i];
    }
    votes = &votes_;
    voteCounts = &voteCounts_;
    cpu_transformations.clear();
    transformation_trans = nullptr;
    transformation_rots = nullptr;
    vote_counts_out = nullptr;
    max_idx = 0;
    if (cpu_clustering) {
      if (P) {
        float poses[3 * 16];
        uint32_t v[3];
        int32_t nc = 0;
        check(ppf_result_cpu_clusters(res_, poses, v, &nc));
        for (int32_t i = 0; i < nc; ++i) {
          PoseWithVotes pw;
          std::memcpy(pw.pose.T.m.data(), poses + 16 * i, 64);
          pw.votes = v[i];
          cpu_transformations.push_back(pw);
        }
      }
    } else {
      trans_.assign(P, float3{0, 0, 0});
      rots_.assign(P, float4{0, 0, 0, 0});
      vc_.assign(P, 0.0f);
      if (P)
        check(ppf_result_clusters(res_, vc_.data(), reinterpret_cast<float *>(trans_.data()),
                                  reinterpret_cast<float *>(rots_.data())));
      transformation_trans = &trans_;
      transformation_rots = &rots_;
      vote_counts_out = &vc_;
      max_idx = (unsigned int)(info_.max_idx < 0 ? 0 : info_.max_idx);
    }
  }

  // row-major float[16] per kept vote (Model::getTransformations, model.cu:308-310)
  std::vector<float> getTransformations() const { return transformations; }
  // final pose as ppf_registration extracts it (src/cuda/ppf.cu:74-93)
  Matrix4f pose() const {
    Matrix4f T;
    if (res_) {
      int rc = ppf_result_pose(res_, T.m.data());
      if (rc != PPF_OK && rc != PPF_ERR_NO_VOTES) check(rc);
    }
    return T;
  }
  const ppf_result_info_t &info() const { return info_; }
  ppf_model get() const { return h_; }

  const Cloud *cloud_ptr;
  float vote_count_threshold;
  bool cpu_clustering;
  bool use_l1_norm;
  bool use_averaged_clusters;

  // results of the last ppf_lookup, as the reference's public members
  // (include/model.h:57-113) hold them, copied to the host
  std::vector<unsigned long> *votes = nullptr;      // kept vote codes [s_r|m_r|alpha]
  std::vector<unsigned int> *voteCounts = nullptr;  // their counts
  std::vector<float> transformations;               // 16 floats per kept vote
  std::vector<float> weightedVoteCounts;
  std::vector<float3> *transformation_trans = nullptr;  // GPU clustering only
  std::vector<float4> *transformation_rots = nullptr;   // GPU clustering only
  std::vector<float> *vote_counts_out = nullptr;        // GPU clustering only
  PoseWithVotesList cpu_transformations;                // CPU clustering only
  std::vector<float> modelPointVoteWeights;
  unsigned int max_idx = 0;

 private:
  ppf_model h_ = nullptr;
  ppf_result res_ = nullptr;
  ppf_result_info_t info_{};
  std::vector<unsigned long> votes_;
  std::vector<unsigned int> voteCounts_;
  std::vector<float3> trans_;
  std::vector<float4> rots_;
  std::vector<float> vc_;
};

// ppf_registration (include/ppf.h:9-15, src/cuda/ppf.cu:29-106): poses
// [scene][model].  model_weights may be NULL (the reference accepts the
// pointer and ignores it); when given it holds one weight per model point,
// the models' points concatenated in model order.  Runs on devUse's default
// context.
inline std::vector<std::vector<Matrix4f>> ppf_registration(
    std::vector<Cloud::Ptr> scene_clouds, std::vector<Cloud::Ptr> model_clouds,
    std::vector<float> model_d_dists, unsigned int ref_point_downsample_factor,
    float vote_count_threshold, bool cpu_clustering, bool use_l1_norm, bool use_averaged_clusters,
    int devUse, float *model_weights) {
  if (model_d_dists.size() != model_clouds.size())
    throw Error(PPF_ERR_ARG, "ppf: each model needs a d_dist");
  Context &ctx = default_context(devUse < 0 ? 0 : devUse);
  std::vector<std::vector<float>> sp(scene_clouds.size()), sn(scene_clouds.size());
  std::vector<std::vector<float>> mp(model_clouds.size()), mn(model_clouds.size());
  std::vector<const float *> spp, snp, mpp, mnp, wp;
  std::vector<int64_t> ss, ms;
  for (size_t i = 0; i < scene_clouds.size(); ++i) {
    detail::split(*scene_clouds[i], sp[i], sn[i]);
    spp.push_back(sp[i].data());
    snp.push_back(sn[i].data());
    ss.push_back((int64_t)scene_clouds[i]->size());
  }
  size_t woff = 0;
  for (size_t j = 0; j < model_clouds.size(); ++j) {
    detail::split(*model_clouds[j], mp[j], mn[j]);
    mpp.push_back(mp[j].data());
    mnp.push_back(mn[j].data());
    ms.push_back((int64_t)model_clouds[j]->size());
    wp.push_back(model_weights ? model_weights + woff : nullptr);
    woff += model_clouds[j]->size();
  }
  std::vector<float> poses(scene_clouds.size() * model_clouds.size() * 16);
  check(::ppf_registration(ctx.get(), (int32_t)scene_clouds.size(), spp.data(), snp.data(),
                           ss.data(), (int32_t)model_clouds.size(), mpp.data(), mnp.data(),
                           ms.data(), model_d_dists.data(), ref_point_downsample_factor,
                           vote_count_threshold, cpu_clustering, use_l1_norm,
                           use_averaged_clusters, model_weights ? wp.data() : nullptr,
                           poses.data(), nullptr));
  std::vector<std::vector<Matrix4f>> out(scene_clouds.size());
  for (size_t i = 0; i < scene_clouds.size(); ++i)
    for (size_t j = 0; j < model_clouds.size(); ++j) {
      Matrix4f T;
      std::memcpy(T.m.data(), &poses[(i * model_clouds.size() + j) * 16], 64);
      out[i].push_back(T);
    }
  return out;
}

}  // namespace ppf
