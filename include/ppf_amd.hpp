// ppf_amd.hpp -- header-only C++ facade over the C ABI of libppf_amd.so.
//
// Mirrors the host API of nicolasavru/ppf-registration-spatial-hashing so a
// caller switches by changing includes and types:
//   Scene(cloud, d_dist, ref_point_downsample_factor)      include/scene.h:15-16
//   Model(cloud, d_dist, vote_count_threshold, cpu_clustering,
//         use_l1_norm, use_averaged_clusters)             include/model.h:17-19
//   Model::ppf_lookup(Scene*), Model::max_idx,
//   Model::getTransformations(), SetModelPointVoteWeights  include/model.h:22-37
//   ppf_registration(scenes, models, d_dists, rpdf, thr,
//                    cpu_clustering, use_l1, use_avg, dev, weights)  include/ppf.h:9-15
// pcl::PointCloud<pcl::PointNormal> becomes ppf::PointCloud (a vector of
// PointNormal with the same field names); Eigen::Matrix4f becomes ppf::Matrix4f
// (row-major float[16], operator()(r, c)).  Errors throw ppf::Error carrying
// the status code and ppf_last_error() (the reference exits the process).
#pragma once

#include <array>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ppf_amd.h"

namespace ppf {

struct PointNormal {
  float x, y, z;
  float normal_x, normal_y, normal_z;
};
using PointCloud = std::vector<PointNormal>;

struct Matrix4f {
  std::array<float, 16> m{};  // row-major
  float &operator()(int r, int c) { return m[4 * r + c]; }
  float operator()(int r, int c) const { return m[4 * r + c]; }
  static Matrix4f Zero() { return Matrix4f{}; }
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != PPF_OK) throw Error(rc, std::string("ppf: ") + ppf_last_error());
}

namespace detail {
inline void split(const PointCloud &c, std::vector<float> &p, std::vector<float> &n) {
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

class Scene {
 public:
  // d_dist is accepted for signature parity: scene features are quantized
  // with the model's d_dist during the lookup.
  Scene(Context &ctx, const PointCloud *cloud_ptr, float d_dist,
        unsigned int ref_point_downsample_factor = 1)
      : d_dist_(d_dist), n_((int)cloud_ptr->size()) {
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

 private:
  ppf_scene h_ = nullptr;
  float d_dist_;
  int n_;
};

class Model {
 public:
  Model(Context &ctx, const PointCloud *cloud, float d_dist, float vote_count_threshold,
        bool cpu_clustering, bool use_l1_norm, bool use_averaged_clusters)
      : prm_{vote_count_threshold, cpu_clustering ? 1 : 0, use_l1_norm ? 1 : 0,
             use_averaged_clusters ? 1 : 0, 0, 0, 0} {
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

  void SetModelPointVoteWeights(const std::vector<float> &w) {
    check(ppf_model_set_weights(h_, w.data(), (int64_t)w.size()));
  }

  // Model::ppf_lookup (src/cuda/model.cu:269-306)
  void ppf_lookup(Scene *scene) {
    ppf_result_destroy(res_);
    res_ = nullptr;
    check(::ppf_lookup(h_, scene->get(), &prm_, &res_));
    check(ppf_result_info(res_, &info_));
    max_idx = (unsigned int)(info_.max_idx < 0 ? 0 : info_.max_idx);
  }

  // row-major float[16] per kept vote (Model::getTransformations)
  std::vector<float> getTransformations() const {
    std::vector<float> t((size_t)info_.n_kept * 16);
    if (res_) check(ppf_result_transforms(res_, t.data()));
    return t;
  }
  // kept vote codes and counts (Model::votes, Model::voteCounts)
  void getVotes(std::vector<uint64_t> &codes, std::vector<uint32_t> &counts) const {
    codes.resize((size_t)info_.n_kept);
    counts.resize((size_t)info_.n_kept);
    if (res_) check(ppf_result_votes(res_, codes.data(), counts.data()));
  }
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

  unsigned int max_idx = 0;

 private:
  ppf_model h_ = nullptr;
  ppf_result res_ = nullptr;
  ppf_lookup_params prm_;
  ppf_result_info_t info_{};
};

// ppf_registration (include/ppf.h:9-15): poses[scene][model].  model_weights
// may be null; when given it holds one weight array per model.
inline std::vector<std::vector<Matrix4f>> ppf_registration(
    const std::vector<const PointCloud *> &scene_clouds,
    const std::vector<const PointCloud *> &model_clouds, const std::vector<float> &model_d_dists,
    unsigned int ref_point_downsample_factor, float vote_count_threshold, bool cpu_clustering,
    bool use_l1_norm, bool use_averaged_clusters, int devUse,
    const std::vector<std::vector<float>> *model_weights = nullptr) {
  if (model_d_dists.size() != model_clouds.size())
    throw Error(PPF_ERR_ARG, "ppf: each model needs a d_dist");
  Context ctx(devUse);
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
  for (size_t j = 0; j < model_clouds.size(); ++j) {
    detail::split(*model_clouds[j], mp[j], mn[j]);
    mpp.push_back(mp[j].data());
    mnp.push_back(mn[j].data());
    ms.push_back((int64_t)model_clouds[j]->size());
    wp.push_back(model_weights ? (*model_weights)[j].data() : nullptr);
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
