// Test program for include/ppf_amd.hpp: the reference's host code, with only
// includes and type names changed, against libppf_amd.so.
//
//   facade_main MODEL.f32 SCENE.f32 D_DIST RPDF THR CPU_CLUSTERING OUT.bin
//
// MODEL/SCENE hold float32 [n][6] (x y z nx ny nz).  OUT receives
// float32[16] pose extracted from Model's public results the way
// src/cuda/ppf.cu:63-93 does, float32[16] pose via ppf_registration (the call
// of src/alignment.cpp:290-298), then int64 n_kept, max_idx, votes,
// n_cpu_transformations, and float32 weightedVoteCounts[0].
#include <cstdio>
#include <cstdlib>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ppf_amd.hpp"

using namespace ppf;

// Stand-in for boost::program_options::variables_map: vm["name"].as<T>().
struct OptionValue {
  std::string text;
  template <class T>
  T as() const {
    std::istringstream is(text);
    T v{};
    is >> v;
    return v;
  }
};

static PointCloud<PointNormal>::Ptr load(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    std::perror(path);
    std::exit(2);
  }
  PointCloud<PointNormal>::Ptr c(new PointCloud<PointNormal>);
  float b[6];
  while (std::fread(b, sizeof(float), 6, f) == 6) c->push_back({b[0], b[1], b[2], b[3], b[4], b[5]});
  std::fclose(f);
  return c;
}

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s MODEL SCENE D_DIST RPDF THR CPU OUT\n", argv[0]);
    return 2;
  }
  std::map<std::string, OptionValue> vm;
  vm["ref_point_df"].text = argv[4];
  vm["vote_count_threshold"].text = argv[5];
  vm["cpu_clustering"].text = argv[6];
  vm["use_l1_norm"].text = "0";
  vm["use_averaged_clusters"].text = "0";
  vm["dev"].text = "0";
  std::vector<PointCloud<PointNormal>::Ptr> model_clouds{load(argv[1])};
  std::vector<PointCloud<PointNormal>::Ptr> scene_clouds{load(argv[2])};
  std::vector<float> model_d_dists{std::strtof(argv[3], nullptr)};
  const bool cpu_clustering = vm["cpu_clustering"].as<bool>();
  try {
    // src/cuda/ppf.cu:63-93, one (scene, model) pair
    Matrix4f T;
    Scene *scene = new Scene(scene_clouds[0].get(), model_d_dists[0], vm["ref_point_df"].as<unsigned int>());
    Model *model = new Model(model_clouds[0].get(), model_d_dists[0], vm["vote_count_threshold"].as<float>(),
                             cpu_clustering, vm["use_l1_norm"].as<bool>(),
                             vm["use_averaged_clusters"].as<bool>());
    model->ppf_lookup(scene);
    if (cpu_clustering) {
      T = model->cpu_transformations[0].pose.matrix();
    } else {
      std::vector<float> transformations = model->getTransformations();
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) T(r, c) = transformations[model->max_idx * 16 + r * 4 + c];
      std::vector<float3> transformation_trans(*model->transformation_trans);
      T(0, 3) = transformation_trans[model->max_idx].x;
      T(1, 3) = transformation_trans[model->max_idx].y;
      T(2, 3) = transformation_trans[model->max_idx].z;
    }
    const int64_t n_kept = (int64_t)model->votes->size();
    if (model->voteCounts->size() != (size_t)n_kept ||
        model->transformations.size() != (size_t)n_kept * 16 ||
        model->weightedVoteCounts.size() != (size_t)n_kept)
      throw Error(1, "result sizes");
    if (!cpu_clustering && (model->vote_counts_out->size() != (size_t)n_kept ||
                            model->transformation_rots->size() != (size_t)n_kept))
      throw Error(1, "clustering result sizes");
    const int64_t st[4] = {n_kept, (int64_t)model->max_idx, model->info().votes,
                           (int64_t)model->cpu_transformations.size()};
    const float w0 = n_kept ? model->weightedVoteCounts[0] : 0.0f;
    delete model;
    delete scene;

    // src/alignment.cpp:290-298
    std::vector<std::vector<Matrix4f>> results =
        ppf_registration(scene_clouds, model_clouds,
                         model_d_dists, vm["ref_point_df"].as<unsigned int>(),
                         vm["vote_count_threshold"].as<float>(),
                         vm["cpu_clustering"].as<bool>(),
                         vm["use_l1_norm"].as<bool>(),
                         vm["use_averaged_clusters"].as<bool>(),
                         vm["dev"].as<int>(),
                         NULL);

    FILE *f = std::fopen(argv[7], "wb");
    if (!f) return 2;
    std::fwrite(T.m.data(), sizeof(float), 16, f);
    std::fwrite(results.at(0).at(0).m.data(), sizeof(float), 16, f);
    std::fwrite(st, sizeof(int64_t), 4, f);
    std::fwrite(&w0, sizeof(float), 1, f);
    std::fclose(f);
  } catch (const Error &e) {
    std::fprintf(stderr, "ppf::Error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
