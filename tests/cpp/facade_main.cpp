// Test program for include/ppf_amd.hpp: the reference's host API shape
// (Scene/Model/ppf_lookup, ppf_registration) driven from C++.
//
//   facade_main MODEL.f32 SCENE.f32 D_DIST RPDF THR CPU_CLUSTERING OUT.bin
//
// MODEL/SCENE hold float32 [n][6] (x y z nx ny nz).  OUT receives
// float32[16] pose via Model::ppf_lookup, float32[16] pose via
// ppf_registration, then int64 n_kept, max_idx, votes.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ppf_amd.hpp"

static ppf::PointCloud load(const char *path) {
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    std::perror(path);
    std::exit(2);
  }
  std::vector<float> v;
  float buf[6];
  while (std::fread(buf, sizeof(float), 6, f) == 6) v.insert(v.end(), buf, buf + 6);
  std::fclose(f);
  ppf::PointCloud c(v.size() / 6);
  for (size_t i = 0; i < c.size(); ++i)
    c[i] = {v[6 * i], v[6 * i + 1], v[6 * i + 2], v[6 * i + 3], v[6 * i + 4], v[6 * i + 5]};
  return c;
}

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s MODEL SCENE D_DIST RPDF THR CPU OUT\n", argv[0]);
    return 2;
  }
  ppf::PointCloud model = load(argv[1]), scene = load(argv[2]);
  float d_dist = std::strtof(argv[3], nullptr);
  unsigned rpdf = (unsigned)std::strtoul(argv[4], nullptr, 10);
  float thr = std::strtof(argv[5], nullptr);
  bool cpu = std::atoi(argv[6]) != 0;
  try {
    ppf::Context ctx(0);
    ppf::Scene s(ctx, &scene, d_dist, rpdf);
    ppf::Model m(ctx, &model, d_dist, thr, cpu, false, false);
    m.ppf_lookup(&s);
    ppf::Matrix4f T1 = m.pose();
    std::vector<float> t = m.getTransformations();
    if (t.size() != (size_t)m.info().n_kept * 16) throw ppf::Error(1, "transform count");
    auto all = ppf::ppf_registration({&scene}, {&model}, {d_dist}, rpdf, thr, cpu, false, false, 0);
    ppf::Matrix4f T2 = all.at(0).at(0);
    FILE *f = std::fopen(argv[7], "wb");
    if (!f) return 2;
    std::fwrite(T1.m.data(), sizeof(float), 16, f);
    std::fwrite(T2.m.data(), sizeof(float), 16, f);
    int64_t st[3] = {m.info().n_kept, m.info().max_idx, m.info().votes};
    std::fwrite(st, sizeof(int64_t), 3, f);
    std::fclose(f);
  } catch (const ppf::Error &e) {
    std::fprintf(stderr, "ppf::Error %d: %s\n", e.code, e.what());
    return 3;
  }
  return 0;
}
