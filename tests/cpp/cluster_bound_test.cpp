// Soundness of the wave bounds of GPU clustering (csrc/ppf_cluster.h): a
// member that cluster_skip() rejects must fail the exact pose-pair test
// (cluster_test) against every pose of the group, or skipping it would change
// a vote sum.  Random groups of 128 poses with concentrated and spread
// rotations, members built to sit just inside and just outside both
// thresholds, non-unit and zero quaternions, NaNs.  Prints the skip rate per
// setting; exits non-zero on the first violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "ppf_cluster.h"
#include "ppf_geom.h"

using ppfc::ClPose;
using ppfc::ClWave;

static float sqrt_below_max(float t) {
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

struct Q {
  double w, x, y, z;
};

static Q qmul(Q a, Q b) {
  return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z, a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
          a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x, a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}

static Q qaxis(double ax, double ay, double az, double ang) {
  const double n = std::sqrt(ax * ax + ay * ay + az * az);
  const double s = std::sin(ang / 2) / n;
  return {std::cos(ang / 2), ax * s, ay * s, az * s};
}

// float quaternion the way the library makes it: rotation matrix in float,
// then rotmat_to_quat (scalar first)
static void quat_of(Q q, float out[4]) {
  const double w = q.w, x = q.x, y = q.y, z = q.z;
  float T[16] = {(float)(1 - 2 * (y * y + z * z)), (float)(2 * (x * y - w * z)), (float)(2 * (x * z + w * y)), 0,
                 (float)(2 * (x * y + w * z)), (float)(1 - 2 * (x * x + z * z)), (float)(2 * (y * z - w * x)), 0,
                 (float)(2 * (x * z - w * y)), (float)(2 * (y * z + w * x)), (float)(1 - 2 * (x * x + y * y)), 0,
                 0, 0, 0, 1};
  ppfg::rotmat_to_quat(T, out);
}

int main() {
  std::mt19937_64 rng(12345);
  std::normal_distribution<double> N(0.0, 1.0);
  std::uniform_real_distribution<double> Uf(0.0, 1.0);
  const double rot_pass = 2.0 * std::acos(1.0 - (double)(ppfm::ROT_THRESH * ppfm::ROT_THRESH) / 8.0);
  long long violations = 0, members = 0, skipped = 0;
  const double spreads[] = {0.002, 0.02, 0.1, 0.4, 3.0};
  for (int setting = 0; setting < 40; ++setting) {
    const double spread = spreads[setting % 5];
    const int use_l1 = (setting / 5) % 2;
    const bool odd = (setting / 10) % 2;  // non-unit / NaN / zero poses mixed in
    const float d = (float)(0.02 + 0.2 * Uf(rng));
    const float s_max = sqrt_below_max(d);
    const double tscale = setting >= 20 ? 50.0 : 2.0;
    // group centre
    Q c = qaxis(N(rng), N(rng), N(rng), Uf(rng) * 3.14159);
    const double cx = tscale * N(rng), cy = tscale * N(rng), cz = tscale * N(rng);
    std::vector<ClPose> P(128);
    std::vector<bool> have(128, true);
    for (int i = 0; i < 128; ++i) {
      Q r = qmul(qaxis(N(rng), N(rng), N(rng), spread * Uf(rng)), c);
      float q[4];
      quat_of(r, q);
      ClPose &p = P[i];
      p.qx = q[0];
      p.qy = q[1];
      p.qz = q[2];
      p.qw = q[3];
      p.tx = (float)(cx + d * (Uf(rng) - 0.5) * 1.5);
      p.ty = (float)(cy + d * (Uf(rng) - 0.5) * 1.5);
      p.tz = (float)(cz + d * (Uf(rng) - 0.5) * 1.5);
      if (odd && i % 17 == 3) {
        const float sc = (float)(0.9 + 0.2 * Uf(rng));
        p.qx *= sc, p.qy *= sc, p.qz *= sc, p.qw *= sc;
      }
      if (odd && i % 29 == 5) p.qx = NAN;
      if (odd && i % 31 == 7) p.tx = NAN;
      if (odd && i % 37 == 9) p.qx = p.qy = p.qz = p.qw = 0.0f;
      if (i % 41 == 11) have[i] = false;
      p.vc = 1.0f;
      p.ox = p.tx, p.oy = p.ty, p.oz = p.tz;
    }
    // centre: the first pose (odd values of setting / 2), or the group's
    // mean quaternion and translation-box midpoint as the GPU kernel builds it
    ClWave W;
    if ((setting / 2) % 2 == 0) {
      float sq[4] = {0, 0, 0, 0}, nq = 0.0f;
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (int i = 0; i < 128; ++i) {
        if (!have[i]) continue;
        const ClPose &p = P[i];
        if (std::isfinite(p.qx) && std::isfinite(p.qy) && std::isfinite(p.qz) && std::isfinite(p.qw)) {
          sq[0] += p.qx, sq[1] += p.qy, sq[2] += p.qz, sq[3] += p.qw;
          nq += 1.0f;
        }
        const float t[3] = {p.tx, p.ty, p.tz};
        for (int k = 0; k < 3; ++k) lo[k] = fminf(lo[k], t[k]), hi[k] = fmaxf(hi[k], t[k]);
      }
      ppfc::cl_wave_center(ppfc::cl_group_centre(sq, nq, lo, hi, P[0]), have[0], W);
    } else {
      ppfc::cl_wave_center(P[0], have[0], W);
    }
    float rho = 0, nmax = 0, rad = 0;
    for (int i = 0; i < 128; ++i)
      if (have[i]) ppfc::cl_wave_add(W, P[i], rho, nmax, rad);
    ppfc::cl_wave_finish(W, rho, nmax, rad, d);
    long long sk = 0, mm = 0;
    for (int m = 0; m < 20000; ++m) {
      // members: around a random group pose, at rotation angles and
      // distances concentrated near the thresholds
      const int i = (int)(Uf(rng) * 128) % 128;
      Q qi = {P[i].qx, P[i].qy, P[i].qz, P[i].qw};
      const double qn = std::sqrt(qi.w * qi.w + qi.x * qi.x + qi.y * qi.y + qi.z * qi.z);
      if (!(qn > 0)) continue;
      qi = {qi.w / qn, qi.x / qn, qi.y / qn, qi.z / qn};
      double ang;
      const double sel = Uf(rng);
      if (sel < 0.4)
        ang = rot_pass * (1.0 + 1e-4 * N(rng));
      else if (sel < 0.7)
        ang = rot_pass + 0.05 * N(rng);
      else
        ang = 3.14159 * Uf(rng);
      const Q r = qmul(qaxis(N(rng), N(rng), N(rng), ang), {qi.w, qi.x, qi.y, qi.z});
      float q[4];
      quat_of(r, q);
      if (m % 13 == 0) {
        const float sc = (float)(0.8 + 0.4 * Uf(rng));
        for (float &x : q) x *= sc;
      }
      if (m % 19 == 0) {
        // the negated quaternion of the same rotation
        for (float &x : q) x = -x;
      }
      double dist;
      const double s2 = Uf(rng);
      if (s2 < 0.4)
        dist = d * (1.0 + 1e-6 * N(rng));
      else if (s2 < 0.7)
        dist = d * (1.0 + 0.3 * N(rng));
      else
        dist = d * 3.0 * Uf(rng);
      double ux = N(rng), uy = N(rng), uz = N(rng);
      const double un = std::sqrt(ux * ux + uy * uy + uz * uz);
      const float tx = (float)(P[i].tx + dist * ux / un), ty = (float)(P[i].ty + dist * uy / un),
                  tz = (float)(P[i].tz + dist * uz / un);
      if (!(tx == tx)) continue;
      ++mm;
      if (!ppfc::cluster_skip(W, q[0], q[1], q[2], q[3], tx, ty, tz, use_l1)) continue;
      ++sk;
      for (int k = 0; k < 128; ++k) {
        if (!have[k]) continue;
        ClPose t = P[k];
        ppfc::cluster_test(t, q[0], q[1], q[2], q[3], tx, ty, tz, 2.0f, s_max, use_l1, 0);
        if (t.vc != P[k].vc) {
          ++violations;
          std::printf("VIOLATION setting %d member %d pose %d\n", setting, m, k);
          return 1;
        }
      }
    }
    members += mm;
    skipped += sk;
    std::printf("setting %2d spread %.3f l1 %d odd %d: skipped %5.1f%% of %lld\n", setting, spread,
                use_l1, (int)odd, 100.0 * (double)sk / (double)(mm ? mm : 1), mm);
  }
  std::printf("members %lld skipped %lld violations %lld\n", members, skipped, violations);
  // the chunk order's Hilbert curve: every point of the 4-D grid [0, 8)^4 has
  // its own position, and consecutive positions are neighbouring cells
  std::vector<int> at(4096, -1);
  for (uint32_t x = 0; x < 4096; ++x) {
    const uint32_t h = ppfc::hilbert4<3>(x & 7, (x >> 3) & 7, (x >> 6) & 7, x >> 9);
    if (h >= 4096 || at[h] >= 0) {
      std::printf("HILBERT not a bijection at %u\n", x);
      return 1;
    }
    at[h] = (int)x;
  }
  for (int h = 1; h < 4096; ++h) {
    int l1 = 0;
    for (int k = 0; k < 4; ++k) l1 += std::abs(((at[h] >> (3 * k)) & 7) - ((at[h - 1] >> (3 * k)) & 7));
    if (l1 != 1) {
      std::printf("HILBERT step %d is not to a neighbour\n", h);
      return 1;
    }
  }
  std::printf("hilbert ok\n");
  return violations ? 1 : 0;
}
