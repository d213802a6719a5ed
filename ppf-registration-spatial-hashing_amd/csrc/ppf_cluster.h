// ppf_cluster.h -- the pose-pair test of GPU spatial-hash clustering and the
// conservative group bounds that let a wave skip members that fail the test
// for all of its poses.  Host/device, so that the bounds can be checked on
// the host against the exact test.
#pragma once

#include "ppf_math.h"

namespace ppfc {

using namespace ppfm;

// The rotation test below can pass only when the float dot product of the two
// quaternions exceeds 1 - ROT_THRESH^2/8 = 0.97807.
constexpr float CL_ROT_DOT_MIN = 0.978f;
constexpr float CL_ANGLE_MARGIN = 5e-3f;  // rad, far above float rounding of the angles
constexpr float CL_DIST_MARGIN = 1e-5f;   // relative
// Members whose quaternion norm lies in [CL_NLO, CL_NHI] (the unit
// quaternions of rotation matrices) get the cheap rotation pre-test.
constexpr float CL_NLO = 0.999f;
constexpr float CL_NHI = 1.001f;

struct ClPose {
  float tx, ty, tz, qx, qy, qz, qw;
  float vc, ox, oy, oz;  // running vote count and (averaged) translation
};

// One member j of a neighbour cell against pose P: rot_clustering_kernel's
// loop body (kernel.cu:735-756).  The rotation test |8(1 - q.q_j)| <
// ROT_THRESH^2 as written there; the translation test sqrtf(s) < d_dist as
// s <= s_max (sqrtf is correctly rounded and monotone, s_max is the largest
// float whose square root rounds below d_dist).  jt.w is the member's weight.
PPF_HD void cluster_test(ClPose &P, float jqx, float jqy, float jqz, float jqw, float jtx, float jty,
                         float jtz, float wj, float s_max, int use_l1, int use_avg) {
  // |8 (1 - d4)| < T is |1 - d4| < T/8: scaling by a power of two is exact
  // in both products (and overflows to infinity on both sides)
  const float rot_thresh_sq = ROT_THRESH * ROT_THRESH;
  const float rot_thresh_sq8 = rot_thresh_sq * 0.125f;
  const float d4 = P.qx * jqx + P.qy * jqy + P.qz * jqz + P.qw * jqw;
  const float qd = fabsf(1.0f - d4);
  if (qd < rot_thresh_sq8) {
    if (!use_l1) {
      const float dx = P.tx - jtx, dy = P.ty - jty, dz = P.tz - jtz;
      const float ss = dx * dx + dy * dy + dz * dz;
      if (!(ss <= s_max)) return;
    }
    if (use_avg) {
      P.ox = P.vc * P.ox;
      P.oy = P.vc * P.oy;
      P.oz = P.vc * P.oz;
      P.ox = P.ox + wj * jtx;
      P.oy = P.oy + wj * jty;
      P.oz = P.oz + wj * jtz;
      const float sc = 1.0f / (P.vc + wj);
      P.ox = sc * P.ox;
      P.oy = sc * P.oy;
      P.oz = sc * P.oz;
    }
    P.vc += wj;
  }
}

PPF_HD float clamp_unit(float x) { return fminf(fmaxf(x, -1.0f), 1.0f); }

// Bounds of a group of poses: a rotation cap (normalized centre c, angle rho,
// largest quaternion norm) and a translation ball (centre, radius).  A member
// whose quaternion angle to c exceeds acos(0.978 / (nmax |q_j|)) + rho cannot
// pass the rotation test with any pose of the group (a pass needs the cosine
// of the angle between q_i and q_j above that bound; triangle inequality on
// the unit 3-sphere); a member farther than d_dist + rad from the ball's
// centre cannot pass the translation test.  ok = false disables skipping.
//
// The per-member test uses two thresholds fixed per group.  Rotation: with
// |q_j| in [NLO, NHI], kappa_j = 0.978/(nmax |q_j|) - 1e-5 >= kappa_lo (the
// same with NHI), so an angle above Theta = acos(kappa_lo) + rho + margin
// exceeds the member's own bound; and c.q_j < N cos(Theta) - 1e-6 <=
// |q_j| cos(Theta) - 1e-6 implies that angle, with N = NLO when cos(Theta)
// >= 0 and N = NHI otherwise (Theta < 3.1).  Translation:
// with M = |centre| + d + rad, a member at |t_j| <= 2M has the margin
// 1e-5 (M + |t_j|) <= 3e-5 M of the distance bound, and a member beyond 2M is
// farther than d + rad + 1e-5 (M + |t_j|) anyway, so the squared distance is
// compared with (d + rad + 4e-5 M)^2.
struct ClWave {
  float cqx, cqy, cqz, cqw;
  float ctx, cty, ctz;
  float rho, nmax, rad, cmag;
  float cos_thr;  // skip when |q_j|^2 in [NLO^2, NHI^2] and c.q_j < cos_thr
  float b2;       // skip when |t_j - centre|^2 > b2 (unless use_l1)
  bool ok;
};

// Centre from the group's first pose.
PPF_HD void cl_wave_center(const ClPose &c, bool have, ClWave &W) {
  const float cn2 = c.qx * c.qx + c.qy * c.qy + c.qz * c.qz + c.qw * c.qw;
  bool ok = have && cn2 > 0.25f && cn2 < 4.0f;
  const float v[7] = {c.qx, c.qy, c.qz, c.qw, c.tx, c.ty, c.tz};
  for (int k = 0; k < 7; ++k) ok = ok && (v[k] == v[k]) && fabsf(v[k]) < 1e30f;
  const float ci = ok ? 1.0f / sqrtf(cn2) : 0.0f;
  W.cqx = c.qx * ci;
  W.cqy = c.qy * ci;
  W.cqz = c.qz * ci;
  W.cqw = c.qw * ci;
  W.ctx = c.tx;
  W.cty = c.ty;
  W.ctz = c.tz;
  W.rho = W.nmax = W.rad = W.cmag = 0.0f;
  W.ok = ok;
}

// Fold one pose of the group into (rho, nmax, rad).  A pose with NaN
// components never passes a test, so fmaxf dropping NaN keeps the bounds
// valid.
PPF_HD void cl_wave_add(const ClWave &W, const ClPose &P, float &rho, float &nmax, float &rad) {
  const float n = sqrtf(P.qx * P.qx + P.qy * P.qy + P.qz * P.qz + P.qw * P.qw);
  if (n > 0.0f) {
    const float d = (W.cqx * P.qx + W.cqy * P.qy + W.cqz * P.qz + W.cqw * P.qw) / n;
    rho = fmaxf(rho, acosf_own(clamp_unit(d)));
  }
  nmax = fmaxf(nmax, n);
  const float dx = P.tx - W.ctx, dy = P.ty - W.cty, dz = P.tz - W.ctz;
  rad = fmaxf(rad, sqrtf(dx * dx + dy * dy + dz * dz));
}

PPF_HD void cl_wave_finish(ClWave &W, float rho, float nmax, float rad, float d_dist) {
  W.rho = rho + CL_ANGLE_MARGIN;
  W.nmax = nmax * (1.0f + CL_DIST_MARGIN);
  W.rad = rad;
  W.cmag = sqrtf(W.ctx * W.ctx + W.cty * W.cty + W.ctz * W.ctz);
  W.cos_thr = -INFINITY;
  W.b2 = INFINITY;
  if (!W.ok) return;
  const float kappa = CL_ROT_DOT_MIN / (W.nmax * CL_NHI) - 1e-5f;
  if (kappa >= 1.0f) {
    W.cos_thr = INFINITY;  // no member in the norm range can pass
  } else if (kappa > -1.0f) {
    const float th = acosf_own(kappa) + W.rho + CL_ANGLE_MARGIN;
    if (th < 3.1f) {
      const float c = cosf_own(th);
      W.cos_thr = (c >= 0.0f ? CL_NLO : CL_NHI) * c - 1e-6f;
    }
  }
  const float M = W.cmag + d_dist + W.rad;
  const float B = d_dist + W.rad + 4e-5f * M;
  if (B < 1e18f) W.b2 = B * B * (1.0f + 1e-6f);
}

// Centre of a group for cl_wave_center: the mean of the group's quaternions
// (sq = their sum over nq poses with finite components) and the midpoint of
// the box [lo, hi] of its translations; `first` (a pose of the group) when
// the mean is too short to normalize reliably.  Any centre gives sound
// bounds (rho and rad are measured from it); a central one gives tighter
// ones than a pose at the group's edge.
PPF_HD ClPose cl_group_centre(const float sq[4], float nq, const float lo[3], const float hi[3],
                              const ClPose &first) {
  ClPose c = first;
  if (nq > 0.0f) {
    const float inv = 1.0f / nq;
    const float mx = sq[0] * inv, my = sq[1] * inv, mz = sq[2] * inv, mw = sq[3] * inv;
    if (mx * mx + my * my + mz * mz + mw * mw > 0.25f) {
      c.qx = mx;
      c.qy = my;
      c.qz = mz;
      c.qw = mw;
    }
  }
  if (lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2]) {
    c.tx = 0.5f * lo[0] + 0.5f * hi[0];
    c.ty = 0.5f * lo[1] + 0.5f * hi[1];
    c.tz = 0.5f * lo[2] + 0.5f * hi[2];
  }
  return c;
}

// Position of a point of the 4-D grid [0, 2^B)^4 along a Hilbert curve
// (Skilling's transpose form, "Programming the Hilbert curve", AIP Conf.
// Proc. 707, 2004): consecutive positions are neighbouring cells, so
// consecutive poses in this order have close rotations.
template <int B>
PPF_HD uint32_t hilbert4(uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
  uint32_t X[4] = {x0, x1, x2, x3};
  for (uint32_t q = 1u << (B - 1); q > 1u; q >>= 1) {
    const uint32_t p = q - 1u;
    for (int i = 0; i < 4; ++i) {
      if (X[i] & q) {
        X[0] ^= p;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & p;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 4; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t q = 1u << (B - 1); q > 1u; q >>= 1)
    if (X[3] & q) t ^= q - 1u;
  for (int i = 0; i < 4; ++i) X[i] ^= t;
  uint32_t h = 0;
  for (int b = B - 1; b >= 0; --b)
    for (int i = 0; i < 4; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

// true: the member fails the test against every pose of the group.
PPF_HD bool cluster_skip(const ClWave &W, float jqx, float jqy, float jqz, float jqw, float jtx,
                         float jty, float jtz, int use_l1) {
  const float n2 = jqx * jqx + jqy * jqy + jqz * jqz + jqw * jqw;
  const float cj = W.cqx * jqx + W.cqy * jqy + W.cqz * jqz + W.cqw * jqw;
  if (n2 >= CL_NLO * CL_NLO && n2 <= CL_NHI * CL_NHI && cj < W.cos_thr) return true;
  if (!use_l1) {
    const float dx = jtx - W.ctx, dy = jty - W.cty, dz = jtz - W.ctz;
    if (dx * dx + dy * dy + dz * dz > W.b2) return true;
  }
  return false;
}

}  // namespace ppfc
