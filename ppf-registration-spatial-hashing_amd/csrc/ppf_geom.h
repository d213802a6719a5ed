// ppf_geom.h -- point-pair feature, local frames, alpha bins and poses.
//
// Host/device functions that the kernels use.  The floating-point operation
// order follows the reference's device functions exactly (citations below), so
// results are bit-identical to a literal evaluation of those functions with
// the numerics of ppf_math.h.  What differs is the organisation: the per-point
// frame T_{x->g} (first half of trans_model_scene) is computed once per point
// and reused for every vote and pose, which is bit-identical because the frame
// depends only on that point.
#pragma once

#include "ppf_math.h"

namespace ppfg {

using namespace ppfm;

struct F3 {
  float x, y, z;
};

// ((a.x*b.x + a.y*b.y) + a.z*b.z)          reference kernel.cu:51-53
PPF_HD float dot3(float ax, float ay, float az, float bx, float by, float bz) {
  return ax * bx + ay * by + az * bz;
}
// (((a0*b0 + a1*b1) + a2*b2) + a3*b3)       reference kernel.cu:55-57
PPF_HD float dot4(const float *a, float b0, float b1, float b2, float b3) {
  return a[0] * b0 + a[1] * b1 + a[2] * b2 + a[3] * b3;
}

// Quantized PPF of (p1,n1) -> (p2,n2), reference kernel.cu:109-122 and
// disc_feature kernel.cu:94-100.  nn1/nn2 are norm(n1)/norm(n2), which the
// caller may precompute (the value is the same every time).
PPF_HD void ppf_feature_q(float p1x, float p1y, float p1z, float n1x, float n1y, float n1z,
                          float nn1, float p2x, float p2y, float p2z, float n2x, float n2y,
                          float n2z, float nn2, float d_dist, float f[4]) {
  float dx = p2x - p1x, dy = p2y - p1y, dz = p2z - p1z;
  float nd = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
  float fy = acosf_own(dot3(n1x, n1y, n1z, dx, dy, dz) / (nn1 * nd));
  float fz = acosf_own(dot3(n2x, n2y, n2z, dx, dy, dz) / (nn2 * nd));
  float fw = acosf_own(dot3(n1x, n1y, n1z, n2x, n2y, n2z) / (nn1 * nn2));
  f[0] = quant_downf(nd, d_dist);
  f[1] = quant_downf(fy, D_ANGLE0);
  f[2] = quant_downf(fz, D_ANGLE0);
  f[3] = quant_downf(fw, D_ANGLE0);
}

PPF_HD float vnorm(float x, float y, float z) { return sqrtf(dot3(x, y, z, x, y, z)); }

// ---- 4x4 helpers in the reference's literal order (kernel.cu:32-49,170-299)
PPF_HD void m4_zero(float T[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i) T[i] = 0.0f;
}
PPF_HD void m4_trans(float vx, float vy, float vz, float T[16]) {
  m4_zero(T);
  T[0] = 1.0f;
  T[5] = 1.0f;
  T[10] = 1.0f;
  T[15] = 1.0f;
  T[3] = vx;
  T[7] = vy;
  T[11] = vz;
}
PPF_HD void m4_rotx(float th, float T[16]) {
  m4_zero(T);
  T[0] = 1.0f;
  T[5] = cosf_own(th);
  T[9] = sinf_own(th);
  T[6] = -1.0f * T[9];
  T[10] = T[5];
  T[15] = 1.0f;
}
PPF_HD void m4_roty(float th, float T[16]) {
  m4_zero(T);
  T[0] = cosf_own(th);
  T[2] = sinf_own(th);
  T[5] = 1.0f;
  T[8] = -1.0f * T[2];
  T[10] = T[0];
  T[15] = 1.0f;
}
PPF_HD void m4_rotz(float th, float T[16]) {
  m4_zero(T);
  T[0] = cosf_own(th);
  T[4] = sinf_own(th);
  T[1] = -1.0f * T[4];
  T[5] = T[0];
  T[10] = 1.0f;
  T[15] = 1.0f;
}
// C = A*B accumulated from zero: ((((0 + a0b0) + a1b1) + a2b2) + a3b3)
PPF_HD void m4_mul(const float A[16], const float B[16], float C[16]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float c = 0.0f;
#pragma unroll
      for (int k = 0; k < 4; ++k) c = c + A[4 * i + k] * B[4 * k + j];
      C[4 * i + j] = c;
    }
  }
}
// invht, reference kernel.cu:254-299
PPF_HD void m4_invht(const float T[16], float Ti[16]) {
  Ti[0] = T[0];
  Ti[1] = T[4];
  Ti[2] = T[8];
  Ti[4] = T[1];
  Ti[5] = T[5];
  Ti[6] = T[9];
  Ti[8] = T[2];
  Ti[9] = T[6];
  Ti[10] = T[10];
  float nr[9];
  nr[0] = -Ti[0];
  nr[1] = -Ti[1];
  nr[2] = -Ti[2];
  nr[3] = -Ti[4];
  nr[4] = -Ti[5];
  nr[5] = -Ti[6];
  nr[6] = -Ti[8];
  nr[7] = -Ti[9];
  nr[8] = -Ti[10];
  float tx = T[3], ty = T[7], tz = T[11];
  Ti[3] = dot3(nr[0], nr[1], nr[2], tx, ty, tz);
  Ti[7] = dot3(nr[3], nr[4], nr[5], tx, ty, tz);
  Ti[11] = dot3(nr[6], nr[7], nr[8], tx, ty, tz);
  Ti[12] = 0.0f;
  Ti[13] = 0.0f;
  Ti[14] = 0.0f;
  Ti[15] = 1.0f;
}

// Frame T_{x->g} of an oriented point (p, n): first half of trans_model_scene
// (reference kernel.cu:310-318), equal to compute_rot_angles + the frame part of
// compute_transforms (kernel.cu:352-395).
PPF_HD void frame_to_global(float px, float py, float pz, float nx, float ny, float nz,
                            float T[16]) {
  float transm[16], rot_y[16], rot_z[16], tmp[16];
  m4_trans(-1.0f * px, -1.0f * py, -1.0f * pz, transm);
  m4_roty(atan2f_own(nz, nx), rot_y);
  float ntx = dot4(rot_y + 0, nx, ny, nz, 1.0f);
  float nty = dot4(rot_y + 4, nx, ny, nz, 1.0f);
  m4_rotz(-1.0f * atan2f_own(nty, ntx), rot_z);
  m4_mul(rot_z, rot_y, tmp);
  m4_mul(tmp, transm, T);
}

// y and z of T * (p, 1): rows 1 and 2 of mat4f_vmul (kernel.cu:234-242).
PPF_HD void frame_apply_yz(const float *row1, const float *row2, float px, float py, float pz,
                           float &oy, float &oz) {
  oy = dot4(row1, px, py, pz, 1.0f);
  oz = dot4(row2, px, py, pz, 1.0f);
}

// Alpha bin from the y/z components of u = T_m*m_i and v = T_s*s_i with the
// x components zeroed (kernel.cu:338-342):
//   alpha = atan2(cross(u,v).x, dot(u,v)), a = quant_downf(alpha + pi, D),
//   idx = lrintf(a / D)  in 0..30.
PPF_HD uint32_t alpha_bin(float uy, float uz, float vy, float vz) {
  float cr = uy * vz - uz * vy;
  float zero = 0.0f;
  float dt = (zero * zero + uy * vy) + uz * vz;
  float a = atan2f_own(cr, dt);
  a = quant_downf(a + PI_F, D_ANGLE0);
  // NaN (non-finite inputs) gives bin 0, as lrintf(NaN) truncated to 32 bits
  // does on x86-64
  return (uint32_t)f2i_trunc(rintf(a / D_ANGLE0));
}

// Pose of one vote: invht(T_s) * rotx(alpha_idx*D - pi) * T_m
// (compute_transforms, kernel.cu:372-401).
PPF_HD void pose_from_frames(const float Tm[16], const float Ts[16], uint32_t angle_idx,
                             float T[16]) {
  float rx[16], tsi[16], tmp[16];
  m4_rotx((float)angle_idx * D_ANGLE0 - PI_F, rx);
  m4_invht(Ts, tsi);
  m4_mul(tsi, rx, tmp);
  m4_mul(tmp, Tm, T);
}

// hrotmat2quat, reference kernel.cu:128-144 (scalar in .x; divided by
// sqrt(norm(q)), as the reference does).  0.5*x in double then rounded to float
// equals 0.5f*x in float.
PPF_HD void rotmat_to_quat(const float T[16], float q[4]) {
  float t = T[0] + T[5] + T[10];
  float r = sqrtf(1.0f + t);
  q[0] = 0.5f * r;
  q[1] = copysignf(0.5f * sqrtf(1.0f + T[0] - T[5] - T[10]), T[9] - T[6]);
  q[2] = copysignf(0.5f * sqrtf(1.0f - T[0] + T[5] - T[10]), T[2] - T[8]);
  q[3] = copysignf(0.5f * sqrtf(1.0f - T[0] - T[5] + T[10]), T[4] - T[1]);
  float nq = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  float n = sqrtf(nq);
  q[0] /= n;
  q[1] /= n;
  q[2] /= n;
  q[3] /= n;
}

// Voxel cell of a translation: (int)(quant_downf(t, d)/d) per axis
// (discretize + trans2idx_kernel, kernel.cu:102-107,675-678).
PPF_HD void cell_of(float tx, float ty, float tz, float d, int32_t c[3]) {
  c[0] = f2i_trunc(quant_downf(tx, d) / d);
  c[1] = f2i_trunc(quant_downf(ty, d) / d);
  c[2] = f2i_trunc(quant_downf(tz, d) / d);
}

}  // namespace ppfg
