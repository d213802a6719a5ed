// ppf_math.h -- scalar numerics shared by the gfx950 kernels and the host code.
//
// Every function here is compiled twice: by hipcc for the device and by the
// host compiler (hipcc host pass or g++) for CPU code.  All arithmetic is plain
// IEEE binary32 with round-to-nearest, no FMA contraction (-ffp-contract=off on
// both sides), correctly rounded division and square root, and no flushing of
// subnormals.  Under those rules the same source produces the same bits on the
// CPU and on the GPU, which is what makes keys, vote bins and poses bit-exact
// between the device path and the CPU oracle.
//
// Contents
//   * algorithm constants (reference include/kernel.h:15-17)
//   * FNV-1a over little-endian bytes read as signed char (reference
//     src/cuda/kernel.cu:20-30)
//   * exact fmodf and quant_downf (reference src/cuda/kernel.cu:90-92)
//   * own single-precision acosf / atan2f / sinf / cosf.  They follow the
//     published Cephes Math Library single-precision method (S. L. Moshier,
//     "Methods and Programs for Mathematical Functions", 1989; netlib cephes
//     single/ atanf.c, asinf.c, sinf.c): Cody-Waite range reduction and short
//     minimax polynomials.  They replace the platform libm so both sides round
//     identically; accuracy is checked against double-precision libm in tests.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PPF_HD __host__ __device__ inline
#else
#define PPF_HD inline
#endif

namespace ppfm {

// ---------------------------------------------------------------- constants
// reference include/kernel.h:15-17 with CUDART_PI_F = 3.141592654f.
constexpr float PI_F = 3.141592654f;
constexpr int N_ANGLE = 30;
constexpr float D_ANGLE0 = (2.0f * PI_F) / float(N_ANGLE);
constexpr float ROT_THRESH = 2 * D_ANGLE0;
constexpr int N_ALPHA_BINS = N_ANGLE + 1;  // alpha_idx in 0..30, see alpha_bin()
constexpr float PIO2_F = 1.5707963267948966f;
constexpr float PIO4_F = 0.7853981633974483f;
constexpr uint32_t FNV_BASIS = 2166136261u;  // reference include/kernel.h:22
constexpr uint32_t FNV_PRIME = 16777619u;    // reference src/cuda/kernel.cu:27
constexpr uint32_t CANON_NAN_BITS = 0x7fc00000u;

PPF_HD uint32_t fbits(float f) { return __builtin_bit_cast(uint32_t, f); }
PPF_HD float bitsf(uint32_t u) { return __builtin_bit_cast(float, u); }
PPF_HD bool is_nan(float f) { return (fbits(f) & 0x7fffffffu) > 0x7f800000u; }
PPF_HD float qnan() { return bitsf(CANON_NAN_BITS); }
PPF_HD bool sign_bit(float f) { return (fbits(f) >> 31) != 0u; }

// ------------------------------------------------------------------- FNV-1a
// The reference walks the object through a `char *`, so every byte is
// sign-extended to 32 bits before the XOR (bytes >= 0x80 flip the top 24 bits).
PPF_HD uint32_t fnv1a_word(uint32_t h, uint32_t w) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    int8_t b = (int8_t)(uint8_t)(w >> (8 * i));
    h ^= (uint32_t)(int32_t)b;
    h *= FNV_PRIME;
  }
  return h;
}

// Key of a quantized feature (reference src/cuda/kernel.cu:460-477):
// 0 when .x is NaN, otherwise FNV-1a over the 16 bytes of the float4.  NaN in
// .y/.z/.w (acos of a ratio outside [-1,1], or 0/0 for coincident points) is
// hashed with the canonical quiet-NaN pattern so both sides agree.
PPF_HD uint32_t ppf_key(float fx, float fy, float fz, float fw) {
  if (is_nan(fx)) return 0u;
  uint32_t h = FNV_BASIS;
  h = fnv1a_word(h, fbits(fx));
  h = fnv1a_word(h, is_nan(fy) ? CANON_NAN_BITS : fbits(fy));
  h = fnv1a_word(h, is_nan(fz) ? CANON_NAN_BITS : fbits(fz));
  h = fnv1a_word(h, is_nan(fw) ? CANON_NAN_BITS : fbits(fw));
  return h;
}

// FNV-1a of an int3 (reference src/cuda/kernel.cu:680,693).
PPF_HD uint32_t cell_hash(int32_t x, int32_t y, int32_t z) {
  uint32_t h = FNV_BASIS;
  h = fnv1a_word(h, (uint32_t)x);
  h = fnv1a_word(h, (uint32_t)y);
  h = fnv1a_word(h, (uint32_t)z);
  return h;
}

// ------------------------------------------------------------- exact fmodf
// fmod is exact by definition, so any correct method gives the same bits.
// General case: binary long division on the significands.
PPF_HD float fmodf_bitwise(float x, float y) {
  uint32_t ux = fbits(x), uy = fbits(y);
  uint32_t sx = ux & 0x80000000u;
  uint32_t ax = ux & 0x7fffffffu, ay = uy & 0x7fffffffu;
  if (ay == 0 || ax >= 0x7f800000u || ay > 0x7f800000u) return qnan();
  if (ax < ay) return x;
  int ex = (int)(ax >> 23), ey = (int)(ay >> 23);
  uint32_t mx, my;
  if (ex == 0) {  // subnormal x: normalise
    mx = ax;
    ex = 1;
    while (!(mx & 0x800000u)) { mx <<= 1; --ex; }
  } else {
    mx = (ax & 0x7fffffu) | 0x800000u;
  }
  if (ey == 0) {
    my = ay;
    ey = 1;
    while (!(my & 0x800000u)) { my <<= 1; --ey; }
  } else {
    my = (ay & 0x7fffffu) | 0x800000u;
  }
  for (; ex > ey; --ex) {
    if (mx >= my) {
      mx -= my;
      if (mx == 0) return bitsf(sx);
    }
    mx <<= 1;
  }
  if (mx >= my) {
    mx -= my;
    if (mx == 0) return bitsf(sx);
  }
  while (!(mx & 0x800000u)) { mx <<= 1; --ex; }
  uint32_t r;
  if (ex > 0) {
    r = (mx - 0x800000u) | ((uint32_t)ex << 23);
  } else {
    r = mx >> (1 - ex);  // subnormal result (exact: low bits are zero)
  }
  return bitsf(r | sx);
}

// Fast exact fmod: k = trunc(|x|/|y|) is off by at most one when the quotient
// is below 2^22, and fma(-k, |y|, |x|) is exact whenever k is right (the true
// remainder is representable); a wrong k shows up as r < 0 or r >= |y|.
PPF_HD float fmodf_exact(float x, float y) {
  float ax = fabsf(x), ay = fabsf(y);
  if (!(ay > 0.0f) || !(ax < INFINITY) || is_nan(y)) return qnan();
  if (ax < ay) return x;
  float q = ax / ay;
  if (q >= 4194304.0f) return fmodf_bitwise(x, y);
  float k = truncf(q);
  float r = fmaf(-k, ay, ax);
  if (r < 0.0f) {
    r = fmaf(-(k - 1.0f), ay, ax);
  } else if (r >= ay) {
    r = fmaf(-(k + 1.0f), ay, ax);
  }
  return copysignf(r, x);
}

// reference src/cuda/kernel.cu:90-92
PPF_HD float quant_downf(float x, float y) { return x - fmodf_exact(x, y); }

// -------------------------------------------------------- transcendentals
// atan on [0, inf): Cephes atanf range reduction at tan(3pi/8), tan(pi/8).
PPF_HD float atanf_pos(float x) {
  float y0;
  if (x > 2.414213562373095f) {
    y0 = PIO2_F;
    x = -1.0f / x;
  } else if (x > 0.4142135623730950f) {
    y0 = PIO4_F;
    x = (x - 1.0f) / (x + 1.0f);
  } else {
    y0 = 0.0f;
  }
  float z = x * x;
  float p = (((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z -
             3.33329491539e-1f) * z * x + x;
  return y0 + p;
}

PPF_HD float atanf_own(float x) {
  if (is_nan(x)) return x;
  return x < 0.0f ? -atanf_pos(-x) : atanf_pos(x);
}

// atan2 with the C99 / IEEE special cases (signed zeros and infinities).
PPF_HD float atan2f_own(float y, float x) {
  if (is_nan(x) || is_nan(y)) return qnan();
  bool ny = sign_bit(y);
  bool nx = sign_bit(x);
  if (y == 0.0f) {
    if (!nx) return y;  // atan2(+-0, +0 or +x) = +-0
    return ny ? -PI_F : PI_F;
  }
  if (x == 0.0f) return ny ? -PIO2_F : PIO2_F;
  bool ix = fabsf(x) == INFINITY, iy = fabsf(y) == INFINITY;
  if (ix) {
    float r;
    if (iy)
      r = nx ? 3.0f * PIO4_F : PIO4_F;
    else
      r = nx ? PI_F : 0.0f;
    return ny ? -r : r;
  }
  if (iy) return ny ? -PIO2_F : PIO2_F;
  float z = atanf_own(y / x);
  if (nx) z = ny ? z - PI_F : z + PI_F;
  return z;
}

// asin on [0, 1] (Cephes asinf).
PPF_HD float asinf_pos(float a) {
  float x, z;
  bool flag;
  if (a > 0.5f) {
    z = 0.5f * (1.0f - a);
    x = sqrtf(z);
    flag = true;
  } else {
    x = a;
    z = x * x;
    flag = false;
  }
  z = ((((4.2163199048e-2f * z + 2.4181311049e-2f) * z + 4.5470025998e-2f) * z +
        7.4953002686e-2f) * z + 1.6666752422e-1f) * z * x + x;
  if (flag) {
    z = z + z;
    z = PIO2_F - z;
  }
  return z;
}

PPF_HD float asinf_own(float x) {
  if (is_nan(x)) return qnan();
  float a = fabsf(x);
  if (a > 1.0f) return qnan();
  float r = asinf_pos(a);
  return x < 0.0f ? -r : r;
}

// Cephes acosf.
PPF_HD float acosf_own(float x) {
  if (is_nan(x)) return qnan();
  if (x < -1.0f || x > 1.0f) return qnan();
  if (x < -0.5f) return PI_F - 2.0f * asinf_pos(sqrtf(0.5f * (1.0f + x)));
  if (x > 0.5f) return 2.0f * asinf_pos(sqrtf(0.5f * (1.0f - x)));
  return PIO2_F - asinf_own(x);
}

// Cephes sinf/cosf: octant j from 4/pi, three-part Cody-Waite reduction.
constexpr float FOPI = 1.27323954473516f;
constexpr float DP1 = 0.78515625f;
constexpr float DP2 = 2.4187564849853515625e-4f;
constexpr float DP3 = 3.77489497744594108e-8f;

PPF_HD float sin_poly(float x, float z) {
  return ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x + x;
}
PPF_HD float cos_poly(float z) {
  return ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z -
         0.5f * z + 1.0f;
}

PPF_HD float sinf_own(float xx) {
  if (is_nan(xx) || fabsf(xx) == INFINITY) return qnan();
  if (xx == 0.0f) return xx;  // sin(+-0) = +-0
  bool neg = false;
  float x = xx;
  if (x < 0.0f) {
    neg = true;
    x = -x;
  }
  int j = (int)(FOPI * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  if (j > 3) {
    neg = !neg;
    j -= 4;
  }
  x = ((x - y * DP1) - y * DP2) - y * DP3;
  float z = x * x;
  float r = (j == 1 || j == 2) ? cos_poly(z) : sin_poly(x, z);
  return neg ? -r : r;
}

PPF_HD float cosf_own(float xx) {
  if (is_nan(xx) || fabsf(xx) == INFINITY) return qnan();
  float x = fabsf(xx);
  int j = (int)(FOPI * x);
  float y = (float)j;
  if (j & 1) {
    j += 1;
    y += 1.0f;
  }
  j &= 7;
  bool neg = false;
  if (j > 3) {
    j -= 4;
    neg = !neg;
  }
  if (j > 1) neg = !neg;
  x = ((x - y * DP1) - y * DP2) - y * DP3;
  float z = x * x;
  float r = (j == 1 || j == 2) ? sin_poly(x, z) : cos_poly(z);
  return neg ? -r : r;
}

// float -> int truncation with a defined result for NaN (0) and saturation
// outside the int32 range, identical on both sides (C leaves these undefined).
PPF_HD int32_t f2i_trunc(float x) {
  if (is_nan(x)) return 0;
  if (x >= 2147483648.0f) return 2147483647;
  if (x <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)x;
}

}  // namespace ppfm
