// Single-precision sine and cosine that return the bits of the C library the
// reference is built against: glibc's sinf / cosf / sincosf (the reference
// calls cosf/sinf at src/propagate-toz-test_OMP.cpp:633-636, 643-644,
// 652-653; gcc merges a sinf and a cosf of one argument into one sincosf
// call, and all three return the same bits).
//
// Algorithm (glibc >= 2.28, sysdeps/ieee754/flt-32/s_sinf.c, s_cosf.c,
// s_sincosf.c, sincosf.h, sincosf_data.c; Arm optimized-routines'
// single-precision sin/cos by Szabolcs Nagy): the float argument is widened
// to double and everything below is double arithmetic, rounded to float once
// at the end.  On x86-64 glibc selects at load time a build of these files
// compiled with FMA (sysdeps/x86_64/fpu/multiarch/s_sinf-fma.c and its
// siblings), so every a*b+c below is one fused multiply-add; the build without
// FMA gives other bits on 12 (sin) and 22 (cos) of the 2^32 inputs.
//
// Ranges, by the top 12 bits of the float (sign masked off: exponent and the
// first 3 mantissa bits), as glibc compares them:
//   |y| < 0.75 (the top bits of pi/4)  no reduction; |y| < 2^-12 returns
//                                      sin = y, cos = 1 exactly;
//   |y| < 120                          quadrant n = round(y * 2/pi) from the
//                                      integer part of y * 2^24*2/pi plus
//                                      2^23, shifted down 24 bits;
//                                      r = fma(-n, pi/2, y);
//   finite                             Payne-Hanek: the 24-bit significand
//                                      (shifted by the exponent's low 3 bits)
//                                      times three 32-bit windows of the
//                                      binary expansion of 2/pi, chosen by the
//                                      exponent's high bits, as a 64-bit
//                                      fixed-point number with 62 fraction
//                                      bits; n = its rounded top 2 bits;
//                                      r = (the signed rest) * pi/2^63;
//   inf, NaN                           NaN.
// On r (|r| <= pi/4 + rounding), with r2 = r*r:
//   sin r = fma(r^5, fma(r2, S3, S2), fma(r^3, S1, r))
//   cos r = fma(r^6, fma(r2, C4, C3), fma(r^4, C2, fma(r2, C1, C0)))
// where r^3 = r2*r, r^5 = r^3*r2, r^4 = r2*r2, r^6 = r^4*r2 (each one rounded
// product), and the quadrant picks sin/cos and their signs.
//
// Exhaustively equal to the host glibc 2.35 on all 2^32 float inputs
// (sinf, cosf and both outputs of sincosf) -- checked by the CPU test of
// this header compiled for the host, and the device build by the GPU test
// over the same 2^32 inputs.
#pragma once
#include <stdint.h>

namespace p2z {
namespace glibc_sincosf {

// Polynomial coefficients of glibc's sincosf_data.c (minimax fits of sin r / r
// and cos r on |r| <= pi/4 published with the algorithm).
constexpr double kC0 = 0x1p0;
constexpr double kC1 = -0x1.ffffffd0c621cp-2;
constexpr double kC2 = 0x1.55553e1068f19p-5;
constexpr double kC3 = -0x1.6c087e89a359dp-10;
constexpr double kC4 = 0x1.99343027bf8c3p-16;
constexpr double kS1 = -0x1.555545995a603p-3;
constexpr double kS2 = 0x1.1107605230bc4p-7;
constexpr double kS3 = -0x1.994eb3774cf24p-13;
constexpr double kTwoOverPiScaled = 0x1.45F306DC9C883p+23;  // 2/pi * 2^24
constexpr double kPiOver2 = 0x1.921FB54442D18p0;
constexpr double kPiOver2To63 = 0x1.921FB54442D18p-62;      // pi * 2^-63

// The binary expansion of 2/pi as overlapping 32-bit big-endian windows:
// word i holds bits 8*(i-3)+1 .. 8*(i-3)+32 after the binary point (words 0-2
// start with zero bytes).  A mathematical constant; the CPU test rebuilds it
// from an integer computation of pi.
constexpr uint32_t kInvPio4[24] = {
    0xa2u,       0xa2f9u,     0xa2f983u,   0xa2f9836eu, 0xf9836e4eu, 0x836e4e44u,
    0x6e4e4415u, 0x4e441529u, 0x441529fcu, 0x1529fc27u, 0x29fc2757u, 0xfc2757d1u,
    0x2757d1f5u, 0x57d1f534u, 0xd1f534ddu, 0xf534ddc0u, 0x34ddc0dbu, 0xddc0db62u,
    0xc0db6295u, 0xdb629599u, 0x6295993cu, 0x95993c43u, 0x993c4390u, 0x3c439041u};

__host__ __device__ inline uint32_t top12(uint32_t bits) { return (bits >> 20) & 0x7ffu; }

// Payne-Hanek reduction of a finite float with |y| >= 120 (bits: its bit
// pattern): r = the reduced argument, n = the quadrant (the low two bits of
// round(y * 2/pi)).  Kept out of line: inline, its 64-bit products would add
// their registers to every layer loop that calls sincosf_glibc, where real
// arguments never reach it.
struct LargeReduction {
  double r;
  int n;
};
__host__ __device__ __attribute__((noinline)) inline LargeReduction reduce_large(uint32_t bits) {
  const uint32_t* w = &kInvPio4[(bits >> 26) & 15u];
  const uint32_t m = ((bits & 0xffffffu) | 0x800000u) << ((bits >> 23) & 7u);
  const uint64_t lo = (uint64_t)m * w[8];
  const uint64_t mid = (uint64_t)m * w[4];
  const uint64_t hi = (uint64_t)(uint32_t)(m * w[0]);  // only its low 32 bits matter
  uint64_t f = ((lo >> 32) | (hi << 32)) + mid;
  const uint64_t n = (f + (1ull << 61)) >> 62;
  f -= n << 62;
  return LargeReduction{(double)(int64_t)f * kPiOver2To63, (int)n};
}

}  // namespace glibc_sincosf

// sin(y) and cos(y) with glibc's bits.
__host__ __device__ inline void sincosf_glibc(float y, float* sp, float* cp) {
#pragma clang fp contract(off)
  using namespace glibc_sincosf;
  const uint32_t bits = __builtin_bit_cast(uint32_t, y);
  const uint32_t top = top12(bits);
  if (top >= 0x7f8u) {  // inf or NaN
    *sp = *cp = y - y;
    return;
  }
  if (top < 0x398u) {   // |y| < 2^-12
    *sp = y;
    *cp = 1.0f;
    return;
  }
  double r;      // reduced argument, |r| <~ pi/4
  int quad = 0;  // quadrant: selects sin or cos of r
  int sgn = 0;   // sign pattern index: bit 0 negates r, bit 1 negates cos r
  if (top < 0x3f4u) {  // |y| < 0.75: no reduction
    r = (double)y;
  } else if (top < 0x42fu) {  // |y| < 120
    const double x = (double)y;
    const int n = ((int32_t)(x * kTwoOverPiScaled) + 0x800000) >> 24;
    r = __builtin_fma(-(double)n, kPiOver2, x);
    quad = n;
    sgn = n;
  } else {
    const LargeReduction red = reduce_large(bits);
    r = red.r;
    quad = red.n;
    sgn = red.n + (int)(bits >> 31);
  }
  // glibc's sign table {+1, -1, -1, +1}, indexed by sgn & 3, applied to r.
  const double rs = ((sgn + 1) & 2) ? -r : r;
  const double r2 = r * r;
  const double r3 = r2 * rs;
  const double r4 = r2 * r2;
  const double r5 = r3 * r2;
  const double r6 = r4 * r2;
  const double sin_tail = __builtin_fma(r2, kS3, kS2);
  const double sin_r = __builtin_fma(r5, sin_tail, __builtin_fma(r3, kS1, rs));
  // glibc's second coefficient table (sign indexes 2, 3) negates every cos
  // coefficient, which negates the cos polynomial exactly.
  const double cos_abs = __builtin_fma(r6, __builtin_fma(r2, kC4, kC3),
                                       __builtin_fma(r4, kC2, __builtin_fma(r2, kC1, kC0)));
  const double cos_r = (sgn & 2) ? -cos_abs : cos_abs;
  const float fs = (float)sin_r, fc = (float)cos_r;
  *sp = (quad & 1) ? fc : fs;
  *cp = (quad & 1) ? fs : fc;
}

}  // namespace p2z
