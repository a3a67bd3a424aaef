// Per-track (per-lane) device math of the propagate-to-z + Kalman-update path.
//
// One GPU lane owns one track: its 6 parameters (x, y, z, ipt, phi, theta), its
// packed symmetric 6x6 covariance P[21] (lower triangle, row-major) and its
// charge stay in VGPRs for the whole layer loop.  Three arithmetic modes:
//
//   kRef      : the reference CPU typing (src/propagate-toz-test_OMP.cpp:
//               propagateToZ 623-682, MultHelixPropEndcap 296-341,
//               MultHelixPropTranspEndcap 343-373, KalmanUpdate_v2 503-620):
//               FP32 with the FP64 sub-expressions the reference's `1.`
//               literals and `(double)` casts create, no contraction, the
//               reference's summation order, and sin/cos with the bits of
//               the C library the reference calls (glibc's sinf/cosf,
//               p2z_libm.hpp).  It reproduces the reference build bit for
//               bit.
//   kFast     : FP32 with FMA contraction, native reciprocals and the sparse
//               structure of the Jacobian and of the covariance exploited.
//   kFp64Gain : kFast propagation; the 2x2 residual-covariance inverse, the
//               gain, the parameter update and a Joseph-form covariance update
//               (I-KH) P (I-KH)^T + K R K^T in FP64.
//
// Structure used by kFast / kFp64Gain (exact for finite inputs):
//   * the propagation Jacobian Jt is the identity plus 11 entries with row 2
//     (z) zero, so after any propagation the z row/column of P is exactly 0;
//   * the update therefore sees P(2,.) = 0, its gain row 2 is 0 and it leaves
//     z and P(2,.) unchanged.
// Only the first propagation of a track can see a non-zero z row (its input
// covariance); later layers use the 15-entry form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "p2z_cov.hpp"
#include "p2z_libm.hpp"

namespace p2z {

enum Math : int { kFast = 0, kRef = 1, kFp64Gain = 2 };

// kfact as the reference declares it: a float initialised from a double
// expression (OMP:622).
constexpr float kKfact = 100. / (-0.299792458 * 3.8112);


struct Track {
  float p[6];   // x y z ipt phi theta
  float P[21];  // packed lower triangle
  int q;
};

struct Hit {
  float x, y, z;     // position
  float xx, xy, yy;  // covariance elements 0..2 (xz, yz, zz are never read)
};

// ---------------------------------------------------------------------------
// kRef: reference typing
// ---------------------------------------------------------------------------

__device__ inline void propagate_ref(Track& t, float hz) {
#pragma clang fp contract(off)
  float* p = t.p;
  float* P = t.P;
  const float zout = hz;
  const float k = t.q * kKfact;
  const float deltaZ = zout - p[2];
  const float ipt = p[3];
  const float pt = (float)(1. / (double)ipt);
  float sinP, cosP, sinT, cosT;
  sincosf_glibc(p[4], &sinP, &cosP);
  sincosf_glibc(p[5], &sinT, &cosT);
  const float pxin = cosP * pt;
  const float pyin = sinP * pt;
  const float icosT = (float)(1.0 / (double)cosT);
  const float icosTk = icosT / k;
  const float alpha = deltaZ * sinT * ipt * icosTk;
  float sina, cosa;
  sincosf_glibc(alpha, &sina, &cosa);
  p[0] = (float)((double)p[0] + (double)k * ((double)(pxin * sina) - (double)pyin * (1. - (double)cosa)));
  p[1] = (float)((double)p[1] + (double)k * ((double)(pyin * sina) + (double)pxin * (1. - (double)cosa)));
  p[2] = zout;
  p[4] = p[4] + alpha;

  float sC, cC;
  sincosf_glibc(cosP * sina, &sC, &cC);
  const double omcC = 1. - (double)cC;
  const float J02 = cosP * sinT * (sinP * cosa * sC - cosa) * icosT;
  const float J03 = (float)((double)(cosP * sinT * deltaZ * cosa) * (1. - (double)(sinP * sC)) * (double)(icosT * pt)
                            - (double)k * ((double)(cosP * sina) - (double)sinP * omcC) * (double)(pt * pt));
  const float J04 = (float)((double)(k * pt) * ((double)(-sinP * sina + sinP * sinP * sina * sC) - (double)cosP * omcC));
  const float J05 = (float)((double)(cosP * deltaZ * cosa) * (1. - (double)(sinP * sC)) * (double)(icosT * icosT));
  const float J12 = cosa * sinT * (cosP * cosP * sC - sinP) * icosT;
  const float J13 = (float)((double)(sinT * deltaZ * cosa * (cosP * cosP * sC + sinP) * (icosT * pt))
                            - (double)k * ((double)(sinP * sina) + (double)cosP * omcC) * (double)(pt * pt));
  const float J14 = (float)((double)(k * pt) * (-(double)sinP * omcC - (double)(sinP * cosP * sina * sC)
                                                + (double)(cosP * sina)));
  const float J15 = deltaZ * cosa * (cosP * cosP * sC + sinP) * (icosT * icosT);
  const float J42 = -ipt * sinT * icosTk;
  const float J43 = sinT * deltaZ * icosTk;
  const float J45 = ipt * deltaZ * (icosT * icosTk);

  // T = Jt P: rows 0, 1, 4 summed in the reference's order; row 2 is zero,
  // rows 3 and 5 copy P.
  float T0[6], T1[6], T4[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const float p0 = P[S(0, c)], p1 = P[S(1, c)], p2 = P[S(2, c)];
    const float p3 = P[S(3, c)], p4 = P[S(4, c)], p5 = P[S(5, c)];
    T0[c] = p0 + J02 * p2 + J03 * p3 + J04 * p4 + J05 * p5;
    T1[c] = p1 + J12 * p2 + J13 * p3 + J14 * p4 + J15 * p5;
    T4[c] = J42 * p2 + J43 * p3 + p4 + J45 * p5;
  }
  float T3[6], T5[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T3[c] = P[S(3, c)];
    T5[c] = P[S(5, c)];
  }
  // P' = T Jt^T.  Entries (2,0) and (2,1) are sums over the zero row of T:
  // they are 0, or NaN when a Jacobian entry is not finite, as in the reference.
  const float z0 = 0.f;
  P[S(0, 0)] = T0[0] + T0[2] * J02 + T0[3] * J03 + T0[4] * J04 + T0[5] * J05;
  P[S(1, 0)] = T1[0] + T1[2] * J02 + T1[3] * J03 + T1[4] * J04 + T1[5] * J05;
  P[S(1, 1)] = T1[1] + T1[2] * J12 + T1[3] * J13 + T1[4] * J14 + T1[5] * J15;
  P[S(2, 0)] = z0 + z0 * J02 + z0 * J03 + z0 * J04 + z0 * J05;
  P[S(2, 1)] = z0 + z0 * J12 + z0 * J13 + z0 * J14 + z0 * J15;
  P[S(2, 2)] = 0.f;
  P[S(3, 0)] = T3[0] + T3[2] * J02 + T3[3] * J03 + T3[4] * J04 + T3[5] * J05;
  P[S(3, 1)] = T3[1] + T3[2] * J12 + T3[3] * J13 + T3[4] * J14 + T3[5] * J15;
  P[S(3, 2)] = 0.f;
  P[S(3, 3)] = T3[3];
  P[S(4, 0)] = T4[0] + T4[2] * J02 + T4[3] * J03 + T4[4] * J04 + T4[5] * J05;
  P[S(4, 1)] = T4[1] + T4[2] * J12 + T4[3] * J13 + T4[4] * J14 + T4[5] * J15;
  P[S(4, 2)] = 0.f;
  P[S(4, 3)] = T4[3];
  P[S(4, 4)] = T4[2] * J42 + T4[3] * J43 + T4[4] + T4[5] * J45;
  P[S(5, 0)] = T5[0] + T5[2] * J02 + T5[3] * J03 + T5[4] * J04 + T5[5] * J05;
  P[S(5, 1)] = T5[1] + T5[2] * J12 + T5[3] * J13 + T5[4] * J14 + T5[5] * J15;
  P[S(5, 2)] = 0.f;
  P[S(5, 3)] = T5[3];
  P[S(5, 4)] = T5[2] * J42 + T5[3] * J43 + T5[4] + T5[5] * J45;
  P[S(5, 5)] = T5[5];
}

__device__ inline void update_ref(Track& t, const Hit& h) {
#pragma clang fp contract(off)
  float* p = t.p;
  float* P = t.P;
  float r00 = P[0] + h.xx;
  float r01 = P[1] + h.xy;
  float r11 = P[2] + h.yy;
  const double det = (double)r00 * (double)r11 - (double)r01 * (double)r01;
  const float s = (float)(1. / det);
  const float tmp = s * r11;
  r01 = r01 * -s;
  r11 = s * r00;
  r00 = tmp;
  float K0[6], K1[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const float a = P[S(i, 0)], b = P[S(i, 1)];
    K0[i] = a * r00 + b * r01;
    K1[i] = a * r01 + b * r11;
  }
  const float res0 = h.x - p[0];
  const float res1 = h.y - p[1];
#pragma unroll
  for (int i = 0; i < 6; ++i) p[i] = p[i] + K0[i] * res0 + K1[i] * res1;
  float C[21];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) C[S(i, j)] = K0[i] * P[S(0, j)] + K1[i] * P[S(1, j)];
#pragma unroll
  for (int e = 0; e < 21; ++e) P[e] = P[e] - C[e];
}

// ---------------------------------------------------------------------------
// kFast: FP32 + FMA, sparse forms
// ---------------------------------------------------------------------------

// sin and cos of a float, rounded to the nearest float except in ~1e-5 of
// cases, at about the cost of the library's single-precision sincosf (FP64
// fused multiply-adds issue at the FP32 rate on CDNA4):
//   k = rint(x * 2/pi), r = x - k*pi/2 with pi/2 split into two doubles,
//   sin r = r + r^3 S(r^2) and cos r = 1 + r^2 C(r^2) in FP64 with S, C the
//   least-squares polynomials of gen_sincos_coeffs.py (max relative error on
//   |r| <= pi/4: sin 1.916e-14, cos 3.129e-13), then the quadrant k mod 4 selects and
//   signs the two values.
// The reduction error is |k| * 1.5e-33 (the next term of pi/2), far below the
// polynomial error while k is exact, i.e. while x * 2/pi rounds to the right
// quadrant: true with a wide margin for |x| < 2^20 (k < 2^20; the product's
// rounding error stays below 2^-32).  Larger |x| (never a track angle) take
// the C-library-faithful sincosf_glibc, whose Payne-Hanek reduction is exact
// at any magnitude; so do infinity and NaN, which give NaN there (the
// reduction below never sees a NaN, whose conversion to an integer would be
// undefined).
// The library's sincosf is faithful only (measured: 10-25% of its results at
// arguments like the track angles phi ~ 1.4 and theta ~ 2.8 are one ulp from
// the nearest float); the tracks of an event share nearly the same angles, so
// those ulps shift every track alike instead of averaging out.
// A double constant placed in scalar registers where it is used.  FP64
// instructions take no literal operands, so the compiler materialises such
// constants in registers; left to itself it puts them in vector register pairs
// and, in the persistent ragged kernel, keeps them live across its whole item
// loop (16 VGPRs for the polynomials below, which run once per track).  The
// empty asm leaves the value unchanged.
__device__ __forceinline__ double sgpr_const(double v) {
  asm volatile("" : "+s"(v));
  return v;
}

__device__ inline void sincos_accurate(float xf, float* sp, float* cp) {
  if (!(__builtin_fabsf(xf) < 0x1p20f)) {  // large, infinite or NaN
    sincosf_glibc(xf, sp, cp);
    return;
  }
  const double x = (double)xf;
  const double k = __builtin_rint(x * sgpr_const(0.63661977236758134308));
  const double r = __builtin_fma(-k, sgpr_const(6.123233995736766036e-17),
                                 __builtin_fma(-k, sgpr_const(1.5707963267948965580), x));
  const double u = r * r;
  double ps = sgpr_const(-2.480613491641723e-08);
  ps = __builtin_fma(ps, u, sgpr_const(2.755599789925388e-06));
  ps = __builtin_fma(ps, u, sgpr_const(-0.0001984126694941195));
  ps = __builtin_fma(ps, u, sgpr_const(0.008333333331134977));
  ps = __builtin_fma(ps, u, sgpr_const(-0.16666666666664143));
  const double sr = __builtin_fma(r * u, ps, r);
  double pc = sgpr_const(-2.7237166798569765e-07);
  pc = __builtin_fma(pc, u, sgpr_const(2.4799862612840642e-05));
  pc = __builtin_fma(pc, u, sgpr_const(-0.0013888885093552119));
  pc = __builtin_fma(pc, u, sgpr_const(0.04166666663743458));
  pc = __builtin_fma(pc, u, sgpr_const(-0.4999999999996406));
  const double cr = __builtin_fma(u, pc, 1.0);
  // quadrant q = k mod 4; rounding commutes with the swap and the signs
  const int q = (int)(long long)k & 3;
  const float sf = (float)sr, cf = (float)cr;
  const float s = (q & 1) ? cf : sf;
  const float c = (q & 1) ? sf : cf;
  *sp = (q & 2) ? -s : s;
  *cp = ((q + 1) & 2) ? -c : c;
}

// 1/x rounded to nearest in all but rare cases: the hardware reciprocal
// (1 ulp) refined by one Newton step with a fused multiply-add.  Every track
// of an event shares nearly the same ipt / cos(theta); a biased reciprocal
// would shift all of them the same way, so the 1-ulp estimate alone is not
// used.
__device__ inline float fast_rcp(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

// a / b through the refined reciprocal plus one residual correction.
__device__ inline float fast_div(float a, float b) {
  const float rb = fast_rcp(b);
  const float q = a * rb;
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), rb, q);
}

// sin x, cos x and 1 - cos x for |x| <= 1, in FP32 without argument
// reduction:
//   sin x = x + x^3 S(x^2),  cos x = 1 + x^2 C(x^2),  1 - cos x = -x^2 C(x^2)
// with S, C the cubic least-squares fits of gen_sincos_coeffs.py --unit,
// rounded to float.  Checked on every float in [-1, 1]: sin within 0.83 ulp
// (0.06% not the nearest float), cos within 1.02 ulp (0.12%), 1 - cos within
// 2.1 ulp of its own value (no cancellation, unlike 1 - cosf(x)).  About a
// third of the library sincosf's cost, which also handles range reduction.
__device__ inline void sincos_unit(float x, float* sp, float* cp, float* omcp) {
  const float u = x * x;
  const float t = u * x;
  float ps = __builtin_fmaf(0x1.6b33ecp-19f, u, -0x1.a00964p-13f);
  ps = __builtin_fmaf(ps, u, 0x1.111104p-7f);
  ps = __builtin_fmaf(ps, u, -0x1.555556p-3f);
  *sp = __builtin_fmaf(t, ps, x);
  float pc = __builtin_fmaf(0x1.96f3acp-16f, u, -0x1.6bfff6p-10f);
  pc = __builtin_fmaf(pc, u, 0x1.55553p-5f);
  pc = __builtin_fmaf(pc, u, -0.5f);
  *omcp = -(u * pc);
  *cp = __builtin_fmaf(u, pc, 1.0f);
}

// sincos_unit where |x| <= 1 (the common case for the turning angle of one
// layer), the library's reduced-argument sincosf elsewhere.  NaN takes the
// library branch.
// The library's sincosf, kept out of line: inline, its reduction of large
// arguments (64-bit products) would add its registers to the register peak of
// every layer loop that calls it, although track angles never need it.
__device__ __noinline__ float2 sincosf_outline(float x) {
  float s, c;
  sincosf(x, &s, &c);
  return make_float2(s, c);
}

__device__ inline void sincos_short(float x, float* sp, float* cp, float* omcp) {
  if (__builtin_fabsf(x) <= 1.0f) {
    sincos_unit(x, sp, cp, omcp);
  } else {
    const float2 v = sincosf_outline(x);
    *sp = v.x;
    *cp = v.y;
    *omcp = 1.0f - *cp;
  }
}

// sin x and 1 - cos x for |x| < 2^-8 (the update's corrections of the track
// angles are typically ~1e-4): x - x^3/6 (next term x^5/120 < 2^-46 |x|) and
// x^2/2 (next term x^4/24 < 2^-35, absolute).  The latter enters only as a
// product with a sine or cosine added to one, so its absolute error is what
// counts.
__device__ inline void sincos_tiny(float x, float* sp, float* omcp) {
  const float u = x * x;
  *sp = __builtin_fmaf(u * x, -0x1.555556p-3f, x);
  *omcp = 0.5f * u;
}

// Rotation increments for a small angle: sincos_tiny below 2^-8, sincos_short
// (sincos_unit or the library) above.
__device__ inline void sincos_increment(float x, float* sp, float* omcp) {
  if (__builtin_fabsf(x) < 0x1p-8f) {
    sincos_tiny(x, sp, omcp);
  } else {
    float c;
    sincos_short(x, sp, &c, omcp);
  }
}

// The 11 Jacobian entries of one track's step.

// sin and cos of a track's phi and theta.  The fused layer loops carry them
// from layer to layer, rotating them by each change of the angles (alpha in
// the propagation, the update's corrections) instead of evaluating sin/cos of
// the updated float angles again: the rotation tracks the exact sum of the
// increments, where sin of the rounded angle carries that angle's rounding
// (measured over 40 steps: median 1.3 ulp either way, p99 4.8 ulp rotating
// against 34 ulp recomputing, the latter near the zeros of sin).
struct Trig {
  float sinP, cosP, sinT, cosT;
};

// kAccurateTrig: sincos_accurate (correctly rounded but for rare cases);
// otherwise the library's sincosf.
template <bool kAccurateTrig>
__device__ inline void trig_init(const Track& t, Trig& g) {
  if constexpr (kAccurateTrig) {
    sincos_accurate(t.p[4], &g.sinP, &g.cosP);
    sincos_accurate(t.p[5], &g.sinT, &g.cosT);
  } else {
    const float2 a = sincosf_outline(t.p[4]), b = sincosf_outline(t.p[5]);
    g.sinP = a.x;
    g.cosP = a.y;
    g.sinT = b.x;
    g.cosT = b.y;
  }
}

// (s, c) = (sin x, cos x) -> (sin, cos)(x + d), given sin d and 1 - cos d:
//   sin(x + d) = s + (c sin d - s (1 - cos d)),  cos(x + d) = c - (s sin d + c (1 - cos d))
// (small corrections added to the values, no product by cos d).
__device__ inline void rotate(float& s, float& c, float sd, float omcd) {
#pragma clang fp contract(fast)
  const float hs = c * sd - s * omcd;
  const float hc = s * sd + c * omcd;
  s = s + hs;
  c = c - hc;
}

// Helix step and Jacobian (OMP:629-666) in FP32 with contraction, with the
// Jacobian regrouped on shared products.  With sC = sin(cosP sina),
// omcC = 1 - cos(cosP sina), B = 1 - sinP sC, C = cosP^2 sC + sinP,
// E = deltaZ cosa / cosT, G = E sinT pt, kpt = k pt:
//   J02 = -cosP B sinT cosa / cosT       J12 = cosa sinT (cosP^2 sC - sinP) / cosT
//   J03 = cosP B G - kpt pt (cosP sina - sinP omcC)
//   J13 = C G      - kpt pt (sinP sina + cosP omcC)
//   J04 = kpt (-sinP sina B - cosP omcC) J14 = kpt (cosP sina B - sinP omcC)
//   J05 = cosP B E / cosT                J15 = C E / cosT
//   J42 = -ipt sinT icosTk, J43 = sinT deltaZ icosTk, J45 = ipt deltaZ icosT icosTk
// which is the reference's expressions with the common factors taken out
// (e.g. -sinP sina + sinP^2 sina sC = -sinP sina (1 - sinP sC)); likewise
//   x' = x + kpt (cosP sina - sinP (1 - cosa)),  y' = y + kpt (sinP sina + cosP (1 - cosa)),
// whose brackets are also the changes of sin phi and cos phi under phi' = phi + alpha
// (g is rotated with them).  The two small angles (alpha and cosP sina,
// |cosP sina| <= 1 always) use sincos_short / sincos_unit, which give 1 - cos
// without cancellation.
// kFirst: J02, J12, J42 are needed (the input z row can be non-zero).
template <bool kFirst>
__device__ inline Jac helix_core(Track& t, float hz, Trig& g) {
#pragma clang fp contract(fast)
  float* p = t.p;
  const float sinP = g.sinP, cosP = g.cosP, sinT = g.sinT, cosT = g.cosT;
  const float k = (float)t.q * kKfact;
  const float rk = fast_rcp(k);
  const float deltaZ = hz - p[2];
  const float ipt = p[3];
  const float pt = fast_rcp(ipt);
  const float icosT = fast_rcp(cosT);
  const float dzc = deltaZ * icosT;
  const float dzk = dzc * rk;
  const float J43 = sinT * dzk;
  const float alpha = J43 * ipt;
  float sina, cosa, omca;
  sincos_short(alpha, &sina, &cosa, &omca);
  const float csa = cosP * sina;
  const float ssa = sinP * sina;
  const float kpt = k * pt;
  const float hx = csa - sinP * omca;
  const float hy = ssa + cosP * omca;
  p[0] = p[0] + kpt * hx;
  p[1] = p[1] + kpt * hy;
  p[2] = hz;
  p[4] = p[4] + alpha;
  g.sinP = sinP + hx;
  g.cosP = cosP - hy;

  float sC, cC, omcC;
  sincos_unit(csa, &sC, &cC, &omcC);
  const float B = __builtin_fmaf(-sinP, sC, 1.0f);
  const float cc = cosP * cosP;
  const float C = __builtin_fmaf(cc, sC, sinP);
  const float E = dzc * cosa;
  const float F = E * icosT;
  const float G = E * sinT * pt;
  const float kpp = kpt * pt;
  const float cO = cosP * omcC, sO = sinP * omcC;
  const float cB = cosP * B;
  Jac J;
  J.J03 = cB * G - kpp * (csa - sO);
  J.J13 = C * G - kpp * (ssa + cO);
  J.J04 = kpt * (-ssa * B - cO);
  J.J14 = kpt * (csa * B - sO);
  J.J05 = cB * F;
  J.J15 = C * F;
  J.J43 = J43;
  J.J45 = ipt * dzk * icosT;
  if constexpr (kFirst) {
    const float sca = sinT * cosa * icosT;
    J.J02 = -cB * sca;
    J.J12 = (cc * sC - sinP) * sca;
    J.J42 = -ipt * sinT * (icosT * rk);
  } else {
    J.J02 = J.J12 = J.J42 = 0.f;
  }
  return J;
}

// One step from the track's own angles (the per-function entry points).
// kAccurateTrig: sin/cos of phi and theta through sincos_accurate.
template <bool kAccurateTrig, bool kFirst = true>
__device__ inline Jac helix_fast(Track& t, float hz) {
  Trig g;
  trig_init<kAccurateTrig>(t, g);
  return helix_core<kFirst>(t, hz, g);
}


// P' = Jt P Jt^T.  kZRow0: the input z row/column is known to be zero.
template <bool kZRow0>
__device__ inline void propagate_cov_fast(float* P, const Jac& J) {
#pragma clang fp contract(fast)
  if constexpr (kZRow0) {
    propagate_cov_fast_z0(P, J);
    return;
  }
  // rows 0, 1, 4 of T = Jt P (rows 3 and 5 are rows of P; row 2 is zero).
  // Column 2 of T is only needed when P's z column can be non-zero, and then
  // only for the (i, j) sums below through T[i][2] * J.j2.
  float T0[6], T1[6], T4[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    if (kZRow0 && c == 2) {
      T0[c] = T1[c] = T4[c] = 0.f;
      continue;
    }
    // the reference's order (OMP:304-339); with a zero z row the p2 terms
    // are exact zeros and are dropped
    const float p0 = P[S(0, c)], p1 = P[S(1, c)];
    const float p3 = P[S(3, c)], p4 = P[S(4, c)], p5 = P[S(5, c)];
    if (kZRow0) {
      T0[c] = p0 + J.J03 * p3 + J.J04 * p4 + J.J05 * p5;
      T1[c] = p1 + J.J13 * p3 + J.J14 * p4 + J.J15 * p5;
      T4[c] = J.J43 * p3 + p4 + J.J45 * p5;
    } else {
      const float p2 = P[S(2, c)];
      T0[c] = p0 + J.J02 * p2 + J.J03 * p3 + J.J04 * p4 + J.J05 * p5;
      T1[c] = p1 + J.J12 * p2 + J.J13 * p3 + J.J14 * p4 + J.J15 * p5;
      T4[c] = J.J42 * p2 + J.J43 * p3 + p4 + J.J45 * p5;
    }
  }
  float T3[6], T5[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T3[c] = (kZRow0 && c == 2) ? 0.f : P[S(3, c)];
    T5[c] = (kZRow0 && c == 2) ? 0.f : P[S(5, c)];
  }
  // (i, 0): Ti0 + Ti2 J02 + Ti3 J03 + Ti4 J04 + Ti5 J05 ; (i, 1) likewise with
  // row 1 of Jt; (i, 3) = Ti3; (i, 4) = Ti2 J42 + Ti3 J43 + Ti4 + Ti5 J45.
#define P2Z_ROW0(T) ((kZRow0 ? T[0] : T[0] + T[2] * J.J02) + T[3] * J.J03 + T[4] * J.J04 + T[5] * J.J05)
#define P2Z_ROW1(T) ((kZRow0 ? T[1] : T[1] + T[2] * J.J12) + T[3] * J.J13 + T[4] * J.J14 + T[5] * J.J15)
#define P2Z_ROW4(T) ((kZRow0 ? T[3] * J.J43 : T[2] * J.J42 + T[3] * J.J43) + T[4] + T[5] * J.J45)
  P[S(0, 0)] = P2Z_ROW0(T0);
  P[S(1, 0)] = P2Z_ROW0(T1);
  P[S(1, 1)] = P2Z_ROW1(T1);
  P[S(3, 0)] = P2Z_ROW0(T3);
  P[S(3, 1)] = P2Z_ROW1(T3);
  P[S(3, 3)] = T3[3];
  P[S(4, 0)] = P2Z_ROW0(T4);
  P[S(4, 1)] = P2Z_ROW1(T4);
  P[S(4, 3)] = T4[3];
  P[S(4, 4)] = P2Z_ROW4(T4);
  P[S(5, 0)] = P2Z_ROW0(T5);
  P[S(5, 1)] = P2Z_ROW1(T5);
  P[S(5, 3)] = T5[3];
  P[S(5, 4)] = P2Z_ROW4(T5);
  P[S(5, 5)] = T5[5];
#undef P2Z_ROW0
#undef P2Z_ROW1
#undef P2Z_ROW4
  P[S(2, 0)] = 0.f;
  P[S(2, 1)] = 0.f;
  P[S(2, 2)] = 0.f;
  P[S(3, 2)] = 0.f;
  P[S(4, 2)] = 0.f;
  P[S(5, 2)] = 0.f;
}

template <bool kZRow0, bool kAccurateTrig = false>
__device__ inline void propagate_fast(Track& t, float hz) {
  const Jac J = helix_fast<kAccurateTrig, !kZRow0>(t, hz);
  propagate_cov_fast<kZRow0>(t.P, J);
}

// Rows of the state that carry information.  After a propagation row 2 is
// zero and the update skips it (kZRow0); a standalone update of an arbitrary
// covariance uses all six rows.
template <bool kZRow0>
struct Live {
  static constexpr int n = kZRow0 ? 5 : 6;
  __device__ static constexpr int at(int a) { return kZRow0 ? (a < 2 ? a : a + 1) : a; }
};

// kCarry: g holds sin/cos of phi and theta, rotated here by the update's
// corrections of the two angles.
// 1 / (r00 r11 - r01^2) with the determinant and the division in double,
// rounded to float.  Out of line: the update takes it only for operands whose
// FP32 determinant overflows, and inline its double division's registers
// would add to the register peak of every layer loop.
__device__ __noinline__ float inv_det_double(float r00, float r01, float r11) {
#pragma clang fp contract(off)
  return (float)(1.0 / ((double)r00 * (double)r11 - (double)r01 * (double)r01));
}

template <bool kZRow0, bool kCarry = false>
__device__ inline void update_fast(Track& t, const Hit& h, Trig* g = nullptr) {
#pragma clang fp contract(fast)
  using L = Live<kZRow0>;
  float* p = t.p;
  float* P = t.P;
  const float r00 = P[0] + h.xx;
  const float r01 = P[1] + h.xy;
  const float r11 = P[2] + h.yy;
  // r00 r11 - r01^2 with Kahan's FMA algorithm: w = r01^2 rounded, e its
  // exact rounding error, then r00 r11 - w in one rounding plus e.  Within
  // 1.5 ulp of the exact determinant, as the reference's FP64 evaluation
  // (OMP:519-520) rounded to float is within 0.5 ulp, so cancellation in it
  // does not dominate; four FP32 operations instead of six with conversions.
  const float w = r01 * r01;
  const float e = __builtin_fmaf(-r01, r01, w);
  const float det = __builtin_fmaf(r00, r11, -w) + e;
  float s;
  if (__builtin_expect(!(__builtin_fabsf(det) < 0x1p126f), 0)) {
    // The FP32 products overflowed (an entry beyond ~1.8e19: a track whose
    // theta is within rounding of pi/2 turns by 1e4 rad and its covariance
    // follows), or 1/det would be subnormal: the determinant and its
    // reciprocal in double, as the reference evaluates them (OMP:519-521).
    s = inv_det_double(r00, r01, r11);
  } else {
    s = fast_rcp(det);
  }
  const float i00 = s * r11, i01 = r01 * -s, i11 = s * r00;
  float K0[6], K1[6];
#pragma unroll
  for (int n = 0; n < L::n; ++n) {
    const int i = L::at(n);
    const float a = P[S(i, 0)], b = P[S(i, 1)];
    K0[i] = a * i00 + b * i01;
    K1[i] = a * i01 + b * i11;
  }
  const float res0 = h.x - p[0];
  const float res1 = h.y - p[1];
#pragma unroll
  for (int n = 0; n < L::n; ++n) {
    const int i = L::at(n);
    if (kCarry && i >= 4) continue;
    p[i] = p[i] + K0[i] * res0 + K1[i] * res1;
  }
  if constexpr (kCarry) {
    const float dphi = K0[4] * res0 + K1[4] * res1;
    const float dtheta = K0[5] * res0 + K1[5] * res1;
    p[4] = p[4] + dphi;
    p[5] = p[5] + dtheta;
    float sd, omcd;
    sincos_increment(dphi, &sd, &omcd);
    rotate(g->sinP, g->cosP, sd, omcd);
    sincos_increment(dtheta, &sd, &omcd);
    rotate(g->sinT, g->cosT, sd, omcd);
  }
  // P - K (H P); H P = rows 0 and 1 of P.
  float HP0[6], HP1[6];
#pragma unroll
  for (int n = 0; n < L::n; ++n) {
    const int j = L::at(n);
    HP0[j] = P[S(0, j)];
    HP1[j] = P[S(1, j)];
  }
#pragma unroll
  for (int a = 0; a < L::n; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      const int i = L::at(a), j = L::at(b);
      P[S(i, j)] = P[S(i, j)] - K0[i] * HP0[j] - K1[i] * HP1[j];
    }
}

// ---------------------------------------------------------------------------
// kFp64Gain: FP64 inverse, gain, parameter update and Joseph-form covariance
// ---------------------------------------------------------------------------

// kCarry: g holds sin/cos of phi and theta, rotated here by the update's
// corrections of the two angles.
template <bool kZRow0, bool kCarry = false>
__device__ inline void update_fp64gain(Track& t, const Hit& h, Trig* g = nullptr) {
#pragma clang fp contract(fast)
  using L = Live<kZRow0>;
  constexpr int n = L::n;
  float* p = t.p;
  float* P = t.P;
  // entry (a, b) of the live block in double (a and b over live indices)
  auto q = [&](int a, int b) -> double { return (double)P[S(L::at(a), L::at(b))]; };
  // C = P H^T: the columns of x and y
  double C0[6], C1[6];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    C0[a] = q(a, 0);
    C1[a] = q(a, 1);
  }
  const double R00 = h.xx, R01 = h.xy, R11 = h.yy;
  // S = H P H^T + R and its inverse: 1/det from the hardware FP64 reciprocal
  // refined by two Newton steps (each doubles the correct bits; the
  // estimate has more than 20), i.e. correct to about one ulp of double
  const double r00 = C0[0] + R00, r01 = C0[1] + R01, r11 = C1[1] + R11;
  const double det = r00 * r11 - r01 * r01;
  double s = __builtin_amdgcn_rcp(det);
  s = __builtin_fma(s, __builtin_fma(-det, s, 1.0), s);
  s = __builtin_fma(s, __builtin_fma(-det, s, 1.0), s);
  const double i00 = s * r11, i01 = -s * r01, i11 = s * r00;
  double K0[6], K1[6];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    K0[a] = C0[a] * i00 + C1[a] * i01;
    K1[a] = C0[a] * i01 + C1[a] * i11;
  }
  const double res0 = (double)h.x - (double)p[0];
  const double res1 = (double)h.y - (double)p[1];
  // p + K res as one FMA chain from p; the two angles keep their increment
  // apart for the rotation of their sin/cos
  float dang[2] = {0.f, 0.f};
#pragma unroll
  for (int a = 0; a < n; ++a) {
    const int i = L::at(a);
    if (i >= 4) {
      const double d = __builtin_fma(K0[a], res0, K1[a] * res1);
      p[i] = (float)((double)p[i] + d);
      dang[i - 4] = (float)d;
    } else {
      p[i] = (float)__builtin_fma(K0[a], res0, __builtin_fma(K1[a], res1, (double)p[i]));
    }
  }
  if constexpr (kCarry) {
    float sd, omcd;
    sincos_increment(dang[0], &sd, &omcd);
    rotate(g->sinP, g->cosP, sd, omcd);
    sincos_increment(dang[1], &sd, &omcd);
    rotate(g->sinT, g->cosT, sd, omcd);
  }
  // Joseph form (I - K H) P (I - K H)^T + K R K^T, expanded:
  //   P - K C^T - C K^T + K S K^T,   C = P H^T,  S = H P H^T + R,
  // grouped symmetrically as P + M K^T + K M^T with M = K S / 2 - C:
  // entry (a, b) = P_ab + M_a.K_b + K_a.M_b, four FMAs.  K S - C vanishes
  // for the exact optimal gain; keeping it (in M) keeps the form's
  // insensitivity to errors in K.  Only K and M stay live over the 15
  // entries (the entries of P are widened one at a time).
  const double h00 = 0.5 * r00, h01 = 0.5 * r01, h11 = 0.5 * r11;  // S / 2, exact
  double M0[6], M1[6];
#pragma unroll
  for (int a = 0; a < n; ++a) {
    M0[a] = __builtin_fma(K0[a], h00, __builtin_fma(K1[a], h01, -C0[a]));
    M1[a] = __builtin_fma(K0[a], h01, __builtin_fma(K1[a], h11, -C1[a]));
  }
#pragma unroll
  for (int a = 0; a < n; ++a)
#pragma unroll
    for (int b = 0; b <= a; ++b) {
      const double v = q(a, b) + M0[a] * K0[b] + M1[a] * K1[b] + K0[a] * M0[b] + K1[a] * M1[b];
      P[S(L::at(a), L::at(b))] = (float)v;
    }
}

// ---------------------------------------------------------------------------
// KalmanUpdate v1 (OMP:375-501): 3-D hit, 3x3 inverse.  Off the reference's
// path (commented out at OMP:743); the SYCL variant calls it.
// ---------------------------------------------------------------------------

// Word of (i, j) in the upper triangle stored row by row, the addressing v1
// uses for both covariances ((0,0) (0,1) .. (0,5) (1,1) ..).
__device__ constexpr int up_idx(int i, int j) {
  return i <= j ? i * 6 - i * (i - 1) / 2 + (j - i) : j * 6 - j * (j - 1) / 2 + (i - j);
}

// v1 as written, including its addressing of the 21 covariance words and the
// 6 hit covariance words as upper triangles (the path stores lower
// triangles), its (0,2) inverse entry m01*m12 - m02*m12 (OMP:395) and its
// write of pt + K*res into the ipt slot (OMP:447, 494).  Typing as written:
// float sums and determinant expression, the determinant widened to double,
// 1/det and its products in double, pt = 1./ipt in double; no contraction.
// hc: the 6 hit covariance words.
__device__ inline void update_v1_ref(Track& t, const float hc[6], float xout, float yout) {
#pragma clang fp contract(off)
  float* a = t.P;
  float* par = t.p;
  const float m00 = a[0] + hc[0], m01 = a[1] + hc[1], m02 = a[2] + hc[2];
  const float m11 = a[6] + hc[3], m12 = a[7] + hc[4], m22 = a[11] + hc[5];
  const double det = (double)((m00 * ((m11 * m22) - (m12 * m12))) - (m01 * ((m01 * m22) - (m12 * m02))) +
                              (m02 * ((m01 * m12) - (m02 * m11))));
  const double invdet = 1.0 / det;
  float c[9];
  c[0] = (float)(invdet * (double)((m11 * m22) - (m12 * m12)));
  c[1] = (float)(-invdet * (double)((m01 * m22) - (m02 * m12)));
  c[2] = (float)(invdet * (double)((m01 * m12) - (m02 * m12)));
  c[3] = (float)(-invdet * (double)((m01 * m22) - (m12 * m02)));
  c[4] = (float)(invdet * (double)((m00 * m22) - (m02 * m02)));
  c[5] = (float)(-invdet * (double)((m00 * m12) - (m02 * m01)));
  c[6] = (float)(invdet * (double)((m01 * m12) - (m02 * m11)));
  c[7] = (float)(-invdet * (double)((m00 * m12) - (m02 * m01)));
  c[8] = (float)(invdet * (double)((m00 * m11) - (m01 * m01)));
  float K[18];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      K[3 * i + j] = a[up_idx(i, 0)] * c[j] + a[up_idx(i, 1)] * c[3 + j] + a[up_idx(i, 2)] * c[6 + j];
  const float rx = xout - par[0], ry = yout - par[1];
  const float ptin = (float)(1.0 / (double)par[3]);
  float pn[6];
  pn[0] = par[0] + (K[0] * rx) + (K[1] * ry);
  pn[1] = par[1] + (K[3] * rx) + (K[4] * ry);
  pn[2] = par[2] + (K[6] * rx) + (K[7] * ry);
  pn[3] = ptin + (K[9] * rx) + (K[10] * ry);
  pn[4] = par[4] + (K[12] * rx) + (K[13] * ry);
  pn[5] = par[5] + (K[15] * rx) + (K[16] * ry);
  float n[21];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      const int e = up_idx(i, j);
      n[e] = a[e] - (K[3 * i] * a[up_idx(0, j)] + K[3 * i + 1] * a[up_idx(1, j)] + K[3 * i + 2] * a[up_idx(2, j)]);
    }
#pragma unroll
  for (int e = 0; e < 21; ++e) a[e] = n[e];
#pragma unroll
  for (int e = 0; e < 6; ++e) par[e] = pn[e];
}

// The update v1 means, in the path's packing: lower triangles for both
// covariances, residual (x, y, z) - (p0, p1, p2), S = P(0..2, 0..2) + R
// inverted by cofactors in double, K = P H^T S^-1 in float, p' = p + K res
// (ipt updated as ipt), P' = P - K H P.  No contraction.
__device__ inline void update_v1_fixed(Track& t, const float hc[6], float hx, float hy, float hz) {
#pragma clang fp contract(off)
  float* P = t.P;
  float* par = t.p;
  const double m00 = P[S(0, 0)] + hc[0], m01 = P[S(1, 0)] + hc[1], m11 = P[S(1, 1)] + hc[2];
  const double m02 = P[S(2, 0)] + hc[3], m12 = P[S(2, 1)] + hc[4], m22 = P[S(2, 2)] + hc[5];
  const double c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22, c02 = m01 * m12 - m02 * m11;
  const double c11 = m00 * m22 - m02 * m02, c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
  const double det = m00 * c00 + m01 * c01 + m02 * c02;
  const double id = 1.0 / det;
  const float inv[3][3] = {{(float)(c00 * id), (float)(c01 * id), (float)(c02 * id)},
                           {(float)(c01 * id), (float)(c11 * id), (float)(c12 * id)},
                           {(float)(c02 * id), (float)(c12 * id), (float)(c22 * id)}};
  float K[6][3];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) K[i][j] = P[S(i, 0)] * inv[0][j] + P[S(i, 1)] * inv[1][j] + P[S(i, 2)] * inv[2][j];
  const float r0 = hx - par[0], r1 = hy - par[1], r2 = hz - par[2];
  float n[21];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j)
      n[S(i, j)] = P[S(i, j)] - (K[i][0] * P[S(0, j)] + K[i][1] * P[S(1, j)] + K[i][2] * P[S(2, j)]);
#pragma unroll
  for (int i = 0; i < 6; ++i) par[i] = par[i] + K[i][0] * r0 + K[i][1] * r1 + K[i][2] * r2;
#pragma unroll
  for (int e = 0; e < 21; ++e) P[e] = n[e];
}

// ---------------------------------------------------------------------------
// one layer
// ---------------------------------------------------------------------------

// kFirst: the covariance may still have a non-zero z row (first layer of a
// track); otherwise the fast forms drop the structurally-zero terms.
// g: sin/cos of phi and theta carried between the layers of one track
// (FAST, FP64GAIN); set up on the first layer.
template <int kMath, bool kFirst>
__device__ inline void layer_step(Track& t, const Hit& h, Trig& g) {
  if constexpr (kMath == kRef) {
    propagate_ref(t, h.z);
    update_ref(t, h);
  } else if constexpr (kMath == kFast) {
    if constexpr (kFirst) trig_init<false>(t, g);
    const Jac J = helix_core<kFirst>(t, h.z, g);
    propagate_cov_fast<!kFirst>(t.P, J);
    update_fast<true, true>(t, h, &g);
  } else {
    if constexpr (kFirst) trig_init<true>(t, g);
    const Jac J = helix_core<kFirst>(t, h.z, g);
    propagate_cov_fast<!kFirst>(t.P, J);
    update_fp64gain<true, true>(t, h, &g);
  }
}

}  // namespace p2z
