// Covariance propagation of the fast arithmetic modes when the input z row
// and column are zero (every layer of a track after its first), and the
// packed-index helper it shares with the lane math.  Plain float code with
// __host__ __device__ annotations and no HIP builtins, so a host program can
// include it (after defining the two annotations) and check it against a
// dense product.
#pragma once

namespace p2z {

// Packed lower-triangular index of (i, j) (reference SymOffsets66, OMP:52-55).
__host__ __device__ constexpr int S(int i, int j) {
  return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
}

// The 11 entries of the propagation Jacobian Jt that differ from the identity
// (OMP:655-666); row 2 (z) of Jt is zero.
struct Jac {
  float J02, J03, J04, J05, J12, J13, J14, J15, J42, J43, J45;
};

// P' = Jt P Jt^T when the input z row/column is zero (every layer after the
// first).  The live indices split into u = (x, y) and v = (ipt, phi, theta);
// Jt restricted to them is [[I, A], [0, B]] with A = [J0v; J1v] (2x3) and B
// the identity but for the phi row (J43, 1, J45).  With Q = Puv + A Pvv:
//   P'uv = Q B^T  (only the phi column differs from Q),
//   P'uu = Puu + A Pvu + Q A^T,
//   P'vv = B Pvv B^T  (only the phi row/column changes),
// 48 FMAs, 6 fewer per layer than the row-by-row product (as compiled, with
// its common terms shared); the z row and column stay exactly zero.
__host__ __device__ inline void propagate_cov_fast_z0(float* P, const Jac& J) {
#pragma clang fp contract(fast)
  const float p00 = P[S(0, 0)], p10 = P[S(1, 0)], p11 = P[S(1, 1)];
  const float p30 = P[S(3, 0)], p31 = P[S(3, 1)], p33 = P[S(3, 3)];
  const float p40 = P[S(4, 0)], p41 = P[S(4, 1)], p43 = P[S(4, 3)], p44 = P[S(4, 4)];
  const float p50 = P[S(5, 0)], p51 = P[S(5, 1)], p53 = P[S(5, 3)], p54 = P[S(5, 4)], p55 = P[S(5, 5)];
  // Q = Puv + A Pvv
  const float q03 = p30 + J.J03 * p33 + J.J04 * p43 + J.J05 * p53;
  const float q04 = p40 + J.J03 * p43 + J.J04 * p44 + J.J05 * p54;
  const float q05 = p50 + J.J03 * p53 + J.J04 * p54 + J.J05 * p55;
  const float q13 = p31 + J.J13 * p33 + J.J14 * p43 + J.J15 * p53;
  const float q14 = p41 + J.J13 * p43 + J.J14 * p44 + J.J15 * p54;
  const float q15 = p51 + J.J13 * p53 + J.J14 * p54 + J.J15 * p55;
  // P'uu = (Puu + A Pvu) + Q A^T; entry (1, 0) as (0, 1) = (Puu + A Pvu)_01 + Q_0 . A_1
  P[S(0, 0)] = p00 + J.J03 * p30 + J.J04 * p40 + J.J05 * p50 + q03 * J.J03 + q04 * J.J04 + q05 * J.J05;
  P[S(1, 0)] = p10 + J.J03 * p31 + J.J04 * p41 + J.J05 * p51 + q03 * J.J13 + q04 * J.J14 + q05 * J.J15;
  P[S(1, 1)] = p11 + J.J13 * p31 + J.J14 * p41 + J.J15 * p51 + q13 * J.J13 + q14 * J.J14 + q15 * J.J15;
  // P'uv = Q B^T
  P[S(3, 0)] = q03;
  P[S(3, 1)] = q13;
  P[S(4, 0)] = J.J43 * q03 + q04 + J.J45 * q05;
  P[S(4, 1)] = J.J43 * q13 + q14 + J.J45 * q15;
  P[S(5, 0)] = q05;
  P[S(5, 1)] = q15;
  // P'vv = B Pvv B^T
  const float t3 = J.J43 * p33 + p43 + J.J45 * p53;  // phi row of B Pvv
  const float t4 = J.J43 * p43 + p44 + J.J45 * p54;
  const float t5 = J.J43 * p53 + p54 + J.J45 * p55;
  P[S(4, 3)] = t3;
  P[S(4, 4)] = J.J43 * t3 + t4 + J.J45 * t5;
  P[S(5, 4)] = t5;
  // (3,3), (5,3), (5,5) unchanged; the z row and column stay zero
  P[S(2, 0)] = 0.f;
  P[S(2, 1)] = 0.f;
  P[S(2, 2)] = 0.f;
  P[S(3, 2)] = 0.f;
  P[S(4, 2)] = 0.f;
  P[S(5, 2)] = 0.f;
}

}  // namespace p2z
