/*
 * p2z oracle -- TEST INFRASTRUCTURE ONLY (see p2z_oracle.h).
 *
 * Restates the reference CPU path of cerati/p2z-tests
 * (src/propagate-toz-test_OMP.cpp, "OMP" below) in plain C, keeping the
 * reference's floating-point typing: every quantity is float except where the
 * reference's literals `1.` / `1.0` or its `(double)` casts promote a
 * sub-expression to double.  Build with -ffp-contract=off so that every product
 * and sum is rounded as written; the GPU "reference typing" kernel is checked
 * against this build.
 *
 * Pinning (what checks each part):
 *   generator (randn, prepare_*)      : bitwise against the reference compiled
 *                                       from its own source (oracle/_ref), and the
 *                                       printed known answer of the reference logs
 *   propagate / update (float typing) : bitwise against oracle/_ref per track,
 *                                       closed-form helix and dense double
 *                                       J*P*J^T / textbook Kalman update in numpy,
 *                                       and the known-answer summary
 *   summary                           : the 11 printed lines of
 *                                       logs/propagate_gcc_omp.log:30-40
 *   fp64 variant                      : against the float oracle (relative 1e-5)
 *                                       and the textbook forms in numpy
 *   ragged driver                     : equals the plain driver when every lane
 *                                       has the same count; per-lane truncation
 *                                       equals a plain run with fewer layers
 */
#include "p2z_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* sin / cos of the path.  Default: the C library's sinf / cosf, as the
 * reference calls them.  With -DORACLE_CR_LIBM: correctly rounded values
 * (evaluated in double, then rounded), i.e. the same algorithm under a
 * different libm; the pair measures the path's sensitivity to the libm. */
#ifdef ORACLE_CR_LIBM
static inline float path_sinf(float x) { return (float)sin((double)x); }
static inline float path_cosf(float x) { return (float)cos((double)x); }
#else
#define path_sinf sinf
#define path_cosf cosf
#endif

/* kfact as the reference declares it: a float initialised from a double
 * expression (OMP:622). */
static const float kKfact = 100. / (-0.299792458 * 3.8112);
static const double kKfact64 = 100. / (-0.299792458 * 3.8112);

/* Packed lower-triangular index of (i,j) in a symmetric 6x6 (OMP:52-55). */
static inline int sym6(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

/* ========================================================================= */
/* generator                                                                 */
/* ========================================================================= */

/* Marsaglia polar method over libc rand() with one cached variate
 * (OMP:127-145).  The reference is C++: pow(float,int) promotes to double, but
 * log(float) and sqrt(float) resolve to the float overloads. */
static int g_randn_have = 0;
static float g_randn_x2 = 0.f;

void oracle_srand(unsigned seed) {
  srand(seed);
  g_randn_have = 0;
}

float oracle_randn(float mu, float sigma) {
  if (g_randn_have) {
    g_randn_have = 0;
    return mu + sigma * g_randn_x2;
  }
  float u1, u2, w;
  do {
    u1 = -1 + ((float)rand() / RAND_MAX) * 2;
    u2 = -1 + ((float)rand() / RAND_MAX) * 2;
    w = (float)(pow(u1, 2) + pow(u2, 2));
  } while (w >= 1 || w == 0);
  const float mult = sqrtf((-2 * logf(w)) / w);
  const float x1 = u1 * mult;
  g_randn_x2 = u2 * mult;
  g_randn_have = 1;
  return mu + sigma * x1;
}

void oracle_prepare_tracks(const oracle_atrk* in, float* out, int64_t nevts, int64_t nb,
                           int W, double smear) {
  /* OMP:243-263: loops ie, ib, it; 6 par draws then 21 cov draws per lane. */
  for (int64_t ie = 0; ie < nevts; ++ie)
    for (int64_t ib = 0; ib < nb; ++ib) {
      float* blk = out + (ib + nb * ie) * (int64_t)ORACLE_TRK_WORDS * W;
      for (int it = 0; it < W; ++it) {
        for (int ip = 0; ip < 6; ++ip)
          blk[ip * W + it] = (float)((1 + smear * oracle_randn(0, 1)) * in->par[ip]);
        for (int ip = 0; ip < 21; ++ip)
          blk[(6 + ip) * W + it] = (float)((1 + smear * oracle_randn(0, 1)) * in->cov[ip] * 100);
        ((int32_t*)blk)[27 * W + it] = in->q;
      }
    }
}

int oracle_prepare_hits(const oracle_ahit* hits, int nhits, float* out, int64_t nevts,
                        int64_t nb, int W, int nlayer, double smear) {
  /* OMP:265-293: layer loop outermost. */
  for (int lay = 0; lay < nlayer; ++lay) {
    if (lay >= nhits) return -1;
    const oracle_ahit* h = &hits[lay];
    for (int64_t ie = 0; ie < nevts; ++ie)
      for (int64_t ib = 0; ib < nb; ++ib) {
        float* blk = out + (lay + nlayer * (ib + nb * ie)) * (int64_t)ORACLE_HIT_WORDS * W;
        for (int it = 0; it < W; ++it) {
          for (int ip = 0; ip < 3; ++ip)
            blk[ip * W + it] = (float)((1 + smear * oracle_randn(0, 1)) * h->pos[ip]);
          for (int ip = 0; ip < 6; ++ip)
            blk[(3 + ip) * W + it] = (float)((1 + smear * oracle_randn(0, 1)) * h->cov[ip]);
        }
      }
  }
  return 0;
}

/* ========================================================================= */
/* per-lane path, reference typing                                            */
/* ========================================================================= */

/* Helix propagation of one lane to z = hz (OMP:627-678) followed by
 * P' = Jt P Jt^T with Jt = J except row 2 zeroed (OMP:296-373).
 * par: x y z ipt phi theta.  P: packed lower triangle (21). */
static void lane_propagate(float P[21], float par[6], int q, float hz) {
  const float zout = hz;
  const float k = q * kKfact;
  const float deltaZ = zout - par[2];
  const float ipt = par[3];
  const float pt = (float)(1. / ipt);                 /* double divide   */
  const float cosP = path_cosf(par[4]);
  const float sinP = path_sinf(par[4]);
  const float cosT = path_cosf(par[5]);
  const float sinT = path_sinf(par[5]);
  const float pxin = cosP * pt;
  const float pyin = sinP * pt;
  const float icosT = (float)(1.0 / cosT);            /* double divide   */
  const float icosTk = icosT / k;                     /* float divide    */
  const float alpha = deltaZ * sinT * ipt * icosTk;
  const float sina = path_sinf(alpha);
  const float cosa = path_cosf(alpha);

  /* x, y: the (1.-cosa) factor promotes the rest of the expression to double */
  par[0] = (float)(par[0] + k * ((double)(pxin * sina) - pyin * (1. - cosa)));
  par[1] = (float)(par[1] + k * ((double)(pyin * sina) + pxin * (1. - cosa)));
  par[2] = zout;
  /* par[3] (ipt) and par[5] (theta) unchanged */
  par[4] = par[4] + alpha;

  const float sCosPsina = path_sinf(cosP * sina);
  const float cCosPsina = path_cosf(cosP * sina);

  /* the eleven off-diagonal Jacobian entries (OMP:656-666) */
  const float J02 = cosP * sinT * (sinP * cosa * sCosPsina - cosa) * icosT;
  const float J03 = (float)((double)(cosP * sinT * deltaZ * cosa) * (1. - sinP * sCosPsina) * (icosT * pt)
                           - k * ((double)(cosP * sina) - sinP * (1. - cCosPsina)) * (pt * pt));
  const float J04 = (float)((k * pt) * ((double)(-sinP * sina + sinP * sinP * sina * sCosPsina)
                                        - cosP * (1. - cCosPsina)));
  const float J05 = (float)((double)(cosP * deltaZ * cosa) * (1. - sinP * sCosPsina) * (icosT * icosT));
  const float J12 = cosa * sinT * (cosP * cosP * sCosPsina - sinP) * icosT;
  const float J13 = (float)((double)(sinT * deltaZ * cosa * (cosP * cosP * sCosPsina + sinP) * (icosT * pt))
                           - k * ((double)(sinP * sina) + cosP * (1. - cCosPsina)) * (pt * pt));
  const float J14 = (float)((k * pt) * (-sinP * (1. - cCosPsina) - (double)(sinP * cosP * sina * sCosPsina)
                                        + (double)(cosP * sina)));
  const float J15 = deltaZ * cosa * (cosP * cosP * sCosPsina + sinP) * (icosT * icosT);
  const float J42 = -ipt * sinT * icosTk;
  const float J43 = sinT * deltaZ * icosTk;
  const float J45 = ipt * deltaZ * (icosT * icosTk);

  /* T = Jt * P (6x6 dense), summation order of OMP:304-339 */
  float T[6][6];
  for (int c = 0; c < 6; ++c) {
    const float p0 = P[sym6(0, c)], p1 = P[sym6(1, c)], p2 = P[sym6(2, c)];
    const float p3 = P[sym6(3, c)], p4 = P[sym6(4, c)], p5 = P[sym6(5, c)];
    T[0][c] = p0 + J02 * p2 + J03 * p3 + J04 * p4 + J05 * p5;
    T[1][c] = p1 + J12 * p2 + J13 * p3 + J14 * p4 + J15 * p5;
    T[2][c] = 0.f;
    T[3][c] = p3;
    T[4][c] = J42 * p2 + J43 * p3 + p4 + J45 * p5;
    T[5][c] = p5;
  }
  /* P' = T * Jt^T, packed lower triangle, summation order of OMP:351-371 */
  for (int i = 0; i < 6; ++i) {
    P[sym6(i, 0)] = T[i][0] + T[i][2] * J02 + T[i][3] * J03 + T[i][4] * J04 + T[i][5] * J05;
    if (i >= 1) P[sym6(i, 1)] = T[i][1] + T[i][2] * J12 + T[i][3] * J13 + T[i][4] * J14 + T[i][5] * J15;
    if (i >= 2) P[sym6(i, 2)] = 0.f;
    if (i >= 3) P[sym6(i, 3)] = T[i][3];
    if (i >= 4) P[sym6(i, 4)] = T[i][2] * J42 + T[i][3] * J43 + T[i][4] + T[i][5] * J45;
    if (i >= 5) P[sym6(i, 5)] = T[i][5];
  }
}

/* 2-D (x,y) Kalman update of one lane (OMP:503-620).
 * hc: hit covariance xx, xy, yy.  Standard form P' = P - K H P. */
static void lane_update(float P[21], float par[6], const float hc[3], float hx, float hy) {
  float r00 = P[0] + hc[0];
  float r01 = P[1] + hc[1];
  float r11 = P[2] + hc[2];
  const double det = (double)r00 * r11 - (double)r01 * r01;
  const float s = (float)(1.f / det);                 /* double divide   */
  const float tmp = s * r11;
  r01 *= -s;
  r11 = s * r00;
  r00 = tmp;

  /* gain K (6x2): rows use P entries (i,0),(i,1) */
  float K[6][2];
  for (int i = 0; i < 6; ++i) {
    const float a = P[sym6(i, 0)], b = P[sym6(i, 1)];
    K[i][0] = a * r00 + b * r01;
    K[i][1] = a * r01 + b * r11;
  }
  const float res0 = hx - par[0];
  const float res1 = hy - par[1];
  for (int i = 0; i < 6; ++i) par[i] = par[i] + K[i][0] * res0 + K[i][1] * res1;

  /* C = K (H P) over the packed lower triangle, then P - C */
  float C[21];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) C[sym6(i, j)] = K[i][0] * P[sym6(0, j)] + K[i][1] * P[sym6(1, j)];
  for (int e = 0; e < 21; ++e) P[e] = P[e] - C[e];
}

/* ========================================================================= */
/* per-lane path, all double                                                 */
/* ========================================================================= */

static void lane_propagate64(double P[21], double par[6], int q, double hz) {
  const double k = q * kKfact64;
  const double deltaZ = hz - par[2];
  const double ipt = par[3];
  const double pt = 1. / ipt;
  const double cosP = cos(par[4]), sinP = sin(par[4]);
  const double cosT = cos(par[5]), sinT = sin(par[5]);
  const double pxin = cosP * pt, pyin = sinP * pt;
  const double icosT = 1. / cosT;
  const double icosTk = icosT / k;
  const double alpha = deltaZ * sinT * ipt * icosTk;
  const double sina = sin(alpha), cosa = cos(alpha);
  par[0] = par[0] + k * (pxin * sina - pyin * (1. - cosa));
  par[1] = par[1] + k * (pyin * sina + pxin * (1. - cosa));
  par[2] = hz;
  par[4] = par[4] + alpha;
  const double sC = sin(cosP * sina), cC = cos(cosP * sina);
  double J[6][6];
  memset(J, 0, sizeof J);
  J[0][0] = J[1][1] = J[3][3] = J[4][4] = J[5][5] = 1.;  /* row 2 stays zero */
  J[0][2] = cosP * sinT * (sinP * cosa * sC - cosa) * icosT;
  J[0][3] = cosP * sinT * deltaZ * cosa * (1. - sinP * sC) * (icosT * pt) - k * (cosP * sina - sinP * (1. - cC)) * (pt * pt);
  J[0][4] = (k * pt) * (-sinP * sina + sinP * sinP * sina * sC - cosP * (1. - cC));
  J[0][5] = cosP * deltaZ * cosa * (1. - sinP * sC) * (icosT * icosT);
  J[1][2] = cosa * sinT * (cosP * cosP * sC - sinP) * icosT;
  J[1][3] = sinT * deltaZ * cosa * (cosP * cosP * sC + sinP) * (icosT * pt) - k * (sinP * sina + cosP * (1. - cC)) * (pt * pt);
  J[1][4] = (k * pt) * (-sinP * (1. - cC) - sinP * cosP * sina * sC + cosP * sina);
  J[1][5] = deltaZ * cosa * (cosP * cosP * sC + sinP) * (icosT * icosT);
  J[4][2] = -ipt * sinT * icosTk;
  J[4][3] = sinT * deltaZ * icosTk;
  J[4][5] = ipt * deltaZ * (icosT * icosTk);
  double T[6][6];
  for (int i = 0; i < 6; ++i)
    for (int c = 0; c < 6; ++c) {
      double acc = 0.;
      for (int m = 0; m < 6; ++m) acc += J[i][m] * P[sym6(m, c)];
      T[i][c] = acc;
    }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      double acc = 0.;
      for (int m = 0; m < 6; ++m) acc += T[i][m] * J[j][m];
      P[sym6(i, j)] = acc;
    }
}

static void lane_update64(double P[21], double par[6], const double hc[3], double hx, double hy) {
  const double r00 = P[0] + hc[0], r01 = P[1] + hc[1], r11 = P[2] + hc[2];
  const double s = 1. / (r00 * r11 - r01 * r01);
  const double i00 = s * r11, i01 = -s * r01, i11 = s * r00;
  double K[6][2];
  for (int i = 0; i < 6; ++i) {
    const double a = P[sym6(i, 0)], b = P[sym6(i, 1)];
    K[i][0] = a * i00 + b * i01;
    K[i][1] = a * i01 + b * i11;
  }
  const double res0 = hx - par[0], res1 = hy - par[1];
  for (int i = 0; i < 6; ++i) par[i] += K[i][0] * res0 + K[i][1] * res1;
  double C[21];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) C[sym6(i, j)] = K[i][0] * P[sym6(0, j)] + K[i][1] * P[sym6(1, j)];
  for (int e = 0; e < 21; ++e) P[e] -= C[e];
}

/* ========================================================================= */
/* block functions                                                           */
/* ========================================================================= */

void oracle_propagate_to_z(const float* inErr, const float* inPar, const int32_t* inChg,
                           const float* msP, float* outErr, float* outPar, int W) {
  for (int it = 0; it < W; ++it) {
    float P[21], par[6];
    for (int e = 0; e < 21; ++e) P[e] = inErr[e * W + it];
    for (int e = 0; e < 6; ++e) par[e] = inPar[e * W + it];
    lane_propagate(P, par, inChg[it], msP[2 * W + it]);
    for (int e = 0; e < 21; ++e) outErr[e * W + it] = P[e];
    for (int e = 0; e < 6; ++e) outPar[e * W + it] = par[e];
  }
}

void oracle_kalman_update_v2(float* trkErr, float* inPar, const float* hitErr, const float* msP,
                             int W) {
  for (int it = 0; it < W; ++it) {
    float P[21], par[6];
    for (int e = 0; e < 21; ++e) P[e] = trkErr[e * W + it];
    for (int e = 0; e < 6; ++e) par[e] = inPar[e * W + it];
    const float hc[3] = {hitErr[0 * W + it], hitErr[1 * W + it], hitErr[2 * W + it]};
    lane_update(P, par, hc, msP[0 * W + it], msP[1 * W + it]);
    for (int e = 0; e < 21; ++e) trkErr[e * W + it] = P[e];
    for (int e = 0; e < 6; ++e) inPar[e * W + it] = par[e];
  }
}

void oracle_propagate_to_z_fp64(const float* inErr, const float* inPar, const int32_t* inChg,
                                const float* msP, float* outErr, float* outPar, int W) {
  for (int it = 0; it < W; ++it) {
    double P[21], par[6];
    for (int e = 0; e < 21; ++e) P[e] = inErr[e * W + it];
    for (int e = 0; e < 6; ++e) par[e] = inPar[e * W + it];
    lane_propagate64(P, par, inChg[it], msP[2 * W + it]);
    for (int e = 0; e < 21; ++e) outErr[e * W + it] = (float)P[e];
    for (int e = 0; e < 6; ++e) outPar[e * W + it] = (float)par[e];
  }
}

void oracle_kalman_update_v2_fp64(float* trkErr, float* inPar, const float* hitErr, const float* msP,
                                  int W) {
  for (int it = 0; it < W; ++it) {
    double P[21], par[6];
    for (int e = 0; e < 21; ++e) P[e] = trkErr[e * W + it];
    for (int e = 0; e < 6; ++e) par[e] = inPar[e * W + it];
    const double hc[3] = {hitErr[0 * W + it], hitErr[1 * W + it], hitErr[2 * W + it]};
    lane_update64(P, par, hc, msP[0 * W + it], msP[1 * W + it]);
    for (int e = 0; e < 21; ++e) trkErr[e * W + it] = (float)P[e];
    for (int e = 0; e < 6; ++e) inPar[e * W + it] = (float)par[e];
  }
}

/* ---- KalmanUpdate v1 (OMP:375-501), off the reference's path ---------------
 * A 3-D hit update with a 3x3 inverse, commented out at its call site
 * (OMP:743); the SYCL variant calls it (src/propagate-toz-test_sycl.cpp:718).
 *
 * fixed == 0 restates it as written, including three properties of that code
 * that a caller has to know about:
 *  - it addresses the 21 covariance words as the UPPER triangle row by row
 *    ((0,0) (0,1) .. (0,5) (1,1) ..), and the 6 hit covariance words likewise
 *    (xx xy xz yy yz zz), while propagateToZ and KalmanUpdate_v2 store the
 *    LOWER triangle row by row (OMP:52-55);
 *  - entry (0,2) of the inverse multiplies m01*m12 - m02*m12 (OMP:395), where
 *    the cofactor is m01*m12 - m02*m11 (entry (2,0), OMP:399, has it right);
 *  - it writes pt + K*res (pt = 1/ipt) into the ipt slot (OMP:447, 494).
 * Typing as written: the sums and the determinant expression in float, the
 * determinant widened to double, 1.0/det and the products with it in double,
 * pt = 1./ipt in double; no contraction.
 *
 * fixed != 0 is the update the code means, in the path's own packing: lower
 * triangles for the track and the hit covariance, residual (x, y, z) -
 * H p with H = [I3 0], S = H P H^T + R inverted by cofactors with the
 * determinant and the inverse in double, K = P H^T S^-1, p' = p + K res (ipt
 * updated as ipt, the state's own parameter), P' = P - K H P; float
 * elsewhere, no contraction. */
static int lo_idx(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
static int up_idx(int i, int j) {
  if (i > j) { const int t = i; i = j; j = t; }
  return i * 6 - i * (i - 1) / 2 + (j - i);
}

static void lane_update_v1_ref(float a[21], float par[6], const float b[6], float xout, float yout) {
  const float m00 = a[0] + b[0], m01 = a[1] + b[1], m02 = a[2] + b[2];
  const float m11 = a[6] + b[3], m12 = a[7] + b[4], m22 = a[11] + b[5];
  const double det = (m00 * ((m11 * m22) - (m12 * m12))) - (m01 * ((m01 * m22) - (m12 * m02))) +
                     (m02 * ((m01 * m12) - (m02 * m11)));
  const double invdet = 1.0 / det;
  float c[9];
  c[0] = invdet * ((m11 * m22) - (m12 * m12));
  c[1] = -1 * invdet * ((m01 * m22) - (m02 * m12));
  c[2] = invdet * ((m01 * m12) - (m02 * m12));
  c[3] = -1 * invdet * ((m01 * m22) - (m12 * m02));
  c[4] = invdet * ((m00 * m22) - (m02 * m02));
  c[5] = -1 * invdet * ((m00 * m12) - (m02 * m01));
  c[6] = invdet * ((m01 * m12) - (m02 * m11));
  c[7] = -1 * invdet * ((m00 * m12) - (m02 * m01));
  c[8] = invdet * ((m00 * m11) - (m01 * m01));
  /* K (OMP:418-436): row i of K uses the words of P(i, 0..2) in the upper
   * addressing: i = 0: 0 1 2, 1: 1 6 7, 2: 2 7 11, 3: 3 8 12, 4: 4 9 13, 5: 5 10 14 */
  static const int row[6][3] = {{0, 1, 2}, {1, 6, 7}, {2, 7, 11}, {3, 8, 12}, {4, 9, 13}, {5, 10, 14}};
  float K[18];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 3; ++j)
      K[3 * i + j] = a[row[i][0]] * c[j] + a[row[i][1]] * c[3 + j] + a[row[i][2]] * c[6 + j];
  const float xin = par[0], yin = par[1], zin = par[2];
  const float ptin = 1. / par[3];
  const float phiin = par[4], thetain = par[5];
  const float rx = xout - xin, ry = yout - yin;
  const float xnew = xin + (K[0] * rx) + (K[1] * ry);
  const float ynew = yin + (K[3] * rx) + (K[4] * ry);
  const float znew = zin + (K[6] * rx) + (K[7] * ry);
  const float ptnew = ptin + (K[9] * rx) + (K[10] * ry);
  const float phinew = phiin + (K[12] * rx) + (K[13] * ry);
  const float thetanew = thetain + (K[15] * rx) + (K[16] * ry);
  /* P' (OMP:451-486): word (i, j) of the upper addressing minus
   * sum_m K(i, m) * word (m, j) */
  float n[21];
  for (int i = 0; i < 6; ++i)
    for (int j = i; j < 6; ++j) {
      const int e = up_idx(i, j);
      n[e] = a[e] - (K[3 * i] * a[up_idx(0, j)] + K[3 * i + 1] * a[up_idx(1, j)] + K[3 * i + 2] * a[up_idx(2, j)]);
    }
  for (int e = 0; e < 21; ++e) a[e] = n[e];
  par[0] = xnew; par[1] = ynew; par[2] = znew;
  par[3] = ptnew; par[4] = phinew; par[5] = thetanew;
}

static void lane_update_v1_fixed(float P[21], float par[6], const float b[6], const float hit[3]) {
  /* S = H P H^T + R, lower packing of both */
  float m[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m[i][j] = P[lo_idx(i, j)] + b[lo_idx(i, j)];
  const double m00 = m[0][0], m01 = m[1][0], m02 = m[2][0], m11 = m[1][1], m12 = m[2][1], m22 = m[2][2];
  const double c00 = m11 * m22 - m12 * m12, c01 = m02 * m12 - m01 * m22, c02 = m01 * m12 - m02 * m11;
  const double c11 = m00 * m22 - m02 * m02, c12 = m01 * m02 - m00 * m12, c22 = m00 * m11 - m01 * m01;
  const double det = m00 * c00 + m01 * c01 + m02 * c02;
  const double id = 1.0 / det;
  const float inv[3][3] = {{(float)(c00 * id), (float)(c01 * id), (float)(c02 * id)},
                           {(float)(c01 * id), (float)(c11 * id), (float)(c12 * id)},
                           {(float)(c02 * id), (float)(c12 * id), (float)(c22 * id)}};
  float K[6][3];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 3; ++j)
      K[i][j] = P[lo_idx(i, 0)] * inv[0][j] + P[lo_idx(i, 1)] * inv[1][j] + P[lo_idx(i, 2)] * inv[2][j];
  const float r[3] = {hit[0] - par[0], hit[1] - par[1], hit[2] - par[2]};
  float n[21];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j)
      n[lo_idx(i, j)] = P[lo_idx(i, j)] -
                        (K[i][0] * P[lo_idx(0, j)] + K[i][1] * P[lo_idx(1, j)] + K[i][2] * P[lo_idx(2, j)]);
  for (int i = 0; i < 6; ++i) par[i] = par[i] + K[i][0] * r[0] + K[i][1] * r[1] + K[i][2] * r[2];
  for (int e = 0; e < 21; ++e) P[e] = n[e];
}

void oracle_kalman_update_v1(float* trkErr, float* inPar, const float* hitErr, const float* msP, int W,
                             int fixed) {
  for (int it = 0; it < W; ++it) {
    float P[21], par[6], hc[6];
    for (int e = 0; e < 21; ++e) P[e] = trkErr[e * W + it];
    for (int e = 0; e < 6; ++e) par[e] = inPar[e * W + it];
    for (int e = 0; e < 6; ++e) hc[e] = hitErr[e * W + it];
    const float hit[3] = {msP[0 * W + it], msP[1 * W + it], msP[2 * W + it]};
    if (fixed)
      lane_update_v1_fixed(P, par, hc, hit);
    else
      lane_update_v1_ref(P, par, hc, hit[0], hit[1]);
    for (int e = 0; e < 21; ++e) trkErr[e * W + it] = P[e];
    for (int e = 0; e < 6; ++e) inPar[e * W + it] = par[e];
  }
}

/* One block through nlayer layers (body of OMP:736-746). */
static void run_block(const float* tb, const float* hb0, float* ob, int W, int nlayer) {
  memcpy(ob, tb, sizeof(float) * ORACLE_TRK_WORDS * W);
  for (int lay = 0; lay < nlayer; ++lay) {
    const float* hb = hb0 + (int64_t)lay * ORACLE_HIT_WORDS * W;
    oracle_propagate_to_z(ob + 6 * W, ob, (const int32_t*)(ob + 27 * W), hb, ob + 6 * W, ob, W);
    oracle_kalman_update_v2(ob + 6 * W, ob, hb + 3 * W, hb, W);
  }
}

void oracle_run(const float* trk, const float* hit, float* out, int64_t nevts, int64_t nb,
                int W, int nlayer, int nthreads) {
  const int64_t nblk = nevts * nb;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t b = 0; b < nblk; ++b)
    run_block(trk + b * ORACLE_TRK_WORDS * W, hit + b * (int64_t)nlayer * ORACLE_HIT_WORDS * W,
              out + b * ORACLE_TRK_WORDS * W, W, nlayer);
  (void)nthreads;
}

void oracle_run_events_only(const float* trk, const float* hit, float* out, int64_t nevts,
                            int64_t nb, int W, int nlayer, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads)
#endif
  for (int64_t ie = 0; ie < nevts; ++ie)
    for (int64_t ib = 0; ib < nb; ++ib) {
      const int64_t b = ib + nb * ie;
      run_block(trk + b * ORACLE_TRK_WORDS * W, hit + b * (int64_t)nlayer * ORACLE_HIT_WORDS * W,
                out + b * ORACLE_TRK_WORDS * W, W, nlayer);
    }
  (void)nthreads;
}

/* Lane l of block tb/hb through n layers (of a block with hit slots at hb). */
static void lane_run_f32(const float* tb, const float* hb0, float* ob, int W, int l, int n) {
  float P[21], par[6];
  for (int e = 0; e < 6; ++e) par[e] = tb[e * W + l];
  for (int e = 0; e < 21; ++e) P[e] = tb[(6 + e) * W + l];
  const int q = ((const int32_t*)tb)[27 * W + l];
  for (int lay = 0; lay < n; ++lay) {
    const float* hb = hb0 + (int64_t)lay * ORACLE_HIT_WORDS * W;
    lane_propagate(P, par, q, hb[2 * W + l]);
    const float hc[3] = {hb[3 * W + l], hb[4 * W + l], hb[5 * W + l]};
    lane_update(P, par, hc, hb[0 * W + l], hb[1 * W + l]);
  }
  for (int e = 0; e < 6; ++e) ob[e * W + l] = par[e];
  for (int e = 0; e < 21; ++e) ob[(6 + e) * W + l] = P[e];
  ((int32_t*)ob)[27 * W + l] = q;
}

static void lane_run_f64(const float* tb, const float* hb0, double* o, int W, int l, int n) {
  double P[21], par[6];
  for (int e = 0; e < 6; ++e) par[e] = tb[e * W + l];
  for (int e = 0; e < 21; ++e) P[e] = tb[(6 + e) * W + l];
  const int q = ((const int32_t*)tb)[27 * W + l];
  for (int lay = 0; lay < n; ++lay) {
    const float* hb = hb0 + (int64_t)lay * ORACLE_HIT_WORDS * W;
    lane_propagate64(P, par, q, hb[2 * W + l]);
    const double hc[3] = {hb[3 * W + l], hb[4 * W + l], hb[5 * W + l]};
    lane_update64(P, par, hc, hb[0 * W + l], hb[1 * W + l]);
  }
  for (int e = 0; e < 6; ++e) o[e] = par[e];
  for (int e = 0; e < 21; ++e) o[6 + e] = P[e];
  o[27] = q;
}

void oracle_run_fp64(const float* trk, const float* hit, double* out64, int64_t nevts, int64_t nb,
                     int W, int nlayer, int nthreads) {
  const int64_t nblk = nevts * nb;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t b = 0; b < nblk; ++b)
    for (int l = 0; l < W; ++l)
      lane_run_f64(trk + b * ORACLE_TRK_WORDS * W, hit + b * (int64_t)nlayer * ORACLE_HIT_WORDS * W,
                   out64 + (b * W + l) * 28, W, l, nlayer);
  (void)nthreads;
}

void oracle_run_ragged(const float* trk, const float* hit, float* out, double* out64,
                       int64_t nblocks, int W, const int64_t* blk_trk, const int64_t* blk_hit,
                       const int32_t* blk_nlay, const int32_t* lane_nlay, int fp64, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
#endif
  for (int64_t b = 0; b < nblocks; ++b) {
    const int64_t tb = blk_trk[b];
    const float* tin = trk + tb * ORACLE_TRK_WORDS * W;
    const float* hb = hit + blk_hit[b] * ORACLE_HIT_WORDS * W;
    for (int l = 0; l < W; ++l) {
      int n = lane_nlay[tb * W + l];
      if (n > blk_nlay[b]) n = blk_nlay[b];
      if (fp64)
        lane_run_f64(tin, hb, out64 + (tb * W + l) * 28, W, l, n);
      else
        lane_run_f32(tin, hb, out + tb * ORACLE_TRK_WORDS * W, W, l, n);
    }
  }
  (void)nthreads;
}

/* ========================================================================= */
/* report                                                                    */
/* ========================================================================= */

/* Statistics of OMP:758-851: double accumulation over the tracks that are
 * finite and within the reference's acceptance window, in event/track order. */
void oracle_summary(const float* out, int64_t nevts, int64_t nb, int W, int64_t ntrks,
                    const float lasthit[3], oracle_summary_t* s) {
  memset(s, 0, sizeof *s);
  const double hx = lasthit[0], hy = lasthit[1], hz = lasthit[2];
  for (int pass = 0; pass < 2; ++pass) {
    for (int64_t ie = 0; ie < nevts; ++ie)
      for (int64_t it = 0; it < ntrks; ++it) {
        const float* blk = out + ((it / W) + nb * ie) * (int64_t)ORACLE_TRK_WORDS * W;
        const int64_t l = it % W;
        const double x = blk[0 * W + l], y = blk[1 * W + l], z = blk[2 * W + l];
        const double pt = 1. / blk[3 * W + l];
        const double phi = blk[4 * W + l], theta = blk[5 * W + l];
        if (!isfinite(x) || !isfinite(y) || !isfinite(z) || !isfinite(pt) || !isfinite(phi) ||
            !isfinite(theta)) {
          if (pass == 0) s->nnans++;
          continue;
        }
        if (fabs((x - hx) / hx) > 1. || fabs((y - hy) / hy) > 1. || fabs((z - hz) / hz) > 1. ||
            fabs((pt - 12.) / 12.) > 1. || fabs((phi - 1.3) / 1.3) > 1. ||
            fabs((theta - 2.8) / 2.8) > 1.) {
          if (pass == 0) s->nfail++;
          continue;
        }
        if (pass == 0) {
          s->avgpt += pt; s->avgphi += phi; s->avgtheta += theta;
          s->avgx += x; s->avgy += y; s->avgz += z;
          s->avgdx += (x - hx) / x; s->avgdy += (y - hy) / y; s->avgdz += (z - hz) / z;
          s->ngood++;
        } else {
          s->stdx += (x - s->avgx) * (x - s->avgx);
          s->stdy += (y - s->avgy) * (y - s->avgy);
          s->stdz += (z - s->avgz) * (z - s->avgz);
          const double ddx = (x - hx) / x - s->avgdx, ddy = (y - hy) / y - s->avgdy,
                       ddz = (z - hz) / z - s->avgdz;
          s->stddx += ddx * ddx; s->stddy += ddy * ddy; s->stddz += ddz * ddz;
        }
      }
    if (pass == 0) {
      const double n = (double)s->ngood;
      s->avgpt /= n; s->avgphi /= n; s->avgtheta /= n;
      s->avgx /= n; s->avgy /= n; s->avgz /= n;
      s->avgdx /= n; s->avgdy /= n; s->avgdz /= n;
    }
  }
  const double n = (double)s->ngood;
  /* the reference takes sqrtf of a double: the result is rounded to float */
  s->stdx = sqrtf(s->stdx / n); s->stdy = sqrtf(s->stdy / n); s->stdz = sqrtf(s->stdz / n);
  s->stddx = sqrtf(s->stddx / n); s->stddy = sqrtf(s->stddy / n); s->stddz = sqrtf(s->stddz / n);
}

int oracle_format_summary(const oracle_summary_t* s, char* buf, int cap) {
  return snprintf(buf, (size_t)cap,
                  "track x avg=%f std/avg=%f\n"
                  "track y avg=%f std/avg=%f\n"
                  "track z avg=%f std/avg=%f\n"
                  "track dx/x avg=%f std=%f\n"
                  "track dy/y avg=%f std=%f\n"
                  "track dz/z avg=%f std=%f\n"
                  "track pt avg=%f\n"
                  "track phi avg=%f\n"
                  "track theta avg=%f\n"
                  "number of tracks with nans=%i\n"
                  "number of tracks failed=%i\n",
                  s->avgx, fabs(s->stdx / s->avgx), s->avgy, fabs(s->stdy / s->avgy), s->avgz,
                  fabs(s->stdz / s->avgz), s->avgdx, s->stddx, s->avgdy, s->stddy, s->avgdz,
                  s->stddz, s->avgpt, s->avgphi, s->avgtheta, s->nnans, s->nfail);
}

/* ---- the C library's sin / cos ----------------------------------------- */
/* Through volatile pointers so that no compiler merges the pair into one
 * sincosf call: this checks sinf and cosf themselves. */
static float (*volatile libm_sinf)(float) = sinf;
static float (*volatile libm_cosf)(float) = cosf;

void oracle_libm_sincosf_range(uint32_t first_bits, int64_t n, float* s, float* c, int nthreads) {
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const uint32_t u = (uint32_t)(first_bits + (uint64_t)i);
    float y;
    memcpy(&y, &u, sizeof y);
    s[i] = libm_sinf(y);
    c[i] = libm_cosf(y);
  }
}
