// Host program for tests/test_lane_cov.py: runs the fast modes' covariance
// propagation (p2z-tests_amd/csrc/p2z_cov.hpp) on the host.  Reads cases of
// 11 Jacobian entries (J02 J03 J04 J05 J12 J13 J14 J15 J42 J43 J45) and 21
// packed covariance words from stdin; prints the 21 propagated words.
#define __host__
#define __device__
#include <cstdio>

#include "p2z_cov.hpp"

int main() {
  p2z::Jac J;
  float P[21];
  float* j = &J.J02;
  for (;;) {
    for (int i = 0; i < 11; ++i)
      if (std::scanf("%f", &j[i]) != 1) return 0;
    for (int i = 0; i < 21; ++i)
      if (std::scanf("%f", &P[i]) != 1) return 1;
    p2z::propagate_cov_fast_z0(P, J);
    for (int i = 0; i < 21; ++i) std::printf("%.9g%c", P[i], i == 20 ? '\n' : ' ');
  }
}
