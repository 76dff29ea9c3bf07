// The local bond order Q6 of mc2's umbrella (Q6Local.F90, q6.F90): the
// spherical harmonics Y_6m of a bond direction as the reference evaluates
// them.  Shared by the order-parameter kernels and the MC trial kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#pragma clang fp contract(off)

namespace g6 {

constexpr int kMaxBond = 10;  // MAXBOND (Q6Local.F90:15): bonds within 3 A per molecule
constexpr double kPiT = 3.1415926535897932;  // PItanaka (PhysConst.F90:18)

// plgndr (q6.F90): associated Legendre function, upward recurrence in l
__device__ inline double q6_plgndr(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    const double somx2 = sqrt((1.0 - x) * (1.0 + x));
    double fact = 1.0;
    for (int i = 1; i <= m; i++) {
      pmm = -pmm * fact * somx2;
      fact = fact + 2.0;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2 * m + 1) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pll = (x * (2 * ll - 1) * pmmp1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

// qlm(6, m, x, y, z) of q6.F90 for m = 0..6 (q[2m], q[2m+1]); m < 0 is
// (-1)^m conjg of these.  As the reference evaluates it: w = cmplx(x, y)
// (rounded to single precision) divided by |(x, y)| as a complex division,
// w**m by repeated squaring, the prefactor times plgndr before the power.
__device__ inline void q6_ylm6(double x, double y, double z, double (&q)[14]) {
  const double prod[13] = {1.0, 1.0, 2.0, 6.0, 24.0, 120.0, 720.0, 5040.0, 40320.0,
                           362880.0, 3628800.0, 39916800.0, 479001600.0};
  const double r = sqrt(x * x + y * y);
  const double fa = (double)(float)x, fb = (double)(float)y, den = r * r + 0.0 * 0.0;
  const double wa = (fa * r + fb * 0.0) / den, wb = (fb * r - fa * 0.0) / den;
  for (int m = 0; m <= 6; m++) {
    const double c = sqrt((2.0 * 6 + 1.0) * prod[6 - m] / (4.0 * kPiT * prod[6 + m])) *
                     q6_plgndr(6, m, z);
    double pa = 1.0, pb = 0.0, xa = wa, xb = wb;
    int k = m;
    while (k > 0) {
      if (k & 1) {
        const double ta = pa * xa - pb * xb, tb = pa * xb + pb * xa;
        pa = ta;
        pb = tb;
      }
      k >>= 1;
      if (k > 0) {
        const double ta = xa * xa - xb * xb, tb = xa * xb + xb * xa;
        xa = ta;
        xb = tb;
      }
    }
    q[2 * m] = c * pa;
    q[2 * m + 1] = c * pb;
  }
}

}  // namespace g6
