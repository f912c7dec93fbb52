// float64 erf for gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "xpg_erf_table.h"

namespace xpg {

// float64 erf from a table of erf and erf' at x = i / 256 (xpg_erf_table.h, generated by gen_erf_table.py): the
// degree-5 Taylor expansion about the nearest node (|x - node| <= 1/512, truncation < 4e-18), one 16-byte table read
// and no data-dependent branch; +-1 beyond |x| = 6, where float64 erf rounds to +-1.  Worst error 1.25 ulp (the
// generator emulates these operations with exact rational arithmetic and compares with 80-digit erf).
__host__ __device__ __forceinline__ double erf_tab(double x) {
  const double a = fabs(x);
  const int i = (int)fmin(rint(a * (double)ERF_TAB_INV_H), (double)(ERF_TAB_N - 1));   // NaN -> the last node
  const double x0 = (double)i / (double)ERF_TAB_INV_H;
  const double d = a - x0;                 // exact (Sterbenz) below the last node
  const ErfNode e = ERF_TAB[i];
  const double x2 = x0 * x0;
  const double c2 = -x0;
  const double c3 = (2.0 * x2 - 1.0) * (1.0 / 3.0);
  const double c4 = (x0 * (3.0 - 2.0 * x2)) * (1.0 / 6.0);
  const double c5 = (x2 * (4.0 * x2 - 12.0) + 3.0) * (1.0 / 30.0);
  double q = fma(c5, d, c4);
  q = fma(q, d, c3);
  q = fma(q, d, c2);
  q = fma(q, d, 1.0);
  const double r = fma(e.f1 * d, q, e.f0);
  return copysign(a > 6.0 ? 1.0 : r, x);
}

// float64 erfc for x >= 0 with relative precision in the tail: the same expansion taken from erfc at the node
// (ERFC_TAB), erfc(x0 + d) = erfc(x0) - erf'(x0) d q(d), whose correction stays below 3% of erfc(x0) up to x = 6, so
// that nothing cancels (worst relative error 1.9e-13, near x = 6; 1 - erf_tab would carry an absolute 1e-16).  Beyond
// the last node 0 (erfc < 2.2e-17 there, and float64 erf rounds to 1, so differences of erf values are 0 as well);
// NaN propagates.
__host__ __device__ __forceinline__ double erfc_tab(double x) {
  const int i = (int)fmin(rint(x * (double)ERF_TAB_INV_H), (double)(ERF_TAB_N - 1));   // NaN -> the last node
  const double x0 = (double)i / (double)ERF_TAB_INV_H;
  const double d = x - x0;
  const ErfcNode e = ERFC_TAB[i];
  const double x2 = x0 * x0;
  const double c2 = -x0;
  const double c3 = (2.0 * x2 - 1.0) * (1.0 / 3.0);
  const double c4 = (x0 * (3.0 - 2.0 * x2)) * (1.0 / 6.0);
  const double c5 = (x2 * (4.0 * x2 - 12.0) + 3.0) * (1.0 / 30.0);
  double q = fma(c5, d, c4);
  q = fma(q, d, c3);
  q = fma(q, d, c2);
  q = fma(q, d, 1.0);
  const double r = fma(-(e.f1 * d), q, e.c0);
  return x <= 6.0 ? r : (x > 6.0 ? 0.0 : x);          // selects, no branch
}

// erf(a) - erf(b) for a >= b, from ea = erfc|a| and eb = erfc|b|: eb - ea with both arguments >= 0, ea - eb with
// both <= 0 (there erf rounds towards +-1 and a difference of erf values would keep only an absolute 1e-16, nothing
// of a tail value of 1e-12), (2 - ea) - eb across zero (no cancellation).  The same two table evaluations on every
// lane.
__host__ __device__ __forceinline__ double erf_diff(double a, double b) {
  const double ea = erfc_tab(fabs(a)), eb = erfc_tab(fabs(b));
  return b >= 0.0 ? eb - ea : (a <= 0.0 ? ea - eb : (2.0 - ea) - eb);
}

}  // namespace xpg
