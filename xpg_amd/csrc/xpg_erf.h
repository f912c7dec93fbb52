// Branch-free float64 erf for gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace xpg {

// float64 erf without data-dependent branches (the lanes of one wavefront spread over every argument range, so a
// piecewise erf runs several of its pieces per call): x sum_n a_n (x^2)^n up to ERF_SWITCH, then
// 1 - exp(-x^2) erfcx(|x|) with erfcx a Chebyshev series in u = affine((|x| - K) / (|x| + K)), +-1 beyond ERF_TMAX.
// generated by gen_erf_coeffs.py: K = 2.5, switch 0.75, saturation 6.75, 20 Chebyshev terms, 16 series terms;
// worst |error| of a float64 model 1.31e-16 (1.44 ulp) at 1500 points in [0, 7]
constexpr double ERF_K = 2.5, ERF_SWITCH = 0.75, ERF_TMAX = 6.75, ERF_U1 = 2.0041666666666664, ERF_U0 = 0.07916666666666661;
constexpr double ERF_SERIES[16] = {1.1283791670955126, -0.37612638903183754, 0.11283791670955126, -0.026866170645131252, 0.005223977625442188, -0.0008548327023450853, 0.00012055332981789664, -1.492565035840625e-05, 1.6462114365889248e-06, -1.6365844691234924e-07, 1.4807192815879218e-08, -1.2290555301717928e-09, 9.422759064650411e-11, -6.7113668551641105e-12, 4.4632242632864775e-13, -2.7835162072109215e-14};
constexpr double ERFCX_CHEB[20] = {0.2600467940043389, -0.2080045476633964, 0.03446404496951066, -0.004111023196646251, 0.000306831125612948, -5.81367467314241e-06, -1.3239378399511994e-06, 8.854924007498369e-08, 7.43838558964106e-09, -7.799167648452338e-10, -6.671900193697053e-11, 6.386684292197406e-12, 8.204969190235308e-13, -3.999179974273952e-14, -1.0816524728010787e-14, -8.048013407626442e-17, 1.2646552748718328e-16, 9.255643960100123e-18, -9.8266808262368e-19, -2.0382673468958777e-19};

__device__ __forceinline__ double erf_bf(double x) {
  const double a = fabs(x), z = x * x;
  double s = ERF_SERIES[15];
#pragma unroll
  for (int n = 14; n >= 0; --n) s = fma(s, z, ERF_SERIES[n]);
  const double small = x * s;
  const double u = fma((a - ERF_K) / (a + ERF_K), ERF_U1, ERF_U0), u2 = u + u;
  double b1 = 0.0, b2 = 0.0;
#pragma unroll
  for (int j = 19; j >= 1; --j) {
    const double t = fma(u2, b1, ERFCX_CHEB[j] - b2);
    b2 = b1;
    b1 = t;
  }
  const double g = fma(u, b1, ERFCX_CHEB[0] - b2);
  const double big = copysign(fma(-exp(-z), g, 1.0), x);
  return a <= ERF_SWITCH ? small : (a > ERF_TMAX ? copysign(1.0, x) : big);
}

}  // namespace xpg
