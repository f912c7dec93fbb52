// Device helpers shared by the integration kernels (gfx950, wave64).
//
// Precision plan: per-reflection, per-pixel and per-frame geometry (Kabsch frame, eps, pixel area J, rocking-curve
// fractions R, masks) in float64, in the operation order of the float64 oracle so that every threshold-derived mask
// is identical; per-voxel arithmetic in float32 with compensated (TwoSum) accumulation.  Compiled with
// -ffp-contract=off: no fused multiply-adds change the float64 roundings.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "xpg_erf.h"
#include "xpg_integrate.h"

namespace xpg {

constexpr int WAVE = 64;
constexpr double PIX0 = 1.0;        // coordinate of the centre of array element 0
constexpr double UMAX = 4.5;        // N_BOX * 0.9: half-width of the profile grid (sigma units)
constexpr int G = XPG_PROFILE_GRID;
constexpr int RCK = XPG_RC_K;
// fixed-point profile grids: value * 2^44 in int64.  A contribution is rounded to 2^-44 (6e-14), so that the learned
// float32 profiles stay within ~1e-12 of exact float64 sums (deblended intensities next to partners 1e5 x brighter
// depend on them to that level); a grid cell holds up to 2^19 (524288) in value units: the weights and densities of
// that many strong-row pixels
constexpr double FX_SCALE = 17592186044416.0;   // 2^44
constexpr double VAR_FLOOR = 1e-3;
constexpr int N_IRLS = 6;
constexpr double IRLS_TOL = 1e-3;

// rocking-curve mixture centres linspace(-3.5, 3.5, 21) as float32 values (the reference's device tensor)
__device__ __constant__ static const float RC_C32[RCK] = {
    -3.5f, -3.1500000953674316f, -2.799999952316284f, -2.450000047683716f, -2.0999999046325684f, -1.75f,
    -1.399999976158142f, -1.0499999523162842f, -0.699999988079071f, -0.3499999940395355f, 0.0f,
    0.3499999940395355f, 0.699999988079071f, 1.0499999523162842f, 1.399999976158142f, 1.75f,
    2.0999999046325684f, 2.450000047683716f, 2.799999952316284f, 3.1500000953674316f, 3.5f};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- cross-lane reductions (DPP within rows of 16 lanes, then the four row results by readlane) -----------------
// Every lane returns the same value; call with all 64 lanes active.

template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dppi(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
__device__ __forceinline__ double dppd(double v) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float rlf(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double rld(double v, int l) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
constexpr int DPP_Q1 = 0xB1, DPP_Q2 = 0x4E, DPP_R4 = 0x124, DPP_R8 = 0x128;

// ---- compensated float32 sums -----------------------------------------------------------------------------------

struct csum {
  float s, c;
};
__device__ __forceinline__ csum cz() { return csum{0.f, 0.f}; }
// TwoSum (Knuth): t + e == a.s + x exactly
__device__ __forceinline__ void cadd(csum &a, float x) {
  float t = a.s + x;
  float bp = t - a.s;
  float e = (a.s - (t - bp)) + (x - bp);
  a.s = t;
  a.c += e;
}
__device__ __forceinline__ void cmerge(csum &a, float s2, float c2) {
  float t = a.s + s2;
  float bp = t - a.s;
  float e = (a.s - (t - bp)) + (s2 - bp);
  a.s = t;
  a.c = (a.c + c2) + e;
}
__device__ __forceinline__ double creduce(csum a) {
  cmerge(a, dppf<DPP_Q1>(a.s), dppf<DPP_Q1>(a.c));
  cmerge(a, dppf<DPP_Q2>(a.s), dppf<DPP_Q2>(a.c));
  cmerge(a, dppf<DPP_R4>(a.s), dppf<DPP_R4>(a.c));
  cmerge(a, dppf<DPP_R8>(a.s), dppf<DPP_R8>(a.c));
  // the four row sums, combined in float64 in a fixed order
  double r0 = (double)rlf(a.s, 0) + (double)rlf(a.c, 0), r1 = (double)rlf(a.s, 16) + (double)rlf(a.c, 16);
  double r2 = (double)rlf(a.s, 32) + (double)rlf(a.c, 32), r3 = (double)rlf(a.s, 48) + (double)rlf(a.c, 48);
  return (r0 + r1) + (r2 + r3);
}
__device__ __forceinline__ double wsum_d(double v) {
  v += dppd<DPP_Q1>(v);
  v += dppd<DPP_Q2>(v);
  v += dppd<DPP_R4>(v);
  v += dppd<DPP_R8>(v);
  return (rld(v, 0) + rld(v, 16)) + (rld(v, 32) + rld(v, 48));
}
// Sums of V float64 values over the wavefront at once (V = 1, 2, 4, 8 or 16; every lane gets every sum): butterflies
// over lane bits 0 .. 3 in which a lane keeps half of its values at each step (so V values cost about as much as two
// wsum_d), then the two row bits; the row sums combine as ((r0 + r1) + (r2 + r3)).
template <int S>
__device__ __forceinline__ double xchg_row(double x) {
  // the value of the lane paired with this one at step S: lane ^ 1, lane ^ 2, then a rotation by 4 within the row of
  // 16 (a partner whose bits 0-1 match and bit 2 differs: the halving steps below stay exact), then lane ^ 8
  if constexpr (S == 0) return dppd<DPP_Q1>(x);
  else if constexpr (S == 1) return dppd<DPP_Q2>(x);
  else if constexpr (S == 2) return dppd<DPP_R4>(x);
  else return dppd<DPP_R8>(x);                                   // rotation by 8 in a row of 16: lane ^ 8
}
// x[lane] + x[lane ^ 16] (and ^ 32) on every lane, by the gfx950 row / half-wave swaps: the pair a, b of a swap holds
// this lane's value and its partner's, and IEEE addition is commutative, so both lanes of a pair get the same bits
__device__ __forceinline__ double add_xor16(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)b, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
  const double a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double c = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return a + c;
}
__device__ __forceinline__ double add_xor32(double x) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)b, (uint32_t)b, false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(b >> 32), (uint32_t)(b >> 32), false, false);
  const double a = __builtin_bit_cast(double, ((uint64_t)hi[0] << 32) | lo[0]);
  const double c = __builtin_bit_cast(double, ((uint64_t)hi[1] << 32) | lo[1]);
  return a + c;
}
template <int S, int N>
__device__ __forceinline__ void wsum_step(double *v, int lane) {
  if constexpr (N == 1) {
    v[0] = v[0] + xchg_row<S>(v[0]);
  } else {
    constexpr int H = N / 2;
    const bool hi = (lane >> S) & 1;                             // keeps values H .. N - 1
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const double keep = hi ? v[i + H] : v[i], send = hi ? v[i] : v[i + H];
      v[i] = keep + xchg_row<S>(send);
    }
  }
}
template <int V>
__device__ __forceinline__ void wsum_multi(double (&v)[V]) {
  static_assert(V == 1 || V == 2 || V == 4 || V == 8 || V == 16, "V: a power of two up to 16");
  const int lane = threadIdx.x & 63;
  wsum_step<0, V>(v, lane);
  wsum_step<1, (V > 1 ? V / 2 : 1)>(v, lane);
  wsum_step<2, (V > 2 ? V / 4 : 1)>(v, lane);
  wsum_step<3, (V > 4 ? V / 8 : 1)>(v, lane);
  double t = v[0];                       // lane l holds the row sum of value bitreverse(l mod V)
  t = add_xor16(t);
  t = add_xor32(t);
  constexpr int B = V == 16 ? 4 : (V == 8 ? 3 : (V == 4 ? 2 : (V == 2 ? 1 : 0)));
#pragma unroll
  for (int j = 0; j < V; ++j) {
    int L = 0;
#pragma unroll
    for (int b = 0; b < B; ++b) L |= ((j >> b) & 1) << (B - 1 - b);
    v[j] = rld(t, L);
  }
}
__device__ __forceinline__ int wsum_i(int v) {
  v += dppi<DPP_Q1>(v);
  v += dppi<DPP_Q2>(v);
  v += dppi<DPP_R4>(v);
  v += dppi<DPP_R8>(v);
  return (__builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16)) +
         (__builtin_amdgcn_readlane(v, 32) + __builtin_amdgcn_readlane(v, 48));
}
__device__ __forceinline__ double wmax_d(double v) {
  v = fmax(v, dppd<DPP_Q1>(v));
  v = fmax(v, dppd<DPP_Q2>(v));
  v = fmax(v, dppd<DPP_R4>(v));
  v = fmax(v, dppd<DPP_R8>(v));
  return fmax(fmax(rld(v, 0), rld(v, 16)), fmax(rld(v, 32), rld(v, 48)));
}
__device__ __forceinline__ bool wany(bool b) { return __ballot(b) != 0; }

// Voxels of a (Dz, P) box visited by one lane: v = lane, lane + 64, ... with (j, p) = divmod(v, P) kept incrementally
struct VoxIter {
  int v, j, p, dj, dp, P, V;
  __device__ __forceinline__ VoxIter(int lane, int P_, int V_) : P(P_), V(V_) {
    v = lane;
    j = lane / P_;
    p = lane - j * P_;
    dj = WAVE / P_;
    dp = WAVE - dj * P_;
  }
  __device__ __forceinline__ bool ok() const { return v < V; }
  __device__ __forceinline__ void next() {
    v += WAVE;
    j += dj;
    p += dp;
    if (p >= P) {
      p -= P;
      ++j;
    }
  }
};

// ---- float64 vector helpers (left-to-right order) ---------------------------------------------------------------

__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---- detector: S of a detector point (xpg/integrate_kernels.py:78-105) -------------------------------------------

__device__ __forceinline__ double bilinear_table(const float *tab, int nx, int ny, double x, double y) {
  double fx = fmin(fmax(x - 1.0, 0.0), (double)nx - 1.000001);
  double fy = fmin(fmax(y - 1.0, 0.0), (double)ny - 1.000001);
  int ix = (int)floor(fx), iy = (int)floor(fy);
  int ix1 = min(ix + 1, nx - 1), iy1 = min(iy + 1, ny - 1);
  double tx = fx - ix, ty = fy - iy;
  double a = (double)tab[(int64_t)iy * nx + ix], b = (double)tab[(int64_t)iy * nx + ix1];
  double c = (double)tab[(int64_t)iy1 * nx + ix], d = (double)tab[(int64_t)iy1 * nx + ix1];
  return (((a * (1 - tx)) * (1 - ty) + (b * tx) * (1 - ty)) + (c * (1 - tx)) * ty) + (d * tx) * ty;
}

__device__ void pixel_S(const xpg_detector &d, double x, double y, double *S) {
  int k = 0;
  if (d.segmap) {
    long ix = (long)floor(x - 0.5), iy = (long)floor(y - 0.5);
    ix = ix < 0 ? 0 : (ix > d.seg_nx - 1 ? d.seg_nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy > d.seg_ny - 1 ? d.seg_ny - 1 : iy);
    k = min(max((int)d.segmap[iy * d.seg_nx + ix], 0), d.n_seg - 1);   // gaps (-1) use segment 0
  }
  const double *P0 = d.P0 + 3 * k, *Bx = d.Bx + 3 * k, *By = d.By + 3 * k;
  double xc = x, yc = y;
  if (d.gx) {
    xc = x + bilinear_table(d.gx, d.geo_nx, d.geo_ny, x, y);
    yc = y + bilinear_table(d.gy, d.geo_nx, d.geo_ny, x, y);
  }
  double X[3];
  {
    double a = xc - d.orgx, b = yc - d.orgy;
    for (int i = 0; i < 3; ++i) X[i] = (P0[i] + a * Bx[i]) + b * By[i];
  }
  if (d.plx != 0.0) {
    const double *U1 = d.U1 + 3 * k, *U2 = d.U2 + 3 * k, *UN = d.UN + 3 * k;
    for (int it = 0; it < 2; ++it) {
      double nrm = sqrt(dot3(X, X));
      double e[3] = {X[0] / nrm, X[1] / nrm, X[2] / nrm};
      double cosa = fmin(fmax(dot3(e, UN), 1e-3), 1.0);
      double L = d.tsens / cosa;
      double muL = d.mu * L;
      double lbar = 1.0 / d.mu - (L * exp(-muL)) / fmax(-expm1(-muL), 1e-12);
      double d1 = lbar * dot3(e, U1), d2 = lbar * dot3(e, U2);
      double a = (xc - (d.plx * d1) / d.qx) - d.orgx, b = (yc - (d.plx * d2) / d.qy) - d.orgy;
      for (int i = 0; i < 3; ++i) X[i] = (P0[i] + a * Bx[i]) + b * By[i];
    }
  }
  double den = sqrt(dot3(X, X)) * d.wl;
  S[0] = X[0] / den;
  S[1] = X[1] / den;
  S[2] = X[2] / den;
}

// Kabsch frame of one reflection: e1 = S x s0 / |.|, e2 = S x e1 / |.|, zeta = e1 . axis (xpg/integrate_kernels.py:234-238)
struct Frame {
  double S[3], e1[3], e2[3], invS, zeta;
};
__device__ __forceinline__ void kabsch(const double *S, const double *s0, const double *axis, Frame &f) {
  double c1[3], c2[3];
  for (int i = 0; i < 3; ++i) f.S[i] = S[i];
  cross3(S, s0, c1);
  double n1 = sqrt(dot3(c1, c1));
  for (int i = 0; i < 3; ++i) f.e1[i] = c1[i] / n1;
  cross3(S, f.e1, c2);
  double n2 = sqrt(dot3(c2, c2));
  for (int i = 0; i < 3; ++i) f.e2[i] = c2[i] / n2;
  f.zeta = dot3(f.e1, axis);
  f.invS = 1.0 / sqrt(dot3(S, S));
}

// eps1, eps2 of detector point (x, y) (xpg/integrate_kernels.py:108-115)
__device__ __forceinline__ void eps12(const xpg_detector &d, const Frame &f, double x, double y, double &e1,
                                      double &e2) {
  double Sp[3], dd[3];
  pixel_S(d, x, y, Sp);
  for (int i = 0; i < 3; ++i) dd[i] = Sp[i] - f.S[i];
  e1 = dot3(dd, f.e1) * f.invS;
  e2 = dot3(dd, f.e2) * f.invS;
}

// ---- rocking curve (xpg/integrate_kernels.py:130-141), one mixture term ------------------------------------------

// effective widths of a launch: from the device widths record when set (adaptive chain on the device)
__device__ __forceinline__ double sd_of(const xpg_call &c) { return c.widths ? c.widths->sd : c.sd; }
__device__ __forceinline__ double sm_of(const xpg_call &c) { return c.widths ? c.widths->sm : c.sm; }
__device__ __forceinline__ bool use_rcw_of(const xpg_call &c) { return c.widths != nullptr || c.use_rcw != 0; }
__device__ __forceinline__ float rcw_of(const xpg_call &c, int j) { return c.widths ? c.widths->rcw[j] : c.rcw[j]; }

// term j of frame: (w_j * 0.5) * (erf(a) - erf(b)) (erf_diff: tail values keep their relative precision); plain
// Gaussian when no mixture weights (j == 0 only).  s: the width in phi, isw = 1 / sw with sw = (s / 4) sqrt(2) the
// mixture components' scale (a product instead of a quotient: the erf arguments move by at most an ulp)
__device__ __forceinline__ double rc_term(const xpg_call &c, double lo, double hi, double phi, double s, double isw,
                                          int j) {
  if (!use_rcw_of(c)) {
    double a = (hi - phi) / (s * sqrt(2.0));
    double b = (lo - phi) / (s * sqrt(2.0));
    return 0.5 * erf_diff(a, b);
  }
  double cj = (double)RC_C32[j];
  double a = ((hi - phi) - cj * s) * isw;
  double b = ((lo - phi) - cj * s) * isw;
  return ((double)rcw_of(c, j) * 0.5) * erf_diff(a, b);
}

__device__ __forceinline__ void frame_bounds(const xpg_call &c, int iz0, int fb, int j, double &lo, double &hi) {
  double k = (double)iz0 + (double)fb * (double)j;
  double plo = c.frames_phi0 + c.osc * k;
  double phi_ = plo + c.osc * (double)fb;
  lo = fmin(plo, phi_);
  hi = fmax(plo, phi_);
}

// R[j] of one frame: the mixture terms summed in component order (the arithmetic of rocking_curve for one frame)
__device__ __forceinline__ double rocking_frame(const xpg_call &c, int iz0, int fb, double phi, double zeta, int j) {
  const int nt = use_rcw_of(c) ? RCK : 1;
  const double s = sm_of(c) / fmax(fabs(zeta), 1e-3);
  const double isw = 1.0 / ((s * 0.25) * sqrt(2.0));
  double lo, hi;
  frame_bounds(c, iz0, fb, j, lo, hi);
  if (nt == 1) return rc_term(c, lo, hi, phi, s, isw, 0);
  double out = 0.0;
  for (int q = 0; q < nt; ++q) out = out + rc_term(c, lo, hi, phi, s, isw, q);
  return out;
}

// R[j] for all frames of a reflection into R (LDS), using scratch (>= Dz * 21 doubles); the mixture terms are summed
// in component order as the oracle does.
__device__ void rocking_curve(const xpg_call &c, int iz0, int fb, double phi, double zeta, int Dz, double *scratch,
                              double *R, int lane) {
  const int nt = use_rcw_of(c) ? RCK : 1;
  const double s = sm_of(c) / fmax(fabs(zeta), 1e-3);
  const double isw = 1.0 / ((s * 0.25) * sqrt(2.0));
  constexpr int RB = 3;                      // terms per lane at a time: their erf table reads overlap
  for (int t0 = lane; t0 < Dz * nt; t0 += RB * WAVE) {
    double v[RB];
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int t = min(t0 + i * WAVE, Dz * nt - 1);
      const int j = nt == RCK ? t / RCK : t, q = t - j * nt;     // a constant divisor: multiply and shift
      double lo, hi;
      frame_bounds(c, iz0, fb, j, lo, hi);
      v[i] = rc_term(c, lo, hi, phi, s, isw, q);
    }
#pragma unroll
    for (int i = 0; i < RB; ++i)
      if (t0 + i * WAVE < Dz * nt) scratch[t0 + i * WAVE] = v[i];
  }
  wave_sync();
  for (int j = lane; j < Dz; j += WAVE) {
    double out = 0.0;
    if (nt == 1) {
      out = scratch[j];
    } else {
      double v[RCK];               // every load in flight at once; the sum keeps component order
#pragma unroll
      for (int q = 0; q < RCK; ++q) v[q] = scratch[j * RCK + q];
#pragma unroll
      for (int q = 0; q < RCK; ++q) out = out + v[q];
    }
    R[j] = out;
  }
  wave_sync();
}

// ---- per-reflection eps / J / seff on the box grid (xpg/integrate_kernels.py:224-272) -----------------------------

__device__ __forceinline__ double pix_x(int ix0, int i, int s) {
  return ((double)ix0 + (double)i) * (double)s + ((double)s - 1.0) * 0.5 + PIX0;
}

// The widths-independent half: eps1, eps2 (float64) of every pixel centre into U1, U2 (P each) and the pixel area J
// into Jd (float64, P).  Small boxes (a side < 3): J by central differences at the pixel edges.
__device__ void box_eps_J(const xpg_detector &d, const Frame &f, int ix0, int iy0, int s, int Dy, int Dx, double *U1,
                          double *U2, double *Jd, int lane) {
  const int P = Dy * Dx;
  for (int p = lane; p < P; p += WAVE) {
    int iy = p / Dx, ix = p - iy * Dx;
    double x = pix_x(ix0, ix, s), y = pix_x(iy0, iy, s);
    double a, b;
    eps12(d, f, x, y, a, b);
    U1[p] = a;
    U2[p] = b;
  }
  wave_sync();
  if (Dx >= 3 && Dy >= 3) {
    for (int p = lane; p < P; p += WAVE) {
      int iy = p / Dx, ix = p - iy * Dx;
      double g1x, g1y, g2x, g2y;
      // d/dx
      if (ix == 0) {
        g1x = (-1.5 * U1[p] + 2.0 * U1[p + 1]) + -0.5 * U1[p + 2];
        g2x = (-1.5 * U2[p] + 2.0 * U2[p + 1]) + -0.5 * U2[p + 2];
      } else if (ix == Dx - 1) {
        g1x = (0.5 * U1[p - 2] + -2.0 * U1[p - 1]) + 1.5 * U1[p];
        g2x = (0.5 * U2[p - 2] + -2.0 * U2[p - 1]) + 1.5 * U2[p];
      } else {
        g1x = (U1[p + 1] - U1[p - 1]) / 2.0;
        g2x = (U2[p + 1] - U2[p - 1]) / 2.0;
      }
      if (iy == 0) {
        g1y = (-1.5 * U1[p] + 2.0 * U1[p + Dx]) + -0.5 * U1[p + 2 * Dx];
        g2y = (-1.5 * U2[p] + 2.0 * U2[p + Dx]) + -0.5 * U2[p + 2 * Dx];
      } else if (iy == Dy - 1) {
        g1y = (0.5 * U1[p - 2 * Dx] + -2.0 * U1[p - Dx]) + 1.5 * U1[p];
        g2y = (0.5 * U2[p - 2 * Dx] + -2.0 * U2[p - Dx]) + 1.5 * U2[p];
      } else {
        g1y = (U1[p + Dx] - U1[p - Dx]) / 2.0;
        g2y = (U2[p + Dx] - U2[p - Dx]) / 2.0;
      }
      Jd[p] = fabs(g1x * g2y - g2x * g1y);
    }
  } else {
    for (int p = lane; p < P; p += WAVE) {
      int iy = p / Dx, ix = p - iy * Dx;
      double x = pix_x(ix0, ix, s), y = pix_x(iy0, iy, s);
      double a1, a2, b1, b2, c1, c2, d1, d2;
      eps12(d, f, x + 0.5, y, a1, a2);
      eps12(d, f, x - 0.5, y, b1, b2);
      eps12(d, f, x, y + 0.5, c1, c2);
      eps12(d, f, x, y - 0.5, d1, d2);
      Jd[p] = fabs((a1 - b1) * (c2 - d2) - (a2 - b2) * (c1 - d1));
    }
  }
  wave_sync();
}

// The sigma_D-dependent half: seff from the area of the centre pixel, then u = eps / seff in place; returns seff.
__device__ double box_u(int Dy, int Dx, double sd, double *U1, double *U2, const double *Jd, int lane) {
  const int P = Dy * Dx;
  double psz = sqrt(Jd[(Dy / 2) * Dx + Dx / 2]);
  double t = psz / sd;
  double seff = sd * sqrt(1.0 + (t * t) / 12.0);
  for (int p = lane; p < P; p += WAVE) {
    U1[p] = U1[p] / seff;
    U2[p] = U2[p] / seff;
  }
  wave_sync();
  return seff;
}

// u = eps / seff in place for a given seff (pass 1's effective width of the row); returns seff
__device__ __forceinline__ double box_u_given(int P, double seff, double *U1, double *U2, int lane) {
  for (int p = lane; p < P; p += WAVE) {
    U1[p] = U1[p] / seff;
    U2[p] = U2[p] / seff;
  }
  wave_sync();
  return seff;
}

// u1 = eps1/seff, u2 = eps2/seff into U1, U2 and J into Jd; returns seff
__device__ __forceinline__ double box_geometry(const xpg_detector &d, const Frame &f, int ix0, int iy0, int s, int Dy,
                                               int Dx, double sd, double *U1, double *U2, double *Jd, int lane) {
  box_eps_J(d, f, ix0, iy0, s, Dy, Dx, U1, U2, Jd, lane);
  return box_u(Dy, Dx, sd, U1, U2, Jd, lane);
}


// ---- bilinear profile grid coordinates (xpg/integrate_kernels.py:154-166) ---------------------------------------

struct Bil {
  bool inside;
  int i1, i2;
  double t1, t2;
};
__device__ __forceinline__ Bil bilinear(double u1, double u2) {
  Bil b;
  double f1 = (u1 + UMAX) / (2 * UMAX) * (double)(G - 1);
  double f2 = (u2 + UMAX) / (2 * UMAX) * (double)(G - 1);
  b.inside = (f1 >= 0) && (f1 <= G - 1) && (f2 >= 0) && (f2 <= G - 1);
  f1 = fmin(fmax(f1, 0.0), (double)(G - 1) - 1e-4);
  f2 = fmin(fmax(f2, 0.0), (double)(G - 1) - 1e-4);
  b.i1 = (int)floor(f1);
  b.i2 = (int)floor(f2);
  b.t1 = f1 - b.i1;
  b.t2 = f2 - b.i2;
  return b;
}

// ---- neighbour ellipsoids (xpg/integrate_kernels.py:186-213) ----------------------------------------------------
// For each neighbour column that reaches the box (separable nearest-point test, exact), ORs bit `bit` of vflag for the
// voxels with ((dx^2 + dy^2) + dz^2) < 1.  tab: scratch of Dx + Dy + Dz doubles.
__device__ void neighbour_mask(const xpg_refl &R, int r, const double *xs, const double *ys, const double *zs,
                               int Dz, int Dy, int Dx, int skip_first, uint8_t *vflag, uint8_t bit, double *tab,
                               const uint8_t *pix_row, int lane) {
  const int K = R.K;
  if (K <= 0) return;
  double lo_x = xs[0], hi_x = xs[Dx - 1], lo_y = ys[0], hi_y = ys[Dy - 1], lo_z = zs[0], hi_z = zs[Dz - 1];
  bool hit = false;
  if (lane < K && !(lane == 0 && skip_first)) {
    double nx = (double)R.nbx[(int64_t)r * K + lane], ny = (double)R.nby[(int64_t)r * K + lane];
    double nz = (double)R.nbz[(int64_t)r * K + lane];
    double sx = (double)R.nbsx[(int64_t)r * K + lane], sz = (double)R.nbsz[(int64_t)r * K + lane];
    double mx = (fmin(fmax(nx, lo_x), hi_x) - nx) / sx;
    double my = (fmin(fmax(ny, lo_y), hi_y) - ny) / sx;
    double mz = (fmin(fmax(nz, lo_z), hi_z) - nz) / sz;
    hit = ((mx * mx + my * my) + mz * mz) < 1.0;
  }
  uint64_t hits = __ballot(hit);
  const int P = Dy * Dx;
  double *dx2 = tab, *dy2 = tab + Dx, *dz2 = tab + Dx + Dy;
  while (hits) {
    int k = __ffsll((unsigned long long)hits) - 1;
    hits &= hits - 1;
    double nx = (double)R.nbx[(int64_t)r * K + k], ny = (double)R.nby[(int64_t)r * K + k];
    double nz = (double)R.nbz[(int64_t)r * K + k];
    double sx = (double)R.nbsx[(int64_t)r * K + k], sz = (double)R.nbsz[(int64_t)r * K + k];
    for (int i = lane; i < Dx + Dy + Dz; i += WAVE) {
      double v;
      if (i < Dx) {
        v = (xs[i] - nx) / sx;
      } else if (i < Dx + Dy) {
        v = (ys[i - Dx] - ny) / sx;
      } else {
        v = (zs[i - Dx - Dy] - nz) / sz;
      }
      tab[i] = v * v;
    }
    wave_sync();
    for (int p0 = 0; p0 < P; p0 += WAVE) {       // lanes over pixels, frames in order
      const int p = p0 + lane;
      if (p < P) {
        const int iy = pix_row[p], ix = p - iy * Dx;
        const double dxy = dx2[ix] + dy2[iy];
        for (int j = 0; j < Dz; ++j)
          if (dxy + dz2[j] < 1.0) vflag[j * P + p] |= bit;
      }
    }
    wave_sync();
  }
}

}  // namespace xpg
