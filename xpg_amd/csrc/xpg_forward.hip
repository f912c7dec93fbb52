// Forward (window) model passes for gfx950: scatter of modelled counts, per-reflection gathers (include/xpg_forward.h).
//
// A reflection's fraction is separable, f(z, y, x) = sum_t R_t[z] P_t[y, x], so nothing per voxel is stored: each lane
// forms f in registers from the row's factors.
//   scatter: one wavefront per reflection, its Dz b b voxels over the lanes, one float atomic per in-window voxel
//            (supports of neighbouring reflections overlap, as in the reference).
//   gather:  one wavefront per reflection, over its voxels, the sums in registers in a fixed order, one wave
//            reduction and one add per row (no per-frame atomics: the result does not depend on scheduling).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xpg_forward.h"
#include "xpg_integrate.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

constexpr int WAVE = 64;

// a support voxel's window index, or -1 outside the window
__device__ __forceinline__ int64_t win_index(const xpg_fm_window &w, int kz, int ky, int kx) {
  if (kz < 0 || kz >= w.nf || ky < 0 || ky >= w.ny || kx < 0 || kx >= w.nx) return -1;
  return ((int64_t)kz * w.ny + ky) * w.nx + kx;
}

// f at voxel (z, p = y b + x) of row r: the T terms in order
__device__ __forceinline__ float frac(const xpg_fm_supports &s, int r, int z, int p) {
  const int64_t bb = (int64_t)s.b * s.b;
  float f = 0.f;
  for (int t = 0; t < s.T; ++t) {
    const int64_t tr = (int64_t)t * s.n + r;
    f += s.R[tr * s.Dz + z] * s.P[tr * bb + p];
  }
  return f;
}

// A lane's walk over the Dz b b voxels of a support, item i = (z b + y) b + x, advancing 64 items a step: one
// integer division at the start, then carries (exact for every b and Dz).
struct Walk {
  int z, y, x;            // current item
  int sz, sy, sx;         // 64 items in (z, y, x) digits
  int b;
  __device__ Walk(int lane, int b_) : b(b_) {
    const int bb = b * b;
    z = lane / bb;
    y = (lane - z * bb) / b;
    x = lane - z * bb - y * b;
    sz = WAVE / bb;
    sy = (WAVE - sz * bb) / b;
    sx = WAVE - sz * bb - sy * b;
  }
  __device__ __forceinline__ void next() {
    x += sx;
    y += sy;
    z += sz;
    if (x >= b) { x -= b; ++y; }
    if (y >= b) { y -= b; ++z; }
  }
};

// One wavefront per reflection over all Dz b b voxels of its support (every lane busy; the reference's program per
// (reflection, frame) leaves the lanes beyond b b idle).
__global__ void __launch_bounds__(256) k_fm_scatter(float *mu, xpg_fm_supports s, const float *I, const float *beta,
                                                    xpg_fm_window w) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int r = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (r >= s.n) return;
  const int kz0 = s.iz0[r] - w.f0, ky0 = s.iy0[r] - w.y0, kx0 = s.ix0[r] - w.x0;
  const float Ir = I[r], br = beta ? beta[r] : 0.f;
  const int nvox = s.Dz * s.b * s.b;
  Walk k(lane, s.b);
  for (int i = lane; i < nvox; i += WAVE, k.next()) {
    const int64_t v = win_index(w, kz0 + k.z, ky0 + k.y, kx0 + k.x);
    if (v < 0) continue;
    const float f = frac(s, r, k.z, k.y * s.b + k.x);
    atomicAdd(mu + v, beta ? f * Ir + br : f * Ir);
  }
}

template <int NS_T>
__device__ __forceinline__ float ns_at(const void *NS, int64_t v) {
  if constexpr (NS_T == XPG_FM_NS_I16) return (float)((const int16_t *)NS)[v];
  else if constexpr (NS_T == XPG_FM_NS_I32) return (float)((const int32_t *)NS)[v];
  else return ((const float *)NS)[v];
}

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
  return v;
}

template <int NS_T, int MODE>
__global__ void __launch_bounds__(256) k_fm_gather(float *g, float *fd, xpg_fm_supports s, const int64_t *rows,
                                                   int64_t n_out, const float *A, const float *M, const void *NS,
                                                   xpg_fm_window w) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int r = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
  if (r >= s.n) return;
  const int64_t out = rows[r];
  if (out < 0 || out >= n_out) return;
  const int kz0 = s.iz0[r] - w.f0, ky0 = s.iy0[r] - w.y0, kx0 = s.ix0[r] - w.x0;
  const int nvox = s.Dz * s.b * s.b;
  float acc = 0.f, accd = 0.f;
  Walk k(lane, s.b);
  for (int i = lane; i < nvox; i += WAVE, k.next()) {
    const int64_t v = win_index(w, kz0 + k.z, ky0 + k.y, kx0 + k.x);
    if (v < 0) continue;
    const float ns = ns_at<NS_T>(NS, v);
    if (!(ns >= 0.f)) continue;
    const float f = frac(s, r, k.z, k.y * s.b + k.x);
    const float mv = M[v];
    if constexpr (MODE == 0) {
      acc += f * (A[v] / mv);
    } else {
      acc += f * (1.0f - fmaxf(ns, 0.f) / mv);
      accd += f * f / fmaxf(mv, 1e-6f);
    }
  }
  acc = wave_sum(acc);
  if (MODE == 1) accd = wave_sum(accd);
  if (lane == 0) {
    atomicAdd(g + out, acc);
    if (MODE == 1) atomicAdd(fd + out, accd);
  }
}

int check(const xpg_fm_supports *s, const xpg_fm_window *w) {
  if (!s || !w || !s->R || !s->P || !s->iz0 || !s->iy0 || !s->ix0)
    return xpg::report_error(XPG_ERR_ARG, "forward model: missing supports or window");
  if (s->T < 1 || s->T > 2 || s->n < 0 || s->Dz < 1 || s->b < 1 || (int64_t)s->Dz * s->b * s->b > (1 << 30))
    return xpg::report_error(XPG_ERR_ARG, "forward model: T in 1..2, Dz >= 1, b >= 1, Dz b b <= 2^30 expected");
  if (w->nf < 0 || w->ny < 0 || w->nx < 0) return xpg::report_error(XPG_ERR_ARG, "forward model: negative window");
  return XPG_OK;
}

int launched() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? XPG_OK : xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
}

}  // namespace

extern "C" {

int xpg_fm_scatter(float *mu, const xpg_fm_supports *sup, const float *I, const float *beta, const xpg_fm_window *win,
                   void *stream) {
  int rc = check(sup, win);
  if (rc) return rc;
  if (!mu || !I) return xpg::report_error(XPG_ERR_ARG, "fm_scatter: mu or I is NULL");
  if (sup->n == 0) return XPG_OK;
  hipLaunchKernelGGL(k_fm_scatter, dim3((uint32_t)((sup->n + 3) / 4)), dim3(256), 0, (hipStream_t)stream, mu, *sup, I,
                     beta, *win);
  return launched();
}

int xpg_fm_gather(float *g, float *fd, const xpg_fm_supports *sup, const int64_t *rows, int64_t n_out, const float *A,
                  const float *M, const void *NS, int32_t ns_type, int32_t mode, const xpg_fm_window *win,
                  void *stream) {
  int rc = check(sup, win);
  if (rc) return rc;
  if (!g || !rows || !M || !NS) return xpg::report_error(XPG_ERR_ARG, "fm_gather: g, rows, M or NS is NULL");
  if (mode == 0 && !A) return xpg::report_error(XPG_ERR_ARG, "fm_gather: mode 0 needs A");
  if (mode == 1 && !fd) return xpg::report_error(XPG_ERR_ARG, "fm_gather: mode 1 needs fd");
  if (mode != 0 && mode != 1) return xpg::report_error(XPG_ERR_ARG, "fm_gather: mode 0 or 1");
  if (ns_type < XPG_FM_NS_I16 || ns_type > XPG_FM_NS_F32) return xpg::report_error(XPG_ERR_ARG, "fm_gather: NS type");
  if (sup->n == 0) return XPG_OK;
  const dim3 grid((uint32_t)((sup->n + 3) / 4)), blk(256);
  hipStream_t st = (hipStream_t)stream;
#define XPG_FM_LAUNCH(NT, MD) hipLaunchKernelGGL((k_fm_gather<NT, MD>), grid, blk, 0, st, g, fd, *sup, rows, n_out, A, M, NS, \
                                                 *win)
  if (mode == 0) {
    if (ns_type == XPG_FM_NS_I16) XPG_FM_LAUNCH(XPG_FM_NS_I16, 0);
    else if (ns_type == XPG_FM_NS_I32) XPG_FM_LAUNCH(XPG_FM_NS_I32, 0);
    else XPG_FM_LAUNCH(XPG_FM_NS_F32, 0);
  } else {
    if (ns_type == XPG_FM_NS_I16) XPG_FM_LAUNCH(XPG_FM_NS_I16, 1);
    else if (ns_type == XPG_FM_NS_I32) XPG_FM_LAUNCH(XPG_FM_NS_I32, 1);
    else XPG_FM_LAUNCH(XPG_FM_NS_F32, 1);
  }
#undef XPG_FM_LAUNCH
  return launched();
}

}  // extern "C"
