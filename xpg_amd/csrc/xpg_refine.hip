// Residuals and normal equations of the geometry refinement (include/xpg_refine.h), the device part of
// xpg/refine.py:151-280.  One thread per (model, spot) predicts the spot's position and spindle angle under that
// model; the normal equations are reduced per (job, matrix entry) by one workgroup in a fixed order.
//
// The reference evaluates a parameter vector and its p forward-difference perturbations in one batched torch call;
// here every (parameter vector, spot) pair is a thread of one launch, and several refinement problems (phi blocks)
// share a launch.  Float64 throughout, built without contraction: the sums of three products are added left to
// right, as numpy and torch add them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_common.h"
#include "xpg_refine.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

using xpg::cross3;
using xpg::dot3;

constexpr double TWO_PI = 6.283185307179586;

// the job of thread t: the last j with jobs[j].t0 <= t
__device__ __forceinline__ int job_of(const xpg_rf_job *jobs, int n_jobs, int64_t t) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].t0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// predict_xyphi_t (xpg/refine.py:253-280) and the residual rows of residuals (xpg/refine.py:154-183) for one spot
__global__ void __launch_bounds__(256) k_rf_residuals(xpg_rf_spots sp, xpg_rf_setup su, const xpg_rf_job *jobs,
                                                      int n_jobs, int64_t total, const double *models, int stride,
                                                      double *r) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= total) return;
  const xpg_rf_job jb = jobs[job_of(jobs, n_jobs, t)];
  const int64_t local = t - jb.t0, nj = jb.n_spots;
  if (nj <= 0) return;                         // (a job without spots owns no threads)
  const int m = (int)(local / nj);
  const int64_t il = local - (int64_t)m * nj, i = jb.spot0 + il;
  const double *md = models + (jb.model0 + m) * stride;
  const double *s0 = md, *ax = md + 3, *A = md + 6;
  const double h[3] = {sp.h[3 * i], sp.h[3 * i + 1], sp.h[3 * i + 2]};
  double p0[3];
  for (int k = 0; k < 3; ++k) p0[k] = (A[3 * k] * h[0] + A[3 * k + 1] * h[1]) + A[3 * k + 2] * h[2];
  const double pm = dot3(p0, ax);
  double par[3], perp[3], mxp[3];
  for (int k = 0; k < 3; ++k) {
    par[k] = pm * ax[k];
    perp[k] = p0[k] - par[k];
  }
  cross3(ax, p0, mxp);
  const double a = dot3(perp, s0), b = dot3(mxp, s0);
  const double c = -0.5 * dot3(p0, p0) - dot3(par, s0);
  const double rr = sqrt(a * a + b * b);
  const double ratio = fmin(fmax(c / rr, -1.0 + 1e-9), 1.0 - 1e-9);
  const double phio = sp.phi[i];
  double phi = atan2(b, a) + sp.branch[i] * acos(ratio);
  phi = phi + TWO_PI * rint((phio - phi) / TWO_PI);
  const double cph = cos(phi), sph = sin(phi);
  double S[3];
  for (int k = 0; k < 3; ++k) S[k] = ((par[k] + cph * perp[k]) + sph * mxp[k]) + s0[k];
  // the plane of the segment the spot was observed on
  const int sg = min(max(sp.seg[i], 0), su.n_seg - 1);
  const double *P0 = md + XPG_RF_HEAD + XPG_RF_SEG * sg, *Bx = P0 + 3, *By = P0 + 6;
  double nrm[3];
  cross3(Bx, By, nrm);
  const double tt = dot3(P0, nrm) / dot3(S, nrm);
  double P[3];
  for (int k = 0; k < 3; ++k) P[k] = tt * S[k] - P0[k];
  double xp = md[15] + dot3(P, Bx) / dot3(Bx, Bx);
  double yp = md[16] + dot3(P, By) / dot3(By, By);
  // thick-sensor parallax: the absorption-weighted mean depth along the ray, projected on the segment's axes
  const double sn = sqrt(dot3(S, S));
  double e[3], u1[3], u2[3], un[3];
  for (int k = 0; k < 3; ++k) {
    e[k] = S[k] / sn;
    u1[k] = Bx[k] / su.qx;
    u2[k] = By[k] / su.qy;
  }
  cross3(u1, u2, un);
  const double cosa = fmin(fmax(dot3(e, un), 1e-3), 1.0);
  const double L = su.tsens / cosa;
  const double lbar = 1.0 / su.mu - (L * exp(-su.mu * L)) / fmax(-expm1(-su.mu * L), 1e-12);
  xp = xp + ((su.plx * lbar) * dot3(e, u1)) / su.qx;
  yp = yp + ((su.plx * lbar) * dot3(e, u2)) / su.qy;
  double *o = r + jb.r0 + (int64_t)m * 3 * nj;
  o[il] = (xp - sp.xo[i]) / su.sxy;
  o[nj + il] = (yp - sp.yo[i]) / su.sxy;
  o[2 * nj + il] = (phi - phio) / su.sphi;
}

constexpr int NT = 256;

// One workgroup per (job, entry): entries 0 .. np - 1 are the pairs k <= l of J^T W J, then the p entries of J^T W r_0.
// Each thread sums its spots in order, then a fixed tree over the workgroup.
__device__ __forceinline__ double spot_weight(const double *wj, const double *r0, int64_t nj, int64_t q, double c) {
  const int64_t i = q % nj;
  const double wi = wj[i];
  if (c <= 0.0) return wi;
  const double x = r0[i], y = r0[nj + i], z = r0[2 * nj + i];
  const double t = sqrt((x * x + y * y) + z * z) / c;
  return wi / (1.0 + t * t);
}

__global__ void __launch_bounds__(NT) k_rf_normal(xpg_rf_spots sp, xpg_rf_setup su, const xpg_rf_job *jobs,
                                                  const double *r, const double *hd, const double *w, double *out,
                                                  double *r0_out) {
  __shared__ double red[NT];
  const int p = su.n_models - 1, j = blockIdx.y;
  const int np = p * (p + 1) / 2;
  int e = blockIdx.x, k = 0, l = -1;
  if (e < np) {
    while (e >= p - k) {
      e -= p - k;
      ++k;
    }
    l = k + e;
  } else {
    k = e - np;
  }
  const xpg_rf_job jb = jobs[j];
  const int64_t nj = jb.n_spots;
  const double hk = hd[(int64_t)j * p + k], hl = l >= 0 ? hd[(int64_t)j * p + l] : 1.0;
  const double *r0 = r + jb.r0, *rk = r0 + (int64_t)(1 + k) * 3 * nj;
  const double *rl = l >= 0 ? r0 + (int64_t)(1 + l) * 3 * nj : r0;
  // the job's 3 n rows (row o: spot o mod n), four rows per thread in flight before any is used
  const int64_t rows = 3 * nj;
  const double *wj = w + jb.w0;
  double acc = 0.0;
  int64_t o = threadIdx.x;
  for (; o + 3 * NT < rows; o += 4 * NT) {
    double vk[4], vl[4], v0[4], vw[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t q = o + u * NT;
      vk[u] = rk[q];
      vl[u] = rl[q];
      v0[u] = r0[q];
      vw[u] = spot_weight(wj, r0, nj, q, su.robust_c);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double jk = (vk[u] - v0[u]) / hk;
      const double other = l >= 0 ? (vl[u] - v0[u]) / hl : v0[u];
      acc += (jk * vw[u]) * other;
    }
  }
  for (; o < rows; o += NT) {
    const double jk = (rk[o] - r0[o]) / hk;
    const double other = l >= 0 ? (rl[o] - r0[o]) / hl : r0[o];
    acc += (jk * spot_weight(wj, r0, nj, o, su.robust_c)) * other;
  }
  if (r0_out && l < 0 && k == 0)                 // one workgroup per job returns the model-0 rows
    for (int64_t q = threadIdx.x; q < rows; q += NT) r0_out[3 * jb.w0 + q] = r0[q];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double *oq = out + (int64_t)j * (p * p + p);
    if (l >= 0) {
      oq[k * p + l] = red[0];
      oq[l * p + k] = red[0];
    } else {
      oq[p * p + k] = red[0];
    }
  }
}

int check_common(const xpg_rf_spots *sp, const xpg_rf_setup *su, const xpg_rf_job *jobs, int n_jobs) {
  if (!sp || !su) return xpg::report_error(XPG_ERR_ARG, "refine: NULL argument");
  if (sp->n < 0 || su->n_seg < 1 || n_jobs < 0) return xpg::report_error(XPG_ERR_ARG, "refine: bad sizes");
  if (n_jobs > 0 && !jobs) return xpg::report_error(XPG_ERR_ARG, "refine: NULL job table");
  if (sp->n > 0 && (!sp->h || !sp->phi || !sp->branch || !sp->xo || !sp->yo || !sp->seg))
    return xpg::report_error(XPG_ERR_ARG, "refine: missing spot array");
  if (!(su->qx > 0.0) || !(su->qy > 0.0) || !(su->sxy > 0.0) || !(su->sphi > 0.0) || !(su->mu > 0.0))
    return xpg::report_error(XPG_ERR_ARG, "refine: qx, qy, sigma_xy, sigma_phi and mu must be positive");
  return XPG_OK;
}

int launched() {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return XPG_OK;
}

}  // namespace

// The job table lives in device memory, so its ranges are the caller's contract (the Python host checks them before
// it uploads the table): spot0 + n_spots <= n, model0 + n_models within models, r0 + 3 n_models n_spots within r,
// w0 + n_spots within w.
extern "C" int xpg_rf_residuals(const xpg_rf_spots *sp, const xpg_rf_setup *su, const xpg_rf_job *jobs, int32_t n_jobs,
                                int64_t total, const double *models, double *r, void *stream) {
  const int rc = check_common(sp, su, jobs, n_jobs);
  if (rc != XPG_OK) return rc;
  if (total < 0) return xpg::report_error(XPG_ERR_ARG, "rf_residuals: negative total");
  if (total == 0 || n_jobs == 0) return XPG_OK;
  if (!models || !r) return xpg::report_error(XPG_ERR_ARG, "rf_residuals: NULL models or output");
  if (total > ((int64_t)1 << 40)) return xpg::report_error(XPG_ERR_ARG, "rf_residuals: too many evaluations");
  const int stride = XPG_RF_HEAD + XPG_RF_SEG * su->n_seg;
  hipLaunchKernelGGL(k_rf_residuals, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *sp,
                     *su, jobs, n_jobs, total, models, stride, r);
  return launched();
}

extern "C" int xpg_rf_normal(const xpg_rf_spots *sp, const xpg_rf_setup *su, const xpg_rf_job *jobs, int32_t n_jobs,
                             const double *r, const double *h, const double *w, double *out, double *r0_out,
                             void *stream) {
  const int rc = check_common(sp, su, jobs, n_jobs);
  if (rc != XPG_OK) return rc;
  if (su->n_models < 2) return xpg::report_error(XPG_ERR_ARG, "rf_normal: needs at least one parameter");
  if (su->n_models > 257) return xpg::report_error(XPG_ERR_ARG, "rf_normal: more than 256 parameters");
  if (n_jobs > 65535) return xpg::report_error(XPG_ERR_ARG, "rf_normal: more than 65535 jobs");
  if (n_jobs == 0) return XPG_OK;
  if (!r || !h || !w || !out) return xpg::report_error(XPG_ERR_ARG, "rf_normal: NULL argument");
  const int p = su->n_models - 1;
  const int entries = p * (p + 1) / 2 + p;
  hipLaunchKernelGGL(k_rf_normal, dim3(entries, n_jobs), dim3(NT), 0, (hipStream_t)stream, *sp, *su, jobs, r, h, w,
                     out, r0_out);
  return launched();
}
