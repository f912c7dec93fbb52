// Per-reflection finalisation after integration (include/xpg_finalise.h): the Lorentz-polarisation and absorption
// corrections and the derived columns of xpg/stage_integrate.py:408-421, one lane per reflection.
//
// The reference computes these on the host after the whole sweep has come back; here they run on the device behind
// each block's pass 2, so the host receives final intensities.  Float64 throughout, in the reference's operation
// order where it has one (numpy's dot products are sums of three products; the device adds them left to right).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <cmath>

#include "xpg_common.h"
#include "xpg_finalise.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

using xpg::dot3;

__device__ __forceinline__ double clip01(double v, double lo) { return fmin(fmax(v, lo), 1.0); }

// Lab position (mm) where the ray recorded at raw pixel (x, y) enters the sensor, as Geometry.pixel_to_lab computes it
// (xpg/geometry.py:235-262): segment k (gaps: 0), geometric correction, the segment's affine map, then four
// parallax iterations when the detector has a sensor model.  (The integration kernels' pixel_S takes two.)
__device__ void lab_point4(const xpg_detector &d, double x, double y, int k, double *X) {
  const double *P0 = d.P0 + 3 * k, *Bx = d.Bx + 3 * k, *By = d.By + 3 * k;
  double xc = x, yc = y;
  if (d.gx) {
    xc = x + xpg::bilinear_table(d.gx, d.geo_nx, d.geo_ny, x, y);
    yc = y + xpg::bilinear_table(d.gy, d.geo_nx, d.geo_ny, x, y);
  }
  double a = xc - d.orgx, b = yc - d.orgy;
  for (int i = 0; i < 3; ++i) X[i] = (P0[i] + a * Bx[i]) + b * By[i];
  if (d.plx == 0.0) return;
  const double *U1 = d.U1 + 3 * k, *U2 = d.U2 + 3 * k, *UN = d.UN + 3 * k;
  for (int it = 0; it < 4; ++it) {
    const double nrm = sqrt(dot3(X, X));
    const double e[3] = {X[0] / nrm, X[1] / nrm, X[2] / nrm};
    const double cosa = fmin(fmax(dot3(e, UN), 1e-3), 1.0);
    const double L = d.tsens / cosa;
    const double muL = d.mu * L;
    const double lbar = 1.0 / d.mu - (L * exp(-muL)) / fmax(-expm1(-muL), 1e-12);
    const double d1 = lbar * dot3(e, U1), d2 = lbar * dot3(e, U2);
    a = (xc - d1 / d.qx) - d.orgx;
    b = (yc - d2 / d.qy) - d.orgy;
    for (int i = 0; i < 3; ++i) X[i] = (P0[i] + a * Bx[i]) + b * By[i];
  }
}

__global__ void __launch_bounds__(256) k_finalise(xpg_lp_geom g, xpg_detector det, int64_t n, xpg_final_in in,
                                                  xpg_final_out out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double S[3] = {in.S[3 * i], in.S[3 * i + 1], in.S[3 * i + 2]};
  const double nS = sqrt(dot3(S, S));
  const double e[3] = {S[0] / nS, S[1] / nS, S[2] / nS};
  // Lorentz: zeta = (e x e0) . axis; polarisation P = f (1 - (e.emain)^2) + (1 - f) (1 - (e.n)^2)
  double c[3];
  xpg::cross3(e, g.e0, c);
  const double zeta = dot3(c, g.axis);
  const double em = dot3(e, g.pm), en = dot3(e, g.pn);
  const double P = g.pfrac * (1.0 - em * em) + (1.0 - g.pfrac) * (1.0 - en * en);
  // absorption: sensor efficiency 1 - exp(-mu t / cos) and air transmission exp(-mu_air path)
  double cosn, path;
  if (g.segmented) {
    double X[3];
    const double x = isnan(in.x[i]) ? 1.0 : in.x[i], y = isnan(in.y[i]) ? 1.0 : in.y[i];
    int k = 0;
    if (det.segmap) {
      long ix = (long)floor(x - 0.5), iy = (long)floor(y - 0.5);
      ix = ix < 0 ? 0 : (ix > det.seg_nx - 1 ? det.seg_nx - 1 : ix);
      iy = iy < 0 ? 0 : (iy > det.seg_ny - 1 ? det.seg_ny - 1 : iy);
      k = min(max((int)det.segmap[iy * det.seg_nx + ix], 0), det.n_seg - 1);
    }
    cosn = dot3(e, det.UN + 3 * k);
    lab_point4(det, x, y, k, X);
    path = sqrt(dot3(X, X));
  } else {
    cosn = dot3(e, g.ed3);
    path = g.distance / clip01(cosn, 1e-3);
  }
  double f = 1.0;
  if (g.tsens > 0 && g.mu > 0) f = f * -expm1(-g.mu * g.tsens / clip01(cosn, 1e-3));
  f = f * exp(-g.mu_air * path);
  const double rlp = fabs(zeta) / P / f;
  const double I = in.iprf[i] * rlp;
  const double sig = sqrt(fmax(in.vprf[i], 0.0)) * rlp;
  const double cov = in.coverage ? (double)in.coverage[i] : 1.0;
  const double frac = in.p_valid[i] / fmax(in.p_total[i], 1e-9);
  const double cr = in.corr[i];
  if (out.rlp) out.rlp[i] = rlp;
  if (out.I) out.I[i] = I;
  if (out.sig) out.sig[i] = sig;
  if (out.Isum) out.Isum[i] = in.isum[i] * rlp;
  if (out.sigsum) out.sigsum[i] = sqrt(fmax(in.vsum[i], 0.0)) * rlp;
  // np.clip keeps NaN (p_total of a reflection that was not integrated)
  if (out.peak) out.peak[i] = 100.0 * (isnan(frac) ? frac : fmin(fmax(frac, 0.0), 1.0)) * cov;
  if (out.corr) out.corr[i] = 100.0 * (isnan(cr) ? 0.0 : (isinf(cr) ? (cr > 0 ? 1.7976931348623157e308
                                                                              : -1.7976931348623157e308) : cr));
  if (out.good) out.good[i] = (isfinite(I) && isfinite(sig) && sig > 0.0) ? 1 : 0;
}

}  // namespace

extern "C" int xpg_finalise(const xpg_lp_geom *g, const xpg_detector *det, int64_t n, const xpg_final_in *in,
                            const xpg_final_out *out, void *stream) {
  if (!g || !in || !out) return xpg::report_error(XPG_ERR_ARG, "finalise: geometry / inputs / outputs is NULL");
  if (n < 0) return xpg::report_error(XPG_ERR_ARG, "finalise: n < 0");
  if (n == 0) return XPG_OK;
  if (!in->S || !in->iprf || !in->vprf || !in->isum || !in->vsum || !in->p_valid || !in->p_total || !in->corr)
    return xpg::report_error(XPG_ERR_ARG, "finalise: missing per-reflection input");
  xpg_detector d{};
  if (g->segmented) {
    if (!det || !in->x || !in->y) return xpg::report_error(XPG_ERR_ARG, "finalise: segmented without det / x / y");
    if (det->n_seg < 1 || !det->P0 || !det->Bx || !det->By || !det->U1 || !det->U2 || !det->UN)
      return xpg::report_error(XPG_ERR_ARG, "finalise: bad detector");
    d = *det;
  }
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return xpg::report_error(XPG_ERR_ARG, "finalise: too many reflections in one launch");
  hipLaunchKernelGGL(k_finalise, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, *g, d, n, *in, *out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return XPG_OK;
}

// ---- INTEGRATE.HKL data records (host) ------------------------------------------------------------------------------

namespace {

// a double as Python's format(v, spec) prints it for the specs used here ("f" / "E" with a precision): glibc's
// correctly rounded conversion, NaN without a sign as Python has it
int put_double(char *p, size_t cap, double v, int prec, char kind) {
  if (std::isnan(v)) return snprintf(p, cap, kind == 'E' ? "NAN" : "nan");
  return snprintf(p, cap, kind == 'E' ? "%.*E" : "%.*f", prec, v);
}

// int(numpy.nan_to_num(v)): NaN -> 0, +-inf -> +-DBL_MAX, then truncation towards zero, printed in full
int put_trunc(char *p, size_t cap, double v) {
  if (std::isnan(v)) v = 0.0;
  if (std::isinf(v)) v = v > 0 ? 1.7976931348623157e308 : -1.7976931348623157e308;
  const double t = std::trunc(v);
  if (t == 0.0) return snprintf(p, cap, "0");
  return snprintf(p, cap, "%.0f", t);
}

}  // namespace

extern "C" int64_t xpg_format_integrate_rows(char *buf, int64_t cap, int64_t n, const int32_t *hkl, const double *I,
                                             const double *sig, const double *x, const double *y, const double *zc,
                                             const double *rlp, const double *peak, const double *corr,
                                             const double *xo, const double *yo, const double *zo, double alf0,
                                             double bet0, const double *alf1, const double *bet1) {
  if (!buf || cap < 0 || n < 0) return -1;
  int64_t o = 0;
  char rec[19 * 330 + 16];               // one record: 19 fields, each at most 330 characters (310-digit numbers)
  for (int64_t i = 0; i < n; ++i) {
    char *p = rec;
    const size_t room = sizeof(rec);
    int k = snprintf(p, room, " %d %d %d ", hkl[3 * i], hkl[3 * i + 1], hkl[3 * i + 2]);
    auto f = [&](double v, int prec, char kind, const char *sep) {
      k += put_double(p + k, room - k, v, prec, kind);
      k += snprintf(p + k, room - k, "%s", sep);
    };
    f(I[i], 3, 'E', " ");
    f(sig[i], 3, 'E', " ");
    f(x[i], 1, 'f', " ");
    f(y[i], 1, 'f', " ");
    f(zc[i], 1, 'f', " ");
    f(rlp[i], 5, 'f', " ");
    f(peak[i], 5, 'f', " ");
    k += put_trunc(p + k, room - k, corr[i]);
    k += snprintf(p + k, room - k, " 0 ");
    f(xo[i], 1, 'f', " ");
    f(yo[i], 1, 'f', " ");
    f(zo[i], 1, 'f', " ");
    f(alf0, 2, 'f', " ");
    f(bet0, 2, 'f', " ");
    f(alf1[i], 2, 'f', " ");
    f(bet1[i], 2, 'f', " 0.00 1\n");
    if (o + k > cap) return -1;
    memcpy(buf + o, rec, (size_t)k);
    o += k;
  }
  return o;
}
