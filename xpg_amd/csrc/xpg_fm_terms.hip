// Separable factors of the forward model's fractions (xpg_fm_terms, include/xpg_forward.h): fraction_terms of
// xpg/forward/model.py:206-275 on the device, one wavefront per reflection, with the integration kernels' float64
// geometry (Kabsch frame, eps at the pixel centres, J by second-order differences, the table-erf rocking curve).
// The factors feed xpg_fm_scatter / xpg_fm_gather directly: R (T, n, Dz) and P (T, n, b, b) float32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_common.h"
#include "xpg_forward.h"
#include "xpg_integrate.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

using namespace xpg;

// LDS of one wavefront: eps1, eps2, J (b b doubles each), the rocking-curve scratch (21 Dz) and R (Dz), R_h (Dz)
__host__ __device__ inline size_t terms_wave_bytes(int b, int Dz) {
  return 8 * (3 * (size_t)b * b + (size_t)(RCK + 2) * Dz);
}

__device__ __forceinline__ double gauss_pixel(double r2, double se, double J) {
  return ((exp(-0.5 * r2 / (se * se)) / (2.0 * 3.141592653589793)) * J) / (se * se);
}

__global__ void __launch_bounds__(256) k_fm_terms(xpg_detector det, xpg_fm_model m, xpg_fm_rows rows, float *Rout,
                                                  float *Pout, int waves, int wave_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int r = blockIdx.x * waves + wid;
  if (r >= rows.n) return;
  const int b = rows.b, Dz = rows.Dz, P = b * b, s = rows.pxs;
  double *U1 = (double *)(smem + (size_t)wid * wave_bytes);
  double *U2 = U1 + P, *Jd = U2 + P, *scr = Jd + P, *Rd = scr + RCK * Dz, *Rh = Rd + Dz;
  const int ix0 = __builtin_amdgcn_readfirstlane(rows.ix0[r]), iy0 = __builtin_amdgcn_readfirstlane(rows.iy0[r]);
  const int iz0 = __builtin_amdgcn_readfirstlane(rows.iz0[r]);
  const double sd = rows.sd ? rows.sd[r] : m.sd, sm = rows.sm ? rows.sm[r] : m.sm;
  Frame f;
  kabsch(rows.S + 3 * (int64_t)r, m.s0, m.axis, f);
  // eps and J at the (super-)pixel centres; J of boxes narrower than 3 is the constant (qx pxs)^2
  box_eps_J(det, f, ix0, iy0, s, b, b, U1, U2, Jd, lane);
  if (b < 3) {
    for (int p = lane; p < P; p += WAVE) Jd[p] = (det.qx * (double)s) * (det.qx * (double)s);
    wave_sync();
  }
  // rocking curves over frames k = iz0 + j, [phi0 + osc k, phi0 + osc k + osc]: the mixture (or a Gaussian), and the
  // halo's Gaussian of kappa sigma_M
  xpg_call c{};
  for (int i = 0; i < 3; ++i) {
    c.axis[i] = m.axis[i];
    c.s0[i] = m.s0[i];
  }
  c.frames_phi0 = m.frames_phi0;
  c.osc = m.osc;
  c.sd = sd;
  c.sm = sm;
  c.use_rcw = m.use_rcw;
  for (int j = 0; j < RCK; ++j) c.rcw[j] = m.rcw[j];
  c.widths = nullptr;
  const double phi = rows.phi[r];
  rocking_curve(c, iz0, 1, phi, f.zeta, Dz, scr, Rd, lane);
  if (m.halo) {
    xpg_call ch = c;
    ch.use_rcw = 0;
    ch.sm = m.halo_kappa * sm;
    rocking_curve(ch, iz0, 1, phi, f.zeta, Dz, scr, Rh, lane);
  }
  // seff from sigma_D and the centre pixel's area; the halo's pixel-convolved width
  const double pix = sqrt(Jd[(b / 2) * b + b / 2]);
  const double t = pix / sd;
  const double seff = sd * sqrt(1.0 + (t * t) / 12.0);
  const double ks = m.halo_kappa * sd;
  const double sh = sqrt(ks * ks + (pix * pix) / 12.0);
  const int64_t n = rows.n;
  for (int j = lane; j < Dz; j += WAVE) {
    if (m.halo) {
      Rout[(int64_t)r * Dz + j] = (float)((1.0 - m.halo_w) * Rd[j]);
      Rout[(n + r) * Dz + j] = (float)(m.halo_w * Rh[j]);
    } else {
      Rout[(int64_t)r * Dz + j] = (float)Rd[j];
    }
  }
  const int region = m.prof && rows.region ? min(max(rows.region[r], 0), m.n_regions - 1) : 0;
  const float *pr = m.prof ? m.prof + (int64_t)region * m.G * m.G : nullptr;
  const double se2 = seff * seff, G1 = (double)(m.G - 1);
  for (int p = lane; p < P; p += WAVE) {
    const double e1 = U1[p], e2 = U2[p], J = Jd[p];
    double v;
    if (pr) {
      // the learnt profile at u = eps / seff, bilinear on the grid of half-width umax
      double f1 = ((e1 / seff + m.umax) / (2.0 * m.umax)) * G1;
      double f2 = ((e2 / seff + m.umax) / (2.0 * m.umax)) * G1;
      const bool inside = f1 >= 0.0 && f1 <= G1 && f2 >= 0.0 && f2 <= G1;
      f1 = fmin(fmax(f1, 0.0), G1 - 1e-4);
      f2 = fmin(fmax(f2, 0.0), G1 - 1e-4);
      const int i1 = (int)floor(f1), i2 = (int)floor(f2);
      const double t1 = f1 - i1, t2 = f2 - i2;
      const double a = pr[i2 * m.G + i1], bq = pr[i2 * m.G + i1 + 1];
      const double cq = pr[(i2 + 1) * m.G + i1], d = pr[(i2 + 1) * m.G + i1 + 1];
      const double val = ((((1 - t1) * (1 - t2)) * a + (t1 * (1 - t2)) * bq) + ((1 - t1) * t2) * cq) + (t1 * t2) * d;
      v = ((inside ? val : 0.0) * J) / se2;
    } else {
      v = gauss_pixel(e1 * e1 + e2 * e2, seff, J);
    }
    Pout[(int64_t)r * P + p] = (float)v;
    if (m.halo) Pout[(n + r) * P + p] = (float)gauss_pixel(e1 * e1 + e2 * e2, sh, J);
  }
}

}  // namespace

extern "C" int xpg_fm_terms(const xpg_detector *det, const xpg_fm_model *m, const xpg_fm_rows *rows, float *R, float *P,
                            void *stream) {
  if (!det || !m || !rows || !R || !P) return xpg::report_error(XPG_ERR_ARG, "fm_terms: NULL argument");
  if (rows->n < 0 || rows->Dz < 1 || rows->Dz > XPG_MAX_DZ || rows->b < 1 || rows->pxs < 1)
    return xpg::report_error(XPG_ERR_ARG, "fm_terms: bad shape (n, Dz, b, pxs)");
  if (rows->n == 0) return XPG_OK;
  if (!rows->S || !rows->phi || !rows->ix0 || !rows->iy0 || !rows->iz0)
    return xpg::report_error(XPG_ERR_ARG, "fm_terms: missing per-reflection array");
  if (det->n_seg < 1 || !det->P0 || !det->Bx || !det->By || !det->U1 || !det->U2 || !det->UN)
    return xpg::report_error(XPG_ERR_ARG, "fm_terms: bad detector");
  if (m->prof && (m->n_regions < 1 || m->G < 2 || !rows->region || !(m->umax > 0.0)))
    return xpg::report_error(XPG_ERR_ARG, "fm_terms: profiles need n_regions, G >= 2, region and umax > 0");
  const size_t wb = terms_wave_bytes(rows->b, rows->Dz);
  int lim = 65536, dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    lim = 65536;
  if (wb > (size_t)lim) return xpg::report_error(XPG_ERR_TOO_LARGE, "fm_terms: support too large for the LDS");
  int waves = (int)(65536 / wb);
  waves = waves < 1 ? 1 : (waves > 4 ? 4 : waves);
  const size_t lds = (size_t)waves * wb;
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void *)k_fm_terms, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
    if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_fm_terms, dim3((rows->n + waves - 1) / waves), dim3(64 * waves), lds, (hipStream_t)stream,
                     *det, *m, *rows, R, P, waves, (int)wb);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return XPG_OK;
}
