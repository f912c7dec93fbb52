/* C ABI of the per-reflection finalisation after integration, on the device (gfx950).
 *
 * Replaces the host post-processing of xpg/stage_integrate.py:408-421: the Lorentz-polarisation factor
 * (Geometry.lp_factor, xpg/geometry.py:394-407), the sensor-efficiency and air-absorption factor
 * (Geometry.absorption_factors, xpg/geometry.py:372-392), and the columns the caller derives from the integrator's
 * results: rlp, I, sig, Isum, sigsum, peak, corr (percent) and the finite-and-positive-sigma selection.
 * Same conventions as xpg_integrate.h: device pointers, asynchronous on the given stream, 0 or a negative
 * xpg_status (xpg_last_error() has the text).
 */
#ifndef XPG_FINALISE_H
#define XPG_FINALISE_H

#include <stdint.h>

#include "xpg_integrate.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Constants of one experiment geometry (host-computed, as Geometry holds them):
 *   e0       the beam direction as Geometry.beam (the reference takes it as a unit vector)
 *   axis     Geometry.axis as stored (lp_factor projects on it without normalising)
 *   pn, pm   the polarisation-plane normal orthogonalised to e0 and normalised, and pn x e0 (lp_factor's n, emain)
 *   pfrac    the polarisation fraction
 *   ed3      the detector normal ED1 x ED2 of a flat detector
 *   distance detector distance (mm)
 *   tsens, mu  sensor thickness (mm) and attenuation (1/mm): the efficiency term applies when both are > 0
 *   mu_air   linear attenuation of air (1/mm) at the wavelength (mu_linear(wavelength, "AIR"))
 *   segmented  non-zero: normals and air paths per segment, from det (xpg_detector with the segment map, the
 *            geometric-correction tables and the parallax model, as the integration kernels take it) */
typedef struct xpg_lp_geom {
  double e0[3], axis[3], pn[3], pm[3];
  double pfrac;
  double ed3[3];
  double distance, tsens, mu, mu_air;
  int32_t segmented;
} xpg_lp_geom;

/* Per-reflection inputs (device float64, n each): the diffracted beam S (n, 3), its detector position x, y (the
 * prediction's; read only for segmented detectors), the integrator's iprf, vprf, isum, vsum, p_valid, p_total, corr,
 * and coverage (NULL: 1). */
typedef struct xpg_final_in {
  const double *S, *x, *y;
  const double *iprf, *vprf, *isum, *vsum, *p_valid, *p_total, *corr;
  const float *coverage;
} xpg_final_in;

/* Outputs (device, n each): float64 rlp, I, sig, Isum, sigsum, peak, corr (100 x corr, NaN -> 0); int32 good =
 * isfinite(I) and isfinite(sig) and sig > 0.  Any pointer may be NULL (not written). */
typedef struct xpg_final_out {
  double *rlp, *I, *sig, *Isum, *sigsum, *peak, *corr;
  int32_t *good;
} xpg_final_out;

/* The finalisation of n reflections (xpg/stage_integrate.py:408-421 with xpg/geometry.py:372-407); det is read
 * only when g->segmented (may be NULL otherwise). */
int xpg_finalise(const xpg_lp_geom *g, const xpg_detector *det, int64_t n, const xpg_final_in *in,
                 const xpg_final_out *out, void *stream);

/* Host: the data records of INTEGRATE.HKL (write_integrate_hkl, xpg/io/legacy.py:168-176) as text into buf (host
 * memory, cap bytes), one line per reflection with the reference's fields and formats: H K L, IOBS and SIGMA (%.3E),
 * XCAL YCAL ZCAL (%.1f), RLP and PEAK (%.5f), CORR (the integer part of the percent correlation), 0, XOBS YOBS ZOBS
 * (%.1f), ALF0 BET0 ALF1 BET1 (%.2f), 0.00 1.  All arrays are host arrays of n (hkl: n x 3).  Returns the bytes
 * written, or -1 when buf is too small (150 bytes per record suffice for finite values of the usual magnitudes). */
int64_t xpg_format_integrate_rows(char *buf, int64_t cap, int64_t n, const int32_t *hkl, const double *I,
                                  const double *sig, const double *x, const double *y, const double *zc,
                                  const double *rlp, const double *peak, const double *corr, const double *xo,
                                  const double *yo, const double *zo, double alf0, double bet0, const double *alf1,
                                  const double *bet1);

#ifdef __cplusplus
}
#endif

#endif
