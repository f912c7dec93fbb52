/* C ABI of the geometry refinement's device work (gfx950): the normalised spot-position residuals of many parameter
 * vectors of many refinement problems in one launch, and the weighted normal equations of their forward-difference
 * Jacobians.
 *
 * Replaces the device part of xpg/refine.py:151-244 (_residual_fns: residuals, jac_r, normal, with predict_xyphi_t of
 * xpg/refine.py:253-280): the per-block refinement before integration (xpg/stage_integrate.py:28-153) runs its
 * Levenberg-Marquardt loops on the host and asks the device for residuals and p x p normal equations, for every phi
 * block that is waiting at once.  Same conventions as xpg_integrate.h: device pointers, asynchronous on the given
 * stream, 0 or a negative xpg_status (xpg_last_error() has the text).  All arithmetic is float64.
 *
 * A "model" is one parameter vector's experiment constants, XPG_RF_HEAD + XPG_RF_SEG n_seg doubles:
 *   [0:3]   s0 = unit(beam) / wavelength        [3:6]  unit spindle axis
 *   [6:15]  A* row-major (columns a*, b*, c*)    [15]   orgx   [16] orgy
 *   then per segment k at XPG_RF_HEAD + XPG_RF_SEG k: P0 (3), Bx (3), By (3) -- the lab affine map of the segment,
 *   X = P0 + (x - orgx) Bx + (y - orgy) By (xpg/detector.py:84-110).
 */
#ifndef XPG_REFINE_H
#define XPG_REFINE_H

#include <stdint.h>

#include "xpg_integrate.h"

#ifdef __cplusplus
extern "C" {
#endif

#define XPG_RF_HEAD 17
#define XPG_RF_SEG 9

/* Observed spots of all problems (device arrays of n rows). */
typedef struct xpg_rf_spots {
  const double *h;         /* (n, 3) Miller indices */
  const double *phi;       /* observed spindle angle (radians) */
  const double *branch;    /* root of the diffraction condition: +1 or -1 (choose_branch) */
  const double *xo, *yo;   /* observed position after the geometric correction (pixels) */
  const int32_t *seg;      /* segment the spot was observed on, 0 .. n_seg - 1 */
  int64_t n;
} xpg_rf_spots;

/* Scalars: sigma_xy (pixels) and sigma_phi (radians) normalise the residuals; tsens, mu and plx the thick-sensor
 * parallax (plx = 0 or 1, as xpg/refine.py:172-178 multiplies by it).  Read by xpg_rf_normal only: n_models (every job
 * of a normal-equations launch has that many models, p = n_models - 1 parameters) and robust_c (> 0: each spot's
 * weight is w / (1 + (|r_0| / robust_c)^2), |r_0| the norm of the spot's three residuals under model 0 -- the Cauchy
 * weights of xpg/refine.py:366-370; 0: w as given). */
typedef struct xpg_rf_setup {
  double qx, qy, sxy, sphi, tsens, mu, plx, robust_c;
  int32_t n_seg, n_models;
} xpg_rf_setup;

/* One job: the spots [spot0, spot0 + n_spots) of one problem under n_models consecutive models starting at model
 * `model0` (models + model0 model_doubles); its residuals go to r + r0 as (n_models, 3, n_spots).  t0 is the job's
 * first thread, the sum of n_models n_spots over the jobs before it (the jobs are in launch order).  w0: where the
 * job's n_spots weights start in xpg_rf_normal's w. */
typedef struct xpg_rf_job {
  int64_t spot0, model0, r0, t0, w0;
  int32_t n_spots, n_models;
} xpg_rf_job;

/* For every job and every model m and spot i of it: ((x - xo) / sxy, (y - yo) / sxy, (phi - phi_obs) / sphi) with
 * (x, y, phi) the spot predicted by model m.  jobs is a device array of n_jobs; total = the sum of n_models n_spots.
 * Replaces residuals (xpg/refine.py:154-183). */
int xpg_rf_residuals(const xpg_rf_spots *sp, const xpg_rf_setup *su, const xpg_rf_job *jobs, int32_t n_jobs,
                     int64_t total, const double *models, double *r, void *stream);

/* The normal equations of forward-difference Jacobians (xpg/refine.py:185-216, 218-222), one per job: r as written by
 * xpg_rf_residuals for the job, model 0 the parameter vector and model 1 + k that vector with parameter k moved by
 * h[j p + k]; J[row][k] = (r_{1+k}[row] - r_0[row]) / h[j p + k];
 * out[j] = (J^T W J (p x p, row-major), J^T W r_0 (p)) over the job's rows, W = the weight of the job's spot i (w[w0 + i],
 * or its Cauchy form, robust_c) on each of its three rows.  Sums run in a fixed order (the same result on every call).
 * out holds n_jobs (p p + p) doubles.  r0_out (or NULL): job j's model-0 residuals copied to r0_out + 3 w0 as
 * (3, n_spots), so one launch answers "residuals at x" and "normal equations at x" together. */
int xpg_rf_normal(const xpg_rf_spots *sp, const xpg_rf_setup *su, const xpg_rf_job *jobs, int32_t n_jobs,
                  const double *r, const double *h, const double *w, double *out, double *r0_out, void *stream);

#ifdef __cplusplus
}
#endif

#endif
