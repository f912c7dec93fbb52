/* Forward (window) model passes on MI355X: the scatter of a window's modelled counts and the two per-reflection
 * gathers of its intensity fit (xpg/forward/kernels.py:34-134, used by WindowModel.mu / grad_ratio / grad_newton,
 * xpg/forward/model.py:477-537).
 *
 * A reflection's fraction in voxel (z, y, x) of its support is separable: f = sum_t R[t][row][z] P[t][row][y][x]
 * (T = 1 term, or 2 with the halo), R (T, n, Dz) and P (T, n, b, b) float32.  The window is frames [f0, f0 + nf),
 * rows [y0, y0 + ny), columns [x0, x0 + nx), its voxels indexed (z - f0) ny nx + (y - y0) nx + (x - x0); a support
 * voxel outside the window is skipped.  All pointers are device pointers; `stream` is a hipStream_t. */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xpg_fm_window {
  int32_t f0, nf, y0, ny, x0, nx;
} xpg_fm_window;

/* The supports of a launch: n reflections of one shape (Dz frames, b x b pixels), T terms. */
typedef struct xpg_fm_supports {
  const float *R;            /* (T, n, Dz) */
  const float *P;            /* (T, n, b, b) */
  const int32_t *iz0, *iy0, *ix0;   /* (n): the support's first frame, row, column (window coordinates + origin) */
  int32_t T, n, Dz, b;
} xpg_fm_supports;

/* NS element types of xpg_fm_gather (the window's counts; negative = no likelihood term) */
enum { XPG_FM_NS_I16 = 0, XPG_FM_NS_I32 = 1, XPG_FM_NS_F32 = 2 };

/* Replaces _scatter_kernel / scatter (xpg/forward/kernels.py:34-61, :107-120): mu[v] += I[row] f + beta[row] over the
 * in-window voxels of every support (beta may be NULL).  Float atomics: the order of additions to a voxel shared by
 * several supports is not fixed, as in the reference. */
int xpg_fm_scatter(float *mu, const xpg_fm_supports *sup, const float *I, const float *beta, const xpg_fm_window *win,
                   void *stream);

/* Replaces _gather_kernel / gather (xpg/forward/kernels.py:63-104, :123-134).  Over the in-window voxels of row r's
 * support with NS >= 0: mode 0 g[rows[r]] += sum f A / M; mode 1 g[rows[r]] += sum f (1 - max(NS, 0) / M) and
 * fd[rows[r]] += sum f^2 / max(M, 1e-6).  g and fd hold n_out values; a reflection whose rows[r] lies outside
 * [0, n_out) adds nothing.  One wavefront per reflection sums its support in a fixed order and adds once (repeated
 * rows in one launch add in an unspecified order). */
int xpg_fm_gather(float *g, float *fd, const xpg_fm_supports *sup, const int64_t *rows, int64_t n_out, const float *A,
                  const float *M, const void *NS, int32_t ns_type, int32_t mode, const xpg_fm_window *win,
                  void *stream);

/* ---- the separable factors themselves (fraction_terms, xpg/forward/model.py:206-275) ------------------------------
 * For each reflection r: the Kabsch frame of S, eps1 / eps2 at the centres of its b x b support pixels (super-pixels of
 * pxs x pxs when pxs > 1), the pixel area J by second-order differences (b >= 3; else (qx pxs)^2), seff from sigma_D
 * and the centre pixel's J, the rocking curve R over frames iz0 .. iz0 + Dz - 1 (the mixture with the given weights,
 * else a Gaussian), and P = the learnt profile of the reflection's region sampled bilinearly at u = eps / seff times
 * J / seff^2 (no profiles: the pixel-integrated Gaussian exp(-r^2 / 2 seff^2) J / (2 pi seff^2)).  With the halo the
 * factors are ((1 - w) R, P) and (w R_h, P_h): R_h the Gaussian rocking curve of kappa sigma_M, P_h the Gaussian of
 * width sqrt((kappa sigma_D)^2 + pix^2 / 12). */
typedef struct xpg_fm_rows {
  const double *S;                   /* (n, 3) diffracted beams, 1/A */
  const double *phi;                 /* (n) predicted rotation angles, radians */
  const int32_t *ix0, *iy0, *iz0;    /* (n) support origin: column, row (super-pixels when pxs > 1), frame */
  const int32_t *region;             /* (n) profile region, or NULL without learnt profiles */
  const double *sd, *sm;             /* (n) per-reflection widths (radians), or NULL: the model's scalars */
  int32_t n, Dz, b, pxs;
} xpg_fm_rows;

typedef struct xpg_fm_model {
  double axis[3], s0[3];
  double frames_phi0, osc;           /* radians: phi at frame coordinate 0, per frame */
  double sd, sm;                     /* radians */
  const float *prof;                 /* (n_regions, G, G) learnt profiles (density in seff units), or NULL */
  int32_t n_regions, G;
  double umax;                       /* half-width of the profile grid, seff units */
  int32_t halo;                      /* non-zero: the halo term, weight halo_w and broadening halo_kappa */
  double halo_w, halo_kappa;
  float rcw[21];                     /* rocking-curve mixture weights when use_rcw */
  int32_t use_rcw;
} xpg_fm_model;

/* R (T, n, Dz) and P (T, n, b, b) float32 (T = 2 with the halo, else 1) of the reflections of `rows`. */
int xpg_fm_terms(const struct xpg_detector *det, const xpg_fm_model *m, const xpg_fm_rows *rows, float *R, float *P,
                 void *stream);

#ifdef __cplusplus
}
#endif
