/* C ABI of the MI355X (gfx950) shoebox-integration kernels.
 *
 * Plain device pointers, sizes and one hipStream_t per call (passed as void*); no torch types.  Every entry point
 * is asynchronous on the given stream, allocates nothing and returns 0 on success or a negative xpg_status; the text
 * of the last error of the calling thread is available from xpg_last_error().
 *
 * Conventions shared by all entry points (they mirror the tensors of xpg/integrate_kernels.py):
 *   - a "chunk" is n reflections of one shoebox shape (Dz frames or phi-bins, Dy rows, Dx columns); shoebox values
 *     are float32 value codes: -3 unrecorded, -2 overloaded, -1 masked, >= 0 counts (xpg/shoebox.py:11,
 *     xpg/integrate_kernels.py:291-293), laid out (n, Dz, Dy, Dx) row-major;
 *   - angles are radians, per-reflection scalars and vectors are float64, neighbour columns float32 (n, K);
 *   - profile grids are (n_regions, 41, 41).
 */
#ifndef XPG_INTEGRATE_H
#define XPG_INTEGRATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPG_PROFILE_GRID 41
#define XPG_RC_K 21
#define XPG_MAX_DZ 65
#define XPG_MAX_K 16

enum xpg_status {
  XPG_OK = 0,
  XPG_ERR_ARG = -1,        /* invalid argument (shape, null pointer, size beyond a limit) */
  XPG_ERR_HIP = -2,        /* a HIP runtime call failed */
  XPG_ERR_TOO_LARGE = -3   /* shoebox does not fit the LDS-staged kernels */
};

/* Detector description: xpg/integrate_kernels.py:48-65 (_detector_consts).  Segment arrays are device float64 (K,3);
 * segmap is device int16 (seg_ny, seg_nx) or NULL for one panel; gx/gy device float32 geometric-correction tables or
 * NULL; bkg the device float32 background template (bkg_ny, bkg_nx) or NULL (flat). */
typedef struct xpg_detector {
  int32_t n_seg;
  const double *P0, *Bx, *By, *U1, *U2, *UN;
  double orgx, orgy, qx, qy, wl, tsens, mu, plx;
  const int16_t *segmap; int32_t seg_nx, seg_ny;
  const float *gx, *gy; int32_t geo_nx, geo_ny;
  const float *bkg; int32_t bkg_nx, bkg_ny;
} xpg_detector;

/* Scalars of one kernel call: axis, s0 (1/A), frames_phi0 = phi at frame coordinate 0, osc per frame, sigma_D,
 * sigma_M (all radians), gain, the 21 rocking-curve mixture weights (use_rcw = 0: plain Gaussian). */
typedef struct xpg_call {
  double axis[3], s0[3];
  double frames_phi0, osc, sd, sm, gain;
  float rcw[XPG_RC_K];
  int32_t use_rcw;
  const struct xpg_widths *widths;   /* device pointer or NULL: when set, sd, sm and rcw are read from it */
} xpg_call;

/* Adaptive widths of one launch, in device memory (written by xpg_adapt_block, read by the kernels when
 * xpg_call.widths is set): sigma_D, sigma_M (radians) and the rocking-curve mixture weights. */
typedef struct xpg_widths {
  double sd, sm;
  float rcw[XPG_RC_K];
  int32_t n_strong;        /* strong reflections the widths were fitted on (0: carried over) */
  uint32_t active;         /* bit j: fitted mixture weight j is positive (the next fit starts from this set; 0: cold) */
  double sd_fit, sm_fit;   /* this block's own estimates before shrinkage (radians; 0 when not fitted) */
} xpg_widths;

/* Per-reflection table of a chunk (device pointers), xpg/integrate.py:509-517.  pxs may be NULL (1); has_p, pS,
 * pphi may be NULL (no partner); K = 0 means no neighbour columns. */
typedef struct xpg_refl {
  int32_t n, Dz, Dy, Dx;
  const double *S, *phi;
  const int32_t *ix0, *iy0, *iz0, *fb, *pxs, *region;
  const float *has_p;
  const double *pS, *pphi;
  int32_t K;
  const float *nbx, *nby, *nbz, *nbsx, *nbsz;
  /* Mixed shapes in one launch (all NULL: every row is Dz x Dy x Dx).  rdz, rdy, rdx: each row's box; Dz, Dy, Dx then
   * hold their maxima (they size the LDS).  voff: offset of a row's voxels in sb and nbm (elements); poff: offset of
   * its pixels (a row's geo / geo_p block starts at element 3 poff, its per-pixel pass-1 outputs at poff).
   * wave_lds: LDS bytes per wavefront of the launch, at least xpg_lds_bytes of every row's shape (0: sized from
   * Dz x Dy x Dx); a row whose layout exceeds it is skipped, its outputs left as they were.
   * wg_waves: at most this many wavefronts (rows) per workgroup (0: as many as fit, up to 4).
   * lds_rows: pass 2 of prepared rows whose boxes all fit the register kernel (at most 8 frames and 128 pixels) runs
   * the rows with an overlapping partner in the LDS kernel; lds_rows > 0 says that none of them lies past the first
   * lds_rows rows of the table (the LDS kernel covers only those), -1 that there is none; 0 (the default)
   * gives no bound. */
  const int32_t *rdz, *rdy, *rdx;
  const int64_t *voff, *poff;
  int32_t wave_lds;
  int32_t wg_waves;
  int32_t lds_rows;
} xpg_refl;

/* pass 1 outputs (xpg/integrate_kernels.py:362-365).  Per-reflection float64 / int32 (n); fz (n, Dz); per-pixel
 * arrays (n, Dy, Dx).  Any per-pixel pointer may be NULL (not written). */
typedef struct xpg_pass1_out {
  double *bg, *nbg, *isum, *vsum, *vsum0, *xo, *yo, *zo, *m_d, *m_m, *frac_rec, *nfg, *zeta, *seff;
  int32_t *n_over, *n_fg_masked, *n_fg_nb;
  float *fz;               /* (n, fz_stride): frames 0..Dz-1 written, the rest of each row zeroed */
  int32_t fz_stride;       /* 0: Dz */
  float *img2, *J, *u1, *u2;
  uint8_t *all_valid2;
  /* Per-voxel classes (n, Dz, Dy, Dx) at each row's voxel offset, or NULL (not written; for mask comparisons):
   * XPG_VM_FG foreground (ellipse x foreground frame), XPG_VM_BGM background candidate, XPG_VM_NBM inside a neighbour
   * ellipsoid, XPG_VM_BGW in the final background set (after the last clipping round), XPG_VM_FGV foreground, valid and
   * not claimed by a neighbour. */
  uint8_t *vmask;
} xpg_pass1_out;

enum { XPG_VM_FG = 1, XPG_VM_BGM = 2, XPG_VM_NBM = 4, XPG_VM_BGW = 8, XPG_VM_FGV = 16, XPG_VM_USE = 32 };

/* pass 2 outputs (xpg/integrate_kernels.py:429, :482), float64 (n). */
typedef struct xpg_pass2_out {
  double *iprf, *vprf, *vprf0, *corr, *p_valid, *p_total, *ipartner, *rho, *p_fg;
  /* Per-voxel classes (n, Dz, Dy, Dx) at each row's voxel offset, or NULL: XPG_VM_USE in the fit mask, XPG_VM_NBM
   * inside a neighbour ellipsoid other than the modelled partner's. */
  uint8_t *vmask;
} xpg_pass2_out;

/* Fixed-point profile accumulation grids: int64 (n_regions, 41, 41) each, value * 2^44, zeroed by the caller (a cell
 * holds up to 2^19 in value units).
 * Integer addition is associative, so the grids are bit-identical for any launch order. */
typedef struct xpg_profile_acc {
  int64_t *num, *den;
  int32_t n_regions;
  /* copies > 1: num and den each hold `copies` consecutive (n_regions, 41, 41) grids; the workgroups of a launch spread
   * their integer increments over them (fewer atomics on one address) and xpg_normalise_profiles adds them up.
   * 0 or 1: one grid. */
  int32_t copies;
} xpg_profile_acc;

/* Device frame store: uint32 raw counts (n_frames, ny, nx) for 0-based frames frame0 .. frame0 + n_frames - 1, and the
 * value codes of xpg/io/images.py:80-96 with the integrator's mask ceiling (xpg/integrate.py:648-702).
 * available: per store frame 1 = read, 0 = excluded / missing (NULL: all read).  src_mask: 0 = untrusted pixel
 * (-1, wins over overload; NULL: none).  ceil_mask: 0 = min(value, -1) (a masked overload keeps -2; NULL: none).
 * Raw values < min_valid or == flag_value are -1; >= overload are -2 (overload 0: none).  Float32 frames (dtype 1)
 * carry the codes already (xpg/io/images.py ImageSet.get) and only src_mask / ceil_mask apply. */
typedef struct xpg_frames {
  const void *data;            /* dtype 0: uint32 raw counts; dtype 1: float32 frames already value-coded */
  int64_t frame_stride;        /* elements between frames (>= ny * nx) */
  int32_t nx, ny, frame0, n_frames;
  const uint8_t *available, *src_mask, *ceil_mask;
  uint32_t min_valid, flag_value, overload;
  int32_t use_flag_value;
  int32_t dtype;
  int32_t ring;                /* > 0: a ring of `ring` frame slots, frame frame0 + t in slot t % ring (the caller keeps
                                * every frame a launch reads resident); 0: frame frame0 + t in slot t */
} xpg_frames;

/* ---- entry points ------------------------------------------------------------------------------------------- */

/* Library version string and the last error of the calling thread. */
const char *xpg_version(void);
const char *xpg_last_error(void);

/* Replaces the host shoebox gather of xpg/integrate.py:704-803 (_lean_task / frame_loop) and the bin decoding of
 * xpg/shoebox.py:264-268 (_decode_bins) with xpg/integrate.py:484-488: writes the value-coded shoeboxes sb
 * (n, Dz, Dy, Dx) of a chunk from the device frame store, with s x s super-pixels (pxs) and phi-bins of fb frames. */
int xpg_extract(const xpg_frames *frames, const xpg_refl *refl, float *sb, void *stream);

/* Replaces _decode_bins (xpg/shoebox.py:264-268): raw IEEE-encoded bin sums (+inf masked, NaN overloaded) and bin
 * counts -> stored codes (value + 3, 0 = unrecorded), in place.  arr (n, B, bx, bx) float32, cnt (n, B) int16. */
int xpg_decode_bins(float *arr, const int16_t *cnt, int64_t n_rows, int32_t B, int32_t bxbx, void *stream);

/* Replaces pass1 of make_kernels (xpg/integrate_kernels.py:287-365).  If acc is not NULL the strong-reflection
 * profile accumulation of xpg/integrate.py:525-543 (accumulate_profiles on rows with I/sigma > 10, no overload,
 * frac_rec > 0.95, no foreground pixel masked or claimed by a neighbour) is fused in. */
int xpg_pass1(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
              const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream);

/* xpg_pass1 reading the shoeboxes straight from the device frame store (extraction fused into the kernel, no sb
 * buffer): same outputs as xpg_extract followed by xpg_pass1. */
int xpg_pass1_frames(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                     const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream);

/* Replaces accumulate_profiles of make_kernels (xpg/integrate_kernels.py:367-380) for the rows with sel != 0.
 * img2, J, u1, u2 (n, Dy, Dx) float32; ok2 (n, Dy, Dx) uint8; region int32, isum / seff float64, sel uint8 (n). */
int xpg_accumulate_profiles(const float *img2, const uint8_t *ok2, const float *J, const float *u1, const float *u2,
                            const int32_t *region, const double *isum, const uint8_t *sel, const double *seff,
                            int32_t n, int32_t Dy, int32_t Dx, const xpg_profile_acc *acc, void *stream);

/* Replaces the float64 profile normalisation of xpg/integrate.py:546-556: fixed-point grids -> float32 profiles
 * (n_regions, 41, 41) regularised towards gauss (41, 41) with 3 pseudo-counts; num64/den64 (may be NULL) receive
 * the float64 grids. */
int xpg_normalise_profiles(const xpg_profile_acc *acc, const float *gauss, float *prof, double *num64, double *den64,
                           void *stream);

/* Replaces pass2 of make_kernels (xpg/integrate_kernels.py:382-483): Poisson-weighted profile fit with the 2x2
 * overlap deconvolution.  bg, nbg from pass 1 (float64, n); prof (n_regions, 41, 41) float32.  irls_rule 0: each
 * reflection stops on its own convergence test (independent of which reflections share the call); 1: the chunk-global
 * rule of xpg/integrate_kernels.py:464-467 evaluated over this call's n reflections. */
int xpg_pass2(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
              const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
              const xpg_pass2_out *out, void *stream);

/* xpg_pass2 reading the shoeboxes straight from the device frame store: same outputs as xpg_extract + xpg_pass2. */
int xpg_pass2_frames(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                     const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
                     const xpg_pass2_out *out, void *stream);

/* Replaces _adapt_block (xpg/integrate.py:404-461) on the device, without a host round trip: from the pass-1 results
 * of one phi block (rows [0, n) of the arrays below, in any fixed order) it selects the strong reflections (I/sigma >
 * 15, no overload, frac_rec > 0.98, fb == 1), fits sigma_M on the 61-point log grid of _fit_rocking_width
 * (xpg/integrate.py:1130-1154) against their per-frame net counts fz (n, 65), shrinks it towards the run's sigma_M with
 * a 300-reflection prior, fits the 21 rocking-curve weights by non-negative least squares on the Cholesky factor of
 * the Gram matrix (_nnls_gram, :1095-1104) regularised towards the Gaussian, and sigma_D from the median of the
 * truncation-corrected second moments.  in: the widths the block's pass 1 used; out_pass2: widths for this block's
 * pass 2; out_next: widths for the next block's pass 1 (in copied when fewer than 100 strong reflections).
 * Per-row inputs (device): isum, vsum, frac_rec, m_d, phi (deg), zeta_k, Xlen (|X| mm), cosa (|X.ed3|/|X|) float64;
 * n_over, fb, iz0 int32.  Scalars: run widths sd_run, sm_run (radians), frames_phi0_deg + osc_deg (phi of frame
 * coordinate z = phi0 + osc z), qx.  workspace: xpg_adapt_workspace_bytes(n) bytes of device memory. */
int xpg_adapt_block(int32_t n, const double *isum, const double *vsum, const int32_t *n_over, const double *frac_rec,
                    const double *m_d, const float *fz, const double *phi_deg, const double *zeta_k,
                    const int32_t *iz0, const int32_t *fb, const double *Xlen, const double *cosa, double sd_run,
                    double sm_run, double phi0_deg, double osc_deg, double qx, const xpg_widths *in,
                    xpg_widths *out_pass2, xpg_widths *out_next, void *workspace, int64_t workspace_bytes,
                    void *stream);
int64_t xpg_adapt_workspace_bytes(int32_t n);

/* Widths-independent per-reflection data of a chunk (device pointers), written by xpg_prepare ahead of the adaptive
 * chain and read by xpg_pass1_prepared / xpg_pass2_prepared. */
typedef struct xpg_prepared {
  float *sb;       /* (n, Dz, Dy, Dx) value-coded shoeboxes, as xpg_extract writes them */
  double *geo;     /* (n, 3, Dy, Dx): eps1, eps2 and the pixel area J at every pixel centre */
  double *zeta;    /* (n,) zeta of the Kabsch frame */
  uint8_t *nbm;    /* (n, Dz, Dy, Dx) bit 0: inside a neighbour ellipsoid; bit 1: the same with column 0 skipped */
  double *geo_p;   /* (n, 3, Dy, Dx): the overlapping partner's eps1, eps2, J on this box (rows with has_p only) */
  double *zeta_p;  /* (n,) the partner's zeta (rows with has_p only) */
  /* (n,) or NULL: the effective spot width seff of each row, given instead of computed from sigma_D and the centre
   * pixel's area (the reference's pass 2 with pass 1's ``pix``, xpg/integrate_kernels.py:259-260: eps = u seff) */
  const double *seff;
  /* (n, r_stride) or NULL: the rows' frame-integrated rocking curves R for the widths of the launch that reads them,
   * written by xpg_rocking (NULL: the pass kernels evaluate them) */
  double *R;
  int32_t r_stride;        /* doubles per row of R (>= every row's Dz) */
} xpg_prepared;

/* Shoebox gather (host gather xpg/integrate.py:704-803 with _decode_bins xpg/shoebox.py:264-268), the widths-independent
 * geometry of geometry_terms (Kabsch frame, eps, J: xpg/integrate_kernels.py:219-272) and the neighbour masks
 * (_neighbour_mask, :186-213) of a chunk, into prep. */
int xpg_prepare(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                const xpg_prepared *prep, void *stream);

/* xpg_prepare from a chunk of value-coded shoeboxes sb (n, Dz, Dy, Dx) instead of the frame store (the kernel boundary
 * of xpg/integrate_kernels.py:216-485, whose closures take the chunk); prep->sb may be sb itself. */
int xpg_prepare_sb(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
                   const xpg_prepared *prep, void *stream);

/* The rocking-curve half of geometry_terms (xpg/integrate_kernels.py:246-252 with _phi_rc :130-141): each row's
 * frame-integrated rocking curve R (Dz values) for the widths of call, into prep->R (one lane per row and frame; the
 * same arithmetic as the pass kernels' own evaluation, so their results do not change).  Reads prep->zeta. */
int xpg_rocking(const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep, void *stream);

/* xpg_pass1 / xpg_pass2 reading the prepared data of xpg_prepare (results identical to xpg_pass1_frames /
 * xpg_pass2_frames on the same frames). */
int xpg_pass1_prepared(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep,
                       const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream);
int xpg_pass2_prepared(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep,
                       const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
                       const xpg_pass2_out *out, void *stream);

/* Bytes of LDS one wavefront of pass 1 / pass 2 uses for a (Dz, Dy, Dx) shoebox (launch planning, autotuning). */
int xpg_lds_bytes(int32_t Dz, int32_t Dy, int32_t Dx);

/* LDS bytes a workgroup may use on the current device.  Shoeboxes whose xpg_lds_bytes exceeds it are integrated from
 * gathered values (xpg_prepare / xpg_extract), read in place; the *_frames entry points reject them. */
int xpg_lds_limit(void);

/* The same byte range of `height` rows of a row-major block, between device and (pinned) host memory or within one
 * of them, as one asynchronous 2-D copy on `stream`: dst + i * dst_pitch <- src + i * src_pitch, `width` bytes each.
 * No counterpart in the reference (its results are host arrays already): a sweep moves each phi block's rows of every
 * per-reflection result to the host with one such copy per dtype, behind the block's pass 2. */
int xpg_copy_rows(void *dst, int64_t dst_pitch, const void *src, int64_t src_pitch, int64_t width, int64_t height,
                  void *stream);

#ifdef __cplusplus
}
#endif
#endif
