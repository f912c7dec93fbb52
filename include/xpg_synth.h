/* Seeded synthetic rotation data generated on the device (workload generator for the benchmarks and the serving
 * loop).  No counterpart in the reference's integration path: it plays the role of xpg/simulate.py (spots as 3D
 * Gaussians with an erf rocking curve on a Poisson background) at full detector size, directly into the device frame
 * store, in float32 arithmetic.  Counter-based random numbers keyed by (seed, frame, pixel) and (seed, reflection,
 * frame, pixel) make every frame reproducible independently of batching and launch order.
 */
#ifndef XPG_SYNTH_H
#define XPG_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Background and static pixel flags of frames [f0, f0 + n) into frames (n, ny, nx) uint32:
 * counts ~ Poisson(bg_centre + (bg_edge - bg_centre) * r / r_max) with r the distance (pixels) from (cx, cy);
 * a fraction dead_frac of pixels (fixed per pixel) holds flag_value, a fraction sat_frac holds sat_value. */
typedef struct xpg_synth_bg {
  uint64_t seed;
  int32_t nx, ny, f0, n;
  double cx, cy, bg_centre, bg_edge;
  double dead_frac, sat_frac;
  uint32_t flag_value, sat_value;
} xpg_synth_bg;

int xpg_synth_background(const xpg_synth_bg *p, uint32_t *frames, void *stream);

/* Spots of n reflections added to frames [f0, f0 + nf): expected counts counts_r * F_r(frame) * g_r(pixel) with F the
 * erf-integrated Gaussian rocking curve (centre z, width sig_z frames) and g an elliptical Gaussian (radial width
 * sig_r, tangential sig_t pixels, radial direction (rdx, rdy)) sampled on 3x3 sub-pixels within +-half pixels;
 * Poisson draws are added to pixels that do not hold a flag.  Arrays are device float64 (n). */
typedef struct xpg_synth_spots {
  uint64_t seed;
  int32_t nx, ny, f0, nf, n, half;
  const double *x, *y, *z, *counts, *sig_r, *sig_t, *sig_z, *rdx, *rdy;
  uint32_t flag_value, sat_value;
  const int32_t *sid;   /* each spot's index in its dataset's spot table, keying its random draws (NULL: 0 .. n-1),
                         * so frames do not depend on which frames a launch renders */
} xpg_synth_spots;

int xpg_synth_spots_add(const xpg_synth_spots *p, uint32_t *frames, void *stream);

#ifdef __cplusplus
}
#endif
#endif
