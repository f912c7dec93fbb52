// Shoebox integration kernels for MI355X (gfx950): extraction from a device frame store, pass 1 (masks, robust
// background, summation, centroids, moments, projected images, fused profile accumulation), profile normalisation
// and pass 2 (Poisson-weighted profile fit with 2x2 overlap deconvolution), behind the C ABI of
// include/xpg_integrate.h.
//
// Work decomposition: one wavefront per reflection, W wavefronts per workgroup, each with its own slice of dynamic LDS
// holding the shoebox values, per-voxel flags and the per-pixel / per-frame geometry.  No workgroup barriers: a
// wavefront synchronises its lanes through LDS with wave_sync().  Reductions are cross-lane.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "xpg_common.h"

using namespace xpg;

// ---------------------------------------------------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------------------------------------------------

static thread_local char g_err[512] = "";

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

namespace xpg {
// error message of the other sources of the library (xpg_decode.hip), read back by xpg_last_error()
int report_error(int code, const char *msg) { return fail(code, "%s", msg); }
}  // namespace xpg

#define HIP_TRY(expr)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(XPG_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
  } while (0)

// ---------------------------------------------------------------------------------------------------------------------
// LDS layout of one wavefront, per pass
//   vals  float[V]    shoebox values                 vflag u8[V]   per-voxel flags
//   scr   double[21 Dz]  rocking-curve terms (and the recorded-frame fractions of pass 1)
//   U1,U2 double[P]   u1, u2                          Jd    double[P] pixel area (pass 2: then the partner's)
//   pa    8P bytes    pass 1: bref, thr (float)    /  pass 2: p2 (double)   (staging: detector offsets, mask bits)
//   pb    8P bytes                                     pass 2: pp2 (double)
//   pc    4P bytes                                     pass 2: bref (float)
//   prow  u8[P]       row of each pixel               pflag u8[P]
//   R     double[Dz]  rocking-curve fractions (staging: store frames)    Rp double[Dz] pass 2: the partner's
//   fr    pass 1: zz, e3 (double[Dz] each), flags u8[Dz] / pass 2: R, Rp (float[Dz] each), flags u8[Dz]
//   co    double[Dx + Dy + Dz] box coordinates        tab   double[Dx + Dy + Dz] neighbour terms
// ---------------------------------------------------------------------------------------------------------------------

__host__ __device__ static inline size_t al8(size_t x) { return (x + 7) & ~(size_t)7; }

struct Layout {
  size_t vals, vflag, scr, U1, U2, Jd, pa, pb, pc, prow, pflag, R, Rp, fr, co, tab, bytes;
};

// vals_lds = false: the shoebox values are read in place from a device buffer (boxes too large for LDS).
// compact: pass-1 register rows over prepared data with their rocking curve given (compact_rows): voxels and their
// bits live in registers, the curve comes ready and the neighbour masks were made ahead, so no LDS for voxels, curve
// terms or neighbour tables (scr keeps Dz doubles, the per-frame scratch) -- a fraction of the full layout, and with
// it more wavefronts per CU.
__host__ __device__ static inline Layout make_layout(int Dz, int Dy, int Dx, int pass, bool vals_lds = true,
                                                     bool compact = false) {
  const size_t P = (size_t)Dy * Dx, V = (size_t)Dz * P, C = (size_t)Dx + Dy + Dz;
  Layout L;
  size_t o = 0;
  L.vals = o; if (vals_lds && !compact) o = al8(o + 4 * V);
  L.vflag = o; if (!compact) o = al8(o + V);
  L.scr = o; o += 8 * (compact ? 1 : 21) * (size_t)Dz;         // compact: the per-frame scratch of pass 1 only
  L.U1 = o; o += 8 * P;
  L.U2 = o; o += 8 * P;
  L.Jd = o; o += 8 * P;
  L.pa = o; o += 8 * P;
  L.pb = o; if (pass == 2) o += 8 * P;
  L.pc = o; if (pass == 2) o = al8(o + 4 * P);
  L.prow = o; o = al8(o + P);
  L.pflag = o; o = al8(o + P);
  L.R = o; o += 8 * Dz;
  L.Rp = o; if (pass == 2) o += 8 * Dz;
  L.fr = o; o = al8(o + (pass == 2 ? 9 : 17) * (size_t)Dz);
  L.co = o; o += 8 * C;
  L.tab = o; if (!compact) o += 8 * C;
  L.bytes = al8(o);
  return L;
}

// pass-1 rows of this shape over prepared data with a rocking-curve table use the compact layout (make_layout)
__host__ __device__ static inline bool compact_rows(int Dz, int Dy, int Dx, bool prep, bool rtab) {
  return prep && rtab && Dy * Dx <= 2 * 64 && Dz <= 8;
}

// frames per batch of LDS reads in the per-pixel frame loops of boxes that are not held in registers: the reads
// of a batch are in flight together, then its voxels are used in frame order
constexpr int LB = 8;

// per-voxel flags: neighbour ellipsoid, fit mask (pass 2), background candidate, foreground (fg & valid & ~nbm),
// foreground ellipse x foreground frame
enum : uint8_t { VF_NBM = 1, VF_USE = 2, VF_BGM = 4, VF_FGV = 8, VF_FG = 16 };
enum : uint8_t { PF_FG2 = 1, PF_BGR = 2 };      // per-pixel flags: r^2 <= FG_R^2, r^2 >= BG_R^2

// ---------------------------------------------------------------------------------------------------------------------
// shoebox values: from a chunk buffer, or extracted from the frame store
// ---------------------------------------------------------------------------------------------------------------------

// value code of one detector pixel: xpg/io/images.py:80-96 (min_valid / flagged / static mask -> -1 and it wins,
// overload -> -2) and the integrator's mask ceiling of xpg/integrate.py:661-664 (min(value, -1))
__device__ __forceinline__ float frame_code(const xpg_frames &F, int64_t t, int iy, int ix) {
  const int64_t p = (int64_t)iy * F.nx + ix;
  float v;
  if (F.dtype == 1) {
    v = ((const float *)F.data)[t * F.frame_stride + p];
    if (F.src_mask && !F.src_mask[p]) v = -1.0f;
  } else {
    const uint32_t raw = ((const uint32_t *)F.data)[t * F.frame_stride + p];
    bool bad = raw < F.min_valid || (F.use_flag_value && raw == F.flag_value);
    if (F.src_mask && !F.src_mask[p]) bad = true;
    v = (float)raw;
    if (F.overload && raw >= F.overload) v = -2.0f;
    if (bad) v = -1.0f;
  }
  if (F.ceil_mask && !F.ceil_mask[p]) v = fminf(v, -1.0f);
  return v;
}

// (super-)pixel (sy, sx) of store frame t: the padded frame's -1 border, s x s super-pixel sums flagged -2 if any
// pixel is overloaded else -1 if any is negative, incomplete super-pixels at the far edges masked (xpg/integrate.py:
// 648-702)
__device__ __forceinline__ float pixel_value(const xpg_frames &F, int64_t t, int sy, int sx, int s) {
  if (s == 1) {
    if (sy < 0 || sx < 0 || sy >= F.ny || sx >= F.nx) return -1.0f;
    return frame_code(F, t, sy, sx);
  }
  if (sy < 0 || sx < 0 || sy >= F.ny / s || sx >= F.nx / s) return -1.0f;
  float sum = 0.f, mn = 0.f;
  for (int a = 0; a < s; ++a)
    for (int b = 0; b < s; ++b) {
      float v = frame_code(F, t, sy * s + a, sx * s + b);
      if (a == 0 && b == 0) {
        sum = v;
        mn = v;
      } else {
        sum = sum + v;
        mn = fminf(mn, v);
      }
    }
  if (mn < 0.f) return mn <= -1.5f ? -2.0f : -1.0f;
  return sum;
}

// frame fi recorded and readable: t = its slot in the store
__device__ __forceinline__ bool frame_read(const xpg_frames &F, int fi, int64_t &t) {
  t = (int64_t)fi - F.frame0;
  if (t < 0 || t >= F.n_frames) return false;
  if (F.available != nullptr && F.available[t] == 0) return false;
  if (F.ring > 0) t %= F.ring;
  return true;
}

// voxel (j, sy, sx): unbinned frames are read or unrecorded (-3); phi-bins of fb frames are summed with the
// reference's IEEE encoding (an overload in any frame wins, then masked), a bin missing a frame is unrecorded
// (xpg/integrate.py:704-716, xpg/shoebox.py:264-268, xpg/integrate.py:484-488)
__device__ float voxel_value(const xpg_frames &F, int iz0, int fb, int j, int sy, int sx, int s) {
  int64_t t;
  if (fb == 1) {
    if (!frame_read(F, iz0 + j, t)) return -3.0f;
    return pixel_value(F, t, sy, sx, s);
  }
  float sum = 0.f;
  bool any_over = false, any_mask = false, first = true;
  for (int q = 0; q < fb; ++q) {
    if (!frame_read(F, iz0 + fb * j + q, t)) return -3.0f;
    float v = pixel_value(F, t, sy, sx, s);
    if (v < 0.f) {
      if (v == -2.0f) any_over = true;
      else any_mask = true;
    } else {
      sum = first ? v : sum + v;
      first = false;
    }
  }
  if (any_over) return -2.0f;
  if (any_mask) return -1.0f;
  return sum;
}

__global__ void k_decode_bins(float *arr, const int16_t *cnt, int64_t n_rows, int B, int bxbx) {
  const int64_t total = n_rows * B * (int64_t)bxbx;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    int64_t rb = g / bxbx;
    float v = arr[g];
    if (isnan(v)) v = -2.0f;
    else if (isinf(v)) v = -1.0f;
    v = v + 3.0f;
    if (cnt[rb] == 0) v = 0.0f;
    arr[g] = v;
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// per-reflection setup shared by pass 1 and pass 2
// ---------------------------------------------------------------------------------------------------------------------

struct Box {
  int r, Dz, Dy, Dx, P, V, ix0, iy0, iz0, fb, s, region;
  int64_t voff, poff;       // offsets of the row's voxels and pixels in the gathered / prepared buffers
  double phi;
};

// a value the whole wavefront holds alike (one row per wavefront), kept in scalar registers
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float uni_f(float v) { return __builtin_bit_cast(float, uni(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ double uni_d(double v) { return __builtin_bit_cast(double, uni64(__builtin_bit_cast(int64_t, v))); }

__device__ __forceinline__ Box box_of(const xpg_refl &R, int r) {
  Box b;
  b.r = r;
  b.Dz = uni(R.rdz ? R.rdz[r] : R.Dz);
  b.Dy = uni(R.rdy ? R.rdy[r] : R.Dy);
  b.Dx = uni(R.rdx ? R.rdx[r] : R.Dx);
  b.P = b.Dy * b.Dx;
  b.V = b.Dz * b.P;
  b.voff = uni64(R.voff ? R.voff[r] : (int64_t)r * b.V);
  b.poff = uni64(R.poff ? R.poff[r] : (int64_t)r * b.P);
  b.ix0 = uni(R.ix0[r]);
  b.iy0 = uni(R.iy0[r]);
  b.iz0 = uni(R.iz0[r]);
  b.fb = uni(R.fb[r]);
  b.s = uni(R.pxs ? R.pxs[r] : 1);
  b.region = uni(R.region ? R.region[r] : 0);
  b.phi = uni_d(R.phi[r]);
  return b;
}

// value code of a raw frame word whose pixel lies on the detector, m: bit 0 untrusted (source mask), bit 1 outside the
// integrator's mask (the same rules as frame_code)
__device__ __forceinline__ float code_raw(const xpg_frames &F, uint32_t raw, uint8_t m) {
  float v;
  if (F.dtype == 1) {
    v = __builtin_bit_cast(float, raw);
    if (m & 1) v = -1.0f;
  } else {
    const bool bad = raw < F.min_valid || (F.use_flag_value && raw == F.flag_value) || (m & 1);
    v = (float)raw;
    if (F.overload && raw >= F.overload) v = -2.0f;
    if (bad) v = -1.0f;
  }
  if (m & 2) v = fminf(v, -1.0f);
  return v;
}

// per-pixel row index (prow), then the shoebox of reflection r into vals (from sb, or extracted from the frame store
// when sb == nullptr), flags cleared.  tmp (8 P bytes) and ftmp (4 Dz bytes): LDS free during staging.
//
// Unbinned boxes of single pixels take the streaming path: the per-pixel detector offsets and mask bits and the
// per-frame store indices are resolved once, then each lane issues NB frame loads before it consumes any, so a
// wavefront waits for memory about V / (64 NB) times instead of V / 64.
__device__ __forceinline__ void stage_values(const Box &b, const float *sb, const xpg_frames &F, float *vals,
                                             uint8_t *vflag, uint8_t *prow, unsigned char *tmp, unsigned char *ftmp,
                                             int lane) {
  for (int p = lane; p < b.P; p += WAVE) prow[p] = (uint8_t)(p / b.Dx);
  wave_sync();
  if (sb) {
    const float *src = sb + b.voff;
    constexpr int NB = 16;
    for (int v0 = lane; v0 < b.V && vals != src; v0 += NB * WAVE) {
      float t[NB];
#pragma unroll
      for (int u = 0; u < NB; ++u) t[u] = v0 + u * WAVE < b.V ? src[v0 + u * WAVE] : 0.f;
#pragma unroll
      for (int u = 0; u < NB; ++u)
        if (v0 + u * WAVE < b.V) vals[v0 + u * WAVE] = t[u];
    }
  } else if (b.fb == 1 && b.s == 1) {
    int32_t *poff = (int32_t *)tmp;                 // P: y nx + x, or -1 off the detector
    uint8_t *pm = tmp + 4 * (size_t)b.P;            // P: mask bits
    int32_t *ft = (int32_t *)ftmp;                  // Dz: store frame, or -1 (not read)
    for (int p = lane; p < b.P; p += WAVE) {
      const int y = b.iy0 + prow[p], x = b.ix0 + (p - prow[p] * b.Dx);
      int32_t off = -1;
      uint8_t m = 0;
      if (y >= 0 && x >= 0 && y < F.ny && x < F.nx) {
        off = y * F.nx + x;
        if (F.src_mask && !F.src_mask[off]) m |= 1;
        if (F.ceil_mask && !F.ceil_mask[off]) m |= 2;
      }
      poff[p] = off;
      pm[p] = m;
    }
    for (int j = lane; j < b.Dz; j += WAVE) {
      int64_t t;
      ft[j] = frame_read(F, b.iz0 + j, t) ? (int32_t)t : -1;
    }
    wave_sync();
    const uint32_t *src = (const uint32_t *)F.data;
    constexpr int NB = 16;
    VoxIter it(lane, b.P, b.V);
    while (it.ok()) {
      uint32_t raw[NB];
      int vv[NB], pp[NB], kind[NB];            // kind: 0 loaded, 1 frame not read (-3), 2 off the detector (-1), 3 none
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        vv[u] = it.v;
        pp[u] = it.p;
        raw[u] = 0;
        kind[u] = 3;
        if (it.ok()) {
          const int32_t t = ft[it.j], off = poff[it.p];
          if (t < 0) {
            kind[u] = 1;
          } else if (off < 0) {
            kind[u] = 2;
          } else {
            kind[u] = 0;
            raw[u] = src[(int64_t)t * F.frame_stride + off];
          }
          it.next();
        }
      }
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        if (kind[u] == 3) continue;
        vals[vv[u]] = kind[u] == 1 ? -3.0f : (kind[u] == 2 ? -1.0f : code_raw(F, raw[u], pm[pp[u]]));
      }
    }
  } else {
    for (VoxIter it(lane, b.P, b.V); it.ok(); it.next()) {
      const int iy = prow[it.p], ix = it.p - iy * b.Dx;
      vals[it.v] = voxel_value(F, b.iz0, b.fb, it.j, b.iy0 + iy, b.ix0 + ix, b.s);
    }
  }
  for (int v = lane; v < b.V; v += WAVE) vflag[v] = 0;
  wave_sync();
}

// shoeboxes of a chunk from the frame store: one wavefront per reflection stages its box in LDS (the same reads as
// the passes' fused path) and writes it out contiguously
__host__ __device__ static inline size_t extract_wave_bytes(int Dz, int Dy, int Dx) {
  const size_t P = (size_t)Dy * Dx, V = (size_t)Dz * P;
  return al8(al8(4 * V) + al8(V) + al8(P) + 8 * P + 4 * (size_t)Dz);
}

__global__ void __launch_bounds__(256) k_extract(xpg_frames F, xpg_refl R, float *sb, int wave_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);   // wave-uniform: row data in SGPRs
  const int r = blockIdx.x * (blockDim.x / WAVE) + wid;
  if (r >= R.n) return;
  const Box b = box_of(R, r);
  const size_t P = (size_t)b.P, V = (size_t)b.V;
  if (extract_wave_bytes(b.Dz, b.Dy, b.Dx) > (size_t)wave_bytes) return;
  unsigned char *base = smem + (size_t)wid * wave_bytes;
  float *vals = (float *)base;
  uint8_t *vflag = base + al8(4 * V);
  uint8_t *prow = vflag + al8(V);
  unsigned char *tmp = prow + al8(P);
  unsigned char *ftmp = tmp + 8 * P;
  stage_values(b, nullptr, F, vals, vflag, prow, tmp, ftmp, lane);
  float *dst = sb + b.voff;
  for (int v = lane; v < b.V; v += WAVE) dst[v] = vals[v];
}

// prepared data of reflection r: shoebox values, neighbour-mask bit `bit` as VF_NBM, eps1 / eps2 / J into U1, U2, Jd;
// returns zeta
// vox = false: the voxel values and mask bits stay in global memory (rows whose voxels are read into registers)
__device__ __forceinline__ double stage_prepared(const Box &b, const xpg_prepared &pp, int bit, float *vals,
                                                 uint8_t *vflag, uint8_t *prow, double *U1, double *U2, double *Jd,
                                                 int lane, bool vox = true) {
  for (int p = lane; p < b.P; p += WAVE) prow[p] = (uint8_t)(p / b.Dx);
  const float *src = pp.sb + b.voff;
  const uint8_t *nsrc = pp.nbm + b.voff;
  constexpr int NB = 16;
  for (int v0 = lane; v0 < b.V && vox; v0 += NB * WAVE) {
    float t[NB];
    uint8_t m[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const bool in = v0 + u * WAVE < b.V;
      t[u] = in ? src[v0 + u * WAVE] : 0.f;
      m[u] = in ? nsrc[v0 + u * WAVE] : 0;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u)
      if (v0 + u * WAVE < b.V) {
        if (vals != src) vals[v0 + u * WAVE] = t[u];
        vflag[v0 + u * WAVE] = ((m[u] >> bit) & 1) ? 1 : 0;     // VF_NBM
      }
  }
  const double *g = pp.geo + 3 * b.poff;
  for (int p = lane; p < b.P; p += WAVE) {
    U1[p] = g[p];
    U2[p] = g[b.P + p];
    Jd[p] = g[2 * b.P + p];
  }
  const double zeta = pp.zeta[b.r];
  wave_sync();
  return zeta;
}

// a row's rocking curve computed ahead by xpg_rocking, into Rd (LDS)
__device__ __forceinline__ void load_rocking(const xpg_prepared &pp, int r, int Dz, double *Rd, int lane) {
  for (int j = lane; j < Dz; j += WAVE) Rd[j] = pp.R[(int64_t)r * pp.r_stride + j];
  wave_sync();
}

// rocking curves of a launch's rows ahead of a pass: one lane per (row, frame)
__global__ void __launch_bounds__(256) k_rocking(xpg_call call, xpg_refl R, xpg_prepared pp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = pp.r_stride;
  const int64_t r = t / stride;
  const int j = (int)(t - r * stride);
  if (r >= R.n) return;
  const int Dz = R.rdz ? R.rdz[r] : R.Dz;
  if (j >= Dz) return;
  pp.R[t] = rocking_frame(call, R.iz0[r], R.fb[r], R.phi[r], pp.zeta[r], j);
}

// box coordinates: co[0..Dx) pixel-centre x, co[Dx..Dx+Dy) y, then the Dz frame centres iz0 + fb (j + 1/2)
__device__ __forceinline__ void box_coords(const Box &b, double *co, int lane) {
  for (int i = lane; i < b.Dx + b.Dy + b.Dz; i += WAVE) {
    double v;
    if (i < b.Dx) v = pix_x(b.ix0, i, b.s);
    else if (i < b.Dx + b.Dy) v = pix_x(b.iy0, i - b.Dx, b.s);
    else v = (double)b.iz0 + (double)b.fb * ((double)(i - b.Dx - b.Dy) + 0.5);
    co[i] = v;
  }
}

// background template at the pixel centres (xpg/integrate_kernels.py:68-75; numpy round = half to even = rint)
__device__ __forceinline__ float template_at(const xpg_detector &det, const double *co, int Dx, int p) {
  if (!det.bkg) return 1.0f;
  int iy = p / Dx, ix = p - iy * Dx;
  int kx = min(max((int)rint(co[ix] - 1.0), 0), det.bkg_nx - 1);
  int ky = min(max((int)rint(co[Dx + iy] - 1.0), 0), det.bkg_ny - 1);
  return det.bkg[(int64_t)ky * det.bkg_nx + kx];
}

__device__ __forceinline__ double frames_fg(const double *Rd, int Dz, uint8_t *fflag, int lane) {
  double Rmax = -INFINITY;
  for (int j = lane; j < Dz; j += WAVE) Rmax = fmax(Rmax, Rd[j]);
  Rmax = wmax_d(Rmax);
  const double thr = 1e-3 * Rmax + 1e-6;
  for (int j = lane; j < Dz; j += WAVE) fflag[j] = (Rd[j] >= thr) ? 1 : 0;
  return Rmax;
}

// ---------------------------------------------------------------------------------------------------------------------
// prepare: the widths-independent part of both passes, ahead of the adaptive chain
// ---------------------------------------------------------------------------------------------------------------------

struct PrepLayout {
  size_t vals, vflag, prow, tmp, ftmp, U1, U2, Jd, co, tab, bytes;
};
// big: values and mask bits are written in place into the output buffers (boxes too large for LDS)
__host__ __device__ static inline PrepLayout prep_layout(int Dz, int Dy, int Dx, bool big = false) {
  const size_t P = (size_t)Dy * Dx, V = (size_t)Dz * P, C = (size_t)Dx + Dy + Dz;
  PrepLayout L;
  size_t o = 0;
  L.vals = o; if (!big) o = al8(o + 4 * V);
  L.vflag = o; if (!big) o = al8(o + V);
  L.prow = o; o = al8(o + P);
  L.tmp = o; o += 8 * P;
  L.ftmp = o; o = al8(o + 4 * (size_t)Dz);
  L.U1 = o; o += 8 * P;
  L.U2 = o; o += 8 * P;
  L.Jd = o; o += 8 * P;
  L.co = o; o += 8 * C;
  L.tab = o; o += 8 * C;
  L.bytes = al8(o);
  return L;
}

// sb: the shoeboxes from a chunk buffer instead of the frame store (NULL: from F)
// at least 4 wavefronts per SIMD (the register budget of 128 VGPRs; a few bytes spill): 8% faster than the 3 that
// its natural 145 VGPRs allow
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_prepare(xpg_detector det, xpg_call call, xpg_refl R, xpg_frames F,
                                                 const float *sb, xpg_prepared pp, int wave_bytes, int big) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);   // wave-uniform: row data in SGPRs
  const int r = blockIdx.x * (blockDim.x / WAVE) + wid;
  if (r >= R.n) return;
  const Box b = box_of(R, r);
  const PrepLayout L = prep_layout(b.Dz, b.Dy, b.Dx, big != 0);
  if (L.bytes > (size_t)wave_bytes) return;
  unsigned char *base = smem + (size_t)wid * wave_bytes;
  float *vals = big ? pp.sb + b.voff : (float *)(base + L.vals);
  uint8_t *vflag = big ? pp.nbm + b.voff : base + L.vflag;
  uint8_t *prow = base + L.prow;
  double *U1 = (double *)(base + L.U1), *U2 = (double *)(base + L.U2), *Jd = (double *)(base + L.Jd);
  double *co = (double *)(base + L.co), *tab = (double *)(base + L.tab);
  const int P = b.P, V = b.V, Dz = b.Dz, Dy = b.Dy, Dx = b.Dx;
  stage_values(b, sb, F, vals, vflag, prow, base + L.tmp, base + L.ftmp, lane);
  Frame f;
  kabsch(R.S + 3 * (int64_t)r, call.s0, call.axis, f);
  box_eps_J(det, f, b.ix0, b.iy0, b.s, Dy, Dx, U1, U2, Jd, lane);
  box_coords(b, co, lane);
  wave_sync();
  neighbour_mask(R, r, co, co + Dx, co + Dx + Dy, Dz, Dy, Dx, 0, vflag, 1, tab, prow, lane);
  neighbour_mask(R, r, co, co + Dx, co + Dx + Dy, Dz, Dy, Dx, 1, vflag, 2, tab, prow, lane);
  float *sdst = pp.sb + b.voff;
  uint8_t *ndst = pp.nbm + b.voff;
  for (int v = lane; v < V && !big; v += WAVE) {
    sdst[v] = vals[v];
    ndst[v] = vflag[v];
  }
  double *g = pp.geo + 3 * b.poff;
  for (int p = lane; p < P; p += WAVE) {
    g[p] = U1[p];
    g[P + p] = U2[p];
    g[2 * P + p] = Jd[p];
  }
  if (lane == 0) pp.zeta[r] = f.zeta;
  if (R.has_p && (double)uni_f(R.has_p[r]) > 0.5) {
    // the overlapping partner's frame and eps, J on this box's pixels (pass 2 models its profile here)
    Frame fp;
    kabsch(R.pS + 3 * (int64_t)r, call.s0, call.axis, fp);
    box_eps_J(det, fp, b.ix0, b.iy0, b.s, Dy, Dx, U1, U2, Jd, lane);
    double *gp = pp.geo_p + 3 * b.poff;
    for (int p = lane; p < P; p += WAVE) {
      gp[p] = U1[p];
      gp[P + p] = U2[p];
      gp[2 * P + p] = Jd[p];
    }
    if (lane == 0) pp.zeta_p[r] = fp.zeta;
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// pass 1 (xpg/integrate_kernels.py:287-365)
// ---------------------------------------------------------------------------------------------------------------------

struct P1Args {
  xpg_detector det;
  xpg_call call;
  xpg_refl R;
  xpg_frames F;
  const float *sb;
  xpg_prepared prep;
  int use_prep;
  int vals_global;          // shoebox values read in place (sb / prep.sb), not staged in LDS
  xpg_pass1_out out;
  xpg_profile_acc acc;
  int use_acc;
  int waves;
  int wave_bytes;
};

__device__ __forceinline__ void splat(int64_t *grid, int idx, double v) {
  long long q = llrint(v * FX_SCALE);
  if (q != 0) atomicAdd((unsigned long long *)(grid + idx), (unsigned long long)q);
}

// the grid copy of this workgroup (integer sums: which copy takes an increment does not change the totals)
__device__ __forceinline__ int acc_copy(const xpg_profile_acc &acc) {
  return acc.copies > 1 ? (int)(blockIdx.x % (unsigned)acc.copies) : 0;
}
__device__ __forceinline__ void splat4(const xpg_profile_acc &acc, int region, double u1, double u2, double dens) {
  Bil q = bilinear(u1, u2);
  if (!q.inside) return;
  region = min(max(region, 0), acc.n_regions - 1);   // memory safety; the host validates region
  const int i00 = (acc_copy(acc) * acc.n_regions + region) * G * G + q.i2 * G + q.i1;
  const double w00 = (1 - q.t1) * (1 - q.t2), w10 = q.t1 * (1 - q.t2), w01 = (1 - q.t1) * q.t2, w11 = q.t1 * q.t2;
  splat(acc.num, i00, w00 * dens);
  splat(acc.den, i00, w00);
  splat(acc.num, i00 + 1, w10 * dens);
  splat(acc.den, i00 + 1, w10);
  splat(acc.num, i00 + G, w01 * dens);
  splat(acc.den, i00 + G, w01);
  splat(acc.num, i00 + G + 1, w11 * dens);
  splat(acc.den, i00 + G + 1, w11);
}

// PREP: the prepared-data form (xpg_pass1_prepared), compiled separately so that it carries none of the fused
// form's register pressure; VG: shoebox values read in place from global memory (boxes too large for LDS), compiled
// separately so that the LDS form's value accesses stay LDS instructions
// The prepared register-row form (the batched sweep's) runs at 6 wavefronts per SIMD (80 VGPRs, a few bytes spilled;
// 4% faster than 5); the others keep their natural budget (they would spill much more).
template <bool PREP, bool VG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PREP && !VG ? 6 : 1))) k_pass1(P1Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);   // wave-uniform: row data in SGPRs
  const int r = blockIdx.x * a.waves + wid;
  if (r >= a.R.n) return;
  const Box b = box_of(a.R, r);
  const Layout L = make_layout(b.Dz, b.Dy, b.Dx, 1, !VG,
                               !VG && compact_rows(b.Dz, b.Dy, b.Dx, PREP, PREP && a.prep.R != nullptr));
  if (L.bytes > (size_t)a.wave_bytes) return;
  unsigned char *base = smem + (size_t)wid * a.wave_bytes;
  const int P = b.P;
  float *vals = VG ? const_cast<float *>((PREP ? a.prep.sb : a.sb) + b.voff) : (float *)(base + L.vals);
  uint8_t *vflag = base + L.vflag;
  double *scr = (double *)(base + L.scr);
  double *U1 = (double *)(base + L.U1), *U2 = (double *)(base + L.U2), *Jd = (double *)(base + L.Jd);
  float *bref = (float *)(base + L.pa), *thr = bref + P;
  uint8_t *prow = base + L.prow, *pflag = base + L.pflag;
  double *Rd = (double *)(base + L.R);
  double *zzd = (double *)(base + L.fr), *e3d = zzd + b.Dz;
  uint8_t *fflag = (uint8_t *)(e3d + b.Dz);
  double *co = (double *)(base + L.co), *tab = (double *)(base + L.tab);

  const int Dz = b.Dz, Dy = b.Dy, Dx = b.Dx;
  const xpg_detector &det = a.det;
  const xpg_call &call = a.call;

  // Boxes of at most NPL pixels x DZM frames per lane (regs) read their voxels once, into registers (value,
  // neighbour bit): from the prepared buffers directly, or from LDS after staging.
  constexpr int NPL = 2, DZM = 8, NK = NPL * DZM;
  const bool regs = P <= NPL * WAVE && Dz <= DZM;
  const bool direct = PREP && !VG && regs;
  double zeta;
  if constexpr (PREP) {
    zeta = stage_prepared(b, a.prep, 0, vals, vflag, prow, U1, U2, Jd, lane, !direct);
  } else {
    stage_values(b, a.sb, a.F, vals, vflag, prow, base + L.pa, base + L.R, lane);
    Frame f;
    kabsch(a.R.S + 3 * (int64_t)r, call.s0, call.axis, f);
    zeta = f.zeta;
    box_eps_J(det, f, b.ix0, b.iy0, b.s, Dy, Dx, U1, U2, Jd, lane);
  }
  float dv[NK];                          // direct rows: values and neighbour bytes in flight during the setup below
  uint8_t dn[NK];
  if (direct) {
    const float *src = a.prep.sb + b.voff;
    const uint8_t *nsrc = a.prep.nbm + b.voff;
#pragma unroll
    for (int q = 0; q < NPL; ++q)
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int v = min(j, Dz - 1) * P + min(lane + q * WAVE, P - 1);   // in the row, whatever the lane
        dv[q * DZM + j] = src[v];
        dn[q * DZM + j] = nsrc[v];
      }
  }
  if (PREP && a.prep.R) load_rocking(a.prep, r, Dz, Rd, lane);
  else rocking_curve(call, b.iz0, b.fb, b.phi, zeta, Dz, scr, Rd, lane);
  const double seff = (PREP && a.prep.seff) ? box_u_given(P, uni_d(a.prep.seff[r]), U1, U2, lane)
                                            : box_u(Dy, Dx, sd_of(call), U1, U2, Jd, lane);
  box_coords(b, co, lane);
  wave_sync();
  for (int p = lane; p < P; p += WAVE) {
    const double u1 = U1[p], u2 = U2[p];
    const double r2 = u1 * u1 + u2 * u2;
    pflag[p] = (r2 <= 9.0 ? PF_FG2 : 0) | (r2 >= 12.25 ? PF_BGR : 0);
    bref[p] = template_at(det, co, Dx, p);
  }
  frames_fg(Rd, Dz, fflag, lane);
  for (int j = lane; j < Dz; j += WAVE) {
    double lo, hi;
    frame_bounds(call, b.iz0, b.fb, j, lo, hi);
    e3d[j] = zeta * (0.5 * (lo + hi) - b.phi);
    zzd[j] = co[Dx + Dy + j];
  }
  wave_sync();
  if constexpr (!PREP) neighbour_mask(a.R, r, co, co + Dx, co + Dx + Dy, Dz, Dy, Dx, 0, vflag, VF_NBM, tab, prow, lane);

  // Lanes run over pixels (p = lane, lane + 64, ...) and each lane over the frames of its pixel: per-pixel terms stay
  // in registers, per-frame terms are broadcasts.  Sums are float64; sums of counts and of template values are exact.
  //
  // Register rows keep the per-voxel classes as bit masks (bit k = q DZM + j for pixel lane + 64 q, frame j); the
  // sums run over the same voxels in the same order as the LDS loops of larger boxes.
  float cv[NK], brq[NPL];
  uint32_t inm = 0, nbmm = 0, fgm = 0, fgvm = 0, bgmm = 0, fg2q = 0;   // in box / nbm / fg / fgv / bg / bit q: fg2
  uint32_t fgz = 0;                                                   // bit j: foreground frame (uniform)
  if (regs) {
#pragma unroll
    for (int j = 0; j < DZM; ++j)
      if (j < Dz && fflag[j]) fgz |= 1u << j;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int p = lane + q * WAVE;
      const bool pin = p < P;
      const uint8_t pfl = pin ? pflag[p] : 0;
      brq[q] = pin ? bref[p] : 0.f;
      if (pfl & PF_FG2) fg2q |= 1u << q;
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        cv[k] = 0.f;
        if (pin && j < Dz) {
          const int v = j * P + p;
          const float c = direct ? dv[k] : vals[v];
          const bool nbm = direct ? (dn[k] & 1) != 0 : (vflag[v] & VF_NBM) != 0;
          cv[k] = c;
          inm |= 1u << k;
          if (nbm) nbmm |= 1u << k;
          const bool fg = (pfl & PF_FG2) && ((fgz >> j) & 1u);
          const bool vn = c >= 0.f && !nbm;
          if (fg) fgm |= 1u << k;
          if (fg && vn) fgvm |= 1u << k;
          if (!fg && vn && (pfl & PF_BGR)) bgmm |= 1u << k;
        }
      }
    }
  }

  // per-voxel classes: fg = fg2 & fgz; fgv = fg & valid & ~nbm; background candidates (xpg/integrate_kernels.py:291-305),
  // with the unclipped background sums of the first round
  double sc0 = 0.0, sT0 = 0.0;
  if (regs) {
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      if (lane + q * WAVE >= P) continue;
      double cs = 0.0;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if ((bgmm >> k) & 1u) {
          cs += (double)cv[k];
          ++cnt;
        }
      }
      sc0 += cs;
      sT0 += (double)cnt * (double)brq[q];
    }
  } else {
    for (int p0 = 0; p0 < P; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const uint8_t pfl = pflag[p];
        const bool fg2 = pfl & PF_FG2, bgr = pfl & PF_BGR;
        double cs = 0.0;
        int cnt = 0;
        for (int j0 = 0; j0 < Dz; j0 += LB) {       // LB frames' reads in flight before any write
          float cc[LB];
          uint8_t ff[LB], fz_[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
            fz_[u] = fflag[j];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            if (j0 + u >= Dz) break;
            const float c = cc[u];
            uint8_t fl = ff[u];
            const bool valid = c >= 0.f, nbm = fl & VF_NBM;
            const bool fg = fg2 && fz_[u];
            if (fg) fl |= VF_FG;
            if (fg && valid && !nbm) fl |= VF_FGV;
            if (!fg && valid && !nbm && bgr) {
              fl |= VF_BGM;
              cs += (double)c;
              ++cnt;
            }
            vflag[(j0 + u) * P + p] = fl;
          }
        }
        sc0 += cs;
        sT0 += (double)cnt * (double)bref[p];
      }
    }
    wave_sync();
  }

  // ---- robust background b = sum(w c) / max(sum(w T), 1e-6) with four Poisson-aware clipping rounds ----
  // A round that leaves b unchanged (bit for bit) sets the same thresholds again, so every later round would repeat
  // it: the loop stops there.
  double s2[2] = {sc0, sT0};
  wsum_multi(s2);
  double Sc = s2[0], nbg = s2[1];
  double bk = Sc / fmax(nbg, 1e-6);
  double bk_thr = bk;                   // the level whose thresholds made the final background set
  for (int round = 1; round < 5 && regs; ++round) {
    bk_thr = bk;
    double sc = 0.0, sT = 0.0;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      if (lane + q * WAVE >= P) continue;
      const double bp = bk * (double)brq[q];
      const float th = __double2float_rd((bp + 3.0 * sqrt(call.gain * fmax(bp, 0.5))) + 1.0);
      double cs = 0.0;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if (((bgmm >> k) & 1u) && cv[k] <= th) {
          cs += (double)cv[k];
          ++cnt;
        }
      }
      sc += cs;
      sT += (double)cnt * (double)brq[q];
    }
    double r2[2] = {sc, sT};
    wsum_multi(r2);
    Sc = r2[0];
    nbg = r2[1];
    const double bk_prev = bk;
    bk = Sc / fmax(nbg, 1e-6);
    if (bk == bk_prev) break;
  }
  for (int round = 1; round < 5 && !regs; ++round) {
    bk_thr = bk;
    for (int p = lane; p < P; p += WAVE) {
      const double bp = bk * (double)bref[p];
      const double t = (bp + 3.0 * sqrt(call.gain * fmax(bp, 0.5))) + 1.0;
      thr[p] = __double2float_rd(t);     // c <= t  <=>  c <= rd(t) for float32 c
    }
    wave_sync();
    double sc = 0.0, sT = 0.0;
    for (int p0 = 0; p0 < P; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const float th = thr[p];
        double cs = 0.0;
        int cnt = 0;
        for (int j0 = 0; j0 < Dz; j0 += LB) {
          float cc[LB];
          uint8_t ff[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            if (j0 + u < Dz && (ff[u] & VF_BGM) && cc[u] <= th) {
              cs += (double)cc[u];
              ++cnt;
            }
          }
        }
        sc += cs;
        sT += (double)cnt * (double)bref[p];
      }
    }
    double r2[2] = {sc, sT};
    wsum_multi(r2);
    Sc = r2[0];
    nbg = r2[1];
    const double bk_prev = bk;
    bk = Sc / fmax(nbg, 1e-6);
    if (bk == bk_prev) break;           // the next threshold equals this one: uniform over the wavefront
  }

  // ---- per-voxel classes for mask comparisons (only when asked for) ----
  if (a.out.vmask) {
    uint8_t *vm = a.out.vmask + b.voff;
    if (regs) {
#pragma unroll
      for (int q = 0; q < NPL; ++q) {
        const int p = lane + q * WAVE;
        if (p >= P) continue;
        const double bp = bk_thr * (double)brq[q];
        const float th = __double2float_rd((bp + 3.0 * sqrt(call.gain * fmax(bp, 0.5))) + 1.0);
#pragma unroll
        for (int j = 0; j < DZM; ++j) {
          const int k = q * DZM + j;
          if (j >= Dz) continue;
          const bool bgm = (bgmm >> k) & 1u;
          vm[j * P + p] = (((fgm >> k) & 1u) ? XPG_VM_FG : 0) | (bgm ? XPG_VM_BGM : 0) |
                          (((nbmm >> k) & 1u) ? XPG_VM_NBM : 0) | ((bgm && cv[k] <= th) ? XPG_VM_BGW : 0) |
                          (((fgvm >> k) & 1u) ? XPG_VM_FGV : 0);
        }
      }
    } else {
      // thr[] holds the thresholds of the last clipping round (those of bk_thr)
      for (int p = lane; p < P; p += WAVE) {
        const float th = thr[p];
        for (int j = 0; j < Dz; ++j) {
          const uint8_t fl = vflag[j * P + p];
          const bool bgm = fl & VF_BGM;
          vm[j * P + p] = ((fl & VF_FG) ? XPG_VM_FG : 0) | (bgm ? XPG_VM_BGM : 0) | ((fl & VF_NBM) ? XPG_VM_NBM : 0) |
                          ((bgm && vals[j * P + p] <= th) ? XPG_VM_BGW : 0) | ((fl & VF_FGV) ? XPG_VM_FGV : 0);
        }
      }
    }
  }

  // ---- sums over the foreground ----
  double Sfgc = 0.0, sfb = 0.0, Swpos = 0.0, Swx = 0.0, Swy = 0.0, Swz = 0.0;
  double Sm1 = 0.0, Sm2 = 0.0, Smdd = 0.0, Sm3 = 0.0, Sm33 = 0.0;
  int n_fg = 0, n_over = 0, n_fgm = 0, n_fgnb = 0;
  auto fg_pixel = [&](int p, double cs, double npos, double nzz, double nall, double n3, double n33, int cnt) {
    const int iy = prow[p], ix = p - iy * Dx;
    const double a1 = U1[p] * seff, a2 = U2[p] * seff;
    n_fg += cnt;
    Sfgc += cs;
    sfb += (double)cnt * (double)bref[p];
    Swpos += npos;
    Swx += npos * co[ix];
    Swy += npos * co[Dx + iy];
    Swz += nzz;
    Sm1 += nall * a1;
    Sm2 += nall * a2;
    Smdd += nall * (a1 * a1 + a2 * a2);
    Sm3 += n3;
    Sm33 += n33;
  };
  if (regs) {
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int p = lane + q * WAVE;
      if (p >= P) continue;
      const double bT = bk * (double)brq[q];
      double cs = 0.0, npos = 0.0, nzz = 0.0, nall = 0.0, n3 = 0.0, n33 = 0.0;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if (!((fgm >> k) & 1u)) continue;
        const float c = cv[k];
        n_over += (c == -2.0f);
        n_fgm += (c < 0.f && c > -2.5f);
        n_fgnb += (nbmm >> k) & 1u;
        if (!((fgvm >> k) & 1u)) continue;
        const double net = (double)c - bT;
        ++cnt;
        cs += (double)c;
        if (net > 0.0) {
          npos += net;
          nzz += net * zzd[j];
        }
        const double e3 = e3d[j];
        nall += net;
        n3 += net * e3;
        n33 += net * (e3 * e3);
      }
      fg_pixel(p, cs, npos, nzz, nall, n3, n33, cnt);
    }
  } else {
    for (int p0 = 0; p0 < P; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const double bT = bk * (double)bref[p];
        double cs = 0.0, npos = 0.0, nzz = 0.0, nall = 0.0, n3 = 0.0, n33 = 0.0;
        int cnt = 0;
        for (int j0 = 0; j0 < Dz; j0 += LB) {
          float cc[LB];
          uint8_t ff[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = j0 + u;
            if (j >= Dz) break;
            const uint8_t fl = ff[u];
            if (!(fl & VF_FG)) continue;
            const float c = cc[u];
            n_over += (c == -2.0f);
            n_fgm += (c < 0.f && c > -2.5f);
            n_fgnb += (fl & VF_NBM) ? 1 : 0;
            if (!(fl & VF_FGV)) continue;
            const double net = (double)c - bT;
            ++cnt;
            cs += (double)c;
            if (net > 0.0) {
              npos += net;
              nzz += net * zzd[j];
            }
            const double e3 = e3d[j];
            nall += net;
            n3 += net * e3;
            n33 += net * (e3 * e3);
          }
        }
        fg_pixel(p, cs, npos, nzz, nall, n3, n33, cnt);
      }
    }
  }
  {                                      // the counts are exact as float64
    double sv[16] = {Sfgc, sfb, Swpos, Swx, Swy, Swz, Sm1, Sm2, Smdd, Sm3, Sm33,
                     (double)n_fg, (double)n_over, (double)n_fgm, (double)n_fgnb, 0.0};
    wsum_multi(sv);
    Sfgc = sv[0];
    sfb = sv[1];
    Swpos = sv[2];
    Swx = sv[3];
    Swy = sv[4];
    Swz = sv[5];
    Sm1 = sv[6];
    Sm2 = sv[7];
    Smdd = sv[8];
    Sm3 = sv[9];
    Sm33 = sv[10];
    n_fg = (int)sv[11];
    n_over = (int)sv[12];
    n_fgm = (int)sv[13];
    n_fgnb = (int)sv[14];
  }

  const double gain = call.gain;
  const double isum = Sfgc - bk * sfb;
  const double vb = ((sfb * sfb) * gain * fmax(bk, VAR_FLOOR)) / fmax(nbg, 1.0);
  const double vsum = gain * Sfgc + vb;
  const double vsum0 = gain * (bk * sfb) + vb;
  const double sw = fmax(Swpos, 1e-6);
  const double sws = fmax(isum, 1e-6);
  const double mu1 = Sm1 / sws, mu2 = Sm2 / sws;
  const double m_d = (Smdd / sws - mu1 * mu1 - mu2 * mu2) / 2.0;
  const double mu3 = Sm3 / sws;
  const double m_m = Sm33 / sws - mu3 * mu3;

  // ---- per frame: fz (net counts in the foreground ellipse), recorded frames -> frac_rec ----
  const int fzs = a.out.fz_stride > 0 ? a.out.fz_stride : Dz;
  float *fz_out = a.out.fz ? a.out.fz + (int64_t)r * fzs : nullptr;
  if (fz_out)
    for (int j = Dz + lane; j < fzs; j += WAVE) fz_out[j] = 0.f;
  if (regs) {
    // each lane's pixels per frame, then one sum over the wavefront per frame
    double fzp[DZM];
#pragma unroll
    for (int j = 0; j < DZM; ++j) fzp[j] = 0.0;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      if (!((fg2q >> q) & 1u)) continue;
      const double bpx = bk * (double)brq[q];
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if (((inm >> k) & 1u) && cv[k] >= 0.f && !((nbmm >> k) & 1u)) fzp[j] += (double)cv[k] - bpx;
      }
    }
    wsum_multi(fzp);
#pragma unroll
    for (int j = 0; j < DZM; ++j) {
      if (j < Dz) {
        const double t = fzp[j];
        bool rec = false;
#pragma unroll
        for (int q = 0; q < NPL; ++q) rec |= ((inm >> (q * DZM + j)) & 1u) && cv[q * DZM + j] > -2.5f;
        const bool anyrec = __ballot(rec) != 0;
        if (lane == 0) {
          if (fz_out) fz_out[j] = (float)t;
          scr[j] = anyrec ? Rd[j] : 0.0;
        }
      }
    }
  } else if (Dz <= 16) {
    // every frame at once: g lanes per frame (the largest of 16, 8, 4 with g Dz <= 64) stride over its pixels, then
    // a sum within each group of g lanes
    const int g = Dz <= 4 ? 16 : (Dz <= 8 ? 8 : 4);
    const int j = lane / g, q = lane - j * g;
    double part = 0.0;
    bool rec = false;
    if (j < Dz) {
      for (int p0 = q; p0 < P; p0 += 4 * g) {      // four pixels' reads in flight together, used in pixel order
        float cc[4], br[4];
        uint8_t pf[4], ff[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int p = min(p0 + u * g, P - 1);
          cc[u] = vals[j * P + p];
          ff[u] = vflag[j * P + p];
          pf[u] = pflag[p];
          br[u] = bref[p];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (p0 + u * g >= P) break;
          const float c = cc[u];
          rec |= c > -2.5f;
          if ((pf[u] & PF_FG2) && c >= 0.f && !(ff[u] & VF_NBM)) part += (double)c - bk * (double)br[u];
        }
      }
    }
    part += dppd<DPP_Q1>(part);
    part += dppd<DPP_Q2>(part);
    if (g >= 8) part += __shfl_xor(part, 4, WAVE);
    if (g >= 16) part += __shfl_xor(part, 8, WAVE);
    const uint64_t rb = __ballot(rec);
    if (q == 0 && j < Dz) {
      const bool anyrec = ((rb >> (j * g)) & ((1ull << g) - 1)) != 0;
      if (fz_out) fz_out[j] = (float)part;
      scr[j] = anyrec ? Rd[j] : 0.0;
    }
  } else {
    for (int j = lane; j < Dz; j += WAVE) {      // one lane per frame
      double fzj = 0.0;
      bool anyrec = false;
      for (int p = 0; p < P; ++p) {
        const float c = vals[j * P + p];
        anyrec |= c > -2.5f;
        if ((pflag[p] & PF_FG2) && c >= 0.f && !(vflag[j * P + p] & VF_NBM)) fzj += (double)c - bk * (double)bref[p];
      }
      if (fz_out) fz_out[j] = (float)fzj;
      scr[j] = anyrec ? Rd[j] : 0.0;
    }
  }
  wave_sync();
  double frac_rec = 0.0;
  for (int j = 0; j < Dz; ++j) frac_rec = frac_rec + scr[j];

  // ---- per pixel: projected image over the foreground frames, all-valid flag, fused profile accumulation ----
  const bool strong = a.use_acc && (isum > 0.0) && (isum / sqrt(fmax(vsum, 1e-6)) > 10.0) && (n_over == 0) &&
                      (frac_rec > 0.95) && (n_fgnb == 0) && (n_fgm == 0);
  const double seff2 = seff * seff;
  const bool pix_out = a.out.img2 || a.out.all_valid2 || a.out.J || a.out.u1 || a.out.u2;
  auto img_pixel = [&](int p, double img, bool okv) {
    const int64_t o = b.poff + p;
    if (a.out.img2) a.out.img2[o] = (float)img;
    if (a.out.all_valid2) a.out.all_valid2[o] = okv ? 1 : 0;
    if (a.out.J) a.out.J[o] = (float)Jd[p];
    if (a.out.u1) a.out.u1[o] = (float)U1[p];
    if (a.out.u2) a.out.u2[o] = (float)U2[p];
    if (strong && okv) {
      // the projected image in float64 (the standalone accumulation takes pass 1's float32 img2, as the reference's
      // accumulate_profiles does; fused, nothing forces the rounding and the grids stay at the float64 sums)
      const double dens = (img / fmax(isum, 1e-6)) / fmax(Jd[p] / seff2, 1e-12);
      splat4(a.acc, b.region, U1[p], U2[p], dens);
    }
  };
  if (regs && (strong || pix_out)) {
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int p = lane + q * WAVE;
      if (p >= P) continue;
      double img = 0.0;
      bool okv = true;
      const double bpx = bk * (double)brq[q];
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if (!((inm >> k) & 1u) || !((fgz >> j) & 1u)) continue;
        if (cv[k] >= 0.f && !((nbmm >> k) & 1u)) img += (double)cv[k] - bpx;
        else okv = false;
      }
      img_pixel(p, img, okv);
    }
  }
  for (int p = lane; p < P && !regs && (strong || pix_out); p += WAVE) {
    double img = 0.0;
    bool okv = true;
    const double bpx = bk * (double)bref[p];
    for (int j0 = 0; j0 < Dz; j0 += LB) {
      float cc[LB];
      uint8_t ff[LB], fz_[LB];
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        const int j = min(j0 + u, Dz - 1);
        cc[u] = vals[j * P + p];
        ff[u] = vflag[j * P + p];
        fz_[u] = fflag[j];
      }
#pragma unroll
      for (int u = 0; u < LB; ++u) {
        if (j0 + u >= Dz || !fz_[u]) continue;
        if (cc[u] >= 0.f && !(ff[u] & VF_NBM)) img += (double)cc[u] - bpx;
        else okv = false;
      }
    }
    img_pixel(p, img, okv);
  }

  if (lane == 0) {
    const xpg_pass1_out &o = a.out;
    o.bg[r] = bk;
    o.nbg[r] = nbg;
    o.isum[r] = isum;
    o.vsum[r] = vsum;
    o.vsum0[r] = vsum0;
    o.xo[r] = Swx / sw;
    o.yo[r] = Swy / sw;
    o.zo[r] = Swz / sw;
    o.m_d[r] = m_d;
    o.m_m[r] = m_m;
    o.frac_rec[r] = frac_rec;
    o.nfg[r] = (double)n_fg;
    o.zeta[r] = zeta;
    o.seff[r] = seff;
    o.n_over[r] = n_over;
    o.n_fg_masked[r] = n_fgm;
    o.n_fg_nb[r] = n_fgnb;
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// accumulate_profiles (xpg/integrate_kernels.py:367-380), standalone form
// ---------------------------------------------------------------------------------------------------------------------

__global__ void k_accumulate(const float *img2, const uint8_t *ok2, const float *J, const float *u1, const float *u2,
                             const int32_t *region, const double *isum, const uint8_t *sel, const double *seff, int n,
                             int P, xpg_profile_acc acc) {
  const int64_t total = (int64_t)n * P;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(g / P);
    if (!sel[r] || !ok2[g]) continue;
    const double s = seff[r];
    const double dens = ((double)img2[g] / fmax(isum[r], 1e-6)) / fmax((double)J[g] / (s * s), 1e-12);
    splat4(acc, region[r], (double)u1[g], (double)u2[g], dens);
  }
}

// profile normalisation in float64 (xpg/integrate.py:546-556): one workgroup of 256 per region
__global__ void __launch_bounds__(256) k_normalise(xpg_profile_acc acc, const float *gauss, float *prof, double *num64,
                                                   double *den64) {
  __shared__ double red[256];
  __shared__ double s_num[G * G], s_den[G * G];
  const int reg = blockIdx.x;
  const int GG = G * G;
  const double du = (2 * UMAX / (G - 1)) * (2 * UMAX / (G - 1));
  const int nc = acc.copies > 1 ? acc.copies : 1;
  const int64_t cs = (int64_t)acc.n_regions * GG;      // stride of the grid copies
  // the grids summed over their copies (integers: exact, order-free), read eight copies at a time
  for (int i = threadIdx.x; i < GG; i += blockDim.x) {
    long long tn = 0, td = 0;
    for (int c0 = 0; c0 < nc; c0 += 8) {
      long long vn[8], vd[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool in = c0 + u < nc;
        vn[u] = in ? acc.num[(c0 + u) * cs + reg * GG + i] : 0;
        vd[u] = in ? acc.den[(c0 + u) * cs + reg * GG + i] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        tn += vn[u];
        td += vd[u];
      }
    }
    s_num[i] = (double)tn / FX_SCALE;
    s_den[i] = (double)td / FX_SCALE;
  }
  __syncthreads();
  double part = 0.0;
  for (int i = threadIdx.x; i < GG; i += blockDim.x) {
    const double num = s_num[i], den = s_den[i];
    if (num64) num64[reg * GG + i] = num;
    if (den64) den64[reg * GG + i] = den;
    const double dens = den > 0 ? num / fmax(den, 1e-12) : 0.0;
    const double pr = (den * fmax(dens, 0.0) + 3.0 * (double)gauss[i]) / (den + 3.0);
    part += pr;
  }
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double tot = fmax(red[0] * du, 1e-12);
  for (int i = threadIdx.x; i < GG; i += blockDim.x) {
    const double num = s_num[i], den = s_den[i];
    const double dens = den > 0 ? num / fmax(den, 1e-12) : 0.0;
    const double pr = (den * fmax(dens, 0.0) + 3.0 * (double)gauss[i]) / (den + 3.0);
    prof[reg * GG + i] = (float)(pr / tot);
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// pass 2 (xpg/integrate_kernels.py:382-483)
// ---------------------------------------------------------------------------------------------------------------------

struct P2Args {
  xpg_detector det;
  xpg_call call;
  xpg_refl R;
  xpg_frames F;
  const float *sb;
  xpg_prepared prep;
  int use_prep;
  int vals_global;          // shoebox values read in place (sb / prep.sb), not staged in LDS
  const float *prof;
  int n_regions;
  const double *bg, *nbg;
  xpg_pass2_out out;
  int waves;
  int wave_bytes;
  int only_partner;         // rows without an overlapping partner are left to k_pass2_small
};

// p2 = P2(u) J / seff^2 per pixel, float64 (xpg/integrate_kernels.py:278-285)
__device__ __forceinline__ void profile_pixels(const float *prof, int region, const double *U1, const double *U2,
                                               const double *Jd, double seff, int P, double *p2, int lane) {
  const float *pr = prof + (int64_t)region * G * G;   // region in [0, n_regions): clamped by the caller
  const double s2 = seff * seff;
  for (int p = lane; p < P; p += WAVE) {
    const Bil q = bilinear(U1[p], U2[p]);
    const double a = (double)pr[q.i2 * G + q.i1], bq = (double)pr[q.i2 * G + q.i1 + 1];
    const double c = (double)pr[(q.i2 + 1) * G + q.i1], d = (double)pr[(q.i2 + 1) * G + q.i1 + 1];
    const double v = ((((1 - q.t1) * (1 - q.t2)) * a + (q.t1 * (1 - q.t2)) * bq) + ((1 - q.t1) * q.t2) * c) +
                     (q.t1 * q.t2) * d;
    p2[p] = (q.inside ? v : 0.0) * (Jd[p] / s2);
  }
}

template <bool PREP, bool VG>
__global__ void __launch_bounds__(256) k_pass2(P2Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);   // wave-uniform: row data in SGPRs
  const int r = blockIdx.x * a.waves + wid;
  if (r >= a.R.n) return;
  Box b = box_of(a.R, r);
  const Layout L = make_layout(b.Dz, b.Dy, b.Dx, 2, !VG);
  if (L.bytes > (size_t)a.wave_bytes) return;
  unsigned char *base = smem + (size_t)wid * a.wave_bytes;
  const int P = b.P;
  float *vals = VG ? const_cast<float *>((PREP ? a.prep.sb : a.sb) + b.voff) : (float *)(base + L.vals);
  uint8_t *vflag = base + L.vflag;
  double *scr = (double *)(base + L.scr);
  double *U1 = (double *)(base + L.U1), *U2 = (double *)(base + L.U2), *Jd = (double *)(base + L.Jd);
  double *p2 = (double *)(base + L.pa), *pp2 = (double *)(base + L.pb);
  float *bref = (float *)(base + L.pc);
  uint8_t *prow = base + L.prow, *pflag = base + L.pflag;
  double *Rd = (double *)(base + L.R), *Rpd = (double *)(base + L.Rp);
  float *Rf = (float *)(base + L.fr), *Rpf = Rf + b.Dz;
  uint8_t *fflag = (uint8_t *)(Rpf + b.Dz);
  double *co = (double *)(base + L.co), *tab = (double *)(base + L.tab);

  b.region = min(max(b.region, 0), a.n_regions - 1);   // memory safety; the host validates region
  const int Dz = b.Dz, Dy = b.Dy, Dx = b.Dx;
  const xpg_detector &det = a.det;
  const xpg_call &call = a.call;
  const double has_p = a.R.has_p ? (double)uni_f(a.R.has_p[r]) : 0.0;
  const bool partner = has_p > 0.5;
  if (a.only_partner && !partner) return;
  const float hpf = (float)has_p;
  const double bgr = a.bg[r];
  const double gain = call.gain;

  // Rows without a partner whose voxels fit NPL pixels x DZM frames per lane (regs) keep, from the fit mask to the
  // tail sums, their net counts and profile values in registers; prepared rows read their voxels straight into them.
  constexpr int NPL = 2, DZM = 8, NK = NPL * DZM;
  const bool regs = !partner && P <= NPL * WAVE && Dz <= DZM;
  const bool direct = PREP && !VG && regs;
  double zeta;
  float dv[NK];
  uint8_t dn[NK];
  if constexpr (PREP) {
    zeta = stage_prepared(b, a.prep, partner ? 1 : 0, vals, vflag, prow, U1, U2, Jd, lane, !direct);
    if (direct) {
      const float *src = a.prep.sb + b.voff;
      const uint8_t *nsrc = a.prep.nbm + b.voff;
#pragma unroll
      for (int q = 0; q < NPL; ++q)
#pragma unroll
        for (int j = 0; j < DZM; ++j) {
          const int v = min(j, Dz - 1) * P + min(lane + q * WAVE, P - 1);   // in the row, whatever the lane
          dv[q * DZM + j] = src[v];
          dn[q * DZM + j] = nsrc[v];
        }
    }
  } else {
    stage_values(b, a.sb, a.F, vals, vflag, prow, base + L.pa, base + L.R, lane);
    Frame f;
    kabsch(a.R.S + 3 * (int64_t)r, call.s0, call.axis, f);
    zeta = f.zeta;
    box_eps_J(det, f, b.ix0, b.iy0, b.s, Dy, Dx, U1, U2, Jd, lane);
  }
  if (PREP && a.prep.R) load_rocking(a.prep, r, Dz, Rd, lane);
  else rocking_curve(call, b.iz0, b.fb, b.phi, zeta, Dz, scr, Rd, lane);
  const double seff = (PREP && a.prep.seff) ? box_u_given(P, uni_d(a.prep.seff[r]), U1, U2, lane)
                                            : box_u(Dy, Dx, sd_of(call), U1, U2, Jd, lane);
  box_coords(b, co, lane);
  profile_pixels(a.prof, b.region, U1, U2, Jd, seff, P, p2, lane);
  wave_sync();
  for (int p = lane; p < P; p += WAVE) {
    const double r2 = U1[p] * U1[p] + U2[p] * U2[p];
    pflag[p] = r2 <= 9.0 ? PF_FG2 : 0;
    bref[p] = template_at(det, co, Dx, p);
    pp2[p] = 0.0;
  }
  frames_fg(Rd, Dz, fflag, lane);
  for (int j = lane; j < Dz; j += WAVE) {
    Rf[j] = (float)Rd[j];
    Rpf[j] = 0.f;
    Rpd[j] = 0.0;
  }
  if (partner) {
    // the overlapping partner's profile on this box's grid, fresh geometry (xpg/integrate_kernels.py:393-400); U1,
    // U2, Jd of this reflection are no longer needed (pflag is set): they take the partner's
    double zeta_p;
    wave_sync();
    if constexpr (PREP) {
      const double *gp = a.prep.geo_p + 3 * b.poff;
      for (int p = lane; p < P; p += WAVE) {
        U1[p] = gp[p];
        U2[p] = gp[P + p];
        Jd[p] = gp[2 * P + p];
      }
      zeta_p = a.prep.zeta_p[r];
    } else {
      Frame fp;
      kabsch(a.R.pS + 3 * (int64_t)r, call.s0, call.axis, fp);
      zeta_p = fp.zeta;
      box_eps_J(det, fp, b.ix0, b.iy0, b.s, Dy, Dx, U1, U2, Jd, lane);
    }
    rocking_curve(call, b.iz0, b.fb, a.R.pphi[r], zeta_p, Dz, scr, Rpd, lane);
    const double seffp = box_u(Dy, Dx, sd_of(call), U1, U2, Jd, lane);
    profile_pixels(a.prof, b.region, U1, U2, Jd, seffp, P, pp2, lane);
    for (int j = lane; j < Dz; j += WAVE) Rpf[j] = (float)Rpd[j];
  }
  wave_sync();
  if constexpr (!PREP)
    neighbour_mask(a.R, r, co, co + Dx, co + Dx + Dy, Dz, Dy, Dx, partner ? 1 : 0, vflag, VF_NBM, tab, prow, lane);

  // Lanes run over pixels, each lane over the frames of its pixel (per-pixel terms in registers, per-frame terms
  // broadcast); sums in float64.

  // ---- fit mask: valid & ~nbm & (p > 0.01 max p | pp > 0.01 max(max pp, 1e-12)), float64 products ----
  // p = p2 R and pp = (pp2 Rp) has_p are products of non-negative factors, so their maxima over the voxels are the
  // products of the factors' maxima (rounding is monotonic).
  double p2m = 0.0, pp2m = 0.0, Rm = 0.0, Rpm = 0.0;
  for (int p = lane; p < P; p += WAVE) {
    p2m = fmax(p2m, p2[p]);
    pp2m = fmax(pp2m, pp2[p]);
  }
  for (int j = lane; j < Dz; j += WAVE) {
    Rm = fmax(Rm, Rd[j]);
    Rpm = fmax(Rpm, Rpd[j]);
  }
  const double pmax = wmax_d(p2m) * wmax_d(Rm);
  const double ppmax = partner ? (wmax_d(pp2m) * wmax_d(Rpm)) * has_p : 0.0;
  const double pthr = 0.01 * pmax, ppthr = 0.01 * fmax(ppmax, 1e-12);
  // Register rows keep their classes as bit masks (bit k = q DZM + j for pixel lane + 64 q, frame j): the same
  // products, summed in the same order as the LDS loops.
  float netr[NK], pfr[NK], b0r[NPL];
  uint32_t usem = 0, inm = 0, validm = 0, fgm = 0;   // bit k: in the box / fit voxel / valid & ~nbm / foreground
  uint32_t nbmr = 0;                                  // bit k: inside a neighbour ellipsoid (mask export only)
  double s_c = 0.0, s_b = 0.0, s_pw = 0.0;
  if (regs) {                             // no partner: pp = +0 and the fit test is p > 0.01 max p
    uint32_t fgz = 0;
#pragma unroll
    for (int j = 0; j < DZM; ++j)
      if (j < Dz && fflag[j]) fgz |= 1u << j;
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int p = lane + q * WAVE;
      const bool pin = p < P;
      const double p2p = pin ? p2[p] : 0.0;
      const bool fg2 = pin && (pflag[p] & PF_FG2);
      const float b0 = pin ? (float)(bgr * (double)bref[p]) : 0.f;
      const float p2f = (float)p2p;
      b0r[q] = b0;
      double cs = 0.0, pw = 0.0;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        netr[k] = 0.f;
        pfr[k] = 0.f;
        if (pin && j < Dz) {
          const int v = j * P + p;
          const float c = direct ? dv[k] : vals[v];
          const bool nbm = direct ? (dn[k] & 1) != 0 : (vflag[v] & VF_NBM) != 0;
          const bool vn = c >= 0.f && !nbm;
          const double pd = p2p * Rd[j];
          inm |= 1u << k;
          if (nbm) nbmr |= 1u << k;
          pfr[k] = p2f * Rf[j];
          netr[k] = c - b0;
          if (vn) validm |= 1u << k;
          if (fg2 && ((fgz >> j) & 1u)) fgm |= 1u << k;
          if (vn && pd > pthr) {
            usem |= 1u << k;
            cs += (double)c;
            pw += pd + 0.0;
            ++cnt;
          }
        }
      }
      if (pin) {
        s_c += cs;
        s_b += (double)cnt * (bgr * (double)bref[p]);
        s_pw += pw;
      }
    }
  } else {
    for (int p0 = 0; p0 < P; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const double p2p = p2[p], pp2p = pp2[p];
        const bool fg2 = pflag[p] & PF_FG2;
        double cs = 0.0, pw = 0.0;
        int cnt = 0;
        for (int j0 = 0; j0 < Dz; j0 += LB) {       // LB frames' reads in flight before any write
          float cc[LB];
          uint8_t ff[LB], fz_[LB];
          double rd[LB], rpd[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
            fz_[u] = fflag[j];
            rd[u] = Rd[j];
            rpd[u] = Rpd[j];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            if (j0 + u >= Dz) break;
            const float c = cc[u];
            const double pd = p2p * rd[u];
            const double ppd = partner ? (pp2p * rpd[u]) * has_p : 0.0;
            uint8_t fl = ff[u];
            if (fg2 && fz_[u]) fl |= VF_FG;
            if (c >= 0.f && !(fl & VF_NBM) && ((pd > pthr) || (ppd > ppthr))) {
              fl |= VF_USE;
              cs += (double)c;
              pw += pd + ppd;
              ++cnt;
            }
            vflag[(j0 + u) * P + p] = fl;
          }
        }
        s_c += cs;
        s_b += (double)cnt * (bgr * (double)bref[p]);
        s_pw += pw;
      }
    }
    wave_sync();
  }
  if (a.out.vmask) {                      // per-voxel classes for mask comparisons (only when asked for)
    uint8_t *vm = a.out.vmask + b.voff;
    for (int p = lane; p < P; p += WAVE) {
      const int q = (p - lane) / WAVE;
      for (int j = 0; j < Dz; ++j) {
        bool use, nbm;
        if (regs) {                       // register rows: their bits (their LDS flags may not be staged)
          use = ((usem >> (q * DZM + j)) & 1u) != 0;
          nbm = ((nbmr >> (q * DZM + j)) & 1u) != 0;
        } else {
          const uint8_t fl = vflag[j * P + p];
          use = (fl & VF_USE) != 0;
          nbm = (fl & VF_NBM) != 0;
        }
        vm[j * P + p] = (use ? XPG_VM_USE : 0) | (nbm ? XPG_VM_NBM : 0);
      }
    }
  }
  double s4[4] = {s_c, s_b, s_pw, 0.0};
  wsum_multi(s4);
  double I = (s4[0] - s4[1]) / fmax(s4[2], 1e-6);
  double Ip = I * has_p;

  // ---- IRLS (xpg/integrate_kernels.py:445-467): each reflection stops on its own test ----
  const float gainf = (float)gain;
  double a11 = 0.0, a12 = 0.0, a22 = 1.0, det2 = 1.0, Iw = I, Ipw = Ip;
  for (int it_ = 0; it_ < N_IRLS; ++it_) {
    Iw = I;
    Ipw = Ip;
    const float Ii = (float)fmax(I, 0.0), Ipi = (float)fmax(Ip, 0.0);
    double s11 = 0.0, s12 = 0.0, s22 = 0.0, sb1 = 0.0, sb2 = 0.0;
    if (regs) {
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        if ((usem >> k) & 1u) {
          const float pf = pfr[k], net = netr[k];
          const float mu = b0r[k / DZM] + Ii * pf;         // the partner term is +0 here
          const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));
          s11 += (double)((pf * pf) * iv);
          sb1 += (double)((pf * net) * iv);
        }
      }
    }
    // rows with an overlapping partner: the 2 x 2 fit in float64 throughout.  With a partner orders of magnitude
    // brighter and overlapping profiles (rho -> 1) the system is nearly singular, and float32 products and weights
    // would move I by more than 1e-3 sigma from the float64 result.
    for (int p0 = 0; p0 < P && !regs && partner; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const double b0 = bgr * (double)bref[p];
        const double p2d = p2[p], pp2d = pp2[p], Iid = fmax(I, 0.0), Ipid = fmax(Ip, 0.0);
        for (int j0 = 0; j0 < Dz; j0 += LB) {
          float cc[LB];
          uint8_t ff[LB];
          double rd[LB], rpd[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
            rd[u] = Rd[j];
            rpd[u] = Rpd[j];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            if (j0 + u >= Dz || !(ff[u] & VF_USE)) continue;
            const double net = (double)cc[u] - b0;
            const double pf = p2d * rd[u];
            const double ppf = (pp2d * rpd[u]) * has_p;
            const double mu = (b0 + Iid * pf) + Ipid * ppf;
            const double iv = 1.0 / (gain * fmax(mu, VAR_FLOOR));
            s11 += (pf * pf) * iv;
            sb1 += (pf * net) * iv;
            s12 += (pf * ppf) * iv;
            s22 += (ppf * ppf) * iv;
            sb2 += (ppf * net) * iv;
          }
        }
      }
    }
    for (int p0 = 0; p0 < P && !regs && !partner; p0 += WAVE) {
      const int p = p0 + lane;
      if (p < P) {
        const float b0 = (float)(bgr * (double)bref[p]);
        const float p2f = (float)p2[p], pp2f = (float)pp2[p];
        for (int j0 = 0; j0 < Dz; j0 += LB) {
          float cc[LB], rf[LB], rpf[LB];
          uint8_t ff[LB];
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            const int j = min(j0 + u, Dz - 1);
            cc[u] = vals[j * P + p];
            ff[u] = vflag[j * P + p];
            rf[u] = Rf[j];
            rpf[u] = Rpf[j];
          }
#pragma unroll
          for (int u = 0; u < LB; ++u) {
            if (j0 + u >= Dz || !(ff[u] & VF_USE)) continue;
            const float net = cc[u] - b0;
            const float pf = p2f * rf[u];
            const float ppf = (pp2f * rpf[u]) * hpf;
            const float mu = (b0 + Ii * pf) + Ipi * ppf;
            const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));   // 1 ulp
            s11 += (double)((pf * pf) * iv);
            sb1 += (double)((pf * net) * iv);
            if (partner) {
              s12 += (double)((pf * ppf) * iv);
              s22 += (double)((ppf * ppf) * iv);
              sb2 += (double)((ppf * net) * iv);
            }
          }
        }
      }
    }
    double b1, b2 = 0.0;
    if (partner) {
      double t8[8] = {s11, sb1, s12, s22, sb2, 0.0, 0.0, 0.0};
      wsum_multi(t8);
      a11 = t8[0];
      b1 = t8[1];
      a12 = t8[2];
      a22 = t8[3] + (1.0 - has_p) * 1.0;
      b2 = t8[4];
    } else {
      double t2[2] = {s11, sb1};
      wsum_multi(t2);
      a11 = t2[0];
      b1 = t2[1];
      a12 = 0.0;
      a22 = 0.0 + (1.0 - has_p) * 1.0;
    }
    det2 = fmax(a11 * a22 - a12 * a12, 1e-30);
    const double I_old = I;
    I = (a22 * b1 - a12 * b2) / det2;
    Ip = (a11 * b2 - a12 * b1) / det2 * has_p;
    if (IRLS_TOL > 0 && it_ >= 1 && it_ < N_IRLS - 1) {
      const double dz = fabs(I - I_old) * sqrt(fmax(det2 / fmax(a22, 1e-30), 0.0));
      if (dz < IRLS_TOL) break;     // uniform: every lane holds the same reduced values
    }
  }

  // ---- tail (xpg/integrate_kernels.py:405-429): weights of the last pass (intensities Iw, Ipw) ----
  const float Ii = (float)fmax(Iw, 0.0), Ipi = (float)fmax(Ipw, 0.0);
  const float Ipos = (float)fmax(Ip, 0.0);
  double sC = 0.0, s110 = 0.0, sC0 = 0.0, swp = 0.0, sres = 0.0;
  int nuse = 0;
  if (regs) {                            // the same sums over the register-held fit voxels (partner terms +0)
#pragma unroll
    for (int q = 0; q < NPL; ++q) {
      const int p = lane + q * WAVE;
      const float br = p < P ? bref[p] : 0.f;
      const float iv0 = __builtin_amdgcn_rcpf(gainf * fmaxf(b0r[q], (float)VAR_FLOOR));
#pragma unroll
      for (int j = 0; j < DZM; ++j) {
        const int k = q * DZM + j;
        if ((usem >> k) & 1u) {
          const float pf = pfr[k], net = netr[k];
          const float mu = b0r[q] + Ii * pf;
          const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));
          const float pb = pf * br;
          sC += (double)(pb * iv);
          s110 += (double)((pf * pf) * iv0);
          sC0 += (double)(pb * iv0);
          swp += (double)pf;
          sres += (double)net;
          nuse += 1;
        }
      }
    }
  }
  for (int p0 = 0; p0 < P && !regs && partner; p0 += WAVE) {       // partner rows: float64 (see the IRLS)
    const int p = p0 + lane;
    if (p < P) {
      const double b0 = bgr * (double)bref[p], br = (double)bref[p];
      const double p2d = p2[p], pp2d = pp2[p];
      const double Iid = fmax(Iw, 0.0), Ipid = fmax(Ipw, 0.0), Iposd = fmax(Ip, 0.0);
      const double iv0 = 1.0 / (gain * fmax(b0, VAR_FLOOR));
      for (int j0 = 0; j0 < Dz; j0 += LB) {
        float cc[LB];
        uint8_t ff[LB];
        double rd[LB], rpd[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int j = min(j0 + u, Dz - 1);
          cc[u] = vals[j * P + p];
          ff[u] = vflag[j * P + p];
          rd[u] = Rd[j];
          rpd[u] = Rpd[j];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          if (j0 + u >= Dz || !(ff[u] & VF_USE)) continue;
          const double net = (double)cc[u] - b0;
          const double pf = p2d * rd[u];
          const double ppf = (pp2d * rpd[u]) * has_p;
          const double mu = (b0 + Iid * pf) + Ipid * ppf;
          const double iv = 1.0 / (gain * fmax(mu, VAR_FLOOR));
          const double pb = pf * br;
          sC += pb * iv;
          s110 += (pf * pf) * iv0;
          sC0 += pb * iv0;
          swp += pf;
          sres += net - Iposd * ppf;
          nuse += 1;
        }
      }
    }
  }
  for (int p0 = 0; p0 < P && !regs && !partner; p0 += WAVE) {
    const int p = p0 + lane;
    if (p < P) {
      const float b0 = (float)(bgr * (double)bref[p]), br = bref[p];
      const float p2f = (float)p2[p], pp2f = (float)pp2[p];
      const float iv0 = __builtin_amdgcn_rcpf(gainf * fmaxf(b0, (float)VAR_FLOOR));
      for (int j0 = 0; j0 < Dz; j0 += LB) {
        float cc[LB], rf[LB], rpf[LB];
        uint8_t ff[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int j = min(j0 + u, Dz - 1);
          cc[u] = vals[j * P + p];
          ff[u] = vflag[j * P + p];
          rf[u] = Rf[j];
          rpf[u] = Rpf[j];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          if (j0 + u >= Dz || !(ff[u] & VF_USE)) continue;
          const float net = cc[u] - b0;
          const float pf = p2f * rf[u];
          const float ppf = (pp2f * rpf[u]) * hpf;
          const float mu = (b0 + Ii * pf) + Ipi * ppf;
          const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));   // 1 ulp
          const float pb = pf * br;
          sC += (double)(pb * iv);
          s110 += (double)((pf * pf) * iv0);
          sC0 += (double)(pb * iv0);
          swp += (double)pf;
          sres += (double)(net - Ipos * ppf);
          nuse += 1;
        }
      }
    }
  }
  double t1[8] = {sC, s110, sC0, swp, sres, (double)nuse, 0.0, 0.0};   // the count is exact as float64
  wsum_multi(t1);
  const double C = t1[0], A110 = t1[1], C0 = t1[2], Swp = t1[3], Sres = t1[4];
  nuse = (int)t1[5];
  const double nu = fmax((double)nuse, 1.0);
  const float pm = (float)(Swp / nu), nm = (float)(Sres / nu);
  const double pmd = Swp / nu, nmd = Sres / nu;
  double sdd = 0.0, sdp = 0.0, snn = 0.0, s_pv = 0.0, s_pt = 0.0, s_pfg = 0.0;
  if (regs) {
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      if (!((inm >> k) & 1u)) continue;
      const float pf = pfr[k];
      s_pt += (double)pf;
      if ((validm >> k) & 1u) s_pv += (double)pf;
      if ((fgm >> k) & 1u) s_pfg += (double)pf;
      if (!((usem >> k) & 1u)) continue;
      const float dp = pf - pm;
      const float dn = netr[k] - nm;       // (net - Ipos pp) - nm with pp = +0
      sdp += (double)(dp * dn);
      sdd += (double)(dp * dp);
      snn += (double)(dn * dn);
    }
  }
  for (int p0 = 0; p0 < P && !regs && partner; p0 += WAVE) {       // partner rows: float64 (see the IRLS)
    const int p = p0 + lane;
    if (p < P) {
      const double b0 = bgr * (double)bref[p];
      const double p2d = p2[p], pp2d = pp2[p], Iposd = fmax(Ip, 0.0);
      for (int j0 = 0; j0 < Dz; j0 += LB) {
        float cc[LB];
        uint8_t ff[LB];
        double rd[LB], rpd[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int j = min(j0 + u, Dz - 1);
          cc[u] = vals[j * P + p];
          ff[u] = vflag[j * P + p];
          rd[u] = Rd[j];
          rpd[u] = Rpd[j];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          if (j0 + u >= Dz) break;
          const uint8_t fl = ff[u];
          const double pf = p2d * rd[u];
          const float c = cc[u];
          s_pt += pf;
          if (c >= 0.f && !(fl & VF_NBM)) s_pv += pf;
          if (fl & VF_FG) s_pfg += pf;
          if (!(fl & VF_USE)) continue;
          const double net = (double)c - b0;
          const double ppf = (pp2d * rpd[u]) * has_p;
          const double dp = pf - pmd;
          const double dn = (net - Iposd * ppf) - nmd;
          sdp += dp * dn;
          sdd += dp * dp;
          snn += dn * dn;
        }
      }
    }
  }
  for (int p0 = 0; p0 < P && !regs && !partner; p0 += WAVE) {
    const int p = p0 + lane;
    if (p < P) {
      const float b0 = (float)(bgr * (double)bref[p]);
      const float p2f = (float)p2[p], pp2f = (float)pp2[p];
      for (int j0 = 0; j0 < Dz; j0 += LB) {
        float cc[LB], rf[LB], rpf[LB];
        uint8_t ff[LB];
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          const int j = min(j0 + u, Dz - 1);
          cc[u] = vals[j * P + p];
          ff[u] = vflag[j * P + p];
          rf[u] = Rf[j];
          rpf[u] = Rpf[j];
        }
#pragma unroll
        for (int u = 0; u < LB; ++u) {
          if (j0 + u >= Dz) break;
          const uint8_t fl = ff[u];
          const float pf = p2f * rf[u];
          const float c = cc[u];
          s_pt += (double)pf;
          if (c >= 0.f && !(fl & VF_NBM)) s_pv += (double)pf;
          if (fl & VF_FG) s_pfg += (double)pf;
          if (!(fl & VF_USE)) continue;
          const float net = c - b0;
          const float ppf = (pp2f * rpf[u]) * hpf;
          const float dp = pf - pm;
          const float dn = (net - Ipos * ppf) - nm;
          sdp += (double)(dp * dn);
          sdd += (double)(dp * dp);
          snn += (double)(dn * dn);
        }
      }
    }
  }
  double t6[8] = {sdp, sdd, snn, s_pv, s_pt, s_pfg, 0.0, 0.0};
  wsum_multi(t6);
  const double Sdp = t6[0], Sdd = t6[1], Snn = t6[2];
  const double Spv = t6[3], Spt = t6[4], Spfg = t6[5];
  if (lane == 0) {
    const double var_b = gain * fmax(bgr, VAR_FLOOR) / fmax(a.nbg[r], 1.0);
    const double a11i = a22 / det2;
    const double varI = a11i + (C * a11i) * (C * a11i) * var_b;
    const double a110 = fmax(A110, 1e-12);
    const double infl = a11i * a11;
    const double var0 = (1.0 / a110 + (C0 / a110) * (C0 / a110) * var_b) * fmax(infl, 1.0);
    const xpg_pass2_out &o = a.out;
    o.iprf[r] = I;
    o.vprf[r] = varI;
    o.vprf0[r] = var0;
    o.corr[r] = Sdp / sqrt(fmax(Sdd * Snn, 1e-20));
    o.p_valid[r] = Spv;
    o.p_total[r] = Spt;
    o.ipartner[r] = Ip;
    o.rho[r] = partner ? a12 / sqrt(fmax(a11 * a22, 1e-30)) : 0.0;
    o.p_fg[r] = Spfg / fmax(Spt, 1e-9);
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// register kernel of pass 2: prepared rows of at most SPL x 64 pixels and SDZ frames without an overlapping partner.
// Each lane holds its pixels' voxels, neighbour bits and geometry in registers, read straight from the prepared
// buffers; LDS holds only the rocking curve and the per-frame values.  The arithmetic is that of the LDS kernel's
// register-row path (same masks, exact integer sums, same per-pixel accumulation order).  (A pass-1 form of it was
// measured slower than the LDS kernel's register rows on full-size blocks and is not kept.)
// ---------------------------------------------------------------------------------------------------------------------

constexpr int SPL = 2, SDZ = 8, SNK = SPL * SDZ;

// LDS of one wavefront: rocking-curve terms (21 Dz doubles), then R, e3 and frame centres (Dz doubles each)
__host__ __device__ static inline size_t small_wave_bytes(int Dz) { return al8(8 * (21 + 3) * (size_t)Dz); }

__host__ __device__ static inline bool small_shape(int Dz, int Dy, int Dx) { return Dz <= SDZ && Dy * Dx <= SPL * WAVE; }

// one row's prepared data in registers: lane l holds pixels l and l + 64 (clamped reads past the box; voxels outside
// it are -3, unrecorded, and belong to no class)
struct SmallRow {
  float c[SNK];
  uint32_t nbm;            // bit k = q SDZ + j
  double e1[SPL], e2[SPL], J[SPL], Jc;
};

__device__ __forceinline__ void small_load(const Box &b, const xpg_prepared &pp, int bit, SmallRow &w, int lane) {
  const float *src = pp.sb + b.voff;
  const uint8_t *nsrc = pp.nbm + b.voff;
  const double *g = pp.geo + 3 * b.poff;
  uint8_t nb[SNK];
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const int pc = min(lane + q * WAVE, b.P - 1);
    w.e1[q] = g[pc];
    w.e2[q] = g[b.P + pc];
    w.J[q] = g[2 * b.P + pc];
#pragma unroll
    for (int j = 0; j < SDZ; ++j) {
      const int v = min(j, b.Dz - 1) * b.P + pc;
      w.c[q * SDZ + j] = src[v];
      nb[q * SDZ + j] = nsrc[v];
    }
  }
  w.Jc = g[2 * b.P + (b.Dy / 2) * b.Dx + b.Dx / 2];
  w.nbm = 0;
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const bool pin = lane + q * WAVE < b.P;
#pragma unroll
    for (int j = 0; j < SDZ; ++j) {
      const int k = q * SDZ + j;
      if (!pin || j >= b.Dz) w.c[k] = -3.0f;
      else if ((nb[k] >> bit) & 1) w.nbm |= 1u << k;
    }
  }
}

// seff and u = eps / seff as box_u / box_u_given compute them
__device__ __forceinline__ double small_u(const xpg_prepared &pp, int r, double sd, const SmallRow &w, double *u1,
                                          double *u2) {
  double seff;
  if (pp.seff) {
    seff = uni_d(pp.seff[r]);
  } else {
    const double psz = sqrt(w.Jc);
    const double t = psz / sd;
    seff = sd * sqrt(1.0 + (t * t) / 12.0);
  }
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    u1[q] = w.e1[q] / seff;
    u2[q] = w.e2[q] / seff;
  }
  return seff;
}

// box row of pixel p by a multiply and shift (p < 128, Dx <= 128: exact)
__device__ __forceinline__ int small_row(int p, uint32_t mul) { return (int)(((uint32_t)p * mul) >> 16); }

// template value of pixel (x, y) (xpg/integrate_kernels.py:68-75)
__device__ __forceinline__ float small_template(const xpg_detector &det, double x, double y) {
  if (!det.bkg) return 1.0f;
  const int kx = min(max((int)rint(x - 1.0), 0), det.bkg_nx - 1);
  const int ky = min(max((int)rint(y - 1.0), 0), det.bkg_ny - 1);
  return det.bkg[(int64_t)ky * det.bkg_nx + kx];
}

// per-frame values into LDS (lane j): e3 = zeta (phi_c - phi) and the frame centre iz0 + fb (j + 1/2); the
// foreground frames (R >= 1e-3 max R + 1e-6) as bits
__device__ __forceinline__ uint32_t small_frames(const xpg_call &call, const Box &b, double zeta, const double *Rd,
                                                 double *e3d, double *zzd, int lane) {
  double Rm = -INFINITY;
  for (int j = 0; j < b.Dz; ++j) Rm = fmax(Rm, Rd[j]);
  const double thr = 1e-3 * Rm + 1e-6;
  uint32_t fgz = 0;
  for (int j = 0; j < b.Dz; ++j)
    if (Rd[j] >= thr) fgz |= 1u << j;
  if (lane < b.Dz) {
    double lo, hi;
    frame_bounds(call, b.iz0, b.fb, lane, lo, hi);
    e3d[lane] = zeta * (0.5 * (lo + hi) - b.phi);
    zzd[lane] = (double)b.iz0 + (double)b.fb * ((double)lane + 0.5);
  }
  wave_sync();
  return fgz;
}

__global__ void __launch_bounds__(256) k_pass2_small(P2Args a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int lane = threadIdx.x & (WAVE - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
  const int r = blockIdx.x * a.waves + wid;
  if (r >= a.R.n) return;
  Box b = box_of(a.R, r);
  if (!small_shape(b.Dz, b.Dy, b.Dx) || small_wave_bytes(b.Dz) > (size_t)a.wave_bytes) return;
  if (a.R.has_p && uni_f(a.R.has_p[r]) > 0.5f) return;     // rows with a partner: the LDS kernel
  unsigned char *base = smem + (size_t)wid * a.wave_bytes;
  double *scr = (double *)base, *Rd = scr + 21 * b.Dz, *e3d = Rd + b.Dz, *zzd = e3d + b.Dz;
  b.region = min(max(b.region, 0), a.n_regions - 1);       // memory safety; the host validates region
  const int P = b.P, Dz = b.Dz, Dx = b.Dx;
  const xpg_detector &det = a.det;
  const xpg_call &call = a.call;
  const double bgr = a.bg[r];
  const double gain = call.gain;

  SmallRow w;
  small_load(b, a.prep, 0, w, lane);
  const double zeta = uni_d(a.prep.zeta[r]);
  if (a.prep.R) load_rocking(a.prep, r, Dz, Rd, lane);
  else rocking_curve(call, b.iz0, b.fb, b.phi, zeta, Dz, scr, Rd, lane);
  double u1[SPL], u2[SPL];
  const double seff = small_u(a.prep, r, sd_of(call), w, u1, u2);
  const uint32_t fgz = small_frames(call, b, zeta, Rd, e3d, zzd, lane);

  // per pixel: profile value P2(u) J / seff^2 (xpg/integrate_kernels.py:278-285), template, foreground ellipse
  const float *pr = a.prof + (int64_t)b.region * G * G;
  const double s2 = seff * seff;
  const uint32_t rmul = (65536u + (uint32_t)Dx - 1u) / (uint32_t)Dx;
  double p2[SPL];
  float brq[SPL];
  uint32_t fg2q = 0;
  double p2m = 0.0, Rm = 0.0;
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const int p = lane + q * WAVE;
    const int pc = min(p, P - 1);
    const int iy = small_row(pc, rmul), ix = pc - iy * Dx;
    brq[q] = p < P ? small_template(det, pix_x(b.ix0, ix, b.s), pix_x(b.iy0, iy, b.s)) : 0.f;
    const Bil bq = bilinear(u1[q], u2[q]);
    const double va = (double)pr[bq.i2 * G + bq.i1], vb = (double)pr[bq.i2 * G + bq.i1 + 1];
    const double vc = (double)pr[(bq.i2 + 1) * G + bq.i1], vd = (double)pr[(bq.i2 + 1) * G + bq.i1 + 1];
    const double v = ((((1 - bq.t1) * (1 - bq.t2)) * va + (bq.t1 * (1 - bq.t2)) * vb) + ((1 - bq.t1) * bq.t2) * vc) +
                     (bq.t1 * bq.t2) * vd;
    p2[q] = (bq.inside ? v : 0.0) * (w.J[q] / s2);
    if (p < P) {
      p2m = fmax(p2m, p2[q]);
      if (u1[q] * u1[q] + u2[q] * u2[q] <= 9.0) fg2q |= 1u << q;
    } else {
      p2[q] = 0.0;
    }
  }
  for (int j = lane; j < Dz; j += WAVE) Rm = fmax(Rm, Rd[j]);
  const double pmax = wmax_d(p2m) * wmax_d(Rm);
  const double pthr = 0.01 * pmax;

  // fit mask and the starting intensity: p > 0.01 max p on valid voxels outside other spots (no partner: pp = +0)
  float netr[SNK], pfr[SNK], b0r[SPL];
  uint32_t usem = 0, validm = 0, fgm = 0;
  double s_c = 0.0, s_b = 0.0, s_pw = 0.0;
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const int p = lane + q * WAVE;
    const bool pin = p < P;
    const bool fg2 = (fg2q >> q) & 1u;
    const float b0 = pin ? (float)(bgr * (double)brq[q]) : 0.f;
    const float p2f = (float)p2[q];
    b0r[q] = b0;
    double cs = 0.0, pw = 0.0;
    int cnt = 0;
#pragma unroll
    for (int j = 0; j < SDZ; ++j) {
      const int k = q * SDZ + j;
      netr[k] = 0.f;
      pfr[k] = 0.f;
      if (pin && j < Dz) {
        const float c = w.c[k];
        const bool vn = c >= 0.f && !((w.nbm >> k) & 1u);
        const double Rj = Rd[j];
        const double pd = p2[q] * Rj;
        pfr[k] = p2f * (float)Rj;
        netr[k] = c - b0;
        if (vn) validm |= 1u << k;
        if (fg2 && ((fgz >> j) & 1u)) fgm |= 1u << k;
        if (vn && pd > pthr) {
          usem |= 1u << k;
          cs += (double)c;
          pw += pd + 0.0;
          ++cnt;
        }
      }
    }
    if (pin) {
      s_c += cs;
      s_b += (double)cnt * (bgr * (double)brq[q]);
      s_pw += pw;
    }
  }
  if (a.out.vmask) {                      // per-voxel classes for mask comparisons (only when asked for)
    uint8_t *vm = a.out.vmask + b.voff;
#pragma unroll
    for (int q = 0; q < SPL; ++q) {
      const int p = lane + q * WAVE;
      if (p >= P) continue;
      for (int j = 0; j < Dz; ++j) {
        const int k = q * SDZ + j;
        vm[j * P + p] = (((usem >> k) & 1u) ? XPG_VM_USE : 0) | (((w.nbm >> k) & 1u) ? XPG_VM_NBM : 0);
      }
    }
  }
  double s4[4] = {s_c, s_b, s_pw, 0.0};
  wsum_multi(s4);
  double I = (s4[0] - s4[1]) / fmax(s4[2], 1e-6);
  double Ip = I * 0.0;

  // ---- IRLS (xpg/integrate_kernels.py:445-467): each reflection stops on its own test ----
  const float gainf = (float)gain;
  double a11 = 0.0, a12 = 0.0, a22 = 1.0, det2 = 1.0, Iw = I;
  for (int it_ = 0; it_ < N_IRLS; ++it_) {
    Iw = I;
    const float Ii = (float)fmax(I, 0.0);
    double s11 = 0.0, sb1 = 0.0;
#pragma unroll
    for (int k = 0; k < SNK; ++k) {
      if ((usem >> k) & 1u) {
        const float pf = pfr[k], net = netr[k];
        const float mu = b0r[k / SDZ] + Ii * pf;         // the partner term is +0 here
        const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));
        s11 += (double)((pf * pf) * iv);
        sb1 += (double)((pf * net) * iv);
      }
    }
    double t2[2] = {s11, sb1};
    wsum_multi(t2);
    a11 = t2[0];
    const double b1 = t2[1];
    a12 = 0.0;
    a22 = 0.0 + (1.0 - 0.0) * 1.0;
    det2 = fmax(a11 * a22 - a12 * a12, 1e-30);
    const double I_old = I;
    I = (a22 * b1 - a12 * 0.0) / det2;
    Ip = (a11 * 0.0 - a12 * b1) / det2 * 0.0;
    if (IRLS_TOL > 0 && it_ >= 1 && it_ < N_IRLS - 1) {
      const double dz = fabs(I - I_old) * sqrt(fmax(det2 / fmax(a22, 1e-30), 0.0));
      if (dz < IRLS_TOL) break;     // uniform: every lane holds the same reduced values
    }
  }

  // ---- tail (xpg/integrate_kernels.py:405-429): weights of the last pass (intensity Iw) ----
  const float Ii = (float)fmax(Iw, 0.0);
  double sC = 0.0, s110 = 0.0, sC0 = 0.0, swp = 0.0, sres = 0.0;
  int nuse = 0;
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    const float br = lane + q * WAVE < P ? brq[q] : 0.f;
    const float iv0 = __builtin_amdgcn_rcpf(gainf * fmaxf(b0r[q], (float)VAR_FLOOR));
#pragma unroll
    for (int j = 0; j < SDZ; ++j) {
      const int k = q * SDZ + j;
      if ((usem >> k) & 1u) {
        const float pf = pfr[k], net = netr[k];
        const float mu = b0r[q] + Ii * pf;
        const float iv = __builtin_amdgcn_rcpf(gainf * fmaxf(mu, (float)VAR_FLOOR));
        const float pb = pf * br;
        sC += (double)(pb * iv);
        s110 += (double)((pf * pf) * iv0);
        sC0 += (double)(pb * iv0);
        swp += (double)pf;
        sres += (double)net;
        nuse += 1;
      }
    }
  }
  double t1[8] = {sC, s110, sC0, swp, sres, (double)nuse, 0.0, 0.0};   // the count is exact as float64
  wsum_multi(t1);
  const double C = t1[0], A110 = t1[1], C0 = t1[2], Swp = t1[3], Sres = t1[4];
  nuse = (int)t1[5];
  const double nu = fmax((double)nuse, 1.0);
  const float pm = (float)(Swp / nu), nm = (float)(Sres / nu);
  double sdd = 0.0, sdp = 0.0, snn = 0.0, s_pv = 0.0, s_pt = 0.0, s_pfg = 0.0;
#pragma unroll
  for (int q = 0; q < SPL; ++q) {
    if (lane + q * WAVE >= P) continue;
#pragma unroll
    for (int j = 0; j < SDZ; ++j) {
      if (j >= Dz) continue;
      const int k = q * SDZ + j;
      const float pf = pfr[k];
      s_pt += (double)pf;
      if ((validm >> k) & 1u) s_pv += (double)pf;
      if ((fgm >> k) & 1u) s_pfg += (double)pf;
      if (!((usem >> k) & 1u)) continue;
      const float dp = pf - pm;
      const float dn = netr[k] - nm;       // (net - Ipos pp) - nm with pp = +0
      sdp += (double)(dp * dn);
      sdd += (double)(dp * dp);
      snn += (double)(dn * dn);
    }
  }
  double t6[8] = {sdp, sdd, snn, s_pv, s_pt, s_pfg, 0.0, 0.0};
  wsum_multi(t6);
  const double Sdp = t6[0], Sdd = t6[1], Snn = t6[2];
  const double Spv = t6[3], Spt = t6[4], Spfg = t6[5];
  if (lane == 0) {
    const double var_b = gain * fmax(bgr, VAR_FLOOR) / fmax(a.nbg[r], 1.0);
    const double a11i = a22 / det2;
    const double varI = a11i + (C * a11i) * (C * a11i) * var_b;
    const double a110 = fmax(A110, 1e-12);
    const double infl = a11i * a11;
    const double var0 = (1.0 / a110 + (C0 / a110) * (C0 / a110) * var_b) * fmax(infl, 1.0);
    const xpg_pass2_out &o = a.out;
    o.iprf[r] = I;
    o.vprf[r] = varI;
    o.vprf0[r] = var0;
    o.corr[r] = Sdp / sqrt(fmax(Sdd * Snn, 1e-20));
    o.p_valid[r] = Spv;
    o.p_total[r] = Spt;
    o.ipartner[r] = Ip;
    o.rho[r] = 0.0;
    o.p_fg[r] = Spfg / fmax(Spt, 1e-9);
  }
}

// ---------------------------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------------------------

static int lds_limit() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (!cached[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
    cached[dev] = v;
  }
  return cached[dev];
}

static int check_refl(const xpg_refl *R, bool need_partner_arrays) {
  if (!R) return fail(XPG_ERR_ARG, "refl is NULL");
  if (R->n < 0 || R->Dz < 1 || R->Dy < 1 || R->Dx < 1) return fail(XPG_ERR_ARG, "bad chunk shape");
  if (R->Dz > XPG_MAX_DZ) return fail(XPG_ERR_ARG, "Dz %d > %d", R->Dz, XPG_MAX_DZ);
  if (R->K < 0 || R->K > XPG_MAX_K) return fail(XPG_ERR_ARG, "K %d outside [0, %d]", R->K, XPG_MAX_K);
  if ((R->Dx < 3 || R->Dy < 3) && (R->Dx < 1 || R->Dy < 1)) return fail(XPG_ERR_ARG, "bad box side");
  if (R->n == 0) return XPG_OK;
  if (!R->S || !R->phi || !R->ix0 || !R->iy0 || !R->iz0 || !R->fb) return fail(XPG_ERR_ARG, "missing refl array");
  if (R->K > 0 && (!R->nbx || !R->nby || !R->nbz || !R->nbsx || !R->nbsz))
    return fail(XPG_ERR_ARG, "missing neighbour columns");
  if (need_partner_arrays && R->has_p && (!R->pS || !R->pphi)) return fail(XPG_ERR_ARG, "has_p without pS/pphi");
  return XPG_OK;
}

// wavefronts per workgroup for a per-wavefront LDS slice of `bytes`: the count that keeps the most wavefronts of the
// launch resident on a CU (its LDS holds floor(lim / (w bytes)) workgroups of w; at most 16 by registers), the larger
// on a tie; cap > 0 limits it
static int pack_waves(size_t bytes, int lim, int cap) {
  int best_w = 1, best = 0;
  for (int w = 4; w >= 1; w >>= 1) {
    if ((cap > 0 && w > cap) || (long long)w * (long long)bytes > lim) continue;
    const int resident = (int)((long long)lim / ((long long)w * (long long)bytes)) * w;
    const int waves = resident < 16 ? resident : 16;
    if (waves > best) {
      best = waves;
      best_w = w;
    }
  }
  return best_w;
}

// in_place_ok: the shoebox values are in a device buffer the kernel may read in place when they do not fit in LDS
static int launch_shape(const xpg_refl *R, int pass, bool in_place_ok, int *waves, int *wave_bytes, size_t *lds,
                        int *vals_global, bool compact = false) {
  // compact: every row of the launch (shapes up to the maxima R->Dz, Dy, Dx) takes the compact layout, which grows
  // with the shape, so that of the maxima bounds them all
  Layout L = make_layout(R->Dz, R->Dy, R->Dx, pass, true, compact);
  const int lim = lds_limit();
  *vals_global = 0;
  if (!compact && R->wave_lds > 0 && R->wave_lds <= lim && (size_t)R->wave_lds < L.bytes)
    L.bytes = al8((size_t)R->wave_lds);
  if ((long long)L.bytes > lim && in_place_ok) {
    L = make_layout(R->Dz, R->Dy, R->Dx, pass, false);
    *vals_global = 1;
  }
  if ((long long)L.bytes > lim)
    return fail(XPG_ERR_TOO_LARGE, "shoebox %dx%dx%d needs %zu bytes of LDS > %d%s", R->Dz, R->Dy, R->Dx, L.bytes,
                lim, in_place_ok ? "" : " (gather it first: xpg_prepare or xpg_extract)");
  const int w = pack_waves(L.bytes, lim, R->wg_waves);
  *waves = w;
  *wave_bytes = (int)L.bytes;
  *lds = (size_t)w * L.bytes;
  return XPG_OK;
}

template <typename K>
static int set_lds(K kernel, size_t lds) {
  if (lds > 65536) HIP_TRY(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return XPG_OK;
}

static int grid_for(int64_t total) {
  int64_t g = (total + 255) / 256;
  if (g > (1 << 20)) g = 1 << 20;
  return (int)(g < 1 ? 1 : g);
}

static const xpg_frames kNoFrames = {};

extern "C" {

const char *xpg_version(void) { return "xpg_amd 0.1 (gfx950)"; }
const char *xpg_last_error(void) { return g_err; }

int xpg_extract(const xpg_frames *frames, const xpg_refl *refl, float *sb, void *stream) {
  int rc = check_refl(refl, false);
  if (rc) return rc;
  if (!frames || !frames->data || !sb) return fail(XPG_ERR_ARG, "frames / sb is NULL");
  if (frames->nx <= 0 || frames->ny <= 0 || frames->frame_stride < (int64_t)frames->nx * frames->ny ||
      (frames->dtype != 0 && frames->dtype != 1) || frames->ring < 0)
    return fail(XPG_ERR_ARG, "bad frame store shape, dtype or ring");
  if (refl->n == 0) return XPG_OK;
  const int lim = lds_limit();
  size_t wb = extract_wave_bytes(refl->Dz, refl->Dy, refl->Dx);
  if (refl->wave_lds > 0 && refl->wave_lds <= lim && (size_t)refl->wave_lds < wb) wb = al8((size_t)refl->wave_lds);
  if ((long long)wb > lim) return fail(XPG_ERR_TOO_LARGE, "shoebox needs %zu bytes of LDS > %d", wb, lim);
  const int w = pack_waves(wb, lim, refl->wg_waves);
  rc = set_lds(k_extract, (size_t)w * wb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_extract, dim3((refl->n + w - 1) / w), dim3(64 * w), (size_t)w * wb, (hipStream_t)stream, *frames,
                     *refl, sb, (int)wb);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_decode_bins(float *arr, const int16_t *cnt, int64_t n_rows, int32_t B, int32_t bxbx, void *stream) {
  if (!arr || !cnt || n_rows < 0 || B < 1 || bxbx < 1) return fail(XPG_ERR_ARG, "bad decode_bins arguments");
  const int64_t total = n_rows * B * (int64_t)bxbx;
  if (total == 0) return XPG_OK;
  hipLaunchKernelGGL(k_decode_bins, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, arr, cnt, n_rows, B, bxbx);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

static int check_frames(const xpg_frames *frames) {
  if (!frames || !frames->data) return fail(XPG_ERR_ARG, "frames is NULL");
  if (frames->nx <= 0 || frames->ny <= 0 || frames->frame_stride < (int64_t)frames->nx * frames->ny ||
      (frames->dtype != 0 && frames->dtype != 1) || frames->ring < 0)
    return fail(XPG_ERR_ARG, "bad frame store shape, dtype or ring");
  return XPG_OK;
}

static int check_prep(const xpg_prepared *prep, const xpg_refl *refl = nullptr) {
  if (!prep || !prep->sb || !prep->geo || !prep->zeta || !prep->nbm || !prep->geo_p || !prep->zeta_p)
    return fail(XPG_ERR_ARG, "incomplete xpg_prepared");
  if (refl && prep->R && prep->r_stride < refl->Dz)
    return fail(XPG_ERR_ARG, "rocking-curve table: r_stride %d < Dz %d", prep->r_stride, refl->Dz);
  return XPG_OK;
}

static int prepare_impl(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                        const float *sb, const xpg_prepared *prep, void *stream) {
  int rc = check_refl(refl, false);
  if (rc) return rc;
  if (!det || !call) return fail(XPG_ERR_ARG, "det / call is NULL");
  if (!sb) {
    rc = check_frames(frames);
    if (rc) return rc;
  }
  rc = check_prep(prep);
  if (rc) return rc;
  if (refl->n == 0) return XPG_OK;
  const int lim = lds_limit();
  int big = 0;
  size_t wb = prep_layout(refl->Dz, refl->Dy, refl->Dx).bytes;
  if (refl->wave_lds > 0 && refl->wave_lds <= lim && (size_t)refl->wave_lds < wb) wb = al8((size_t)refl->wave_lds);
  if ((long long)wb > lim) {
    big = 1;
    wb = prep_layout(refl->Dz, refl->Dy, refl->Dx, true).bytes;
  }
  if ((long long)wb > lim) return fail(XPG_ERR_TOO_LARGE, "shoebox needs %zu bytes of LDS > %d", wb, lim);
  const int w = pack_waves(wb, lim, refl->wg_waves);
  rc = set_lds(k_prepare, (size_t)w * wb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_prepare, dim3((refl->n + w - 1) / w), dim3(64 * w), (size_t)w * wb, (hipStream_t)stream, *det,
                     *call, *refl, sb ? kNoFrames : *frames, sb, *prep, (int)wb, big);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_prepare(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                const xpg_prepared *prep, void *stream) {
  return prepare_impl(det, call, refl, frames, nullptr, prep, stream);
}

int xpg_prepare_sb(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
                   const xpg_prepared *prep, void *stream) {
  if (!sb) return fail(XPG_ERR_ARG, "sb is NULL");
  return prepare_impl(det, call, refl, nullptr, sb, prep, stream);
}

static int pass1_impl(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
                      const xpg_frames *frames, const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream,
                      const xpg_prepared *prep = nullptr) {
  int rc = check_refl(refl, false);
  if (rc) return rc;
  if (!det || !call || !out) return fail(XPG_ERR_ARG, "det / call / out is NULL");
  if (!sb && (!frames || !frames->data)) return fail(XPG_ERR_ARG, "neither sb nor a frame store");
  if (out->fz && out->fz_stride != 0 && out->fz_stride < refl->Dz) return fail(XPG_ERR_ARG, "fz_stride < Dz");
  if (!out->bg || !out->nbg || !out->isum || !out->vsum || !out->vsum0 || !out->xo || !out->yo || !out->zo ||
      !out->m_d || !out->m_m || !out->frac_rec || !out->nfg || !out->zeta || !out->seff || !out->n_over ||
      !out->n_fg_masked || !out->n_fg_nb)
    return fail(XPG_ERR_ARG, "missing per-reflection output");
  if (acc && (!acc->num || !acc->den || acc->n_regions < 1 || acc->copies < 0 || acc->copies > 64))
    return fail(XPG_ERR_ARG, "bad accumulation grids");
  if (det->n_seg < 1 || !det->P0 || !det->Bx || !det->By || !det->U1 || !det->U2 || !det->UN)
    return fail(XPG_ERR_ARG, "bad detector");
  if (refl->n == 0) return XPG_OK;
  if (acc && refl->region == nullptr) return fail(XPG_ERR_ARG, "accumulation needs region");
  P1Args a;
  a.det = *det;
  a.call = *call;
  a.R = *refl;
  a.F = frames ? *frames : kNoFrames;
  a.sb = sb;
  a.use_prep = prep != nullptr;
  if (prep) a.prep = *prep;
  else memset(&a.prep, 0, sizeof(a.prep));
  a.out = *out;
  a.use_acc = acc != nullptr;
  if (acc) a.acc = *acc;
  else memset(&a.acc, 0, sizeof(a.acc));
  size_t lds;
  rc = launch_shape(refl, 1, sb != nullptr, &a.waves, &a.wave_bytes, &lds, &a.vals_global,
                    compact_rows(refl->Dz, refl->Dy, refl->Dx, prep != nullptr, prep && prep->R));
  if (rc) return rc;
  void (*kern)(P1Args) = prep ? (a.vals_global ? k_pass1<true, true> : k_pass1<true, false>)
                               : (a.vals_global ? k_pass1<false, true> : k_pass1<false, false>);
  rc = set_lds(kern, lds);
  if (rc) return rc;
  const int blocks = (refl->n + a.waves - 1) / a.waves;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * a.waves), lds, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_pass1(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
              const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream) {
  if (!sb) return fail(XPG_ERR_ARG, "sb is NULL");
  return pass1_impl(det, call, refl, sb, nullptr, out, acc, stream);
}

int xpg_pass1_frames(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                     const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream) {
  if (!frames || !frames->data) return fail(XPG_ERR_ARG, "frames is NULL");
  return pass1_impl(det, call, refl, nullptr, frames, out, acc, stream);
}

int xpg_accumulate_profiles(const float *img2, const uint8_t *ok2, const float *J, const float *u1, const float *u2,
                            const int32_t *region, const double *isum, const uint8_t *sel, const double *seff,
                            int32_t n, int32_t Dy, int32_t Dx, const xpg_profile_acc *acc, void *stream) {
  if (!img2 || !ok2 || !J || !u1 || !u2 || !region || !isum || !sel || !seff || !acc || !acc->num || !acc->den)
    return fail(XPG_ERR_ARG, "NULL argument");
  if (n < 0 || Dy < 1 || Dx < 1) return fail(XPG_ERR_ARG, "bad shape");
  const int64_t total = (int64_t)n * Dy * Dx;
  if (total == 0) return XPG_OK;
  hipLaunchKernelGGL(k_accumulate, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, img2, ok2, J, u1, u2,
                     region, isum, sel, seff, n, Dy * Dx, *acc);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_normalise_profiles(const xpg_profile_acc *acc, const float *gauss, float *prof, double *num64, double *den64,
                           void *stream) {
  if (!acc || !acc->num || !acc->den || acc->n_regions < 1 || !gauss || !prof) return fail(XPG_ERR_ARG, "NULL argument");
  if (acc->copies < 0 || acc->copies > 64) return fail(XPG_ERR_ARG, "copies %d outside [0, 64]", acc->copies);
  hipLaunchKernelGGL(k_normalise, dim3(acc->n_regions), dim3(256), 0, (hipStream_t)stream, *acc, gauss, prof, num64,
                     den64);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

static int pass2_impl(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
                      const xpg_frames *frames, const float *prof, int32_t n_regions, const double *bg,
                      const double *nbg, int32_t irls_rule, const xpg_pass2_out *out, void *stream,
                      const xpg_prepared *prep = nullptr) {
  int rc = check_refl(refl, true);
  if (rc) return rc;
  if (!det || !call || !out || !prof || !bg || !nbg) return fail(XPG_ERR_ARG, "NULL argument");
  if (!sb && (!frames || !frames->data)) return fail(XPG_ERR_ARG, "neither sb nor a frame store");
  if (!out->iprf || !out->vprf || !out->vprf0 || !out->corr || !out->p_valid || !out->p_total || !out->ipartner ||
      !out->rho || !out->p_fg)
    return fail(XPG_ERR_ARG, "missing output");
  if (irls_rule != 0) return fail(XPG_ERR_ARG, "irls_rule %d: only the per-reflection rule (0) is built", irls_rule);
  if (n_regions < 1) return fail(XPG_ERR_ARG, "n_regions < 1");
  if (det->n_seg < 1 || !det->P0 || !det->Bx || !det->By || !det->U1 || !det->U2 || !det->UN)
    return fail(XPG_ERR_ARG, "bad detector");
  if (refl->n == 0) return XPG_OK;
  P2Args a;
  a.det = *det;
  a.call = *call;
  a.R = *refl;
  a.F = frames ? *frames : kNoFrames;
  a.sb = sb;
  a.use_prep = prep != nullptr;
  if (prep) a.prep = *prep;
  else memset(&a.prep, 0, sizeof(a.prep));
  a.prof = prof;
  a.n_regions = n_regions;
  a.bg = bg;
  a.nbg = nbg;
  a.out = *out;
  a.only_partner = 0;
  if (prep && small_shape(refl->Dz, refl->Dy, refl->Dx)) {
    // prepared rows of at most 128 pixels x 8 frames: rows without a partner in the register kernel, rows with one
    // (if the table has partners) in the LDS kernel, launched over the same rows after it
    P2Args s = a;
    s.waves = refl->wg_waves > 0 && refl->wg_waves < 4 ? refl->wg_waves : 4;
    s.wave_bytes = (int)small_wave_bytes(refl->Dz);
    s.vals_global = 0;
    const size_t slds = (size_t)s.waves * s.wave_bytes;
    hipLaunchKernelGGL(k_pass2_small, dim3((refl->n + s.waves - 1) / s.waves), dim3(64 * s.waves), slds,
                       (hipStream_t)stream, s);
    HIP_TRY(hipGetLastError());
    if (!refl->has_p || refl->lds_rows < 0) return XPG_OK;
    a.only_partner = 1;
  }
  size_t lds;
  rc = launch_shape(refl, 2, sb != nullptr, &a.waves, &a.wave_bytes, &lds, &a.vals_global);
  if (rc) return rc;
  void (*kern)(P2Args) = prep ? (a.vals_global ? k_pass2<true, true> : k_pass2<true, false>)
                               : (a.vals_global ? k_pass2<false, true> : k_pass2<false, false>);
  rc = set_lds(kern, lds);
  if (rc) return rc;
  // partner rows only: the leading rows that may hold them (the caller's lds_rows), not the whole table
  const int nl = a.only_partner && refl->lds_rows > 0 ? (refl->lds_rows < refl->n ? refl->lds_rows : refl->n) : refl->n;
  const int blocks = (nl + a.waves - 1) / a.waves;
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * a.waves), lds, (hipStream_t)stream, a);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_pass2(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const float *sb,
              const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
              const xpg_pass2_out *out, void *stream) {
  if (!sb) return fail(XPG_ERR_ARG, "sb is NULL");
  return pass2_impl(det, call, refl, sb, nullptr, prof, n_regions, bg, nbg, irls_rule, out, stream);
}

int xpg_pass2_frames(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_frames *frames,
                     const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
                     const xpg_pass2_out *out, void *stream) {
  if (!frames || !frames->data) return fail(XPG_ERR_ARG, "frames is NULL");
  return pass2_impl(det, call, refl, nullptr, frames, prof, n_regions, bg, nbg, irls_rule, out, stream);
}

int xpg_pass1_prepared(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep,
                       const xpg_pass1_out *out, const xpg_profile_acc *acc, void *stream) {
  const int rc = check_prep(prep, refl);
  if (rc) return rc;
  return pass1_impl(det, call, refl, prep->sb, nullptr, out, acc, stream, prep);
}

int xpg_pass2_prepared(const xpg_detector *det, const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep,
                       const float *prof, int32_t n_regions, const double *bg, const double *nbg, int32_t irls_rule,
                       const xpg_pass2_out *out, void *stream) {
  const int rc = check_prep(prep, refl);
  if (rc) return rc;
  return pass2_impl(det, call, refl, prep->sb, nullptr, prof, n_regions, bg, nbg, irls_rule, out, stream, prep);
}

int xpg_rocking(const xpg_call *call, const xpg_refl *refl, const xpg_prepared *prep, void *stream) {
  int rc = check_refl(refl, false);
  if (rc) return rc;
  if (!call || !prep || !prep->R || !prep->zeta) return fail(XPG_ERR_ARG, "call / prep / prep->R / prep->zeta is NULL");
  if (prep->r_stride < refl->Dz) return fail(XPG_ERR_ARG, "r_stride %d < Dz %d", prep->r_stride, refl->Dz);
  if (refl->n == 0) return XPG_OK;
  const int64_t total = (int64_t)refl->n * prep->r_stride;
  if (total > ((int64_t)1 << 31) - 256) return fail(XPG_ERR_ARG, "too many rocking-curve values in one launch");
  hipLaunchKernelGGL(k_rocking, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, *call, *refl,
                     *prep);
  HIP_TRY(hipGetLastError());
  return XPG_OK;
}

int xpg_copy_rows(void *dst, int64_t dst_pitch, const void *src, int64_t src_pitch, int64_t width, int64_t height,
                  void *stream) {
  if (!dst || !src || width < 0 || height < 0 || dst_pitch < width || src_pitch < width)
    return fail(XPG_ERR_ARG, "copy_rows: bad pointers or sizes");
  if (width == 0 || height == 0) return XPG_OK;
  HIP_TRY(hipMemcpy2DAsync(dst, (size_t)dst_pitch, src, (size_t)src_pitch, (size_t)width, (size_t)height,
                           hipMemcpyDefault, (hipStream_t)stream));
  return XPG_OK;
}

int xpg_lds_limit(void) { return lds_limit(); }

int xpg_lds_bytes(int32_t Dz, int32_t Dy, int32_t Dx) {
  const size_t a = make_layout(Dz, Dy, Dx, 1).bytes, b = make_layout(Dz, Dy, Dx, 2).bytes;
  return (int)(a > b ? a : b);
}

}  // extern "C"
