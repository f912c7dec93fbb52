// Seeded synthetic rotation frames generated on the device (include/xpg_synth.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_synth.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// uniform in (0, 1) from a key
__device__ __forceinline__ double uni(uint64_t key) {
  return ((double)(mix64(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ uint64_t key3(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(seed ^ 0x5851F42D4C957F2Dull) ^ mix64(a * 0x2545F4914F6CDD1Dull + b));
}

// Poisson draw: inversion for small means, rounded normal approximation (Box-Muller) above 30
__device__ uint32_t poisson(double mu, uint64_t key) {
  if (mu <= 0.0) return 0;
  double u = uni(key);
  if (mu < 30.0) {
    double p = exp(-mu), cdf = p;
    uint32_t k = 0;
    while (u > cdf && k < 200) {
      ++k;
      p *= mu / (double)k;
      cdf += p;
    }
    return k;
  }
  double u2 = uni(key ^ 0xD6E8FEB86659FD93ull);
  double z = sqrt(-2.0 * log(u)) * cos(6.283185307179586 * u2);
  double v = floor(mu + sqrt(mu) * z + 0.5);
  return v < 0.0 ? 0u : (uint32_t)fmin(v, 4.0e9);
}

__global__ void k_background(xpg_synth_bg p, uint32_t *frames) {
  const int64_t npx = (int64_t)p.nx * p.ny;
  const int64_t total = npx * p.n;
  const double rmax = sqrt(fmax(p.cx * p.cx, (p.nx - p.cx) * (p.nx - p.cx)) + fmax(p.cy * p.cy, (p.ny - p.cy) * (p.ny - p.cy)));
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = g / npx, px = g - t * npx;
    const int iy = (int)(px / p.nx), ix = (int)(px - (int64_t)iy * p.nx);
    const double us = uni(key3(p.seed, 0xABCDull, (uint64_t)px));      // static per pixel
    uint32_t v;
    if (us < p.dead_frac) {
      v = p.flag_value;
    } else if (us < p.dead_frac + p.sat_frac) {
      v = p.sat_value;
    } else {
      const double dx = ix + 1.0 - p.cx, dy = iy + 1.0 - p.cy;
      const double lam = p.bg_centre + (p.bg_edge - p.bg_centre) * sqrt(dx * dx + dy * dy) / rmax;
      v = poisson(lam, key3(p.seed, (uint64_t)(p.f0 + t) + 1, (uint64_t)px));
    }
    frames[g] = v;
  }
}

__global__ void __launch_bounds__(256) k_spots(xpg_synth_spots p, uint32_t *frames) {
  const int r = blockIdx.x;
  if (r >= p.n) return;
  const double x = p.x[r], y = p.y[r], z = p.z[r], cnt = p.counts[r];
  const double sr = p.sig_r[r], st = p.sig_t[r], sz = p.sig_z[r], rdx = p.rdx[r], rdy = p.rdy[r];
  int za = (int)floor(z - 5.0 * sz), zb = (int)floor(z + 5.0 * sz);
  za = max(za, p.f0);
  zb = min(zb, p.f0 + p.nf - 1);
  if (zb < za) return;
  const int ix0 = (int)floor(x - 0.5), iy0 = (int)floor(y - 0.5);
  const int w = 2 * p.half + 1, npx = w * w;
  const int64_t npix = (int64_t)p.nx * p.ny;
  const double norm = 1.0 / (9.0 * 6.283185307179586 * sr * st);
  const double s2z = 1.0 / (sqrt(2.0) * sz);
  for (int q = threadIdx.x; q < (zb - za + 1) * npx; q += blockDim.x) {
    const int fi = za + q / npx, pp = q % npx;
    const int iy = iy0 - p.half + pp / w, ix = ix0 - p.half + pp % w;
    if (ix < 0 || iy < 0 || ix >= p.nx || iy >= p.ny) continue;
    const double fz = 0.5 * (erf((fi + 1.0 - z) * s2z) - erf((fi - z) * s2z));
    if (fz < 1e-9) continue;
    double g = 0.0;
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b) {
        const double dxs = ix + 1.0 + b / 3.0 - x, dys = iy + 1.0 + a / 3.0 - y;
        const double ur = dxs * rdx + dys * rdy, ut = -dxs * rdy + dys * rdx;
        g += exp(-0.5 * ((ur / sr) * (ur / sr) + (ut / st) * (ut / st)));
      }
    const double mu = cnt * fz * g * norm;
    if (mu < 1e-7) continue;
    const int64_t o = (int64_t)(fi - p.f0) * npix + (int64_t)iy * p.nx + ix;
    const uint32_t cur = frames[o];
    if (cur == p.flag_value || cur == p.sat_value) continue;
    const uint64_t id = p.sid ? (uint64_t)p.sid[r] : (uint64_t)r;
    const uint32_t k = poisson(mu, key3(p.seed ^ 0x1234567ull, id, ((uint64_t)fi << 32) | (uint32_t)pp));
    if (k) atomicAdd(frames + o, k);
  }
}

}  // namespace

extern "C" {

int xpg_synth_background(const xpg_synth_bg *p, uint32_t *frames, void *stream) {
  if (!p || !frames || p->nx <= 0 || p->ny <= 0 || p->n < 0) return -1;
  const int64_t total = (int64_t)p->nx * p->ny * p->n;
  if (!total) return 0;
  int64_t grid = (total + 255) / 256;
  if (grid > (1 << 20)) grid = 1 << 20;
  hipLaunchKernelGGL(k_background, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream, *p, frames);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int xpg_synth_spots_add(const xpg_synth_spots *p, uint32_t *frames, void *stream) {
  if (!p || !frames || p->n < 0 || p->half < 0 || p->half > 64) return -1;
  if (!p->n) return 0;
  if (!p->x || !p->y || !p->z || !p->counts || !p->sig_r || !p->sig_t || !p->sig_z || !p->rdx || !p->rdy) return -1;
  hipLaunchKernelGGL(k_spots, dim3(p->n), dim3(256), 0, (hipStream_t)stream, *p, frames);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
