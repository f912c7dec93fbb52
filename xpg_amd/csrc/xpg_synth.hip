// Seeded synthetic rotation frames generated on the device (include/xpg_synth.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "xpg_synth.h"

namespace {

// 32-bit counter-based hashing (a low-bias integer finaliser); a draw is keyed by (seed, stream, index) only
__device__ __forceinline__ uint32_t fmix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ uint32_t hash3(uint32_t seed, uint32_t a, uint32_t b) {
  return fmix(fmix(fmix(seed ^ 0x9e3779b9u) ^ a) + 0x85ebca6bu * b);
}
// uniform in (0, 1)
__device__ __forceinline__ float unif(uint32_t h) { return ((float)(h >> 8) + 0.5f) * (1.0f / 16777216.0f); }

// Poisson draw of mean mu given e^-mu: inversion below 30, rounded normal approximation (Box-Muller) above
__device__ __forceinline__ uint32_t poisson(float mu, float emu, uint32_t h) {
  if (mu <= 0.f) return 0;
  const float u = unif(h);
  if (mu < 30.f) {
    float p = emu, cdf = p;
    uint32_t k = 0;
    while (u > cdf && k < 200) {
      ++k;
      p *= mu / (float)k;
      cdf += p;
    }
    return k;
  }
  const float u2 = unif(fmix(h ^ 0xd6e8feb8u));
  const float z = sqrtf(-2.f * logf(u)) * cosf(6.2831853f * u2);
  const float v = floorf(mu + sqrtf(mu) * z + 0.5f);
  return v < 0.f ? 0u : (uint32_t)fminf(v, 4.0e9f);
}

// one thread per pixel, over the frames of the call: the pixel's static flag, mean and e^-mean are computed once
__global__ void __launch_bounds__(256) k_background(xpg_synth_bg p, uint32_t *frames) {
  const uint32_t npx = (uint32_t)p.nx * (uint32_t)p.ny;
  const uint32_t seed = (uint32_t)p.seed ^ (uint32_t)(p.seed >> 32);
  const float cx = (float)p.cx, cy = (float)p.cy;
  const float rmax = sqrtf(fmaxf(cx * cx, (p.nx - cx) * (p.nx - cx)) + fmaxf(cy * cy, (p.ny - cy) * (p.ny - cy)));
  for (uint32_t px = blockIdx.x * blockDim.x + threadIdx.x; px < npx; px += gridDim.x * blockDim.x) {
    const uint32_t iy = px / (uint32_t)p.nx, ix = px - iy * (uint32_t)p.nx;
    const float us = unif(hash3(seed, 0xabcdu, px));                  // static per pixel
    const bool dead = us < (float)p.dead_frac, sat = !dead && us < (float)(p.dead_frac + p.sat_frac);
    const float dx = ix + 1.f - cx, dy = iy + 1.f - cy;
    const float lam = (float)p.bg_centre + (float)(p.bg_edge - p.bg_centre) * sqrtf(dx * dx + dy * dy) / rmax;
    const float elam = expf(-lam);
    for (int t = 0; t < p.n; ++t) {
      uint32_t v;
      if (dead) v = p.flag_value;
      else if (sat) v = p.sat_value;
      else v = poisson(lam, elam, hash3(seed, (uint32_t)(p.f0 + t) + 1u, px));
      frames[(size_t)t * npx + px] = v;
    }
  }
}

__global__ void __launch_bounds__(256) k_spots(xpg_synth_spots p, uint32_t *frames) {
  const int r = blockIdx.x;
  if (r >= p.n) return;
  const float x = (float)p.x[r], y = (float)p.y[r], z = (float)p.z[r], cnt = (float)p.counts[r];
  const float sr = (float)p.sig_r[r], st = (float)p.sig_t[r], sz = (float)p.sig_z[r];
  const float rdx = (float)p.rdx[r], rdy = (float)p.rdy[r];
  int za = (int)floorf(z - 5.f * sz), zb = (int)floorf(z + 5.f * sz);
  za = max(za, p.f0);
  zb = min(zb, p.f0 + p.nf - 1);
  if (zb < za) return;
  const int ix0 = (int)floorf(x - 0.5f), iy0 = (int)floorf(y - 0.5f);
  const int w = 2 * p.half + 1, npx = w * w;
  const size_t npix = (size_t)p.nx * p.ny;
  const float norm = 1.f / (9.f * 6.2831853f * sr * st);
  const float s2z = 1.f / (1.41421356f * sz);
  const float isr = 1.f / sr, ist = 1.f / st;
  const uint32_t seed = ((uint32_t)p.seed ^ (uint32_t)(p.seed >> 32)) ^ 0x1234567u;
  const uint32_t id = p.sid ? (uint32_t)p.sid[r] : (uint32_t)r;
  for (int pp = threadIdx.x; pp < npx; pp += blockDim.x) {
    const int iy = iy0 - p.half + pp / w, ix = ix0 - p.half + pp % w;
    if (ix < 0 || iy < 0 || ix >= p.nx || iy >= p.ny) continue;
    float g = 0.f;                        // 3 x 3 sub-pixel samples of the elliptical Gaussian
    for (int a = -1; a <= 1; ++a)
      for (int b = -1; b <= 1; ++b) {
        const float dxs = ix + 1.f + b / 3.f - x, dys = iy + 1.f + a / 3.f - y;
        const float ur = (dxs * rdx + dys * rdy) * isr, ut = (-dxs * rdy + dys * rdx) * ist;
        g += expf(-0.5f * (ur * ur + ut * ut));
      }
    const float gn = cnt * g * norm;
    if (gn < 1e-7f) continue;
    const size_t o0 = (size_t)iy * p.nx + ix;
    float ehi = erff((za - z) * s2z);
    for (int fi = za; fi <= zb; ++fi) {
      const float elo = ehi;
      ehi = erff((fi + 1.f - z) * s2z);
      const float mu = gn * 0.5f * (ehi - elo);
      if (mu < 1e-7f) continue;
      const size_t o = (size_t)(fi - p.f0) * npix + o0;
      const uint32_t cur = frames[o];
      if (cur == p.flag_value || cur == p.sat_value) continue;
      const uint32_t k = poisson(mu, expf(-mu), hash3(seed ^ id, (uint32_t)fi, (uint32_t)pp));
      if (k) atomicAdd(frames + o, k);
    }
  }
}

}  // namespace

extern "C" {

int xpg_synth_background(const xpg_synth_bg *p, uint32_t *frames, void *stream) {
  if (!p || !frames || p->nx <= 0 || p->ny <= 0 || p->n < 0) return -1;
  const int64_t npx = (int64_t)p->nx * p->ny;
  if (!npx || !p->n) return 0;
  if (npx >= (1ll << 32)) return -1;
  int64_t grid = (npx + 255) / 256;
  if (grid > (1 << 16)) grid = 1 << 16;
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
