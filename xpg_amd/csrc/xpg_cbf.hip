// CBF byte-offset decode on MI355X (gfx950) behind include/xpg_cbf.h.
//
// decode_byte_offset (xpg/io/cbf.py:12-44) is a sequential walk: every element's position depends on the lengths of
// all elements before it (1, 3, 7 or 15 bytes) and its value on the sum of all differences before it.  Both are
// recovered in parallel from 256-byte chunks:
//
//   k_cbf_maps    one lane per (chunk, entry offset e = 0 .. 14): parse the elements that start in the chunk when the
//                 first one starts e bytes into it; record the exit offset into the next chunk, the element count and
//                 the sum of the differences (the chunk's map).
//   k_cbf_groups  one lane per (group of 64 chunks, entry offset): compose the group's chunk maps.
//   k_cbf_frames  one lane per frame: walk the frame's groups from entry 0 -- each group's true entry, element base
//                 and value base; the frame's element total.
//   k_cbf_chunks  one lane per group: walk its chunks from the group's entry -- each chunk's entry and bases.
//   k_cbf_write   one lane per chunk: parse from the true entry and write the running sums (int64, stored as int32).
//
// HBM traffic: the stream read about twice (the maps' 15 parses hit the cache), the values written once.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#include "xpg_cbf.h"
#include "xpg_integrate.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

constexpr int C = XPG_CBF_CHUNK;
constexpr int G = XPG_CBF_GROUP;
constexpr int NE = 15;          // entry offsets 0 .. 14 (an element is at most 15 bytes)
constexpr int NES = 16;         // map stride
constexpr uint8_t STOP = 255;   // exit of a map whose last element runs past the stream's end

struct Map {                    // per (chunk or group, entry): structure of arrays in the workspace
  uint8_t *exit;
  int32_t *cnt;
  int64_t *sum;
};

struct Work {
  Map cm, gm;                   // chunk maps, group maps
  uint8_t *c_entry;             // per chunk: true entry offset
  int64_t *c_base, *c_vbase;    // per chunk: elements and value sum before it
  uint8_t *g_entry;
  int64_t *g_base, *g_vbase;
};

__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

__host__ __device__ inline Work carve(void *w, int64_t nc, int64_t ng) {
  uint8_t *p = static_cast<uint8_t *>(w);
  Work k;
  k.cm.sum = reinterpret_cast<int64_t *>(p); p += align256(8 * NES * nc);
  k.gm.sum = reinterpret_cast<int64_t *>(p); p += align256(8 * NES * ng);
  k.c_base = reinterpret_cast<int64_t *>(p); p += align256(8 * nc);
  k.c_vbase = reinterpret_cast<int64_t *>(p); p += align256(8 * nc);
  k.g_base = reinterpret_cast<int64_t *>(p); p += align256(8 * ng);
  k.g_vbase = reinterpret_cast<int64_t *>(p); p += align256(8 * ng);
  k.cm.cnt = reinterpret_cast<int32_t *>(p); p += align256(4 * NES * nc);
  k.gm.cnt = reinterpret_cast<int32_t *>(p); p += align256(4 * NES * ng);
  k.cm.exit = p; p += align256(NES * nc);
  k.gm.exit = p; p += align256(NES * ng);
  k.c_entry = p; p += align256(nc);
  k.g_entry = p; p += align256(ng);
  return k;
}

struct Batch {
  const uint8_t *src;
  const int64_t *soff, *slen, *nel, *chunk0, *group0;
  int32_t nf;
  int64_t nc, ng;
  Work w;
  int32_t *out;
  const int64_t *ooff;
  int32_t *status;
};

__device__ __forceinline__ uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}

// One element at p (p < L): its length (1, 3, 7 or 15 bytes) and difference, or 0 when its encoding runs past L
// (decode_byte_offset, xpg/io/cbf.py:24-40).
__device__ __forceinline__ int element(const uint8_t *s, int64_t L, int64_t p, int64_t &d) {
  const uint8_t b0 = s[p];
  if (b0 != 0x80) {
    d = (int8_t)b0;
    return 1;
  }
  if (p + 3 > L) return 0;
  const int v16 = s[p + 1] | s[p + 2] << 8;
  if (v16 != 0x8000) {
    d = (int16_t)v16;
    return 3;
  }
  if (p + 7 > L) return 0;
  const int32_t v32 = (int32_t)le32(s + p + 3);
  if (v32 != INT32_MIN) {
    d = v32;
    return 7;
  }
  if (p + 15 > L) return 0;
  d = (int64_t)((uint64_t)le32(s + p + 7) | (uint64_t)le32(s + p + 11) << 32);
  return 15;
}

// frame of a global chunk (or group) index: the last f with first[f] <= i (binary search over n_frames + 1 bounds)
__device__ __forceinline__ int frame_of(const int64_t *first, int nf, int64_t i) {
  int lo = 0, hi = nf - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) k_cbf_maps(Batch b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t c = t / NES;
  const int e = (int)(t % NES);
  if (c >= b.nc || e >= NE) return;
  const int f = frame_of(b.chunk0, b.nf, c);
  const uint8_t *s = b.src + b.soff[f];
  const int64_t L = b.slen[f];
  const int64_t c0 = (c - b.chunk0[f]) * C;
  int64_t p = c0 + e;
  int32_t n = 0;
  int64_t sum = 0;
  uint8_t ex = 0;
  while (p < c0 + C && p < L) {
    int64_t d;
    const int len = element(s, L, p, d);
    if (!len) {
      ex = STOP;
      break;
    }
    sum += d;
    ++n;
    p += len;
  }
  if (ex != STOP) ex = (uint8_t)(p >= c0 + C ? p - (c0 + C) : 0);
  b.w.cm.exit[c * NES + e] = ex;
  b.w.cm.cnt[c * NES + e] = n;
  b.w.cm.sum[c * NES + e] = sum;
}

__global__ void __launch_bounds__(256) k_cbf_groups(Batch b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = t / NES;
  const int e0 = (int)(t % NES);
  if (g >= b.ng || e0 >= NE) return;
  const int f = frame_of(b.group0, b.nf, g);
  const int64_t cfirst = b.chunk0[f] + (g - b.group0[f]) * G;
  const int64_t cend = std::min<int64_t>(cfirst + G, b.chunk0[f + 1]);
  int e = e0;
  int32_t n = 0;
  int64_t sum = 0;
  for (int64_t c = cfirst; c < cend && e != STOP; ++c) {
    n += b.w.cm.cnt[c * NES + e];
    sum += b.w.cm.sum[c * NES + e];
    e = b.w.cm.exit[c * NES + e];
  }
  b.w.gm.exit[g * NES + e0] = (uint8_t)e;
  b.w.gm.cnt[g * NES + e0] = n;
  b.w.gm.sum[g * NES + e0] = sum;
}

__global__ void __launch_bounds__(64) k_cbf_frames(Batch b) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= b.nf) return;
  int e = 0;
  int64_t n = 0, sum = 0;
  for (int64_t g = b.group0[f]; g < b.group0[f + 1]; ++g) {
    b.w.g_entry[g] = (uint8_t)e;
    b.w.g_base[g] = n;
    b.w.g_vbase[g] = sum;
    if (e == STOP) continue;
    n += b.w.gm.cnt[g * NES + e];
    sum += b.w.gm.sum[g * NES + e];
    e = b.w.gm.exit[g * NES + e];
  }
  // the stream must hold n_elem whole elements (an element cut by the stream's end stops the count)
  b.status[f] = n < b.nel[f] ? 1 : 0;
}

__global__ void __launch_bounds__(64) k_cbf_chunks(Batch b) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= b.ng) return;
  const int f = frame_of(b.group0, b.nf, g);
  const int64_t cfirst = b.chunk0[f] + (g - b.group0[f]) * G;
  const int64_t cend = std::min<int64_t>(cfirst + G, b.chunk0[f + 1]);
  int e = b.w.g_entry[g];
  int64_t n = b.w.g_base[g], sum = b.w.g_vbase[g];
  for (int64_t c = cfirst; c < cend; ++c) {
    b.w.c_entry[c] = (uint8_t)e;
    b.w.c_base[c] = n;
    b.w.c_vbase[c] = sum;
    if (e == STOP) continue;
    n += b.w.cm.cnt[c * NES + e];
    sum += b.w.cm.sum[c * NES + e];
    e = b.w.cm.exit[c * NES + e];
  }
}

__global__ void __launch_bounds__(256) k_cbf_write(Batch b) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= b.nc) return;
  const int f = frame_of(b.chunk0, b.nf, c);
  if (b.status[f]) return;
  const int e = b.w.c_entry[c];
  const int64_t nmax = b.nel[f];
  int64_t i = b.w.c_base[c];
  if (e == STOP || i >= nmax) return;
  const uint8_t *s = b.src + b.soff[f];
  const int64_t L = b.slen[f];
  const int64_t c0 = (c - b.chunk0[f]) * C;
  int32_t *o = b.out + b.ooff[f];
  int64_t sum = b.w.c_vbase[c];
  for (int64_t p = c0 + e; p < c0 + C && p < L && i < nmax;) {
    int64_t d;
    const int len = element(s, L, p, d);
    if (!len) break;              // cannot happen for i < n_elem once the frame's status is 0
    sum += d;
    o[i++] = (int32_t)(uint32_t)(uint64_t)sum;
    p += len;
  }
}

}  // namespace

extern "C" {

int64_t xpg_cbf_workspace_bytes(int64_t n_chunks, int64_t n_groups) {
  if (n_chunks < 0 || n_groups < 0) return -1;
  const Work k = carve(nullptr, n_chunks, n_groups);
  return (int64_t)(reinterpret_cast<uintptr_t>(k.g_entry) + align256(n_groups));
}

int xpg_cbf_decode(const xpg_cbf_batch *p, void *stream) {
  if (!p || !p->src || !p->stream_off || !p->stream_len || !p->n_elem || !p->chunk0 || !p->group0 || !p->work ||
      !p->out || !p->out_off || !p->status || p->n_frames < 0 || p->n_chunks < 0 || p->n_groups < 0)
    return xpg::report_error(XPG_ERR_ARG, "cbf batch: missing pointer or negative size");
  if (p->n_frames == 0) return 0;
  Batch b;
  b.src = p->src;
  b.soff = p->stream_off;
  b.slen = p->stream_len;
  b.nel = p->n_elem;
  b.chunk0 = p->chunk0;
  b.group0 = p->group0;
  b.nf = p->n_frames;
  b.nc = p->n_chunks;
  b.ng = p->n_groups;
  b.w = carve(p->work, p->n_chunks, p->n_groups);
  b.out = p->out;
  b.ooff = p->out_off;
  b.status = p->status;
  hipStream_t s = (hipStream_t)stream;
  const int64_t tm = b.nc * NES, tg = b.ng * NES;
  if (tm > 0) hipLaunchKernelGGL(k_cbf_maps, dim3((uint32_t)((tm + 255) / 256)), dim3(256), 0, s, b);
  if (tg > 0) hipLaunchKernelGGL(k_cbf_groups, dim3((uint32_t)((tg + 255) / 256)), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_cbf_frames, dim3((uint32_t)((b.nf + 63) / 64)), dim3(64), 0, s, b);
  if (b.ng > 0) hipLaunchKernelGGL(k_cbf_chunks, dim3((uint32_t)((b.ng + 63) / 64)), dim3(64), 0, s, b);
  if (b.nc > 0) hipLaunchKernelGGL(k_cbf_write, dim3((uint32_t)((b.nc + 255) / 256)), dim3(256), 0, s, b);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return 0;
}

int64_t xpg_cbf_encode_host(const int32_t *values, int64_t n, uint8_t *out, int64_t cap) {
  if (!values || !out || n < 0) return -1;
  int64_t pos = 0, prev = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t d = (int64_t)values[i] - prev;
    prev = values[i];
    if (d >= -127 && d <= 127) {
      if (pos + 1 > cap) return -1;
      out[pos++] = (uint8_t)(int8_t)d;
    } else if (d >= -32767 && d <= 32767) {
      if (pos + 3 > cap) return -1;
      const uint16_t v = (uint16_t)(int16_t)d;
      out[pos] = 0x80;
      out[pos + 1] = (uint8_t)v;
      out[pos + 2] = (uint8_t)(v >> 8);
      pos += 3;
    } else {
      // int32 differences (values are int32, so |d| < 2^32; beyond int32 the reference raises too)
      if (d < -2147483647LL || d > 2147483647LL) return -1;
      if (pos + 7 > cap) return -1;
      const uint32_t v = (uint32_t)(int32_t)d;
      out[pos] = 0x80;
      out[pos + 1] = 0x00;
      out[pos + 2] = 0x80;
      for (int k = 0; k < 4; ++k) out[pos + 3 + k] = (uint8_t)(v >> (8 * k));
      pos += 7;
    }
  }
  return pos;
}

}  // extern "C"
