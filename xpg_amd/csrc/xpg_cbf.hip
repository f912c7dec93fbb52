// CBF byte-offset decode on MI355X (gfx950) behind include/xpg_cbf.h.
//
// decode_byte_offset (xpg/io/cbf.py:12-44) is a sequential walk: every element's position depends on the lengths of
// all elements before it (1, 3, 7 or 15 bytes) and its value on the sum of all differences before it.  Both are
// recovered in parallel from 256-byte chunks:
//
//   k_cbf_maps    one wavefront per chunk, staged in LDS: for each entry offset e = 0 .. 14 (the first element
//                 starting e bytes into the chunk) the exit offset into the next chunk, the element count and the sum
//                 of the differences (the chunk's map).  Only the 0x80 escape bytes carry a dependency: a scalar walk
//                 over them finds the escaped elements; counts and sums follow from the bytes' int8 prefix sums.
//   k_cbf_groups  one lane per (group of 64 chunks, entry offset): compose the group's chunk maps.
//   k_cbf_frames  one lane per frame: walk the frame's groups from entry 0 -- each group's true entry, element base
//                 and value base; the frame's element total.
//   k_cbf_chunks  one lane per group: walk its chunks from the group's entry -- each chunk's entry and bases.
//   k_cbf_write   one wavefront per chunk: from the true entry, the element starts (bytes not inside an escaped
//                 element) and their running sums (int64, stored as int32) by wave scans in position order; stores
//                 contiguous across lanes.
//
// HBM traffic: the stream read twice (maps and write), the values written once.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <stdio.h>

#include <algorithm>
#include <mutex>

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
constexpr int WAVE64 = 64;

struct Map {                    // per (chunk or group, entry): structure of arrays in the workspace
  uint8_t *exit;
  int32_t *cnt;
  uint32_t *sum;               // mod 2^32: the values are the int32 cast of the running int64 sum
};

struct Work {
  Map cm, gm;                   // chunk maps, group maps
  uint8_t *c_entry;             // per chunk: true entry offset
  int64_t *c_base;             // per chunk: elements before it
  uint32_t *c_vbase;           // per chunk: value sum before it, mod 2^32
  uint8_t *g_entry;
  int64_t *g_base;
  uint32_t *g_vbase;
};

__host__ __device__ inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

__host__ __device__ inline Work carve(void *w, int64_t nc, int64_t ng) {
  uint8_t *p = static_cast<uint8_t *>(w);
  Work k;
  k.cm.sum = reinterpret_cast<uint32_t *>(p); p += align256(4 * NES * nc);
  k.gm.sum = reinterpret_cast<uint32_t *>(p); p += align256(4 * NES * ng);
  k.c_base = reinterpret_cast<int64_t *>(p); p += align256(8 * nc);
  k.c_vbase = reinterpret_cast<uint32_t *>(p); p += align256(4 * nc);
  k.g_base = reinterpret_cast<int64_t *>(p); p += align256(8 * ng);
  k.g_vbase = reinterpret_cast<uint32_t *>(p); p += align256(4 * ng);
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

// One dword of the bytes [c0, c0 + STAGE) of a frame's stream (zero past its end L) into LDS: an aligned dword load
// (streams placed on 4-byte boundaries, c0 a multiple of 256), else byte loads.
constexpr int STAGE = C + 16;   // a chunk plus the longest element that starts in it and ends past it

__device__ __forceinline__ void stage_word(const uint8_t *s, int64_t L, int64_t c0, int w, uint8_t *dst) {
  const int64_t q = c0 + 4 * w;
  uint32_t v = 0;
  if (q + 4 <= L && (reinterpret_cast<uintptr_t>(s + q) & 3) == 0) {
    v = *reinterpret_cast<const uint32_t *>(s + q);
  } else {
    for (int k = 0; k < 4; ++k)
      if (q + k < L) v |= (uint32_t)s[q + k] << (8 * k);
  }
  *reinterpret_cast<uint32_t *>(dst + 4 * w) = v;
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

// ---- byte-parallel chunk parse: one wavefront per chunk ------------------------------------------------------------
// Inside a chunk every byte that is not 0x80 and not inside an escaped element is a one-byte element, so only the
// escape bytes carry the sequential dependency: a scalar walk over the chunk's 0x80 bytes in order (skipping those
// inside the previous escaped element) finds the escaped elements that start there; everything else follows from
// per-byte prefix sums.
constexpr int WPB = 4;   // wavefronts per workgroup
constexpr int CPW = 8;   // consecutive chunks per wavefront in k_cbf_maps and k_cbf_write

struct ChunkView {
  const uint8_t *bytes;   // LDS: STAGE bytes from the chunk's start, zero past the stream's end
  int E;                  // bytes of the chunk inside the stream: min(C, L - c0)
  int Lr;                 // stream end relative to the chunk (may exceed C)
  uint64_t esc[4];        // 0x80 bytes among positions 64 k + lane < E
  int32_t *P;             // LDS (C + 1): P[q] = sum of int8(b[i]) for i < q, i < E
};

// inclusive scan over the wave with DPP: within rows of 16 lanes by row_shr 1, 2, 4, 8 (lanes without a source add
// 0), then row 0's last lane into row 1 and row 2's into row 3 (row_bcast:15), then lane 31 into rows 2 and 3
// (row_bcast:31)
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_add(int v) {
  return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}

__device__ __forceinline__ int wscan32(int v, int lane) {
  (void)lane;
  v = dpp_add<0x111, 0xF>(v);   // row_shr:1
  v = dpp_add<0x112, 0xF>(v);   // row_shr:2
  v = dpp_add<0x114, 0xF>(v);   // row_shr:4
  v = dpp_add<0x118, 0xF>(v);   // row_shr:8
  v = dpp_add<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
  v = dpp_add<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
  return v;
}

// the escaped element at a (bytes[a] == 0x80, a < E): its length (3, 7, 15; 0 when it runs past the stream's end)
// and difference (decode_byte_offset, xpg/io/cbf.py:30-40), read by lanes 0 .. 14 and gathered uniformly
__device__ __forceinline__ int escaped(const ChunkView &v, int a, int64_t &d, int lane) {
  const int x = lane < 15 ? v.bytes[a + lane] : 0;
  const int b1 = __builtin_amdgcn_readlane(x, 1), b2 = __builtin_amdgcn_readlane(x, 2);
  const int v16 = b1 | b2 << 8;
  if (a + 3 > v.Lr) return 0;
  if (v16 != 0x8000) {
    d = (int16_t)v16;
    return 3;
  }
  if (a + 7 > v.Lr) return 0;
  const uint32_t w32 = (uint32_t)__builtin_amdgcn_readlane(x, 3) | (uint32_t)__builtin_amdgcn_readlane(x, 4) << 8 |
                       (uint32_t)__builtin_amdgcn_readlane(x, 5) << 16 | (uint32_t)__builtin_amdgcn_readlane(x, 6) << 24;
  if ((int32_t)w32 != INT32_MIN) {
    d = (int32_t)w32;
    return 7;
  }
  if (a + 15 > v.Lr) return 0;
  uint64_t w = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) w |= (uint64_t)(uint32_t)__builtin_amdgcn_readlane(x, 7 + k) << (8 * k);
  d = (int64_t)w;
  return 15;
}

// Stage the chunk (by the wave), find its 0x80 bytes and (PREFIX) the int8 prefix sums
template <bool PREFIX>
__device__ __forceinline__ void view_chunk(ChunkView &v, uint8_t *buf, int32_t *P, const uint8_t *s, int64_t L,
                                           int64_t c0, int lane) {
  for (int w = lane; w < STAGE / 4; w += WAVE64) stage_word(s, L, c0, w, buf);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  v.bytes = buf;
  v.P = P;
  v.Lr = (int)std::min<int64_t>(L - c0, 1 << 20);
  v.E = std::min(C, v.Lr);
  int carry = 0;
  if (PREFIX && lane == 0) P[0] = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int q = 64 * k + lane;
    const int byte = q < v.E ? buf[q] : 0;
    v.esc[k] = __ballot(q < v.E && byte == 0x80);
    if constexpr (PREFIX) {
      const int inc = wscan32((int)(int8_t)byte, lane) + carry;
      P[q + 1] = inc;
      carry = __builtin_amdgcn_readlane(inc, 63);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Walk from entry e: the escaped elements that start in [e, E), in order; `fn(a, len, d)` for each; stops at one that
// runs past the stream's end (returns false then).  `pos` ends at the position after the last element.
template <typename F>
__device__ __forceinline__ bool walk(const ChunkView &v, int e, int &pos, int lane, F fn) {
  pos = e;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    uint64_t m = v.esc[k];
    while (m) {
      const int a = 64 * k + __builtin_ctzll(m);
      m &= m - 1;
      if (a < pos) continue;               // inside the previous escaped element (or before the entry)
      int64_t d = 0;
      const int len = escaped(v, a, d, lane);
      if (!len) {
        fn(a, 0, 0);
        return false;
      }
      fn(a, len, d);
      pos = a + len;
    }
  }
  if (pos < v.E) pos = v.E;                // one-byte elements up to the chunk's (or stream's) end
  return true;
}

__global__ void __launch_bounds__(64 * WPB) k_cbf_maps(Batch b) {
  __shared__ __align__(16) uint8_t buf[WPB][STAGE];
  __shared__ int32_t Ps[WPB][C + 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t cfirst = ((int64_t)blockIdx.x * WPB + wv) * CPW;
  int f = 0;
  int64_t flo = 0, fhi = 0;                 // the chunks of frame f: [flo, fhi)
  for (int t = 0; t < CPW; ++t) {
    const int64_t c = cfirst + t;
    if (c >= b.nc) break;
    if (c < flo || c >= fhi) {
      f = frame_of(b.chunk0, b.nf, c);
      flo = b.chunk0[f];
      fhi = b.chunk0[f + 1];
    }
    const int64_t c0 = (c - flo) * C;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the previous chunk's LDS reads precede the new stage
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    ChunkView v;
    view_chunk<true>(v, buf[wv], Ps[wv], b.src + b.soff[f], b.slen[f], c0, lane);
    // lane e < 15 follows entry e through one pass over the 0x80 bytes: each escaped element's length and difference
    // are the same for every entry; an entry takes the ones at or past its position.  Starts in [e, end) = bytes there
    // minus the escaped elements' other bytes; sums from the prefix sums, each escaped element's bytes replaced by its
    // difference.
    const int e = lane;
    int pos = e, end = v.E, payload = 0;
    uint32_t corr = 0;
    bool cut = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint64_t m = v.esc[k];
      while (m) {
        const int a = 64 * k + __builtin_ctzll(m);
        m &= m - 1;
        int64_t d = 0;
        const int len = escaped(v, a, d, lane);
        const int hi = std::min(a + (len ? len : 1), v.E);
        const uint32_t pa = (uint32_t)(v.P[hi] - v.P[a]);
        if (e < NE && !cut && a >= pos) {
          if (!len) {
            end = a;
            cut = true;
          } else {
            payload += hi - a - 1;
            corr += (uint32_t)d - pa;
            pos = a + len;
          }
        }
      }
    }
    if (e < NE) {
      int n = 0;
      uint32_t sum = 0;
      uint8_t ex = 0;
      if (e < v.E) {
        if (pos < v.E) pos = v.E;
        n = end - e - payload;
        sum = (uint32_t)(v.P[end] - v.P[e]) + corr;
        ex = cut ? STOP : (uint8_t)(pos > C ? pos - C : 0);
      }
      b.w.cm.exit[c * NES + e] = ex;
      b.w.cm.cnt[c * NES + e] = n;
      b.w.cm.sum[c * NES + e] = sum;
    }
  }
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
  uint32_t sum = 0;
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
  int64_t n = 0;
  uint32_t sum = 0;
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
  int64_t n = b.w.g_base[g];
  uint32_t sum = b.w.g_vbase[g];
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

__global__ void __launch_bounds__(64 * WPB) k_cbf_write(Batch b) {
  // one wavefront per chunk (CPW consecutive chunks in turn): the starts of elements are the bytes from the entry on
  // that are not inside an escaped element; their indices and running sums are wave scans in position order, so the
  // stores are contiguous
  __shared__ __align__(16) uint8_t buf[WPB][STAGE];
  __shared__ uint32_t dq[WPB][C];
  __shared__ uint8_t pay[WPB][C];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t cfirst = ((int64_t)blockIdx.x * WPB + wv) * CPW;
  int f = 0;
  int64_t flo = 0, fhi = 0;                 // the chunks of frame f: [flo, fhi)
  for (int t = 0; t < CPW; ++t) {
    const int64_t c = cfirst + t;
    if (c >= b.nc) break;
    if (c < flo || c >= fhi) {
      f = frame_of(b.chunk0, b.nf, c);
      flo = b.chunk0[f];
      fhi = b.chunk0[f + 1];
    }
    const int e = b.w.c_entry[c];
    const int64_t nmax = b.nel[f];
    const int64_t i0 = b.w.c_base[c];
    if (b.status[f] || e == STOP || i0 >= nmax) continue;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");   // the previous chunk's LDS reads precede the new stage
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t c0 = (c - flo) * C;
    ChunkView v;
    view_chunk<false>(v, buf[wv], nullptr, b.src + b.soff[f], b.slen[f], c0, lane);
#pragma unroll
    for (int k = 0; k < 4; ++k) pay[wv][64 * k + lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int end = v.E, pos = 0;
    walk(v, e, pos, lane, [&](int a, int len, int64_t d) {
      if (!len) {
        end = a;
        return;
      }
      if (lane < len - 1 && a + 1 + lane < C) pay[wv][a + 1 + lane] = 1;
      if (lane == 0) dq[wv][a] = (uint32_t)d;
    });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int32_t *o = b.out + b.ooff[f];
    int64_t cn = 0;
    uint32_t csum = b.w.c_vbase[c];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int q = 64 * k + lane;
      const bool start = q >= e && q < end && !pay[wv][q];
      const int byte = v.bytes[q];
      const uint32_t d = start ? (byte == 0x80 ? dq[wv][q] : (uint32_t)(int32_t)(int8_t)byte) : 0u;
      const int rn = wscan32(start ? 1 : 0, lane);
      const uint32_t rs = (uint32_t)wscan32((int)d, lane);   // wraps mod 2^32 like the int32 cast of the int64 sum
      const int64_t idx = i0 + cn + rn - 1;
      if (start && idx < nmax) o[idx] = (int32_t)(csum + rs);
      cn += __builtin_amdgcn_readlane(rn, 63);
      csum += (uint32_t)__builtin_amdgcn_readlane((int)rs, 63);
    }
  }
}

// The workspace of xpg_cbf_decode_static: one array per GPU, held until the process ends.  Host code never passes
// the array's name to a kernel (on the host it is a stand-in, not a device address): it takes the device address
// from hipGetSymbolAddress once per device and uses that.
static __device__ __align__(256) uint8_t scratch[XPG_CBF_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[MAX_DEVICES] = {};

// the current device's scratch address, looked up on the device's first call (device current)
int scratch_address(void **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return xpg::report_error(XPG_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return xpg::report_error(XPG_ERR_HIP, e != hipSuccess ? hipGetErrorString(e)
                                                                                   : "scratch has no device address");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return 0;
}

int check_batch(const xpg_cbf_batch *p, bool need_work) {
  if (!p || !p->src || !p->stream_off || !p->stream_len || !p->n_elem || !p->chunk0 || !p->group0 ||
      (need_work && !p->work) || !p->out || !p->out_off || !p->status || p->n_frames < 0 || p->n_chunks < 0 ||
      p->n_groups < 0)
    return xpg::report_error(XPG_ERR_ARG, "cbf batch: missing pointer or negative size");
  return 0;
}

int launch_decode(const xpg_cbf_batch *p, void *work, void *stream) {
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
  b.w = carve(work, p->n_chunks, p->n_groups);
  b.out = p->out;
  b.ooff = p->out_off;
  b.status = p->status;
  hipStream_t s = (hipStream_t)stream;
  const int64_t tm = b.nc * NES, tg = b.ng * NES;
  if (tm > 0)
    hipLaunchKernelGGL(k_cbf_maps, dim3((uint32_t)((b.nc + WPB * CPW - 1) / (WPB * CPW))), dim3(64 * WPB), 0, s, b);
  if (tg > 0) hipLaunchKernelGGL(k_cbf_groups, dim3((uint32_t)((tg + 255) / 256)), dim3(256), 0, s, b);
  hipLaunchKernelGGL(k_cbf_frames, dim3((uint32_t)((b.nf + 63) / 64)), dim3(64), 0, s, b);
  if (b.ng > 0) hipLaunchKernelGGL(k_cbf_chunks, dim3((uint32_t)((b.ng + 63) / 64)), dim3(64), 0, s, b);
  if (b.nc > 0)
    hipLaunchKernelGGL(k_cbf_write, dim3((uint32_t)((b.nc + WPB * CPW - 1) / (WPB * CPW))), dim3(64 * WPB), 0, s, b);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int64_t xpg_cbf_workspace_bytes(int64_t n_chunks, int64_t n_groups) {
  if (n_chunks < 0 || n_groups < 0) return -1;
  const Work k = carve(nullptr, n_chunks, n_groups);
  return (int64_t)(reinterpret_cast<uintptr_t>(k.g_entry) + align256(n_groups));
}

int xpg_cbf_decode(const xpg_cbf_batch *p, void *stream) {
  const int rc = check_batch(p, true);
  if (rc) return rc;
  if (p->n_frames == 0) return 0;
  return launch_decode(p, p->work, stream);
}

int xpg_cbf_decode_static(const xpg_cbf_batch *p, void *stream) {
  int rc = check_batch(p, false);
  if (rc) return rc;
  const int64_t need = xpg_cbf_workspace_bytes(p->n_chunks, p->n_groups);
  if (need < 0 || need > XPG_CBF_STATIC_SCRATCH_BYTES) {
    char msg[160];
    snprintf(msg, sizeof(msg), "cbf batch of %lld chunks needs %lld bytes of workspace > the %lld-byte static scratch",
             (long long)p->n_chunks, (long long)need, (long long)XPG_CBF_STATIC_SCRATCH_BYTES);
    return xpg::report_error(XPG_ERR_TOO_LARGE, msg);
  }
  if (p->n_frames == 0) return 0;
  void *work = nullptr;
  rc = scratch_address(&work);
  if (rc) return rc;
  return launch_decode(p, work, stream);
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
