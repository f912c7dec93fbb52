// Bitshuffle-LZ4 frame decode on MI355X (gfx950) behind include/xpg_decode.h.
//
// The reference decodes in two places: host worker processes run the LZ4 stage (lz4_stage, xpg/io/bslz4.py:34-63)
// and the device undoes the bit shuffle and applies the value codes (unshuffle_frame_torch, xpg/io/bslz4.py:80-108).
// Here both run in one kernel over chunks resident in HBM:
//
//   k_bslz4_decode   one wavefront per LZ4 block (block_elems elements, 8 KB for the default block).  The wave stages
//                    the compressed block in its LDS slice with aligned 16-byte loads, parses the sequences with
//                    wave-uniform control flow (token, lengths and offsets in scalar registers), copies each literal
//                    run and match with all 64 lanes, then undoes the bit shuffle from LDS -- lane l takes the groups
//                    of 8 elements g = l, l + 64, ...: the 8 x 8 bit matrix of each output byte is gathered from 8
//                    planes and transposed in registers -- and writes the block's pixels (raw uint32 counts for the
//                    integrator's frame store, or value-coded float32).
//   k_bslz4_index    one lane per frame walks the chain of block size fields into the block table (the host form,
//                    xpg_bslz4_index_host, is the same walk over a chunk in host memory).
//
// HBM traffic per frame: the compressed chunk read once, the frame written once (4 bytes per pixel).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "xpg_decode.h"
#include "xpg_integrate.h"

namespace xpg {
int report_error(int code, const char *msg);
}

namespace {

constexpr int WAVE = 64;
constexpr int MINMATCH = 4;       // LZ4 block format: shortest match
constexpr int LASTLITERALS = 5;   // the last 5 output bytes are literals
constexpr int MFLIMIT = 12;       // a match may not start in the last 12 output bytes
constexpr int MAX_BLOCK_BYTES = 65536;

__host__ __device__ inline uint32_t be32(const uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

__host__ __device__ inline int64_t lz4_bound(int64_t n) { return n + n / 255 + 16; }

__host__ __device__ inline int32_t block_elems_of(uint32_t block_bytes, int32_t esize) {
  const int64_t bb = block_bytes ? block_bytes : 8192;
  int64_t be = bb / esize;
  be -= be % 8;
  return (int32_t)std::min<int64_t>(be, INT32_MAX);
}

// lz4_stage's walk over one chunk (xpg/io/bslz4.py:36-62): header checks, then per block its size field, its LZ4 data
// and, for the last block, the raw tail; table rows as include/xpg_decode.h describes them.
__host__ __device__ inline int32_t walk_chunk(const uint8_t *c, int64_t len, int64_t nel, int32_t esize, int32_t belem,
                                              int32_t nbf, uint32_t *t) {
  for (int k = 0; k <= nbf; ++k) t[2 * k] = t[2 * k + 1] = 0;
  if (len < 12) return XPG_BSLZ4_HEADER;
  if (len > (int64_t)UINT32_MAX) return XPG_BSLZ4_TRUNCATED;
  const uint64_t total = (uint64_t)be32(c) << 32 | be32(c + 4);
  if (total != (uint64_t)nel * (uint64_t)esize) return XPG_BSLZ4_SIZE;
  const int32_t be = block_elems_of(be32(c + 8), esize);
  if (be <= 0 || be != belem) return XPG_BSLZ4_BLOCK;
  int64_t pos = 12, done = 0;
  for (int k = 0; done < nel; ++k) {
    const int64_t n = nel - done < belem ? nel - done : belem;
    const int64_t n8 = n - n % 8;
    if (n8 > 0) {
      if (pos + 4 > len) return XPG_BSLZ4_TRUNCATED;
      const uint32_t cs = be32(c + pos);
      pos += 4;
      if ((int64_t)cs > lz4_bound(n8 * esize)) return XPG_BSLZ4_CSIZE;
      if (pos + cs > len) return XPG_BSLZ4_TRUNCATED;
      t[2 * k] = (uint32_t)pos;
      t[2 * k + 1] = cs;
      pos += cs;
    }
    if (n8 < n) {
      const int64_t rest = (n - n8) * esize;
      if (pos + rest > len) return XPG_BSLZ4_TRUNCATED;
      t[2 * nbf] = (uint32_t)pos;
      t[2 * nbf + 1] = (uint32_t)rest;
      pos += rest;
    }
    done += n;
  }
  return XPG_BSLZ4_OK;
}

// 8 x 8 bit-matrix transpose of a 64-bit word, bit 8 r + c <-> bit 8 c + r (three delta swaps; Warren, Hacker's
// Delight, section 7-3)
__host__ __device__ inline uint64_t transpose8(uint64_t x) {
  uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x ^= t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x ^= t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x ^= t ^ (t << 28);
  return x;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

struct Batch {
  const uint8_t *src;
  const int64_t *coff, *clen;
  int32_t nf, esize;
  int64_t nel;
  int32_t belem, nbf;
  uint32_t *table;
  int32_t *status;
  void *out;
  int64_t ostride;
  const int64_t *oslot;
  int32_t use_min, use_ovl;
  int64_t min_ceil, ovl_ceil;
  const uint8_t *sbad;
  int32_t cap_c, wave_lds;  // LDS: staging region (16-aligned, + 16 for the source misalignment), then the block
  int64_t f0;               // first frame of this launch
};

__global__ void __launch_bounds__(64) k_bslz4_index(Batch b) {
  const int64_t f = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (f >= b.nf) return;
  const int32_t st = walk_chunk(b.src + b.coff[f], b.clen[f], b.nel, b.esize, b.belem, b.nbf,
                                b.table + f * (int64_t)(b.nbf + 1) * 2);
  b.status[f] = st;
}

// Extended length (LZ4 block format: 255-valued bytes add up until a smaller byte) read 64 bytes at a time.  The
// sequential reading stops after the byte that brings ip to `lim` (LZ4_decompress_safe's bound checks); `stopped`
// reports that.  Returns the added length; ip advances past the bytes read.
__device__ __forceinline__ int ext_len(const uint8_t *cb, int cs, int &ip, int lim, bool &stopped, int lane) {
  int add = 0;
  for (;;) {
    const int p = ip + lane;
    const int v = p < cs ? cb[p] : 0;
    const uint64_t nz = __ballot(v != 255);
    const int q = nz ? __builtin_ctzll(nz) : WAVE;
    const int istop = lim - ip - 1;   // index of the byte after which ip reaches lim
    const int c = q < istop ? q : istop;
    if (c >= WAVE) {                  // 64 bytes of 255 and no stop
      add += 255 * WAVE;
      ip += WAVE;
      continue;
    }
    add += 255 * c + __builtin_amdgcn_readlane(v, c);
    ip += c + 1;
    stopped = c == istop;
    return add;
  }
}

// One LZ4 block of cs bytes (cb, LDS) into exactly dn bytes (db, LDS), by the whole wave; 0 or XPG_BSLZ4_LZ4.
// Rules of the LZ4 block format as LZ4_decompress_safe applies them: literals that reach the last MFLIMIT output
// bytes or the last 8 input bytes end the block (they must consume the input exactly); a match must lie in the
// decoded output and end before the last LASTLITERALS bytes; offset 0 is rejected.
__device__ int lz4_block(const uint8_t *cb, int cs, uint8_t *db, int dn, int lane) {
  int ip = 0, op = 0;
  for (;;) {
    if (ip >= cs) return XPG_BSLZ4_LZ4;
    const int tok = uni(cb[ip]);
    ++ip;
    int lit = tok >> 4;
    if (lit == 15) {
      if (ip >= cs - 15) return XPG_BSLZ4_LZ4;
      bool stopped = false;
      lit += ext_len(cb, cs, ip, cs - 15, stopped, lane);
      ip = uni(ip);
      lit = uni(lit);
    }
    if (lit > cs - ip) return XPG_BSLZ4_LZ4;
    const bool last = lit > dn - MFLIMIT - op || lit > cs - 8 - ip;
    if (last && (ip + lit != cs || lit > dn - op)) return XPG_BSLZ4_LZ4;
    for (int j = lane; j < lit; j += WAVE) db[op + j] = cb[ip + j];
    ip += lit;
    op += lit;
    if (last) break;
    const int off = uni(cb[ip] | cb[ip + 1] << 8);
    ip += 2;
    int ml = tok & 15;
    if (ml == 15) {
      bool stopped = false;
      ml += ext_len(cb, cs, ip, cs - LASTLITERALS + 1, stopped, lane);
      ip = uni(ip);
      ml = uni(ml);
      if (stopped) return XPG_BSLZ4_LZ4;
    }
    ml += MINMATCH;
    if (off == 0 || off > op || ml > dn - LASTLITERALS - op) return XPG_BSLZ4_LZ4;
    wave_sync();
    if (off >= WAVE || off >= ml) {
      // every source byte was written before the pass of 64 that reads it
      for (int j = lane; j < ml; j += WAVE) {
        db[op + j] = db[op - off + j];
        wave_sync();
      }
    } else {
      // overlapping match of period off < 64: byte j repeats byte j mod off of the last off bytes
      int r = lane % off;
      const int step = WAVE % off;
      for (int j = lane; j < ml; j += WAVE) {
        db[op + j] = db[op - off + r];
        r += step;
        if (r >= off) r -= off;
      }
    }
    op += ml;
    wave_sync();
  }
  return op == dn ? 0 : XPG_BSLZ4_LZ4;
}

// Global -> LDS copy of n bytes from s with aligned 16-byte loads: byte i lands at base + (s mod 16) + i, so the
// aligned middle of the source maps to aligned LDS words.  Returns the LDS address of byte 0.
__device__ __forceinline__ uint8_t *stage(const uint8_t *s, int n, uint8_t *base, int lane) {
  const int mis = (int)((uintptr_t)s & 15);
  uint8_t *d = base + mis;
  const int head = std::min(n, (16 - mis) & 15);
  for (int j = lane; j < head; j += WAVE) d[j] = s[j];
  const int nw = (n - head) >> 4;
  const uint4 *sw = reinterpret_cast<const uint4 *>(s + head);
  uint4 *dw = reinterpret_cast<uint4 *>(d + head);
  for (int w = lane; w < nw; w += WAVE) dw[w] = sw[w];
  for (int j = head + (nw << 4) + lane; j < n; j += WAVE) d[j] = s[j];
  return d;
}

template <int KIND>
__device__ __forceinline__ void put(const Batch &b, void *out, int64_t e, uint32_t v) {
  if constexpr (KIND == 0) {
    static_cast<uint32_t *>(out)[e] = v;
  } else {
    // unshuffle_frame_torch's codes (xpg/io/bslz4.py:100-108): all ones (2^n - 1) -> -1, below min_valid -> -1,
    // static mask -> -1, at or above the overload -> -2; -1 wins
    const uint32_t ones = b.esize == 4 ? 0xFFFFFFFFu : 0xFFFFu;
    bool bad = v == ones;
    if (b.use_min) bad |= (int64_t)v < b.min_ceil;
    if (b.sbad) bad |= b.sbad[e] != 0;
    float f = (float)v;
    if (b.use_ovl && (int64_t)v >= b.ovl_ceil) f = -2.0f;
    static_cast<float *>(out)[e] = bad ? -1.0f : f;
  }
}

template <int ES, int KIND>
__global__ void __launch_bounds__(64) k_bslz4_decode(Batch b) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const int64_t gw = b.f0 * b.nbf + blockIdx.x;
  const int64_t f = gw / b.nbf;
  const int k = (int)(gw - f * b.nbf);
  if (f >= b.nf || b.status[f] != 0) return;
  uint8_t *cbase = lds;
  uint8_t *db = lds + b.cap_c;
  const uint32_t *t = b.table + f * (int64_t)(b.nbf + 1) * 2;
  const uint8_t *chunk = b.src + b.coff[f];
  const int64_t e0 = (int64_t)k * b.belem;
  const int n = (int)std::min<int64_t>(b.belem, b.nel - e0);
  const int n8 = n - n % 8;
  if (n8 > 0) {
    const int cs = (int)t[2 * k + 1];
    const uint8_t *cb = stage(chunk + t[2 * k], cs, cbase, lane);
    wave_sync();
    const int err = lz4_block(cb, cs, db, n8 * ES, lane);
    if (err) {
      if (lane == 0) atomicOr(&b.status[f], err);
      return;
    }
  }
  if (n8 < n) {
    const uint8_t *tail = chunk + t[2 * b.nbf];
    for (int j = lane; j < (n - n8) * ES; j += WAVE) db[n8 * ES + j] = tail[j];
  }
  wave_sync();
  const int64_t slot = b.oslot ? b.oslot[f] : f;
  void *out = static_cast<uint8_t *>(b.out) + slot * b.ostride * 4;
  // bit-shuffled part: plane p (bit p of every element) holds m = n8 / 8 bytes, element i's bit at byte i / 8, bit
  // i % 8; the 8 planes 8 j .. 8 j + 7 of byte g form the 8 x 8 matrix of output byte j of elements 8 g .. 8 g + 7
  const int m = n8 >> 3;
  for (int g = lane; g < m; g += WAVE) {
    uint32_t v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < ES; ++j) {
      const uint8_t *pl = db + 8 * j * m + g;
      const uint32_t lo = (uint32_t)pl[0] | (uint32_t)pl[m] << 8 | (uint32_t)pl[2 * m] << 16 | (uint32_t)pl[3 * m] << 24;
      const uint32_t hi =
          (uint32_t)pl[4 * m] | (uint32_t)pl[5 * m] << 8 | (uint32_t)pl[6 * m] << 16 | (uint32_t)pl[7 * m] << 24;
      const uint64_t x = transpose8((uint64_t)hi << 32 | lo);
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] |= (uint32_t)((x >> (8 * c)) & 0xFF) << (8 * j);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) put<KIND>(b, out, e0 + 8 * g + c, v[c]);
  }
  // raw tail: little-endian elements
  for (int i = n8 + lane; i < n; i += WAVE) {
    const uint8_t *p = db + i * ES;
    uint32_t v = (uint32_t)p[0] | (uint32_t)p[1] << 8;
    if (ES == 4) v |= (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
    put<KIND>(b, out, e0 + i, v);
  }
}

int check_batch(const xpg_bslz4_batch *p) {
  if (!p || !p->src || !p->chunk_off || !p->chunk_len || !p->table || !p->status)
    return xpg::report_error(XPG_ERR_ARG, "bslz4 batch: missing pointer");
  if (p->esize != 2 && p->esize != 4) return xpg::report_error(XPG_ERR_ARG, "bslz4 batch: esize must be 2 or 4");
  if (p->n_frames < 0 || p->nel < 1 || p->block_elems < 8 || p->block_elems % 8)
    return xpg::report_error(XPG_ERR_ARG, "bslz4 batch: bad n_frames, nel or block_elems");
  if ((int64_t)p->block_elems * p->esize > MAX_BLOCK_BYTES)
    return xpg::report_error(XPG_ERR_ARG, "bslz4 batch: block larger than 64 KiB");
  if (p->n_blocks != xpg_bslz4_n_blocks(p->nel, p->block_elems))
    return xpg::report_error(XPG_ERR_ARG, "bslz4 batch: n_blocks != ceil(nel / block_elems)");
  return 0;
}

Batch make_batch(const xpg_bslz4_batch *p) {
  Batch b;
  b.src = p->src;
  b.coff = p->chunk_off;
  b.clen = p->chunk_len;
  b.nf = p->n_frames;
  b.esize = p->esize;
  b.nel = p->nel;
  b.belem = p->block_elems;
  b.nbf = p->n_blocks;
  b.table = p->table;
  b.status = p->status;
  b.out = p->out;
  b.ostride = p->out_stride;
  b.oslot = p->out_slot;
  b.use_min = p->use_min_valid;
  b.use_ovl = p->use_overload;
  b.min_ceil = p->min_valid_ceil;
  b.ovl_ceil = p->overload_ceil;
  b.sbad = p->static_bad;
  const int64_t bb = (int64_t)p->block_elems * p->esize;
  b.cap_c = (int32_t)((lz4_bound(bb) + 16 + 15) / 16 * 16);
  b.wave_lds = b.cap_c + (int32_t)((bb + 15) / 16 * 16);
  b.f0 = 0;
  return b;
}

}  // namespace

extern "C" {

int32_t xpg_bslz4_block_elems(uint32_t block_bytes, int32_t esize) {
  if (esize < 1) return 0;
  return block_elems_of(block_bytes, esize);
}

int32_t xpg_bslz4_n_blocks(int64_t nel, int32_t block_elems) {
  if (block_elems <= 0 || nel < 0) return 0;
  return (int32_t)((nel + block_elems - 1) / block_elems);
}

int32_t xpg_bslz4_index_host(const uint8_t *chunk, int64_t len, int64_t nel, int32_t esize, int32_t block_elems,
                             uint32_t *table) {
  if (!chunk || !table || (esize != 2 && esize != 4) || block_elems < 8 || nel < 1) return XPG_BSLZ4_BLOCK;
  return walk_chunk(chunk, len, nel, esize, block_elems, xpg_bslz4_n_blocks(nel, block_elems), table);
}

int xpg_bslz4_index(const xpg_bslz4_batch *p, void *stream) {
  if (int e = check_batch(p)) return e;
  if (p->n_frames == 0) return 0;
  Batch b = make_batch(p);
  hipLaunchKernelGGL(k_bslz4_index, dim3((p->n_frames + 63) / 64), dim3(64), 0, (hipStream_t)stream, b);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  return 0;
}

int xpg_bslz4_decode(const xpg_bslz4_batch *p, void *stream) {
  if (int e = check_batch(p)) return e;
  if (!p->out || (p->out_kind != 0 && p->out_kind != 1) || p->out_stride < p->nel)
    return xpg::report_error(XPG_ERR_ARG, "bslz4 decode: bad output");
  if (p->n_frames == 0) return 0;
  Batch b = make_batch(p);
  using K = void (*)(Batch);
  const K kern = p->esize == 4 ? (p->out_kind == 0 ? k_bslz4_decode<4, 0> : k_bslz4_decode<4, 1>)
                               : (p->out_kind == 0 ? k_bslz4_decode<2, 0> : k_bslz4_decode<2, 1>);
  if (b.wave_lds > 65536) {
    const hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             b.wave_lds);
    if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  }
  // one workgroup (one wavefront) per block; launches of at most 2^24 workgroups
  const int64_t per = std::max<int64_t>(1, ((int64_t)1 << 24) / b.nbf);
  for (int64_t f0 = 0; f0 < p->n_frames; f0 += per) {
    const int64_t nfr = std::min<int64_t>(per, p->n_frames - f0);
    b.f0 = f0;
    hipLaunchKernelGGL(kern, dim3((uint32_t)(nfr * b.nbf)), dim3(64), b.wave_lds, (hipStream_t)stream, b);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xpg::report_error(XPG_ERR_HIP, hipGetErrorString(e));
  }
  return 0;
}

int64_t xpg_bslz4_encode_host(const void *elems, int64_t nel, int32_t esize, int32_t block_bytes,
                              xpg_lz4_compress_fn compress, uint8_t *out, int64_t cap) {
  if (!elems || !compress || !out || (esize != 2 && esize != 4) || nel < 1 || block_bytes < 0 || cap < 12) return -1;
  const int32_t be = block_elems_of((uint32_t)block_bytes, esize);
  if (be < 8 || (int64_t)be * esize > MAX_BLOCK_BYTES) return -1;
  const uint64_t total = (uint64_t)nel * esize;
  for (int i = 0; i < 8; ++i) out[i] = (uint8_t)(total >> (56 - 8 * i));
  for (int i = 0; i < 4; ++i) out[8 + i] = (uint8_t)((uint32_t)block_bytes >> (24 - 8 * i));
  const uint8_t *e = static_cast<const uint8_t *>(elems);
  std::vector<uint8_t> sh((size_t)be * esize);
  int64_t pos = 12;
  for (int64_t done = 0; done < nel;) {
    const int64_t n = std::min<int64_t>(be, nel - done), n8 = n - n % 8, m = n8 / 8;
    if (n8 > 0) {
      // bit shuffle: plane 8 j + t, byte g = bit t of byte j of elements 8 g .. 8 g + 7 (the transpose of the
      // decoder's matrix)
      for (int64_t g = 0; g < m; ++g) {
        for (int j = 0; j < esize; ++j) {
          uint64_t x = 0;
          for (int c = 0; c < 8; ++c) x |= (uint64_t)e[(done + 8 * g + c) * esize + j] << (8 * c);
          x = transpose8(x);
          for (int tb = 0; tb < 8; ++tb) sh[(size_t)(8 * j + tb) * m + g] = (uint8_t)(x >> (8 * tb));
        }
      }
      if (pos + 4 > cap) return -1;
      const int64_t room = std::min<int64_t>(cap - pos - 4, INT32_MAX);
      const int c = compress(reinterpret_cast<const char *>(sh.data()), reinterpret_cast<char *>(out + pos + 4),
                             (int)(n8 * esize), (int)room);
      if (c <= 0) return -1;
      for (int i = 0; i < 4; ++i) out[pos + i] = (uint8_t)((uint32_t)c >> (24 - 8 * i));
      pos += 4 + c;
    }
    if (n8 < n) {
      const int64_t rest = (n - n8) * esize;
      if (pos + rest > cap) return -1;
      memcpy(out + pos, e + (done + n8) * esize, (size_t)rest);
      pos += rest;
    }
    done += n;
  }
  return pos;
}

}  // extern "C"
