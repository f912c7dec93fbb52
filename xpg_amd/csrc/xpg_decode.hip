// Bitshuffle-LZ4 frame decode on MI355X (gfx950) behind include/xpg_decode.h.
//
// The reference decodes in two places: host worker processes run the LZ4 stage (lz4_stage, xpg/io/bslz4.py:34-63)
// and the device undoes the bit shuffle and applies the value codes (unshuffle_frame_torch, xpg/io/bslz4.py:80-108).
// Here both run in one kernel over chunks resident in HBM:
//
//   k_bslz4_decode   one wavefront per LZ4 block (block_elems elements, 8 KB for the default block; BPW consecutive
//                    blocks per wavefront in turn), decoding into
//                    its LDS slice (the block and a 256-byte dummy area, so that 19 waves fit a CU).  The compressed
//                    bytes stay in registers: two 256-byte windows, lane l holding a dword of each, the next one loaded
//                    ahead.  The parse is wave-uniform (token, lengths and offset read from the window with readlane
//                    into scalar registers); a literal run is written by the lanes that hold its bytes, with no read; a
//                    match is copied through LDS in steps free of dependencies (short matches in one load and store).
//                    Then the bit shuffle is undone from LDS: for a full block lane l reads one dword per bit plane for
//                    its 4 groups of 8 elements, 4 x 4 byte transposes (v_perm) and one 8 x 8 bit transpose per output
//                    byte give the 32 elements, written with 16-byte stores (raw uint32 counts for the integrator's
//                    frame store, or value-coded float32); a last, shorter block goes one group per lane.
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
constexpr int DUMMY_PAD = 16;       // the dummy area starts 16 bytes past the block's bytes (dword aligned)
constexpr int DUMMY_BYTES = 256;

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
  int32_t wave_lds;         // LDS per wave: the decoded block, then the dummy area
  int64_t f0, f1;           // frames of this launch: [f0, f1)
};

__global__ void __launch_bounds__(64) k_bslz4_index(Batch b) {
  const int64_t f = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (f >= b.nf) return;
  const int32_t st = walk_chunk(b.src + b.coff[f], b.clen[f], b.nel, b.esize, b.belem, b.nbf,
                                b.table + f * (int64_t)(b.nbf + 1) * 2);
  b.status[f] = st;
}

// The compressed block as seen by the wave: two 256-byte windows in registers, lane l holding the dword at window
// offset 4 l.  `wb` is the current window's start relative to the block (the block's first byte rounded down to a
// dword, so wb starts in -3 .. 0); the next window is loaded ahead.  Bytes outside the chunk read as 0 and are never
// used (every read below stays inside the block).
struct Window {
  const uint8_t *blk, *lo, *hi;   // block start; the chunk's bytes [lo, hi)
  int wb;
  uint32_t cur, nxt;

  __device__ __forceinline__ uint32_t load(int base, int lane) const {
    const uint8_t *a = blk + base + 4 * lane;
    if (a >= lo && a + 4 <= hi) return *reinterpret_cast<const uint32_t *>(a);
    uint32_t v = 0;
    for (int b = 0; b < 4; ++b)
      if (a + b >= lo && a + b < hi) v |= (uint32_t)a[b] << (8 * b);
    return v;
  }
  __device__ __forceinline__ void init(const uint8_t *block, const uint8_t *chunk_lo, const uint8_t *chunk_hi,
                                       int lane) {
    blk = block;
    lo = chunk_lo;
    hi = chunk_hi;
    wb = -(int)((uintptr_t)block & 3);
    cur = load(wb, lane);
    nxt = load(wb + 256, lane);
  }
  // make position q (uniform, q >= wb) part of the current window
  __device__ __forceinline__ void reach(int q, int lane) {
    while (q >= wb + 256) {
      cur = nxt;
      wb += 256;
      nxt = load(wb + 256, lane);
    }
  }
  // bytes q .. q + 4 as the low 40 bits, after reach(q) (so wb <= q < wb + 256): dword d of the current window and
  // the next dword (lane 0 of the next window for d = 63), both read and one selected
  __device__ __forceinline__ uint64_t at5(int q) const {
    const int i = q - wb;
    const int d = i >> 2;
    const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)cur, d);
    const uint32_t a1 = (uint32_t)__builtin_amdgcn_readlane((int)cur, (d + 1) & 63);
    const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)nxt, 0);
    const uint32_t w1 = d == 63 ? b1 : a1;
    return ((uint64_t)w1 << 32 | w0) >> (8 * (i & 3));
  }
  // byte at position q, wb <= q < wb + 512 (uniform)
  __device__ __forceinline__ int at(int q) const {
    const int i = q - wb;
    const uint32_t w = i < 256 ? (uint32_t)__builtin_amdgcn_readlane((int)cur, i >> 2)
                               : (uint32_t)__builtin_amdgcn_readlane((int)nxt, (i - 256) >> 2);
    return (int)((w >> (8 * (i & 3))) & 0xFF);
  }
};

// Extended length (LZ4 block format: 255-valued bytes add up until a smaller byte).  The reading stops after the byte
// that brings ip to `lim` (LZ4_decompress_safe's bound checks); `stopped` reports that.
__device__ __forceinline__ int ext_len(Window &w, int &ip, int lim, bool &stopped, int lane) {
  int add = 0;
  for (;;) {
    w.reach(ip, lane);
    const int v = w.at(ip);
    ++ip;
    add += v;
    if (ip >= lim) {
      stopped = true;
      return add;
    }
    if (v != 255) return add;
  }
}

// Literal run [ip, ip + lit) of the block to db[op ..]: every lane writes the bytes of its own window dword that fall
// in the run (the others store to their own byte of the dummy area), one window at a time.  Leaves the window at
// ip + lit.
__device__ __forceinline__ void literals(Window &w, int ip, int lit, uint8_t *db, int op, int dummy, int lane) {
  const int end = ip + lit;
  while (ip < end) {
    w.reach(ip, lane);
    const int p0 = w.wb + 4 * lane;          // position of this lane's first byte
    const int seg_end = min(end, w.wb + 256);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int p = p0 + b;
      db[(p >= ip && p < seg_end) ? op + (p - ip) : dummy + lane] = (uint8_t)(w.cur >> (8 * b));
    }
    op += seg_end - ip;
    ip = seg_end;
  }
}

// Match: db[op + j] = db[op - off + j] for j = 0 .. ml - 1 in order (an overlap repeats the period), off <= op.
// Byte op + j depends on byte op + j - off, so steps of up to off bytes are free of dependencies inside the step;
// LDS executes the wave's instructions in order, so each step reads what the earlier steps wrote.  Loops have uniform
// trip counts and every lane stores each step: lanes with nothing to copy write their own slot of the dummy area
// at `dummy` (256 bytes after the block), so no store needs an exec-mask branch.
__device__ __forceinline__ void match_copy(uint8_t *db, int op, int off, int ml, int dummy, int lane) {
  wave_sync();
  if (ml <= WAVE && off >= ml) {
    // the common case: at most 64 bytes, no overlap -- one load and one store per lane
    const uint8_t v = db[op - off + min(lane, ml - 1)];
    db[lane < ml ? op + lane : dummy + lane] = v;
  } else if (off < WAVE) {
    // period below 64: lane i < T = off * floor(64 / off) writes bytes i, i + T, ... with pattern byte i mod off
    // (floor((x + 0.5) / off) for integers x, off < 65 is exact in float: the quotient is at least 0.5 / 64 from an
    // integer)
    const float inv = 1.0f / (float)off;
    const int T = off * (int)(64.5f * inv);
    const int r = lane - off * (int)(((float)lane + 0.5f) * inv);
    const uint8_t v = db[op - off + r];
    for (int base = 0; base < ml; base += T) {
      const int j = base + lane;
      db[(lane < T && j < ml) ? op + j : dummy + lane] = v;
    }
  } else if ((off & 3) == 0) {
    // whole dwords (source and destination share the alignment): up to 64 dwords per step, at most off / 4
    const int h = min(ml, (4 - (op & 3)) & 3);
    {
      const uint8_t v = db[op - off + min(lane, 3)];
      db[lane < h ? op + lane : dummy + lane] = v;
    }
    const int nw = (ml - h) >> 2;
    const int D = min(off >> 2, WAVE);
    const int a0 = op + h;   // dword aligned, as is a0 - off
    for (int s0 = 0; s0 < nw; s0 += D) {
      wave_sync();
      const int i = s0 + lane;
      const bool ok = lane < D && i < nw;
      const uint32_t v = *reinterpret_cast<const uint32_t *>(db + (ok ? a0 - off + 4 * i : 0));
      *reinterpret_cast<uint32_t *>(db + (ok ? a0 + 4 * i : dummy + 4 * lane)) = v;
    }
    wave_sync();
    const int t0 = h + 4 * nw;
    const bool ok = lane < ml - t0;
    const uint8_t v = db[ok ? op + t0 + lane - off : 0];
    db[ok ? op + t0 + lane : dummy + lane] = v;
  } else {
    // bytes: steps of min(off, 256), up to 4 bytes per lane, loads before stores
    const int S = min(off, 4 * WAVE);
    for (int s0 = 0; s0 < ml; s0 += S) {
      wave_sync();
      uint8_t v[4];
      bool ok[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int j = s0 + lane + WAVE * b;
        ok[b] = lane + WAVE * b < S && j < ml;
        v[b] = db[ok[b] ? op + j - off : 0];
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) db[ok[b] ? op + s0 + lane + WAVE * b : dummy + 4 * lane + b] = v[b];
    }
  }
  wave_sync();
}

// One LZ4 block of cs bytes (window over the chunk) into exactly dn bytes (db, LDS), by the whole wave; 0 or
// XPG_BSLZ4_LZ4.  Rules of the LZ4 block format as LZ4_decompress_safe applies them: literals that reach the last
// MFLIMIT output bytes or the last 8 input bytes end the block (they must consume the input exactly); a match must lie
// in the decoded output and end before the last LASTLITERALS bytes; offset 0 is rejected.
__device__ int lz4_block(Window &w, int cs, uint8_t *db, int dn, int lane) {
  int ip = 0, op = 0;
  // the parse reads through a 5-byte cache: the read at a match's offset also holds the next token
  int px = -8;
  uint64_t x = 0;
  auto get = [&](int q) -> int {
    if (q - px > 4) {
      w.reach(q, lane);
      x = w.at5(q);
      px = q;
    }
    return (int)((x >> (8 * (q - px))) & 0xFF);
  };
  for (;;) {
    if (ip >= cs) return XPG_BSLZ4_LZ4;
    const int tok = get(ip);
    ++ip;
    int lit = tok >> 4;
    if (lit == 15) {
      if (ip >= cs - 15) return XPG_BSLZ4_LZ4;
      bool stopped = false;
      lit += ext_len(w, ip, cs - 15, stopped, lane);
    }
    if (lit > cs - ip) return XPG_BSLZ4_LZ4;
    const bool last = lit > dn - MFLIMIT - op || lit > cs - 8 - ip;
    if (last && (ip + lit != cs || lit > dn - op)) return XPG_BSLZ4_LZ4;
    if (lit) literals(w, ip, lit, db, op, dn + DUMMY_PAD, lane);
    ip += lit;
    op += lit;
    if (last) break;
    const int off = get(ip) | get(ip + 1) << 8;
    ip += 2;
    int ml = tok & 15;
    if (ml == 15) {
      bool stopped = false;
      ml += ext_len(w, ip, cs - LASTLITERALS + 1, stopped, lane);
      if (stopped) return XPG_BSLZ4_LZ4;
    }
    ml += MINMATCH;
    if (off == 0 || off > op || ml > dn - LASTLITERALS - op) return XPG_BSLZ4_LZ4;
    match_copy(db, op, off, ml, dn + DUMMY_PAD, lane);
    op += ml;
  }
  wave_sync();
  return op == dn ? 0 : XPG_BSLZ4_LZ4;
}

template <int KIND>
__device__ __forceinline__ void put(const Batch &b, void *out, int64_t e, uint32_t v) {
  if constexpr (KIND == 0) {
    static_cast<uint32_t *>(out)[e] = v;
  } else {
    static_cast<float *>(out)[e] = __builtin_bit_cast(float, v);
  }
}

// unshuffle_frame_torch's codes (xpg/io/bslz4.py:100-108): all ones (2^n - 1) -> -1, below min_valid -> -1, static
// mask -> -1, at or above the overload -> -2; -1 wins.  Returns the float32's bits.
__device__ __forceinline__ uint32_t coded(const Batch &b, int64_t e, uint32_t v) {
  const uint32_t ones = b.esize == 4 ? 0xFFFFFFFFu : 0xFFFFu;
  bool bad = v == ones;
  if (b.use_min) bad |= (int64_t)v < b.min_ceil;
  if (b.sbad) bad |= b.sbad[e] != 0;
  float f = (float)v;
  if (b.use_ovl && (int64_t)v >= b.ovl_ceil) f = -2.0f;
  return __builtin_bit_cast(uint32_t, bad ? -1.0f : f);
}

// 4 x 4 byte transpose: t[c] = byte c of w0, w1, w2, w3 (v_perm_b32: selector bytes 0-3 pick from the second
// operand, 4-7 from the first)
__device__ __forceinline__ void bt4(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t t[4]) {
  const uint32_t x01 = __builtin_amdgcn_perm(w1, w0, 0x05010400u), y01 = __builtin_amdgcn_perm(w1, w0, 0x07030602u);
  const uint32_t x23 = __builtin_amdgcn_perm(w3, w2, 0x05010400u), y23 = __builtin_amdgcn_perm(w3, w2, 0x07030602u);
  t[0] = __builtin_amdgcn_perm(x23, x01, 0x05040100u);
  t[1] = __builtin_amdgcn_perm(x23, x01, 0x07060302u);
  t[2] = __builtin_amdgcn_perm(y23, y01, 0x05040100u);
  t[3] = __builtin_amdgcn_perm(y23, y01, 0x07060302u);
}

// Elements 32 q .. 32 q + 31 of a block whose planes hold m bytes, m a multiple of 4: one dword per plane gives the
// plane's bytes of groups 4 q .. 4 q + 3.
template <int ES>
__device__ __forceinline__ void unshuffle32(const uint8_t *db, int m, int q, uint32_t v[32]) {
  uint32_t X[ES][4][2];   // [output byte j][group c][half]: byte e of the 64-bit word = byte j of element 8 c + e
#pragma unroll
  for (int j = 0; j < ES; ++j) {
    uint32_t w[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) w[r] = *reinterpret_cast<const uint32_t *>(db + (8 * j + r) * m + 4 * q);
    uint32_t lo[4], hi[4];
    bt4(w[0], w[1], w[2], w[3], lo);
    bt4(w[4], w[5], w[6], w[7], hi);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint64_t x = transpose8((uint64_t)hi[c] << 32 | lo[c]);
      X[j][c][0] = (uint32_t)x;
      X[j][c][1] = (uint32_t)(x >> 32);
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t t[4];
      if constexpr (ES == 4) {
        bt4(X[0][c][h], X[1][c][h], X[2][c][h], X[3][c][h], t);
      } else {
        t[0] = __builtin_amdgcn_perm(X[1][c][h], X[0][c][h], 0x0C0C0400u);
        t[1] = __builtin_amdgcn_perm(X[1][c][h], X[0][c][h], 0x0C0C0501u);
        t[2] = __builtin_amdgcn_perm(X[1][c][h], X[0][c][h], 0x0C0C0602u);
        t[3] = __builtin_amdgcn_perm(X[1][c][h], X[0][c][h], 0x0C0C0703u);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) v[8 * c + 4 * h + e] = t[e];
    }
}

constexpr int BPW = 4;   // blocks per wavefront

template <int ES, int KIND>
__device__ __forceinline__ void decode_block(const Batch &b, uint8_t *lds, int64_t f, int k, int lane);

template <int ES, int KIND>
__global__ void __launch_bounds__(64) k_bslz4_decode(Batch b) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  // BPW consecutive blocks per wavefront, in turn
  const int64_t gw0 = b.f0 * b.nbf + (int64_t)blockIdx.x * BPW;
  const int64_t f = gw0 / b.nbf;
  const int k0 = (int)(gw0 - f * b.nbf);
  for (int t = 0; t < BPW; ++t) {
    int64_t fb = f;
    int k = k0 + t;
    if (k >= b.nbf) {
      fb += k / b.nbf;
      k %= b.nbf;
    }
    if (fb >= b.f1) break;   // past this launch's frames
    decode_block<ES, KIND>(b, lds, fb, k, lane);
  }
}

template <int ES, int KIND>
__device__ __forceinline__ void decode_block(const Batch &b, uint8_t *lds, int64_t f, int k, int lane) {
  if (f >= b.nf || b.status[f] != 0) return;
  wave_sync();   // the previous block's LDS reads precede this block's writes
  uint8_t *db = lds;
  const uint32_t *t = b.table + f * (int64_t)(b.nbf + 1) * 2;
  const uint8_t *chunk = b.src + b.coff[f];
  const int64_t e0 = (int64_t)k * b.belem;
  const int n = (int)std::min<int64_t>(b.belem, b.nel - e0);
  const int n8 = n - n % 8;
  if (n8 > 0) {
    Window w;
    w.init(chunk + t[2 * k], chunk, chunk + b.clen[f], lane);
    const int err = lz4_block(w, (int)t[2 * k + 1], db, n8 * ES, lane);
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
  uint32_t *out = reinterpret_cast<uint32_t *>(static_cast<uint8_t *>(b.out) + slot * b.ostride * 4) + e0;
  const bool vec = ((uintptr_t)out & 15) == 0;
  // bit-shuffled part: plane p (bit p of every element) holds m = n8 / 8 bytes, element i's bit at byte i / 8, bit
  // i % 8; the 8 planes 8 j .. 8 j + 7 of byte g form the 8 x 8 matrix of output byte j of elements 8 g .. 8 g + 7
  const int m = n8 >> 3;
  if ((m & 3) == 0) {
    for (int q = lane; 4 * q < m; q += WAVE) {
      uint32_t v[32];
      unshuffle32<ES>(db, m, q, v);
      if constexpr (KIND == 1) {
#pragma unroll
        for (int i = 0; i < 32; ++i) v[i] = coded(b, e0 + 32 * q + i, v[i]);
      }
      if (vec) {
        uint4 *o = reinterpret_cast<uint4 *>(out + 32 * q);
#pragma unroll
        for (int i = 0; i < 8; ++i) o[i] = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
      } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) out[32 * q + i] = v[i];
      }
    }
  } else
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
    if constexpr (KIND == 1) {
#pragma unroll
      for (int c = 0; c < 8; ++c) v[c] = coded(b, e0 + 8 * g + c, v[c]);
    }
    if (vec) {
      uint4 *o = reinterpret_cast<uint4 *>(out + 8 * g);
      o[0] = make_uint4(v[0], v[1], v[2], v[3]);
      o[1] = make_uint4(v[4], v[5], v[6], v[7]);
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) out[8 * g + c] = v[c];
    }
  }
  // raw tail: little-endian elements
  for (int i = n8 + lane; i < n; i += WAVE) {
    const uint8_t *p = db + i * ES;
    uint32_t v = (uint32_t)p[0] | (uint32_t)p[1] << 8;
    if (ES == 4) v |= (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
    out[i] = KIND == 1 ? coded(b, e0 + i, v) : v;
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
  b.wave_lds = (int32_t)((bb + DUMMY_PAD + DUMMY_BYTES + 15) / 16 * 16);
  b.f0 = 0;
  b.f1 = p->n_frames;
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
    b.f1 = f0 + nfr;
    hipLaunchKernelGGL(kern, dim3((uint32_t)((nfr * b.nbf + BPW - 1) / BPW)), dim3(64), b.wave_lds,
                       (hipStream_t)stream, b);
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
