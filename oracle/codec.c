/* Oracle (test infrastructure only): CPU restatement of the bitshuffle-LZ4 and CBF byte-offset frame decodes.
 *
 * Only the tests, the smoke check and bench.py's CPU-baseline leg load this library, as the checker; the product
 * decodes are xpg_amd/csrc/xpg_decode.hip and xpg_amd/csrc/xpg_cbf.hip.
 *
 *   lz4_block_decode   the LZ4 block format (lz4 1.9.3, doc/lz4_Block_format.md), with the end-of-block rules of
 *                      LZ4_decompress_safe, the routine python-lz4's lz4.block.decompress calls for the reference's
 *                      LZ4 stage (xpg/io/bslz4.py:57-58).  The third-party library is not vendored by the reference
 *                      (no lock file pins it); liblz4 1.9.3 is the system library of this image and the tests check this
 *                      restatement against it.
 *   bslz4_decode_chunk lz4_stage's chunk walk (xpg/io/bslz4.py:34-63) followed by the bit unshuffle of
 *                      _bitunshuffle / unshuffle_frame_torch (xpg/io/bslz4.py:28-31, :80-99), written bit by bit.
 *   bslz4_codes        the float32 value codes of unshuffle_frame_torch (xpg/io/bslz4.py:100-108).
 *   cbf_decode         decode_byte_offset (xpg/io/cbf.py:12-44): byte-offset differences (int8, or after the escape
 *                      0x80 an int16, int32, int64) summed in int64 and stored as int32 (read_cbf, :97).
 *
 * Parity pins: hand-built LZ4 blocks (literal-only, offset-1 runs, overlapping periodic matches, extended lengths,
 * each error rule), liblz4's compressor and decoder on random and bit-shuffled data (including LZ4_decompress_safe's
 * accept / reject decision on corrupted blocks), numpy's unpackbits / packbits for the bit transpose, and closed forms
 * for the codes; cbf_decode by hand-built streams of each escape class with the values the format defines, int32 wrap
 * of the int64 sum, cut and short streams, and the reference test's vector (tests/test_core.py:15-22 of the
 * reference) through the product's encoder.
 */
#include <stdint.h>
#include <string.h>
#include <limits.h>

/* Decode one LZ4 block of cs bytes into exactly dn bytes.  Returns dn, -1 for a malformed block, -2 for a match of
 * offset 0 (the format calls it invalid; LZ4_decompress_safe 1.9.3 does not test for it). */
int lz4_block_decode(const uint8_t *s, int cs, uint8_t *d, int dn) {
  int ip = 0, op = 0;
  for (;;) {
    if (ip >= cs) return -1;
    const int tok = s[ip++];
    int lit = tok >> 4;
    if (lit == 15) {
      if (ip >= cs - 15) return -1;               /* length bytes must leave room for the literals */
      int b;
      do {
        b = s[ip++];
        lit += b;
        if (ip >= cs - 15) break;                 /* LZ4_decompress_safe stops reading the length here */
      } while (b == 255);
    }
    if (lit > cs - ip) return -1;
    /* literals reaching the last 12 output bytes or the last 8 input bytes end the block */
    if (lit > dn - 12 - op || lit > cs - 8 - ip) {
      if (ip + lit != cs || lit > dn - op) return -1;
      memcpy(d + op, s + ip, (size_t)lit);
      op += lit;
      break;
    }
    memcpy(d + op, s + ip, (size_t)lit);
    ip += lit;
    op += lit;
    const int off = s[ip] | s[ip + 1] << 8;
    ip += 2;
    int ml = tok & 15;
    if (ml == 15) {
      int b;
      do {
        b = s[ip++];
        ml += b;
        if (ip >= cs - 4) return -1;              /* a match needs the final literals after it */
      } while (b == 255);
    }
    ml += 4;
    if (off == 0) return -2;
    if (off > op) return -1;                      /* before the start of the block */
    if (ml > dn - 5 - op) return -1;              /* the last 5 output bytes are literals */
    for (int j = 0; j < ml; ++j) d[op + j] = d[op - off + j];   /* byte by byte: an overlap repeats the period */
    op += ml;
  }
  return op == dn ? op : -1;
}

static uint32_t rd_be32(const uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

/* Undo the bit shuffle of n elements (n a multiple of 8) of esize bytes: plane p = 8 j + t holds bit t of byte j of
 * every element, element i's bit at byte i / 8, bit i % 8 (numpy's little bit order). */
void bitunshuffle(const uint8_t *in, int64_t n, int esize, uint8_t *out) {
  const int64_t m = n / 8;
  for (int64_t i = 0; i < n; ++i)
    for (int j = 0; j < esize; ++j) {
      unsigned byte = 0;
      for (int t = 0; t < 8; ++t) byte |= (unsigned)((in[(int64_t)(8 * j + t) * m + i / 8] >> (i % 8)) & 1) << t;
      out[i * esize + j] = (uint8_t)byte;
    }
}

/* Decode a chunk of nel elements into out (nel * esize bytes, little-endian elements).  Returns the block size in
 * elements (lz4_stage's return value) or a negative code: -1 header / size / block size, -2 truncated chunk or
 * compressed size beyond the LZ4 bound, -3 malformed LZ4 block, -4 LZ4 match of offset 0.  scratch: block_bytes. */
int64_t bslz4_decode_chunk(const uint8_t *c, int64_t len, int64_t nel, int esize, uint8_t *out, uint8_t *scratch,
                           int64_t scratch_bytes) {
  if (len < 12) return -1;
  const uint64_t total = (uint64_t)rd_be32(c) << 32 | rd_be32(c + 4);
  uint32_t block = rd_be32(c + 8);
  if (block == 0) block = 8192;
  if (total != (uint64_t)nel * (uint64_t)esize) return -1;
  int64_t belem = block / esize;
  belem -= belem % 8;
  if (belem <= 0 || belem * esize > scratch_bytes) return -1;
  int64_t pos = 12, done = 0;
  while (done < nel) {
    const int64_t n = nel - done < belem ? nel - done : belem;
    const int64_t n8 = n - n % 8;
    if (n8 > 0) {
      if (pos + 4 > len) return -2;
      const int64_t cs = rd_be32(c + pos);
      pos += 4;
      if (cs > n8 * esize + n8 * esize / 255 + 16 || pos + cs > len) return -2;
      const int r = lz4_block_decode(c + pos, (int)cs, scratch, (int)(n8 * esize));
      if (r == -2) return -4;
      if (r < 0) return -3;
      bitunshuffle(scratch, n8, esize, out + done * esize);
      pos += cs;
    }
    if (n8 < n) {
      const int64_t rest = (n - n8) * esize;
      if (pos + rest > len) return -2;
      memcpy(out + (done + n8) * esize, c + pos, (size_t)rest);
      pos += rest;
    }
    done += n;
  }
  return belem;
}

/* Value codes: counts v (unsigned, esize bytes) to float32, v = 2^(8 esize) - 1 -> -1, v < min_valid_ceil (when
 * use_min) -> -1, static_bad -> -1, v >= overload_ceil (when use_ovl) -> -2; -1 wins. */
void bslz4_codes(const uint8_t *bytes, int64_t nel, int esize, int use_min, int64_t min_ceil, int use_ovl,
                 int64_t ovl_ceil, const uint8_t *static_bad, float *out) {
  const uint64_t ones = esize == 4 ? 0xFFFFFFFFull : 0xFFFFull;
  for (int64_t i = 0; i < nel; ++i) {
    uint64_t v = 0;
    for (int j = 0; j < esize; ++j) v |= (uint64_t)bytes[i * esize + j] << (8 * j);
    int bad = v == ones;
    if (use_min && (int64_t)v < min_ceil) bad = 1;
    if (static_bad && static_bad[i]) bad = 1;
    float f = (float)v;
    if (use_ovl && (int64_t)v >= ovl_ceil) f = -2.0f;
    out[i] = bad ? -1.0f : f;
  }
}

/* CBF byte offset: n values from a stream of len bytes.  Returns 0, or 1 when the stream ends before n elements or
 * inside the n-th (the reference raises IndexError there, or for a cut int32 / int64 escape reads the bytes that are
 * there). */
int cbf_decode(const uint8_t *b, int64_t len, int64_t n, int32_t *out) {
  int64_t pos = 0, sum = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (pos >= len) return 1;
    int64_t d;
    if (b[pos] != 0x80) {
      d = (int8_t)b[pos];
      pos += 1;
    } else {
      if (pos + 3 > len) return 1;
      const int v16 = b[pos + 1] | b[pos + 2] << 8;
      if (v16 != 0x8000) {
        d = (int16_t)v16;
        pos += 3;
      } else {
        if (pos + 7 > len) return 1;
        const int32_t v32 = (int32_t)((uint32_t)b[pos + 3] | (uint32_t)b[pos + 4] << 8 | (uint32_t)b[pos + 5] << 16 |
                                      (uint32_t)b[pos + 6] << 24);
        if (v32 != INT32_MIN) {
          d = v32;
          pos += 7;
        } else {
          if (pos + 15 > len) return 1;
          uint64_t v = 0;
          for (int k = 0; k < 8; ++k) v |= (uint64_t)b[pos + 7 + k] << (8 * k);
          d = (int64_t)v;
          pos += 15;
        }
      }
    }
    sum += d;
    out[i] = (int32_t)(uint32_t)(uint64_t)sum;
  }
  return 0;
}
