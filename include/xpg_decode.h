/* Device decode of bitshuffle-LZ4 frame chunks (HDF5 filter 32008, the Eiger / NeXus format) into the device frame
 * store, for MI355X (gfx950).
 *
 * Replaces the reference's two-stage decode: the LZ4 stage on host worker processes (lz4_stage,
 * xpg/io/bslz4.py:34-63, run by _worker_bitshuffled, xpg/io/images.py:466-477) followed by the device bit unshuffle and
 * value codes (unshuffle_frame_torch, xpg/io/bslz4.py:80-108, called from xpg/spotfind.py:243-248).  Here the chunks'
 * compressed bytes are resident in HBM and both stages run on the device: one wavefront per LZ4 block stages the block
 * in LDS, decodes it, undoes the bit shuffle and writes the frame's pixels once.
 *
 * Chunk layout (as lz4_stage reads it): 8-byte big-endian uncompressed size, 4-byte big-endian block size in bytes
 * (0 = 8192), then per block of block_elems = (block bytes / esize) rounded down to a multiple of 8 elements: a 4-byte
 * big-endian compressed size and an LZ4 block holding the block's n8 = n - n % 8 elements bit-shuffled; the last
 * n % 8 elements of the last block are stored raw after it.
 *
 * Block table: the blocks' positions form a chain through the chunk (each size field locates the next), so they are
 * found once per chunk, on the host while the chunk is in host memory (xpg_bslz4_index_host) or on the device with one
 * lane per frame (xpg_bslz4_index), into a table of (n_blocks + 1) pairs of uint32 per frame: (offset of the block's
 * LZ4 data from the chunk start, compressed size) per block, (0, 0) for a last block with no full group of 8
 * elements, then (offset, bytes) of the raw tail (0, 0 when there is none).
 *
 * Every entry point returns 0 or a negative XPG_ERR_* code (include/xpg_integrate.h; message from xpg_last_error()).
 * Malformed chunks are reported per frame in `status` (XPG_BSLZ4_*), never read or written out of bounds.
 */
#ifndef XPG_DECODE_H
#define XPG_DECODE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* per-frame status codes (0 = decoded) */
enum {
  XPG_BSLZ4_OK = 0,
  XPG_BSLZ4_HEADER = 1,     /* chunk shorter than its 12-byte header */
  XPG_BSLZ4_SIZE = 2,       /* uncompressed size is not nel * esize (lz4_stage: ValueError) */
  XPG_BSLZ4_BLOCK = 4,      /* block size differs from the batch's block_elems (or gives 0 elements) */
  XPG_BSLZ4_TRUNCATED = 8,  /* a block or the raw tail runs past the chunk's end */
  XPG_BSLZ4_CSIZE = 16,     /* compressed size beyond the LZ4 bound of the block */
  XPG_BSLZ4_LZ4 = 32        /* malformed LZ4 block, or one that does not decode to exactly the block's bytes */
};

/* One batch of frame chunks, all of nel elements of esize (2 or 4) bytes and one bitshuffle block size. */
typedef struct xpg_bslz4_batch {
  const uint8_t *src;        /* device: the chunks' bytes */
  const int64_t *chunk_off;  /* device (n_frames): start of each frame's chunk in src */
  const int64_t *chunk_len;  /* device (n_frames): its length in bytes (< 2^32) */
  int32_t n_frames;
  int32_t esize;
  int64_t nel;
  int32_t block_elems;       /* elements per block, from the chunks' header: xpg_bslz4_block_elems */
  int32_t n_blocks;          /* xpg_bslz4_n_blocks(nel, block_elems) */
  uint32_t *table;           /* device (n_frames, n_blocks + 1, 2) */
  int32_t *status;           /* device (n_frames): per-frame XPG_BSLZ4_* bits; frames with status != 0 are skipped */
  /* output: frame f goes to out + slot * out_stride elements, slot = out_slot ? out_slot[f] : f */
  void *out;
  int32_t out_kind;          /* 0: raw counts as uint32 (the integrator's frame store); 1: value-coded float32 */
  int64_t out_stride;
  const int64_t *out_slot;   /* device (n_frames) or NULL */
  /* value codes of out_kind 1 (unshuffle_frame_torch, xpg/io/bslz4.py:100-108): all-ones -> -1; v < min_valid_ceil
   * (when use_min_valid) -> -1; static_bad[pixel] != 0 -> -1; v >= overload_ceil (when use_overload) -> -2; -1 wins */
  int32_t use_min_valid, use_overload;
  int64_t min_valid_ceil, overload_ceil;
  const uint8_t *static_bad; /* device (nel) or NULL */
} xpg_bslz4_batch;

/* block_elems of a chunk header's block size field (bytes; 0 = 8192), as lz4_stage computes it
 * (xpg/io/bslz4.py:41-45); 0 when the block holds fewer than 8 elements. */
int32_t xpg_bslz4_block_elems(uint32_t block_bytes, int32_t esize);

/* blocks per frame: ceil(nel / block_elems) */
int32_t xpg_bslz4_n_blocks(int64_t nel, int32_t block_elems);

/* Host: walk one chunk in host memory (lz4_stage's parse, xpg/io/bslz4.py:34-63) into its table rows
 * (n_blocks + 1 pairs).  Returns its XPG_BSLZ4_* status (0 = well formed). */
int32_t xpg_bslz4_index_host(const uint8_t *chunk, int64_t len, int64_t nel, int32_t esize, int32_t block_elems,
                             uint32_t *table);

/* Device: the same walk for every chunk of the batch, one lane per frame; writes table and status. */
int xpg_bslz4_index(const xpg_bslz4_batch *b, void *stream);

/* Device: decode every block of the batch (frames with status 0) into out; LZ4 errors are or-ed into status. */
int xpg_bslz4_decode(const xpg_bslz4_batch *b, void *stream);

/* Host: encode one frame (nel elements of esize bytes) as a bitshuffle-LZ4 chunk with the given block size in bytes,
 * compressing each block with `compress` (the LZ4 library's LZ4_compress_default signature: src, dst, src_size,
 * dst_capacity -> compressed size or <= 0).  Returns the chunk's length, or -1 when cap is too small or a block fails.
 * The writer side of the format (the reference's simulator writes frames through hdf5plugin's Bitshuffle filter,
 * xpg/simulate.py:186-192); used to make workloads. */
typedef int (*xpg_lz4_compress_fn)(const char *src, char *dst, int src_size, int dst_capacity);
int64_t xpg_bslz4_encode_host(const void *elems, int64_t nel, int32_t esize, int32_t block_bytes,
                              xpg_lz4_compress_fn compress, uint8_t *out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* XPG_DECODE_H */
