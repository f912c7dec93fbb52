/* Device decode of CBF byte-offset streams (x-CBF_BYTE_OFFSET, the PILATUS / miniCBF format) for MI355X (gfx950).
 *
 * Replaces decode_byte_offset (xpg/io/cbf.py:12-44), called by read_cbf (:74-98) on the host for every frame
 * (CBFImageSet, xpg/io/images.py:213-235): a sequential walk over the stream that reads each pixel's difference to the previous one -- one signed byte, or after
 * the escape byte 0x80 a little-endian int16, or after 0x80 0x00 0x80 an int32, or after a further int32 escape
 * (0x80000000) an int64 -- and returns their running sum (int64, cast to int32 by read_cbf).
 *
 * On the device the walk is made parallel over 256-byte chunks of the stream: an element starting in a chunk begins
 * at one of 15 offsets past the previous chunk's boundary (an element is at most 15 bytes), so every chunk is parsed
 * speculatively from each of the 15 entry offsets (exit offset, element count, sum of differences), the maps are
 * composed per group of 64 chunks, one lane per frame walks its groups from entry 0 and one lane per group its chunks,
 * and each chunk is parsed once more from its true entry to write the running sums.
 *
 * Returns 0 or a negative XPG_ERR_* code (include/xpg_integrate.h; message from xpg_last_error()).  Per frame,
 * status = 1 when the stream holds fewer than n elements, or its last element runs past the stream's end (the
 * reference raises IndexError, or for a truncated int32 / int64 escape reads the bytes that are there).
 */
#ifndef XPG_CBF_H
#define XPG_CBF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPG_CBF_CHUNK 256     /* bytes per chunk */
#define XPG_CBF_GROUP 64      /* chunks per group */

typedef struct xpg_cbf_batch {
  const uint8_t *src;          /* device: the frames' streams */
  const int64_t *stream_off;   /* device (n_frames): start of frame f's stream in src */
  const int64_t *stream_len;   /* device (n_frames): its length in bytes */
  const int64_t *n_elem;       /* device (n_frames): elements (pixels) of frame f */
  const int64_t *chunk0;       /* device (n_frames + 1): first chunk of frame f; chunks = ceil(len / 256) */
  const int64_t *group0;       /* device (n_frames + 1): first group of frame f; groups = ceil(chunks / 64) */
  int32_t n_frames;
  int64_t n_chunks, n_groups;  /* chunk0[n_frames], group0[n_frames] */
  void *work;                  /* device workspace of xpg_cbf_workspace_bytes(n_chunks, n_groups) bytes */
  int32_t *out;                /* device: frame f's values at out + out_off[f] */
  const int64_t *out_off;      /* device (n_frames) */
  int32_t *status;             /* device (n_frames), written */
} xpg_cbf_batch;

int64_t xpg_cbf_workspace_bytes(int64_t n_chunks, int64_t n_groups);

int xpg_cbf_decode(const xpg_cbf_batch *b, void *stream);

/* xpg_cbf_decode without a caller workspace and without any allocation: no counterpart in the reference (it decodes
 * on the host); for callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a file-scope device array of XPG_CBF_STATIC_SCRATCH_BYTES
 * (64 MiB) on each GPU the library runs on, held until the process ends; b->work is ignored (may be NULL).  A batch
 * whose xpg_cbf_workspace_bytes exceeds it is rejected with XPG_ERR_TOO_LARGE (about 420,000 chunks of 256 bytes:
 * some 107 MB of byte-offset streams per call).  Every call on a GPU shares that GPU's array, so calls that overlap
 * on one GPU (on different streams, or from several threads) race: serialise them.  Nothing in the array needs to
 * start at zero: every call writes each workspace entry before reading it. */
#define XPG_CBF_STATIC_SCRATCH_BYTES (64LL << 20)
int xpg_cbf_decode_static(const xpg_cbf_batch *b, void *stream);

/* Host: encode_byte_offset (xpg/io/cbf.py:47-70) of n int32 values into out (at most 7 n bytes); returns the length.
 * The writer side of the format (write_cbf, :150-194); used to make workloads. */
int64_t xpg_cbf_encode_host(const int32_t *values, int64_t n, uint8_t *out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* XPG_CBF_H */
