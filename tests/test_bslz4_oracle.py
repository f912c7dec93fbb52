"""Oracle of the bitshuffle-LZ4 frame decode (oracle/codec.c) pinned without a GPU:

- hand-built LZ4 blocks from the block format (literal-only, offset-1 runs, overlapping periodic matches, extended
  literal and match lengths) and each end-of-block rule, also decoded by the system's liblz4 1.9.3 (the library
  python-lz4 wraps for the reference's lz4_stage);
- liblz4's compressors (fast and high-compression) on random, run-heavy and bit-shuffled Poisson data, and
  LZ4_decompress_safe's accept / reject decision on corrupted blocks;
- the bit transpose against numpy's unpackbits / packbits (the reference's _bitunshuffle, xpg/io/bslz4.py:28-31);
- the chunk walk and the product's host encoder / block index on the reference test's shapes (a 37 x 61 frame: one
  2048-element block + 208 + 1, uint32 and uint16), block sizes, truncated and mislabelled chunks;
- the float32 codes against ImageSet.get's rules (xpg/io/images.py:80-96 restated on raw counts)."""
import ctypes as C

import numpy as np
import pytest

from oracle import codec as oc

try:
    _LZ4 = C.CDLL("liblz4.so.1")
    for _n in ("LZ4_compress_default", "LZ4_decompress_safe"):
        getattr(_LZ4, _n).restype = C.c_int
    _LZ4.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    _LZ4.LZ4_compress_HC.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int]
    _LZ4.LZ4_compress_HC.restype = C.c_int
    _LZ4.LZ4_decompress_safe.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
except OSError:
    _LZ4 = None

need_lz4 = pytest.mark.skipif(_LZ4 is None, reason="liblz4.so.1 not installed")


def lz4_compress(data: bytes, hc: int = 0) -> bytes:
    cap = len(data) + len(data) // 255 + 16
    dst = C.create_string_buffer(cap)
    n = (_LZ4.LZ4_compress_HC(data, dst, len(data), cap, hc) if hc else
         _LZ4.LZ4_compress_default(data, dst, len(data), cap))
    assert n > 0
    return dst.raw[:n]


def lz4_safe(block: bytes, n: int):
    dst = C.create_string_buffer(max(n, 1))
    r = _LZ4.LZ4_decompress_safe(block, dst, len(block), n)
    return r, dst.raw[:n]


def seq(lit: bytes, off=None, ml=None) -> bytes:
    """One LZ4 sequence: token, extended lengths, literals, offset (little-endian)."""
    ll = len(lit)
    tok_l = min(ll, 15)
    out = bytearray()
    tok_m = 0 if ml is None else min(ml - 4, 15)
    out.append(tok_l << 4 | tok_m)
    if ll >= 15:
        r = ll - 15
        while r >= 255:
            out.append(255)
            r -= 255
        out.append(r)
    out += lit
    if off is not None:
        out += bytes([off & 255, off >> 8])
        if ml - 4 >= 15:
            r = ml - 4 - 15
            while r >= 255:
                out.append(255)
                r -= 255
            out.append(r)
    return bytes(out)


# (block, expected output or None for a malformed block)
KNOWN = [
    (seq(b"hello"), b"hello"),                                                          # literal-only block
    (seq(b"a", 1, 19) + seq(b"bcdef"), b"a" * 20 + b"bcdef"),                           # offset-1 run
    (seq(b"abc", 3, 9) + seq(b"XYZWV"), b"abc" * 4 + b"XYZWV"),                          # periodic overlap, period 3
    (seq(bytes(range(100)), 100, 300) + seq(b"12345"),
     bytes(range(100)) * 4 + b"12345"),                                                 # long match, offset 100
    (seq(bytes(range(256)) * 2, 70, 1000) + seq(b"tail!"),
     bytes(range(256)) * 2 + (bytes(range(256)) * 2)[-70:] * 14 + (bytes(range(256)) * 2)[-70:][:20] + b"tail!"),
    (seq(bytes(np.random.default_rng(1).integers(0, 256, 300, dtype=np.uint8))), None),  # filled below
    (seq(b"ab", 0, 6) + seq(b"12345"), "off0"),                                          # offset 0
    (seq(b"ab", 3, 6) + seq(b"12345"), None),                                            # offset before the block
    (seq(b"abcdefghijklmnopqrst", 1, 8) + seq(b"xyz"), None),                            # match into the last 5 bytes
    (seq(b"hello")[:-1], None),                                                          # truncated literals
    (seq(b"abcdefgh", 2, 40), None),                                                     # block ends after a match
]
KNOWN[5] = (KNOWN[5][0], bytes(np.random.default_rng(1).integers(0, 256, 300, dtype=np.uint8)))  # 300 literals


@pytest.mark.parametrize("i", range(len(KNOWN)))
def test_lz4_known_blocks(i):
    block, want = KNOWN[i]
    n = len(want) if isinstance(want, bytes) else 64
    r, got = oc.lz4_decode(block, n)
    if isinstance(want, bytes):
        assert r == n and got == want
    elif want == "off0":
        assert r == -2
    else:
        assert r == -1
    if _LZ4 is not None and want != "off0":
        lr, lgot = lz4_safe(block, n)
        assert (lr == n) == (r == n)
        if r == n:
            assert lgot == got


def _payloads(rng):
    out = [bytes(rng.integers(0, 256, n, dtype=np.uint8)) for n in (1, 7, 13, 64, 1000, 8192)]
    out += [bytes(np.repeat(rng.integers(0, 4, 400, dtype=np.uint8), rng.integers(1, 40, 400))[:8192])]
    words = rng.integers(0, 2 ** 16, 50, dtype=np.uint32).tobytes()
    out += [b"".join(words[rng.integers(0, 40):][:rng.integers(4, 60)] for _ in range(300))[:16384]]
    # bit-shuffled Poisson counts, as the detector writes them
    v = rng.poisson(3.0, 2048).astype(np.uint32)
    v[rng.integers(0, 2048, 10)] = 2 ** 32 - 1
    bits = np.unpackbits(v.view(np.uint8).reshape(2048, 4), axis=1, bitorder="little")     # (n, 32)
    out += [np.packbits(bits.T, axis=1, bitorder="little").tobytes()]
    return out


@need_lz4
@pytest.mark.parametrize("hc", [0, 9])
def test_lz4_oracle_decodes_liblz4_blocks(hc):
    rng = np.random.default_rng(7 + hc)
    for data in _payloads(rng):
        block = lz4_compress(data, hc)
        r, got = oc.lz4_decode(block, len(data))
        assert r == len(data) and got == data
        # a wrong expected size is an error, as in lz4_stage
        assert oc.lz4_decode(block, len(data) + 1)[0] < 0


@need_lz4
def test_lz4_error_rules_match_liblz4():
    """Corrupted blocks: the oracle accepts exactly what LZ4_decompress_safe decodes to the full size (matches of
    offset 0, which liblz4 does not test for, excluded), with the same bytes."""
    rng = np.random.default_rng(11)
    data = _payloads(rng)
    seen = {"accept": 0, "reject": 0, "off0": 0}
    for k in range(3000):
        d = data[k % len(data)]
        block = bytearray(lz4_compress(d, 9 if k % 2 else 0))
        how = k % 3
        if how == 0:
            for _ in range(rng.integers(1, 4)):
                block[rng.integers(0, len(block))] = rng.integers(0, 256)
        elif how == 1:
            block = block[:rng.integers(0, len(block))]
        else:
            block[rng.integers(0, len(block))] ^= 1 << int(rng.integers(0, 8))
        block = bytes(block)
        n = len(d)
        r, got = oc.lz4_decode(block, n)
        if r == -2:
            seen["off0"] += 1
            continue
        lr, lgot = lz4_safe(block, n) if block else (-1, b"")
        assert (r == n) == (lr == n), (k, r, lr)
        if r == n:
            assert got == lgot
        seen["accept" if r == n else "reject"] += 1
    assert seen["accept"] > 50 and seen["reject"] > 1000


@pytest.mark.parametrize("n,esize", [(8, 4), (2048, 4), (208, 4), (4096, 2), (24, 2), (1024, 4)])
def test_bitunshuffle_matches_numpy(n, esize):
    rng = np.random.default_rng(n + esize)
    buf = rng.integers(0, 256, n * esize, dtype=np.uint8)
    bits = np.unpackbits(buf.reshape(8 * esize, n // 8), axis=1, bitorder="little")      # (planes, n)
    want = np.packbits(bits.T, axis=1, bitorder="little").reshape(-1)
    np.testing.assert_array_equal(oc.bitunshuffle(buf, n, esize), want)


def _frames(rng, ny, nx, dtype):
    fr = rng.poisson(5.0, (ny, nx)).astype(dtype)
    top = np.iinfo(dtype).max
    k = min(20, ny * nx // 4)
    fr.flat[rng.choice(ny * nx, k, replace=False)] = top                        # saturated / module gaps
    fr.flat[rng.choice(ny * nx, k, replace=False)] = min(2 ** 20 + 7, top - 1)   # above a count cut-off
    return fr


@need_lz4
@pytest.mark.parametrize("ny,nx,dtype,block", [(37, 61, np.uint32, 0), (37, 61, np.uint16, 0), (64, 48, np.uint32, 0),
                                              (37, 61, np.uint32, 1000), (1, 5, np.uint32, 0), (1, 2049, np.uint32, 0),
                                              (16, 128, np.uint16, 4096), (3, 7, np.uint16, 16)])
def test_chunk_roundtrip_and_block_table(ny, nx, dtype, block):
    from xpg_amd import bslz4, _lib as L
    rng = np.random.default_rng(ny * nx)
    fr = _frames(rng, ny, nx, dtype)
    chunk = bslz4.encode_frame(fr, block)
    esize, nel = fr.dtype.itemsize, fr.size
    total, bb = bslz4.read_header(chunk)
    assert total == nel * esize and bb == block
    be, raw = oc.decode_chunk(chunk, nel, esize)
    assert be == bslz4.block_elems(block, esize) and be % 8 == 0
    np.testing.assert_array_equal(raw.view(dtype), fr.reshape(-1))
    # the product's host block index: each (offset, size) is an LZ4 block the oracle decodes to its block's bytes
    nb = L.lib().xpg_bslz4_n_blocks(nel, be)
    tab = np.zeros((nb + 1, 2), np.uint32)
    src = np.frombuffer(chunk, np.uint8).copy()
    assert L.lib().xpg_bslz4_index_host(src.ctypes.data_as(C.c_void_p), len(chunk), nel, esize, be,
                                        tab.ctypes.data_as(C.c_void_p)) == 0
    for k in range(nb):
        n = min(be, nel - k * be)
        n8 = n - n % 8
        off, cs = tab[k]
        if n8 == 0:
            assert off == 0 and cs == 0
            continue
        r, blk = oc.lz4_decode(chunk[off:off + cs], n8 * esize)
        assert r == n8 * esize
        np.testing.assert_array_equal(oc.bitunshuffle(np.frombuffer(blk, np.uint8), n8, esize),
                                      fr.reshape(-1)[k * be:k * be + n8].view(np.uint8))
    rest = (nel % be) % 8
    assert tab[nb, 1] == rest * esize
    if rest:
        assert chunk[tab[nb, 0]:tab[nb, 0] + tab[nb, 1]] == fr.reshape(-1)[nel - rest:].tobytes()


@need_lz4
def test_malformed_chunks_rejected_by_oracle_and_index():
    from xpg_amd import bslz4, _lib as L
    rng = np.random.default_rng(3)
    fr = _frames(rng, 37, 61, np.uint32)
    chunk = bslz4.encode_frame(fr)
    nel, be = fr.size, 2048

    def idx(c, nel=nel, be=be):
        tab = np.zeros((L.lib().xpg_bslz4_n_blocks(nel, be) + 1, 2), np.uint32)   # the table of this block size
        src = np.frombuffer(c, np.uint8).copy() if c else np.zeros(1, np.uint8)
        return L.lib().xpg_bslz4_index_host(src.ctypes.data_as(C.c_void_p), len(c), nel, 4, be,
                                            tab.ctypes.data_as(C.c_void_p))
    assert idx(chunk) == 0
    assert idx(chunk[:11]) == 1 and oc.decode_chunk(chunk[:11], nel, 4)[0] == -1             # header
    assert idx(chunk, nel=nel - 1) == 2 and oc.decode_chunk(chunk, nel - 1, 4)[0] == -1      # wrong total
    assert idx(chunk, be=1024) == 4                                                         # other block size
    cut = chunk[:len(chunk) // 2]
    assert idx(cut) == 8 and oc.decode_chunk(cut, nel, 4)[0] == -2                         # truncated
    big = bytearray(chunk)
    big[12:16] = (10 ** 6).to_bytes(4, "big")
    assert idx(bytes(big)) == 16 and oc.decode_chunk(bytes(big), nel, 4)[0] == -2          # beyond the LZ4 bound
    bad = bytearray(chunk)
    bad[16:24] = b"\xff" * 8                                                                # corrupt first block
    assert idx(bytes(bad)) == 0 and oc.decode_chunk(bytes(bad), nel, 4)[0] in (-3, -4)


@pytest.mark.parametrize("esize", [2, 4])
def test_codes_match_imageset_rules(esize):
    """unshuffle_frame_torch's codes against ImageSet.get's (xpg/io/images.py:80-96) on raw counts: raw < min_valid,
    flagged (2^n - 1) or outside the static mask -> -1, raw >= overload -> -2, -1 wins."""
    rng = np.random.default_rng(esize)
    dt = np.uint32 if esize == 4 else np.uint16
    raw = rng.integers(0, 60, 5000).astype(dt)
    raw[rng.integers(0, 5000, 80)] = np.iinfo(dt).max
    raw[rng.integers(0, 5000, 80)] = 50
    static = rng.random(5000) < 0.05                                           # True = untrusted
    for overload, min_valid in ((None, 0.0), (49.5, 0.0), (50.0, 2.5), (40.2, 1.0)):
        got = oc.codes(raw.view(np.uint8), esize, overload, min_valid, static)
        img = raw.astype(np.float32)
        bad = raw < min_valid
        bad |= raw == np.iinfo(dt).max
        bad |= static
        if overload is not None:
            img[raw >= overload] = -2.0
        img[bad] = -1.0
        np.testing.assert_array_equal(got, img)


def test_decode_output_capacity_checked_before_launch():
    # ChunkBatch.decode_into validates the output with output_frames() before any launch: the kernel writes nel
    # elements per frame at slot * stride and has no capacity argument of its own
    from xpg_amd.bslz4 import output_frames
    assert output_frames((4, 37, 61), 37 * 61, 4, False) == (37 * 61, 4)
    assert output_frames((37 * 61,), 37 * 61, 1, False) == (37 * 61, 1)
    assert output_frames((2, 3000), 2257, 9, True) == (3000, 2)            # slots pick the frames
    with pytest.raises(ValueError):
        output_frames((37 * 61 - 1,), 37 * 61, 1, False)                   # 1-D output one element short
    with pytest.raises(ValueError):
        output_frames((0,), 16, 1, False)
    with pytest.raises(ValueError):
        output_frames((4, 100), 101, 4, False)                             # frames smaller than the chunks'
    with pytest.raises(ValueError):
        output_frames((2, 37 * 61), 37 * 61, 3, False)                     # more chunks than frames
