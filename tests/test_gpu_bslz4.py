"""Device bitshuffle-LZ4 decode (xpg_amd/csrc/xpg_decode.hip through include/xpg_decode.h) against the oracle
(oracle/codec.c), bit-exact: raw counts and value-coded float32 frames, the reference test's shapes and codes, blocks
from liblz4's compressors, hand-built LZ4 blocks and their error rules, block tables walked on the device, a full
4148 x 4362 detector frame, and integration from decoded frames equal to integration from the original frames."""
import numpy as np
import pytest
import torch

from oracle import codec as oc

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _frames(rng, n, ny, nx, dtype, lam=5.0):
    fr = rng.poisson(lam, (n, ny, nx)).astype(dtype)
    top = np.iinfo(dtype).max
    for f in fr:
        k = min(20, ny * nx // 4)
        f.flat[rng.choice(ny * nx, k, replace=False)] = top
        f.flat[rng.choice(ny * nx, k, replace=False)] = min(2 ** 20 + 7, top - 1)
    return fr


def _decode_all(chunks, nel, esize, kind, index="host", **codes):
    from xpg_amd import bslz4
    cb = bslz4.ChunkBatch(chunks, nel, esize, DEV, index=index)
    out = torch.full((len(chunks), nel), -7, dtype=torch.int32 if kind == "raw" else torch.float32, device=DEV)
    cb.decode_into(out, kind, **codes)
    return cb, out


@pytest.mark.parametrize("ny,nx,dtype,block", [(37, 61, np.uint32, 0), (37, 61, np.uint16, 0), (64, 48, np.uint32, 0),
                                              (37, 61, np.uint32, 1000), (1, 5, np.uint32, 0),
                                              (16, 128, np.uint16, 4096), (3, 7, np.uint16, 16),
                                              (200, 173, np.uint32, 32768)])
def test_decode_matches_oracle(ny, nx, dtype, block):
    from xpg_amd import bslz4
    print(f"bslz4 decode {ny}x{nx} {np.dtype(dtype).name} block {block}", flush=True)
    rng = np.random.default_rng(ny * nx + block)
    fr = _frames(rng, 3, ny, nx, dtype)
    chunks = bslz4.encode_frames(list(fr), block, threads=3)
    esize, nel = fr.itemsize, ny * nx
    static = rng.random(nel) < 0.05
    for index in ("host", "device"):
        cb, raw = _decode_all(chunks, nel, esize, "raw", index=index)
        cb.check()
        want = np.stack([oc.decode_chunk(c, nel, esize)[1].view(dtype) for c in chunks]).astype(np.uint32)
        np.testing.assert_array_equal(raw.cpu().numpy().view(np.uint32), want)
        for overload, min_valid, sb in ((None, 0.0, None), (2 ** 20 + 7 - 0.5, 1.0, static), (50.0, 0.0, static)):
            cb, img = _decode_all(chunks, nel, esize, "coded", index=index, overload=overload, min_valid=min_valid,
                                  static_bad=sb)
            cb.check()
            for i, c in enumerate(chunks):
                ref = oc.codes(oc.decode_chunk(c, nel, esize)[1], esize, overload, min_valid, sb)
                np.testing.assert_array_equal(img[i].cpu().numpy(), ref)
    # the device walk writes the host walk's table
    hb, db = bslz4.ChunkBatch(chunks, nel, esize, DEV, "host"), bslz4.ChunkBatch(chunks, nel, esize, DEV, "device")
    assert torch.equal(hb.table, db.table) and torch.equal(hb.status, db.status)


def test_decode_frame_matches_reference_codes():
    """decode_frame (lz4_stage + unshuffle_frame_torch in one call) on the reference test's frame
    (tests/test_robustness.py:73-104 of the reference: 37 x 61 uint32, 20 saturated, 20 above the cut-off, 5% masked)
    against ImageSet.get's rules on the raw counts."""
    from xpg_amd import bslz4
    rng = np.random.default_rng(3)
    ny, nx = 37, 61
    mask = rng.random((ny, nx)) < 0.05
    for k in range(3):
        fr = rng.poisson(5.0 * (k + 1), (ny, nx)).astype(np.uint32)
        fr.flat[rng.choice(ny * nx, 20, replace=False)] = 2 ** 32 - 1
        fr.flat[rng.choice(ny * nx, 20, replace=False)] = 2 ** 20 + 7
        img = bslz4.decode_frame(bslz4.encode_frame(fr), ny * nx, 4, overload=2 ** 20, min_valid=0.0,
                                 static_bad=mask).cpu().numpy().reshape(ny, nx)
        want = fr.astype(np.float32)
        want[fr >= 2 ** 20] = -2.0
        want[(fr == 2 ** 32 - 1) | mask] = -1.0
        np.testing.assert_array_equal(img, want)
    a = rng.integers(0, 2 ** 16, (ny, nx), dtype=np.uint16)
    img = bslz4.decode_frame(bslz4.encode_frame(a), ny * nx, 2).cpu().numpy().reshape(ny, nx)
    want = a.astype(np.float32)
    want[a == 2 ** 16 - 1] = -1.0
    np.testing.assert_array_equal(img, want)


def _chunk_of_block(block: bytes, nel: int, esize: int) -> bytes:
    """A one-block chunk around a raw LZ4 block of nel (a multiple of 8) elements."""
    return ((nel * esize).to_bytes(8, "big") + (nel * esize).to_bytes(4, "big") + len(block).to_bytes(4, "big") + block)


def test_hand_built_and_liblz4_blocks_match_oracle():
    """Single-block chunks around LZ4 blocks: hand-built blocks (literal-only, offset-1 run, periodic overlap, long
    match, extended lengths) and each error rule, liblz4's fast and high-compression output on run-heavy, periodic,
    random and bit-shuffled data, and corrupted copies -- the device decodes exactly what the oracle decodes and flags
    exactly the blocks the oracle rejects."""
    from test_bslz4_oracle import _LZ4, _payloads, lz4_compress, seq
    from xpg_amd import bslz4
    rng = np.random.default_rng(5)
    r300 = bytes(rng.integers(0, 256, 320, dtype=np.uint8))
    cases = [(seq(bytes(range(64))), 64), (seq(b"a", 1, 53) + seq(b"bcdefghijk"), 64),
             (seq(b"abc", 3, 45) + seq(b"0123456789abcdef"), 64),
             (seq(bytes(range(100)), 100, 300) + seq(b"0123456789abcdef"), 416), (seq(r300), 320),
             (seq(bytes(range(256)), 70, 1000) + seq(b"0123456789abcdefghijklmn"), 1280),
             (seq(b"ab", 0, 46) + seq(b"0123456789abcdef"), 64), (seq(b"ab", 3, 46) + seq(b"0123456789abcdef"), 64),
             (seq(bytes(range(40)), 1, 22) + seq(b"xy"), 64), (seq(bytes(range(64)))[:-1], 64),
             (seq(bytes(range(16)), 2, 48), 64)]
    if _LZ4 is not None:
        for hc in (0, 9):
            for data in _payloads(rng):
                if len(data) % 32 == 0:
                    blk = lz4_compress(data, hc)
                    cases.append((blk, len(data)))
                    for _ in range(3):
                        bad = bytearray(blk)
                        bad[rng.integers(0, len(bad))] = rng.integers(0, 256)
                        cases.append((bytes(bad), len(data)))
    print(f"{len(cases)} single-block chunks", flush=True)
    n_ok = n_bad = 0
    for blk, n_out in cases:
        for esize in (4, 2):
            if n_out % (8 * esize):
                continue
            nel = n_out // esize
            chunk = _chunk_of_block(blk, nel, esize)
            be, raw = oc.decode_chunk(chunk, nel, esize)
            cb = bslz4.ChunkBatch([chunk], nel, esize, DEV)
            out = torch.zeros((1, nel), dtype=torch.int32, device=DEV)
            cb.decode_into(out, "raw")
            st = int(cb.status.cpu()[0])
            if be > 0:
                assert st == 0, (n_out, esize)
                want = raw.view(np.uint16 if esize == 2 else np.uint32).astype(np.uint32)
                np.testing.assert_array_equal(out[0].cpu().numpy().view(np.uint32), want)
                n_ok += 1
            else:
                assert st != 0, (be, n_out, esize)
                assert (be in (-3, -4)) == bool(st & 32), (be, st)
                n_bad += 1
    assert n_ok > 20 and n_bad > 10


def test_full_size_frames_and_slots():
    """Two 4148 x 4362 frames of the synthetic sweep (background, spots, flagged and saturated pixels): decoded into
    a store in swapped slots, bit-identical to the originals; a truncated and a corrupted chunk raise ValueError and
    leave the other frames decoded."""
    from xpg_amd import bslz4, synth
    spec = synth.replace(synth.spec_full(), nframes=4, block_frames=2) if hasattr(synth, "spec_full") else None
    if spec is None:
        pytest.skip("no full-size spec")
    frames, _ = synth.generate(spec, DEV)
    src = frames.data[:2].contiguous()
    host = src.cpu().numpy().view(np.uint32)
    print("full-size frames", host.shape, flush=True)
    chunks = bslz4.encode_frames(list(host), 0, threads=2)
    nel = host[0].size
    cb = bslz4.ChunkBatch(chunks, nel, 4, DEV)
    out = torch.zeros((2, spec.ny, spec.nx), dtype=torch.int32, device=DEV)
    cb.decode_into(out, "raw", slots=[1, 0])
    cb.check()
    assert torch.equal(out[1], src[0]) and torch.equal(out[0], src[1])
    cr = bytearray(chunks[1])
    cs0 = int.from_bytes(cr[12:16], "big")           # first block one byte short: its LZ4 data or the walk fails
    cr[12:16] = (cs0 - 1).to_bytes(4, "big")
    cb = bslz4.ChunkBatch([chunks[0], chunks[0][:len(chunks[0]) // 3], bytes(cr)], nel, 4, DEV)
    out = torch.zeros((3, spec.ny, spec.nx), dtype=torch.int32, device=DEV)
    cb.decode_into(out, "raw")
    with pytest.raises(ValueError):
        cb.check()
    st = cb.status.cpu().numpy()
    assert st[0] == 0 and st[1] & 8 and st[2] != 0
    assert torch.equal(out[0], src[0])


def test_integrate_from_decoded_frames():
    """Frames written as bitshuffle-LZ4 chunks and decoded on the device into the frame store integrate to exactly the
    results of the original frames."""
    from xpg_amd import bslz4, integrate as I, synth
    spec = synth.spec_small()
    frames, spots = synth.generate(spec, DEV)
    exp = spots["exp"]
    edges = synth.block_edges(spec)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    blocks = [exp] * (len(edges) - 1)
    host = frames.data.cpu().numpy().view(np.uint32)
    chunks = bslz4.encode_frames(list(host), 0)
    dec = bslz4.frames_from_chunks(chunks, spec.ny, spec.nx, 4, DEV, first=frames.first, overload=frames.overload,
                                   min_valid=frames.min_valid, static_mask=None)
    dec.flag_value = frames.flag_value
    assert torch.equal(dec.data, frames.data)
    r0 = I.integrate(frames, blocks, edges, model, spots["dmin"], verbose=False).refl
    r1 = I.integrate(dec, blocks, edges, model, spots["dmin"], verbose=False).refl
    for k in ("iprf", "vprf", "isum", "bg"):
        assert np.array_equal(r0[k], r1[k], equal_nan=True), k
