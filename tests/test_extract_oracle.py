"""Pins of the extraction oracle (oracle/extract.py): the reference's golden vector for the bin decoding and a
brute-force definition of every shoebox voxel on small frames."""
import numpy as np

from oracle import extract as ox


def test_decode_bins_golden_vector():
    """tests/test_robustness.py:44-54 of the reference: overload (NaN) wins, masked (+inf) -> -1, empty bins
    unrecorded (0), values stored + SB_OFF."""
    arr = np.array([[[[5.0, np.inf], [np.nan, 0.0]], [[2.0, 3.0], [4.0, 1.0]]]], np.float32)
    cnt = np.array([[2, 0]], np.int16)
    out = ox.decode_bins(arr, cnt)
    np.testing.assert_allclose(out[0, 0], np.array([[5.0 + 3.0, 3.0 - 1.0], [3.0 - 2.0, 3.0]], np.float32))
    assert np.all(out[0, 1] == 0.0)
    assert np.isnan(np.float32(np.nan) + np.float32(np.inf)) and np.isinf(np.float32(np.inf) + np.float32(3.0))


def test_frame_codes():
    raw = np.array([[0, 5, 2 ** 32 - 1, 70000]], np.uint32)
    sm = np.array([[True, True, True, False]])
    c = ox.frame_codes(raw, min_valid=1, flag_value=2 ** 32 - 1, overload=65000, static_mask=sm)
    assert c.tolist() == [[-1.0, 5.0, -1.0, -1.0]]          # masked overload: -1 (written last)
    c = ox.frame_codes(raw, overload=65000)
    assert c.tolist() == [[0.0, 5.0, -2.0, -2.0]]


def _voxel(frames, read, mask, fi_list, sy, sx, s):
    """Brute force: value of (super-)pixel (sy, sx) summed over the frames of one bin."""
    vals = []
    for fi in fi_list:
        if fi not in read:
            return -3.0
        img = frames[fi]
        ny, nx = img.shape
        if s == 1:
            if not (0 <= sy < ny and 0 <= sx < nx):
                v = -1.0
            else:
                v = float(img[sy, sx])
                if mask is not None and not mask[sy, sx]:
                    v = min(v, -1.0)
        else:
            if not (0 <= sy < ny // s and 0 <= sx < nx // s):
                v = -1.0
            else:
                px = []
                for a in range(s):
                    for b in range(s):
                        w = float(img[sy * s + a, sx * s + b])
                        if mask is not None and not mask[sy * s + a, sx * s + b]:
                            w = min(w, -1.0)
                        px.append(w)
                mn = min(px)
                v = (-2.0 if mn <= -1.5 else -1.0) if mn < 0 else float(np.float32(sum(px)))
        vals.append(v)
    if len(vals) == 1:
        return vals[0]
    if any(v == -2.0 for v in vals):
        return -2.0
    if any(v < 0 for v in vals):
        return -1.0
    return float(np.float32(sum(vals)))


def _frames(rng, n, ny, nx):
    frames = {}
    for fi in range(n):
        raw = rng.poisson(3.0, (ny, nx)).astype(np.uint32)
        raw[rng.random((ny, nx)) < 0.03] = 2 ** 32 - 1
        raw[rng.random((ny, nx)) < 0.02] = 70000
        frames[fi] = ox.frame_codes(raw, flag_value=2 ** 32 - 1, overload=65000)
    return frames


def test_gather_matches_brute_force():
    rng = np.random.default_rng(4)
    ny, nx, nfr = 23, 31, 14
    frames = _frames(rng, nfr, ny, nx)
    mask = rng.random((ny, nx)) > 0.05
    read = [f for f in range(nfr) if f not in (6,)]          # frame 6 excluded
    for (fbv, s, B, bx) in ((1, 1, 5, 7), (2, 1, 3, 5), (1, 2, 3, 5), (3, 2, 3, 5)):
        n = 12
        iz0 = rng.integers(-3, nfr - 2, n)
        # origins the reference can produce (box centre on the detector): floor(x - 0.5) // s - half
        h = bx // 2
        ix0 = rng.integers(0, (nx - 1) // s + 1, n) - h
        iy0 = rng.integers(0, (ny - 1) // s + 1, n) - h
        fb = np.full(n, fbv)
        sb = ox.gather(frames, read, ix0, iy0, iz0, fb, s, B, bx, mask)
        for r in range(n):
            for j in range(B):
                for y in range(bx):
                    for x in range(bx):
                        fl = [iz0[r] + fbv * j + q for q in range(fbv)]
                        want = _voxel(frames, set(read), mask, fl, iy0[r] + y, ix0[r] + x, s)
                        assert sb[r, j, y, x] == np.float32(want), (fbv, s, r, j, y, x, sb[r, j, y, x], want)
