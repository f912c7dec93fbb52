"""Device CBF byte-offset decode (xpg_amd/csrc/xpg_cbf.hip through include/xpg_cbf.h) against the oracle
(oracle/codec.c: cbf_decode), bit-exact: the reference test's vector, hand-built streams of every escape class,
escape-dense streams whose elements straddle the 256-byte chunks (every entry offset 0 .. 14 taken), int64
escapes, a PILATUS 6M frame, batches of frames of different sizes with trailing padding, and short streams."""
import numpy as np
import pytest
import torch

from oracle import codec as oc

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _dense_stream(rng, n, p=(0.55, 0.2, 0.15, 0.1)):
    """n elements drawn from the four classes (int8, int16, int32, int64 escapes); returns bytes and the values."""
    cls = rng.choice(4, n, p=p)
    parts, d = [], np.zeros(n, np.int64)
    for i, c in enumerate(cls):
        if c == 0:
            v = int(rng.integers(-127, 128))
            parts.append(int(v & 0xFF).to_bytes(1, "little"))
        elif c == 1:
            v = int(rng.integers(-32767, 32768))
            if v == -32768:
                v = 0
            parts.append(b"\x80" + v.to_bytes(2, "little", signed=True))
        elif c == 2:
            v = int(rng.integers(-2 ** 31 + 1, 2 ** 31))
            parts.append(b"\x80\x00\x80" + v.to_bytes(4, "little", signed=True))
        else:
            v = int(rng.integers(-2 ** 62, 2 ** 62))
            parts.append(b"\x80\x00\x80\x00\x00\x00\x80" + v.to_bytes(8, "little", signed=True))
        d[i] = v
    vals = np.cumsum(d.astype(np.int64)).astype(np.int32)   # wraps like read_cbf's astype
    return b"".join(parts), vals


def _decode(streams, ns, out_off=None, total=None):
    from xpg_amd import cbf
    sb = cbf.StreamBatch(streams, ns, DEV)
    n_out = total if total is not None else int(np.sum(ns))
    out = torch.full((max(n_out, 1),), -77, dtype=torch.int32, device=DEV)
    sb.decode_into(out, out_off)
    return sb, out


def test_reference_vector_and_hand_built():
    from test_cbf_oracle import HAND
    from xpg_amd import cbf
    rng = np.random.default_rng(0)
    a = rng.poisson(3, (37, 53)).astype(np.int32)
    a[3, 4], a[5, 6], a[7, 8], a[9, 10] = 1_000_000, -1, 40_000, -2_000_000
    got = cbf.decode_byte_offset(cbf.encode_byte_offset(a), a.size, DEV).cpu().numpy()
    np.testing.assert_array_equal(got.reshape(a.shape), a)
    streams = [buf + b"\x00" * 4095 for buf, _ in HAND]
    ns = [len(w) for _, w in HAND]
    sb, out = _decode(streams, ns)
    sb.check()
    want = np.concatenate([np.array(w, np.int64).astype(np.int32) for _, w in HAND])
    np.testing.assert_array_equal(out.cpu().numpy()[:want.size], want)


@pytest.mark.parametrize("n,seed", [(1, 1), (17, 2), (500, 3), (20000, 4), (150000, 5)])
def test_escape_dense_streams(n, seed):
    print(f"cbf dense n={n}", flush=True)
    rng = np.random.default_rng(seed)
    buf, vals = _dense_stream(rng, n)
    st, ref = oc.cbf_decode(buf, n)
    assert st == 0 and np.array_equal(ref, vals)
    sb, out = _decode([buf], [n])
    sb.check()
    np.testing.assert_array_equal(out.cpu().numpy()[:n], vals)


def test_batch_of_frames_with_offsets_and_short_streams():
    from xpg_amd import cbf
    rng = np.random.default_rng(9)
    streams, ns, wants = [], [], []
    for k in range(12):
        n = int(rng.integers(1, 30000))
        if k % 3 == 0:
            buf, vals = _dense_stream(rng, n, p=(0.9, 0.05, 0.04, 0.01))
        else:
            vals = rng.poisson(4.0, n).astype(np.int32)
            vals[rng.integers(0, n, max(1, n // 100))] = -1
            vals[rng.integers(0, n, max(1, n // 500))] = 1_000_000
            buf = cbf.encode_byte_offset(vals)
        streams.append(buf + b"\x00" * int(rng.integers(0, 600)))
        ns.append(n)
        wants.append(vals)
    # short streams: one element more than held, and one cut inside its last element
    streams.append(streams[1][:len(cbf.encode_byte_offset(wants[1]))])
    ns.append(ns[1] + 1)
    wants.append(None)
    cut = cbf.encode_byte_offset(np.array([1, 2, 100000], np.int32))[:-2]
    streams.append(cut)
    ns.append(3)
    wants.append(None)
    ns = np.array(ns, np.int64)
    gap = 37
    out_off = np.concatenate([[5], 5 + np.cumsum(ns[:-1] + gap)]).astype(np.int64)
    sb, out = _decode(streams, ns, out_off, int(out_off[-1] + ns[-1] + 3))
    st = sb.status.cpu().numpy()
    assert list(st[:12]) == [0] * 12 and st[12] == 1 and st[13] == 1
    with pytest.raises(ValueError):
        sb.check()
    o = out.cpu().numpy()
    for i in range(12):
        np.testing.assert_array_equal(o[out_off[i]:out_off[i] + ns[i]], wants[i])
        assert o[out_off[i] + ns[i]] == -77          # nothing written past a frame
    assert (o[:5] == -77).all()


def test_pilatus_6m_frame():
    from xpg_amd import cbf
    rng = np.random.default_rng(11)
    ny, nx = 2527, 2463
    a = rng.poisson(rng.uniform(0.5, 6.0), (ny, nx)).astype(np.int32)
    a[:, 487::494] = -1                                # module gaps
    a[195::212, :] = -1
    a.flat[rng.choice(a.size, 3000, replace=False)] = rng.integers(1000, 1_048_500, 3000)
    a.flat[rng.choice(a.size, 50, replace=False)] = -2
    buf = cbf.encode_byte_offset(a)
    print(f"pilatus 6M: {len(buf) / a.size:.3f} bytes per pixel", flush=True)
    got = cbf.decode_byte_offset(buf + b"\x00" * 4095, a.size, DEV).cpu().numpy()
    np.testing.assert_array_equal(got.reshape(ny, nx), a)
