"""The allocation-free CBF decode (xpg_cbf_decode_static, include/xpg_cbf.h) against the workspace entry point
(xpg_cbf_decode) through the C ABI, bit for bit: two calls in a row on the same GPU, the second with a smaller batch
(the scratch keeps the first call's contents), and the source values themselves.  Oversized batches are refused
before any launch (tests/test_lib_abi.py)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _frames(rng, shapes):
    from xpg_amd import cbf
    vals, streams = [], []
    for ny, nx in shapes:
        a = rng.poisson(2.0, (ny, nx)).astype(np.int32)
        hot = rng.random((ny, nx)) < 1e-3
        a[hot] = rng.integers(200, 1_048_500, int(hot.sum()))      # int16 and int32 escapes
        a[rng.random((ny, nx)) < 1e-4] = -1
        vals.append(a.reshape(-1))
        streams.append(cbf.encode_byte_offset(a) + b"\x00" * 13)    # trailing padding, as a binary section has
    return vals, streams


def _decode_both(streams, ns):
    from xpg_amd import cbf
    sb = cbf.StreamBatch(streams, ns, DEV)
    n = int(np.sum(ns))
    ref = torch.full((n,), -77, dtype=torch.int32, device=DEV)
    got = torch.full((n,), -77, dtype=torch.int32, device=DEV)
    sb.decode_into(ref)
    sb.decode_into(got, static_scratch=True)
    torch.cuda.synchronize()
    sb.check()
    return sb, ref.cpu().numpy(), got.cpu().numpy()


def test_static_scratch_decode_matches_the_workspace_decode():
    rng = np.random.default_rng(41)
    for call, shapes in enumerate([[(2527, 2463)] * 4 + [(1043, 981)], [(400, 300), (77, 61)]]):
        print("static-scratch CBF decode, call", call + 1, "frames", shapes, flush=True)
        vals, streams = _frames(rng, shapes)
        ns = [v.size for v in vals]
        sb, ref, got = _decode_both(streams, ns)
        if call == 0:
            assert sb.n_chunks > 40_000          # the first call fills a good part of the scratch
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, np.concatenate(vals))
