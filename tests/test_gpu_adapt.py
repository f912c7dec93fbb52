"""Device adaptive-width fits (xpg_adapt_block) against the oracle restatement of _adapt_block."""
import ctypes as C

import numpy as np
import pytest
import torch

from adapt_inputs import make_block
from oracle import adapt as oa

pytestmark = pytest.mark.gpu


def _run_device(b, sd, sm, dev, active=0):
    from xpg_amd import _lib as L
    from xpg_amd.integrate_kernels import rc_default_weights
    n = len(b["isum"])
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
    arrs = dict(isum=T(b["isum"], torch.float64), vsum=T(b["vsum"], torch.float64),
                n_over=T(b["n_over"], torch.int32), frac_rec=T(b["frac_rec"], torch.float64),
                m_d=T(b["m_d"], torch.float64), fz=T(b["fz"], torch.float32), phi=T(b["phi_deg"], torch.float64),
                zeta_k=T(b["zeta_k"], torch.float64), iz0=T(b["iz0"], torch.int32), fb=T(b["fb"], torch.int32),
                Xlen=T(b["Xlen"], torch.float64), cosa=T(b["cosa"], torch.float64))
    w = L.Widths()
    w.sd, w.sm = sd, sm
    w.rcw[:] = [float(v) for v in rc_default_weights()]
    w.active = active
    sz = C.sizeof(L.Widths)
    buf = torch.frombuffer(bytearray(bytes(w) * 3), dtype=torch.uint8).to(dev)
    wsb = int(L.lib().xpg_adapt_workspace_bytes(n))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    p = lambda t: L.ptr(t)
    print(f"adapt_block n={n}", flush=True)
    L.check(L.lib().xpg_adapt_block(n, p(arrs["isum"]), p(arrs["vsum"]), p(arrs["n_over"]), p(arrs["frac_rec"]),
                                    p(arrs["m_d"]), p(arrs["fz"]), p(arrs["phi"]), p(arrs["zeta_k"]), p(arrs["iz0"]),
                                    p(arrs["fb"]), p(arrs["Xlen"]), p(arrs["cosa"]), sd, sm, b["phi0_deg"],
                                    b["osc_deg"], b["qx"], C.c_void_p(buf.data_ptr()),
                                    C.c_void_p(buf.data_ptr() + sz), C.c_void_p(buf.data_ptr() + 2 * sz), p(ws), wsb,
                                    L.stream_ptr(dev)))
    raw = buf.cpu().numpy().tobytes()
    return L.Widths.from_buffer_copy(raw[sz:2 * sz]), L.Widths.from_buffer_copy(raw[2 * sz:3 * sz])


@pytest.mark.parametrize("seed,n,sm_true,levels", [
    (3, 900, 0.15, 0), (4, 3000, 0.08, 0), (5, 150, 0.2, 0), (6, 60, 0.1, 0),
    (7, 12000, 0.12, 0),      # 6469 strong rows, more than the median keeps in registers: keys read from memory
    (8, 2500, 0.1, 7),        # seven distinct sigma_D terms: the middle rank among many equal keys (1367 strong)
    (10, 2500, 0.1, 7),       # the same with an even strong count (1374): two middle ranks among equal keys
    (10, 2501, 0.1, 0)])      # an even strong count (1380) with distinct keys
def test_adapt_block_matches_oracle(seed, n, sm_true, levels):
    b = make_block(n=n, seed=seed, sm_true_deg=sm_true, md_levels=levels)
    sd, sm = np.radians(0.02), np.radians(0.1)
    w2, wn = _run_device(b, sd, sm, torch.device("cuda:0"))
    m, sm_new, w, sd_new, sd_fit, sm_fit = oa.adapt_block(
        b["isum"], b["vsum"], b["n_over"], b["frac_rec"], b["m_d"], b["fz"], b["phi_deg"], b["zeta_k"], b["iz0"],
        b["fb"], b["Xlen"], b["cosa"], sd, sm, b["phi0_deg"], b["osc_deg"], b["qx"])
    assert w2.n_strong == m
    if m < 100:
        assert w2.sm == sm and wn.sd == sd and w2.sm_fit == 0
        return
    assert abs(w2.sm_fit - sm_fit) <= 1e-12 * sm_fit
    assert abs(w2.sm - sm_new) <= 1e-12 * sm_new and wn.sm == w2.sm
    assert abs(wn.sd - sd_new) <= 1e-12 * sd_new and abs(w2.sd_fit - sd_fit) <= 1e-12 * sd_fit
    assert w2.sd == sd                                      # this block's pass 2 keeps its own sigma_D
    np.testing.assert_allclose(np.array(w2.rcw[:]), w, atol=2e-6)
    assert w2.active == wn.active and w2.active != 0
    # a fit started from the positive set of the previous one (the next block's warm start) gives the same weights
    w2w, _ = _run_device(b, sd, sm, torch.device("cuda:0"), active=w2.active)
    np.testing.assert_allclose(np.array(w2w.rcw[:]), np.array(w2.rcw[:]), atol=1e-6)
    # and so does a start from an unrelated set
    w2x, _ = _run_device(b, sd, sm, torch.device("cuda:0"), active=0b101010101010101010101)
    np.testing.assert_allclose(np.array(w2x.rcw[:]), np.array(w2.rcw[:]), atol=1e-6)
