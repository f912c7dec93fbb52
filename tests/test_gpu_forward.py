"""The forward model's HIP scatter / gather (include/xpg_forward.h, xpg_amd/forward.py) against the oracle
(oracle/forward.py, xpg/forward/kernels.py:34-134): float32 atomics and sums, so within a rounding bound taken
against the sum of |terms| each output adds (2e-5 of it: a few hundred float32 ulps)."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import forward as FO

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
RTOL = 2e-5


def _case(seed, T, n, Dz, b, win):
    rng = np.random.default_rng(seed)
    nv = (win.f1 - win.f0) * win.ny * win.nx
    R = rng.uniform(0.05, 1.0, (T, n, Dz)).astype(np.float32)
    P = rng.uniform(0.0, 1.0, (T, n, b, b)).astype(np.float32)
    iz0 = rng.integers(win.f0 - Dz + 1, win.f1, n)
    iy0 = rng.integers(win.y0 - b + 1, win.y0 + win.ny, n)
    ix0 = rng.integers(win.x0 - b + 1, win.x0 + win.nx, n)
    return SimpleNamespace(R=R, P=P, iz0=iz0, iy0=iy0, ix0=ix0, nv=nv, rng=rng,
                           I=rng.uniform(1, 500, n).astype(np.float32), beta=rng.uniform(0, 3, n).astype(np.float32))


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV, dtype=dtype)


WIN_SMALL = SimpleNamespace(f0=3, f1=11, y0=40, ny=70, x0=25, nx=90)
WIN_LARGE = SimpleNamespace(f0=100, f1=130, y0=512, ny=512, x0=1024, nx=768)


@pytest.mark.parametrize("T,n,Dz,b,win,beta", [(1, 300, 5, 7, WIN_SMALL, True), (2, 300, 5, 7, WIN_SMALL, False),
                                                 (2, 20000, 9, 13, WIN_LARGE, True), (1, 500, 3, 9, WIN_SMALL, True)])
def test_scatter_matches_oracle(T, n, Dz, b, win, beta):
    from xpg_amd import forward as F
    c = _case(5 + n + T, T, n, Dz, b, win)
    mu0 = c.rng.uniform(0, 2, c.nv).astype(np.float32)
    want = FO.scatter(mu0, c.R, c.P, c.I, c.beta if beta else None, c.iz0, c.iy0, c.ix0, win)
    mu = _t(mu0)
    print(f"scatter T={T} n={n} Dz={Dz} b={b} nv={c.nv}", flush=True)
    F.scatter(mu, _t(c.R), _t(c.P), _t(c.I), _t(c.beta) if beta else None, _t(c.iz0), _t(c.iy0), _t(c.ix0), win)
    got = mu.cpu().numpy().astype(np.float64)
    # every term is positive: the oracle's sum is the scale of the rounding bound
    np.testing.assert_allclose(got, want, rtol=RTOL, atol=0)
    assert np.count_nonzero(got != mu0) > c.nv // 20


@pytest.mark.parametrize("ns_dtype", [torch.int16, torch.int32, torch.float32])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("T,n,Dz,b,win", [(2, 400, 5, 7, WIN_SMALL), (2, 20000, 9, 13, WIN_LARGE),
                                          (1, 300, 4, 11, WIN_SMALL)])
def test_gather_matches_oracle(ns_dtype, mode, T, n, Dz, b, win):
    from xpg_amd import forward as F
    c = _case(17 + n + mode, T, n, Dz, b, win)
    rng = c.rng
    A = rng.normal(0, 5, c.nv).astype(np.float32)
    M = rng.uniform(0.5, 40, c.nv).astype(np.float32)
    NS = rng.integers(-1, 60, c.nv)
    NSn = NS.astype({torch.int16: np.int16, torch.int32: np.int32, torch.float32: np.float32}[ns_dtype])
    if ns_dtype == torch.float32:
        NSn[::11] = np.nan
    n_out = n + 7
    rows = rng.permutation(n_out)[:n]
    rows[: n // 10] = rows[n // 10: 2 * (n // 10)]           # some rows receive two reflections
    g0 = rng.normal(0, 1, n_out).astype(np.float32)
    fd0 = rng.uniform(0, 1, n_out).astype(np.float32)
    want_g, want_fd = FO.gather(g0, fd0 if mode == 1 else None, c.R, c.P, rows, c.iz0, c.iy0, c.ix0, win,
                                A if mode == 0 else None, M, NSn, mode)
    scale = FO.term_scale(c.R, c.P, rows, c.iz0, c.iy0, c.ix0, win, A, M, NSn, mode, n_out)
    g, fd = _t(g0), _t(fd0)
    print(f"gather mode={mode} ns={ns_dtype} T={T} n={n} Dz={Dz} b={b}", flush=True)
    F.gather(g, fd if mode == 1 else None, _t(c.R), _t(c.P), _t(rows), _t(c.iz0), _t(c.iy0), _t(c.ix0), win,
             _t(A) if mode == 0 else None, _t(M), _t(NSn), mode)
    got = g.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(got - want_g) <= RTOL * (scale + np.abs(g0)) + 1e-6), np.max(np.abs(got - want_g) / (scale + 1))
    assert np.count_nonzero(scale) > n // 2
    if mode == 1:
        scale_d = FO.gather(np.zeros(n_out), np.zeros(n_out), c.R, c.P, rows, c.iz0, c.iy0, c.ix0, win, None, M, NSn,
                            1)[1]
        got_d = fd.cpu().numpy().astype(np.float64)
        assert np.all(np.abs(got_d - want_fd) <= RTOL * (scale_d + np.abs(fd0)) + 1e-6)
    else:
        assert torch.equal(fd, _t(fd0))


def test_gather_is_deterministic_for_distinct_rows():
    """One wavefront sums a reflection in a fixed order and adds once: with distinct rows two runs agree bit for bit."""
    from xpg_amd import forward as F
    win = WIN_LARGE
    c = _case(99, 2, 5000, 9, 13, win)
    M = _t(c.rng.uniform(0.5, 40, c.nv).astype(np.float32))
    NS = _t(c.rng.integers(-1, 60, c.nv).astype(np.int16))
    args = (_t(c.R), _t(c.P), _t(np.arange(5000)), _t(c.iz0), _t(c.iy0), _t(c.ix0), win)
    out = []
    for _ in range(2):
        g = torch.zeros(5000, device=DEV)
        fd = torch.zeros(5000, device=DEV)
        F.gather(g, fd, *args, None, M, NS, 1)
        out.append((g, fd))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_empty_and_bad_arguments():
    from xpg_amd import _lib as L
    from xpg_amd import forward as F
    win = WIN_SMALL
    mu = torch.zeros((win.f1 - win.f0) * win.ny * win.nx, device=DEV)
    e = torch.zeros((1, 0, 5), device=DEV)
    F.scatter(mu, e, torch.zeros((1, 0, 7, 7), device=DEV), torch.zeros(0, device=DEV), None,
              *(torch.zeros(0, dtype=torch.int32, device=DEV),) * 3, win)
    assert int(torch.count_nonzero(mu)) == 0
    c = _case(1, 1, 10, 5, 7, win)
    args = (_t(c.R), _t(c.P), _t(np.arange(10)), _t(c.iz0), _t(c.iy0), _t(c.ix0), win)
    g = torch.zeros(10, device=DEV)
    with pytest.raises(L.XpgError):
        F.gather(g, None, *args, None, mu, mu, 2)
    with pytest.raises(L.XpgError):
        F.gather(g, g.clone(), *args, None, mu, mu.double(), 1)
    with pytest.raises(L.XpgError):
        F.gather(g, g[:5].clone(), *args, None, mu, mu, 1)
    with pytest.raises(L.XpgError):
        F.scatter(mu[:10], _t(c.R), _t(c.P), _t(c.I), None, _t(c.iz0), _t(c.iy0), _t(c.ix0), win)
    assert F.available(DEV) and not F.available(torch.device("cpu"))


def test_rows_outside_g():
    """Device rows outside g add nothing (the kernel bounds them, no host sync); host rows outside g raise."""
    from xpg_amd import _lib as L
    from xpg_amd import forward as F
    win = WIN_SMALL
    c = _case(2, 2, 40, 5, 7, win)
    M = _t(c.rng.uniform(0.5, 40, c.nv).astype(np.float32))
    NS = _t(c.rng.integers(0, 60, c.nv).astype(np.int32))
    sup = (_t(c.R), _t(c.P))
    org = (_t(c.iz0), _t(c.iy0), _t(c.ix0))
    ref = torch.zeros(40, device=DEV)
    F.gather(ref, torch.zeros(40, device=DEV), *sup, _t(np.arange(40)), *org, win, None, M, NS, 1)
    rows = np.arange(40)
    rows[3], rows[4] = 40, -1
    g = torch.zeros(40, device=DEV)
    F.gather(g, torch.zeros(40, device=DEV), *sup, _t(rows), *org, win, None, M, NS, 1)
    keep = torch.ones(40, dtype=torch.bool, device=DEV)
    keep[3:5] = False
    assert torch.equal(g[keep], ref[keep]) and float(g[3]) == 0.0 and float(g[4]) == 0.0
    assert int(torch.count_nonzero(ref)) > 10
    with pytest.raises(L.XpgError):
        F.gather(g, torch.zeros(40, device=DEV), *sup, rows, *org, win, None, M, NS, 1)
