"""torch.ops.xint on the device: every operator gives, bit for bit, what the C-ABI entry points give through
make_kernels' closures and the chunk helpers (the same kernels on the same values), on a chunk with partners,
neighbours and mixture weights."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _same(a, b, key):
    a, b = torch.nan_to_num(a.double(), nan=-7.0), torch.nan_to_num(b.double(), nan=-7.0)
    assert torch.equal(a, b), key


def test_operators_equal_the_entry_points():
    from chunks import make_chunk
    from xpg_amd import integrate_kernels as K, ops
    from xpg_amd.frames import DeviceFrames
    xint = ops.load()
    dev = torch.device("cuda:0")
    c = make_chunk(n=96, seed=23)
    assert np.asarray(c["rc"]["has_p"]).any()
    print("xint ops n=96 7x11x11", flush=True)
    a = (c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"], c["rcw"])
    rows = ops.rows(c["rc"], c["nb"], dev)
    det, dsc = ops.detector(c["det"], dev)
    cs, w = ops.call(*a)
    p1, acc, p2 = K.make_kernels(7, 11, 11, 9)
    sb = torch.as_tensor(c["sb"], dtype=torch.float32, device=dev)
    args = (c["rc"], c["det"], c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"])
    r1 = p1(sb, *args, c["prof"], c["nb"], c["rcw"])
    o1 = xint.pass1(sb, rows, det, dsc, cs, w)
    for k, t in zip(ops.P1_OUT, o1):
        _same(t, r1[k], k)
    prof = torch.as_tensor(c["prof"], device=dev)
    r2 = p2(sb, *args, c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"])
    o2 = xint.pass2(sb, rows, det, dsc, cs, w, prof, o1[0], o1[1])
    for k, t in zip(ops.P2_OUT, o2):
        _same(t, r2[k], k)
    # accumulation and normalisation against the closures' accumulate_profiles and the library's normalisation
    sel = (r1["isum"] > 0).to(torch.uint8)
    region = torch.as_tensor(np.asarray(c["rc"]["region"]), dtype=torch.int32, device=dev)
    num, den = xint.accumulate_profiles(o1[18], o1[22], o1[19], o1[20], o1[21], region, o1[2], sel, o1[13], 9)
    n2, d2 = acc(r1["img2"], r1["all_valid2"], r1["J"], r1["u1"], r1["u2"], region, r1["isum"], sel, r1["seff"])
    _same(num, n2, "num")
    _same(den, d2, "den")
    nfx = torch.round(num * K.FX_SCALE).to(torch.int64)
    dfx = torch.round(den * K.FX_SCALE).to(torch.int64)
    pr, n64, d64 = xint.normalise_profiles(nfx, dfx, torch.as_tensor(K.gaussian_profile(), device=dev))
    pr2, _, _ = K.normalise_profiles(nfx, dfx, dev)
    _same(pr, pr2, "prof")
    # extraction from a device frame store
    nfr, ny, nx = 420, 1024, 1024
    raw = np.random.default_rng(2).poisson(3.0, (nfr, ny, nx)).astype(np.uint32)
    frames = DeviceFrames(torch.as_tensor(raw.view(np.int32), device=dev), first=1)
    refl = K.DeviceRefl(c["rc"], c["nb"], (7, 11, 11), dev)
    st = frames.store()
    sb2 = K.run_extract(st, refl, dev)
    sbo = xint.extract(frames.data, st.frame0, st.min_valid, st.flag_value, st.overload, bool(st.use_flag_value), rows,
                       [7, 11, 11])
    _same(sbo, sb2, "sb")
    # errors raise before any launch
    with pytest.raises(RuntimeError, match="rows"):
        xint.pass1(sb, rows[:10], det, dsc, cs, w)
    with pytest.raises(RuntimeError, match="rcw"):
        xint.pass1(sb, rows, det, dsc, cs, w[:5])
