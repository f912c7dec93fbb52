"""Device shoebox extraction (xpg_extract and the extraction fused into pass 1 / pass 2) against the extraction oracle:
bit-exact value codes, with dead / overloaded pixels, the integrator's mask, an excluded frame, phi-bins and
super-pixels."""
import numpy as np
import pytest
import torch

from oracle import extract as ox

pytestmark = pytest.mark.gpu


def _setup(seed=4, ny=23, nx=31, nfr=14):
    rng = np.random.default_rng(seed)
    raw = rng.poisson(3.0, (nfr, ny, nx)).astype(np.uint32)
    raw[rng.random(raw.shape) < 0.03] = 2 ** 32 - 1
    raw[rng.random(raw.shape) < 0.02] = 70000
    mask = rng.random((ny, nx)) > 0.05
    coded = {fi: ox.frame_codes(raw[fi], flag_value=2 ** 32 - 1, overload=65000) for fi in range(nfr)}
    return rng, raw, mask, coded


@pytest.mark.parametrize("fbv,s,B,bx", [(1, 1, 5, 7), (2, 1, 3, 5), (1, 2, 3, 5), (3, 2, 3, 5), (1, 1, 9, 9)])
def test_extract_matches_oracle(fbv, s, B, bx):
    from xpg_amd.frames import DeviceFrames
    from xpg_amd.integrate_kernels import DeviceRefl, run_extract
    rng, raw, mask, coded = _setup()
    nfr, ny, nx = raw.shape
    dev = torch.device("cuda:0")
    frames = DeviceFrames(torch.as_tensor(raw.view(np.int32), device=dev), first=1, overload=65000,
                          flag_value=2 ** 32 - 1)
    store = frames.store(mask=mask, exclude_frames=(7,))          # 1-based frame 7 = index 6
    n, h = 40, bx // 2
    iz0 = rng.integers(-3, nfr - 2, n)
    ix0 = rng.integers(0, (nx - 1) // s + 1, n) - h
    iy0 = rng.integers(0, (ny - 1) // s + 1, n) - h
    fb = np.full(n, fbv)
    rc = dict(S=np.zeros((n, 3)), phi=np.zeros(n), ix0=ix0, iy0=iy0, iz0=iz0, fb=fb, pxs=np.full(n, s),
              region=np.zeros(n, int))
    refl = DeviceRefl(rc, None, (B, bx, bx), dev)
    print(f"extract fb={fbv} s={s} B={B} bx={bx} n={n}", flush=True)
    sb = run_extract(store, refl, dev).cpu().numpy()
    read = [f for f in range(nfr) if f != 6]
    want = ox.gather(coded, read, ix0, iy0, iz0, fb, s, B, bx, mask)
    assert np.array_equal(sb, want)


def test_fused_extraction_gives_identical_results():
    """pass 1 and pass 2 reading the frame store directly == extraction into sb followed by the sb kernels."""
    import sys
    from chunks import make_chunk
    from xpg_amd import integrate_kernels as K
    from xpg_amd.frames import DeviceFrames
    dev = torch.device("cuda:0")
    c = make_chunk(n=64, seed=21)
    # frames that contain the chunk's boxes: place each box's voxels into a frame store
    nfr, ny, nx = 420, 1024, 1024
    store_np = np.random.default_rng(1).poisson(2.0, (nfr, ny, nx)).astype(np.uint32)
    frames = DeviceFrames(torch.as_tensor(store_np.view(np.int32), device=dev), first=1)
    store = frames.store()
    det = K.device_detector(c["det"], dev)
    call = K.make_call(c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"], c["rcw"])
    refl = K.DeviceRefl(c["rc"], c["nb"], (7, 11, 11), dev)
    sb = K.run_extract(store, refl, dev)
    a = K.run_pass1(det, call, refl, sb=sb)
    b = K.run_pass1(det, call, refl, frames=store)
    for k in K.P1_F64 + K.P1_I32 + ("fz", "img2"):
        assert torch.equal(a[k], b[k]), k
    prof = torch.as_tensor(c["prof"], device=dev)
    a2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], sb=sb)
    b2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], frames=store)
    for k in K.P2_F64:
        assert torch.equal(a2[k], b2[k]), k


def test_prepared_passes_give_identical_results():
    """xpg_prepare (gather, Kabsch frame, eps, J, neighbour masks ahead) + the prepared passes == the fused passes,
    bit for bit, with partners and neighbours."""
    from chunks import make_chunk
    from xpg_amd import integrate_kernels as K
    from xpg_amd.frames import DeviceFrames
    dev = torch.device("cuda:0")
    c = make_chunk(n=96, seed=23)
    assert np.asarray(c["rc"]["has_p"]).any() and np.isfinite(np.asarray(c["nb"]["x"])).any()
    nfr, ny, nx = 420, 1024, 1024
    store_np = np.random.default_rng(2).poisson(3.0, (nfr, ny, nx)).astype(np.uint32)
    frames = DeviceFrames(torch.as_tensor(store_np.view(np.int32), device=dev), first=1)
    store = frames.store()
    det = K.device_detector(c["det"], dev)
    call = K.make_call(c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"], c["rcw"])
    refl = K.DeviceRefl(c["rc"], c["nb"], (7, 11, 11), dev)
    print("prepare n=96 7x11x11", flush=True)
    ps, bufs = K.run_prepare(det, call, refl, store, dev)
    assert torch.equal(bufs["sb"], K.run_extract(store, refl, dev))
    a = K.run_pass1(det, call, refl, prep=ps)
    b = K.run_pass1(det, call, refl, frames=store)
    for k in K.P1_F64 + K.P1_I32 + ("fz", "img2", "J", "u1", "u2", "all_valid2"):
        assert torch.equal(a[k], b[k]), k
    prof = torch.as_tensor(c["prof"], device=dev)
    a2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], prep=ps)
    b2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], frames=store)
    for k in K.P2_F64:
        assert torch.equal(a2[k], b2[k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("mixture", [True, False])
def test_rocking_table_gives_identical_results(mixture):
    """xpg_rocking's per-row rocking-curve table read by the prepared passes == the passes' own evaluation, bit for bit
    (mixture weights and the plain Gaussian; a table wider than the boxes)."""
    import ctypes as C
    from chunks import make_chunk
    from xpg_amd import _lib as L
    from xpg_amd import integrate_kernels as K
    from xpg_amd.frames import DeviceFrames
    dev = torch.device("cuda:0")
    c = make_chunk(n=80, seed=31)
    nfr, ny, nx = 420, 1024, 1024
    store_np = np.random.default_rng(5).poisson(3.0, (nfr, ny, nx)).astype(np.uint32)
    frames = DeviceFrames(torch.as_tensor(store_np.view(np.int32), device=dev), first=1)
    store = frames.store()
    det = K.device_detector(c["det"], dev)
    call = K.make_call(c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"],
                       c["rcw"] if mixture else None)
    refl = K.DeviceRefl(c["rc"], c["nb"], (7, 11, 11), dev)
    print(f"rocking table n=80 7x11x11 mixture={mixture}", flush=True)
    ps, bufs = K.run_prepare(det, call, refl, store, dev)
    R = torch.full((80, 9), float("nan"), dtype=torch.float64, device=dev)
    pt = L.Prepared.from_buffer_copy(ps)
    pt.R, pt.r_stride = L.ptr(R), 9
    L.check(L.lib().xpg_rocking(C.byref(call), C.byref(refl.s), C.byref(pt), L.stream_ptr(dev)))
    assert torch.isfinite(R[:, :7]).all() and torch.isnan(R[:, 7:]).all()
    a = K.run_pass1(det, call, refl, prep=pt)
    b = K.run_pass1(det, call, refl, prep=ps)
    for k in K.P1_F64 + K.P1_I32 + ("fz", "img2"):
        assert torch.equal(a[k], b[k]), k
    prof = torch.as_tensor(c["prof"], device=dev)
    a2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], prep=pt)
    b2 = K.run_pass2(det, call, refl, prof, a["bg"], a["nbg"], prep=ps)
    for k in K.P2_F64:
        assert torch.equal(torch.nan_to_num(a2[k], nan=-7.0), torch.nan_to_num(b2[k], nan=-7.0)), k
    pt.r_stride = 6
    with pytest.raises(L.XpgError, match="r_stride"):
        K.run_pass1(det, call, refl, prep=pt)
