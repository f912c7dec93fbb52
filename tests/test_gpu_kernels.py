"""Parity of the HIP kernels (through the C ABI) with the FP64 oracle, on the golden chunks and seeded synthetic ones.

Bars: integer masks and counts bit-exact; |dI|/sigma <= 1e-3 for the profile-fitted and summation intensities
(observed ~1e-6); other float outputs within the relative tolerances stated per key.
"""
import numpy as np
import pytest
import torch

from golden_io import cases, load
from oracle import kernels as ok

pytestmark = pytest.mark.gpu

P1_TOL = dict(bg=1e-6, nbg=1e-9, isum=None, vsum=1e-6, vsum0=1e-6, xo=1e-5, yo=1e-5, zo=1e-5, frac_rec=1e-9,
              zeta=1e-12, seff=1e-12)
P2_TOL = dict(vprf=1e-5, vprf0=1e-5, corr=1e-4, p_valid=1e-5, p_total=1e-5, rho=1e-5, p_fg=1e-5)


def _args(c):
    return (c["rc"], c["det"], c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"])


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = np.maximum(np.abs(b), np.max(np.abs(b)) * 1e-3 + 1e-300)
    return float(np.max(np.abs(a - b) / scale))


def _run(c, dev):
    from xpg_amd.integrate_kernels import make_kernels
    p1, acc, p2 = make_kernels(c["Dz"], c["Dy"], c["Dx"], 9)
    sb = torch.as_tensor(c["sb"], dtype=torch.float32, device=dev)
    r1 = p1(sb, *_args(c), c["prof"], c["nb"], c["rcw"])
    r2 = p2(sb, *_args(c), c["prof"], r1["bg"], r1["nbg"], c["nb"], c["rcw"])
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in r1.items()}, {k: v.cpu().numpy() for k, v in r2.items()}


def _check_pass1(h1, o1, name):
    for k in ("n_over", "n_fg_masked", "n_fg_nb"):
        assert np.array_equal(h1[k], np.asarray(o1[k])), (name, k)
    assert np.array_equal(h1["nfg"], o1["nfg"]), name
    assert np.array_equal(h1["all_valid2"], o1["all_valid2"]), name
    for k, tol in P1_TOL.items():
        if tol is not None:
            assert _rel(h1[k], o1[k]) < tol, (name, k, _rel(h1[k], o1[k]))
    sig = np.sqrt(np.maximum(o1["vsum"], 1e-6))
    assert np.max(np.abs(h1["isum"] - o1["isum"]) / sig) < 1e-4, name
    for k in ("J", "u1", "u2"):
        assert _rel(h1[k], o1[k]) < 1e-6, (name, k)
    scale = np.max(np.abs(o1["img2"]))
    assert np.max(np.abs(h1["img2"] - o1["img2"])) < 1e-5 * scale, name
    assert np.max(np.abs(h1["fz"] - o1["fz"])) < 1e-5 * np.max(np.abs(o1["fz"])), name
    for k in ("m_d", "m_m"):
        assert np.max(np.abs(h1[k] - o1[k])) < 1e-4 * np.max(np.abs(o1[k])), (name, k)


def _check_pass2(h2, o2, name):
    vp = np.asarray(o2["vprf"])
    pos = vp > 0      # a reflection with no fit voxel gives I = 0, var = 0 (xpg/integrate_kernels.py:445-467)
    assert np.all(np.abs(h2["iprf"][~pos]) <= 1e-12) and np.all(np.abs(h2["vprf"][~pos]) <= 1e-12), name
    sig = np.sqrt(np.where(pos, vp, 1.0))
    dI = np.where(pos, np.abs(h2["iprf"] - o2["iprf"]) / sig, 0.0)
    assert np.isfinite(h2["iprf"]).all() and dI.max() < 1e-3, (name, dI.max())
    for k, tol in P2_TOL.items():
        assert _rel(h2[k], o2[k]) < tol, (name, k, _rel(h2[k], o2[k]))
    hp = np.asarray(o2["ipartner"]) != 0
    if hp.any():      # the partner's own scale: float32-level relative agreement
        err = np.abs(h2["ipartner"][hp] - o2["ipartner"][hp])
        assert np.all(err <= 1e-5 * np.abs(o2["ipartner"][hp]) + 1e-3 * sig[hp]), name
    return dI.max()


@pytest.mark.parametrize("name", cases())
def test_kernels_match_oracle_on_golden_chunks(name):
    c, _ = load(name)
    h1, h2 = _run(c, torch.device("cuda:0"))
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    _check_pass1(h1, o1, name)
    o2 = ok.pass2(c["sb"], *_args(c), c["prof"], o1["bg"], o1["nbg"], c["nb"], c["rcw"], irls_rule="reflection")
    d = _check_pass2(h2, o2, name)
    print(f"{name}: max |dI|/sigma {d:.2e}")


@pytest.mark.parametrize("seed,shape,kw", [
    (101, (7, 11, 11), {}),
    (102, (9, 9, 9), dict(fb=3)),
    (103, (5, 13, 13), dict(pxs=3)),
    (104, (17, 21, 21), dict(intensity=(10.0, 1e6))),
    (105, (33, 9, 9), dict(sm_deg=0.4)),
    (106, (5, 9, 9), dict(intensity=(1.0, 2 ** 20 * 20), background=0.5)),
    (107, (33, 31, 31), dict(n=12, sd_deg=0.06, sm_deg=0.5)),    # too large for LDS: values read in place
])
def test_kernels_match_oracle_synthetic(seed, shape, kw):
    from chunks import make_chunk
    kw = dict(kw)
    c = make_chunk(n=kw.pop("n", 96), Dz=shape[0], Dy=shape[1], Dx=shape[2], seed=seed, **kw)
    h1, h2 = _run(c, torch.device("cuda:0"))
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    _check_pass1(h1, o1, seed)
    o2 = ok.pass2(c["sb"], *_args(c), c["prof"], o1["bg"], o1["nbg"], c["nb"], c["rcw"], irls_rule="reflection")
    _check_pass2(h2, o2, seed)


def test_accumulate_and_normalise_match_oracle():
    from xpg_amd.integrate_kernels import FX_SCALE, make_kernels, new_profile_acc, normalise_profiles
    c, _ = load("a_7x11x11")
    dev = torch.device("cuda:0")
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    sel = np.asarray(o1["isum"]) > 0
    f32 = lambda a: np.asarray(a, np.float32)
    _, acc, _ = make_kernels(c["Dz"], c["Dy"], c["Dx"], 9)
    T = lambda a, dt: torch.as_tensor(np.asarray(a), dtype=dt, device=dev)
    num, den = acc(T(f32(o1["img2"]), torch.float32), T(o1["all_valid2"], torch.uint8), T(f32(o1["J"]), torch.float32),
                   T(f32(o1["u1"]), torch.float32), T(f32(o1["u2"]), torch.float32),
                   T(c["rc"]["region"], torch.int32), T(o1["isum"], torch.float64), T(sel, torch.uint8),
                   T(o1["seff"], torch.float64))
    # oracle on the same float32-rounded per-pixel inputs
    onum, oden = ok.accumulate_profiles(f32(o1["img2"]), o1["all_valid2"], f32(o1["J"]), f32(o1["u1"]), f32(o1["u2"]),
                                        c["rc"]["region"], o1["isum"], sel, o1["seff"], 9)
    num, den = num.cpu().numpy(), den.cpu().numpy()
    assert np.max(np.abs(num - onum)) < 1e-8 * max(np.max(np.abs(onum)), 1)
    assert np.max(np.abs(den - oden)) < 1e-8 * max(np.max(np.abs(oden)), 1)
    # normalisation on the device against the float64 recipe
    num_fx = torch.as_tensor(np.rint(onum * FX_SCALE).astype(np.int64), device=dev)
    den_fx = torch.as_tensor(np.rint(oden * FX_SCALE).astype(np.int64), device=dev)
    prof, _, _ = normalise_profiles(num_fx, den_fx, dev)
    oprof = ok.normalise_profiles(onum, oden)
    assert np.max(np.abs(prof.cpu().numpy() - oprof)) < 1e-6 * np.max(oprof)


def test_fused_accumulation_matches_strong_selection():
    """pass1 with the profile grids fused in: the strong rows of xpg/integrate.py:530-541 splatted on the device."""
    from chunks import make_chunk
    from xpg_amd import integrate_kernels as K
    c = make_chunk(n=200, seed=77, intensity=(500.0, 1e5), codes=False, neighbours=False)
    dev = torch.device("cuda:0")
    det = K.device_detector(c["det"], dev)
    call = K.make_call(c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"], c["rcw"])
    refl = K.DeviceRefl(c["rc"], c["nb"], (7, 11, 11), dev)
    num_fx, den_fx, acc = K.new_profile_acc(9, dev)
    sb = torch.as_tensor(c["sb"], dtype=torch.float32, device=dev)
    K.run_pass1(det, call, refl, sb=sb, acc=acc)
    torch.cuda.synchronize()
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    strong = ((o1["isum"] > 0) & (o1["isum"] / np.sqrt(np.maximum(o1["vsum"], 1e-6)) > 10) & (o1["n_over"] == 0)
              & (o1["frac_rec"] > 0.95) & (o1["n_fg_nb"] == 0) & (o1["n_fg_masked"] == 0))
    assert strong.sum() > 20
    onum, oden = ok.accumulate_profiles(o1["img2"], o1["all_valid2"], o1["J"], o1["u1"], o1["u2"], c["rc"]["region"],
                                        o1["isum"], strong, o1["seff"], 9)
    num = num_fx.cpu().numpy() / K.FX_SCALE
    den = den_fx.cpu().numpy() / K.FX_SCALE
    assert np.max(np.abs(den - oden)) < 1e-9 * np.max(oden)
    assert np.max(np.abs(num - onum)) < 1e-9 * np.max(np.abs(onum))
    # bit-identical on a second launch (integer accumulation is order-independent)
    num2, den2, acc2 = K.new_profile_acc(9, dev)
    K.run_pass1(det, call, refl, sb=sb, acc=acc2)
    torch.cuda.synchronize()
    assert torch.equal(num2, num_fx) and torch.equal(den2, den_fx)


def _masks(c, dev):
    """Per-voxel classes of pass 1 and pass 2 from the device (XPG_VM_* bits), through the C ABI."""
    from xpg_amd import integrate_kernels as K
    det = K.device_detector(c["det"], dev)
    call = K.make_call(c["axis"], c["s0"], c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"], c["rcw"])
    refl = K.DeviceRefl(c["rc"], c["nb"], (c["Dz"], c["Dy"], c["Dx"]), dev)
    sb = torch.as_tensor(c["sb"], dtype=torch.float32, device=dev)
    r1 = K.run_pass1(det, call, refl, sb=sb, masks=True)
    r2 = K.run_pass2(det, call, refl, c["prof"], r1["bg"], r1["nbg"], sb=sb, masks=True)
    torch.cuda.synchronize()
    return r1["vmask"].cpu().numpy(), r2["vmask"].cpu().numpy()


MASK_CASES = [("golden", n) for n in cases()] + [
    ("synthetic", (201, (7, 11, 11), {})), ("synthetic", (202, (9, 9, 9), dict(fb=3))),
    ("synthetic", (203, (5, 13, 13), dict(pxs=3))), ("synthetic", (204, (17, 21, 21), dict(intensity=(10.0, 1e6)))),
    ("synthetic", (205, (33, 31, 31), dict(n=12, sd_deg=0.06, sm_deg=0.5))),
    ("synthetic", (206, (5, 9, 9), dict(intensity=(1.0, 2 ** 20 * 20), background=0.5)))]


@pytest.mark.parametrize("kind,case", MASK_CASES, ids=[str(c[1] if c[0] == "golden" else c[1][0]) for c in MASK_CASES])
def test_voxel_masks_bit_exact(kind, case):
    """Every voxel's class, bit for bit against the oracle's masks: foreground (ellipse x foreground frame),
    background candidate, neighbour ellipsoid, final background set after the clipping rounds, valid foreground;
    pass 2's fit mask and neighbour mask (partner column skipped).  Boxes too large for LDS (33 x 31 x 31) included."""
    from xpg_amd.integrate_kernels import P2_F64  # noqa: F401  (the library loads)
    if kind == "golden":
        c, _ = load(case)
    else:
        from chunks import make_chunk
        seed, shape, kw = case
        kw = dict(kw)
        c = make_chunk(n=kw.pop("n", 96), Dz=shape[0], Dy=shape[1], Dx=shape[2], seed=seed, **kw)
    print("voxel masks", kind, case if kind == "golden" else case[0], flush=True)
    m1, m2 = _masks(c, torch.device("cuda:0"))
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    o2 = ok.pass2(c["sb"], *_args(c), c["prof"], o1["bg"], o1["nbg"], c["nb"], c["rcw"], irls_rule="reflection")
    valid = np.asarray(c["sb"]) >= 0
    want1 = {1: o1["_fg"], 2: o1["_bgm"], 4: o1["_nbm"], 8: o1["_w"], 16: o1["_fg"] & valid & ~o1["_nbm"]}
    for bit, w in want1.items():
        got = (m1 & bit) != 0
        assert np.array_equal(got, np.broadcast_to(w, got.shape)), (case, bit, int((got != w).sum()))
    assert np.array_equal((m2 & 32) != 0, o2["_use"]), (case, int((((m2 & 32) != 0) != o2["_use"]).sum()))
    assert np.array_equal((m2 & 4) != 0, np.broadcast_to(o2["_nbm"], m2.shape)), case
    assert o1["_fg"].any() and o2["_use"].any()
    assert o1["_w"].any() or not o1["_bgm"].any()      # (the 3 x 2 x 3 boxes have no background ring at all)


def test_make_kernels_with_the_reference_argument_types():
    """The closures of make_kernels called as xpg/integrate.py:466-563 calls them: a chunk on the device, per-reflection
    arrays as float32 / int64 device tensors, the detector dict of _detector_consts (float32 device tensors, (1, 1)
    placeholders, Python scalars), axis / s0 / rcw as float32 device tensors, numpy float32 scalars, and pass 2 given
    pass 1's per-pixel geometry ``pix``.  Results against the oracle run on the same float32 inputs."""
    from chunks import make_chunk
    from xpg_amd.integrate_kernels import make_kernels
    dev = torch.device("cuda:0")
    c = make_chunk(n=96, seed=211)
    f32d = lambda a: torch.as_tensor(np.asarray(a, np.float32), device=dev)
    rc = dict(S=f32d(c["rc"]["S"]), phi=f32d(c["rc"]["phi"]), ix0=torch.as_tensor(c["rc"]["ix0"], device=dev),
              iy0=torch.as_tensor(c["rc"]["iy0"], device=dev), iz0=torch.as_tensor(c["rc"]["iz0"], device=dev),
              fb=f32d(c["rc"]["fb"]), region=torch.as_tensor(c["rc"]["region"], device=dev),
              has_p=f32d(c["rc"]["has_p"]), pS=f32d(c["rc"]["pS"]), pphi=f32d(c["rc"]["pphi"]))
    d = c["det"]
    det = dict(orgx=d["orgx"], orgy=d["orgy"], qx=d["qx"], qy=d["qy"], wl=d["wl"], tsens=d["tsens"], mu=d["mu"],
               plx=d["plx"], segmap=torch.zeros((1, 1), dtype=torch.int16, device=dev), gx=f32d(np.zeros((1, 1))),
               gy=f32d(np.zeros((1, 1))), bkg=f32d(np.ones((1, 1))),
               **{k: f32d(d[k]) for k in ("P0", "Bx", "By", "U1", "U2", "UN")})
    nb = {k: f32d(v) for k, v in c["nb"].items()}
    scal = [np.float32(v) for v in (c["frames_phi0"], c["osc"], c["sd"], c["sm"], c["gain"])]
    p1, acc, p2 = make_kernels(7, 11, 11, 9)
    sb = torch.as_tensor(c["sb"], dtype=torch.float32, device=dev)
    prof0 = f32d(c["prof"])
    r1 = p1(sb, rc, det, f32d(c["axis"]), f32d(c["s0"]), *scal, prof0, nb, f32d(c["rcw"]))
    pix = {k: r1[k] for k in ("zeta", "J", "u1", "u2", "seff")}
    r2 = p2(sb, rc, det, f32d(c["axis"]), f32d(c["s0"]), *scal, prof0, r1["bg"], r1["nbg"], nb, f32d(c["rcw"]), pix)
    torch.cuda.synchronize()
    # the oracle on the same float32-rounded inputs
    r32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    orc = dict(c["rc"])
    for k in ("S", "phi", "pS", "pphi"):
        orc[k] = r32(orc[k])
    odet = dict(d)
    for k in ("P0", "Bx", "By", "U1", "U2", "UN"):
        odet[k] = r32(d[k])
    oargs = (orc, odet, r32(c["axis"]), r32(c["s0"]), *[float(v) for v in scal])
    o1 = ok.pass1(c["sb"], *oargs, c["nb"], c["rcw"])
    for k in ("n_over", "n_fg_masked", "n_fg_nb"):
        assert np.array_equal(r1[k].cpu().numpy(), np.asarray(o1[k])), k
    opix = dict(zeta=r1["zeta"].cpu().numpy(), J=r1["J"].cpu().numpy().astype(np.float64),
                u1=r1["u1"].cpu().numpy().astype(np.float64), u2=r1["u2"].cpu().numpy().astype(np.float64),
                seff=r1["seff"].cpu().numpy().astype(np.float32).astype(np.float64))
    o2 = ok.pass2(c["sb"], *oargs, c["prof"], r1["bg"].cpu().numpy(), r1["nbg"].cpu().numpy(), c["nb"], c["rcw"],
                  pix=opix, irls_rule="reflection")
    _check_pass2({k: v.cpu().numpy() for k, v in r2.items() if torch.is_tensor(v)}, o2, "reference argument types")
    # pix is honoured: a pass 2 given another width than its own sigma_D follows the given one
    pix_w = dict(pix, seff=pix["seff"] * 1.25, u1=pix["u1"] / 1.25, u2=pix["u2"] / 1.25)
    r2w = p2(sb, rc, det, f32d(c["axis"]), f32d(c["s0"]), *scal, prof0, r1["bg"], r1["nbg"], nb, f32d(c["rcw"]),
             pix_w)
    torch.cuda.synchronize()
    ow = ok.pass2(c["sb"], *oargs, c["prof"], r1["bg"].cpu().numpy(), r1["nbg"].cpu().numpy(), c["nb"], c["rcw"],
                  pix=dict(opix, seff=(pix["seff"] * 1.25).cpu().numpy().astype(np.float32).astype(np.float64),
                           u1=(r1["u1"] / 1.25).cpu().numpy().astype(np.float64),
                           u2=(r1["u2"] / 1.25).cpu().numpy().astype(np.float64)), irls_rule="reflection")
    _check_pass2({k: v.cpu().numpy() for k, v in r2w.items() if torch.is_tensor(v)}, ow, "pix with another width")
    assert not np.allclose(r2w["iprf"].cpu().numpy(), r2["iprf"].cpu().numpy(), rtol=1e-3)


def test_kernels_on_a_segmented_detector_with_parallax_and_distortion():
    """Two tilted, shifted segments with a gap, 0.45 mm sensor parallax and geometric-correction tables, routed from
    the host Geometry through detector_consts to the kernels (segment map, per-segment maps, gx / gy): pass 1 and
    pass 2 against the oracle, masks bit-exact, boxes that straddle the gap included."""
    from chunks import make_chunk
    from xpg_amd.geometry import Geometry, Segment, detector_consts
    th = np.radians(0.4)
    segs = [Segment(1, 2068, 1, 1000),
            Segment(1, 2068, 1040, 2162, orgx=0.4, orgy=-0.3, distance=0.2,
                    ex=np.array([np.cos(th), np.sin(th), 0.0]), ey=np.array([-np.sin(th), np.cos(th), 0.0]))]
    yy, xx = np.mgrid[0:2162, 0:2068]
    geo = ((0.5 * np.sin(xx / 150.0)).astype(np.float32), (0.4 * np.cos(yy / 200.0)).astype(np.float32))
    g = Geometry(wavelength=1.0, beam=np.array([0, 0, 1.0]), axis=np.array([1.0, 0, 0]), orgx=1030, orgy=1080,
                 distance=150, ed1=np.array([1.0, 0, 0]), ed2=np.array([0, 1.0, 0]), qx=0.075, qy=0.075, nx=2068,
                 ny=2162, phi0=0, osc=0.1, sensor_thickness=0.45, mu=3.7, segments=segs, geo_corr=geo)
    det = dict(detector_consts(g), nx=g.nx, ny=g.ny)
    c = make_chunk(n=128, seed=221, det=det)
    print("segmented detector chunk", flush=True)
    h1, h2 = _run(c, torch.device("cuda:0"))
    o1 = ok.pass1(c["sb"], *_args(c), c["nb"], c["rcw"])
    _check_pass1(h1, o1, "segmented")
    o2 = ok.pass2(c["sb"], *_args(c), c["prof"], o1["bg"], o1["nbg"], c["nb"], c["rcw"], irls_rule="reflection")
    _check_pass2(h2, o2, "segmented")
    m1, m2 = _masks(c, torch.device("cuda:0"))
    assert np.array_equal((m1 & 1) != 0, o1["_fg"]) and np.array_equal((m1 & 8) != 0, o1["_w"])
    assert np.array_equal((m2 & 32) != 0, o2["_use"])
    # the detector really is segmented where the boxes are: some boxes reach into the second segment or the gap
    y_box = c["rc"]["iy0"][:, None] + np.arange(c["Dy"])[None] + 1.0
    assert (y_box > 1000).any()
