"""Write golden vectors of XPG's integration kernels by running the reference's own code on CPU.

Run in the development container only (the reference tree is not available elsewhere):

    python tests/golden/make_golden.py /root/reference

It imports xpg.integrate_kernels from the reference (torch/numpy/scipy only; no gemmi), builds seeded synthetic
chunks with tests/chunks.py, and records for each case the inputs and the outputs of the reference's pass1,
accumulate_profiles and pass2 (xpg/integrate_kernels.py:287-483) in two modes:

  * out64: every floating input given as float64, so the reference evaluates its own algorithm in double
    precision -- this pins the FP64 oracle (agreement ~1e-12, identical masks);
  * out32: the float32 inputs xpg/integrate.py:509-519 actually passes -- this measures the reference's own
    float32 drift against the FP64 oracle.

It also records _phi_rc, rc_basis_np, _bilinear, _neighbour_mask, gaussian_profile and rc_default_weights on random
inputs.  The .npz files hold data only (inputs and expected outputs).
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

CASES = dict(
    a_7x11x11=dict(n=48, Dz=7, Dy=11, Dx=11, seed=1),
    b_5x9x9_template=dict(n=40, Dz=5, Dy=9, Dx=9, seed=2, bkg_template=True, rcw="random"),
    c_17x11x11=dict(n=24, Dz=17, Dy=11, Dx=11, seed=3, intensity=(20.0, 2e5)),
    d_9x9x9_binned=dict(n=24, Dz=9, Dy=9, Dx=9, seed=4, fb=2),
    e_5x9x9_superpixel=dict(n=24, Dz=5, Dy=9, Dx=9, seed=5, pxs=2),
    f_3x2x3_smallbox=dict(n=16, Dz=3, Dy=2, Dx=3, seed=6, partners=False),
    g_7x11x11_parallax=dict(n=24, Dz=7, Dy=11, Dx=11, seed=7, parallax=True),
    h_5x9x9_weak=dict(n=64, Dz=5, Dy=9, Dx=9, seed=8, intensity=(1.0, 60.0), background=0.5),
)

RC_KEYS = ("S", "phi", "ix0", "iy0", "iz0", "fb", "region", "has_p", "pS", "pphi", "pxs")
P1_KEYS = ("bg", "nbg", "isum", "vsum", "vsum0", "xo", "yo", "zo", "m_d", "m_m", "fz", "img2", "all_valid2",
           "frac_rec", "n_over", "n_fg_masked", "nfg", "zeta", "J", "u1", "u2", "seff", "n_fg_nb")
P2_KEYS = ("iprf", "vprf", "vprf0", "corr", "p_valid", "p_total", "ipartner", "rho", "p_fg")


def _chunk(case):
    from chunks import flat_detector, make_chunk
    kw = dict(case)
    det = None
    if kw.pop("parallax", False):
        det = flat_detector(sensor_thickness=0.45, mu=3.7)
    return make_chunk(det=det, **kw)


def run_reference(ref_root, c, mode):
    import torch
    sys.path.insert(0, str(ref_root))
    from xpg import tnp as jnp
    from xpg.integrate_kernels import make_kernels

    f = torch.float64 if mode == 64 else torch.float32
    npf = np.float64 if mode == 64 else np.float32
    T = lambda a, dt=f: torch.as_tensor(np.asarray(a), dtype=dt)
    det = c["det"]
    d = dict(orgx=det["orgx"], orgy=det["orgy"], P0=T(det["P0"]), Bx=T(det["Bx"]), By=T(det["By"]), U1=T(det["U1"]),
             U2=T(det["U2"]), UN=T(det["UN"]), qx=det["qx"], qy=det["qy"], segmap=torch.zeros((1, 1), dtype=torch.int16),
             gx=T(np.zeros((1, 1))), gy=T(np.zeros((1, 1))), wl=det["wl"], tsens=det["tsens"], mu=det["mu"],
             plx=det["plx"], bkg=T(det["bkg"]) if det["bkg"] is not None else T(np.ones((1, 1)), torch.float32))
    rc = c["rc"]
    r = dict(S=T(rc["S"]), phi=T(rc["phi"]), ix0=torch.as_tensor(rc["ix0"]), iy0=torch.as_tensor(rc["iy0"]),
             iz0=torch.as_tensor(rc["iz0"]), fb=T(rc["fb"]), region=torch.as_tensor(rc["region"].astype(np.int32)),
             has_p=T(rc["has_p"]), pS=T(rc["pS"]), pphi=T(rc["pphi"]))
    if "pxs" in rc:
        r["pxs"] = T(rc["pxs"])
    nb = {k: torch.as_tensor(np.asarray(v, np.float32)) for k, v in c["nb"].items()}
    sc = lambda v: float(v) if mode == 64 else np.float32(v)
    with jnp.on("cpu"):
        p1, acc, p2 = make_kernels(c["Dz"], c["Dy"], c["Dx"], 9)
        sb = T(c["sb"])
        prof0 = torch.as_tensor(np.repeat(np.asarray(c["prof"])[:1] * 0 + np.asarray(c["prof"])[:1], 9, 0))
        rcw = torch.as_tensor(np.asarray(c["rcw"], np.float32))
        r1 = p1(sb, r, d, T(c["axis"]), T(c["s0"]), sc(c["frames_phi0"]), sc(c["osc"]), sc(c["sd"]), sc(c["sm"]),
                sc(c["gain"]), prof0, nb, rcw)
        sel = torch.as_tensor(np.asarray(r1["isum"]) > 0)
        num, den = acc(r1["img2"], r1["all_valid2"], r1["J"], r1["u1"], r1["u2"], r["region"], r1["isum"], sel,
                       r1["seff"])
        prof = torch.as_tensor(np.asarray(c["prof"], np.float32))
        pix = {k: r1[k] for k in ("zeta", "J", "u1", "u2", "seff")}
        r2 = p2(sb, r, d, T(c["axis"]), T(c["s0"]), sc(c["frames_phi0"]), sc(c["osc"]), sc(c["sd"]), sc(c["sm"]),
                sc(c["gain"]), prof, r1["bg"], r1["nbg"], nb, rcw, pix)
    out = {}
    for k in P1_KEYS:
        out["p1_" + k] = np.asarray(r1[k].numpy())
    out["acc_num"], out["acc_den"], out["acc_sel"] = num.numpy(), den.numpy(), sel.numpy()
    for k in P2_KEYS:
        out["p2_" + k] = np.asarray(r2[k].numpy())
    return out


def helpers(ref_root, rng):
    import torch
    sys.path.insert(0, str(ref_root))
    from xpg import tnp as jnp
    from xpg.integrate_kernels import (_bilinear, _neighbour_mask, _phi_rc, gaussian_profile, rc_basis_np,
                                       rc_default_weights)
    n, Dz = 40, 9
    lo = rng.uniform(-0.02, 0.02, (n, Dz))
    hi = lo + rng.uniform(1e-4, 5e-3, (n, Dz))
    phic = rng.uniform(-0.02, 0.02, (n, Dz))
    zeta = rng.uniform(-1, 1, (n, Dz))
    zeta[0, 0] = 1e-5
    sm = np.radians(0.15)
    rcw = rng.gamma(2.0, 1.0, 21).astype(np.float32)
    rcw /= rcw.sum()
    T = lambda a: torch.as_tensor(a, dtype=torch.float64)
    with jnp.on("cpu"):
        R_mix = _phi_rc(T(lo), T(hi), T(phic), T(zeta), sm, torch.as_tensor(rcw)).numpy()
        R_gauss = _phi_rc(T(lo), T(hi), T(phic), T(zeta), sm, None).numpy()
        basis = rc_basis_np(lo, hi, phic[:, 0], zeta[:, 0], sm)
        u1 = rng.uniform(-6, 6, (n, 7, 7))
        u2 = rng.uniform(-6, 6, (n, 7, 7))
        u1[0, 0, 0], u2[0, 0, 0] = 4.5, -4.5                  # on the grid edges
        inside, i1, i2, t1, t2 = (a.numpy() for a in _bilinear(torch.zeros((1, 41, 41)), T(u1), T(u2), 4.5))
        m, Dy, Dx = 20, 9, 11
        xx = T(rng.integers(50, 60, (m, 1, 1, 1)) + np.arange(Dx)[None, None, None, :] + 1.0)
        yy = T(rng.integers(50, 60, (m, 1, 1, 1)) + np.arange(Dy)[None, None, :, None] + 1.0)
        zz = T(rng.integers(5, 9, (m, 1, 1, 1)) + np.arange(7)[None, :, None, None] + 0.5)
        nb = dict(x=rng.uniform(45, 75, (m, 8)).astype(np.float32), y=rng.uniform(45, 75, (m, 8)).astype(np.float32),
                  z=rng.uniform(2, 18, (m, 8)).astype(np.float32), sx=rng.uniform(1, 4, (m, 8)).astype(np.float32),
                  sz=rng.uniform(0.5, 3, (m, 8)).astype(np.float32))
        nb["x"][:, 5:] = -1e6
        skip = rng.random(m) < 0.5
        nbt = {k: torch.as_tensor(v) for k, v in nb.items()}
        nbm = _neighbour_mask(xx, yy, zz, nbt, (m, 7, Dy, Dx)).numpy()
        nbm_skip = _neighbour_mask(xx, yy, zz, nbt, (m, 7, Dy, Dx), skip_first=torch.as_tensor(skip)).numpy()
    return dict(lo=lo, hi=hi, phic=phic, zeta=zeta, sm=np.float64(sm), rcw=rcw, R_mix=R_mix, R_gauss=R_gauss,
                basis=basis, u1=u1, u2=u2, inside=inside, i1=i1, i2=i2, t1=t1, t2=t2, xx=xx.numpy(), yy=yy.numpy(),
                zz=zz.numpy(), nb_x=nb["x"], nb_y=nb["y"], nb_z=nb["z"], nb_sx=nb["sx"], nb_sz=nb["sz"], skip=skip,
                nbm=nbm, nbm_skip=nbm_skip, gauss=gaussian_profile(), rcw_default=rc_default_weights())


def main(ref_root):
    os.environ["XPG_DEVICE"] = "cpu"
    os.environ.setdefault("XPG_TORCH_THREADS", "4")
    out_dir = HERE
    for name, case in CASES.items():
        c = _chunk(case)
        rec = dict(Dz=c["Dz"], Dy=c["Dy"], Dx=c["Dx"], sb=c["sb"].astype(np.float32), axis=c["axis"], s0=c["s0"],
                   frames_phi0=np.float64(c["frames_phi0"]), osc=np.float64(c["osc"]), sd=np.float64(c["sd"]),
                   sm=np.float64(c["sm"]), gain=np.float64(c["gain"]), rcw=np.asarray(c["rcw"], np.float32),
                   prof=c["prof"])
        for k in RC_KEYS:
            if k in c["rc"]:
                rec["rc_" + k] = np.asarray(c["rc"][k])
        for k, v in c["nb"].items():
            rec["nb_" + k] = v
        det = c["det"]
        for k in ("orgx", "orgy", "qx", "qy", "wl", "tsens", "mu", "plx"):
            rec["det_" + k] = np.float64(det[k])
        for k in ("P0", "Bx", "By", "U1", "U2", "UN"):
            rec["det_" + k] = det[k]
        rec["det_bkg_analytic"] = np.int64(det["bkg"] is not None)     # tests/chunks.py:analytic_template
        for mode in (64, 32):
            o = run_reference(ref_root, c, mode)
            for k, v in o.items():
                rec[f"out{mode}_{k}"] = v
        np.savez_compressed(out_dir / f"kernels_{name}.npz", **rec)
        print("wrote", name)
    np.savez_compressed(out_dir / "helpers.npz", **helpers(ref_root, np.random.default_rng(11)))
    print("wrote helpers")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
