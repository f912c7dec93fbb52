"""CPU oracle of a whole integration sweep -- TEST INFRASTRUCTURE ONLY.

Runs the block loop of xpg/integrate.py:463-574 with the oracle's pieces, in float64: for each phi block in the frame
loop's order, gather the shoeboxes of every shape group (oracle/extract.py), pass 1, accumulate the strong
reflections' projected images (xpg/integrate.py:525-543), fit the adaptive widths (oracle/adapt.py), normalise the
profiles, pass 2 (per-reflection IRLS rule).  The host plan (predictions, boxes, neighbours, block order) is taken
from the product's planning, which tests/test_planning.py pins separately.
"""
from __future__ import annotations

import numpy as np

from . import adapt as oa
from . import extract as ox
from . import kernels as ok

STRONG_KEYS = ("isum", "vsum", "n_over", "frac_rec", "n_fg_nb", "n_fg_masked")


def profile_strong(o1):
    """The rows whose projected images build the reference profiles (xpg/integrate.py:530-533): I/sigma > 10 by
    summation, no overload, > 95% of the rocking curve recorded, no foreground pixel claimed by a neighbour or masked.
    Pinned by rows just over and under each threshold (tests/test_oracle_pins.py)."""
    return ((o1["isum"] > 0) & (o1["isum"] / np.sqrt(np.maximum(o1["vsum"], 1e-6)) > 10) & (o1["n_over"] == 0)
            & (o1["frac_rec"] > 0.95) & (o1["n_fg_nb"] == 0) & (o1["n_fg_masked"] == 0))


def integrate_sweep(coded_frames, read, hp, exp_blocks, model, det, adaptive=True, learn_profiles=True, mask=None,
                    n_regions=9):
    """coded_frames: {0-based frame: float32 value-coded frame}; read: 0-based frames read; hp: the host plan of
    xpg_amd.integrate._integrate(plan_only=True).  Returns per-reflection outputs (prediction order)."""
    refl = hp["refl"]
    N = refl["x"].size
    out = {k: np.full(N, np.nan) for k in ("bg", "isum", "vsum", "iprf", "vprf", "corr", "frac_rec", "m_d")}
    out["n_over"] = np.zeros(N, np.int32)
    sd0, sm0 = np.radians(model.sigma_d), np.radians(model.sigma_m)
    adapt = dict(sd=sd0, sm=sm0, rcw=ok.rc_default_weights())
    block_params = {}
    for b in hp["block_order"]:
        rows = hp["blocks_refl"][b]
        if not rows.size:
            continue
        g = exp_blocks[b].geom
        sd_b, sm_b, rcw_b = (adapt["sd"], adapt["sm"], adapt["rcw"]) if adaptive else (sd0, sm0,
                                                                                       ok.rc_default_weights())
        frames_phi0, osc = np.radians(g.z_to_phi(0.0)), np.radians(g.osc)
        keys = sorted({(int(hp["bucket"][r]), int(hp["bxs"][r]), int(hp["pxs"][r])) for r in rows})
        groups = []
        for gk in keys:
            sel = rows[(hp["bucket"][rows] == gk[0]) & (hp["bxs"][rows] == gk[1]) & (hp["pxs"][rows] == gk[2])]
            sel = sel[np.argsort(~hp["has_partner"][sel], kind="stable")]
            B, bx, sp = gk
            sb = ox.gather(coded_frames, read, hp["ix0"][sel], hp["iy0"][sel], hp["iz0"][sel], hp["fb"][sel], sp, B,
                           bx, mask).astype(np.float64)
            pj = hp["pj"][sel]
            rc = dict(S=refl["S"][sel], phi=np.radians(refl["phi"][sel]), ix0=hp["ix0"][sel], iy0=hp["iy0"][sel],
                      iz0=hp["iz0"][sel], fb=hp["fb"][sel].astype(float), region=hp["region"][sel],
                      has_p=hp["has_partner"][sel].astype(float), pS=refl["S"][pj],
                      pphi=np.radians(refl["phi"][pj]))
            if sp > 1:
                rc["pxs"] = np.full(sel.size, float(sp))
            nb = {k: v[sel] for k, v in hp["nb"].items()}
            o1 = ok.pass1(sb, rc, det, g.axis, g.s0, frames_phi0, osc, sd_b, sm_b, model.gain, nb, rcw_b)
            groups.append((sel, sb, rc, nb, o1, B))
        num = np.zeros((n_regions, 41, 41))
        den = np.zeros((n_regions, 41, 41))
        if learn_profiles:
            for sel, sb, rc, nb, o1, B in groups:
                strong = profile_strong(o1)
                if strong.any():
                    nn, dd = ok.accumulate_profiles(o1["img2"], o1["all_valid2"], o1["J"], o1["u1"], o1["u2"],
                                                    rc["region"], o1["isum"], strong, o1["seff"], n_regions)
                    num += nn
                    den += dd
        if adaptive:
            cat = lambda k: np.concatenate([o1[k] for *_, o1, _ in groups])
            sel_all = np.concatenate([sel for sel, *_ in groups])
            fz = np.zeros((sel_all.size, 65), np.float32)
            o = 0
            for sel, sb, rc, nb, o1, B in groups:
                fz[o:o + sel.size, :B] = o1["fz"]
                o += sel.size
            Xv = g.pixel_to_lab(refl["x"][sel_all], refl["y"][sel_all])
            Xlen = np.linalg.norm(Xv, axis=1)
            cosa = np.abs(Xv @ g.ed3) / Xlen
            n, sm_new, w, sd_new, sd_fit, sm_fit = oa.adapt_block(
                cat("isum"), cat("vsum"), cat("n_over"), cat("frac_rec"), cat("m_d"), fz, refl["phi"][sel_all],
                refl["zeta_k"][sel_all], hp["iz0"][sel_all], hp["fb"][sel_all], Xlen, cosa, sd0, sm0,
                g.z_to_phi(0.0), g.osc, g.qx)
            block_params[b] = dict(n=n)
            if n >= 100:
                sm_b, rcw_b = sm_new, w
                adapt.update(sd=sd_new, sm=sm_new, rcw=w)
                block_params[b].update(sigma_m=np.degrees(sm_new), sigma_d_next=np.degrees(sd_new))
        prof = ok.normalise_profiles(num, den) if learn_profiles else \
            np.repeat(ok.gaussian_profile()[None], n_regions, 0)
        for sel, sb, rc, nb, o1, B in groups:
            o2 = ok.pass2(sb, rc, det, g.axis, g.s0, frames_phi0, osc, sd_b, sm_b, model.gain, prof, o1["bg"],
                          o1["nbg"], nb, rcw_b, irls_rule="reflection")
            for k in ("bg", "isum", "vsum", "frac_rec", "m_d"):
                out[k][sel] = o1[k]
            out["n_over"][sel] = o1["n_over"]
            for k in ("iprf", "vprf", "corr"):
                out[k][sel] = o2[k]
    return out, block_params
