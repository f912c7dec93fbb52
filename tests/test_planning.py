"""Host planning of the integration (box planning, neighbour lists, block order, prediction) against the per-reflection
oracle restatements and hand-computed cases.  CPU only."""
from types import SimpleNamespace

import numpy as np

from oracle import planning as op
from xpg_amd import geometry as G
from xpg_amd import integrate as I
from xpg_amd import synth


def _refl(rng, n, g, sigma_m):
    x = rng.uniform(1, g.nx, n)
    y = rng.uniform(1, g.ny, n)
    z = rng.uniform(0, 300, n)
    zeta_k = rng.choice([-1, 1], n) * np.exp(rng.uniform(np.log(0.05), 0, n))
    return dict(x=x, y=y, z=z, zeta_k=zeta_k)


def test_depth_rule_example():
    """sigma_M = osc = 0.1 deg: hz = ceil(4 / |zeta| + 0.5) >= 5, so dz >= 11 and the bucket is 17."""
    hz, fb, iz0, dz, bucket = op.box_depth(100.3, 1.0, 0.1, 0.1)
    assert (hz, fb, dz, bucket) == (5, 1, 11, 17) and iz0 == 100 - 5


def test_box_planning_matches_oracle():
    rng = np.random.default_rng(0)
    g = synth.make_geometry(synth.spec_full())
    for sigma_m, sigma_d in ((0.1, 0.02), (0.8, 0.05), (2.0, 0.2)):
        refl = _refl(rng, 400, g, sigma_m)
        model = I.IntegrationModel(sigma_d, sigma_m)
        keep, pl = I.plan_boxes(refl, g, model)
        for i in range(400):
            hz, fb, iz0, dz, bucket = op.box_depth(refl["z"][i], refl["zeta_k"][i], sigma_m, g.osc)
            assert (pl["fb"][i], pl["iz0"][i], pl["dz"][i], pl["bucket"][i]) == (fb, iz0, dz, bucket)
        r = {k: v[keep] for k, v in refl.items()}
        bxs, pxs, ix0, iy0 = I.plan_sides(r, g, model, pl["box"])
        Xv = g.pixel_to_lab(r["x"], r["y"])
        Lr = np.linalg.norm(Xv, axis=1)
        cosr = np.abs(Xv @ g.ed3) / Lr
        spx = np.sqrt((Lr * np.radians(sigma_d * 1.3) / g.qx / np.maximum(cosr, 0.4)) ** 2 + 1 / 12)
        for i in range(len(r["x"])):
            want = op.box_side(r["x"][i], r["y"][i], spx[i], pl["box"], 0)
            assert (bxs[i], pxs[i], ix0[i], iy0[i]) == want


def test_fixed_box_override():
    g = synth.make_geometry(synth.spec_full())
    refl = _refl(np.random.default_rng(1), 50, g, 0.1)
    keep, pl = I.plan_boxes(refl, g, I.IntegrationModel(0.02, 0.1, box_xy=11, box_z=7))
    assert keep.all() and (pl["bucket"] == 7).all() and (pl["fb"] == 1).all()
    assert np.array_equal(pl["iz0"], np.floor(refl["z"]).astype(int) - 3)


def test_neighbours_match_brute_force():
    rng = np.random.default_rng(2)
    x, y, z = rng.uniform(0, 60, 80), rng.uniform(0, 60, 80), rng.uniform(0, 30, 80)
    j, has, d = I._neighbours(x, y, z, 2.0, 1.5)
    want = op.neighbours(x, y, z, 2.0, 1.5)
    for i in range(80):
        assert list(j[i][has[i]]) == want[i]


def test_block_order_matches_frame_walk():
    rng = np.random.default_rng(3)
    for _ in range(20):
        nb = 12
        lo = sorted(rng.integers(0, 200, nb))
        hi = [l_ + int(rng.integers(0, 80)) for l_ in lo]
        blocks = [np.arange(int(rng.integers(0, 3))) for _ in range(nb)]
        frames = [f for f in range(1, 251) if rng.random() > 0.05]
        got = I.block_order(blocks, lo, hi, frames)
        assert got == op.frame_loop(lo, hi, [b.size > 0 for b in blocks], frames)


def test_prediction_geometry_consistent():
    """Predicted (x, y, z) map back to integer hkl through the inverse geometry (reference test_core.py:92-98)."""
    spec = synth.spec_small()
    exp = synth.make_experiment(spec)
    g = exp.geom
    pr = G.predict_reflections(g, exp.crystal.Astar, 3.0, z_range=(0, 100))
    S = g.pixel_to_s(pr["x"], pr["y"])
    p = S - g.s0
    R = G.rotation_matrices(g.axis, -np.radians(g.z_to_phi(pr["z"])))
    p0 = np.einsum("nij,nj->ni", R, p)
    hkl = p0 @ np.linalg.inv(exp.crystal.Astar).T
    assert np.max(np.abs(hkl - pr["hkl"])) < 1e-6
    assert np.all(np.abs(np.linalg.norm(S, axis=1) - 1 / g.wavelength) < 1e-12)


def test_centring_conditions():
    h = np.array([[1, 0, 0], [1, 1, 0], [1, 1, 1], [2, 0, 0], [1, 2, 3]])
    assert G.centring_allowed(h, "C 2 2 21").tolist() == [False, True, True, True, False]
    assert G.centring_allowed(h, "I 2 2 2").tolist() == [False, True, False, True, True]
    assert G.centring_allowed(h, "F 2 3").tolist() == [False, False, True, True, False]
    assert G.centring_allowed(h, "P 21 21 21").all()


def test_plan_only_full_workload_shapes():
    """Host plan of a short slice of the full-size workload: fixed 11 x 11 x 7 boxes, partners first per group."""
    spec = synth.replace(synth.spec_full(), nframes=100)
    exp = synth.make_experiment(spec)
    edges = synth.block_edges(spec)
    img = SimpleNamespace(first=1, last=100, nx=spec.nx, ny=spec.ny)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=11, box_z=7)
    hp = I._integrate(img, [exp] * (len(edges) - 1), edges, model, synth.edge_resolution(exp.geom), plan_only=True,
                      verbose=False)
    assert (hp["bxs"] == 11).all() and (hp["bucket"] == 7).all()
    assert len(hp["block_order"]) == 2 and hp["refl"]["x"].size > 20000


def test_drawn_boxes_and_size_bins():
    """Drawn per-reflection boxes (odd side log-U[5, 21], depth log-U[3, 15]) and the sweep plan's size bins: every
    block's rows are covered by consecutive bin launches, each bin's largest depth x largest side fits the bin's LDS
    bound, and the per-row voxel / pixel offsets pack each bin's rows back to back."""
    import torch
    from xpg_amd import _lib as L
    spec = synth.replace(synth.spec_uneven(), nframes=100)
    exp = synth.make_experiment(spec)
    edges = synth.block_edges(spec)
    img = SimpleNamespace(first=1, last=100, nx=spec.nx, ny=spec.ny)
    model = I.IntegrationModel(spec.sigma_d, spec.sigma_m)
    hp = I._integrate(img, [exp] * (len(edges) - 1), edges, model, synth.edge_resolution(exp.geom), plan_only=True,
                      verbose=False, _box_draw=synth.box_draw_for(spec))
    dz, bx = hp["dz"], hp["bxs"]
    assert (dz % 2 == 1).all() and (bx % 2 == 1).all()
    assert dz.min() == 3 and dz.max() == 15 and bx.min() == 5 and bx.max() == 21
    assert (hp["fb"] == 1).all() and (hp["pxs"] == 1).all() and np.array_equal(hp["bucket"], dz)
    # log-uniform: E[side^2] = (21^2 - 5^2) / (2 ln(21 / 5)) ~ 145 and E[depth] = 12 / ln 5 ~ 7.46 before rounding
    assert abs(np.mean(bx.astype(float) ** 2) / 145.0 - 1) < 0.05 and abs(np.mean(dz) / 7.46 - 1) < 0.05
    # the same reflection gets the same box whatever else is predicted
    d2, s2 = synth.draw_boxes(spec, {"hkl": hp["refl"]["hkl"][::-1]})
    assert np.array_equal(d2[::-1], dz) and np.array_equal(s2[::-1], bx)
    # partials: boxes reaching into the next or previous 50-frame block
    blk_lo = np.asarray(edges)[hp["refl"]["block"]]
    part = (hp["iz0"] < blk_lo) | (hp["iz0"] + dz > blk_lo + spec.block_frames)
    assert 0.05 < part.mean() < 0.3
    sp = I.SweepPlan(hp["refl"], hp["iz0"], hp["fb"], hp["bucket"], hp["bxs"], hp["pxs"], hp["ix0"], hp["iy0"],
                     hp["region"], hp["has_partner"], hp["pj"], hp["nb"], hp["blocks_refl"], hp["block_order"],
                     torch.device("cpu"))
    assert sorted(sp.perm.tolist()) == list(range(hp["refl"]["x"].size))
    voff, poff = sp.t["voff"].numpy(), sp.t["poff"].numpy()
    pos = 0
    for bi, (b, lb, start, m, (mB, mbx), wl) in enumerate(sp.bins):
        assert start == pos and m > 0
        pos += m
        rows = sp.perm[start:start + m]
        assert (hp["refl"]["block"][rows] == b).all()
        assert mB == dz[rows].max() and mbx == bx[rows].max()
        need = [L.lib().xpg_lds_bytes(int(B), int(s), int(s)) for B, s in set(zip(dz[rows], bx[rows]))]
        assert wl == max(need) and (lb == len(sp.lds_bins) or wl <= sp.lds_bins[lb])
        assert lb == 0 or min(need) > sp.lds_bins[lb - 1]
        V = (dz[rows] * bx[rows] ** 2).astype(np.int64)
        assert np.array_equal(voff[start:start + m], np.cumsum(V) - V)
        assert np.array_equal(poff[start:start + m], np.cumsum(bx[rows] ** 2) - bx[rows] ** 2)
        assert sp.bin_sizes[bi] == (int(V.sum()), int((bx[rows] ** 2).sum()))
    assert pos == sp.n
    # shape groups nest in bins, with partner rows first in each group
    for b, (B, s, _), start, m in sp.groups:
        rows = sp.perm[start:start + m]
        assert (dz[rows] == B).all() and (bx[rows] == s).all()
        hpr = hp["has_partner"][rows]
        assert not (~hpr[:-1] & hpr[1:]).any()
        assert any(s2 <= start and start + m <= s2 + m2 for _, _, s2, m2, _, _ in sp.bins)
    assert len(sp.bins) < len(sp.groups)


def test_row_order_does_not_depend_on_the_bins():
    """The autotuner's bin ladders only cut the same row order into different launches (so the adaptive fits reduce
    the rows in the same order and every candidate gives the same results)."""
    import torch
    from xpg_amd import autotune
    spec = synth.replace(synth.spec_uneven(), nframes=60)
    exp = synth.make_experiment(spec)
    edges = synth.block_edges(spec)
    img = SimpleNamespace(first=1, last=60, nx=spec.nx, ny=spec.ny)
    hp = I._integrate(img, [exp] * (len(edges) - 1), edges, I.IntegrationModel(spec.sigma_d, spec.sigma_m),
                      synth.edge_resolution(exp.geom), plan_only=True, verbose=False,
                      _box_draw=synth.box_draw_for(spec))
    perms = []
    for r in (1.25, 1.5, 2.0, None):
        sp = I.SweepPlan(hp["refl"], hp["iz0"], hp["fb"], hp["bucket"], hp["bxs"], hp["pxs"], hp["ix0"], hp["iy0"],
                         hp["region"], hp["has_partner"], hp["pj"], hp["nb"], hp["blocks_refl"], hp["block_order"],
                         torch.device("cpu"), lds_bins=autotune.ladder(r))
        perms.append(sp.perm)
        assert len(sp.bins) <= len(sp.groups)
    for p in perms[1:]:
        assert np.array_equal(p, perms[0])
