"""After integration: the corrections and outputs of the reference's INTEGRATE stage and SCALE's first step.

- ``lp_geom`` / ``finalise_device``: the Lorentz-polarisation factor, sensor efficiency and air absorption, and the
  columns rlp, I, sig, Isum, sigsum, peak, corr of xpg/stage_integrate.py:408-421 (with Geometry.lp_factor and
  Geometry.absorption_factors, xpg/geometry.py:372-407) on the device: xpg_finalise (include/xpg_finalise.h), one lane
  per reflection.  The batched sweep runs it behind every block's pass 2, so integrate() returns these columns too.
- ``finalise_refl``: the stage's handling of integrate()'s result -- the columns (from the sweep, or computed on the
  device here), the selection of finite reflections with sigma > 0, and the split of extra lattices
  (xpg/stage_integrate.py:408-432).
- ``write_integrated_npz`` / ``write_integrate_hkl``: xpg_integrated.npz and INTEGRATE.HKL
  (xpg/stage_integrate.py:442-447, xpg/io/legacy.py:140-176); the data records are formatted by the library's host
  code (xpg_format_integrate_rows).
- ``combine_prf_sum``: SCALE's profile / summation combination (xpg/stage_scale.py:623-695) for given symmetry-unique
  reflection ids (the tNCS class correction, which needs the tNCS analysis, is not part of it).
"""
from __future__ import annotations

import ctypes as C
import datetime

import numpy as np
import torch

from .geometry import Experiment, Geometry, mu_linear, unit

FINAL_F64 = ("rlp", "I", "sig", "Isum", "sigsum", "peak")


def lp_geom(g: Geometry):
    """The per-geometry constants of xpg_finalise (xpg_lp_geom)."""
    from . import _lib as L
    s = L.LpGeom()
    e0 = np.asarray(g.beam, np.float64)
    pn = np.asarray(g.polarization_normal, np.float64)
    n = pn - np.dot(pn, e0) * e0
    n = n / np.linalg.norm(n)
    for k, v in (("e0", e0), ("axis", np.asarray(g.axis, np.float64)), ("pn", n), ("pm", np.cross(n, e0)),
                 ("ed3", np.asarray(g.ed3, np.float64))):
        getattr(s, k)[:] = [float(t) for t in v]
    s.pfrac = float(g.polarization_fraction)
    s.distance = float(g.distance)
    s.tsens, s.mu = float(g.sensor_thickness), float(g.mu)
    s.mu_air = float(mu_linear(g.wavelength, "AIR"))
    s.segmented = int(bool(g.segmented))
    return s


def finalise_device(geom, det, n, S, x, y, cols, out, stream=None):
    """xpg_finalise on device tensors: ``cols`` the integrator's iprf, vprf, isum, vsum, p_valid, p_total, corr (and
    optionally coverage, float32), ``out`` the float64 output tensors by name (FINAL_F64, "corr") and "good" (int32);
    ``geom`` from lp_geom, ``det`` the kernels' DeviceDetector (segmented detectors) or None."""
    from . import _lib as L
    fin = L.FinalIn()
    fin.S, fin.x, fin.y = L.ptr(S), L.ptr(x), L.ptr(y)
    for k in ("iprf", "vprf", "isum", "vsum", "p_valid", "p_total", "corr", "coverage"):
        setattr(fin, k, L.ptr(cols.get(k)))
    fo = L.FinalOut()
    for k in FINAL_F64 + ("corr", "good"):
        setattr(fo, k, L.ptr(out.get(k)))
    st = C.c_void_p(stream) if stream is not None else L.stream_ptr()
    L.check(L.lib().xpg_finalise(C.byref(geom), C.byref(det.s) if det is not None else None, int(n), C.byref(fin),
                                 C.byref(fo), st))


def finalise(refl, g: Geometry, device=None):
    """The columns of xpg/stage_integrate.py:408-421 for a host result dict (integrate()'s refl), computed on the
    device: {rlp, I, sig, Isum, sigsum, peak, corr (percent), good} as host arrays."""
    from .geometry import detector_consts
    from .integrate_kernels import DeviceDetector
    device = torch.device(device or "cuda")
    n = len(refl["iprf"])
    dev = lambda k, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(refl[k]), dtype=dt, device=device)
    cols = {k: dev(k) for k in ("iprf", "vprf", "isum", "vsum", "p_valid", "p_total", "corr")}
    if "coverage" in refl:
        cols["coverage"] = dev("coverage", torch.float32)
    out = {k: torch.empty(n, dtype=torch.float64, device=device) for k in FINAL_F64 + ("corr",)}
    out["good"] = torch.empty(n, dtype=torch.int32, device=device)
    det = DeviceDetector(detector_consts(g), device) if g.segmented else None
    x = dev("x") if g.segmented else None
    y = dev("y") if g.segmented else None
    with torch.cuda.device(device):
        finalise_device(lp_geom(g), det, n, dev("S"), x, y, cols, out)
        res = {k: v.cpu().numpy() for k, v in out.items()}
    res["good"] = res["good"].astype(bool)
    return res


def finalise_refl(refl, g: Geometry, device=None):
    """The INTEGRATE stage's use of integrate()'s result (xpg/stage_integrate.py:408-432): the corrected columns
    (those the sweep computed on the device, else computed here), corr in percent, the finite reflections with
    sigma > 0, split into the first lattice's table and {k - 1: table} of the extra lattices."""
    r = dict(refl)
    if all(k in r for k in FINAL_F64):
        r["corr"] = 100 * np.nan_to_num(r["corr"])
        good = np.isfinite(r["I"]) & np.isfinite(r["sig"]) & (r["sig"] > 0)
    else:
        f = finalise(r, g, device)
        good = f.pop("good")
        r.update(f)
    n = good.shape[0]
    pick = lambda d, m: {k: (v[m] if isinstance(v, np.ndarray) and v.shape[:1] == (n,) else v) for k, v in d.items()}
    r = pick(r, good)
    n = int(good.sum())
    lat = r.get("lattice", np.zeros(n, int))
    pick2 = lambda m: {k: (v[m] if isinstance(v, np.ndarray) and v.shape[:1] == m.shape else v) for k, v in r.items()}
    extra = {k - 1: pick2(lat == k) for k in range(1, int(lat.max()) + 1 if lat.size else 1)}
    return pick2(lat == 0), extra


def write_integrated_npz(path, refl):
    """xpg_integrated.npz: every per-reflection array of the table except the per-frame profiles ``fz``
    (xpg/stage_integrate.py:446-447)."""
    np.savez_compressed(path, **{k: v for k, v in refl.items() if isinstance(v, np.ndarray) and k != "fz"})


# International Tables numbers of the 65 space groups without inversion or mirror symmetry (those of macromolecular
# crystals), by Hermann-Mauguin symbol without spaces or screw parentheses
_SG_NUMBERS = dict(
    P1=1, P2=3, P21=4, C2=5, P222=16, P2221=17, P21212=18, P212121=19, C2221=20, C222=21, F222=22, I222=23,
    I212121=24, P4=75, P41=76, P42=77, P43=78, I4=79, I41=80, P422=89, P4212=90, P4122=91, P41212=92, P4222=93,
    P42212=94, P4322=95, P43212=96, I422=97, I4122=98, P3=143, P31=144, P32=145, R3=146, H3=146, P312=149, P321=150,
    P3112=151, P3121=152, P3212=153, P3221=154, R32=155, H32=155, P6=168, P61=169, P65=170, P62=171, P64=172,
    P63=173, P622=177, P6122=178, P6522=179, P6222=180, P6422=181, P6322=182, P23=195, F23=196, I23=197, P213=198,
    I213=199, P432=207, P4232=208, F432=209, F4132=210, I432=211, P4332=212, P4132=213, I4132=214)


def space_group_number(symbol):
    """The number of a space-group symbol (1 when unknown, as the reference falls back without gemmi)."""
    if not symbol:
        return 1
    key = "".join(ch for ch in str(symbol) if ch not in " ()_")
    return _SG_NUMBERS.get(key.upper() if key[:1].islower() else key, 1)


def _cell_of(R):
    a, b, c = (np.linalg.norm(v) for v in R)
    ang = lambda u, v: np.degrees(np.arccos(np.clip(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)), -1, 1)))
    return a, b, c, ang(R[1], R[2]), ang(R[0], R[2]), ang(R[0], R[1])


def _frame_safe(z):
    """z in [n - 1, n) printed with one decimal stays in frame n (xpg/io/legacy.py:102-106)."""
    z = np.asarray(z, float)
    return np.minimum(z, np.floor(z) + 0.94)


def integrate_hkl_header(exp: Experiment, data_range, sigma_m, sigma_d, template=""):
    """Header lines of INTEGRATE.HKL (xpg/io/legacy.py:71-98, :143-162): the XDS keywords of the geometry, cell and
    widths, then the item list of the 21-column records."""
    g, cr = exp.geom, exp.crystal
    R = np.linalg.inv(np.asarray(cr.Astar, np.float64))       # rows: real-space a, b, c
    cell = _cell_of(R)
    s0 = g.s0 * g.wavelength
    head = ["!OUTPUT_FILE=INTEGRATE.HKL    DATE=" + datetime.date.today().strftime("%d-%b-%Y"),
            "!Generated by xpg_amd",
            f"!NAME_TEMPLATE_OF_DATA_FRAMES={template}",
            f"!DATA_RANGE={data_range[0]:8d}{data_range[1]:8d}",
            f"!ROTATION_AXIS={g.axis[0]:10.6f}{g.axis[1]:10.6f}{g.axis[2]:10.6f}",
            f"!OSCILLATION_RANGE={g.osc:10.6f}",
            f"!STARTING_ANGLE={g.phi0:10.3f}",
            f"!STARTING_FRAME={g.starting_frame:8d}",
            f"!SPACE_GROUP_NUMBER={space_group_number(getattr(cr, 'space_group', None)):5d}",
            f"!UNIT_CELL_CONSTANTS={cell[0]:10.3f}{cell[1]:10.3f}{cell[2]:10.3f}{cell[3]:8.3f}{cell[4]:8.3f}"
            f"{cell[5]:8.3f}"]
    head += [f"!UNIT_CELL_{ax}-AXIS={R[i, 0]:10.3f}{R[i, 1]:10.3f}{R[i, 2]:10.3f}" for i, ax in enumerate("ABC")]
    head += [f"!REFLECTING_RANGE_E.S.D.={sigma_m:10.3f}",
             f"!BEAM_DIVERGENCE_E.S.D.={sigma_d:10.3f}",
             f"!X-RAY_WAVELENGTH={g.wavelength:10.6f}",
             f"!INCIDENT_BEAM_DIRECTION={s0[0]:10.6f}{s0[1]:10.6f}{s0[2]:10.6f}",
             f"!FRACTION_OF_POLARIZATION={g.polarization_fraction:8.3f}",
             "!POLARIZATION_PLANE_NORMAL=" + "".join(f"{v:10.6f}" for v in g.polarization_normal),
             "!DETECTOR=PILATUS",
             f"!NX={g.nx:6d}  NY={g.ny:6d}    QX={g.qx:10.6f}  QY={g.qy:10.6f}",
             f"!ORGX={g.orgx:10.2f}  ORGY={g.orgy:10.2f}",
             f"!DETECTOR_DISTANCE={g.distance:10.3f}",
             "!DIRECTION_OF_DETECTOR_X-AXIS=" + "".join(f"{v:10.5f}" for v in g.ed1),
             "!DIRECTION_OF_DETECTOR_Y-AXIS=" + "".join(f"{v:10.5f}" for v in g.ed2),
             "!NUMBER_OF_ITEMS_IN_EACH_DATA_RECORD=21",
             "!H,K,L,IOBS,SIGMA,XCAL,YCAL,ZCAL,RLP,PEAK,CORR,MAXC,",
             "!             XOBS,YOBS,ZOBS,ALF0,BET0,ALF1,BET1,PSI,ISEG",
             "!Items are separated by a blank and can be read in free-format",
             "!END_OF_HEADER"]
    return head


def integrate_hkl_records(g: Geometry, refl):
    """The data records of INTEGRATE.HKL as text (xpg/io/legacy.py:163-176): the beam directions ALF / BET as
    spherical angles of the incident and diffracted beams, observed positions falling back to the calculated ones,
    frame coordinates kept in their frame; formatted by xpg_format_integrate_rows."""
    from . import _lib as L
    n = len(refl["I"])
    u0 = unit(g.s0)
    alf0, bet0 = np.degrees(np.arctan2(u0[1], u0[0])), np.degrees(np.arccos(np.clip(u0[2], -1, 1)))
    if "S" in refl:
        u1 = np.asarray(refl["S"], float)
        u1 = u1 / np.linalg.norm(u1, axis=1, keepdims=True)
        alf1, bet1 = np.degrees(np.arctan2(u1[:, 1], u1[:, 0])), np.degrees(np.arccos(np.clip(u1[:, 2], -1, 1)))
    else:
        alf1 = bet1 = np.zeros(n)
    fin = lambda k, fb: np.where(np.isfinite(refl[k]), refl[k], refl[fb])
    a = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    cols = [a(refl["I"]), a(refl["sig"]), a(refl["x"]), a(refl["y"]), a(_frame_safe(refl["z"])), a(refl["rlp"]),
            a(refl["peak"]), a(refl["corr"]), a(fin("xo", "x")), a(fin("yo", "y")), a(_frame_safe(fin("zo", "z")))]
    hkl = np.ascontiguousarray(refl["hkl"], dtype=np.int32).reshape(n, 3)
    a1, b1 = a(alf1), a(bet1)
    cap = max(200 * n, 4096)
    while True:
        buf = C.create_string_buffer(cap)
        used = L.lib().xpg_format_integrate_rows(buf, cap, n, hkl.ctypes.data, *[c.ctypes.data for c in cols],
                                                 float(alf0), float(bet0), a1.ctypes.data, b1.ctypes.data)
        if used >= 0:
            return buf.raw[:used].decode()
        cap *= 4


def write_integrate_hkl(path, exp: Experiment, refl, data_range, sigma_m, sigma_d, template=""):
    """INTEGRATE.HKL (xpg/io/legacy.py:140-176): LP-corrected profile-fitted intensities of the stage's table."""
    head = integrate_hkl_header(exp, data_range, sigma_m, sigma_d, template)
    with open(path, "w") as f:
        f.write("\n".join(head) + "\n")
        f.write(integrate_hkl_records(exp.geom, refl))
        f.write("!END_OF_DATA\n")


# ---- SCALE: profile-fitted / summation combination -----------------------------------------------------------------

def _group_medians(v, inv, ng):
    """Median of v within each of ng groups (inv: group of each element; every group non-empty), vectorised: one sort
    by (group, value), then the middle element or the mean of the two middle ones."""
    order = np.lexsort((v, inv))
    sv = v[order]
    cnt = np.bincount(inv, minlength=ng)
    start = np.concatenate(([0], np.cumsum(cnt)[:-1]))
    lo = start + (cnt - 1) // 2
    hi = start + cnt // 2
    return 0.5 * (sv[lo] + sv[hi])


def _rmeas(I, inv, n):
    m = np.bincount(inv, I) / n
    mm = n[inv] >= 2
    dev = np.abs(I - m[inv]) * np.sqrt(n[inv] / np.maximum(n[inv] - 1, 1))
    return float(dev[mm].sum() / max(np.abs(I[mm]).sum(), 1e-30))


def combine_prf_sum(r, uid, nx, ny, block_frames=100, log=None):
    """SCALE's combination of profile-fitted and summation intensities (xpg/stage_scale.py:623-695):
    I = w I_prf + (1 - w) I_sum / f, w = 1 / (1 + max(I_sum / f, 0) / I_mid), with f the recorded fraction inside the
    summation foreground (PEAK, times the median profile fraction p_fg of reflections of the same detector third,
    |zeta| quintile and 100-frame block when p_fg is known for 90% of them), and I_mid the candidate (inf, then the
    50 / 75 / 90 / 97 / 99% quantiles of |I_sum / f|) with the lowest Rmeas of the symmetry mates ``uid`` within
    blocks of ``block_frames`` frames; inf is dropped when strong, fully recorded reflections disagree with the profile
    model by more than 10%.  Returns (I, sig, info)."""
    rec = np.clip(r["peak"] / 100.0, 0.5, 1.0)
    frac = rec
    if "p_fg" in r and np.isfinite(r["p_fg"]).mean() >= 0.9:
        v = np.clip(np.nan_to_num(r["p_fg"], nan=1.0), 0.5, 1.0)
        reg = np.clip((r["x"] / nx * 3).astype(int), 0, 2) * 3 + np.clip((r["y"] / ny * 3).astype(int), 0, 2)
        zk = np.abs(r.get("zeta_k", np.ones(len(v))))
        zb = np.clip(np.searchsorted(np.quantile(zk, [0.2, 0.4, 0.6, 0.8]), zk), 0, 4)
        key_ = (reg * 5 + zb) * 100000 + (r["z"] // 100).astype(int)
        uk_, inv_ = np.unique(key_, return_inverse=True)
        frac = np.clip(rec * _group_medians(v, inv_.ravel(), uk_.size)[inv_], 0.5, 1.0)
    Ip, sp, Is_, ss_ = r["I"], r["sig"], r["Isum"] / frac, r["sigsum"] / frac
    okc = np.isfinite(Is_) & (ss_ > 0)
    if "sum_ok" in r:
        okc &= np.asarray(r["sum_ok"], bool)
    key = np.asarray(uid, np.int64) * 10000 + (r["z"] // block_frames).astype(np.int64)
    _, inv = np.unique(key, return_inverse=True)
    inv = inv.ravel()
    n = np.bincount(inv)
    cands = [np.inf] + list(np.quantile(np.abs(Is_[okc]), [0.5, 0.75, 0.9, 0.97, 0.99]))
    strong = okc & (Is_ / np.maximum(ss_, 1e-9) > 20) & (r["peak"] > 95) & (Ip > 0) & (r["n_over"] == 0)
    ratio = None
    if strong.sum() > 100:
        ratio = float(np.exp(np.median(np.log(np.maximum(Is_[strong], 1e-9) / Ip[strong]))))
        if abs(ratio - 1.0) > 0.1:
            cands = cands[1:]
    best = None
    Is0, ss0 = np.where(okc, Is_, 0.0), np.where(okc, ss_, 0.0)
    for imid in cands:
        w = np.where(okc, 1.0 / (1.0 + np.maximum(Is_, 0) / imid), 1.0)
        rm = _rmeas(w * Ip + (1 - w) * Is0, inv, n)
        if best is None or rm < best[0] - 1e-5:
            best = (rm, imid, w)
    rm, imid, w = best
    I = w * Ip + (1 - w) * Is0
    sig = np.sqrt((w * sp) ** 2 + ((1 - w) * ss0) ** 2)
    info = dict(i_mid=float(imid), rmeas_prf=_rmeas(Ip, inv, n), rmeas=float(rm), strong_ratio=ratio)
    if log is not None:
        log(f"  SCALE: profile-fitted / summation combination: I_mid = {imid:.0f}; block Rmeas "
            f"{100 * info['rmeas_prf']:.2f}% (profile) -> {100 * rm:.2f}%")
    return I, sig, info
