"""Load the golden kernel fixtures (tests/golden/kernels_*.npz) back into kernel-call inputs."""
from __future__ import annotations

from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
RC_KEYS = ("S", "phi", "ix0", "iy0", "iz0", "fb", "region", "has_p", "pS", "pphi", "pxs")


def cases():
    return sorted(p.stem[len("kernels_"):] for p in GOLDEN.glob("kernels_*.npz"))


def load(name):
    from chunks import analytic_template
    z = np.load(GOLDEN / f"kernels_{name}.npz")
    c = dict(Dz=int(z["Dz"]), Dy=int(z["Dy"]), Dx=int(z["Dx"]), sb=z["sb"].astype(np.float64), axis=z["axis"],
             s0=z["s0"], frames_phi0=float(z["frames_phi0"]), osc=float(z["osc"]), sd=float(z["sd"]),
             sm=float(z["sm"]), gain=float(z["gain"]), rcw=z["rcw"], prof=z["prof"])
    c["rc"] = {k: z["rc_" + k] for k in RC_KEYS if "rc_" + k in z}
    c["nb"] = {k: z["nb_" + k] for k in ("x", "y", "z", "sx", "sz")}
    det = {k: float(z["det_" + k]) for k in ("orgx", "orgy", "qx", "qy", "wl", "tsens", "mu", "plx")}
    for k in ("P0", "Bx", "By", "U1", "U2", "UN"):
        det[k] = z["det_" + k]
    det.update(segmap=None, gx=None, gy=None, nx=1024, ny=1024,
               bkg=analytic_template() if int(z["det_bkg_analytic"]) else None)
    c["det"] = det
    out = {}
    for mode in (64, 32):
        out[mode] = {k[len(f"out{mode}_"):]: z[k] for k in z.files if k.startswith(f"out{mode}_")}
    return c, out
