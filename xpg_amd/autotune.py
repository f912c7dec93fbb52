"""Autotuner of the batched form's launch parameters: the LDS size bins that group shoeboxes of different shapes
into one launch, and the wavefronts (rows) per workgroup.

Both parameters only regroup rows into launches: the table keeps its row order (shapes in order of their LDS need)
whatever the bins, and every row is computed by the same per-row code whatever launch it is in (each reflection
stops its own fit), so every candidate gives bit-identical results -- checked for each -- and the tuner picks on
time alone.  A candidate is timed as whole batched sweeps of the caller's workload (eager launches, best of
``repeats`` after one warm-up); the winner can be written to xpg_amd/tuned.json, which sweep_plan() reads.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np
import torch

from .integrate import LDS_BINS, Sweep

LO, HI = 8192, 163840


def ladder(ratio):
    """LDS bin bounds from 8 KB up by ``ratio`` (rounded to 4 KB), closed by the 160 KB limit; ratio None: one bin."""
    if ratio is None:
        return (HI,)
    out, b = [], float(LO)
    while b < HI:
        out.append(int(round(b / 4096.0)) * 4096)
        b *= ratio
    out.append(HI)
    return tuple(sorted(set(out)))


CANDIDATES = [dict(lds_bins=LDS_BINS, wg_waves=0)] + \
             [dict(lds_bins=ladder(r), wg_waves=w) for r in (1.25, 1.5, 2.0, 4.0, None) for w in (0, 2, 1)]


def tune(frames, exp_blocks, model, hp, opts, candidates=None, repeats=2, log=None):
    """Time every candidate on this workload; returns (best candidate, table of results)."""
    candidates = candidates or CANDIDATES
    seen, table = set(), []
    ref = None
    for cand in candidates:
        key = (tuple(cand["lds_bins"]), int(cand["wg_waves"]))
        if key in seen:
            continue
        seen.add(key)
        o = dict(opts, form="batched", lds_bins=key[0], wg_waves=key[1])
        sw = Sweep(frames, exp_blocks, model, hp, o)
        out = sw.run()
        iprf = out["o2"]["iprf"]
        if ref is None:
            ref = iprf.clone()
        same = bool(torch.equal(torch.nan_to_num(iprf, nan=-1.0), torch.nan_to_num(ref, nan=-1.0)))
        best = np.inf
        for _ in range(repeats):
            torch.cuda.synchronize(sw.device)
            t0 = time.perf_counter()
            sw.run()
            best = min(best, time.perf_counter() - t0)
        rec = dict(lds_bins=list(key[0]), wg_waves=key[1], launches=len(sw.sp.bins), ms=round(best * 1e3, 3),
                   identical=same)
        table.append(rec)
        if log:
            log(f"autotune: bins {len(key[0])} wg_waves {key[1]}: {rec['ms']} ms ({rec['launches']} launches)"
                f"{'' if same else ' RESULTS DIFFER'}")
        del sw, out
        torch.cuda.empty_cache()
    ok = [r for r in table if r["identical"]]
    best = min(ok, key=lambda r: r["ms"])
    return best, table


def save(best, table, workload, path=None):
    """Record the winner for sweep_plan() (xpg_amd/tuned.json) with the table it was chosen from."""
    path = Path(path) if path else Path(__file__).with_name("tuned.json")
    rec = dict(device="gfx950", workload=workload, lds_bins=best["lds_bins"], wg_waves=best["wg_waves"],
               ms=best["ms"], table=table)
    path.write_text(json.dumps(rec, indent=1) + "\n")
    return path
