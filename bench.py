#!/usr/bin/env python
"""Benchmark: profile-fitted reflections/s of the MI355X integration path on one GPU.

A step is one pass of the integration hot path over the whole sweep with the frames resident in HBM: for every phi
block, the shoebox gather, pass 1 with fused profile accumulation, the adaptive-width fits, the float64 profile
normalisation and pass 2 (Poisson profile fit), then every per-reflection result copied to pinned host memory.  Host
planning (prediction, box plan, neighbour lists) and the upload of the reflection table happen once before the timed
region, like the frames; their time is reported beside the value (``plan``).

Prints one JSON line: metric, value = reflections/s, the roofline of the kernel with the most device time (HIP events
on its launch stream, algorithmic bytes per launch), the CPU baseline (the reference CPU path restated: host shoebox
gather plus the oracle's passes in float32 on every host core, over a bounded sample) and the parity of that sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, 8 TB/s
FORMS = dict(
    plain="plain: one stream; per phi block one pass-1 launch per shoebox group (voxels read from the frame store), "
          "the adaptive fits, normalisation, one pass-2 launch per group",
    batched="batched: widths-independent work (shoebox gather, Kabsch frame, eps, J, neighbour masks) ahead on a "
            "second stream, the serial pass-1/fits chain on one stream, pass 2 overlapped on a third; timed as eager "
            "launches and as replays of the captured HIP graph (batched_forms)")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def per_refl_bytes(V, Dz, kind):
    """Algorithmic HBM bytes of one reflection in one launch.  Shoebox gather (prepare), pass 1 and pass 2: 4 V + 80,
    the shoebox's V voxels read once as 4-byte values plus 80 B of parameters in and results out (each pass is half of
    the step's per-reflection figure 8 V + 160 below; the gather reads what the fused form's pass 1 reads).  Rocking
    curves: 8 Dz + 48, the curve written and the row's parameters read."""
    if kind == "rocking":
        return 8 * Dz + 48
    return 4 * V + 80


def step_bytes(hp, nframes, ny, nx, device):
    """Algorithmic bytes of one whole step.  B_alg = 4 nx ny N_frames + sum over reflections (8 V + 160): every frame
    pixel read once as uint32 plus two passes over each shoebox.  This design reads only shoebox voxels, so its bound
    is B_min = 4 U + sum (4 V + 160), U = the exact union of shoebox voxels (frame x pixel) on the detector and in
    the frame range, counted on the device with a bitmap per window of frames."""
    import numpy as np
    import torch
    dz, fb, pxs, bx = (np.asarray(hp[k], np.int64) for k in ("dz", "fb", "pxs", "bxs"))
    nz, side = dz * fb, bx * pxs                        # frames and detector pixels a box spans
    z0 = np.asarray(hp["iz0"], np.int64)
    y0 = np.asarray(hp["iy0"], np.int64) * pxs
    x0 = np.asarray(hp["ix0"], np.int64) * pxs
    V = dz * bx * bx
    W = 100
    U = 0
    for w0 in range(0, nframes, W):
        w1 = min(w0 + W, nframes)
        bm = torch.zeros((w1 - w0) * ny * nx, dtype=torch.bool, device=device)
        hit = (z0 < w1) & (z0 + nz > w0)
        for a, b_ in sorted({(int(p), int(q)) for p, q in zip(nz[hit], side[hit])}):
            sel = np.nonzero(hit & (nz == a) & (side == b_))[0]
            for c0 in range(0, sel.size, 20000):
                rows = sel[c0:c0 + 20000]
                T = lambda v: torch.as_tensor(v[rows], device=device)
                z = T(z0)[:, None, None, None] + torch.arange(a, device=device)[None, :, None, None]
                y = T(y0)[:, None, None, None] + torch.arange(b_, device=device)[None, None, :, None]
                x = T(x0)[:, None, None, None] + torch.arange(b_, device=device)[None, None, None, :]
                ok = (z >= w0) & (z < w1) & (y >= 0) & (y < ny) & (x >= 0) & (x < nx)
                idx = ((z - w0) * ny + y) * nx + x
                bm[idx[ok]] = True
        U += int(bm.sum().item())
        del bm
    n = V.size
    b_alg = 4 * nx * ny * nframes + int(np.sum(8 * V)) + 160 * n
    b_min = 4 * U + int(np.sum(4 * V)) + 160 * n
    return dict(B_alg=b_alg, B_min=b_min, union_voxels=U, sum_V=int(V.sum()), reflections=n)


def host_cores():
    """Host cores this process may use (affinity), at most the 16 of one GPU's share of the machine."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


CHUNK_ELEMENTS = 15_000_000       # voxels per kernel call of the reference (xpg/integrate.py:56)


def cpu_baseline(sweep, hp, spec, frames, n_sample, seconds, threads):
    """The reference's CPU path restated and timed on this host: the host shoebox gather (frame codes, padded frames,
    shoebox windows; xpg/integrate.py:648-803, on min(8, cores) gather threads as :671) and the oracle's pass 1,
    profile accumulation, normalisation and pass 2 in float32, the precision of the reference's tensor program, over
    chunks of at most 1.5e7 voxels (xpg/integrate.py:56), each chunk's rows split over ``threads`` threads (the
    reference's torch intra-op threads, xpg/backend.py:60-69).  The sample is one block's largest shoebox group (all
    of its frames, so that the per-frame host work is amortised as in a whole run).

    Also returns the parity record of the same rows: the float64 oracle against the HIP passes (same shoeboxes, default
    widths, the oracle's profiles), the float32 restatement against the float64 oracle (the reference precision's own
    drift), and whether the host gather equals the device extraction."""
    import numpy as np
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from oracle import extract as ox
    from oracle import kernels as ok
    from xpg_amd import integrate_kernels as K
    from xpg_amd.geometry import detector_consts

    sp = sweep.sp
    b_mid = hp["block_order"][len(hp["block_order"]) // 2]
    gi, (B, bx, s), start, m = max(sweep.by_block[b_mid], key=lambda t: t[3])
    if n_sample:
        m = min(m, n_sample)
    view = sp.refl_view(start, m, (B, bx, bx))
    rows = sp.perm[start:start + m]
    refl = hp["refl"]
    exp = sweep.exp_blocks[b_mid]
    g = exp.geom
    det = detector_consts(g)
    det.update(nx=g.nx, ny=g.ny)
    ix0, iy0, iz0, fb = hp["ix0"][rows], hp["iy0"][rows], hp["iz0"][rows], hp["fb"][rows]
    z0 = int(max(0, iz0.min()))                  # boxes may start before the first frame or end after the last
    z1 = int(min(frames.n, (iz0 + B * fb).max()))
    # the block's frames in host memory as the reference reads them (raw uint32 counts): outside the timed region
    raw = {z: frames.data[z].cpu().numpy().view(np.uint32) for z in range(z0, z1)}
    rc = dict(S=refl["S"][rows], phi=np.radians(refl["phi"][rows]), ix0=ix0, iy0=iy0, iz0=iz0,
              fb=fb.astype(float), region=hp["region"][rows], has_p=hp["has_partner"][rows].astype(float),
              pS=refl["S"][hp["pj"][rows]], pphi=np.radians(refl["phi"][hp["pj"][rows]]))
    if s > 1:
        rc["pxs"] = np.full(m, float(s))
    nb = {k: v[rows] for k, v in hp["nb"].items()}
    sd, sm = np.radians(sweep.model.sigma_d), np.radians(sweep.model.sigma_m)
    frames_phi0, osc = np.radians(g.z_to_phi(0.0)), np.radians(g.osc)
    rcw = ok.rc_default_weights()
    a = (det, g.axis, g.s0, frames_phi0, osc, sd, sm, 1.0)
    pool = ThreadPoolExecutor(threads)
    gpool = ThreadPoolExecutor(min(8, threads))
    # row ranges: chunks of the reference's size, each split over the threads
    per_chunk = max(1, CHUNK_ELEMENTS // (B * bx * bx))
    parts = []
    for c0 in range(0, m, per_chunk):
        c1 = min(m, c0 + per_chunk)
        step = -(-(c1 - c0) // threads)
        parts.append([(r0, min(c1, r0 + step)) for r0 in range(c0, c1, step)])
    sub = lambda d, r0, r1: {k: (v[r0:r1] if np.ndim(v) else v) for k, v in d.items()}
    pad = bx // 2 + 1

    def host_gather():
        def one(z):
            P = ox.prepared_frame(ox.frame_codes(raw[z], flag_value=frames.flag_value, overload=frames.overload), pad)
            if s > 1:
                P = ox.superpixel_frame(P[pad:pad + g.ny, pad:pad + g.nx], s, pad)
            return z, P
        padded = dict(gpool.map(one, range(z0, z1)))
        sb = np.zeros((m, B, bx, bx), np.float32)
        ny_p, nx_p = next(iter(padded.values())).shape

        def rows_of(r0, r1):
            for r in range(r0, r1):
                iy = int(np.clip(iy0[r] + pad, 0, ny_p - bx))
                ix = int(np.clip(ix0[r] + pad, 0, nx_p - bx))
                for k in range(B):
                    z = int(iz0[r]) + k
                    if fb[r] == 1 and z in padded:
                        sb[r, k] = padded[z][iy:iy + bx, ix:ix + bx] + ox.SB_OFF
            return None
        list(gpool.map(lambda t: rows_of(*t), [t for ch in parts for t in ch]))
        if (fb > 1).any():                        # binned rows: the oracle's own gather
            sel = np.flatnonzero(fb > 1)
            sb[sel] = ox.gather({z: ox.frame_codes(raw[z], flag_value=frames.flag_value, overload=frames.overload)
                                 for z in range(z0, z1)}, range(z0, z1), ix0[sel], iy0[sel], iz0[sel], fb[sel], s, B,
                                bx) + ox.SB_OFF
        return sb - ox.SB_OFF

    def one_run(ft):
        ok.FT = ft
        try:
            sb = host_gather()
            o1s = [None] * m
            res1 = {}
            for ch in parts:
                for (r0, r1), o1 in zip(ch, pool.map(lambda t: ok.pass1(sb[t[0]:t[1]], sub(rc, *t), *a,
                                                                         sub(nb, *t), rcw), ch)):
                    res1[(r0, r1)] = o1
            cat = lambda k: np.concatenate([res1[t][k] for ch in parts for t in ch])
            o1 = {k: cat(k) for k in ("isum", "vsum", "n_over", "frac_rec", "n_fg_nb", "n_fg_masked", "bg", "nbg")}
            strong = ((o1["isum"] > 0) & (o1["isum"] / np.sqrt(np.maximum(o1["vsum"], 1e-6)) > 10)
                      & (o1["n_over"] == 0) & (o1["frac_rec"] > 0.95) & (o1["n_fg_nb"] == 0)
                      & (o1["n_fg_masked"] == 0))

            def acc(t):
                r = res1[t]
                return ok.accumulate_profiles(r["img2"], r["all_valid2"], r["J"], r["u1"], r["u2"],
                                              rc["region"][t[0]:t[1]], r["isum"], strong[t[0]:t[1]], r["seff"], 9)
            grids = list(pool.map(acc, [t for ch in parts for t in ch]))
            prof = ok.normalise_profiles(sum(n_ for n_, _ in grids), sum(d_ for _, d_ in grids))
            res2 = {}
            for ch in parts:
                for t, o2 in zip(ch, pool.map(lambda t: ok.pass2(sb[t[0]:t[1]], sub(rc, *t), *a, prof,
                                                                  o1["bg"][t[0]:t[1]], o1["nbg"][t[0]:t[1]],
                                                                  sub(nb, *t), rcw, irls_rule="reflection"), ch)):
                    res2[t] = o2
            o2 = {k: np.concatenate([res2[t][k] for ch in parts for t in ch]) for k in ("iprf", "vprf")}
            return sb, o1, prof, o2
        finally:
            ok.FT = np.float64

    # the float64 oracle once (parity), then the float32 restatement timed
    sb64, o1_64, prof64, o2_64 = one_run(np.float64)
    t0 = time.perf_counter()
    reps = 0
    while True:
        sb32, o1_32, prof32, o2_32 = one_run(np.float32)
        reps += 1
        el = time.perf_counter() - t0
        if el > seconds or reps >= 20:
            break
    pool.shutdown()
    gpool.shutdown()
    rate = reps * m / el
    # the same rows on the device with the same inputs (default widths, the oracle's profiles)
    dev = sweep.device
    sb_dev = K.run_extract(sweep.store, view, dev).cpu().numpy()
    detd = K.device_detector(det, dev)
    call = K.make_call(g.axis, g.s0, frames_phi0, osc, sd, sm, 1.0, rcw)
    r1 = K.run_pass1(detd, call, view, frames=sweep.store, pixel_outputs=False)
    r2 = K.run_pass2(detd, call, view, torch.as_tensor(prof64, device=dev), r1["bg"], r1["nbg"], frames=sweep.store)
    torch.cuda.synchronize(dev)
    hI = r2["iprf"].cpu().numpy()
    hV = r2["vprf"].cpu().numpy()
    sig = np.sqrt(np.maximum(o2_64["vprf"], 1e-30))
    ok_ = np.isfinite(hI) & np.isfinite(o2_64["iprf"]) & (o2_64["vprf"] > 0)
    dI = np.abs(hI - o2_64["iprf"]) / sig
    d32 = np.abs(o2_32["iprf"] - o2_64["iprf"]) / sig
    ok32 = ok_ & np.isfinite(o2_32["iprf"])
    deciles = []
    if ok_.sum() >= 10:
        I0 = o2_64["iprf"][ok_]
        zs = o2_64["iprf"][ok_] / sig[ok_] - hI[ok_] / np.sqrt(np.maximum(hV[ok_], 1e-30))
        edges = np.quantile(I0, np.linspace(0, 1, 11))
        idx = np.clip(np.searchsorted(edges, I0, side="right") - 1, 0, 9)
        for d in range(10):
            sel = idx == d
            if sel.any():
                deciles.append(dict(I_lo=float(edges[d]), I_hi=float(edges[d + 1]), n=int(sel.sum()),
                                    max_dI_over_sigma=float(dI[ok_][sel].max()),
                                    median_dI_over_sigma=float(np.median(dI[ok_][sel])),
                                    max_d_I_over_sigma=float(np.abs(zs[sel]).max())))
    parity = dict(sample=int(m), max_dI_over_sigma=float(np.max(dI[ok_])) if ok_.any() else None,
                  median_dI_over_sigma=float(np.median(dI[ok_])) if ok_.any() else None,
                  host_gather_equal_device=bool(np.array_equal(sb64, sb_dev)),
                  n_over_equal=bool(np.array_equal(r1["n_over"].cpu().numpy(), o1_64["n_over"])),
                  n_fg_nb_equal=bool(np.array_equal(r1["n_fg_nb"].cpu().numpy(), o1_64["n_fg_nb"])),
                  n_fg_masked_equal=bool(np.array_equal(r1["n_fg_masked"].cpu().numpy(), o1_64["n_fg_masked"])),
                  cpu_float32_vs_oracle_max_dI_over_sigma=float(np.max(d32[ok32])) if ok32.any() else None,
                  by_intensity_decile=deciles)
    base = dict(value=round(rate, 2), unit="reflections/s", cores=threads, kind="port",
                sample=f"block {b_mid}: {m} reflections ({B}x{bx}x{bx} shoeboxes) x {reps} repetitions; host gather "
                       f"from {z1 - z0} raw frames in host memory ({min(8, threads)} threads) + float32 oracle pass 1, "
                       f"profile accumulation, normalisation, pass 2 in {len(parts)} chunk(s) of <= 1.5e7 voxels over "
                       f"{threads} threads; {el:.1f} s")
    return base, parity


def summarize_stats(stats_csv, command, out_path):
    """Markdown summary of a `rocprofv3 --kernel-trace --stats --output-format csv` run (its *_kernel_stats.csv)."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    lines = [f"# Kernel time: `{command}`", "", "| kernel | calls | total ms | average us | min us | max us | % |",
             "|---|---|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"].replace("(anonymous namespace)::", "").split("(")[0]
        lines.append(f"| {name} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                     f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
    Path(out_path).write_text("\n".join(lines) + "\n")


def summarize_pmc(fetch_csv, write_csv, copy_mib, kernel, workload, command):
    """HBM bytes per dispatch of `kernel` from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch), each
    scaled by what the same counter reads for the copy-like calibration kernel of known size (bench.py
    --calibrate-copy MIB: MIB MiB read, MIB MiB written).  Writes profiles/traffic_<kernel>_<workload>.json and .md."""
    import csv
    import statistics

    def per_kernel(path, counter):
        vals = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            vals.setdefault(r["Kernel_Name"], {}).setdefault(r["Dispatch_Id"], 0.0)
            vals[r["Kernel_Name"]][r["Dispatch_Id"]] += float(r["Counter_Value"])
        return {k: list(v.values()) for k, v in vals.items()}

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    import re
    calib = next(k for k in fetch if "neg" in k.lower())
    nbytes = copy_mib * 2 ** 20
    f_fac = statistics.mean(fetch[calib]) * 1024 / nbytes
    w_fac = statistics.mean(write[calib]) * 1024 / nbytes
    pat = re.compile(r"\b" + re.escape(kernel) + (r"\b" if re.match(r"\w", kernel[-1]) else ""))
    fk = [v for k, vs in fetch.items() if pat.search(k) for v in vs]
    wk = [v for k, vs in write.items() if pat.search(k) for v in vs]
    f_b = statistics.mean(fk) * 1024 / f_fac
    w_b = statistics.mean(wk) * 1024 / w_fac
    rec = dict(kernel=kernel, workload=workload, bytes_per_launch=round(f_b + w_b), fetch_bytes=round(f_b),
               write_bytes=round(w_b), dispatches=len(fk), copy_calibration=dict(
                   mib=copy_mib, fetch_factor=round(f_fac, 4), write_factor=round(w_fac, 4)), command=command)
    label = re.sub(r"\W+", "_", kernel[2:] if kernel.startswith("k_") else kernel).strip("_")
    base = ROOT / "profiles" / f"traffic_{label}_{workload}"
    base.with_suffix(".json").write_text(json.dumps(rec, indent=1) + "\n")
    base.with_suffix(".md").write_text(
        f"# HBM traffic of {kernel} ({workload})\n\nCommand (each counter in its own run): `{command}`\n\n"
        f"- FETCH_SIZE per dispatch: {f_b / 1e6:.2f} MB (raw / copy factor {f_fac:.4f})\n"
        f"- WRITE_SIZE per dispatch: {w_b / 1e6:.2f} MB (raw / copy factor {w_fac:.4f})\n"
        f"- dispatches: {len(fk)}\n- calibration: one copy-like kernel over {copy_mib} MiB read and "
        f"written (`--calibrate-copy {copy_mib}`), counters read {f_fac:.4f} x and {w_fac:.4f} x the known bytes\n")
    return rec


def summarize_sq(csv_path, workload, command, kernels):
    """Per-wavefront instruction counts of the sweep's kernels from one rocprofv3 --pmc run of SQ counters
    (SQ_WAVES, SQ_INSTS_VALU, _SALU, _LDS, _VMEM_RD, _VMEM_WR, SQ_WAVE_CYCLES, SQ_BUSY_CYCLES): writes
    profiles/counters_<workload>.md and .json."""
    import csv
    import re
    tot, disp, meta = {}, {}, {}
    for r in csv.DictReader(open(csv_path)):
        name = r["Kernel_Name"]
        k = next((k for k in kernels if re.search(r"\b" + re.escape(k) + r"\b", name)), None)
        if k is None:
            continue
        tot.setdefault(k, {}).setdefault(r["Counter_Name"], 0.0)
        tot[k][r["Counter_Name"]] += float(r["Counter_Value"])
        disp.setdefault(k, set()).add(r["Dispatch_Id"])
        meta[k] = dict(vgpr=int(r["VGPR_Count"]), sgpr=int(r["SGPR_Count"]), workgroup=int(r["Workgroup_Size"]))
    rows = []
    for k in kernels:
        if k not in tot:
            continue
        t, n = tot[k], len(disp[k])
        w = max(t.get("SQ_WAVES", 0.0), 1.0)
        rows.append(dict(kernel=k, dispatches=n, waves_per_dispatch=round(w / n),
                         valu_per_wave=round(t.get("SQ_INSTS_VALU", 0) / w), salu_per_wave=round(t.get("SQ_INSTS_SALU", 0) / w),
                         lds_per_wave=round(t.get("SQ_INSTS_LDS", 0) / w), vmem_rd_per_wave=round(t.get("SQ_INSTS_VMEM_RD", 0) / w),
                         vmem_wr_per_wave=round(t.get("SQ_INSTS_VMEM_WR", 0) / w),
                         wave_cycles_per_wave=round(t.get("SQ_WAVE_CYCLES", 0) / w), **meta[k]))
    base = ROOT / "profiles" / f"counters_{workload}"
    base.with_suffix(".json").write_text(json.dumps(dict(workload=workload, command=command, kernels=rows), indent=1) + "\n")
    hdr = "| kernel | dispatches | waves / dispatch | VALU / wave | SALU / wave | LDS / wave | VMEM rd / wave | VMEM wr / wave | SQ_WAVE_CYCLES / wave | VGPR | SGPR | workgroup |"
    lines = [f"# Instruction counters per wavefront ({workload})", "", f"Command: `{command}`", "",
             "One wavefront per reflection in every kernel but k_rocking (one lane per reflection and frame).  "
             "SQ_WAVE_CYCLES per wavefront is the counter as reported, summed over the dispatch's wavefronts and divided by their number.", "", hdr, "|" + "---|" * 12]
    for r in rows:
        lines.append(f"| {r['kernel']} | {r['dispatches']} | {r['waves_per_dispatch']} | {r['valu_per_wave']} | "
                     f"{r['salu_per_wave']} | {r['lds_per_wave']} | {r['vmem_rd_per_wave']} | {r['vmem_wr_per_wave']} | "
                     f"{r['wave_cycles_per_wave']} | {r['vgpr']} | {r['sgpr']} | {r['workgroup']} |")
    base.with_suffix(".md").write_text("\n".join(lines) + "\n")
    return rows


def bench_serve(args, dev):
    """Serving workload: datasets arriving as a Poisson process at args.serve_load of the measured capacity, frames
    generated on the device per 10-frame batch, continuous batching under a 200 GB HBM budget; p50 / p99 per-batch
    latency and aggregate reflections/s."""
    import concurrent.futures as cf
    import multiprocessing as mp

    import numpy as np
    import torch
    from xpg_amd import serve

    torch.cuda.set_device(dev)
    reqs = serve.serving_requests(n=args.serve_datasets, frame_scale=args.serve_frame_scale)
    t = time.time()
    # host plans (predictions, boxes, neighbours, block order, the generator's spot table) are made when a request
    # is submitted, before its frames arrive: worker processes, outside the served window
    with cf.ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1), mp_context=mp.get_context("spawn")) as ex:
        plans = list(ex.map(serve.plan_request, reqs))
    n_refl = sum(p.hp["refl"]["x"].size for p in plans)
    frames = sum(p.req.spec.nframes for p in plans)
    fbytes = sum(4 * p.req.spec.nx * p.req.spec.ny * p.req.spec.nframes for p in plans)
    log(f"planned {len(plans)} datasets, {frames} frames ({fbytes / 1e9:.1f} GB), {n_refl} reflections in "
        f"{time.time() - t:.1f}s")
    server = serve.Server(dev, hbm_budget=200e9, slots=args.serve_slots)
    small = min(plans, key=lambda p: p.req.spec.nframes * p.req.spec.nx * p.req.spec.ny)
    server.run([small])                                          # warm-up
    for p in plans:
        p.req.arrival = 0.0
    cap = server.run(plans)                                      # capacity: closed loop, every request at t = 0
    log(f"capacity run: {cap['wall_s']:.2f}s, {cap['reflections_per_s']:.0f} reflections/s, p50 {cap['p50_ms']} ms "
        f"p99 {cap['p99_ms']} ms; host {cap['host_s']}")
    rate = args.serve_load * len(plans) / cap["wall_s"]
    serve.poisson_arrivals([p.req for p in plans], rate)
    rep = server.run(plans)
    log(f"served: {rep['wall_s']:.2f}s, {rep['reflections_per_s']:.0f} reflections/s, p50 {rep['p50_ms']} ms "
        f"p99 {rep['p99_ms']} ms; host {rep['host_s']}")
    b_alg = 0
    for p in plans:
        V = p.hp["dz"] * p.hp["bxs"] ** 2
        b_alg += 4 * p.req.spec.nx * p.req.spec.ny * p.req.spec.nframes + int(np.sum(8 * V)) + 160 * V.size
    per = rep.pop("per_dataset")
    out = dict(metric="profile-fitted reflections/s and p50/p99 per-batch latency, serving on 1 MI355X",
               value=rep["reflections_per_s"], unit="reflections/s", n_gpus=1, steps=1, warmup=1,
               ms_per_step=round(rep["wall_s"] * 1e3, 3), higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f32", data="synthetic (frames generated on the device per batch)",
               config=dict(workload=f"serve: {len(plans)} datasets, frame counts U{{50..3600}} x "
                                    f"{args.serve_frame_scale}, 10-frame batches, 200 GB HBM budget, "
                                    f"{args.serve_slots} slots",
                           frames=frames, frame_gb=round(fbytes / 1e9, 2), reflections=int(n_refl),
                           parallelism="single GPU"),
               host_s=rep["host_s"],
               latency=dict(p50_ms=rep["p50_ms"], p99_ms=rep["p99_ms"], max_ms=rep["max_ms"],
                            batches=rep["batches_with_results"],
                            definition="batch frames resident on the device -> every reflection the batch completes "
                                       "integrated (HIP events on the dataset's stream)"),
               capacity=dict(wall_s=cap["wall_s"], reflections_per_s=cap["reflections_per_s"],
                             datasets_per_s=round(len(plans) / cap["wall_s"], 3), p50_ms=cap["p50_ms"],
                             p99_ms=cap["p99_ms"]),
               offered=dict(load=args.serve_load, arrivals_per_s=round(rate, 3)),
               roofline=dict(bound="hbm", B_alg_GB=round(b_alg / 1e9, 3), peak=PEAK_HBM_GBS, unit="GB/s",
                             achieved=round(b_alg / rep["wall_s"] / 1e9, 1),
                             frac=round(b_alg / rep["wall_s"] / (PEAK_HBM_GBS * 1e9), 4)),
               per_dataset=per)
    print(json.dumps(out), flush=True)


def bench_decode(args, dev):
    """Bitshuffle-LZ4 decode of a sweep's frames: frames of the full-size detector (4148 x 4362 uint32, the synthetic
    C2 sweep's first frames) written as bitshuffle-LZ4 chunks by the system's LZ4 library, the chunks and their block
    tables resident in HBM; a step decodes every chunk into the frame store (raw counts).  Roofline: compressed bytes
    read + 4 bytes per pixel written, per launch of k_bslz4_decode, over its HIP-event time on the launch stream."""
    import ctypes as C

    import numpy as np
    import torch
    from xpg_amd import _lib as L, bslz4, synth

    torch.cuda.set_device(dev)
    nfr = args.frames or 200
    spec = synth.replace(synth.spec_full(), nframes=nfr, name=f"full[:{nfr}]")
    t = time.time()
    frames, _ = synth.generate(spec, dev, log=log)
    host = frames.data.cpu().numpy().view(np.uint32)
    log(f"generated {nfr} frames {spec.ny}x{spec.nx} in {time.time() - t:.1f}s")
    t = time.time()
    chunks = bslz4.encode_frames(list(host), 0, threads=min(16, os.cpu_count() or 1))
    t_enc = time.time() - t
    t = time.time()
    cb = bslz4.ChunkBatch(chunks, spec.ny * spec.nx, 4, dev, index="host")
    torch.cuda.synchronize(dev)
    t_index = time.time() - t
    comp = sum(len(c) for c in chunks)
    raw_bytes = host.nbytes
    log(f"encoded {raw_bytes / 1e9:.2f} GB -> {comp / 1e9:.2f} GB (ratio {raw_bytes / comp:.2f}) in {t_enc:.1f}s; "
        f"host block index + upload {t_index:.2f}s")
    out = torch.zeros_like(frames.data)
    st = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        cb.decode_into(out, "raw")
    cb.check()
    assert torch.equal(out, frames.data), "decoded frames differ from the originals"
    # device walk of the block chains (the form for chunks that arrive without a table)
    db = bslz4.ChunkBatch(chunks, spec.ny * spec.nx, 4, dev, index="device")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(3):
        db.status.zero_()
        L.check(L.lib().xpg_bslz4_index(C.byref(db._batch()), L.stream_ptr(dev)))
    e1.record(st)
    torch.cuda.synchronize(dev)
    index_ms = e0.elapsed_time(e1) / 3
    assert torch.equal(db.table, cb.table)
    del db
    out.zero_()
    torch.cuda.synchronize(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.time()
    for a, b in ev:
        a.record(st)
        cb.decode_into(out, "raw")
        b.record(st)
    torch.cuda.synchronize(dev)
    wall = (time.time() - t0) / args.steps
    ms = [a.elapsed_time(b) for a, b in ev]
    cb.check()
    assert torch.equal(out, frames.data), "decoded frames differ from the originals"
    avg = sum(ms) / len(ms)
    alg = comp + raw_bytes
    achieved = alg / (avg * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic_bslz4_decode_decode.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    roof = dict(bound="hbm", kernel="k_bslz4_decode", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic, avg_launch_ms=round(avg, 4),
                alg_bytes_per_launch=alg, launches_per_step=1, compressed_bytes=comp, decoded_bytes=raw_bytes,
                device_index_ms=round(index_ms, 3))
    cpu = None
    if not args.no_cpu:
        # the oracle (C, one core): LZ4 stage + bit unshuffle of whole frames, a bounded sample
        from oracle import codec as oc
        t = time.time()
        n = 0
        while n < len(chunks) and (time.time() - t < args.cpu_seconds or n < 2):
            r, _ = oc.decode_chunk(chunks[n], spec.ny * spec.nx, 4)
            assert r > 0
            n += 1
        dt = time.time() - t
        cpu = dict(value=round(n / dt, 3), unit="frames/s", cores=1, kind="port",
                   sample=f"{n} frames of {spec.ny}x{spec.nx} uint32 (the first of the same chunks), oracle/codec.c "
                          f"LZ4 stage + bit unshuffle, {dt:.1f} s")
        log(f"cpu baseline {cpu}")
    value = nfr / (avg * 1e-3)
    outj = dict(metric="bitshuffle-LZ4 frames decoded/s on 1 MI355X", value=round(value, 1), unit="frames/s",
                n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(avg, 3), higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="u8", data="synthetic",
                config=dict(workload=f"bslz4_decode_{spec.nx}x{spec.ny}x{nfr}", frames=nfr,
                            frame_shape=[spec.ny, spec.nx], element="uint32", block_bytes=8192,
                            compression_ratio=round(raw_bytes / comp, 3), encoder="liblz4 LZ4_compress_default",
                            decoded_GBps=round(raw_bytes / (avg * 1e-3) / 1e9, 1), wall_ms_per_step=round(wall * 1e3, 3),
                            identical_to_source=True, parallelism="single GPU"),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(outj), flush=True)


def bench_forward(args, dev):
    """Forward (window) model passes (xpg/forward/kernels.py:34-134) at the full-size sweep's density: a window of the
    whole 4148 x 4362 detector over 100 frames (1.81 G voxels: mu, A, M float32 and NS int16 resident, 25 GB), the
    reflections whose supports meet it (896k over 1800 frames -> 52.8k), supports 7 frames x 9 x 9 pixels (the central
    0.8 of the 7 x 11 x 11 shoeboxes), two terms (halo).  A step is one iteration of the intensity solve: scatter
    (mu += I f + beta), gather mode 1 (Newton gradient and Fisher diagonal), gather mode 0 (Fisher-vector product).
    Roofline per kernel: algorithmic bytes = factors (R, P) + per-row scalars + per in-window voxel 8 B (scatter:
    the float atomic's read and write), 6 B (mode 1: M, NS) or 10 B (mode 0: A, M, NS); over the kernel's HIP-event
    time on its launch stream."""
    from types import SimpleNamespace

    import numpy as np
    import torch
    from xpg_amd import forward as F

    torch.cuda.set_device(dev)
    nf = args.frames or 100
    ny, nx, Dz, b, T = 4362, 4148, 7, 9, 2
    n = int(round(896_000 * (nf + Dz - 1) / 1800))
    win = SimpleNamespace(f0=0, f1=nf, y0=0, ny=ny, x0=0, nx=nx)
    nv = nf * ny * nx
    rng = np.random.default_rng(7)
    iz0 = rng.integers(-Dz + 1, nf, n)
    iy0 = rng.integers(-b // 2, ny - b // 2, n)
    ix0 = rng.integers(-b // 2, nx - b // 2, n)
    o = np.argsort(iz0 * (ny * nx) + iy0 * nx + ix0, kind="stable")    # the model's row order: by support origin
    iz0, iy0, ix0 = iz0[o], iy0[o], ix0[o]
    kz = np.clip(iz0[:, None] + np.arange(Dz), -1, nf)
    ky = iy0[:, None] + np.arange(b)
    kx = ix0[:, None] + np.arange(b)
    in_vox = (((kz >= 0) & (kz < nf)).sum(1) * ((ky >= 0) & (ky < ny)).sum(1) * ((kx >= 0) & (kx < nx)).sum(1))
    n_in = int(in_vox.sum())
    g = torch.Generator(device=dev).manual_seed(11)
    # the factors from the full-size geometry (xpg_fm_terms): each support centred on its reflection, learnt profiles
    # of 9 regions, the 21-component rocking curve and the halo
    from xpg_amd import synth
    from xpg_amd.geometry import detector_consts
    geo = synth.make_geometry(synth.spec_full())
    xc, yc = (ix0 + b / 2 + 0.5).astype(float), (iy0 + b / 2 + 0.5).astype(float)
    S_t = torch.as_tensor(geo.pixel_to_s(xc, yc), device=dev)
    phi_t = torch.as_tensor(np.radians((iz0 + Dz / 2) * geo.osc), device=dev)
    from xpg_amd.integrate_kernels import DeviceDetector, gaussian_profile, rc_default_weights
    prof_t = torch.as_tensor(np.repeat(np.asarray(gaussian_profile(), np.float32)[None], 9, 0), device=dev)
    reg_t = torch.as_tensor(np.clip((xc / nx * 3).astype(int), 0, 2) * 3 + np.clip((yc / ny * 3).astype(int), 0, 2),
                            device=dev)
    F_det = DeviceDetector(detector_consts(geo), dev)
    org = [torch.as_tensor(a, dtype=torch.int32, device=dev) for a in (ix0, iy0, iz0)]
    terms_args = (F_det, geo.axis, geo.s0, 0.0, np.radians(geo.osc), np.radians(0.02), np.radians(0.1), S_t, phi_t,
                  *org, Dz, b)
    terms_kw = dict(prof=prof_t, region=reg_t, umax=4.5, halo=(0.1, 3.0), rcw=rc_default_weights())
    R, P = F.fraction_factors(*terms_args, **terms_kw)
    torch.cuda.synchronize(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        F.fraction_factors(*terms_args, **terms_kw)
    e1.record()
    torch.cuda.synchronize(dev)
    ms_terms = e0.elapsed_time(e1) / 5
    I = torch.rand(n, device=dev, generator=g) * 1000
    beta = torch.rand(n, device=dev, generator=g)
    t_iz, t_iy, t_ix = (torch.as_tensor(a, dtype=torch.int32, device=dev) for a in (iz0, iy0, ix0))
    rows = torch.arange(n, device=dev)
    B = torch.full((nv,), 2.0, device=dev)
    NS = torch.randint(0, 9, (nv,), dtype=torch.int16, device=dev, generator=g)
    A = torch.rand(nv, device=dev, generator=g)
    mu = torch.empty(nv, device=dev)
    gN = torch.zeros(n, device=dev)
    fd = torch.zeros(n, device=dev)
    gR = torch.zeros(n, device=dev)
    torch.cuda.synchronize(dev)
    log(f"forward window {nf}x{ny}x{nx} ({nv / 1e9:.2f} G voxels), {n} reflections, {n_in / 1e6:.1f} M in-window "
        f"support voxels")
    st = torch.cuda.current_stream(dev)

    def step(ev=None):
        mu.copy_(B)
        gN.zero_(), fd.zero_(), gR.zero_()
        if ev:
            ev[0].record(st)
        F.scatter(mu, R, P, I, beta, t_iz, t_iy, t_ix, win)
        if ev:
            ev[1].record(st)
        F.gather(gN, fd, R, P, rows, t_iz, t_iy, t_ix, win, None, mu, NS, 1)
        if ev:
            ev[2].record(st)
        F.gather(gR, None, R, P, rows, t_iz, t_iy, t_ix, win, A, mu, NS, 0)
        if ev:
            ev[3].record(st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    first = (gN.clone(), fd.clone(), gR.clone())
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    t0 = time.time()
    for ev in evs:
        step(ev)
    torch.cuda.synchronize(dev)
    wall = (time.time() - t0) / args.steps
    # mu's float atomics add in a varying order, so the gathers agree between steps to rounding only
    assert args.warmup == 0 or all(torch.allclose(a, c, rtol=1e-4, atol=1e-6) for a, c in zip(first, (gN, fd, gR))), \
        "gathers differ between steps"
    ms = {k: sum(ev[i].elapsed_time(ev[i + 1]) for ev in evs) / args.steps
          for i, k in enumerate(("k_fm_scatter", "k_fm_gather<newton>", "k_fm_gather<ratio>"))}
    fac = T * (Dz + b * b) * 4 + 20
    alg = {"k_fm_scatter": n * fac + 8 * n_in, "k_fm_gather<newton>": n * (fac + 8) + 6 * n_in,
           "k_fm_gather<ratio>": n * (fac + 4) + 10 * n_in}
    kern = {k: dict(avg_launch_ms=round(ms[k], 4), alg_bytes_per_launch=alg[k],
                    achieved=round(alg[k] / (ms[k] * 1e-3) / 1e9, 1),
                    frac=round(alg[k] / (ms[k] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)) for k in ms}
    # HBM bytes per launch from the committed counter summaries (the int16-count instantiations k_fm_gather<0, MODE>)
    for k, f in (("k_fm_scatter", "fm_scatter"), ("k_fm_gather<newton>", "fm_gather_0_1"),
                 ("k_fm_gather<ratio>", "fm_gather_0_0")):
        tf = ROOT / "profiles" / f"traffic_{f}_forward.json"
        kern[k]["traffic"] = json.loads(tf.read_text()).get("bytes_per_launch") if tf.exists() else None
    dom = max(ms, key=ms.get)
    # the factors: S, phi, origins in (40 B per reflection), R and P out, the profile grids (L2-resident)
    alg_t = n * (40 + T * (Dz + b * b) * 4)
    kern["k_fm_terms"] = dict(avg_launch_ms=round(ms_terms, 4), alg_bytes_per_launch=alg_t,
                              achieved=round(alg_t / (ms_terms * 1e-3) / 1e9, 1),
                              frac=round(alg_t / (ms_terms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                              note="per change of widths or geometry (the reference caches the factors between)")
    roof = dict(bound="hbm", kernel=dom, achieved=kern[dom]["achieved"], peak=PEAK_HBM_GBS, unit="GB/s",
                frac=kern[dom]["frac"], traffic=kern[dom]["traffic"], kernels=kern)
    t_step = sum(ms.values())
    cpu = None
    if not args.no_cpu:
        # the oracle (numpy float64, one core) on a 512 x 512 tile of the same window at the same density
        from oracle import forward as FO
        ty, tx = 512, 512
        m = (iy0 >= 0) & (iy0 < ty) & (ix0 >= 0) & (ix0 < tx)
        sw = SimpleNamespace(f0=0, f1=nf, y0=0, ny=ty, x0=0, nx=tx)
        Rh, Ph = R[:, torch.as_tensor(m, device=dev)].cpu().numpy(), P[:, torch.as_tensor(m, device=dev)].cpu().numpy()
        Ih, bh = I.cpu().numpy()[m], beta.cpu().numpy()[m]
        nvs = nf * ty * tx
        As = rng.random(nvs).astype(np.float32)
        NSs = rng.integers(0, 9, nvs).astype(np.int16)
        ns = int(m.sum())
        t = time.time()
        reps = 0
        while reps < 1 or time.time() - t < min(args.cpu_seconds, 10.0):
            mus = FO.scatter(np.full(nvs, 2.0), Rh, Ph, Ih, bh, iz0[m], iy0[m], ix0[m], sw)
            FO.gather(np.zeros(ns), np.zeros(ns), Rh, Ph, np.arange(ns), iz0[m], iy0[m], ix0[m], sw, None, mus, NSs, 1)
            FO.gather(np.zeros(ns), None, Rh, Ph, np.arange(ns), iz0[m], iy0[m], ix0[m], sw, As, mus, NSs, 0)
            reps += 1
        dt = (time.time() - t) / reps
        cpu = dict(value=round(ns / dt, 1), unit="reflections/s", cores=1, kind="port",
                   sample=f"{ns} reflections of a 512x512x{nf} tile of the same window, oracle/forward.py (numpy "
                          f"float64) scatter + both gathers, {reps} reps, {dt:.2f} s each")
        log(f"cpu baseline {cpu}")
    outj = dict(metric="forward-model window passes: reflections/s on 1 MI355X", value=round(n / (t_step * 1e-3), 1),
                unit="reflections/s", n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(t_step, 4),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f32", data="synthetic",
                config=dict(workload=f"forward_window_{nx}x{ny}x{nf}", reflections=n, support=[Dz, b, b], terms=T,
                            in_window_voxels=n_in, window_voxels=nv, ns_dtype="int16",
                            wall_ms_per_step=round(wall * 1e3, 3), parallelism="single GPU"),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(outj), flush=True)


def refine_sweep(n_blocks=36, frames_per_block=50, every=6, seed=11):
    """The block stage's input on the full-size sweep's geometry: spots of a crystal turning 0.02 deg per 5-deg block
    (slippage) about a fixed axis, every `every`-th predicted reflection to the detector edge (the strong spots a spot
    finder reports), 0.1 px / 0.05 frame noise, 2% displaced by 5-20 px; the starting model is the global one, its
    detector origin 0.3 / -0.2 px and distance 0.05 mm off.  Returns (start, spots, data_range, delphi)."""
    from dataclasses import replace

    import numpy as np
    from scipy.spatial.transform import Rotation
    from xpg_amd import synth
    from xpg_amd.geometry import Experiment, predict_reflections
    from xpg_amd.models import Crystal, b_matrix

    spec = synth.spec_full()
    g = synth.make_geometry(spec)
    c0 = Crystal.from_Astar(synth.make_experiment(spec).crystal.Astar, "o", space_group=spec.space_group)
    dmin = synth.edge_resolution(g)
    ax = np.array([0.3, -0.5, 0.8]) / np.linalg.norm([0.3, -0.5, 0.8])
    rng = np.random.default_rng(seed)
    xs, ys, zs = [], [], []
    for b in range(n_blocks):
        U = Rotation.from_rotvec(np.radians(0.02 * (b + 0.5)) * ax).as_matrix() @ c0.U
        z0, z1 = b * frames_per_block, (b + 1) * frames_per_block
        pr = predict_reflections(g, U @ b_matrix(c0.cell), dmin, z_range=(z0, z1))
        k = np.arange(pr["x"].size)[(b % every)::every]
        n = k.size
        x, y = pr["x"][k] + rng.normal(scale=0.1, size=n), pr["y"][k] + rng.normal(scale=0.1, size=n)
        out = rng.random(n) < 0.02
        ang = rng.uniform(0, 2 * np.pi, n)
        d = rng.uniform(5, 20, n)
        x, y = np.where(out, x + d * np.cos(ang), x), np.where(out, y + d * np.sin(ang), y)
        xs.append(x)
        ys.append(y)
        zs.append(np.clip(pr["z"][k] + rng.normal(scale=0.05, size=n), z0, z1 - 1e-6))
    start = Experiment(replace(g, orgx=g.orgx + 0.3, orgy=g.orgy - 0.2, distance=g.distance + 0.05), c0)
    return start, (np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)), \
        (1, n_blocks * frames_per_block), frames_per_block * spec.osc


def bench_refine(args, dev):
    """Per-phi-block geometry before integration (xpg/stage_integrate.py:28-212 with xpg/refine.py:26-443) on the
    full-size sweep's geometry: 36 blocks of 5 deg (1800 frames of 0.1 deg), ~4k spots per block (refine_sweep).  A step
    is the whole stage: the seed block, the tracking outwards with the weak position prior (three tolerances, up to 15
    Levenberg-Marquardt iterations and 2 rejection cycles each), the empirical-Bayes prior, the restrained pass over all
    blocks, then the smoothing of the orientation trajectory; flags POSITION, BEAM, ORIENTATION (8 parameters).
    Roofline: k_rf_residuals at the restrained pass's shape (every block's parameter vector and its 8 perturbations in
    one launch): algorithmic bytes = the spots read once (60 B each) + 24 B of residuals written per (spot, vector) +
    the model rows, over its HIP-event time; the stage itself is bound by the host's round trips (the device share is
    reported beside the wall time)."""
    import ctypes as C

    import numpy as np
    import torch
    from xpg_amd import _lib
    from xpg_amd import blocks as xb
    from xpg_amd import refine as xr

    torch.cuda.set_device(dev)
    nb = max(4, min(36, args.frames // 50)) if args.frames else 36
    start, spots, drange, delphi = refine_sweep(n_blocks=nb)
    n_spots = spots[0].size
    flags = ["POSITION", "BEAM", "ORIENTATION"]
    log(f"refine workload: {nb} blocks, {n_spots} spots")

    def step():
        state = {}
        blocks, edges = xb.block_models(start, spots, drange, delphi, flags, state=state, device=dev)
        return xb.smooth_blocks(blocks, start), state

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize(dev)
    xr.TIMING = []
    t = time.time()
    for _ in range(args.steps):
        out, state = step()
    torch.cuda.synchronize(dev)
    wall = (time.time() - t) / args.steps
    rounds = xr.TIMING
    xr.TIMING = None
    dev_ms = sum(a.elapsed_time(b) for a, b, _ in rounds) / args.steps
    launches = sum(n for _, _, n in rounds) / args.steps
    log(f"refine: {wall * 1e3:.1f} ms per stage, {len(rounds) / args.steps:.0f} device rounds, {launches:.0f} launches, "
        f"device {dev_ms:.1f} ms; seed block {state['seed']}, {state['n_fallback']} fallbacks, prior sd "
        f"{state.get('pos_prior_sd')}")
    # the residual kernel alone at the restrained pass's shape: every block's vector and its 8 perturbations
    edges = xb.block_edges(start.geom, drange, delphi)
    par = xr.Parameterisation(start, flags)
    reqs = []
    for i in range(nb):
        m = (spots[2] >= edges[i]) & (spots[2] < edges[i + 1])
        p0 = start.geom.spot_to_p0(spots[0][m], spots[1][m], spots[2][m])
        hkl, ok = xr.assign_indices(start.crystal.Astar, p0, tol=0.3)
        sp, su_, _ = xr._spot_data(start, hkl[ok], spots[0][m][ok], spots[1][m][ok], spots[2][m][ok], 0.5, 0.05)
        reqs.append(xr._Normal(xr._Problem(par, sp, su_), par.x0, np.ones(int(ok.sum()))))
    pool = xr._Pool(dev)
    M = par.x0.size + 1
    jobs = np.zeros(nb, xr._JOB)
    t0 = r0 = w0 = 0
    for j, rq in enumerate(reqs):
        jobs[j] = (pool.start_of(rq.prob), j * M, r0, t0, w0, rq.prob.n, M)
        w0 += rq.prob.n
        t0, r0 = t0 + M * rq.prob.n, r0 + 3 * M * rq.prob.n
    steps_ = np.concatenate([xr.steps_of(rq) for rq in reqs])
    X = np.vstack([par.x0[None], par.x0 + np.diag(xr.steps_of(reqs[0]))])
    models = torch.as_tensor(np.concatenate([par.models(X)] * nb), device=dev)
    jobs_d = torch.from_numpy(jobs.view(np.uint8)).to(dev)
    r_d = torch.empty(r0, dtype=torch.float64, device=dev)
    w_d = torch.ones(pool.n, dtype=torch.float64, device=dev)
    h_d = torch.as_tensor(steps_, device=dev)
    o_d = torch.empty(nb * (M * M - M), dtype=torch.float64, device=dev)
    st, su = pool.spots(), xr._setup(reqs[0].prob, M)
    h = _lib.lib()
    stream = _lib.stream_ptr(dev)
    ms = {}
    for name, call in (("k_rf_residuals", lambda: h.xpg_rf_residuals(C.byref(st), C.byref(su), _lib.ptr(jobs_d), nb,
                                                                     t0, _lib.ptr(models), _lib.ptr(r_d), stream)),
                       ("k_rf_normal", lambda: h.xpg_rf_normal(C.byref(st), C.byref(su), _lib.ptr(jobs_d), nb,
                                                               _lib.ptr(r_d), _lib.ptr(h_d), _lib.ptr(w_d),
                                                               _lib.ptr(o_d), None, stream))):
        _lib.check(call())
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            _lib.check(call())
        e1.record()
        torch.cuda.synchronize(dev)
        ms[name] = e0.elapsed_time(e1) / 20
    n_res = pool.n
    alg = {"k_rf_residuals": n_res * 60 + t0 * 24 + models.numel() * 8,
           "k_rf_normal": 3 * n_res * M * 8 + n_res * 8 + nb * (M * M - M) * 8}
    kern = {k: dict(avg_launch_ms=round(ms[k], 4), alg_bytes_per_launch=int(alg[k]),
                    achieved=round(alg[k] / (ms[k] * 1e-3) / 1e9, 1),
                    frac=round(alg[k] / (ms[k] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)) for k in ms}
    kern["k_rf_residuals"]["evaluations"] = int(t0)
    kern["k_rf_residuals"]["Gevals_per_s"] = round(t0 / (ms["k_rf_residuals"] * 1e-3) / 1e9, 2)
    roof = dict(bound="hbm", kernel="k_rf_residuals", achieved=kern["k_rf_residuals"]["achieved"], peak=PEAK_HBM_GBS,
                unit="GB/s", frac=kern["k_rf_residuals"]["frac"], traffic=None, kernels=kern,
                stage=dict(wall_ms=round(wall * 1e3, 2), device_ms=round(dev_ms, 2), rounds=len(rounds) // args.steps,
                           launches=int(launches),
                           note="each round trip carries every waiting problem's requests; the host's p x p solves, "
                                "index assignment and waits bound the stage"))
    cpu = None
    if not args.no_cpu:
        # the oracle (numpy float64, one core) on the first 4 blocks of the same sweep: a whole stage
        from oracle import refine as orf
        n4 = 4
        sel = spots[2] < n4 * 50
        t = time.time()
        ob, _ = orf.block_models(start, tuple(v[sel] for v in spots), (1, n4 * 50), delphi, flags)
        orf.smooth_blocks(ob, start)
        dt = time.time() - t
        cpu = dict(value=round(n4 / dt, 3), unit="blocks/s", cores=1, kind="port",
                   sample=f"the whole block stage on the first {n4} blocks ({int(sel.sum())} spots), oracle/refine.py "
                          f"(numpy float64, one core), {dt:.1f} s")
        log(f"cpu baseline {cpu}")
    outj = dict(metric="per-block geometry refinement: phi blocks/s on 1 MI355X", value=round(nb / wall, 3),
                unit="blocks/s", n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(wall * 1e3, 2),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload=f"refine_blocks_4148x4362x{nb * 50}", blocks=nb, spots=int(n_spots),
                            flags=flags, parameters=int(par.x0.size), seed_block=int(state["seed"]),
                            fallbacks=int(state["n_fallback"]), parallelism="single GPU"),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(outj), flush=True)


def bench_cbf(args, dev):
    """CBF byte-offset decode of PILATUS 6M frames (2463 x 2527 int32: Poisson background, module gaps at -1, bad
    pixels at -2, hot pixels up to the count cut-off), encoded as read_cbf reads them (encode_byte_offset); the streams
    resident in HBM; a step decodes every frame into device int32 values (5 kernel launches).  Roofline: stream bytes
    read + 4 bytes per pixel written, over the step's HIP-event time on the launch stream."""
    import numpy as np
    import torch
    from concurrent.futures import ThreadPoolExecutor
    from xpg_amd import cbf

    torch.cuda.set_device(dev)
    nfr = args.frames or 200
    ny, nx = 2527, 2463
    g = torch.Generator(device=dev).manual_seed(1005)
    t = time.time()
    host = np.empty((nfr, ny, nx), np.int32)
    for f in range(nfr):
        lam = torch.full((ny, nx), 0.5 + 5.5 * (f % 17) / 16, device=dev)
        a = torch.poisson(lam, generator=g).to(torch.int32)
        a[:, 487::494] = -1
        a[195::212, :] = -1
        idx = torch.randint(0, ny * nx, (3000,), device=dev, generator=g)
        a.view(-1)[idx] = torch.randint(1000, 1_048_500, (3000,), device=dev, generator=g, dtype=torch.int32)
        host[f] = a.cpu().numpy()
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        streams = list(ex.map(cbf.encode_byte_offset, list(host)))
    log(f"generated and encoded {nfr} frames {ny}x{nx} in {time.time() - t:.1f}s")
    sb = cbf.StreamBatch(streams, [ny * nx] * nfr, dev)
    ref = torch.from_numpy(host.reshape(-1)).to(dev)
    out = torch.zeros(nfr * ny * nx, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        sb.decode_into(out)
    sb.check()
    assert torch.equal(out, ref), "decoded frames differ from the originals"
    out.zero_()
    torch.cuda.synchronize(dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.time()
    for a, b in ev:
        a.record(st)
        sb.decode_into(out)
        b.record(st)
    torch.cuda.synchronize(dev)
    wall = (time.time() - t0) / args.steps
    ms = sum(a.elapsed_time(b) for a, b in ev) / args.steps
    sb.check()
    assert torch.equal(out, ref), "decoded frames differ from the originals"
    raw = nfr * ny * nx * 4
    alg = sb.stream_bytes + raw
    achieved = alg / (ms * 1e-3) / 1e9
    # measured HBM bytes of the decode's two large kernels (maps and write; the three walks move a few MB)
    tfs = [ROOT / "profiles" / f"traffic_cbf_{k}_cbf.json" for k in ("maps", "write")]
    traffic = (sum(json.loads(t.read_text())["bytes_per_launch"] for t in tfs) if all(t.exists() for t in tfs)
               else None)
    roof = dict(bound="hbm", kernel="xpg_cbf_decode (k_cbf_maps, k_cbf_groups, k_cbf_frames, k_cbf_chunks, k_cbf_write)",
                achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s", frac=round(achieved / PEAK_HBM_GBS, 4),
                traffic=traffic, avg_launch_ms=round(ms, 4), alg_bytes_per_launch=alg, launches_per_step=5,
                stream_bytes=sb.stream_bytes, decoded_bytes=raw)
    cpu = None
    if not args.no_cpu:
        from oracle import codec as oc
        t = time.time()
        n = 0
        while n < nfr and (time.time() - t < args.cpu_seconds or n < 2):
            r, _ = oc.cbf_decode(streams[n], ny * nx)
            assert r == 0
            n += 1
        dt = time.time() - t
        cpu = dict(value=round(n / dt, 3), unit="frames/s", cores=1, kind="port",
                   sample=f"{n} frames of {ny}x{nx} (the first of the same streams), oracle/codec.c cbf_decode, "
                          f"{dt:.1f} s")
        log(f"cpu baseline {cpu}")
    outj = dict(metric="CBF byte-offset frames decoded/s on 1 MI355X", value=round(nfr / (ms * 1e-3), 1),
                unit="frames/s", n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 3),
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="int64", data="synthetic",
                config=dict(workload=f"cbf_decode_{nx}x{ny}x{nfr}", frames=nfr, frame_shape=[ny, nx],
                            bytes_per_pixel=round(sb.stream_bytes / (nfr * ny * nx), 3),
                            decoded_GBps=round(raw / (ms * 1e-3) / 1e9, 1), wall_ms_per_step=round(wall * 1e3, 3),
                            identical_to_source=True, parallelism="single GPU"),
                roofline=roof, cpu_baseline=cpu)
    print(json.dumps(outj), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="full", choices=["full", "small", "uneven", "stress", "serve", "decode", "cbf",
                                                                 "forward", "refine"])
    ap.add_argument("--serve-datasets", type=int, default=32)
    ap.add_argument("--serve-frame-scale", type=float, default=0.1,
                    help="serving workload: frame counts U{50..3600} times this factor (at least 20 frames)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=0.8, help="offered load as a fraction of measured capacity")
    ap.add_argument("--frames", type=int, default=0, help="override the number of frames (smaller runs)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="CPU baseline: at most this many reflections of the sample group (0: the whole group)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: every host core, <= 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", default="both", choices=["both", "batched", "plain"],
                    help="batched (the reported value) and plain (reported beside it), or one of them")
    ap.add_argument("--autotune", action="store_true",
                    help="time the batched form's launch-parameter candidates (xpg_amd/autotune.py) on this workload "
                         "first and run with the fastest")
    ap.add_argument("--autotune-save", default="", help="write the autotuner's choice to this file (tuned.json form)")
    ap.add_argument("--defer-pass2", type=int, default=1,
                    help="batched form: launch each block's pass 2 behind the next block's pass 1 (1) or right after "
                         "its own chain (0)")
    ap.add_argument("--prep-ahead", type=int, default=-1,
                    help="batched form: prepare blocks this many ahead of the chain (-1: all up front)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one copy-like kernel (out = -in, float32) over MIB MiB, outside the timed region: "
                         "the reference of known traffic for hardware byte counters")
    ap.add_argument("--summarize-stats", nargs=3, metavar=("STATS_CSV", "COMMAND", "OUT_MD"),
                    help="write a markdown summary of a rocprofv3 kernel-stats csv and exit")
    ap.add_argument("--summarize-sq", nargs=3, metavar=("PMC_CSV", "WORKLOAD", "COMMAND"),
                    help="write profiles/counters_<workload>.md/.json from a rocprofv3 --pmc run of SQ counters and exit")
    ap.add_argument("--summarize-pmc", nargs=6, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_MIB", "KERNEL", "WORKLOAD",
                                                          "COMMAND"),
                    help="write profiles/traffic_<kernel>_<workload>.json/.md from two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.summarize_stats:
        summarize_stats(*args.summarize_stats)
        return
    if args.summarize_sq:
        print(json.dumps(summarize_sq(*args.summarize_sq, kernels=("k_prepare", "k_rocking", "k_pass1", "k_pass2_small",
                                                                    "k_pass2", "k_normalise", "k_extract", "k_finalise"))))
        return
    if args.summarize_pmc:
        f, w, mib, kern, wl, cmd = args.summarize_pmc
        print(json.dumps(summarize_pmc(f, w, int(mib), kern, wl, cmd)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: --gpus must be 1 (the sweep shards by phi range and serving by dataset "
              "across GPUs; neither is run here)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from xpg_amd import synth
    from xpg_amd.integrate import IntegrationModel, Sweep, _integrate

    dev = torch.device("cuda:0")
    if args.workload == "serve":
        bench_serve(args, dev)
        return
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy * 2 ** 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize(dev)
        log(f"copy calibration: {src.numel() * 4} bytes read, {dst.numel() * 4} bytes written")
        del src, dst
    if args.workload == "decode":
        bench_decode(args, dev)
        return
    if args.workload == "cbf":
        bench_cbf(args, dev)
        return
    if args.workload == "forward":
        bench_forward(args, dev)
        return
    if args.workload == "refine":
        bench_refine(args, dev)
        return
    spec = dict(full=synth.spec_full, small=synth.spec_small, uneven=synth.spec_uneven,
                stress=synth.spec_stress)[args.workload]()
    if args.frames:
        spec = synth.replace(spec, nframes=args.frames, name=f"{spec.name}[:{args.frames}]")
    t = time.time()
    frames, spots = synth.generate(spec, dev, log=log)
    log(f"generated {spec.nframes} frames {spec.ny}x{spec.nx} ({frames.nbytes() / 1e9:.1f} GB) and "
        f"{len(spots['x'])} spots in {time.time() - t:.1f}s")
    exp = spots["exp"]
    edges = synth.block_edges(spec)
    model = IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    t = time.time()
    hp = _integrate(frames, [exp] * (len(edges) - 1), edges, model, spots["dmin"], plan_only=True, verbose=False,
                    _box_draw=synth.box_draw_for(spec))
    t_host_plan = time.time() - t
    from xpg_amd.integrate import DZ_BUCKETS  # noqa: F401
    opts = dict(frames=(1, spec.nframes), mask=None, exclude_frames=(), learn_profiles=True, adaptive=True,
                bkg_template=None, moments_only=False, nreg=9, sd=np.radians(model.sigma_d),
                sm=np.radians(model.sigma_m), prep_ahead=None if args.prep_ahead < 0 else args.prep_ahead,
                defer_pass2=bool(args.defer_pass2))
    tuned = None
    if args.autotune:
        from xpg_amd import autotune
        best, table = autotune.tune(frames, [exp] * (len(edges) - 1), model, hp, opts, log=log)
        opts.update(lds_bins=tuple(best["lds_bins"]), wg_waves=best["wg_waves"])
        tuned = dict(best=best, candidates=len(table), worst_ms=max(r["ms"] for r in table),
                     all_identical=all(r["identical"] for r in table))
        if args.autotune_save:
            autotune.save(best, table, spec.name, args.autotune_save)
    t = time.time()
    sweep = Sweep(frames, [exp] * (len(edges) - 1), model, hp, opts)
    torch.cuda.synchronize(dev)
    t_setup = time.time() - t
    N = sweep.sp.n
    # pinned host memory for the results: every timed step ends with them copied there
    hostbuf = sweep.host_results()
    log(f"planned {N} reflections in {len(sweep.sp.groups)} shoebox groups ({len(sweep.sp.bins)} size-bin launches) "
        f"over {len(edges) - 1} blocks (host plan {t_host_plan:.1f}s, device setup {t_setup:.1f}s)")

    events = []
    iso_events = []

    class Timer:
        """HIP events around one launch, recorded on the stream the launch goes to."""
        def __init__(self, kind, b, rows, st):
            self.kind, self.gi, self.st = kind, rows, st

        def __enter__(self):
            self.e0 = torch.cuda.Event(enable_timing=True)
            self.e1 = torch.cuda.Event(enable_timing=True)
            self.e0.record(self.st)

        def __exit__(self, *a):
            self.e1.record(self.st)
            events.append((self.kind, self.gi, self.e0, self.e1))
            return False

    def time_form(form, steps, warmup, timer):
        sweep.o["form"] = form
        for i in range(warmup):
            r = sweep.run()
            log(f"{form} warmup {i + 1}/{warmup}: {r['t_gpu'] * 1e3:.1f} ms")
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        last = None
        for i in range(steps):
            last = sweep.enqueue(timer, hostbuf)       # each block's results to pinned host memory behind its pass 2
            torch.cuda.synchronize(dev)
        el = time.perf_counter() - t0
        integrated = int(torch.isfinite(last["o2"]["iprf"]).sum().item())
        log(f"{form}: timed {steps} steps: {el / steps * 1e3:.1f} ms/step, {integrated} profile-fitted "
            f"reflections/step")
        return el / steps, integrated, last

    plain = None
    graph_info = None
    form = "plain" if args.form == "plain" else "batched"
    if form == "plain":
        t_plain, n_plain, last_plain = time_form("plain", args.steps, args.warmup, None)
        plain = dict(value=round(n_plain / t_plain, 1), ms_per_step=round(t_plain * 1e3, 3))
        t_step, integrated, last = t_plain, n_plain, last_plain
        events.clear()
        time_form("plain", 1, 0, Timer)
    else:
        # one eager batched run with HIP events around every launch (per-kernel durations for the roofline); the
        # batched form timed with eager launches on its three streams and as replays of one captured HIP graph: the
        # faster of the two is the reported value, the other is reported beside it
        _, _, eager = time_form("batched", 1, args.warmup, Timer)
        t_eager, n_eager, last_eager = time_form("batched", args.steps, 0, None)
        sweep.capture(hostbuf)                         # the per-block result copies are graph nodes too
        for i in range(args.warmup):
            sweep.replay()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        for i in range(args.steps):
            sweep.graph.replay()
            torch.cuda.synchronize(dev)
        t_graph = (time.perf_counter() - t0) / args.steps
        last_graph = dict(sweep.graph_out)
        n_graph = int(torch.isfinite(last_graph["o2"]["iprf"]).sum().item())
        log(f"batched (graph): timed {args.steps} replays: {t_graph * 1e3:.1f} ms/step, {n_graph} profile-fitted "
            f"reflections/step")
        graph_info = dict(eager_ms_per_step=round(t_eager * 1e3, 3), graph_ms_per_step=round(t_graph * 1e3, 3))
        if t_graph < t_eager:
            t_step, integrated, last = t_graph, n_graph, last_graph
            graph_info["reported"] = "graph replay"
        else:
            t_step, integrated, last = t_eager, n_eager, last_eager
            graph_info["reported"] = "eager launches"
        b_ = last["o2"]["iprf"]
        for other in (eager, last_eager, last_graph):
            if not torch.equal(torch.nan_to_num(other["o2"]["iprf"], nan=-1.0), torch.nan_to_num(b_, nan=-1.0)):
                graph_info["identical_results"] = False
        graph_info.setdefault("identical_results", True)
        # one more eager run, outside the timed region, with the device synchronised around every launch: each
        # kernel's duration alone (the timed sweep's overlapping streams stretch every launch by the kernels beside it)
        class IsoTimer(Timer):
            def __enter__(self):
                torch.cuda.synchronize(dev)
                super().__enter__()

            def __exit__(self, *a):
                self.e1.record(self.st)
                torch.cuda.synchronize(dev)
                iso_events.append((self.kind, self.gi, self.e0, self.e1))
                return False
        sweep.enqueue(IsoTimer, hostbuf)
        torch.cuda.synchronize(dev)
    ms = t_step * 1e3
    value = integrated / t_step

    # the last step's results as copied to pinned host memory
    fin = lambda t: torch.nan_to_num(t.cpu(), nan=-1.0)
    src = {**last["o1"], **last["o2"]}
    host_copy_identical = all(torch.equal(fin(hostbuf[k]), fin(src[k])) for k in sweep.RESULT_KEYS)

    # the plain form beside it, timed after the batched one with the batched form's prepared data and graph released
    # (each form is timed with only its own buffers allocated: with both held, the largest workloads brought the
    # allocator close to the device's memory, and the form timed second ran up to 15% slower)
    if form == "batched" and args.form == "both":
        sweep.prep, sweep.prep_g = {}, {}
        del sweep.graph, sweep.graph_out, eager, last_eager, last_graph
        torch.cuda.synchronize(dev)
        torch.cuda.empty_cache()
        t_plain, n_plain, last_plain = time_form("plain", args.steps, args.warmup, None)
        a = last_plain["o2"]["iprf"]
        plain = dict(value=round(n_plain / t_plain, 1), ms_per_step=round(t_plain * 1e3, 3),
                     identical_results=bool(torch.equal(torch.nan_to_num(a, nan=-1.0),
                                                        torch.nan_to_num(b_, nan=-1.0))))

    # dominant kernel and its roofline, over every kernel of the sweep (HIP events on the launch stream, algorithmic
    # bytes per launch)
    def per_kind(evs):
        per = {}
        for kind, (start, m), e0, e1 in evs:
            Bs = sweep.sp.rows_dz[start:start + m]
            bxs = sweep.sp.rows_bx[start:start + m]
            d = per.setdefault(kind, dict(ms=0.0, bytes=0.0, n=0))
            d["ms"] += e0.elapsed_time(e1)
            d["bytes"] += float(np.sum(per_refl_bytes(Bs * bxs * bxs, Bs, kind)))
            d["n"] += 1
        return per
    per, iso = per_kind(events), per_kind(iso_events)
    n_timed = 1
    dom = max(per, key=lambda k: per[k]["ms"])
    dd = per[dom]
    avg_ms = dd["ms"] / dd["n"]
    avg_bytes = dd["bytes"] / dd["n"]
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{dom}_{args.workload}.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
    roof = dict(bound="hbm", kernel=f"k_{dom}", achieved=round(achieved, 1), peak=PEAK_HBM_GBS, unit="GB/s",
                frac=round(achieved / PEAK_HBM_GBS, 4), traffic=traffic, avg_launch_ms=round(avg_ms, 4),
                alg_bytes_per_launch=round(avg_bytes), launches_per_step=dd["n"] // n_timed,
                kernels={k: dict(ms_per_step=round(v["ms"] / n_timed, 3),
                                 GBps=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1)) for k, v in per.items()})
    if iso:
        # the same kernels timed alone (IsoTimer): what each achieves without the kernels that share the device in the
        # sweep; frac stays the in-sweep figure
        for k, v in iso.items():
            roof["kernels"].setdefault(k, {}).update(isolated_ms_per_step=round(v["ms"], 3),
                                                     isolated_GBps=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1))
        if dom in iso:
            roof["frac_isolated"] = round(iso[dom]["bytes"] / (iso[dom]["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)

    # the whole step against the HBM roofline: B_alg (every frame pixel once + two passes per shoebox) and the bound of
    # a design that reads only shoebox voxels, B_min (their exact union once + two passes)
    sb_ = step_bytes(hp, spec.nframes, spec.ny, spec.nx, dev)
    roof["step"] = dict(B_alg_GB=round(sb_["B_alg"] / 1e9, 3), B_min_GB=round(sb_["B_min"] / 1e9, 3),
                        union_voxels=sb_["union_voxels"], sum_V=sb_["sum_V"],
                        frac_B_alg=round(sb_["B_alg"] / t_step / (PEAK_HBM_GBS * 1e9), 4),
                        frac_B_min=round(sb_["B_min"] / t_step / (PEAK_HBM_GBS * 1e9), 4),
                        t_roof_B_min_ms=round(sb_["B_min"] / (PEAK_HBM_GBS * 1e9) * 1e3, 3))
    log(f"step bytes {roof['step']}")

    cpu, parity = (None, None)
    if not args.no_cpu:
        t = time.time()
        cpu, parity = cpu_baseline(sweep, hp, spec, frames, args.cpu_sample, args.cpu_seconds,
                                   args.cpu_threads or host_cores())
        log(f"cpu baseline {cpu['value']} refl/s ({time.time() - t:.1f}s); parity {parity}")

    out = dict(metric="profile-fitted reflections/s on 1 MI355X", value=round(value, 1), unit="reflections/s",
               n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=round(ms, 3), higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
               timed_region="every launch of the sweep from shoebox gather to pass 2 and the per-reflection results "
                            "copied to pinned host memory (each block's behind its pass 2, on a copy stream; "
                            "host_copy_identical); frames and the plan resident before",
               plan=dict(host_plan_s=round(t_host_plan, 3), device_setup_s=round(t_setup, 3),
                         value_incl_plan=round(integrated / (t_step + t_host_plan + t_setup), 1)),
               host_copy_identical=host_copy_identical,
               config=dict(workload=spec.name, form=FORMS[form],
                           frames=spec.nframes, frame_shape=[spec.ny, spec.nx], reflections=N,
                           integrated=integrated, blocks=len(edges) - 1, box=[spec.box_z, spec.box_xy, spec.box_xy]
                           if spec.box_xy else (f"drawn: side log-U{list(spec.box_draw[0])}, depth log-U"
                                                f"{list(spec.box_draw[1])}, odd" if spec.box_draw else "reference rules"), parallelism="single GPU"),
               roofline=roof, cpu_baseline=cpu, parity=parity,
               plain=plain, vs_plain=round(value / plain["value"], 3) if plain and form == "batched" else None,
               batched_forms=graph_info if form == "batched" else None,
               launch=dict(lds_bins=list(sweep.sp.lds_bins), wg_waves=sweep.sp.wg_waves, autotune=tuned))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
