"""Serving: frame batches of several concurrent datasets integrated as they arrive, with continuous batching.

Each request is a rotation dataset whose frames arrive in batches of ``batch_frames`` frames.  A dataset is admitted
when a slot and its HBM share are free (the budget covers its frame ring, reflection table, outputs and working
buffers); its per-dataset state lives on the device for as long as it is active:

  frame ring     the last R frames (R from the plan: every phi block's frames are resident when the block runs)
  table          the reflection table of the whole dataset (SweepPlan), uploaded once at admission
  widths chain   the adaptive sigma_D / sigma_M / rocking-curve records carried block to block (xpg_adapt_block)
  profiles       the learnt reference profiles of every block
  partials       reflections whose frames reach past a block's own frames wait, on the device, until the frames they
                 need have arrived: a block runs once its last needed frame is resident

Each admitted dataset owns one stream; its batches are written into the ring on that stream and every block that a
batch completes is enqueued right behind it (gather, pass 1, fits, normalisation, pass 2: Sweep.enqueue_block), so
the work of all active datasets overlaps on the device while the host loop admits new requests as slots and memory
free up.  A block is processed exactly as in the offline sweep (the same kernels on the same values, in the same
order), so a dataset's results are the same alone and in any mix.

Per-batch latency is measured on the device timeline: from the end of a batch's arrival (its frames resident) to
the completion of every result it completes (the reflections whose last frame is in the batch), with HIP events.
The per-block sequence follows xpg/integrate.py:463-574; the frame walk and block readiness follow
xpg/integrate.py:463-520 and :953-957.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import synth
from .frames import DeviceFrames
from .integrate import IntegrationModel, RunBufs, Sweep, _integrate, sweep_plan
from .passes import ring_frames  # noqa: F401  (the ring of frames a dataset needs; re-exported)


@dataclass
class Request:
    """One dataset: its synthetic spec (geometry, cell, frame count, seed) and arrival time (s)."""
    spec: synth.SynthSpec
    arrival: float = 0.0
    rid: int = 0


@dataclass
class Plan:
    """Host plan of a request (computed when it is submitted): spot table of the generator, integration plan and
    the frame ring it needs."""
    req: Request
    spots: dict
    exp: object
    edges: np.ndarray
    model: IntegrationModel
    hp: dict
    ring: int
    order: list
    ready_batch: list            # per block of ``order``: the batch after which it is enqueued
    done_batch: np.ndarray       # per reflection (table order of hp): the batch that completes it
    spot_idx: np.ndarray = None  # the generator's spots reaching each batch (batch_spot_index)
    spot_off: np.ndarray = None


def plan_request(req: Request, batch_frames=10):
    spec = req.spec
    exp = synth.make_experiment(spec)
    spots = synth.spot_table(spec, exp)
    edges = synth.block_edges(spec)
    model = IntegrationModel(spec.sigma_d, spec.sigma_m, box_xy=spec.box_xy, box_z=spec.box_z)
    img = _ImageStub(spec)
    blocks = [exp] * (len(edges) - 1)
    hp = _integrate(img, blocks, edges, model, spots["dmin"], plan_only=True, verbose=False,
                    _box_draw=synth.box_draw_for(spec))
    order = [b for b in hp["block_order"] if hp["blocks_refl"][b].size]
    ring, ready = ring_frames(order, hp["need_lo"], hp["need_hi"], batch_frames, spec.nframes)
    zend = np.minimum(hp["iz0"] + hp["dz"] * hp["fb"] - 1, spec.nframes - 1)
    done = np.maximum(zend, 0) // batch_frames
    hp["sweep_plan"] = sweep_plan(hp, blocks)
    idx, off = synth.batch_spot_index(spots, spec.nframes, batch_frames)
    return Plan(req, spots, exp, edges, model, hp, ring, order, ready, done, idx, off)


class _ImageStub:
    def __init__(self, spec):
        self.first, self.last, self.nx, self.ny = 1, spec.nframes, spec.nx, spec.ny


@dataclass
class _Active:
    plan: Plan
    stream: torch.cuda.Stream
    sweep: Sweep
    ring: torch.Tensor
    dspots: synth.DeviceSpots
    arena: dict
    rb: RunBufs
    nbytes: int
    next_batch: int = 0
    next_block: int = 0
    ev_res: list = field(default_factory=list)
    ev_blk: list = field(default_factory=list)
    t_admit: float = 0.0
    t_first: float = 0.0
    spot_idx: torch.Tensor = None


def dataset_bytes(plan: Plan, sweep: Sweep | None = None):
    """HBM bytes a dataset holds while active: frame ring, spot table, reflection table and outputs, fits
    workspace, the per-block arena of prepared data."""
    spec = plan.req.spec
    n = plan.hp["refl"]["x"].size
    ring = plan.ring * spec.nx * spec.ny * 4
    spots = 8 * 9 * len(plan.spots["x"])
    table = n * (24 + 8 + 7 * 4 + 24 + 8 + 5 * 4 * 8 + 3 * 4 + 2 * 8 + 4 * 8)     # refl view, neighbours, fit inputs
    outputs = n * (8 * 16 + 4 * 3 + 8 * 9 + 4 * 65)                                 # pass-1 / pass-2 outputs, fz
    arena = sweep.arena_bytes() if sweep is not None else 0
    return int(ring + spots + table + outputs + arena + (64 << 20))


class Server:
    """Continuous-batching request loop on one device.

    slots: datasets active at once (each on its own stream); hbm_budget: bytes the active datasets may hold;
    batch_frames: frames per arriving batch; lookahead: batches a dataset may have in flight on the device."""

    def __init__(self, device, hbm_budget=200e9, slots=4, batch_frames=10, lookahead=8, log=None):
        self.device = torch.device(device)
        self.budget = float(hbm_budget)
        self.slots = int(slots)
        self.batch = int(batch_frames)
        self.lookahead = int(lookahead)
        self.log = log or (lambda *a: None)

    # --- one dataset -------------------------------------------------------------------------------------------------

    def _stream(self):
        """A free stream of the pool (one per slot, default priority): a dataset's allocations are made on its stream,
        so the memory a finished dataset frees is reused, in stream order, by the next dataset on that stream."""
        if not hasattr(self, "_free"):
            self._free = [torch.cuda.Stream(self.device) for _ in range(self.slots)]
        return self._free.pop()

    def _admit(self, plan: Plan):
        spec = plan.req.spec
        st = self._stream()
        if not hasattr(self, "_up"):
            self._up = torch.cuda.Stream(self.device)       # uploads: an otherwise idle stream, so that the host
        with torch.cuda.stream(self._up):                    # copies do not wait for the slot stream's earlier work
            sweep_t = self._upload(plan)
        st.wait_stream(self._up)
        with torch.cuda.stream(st):
            ring = torch.empty((plan.ring, spec.ny, spec.nx), dtype=torch.int32, device=self.device)
            frames = DeviceFrames(ring, first=1, overload=synth.OVERLOAD, flag_value=synth.FLAG)
            opts = dict(frames=(1, spec.nframes), mask=None, exclude_frames=(), learn_profiles=True, adaptive=True,
                        bkg_template=None, moments_only=False, nreg=9, sd=np.radians(plan.model.sigma_d),
                        sm=np.radians(plan.model.sigma_m))
            sweep = Sweep(frames, [plan.exp] * (len(plan.edges) - 1), plan.model, plan.hp, opts, uploaded=sweep_t)
            sweep.store.ring = plan.ring
            sweep.store.n_frames = 0          # frames arrived so far (every frame of the dataset is read)
            sweep.store.available = None
            if sweep.order != plan.order:
                raise L.XpgError("block order of the sweep differs from the plan")
            arena = sweep.arena()
            rb = RunBufs(sweep)
            dspots, spot_idx = sweep_t["spots"], sweep_t["spot_idx"]
            sweep._init_widths()
        return _Active(plan, st, sweep, ring, dspots, arena, rb, dataset_bytes(plan, sweep), spot_idx=spot_idx)

    def _upload(self, plan: Plan):
        """Everything a dataset uploads once (reflection table, detector, spot table and per-batch spot index), on the
        current stream."""
        sp = plan.hp["sweep_plan"].to(self.device)
        dets = {}
        from .integrate import detector_consts
        from . import integrate_kernels as K
        g = plan.exp.geom
        dets[id(g)] = (g, K.DeviceDetector(detector_consts(g, None), self.device))
        return dict(sp=sp, dets=dets, spots=synth.DeviceSpots(plan.spots, self.device),
                    spot_idx=torch.as_tensor(plan.spot_idx, device=self.device))

    def _step(self, a: _Active):
        """Arrival of a dataset's next batch (written into the ring on its stream) and every block it makes ready."""
        spec = a.plan.req.spec
        k = a.next_batch
        f0 = k * self.batch
        nf = min(self.batch, spec.nframes - f0)
        slot = f0 % a.plan.ring
        t0 = time.perf_counter()
        with torch.cuda.stream(a.stream):
            o0, o1_ = int(a.plan.spot_off[k]), int(a.plan.spot_off[k + 1])
            synth.render(spec, a.dspots, a.ring[slot:slot + nf], f0, nf, sync=False, sel=a.spot_idx[o0:o1_])
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(a.stream)
            a.ev_res.append(ev)
            a.sweep.store.n_frames = f0 + nf
            t1 = time.perf_counter()
            self.host["render"] += t1 - t0
            while a.next_block < len(a.plan.order) and a.plan.ready_batch[a.next_block] <= k:
                a.sweep.enqueue_block(a.next_block, a.sweep.store, a.arena, a.rb)
                e = torch.cuda.Event(enable_timing=True)
                e.record(a.stream)
                a.ev_blk.append(e)
                a.next_block += 1
            self.host["blocks"] += time.perf_counter() - t1
        a.next_batch += 1

    def _in_flight(self, a: _Active):
        return sum(1 for e in a.ev_res[-self.lookahead - 1:] if not e.query())

    def _finished(self, a: _Active):
        n_batches = -(-a.plan.req.spec.nframes // self.batch)
        return a.next_batch >= n_batches and a.next_block >= len(a.plan.order)

    def results(self, a: _Active):
        """Host results in prediction order (after the dataset's stream has drained)."""
        res = a.sweep.finish(dict(a.rb.result(), t_gpu=float("nan")))
        return res

    def latencies(self, a: _Active):
        """Per result-bearing batch: ms from its frames resident to the completion of every reflection it completes."""
        sp = a.sweep.sp
        blk_of_row = np.empty(sp.n, np.int64)
        for i, b in enumerate(a.sweep.order):
            for gi, gk, start, m in a.sweep.by_block[b]:
                blk_of_row[start:start + m] = i
        done = a.plan.done_batch[sp.perm]
        last = np.full(len(a.ev_res), -1, np.int64)
        np.maximum.at(last, done, blk_of_row)
        return {k_: a.ev_res[k_].elapsed_time(a.ev_blk[i]) for k_, i in enumerate(last.tolist()) if i >= 0}

    # --- the loop ------------------------------------------------------------------------------------------------------

    def run(self, plans, keep_results=False):
        """Serve ``plans`` (arrival times from their requests, seconds from the start).  Returns a report: per-batch
        latencies, per-dataset timings and, with keep_results, each dataset's IntegrationResult."""
        pending = sorted(plans, key=lambda p: (p.req.arrival, p.req.rid))
        queue, active, done = [], [], []
        used = 0.0
        n_refl = 0
        lat = []
        per = {}
        results = {}
        t0 = time.perf_counter()
        first_gpu = None
        self.host = dict(admit=0.0, render=0.0, blocks=0.0, retire=0.0, idle=0.0)
        while pending or queue or active:
            now = time.perf_counter() - t0
            while pending and pending[0].req.arrival <= now:
                queue.append(pending.pop(0))
            while queue and len(active) < self.slots:
                need = dataset_bytes(queue[0])
                if active and used + need > self.budget:
                    break
                p = queue.pop(0)
                ta = time.perf_counter()
                a = self._admit(p)
                self.host["admit"] += time.perf_counter() - ta
                a.t_admit = time.perf_counter() - t0
                used += a.nbytes
                active.append(a)
                per[p.req.rid] = dict(arrival=p.req.arrival, admitted=round(a.t_admit, 4), ring=p.ring,
                                      reflections=int(p.hp["refl"]["x"].size), frames=p.req.spec.nframes,
                                      frame_shape=[p.req.spec.ny, p.req.spec.nx], hbm_gb=round(a.nbytes / 1e9, 3))
            progressed = False
            for a in active:
                if not self._finished(a) and self._in_flight(a) < self.lookahead:
                    if first_gpu is None:
                        first_gpu = time.perf_counter() - t0
                    self._step(a)
                    progressed = True
            for a in list(active):
                if self._finished(a) and (not a.ev_blk or a.ev_blk[-1].query()) and a.ev_res[-1].query():
                    tr = time.perf_counter()      # the dataset's stream has drained (its last events completed)
                    t_end = time.perf_counter() - t0
                    ls = self.latencies(a)
                    lat += list(ls.values())
                    rid = a.plan.req.rid
                    n_r = int(np.isfinite(a.rb.o2["iprf"].cpu().numpy()).sum())
                    n_refl += n_r
                    per[rid].update(finished=round(t_end, 4), integrated=n_r, batches=len(a.ev_res),
                                    p50_ms=round(float(np.percentile(list(ls.values()), 50)), 3) if ls else None)
                    if keep_results:
                        results[rid] = self.results(a)
                    active.remove(a)
                    used -= a.nbytes
                    self._free.append(a.stream)
                    self.host["retire"] += time.perf_counter() - tr
                    done.append(rid)
                    del a
            ti = time.perf_counter()
            if not progressed and (active or queue):
                time.sleep(0.0002)
            elif not active and not queue and pending:
                time.sleep(max(0.0, min(0.01, pending[0].req.arrival - (time.perf_counter() - t0))))
            self.host["idle"] += time.perf_counter() - ti
        wall = time.perf_counter() - t0
        lat = np.asarray(lat, float)
        rep = dict(datasets=len(done), reflections=n_refl, wall_s=round(wall, 4),
                   reflections_per_s=round(n_refl / wall, 1) if wall > 0 else None,
                   batches_with_results=int(lat.size),
                   p50_ms=round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                   p99_ms=round(float(np.percentile(lat, 99)), 3) if lat.size else None,
                   max_ms=round(float(lat.max()), 3) if lat.size else None,
                   host_s={k: round(v, 3) for k, v in self.host.items()}, per_dataset=per)
        if keep_results:
            rep["results"] = results
        return rep


def serving_requests(n=32, seed=1005, frame_scale=1.0, min_frames=20, sizes=None):
    """The serving workload's datasets: frame size from {1024 x 1024, 2070 x 2167, 4148 x 4362} uniformly, frame
    count uniform in 50 .. 3600 (times frame_scale, at least min_frames), cell axes uniform in 30 .. 150 A in a
    random orthorhombic (P 21 21 21) or tetragonal (P 43 21 2) setting; geometry and physics of the full-size
    stream otherwise; seed + dataset id per dataset."""
    sizes = sizes or [(1024, 1024), (2167, 2070), (4362, 4148)]
    out = []
    for i in range(n):
        rng = np.random.default_rng(seed + i)
        ny, nx = sizes[int(rng.integers(len(sizes)))]
        nfr = max(min_frames, int(round(int(rng.integers(50, 3601)) * frame_scale)))
        a, b, c = rng.uniform(30.0, 150.0, 3)
        if rng.random() < 0.5:
            cell, sg = (a, b, c, 90.0, 90.0, 90.0), "P 21 21 21"
        else:
            cell, sg = (a, a, c, 90.0, 90.0, 90.0), "P 43 21 2"
        spec = synth.replace(synth.spec_full(), name=f"serve{i}_{nx}x{ny}x{nfr}", nx=nx, ny=ny, nframes=nfr,
                             cell=cell, space_group=sg, seed=seed + i)
        out.append(Request(spec, 0.0, i))
    return out


def poisson_arrivals(reqs, rate, seed=1005):
    """Arrival times of a Poisson process of ``rate`` requests per second, in request order."""
    rng = np.random.default_rng(seed + 7919)
    t = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    t -= t[0]
    for r, ti in zip(reqs, t):
        r.arrival = float(ti)
    return reqs
