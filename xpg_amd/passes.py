"""HBM plan of an offline sweep, and the streamed form for sweeps larger than it.

The reference bounds its host memory by splitting the sweep into passes of phi blocks whose shoeboxes fit
(``_plan_passes`` / ``_split_blocks``, xpg/shoebox.py:134-196, with the block queue of xpg/integrate.py:588-646).
Here the frames themselves live in HBM, so the question is whether the sweep's frames and the prepared data of the
batched form fit the budget:

  resident  every frame uploaded once (two pinned buffers, frames.from_imageset), then the batched form
            (Sweep.run: every block's shoeboxes gathered ahead, the adaptive chain, pass 2 overlapped);
  streamed  the frames pass through a ring of R frames on the device, R = the most frames any block needs when it
            runs (ring_frames over the blocks in chain order); batches of frames are uploaded through two
            pinned buffers, and each block runs (Sweep.enqueue_block: gather into a one-block arena, pass 1, fits,
            normalisation, pass 2) as soon as its last frame has arrived.  HBM: the ring, the arena of the largest
            block, the table and outputs -- independent of the number of frames.

Both forms run the same kernels on the same values in the same chain order, so their results are bit-identical.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import _lib as L
from .frames import DeviceFrames, FrameUploader


@dataclass
class MemoryPlan:
    streamed: bool
    batch: int                    # frames per upload
    ring: int                     # frame slots of the ring (streamed)
    ready: list = field(default_factory=list)   # per block in chain order: the batch after which it runs
    frame_bytes: int = 0
    resident_bytes: int = 0       # frames + prepared data of the batched form
    streamed_bytes: int = 0       # ring + one block's arena
    budget: int = 0


def ring_frames(order, need_lo, need_hi, batch, nframes):
    """Frames a ring must hold, and each block's upload batch: block i (chain order) runs after the batch holding its
    last needed frame and after the blocks before it; when it runs after batch j the frames uploaded so far are
    [0, (j + 1) batch) and it still needs need_lo[i] onwards."""
    ready, last = [], -1
    ring = batch
    for b in order:
        j = max(last, min(need_hi[b], nframes - 1) // batch)
        last = j
        ready.append(j)
        ring = max(ring, (j + 1) * batch - max(need_lo[b], 0))
    ring = -(-ring // batch) * batch
    return min(ring, -(-nframes // batch) * batch), ready


def plan_memory(hp, sp, frames, ny, nx, budget, batch=10, frame_offset=0):
    """MemoryPlan of a sweep: ``hp`` the host plan, ``sp`` its SweepPlan, ``frames`` (first, last), ``budget`` HBM
    bytes for frames and prepared data (the table and outputs come on top).  ``frame_offset``: frame index of the
    first frame in hp's frame coordinates (need_lo / need_hi count from frame 1 at 0)."""
    f0, f1 = frames
    n = f1 - f0 + 1
    fb = 4 * ny * nx
    # prepared data per row: values and neighbour bits (5 B per voxel), eps1, eps2, J and the partner's (48 B per
    # pixel), zeta and the partner's, and the rocking-curve table (8 B per frame of the deepest box of the launch)
    prep = sum(5 * int(v) + 48 * int(p) + 16 * int(m) + 8 * int(m) * int(mx[0])
               for (v, p), (_, _, _, m, mx, _) in zip(sp.bin_sizes, sp.bins))
    nV, nP, nR = sp.block_arena
    arena = 5 * nV + 48 * nP + 16 * nR + 8 * nR * max((int(b[4][0]) for b in sp.bins), default=1)
    order = [b for b in hp["block_order"] if hp["blocks_refl"][b].size]
    lo = [hp["need_lo"][b] - frame_offset for b in range(len(hp["need_lo"]))]
    hi = [hp["need_hi"][b] - frame_offset for b in range(len(hp["need_hi"]))]
    ring, ready = ring_frames(order, lo, hi, batch, n)
    mp = MemoryPlan(streamed=False, batch=batch, ring=ring, ready=ready, frame_bytes=fb, resident_bytes=n * fb + prep,
                    streamed_bytes=ring * fb + arena, budget=int(budget))
    if mp.resident_bytes > budget:
        if mp.streamed_bytes > budget:
            raise L.XpgError(f"HBM budget {budget / 1e9:.2f} GB too small: the streamed sweep needs "
                             f"{mp.streamed_bytes / 1e9:.2f} GB ({ring} frames of {fb / 1e6:.1f} MB and a block's "
                             f"prepared data)")
        mp.streamed = True
    return mp


def run_streamed(sweep, images, mp: MemoryPlan, frames):
    """The streamed form of ``sweep`` (built over a ring store, stream_store) fed from the host ImageSet ``images``:
    per batch, the frames to the ring through two pinned buffers, then every block the batch completes.  Returns the
    run record (as Sweep.run)."""
    import time
    f0, f1 = frames
    n = f1 - f0 + 1
    dev = sweep.device
    t0 = time.time()
    st = torch.cuda.current_stream(dev)
    arena = sweep.arena()
    from .integrate import RunBufs
    rb = RunBufs(sweep)
    sweep._init_widths()
    ring = sweep.frames_src.data
    up = FrameUploader(images, dev, mp.batch)
    nb = -(-n // mp.batch)
    i = 0
    for k in range(nb):
        a = k * mp.batch
        nf = min(mp.batch, n - a)
        slot = a % mp.ring
        up.upload(f0 + a, nf, ring[slot:slot + nf])       # in stream order behind the blocks that read the slots
        while i < len(mp.ready) and mp.ready[i] <= k:
            sweep.enqueue_block(i, sweep.store, arena, rb)
            i += 1
    st.synchronize()
    out = rb.result()
    out["t_gpu"] = time.time() - t0
    return out


def stream_store(images, mp: MemoryPlan, frames, device):
    """A ring DeviceFrames of mp.ring frame slots for the value-coded float32 frames of ``images``."""
    f0, f1 = frames
    ring = torch.empty((mp.ring, images.ny, images.nx), dtype=torch.float32, device=device)
    return DeviceFrames(ring, first=f0, n_frames=f1 - f0 + 1)
