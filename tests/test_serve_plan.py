"""Host side of the serving loop: the frame ring and block readiness of a dataset plan (CPU only)."""
import numpy as np

from xpg_amd import serve, synth


def _small(seed, nframes=83, nx=1024, ny=1024, cell=(61.0, 47.0, 88.0, 90.0, 90.0, 90.0)):
    spec = synth.replace(synth.spec_full(), name=f"t{seed}", nx=nx, ny=ny, nframes=nframes, cell=cell,
                         space_group="P 21 21 21", seed=seed)
    return serve.Request(spec, 0.0, seed)


def test_ring_holds_every_block_when_it_runs():
    for seed, nfr in ((11, 83), (12, 50), (13, 147)):
        p = serve.plan_request(_small(seed, nfr), batch_frames=10)
        hp = p.hp
        assert p.ring % 10 == 0 and p.ring >= 10
        last = -1
        for b, j in zip(p.order, p.ready_batch):
            assert j >= last                                   # blocks run in the chain's order
            last = j
            written = min((j + 1) * 10, nfr)
            lo, hi = max(hp["need_lo"][b], 0), min(hp["need_hi"][b], nfr - 1)
            assert hi < written                                # every frame the block reads has arrived ...
            assert lo >= (j + 1) * 10 - p.ring                 # ... and is still in the ring
            assert j >= hi // 10
        # every reflection completes in a batch no later than its block runs
        blk = np.empty(hp["refl"]["x"].size, int)
        for i, b in enumerate(p.order):
            blk[hp["blocks_refl"][b]] = i
        ready = np.asarray(p.ready_batch)[blk]
        assert (p.done_batch <= ready).all()


def test_ring_frames_closed_form():
    # three 50-frame blocks whose boxes reach 3 frames past their edges, batches of 10
    ring, ready = serve.ring_frames([0, 1, 2], [-3, 47, 97], [52, 102, 149], 10, 150)
    assert ready == [5, 10, 14]
    assert ring == 70            # block 1 runs after frame 109 is written and still needs frame 47: 110 - 47 -> 70


def test_serving_requests_follow_the_workload():
    reqs = serve.serving_requests(n=32, frame_scale=1.0)
    sizes = {(r.spec.ny, r.spec.nx) for r in reqs}
    assert sizes <= {(1024, 1024), (2167, 2070), (4362, 4148)} and len(sizes) == 3
    nf = np.array([r.spec.nframes for r in reqs])
    assert nf.min() >= 50 and nf.max() <= 3600
    for r in reqs:
        a, b, c = r.spec.cell[:3]
        assert 30 <= a <= 150 and 30 <= b <= 150 and 30 <= c <= 150
        assert (r.spec.space_group == "P 43 21 2") == (a == b)
    t = [r.arrival for r in serve.poisson_arrivals(reqs, rate=4.0)]
    assert t[0] == 0 and np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 0.25) < 0.15
