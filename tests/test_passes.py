"""The HBM plan of an offline sweep (xpg_amd/passes.py): the ring of frames a streamed sweep needs and the choice
between the resident and the streamed form, on hand-made plans."""
from types import SimpleNamespace

import numpy as np
import pytest

from xpg_amd import passes


def test_ring_frames_hand_case():
    # three blocks in chain order; batches of 10 frames
    need_lo, need_hi = [0, 18, 37], [24, 45, 59]
    ring, ready = passes.ring_frames([0, 1, 2], need_lo, need_hi, 10, 60)
    assert ready == [2, 4, 5]                       # the batch holding each block's last frame
    # block 1 runs after batch 4 (frames 0..49 written) and still needs frame 18: 32 frames -> 40 (whole batches)
    assert ring == 40
    # a block that runs late (after its successor's last frame) widens the ring
    ring2, ready2 = passes.ring_frames([1, 0, 2], need_lo, need_hi, 10, 60)
    assert ready2 == [4, 4, 5] and ring2 == 50      # block 0 runs after batch 4 and still needs frame 0
    # never more than the sweep
    assert passes.ring_frames([0], [0], [5], 10, 7) == (10, [0])


def _plan(n_blocks, rows_per_block, V, P, frames):
    bl = [np.arange(i * rows_per_block, (i + 1) * rows_per_block) for i in range(n_blocks)]
    per = frames // n_blocks
    hp = dict(block_order=list(range(n_blocks)), blocks_refl=bl, need_lo=[max(0, i * per - 3) for i in range(n_blocks)],
              need_hi=[min(frames - 1, (i + 1) * per + 3) for i in range(n_blocks)])
    sp = SimpleNamespace(bins=[(i, 0, i * rows_per_block, rows_per_block, (7, 11), 0) for i in range(n_blocks)],
                         bin_sizes=[(rows_per_block * V, rows_per_block * P)] * n_blocks,
                         block_arena=(rows_per_block * V, rows_per_block * P, rows_per_block))
    return hp, sp


def test_plan_memory_chooses_the_form_from_the_budget():
    hp, sp = _plan(24, 1000, 847, 121, 1800)
    fb = 4 * 100 * 100
    mp = passes.plan_memory(hp, sp, (1, 1800), 100, 100, 1 << 40)
    row = 5 * 847 + 48 * 121 + 16 + 8 * 7               # prepared bytes per row of a 7 x 11 x 11 box
    assert not mp.streamed and mp.resident_bytes == 1800 * fb + 24 * 1000 * row
    assert mp.ring < 1800 and mp.streamed_bytes == mp.ring * fb + 1000 * row
    mp2 = passes.plan_memory(hp, sp, (1, 1800), 100, 100, mp.resident_bytes - 1)
    assert mp2.streamed and mp2.ring == mp.ring and len(mp2.ready) == 24
    with pytest.raises(Exception, match="budget"):
        passes.plan_memory(hp, sp, (1, 1800), 100, 100, mp.streamed_bytes - 1)
