"""The pass planner of xpg_amd.shoebox against the reference's own tests (tests/test_new_components.py:216-269 of the
reference, restated: same inputs, same assertions) -- CPU only."""
import numpy as np

from xpg_amd.shoebox import KERNEL_RESERVE, SB_OFF, _plan_passes, _split_blocks


def test_sb_off():
    assert SB_OFF == np.float32(3.0)


def test_integration_passes_fit_memory_budget():
    n_lat, n_blk, n_rows, nf = 2, 10, 100, 100
    blocks_refl, lattice_of, need_hi, t_row = [], [], [], []
    for li in range(n_lat):
        for j in range(n_blk):
            r0 = len(t_row)
            t_row += list(10 * j + np.arange(n_rows) % 10)
            blocks_refl.append(np.arange(r0, r0 + n_rows))
            lattice_of.append(li)
            need_hi.append(min(10 * j + 50, nf - 1))
    t_row = np.array(t_row)
    w_row = np.full(t_row.size, 2.0 ** 20)
    order = sorted(range(len(blocks_refl)), key=lambda b: (10 * (b % n_blk), b))
    plan = lambda budget, ord_=order: _plan_passes(ord_, blocks_refl, lattice_of, need_hi, t_row, w_row, nf, budget)
    passes, info = plan(None)
    assert passes == [order] and info["peak"] == info["one"]
    one = info["one"] - KERNEL_RESERVE
    assert 0.9e9 < one < 1.4e9
    passes, info = plan(KERNEL_RESERVE + 0.7 * one)
    assert len(passes) == 2 and all(len({lattice_of[b] for b in p}) == 1 for p in passes)
    assert sorted(sum(passes, [])) == sorted(order) and all(p == [b for b in order if b in p] for p in passes)
    assert info["peak"] <= KERNEL_RESERVE + 0.7 * one
    lat0 = [b for b in order if lattice_of[b] == 0]
    passes, info = plan(KERNEL_RESERVE + 0.4 * one, lat0)
    assert len(passes) >= 2 and info["peak"] <= KERNEL_RESERVE + 0.4 * one and sorted(sum(passes, [])) == sorted(lat0)


def test_block_parts_fit_memory_budget():
    n_blk, n_rows, nf = 6, 200, 120
    t_row = np.repeat(20 * np.arange(n_blk), n_rows) + np.tile(np.arange(n_rows) % 20, n_blk)
    t_end = np.minimum(t_row + 10, nf - 1)
    w_row = np.full(t_row.size, 2.0 ** 20)
    blocks_refl = [np.arange(b * n_rows, (b + 1) * n_rows) for b in range(n_blk)]
    budget = KERNEL_RESERVE + 1.5 * n_rows * 2.0 ** 20

    def plan(blocks, par):
        hi = [int(t_end[r].max()) for r in blocks]
        order = sorted(range(len(blocks)), key=lambda v: int(t_row[blocks[v]].min()))
        return _plan_passes(order, blocks, [0] * len(par), hi, t_row, w_row, nf, budget)[1]["peak"]
    assert plan(blocks_refl, range(n_blk)) > budget
    fits = []
    for k in range(2, 9):
        parts, par = _split_blocks(blocks_refl, t_row + 0.5, w_row, k)
        assert np.array_equal(np.sort(np.concatenate(parts)), np.arange(t_row.size))
        assert all(np.isin(p_, blocks_refl[b]).all() for p_, b in zip(parts, par))
        assert max(w_row[p_].sum() for p_ in parts) <= n_rows * 2.0 ** 20 / k + 2.0 ** 20
        fits.append(plan(parts, par) <= budget)
    assert any(fits)
