"""Index arithmetic of the 4x4-block stack kernels (libxstream_amd/csrc/smm_mfma.hip, ep_body with B4 = 1 and 2),
checked by brute force on the CPU over every shape the runtime-shape kernels accept (lxs_launch.hip, mfma_fits):
every element of the C block is produced by exactly one (lane, instruction) pair of the flush, every element
load stays inside its block, the element a lane uses for a real row or column is that row or column, and the
k values of one 16-wide step are fed exactly once per row and column position.

The lane layouts are the gfx950 ones the kernels rely on (v_mfma_f64_4x4x4_4b: lane l feeds k = l / 16 of
block (l / 4) % 4, A row and B column l % 4, result (i, j) of block g in lane 16 i + 4 g + j;
v_mfma_f32_4x4x1_16b: lane l feeds block l / 4, result column l % 4 in registers 0..3); the GPU parity tests in
tests/test_gpu_ep.py check them end to end.
"""
import itertools

import pytest


def lane_rows(first, t, extent):
    """the t elements a lane uses from `first` on: element pairs start at min(row, extent - 2) and a pair whose
    first row is the last row of the block holds it in its second element; an odd t adds a single element at
    min(row, extent - 1). Returns (element index used for each row, every index read)."""
    used, read = [], []
    for p in range(t // 2):
        s = min(first + 2 * p, extent - 2)
        read += [s, s + 1]
        x = s + 1 if first + 2 * p == extent - 1 else s
        used += [x, s + 1]
    if t % 2:
        s = min(first + t - 1, extent - 1)
        read.append(s)
        used.append(s)
    return used, read


def check_loads(first, t, extent):
    used, read = lane_rows(first, t, extent)
    assert all(0 <= x < extent for x in read), (first, t, extent, read)
    for r in range(t):  # a real row or column must come from its own element
        if first + r < extent:
            assert used[r] == first + r, (first, t, extent, r, used)


@pytest.mark.parametrize("ti,tj", list(itertools.product(range(1, 5), range(1, 5))))
def test_quadrant_blocks(ti, tj):
    """B4 = 1 (variant 3): C rows 4 TI (g % 2) + TI i + r, columns 4 TJ (g / 2) + TJ j + c, for every
    m in (8 (TI - 1), 8 TI] and n in (8 (TJ - 1), 8 TJ]"""
    for m in range(8 * (ti - 1) + 1, 8 * ti + 1):
        for n in range(8 * (tj - 1) + 1, 8 * tj + 1):
            seen = {}
            for lane in range(64):
                g = (lane >> 2) & 3
                r4 = 4 * ti * (g & 1) + ti * (lane & 3)
                c4 = 4 * tj * (g >> 1) + tj * (lane & 3)
                check_loads(r4, ti, m)
                check_loads(c4, tj, n)
                # flush: lane holds (i, j) = (lane / 16, lane % 4) of block g
                rb = 4 * ti * (g & 1) + ti * (lane >> 4)
                cb = 4 * tj * (g >> 1) + tj * (lane & 3)
                for r in range(ti):
                    for c in range(tj):
                        i, j = rb + r, cb + c
                        if i < m and j < n:
                            assert (i, j) not in seen
                            seen[(i, j)] = (lane, r, c)
            assert len(seen) == m * n, (m, n)


def test_quadrant_block_operands_match_results():
    """the A row a lane feeds for block g and the result row of that block: a lane with A row i of block g
    computes C row 4 TI (g % 2) + TI i + r, the same row the flush writes for result row i of block g"""
    ti = 3
    for lane in range(64):
        g, i = (lane >> 2) & 3, lane & 3
        a_row = 4 * ti * (g & 1) + ti * i
        out_lane = 16 * i + 4 * g  # where result row i of block g sits (column 0)
        assert 4 * ti * (((out_lane >> 2) & 3) & 1) + ti * (out_lane >> 4) == a_row


@pytest.mark.parametrize("ti,tj", list(itertools.product(range(1, 5), range(1, 5))))
@pytest.mark.parametrize("fp32", [False, True], ids=["f64", "f32"])
def test_k_split_blocks(ti, tj, fp32):
    """B4 = 2 (variant 4): C rows TI i + r and columns TJ j + c with i = j = l % 4; the sixteen k of a step are
    fed once for every row (and column) position; the flush covers the block exactly once"""
    for lane_pos in range(4):  # lanes with A row position i = l % 4
        ks = sorted((4 * ((l >> 2) & 3) + (l >> 4)) if not fp32 else (l >> 2) for l in range(64) if l % 4 == lane_pos)
        assert ks == list(range(16)), ks
    for m in range(4 * (ti - 1) + 1, 4 * ti + 1):
        for n in range(4 * (tj - 1) + 1, 4 * tj + 1):
            for lane in range(64):
                check_loads(ti * (lane & 3), ti, m)
                check_loads(tj * (lane & 3), tj, n)
            seen = set()
            for t in range(ti * tj):  # slot t = (r, c), element x = 4 i + j after the fold
                r, c = divmod(t, tj)
                for x in range(16):
                    i, j = ti * (x >> 2) + r, tj * (x & 3) + c
                    if i < m and j < n:
                        assert (i, j) not in seen
                        seen.add((i, j))
            assert len(seen) == m * n, (m, n)
