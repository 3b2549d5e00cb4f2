"""Launch-shape invariants of the block-Jacobi CG kernels (k_cg.hip), replayed
on the CPU for the shapes Solver::setup_cg builds: the matrix-product table
(workgroup -> (tile row, run), runs of up to tpw tiles listed row by row, and
each tile row's first run) covers every lower tile exactly once, and the
partial-sum terms k_cg_vector adds for each row -- runs w, w + 4, ... of its
tile row (ceil((nc - w) / 4) of them) and the column partials of the tiles
(J, I), J = I + w, I + w + 4, ... -- are each of the row's partials exactly
once, with every index inside the buffers the solver allocates."""
import pytest


def tables(nblk, tpw):  # as Solver::setup_cg builds wgtab and rowstart
    tab, rowstart = [], []
    for i in range(nblk):
        rowstart.append(len(tab))
        tab.extend((i, c) for c in range((i + tpw) // tpw))
    rowstart.append(len(tab))
    return tab, rowstart


@pytest.mark.parametrize("C", [1, 2, 7, 8, 49, 50, 120, 356, 1936])
@pytest.mark.parametrize("tpw", [1, 2, 3, 8])
def test_cg_launch_shapes(C, tpw):
    npad = (9 * C + 63) // 64 * 64
    nblk = npad // 64
    tab, rowstart = tables(nblk, tpw)
    nwg_m = len(tab)
    seen = []
    for i, c in tab:  # k_cg_matvec: j0 = c tpw, j1 = min(j0 + tpw, i + 1)
        j0 = c * tpw
        j1 = min(j0 + tpw, i + 1)
        assert j0 < j1
        seen.extend((i, j) for j in range(j0, j1))
    assert sorted(seen) == [(i, j) for i in range(nblk) for j in range(i + 1)]
    n_cp = nblk * (nblk + 1) // 2 * 64
    nwg_v = (npad + 62) // 63  # cg_vector_groups: 63 rows (seven cameras) each
    for blk in range(nwg_v):
        for lane in range(63):
            row = 63 * blk + lane
            if row >= npad:
                continue
            I, rr = row >> 6, row & 63
            nc = (I + tpw) // tpw
            assert nc == rowstart[I + 1] - rowstart[I]
            runs, cols = [], []
            for w in range(4):
                for q in range((nc - w + 3) // 4 if nc > w else 0):
                    runs.append(w + 4 * q)
                    assert (rowstart[I] + w + 4 * q) * 64 + rr < nwg_m * 64
                for J in range(I + w, nblk, 4):
                    cols.append(J)
                    assert (J * (J + 1) // 2 + I) * 64 + rr < n_cp
            assert sorted(runs) == list(range(nc))
            assert sorted(cols) == list(range(I, nblk))
            if row < 9 * C:  # the camera's nine rows lie inside the workgroup
                assert 9 * (lane // 9) + 8 < 63 and (63 * blk) // 9 + lane // 9 == row // 9


@pytest.mark.parametrize("C", [1, 2, 49, 120, 356, 1936])
def test_cg_gauge_launch_shapes(C):
    """The deflation set-up kernels (k_cg_gauge, k_cg_gauge_product,
    k_cg_gauge_rows, k_cg_gauge_finish) and the deflated start read and
    write inside the buffers Solver::setup_cg carves out of one allocation,
    and the runs of k_cg_gauge_product cover every tile of every tile row
    exactly once (runs of kGaugeRun = 4 tiles, grid (runs, nblk))."""
    npad = (9 * C + 63) // 64 * 64
    nblk = npad // 64
    nwg_v = (npad + 62) // 63
    nruns = (nblk + 3) // 4
    # carve-up: zraw, zd, azd (npad x 8 each), npart, zbpart (nwg_v x 8 each),
    # epart (nblk x 49), einv (64), apart (nblk x nruns x 64 x 8)
    off = {"zraw": 0, "zd": 8 * npad, "azd": 16 * npad, "npart": 24 * npad, "zbpart": 24 * npad + 8 * nwg_v,
           "epart": 24 * npad + 16 * nwg_v}
    off["einv"] = off["epart"] + 49 * nblk
    off["apart"] = off["einv"] + 64
    total = 3 * 8 * npad + 2 * 8 * nwg_v + 49 * nblk + 64 + nblk * nruns * 64 * 8
    size = {"zraw": 8 * npad, "zd": 8 * npad, "azd": 8 * npad, "npart": 8 * nwg_v, "zbpart": 8 * nwg_v,
            "epart": 49 * nblk, "einv": 64, "apart": nblk * nruns * 512}
    ends = sorted((off[k], off[k] + size[k]) for k in off)
    assert ends[0][0] == 0 and ends[-1][1] == total
    assert all(a[1] == b[0] for a, b in zip(ends, ends[1:]))  # contiguous, no overlap
    # k_cg_gauge / start partials: rows 63 blk + lane < npad, records 8 blk + j
    assert 63 * (nwg_v - 1) < npad <= 63 * nwg_v
    assert 8 * (nwg_v - 1) + 7 < size["zbpart"]
    # the matvec's b.b at k = 0 reads zbpart[7 + 8 q], q < nwg_v
    assert 7 + 8 * (nwg_v - 1) < size["zbpart"]
    # k_cg_gauge_product: tiles J0 = 4 run .. min(J0 + 4, nblk) - 1, zraw rows of tile J
    for I in range(nblk):
        seen = []
        for run in range(nruns):
            J0 = 4 * run
            J1 = min(J0 + 4, nblk)
            assert J0 < J1
            seen.extend(range(J0, J1))
            for J in range(J0, J1):
                assert 8 * J * 64 + 255 + 256 < size["zraw"]
            assert (I * nruns + run) * 512 + 8 * 63 + 7 < size["apart"]
        assert seen == list(range(nblk))
    # gauge_finish: epart[49 I + t], einv[49]
    assert 49 * (nblk - 1) + 48 < size["epart"] and 49 <= size["einv"]
