"""Randomized parity sweep on small graphs: every SSSP and BFS solution and
the tuned colouring, exact against the oracle, over seeded graphs with the
edge cases of the domain mixed in (empty rows, self-loops, duplicate arcs,
zero and large weights, unreached parts, heavy vertices above the split
degree).  Seeds are fixed, so the sweep is the same on every run."""
import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph, symmetrize

pytestmark = pytest.mark.gpu

SSSP_SOLS = [smdp.SOL_FLAT, smdp.SOL_NAIVE, smdp.SOL_WAVE, smdp.SOL_BLOCK, smdp.SOL_GRID, smdp.SOL_TUNED]
BFS_SOLS = SSSP_SOLS + [smdp.SOL_HIER]


def _case(seed):
    """A small graph of seeded shape: sizes from 1 to ~3000 vertices, sparse
    to dense, with a hub of extra arcs, weights from [0, 1] up to [1, 2^16]."""
    rng = np.random.default_rng(1000 + seed)
    n = int(rng.choice([1, 2, 7, 64, 300, 1500, 3000]))
    m = int(rng.integers(0, 12 * n + 1))
    wlo, whi = [(0, 1), (1, 100), (1, 1 << 16)][seed % 3]  # (n - 1) * 2^16 stays below the 2^30 sentinel
    row, col, w = random_graph(n, m, seed, wlo=wlo, whi=whi)
    if n > 8 and seed % 2 == 0:
        # a hub: vertex h gets 4n extra arcs (above every split degree used below)
        h = int(rng.integers(0, n))
        extra = int(4 * n)
        src = np.repeat(np.arange(n), np.diff(row)).tolist() + [h] * extra
        dst = col.tolist() + rng.integers(0, n, extra).tolist()
        ww = w.tolist() + rng.integers(wlo, whi + 1, extra).tolist()
        order = np.argsort(np.asarray(src), kind="stable")
        src = np.asarray(src)[order]
        col = np.asarray(dst, np.int32)[order]
        w = np.asarray(ww, np.int32)[order]
        row = np.zeros(n + 1, np.int32)
        np.add.at(row, src + 1, 1)
        row = np.cumsum(row).astype(np.int32)
    src_v = int(rng.integers(0, n))
    return row, col, w, src_v


@pytest.mark.parametrize("seed", range(24))
def test_random_sssp_bfs_all_solutions(ctx, seed):
    row, col, w, s = _case(seed)
    print(f"sweep seed {seed}: n={len(row) - 1} m={len(col)} src={s}", flush=True)
    g = smdp.HostGraph(row, col, w, s)
    dref = orc.dijkstra(row, col, w, s)
    lref = orc.bfs(row, col, s)
    old_s, old_b = ctx.get_params(smdp.APP_SSSP), ctx.get_params(smdp.APP_BFS)
    try:
        for small in (False, True):
            if small:  # tiny thresholds push most vertices through the queues and the grid chunks
                for app in (smdp.APP_SSSP, smdp.APP_BFS):
                    ctx.set_params(app, heavy_threshold=2, split_threshold=4, chunk=64)
            for sol in SSSP_SOLS:
                d, _ = ctx.sssp(g, s, sol)
                assert np.array_equal(d, dref), (seed, small, sol)
            for sol in BFS_SOLS:
                lv, _ = ctx.bfs(g, s, sol)
                assert np.array_equal(lv, lref), (seed, small, sol)
    finally:
        ctx.set_params(smdp.APP_SSSP, **old_s.as_dict())
        ctx.set_params(smdp.APP_BFS, **old_b.as_dict())


@pytest.mark.parametrize("seed", range(12))
def test_random_color_tuned(ctx, seed):
    row, col, _, _ = _case(seed)
    srow, scol = symmetrize(row, col)
    g = smdp.HostGraph(srow, scol)
    print(f"sweep colour seed {seed}: n={len(srow) - 1} m={len(scol)}", flush=True)
    for prio in (None, smdp.gen_prio(len(srow) - 1, seed)):
        ref, _ = orc.color_closed(srow, scol, prio)
        c, _ = ctx.color(g, prio, smdp.SOL_TUNED)
        assert np.array_equal(c, ref), (seed, prio is None)
