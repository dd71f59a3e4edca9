"""Colouring on the device against the oracle's closed form, every solution.

Colours are integers and must match bit for bit.  The closed form is itself
pinned against a round-by-round restatement of the reference kernel in
tests/test_oracle.py.
"""
import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph, symmetrize

pytestmark = pytest.mark.gpu

SOLS = [smdp.SOL_FLAT, smdp.SOL_NAIVE, smdp.SOL_WAVE, smdp.SOL_BLOCK, smdp.SOL_GRID, smdp.SOL_TUNED]


def _graphs():
    out = {}
    out["rmat12_sym"] = smdp.gen_rmat(12, 8, symmetrize=True, seed=5)
    out["rmat11_sym_noperm"] = smdp.gen_rmat(11, 16, symmetrize=True, permute=False, seed=5)
    row, col, _ = random_graph(4000, 20000, 8)
    out["random_directed"] = smdp.HostGraph(row, col)
    out["uniform"] = smdp.gen_uniform(8192, 8, 1, 1, 3)
    return out


GRAPHS = _graphs()


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("use_prio", [False, True])
@pytest.mark.parametrize("sol", SOLS)
def test_color_parity(ctx, name, use_prio, sol):
    g = GRAPHS[name]
    print(f"color {name} prio={use_prio} sol={sol}", flush=True)
    prio = smdp.gen_prio(g.n, 5) if use_prio else None
    ref, k = orc.color_closed(g.row, g.col, prio)
    c, st = ctx.color(g, prio, sol)
    assert np.array_equal(c, ref)
    assert st.rounds == k
    if sol == smdp.SOL_TUNED:
        # layering queues every vertex of colour >= 2 once, when its last preceding neighbour is coloured
        assert st.q_push[0] == int((ref >= 2).sum())


@pytest.mark.parametrize("sol", SOLS)
def test_color_known_answer(ctx, golden, data_test_path, sol):
    g = smdp.read_graph(data_test_path, smdp.FMT_DIMACS9)
    c, st = ctx.color(g, None, sol)
    assert c.tolist() == golden["color"]


def test_color_proper_and_ties(ctx):
    g = GRAPHS["rmat12_sym"]
    prio = (smdp.gen_prio(g.n, 7) % 3).astype(np.uint32)  # heavy ties: broken by id
    for sol in SOLS:
        c, _ = ctx.color(g, prio, sol)
        assert np.array_equal(c, orc.color_closed(g.row, g.col, prio)[0])
        src = np.repeat(np.arange(g.n), np.diff(g.row))
        loop = src == g.col
        assert not np.any(c[src[~loop]] == c[g.col[~loop]])


def test_color_tiny_thresholds(ctx):
    g = GRAPHS["rmat11_sym_noperm"]
    prio = smdp.gen_prio(g.n, 9)
    ref, _ = orc.color_closed(g.row, g.col, prio)
    old = ctx.get_params(smdp.APP_COLOR)
    try:
        ctx.set_params(smdp.APP_COLOR, heavy_threshold=2, split_threshold=8, chunk=64, wg_per_cu=1)
        for sol in SOLS:
            c, st = ctx.color(g, prio, sol)
            assert np.array_equal(c, ref), sol
    finally:
        ctx.set_params(smdp.APP_COLOR, **old.as_dict())


def test_color_edge_cases(ctx):
    cases = [smdp.HostGraph([0, 0], []), smdp.HostGraph([0, 1, 2], [0, 1]),
             smdp.HostGraph([0, 0, 1, 2, 3], [0, 1, 2])]
    for g in cases:
        ref, _ = orc.color_closed(g.row, g.col)
        for sol in SOLS:
            c, _ = ctx.color(g, None, sol)
            assert np.array_equal(c, ref)


@pytest.mark.parametrize("use_prio", [False, True])
def test_color_layers_long_and_short(ctx, use_prio):
    """RMAT-15 symmetrised: the first layers exceed 2048 vertices (edge-balanced tiles with
    grid chunks for hubs), the later ones are expanded arc-balanced without chunks."""
    g = smdp.gen_rmat(15, 8, symmetrize=True, seed=9)
    prio = smdp.gen_prio(g.n, 9) if use_prio else None
    ref, k = orc.color_closed(g.row, g.col, prio)
    assert int((ref == 1).sum()) > 2048 and int(np.bincount(ref).max()) > 2048
    c, st = ctx.color(g, prio, smdp.SOL_TUNED)
    assert np.array_equal(c, ref)
    assert st.rounds == k
    assert st.q_push[0] == int((ref >= 2).sum())


def test_color_new_priorities_on_a_resident_graph(ctx):
    """The tuned solution keeps each arc's priority beside the arc, built once per
    priority upload: every new upload must be seen, with id order in between."""
    g = smdp.gen_rmat(13, 8, symmetrize=True, seed=7)
    with smdp.Graph(ctx, g) as dg:
        for seed in (1, 2, 1):
            prio = smdp.gen_prio(g.n, seed)
            dg.set_prio(prio)
            for use_prio in (True, False, True):
                dg.color(use_prio, smdp.SOL_TUNED)
                ref, _ = orc.color_closed(g.row, g.col, prio if use_prio else None)
                assert np.array_equal(dg.read(smdp.RES_COLOR), ref), (seed, use_prio)


def test_color_short_long_short_layers(ctx):
    """Vertex-id priority on a symmetric graph whose layers go short (a path of
    300 vertices, one vertex per layer), long (3000 vertices hanging off the
    path's end: one layer above the 2048-vertex limit of the short-layer run),
    and short again (a second path after them): the short-layer launch must hand
    the long layer back to the tile kernels and resume after it."""
    p1, fan, p2 = 300, 3000, 200
    n = p1 + fan + p2
    src, dst = [], []
    src += list(range(p1 - 1)); dst += list(range(1, p1))            # first path
    src += [p1 - 1] * fan; dst += list(range(p1, p1 + fan))           # the fan
    src += list(range(p1, p1 + fan)); dst += [p1 + fan] * fan         # all join the second path
    src += list(range(p1 + fan, n - 1)); dst += list(range(p1 + fan + 1, n))
    src, dst = np.array(src), np.array(dst)
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    srow, scol = symmetrize(np.cumsum(row).astype(np.int32), dst[order].astype(np.int32))
    g = smdp.HostGraph(srow, scol)
    ref, k = orc.color_closed(g.row, g.col, None)
    assert int(np.bincount(ref).max()) > 2048 and k > 400
    c, st = ctx.color(g, None, smdp.SOL_TUNED)
    assert np.array_equal(c, ref)
    assert st.rounds == k
    assert st.q_push[0] == int((ref >= 2).sum())
