"""Parity at the configured full sizes (RMAT-22 SSSP and PageRank, RMAT-20
SpMV, symmetric RMAT-23 colouring, the 4M-node Zipf tree, batched RMAT-21 and
256 small graphs).  Exact for integer outputs, the stated tolerances for FP;
plus size-independent properties (triangle inequality and tight in-arcs for
SSSP, depth for the tree, proper colouring, rank mass)."""
import numpy as np
import pytest

import orc
import smdp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rmat22():
    return smdp.gen_rmat(22, 16, wlo=1, whi=255, seed=3)


def test_sssp_rmat22_exact_and_properties(ctx, rmat22):
    g = rmat22
    print("sssp rmat22", flush=True)
    with smdp.Graph(ctx, g) as dg:
        dg.sssp(g.source, smdp.SOL_TUNED)
        d = dg.read(smdp.RES_DIST)
        dg.sssp(g.source, smdp.SOL_GRID)
        d_grid = dg.read(smdp.RES_DIST)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    assert np.array_equal(d, ref) and np.array_equal(d_grid, ref)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    reach = d[src] != orc.INF
    alt = d[src][reach].astype(np.int64) + g.w[reach]
    assert np.all(d[g.col[reach]] <= alt)                       # no arc can still relax
    tight = np.zeros(g.n, bool)
    tight[g.col[reach][d[g.col[reach]] == alt]] = True
    reached = (d != orc.INF)
    reached[g.source] = False
    assert np.all(tight[reached])                               # every distance is realised by an arc


def test_pagerank_rmat22(ctx, rmat22):
    g = rmat22
    print("pagerank rmat22", flush=True)
    ig = smdp.invert(g)
    od = g.outdegree()
    r, st = ctx.pagerank(ig, od, smdp.pr_opts(0.85, 1.0, 1e-6, 100), smdp.SOL_TUNED)
    ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, fixed_iters=int(st.rounds))
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6
    assert st.residual < 1e-6 and abs(float(r.sum(dtype=np.float64)) - 1.0) < 1e-5


@pytest.mark.parametrize("sol,layout", [(smdp.SOL_FLAT, smdp.LAYOUT_AUTO), (smdp.SOL_GRID, smdp.LAYOUT_AUTO),
                                        (smdp.SOL_TUNED, smdp.LAYOUT_CSR), (smdp.SOL_TUNED, smdp.LAYOUT_XCD8),
                                        (smdp.SOL_TUNED, smdp.LAYOUT_BINNED)])
def test_pagerank_rmat22_solutions_and_layouts(ctx, rmat22, sol, layout):
    """Every PageRank solution and tuned layout on the full RMAT-22 graph,
    against the FP64 iterates of the same iteration count (OpenMP oracle,
    whose reductions may differ from the sequential one in the last bits)."""
    g = rmat22
    print(f"pagerank rmat22 sol {sol} layout {layout}", flush=True)
    ig = smdp.invert(g)
    od = g.outdegree()
    old = ctx.get_params(smdp.APP_PAGERANK)
    try:
        ctx.set_params(smdp.APP_PAGERANK, layout=layout)
        with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
            dg.set_outdeg(od)
            st = dg.pagerank(smdp.pr_opts(0.85, 1.0, 1e-6, 100), sol)
            r = dg.read(smdp.RES_RANK, dtype=np.float32)
    finally:
        ctx.set_params(smdp.APP_PAGERANK, **old.as_dict())
    ref, _, _ = orc.pagerank_par(ig.row, ig.col, od, int(st.rounds), d=0.85, T=1.0)
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6, sol
    assert st.residual < 1e-6


def test_spmv_rmat20(ctx):
    g = smdp.gen_rmat(20, 16, seed=2)
    print("spmv rmat20", flush=True)
    val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 2, 6)
    x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 2, 9)
    y64 = orc.spmv(g.row, g.col, val, x)
    prod = np.abs(val.astype(np.float64) * x.astype(np.float64)[g.col])
    S = np.zeros(g.n)
    np.add.at(S, np.repeat(np.arange(g.n), np.diff(g.row)), prod)
    nr = np.diff(g.row)
    bound = 2 * 2.0 ** -24 * np.abs(y64) + 2 * (nr * 2.0 ** -24) ** 2 * S
    for sol in (smdp.SOL_TUNED, smdp.SOL_GRID, smdp.SOL_FLAT):
        y, _ = ctx.spmv(g, val, x, sol)
        assert np.all(np.abs(y - y64) <= bound), sol


def test_color_rmat23_symmetric(ctx):
    g = smdp.gen_rmat(23, 8, symmetrize=True, seed=5)
    print("color rmat23", flush=True)
    prio = smdp.gen_prio(g.n, 5)
    c, st = ctx.color(g, prio, smdp.SOL_TUNED)
    ref, k = orc.color_closed(g.row, g.col, prio)
    assert np.array_equal(c, ref) and st.rounds == k
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    loop = src == g.col
    assert not np.any(c[src[~loop]] == c[g.col[~loop]])


@pytest.mark.parametrize("sol", [smdp.SOL_GRID, smdp.SOL_BLOCK])
def test_color_rmat23_round_solutions(ctx, sol):
    """The reference's round rule (graph_color_kernel.cu:28-49) with the
    grid- and block-level consolidation on the full symmetric RMAT-23 graph:
    the same colours as the closed form."""
    g = smdp.gen_rmat(23, 8, symmetrize=True, seed=5)
    print(f"color rmat23 sol {sol}", flush=True)
    prio = smdp.gen_prio(g.n, 5)
    c, _ = ctx.color(g, prio, sol)
    ref, _ = orc.color_closed(g.row, g.col, prio)
    assert np.array_equal(c, ref), sol


def test_bfs_zipf_tree_4m(ctx):
    t = smdp.gen_zipf_tree(4194304, 1.5, 4096, 5)
    print("bfs tree 4M", flush=True)
    for sol in (smdp.SOL_TUNED, smdp.SOL_GRID):
        lv, _ = ctx.bfs(t, 0, sol)
        parent = np.repeat(np.arange(t.n), np.diff(t.row))
        assert lv[0] == 0 and np.array_equal(lv[t.col], lv[parent] + 1)
        assert np.array_equal(lv, orc.bfs(t.row, t.col, 0))
    # resident: the renumbered copy, the forest path (one in-arc per vertex) and no queue entries for leaves
    ref = orc.bfs(t.row, t.col, 0)
    with smdp.Graph(ctx, t) as dg:
        for _ in range(2):
            dg.bfs(0, smdp.SOL_TUNED)
            assert np.array_equal(dg.read(smdp.RES_LEVEL), ref)


def test_batched_rmat21_and_small_graphs(ctx):
    g = smdp.gen_rmat(21, 16, wlo=1, whi=255, seed=4)
    print("batched rmat21", flush=True)
    srcs = smdp.gen_sources(g, 64, 4)
    outs, st = ctx.sssp_multi_source(g, srcs)
    # every one of the 64 sources (the oracle's Dijkstra runs on host threads)
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=16) as ex:
        refs = list(ex.map(lambda s: orc.dijkstra(g.row, g.col, g.w, int(s)), srcs))
    for i in range(64):
        assert np.array_equal(outs[i], refs[i]), i
    graphs = [smdp.gen_small_graph(i, seed=4) for i in range(256)]
    outs, _ = ctx.sssp_multi_graph(graphs, np.zeros(256, np.int32))
    for gi, di in zip(graphs, outs):
        assert np.array_equal(di, orc.dijkstra(gi.row, gi.col, gi.w, 0))
