"""Batched SSSP on the device: many sources on one graph, and many small
graphs in one captured hipGraph launch.  Distances must equal Dijkstra's."""
import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k", [1, 5, 64, 70])
def test_multi_source_parity(ctx, k):
    g = smdp.gen_rmat(12, 16, wlo=1, whi=255, seed=4)
    srcs = smdp.gen_sources(g, k, 11)
    d, st = ctx.sssp_multi_source(g, srcs)
    assert d.shape == (k, g.n)
    for i, s in enumerate(srcs):
        assert np.array_equal(d[i], orc.dijkstra(g.row, g.col, g.w, int(s))), i
    assert st.rounds >= 1


def test_multi_source_resident_read_and_replays(ctx):
    """The resident call keeps [k][n] on the device for smdp_graph_read; later
    calls with the same k replay the captured graph with new sources, with
    eager work (which may reallocate the graph's shared grid queue) between
    replays."""
    g = smdp.gen_rmat(13, 16, wlo=1, whi=255, seed=7)
    with smdp.Graph(ctx, g) as dg:
        for rep in range(4):
            srcs = smdp.gen_sources(g, 40, 100 + rep)
            st = dg.sssp_multi_source(srcs)
            d = dg.read(smdp.RES_MDIST, count=40 * g.n).reshape(40, g.n)
            for i, s in enumerate(srcs):
                assert np.array_equal(d[i], orc.dijkstra(g.row, g.col, g.w, int(s))), (rep, i)
            assert st.rounds >= 1 and st.edges_relaxed > 0
            # eager work on the same graph between replays: an SSSP with small
            # chunks needs a larger grid queue than the batch captured
            old = ctx.get_params(smdp.APP_SSSP)
            ctx.set_params(smdp.APP_SSSP, layout=smdp.LAYOUT_HOT, split_threshold=64, chunk=64)
            dg.sssp(int(srcs[0]), smdp.SOL_TUNED)
            ctx.set_params(smdp.APP_SSSP, **old.as_dict())
            assert np.array_equal(dg.read(smdp.RES_DIST), orc.dijkstra(g.row, g.col, g.w, int(srcs[0])))
            with pytest.raises(smdp.SmdpError):
                dg.read(smdp.RES_MDIST, count=40 * g.n)  # the last run was an SSSP


def test_results_of_another_run_are_not_readable(ctx):
    """A run invalidates what an earlier run of another kind left readable,
    and a failed run leaves nothing readable."""
    t = smdp.gen_zipf_tree(5000, 1.5, 64, 3)
    row, col = t.row, t.col
    w = np.ones(len(col), np.int32)
    with smdp.Graph(ctx, smdp.HostGraph(row, col, w)) as dg:
        dg.tree(0)
        dg.read(smdp.RES_DESC)
        dg.sssp_multi_source(np.array([0, 1], np.int32))
        with pytest.raises(smdp.SmdpError):
            dg.read(smdp.RES_DESC)
    # colour, then an SSSP that fails on a negative cycle: no colours to read
    rowc, colc, wc = random_graph(500, 3000, 5, wlo=1, whi=9)
    wc = wc.copy()
    wc[:] = -1  # every cycle is negative
    old = ctx.get_params(smdp.APP_SSSP)
    with smdp.Graph(ctx, smdp.HostGraph(rowc, colc, wc)) as dg:
        dg.color(False, smdp.SOL_TUNED)
        dg.read(smdp.RES_COLOR)
        try:
            ctx.set_params(smdp.APP_SSSP, max_rounds=50)
            with pytest.raises(smdp.SmdpError) as e:
                dg.sssp(0, smdp.SOL_FLAT)
            assert e.value.code == smdp.ERR_OVERFLOW
        finally:
            ctx.set_params(smdp.APP_SSSP, **old.as_dict())
        with pytest.raises(smdp.SmdpError):
            dg.read(smdp.RES_COLOR)


def test_multi_source_duplicates_and_chunks(ctx):
    row, col, w = random_graph(3000, 40000, 12, 1, 30)
    g = smdp.HostGraph(row, col, w)
    srcs = np.array([0, 0, 5, 17, 5, 2999], dtype=np.int32)
    old = ctx.get_params(smdp.APP_MSSSP)
    try:
        for split, chunk in [(512, 256), (4, 64)]:
            ctx.set_params(smdp.APP_MSSSP, split_threshold=split, chunk=chunk)
            d, _ = ctx.sssp_multi_source(g, srcs)
            for i, s in enumerate(srcs):
                assert np.array_equal(d[i], orc.dijkstra(row, col, w, int(s)))
    finally:
        ctx.set_params(smdp.APP_MSSSP, **old.as_dict())


def _small_graphs(count, seed=4):
    gs = [smdp.gen_small_graph(i, vlo=64, vhi=4096, seed=seed) for i in range(count)]
    gs.append(smdp.HostGraph([0, 0], [], [], 0))           # one vertex
    gs.append(smdp.HostGraph([0, 1, 1, 1], [2], [9], 0))   # unreachable vertex
    return gs


def test_multi_graph_parity(ctx):
    graphs = _small_graphs(40)
    srcs = np.zeros(len(graphs), dtype=np.int32)
    srcs[3] = graphs[3].n - 1
    outs, st = ctx.sssp_multi_graph(graphs, srcs)
    for i, (g, d) in enumerate(zip(graphs, outs)):
        assert np.array_equal(d, orc.dijkstra(g.row, g.col, g.w, int(srcs[i]))), i


def test_batch_graph_replays(ctx):
    """The captured hipGraph is replayed with eager work in between; every
    replay must give the same exact distances."""
    graphs = _small_graphs(24, seed=9)
    srcs = np.zeros(len(graphs), dtype=np.int32)
    refs = [orc.dijkstra(g.row, g.col, g.w, 0) for g in graphs]
    with smdp.Batch(ctx, graphs, srcs) as b:
        for rep in range(4):
            st = b.run()
            for i in range(len(graphs)):
                assert np.array_equal(b.read(i), refs[i]), (rep, i)
            ctx.sssp(graphs[0], 0, smdp.SOL_TUNED)  # eager work between replays
        assert st.edges_relaxed > 0


def test_multi_graph_bad_args(ctx):
    g = smdp.HostGraph([0, 1, 1], [1], [3])
    with pytest.raises(smdp.SmdpError) as e:
        ctx.sssp_multi_graph([g], np.array([5], dtype=np.int32))
    assert e.value.code == smdp.ERR_ARG


def test_batch_big_graph_solved_by_a_group(ctx):
    """A graph far above a workgroup's share of the batch is solved by a group
    of workgroups that meet at a barrier after each round; the small ones are
    claimed one per workgroup.  Exact on every replay."""
    big = smdp.gen_small_graph(0, vlo=60000, vhi=60000, dlo=16.0, dhi=16.0, seed=11)
    graphs = [big] + [smdp.gen_small_graph(i, vlo=64, vhi=2048, seed=12) for i in range(1, 31)]
    srcs = np.zeros(len(graphs), dtype=np.int32)
    srcs[0] = 17
    refs = [orc.dijkstra(g.row, g.col, g.w, int(s)) for g, s in zip(graphs, srcs)]
    print(f"group batch: big graph {big.n} vertices {big.m} arcs", flush=True)
    with smdp.Batch(ctx, graphs, srcs) as b:
        for rep in range(3):
            b.run()
            for i in range(len(graphs)):
                assert np.array_equal(b.read(i), refs[i]), (rep, i)


def test_batch_group_negative_cycle(ctx):
    """A negative cycle in a group-solved graph ends after n rounds with the
    overflow error, like one solved by a single workgroup."""
    big = smdp.gen_small_graph(0, vlo=4000, vhi=4000, dlo=16.0, dhi=16.0, seed=13)
    row = np.asarray(big.row, np.int64).copy()
    col = np.asarray(big.col).tolist()
    w = np.asarray(big.w).tolist()
    # arcs 0 -> 1 and 1 -> 0 of weight -5: a cycle reached from source 0
    for u, v in ((0, 1), (1, 0)):
        col.insert(int(row[u + 1]), v)
        w.insert(int(row[u + 1]), -5)
        row[u + 1:] += 1
    neg = smdp.HostGraph(row.astype(np.int32), np.asarray(col, np.int32), np.asarray(w, np.int32), 0)
    graphs = [neg] + [smdp.gen_small_graph(i, vlo=64, vhi=512, seed=14) for i in range(1, 9)]
    print("group batch: negative cycle", flush=True)
    with pytest.raises(smdp.SmdpError) as e:
        ctx.sssp_multi_graph(graphs, np.zeros(len(graphs), dtype=np.int32))
    assert e.value.code == smdp.ERR_OVERFLOW


def test_batch_autotune_picks_a_candidate_plan_and_stays_exact(ctx):
    """smdp_batch_autotune re-plans the batch for each candidate (workgroups
    per CU 1 or 2, group share 50/100/200 percent), keeps the fastest and
    installs it in the multi-source parameters; the re-captured batch is
    exact on every replay."""
    big = smdp.gen_small_graph(0, vlo=40000, vhi=40000, dlo=16.0, dhi=16.0, seed=21)
    graphs = [big] + [smdp.gen_small_graph(i, vlo=64, vhi=4096, seed=22) for i in range(1, 48)]
    srcs = np.zeros(len(graphs), dtype=np.int32)
    refs = [orc.dijkstra(g.row, g.col, g.w, 0) for g in graphs]
    old = ctx.get_params(smdp.APP_MSSSP)
    try:
        with smdp.Batch(ctx, graphs, srcs) as b:
            chosen = b.autotune()
            assert chosen.wg_per_cu in (1, 2) and chosen.group_share in (50, 100, 200), chosen.as_dict()
            now = ctx.get_params(smdp.APP_MSSSP)
            assert (now.wg_per_cu, now.group_share) == (chosen.wg_per_cu, chosen.group_share)
            for rep in range(3):
                b.run()
                for i in range(len(graphs)):
                    assert np.array_equal(b.read(i), refs[i]), (rep, i)
    finally:
        ctx.set_params(smdp.APP_MSSSP, **old.as_dict())


def test_group_share_is_validated(ctx):
    old = ctx.get_params(smdp.APP_MSSSP)
    try:
        for bad in (5, 1001):
            with pytest.raises(smdp.SmdpError) as e:
                ctx.set_params(smdp.APP_MSSSP, group_share=bad)
            assert e.value.code == smdp.ERR_ARG
        ctx.set_params(smdp.APP_MSSSP, group_share=0)
        assert ctx.get_params(smdp.APP_MSSSP).group_share == 100
    finally:
        ctx.set_params(smdp.APP_MSSSP, **old.as_dict())
