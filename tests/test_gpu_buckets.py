"""Bucketed SSSP (SMDP_LAYOUT_BUCKETS, the tuned solution's default on large
resident graphs) against the oracle's Dijkstra.

Distances are unique whatever order arcs are relaxed in, so every bucket
width must give Dijkstra's distances bit for bit: widths from 1 (one distance
per bucket) to beyond the largest distance (one bucket, every arc relaxed in
light phases), zero weights (arcs that stay in their bucket), large weights
that jump many buckets, hubs cut into grid chunks, unreachable vertices, and a
graph with a negative weight (which must fall back to unordered rounds).
"""
import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bctx():
    c = smdp.Context(0)
    yield c
    c.close()


def _solve(bctx, g, src, delta, persistent=0, **params):
    bctx.set_params(smdp.APP_SSSP, layout=smdp.LAYOUT_BUCKETS, delta=delta, persistent=persistent, **params)
    with smdp.Graph(bctx, g) as dg:
        st = dg.sssp(src, smdp.SOL_TUNED)
        d = dg.read(smdp.RES_DIST)
        st2 = dg.sssp(src, smdp.SOL_TUNED)  # a second run on the same resident graph
        d2 = dg.read(smdp.RES_DIST)
    assert np.array_equal(d, d2)
    return d, st, st2


def _graphs():
    out = {}
    out["uniform16k"] = smdp.gen_uniform(16384, 8, 1, 100, 1)
    out["rmat14"] = smdp.gen_rmat(14, 16, wlo=1, whi=255, seed=3)
    row, col, w = random_graph(3000, 12000, 31, wlo=0, whi=20)
    out["zero_weights"] = smdp.HostGraph(row, col, w, 0)
    row, col, w = random_graph(4000, 16000, 32, wlo=1, whi=100000)
    out["large_weights"] = smdp.HostGraph(row, col, w, 5)
    row, col, w = random_graph(5000, 6000, 33)  # sparse: many vertices unreachable
    out["sparse_unreachable"] = smdp.HostGraph(row, col, w, 1)
    return out


GRAPHS = _graphs()


@pytest.mark.parametrize("delta", [1, 8, 64, 1 << 20])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_buckets_parity(bctx, name, delta):
    g = GRAPHS[name]
    print(f"buckets {name} delta={delta}", flush=True)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, _ = _solve(bctx, g, g.source, delta)
    assert np.array_equal(d, ref)
    reached = ref != orc.INF
    assert st.m_reached == int(np.diff(g.row)[reached].sum())
    # every arc of a reached vertex is read, and relaxed by at most one pass per
    # expansion; arcs left to the bucket's other pass are not counted as relaxed
    assert st.edges_relaxed + st.arcs_skipped >= st.m_reached


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_buckets_persistent_parity(bctx, name):
    """The same phases run in one launch whose workgroups are all resident,
    with a grid barrier between passes."""
    g = GRAPHS[name]
    print(f"buckets persistent {name}", flush=True)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, _ = _solve(bctx, g, g.source, 64, persistent=0)
    assert np.array_equal(d, ref)
    d2, st2, _ = _solve(bctx, g, g.source, 64, persistent=1)
    assert np.array_equal(d2, ref)
    # persistent: the initialisation, one launch per 1024 phases and a barrier
    # reset between two launches; per pass: two launches per phase
    assert st2.kernel_launches <= 2 * (st2.rounds // 1024 + 1) < st.kernel_launches


def test_buckets_known_answer(bctx, golden, data_test_path):
    g = smdp.read_graph(data_test_path, smdp.FMT_DIMACS9)
    for delta in (1, 2, 1024, 1 << 16):
        d, _, _ = _solve(bctx, g, g.source, delta)
        assert d.tolist() == golden["dist"], delta


def test_buckets_chunks_forced(bctx):
    """Split degree 2 and 64-arc chunks send every vertex of two or more arcs
    through the grid queue in every pass."""
    g = GRAPHS["rmat14"]
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    for persistent in (0, 1):
        d, st, _ = _solve(bctx, g, g.source, 32, persistent=persistent, split_threshold=2, chunk=64)
        assert np.array_equal(d, ref)
        assert st.q_push[3] > 0 and st.edges_phase2 > 0
    bctx.set_params(smdp.APP_SSSP, split_threshold=2048, chunk=2048)


def test_buckets_work_is_near_one_pass(bctx):
    """With delta = 64 on RMAT weights in [1,255] each reachable arc is relaxed
    about once (unordered rounds relax it about three times)."""
    g = smdp.gen_rmat(16, 16, wlo=1, whi=255, seed=3)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, _ = _solve(bctx, g, g.source, 64)
    assert np.array_equal(d, ref)
    assert st.edges_relaxed <= 1.6 * st.m_reached, (st.edges_relaxed, st.m_reached)


def test_buckets_single_vertex_and_no_arcs(bctx):
    one = smdp.HostGraph(np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    d, st, _ = _solve(bctx, one, 0, 64)
    assert d.tolist() == [0] and st.edges_relaxed == 0
    row = np.zeros(101, np.int32)
    none = smdp.HostGraph(row, np.zeros(0, np.int32), None)
    d, _, _ = _solve(bctx, none, 7, 64)
    assert d[7] == 0 and (np.delete(d, 7) == orc.INF).all()


def test_buckets_source_without_arcs(bctx):
    g = GRAPHS["uniform16k"]
    row = g.row.copy()
    # vertex 0 loses its arcs: only it is reached
    col = g.col[row[1]:]
    w = g.w[row[1]:]
    row = np.concatenate([[0], row[1:] - row[1]]).astype(np.int32)
    h = smdp.HostGraph(row, col, w)
    d, _, _ = _solve(bctx, h, 0, 64)
    assert d[0] == 0 and (d[1:] == orc.INF).all()


def test_buckets_negative_weight_falls_back(bctx):
    """A negative weight (no negative cycle: arcs only go to larger ids) is
    not bucketed; the unordered rounds still give exact distances."""
    rng = np.random.default_rng(7)
    n, m = 2000, 10000
    src = np.sort(rng.integers(0, n - 1, m))
    dst = np.minimum(n - 1, src + 1 + rng.integers(0, 50, m))
    w = rng.integers(-5, 100, m).astype(np.int32)
    row = np.zeros(n + 1, np.int64)
    np.add.at(row, src + 1, 1)
    row = np.cumsum(row).astype(np.int32)
    g = smdp.HostGraph(row, dst.astype(np.int32), w, 0)
    ref = orc.bellman_ford(g.row, g.col, g.w, 0)
    d, _, _ = _solve(bctx, g, 0, 64)
    assert np.array_equal(d, ref)


def test_buckets_then_other_results_are_not_readable(bctx):
    """Reading a kind of result the last run did not produce is an error."""
    g = GRAPHS["uniform16k"]
    bctx.set_params(smdp.APP_SSSP, layout=smdp.LAYOUT_BUCKETS, delta=64)
    with smdp.Graph(bctx, g) as dg:
        dg.sssp(g.source, smdp.SOL_TUNED)
        with pytest.raises(smdp.SmdpError):
            dg.read(smdp.RES_LEVEL)
        assert np.array_equal(dg.read(smdp.RES_DIST), orc.dijkstra(g.row, g.col, g.w, g.source))


def _chain(n, w):
    """Vertex i -> i + 1 for i < n - 1, every weight w: the solve's rounds and
    phases do not depend on the order of relaxations."""
    row = np.minimum(np.arange(n + 1), n - 1).astype(np.int32)
    col = np.arange(1, n, dtype=np.int32)
    return smdp.HostGraph(row, col, np.full(n - 1, w, np.int32), 0)


@pytest.mark.parametrize("resident", [False, True])
def test_round_cap_is_exact(bctx, resident):
    """The host keeps rounds (phases) queued ahead of the device, but never more
    than max_rounds: a cap of one more than the steps with work (the last one
    finds none) succeeds, and the steps with work alone are ERR_OVERFLOW."""
    g = _chain(20, 100)
    ref = orc.dijkstra(g.row, g.col, g.w, 0)

    def solve(cap):
        if resident:
            bctx.set_params(smdp.APP_SSSP, layout=smdp.LAYOUT_BUCKETS, delta=64, max_rounds=cap)
            with smdp.Graph(bctx, g) as dg:
                st = dg.sssp(0, smdp.SOL_TUNED)
                return dg.read(smdp.RES_DIST), st
        bctx.set_params(smdp.APP_SSSP, max_rounds=cap)
        return bctx.sssp(g, 0, smdp.SOL_TUNED)

    old = bctx.get_params(smdp.APP_SSSP)
    try:
        d, st = solve(1 << 20)
        assert np.array_equal(d, ref)
        steps = int(st.rounds)
        print(f"chain of 20, resident={resident}: {steps} steps with work", flush=True)
        d, _ = solve(steps + 1)
        assert np.array_equal(d, ref)
        with pytest.raises(smdp.SmdpError) as e:
            solve(steps)
        assert e.value.code == smdp.ERR_OVERFLOW
    finally:
        bctx.set_params(smdp.APP_SSSP, **old.as_dict())
