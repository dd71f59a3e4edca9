"""Bucketed SSSP (SMDP_LAYOUT_BUCKETS, the tuned solution's default on large
resident graphs) against the oracle's Dijkstra.

Distances are unique whatever order arcs are relaxed in, so every bucket
width must give Dijkstra's distances bit for bit: widths from 1 (one distance
per bucket) to beyond the largest distance (one bucket, every arc relaxed in
light phases), zero weights (arcs that stay in their bucket), large weights
that jump many buckets, hubs cut into grid chunks, unreachable vertices, and a
graph with a negative weight (which must fall back to unordered rounds).  The
same holds for each way a settle pass reaches the arcs that leave its bucket:
push (the bucket's vertices relax them), pull (the vertices beyond the bucket
read their in-arcs) and the automatic choice between the two.
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


SETTLES = {"auto": smdp.SETTLE_AUTO, "push": smdp.SETTLE_PUSH, "pull": smdp.SETTLE_PULL}


def _solve(bctx, g, src, delta, persistent=0, settle=smdp.SETTLE_AUTO, **params):
    bctx.set_params(smdp.APP_SSSP, layout=smdp.LAYOUT_BUCKETS, delta=delta, persistent=persistent, settle=settle,
                    **params)
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


@pytest.mark.parametrize("settle", sorted(SETTLES))
@pytest.mark.parametrize("delta", [1, 8, 64, 1 << 20])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_buckets_parity(bctx, name, delta, settle):
    g = GRAPHS[name]
    print(f"buckets {name} delta={delta} settle={settle}", flush=True)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, _ = _solve(bctx, g, g.source, delta, settle=SETTLES[settle])
    assert np.array_equal(d, ref)
    reached = ref != orc.INF
    assert st.m_reached == int(np.diff(g.row)[reached].sum())
    if settle == "push":
        # every arc of a reached vertex is read, and relaxed by at most one pass
        # per expansion; arcs left to the bucket's other pass are not counted
        # as relaxed (a pull settle reads in-arcs instead)
        assert st.edges_relaxed + st.arcs_skipped >= st.m_reached


@pytest.mark.parametrize("settle", ["push", "pull"])
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_buckets_persistent_parity(bctx, name, settle):
    """The same phases run in one launch whose workgroups are all resident,
    with a grid barrier between passes."""
    g = GRAPHS[name]
    print(f"buckets persistent {name} settle={settle}", flush=True)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, _ = _solve(bctx, g, g.source, 64, persistent=0, settle=SETTLES[settle])
    assert np.array_equal(d, ref)
    d2, st2, _ = _solve(bctx, g, g.source, 64, persistent=1, settle=SETTLES[settle])
    assert np.array_equal(d2, ref)
    # persistent: the initialisation, one launch per 1024 phases and a barrier
    # reset between two launches; per pass: two launches per phase
    assert st2.kernel_launches <= 2 * (st2.rounds // 1024 + 1) < st.kernel_launches


def test_buckets_known_answer(bctx, golden, data_test_path):
    g = smdp.read_graph(data_test_path, smdp.FMT_DIMACS9)
    for settle in SETTLES.values():
        for delta in (1, 2, 1024, 1 << 16):
            d, _, _ = _solve(bctx, g, g.source, delta, settle=settle)
            assert d.tolist() == golden["dist"], (delta, settle)


def test_buckets_chunks_forced(bctx):
    """Split degree 2 and 64-arc chunks send every vertex of two or more arcs
    through the grid queue in every pass."""
    g = GRAPHS["rmat14"]
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    for settle in (smdp.SETTLE_PUSH, smdp.SETTLE_PULL):
        for persistent in (0, 1):
            d, st, _ = _solve(bctx, g, g.source, 32, persistent=persistent, settle=settle, split_threshold=2,
                              chunk=64)
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


def _settle_modes(bctx, g, delta, settle):
    """Distances and stats of one solve with the per-phase log on, and the
    log's mode codes of its settle phases (2 = push, 4 = pull)."""
    bctx.set_profiling(True)
    try:
        d, st, _ = _solve(bctx, g, g.source, delta, settle=settle)
        q, _, _ = bctx.round_log()
    finally:
        bctx.set_profiling(False)
    modes = (q.astype(np.uint64) >> np.uint64(60)).astype(int)
    return d, st, [m for m in modes.tolist() if m in (2, 4)]


def test_auto_settle_pulls_on_rmat(bctx):
    """RMAT's first full bucket holds the hubs: their out-arcs far outnumber
    the in-arcs of the vertices still beyond the bucket (on this graph bucket
    0 is the source alone, with 1 arc against 1,048,574 in-arcs beyond it; in
    bucket 1, 977,210 out-arcs against 62,702), so the automatic choice pushes
    the first settle, pulls the next, reads fewer arcs than push, and gives the
    same distances."""
    g = smdp.gen_rmat(16, 16, wlo=1, whi=255, seed=3)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, st, modes = _settle_modes(bctx, g, 64, smdp.SETTLE_AUTO)
    assert np.array_equal(d, ref)
    assert len(modes) >= 2 and modes[:2] == [2, 4], modes
    dp, stp, pmodes = _settle_modes(bctx, g, 64, smdp.SETTLE_PUSH)
    assert np.array_equal(dp, ref)
    assert set(pmodes) == {2}
    read, read_push = st.edges_relaxed + st.arcs_skipped, stp.edges_relaxed + stp.arcs_skipped
    print(f"rmat16: arcs read auto {read} push {read_push}", flush=True)
    assert read < read_push


def _grid(k, seed):
    """k x k grid, arcs both ways between 4-neighbours, weights in [1, 100]."""
    rng = np.random.default_rng(seed)
    ids = np.arange(k * k).reshape(k, k)
    src = np.concatenate([ids[:, :-1].ravel(), ids[:, 1:].ravel(), ids[:-1, :].ravel(), ids[1:, :].ravel()])
    dst = np.concatenate([ids[:, 1:].ravel(), ids[:, :-1].ravel(), ids[1:, :].ravel(), ids[:-1, :].ravel()])
    order = np.argsort(src, kind="stable")
    src, dst = src[order], dst[order]
    row = np.searchsorted(src, np.arange(k * k + 1)).astype(np.int32)
    w = rng.integers(1, 101, len(dst)).astype(np.int32)
    return smdp.HostGraph(row, dst.astype(np.int32), w, 0)


def test_auto_settle_pushes_on_long_paths(bctx):
    """On a grid most vertices stay beyond each bucket, so their in-arcs
    outnumber the bucket's out-arcs: the automatic choice never pulls."""
    g = _grid(120, 5)
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    d, _, modes = _settle_modes(bctx, g, 32, smdp.SETTLE_AUTO)
    assert np.array_equal(d, ref)
    assert modes and 4 not in modes, modes
    d, _, modes = _settle_modes(bctx, g, 32, smdp.SETTLE_PULL)
    assert np.array_equal(d, ref)
    assert set(modes) == {4}


@pytest.mark.parametrize("settle", sorted(SETTLES))
def test_buckets_single_vertex_and_no_arcs(bctx, settle):
    one = smdp.HostGraph(np.array([0, 0], np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    d, st, _ = _solve(bctx, one, 0, 64, settle=SETTLES[settle])
    assert d.tolist() == [0] and st.edges_relaxed == 0
    row = np.zeros(101, np.int32)
    none = smdp.HostGraph(row, np.zeros(0, np.int32), None)
    d, _, _ = _solve(bctx, none, 7, 64, settle=SETTLES[settle])
    assert d[7] == 0 and (np.delete(d, 7) == orc.INF).all()


def test_buckets_source_without_arcs(bctx):
    g = GRAPHS["uniform16k"]
    row = g.row.copy()
    # vertex 0 loses its arcs: only it is reached
    col = g.col[row[1]:]
    w = g.w[row[1]:]
    row = np.concatenate([[0], row[1:] - row[1]]).astype(np.int32)
    h = smdp.HostGraph(row, col, w)
    for settle in SETTLES.values():
        d, _, _ = _solve(bctx, h, 0, 64, settle=settle)
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
