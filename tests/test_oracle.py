"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Every oracle function is checked against something it shares no code with:
known answers computed by library routines (tests/golden), scipy/networkx,
a second algorithm, a closed form, or brute force on tiny inputs.  Each check
is chosen so that a plausible slip (wrong sign, dropped term, off-by-one,
transposed operand) fails it.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra, shortest_path

import orc
from conftest import random_graph, symmetrize


def _scipy_dist(row, col, w, src):
    n = len(row) - 1
    srcs = np.repeat(np.arange(n), np.diff(row))
    best = {}
    for u, v, ww in zip(srcs.tolist(), col.tolist(), w.tolist()):
        if u == v:
            continue
        best[(u, v)] = min(ww, best.get((u, v), ww))
    if not best:
        d = np.full(n, np.inf)
        d[src] = 0
    else:
        k = list(best)
        A = sp.csr_matrix(([best[x] for x in k], ([x[0] for x in k], [x[1] for x in k])), shape=(n, n))
        d = sp_dijkstra(A, directed=True, indices=src)
    return np.where(np.isfinite(d), d, orc.INF).astype(np.int64)


def _golden_graph(data_test_path):
    import smdp
    return smdp.read_graph(data_test_path, smdp.FMT_DIMACS9)


# ---- SSSP ------------------------------------------------------------------

def test_dijkstra_known_answer(golden, data_test_path):
    g = _golden_graph(data_test_path)
    assert orc.dijkstra(g.row, g.col, g.w, golden["source"]).tolist() == golden["dist"]


@pytest.mark.parametrize("seed", range(6))
def test_dijkstra_vs_scipy(seed):
    row, col, w = random_graph(300, 1500 + 200 * seed, seed, 1, 50 + 100 * seed)
    d = orc.dijkstra(row, col, w, seed)
    assert np.array_equal(d, _scipy_dist(row, col, w, seed))


@pytest.mark.parametrize("seed", range(10))
def test_dijkstra_vs_brute_force(seed):
    row, col, w = random_graph(12, 30, 100 + seed, 1, 9)
    assert np.array_equal(orc.dijkstra(row, col, w, 0), orc.bellman_ford(row, col, w, 0))


@pytest.mark.parametrize("seed", range(6))
def test_worklist_matches_dijkstra(seed):
    row, col, w = random_graph(2000, 16000, 7 + seed, 1, 255)
    d, rounds, edges = orc.sssp_worklist(row, col, w, 3)
    assert np.array_equal(d, orc.dijkstra(row, col, w, 3))
    assert rounds >= 1 and edges >= np.diff(row)[3]


@pytest.mark.parametrize("seed", range(3))
def test_parallel_worklist_matches_dijkstra(seed):
    # the multi-core CPU baseline: same rounds over all host threads, atomic minimum
    row, col, w = random_graph(20000, 200000, 11 + seed, 1, 255)
    d, rounds, edges, threads = orc.sssp_worklist_par(row, col, w, 5)
    assert np.array_equal(d, orc.dijkstra(row, col, w, 5))
    assert rounds >= 1 and edges >= np.diff(row)[5] and threads >= 1


def test_sssp_unreachable_sentinel():
    row = np.array([0, 1, 1, 1], dtype=np.int32)
    col = np.array([1], dtype=np.int32)
    w = np.array([5], dtype=np.int32)
    assert orc.dijkstra(row, col, w, 0).tolist() == [0, 5, orc.INF]
    assert orc.sssp_worklist(row, col, w, 0)[0].tolist() == [0, 5, orc.INF]
    assert orc.sssp_worklist_par(row, col, w, 0)[0].tolist() == [0, 5, orc.INF]


# ---- BFS -------------------------------------------------------------------

def test_bfs_known_answer(golden, data_test_path):
    g = _golden_graph(data_test_path)
    assert orc.bfs(g.row, g.col, golden["source"]).tolist() == golden["level"]


@pytest.mark.parametrize("seed", range(6))
def test_bfs_vs_scipy(seed):
    row, col, _ = random_graph(500, 1200 + 300 * seed, 50 + seed)
    n = len(row) - 1
    A = sp.csr_matrix((np.ones(len(col)), col, row), shape=(n, n))
    ref = shortest_path(A, directed=True, unweighted=True, indices=seed)
    ref = np.where(np.isfinite(ref), ref, orc.UNDEF).astype(np.int64)
    assert np.array_equal(orc.bfs(row, col, seed), ref)


# ---- coloring --------------------------------------------------------------

def test_color_known_answer(golden, data_test_path):
    g = _golden_graph(data_test_path)
    c1, k1 = orc.color_rounds(g.row, g.col)
    c2, k2 = orc.color_closed(g.row, g.col)
    assert c1.tolist() == golden["color"] and c2.tolist() == golden["color"]
    assert k1 == k2 == max(golden["color"])


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("use_prio", [False, True])
def test_color_rounds_equals_closed_form(seed, use_prio):
    row, col, _ = random_graph(400, 2500, 200 + seed)
    prio = np.random.default_rng(seed).integers(0, 2**32, 400, dtype=np.uint64).astype(np.uint32) if use_prio else None
    if use_prio and seed % 2:
        prio = (prio % 7).astype(np.uint32)  # many ties: broken by vertex id
    a, ka = orc.color_rounds(row, col, prio)
    b, kb = orc.color_closed(row, col, prio)
    assert np.array_equal(a, b) and ka == kb


@pytest.mark.parametrize("seed", range(4))
def test_color_proper_on_symmetric(seed):
    row, col, _ = random_graph(300, 1500, 300 + seed, self_loops=False)
    row, col = symmetrize(row, col)
    prio = np.random.default_rng(seed).integers(0, 2**32, 300, dtype=np.uint64).astype(np.uint32)
    c, _ = orc.color_closed(row, col, prio)
    src = np.repeat(np.arange(300), np.diff(row))
    assert np.all(c > 0)
    assert not np.any(c[src] == c[col])


def test_color_chain_brute_force():
    # path 0-1-2-...-9 with identity priority: colours alternate 1,2
    n = 10
    src = list(range(n - 1)) + list(range(1, n))
    dst = list(range(1, n)) + list(range(n - 1))
    order = np.argsort(src, kind="stable")
    src = np.array(src)[order]
    dst = np.array(dst)[order]
    row = np.searchsorted(src, np.arange(n + 1)).astype(np.int32)
    # identity priority: the path in id order is one precedence chain
    c, _ = orc.color_rounds(row, dst)
    assert c.tolist() == list(range(1, n + 1))
    # even vertices first: two colours alternate
    prio = np.array([i % 2 for i in range(n)], dtype=np.uint32)
    c, k = orc.color_rounds(row, dst, prio)
    assert c.tolist() == [1, 2] * 5 and k == 2
    # directed chain i -> i-1 only: colour(i) = i + 1
    row2 = np.array([0] + list(range(0, n)), dtype=np.int32)
    col2 = np.arange(0, n - 1, dtype=np.int32)
    c2, k2 = orc.color_closed(row2, col2)
    assert c2.tolist() == list(range(1, n + 1)) and k2 == n


# ---- SpMV ------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(4))
def test_spmv_vs_scipy(seed):
    row, col, _ = random_graph(700, 9000, 400 + seed)
    rng = np.random.default_rng(seed)
    val = rng.uniform(-1, 1, len(col)).astype(np.float32)
    x = rng.uniform(0, 1, 700).astype(np.float32)
    A = sp.csr_matrix((val.astype(np.float64), col, row), shape=(700, 700))
    ref = A @ x.astype(np.float64)
    y = orc.spmv(row, col, val, x)
    assert np.max(np.abs(y - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_spmv_reference_inputs_exact():
    # data[i] = i%13+0.5, x[i] = i%17+0.5 (SpMV_wrapper.cu:105-110): 4*y is an
    # integer, computed exactly with Python ints.
    row, col, _ = random_graph(500, 20000, 9)
    m = len(col)
    data = (np.arange(m) % 13 + 0.5).astype(np.float32)
    x = (np.arange(500) % 17 + 0.5).astype(np.float32)
    y = orc.spmv(row, col, data, x)
    for r in range(500):
        s4 = sum((2 * (i % 13) + 1) * (2 * (col[i] % 17) + 1) for i in range(row[r], row[r + 1]))
        assert y[r] * 4 == s4


# ---- PageRank --------------------------------------------------------------

def _pr_direct(row, col, d, T):
    """Fixed point of the reference update by a dense linear solve."""
    n = len(row) - 1
    outdeg = np.diff(row)
    M = np.zeros((n, n))
    src = np.repeat(np.arange(n), outdeg)
    for u, v in zip(src, col):
        M[v, u] += 1.0 / outdeg[u]
    dang = (outdeg == 0).astype(float)
    A = np.eye(n) - d * M - d * np.outer(np.ones(n), dang) / n
    b = np.full(n, (1 - d) * T / n)
    return np.linalg.solve(A, b)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("d,T", [(0.85, 1.0), (0.5, 1000.0)])
def test_pagerank_fixed_point(seed, d, T):
    import smdp
    row, col, _ = random_graph(60, 200, 500 + seed)  # has dangling vertices
    g = smdp.HostGraph(row, col)
    ig = smdp.invert(g)
    r, it, res = orc.pagerank(ig.row, ig.col, np.diff(row), d=d, T=T, tol=1e-15, max_iter=2000)
    ref = _pr_direct(row, col, d, T)
    assert np.max(np.abs(r - ref)) <= 1e-10 * T
    assert abs(r.sum() - T) <= 1e-9 * T


def test_pagerank_iteration_matched_and_init():
    import smdp
    row, col, _ = random_graph(50, 150, 77)
    ig = smdp.invert(smdp.HostGraph(row, col))
    od = np.diff(row)
    r1, it1, _ = orc.pagerank(ig.row, ig.col, od, fixed_iters=1, init_zero=True, d=0.5, T=1000.0)
    # one step from r = 0: every rank is (1-d)T/N (no dangling mass yet)
    assert np.allclose(r1, 0.5 * 1000.0 / 50, rtol=0, atol=1e-12) and it1 == 1
    r2, it2, res2 = orc.pagerank(ig.row, ig.col, od, tol=1e-6, max_iter=100)
    assert it2 < 100 and res2 < 1e-6


@pytest.mark.parametrize("seed,d", [(0, 0.85), (1, 0.5), (2, 0.99)])
def test_pagerank_k_steps_are_powers_of_the_google_matrix(seed, d):
    """From ranks summing to T, k pull iterations with dangling mass spread
    uniformly are G^k r0 for the Google matrix G = d P + (1 - d)/N 11^T, where
    column u of P is 1/outdeg(u) on u's targets (arcs counted with their
    multiplicity) and 1/N everywhere when u has no arcs.  A dense matrix power
    pins every iterate, not only the fixed point: a dangling mass taken from
    the new ranks, or damping left off it, moves them."""
    import smdp
    row, col, _ = random_graph(40, 120, 900 + seed)  # dangling vertices and duplicate arcs
    n = len(row) - 1
    assert (np.diff(row) == 0).any()
    P = np.zeros((n, n))
    for u in range(n):
        deg = row[u + 1] - row[u]
        if deg == 0:
            P[:, u] = 1.0 / n
        for e in range(row[u], row[u + 1]):
            P[col[e], u] += 1.0 / deg
    G = d * P + (1.0 - d) / n * np.ones((n, n))
    T = 3.0
    r0 = np.full(n, T / n)
    ig = smdp.invert(smdp.HostGraph(row, col))
    for k in range(1, 7):
        rk, it, _ = orc.pagerank(ig.row, ig.col, np.diff(row), d=d, T=T, fixed_iters=k)
        assert it == k
        assert np.max(np.abs(rk - np.linalg.matrix_power(G, k) @ r0)) <= 1e-13 * T, k


# ---- OpenMP twins (CPU baseline timing) -----------------------------------

@pytest.mark.parametrize("seed", [0, 1])
def test_parallel_twins_match_sequential(seed):
    rng = np.random.default_rng(40 + seed)
    row, col, w = random_graph(5000, 40000, 50 + seed, 1, 9)
    n = len(row) - 1
    val = rng.uniform(-1, 1, len(col)).astype(np.float32)
    x = rng.uniform(0, 1, n).astype(np.float32)
    y, nt = orc.spmv_par(row, col, val, x)
    assert nt >= 1 and np.array_equal(y, orc.spmv(row, col, val, x))  # same per-row sums
    lv, _ = orc.bfs_par(row, col, 3)
    assert np.array_equal(lv, orc.bfs(row, col, 3))
    # colouring: symmetric and directed inputs, id and random priorities (with ties)
    sym_r, sym_c = _symmetrize(row, col)
    for r_, c_ in ((row, col), (sym_r, sym_c)):
        for prio in (None, rng.integers(0, 50, n).astype(np.uint32)):
            c, k, _ = orc.color_layers_par(r_, c_, prio)
            ref, kref = orc.color_closed(r_, c_, prio)
            assert np.array_equal(c, ref) and k == kref
    # PageRank: same iterations, reductions in another order
    od = np.diff(row).astype(np.int32)
    inv_r, inv_c = _invert(row, col)
    ref, it, res = orc.pagerank(inv_r, inv_c, od, fixed_iters=12)
    r, res_p, _ = orc.pagerank_par(inv_r, inv_c, od, 12)
    assert np.abs(r - ref).max() <= 1e-14 and abs(res_p - res) <= 1e-12


def _symmetrize(row, col):
    n = len(row) - 1
    src = np.repeat(np.arange(n), np.diff(row))
    s = np.concatenate([src, col]); d = np.concatenate([col, src])
    o = np.lexsort((d, s))
    r = np.zeros(n + 1, np.int64)
    np.add.at(r, s[o] + 1, 1)
    return np.cumsum(r).astype(np.int32), d[o].astype(np.int32)


def _invert(row, col):
    n = len(row) - 1
    src = np.repeat(np.arange(n), np.diff(row))
    o = np.lexsort((src, col))
    r = np.zeros(n + 1, np.int64)
    np.add.at(r, col[o] + 1, 1)
    return np.cumsum(r).astype(np.int32), src[o].astype(np.int32)


# ---- Tree Height / Tree Descendant (README.md:12-13 names them; no reference code)

def _tree_csr(parent, order_seed=None):
    """Child lists of the tree given by parent[] (parent[root] = -1), children
    of each vertex in a seeded random order."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    for v, p in enumerate(parent):
        if p >= 0:
            kids[p].append(v)
    rng = np.random.default_rng(order_seed)
    for k in kids:
        rng.shuffle(k)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum([len(k) for k in kids])
    col = np.array([c for k in kids for c in k], np.int32)
    return row, col


def _random_parent(n, seed):
    """A random tree on a shuffled id space: vertex i > 0 of the hidden order
    hangs below a random earlier one."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    parent = np.full(n, -1)
    for i in range(1, n):
        parent[perm[i]] = perm[rng.integers(0, i)]
    return parent, int(perm[0])


@pytest.mark.parametrize("k,h", [(2, 5), (3, 4), (5, 2), (1, 7)])
def test_tree_complete_kary_closed_form(k, h):
    """Complete k-ary tree of height h, heap-numbered: a vertex at depth t has
    height h - t and (k^(h-t+1) - 1)/(k - 1) - 1 descendants (h - t for a chain)."""
    n = sum(k ** t for t in range(h + 1))
    parent = np.array([-1] + [(v - 1) // k for v in range(1, n)])
    row, col = _tree_csr(parent, 1)
    depth = np.zeros(n, int)
    for v in range(1, n):
        depth[v] = depth[parent[v]] + 1
    hh, dd, th = orc.tree(row, col, 0)
    assert th == h
    assert np.array_equal(hh, h - depth)
    size = (h - depth + 1) if k == 1 else (k ** (h - depth + 1) - 1) // (k - 1)
    assert np.array_equal(dd, size - 1)


def test_tree_star_and_single_vertex():
    n = 50
    row, col = _tree_csr(np.array([-1] + [0] * (n - 1)), 2)
    hh, dd, th = orc.tree(row, col, 0)
    assert th == 1 and hh[0] == 1 and dd[0] == n - 1
    assert np.all(hh[1:] == 0) and np.all(dd[1:] == 0)
    hh, dd, th = orc.tree(np.array([0, 0], np.int32), np.zeros(0, np.int32), 0)
    assert th == 0 and hh[0] == 0 and dd[0] == 0


@pytest.mark.parametrize("seed", range(6))
def test_tree_vs_brute_force(seed):
    """Against a separate BFS from every vertex: desc = reached - 1 and
    height = the deepest BFS level; vertices outside the root's tree get -1."""
    n = 60 + 17 * seed
    parent, root = _random_parent(n, seed)
    # a second tree on other ids stays unreached
    extra = 9
    parent = np.concatenate([parent, [-1] + [n + int(i) for i in np.random.default_rng(seed).integers(0, np.arange(1, extra))]])
    row, col = _tree_csr(parent, seed)
    hh, dd, th = orc.tree(row, col, root)
    for v in range(n + extra):
        if v >= n:
            assert hh[v] == -1 and dd[v] == -1
            continue
        lv = orc.bfs(row, col, v)
        reached = lv != orc.UNDEF
        assert dd[v] == reached.sum() - 1
        assert hh[v] == lv[reached].max()
    assert th == hh[root]


def test_tree_rejects_two_in_arcs_and_ignores_arcs_into_root():
    row, col = _tree_csr(np.array([-1, 0, 0, 1]), 3)
    # vertex 3 gets a second in-arc (from 2)
    row2 = row.copy()
    row2[3:] += 1
    col2 = np.insert(col, row[2], 3)
    with pytest.raises(ValueError):
        orc.tree(row2, col2, 0)
    # an arc 3 -> 0 (into the root) closes a cycle and is ignored
    row3 = row.copy()
    row3[4:] += 1
    col3 = np.append(col, 0).astype(np.int32)
    a = orc.tree(row, col, 0)
    b = orc.tree(row3, col3, 0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2] == 2
