"""Tree Height / Tree Descendant (listed in the reference's README.md:12-13,
no reference code): exact against the CPU oracle (orc.tree, pinned in
test_oracle.py by closed forms and brute force) on trees with hubs, chains,
stars, a cycle through the root and unreached parts; non-trees rejected."""
import numpy as np
import pytest

import orc
import smdp

pytestmark = pytest.mark.gpu


def _tree_graph(parent, seed, extra_arcs=()):
    """HostGraph of the child lists of parent[] (children shuffled), plus
    extra (u, v) arcs appended to row u."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    for v, p in enumerate(parent):
        if p >= 0:
            kids[int(p)].append(v)
    for u, v in extra_arcs:
        kids[u].append(v)
    rng = np.random.default_rng(seed)
    for k in kids:
        rng.shuffle(k)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum([len(k) for k in kids])
    col = np.array([c for k in kids for c in k], np.int32)
    return smdp.HostGraph(row, col)


def _random_parent(n, seed, hub_share=0.0):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    parent = np.full(n, -1)
    for i in range(1, n):
        # a share of the vertices hang below the first one: a hub
        j = 0 if rng.random() < hub_share else rng.integers(0, i)
        parent[perm[i]] = perm[j]
    return parent, int(perm[0])


def _check(ctx, g, root):
    h, d, th, st = ctx.tree(g, root)
    rh, rd, rth = orc.tree(g.row, g.col, root)
    assert np.array_equal(h, rh)
    assert np.array_equal(d, rd)
    assert th == rth
    return st


@pytest.mark.parametrize("n,seed,hub", [(1, 0, 0.0), (2, 1, 0.0), (257, 2, 0.0), (5000, 3, 0.0), (20000, 4, 0.5),
                                        (70000, 5, 0.02)])
def test_tree_random(ctx, n, seed, hub):
    print(f"tree random n={n} seed={seed}", flush=True)
    parent, root = _random_parent(n, seed, hub)
    st = _check(ctx, _tree_graph(parent, seed), root)
    assert st.vertices_expanded == n and st.m_reached == n - 1


def test_tree_chain_star_and_kary(ctx):
    print("tree chain/star/kary", flush=True)
    n = 3000
    _check(ctx, _tree_graph(np.arange(-1, n - 1), 1), 0)  # chain: 3000 levels
    _check(ctx, _tree_graph(np.arange(-1, 499), 4, extra_arcs=[(499, 0)]), 0)  # chain closed through the root
    _check(ctx, _tree_graph(np.array([-1] + [0] * 40000), 2), 0)  # one hub of 40000 children
    k, h = 3, 8
    m = sum(k ** t for t in range(h + 1))
    _check(ctx, _tree_graph(np.array([-1] + [(v - 1) // k for v in range(1, m)]), 3), 0)


@pytest.mark.parametrize("n", [1, 2, 6, 7, 8, 9, 17, 24, 25])
def test_tree_depths_around_the_batches(ctx, n):
    """Chains of n levels: the down launches go out in batches (8, then 16, ...)
    with a host check of the depth between batches, so the empty level that
    ends the down pass falls inside, at the end of or just past a batch."""
    print(f"tree chain n={n}", flush=True)
    _check(ctx, _tree_graph(np.arange(-1, n - 1), 1), 0)


def test_tree_unreached_part_and_cycle_through_root(ctx):
    print("tree unreached and cycle", flush=True)
    parent, root = _random_parent(3000, 7)
    # a second tree on ids 3000.. is not reached; a leaf points back at the root
    extra = np.array([-1] + [3000 + i // 2 for i in range(199)])
    parent = np.concatenate([parent, extra])
    leaf = int(np.setdiff1d(np.arange(3000), parent[parent >= 0])[0])
    g = _tree_graph(parent, 7, extra_arcs=[(leaf, root)])
    _check(ctx, g, root)
    h, d, th, _ = ctx.tree(g, root)
    assert np.all(h[3000:] == -1) and np.all(d[3000:] == -1)
    # rooted in the other tree: the first one is unreached
    _check(ctx, g, 3000)


def test_tree_rejects_non_trees_and_bad_roots(ctx):
    g = _tree_graph(np.array([-1, 0, 0, 1]), 1, extra_arcs=[(2, 3)])  # vertex 3 has two in-arcs
    with pytest.raises(smdp.SmdpError) as e:
        ctx.tree(g, 0)
    assert e.value.code == smdp.ERR_ARG
    ok = _tree_graph(np.array([-1, 0, 0, 1]), 1)
    for root in (-1, 4):
        with pytest.raises(smdp.SmdpError) as e:
            ctx.tree(ok, root)
        assert e.value.code == smdp.ERR_ARG


def test_tree_resident_zipf_4m(ctx):
    """The 4M-node Zipf(1.5) tree: exact against the oracle from every root
    tried, and consistent with BFS (tree height = deepest level, root's
    descendants = vertices reached - 1)."""
    print("tree zipf 4m", flush=True)
    g = smdp.gen_zipf_tree(1 << 22, 1.5, 4096, 5)
    with smdp.Graph(ctx, g) as dg:
        for root in (0, 1, 12345):
            st = dg.tree(root)
            h = dg.read(smdp.RES_HEIGHT)
            d = dg.read(smdp.RES_DESC)
            rh, rd, rth = orc.tree(g.row, g.col, root)
            assert np.array_equal(h, rh) and np.array_equal(d, rd), root
            dg.bfs(root, smdp.SOL_TUNED)
            lv = dg.read(smdp.RES_LEVEL)
            reached = lv != orc.UNDEF
            assert rth == lv[reached].max() and rd[root] == reached.sum() - 1
            assert st.rounds == rth  # levels holding vertices with children (the root has some)
        with pytest.raises(smdp.SmdpError):
            dg.read(smdp.RES_DESC)  # the last run was a BFS
