"""SpMV and PageRank on the device against the FP64 oracle, every solution.

SpMV is FP32 with compensated (Dot2) accumulation.  Tolerance per row,
written here once: |y - y64| <= 2u|y64| + 2 (n_r u)^2 S_r with u = 2^-24,
n_r the row length and S_r = sum |val*x| over the row (the Dot2 error bound
of Ogita, Rump and Oishi with a factor 2 of margin).  With the reference's own
inputs (data[i] = i%13+0.5, x[i] = i%17+0.5) every partial sum is a multiple
of 0.25 below 2^22 here, so y must be exact.

PageRank runs FP32 ranks with compensated row sums; its L1 error against an
iteration-matched FP64 oracle (same iteration count) must be <= 1e-6 * T, and
the FP64 residual stop rule must fire as on the device.
"""
import numpy as np
import pytest

import orc
import smdp
from conftest import random_graph

pytestmark = pytest.mark.gpu

SOLS = [smdp.SOL_FLAT, smdp.SOL_NAIVE, smdp.SOL_WAVE, smdp.SOL_BLOCK, smdp.SOL_GRID, smdp.SOL_TUNED]
U = 2.0 ** -24


def _mats():
    out = {}
    out["rmat12"] = smdp.gen_rmat(12, 16, seed=2)
    out["rmat10_noperm"] = smdp.gen_rmat(10, 32, permute=False, seed=7)
    row, col, _ = random_graph(5000, 30000, 3)
    out["random"] = smdp.HostGraph(row, col)
    out["uniform"] = smdp.gen_uniform(8192, 8, 1, 1, 5)
    return out


MATS = _mats()


def _spmv_bound(g, val, x, y64):
    prod = np.abs(val.astype(np.float64) * x.astype(np.float64)[g.col])
    S = np.zeros(g.n)
    np.add.at(S, np.repeat(np.arange(g.n), np.diff(g.row)), prod)
    nr = np.diff(g.row)
    return 2 * U * np.abs(y64) + 2 * (nr * U) ** 2 * S


@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("sol", SOLS)
def test_spmv_parity(ctx, name, sol):
    g = MATS[name]
    print(f"spmv {name} sol={sol}", flush=True)
    val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 2, 6)
    x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 2, 9)
    y, st = ctx.spmv(g, val, x, sol)
    y64 = orc.spmv(g.row, g.col, val, x)
    err = np.abs(y.astype(np.float64) - y64)
    assert np.all(err <= _spmv_bound(g, val, x, y64)), float(err.max())
    assert st.edges_relaxed == g.m


@pytest.mark.parametrize("sol", SOLS)
def test_spmv_reference_inputs_exact(ctx, sol):
    g = MATS["rmat12"]
    data, x = smdp.gen_spmv_reference_inputs(g.n, g.m)
    y, _ = ctx.spmv(g, data, x, sol)
    four_y = np.zeros(g.n, dtype=np.int64)
    terms = (2 * (np.arange(g.m) % 13) + 1) * (2 * (g.col % 17) + 1)
    src = np.repeat(np.arange(g.n), np.diff(g.row))
    np.add.at(four_y, src, terms)
    assert np.all(four_y < 4 * 2 ** 22)
    assert np.array_equal(y.astype(np.float64) * 4, four_y.astype(np.float64))


def test_spmv_tiny_thresholds(ctx):
    g = MATS["rmat12"]
    val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 3, 6)
    x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 3, 9)
    y64 = orc.spmv(g.row, g.col, val, x)
    old = ctx.get_params(smdp.APP_SPMV)
    try:
        ctx.set_params(smdp.APP_SPMV, heavy_threshold=2, split_threshold=64, chunk=64, wg_per_cu=1)
        for sol in SOLS:
            y, _ = ctx.spmv(g, val, x, sol)
            assert np.all(np.abs(y - y64) <= _spmv_bound(g, val, x, y64)), sol
    finally:
        ctx.set_params(smdp.APP_SPMV, **old.as_dict())


def test_spmv_edge_cases(ctx):
    for g in [smdp.HostGraph([0, 0], []), smdp.HostGraph([0, 0, 3, 3], [0, 2, 2]),
              smdp.HostGraph([0, 1], [0])]:
        val = np.linspace(-1, 1, max(g.m, 1)).astype(np.float32)[:g.m]
        x = np.linspace(0.1, 1, g.n).astype(np.float32)
        for sol in SOLS:
            y, _ = ctx.spmv(g, val, x, sol)
            assert np.allclose(y, orc.spmv(g.row, g.col, val, x), rtol=1e-6, atol=1e-7)


def _merge_mats():
    """Row layouts around the tuned SpMV's tiles of 2048 merged items (row ends + nonzeros)."""
    rng = np.random.default_rng(11)
    out = []
    deg = np.zeros(20000, np.int64)
    deg[7] = 60000                       # crosses ~30 tiles
    deg[5000] = 4095
    deg[12000:12010] = 1500
    deg[15000:] = rng.integers(0, 5, 5000)
    out.append(deg)                      # long empty runs between
    out.append(np.full(3 * 1024, 1))     # every tile ends on a row end
    out.append(np.full(6000, 2047))      # rows just under a tile
    out.append(np.zeros(10000, np.int64))  # no nonzeros at all
    out.append(rng.geometric(0.02, 4000) - 1)
    mats = []
    for i, d in enumerate(out):
        row = np.concatenate([[0], np.cumsum(d)]).astype(np.int32)
        n = len(d)
        col = rng.integers(0, n, int(row[-1])).astype(np.int32)
        mats.append(smdp.HostGraph(row, col))
    return mats


@pytest.mark.parametrize("sol", SOLS)
def test_spmv_columns_that_never_occur(ctx, sol):
    """Columns are renumbered by how often they occur; x entries of columns no
    nonzero uses must never be read: NaN there leaves y finite and exact."""
    rng = np.random.default_rng(21)
    n = 6000
    deg = rng.integers(0, 9, n)
    row = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    hot = rng.integers(0, 40, int(row[-1]))            # a few frequent columns
    cold = rng.integers(40, n // 2, int(row[-1]))      # columns >= n/2 never occur
    col = np.where(rng.random(int(row[-1])) < 0.5, hot, cold).astype(np.int32)
    g = smdp.HostGraph(row, col)
    val = ((np.arange(g.m) % 7) - 3).astype(np.float32)
    x = ((np.arange(n) % 5) + 1).astype(np.float32)
    x[n // 2:] = np.nan
    exact = np.zeros(n, np.int64)
    np.add.at(exact, np.repeat(np.arange(n), deg), val.astype(np.int64) * x[:n // 2].astype(np.int64)[col])
    y, _ = ctx.spmv(g, val, x, sol)
    assert np.isfinite(y).all()
    assert np.array_equal(y.astype(np.int64), exact)


@pytest.mark.parametrize("sol", SOLS)
def test_spmv_one_column(ctx, sol):
    """Every nonzero in column 7 (the renumbering's degree range is then the
    whole nonzero count, and one x value is read): exact integer sums."""
    rng = np.random.default_rng(23)
    n = 3000
    deg = rng.integers(0, 30, n)
    row = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
    g = smdp.HostGraph(row, np.full(int(row[-1]), 7, np.int32))
    val = ((np.arange(g.m) % 5) - 2).astype(np.float32)
    x = np.full(n, np.nan, np.float32)
    x[7] = 3.0
    exact = np.zeros(n, np.int64)
    np.add.at(exact, np.repeat(np.arange(n), deg), val.astype(np.int64) * 3)
    y, _ = ctx.spmv(g, val, x, sol)
    assert np.array_equal(y.astype(np.int64), exact)


def test_spmv_tuned_tile_boundaries_exact(ctx):
    """Small integer values make every sum exact: a partial dropped or added twice changes y."""
    for i, g in enumerate(_merge_mats()):
        print(f"merge case {i}: n={g.n} m={g.m}", flush=True)
        val = ((np.arange(g.m) % 7) - 3).astype(np.float32)
        x = ((np.arange(g.n) % 5) + 1).astype(np.float32)
        exact = np.zeros(g.n, np.int64)
        np.add.at(exact, np.repeat(np.arange(g.n), np.diff(g.row)),
                  val.astype(np.int64) * x.astype(np.int64)[g.col])
        assert np.abs(exact).max(initial=0) < 2 ** 24
        y, st = ctx.spmv(g, val, x, smdp.SOL_TUNED)
        assert np.array_equal(y.astype(np.int64), exact)
        assert st.edges_relaxed == g.m and st.vertices_expanded == g.n
        rv = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 4, 6)
        rx = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 4, 9)
        y1, _ = ctx.spmv(g, rv, rx, smdp.SOL_TUNED)
        y2, _ = ctx.spmv(g, rv, rx, smdp.SOL_TUNED)
        y64 = orc.spmv(g.row, g.col, rv, rx)
        assert np.all(np.abs(y1 - y64) <= _spmv_bound(g, rv, rx, y64))
        assert np.array_equal(y1.view(np.uint32), y2.view(np.uint32))


def _pr_graph(name):
    g = MATS[name]
    ig = smdp.invert(g)
    return g, ig, g.outdegree()


@pytest.mark.parametrize("name", ["rmat12", "random", "rmat10_noperm"])
@pytest.mark.parametrize("sol", SOLS)
def test_pagerank_parity(ctx, name, sol):
    g, ig, od = _pr_graph(name)
    print(f"pagerank {name} sol={sol}", flush=True)
    opts = smdp.pr_opts(0.85, 1.0, 1e-6, 100)
    r, st = ctx.pagerank(ig, od, opts, sol)
    it = int(st.rounds)
    ref, it_ref, res_ref = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, fixed_iters=it)
    l1 = np.abs(r.astype(np.float64) - ref).sum()
    assert l1 <= 1e-6, l1
    assert st.residual < 1e-6 or it == 100
    # the FP64 oracle with the same stop rule stops within one iteration
    _, it_conv, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, tol=1e-6, max_iter=100)
    assert abs(it_conv - it) <= 1
    assert abs(r.astype(np.float64).sum() - 1.0) < 1e-5


@pytest.mark.parametrize("sol", SOLS)
def test_pagerank_reference_parameters(ctx, sol):
    # damping 0.5, total rank 1000, ranks starting at 0 (pagerank_wrapper.cu:7,115; pagerank.h:15)
    g, ig, od = _pr_graph("random")
    opts = smdp.pr_opts(0.5, 1000.0, 1e-6, 100, init_zero=True)
    r, st = ctx.pagerank(ig, od, opts, sol)
    ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.5, T=1000.0, init_zero=True, fixed_iters=int(st.rounds))
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6 * 1000.0


def test_pagerank_deterministic_and_tiny_thresholds(ctx):
    g, ig, od = _pr_graph("rmat12")
    opts = smdp.pr_opts(0.85, 1.0, 1e-6, 100)
    a, _ = ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    b, _ = ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    assert np.array_equal(a, b)
    old = ctx.get_params(smdp.APP_PAGERANK)
    try:
        ctx.set_params(smdp.APP_PAGERANK, heavy_threshold=4, split_threshold=64, chunk=64, wg_per_cu=1)
        for sol in SOLS:
            r, st = ctx.pagerank(ig, od, opts, sol)
            ref, _, _ = orc.pagerank(ig.row, ig.col, od, fixed_iters=int(st.rounds))
            assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6, sol
    finally:
        ctx.set_params(smdp.APP_PAGERANK, **old.as_dict())


def test_pagerank_max_iter_and_resident(ctx):
    g, ig, od = _pr_graph("random")
    with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
        dg.set_outdeg(od)
        st = dg.pagerank(smdp.pr_opts(0.85, 1.0, 0.0, 7), smdp.SOL_TUNED)
        assert st.rounds == 7
        r = dg.read(smdp.RES_RANK, np.float32)
        ref, _, _ = orc.pagerank(ig.row, ig.col, od, fixed_iters=7)
        assert np.abs(r - ref).sum() <= 1e-6


@pytest.mark.parametrize("name", sorted(MATS))
def test_column_blocked_layout(ctx, name):
    """Tuned solution with the sources split into 8 XCD blocks (forced on
    small inputs; chosen automatically once the gathered vector outgrows an
    L2)."""
    g = MATS[name]
    old_s, old_p = ctx.get_params(smdp.APP_SPMV), ctx.get_params(smdp.APP_PAGERANK)
    try:
        for app in (smdp.APP_SPMV, smdp.APP_PAGERANK):
            ctx.set_params(app, layout=smdp.LAYOUT_XCD8, split_threshold=256, heavy_threshold=16)
        val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, 4, 6)
        x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 4, 9)
        y, _ = ctx.spmv(g, val, x, smdp.SOL_TUNED)
        y64 = orc.spmv(g.row, g.col, val, x)
        assert np.all(np.abs(y - y64) <= _spmv_bound(g, val, x, y64))
        data, xr = smdp.gen_spmv_reference_inputs(g.n, g.m)
        yr, _ = ctx.spmv(g, data, xr, smdp.SOL_TUNED)
        assert np.array_equal(yr.astype(np.float64), orc.spmv(g.row, g.col, data, xr))
        ig = smdp.invert(g)
        od = g.outdegree()
        r1, st = ctx.pagerank(ig, od, smdp.pr_opts(), smdp.SOL_TUNED)
        ref, _, _ = orc.pagerank(ig.row, ig.col, od, fixed_iters=int(st.rounds))
        assert np.abs(r1 - ref).sum() <= 1e-6
        r2, _ = ctx.pagerank(ig, od, smdp.pr_opts(), smdp.SOL_TUNED)
        assert np.array_equal(r1, r2)
    finally:
        ctx.set_params(smdp.APP_SPMV, **old_s.as_dict())
        ctx.set_params(smdp.APP_PAGERANK, **old_p.as_dict())


def test_values_change_rebuilds_blocked_layout(ctx):
    g = MATS["rmat12"]
    old = ctx.get_params(smdp.APP_SPMV)
    try:
        ctx.set_params(smdp.APP_SPMV, layout=smdp.LAYOUT_XCD8)
        x = smdp.gen_uniform_f32(g.n, 0.0, 1.0, 5, 9)
        with smdp.Graph(ctx, g) as dg:
            dg.set_x(x)
            for seed in (1, 2):
                val = smdp.gen_uniform_f32(g.m, -1.0, 1.0, seed, 6)
                dg.set_values(val)
                dg.spmv(smdp.SOL_TUNED)
                y = dg.read(smdp.RES_Y, np.float32)
                y64 = orc.spmv(g.row, g.col, val, x)
                assert np.all(np.abs(y - y64) <= _spmv_bound(g, val, x, y64))
    finally:
        ctx.set_params(smdp.APP_SPMV, **old.as_dict())


def _wide_graph():
    """Four source chunks of 2^15 ids (the last one partial), a row with 70000
    in-arcs (cut into several pieces), a row with 300000 in-arcs all from the
    first chunk (one (chunk, row) segment longer than a scatter work item of
    262144 arcs), a
    row at and one just below the wide-row degree 4096, empty rows and
    dangling vertices."""
    rng = np.random.default_rng(11)
    n = 3 * 32768 + 1000
    parts_src = [rng.integers(0, n, 300000), rng.integers(0, 32768, 300000)]
    parts_dst = [rng.integers(0, n, 300000), np.full(300000, 6)]
    for v, k in ((5, 70000), (n - 1, 4096), (7, 4095)):
        parts_src.append(rng.integers(0, n, k))
        parts_dst.append(np.full(k, v))
    src = np.concatenate(parts_src)
    dst = np.concatenate(parts_dst)
    keep = (src % 97) != 3  # about 1% of the vertices keep no out-arc: dangling
    src, dst = src[keep], dst[keep]
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row, src + 1, 1)
    return smdp.HostGraph(np.cumsum(row).astype(np.int32), dst[order].astype(np.int32))


def _empty_head_graph():
    """Every arc points into the last 100 vertices: the first 4900 rows of the
    in-CSR are empty, so the first bins hold rows but no arcs, the first of
    them starting at arc 0."""
    rng = np.random.default_rng(12)
    n = 5000
    src = rng.integers(0, n, 20000)
    dst = rng.integers(n - 100, n, 20000)
    order = np.argsort(src, kind="stable")
    row = np.zeros(n + 1, dtype=np.int32)
    np.add.at(row, src + 1, 1)
    return smdp.HostGraph(np.cumsum(row).astype(np.int32), dst[order].astype(np.int32))


PR_BIN = {"wide": _wide_graph(), "empty_head": _empty_head_graph(),
          **{k: MATS[k] for k in ("rmat12", "random", "rmat10_noperm", "uniform")}}


@pytest.fixture
def binned_ctx(ctx):
    old = ctx.get_params(smdp.APP_PAGERANK)
    ctx.set_params(smdp.APP_PAGERANK, layout=smdp.LAYOUT_BINNED)
    yield ctx
    ctx.set_params(smdp.APP_PAGERANK, **old.as_dict())


@pytest.mark.parametrize("name", sorted(PR_BIN))
def test_pagerank_binned_layout(binned_ctx, name):
    """Tuned PageRank on the binned layout (scatter + reduce passes) against
    the iteration-matched FP64 oracle; two runs are bit-identical."""
    ctx = binned_ctx
    g = PR_BIN[name]
    ig, od = smdp.invert(g), g.outdegree()
    print(f"pagerank binned {name} n={g.n} m={g.m}", flush=True)
    opts = smdp.pr_opts(0.85, 1.0, 1e-6, 100)
    r, st = ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    it = int(st.rounds)
    ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, fixed_iters=it)
    l1 = np.abs(r.astype(np.float64) - ref).sum()
    assert l1 <= 1e-6, l1
    assert st.residual < 1e-6 or it == 100
    _, it_conv, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, tol=1e-6, max_iter=100)
    assert abs(it_conv - it) <= 1
    assert st.edges_relaxed == g.m * it
    assert st.vertices_expanded == g.n * it
    r2, _ = ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    assert np.array_equal(r, r2)


def test_pagerank_binned_reference_parameters(binned_ctx):
    # damping 0.5, total rank 1000, ranks starting at 0 (pagerank_wrapper.cu:7,115; pagerank.h:15)
    g = PR_BIN["wide"]
    ig, od = smdp.invert(g), g.outdegree()
    opts = smdp.pr_opts(0.5, 1000.0, 1e-6, 100, init_zero=True)
    r, st = binned_ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.5, T=1000.0, init_zero=True, fixed_iters=int(st.rounds))
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6 * 1000.0


def test_pagerank_binned_resident_reuse(binned_ctx):
    """The layout is built once per resident graph; later solves reuse it and
    give the same ranks as a fresh drop-in call."""
    ctx = binned_ctx
    g = PR_BIN["wide"]
    ig, od = smdp.invert(g), g.outdegree()
    opts = smdp.pr_opts(0.85, 1.0, 0.0, 9)
    fresh, _ = ctx.pagerank(ig, od, opts, smdp.SOL_TUNED)
    with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
        dg.set_outdeg(od)
        for _ in range(3):
            st = dg.pagerank(opts, smdp.SOL_TUNED)
            assert st.rounds == 9
            assert np.array_equal(dg.read(smdp.RES_RANK, np.float32), fresh)


def test_binned_layout_is_pagerank_only(ctx):
    with pytest.raises(smdp.SmdpError):
        ctx.set_params(smdp.APP_SPMV, layout=smdp.LAYOUT_BINNED)
    with pytest.raises(smdp.SmdpError):
        ctx.set_params(smdp.APP_PAGERANK, layout=3)


def test_pagerank_iterations_are_fed_ahead(binned_ctx):
    """The last workgroup of each iteration reports it to host memory and the
    host keeps at most 3 + 4 iterations queued ahead of the one reported, so
    a converging solve enqueues at most 7 iterations past the one where its
    stop rule fires (doubling batches of 4, 8, 16 would enqueue 28 for an
    18-iteration solve).  Launches: one initialisation, then a scatter and a
    reduce per enqueued iteration."""
    g = smdp.gen_rmat(16, 16, seed=3)
    ig, od = smdp.invert(g), g.outdegree()
    r, st = binned_ctx.pagerank(ig, od, smdp.pr_opts(0.85, 1.0, 1e-6, 100), smdp.SOL_TUNED)
    it = int(st.rounds)
    enqueued = (int(st.kernel_launches) - 1) // 2
    print(f"pagerank rmat16: {it} iterations, {enqueued} enqueued", flush=True)
    assert 12 <= it < 100
    assert it <= enqueued <= it + 7, (it, enqueued)
    ref, _, _ = orc.pagerank(ig.row, ig.col, od, d=0.85, T=1.0, fixed_iters=it)
    assert np.abs(r.astype(np.float64) - ref).sum() <= 1e-6
