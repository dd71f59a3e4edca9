"""Autotuner: picks a candidate parameter set, installs it, caches it by GPU
arch + degree histogram, and the tuned solution stays exact."""
import json

import numpy as np
import pytest

import orc
import smdp

pytestmark = pytest.mark.gpu


def test_autotune_sssp_cache_and_parity(ctx, tmp_path):
    g = smdp.gen_rmat(13, 16, wlo=1, whi=255, seed=3)
    cache = tmp_path / "tune.json"
    old = ctx.get_params(smdp.APP_SSSP)
    try:
        with smdp.Graph(ctx, g) as dg:
            p1 = dg.autotune(smdp.APP_SSSP, g.source, cache)
            # one of the SSSP candidates of csrc/tune.hip: bucketed or unordered rounds on the renumbered copy
            if p1.layout == smdp.LAYOUT_BUCKETS:
                assert p1.split_threshold in (128, 256, 512) and p1.wg_per_cu in (2, 4), p1.as_dict()
                assert p1.delta in (32, 64, 128), p1.as_dict()
            else:
                assert p1.layout == smdp.LAYOUT_HOT, p1.as_dict()
                assert p1.split_threshold in (512, 2048) and p1.wg_per_cu in (4, 8), p1.as_dict()
            data = json.loads(cache.read_text())
            assert len(data) == 1 and list(data)[0].startswith("gfx")
            p2 = dg.autotune(smdp.APP_SSSP, g.source, cache)  # cache hit
            assert p2.as_dict() == p1.as_dict()
            assert ctx.get_params(smdp.APP_SSSP).as_dict() == p1.as_dict()
            dg.sssp(g.source, smdp.SOL_TUNED)
            assert np.array_equal(dg.read(smdp.RES_DIST), orc.dijkstra(g.row, g.col, g.w, g.source))
    finally:
        ctx.set_params(smdp.APP_SSSP, **old.as_dict())


@pytest.mark.parametrize("app", [smdp.APP_BFS, smdp.APP_SPMV, smdp.APP_PAGERANK, smdp.APP_COLOR, smdp.APP_MSSSP])
def test_autotune_other_apps(ctx, app, tmp_path):
    g = smdp.gen_rmat(11, 16, wlo=1, whi=255, seed=5, symmetrize=(app == smdp.APP_COLOR))
    old = ctx.get_params(app)
    try:
        if app == smdp.APP_PAGERANK:
            ig = smdp.invert(g)
            with smdp.Graph(ctx, smdp.HostGraph(ig.row, ig.col)) as dg:
                dg.set_outdeg(g.outdegree())
                dg.autotune(app, 0, tmp_path / "t.json")
                st = dg.pagerank(smdp.pr_opts(), smdp.SOL_TUNED)
                r = dg.read(smdp.RES_RANK, np.float32)
                ref, _, _ = orc.pagerank(ig.row, ig.col, g.outdegree(), fixed_iters=int(st.rounds))
                assert np.abs(r - ref).sum() <= 1e-6
            return
        with smdp.Graph(ctx, g) as dg:
            if app == smdp.APP_SPMV:
                val = smdp.gen_uniform_f32(g.m, -1, 1, 2, 6)
                x = smdp.gen_uniform_f32(g.n, 0, 1, 2, 9)
                dg.set_values(val)
                dg.set_x(x)
            dg.autotune(app, g.source, tmp_path / "t.json")
            if app == smdp.APP_BFS:
                dg.bfs(g.source, smdp.SOL_TUNED)
                assert np.array_equal(dg.read(smdp.RES_LEVEL), orc.bfs(g.row, g.col, g.source))
            elif app == smdp.APP_SPMV:
                dg.spmv(smdp.SOL_TUNED)
                y = dg.read(smdp.RES_Y, np.float32)
                assert np.allclose(y, orc.spmv(g.row, g.col, val, x), rtol=1e-6, atol=1e-6)
            elif app == smdp.APP_COLOR:
                dg.color(False, smdp.SOL_TUNED)
                assert np.array_equal(dg.read(smdp.RES_COLOR), orc.color_closed(g.row, g.col)[0])
            else:
                srcs = smdp.gen_sources(g, 8, 1)
                dg.sssp_multi_source(srcs)
                d = dg.read(smdp.RES_MDIST, count=8 * g.n).reshape(8, g.n)
                for i, s in enumerate(srcs):
                    assert np.array_equal(d[i], orc.dijkstra(g.row, g.col, g.w, int(s))), i
    finally:
        ctx.set_params(app, **old.as_dict())
