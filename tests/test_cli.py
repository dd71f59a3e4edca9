"""The command-line drivers: reference flags, messages and output formats."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG

BIN = os.path.join(PKG, "bin")


def run(app, *args):
    return subprocess.run([os.path.join(BIN, app), *args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("app", ["gpu-sssp", "gpu-SpMV", "gpu-pg", "gpu-graph-color", "gpu-bfs-rec"])
def test_usage_and_input_errors(app, tmp_path):
    r = run(app)
    assert r.returncode == 0 and "Usage:" in r.stderr          # no arguments: usage, exit 0
    assert "Usage:" in run(app, "-h").stderr
    r = run(app, "-i", str(tmp_path / "missing.gr"))
    assert r.returncode == 1 and "Error: NULL file pointer." in r.stderr
    bad = tmp_path / "g.gr"
    bad.write_text("p sp 1 0\n")
    r = run(app, "-i", str(bad), "-f", "9")
    assert r.returncode == 1 and "Wrong code for dataset" in r.stderr


def test_smdp_gen_writes_dimacs(tmp_path):
    out = tmp_path / "u.gr"
    r = run("smdp-gen", "uniform", "1000", "8", "1", "100", "1", str(out))
    assert r.returncode == 0
    text = out.read_text()
    assert "p sp 1000 8000" in text and "s 1" in text


def _gpu_run(app, path, *extra):
    r = run(app, "-i", path, "-f", "0", *extra)
    assert r.returncode == 0, r.stderr
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("sol", ["0", "1", "2", "3", "4", "6"])
def test_gpu_sssp_cli(golden, data_test_path, sol):
    r = _gpu_run("gpu-sssp", data_test_path, "-s", sol, "-d")
    got = re.findall(r"From Node (\d+) -> (\d+) is (\d+)", r.stdout)
    assert len(got) == 20 and all(int(a) == golden["source"] + 1 for a, _, _ in got)
    assert [int(c) for _, _, c in got] == golden["dist"]


@pytest.mark.gpu
@pytest.mark.parametrize("sol", ["0", "1", "2", "3", "4", "5", "6"])
def test_gpu_bfs_cli(golden, data_test_path, sol):
    r = _gpu_run("gpu-bfs-rec", data_test_path, "-s", sol, "-d")
    assert "PASS: GPU bfs rec !" in r.stdout
    lv = [int(x) for x in re.findall(r"Node \d+ is in level (\d+)", r.stdout)]
    assert lv == golden["level"]


@pytest.mark.gpu
def test_gpu_color_cli(golden, data_test_path):
    r = _gpu_run("gpu-graph-color", data_test_path, "-s", "3", "-d")
    assert [int(x) for x in re.findall(r"Node \d+: (\d+)", r.stdout)] == golden["color"]


@pytest.mark.gpu
def test_gpu_spmv_and_pagerank_cli(data_test_path):
    r = _gpu_run("gpu-SpMV", data_test_path, "-s", "4", "-v")
    err = float(re.search(r"L-inf error vs FP64: (\S+)", r.stdout).group(1))
    assert err == 0.0  # the reference's inputs make every sum exact
    for sol in ("0", "4", "6"):
        r = _gpu_run("gpu-pg", data_test_path, "-s", sol, "-v", "-d")
        ranks = np.array([float(x) for x in re.findall(r"Node \d+: (\S+)", r.stdout)])
        assert len(ranks) == 20 and abs(ranks.sum() - 1000.0) < 1e-2
        # FP32 ranks of size ~50 against the FP64 host iterates: a few ulps each
        m = re.search(r"vs FP64 host \((\d+) iterations\): L1 (\S+)  L-inf (\S+)", r.stdout)
        assert m and int(m.group(1)) > 0, r.stdout
        assert float(m.group(2)) < 1e-3 and float(m.group(3)) < 1e-4, m.group(0)


@pytest.mark.gpu
@pytest.mark.parametrize("sol", ["2", "3", "4"])
def test_gpu_sssp_cli_counters(golden, data_test_path, sol):
    """-v prints the heavy-work queue counters (the reference's GPU_PROFILE
    nested-call count, apps/sssp/src/sssp_kernel.cu:14-25); a heavy degree of
    1 sends every expanded vertex with arcs through the queues."""
    r = _gpu_run("gpu-sssp", data_test_path, "-s", sol, "-v", "-d", "--threshold", "1")
    got = re.findall(r"From Node (\d+) -> (\d+) is (\d+)", r.stdout)
    assert [int(c) for _, _, c in got] == golden["dist"]
    m = re.search(r"Heavy work queued: wave (\d+)  workgroup (\d+)  grid (\d+)  \(next-frontier pushes (\d+)\)",
                  r.stdout)
    assert m, r.stdout
    assert sum(int(x) for x in m.groups()[:3]) > 0
    k = re.search(r"Kernel launches: (\d+)  atomics tried: (\d+)  won: (\d+)", r.stdout)
    assert k and int(k.group(1)) > 0 and int(k.group(3)) <= int(k.group(2))


@pytest.mark.gpu
@pytest.mark.parametrize("sol", ["0", "3", "6"])
def test_gpu_sssp_cli_uniform16k_gr(tmp_path, sol):
    """The reference's small SSSP case end to end through files: 16384
    vertices with 8 uniform out-arcs each, weights U[1,100], written as a
    DIMACS9 .gr by smdp-gen and read back by gpu-sssp -f 0 (sssp.cpp:100-154);
    every distance printed matches Dijkstra on the same file."""
    import orc
    import smdp
    path = tmp_path / "uniform16k.gr"
    assert run("smdp-gen", "uniform", "16384", "8", "1", "100", "1", str(path)).returncode == 0
    g = smdp.read_graph(str(path), 0)
    assert (g.n, g.m, g.source) == (16384, 131072, 0)
    r = _gpu_run("gpu-sssp", str(path), "-s", sol, "-d")
    got = re.findall(r"From Node (\d+) -> (\d+) is (\d+)", r.stdout)
    assert len(got) == g.n
    ref = orc.dijkstra(g.row, g.col, g.w, g.source)
    dist = np.array([int(c) for _, _, c in got])
    to = np.array([int(b) for _, b, _ in got]) - 1
    assert np.array_equal(dist, ref[to])
