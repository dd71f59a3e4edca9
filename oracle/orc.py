"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_GLOBAL = os.path.join(HERE, "_ref", "libref_global.so")

INF = 1073741824
UNDEF = 0x7FFFFFFF

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i = C.c_void_p, C.c_int
        L.orc_dijkstra.argtypes = [i, vp, vp, vp, i, vp]
        L.orc_bellman_ford.argtypes = [i, vp, vp, vp, i, vp]
        L.orc_sssp_worklist.argtypes = [i, vp, vp, vp, i, vp, C.POINTER(C.c_longlong)]
        L.orc_sssp_worklist.restype = C.c_longlong
        L.orc_sssp_worklist_par.argtypes = [i, vp, vp, vp, i, vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.orc_sssp_worklist_par.restype = C.c_longlong
        L.orc_bfs_fifo.argtypes = [i, vp, vp, i, vp]
        L.orc_color_rounds.argtypes = [i, vp, vp, vp, vp]
        L.orc_color_closed.argtypes = [i, vp, vp, vp, vp]
        L.orc_spmv_f64.argtypes = [i, vp, vp, vp, vp, vp]
        L.orc_spmv_f64_par.argtypes = [i, vp, vp, vp, vp, vp]
        L.orc_pagerank_f64_par.argtypes = [i, vp, vp, vp, C.c_double, C.c_double, i, i, vp, C.POINTER(C.c_double)]
        L.orc_bfs_par.argtypes = [i, vp, vp, i, vp]
        L.orc_tree.argtypes = [i, vp, vp, i, vp, vp]
        L.orc_color_layers_par.argtypes = [i, vp, vp, vp, vp, C.POINTER(C.c_int)]
        L.orc_pagerank_f64.argtypes = [i, vp, vp, vp, C.c_double, C.c_double, i, C.c_double, i, i, vp,
                                       C.POINTER(C.c_int), C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def dijkstra(row, col, w, src):
    row, col, w = _i32(row), _i32(col), _i32(w)
    d = np.zeros(len(row) - 1, dtype=np.int32)
    assert lib().orc_dijkstra(len(row) - 1, _p(row), _p(col), _p(w), src, _p(d)) == 0
    return d


def bellman_ford(row, col, w, src):
    row, col, w = _i32(row), _i32(col), _i32(w)
    d = np.zeros(len(row) - 1, dtype=np.int32)
    assert lib().orc_bellman_ford(len(row) - 1, _p(row), _p(col), _p(w), src, _p(d)) == 0
    return d


def sssp_worklist(row, col, w, src):
    """Reference algorithm (worklist Bellman-Ford); returns (dist, rounds, edges)."""
    row, col, w = _i32(row), _i32(col), _i32(w)
    d = np.zeros(len(row) - 1, dtype=np.int32)
    edges = C.c_longlong()
    rounds = lib().orc_sssp_worklist(len(row) - 1, _p(row), _p(col), _p(w), src, _p(d), C.byref(edges))
    assert rounds >= 0
    return d, int(rounds), int(edges.value)


def sssp_worklist_par(row, col, w, src):
    """The same rounds on all host threads; returns (dist, rounds, edges, threads)."""
    row, col, w = _i32(row), _i32(col), _i32(w)
    d = np.zeros(len(row) - 1, dtype=np.int32)
    edges = C.c_longlong()
    threads = C.c_int()
    rounds = lib().orc_sssp_worklist_par(len(row) - 1, _p(row), _p(col), _p(w), src, _p(d), C.byref(edges),
                                         C.byref(threads))
    assert rounds >= 0
    return d, int(rounds), int(edges.value), int(threads.value)


def bfs(row, col, src):
    row, col = _i32(row), _i32(col)
    lv = np.zeros(len(row) - 1, dtype=np.int32)
    assert lib().orc_bfs_fifo(len(row) - 1, _p(row), _p(col), src, _p(lv)) == 0
    return lv


def tree(row, col, root):
    """Subtree heights and descendant counts of a rooted tree (child lists);
    returns (height, desc, tree_height), or raises ValueError when some vertex
    has two in-arcs."""
    row, col = _i32(row), _i32(col)
    n = len(row) - 1
    h = np.zeros(n, dtype=np.int32)
    d = np.zeros(n, dtype=np.int32)
    r = lib().orc_tree(n, _p(row), _p(col), root, _p(h), _p(d))
    if r == -3:
        raise ValueError("not a tree: a vertex has two in-arcs")
    assert r >= 0
    return h, d, int(r)


def color_rounds(row, col, prio=None):
    row, col = _i32(row), _i32(col)
    prio = None if prio is None else np.ascontiguousarray(prio, dtype=np.uint32)
    c = np.zeros(len(row) - 1, dtype=np.int32)
    k = lib().orc_color_rounds(len(row) - 1, _p(row), _p(col), _p(prio), _p(c))
    assert k >= 0
    return c, k


def color_closed(row, col, prio=None):
    row, col = _i32(row), _i32(col)
    prio = None if prio is None else np.ascontiguousarray(prio, dtype=np.uint32)
    c = np.zeros(len(row) - 1, dtype=np.int32)
    k = lib().orc_color_closed(len(row) - 1, _p(row), _p(col), _p(prio), _p(c))
    assert k >= 0
    return c, k


def spmv(row, col, val, x):
    row, col = _i32(row), _i32(col)
    val = np.ascontiguousarray(val, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(len(row) - 1, dtype=np.float64)
    assert lib().orc_spmv_f64(len(row) - 1, _p(row), _p(col), _p(val), _p(x), _p(y)) == 0
    return y


def pagerank(in_row, in_col, outdeg, d=0.85, T=1.0, init_zero=False, tol=1e-6, max_iter=100, fixed_iters=0):
    """FP64 pull PageRank; returns (rank, iterations, last residual)."""
    in_row, in_col, outdeg = _i32(in_row), _i32(in_col), _i32(outdeg)
    n = len(in_row) - 1
    r = np.zeros(n, dtype=np.float64)
    it = C.c_int()
    res = C.c_double()
    assert lib().orc_pagerank_f64(n, _p(in_row), _p(in_col), _p(outdeg), d, T, int(init_zero), tol, max_iter,
                                  fixed_iters, _p(r), C.byref(it), C.byref(res)) == 0
    return r, it.value, res.value


# ---- OpenMP twins, for timing the CPU baseline on all host cores -----------

def spmv_par(row, col, val, x):
    """orc_spmv_f64 on all host threads; returns (y, threads)."""
    row, col = _i32(row), _i32(col)
    val = np.ascontiguousarray(val, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros(len(row) - 1, dtype=np.float64)
    nt = lib().orc_spmv_f64_par(len(row) - 1, _p(row), _p(col), _p(val), _p(x), _p(y))
    assert nt > 0
    return y, nt


def pagerank_par(in_row, in_col, outdeg, iters, d=0.85, T=1.0, init_zero=False):
    """`iters` FP64 pull iterations on all host threads; returns (rank, last residual, threads)."""
    in_row, in_col, outdeg = _i32(in_row), _i32(in_col), _i32(outdeg)
    n = len(in_row) - 1
    r = np.zeros(n, dtype=np.float64)
    res = C.c_double()
    nt = lib().orc_pagerank_f64_par(n, _p(in_row), _p(in_col), _p(outdeg), d, T, int(init_zero), iters, _p(r),
                                    C.byref(res))
    assert nt > 0
    return r, res.value, nt


def bfs_par(row, col, src):
    """Level-synchronous BFS on all host threads; returns (level, threads)."""
    row, col = _i32(row), _i32(col)
    lv = np.zeros(len(row) - 1, dtype=np.int32)
    nt = lib().orc_bfs_par(len(row) - 1, _p(row), _p(col), src, _p(lv))
    assert nt > 0
    return lv, nt


def color_layers_par(row, col, prio=None):
    """The closed-form colouring by layers on all host threads; returns (color, ncolors, threads)."""
    row, col = _i32(row), _i32(col)
    prio = None if prio is None else np.ascontiguousarray(prio, dtype=np.uint32)
    c = np.zeros(len(row) - 1, dtype=np.int32)
    k = C.c_int()
    nt = lib().orc_color_layers_par(len(row) - 1, _p(row), _p(col), _p(prio), _p(c), C.byref(k))
    assert nt > 0
    return c, k.value, nt


# ---- the reference's own host graph library (built from its sources) --------

class _RefGraph(C.Structure):
    # struct _GRAPH_, common/global.h:28-42
    _fields_ = [("vertexArray", C.POINTER(C.c_int)), ("costArray", C.POINTER(C.c_int)),
                ("levelArray", C.POINTER(C.c_int)), ("edgeArray", C.POINTER(C.c_int)),
                ("weightArray", C.POINTER(C.c_int)), ("frontier", C.c_void_p), ("visited", C.c_void_p),
                ("update", C.c_void_p), ("childVertexArray", C.POINTER(C.c_int)),
                ("rEdgeArray", C.POINTER(C.c_int)), ("outdegreeArray", C.POINTER(C.c_int)),
                ("rankArray", C.c_void_p), ("colorArray", C.c_void_p)]


def ref_available():
    return os.path.exists(REF_GLOBAL)


def ref_read(path, fmt, invert=False):
    """Runs the reference's readInput* + convertCSR (+ invertGraph) on a file.

    Loads a fresh copy of the library per call (its state lives in globals).
    Returns dict(n, m, source, row, col, w[, in_row, in_col]).
    """
    import shutil
    import tempfile
    tmpdir = tempfile.mkdtemp()
    try:
        copy = os.path.join(tmpdir, "libref.so")
        shutil.copy(REF_GLOBAL, copy)
        L = C.CDLL(copy)
        libc = C.CDLL(None)
        libc.fopen.restype = C.c_void_p
        libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        libc.fclose.argtypes = [C.c_void_p]
        f = libc.fopen(str(path).encode(), b"r")
        assert f
        C.c_void_p.in_dll(L, "fp").value = f
        reader = {0: "_Z16readInputDIMACS9v", 1: "_Z17readInputDIMACS10v", 2: "_Z14readInputSLNDCv"}[fmt]
        getattr(L, reader)()
        libc.fclose(f)
        getattr(L, "_Z10convertCSRv")()
        n = C.c_int.in_dll(L, "noNodeTotal").value
        m = C.c_int.in_dll(L, "noEdgeTotal").value
        src = C.c_int.in_dll(L, "source").value
        g = _RefGraph.in_dll(L, "graph")
        out = {"n": n, "m": m, "source": src,
               "row": np.ctypeslib.as_array(g.vertexArray, (n + 1,)).copy(),
               "col": np.ctypeslib.as_array(g.edgeArray, (max(m, 1),))[:m].copy(),
               "w": np.ctypeslib.as_array(g.weightArray, (max(m, 1),))[:m].copy()}
        if invert:
            g.childVertexArray = C.POINTER(C.c_int)()
            g.rEdgeArray = C.POINTER(C.c_int)()
            getattr(L, "_Z11invertGraphv")()
            g = _RefGraph.in_dll(L, "graph")
            out["in_row"] = np.ctypeslib.as_array(g.childVertexArray, (n + 1,)).copy()
            out["in_col"] = np.ctypeslib.as_array(g.rEdgeArray, (max(m, 1),))[:m].copy()
        return out
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
