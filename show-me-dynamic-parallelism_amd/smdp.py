"""ctypes binding of libsmdp.so (the C ABI declared in include/smdp.h).

Used by the tests and the benchmark.  It loads the in-tree library and fails
loudly when it has not been built: there is no CPU fallback for any compute
call.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmdp.so")

OK = 0
ERR_ARG, ERR_NOMEM, ERR_DEVICE, ERR_SOLUTION, ERR_OVERFLOW, ERR_IO, ERR_FORMAT, ERR_NOT_BUILT = (
    -1, -2, -3, -4, -5, -6, -7, -8)

SOL_FLAT, SOL_NAIVE, SOL_WAVE, SOL_BLOCK, SOL_GRID, SOL_HIER, SOL_TUNED = range(7)
SOLUTIONS = {"flat": 0, "naive": 1, "wave": 2, "block": 3, "grid": 4, "hier": 5, "tuned": 6}
APP_SSSP, APP_BFS, APP_SPMV, APP_PAGERANK, APP_COLOR, APP_MSSSP = range(6)
LAYOUT_AUTO, LAYOUT_CSR, LAYOUT_BINNED, LAYOUT_HOT, LAYOUT_BUCKETS, LAYOUT_XCD8 = 0, 1, 2, 3, 4, 8
SETTLE_AUTO, SETTLE_PUSH, SETTLE_PULL = 0, 1, 2
RES_DIST, RES_LEVEL, RES_Y, RES_RANK, RES_COLOR, RES_MDIST, RES_HEIGHT, RES_DESC = range(8)
FMT_DIMACS9, FMT_DIMACS10, FMT_SLNDC = range(3)

SSSP_INF = 1073741824
BFS_UNDEF = 0x7FFFFFFF


class SmdpError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code}) {last_hip_error()}")


class Csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("row", C.c_void_p), ("col", C.c_void_p),
                ("w", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [("heavy_threshold", C.c_int32), ("split_threshold", C.c_int32), ("chunk", C.c_int32),
                ("wg_per_cu", C.c_int32), ("batch_rounds", C.c_int32), ("max_rounds", C.c_int32),
                ("layout", C.c_int32), ("delta", C.c_int32), ("persistent", C.c_int32),
                ("group_share", C.c_int32), ("settle", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("ms_kernel", C.c_double), ("ms_total", C.c_double), ("edges_relaxed", C.c_uint64),
                ("vertices_expanded", C.c_uint64), ("rounds", C.c_uint64), ("q_push", C.c_uint64 * 4),
                ("m_reached", C.c_uint64), ("residual", C.c_double), ("dangling", C.c_double),
                ("kernel_launches", C.c_uint64), ("ms_dominant", C.c_double),
                ("dominant_launches", C.c_uint64), ("ms_phase2", C.c_double), ("phase2_launches", C.c_uint64),
                ("edges_phase2", C.c_uint64), ("atomics_tried", C.c_uint64), ("atomics_won", C.c_uint64),
                ("arcs_skipped", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["q_push"] = list(self.q_push)
        return d


class PrOpts(C.Structure):
    _fields_ = [("damping", C.c_float), ("total_rank", C.c_float), ("tol_l1", C.c_double),
                ("max_iter", C.c_int32), ("init_zero", C.c_int32)]


class HGraphStruct(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("row", C.POINTER(C.c_int32)),
                ("col", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_int32)), ("source", C.c_int32)]


_lib = None


def lib():
    """The loaded library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or make -C "
                               f"{HERE}")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    i32, vp, cp = C.c_int32, C.c_void_p, C.c_char_p
    P = C.POINTER
    sig = {
        "smdp_strerror": (cp, [C.c_int]),
        "smdp_last_hip_error": (cp, []),
        "smdp_create": (C.c_int, [C.c_int, P(vp)]),
        "smdp_destroy": (None, [vp]),
        "smdp_get_params": (C.c_int, [vp, C.c_int, P(Params)]),
        "smdp_set_params": (C.c_int, [vp, C.c_int, P(Params)]),
        "smdp_device_info": (C.c_int, [vp, cp, C.c_int, cp, C.c_int, P(i32)]),
        "smdp_set_profiling": (C.c_int, [vp, C.c_int]),
        "smdp_round_log": (C.c_int, [vp, vp, vp, vp, i32]),
        "smdp_sssp": (C.c_int, [vp, P(Csr), i32, C.c_int, vp, P(Stats)]),
        "smdp_bfs": (C.c_int, [vp, P(Csr), i32, C.c_int, vp, P(Stats)]),
        "smdp_tree": (C.c_int, [vp, P(Csr), i32, vp, vp, vp, P(Stats)]),
        "smdp_spmv_f32": (C.c_int, [vp, P(Csr), vp, vp, C.c_int, vp, P(Stats)]),
        "smdp_pagerank_f32": (C.c_int, [vp, P(Csr), vp, P(PrOpts), C.c_int, vp, P(Stats)]),
        "smdp_color": (C.c_int, [vp, P(Csr), vp, C.c_int, vp, P(Stats)]),
        "smdp_sssp_multi_source": (C.c_int, [vp, P(Csr), vp, i32, vp, P(Stats)]),
        "smdp_sssp_multi_graph": (C.c_int, [vp, P(Csr), i32, vp, vp, P(Stats)]),
        "smdp_graph_create": (C.c_int, [vp, P(Csr), P(vp)]),
        "smdp_graph_destroy": (None, [vp]),
        "smdp_graph_set_values": (C.c_int, [vp, vp]),
        "smdp_graph_set_x": (C.c_int, [vp, vp]),
        "smdp_graph_set_outdeg": (C.c_int, [vp, vp]),
        "smdp_graph_set_prio": (C.c_int, [vp, vp]),
        "smdp_sssp_run": (C.c_int, [vp, i32, C.c_int, P(Stats)]),
        "smdp_bfs_run": (C.c_int, [vp, i32, C.c_int, P(Stats)]),
        "smdp_tree_run": (C.c_int, [vp, i32, P(Stats)]),
        "smdp_spmv_run": (C.c_int, [vp, C.c_int, P(Stats)]),
        "smdp_pagerank_run": (C.c_int, [vp, P(PrOpts), C.c_int, P(Stats)]),
        "smdp_color_run": (C.c_int, [vp, C.c_int, C.c_int, P(Stats)]),
        "smdp_sssp_multi_source_run": (C.c_int, [vp, vp, i32, P(Stats)]),
        "smdp_sssp_device_noalloc": (C.c_int, [vp, i32, i32, vp, vp, vp, i32, C.c_int, vp, P(Stats)]),
        "smdp_graph_read": (C.c_int, [vp, C.c_int, vp, C.c_int64]),
        "smdp_batch_create": (C.c_int, [vp, P(Csr), i32, vp, P(vp)]),
        "smdp_batch_run": (C.c_int, [vp, P(Stats)]),
        "smdp_batch_read": (C.c_int, [vp, i32, vp]),
        "smdp_batch_destroy": (None, [vp]),
        "smdp_batch_autotune": (C.c_int, [vp, P(Params)]),
        "smdp_autotune": (C.c_int, [vp, C.c_int, i32, cp, P(Params)]),
        "smdp_read_graph": (C.c_int, [cp, C.c_int, P(P(HGraphStruct))]),
        "smdp_invert": (C.c_int, [P(HGraphStruct), P(P(HGraphStruct))]),
        "smdp_write_dimacs9": (C.c_int, [cp, P(HGraphStruct)]),
        "smdp_hgraph_free": (None, [P(HGraphStruct)]),
        "smdp_gen_uniform": (C.c_int, [i32, i32, i32, i32, C.c_uint64, P(P(HGraphStruct))]),
        "smdp_gen_rmat": (C.c_int, [i32, i32, C.c_double, C.c_double, C.c_double, i32, i32, C.c_int,
                                    C.c_int, C.c_uint64, P(P(HGraphStruct))]),
        "smdp_gen_zipf_tree": (C.c_int, [i32, C.c_double, i32, C.c_uint64, P(P(HGraphStruct))]),
        "smdp_gen_small_graph": (C.c_int, [i32, i32, i32, C.c_double, C.c_double, i32, i32, C.c_uint64,
                                           P(P(HGraphStruct))]),
        "smdp_gen_sources": (C.c_int, [P(HGraphStruct), i32, C.c_uint64, vp]),
        "smdp_gen_uniform_f32": (C.c_int, [C.c_int64, C.c_float, C.c_float, C.c_uint64, C.c_uint32, vp]),
        "smdp_gen_prio": (C.c_int, [i32, C.c_uint64, vp]),
        "smdp_gen_spmv_reference_inputs": (C.c_int, [i32, i32, vp, vp]),
        "smdp_outdegree": (C.c_int, [P(HGraphStruct), vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def strerror(code):
    return lib().smdp_strerror(code).decode()


def last_hip_error():
    return lib().smdp_last_hip_error().decode()


def _check(rc, what):
    if rc != OK:
        raise SmdpError(rc, what)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class HostGraph:
    """A CSR graph in host memory: numpy arrays row, col, w (or None), source."""

    def __init__(self, row, col, w=None, source=0):
        self.row = _i32(row)
        self.col = _i32(col)
        self.w = None if w is None else _i32(w)
        self.source = int(source)
        self.n = len(self.row) - 1
        self.m = len(self.col)

    def csr(self):
        c = Csr(self.n, self.m, _ptr(self.row), _ptr(self.col), _ptr(self.w))
        c._keep = self  # keep arrays alive
        return c

    def hstruct(self):
        h = HGraphStruct(self.n, self.m, self.row.ctypes.data_as(C.POINTER(C.c_int32)),
                         self.col.ctypes.data_as(C.POINTER(C.c_int32)),
                         None if self.w is None else self.w.ctypes.data_as(C.POINTER(C.c_int32)),
                         self.source)
        h._keep = self
        return h

    def outdegree(self):
        return np.diff(self.row).astype(np.int32)

    @staticmethod
    def _from_owned(p):
        h = p.contents
        n, m = h.n, h.m
        row = np.ctypeslib.as_array(h.row, shape=(n + 1,)).copy()
        col = np.ctypeslib.as_array(h.col, shape=(max(m, 1),))[:m].copy()
        w = None if not h.w else np.ctypeslib.as_array(h.w, shape=(max(m, 1),))[:m].copy()
        g = HostGraph(row, col, w, h.source)
        lib().smdp_hgraph_free(p)
        return g


def _owned_call(fn, *args, what):
    out = C.POINTER(HGraphStruct)()
    _check(fn(*args, C.byref(out)), what)
    return HostGraph._from_owned(out)


def read_graph(path, fmt):
    return _owned_call(lib().smdp_read_graph, str(path).encode(), fmt, what="smdp_read_graph")


def invert(g):
    return _owned_call(lib().smdp_invert, C.byref(g.hstruct()), what="smdp_invert")


def write_dimacs9(path, g):
    _check(lib().smdp_write_dimacs9(str(path).encode(), C.byref(g.hstruct())), "smdp_write_dimacs9")


def gen_uniform(n, deg, wlo, whi, seed):
    return _owned_call(lib().smdp_gen_uniform, n, deg, wlo, whi, seed, what="smdp_gen_uniform")


def gen_rmat(scale, ef, a=0.57, b=0.19, c=0.19, wlo=0, whi=0, symmetrize=False, permute=True, seed=1):
    return _owned_call(lib().smdp_gen_rmat, scale, ef, a, b, c, wlo, whi, int(symmetrize), int(permute),
                       seed, what="smdp_gen_rmat")


def gen_zipf_tree(n, s=1.5, cap=4096, seed=5):
    return _owned_call(lib().smdp_gen_zipf_tree, n, s, cap, seed, what="smdp_gen_zipf_tree")


def gen_small_graph(i, vlo=1024, vhi=65536, dlo=4.0, dhi=32.0, wlo=1, whi=100, seed=4):
    return _owned_call(lib().smdp_gen_small_graph, i, vlo, vhi, dlo, dhi, wlo, whi, seed,
                       what="smdp_gen_small_graph")


def gen_sources(g, k, seed):
    out = np.zeros(k, dtype=np.int32)
    _check(lib().smdp_gen_sources(C.byref(g.hstruct()), k, seed, _ptr(out)), "smdp_gen_sources")
    return out


def gen_uniform_f32(count, lo, hi, seed, stream):
    out = np.zeros(count, dtype=np.float32)
    _check(lib().smdp_gen_uniform_f32(count, lo, hi, seed, stream, _ptr(out)), "smdp_gen_uniform_f32")
    return out


def gen_prio(n, seed):
    out = np.zeros(n, dtype=np.uint32)
    _check(lib().smdp_gen_prio(n, seed, _ptr(out)), "smdp_gen_prio")
    return out


def gen_spmv_reference_inputs(n, m):
    data = np.zeros(m, dtype=np.float32)
    x = np.zeros(n, dtype=np.float32)
    _check(lib().smdp_gen_spmv_reference_inputs(n, m, _ptr(data), _ptr(x)), "smdp_gen_spmv_reference_inputs")
    return data, x


def pr_opts(damping=0.85, total_rank=1.0, tol_l1=1e-6, max_iter=100, init_zero=False):
    return PrOpts(damping, total_rank, tol_l1, max_iter, int(init_zero))


class Context:
    """A device context (one HIP stream, tuned parameters)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().smdp_create(device, C.byref(self.h)), "smdp_create")

    def close(self):
        if self.h:
            lib().smdp_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        name = C.create_string_buffer(256)
        arch = C.create_string_buffer(64)
        cus = C.c_int32()
        _check(lib().smdp_device_info(self.h, name, 256, arch, 64, C.byref(cus)), "smdp_device_info")
        return {"name": name.value.decode(), "arch": arch.value.decode(), "cus": cus.value}

    def get_params(self, app):
        p = Params()
        _check(lib().smdp_get_params(self.h, app, C.byref(p)), "smdp_get_params")
        return p

    def set_params(self, app, **kw):
        p = self.get_params(app)
        for k, v in kw.items():
            setattr(p, k, v)
        _check(lib().smdp_set_params(self.h, app, C.byref(p)), "smdp_set_params")

    def set_profiling(self, on):
        _check(lib().smdp_set_profiling(self.h, int(on)), "smdp_set_profiling")

    def round_log(self, cap=4096):
        q = np.zeros(cap, np.int64)
        a = np.zeros(cap, np.int64)
        t = np.zeros(cap, np.float64)
        k = lib().smdp_round_log(self.h, _ptr(q), _ptr(a), _ptr(t), cap)
        if k < 0:
            raise SmdpError(k, "smdp_round_log")
        return q[:k], a[:k], t[:k]

    # drop-in calls ---------------------------------------------------------
    def sssp(self, g, src, sol):
        out = np.zeros(g.n, dtype=np.int32)
        st = Stats()
        _check(lib().smdp_sssp(self.h, C.byref(g.csr()), src, sol, _ptr(out), C.byref(st)), "smdp_sssp")
        return out, st

    def bfs(self, g, src, sol):
        out = np.zeros(g.n, dtype=np.int32)
        st = Stats()
        _check(lib().smdp_bfs(self.h, C.byref(g.csr()), src, sol, _ptr(out), C.byref(st)), "smdp_bfs")
        return out, st

    def tree(self, g, root):
        """Tree Height / Tree Descendant: (height, desc, tree_height, stats)."""
        h = np.zeros(g.n, dtype=np.int32)
        d = np.zeros(g.n, dtype=np.int32)
        th = C.c_int32()
        st = Stats()
        _check(lib().smdp_tree(self.h, C.byref(g.csr()), root, _ptr(h), _ptr(d), C.byref(th), C.byref(st)),
               "smdp_tree")
        return h, d, int(th.value), st

    def spmv(self, a, val, x, sol):
        val = np.ascontiguousarray(val, dtype=np.float32)
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(a.n, dtype=np.float32)
        st = Stats()
        _check(lib().smdp_spmv_f32(self.h, C.byref(a.csr()), _ptr(val), _ptr(x), sol, _ptr(y), C.byref(st)),
               "smdp_spmv_f32")
        return y, st

    def pagerank(self, in_g, outdeg, opts, sol):
        outdeg = _i32(outdeg)
        r = np.zeros(in_g.n, dtype=np.float32)
        st = Stats()
        _check(lib().smdp_pagerank_f32(self.h, C.byref(in_g.csr()), _ptr(outdeg), C.byref(opts), sol, _ptr(r),
                                       C.byref(st)), "smdp_pagerank_f32")
        return r, st

    def color(self, g, prio, sol):
        prio = None if prio is None else np.ascontiguousarray(prio, dtype=np.uint32)
        out = np.zeros(g.n, dtype=np.int32)
        st = Stats()
        _check(lib().smdp_color(self.h, C.byref(g.csr()), _ptr(prio), sol, _ptr(out), C.byref(st)), "smdp_color")
        return out, st

    def sssp_multi_source(self, g, srcs):
        srcs = _i32(srcs)
        out = np.zeros((len(srcs), g.n), dtype=np.int32)
        st = Stats()
        _check(lib().smdp_sssp_multi_source(self.h, C.byref(g.csr()), _ptr(srcs), len(srcs), _ptr(out),
                                            C.byref(st)), "smdp_sssp_multi_source")
        return out, st

    def sssp_multi_graph(self, graphs, srcs):
        srcs = _i32(srcs)
        csrs = (Csr * len(graphs))(*[g.csr() for g in graphs])
        outs = [np.zeros(g.n, dtype=np.int32) for g in graphs]
        ptrs = (C.c_void_p * len(graphs))(*[o.ctypes.data for o in outs])
        st = Stats()
        _check(lib().smdp_sssp_multi_graph(self.h, csrs, len(graphs), _ptr(srcs), ptrs, C.byref(st)),
               "smdp_sssp_multi_graph")
        return outs, st


class Graph:
    """A graph resident in device memory (upload once, run many times)."""

    def __init__(self, ctx, g):
        self.ctx = ctx
        self.host = g
        self.n, self.m = g.n, g.m
        self.h = C.c_void_p()
        _check(lib().smdp_graph_create(ctx.h, C.byref(g.csr()), C.byref(self.h)), "smdp_graph_create")

    def close(self):
        if self.h:
            lib().smdp_graph_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_values(self, val):
        self._val = np.ascontiguousarray(val, dtype=np.float32)
        _check(lib().smdp_graph_set_values(self.h, _ptr(self._val)), "smdp_graph_set_values")

    def set_x(self, x):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        _check(lib().smdp_graph_set_x(self.h, _ptr(self._x)), "smdp_graph_set_x")

    def set_outdeg(self, d):
        self._od = _i32(d)
        _check(lib().smdp_graph_set_outdeg(self.h, _ptr(self._od)), "smdp_graph_set_outdeg")

    def set_prio(self, p):
        self._prio = np.ascontiguousarray(p, dtype=np.uint32)
        _check(lib().smdp_graph_set_prio(self.h, _ptr(self._prio)), "smdp_graph_set_prio")

    def sssp(self, src, sol):
        st = Stats()
        _check(lib().smdp_sssp_run(self.h, src, sol, C.byref(st)), "smdp_sssp_run")
        return st

    def bfs(self, src, sol):
        st = Stats()
        _check(lib().smdp_bfs_run(self.h, src, sol, C.byref(st)), "smdp_bfs_run")
        return st

    def tree(self, root):
        st = Stats()
        _check(lib().smdp_tree_run(self.h, root, C.byref(st)), "smdp_tree_run")
        return st

    def spmv(self, sol):
        st = Stats()
        _check(lib().smdp_spmv_run(self.h, sol, C.byref(st)), "smdp_spmv_run")
        return st

    def pagerank(self, opts, sol):
        st = Stats()
        _check(lib().smdp_pagerank_run(self.h, C.byref(opts), sol, C.byref(st)), "smdp_pagerank_run")
        return st

    def color(self, use_prio, sol):
        st = Stats()
        _check(lib().smdp_color_run(self.h, int(use_prio), sol, C.byref(st)), "smdp_color_run")
        return st

    def sssp_multi_source(self, srcs):
        self._srcs = _i32(srcs)
        st = Stats()
        _check(lib().smdp_sssp_multi_source_run(self.h, _ptr(self._srcs), len(self._srcs), C.byref(st)),
               "smdp_sssp_multi_source_run")
        return st

    def read(self, what, dtype=np.int32, count=None):
        k = self.n if count is None else int(count)
        out = np.zeros(k, dtype=dtype)
        _check(lib().smdp_graph_read(self.h, what, _ptr(out), k), "smdp_graph_read")
        return out

    def autotune(self, app, src=0, cache_path=None):
        p = Params()
        _check(lib().smdp_autotune(self.h, app, src, None if cache_path is None else str(cache_path).encode(),
                                   C.byref(p)), "smdp_autotune")
        return p


class Batch:
    """Small-graph batch resident on the device, run as one captured hipGraph."""

    def __init__(self, ctx, graphs, srcs):
        self.ctx = ctx
        self.graphs = graphs
        srcs = _i32(srcs)
        csrs = (Csr * len(graphs))(*[g.csr() for g in graphs])
        self.h = C.c_void_p()
        _check(lib().smdp_batch_create(ctx.h, csrs, len(graphs), _ptr(srcs), C.byref(self.h)), "smdp_batch_create")

    def run(self):
        st = Stats()
        _check(lib().smdp_batch_run(self.h, C.byref(st)), "smdp_batch_run")
        return st

    def read(self, i):
        out = np.zeros(self.graphs[i].n, dtype=np.int32)
        _check(lib().smdp_batch_read(self.h, i, _ptr(out)), "smdp_batch_read")
        return out

    def autotune(self):
        p = Params()
        _check(lib().smdp_batch_autotune(self.h, C.byref(p)), "smdp_batch_autotune")
        return p

    def close(self):
        if self.h:
            lib().smdp_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
