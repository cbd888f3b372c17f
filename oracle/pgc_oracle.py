"""ctypes wrapper of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY.

Imported by tests/, the smoke check and the benchmark's CPU-baseline leg; never
by the product package. Mirrors the reference's CPU path:
product_graph_generators/product_spaces/product_space.cpp:19-132,
product_graph_generators/energy/deformationEnergy.cpp:11-103,
product_graph_generators/cost_time_ratio_solver/cost_time_ratio_solver.cpp:42-905,
product_graph_generators/dijkstra_solver/dijkstra_solver.cpp:41-579,
helper/utils.cpp:86-157 (normals).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.c_void_p
        L.orc_product_space.restype = C.c_long
        L.orc_product_space.argtypes = [P, C.c_long, P, C.c_long, C.c_int, P, P, P]
        L.orc_energy.restype = C.c_int
        L.orc_energy.argtypes = [P, C.c_long, P, C.c_long, P, C.c_long, P, C.c_int, C.c_char_p, C.c_char_p, P]
        L.orc_median.restype = C.c_double
        L.orc_median.argtypes = [P, C.c_long]
        L.orc_quantize.restype = None
        L.orc_quantize.argtypes = [P, C.c_long, C.c_int]
        L.orc_ctr_solve.restype = C.c_int
        L.orc_ctr_solve.argtypes = [P, P, C.c_long, P, C.c_int, C.c_long, C.c_int, C.c_int, P, C.c_long,
                                    P, P, P, P]
        L.orc_dijkstra_solve.restype = C.c_int
        L.orc_dijkstra_solve.argtypes = [P, P, C.c_long, P, C.c_long, C.c_int, P, P, C.c_int, P, C.c_long,
                                         P, P, P, P]
        L.orc_last_error.restype = C.c_char_p
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_conj_product_space.restype = C.c_long
        L.orc_conj_product_space.argtypes = [P, C.c_long, P, C.c_long, C.c_int, P, P, P]
        L.orc_conj_branch_graph.restype = C.c_long
        L.orc_conj_branch_graph.argtypes = [P, C.c_long, C.c_int, P, P]
        L.orc_conj_quaternions.argtypes = [P, P, P, P, P, C.c_long, P, C.c_long, C.c_int, P]
        L.orc_conj_energy.argtypes = [P, P, P, C.c_long, P, P, C.c_long, P, C.c_long, P, P, P, C.c_long, P,
                                      C.c_long, C.c_int, C.c_char_p, C.c_char_p, C.c_double, C.c_double,
                                      C.c_double, C.c_int, P]
        L.orc_normals2d.restype = C.c_int
        L.orc_normals2d.argtypes = [P, C.c_long, P, C.c_long, P, P]
        L.orc_normals3d.restype = None
        L.orc_normals3d.argtypes = [P, C.c_long, P, C.c_long, P]
        L.orc_exact_ratio.restype = C.c_int
        L.orc_exact_ratio.argtypes = [P, P, C.c_int64, P, C.c_int64, C.c_int, C.c_int64, C.c_int64, P, P, P, P,
                                      C.c_int64, P, P]
        L.orc_exact_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def product_space(EX, EY, prune=False):
    EX = np.ascontiguousarray(EX, dtype=np.int32)
    EY = np.ascontiguousarray(EY, dtype=np.int32)
    nVX = int(EX.max()) + 1
    n = len(EY) * ((len(EX) if not prune else 0) + len(EX) + nVX)
    rows = np.zeros((n, 4), np.int32)
    src = np.zeros(n, np.int32)
    trgt = np.zeros(n, np.int32)
    m = lib().orc_product_space(_p(EX), len(EX), _p(EY), len(EY), int(prune), _p(rows), _p(src), _p(trgt))
    assert m == n
    return rows, src, trgt


def energy(VX, VY, rows, FD, reg=False, cost_mode="vanilla", time_mode=""):
    VX = np.ascontiguousarray(VX, dtype=np.float64)
    VY = np.ascontiguousarray(VY, dtype=np.float64)
    FD = np.ascontiguousarray(FD, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    E = len(rows)
    out = np.zeros(2 * E, np.float64)
    nc = lib().orc_energy(_p(VX), len(VX), _p(VY), len(VY), _p(rows), E, _p(FD), int(reg),
                          cost_mode.encode(), time_mode.encode(), _p(out))
    return out[: nc * E].reshape(nc, E).T.copy()  # [E, ncols]


def median(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_median(_p(x), len(x))


def quantize(cost):
    """cost [E, ncols] float64 -> quantized copy."""
    cm = np.ascontiguousarray(np.asarray(cost, np.float64).T).copy()
    lib().orc_quantize(_p(cm), cm.shape[1], cm.shape[0])
    return cm.T.copy()


# Branch-and-bound events counted by orc_ctr_solve (stats[5:16]); see Stats in
# pgc_oracle.cpp. The first ten are the ways the reference's algorithm can end
# above the once-wrapping optimum.
CTR_EVENTS = ("push_id", "push_present", "single", "rise", "stale", "cap", "walk", "empty", "maxiter",
              "ub_break", "valid_pop")


def ctr_events(stats):
    return {k: int(stats[5 + i]) for i, k in enumerate(CTR_EVENTS)}


def ctr_solve(src, trgt, cost, npl, K, gpu_pick=False, corrected=False):
    """Lawler ratio solve. Returns (edge_ids, objective float32, passes, stats, quantized cost).
    corrected: fold the closing layer's depth-0 row for odd L (Q1 fixed).
    stats: [bf_calls, sweeps, lawler_iters, branches, relaxations] and the
    branch-and-bound events of CTR_EVENTS at indices 5..15."""
    src = np.ascontiguousarray(src, np.int32)
    trgt = np.ascontiguousarray(trgt, np.int32)
    cost = np.asarray(cost, np.float64)
    if cost.ndim == 1:
        cost = cost[:, None]
    cm = np.ascontiguousarray(cost.T).copy()
    E = len(src)
    L = (int(src.max()) + 1) // npl
    cap = max(64, 4 * (L + 1) * K * npl)
    out = np.zeros(cap, np.int32)
    n = C.c_long(0)
    obj = C.c_float(0)
    passes = C.c_int(0)
    stats = np.zeros(16, np.int64)
    flags = int(gpu_pick) | (2 if corrected else 0)
    rc = lib().orc_ctr_solve(_p(src), _p(trgt), E, _p(cm), cm.shape[0], npl, K, flags, _p(out), cap,
                             C.byref(n), C.byref(obj), C.byref(passes), _p(stats))
    if rc != 0:
        raise RuntimeError("oracle ctr solve failed: %s" % lib().orc_last_error().decode())
    return out[: n.value].copy(), np.float32(obj.value), bool(passes.value), stats, cm.T.copy()


def dijkstra_solve(src, trgt, cost0, npl, K, branch_graph=None, cpu_rules=False, corrected=False):
    """Minimum-cost solve on raw column-0 costs. branch_graph: list of value lists or None.
    corrected: lower bounds from the closing layer's depth-0 row for odd L (Q1 fixed)."""
    src = np.ascontiguousarray(src, np.int32)
    trgt = np.ascontiguousarray(trgt, np.int32)
    cost0 = np.ascontiguousarray(np.asarray(cost0, np.float64).reshape(-1))
    E = len(src)
    L = (int(src.max()) + 1) // npl
    cap = max(64, 4 * (L + 1) * K)
    out = np.zeros(cap, np.int32)
    n = C.c_long(0)
    obj = C.c_float(0)
    passes = C.c_int(0)
    stats = np.zeros(5, np.int64)
    if branch_graph is not None:
        off = np.zeros(npl + 1, np.int64)
        vals = []
        for v in range(npl):
            s = sorted(branch_graph[v]) if v < len(branch_graph) else []
            vals += s
            off[v + 1] = len(vals)
        val = np.array(vals if vals else [0], np.int64)
        bo, bv = _p(off), _p(val)
    else:
        bo = bv = None
    flags = int(cpu_rules) | (2 if corrected else 0)
    rc = lib().orc_dijkstra_solve(_p(src), _p(trgt), E, _p(cost0), npl, K, bo, bv, flags, _p(out), cap,
                                  C.byref(n), C.byref(obj), C.byref(passes), _p(stats))
    if rc != 0:
        raise RuntimeError("oracle dijkstra solve failed: %s" % lib().orc_last_error().decode())
    return out[: n.value].copy(), np.float32(obj.value), bool(passes.value), stats


# ------------------------------------------------------------ conjugate graph
def conj_product_space(EX, EY, prune=False):
    """Rows [E, 6] (src2d, mid2d, trgt2d, src3d, mid3d, trgt3d), SRC, TRGT."""
    EX = np.ascontiguousarray(EX, dtype=np.int32)
    EY = np.ascontiguousarray(EY, dtype=np.int32)
    n = lib().orc_conj_product_space(_p(EX), len(EX), _p(EY), len(EY), int(prune), None, None, None)
    rows = np.zeros((n, 6), np.int32)
    src = np.zeros(n, np.int32)
    trgt = np.zeros(n, np.int32)
    m = lib().orc_conj_product_space(_p(EX), len(EX), _p(EY), len(EY), int(prune), _p(rows), _p(src), _p(trgt))
    assert m == n
    return rows, src, trgt


def conj_branch_graph(EX, prune=False):
    """List (per layer-0 node) of sorted branch-graph values."""
    EX = np.ascontiguousarray(EX, dtype=np.int32)
    npl = 3 * len(EX)
    off = np.zeros(npl + 1, np.int64)
    n = lib().orc_conj_branch_graph(_p(EX), len(EX), int(prune), _p(off), None)
    vals = np.zeros(max(n, 1), np.int64)
    lib().orc_conj_branch_graph(_p(EX), len(EX), int(prune), _p(off), _p(vals))
    return [list(vals[off[v]:off[v + 1]]) for v in range(npl)]


def conj_quaternions(VX, NX, VY, NY, EX, EY, prune=False):
    """(w, x, y, z) per product node, [3*|EX|*|EY|, 4]."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (VX, NX, VY, NY)]
    EX = np.ascontiguousarray(EX, dtype=np.int32)
    EY = np.ascontiguousarray(EY, dtype=np.int32)
    q = np.zeros((3 * len(EX) * len(EY), 4), np.float64)
    lib().orc_conj_quaternions(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), _p(EX), len(EX), _p(EY), len(EY),
                               int(prune), _p(q))
    return q


def conj_energy(VX, NX, VY, NY, EX, EY, rows, src, trgt, FD, reg=False, cost_mode="vanilla", time_mode="",
                rl=(0.7, 0.6, 4.0), prune=False):
    """Energy [E, ncols]; raises ValueError for an unsupported (cost, time) pair."""
    VX, NX, VY, NY, FD = [np.ascontiguousarray(x, dtype=np.float64) for x in (VX, NX, VY, NY, FD)]
    EX = np.ascontiguousarray(EX, dtype=np.int32)
    EY = np.ascontiguousarray(EY, dtype=np.int32)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    src = np.ascontiguousarray(src, dtype=np.int32)
    trgt = np.ascontiguousarray(trgt, dtype=np.int32)
    E = len(rows)
    out = np.zeros(2 * E, np.float64)
    nc = lib().orc_conj_energy(_p(VX), _p(NX), _p(VY), len(VY), _p(NY), _p(EX), len(EX), _p(EY), len(EY),
                               _p(rows), _p(src), _p(trgt), E, _p(FD), len(VX), int(reg), cost_mode.encode(),
                               time_mode.encode(), float(rl[0]), float(rl[1]), float(rl[2]), int(prune), _p(out))
    if nc < 0:
        raise ValueError("unsupported (cost, time) mode pair for the conjugate energy: (%s, %s)"
                         % (cost_mode, time_mode))
    return out[: nc * E].reshape(nc, E).T.copy()



def normals2d(V, E):
    """utils::normals2d (helper/utils.cpp:93-157). Returns (status, zero_dim,
    normals): status 0, or 1 / 2 for the reference's two error returns."""
    V = np.ascontiguousarray(V, np.float64)
    E = np.ascontiguousarray(E, np.int32)
    out = np.zeros((len(V), 3))
    zd = np.zeros(1, np.int32)
    st = lib().orc_normals2d(_p(V), len(V), _p(E), len(E), _p(out), _p(zd))
    return int(st), int(zd[0]), out


def normals3d(V, F):
    """utils::normals3d = igl::per_vertex_normals, default weighting (helper/utils.cpp:86-90)."""
    V = np.ascontiguousarray(V, np.float64)
    F = np.ascontiguousarray(F, np.int32)
    out = np.zeros((len(V), 3))
    lib().orc_normals3d(_p(V), len(V), _p(F), len(F), _p(out))
    return out


def exact_ratio(src, trgt, q, npl, K, a0=None, b0=None):
    """Exact minimum ratio of the once-wrapping cycles with integer costs q and
    unit times (oracle/exact_ratio.cpp, Dinkelbach in int64). (a0, b0): a ratio
    not below the minimum (default: max(q) + 1 over 1). Returns (a, b, start,
    edges, iterations); edges is empty when no cycle is below a0/b0, which then
    is the minimum (certified)."""
    src = np.ascontiguousarray(src, np.int32)
    trgt = np.ascontiguousarray(trgt, np.int32)
    qi = np.ascontiguousarray(np.rint(np.asarray(q, np.float64)).astype(np.int64))
    assert np.array_equal(qi.astype(np.float64), np.asarray(q, np.float64)), "costs must be integers"
    if a0 is None:
        a0, b0 = int(qi.max()) + 1, 1
    n_nodes = int(src.max()) + 1
    cap = (n_nodes // npl) * K + 8
    path = np.zeros(cap, np.int32)
    a, b, st, pl, it = (C.c_int64(), C.c_int64(), C.c_int32(), C.c_int64(), C.c_int64())
    rc = lib().orc_exact_ratio(_p(src), _p(trgt), len(src), _p(qi), int(npl), int(K), int(a0), int(b0), C.byref(a),
                               C.byref(b), C.byref(st), _p(path), cap, C.byref(pl), C.byref(it))
    if rc != 0:
        raise RuntimeError(lib().orc_exact_error().decode())
    return a.value, b.value, st.value, path[:pl.value].copy(), it.value
