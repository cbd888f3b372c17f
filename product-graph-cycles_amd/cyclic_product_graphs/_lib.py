"""ctypes binding of libpgc.so (include/pgc.h).

The library is built in-tree (product-graph-cycles_amd/Makefile). Loading
fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgc.so")

PGC_LAWLER_PICK_LAST = 1
PGC_MINCOST_CPU_RULES = 2
PGC_CORRECTED_ODD_L = 4
# status codes (include/pgc.h)
PGC_OK, PGC_ERR_ARG, PGC_ERR_HIP, PGC_ERR_SOLVER, PGC_ERR_CAPACITY, PGC_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5


class PgcError(RuntimeError):
    pass


class pgc_result(C.Structure):
    _fields_ = [
        ("passes_layers_once", C.c_int32),
        ("objective", C.c_float),
        ("n_edges", C.c_int64),
        ("sweeps", C.c_int64),
        ("bf_calls", C.c_int64),
        ("lawler_iters", C.c_int64),
        ("branches", C.c_int64),
        ("relaxations", C.c_double),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("phase_ms", C.c_double * 5),
        ("status", C.c_int32),
        ("launches", C.c_int32),
        ("implicit_costs", C.c_int32),
        ("reserved", C.c_int32),
    ]

    PHASES = ("sweep", "fold", "chase", "walk", "extract")

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("phase_ms", "reserved")}
        d["phase_ms"] = dict(zip(self.PHASES, list(self.phase_ms)))
        return d


class pgc_tuning(C.Structure):
    _fields_ = [
        ("grid_chunks_per_wave", C.c_int32),
        ("pred_in_lds", C.c_int32),
        ("wrap_in_lds", C.c_int32),
        ("origins", C.c_int32),
        ("cached_sources", C.c_int32),
        ("short_slots", C.c_int32),
        ("global_table", C.c_int32),
        ("grid_workgroups", C.c_int32),
        ("implicit_costs", C.c_int32),
    ]
    FIELDS = ("grid_chunks_per_wave", "pred_in_lds", "wrap_in_lds", "origins", "cached_sources", "short_slots",
              "global_table", "grid_workgroups", "implicit_costs")

    def as_dict(self):
        return {k: int(getattr(self, k)) for k in self.FIELDS}


EXPORTED = [
    "pgc_last_error", "pgc_abi_version", "pgc_context_create", "pgc_context_destroy", "pgc_graph_generate",
    "pgc_graph_generate_device",
    "pgc_graph_from_arrays", "pgc_graph_set_branch_graph", "pgc_graph_info", "pgc_graph_get_product_space",
    "pgc_graph_get_cost", "pgc_graph_set_cost", "pgc_graph_destroy", "pgc_solve_ratio", "pgc_solve_min_cost",
    "pgc_solve_ratio_batch", "pgc_graph_set_feat_diff", "pgc_graphs_set_feat_diff", "pgc_graph_derive",
    "pgc_cycle_ratio_values", "pgc_cycle_ratio_grad", "pgc_solve_ratio_noalloc", "pgc_context_set_tuning",
    "pgc_context_get_tuning", "pgc_graph_get_layout",
]

_lib = None


def load():
    """Load libpgc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PgcError("libpgc.so is not built (%s); run `make -C product-graph-cycles_amd`" % LIB_PATH)
    # PyTorch ships its own HIP runtime. In a process that uses both, torch's
    # must be loaded first: loaded after libpgc's (ROCm's own), it finds no
    # device. Importing torch here fixes the order.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P, I64, I32, D = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    L.pgc_last_error.restype = C.c_char_p
    L.pgc_abi_version.restype = C.c_int
    L.pgc_context_create.argtypes = [C.c_int, C.POINTER(P)]
    L.pgc_context_destroy.argtypes = [P]
    L.pgc_context_destroy.restype = None
    L.pgc_context_set_tuning.argtypes = [P, C.POINTER(pgc_tuning)]
    L.pgc_context_get_tuning.argtypes = [P, C.POINTER(pgc_tuning)]
    L.pgc_graph_generate.argtypes = [P, P, I64, P, I64, P, I64, P, I64, P, C.c_int, C.c_int, C.c_int,
                                     C.c_char_p, C.c_char_p, P, P, D, D, D, C.POINTER(P)]
    L.pgc_graph_generate_device.argtypes = L.pgc_graph_generate.argtypes
    L.pgc_graph_from_arrays.argtypes = [P, P, C.c_int, P, P, I64, P, C.c_int, I64, C.POINTER(P)]
    L.pgc_graph_set_branch_graph.argtypes = [P, P, P]
    L.pgc_graph_info.argtypes = [P, C.POINTER(I64), C.POINTER(I32), C.POINTER(I32), C.POINTER(I64),
                                 C.POINTER(I64)]
    L.pgc_graph_get_product_space.argtypes = [P, P]
    L.pgc_graph_get_cost.argtypes = [P, P]
    L.pgc_graph_set_cost.argtypes = [P, P]
    L.pgc_graph_destroy.argtypes = [P]
    L.pgc_graph_destroy.restype = None
    L.pgc_solve_ratio.argtypes = [P, P, C.c_int, C.c_uint32, P, P, I64, C.POINTER(pgc_result)]
    L.pgc_solve_min_cost.argtypes = [P, P, C.c_int, C.c_uint32, P, P, I64, C.POINTER(pgc_result)]
    L.pgc_solve_ratio_noalloc.argtypes = [P, P, C.c_int, C.c_uint32, P, P, I64, C.POINTER(pgc_result)]
    L.pgc_solve_ratio_batch.argtypes = [P, P, P, I64, C.c_uint32, P, P, P, P]
    L.pgc_graph_set_feat_diff.argtypes = [P, P, C.c_int]
    L.pgc_graphs_set_feat_diff.argtypes = [P, P, I64, C.c_int]
    L.pgc_graph_derive.argtypes = [P, P, C.c_int, C.POINTER(P)]
    L.pgc_cycle_ratio_values.argtypes = [P, P, I64, P, P, P, P, P]
    L.pgc_cycle_ratio_grad.argtypes = [P, P, I64, P, P, P, P, P]
    L.pgc_graph_get_layout.argtypes = [P, C.c_int, P, P, I64, C.POINTER(I64), C.POINTER(I32)]
    for name in EXPORTED:
        getattr(L, name)
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise PgcError("libpgc error %d: %s" % (rc, load().pgc_last_error().decode()))


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Context:
    """A device + HIP stream owned by libpgc."""

    def __init__(self, device: int = 0):
        L = load()
        self._h = C.c_void_p()
        check(L.pgc_context_create(int(device), C.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        return self._h

    def set_tuning(self, choices=None):
        """Launch choices (a dict with pgc_tuning's fields; missing fields and
        None keep the built-in rule, -1)."""
        t = pgc_tuning()
        for k in pgc_tuning.FIELDS:
            setattr(t, k, -1)
        for k, v in (choices or {}).items():
            if k not in pgc_tuning.FIELDS:
                raise ValueError("unknown tuning field %r" % k)
            setattr(t, k, -1 if v is None else int(v))
        check(load().pgc_context_set_tuning(self._h, C.byref(t)))

    def tuning(self):
        t = pgc_tuning()
        check(load().pgc_context_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def close(self):
        if self._h:
            load().pgc_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    if device not in _default_ctx:
        _default_ctx[device] = Context(device)
    return _default_ctx[device]


class Graph:
    """A device-resident product graph (rows, SRC/TRGT, per-edge cost/time)."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self._h = handle

    @classmethod
    def generate(cls, ctx, VX, EX, VY, EY, FD, conjugate, regularising, prune, cost_mode, time_mode,
                 NX=None, NY=None, rl=(0.7, 0.6, 4.0)):
        L = load()
        VX = np.ascontiguousarray(VX, np.float64)
        VY = np.ascontiguousarray(VY, np.float64)
        EX = np.ascontiguousarray(EX, np.int32)
        EY = np.ascontiguousarray(EY, np.int32)
        FD = np.ascontiguousarray(FD, np.float64)
        NX = None if NX is None else np.ascontiguousarray(NX, np.float64)
        NY = None if NY is None else np.ascontiguousarray(NY, np.float64)
        h = C.c_void_p()
        check(L.pgc_graph_generate(ctx.handle, ptr(VX), len(VX), ptr(EX), len(EX), ptr(VY), len(VY), ptr(EY),
                                   len(EY), ptr(FD), int(conjugate), int(regularising), int(prune),
                                   cost_mode.encode(), time_mode.encode(), ptr(NX), ptr(NY), float(rl[0]),
                                   float(rl[1]), float(rl[2]), C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def generate_device(cls, ctx, VX, EX, VY, EY, FD, regularising=False, prune=False, cost_mode="vanilla",
                        time_mode=""):
        """Inputs are device tensors (torch, on this context's device): VX/VY/FD float64, EX/EY int32."""
        L = load()
        for t, dt in ((VX, "float64"), (VY, "float64"), (FD, "float64"), (EX, "int32"), (EY, "int32")):
            if not t.is_cuda or str(t.dtype) != "torch." + dt or not t.is_contiguous():
                raise ValueError("generate_device expects contiguous device tensors (float64 / int32)")
        h = C.c_void_p()
        check(L.pgc_graph_generate_device(ctx.handle, C.c_void_p(VX.data_ptr()), VX.shape[0],
                                          C.c_void_p(EX.data_ptr()), EX.shape[0], C.c_void_p(VY.data_ptr()),
                                          VY.shape[0], C.c_void_p(EY.data_ptr()), EY.shape[0],
                                          C.c_void_p(FD.data_ptr()), 0, int(regularising), int(prune),
                                          cost_mode.encode(), time_mode.encode(), None, None, 0.7, 0.6, 4.0,
                                          C.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_arrays(cls, ctx, productspace, src, trgt, cost, nodes_per_layer):
        L = load()
        src = np.ascontiguousarray(src, np.int32).reshape(-1)
        trgt = np.ascontiguousarray(trgt, np.int32).reshape(-1)
        cost = np.asarray(cost, np.float64)
        if cost.ndim == 1:
            cost = cost[:, None]
        cm = np.ascontiguousarray(cost.T)
        ps = None if productspace is None else np.ascontiguousarray(productspace, np.int32)
        cols = 4 if ps is None else ps.shape[1]
        h = C.c_void_p()
        check(L.pgc_graph_from_arrays(ctx.handle, ptr(ps), cols, ptr(src), ptr(trgt), len(src), ptr(cm),
                                      cm.shape[0], int(nodes_per_layer), C.byref(h)))
        return cls(ctx, h)

    @staticmethod
    def _fd_arg(fd):
        """(pointer, on_device, keep-alive) of a feat_diff given as a device
        tensor (float64, contiguous) or a host array."""
        if hasattr(fd, "is_cuda") and fd.is_cuda:
            if str(fd.dtype) != "torch.float64" or not fd.is_contiguous():
                raise ValueError("device feat_diff must be a contiguous float64 tensor")
            return C.c_void_p(fd.data_ptr()), 1, fd
        a = np.ascontiguousarray(fd, np.float64)
        return ptr(a), 0, a

    def set_feat_diff(self, fd):
        """Recompute the energy from another feat_diff (same graph)."""
        p, dev, keep = self._fd_arg(fd)
        check(load().pgc_graph_set_feat_diff(self._h, p, dev))

    def derive(self, fd):
        """A graph sharing this one's product space and layout with the energy of fd."""
        p, dev, keep = self._fd_arg(fd)
        h = C.c_void_p()
        check(load().pgc_graph_derive(self._h, p, dev, C.byref(h)))
        return Graph(self.ctx, h)

    def info(self):
        # the sizes are fixed when the graph is built: read them once (batched
        # solves ask for every item's sizes on every call)
        cached = getattr(self, "_info", None)
        if cached is None or cached[0] != self._h.value:
            e, npl, nl = C.c_int64(), C.c_int64(), C.c_int64()
            pc, cc = C.c_int32(), C.c_int32()
            check(load().pgc_graph_info(self._h, C.byref(e), C.byref(pc), C.byref(cc), C.byref(npl), C.byref(nl)))
            cached = (self._h.value, dict(n_edges=e.value, ps_cols=pc.value, n_cost_cols=cc.value,
                                          nodes_per_layer=npl.value, n_layers=nl.value))
            self._info = cached
        return dict(cached[1])

    def product_space(self):
        i = self.info()
        out = np.empty((i["n_edges"], i["ps_cols"]), np.int32)
        check(load().pgc_graph_get_product_space(self._h, ptr(out)))
        return out

    def cost(self):
        i = self.info()
        out = np.empty((i["n_cost_cols"], i["n_edges"]), np.float64)
        check(load().pgc_graph_get_cost(self._h, ptr(out)))
        return out.T.copy()

    def layout(self, cls):
        """(src, eid, width) of the solver's in-edge slots of class cls (0
        intralayer, 1 interlayer), built if needed: slot j of lane i of chunk c
        on layer l at l*chunks*width*64 + (c*width + j)*64 + i."""
        L = load()
        n, w = C.c_int64(), C.c_int32()
        check(L.pgc_graph_get_layout(self._h, int(cls), None, None, 0, C.byref(n), C.byref(w)))
        src = np.empty(n.value, np.int32)
        eid = np.empty(n.value, np.int32)
        check(L.pgc_graph_get_layout(self._h, int(cls), ptr(src), ptr(eid), n.value, C.byref(n), C.byref(w)))
        return src, eid, w.value

    def set_cost(self, cost):
        cost = np.asarray(cost, np.float64)
        if cost.ndim == 1:
            cost = cost[:, None]
        cm = np.ascontiguousarray(cost.T)
        check(load().pgc_graph_set_cost(self._h, ptr(cm)))

    def set_branch_graph(self, lists):
        off = np.zeros(len(lists) + 1, np.int64)
        vals = []
        for i, l in enumerate(lists):
            vals += sorted(int(x) for x in l)
            off[i + 1] = len(vals)
        v = np.array(vals if vals else [0], np.int64)
        check(load().pgc_graph_set_branch_graph(self._h, ptr(off), ptr(v)))

    def _solve(self, fn, max_depth, flags):
        i = self.info()
        cap = max(64, 101 * i["n_layers"] * max_depth + 8)
        eids = np.empty(cap, np.int32)
        rows = np.empty((cap, i["ps_cols"]), np.int32)
        res = pgc_result()
        check(fn(self.ctx.handle, self._h, int(max_depth), int(flags), ptr(eids), ptr(rows), cap, C.byref(res)))
        n = res.n_edges
        return eids[:n].copy(), rows[:n].copy(), res

    def solve_ratio(self, max_depth, flags=PGC_LAWLER_PICK_LAST):
        return self._solve(load().pgc_solve_ratio, max_depth, flags)

    def solve_ratio_noalloc(self, max_depth, flags=PGC_LAWLER_PICK_LAST):
        """solve_ratio with no device allocation (the library's static scratch)."""
        return self._solve(load().pgc_solve_ratio_noalloc, max_depth, flags)

    def solve_min_cost(self, max_depth, flags=0):
        return self._solve(load().pgc_solve_min_cost, max_depth, flags)

    def close(self):
        if self._h:
            load().pgc_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_ratio_batch(ctx, graphs, max_depth, flags=PGC_LAWLER_PICK_LAST, with_rows=True, raise_on_error=True,
                      capacity=None):
    """Complete lawlergpu solves of a list of Graphs in persistent launches.

    max_depth: one int for all graphs or one per graph. Returns a list of
    (edge_ids, rows or None, pgc_result) in the order of `graphs`. With
    raise_on_error=False a failing graph does not raise: its result's status
    is its error code (the graphs that solved are returned as usual).
    capacity: output edges per graph (default: the reference's cycle buffer).
    """
    L = load()
    n = len(graphs)
    if n == 0:
        return []
    K = np.broadcast_to(np.asarray(max_depth, np.int32), (n,)).copy()
    infos = [g.info() for g in graphs]
    if capacity is None:
        caps = np.array([max(64, 101 * i["n_layers"] * int(k) + 8) for i, k in zip(infos, K)], np.int64)
    else:
        caps = np.broadcast_to(np.asarray(capacity, np.int64), (n,)).copy()
    off = np.zeros(n + 1, np.int64)
    np.cumsum(caps, out=off[1:])
    eids = np.empty(int(off[-1]), np.int32)
    cols = infos[0]["ps_cols"]
    rows = np.empty((int(off[-1]), cols), np.int32) if with_rows else None
    handles = (C.c_void_p * n)(*[g._h.value for g in graphs])
    res = (pgc_result * n)()
    rc = L.pgc_solve_ratio_batch(ctx.handle, handles, ptr(K), n, int(flags), ptr(eids), ptr(rows), ptr(off), res)
    if rc != 0 and (raise_on_error or all(res[i].status == 0 for i in range(n))):
        check(rc)
    out = []
    for i in range(n):
        m = res[i].n_edges
        a = int(off[i])
        out.append((eids[a:a + m].copy(), None if rows is None else rows[a:a + m].copy(), res[i]))
    return out


def _cycles(cycles):
    n = len(cycles)
    lengths = np.array([len(c) for c in cycles], np.int64)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(lengths, out=off[1:])
    cat = np.concatenate([np.asarray(c, np.int32) for c in cycles]) if int(off[-1]) else np.zeros(1, np.int32)
    return np.ascontiguousarray(cat, np.int32), off, lengths


def _dptr(x):
    return x if isinstance(x, int) else x.data_ptr()


def set_feat_diff_batch(graphs, fds, dtype=0):
    """pgc_graphs_set_feat_diff: new energies of graphs from feat_diffs fds
    (contiguous device tensors or device addresses; float64 when dtype is 0,
    float32 when 1), in one launch."""
    n = len(graphs)
    if n == 0:
        return
    handles = (C.c_void_p * n)(*[g._h.value for g in graphs])
    fdp = (C.c_void_p * n)(*[_dptr(f) for f in fds])
    check(load().pgc_graphs_set_feat_diff(handles, fdp, n, int(dtype)))


def cycle_ratio_values(ctx, graphs, fds, cycles):
    """Raw-unit ratios sum c / sum t of cycles (edge-id arrays) of graphs, for
    feat_diffs fds (contiguous float64 device tensors or device addresses)."""
    n = len(graphs)
    cat, off, lengths = _cycles(cycles)
    handles = (C.c_void_p * n)(*[g._h.value for g in graphs])
    fdp = (C.c_void_p * n)(*[_dptr(f) for f in fds])
    vals = np.zeros(n, np.float64)
    check(load().pgc_cycle_ratio_values(ctx.handle, handles, n, fdp, ptr(cat), ptr(off), ptr(lengths), ptr(vals)))
    return vals


def cycle_ratio_grad(ctx, graphs, cycles, grad_out, dfds):
    """dfds[i] += grad_out[i] * d ratio_i / d feat_diff (float64 device tensors
    or device addresses)."""
    n = len(graphs)
    cat, off, lengths = _cycles(cycles)
    handles = (C.c_void_p * n)(*[g._h.value for g in graphs])
    dp = (C.c_void_p * n)(*[_dptr(d) for d in dfds])
    go = np.ascontiguousarray(grad_out, np.float64)
    check(load().pgc_cycle_ratio_grad(ctx.handle, handles, n, ptr(cat), ptr(off), ptr(lengths), ptr(go), dp))

