"""MI355X-native product-graph cycle solvers with the reference's Python API.

Mirrors the pybind11 module `cyclic_product_graphs` of the reference
(pythonbinding/bindings.cpp:22-50) and the state semantics of
ProductGraphGenerators (product_graph_generators.cpp:93-333):

  * the graph is built lazily on the first solve / getter and then cached, so
    set_cost_time_ratio_mode after that has no effect (:289-291, :309-312);
  * a Lawler solve quantizes the generator's energy in place to q-units
    (cost_time_ratio_solver.cpp:875-882); later get_cost_vector / min-cost
    solves see the quantized values;
  * max_depth defaults to 4 (:109);
  * "export" writes VX/VY/EX/EY/FeatDiffMatrix(/NX/NY).csv to the working
    directory (:272-282).

Solvers run on the GPU through libpgc.so (C-ABI, include/pgc.h):
  lawlergpu   - Lawler ratio cycle, candidate pick of the reference's GPU path
  lawlercpu   - same device solver with the reference CPU path's candidate order
  dijkstragpu - minimum-cost once-wrapping cycle, rules of the GPU path
  dijkstracpu - same device solver with the CPU path's rules
An unknown Lawler name raises ValueError (the reference exits the process);
an unknown min-cost name falls back to dijkstracpu rules, as the reference does
(dijkstra_solver.cpp:576-578).
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import PgcError  # noqa: F401

__all__ = ["product_graph_generator", "compute_normals_3d", "compute_normals_2d", "compute_local_thickness_3d",
           "compute_local_thickness_2d", "PgcError"]

_LAWLER = {"lawlergpu": _lib.PGC_LAWLER_PICK_LAST, "lawlercpu": 0}
_MINCOST = {"dijkstragpu": 0, "dijkstracpu": _lib.PGC_MINCOST_CPU_RULES}


def _write_csv(mat, name):
    # Eigen::IOFormat(FullPrecision, DontAlignCols, ", ", "\n") (helper/utils.hpp:24-32)
    mat = np.asarray(mat)
    with open(name + ".csv", "w") as f:
        lines = []
        for row in np.atleast_2d(mat):
            if np.issubdtype(mat.dtype, np.integer):
                lines.append(", ".join(str(int(x)) for x in row))
            else:  # std::ostream with precision max_digits10 (17), general format
                lines.append(", ".join("%.17g" % float(x) for x in row))
        f.write("\n".join(lines))


class product_graph_generator:
    """product_graph_generator(VX, EX, VY, EY, feat_diff, use_conjugate_graph,
    use_regularising_cost_term[, prune_intralayer_edges])"""

    def __init__(self, VX, EX, VY, EY, feat_diff, use_conjugate_graph, use_regularising_cost_term,
                 prune_intralayer_edges=False, device=0):
        self.VX = np.array(VX, dtype=np.float64)
        self.EX = np.array(EX, dtype=np.int32)
        self.VY = np.array(VY, dtype=np.float64)
        self.EY = np.array(EY, dtype=np.int32)
        self.feat_diff = np.array(feat_diff, dtype=np.float64)
        self.conjugate = bool(use_conjugate_graph)
        self.regularising = bool(use_regularising_cost_term)
        self.prune = bool(prune_intralayer_edges)
        self.rl = (0.7, 0.6, 4.0)
        self.max_depth = 4
        self.cost_name = "vanilla"
        self.time_name = ""
        self.NX = None
        self.NY = None
        self.device = device
        self._graph = None
        self.last_result = None
        if self.VX.ndim != 2 or self.VX.shape[1] != 3:
            raise ValueError("VX must be |VX| x 3")
        if self.VY.ndim != 2 or self.VY.shape[1] != 3:
            raise ValueError("VY must be |VY| x 3 (z = 0 for planar curves)")
        if self.EX.ndim != 2 or self.EX.shape[1] != 2 or self.EY.ndim != 2 or self.EY.shape[1] != 2:
            raise ValueError("EX and EY must have two columns")
        if self.feat_diff.shape != (len(self.VX), len(self.VY)):
            raise ValueError("feat_diff must be |VX| x |VY|")

    # ------------------------------------------------------------- setters
    def set_normals(self, NX, NY):
        NX = np.array(NX, dtype=np.float64)
        NY = np.array(NY, dtype=np.float64)
        if NX.shape != (len(self.VX), 3):
            print("[ShapeMM] error: normals for shape X do not contain as many entries as expected. "
                  "Assumed shape = |VX| x 3")
            return
        if NY.shape != (len(self.VY), 3):
            print("[ShapeMM] error: normals for shape Y do not contain as many entries as expected. "
                  "Assumed shape = |VY| x 3")
            return
        self.NX, self.NY = NX, NY

    def set_max_depth_cost_time_ratio(self, max_depth):
        self.max_depth = int(max_depth)

    def set_cost_time_ratio_mode(self, cost_mode, time_mode):
        self.cost_name = str(cost_mode)
        self.time_name = str(time_mode)

    def update_robust_loss_params(self, alpha, c, pwr):
        if not self.conjugate:
            return
        if self._graph is not None:
            print("[ShapeMM] error: cannot update robust loss params after model has been generated")
            return
        self.rl = (float(alpha), float(c), float(pwr))

    # ------------------------------------------------------------- graph
    def _ctx(self):
        return _lib.default_context(self.device)

    def generate(self):
        if self.conjugate and (self.NX is None or self.NY is None):
            raise ValueError("conjugate product graph needs normals: call set_normals(NX, NY) first")
        self._graph = _lib.Graph.generate(self._ctx(), self.VX, self.EX, self.VY, self.EY, self.feat_diff,
                                          self.conjugate, self.regularising, self.prune, self.cost_name,
                                          self.time_name, self.NX, self.NY, self.rl)

    def _ensure(self):
        if self._graph is None:
            self.generate()
        return self._graph

    def get_cost_vector(self):
        return self._ensure().cost()

    def get_product_space(self):
        return self._ensure().product_space()

    def _export(self):
        _write_csv(self.VX, "VX")
        _write_csv(self.VY, "VY")
        _write_csv(self.EX, "EX")
        _write_csv(self.EY, "EY")
        _write_csv(self.feat_diff, "FeatDiffMatrix")
        if self.conjugate and self.NX is not None:
            _write_csv(self.NX, "NX")
            _write_csv(self.NY, "NY")
        return False, -1.0, np.zeros((4, 1), np.int32)

    # ------------------------------------------------------------- solvers
    def solve_with_cost_time_ratio(self, solvername, max_depth=None, corrected=False):
        """corrected (an addition to the reference's signature): the odd-L fold
        reads the closing layer's depth-0 row (PGC_CORRECTED_ODD_L); the
        default keeps the reference's row L % 2."""
        if solvername == "export":
            return self._export()
        if solvername not in _LAWLER:
            raise ValueError("Solver >>%s<< is not implemented" % solvername)
        g = self._ensure()
        K = self.max_depth if max_depth is None else int(max_depth)
        flags = _LAWLER[solvername] | (_lib.PGC_CORRECTED_ODD_L if corrected else 0)
        _, rows, res = g.solve_ratio(K, flags)
        self.last_result = res.as_dict()
        return bool(res.passes_layers_once), float(np.float32(res.objective)), rows

    def solve_with_minimum_cost(self, solvername, max_depth=None, corrected=False):
        """corrected: odd-L lower bounds from the closing layer's depth-0 row
        (PGC_CORRECTED_ODD_L), an addition to the reference's signature."""
        if solvername == "export":
            return self._export()
        if solvername not in _MINCOST:
            print("Solver >>%s<< is not implemented\nFalling back to >> dijkstracpu <<" % solvername)
            flags = _MINCOST["dijkstracpu"]
        else:
            flags = _MINCOST[solvername]
        if corrected:
            flags |= _lib.PGC_CORRECTED_ODD_L
        g = self._ensure()
        K = self.max_depth if max_depth is None else int(max_depth)
        _, rows, res = g.solve_min_cost(K, flags)
        self.last_result = res.as_dict()
        return bool(res.passes_layers_once), float(np.float32(res.objective)), rows


# ------------------------------------------------------------- helpers
def _normalize_rows(M):
    """Each row divided by its Euclidean norm, with no guard: a zero row
    becomes NaN, as Eigen 3.4's rowwise normalize() divides by the norm."""
    nrm = np.sqrt(np.add.reduce(M * M, axis=1))
    with np.errstate(invalid="ignore", divide="ignore"):
        return M / nrm[:, None]


def compute_normals_3d(V, F):
    """Per-vertex normals (utils::normals3d, helper/utils.cpp:86-90), which is
    igl::per_vertex_normals(V, F, N) of libigl b1bd5b1 with its default
    weighting: every face adds its unit normal ((v1-v0) x (v2-v0), zero for a
    degenerate face) times its double area to its three corners, face by face;
    the sums are normalised (a vertex on no face gets NaN)."""
    V = np.asarray(V, np.float64)
    F = np.asarray(F, np.int64)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    r = np.sqrt(np.add.reduce(n * n, axis=1))
    with np.errstate(invalid="ignore", divide="ignore"):
        fn = np.where(r[:, None] == 0.0, 0.0, n / r[:, None])
    # double area: the double areas projected on (x,y), (y,z), (z,x), each
    # from corner 2, combined as the root of their summed squares
    w = np.zeros(len(F))
    for d in range(3):
        e = (d + 1) % 3
        pa = (p0[:, d] - p2[:, d]) * (p1[:, e] - p2[:, e]) - (p0[:, e] - p2[:, e]) * (p1[:, d] - p2[:, d])
        w = w + pa * pa
    w = np.sqrt(w)
    N = np.zeros_like(V)
    # corners in face order (np.add.at applies repeated indices in order)
    np.add.at(N, F.reshape(-1), np.repeat(w[:, None] * fn, 3, axis=0))
    return _normalize_rows(N)


def compute_normals_2d(V, E):
    """Outward vertex normals of a closed planar curve (utils::normals2d,
    helper/utils.cpp:93-157).

    The curve must lie in a coordinate plane: the first coordinate column
    whose sum is below 1e-4 in magnitude is taken as the dropped one (a curve
    centred on the origin can make the wrong column qualify, as in the
    reference), and the columns are permuted so that it comes last; the
    normals are returned in that permuted frame, as the reference returns
    them. Edge normals are the edge vectors turned by -90 degrees about z,
    flipped when moving the points by a tenth of the smaller bounding-box side
    along them does not grow the box in +x (the box values and that tenth are
    float, as declared in the reference); a vertex normal is the
    length-weighted sum of its two edges' normals, normalised (NaN for a zero
    sum, as in Eigen's rowwise normalisation)."""
    V = np.array(V, np.float64)
    E = np.asarray(E, np.int64)
    if len(V) != len(E):
        print("error: V.rows() != E.rows()")
        return np.zeros((1, 1))
    zero = -1
    for i in range(3):
        if abs(V[:, i].sum()) < 1e-4:
            zero = i
            break
    if zero == -1:
        print("error: points dont lie on xy, xz or yz plane")
        return np.zeros((1, 1))
    if zero != 2:
        V = V[:, [0, 2, 1]] if zero == 1 else V[:, [1, 2, 0]]
    ev = V[E[:, 0]] - V[E[:, 1]]
    el = np.sqrt(np.add.reduce(ev * ev, axis=1))
    R = np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1]], np.float64)
    en = _normalize_rows(ev @ R.T)
    maxX, minX = np.float32(V[:, 0].max()), np.float32(V[:, 0].min())
    maxY, minY = np.float32(V[:, 1].max()), np.float32(V[:, 1].min())
    # float - float, then 0.1 * float in double, narrowed to float
    side = min(abs(maxX - minX), abs(maxY - minY))
    mult = np.float32(0.1 * float(side))
    if np.float32((V + float(mult) * en)[:, 0].max()) < maxX:
        en = -en
    n = len(V)
    vn = np.zeros((n, 3))
    for i in range(1, n + 1):
        j = i % n
        vn[j] = el[i - 1] * en[i - 1] + el[j] * en[j]
    return _normalize_rows(vn)


def compute_local_thickness_3d():
    print("TODO: IMPL")


def compute_local_thickness_2d():
    print("TODO: IMPL")
