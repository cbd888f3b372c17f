"""The solver's in-edge layout: the closed-form builds of generated product
graphs against the generic build from SRC/TRGT and against a numpy model of
the layout.

The reference keeps, per node and class, the set of in-edges
(buildingInGraph, cost_time_ratio_solver.cpp:42-66; VECSET,
ctr_bnb_subroutine.cuh:232-287). Here every node's in-edges of a class sit in
the slots of its lane in a 64-node chunk, ordered by (source index inside
its layer, edge id), padding pointing at the sentinel npl with edge id -1;
an edge into layer 0 from a source id above npl moves to the closing layer L
(the closing copies), an edge that is neither intralayer nor interlayer is
dropped. The model below states that rule directly in numpy. Generated
non-conjugate graphs with three or more Y edges take the closed-form build
(the X graph's slots copied to every layer), conjugate graphs with three or
more layers a build from their first two layers (every layer repeats one
pattern); PGC_GENERIC_LAYOUT=1 forces the generic build.
Bar: slot arrays equal element for element (integer data), and the solves of
both builds equal (rows, objective, counters).
"""
import numpy as np
import pytest

import pgc_oracle as O
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu


def model_layout(src, trgt, npl, W):
    """(src slots, eid slots) of both classes, as the layout rule states."""
    src = src.astype(np.int64)
    trgt = trgt.astype(np.int64)
    numNodes = int(src.max()) + 1
    L = numNodes // npl
    nLayers = L + 1
    nC = (npl + 63) // 64
    t = np.where((trgt < npl) & (src > npl), trgt + numNodes, trgt)
    layer = t // npl
    ls = src // npl
    cls = np.where(ls == layer, 0, np.where(ls + 1 == layer, 1, -1))
    keep = (cls >= 0) & (t < numNodes + npl)
    e = np.nonzero(keep)[0]
    cls, t, layer, sl = cls[keep], t[keep], layer[keep], (src - np.where(cls == 0, layer, layer - 1) * npl)[keep]
    order = np.lexsort((e, sl, t, cls))
    cls, t, layer, sl, e = cls[order], t[order], layer[order], sl[order], e[order]
    key = cls * (nLayers * npl) + t
    first = np.r_[True, key[1:] != key[:-1]]
    start = np.maximum.accumulate(np.where(first, np.arange(len(key)), 0))
    pos = np.arange(len(key)) - start
    assert pos.max() < W
    v = t - layer * npl
    slot = ((layer * nC + v // 64) * W + pos) * 64 + v % 64
    out = []
    for c in (0, 1):
        n = nLayers * nC * W * 64
        s_ = np.full(n, npl, np.int32)
        e_ = np.full(n, -1, np.int32)
        m = cls == c
        s_[slot[m]] = sl[m]
        e_[slot[m]] = e[m]
        out.append((s_, e_))
    return out


def _graph(p, prune=False):
    ctx = _lib.default_context(0)
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                               False, prune, "vanilla", "nomode")


def _ring_with_extras(n_x=7, n_y=5, seed=0):
    """A ring X graph with a repeated edge and a self-loop (multi-edges keep
    their edge-id order inside a node's slots)."""
    p = S.tiny_pair(n_x, n_y, seed=seed)
    EX = np.concatenate([p["EX"], [[2, 3], [4, 4], [0, 1]]]).astype(np.int32)
    p["EX"] = EX
    return p


CASES = [
    ("ico1 y3", lambda: S.make_pair(1, 3, seed=1), False),
    ("ico1 y5 pruned", lambda: S.make_pair(1, 5, seed=2), True),
    ("ico2 y12", lambda: S.make_pair(2, 12, seed=3), False),
    ("ico2 y9 pruned", lambda: S.make_pair(2, 9, seed=4), True),
    ("ring extras y5", lambda: _ring_with_extras(), False),
    ("ring extras y4 pruned", lambda: _ring_with_extras(9, 4, 1), True),
    ("bipyramid40 y4", lambda: S.make_pair(0, 4, seed=5, mesh=S.bipyramid(40)), False),
    ("ico1 y2 (generic: wrap test)", lambda: S.make_pair(1, 2, seed=6), False),
]


@pytest.mark.parametrize("name,make,prune", CASES, ids=[c[0] for c in CASES])
def test_closed_form_layout_equals_generic_and_model(name, make, prune, monkeypatch):
    print("case", name, flush=True)
    p = make()
    g_closed = _graph(p, prune)
    monkeypatch.setenv("PGC_GENERIC_LAYOUT", "1")
    g_generic = _graph(p, prune)
    lay_g = [g_generic.layout(c) for c in (0, 1)]
    monkeypatch.delenv("PGC_GENERIC_LAYOUT")
    lay_c = [g_closed.layout(c) for c in (0, 1)]
    _, src, trgt = O.product_space(p["EX"], p["EY"], prune)
    npl = len(p["VX"])
    W = lay_c[0][2]
    assert lay_c[1][2] == W and lay_g[0][2] == W and lay_g[1][2] == W
    model = model_layout(src, trgt, npl, W)
    for c in (0, 1):
        assert np.array_equal(lay_c[c][0], lay_g[c][0]), "closed-form sources differ from the generic build"
        assert np.array_equal(lay_c[c][1], lay_g[c][1]), "closed-form edge ids differ from the generic build"
        assert np.array_equal(lay_c[c][0], model[c][0]), "sources differ from the layout rule"
        assert np.array_equal(lay_c[c][1], model[c][1]), "edge ids differ from the layout rule"


def _conj_graph(p, prune=False):
    ctx = _lib.default_context(0)
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), True,
                               True, prune, "vanilla", "", NX=p["NX"].astype(np.float64),
                               NY=p["NY"].astype(np.float64))


CONJ = [
    ("conj ico0 y3", lambda: S.make_pair(0, 3, seed=11, with_normals=True)),
    ("conj ico1 y5", lambda: S.make_pair(1, 5, seed=12, with_normals=True)),
    ("conj ico1 y8", lambda: S.make_pair(1, 8, seed=13, with_normals=True)),
    ("conj bipyramid12 y4", lambda: S.make_pair(0, 4, seed=14, mesh=S.bipyramid(12), with_normals=True)),
    ("conj ico0 y2 (generic)", lambda: S.make_pair(0, 2, seed=15, with_normals=True)),
]


@pytest.mark.parametrize("name,make", CONJ, ids=[c[0] for c in CONJ])
def test_conjugate_layout_from_two_layers(name, make, monkeypatch):
    # conjugate graphs repeat one layer's in-edge pattern: the layout is built
    # from layers 0 and 1 and copied to every layer
    print("case", name, flush=True)
    p = make()
    g_rep = _conj_graph(p)
    monkeypatch.setenv("PGC_GENERIC_LAYOUT", "1")
    g_gen = _conj_graph(p)
    lay_g = [g_gen.layout(c) for c in (0, 1)]
    monkeypatch.delenv("PGC_GENERIC_LAYOUT")
    lay_r = [g_rep.layout(c) for c in (0, 1)]
    _, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    npl = 3 * len(p["EX"])
    W = lay_r[0][2]
    model = model_layout(src, trgt, npl, W)
    for c in (0, 1):
        assert lay_g[c][2] == W
        assert np.array_equal(lay_r[c][0], lay_g[c][0]) and np.array_equal(lay_r[c][1], lay_g[c][1])
        assert np.array_equal(lay_r[c][0], model[c][0]) and np.array_equal(lay_r[c][1], model[c][1])


@pytest.mark.parametrize("K", [1, 2, 3])
def test_solves_equal_under_both_layouts(K, monkeypatch):
    p = S.make_pair(2, 10, seed=7, mesh=None)
    out = []
    for generic in (False, True):
        if generic:
            monkeypatch.setenv("PGC_GENERIC_LAYOUT", "1")
        g = _graph(p)
        e, rows, res = g.solve_ratio(K)
        g2 = _graph(p)
        e2, rows2, res2 = g2.solve_min_cost(K)
        out.append((e, rows, res.objective, res.sweeps, res.bf_calls, e2, res2.objective, res2.sweeps))
        monkeypatch.delenv("PGC_GENERIC_LAYOUT", raising=False)
    a, b = out
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert np.array_equal(x, y)
        else:
            assert x == y
