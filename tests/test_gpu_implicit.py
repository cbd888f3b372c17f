"""Implicit costs: the max_depth-1 sweep recomputes every slot's q-cost from
feat_diff instead of loading the materialized slot costs.

With pgc_tuning.implicit_costs = 1 a generated non-conjugate graph with the
vanilla energy keeps its feat_diff by columns, and a ratio solve at max_depth 1
whose sweep holds its rows, predecessors and chain origins in LDS stages the
feat_diff column of each layer's sources in LDS and computes each slot's cost
as min(round(s * 0.5 (FD[x0, ys] + FD[x1, yd])), 1e7) with the graph's
quantization scale s (deformationEnergy.cpp:66-69, cost_time_ratio_solver.cpp
:709-721): the reference's energy and quantization, in its order of
operations. Bar: the cycle (edge ids, rows) bit-exact, the fp32 objective equal
(1e-5 relative written as the tolerance) and the Bellman-Ford call, sweep,
Lawler iteration and branch counters equal, against the same solve with slot
costs and against the oracle; the result reports which path ran.
"""
import os

import numpy as np
import pytest

import pgc_oracle as O
from branching_cases import BRANCHING
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
OBJ_RTOL = 1e-5

_ctx = {}


def _contexts():
    if not _ctx:
        ci = _lib.Context(0)
        ci.set_tuning({"implicit_costs": 1})
        _ctx["implicit"] = ci
        _ctx["slots"] = _lib.default_context(0)
    return _ctx["implicit"], _ctx["slots"]


def _gen(ctx, p, reg=False, cost="vanilla", time="", prune=False, fd=None):
    fd = p["feat_diff"] if fd is None else fd
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], np.asarray(fd, np.float64), False, reg,
                               prune, cost, time)


def _oracle(p, K, gpu_pick, fd=None, prune=False):
    fd = p["feat_diff"] if fd is None else fd
    rows, src, trgt = O.product_space(p["EX"], p["EY"], prune)
    en = O.energy(p["VX"], p["VY"], rows, np.asarray(fd, np.float64))
    e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=gpu_pick)
    return rows, e, obj, stats


def _same(a, b):
    (ea, ra, sa), (eb, rb, sb) = a, b
    assert np.array_equal(ea, eb), "cycle differs"
    assert np.array_equal(ra, rb)
    assert abs(sa.objective - sb.objective) <= OBJ_RTOL * abs(sb.objective)
    assert sa.objective == sb.objective
    assert (sa.bf_calls, sa.sweeps, sa.lawler_iters, sa.branches) == (sb.bf_calls, sb.sweeps, sb.lawler_iters,
                                                                      sb.branches)


def _vs_oracle(out, orc):
    e, rows, res = out
    orow, oe, oobj, stats = orc
    assert np.array_equal(e, oe), "cycle differs from the oracle"
    assert np.array_equal(rows, orow[oe])
    assert res.objective == float(oobj)
    assert (res.bf_calls, res.sweeps, res.lawler_iters, res.branches) == tuple(int(x) for x in stats[:4])


PAIRS = [
    ("ico1 y8", lambda: S.make_pair(1, 8, seed=1)),
    ("ico2 y16", lambda: S.make_pair(2, 16, seed=2)),
    ("ico2 y33", lambda: S.make_pair(2, 33, seed=3)),
    ("ico3 y41", lambda: S.make_pair(3, 41, seed=4)),
    ("ring9 y7", lambda: S.tiny_pair(9, 7, seed=5)),
    ("complete8 y6", lambda: S.tiny_pair(8, 6, seed=6, ring=False)),
    ("bipyramid6 y5", lambda: S.make_pair(0, 5, seed=7, mesh=S.bipyramid(6))),
]


@pytest.mark.parametrize("rule", ["lawlergpu", "lawlercpu"])
@pytest.mark.parametrize("name,make", PAIRS, ids=[c[0] for c in PAIRS])
def test_implicit_equals_slot_costs_and_oracle(name, make, rule):
    print("case implicit", name, rule, flush=True)
    ci, cs = _contexts()
    p = make()
    flags = _lib.PGC_LAWLER_PICK_LAST if rule == "lawlergpu" else 0
    a = _gen(ci, p).solve_ratio(1, flags)
    b = _gen(cs, p).solve_ratio(1, flags)
    assert a[2].implicit_costs == 1 and b[2].implicit_costs == 0
    _same(a, b)
    _vs_oracle(a, _oracle(p, 1, rule == "lawlergpu"))


@pytest.mark.parametrize("driver", ["device", "host"])
def test_implicit_branching(driver, monkeypatch):
    if driver == "host":
        monkeypatch.setenv("PGC_HOST_BNB", "1")
    ci, _ = _contexts()
    n = 0
    for seed, nx, ny, K in BRANCHING:
        if K != 1:
            continue
        print("case implicit branching", driver, seed, nx, ny, flush=True)
        p = S.tiny_pair(nx, ny, seed=seed, ring=True)
        out = _gen(ci, p).solve_ratio(1)
        assert out[2].implicit_costs == 1
        _vs_oracle(out, _oracle(p, 1, True))
        n += out[2].branches > 1
    assert n >= 5


def test_implicit_lifecycle():
    # a second solve re-quantizes the quantized energy (the reference's
    # in-place quantization): it loads slot costs; a new feat_diff makes the
    # energy fresh again; a derived graph keeps its own feat_diff
    ci, cs = _contexts()
    p = S.make_pair(2, 20, seed=8)
    fd2 = p["feat_diff"] * 1.25 + 0.01
    gi, gs = _gen(ci, p), _gen(cs, p)
    for _ in range(2):
        a, b = gi.solve_ratio(1), gs.solve_ratio(1)
        _same(a, b)
    gi.set_feat_diff(fd2)
    gs.set_feat_diff(fd2)
    a, b = gi.solve_ratio(1), gs.solve_ratio(1)
    assert a[2].implicit_costs == 1 and b[2].implicit_costs == 0
    _same(a, b)
    fd3 = p["feat_diff"] * 0.8
    di, ds = gi.derive(fd3), gs.derive(fd3)
    a, b = di.solve_ratio(1), ds.solve_ratio(1)
    assert a[2].implicit_costs == 1
    _same(a, b)
    _vs_oracle(a, _oracle(p, 1, True, fd=fd3))


def test_implicit_not_applied_where_it_does_not_hold():
    # max_depth 2, the length term, a time column, an explicitly set cost:
    # slot costs, same results as the default context
    ci, cs = _contexts()
    p = S.make_pair(1, 10, seed=9)
    for kw, K in (({}, 2), ({"reg": True}, 1), ({"cost": "feature", "time": "lengthnormalisation"}, 1)):
        a, b = _gen(ci, p, **kw).solve_ratio(K), _gen(cs, p, **kw).solve_ratio(K)
        assert a[2].implicit_costs == 0
        _same(a, b)
    g = _gen(ci, p)
    g.set_cost(g.cost())
    a = g.solve_ratio(1)
    assert a[2].implicit_costs == 0
    _same(a, _gen(cs, p).solve_ratio(1))


def test_implicit_sweep_cap_fixture():
    # the C4 batch item whose Bellman-Ford calls run to the sweep cap (26,363
    # sweeps), against its committed cycle
    ci, _ = _contexts()
    z = np.load(os.path.join(HERE, "golden", "bf_cap_case.npz"))
    p, K = S.make_uneven_pairs(256, seed=0)[int(z["pair_index"])]
    e, rows, res = _gen(ci, p, fd=z["feat_diff"]).solve_ratio(K)
    assert K == 1 and res.implicit_costs == 1
    assert np.array_equal(e, z["edges"])
    assert res.sweeps == int(z["stats"][1])
