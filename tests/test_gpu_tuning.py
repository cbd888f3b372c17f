"""Launch choices (pgc_tuning, set per context) never change a result.

Every knob of the tuning struct is set, one at a time, to each value the
autotuner tries, and graphs that reach it are solved by both solvers against
the oracle: rows bit-exact, fp32 objective equal (1e-5 relative written),
sweep, Bellman-Ford call and branch counters equal. Graphs: a max_depth-1
ring (chain-origin rows, the unrolled sweep), an icosphere pair at max_depth
3 (cached sources, short slots, predecessor table and wrap map in LDS), and a
conjugate graph beyond the LDS (the device-wide solve with 1..4 chunks per
wave, or one workgroup). Argument checks of pgc_context_set_tuning.
"""
import numpy as np
import pytest

import pgc_oracle as O
from cyclic_product_graphs import _lib, autotune
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu


def _cases():
    return [("ring K1", S.make_pair(2, 24, seed=3), 1, False),
            ("ico K3", S.make_pair(2, 30, seed=5), 3, False),
            ("conj", S.make_pair(3, 12, seed=2, with_normals=True), 2, True)]


def _oracle(p, K, conj, en, mincost):
    if conj:
        rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
        npl = 3 * len(p["EX"])
    else:
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        npl = len(p["VX"])
    if mincost:
        bg = O.conj_branch_graph(p["EX"]) if conj else None
        e, obj, _, stats = O.dijkstra_solve(src, trgt, en[:, 0], npl, K, branch_graph=bg)
    else:
        e, obj, _, stats, _ = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=True)
    return rows, e, obj, stats


def _graph(ctx, p, conj):
    if conj:
        return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), True,
                                   True, False, "vanilla", "nomode", NX=p["NX"].astype(np.float64),
                                   NY=p["NY"].astype(np.float64))
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                               False, False, "vanilla", "")


KNOB_VALUES = [(k, v) for k, vs in autotune.KNOBS.items() for v in vs if v != -1]


@pytest.mark.parametrize("knob,value", KNOB_VALUES)
def test_knob_keeps_results(knob, value):
    ctx = _lib.Context(0)
    ctx.set_tuning({knob: value})
    assert ctx.tuning()[knob] == value
    for name, p, K, conj in _cases():
        for mincost in (False, True):
            print("case", knob, value, name, "mincost" if mincost else "ratio", flush=True)
            g = _graph(ctx, p, conj)
            en = g.cost()
            rows, e, obj, stats = _oracle(p, K, conj, en, mincost)
            if mincost:
                eids, rws, res = g.solve_min_cost(K)
            else:
                eids, rws, res = g.solve_ratio(K)
            assert np.array_equal(rws, rows[e]), "matching differs"
            assert abs(res.objective - float(obj)) <= 1e-5 * abs(float(obj)) and res.objective == float(obj)
            assert (res.sweeps, res.branches) == (stats[1], stats[3])
            g.close()
    ctx.set_tuning(None)
    assert ctx.tuning() == {k: -1 for k in _lib.pgc_tuning.FIELDS}
    ctx.close()


def test_batch_under_each_knob():
    ctx = _lib.Context(0)
    cases = [(S.make_pair(2, 24, seed=3), 1), (S.make_pair(2, 30, seed=5), 3), (S.make_pair(1, 9, seed=1), 2)]
    ref = None
    for knob, value in [(None, None)] + KNOB_VALUES:
        ctx.set_tuning({knob: value} if knob else None)
        graphs = [_graph(ctx, p, False) for p, _ in cases]
        outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K in cases])
        got = [(o[0].tobytes(), o[2].objective, o[2].sweeps) for o in outs]
        if ref is None:
            ref = got
            for (p, K), (e, rws, res) in zip(cases, outs):
                rows, oe, oobj, stats = _oracle(p, K, False, O.energy(p["VX"], p["VY"], O.product_space(
                    p["EX"], p["EY"])[0], p["feat_diff"].astype(np.float64)), False)
                assert np.array_equal(rws, rows[oe]) and res.objective == float(oobj)
        assert got == ref, (knob, value)
        for g in graphs:
            g.close()
    ctx.close()


def test_tuning_argument_checks():
    ctx = _lib.Context(0)
    with pytest.raises(ValueError):
        ctx.set_tuning({"no_such_knob": 0})
    with pytest.raises(_lib.PgcError):
        ctx.set_tuning({"pred_in_lds": 1})
    with pytest.raises(_lib.PgcError):
        ctx.set_tuning({"grid_chunks_per_wave": -2})
    with pytest.raises(_lib.PgcError):
        ctx.set_tuning({"global_table": 2})
    with pytest.raises(_lib.PgcError):
        ctx.set_tuning({"grid_workgroups": 257})
    ctx.set_tuning({"global_table": 1, "grid_workgroups": 3})
    assert ctx.tuning()["global_table"] == 1 and ctx.tuning()["grid_workgroups"] == 3
    ctx.set_tuning(None)
    assert set(ctx.tuning().values()) == {-1}
    ctx.close()
