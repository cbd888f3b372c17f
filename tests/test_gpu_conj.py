"""Conjugate product graph on the GPU (through libpgc.so) against the oracle.

Bar: product space, SRC/TRGT bit-exact (same deterministic edge order);
energies within 1e-12 relative (the device's acos/pow may differ from the
host's libm in the last bit; without the rotation term they are bit-exact);
solves bit-exact against the oracle run on the same (device) energies: rows,
fp32 objective (1e-5 relative written; equal observed), sweep and branch
counters; branch graph equal to the oracle's.
"""
import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5


def _pair(sub, ny, seed):
    p = S.make_pair(sub, ny, seed=seed, with_normals=True)
    p["NX"] = p["NX"].astype(np.float64)
    p["NY"] = p["NY"].astype(np.float64)
    return p


def _gen(p, reg=False, cost="vanilla", time="", prune=False):
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, reg, prune)
    g.set_normals(p["NX"], p["NY"])
    g.set_cost_time_ratio_mode(cost, time)
    g.generate()
    return g


@pytest.mark.parametrize("prune", [False, True])
@pytest.mark.parametrize("reg,cost,time", [(False, "vanilla", "vanilla"), (True, "vanilla", "vanilla"),
                                           (True, "vanilla", ""), (True, "vanilla", "lengthnormalisation"),
                                           (True, "feature", "robustso3"), (True, "plainso3", "feature"),
                                           (True, "lengthnormalisation", "vanilla")])
def test_conj_graph_construction(reg, cost, time, prune):
    p = _pair(1, 7, seed=3)
    print("case conj construction", reg, cost, time, prune, flush=True)
    g = _gen(p, reg, cost, time, prune)
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"], prune=prune)
    assert np.array_equal(g.get_product_space(), rows)
    # numNodes = max(SRC) + 1 (taken in closed form for conjugate graphs):
    # numLayers = numNodes // npl, the reference's integer division
    info = g._graph.info()
    assert info["n_layers"] == (int(src.max()) + 1) // (3 * len(p["EX"]))
    en = O.conj_energy(p["VX"], p["NX"], p["VY"], p["NY"], p["EX"], p["EY"], rows, src, trgt,
                       p["feat_diff"].astype(np.float64), reg=reg, cost_mode=cost, time_mode=time, prune=prune)
    got = g.get_cost_vector()
    assert got.shape == en.shape
    if reg:
        np.testing.assert_allclose(got, en, rtol=1e-12, atol=1e-13)
    else:
        assert np.array_equal(got, en)


def test_conj_unsupported_mode_and_missing_normals():
    p = _pair(0, 4, seed=0)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
    with pytest.raises(ValueError):
        g.generate()  # no normals
    g.set_normals(p["NX"], p["NY"])
    g.set_cost_time_ratio_mode("feature", "feature")
    with pytest.raises(cpg.PgcError):
        g.generate()


def _oracle_solve_on(g, p, K, name):
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    en = g.get_cost_vector()  # the device energy, before any quantization
    npl = 3 * len(p["EX"])
    if name.startswith("lawler"):
        e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=name == "lawlergpu")
    else:
        bg = O.conj_branch_graph(p["EX"])
        e, obj, passes, stats = O.dijkstra_solve(src, trgt, en[:, 0], npl, K, branch_graph=bg,
                                                 cpu_rules=name == "dijkstracpu")
    return rows, e, obj, passes, stats


@pytest.mark.parametrize("sub,ny,K,seed", [(0, 5, 1, 0), (0, 6, 2, 1), (1, 6, 2, 2), (1, 5, 1, 3), (0, 7, 3, 4)])
@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_conj_solves(sub, ny, K, seed, name):
    p = _pair(sub, ny, seed)
    print("case conj solve", name, sub, ny, K, seed, flush=True)
    g = _gen(p, reg=True, cost="vanilla", time="vanilla")
    try:
        rows, e, obj, passes, stats = _oracle_solve_on(g, p, K, name)
    except RuntimeError as ex:
        assert "no nodes in branch" in str(ex)
        with pytest.raises(cpg.PgcError):
            g.solve_with_minimum_cost(name, K)
        return
    if name == "lawlergpu":
        gp, gobj, grows = g.solve_with_cost_time_ratio(name, K)
    else:
        gp, gobj, grows = g.solve_with_minimum_cost(name, K)
    assert gp == passes
    assert grows.shape[1] == 6
    assert np.array_equal(grows, rows[e]), "matching differs"
    assert abs(gobj - float(obj)) <= OBJ_RTOL * abs(float(obj))
    assert gobj == float(obj)
    st = g.last_result
    assert st["sweeps"] == stats[1] and st["branches"] == stats[3]


def test_conj_default_modes_two_columns():
    # reference defaults: cost "vanilla", time "" -> (vanilla, penalisedegenerate): two columns
    p = _pair(0, 5, seed=5)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
    g.set_normals(p["NX"], p["NY"])
    g.generate()
    assert g.get_cost_vector().shape[1] == 2
    rows, e, obj, passes, stats = _oracle_solve_on(g, p, 2, "lawlergpu")
    _, gobj, grows = g.solve_with_cost_time_ratio("lawlergpu", 2)
    assert np.array_equal(grows, rows[e]) and gobj == float(obj)
