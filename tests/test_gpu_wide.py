"""Nodes with more than 32 in-edges of one class: the runtime-width solver
variant against the oracle.

Above 32 in-edges the layout's slot width is the largest in-degree rounded up
to a multiple of 8, and the sweep scans each chunk's slots in groups of eight,
as many as the chunk's own largest in-degree needs, with the same
first-minimum rule (the reference flattens any in-degree, VECSET,
ctr_bnb_subroutine.cuh:232-287). X is a bipyramid over a 40-gon: its two apexes
have valence 40, so 41 interlayer in-edges (slot width 48); its conjugate
product graph has nodes with more than 80 in-edges. Same bar as the parity
tests: rows bit-exact, fp32 objective equal (1e-5 relative written), counters
equal. Cases: the non-conjugate graph at max_depth 1..3 under both rule sets
of both solvers, in the LDS variant and (subprocess) the forced global-table
variant; the conjugate graph with both solvers; a batch mixing the wide graph
with icosphere graphs.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S
from test_gpu_parity import LAWLER, MINCOST, _check_mincost, _check_ratio

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _wide(n_y=12, seed=1, normals=False):
    return S.make_pair(0, n_y, seed=seed, mesh=S.bipyramid(40), with_normals=normals)


@pytest.mark.parametrize("K", [1, 2, 3])
@pytest.mark.parametrize("name", list(LAWLER))
def test_noncon_wide_ratio(K, name):
    _check_ratio(_wide(), K, name)


@pytest.mark.parametrize("K", [1, 2, 3])
@pytest.mark.parametrize("name", list(MINCOST))
def test_noncon_wide_mincost(K, name):
    _check_mincost(_wide(), K, name)


@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_conj_wide(name):
    p = _wide(n_y=8, seed=2, normals=True)
    K = 2
    print("case wide conj", name, flush=True)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
    g.set_normals(p["NX"].astype(np.float64), p["NY"].astype(np.float64))
    g.set_cost_time_ratio_mode("vanilla", "nomode")
    g.generate()
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    npl = 3 * len(p["EX"])
    assert np.bincount(trgt).max() > 64
    en = g.get_cost_vector()
    if name == "lawlergpu":
        e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=True)
        gp_, gobj, grows = g.solve_with_cost_time_ratio(name, K)
    else:
        e, obj, passes, stats = O.dijkstra_solve(src, trgt, en[:, 0], npl, K, branch_graph=O.conj_branch_graph(p["EX"]))
        gp_, gobj, grows = g.solve_with_minimum_cost(name, K)
    assert gp_ == passes
    assert np.array_equal(grows, rows[e]), "matching differs"
    assert gobj == float(obj)
    st = g.last_result
    assert (st["sweeps"], st["branches"]) == (stats[1], stats[3])


def test_batch_with_wide_graph():
    ctx = _lib.default_context(0)
    cases = [(_wide(seed=3), 2), (S.make_pair(2, 20, seed=3), 2), (_wide(n_y=9, seed=4), 1), (S.make_pair(1, 8, seed=2), 3)]
    graphs = [_lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                                  False, False, "vanilla", "") for p, _ in cases]
    outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K in cases])
    for (p, K), (e, rows, res) in zip(cases, outs):
        orows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], orows, p["feat_diff"].astype(np.float64))
        oe, oobj, _, stats, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
        assert np.array_equal(rows, orows[oe]) and res.objective == float(oobj)
        assert (res.sweeps, res.branches) == (stats[1], stats[3])
    for g in graphs:
        g.close()


def test_wide_with_global_table_forced():
    env = dict(os.environ, PGC_GLOBAL_TABLE="1")
    cmd = [sys.executable, "-u", "-m", "pytest", "-x", "-q", "-m", "gpu", "-p", "no:cacheprovider",
           os.path.join(HERE, "test_gpu_wide.py"), "-k", "noncon or conj_wide or batch_with"]
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(HERE), timeout=600)
    assert r.returncode == 0
