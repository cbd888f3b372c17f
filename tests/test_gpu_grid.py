"""The device-wide solve against the oracle.

Graphs whose distance table lives in global memory and that have at least
128 chunks of 64 nodes per layer are solved by ceil(chunks / 16) workgroups
of one cooperative launch (one chunk per wave, device-wide barriers between
the steps of a sweep). Same bar as the parity tests: rows bit-exact, fp32
objective equal (1e-5 relative written), sweep, Bellman-Ford call and branch
counters equal.

Cases: conjugate graphs of 11,520 and 46,080 nodes per layer (12 and 45
workgroups) with both solvers; and the small parity cases and the branching
instances in a subprocess that forces the global table and a three-workgroup
device-wide solve for every single solve (PGC_GLOBAL_TABLE=1, PGC_GRID=3).
The full-size conjugate configuration (46,080 nodes per layer, 45
workgroups) is in tests/test_gpu_configs.py.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("sub,ny,K,cost,time", [(3, 24, 2, "vanilla", "nomode"), (3, 10, 3, "vanilla", "vanilla"),
                                                 (4, 16, 2, "vanilla", "nomode")])
@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_conjugate_device_wide(sub, ny, K, cost, time, name):
    p = S.make_pair(sub, ny, seed=sub + ny, with_normals=True)
    print("case grid conj", name, sub, ny, K, cost, time, flush=True)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
    g.set_normals(p["NX"].astype(np.float64), p["NY"].astype(np.float64))
    g.set_cost_time_ratio_mode(cost, time)
    g.generate()
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    npl = 3 * len(p["EX"])
    en = g.get_cost_vector()
    if name == "lawlergpu":
        e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=True)
        gp_, gobj, grows = g.solve_with_cost_time_ratio(name, K)
    else:
        bg = O.conj_branch_graph(p["EX"])
        e, obj, passes, stats = O.dijkstra_solve(src, trgt, en[:, 0], npl, K, branch_graph=bg)
        gp_, gobj, grows = g.solve_with_minimum_cost(name, K)
    assert gp_ == passes
    assert np.array_equal(grows, rows[e]), "matching differs"
    assert abs(gobj - float(obj)) <= 1e-5 * abs(float(obj)) and gobj == float(obj)
    st = g.last_result
    assert (st["sweeps"], st["bf_calls"], st["branches"]) == (stats[1], stats[0], stats[3])


def test_small_cases_device_wide_forced():
    env = dict(os.environ, PGC_GLOBAL_TABLE="1", PGC_GRID="3")
    cmd = [sys.executable, "-u", "-m", "pytest", "-x", "-q", "-m", "gpu", "-p", "no:cacheprovider",
           os.path.join(HERE, "test_gpu_parity.py"), os.path.join(HERE, "test_gpu_bnb.py"),
           "-k", "tiny or meshes or modes or pruned or full_slot_width or corrected or single_solves"]
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(HERE), timeout=900)
    assert r.returncode == 0
