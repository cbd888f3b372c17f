"""The solver variant with the distance table in global memory (graphs whose
table exceeds the LDS or that have more than 8 64-vertex chunks per wave)
against the oracle. Same bar as the parity tests: rows bit-exact, fp32
objective equal (1e-5 relative written), sweep and branch counters equal.

Cases: graphs that need the variant (non-conjugate max_depth 4 on an
icosphere s=4, conjugate graphs of 2880 and 11520 nodes per layer), a batch
mixing both variants, and the small parity cases in a subprocess that forces
the variant for every graph (PGC_GLOBAL_TABLE=1). Also the single-solve median
of columns of 1 M keys and more (gathered radix passes; regions of the whole
column below 8 M keys, of an eighth above), whose quantized costs must equal
the oracle's bit for bit.
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

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _check(g, p, K, name, conj=False):
    if conj:
        rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
        npl = 3 * len(p["EX"])
    else:
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        npl = len(p["VX"])
    en = g.get_cost_vector()
    if name == "lawlergpu":
        e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=True)
        gp_, gobj, grows = g.solve_with_cost_time_ratio(name, K)
    else:
        bg = O.conj_branch_graph(p["EX"]) if conj else None
        e, obj, passes, stats = O.dijkstra_solve(src, trgt, en[:, 0], npl, K, branch_graph=bg)
        gp_, gobj, grows = g.solve_with_minimum_cost(name, K)
    assert gp_ == passes
    assert np.array_equal(grows, rows[e]), "matching differs"
    assert abs(gobj - float(obj)) <= 1e-5 * abs(float(obj)) and gobj == float(obj)
    st = g.last_result
    assert st["sweeps"] == stats[1] and st["branches"] == stats[3]


@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_noncon_depth4_exceeds_lds(name):
    # (2K+1)*(npl+1)*8 = 9*2563*8 B = 184 KB > 160 KB of LDS
    p = S.make_pair(4, 12, seed=1)
    print("case big noncon", name, flush=True)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], False, False)
    _check(g, p, 4, name)


@pytest.mark.parametrize("sub,ny,K", [(2, 6, 3), (3, 4, 2)])
@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_conj_exceeds_lds(sub, ny, K, name):
    p = S.make_pair(sub, ny, seed=sub, with_normals=True)
    print("case big conj", name, sub, ny, K, flush=True)
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
    g.set_normals(p["NX"].astype(np.float64), p["NY"].astype(np.float64))
    g.set_cost_time_ratio_mode("vanilla", "vanilla")
    _check(g, p, K, name, conj=True)


def test_batch_mixing_both_variants():
    ctx = _lib.default_context(0)
    cases = [(S.make_pair(4, 12, seed=1), 4), (S.make_pair(2, 20, seed=3), 2), (S.make_pair(4, 16, seed=2), 4)]
    graphs = [_lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                                  False, False, "vanilla", "") for p, _ in cases]
    outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K in cases])
    for (p, K), (e, rows, res) in zip(cases, outs):
        orows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], orows, p["feat_diff"].astype(np.float64))
        oe, oobj, _, stats, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
        assert np.array_equal(rows, orows[oe]) and res.objective == float(oobj)
        assert res.sweeps == stats[1]
    for g in graphs:
        g.close()


def test_small_cases_with_global_table_forced():
    env = dict(os.environ, PGC_GLOBAL_TABLE="1")
    cmd = [sys.executable, "-m", "pytest", "-x", "-q", "-m", "gpu", "-p", "no:cacheprovider",
           os.path.join(HERE, "test_gpu_parity.py"), os.path.join(HERE, "test_gpu_batch.py"),
           "-k", "tiny or meshes or modes or pruned or matches_single or requantizes or sweep_cap or full_slot_width"]
    # output goes straight to this process's stdout (progress stays visible)
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(HERE), timeout=900)
    assert r.returncode == 0


@pytest.mark.parametrize("kind", ["normal", "overflow", "odd"])
@pytest.mark.parametrize("pair", [(4, 256), (3, 256)], ids=["8M", "2M"])
def test_large_column_median_gather(kind, pair):
    # single solves of columns of 1 M keys and more: the median's third radix
    # pass gathers the keys that share the median's top 16 bits and the later
    # passes read them (each workgroup its own region). Above 8 M keys a
    # region holds an eighth of the column and can overflow (almost all keys
    # equal: every later pass then reads the column); below, it holds the
    # whole column. The quantized costs must equal the oracle's quantization
    # (numpy-pinned median and rounding)
    p = S.make_pair(*pair, seed=11)
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    npl = len(p["VX"])
    E = len(src)
    assert E >= (8 << 20 if pair[0] == 4 else 1 << 20)
    rng = np.random.default_rng(4)
    if kind == "overflow":
        col = np.full(E, 3.0)
        col[rng.integers(0, E, 1000)] = 7.0
    else:
        col = rng.uniform(0.1, 20.0, E)
    if kind == "odd":
        rows, src, trgt, col = rows[:-1], src[:-1], trgt[:-1], col[:-1]
    print("case large column", kind, len(col), flush=True)
    ctx = _lib.default_context(0)
    g = _lib.Graph.from_arrays(ctx, rows, src, trgt, col, npl)
    g.solve_ratio(1)
    assert np.array_equal(g.cost()[:, 0], O.quantize(col[:, None])[:, 0])
    g.close()
