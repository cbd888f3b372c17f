"""The device-wide solve against the oracle.

Graphs whose distance table lives in global memory and that have at least
128 chunks of 64 nodes per layer are solved by ceil(chunks / 16) workgroups
of one cooperative launch (one chunk per wave, device-wide barriers between
the steps of a sweep). Same bar as the parity tests: rows bit-exact, fp32
objective equal (1e-5 relative written), sweep, Bellman-Ford call and branch
counters equal.

Cases: conjugate graphs of 11,520 and 46,080 nodes per layer (12 and 45
workgroups) with both solvers; and the small parity cases and the branching
instances with the context's launch choices forcing the global table and a
three-workgroup device-wide solve for every single solve (global_table = 1,
grid_workgroups = 3), in this process: a second process on the GPU cannot
make cooperative launches while this one holds the device. The full-size
conjugate configuration (46,080 nodes per layer, 45 workgroups) is in
tests/test_gpu_configs.py.
"""
import inspect
import itertools

import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
import test_gpu_bnb
import test_gpu_parity
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu


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


def _cases(fn):
    """The argument sets of a test function's parametrize marks (their product)."""
    axes = []
    for m in getattr(fn, "pytestmark", []):
        if m.name != "parametrize":
            continue
        names = [n.strip() for n in m.args[0].split(",")] if isinstance(m.args[0], str) else list(m.args[0])
        vals = [getattr(v, "values", v) for v in m.args[1]]
        axes.append([dict(zip(names, v if len(names) > 1 else (v,))) for v in vals])
    for combo in itertools.product(*axes):
        kw = {}
        for d in combo:
            kw.update(d)
        yield kw


SELECT = ("tiny", "meshes", "modes", "pruned", "full_slot_width", "corrected", "single_solves")


def test_small_cases_device_wide_forced():
    ctx = _lib.default_context(0)
    saved = ctx.tuning()
    ctx.set_tuning({"global_table": 1, "grid_workgroups": 3})
    n = 0
    try:
        for mod in (test_gpu_parity, test_gpu_bnb):
            for name in sorted(dir(mod)):
                fn = getattr(mod, name)
                if not (name.startswith("test_") and callable(fn) and any(k in name for k in SELECT)):
                    continue
                takes_mp = "monkeypatch" in inspect.signature(fn).parameters
                for kw in _cases(fn):
                    # device-wide solves always take the host branch and
                    # bound: the other driver settings would repeat the case
                    if kw.get("driver", "device") != "device":
                        continue
                    print("case forced grid", name, kw, flush=True)
                    if takes_mp:
                        mp = pytest.MonkeyPatch()
                        try:
                            fn(monkeypatch=mp, **kw)
                        finally:
                            mp.undo()
                    else:
                        fn(**kw)
                    n += 1
    finally:
        ctx.set_tuning(saved)
    assert n >= 150


@pytest.mark.parametrize("nblk", [3, 6, 12])
def test_c1_size_with_idle_waves(nblk):
    # a C1-size pair (icosphere s=4 x 128-point curve, max_depth 3; 41 chunks
    # per layer) solved device-wide on 3 to 12 workgroups (48 to 192 waves,
    # most of them without a chunk), against the one-workgroup LDS solve.
    # Regression: the predecessors that the last steps of a sweep store must
    # reach every workgroup before the leader walks them (a full barrier
    # after the sweep, and every wave's stores drained before a barrier).
    ctx = _lib.default_context(0)
    p = S.make_pair(4, 128, seed=0)
    fd = p["feat_diff"].astype(np.float64)

    def solve():
        g = _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], fd, False, False, False, "vanilla", "")
        out = g.solve_ratio(3, _lib.PGC_LAWLER_PICK_LAST)
        g.close()
        return out

    e0, r0, res0 = solve()
    saved = ctx.tuning()
    ctx.set_tuning({"global_table": 1, "grid_workgroups": nblk})
    try:
        print("case c1 grid", nblk, flush=True)
        e1, r1, res1 = solve()
    finally:
        ctx.set_tuning(saved)
    assert np.array_equal(e1, e0) and np.array_equal(r1, r0)
    assert res1.objective == res0.objective
    assert (res1.sweeps, res1.bf_calls, res1.branches) == (res0.sweeps, res0.bf_calls, res0.branches)
