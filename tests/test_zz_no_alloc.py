"""pgc_solve_ratio_noalloc against pgc_solve_ratio through the C-ABI.

The allocation-free entry point runs the whole solve in one launch with every
buffer beyond the graph's arrays and the outputs in the library's static
device scratch. Bar: on the same inputs (two graphs built from one pair, as
both entry points quantize the energy in place) the cycle's edge ids and rows
are bit-exact, the fp32 objective equal (1e-5 relative written, observed 0),
the quantized energies bit-exact and the sweep, Bellman-Ford call, Lawler
iteration and branch counters equal; on two calls in a row, the second on a
smaller graph (the scratch keeps the first call's contents). A graph too
large for the scratch is refused before its energy is touched.
"""
import numpy as np
import pytest

from branching_cases import BRANCHING
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5


def _graph(ctx, p, **kw):
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                               kw.get("reg", False), False, kw.get("cost", "vanilla"), kw.get("time", ""))


def _compare(ctx, p, K, flags=_lib.PGC_LAWLER_PICK_LAST, **kw):
    ga, gb = _graph(ctx, p, **kw), _graph(ctx, p, **kw)
    e1, r1, res1 = ga.solve_ratio(K, flags)
    e2, r2, res2 = gb.solve_ratio_noalloc(K, flags)
    assert np.array_equal(e1, e2) and np.array_equal(r1, r2), "cycle differs"
    assert abs(res2.objective - res1.objective) <= OBJ_RTOL * abs(res1.objective)
    assert res2.objective == res1.objective
    assert (res2.sweeps, res2.bf_calls, res2.lawler_iters, res2.branches) == \
        (res1.sweeps, res1.bf_calls, res1.lawler_iters, res1.branches)
    assert res2.status == 0
    assert np.array_equal(ga.cost(), gb.cost()), "quantized energy differs"
    ga.close()
    gb.close()


def test_noalloc_two_calls_second_smaller():
    ctx = _lib.default_context(0)
    for sub, ny, K, seed in [(3, 64, 2, 0), (2, 24, 3, 5)]:
        print("case noalloc", sub, ny, K, seed, flush=True)
        _compare(ctx, S.make_pair(sub, ny, seed=seed), K)


def test_noalloc_time_column_and_global_table_sizes():
    ctx = _lib.default_context(0)
    p = S.make_pair(2, 20, seed=7)
    print("case noalloc time column", flush=True)
    _compare(ctx, p, 2, cost="feature", time="lengthnormalisation")
    print("case noalloc C1 size", flush=True)
    _compare(ctx, S.make_pair(4, 128, seed=0), 3)


@pytest.mark.parametrize("flags", [_lib.PGC_LAWLER_PICK_LAST, 0])
def test_noalloc_branching(flags):
    ctx = _lib.default_context(0)
    for seed, nx, ny, K in BRANCHING[:12]:
        print("case noalloc branching", seed, nx, ny, K, flags, flush=True)
        _compare(ctx, S.tiny_pair(nx, ny, seed=seed, ring=True), K, flags)


def test_noalloc_too_large_is_refused_untouched():
    ctx = _lib.default_context(0)
    p = S.make_pair(4, 300, seed=1)
    g = _graph(ctx, p)
    before = g.cost()
    with pytest.raises(_lib.PgcError, match="static scratch"):
        g.solve_ratio_noalloc(3)
    assert np.array_equal(g.cost(), before)
    g.close()
