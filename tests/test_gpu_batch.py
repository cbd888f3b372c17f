"""Batched complete solves (pgc_solve_ratio_batch: device-resident branch and
bound, persistent launches) against the single-graph solve and the oracle.

Bar: for every graph of a batch, the cycle (edge ids, walk order) and its
product-space rows are bit-exact with the oracle's and with pgc_solve_ratio's;
the fp32 objective is equal (1e-5 relative written as the tolerance, observed
difference 0); the sweep, Bellman-Ford call and branch counters are equal.
"""
import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5


def _graph(ctx, p, reg=False, cost="vanilla", time=""):
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                               reg, False, cost, time)


def _oracle(p, K, reg=False, cost="vanilla", time=""):
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64), reg=reg, cost_mode=cost,
                  time_mode=time)
    e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
    return rows, e, obj, stats


def _cases():
    out = []
    for s, (nx, ny, ring) in enumerate([(4, 4, True), (5, 6, True), (6, 5, False), (4, 7, False)]):
        for K in (1, 2, 3):
            out.append((S.tiny_pair(nx, ny, seed=s, ring=ring), K, {}))
    for sub, ny, K, seed in [(1, 8, 2, 0), (2, 33, 2, 1), (2, 41, 3, 4), (3, 64, 2, 0), (2, 20, 1, 3)]:
        out.append((S.make_pair(sub, ny, seed=seed), K, {}))
    p = S.make_pair(2, 24, seed=7)
    out.append((p, 2, {"reg": True}))
    out.append((p, 2, {"cost": "feature", "time": "lengthnormalisation"}))
    return out


def test_batch_matches_single_and_oracle():
    ctx = _lib.default_context(0)
    cases = _cases()
    print("batch of", len(cases), "graphs", flush=True)
    graphs = [_graph(ctx, p, **kw) for p, K, kw in cases]
    outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K, _ in cases])
    for (p, K, kw), (eids, rows, res) in zip(cases, outs):
        orow, oe, oobj, stats = _oracle(p, K, **kw)
        assert np.array_equal(rows, orow[oe]), "matching differs from the oracle"
        assert abs(res.objective - float(oobj)) <= OBJ_RTOL * abs(float(oobj))
        assert res.objective == float(oobj)
        assert (res.bf_calls, res.sweeps, res.branches) == (stats[0], stats[1], stats[3])
        g1 = _graph(ctx, p, **kw)
        e1, r1, res1 = g1.solve_ratio(K)
        assert np.array_equal(eids, e1) and np.array_equal(rows, r1)
        assert res.objective == res1.objective
        assert (res.sweeps, res.bf_calls, res.lawler_iters, res.branches) == \
            (res1.sweeps, res1.bf_calls, res1.lawler_iters, res1.branches)
        g1.close()
    for g in graphs:
        g.close()


def test_batch_repeated_requantizes_like_single():
    # a second batch over the same graphs re-quantizes their quantized costs
    # in place, as a second pgc_solve_ratio does
    ctx = _lib.default_context(0)
    p = S.make_pair(2, 16, seed=9)
    gb = _graph(ctx, p)
    gs = _graph(ctx, p)
    for _ in range(2):
        (eb, rb, resb), = _lib.solve_ratio_batch(ctx, [gb], 2)
        es, rs, ress = gs.solve_ratio(2)
        assert np.array_equal(eb, es) and resb.objective == ress.objective
        assert np.array_equal(gb.cost(), gs.cost())


def test_batch_c1_full_size_with_duplicated_pairs():
    # eight independent copies of the full-size pair (icosphere s=4 x 128, K=3)
    ctx = _lib.default_context(0)
    p = S.make_pair(4, 128, seed=0)
    graphs = [_graph(ctx, p) for _ in range(8)]
    outs = _lib.solve_ratio_batch(ctx, graphs, 3)
    orow, oe, oobj, stats = _oracle(p, 3)
    for eids, rows, res in outs:
        assert np.array_equal(rows, orow[oe])
        assert res.objective == float(oobj)
        assert res.sweeps == stats[1]
    for g in graphs:
        g.close()


def test_batch_argument_errors():
    ctx = _lib.default_context(0)
    p = S.tiny_pair(4, 4, seed=0, ring=True)
    g = _graph(ctx, p)
    with pytest.raises(_lib.PgcError):
        _lib.solve_ratio_batch(ctx, [g, g], 2)
    with pytest.raises(_lib.PgcError):
        _lib.solve_ratio_batch(ctx, [g], 0)
    g.close()


def test_sweep_cap_case_golden():
    # a perturbed pair whose Bellman-Ford calls run to the numNodes-1 sweep
    # cap (26363 sweeps): single and batched solves against the oracle's
    # answer stored in tests/golden/bf_cap_case.npz
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "bf_cap_case.npz"))
    p, K = S.make_uneven_pairs(256, seed=0)[int(z["pair_index"])]
    assert K == int(z["max_depth"])
    p = dict(p, feat_diff=z["feat_diff"])
    ctx = _lib.default_context(0)
    rows, _, _ = O.product_space(p["EX"], p["EY"])
    g1 = _graph(ctx, p)
    e1, r1, res1 = g1.solve_ratio(K)
    gb = _graph(ctx, p)
    (eb, rb, resb), = _lib.solve_ratio_batch(ctx, [gb], K)
    for e, r, res in ((e1, r1, res1), (eb, rb, resb)):
        assert np.array_equal(r, rows[z["edges"]])
        assert res.objective == float(z["objective"])
        assert [res.bf_calls, res.sweeps, res.lawler_iters, res.branches] == list(z["stats"])
    g1.close()
    gb.close()


def test_batch_heavy_item_first_phase():
    # A batch of graphs that were all solved before, one of which ran far more
    # sweeps (the sweep-cap pair): that item is prepared and launched first and
    # the others are prepared while it runs. Results equal those of single
    # solves on copies of the graphs (both paths quantize the quantized costs
    # again on the second solve, as the reference's in-place quantization does).
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "bf_cap_case.npz"))
    pc, Kc = S.make_uneven_pairs(256, seed=0)[int(z["pair_index"])]
    pc = dict(pc, feat_diff=z["feat_diff"])
    items = [(pc, Kc, {})] + [c for c in _cases() if c[2] == {}][:17]
    assert len(items) >= 16
    ctx = _lib.default_context(0)
    gb = [_graph(ctx, p, **kw) for p, K, kw in items]
    gs = [_graph(ctx, p, **kw) for p, K, kw in items]
    Ks = [K for _, K, _ in items]
    for rnd in range(2):
        print("heavy-first batch round", rnd, flush=True)
        outs = _lib.solve_ratio_batch(ctx, gb, Ks)
        for (eb, rb, resb), g, K in zip(outs, gs, Ks):
            es, rs, ress = g.solve_ratio(K)
            assert np.array_equal(eb, es) and np.array_equal(rb, rs)
            assert resb.objective == ress.objective
            assert (resb.sweeps, resb.bf_calls, resb.lawler_iters, resb.branches) == \
                (ress.sweeps, ress.bf_calls, ress.lawler_iters, ress.branches)
        if rnd == 0:
            assert outs[0][2].sweeps > 8 * max(o[2].sweeps for o in outs[1:])
    for g in gb + gs:
        g.close()


def test_batch_median_bracket_miss_and_overflow():
    # the batched median samples 8192 strided keys per column and brackets the
    # median's expected position; these columns defeat the sample (every
    # sampled key is the smallest) or overflow the bracket (all keys equal),
    # so they take the full radix select; the quantized costs must still equal
    # the oracle's quantization (numpy-pinned median and rounding)
    ctx = _lib.default_context(0)
    p = S.make_pair(2, 33, seed=5)
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    npl = len(p["VX"])
    E = len(src)
    assert E > 8192
    rng = np.random.default_rng(3)
    sampled = (np.arange(8192, dtype=np.int64) * E) // 8192
    miss = 2.0 + np.arange(E) * 1e-6
    miss[sampled] = 1.0
    flat = np.full(E, 3.0)
    flat[rng.integers(0, E, 50)] = 7.0
    normal = rng.uniform(0.1, 20.0, E)
    cols = {"miss": miss, "overflow": flat, "normal": normal}
    graphs = [_lib.Graph.from_arrays(ctx, rows, src, trgt, c, npl) for c in cols.values()]
    # one more column of odd length (the two median ranks coincide)
    graphs.append(_lib.Graph.from_arrays(ctx, rows[:-1], src[:-1], trgt[:-1], normal[:-1], npl))
    print("batch of", len(graphs), "array graphs, E", E, flush=True)
    _lib.solve_ratio_batch(ctx, graphs, 1)
    expect = [O.quantize(c[:, None])[:, 0] for c in cols.values()] + [O.quantize(normal[:-1, None])[:, 0]]
    for name, g, q in zip(list(cols) + ["odd"], graphs, expect):
        assert np.array_equal(g.cost()[:, 0], q), "quantized costs differ (%s column)" % name
    for g in graphs:
        g.close()


def test_batch_corrected_odd_L_matches_single_and_oracle():
    # PGC_CORRECTED_ODD_L through the batch path: odd layer counts with K >= 2
    # (where the corrected fold differs from the reference's), against
    # pgc_solve_ratio with the same flag and the oracle's corrected solver
    ctx = _lib.default_context(0)
    flags = _lib.PGC_LAWLER_PICK_LAST | _lib.PGC_CORRECTED_ODD_L
    cases = [(S.tiny_pair(nx, ny, seed=s + 60, ring=ring), K)
             for s, (nx, ny, ring) in enumerate([(4, 5, True), (5, 7, True), (4, 5, False), (6, 3, True)])
             for K in (2, 3)]
    cases += [(S.make_pair(2, 33, seed=1), 2), (S.make_pair(2, 41, seed=4), 3)]
    print("batch of", len(cases), "odd-L graphs, corrected", flush=True)
    graphs = [_graph(ctx, p) for p, K in cases]
    outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K in cases], flags)
    for (p, K), (eids, rows, res) in zip(cases, outs):
        orow, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], orow, p["feat_diff"].astype(np.float64))
        oe, oobj, _, stats, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True, corrected=True)
        assert np.array_equal(rows, orow[oe]), "matching differs from the oracle"
        assert res.objective == float(oobj)
        assert (res.bf_calls, res.sweeps, res.branches) == (stats[0], stats[1], stats[3])
        g1 = _graph(ctx, p)
        e1, r1, res1 = g1.solve_ratio(K, flags)
        assert np.array_equal(eids, e1) and res.objective == res1.objective
        g1.close()
    for g in graphs:
        g.close()
