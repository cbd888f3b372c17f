"""The ratio branch and bound on the GPU, on instances that branch.

tests/branching_cases.py holds 29 seeded tiny rings on which the Lawler solve
splits branches (pinned against brute force in tests/test_oracle.py). Each is
solved through the C-ABI by pgc_solve_ratio, whose branch and bound runs on the
device (one launch) or, with PGC_HOST_BNB=1, on the host around the device
Lawler launches, and, with the device store of valid-branch cycles capped
(PGC_VALID_STORE), on the device until the store overflows and then on the
host; and by pgc_solve_ratio_batch (device branch and bound in the persistent
launch), under both candidate pick rules, and compared with the oracle. Bar: edge ids and product-space rows
bit-exact, the fp32 objective equal (1e-5 relative written as the tolerance,
observed difference 0), and the Bellman-Ford call, sweep, Lawler iteration and
branch counters equal.
"""
import os

import numpy as np
import pytest

import pgc_oracle as O
from branching_cases import BRANCHING
from cyclic_product_graphs import _lib
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5
RULES = {"lawlergpu": _lib.PGC_LAWLER_PICK_LAST, "lawlercpu": 0}


def _pair(seed, nx, ny):
    return S.tiny_pair(nx, ny, seed=seed, ring=True)


def _graph(ctx, p):
    return _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                               False, False, "vanilla", "")


def _oracle(p, K, gpu_pick):
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
    e, obj, passes, stats, q = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=gpu_pick)
    return rows, e, obj, stats


def _same(eids, rows, res, orow, oe, oobj, stats):
    assert np.array_equal(eids, oe), "cycle differs from the oracle"
    assert np.array_equal(rows, orow[oe])
    assert abs(res.objective - float(oobj)) <= OBJ_RTOL * abs(float(oobj))
    assert res.objective == float(oobj)
    assert (res.bf_calls, res.sweeps, res.lawler_iters, res.branches) == \
        (stats[0], stats[1], stats[2], stats[3])


@pytest.mark.parametrize("driver", ["device", "host", "device-store-0", "device-store-1"])
@pytest.mark.parametrize("name", list(RULES))
def test_single_solves_branching(name, driver, monkeypatch):
    if driver == "host":
        monkeypatch.setenv("PGC_HOST_BNB", "1")
    elif driver.startswith("device-store-"):
        monkeypatch.setenv("PGC_VALID_STORE", driver[-1])
    ctx = _lib.default_context(0)
    branched = 0
    for seed, nx, ny, K in BRANCHING:
        print("case", name, driver, seed, nx, ny, K, flush=True)
        p = _pair(seed, nx, ny)
        orow, oe, oobj, stats = _oracle(p, K, RULES[name] != 0)
        g = _graph(ctx, p)
        eids, rows, res = g.solve_ratio(K, RULES[name])
        _same(eids, rows, res, orow, oe, oobj, stats)
        branched += res.branches > 1
        g.close()
    assert branched >= 20


@pytest.mark.parametrize("store", [None, "0", "1"])
@pytest.mark.parametrize("name", list(RULES))
def test_batch_solves_branching(name, store, monkeypatch):
    # store: PGC_VALID_STORE caps the device store of valid-branch cycles; an
    # item that outgrows it is solved again by the host branch and bound
    # (same quantized costs), so the results must not change
    if store is not None:
        monkeypatch.setenv("PGC_VALID_STORE", store)
    ctx = _lib.default_context(0)
    pairs = [(_pair(seed, nx, ny), K) for seed, nx, ny, K in BRANCHING]
    graphs = [_graph(ctx, p) for p, _ in pairs]
    print("batch", name, "store", store, len(graphs), "graphs", flush=True)
    outs = _lib.solve_ratio_batch(ctx, graphs, [K for _, K in pairs], flags=RULES[name])
    branched = 0
    for (p, K), (eids, rows, res) in zip(pairs, outs):
        orow, oe, oobj, stats = _oracle(p, K, RULES[name] != 0)
        assert res.status == 0
        _same(eids, rows, res, orow, oe, oobj, stats)
        branched += res.branches > 1
    assert branched >= 20
    for g in graphs:
        g.close()


def test_batch_duplicate_beyond_64_refused():
    ctx = _lib.default_context(0)
    p = S.tiny_pair(4, 4, seed=0, ring=True)
    graphs = [_graph(ctx, p) for _ in range(70)]
    with pytest.raises(_lib.PgcError, match="twice"):
        _lib.solve_ratio_batch(ctx, graphs + [graphs[66]], 2)
    # nothing was quantized: the energies are still the raw ones
    rows, _, _ = O.product_space(p["EX"], p["EY"])
    en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
    assert np.array_equal(graphs[66].cost(), en)
    for g in graphs:
        g.close()


def test_batch_per_item_status():
    # one item's output capacity is too small: that item fails with its own
    # status, the others are solved and written as usual
    ctx = _lib.default_context(0)
    cases = [(S.tiny_pair(5, 6, seed=s, ring=True), 2) for s in range(4)]
    graphs = [_graph(ctx, p) for p, _ in cases]
    caps = [64, 3, 64, 64]
    with pytest.raises(_lib.PgcError):
        _lib.solve_ratio_batch(ctx, graphs, 2, capacity=caps)
    graphs = [_graph(ctx, p) for p, _ in cases]
    outs = _lib.solve_ratio_batch(ctx, graphs, 2, capacity=caps, raise_on_error=False)
    for i, ((p, K), (eids, rows, res)) in enumerate(zip(cases, outs)):
        if i == 1:
            assert res.status == -4  # PGC_ERR_CAPACITY
            continue
        assert res.status == 0
        orow, oe, oobj, stats = _oracle(p, K, True)
        _same(eids, rows, res, orow, oe, oobj, stats)
