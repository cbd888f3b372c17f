"""The workload configurations at their full sizes, on the GPU.

C2: the conjugate product graph with normals (icosphere s=4 x 100-point
    curve, max_depth 2, regularising term, vanilla/nomode; 41.4 M edges):
    rows bit-exact with the oracle's product space, energies within 1e-12
    relative of the oracle's (the rotation term's acos/pow may differ in the
    last bit), and the lawlergpu solve bit-exact with the oracle's solve on
    the device's own energies (rows, fp32 objective, counters).
C4: 256 uneven pairs x (1 + 8 perturbations) = 2304 batched lawlergpu
    solves. Every item: status PGC_OK, a closed walk that passes every layer
    once (from its product-space rows), and the same cycle, objective and
    counters as pgc_solve_ratio on the same item. A seeded subset of 72 items
    (subdivision 4, max_depth 3 and odd layer counts among them) against the
    oracle, under the reference's fold and the corrected one.
C3: 16 pairs (icosphere s=3 x 64-point curve, max_depth 2): lambda* and
    d lambda*/d feat_diff from the autograd Function; FP64 central finite
    differences (eps 1e-4) on 200 entries per pair, 100 on the cycle and 100
    off it, compared wherever both perturbed solves keep the cycle (lambda*
    is then linear in feat_diff, so they agree to 1e-8); the backward output
    against the gradient assembled from the oracle's cycle; the smoothed
    estimator (32 samples, sigma 0.05) against the same samples solved in
    one batch.
Tolerances: bit-exact for ids, rows, counters and quantized costs; the fp32
objective equal (1e-5 relative written, observed 0).
"""
import numpy as np
import pytest
import torch

import pgc_oracle as O
from cyclic_product_graphs import _lib
from cyclic_product_graphs import autograd as A
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5


def _same_obj(a, b):
    assert abs(a - float(b)) <= OBJ_RTOL * abs(float(b))
    assert a == float(b)


# ------------------------------------------------------------------ C2
def test_c2_conjugate_full_size():
    ctx = _lib.default_context(0)
    p = S.make_pair(4, 100, seed=0, with_normals=True)
    NX, NY = p["NX"].astype(np.float64), p["NY"].astype(np.float64)
    print("C2: generate on the device", flush=True)
    g = _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), True, True,
                            False, "vanilla", "nomode", NX=NX, NY=NY)
    info = g.info()
    assert info["n_edges"] == 41442000 and info["nodes_per_layer"] == 46080
    rows_d = g.product_space()
    en_d = g.cost()
    print("C2: oracle product space and energy", flush=True)
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    assert np.array_equal(rows_d, rows), "conjugate product space differs"
    en = O.conj_energy(p["VX"], NX, p["VY"], NY, p["EX"], p["EY"], rows, src, trgt, p["feat_diff"], reg=True,
                       cost_mode="vanilla", time_mode="nomode")
    np.testing.assert_allclose(en_d, en, rtol=1e-12, atol=0)
    del rows_d, en
    print("C2: lawlergpu solve", flush=True)
    eids, rws, res = g.solve_ratio(2, _lib.PGC_LAWLER_PICK_LAST)
    qd = g.cost()
    print("C2: oracle solve on the device energies", flush=True)
    e, oobj, passes, stats, q = O.ctr_solve(src, trgt, en_d, info["nodes_per_layer"], 2, gpu_pick=True)
    assert np.array_equal(qd, q), "quantized cost differs"
    assert np.array_equal(eids, e) and np.array_equal(rws, rows[e])
    _same_obj(res.objective, oobj)
    assert (res.bf_calls, res.sweeps, res.lawler_iters, res.branches) == tuple(int(x) for x in stats[:4])
    print("C2 solve: %d sweeps, %.1f ms solver kernels" % (res.sweeps, res.kernel_ms), flush=True)
    g.close()


# ------------------------------------------------------------------ C4
def _c4_items():
    """The serving-scale batch: 256 uneven pairs, each with its unperturbed
    feat_diff and 8 Gaussian perturbations (sigma 0.05, fp32)."""
    pairs = S.make_uneven_pairs(256, seed=0)
    gen = torch.Generator(device="cuda:0").manual_seed(1)
    items = []
    for i, (p, K) in enumerate(pairs):
        f = torch.tensor(p["feat_diff"], dtype=torch.float32, device="cuda:0")
        items.append((i, f.double().cpu().numpy()))
        for _ in range(8):
            items.append((i, (f + 0.05 * torch.randn(f.shape, generator=gen, device="cuda:0")).double().cpu().numpy()))
    return pairs, items


def _walk_props(rows, L):
    # product-space rows (y0, y1, x0, x1): an edge goes from (y0, x0) to (y1, x1)
    src = rows[:, [0, 2]]
    dst = rows[:, [1, 3]]
    assert np.array_equal(dst[:-1], src[1:]), "not a walk"
    assert np.array_equal(dst[-1], src[0]), "not closed"
    inter = rows[rows[:, 0] != rows[:, 1]]
    assert len(inter) == L, "does not pass every layer once"
    assert len({(int(a), int(b)) for a, b in inter[:, :2]}) == L


def test_c4_full_batch():
    ctx = _lib.default_context(0)
    pairs, items = _c4_items()
    bases = [_lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], np.zeros((len(p["VX"]), len(p["VY"]))),
                                 False, False, False, "vanilla", "") for p, _ in pairs]
    graphs = [bases[i].derive(fd) for i, fd in items]
    Ks = [pairs[i][1] for i, _ in items]
    print("C4: batch of %d items" % len(graphs), flush=True)
    outs = _lib.solve_ratio_batch(ctx, graphs, Ks)
    for (i, _), (eids, rws, res) in zip(items, outs):
        assert res.status == 0 and res.passes_layers_once == 1
        _walk_props(rws, len(pairs[i][0]["EY"]))
    print("C4: plain per-item loop", flush=True)
    for k, ((i, fd), (eids, rws, res)) in enumerate(zip(items, outs)):
        g1 = bases[i].derive(fd)
        e1, r1, res1 = g1.solve_ratio(Ks[k])
        assert np.array_equal(eids, e1) and np.array_equal(rws, r1)
        assert res.objective == res1.objective
        assert (res.sweeps, res.bf_calls, res.lawler_iters, res.branches) == \
            (res1.sweeps, res1.bf_calls, res1.lawler_iters, res1.branches)
        g1.close()
    for g in graphs:
        g.close()
    # seeded oracle subset: 24 items of subdivision-4 pairs, 16 of max_depth 3,
    # 16 of odd L with max_depth >= 2, 16 more at random
    rng = np.random.default_rng(7)
    nx4 = [k for k, (i, _) in enumerate(items) if len(pairs[i][0]["VX"]) == 2562]
    k3 = [k for k, (i, _) in enumerate(items) if pairs[i][1] == 3]
    odd = [k for k, (i, _) in enumerate(items) if len(pairs[i][0]["EY"]) % 2 == 1 and pairs[i][1] >= 2]
    pick = list(rng.choice(nx4, 24, replace=False)) + list(rng.choice(k3, 16, replace=False)) + \
        list(rng.choice(odd, 16, replace=False)) + list(rng.choice(len(items), 16, replace=False))
    pick = sorted(set(int(k) for k in pick))
    assert len(pick) >= 64
    for corrected in (False, True):
        flags = _lib.PGC_LAWLER_PICK_LAST | (_lib.PGC_CORRECTED_ODD_L if corrected else 0)
        sub = [bases[items[k][0]].derive(items[k][1]) for k in pick]
        print("C4: oracle subset of %d items, corrected=%s" % (len(sub), corrected), flush=True)
        souts = _lib.solve_ratio_batch(ctx, sub, [Ks[k] for k in pick], flags)
        for k, (eids, rws, res) in zip(pick, souts):
            i, fd = items[k]
            p, K = pairs[i]
            rows, src, trgt = O.product_space(p["EX"], p["EY"])
            en = O.energy(p["VX"], p["VY"], rows, fd)
            e, oobj, _, stats, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True, corrected=corrected)
            assert np.array_equal(eids, e) and np.array_equal(rws, rows[e]), (k, corrected)
            _same_obj(res.objective, oobj)
            assert (res.bf_calls, res.sweeps, res.lawler_iters, res.branches) == tuple(int(x) for x in stats[:4])
        for g in sub:
            g.close()
    for g in bases:
        g.close()


# ------------------------------------------------------------------ C3
EPS = 1e-4
FD_ATOL = 1e-8


def test_c3_sixteen_pairs_finite_differences_and_smoothing():
    pairs = [S.make_pair(3, 64, seed=s) for s in range(16)]
    K = 2
    prob = A.RatioProblem([(p["VX"], p["EX"], p["VY"], p["EY"]) for p in pairs], K)
    fds = [torch.tensor(p["feat_diff"].astype(np.float64), device="cuda", requires_grad=True) for p in pairs]
    lam = A.cycle_ratio(prob, fds)
    lam.sum().backward()
    cycles = [c.copy() for c in prob.last_cycles]
    rng = np.random.default_rng(3)
    full_pairs, used_total, tried_total = 0, 0, 0
    for i, p in enumerate(pairs):
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
        e, _, _, _, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
        assert np.array_equal(cycles[i], e)
        s = 0.0
        for x in en[e, 0]:
            s += float(x)
        assert float(lam[i].detach()) == s / len(e)
        grad = fds[i].grad.cpu().numpy()
        # the gradient assembled from the lawlercpu cycle: 0.5 / T at both
        # endpoints of every cycle edge (lawlergpu's cycle where they differ)
        ec = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=False)[0]
        ref_cycle = ec if np.array_equal(ec, e) else e
        gref = np.zeros(p["feat_diff"].shape)
        for y0, y1, x0, x1 in rows[ref_cycle]:
            gref[x0, y0] += 0.5 / len(ref_cycle)
            gref[x1, y1] += 0.5 / len(ref_cycle)
        np.testing.assert_allclose(grad, gref, rtol=1e-12, atol=0)
        # 100 entries on the cycle (with repetition when it touches fewer) and
        # 100 off it
        on_set = sorted({(int(r[2]), int(r[0])) for r in rows[e]} | {(int(r[3]), int(r[1])) for r in rows[e]})
        on = [on_set[j] for j in rng.choice(len(on_set), size=100, replace=len(on_set) < 100)]
        nx, ny = p["feat_diff"].shape
        off = []
        while len(off) < 100:
            x, y = int(rng.integers(nx)), int(rng.integers(ny))
            if (x, y) not in on_set:
                off.append((x, y))
        entries = on + off
        base = fds[i].detach()
        sub = A.RatioProblem([(p["VX"], p["EX"], p["VY"], p["EY"])], K)
        pert = []
        for (x, y) in entries:
            for sgn in (1.0, -1.0):
                f = base.clone()
                f[x, y] += sgn * EPS
                pert.append(f)
        with torch.no_grad():
            lams = A.cycle_ratio(sub, pert, [0] * len(pert)).cpu().numpy()
        pc = sub.last_cycles
        used = 0
        for k, (x, y) in enumerate(entries):
            if not (np.array_equal(pc[2 * k], e) and np.array_equal(pc[2 * k + 1], e)):
                continue
            used += 1
            fdiff = (lams[2 * k] - lams[2 * k + 1]) / (2 * EPS)
            assert abs(fdiff - grad[x, y]) <= FD_ATOL, (i, x, y, fdiff, grad[x, y])
        sub.close()
        print("pair %d: finite differences compared on %d of %d entries" % (i, used, len(entries)), flush=True)
        full_pairs += used == len(entries)
        used_total += used
        tried_total += len(entries)
    # every pair keeps its cycle under all 400 perturbed solves: the 16
    # optima are separated from the next cycle by more than a 1e-4 change of
    # one feat_diff entry moves them
    assert full_pairs == 16 and used_total == tried_total
    # smoothed estimator: 32 samples per pair, all 16 pairs in one batch,
    # against the same samples solved as one plain batch
    g = torch.Generator().manual_seed(5)
    f2 = [f.detach().clone().requires_grad_(True) for f in fds]
    sm = A.smoothed_cycle_ratio(prob, f2, sigma=0.05, samples=32, generator=g)
    sm.sum().backward()
    g = torch.Generator().manual_seed(5)
    xs, owner = [], []
    for i, f in enumerate(f2):
        for _ in range(32):
            z = torch.randn(f.shape, generator=g, dtype=f.dtype).to(f.device)
            xs.append((f.detach() + 0.05 * z).requires_grad_(True))
            owner.append(i)
    v = A.cycle_ratio(prob, xs, owner)
    v.sum().backward()
    vals = v.detach().cpu().numpy()
    for i in range(16):
        assert abs(float(sm[i]) - vals[32 * i:32 * i + 32].mean()) <= 1e-12 * abs(float(sm[i]))
        ref = torch.stack([xs[32 * i + s].grad for s in range(32)]).mean(0)
        torch.testing.assert_close(f2[i].grad, ref / 1.0, rtol=1e-12, atol=1e-15)
    prob.close()
