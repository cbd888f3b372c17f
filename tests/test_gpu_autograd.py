"""Differentiable lambda* (cyclic_product_graphs.autograd) on the GPU.

lambda*(FD) = sum_{e in C*} c_e / sum_{e in C*} t_e in raw energy units, C*
the cycle lawlergpu returns. Checks:
  * the forward value equals the mean raw energy of the oracle's cycle
    (sequential fp64 sum in walk order: exact);
  * the backward kernel equals the gradient assembled on the CPU from the
    lawlercpu (oracle) cycle (rtol 1e-12: fp64 atomics may add in any order);
  * FP64 central finite differences (eps = 1e-4) on on-cycle and off-cycle
    entries agree with the backward output to 1e-8 absolute wherever both
    perturbed solves return the unperturbed cycle (lambda* is then linear);
  * the perturbation-smoothed estimator is the mean over its samples.
"""
import numpy as np
import pytest
import torch

import pgc_oracle as O
from cyclic_product_graphs import autograd as A
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

EPS = 1e-4
FD_ATOL = 1e-8


def _problem(pairs, K):
    return A.RatioProblem([(p["VX"], p["EX"], p["VY"], p["EY"]) for p in pairs], K)


def _cpu_cycle(p, K, gpu_pick):
    rows, src, trgt = O.product_space(p["EX"], p["EY"])
    en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
    e, obj, _, _, _ = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=gpu_pick)
    return rows, en[:, 0], e


def test_forward_and_backward_match_cpu_cycles():
    pairs = [S.make_pair(3, 64, seed=s) for s in range(3)]
    K = 2
    prob = _problem(pairs, K)
    fds = [torch.tensor(p["feat_diff"], dtype=torch.float32, device="cuda", requires_grad=True) for p in pairs]
    lam = A.cycle_ratio(prob, fds)
    lam.sum().backward()
    for i, p in enumerate(pairs):
        rows, en, e = _cpu_cycle(p, K, gpu_pick=True)
        s = 0.0
        for x in en[e]:
            s += float(x)
        assert float(lam[i].detach()) == s / len(e)
        # gradient from the lawlercpu cycle: 0.5 / T at both endpoints of every edge
        _, _, ec = _cpu_cycle(p, K, gpu_pick=False)
        g = np.zeros(p["feat_diff"].shape)
        for y0, y1, x0, x1 in rows[ec]:
            g[x0, y0] += 0.5 / len(ec)
            g[x1, y1] += 0.5 / len(ec)
        got = fds[i].grad.double().cpu().numpy()
        np.testing.assert_allclose(got, g.astype(np.float32).astype(np.float64), rtol=1e-6, atol=0)
    prob.close()


def test_backward_fp64_exact_against_assembled_gradient():
    p = S.make_pair(2, 40, seed=5)
    prob = _problem([p], 2)
    fd = torch.tensor(p["feat_diff"].astype(np.float64), device="cuda", requires_grad=True)
    lam = A.cycle_ratio(prob, [fd])
    (3.0 * lam).sum().backward()
    rows, en, e = _cpu_cycle(p, 2, gpu_pick=True)
    g = np.zeros(p["feat_diff"].shape)
    for y0, y1, x0, x1 in rows[e]:
        g[x0, y0] += 3.0 * 0.5 / len(e)
        g[x1, y1] += 3.0 * 0.5 / len(e)
    np.testing.assert_allclose(fd.grad.cpu().numpy(), g, rtol=1e-12, atol=0)
    prob.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_central_finite_differences(seed):
    p = S.make_pair(3, 64, seed=seed)
    K = 2
    prob = _problem([p], K)
    fd0 = torch.tensor(p["feat_diff"].astype(np.float64), device="cuda", requires_grad=True)
    lam = A.cycle_ratio(prob, [fd0])
    lam.sum().backward()
    grad = fd0.grad.cpu().numpy()
    cyc = prob.last_cycles[0]
    rows, _, _ = O.product_space(p["EX"], p["EY"])
    on = sorted({(int(r[2]), int(r[0])) for r in rows[cyc]} | {(int(r[3]), int(r[1])) for r in rows[cyc]})
    rng = np.random.default_rng(seed)
    on = [on[i] for i in rng.choice(len(on), size=min(10, len(on)), replace=False)]
    nx, ny = p["feat_diff"].shape
    off = []
    while len(off) < 10:
        x, y = int(rng.integers(nx)), int(rng.integers(ny))
        if (x, y) not in on and grad[x, y] == 0:
            off.append((x, y))
    entries = on + off
    base = fd0.detach()
    fds = []
    for (x, y) in entries:
        for sgn in (1.0, -1.0):
            f = base.clone()
            f[x, y] += sgn * EPS
            fds.append(f)
    with torch.no_grad():
        lams = A.cycle_ratio(prob, fds, [0] * len(fds)).cpu().numpy()
    cycles = prob.last_cycles
    used = 0
    for k, (x, y) in enumerate(entries):
        if not (np.array_equal(cycles[2 * k], cyc) and np.array_equal(cycles[2 * k + 1], cyc)):
            continue  # the perturbation changed the optimal cycle: lambda* is not locally linear here
        used += 1
        fdiff = (lams[2 * k] - lams[2 * k + 1]) / (2 * EPS)
        assert abs(fdiff - grad[x, y]) <= FD_ATOL, (x, y, fdiff, grad[x, y])
    print("finite-difference entries used:", used, "of", len(entries), flush=True)
    assert used >= len(entries) * 0.8
    prob.close()


def test_sixteen_pair_batch_and_smoothed_gradient():
    pairs = [S.make_pair(3, 64, seed=100 + s) for s in range(16)]
    prob = _problem(pairs, 2)
    fds = [torch.tensor(p["feat_diff"], dtype=torch.float32, device="cuda", requires_grad=True) for p in pairs]
    lam = A.cycle_ratio(prob, fds)
    lam.sum().backward()
    for i, p in enumerate(pairs):
        rows, en, e = _cpu_cycle(p, 2, gpu_pick=True)
        assert np.array_equal(prob.last_cycles[i], e)
        assert fds[i].grad is not None and torch.isfinite(fds[i].grad).all()
    # smoothed estimator over 32 perturbations of two pairs: its value and
    # gradient are the means over the same samples solved one by one
    g = torch.Generator().manual_seed(3)
    f2 = [fds[0].detach().clone().requires_grad_(True), fds[1].detach().clone().requires_grad_(True)]
    sm = A.smoothed_cycle_ratio(prob, f2, sigma=0.05, samples=32, generator=g, pair_index=[0, 1])
    sm.sum().backward()
    g = torch.Generator().manual_seed(3)
    vals, grads = [], []
    for i in range(2):
        for _ in range(32):
            z = torch.randn(f2[i].shape, generator=g, dtype=f2[i].dtype).to(f2[i].device)
            x = (f2[i].detach() + 0.05 * z).requires_grad_(True)
            v = A.cycle_ratio(prob, [x], [i])
            v.sum().backward()
            vals.append(float(v[0]))
            grads.append(x.grad.double())
    for i in range(2):
        assert abs(float(sm[i]) - np.mean(vals[32 * i:32 * i + 32])) <= 1e-12 * abs(float(sm[i]))
        ref = torch.stack(grads[32 * i:32 * i + 32]).mean(0)
        torch.testing.assert_close(f2[i].grad.double(), ref, rtol=1e-5, atol=1e-9)
    prob.close()


@pytest.mark.parametrize("reg,prune,cost,time", [(False, False, "vanilla", ""), (True, False, "vanilla", ""),
                                                 (False, True, "vanilla", ""), (True, True, "vanilla", ""),
                                                 (False, False, "feature", "lengthnormalisation")])
def test_batched_energies_equal_single_graph_energies(reg, prune, cost, time):
    # pgc_graphs_set_feat_diff decodes each edge's row from its id (X and Y
    # edge lists) instead of reading the product space; its energies must be
    # bit-exact with the single-graph energy (which reads the rows), for
    # float64 and float32 feat_diff
    from cyclic_product_graphs import _lib
    ctx = _lib.default_context(0)
    dev = torch.device("cuda:0")
    pairs = [S.make_pair(2, 33, seed=3), S.make_pair(3, 40, seed=4)]
    gen = torch.Generator().manual_seed(5)
    for p in pairs:
        print("pair nx", len(p["VX"]), "ny", len(p["VY"]), "reg", reg, "prune", prune, cost, time, flush=True)
        base = _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64), False,
                                   reg, prune, cost, time)
        fds = [torch.rand(p["feat_diff"].shape, generator=gen, dtype=torch.float64) * 3.0 for _ in range(3)]
        single = []
        for f in fds:
            g = base.derive(f.numpy())
            single.append(g.cost())
            g.close()
        items = [base.derive(p["feat_diff"].astype(np.float64)) for _ in fds]
        d64 = [f.to(dev).contiguous() for f in fds]
        _lib.set_feat_diff_batch(items, d64, 0)
        torch.cuda.synchronize()
        for g, want in zip(items, single):
            assert np.array_equal(g.cost(), want), "float64 feat_diff: batched energy differs"
        d32 = [f.to(torch.float32).to(dev).contiguous() for f in fds]
        _lib.set_feat_diff_batch(items, d32, 1)
        torch.cuda.synchronize()
        for g, f in zip(items, fds):
            h = base.derive(f.to(torch.float32).double().numpy())
            assert np.array_equal(g.cost(), h.cost()), "float32 feat_diff: batched energy differs"
            h.close()
        for g in items:
            g.close()
        base.close()
