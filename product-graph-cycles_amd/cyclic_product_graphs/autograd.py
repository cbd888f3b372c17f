"""Differentiable minimum cost/time-ratio cycles (PyTorch).

`RatioProblem` holds the product graphs of a set of shape pairs on the
device; `cycle_ratio(problem, fds, pair_index)` solves every item (one
feat_diff per item, several items may share a pair, e.g. perturbation
samples) in one batched lawlergpu call and returns lambda* per item in raw
energy units:

    lambda*(FD) = sum_{e in C*} c_e(FD) / sum_{e in C*} t_e,

with C* the cycle the solver returns for the quantized problem, c_e the raw
edge energy and t_e the time column (1 per edge without one). lambda* is
piecewise linear in FD (C* is locally constant), and the backward pass is the
hand-written scatter kernel of libpgc: every cycle edge adds dc_e/dFD / T(C*)
at its two (x, y) entries of dL/dfeat_diff. The solver's discrete arithmetic
is unchanged. `smoothed_cycle_ratio` averages lambda* and its gradient over
Gaussian perturbations of FD (a perturbation-smoothed estimator).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib


class RatioProblem:
    """Product graphs of shape pairs (VX, EX, VY, EY) on one device.

    Items are created on first use: item i uses pair pair_index[i] with its
    own feat_diff, energy, quantized costs and solver scratch (a graph derived
    from the pair's graph), so all items of a call are solved concurrently.
    """

    def __init__(self, pairs, max_depth, device=0, regularising=False, cost_mode="vanilla", time_mode="",
                 flags=_lib.PGC_LAWLER_PICK_LAST):
        self.ctx = _lib.default_context(device)
        self.device = torch.device("cuda", device)
        self.flags = flags
        self.K = list(np.broadcast_to(np.asarray(max_depth, np.int64), (len(pairs),)))
        self.shapes = []
        self.bases = []
        for (VX, EX, VY, EY) in pairs:
            VX = np.asarray(VX, np.float64)
            VY = np.asarray(VY, np.float64)
            fd0 = np.zeros((len(VX), len(VY)), np.float64)
            g = _lib.Graph.generate(self.ctx, VX, EX, VY, EY, fd0, False, regularising, False, cost_mode, time_mode)
            self.bases.append(g)
            self.shapes.append((len(VX), len(VY)))
        self._items = {}

    def item_graph(self, slot, pair, fd64):
        key = (slot, pair)
        g = self._items.get(key)
        if g is None:
            g = self.bases[pair].derive(fd64)
            self._items[key] = g
        else:
            g.set_feat_diff(fd64)
        return g

    def close(self):
        for g in self._items.values():
            g.close()
        self._items.clear()
        for g in self.bases:
            g.close()
        self.bases.clear()


class _CycleRatio(torch.autograd.Function):
    @staticmethod
    def forward(ctx, problem, pair_index, *fds):
        n = len(fds)
        for i, f in enumerate(fds):
            if tuple(f.shape) != problem.shapes[pair_index[i]]:
                raise ValueError("feat_diff %d has shape %s, pair %d needs %s"
                                 % (i, tuple(f.shape), pair_index[i], problem.shapes[pair_index[i]]))
        # all feat_diffs in one float64 buffer (one copy and one conversion for
        # the whole call), item i at offset off[i]
        sizes = [int(f.numel()) for f in fds]
        off = np.zeros(n + 1, np.int64)
        np.cumsum(sizes, out=off[1:])
        flat = torch.cat([f.detach().reshape(-1) for f in fds]).to(problem.device, torch.float64)
        base = flat.data_ptr()
        ptrs = [base + 8 * int(off[i]) for i in range(n)]
        torch.cuda.current_stream(problem.device).synchronize()  # flat is read on libpgc's stream
        # one graph per (position, pair): items of one call never share a graph
        graphs, update = [], []
        for i in range(n):
            key = (i, pair_index[i])
            g = problem._items.get(key)
            if g is None:
                g = problem.bases[pair_index[i]].derive(flat[off[i]:off[i + 1]].view(problem.shapes[pair_index[i]]))
                problem._items[key] = g
            else:
                update.append(i)
            graphs.append(g)
        _lib.set_feat_diff_batch([graphs[i] for i in update], [ptrs[i] for i in update], 0)
        Ks = [problem.K[pair_index[i]] for i in range(n)]
        out = _lib.solve_ratio_batch(problem.ctx, graphs, Ks, problem.flags, with_rows=False)
        cycles = [o[0] for o in out]
        vals = _lib.cycle_ratio_values(problem.ctx, graphs, ptrs, cycles)
        ctx.problem = problem
        ctx.pair_index = list(pair_index)
        ctx.graphs = graphs
        ctx.cycles = cycles
        ctx.off = off
        ctx.dtypes = [f.dtype for f in fds]
        problem.last_cycles = cycles
        problem.last_results = [o[2] for o in out]
        return torch.as_tensor(vals, dtype=torch.float64, device=problem.device)

    @staticmethod
    def backward(ctx, grad):
        problem = ctx.problem
        go = grad.detach().to(torch.float64).cpu().numpy()
        off = ctx.off
        n = len(ctx.graphs)
        # all gradients in one zeroed float64 buffer, converted in one step
        flat = torch.zeros(int(off[-1]), dtype=torch.float64, device=problem.device)
        base = flat.data_ptr()
        torch.cuda.current_stream(problem.device).synchronize()
        _lib.cycle_ratio_grad(problem.ctx, ctx.graphs, ctx.cycles, go, [base + 8 * int(off[i]) for i in range(n)])
        shapes = [problem.shapes[p] for p in ctx.pair_index]
        if len(set(ctx.dtypes)) == 1:
            conv = flat.to(ctx.dtypes[0])
            outs = [conv[off[i]:off[i + 1]].view(shapes[i]) for i in range(n)]
        else:
            outs = [flat[off[i]:off[i + 1]].view(shapes[i]).to(ctx.dtypes[i]) for i in range(n)]
        return (None, None) + tuple(outs)


def cycle_ratio(problem: RatioProblem, fds, pair_index=None):
    """lambda* per item (float64 tensor [n]); fds: list of [nVX, nVY] tensors,
    item i on pair pair_index[i] (default: item i on pair i)."""
    fds = list(fds)
    if pair_index is None:
        pair_index = list(range(len(fds)))
    pair_index = tuple(int(i) for i in pair_index)
    if len(pair_index) != len(fds):
        raise ValueError("one pair index per feat_diff")
    return _CycleRatio.apply(problem, pair_index, *fds)


def smoothed_cycle_ratio(problem: RatioProblem, fds, sigma=0.05, samples=32, generator=None, pair_index=None):
    """Perturbation-smoothed lambda*: the mean over `samples` Gaussian
    perturbations FD + sigma * N(0, 1) per item, all solved in one batch; its
    gradient is the mean of the per-sample gradients."""
    fds = list(fds)
    if pair_index is None:
        pair_index = list(range(len(fds)))
    noisy, owner = [], []
    for i, f in enumerate(fds):
        for _ in range(samples):
            z = torch.randn(f.shape, generator=generator, dtype=f.dtype, device="cpu").to(f.device)
            noisy.append(f + sigma * z)
            owner.append(pair_index[i])
    lam = cycle_ratio(problem, noisy, owner)
    return lam.view(len(fds), samples).mean(dim=1)
