"""Autotuner of the solvers' launch choices (pgc_tuning).

The built-in rules pick, per graph, how a solve is laid out on the device:
whether the predecessor table, the wrap map and (max_depth 1) the rows of
chain origins live in LDS, whether layer-invariant source slots are held in
registers, whether slots that are padding on every node are skipped, and for
graphs whose distance table exceeds the LDS how many workgroups of a
cooperative launch share a sweep. None of these changes a result; they only
change the time. The tuner times alternatives on a workload by coordinate
descent from the rules (each knob in turn is set to each of its values while
the others stay; a change is kept when it is faster by more than the noise
margin; passes repeat until none is), checks that every candidate returns the
same cycles, and records the fastest choice per workload key in a JSON cache.

Workloads are timed on the device (the solver kernels' HIP-event time): a
single ratio solve of a pair on a freshly derived graph per repetition (the
energy is quantized in place, so a graph is solved once), or one batched
solve of a list of (pair, feat_diff, max_depth) items.

Usage (on the GPU): python -m cyclic_product_graphs.autotune --out tuning.json
tunes the benchmark's three workloads (the single pair, the conjugate pair
and the serving batch) and writes the cache; bench.py applies the cached
choice of each workload and prints it.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import _lib

# knob -> values tried (-1 is the built-in rule)
KNOBS = {
    "grid_chunks_per_wave": (-1, 0, 1, 2, 3, 4),
    "pred_in_lds": (-1, 0),
    "wrap_in_lds": (-1, 0),
    "origins": (-1, 0),
    "cached_sources": (-1, 0),
    "short_slots": (-1, 0),
}
RULES = {k: -1 for k in KNOBS}
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")


def graph_key(info, max_depth, conjugate=False):
    """Cache key of a single solve: the graph's shape and depth."""
    return "%s npl=%d L=%d K=%d cols=%d" % ("conj" if conjugate else "ratio", info["nodes_per_layer"],
                                           info["n_layers"], int(max_depth), info["n_cost_cols"])


class SingleSolve:
    """One lawlergpu solve of a base graph's energy (a derived graph per run)."""

    def __init__(self, ctx, base, feat_diff, max_depth, flags=_lib.PGC_LAWLER_PICK_LAST, conjugate=False):
        self.ctx, self.base, self.fd, self.K, self.flags = ctx, base, feat_diff, max_depth, flags
        self.conjugate = conjugate
        self.key = graph_key(base.info(), max_depth, conjugate)

    def run(self):
        if self.conjugate:
            # conjugate graphs have no derived graphs: regenerate through the factory
            g = self.base.regenerate()
        else:
            g = self.base.derive(self.fd)
        e, _, res = g.solve_ratio(self.K, self.flags)
        g.close()
        return res.kernel_ms, e.tobytes()


class BatchSolve:
    """One pgc_solve_ratio_batch over items (base graph index, feat_diff, max_depth)."""

    def __init__(self, ctx, bases, items, flags=_lib.PGC_LAWLER_PICK_LAST, key=None):
        self.ctx, self.bases, self.items, self.flags = ctx, bases, items, flags
        tot = sum(bases[i].info()["n_edges"] for i, _, _ in items)
        self.key = key or "batch n=%d edges=%d" % (len(items), tot)

    def run(self):
        graphs = [self.bases[i].derive(fd) for i, fd, _ in self.items]
        outs = _lib.solve_ratio_batch(self.ctx, graphs, [k for _, _, k in self.items], self.flags, with_rows=False)
        ms = outs[0][2].kernel_ms
        sig = b"".join(o[0].tobytes() for o in outs)
        for g in graphs:
            g.close()
        return ms, sig


def _time(ctx, workload, choices, reps):
    ctx.set_tuning(choices)
    ms, sig = [], None
    for _ in range(reps):
        t, s = workload.run()
        ms.append(t)
        if sig is None:
            sig = s
        elif s != sig:
            raise RuntimeError("a tuning candidate changed the result of its own repetitions")
    return float(np.median(ms)), sig


def tune(ctx, workload, knobs=None, reps=3, margin=0.02, log=print):
    """Coordinate descent from the built-in rules; returns the cache entry."""
    knobs = knobs or KNOBS
    best = dict(RULES)
    rule_ms, ref = _time(ctx, workload, best, reps)
    best_ms = rule_ms
    tried = [{"choices": dict(best), "ms": rule_ms}]
    log("  rules: %.3f ms" % rule_ms)
    improved = True
    while improved:
        improved = False
        for k, values in knobs.items():
            for v in values:
                if v == best[k]:
                    continue
                cand = dict(best, **{k: v})
                ms, sig = _time(ctx, workload, cand, reps)
                if sig != ref:
                    raise RuntimeError("tuning %s=%s changed the solver's result" % (k, v))
                tried.append({"choices": cand, "ms": ms})
                log("  %s=%d: %.3f ms" % (k, v, ms))
                if ms < best_ms * (1.0 - margin):
                    best, best_ms, improved = cand, ms, True
    ctx.set_tuning(None)
    return {"choices": best, "ms": best_ms, "rule_ms": rule_ms, "candidates": tried}


def load_cache(path=DEFAULT_CACHE):
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def choices_for(key, cache=None):
    """The cached choice of a workload key, or of the first key starting with
    `key` (None: the built-in rules)."""
    cache = load_cache() if cache is None else cache
    wl = cache.get("workloads", {})
    entry = wl.get(key)
    if entry is None:
        entry = next((v for k, v in wl.items() if k.startswith(key)), None)
    return None if entry is None else entry["choices"]


def benchmark_workloads(ctx):
    """The benchmark's three workloads, each with the knobs that apply to it."""
    from . import synthetic as S
    out = []
    p = S.make_pair(4, 128, seed=0)
    base = _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], np.zeros(p["feat_diff"].shape), False,
                               False, False, "vanilla", "")
    w = SingleSolve(ctx, base, p["feat_diff"].astype(np.float64), 3)
    w.key = "C1 " + w.key
    out.append((w, {k: KNOBS[k] for k in ("pred_in_lds", "wrap_in_lds", "cached_sources", "short_slots")}))
    pc = S.make_pair(4, 100, seed=0, with_normals=True)
    conj = _ConjFactory(ctx, pc)
    w = SingleSolve(ctx, conj, None, 2, conjugate=True)
    w.key = "C2 " + w.key
    out.append((w, {k: KNOBS[k] for k in ("grid_chunks_per_wave", "cached_sources", "short_slots")}))
    out.append((c4_workload(ctx), {k: KNOBS[k] for k in ("origins", "pred_in_lds", "wrap_in_lds", "cached_sources",
                                                         "short_slots")}))
    return out


def c4_workload(ctx, n_pairs=256, samples=8):
    """The serving batch of bench.py: 256 uneven pairs, each with its
    feat_diff and 8 Gaussian perturbations (sigma 0.05, fp32, drawn on the
    device with the same generator and seed as the benchmark)."""
    import torch
    from . import synthetic as S
    pairs = S.make_uneven_pairs(n_pairs, seed=0)
    bases = [_lib.Graph.generate(ctx, q["VX"], q["EX"], q["VY"], q["EY"], np.zeros(q["feat_diff"].shape), False,
                                 False, False, "vanilla", "") for q, _ in pairs]
    gen = torch.Generator(device="cuda:0").manual_seed(1)
    items = []
    for i, (q, K) in enumerate(pairs):
        f = torch.tensor(q["feat_diff"], dtype=torch.float32, device="cuda:0")
        items.append((i, f.double().contiguous(), K))
        for _ in range(samples):
            items.append((i, (f + 0.05 * torch.randn(f.shape, generator=gen, device="cuda:0")).double().contiguous(),
                          K))
    torch.cuda.synchronize()
    return BatchSolve(ctx, bases, items, key="C4 batch %d pairs x %d" % (n_pairs, samples + 1))


class _ConjFactory:
    """A conjugate graph regenerated per run (its energy is quantized in place)."""

    def __init__(self, ctx, p):
        self.ctx, self.p = ctx, p
        g = self.regenerate()
        self._info = g.info()
        g.close()

    def regenerate(self):
        p = self.p
        return _lib.Graph.generate(self.ctx, p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"].astype(np.float64),
                                   True, True, False, "vanilla", "nomode", NX=p["NX"].astype(np.float64),
                                   NY=p["NY"].astype(np.float64))

    def info(self):
        return dict(self._info)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_CACHE)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    ctx = _lib.default_context(0)
    cache = {"device": "MI355X (gfx950)", "method": "coordinate descent from the built-in rules, median of %d "
             "repetitions, 2%% margin" % args.reps, "workloads": {}}
    t0 = time.time()
    for w, knobs in benchmark_workloads(ctx):
        print("tuning", w.key, flush=True)
        reps = 1 if isinstance(w, BatchSolve) else args.reps
        cache["workloads"][w.key] = tune(ctx, w, knobs, reps=reps, log=lambda s: print(s, flush=True))
    cache["seconds"] = time.time() - t0
    with open(args.out, "w") as f:
        json.dump(cache, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
