#!/usr/bin/env python3
"""Benchmark of the MI355X product-graph cycle solvers.

Workload (one step = one complete `lawlergpu` solve of one shape pair, graph
construction and quantization included): icosphere subdivision 4 (2562
vertices, 15360 directed edges) against a 128-point closed curve, max_depth 3,
cost vanilla / time nomode (mean problem), non-conjugate, no regulariser, no
pruning; seeded synthetic features. The inputs (VX, EX, VY, EY, feat_diff) are
device tensors before the timed region starts.

Reported: solves/s (value), ms per solve, min-cost (`dijkstragpu`) solves/s on
the same pair, Bellman-Ford GTEPS (edge relaxations / solver-kernel time), the
roofline of the dominant kernel (the persistent solver kernel, HBM-bound:
algorithmic bytes per sweep = E*(4 src + 4 cost [+ 4 time]) + V_x*4 parent,
summed over the sweeps of each launch, divided by the launch's HIP-event
duration), and both CPU solvers of the oracle (the restated lawlercpu and
dijkstracpu paths) timed on the host beside their GPU counterparts.

The "batched" object is the serving-scale workload: 256 shape pairs of uneven
size (icosphere subdivision in {2,3,4}, |VY| in [32,160], max_depth in
{1,2,3}), 8 Gaussian feat_diff perturbations (sigma 0.05) per pair = 2304
lawlergpu solves, each step one forward (energy from the perturbed feat_diff,
quantization, solve, lambda*) and backward (dL/dfeat_diff) of all 2304 items
through cyclic_product_graphs.autograd; against the plain per-item loop
over the same 2304 items (energy + one pgc_solve_ratio call per item), with
the cycles of both compared item by item. Its "corrected" object times the
same steps with the corrected odd-layer fold (PGC_CORRECTED_ODD_L) and counts
the odd-layer items whose cycle and lambda* change; the reference's fold is
the one timed everywhere else.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "product-graph-cycles_amd"))

HBM_PEAK_GBS = 8000.0


def _pair():
    from cyclic_product_graphs import synthetic as S
    return S.make_pair(4, 128, seed=0)


def cpu_baseline(p, K, seconds):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pgc_oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    O.set_threads(threads)
    n = 0
    t0 = time.perf_counter()
    while True:
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
        O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=True)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 2:
            break
    return {"value": n / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": "%d complete lawler solves (product space + energy + quantization + BnB) of the same pair, "
                      "%.1f s, oracle/pgc_oracle.cpp with OpenMP" % (n, el)}


def cpu_baseline_mincost(p, K, seconds):
    """The oracle's dijkstracpu restatement (CPU rules) on the same pair."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pgc_oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    O.set_threads(threads)
    n = 0
    t0 = time.perf_counter()
    while True:
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
        O.dijkstra_solve(src, trgt, en[:, 0], len(p["VX"]), K, cpu_rules=True)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds and n >= 2:
            break
    return {"value": n / el, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": "%d complete dijkstracpu solves (product space + energy + BnB sweep) of the same pair, "
                      "%.1f s, oracle/pgc_oracle.cpp with OpenMP" % (n, el)}


def _profile_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def copy_calibration(mib=256, reps=4):
    """A plain device-to-device copy of `mib` MiB, `reps` times: the per-dispatch
    FETCH_SIZE / WRITE_SIZE of this kernel against its known byte count gives the
    counters' scale on this GPU."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.int32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps}))


def _tuning(prefix):
    """The autotuned launch choices of a workload (cyclic_product_graphs/
    tuning_mi355x.json, written by python -m cyclic_product_graphs.autotune),
    applied to the default context; None: the built-in rules."""
    from cyclic_product_graphs import _lib, autotune
    choices = autotune.choices_for(prefix)
    _lib.default_context(0).set_tuning(choices)
    return choices


def batched(args):
    """Batched forward+backward over uneven pairs vs the plain per-item loop."""
    import numpy as np
    import torch
    from cyclic_product_graphs import _lib, synthetic as S
    from cyclic_product_graphs import autograd as A

    tuned = _tuning("C4")

    pairs = S.make_uneven_pairs(args.batch_pairs, seed=0)
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    prob = A.RatioProblem([(p["VX"], p["EX"], p["VY"], p["EY"]) for p, _ in pairs], [K for _, K in pairs])
    gen = torch.Generator(device=dev).manual_seed(1)
    base = [torch.tensor(p["feat_diff"], dtype=torch.float32, device=dev) for p, _ in pairs]
    S_ = args.batch_samples
    fds, owner = [], []
    for i, f in enumerate(base):
        # the unperturbed feat_diff and S_ Gaussian perturbations of it
        fds.append(f.clone().requires_grad_(True))
        owner.append(i)
        for _ in range(S_):
            fds.append((f + 0.05 * torch.randn(f.shape, generator=gen, device=dev)).requires_grad_(True))
            owner.append(i)
    n = len(fds)

    def step():
        for f in fds:
            f.grad = None
        lam = A.cycle_ratio(prob, fds, owner)
        lam.sum().backward()
        torch.cuda.synchronize()
        return lam

    step()  # allocates every item's graph and scratch
    setup_s = time.perf_counter() - t0
    torch.cuda.synchronize()
    b0 = time.perf_counter()
    for _ in range(args.batch_steps):
        lam = step()
    b_el = time.perf_counter() - b0
    res = prob.last_results
    cycles = [c.copy() for c in prob.last_cycles]
    kms = res[0].kernel_ms
    relax = sum(r.relaxations for r in res)
    sweeps = sum(r.sweeps for r in res)
    # algorithmic bytes of every item's sweeps, per sweep as for the single
    # pair: 4 B source + 4 B cost per product edge, 4 B parent per node
    infos = [b.info() for b in prob.bases]
    alg_bytes = 0.0
    for r, i in zip(res, owner):
        inf = infos[i]
        vx = inf["nodes_per_layer"] * (inf["n_layers"] + 1) * int(pairs[i][1])
        alg_bytes += r.sweeps * (inf["n_edges"] * 8 + vx * 4)
    b_achieved = alg_bytes / (kms * 1e-3) / 1e9
    corrected = None
    if not args.no_corrected:
        corrected = batched_corrected(args, prob, step, cycles, lam.detach().cpu().numpy(), pairs, owner)
    tr = _profile_traffic("batch_step_c4_traffic.json")
    b_roof = {"bound": "hbm", "achieved": b_achieved, "peak": 8000.0, "unit": "GB/s",
              "frac": b_achieved / 8000.0, "algorithmic_bytes": alg_bytes,
              "traffic": tr["traffic_bytes_per_step"] if tr else None}
    if args.no_plain:
        prob.close()
        return {"workload": "%d uneven pairs x (1 + %d perturbations) = %d lawlergpu solves, forward+backward" % (
            len(pairs), S_, n), "batched_ms_per_step": b_el / args.batch_steps * 1e3,
            "batched_solver_kernel_ms": kms, "batched_roofline": b_roof}

    # plain loop over the same items: energy from the item's feat_diff, then
    # one pgc_solve_ratio call per item, on one graph per pair
    plain = [prob.bases[i].derive(fds[i * (S_ + 1)].detach().double().contiguous()) for i in range(len(pairs))]
    fd64 = [f.detach().double().contiguous() for f in fds]
    torch.cuda.synchronize()
    same = True
    p0 = time.perf_counter()
    for i in range(n):
        g = plain[owner[i]]
        g.set_feat_diff(fd64[i])
        e, _, r = g.solve_ratio(pairs[owner[i]][1])
        same &= bool(np.array_equal(e, cycles[i]))
    p_el = time.perf_counter() - p0
    for g in plain:
        g.close()
    prob.close()
    return {
        "workload": "%d uneven pairs x (1 + %d perturbations) = %d lawlergpu solves, forward+backward" % (
            len(pairs), S_, n),
        "tuning": tuned,
        "batched_solves_per_s": n * args.batch_steps / b_el,
        "batched_ms_per_step": b_el / args.batch_steps * 1e3,
        "batched_solver_kernel_ms": kms,
        "batched_gteps": relax / (kms * 1e-3) / 1e9,
        "batched_sweeps": sweeps,
        # the solver launches of one step against the HBM roof; their time is
        # the makespan of the longest item, while most CUs idle near the end
        "batched_roofline": b_roof,
        "plain_loop_solves_per_s": n / p_el,
        "plain_loop_s": p_el,
        "batched_vs_plain": (n * args.batch_steps / b_el) / (n / p_el),
        "same_cycles_as_plain_loop": same,
        "setup_s": setup_s,
        "corrected": corrected,
    }


def batched_corrected(args, prob, step, cycles, lam, pairs, owner):
    """The same items with the corrected odd-layer fold (PGC_CORRECTED_ODD_L):
    time per forward+backward step, and, over the items with an odd layer
    count and max_depth >= 2 (the only ones whose fold differs), how many
    cycles and lambda* change and by how much lambda* drops on average."""
    import numpy as np
    import torch
    from cyclic_product_graphs import _lib

    prob.flags |= _lib.PGC_CORRECTED_ODD_L
    try:
        step()
        torch.cuda.synchronize()
        c0 = time.perf_counter()
        for _ in range(args.batch_steps):
            lam_c = step()
        c_el = time.perf_counter() - c0
        lam_c = lam_c.detach().cpu().numpy()
        odd = [i for i in range(len(owner))
               if len(pairs[owner[i]][0]["EY"]) % 2 == 1 and pairs[owner[i]][1] >= 2]
        changed = [i for i in odd if not np.array_equal(prob.last_cycles[i], cycles[i])]
        odd_set = set(odd)
        even_same = all(np.array_equal(prob.last_cycles[i], cycles[i])
                        for i in range(len(owner)) if i not in odd_set)
        rel = [(lam_c[i] - lam[i]) / abs(lam[i]) for i in changed if lam[i] != 0]
        return {
            "ms_per_step": c_el / args.batch_steps * 1e3,
            "odd_L_items": len(odd),
            "odd_L_cycles_changed": len(changed),
            "odd_L_lambda_lower": int(sum(1 for i in changed if lam_c[i] < lam[i])),
            "odd_L_lambda_higher": int(sum(1 for i in changed if lam_c[i] > lam[i])),
            "mean_rel_lambda_change_of_changed": float(np.mean(rel)) if rel else 0.0,
            "other_items_same_cycle": bool(even_same),
        }
    finally:
        prob.flags &= ~_lib.PGC_CORRECTED_ODD_L


def conjugate(steps=2):
    """Conjugate product graph with normals: icosphere subdivision 4 (radial
    noise 0.05) x 100-point curve, max_depth 2, regularising term on, cost
    vanilla / time nomode (the one-column mean problem). One step = graph
    construction (product space, node rotations, energy) + one lawlergpu solve."""
    import numpy as np
    import cyclic_product_graphs as cpg
    from cyclic_product_graphs import synthetic as S
    tuned = _tuning("C2")
    p = S.make_pair(4, 100, seed=0, with_normals=True)
    NX, NY = p["NX"].astype(np.float64), p["NY"].astype(np.float64)

    def one():
        g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], True, True)
        g.set_normals(NX, NY)
        g.set_cost_time_ratio_mode("vanilla", "nomode")
        t0 = time.perf_counter()
        g.generate()
        t1 = time.perf_counter()
        g.solve_with_cost_time_ratio("lawlergpu", 2)
        t2 = time.perf_counter()
        return g, t1 - t0, t2 - t1

    one()
    runs = [one() for _ in range(steps)]
    g = runs[-1][0]
    info = g._graph.info()
    r = g.last_result
    # algorithmic bytes per sweep as for the single pair: 4 B source + 4 B
    # q-cost per product edge, 4 B parent per expanded node
    vx = info["nodes_per_layer"] * (info["n_layers"] + 1) * 2
    bps = info["n_edges"] * 8 + vx * 4
    achieved = bps * r["sweeps"] / (r["kernel_ms"] * 1e-3) / 1e9
    tr = _profile_traffic("solve_kernel_c2_traffic.json")
    launches = max(1, r["launches"])
    return {"workload": "conjugate, icosphere s=4 x 100-point curve, max_depth 2, regularised, vanilla/nomode",
            "tuning": tuned, "n_edges": info["n_edges"], "nodes_per_layer": info["nodes_per_layer"],
            "generate_s": sum(x[1] for x in runs) / steps, "lawler_solve_s": sum(x[2] for x in runs) / steps,
            "sweeps": r["sweeps"], "solver_kernel_ms": r["kernel_ms"], "solver_phase_ms": r["phase_ms"],
            "gteps": r["relaxations"] / (r["kernel_ms"] * 1e-3) / 1e9,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_sweep": bps,
                         "algorithmic_bytes_per_launch": bps * r["sweeps"] / launches,
                         "avg_launch_ms": r["kernel_ms"] / launches,
                         "traffic": tr["traffic_bytes_per_launch"] if tr else None},
            "distance_table": "global memory (46080 nodes per layer exceed the LDS)"}


def cap_case(reps=2, implicit=False):
    """The C4 item that sets the batch makespan: pair 111 of the uneven batch
    (icosphere s=3, |VY| = 41, max_depth 1) with the feat_diff perturbation of
    tests/golden/bf_cap_case.npz, whose Bellman-Ford calls run to the sweep cap
    (26,363 sweeps). One lawlergpu solve per repetition on a fresh graph; prints
    one JSON line per solve with the solver's phase times and whether the cycle
    equals the fixture's."""
    import numpy as np
    from cyclic_product_graphs import _lib, synthetic as S
    z = np.load(os.path.join(ROOT, "tests", "golden", "bf_cap_case.npz"))
    p, K = S.make_uneven_pairs(256, seed=0)[int(z["pair_index"])]
    ctx = _lib.default_context(0)
    if implicit:
        ctx = _lib.Context(0)
        ctx.set_tuning({"implicit_costs": 1})
    for _ in range(reps):
        g = _lib.Graph.generate(ctx, p["VX"], p["EX"], p["VY"], p["EY"], z["feat_diff"].astype(np.float64), False,
                                False, False, "vanilla", "")
        e, _, res = g.solve_ratio(K)
        d = res.as_dict()
        print(json.dumps({"workload": "C4 sweep-cap item (pair 111, npl 642, L 41, K 1)", "kernel_ms": res.kernel_ms,
                          "sweeps": res.sweeps, "phase_ms": d["phase_ms"], "implicit_costs": res.implicit_costs,
                          "same_cycle_as_fixture": bool(np.array_equal(e, z["edges"]))}), flush=True)
        g.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched serving-scale workload")
    ap.add_argument("--no-corrected", action="store_true",
                    help="skip the batched workload's second pass with the corrected odd-layer fold")
    ap.add_argument("--batch-pairs", type=int, default=256)
    ap.add_argument("--batch-samples", type=int, default=8, help="perturbations per pair (plus the unperturbed one)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--no-conj", action="store_true", help="skip the conjugate-graph measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a device copy of known size (for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--implicit", action="store_true",
                    help="with --cap-case: recompute the slot costs from feat_diff (pgc_tuning.implicit_costs)")
    ap.add_argument("--cap-case", action="store_true",
                    help="run only the batch item that runs to the sweep cap (for profiling its solver launch)")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched workload (its steps and the plain loop), print its JSON object")
    ap.add_argument("--no-plain", action="store_true", help="batched workload without the plain per-item loop")
    ap.add_argument("--conj-only", action="store_true", help="run only the conjugate-graph measurement")
    ap.add_argument("--conj-steps", type=int, default=2)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.cap_case:
        cap_case(implicit=args.implicit)
        return
    if args.batch_only:
        print(json.dumps(batched(args)))
        return
    if args.conj_only:
        print(json.dumps(conjugate(args.conj_steps)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path in this build "
              "(independent pairs would shard across GPUs with no collective). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cyclic_product_graphs import _lib

    K = 3
    p = _pair()
    dev = torch.device("cuda:0")
    VX = torch.tensor(p["VX"], dtype=torch.float64, device=dev)
    VY = torch.tensor(p["VY"], dtype=torch.float64, device=dev)
    EX = torch.tensor(p["EX"], dtype=torch.int32, device=dev)
    EY = torch.tensor(p["EY"], dtype=torch.int32, device=dev)
    FD = torch.tensor(p["feat_diff"].astype(np.float64), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ctx = _lib.default_context(0)
    tuned_c1 = _tuning("C1")

    def step(mincost=False):
        g = _lib.Graph.generate_device(ctx, VX, EX, VY, EY, FD)
        if mincost:
            out = g.solve_min_cost(K, 0)
        else:
            out = g.solve_ratio(K, _lib.PGC_LAWLER_PICK_LAST)
        info = g.info()
        g.close()
        return out, info

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = [step() for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    el = t1 - t0

    # min-cost solver on the same pair (fewer steps)
    nd = max(3, args.steps // 2)
    step(mincost=True)
    torch.cuda.synchronize()
    d0 = time.perf_counter()
    dres = [step(mincost=True) for _ in range(nd)]
    torch.cuda.synchronize()
    d_el = time.perf_counter() - d0

    info = results[-1][1]
    E = info["n_edges"]
    npl = info["nodes_per_layer"]
    L = info["n_layers"]
    Vx = npl * (L + 1) * K
    bytes_per_sweep = E * 8 + Vx * 4  # no time column: cost 4 B + src 4 B per edge, parent 4 B per node
    sweeps = sum(r[0][2].sweeps for r in results)
    kms = sum(r[0][2].kernel_ms for r in results)
    relax = sum(r[0][2].relaxations for r in results)
    launches = sum(max(1, r[0][2].launches) for r in results)
    phases = {}
    for r in results:
        for k, v in r[0][2].as_dict()["phase_ms"].items():
            phases[k] = phases.get(k, 0.0) + v / args.steps
    achieved = bytes_per_sweep * sweeps / (kms * 1e-3) / 1e9
    # HBM bytes per solver launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE summary of this workload (scaled by the copy calibration)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "solve_kernel_c1_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["traffic_bytes_per_launch"]
    gteps = relax / (kms * 1e-3) / 1e9

    # parity of the last step with the CPU oracle (outside the timed region)
    parity = None
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pgc_oracle as O
        rows, src, trgt = O.product_space(p["EX"], p["EY"])
        en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64))
        e, oobj, _, _, _ = O.ctr_solve(src, trgt, en, npl, K, gpu_pick=True)
        (eids, rws, res), _ = results[-1]
        parity = bool(np.array_equal(rws, rows[e]) and np.float32(res.objective) == oobj)
    except Exception as ex:  # the checker is optional at run time
        print("parity check skipped: %s" % ex, file=sys.stderr)

    _lib.default_context(0).set_tuning(None)
    batch = None if args.no_batch else batched(args)
    conj = None if args.no_conj else conjugate(args.conj_steps)

    cpu = cpu_mc = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p, K, args.cpu_seconds)
        cpu_mc = cpu_baseline_mincost(p, K, args.cpu_seconds / 2)

    line = {
        "metric": "solves/sec & Bellman-Ford GTEPS; batched fwd+bwd vs plain loop, same cycle",
        "value": args.steps / el,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded icosphere/curve pair, planted descriptors)",
        "config": {"workload": "lawlergpu single pair, icosphere s=4 (2562 V, 15360 directed E) x 128-point curve, "
                               "max_depth=3, vanilla/nomode, inputs resident in HBM",
                   "n_edges": E, "nodes_per_layer": npl, "layers": L, "expanded_nodes": Vx},
        "gteps": gteps,
        "tuning": tuned_c1,
        "sweeps_per_solve": sweeps / args.steps,
        "solver_launches_per_solve": launches / args.steps,
        "solver_kernel_ms_per_solve": kms / args.steps,
        "solver_phase_ms_per_solve": phases,
        "mincost_solves_per_s": nd / d_el,
        "mincost_ms_per_solve": d_el / nd * 1e3,
        "mincost_cpu_baseline": cpu_mc,
        "parity_with_oracle": parity,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "pgc::solve_kernel (ratio mode)",
                     "bytes_per_sweep": bytes_per_sweep,
                     "algorithmic_bytes_per_launch": bytes_per_sweep * sweeps / launches,
                     "avg_launch_ms": kms / launches},
        "cpu_baseline": cpu,
        "batched": batch,
        "conjugate": conj,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
