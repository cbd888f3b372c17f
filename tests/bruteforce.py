"""Exact optima of once-wrapping cycles on tiny layered product graphs.

Independent of the oracle and of the device solver: per start vertex s it runs
a forward dynamic program over the depth-expanded layered DAG taken straight
from the definition (a cycle enters layer 0 at s, takes at most K-1
intralayer steps per layer and one interlayer step per layer, and returns to
s on the copy of layer 0). For the ratio it keeps, per DAG state and per total
time T (an integer in q-units), the minimum integer cost C, so min C/T is
exact (rational arithmetic, no parametric search); for the raw minimum cost
it keeps the minimum path-order sum.
"""
from fractions import Fraction


def _adjacency(src, trgt, npl):
    n_nodes = int(max(src)) + 1
    L = n_nodes // npl
    out_intra, out_inter = {}, {}
    for e, (s, t) in enumerate(zip(src, trgt)):
        s, t = int(s), int(t)
        if s // npl == t // npl:
            out_intra.setdefault(s, []).append(e)
        else:
            out_inter.setdefault(s, []).append(e)
    return L, out_intra, out_inter


def _dp(src, trgt, npl, K, start, init, extend, better):
    """Generic forward DP; states (layer, depth, v) -> table; returns the table at (L, 0, start)."""
    L, out_intra, out_inter = _adjacency(src, trgt, npl)
    cur = {(0, start): init}  # depth, v on the current layer
    for layer in range(L):
        # intralayer steps in depth order
        for depth in range(K - 1):
            for (d, v), tab in list(cur.items()):
                if d != depth:
                    continue
                for e in out_intra.get(layer * npl + v, []):
                    u = int(trgt[e]) - layer * npl
                    key = (depth + 1, u)
                    cur[key] = better(cur.get(key), extend(tab, e))
        nxt = {}
        for (d, v), tab in cur.items():
            for e in out_inter.get(layer * npl + v, []):
                t = int(trgt[e])
                lt = t // npl
                if lt != (layer + 1) % L:
                    continue
                u = t - lt * npl
                if layer + 1 == L and u != start:
                    continue
                nxt[(0, u)] = better(nxt.get((0, u)), extend(tab, e))
        cur = nxt
    return cur.get((0, start))


def best_ratio(src, trgt, npl, K, cost_q, time_q):
    """(ratio Fraction, C, T, edge list) of the minimum cost/time-ratio once-wrapping cycle."""

    def extend(tab, e):
        c, t = int(cost_q[e]), int(time_q[e])
        return {T + t: (C + c, path + (e,)) for T, (C, path) in tab.items()}

    def better(a, b):
        if a is None:
            return dict(b)
        out = dict(a)
        for T, v in b.items():
            if T not in out or v[0] < out[T][0]:
                out[T] = v
        return out

    best = None
    for s in range(npl):
        tab = _dp(src, trgt, npl, K, s, {0: (0, ())}, extend, better)
        if not tab:
            continue
        for T, (C, path) in tab.items():
            r = Fraction(C, T)
            if best is None or r < best[0]:
                best = (r, C, T, list(path))
    return best


def best_cost(src, trgt, npl, K, cost):
    """(cost, edge list) of the minimum raw-cost once-wrapping cycle (path-order fp64 sums)."""

    def extend(tab, e):
        return (tab[0] + float(cost[e]), tab[1] + (e,))

    def better(a, b):
        return b if a is None or b[0] < a[0] else a

    best = None
    for s in range(npl):
        r = _dp(src, trgt, npl, K, s, (0.0, ()), extend, better)
        if r is not None and (best is None or r[0] < best[0]):
            best = (r[0], list(r[1]))
    return best
