"""Parity of the HIP solvers (through libpgc.so's C-ABI) with the CPU oracle.

Bar: product space and quantized costs bit-exact; the matching (product-space
rows of the returned cycle, walk order) bit-exact; passes_layers_just_once
equal; the fp32 objective equal (tolerance written below: 1e-5 relative, the
observed difference is 0). Raw fp64 energies: bit-exact without the length
term, within 1e-15 relative with it (the reference squares with std::pow,
the device with a multiply).
"""
import numpy as np
import pytest

import pgc_oracle as O
import cyclic_product_graphs as cpg
from cyclic_product_graphs import synthetic as S

pytestmark = pytest.mark.gpu

OBJ_RTOL = 1e-5

LAWLER = {"lawlergpu": True, "lawlercpu": False}
MINCOST = {"dijkstragpu": False, "dijkstracpu": True}


def _gen(p, reg=False, cost="vanilla", time="", prune=False):
    g = cpg.product_graph_generator(p["VX"], p["EX"], p["VY"], p["EY"], p["feat_diff"], False, reg, prune)
    g.set_cost_time_ratio_mode(cost, time)
    return g


def _oracle_graph(p, reg=False, cost="vanilla", time="", prune=False):
    rows, src, trgt = O.product_space(p["EX"], p["EY"], prune=prune)
    en = O.energy(p["VX"], p["VY"], rows, p["feat_diff"].astype(np.float64), reg=reg, cost_mode=cost,
                  time_mode=time)
    return rows, src, trgt, en


def _check_ratio(p, K, name, reg=False, cost="vanilla", time="", prune=False, corrected=False):
    print("case ratio", name, "nx", len(p["VX"]), "ny", len(p["VY"]), "K", K, "corrected", corrected, flush=True)
    g = _gen(p, reg, cost, time, prune)
    passes, obj, rows_gpu = g.solve_with_cost_time_ratio(name, K, corrected=corrected)
    rows, src, trgt, en = _oracle_graph(p, reg, cost, time, prune)
    e, oobj, opasses, stats, q = O.ctr_solve(src, trgt, en, len(p["VX"]), K, gpu_pick=LAWLER[name],
                                             corrected=corrected)
    assert passes == opasses
    assert np.array_equal(rows_gpu, rows[e]), "matching differs"
    assert abs(obj - float(oobj)) <= OBJ_RTOL * abs(float(oobj))
    assert obj == float(oobj)
    assert np.array_equal(g.get_cost_vector(), q), "quantized cost differs"
    st = g.last_result
    assert st["sweeps"] == stats[1] and st["bf_calls"] == stats[0] and st["branches"] == stats[3]
    return g


def _check_mincost(p, K, name, prune=False, corrected=False):
    print("case mincost", name, "nx", len(p["VX"]), "ny", len(p["VY"]), "K", K, "corrected", corrected, flush=True)
    g = _gen(p, prune=prune)
    rows, src, trgt, en = _oracle_graph(p, prune=prune)
    bg = None
    if prune:
        bg = [[] for _ in range(len(p["VX"]))]
        for a, b in p["EX"]:
            bg[a].append(b)
    try:
        e, oobj, opasses, stats = O.dijkstra_solve(src, trgt, en[:, 0], len(p["VX"]), K, branch_graph=bg,
                                                   cpu_rules=MINCOST[name], corrected=corrected)
    except RuntimeError as ex:
        assert "no nodes in branch" in str(ex)
        with pytest.raises(cpg.PgcError):
            g.solve_with_minimum_cost(name, K, corrected=corrected)
        return
    passes, obj, rows_gpu = g.solve_with_minimum_cost(name, K, corrected=corrected)
    assert passes == opasses
    assert np.array_equal(rows_gpu, rows[e]), "matching differs"
    assert obj == float(oobj)
    st = g.last_result
    assert st["sweeps"] == stats[1] and st["branches"] == stats[3]


TINY = [(s, nx, ny, ring, K) for s, (nx, ny, ring) in enumerate([(4, 4, True), (4, 4, False), (5, 6, True),
                                                                  (3, 4, False), (6, 5, True), (4, 7, False)])
        for K in (1, 2, 3)]


@pytest.mark.parametrize("seed,nx,ny,ring,K", TINY)
@pytest.mark.parametrize("name", list(LAWLER))
def test_ratio_tiny(seed, nx, ny, ring, K, name):
    _check_ratio(S.tiny_pair(nx, ny, seed=seed, ring=ring), K, name)


@pytest.mark.parametrize("seed,nx,ny,ring,K", TINY)
@pytest.mark.parametrize("name", list(MINCOST))
def test_mincost_tiny(seed, nx, ny, ring, K, name):
    _check_mincost(S.tiny_pair(nx, ny, seed=seed, ring=ring), K, name)


# odd layer counts (the Y curve is closed, so L = |VY|) and K >= 2: the only
# cases where the corrected fold differs from the reference's
ODD_L = [(s, nx, ny, ring, K) for s, (nx, ny, ring) in enumerate([(4, 5, True), (5, 5, True), (3, 7, True),
                                                                  (4, 7, False), (6, 3, True)])
         for K in (2, 3)]


@pytest.mark.parametrize("seed,nx,ny,ring,K", ODD_L)
@pytest.mark.parametrize("name", list(LAWLER))
def test_ratio_corrected_odd_L(seed, nx, ny, ring, K, name):
    _check_ratio(S.tiny_pair(nx, ny, seed=seed + 40, ring=ring), K, name, corrected=True)


@pytest.mark.parametrize("seed,nx,ny,ring,K", ODD_L)
@pytest.mark.parametrize("name", list(MINCOST))
def test_mincost_corrected_odd_L(seed, nx, ny, ring, K, name):
    _check_mincost(S.tiny_pair(nx, ny, seed=seed + 40, ring=ring), K, name, corrected=True)


@pytest.mark.parametrize("sub,ny,K,seed", [(2, 33, 2, 1), (2, 41, 3, 4)])
def test_ratio_corrected_meshes(sub, ny, K, seed):
    _check_ratio(S.make_pair(sub, ny, seed=seed), K, "lawlergpu", corrected=True)


@pytest.mark.parametrize("sub,ny,K,seed", [(1, 8, 2, 0), (2, 33, 2, 1), (2, 33, 3, 2), (2, 20, 1, 3),
                                           (2, 41, 3, 4), (3, 64, 2, 0)])
@pytest.mark.parametrize("name", list(LAWLER))
def test_ratio_meshes(sub, ny, K, seed, name):
    # odd |VY| exercises the reference's fold row numLayers % 2 for K >= 2
    _check_ratio(S.make_pair(sub, ny, seed=seed), K, name)


@pytest.mark.parametrize("sub,ny,K,seed", [(1, 8, 2, 0), (2, 33, 2, 1), (2, 20, 1, 3), (3, 64, 2, 0)])
@pytest.mark.parametrize("name", list(MINCOST))
def test_mincost_meshes(sub, ny, K, seed, name):
    _check_mincost(S.make_pair(sub, ny, seed=seed), K, name)


def test_ratio_modes():
    p = S.make_pair(2, 24, seed=7)
    _check_ratio(p, 2, "lawlergpu", reg=True)
    _check_ratio(p, 2, "lawlergpu", cost="feature", time="lengthnormalisation")
    _check_ratio(p, 2, "lawlergpu", prune=True)


def test_mincost_pruned():
    _check_mincost(S.make_pair(2, 24, seed=8), 2, "dijkstragpu", prune=True)


@pytest.mark.parametrize("reg,cost,time", [(False, "vanilla", ""), (True, "vanilla", ""),
                                           (False, "feature", "lengthnormalisation")])
def test_graph_construction(reg, cost, time):
    p = S.make_pair(3, 48, seed=4)
    g = _gen(p, reg, cost, time)
    rows, src, trgt, en = _oracle_graph(p, reg, cost, time)
    assert np.array_equal(g.get_product_space(), rows)
    got = g.get_cost_vector()
    assert got.shape == en.shape
    if reg or cost == "feature":
        np.testing.assert_allclose(got, en, rtol=1e-15, atol=0)
    else:
        assert np.array_equal(got, en)


def test_state_semantics():
    p = S.make_pair(2, 16, seed=9)
    g = _gen(p)
    e0 = g.get_cost_vector()
    g.set_cost_time_ratio_mode("feature", "lengthnormalisation")  # no effect after generation
    g.solve_with_cost_time_ratio("lawlergpu", 2)
    q1 = g.get_cost_vector()
    assert q1.shape == e0.shape and np.all(q1 == np.round(q1))
    # a second Lawler solve re-quantizes the quantized costs, as the reference does
    rows, src, trgt, en = _oracle_graph(p)
    _, _, _, _, qa = O.ctr_solve(src, trgt, en, len(p["VX"]), 2)
    _, obj_b, _, _, qb = O.ctr_solve(src, trgt, qa, len(p["VX"]), 2, gpu_pick=True)
    _, obj2, _ = g.solve_with_cost_time_ratio("lawlergpu", 2)
    assert np.array_equal(g.get_cost_vector(), qb)
    assert obj2 == float(obj_b)


@pytest.mark.parametrize("name", ["lawlergpu", "dijkstragpu"])
def test_c1_full_size(name):
    # icosphere subdivision 4 (2562 vertices), 128 curve points, max_depth 3
    p = S.make_pair(4, 128, seed=0)
    if name == "lawlergpu":
        _check_ratio(p, 3, name)
    else:
        _check_mincost(p, 3, name)


def _random_cases(n, seed):
    # seeded draws over the configuration space (deterministic case list)
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        sub = int(rng.integers(1, 3))
        ny = int(rng.integers(6, 41))
        K = int(rng.integers(1, 4))
        mode = int(rng.integers(0, 3))  # 0 vanilla, 1 vanilla + regulariser, 2 feature / lengthnormalisation
        prune = bool(rng.integers(0, 4) == 0)
        corrected = bool(rng.integers(0, 2))
        out.append((i, sub, ny, K, mode, prune, corrected, int(rng.integers(1 << 30))))
    return out


@pytest.mark.parametrize("i,sub,ny,K,mode,prune,corrected,seed", _random_cases(16, 2024))
@pytest.mark.parametrize("name", list(LAWLER))
def test_ratio_random_configurations(i, sub, ny, K, mode, prune, corrected, seed, name):
    kw = [{}, {"reg": True}, {"cost": "feature", "time": "lengthnormalisation"}][mode]
    _check_ratio(S.make_pair(sub, ny, seed=seed), K, name, prune=prune, corrected=corrected, **kw)


@pytest.mark.parametrize("i,sub,ny,K,mode,prune,corrected,seed", _random_cases(12, 77))
@pytest.mark.parametrize("name", list(MINCOST))
def test_mincost_random_configurations(i, sub, ny, K, mode, prune, corrected, seed, name):
    _check_mincost(S.make_pair(sub, ny, seed=seed), K, name, prune=prune, corrected=corrected)


# Graphs whose in-degrees reach the slot width, where the sweeps read every
# slot (the icosphere and ring cases above skip the slots that are padding on
# every node): complete X graphs (8 vertices: interlayer in-degree 8 = the
# width; 13 vertices: 13 of 16), and an icosphere with one added edge (a
# vertex of valence 7, interlayer in-degree 8) large enough for the
# several-chunks-per-wave sweep.
WIDE = [(s, nx, ny, K) for s, (nx, ny) in enumerate([(8, 5), (13, 4)]) for K in (1, 2, 3)]


def _with_extra_edge(p):
    EX = np.asarray(p["EX"])
    deg = np.bincount(EX[:, 0], minlength=len(p["VX"]))
    a = int(np.argmax(deg))
    nb = set(EX[EX[:, 0] == a, 1].tolist())
    b = next(v for v in range(len(p["VX"])) if v != a and v not in nb and deg[v] == deg.max())
    EX2 = np.unique(np.vstack([EX, [[a, b], [b, a]]]).astype(np.int32), axis=0)
    return dict(p, EX=EX2)


@pytest.mark.parametrize("seed,nx,ny,K", WIDE)
@pytest.mark.parametrize("name", list(LAWLER))
def test_ratio_full_slot_width(seed, nx, ny, K, name):
    _check_ratio(S.tiny_pair(nx, ny, seed=seed + 60, ring=False), K, name)


@pytest.mark.parametrize("seed,nx,ny,K", WIDE)
@pytest.mark.parametrize("name", list(MINCOST))
def test_mincost_full_slot_width(seed, nx, ny, K, name):
    _check_mincost(S.tiny_pair(nx, ny, seed=seed + 60, ring=False), K, name)


@pytest.mark.parametrize("K", [1, 2, 3])
def test_full_slot_width_many_chunks(K):
    p = _with_extra_edge(S.make_pair(4, 8, seed=5))
    assert np.bincount(p["EX"][:, 0]).max() == 7
    _check_ratio(p, K, "lawlergpu")
    _check_mincost(p, K, "dijkstragpu")
