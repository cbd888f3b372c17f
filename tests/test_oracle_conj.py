"""Pins the conjugate-graph part of the oracle (oracle/conj_oracle.cpp).

The reference ships no tests or fixtures for it either; the checks are:
  * the edge count equals the reference's own formula
    (conj_product_space.cpp:16-39 numOutgoingEdges, :80-82 per-layer counts);
  * every row is reconstructed from the semantics of its SRC and TRGT nodes
    (a node (layer i, k) is the 2D edge EY[i] (degenerate for k in the 2D
    block) times the 3D edge EX[k mod |EX|] (degenerate for the 3D block));
  * the edge set equals an independent enumeration (Python sets) of the
    reference's families, and the edge order is ascending per source edge;
  * the node rotations equal the Kabsch solution U diag(1,1,det(UV^T)) V^T
    of numpy's SVD (the reference's JacobiSVD formula) to 1e-12;
  * the energy table, recomputed in numpy from the pinned rotations.
"""
import numpy as np
import pytest

import pgc_oracle as O
from cyclic_product_graphs import synthetic as S


def _pair(sub=1, ny=6, seed=0):
    return S.make_pair(sub, ny, seed=seed, with_normals=True)


def _node(k, i, p, prune):
    """(2D pair, 3D pair) a conjugate node represents."""
    EX, EY = p["EX"], p["EY"]
    nEX = len(EX)
    L = len(EY)
    a, b = int(EY[i % L, 0]), int(EY[i % L, 1])
    off3d = nEX if prune else 2 * nEX
    if k < nEX:
        return (a, b), (int(EX[k, 0]), int(EX[k, 1]))
    if not prune and k < 2 * nEX:
        return (a, a), (int(EX[k - nEX, 0]), int(EX[k - nEX, 1]))
    e = k - off3d
    v = int(EX[e, 1])
    return (a, b), (v, v)


@pytest.mark.parametrize("prune", [False, True])
def test_conj_rows_match_node_semantics_and_counts(prune):
    p = _pair()
    EX, EY = p["EX"], p["EY"]
    rows, src, trgt = O.conj_product_space(EX, EY, prune=prune)
    nEX, L = len(EX), len(EY)
    npl = 3 * nEX
    outdeg = np.bincount(EX[:, 0], minlength=EX.max() + 1)
    P = int(sum(outdeg[EX[j, 1]] - 1 for j in range(nEX)))
    per_layer = (2 * P if not prune else 0) + (3 * P if not prune else 2 * P) + 2 * nEX
    assert len(rows) == L * per_layer
    for r, s, t in zip(rows, src, trgt):
        ls, ks = divmod(int(s), npl)
        lt, kt = divmod(int(t), npl)
        (a0, a1), (b0, b1) = _node(ks, ls, p, prune)
        (c0, c1), (d0, d1) = _node(kt, lt, p, prune)
        assert (r[0], r[1]) == (a0, a1) and (r[3], r[4]) == (b0, b1)
        assert (r[1], r[2]) == (c0, c1) and (r[4], r[5]) == (d0, d1)
        assert lt == ls or lt == (ls + 1) % L
    # the 3D part never walks straight back (src3d != trgt3d unless degenerate)
    nd = (rows[:, 3] != rows[:, 4]) & (rows[:, 4] != rows[:, 5])
    assert np.all(rows[nd, 3] != rows[nd, 5])


def test_conj_edge_set_and_order():
    p = _pair(1, 5, seed=2)
    EX, EY = p["EX"], p["EY"]
    rows, src, trgt = O.conj_product_space(EX, EY)
    nEX, L = len(EX), len(EY)
    npl = 3 * nEX
    out = {}
    for e, (a, b) in enumerate(EX):
        out.setdefault(int(a), []).append(e)
    want = set()
    for i in range(L):
        n = (i + 1) % L
        s2, m2, t2 = int(EY[i, 0]), int(EY[i, 1]), int(EY[n, 1])
        c, x = i * npl, n * npl
        for j, (s3, m3) in enumerate(EX):
            s3, m3 = int(s3), int(m3)
            for it in out[m3]:
                t3 = int(EX[it, 1])
                if t3 == s3:
                    continue
                want |= {((s2, s2, s2, s3, m3, t3), c + nEX + j, c + nEX + it),
                         ((s2, s2, m2, s3, m3, t3), c + nEX + j, c + it),
                         ((s2, m2, t2, s3, m3, t3), c + j, x + it),
                         ((s2, m2, m2, s3, m3, t3), c + j, x + nEX + it),
                         ((s2, m2, t2, m3, m3, t3), c + 2 * nEX + j, x + it)}
            want |= {((s2, m2, t2, s3, m3, m3), c + j, x + 2 * nEX + j),
                     ((s2, m2, t2, m3, m3, m3), c + 2 * nEX + j, x + 2 * nEX + j)}
    got = {(tuple(int(v) for v in r), int(s), int(t)) for r, s, t in zip(rows, src, trgt)}
    assert got == want and len(got) == len(rows)
    # (source, target) node pairs are unique: edge order cannot change a solve
    assert len({(int(s), int(t)) for s, t in zip(src, trgt)}) == len(src)


def _kabsch(v2, n2, c2, v3, n3, c3):
    def nz(v):
        n = np.linalg.norm(v)
        return v / n if n > 0 else v
    X = np.stack([nz(n2), nz(v2), nz(c2)])
    Y = np.stack([nz(n3), nz(v3), nz(c3)])
    U, s, Vt = np.linalg.svd(X.T @ Y)
    d = np.diag([1, 1, np.linalg.det(U @ Vt)])
    return U @ d @ Vt


def _qmat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("prune", [False, True])
def test_conj_rotations_match_svd_kabsch(prune):
    p = _pair(1, 7, seed=4)
    VX, NX, VY, NY, EX, EY = p["VX"], p["NX"].astype(np.float64), p["VY"], p["NY"].astype(np.float64), p["EX"], p["EY"]
    q = O.conj_quaternions(VX, NX, VY, NY, EX, EY, prune=prune)
    nEX, L = len(EX), len(EY)
    npl = 3 * nEX
    rng = np.random.default_rng(0)
    for node in rng.choice(len(q), 300, replace=False):
        i, j = divmod(int(node), npl)
        ie = (i + L - 1) % L if (nEX <= j < 2 * nEX and not prune) else i
        e2 = VY[EY[ie, 1]] - VY[EY[ie, 0]]
        n2 = NY[EY[ie, 0]]
        ex = j % nEX
        e3 = VX[EX[ex, 1]] - VX[EX[ex, 0]]
        n3 = (NX[EX[ex, 0]] + NX[EX[ex, 1]]) / 2
        R = _kabsch(e2, n2, np.cross(e2, n2), e3, n3, np.cross(e3, n3))
        assert np.abs(_qmat(q[node]) - R).max() < 1e-12


def _rho(x, a=0.7, c=0.6, pw=4.0):
    return abs(a - 2) / a * (((x / c) ** pw / abs(a - 2) + 1) ** (0.5 * a) - 1)


MODES = [("vanilla", "vanilla"), ("vanilla", "nomode"), ("vanilla", ""), ("vanilla", "lengthnormalisation"),
         ("lengthnormalisation", "vanilla"), ("feature", "plainso3"), ("feature", "robustso3"),
         ("robustso3", "feature"), ("plainso3", "feature"), ("", "vanilla")]


@pytest.mark.parametrize("cost,time", MODES)
@pytest.mark.parametrize("reg", [False, True])
def test_conj_energy_table(cost, time, reg):
    p = _pair(1, 5, seed=6)
    VX, NX, VY, NY, EX, EY = p["VX"], p["NX"].astype(np.float64), p["VY"], p["NY"].astype(np.float64), p["EX"], p["EY"]
    FD = p["feat_diff"].astype(np.float64)
    rows, src, trgt = O.conj_product_space(EX, EY)
    en = O.conj_energy(VX, NX, VY, NY, EX, EY, rows, src, trgt, FD, reg=reg, cost_mode=cost, time_mode=time)
    q = O.conj_quaternions(VX, NX, VY, NY, EX, EY)
    if reg:
        ip = np.minimum(1.0, np.abs((q[src] * q[trgt]).sum(1)))
        so3 = 2 * np.arccos(ip)
    else:
        so3 = np.zeros(len(rows))
    fd = FD[rows[:, 5], rows[:, 2]]
    ln = np.sqrt(((VY[rows[:, 0]] - VY[rows[:, 1]]) ** 2).sum(1) + ((VX[rows[:, 3]] - VX[rows[:, 4]]) ** 2).sum(1))
    pen = np.where((rows[:, 2] == rows[:, 1]) | (rows[:, 5] == rows[:, 4]), 0.1, 1.0)
    r = _rho(so3)
    mso3 = float(np.float32(_rho(2 * np.pi)))
    mfd = float(np.float32(FD.max()))
    expect = {("vanilla", "vanilla"): [r + fd], ("vanilla", "nomode"): [r + fd], ("vanilla", ""): [r + fd, pen],
              ("vanilla", "lengthnormalisation"): [r + fd, ln], ("lengthnormalisation", "vanilla"): [ln, r + fd],
              ("feature", "plainso3"): [fd, np.maximum(0, 2 * np.pi - so3)],
              ("feature", "robustso3"): [fd, np.maximum(0, mso3 - r)],
              ("robustso3", "feature"): [r, np.maximum(0, mfd - fd)],
              ("plainso3", "feature"): [so3, np.maximum(0, mfd - fd)],
              ("", "vanilla"): [pen, r + fd]}[(cost, time)]
    assert en.shape == (len(rows), len(expect))
    for k, col in enumerate(expect):
        np.testing.assert_allclose(en[:, k], col, rtol=1e-11, atol=1e-12)


def test_conj_unsupported_mode():
    p = _pair(0, 4)
    rows, src, trgt = O.conj_product_space(p["EX"], p["EY"])
    with pytest.raises(ValueError):
        O.conj_energy(p["VX"], p["NX"], p["VY"], p["NY"], p["EX"], p["EY"], rows, src, trgt, p["feat_diff"],
                      cost_mode="feature", time_mode="feature")


def test_conj_branch_graph():
    p = _pair(1, 4)
    EX = p["EX"]
    nEX = len(EX)
    for prune in (False, True):
        bg = O.conj_branch_graph(EX, prune=prune)
        assert len(bg) == 3 * nEX
        for j in range(nEX):
            outs = [e for e in range(nEX) if EX[e, 0] == EX[j, 0]]
            assert sorted(set([j + nEX] + outs)) == bg[j]
            if not prune:
                assert sorted(set([j + 2 * nEX] + [e + nEX for e in outs])) == bg[j + nEX]
                assert sorted(set([j] + [e + 2 * nEX for e in outs])) == bg[j + 2 * nEX]
            else:
                assert sorted(set([j] + [e + nEX for e in outs])) == bg[j + nEX]
