"""Normal helpers: the oracle's restatements (oracle/normals_oracle.cpp) pinned
by closed forms, and the package's compute_normals_2d / compute_normals_3d
against the oracle.

References: helper/utils.cpp:93-157 (normals2d), helper/utils.cpp:86-90
(normals3d = igl::per_vertex_normals of libigl b1bd5b1, default weighting;
the library is not in the reference tree, so the 3D closed forms below are
its only pins)."""
import numpy as np
import pytest

import cyclic_product_graphs as cpg
from cyclic_product_graphs import synthetic as S
from oracle import pgc_oracle as O


def ellipse(n, a, b, c0, c1, plane="xy", cw=False):
    th = 2.0 * np.pi * np.arange(n) / n
    if cw:
        th = -th
    u, v = c0 + a * np.cos(th), c1 + b * np.sin(th)
    V = np.zeros((n, 3))
    cols = {"xy": (0, 1), "xz": (0, 2), "yz": (1, 2)}[plane]
    V[:, cols[0]], V[:, cols[1]] = u, v
    E = np.stack([np.arange(n), (np.arange(n) + 1) % n], axis=1).astype(np.int32)
    return V, E, th


def ellipse_normals(th, a, b):
    # the length-weighted sum of the two edge normals at point j is the
    # central difference V[j-1] - V[j+1] turned by -90 degrees, which for
    # points on an ellipse is parallel to (b cos t, a sin t)
    N = np.stack([b * np.cos(th), a * np.sin(th), np.zeros_like(th)], axis=1)
    return N / np.linalg.norm(N, axis=1, keepdims=True)


# ------------------------------------------------------------ oracle, 2D
@pytest.mark.parametrize("n,a,b", [(16, 1.0, 1.0), (41, 2.0, 0.5), (100, 0.3, 1.7), (7, 1.0, 3.0)])
@pytest.mark.parametrize("cw", [False, True])
def test_oracle_2d_ellipse_closed_form(n, a, b, cw):
    V, E, th = ellipse(n, a, b, 3.0, -2.0, cw=cw)
    st, zd, N = O.normals2d(V, E)
    assert st == 0 and zd == 2
    np.testing.assert_allclose(N, ellipse_normals(th, a, b), rtol=0, atol=1e-12)


@pytest.mark.parametrize("plane,zd_expected", [("xz", 1), ("yz", 0)])
def test_oracle_2d_coordinate_planes_return_the_permuted_frame(plane, zd_expected):
    # the dropped column moves last; the normals stay in that frame
    V, E, th = ellipse(30, 1.5, 0.7, 2.0, 1.0, plane=plane)
    st, zd, N = O.normals2d(V, E)
    assert st == 0 and zd == zd_expected
    np.testing.assert_allclose(N, ellipse_normals(th, 1.5, 0.7), rtol=0, atol=1e-12)


def test_oracle_2d_centred_circle_drops_the_x_column():
    # a circle centred on the origin: the x column already sums to ~0, so the
    # reference drops x (zeroDimension 0), works in the (y, z) frame with x as
    # the third column, and flips nothing (the box has no y extent, so the
    # tenth of its smaller side is 0); the result is (0, cos t, sin t)
    V, E, th = ellipse(24, 1.0, 1.0, 0.0, 0.0)
    st, zd, N = O.normals2d(V, E)
    assert st == 0 and zd == 0
    ref = np.stack([np.zeros_like(th), np.cos(th), np.sin(th)], axis=1)
    np.testing.assert_allclose(N, ref, rtol=0, atol=1e-12)


def test_oracle_2d_error_returns():
    V, E, _ = ellipse(10, 1.0, 1.0, 2.0, 2.0)
    assert O.normals2d(V, E[:9])[0] == 1
    V3 = V + np.array([0.0, 0.0, 1.0])  # no coordinate column sums to ~0
    st, zd, _ = O.normals2d(V3, E)
    assert st == 2 and zd == -1


# ------------------------------------------------------------ oracle, 3D
def octahedron():
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
    F = []
    for x in (0, 1):
        for y in (2, 3):
            for z in (4, 5):
                f = [x, y, z]
                n = np.cross(V[y] - V[x], V[z] - V[x])
                if np.dot(n, V[x] + V[y] + V[z]) < 0:
                    f = [x, z, y]
                F.append(f)
    return V, np.array(F, np.int32)


@pytest.mark.parametrize("mesh", ["octahedron", "icosahedron", "icosphere2"])
def test_oracle_3d_symmetric_polyhedra_are_radial(mesh):
    if mesh == "octahedron":
        V, F = octahedron()
    else:
        V, F = S.icosphere(0 if mesh == "icosahedron" else 2)
    N = O.normals3d(V, F)
    if mesh == "icosphere2":
        # not vertex-transitive: only the 12 valence-5 vertices are fixed by symmetry
        N, V = N[:12], V[:12]
    np.testing.assert_allclose(N, V / np.linalg.norm(V, axis=1, keepdims=True), rtol=0, atol=1e-12)


def test_oracle_3d_area_weighting_corner():
    # two right-angled faces at vertex 0: (0,1,2) in z=0 with double area
    # s^2 and normal +z, (0,3,4) in y=0 with double area t^2 and normal +y.
    # Area weighting gives (0, t^2, s^2); uniform or angle weighting would
    # give (0, 1, 1). A collinear face (0,1,4) adds nothing; vertex 5 is on
    # no face (NaN).
    s, t = 1.0, 2.0
    V = np.array([[0, 0, 0], [s, 0, 0], [0, s, 0], [0, 0, t], [t, 0, 0], [5, 5, 5]], np.float64)
    F = np.array([[0, 1, 2], [0, 3, 4], [0, 1, 4]], np.int32)
    N = O.normals3d(V, F)
    np.testing.assert_allclose(N[0], np.array([0.0, t * t, s * s]) / np.hypot(t * t, s * s), rtol=0, atol=1e-15)
    np.testing.assert_allclose(N[2], [0.0, 0.0, 1.0], atol=1e-15)
    np.testing.assert_allclose(N[3], [0.0, 1.0, 0.0], atol=1e-15)
    assert np.isnan(N[5]).all()
    # reversed orientation flips the normal
    N2 = O.normals3d(V, F[:, ::-1].copy())
    np.testing.assert_allclose(N2[0], -N[0], atol=1e-15)


def test_oracle_3d_bipyramid_apexes():
    V, F = S.bipyramid(7)
    N = O.normals3d(V, F)
    np.testing.assert_allclose(N[0], [0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(N[-1], [0, 0, -1], atol=1e-15)


# ------------------------------------------------------------ package vs oracle
@pytest.mark.parametrize("seed", range(6))
def test_package_normals_2d_match_oracle(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = int(rng.integers(5, 200))
    plane = ["xy", "xz", "yz"][seed % 3]
    V, E, _ = ellipse(n, 1 + rng.random(), 1 + rng.random(), 1 + rng.random(), -1 - rng.random(), plane=plane,
                      cw=bool(seed & 1))
    cols = [c for c in range(3) if np.abs(V[:, c]).sum() > 0]
    V[:, cols] += 0.05 * rng.standard_normal((n, 2))
    st, zd, ref = O.normals2d(V, E)
    assert st == 0
    got = cpg.compute_normals_2d(V, E)
    np.testing.assert_array_equal(got, ref)


def test_package_normals_2d_quirks_match_oracle(capsys):
    V, E, _ = ellipse(24, 1.0, 1.0, 0.0, 0.0)
    np.testing.assert_array_equal(cpg.compute_normals_2d(V, E), O.normals2d(V, E)[2])
    V, E, _ = ellipse(10, 1.0, 1.0, 2.0, 2.0)
    assert cpg.compute_normals_2d(V, E[:9]).shape == (1, 1)
    assert cpg.compute_normals_2d(V + np.array([0.0, 0.0, 1.0]), E).shape == (1, 1)
    out = capsys.readouterr().out
    assert "V.rows() != E.rows()" in out and "xy, xz or yz plane" in out


@pytest.mark.parametrize("seed", range(4))
def test_package_normals_3d_match_oracle(seed):
    p = S.make_pair(1 + seed % 3, 16, seed=seed)
    V, F = p["VX"], p["F"]
    np.testing.assert_array_equal(cpg.compute_normals_3d(V, F), O.normals3d(V, F))


def test_package_normals_3d_corner_cases_match_oracle():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 2], [2, 0, 0], [5, 5, 5]], np.float64)
    F = np.array([[0, 1, 2], [0, 3, 4], [0, 1, 4]], np.int32)
    got, ref = cpg.compute_normals_3d(V, F), O.normals3d(V, F)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(got[:5], ref[:5])
