"""Seeded synthetic shape pairs for the solver workloads.

A pair is a triangle mesh X (an icosphere with radial noise, both orientations
of every edge in lexicographic order, as the reference requires of EX,
README.md:51) and a closed planar curve Y (ordered edges [i, i+1 mod n],
README.md:53) with a dense feature-difference matrix feat_diff[|VX|, |VY|].

Construction:
  * X: icosphere at subdivision s on the unit sphere, each vertex scaled
    radially by 1 + N(0, sigma_radial).
  * Y: n points evenly spaced in angle on the unit circle z=0; the planted X
    vertex of point i is the icosphere vertex nearest to it; the Y point is the
    xy-projection of that (noisy) X vertex plus N(0, 0.01) in x and y, z = 0.
  * descriptors: each X vertex N(0,1)^32; each Y point its planted vertex's
    descriptor + N(0, 0.1^2); feat_diff = Euclidean descriptor distance (fp32).
"""
from __future__ import annotations

import numpy as np
from scipy.spatial.distance import cdist


def icosphere(subdiv: int):
    """Unit icosphere vertices (float64 [n,3]) and triangles (int32 [m,3])."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0),
         (0, -1, phi), (0, 1, phi), (0, -1, -phi), (0, 1, -phi),
         (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdiv):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return np.array(verts), np.array(faces, dtype=np.int32)


def directed_edges(F: np.ndarray) -> np.ndarray:
    """Both orientations of every mesh edge, sorted lexicographically (int32 [2m,2])."""
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    e = np.concatenate([e, e[:, ::-1]], axis=0)
    e = np.unique(e, axis=0)
    return e.astype(np.int32)


def vertex_normals_area_weighted(V: np.ndarray, F: np.ndarray) -> np.ndarray:
    fn = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])  # length = 2*area
    n = np.zeros_like(V)
    for k in range(3):
        np.add.at(n, F[:, k], fn)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n


def curve_normals_2d(VY: np.ndarray) -> np.ndarray:
    """In-plane normals (nx, ny, 0) of a closed curve from neighbouring points."""
    t = np.roll(VY, -1, axis=0) - np.roll(VY, 1, axis=0)
    n = np.stack([t[:, 1], -t[:, 0], np.zeros(len(VY))], axis=1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n


def bipyramid(n_ring: int):
    """Unit bipyramid over an n_ring-gon: two apexes (valence n_ring) and an
    equator of valence-4 vertices; vertices (float64 [n+2,3]), triangles."""
    ang = 2.0 * np.pi * np.arange(n_ring) / n_ring
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(n_ring)], axis=1)
    V = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]], axis=0)
    top, bot = 0, n_ring + 1
    F = []
    for i in range(n_ring):
        a, b = 1 + i, 1 + (i + 1) % n_ring
        F += [(top, a, b), (bot, b, a)]
    return V, np.array(F, dtype=np.int32)


def make_pair(subdiv: int = 3, n_y: int = 64, seed: int = 0, dim: int = 32,
              sigma_radial: float = 0.05, sigma_curve: float = 0.01, sigma_desc: float = 0.1,
              with_normals: bool = False, mesh=None):
    """mesh: (vertices, triangles) of X instead of the icosphere of `subdiv`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    V0, F = icosphere(subdiv) if mesh is None else mesh
    VX = V0 * (1.0 + sigma_radial * rng.standard_normal(len(V0)))[:, None]
    EX = directed_edges(F)
    ang = 2.0 * np.pi * np.arange(n_y) / n_y
    targets = np.stack([np.cos(ang), np.sin(ang), np.zeros(n_y)], axis=1)
    planted = np.argmin(((V0[None, :, :] - targets[:, None, :]) ** 2).sum(-1), axis=1)
    VY = np.zeros((n_y, 3))
    VY[:, :2] = VX[planted, :2] + sigma_curve * rng.standard_normal((n_y, 2))
    EY = np.stack([np.arange(n_y), (np.arange(n_y) + 1) % n_y], axis=1).astype(np.int32)
    dX = rng.standard_normal((len(VX), dim))
    dY = dX[planted] + sigma_desc * rng.standard_normal((n_y, dim))
    FD = cdist(dX, dY).astype(np.float32)
    out = dict(VX=VX, EX=EX, VY=VY, EY=EY, feat_diff=FD, planted=planted, F=F)
    if with_normals:
        out["NX"] = vertex_normals_area_weighted(VX, F).astype(np.float32)
        out["NY"] = curve_normals_2d(VY).astype(np.float32)
    return out


def tiny_pair(n_x: int = 4, n_y: int = 4, seed: int = 0, ring: bool = True):
    """A tiny X graph (a ring or a tetrahedron-like complete graph) for brute force."""
    rng = np.random.Generator(np.random.PCG64(seed))
    VX = rng.standard_normal((n_x, 3))
    if ring:
        e = [(i, (i + 1) % n_x) for i in range(n_x)]
    else:
        e = [(i, j) for i in range(n_x) for j in range(i + 1, n_x)]
    e = e + [(b, a) for a, b in e]
    EX = np.unique(np.array(e, dtype=np.int32), axis=0)
    VY = np.zeros((n_y, 3))
    ang = 2 * np.pi * np.arange(n_y) / n_y
    VY[:, 0], VY[:, 1] = np.cos(ang), np.sin(ang)
    EY = np.stack([np.arange(n_y), (np.arange(n_y) + 1) % n_y], axis=1).astype(np.int32)
    FD = rng.uniform(0.1, 2.0, (n_x, n_y)).astype(np.float32)
    return dict(VX=VX, EX=EX, VY=VY, EY=EY, feat_diff=FD)


def make_uneven_pairs(n_pairs: int = 256, seed: int = 0, subdivisions=(2, 3, 4), n_y_range=(32, 160),
                      depths=(1, 2, 3)):
    """Serving-scale batch of shape pairs of uneven size: icosphere subdivision
    drawn uniformly from `subdivisions`, |VY| uniform in n_y_range (inclusive),
    max_depth uniform in `depths`, each pair from make_pair's generator with its
    own seed. Returns a list of (pair, max_depth)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_pairs):
        sub = int(rng.choice(subdivisions))
        ny = int(rng.integers(n_y_range[0], n_y_range[1] + 1))
        K = int(rng.choice(depths))
        out.append((make_pair(sub, ny, seed=int(rng.integers(1 << 31))), K))
    return out
