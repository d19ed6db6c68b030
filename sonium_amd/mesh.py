"""Surface meshes for the dense Burton–Miller path (Tri3, P0 collocation).

These generators produce the (nodes, triangles) inputs the assembly consumes.
They restate the reference's mesh builders so that the panel ordering and the
node coordinates are reproducible:

* ``icosphere`` — crates/sonium-bem/src/core/mesh/generators.rs:309-397
  (``generate_icosphere_mesh``) with ``get_midpoint`` (:400-427): same base
  icosahedron, same face split order ``[v0,m01,m20], [v1,m12,m01],
  [v2,m20,m12], [m01,m12,m20]``, same midpoint numbering by first encounter.
* ``fibonacci_sphere`` — generators.rs:35-77 (``generate_sphere_mesh``):
  Fibonacci points on the sphere and their convex hull.  The hull is taken
  with scipy's Qhull; the triangulation therefore need not equal the
  reference's own hull bit for bit, which does not matter for parity because
  the HIP path and the oracle are always compared on the same mesh.
* ``ellipsoid`` — an icosphere whose nodes are scaled by the semi-axes.

Every face is wound so that its vertex-order normal points outward, so the
normal used for field panels (vertex order, regular.rs:636-714) equals the
flipped element normal used for collocation points (generators.rs:852-866).
"""

from __future__ import annotations

import math

import numpy as np

_PHI = (1.0 + math.sqrt(5.0)) / 2.0

_ICO_VERTICES = [
    [-1.0, _PHI, 0.0],
    [1.0, _PHI, 0.0],
    [-1.0, -_PHI, 0.0],
    [1.0, -_PHI, 0.0],
    [0.0, -1.0, _PHI],
    [0.0, 1.0, _PHI],
    [0.0, -1.0, -_PHI],
    [0.0, 1.0, -_PHI],
    [_PHI, 0.0, -1.0],
    [_PHI, 0.0, 1.0],
    [-_PHI, 0.0, -1.0],
    [-_PHI, 0.0, 1.0],
]

_ICO_FACES = [
    [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
    [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
    [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
    [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
]


class Mesh:
    """A closed Tri3 surface: ``nodes`` (V, 3) float64, ``triangles`` (N, 3) int64."""

    def __init__(self, nodes: np.ndarray, triangles: np.ndarray):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int64)
        if self.nodes.ndim != 2 or self.nodes.shape[1] != 3:
            raise ValueError("nodes must be (V, 3)")
        if self.triangles.ndim != 2 or self.triangles.shape[1] != 3:
            raise ValueError("triangles must be (N, 3)")
        if self.triangles.size and (self.triangles.min() < 0 or self.triangles.max() >= len(self.nodes)):
            raise ValueError("triangle index out of range")

    @property
    def num_elements(self) -> int:
        return int(self.triangles.shape[0])

    def element_geometry(self):
        """Centroid, unit outward normal and area per panel.

        Restates compute_element_geometry (generators.rs:779-870): centre is
        the vertex mean, area half the cross-product norm, normal the
        normalised cross product flipped when n.c < 0.
        """
        p = self.nodes[self.triangles]
        center = ((p[:, 0] + p[:, 1]) + p[:, 2]) / 3.0
        v1 = p[:, 1] - p[:, 0]
        v2 = p[:, 2] - p[:, 0]
        cr = np.stack(
            [
                v1[:, 1] * v2[:, 2] - v1[:, 2] * v2[:, 1],
                v1[:, 2] * v2[:, 0] - v1[:, 0] * v2[:, 2],
                v1[:, 0] * v2[:, 1] - v1[:, 1] * v2[:, 0],
            ],
            axis=1,
        )
        ln = np.sqrt((cr[:, 0] * cr[:, 0] + cr[:, 1] * cr[:, 1]) + cr[:, 2] * cr[:, 2])
        area = ln / 2.0
        normal = np.where(ln[:, None] > 1e-15, cr / np.where(ln > 1e-15, ln, 1.0)[:, None], 0.0)
        ndc = (normal[:, 0] * center[:, 0] + normal[:, 1] * center[:, 1]) + normal[:, 2] * center[:, 2]
        normal = np.where(ndc[:, None] < 0.0, -normal, normal)
        return center, normal, area


def icosphere(radius: float = 1.0, subdivisions: int = 3) -> Mesh:
    """Icosphere with 20*4**subdivisions panels (generators.rs:309-397)."""
    verts = []
    for v in _ICO_VERTICES:
        ln = math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
        verts.append([v[0] / ln, v[1] / ln, v[2] / ln])
    faces = [list(f) for f in _ICO_FACES]
    for _ in range(subdivisions):
        cache: dict = {}

        def mid(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            idx = cache.get(key)
            if idx is not None:
                return idx
            va, vb = verts[a], verts[b]
            m = [(va[0] + vb[0]) / 2.0, (va[1] + vb[1]) / 2.0, (va[2] + vb[2]) / 2.0]
            ln = math.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
            verts.append([m[0] / ln, m[1] / ln, m[2] / ln])
            cache[key] = len(verts) - 1
            return len(verts) - 1

        new_faces = []
        for v0, v1, v2 in faces:
            m01 = mid(v0, v1)
            m12 = mid(v1, v2)
            m20 = mid(v2, v0)
            new_faces.append([v0, m01, m20])
            new_faces.append([v1, m12, m01])
            new_faces.append([v2, m20, m12])
            new_faces.append([m01, m12, m20])
        faces = new_faces
    nodes = np.array(verts, dtype=np.float64) * radius
    return Mesh(nodes, np.array(faces, dtype=np.int64))


def _orient_outward(nodes: np.ndarray, tris: np.ndarray, center: np.ndarray) -> np.ndarray:
    p = nodes[tris]
    cr = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    c = p.mean(axis=1) - center
    flip = np.einsum("ij,ij->i", cr, c) < 0.0
    out = tris.copy()
    out[flip, 1], out[flip, 2] = tris[flip, 2], tris[flip, 1]
    return out


def fibonacci_sphere(radius: float, n_theta: int, n_phi: int) -> Mesh:
    """Convex hull of (n_theta-1)*n_phi+2 Fibonacci points (generators.rs:35-77).

    n_theta=65, n_phi=80 gives 10240 panels; n_theta=129, n_phi=160 gives 40960.
    """
    from scipy.spatial import ConvexHull

    if not (radius > 0.0 and math.isfinite(radius)):
        raise ValueError("radius must be positive")
    if n_theta < 3 or n_phi < 3:
        raise ValueError("n_theta and n_phi must be at least 3")
    count = (n_theta - 1) * n_phi + 2
    golden = math.pi * (3.0 - math.sqrt(5.0))
    idx = np.arange(count, dtype=np.float64)
    z = 1.0 - 2.0 * (idx + 0.5) / count
    radial = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    az = golden * idx
    pts = np.stack([radius * radial * np.cos(az), radius * radial * np.sin(az), radius * z], axis=1)
    hull = ConvexHull(pts)
    tris = _orient_outward(pts, hull.simplices.astype(np.int64), np.zeros(3))
    return Mesh(pts, tris)


def ellipsoid(semi_axes, subdivisions: int) -> Mesh:
    """Icosphere mapped onto an ellipsoid with the given semi-axes."""
    a = np.asarray(semi_axes, dtype=np.float64).reshape(3)
    if np.any(a <= 0.0):
        raise ValueError("semi-axes must be positive")
    base = icosphere(1.0, subdivisions)
    return Mesh(base.nodes * a[None, :], base.triangles)


def fibonacci_directions(count: int) -> np.ndarray:
    """Unit vectors on a Fibonacci sphere: z_i = 1-(2i+1)/count, golden-angle azimuth."""
    golden = math.pi * (3.0 - math.sqrt(5.0))
    i = np.arange(count, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / count
    r = np.sqrt(np.maximum(1.0 - z * z, 0.0))
    return np.stack([r * np.cos(golden * i), r * np.sin(golden * i), z], axis=1)


def uniform_directions(rng: np.random.Generator, count: int) -> np.ndarray:
    """count unit vectors uniform on the sphere (normalised Gaussian triples)."""
    d = rng.standard_normal((count, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def random_ellipsoid_batch(count: int = 256, seed: int = 20261202, levels=(2, 3, 4), k_range=(1.0, 8.0),
                           max_directions: int = 16):
    """Seeded batch of rigid ellipsoid scatterers of uneven sizes.

    Per instance, drawn in this order from ``numpy.random.Generator(PCG64(seed))``:
    semi-axes U[0.5, 1.5]^3, icosphere level from ``levels`` (320, 1280, 5120
    panels for 2, 3, 4), wavenumber U[k_range], number of plane-wave
    directions U{1..max_directions}, then the directions (uniform on the
    sphere).  Returns a list of dicts with keys mesh, semi_axes, level, k,
    directions.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        axes = rng.uniform(0.5, 1.5, size=3)
        level = int(rng.choice(np.asarray(levels)))
        k = float(rng.uniform(k_range[0], k_range[1]))
        ndir = int(rng.integers(1, max_directions + 1))
        dirs = uniform_directions(rng, ndir)
        out.append({"mesh": ellipsoid(axes, level), "semi_axes": axes, "level": level, "k": k, "directions": dirs})
    return out
