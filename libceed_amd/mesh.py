"""Synthetic structured hex meshes on the unit cube (the benchmark inputs).

``n`` elements per direction, element ``e = ex + n*(ey + n*ez)``, local node
``i + P*(j + P*k)`` (x fastest), global node numbering lexicographic on the
``(n*p + 1)^3`` grid, as the Cartesian restrictions of
examples/ceed/ex1-volume.c:356-400 build them.  The geometry is trilinear
(mesh basis P = 2) with interior vertices moved by a seeded uniform
perturbation of +-``perturb`` * h per coordinate; boundary vertices stay put,
so the domain is exactly the unit cube.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HexMesh:
    n: int                     # elements per direction
    p: int                     # solution polynomial degree (P = p + 1)
    num_elem: int
    num_nodes: int             # scalar L-vector length, (n*p+1)^3
    offsets: np.ndarray        # int32 [num_elem, P^3], solution restriction
    mesh_offsets: np.ndarray   # int32 [num_elem, 8], trilinear geometry restriction
    coords: np.ndarray         # float64 [3 * (n+1)^3], component stride (n+1)^3
    num_vertices: int

    @property
    def P(self) -> int:
        return self.p + 1


def _lex_offsets(n: int, p: int) -> np.ndarray:
    """Element offsets [n^3, (p+1)^3] of the lexicographic node numbering: element k = ex + n (ey + n ez) and local node
    i + P (j + P k), both x fastest, hold global node (ex p + i) + M ((ey p + j) + M (ez p + k)), M = n p + 1."""
    P, M = p + 1, n * p + 1
    k = np.arange(n ** 3, dtype=np.int64)
    base = p * ((k % n) + M * ((k // n) % n + M * (k // (n * n))))   # global node of each element's local node 0
    i = np.arange(P, dtype=np.int64)
    local = (i[None, None, :] + M * (i[None, :, None] + M * i[:, None, None])).reshape(-1)  # [z][y][x], x fastest
    out = np.empty((n ** 3, P ** 3), dtype=np.int32)
    np.add(base[:, None], local[None, :], out=out, casting="unsafe")
    return out


def hex_mesh(n: int, p: int, seed: int = 0, perturb: float = 0.1) -> HexMesh:
    """Structured n^3 hex mesh of degree p on [0,1]^3 with perturbed interior vertices."""
    if n < 1 or p < 1:
        raise ValueError("need n >= 1 and p >= 1")
    nv1 = n + 1
    h = 1.0 / n
    g = np.linspace(0.0, 1.0, nv1)
    # vertex ix + nv1 (iy + nv1 iz) at (g[ix], g[iy], g[iz])
    X, Y, Z = np.tile(g, nv1 * nv1), np.tile(np.repeat(g, nv1), nv1), np.repeat(g, nv1 * nv1)
    rng = np.random.default_rng(seed)
    d = rng.uniform(-1.0, 1.0, size=(3, nv1 ** 3)) * perturb * h
    inner = np.zeros(nv1, dtype=bool)
    inner[1:n] = True
    interior = np.tile(inner, nv1 * nv1) & np.tile(np.repeat(inner, nv1), nv1) & np.repeat(inner, nv1 * nv1)
    for a, dd in zip((X, Y, Z), d):
        a[interior] += dd[interior]
    coords = np.concatenate([X, Y, Z]).astype(np.float64)
    offsets = _lex_offsets(n, p)
    return HexMesh(n=n, p=p, num_elem=n ** 3, num_nodes=(n * p + 1) ** 3, offsets=offsets,
                   mesh_offsets=offsets if p == 1 else _lex_offsets(n, 1), coords=coords, num_vertices=nv1 ** 3)


def node_coordinates(mesh: HexMesh, gll_nodes: np.ndarray) -> np.ndarray:
    """Physical coordinates [3, num_nodes] of the solution nodes (trilinear map of the GLL points)."""
    P = mesh.P
    r = (np.asarray(gll_nodes) + 1.0) / 2.0          # reference [0, 1]
    rx, ry, rz = np.meshgrid(r, r, r, indexing="ij")
    rx, ry, rz = (a.transpose(2, 1, 0).reshape(-1) for a in (rx, ry, rz))  # local node x fastest
    # trilinear shape functions of the 8 vertices (local vertex a + 2*(b + 2*c))
    shp = np.empty((P ** 3, 8))
    for c in range(2):
        for b in range(2):
            for a in range(2):
                shp[:, a + 2 * (b + 2 * c)] = (rx if a else 1 - rx) * (ry if b else 1 - ry) * (rz if c else 1 - rz)
    nv = mesh.num_vertices
    out = np.empty((3, mesh.num_nodes))
    for d in range(3):
        vc = mesh.coords[d * nv:(d + 1) * nv][mesh.mesh_offsets]   # [E, 8]
        out[d, mesh.offsets.reshape(-1)] = (vc @ shp.T).reshape(-1)
    return out


def mesh_size_for_dofs(p: int, target_dofs: float) -> int:
    """Elements per direction giving about `target_dofs` scalar nodes at degree p."""
    return max(1, int(round((target_dofs ** (1.0 / 3.0) - 1.0) / p)))


def quadrature_points(mesh: HexMesh, q_ref: np.ndarray) -> np.ndarray:
    """Physical coordinates [3, num_elem * Q^3] of the tensor quadrature points (x fastest within an element)."""
    r = (np.asarray(q_ref) + 1.0) / 2.0
    rx, ry, rz = np.meshgrid(r, r, r, indexing="ij")
    rx, ry, rz = (a.transpose(2, 1, 0).reshape(-1) for a in (rx, ry, rz))
    shp = np.empty((rx.size, 8))
    for c in range(2):
        for b in range(2):
            for a in range(2):
                shp[:, a + 2 * (b + 2 * c)] = (rx if a else 1 - rx) * (ry if b else 1 - ry) * (rz if c else 1 - rz)
    nv = mesh.num_vertices
    return np.stack([(mesh.coords[d * nv:(d + 1) * nv][mesh.mesh_offsets] @ shp.T).reshape(-1) for d in range(3)])


def smooth_random_field(points: np.ndarray, seed: int = 0, modes: int = 16, amplitude: float = 0.5) -> np.ndarray:
    """g(x) = sum_k c_k cos(2 pi m_k . x + phi_k), m_k integer wave vectors in [-3, 3]^3, with
    sum |c_k| = amplitude so that |g| <= amplitude; evaluated at points [3, n]."""
    rng = np.random.default_rng(seed)
    m = rng.integers(-3, 4, size=(modes, 3))
    phi = rng.uniform(0.0, 2.0 * np.pi, modes)
    c = rng.uniform(-1.0, 1.0, modes)
    c *= amplitude / np.abs(c).sum()
    g = np.zeros(points.shape[1])
    for k in range(modes):
        g += c[k] * np.cos(2.0 * np.pi * (m[k] @ points) + phi[k])
    return g
