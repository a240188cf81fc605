"""Synthetic problem generators: meshes, rest data, seeded noise.

Every input the HIP library, the CPU oracle and the benchmark see is built
here, on the host, from an explicit counter-based PRNG (splitmix64) and an
explicit Box-Muller transform, so the same seed gives the same bits on every
machine.

Workloads (names used by bench.py and the tests):

* ``grid_param`` -- 2D parametrization: n x n vertices on [0,1]^2, each quad
  split into two triangles; the rest surface is the height field
  z = 0.1 sin(2 pi x) sin(2 pi y); initial UV is (x, y) plus uniform noise in
  +-0.2 h; symmetric Dirichlet per triangle (100 x 100 -> 19,602 triangles).
* ``twist_cube`` -- 3D deformation: unit cube of m^3 cells, each split into
  6 Kuhn tetrahedra (m = 60 -> 1,296,000 tets, 226,981 vertices); bottom face
  pinned at rest, top face pinned to its rest position rotated 60 degrees about
  the z axis through the face centre (quadratic penalty); initial x is rest
  plus Gaussian noise with sigma = 0.01 h; symmetric Dirichlet per tet.
* ``disk_batch`` -- many radius-clipped grid disks of seeded sizes.
* ``planar_test_mesh`` -- the 6-vertex / 4-face fixture of the reference's
  Newton test (tests/Meshes.hh:12-42).

Energy parameter layouts (shared with the library, include/tinyad_amd.h):
  symmetric Dirichlet triangle: [Mr^-1 (2x2 row-major), w]
  symmetric Dirichlet tet:      [Mr^-1 (3x3 row-major), w]
  position penalty:             [target (d), w]
  quadratic:                    [A (KxK row-major), b (K), c]
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

E_SYMDIR_TRI = 1
E_SYMDIR_TET = 2
E_PENALTY = 3
E_QUADRATIC = 4
# residual elements of vector functions (sum-of-squares / Gauss-Newton):
#   symmetric Dirichlet triangle residuals: [Mr^-1 (2x2 row-major), s] -> 8 outputs
#     s (J00, J01, J10, J11, Ji00, Ji01, Ji10, Ji11), J = M Mr^-1, Ji = Mr M^-1
#   pin residuals:                          [target (d), s] -> d outputs s (target - x)
R_SYMDIR_TRI = 101
R_PIN = 102

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n consecutive outputs of splitmix64 started at (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) ^ (np.uint64(stream) * np.uint64(0xD1B54A32D192ED03))
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Standard normal samples by Box-Muller (cosine branch only)."""
    u = uniform(seed, 2 * n, stream)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class Term:
    energy: int
    conn: np.ndarray      # (n_elem, N) int32
    params: np.ndarray    # (n_elem, n_params) float64

    @property
    def valence(self) -> int:
        return int(self.conn.shape[1])


@dataclass
class Problem:
    name: str
    d: int
    n_handles: int
    terms: list
    x0: np.ndarray                 # AoS, length d * n_handles
    info: dict = field(default_factory=dict)

    @property
    def n_vars(self) -> int:
        return self.d * self.n_handles

    @property
    def n_elements(self) -> int:
        return int(sum(t.conn.shape[0] for t in self.terms))


# ---------------------------------------------------------------------------
# Triangle meshes
# ---------------------------------------------------------------------------

def grid_triangles(n: int):
    """n x n vertices on [0,1]^2; quad (i, j) split along its (i,j)-(i+1,j+1) diagonal."""
    h = 1.0 / (n - 1)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    V = np.stack([ii.ravel() * h, jj.ravel() * h], axis=1)
    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    ci = ci.ravel(); cj = cj.ravel()
    v00 = cj * n + ci
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    F = np.empty((2 * ci.size, 3), dtype=np.int32)
    F[0::2] = np.stack([v00, v10, v11], axis=1)
    F[1::2] = np.stack([v00, v11, v01], axis=1)
    return V, F, h


def height_field(V2: np.ndarray) -> np.ndarray:
    z = 0.1 * np.sin(2 * math.pi * V2[:, 0]) * np.sin(2 * math.pi * V2[:, 1])
    return np.concatenate([V2, z[:, None]], axis=1)


def tri_rest_params(V3: np.ndarray, F: np.ndarray, weights: np.ndarray | None = None):
    """Per-triangle [Mr^-1 (2x2), w] in an orthonormal frame of the 3D rest triangle.

    e1 is along b - a, e2 is the in-plane unit vector orthogonal to e1 on c's
    side, so Mr = [[|b-a|, (c-a).e1], [0, (c-a).e2]] has positive determinant.
    Default weight: rest area / total rest area.
    """
    a, b, c = V3[F[:, 0]], V3[F[:, 1]], V3[F[:, 2]]
    eb, ec = b - a, c - a
    lb = np.linalg.norm(eb, axis=1)
    e1 = eb / lb[:, None]
    cx = np.einsum("ij,ij->i", ec, e1)
    perp = ec - cx[:, None] * e1
    cy = np.linalg.norm(perp, axis=1)
    Mr = np.zeros((F.shape[0], 2, 2))
    Mr[:, 0, 0] = lb
    Mr[:, 0, 1] = cx
    Mr[:, 1, 1] = cy
    R = np.linalg.inv(Mr)
    area = 0.5 * lb * cy
    w = area / area.sum() if weights is None else weights
    return np.concatenate([R.reshape(-1, 4), w[:, None]], axis=1), area


def planar_rest_params(V2: np.ndarray, F: np.ndarray, w: float):
    """[Mr^-1, w] with Mr = col_mat(b - a, c - a) taken straight from 2D rest positions."""
    a, b, c = V2[F[:, 0]], V2[F[:, 1]], V2[F[:, 2]]
    Mr = np.stack([b - a, c - a], axis=2)  # columns
    R = np.linalg.inv(Mr)
    return np.concatenate([R.reshape(-1, 4), np.full((F.shape[0], 1), w)], axis=1)


def grid_param(n: int = 100, seed: int = 1, noise: float = 0.2) -> Problem:
    """2D parametrization of the height-field grid (19,602 triangles at n = 100)."""
    V2, F, h = grid_triangles(n)
    V3 = height_field(V2)
    params, _ = tri_rest_params(V3, F)
    u = uniform(seed, V2.size, stream=1).reshape(V2.shape)
    x0 = (V2 + (2.0 * u - 1.0) * noise * h).ravel()
    return Problem(f"grid_param_{n}", 2, V2.shape[0], [Term(E_SYMDIR_TRI, F, params)], x0,
                   {"h": h, "seed": seed, "noise": noise})


def planar_test_mesh():
    """The reference Newton-test fixture: rest, stretched init, pins b and targets bc."""
    V_rest = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [1, 2]], dtype=np.float64)
    V_init = V_rest.copy()
    V_init[:, 0] *= 0.5
    V_init[:, 1] *= 0.25
    F = np.array([[0, 1, 2], [1, 3, 2], [2, 3, 5], [2, 5, 4]], dtype=np.int32)
    b = np.array([0, 4], dtype=np.int32)
    bc = np.array([[0.0, 0.0], [-2.0, 0.0]])
    return V_rest, V_init, F, b, bc


def newton_test_problem() -> Problem:
    """Symmetric Dirichlet / F.rows() plus unit-weight pins, as in the Newton test."""
    V_rest, V_init, F, b, bc = planar_test_mesh()
    tri = Term(E_SYMDIR_TRI, F, planar_rest_params(V_rest, F, 1.0 / F.shape[0]))
    pen = Term(E_PENALTY, b[:, None].copy(), np.concatenate([bc, np.ones((2, 1))], axis=1))
    return Problem("newton_test", 2, 6, [tri, pen], V_init.ravel().copy())


def gauss_newton_test_problem() -> Problem:
    """The residual form of newton_test_problem, as in the reference's
    Gauss-Newton test (tests/GaussNewtonTest.cc:30-74): 8 residuals per
    triangle scaled by 1 / sqrt(F.rows()) and 2 unit pin residuals per pinned
    vertex, so the sum of squares equals newton_test_problem's energy."""
    V_rest, V_init, F, b, bc = planar_test_mesh()
    p = planar_rest_params(V_rest, F, 1.0 / math.sqrt(F.shape[0]))
    tri = Term(R_SYMDIR_TRI, F, p)
    pin = Term(R_PIN, b[:, None].copy(), np.concatenate([bc, np.ones((2, 1))], axis=1))
    return Problem("gauss_newton_test", 2, 6, [tri, pin], V_init.ravel().copy())


def grid_residuals(n: int = 100, seed: int = 1, noise: float = 0.2, pins: int = 2) -> Problem:
    """grid_param's energy as residuals: per triangle s = sqrt(w) times the 8
    entries of J and J^-1, plus `pins` pinned corner vertices at their noisy
    start (s = 1)."""
    P = grid_param(n, seed, noise)
    t = P.terms[0]
    p = t.params.copy()
    p[:, 4] = np.sqrt(p[:, 4])
    corners = np.array([0, n - 1, n * n - n, n * n - 1][:pins], dtype=np.int32)
    x0 = P.x0.reshape(-1, 2)
    pin = Term(R_PIN, corners[:, None].copy(), np.concatenate([x0[corners], np.ones((pins, 1))], axis=1))
    return Problem(f"grid_residuals_{n}", 2, P.n_handles, [Term(R_SYMDIR_TRI, t.conn, p), pin], P.x0.copy(),
                   dict(P.info))


# ---------------------------------------------------------------------------
# Tetrahedral meshes
# ---------------------------------------------------------------------------

# Kuhn split of a cube cell: one tet per permutation of the axes, walking
# 000 -> 111 along the permutation; odd permutations get two vertices swapped
# so every tet is positively oriented.
_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PARITY = [1, -1, -1, 1, 1, -1]


def kuhn_cube(m: int):
    """Unit cube with m^3 cells, 6 m^3 positively oriented tets, (m+1)^3 vertices."""
    n = m + 1
    h = 1.0 / m
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    V = np.stack([ii.ravel() * h, jj.ravel() * h, kk.ravel() * h], axis=1)

    def vid(i, j, k):
        return i + n * (j + n * k)

    ck, cj, ci = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    ci = ci.ravel(); cj = cj.ravel(); ck = ck.ravel()
    T = np.empty((6 * ci.size, 4), dtype=np.int32)
    for t, (perm, par) in enumerate(zip(_PERMS, _PARITY)):
        off = np.zeros(3, dtype=np.int64)
        verts = [vid(ci, cj, ck)]
        for ax in perm:
            off[ax] += 1
            verts.append(vid(ci + off[0], cj + off[1], ck + off[2]))
        if par < 0:
            verts[1], verts[2] = verts[2], verts[1]
        T[t::6] = np.stack(verts, axis=1)
    return V, T, h


def tet_rest_params(V: np.ndarray, T: np.ndarray, weights: np.ndarray | None = None):
    """Per-tet [Mr^-1 (3x3), w], Mr = col_mat(b - a, c - a, d - a); w = volume / total."""
    a = V[T[:, 0]]
    Mr = np.stack([V[T[:, 1]] - a, V[T[:, 2]] - a, V[T[:, 3]] - a], axis=2)
    vol = np.linalg.det(Mr) / 6.0
    R = np.linalg.inv(Mr)
    w = vol / vol.sum() if weights is None else weights
    return np.concatenate([R.reshape(-1, 9), w[:, None]], axis=1), vol


def twist_cube(m: int = 60, seed: int = 7, angle_deg: float = 60.0, noise: float = 0.01,
               pin_weight: float = 1.0) -> Problem:
    """Twisted-cube deformation problem (1,296,000 tets at m = 60)."""
    V, T, h = kuhn_cube(m)
    params, _ = tet_rest_params(V, T)
    z = V[:, 2]
    bottom = np.nonzero(z < 0.5 * h)[0]
    top = np.nonzero(z > 1.0 - 0.5 * h)[0]
    th = math.radians(angle_deg)
    c, s = math.cos(th), math.sin(th)
    tgt_top = V[top].copy()
    px, py = V[top, 0] - 0.5, V[top, 1] - 0.5
    tgt_top[:, 0] = 0.5 + c * px - s * py
    tgt_top[:, 1] = 0.5 + s * px + c * py
    pins = np.concatenate([bottom, top]).astype(np.int32)
    targets = np.concatenate([V[bottom], tgt_top])
    pen = Term(E_PENALTY, pins[:, None].copy(),
               np.concatenate([targets, np.full((pins.size, 1), pin_weight)], axis=1))
    x0 = (V + noise * h * normal(seed, V.size, stream=2).reshape(V.shape)).ravel()
    return Problem(f"twist_cube_{m}", 3, V.shape[0], [Term(E_SYMDIR_TET, T, params), pen], x0,
                   {"h": h, "seed": seed, "angle_deg": angle_deg, "noise": noise,
                    "pin_weight": pin_weight, "rest": V.ravel().copy(), "n_pins": int(pins.size)})


# ---------------------------------------------------------------------------
# Disk batch
# ---------------------------------------------------------------------------

def clipped_disk(n: int, seed: int, noise: float = 0.2) -> Problem:
    """Height-field grid disk: vertices within (n-1)h/2 of the centre, triangles fully inside."""
    V2, F, h = grid_triangles(n)
    r = np.linalg.norm(V2 - 0.5, axis=1)
    keep = r <= 0.5 + 1e-12
    Fk = F[keep[F].all(axis=1)]
    used = np.zeros(V2.shape[0], dtype=bool)
    used[Fk.ravel()] = True
    remap = -np.ones(V2.shape[0], dtype=np.int64)
    remap[used] = np.arange(int(used.sum()))
    V2k = V2[used]
    Fk = remap[Fk].astype(np.int32)
    V3 = height_field(V2k)
    params, _ = tri_rest_params(V3, Fk)
    u = uniform(seed, V2k.size, stream=3).reshape(V2k.shape)
    x0 = (V2k + (2.0 * u - 1.0) * noise * h).ravel()
    return Problem(f"disk_{n}", 2, V2k.shape[0], [Term(E_SYMDIR_TRI, Fk, params)], x0, {"h": h, "n": n})


def disk_batch(count: int = 256, seed: int = 3, n_min: int = 20, n_max: int = 160):
    """`count` independent disks with grid sizes drawn uniformly from [n_min, n_max]."""
    ns = n_min + (splitmix64(seed, count, stream=4) % np.uint64(n_max - n_min + 1)).astype(np.int64)
    return [clipped_disk(int(n), seed * 1000003 + i) for i, n in enumerate(ns)]


def min_signed_volume(problem: Problem, x: np.ndarray | None = None) -> float:
    """Smallest det(M) over the symmetric-Dirichlet elements (positive = no inversions)."""
    x = problem.x0 if x is None else x
    X = x.reshape(-1, problem.d)
    out = math.inf
    for t in problem.terms:
        if t.energy in (E_SYMDIR_TRI, E_SYMDIR_TET):
            a = X[t.conn[:, 0]]
            M = np.stack([X[t.conn[:, k]] - a for k in range(1, t.valence)], axis=2)
            out = min(out, float(np.linalg.det(M).min()))
    return out


def concat(problems, name: str = "batch"):
    """Concatenates independent problems of equal d into one; returns it and
    the handle offsets of the pieces (the segment boundaries)."""
    d = problems[0].d
    offsets = [0]
    terms_by_energy = {}
    xs = []
    for P in problems:
        assert P.d == d
        for t in P.terms:
            terms_by_energy.setdefault((t.energy, t.valence), []).append((t.conn + offsets[-1], t.params))
        xs.append(P.x0)
        offsets.append(offsets[-1] + P.n_handles)
    terms = [Term(e, np.concatenate([c for c, _ in v]).astype(np.int32), np.concatenate([p for _, p in v]))
             for (e, _), v in terms_by_energy.items()]
    return Problem(name, d, offsets[-1], terms, np.concatenate(xs)), np.array(offsets, dtype=np.int64)
