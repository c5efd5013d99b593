"""Synthetic scene generation for the FLIP step.

Every workload in this repository is seeded the same way on the CPU and the
GPU so results can be compared particle by particle:

* marker particles are placed 8 per fluid cell at the sub-cell centres
  ``c +/- dx/4`` in the order the reference uses for its sub-cell offsets
  (src/engine/fluidsimulation.cpp:2723-2732);
* each coordinate gets a uniform jitter in ``[-j, j)`` with
  ``j = 0.25 * (f - 1e-3) * dx`` and jitter factor ``f = 1``
  (fluidsimulation.cpp:2616-2626), drawn from SplitMix64 keyed by
  ``(seed, particle index)`` instead of the C library ``rand()`` the
  reference calls (fluidsimulation.h:1501-1503), so the stream is
  platform-independent;
* particles outside the domain's wall box -- the domain inset by
  ``1.5 dx + 5e-5`` per side (fluidsimulation.cpp:2834-2839, aabb.cpp:122-133)
  -- are rejected;
* initial velocities are zero.

The particles are then handed to the engine the way the reference's
``loadMarkerParticleData`` expects them (fluidsimulation.cpp:2488-2509): flat
``float32`` xyz triples.
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Sequence

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied elementwise to uint64 counters."""
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _MASK64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _MASK64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _MASK64
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, index: np.ndarray, stream: int) -> np.ndarray:
    """Uniform doubles in [0, 1) keyed by (seed, particle index, axis)."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0xD1B54A32D192ED03)) & _MASK64
        ctr = (index.astype(np.uint64) * np.uint64(4) + np.uint64(stream)) & _MASK64
        bits = splitmix64(splitmix64(ctr ^ key))
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# Sub-cell offsets in the reference's order (fluidsimulation.cpp:2723-2732).
_OFFSETS = np.array(
    [[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [1, 1, -1],
     [-1, -1, 1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]], dtype=np.float64)


def boundary_box(isize: int, jsize: int, ksize: int, dx: float):
    """Domain box inset by 1.5 dx + 5e-5 per side, as float32 min corner and
    double extents (FluidSimulation::_getBoundaryAABB, fluidsimulation.cpp:2834-2839;
    AABB::expand, aabb.cpp:122-128)."""
    eps = 1e-4
    v = -3.0 * dx - eps
    h = 0.5 * v
    pos = np.array([0.0 - h, 0.0 - h, 0.0 - h], dtype=np.float32)
    ext = np.array([isize * dx + v, jsize * dx + v, ksize * dx + v], dtype=np.float64)
    return pos, ext


def inside_box(p: np.ndarray, pos: np.ndarray, ext: np.ndarray) -> np.ndarray:
    """AABB::isPointInside (aabb.cpp:130-133): float32 lower test, double upper test."""
    lo = np.all(p >= pos[None, :], axis=1)
    hi = np.all(p.astype(np.float64) < (pos.astype(np.float64) + ext)[None, :], axis=1)
    return lo & hi


@dataclasses.dataclass
class Scene:
    isize: int
    jsize: int
    ksize: int
    dx: float
    positions: np.ndarray  # (N, 3) float32
    velocities: np.ndarray  # (N, 3) float32
    name: str = ""

    @property
    def num_particles(self) -> int:
        return int(self.positions.shape[0])


def seed_cells(isize: int, jsize: int, ksize: int, dx: float,
               inside: Callable[[np.ndarray, np.ndarray, np.ndarray], np.ndarray],
               seed: int, jitter_factor: float = 1.0, name: str = "") -> Scene:
    """Seeds 8 jittered particles in every cell whose centre satisfies
    ``inside(x, y, z)`` (cell centres in double precision)."""
    kk, jj, ii = np.meshgrid(np.arange(ksize), np.arange(jsize), np.arange(isize), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel(); kk = kk.ravel()
    cx = ii * dx + 0.5 * dx
    cy = jj * dx + 0.5 * dx
    cz = kk * dx + 0.5 * dx
    mask = inside(cx, cy, cz)
    ii, jj, kk = ii[mask], jj[mask], kk[mask]
    ncell = ii.shape[0]
    # Cell centre as the reference forms it: (float)i*dx + 0.5*dx, stored float32
    # (grid3d.h:106-109).
    centre = np.stack([
        (ii.astype(np.float32).astype(np.float64) * dx + 0.5 * dx),
        (jj.astype(np.float32).astype(np.float64) * dx + 0.5 * dx),
        (kk.astype(np.float32).astype(np.float64) * dx + 0.5 * dx)], axis=1).astype(np.float32)
    q = np.float32(0.25 * dx)
    p = (centre[:, None, :] + (_OFFSETS.astype(np.float32) * q)[None, :, :]).reshape(-1, 3)
    idx = np.arange(ncell * 8, dtype=np.uint64)
    jit = 0.25 * (jitter_factor - 1e-3) * dx
    out = np.empty_like(p)
    for ax in range(3):
        r = -jit + uniform01(seed, idx, ax) * (2.0 * jit)
        out[:, ax] = (p[:, ax].astype(np.float64) + r).astype(np.float32)
    bpos, bext = boundary_box(isize, jsize, ksize, dx)
    keep = inside_box(out, bpos, bext)
    pos = np.ascontiguousarray(out[keep])
    vel = np.zeros_like(pos)
    return Scene(isize, jsize, ksize, dx, pos, vel, name)


def dam_break(isize: int, jsize: int, ksize: int, dx: float, seed: int,
              box: Sequence[float] = (0.0, 0.0, 0.0, 0.4, 0.5, 1.0), jitter_factor: float = 1.0) -> Scene:
    """Single fluid cuboid [x0,x1]x[y0,y1]x[z0,z1] given as fractions of the
    domain extent (config: dam break, 64^3 at dx=1/64, cuboid
    [0,0.4]x[0,0.5]x[0,1])."""
    W, H, D = isize * dx, jsize * dx, ksize * dx
    x0, y0, z0, x1, y1, z1 = box

    def inside(x, y, z):
        return ((x >= x0 * W) & (x < x1 * W) & (y >= y0 * H) & (y < y1 * H)
                & (z >= z0 * D) & (z < z1 * D))

    return seed_cells(isize, jsize, ksize, dx, inside, seed, jitter_factor, name="dam_break")


def double_dam_break(isize: int, jsize: int, ksize: int, dx: float, seed: int,
                     width: float = 0.25, height: float = 0.6, jitter_factor: float = 1.0) -> Scene:
    """Two cuboids at opposite x ends, each ``width`` of the x extent wide and
    ``height`` of the y extent tall, full z (config: 256x128x128 mixed
    precision)."""
    W, H = isize * dx, jsize * dx

    def inside(x, y, z):
        return (((x < width * W) | (x >= (1.0 - width) * W)) & (y < height * H))

    return seed_cells(isize, jsize, ksize, dx, inside, seed, jitter_factor, name="double_dam_break")


def pool_drop(isize: int, jsize: int, ksize: int, dx: float, seed: int,
              depth: float = 0.3, centre=(0.5, 0.7, 0.5), radius: float = 0.15,
              jitter_factor: float = 1.0) -> Scene:
    """Pool filling the bottom ``depth`` of the domain plus a sphere (config:
    128^3 drop into pool)."""
    W, H, D = isize * dx, jsize * dx, ksize * dx
    cx, cy, cz = centre[0] * W, centre[1] * H, centre[2] * D
    r = radius * min(W, H, D)

    def inside(x, y, z):
        return (y < depth * H) | ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2 < r * r)

    return seed_cells(isize, jsize, ksize, dx, inside, seed, jitter_factor, name="pool_drop")


# ---------------------------------------------------------------------------
# Closed triangle meshes for mesh inflow sources (MeshFluidSource with a
# static mesh, src/engine/meshfluidsource.cpp).  Vertices are float32 xyz
# triples and triangles int32 index triples, as the reference's
# TriangleMesh holds them (trianglemesh.h:80-84).
# ---------------------------------------------------------------------------

def cylinder_mesh(x0: float, x1: float, centre_yz: Sequence[float], radius: float,
                  segments: int = 32, phase: float = 0.0):
    """Closed cylinder along +x from x0 to x1 around (y, z) = centre_yz: two
    rings of ``segments`` vertices, quad sides split in two triangles and a
    triangle fan per cap."""
    n = int(segments)
    th = phase + 2.0 * np.pi * np.arange(n) / n
    yc, zc = float(centre_yz[0]), float(centre_yz[1])
    ring = np.stack([yc + radius * np.cos(th), zc + radius * np.sin(th)], axis=1)
    v = np.zeros((2 * n + 2, 3), np.float64)
    v[:n, 0], v[:n, 1:] = x0, ring
    v[n:2 * n, 0], v[n:2 * n, 1:] = x1, ring
    v[2 * n] = (x0, yc, zc)
    v[2 * n + 1] = (x1, yc, zc)
    t = []
    for i in range(n):
        a, b = i, (i + 1) % n
        t += [(a, b, n + b), (a, n + b, n + a), (2 * n, b, a), (2 * n + 1, n + a, n + b)]
    return v.astype(np.float32), np.asarray(t, np.int32)


def box_mesh(lo: Sequence[float], hi: Sequence[float], rotation=None, centre=None):
    """Closed box [lo, hi] as 12 triangles, optionally rotated by a 3x3 matrix
    about ``centre`` (default: the box centre)."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], np.float64)
    v = lo + c * (hi - lo)
    if rotation is not None:
        ctr = (lo + hi) / 2 if centre is None else np.asarray(centre, np.float64)
        v = (v - ctr) @ np.asarray(rotation, np.float64).T + ctr
    # corner index = x + 2y + 4z
    t = [(0, 2, 3), (0, 3, 1), (4, 5, 7), (4, 7, 6), (0, 1, 5), (0, 5, 4),
         (2, 6, 7), (2, 7, 3), (0, 4, 6), (0, 6, 2), (1, 3, 7), (1, 7, 5)]
    return v.astype(np.float32), np.asarray(t, np.int32)


def rotation_matrix(seed: int) -> np.ndarray:
    """A rotation drawn from a seeded generator (for test meshes)."""
    rng = np.random.default_rng(seed)
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


@dataclasses.dataclass
class Inflow:
    """A mesh inflow source: closed mesh and fluid velocity."""
    vertices: np.ndarray
    triangles: np.ndarray
    velocity: tuple


def pool_jet(isize: int, jsize: int, ksize: int, dx: float, seed: int, depth: float = 0.25,
             radius: float = 0.05, speed: float = 2.0, height: float = 0.6,
             x_span=(0.05, 0.15), segments: int = 32, phase: float = 0.1):
    """Pool filling the bottom ``depth`` of the domain plus an inflow jet: a
    cylinder of ``radius`` (metres) along +x over ``x_span`` (fractions of
    the domain width) at ``height`` (fraction of the domain height) and mid
    depth, emitting fluid at ``speed`` m/s along +x."""
    pool = pool_drop(isize, jsize, ksize, dx, seed, depth=depth, radius=0.0)
    W, H, D = isize * dx, jsize * dx, ksize * dx
    v, t = cylinder_mesh(x_span[0] * W, x_span[1] * W, (height * H, 0.5 * D), radius, segments, phase)
    return pool, Inflow(v, t, (speed, 0.0, 0.0))
