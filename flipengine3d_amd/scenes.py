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
