"""The CPU oracle against the reference engine compiled from its own sources.

oracle/_ref/libflipref.so is the reference's src/engine built by
oracle/Makefile with a small C harness (oracle/ref_harness.cpp).  Each stage
of the FLIP step is run on both with identical inputs and compared bit for
bit; the whole update(dt) is compared over several frames with the reference
on one thread (its PCG dot products and removal are then in a fixed order).
Skipped where the reference build is absent (it is built only where the
reference sources are present).
"""

import numpy as np
import pytest

from tests.helpers import bits_equal, cell_histogram, dam, fast_particle_scene, ndiff, radius, random_velocities

SIZES = [(16, 16, 16), (23, 17, 20)]


@pytest.mark.parametrize("dims", SIZES + [(12, 31, 9)])
def test_domain_solid_and_weights(oracle, reference, dims):
    I, J, K = dims
    dx = 1.0 / max(dims)
    po, wo = oracle.domain_solid(I, J, K, dx)
    pr, wr = reference.domain_solid(I, J, K, dx)
    assert bits_equal(po, pr), f"solid phi ndiff={ndiff(po, pr)}"
    for a, b in zip(wo, wr):
        assert bits_equal(a, b)


@pytest.mark.parametrize("dims", SIZES)
def test_liquid_sdf(oracle, reference, dims):
    I, J, K = dims
    s = dam(I, J, K, seed=3)
    solid, _ = oracle.domain_solid(I, J, K, s.dx)
    for sp in (None, solid):
        a = oracle.liquid_sdf(I, J, K, s.dx, s.positions, radius(s.dx), sp)
        b = reference.liquid_sdf(I, J, K, s.dx, s.positions, radius(s.dx), sp)
        assert bits_equal(a, b), f"ndiff={ndiff(a, b)}"


@pytest.mark.parametrize("shuffled", [False, True])
@pytest.mark.parametrize("dims", SIZES)
def test_p2g(oracle, reference, dims, shuffled):
    # Shuffled: particle index order unrelated to position, so every face's
    # sum depends on summing in the reference's ascending-index order.
    I, J, K = dims
    s = dam(I, J, K, seed=5)
    pos = s.positions
    vel = random_velocities(s.num_particles, seed=1)
    if shuffled:
        perm = np.random.default_rng(11).permutation(s.num_particles)
        pos, vel = pos[perm], vel[perm]
    (mo, vo) = oracle.p2g(I, J, K, s.dx, pos, vel, radius(s.dx))
    (mr, vr) = reference.p2g(I, J, K, s.dx, pos, vel, radius(s.dx))
    for a, b in zip(mo + vo, mr + vr):
        assert bits_equal(a, b), f"ndiff={ndiff(a, b)}"


@pytest.mark.parametrize("layers", [1, 2, 7])
def test_extrapolation(oracle, reference, layers):
    I, J, K = 18, 14, 16
    s = dam(I, J, K, seed=2)
    vel = random_velocities(s.num_particles, seed=2)
    (mo, vo) = oracle.p2g(I, J, K, s.dx, s.positions, vel, radius(s.dx))
    for g, v in zip(mo, vo):
        assert bits_equal(oracle.extrapolate(g, v, layers), reference.extrapolate(g, v, layers))


def _projection_inputs(oracle, I, J, K, seed=4):
    s = dam(I, J, K, seed=seed)
    solid, weights = oracle.domain_solid(I, J, K, s.dx)
    liquid = oracle.liquid_sdf(I, J, K, s.dx, s.positions, radius(s.dx), solid)
    vel = random_velocities(s.num_particles, seed=seed, scale=0.3)
    mac, valid = oracle.p2g(I, J, K, s.dx, s.positions, vel, radius(s.dx))
    mac = [oracle.extrapolate(g, v, 7) for g, v in zip(mac, valid)]
    mac[1] = (mac[1] + np.float32(-9.81 / 60)).astype(np.float32)
    return s, solid, weights, liquid, mac, valid


@pytest.mark.parametrize("dims", SIZES)
def test_mic0(oracle, reference, dims):
    I, J, K = dims
    s, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    n = int(np.count_nonzero(liquid[1:-1, 1:-1, 1:-1] < 0))
    r = np.random.default_rng(3).standard_normal(n)
    io, zo = oracle.mic0_apply(I, J, K, s.dx, 1 / 60, liquid, weights, r)
    ir, zr = reference.mic0_apply(I, J, K, s.dx, 1 / 60, liquid, weights, r)
    assert bits_equal(io, ir)
    assert bits_equal(zo, zr)


@pytest.mark.parametrize("dims", SIZES)
def test_pressure_solve_reference_exact(oracle, reference, dims):
    I, J, K = dims
    s, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    reference.set_threads(1)
    a = oracle.pressure_solve(I, J, K, s.dx, 1 / 60, liquid, solid, weights, mac, valid, mode=0)
    b = reference.pressure_solve(I, J, K, s.dx, 1 / 60, liquid, solid, weights, mac, valid)
    assert a["ok"] == b["ok"] == 1
    assert a["iters"] == b["iters"]
    assert bits_equal(a["rhs"], b["rhs"])
    assert bits_equal(a["soln"], b["soln"])
    for x, y in zip(a["mac"] + a["valid"], b["mac"] + b["valid"]):
        assert bits_equal(x, y)


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_particle_advector(oracle, reference, order):
    I, J, K = 20, 18, 22
    s, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    rng = np.random.default_rng(order)
    pos = (rng.random((3000, 3)) * np.array([I, J, K]) * s.dx * 1.1 - 0.05 * s.dx).astype(np.float32)
    assert bits_equal(oracle.advect_particles(I, J, K, s.dx, mac, pos, 1 / 30, order),
                      reference.advect_particles(I, J, K, s.dx, mac, pos, 1 / 30, order))
    assert bits_equal(oracle.trilinear_velocity(I, J, K, s.dx, mac, pos),
                      reference.trilinear_velocity(I, J, K, s.dx, mac, pos))


@pytest.mark.parametrize("case", [
    # (dims, cuboid, frames, frame dt, substep bounds)
    ((20, 20, 20), (0.0, 0.0, 0.0, 0.4, 0.5, 1.0), 8, 1 / 30, (1, 6)),
    ((24, 16, 12), (0.0, 0.0, 0.0, 0.5, 0.7, 1.0), 5, 1 / 60, (2, 4)),
])
def test_whole_update(oracle, reference, case):
    """FluidSimulation::update over several frames: particle count, positions,
    velocities and substep counts agree bit for bit (reference on 1 thread)."""
    (I, J, K), box, frames, dt, (mn, mx) = case
    s = dam(I, J, K, seed=1, box=box)
    reference.set_threads(1)
    sims = []
    for lib in (oracle, reference):
        h = lib.sim(I, J, K, s.dx)
        h.add_body_force(0.0, -9.81, 0.0)
        h.set_substeps(mn, mx)
        h.load_particles(s.positions, s.velocities)
        h.initialize()
        sims.append(h)
    for f in range(frames):
        for h in sims:
            h.update(dt)
        a, b = sims
        assert a.substeps() == b.substeps(), f"frame {f}"
        assert a.num_particles == b.num_particles, f"frame {f}"
        pa, va = a.particles()
        pb, vb = b.particles()
        assert bits_equal(pa, pb), f"frame {f}: positions ndiff={ndiff(pa, pb)}"
        assert bits_equal(va, vb), f"frame {f}: velocities ndiff={ndiff(va, vb)}"
    for which in (0, 2, 3, 4):
        assert bits_equal(sims[0].float_grid(which), sims[1].float_grid(which))


def test_pressure_solve_pockets_reference_exact(oracle, reference):
    # Nonzero solid face velocities with an isolated multi-cell pocket and a
    # one-cell pocket: the solver's pocket conditioning, the solve and the
    # conditioned solid velocities agree bit for bit.
    from tests.helpers import pocket_inputs
    I, J, K = 20, 18, 17
    dx, solid, weights, liquid, mac, valid, svel, cavity, single = pocket_inputs(oracle, I, J, K)
    reference.set_threads(1)
    a = oracle.pressure_solve(I, J, K, dx, 1 / 60, liquid, solid, weights, mac, valid, mode=0, solid_vel=svel)
    b = reference.pressure_solve(I, J, K, dx, 1 / 60, liquid, solid, weights, mac, valid, solid_vel=svel)
    assert a["ok"] == b["ok"] == 1
    assert a["iters"] == b["iters"]
    assert bits_equal(a["rhs"], b["rhs"])
    assert bits_equal(a["soln"], b["soln"])
    for x, y in zip(a["solid_vel"] + a["mac"] + a["valid"], b["solid_vel"] + b["mac"] + b["valid"]):
        assert bits_equal(x, y)


def _domain_particles(s, offset, scale, seed=9):
    """The scene's particles in domain coordinates (p * scale + offset in
    float32, as getMarkerParticlePositionData writes them), plus points that
    fall outside the grid box once mapped back."""
    off = np.asarray(offset, np.float32)
    pd = (s.positions * np.float32(scale) + off).astype(np.float32)
    box = np.array([s.isize, s.jsize, s.ksize], np.float32) * np.float32(s.dx)
    outside = np.stack([off - np.float32(0.01), off + box * np.float32(scale) + np.float32(0.01),
                        off + np.float32(0.5 * scale) * box * np.array([1, 1, 3], np.float32)]).astype(np.float32)
    pos = np.concatenate([pd, outside]).astype(np.float32)
    vel = random_velocities(pos.shape[0], seed=seed, scale=0.3)
    return pos, vel


@pytest.mark.parametrize("offset,scale", [((0.3, -0.2, 0.15), 2.5), ((0.0, 0.0, 0.0), 1.0), ((-1.0, 4.0, 0.5), 0.37)])
def test_particle_io_domain_coordinates(oracle, reference, offset, scale):
    """setDomainOffset / setDomainScale, loadMarkerParticleData admitted at
    initialize (domain to grid coordinates, then the grid-box filter) and the
    ranged vec3 read-back, bit for bit after one frame; invalid ranges are
    rejected by both."""
    I, J, K = 18, 14, 12
    s = dam(I, J, K, seed=4)
    pos, vel = _domain_particles(s, offset, scale)
    reference.set_threads(1)
    sims = []
    for lib in (oracle, reference):
        h = lib.sim(I, J, K, s.dx)
        h.add_body_force(0.0, -9.81, 0.0)
        h.set_domain(offset, scale)
        h.load_particles(pos, vel)
        assert h.num_particles == 0  # queued until initialize
        h.initialize()
        sims.append(h)
    a, b = sims
    assert a.num_particles == b.num_particles
    assert a.num_particles <= s.num_particles  # the outside points were dropped
    for h in sims:
        h.update(1 / 60)
    n = a.num_particles
    assert n == b.num_particles
    for rng in ((0, n), (3, 17), (n - 5, n), (7, 7)):
        ra, rb = a.particle_range(*rng), b.particle_range(*rng)
        assert ra is not None and rb is not None
        assert bits_equal(ra[0], rb[0]) and bits_equal(ra[1], rb[1]), f"range {rng}"
    for bad in ((-1, 3), (0, n + 1), (5, 4)):
        assert a.particle_range(*bad) is None and b.particle_range(*bad) is None


def test_speed_removal_and_count_grid(oracle, reference):
    """Extreme-velocity removal (fluidsimulation.cpp:4298-4353): the oracle
    removes the same particles as the reference, frame by frame, bit for bit.
    Its count grid of the removal pass (fluidsimulation.cpp:4325-4343) is
    pinned by the reference's surviving particles: every counted particle
    survives except the ones the speed test removed in that pass."""
    I, J, K = 20, 20, 20
    s, vel, ids = fast_particle_scene(I, J, K, nfast=3, speed=40.0)
    # maxRemoval = min(int(0.0005 N), 35) must admit the fast particles.
    assert min(int(s.num_particles * 0.0005), 35) >= 3
    reference.set_threads(1)
    sims = []
    for lib in (oracle, reference):
        h = lib.sim(I, J, K, s.dx)
        h.add_body_force(0.0, -9.81, 0.0)
        h.load_particles(s.positions, vel)
        h.initialize()
        sims.append(h)
    a, b = sims
    n0 = s.num_particles
    for f in range(3):
        for h in sims:
            h.update(1 / 30)
        assert a.num_particles == b.num_particles, f"frame {f}"
        pa, va = a.particles()
        pb, vb = b.particles()
        assert bits_equal(pa, pb) and bits_equal(va, vb), f"frame {f}"
        counts, last, total = a.count_grid()
        hist = cell_histogram(pb, I, J, K, s.dx)
        assert counts.max() <= 250
        assert np.all(counts >= hist), f"frame {f}"
        assert int(counts.sum() - hist.sum()) == last, f"frame {f}"
        if last == 0:
            assert np.array_equal(counts, hist), f"frame {f}"
    # The speed test did run: the three fast particles are gone and were the
    # only speed removals; nothing else left (no solid or cap removals here).
    assert total == 3
    assert a.num_particles == n0 - 3
    fast = s.positions[ids]
    assert not np.any(np.all(np.isin(pb, fast), axis=1))


def test_removal_count_stage(oracle, reference):
    """orc_removal_counts (the counting loop alone) against the reference's
    survivors: with no speed removal, the count grid of a removal pass is the
    cell histogram of the particles that survive it."""
    I, J, K = 18, 14, 16
    s = dam(I, J, K, seed=8)
    reference.set_threads(1)
    h = reference.sim(I, J, K, s.dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.load_particles(s.positions, s.velocities)
    h.initialize()
    h.update(1 / 30)
    pb, _ = h.particles()
    got = oracle.removal_counts(I, J, K, s.dx, pb)
    assert np.array_equal(got, cell_histogram(pb, I, J, K, s.dx))
    # Flagged particles are skipped and a cell stops at 250.
    rm = np.zeros(pb.shape[0], np.uint8)
    rm[::3] = 1
    assert np.array_equal(oracle.removal_counts(I, J, K, s.dx, pb, rm),
                          cell_histogram(pb[rm == 0], I, J, K, s.dx))
    crowd = np.repeat(pb[:1], 300, axis=0)
    c = oracle.removal_counts(I, J, K, s.dx, np.concatenate([crowd, pb[1:]]))
    g = np.floor(pb[0].astype(np.float64) / s.dx).astype(int)
    assert c[g[2], g[1], g[0]] == 250
