"""Shared inputs for the parity tests (seeded, small enough for the oracle)."""

import numpy as np

from flipengine3d_amd import scenes


def radius(dx):
    # FluidSimulation::_initializeParticleRadii (fluidsimulation.cpp:2644-2650)
    return 0.5 * 1.0 * dx * np.sqrt(3.0)


def dam(I, J, K, seed=1, box=(0.0, 0.0, 0.0, 0.4, 0.5, 1.0)):
    dx = 1.0 / max(I, J, K)
    return scenes.dam_break(I, J, K, dx, seed=seed, box=box)


def random_velocities(n, seed=0, scale=1.0):
    return (np.random.default_rng(seed).standard_normal((n, 3)) * scale).astype(np.float32)


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def ndiff(a, b):
    return int(np.count_nonzero(np.asarray(a) != np.asarray(b)))


def pocket_inputs(oracle, I, J, K, seed=6):
    """Projection inputs with nonzero solid face velocities and fluid pockets.

    Besides a dam break (which borders air), the liquid field holds
      * a 3 x 3 x 3 cavity in the air region whose boundary faces all have
        weight 0: a multi-cell region that borders no air, so the pressure
        solver zeroes the solid velocities around it
        (pressuresolver.cpp:124-268);
      * a single fluid cell with closed faces and zero cell weight: a region of
        one cell, left alone by that rule (its row is empty).
    Returns (dx, solid_phi, weights, liquid, mac, valid, solid_vel, cavity, single).
    """
    s = dam(I, J, K, seed=seed)
    dx = s.dx
    solid, weights = oracle.domain_solid(I, J, K, dx)
    wu, wv, ww, wc = [w.copy() for w in weights]
    liquid = oracle.liquid_sdf(I, J, K, dx, s.positions, radius(dx), solid)
    i0, j0, k0 = int(0.65 * I), int(0.7 * J), K // 2 - 1
    cavity = (slice(k0, k0 + 3), slice(j0, j0 + 3), slice(i0, i0 + 3))
    liquid[cavity] = np.float32(-0.5 * dx)
    wu[k0:k0 + 3, j0:j0 + 3, i0] = 0.0
    wu[k0:k0 + 3, j0:j0 + 3, i0 + 3] = 0.0
    wv[k0:k0 + 3, j0, i0:i0 + 3] = 0.0
    wv[k0:k0 + 3, j0 + 3, i0:i0 + 3] = 0.0
    ww[k0, j0:j0 + 3, i0:i0 + 3] = 0.0
    ww[k0 + 3, j0:j0 + 3, i0:i0 + 3] = 0.0
    si, sj, sk = int(0.85 * I), int(0.3 * J), K // 3
    liquid[sk, sj, si] = np.float32(-0.5 * dx)
    wc[sk, sj, si] = 0.0
    wu[sk, sj, si] = wu[sk, sj, si + 1] = 0.0
    wv[sk, sj, si] = wv[sk, sj + 1, si] = 0.0
    ww[sk, sj, si] = ww[sk + 1, sj, si] = 0.0
    vel = random_velocities(s.num_particles, seed=seed, scale=0.3)
    mac, valid = oracle.p2g(I, J, K, dx, s.positions, vel, radius(dx))
    mac = [oracle.extrapolate(g, v, 7) for g, v in zip(mac, valid)]
    rng = np.random.default_rng(seed)
    solid_vel = tuple((rng.standard_normal(g.shape) * 0.2).astype(np.float32) for g in mac)
    return dx, solid, (wu, wv, ww, wc), liquid, mac, valid, solid_vel, cavity, (sk, sj, si)


def fast_particle_scene(I, J, K, nfast, speed, seed=7):
    """A dam break in which a few particles start far faster than
    CFL * dx / dt, so _getMarkerParticleSpeedLimit (fluidsimulation.cpp:
    4298-4322) sets a limit that removes them."""
    s = dam(I, J, K, seed=seed)
    vel = s.velocities.copy()
    ids = np.linspace(17, s.num_particles - 17, nfast).astype(np.int64)
    vel[ids] = np.array([speed, 0.3 * speed, -0.2 * speed], np.float32)
    return s, vel, ids


def cell_histogram(pos, I, J, K, dx):
    g = np.floor(pos.astype(np.float64) * (1.0 / dx)).astype(np.int64)
    h = np.zeros((K, J, I), np.int64)
    np.add.at(h, (g[:, 2], g[:, 1], g[:, 0]), 1)
    return h
