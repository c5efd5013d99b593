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
