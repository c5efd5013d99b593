"""The 250-particle per-cell cap of FluidSimulation::_removeMarkerParticles
(fluidsimulation.cpp:4324-4353) on the device against the oracle (which
tests/test_oracle_math.py pins: the first 250 by particle order stay).

The removal pass builds its per-cell order only when some cell is over the
cap; both sides of that switch are covered: 300 particles in one cell (250
kept, in order) and 240 (all kept)."""

import numpy as np
import pytest

import flipengine3d_amd as fa

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("count", [300, 240])
def test_per_cell_cap_against_oracle(gpu, oracle, count):
    n, dx = 16, 1.0 / 16
    rng = np.random.default_rng(0)
    c = np.array([8.5, 8.5, 8.5]) * dx
    pos = (c + (rng.random((count, 3)) - 0.5) * 0.9 * dx).astype(np.float32)
    vel = np.zeros_like(pos)
    sim = fa.FluidSimulation(n, n, n, dx, ctx=gpu, precision=fa.PCG_FP64)
    sim.loadMarkerParticleData(pos, vel)
    sim.initialize()
    try:
        sim.update(1.0 / 60)
        pg, _ = sim.particles()
    finally:
        sim.close()
    h = oracle.sim(n, n, n, dx)
    h.load_particles(pos, vel)
    h.initialize()
    h.update(1.0 / 60)
    po, _ = h.particles()
    assert pg.shape == po.shape == (min(count, 250), 3)
    assert float(np.abs(pg - po).max()) <= 1e-4 * dx
    assert float(np.abs(pg - pos[:pg.shape[0]]).max()) <= 0.5 * dx
