"""Particle I/O on the device: domain coordinates (setDomainOffset /
setDomainScale), loads admitted at initialize and the ranged vec3 read-back
(fluidsimulation.cpp:303-330, 2314-2338, 2488-2509, 2773-2785).

The bar: right after initialize the admitted particles read back bit-exact
against the oracle (the mapping is host arithmetic); after a frame,
positions within 1e-4 dx times the domain scale and velocities within 1e-4
of the largest speed, the whole-step bar.  Invalid ranges raise, as the
reference throws domain_error.
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from tests.helpers import bits_equal, dam, random_velocities

pytestmark = pytest.mark.gpu

CASES = [((0.3, -0.2, 0.15), 2.5), ((-1.0, 4.0, 0.5), 0.37)]


def _domain_particles(s, offset, scale):
    off = np.asarray(offset, np.float32)
    pd = (s.positions * np.float32(scale) + off).astype(np.float32)
    box = np.array([s.isize, s.jsize, s.ksize], np.float32) * np.float32(s.dx)
    outside = np.stack([off - np.float32(0.01), off + box * np.float32(scale) + np.float32(0.01)]).astype(np.float32)
    pos = np.concatenate([pd, outside]).astype(np.float32)
    return pos, random_velocities(pos.shape[0], seed=9, scale=0.3)


@pytest.mark.parametrize("offset,scale", CASES)
def test_domain_coordinates_against_oracle(gpu, oracle, offset, scale):
    I, J, K = 18, 14, 12
    s = dam(I, J, K, seed=4)
    pos, vel = _domain_particles(s, offset, scale)
    g = fa.FluidSimulation(I, J, K, s.dx, ctx=gpu, precision=fa.PCG_FP64)
    g.addBodyForce(0.0, -9.81, 0.0)
    g.setDomainOffset(*offset)
    g.setDomainScale(scale)
    g.loadMarkerParticleData(pos, vel)
    assert g.getNumMarkerParticles() == 0  # admitted at initialize
    g.initialize()
    o = oracle.sim(I, J, K, s.dx)
    o.add_body_force(0.0, -9.81, 0.0)
    o.set_domain(offset, scale)
    o.load_particles(pos, vel)
    o.initialize()
    n = g.getNumMarkerParticles()
    assert n == o.num_particles and n < pos.shape[0]
    for rng in ((0, n), (3, 17), (n - 5, n)):
        pg, vg = g.particles_range(*rng)
        po, vo = o.particle_range(*rng)
        assert bits_equal(pg, po) and bits_equal(vg, vo), f"range {rng}"
    allp, allv = g.particles()
    assert bits_equal(allp, o.particle_range(0, n)[0])
    g.update(1 / 60)
    o.update(1 / 60)
    n = g.getNumMarkerParticles()
    assert n == o.num_particles
    pg, vg = g.particles_range(0, n)
    po, vo = o.particle_range(0, n)
    assert np.abs(pg - po).max() <= 1e-4 * s.dx * scale
    assert np.abs(vg - vo).max() <= 1e-4 * max(1.0, float(np.abs(vo).max()))
    g.close()


def test_invalid_ranges_raise(gpu):
    s = dam(12, 12, 12, seed=1)
    g = fa.FluidSimulation(12, 12, 12, s.dx, ctx=gpu)
    g.loadMarkerParticleData(s.positions, s.velocities)
    with pytest.raises(Exception):
        g.particles_range(0, 1)  # not initialized
    g.initialize()
    n = g.getNumMarkerParticles()
    assert n == s.num_particles
    for bad in ((-1, 3), (0, n + 1), (5, 4)):
        with pytest.raises(ValueError):
            g.particles_range(*bad)
    p, v = g.particles_range(4, 4)
    assert p.shape == (0, 3) and v.shape == (0, 3)
    g.close()
