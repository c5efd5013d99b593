"""Static mesh obstacles on the device against the oracle (which
tests/test_oracle_obstacles.py pins against the reference).

* The solid SDF with obstacles and the weight grid are bit-exact against the
  oracle after the first substep, for a rotated box with a cylinder and for
  two rotated boxes (ray jitter keyed by the same seed on both sides).
* Disabling an obstacle gives back the wall-only solid, bit-exact;
  enabling it again gives the obstacle solid.
* Whole steps with obstacles: particle counts equal to the oracle over three
  frames, positions within 1e-4 dx, velocities within 1e-4 max(1, vmax).
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from flipengine3d_amd import scenes
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

I = J = K = 32
DX = 1.0 / 32
SETS = {
    "rotated_box_and_cylinder": [scenes.box_mesh((0.45, 0.2, 0.3), (0.65, 0.5, 0.7), scenes.rotation_matrix(4)),
                                 scenes.cylinder_mesh(0.2, 0.8, (0.75, 0.5), 0.1, segments=20, phase=0.3)],
    "two_rotated_boxes": [scenes.box_mesh((0.3, 0.25, 0.3), (0.5, 0.45, 0.55), scenes.rotation_matrix(7)),
                          scenes.box_mesh((0.6, 0.5, 0.4), (0.8, 0.7, 0.6), scenes.rotation_matrix(8))],
}
SEED = 31


def _gpu(gpu, meshes):
    sc = scenes.dam_break(I, J, K, DX, seed=3)
    sim = fa.FluidSimulation(I, J, K, DX, ctx=gpu, precision=fa.PCG_FP64)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.loadMarkerParticleData(sc.positions, sc.velocities)
    sim.setMarkerParticleJitter(1.0, False, SEED)
    ids = [sim.addMeshObstacle(v, t) for v, t in meshes]
    sim.initialize()
    return sim, ids


def _orc(oracle, meshes):
    sc = scenes.dam_break(I, J, K, DX, seed=3)
    h = oracle.sim(I, J, K, DX)
    h.add_body_force(0.0, -9.81, 0.0)
    h.load_particles(sc.positions, sc.velocities)
    h.set_jitter(1.0, False, SEED)
    for v, t in meshes:
        h.add_obstacle(v, t)
    h.initialize()
    return h


@pytest.mark.parametrize("name", sorted(SETS))
def test_obstacle_solid_and_weights_bitexact(gpu, oracle, name):
    sim, _ = _gpu(gpu, SETS[name])
    h = _orc(oracle, SETS[name])
    try:
        for frame in range(3):
            sim.update(1 / 30)
            h.update(1 / 30)
            if frame == 0:
                assert bits_equal(sim.grid(1), h.float_grid(1))
                for w in (5, 6, 7, 8):
                    assert bits_equal(sim.grid(w), h.float_grid(w)), w
            pg, vg = sim.particles()
            po, vo = h.particles()
            assert pg.shape == po.shape, frame
            assert float(np.abs(pg - po).max()) <= 1e-4 * DX
            assert float(np.abs(vg - vo).max()) <= 1e-4 * max(1.0, float(np.abs(vo).max()))
    finally:
        sim.close()


def test_obstacle_toggle(gpu, oracle):
    meshes = SETS["rotated_box_and_cylinder"]
    sim, ids = _gpu(gpu, meshes)
    plain, _ = _gpu(gpu, [])
    try:
        sim.update(1 / 60)
        plain.update(1 / 60)
        with_obstacles = sim.grid(1)
        for oid in ids:
            sim.setMeshObstacle(oid, False)
        sim.update(1 / 60)
        assert bits_equal(sim.grid(1), plain.grid(1))
        for oid in ids:
            sim.setMeshObstacle(oid, True)
        sim.update(1 / 60)
        assert bits_equal(sim.grid(1), with_obstacles)
    finally:
        sim.close()
        plain.close()


def test_inversed_obstacle_bitexact(gpu, oracle):
    """A container (inversed obstacle) with a rotated box inside: the solid
    SDF and weights bit-exact against the oracle, whole steps within the bar."""
    container = scenes.box_mesh((0.06, 0.04, 0.07), (0.93, 0.91, 0.95))
    inner = SETS["rotated_box_and_cylinder"][0]
    sc = scenes.dam_break(I, J, K, DX, seed=3)
    sim = fa.FluidSimulation(I, J, K, DX, ctx=gpu, precision=fa.PCG_FP64)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.loadMarkerParticleData(sc.positions, sc.velocities)
    sim.setMarkerParticleJitter(1.0, False, SEED)
    cid = sim.addMeshObstacle(*container)
    sim.addMeshObstacle(*inner)
    sim.setMeshObstacleInversed(cid, True)
    sim.initialize()
    h = oracle.sim(I, J, K, DX)
    h.add_body_force(0.0, -9.81, 0.0)
    h.load_particles(sc.positions, sc.velocities)
    h.set_jitter(1.0, False, SEED)
    h.add_obstacle(*container)
    h.add_obstacle(*inner)
    h.set_obstacle_inversed(0, True)
    h.initialize()
    try:
        for frame in range(3):
            sim.update(1 / 30)
            h.update(1 / 30)
            if frame == 0:
                assert bits_equal(sim.grid(1), h.float_grid(1))
                for w in (5, 6, 7, 8):
                    assert bits_equal(sim.grid(w), h.float_grid(w)), w
            pg, vg = sim.particles()
            po, vo = h.particles()
            assert pg.shape == po.shape, frame
            assert float(np.abs(pg - po).max()) <= 1e-4 * DX
            assert float(np.abs(vg - vo).max()) <= 1e-4 * max(1.0, float(np.abs(vo).max()))
    finally:
        sim.close()
