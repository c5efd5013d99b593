"""Surface reconstruction on the GPU against the oracle (which equals the
reference's getIsomesh bit for bit, tests/test_oracle_surface.py).

* flipmi_mesh_particles on fixed particle sets: vertices and triangles
  bit-exact, over subdivision levels 1-3, a partial compute chunk (liquid in
  a corner: solid nodes from trilinear samples), a solid obstacle, the
  optional post-processing and smoothing settings, particle scale, a domain
  transform, a lone particle, and no particles.
* A scene with reconstruction on: after every frame getIsomesh() equals the
  oracle's mesh of the scene's own particles and solid (one substep per
  frame, so the frame's particles are the ones meshed), bit for bit; with
  several substeps per frame the mesh is the one after the first substep
  (checked against the oracle simulation, whose particles differ from the
  device's only by the pressure solve's rounding).
* Settings the reference rejects, and the ones not built here, raise.
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from tests.helpers import bits_equal, dam

pytestmark = pytest.mark.gpu


def _check(gpu_mesh, orc_mesh, label):
    (vg, tg), (vo, to) = gpu_mesh, orc_mesh
    print(f"{label}: {vo.shape[0]} vertices, {to.shape[0]} triangles", flush=True)
    assert vg.shape == vo.shape and tg.shape == to.shape, (label, vg.shape, vo.shape, tg.shape, to.shape)
    assert bits_equal(vg, vo), label
    assert np.array_equal(tg, to), label


MESH_CASES = [
    # dims, liquid box, subdivisions, settings
    ((20, 20, 20), (0.0, 0.0, 0.0, 0.4, 0.5, 1.0), 1, {}),
    ((24, 16, 12), (0.0, 0.0, 0.0, 0.4, 0.5, 1.0), 2, {}),
    ((48, 40, 36), (0.0, 0.0, 0.0, 0.3, 0.4, 0.35), 1, {}),
    ((48, 40, 36), (0.0, 0.0, 0.0, 0.3, 0.4, 0.35), 2, {}),
    ((40, 32, 32), (0.2, 0.0, 0.3, 0.6, 0.5, 0.7), 3, dict(smoothing_value=0.3, smoothing_iterations=3)),
    ((20, 20, 20), (0.0, 0.0, 0.0, 0.4, 0.5, 1.0), 1,
     dict(remove_near_domain=1, remove_near_domain_distance=1, invert_contact_normals=1,
          obstacle_meshing_offset=0.5)),
    ((32, 32, 32), (0.0, 0.0, 0.0, 0.5, 0.5, 1.0), 1, dict(particle_scale=2.0, smoothing_iterations=0)),
    ((64, 48, 40), (0.0, 0.0, 0.0, 0.5, 0.6, 1.0), 1, {}),
]


@pytest.mark.parametrize("case", MESH_CASES)
def test_mesh_particles_bitexact(gpu, oracle, case):
    (I, J, K), box, sub, settings = case
    s = dam(I, J, K, seed=1, box=box)
    solid, _ = oracle.domain_solid(I, J, K, s.dx)
    orc_kw = dict(subdivisions=sub)
    names = dict(smoothing_value="smooth_value", smoothing_iterations="smooth_iterations",
                 remove_near_domain="remove_near_domain", remove_near_domain_distance="remove_distance",
                 invert_contact_normals="invert_contact", obstacle_meshing_offset="obstacle_offset",
                 particle_scale="particle_scale")
    for k, v in settings.items():
        orc_kw[names[k]] = v
    ref = oracle.surface_mesh(I, J, K, s.dx, s.positions, solid, **orc_kw)
    got = gpu.mesh_particles(I, J, K, s.dx, s.positions, solid, subdivisions=sub, **settings)
    assert ref[1].shape[0] > 100
    _check(got, ref, f"{(I, J, K)} sub {sub} {settings}")


def test_mesh_with_an_obstacle_and_a_domain_transform(gpu, oracle):
    # A solid block inside the liquid: the solid clamp of the field and of
    # the edge interpolation both act.
    I, J, K = 32, 24, 24
    s = dam(I, J, K, seed=2, box=(0.0, 0.0, 0.0, 0.6, 0.7, 1.0))
    solid, _ = oracle.domain_solid(I, J, K, s.dx)
    k, j, i = np.meshgrid(np.arange(K + 1), np.arange(J + 1), np.arange(I + 1), indexing="ij")
    box = np.maximum.reduce([np.abs(i * s.dx - 0.35) - 0.08, np.abs(j * s.dx - 0.2) - 0.1,
                             np.abs(k * s.dx - 0.5) - 0.12]).astype(np.float32)
    solid = np.minimum(solid, box).astype(np.float32)
    ref = oracle.surface_mesh(I, J, K, s.dx, s.positions, solid, domain_offset=(0.25, -1.0, 3.0), domain_scale=2.5)
    got = gpu.mesh_particles(I, J, K, s.dx, s.positions, solid, domain_offset=(0.25, -1.0, 3.0), domain_scale=2.5)
    _check(got, ref, "obstacle + domain transform")


def test_lone_particle_and_no_particles(gpu, oracle):
    I = J = K = 24
    dx = 1 / 24
    solid, _ = oracle.domain_solid(I, J, K, dx)
    p = np.array([[0.51, 0.47, 0.53]], np.float32)
    _check(gpu.mesh_particles(I, J, K, dx, p, solid), oracle.surface_mesh(I, J, K, dx, p, solid), "one particle")
    v, t = gpu.mesh_particles(I, J, K, dx, np.zeros((0, 3), np.float32), solid)
    assert v.shape == (0, 3) and t.shape == (0, 3)


def _scene(I, J, K, s, gpu, steps_per_frame):
    sim = fa.FluidSimulation(I, J, K, s.dx, ctx=gpu, precision=fa.PCG_FP64)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(steps_per_frame[0])
    sim.setMaxTimeStepsPerFrame(steps_per_frame[1])
    sim.loadMarkerParticleData(s.positions, s.velocities)
    sim.enableSurfaceReconstruction()
    sim.initialize()
    return sim


def test_scene_isomesh_is_the_mesh_of_its_particles(gpu, oracle):
    I, J, K = 32, 32, 32
    s = dam(I, J, K, seed=1)
    sim = _scene(I, J, K, s, gpu, (1, 1))
    v0, t0 = sim.getIsomesh()
    assert v0.shape == (0, 3) and t0.shape == (0, 3)
    for f in range(4):
        sim.update(1 / 120)
        pos, _ = sim.particles()
        solid = sim.grid(1)  # FLIPMI_GRID_SOLID_PHI
        _check(sim.getIsomesh(), oracle.surface_mesh(I, J, K, s.dx, pos, solid), f"scene frame {f}")
    sim.disableSurfaceReconstruction()
    before = sim.getIsomesh()
    sim.update(1 / 120)
    after = sim.getIsomesh()
    assert bits_equal(before[0], after[0]) and np.array_equal(before[1], after[1])  # the last mesh stays
    sim.close()


def test_scene_meshes_after_the_first_substep(gpu, oracle):
    I, J, K = 24, 24, 24
    s = dam(I, J, K, seed=1)
    sim = _scene(I, J, K, s, gpu, (3, 3))
    h = oracle.sim(I, J, K, s.dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(3, 3)
    h.load_particles(s.positions, s.velocities)
    h.set_surface(True)
    h.initialize()
    for f in range(2):
        sim.update(1 / 30)
        h.update(1 / 30)
        vg, tg = sim.getIsomesh()
        vo, to = h.isomesh()
        print(f"frame {f}: gpu {vg.shape[0]} / {tg.shape[0]}, oracle {vo.shape[0]} / {to.shape[0]}", flush=True)
        # Particles differ from the oracle's by the pressure solve's rounding
        # only, so the meshes nearly coincide; a mesh of the frame's end
        # state (two more substeps) would be far off.
        assert abs(vg.shape[0] - vo.shape[0]) <= max(4, vo.shape[0] // 200)
        if vg.shape == vo.shape and np.array_equal(tg, to):
            assert np.abs(vg - vo).max() <= 1e-3 * s.dx
        else:
            from scipy.spatial import cKDTree
            d, _ = cKDTree(vo.astype(np.float64)).query(vg.astype(np.float64))
            assert np.percentile(d, 99) <= 1e-2 * s.dx and d.max() <= 0.5 * s.dx
    sim.close()


def test_surface_settings_are_checked(gpu):
    sim = fa.FluidSimulation(16, 16, 16, 1 / 16, ctx=gpu)
    sim.loadMarkerParticleData(np.full((8, 3), 0.5, np.float32))
    for bad in (lambda: sim.setSurfaceSubdivisionLevel(0), lambda: sim.setNumPolygonizerSlices(0),
                lambda: sim.setSurfaceSmoothingIterations(-1), lambda: sim.setMarkerParticleScale(-1.0),
                lambda: sim.setMinPolyhedronTriangleCount(-1), lambda: sim.setRemoveSurfaceNearDomainDistance(-1),
                lambda: sim.setNumPolygonizerSlices(2), lambda: sim.setMinPolyhedronTriangleCount(3),
                lambda: sim.setMarkerParticleScale(50.0)):
        with pytest.raises(ValueError):  # FLIPMI_ERR_INVALID_ARG, the reference's domain_error
            bad()
    assert sim.surface.subdivisions == 1 and sim.surface.polygonizer_slices == 1
    sim.close()
