"""The C++ engine classes (include/flipmi_engine.hpp) against the oracle.

tests/cpp/wrapper_check drives FluidSimulation, PressureSolver,
VelocityAdvector, ParticleAdvector, ParticleLevelSet and MeshLevelSet with the
reference engine's signatures (pressuresolver.h:63-92, velocityadvector.h:62-75,
particleadvector.h:53-79, fluidsimulation.h:127-153, particlelevelset.h:66-68)
and writes raw arrays; each is held to the bar of the matching C-ABI parity
test: bit-exact for level sets, weights, P2G (values and valid flags),
particle advection and the surface mesh, pressure to 1e-6 relative, and the
whole step to float rounding after one substep.
"""

import os
import subprocess

import numpy as np
import pytest

from tests.helpers import bits_equal, dam, ndiff, radius, random_velocities

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "wrapper_check")


def _run(mode, d):
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} is missing: run flipengine3d_amd/build.py")
    r = subprocess.run([BIN, mode, str(d)], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr


def _dims(d, I, J, K, dx, dt, extra):
    np.array([I, J, K, dx, dt, extra], np.float64).tofile(d / "dims.f64")


def _read(d, name, shape, dtype):
    return np.fromfile(d / name, dtype).reshape(shape)


def _particles(d, pos, vel):
    np.concatenate([pos, vel], axis=1).astype(np.float32).tofile(d / "particles.f32")


def test_levelset_and_walls(tmp_path, oracle):
    I, J, K = 26, 21, 23
    s = dam(I, J, K, seed=3)
    dx = s.dx
    _dims(tmp_path, I, J, K, dx, 0.0, radius(dx))
    _particles(tmp_path, s.positions, s.velocities)
    _run("levelset", tmp_path)
    solid, weights = oracle.domain_solid(I, J, K, dx)
    assert bits_equal(_read(tmp_path, "solid.f32", solid.shape, np.float32), solid)
    for name, w in zip(("wu", "wv", "ww", "wc"), weights):
        assert bits_equal(_read(tmp_path, name + ".f32", w.shape, np.float32), w)
    raw = oracle.liquid_sdf(I, J, K, dx, s.positions, radius(dx))
    post = oracle.liquid_sdf(I, J, K, dx, s.positions, radius(dx), solid)
    assert bits_equal(_read(tmp_path, "phi_raw.f32", raw.shape, np.float32), raw)
    got = _read(tmp_path, "phi.f32", post.shape, np.float32)
    assert bits_equal(got, post), f"ndiff={ndiff(got, post)}"


def _projection_inputs(oracle, I, J, K, seed=4):
    s = dam(I, J, K, seed=seed)
    dx = s.dx
    solid, weights = oracle.domain_solid(I, J, K, dx)
    liquid = oracle.liquid_sdf(I, J, K, dx, s.positions, radius(dx), solid)
    vel = random_velocities(s.num_particles, seed=seed, scale=0.3)
    mac, valid = oracle.p2g(I, J, K, dx, s.positions, vel, radius(dx))
    mac = [oracle.extrapolate(g, v, 7) for g, v in zip(mac, valid)]
    mac[1] = (mac[1] + np.float32(-9.81 / 60)).astype(np.float32)
    return s, vel, dx, solid, weights, liquid, mac, valid


@pytest.mark.parametrize("precision", [0, 1])
def test_pressure_solver(tmp_path, oracle, precision):
    I, J, K = 24, 28, 21
    s, vel, dx, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    dt = 1 / 60
    _dims(tmp_path, I, J, K, dx, dt, precision)
    liquid.astype(np.float32).tofile(tmp_path / "phi.f32")
    for name, w in zip(("wu", "wv", "ww", "wc"), weights):
        w.astype(np.float32).tofile(tmp_path / (name + ".f32"))
    for name, a in zip(("u", "v", "w"), mac):
        a.astype(np.float32).tofile(tmp_path / (name + ".f32"))
    for name, a in zip(("vu", "vv", "vw"), valid):
        a.astype(np.uint8).tofile(tmp_path / (name + ".u8"))
    _run("pressure", tmp_path)
    status = (tmp_path / "status.txt").read_text().splitlines()
    assert status[0] == "1"
    assert status[1].startswith("Pressure Solver Iterations: ")
    assert status[2].startswith("Estimated Error: ")
    po = oracle.pressure_solve(I, J, K, dx, dt, liquid, solid, weights, mac, valid, mode=1 if precision else 0)
    assert po["ok"] == 1
    vmax = max(float(np.abs(g).max()) for g in po["mac"])
    for name, vname, a, va in zip(("u", "v", "w"), ("vu", "vv", "vw"), po["mac"], po["valid"]):
        got = _read(tmp_path, name + "_out.f32", a.shape, np.float32)
        assert np.abs(got.astype(np.float64) - a).max() <= 1e-5 * vmax
        assert bits_equal(_read(tmp_path, vname + "_out.u8", va.shape, np.uint8), va)


def test_velocity_advector(tmp_path, oracle):
    I, J, K = 22, 19, 25
    s = dam(I, J, K, seed=5)
    dx = s.dx
    vel = random_velocities(s.num_particles, seed=5, scale=0.5)
    _dims(tmp_path, I, J, K, dx, 0.0, radius(dx))
    _particles(tmp_path, s.positions, vel)
    _run("p2g", tmp_path)
    (mo, vo) = oracle.p2g(I, J, K, dx, s.positions, vel, radius(dx))
    for name, vname, a, va in zip(("u", "v", "w"), ("vu", "vv", "vw"), mo, vo):
        got = _read(tmp_path, name + ".f32", a.shape, np.float32)
        gv = _read(tmp_path, vname + ".u8", va.shape, np.uint8)
        assert bits_equal(gv, va), f"{vname}: ndiff={ndiff(gv, va)}"
        assert bits_equal(got, a), f"{name}: ndiff={ndiff(got, a)}"


def test_particle_advector(tmp_path, oracle):
    I, J, K = 20, 18, 22
    s, vel, dx, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    rng = np.random.default_rng(9)
    pos = (rng.random((3000, 3)) * np.array([I, J, K]) * dx * 1.1 - 0.05 * dx).astype(np.float32)
    dt = 1 / 30
    _dims(tmp_path, I, J, K, dx, dt, 0.0)
    for name, a in zip(("u", "v", "w"), mac):
        a.astype(np.float32).tofile(tmp_path / (name + ".f32"))
    pos.tofile(tmp_path / "points.f32")
    _run("advect", tmp_path)
    for order in (1, 2, 3, 4):
        want = oracle.advect_particles(I, J, K, dx, mac, pos, dt, order)
        assert bits_equal(_read(tmp_path, f"rk{order}.f32", want.shape, np.float32), want)
    want = oracle.trilinear_velocity(I, J, K, dx, mac, pos)
    assert bits_equal(_read(tmp_path, "vel.f32", want.shape, np.float32), want)
    assert bits_equal(_read(tmp_path, "vel_inplace.f32", want.shape, np.float32), want)


def test_fluid_simulation(tmp_path, oracle):
    s = dam(24, 24, 24, seed=1)
    dx = s.dx
    dt = 1 / 60
    frames = 3
    _dims(tmp_path, s.isize, s.jsize, s.ksize, dx, dt, frames)
    s.positions.astype(np.float32).tofile(tmp_path / "positions.f32")
    s.velocities.astype(np.float32).tofile(tmp_path / "velocities.f32")
    _run("sim", tmp_path)
    h = oracle.sim(s.isize, s.jsize, s.ksize, dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(1, 1)
    h.load_particles(s.positions, s.velocities)
    h.initialize()
    rows = [line.split() for line in (tmp_path / "stats.txt").read_text().splitlines()]
    assert [int(r[0]) for r in rows] == list(range(frames))
    for r in rows:
        h.update(dt)
        assert int(r[1]) == 1 and int(r[4]) == 1
    po, vo = h.particles()
    pg = np.fromfile(tmp_path / "positions_out.f32", np.float32).reshape(-1, 3)
    vg = np.fromfile(tmp_path / "velocities_out.f32", np.float32).reshape(-1, 3)
    assert int(rows[-1][2]) == pg.shape[0]
    assert pg.shape == po.shape
    assert np.abs(pg - po).max() <= 1e-3 * dx
    assert np.abs(vg - vo).max() <= 1e-3 * max(1.0, float(np.abs(vo).max()))
    # getMarkerParticle{Position,Velocity}DataRange(2, n - 3): the same bytes.
    n = pg.shape[0]
    pr = np.fromfile(tmp_path / "positions_range.f32", np.float32).reshape(-1, 3)
    vr = np.fromfile(tmp_path / "velocities_range.f32", np.float32).reshape(-1, 3)
    assert bits_equal(pr, pg[2:n - 3]) and bits_equal(vr, vg[2:n - 3])


def test_obstacle_and_inflow_source(tmp_path, oracle):
    """MeshObject / addMeshObstacle and MeshFluidSource / addMeshFluidSource
    through the C++ classes, against the oracle with the same jitter seed."""
    from flipengine3d_amd import scenes
    I, J, K, dx = 48, 24, 24, 1 / 24
    pool, jet = scenes.pool_jet(I, J, K, dx, seed=5)
    obst = scenes.box_mesh((0.9, 0.15, 0.3), (1.2, 0.45, 0.7), scenes.rotation_matrix(2))
    frames, dt = 3, 1 / 60
    _dims(tmp_path, I, J, K, dx, dt, frames)
    pool.positions.astype(np.float32).tofile(tmp_path / "positions.f32")
    pool.velocities.astype(np.float32).tofile(tmp_path / "velocities.f32")
    obst[0].astype(np.float32).tofile(tmp_path / "obst_verts.f32")
    obst[1].astype(np.int32).tofile(tmp_path / "obst_tris.i32")
    jet.vertices.astype(np.float32).tofile(tmp_path / "src_verts.f32")
    jet.triangles.astype(np.int32).tofile(tmp_path / "src_tris.i32")
    np.array(jet.velocity, np.float64).tofile(tmp_path / "src_vel.f64")
    _run("sources", tmp_path)
    h = oracle.sim(I, J, K, dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(1, 6)
    h.load_particles(pool.positions, pool.velocities)
    h.set_jitter(1.0, False, 7)
    h.add_obstacle(*obst)
    h.add_inflow(jet.vertices, jet.triangles, jet.velocity)
    h.initialize()
    rows = [line.split() for line in (tmp_path / "stats.txt").read_text().splitlines()]
    for r in rows:
        h.update(dt)
        assert int(r[1]) == h.substeps() and int(r[2]) == h.num_particles
    po, vo = h.particles()
    pg = np.fromfile(tmp_path / "positions_out.f32", np.float32).reshape(-1, 3)
    vg = np.fromfile(tmp_path / "velocities_out.f32", np.float32).reshape(-1, 3)
    assert pg.shape == po.shape and pg.shape[0] > 0  # pool particles inside the obstacle are removed
    assert np.abs(pg - po).max() <= 1e-4 * dx
    assert np.abs(vg - vo).max() <= 1e-4 * max(1.0, float(np.abs(vo).max()))


def test_surface_through_the_engine_classes(tmp_path, oracle):
    """FluidSimulation::getIsomesh read as FluidManager reads it, and
    ParticleMesher::meshParticles, against the oracle."""
    s = dam(20, 20, 20, seed=1)
    dx = s.dx
    _dims(tmp_path, s.isize, s.jsize, s.ksize, dx, 1 / 60, 2)
    s.positions.astype(np.float32).tofile(tmp_path / "positions.f32")
    s.velocities.astype(np.float32).tofile(tmp_path / "velocities.f32")
    _run("surface", tmp_path)
    solid, _ = oracle.domain_solid(s.isize, s.jsize, s.ksize, dx)
    pos = np.fromfile(tmp_path / "positions_out.f32", np.float32).reshape(-1, 3)
    vo, to = oracle.surface_mesh(s.isize, s.jsize, s.ksize, dx, pos, solid, subdivisions=2)
    vg = np.fromfile(tmp_path / "iso_verts.f32", np.float32).reshape(-1, 3)
    tg = np.fromfile(tmp_path / "iso_tris.i32", np.int32).reshape(-1, 3)
    assert to.shape[0] > 100
    assert vg.shape == vo.shape and bits_equal(vg, vo)
    assert np.array_equal(tg, to)
    # Vertex normals as the viewer forms them: unit length, pointing out of
    # the liquid on the free surface (up, at the top of the column).
    nrm = np.fromfile(tmp_path / "iso_normals.f32", np.float32).reshape(-1, 3)
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    top = vg[:, 1] > vg[:, 1].max() - 0.5 * dx
    assert np.all(nrm[top, 1] > 0.0) and nrm[top, 1].mean() > 0.8
    vm, tm = oracle.surface_mesh(s.isize, s.jsize, s.ksize, dx, s.positions, solid, smooth_iterations=0)
    vgm = np.fromfile(tmp_path / "mesher_verts.f32", np.float32).reshape(-1, 3)
    tgm = np.fromfile(tmp_path / "mesher_tris.i32", np.int32).reshape(-1, 3)
    assert vgm.shape == vm.shape and bits_equal(vgm, vm)
    assert np.array_equal(tgm, tm)
