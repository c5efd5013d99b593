"""Mesh inflow sources on the device against the oracle (which
tests/test_oracle_inflow.py pins against the reference).

* Set-up: the source SDF on its sub-grid and the source cells are bit-exact
  against the oracle, for a box and for the large workload's jet cylinder
  at full size (512x256x256).
* Emission: from an empty domain the first substep only emits; the new
  particles (positions, velocities, order) are bit-exact against the
  oracle, for two sources with jitter factor 1 and surface jitter off/on.
* With a pool: the emitted particles of the first substep are bit-exact and
  carry the source velocity, the pool particles are within the whole-step
  bar (1e-4 dx, velocities 1e-4 of the largest speed); over three frames
  the particle counts match the oracle exactly.
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from flipengine3d_amd import scenes
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

BOX = scenes.box_mesh((0.31, 0.55, 0.36), (0.47, 0.71, 0.58))


def _gpu_sim(gpu, dims, dx, sources, pool=None, jitter=(1.0, False, 77), substeps=(1, 1), precision=fa.PCG_FP64):
    sim = fa.FluidSimulation(*dims, dx, ctx=gpu, precision=precision)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(substeps[0])
    sim.setMaxTimeStepsPerFrame(substeps[1])
    if pool is not None:
        sim.loadMarkerParticleData(pool.positions, pool.velocities)
    sim.setMarkerParticleJitter(*jitter)
    ids = [sim.addMeshFluidSource(v, t, vel) for v, t, vel in sources]
    sim.initialize()
    return sim, ids


def _orc_sim(oracle, dims, dx, sources, pool=None, jitter=(1.0, False, 77), substeps=(1, 1)):
    h = oracle.sim(*dims, dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(*substeps)
    if pool is not None:
        h.load_particles(pool.positions, pool.velocities)
    h.set_jitter(*jitter)
    for v, t, vel in sources:
        h.add_inflow(v, t, vel)
    h.initialize()
    return h


def _jet(I, J, K, dx):
    pool, jet = scenes.pool_jet(I, J, K, dx, seed=5)
    return pool, (jet.vertices, jet.triangles, jet.velocity)


@pytest.mark.parametrize("case", ["box_24", "jet_128", "jet_512"])
def test_source_setup_bitexact(gpu, oracle, case):
    if case == "box_24":
        dims, dx, src = (24, 24, 24), 1 / 24, (BOX[0], BOX[1], (1.5, 0.0, 0.5))
    else:
        I = int(case.split("_")[1])
        dims, dx = (I, I // 2, I // 2), 1.0 / (I // 2)
        src = _jet(*dims, dx)[1]
    sim, ids = _gpu_sim(gpu, dims, dx, [src])
    off, sdf, cells = sim.meshFluidSource(ids[0])
    sim.close()
    want = oracle.inflow_build(src[0], src[1], dims, dx, seed=77, sid=0)
    assert off == want[0]
    assert sdf.shape == want[1].shape and bits_equal(sdf, want[1])
    assert np.array_equal(cells, want[2]) and int(cells.sum()) > 0
    assert (sdf < 0).any()


@pytest.mark.parametrize("surface", [False, True])
def test_first_emission_bitexact(gpu, oracle, surface):
    dims, dx = (48, 24, 24), 1 / 24
    srcs = [_jet(*dims, dx)[1], (BOX[0] * np.float32([1.6, 1, 1]), BOX[1], (0.0, -0.5, 1.0))]
    jit = (1.0, surface, 91)
    sim, _ = _gpu_sim(gpu, dims, dx, srcs, jitter=jit)
    st = sim.update(1e-3)
    pg, vg = sim.particles()
    sim.close()
    h = _orc_sim(oracle, dims, dx, srcs, jitter=jit)
    h.update(1e-3)
    po, vo = h.particles()
    assert st.num_particles == po.shape[0] > 1000
    assert bits_equal(pg, po) and bits_equal(vg, vo)


def test_pool_and_jet_against_oracle(gpu, oracle):
    dims, dx = (64, 32, 32), 1 / 32
    pool, src = _jet(*dims, dx)
    n0 = pool.num_particles
    sim, _ = _gpu_sim(gpu, dims, dx, [src], pool=pool, substeps=(1, 6))
    h = _orc_sim(oracle, dims, dx, [src], pool=pool, substeps=(1, 6))
    try:
        for frame in range(3):
            st = sim.update(1 / 60)
            h.update(1 / 60)
            pg, vg = sim.particles()
            po, vo = h.particles()
            assert pg.shape == po.shape, (frame, pg.shape, po.shape)
            if frame == 0:
                assert st.substeps == 1 and pg.shape[0] > n0
                # the first emission is exact and carries the source velocity
                assert bits_equal(pg[n0:], po[n0:]) and bits_equal(vg[n0:], vo[n0:])
                assert np.all(vg[n0:] == np.float32(src[2]))
            assert float(np.abs(pg - po).max()) <= 1e-4 * dx
            assert float(np.abs(vg - vo).max()) <= 1e-4 * max(1.0, float(np.abs(vo).max()))
    finally:
        sim.close()


OUTFLOW_BOX = scenes.box_mesh((0.13, 0.07, 0.21), (0.36, 0.31, 0.79))


@pytest.mark.parametrize("inversed", [False, True])
def test_outflow_against_oracle(gpu, oracle, inversed):
    """An outflow source (setOutflow, outflowInverse) in a dam break: the
    particle counts after each of three frames equal the oracle's, positions
    within 1e-4 dx and velocities within 1e-4 max(1, vmax)."""
    dims, dx = (24, 24, 24), 1 / 24
    sc = scenes.dam_break(*dims, dx, seed=3)
    sim = fa.FluidSimulation(*dims, dx, ctx=gpu, precision=fa.PCG_FP64)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.loadMarkerParticleData(sc.positions, sc.velocities)
    sid = sim.addMeshFluidSource(OUTFLOW_BOX[0], OUTFLOW_BOX[1], (0.0, 0.0, 0.0))
    sim.setMeshFluidSourceOutflow(sid, True, inversed)
    sim.initialize()
    h = oracle.sim(*dims, dx)
    h.add_body_force(0.0, -9.81, 0.0)
    h.load_particles(sc.positions, sc.velocities)
    hid = h.add_inflow(OUTFLOW_BOX[0], OUTFLOW_BOX[1], (0.0, 0.0, 0.0))
    h.set_outflow(hid, True, inversed)
    h.initialize()
    try:
        for frame in range(3):
            sim.update(1 / 30)
            h.update(1 / 30)
            pg, vg = sim.particles()
            po, vo = h.particles()
            assert pg.shape == po.shape, (frame, pg.shape, po.shape)
            assert float(np.abs(pg - po).max()) <= 1e-4 * dx
            assert float(np.abs(vg - vo).max()) <= 1e-4 * max(1.0, float(np.abs(vo).max()))
        assert pg.shape[0] < sc.num_particles
    finally:
        sim.close()
