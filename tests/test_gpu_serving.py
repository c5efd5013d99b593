"""The serving loop (flipmi_server): concurrent scenes, scene reuse, memory
admission.

A request's result must be the same alone and in the mix: every scene owns
its buffers and its stream, and a re-armed resident scene starts from the
state of a freshly initialised one, so the bar is bit-exact equality with the
same request stepped through a fresh FluidSimulation on its own.
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from flipengine3d_amd import scenes
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu

DT = 1 / 240


def _request(I, J, K, seed, fill):
    dx = 1.0 / max(I, J, K)
    s = scenes.dam_break(I, J, K, dx, seed=seed, box=(0.0, 0.0, 0.0, fill, fill, 1.0))
    desc = fa.scene_desc(I, J, K, dx, precision=fa.PCG_MIXED, substeps=(1, 1))
    return desc, s


def _alone(gpu, desc, s, steps):
    sim = fa.FluidSimulation(desc.isize, desc.jsize, desc.ksize, desc.dx, ctx=gpu, precision=desc.pcg_precision)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(1)
    sim.setMaxTimeStepsPerFrame(1)
    sim.loadMarkerParticleData(s.positions, s.velocities)
    sim.initialize()
    for _ in range(steps):
        st = sim.update(DT)
        assert st.pressure_ok == 1
    pos, vel = sim.particles()
    sim.close()
    return pos, vel


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    steps = 5
    reqs = [_request(24, 24, 24, 11, 0.45), _request(32, 20, 28, 12, 0.55), _request(24, 24, 24, 13, 0.3),
            _request(20, 30, 18, 14, 0.5), _request(32, 20, 28, 15, 0.35), _request(24, 24, 24, 16, 0.6)]
    server = fa.Server(gpu, max_active=3)
    ids = [server.submit(d, s.positions, s.velocities, DT, steps) for d, s in reqs]
    results = [server.result(i, timeout=100.0) for i in ids]
    # A second wave of the same kinds re-arms the resident scenes.
    ids2 = [server.submit(d, s.positions, s.velocities, DT, steps) for d, s in reqs[:3]]
    results2 = [server.result(i, timeout=100.0) for i in ids2]
    stats = server.stats()
    server.close()
    assert stats["completed"] == len(reqs) + 3 and stats["failed"] == 0
    assert stats["scenes_reused"] >= 3
    assert sum(r[0]["reused_scene"] for r in results2) == 3
    for (d, s), (rep, pos, vel) in zip(reqs + reqs[:3], results + results2):
        assert rep["status"] == 0 and rep["steps_done"] == steps and rep["pressure_failures"] == 0
        ap, av = _alone(gpu, d, s, steps)
        assert bits_equal(pos, ap), f"positions differ for {d.isize}x{d.jsize}x{d.ksize}"
        assert bits_equal(vel, av)
        assert rep["particle_substeps"] >= steps * pos.shape[0]


def test_memory_budget_admission(gpu):
    desc, s = _request(28, 28, 28, 21, 0.5)
    probe = fa.Server(gpu, max_active=1)
    probe.result(probe.submit(desc, s.positions, s.velocities, DT, 1), timeout=100.0)
    one = probe.stats()["peak_resident_bytes"]
    probe.close()
    assert one > 0
    # Room for one scene of this kind: three requests with two slots run one
    # after another, and the resident memory never exceeds the budget.
    budget = int(one * 1.5)
    server = fa.Server(gpu, max_active=2, memory_budget=budget)
    ids = [server.submit(desc, s.positions, s.velocities, DT, 2) for _ in range(3)]
    reps = [server.result(i, timeout=100.0)[0] for i in ids]
    st = server.stats()
    assert all(r["status"] == 0 for r in reps)
    assert st["peak_resident_bytes"] <= budget
    assert st["scenes_created"] == 1 and st["scenes_reused"] == 2
    # A scene that can never fit is refused with FLIPMI_ERR_BUDGET.
    big_desc, big = _request(64, 64, 64, 22, 0.4)
    rep = server.wait(server.submit(big_desc, big.positions, big.velocities, DT, 1), timeout=100.0)
    assert rep["status"] == fa.FLIPMI_ERR_BUDGET
    server.close()
