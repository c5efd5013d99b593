"""PCG iteration batches launched as captured HIP graphs give bit-identical
results to launching every kernel.

Each solve runs its first batch as kernels, captures each later (batch size,
starting buffer) the first time it occurs and replays it afterwards, with the
scene's own kernels and host synchronizations between replays; every frame
of the graph run is compared with the kernel-launch run, bit for bit, for
both PCG precisions.
"""

import numpy as np
import pytest

import flipengine3d_amd as fa
from flipengine3d_amd import scenes
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def _scene(gpu, s, precision, graphs):
    sim = fa.FluidSimulation(s.isize, s.jsize, s.ksize, s.dx, ctx=gpu, precision=precision)
    sim.addBodyForce(0.0, -9.81, 0.0)
    sim.setMinTimeStepsPerFrame(1)
    sim.setMaxTimeStepsPerFrame(3)
    sim.loadMarkerParticleData(s.positions, s.velocities)
    sim.initialize()
    sim.set_launch_graphs(graphs)
    return sim


@pytest.mark.parametrize("precision", [fa.PCG_MIXED, fa.PCG_FP64])
def test_graph_batches_match_kernel_launches(gpu, precision):
    s = scenes.double_dam_break(64, 32, 32, 1 / 32, seed=3)
    a = _scene(gpu, s, precision, False)
    b = _scene(gpu, s, precision, True)
    for f in range(6):
        sa = a.update(1 / 60)
        sb = b.update(1 / 60)
        print(f"frame {f}: {sa.substeps} substeps, {sa.pcg_iterations} / {sb.pcg_iterations} iterations", flush=True)
        assert sa.pcg_iterations == sb.pcg_iterations and sa.substeps == sb.substeps
        assert sa.pressure_ok == 1 and sb.pressure_ok == 1
        pa, va = a.particles()
        pb, vb = b.particles()
        assert bits_equal(pa, pb) and bits_equal(va, vb), f"frame {f}"
        for which in (2, 3, 4):
            assert bits_equal(a.grid(which), b.grid(which)), f"frame {f} grid {which}"
    a.close()
    b.close()
