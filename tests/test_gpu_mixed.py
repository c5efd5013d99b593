"""Mixed-precision PCG against the FP64 oracle.

The mixed path iterates in FP32 (matrix, MIC(0) factor, vectors, SpMV and
triangular solves), accumulates every dot product and the solution x in
FP64, and replaces the FP32 recursive residual by the FP64 true residual
b - A x (FP64 A) whenever it has fallen 100-fold or reached the tolerance.
It stops only when the true residual satisfies |b - A x|_inf <= 1e-9 |b|_inf,
the reference's criterion.

Stated tolerances (per test case): relative pressure error
max|p - p_oracle| / max|p_oracle| <= 1e-6 against the oracle's FP64 solve at
the same 1e-9 tolerance; projected face velocities within 1e-5 of the
largest velocity; valid flags bit-exact.
"""

import numpy as np
import pytest

from tests.helpers import bits_equal, dam, radius, random_velocities

pytestmark = pytest.mark.gpu

PCG_MIXED = 1


def _projection_inputs(oracle, I, J, K, seed=4, box=(0.0, 0.0, 0.0, 0.4, 0.5, 1.0)):
    s = dam(I, J, K, seed=seed, box=box)
    dx = s.dx
    solid, weights = oracle.domain_solid(I, J, K, dx)
    liquid = oracle.liquid_sdf(I, J, K, dx, s.positions, radius(dx), solid)
    vel = random_velocities(s.num_particles, seed=seed, scale=0.3)
    mac, valid = oracle.p2g(I, J, K, dx, s.positions, vel, radius(dx))
    mac = [oracle.extrapolate(g, v, 7) for g, v in zip(mac, valid)]
    mac[1] = (mac[1] + np.float32(-9.81 / 60)).astype(np.float32)
    return dx, solid, weights, liquid, mac, valid


@pytest.mark.parametrize("dims", [(16, 16, 16), (24, 28, 21), (37, 22, 30), (48, 40, 36)])
def test_mixed_pressure_matches_fp64_oracle(gpu, oracle, dims):
    I, J, K = dims
    dx, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K)
    dt = 1 / 60
    pm = gpu.pressure_solve(I, J, K, dx, dt, liquid, weights, mac, valid, precision=PCG_MIXED)
    po = oracle.pressure_solve(I, J, K, dx, dt, liquid, solid, weights, mac, valid, mode=1)
    assert pm["ok"] == po["ok"] == 1
    assert pm["rows"] == len(po["soln"])
    assert pm["refinements"] >= 1
    # The reported residual is the FP64 true residual at exit.
    bnorm = np.abs(po["rhs"]).max()
    assert pm["err"] <= 1e-9 * bnorm
    rel = np.abs(pm["soln"] - po["soln"]).max() / np.abs(po["soln"]).max()
    print(f"{dims}: iterations mixed {pm['iters']} fp64 {po['iters']}, refinements {pm['refinements']}, "
          f"relative pressure error {rel:.3e}")
    assert rel <= 1e-6
    for a, b in zip(pm["valid"], po["valid"]):
        assert bits_equal(a, b)
    vmax = max(float(np.abs(g).max()) for g in po["mac"])
    for a, b in zip(pm["mac"], po["mac"]):
        assert np.abs(a.astype(np.float64) - b).max() <= 1e-5 * vmax


@pytest.mark.parametrize("dims,box", [((1100, 16, 16), (0.0, 0.0, 0.0, 0.9, 0.6, 1.0)),
                                      ((128, 24, 16), (0.45, 0.3, 0.0, 0.6, 0.8, 1.0))])
def test_mixed_pressure_long_tiles_and_rows_far_along_i(gpu, oracle, dims, box):
    # The mixed iteration's forward sweep also does the axpy (a third wave
    # updates x) and the backward sweep's z.r is taken by a third wave; both
    # walk a tile's active slot groups.  Tiles of more than 256 groups
    # (I > 1010) have no group mask and take every group; rows only at large
    # i move each tile's march range away from step 0.  Same bounds as above.
    I, J, K = dims
    dx, solid, weights, liquid, mac, valid = _projection_inputs(oracle, I, J, K, seed=6, box=box)
    dt = 1 / 60
    pm = gpu.pressure_solve(I, J, K, dx, dt, liquid, weights, mac, valid, precision=PCG_MIXED)
    po = oracle.pressure_solve(I, J, K, dx, dt, liquid, solid, weights, mac, valid, mode=1)
    assert pm["ok"] == po["ok"] == 1
    assert pm["rows"] == len(po["soln"])
    bnorm = np.abs(po["rhs"]).max()
    assert pm["err"] <= 1e-9 * bnorm
    rel = np.abs(pm["soln"] - po["soln"]).max() / np.abs(po["soln"]).max()
    print(f"{dims}: iterations mixed {pm['iters']} fp64 {po['iters']}, refinements {pm['refinements']}, "
          f"relative pressure error {rel:.3e}", flush=True)
    assert rel <= 1e-6
    for a, b in zip(pm["valid"], po["valid"]):
        assert bits_equal(a, b)


def test_mixed_scene_tracks_fp64_scene(gpu):
    import flipengine3d_amd as fa
    n = 24
    s = dam(n, n, n, seed=2)
    sims = []
    for prec in (fa.PCG_FP64, fa.PCG_MIXED):
        sim = fa.FluidSimulation(n, n, n, s.dx, ctx=gpu, precision=prec)
        sim.addBodyForce(0.0, -9.81, 0.0)
        sim.loadMarkerParticleData(s.positions, s.velocities)
        sim.initialize()
        sims.append(sim)
    for f in range(4):
        st = [sim.update(1 / 60) for sim in sims]
        assert st[0].pressure_ok == st[1].pressure_ok == 1
        assert st[0].substeps == st[1].substeps
        assert st[0].num_particles == st[1].num_particles
    (p0, v0), (p1, v1) = sims[0].particles(), sims[1].particles()
    assert np.abs(p0 - p1).max() <= 1e-3 * s.dx
    assert np.abs(v0 - v1).max() <= 1e-3
