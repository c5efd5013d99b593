"""The CPU oracle against committed reference vectors (tests/golden).

tests/golden/reference_small.npz holds inputs and outputs of the reference
engine compiled from its own sources (written by tests/golden/make_golden.py).
These checks pin the oracle bit for bit without needing the reference build.
"""

import os

import numpy as np
import pytest

from tests.helpers import bits_equal, ndiff, radius

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_small.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_domain_solid(oracle, g):
    I, J, K = (int(x) for x in g["solid_dims"])
    phi, w = oracle.domain_solid(I, J, K, float(g["solid_dx"]))
    assert bits_equal(phi, g["solid_phi"]), ndiff(phi, g["solid_phi"])
    for name, a in zip(("wU", "wV", "wW", "wC"), w):
        assert bits_equal(a, g["solid_" + name]), name


def _dam(g):
    I, J, K = (int(x) for x in g["dam_dims"])
    return I, J, K, float(g["dam_dx"])


def test_liquid_sdf(oracle, g):
    I, J, K, dx = _dam(g)
    solid, _ = oracle.domain_solid(I, J, K, dx)
    phi = oracle.liquid_sdf(I, J, K, dx, g["dam_pos"], radius(dx), solid)
    assert bits_equal(phi, g["dam_liquid_phi"]), ndiff(phi, g["dam_liquid_phi"])


def test_p2g_and_extrapolation(oracle, g):
    I, J, K, dx = _dam(g)
    mac, valid = oracle.p2g(I, J, K, dx, g["dam_pos"], g["dam_vel"], radius(dx))
    for c, m, v in zip("uvw", mac, valid):
        assert bits_equal(m, g["dam_p2g_" + c]), c
        assert bits_equal(v, g["dam_valid_" + c]), c
        e = oracle.extrapolate(m, v, 7)
        if c == "v":
            e = (e + np.float32(-9.81 / 60)).astype(np.float32)
        assert bits_equal(e, g["dam_ext_" + c]), c


def test_mic0(oracle, g):
    I, J, K, dx = _dam(g)
    _, weights = oracle.domain_solid(I, J, K, dx)
    inv, z = oracle.mic0_apply(I, J, K, dx, 1 / 60, g["dam_liquid_phi"], weights, g["dam_mic_r"])
    assert bits_equal(inv, g["dam_mic_invdiag"])
    assert bits_equal(z, g["dam_mic_z"])


def test_pressure_solve(oracle, g):
    I, J, K, dx = _dam(g)
    solid, weights = oracle.domain_solid(I, J, K, dx)
    mac = [g["dam_ext_" + c] for c in "uvw"]
    valid = [g["dam_valid_" + c] for c in "uvw"]
    p = oracle.pressure_solve(I, J, K, dx, 1 / 60, g["dam_liquid_phi"], solid, weights, mac, valid, mode=0)
    assert p["ok"] == int(g["dam_ok"]) == 1
    assert p["iters"] == int(g["dam_iters"])
    assert bits_equal(p["rhs"], g["dam_rhs"])
    assert bits_equal(p["soln"], g["dam_soln"])
    for c, m, v in zip("uvw", p["mac"], p["valid"]):
        assert bits_equal(m, g["dam_proj_" + c]), c
        assert bits_equal(v, g["dam_proj_valid_" + c]), c


def test_whole_update(oracle, g):
    I, J, K = (int(x) for x in g["sim_dims"])
    h = oracle.sim(I, J, K, float(g["sim_dx"]))
    h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(1, 6)
    pos0 = g["sim_pos0"]
    h.load_particles(pos0, np.zeros_like(pos0))
    h.initialize()
    for f in range(len(g["sim_substeps"])):
        h.update(1 / 30)
        assert h.substeps() == int(g["sim_substeps"][f]), f"frame {f}"
    p, v = h.particles()
    assert bits_equal(p, g["sim_pos"]), ndiff(p, g["sim_pos"])
    assert bits_equal(v, g["sim_vel"]), ndiff(v, g["sim_vel"])
