"""Static mesh obstacles in the CPU oracle, pinned against the compiled
reference (oracle/_ref) and a closed form.

* The solid SDF with obstacles (MeshObject::getMeshLevelSet unioned into the
  domain SDF, fluidsimulation.cpp:2932-2985, meshobject.cpp:239-262,
  519-553), the weights and the particles after three frames are bit-exact
  against the reference.  The reference jitters the rays of every inside
  test with rand(); the harness seeds rand() before the first update and
  predicts the draws (domain box first, then each obstacle), which the
  oracle is given.  A wrong jitter changes the solid on these meshes.
* Disabling an obstacle rebuilds the solid without it (fluidsimulation.cpp:
  3028-3036), as in the reference.
* Closed form: within the exact band of an axis-aligned box obstacle the
  solid SDF is the signed distance to the box (negative inside), up to float
  rounding.
"""

import numpy as np
import pytest

from flipengine3d_amd import scenes
from oracle.bindings import _SimHandle

I = J = K = 32
DX = 1.0 / 32

SETS = {
    "rotated_box_and_cylinder": [scenes.box_mesh((0.45, 0.2, 0.3), (0.65, 0.5, 0.7), scenes.rotation_matrix(4)),
                                 scenes.cylinder_mesh(0.2, 0.8, (0.75, 0.5), 0.1, segments=20, phase=0.3)],
    "two_rotated_boxes": [scenes.box_mesh((0.3, 0.25, 0.3), (0.5, 0.45, 0.55), scenes.rotation_matrix(7)),
                          scenes.box_mesh((0.6, 0.5, 0.4), (0.8, 0.7, 0.6), scenes.rotation_matrix(8))],
}


def _sim(L, meshes, jitters=None):
    sc = scenes.dam_break(I, J, K, DX, seed=3)
    s = _SimHandle(L, I, J, K, DX)
    s.add_body_force(0.0, -9.81, 0.0)
    s.load_particles(sc.positions, sc.velocities)
    for q, (v, t) in enumerate(meshes):
        s.add_obstacle(v, t) if jitters is None else s.add_obstacle(v, t, jitters[q])
    s.initialize()
    return s


def _pair(oracle, reference, meshes, seed):
    b = _sim(reference, meshes)
    jit = reference.predict_ray_jitters(seed, DX, 1 + len(meshes))  # domain box, then the obstacles
    a = _sim(oracle, meshes, jitters=jit[1:])
    return a, b, jit


@pytest.mark.parametrize("name", sorted(SETS))
def test_obstacle_solid_against_reference(oracle, reference, name):
    a, b, _ = _pair(oracle, reference, SETS[name], 4242)
    for frame in range(3):
        a.update(1 / 30)
        b.update(1 / 30)
        if frame == 0:
            sa, sb = a.float_grid(1), b.float_grid(1)
            assert sa.tobytes() == sb.tobytes()
            plain = _sim(oracle, [])
            plain.update(1 / 30)
            assert int((sa < 0).sum()) > int((plain.float_grid(1) < 0).sum()) + 500
            for w in (5, 6, 7, 8):
                assert a.float_grid(w).tobytes() == b.float_grid(w).tobytes()
        pa, va = a.particles()
        pb, vb = b.particles()
        assert pa.shape == pb.shape and pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes(), frame


def test_obstacle_solid_depends_on_the_ray_jitter(oracle, reference):
    meshes = SETS["rotated_box_and_cylinder"]
    b = _sim(reference, meshes)
    reference.predict_ray_jitters(4242, DX, 3)
    b.update(1 / 30)
    a = _sim(oracle, meshes, jitters=np.zeros((2, 3), np.float32))
    a.update(1 / 30)
    assert a.float_grid(1).tobytes() != b.float_grid(1).tobytes()


def test_disabling_an_obstacle_rebuilds_the_solid(oracle, reference):
    meshes = SETS["rotated_box_and_cylinder"][:1]
    a, b, _ = _pair(oracle, reference, meshes, 99)
    a.update(1 / 30)
    b.update(1 / 30)
    a.set_obstacle(0, False)
    b.set_obstacle(0, False)
    a.update(1 / 30)
    b.update(1 / 30)
    plain = _sim(oracle, [])
    plain.update(1 / 30)
    assert a.float_grid(1).tobytes() == b.float_grid(1).tobytes() == plain.float_grid(1).tobytes()
    pa, _ = a.particles()
    pb, _ = b.particles()
    assert pa.tobytes() == pb.tobytes()


def test_axis_box_obstacle_closed_form(oracle):
    lo, hi = np.array([0.3071, 0.2813, 0.3342]), np.array([0.6417, 0.5229, 0.6608])
    v, t = scenes.box_mesh(lo, hi)
    s = _sim(oracle, [(v, t)], jitters=np.zeros((1, 3), np.float32))
    s.update(1 / 60)
    phi = s.float_grid(1).astype(np.float64)
    k, j, i = np.meshgrid(*(np.arange(I + 1) * DX,) * 3, indexing="ij")
    p = np.stack([i, j, k], axis=-1)
    q = np.abs(p - (lo + hi) / 2) - (hi - lo) / 2
    outside = np.linalg.norm(np.maximum(q, 0), axis=-1)
    inside = np.minimum(q.max(axis=-1), 0)
    sd = outside + inside  # signed distance to the box, negative inside
    near = np.abs(sd) < 2.5 * DX  # inside the exact band of 3 cells
    assert near.sum() > 1000
    assert np.abs(phi[near] - sd[near]).max() < 1e-5
    assert ((phi < 0) == (sd < 0))[near].all()


def test_inversed_obstacle_against_reference(oracle, reference):
    """An inversed obstacle (MeshObject::inverse) is a container: solid
    outside its mesh (fluidsimulation.cpp:2968-2984).  With a rotated box
    inside it, the solid SDF, weights and particles over 3 frames are
    bit-exact against the reference.  The reference draws the ray jitters
    for the normal obstacles first, then the inversed ones."""
    container = scenes.box_mesh((0.06, 0.04, 0.07), (0.93, 0.91, 0.95))
    inner = SETS["rotated_box_and_cylinder"][0]
    sc = scenes.dam_break(I, J, K, DX, seed=3)

    def mk(L, jit=None):
        s = _SimHandle(L, I, J, K, DX)
        s.add_body_force(0.0, -9.81, 0.0)
        s.load_particles(sc.positions, sc.velocities)
        if jit is None:
            s.add_obstacle(*container)
            s.add_obstacle(*inner)
        else:
            s.add_obstacle(*container, jitter=jit[2])
            s.add_obstacle(*inner, jitter=jit[1])
        s.set_obstacle_inversed(0, True)
        s.initialize()
        return s
    b = mk(reference)
    jit = reference.predict_ray_jitters(777, DX, 3)
    a = mk(oracle, jit)
    for frame in range(3):
        a.update(1 / 30)
        b.update(1 / 30)
        if frame == 0:
            sa = a.float_grid(1)
            assert sa.tobytes() == b.float_grid(1).tobytes()
            for w in (5, 6, 7, 8):
                assert a.float_grid(w).tobytes() == b.float_grid(w).tobytes()
            # outside the container is solid
            assert sa[0, 0, 0] < 0 and sa[K // 2, J // 2, 1] < 0  # (k, j, i): x = dx is outside
        pa, va = a.particles()
        pb, vb = b.particles()
        assert pa.shape == pb.shape and pa.tobytes() == pb.tobytes() and va.tobytes() == vb.tobytes(), frame
