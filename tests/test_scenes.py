"""Host-side scene seeding: counts, jitter bounds, wall rejection, determinism."""

import numpy as np

from flipengine3d_amd import scenes


def test_unjittered_dam_break_count_is_closed_form():
    # 64^3, dx = 1/64, cuboid [0,0.4]x[0,0.5]x[0,1]: 26 x 32 x 64 cells.  Without
    # jitter the sub-cell centres sit at (i + 1/4) dx and (i + 3/4) dx; the
    # wall box [1.5 dx + 5e-5, (n - 1.5) dx - 5e-5) drops the first three
    # per axis at the low walls and three per axis at the high walls.
    s = scenes.dam_break(64, 64, 64, 1 / 64, seed=1, jitter_factor=1e-3)
    assert s.num_particles == (2 * 26 - 3) * (2 * 32 - 3) * (2 * 64 - 6) == 364658


def test_unjittered_full_domain_count():
    for n in (8, 11):
        s = scenes.dam_break(n, n, n, 1 / n, seed=0, box=(0, 0, 0, 1, 1, 1), jitter_factor=1e-3)
        assert s.num_particles == (2 * n - 6) ** 3


def test_double_dam_break_cells():
    # Two cuboids of 0.25 of the x extent, 0.6 of the height, full depth.
    s = scenes.double_dam_break(32, 16, 16, 1 / 16, seed=3, jitter_factor=1e-3)
    per_axis_x = (2 * 8 - 3) * 2       # 8 cells per cuboid, 3 sub-cells lost at each outer wall
    per_axis_y = 2 * 10 - 3            # cells with centre below 0.6 * 16 = 9.6: 10
    per_axis_z = 2 * 16 - 6
    assert s.num_particles == per_axis_x * per_axis_y * per_axis_z


def test_jitter_bounds_and_walls():
    n, dx = 16, 1 / 16
    s = scenes.dam_break(n, n, n, dx, seed=7, box=(0, 0, 0, 1, 1, 1))
    p = s.positions.astype(np.float64)
    j = 0.25 * (1.0 - 1e-3) * dx
    sub = (np.floor(p / (0.5 * dx)) + 0.5) * 0.5 * dx  # nearest sub-cell centre
    assert np.abs(p - sub).max() <= j + 1e-7
    lo = 1.5 * dx + 5e-5
    assert p.min() >= lo - 1e-7 and p.max() < n * dx - lo
    # Sub-cells centred 0.25 dx - 5e-5 inside the wall box can be jittered
    # out (the jitter reaches 0.2495 dx); those centred outside stay out.
    s0 = scenes.dam_break(n, n, n, dx, seed=7, box=(0, 0, 0, 1, 1, 1), jitter_factor=1e-3)
    assert 0.99 * s0.num_particles <= s.num_particles <= s0.num_particles
    assert np.all(s.velocities == 0)


def test_seeding_is_deterministic_and_seed_dependent():
    a = scenes.dam_break(12, 12, 12, 1 / 12, seed=5)
    b = scenes.dam_break(12, 12, 12, 1 / 12, seed=5)
    c = scenes.dam_break(12, 12, 12, 1 / 12, seed=6)
    assert a.positions.tobytes() == b.positions.tobytes()
    assert a.positions.shape != c.positions.shape or a.positions.tobytes() != c.positions.tobytes()


def test_uniform01_range_and_mean():
    u = scenes.uniform01(3, np.arange(200000, dtype=np.uint64), 1)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005
