"""Closed-form and brute-force pins for the CPU oracle.

Each check compares the oracle with what the mathematics fixes, computed
independently here (closed forms, textbook algorithms, brute force on tiny
inputs), and is chosen so that a plausible slip in the oracle -- a dropped
term, a wrong sign, offset or index, a transposed operand -- fails it:

  domain solid SDF ... distance to the inset wall box (closed form)
  weights ............ open fractions of faces / cells cut by a wall plane
  liquid SDF ......... brute-force min distance minus radius; labels
  P2G ................ brute-force Wyvill-kernel weighted average per face
  extrapolation ...... hand-computed breadth-first layers
  Poisson system ..... rhs = negative weighted divergence; solution vs a
                       dense direct solve; projected field divergence-free
  MIC(0) ............. Bridson's textbook MIC(0) recurrence + triangular solves
  trilinear sampling . exact for fields linear in x, y, z
  RK1-RK4 ............ truncated exponential series on u = a x
  whole update ....... free fall of a blob; at-rest scene; 250 per-cell cap
"""

import numpy as np
import pytest

from tests.helpers import radius


# ---------------------------------------------------------------------------
# Domain walls
# ---------------------------------------------------------------------------
def _wall_box(I, J, K, dx):
    # The fluid region is the domain inset by 1.5 dx + 5e-5 on every side
    # (FluidSimulation::_getBoundaryAABB expands by -(3 dx + 1e-4)).
    lo = 1.5 * dx + 5e-5
    return np.array([lo, lo, lo]), np.array([I * dx - lo, J * dx - lo, K * dx - lo])


def _box_signed_distance(p, lo, hi):
    """Positive inside the box: distance to the nearest face; negative outside:
    minus the Euclidean distance to the box."""
    inside = np.all((p > lo) & (p < hi), axis=-1)
    d_in = np.min(np.concatenate([p - lo, hi - p], axis=-1), axis=-1)
    q = np.maximum(np.maximum(lo - p, p - hi), 0.0)
    d_out = np.sqrt(np.sum(q * q, axis=-1))
    return np.where(inside, d_in, -d_out)


@pytest.mark.parametrize("dims", [(14, 11, 13), (20, 20, 20)])
def test_domain_solid_is_box_distance(oracle, dims):
    I, J, K = dims
    dx = 1.0 / max(dims)
    phi, _ = oracle.domain_solid(I, J, K, dx)
    k, j, i = np.meshgrid(np.arange(K + 1), np.arange(J + 1), np.arange(I + 1), indexing="ij")
    nodes = np.stack([i * dx, j * dx, k * dx], axis=-1)
    lo, hi = _wall_box(I, J, K, dx)
    d = _box_signed_distance(nodes, lo, hi)
    # Sign everywhere: positive exactly on the fluid side.
    assert np.array_equal(phi > 0, d > 0)
    # Value inside the exact band (3 cells around the walls).
    band = np.abs(d) <= 2.5 * dx
    assert band.sum() > 0
    assert np.abs(phi[band] - d[band]).max() <= 2e-6


def test_weights_of_cut_faces_and_cells(oracle):
    I, J, K = 12, 12, 12
    dx = 1.0 / 12
    _, (wu, wv, ww, wc) = oracle.domain_solid(I, J, K, dx)
    # The wall plane x = 1.5 dx + 5e-5 cuts the cells i = 1; its open part is
    # (2 dx - x0) / dx = 0.5 - 5e-5/dx of every face / cell it crosses.
    frac = 0.5 - 5e-5 / dx
    mid = slice(3, 9)
    assert np.allclose(wv[mid, mid, 1], frac, atol=1e-5)  # V faces (k, j, i=1)
    assert np.allclose(ww[mid, mid, 1], frac, atol=1e-5)
    assert np.allclose(wc[mid, mid, 1], frac, atol=1e-5)
    # U faces are parallel to the plane: closed at x = dx, open at x = 2 dx.
    assert np.all(wu[mid, mid, 1] == 0.0)
    assert np.all(wu[mid, mid, 2] == 1.0)
    assert np.all(wu[mid, mid, 0] == 0.0)
    # Interior faces and cells are fully open.
    assert np.all(wc[3:9, 3:9, 3:9] == 1.0)
    assert np.all(wv[3:9, 3:9, 3:9] == 1.0)
    # Mirror wall at x = I dx - x0 cuts cells i = I - 2.
    assert np.allclose(wc[mid, mid, I - 2], frac, atol=1e-5)
    assert np.all(wu[mid, mid, I - 1] == 0.0)


# ---------------------------------------------------------------------------
# Liquid signed distance field
# ---------------------------------------------------------------------------
def test_liquid_sdf_brute_force(oracle):
    I, J, K = 12, 10, 11
    dx = 1.0 / 12
    r = radius(dx)
    rng = np.random.default_rng(5)
    pos = (rng.random((60, 3)) * np.array([I, J, K]) * dx * 0.8 + 0.1 * np.array([I, J, K]) * dx)
    pos = pos.astype(np.float32)
    phi = oracle.liquid_sdf(I, J, K, dx, pos, r)
    k, j, i = np.meshgrid(np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    c = np.stack([(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx], axis=-1)
    dmin = np.min(np.linalg.norm(c[..., None, :] - pos.astype(np.float64)[None, None, None], axis=-1),
                  axis=-1)
    assert phi.min() >= -r - 1e-6
    assert phi.max() <= 3 * dx * (1 + 1e-6)
    near = dmin < 1.9 * r
    assert near.sum() > 100
    assert np.abs(phi[near] - (dmin[near] - r)).max() <= 2e-6
    sure = np.abs(dmin - r) > 1e-5
    assert np.array_equal((phi < 0)[sure], (dmin < r)[sure])
    # Cells no particle reaches keep the 3 dx cap.
    far = dmin > 2 * r * np.sqrt(3) + dx
    assert far.sum() > 0 and np.all(phi[far] == np.float32(3 * dx))


def test_liquid_sdf_postprocess_near_walls(oracle):
    I, J, K = 12, 12, 12
    dx = 1.0 / 12
    solid, _ = oracle.domain_solid(I, J, K, dx)
    rng = np.random.default_rng(2)
    pos = (rng.random((400, 3)) * 0.3 + 0.12).astype(np.float32)  # touches the low walls
    raw = oracle.liquid_sdf(I, J, K, dx, pos, radius(dx))
    post = oracle.liquid_sdf(I, J, K, dx, pos, radius(dx), solid)
    k, j, i = np.meshgrid(np.arange(K), np.arange(J), np.arange(I), indexing="ij")
    c = np.stack([(i + 0.5) * dx, (j + 0.5) * dx, (k + 0.5) * dx], axis=-1)
    lo, hi = _wall_box(I, J, K, dx)
    dsolid = _box_signed_distance(c, lo, hi)
    in_solid = (dsolid < -0.1 * dx) & (raw < 0.5 * dx)
    assert in_solid.sum() > 0
    assert np.all(post[in_solid] == np.float32(-0.5 * dx))
    free = dsolid > 0.1 * dx
    tiny = np.abs(raw) < 0.005 * dx
    assert np.array_equal(post[free & ~tiny], raw[free & ~tiny])
    assert np.all(np.abs(post[free & tiny]) == np.float32(0.005 * dx))


# ---------------------------------------------------------------------------
# P2G
# ---------------------------------------------------------------------------
def _wyvill(d2, r):
    q = d2 / (r * r)
    return 1.0 - (4.0 / 9.0) * q ** 3 + (17.0 / 9.0) * q ** 2 - (22.0 / 9.0) * q


def test_p2g_brute_force(oracle):
    I, J, K = 8, 7, 9
    dx = 1.0 / 9
    r = radius(dx)
    rng = np.random.default_rng(9)
    n = 300
    pos = (rng.random((n, 3)) * np.array([I, J, K]) * dx).astype(np.float32)
    vel = rng.standard_normal((n, 3)).astype(np.float32)
    (mu, mv, mw), (vu, vv, vw) = oracle.p2g(I, J, K, dx, pos, vel, r)
    p64 = pos.astype(np.float64)
    for comp, (m, vd) in enumerate(((mu, vu), (mv, vv), (mw, vw))):
        nk, nj, ni = m.shape
        k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
        f = np.stack([i + 0.5, j + 0.5, k + 0.5], axis=-1) * dx
        f[..., comp] -= 0.5 * dx  # the face of component c sits at a whole index along c
        d2 = np.sum((f[..., None, :] - p64[None, None, None]) ** 2, axis=-1)
        w = np.where(d2 < r * r, _wyvill(d2, r), 0.0)
        sw = w.sum(-1)
        swv = (w * vel[:, comp].astype(np.float64)).sum(-1)
        sure = np.abs(sw - 1e-6) > 1e-4
        assert np.array_equal((vd == 1)[sure], (sw > 1e-6)[sure])
        ok = (vd == 1) & (sw > 1e-3)
        assert ok.sum() > 50
        ref = swv[ok] / sw[ok]
        # The kernel is evaluated in float: its absolute error (~1e-7) is
        # amplified by 1 / sum(w) in the average.
        vmax = float(np.abs(vel[:, comp]).max())
        assert np.all(np.abs(m[ok] - ref) <= 2e-6 * vmax * (1.0 + 1.0 / sw[ok]))
        assert np.all(m[vd == 0] == 0.0)


def test_p2g_uniform_velocity_is_reproduced(oracle):
    I, J, K = 10, 10, 10
    dx = 0.1
    rng = np.random.default_rng(1)
    pos = (rng.random((2000, 3)) * 1.0).astype(np.float32)
    vel = np.tile(np.array([[0.3, -1.25, 2.0]], np.float32), (2000, 1))
    mac, valid = oracle.p2g(I, J, K, dx, pos, vel, radius(dx))
    for c, (m, v) in enumerate(zip(mac, valid)):
        assert np.abs(m[v == 1] - vel[0, c]).max() <= 4e-7 * abs(vel[0, c]) * 4


# ---------------------------------------------------------------------------
# Extrapolation
# ---------------------------------------------------------------------------
def test_extrapolation_hand_cases(oracle):
    n = 9
    g = np.zeros((n, n, n), np.float32)
    v = np.zeros((n, n, n), np.uint8)
    g[4, 4, 4] = 5.0
    v[4, 4, 4] = 1
    e1 = oracle.extrapolate(g, v, 1)
    for kk, jj, ii in ((4, 4, 5), (4, 4, 3), (4, 5, 4), (4, 3, 4), (5, 4, 4), (3, 4, 4)):
        assert e1[kk, jj, ii] == 5.0
    assert e1[4, 4, 6] == 0.0 and e1[4, 5, 5] == 0.0
    e2 = oracle.extrapolate(g, v, 2)
    assert e2[4, 4, 6] == 5.0 and e2[4, 5, 5] == 5.0 and e2[4, 4, 7] == 0.0
    e4 = oracle.extrapolate(g, v, 4)
    # (4, 4, 7) averages its seed-side neighbour (5) with the border face (0).
    assert e4[4, 4, 7] == 2.5 and e4[4, 4, 8] == 0.0
    n = 7

    # Mean of the finished neighbours: (1 + 3) / 2 between two seeds.
    g = np.zeros((n, n, n), np.float32)
    v = np.zeros((n, n, n), np.uint8)
    g[3, 3, 2], g[3, 3, 4] = 1.0, 3.0
    v[3, 3, 2] = v[3, 3, 4] = 1
    e = oracle.extrapolate(g, v, 2)
    assert e[3, 3, 3] == 2.0
    assert e[3, 4, 3] == 2.0  # (3 + 1 + 2) / 3 from +i, -i, -j

    # Border faces are finished from the start: they are never overwritten,
    # never seed a layer, but enter the mean of their neighbours.
    g = np.full((n, n, n), 100.0, np.float32)
    g[1:-1, 1:-1, 1:-1] = 0.0
    v = np.zeros((n, n, n), np.uint8)
    g[3, 3, 1] = 1.0
    v[3, 3, 1] = 1
    e = oracle.extrapolate(g, v, 1)
    assert e[3, 3, 2] == 1.0
    assert e[3, 4, 1] == 50.5  # (100 border + 1 seed) / 2
    assert e[3, 3, 0] == 100.0
    assert e[3, 5, 1] == 0.0   # reached by no seed in one layer


# ---------------------------------------------------------------------------
# Poisson system: assembly, MIC(0), PCG, projection
# ---------------------------------------------------------------------------
def _box_system(n=8, lo=2, hi=5, seed=0):
    """Liquid box [lo, hi]^3 in an n^3 grid with random open fractions; returns
    the inputs and the matrix assembled from the variational ghost-fluid
    discretisation."""
    dx = 1.0 / n
    dt = 1.0 / 60
    rng = np.random.default_rng(seed)
    phi = np.full((n, n, n), 0.5 * dx, np.float32)
    phi[lo:hi + 1, lo:hi + 1, lo:hi + 1] = -0.5 * dx
    wu = rng.uniform(0.3, 1.0, (n, n, n + 1)).astype(np.float32)
    wv = rng.uniform(0.3, 1.0, (n, n + 1, n)).astype(np.float32)
    ww = rng.uniform(0.3, 1.0, (n + 1, n, n)).astype(np.float32)
    wc = rng.uniform(0.3, 1.0, (n, n, n)).astype(np.float32)
    cells = [(i, j, k) for k in range(1, n - 1) for j in range(1, n - 1) for i in range(1, n - 1)
             if phi[k, j, i] < 0]
    index = {c: q for q, c in enumerate(cells)}
    f = dt / (dx * dx)
    A = np.zeros((len(cells), len(cells)))
    for q, (i, j, k) in enumerate(cells):
        faces = [((i + 1, j, k), wu[k, j, i + 1]), ((i - 1, j, k), wu[k, j, i]),
                 ((i, j + 1, k), wv[k, j + 1, i]), ((i, j - 1, k), wv[k, j, i]),
                 ((i, j, k + 1), ww[k + 1, j, i]), ((i, j, k - 1), ww[k, j, i])]
        diag = f * sum(float(w) for _, w in faces)
        for (a, b, c), w in faces:
            pn = float(phi[c, b, a])
            if pn < 0:
                A[q, index[(a, b, c)]] = -f * float(w)
            else:
                theta = np.clip(pn / (float(phi[k, j, i]) + 1e-9), -25, 25)
                diag -= f * float(w) * theta
        A[q, q] = max(diag, 0.0)
    return dict(n=n, dx=dx, dt=dt, phi=phi, weights=(wu, wv, ww, wc), cells=cells, index=index, A=A)


def _textbook_mic0(A, cells, index, tau=0.97, sigma=0.25):
    """Bridson, Fluid Simulation for Computer Graphics (2nd ed.), modified
    incomplete Cholesky MIC(0) for the 7-point stencil, and its application
    by forward / backward substitution."""
    m = len(cells)
    pre = np.zeros(m)

    def up(q, d):  # coupling of row q to its +d neighbour
        i, j, k = cells[q]
        nb = (i + (d == 0), j + (d == 1), k + (d == 2))
        return A[q, index[nb]] if nb in index else 0.0

    def low(q, d):  # row index of the -d neighbour
        i, j, k = cells[q]
        nb = (i - (d == 0), j - (d == 1), k - (d == 2))
        return index.get(nb)

    for q in range(m):
        e = A[q, q]
        for d in range(3):
            p = low(q, d)
            if p is None:
                continue
            a = up(p, d)
            e -= (a * pre[p]) ** 2
            others = sum(up(p, o) for o in range(3) if o != d)
            e -= tau * a * others * pre[p] ** 2
        if e < sigma * A[q, q]:
            e = A[q, q]
        pre[q] = 1.0 / np.sqrt(e)

    def apply(r):
        y = np.zeros(m)
        for q in range(m):
            t = r[q]
            for d in range(3):
                p = low(q, d)
                if p is not None:
                    t -= up(p, d) * pre[p] * y[p]
            y[q] = t * pre[q]
        z = np.zeros(m)
        for q in reversed(range(m)):
            t = y[q]
            for d in range(3):
                i, j, k = cells[q]
                nb = (i + (d == 0), j + (d == 1), k + (d == 2))
                if nb in index:
                    t -= up(q, d) * pre[q] * z[index[nb]]
            z[q] = t * pre[q]
        return z

    return pre, apply


def test_mic0_matches_textbook(oracle):
    s = _box_system(n=9, lo=2, hi=6, seed=1)
    assert np.allclose(s["A"], s["A"].T)
    pre, apply = _textbook_mic0(s["A"], s["cells"], s["index"])
    r = np.random.default_rng(4).standard_normal(len(s["cells"]))
    inv, z = oracle.mic0_apply(s["n"], s["n"], s["n"], s["dx"], s["dt"], s["phi"], s["weights"], r)
    assert np.allclose(inv, pre, rtol=1e-12, atol=0)
    assert np.allclose(z, apply(r), rtol=1e-10, atol=1e-12 * np.abs(z).max())


def _mac_random(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((n, n, n + 1)).astype(np.float32)
    v = rng.standard_normal((n, n + 1, n)).astype(np.float32)
    w = rng.standard_normal((n + 1, n, n)).astype(np.float32)
    return [u, v, w]


@pytest.mark.parametrize("mode,tol,sol_bar,res_bar", [
    (0, 1e-9, 1e-6, 1e-8),    # reference arithmetic: float alpha/beta, sequential dots
    (1, 1e-9, 1e-6, 1e-8),    # FP64 alpha/beta, stop on the recursive residual
    (2, 1e-12, 1e-9, 1e-11),  # converged error oracle: stop on the true residual
])
def test_pressure_solve_against_dense_solve(oracle, mode, tol, sol_bar, res_bar):
    s = _box_system(n=9, lo=2, hi=6, seed=2)
    n, dx, dt = s["n"], s["dx"], s["dt"]
    wu, wv, ww, wc = [a.astype(np.float64) for a in s["weights"]]
    mac = _mac_random(n, 3)
    u, v, w = [a.astype(np.float64) for a in mac]
    valid = [np.ones(a.shape, np.uint8) for a in mac]
    solid = np.ones((n + 1, n + 1, n + 1), np.float32)  # no solid anywhere
    res = oracle.pressure_solve(n, n, n, dx, dt, s["phi"], solid, s["weights"], mac, valid, tol=tol,
                                maxiter=10000, mode=mode)
    assert res["ok"] == 1
    # rhs = -(1/dx) * weighted divergence (static solid: zero face velocities).
    rhs = np.zeros(len(s["cells"]))
    for q, (i, j, k) in enumerate(s["cells"]):
        div = (wu[k, j, i + 1] * u[k, j, i + 1] - wu[k, j, i] * u[k, j, i]
               + wv[k, j + 1, i] * v[k, j + 1, i] - wv[k, j, i] * v[k, j, i]
               + ww[k + 1, j, i] * w[k + 1, j, i] - ww[k, j, i] * w[k, j, i])
        rhs[q] = -div / dx
    assert np.allclose(res["rhs"], rhs, rtol=1e-12, atol=1e-12 * np.abs(rhs).max())
    x = np.linalg.solve(s["A"], rhs)
    assert np.abs(res["soln"] - x).max() <= sol_bar * np.abs(x).max()
    assert np.abs(s["A"] @ res["soln"] - rhs).max() <= res_bar * np.abs(rhs).max()
    # The projected field is divergence-free (weighted) in every fluid cell.
    pu, pv, pw = [a.astype(np.float64) for a in res["mac"]]
    worst = 0.0
    for (i, j, k) in s["cells"]:
        div = (wu[k, j, i + 1] * pu[k, j, i + 1] - wu[k, j, i] * pu[k, j, i]
               + wv[k, j + 1, i] * pv[k, j + 1, i] - wv[k, j, i] * pv[k, j, i]
               + ww[k + 1, j, i] * pw[k + 1, j, i] - ww[k, j, i] * pw[k, j, i])
        worst = max(worst, abs(div))
    assert worst <= 1e-5 * dx * np.abs(rhs).max()


# ---------------------------------------------------------------------------
# Sampling and particle advection
# ---------------------------------------------------------------------------
def _linear_mac(I, J, K, dx, coef):
    out = []
    for comp in range(3):
        shape = (K + (comp == 2), J + (comp == 1), I + (comp == 0))
        k, j, i = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
        x = (i + 0.5 * (comp != 0)) * dx
        y = (j + 0.5 * (comp != 1)) * dx
        z = (k + 0.5 * (comp != 2)) * dx
        a, b, c, d = coef[comp]
        out.append((a + b * x + c * y + d * z).astype(np.float32))
    return out


def test_trilinear_exact_for_linear_fields(oracle):
    I, J, K = 12, 10, 14
    dx = 1.0 / 14
    coef = [(0.1, 1.0, -2.0, 0.5), (-0.3, 0.25, 0.75, -1.5), (0.2, -0.5, 0.0, 2.0)]
    mac = _linear_mac(I, J, K, dx, coef)
    rng = np.random.default_rng(2)
    pos = (1.5 * dx + rng.random((2000, 3)) * (np.array([I, J, K]) - 3.0) * dx).astype(np.float32)
    out = oracle.trilinear_velocity(I, J, K, dx, mac, pos)
    p = pos.astype(np.float64)
    for comp in range(3):
        a, b, c, d = coef[comp]
        exact = a + b * p[:, 0] + c * p[:, 1] + d * p[:, 2]
        assert np.abs(out[:, comp] - exact).max() <= 2e-6


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_rk_order_on_exponential_flow(oracle, order):
    """For u = a x (v = w = 0), an explicit s-stage method of order s = 1..4
    advances x by the exponential series truncated after h^order / order!."""
    I, J, K = 16, 16, 16
    dx = 1.0 / 16
    a = 4.0
    mac = _linear_mac(I, J, K, dx, [(0.0, a, 0.0, 0.0), (0.0, 0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 0.0)])
    rng = np.random.default_rng(order)
    pos = np.stack([0.3 + 0.2 * rng.random(500), 0.2 + 0.6 * rng.random(500), 0.2 + 0.6 * rng.random(500)],
                   axis=1).astype(np.float32)
    dt = 0.1
    h = a * dt
    series = sum(h ** m / np.prod(range(1, m + 1)) for m in range(order + 1))
    out = oracle.advect_particles(I, J, K, dx, mac, pos, dt, order)
    x = pos[:, 0].astype(np.float64)
    assert np.abs(out[:, 0] - x * series).max() <= 1e-6
    assert np.array_equal(out[:, 1:], pos[:, 1:])
    # The next term of the series is far above float rounding, so a method of
    # another order fails the check above.
    assert (h ** (order + 1) / np.prod(range(1, order + 2))) * x.min() > 1e-5


# ---------------------------------------------------------------------------
# Whole update(dt)
# ---------------------------------------------------------------------------
def _blob(n, dx, centre, half, per_axis=4):
    g = (np.arange(per_axis) + 0.5) / per_axis * 2 * half - half
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1) + np.array(centre)
    return p.astype(np.float32)


def _sim(oracle, n, dx, pos, gravity, substeps=(1, 1)):
    h = oracle.sim(n, n, n, dx)
    if gravity:
        h.add_body_force(0.0, -9.81, 0.0)
    h.set_substeps(*substeps)
    h.load_particles(pos, np.zeros_like(pos))
    h.initialize()
    return h


def test_free_fall_of_a_blob(oracle):
    """A blob at rest far from the walls under gravity: the grid velocity is
    uniform, the pressure vanishes, and one substep moves every particle by
    -g dt^2 and sets its velocity to -g dt."""
    n, dx, dt = 24, 1.0 / 24, 1.0 / 60
    pos = _blob(n, dx, (0.5, 0.6, 0.5), 2.0 * dx, per_axis=8)
    h = _sim(oracle, n, dx, pos, gravity=True)
    h.update(dt)
    assert h.substeps() == 1
    p, v = h.particles()
    assert p.shape == pos.shape
    assert np.abs(v[:, 1] - (-9.81 * dt)).max() <= 1e-6
    assert np.all(v[:, [0, 2]] == 0.0)
    assert np.abs((p[:, 1].astype(np.float64) - pos[:, 1]) - (-9.81 * dt * dt)).max() <= 1e-6
    assert np.array_equal(p[:, [0, 2]], pos[:, [0, 2]])


def test_scene_at_rest_stays_at_rest(oracle):
    n, dx = 16, 1.0 / 16
    pos = _blob(n, dx, (0.5, 0.5, 0.5), 3.0 * dx, per_axis=12)
    h = _sim(oracle, n, dx, pos, gravity=False, substeps=(1, 3))
    for _ in range(3):
        h.update(1.0 / 30)
    p, v = h.particles()
    assert np.array_equal(p, pos)
    assert np.all(v == 0.0)


def test_per_cell_cap_keeps_the_first_250(oracle):
    n, dx = 16, 1.0 / 16
    rng = np.random.default_rng(0)
    c = np.array([8.5, 8.5, 8.5]) * dx
    pos = (c + (rng.random((300, 3)) - 0.5) * 0.9 * dx).astype(np.float32)
    h = _sim(oracle, n, dx, pos, gravity=False)
    h.update(1.0 / 60)
    p, _ = h.particles()
    assert p.shape[0] == 250
    assert np.array_equal(p, pos[:250])


def test_pocket_conditioning_zeroes_only_closed_multi_cell_regions(oracle):
    # pressuresolver.cpp:124-268: solid face velocities are zeroed on the six
    # faces of every cell of a fluid region that has more than one cell and
    # borders no air through an open face.  The cavity of pocket_inputs is
    # such a region; the one-cell pocket and the dam (which borders air) are
    # not, so every other face keeps its velocity.
    from tests.helpers import pocket_inputs
    I, J, K = 20, 18, 17
    dx, solid, weights, liquid, mac, valid, svel, cavity, single = pocket_inputs(oracle, I, J, K)
    out = oracle.pressure_solve(I, J, K, dx, 1 / 60, liquid, solid, weights, mac, valid, solid_vel=svel)
    assert out["ok"] == 1
    ks, js, is_ = cavity
    want = [a.copy() for a in svel]
    want[0][ks, js, is_.start:is_.stop + 1] = 0.0
    want[1][ks, js.start:js.stop + 1, is_] = 0.0
    want[2][ks.start:ks.stop + 1, js, is_] = 0.0
    for got, w in zip(out["solid_vel"], want):
        assert np.array_equal(got, w)
    # The conditioned system is consistent: the solve converged to 1e-9 |b|.
    assert out["err"] <= 1e-9 * np.abs(out["rhs"]).max()


def test_domain_transform_closed_form(oracle):
    """With scale 2 and an integer offset every mapping is exact: particles
    loaded at 2 p + o are admitted at p, read back at 2 p + o, and the grid
    box [0, I dx) keeps p = 0 and drops p = I dx (AABB::isPointInside)."""
    I, J, K, dx = 8, 8, 8, 0.125
    p = np.array([[0.0, 0.0, 0.0], [0.5, 0.25, 0.75], [0.9375, 0.5, 0.5], [1.0, 0.5, 0.5], [0.5, 0.5, -0.125]],
                 np.float32)
    off = np.array([1.0, -2.0, 3.0], np.float32)
    pd = (p * np.float32(2.0) + off).astype(np.float32)
    h = oracle.sim(I, J, K, dx)
    h.set_domain(off, 2.0)
    h.load_particles(pd, np.zeros_like(pd))
    h.initialize()
    assert h.num_particles == 3
    inner, _ = h.particles()
    assert np.array_equal(inner, p[:3])
    back, _ = h.particle_range(0, 3)
    assert np.array_equal(back, pd[:3])
    # The reverse map is vmath's: (p - o) * (float)(1.0 / (float)s).
    s = np.float32(0.37)
    inv = np.float32(1.0 / np.float64(s))
    q = np.array([[0.31, 0.17, 0.05]], np.float32)
    h2 = oracle.sim(I, J, K, dx)
    h2.set_domain((0.0, 0.0, 0.0), float(s))
    h2.load_particles(q, np.zeros_like(q))
    h2.initialize()
    got, _ = h2.particles()
    assert np.array_equal(got, (q * inv).astype(np.float32))
