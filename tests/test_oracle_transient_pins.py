"""Pins of the oracle's gradient limiters (navier_gradient.cpp:126-345) and
unsteady terms (navier.cpp:52-141, 700-742; navier_velocity.cpp:107-123;
navier_scalar.cpp:71-90, 290-304) against what the mathematics fixes:
closed-form limiting factors of linear fields on a uniform hex mesh, MLP's
preservation of linear fields, Barth's zero slope at a discrete extremum,
WENO as a convex combination, the time term's diagonal rho V / dt and right
hand side, BDF2's old levels, the time-loop bounds, and the steady pass as
the dt -> infinity limit of a time step."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import conduction_column, params, push, random_state, rel_err
from oracle.pyoracle import Oracle

N = 6  # cavity(6): h = 1/6, cells not touching the boundary have all 8 vertices inside


def linear_gradient(o, g, limiter_case=None):
    """Green-Gauss gradient of the linear field x.g + 0.4 with exact boundary values."""
    z = o.sizes
    x = o.dbls("CELL_X").reshape(-1, 3)
    fx = o.dbls("FACE_X").reshape(-1, 3)
    o.set("P", x @ g + 0.4)
    o.set("BPRE", fx[: z.n_bnd] @ g + 0.4)
    o.gradient("P", "BPRE", "DPDX")
    return o.get("DPDX").reshape(-1, 3), x


def lim_params(limiter):
    p = params()
    p.limiter = limiter
    return p


def inner(x, depth=1):
    h = 1.0 / N
    return np.all((x > depth * h) & (x < 1 - depth * h), axis=1)


def test_barth_closed_form_factors():
    o = Oracle(s.cavity(N), lim_params(1))
    # one-directional slope: interior and y/z-boundary cells keep it, cells on
    # an x face lose it (no neighbour below: umin = 0 against us < 0)
    g, x = linear_gradient(o, np.array([1.0, 0.0, 0.0]))
    h = 1.0 / N
    xface = (x[:, 0] < h) | (x[:, 0] > 1 - h)
    assert np.allclose(g[~xface], [1.0, 0.0, 0.0], rtol=0, atol=1e-12)
    # (to the rounding of the y/z neighbours' values, which equal u_i up to an ulp)
    assert np.all(np.abs(g[xface]) < 1e-14)
    # diagonal slope: neighbours reach u + h, the corners u + 1.5 h -> 2/3
    a = np.array([1.0, 1.0, 1.0])
    g, x = linear_gradient(o, a)
    assert np.allclose(g[inner(x)], 2.0 / 3.0 * a, rtol=1e-13, atol=0)


def test_barth_zeroes_the_slope_at_an_extremum():
    case = s.channel(5, 5, 5, lx=1, ly=1, lz=1)
    o = Oracle(case, lim_params(1))
    z = o.sizes
    x = o.dbls("CELL_X").reshape(-1, 3)
    phi = 0.1 * x[:, 0] + 0.05 * x[:, 1]
    c = int(np.argmin(np.sum((x - 0.5) ** 2, axis=1)))
    phi[c] = 10.0  # a discrete maximum
    o.set("P", phi)
    o.set("BPRE", np.zeros(z.n_bnd))
    o.gradient("P", "BPRE", "DPDX")
    g = o.get("DPDX").reshape(-1, 3)
    assert np.all(g[c] == 0.0)


def test_mlp_keeps_linear_fields_inside():
    o = Oracle(s.cavity(N), lim_params(2))
    a = np.array([0.7, -1.3, 2.1])
    g, x = linear_gradient(o, a)
    # every vertex of these cells has its full ring of 8 cells
    assert np.allclose(g[inner(x)], a, rtol=1e-13, atol=0)
    o0 = Oracle(s.cavity(N), lim_params(0))
    g0, _ = linear_gradient(o0, a)
    assert np.array_equal(g[inner(x)], g0[inner(x)])
    # and only ever scales a gradient down
    assert np.all(np.abs(g) <= np.abs(g0) * (1 + 1e-15))


def test_weno_convex_combination():
    case = s.channel(6, 4, 4, jitter=0.15, split=0.2, seed=3)
    o0, o3 = Oracle(case, lim_params(0)), Oracle(case, lim_params(3))
    z = o0.sizes
    rng = np.random.default_rng(4)
    phi, bphi = rng.normal(size=z.n_cell), rng.normal(size=z.n_bnd)
    for o in (o0, o3):
        o.set("U", phi)
        o.set("BU", bphi)
        o.gradient("U", "BU", "DUDX")
    g0 = o0.get("DUDX").reshape(-1, 3)
    g3 = o3.get("DUDX").reshape(-1, 3)
    nb = o0.ints("CELL_CELL").reshape(z.n_cell, -1)
    for i in range(z.n_cell):
        ring = np.vstack([g0[nb[i][nb[i] >= 0]], g0[i]])
        assert np.all(g3[i] >= ring.min(0) - 1e-14) and np.all(g3[i] <= ring.max(0) + 1e-14)
    # a uniform gradient is left as it is
    o = Oracle(s.cavity(N), lim_params(3))
    a = np.array([0.7, -1.3, 2.1])
    g, _ = linear_gradient(o, a)
    assert np.allclose(g, a, rtol=1e-13, atol=0)


def test_weno_two_cells_known_answer():
    # two cells, one neighbour each: (w1 g1 + w0 g0) / (w1 + w0), w = 1/(|g|^2 + 1e-16)
    case = s.channel(2, 1, 1, lx=2, ly=1, lz=1)
    o0, o3 = Oracle(case, lim_params(0)), Oracle(case, lim_params(3))
    z = o0.sizes
    for o in (o0, o3):
        o.set("U", [1.0, 3.0])
        o.set("BU", np.where(o.dbls("FACE_X").reshape(-1, 3)[: z.n_bnd, 0] < 1, 0.0, 7.0))
        o.gradient("U", "BU", "DUDX")
    g0 = o0.get("DUDX").reshape(-1, 3)
    assert not np.allclose(g0[0], g0[1])  # two different slopes to weigh
    w = 1.0 / (np.sum(g0 ** 2, axis=1) + 1e-16)
    want0 = (w[1] * g0[1] + w[0] * g0[0]) / (w[1] + w[0])
    want1 = (w[0] * g0[0] + w[1] * g0[1]) / (w[0] + w[1])
    g3 = o3.get("DUDX").reshape(-1, 3)
    assert np.allclose(g3[0], want0, rtol=1e-15, atol=1e-15)
    assert np.allclose(g3[1], want1, rtol=1e-15, atol=1e-15)


# ---------------------------------------------------------------- time terms
def unsteady(scheme, dt=0.05, total=1.0, outer=3):
    p = params()
    p.steady = 0
    p.time_scheme = scheme
    p.dt = dt
    p.total_time = total
    p.max_outer_step = outer
    return p


@pytest.mark.parametrize("scheme,factor", [(1, 1.0), (2, 1.5)])
def test_momentum_time_term_diagonal_and_rhs(scheme, factor):
    case = s.cavity(4, jitter=0.1, seed=2)
    ps, pu = params(), unsteady(scheme)
    os_, ou = Oracle(case, ps), Oracle(case, pu)
    z = os_.sizes
    st = random_state(z.n_cell, z.n_bnd, z.n_face, 3)
    for o in (os_, ou):
        push(o, st)
    rng = np.random.default_rng(9)
    unp, unp2 = rng.normal(size=z.n_cell), rng.normal(size=z.n_cell)
    ou.set("UNP", unp)
    ou.set("UNP2", unp2)
    for o in (os_, ou):
        o.call("build_velocity_matrix")
    dg = o.ints("CSR_DIAG")
    coef = st["R"] / pu.dt * os_.dbls("CELL_VOL")
    urf = ps.urf[0]
    dA = (ou.get("AU")[dg] - os_.get("AU")[dg]) * urf
    assert np.allclose(dA, factor * coef, rtol=1e-9, atol=0)
    # the rhs gains the old-level source plus the relaxation share of the larger diagonal
    db = ou.get("BU_RHS") - os_.get("BU_RHS")
    src = coef * unp if scheme == 1 else coef * (2 * unp - 0.5 * unp2)
    assert np.allclose(db, src + (1 - urf) * (dA / urf) * st["U"], rtol=1e-8, atol=1e-12)


def test_bdf2_sets_both_old_levels():
    o = Oracle(s.cavity(3), unsteady(2))
    z = o.sizes
    u = np.random.default_rng(1).normal(size=z.n_cell)
    o.set("U", u)
    o.call("save_transient_old_data")
    assert np.array_equal(o.get("UNP"), u) and np.array_equal(o.get("UNP2"), u)
    o1 = Oracle(s.cavity(3), unsteady(1))
    o1.set("UNP2", np.full(z.n_cell, 7.0))
    o1.set("U", u)
    o1.call("save_transient_old_data")
    assert np.array_equal(o1.get("UNP"), u) and np.all(o1.get("UNP2") == 7.0)


def test_time_loop_bounds():
    # MaxStep = (int)(total_time / dt): 0.35 / 0.1 -> 3 steps
    o = Oracle(s.cavity(3), unsteady(1, dt=0.1, total=0.35, outer=2))
    h, done, _ = o.iterate(10)
    t, n = o.time_state()
    assert (done, n) == (3, 3)
    assert t == 0.1 + 0.1 + 0.1
    # a second call continues, and stops at once
    _, done, _ = o.iterate(5)
    assert done == 0 and o.time_state() == (t, 3)


@pytest.mark.parametrize("scheme", [1, 2])
def test_steady_pass_is_the_long_time_step_limit(scheme):
    case = s.cavity(5, re=100.0, jitter=0.1, seed=1)
    a = Oracle(case, params())
    b = Oracle(case, unsteady(scheme, dt=1e200, total=1e201, outer=2))
    ha, _, _ = a.iterate(3)
    hb, _, _ = b.iterate(3)
    assert np.allclose(hb, ha, rtol=1e-12, atol=1e-15)
    for n in ("U", "V", "W", "P"):
        assert rel_err(b.get(n), a.get(n)) < 1e-12, n


def test_energy_time_term_drives_toward_the_steady_column():
    # implicit Euler on the conduction column: each step moves the mean
    # temperature monotonically from its start toward the steady profile
    case = conduction_column(n=8)
    p = unsteady(1, dt=0.5, total=100.0, outer=2)
    p.solve_energy = 1
    o = Oracle(case, p)
    z = o.sizes
    t_exact = 300.0 + 100.0 * o.dbls("CELL_X").reshape(-1, 3)[:, 2]
    o.set("T", np.full(z.n_cell, 300.0))
    err = []
    for _ in range(6):
        o.iterate(1)
        err.append(np.max(np.abs(o.get("T") - t_exact)))
    assert all(e1 < e0 for e0, e1 in zip(err, err[1:]))
