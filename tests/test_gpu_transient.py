"""Gradient limiters and unsteady runs on the GPU against the oracle: the
limited gradients (Barth, MLP, WENO; navier_gradient.cpp:126-345) and the
momentum time term (navier_velocity.cpp:107-123) to the last bit, and
time-stepped SIMPLE runs (Euler and BDF2; laminar, k-epsilon, energy) to the
tolerance of their Krylov solves."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import heated, hybrid_channel, params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "cavity_jitter": lambda: s.cavity(6, jitter=0.15, seed=1),
    "channel_split": lambda: s.channel(8, 5, 6, jitter=0.15, split=0.15, seed=2),
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2, turbulent=True),
    "hybrid_prisms_pyramids": lambda: hybrid_channel(6, 4, 5, prism_frac=0.3, pyramid_frac=0.3, jitter=0.15, seed=5),
}


def unsteady(p, scheme, dt=0.05, total=10.0, outer=4):
    p.steady = 0
    p.time_scheme = scheme
    p.dt = dt
    p.total_time = total
    p.max_outer_step = outer
    return p


@pytest.mark.parametrize("limiter", [1, 2, 3])
@pytest.mark.parametrize("name", sorted(CASES))
def test_limited_gradient_bitwise(name, limiter):
    case = CASES[name]()
    p = params()
    p.limiter = limiter
    with s.Solver(case, p) as d:
        o = Oracle(case, p)
        z = d.sizes
        st = random_state(z.n_cell, z.n_bnd, z.n_face, 7)
        sel = {k: st[k] for k in ("U", "P", "BU", "BPRE")}
        push(d, sel)
        push(o, sel)
        for x in (d, o):
            x.gradient("U", "BU", "DUDX")
            x.gradient("P", "BPRE", "DPDX")
        for g in ("DUDX", "DPDX"):
            a, b = d.get(g), o.get(g)
            assert np.array_equal(a, b), f"{g}: rel err {rel_err(a, b):.3e}"
        # the limiter acted
        p0 = params()
        o0 = Oracle(case, p0)
        push(o0, sel)
        o0.gradient("U", "BU", "DUDX")
        assert not np.array_equal(o0.get("DUDX"), d.get("DUDX"))


@pytest.mark.parametrize("scheme", [1, 2])
def test_momentum_time_term_bitwise(scheme):
    case = s.cavity(6, jitter=0.15, seed=1)
    p = unsteady(params(), scheme)
    with s.Solver(case, p) as d:
        o = Oracle(case, p)
        z = d.sizes
        st = random_state(z.n_cell, z.n_bnd, z.n_face, 4)
        sel = {k: st[k] for k in ("U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE")}
        push(d, sel)
        push(o, sel)
        with pytest.raises(s.S355Error):
            d.get("UNP")  # no old level before the first save
        d.call("save_transient_old_data")
        o.call("save_transient_old_data")
        rng = np.random.default_rng(3)
        old = {n: rng.normal(size=z.n_cell) for n in ("UNP", "VNP", "WNP", "UNP2", "VNP2", "WNP2")}
        push(d, old)
        push(o, old)
        for x in (d, o):
            x.call("build_velocity_matrix")
        for n in ("AU", "BU_RHS", "BV_RHS", "BW_RHS", "APR"):
            a, b = d.get(n), o.get(n)
            assert np.array_equal(a, b), f"{n}: rel err {rel_err(a, b):.3e}"


def test_save_old_levels():
    case = s.step(2, upstream=2, downstream=4, span=2, turbulent=True)
    for scheme in (1, 2):
        p = unsteady(params(turb=1), scheme)
        with s.Solver(case, p) as d:
            d.call("save_transient_old_data")
            for cur, old in (("U", "UNP"), ("W", "WNP"), ("T", "TNP"), ("TE", "TEP"), ("ED", "EDP")):
                assert np.array_equal(d.get(old), d.get(cur)), old
                if scheme == 2:
                    assert np.array_equal(d.get(old + "2"), d.get(cur)), old + "2"


def loop_pair(case, p, steps):
    with s.Solver(case, p) as d:
        o = Oracle(case, p)
        hd, nd = d.iterate(steps)
        ho, no, _ = o.iterate(steps)
        fields = {n: (d.get(n), o.get(n)) for n in ("U", "V", "W", "P", "T", "TE", "ED", "VISTUR")}
        td, to = d.time_state(), o.time_state()
    return hd, ho, nd, no, fields, td, to


@pytest.mark.parametrize("scheme", [1, 2])
def test_unsteady_cavity(scheme):
    p = unsteady(params(rtol=1e-12), scheme, dt=0.02, outer=4)
    hd, ho, nd, no, f, td, to = loop_pair(s.cavity(8, re=100.0, jitter=0.1, seed=3), p, 4)
    assert nd == no == 4 and td == to and td[1] == 4
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    for n in ("U", "V", "W", "P"):
        assert rel_err(*f[n]) < 1e-8, n


def test_unsteady_kepsilon_step():
    case = s.step(2, upstream=2, downstream=6, span=2, turbulent=True, re=5000)
    p = unsteady(params(rtol=1e-12, turb=1), 1, dt=0.05, outer=3)
    hd, ho, nd, no, f, td, to = loop_pair(case, p, 3)
    assert nd == no == 3 and td == to
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    for n in ("U", "V", "W", "P", "TE", "ED", "VISTUR"):
        assert rel_err(*f[n]) < 1e-6, n


def test_unsteady_energy_with_limiter():
    case = heated(s.channel(10, 5, 6, jitter=0.1, seed=3, re=400.0))
    p = unsteady(params(rtol=1e-12), 2, dt=0.1, outer=3)
    p.solve_energy = 1
    p.limiter = 2
    hd, ho, nd, no, f, td, to = loop_pair(case, p, 3)
    assert nd == no == 3
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    for n in ("U", "V", "W", "P", "T"):
        assert rel_err(*f[n]) < 1e-7, n


@pytest.mark.parametrize("limiter", [1, 3])
def test_steady_loop_with_limiter(limiter):
    p = params(rtol=1e-12)
    p.limiter = limiter
    hd, ho, nd, no, f, _, _ = loop_pair(s.cavity(8, re=100.0, jitter=0.1, seed=2), p, 8)
    assert nd == no == 8
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    for n in ("U", "V", "W", "P"):
        assert rel_err(*f[n]) < 1e-7, n
