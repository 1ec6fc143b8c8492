"""Mixed precision (FP32 inner Krylov solves inside FP64 iterative
refinement) against the FP64 device path and the FP64 oracle."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def l2_rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixed(p):
    p.precision = s.MIXED
    p.ir_max = 50
    p.ir_inner_rtol = 1e-3
    return p


@pytest.mark.parametrize("ksp", [s.KSP_GMRES, s.KSP_CG])
def test_mixed_pressure_solve_reaches_fp64_tolerance(ksp):
    case = s.cavity(10, jitter=0.1, seed=5)
    d64 = s.Solver(case, params(restart=30, ksp=ksp))
    d32 = s.Solver(case, mixed(params(restart=30, ksp=ksp)))
    st = random_state(d64.sizes.n_cell, d64.sizes.n_bnd, d64.sizes.n_face, 7)
    for d in (d64, d32):
        push(d, {k: st[k] for k in ("U", "V", "W", "P", "R", "APR", "DPDX", "RUFACE")})
        for k in ("bc_velocity", "cal_ru_face", "build_pressure_matrix"):
            d.call(k)
    a = d64.solve_pressure(1e-10, 4000)
    b = d32.solve_pressure(1e-10, 4000)
    assert a["reason"] > 0 and b["reason"] > 0
    assert b["refinements"] >= 2
    # both converge the true preconditioned residual to 1e-10 of ||M^-1 b||
    assert b["rnorm"] <= 1e-10 * b["rnorm0"] * 1.0000001
    assert rel_err(d32.get("DELTAP"), d64.get("DELTAP")) < 1e-7


def test_mixed_velocity_solve_matches_oracle():
    case = s.channel(12, 8, 10, jitter=0.12, split=0.1, seed=3)
    p = mixed(params(restart=30))
    d = s.Solver(case, p)
    o = Oracle(case, params(restart=30))
    st = random_state(d.sizes.n_cell, d.sizes.n_bnd, d.sizes.n_face, 6)
    for x in (d, o):
        push(x, {k: st[k] for k in ("U", "V", "W", "P", "R", "VISLAM", "RUFACE")})
        x.call("build_velocity_matrix")
    info = d.solve_velocity(1e-10, 2000)
    o.solve_velocity(1e-10, 2000)
    assert all(i["reason"] > 0 for i in info)
    for n in "UVW":
        assert rel_err(d.get(n), o.get(n)) < 1e-8


@pytest.mark.parametrize("case_fn,steps", [
    (lambda: s.cavity(12, re=100), 30),
    (lambda: s.channel(16, 8, 12, jitter=0.15, split=0.1, seed=1, re=1000), 15),
])
def test_mixed_simple_fields_within_1e6_of_fp64(case_fn, steps):
    case = case_fn()
    p64 = params(restart=50, rtol=1e-8)
    p64.ksp[s.KSP_PRESSURE].rtol = 1e-6
    d64 = s.Solver(case, p64)
    p32 = mixed(params(restart=50, rtol=1e-8))
    p32.ksp[s.KSP_PRESSURE].rtol = 1e-6
    d32 = s.Solver(case, p32)
    h64, _ = d64.iterate(steps)
    h32, _ = d32.iterate(steps)
    errs = {n: l2_rel(d32.get(n), d64.get(n)) for n in ("U", "V", "W", "P")}
    print("mixed vs fp64 relative L2 field error:", errs)
    assert max(errs.values()) < 1e-6, errs
    assert np.allclose(h32[:, :4], h64[:, :4], rtol=1e-3, atol=1e-12)
