"""The mixed path's FP32 assembly option (params.assembly_fp32): the momentum,
pressure-correction and scalar coefficients of BuildVelocityMatrix
(navier_velocity.cpp:13-212), BuildPressureMatrix (navier_pressure.cpp:
140-190) and BuildScalarMatrix (navier_scalar.cpp:407-592) evaluated and
accumulated in FP32, against the FP64 oracle on the same seeded state.  The
coefficients agree to FP32 rounding; the pressure-correction diagonal sums
the stored FP32 off-diagonals in FP64, so its rows still sum to zero (here
exactly: a handful of FP32 values of one sign add without rounding in FP64),
and both rows of a face store the same value, so it is exactly symmetric;
the option does nothing in FP64 mode; and a SIMPLE loop converges to the FP64
loop's fields to the FP32 perturbation of its coefficients (compared after a
fixed number of steps, with the residuals still above the level where the
pressure-correction right-hand side is rounding noise and its solve can no
longer reach the tolerance)."""
import numpy as np
import pytest
import scipy.sparse as sp

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "cavity_jitter": lambda: s.cavity(6, jitter=0.15, seed=1),
    "channel_split": lambda: s.channel(8, 5, 6, jitter=0.15, split=0.15, seed=2),
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2, turbulent=True),
}
MOMENTUM = ("AU", "BU_RHS", "BV_RHS", "BW_RHS", "APR")


def fp32_params(turb=0, rtol=1e-12):
    p = params(turb=turb, rtol=rtol, maxit=3000)
    p.precision = s.MIXED
    p.assembly_fp32 = 1
    return p


def seeded(d, o, names, seed):
    z = d.sizes
    st = random_state(z.n_cell, z.n_bnd, z.n_face, seed)
    sel = {k: st[k] for k in names}
    push(d, sel)
    push(o, sel)


def csr(d, name):
    z = d.sizes
    return sp.csr_matrix((d.get(name), d.ints("CSR_COL"), d.ints("CSR_ROWPTR")), shape=(z.n_cell, z.n_cell))


def test_option_does_nothing_in_fp64():
    case = CASES["channel_split"]()
    p = params()
    p.assembly_fp32 = 1
    d, o = s.Solver(case, p), Oracle(case, p)
    seeded(d, o, ["U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE", "APR", "DELTAP"], 4)
    for k in ("build_velocity_matrix", "cal_ru_face", "build_pressure_matrix"):
        d.call(k)
        o.call(k)
    for n in ("AU", "BU_RHS", "AP", "BP_RHS"):
        assert rel_err(d.get(n), o.get(n)) <= 1e-13, n


@pytest.mark.parametrize("name", sorted(CASES))
def test_momentum_and_pressure_coefficients_to_fp32_rounding(name):
    case = CASES[name]()
    d, o = s.Solver(case, fp32_params()), Oracle(case, params())
    seeded(d, o, ["U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE"], 4)
    d.call("build_velocity_matrix")
    o.call("build_velocity_matrix")
    errs = {n: rel_err(d.get(n), o.get(n)) for n in MOMENTUM}
    print(f"[{name}] momentum", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for n, e in errs.items():
        assert 0.0 < e <= 2e-6, f"{n}: rel err {e:.3e}"
    # the FP32 coefficients are what is stored: FP32 values in FP64 arrays
    au = d.get("AU")
    assert np.array_equal(au, au.astype(np.float32).astype(np.float64))
    # the pressure correction from the oracle's (FP64) face fluxes
    push(d, {k: o.get(k) for k in ("APR", "DPDX")})
    for x in (d, o):
        x.call("cal_ru_face")
    push(d, {"RUFACE": o.get("RUFACE")})
    for x in (d, o):
        x.call("build_pressure_matrix")
    for n in ("AP", "BP_RHS"):
        e = rel_err(d.get(n), o.get(n))
        print(f"[{name}] {n} rel err {e:.2e}", flush=True)
        assert e <= 2e-6, f"{n}: rel err {e:.3e}"
    A = csr(d, "AP")
    off = A - sp.diags(A.diagonal())
    assert np.array_equal(off.data, off.data.astype(np.float32).astype(np.float64))
    assert np.all(A.diagonal() < 0)
    # both rows of a face store one value (the singular system of a closed
    # domain stays consistent)
    assert (A - A.T).count_nonzero() == 0
    rows = np.abs(A @ np.ones(A.shape[0]))
    if case.meta["kind"] == "cavity":
        assert rows.max() == 0.0
    else:
        assert rows.max() <= 1e-6 * np.abs(A.data).max()


def test_scalar_coefficients_to_fp32_rounding():
    case = CASES["step_tets"]()
    d, o = s.Solver(case, fp32_params(turb=1)), Oracle(case, params(turb=1))
    seeded(d, o, ["U", "V", "W", "R", "VISLAM", "VISTUR", "TE", "ED", "DUDX", "DVDX", "DWDX", "RUFACE"], 7)
    ru = o.get("RUFACE")
    z = d.sizes
    inlet = np.isin(d.ints("BND_RID"), [r.rid for r in case.regions if r.type1 == 2])
    ru[: z.n_bnd][inlet] = -np.abs(ru[: z.n_bnd][inlet])
    push(d, {"RUFACE": ru})
    push(o, {"RUFACE": ru})
    for x in (d, o):
        x.call("kepsilon_sources")
        x.call("bc_kepsilon")
    push(d, {k: o.get(k) for k in ("ED", "BED", "KE_APTE")})
    for isca in (2, 3):
        for x in (d, o):
            x.call("build_scalar_matrix", isca)
        for n in ("AS", "BS_RHS"):
            e = rel_err(d.get(n), o.get(n))
            print(f"[isca {isca}] {n} rel err {e:.2e}", flush=True)
            assert 0.0 < e <= 2e-6, f"{n}: rel err {e:.3e}"


@pytest.mark.parametrize("case_fn,steps,turb,tol", [
    (lambda: s.cavity(8, re=100, jitter=0.1, seed=2), 60, 0, 1e-5),
    (lambda: s.channel(10, 6, 8, jitter=0.15, split=0.1, seed=1, re=200), 40, 0, 1e-5),
    (lambda: s.step(2, upstream=2, downstream=6, span=2, turbulent=True, re=5000), 5, 1, 1e-4),
])
def test_simple_loop_converges_to_the_fp64_fields(case_fn, steps, turb, tol):
    case = case_fn()
    ref = params(turb=turb, rtol=1e-9, maxit=3000)
    p32 = fp32_params(turb, rtol=1e-9)
    ref.residual_steady = p32.residual_steady = 0.0
    d64, d32 = s.Solver(case, ref), s.Solver(case, p32)
    h64, n64 = d64.iterate(steps)
    h32, n32 = d32.iterate(steps)
    print(f"steps {n64} / {n32}, last residuals {h64[-1, 1:6]} / {h32[-1, 1:6]}", flush=True)
    names = ["U", "V", "W", "P"] + (["TE", "ED"] if turb else [])
    errs = {n: rel_err(d32.get(n), d64.get(n)) for n in names}
    print({k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for n, e in errs.items():
        assert e <= tol, f"{n}: rel err {e:.3e}"
