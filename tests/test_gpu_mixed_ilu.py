"""Mixed precision with the ILU(0) preconditioner (the 20 M-cell k-epsilon
channel's solver form: FP32 Krylov work with FP32 ILU(0) factors inside
FP64 iterative refinement) against the FP64 device path and the FP64
oracle: single solves with the stopping test made explicit on the host
(||M^-1 (b - A x)|| <= rtol ||M^-1 b|| with the oracle's FP64 ILU(0)), and a
small run of the same form as the 20 M-cell case (jittered hex channel,
k-epsilon, GMRES(50) + ILU(0) on every solve, captured ILU sweeps)."""
import numpy as np
import pytest
import scipy.sparse as sp

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state, rel_err
from oracle.pyoracle import Oracle, pc_apply

pytestmark = pytest.mark.gpu


def l2_rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def mixed(p):
    p.precision = s.MIXED
    p.ir_max = 50
    p.ir_inner_rtol = 1e-3
    return p


def csr(x, name):
    z = x.sizes
    return sp.csr_matrix((x.get(name), x.ints("CSR_COL"), x.ints("CSR_ROWPTR")), shape=(z.n_cell, z.n_cell))


def pc_norm(x, name, r):
    return float(np.linalg.norm(pc_apply(x.ints("CSR_ROWPTR"), x.ints("CSR_COL"), x.get(name), s.PC_ILU0, r)))


PRESSURE_CASES = {
    "cavity_jitter": lambda: s.cavity(10, jitter=0.1, seed=5),
    "channel_split": lambda: s.channel(12, 7, 8, jitter=0.12, split=0.25, seed=6, re=400.0),
}


@pytest.mark.parametrize("name", sorted(PRESSURE_CASES))
@pytest.mark.parametrize("ksp", [s.KSP_GMRES, s.KSP_CG])
def test_mixed_ilu0_pressure_solve(name, ksp):
    case = PRESSURE_CASES[name]()
    p = params(restart=30, ksp=ksp, pc=s.PC_ILU0)
    d64, d32, o = s.Solver(case, p), s.Solver(case, mixed(params(restart=30, ksp=ksp, pc=s.PC_ILU0))), Oracle(case, p)
    st = random_state(d64.sizes.n_cell, d64.sizes.n_bnd, d64.sizes.n_face, 7)
    for x in (d64, d32, o):
        push(x, {k: st[k] for k in ("U", "V", "W", "P", "R", "APR", "DPDX", "RUFACE")})
        for k in ("bc_velocity", "cal_ru_face", "build_pressure_matrix"):
            x.call(k)
    rtol = 1e-10
    a = d64.solve_pressure(rtol, 4000)
    b = d32.solve_pressure(rtol, 4000)
    c = o.solve_pressure(rtol, 4000)
    assert a["reason"] > 0 and b["reason"] > 0 and c["reason"] > 0, (a, b, c)
    assert b["refinements"] >= 1
    A, rhs = csr(o, "AP"), o.get("BP_RHS")
    x32 = d32.get("DELTAP")
    # the stopping test of the mixed solve, on the FP64 ILU(0) norm: the
    # outer test used the FP32 factors, which move the norm by far less than 1 %
    assert pc_norm(o, "AP", rhs - A @ x32) <= rtol * pc_norm(o, "AP", rhs) * 1.01
    # the same solution as the FP64 path and the oracle, up to the null space
    # of the (singular, Neumann) pressure-correction matrix
    def centred(v):
        return v - v.mean()
    ref = centred(o.get("DELTAP"))
    assert l2_rel(centred(d64.get("DELTAP")), ref) < 1e-7
    assert l2_rel(centred(x32), ref) < 1e-7


def test_mixed_ilu0_velocity_solve_matches_oracle():
    case = s.channel(12, 8, 10, jitter=0.12, split=0.1, seed=3)
    d = s.Solver(case, mixed(params(restart=30, pc=s.PC_ILU0)))
    o = Oracle(case, params(restart=30, pc=s.PC_ILU0))
    st = random_state(d.sizes.n_cell, d.sizes.n_bnd, d.sizes.n_face, 6)
    for x in (d, o):
        push(x, {k: st[k] for k in ("U", "V", "W", "P", "R", "VISLAM", "RUFACE")})
        x.call("build_velocity_matrix")
    info = d.solve_velocity(1e-10, 2000)
    o.solve_velocity(1e-10, 2000)
    assert all(i["reason"] > 0 and i["refinements"] >= 1 for i in info)
    for n in "UVW":
        assert rel_err(d.get(n), o.get(n)) < 1e-8


def turbulent_channel_form(nx=20, ny=10, nz=12):
    # the 20 M-cell case's form at 2,400 cells: 2 x 1 x 2 box, 10 % jitter,
    # k-epsilon, Re_b = 13750
    return s.channel(nx, ny, nz, lx=2.0, ly=1.0, lz=2.0, jitter=0.10, split=0.0, seed=0, re=13750.0, turbulent=True)


def turbulent_params(prec, rtol, graphs=1):
    p = s.set_all_ksp(s.default_params(), s.KSP_GMRES, 50, s.PC_ILU0, rtol=rtol, maxit=20000)
    p.turb_model = 1
    p.residual_steady = 0.0
    p.graphs = graphs
    return mixed(p) if prec == s.MIXED else p


def test_kepsilon_channel_mixed_ilu0_against_fp64_and_oracle():
    case = turbulent_channel_form()
    steps = 4
    names = ("U", "V", "W", "P", "TE", "ED", "VISTUR")
    runs = {}
    for key, prec in (("fp64", s.FP64), ("mixed", s.MIXED)):
        d = s.Solver(case, turbulent_params(prec, 1e-12))
        h, done = d.iterate(steps)
        assert done == steps
        info = d.last_solve_info()
        assert all(info[n]["reason"] > 0 for n in ("U", "V", "W", "P", "TE", "ED"))
        if prec == s.MIXED:
            assert info["P"]["refinements"] >= 1
        runs[key] = ({n: d.get(n) for n in names}, h)
        d.close()
    o = Oracle(case, turbulent_params(s.FP64, 1e-12, graphs=0))
    ho, done, rc = o.iterate(steps)
    assert rc == 0 and done == steps
    fo = {n: o.get(n) for n in names}
    err_mixed = {n: l2_rel(runs["mixed"][0][n], runs["fp64"][0][n]) for n in names}
    err_fp64 = {n: l2_rel(runs["fp64"][0][n], fo[n]) for n in names}
    print("mixed vs FP64 device:", err_mixed)
    print("FP64 device vs oracle:", err_fp64)
    assert max(err_mixed.values()) <= 1e-6, err_mixed
    assert max(err_fp64.values()) <= 1e-8, err_fp64
    # the residual histories (res[0..3], ResMax) agree as well
    for key in ("fp64", "mixed"):
        h = runs[key][1]
        assert np.allclose(h[:, [0, 1, 2, 3, 10]], ho[:, [0, 1, 2, 3, 10]], rtol=1e-5, atol=1e-14), key
