"""GPU parity: every device kernel against the FP64 CPU oracle on the same
seeded inputs.  Assembly, gradient and flux kernels evaluate the same
expressions in the same order, so they agree to the last bit or to rounding
of a different reduction order (outlet mass-flow sums); Krylov solves agree
to their tolerance (PETSc's reduction order is itself unspecified)."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import hybrid_channel, params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def pair(case, p):
    d = s.Solver(case, p)
    o = Oracle(case, p)
    return d, o


def same_state(d, o, names, seed=0):
    z = d.sizes
    st = random_state(z.n_cell, z.n_bnd, z.n_face, seed)
    sel = {k: st[k] for k in names}
    push(d, sel)
    push(o, sel)


def assert_close(d, o, names, tol):
    for n in names:
        a, b = d.get(n), o.get(n)
        e = rel_err(a, b)
        assert e <= tol, f"{n}: rel err {e:.3e} > {tol:.1e}"


CASES = {
    "cavity_jitter": lambda: s.cavity(6, jitter=0.15, seed=1),
    "channel_split": lambda: s.channel(8, 5, 6, jitter=0.15, split=0.15, seed=2),
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2, turbulent=True),
    "hybrid_prisms_pyramids": lambda: hybrid_channel(6, 4, 5, prism_frac=0.3, pyramid_frac=0.3, jitter=0.15, seed=3),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gradient_bitwise(name):
    d, o = pair(CASES[name](), params())
    same_state(d, o, ["U", "V", "W", "P", "BU", "BV", "BW", "BPRE"], seed=3)
    for f, b, g in [("P", "BPRE", "DPDX"), ("U", "BU", "DUDX"), ("TE", "BTE", "DPHIDX")]:
        d.gradient(f, b, g)
        o.gradient(f, b, g)
        assert np.array_equal(d.get(g), o.get(g)), f"{g}: rel err {rel_err(d.get(g), o.get(g)):.3e}"


@pytest.mark.parametrize("name", sorted(CASES))
def test_least_squares_gradient_bitwise(name):
    p = params()
    p.grad_scheme = 1
    d, o = pair(CASES[name](), p)
    same_state(d, o, ["U", "V", "W", "P", "BU", "BV", "BW", "BPRE"], seed=3)
    for f, b, g in [("P", "BPRE", "DPDX"), ("U", "BU", "DUDX")]:
        d.gradient(f, b, g)
        o.gradient(f, b, g)
        assert np.array_equal(d.get(g), o.get(g)), f"{g}: rel err {rel_err(d.get(g), o.get(g)):.3e}"
    # the four velocity/pressure gradients of the momentum assembly, fused
    d.call("build_velocity_matrix")
    o.call("build_velocity_matrix")
    assert_close(d, o, ["DUDX", "DVDX", "DWDX", "DPDX", "AU", "BU_RHS"], 1e-13)


@pytest.mark.parametrize("case_fn,steps", [
    (lambda: s.cavity(8, re=100, jitter=0.1, seed=2), 15),
    (lambda: s.channel(10, 6, 8, jitter=0.15, split=0.1, seed=1, re=200), 10),
])
def test_simple_loop_least_squares_gradients(case_fn, steps):
    p = params(rtol=1e-12)
    p.grad_scheme = 1
    d, o, hd, ho = _loop_pair(case_fn(), p, steps)
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P"], 1e-7)


@pytest.mark.parametrize("name", sorted(CASES))
def test_build_velocity_matrix(name):
    d, o = pair(CASES[name](), params())
    same_state(d, o, ["U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE"], seed=4)
    d.call("build_velocity_matrix")
    o.call("build_velocity_matrix")
    d.call("synchronize")
    # the outlet rescale sums the boundary mass flow in a different order, so
    # cases with an outlet agree to rounding instead of bitwise
    tol = 1e-13 if d.case.meta["kind"] in ("channel", "step") else 0.0
    assert_close(d, o, ["BRO", "BU", "BV", "BW", "BPRE", "DUDX", "DVDX", "DWDX", "DPDX", "AU", "BU_RHS", "BV_RHS",
                        "BW_RHS", "APR", "SHEAR"], tol)


@pytest.mark.parametrize("name", sorted(CASES))
def test_flux_pressure_matrix_and_corrections(name):
    d, o = pair(CASES[name](), params())
    same_state(d, o, ["U", "V", "W", "P", "R", "APR", "DPDX", "BU", "BV", "BW", "BRO", "DELTAP"], seed=5)
    for k in ("cal_ru_face", "build_pressure_matrix"):
        d.call(k)
        o.call(k)
    assert_close(d, o, ["RUFACE", "AP", "BP_RHS"], 0.0)
    rd, ro = d.correct_fields(), o.correct_fields()
    assert rd == pytest.approx(ro, rel=1e-13)
    assert_close(d, o, ["BPRE", "DPDX", "P", "U", "V", "W"], 0.0)
    d.call("correct_ru_face")
    o.call("correct_ru_face")
    assert_close(d, o, ["RUFACE"], 0.0)


@pytest.mark.parametrize("pc", [s.PC_JACOBI, s.PC_NONE, s.PC_ILU0])
def test_velocity_solve_matches_oracle(pc):
    case = s.channel(10, 6, 8, jitter=0.1, seed=3)
    p = params(restart=20, pc=pc)
    d, o = pair(case, p)
    same_state(d, o, ["U", "V", "W", "P", "R", "VISLAM", "RUFACE"], seed=6)
    d.call("build_velocity_matrix")
    o.call("build_velocity_matrix")
    idd = d.solve_velocity(1e-10, 1000)
    ido = o.solve_velocity(1e-10, 1000)
    for a, b in zip(idd, ido):
        assert a["reason"] > 0 and b["reason"] > 0
        assert abs(a["iters"] - b["iters"]) <= 1
        assert a["rnorm0"] == pytest.approx(b["rnorm0"], rel=1e-12)
    assert_close(d, o, ["U", "V", "W"], 1e-9)


@pytest.mark.parametrize("ksp,pc", [(s.KSP_GMRES, s.PC_JACOBI), (s.KSP_CG, s.PC_JACOBI), (s.KSP_GMRES, s.PC_ILU0),
                                     (s.KSP_CG, s.PC_ILU0)])
def test_pressure_solve_matches_oracle(ksp, pc):
    case = s.cavity(8, jitter=0.1, seed=5)
    p = params(restart=30, ksp=ksp, pc=pc)
    d, o = pair(case, p)
    same_state(d, o, ["U", "V", "W", "P", "R", "APR", "DPDX", "RUFACE"], seed=7)
    for k in ("bc_velocity", "cal_ru_face", "build_pressure_matrix"):
        d.call(k)
        o.call(k)
    a = d.solve_pressure(1e-10, 2000)
    b = o.solve_pressure(1e-10, 2000)
    assert a["reason"] > 0 and b["reason"] > 0
    assert abs(a["iters"] - b["iters"]) <= 2
    assert_close(d, o, ["DELTAP"], 1e-8)


@pytest.mark.parametrize("graphs", [0, 1])
@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
@pytest.mark.parametrize("restart", [1, 2, 64])
def test_restart_length_extremes_match_oracle(restart, prec, graphs):
    """GMRES(1), GMRES(2) and GMRES(64), the longest cycle the workspace
    holds (a 65-column Hessenberg matrix, one wavefront's worth of back
    substitution), on a pressure correction that needs several cycles;
    launched kernel by kernel and as captured cycles."""
    case = s.cavity(10, jitter=0.1, seed=8)
    p = params(restart=restart, pc=s.PC_JACOBI)
    p.precision = prec
    p.graphs = graphs
    d, o = pair(case, p)
    same_state(d, o, ["U", "V", "W", "P", "R", "APR", "DPDX", "RUFACE"], seed=9)
    for k in ("bc_velocity", "cal_ru_face", "build_pressure_matrix"):
        d.call(k)
        o.call(k)
    a = d.solve_pressure(1e-8, 20000)
    b = o.solve_pressure(1e-8, 20000)
    assert a["reason"] > 0 and b["reason"] > 0
    assert b["iters"] > 64  # more than one full cycle of the longest restart
    if prec == s.FP64:
        assert abs(a["iters"] - b["iters"]) <= max(2, b["iters"] // 100)
        assert_close(d, o, ["DELTAP"], 1e-6)
    else:
        assert_close(d, o, ["DELTAP"], 1e-5)


def test_solve_cap_raises_not_converged():
    d = s.Solver(s.cavity(6), params(restart=5, pc=s.PC_NONE))
    st = random_state(d.sizes.n_cell, d.sizes.n_bnd, d.sizes.n_face, 1)
    push(d, {k: st[k] for k in ("U", "V", "W", "P", "RUFACE")})
    d.call("build_velocity_matrix")
    with pytest.raises(s.NotConverged):
        d.solve_velocity(1e-14, 2)


def test_bad_arguments_are_errors():
    d = s.Solver(s.cavity(3), params())
    with pytest.raises(s.S355Error):
        d.gradient("U", "BU", "U")  # output must be a 3-vector field
    with pytest.raises(s.S355Error):
        d.set("U", np.zeros(5))


def _loop_pair(case, p, n):
    d, o = pair(case, p)
    hd, nd = d.iterate(n)
    ho, no, _ = o.iterate(n)
    return d, o, hd, ho


def test_simple_loop_cavity_fp64():
    d, o, hd, ho = _loop_pair(s.cavity(8, re=100), params(rtol=1e-12), 25)
    assert hd.shape == ho.shape == (25, 11)
    assert np.allclose(hd, ho, rtol=1e-7, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P", "RUFACE"], 1e-8)


def test_simple_loop_channel_split_fp64():
    d, o, hd, ho = _loop_pair(s.channel(10, 6, 8, jitter=0.15, split=0.1, seed=1, re=200), params(rtol=1e-12), 10)
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P"], 1e-7)


def test_simple_loop_hybrid_prisms_pyramids_fp64():
    case = hybrid_channel(8, 5, 6, prism_frac=0.3, pyramid_frac=0.3, jitter=0.1, seed=4, re=200.0)
    d, o, hd, ho = _loop_pair(case, params(rtol=1e-12), 10)
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P"], 1e-7)


def test_simple_loop_step_kepsilon_fp64():
    case = s.step(2, upstream=2, downstream=6, span=2, turbulent=True, re=5000)
    p = params(rtol=1e-12, turb=1)
    d, o, hd, ho = _loop_pair(case, p, 5)
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P", "TE", "ED", "VISTUR"], 1e-6)


def test_step_tets_kepsilon_converges_on_device_and_oracle():
    # the backward-facing-step workload's form (jittered Kuhn tetrahedra,
    # k-epsilon, Re 5000) with the pressure correction relaxed by 0.3: the
    # outer residual falls on both paths, which agree step by step
    case = s.step(2, upstream=2, downstream=6, span=2, turbulent=True, re=5000)
    p = params(rtol=1e-12, turb=1, maxit=5000)
    p.residual_steady = 0.0
    p.urf[3] = 0.3
    d, o, hd, ho = _loop_pair(case, p, 30)
    for h in (hd, ho):
        assert h[-1, 10] < 0.05 * h[:, 10].max()
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    assert_close(d, o, ["U", "V", "W", "P", "TE", "ED", "VISTUR"], 1e-6)


def test_reference_tolerances_agree_to_solve_accuracy():
    # the reference's own settings: GMRES(30)+Jacobi, rtol 1e-8 everywhere (config 1 form)
    d, o, hd, ho = _loop_pair(s.cavity(10, re=100), params(rtol=1e-8), 20)
    for n in ("U", "V", "W"):
        assert rel_err(d.get(n), o.get(n)) < 1e-5


def test_bitwise_rerun_determinism():
    case = s.channel(8, 5, 6, jitter=0.1, split=0.1, seed=9)
    runs = []
    for _ in range(2):
        d = s.Solver(case, params(rtol=1e-9))
        d.iterate(4)
        runs.append([d.get(n) for n in ("U", "V", "W", "P", "RUFACE")])
        d.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def test_log_stages_accumulate():
    d = s.Solver(s.cavity(6), params())
    d.stage_times(reset=True)
    d.iterate(2)
    t = d.stage_times()
    assert t[0] > 0 and all(t[k] > 0 for k in (1, 2, 3, 4))
    assert t[0] == pytest.approx(t[1:].sum(), rel=1e-9)
