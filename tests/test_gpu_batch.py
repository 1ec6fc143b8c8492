"""Batched cases on the GPU: one BatchSolver holding independent cases of
uneven size and Reynolds number against one Solver per case.  Every case of
the batch is reduced over the same tiles, relative to its own first row, as
when it runs alone, so the batch reproduces the single-case runs bit for
bit (FP64 and mixed, Jacobi and ILU(0), the mixed cycle in the natural and
in the input cell order)."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params

pytestmark = pytest.mark.gpu


def sweep_cases():
    return [s.cavity(6, re=100.0), s.cavity(9, re=400.0, jitter=0.1, seed=3),
            s.channel(10, 5, 6, jitter=0.1, split=0.2, seed=4, re=500.0), s.cavity(4, re=1000.0, tets=True)]


def run_alone(cases, p, steps):
    out = []
    for c in cases:
        d = s.Solver(c, p)
        h, done = d.iterate(steps)
        out.append(({n: d.get(n) for n in ("U", "V", "W", "P", "RUFACE", "BU")}, h, done))
        d.close()
    return out


@pytest.mark.parametrize("prec,pc,order", [(s.FP64, s.PC_JACOBI, 0), (s.MIXED, s.PC_JACOBI, 0), (s.FP64, s.PC_ILU0, 0),
                                           (s.MIXED, s.PC_JACOBI, 2)])
def test_batch_reproduces_single_case_runs(prec, pc, order):
    # order 2: each case's FP32 cycle on its rows in the input cell order, in
    # its own position range of the batch
    cases = sweep_cases()
    p = params(restart=30, pc=pc, rtol=1e-8, maxit=2000)
    p.ksp[s.KSP_PRESSURE].rtol = 1e-6
    p.residual_steady = 0.0
    p.precision = prec
    p.inner_order = order
    steps = 6
    alone = run_alone(cases, p, steps)
    b = s.BatchSolver(cases, p)
    hb, done = b.iterate(steps)
    assert done == steps
    for k, (fields, h, dk) in enumerate(alone):
        assert dk == steps
        for n in ("U", "V", "W", "P"):
            assert np.array_equal(b.case_field(k, n), fields[n]), (k, n)
        assert np.array_equal(b.case_field(k, "BU"), fields["BU"]), k
    # the batch history is the maximum over the cases, step by step
    hmax = np.max(np.stack([a[1] for a in alone]), axis=0)
    assert np.array_equal(hb[:, :4], hmax[:, :4])
    info = b.case_solve_info()
    assert all(i["reason"] > 0 for i in info["P"])
    assert b.kernel_shapes()["fp32"]["order"] == (1 if order else 0)
    b.close()


def test_batch_per_case_solve_records():
    cases = sweep_cases()
    p = params(restart=30, rtol=1e-8, maxit=2000)
    p.ksp[s.KSP_PRESSURE].rtol = 1e-6
    b = s.BatchSolver(cases, p)
    b.iterate(2)
    info = b.case_solve_info()
    its = [i["iters"] for i in info["P"]]
    assert len(its) == len(cases) and min(its) > 0
    # uneven cases converge after different iteration counts
    assert len(set(its)) > 1
    for i in info["P"]:
        assert i["rnorm"] <= 1e-6 * i["rnorm0"] * (1 + 1e-12)
    r = b.case_residuals()
    assert r.shape == (len(cases), 10)


def test_batch_rejects_cg():
    p = params(restart=30)
    p.ksp[s.KSP_PRESSURE].type = s.KSP_CG
    with pytest.raises(s.S355Error):
        s.BatchSolver(sweep_cases()[:2], p)


@pytest.mark.parametrize("scheme,limiter", [(1, 1), (2, 3)])
def test_batch_unsteady_with_limiter_reproduces_single_runs(scheme, limiter):
    # one outer pass per time step (MaxOuterStep 2), so that the batch's
    # shared outer-loop exit cannot differ from a case's own
    cases = sweep_cases()
    p = params(restart=30, rtol=1e-8, maxit=2000)
    p.ksp[s.KSP_PRESSURE].rtol = 1e-6
    p.steady, p.time_scheme, p.dt, p.total_time, p.max_outer_step = 0, scheme, 0.05, 10.0, 2
    p.limiter = limiter
    steps = 3
    alone = run_alone(cases, p, steps)
    with s.BatchSolver(cases, p) as b:
        _, done = b.iterate(steps)
        assert done == steps
        assert b.time_state() == (0.05 + 0.05 + 0.05, 3)
        for k, (fields, _, dk) in enumerate(alone):
            assert dk == steps
            for n in ("U", "V", "W", "P"):
                assert np.array_equal(b.case_field(k, n), fields[n]), (k, n)
