"""The SIMPLE phases between the solves (s355_run_phase) launched from
captured hipGraphs (params.graphs = 1) against the same member calls launched
kernel by kernel: a phase is captured on its second call and replayed after,
with the program's own eager work (solves, field updates, reads) between the
replays, and every replay must give the eager result bit for bit -- k-epsilon
phases, unsteady BDF2 time terms with a limiter and least-squares gradients,
a parameter change in mid-run (the captured phases hold the old arguments and
are dropped), and single phases called directly on changing fields."""
import ctypes as C

import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params

pytestmark = pytest.mark.gpu

TURB = ("U", "V", "W", "P", "RUFACE", "TE", "ED", "VISTUR")


def pair(case, p, batch=False):
    q = s.Params.from_buffer_copy(p)
    q.graphs = 1
    mk = s.BatchSolver if batch else s.Solver
    return mk(case, p), mk(case, q)


def same(a, b, names):
    for n in names:
        x, y = a.get(n), b.get(n)
        assert np.array_equal(x, y), f"{n}: {np.count_nonzero(x != y)} values differ"


def step_case():
    return s.step(2, upstream=2, downstream=6, span=2, turbulent=True, re=5000)


def test_kepsilon_phases():
    p = params(rtol=1e-10, turb=1, maxit=3000)
    p.residual_steady = 0.0
    d0, d1 = pair(step_case(), p)
    for k in range(3):
        h0, _ = d0.iterate(2)
        h1, _ = d1.iterate(2)
        assert np.array_equal(h0, h1)
        same(d0, d1, TURB)


def test_unsteady_bdf2_limiter_least_squares_phases():
    p = params(rtol=1e-10, turb=1, maxit=3000)
    p.steady, p.time_scheme, p.dt, p.total_time, p.max_outer_step = 0, 2, 0.05, 10.0, 3
    p.limiter, p.grad_scheme = 1, 1
    d0, d1 = pair(step_case(), p)
    h0, _ = d0.iterate(3)
    h1, _ = d1.iterate(3)
    assert np.array_equal(h0, h1)
    same(d0, d1, TURB)


def test_parameter_change_drops_the_captured_phases():
    case = s.channel(12, 6, 6, jitter=0.1, split=0.2, seed=1, re=500.0)
    p = params(rtol=1e-10, maxit=3000)
    p.residual_steady = 0.0
    d0, d1 = pair(case, p)
    d0.iterate(3)
    d1.iterate(3)
    for d in (d0, d1):
        q = s.Params.from_buffer_copy(d.params)
        q.urf[0], q.urf[3] = 0.5, 0.2
        d.set_params(q)
    h0, _ = d0.iterate(3)
    h1, _ = d1.iterate(3)
    assert np.array_equal(h0, h1)
    same(d0, d1, ("U", "V", "W", "P", "RUFACE"))


@pytest.mark.parametrize("phase,members,outputs", [
    (s.PHASE_MOMENTUM, [("build_velocity_matrix",), ("save_velocity",)],
     ("AU", "BU_RHS", "BV_RHS", "BW_RHS", "APR", "DUDX", "DPDX", "BU", "BPRE")),
    (s.PHASE_PRESSURE, [("bc_velocity",), ("cal_ru_face",), ("build_pressure_matrix",)], ("RUFACE", "AP", "BP_RHS")),
    (s.PHASE_CORRECT, [("correct_fields", C.POINTER(C.c_double)()), ("bc_velocity",), ("correct_ru_face",)],
     ("P", "U", "V", "W", "RUFACE", "DPDX")),
])
def test_phase_replays_on_changing_fields(phase, members, outputs):
    case = s.channel(10, 6, 8, jitter=0.15, split=0.1, seed=3, re=300.0)
    p = params(rtol=1e-10, maxit=3000)
    d0, d1 = pair(case, p)
    rng = np.random.default_rng(7)
    z = d0.sizes
    for rep in range(4):
        # eager work between the replays: new fields pushed to both contexts
        st = {"U": rng.uniform(-1, 1, z.n_cell), "V": rng.uniform(-1, 1, z.n_cell), "W": rng.uniform(-1, 1, z.n_cell),
              "P": rng.uniform(-1, 1, z.n_cell), "DELTAP": rng.uniform(-0.1, 0.1, z.n_cell),
              "APR": rng.uniform(0.5, 2.0, z.n_cell)}
        for d in (d0, d1):
            for k, v in st.items():
                d.set(k, v)
        for m in members:
            d0.call(*m)
        d1.call("run_phase", phase)
        d1.call("synchronize")
        same(d0, d1, outputs)
