"""The batched sweep at its full size: the 64 lid-driven cavities of the
benchmark's seeded sweep (edges drawn from 16..64, Re log-uniform in
[100, 1000], seed 2024; 5,328,605 cells) as one batch in mixed precision
with captured GMRES cycles, against each case run alone, bit for bit over 2
SIMPLE steps, and the three smallest cases against the FP64 oracle at the
same solve tolerances."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from bench import Sweep
from oracle.pyoracle import Oracle

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

STEPS = 2
NAMES = ("U", "V", "W", "P")


def shapes(d):
    # one set of kernel shapes for the batch and every single case, so that
    # every reduction runs over the same tiles (the autotuner would pick by
    # system size)
    for prec in (0, 1):
        d.set_kernel_shapes(prec, 0, nt=0, rows=4, sigma=512)


@pytest.fixture(scope="module")
def sweep():
    sw = Sweep("full")
    cases = sw.case()
    p = sw.params(s.MIXED)
    p.graphs = 1
    b = s.BatchSolver(cases, p)
    shapes(b)
    print(f"[sweep] batch of {len(cases)} cavities, {b.sizes.n_cell} cells", flush=True)
    hb, done = b.iterate(STEPS)
    assert done == STEPS
    out = {"sw": sw, "cases": cases, "p": p, "hist": hb,
           "fields": [{n: b.case_field(k, n) for n in NAMES} for k in range(len(cases))],
           "res": b.case_residuals(), "info": b.case_solve_info()}
    b.close()
    return out


def test_sweep_layout_and_solves(sweep):
    sw, cases = sweep["sw"], sweep["cases"]
    assert len(cases) == 64 and sum(e ** 3 for e in sw.edges) == 5_328_605
    assert min(sw.edges) >= 16 and max(sw.edges) <= 64
    assert all(100.0 <= r <= 1000.0 for r in sw.res)
    info = sweep["info"]
    assert all(i["reason"] > 0 for i in info["P"]) and all(i["reason"] > 0 for i in info["U"])


def test_sweep_batch_equals_each_case_alone(sweep):
    cases, p = sweep["cases"], sweep["p"]
    hist_max = None
    for k, c in enumerate(cases):
        d = s.Solver(c, p)
        shapes(d)
        h, done = d.iterate(STEPS)
        assert done == STEPS
        for n in NAMES:
            assert np.array_equal(d.get(n), sweep["fields"][k][n]), (k, n)
        hist_max = h if hist_max is None else np.maximum(hist_max, h)
        d.close()
    # the batch history is the element-wise maximum over its cases
    assert np.array_equal(sweep["hist"][:, :4], hist_max[:, :4])


def test_smallest_sweep_cases_against_the_oracle(sweep):
    sw, cases = sweep["sw"], sweep["cases"]
    for k in np.argsort(sw.edges, kind="stable")[:3]:
        o = Oracle(cases[k], sw.params(s.FP64))
        ho, done, rc = o.iterate(STEPS)
        assert rc == 0 and done == STEPS
        for n in NAMES:
            a, b = sweep["fields"][k][n], o.get(n)
            err = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
            # both stop their solves at the sweep's tolerances (1e-8 momentum,
            # 1e-6 pressure), which leave each run ~1e-6 from its converged step
            assert err < 5e-5, (k, n, err)
