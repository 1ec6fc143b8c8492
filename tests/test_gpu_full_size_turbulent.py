"""Size-independent properties at the full size of the 20 M-cell k-epsilon
channel (400 x 200 x 250 hexes, 10 % jitter, Re_b = 13750, GMRES(50) +
ILU(0) on every solve, FP32 inner work in FP64 refinement, captured ILU
sweeps), too large for the oracle: the mesh and CSR invariants, the
pressure-correction solve meeting PETSc's stopping test with left ILU(0)
preconditioning, checked on the host with the oracle's FP64 ILU(0)
(||M^-1 (b - A x)|| <= rtol ||M^-1 b||), and a bitwise rerun of that solve."""
import numpy as np
import pytest
import scipy.sparse as sp

import sysu_cfd_parallel_amd as s
from oracle.pyoracle import pc_apply

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

CELLS = 20_000_000


def turbulent_params():
    p = s.set_all_ksp(s.default_params(), s.KSP_GMRES, 50, s.PC_ILU0, maxit=20000)
    p.turb_model = 1
    p.residual_steady = 0.0
    p.precision = s.MIXED
    p.ir_max = 50
    p.graphs = 1
    return p


@pytest.fixture(scope="module")
def full():
    case = s.channel(400, 200, 250, lx=2.0, ly=1.0, lz=2.0, jitter=0.10, split=0.0, seed=0, re=13750.0,
                     turbulent=True)
    d = s.Solver(case, turbulent_params())
    print(f"[full] turbulent channel context: {d.sizes.n_cell} cells, nnz {d.sizes.nnz}", flush=True)
    h, done = d.iterate(1)  # one SIMPLE step from rest, k-epsilon included
    print(f"[full] first SIMPLE step: ResMax {h[-1, 10]:.3e}", flush=True)
    assert done == 1
    yield d
    d.close()


def test_mesh_and_csr_invariants(full):
    d = full
    z = d.sizes
    assert z.n_cell == CELLS
    assert int(d.ints("CELL_NFACE").sum()) == 2 * z.n_interior + z.n_bnd
    rp, col, dg = d.ints("CSR_ROWPTR"), d.ints("CSR_COL"), d.ints("CSR_DIAG")
    assert rp[-1] == z.nnz == z.n_cell + 2 * z.n_interior
    assert z.nnz == 139_540_000  # N + 2 (399*200*250 + 400*199*250 + 400*200*249)
    rows = np.repeat(np.arange(z.n_cell, dtype=np.int32), np.diff(rp))
    inner = np.diff(col) > 0
    assert np.all(inner | (np.diff(rows) > 0))
    assert np.array_equal(col[dg], np.arange(z.n_cell, dtype=np.int32))
    for n in ("TE", "ED", "VISTUR"):
        v = d.get(n)
        assert np.all(np.isfinite(v)) and v.min() >= 0.0, n


def test_mixed_ilu0_pressure_solve_stopping_test_and_rerun(full):
    d = full
    d.call("build_velocity_matrix")
    info = d.solve_velocity(1e-8, 20000)
    assert all(i["reason"] > 0 for i in info)
    d.call("cal_ru_face")
    d.call("build_pressure_matrix")
    z = d.sizes
    rp, col = d.ints("CSR_ROWPTR"), d.ints("CSR_COL")
    val, b = d.get("AP"), d.get("BP_RHS")
    sols = []
    for _ in range(2):
        d.set("AP", val)
        d.set("BP_RHS", b)
        info = d.solve_pressure(1e-6, 20000)
        assert info["reason"] > 0 and info["refinements"] >= 1, info
        sols.append(d.get("DELTAP"))
        print(f"[full] pressure solve: {info}", flush=True)
    assert np.array_equal(sols[0], sols[1])
    A = sp.csr_matrix((val, col, rp), shape=(z.n_cell, z.n_cell))
    r = b - A @ sols[0]
    num = np.linalg.norm(pc_apply(rp, col, val, s.PC_ILU0, r))
    den = np.linalg.norm(pc_apply(rp, col, val, s.PC_ILU0, b))
    print(f"[full] ||M^-1 r|| / ||M^-1 b|| with FP64 ILU(0): {num / den:.3e}", flush=True)
    # the outer test used the FP32 factors: their norm differs from the FP64
    # factors' by far less than 1 %
    assert num <= 1e-6 * den * 1.01
