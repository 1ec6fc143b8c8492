"""Pins of the oracle's k-epsilon pieces, scalar convection and CG against
closed forms and dense solves.

* CG (the pressure correction's KSPCG, MPIStructure.cpp:462-465): the
  assembled pressure-correction matrix is negative semidefinite with the
  constant null space (navier_pressure.cpp:154-175); CG solves it made
  definite against numpy's dense solve, and the singular system against the
  minimum-norm least-squares solution (no preconditioner: the Krylov space of
  a consistent right-hand side stays in the range) or up to the null space
  (Jacobi, ILU(0)); the stopping test is checked with an explicit
  ||M^-1 (b - A x)||; with a tolerance placed between the residuals of two
  steps, CG stops at the step and with the iterate that the Galerkin
  condition over an explicit Krylov basis gives.
* Production P_k = mu_t (grad u + grad u^T) : grad u (navier_scalar.cpp:46-50)
  for an asymmetric velocity gradient, against numpy.
* The wall function (navier_bc.cpp:262-285) in both y+ branches on two wall
  cells: u_tau = sqrt(tau_w / rho) below y+ = 11, C_mu^1/4 sqrt(k) above,
  P_k = tau_w u_tau / (kappa d), S_k = P_k vol, Ap_k's dissipation swapped to
  epsilon_w, epsilon := epsilon_w.
* BuildScalarMatrix's boundary convection (navier_scalar.cpp:474-492) and
  boundary diffusion (451-463) on one cube cell: inflow through the inlet is
  implicit in a_P with the inlet value on the right-hand side, outflow
  through the outlet adds nothing, reverse flow at an outlet brings the
  outlet value in, and reverse flow at an inlet is refused (476-479).
"""

import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state
from oracle.pyoracle import Oracle, OracleError, pc_apply

KAPPA, CMU, SIGMA_K, SIGMA_E = 0.419, 0.09, 1.0, 1.3


def dense(o, name):
    z = o.sizes
    rp, col = o.ints("CSR_ROWPTR"), o.ints("CSR_COL")
    A = np.zeros((z.n_cell, z.n_cell))
    v = o.get(name)
    for i in range(z.n_cell):
        A[i, col[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return A


def set_dense(o, name, A):
    rp, col = o.ints("CSR_ROWPTR"), o.ints("CSR_COL")
    o.set(name, np.concatenate([A[i, col[rp[i]:rp[i + 1]]] for i in range(A.shape[0])]))


# ---------------------------------------------------------------- CG
def pressure_system(pc, seed=3):
    p = params(restart=30, ksp=s.KSP_CG, pc=pc)
    o = Oracle(s.cavity(5, jitter=0.12, seed=seed), p)
    z = o.sizes
    push(o, {k: v for k, v in random_state(z.n_cell, z.n_bnd, z.n_face, seed).items()
             if k in ("U", "V", "W", "P", "R", "APR", "DPDX", "RUFACE")})
    o.call("bc_velocity")
    o.call("cal_ru_face")
    o.call("build_pressure_matrix")
    return o, dense(o, "AP")


@pytest.mark.parametrize("pc", [s.PC_NONE, s.PC_JACOBI, s.PC_ILU0])
def test_cg_matches_dense_solve_on_definite_pressure_matrix(pc):
    o, A = pressure_system(pc)
    n = A.shape[0]
    # the matrix is negative semidefinite with A 1 = 0; shifting the
    # diagonal makes it negative definite (CG's sign checks accept it)
    assert np.abs(A @ np.ones(n)).max() < 1e-12 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 1e-10 * np.abs(A).max()
    A = A - np.diag(0.05 * np.abs(np.diag(A)))
    set_dense(o, "AP", A)
    b = np.random.default_rng(4).normal(size=n)
    o.set("BP_RHS", b)
    info = o.solve_pressure(1e-13, 2000)
    assert info["reason"] > 0, info
    x = o.get("DELTAP")
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) <= 1e-10 * np.linalg.norm(xd)
    # the stopping test on the preconditioned residual, made explicit
    rp, col = o.ints("CSR_ROWPTR"), o.ints("CSR_COL")
    val = o.get("AP")
    r = pc_apply(rp, col, val, pc, b - A @ x)
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(pc_apply(rp, col, val, pc, b)) * (1 + 1e-6)
    # the recursively updated residual drifts from the true one by rounding
    assert info["rnorm"] == pytest.approx(np.linalg.norm(r), rel=1e-2, abs=1e-30)
    # CG on n unknowns: no more than a few n iterations in rounding
    assert 0 < info["iters"] <= 3 * n


@pytest.mark.parametrize("pc", [s.PC_NONE, s.PC_JACOBI, s.PC_ILU0])
def test_cg_on_the_singular_pressure_system(pc):
    o, A = pressure_system(pc, seed=5)
    n = A.shape[0]
    x_true = np.random.default_rng(6).normal(size=n)
    b = A @ x_true  # consistent: in the range of A
    o.set("BP_RHS", b)
    info = o.solve_pressure(1e-12, 4000)
    assert info["reason"] > 0, info
    x = o.get("DELTAP")
    assert np.linalg.norm(b - A @ x) <= 1e-9 * np.linalg.norm(b)
    xmin = np.linalg.lstsq(A, b, rcond=1e-12)[0]  # minimum-norm solution
    if pc == s.PC_NONE:
        # the Krylov space of b stays in the range of A (orthogonal to 1 up to
        # the rounding asymmetry of A), so CG from zero finds x_min itself
        assert np.linalg.norm(x - xmin) <= 1e-7 * np.linalg.norm(xmin)
    # every preconditioner: x_min up to the null space (the constant)
    d = (x - x.mean()) - (xmin - xmin.mean())
    assert np.linalg.norm(d) <= 1e-7 * np.linalg.norm(xmin)


def _galerkin_history(A, b, M_inv, kmax):
    """Preconditioned CG from zero, restated densely: x_k is the Galerkin
    solution over K_k(M^-1 A, M^-1 b) (Q^T (b - A x_k) = 0 on an explicit,
    twice re-orthogonalised basis Q; no recurrences), returned with
    ||M^-1 (b - A x_k)|| for k = 0..kmax."""
    n = A.shape[0]
    Q = []
    v = M_inv @ b
    xs, h = [np.zeros(n)], [np.linalg.norm(M_inv @ b)]
    for _ in range(kmax):
        for _ in range(2):
            for q in Q:
                v = v - (q @ v) * q
        Q.append(v / np.linalg.norm(v))
        Qm = np.array(Q).T
        x = Qm @ np.linalg.solve(Qm.T @ A @ Qm, Qm.T @ b)
        xs.append(x)
        h.append(np.linalg.norm(M_inv @ (b - A @ x)))
        v = M_inv @ (A @ Q[-1])
    return np.array(h), xs


def test_cg_stops_at_the_galerkin_step_with_its_iterate():
    # PETSc's CG (preconditioned norm, zero guess) stops at the first step
    # whose ||M^-1 r_k|| is at most rtol ||M^-1 b||; with the tolerance placed
    # where every earlier residual is 5 % above it and that step's 5 % below,
    # the oracle must stop there with the Galerkin iterate of the Krylov
    # space (a wrong step length, search direction or norm changes the step
    # or the iterate)
    o, A = pressure_system(s.PC_JACOBI, seed=7)
    n = A.shape[0]
    A = A - np.diag(0.05 * np.abs(np.diag(A)))
    A = 0.5 * (A + A.T)  # symmetric to rounding already; exactly, for the restatement
    set_dense(o, "AP", A)
    b = np.random.default_rng(8).normal(size=n)
    o.set("BP_RHS", b)
    M_inv = np.diag(1.0 / np.diag(A))
    h, xs = _galerkin_history(A, b, M_inv, 40)
    ks = next(j for j in range(5, 41) if h[:j].min() > 1.1025 * h[j])
    tol = np.sqrt(h[:ks].min() * h[ks]) / h[0]
    info = o.solve_pressure(tol, 2000)
    assert info["reason"] == 2, info
    assert info["iters"] == ks, (info["iters"], ks)
    assert abs(info["rnorm0"] - h[0]) <= 1e-12 * h[0]
    x = o.get("DELTAP")
    assert np.linalg.norm(x - xs[ks]) <= 1e-8 * np.linalg.norm(xs[ks])
    assert abs(info["rnorm"] - h[ks]) <= 1e-6 * h[ks]


def test_cg_reports_an_indefinite_matrix():
    # a symmetric indefinite matrix breaks CG's (p, Ap) sign consistency
    o, A = pressure_system(s.PC_NONE, seed=7)
    n = A.shape[0]
    A = A - np.diag(0.05 * np.abs(np.diag(A)))
    A[np.arange(0, n, 2), np.arange(0, n, 2)] *= -1.0
    set_dense(o, "AP", A)
    o.set("BP_RHS", np.random.default_rng(8).normal(size=n))
    info = o.solve_pressure(1e-12, 2000)
    assert info["reason"] < 0


# ---------------------------------------------------------------- k-epsilon
def two_wall_cells(mu=(1e-2, 1e-4)):
    """Two cells of 1 x 0.5 x 1 between the walls y = 0 and y = 1 (wall
    distance 0.25), inlet x = 0, outlet x = 1, symmetry z = 0, 1."""
    case = s.channel(1, 2, 1, lx=1, ly=1, lz=1, turbulent=True, re=1000)
    o = Oracle(case, params(turb=1, rtol=1e-13))
    o.set("VISLAM", np.array(mu))
    return o


def test_production_term_closed_form():
    o = two_wall_cells()
    rng = np.random.default_rng(11)
    G = rng.normal(size=(2, 3, 3))  # G[c, i, j] = d u_i / d x_j, asymmetric
    mut = np.array([0.03, 0.002])
    k, eps, rho, mu = np.array([0.02, 0.05]), np.array([0.3, 0.07]), np.array([1.1, 0.9]), np.array([1e-3, 2e-3])
    for c, n in enumerate(("DUDX", "DVDX", "DWDX")):
        o.set(n, G[:, c, :].ravel())
    push(o, {"VISTUR": mut, "TE": k, "ED": eps, "R": rho, "VISLAM": mu})
    o.call("kepsilon_sources")
    pk = mut * np.einsum("cij,cij->c", G + G.transpose(0, 2, 1), G)
    vol = o.dbls("CELL_VOL")
    assert np.allclose(o.get("KE_PROD"), pk, rtol=1e-14, atol=0)
    assert np.allclose(o.get("KE_TESRC"), pk * vol, rtol=1e-14, atol=0)
    assert np.allclose(o.get("KE_APTE"), rho * eps / k * vol, rtol=1e-14, atol=0)
    assert np.allclose(o.get("KE_VISTE"), mu + mut / SIGMA_K, rtol=1e-15, atol=0)
    assert np.allclose(o.get("KE_VISED"), mu + mut / SIGMA_E, rtol=1e-15, atol=0)
    assert np.all(o.get("KE_EDSRC") == 0) and np.all(o.get("KE_APED") == 0)


def test_wall_function_both_branches():
    mu = np.array([1e-2, 1e-4])  # cell 0: y+ < 11 (viscous sublayer), cell 1: y+ > 11
    o = two_wall_cells(tuple(mu))
    k, eps, rho = np.array([0.02, 0.02]), np.array([0.05, 0.04]), np.array([1.0, 1.2])
    u, v, w = np.array([0.3, -0.2]), np.array([0.7, 0.4]), np.array([0.1, 0.25])
    push(o, {"TE": k, "ED": eps, "R": rho, "U": u, "V": v, "W": w, "VISTUR": np.zeros(2)})
    for n in ("DUDX", "DVDX", "DWDX"):
        o.set(n, np.zeros(6))
    o.call("kepsilon_sources")
    ap0 = o.get("KE_APTE").copy()
    o.call("bc_kepsilon")
    d, vol = 0.25, 0.5
    vel = np.hypot(u, w)  # the walls' normal is y: (u, w) is tangential
    tauw = mu * vel / d
    utau_k = CMU ** 0.25 * np.sqrt(k)
    yplus = d * utau_k * rho / (mu + 1e-16)
    assert yplus[0] < 11.0 < yplus[1]
    utau = np.where(yplus < 11.0, np.sqrt(tauw / rho), utau_k)
    prod = tauw * utau / (KAPPA * d)
    eps_w = CMU ** 0.75 * k ** 1.5 / (KAPPA * d)
    assert np.allclose(o.get("KE_PROD"), prod, rtol=1e-14, atol=0)
    assert np.allclose(o.get("KE_TESRC"), prod * vol, rtol=1e-14, atol=0)
    # Ap_k: the cell's dissipation rho eps / k vol exchanged for rho eps_w / k vol
    assert np.allclose(o.get("KE_APTE"), ap0 - rho * eps / k * vol + rho * eps_w / k * vol, rtol=1e-13, atol=0)
    assert np.allclose(o.get("ED"), eps_w, rtol=1e-14, atol=0)
    brid, c1 = o.ints("BND_RID"), o.ints("FACE_C1")[: o.sizes.n_bnd]
    wall = brid == 1
    assert np.allclose(o.get("BED")[wall], eps_w[c1[wall]], rtol=1e-14, atol=0)
    assert np.array_equal(o.get("BTE")[wall], k[c1[wall]])
    # outlet and symmetry faces take the cell values (eps already replaced)
    zg = (brid == 3) | (brid == 4)
    assert np.array_equal(o.get("BTE")[zg], k[c1[zg]])
    assert np.allclose(o.get("BED")[zg], eps_w[c1[zg]], rtol=1e-14, atol=0)


def test_eps_wall_rows_are_identity_rows():
    o = two_wall_cells()
    o.iterate(1)
    o.call("kepsilon_sources")
    o.call("bc_kepsilon")
    o.call("build_scalar_matrix", 3)
    o.call("eps_wall_rows")
    a, dg = o.get("AS"), o.ints("CSR_DIAG")
    assert np.all(a[dg] == 1.0) and np.count_nonzero(a) == 2
    assert np.array_equal(o.get("BS_RHS"), o.get("ED"))


# ---------------------------------------------------------------- scalar boundary terms
def one_cell(turbulent=True):
    case = s.channel(1, 1, 1, lx=1, ly=1, lz=1, turbulent=turbulent, re=1000)
    return Oracle(case, params(turb=1, rtol=1e-13))


def scalar_row(o, phi, bphi, gamma, flux_in, flux_out):
    """Assemble the k row of the one-cell case: diffusion gamma, uniform
    cell value phi, boundary values bphi, face fluxes through the inlet
    (bnd 0) and the outlet (bnd 1) as given (outward positive)."""
    z = o.sizes
    ru = np.zeros(z.n_face)
    brid = o.ints("BND_RID")
    ru[brid == 2] = flux_in
    ru[brid == 3] = flux_out
    push(o, {"TE": [phi], "BTE": bphi, "KE_VISTE": [gamma], "KE_TESRC": [0.0], "KE_APTE": [0.0], "RUFACE": ru})
    o.call("build_scalar_matrix", 2)
    return o.get("AS")[0], o.get("BS_RHS")[0]


def test_scalar_inlet_inflow_and_outlet_outflow():
    o = one_cell()
    urf = o.params.urf[5]
    phi0, phi_in, F = 0.4, 1.3, 0.7
    brid = o.ints("BND_RID")
    bphi = np.full(o.sizes.n_bnd, phi0)
    bphi[brid == 2] = phi_in
    a, b = scalar_row(o, phi0, bphi, 0.0, -F, F)
    # inflow F through the inlet: a_P = F, right-hand side F phi_in; the
    # outflow through the outlet adds nothing (first-order upwind)
    assert a == pytest.approx(F / urf, rel=1e-15)
    assert b == pytest.approx(F * phi_in + (1 - urf) * (F / urf) * phi0, rel=1e-14)


def test_scalar_reverse_flow_at_the_outlet():
    o = one_cell()
    urf = o.params.urf[5]
    phi0, phi_out, F = 0.4, 0.9, 0.3
    brid = o.ints("BND_RID")
    bphi = np.full(o.sizes.n_bnd, phi0)
    bphi[brid == 3] = phi_out
    a, b = scalar_row(o, phi0, bphi, 0.0, 0.0, -F)
    assert a == pytest.approx(F / urf, rel=1e-15)
    assert b == pytest.approx(F * phi_out + (1 - urf) * (F / urf) * phi0, rel=1e-14)


def test_scalar_reverse_flow_at_the_inlet_is_refused():
    o = one_cell()
    with pytest.raises(OracleError):
        scalar_row(o, 0.4, np.full(o.sizes.n_bnd, 0.4), 0.0, 0.2, -0.2)


def test_scalar_boundary_diffusion():
    # uniform phi everywhere: no gradient, so each of the 6 faces adds
    # gamma * rlencos = gamma * area / (h / 2) = 2 gamma to a_P and
    # 2 gamma phi_b to the right-hand side
    o = one_cell()
    urf = o.params.urf[5]
    phi0, gamma = 0.6, 0.05
    a, b = scalar_row(o, phi0, np.full(o.sizes.n_bnd, phi0), gamma, 0.0, 0.0)
    assert a == pytest.approx(12 * gamma / urf, rel=1e-14)
    assert b == pytest.approx(12 * gamma * phi0 + (1 - urf) * (12 * gamma / urf) * phi0, rel=1e-14)
