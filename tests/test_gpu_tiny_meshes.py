"""The smallest meshes through the device path, against the FP64 oracle: one
hex cell (no interior face), two and three hexes in a row, the 2^3 cavity,
one and two tetrahedra, and k-epsilon on 12 and 30 cells.  Every row kernel, half-face block, SELL slice and
reduction tile is then mostly padding: geometry equals the host's bit for
bit, the assembly equals the oracle's (bitwise, outlet-rescaled arrays to
rounding), and SIMPLE loops in FP64 and mixed precision follow the oracle.
The one-cell channel stays out of the loops: its pressure correction is the
1 x 1 system 0 * dp = b, whose right-hand side (inlet plus rescaled outlet
flow) is exactly 0 in the oracle and a rounding residue on the device, which
sums the outlet flow in another order -- GMRES then breaks down, as PETSc's
would.  That breakdown must end the solve in both precisions (the mixed
refinement once restarted from the same residual without end).  A ragged
batch puts the two-cell channel beside larger cases and must reproduce each
case's single runs bit for bit."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state, rel_err, tet_cases
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

GEOM = ("CELL_VOL", "CELL_X", "FACE_X", "FACE_N", "FACE_AREA", "FACE_LAMBDA", "FACE_RLENCOS", "BND_DIST")

CASES = {
    "hex1": lambda: s.channel(1, 1, 1, re=100.0),
    "hex2": lambda: s.channel(2, 1, 1, re=100.0),
    "hex3_jitter": lambda: s.channel(3, 1, 1, jitter=0.1, seed=3, re=100.0),
    "cavity2": lambda: s.cavity(2, re=100.0),
    "tet1": lambda: tet_cases()[0],
    "tet2": lambda: tet_cases()[1],
}


def log(msg):
    print(f"[tiny] {msg}", flush=True)


@pytest.mark.parametrize("name", sorted(CASES))
def test_geometry_and_assembly(name):
    log(name)
    case = CASES[name]()
    host = s.HostMesh(case)
    d, o = s.Solver(case, params()), Oracle(case, params())
    try:
        for n in GEOM:
            assert np.array_equal(d.dbls(n), host.dbls(n)), n
        z = d.sizes
        st = random_state(z.n_cell, z.n_bnd, z.n_face, 7)
        names = ("U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE")
        push(d, {k: st[k] for k in names})
        push(o, {k: st[k] for k in names})
        for x in (d, o):
            x.call("build_velocity_matrix")
        for n in ("DUDX", "DPDX", "AU", "BU_RHS", "BV_RHS", "BW_RHS", "APR"):
            e = rel_err(d.get(n), o.get(n))
            assert e <= 1e-13, f"{n}: {e:.3e}"
        assert np.array_equal(d.get("AU"), o.get("AU"))
        push(d, {k: o.get(k) for k in ("BU", "BV", "BW", "BRO", "DPDX", "APR")})
        for x in (d, o):
            x.call("cal_ru_face")
            x.call("build_pressure_matrix")
        for n in ("RUFACE", "AP", "BP_RHS"):
            assert np.array_equal(d.get(n), o.get(n)), n
    finally:
        d.close()
        host.close()


@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
@pytest.mark.parametrize("name", sorted(set(CASES) - {"hex1"}))
def test_simple_loop(name, prec):
    log(f"{name} precision {prec}")
    p = params(rtol=1e-12)
    p.precision = prec
    case = CASES[name]()
    d, o = s.Solver(case, p), Oracle(case, params(rtol=1e-12))
    try:
        hd, _ = d.iterate(5)
        ho, _, _ = o.iterate(5)
        tol = 1e-7 if prec == s.FP64 else 1e-5
        assert np.all(np.isfinite(hd))
        assert np.allclose(hd, ho, rtol=tol, atol=1e-13)
        for n in ("U", "V", "W", "P"):
            a, b = d.get(n), o.get(n)
            assert np.allclose(a, b, rtol=tol, atol=tol * max(1.0, float(np.abs(b).max()))), n
    finally:
        d.close()


TURB = {
    "step_30_cells": lambda: s.step(1, upstream=1, downstream=2, span=1, turbulent=True, re=5000.0),
    "channel_12_cells": lambda: s.channel(3, 2, 2, turbulent=True, re=13750.0),
}


@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
@pytest.mark.parametrize("name", sorted(TURB))
def test_kepsilon_loop(name, prec):
    """k-epsilon with wall functions on a few cells: every cell touches a
    wall, an inlet or an outlet."""
    log(f"k-epsilon {name} precision {prec}")
    p = params(rtol=1e-12, turb=1)
    p.precision = prec
    case = TURB[name]()
    d, o = s.Solver(case, p), Oracle(case, params(rtol=1e-12, turb=1))
    try:
        hd, _ = d.iterate(5)
        ho, _, _ = o.iterate(5)
        tol = 1e-7 if prec == s.FP64 else 1e-5
        assert np.all(np.isfinite(hd))
        assert np.allclose(hd, ho, rtol=tol, atol=1e-13)
        # velocity components against the velocity scale (W is round-off on
        # the channel, whose flow has no z component)
        vref = max(float(np.abs(o.get(n)).max()) for n in ("U", "V", "W"))
        for n in ("U", "V", "W", "P", "TE", "ED", "VISTUR"):
            a, b = d.get(n), o.get(n)
            ref = vref if n in ("U", "V", "W") else float(np.abs(b).max())
            assert np.allclose(a, b, rtol=tol, atol=tol * ref), n
    finally:
        d.close()


@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
def test_ragged_batch_with_a_two_cell_case(prec):
    cases = [s.cavity(4, re=300.0), s.channel(2, 1, 1, re=100.0), s.channel(5, 3, 4, jitter=0.1, split=0.2, seed=2)]
    p = params()
    p.precision = prec
    log(f"batch precision {prec}")
    b = s.BatchSolver(cases, p)
    try:
        b.iterate(4)
        for k, case in enumerate(cases):
            with s.Solver(case, p) as d:
                d.iterate(4)
                for n in ("U", "V", "W", "P"):
                    assert np.array_equal(b.case_field(k, n), d.get(n)), (k, n)
    finally:
        b.close()


@pytest.mark.parametrize("graphs", [0, 1])
@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
def test_inconsistent_one_cell_pressure_system_ends_with_breakdown(prec, graphs):
    log(f"breakdown precision {prec} graphs {graphs}")
    p = params()
    p.precision = prec
    p.graphs = graphs
    with s.Solver(CASES["hex1"](), p) as d:
        d.set("AP", np.zeros(1))
        d.set("BP_RHS", np.full(1, 1e-3))
        with pytest.raises(s.NotConverged) as e:
            d.solve_pressure(1e-8, 1000)
        info = e.value.info
        assert info["reason"] == -5 and info["iters"] <= 30, info
        # a consistent system of the same shape still solves
        d.set("AP", np.full(1, 2.0))
        d.set("BP_RHS", np.full(1, 1e-3))
        info = d.solve_pressure(1e-8, 1000)
        assert info["reason"] > 0 and d.get("DELTAP")[0] == pytest.approx(5e-4, rel=1e-8)


@pytest.mark.parametrize("prec,order,pc", [(s.FP64, 0, s.PC_JACOBI), (s.MIXED, 0, s.PC_JACOBI),
                                           (s.MIXED, 2, s.PC_JACOBI)])
def test_batch_of_many_tiny_cases(prec, order, pc):
    """40 two- and three-cell channels at different Reynolds numbers around a
    3^3 cavity: SELL slices and reduction tiles hold many cases each (the
    slices look their case up per lane), and every case must still equal its
    run alone bit for bit.  (Not with ILU(0): on two or three cells it is the
    exact LU factorisation of a singular pressure-correction matrix and meets
    a zero pivot, in the oracle as in PETSc.)"""
    cases = [s.channel(2 + k % 2, 1, 1, re=50.0 + 10.0 * k) for k in range(20)]
    cases += [s.cavity(3, re=200.0)]
    cases += [s.channel(2 + k % 2, 1, 1, jitter=0.05, seed=k, re=300.0 + 5.0 * k) for k in range(20)]
    p = params(pc=pc)
    p.precision = prec
    p.inner_order = order
    log(f"batch of {len(cases)} tiny cases precision {prec} order {order} pc {pc}")
    b = s.BatchSolver(cases, p)
    try:
        b.iterate(3)
        for k, case in enumerate(cases):
            with s.Solver(case, p) as d:
                d.iterate(3)
                for n in ("U", "V", "W", "P"):
                    assert np.array_equal(b.case_field(k, n), d.get(n)), (k, n)
    finally:
        b.close()
