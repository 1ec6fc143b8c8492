"""Kernel parity at the full size of the channel workload (200 x 100 x 200 box,
15 % jitter, 10 % of the hexes split into 6 tets: 5,995,990 cells) against
the FP64 oracle on the same seeded state: the device geometry, the fused
Green-Gauss gradients and the momentum assembly (navier_velocity.cpp:63-393),
the Rhie-Chow face flux (425-476), the pressure-correction assembly
(navier_pressure.cpp:99-232) and the correction block (39-78).  The half-face
kernels add each cell's terms in the reference's order, so every array
agrees bit for bit except where the outlet mass-flow rescale sums the
outlet faces in another order (1e-13)."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[full-parity] {msg}", flush=True)


@pytest.fixture(scope="module")
def pair():
    case = s.channel(200, 100, 200, lx=2.0, ly=1.0, lz=2.0, jitter=0.15, split=0.10, seed=0, re=1000.0)
    log("device context")
    d = s.Solver(case, params())
    log("oracle context (host, ~1 min)")
    o = Oracle(case, params())
    log("contexts ready")
    yield d, o
    d.close()


def close(d, o, names, tol):
    for n in names:
        a, b = d.get(n), o.get(n)
        if tol == 0.0:
            assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} values differ, rel {rel_err(a, b):.3e}"
        else:
            e = rel_err(a, b)
            assert e <= tol, f"{n}: rel err {e:.3e}"


def test_full_size_geometry_matches_the_oracle(pair):
    d, o = pair
    assert d.sizes.n_cell == 5_995_990
    for n in ("CELL_VOL", "CELL_X", "FACE_AREA", "FACE_LAMBDA", "FACE_RLENCOS"):
        assert np.array_equal(d.dbls(n), o.dbls(n)), n


def test_full_size_assembly_bitwise(pair):
    d, o = pair
    z = d.sizes
    st = random_state(z.n_cell, z.n_bnd, z.n_face, 11)
    push(d, {k: st[k] for k in ("U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE")})
    push(o, {k: st[k] for k in ("U", "V", "W", "P", "R", "VISLAM", "VISTUR", "RUFACE")})
    log("momentum assembly")
    for x in (d, o):
        x.call("build_velocity_matrix")
    close(d, o, ["DUDX", "DVDX", "DWDX", "DPDX"], 1e-13)
    close(d, o, ["AU", "BU_RHS", "BV_RHS", "BW_RHS", "APR"], 1e-13)
    # the matrix itself does not see the outlet rescale: bit for bit
    close(d, o, ["AU"], 0.0)
    log("face flux and pressure-correction assembly")
    push(d, {k: o.get(k) for k in ("BU", "BV", "BW", "BRO", "DPDX", "APR")})
    for x in (d, o):
        x.call("cal_ru_face")
        x.call("build_pressure_matrix")
    close(d, o, ["RUFACE", "AP", "BP_RHS"], 0.0)
    log("corrections")
    dp = np.random.default_rng(12).uniform(-0.1, 0.1, z.n_cell)
    push(d, {"DELTAP": dp})
    push(o, {"DELTAP": dp})
    rd, ro = d.correct_fields(), o.correct_fields()
    assert rd == pytest.approx(ro, rel=1e-12)
    close(d, o, ["P", "U", "V", "W", "DPDX"], 0.0)
    for x in (d, o):
        x.call("correct_ru_face")
    close(d, o, ["RUFACE"], 0.0)
