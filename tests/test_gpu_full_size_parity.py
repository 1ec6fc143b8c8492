"""Kernel parity at the full size of the channel workload (200 x 100 x 200 box,
15 % jitter, 10 % of the hexes split into 6 tets: 5,995,990 cells) against
the FP64 oracle on the same seeded state: the device geometry, the fused
Green-Gauss gradients and the momentum assembly (navier_velocity.cpp:63-393),
the Rhie-Chow face flux (425-476), the pressure-correction assembly
(navier_pressure.cpp:99-232) and the correction block (39-78).  The half-face
kernels add each cell's terms in the reference's order, so every array
agrees bit for bit except where the outlet mass-flow rescale sums the
outlet faces in another order (1e-13).  The k-epsilon pieces are checked
the same way at one GPU's share of the 20 M-cell turbulent channel."""
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


def test_kepsilon_pieces_at_one_gpus_share_of_the_20m_channel():
    """The k-epsilon pieces (navier_scalar.cpp:11-197, navier_bc.cpp:220-345)
    on the turbulent channel at one GPU's share of the 20 M-cell workload
    (200 x 100 x 125 hexes, 10 % jitter, 2.5 M cells): sources, wall function,
    the k and epsilon systems and the epsilon wall rows bit for bit, the wall
    epsilon's k^1.5 to the ulp (the device rounds it correctly, the host's
    pow is 1 ulp off for a few arguments)."""
    case = s.channel(200, 100, 125, lx=2.0, ly=1.0, lz=2.0, jitter=0.10, split=0.0, seed=0, re=13750.0,
                     turbulent=True)
    p = params(turb=1)
    log("turbulent channel 200x100x125: contexts")
    d, o = s.Solver(case, p), Oracle(case, p)
    z = d.sizes
    assert z.n_cell == 2_500_000
    rng = np.random.default_rng(5)
    st = random_state(z.n_cell, z.n_bnd, z.n_face, 5)
    st["VISLAM"] = 10.0 ** rng.uniform(-4.5, -1.0, z.n_cell)
    st["TE"] = 10.0 ** rng.uniform(-3, -1, z.n_cell)
    st["ED"] = 10.0 ** rng.uniform(-3, 0, z.n_cell)
    ru = st["RUFACE"]
    inlet = np.isin(d.ints("BND_RID"), [r.rid for r in case.regions if r.type1 == 2])
    ru[: z.n_bnd][inlet] = -np.abs(ru[: z.n_bnd][inlet])
    names = ("U", "V", "W", "R", "VISLAM", "VISTUR", "TE", "ED", "DUDX", "DVDX", "DWDX", "BTE", "BED", "RUFACE")
    push(d, {k: st[k] for k in names})
    push(o, {k: st[k] for k in names})
    log("sources and wall function")
    for x in (d, o):
        x.call("kepsilon_sources")
        x.call("bc_kepsilon")
    close(d, o, ["KE_PROD", "KE_TESRC", "KE_EDSRC", "KE_APED", "KE_VISTE", "KE_VISED", "BTE"], 0.0)
    for n in ("ED", "BED", "KE_APTE"):
        a, b = d.get(n), o.get(n)
        assert np.all(np.abs(a - b) <= 4 * np.spacing(np.abs(b))), n
        d.set(n, b)
    log("k and epsilon systems")
    for isca in (2, 3):
        for x in (d, o):
            x.call("build_scalar_matrix", isca)
        if isca == 3:
            for x in (d, o):
                x.call("eps_wall_rows")
        close(d, o, ["DPHIDX", "AS", "BS_RHS"], 0.0)
    d.close()
