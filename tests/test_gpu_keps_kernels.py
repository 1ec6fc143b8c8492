"""Kernel-level parity of the k-epsilon pieces of UpdateTurKEpsilon
(navier_scalar.cpp:11-197) against the FP64 oracle, piece by piece on the
same seeded state: sources and viscosities (33-64), the wall function and
zero-gradient boundary values (SetBCKEpsilon, navier_bc.cpp:220-345), the k
and epsilon systems with inlet convection (BuildScalarMatrix, 407-592), the
epsilon wall rows (127-154), the clips and epsilon sources (105-121), and
the clip and mu_t (165-178).  Every piece evaluates the reference's
expressions in its order, so each is compared bit for bit, except for the
wall epsilon's k^1.5 (the device rounds it correctly, the host libm's pow is
1 ulp off for ~0.1 % of arguments), which is compared to the ulp; the solves
in between agree to their tolerance, so the oracle's solution is copied to
the device before the next piece."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import hybrid_channel, params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2, turbulent=True),
    "channel_split": lambda: s.channel(8, 6, 5, jitter=0.15, split=0.2, seed=4, turbulent=True, re=8000.0),
    "hybrid_prisms_pyramids": lambda: hybrid_channel(6, 4, 5, prism_frac=0.3, pyramid_frac=0.3, jitter=0.15, seed=5,
                                                     re=8000.0, turbulent=True),
}
KE = ("KE_PROD", "KE_TESRC", "KE_EDSRC", "KE_APTE", "KE_APED", "KE_VISTE", "KE_VISED")


def same(d, o, names):
    for n in names:
        a, b = d.get(n), o.get(n)
        assert np.array_equal(a, b), f"{n}: rel err {rel_err(a, b):.3e}"


def within_ulps(d, o, names, ulps):
    # then continue from the oracle's values, so later pieces compare bitwise
    for n in names:
        a, b = d.get(n), o.get(n)
        tol = ulps * np.spacing(np.abs(b))
        assert np.all(np.abs(a - b) <= tol), f"{n}: {np.count_nonzero(a != b)} values off, rel err {rel_err(a, b):.3e}"
        d.set(n, b)


def keps_state(d, seed):
    z = d.sizes
    rng = np.random.default_rng(seed)
    st = random_state(z.n_cell, z.n_bnd, z.n_face, seed)
    # laminar viscosity over three decades, so wall cells fall on both sides
    # of y+ = 11
    st["VISLAM"] = 10.0 ** rng.uniform(-4.5, -1.0, z.n_cell)
    st["TE"] = 10.0 ** rng.uniform(-3, -1, z.n_cell)
    st["ED"] = 10.0 ** rng.uniform(-3, 0, z.n_cell)
    # no outflow through an inlet (the reference stops there, 476-479)
    ru = st["RUFACE"]
    brid = d.ints("BND_RID")
    inlet = {r.rid for r in d.case.regions if r.type1 == 2}
    isin = np.isin(brid, list(inlet))
    ru[: z.n_bnd][isin] = -np.abs(ru[: z.n_bnd][isin])
    return {k: st[k] for k in ("U", "V", "W", "R", "VISLAM", "VISTUR", "TE", "ED", "DUDX", "DVDX", "DWDX", "BTE",
                               "BED", "RUFACE")}


def yplus(d, st):
    brid = d.ints("BND_RID")
    walls = {r.rid for r in d.case.regions if r.type1 == 1}
    w = np.isin(brid, list(walls))
    c1 = d.ints("FACE_C1")[: d.sizes.n_bnd][w]
    dist = d.dbls("BND_DIST")[w]
    return dist * 0.09 ** 0.25 * np.sqrt(st["TE"][c1]) * st["R"][c1] / (st["VISLAM"][c1] + 1e-16)


@pytest.mark.parametrize("name", sorted(CASES))
def test_kepsilon_pieces_bitwise(name):
    case = CASES[name]()
    p = params(turb=1, rtol=1e-12, maxit=3000)
    d, o = s.Solver(case, p), Oracle(case, p)
    st = keps_state(d, 21)
    yp = yplus(d, st)
    assert np.any(yp < 11.0) and np.any(yp > 11.0), "both wall-function branches"
    push(d, st)
    push(o, st)
    # sources, viscosities, production
    for x in (d, o):
        x.call("kepsilon_sources")
    same(d, o, KE)
    # wall function (in place on epsilon) and boundary values.  The wall
    # epsilon holds k^1.5: the device rounds it correctly, the oracle calls
    # the host's pow as the reference does (glibc: 1 ulp off for ~0.1 % of
    # arguments), so epsilon and what it enters may differ by an ulp or two
    for x in (d, o):
        x.call("bc_kepsilon")
    same(d, o, ("KE_PROD", "KE_TESRC", "KE_EDSRC", "KE_APED", "KE_VISTE", "KE_VISED", "BTE"))
    within_ulps(d, o, ("ED", "BED"), 1)
    within_ulps(d, o, ("KE_APTE",), 4)
    # the k system, inlet convection included
    for x in (d, o):
        x.call("build_scalar_matrix", 2)
    same(d, o, ("DPHIDX", "AS", "BS_RHS"))
    a = d.solve_scalar(2, 1e-12, 3000)
    b = o.solve_scalar(2, 1e-12, 3000)
    assert a["reason"] > 0 and b["reason"] > 0
    assert rel_err(d.get("TE"), o.get("TE")) < 1e-9
    # the same k on both, with a few cells below the clip
    te = o.get("TE")
    te[:: 7] = -1e-3
    te[3:: 11] = 5e-10
    d.set("TE", te)
    o.set("TE", te)
    for x in (d, o):
        x.call("kepsilon_post_k")
    same(d, o, ("TE", "KE_EDSRC", "KE_APED"))
    # the epsilon system and its wall rows
    for x in (d, o):
        x.call("build_scalar_matrix", 3)
    same(d, o, ("DPHIDX", "AS", "BS_RHS"))
    for x in (d, o):
        x.call("eps_wall_rows")
    same(d, o, ("AS", "BS_RHS"))
    a = d.solve_scalar(3, 1e-12, 3000)
    b = o.solve_scalar(3, 1e-12, 3000)
    assert a["reason"] > 0 and b["reason"] > 0
    assert rel_err(d.get("ED"), o.get("ED")) < 1e-9
    ed = o.get("ED")
    ed[:: 5] = -1.0
    ed[2:: 9] = 0.5e-12
    d.set("ED", ed)
    o.set("ED", ed)
    for x in (d, o):
        x.call("kepsilon_post_eps")
    same(d, o, ("ED", "VISTUR"))
    assert np.all(d.get("ED") >= 1e-12) and np.all(d.get("TE") >= 1e-9)


def test_reverse_flow_at_an_inlet_is_an_error_on_both():
    case = CASES["step_tets"]()
    p = params(turb=1)
    d, o = s.Solver(case, p), Oracle(case, p)
    st = keps_state(d, 3)
    brid = d.ints("BND_RID")
    inlet = [r.rid for r in case.regions if r.type1 == 2]
    st["RUFACE"][: d.sizes.n_bnd][np.isin(brid, inlet)] = 0.01  # outflow through the inlet
    push(d, st)
    push(o, st)
    for x in (d, o):
        x.call("kepsilon_sources")
        x.call("bc_kepsilon")
    with pytest.raises(Exception):
        o.call("build_scalar_matrix", 2)
    with pytest.raises(s.S355Error):
        d.call("build_scalar_matrix", 2)
        d.call("synchronize")
