"""Energy equation (UpdateEnergy + SetBCTemperature, navier_scalar.cpp:204-326,
navier_bc.cpp:183-214) on the GPU against the oracle: bitwise assembly of the
temperature system, the conduction-column closed form, and SIMPLE loops of a
heated channel with the energy step on."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import conduction_column, heated, params, push, random_state, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu


def energy_params(**kw):
    p = params(rtol=1e-12, **kw)
    p.solve_energy = 1
    return p


def test_temperature_assembly_is_bitwise():
    case = heated(s.channel(10, 5, 6, jitter=0.15, split=0.2, seed=3, re=400.0))
    p = energy_params()
    d, o = s.Solver(case, p), Oracle(case, p)
    st = random_state(d.sizes.n_cell, d.sizes.n_bnd, d.sizes.n_face, 11)
    st["T"] = 300.0 + 50.0 * np.random.default_rng(2).random(d.sizes.n_cell)
    st["RUFACE"] = np.zeros(d.sizes.n_face)  # no inflow through the inlet
    keys = ("U", "V", "W", "R", "VISLAM", "VISTUR", "T", "RUFACE")
    for x in (d, o):
        push(x, {k: st[k] for k in keys})
    d.call("bc_temperature")
    o.call("bc_temperature")
    assert np.array_equal(d.get("BTEM"), o.get("BTEM"))
    d.call("energy_coefficients")
    d.call("build_scalar_matrix", 1)
    # the oracle's UpdateEnergy assembles the same system before solving it
    o.update_energy()
    assert np.array_equal(d.get("AS"), o.get("AS"))
    assert np.array_equal(d.get("BS_RHS"), o.get("BS_RHS"))


def test_conduction_column_on_the_device():
    case = conduction_column(8)
    d = s.Solver(case, energy_params())
    for _ in range(60):
        info = d.update_energy()
        assert info["reason"] > 0
    z = d.dbls("CELL_X").reshape(-1, 3)[:, 2]
    assert np.max(np.abs(d.get("T") - (300.0 + 100.0 * z))) < 1e-8


@pytest.mark.parametrize("turb", [0, 1])
def test_heated_channel_simple_loop_matches_oracle(turb):
    case = heated(s.channel(12, 6, 8, jitter=0.1, split=0.1, seed=5, re=500.0, turbulent=bool(turb)))
    p = energy_params(turb=turb)
    d, o = s.Solver(case, p), Oracle(case, p)
    hd, _ = d.iterate(8)
    ho, _, _ = o.iterate(8)
    assert np.allclose(hd, ho, rtol=1e-6, atol=1e-14)
    for n in ("U", "V", "W", "P", "T", "BTEM"):
        assert rel_err(d.get(n), o.get(n)) < 1e-7, n
    t = d.get("T")
    # the heated walls warm the fluid (the deferred second-order upwind
    # correction lets single cells undershoot the inlet value slightly)
    assert 300.0 < t.mean() < 350.0
    assert d.last_solve_info()["T"]["reason"] > 0


def test_energy_in_a_batch_matches_single_runs():
    cases = [heated(s.cavity(6, re=100.0), t_wall=320.0), heated(s.channel(8, 4, 5, re=300.0), t_wall=340.0)]
    p = energy_params()
    p.residual_steady = 0.0
    alone = []
    for c in cases:
        d = s.Solver(c, p)
        d.iterate(4)
        alone.append(d.get("T"))
    b = s.BatchSolver(cases, p)
    b.iterate(4)
    for k, t in enumerate(alone):
        assert np.array_equal(b.case_field(k, "T"), t)
