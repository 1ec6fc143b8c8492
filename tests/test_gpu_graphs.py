"""Captured hipGraphs of GMRES restart cycles and the kernel-shape
autotuner.  A graph replays exactly the kernels of the launch-by-launch path,
so results agree bit for bit on every replay, with the program's own eager
work (assembly, corrections, other solves) between the replays.  Tuned shapes
change only the order of the reductions: results stay within rounding, and
two contexts of one process share the tuned shapes and agree bit for bit."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "P", "RUFACE")


def run(case, p, steps, batch=False):
    d = s.BatchSolver(case, p) if batch else s.Solver(case, p)
    h, done = d.iterate(steps)
    out = {n: d.get(n) for n in FIELDS}
    info = d.last_solve_info()
    shapes = d.kernel_shapes()
    d.close()
    return out, h, info, shapes


def base(prec=s.FP64, restart=20):
    p = params(restart=restart, rtol=1e-8, maxit=3000)
    p.ksp[s.KSP_PRESSURE].rtol = 1e-6
    p.residual_steady = 0.0
    p.precision = prec
    return p


@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
@pytest.mark.parametrize("name,make", [
    ("cavity", lambda: s.cavity(8, re=400.0, jitter=0.1, seed=2)),
    ("channel", lambda: s.channel(12, 6, 6, jitter=0.1, split=0.2, seed=1, re=500.0)),
])
def test_graph_replays_match_eager_launches(prec, name, make):
    case = make()
    p0 = base(prec)
    p1 = base(prec)
    p1.graphs = 1
    a, ha, ia, _ = run(case, p0, 8)
    b, hb, ib, _ = run(case, p1, 8)
    for n in FIELDS:
        assert np.array_equal(a[n], b[n]), n
    assert np.array_equal(ha, hb)
    for k in ("U", "P"):
        assert ia[k]["iters"] == ib[k]["iters"] and ia[k]["reason"] == ib[k]["reason"]


@pytest.mark.parametrize("order", [0, 2])
def test_graph_replays_in_a_batch(order):
    # order 2 with graphs: the sweep workload's path (captured GMRES cycles
    # and SIMPLE phases, each case's FP32 cycle in the input cell order)
    cases = [s.cavity(6, re=100.0), s.cavity(9, re=700.0, jitter=0.1, seed=5), s.cavity(5, re=300.0, tets=True)]
    p0 = base(s.MIXED)
    p1 = base(s.MIXED)
    p0.inner_order = p1.inner_order = order
    p1.graphs = 1
    a, ha, _, _ = run(cases, p0, 5, batch=True)
    b, hb, _, _ = run(cases, p1, 5, batch=True)
    for n in FIELDS:
        assert np.array_equal(a[n], b[n]), n
    assert np.array_equal(ha, hb)


@pytest.mark.parametrize("graphs", [0, 1])
def test_iteration_cap_inside_a_cycle(graphs):
    case = s.cavity(10, re=400.0, jitter=0.1, seed=4)
    p = base(s.FP64, restart=5)
    p.graphs = graphs
    d = s.Solver(case, p)
    d.iterate(1)  # a moving lid: a nonzero mass imbalance to solve for
    d.call("build_velocity_matrix")
    d.call("cal_ru_face")
    d.call("build_pressure_matrix")
    assert np.abs(d.get("BP_RHS")).max() > 0
    with pytest.raises(s.NotConverged) as err:
        d.solve_pressure(1e-14, 7)
    info = err.value.info
    assert info["reason"] == -3 and info["iters"] == 7
    assert "did not converge" in str(err.value)


def test_autotuned_shapes_are_shared_within_a_process():
    case = s.channel(16, 8, 8, jitter=0.1, split=0.1, seed=6, re=800.0)
    p = base(s.MIXED)
    p.ksp[s.KSP_PRESSURE].rtol = 1e-10
    p.autotune = 1
    a, ha, _, sa = run(case, p, 4)
    b, hb, _, sb = run(case, p, 4)
    assert sa["fp32"]["tuned"] and sa == sb
    # any shape the tuner offers may win on a case this small (0 = SELL-64)
    assert sa["fp32"]["lanes"] in (0, 1, 2, 4, 8, 16) and sa["fp32"]["rows"] in (1, 2, 4, 8, 16)
    assert sa["fp32"]["nt"] in (0, 1)
    for n in FIELDS:
        assert np.array_equal(a[n], b[n]), n
    q = base(s.MIXED)
    q.ksp[s.KSP_PRESSURE].rtol = 1e-10
    c, _, _, sc = run(case, q, 4)
    assert not sc["fp32"]["tuned"]
    for n in ("U", "V", "W", "P"):
        assert rel_err(a[n], c[n]) < 1e-7, n


@pytest.mark.parametrize("prec", [s.FP64, s.MIXED])
def test_sell_spmv_equals_single_lane_csr(prec):
    # all sum each row's entries in CSR order, one row per lane, whatever
    # window the SELL copy sorts its rows in
    case = s.channel(12, 6, 6, jitter=0.15, split=0.2, seed=8, re=600.0)
    k = 1 if prec == s.MIXED else 0
    out = []
    for lanes, sigma in ((1, 0), (0, 64), (0, 512), (0, 4096)):
        d = s.Solver(case, base(prec))
        d.set_kernel_shapes(k, lanes, sigma=sigma)
        h, _ = d.iterate(5)
        shape = d.kernel_shapes()["fp32" if k else "fp64"]
        assert shape["lanes"] == lanes and (lanes or shape["sigma"] == sigma)
        out.append(({n: d.get(n) for n in FIELDS}, h))
        d.close()
    for o in out[1:]:
        for n in FIELDS:
            assert np.array_equal(out[0][0][n], o[0][n]), n
        assert np.array_equal(out[0][1], o[1])


def test_two_slice_sell_equals_single_lane_csr():
    # above 2^20 rows the FP32 SELL product takes two slices per wave (one row
    # of each per lane); every row still sums its entries in CSR order
    case = s.channel(112, 56, 112, jitter=0.15, split=0.12, seed=8, re=600.0)
    p = base(s.MIXED, restart=50)
    for k in range(3):
        p.ksp[k].maxit = 20000
    out = []
    for lanes, sigma in ((1, 0), (0, 512)):
        d = s.Solver(case, p)
        assert d.sizes.n_cell >= 1 << 20
        d.set_kernel_shapes(1, lanes, sigma=sigma)
        h, _ = d.iterate(2)
        out.append(({n: d.get(n) for n in FIELDS}, h))
        d.close()
    for n in FIELDS:
        assert np.array_equal(out[0][0][n], out[1][0][n]), n
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("sigma", [64, 512, 4096])
def test_forced_slices_per_wave_equal_single_lane_csr(sigma):
    # 1040 rows = 17 slices of 64 (an odd count: the two-slice kernel's last
    # wave has one slice), rows of up to 12 entries (polyhedral cells next to
    # split hexes: slices wider than one batch of 8 slots)
    case = s.channel(11, 7, 5, jitter=0.15, split=0.35, seed=3, re=600.0)
    out = []
    for lanes in (1, -1, -2):
        d = s.Solver(case, base(s.MIXED))
        assert d.sizes.n_cell == 1040
        rp = d.ints("CSR_ROWPTR")
        assert np.diff(rp).max() > 8
        d.set_kernel_shapes(1, lanes, sigma=0 if lanes == 1 else sigma)
        shape = d.kernel_shapes()["fp32"]
        assert shape["slices"] == {1: 0, -1: 1, -2: 2}[lanes]
        h, _ = d.iterate(5)
        out.append(({n: d.get(n) for n in FIELDS}, h))
        d.close()
    for o in out[1:]:
        for n in FIELDS:
            assert np.array_equal(out[0][0][n], o[0][n]), n
        assert np.array_equal(out[0][1], o[1])


def ordered(prec=s.MIXED, restart=20, graphs=0, kind=1):
    p = base(prec, restart)
    p.inner_order = kind
    p.graphs = graphs
    return p


@pytest.mark.parametrize("graphs,kind", [(0, 1), (1, 1), (0, 2)])
def test_inner_order_every_spmv_shape_bitwise(graphs, kind):
    # the FP32 cycle on the rows in reverse Cuthill-McKee order: CSR and the
    # SELL copies with 16-bit column offsets (one and two slices per wave,
    # every window) sum each row in the same entry order at the same
    # positions, so every shape gives the same bits
    case = s.channel(11, 7, 5, jitter=0.15, split=0.35, seed=3, re=600.0)
    out = []
    for lanes, sigma in ((1, 0), (0, 64), (0, 512), (0, 4096), (-1, 512), (-2, 512), (-2, 64)):
        d = s.Solver(case, ordered(graphs=graphs, kind=kind))
        d.set_kernel_shapes(1, lanes, sigma=sigma)
        h, _ = d.iterate(5)
        assert d.kernel_shapes()["fp32"]["order"] == 1
        out.append(({n: d.get(n) for n in FIELDS}, h))
        d.close()
    for o in out[1:]:
        for n in FIELDS:
            assert np.array_equal(out[0][0][n], o[0][n]), n
        assert np.array_equal(out[0][1], o[1])


def test_inner_order_agrees_with_natural_order_and_oracle():
    from oracle.pyoracle import Oracle

    case = s.channel(12, 8, 10, jitter=0.15, split=0.15, seed=4, re=400.0)
    res = {}
    for order in (0, 1, 2):
        p = params(restart=30, rtol=1e-12, maxit=5000)
        p.precision, p.inner_order, p.residual_steady = s.MIXED, order, 0.0
        d = s.Solver(case, p)
        d.set_kernel_shapes(1, 0, sigma=512)
        h, _ = d.iterate(6)
        assert d.kernel_shapes()["fp32"]["order"] == (order > 0)
        res[order] = ({n: d.get(n) for n in FIELDS}, h)
        d.close()
    p = params(restart=30, rtol=1e-12, maxit=5000)
    p.residual_steady = 0.0
    o = Oracle(case, p)
    ho, _, _ = o.iterate(6)
    for k in (1, 2):
        for n in ("U", "V", "W", "P"):
            assert rel_err(res[k][0][n], res[0][0][n]) < 1e-8, (k, n)
            assert rel_err(res[k][0][n], o.get(n)) < 1e-8, (k, n)
        assert np.allclose(res[k][1], ho, rtol=1e-6, atol=1e-14)


def test_inner_order_kind_fixed_once_built():
    case = s.channel(8, 5, 6, jitter=0.1, split=0.1, seed=2)
    d = s.Solver(case, ordered(kind=2))
    d.iterate(1)
    p = ordered(kind=1)
    with pytest.raises(s.S355Error):
        d.set_params(p)
    d.set_params(ordered(kind=0))  # off is always allowed
    d.close()


def test_inner_order_is_for_mixed_gmres_only():
    case = s.channel(8, 5, 6, jitter=0.1, split=0.1, seed=2)
    for prec, pc in ((s.FP64, s.PC_JACOBI), (s.MIXED, s.PC_ILU0)):
        p = params(restart=20, pc=pc)
        p.precision, p.inner_order = prec, 1
        d = s.Solver(case, p)
        d.iterate(2)
        assert d.kernel_shapes()["fp32"]["order"] == 0
        d.close()


def test_two_slices_per_wave_only_for_one_fp32_case():
    d = s.Solver(s.cavity(4), base(s.MIXED))
    with pytest.raises(s.S355Error):
        d.set_kernel_shapes(0, -2)


@pytest.mark.parametrize("lanes,sigma", [(0, 64), (0, 4096), (2, 0), (16, 0)])
def test_every_spmv_shape_matches_the_oracle(lanes, sigma):
    from oracle.pyoracle import Oracle

    case = s.cavity(8, re=100.0, jitter=0.1, seed=1)
    p = params(rtol=1e-12)
    d = s.Solver(case, p)
    d.set_kernel_shapes(0, lanes, nt=1 if lanes == 16 else 0, rows=2, sigma=sigma)
    o = Oracle(case, p)
    hd, _ = d.iterate(10)
    ho, _, _ = o.iterate(10)
    assert np.allclose(hd, ho, rtol=1e-7, atol=1e-14)
    for n in ("U", "V", "W", "P"):
        assert rel_err(d.get(n), o.get(n)) < 1e-8, n


def test_in_solve_spmv_timing():
    case = s.cavity(10, re=400.0)
    d = s.Solver(case, base(s.FP64))
    d.spmv_profile(0, 1)
    d.iterate(2)
    ms, n = d.spmv_profile(0, -1)
    its = sum(v["iters"] for v in d.last_solve_info().values())
    assert n > 0 and n <= its * 2 and 0.0 < ms < 50.0
    # stopped: further solves add nothing
    d.iterate(1)
    assert d.spmv_profile(0, 0)[1] == n


@pytest.mark.parametrize("ksp", [s.KSP_GMRES, s.KSP_CG])
def test_ilu_sweeps_replayed_from_graphs(ksp):
    case = s.channel(10, 6, 6, jitter=0.1, split=0.2, seed=2, re=500.0)
    out = []
    for graphs in (0, 1):
        p = params(restart=20, pc=s.PC_ILU0, rtol=1e-8, maxit=3000)
        p.ksp[s.KSP_PRESSURE].type = ksp
        p.ksp[s.KSP_PRESSURE].rtol = 1e-6
        p.graphs = graphs
        d = s.Solver(case, p)
        h, _ = d.iterate(5)
        out.append(({n: d.get(n) for n in FIELDS}, h))
        d.close()
    for n in FIELDS:
        assert np.array_equal(out[0][0][n], out[1][0][n]), n
    assert np.array_equal(out[0][1], out[1][1])


def shuffled(case, seed):
    """The same mesh with its element stream in a random order."""
    rng = np.random.default_rng(seed)
    ne = case.etype.size
    perm = rng.permutation(ne)
    lens = np.diff(case.vptr)[perm]
    vptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    vidx = np.concatenate([case.vidx[case.vptr[e]:case.vptr[e + 1]] for e in perm]).astype(np.int32)
    return s.CaseData(case.xyz.copy(), case.etype[perm].copy(), case.tag[perm].copy(), vptr, vidx,
                      list(case.regions), dict(case.meta))


def test_inner_order_falls_back_when_offsets_do_not_fit():
    # 48 K cells in a random input order: some column lies more than 32767
    # positions from its row, so the FP32 cycle stays in the natural order,
    # and the solves are those of the natural-order run
    case = shuffled(s.channel(40, 20, 40, jitter=0.1, split=0.1, seed=3, re=300.0), 9)
    out = {}
    for order in (0, 2):
        p = base(s.MIXED)
        p.inner_order = order
        d = s.Solver(case, p)
        h, _ = d.iterate(2)
        assert d.kernel_shapes()["fp32"]["order"] == 0
        out[order] = (h, d.get("P"))
        d.close()
    assert np.array_equal(out[0][0], out[2][0]) and np.array_equal(out[0][1], out[2][1])
