"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol of include/simple355.h, and its mesh set-up (face builder,
metrics, CSR pattern) equals the oracle's restatement bit for bit."""
import re
from pathlib import Path

import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, tet_cases
from oracle.pyoracle import Oracle
from sysu_cfd_parallel_amd._abi import MESH_DBL, MESH_INT

HEADER = Path(__file__).resolve().parents[1] / "include" / "simple355.h"


def declared_symbols():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(s355_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_default_params_match_reference_defaults(lib):
    p = s.default_params()
    assert list(p.urf) == [0.6, 0.6, 0.6, 0.5, 0.8, 0.8, 0.8, 0.8]  # navier.cpp:500-507
    assert p.residual_steady == 1e-6 and p.steady == 1 and p.cell_pressure_ref == 0
    v, pr, sc = p.ksp
    assert (v.type, v.rtol, v.maxit) == (s.KSP_GMRES, 1e-8, 1000)  # navier_velocity.cpp:28
    assert (pr.type, pr.rtol, pr.maxit) == (s.KSP_CG, 1e-6, 1000)  # navier_pressure.cpp:28, MPIStructure.cpp:464
    assert (sc.type, sc.rtol, sc.maxit) == (s.KSP_GMRES, 1e-8, 500)  # navier_scalar.cpp:98
    assert v.restart == 30 and v.dtol == 1e5 and v.abstol == 1e-50


CASES = {
    "cavity_uniform": lambda: s.cavity(6),
    "cavity_jitter": lambda: s.cavity(5, jitter=0.15, seed=11),
    "cavity_tets": lambda: s.cavity(3, tets=True, jitter=0.1, seed=2),
    "channel_split": lambda: s.channel(8, 4, 6, jitter=0.15, split=0.1, seed=7),
    "channel_split_heavy": lambda: s.channel(5, 3, 4, jitter=0.1, split=0.5, seed=3),
    "step_tets": lambda: s.step(2, upstream=2, downstream=4, span=2),
    "one_tet": lambda: tet_cases()[0],
    "two_tets": lambda: tet_cases()[1],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_ingest_bitwise_equal_to_oracle(name):
    case = CASES[name]()
    hm = s.HostMesh(case)
    o = Oracle(case, params())
    assert hm.sizes.as_dict() == o.sizes.as_dict()
    for n in MESH_INT:
        assert np.array_equal(hm.ints(n), o.ints(n)), n
    for n in MESH_DBL:
        a, b = hm.dbls(n), o.dbls(n)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), n


def test_split_hex_neighbours_become_polyhedra():
    case = s.channel(6, 4, 5, jitter=0.1, split=0.2, seed=4)
    hm = s.HostMesh(case)
    nface = hm.ints("CELL_NFACE")
    assert nface.max() > 6 and nface.max() <= 12
    assert set(np.unique(nface)) <= set(range(4, 13))


def test_bad_meshes_are_rejected(lib):
    one, _ = tet_cases()
    # dangling boundary element (its face matches no cell)
    bad = s.CaseData(np.concatenate([one.xyz, [5.0, 5.0, 5.0]]), np.append(one.etype, 2).astype(np.int32),
                     np.append(one.tag, 1).astype(np.int32),
                     np.append(one.vptr, one.vptr[-1] + 3).astype(np.int32),
                     np.concatenate([one.vidx, [0, 1, 4]]).astype(np.int32), one.regions)
    with pytest.raises(s.S355Error):
        s.HostMesh(bad)
    # unknown boundary region
    regs = [r for r in one.regions if r.rid != 1]
    with pytest.raises(s.S355Error):
        s.HostMesh(s.CaseData(one.xyz, one.etype, one.tag, one.vptr, one.vidx, regs))


def _elements(case, keep):
    vp, vi = [0], []
    for k in np.nonzero(keep)[0]:
        vi += list(case.vidx[case.vptr[k]:case.vptr[k + 1]])
        vp.append(len(vi))
    return case.etype[keep], case.tag[keep], np.array(vp, np.int32), np.array(vi, np.int32)


def test_malformed_element_streams_are_errors(lib):
    """Empty streams, boundary elements without cells, vertex ids out of
    range, unknown element types and short vertex lists are refused with an
    error, never read past their arrays."""
    one, _ = tet_cases()
    z = np.zeros(0, np.int32)
    cases = [s.CaseData(one.xyz, z, z, np.zeros(1, np.int32), z, one.regions),
             s.CaseData(one.xyz, *_elements(one, one.etype < 4), one.regions)]
    for bad_id in (99, -1):
        vidx = one.vidx.copy()
        vidx[-1] = bad_id
        cases.append(s.CaseData(one.xyz, one.etype, one.tag, one.vptr, vidx, one.regions))
    cases.append(s.CaseData(one.xyz, np.where(one.etype == 4, 9, one.etype).astype(np.int32), one.tag, one.vptr,
                            one.vidx, one.regions))
    vptr = one.vptr.copy()
    vptr[-1] -= 1
    cases.append(s.CaseData(one.xyz, one.etype, one.tag, vptr, one.vidx[:-1].copy(), one.regions))
    for c in cases:
        with pytest.raises(s.S355Error):
            s.HostMesh(c)


def test_ctx_create_fails_loudly_without_gpu(lib):
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(s.S355Error):
        s.Solver(s.cavity(2), s.default_params())
