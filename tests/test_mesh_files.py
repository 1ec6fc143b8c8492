"""Mesh files in the reference's formats (gmsh 2 ASCII, RootStructure.cpp:
428-487, and the binary form of readBin, 61-116): a hand-written two-hex
file with an unreferenced vertex, and write/read round trips of generated
cases whose ingest must equal the in-memory case bit for bit."""
from pathlib import Path

import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import region
from sysu_cfd_parallel_amd._abi import MESH_DBL, MESH_INT

GOLDEN = Path(__file__).resolve().parent / "golden"


def two_hex_regions():
    return [region(1, 1), region(2, 2, [1.0, 0, 0, 0, 1.0, 300, 0, 0]), region(3, 3),
            region(100, 5, [0, 0, 0, 1.0, 300, 0.01, 0, 0])]


def test_handwritten_gmsh_two_hexes():
    case = s.read_mesh(GOLDEN / "two_hex.msh", two_hex_regions())
    # the 13th vertex is referenced by no element and is dropped
    assert case.xyz.size == 3 * 12
    assert list(case.etype) == [3] * 10 + [5, 5]
    assert list(case.tag) == [1, 1, 2, 3, 1, 1, 1, 1, 1, 1, 100, 100]
    assert list(case.vidx[-8:]) == [1, 2, 5, 4, 7, 8, 11, 10]
    m = s.HostMesh(case)
    assert (m.sizes.n_cell, m.sizes.n_face, m.sizes.n_bnd, m.sizes.n_interior) == (2, 11, 10, 1)
    assert np.isclose(m.dbls("CELL_VOL").sum(), 2.0)


@pytest.mark.parametrize("binary", [False, True])
@pytest.mark.parametrize("make", [lambda: s.channel(6, 3, 4, jitter=0.15, split=0.3, seed=2),
                                  lambda: s.step(2, upstream=2, downstream=3, span=2)])
def test_round_trip_ingest_is_bitwise(tmp_path, binary, make):
    case = make()
    path = tmp_path / ("case.bin" if binary else "case.msh")
    s.write_mesh(case, path, binary=binary)
    back = s.read_mesh(path, case.regions, binary=binary)
    assert np.array_equal(back.xyz, case.xyz)
    assert np.array_equal(back.vidx, case.vidx) and np.array_equal(back.etype, case.etype)
    a, b = s.HostMesh(case), s.HostMesh(back)
    for n in MESH_INT:
        assert np.array_equal(a.ints(n), b.ints(n)), n
    for n in MESH_DBL:
        assert np.array_equal(a.dbls(n), b.dbls(n)), n


def test_bad_files_are_errors(tmp_path):
    with pytest.raises(s.S355Error):
        s.read_mesh(tmp_path / "missing.msh", [])
    bad = tmp_path / "bad.msh"
    bad.write_text("$Nodes\n2\n1 0 0 0\n2 1 0 0\n$EndNodes\n$Elements\n1\n1 4 1 100 1 2 3 4\n$EndElements\n")
    with pytest.raises(s.S355Error):
        s.read_mesh(bad, [])
