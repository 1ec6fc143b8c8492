"""Meshes read from files on the device path: a generated case written as a
gmsh 2 ASCII file and in readBin's binary form (RootStructure.cpp:428-487,
61-116), read back, and run for a few SIMPLE steps on the GPU against the
oracle running the in-memory case -- the file ingest, the device geometry
and the loop together; and the hand-written two-hex file (an unreferenced
vertex, boundary quads of three regions) stepped on both."""
from pathlib import Path

import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, region, rel_err
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("binary", [False, True])
def test_file_mesh_steps_like_the_in_memory_case(tmp_path, binary):
    case = s.channel(10, 5, 6, jitter=0.15, split=0.2, seed=6, re=300.0)
    path = tmp_path / ("case.bin" if binary else "case.msh")
    s.write_mesh(case, path, binary=binary)
    back = s.read_mesh(path, case.regions, binary=binary)
    p = params(rtol=1e-12)
    p.residual_steady = 0.0
    d, o = s.Solver(back, p), Oracle(case, p)
    hd, _ = d.iterate(6)
    ho, _, _ = o.iterate(6)
    assert np.allclose(hd, ho, rtol=1e-7, atol=1e-14)
    for n in ("U", "V", "W", "P"):
        assert rel_err(d.get(n), o.get(n)) <= 1e-8, n
    d.close()


def test_handwritten_two_hex_file_on_the_device():
    regions = [region(1, 1), region(2, 2, [1.0, 0, 0, 0, 1.0, 300, 0, 0]), region(3, 3),
               region(100, 5, [0, 0, 0, 1.0, 300, 0.01, 0, 0])]
    case = s.read_mesh(GOLDEN / "two_hex.msh", regions)
    p = params(rtol=1e-12)
    p.residual_steady = 0.0
    d, o = s.Solver(case, p), Oracle(case, p)
    assert d.sizes.n_cell == 2 and d.sizes.n_face == 11
    assert np.isclose(d.dbls("CELL_VOL").sum(), 2.0)
    hd, _ = d.iterate(3)
    ho, _, _ = o.iterate(3)
    assert np.allclose(hd, ho, rtol=1e-7, atol=1e-14)
    for n in ("U", "V", "W", "P"):
        assert rel_err(d.get(n), o.get(n)) <= 1e-8, n
    d.close()
