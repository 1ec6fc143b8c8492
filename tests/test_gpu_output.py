"""Result and restart files from a device context: a snapshot copies the
cell fields to pinned host memory in stream order, a writer thread formats
the files while the device keeps iterating, a backup restores the fields
it holds, and NSSolve's output schedule (navier.cpp:113-122, 149) writes
the backup and the Tecplot result every noutput steps."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s
from helpers import params, parse_tecplot
from test_output_files import read_postcycas

pytestmark = pytest.mark.gpu

CELL = ("R", "U", "V", "W", "P", "T", "VISLAM", "VISTUR", "TE", "ED")


def fields(d):
    return {k: d.get(k) for k in CELL}


def g6(a):
    return np.array([float("%g" % v) for v in a])


def laminar():
    return s.cavity(6, re=100.0, jitter=0.1, seed=4), params()


def turbulent():
    return s.channel(8, 4, 4, jitter=0.1, seed=6, re=5000.0, turbulent=True), params(turb=1)


def test_snapshot_is_taken_in_stream_order(tmp_path):
    case, p = laminar()
    with s.Solver(case, p) as d:
        d.iterate(2)
        want = fields(d)
        d.snapshot()
        d.write_output("tec_data", tmp_path / "data", async_=True)
        d.iterate(2)  # runs while (or after) the writer formats the snapshot
        d.write_output("tec_geo", tmp_path / "geo")
        d.wait_output()
        later = fields(d)
        head, xyz, conn, vals = read_postcycas(tmp_path / "geo", tmp_path / "data")
        assert head == (0, 1, d.sizes.n_cell, d.sizes.n_vert, 0, 6)
        for k, nm in enumerate(("P", "U", "V", "W", "R", "T")):
            assert np.array_equal(vals[k], want[nm]), nm
        assert not np.array_equal(later["U"], want["U"])
        assert np.array_equal(conn.ravel(), d.ints("CELL_VERT") + 1)
        assert np.array_equal(xyz.ravel(), d.dbls("VERT"))


@pytest.mark.parametrize("make", [laminar, turbulent])
def test_backup_round_trip_restores_the_fields(tmp_path, make):
    case, p = make()
    turb = p.turb_model == 1
    with s.Solver(case, p) as a:
        a.iterate(3)
        want = fields(a)
        a.snapshot()
        a.write_output("backup", tmp_path / "res.sav", async_=True)
        a.wait_output()
    with s.Solver(case, p) as b:
        init = fields(b)
        b.read_backup(tmp_path / "res.sav")
        got = fields(b)
    read = ("R", "U", "V", "W", "P", "T", "VISLAM") + (("VISTUR", "TE", "ED") if turb else ())
    for k in CELL:
        assert np.array_equal(got[k], g6(want[k]) if k in read else init[k]), k


def test_tecplot_from_the_device_matches_the_fields(tmp_path):
    case, p = turbulent()
    with s.Solver(case, p) as d:
        d.iterate(2)
        f = fields(d)
        d.snapshot()
        d.write_output("tecplot", tmp_path / "res0000.dat")
        t = parse_tecplot(tmp_path / "res0000.dat")
    assert t["names"] == ["x", "y", "z", "p", "u", "v", "w", "ro", "T", "Ma", "mu", "te", "ed"]
    for col, k in (("p", "P"), ("u", "U"), ("v", "V"), ("w", "W"), ("ro", "R"), ("T", "T"), ("te", "TE"),
                   ("ed", "ED")):
        assert np.allclose(t[col], f[k], atol=5.1e-6, rtol=0), col
    assert np.allclose(t["mu"], f["VISLAM"] + f["VISTUR"], atol=5.1e-6, rtol=0)


@pytest.mark.parametrize("binary", [False, True])
def test_output_schedule_of_nssolve(tmp_path, binary):
    case, p = laminar()
    with s.Solver(case, p) as d:
        d.output_schedule(2, tmp_path, binary=binary)
        assert (tmp_path / "tec" / ("geo" if binary else "grid.dat")).exists()
        d.iterate(3)  # output after steps 1 and 3, and a last Tecplot file
        f = fields(d)
    back = s.read_backup_file(tmp_path / "res.sav", len(f["U"]))
    for k in ("R", "U", "V", "W", "P", "T", "VISLAM"):
        assert np.array_equal(back[k], g6(f[k])), k
    t = parse_tecplot(tmp_path / "tec" / "res0000.dat")
    assert np.allclose(t["u"], f["U"], atol=5.1e-6, rtol=0)
    if binary:
        _, _, _, vals = read_postcycas(tmp_path / "tec" / "geo", tmp_path / "tec" / "data")
        assert np.array_equal(vals[1], f["U"])


def test_file_errors_leave_the_context_usable(tmp_path):
    case, p = laminar()
    with s.Solver(case, p) as d:
        with pytest.raises(s.S355Error) as err:
            d.read_backup(tmp_path / "missing.sav")
        assert err.value.code == s._abi.S355_E_IO
        with pytest.raises(s.S355Error):
            d.write_output("tecplot", tmp_path / "t.dat")  # no snapshot yet
        d.snapshot()
        d.write_output("backup", tmp_path / "no_dir" / "res.sav", async_=True)
        with pytest.raises(s.S355Error) as err:
            d.wait_output()
        assert err.value.code == s._abi.S355_E_IO
        d.iterate(1)
        d.wait_output()  # the error was reported once
