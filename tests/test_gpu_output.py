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


def test_monitor_records_and_file(tmp_path):
    # OutputMoniter: wall force = summed wall shear, fluid record = sum W^2 vol
    case = s.cavity(6, re=100.0, jitter=0.1, seed=4)
    p = params()
    p.dt = 0.5
    with s.Solver(case, p) as d:
        d.output_schedule(5, tmp_path)
        d.iterate(3)
        rec = d.monitor_records()
        shear = d.get("SHEAR").reshape(-1, 3)
        brid = d.ints("BND_RID")
        w, vol, crid = d.get("W"), d.dbls("CELL_VOL"), d.ints("CELL_RID")
        t, _ = d.time_state()
    assert t == 1.5  # steady steps advance cur_time by dt too
    types = {r.rid: (r.type1, r.type2) for r in case.regions}
    want = [r for r in sorted(types) if types[r][0] in (1, 5) and r != 0]
    assert [r["rid"] for r in rec] == want
    for r in rec:
        if r["type1"] == 1:
            assert np.allclose(r["force"], shear[brid == r["rid"]].sum(0), rtol=1e-12, atol=1e-15)
            assert r["scalar"] == 0.0
        else:
            assert np.allclose(r["scalar"], np.sum(w[crid == r["rid"]] ** 2 * vol[crid == r["rid"]]), rtol=1e-12)
            assert np.all(r["force"] == 0.0)
    lines = (tmp_path / "monitorResult.mon").read_text().split("\n")
    assert lines[0].startswith("%20s || " % "time")
    assert "drag/lift force" in lines[0] and "w energy" in lines[0]
    assert len(lines) == 1 + 3 + 1  # header, one line per step, the final newline
    assert lines[3].startswith("%20e || " % 1.5)
    vals = [float(v) for v in lines[3].replace("||", ",").replace("|", ",").split(",")[1:] if v.strip()]
    flat = [x for r in rec for x in (*r["force"], r["scalar"])]
    assert np.allclose(vals, flat, rtol=1e-6, atol=0)
