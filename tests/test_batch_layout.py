"""Batch layout of independent cases (host side, no GPU): every array of the
concatenated mesh is the cases' own arrays with their index offsets, the
boundary faces of all cases come first, and each case keeps its own face
order.  The device batch runs on exactly this layout."""
import numpy as np
import pytest

import sysu_cfd_parallel_amd as s


def cases():
    return [s.cavity(4, re=100.0), s.channel(6, 3, 4, jitter=0.1, split=0.2, seed=5), s.cavity(3, tets=True, jitter=0.1),
            s.step(2, upstream=2, downstream=4, span=2)]


@pytest.fixture(scope="module")
def built():
    cs = cases()
    parts = [s.HostMesh(c) for c in cs]
    batch = s.HostMesh(cs)
    return parts, batch


def offsets(parts, attr):
    return np.concatenate([[0], np.cumsum([getattr(p.sizes, attr) for p in parts])])


def face_map(parts):
    """global face id of (case, local face)"""
    nb = offsets(parts, "n_bnd")
    ni = np.concatenate([[0], np.cumsum([p.sizes.n_face - p.sizes.n_bnd for p in parts])])
    total_nb = nb[-1]
    maps = []
    for k, p in enumerate(parts):
        f = np.arange(p.sizes.n_face)
        maps.append(np.where(f < p.sizes.n_bnd, nb[k] + f, total_nb + ni[k] + (f - p.sizes.n_bnd)))
    return maps


def test_sizes_add_up(built):
    parts, batch = built
    for a in ("n_vert", "n_cell", "n_face", "n_bnd", "n_interior", "nnz", "n_elem", "n_warped"):
        assert getattr(batch.sizes, a) == sum(getattr(p.sizes, a) for p in parts), a


def test_face_arrays_follow_the_face_map(built):
    parts, batch = built
    fm = face_map(parts)
    c0 = offsets(parts, "n_cell")
    c1, c2, fb = batch.ints("FACE_C1"), batch.ints("FACE_C2"), batch.ints("FACE_BND")
    seen = np.zeros(batch.sizes.n_face, dtype=bool)
    for k, p in enumerate(parts):
        g = fm[k]
        assert not seen[g].any()
        seen[g] = True
        pc1, pc2, pb = p.ints("FACE_C1"), p.ints("FACE_C2"), p.ints("FACE_BND")
        assert np.array_equal(c1[g], pc1 + c0[k])
        assert np.array_equal(c2[g], np.where(pc2 >= 0, pc2 + c0[k], pc2))
        assert np.array_equal(fb[g], np.where(pb >= 0, pb + offsets(parts, "n_bnd")[k], pb))
        for name in ("FACE_X", "FACE_N", "FACE_AREA", "FACE_LAMBDA", "FACE_RLENCOS", "FACE_XPAC", "FACE_XNAC"):
            w = 3 if name in ("FACE_X", "FACE_N", "FACE_XPAC", "FACE_XNAC") else 1
            a = batch.dbls(name).reshape(-1, w)
            assert np.array_equal(a[g], p.dbls(name).reshape(-1, w)), name
    assert seen.all()
    # boundary faces of every case first: face f is boundary record f
    assert np.array_equal(fb[: batch.sizes.n_bnd], np.arange(batch.sizes.n_bnd))
    assert (fb[batch.sizes.n_bnd:] < 0).all()


def test_cell_arrays_and_csr_are_offset_copies(built):
    parts, batch = built
    fm = face_map(parts)
    c0, z0 = offsets(parts, "n_cell"), offsets(parts, "nnz")
    rp, col, dg = batch.ints("CSR_ROWPTR"), batch.ints("CSR_COL"), batch.ints("CSR_DIAG")
    cface, ccell, nface = batch.ints("CELL_FACE"), batch.ints("CELL_CELL"), batch.ints("CELL_NFACE")
    kmax = cface.size // batch.sizes.n_cell
    for k, p in enumerate(parts):
        cs = slice(c0[k], c0[k + 1])
        assert np.array_equal(rp[c0[k]: c0[k + 1] + 1], p.ints("CSR_ROWPTR") + z0[k])
        assert np.array_equal(col[z0[k]: z0[k + 1]], p.ints("CSR_COL") + c0[k])
        assert np.array_equal(dg[cs], p.ints("CSR_DIAG") + z0[k])
        assert np.array_equal(nface[cs], p.ints("CELL_NFACE"))
        pf = p.ints("CELL_FACE").reshape(-1, kmax)
        pc = p.ints("CELL_CELL").reshape(-1, kmax)
        bf = cface.reshape(-1, kmax)[cs]
        bc = ccell.reshape(-1, kmax)[cs]
        assert np.array_equal(bf, np.where(pf >= 0, fm[k][np.maximum(pf, 0)], pf))
        assert np.array_equal(bc, np.where(pc >= 0, pc + c0[k], pc))
        for name in ("CELL_VOL", "CELL_X"):
            w = 3 if name == "CELL_X" else 1
            assert np.array_equal(batch.dbls(name).reshape(-1, w)[cs], p.dbls(name).reshape(-1, w))
    # rows never reach into another case: the matrix is block diagonal
    rows = np.repeat(np.arange(batch.sizes.n_cell), np.diff(rp))
    case_of = np.searchsorted(c0, rows, side="right") - 1
    assert np.array_equal(np.searchsorted(c0, col, side="right") - 1, case_of)


def test_region_ids_are_shifted_per_case(built):
    parts, batch = built
    b0, c0 = offsets(parts, "n_bnd"), offsets(parts, "n_cell")
    stride = 1 + max(max(r.rid for r in c.regions) for c in cases())
    brid, crid = batch.ints("BND_RID"), batch.ints("CELL_RID")
    for k, p in enumerate(parts):
        assert np.array_equal(brid[b0[k]: b0[k + 1]], p.ints("BND_RID") + k * stride)
        assert np.array_equal(crid[c0[k]: c0[k + 1]], p.ints("CELL_RID") + k * stride)


def test_empty_batch_is_rejected(lib):
    with pytest.raises(s.S355Error):
        s.HostMesh([])
