"""Shared helpers for the tests: hand-built element streams, parameter sets
and field initialisation used by both the oracle and the HIP library."""
from __future__ import annotations

import numpy as np

import sysu_cfd_parallel_amd as s
from sysu_cfd_parallel_amd._abi import Region


def region(rid, type1, init=None, type2=0):
    r = Region()
    r.rid, r.type1, r.type2 = rid, type1, type2
    for k, v in enumerate(init or []):
        r.init[k] = v
    return r


def manual_case(xyz, elems, regions) -> s.CaseData:
    """elems: list of (gmsh_type, tag, [vertex ids])"""
    et = np.array([e[0] for e in elems], dtype=np.int32)
    tg = np.array([e[1] for e in elems], dtype=np.int32)
    vp = np.zeros(len(elems) + 1, dtype=np.int32)
    vi = []
    for k, e in enumerate(elems):
        vi.extend(e[2])
        vp[k + 1] = len(vi)
    return s.CaseData(np.asarray(xyz, dtype=np.float64).ravel(), et, tg, vp, np.asarray(vi, dtype=np.int32),
                      list(regions), {"kind": "manual"})


def tet_cases():
    """A single tetrahedron and two tetrahedra sharing a face, all walls."""
    regs = [region(1, 1), region(100, 5, [0, 0, 0, 1, 300, 0.01, 0, 0])]
    x1 = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]
    one = manual_case(x1, [(2, 1, [0, 2, 1]), (2, 1, [0, 1, 3]), (2, 1, [0, 3, 2]), (2, 1, [1, 2, 3]),
                           (4, 100, [0, 1, 2, 3])], regs)
    x2 = x1 + [[1, 1, 1]]
    two = manual_case(x2, [(2, 1, [0, 2, 1]), (2, 1, [0, 1, 3]), (2, 1, [0, 3, 2]), (2, 1, [1, 4, 2]),
                           (2, 1, [1, 3, 4]), (2, 1, [2, 4, 3]), (4, 100, [0, 1, 2, 3]), (4, 100, [1, 2, 3, 4])], regs)
    return one, two


def params(restart=30, pc=s.PC_JACOBI, rtol=1e-8, maxit=1000, ksp=s.KSP_GMRES, turb=0):
    p = s.default_params()
    s.set_all_ksp(p, ksp, restart, pc, rtol, maxit)
    p.turb_model = turb
    return p


def random_state(n_cell, n_bnd, n_face, seed=0):
    """Random but physically shaped fields for kernel-level parity checks."""
    rng = np.random.default_rng(seed)
    f = {
        "U": rng.uniform(-1, 1, n_cell), "V": rng.uniform(-1, 1, n_cell), "W": rng.uniform(-1, 1, n_cell),
        "P": rng.uniform(-1, 1, n_cell), "R": rng.uniform(0.8, 1.2, n_cell),
        "VISLAM": rng.uniform(0.005, 0.02, n_cell), "VISTUR": rng.uniform(0.0, 0.01, n_cell),
        "TE": rng.uniform(0.01, 0.1, n_cell), "ED": rng.uniform(0.01, 0.1, n_cell),
        "DELTAP": rng.uniform(-0.1, 0.1, n_cell), "APR": rng.uniform(0.5, 2.0, n_cell),
        "BU": rng.uniform(-1, 1, n_bnd), "BV": rng.uniform(-1, 1, n_bnd), "BW": rng.uniform(-1, 1, n_bnd),
        "BRO": rng.uniform(0.8, 1.2, n_bnd), "BPRE": rng.uniform(-1, 1, n_bnd),
        "BTE": rng.uniform(0.01, 0.1, n_bnd), "BED": rng.uniform(0.01, 0.1, n_bnd),
        "RUFACE": rng.uniform(-0.01, 0.01, n_face),
        "DUDX": rng.uniform(-1, 1, 3 * n_cell), "DVDX": rng.uniform(-1, 1, 3 * n_cell),
        "DWDX": rng.uniform(-1, 1, 3 * n_cell), "DPDX": rng.uniform(-1, 1, 3 * n_cell),
    }
    return f


def push(obj, state):
    for k, v in state.items():
        obj.set(k, v)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)


def conduction_column(n=6, t_cold=300.0, t_hot=400.0, t0=350.0, mu=0.01):
    """n unit-width hexes stacked in z between an isothermal cold wall (z = 0)
    and an isothermal hot wall (z = 1); symmetry planes on the four sides."""
    h = 1.0 / n
    xyz = []
    for j in range(n + 1):
        z = j * h
        xyz += [[0, 0, z], [1, 0, z], [1, 1, z], [0, 1, z]]
    elems = [(3, 1, [0, 1, 2, 3]), (3, 6, [4 * n, 4 * n + 1, 4 * n + 2, 4 * n + 3])]
    for j in range(n):
        a, b = 4 * j, 4 * (j + 1)
        for q in ([a, a + 1, b + 1, b], [a + 1, a + 2, b + 2, b + 1], [a + 2, a + 3, b + 3, b + 2],
                  [a + 3, a, b, b + 3]):
            elems.append((3, 4, q))
    for j in range(n):
        a, b = 4 * j, 4 * (j + 1)
        elems.append((5, 100, [a, a + 1, a + 2, a + 3, b, b + 1, b + 2, b + 3]))
    cold, hot = region(1, 1), region(6, 1)
    cold.fixed_value, hot.fixed_value = t_cold, t_hot
    regs = [cold, hot, region(4, 4), region(100, 5, [0, 0, 0, 1, t0, mu, 0, 0])]
    return manual_case(xyz, elems, regs)


def heated(case, t_wall=350.0, t_in=300.0, t0=300.0):
    """Isothermal walls at t_wall, inlet fluid at t_in, initial fluid at t0."""
    for r in case.regions:
        if r.type1 == 1:
            r.type2, r.fixed_value = 0, t_wall
        elif r.type1 == 2:
            r.init[5] = t_in
        elif r.type1 == 5:
            r.init[4] = t0
    return case


def parse_tecplot(path) -> dict:
    """A Tecplot ASCII result of Output2Tecplot (navierIO.cpp:161-246) read
    back: the variable names, the vertex coordinates, the cell-centred
    arrays and the 8 vertex ids (1-based) of each cell."""
    text = open(path).read()
    head, zone, body = text.split("\n", 2)
    names = [t.strip().strip('"') for t in head.split("=", 1)[1].replace('" "', '","').split(",")]
    names = [n for n in names if n]
    kv = dict(part.split("=", 1) for part in zone[len("ZONE "):].split(", ")[:2])
    nv, ne = int(kv["N"]), int(kv["E"])
    tok = body.split()
    xyz = np.array(tok[:3 * nv], dtype=float).reshape(3, nv).T
    off = 3 * nv
    out = {"names": names, "xyz": xyz}
    for name in names[3:]:
        out[name] = np.array(tok[off:off + ne], dtype=float)
        off += ne
    out["conn"] = np.array(tok[off:off + 8 * ne], dtype=np.int64).reshape(ne, 8)
    assert off + 8 * ne == len(tok)
    return out


# hex corner bits a | b << 1 | c << 2 in gmsh hex order
_HEX = (0, 1, 3, 2, 4, 5, 7, 6)
# the six hex faces as quads whose right-hand normal points into the cell
_INWARD = ((0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3))


def hybrid_channel(nx, ny, nz, prism_frac=0.3, pyramid_frac=0.3, jitter=0.0, seed=0, re=1000.0, lx=2.0, ly=1.0,
                   lz=1.0, turbulent=False):
    """The channel of ``s.channel`` (inlet x = 0, outlet x = lx, walls y = 0, ly,
    symmetry z = 0, lz) on a hybrid mesh: whole z-columns of the hex lattice
    split into two prisms along the (0,0)-(1,1) diagonal (their z faces
    triangles, conforming from layer to layer), and a random share of the
    other hexes split into six pyramids around a vertex at the cell's corner
    average (their faces stay the hex's quads); nodes jittered by up to
    jitter times the spacing, boundary nodes kept on their planes."""
    rng = np.random.default_rng(seed)
    h = np.array([lx / nx, ly / ny, lz / nz])
    I, J, K = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    ijk = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    xyz = ijk * h
    n = np.array([nx, ny, nz])
    move = (ijk > 0) & (ijk < n)
    xyz = xyz + move * rng.uniform(-jitter, jitter, xyz.shape) * h
    nid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    pts = [tuple(p) for p in xyz]
    prism_col = rng.uniform(size=(nx, ny)) < prism_frac
    pyramid = rng.uniform(size=(nx, ny, nz)) < pyramid_frac
    tag_of = {(0, 0): 2, (0, 1): 3, (1, 0): 1, (1, 1): 1, (2, 0): 4, (2, 1): 4}
    bnd, vol = [], []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                v = [nid(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) for c in range(8)]
                cell = (i, j, k)
                for f, quad in enumerate(_INWARD):
                    axis = (2, 2, 1, 1, 0, 0)[f]
                    side = (0, 1, 0, 1, 0, 1)[f]
                    if 0 <= cell[axis] + (2 * side - 1) < n[axis]:
                        continue
                    out = [v[c] for c in reversed(quad)]  # outward
                    if axis == 2 and prism_col[i, j]:
                        if side == 0:
                            bnd += [(2, tag_of[(2, 0)], [v[0], v[3], v[1]]), (2, tag_of[(2, 0)], [v[0], v[2], v[3]])]
                        else:
                            bnd += [(2, tag_of[(2, 1)], [v[4], v[5], v[7]]), (2, tag_of[(2, 1)], [v[4], v[7], v[6]])]
                    else:
                        bnd.append((3, tag_of[(axis, side)], out))
                if prism_col[i, j]:
                    vol.append((6, 100, [v[0], v[1], v[3], v[4], v[5], v[7]]))
                    vol.append((6, 100, [v[0], v[3], v[2], v[4], v[7], v[6]]))
                elif pyramid[i, j, k]:
                    apex = len(pts)
                    pts.append(tuple(np.mean([pts[c] for c in v], axis=0)))
                    for quad in _INWARD:
                        vol.append((7, 100, [v[c] for c in quad] + [apex]))
                else:
                    vol.append((5, 100, [v[c] for c in _HEX]))
    regions = s.channel(1, 1, 1, re=re, turbulent=turbulent).regions
    case = manual_case(np.array(pts), bnd + vol, regions)
    case.meta = {"kind": "channel", "mesh": "hybrid", "nx": nx, "ny": ny, "nz": nz, "jitter": jitter, "seed": seed}
    return case
