"""MI355X SIMPLE / k-epsilon hot path: Python access to the HIP library.

The library (``_build/libsimple355.so``, built by ``sysu_cfd_parallel_amd.build``)
exports the C ABI of ``include/simple355.h``; this module wraps it with
ctypes for case generation, the device context and the SIMPLE loop.  There
is no CPU fallback: without the built library or a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path
from typing import Optional

import numpy as np

from . import _abi
from ._abi import (FIELD_ID, FIELD_KIND, MESH_DBL_ID, MESH_INT_ID, FP64, KSP_CG, KSP_GMRES, KSP_PRESSURE,
                   KSP_SCALAR, KSP_VELOCITY, MIXED, PC_ILU0, PC_JACOBI, PC_NONE, PHASE_CORRECT, PHASE_KEPS_EPS,
                   PHASE_KEPS_K, PHASE_KEPS_POST, PHASE_MOMENTUM, PHASE_PRESSURE, PHASE_VELOCITY_RESIDUAL, KspOpts,
                   MeshDesc, Params, Region, Sizes, SolveInfo, field_len, mesh_dbl_len, mesh_int_len)

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libsimple355.so"

__all__ = [
    "LIB_PATH", "library", "CaseData", "cavity", "channel", "step", "read_mesh", "write_mesh", "default_params",
    "set_all_ksp", "HostMesh",
    "Solver", "BatchSolver", "S355Error", "NotConverged", "FP64", "MIXED", "KSP_GMRES", "KSP_CG", "PC_NONE", "PC_JACOBI",
    "PC_ILU0", "KSP_VELOCITY", "KSP_PRESSURE", "KSP_SCALAR", "PHASE_MOMENTUM", "PHASE_VELOCITY_RESIDUAL",
    "PHASE_PRESSURE", "PHASE_CORRECT", "PHASE_KEPS_K", "PHASE_KEPS_EPS", "PHASE_KEPS_POST",
]


class S355Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NotConverged(S355Error):
    """A Krylov solve ended without converging (the reference's ConvergeError);
    `info` holds the solve record(s) when the call returns them."""

    def __init__(self, code: int, msg: str, info=None):
        super().__init__(code, msg)
        self.info = info


_lib: Optional[C.CDLL] = None


def library(path: Optional[Path] = None) -> C.CDLL:
    """Load the HIP library once; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise S355Error(-1, f"HIP library not built: {p} (run python -m sysu_cfd_parallel_amd.build)")
    lib = C.CDLL(str(p))
    _abi.declare(lib, "s355")
    I, V = C.c_int, C.c_void_p
    pI, pD = C.POINTER(C.c_int), C.POINTER(C.c_double)
    more = {
        "default_params": (None, [C.POINTER(Params)]),
        "case_build": (I, [I, pI, pD, C.POINTER(V)]),
        "case_sizes": (I, [V, pI]),
        "case_copy": (I, [V, pD, pI, pI, pI, pI, C.POINTER(Region)]),
        "case_free": (None, [V]),
        "case_read": (I, [C.c_char_p, I, C.POINTER(V)]),
        "case_write": (I, [C.POINTER(MeshDesc), C.c_char_p, I]),
        "mesh_build": (I, [C.POINTER(MeshDesc), C.POINTER(V)]),
        "mesh_build_batch": (I, [I, C.POINTER(MeshDesc), C.POINTER(V)]),
        "mesh_sizes": (I, [V, C.POINTER(Sizes)]),
        "mesh_get_int": (I, [V, I, pI, C.c_size_t]),
        "mesh_get_dbl": (I, [V, I, pD, C.c_size_t]),
        "mesh_free": (None, [V]),
        "ctx_create": (I, [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(V)]),
        "ctx_destroy": (I, [V]),
        "batch_create": (I, [I, C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(V)]),
        "batch_size": (I, [V, pI]),
        "batch_layout": (I, [V, I, pI, C.c_size_t]),
        "batch_residuals": (I, [V, pD, I]),
        "batch_solve_info": (I, [V, I, C.POINTER(SolveInfo), I]),
        "ctx_sizes": (I, [V, C.POINTER(Sizes)]),
        "ctx_set_params": (I, [V, C.POINTER(Params)]),
        "synchronize": (I, [V]),
        "save_velocity": (I, [V]),
        "velocity_residual": (I, [V]),
        "run_phase": (I, [V, I]),
        "residuals": (I, [V, pD, I]),
        "gradient_noalloc": (I, [V, I, I, I]),
        "kepsilon_sources": (I, [V]),
        "bc_kepsilon": (I, [V]),
        "build_scalar_matrix": (I, [V, I]),
        "eps_wall_rows": (I, [V]),
        "solve_scalar": (I, [V, I, C.c_double, I, C.POINTER(SolveInfo)]),
        "kepsilon_post_k": (I, [V]),
        "kepsilon_post_eps": (I, [V]),
        "log_stage_push": (I, [V, I]),
        "log_stage_pop": (I, [V]),
        "log_stage_times": (I, [V, pD, I]),
        "calculate_velocity": (I, [V, C.POINTER(SolveInfo)]),
        "calculate_pressure": (I, [V, C.POINTER(SolveInfo)]),
        "update_tur_kepsilon": (I, [V, C.POINTER(SolveInfo)]),
        "update_energy": (I, [V, C.POINTER(SolveInfo)]),
        "energy_coefficients": (I, [V]),
        "bc_temperature": (I, [V]),
        "spmv_time": (I, [V, I, I, I, pD]),
        "last_solve_info": (I, [V, C.POINTER(SolveInfo)]),
        "copy_time": (I, [V, C.c_size_t, I, pD]),
        "kernel_shapes": (I, [V, pI, pI, pI, pI, pI, pI, pI]),
        "tune_report": (I, [V, I, pD, pD, pI]),
        "set_kernel_shapes": (I, [V, I, I, I, I, I]),
        "spmv_profile": (I, [V, I, I, pD, pI]),
        "mesh_write_result": (I, [V, I, C.c_char_p, C.POINTER(_abi.Result)]),
        "backup_read": (I, [C.c_char_p, I, I, pD]),
        "output_snapshot": (I, [V]),
        "output_write": (I, [V, I, C.c_char_p, I]),
        "output_wait": (I, [V]),
        "read_backup": (I, [V, C.c_char_p]),
        "output_schedule": (I, [V, I, I, C.c_char_p]),
        "save_transient_old_data": (I, [V]),
        "unsteady_sources": (I, [V, I]),
        "time_state": (I, [V, pD, pI]),
        "monitor_records": (I, [V, pD, I, pI]),
        "monitor_write": (I, [V, C.c_char_p, C.c_double, I]),
    }
    for name, (res, args) in more.items():
        fn = getattr(lib, f"s355_{name}")
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(lib: C.CDLL, rc: int, what: str, info=None) -> int:
    if rc == 0:
        return rc
    msg = (lib.s355_last_error() or b"").decode(errors="replace")
    if rc == _abi.S355_NOT_CONVERGED:
        raise NotConverged(rc, f"{what}: {msg}", info)
    raise S355Error(rc, f"{what}: {msg}")


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class CaseData:
    """A mesh in element-stream form plus its region table."""
    xyz: np.ndarray
    etype: np.ndarray
    tag: np.ndarray
    vptr: np.ndarray
    vidx: np.ndarray
    regions: list
    meta: dict = field(default_factory=dict)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_cell_elems(self) -> int:
        return int(np.count_nonzero(self.etype >= 4))

    def desc(self) -> MeshDesc:
        regs = (Region * len(self.regions))(*self.regions)
        self._keep = [regs]
        return MeshDesc(len(self.xyz) // 3, _dp(self.xyz), len(self.etype), _ip(self.etype), _ip(self.tag),
                        _ip(self.vptr), _ip(self.vidx), len(self.regions), regs)

    def region(self, rid: int) -> Region:
        for r in self.regions:
            if r.rid == rid:
                return r
        raise KeyError(rid)


def build_case(kind: int, ip: list, dp: list, meta: Optional[dict] = None) -> CaseData:
    lib = library()
    ipa = np.asarray(ip, dtype=np.int32)
    dpa = np.asarray(dp, dtype=np.float64)
    h = C.c_void_p()
    _check(lib, lib.s355_case_build(kind, _ip(ipa), _dp(dpa), C.byref(h)), "case_build")
    return _take_case(lib, h, meta)


def read_mesh(path, regions: list, binary: bool = False) -> CaseData:
    """A mesh file in the reference's formats (gmsh 2 ASCII, or the binary
    form of RootProcess::readBin) with the region table `regions` (what the
    reference reads from param.in)."""
    lib = library()
    h = C.c_void_p()
    _check(lib, lib.s355_case_read(str(path).encode(), int(binary), C.byref(h)), "case_read")
    case = _take_case(lib, h, {"kind": "file", "path": str(path)})
    case.regions = list(regions)
    return case


def read_backup_file(path, n_cell: int, turbulent: bool = False) -> dict:
    """ReadBackupFile on the host: the cell arrays of a res.sav."""
    lib = library()
    names = _abi.RESULT_FIELDS[:10 if turbulent else 7]
    out = np.empty(len(names) * n_cell)
    _check(lib, lib.s355_backup_read(str(path).encode(), int(n_cell), int(turbulent), _dp(out)), "backup_read")
    return {n: out[k * n_cell:(k + 1) * n_cell].copy() for k, n in enumerate(names)}


def write_mesh(case: CaseData, path, binary: bool = False) -> None:
    """Write a case's mesh in the reference's gmsh 2 ASCII or binary format."""
    lib = library()
    d = case.desc()
    _check(lib, lib.s355_case_write(C.byref(d), str(path).encode(), int(binary)), "case_write")


def _take_case(lib, h, meta) -> CaseData:
    try:
        sz = np.zeros(4, dtype=np.int32)
        _check(lib, lib.s355_case_sizes(h, _ip(sz)), "case_sizes")
        nv, ne, nvi, nr = (int(x) for x in sz)
        xyz = np.empty(3 * nv)
        et = np.empty(ne, dtype=np.int32)
        tg = np.empty(ne, dtype=np.int32)
        vp = np.empty(ne + 1, dtype=np.int32)
        vi = np.empty(nvi, dtype=np.int32)
        regs = (Region * nr)()
        _check(lib, lib.s355_case_copy(h, _dp(xyz), _ip(et), _ip(tg), _ip(vp), _ip(vi), regs), "case_copy")
    finally:
        lib.s355_case_free(h)
    return CaseData(xyz, et, tg, vp, vi, [regs[i] for i in range(nr)], dict(meta or {}))


def cavity(n: int, re: float = 100.0, lid_u: float = 1.0, jitter: float = 0.0, tets: bool = False,
           seed: int = 0) -> CaseData:
    """Lid-driven cavity in the unit cube, n^3 hexes (or 6 n^3 tets)."""
    return build_case(0, [n, int(tets), seed], [re, lid_u, jitter],
                      {"kind": "cavity", "n": n, "re": re, "jitter": jitter, "tets": tets})


def channel(nx: int, ny: int, nz: int, lx: float = 2.0, ly: float = 1.0, lz: float = 2.0, jitter: float = 0.0,
            split: float = 0.0, seed: int = 0, re: float = 1000.0, u_in: float = 1.0, turbulent: bool = False,
            intensity: float = 0.05) -> CaseData:
    """Channel: inlet x=0, outlet x=lx, walls y=0,ly, symmetry z=0,lz."""
    return build_case(1, [nx, ny, nz, int(turbulent), seed], [lx, ly, lz, jitter, split, re, u_in, intensity],
                      {"kind": "channel", "nx": nx, "ny": ny, "nz": nz, "jitter": jitter, "split": split, "re": re,
                       "turbulent": turbulent})


def step(n_per_h: int, upstream: float = 4.0, downstream: float = 30.0, span: float = 4.0, jitter: float = 0.10,
         tets: bool = True, turbulent: bool = True, seed: int = 0, re: float = 5000.0, u_in: float = 1.0,
         intensity: float = 0.05) -> CaseData:
    """Backward-facing step, expansion ratio 2 (step height 1)."""
    return build_case(2, [n_per_h, int(tets), int(turbulent), seed],
                      [upstream, downstream, span, jitter, re, u_in, intensity],
                      {"kind": "step", "n_per_h": n_per_h, "tets": tets, "turbulent": turbulent})


def default_params() -> Params:
    p = Params()
    library().s355_default_params(C.byref(p))
    return p


def set_all_ksp(p: Params, ksp_type: int = KSP_GMRES, restart: int = 30, pc: int = PC_JACOBI,
                rtol: Optional[float] = None, maxit: Optional[int] = None) -> Params:
    """Apply one KSP setting to every solve, as the reference's unprefixed
    KSPSetFromOptions does for command-line options (MPIStructure.cpp:261)."""
    for k in range(3):
        o = p.ksp[k]
        o.type, o.restart, o.pc = ksp_type, restart, pc
        if rtol is not None:
            o.rtol = rtol
        if maxit is not None:
            o.maxit = maxit
    return p


def _descs(cases: list):
    arr = (MeshDesc * len(cases))(*[c.desc() for c in cases])
    return arr


class HostMesh:
    """Host-only mesh set-up (no GPU needed); a list of cases gives the
    batch layout of BatchSolver."""

    def __init__(self, case):
        self.lib = library()
        self.h = C.c_void_p()
        if isinstance(case, (list, tuple)):
            self._d = _descs(list(case))
            _check(self.lib, self.lib.s355_mesh_build_batch(len(case), self._d, C.byref(self.h)), "mesh_build_batch")
        else:
            d = case.desc()
            _check(self.lib, self.lib.s355_mesh_build(C.byref(d), C.byref(self.h)), "mesh_build")
        s = Sizes()
        _check(self.lib, self.lib.s355_mesh_sizes(self.h, C.byref(s)), "mesh_sizes")
        self.sizes = s

    def ints(self, name: str) -> np.ndarray:
        out = np.empty(mesh_int_len(name, self.sizes), dtype=np.int32)
        _check(self.lib, self.lib.s355_mesh_get_int(self.h, MESH_INT_ID[name], _ip(out), out.size), name)
        return out

    def dbls(self, name: str) -> np.ndarray:
        out = np.empty(mesh_dbl_len(name, self.sizes))
        _check(self.lib, self.lib.s355_mesh_get_dbl(self.h, MESH_DBL_ID[name], _dp(out), out.size), name)
        return out

    def write_result(self, fmt: str, path, fields: Optional[dict] = None, turbulent: bool = False) -> None:
        """One result file of this mesh from host arrays (fmt: tecplot,
        tec_data, tec_geo, grid, backup; the grid formats need no fields)."""
        keep, r = [], None
        if fields is not None:
            r = _abi.Result()
            for n in _abi.RESULT_FIELDS:
                if fields.get(n) is not None:
                    a = np.ascontiguousarray(fields[n], dtype=np.float64)
                    if a.size != self.sizes.n_cell:
                        raise ValueError(f"{n}: expected {self.sizes.n_cell} values")
                    keep.append(a)
                    setattr(r, n.lower(), _dp(a))
            r.turbulent = int(turbulent)
        _check(self.lib, self.lib.s355_mesh_write_result(self.h, _abi.OUTPUT_FORMATS[fmt], str(path).encode(),
                                                         C.byref(r) if r is not None else None), "mesh_write_result")

    def close(self) -> None:
        if self.h:
            self.lib.s355_mesh_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """One case resident on the GPU (an s355_ctx)."""

    def __init__(self, case: CaseData, params: Params, init: bool = True):
        self.lib = library()
        self.case = case
        self.params = params
        self.ctx = C.c_void_p()
        self._create(case, params)
        self._finish(init)

    def _create(self, case, params) -> None:
        d = case.desc()
        _check(self.lib, self.lib.s355_ctx_create(C.byref(d), C.byref(params), C.byref(self.ctx)), "ctx_create")

    def _finish(self, init: bool) -> None:
        s = Sizes()
        _check(self.lib, self.lib.s355_ctx_sizes(self.ctx, C.byref(s)), "ctx_sizes")
        self.sizes = s
        if init:
            self.call("init_flow_field")

    # --- plumbing
    def call(self, name: str, *args) -> int:
        fn = getattr(self.lib, f"s355_{name}")
        return _check(self.lib, fn(self.ctx, *args), name)

    def get(self, name: str) -> np.ndarray:
        n = field_len(FIELD_KIND[name], self.sizes)
        out = np.empty(n)
        _check(self.lib, self.lib.s355_get_field(self.ctx, FIELD_ID[name], _dp(out), n), f"get {name}")
        return out

    def set(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _check(self.lib, self.lib.s355_set_field(self.ctx, FIELD_ID[name], _dp(a), a.size), f"set {name}")

    def ints(self, name: str) -> np.ndarray:
        out = np.empty(mesh_int_len(name, self.sizes), dtype=np.int32)
        _check(self.lib, self.lib.s355_get_mesh_int(self.ctx, MESH_INT_ID[name], _ip(out), out.size), name)
        return out

    def dbls(self, name: str) -> np.ndarray:
        out = np.empty(mesh_dbl_len(name, self.sizes))
        _check(self.lib, self.lib.s355_get_mesh_dbl(self.ctx, MESH_DBL_ID[name], _dp(out), out.size), name)
        return out

    def set_params(self, params: Params) -> None:
        _check(self.lib, self.lib.s355_ctx_set_params(self.ctx, C.byref(params)), "set_params")
        self.params = params

    def time_state(self) -> tuple:
        """(cur_time, time steps taken) of an unsteady run."""
        t, n = C.c_double(), C.c_int()
        _check(self.lib, self.lib.s355_time_state(self.ctx, C.byref(t), C.byref(n)), "time_state")
        return t.value, n.value

    # --- result and restart files
    def snapshot(self) -> None:
        """Start copying the cell fields to pinned host staging (no wait)."""
        self.call("output_snapshot")

    def write_output(self, fmt: str, path, async_: bool = False) -> None:
        """Write one file (tecplot, tec_data, tec_geo, grid, backup) from the
        latest snapshot; async_: on a writer thread (see wait_output)."""
        self.call("output_write", _abi.OUTPUT_FORMATS[fmt], str(path).encode(), int(async_))

    def wait_output(self) -> None:
        self.call("output_wait")

    def read_backup(self, path) -> None:
        """Restart from a res.sav (after init_flow_field, which the
        constructor runs)."""
        self.call("read_backup", str(path).encode())

    def monitor_records(self) -> list:
        """OutputMoniter's records: wall force per wall region, W^2 vol per
        fluid region (region id 0 left out), ascending region id."""
        out, n = np.empty(6 * 256), C.c_int()
        _check(self.lib, self.lib.s355_monitor_records(self.ctx, _dp(out), 256, C.byref(n)), "monitor_records")
        r = out[: 6 * n.value].reshape(-1, 6)
        return [{"rid": int(x[0]), "type1": int(x[1]), "force": x[2:5].copy(), "scalar": float(x[5])} for x in r]

    def output_schedule(self, noutput: int, directory, binary: bool = False) -> None:
        """NSSolve's output: every noutput steps res.sav and the Tecplot
        result under `directory` (0 turns it off)."""
        self.call("output_schedule", int(noutput), int(binary), str(directory).encode())

    # --- phases
    def gradient(self, field_name: str, bfield: str, out: str) -> None:
        self.call("gradient", FIELD_ID[field_name], FIELD_ID[bfield], FIELD_ID[out])

    def solve_velocity(self, rtol: float, maxit: int) -> list:
        info = (SolveInfo * 3)()
        rc = self.lib.s355_solve_velocity(self.ctx, rtol, maxit, info)
        _check(self.lib, rc, "solve_velocity", [i.as_dict() for i in info])
        return [i.as_dict() for i in info]

    def solve_pressure(self, rtol: float, maxit: int) -> dict:
        info = SolveInfo()
        rc = self.lib.s355_solve_pressure(self.ctx, rtol, maxit, C.byref(info))
        _check(self.lib, rc, "solve_pressure", info.as_dict())
        return info.as_dict()

    def solve_scalar(self, isca: int, rtol: float, maxit: int) -> dict:
        """solveScarlar_GMRES on the assembled scalar system into T (1), k (2)
        or epsilon (3)."""
        info = SolveInfo()
        rc = self.lib.s355_solve_scalar(self.ctx, isca, rtol, maxit, C.byref(info))
        _check(self.lib, rc, "solve_scalar", info.as_dict())
        return info.as_dict()

    def correct_fields(self) -> float:
        r = C.c_double()
        self.call("correct_fields", C.byref(r))
        return r.value

    def residuals(self, reset: bool = False) -> np.ndarray:
        out = np.empty(10)
        _check(self.lib, self.lib.s355_residuals(self.ctx, _dp(out), int(reset)), "residuals")
        return out

    def iterate(self, n_steps: int) -> tuple:
        """Run the NSSolve loop; returns (history[steps, 11], steps_done)."""
        hist = np.zeros(11 * max(n_steps, 1))
        done = C.c_int(0)
        rc = self.lib.s355_simple_iterate(self.ctx, n_steps, _dp(hist), C.byref(done))
        h = hist.reshape(-1, 11)[: done.value]
        _check(self.lib, rc, f"simple_iterate (after {done.value} steps)")
        return h, done.value

    def update_energy(self) -> dict:
        """UpdateEnergy (navier_scalar.cpp:204-326): one energy solve into T."""
        info = SolveInfo()
        _check(self.lib, self.lib.s355_update_energy(self.ctx, C.byref(info)), "update_energy")
        return info.as_dict()

    def last_solve_info(self) -> dict:
        """Solve records of the last phases: U, V, W, pressure, k, epsilon."""
        info = (SolveInfo * 7)()
        _check(self.lib, self.lib.s355_last_solve_info(self.ctx, info), "last_solve_info")
        return {n: info[k].as_dict() for k, n in enumerate(("U", "V", "W", "P", "TE", "ED", "T"))}

    def stage_times(self, reset: bool = False) -> np.ndarray:
        out = np.zeros(6)
        _check(self.lib, self.lib.s355_log_stage_times(self.ctx, _dp(out), int(reset)), "log_stage_times")
        return out

    def spmv_time(self, which: int, precision: int, repeats: int) -> float:
        ms = C.c_double()
        _check(self.lib, self.lib.s355_spmv_time(self.ctx, which, precision, repeats, C.byref(ms)), "spmv_time")
        return ms.value

    def kernel_shapes(self) -> dict:
        """SpMV lanes per row (0 = SELL-64) / non-temporal loads / SELL
        sorting window and Gram-Schmidt rows per thread, FP64 and FP32."""
        lanes, nt, rows, sigma, tuned, slices, order = ((C.c_int * 2)() for _ in range(7))
        _check(self.lib, self.lib.s355_kernel_shapes(self.ctx, lanes, nt, rows, sigma, tuned, slices, order),
               "kernel_shapes")
        return {p: {"lanes": lanes[k], "nt": nt[k], "rows": rows[k], "sigma": sigma[k], "tuned": bool(tuned[k]),
                    "slices": slices[k], "order": order[k]}
                for k, p in enumerate(("fp64", "fp32"))}

    def tune_report(self, precision: int) -> dict:
        """Autotuner timings (ms): SpMV per (lanes, nt), Gram-Schmidt step per rows."""
        sp, gs, it = (C.c_double * 16)(), (C.c_double * 5)(), C.c_int()
        _check(self.lib, self.lib.s355_tune_report(self.ctx, precision, sp, gs, C.byref(it)), "tune_report")
        names = ["L1", "L2", "L4", "L8", "L16", "SELL64w64", "SELL64w512", "SELL64w4096"]
        spmv = {f"{names[v // 2]}{'_nt' if v % 2 else ''}": sp[v] for v in range(16) if sp[v] >= 0}
        return {"spmv_ms": spmv, "gs_ms": {f"R{1 << r}": gs[r] for r in range(5) if gs[r] >= 0}, "gs_it": it.value}

    def spmv_profile(self, precision: int, mode: int) -> tuple:
        """In-solve SpMV timing: mode 1 start, 0 read, -1 stop; (mean ms, launches)."""
        ms, n = C.c_double(), C.c_int()
        _check(self.lib, self.lib.s355_spmv_profile(self.ctx, precision, mode, C.byref(ms), C.byref(n)), "spmv_profile")
        return ms.value, n.value

    def set_kernel_shapes(self, precision: int, lanes: int, nt: int = 0, rows: int = 0, sigma: int = 0) -> None:
        """Force the GMRES kernel shapes of one precision (lanes 0 = SELL-64
        sorted in windows of sigma rows: 64, 512, 4096; 0 = 4096)."""
        _check(self.lib, self.lib.s355_set_kernel_shapes(self.ctx, precision, lanes, nt, rows, sigma),
               "set_kernel_shapes")

    def copy_time(self, nbytes: int, repeats: int) -> float:
        """Mean device ms of a plain copy of nbytes (multiple of 16)."""
        ms = C.c_double()
        _check(self.lib, self.lib.s355_copy_time(self.ctx, nbytes, repeats, C.byref(ms)), "copy_time")
        return ms.value

    def close(self) -> None:
        if self.ctx:
            self.lib.s355_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_ARRAYS = {"cell": 0, "bnd": 1, "iface": 2, "nnz": 3, "vert": 4, "elem": 5}
SOLVE_KINDS = ("U", "V", "W", "P", "TE", "ED", "T")


class BatchSolver(Solver):
    """Independent cases resident on the GPU as one batch: every SIMPLE phase
    runs all of them with one launch per kernel, each case keeping its own
    Krylov solves, residuals and outlet correction.  get/set act on the
    concatenated arrays; case(k, name) slices one case's part."""

    def __init__(self, cases: list, params: Params, init: bool = True):
        self.lib = library()
        self.cases = list(cases)
        self.case = None
        self.params = params
        self.ctx = C.c_void_p()
        self._d = _descs(self.cases)
        _check(self.lib, self.lib.s355_batch_create(len(self.cases), self._d, C.byref(params), C.byref(self.ctx)),
               "batch_create")
        self._finish(init)
        self.n_case = len(self.cases)
        self.layout = {}
        for name, w in BATCH_ARRAYS.items():
            a = np.empty(self.n_case + 1, dtype=np.int32)
            _check(self.lib, self.lib.s355_batch_layout(self.ctx, w, _ip(a), a.size), "batch_layout")
            self.layout[name] = a

    def case_slice(self, k: int, kind: str) -> slice:
        if kind in ("c", "c3"):
            a, b = self.layout["cell"][k], self.layout["cell"][k + 1]
        elif kind in ("b", "b3"):
            a, b = self.layout["bnd"][k], self.layout["bnd"][k + 1]
        elif kind == "z":
            a, b = self.layout["nnz"][k], self.layout["nnz"][k + 1]
        else:
            raise ValueError(f"no per-case slice for field kind {kind}")
        w = 3 if kind in ("c3", "b3") else 1
        return slice(w * int(a), w * int(b))

    def case_field(self, k: int, name: str) -> np.ndarray:
        return self.get(name)[self.case_slice(k, FIELD_KIND[name])]

    def case_residuals(self, reset: bool = False) -> np.ndarray:
        out = np.zeros(10 * self.n_case)
        _check(self.lib, self.lib.s355_batch_residuals(self.ctx, _dp(out), int(reset)), "batch_residuals")
        return out.reshape(self.n_case, 10)

    def case_solve_info(self) -> dict:
        res = {}
        for kind, name in enumerate(SOLVE_KINDS):
            info = (SolveInfo * self.n_case)()
            _check(self.lib, self.lib.s355_batch_solve_info(self.ctx, kind, info, self.n_case), "batch_solve_info")
            res[name] = [i.as_dict() for i in info]
        return res
