"""ORACLE — test infrastructure only.

ctypes binding of the FP64 CPU restatement (oracle/_build/liboracle_s355.so).
Imported by the tests, the smoke check and the benchmark's CPU-baseline leg,
never by the product package."""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

from sysu_cfd_parallel_amd import _abi
from sysu_cfd_parallel_amd._abi import (FIELD_ID, FIELD_KIND, MESH_DBL_ID, MESH_INT_ID, MeshDesc, Params, Sizes,
                                        SolveInfo, field_len, mesh_dbl_len, mesh_int_len)

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle_s355.so"
_lib: Optional[C.CDLL] = None


def build() -> Path:
    r = subprocess.run(["make", "-C", str(HERE), "-j4"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return LIB


def library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        _abi.declare(lib, "orc")
        V, I = C.c_void_p, C.c_int
        lib.orc_create.restype = V
        lib.orc_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params)]
        lib.orc_destroy.restype = None
        lib.orc_destroy.argtypes = [V]
        lib.orc_sizes.restype = I
        lib.orc_sizes.argtypes = [V, C.POINTER(Sizes)]
        lib.orc_set_threads.restype = None
        lib.orc_set_threads.argtypes = [I]
        lib.orc_max_threads.restype = I
        lib.orc_max_threads.argtypes = []
        lib.orc_calculate_velocity.restype = I
        lib.orc_calculate_velocity.argtypes = [V, C.POINTER(C.c_double), C.POINTER(SolveInfo)]
        lib.orc_calculate_pressure.restype = I
        lib.orc_calculate_pressure.argtypes = [V, C.POINTER(C.c_double), C.POINTER(SolveInfo)]
        lib.orc_update_tur_kepsilon.restype = I
        lib.orc_update_tur_kepsilon.argtypes = [V, C.POINTER(SolveInfo)]
        lib.orc_update_energy.restype = I
        lib.orc_update_energy.argtypes = [V, C.POINTER(SolveInfo)]
        lib.orc_pc_apply.restype = I
        lib.orc_pc_apply.argtypes = [I, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), I, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.orc_kepsilon_time_terms.restype = I
        lib.orc_kepsilon_time_terms.argtypes = [V]
        lib.orc_bc_temperature.restype = I
        lib.orc_bc_temperature.argtypes = [V]
        _lib = lib
    return _lib


def pc_apply(rowptr, col, val, pc: int, r) -> np.ndarray:
    """z = M^-1 r for a CSR matrix (rows with sorted columns and a diagonal
    entry) with the oracle's Jacobi or natural-order ILU(0) preconditioner
    (PETSc's PCJACOBI / PCILU with zero fill)."""
    lib = library()
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    n = rp.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(rp))
    dg = np.flatnonzero(cl == rows).astype(np.int32)
    if dg.size != n:
        raise ValueError("every row needs its diagonal entry")
    v = np.ascontiguousarray(val, dtype=np.float64)
    rr = np.ascontiguousarray(r, dtype=np.float64)
    z = np.empty(n)
    rc = lib.orc_pc_apply(n, _ip(rp), _ip(cl), _ip(dg), _dp(v), int(pc), _dp(rr), _dp(z))
    if rc != 0:
        raise OracleError(rc, lib.orc_last_error().decode())
    return z


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """FP64 CPU restatement of one case."""

    def __init__(self, case, params: Params, init: bool = True, threads: Optional[int] = None):
        self.lib = library()
        if threads:
            self.lib.orc_set_threads(int(threads))
        d = case.desc()
        self.h = self.lib.orc_create(C.byref(d), C.byref(params))
        if not self.h:
            raise OracleError(-1, self.lib.orc_last_error().decode())
        s = Sizes()
        self.lib.orc_sizes(self.h, C.byref(s))
        self.sizes = s
        self.params = params
        if init:
            self.call("init_flow_field")

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise OracleError(rc, f"{what}: {self.lib.orc_last_error().decode()}")
        return rc

    def call(self, name: str, *args) -> int:
        return self._check(getattr(self.lib, f"orc_{name}")(self.h, *args), name)

    def get(self, name: str) -> np.ndarray:
        n = field_len(FIELD_KIND[name], self.sizes)
        out = np.empty(n)
        self._check(self.lib.orc_get_field(self.h, FIELD_ID[name], _dp(out), n), f"get {name}")
        return out

    def set(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self.lib.orc_set_field(self.h, FIELD_ID[name], _dp(a), a.size), f"set {name}")

    def ints(self, name: str) -> np.ndarray:
        out = np.empty(mesh_int_len(name, self.sizes), dtype=np.int32)
        self._check(self.lib.orc_get_mesh_int(self.h, MESH_INT_ID[name], _ip(out), out.size), name)
        return out

    def dbls(self, name: str) -> np.ndarray:
        out = np.empty(mesh_dbl_len(name, self.sizes))
        self._check(self.lib.orc_get_mesh_dbl(self.h, MESH_DBL_ID[name], _dp(out), out.size), name)
        return out

    def gradient(self, field_name: str, bfield: str, out: str) -> None:
        self.call("gradient", FIELD_ID[field_name], FIELD_ID[bfield], FIELD_ID[out])

    def solve_velocity(self, rtol: float, maxit: int) -> list:
        info = (SolveInfo * 3)()
        self._check(self.lib.orc_solve_velocity(self.h, rtol, maxit, info), "solve_velocity")
        return [i.as_dict() for i in info]

    def solve_pressure(self, rtol: float, maxit: int) -> dict:
        info = SolveInfo()
        self._check(self.lib.orc_solve_pressure(self.h, rtol, maxit, C.byref(info)), "solve_pressure")
        return info.as_dict()

    def solve_scalar(self, isca: int, rtol: float, maxit: int) -> dict:
        info = SolveInfo()
        self._check(self.lib.orc_solve_scalar(self.h, isca, rtol, maxit, C.byref(info)), "solve_scalar")
        return info.as_dict()

    def correct_fields(self) -> float:
        r = C.c_double()
        self.call("correct_fields", C.byref(r))
        return r.value

    def calculate_velocity(self):
        res = np.zeros(3)
        info = (SolveInfo * 3)()
        rc = self._check(self.lib.orc_calculate_velocity(self.h, _dp(res), info), "calculate_velocity")
        return rc, res, [i.as_dict() for i in info]

    def calculate_pressure(self):
        res = C.c_double()
        info = SolveInfo()
        rc = self._check(self.lib.orc_calculate_pressure(self.h, C.byref(res), C.byref(info)), "calculate_pressure")
        return rc, res.value, info.as_dict()

    def update_tur_kepsilon(self):
        info = (SolveInfo * 2)()
        rc = self._check(self.lib.orc_update_tur_kepsilon(self.h, info), "update_tur_kepsilon")
        return rc, [i.as_dict() for i in info]

    def update_energy(self):
        info = SolveInfo()
        rc = self._check(self.lib.orc_update_energy(self.h, C.byref(info)), "update_energy")
        return rc, info.as_dict()

    def time_state(self) -> tuple:
        t, n = C.c_double(), C.c_int()
        self._check(self.lib.orc_time_state(self.h, C.byref(t), C.byref(n)), "time_state")
        return t.value, n.value

    def iterate(self, n_steps: int):
        hist = np.zeros(11 * max(n_steps, 1))
        done = C.c_int(0)
        rc = self._check(self.lib.orc_simple_iterate(self.h, n_steps, _dp(hist), C.byref(done)), "simple_iterate")
        return hist.reshape(-1, 11)[: done.value], done.value, rc

    def close(self) -> None:
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
