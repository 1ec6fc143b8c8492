"""ctypes mirror of include/simple355.h (structs, enums, prototypes).

Shared by the HIP library binding and the CPU oracle binding, which exports
the same contract with an ``orc_`` prefix."""
from __future__ import annotations

import ctypes as C

S355_OK = 0
S355_NOT_CONVERGED = 1
S355_E_ARG = -1
S355_E_HIP = -2
S355_E_MESH = -3
S355_E_STATE = -4
S355_E_NOMEM = -5
S355_E_IO = -6
MAX_CELL_FACES = 12

KSP_GMRES, KSP_CG = 0, 1
PC_NONE, PC_JACOBI, PC_ILU0 = 0, 1, 2
FP64, MIXED = 0, 1
KSP_VELOCITY, KSP_PRESSURE, KSP_SCALAR = 0, 1, 2
# s355_run_phase
(PHASE_MOMENTUM, PHASE_VELOCITY_RESIDUAL, PHASE_PRESSURE, PHASE_CORRECT, PHASE_KEPS_K, PHASE_KEPS_EPS,
 PHASE_KEPS_POST) = range(7)


class Region(C.Structure):
    _fields_ = [
        ("rid", C.c_int),
        ("type1", C.c_int),
        ("type2", C.c_int),
        ("fixed_value", C.c_double),
        ("init", C.c_double * 10),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_vert", C.c_int),
        ("xyz", C.POINTER(C.c_double)),
        ("n_elem", C.c_int),
        ("elem_type", C.POINTER(C.c_int)),
        ("elem_tag", C.POINTER(C.c_int)),
        ("elem_vptr", C.POINTER(C.c_int)),
        ("elem_vidx", C.POINTER(C.c_int)),
        ("n_region", C.c_int),
        ("regions", C.POINTER(Region)),
    ]


class KspOpts(C.Structure):
    _fields_ = [
        ("type", C.c_int),
        ("restart", C.c_int),
        ("pc", C.c_int),
        ("maxit", C.c_int),
        ("rtol", C.c_double),
        ("abstol", C.c_double),
        ("dtol", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("turb_model", C.c_int),
        ("density_model", C.c_int),
        ("steady", C.c_int),
        ("cell_pressure_ref", C.c_int),
        ("urf", C.c_double * 8),
        ("gravity", C.c_double * 3),
        ("residual_steady", C.c_double),
        ("ksp", KspOpts * 3),
        ("precision", C.c_int),
        ("ir_max", C.c_int),
        ("ir_inner_rtol", C.c_double),
        ("autotune", C.c_int),
        ("graphs", C.c_int),
        ("solve_energy", C.c_int),
        ("inner_order", C.c_int),
        ("cp", C.c_double),
        ("prl", C.c_double),
        ("prte", C.c_double),
        ("limiter", C.c_int),
        ("time_scheme", C.c_int),
        ("dt", C.c_double),
        ("total_time", C.c_double),
        ("max_outer_step", C.c_int),
        ("grad_scheme", C.c_int),
        ("assembly_fp32", C.c_int),
        ("assembly_pad", C.c_int),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("reason", C.c_int),
        ("rnorm", C.c_double),
        ("rnorm0", C.c_double),
        ("refinements", C.c_int),
        ("pad", C.c_int),
    ]

    def as_dict(self) -> dict:
        return {"iters": self.iters, "reason": self.reason, "rnorm": self.rnorm, "rnorm0": self.rnorm0,
                "refinements": self.refinements}


class Sizes(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("n_vert", "n_cell", "n_face", "n_bnd", "n_interior", "nnz", "n_elem", "n_warped")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


# result file formats (enum s355_output_format)
OUTPUT_FORMATS = {"tecplot": 0, "tec_data": 1, "tec_geo": 2, "grid": 3, "backup": 4, "vtk": 5}
# cell arrays of a result, in s355_result order
RESULT_FIELDS = ("R", "U", "V", "W", "P", "T", "VISLAM", "VISTUR", "TE", "ED")


class Result(C.Structure):
    _fields_ = [(n.lower(), C.POINTER(C.c_double)) for n in RESULT_FIELDS] + [("turbulent", C.c_int),
                                                                               ("pad", C.c_int)]


# field ids (enum s355_field)
FIELDS = [
    "R", "U", "V", "W", "P", "T", "TE", "ED", "DELTAP", "VISLAM", "VISTUR", "APR",
    "DUDX", "DVDX", "DWDX", "DPDX", "DPHIDX",
    "BRO", "BU", "BV", "BW", "BPRE", "BTE", "BED",
    "SHEAR", "RUFACE", "AU", "AP", "AS",
    "BU_RHS", "BV_RHS", "BW_RHS", "BP_RHS", "BS_RHS", "BTEM",
    "UNP", "VNP", "WNP", "TNP", "TEP", "EDP", "UNP2", "VNP2", "WNP2", "TNP2", "TEP2", "EDP2",
    "KE_PROD", "KE_TESRC", "KE_EDSRC", "KE_APTE", "KE_APED", "KE_VISTE", "KE_VISED",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
# which size each field has: c = cells, c3 = 3*cells, b = boundary, b3, f = faces, z = nnz
FIELD_KIND = {
    **{n: "c" for n in ("R", "U", "V", "W", "P", "T", "TE", "ED", "DELTAP", "VISLAM", "VISTUR", "APR",
                        "BU_RHS", "BV_RHS", "BW_RHS", "BP_RHS", "BS_RHS")},
    **{n: "c3" for n in ("DUDX", "DVDX", "DWDX", "DPDX", "DPHIDX")},
    **{n: "b" for n in ("BRO", "BU", "BV", "BW", "BPRE", "BTE", "BED", "BTEM")},
    "SHEAR": "b3", "RUFACE": "f", "AU": "z", "AP": "z", "AS": "z",
    **{n: "c" for n in ("UNP", "VNP", "WNP", "TNP", "TEP", "EDP", "UNP2", "VNP2", "WNP2", "TNP2", "TEP2", "EDP2")},
    **{n: "c" for n in ("KE_PROD", "KE_TESRC", "KE_EDSRC", "KE_APTE", "KE_APED", "KE_VISTE", "KE_VISED")},
}

MESH_INT = ["ELEM_PERM", "FACE_C1", "FACE_C2", "FACE_BND", "FACE_VERT", "CELL_NFACE", "CELL_FACE", "CELL_CELL",
            "BND_FACE", "BND_RID", "CELL_RID", "CSR_ROWPTR", "CSR_COL", "CSR_DIAG", "CELL_VERT"]
MESH_INT_ID = {n: i for i, n in enumerate(MESH_INT)}
MESH_DBL = ["CELL_VOL", "CELL_X", "FACE_X", "FACE_N", "FACE_AREA", "FACE_LAMBDA", "FACE_RLENCOS", "FACE_XPAC",
            "FACE_XNAC", "BND_DIST", "VERT"]
MESH_DBL_ID = {n: i for i, n in enumerate(MESH_DBL)}


def field_len(kind: str, sizes: Sizes) -> int:
    return {
        "c": sizes.n_cell, "c3": 3 * sizes.n_cell, "b": sizes.n_bnd, "b3": 3 * sizes.n_bnd,
        "f": sizes.n_face, "z": sizes.nnz,
    }[kind]


def mesh_int_len(name: str, s: Sizes) -> int:
    K = MAX_CELL_FACES
    return {
        "ELEM_PERM": s.n_elem, "FACE_C1": s.n_face, "FACE_C2": s.n_face, "FACE_BND": s.n_face,
        "FACE_VERT": 4 * s.n_face, "CELL_NFACE": s.n_cell, "CELL_FACE": K * s.n_cell, "CELL_CELL": K * s.n_cell,
        "BND_FACE": s.n_bnd, "BND_RID": s.n_bnd, "CELL_RID": s.n_cell, "CSR_ROWPTR": s.n_cell + 1,
        "CSR_COL": s.nnz, "CSR_DIAG": s.n_cell, "CELL_VERT": 8 * s.n_cell,
    }[name]


def mesh_dbl_len(name: str, s: Sizes) -> int:
    return {
        "CELL_VOL": s.n_cell, "CELL_X": 3 * s.n_cell, "FACE_X": 3 * s.n_face, "FACE_N": 3 * s.n_face,
        "FACE_AREA": s.n_face, "FACE_LAMBDA": s.n_face, "FACE_RLENCOS": s.n_face, "FACE_XPAC": 3 * s.n_face,
        "FACE_XNAC": 3 * s.n_face, "BND_DIST": s.n_bnd, "VERT": 3 * s.n_vert,
    }[name]


def _vp():
    return C.c_void_p


def declare(lib: C.CDLL, prefix: str, ctx_type=C.c_void_p) -> None:
    """Declare the phase/field entry points shared by both libraries."""
    P = prefix
    I, D, V = C.c_int, C.c_double, C.c_void_p
    pD = C.POINTER(C.c_double)
    pI = C.POINTER(C.c_int)
    pInfo = C.POINTER(SolveInfo)
    spec = {
        "last_error": (C.c_char_p, []),
        "get_mesh_int": (I, [V, I, pI, C.c_size_t]),
        "get_mesh_dbl": (I, [V, I, pD, C.c_size_t]),
        "set_field": (I, [V, I, pD, C.c_size_t]),
        "get_field": (I, [V, I, pD, C.c_size_t]),
        "init_flow_field": (I, [V]),
        "bc_velocity": (I, [V]),
        "bc_pressure": (I, [V]),
        "bc_deltap": (I, [V]),
        "gradient": (I, [V, I, I, I]),
        "build_velocity_matrix": (I, [V]),
        "solve_velocity": (I, [V, D, I, pInfo]),
        "cal_ru_face": (I, [V]),
        "build_pressure_matrix": (I, [V]),
        "solve_pressure": (I, [V, D, I, pInfo]),
        "correct_fields": (I, [V, pD]),
        "correct_ru_face": (I, [V]),
        "simple_iterate": (I, [V, I, pD, pI]),
        "save_transient_old_data": (I, [V]),
        "time_state": (I, [V, pD, pI]),
        "kepsilon_sources": (I, [V]),
        "bc_kepsilon": (I, [V]),
        "build_scalar_matrix": (I, [V, I]),
        "eps_wall_rows": (I, [V]),
        "solve_scalar": (I, [V, I, D, I, pInfo]),
        "kepsilon_post_k": (I, [V]),
        "kepsilon_post_eps": (I, [V]),
    }
    for name, (res, args) in spec.items():
        fn = getattr(lib, f"{P}_{name}")
        fn.restype = res
        fn.argtypes = args
