"""ctypes binding of libflgpu.so (the C ABI declared in include/flgpu.h).

The library is the product path: there is no CPU fallback.  Loading fails
loudly when libflgpu.so has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflgpu.so")

FLGPU_OK = 0
FLGPU_FP64, FLGPU_FP32, FLGPU_MIXED = 0, 1, 1
FLGPU_ASM_AUTO, FLGPU_ASM_COLOUR, FLGPU_ASM_BLOCKED = 0, 1, 2
FLGPU_KSP_CG, FLGPU_KSP_GMRES, FLGPU_KSP_FCG, FLGPU_KSP_PIPECG = 0, 1, 2, 3
FLGPU_PC_NONE, FLGPU_PC_JACOBI, FLGPU_PC_LEARNED, FLGPU_PC_SOR, FLGPU_PC_FSAI = 0, 1, 2, 3, 4
FLGPU_NORM_PRECONDITIONED, FLGPU_NORM_UNPRECONDITIONED = 0, 1
FIELD_T, FIELD_VELOCITY, FIELD_DIFFUSIVITY, FIELD_SOURCE, FIELD_ABSORPTION = range(5)
FIELD_DENSITY, FIELD_OLD_DENSITY, FIELD_PRESSURE, FIELD_VOLUME_FRACTION = range(5, 9)
FLGPU_EQ = {"advection_diffusion": 0, "internal_energy": 1, "k_epsilon": 2}
BC_NONE, BC_NEUMANN, BC_WEAKDIRICHLET, BC_INTERNAL, BC_ROBIN = range(5)
FLGPU_STAB = {"none": 0, "streamline_upwind": 1, "supg": 2}
FLGPU_NU_BAR = {"optimal": 1, "doubly_asymptotic": 2, "critical_rule": 3, "unity": 4}

# Every symbol include/flgpu.h declares (checked by the CPU tests).
EXPORTS = (
    "flgpu_init", "flgpu_finalize", "flgpu_last_error", "flgpu_mesh_bind",
    "flgpu_field_mesh_bind", "flgpu_sparsity_bind", "flgpu_colouring_bind",
    "flgpu_make_sparsity", "flgpu_colour_mesh", "flgpu_sparsity_build", "flgpu_sparsity_get",
    "flgpu_colouring_build", "flgpu_colouring_get", "flgpu_field_set", "flgpu_mesh_motion_set",
    "flgpu_advdif_assemble_resident", "flgpu_advdif_assemble", "flgpu_rhs_add_field", "flgpu_surface_bind",
    "flgpu_advdif_surface", "flgpu_dirichlet_apply",
    "flgpu_matrix_get", "flgpu_matrix_set", "flgpu_solve", "flgpu_solve_noalloc", "flgpu_pressure_bind", "flgpu_ct_assemble", "flgpu_cmc_assemble", "flgpu_time_pressure", "flgpu_petsc_solve", "flgpu_pc_apply",
    "flgpu_solution_get", "flgpu_csr_mult", "flgpu_matrixdump_read", "flgpu_matrixdump_write",
    "flgpu_learned_param_count",
    "flgpu_learned_set_params", "flgpu_learned_get_params", "flgpu_learned_init_params",
    "flgpu_learned_train_step", "flgpu_learned_loss_grad", "flgpu_learned_factor",
    "flgpu_time_assemble", "flgpu_assembly_plan_info", "flgpu_time_spmv", "flgpu_time_spmv_cold", "flgpu_solver_prepare", "flgpu_sizes",
)


class AdvdifParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("theta", C.c_double), ("beta", C.c_double),
        ("have_mass", C.c_int32), ("lump_mass", C.c_int32),
        ("have_advection", C.c_int32), ("integrate_by_parts", C.c_int32),
        ("have_diffusivity", C.c_int32), ("isotropic_diffusivity", C.c_int32),
        ("have_source", C.c_int32), ("have_absorption", C.c_int32),
        ("precision", C.c_int32), ("scheme", C.c_int32),
        ("stabilisation", C.c_int32), ("nu_bar_scheme", C.c_int32), ("nu_bar_scale", C.c_double),
        ("density_theta", C.c_double), ("equation_type", C.c_int32), ("compressible", C.c_int32),
        ("multiphase", C.c_int32), ("move_mesh", C.c_int32),
    ]


class SolverParams(C.Structure):
    _fields_ = [
        ("ksp", C.c_int32), ("pc", C.c_int32),
        ("rtol", C.c_double), ("atol", C.c_double),
        ("max_its", C.c_int32), ("restart", C.c_int32),
        ("precision", C.c_int32), ("start_from_zero", C.c_int32),
        ("check_every", C.c_int32), ("stall_factor", C.c_double),
        ("stall_window", C.c_int32), ("norm_type", C.c_int32),
    ]


class SolveInfo(C.Structure):
    _fields_ = [
        ("its", C.c_int32), ("reason", C.c_int32),
        ("true_relres", C.c_double), ("precond_relres", C.c_double),
        ("fallback", C.c_int32), ("refinements", C.c_int32),
        ("seconds", C.c_double),
    ]


class FlgpuError(RuntimeError):
    def __init__(self, code: int, fn: str, msg: str):
        super().__init__(f"{fn} failed (code {code}): {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    """Load libflgpu.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    P, I32, I64, D, F = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_float
    sig = {
        "flgpu_init": [C.POINTER(C.c_void_p), I32],
        "flgpu_finalize": [P],
        "flgpu_mesh_bind": [P, I64, I64, I32, I32, I32, I32, P, P],
        "flgpu_field_mesh_bind": [P, I32, I32, I32, P, P, P, P, P],
        "flgpu_sparsity_bind": [P, I32, I32, P, P, P],
        "flgpu_colouring_bind": [P, I32, P, P],
        "flgpu_make_sparsity": [I32, I32, I32, P, P, P, P, C.POINTER(C.c_int64)],
        "flgpu_colour_mesh": [I32, I32, I32, P, P, C.POINTER(C.c_int32)],
        "flgpu_field_set": [P, I32, I32, P],
        "flgpu_mesh_motion_set": [P, P, P, P],
        "flgpu_advdif_assemble_resident": [P, C.POINTER(AdvdifParams)],
        "flgpu_advdif_assemble": [P, C.POINTER(AdvdifParams), P, P, P, P, P, P, P],
        "flgpu_rhs_add_field": [P, I32],
        "flgpu_surface_bind": [P, I32, I32, I32, P, P, P, P, P, P],
        "flgpu_advdif_surface": [P, C.POINTER(AdvdifParams), P, P, P],
        "flgpu_dirichlet_apply": [P, I32, P, P, D],
        "flgpu_matrix_get": [P, P, P, P],
        "flgpu_matrix_set": [P, P, P, P],
        "flgpu_solve": [P, C.POINTER(SolverParams), P, C.POINTER(SolveInfo)],
        "flgpu_solve_noalloc": [P, C.POINTER(SolverParams), P, C.POINTER(SolveInfo)],
        "flgpu_pressure_bind": [P, I32, I32, P, P, P, I32, P, P],
        "flgpu_ct_assemble": [P, I32, I32, P, P, P],
        "flgpu_cmc_assemble": [P, P, I32, P, P, P],
        "flgpu_time_pressure": [P, I32, C.POINTER(C.c_double), C.POINTER(C.c_double)],
        "flgpu_pc_apply": [P, C.POINTER(SolverParams), P, P],
        "flgpu_petsc_solve": [P, C.POINTER(SolverParams), P, P, P, P,
                              C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)],
        "flgpu_solution_get": [P, P],
        "flgpu_csr_mult": [P, P, P],
        "flgpu_sparsity_build": [P, C.POINTER(C.c_int64)],
        "flgpu_sparsity_get": [P, P, P, P],
        "flgpu_colouring_build": [P, C.POINTER(C.c_int32)],
        "flgpu_colouring_get": [P, P, P],
        "flgpu_matrixdump_read": [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), P, P, P, P, P,
                                  C.POINTER(C.c_int32)],
        "flgpu_matrixdump_write": [C.c_char_p, I32, P, P, P, P, P],
        "flgpu_learned_param_count": [],
        "flgpu_learned_set_params": [P, P],
        "flgpu_learned_get_params": [P, P],
        "flgpu_learned_init_params": [P, C.c_uint64],
        "flgpu_learned_train_step": [P, I32, C.c_uint64, F, C.POINTER(C.c_float)],
        "flgpu_learned_loss_grad": [P, I32, C.c_uint64, C.POINTER(C.c_float), P],
        "flgpu_learned_factor": [P, P, C.POINTER(C.c_int64)],
        "flgpu_time_assemble": [P, C.POINTER(AdvdifParams), I32, C.POINTER(C.c_double)],
        "flgpu_assembly_plan_info": [P, P, I32, C.POINTER(C.c_double)],
        "flgpu_time_spmv": [P, I32, I32, C.POINTER(C.c_double)],
        "flgpu_time_spmv_cold": [P, I32, I32, I32, C.POINTER(C.c_double)],
        "flgpu_solver_prepare": [P, I32],
        "flgpu_sizes": [P, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int32)],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.flgpu_last_error.argtypes = []
    lib.flgpu_last_error.restype = C.c_char_p
    _lib = lib
    return lib


def ptr(a: np.ndarray | None):
    """Raw data pointer of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    p = C.c_void_p(a.ctypes.data)
    p._keep = a  # the array lives as long as the argument object
    return p


def check(rc: int, fn: str) -> None:
    if rc != FLGPU_OK:
        msg = load().flgpu_last_error().decode(errors="replace")
        raise FlgpuError(rc, fn, msg)


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)
