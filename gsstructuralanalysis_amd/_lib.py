"""ctypes binding of libksh.so (declarations mirror include/ksh.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libksh.so")

SUCCESS, NOT_CONVERGED, ASSEMBLY_ERROR, SOLVER_ERROR, NOT_STARTED, OTHER_ERROR = range(6)
STATUS_NAMES = ["Success", "NotConverged", "AssemblyError", "SolverError", "NotStarted", "OtherError"]
RES_STATIC, RES_ALM = 0, 1
FP64, FP32_IR, FP32 = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_i64p = C.POINTER(C.c_int64)
_fp = C.POINTER(C.c_float)


class Patch(C.Structure):
    _fields_ = [("p", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("ctrl", _dp)]


class Material(C.Structure):
    _fields_ = [("t", C.c_double), ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double)]


class BC(C.Structure):
    _fields_ = [("fixed", _ip)]


class Loads(C.Structure):
    _fields_ = [("npoint", C.c_int), ("point_xi", _dp), ("point_F", _dp), ("fsurf", C.c_double * 3)]


class AlmOptions(C.Structure):
    _fields_ = [("method", C.c_int), ("precision", C.c_int), ("nsteps", C.c_int), ("length", C.c_double),
                ("max_iter", C.c_int), ("tol", C.c_double), ("tol_u", C.c_double), ("tol_f", C.c_double),
                ("scaling", C.c_double), ("angle_method", C.c_int), ("bifurcation_method", C.c_int),
                ("relaxation", C.c_double), ("quasi", C.c_int), ("quasi_iterations", C.c_int),
                ("adaptive", C.c_int), ("adaptive_iterations", C.c_int), ("max_refine", C.c_int),
                ("refine_tol", C.c_double), ("max_halvings", C.c_int), ("verbose", C.c_int),
                ("inertia_guard", C.c_double), ("learned", C.c_void_p), ("refine_accept", C.c_double),
                ("learned_mode", C.c_int), ("lengths", _dp), ("singular_point", C.c_int),
                ("sp_test_tol", C.c_double), ("sp_test_it", C.c_int), ("sp_tol_e", C.c_double),
                ("sp_tol_b", C.c_double), ("perturbation", C.c_double), ("sp_fail", C.c_int),
                ("length_after_sp", C.c_double)]


class AlmResult(C.Structure):
    _fields_ = [("steps_done", C.c_int), ("halvings", C.c_int), ("final_status", C.c_int),
                ("lambda_", _dp), ("unorm", _dp), ("indicator", _dp), ("negatives", _i64p),
                ("iterations", _ip), ("residue_f", _dp), ("length", _dp), ("U", _dp), ("w", _dp),
                ("probe_dof", C.c_int), ("wall_seconds", C.c_double), ("phi", C.c_double),
                ("jacobians", C.c_int64), ("residuals", C.c_int64), ("solves", C.c_int64),
                ("features", _dp), ("learned_steps", C.c_int64), ("fallbacks", C.c_int64), ("singular", _ip)]


class StaticOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("nsteps", C.c_int), ("dlambda", C.c_double), ("max_iter", C.c_int),
                ("tol", C.c_double), ("max_refine", C.c_int), ("refine_tol", C.c_double), ("verbose", C.c_int),
                ("inertia_guard", C.c_double), ("refine_accept", C.c_double)]


class StaticResult(C.Structure):
    _fields_ = [("steps_done", C.c_int), ("final_status", C.c_int), ("lambda_", _dp), ("unorm", _dp),
                ("iterations", _ip), ("negatives", _i64p), ("residual", _dp), ("U", _dp),
                ("wall_seconds", C.c_double), ("jacobians", C.c_int64)]


class DrOptions(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("tol_f", C.c_double), ("tol_u", C.c_double),
                ("damping", C.c_double), ("alpha", C.c_double), ("tol_e", C.c_double), ("reset_it", C.c_int),
                ("verbose", C.c_int)]


class DrResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("final_status", C.c_int), ("peaks", C.c_int), ("residual", C.c_double),
                ("ek", C.c_double), ("ek0", C.c_double), ("U", _dp), ("ek_history", _dp),
                ("wall_seconds", C.c_double), ("residuals", C.c_int64)]


class NewmarkOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("nsteps", C.c_int), ("dt", C.c_double), ("beta", C.c_double),
                ("gamma", C.c_double), ("damping_mass", C.c_double), ("max_iter", C.c_int), ("tol_f", C.c_double),
                ("tol_u", C.c_double), ("max_refine", C.c_int), ("refine_tol", C.c_double), ("verbose", C.c_int)]


class NewmarkResult(C.Structure):
    _fields_ = [("steps_done", C.c_int), ("final_status", C.c_int), ("time", _dp), ("unorm", _dp),
                ("iterations", _ip), ("U", _dp), ("V", _dp), ("A", _dp), ("w", _dp), ("probe_dof", C.c_int),
                ("wall_seconds", C.c_double), ("jacobians", C.c_int64)]


EIG_SHIFT_INVERT, EIG_BUCKLING = 0, 1


class EigOptions(C.Structure):
    _fields_ = [("mode", C.c_int), ("number", C.c_int), ("ncv_fac", C.c_int), ("shift", C.c_double),
                ("tol", C.c_double), ("max_restarts", C.c_int), ("precision", C.c_int), ("max_refine", C.c_int),
                ("refine_tol", C.c_double), ("seed", C.c_uint64)]


class EigResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("restarts", C.c_int), ("solves", C.c_int64), ("products", C.c_int64),
                ("wall_seconds", C.c_double)]


# every symbol declared in include/ksh.h: name -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    "ksh_last_error": (C.c_char_p, []),
    "ksh_device_count": (C.c_int, [_ip]),
    "ksh_problem_create": (C.c_int, [C.POINTER(Patch), C.POINTER(Material), C.POINTER(BC), C.POINTER(Loads),
                                     C.POINTER(_VP)]),
    "ksh_problem_destroy": (C.c_int, [_VP]),
    "ksh_problem_set_pressure": (C.c_int, [_VP, C.c_double]),
    "ksh_num_free_dofs": (C.c_int, [_VP, _i64p]),
    "ksh_bandwidth": (C.c_int, [_VP, _i64p]),
    "ksh_synchronize": (C.c_int, [_VP]),
    "ksh_assemble_force": (C.c_int, [_VP, _dp]),
    "ksh_assemble_residual": (C.c_int, [_VP, _dp, C.c_double, C.c_int, _dp]),
    "ksh_assemble_residual_dev": (C.c_int, [_VP, _VP, C.c_double, C.c_int, _VP]),
    "ksh_assemble_residual_dev_noalloc": (C.c_int, [_VP, _VP, C.c_double, C.c_int, _VP]),
    "ksh_matrix_create": (C.c_int, [_VP, C.c_int, C.POINTER(_VP)]),
    "ksh_matrix_destroy": (C.c_int, [_VP]),
    "ksh_matrix_dims": (C.c_int, [_VP, _i64p, _i64p, _ip]),
    "ksh_assemble_tangent": (C.c_int, [_VP, _dp, _VP]),
    "ksh_assemble_tangent_dev": (C.c_int, [_VP, _VP, _VP]),
    "ksh_assemble_mass": (C.c_int, [_VP, _VP]),
    "ksh_matrix_download": (C.c_int, [_VP, _dp]),
    "ksh_matrix_upload": (C.c_int, [_VP, _dp]),
    "ksh_matrix_axpby": (C.c_int, [_VP, C.c_double, C.c_double, _VP]),
    "ksh_matrix_copy": (C.c_int, [_VP, _VP]),
    "ksh_matvec": (C.c_int, [_VP, _dp, _dp]),
    "ksh_matvec_dev": (C.c_int, [_VP, _VP, _VP]),
    "ksh_factor": (C.c_int, [_VP]),
    "ksh_solve": (C.c_int, [_VP, _dp, _dp, C.c_int, C.c_double]),
    "ksh_solve_dev": (C.c_int, [_VP, _VP, _VP, C.c_int, C.c_double]),
    "ksh_inertia": (C.c_int, [_VP, _i64p, _dp]),
    "ksh_vector_d": (C.c_int, [_VP, _dp]),
    "ksh_set_inertia_guard": (C.c_int, [_VP, C.c_double]),
    "ksh_profile_enable": (C.c_int, [C.c_int]),
    "ksh_profile_reset": (C.c_int, []),
    "ksh_profile_query": (C.c_int, [C.c_char_p, _i64p, _dp]),
    "ksh_profile_query_work": (C.c_int, [C.c_char_p, _dp]),
    "ksh_tune_choice": (C.c_int, [C.c_char_p, _i64p, _i64p]),
    "ksh_run_alm": (C.c_int, [_VP, C.POINTER(AlmOptions), C.POINTER(AlmResult)]),
    "ksh_batch_run_alm": (C.c_int, [C.POINTER(_VP), C.c_int, C.POINTER(AlmOptions), C.POINTER(AlmResult), _dp]),
    "ksh_run_static": (C.c_int, [_VP, C.POINTER(StaticOptions), C.POINTER(StaticResult)]),
    "ksh_run_dr": (C.c_int, [_VP, C.POINTER(DrOptions), C.POINTER(DrResult)]),
    "ksh_batch_run_dr": (C.c_int, [C.POINTER(_VP), C.c_int, C.POINTER(DrOptions), C.POINTER(DrResult), _dp]),
    "ksh_run_newmark": (C.c_int, [_VP, C.POINTER(NewmarkOptions), _dp, _dp, _dp, C.POINTER(NewmarkResult)]),
    "ksh_batch_run_newmark": (C.c_int, [C.POINTER(_VP), C.c_int, C.POINTER(NewmarkOptions), C.POINTER(NewmarkResult),
                                        _dp]),
    "ksh_eigen": (C.c_int, [_VP, _VP, C.POINTER(EigOptions), _dp, _dp, C.POINTER(EigResult)]),
    "ksh_run_buckling": (C.c_int, [_VP, C.POINTER(EigOptions), _dp, _dp, C.POINTER(EigResult)]),
    "ksh_run_modal": (C.c_int, [_VP, C.POINTER(EigOptions), _dp, _dp, C.POINTER(EigResult)]),
    "ksh_mlp_create": (C.c_int, [_ip, C.c_uint64, C.POINTER(_VP)]),
    "ksh_mlp_destroy": (C.c_int, [_VP]),
    "ksh_mlp_num_params": (C.c_int, [_VP, _i64p]),
    "ksh_mlp_get_params": (C.c_int, [_VP, _fp]),
    "ksh_mlp_set_params": (C.c_int, [_VP, _fp]),
    "ksh_mlp_forward": (C.c_int, [_VP, _fp, C.c_int, _fp]),
    "ksh_mlp_gradients": (C.c_int, [_VP, _fp, _fp, C.c_int, _fp, _dp]),
    "ksh_mlp_train": (C.c_int, [_VP, _fp, _fp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_uint64, _dp]),
    "ksh_mlp_set_normalization": (C.c_int, [_VP, _dp, _dp, _dp, _dp]),
    "ksh_mlp_predict": (C.c_int, [_VP, _dp, C.c_int, _dp]),
    "ksh_predictor_features": (C.c_int, [_VP, _dp, C.c_double, _dp]),
    "ksh_predictor_decode": (C.c_int, [_VP, _dp, C.c_double, _dp]),
}

_lib = None


class KshError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES[code] if 0 <= code < 6 else code}: {msg}")
        self.code = code


def load(path: str = LIB_PATH):
    """Load libksh.so; raises if it is missing (there is no non-GPU execution path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise ImportError(f"libksh.so not built ({path}); run __graft_entry__.build()")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int):
    if code != SUCCESS:
        raise KshError(code, load().ksh_last_error().decode(errors="replace"))
    return code
