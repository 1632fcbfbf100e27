"""ctypes binding of libcfdgpu.so (the C-ABI declared in include/cfdgpu.h).

The library is built in-tree into cfd_amd/lib/ by cfd_amd/build.py. Loading fails loudly when it
is missing: there is no fallback implementation of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libcfdgpu.so"

CFDG_OK = 0
_STATUS_NAMES = {
    1: "CFDG_ERR_INVALID_ARGUMENT",
    2: "CFDG_ERR_NO_DEVICE",
    3: "CFDG_ERR_HIP",
    4: "CFDG_ERR_OUT_OF_MEMORY",
    5: "CFDG_ERR_INTERNAL",
    6: "CFDG_ERR_CAPACITY",
}


class CfdgError(RuntimeError):
    """A libcfdgpu call failed; .status holds the cfdg_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class InvalidArgument(CfdgError, ValueError):
    """Counterpart of the reference's std::invalid_argument."""


class NoDevice(CfdgError):
    """No HIP device is visible; the library never falls back to the host."""


class Csr(C.Structure):
    _fields_ = [
        ("n_rows", C.c_int),
        ("n_cols", C.c_int),
        ("nnz", C.c_int),
        ("row_ptr", C.POINTER(C.c_int)),
        ("col", C.POINTER(C.c_int)),
        ("val", C.POINTER(C.c_double)),
    ]


class SolveControl(C.Structure):
    _fields_ = [("abs_tolerance", C.c_double), ("max_iterations", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("on_device", C.c_int),
        ("exit_reason", C.c_int),
        ("residual", C.c_double),
        ("true_residual", C.c_double),
        ("rhs_norm", C.c_double),
        ("solve_ms", C.c_double),
    ]


class AmgLevelView(C.Structure):
    _fields_ = [("a", Csr), ("p", Csr), ("r", Csr)]


class AmgView(C.Structure):
    _fields_ = [
        ("n_levels", C.c_int),
        ("levels", C.POINTER(AmgLevelView)),
        ("coarse_inverse", C.POINTER(C.c_double)),
        ("smooth_sweeps", C.c_int),
        ("jacobi_weight", C.c_double),
    ]


class LduPattern(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("n_faces", C.c_int),
        ("nnz", C.c_int),
        ("row_ptr", C.POINTER(C.c_int)),
        ("col", C.POINTER(C.c_int)),
        ("diag_pos", C.POINTER(C.c_int)),
        ("upper_pos", C.POINTER(C.c_int)),
        ("lower_pos", C.POINTER(C.c_int)),
    ]


class ServerOptions(C.Structure):
    _fields_ = [
        ("max_active_solves", C.c_int),
        ("method", C.c_int),
        ("use_amg", C.c_int),
        ("precision", C.c_int),
        ("smooth_sweeps", C.c_int),
        ("memory_budget_bytes", C.c_size_t),
        ("staging_threads", C.c_int),
    ]


class ServerRequest(C.Structure):
    _fields_ = [
        ("pattern", LduPattern),
        ("diag", C.POINTER(C.c_double)),
        ("upper", C.POINTER(C.c_double)),
        ("lower", C.POINTER(C.c_double)),
        ("amg", C.POINTER(AmgView)),
        ("n_steps", C.c_int),
        ("rhs", C.POINTER(C.POINTER(C.c_double))),
        ("x0", C.POINTER(C.c_double)),
        ("rel_tolerance", C.c_double),
        ("max_iterations", C.c_int),
        ("x_out", C.POINTER(C.POINTER(C.c_double))),
        ("stats_out", C.POINTER(SolveStats)),
        ("step_latency_ms", C.POINTER(C.c_double)),
    ]


class FvMesh(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int),
        ("n_internal_faces", C.c_int),
        ("n_faces", C.c_int),
        ("owner", C.POINTER(C.c_int)),
        ("neighbour", C.POINTER(C.c_int)),
        ("cell_offsets", C.POINTER(C.c_int)),
        ("cell_faces", C.POINTER(C.c_int)),
        ("weight", C.POINTER(C.c_double)),
        ("sf", C.POINTER(C.c_double)),
        ("face_centre", C.POINTER(C.c_double)),
        ("cell_centre", C.POINTER(C.c_double)),
        ("volume", C.POINTER(C.c_double)),
        ("correction", C.POINTER(C.c_double)),
        ("lsq_d", C.POINTER(C.c_double)),
        ("lsq_inverse", C.POINTER(C.c_double)),
    ]


class PressureInputs(C.Structure):
    _fields_ = [
        ("upper", C.POINTER(C.c_double)),
        ("lower", C.POINTER(C.c_double)),
        ("diag", C.POINTER(C.c_double)),
        ("source", C.POINTER(C.c_double)),
        ("u", C.POINTER(C.c_double)),
        ("delta_coeffs", C.POINTER(C.c_double)),
        ("boundary_kind", C.POINTER(C.c_int)),
        ("fixed_velocity", C.POINTER(C.c_double)),
        ("outlet_pressure", C.POINTER(C.c_double)),
        ("u_prev", C.POINTER(C.c_double)),
        ("phi_prev", C.POINTER(C.c_double)),
        ("alpha_u", C.c_double),
        ("pin_reference", C.c_int),
    ]


class PressureOutputs(C.Structure):
    _fields_ = [(name, C.POINTER(C.c_double))
                for name in ("diag", "upper", "lower", "source", "phi_hbya", "rau_f", "hbya", "rau")]


class ServerCounters(C.Structure):
    _fields_ = [
        ("submitted", C.c_ulonglong),
        ("completed", C.c_ulonglong),
        ("solves", C.c_ulonglong),
        ("resident_bytes", C.c_size_t),
        ("peak_resident_bytes", C.c_size_t),
        ("budget_bytes", C.c_size_t),
        ("active_solves", C.c_int),
        ("peak_active_solves", C.c_int),
        ("waiting", C.c_int),
        ("staging_ms", C.c_double),
        ("retiring_ms", C.c_double),
        ("engine_busy_ms", C.c_double),
    ]


class LayoutRule(C.Structure):
    _fields_ = [
        ("min_rows", C.c_int),
        ("max_rows", C.c_int),
        ("min_mean", C.c_double),
        ("max_mean", C.c_double),
        ("lanes", C.c_int),
        ("csr", C.c_int),
    ]


class AmgSettings(C.Structure):
    _fields_ = [
        ("strength_threshold", C.c_double),
        ("max_levels", C.c_int),
        ("coarse_size", C.c_int),
        ("smooth_sweeps", C.c_int),
        ("jacobi_weight", C.c_double),
        ("strength_decay", C.c_double),
    ]


class AmgCacheOptions(C.Structure):
    _fields_ = [("max_age", C.c_int), ("stale_rebuild", C.c_int), ("amg", AmgSettings), ("device_setup", C.c_int)]


class LinearSettings(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("tolerance", C.c_double),
        ("rel_tolerance", C.c_double),
        ("max_iterations", C.c_int),
        ("precision", C.c_int),
    ]


class LinearResult(C.Structure):
    _fields_ = [
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("iterations", C.c_int),
        ("solved", C.c_int),
        ("used_amg", C.c_int),
        ("setups", C.c_int),
        ("target", C.c_double),
        ("stats", SolveStats),
    ]


class BoxSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("nz", C.c_int),
        ("lo", C.c_double * 3),
        ("hi", C.c_double * 3),
        ("jitter", C.c_double),
        ("seed", C.c_uint64),
    ]


P_INT = C.POINTER(C.c_int)
P_DBL = C.POINTER(C.c_double)
VP = C.c_void_p
ST = C.c_int

# name: (restype, argtypes)
_SIGNATURES = {
    "cfdg_version": (C.c_int, []),
    "cfdg_backend_name": (ST, [C.c_char_p, C.c_size_t]),
    "cfdg_device_available": (ST, [P_INT]),
    "cfdg_device_description": (ST, [C.c_char_p, C.c_size_t]),
    "cfdg_release_cached_memory": (ST, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "cfdg_last_error": (ST, [C.c_char_p, C.c_size_t]),
    "cfdg_solve_csr": (ST, [C.POINTER(Csr), P_DBL, P_DBL, C.c_int, C.POINTER(SolveControl),
                            C.POINTER(SolveStats)]),
    "cfdg_solve_csr_noalloc": (ST, [C.POINTER(Csr), VP, VP, C.c_int, C.POINTER(SolveControl),
                                    C.POINTER(SolveStats)]),
    "cfdg_solve_amg_cg": (ST, [C.POINTER(AmgView), P_DBL, P_DBL, C.POINTER(SolveControl),
                               C.POINTER(SolveStats)]),
    "cfdg_solve_amg_cg_prec": (ST, [C.POINTER(AmgView), P_DBL, P_DBL, C.c_int,
                                    C.POINTER(SolveControl), C.POINTER(SolveStats)]),
    "cfdg_assemble_convection_diffusion": (ST, [C.c_int, C.c_int, P_INT, P_INT, P_DBL, P_DBL,
                                                P_DBL, C.c_double, P_DBL, P_DBL, P_DBL, P_DBL]),
    "cfdg_system_create": (ST, [C.POINTER(LduPattern), C.POINTER(VP)]),
    "cfdg_system_destroy": (ST, [VP]),
    "cfdg_system_on_device": (ST, [VP, P_INT]),
    "cfdg_system_set_amg": (ST, [VP, C.POINTER(AmgView)]),
    "cfdg_system_solve": (ST, [VP, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL, C.c_int, C.c_int,
                               C.POINTER(SolveControl), C.POINTER(SolveStats)]),
    "cfdg_system_assemble": (ST, [VP, P_DBL, P_DBL, P_DBL]),
    "cfdg_system_set_rhs": (ST, [VP, P_DBL]),
    "cfdg_system_set_x": (ST, [VP, P_DBL]),
    "cfdg_system_get_x": (ST, [VP, P_DBL]),
    "cfdg_system_solve_resident": (ST, [VP, C.c_int, C.c_int, C.c_int, C.POINTER(SolveControl),
                                        C.POINTER(SolveStats)]),
    "cfdg_system_footprint": (ST, [VP, C.POINTER(C.c_size_t)]),
    "cfdg_batch_create": (ST, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_int,
                               C.POINTER(VP)]),
    "cfdg_batch_destroy": (ST, [VP]),
    "cfdg_batch_solve": (ST, [VP, C.POINTER(SolveControl), C.POINTER(SolveStats)]),
    "cfdg_batch_zero_x": (ST, [VP]),
    "cfdg_set_layout_rules": (ST, [C.POINTER(LayoutRule), C.c_int]),
    "cfdg_layout_of": (ST, [C.c_int, P_INT, P_INT, P_INT]),
    "cfdg_set_mixed_smoother": (ST, [C.c_int, C.c_double]),
    "cfdg_spmv_time": (ST, [C.POINTER(Csr), C.c_int, C.c_int, C.c_int, C.c_int, P_DBL]),
    "cfdg_estimate_footprint": (ST, [C.POINTER(LduPattern), C.POINTER(AmgView),
                                     C.POINTER(C.c_size_t)]),
    "cfdg_server_default_options": (ST, [C.POINTER(ServerOptions)]),
    "cfdg_server_create": (ST, [C.POINTER(ServerOptions), C.POINTER(VP)]),
    "cfdg_server_submit": (ST, [VP, C.POINTER(ServerRequest), C.POINTER(C.c_ulonglong)]),
    "cfdg_server_wait": (ST, [VP, C.c_ulonglong, P_DBL]),
    "cfdg_server_query": (ST, [VP, C.POINTER(ServerCounters)]),
    "cfdg_server_destroy": (ST, [VP]),
    "cfdg_system_probe": (ST, [VP, C.c_int]),
    "cfdg_system_probe_read": (ST, [VP, C.c_int, P_DBL, C.POINTER(C.c_longlong), P_DBL]),
    "cfdg_amg_default_settings": (ST, [C.POINTER(AmgSettings)]),
    "cfdg_amg_build": (ST, [C.POINTER(Csr), C.POINTER(AmgSettings), C.POINTER(VP)]),
    "cfdg_amg_build_device": (ST, [C.POINTER(Csr), C.POINTER(AmgSettings), C.POINTER(VP)]),
    "cfdg_amg_destroy": (ST, [VP]),
    "cfdg_amg_n_levels": (ST, [VP, P_INT]),
    "cfdg_amg_level_size": (ST, [VP, C.c_int, C.c_int, P_INT, P_INT, P_INT]),
    "cfdg_amg_level_copy": (ST, [VP, C.c_int, C.c_int, P_INT, P_INT, P_DBL]),
    "cfdg_amg_aggregates": (ST, [VP, C.c_int, P_INT]),
    "cfdg_amg_coarse_inverse": (ST, [VP, P_DBL]),
    "cfdg_amg_view_of": (ST, [VP, C.POINTER(AmgLevelView), C.POINTER(AmgView)]),
    "cfdg_amg_setup_ms": (ST, [VP, P_DBL]),
    "cfdg_amg_coarse_on_device": (ST, [VP, P_INT]),
    "cfdg_dense_inverse_batch": (ST, [C.c_int, P_INT, C.POINTER(VP), C.POINTER(VP), P_INT]),
    "cfdg_amg_cache_default_options": (ST, [C.POINTER(AmgCacheOptions)]),
    "cfdg_amg_cache_create": (ST, [C.POINTER(LduPattern), C.POINTER(AmgCacheOptions), C.POINTER(VP)]),
    "cfdg_amg_cache_destroy": (ST, [VP]),
    "cfdg_amg_cache_state": (ST, [VP, P_INT, P_INT, P_INT]),
    "cfdg_ldu_solve": (ST, [VP, P_DBL, P_DBL, P_DBL, P_DBL, P_DBL, C.POINTER(LinearSettings),
                            C.POINTER(LinearResult)]),
    "cfdg_system_normalised_residual": (ST, [VP, P_DBL, P_DBL, P_DBL]),
    "cfdg_fv_create": (ST, [C.POINTER(FvMesh), C.POINTER(VP)]),
    "cfdg_fv_destroy": (ST, [VP]),
    "cfdg_fv_gradient": (ST, [VP, P_DBL, P_DBL, C.c_int, P_DBL]),
    "cfdg_fv_vector_gradient": (ST, [VP, P_DBL, P_DBL, C.c_int, P_DBL]),
    "cfdg_fv_scalar_corrections": (ST, [VP, P_DBL, P_DBL, C.c_int, P_DBL, P_DBL, C.c_int, P_DBL]),
    "cfdg_fv_momentum_corrections": (ST, [VP, P_DBL, P_DBL, C.c_int, C.c_int, C.c_int, P_DBL, C.c_double,
                                          C.c_int, P_DBL]),
    "cfdg_fv_pressure_equation": (ST, [VP, C.POINTER(PressureInputs), C.POINTER(PressureOutputs)]),
    "cfdg_system_assemble_pressure": (ST, [VP, VP]),
    "cfdg_fv_on_device": (ST, [VP, P_INT]),
    "cfdg_fv_timing": (ST, [VP, C.c_int]),
    "cfdg_fv_last_kernel_ms": (ST, [VP, P_DBL]),
    "cfdg_box_sizes": (ST, [C.POINTER(BoxSpec), P_INT, P_INT]),
    "cfdg_box_build": (ST, [C.POINTER(BoxSpec), P_INT, P_INT, P_DBL, P_DBL, P_DBL, P_DBL]),
    "cfdg_uniform": (ST, [C.c_uint64, C.c_int64, C.c_double, C.c_double, P_DBL]),
    "cfdg_csr_pattern": (ST, [C.c_int, C.c_int, P_INT, P_INT, P_INT, P_INT, P_INT, P_INT, P_INT]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads libcfdgpu.so once; raises ImportError if it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("CFDG_LIBRARY", LIB_PATH))
        if not path.exists():
            raise ImportError(
                f"libcfdgpu.so not found at {path}; build it with `python cfd_amd/build.py` "
                "(there is no fallback implementation)")
        h = C.CDLL(str(path))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().cfdg_last_error(buf, len(buf))
    return buf.value.decode()


def check(status: int) -> None:
    if status == CFDG_OK:
        return
    msg = last_error()
    if status == 1:
        raise InvalidArgument(status, msg)
    if status == 2:
        raise NoDevice(status, msg)
    raise CfdgError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(P_INT)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(P_DBL)


def opt_dptr(a):
    return None if a is None else dptr(a)
