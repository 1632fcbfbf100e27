"""Python mirror of the reference accelerator API (namespace cfd::gpu,
src/gpu/include/cfd/gpu/krylov.hpp) over libcfdgpu's C-ABI.

Names, argument meaning and error behaviour follow the reference:
  * solve / solve_amg_cg take x as the initial guess and overwrite it with the solution
    (krylov.hpp:33-64);
  * non-convergence is reported in SolveStats.converged, never raised (krylov.hpp:26-31);
  * an empty hierarchy raises InvalidArgument, the reference's std::invalid_argument
    (device_solver.cpp:88-90, krylov_openmp.cpp:189-191).
The reference's host paths (solve_host, solve_amg_cg_host) stay in the reference; this backend
only provides the device paths and raises NoDevice when no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L


class Method(IntEnum):
    """Method (krylov.hpp:19)."""

    Cg = 0
    BiCgStab = 1


class Precision(IntEnum):
    """Arithmetic of the AMG-PCG solve (FP64 = the reference algorithm)."""

    FP64 = 0
    MIXED_FP32 = 1
    MIXED_BF16 = 2


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


@dataclass
class CsrView:
    """CsrView (krylov.hpp:10-16); n_cols defaults to n (square)."""

    n: int
    nnz: int
    row_ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    n_cols: int = 0

    def __post_init__(self):
        self.row_ptr = _i32(self.row_ptr)
        self.col = _i32(self.col)
        self.val = _f64(self.val)

    @staticmethod
    def from_scipy(m) -> "CsrView":
        m = m.tocsr()
        return CsrView(m.shape[0], m.nnz, m.indptr, m.indices, m.data, m.shape[1])

    def to_scipy(self):
        import scipy.sparse as sp

        cols = self.n_cols if self.n_cols > 0 else self.n
        return sp.csr_matrix((self.val, self.col, self.row_ptr), shape=(self.n, cols))

    def c(self) -> L.Csr:
        s = L.Csr()
        s.n_rows = self.n
        s.n_cols = self.n_cols
        s.nnz = self.nnz
        s.row_ptr = L.iptr(self.row_ptr)
        s.col = L.iptr(self.col)
        s.val = L.dptr(self.val)
        return s


@dataclass
class SolveControl:
    """SolveControl (krylov.hpp:21-24)."""

    abs_tolerance: float = 1e-10
    max_iterations: int = 1000

    def c(self) -> L.SolveControl:
        return L.SolveControl(float(self.abs_tolerance), int(self.max_iterations))


@dataclass
class SolveStats:
    """SolveStats (krylov.hpp:26-31) plus the true residual, |b| and device time."""

    iterations: int = 0
    residual: float = 0.0
    converged: bool = False
    on_device: bool = False
    exit_reason: int = 0
    true_residual: float = 0.0
    rhs_norm: float = 0.0
    solve_ms: float = 0.0

    @staticmethod
    def from_c(s: L.SolveStats) -> "SolveStats":
        return SolveStats(s.iterations, s.residual, bool(s.converged), bool(s.on_device),
                          s.exit_reason, s.true_residual, s.rhs_norm, s.solve_ms)


@dataclass
class AmgLevelView:
    """AmgLevelView (krylov.hpp:48-52)."""

    a: Optional[CsrView] = None
    p: Optional[CsrView] = None
    r: Optional[CsrView] = None


@dataclass
class AmgView:
    """AmgView (krylov.hpp:53-58)."""

    levels: List[AmgLevelView] = field(default_factory=list)
    coarse_inverse: Optional[np.ndarray] = None
    smooth_sweeps: int = 2
    jacobi_weight: float = 0.7

    def c(self):
        """C view plus the objects that must stay alive while it is used."""
        n = len(self.levels)
        arr = (L.AmgLevelView * max(n, 1))()
        for i, lv in enumerate(self.levels):
            if lv.a is not None:
                arr[i].a = lv.a.c()
            if lv.p is not None:
                arr[i].p = lv.p.c()
            if lv.r is not None:
                arr[i].r = lv.r.c()
        v = L.AmgView()
        v.n_levels = n
        v.levels = C.cast(arr, C.POINTER(L.AmgLevelView))
        ci = None
        if self.coarse_inverse is not None:
            ci = _f64(self.coarse_inverse).reshape(-1)
            v.coarse_inverse = L.dptr(ci)
        v.smooth_sweeps = int(self.smooth_sweeps)
        v.jacobi_weight = float(self.jacobi_weight)
        return v, (arr, ci, self)


@dataclass
class LduPattern:
    """LduPattern (krylov.hpp:68-77)."""

    n: int
    n_faces: int
    nnz: int
    row_ptr: np.ndarray
    col: np.ndarray
    diag_pos: np.ndarray
    upper_pos: np.ndarray
    lower_pos: np.ndarray

    def __post_init__(self):
        for k in ("row_ptr", "col", "diag_pos", "upper_pos", "lower_pos"):
            setattr(self, k, _i32(getattr(self, k)))

    def c(self) -> L.LduPattern:
        p = L.LduPattern()
        p.n, p.n_faces, p.nnz = self.n, self.n_faces, self.nnz
        p.row_ptr = L.iptr(self.row_ptr)
        p.col = L.iptr(self.col)
        p.diag_pos = L.iptr(self.diag_pos)
        p.upper_pos = L.iptr(self.upper_pos)
        p.lower_pos = L.iptr(self.lower_pos)
        return p


def _check_vec(v: np.ndarray, n: int, name: str) -> None:
    if not (isinstance(v, np.ndarray) and v.dtype == np.float64 and v.flags.c_contiguous
            and v.shape == (n,)):
        raise L.InvalidArgument(1, f"{name} must be a contiguous float64 array of length {n}")


def solve(a: CsrView, b: np.ndarray, x: np.ndarray, method: Method,
          control: SolveControl) -> SolveStats:
    """solve() (krylov.hpp:37-38): Jacobi-preconditioned CG / BiCGStab on the GPU."""
    _check_vec(b, a.n, "b")
    _check_vec(x, a.n, "x")
    st = L.SolveStats()
    ca, cc = a.c(), control.c()
    L.call("cfdg_solve_csr", C.byref(ca), L.dptr(b), L.dptr(x), int(method), C.byref(cc),
           C.byref(st))
    return SolveStats.from_c(st)


# Limits of the allocation-free solve (cfdgpu.h).
NOALLOC_MAX_ROWS = 524288
NOALLOC_SCRATCH_BYTES = 51380224


def solve_device_noalloc(n: int, nnz: int, row_ptr: int, col: int, val: int, b: int, x: int,
                         method: Method, control: SolveControl) -> SolveStats:
    """cfdg_solve_csr_noalloc (no counterpart in the reference): solve() on arrays that already
    live in device memory, given as integer device addresses (row_ptr int32[n + 1], col
    int32[nnz], val float64[nnz], b and x float64[n], x in/out), with no device allocation: the
    work space is the library's fixed scratch area (at most NOALLOC_MAX_ROWS rows)."""
    ca = L.Csr()
    ca.n_rows, ca.n_cols, ca.nnz = int(n), 0, int(nnz)
    ca.row_ptr = C.cast(C.c_void_p(row_ptr), C.POINTER(C.c_int))
    ca.col = C.cast(C.c_void_p(col), C.POINTER(C.c_int))
    ca.val = C.cast(C.c_void_p(val), C.POINTER(C.c_double))
    st = L.SolveStats()
    cc = control.c()
    L.call("cfdg_solve_csr_noalloc", C.byref(ca), C.c_void_p(b), C.c_void_p(x), int(method),
           C.byref(cc), C.byref(st))
    return SolveStats.from_c(st)


def solve_amg_cg(amg: AmgView, b: np.ndarray, x: np.ndarray, control: SolveControl,
                 precision: Precision = Precision.FP64) -> SolveStats:
    """solve_amg_cg() (krylov.hpp:61-62): CG preconditioned by one V-cycle per iteration."""
    if not amg.levels:
        raise L.InvalidArgument(1, "solve_amg_cg: empty hierarchy")
    n = amg.levels[0].a.n
    _check_vec(b, n, "b")
    _check_vec(x, n, "x")
    v, keep = amg.c()
    st = L.SolveStats()
    cc = control.c()
    L.call("cfdg_solve_amg_cg_prec", C.byref(v), L.dptr(b), L.dptr(x), int(precision),
           C.byref(cc), C.byref(st))
    del keep
    return SolveStats.from_c(st)


class DeviceSolver:
    """class DeviceSolver (krylov.hpp:79-103): persistent system on one LDU pattern."""

    def __init__(self, pattern: LduPattern):
        self._pattern = pattern
        self.n = pattern.n
        self.n_faces = pattern.n_faces
        h = C.c_void_p()
        p = pattern.c()
        L.call("cfdg_system_create", C.byref(p), C.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().cfdg_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_amg(self, amg: AmgView) -> None:
        """set_amg() (device_solver.cpp:87-115); levels[0].a is ignored."""
        if not amg.levels:
            raise L.InvalidArgument(1, "DeviceSolver::set_amg: empty hierarchy")
        v, keep = amg.c()
        L.call("cfdg_system_set_amg", self._h, C.byref(v))
        del keep

    def solve(self, diag, upper, lower, b, x, method: Method, use_amg: bool,
              control: SolveControl) -> SolveStats:
        """solve() (device_solver.cpp:117-150): assemble from LDU and solve, x in/out."""
        diag, upper, lower, b = _f64(diag), _f64(upper), _f64(lower), _f64(b)
        _check_vec(x, self.n, "x")
        st = L.SolveStats()
        cc = control.c()
        L.call("cfdg_system_solve", self._h, L.dptr(diag), L.dptr(upper), L.dptr(lower),
               L.dptr(b), L.dptr(x), int(method), int(bool(use_amg)), C.byref(cc), C.byref(st))
        return SolveStats.from_c(st)

    def on_device(self) -> bool:
        v = C.c_int()
        L.call("cfdg_system_on_device", self._h, C.byref(v))
        return bool(v.value)

    # --- resident use -------------------------------------------------------------------
    def assemble(self, diag, upper, lower) -> None:
        diag, upper, lower = _f64(diag), _f64(upper), _f64(lower)
        L.call("cfdg_system_assemble", self._h, L.dptr(diag), L.dptr(upper), L.dptr(lower))

    def assemble_pressure(self, fv) -> None:
        """Matrix and right-hand side from the pressure equation last assembled by `fv` (a
        DeviceFv on this system's mesh), copied within HBM."""
        L.call("cfdg_system_assemble_pressure", self._h, fv._h)

    def set_rhs(self, b) -> None:
        b = _f64(b)
        L.call("cfdg_system_set_rhs", self._h, L.dptr(b))

    def set_x(self, x=None) -> None:
        if x is None:
            L.call("cfdg_system_set_x", self._h, None)
        else:
            x = _f64(x)
            L.call("cfdg_system_set_x", self._h, L.dptr(x))

    def get_x(self) -> np.ndarray:
        x = np.empty(self.n)
        L.call("cfdg_system_get_x", self._h, L.dptr(x))
        return x

    def normalised_residual(self):
        """LduMatrix::normalised_residual / residual_norms (ldu_matrix.cpp:65-103) of the
        resident matrix, right-hand side and iterate, on the device: (res / (norm + 1e-20), res,
        norm) with res = |b - A x|_1."""
        v, r, m = C.c_double(), C.c_double(), C.c_double()
        L.call("cfdg_system_normalised_residual", self._h, C.byref(v), C.byref(r), C.byref(m))
        return v.value, r.value, m.value

    def solve_resident(self, method: Method, use_amg: bool, precision: Precision,
                       control: SolveControl) -> SolveStats:
        st = L.SolveStats()
        cc = control.c()
        L.call("cfdg_system_solve_resident", self._h, int(method), int(bool(use_amg)),
               int(precision), C.byref(cc), C.byref(st))
        return SolveStats.from_c(st)

    def probe(self, kind: int) -> None:
        """Times one kernel class with HIP events during the next solves (1: last level-0
        Jacobi sweep, 2: outer FP64 SpMV, 0: off)."""
        L.call("cfdg_system_probe", self._h, int(kind))

    def probe_read(self, precision: Precision):
        """(average launch ms, launches, algorithmic bytes per launch)."""
        ms, n, by = C.c_double(), C.c_longlong(), C.c_double()
        L.call("cfdg_system_probe_read", self._h, int(precision), C.byref(ms), C.byref(n),
               C.byref(by))
        return ms.value, n.value, by.value

    def footprint(self) -> int:
        v = C.c_size_t()
        L.call("cfdg_system_footprint", self._h, C.byref(v))
        return int(v.value)


class Batch:
    """Resident systems solved together: one kernel launch per solver stage covers every
    member, each keeping its own reductions and stopping test. No counterpart in the
    reference, whose DeviceSolver solves one system per call (device_solver.cpp:117-150)."""

    def __init__(self, systems: Sequence[DeviceSolver], method: Method = Method.Cg,
                 use_amg: bool = True, precision: Precision = Precision.FP64):
        self._systems = list(systems)  # keeps the members alive while the batch exists
        arr = (C.c_void_p * max(len(self._systems), 1))(*[s._h for s in self._systems])
        h = C.c_void_p()
        L.call("cfdg_batch_create", arr, len(self._systems), int(method), int(bool(use_amg)),
               int(precision), C.byref(h))
        self._h = h

    def __len__(self) -> int:
        return len(self._systems)

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.lib().cfdg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def zero_x(self) -> None:
        """Sets every member's initial guess to zero with one launch (the batched form of
        DeviceSolver.set_x(None))."""
        L.call("cfdg_batch_zero_x", self._h)

    def solve(self, controls: Sequence[SolveControl]) -> List[SolveStats]:
        """Solves every member from its resident right-hand side and iterate."""
        n = len(self._systems)
        if len(controls) != n:
            raise L.InvalidArgument(1, "batch: one control per system")
        cc = (L.SolveControl * n)(*[c.c() for c in controls])
        st = (L.SolveStats * n)()
        L.call("cfdg_batch_solve", self._h, cc, st)
        return [SolveStats.from_c(s) for s in st]


def assemble_convection_diffusion(n_cells: int, n_faces: int, owner, neighbour, flux, gamma,
                                  delta_coeff, convection: float, upper: np.ndarray,
                                  lower: np.ndarray, diag: np.ndarray,
                                  divergence: np.ndarray) -> None:
    """assemble_convection_diffusion() (krylov.hpp:113-116): upper / lower written, diag and
    divergence accumulated in place."""
    owner, neighbour = _i32(owner), _i32(neighbour)
    flux, gamma, delta_coeff = _f64(flux), _f64(gamma), _f64(delta_coeff)
    for name, v, n in (("upper", upper, n_faces), ("lower", lower, n_faces),
                       ("diag", diag, n_cells), ("divergence", divergence, n_cells)):
        _check_vec(v, n, name)
    L.call("cfdg_assemble_convection_diffusion", int(n_cells), int(n_faces), L.iptr(owner),
           L.iptr(neighbour), L.dptr(flux), L.dptr(gamma), L.dptr(delta_coeff),
           float(convection), L.dptr(upper), L.dptr(lower), L.dptr(diag), L.dptr(divergence))


def backend_name() -> str:
    """backend_name() (krylov.hpp:119)."""
    buf = C.create_string_buffer(64)
    L.call("cfdg_backend_name", buf, len(buf))
    return buf.value.decode()


def device_available() -> bool:
    """device_available() (krylov.hpp:121)."""
    v = C.c_int()
    L.call("cfdg_device_available", C.byref(v))
    return bool(v.value)


def device_description() -> str:
    """device_description() (krylov.hpp:123)."""
    buf = C.create_string_buffer(256)
    L.call("cfdg_device_description", buf, len(buf))
    return buf.value.decode()
