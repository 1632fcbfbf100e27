"""The reference's caller policy for the device solver (cfdg_amg_cache_* / cfdg_ldu_solve):
LduMatrix::solve with a CSR pattern and an AmgCache (src/fv/src/ldu_matrix.cpp:111-131,
solve_persistent 287-366; src/fv/include/cfd/fv/ldu_matrix.hpp:23-34, 61-75)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .amg import AmgSettings
from .krylov import LduPattern, Method, Precision, SolveStats, _check_vec, _f64


@dataclass
class LinearSolverSettings:
    """LinearSolverSettings (ldu_matrix.hpp:23-28) of the device solver types: Method.Cg is
    GpuCg, Method.BiCgStab is GpuBiCgStab. precision selects the V-cycle arithmetic of AMG-PCG
    (FP64 is the reference)."""

    method: Method = Method.BiCgStab
    tolerance: float = 1e-10
    rel_tolerance: float = 0.01
    max_iterations: int = 1000
    precision: Precision = Precision.FP64

    def c(self) -> L.LinearSettings:
        return L.LinearSettings(int(self.method), float(self.tolerance), float(self.rel_tolerance),
                                int(self.max_iterations), int(self.precision))


@dataclass
class LinearSolveResult:
    """LinearSolveResult (ldu_matrix.hpp:30-34) plus what the cache policy did."""

    initial_residual: float = 0.0
    final_residual: float = 0.0
    iterations: int = 0
    solved: bool = False
    used_amg: bool = False
    setups: int = 0
    target: float = 0.0
    stats: SolveStats = field(default_factory=SolveStats)


class AmgCache:
    """AmgCache (ldu_matrix.hpp:61-75) with the device system it drives: solve() is
    LduMatrix::solve(x, settings, &cache) for GpuCg / GpuBiCgStab with a CSR pattern.
    stale_rebuild adds the host AmgCg path's stale-hierarchy rule (ldu_matrix.cpp:242-256).
    device_setup: rebuilds with the sparse products and the coarsest inverse on the GPU (the
    default; the same hierarchy, the coarsest inverse equal to rounding), else on the host."""

    def __init__(self, pattern: LduPattern, max_age: int = 20, stale_rebuild: bool = False,
                 amg: AmgSettings | None = None, device_setup: bool = True):
        o = L.AmgCacheOptions()
        L.call("cfdg_amg_cache_default_options", C.byref(o))
        o.max_age = int(max_age)
        o.stale_rebuild = int(bool(stale_rebuild))
        o.device_setup = int(bool(device_setup))
        if amg is not None:
            o.amg = amg.c()
        self.pattern = pattern
        self.n = pattern.n
        h = C.c_void_p()
        cp = pattern.c()
        L.call("cfdg_amg_cache_create", C.byref(cp), C.byref(o), C.byref(h))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            L.call("cfdg_amg_cache_destroy", self._h)
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

    def state(self):
        """(age, levels of the current hierarchy, hierarchies built so far)."""
        a, l, s = C.c_int(), C.c_int(), C.c_int()
        L.call("cfdg_amg_cache_state", self._h, C.byref(a), C.byref(l), C.byref(s))
        return a.value, l.value, s.value

    @property
    def age(self) -> int:
        return self.state()[0]

    def solve(self, diag, upper, lower, source, x: np.ndarray,
              settings: LinearSolverSettings) -> LinearSolveResult:
        diag, upper, lower, source = _f64(diag), _f64(upper), _f64(lower), _f64(source)
        _check_vec(x, self.n, "x")
        r = L.LinearResult()
        cs = settings.c()
        L.call("cfdg_ldu_solve", self._h, L.dptr(diag), L.dptr(upper), L.dptr(lower), L.dptr(source),
               L.dptr(x), C.byref(cs), C.byref(r))
        return LinearSolveResult(r.initial_residual, r.final_residual, r.iterations, bool(r.solved),
                                 bool(r.used_amg), r.setups, r.target, SolveStats.from_c(r.stats))
