"""Serving loop over libcfdgpu's cfdg_server_* C-ABI (no counterpart in the reference, whose
DeviceSolver serves one solve per call, src/gpu/src/device_solver.cpp:117-150).

A request is one mesh's system (LDU pattern + coefficients + AMG hierarchy) with a run of
time-step right-hand sides. Requests are admitted in arrival order while their device footprint
fits the memory budget; each step starts from the previous step's solution and shares every
kernel launch with the other running solves (continuous batching). Blocking calls release the
GIL, so several Python threads may submit and wait at once.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .krylov import AmgView, LduPattern, Method, Precision, SolveStats, _check_vec, _f64


def estimate_footprint(pattern: LduPattern, amg: Optional[AmgView] = None) -> int:
    """Upper bound of the device bytes of a system of this pattern and hierarchy (host-only)."""
    out = C.c_size_t()
    p = pattern.c()
    if amg is None:
        L.call("cfdg_estimate_footprint", C.byref(p), None, C.byref(out))
    else:
        v, keep = amg.c()
        L.call("cfdg_estimate_footprint", C.byref(p), C.byref(v), C.byref(out))
        del keep
    return int(out.value)


@dataclass
class ServerCounters:
    submitted: int
    completed: int
    solves: int
    resident_bytes: int
    peak_resident_bytes: int
    budget_bytes: int
    active_solves: int
    peak_active_solves: int
    waiting: int
    staging_ms: float = 0.0
    retiring_ms: float = 0.0
    engine_busy_ms: float = 0.0


@dataclass
class ServedRequest:
    """Handle of a submitted request; filled in by Server.wait."""

    id: int
    n_steps: int
    stats: List[SolveStats] = field(default_factory=list)
    step_latency_ms: Optional[np.ndarray] = None
    solutions: Optional[List[np.ndarray]] = None
    latency_ms: float = 0.0
    _keep: tuple = ()


class Server:
    def __init__(self, max_active_solves: int = 64, method: Method = Method.Cg,
                 use_amg: bool = True, precision: Precision = Precision.MIXED_FP32,
                 smooth_sweeps: int = 2, memory_budget_bytes: int = 0, staging_threads: int = 4):
        o = L.ServerOptions()
        L.call("cfdg_server_default_options", C.byref(o))
        o.max_active_solves = int(max_active_solves)
        o.method = int(method)
        o.use_amg = int(bool(use_amg))
        o.precision = int(precision)
        o.smooth_sweeps = int(smooth_sweeps)
        o.memory_budget_bytes = int(memory_budget_bytes)
        o.staging_threads = int(staging_threads)
        h = C.c_void_p()
        L.call("cfdg_server_create", C.byref(o), C.byref(h))
        self._h = h
        self._open = {}

    def submit(self, pattern: LduPattern, diag, upper, lower, amg: Optional[AmgView],
               rhs: Sequence[np.ndarray], rel_tolerance: float = 1e-8, max_iterations: int = 500,
               x0: Optional[np.ndarray] = None, keep_solutions: bool = True) -> ServedRequest:
        """Queues a request; the arrays are kept referenced until wait() returns for it."""
        n, nf = pattern.n, pattern.n_faces
        diag, upper, lower = _f64(diag), _f64(upper), _f64(lower)
        _check_vec(diag, n, "diag")
        _check_vec(upper, nf, "upper")
        _check_vec(lower, nf, "lower")
        rhs = [_f64(b) for b in rhs]
        for b in rhs:
            _check_vec(b, n, "rhs")
        steps = len(rhs)
        req = L.ServerRequest()
        req.pattern = pattern.c()
        req.diag, req.upper, req.lower = L.dptr(diag), L.dptr(upper), L.dptr(lower)
        view_keep = None
        if amg is not None:
            v, view_keep = amg.c()
            vbox = C.pointer(v)
            req.amg = vbox
        else:
            vbox = None
        rhs_arr = (C.POINTER(C.c_double) * max(steps, 1))(*[L.dptr(b) for b in rhs])
        req.n_steps = steps
        req.rhs = C.cast(rhs_arr, C.POINTER(C.POINTER(C.c_double)))
        if x0 is not None:
            x0 = _f64(x0)
            _check_vec(x0, n, "x0")
            req.x0 = L.dptr(x0)
        req.rel_tolerance = float(rel_tolerance)
        req.max_iterations = int(max_iterations)
        sols = [np.empty(n) for _ in range(steps)] if keep_solutions else None
        xo = None
        if sols is not None:
            xo = (C.POINTER(C.c_double) * steps)(*[L.dptr(x) for x in sols])
            req.x_out = C.cast(xo, C.POINTER(C.POINTER(C.c_double)))
        stats = (L.SolveStats * steps)()
        lat = np.zeros(steps)
        req.stats_out = C.cast(stats, C.POINTER(L.SolveStats))
        req.step_latency_ms = L.dptr(lat)
        rid = C.c_ulonglong()
        keep = (pattern, diag, upper, lower, view_keep, vbox, rhs, rhs_arr, x0, sols, xo, stats, lat, req)
        L.call("cfdg_server_submit", self._h, C.byref(req), C.byref(rid))
        r = ServedRequest(int(rid.value), steps, step_latency_ms=lat, solutions=sols, _keep=keep)
        self._open[r.id] = r
        return r

    def wait(self, r: ServedRequest) -> ServedRequest:
        """Blocks until the request has completed; raises its error if it failed."""
        lat = C.c_double()
        try:
            L.call("cfdg_server_wait", self._h, C.c_ulonglong(r.id), C.byref(lat))
        finally:
            self._open.pop(r.id, None)
        r.latency_ms = lat.value
        stats = r._keep[11]
        r.stats = [SolveStats.from_c(s) for s in stats]
        r._keep = ()
        return r

    def counters(self) -> ServerCounters:
        c = L.ServerCounters()
        L.call("cfdg_server_query", self._h, C.byref(c))
        return ServerCounters(c.submitted, c.completed, c.solves, c.resident_bytes,
                              c.peak_resident_bytes, c.budget_bytes, c.active_solves,
                              c.peak_active_solves, c.waiting, c.staging_ms, c.retiring_ms,
                              c.engine_busy_ms)

    def close(self) -> None:
        """Completes every submitted request, then stops the server."""
        if getattr(self, "_h", None):
            h, self._h = self._h, None
            L.call("cfdg_server_destroy", h)
            self._open.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
