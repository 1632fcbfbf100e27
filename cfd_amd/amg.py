"""Smoothed-aggregation AMG hierarchy built by libcfdgpu on the host.

Counterpart of fv::AmgPreconditioner (src/fv/include/cfd/fv/amg.hpp:17-69, setup
src/fv/src/amg.cpp:122-190): the same aggregation, Jacobi-smoothed prolongation, Galerkin
coarse operators and exact-zero pruning, in C++ without Eigen.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List

import numpy as np

from . import _lib as L
from .krylov import AmgLevelView, AmgView, CsrView


@dataclass
class AmgSettings:
    """AmgSettings (amg.hpp:17-23) plus the V-cycle parameters of AmgView (krylov.hpp:56-57)."""

    strength_threshold: float = 0.08
    max_levels: int = 25
    coarse_size: int = 400
    smooth_sweeps: int = 2
    jacobi_weight: float = 0.7
    strength_decay: float = 1.0  # 1.0 reference (fixed threshold); 0.5 Vanek et al. (1996)

    def c(self) -> L.AmgSettings:
        return L.AmgSettings(self.strength_threshold, self.max_levels, self.coarse_size,
                             self.smooth_sweeps, self.jacobi_weight, self.strength_decay)


class AmgHierarchy:
    """Owns a cfdg_amg handle; exports levels as CsrView copies."""

    def __init__(self, a: CsrView, settings: AmgSettings | None = None, device: bool = False):
        """device: the sparse products of the setup run on the GPU (cfdg_amg_build_device; the
        hierarchy is bit-identical to the host setup's), the aggregation on the host."""
        self.settings = settings or AmgSettings()
        h = C.c_void_p()
        ca = a.c()
        cs = self.settings.c()
        L.call("cfdg_amg_build_device" if device else "cfdg_amg_build", C.byref(ca), C.byref(cs), C.byref(h))
        self._h = h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                L.lib().cfdg_amg_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def n_levels(self) -> int:
        v = C.c_int()
        L.call("cfdg_amg_n_levels", self._h, C.byref(v))
        return v.value

    def level_size(self, level: int, which: int = 0):
        r, c, z = C.c_int(), C.c_int(), C.c_int()
        L.call("cfdg_amg_level_size", self._h, level, which, C.byref(r), C.byref(c), C.byref(z))
        return r.value, c.value, z.value

    def matrix(self, level: int, which: int = 0) -> CsrView:
        """which: 0 operator, 1 prolongation, 2 restriction."""
        rows, cols, nnz = self.level_size(level, which)
        rp = np.zeros(rows + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        va = np.zeros(max(nnz, 1), np.float64)
        L.call("cfdg_amg_level_copy", self._h, level, which, L.iptr(rp), L.iptr(ci), L.dptr(va))
        return CsrView(rows, nnz, rp, ci[:nnz], va[:nnz], cols)

    def aggregates(self, level: int) -> np.ndarray:
        rows, _, _ = self.level_size(level, 0)
        agg = np.zeros(rows, np.int32)
        L.call("cfdg_amg_aggregates", self._h, level, L.iptr(agg))
        return agg

    def coarse_inverse(self) -> np.ndarray:
        nc, _, _ = self.level_size(self.n_levels - 1, 0)
        out = np.zeros(nc * nc, np.float64)
        L.call("cfdg_amg_coarse_inverse", self._h, L.dptr(out))
        return out.reshape(nc, nc)

    @property
    def coarse_on_device(self) -> bool:
        """True when the coarsest inverse was computed on the matrix cores (device setup)."""
        v = C.c_int()
        L.call("cfdg_amg_coarse_on_device", self._h, C.byref(v))
        return bool(v.value)

    @property
    def setup_ms(self) -> float:
        v = C.c_double()
        L.call("cfdg_amg_setup_ms", self._h, C.byref(v))
        return v.value

    def level_sizes(self) -> List[int]:
        return [self.level_size(lvl, 0)[0] for lvl in range(self.n_levels)]

    def operator_complexity(self) -> float:
        """Sum of non-zeros of all levels / non-zeros of the finest (amg.cpp:242-251)."""
        nnz = [self.level_size(lvl, 0)[2] for lvl in range(self.n_levels)]
        return float(sum(nnz)) / float(nnz[0])

    def view(self) -> AmgView:
        """AmgView of copies of every level (safe to keep after the hierarchy is freed)."""
        nl = self.n_levels
        levels = []
        for lvl in range(nl):
            lv = AmgLevelView(a=self.matrix(lvl, 0))
            if lvl + 1 < nl:
                lv.p = self.matrix(lvl, 1)
                lv.r = self.matrix(lvl, 2)
            levels.append(lv)
        return AmgView(levels, self.coarse_inverse(), self.settings.smooth_sweeps,
                       self.settings.jacobi_weight)


def dense_inverse_batch(mats):
    """cfdg_dense_inverse_batch: inverts square float64 matrices on the device (block
    Gauss-Jordan on the FP64 matrix cores, orders up to 512), the host path for the rest. Returns
    (inverses, on_device flags)."""
    mats = [np.ascontiguousarray(m, np.float64) for m in mats]
    for m in mats:
        if m.ndim != 2 or m.shape[0] != m.shape[1]:
            raise L.InvalidArgument(1, "dense_inverse_batch: square matrices required")
    count = len(mats)
    outs = [np.zeros_like(m) for m in mats]
    order = (C.c_int * max(count, 1))(*[m.shape[0] for m in mats])
    a = (C.c_void_p * max(count, 1))(*[m.ctypes.data for m in mats])
    inv = (C.c_void_p * max(count, 1))(*[o.ctypes.data for o in outs])
    flags = (C.c_int * max(count, 1))()
    L.call("cfdg_dense_inverse_batch", count, order, a, inv, flags)
    return outs, [bool(flags[i]) for i in range(count)]
