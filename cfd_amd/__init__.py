"""MI355X (gfx950) linear-solver backend for the Yipexnn/CFD finite-volume solver.

The compute path is libcfdgpu.so (C++ host code + HIP kernels, C-ABI in include/cfdgpu.h);
this package is its Python mirror of the reference's cfd::gpu interface. Importing it loads
the library and fails loudly if it has not been built (python cfd_amd/build.py).
"""
from . import _lib

_lib.lib()  # fail at import time when the native library is missing

from .amg import AmgHierarchy, AmgSettings  # noqa: E402
from .fv import DeviceFv, FvMesh, PressureEquation, PressureInputs, fv_box, synthetic_pressure_inputs  # noqa: E402
from .server import ServedRequest, Server, ServerCounters, estimate_footprint  # noqa: E402
from .krylov import (AmgLevelView, AmgView, Batch, CsrView, DeviceSolver, LduPattern, Method,  # noqa: E402
                     Precision, SolveControl, SolveStats, assemble_convection_diffusion,
                     backend_name, device_available, device_description, solve, solve_amg_cg)
from .ldu import AmgCache, LinearSolveResult, LinearSolverSettings  # noqa: E402


def _apply_tuning() -> None:
    """Installs cfd_amd/tuning.json (written by python -m cfd_amd.autotune on the target GPU)
    unless CFDG_NO_TUNING is set."""
    import json
    import os
    from pathlib import Path

    path = Path(__file__).resolve().parent / "tuning.json"
    if os.environ.get("CFDG_NO_TUNING") or not path.exists():
        _lib.call("cfdg_set_layout_rules", None, 0)
        return
    from .autotune import apply

    apply(json.loads(path.read_text()))


_apply_tuning()

__all__ = [
    "AmgHierarchy", "AmgSettings", "AmgLevelView", "AmgView", "Batch", "CsrView", "DeviceSolver",
    "LduPattern", "Method", "Precision", "SolveControl", "SolveStats",
    "assemble_convection_diffusion", "backend_name", "device_available", "device_description",
    "solve", "solve_amg_cg", "Server", "ServedRequest", "ServerCounters", "estimate_footprint",
    "DeviceFv", "FvMesh", "fv_box", "PressureInputs", "PressureEquation", "synthetic_pressure_inputs",
    "AmgCache", "LinearSolverSettings", "LinearSolveResult",
]
