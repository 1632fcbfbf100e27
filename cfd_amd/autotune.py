"""Layout and smoother autotuner for libcfdgpu (run on the target GPU).

    python -m cfd_amd.autotune [--out cfd_amd/tuning.json]

1. Sparse layouts. It builds the smoothed-aggregation hierarchies of a set of box meshes
   (cubes and elongated cells, uniform and jittered, the shapes of the benchmark workloads),
   times y = A x for every coarse operator, prolongation and restriction in every candidate
   layout (lanes per row 1..64, sliced ELL or CSR; FP32 plus FP64 time) with cfdg_spmv_time,
   and writes a rule for a matrix class (rows within x1.5, mean row length within x1.25) where
   the best layout beats the library's built-in choice by more than 10%.
2. Mixed-precision smoother. It times resident mixed-precision solves of a jittered pressure
   system (the 128^3 benchmark box) and systems of the batch workload's shapes for Chebyshev
   smoothers of degree 1 to 3
   per side (degree k: k damped-Jacobi sweeps weighted by the reciprocals of the Chebyshev
   polynomial's roots over [h / ratio, h]) and several interval ratios, and keeps the fastest.

The result is a JSON table; `cfd_amd` applies it at import (cfdg_set_layout_rules,
cfdg_set_mixed_smoother) unless CFDG_NO_TUNING is set. A table changes which layouts and which
smoother are used, so results are reproducible for a given table.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time
from pathlib import Path

import numpy as np

from . import _lib as L

DEFAULT_PATH = Path(__file__).resolve().parent / "tuning.json"


def time_layout(m, lanes: int, csr: int, fp32: int, reps: int = 20) -> float:
    a = m.c()
    ms = C.c_double()
    L.call("cfdg_spmv_time", C.byref(a), lanes, csr, fp32, reps, C.byref(ms))
    return ms.value


def default_layout(m):
    lanes, csr = C.c_int(), C.c_int()
    L.call("cfdg_layout_of", m.n, L.iptr(m.row_ptr), C.byref(lanes), C.byref(csr))
    return lanes.value, csr.value


SHAPES = [  # (nx, ny, nz, cell aspect, jitter)
    (128, 128, 128, 1.0, 0.2), (96, 96, 96, 1.0, 0.0), (64, 64, 64, 1.0, 0.2), (40, 40, 40, 1.0, 0.0),
    (24, 24, 24, 1.0, 0.2), (48, 48, 150, 4.0, 0.0), (30, 30, 120, 8.0, 0.2), (80, 80, 40, 2.0, 0.2),
]


def hierarchy_matrices():
    from . import AmgHierarchy, AmgSettings, cases

    out = []
    for nx, ny, nz, r, jit in SHAPES:
        mesh = cases.box_mesh(nx, ny, nz, hi=(1.0, 1.0, nz * r / nx), jitter=jit, seed=nx + nz)
        s = cases.pressure_system(mesh, rhs_seed=1)
        pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
        h = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5))
        view = h.view()
        last = len(view.levels) - 1
        for l, lv in enumerate(view.levels):
            if l >= 1 and l < last:
                out.append((f"{nx}x{ny}x{nz} A{l}", lv.a))
            if l < last:
                out.append((f"{nx}x{ny}x{nz} P{l}", lv.p))
                out.append((f"{nx}x{ny}x{nz} R{l}", lv.r))
    return out


def tune_layouts():
    rules, log = [], []
    for name, m in hierarchy_matrices():
        nnz = int(m.row_ptr[-1])
        mean = nnz / max(1, m.n)
        d_lanes, d_csr = default_layout(m)
        cand = {}
        for lanes in (1, 2, 4, 8, 16, 32, 64):
            if lanes > 2 * mean and lanes > 1:
                break
            for csr in (0, 1):
                t = time_layout(m, lanes, csr, 1) + time_layout(m, lanes, csr, 0)
                cand[(lanes, csr)] = t
        if (d_lanes, d_csr) not in cand:
            cand[(d_lanes, d_csr)] = time_layout(m, d_lanes, d_csr, 1) + time_layout(m, d_lanes, d_csr, 0)
        best = min(cand, key=cand.get)
        t_def, t_best = cand[(d_lanes, d_csr)], cand[best]
        log.append({"matrix": name, "rows": m.n, "mean": mean, "default": [d_lanes, d_csr], "default_ms": t_def,
                    "best": list(best), "best_ms": t_best})
        print(f"{name:22s} rows {m.n:8d} mean {mean:7.1f}: built-in {d_lanes:2d}/{'csr' if d_csr else 'ell'} "
              f"{t_def * 1e3:8.1f} us, best {best[0]:2d}/{'csr' if best[1] else 'ell'} {t_best * 1e3:8.1f} us",
              flush=True)
        if t_best < 0.9 * t_def:
            rules.append({"min_rows": int(m.n / 1.5), "max_rows": int(m.n * 1.5), "min_mean": mean / 1.25,
                          "max_mean": mean * 1.25, "lanes": best[0], "csr": best[1], "gain": t_def / t_best})
    # larger gains first, so that overlapping classes take the more certain rule
    rules.sort(key=lambda r: -r["gain"])
    return rules, log


def tune_smoother(n_side: int = 128, reps: int = 3, sample: int = 12):
    """Mixed-precision smoother: Chebyshev degree per side (the sweeps of cfdg_set_mixed_smoother)
    and the ratio of the damping interval, timed as the total mixed solve time of one n_side^3
    jittered box and `sample` systems of the batch workload's shapes (cases.batch_shapes)."""
    from . import AmgHierarchy, AmgSettings, DeviceSolver, Method, Precision, SolveControl, cases

    systems = []
    shapes = [dict(nx=n_side, ny=n_side, nz=n_side, lz=1.0, jitter=0.2, seed=7)]
    shapes += cases.batch_shapes(256)[::max(1, 256 // sample)][:sample]
    for sh in shapes:
        mesh = cases.box_mesh(sh["nx"], sh["ny"], sh["nz"], hi=(1.0, 1.0, sh["lz"]), jitter=sh["jitter"],
                              seed=sh["seed"])
        s = cases.pressure_system(mesh, rhs_seed=sh["seed"] + 1)
        pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
        a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
        systems.append((s, pat, AmgHierarchy(a, AmgSettings(strength_decay=0.5)).view()))
    result = {}
    for sweeps, ratio in ((1, 8.0), (1, 6.0), (1, 5.0), (2, 8.0), (2, 5.0), (2, 4.0), (2, 3.0), (3, 12.0),
                          (3, 8.0)):
        L.call("cfdg_set_mixed_smoother", sweeps, ratio)
        total = 0.0
        its = 0
        for s, pat, view in systems:
            dev = DeviceSolver(pat)
            dev.assemble(s.diag, s.upper, s.lower)
            dev.set_amg(view)  # the weights follow the smoother set before
            dev.set_rhs(s.source)
            tol = 1e-8 * float(np.linalg.norm(s.source))
            times = []
            for _ in range(reps + 1):
                dev.set_x(None)
                st = dev.solve_resident(Method.Cg, True, Precision.MIXED_FP32, SolveControl(tol, 500))
                times.append(st.solve_ms)
            total += float(np.median(times[1:]))
            its += st.iterations
            dev.close()
        result[(sweeps, ratio)] = (total, its)
        print(f"mixed smoother, Chebyshev degree {sweeps}, ratio {ratio:4.1f}: {total:.3f} ms, {its} iterations in all",
              flush=True)
    best = min(result, key=lambda k: result[k][0])
    return {"sweeps": best[0], "ratio": best[1],
            "timings_ms": {f"{k[0]}/{k[1]:g}": round(v[0], 4) for k, v in result.items()},
            "iterations": {f"{k[0]}/{k[1]:g}": v[1] for k, v in result.items()}}


def apply(table: dict) -> None:
    """Installs a tuning table in the library (process-wide)."""
    rules = table.get("layout_rules", [])
    arr = (L.LayoutRule * max(1, len(rules)))()
    for i, r in enumerate(rules):
        arr[i] = L.LayoutRule(int(r["min_rows"]), int(r["max_rows"]), float(r["min_mean"]),
                              float(r["max_mean"]), int(r["lanes"]), int(r["csr"]))
    L.call("cfdg_set_layout_rules", arr, len(rules))
    sm = table.get("mixed_smoother")
    # the environment's smoother settings, where given, take precedence over the table's
    if sm and not (os.environ.get("CFDG_MIXED_SWEEPS") or os.environ.get("CFDG_MIXED_RATIO")):
        L.call("cfdg_set_mixed_smoother", int(sm["sweeps"]), float(sm["ratio"]))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    ap.add_argument("--smoother-only", action="store_true",
                    help="re-tune the mixed smoother only and keep the table's layout rules")
    args = ap.parse_args(argv)
    if args.smoother_only:
        table = json.loads(Path(args.out).read_text()) if Path(args.out).exists() else {"layout_rules": []}
        apply({"layout_rules": table.get("layout_rules", [])})
        table["mixed_smoother"] = tune_smoother()
        Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
        print(f"wrote {args.out}")
        return 0
    L.call("cfdg_set_layout_rules", None, 0)  # compare with the built-in rule
    t0 = time.time()
    rules, log = tune_layouts()
    smoother = tune_smoother()
    buf = C.create_string_buffer(256)
    L.lib().cfdg_device_description(buf, len(buf))
    table = {"device": buf.value.decode(), "layout_rules": rules, "mixed_smoother": smoother,
             "measurements": log, "seconds": round(time.time() - t0, 1)}
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    print(f"wrote {args.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
