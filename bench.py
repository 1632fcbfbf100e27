"""Benchmark: pressure solves/s to 1e-8 relative residual on one MI355X.

Workload (one step = one pressure solve from x0 = 0): the 128^3 jittered hex box
(2,097,152 cells, interior vertices moved by +-20% of the spacing), pressure Poisson system with
coefficients |S_f| / |C_N - C_P|, zero-gradient walls, pressure pinned at cell 0, seeded
zero-mean right-hand side; smoothed-aggregation AMG (threshold 0.08 halved per level) built once
on the host; solved to |b - A x| <= 1e-8 |b| with
  * value: mixed precision - FP32 Jacobi V-cycle inside FP64 flexible CG (the metric's "mixed"),
  * beside it: the plain FP64 AMG-PCG (the reference algorithm) and the FP64 Jacobi-BiCGStab
    momentum system of the same mesh.
Inputs are resident in HBM before the timed region. The CPU baseline is the reference's own host
AMG-PCG (solve_amg_cg_host, compiled from the reference sources into oracle/_ref) on the same
hierarchy and right-hand side, timed on this machine's host cores for a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 128] [--skip-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pressure solves/s @1e-8 rel-res, mixed vs FP64 AMG-PCG; HBM GB/s; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(hier_view, b, tol, budget_s):
    """Times the reference's solve_amg_cg_host (oracle/_ref) or, where that library was not
    built, the oracle's restatement, on the host cores for about budget_s seconds."""
    from oracle import bind as ob

    levels = hier_view.levels
    a_l = [ob.Csr(lv.a.n, lv.a.n, lv.a.row_ptr, lv.a.col, lv.a.val) for lv in levels]
    p_l = [None if lv.p is None else ob.Csr(lv.p.n, lv.p.n_cols, lv.p.row_ptr, lv.p.col, lv.p.val)
           for lv in levels]
    r_l = [None if lv.r is None else ob.Csr(lv.r.n, lv.r.n_cols, lv.r.row_ptr, lv.r.col, lv.r.val)
           for lv in levels]
    ci = hier_view.coarse_inverse
    cores = os.cpu_count() or 1
    ob.set_threads(cores)
    kind = "reference" if ob.ref_available() else "port"
    times, iters = [], []
    t_start = time.perf_counter()
    while True:
        x = np.zeros(b.shape[0])
        t0 = time.perf_counter()
        if kind == "reference":
            st = ob.ref_solve_amg_cg_host(a_l, p_l, r_l, ci, b, x, tol, 500)
        else:
            arrs = ob._LevelArrays(a_l, p_l, r_l)
            cinv = np.ascontiguousarray(ci, np.float64).reshape(-1)
            res = ob.orc().orc_amg_cg(len(a_l), ob.ip(arrs.n), *arrs.a, *arrs.p, *arrs.r,
                                      ob.dp(cinv), 2, 0.7, ob.dp(b), ob.dp(x), tol, 500)
            st = ob.Stats(res.iterations, bool(res.converged), res.residual)
        times.append(time.perf_counter() - t0)
        iters.append(st.iterations)
        if time.perf_counter() - t_start > budget_s:
            break
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "solves/s", "cores": cores, "kind": kind,
            "sample": f"{len(times)} FP64 Jacobi-V-cycle AMG-PCG solves of the same 128^3 system "
                      f"(same hierarchy, x0 = 0, 1e-8), median {med * 1e3:.1f} ms, "
                      f"{int(np.median(iters))} iterations; reference threads loops with n > 20000"}


def batched(steps: int, count: int = 256) -> dict:
    """Config of 256 independent pressure systems (cell counts log-uniform in [4096, 500000],
    half uniform and half jittered boxes, cell aspect 1:1:1 to 1:1:8, seeded zero-mean right-hand
    sides) solved together to 1e-8: one launch per solver stage covers every system. Reported:
    solves/s of the mixed (FP32 V-cycle, FP64 flexible CG) batch, of the FP64 batch, and of the
    plain FP64 path solving the same systems one after another; the largest relative difference
    between each mixed solution and its FP64 solution."""
    import cfd_amd
    from cfd_amd import AmgHierarchy, AmgSettings, DeviceSolver, Method, Precision, SolveControl, cases

    t0 = time.perf_counter()
    members, ctls = [], []
    cells = 0
    for s in cases.batch_shapes(count):
        mesh = cases.box_mesh(s["nx"], s["ny"], s["nz"], hi=(1.0, 1.0, s["lz"]),
                              jitter=s["jitter"], seed=s["seed"])
        p = cases.pressure_system(mesh, rhs_seed=s["seed"] + 1)
        pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
        a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
        dev = DeviceSolver(pat)
        dev.assemble(p.diag, p.upper, p.lower)
        dev.set_amg(AmgHierarchy(a, AmgSettings(strength_decay=0.5)).view())
        dev.set_rhs(p.source)
        members.append(dev)
        ctls.append(SolveControl(1e-8 * float(np.linalg.norm(p.source)), 500))
        cells += p.n
    log(f"batch: {count} systems, {cells} cells, set up in {time.perf_counter() - t0:.1f} s")

    def timed_batch(prec):
        batch = cfd_amd.Batch(members, Method.Cg, True, prec)
        for rep in range(1 + steps):  # one untimed solve (graph capture), then the timed ones
            if rep == 1:
                t = time.perf_counter()
            for d in members:
                d.set_x(None)
            st = batch.solve(ctls)
        dt = time.perf_counter() - t
        if not all(x.converged for x in st):
            raise RuntimeError(f"batched {prec.name}: a system did not converge")
        xs = [d.get_x() for d in members]
        batch.close()
        return count * steps / dt, dt / steps, st, xs

    sps_m, t_m, st_m, x_m = timed_batch(Precision.MIXED_FP32)
    sps_d, t_d, st_d, x_d = timed_batch(Precision.FP64)
    sps_b, t_b, st_b, x_b = timed_batch(Precision.MIXED_BF16)
    # Plain FP64 path: the same systems, one resident solve after another.
    for d, c in zip(members, ctls):
        d.set_x(None)
        d.solve_resident(Method.Cg, True, Precision.FP64, c)
    t = time.perf_counter()
    for d, c in zip(members, ctls):
        d.set_x(None)
        d.solve_resident(Method.Cg, True, Precision.FP64, c)
    t_seq = time.perf_counter() - t
    err = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(x_m, x_d))
    err_b = max(float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(x_b, x_d))
    footprint = sum(d.footprint() for d in members)
    for d in members:
        d.close()
    its = lambda st: [int(min(x.iterations for x in st)), int(max(x.iterations for x in st))]
    return {
        "workload": f"{count} pressure systems, {cells} cells in all, log-uniform 4096..500000 "
                    "cells, SA-AMG-PCG to 1e-8 rel. residual from x0 = 0",
        "mixed_solves_per_s": round(sps_m, 2),
        "fp64_batched_solves_per_s": round(sps_d, 2),
        "fp64_one_by_one_solves_per_s": round(count / t_seq, 2),
        "bf16_batched_solves_per_s": round(sps_b, 2),
        "speedup_mixed_vs_fp64_batched": round(sps_m / sps_d, 3),
        "speedup_mixed_vs_fp64_one_by_one": round(sps_m * t_seq / count, 3),
        "ms_per_batch_mixed": round(1e3 * t_m, 2),
        "ms_per_batch_fp64": round(1e3 * t_d, 2),
        "iterations_mixed_min_max": its(st_m),
        "iterations_fp64_min_max": its(st_d),
        "max_true_rel_residual_mixed": max(x.true_residual / x.rhs_norm for x in st_m),
        "max_rel_error_mixed_vs_fp64": err,
        "max_rel_error_bf16_vs_fp64": err_b,
        "footprint_bytes": footprint,
    }


def stress(n_side: int = 160, reps: int = 3) -> dict:
    """Precision stress: 160^3 jittered box, rho = 1000 inside a seeded sphere of radius 0.2 and
    1 outside (face coefficients with harmonic averaging), solved to 1e-10 relative residual by
    the FP64 AMG-PCG and by the mixed solvers (FP32 V-cycle, optionally BF16 level-0
    off-diagonals, inside FP64 flexible CG). Reported per solver: iterations, true residual,
    relative error against the FP64 solution, solves/s."""
    import cfd_amd
    from cfd_amd import AmgHierarchy, AmgSettings, DeviceSolver, Method, Precision, SolveControl, cases

    t0 = time.perf_counter()
    mesh, p = cases.config3(n_side)
    pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
    a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    dev = DeviceSolver(pat)
    dev.assemble(p.diag, p.upper, p.lower)
    dev.set_amg(h.view())
    dev.set_rhs(p.source)
    log(f"stress: {p.n} cells, levels {h.level_sizes()}, set up in {time.perf_counter() - t0:.1f} s")
    bnorm = float(np.linalg.norm(p.source))
    ctl = SolveControl(1e-10 * bnorm, 1000)
    out, x64 = {}, None
    for prec in (Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16):
        times = []
        for _ in range(reps + 1):
            dev.set_x(None)
            t = time.perf_counter()
            st = dev.solve_resident(Method.Cg, True, prec, ctl)
            times.append(time.perf_counter() - t)
        x = dev.get_x()
        if prec == Precision.FP64:
            x64 = x
        out[prec.name.lower()] = {
            "iterations": st.iterations, "converged": st.converged,
            "true_rel_residual": st.true_residual / bnorm,
            "rel_error_vs_fp64": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
            "solves_per_s": round(1.0 / float(np.median(times[1:])), 3)}
    # Level-0 FP32 sweep at a working set larger than the 256 MB Infinity Cache.
    dev.probe(1)
    for _ in range(2):
        dev.set_x(None)
        dev.solve_resident(Method.Cg, True, Precision.MIXED_FP32, ctl)
    ms, launches, nbytes = dev.probe_read(Precision.MIXED_FP32)
    dev.probe(0)
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out["roofline_l0_fp32_sweep"] = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                                     "bytes_per_launch": int(nbytes), "avg_us": round(ms * 1e3, 3),
                                     "launches": int(launches),
                                     "working_set_mb": round(dev.footprint() / 1e6, 1)}
    out["workload"] = (f"{n_side}^3 jittered box ({p.n} cells), rho 1000 in a sphere of radius "
                       "0.2 / 1 outside, harmonic faces, SA-AMG-PCG to 1e-10")
    return out


def served(count: int = 24, rate: float = 200.0, seed: int = 20240607) -> dict:
    """Serving workload: requests with Poisson arrivals (rate per second, seeded), each a
    jittered box of log-uniform 10,000..4,000,000 cells with a run of 1..20 time-step right-hand
    sides (each the previous one plus 5% seeded noise), solved to 1e-8 from the previous step's
    solution with the hierarchy kept resident. The same arrival sequence is replayed against
    the mixed-precision server (continuous batching), the FP64 server, and the plain FP64 path
    (one solve at a time). Reported: solves/s over the makespan and p50 / p99 latencies."""
    import threading

    import cfd_amd
    from cfd_amd import AmgHierarchy, AmgSettings, Precision, cases

    t0 = time.perf_counter()
    gen, arrivals = cases.serving_requests(count, rate=rate, seed=seed)
    reqs, cells, solves = [], 0, 0
    for g in gen:
        s = g["system"]
        view = AmgHierarchy(g["csr"], AmgSettings(strength_decay=0.5)).view()
        reqs.append(dict(pattern=g["pattern"], diag=s.diag, upper=s.upper, lower=s.lower, amg=view,
                         rhs=g["rhs"]))
        cells += s.n
        solves += len(g["rhs"])
    log(f"serve: {count} requests, {cells} cells, {solves} solves, set up in "
        f"{time.perf_counter() - t0:.1f} s")

    def replay(prec, slots):
        with cfd_amd.Server(max_active_solves=slots, precision=prec) as srv:
            handles = [None] * count

            def submitter():
                start = time.perf_counter()
                for i, q in enumerate(reqs):
                    delay = arrivals[i] - (time.perf_counter() - start)
                    if delay > 0:
                        time.sleep(delay)
                    handles[i] = srv.submit(**q, keep_solutions=False)

            t = time.perf_counter()
            th = threading.Thread(target=submitter)
            th.start()
            th.join()
            done = [srv.wait(h) for h in handles]
            span = time.perf_counter() - t
            c = srv.counters()
        lat = np.concatenate([d.step_latency_ms for d in done])
        req_lat = np.array([d.latency_ms for d in done])
        ok = all(st.converged for d in done for st in d.stats)
        worst = max(st.true_residual / st.rhs_norm for d in done for st in d.stats)
        return {"solves_per_s": round(solves / span, 2), "makespan_s": round(span, 3),
                "p50_solve_ms": round(float(np.percentile(lat, 50)), 3),
                "p99_solve_ms": round(float(np.percentile(lat, 99)), 3),
                "p50_request_ms": round(float(np.percentile(req_lat, 50)), 3),
                "p99_request_ms": round(float(np.percentile(req_lat, 99)), 3),
                "all_converged": ok, "max_true_rel_residual": worst,
                "peak_active_solves": c.peak_active_solves,
                "peak_resident_gb": round(c.peak_resident_bytes / 1e9, 2),
                "staging_thread_s": round(c.staging_ms / 1e3, 3), "retiring_thread_s": round(c.retiring_ms / 1e3, 3),
                "engine_busy_s": round(c.engine_busy_ms / 1e3, 3)}

    mixed = replay(Precision.MIXED_FP32, 64)
    fp64 = replay(Precision.FP64, 64)
    plain = replay(Precision.FP64, 1)
    return {"workload": f"{count} requests, Poisson arrivals at {rate:g}/s, {cells} cells, "
                        f"{solves} time-step solves to 1e-8",
            "mixed_served": mixed, "fp64_served": fp64, "fp64_one_at_a_time": plain,
            "speedup_mixed_vs_fp64_served": round(mixed["solves_per_s"] / fp64["solves_per_s"], 3),
            "speedup_mixed_vs_fp64_one_at_a_time":
                round(mixed["solves_per_s"] / plain["solves_per_s"], 3)}


def fv_bytes(m, op: str) -> float:
    """Algorithmic bytes of one finite-volume operation: every array element it needs, read or
    written once (the neighbour values a cell gathers are counted once, with their array)."""
    nc, nif, nf = m.n_cells, m.n_internal_faces, m.n_faces
    nb = nf - nif
    lists = 4 * (nc + 1) + 4 * (2 * nif + nb)          # cell_offsets + cell_faces
    inner = 4 * (nc + 1) + 4 * (2 * nif + nc)          # lists read up to the first boundary face
    own_nei = 8 * nif

    def gauss(k):                                        # k components
        return lists + own_nei + 8 * nif + 24 * nf + 8 * k * (nc + nb) + 8 * nc + 24 * k * nc

    def lsq(k):
        return lists + own_nei + 24 * nf + 8 * k * (nc + nb) + 72 * nc + 24 * k * nc

    def corrected(k):
        second = lists + own_nei + 8 * nif + 24 * nf + 24 * nif + 24 * nc + 8 * k * (nc + nb) + 24 * k * nc \
            + 8 * nc + 24 * k * nc
        return lsq(k) + second

    grad = {"gradient_gauss": gauss(1), "gradient_least_squares": lsq(1), "gradient_corrected": corrected(1),
            "vector_gradient_least_squares": lsq(3), "vector_gradient_corrected": corrected(3)}
    if op in grad:
        return float(grad[op])
    if op == "scalar_corrections":
        corr = inner + own_nei + 8 * nif + 24 * nif + 24 * nc + 24 * nc + 8 * nif + 8 * nif + 24 * nif + 16 * nc
        return float(lsq(1) + corr)
    if op == "momentum_corrections":
        limiter = lists + own_nei + 24 * nc + 24 * nb + 72 * nc + 24 * nf + 24 * nc + 24 * nc
        mom = inner + own_nei + 8 * nif + 24 * nif + 24 * nc + 72 * nc + 24 * nc + 24 * nc + 8 * nif + 8 * nif \
            + 24 * nif + 24 * nif + 48 * nc
        return float(limiter + mom)
    if op == "pressure_equation":  # with the relaxation-independent flux term
        hbya = inner + own_nei + 16 * nif + 24 * nc + 48 * nc + 8 * nc + 32 * nc
        press = lists + own_nei + 8 * nif + 24 * nf + 8 * nf + 32 * nc + 36 * nb + 24 * nc + 8 * nf + 16 * nf \
            + 16 * nif + 16 * nc
        return float(hbya + press)
    raise KeyError(op)


def fv_section(n_side: int = 128, reps: int = 10, cpu: bool = True) -> dict:
    """Finite-volume cell kernels (cfdg_fv_*) on an n_side^3 jittered hex box: device time of each
    operation's kernels (HIP events in the library), algorithmic GB/s against the 8 TB/s HBM
    roofline, the PCIe-inclusive time of the call, and the reference's gpu::DeviceFv host path on
    the same mesh (oracle/_ref, all host cores) as the CPU baseline."""
    from cfd_amd.fv import DeviceFv, fv_box, synthetic_pressure_inputs

    t0 = time.perf_counter()
    mesh = fv_box(n_side, n_side, n_side, jitter=0.2, seed=11)
    rng = np.random.default_rng(5)
    nc, nb, nif = mesh.n_cells, mesh.n_boundary, mesh.n_internal_faces
    phi, bv = rng.standard_normal(nc), rng.standard_normal(nb)
    u, ub = rng.standard_normal((nc, 3)), rng.standard_normal((nb, 3))
    flux, nu = rng.standard_normal(nif), rng.uniform(0.01, 0.2, nif)
    src, src3 = np.zeros(nc), np.zeros((nc, 3))
    eps2 = np.array([1e-3, 1e-3, 1e-3])
    log(f"fv: {nc} cells, {mesh.n_faces} faces, mesh built in {time.perf_counter() - t0:.1f} s")
    ops = {
        "gradient_gauss": lambda d: d.gradient(phi, bv, 0),
        "gradient_least_squares": lambda d: d.gradient(phi, bv, 1),
        "gradient_corrected": lambda d: d.gradient(phi, bv, 2),
        "vector_gradient_least_squares": lambda d: d.vector_gradient(u, ub, 1),
        "vector_gradient_corrected": lambda d: d.vector_gradient(u, ub, 2),
        "scalar_corrections": lambda d: d.scalar_corrections(phi, bv, 1, flux, nu, True, src),
        "momentum_corrections": lambda d: d.momentum_corrections(flux, nu, True, True, False, eps2, 0.0, True, src3),
    }
    pin = synthetic_pressure_inputs(mesh, 21, relax=True)
    ops["pressure_equation"] = lambda d: d.pressure_equation(pin)
    out = {"workload": f"{n_side}^3 jittered hex box ({nc} cells, {mesh.n_faces} faces), FP64",
           "roofline_peak_gbs": 8000.0, "ops": {}}
    with DeviceFv(mesh) as dev:
        dev.vector_gradient(u, ub, 1)
        for name, f in ops.items():
            f(dev)  # warm-up
            dev.timing(True)
            dms = []
            for _ in range(reps):
                f(dev)
                dms.append(dev.last_kernel_ms())
            dev.timing(False)
            t = time.perf_counter()
            for _ in range(reps):
                f(dev)
            wall = (time.perf_counter() - t) / reps
            ms = float(np.median(dms))
            nbytes = fv_bytes(mesh, name)
            gbs = nbytes / (ms * 1e-3) / 1e9
            out["ops"][name] = {"device_ms": round(ms, 4), "bytes": int(nbytes), "achieved_gbs": round(gbs, 1),
                                "frac": round(gbs / 8000.0, 4), "call_ms_with_pcie": round(1e3 * wall, 2)}
            log(f"fv {name}: {ms:.3f} ms device, {gbs:.0f} GB/s ({gbs / 80.0:.1f}% of HBM), "
                f"{1e3 * wall:.1f} ms with PCIe")
        out["pressure_handover"] = pressure_handover(mesh, dev, reps)
    dom = max(out["ops"], key=lambda k: out["ops"][k]["device_ms"])
    out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": out["ops"][dom]["achieved_gbs"], "peak": 8000.0,
                       "unit": "GB/s", "frac": out["ops"][dom]["frac"]}
    if cpu:
        out["cpu_baseline"] = fv_cpu_baseline(mesh, ops)
    return out


def pressure_handover(mesh, dev, reps: int) -> dict:
    """A solver system takes the pressure equation the DeviceFv object assembled, within HBM
    (cfdg_system_assemble_pressure), against the same equation as host arrays (assemble +
    set_rhs over PCIe): wall ms per hand-over."""
    import cfd_amd
    from cfd_amd import cases
    from cfd_amd.fv import synthetic_pressure_inputs

    nif = mesh.n_internal_faces
    pat = cases.csr_pattern(mesh.n_cells, mesh.owner[:nif].astype(np.int32), mesh.neighbour.astype(np.int32))
    eq = dev.pressure_equation(synthetic_pressure_inputs(mesh, 21, relax=True))
    sysd = cfd_amd.DeviceSolver(pat)
    try:
        res = {}
        for name, f in (("resident_ms", lambda: sysd.assemble_pressure(dev)),
                        ("host_arrays_ms", lambda: (sysd.assemble(eq.diag, eq.upper, eq.lower),
                                                    sysd.set_rhs(eq.source)))):
            f()
            t = time.perf_counter()
            for _ in range(reps):
                f()
            res[name] = round(1e3 * (time.perf_counter() - t) / reps, 3)
        res["host_bytes"] = int(8 * (2 * mesh.n_cells + 2 * nif))
    finally:
        sysd.close()
    log(f"pressure hand-over: resident {res['resident_ms']} ms, host arrays {res['host_arrays_ms']} ms")
    return res


def fv_cpu_baseline(mesh, ops) -> dict:
    """The reference's gpu::DeviceFv on host threads (oracle/_ref), one call of each operation on
    the same mesh and fields, in the same order (the momentum corrections use the vector gradient
    of the call before them)."""
    from oracle import bind as ob

    kind = "reference" if ob.ref_available() else "port"
    host = ob.FvReference(mesh) if kind == "reference" else None
    res = {}
    for name, f in ops.items():
        if host is None or name == "pressure_equation":
            continue
        t = time.perf_counter()
        f(host)
        res[name] = round(1e3 * (time.perf_counter() - t), 2)
    cores = os.cpu_count() or 1
    out = {"unit": "ms per call", "ms": res, "cores": int(os.environ.get("OMP_NUM_THREADS", cores)), "kind": kind,
           "sample": "one call of each operation on the same mesh (host arrays, no transfers)"}
    if "pressure_equation" in ops:
        # The reference's flow solver needs Eigen (absent): its face loops are timed as the
        # oracle's single-threaded C restatement.
        port = ob.FvOracle(mesh)
        t = time.perf_counter()
        ops["pressure_equation"](_OracleAsDevice(port))
        out["pressure_equation"] = {"ms": round(1e3 * (time.perf_counter() - t), 2), "cores": 1, "kind": "port"}
    return out


class _OracleAsDevice:
    """Lets the pressure-equation op run on the oracle (same call shape as DeviceFv)."""

    def __init__(self, port):
        self.port = port

    def pressure_equation(self, inp):
        return self.port.pressure_equation(inp)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-stress", action="store_true")
    ap.add_argument("--skip-fv", action="store_true")
    ap.add_argument("--fv-only", action="store_true", help="only the finite-volume kernel section")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 10 device copies of 256 MiB (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.copy_calibration:
        import torch

        src = torch.ones(32 * 2 ** 20, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(10):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8,
                          "copies": 10}))
        return 0
    if args.fv_only:
        print(json.dumps({"fv": fv_section(cpu=not args.skip_cpu)}), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(batched and served systems shard across GPUs by system)"}))
        return 2

    import cfd_amd
    from cfd_amd import AmgHierarchy, AmgSettings, DeviceSolver, Method, Precision, SolveControl, cases

    if not cfd_amd.device_available():
        raise RuntimeError("no HIP device visible")
    log(cfd_amd.device_description())
    t0 = time.perf_counter()
    mesh, p_sys, m_sys = cases.config2(args.size)
    pat = cases.csr_pattern(p_sys.n, p_sys.owner, p_sys.neighbour)
    a = cases.ldu_to_csr(pat, p_sys.diag, p_sys.upper, p_sys.lower)
    log(f"generated {p_sys.n} cells / {p_sys.n_faces} faces in {time.perf_counter() - t0:.1f} s")
    hier = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    # the same setup with its sparse products on the GPU (second build: the first call in a
    # process also loads the sort kernels); the hierarchy is bit-identical
    AmgHierarchy(a, AmgSettings(strength_decay=0.5), device=True)
    setup_dev_ms = AmgHierarchy(a, AmgSettings(strength_decay=0.5), device=True).setup_ms
    log(f"AMG setup {hier.setup_ms:.0f} ms (host), {setup_dev_ms:.0f} ms (products on the GPU), "
        f"levels {hier.level_sizes()}, operator complexity {hier.operator_complexity():.3f}")
    view = hier.view()
    dev = DeviceSolver(pat)
    dev.assemble(p_sys.diag, p_sys.upper, p_sys.lower)
    dev.set_amg(view)
    b = p_sys.source
    dev.set_rhs(b)
    bnorm = float(np.linalg.norm(b))
    tol = 1e-8 * bnorm
    ctl = SolveControl(tol, 500)

    def run(prec, k):
        lat, dev_ms, its = [], [], []
        for _ in range(k):
            t = time.perf_counter()
            dev.set_x(None)
            st = dev.solve_resident(Method.Cg, True, prec, ctl)
            lat.append(time.perf_counter() - t)
            dev_ms.append(st.solve_ms)
            its.append(st.iterations)
            if not st.converged:
                raise RuntimeError(f"{prec.name} solve did not converge: {st}")
        return np.array(lat), np.array(dev_ms), its, st

    run(Precision.MIXED_FP32, args.warmup)
    run(Precision.FP64, max(2, args.warmup))
    # Timed region (the value): HIP-graph solves, no instrumentation.
    t_mixed = time.perf_counter()
    lat_m, dms_m, its_m, st_m = run(Precision.MIXED_FP32, args.steps)
    t_mixed = time.perf_counter() - t_mixed
    x_mixed = dev.get_x()
    t_fp64 = time.perf_counter()
    lat_d, dms_d, its_d, st_d = run(Precision.FP64, args.steps)
    t_fp64 = time.perf_counter() - t_fp64
    x_fp64 = dev.get_x()
    run(Precision.MIXED_BF16, max(2, args.warmup))
    t_bf16 = time.perf_counter()
    lat_b, dms_b, its_b, st_b = run(Precision.MIXED_BF16, args.steps)
    t_bf16 = time.perf_counter() - t_bf16
    x_bf16 = dev.get_x()
    log(f"bf16: {args.steps / t_bf16:.2f} solves/s, it {its_b[-1]}, true rel res "
        f"{st_b.true_residual / bnorm:.2e}")
    # Instrumented pass over the same solves: HIP events on the solve stream bracket every
    # launch of the dominant kernel (the last level-0 Jacobi sweep of the V-cycle, fused with the
    # FCG dot products), and of the outer FP64 SpMV. Kernel-by-kernel launches, same kernels.
    k_probe = max(2, args.steps // 2)
    dev.probe(1)
    run(Precision.MIXED_FP32, k_probe)
    probe_m = dev.probe_read(Precision.MIXED_FP32)
    dev.probe(1)
    run(Precision.FP64, k_probe)
    probe_d = dev.probe_read(Precision.FP64)
    dev.probe(2)
    run(Precision.FP64, k_probe)
    probe_spmv = dev.probe_read(Precision.FP64)
    dev.probe(0)

    # HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE) for this workload
    traffic_path = ROOT / "profiles" / "traffic.json"
    traffic = json.loads(traffic_path.read_text())["hbm_bytes_per_launch"] if traffic_path.exists() else {}

    def roof(p, name, key):
        ms, launches, nbytes = p
        gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get(key), "kernel": name,
                "bytes_per_launch": int(nbytes), "avg_us": round(ms * 1e3, 3),
                "launches": int(launches)}
    log(f"mixed: {args.steps / t_mixed:.2f} solves/s, it {its_m[-1]}, true rel res "
        f"{st_m.true_residual / bnorm:.2e}; fp64: {args.steps / t_fp64:.2f} solves/s, it {its_d[-1]}, "
        f"true rel res {st_d.true_residual / bnorm:.2e}")

    # Momentum system of the same mesh: Jacobi-BiCGStab in FP64 (config 2b).
    mdev = DeviceSolver(pat)
    mdev.assemble(m_sys.diag, m_sys.upper, m_sys.lower)
    mdev.set_rhs(m_sys.source)
    mtol = 1e-8 * float(np.linalg.norm(m_sys.source))
    mdev.set_x(None)
    mdev.solve_resident(Method.BiCgStab, False, Precision.FP64, SolveControl(mtol, 2000))
    t_mom = time.perf_counter()
    k_mom = max(1, args.steps // 2)
    for _ in range(k_mom):
        mdev.set_x(None)
        st_mom = mdev.solve_resident(Method.BiCgStab, False, Precision.FP64, SolveControl(mtol, 2000))
    t_mom = time.perf_counter() - t_mom

    err = float(np.linalg.norm(x_mixed - x_fp64) / np.linalg.norm(x_fp64))
    value = args.steps / t_mixed
    fp64_value = args.steps / t_fp64
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"pressure Poisson, {args.size}^3 jittered hex box ({p_sys.n} cells), "
                        "SA-AMG-PCG to 1e-8 rel. residual from x0 = 0",
            "cells": p_sys.n,
            "nnz": int(a.nnz),
            "amg_levels": hier.level_sizes(),
            "operator_complexity": round(hier.operator_complexity(), 4),
            "strength_threshold": "0.08 * 0.5^level",
            "vcycle": "Jacobi w=0.7, 2+2 sweeps, dense FP64 coarsest inverse",
        },
        "iterations_mixed": int(its_m[-1]),
        "iterations_fp64": int(its_d[-1]),
        "fp64_solves_per_s": round(fp64_value, 3),
        "speedup_vs_fp64": round(value / fp64_value, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat_m, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat_m, 99)), 3),
        "fp64_p99_ms": round(1e3 * float(np.percentile(lat_d, 99)), 3),
        "device_ms_mixed": round(float(np.median(dms_m)), 3),
        "device_ms_fp64": round(float(np.median(dms_d)), 3),
        "true_rel_residual_mixed": st_m.true_residual / bnorm,
        "true_rel_residual_fp64": st_d.true_residual / bnorm,
        "mixed_vs_fp64_rel_error": err,
        "mixed_bf16": {"solves_per_s": round(args.steps / t_bf16, 3), "iterations": int(its_b[-1]),
                       "device_ms": round(float(np.median(dms_b)), 3),
                       "p99_ms": round(1e3 * float(np.percentile(lat_b, 99)), 3),
                       "true_rel_residual": st_b.true_residual / bnorm,
                       "rel_error_vs_fp64": float(np.linalg.norm(x_bf16 - x_fp64) / np.linalg.norm(x_fp64))},
        "momentum_bicgstab_fp64": {"solves_per_s": round(k_mom / t_mom, 3),
                                   "iterations": st_mom.iterations,
                                   "true_rel_residual": st_mom.true_residual / np.linalg.norm(m_sys.source)},
        "setup_ms": round(hier.setup_ms, 1),
        "setup_ms_device_products": round(setup_dev_ms, 1),
        "footprint_bytes": dev.footprint(),
        "roofline": roof(probe_m, "k_jacobi_pair<float,2>: last level-0 FP32 Jacobi sweep + r.z, z.q",
                         "k_jacobi_pair<float, 2>@8192"),
        "roofline_fp64_sweep": roof(probe_d, "k_jacobi<double,1>: last level-0 FP64 Jacobi sweep + r.z",
                                    "k_jacobi<double, 1>@8192"),
        "roofline_fp64_spmv": roof(probe_spmv, "k_spmv_pz<double>: FP64 q = A p_new, p_new = z + beta p_old, p.q",
                                   "k_spmv_pz<double>@8192"),
        "working_set_mb": round(dev.footprint() / 1e6, 1),
        "cpu_baseline": None,
    }
    if not args.skip_batch:
        del dev, mdev
        out["batched"] = batched(max(2, args.steps // 4))
    if not args.skip_serve:
        out["served"] = served()
    if not args.skip_stress:
        out["precision_stress"] = stress()
    if not args.skip_fv:
        out["fv"] = fv_section(cpu=not args.skip_cpu)
    if not args.skip_cpu:
        out["cpu_baseline"] = cpu_baseline(view, b, tol, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
