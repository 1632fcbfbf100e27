"""Parity of the benchmark configurations against the oracle (CPU restatement of the
reference's solvers), at full size where a property carries the check and at sizes the oracle
finishes in seconds otherwise.

  * config 1: 32^3 uniform box, the reference's own hierarchy (fixed threshold 0.08), FP64
    AMG-PCG: GPU, oracle and the reference's compiled solve_amg_cg_host (oracle/_ref, when built).
  * config 2: the 128^3 jittered box at full size: (a) pressure AMG-PCG, (b) momentum
    Jacobi-BiCGStab (src/gpu/src/krylov_kernels.inc:96-169).
  * config 3: the 160^3 precision stress at full size, mixed against the FP64 oracle.
  * config 4: every member of a batch of the batch workload's shapes (uniform and jittered,
    cell aspect up to 1:8), solved together, against the oracle's solve of that member alone.
  * config 5: every step of every request served with continuous batching, against the oracle's
    chain of solves (each step from the previous step's solution).

Bars: FP64 within +-1 iteration and 1e-10 relative of the oracle; mixed precision (FP32 V-cycle
in FP64 flexible CG) at the tolerance-implied bound: its true residual, recomputed here by the
oracle's FP64 product, meets the tolerance, and since A (x_m - x_o) = r_o - r_m,
|x_m - x_o| <= (|r_m| + |r_o|) / lambda_min(A), with lambda_min from scipy's shift-invert
Lanczos (shrunk by 10% as a margin) on the systems small enough for it.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

# Full-size systems and the oracle's single-threaded solves take longer than the suite's
# per-test limit.
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(400)]


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _ocsr(oracle, a):
    return oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val)


def _true_res(oracle, oa, b, x):
    return float(np.linalg.norm(b - oracle.spmv(oa, x)))


def _lambda_min(a) -> float:
    import scipy.sparse.linalg as sla

    m = a.to_scipy().tocsc()
    return float(sla.eigsh(m, k=1, sigma=0, which="LM", return_eigenvectors=False)[0])


def _check_mixed(oracle, oa, b, x_m, x_o, tol, lam_min=None, label=""):
    r_m = _true_res(oracle, oa, b, x_m)
    r_o = _true_res(oracle, oa, b, x_o)
    err = float(np.linalg.norm(x_m - x_o))
    bound = None if lam_min is None else (r_m + r_o) / (0.9 * lam_min)
    print(f"{label} mixed: true res {r_m / np.linalg.norm(b):.2e} rel err vs oracle "
          f"{err / np.linalg.norm(x_o):.2e}" + ("" if bound is None else f" (bound {bound / np.linalg.norm(x_o):.2e})"))
    assert r_m <= 1.05 * tol
    if bound is not None:
        assert err <= bound


def _device(pat, s, view=None):
    dev = cfd_amd.DeviceSolver(pat)
    dev.assemble(s.diag, s.upper, s.lower)
    if view is not None:
        dev.set_amg(view)
    dev.set_rhs(s.source)
    return dev


# ---------------------------------------------------------------------------------- config 1
def test_config1_fixed_threshold_against_oracle_and_reference(oracle):
    _, s = cases.config1()
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    h = AmgHierarchy(a)  # strength_decay 1.0: the reference's fixed threshold
    oh = oracle.Hierarchy(_ocsr(oracle, a))
    assert h.level_sizes() == oh.level_sizes()
    b = s.source
    tol = 1e-8 * np.linalg.norm(b)
    x = np.zeros(s.n)
    st = cfd_amd.solve_amg_cg(h.view(), b, x, SolveControl(tol, 500))
    xo = np.zeros(s.n)
    so = oh.amg_cg(b, xo, tol, 500)
    print(f"config 1: levels {h.level_sizes()} gpu it {st.iterations} oracle it {so.iterations} "
          f"rel {_rel(x, xo):.2e}")
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10
    if oracle.ref_available():
        a_l, p_l, r_l = oracle.hierarchy_lists(oh)
        xr = np.zeros(s.n)
        sr = oracle.ref_solve_amg_cg_host(a_l, p_l, r_l, oh.coarse_inverse(), b, xr, tol, 500)
        print(f"config 1: reference solve_amg_cg_host it {sr.iterations} rel {_rel(x, xr):.2e}")
        assert sr.converged and abs(st.iterations - sr.iterations) <= 1 and _rel(x, xr) <= 1e-10


# ---------------------------------------------------------------------------------- config 2
@pytest.fixture(scope="module")
def config2_full():
    mesh, p, m = cases.config2(128)
    pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
    return pat, p, m


def test_config2a_pressure_full_size_against_oracle(oracle, config2_full):
    pat, p, _ = config2_full
    a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    oh = oracle.Hierarchy(_ocsr(oracle, a), decay=0.5)
    b = p.source
    tol = 1e-8 * np.linalg.norm(b)
    xo = np.zeros(p.n)
    so = oh.amg_cg(b, xo, tol, 500)
    dev = _device(pat, p, h.view())
    out = {}
    for prec in (Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16):
        dev.set_x(None)
        st = dev.solve_resident(Method.Cg, True, prec, SolveControl(tol, 500))
        assert st.converged
        out[prec] = (st, dev.get_x())
    dev.close()
    st, x = out[Precision.FP64]
    print(f"config 2a FP64: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e}")
    assert so.converged and abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10
    oa = _ocsr(oracle, a)
    for prec in (Precision.MIXED_FP32, Precision.MIXED_BF16):
        st, x = out[prec]
        _check_mixed(oracle, oa, b, x, xo, tol, label=f"config 2a {prec.name} it {st.iterations}")


def test_config2b_momentum_bicgstab_full_size_against_oracle(oracle, config2_full):
    pat, _, m = config2_full
    a = cases.ldu_to_csr(pat, m.diag, m.upper, m.lower)
    b = m.source
    tol = 1e-8 * np.linalg.norm(b)
    dev = _device(pat, m)
    dev.set_x(None)
    st = dev.solve_resident(Method.BiCgStab, False, Precision.FP64, SolveControl(tol, 2000))
    x = dev.get_x()
    dev.close()
    xo = np.zeros(m.n)
    so = oracle.solve(_ocsr(oracle, a), b, xo, 1, tol, 2000)
    print(f"config 2b: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e} "
          f"true res {st.true_residual / np.linalg.norm(b):.2e}")
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10
    assert _true_res(oracle, _ocsr(oracle, a), b, x) <= 1.05 * tol


def test_config2b_momentum_through_the_reference_seam(oracle):
    """The momentum system solved through DeviceSolver::solve's host-array seam (fresh LDU per
    call, BiCGStab without AMG) on a 40^3 jittered box, against the oracle."""
    mesh = cases.box_mesh(40, 40, 40, jitter=0.2, seed=cases.SEED_BASE + 1)
    m = cases.momentum_system(mesh, 1.0, cases.SEED_BASE + 3)
    pat = cases.csr_pattern(m.n, m.owner, m.neighbour)
    a = cases.ldu_to_csr(pat, m.diag, m.upper, m.lower)
    tol = 1e-8 * np.linalg.norm(m.source)
    with cfd_amd.DeviceSolver(pat) as dev:
        x = np.zeros(m.n)
        st = dev.solve(m.diag, m.upper, m.lower, m.source, x, Method.BiCgStab, False,
                       SolveControl(tol, 2000))
    xo = np.zeros(m.n)
    so = oracle.solve(_ocsr(oracle, a), m.source, xo, 1, tol, 2000)
    print(f"40^3 momentum: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e}")
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10


# ---------------------------------------------------------------------------------- config 3
def test_config3_precision_stress_full_size_against_oracle(oracle):
    _, s = cases.config3(160)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    oh = oracle.Hierarchy(_ocsr(oracle, a), decay=0.5)
    b = s.source
    tol = 1e-10 * np.linalg.norm(b)
    oracle.set_threads(8)  # the oracle's threaded loops change only its reduction order
    try:
        xo = np.zeros(s.n)
        so = oh.amg_cg(b, xo, tol, 1000)
    finally:
        oracle.set_threads(1)
    dev = _device(pat, s, h.view())
    oa = _ocsr(oracle, a)
    for prec in (Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16):
        dev.set_x(None)
        st = dev.solve_resident(Method.Cg, True, prec, SolveControl(tol, 1000))
        x = dev.get_x()
        assert st.converged
        if prec == Precision.FP64:
            print(f"config 3 FP64: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e}")
            assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10
        else:
            _check_mixed(oracle, oa, b, x, xo, tol, label=f"config 3 {prec.name} it {st.iterations}")
            assert st.iterations <= so.iterations + 3
    dev.close()


# ---------------------------------------------------------------------------------- config 4
BATCH = 16
BATCH_MAX_CELLS = 30_000


@pytest.fixture(scope="module")
def batch_members():
    members = []
    for sh in cases.batch_shapes(BATCH, n_max=BATCH_MAX_CELLS):
        mesh = cases.box_mesh(sh["nx"], sh["ny"], sh["nz"], hi=(1.0, 1.0, sh["lz"]),
                              jitter=sh["jitter"], seed=sh["seed"])
        p = cases.pressure_system(mesh, rhs_seed=sh["seed"] + 1)
        pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
        a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
        members.append(dict(s=p, pat=pat, a=a, h=AmgHierarchy(a, AmgSettings(strength_decay=0.5)),
                            lam=_lambda_min(a)))
    return members


@pytest.fixture(scope="module")
def batch_oracle(oracle, batch_members):
    out = []
    for mb in batch_members:
        oh = oracle.Hierarchy(_ocsr(oracle, mb["a"]), decay=0.5)
        b = mb["s"].source
        xo = np.zeros(mb["s"].n)
        so = oh.amg_cg(b, xo, 1e-8 * np.linalg.norm(b), 500)
        assert so.converged
        out.append((so, xo))
    return out


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16])
def test_config4_every_batch_member_against_oracle(oracle, batch_members, batch_oracle, prec):
    devs = [_device(mb["pat"], mb["s"], mb["h"].view()) for mb in batch_members]
    ctls = [SolveControl(1e-8 * np.linalg.norm(mb["s"].source), 500) for mb in batch_members]
    try:
        with cfd_amd.Batch(devs, Method.Cg, True, prec) as batch:
            for rep in range(2):  # the second solve replays the captured graphs
                for d in devs:
                    d.set_x(None)
                stats = batch.solve(ctls)
                for i, (mb, d, st, (so, xo)) in enumerate(zip(batch_members, devs, stats, batch_oracle)):
                    x = d.get_x()
                    label = f"rep {rep} member {i} (n {mb['s'].n}, {'jittered' if i % 2 else 'uniform'})"
                    assert st.converged, label
                    if prec == Precision.FP64:
                        print(f"{label}: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e}")
                        assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10, label
                    else:
                        _check_mixed(oracle, _ocsr(oracle, mb["a"]), mb["s"].source, x, xo,
                                     ctls[i].abs_tolerance, mb["lam"], label)
    finally:
        for d in devs:
            d.close()


# ---------------------------------------------------------------------------------- config 5
@pytest.fixture(scope="module")
def served_requests():
    gen, _ = cases.serving_requests(6, n_min=5e3, n_max=3e4, max_steps=5, seed=20240611)
    for g in gen:
        g["view"] = AmgHierarchy(g["csr"], AmgSettings(strength_decay=0.5)).view()
        g["lam"] = _lambda_min(g["csr"])
    return gen


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32])
def test_config5_every_served_step_against_oracle(oracle, served_requests, prec):
    import threading

    reqs = served_requests
    with cfd_amd.Server(max_active_solves=3, precision=prec) as srv:
        first = [srv.submit(pattern=g["pattern"], diag=g["system"].diag, upper=g["system"].upper,
                            lower=g["system"].lower, amg=g["view"], rhs=g["rhs"]) for g in reqs[:3]]
        late = []

        def arrive():  # requests that arrive while others are on the device
            for g in reqs[3:]:
                late.append(srv.submit(pattern=g["pattern"], diag=g["system"].diag,
                                       upper=g["system"].upper, lower=g["system"].lower,
                                       amg=g["view"], rhs=g["rhs"]))

        t = threading.Thread(target=arrive)
        t.start()
        done = [srv.wait(r) for r in first]
        t.join()
        done += [srv.wait(r) for r in late]
    for i, (g, r) in enumerate(zip(reqs, done)):
        oa = _ocsr(oracle, g["csr"])
        oh = oracle.Hierarchy(oa, decay=0.5)
        xo = np.zeros(g["system"].n)
        for k, b in enumerate(g["rhs"]):
            tol = 1e-8 * np.linalg.norm(b)
            so = oh.amg_cg(b, xo, tol, 500)  # xo: the previous step's oracle solution
            st, x = r.stats[k], r.solutions[k]
            label = f"request {i} (n {g['system'].n}) step {k}"
            assert st.converged and so.converged, label
            if prec == Precision.FP64:
                print(f"{label}: gpu it {st.iterations} oracle it {so.iterations} rel {_rel(x, xo):.2e}")
                assert abs(st.iterations - so.iterations) <= 1 and _rel(x, xo) <= 1e-10, label
            else:
                _check_mixed(oracle, oa, b, x, xo, tol, g["lam"], label)
