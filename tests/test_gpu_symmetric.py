"""Symmetric stencil storage of the finest operator (one stencil for all rows, only the entries
with offset >= 0 stored, the others read from the mirrored row; used while an assembly is
exactly symmetric, upper == lower): solves are bitwise identical to the per-slice stencil layout
(CFDG_NO_SYM) in every precision, batched or alone; an asymmetric assembly of the same system
(momentum-like, or a perturbed pressure matrix) switches the system back to the full layout,
and a later symmetric one switches it again, each time with results identical to a system that
never had the symmetric storage.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

pytestmark = pytest.mark.gpu


def _pressure(shape, jitter, seed):
    mesh = cases.box_mesh(*shape, jitter=jitter, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed + 1)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    view = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5)).view()
    return s, pat, view


def _device(monkeypatch, sym, pat):
    if sym:
        monkeypatch.delenv("CFDG_NO_SYM", raising=False)
    else:
        monkeypatch.setenv("CFDG_NO_SYM", "1")
    return cfd_amd.DeviceSolver(pat)  # the layout is chosen when the system is created


def _solve(dev, diag, upper, lower, b, view, method, prec, use_amg, tol):
    dev.assemble(diag, upper, lower)
    if view is not None:
        dev.set_amg(view)
    dev.set_rhs(b)
    dev.set_x(None)
    st = dev.solve_resident(method, use_amg, prec, SolveControl(tol, 400))
    return st, dev.get_x()


@pytest.mark.parametrize("shape,jitter", [((24, 20, 18), 0.2), ((9, 13, 40), 0.15), ((64, 3, 5), 0.0)])
@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16])
def test_symmetric_storage_solves_match_bitwise(monkeypatch, shape, jitter, prec):
    s, pat, view = _pressure(shape, jitter, 5)
    tol = 1e-9 * np.linalg.norm(s.source)
    out = {}
    for sym in (True, False):
        dev = _device(monkeypatch, sym, pat)
        try:
            st, x = _solve(dev, s.diag, s.upper, s.lower, s.source, view, Method.Cg, prec, True, tol)
            st2, x2 = _solve(dev, s.diag, s.upper, s.lower, s.source, view, Method.Cg, prec, True, tol)  # graphs
            out[sym] = (st, x, st2, x2, dev.footprint())
        finally:
            dev.close()
    (st1, x1, st1b, x1b, f1), (st0, x0, st0b, x0b, f0) = out[True], out[False]
    print(f"{shape} {prec.name}: it {st1.iterations} / {st0.iterations}, footprint {f1} / {f0}")
    assert st1.converged and st0.converged
    assert st1.iterations == st0.iterations and st1b.iterations == st0b.iterations
    assert np.array_equal(x1, x0) and np.array_equal(x1b, x0b) and np.array_equal(x1, x1b)
    assert f1 > f0  # the symmetric storage was built


def test_jacobi_cg_and_normalised_residual_match_bitwise(monkeypatch):
    s, pat, _ = _pressure((21, 17, 15), 0.2, 11)
    tol = 1e-9 * np.linalg.norm(s.source)
    res = {}
    for sym in (True, False):
        dev = _device(monkeypatch, sym, pat)
        try:
            st, x = _solve(dev, s.diag, s.upper, s.lower, s.source, None, Method.Cg, Precision.FP64, False, tol)
            res[sym] = (st.iterations, x, dev.normalised_residual())
        finally:
            dev.close()
    assert res[True][0] == res[False][0]
    assert np.array_equal(res[True][1], res[False][1])
    assert res[True][2] == res[False][2]


def test_asymmetric_assembly_switches_layouts(monkeypatch):
    """One system assembled symmetric, asymmetric (lower perturbed), then symmetric again; each
    solve against a fresh system without the symmetric storage."""
    s, pat, view = _pressure((20, 18, 16), 0.2, 21)
    rng = np.random.default_rng(3)
    lower_p = s.lower * (1.0 + 0.05 * rng.random(s.lower.shape))
    diag_p = 2.0 * s.diag  # diagonally dominant: Jacobi-BiCGStab converges quickly
    tol = 1e-9 * np.linalg.norm(s.source)
    steps = [(s.diag, s.lower, Method.Cg, True), (diag_p, lower_p, Method.BiCgStab, False),
             (s.diag, s.lower, Method.Cg, True)]
    dev = _device(monkeypatch, True, pat)
    try:
        got = [_solve(dev, d, s.upper, lo, s.source, view if amg else None, m, Precision.MIXED_FP32 if amg
                      else Precision.FP64, amg, tol) for d, lo, m, amg in steps]
    finally:
        dev.close()
    for (d, lo, m, amg), (st, x) in zip(steps, got):
        ref = _device(monkeypatch, False, pat)
        try:
            st0, x0 = _solve(ref, d, s.upper, lo, s.source, view if amg else None, m,
                             Precision.MIXED_FP32 if amg else Precision.FP64, amg, tol)
        finally:
            ref.close()
        print(f"{m.name} amg={amg}: it {st.iterations} / {st0.iterations}")
        assert st.converged and st0.converged
        assert st.iterations == st0.iterations
        assert np.array_equal(x, x0)


def test_momentum_system_keeps_the_full_layout_bitwise(monkeypatch):
    mesh = cases.box_mesh(17, 11, 23, jitter=0.2, seed=8)
    s = cases.momentum_system(mesh, dt=0.5, rhs_seed=9)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    tol = 1e-9 * np.linalg.norm(s.source)
    xs = []
    for sym in (True, False):
        dev = _device(monkeypatch, sym, pat)
        try:
            st, x = _solve(dev, s.diag, s.upper, s.lower, s.source, None, Method.BiCgStab, Precision.FP64, False, tol)
            assert st.converged
            xs.append((st.iterations, x))
        finally:
            dev.close()
    assert xs[0][0] == xs[1][0] and np.array_equal(xs[0][1], xs[1][1])


def test_batch_with_symmetric_storage_matches_bitwise(monkeypatch):
    systems = [_pressure(shape, 0.2, 30 + k) for k, shape in enumerate([(16, 16, 16), (30, 12, 9), (11, 11, 40)])]
    out = {}
    for sym in (True, False):
        devs = [_device(monkeypatch, sym, pat) for _, pat, _ in systems]
        try:
            for dev, (s, _, view) in zip(devs, systems):
                dev.assemble(s.diag, s.upper, s.lower)
                dev.set_amg(view)
                dev.set_rhs(s.source)
                dev.set_x(None)
            ctl = [SolveControl(1e-9 * np.linalg.norm(s.source), 400) for s, _, _ in systems]
            with cfd_amd.Batch(devs, Method.Cg, True, Precision.MIXED_FP32) as batch:
                sts = batch.solve(ctl)
            out[sym] = [(st.iterations, dev.get_x()) for st, dev in zip(sts, devs)]
        finally:
            for dev in devs:
                dev.close()
    for (i1, x1), (i0, x0) in zip(out[True], out[False]):
        assert i1 == i0 and np.array_equal(x1, x0)
