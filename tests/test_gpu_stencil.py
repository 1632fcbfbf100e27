"""The per-slice stencil layout of the finest operator (one offset table per 64-row slice,
values only per slot) against the 16-bit row-offset layout it replaces: solves are bitwise
identical in every precision and method, the stencil is taken where the slices' offset unions
are no wider than the rows, and matrices whose columns scatter or do not ascend keep the
per-slot column layouts.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, CsrView, Method, Precision, SolveControl, cases

from .conftest import poisson_2d_csr

pytestmark = pytest.mark.gpu


def _solve_pattern(monkeypatch, stencil, s, pat, view, method, prec, use_amg):
    monkeypatch.setenv("CFDG_NO_SYM", "1")  # the per-slice stencil itself (symmetric storage: test_gpu_symmetric)
    if stencil:
        monkeypatch.delenv("CFDG_NO_STENCIL", raising=False)
    else:
        monkeypatch.setenv("CFDG_NO_STENCIL", "1")
    dev = cfd_amd.DeviceSolver(pat)
    try:
        dev.assemble(s.diag, s.upper, s.lower)
        if view is not None:
            dev.set_amg(view)
        dev.set_rhs(s.source)
        dev.set_x(None)
        tol = 1e-9 * np.linalg.norm(s.source)
        st = dev.solve_resident(method, use_amg, prec, SolveControl(tol, 400))
        return st, dev.get_x(), dev.footprint()
    finally:
        dev.close()


@pytest.mark.parametrize("shape,jitter", [((20, 20, 20), 0.2), ((9, 13, 40), 0.15), ((70, 3, 5), 0.0)])
@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32, Precision.MIXED_BF16])
def test_stencil_amg_solves_match_row_offsets_bitwise(monkeypatch, shape, jitter, prec):
    mesh = cases.box_mesh(*shape, jitter=jitter, seed=5)
    s = cases.pressure_system(mesh, rhs_seed=6)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    view = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5)).view()
    st1, x1, f1 = _solve_pattern(monkeypatch, True, s, pat, view, Method.Cg, prec, True)
    st0, x0, f0 = _solve_pattern(monkeypatch, False, s, pat, view, Method.Cg, prec, True)
    print(f"{shape} {prec.name}: it {st1.iterations} / {st0.iterations}, footprint {f1} / {f0}")
    assert st1.converged and st0.converged
    assert st1.iterations == st0.iterations
    assert np.array_equal(x1, x0)
    assert f1 < f0  # the stencil replaced the per-slot columns


def test_stencil_bicgstab_momentum_matches_bitwise(monkeypatch):
    mesh = cases.box_mesh(17, 11, 23, jitter=0.2, seed=8)
    s = cases.momentum_system(mesh, dt=0.5, rhs_seed=9)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    st1, x1, f1 = _solve_pattern(monkeypatch, True, s, pat, None, Method.BiCgStab, Precision.FP64, False)
    st0, x0, f0 = _solve_pattern(monkeypatch, False, s, pat, None, Method.BiCgStab, Precision.FP64, False)
    assert st1.converged and st1.iterations == st0.iterations
    assert np.array_equal(x1, x0)
    assert f1 < f0


def _csr_solve(monkeypatch, stencil, a, b, method):
    if stencil:
        monkeypatch.delenv("CFDG_NO_STENCIL", raising=False)
    else:
        monkeypatch.setenv("CFDG_NO_STENCIL", "1")
    x = np.zeros(a.n)
    st = cfd_amd.solve(a, b, x, method, SolveControl(1e-10, 4000))
    return st, x


@pytest.mark.parametrize("c,method", [(0.0, Method.Cg), (0.4, Method.BiCgStab)])
def test_stencil_csr_entry_point_matches_bitwise(monkeypatch, c, method):
    rp, ci, va = poisson_2d_csr(45, c)
    a = CsrView(45 * 45, len(va), rp, ci, va)
    b = np.cos(0.07 * np.arange(a.n))
    st1, x1 = _csr_solve(monkeypatch, True, a, b, method)
    st0, x0 = _csr_solve(monkeypatch, False, a, b, method)
    assert st1.converged and st1.iterations == st0.iterations
    assert np.array_equal(x1, x0)


def test_scattered_and_unsorted_columns_keep_slot_columns(monkeypatch):
    """A symmetric permutation scatters the columns (offset unions far wider than the rows);
    shuffled columns within rows are not ascending. Both are solved through the per-slot column
    layouts, identically with and without the stencil switch, and agree with the plain matrix."""
    m = 40
    rp, ci, va = poisson_2d_csr(m)
    n = m * m
    perm = np.random.default_rng(3).permutation(n)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    pr, pc = inv[rows], inv[ci]
    order = np.lexsort((pc, pr))
    prp = np.zeros(n + 1, np.int32)
    np.add.at(prp, pr + 1, 1)
    prp = np.cumsum(prp).astype(np.int32)
    ap = CsrView(n, len(va), prp, pc[order].astype(np.int32), va[order])
    b = np.sin(0.05 * np.arange(n))
    st1, x1 = _csr_solve(monkeypatch, True, ap, b, Method.Cg)
    st0, x0 = _csr_solve(monkeypatch, False, ap, b, Method.Cg)
    assert st1.converged and st1.iterations == st0.iterations
    assert np.array_equal(x1, x0)
    # Unsorted rows: reverse each row's entries.
    uci = ci.copy()
    uva = va.copy()
    for r in range(n):
        uci[rp[r]:rp[r + 1]] = ci[rp[r]:rp[r + 1]][::-1]
        uva[rp[r]:rp[r + 1]] = va[rp[r]:rp[r + 1]][::-1]
    au = CsrView(n, len(va), rp, uci, uva)
    b2 = np.sin(0.05 * np.arange(n))
    su1, xu1 = _csr_solve(monkeypatch, True, au, b2, Method.Cg)
    su0, xu0 = _csr_solve(monkeypatch, False, au, b2, Method.Cg)
    assert su1.converged and su1.iterations == su0.iterations
    assert np.array_equal(xu1, xu0)
    a = CsrView(n, len(va), rp, ci, va)
    sp, xp = _csr_solve(monkeypatch, True, a, b2, Method.Cg)
    assert np.linalg.norm(xu1 - xp) <= 1e-8 * np.linalg.norm(xp)
