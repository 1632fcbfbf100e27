"""The oracle's restatement of the reference's caller policy for the device solver
(LduMatrix::solve with a CSR pattern and an AmgCache, src/fv/src/ldu_matrix.cpp:111-131,
287-366, and the host AmgCg path's stale rule :242-256), pinned by known answers: the counterpart
of Amg.AmgCacheReuse (tests/test_amg.cpp:232-252), the rebuild points of max_age, the one-level
rule, the early return, the tolerance mapping against an independent dense computation, and a
case where the stale rule must fire."""
from __future__ import annotations

import numpy as np
import pytest

from cfd_amd import cases

from .conftest import build_poisson_ldu


def _pattern(oracle, mesh):
    return oracle.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)


def _manufactured(n, s):
    return np.sin(s * np.arange(n)) + 1.0  # tests/test_amg.cpp:94-101


def test_amg_cache_reuse(oracle):
    """Amg.AmgCacheReuse with the device solver type: five solves on one cache, each within 1e-6
    of its manufactured solution, age 5, one hierarchy."""
    mesh = cases.box_mesh(40, 40, 1)
    diag, up, lo = build_poisson_ldu(mesh, True)
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState()
    for k in range(5):
        xe = _manufactured(mesh.n_cells, 0.01 * (k + 1))
        b = oracle.ldu_multiply(mesh.owner, mesh.neighbour, diag, up, lo, xe)
        x = np.zeros(mesh.n_cells)
        out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 0, 1e-12, 0.0, 200)
        assert out["solved"] and out["used_amg"] and out["setups"] == (1 if k == 0 else 0)
        assert np.abs(x - xe).max() < 1e-6
    assert cache.h is not None and cache.age == 5 and cache.setups == 1


def test_hierarchy_rebuilt_every_max_age_solves(oracle):
    mesh = cases.box_mesh(20, 20, 2)
    diag0, up, lo = build_poisson_ldu(mesh, True)
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState(max_age=20)
    setups = []
    for k in range(45):
        diag = diag0 + 0.01 * k
        b = np.cos(0.1 * np.arange(mesh.n_cells) + k)
        x = np.zeros(mesh.n_cells)
        out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 0, 1e-10, 0.0, 500)
        assert out["used_amg"]
        setups.append(out["setups"])
    assert [k for k, s in enumerate(setups) if s] == [0, 20, 40]
    assert cache.age == 5 and cache.setups == 3


def test_bicgstab_never_builds_a_hierarchy(oracle):
    mesh = cases.box_mesh(20, 20, 2)
    diag, up, lo = build_poisson_ldu(mesh, True)
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState()
    b = np.cos(0.1 * np.arange(mesh.n_cells))
    x = np.zeros(mesh.n_cells)
    out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 1, 1e-10, 0.0, 500)
    assert out["solved"] and not out["used_amg"] and cache.h is None and cache.age == 0
    xr = np.zeros(mesh.n_cells)
    st = oracle.solve(oracle.ldu_csr(pat, diag, up, lo), b, xr, 1, out["target"], 500)
    assert st.iterations == out["iterations"] and np.array_equal(x, xr)


def test_one_level_hierarchy_runs_jacobi_cg(oracle):
    """225 cells are within the coarse size (400): the hierarchy has one level, so GpuCg runs
    Jacobi CG (ldu_matrix.cpp:358), bit for bit the oracle's Jacobi CG at the mapped target."""
    mesh = cases.box_mesh(15, 15, 1)
    diag, up, lo = build_poisson_ldu(mesh, True)
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState()
    b = np.cos(0.1 * np.arange(mesh.n_cells))
    x = np.zeros(mesh.n_cells)
    out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 0, 1e-10, 0.0, 500)
    assert cache.h.n_levels == 1 and out["setups"] == 1 and not out["used_amg"] and cache.age == 1
    xr = np.zeros(mesh.n_cells)
    st = oracle.solve(oracle.ldu_csr(pat, diag, up, lo), b, xr, 0, out["target"], 500)
    assert st.iterations == out["iterations"] and np.array_equal(x, xr)


def test_converged_start_returns_before_solving(oracle):
    mesh = cases.box_mesh(20, 20, 2)
    diag, up, lo = build_poisson_ldu(mesh, True)
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState()
    xe = _manufactured(mesh.n_cells, 0.03)
    b = oracle.ldu_multiply(mesh.owner, mesh.neighbour, diag, up, lo, xe)
    x = xe.copy()
    out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 0, 1e-10, 0.0, 500)
    assert not out["solved"] and out["iterations"] == 0 and out["initial"] < 1e-10
    assert np.array_equal(x, xe) and cache.h is None and cache.age == 0


def _dense_normalised_residual(mesh, diag, up, lo, b, x):
    """ldu_matrix.cpp:65-103 from a dense matrix (numpy): |b - Ax|_1 / (sum |Ax - A xref| +
    |b - A xref| + 1e-20), A xref = A 1 mean(x); also |b - Ax|_2."""
    n = mesh.n_cells
    a = np.diag(diag.astype(float))
    a[mesh.owner, mesh.neighbour] = up
    a[mesh.neighbour, mesh.owner] = lo
    ax = a @ x
    ax_ref = a.sum(axis=1) * x.mean()
    res = np.abs(b - ax).sum()
    norm = (np.abs(ax - ax_ref) + np.abs(b - ax_ref)).sum()
    assert n == len(x)
    return res / (norm + 1e-20), float(np.linalg.norm(b - ax))


@pytest.mark.parametrize("tol,rel", [(1e-10, 0.0), (1e-6, 0.05), (1e-3, 1e-4)])
def test_tolerance_mapping(oracle, tol, rel):
    """target = max(rel, tol / initial) |b - A x0|_2 (ldu_matrix.cpp:297-300), with the
    normalised residual of a non-zero initial guess computed independently."""
    mesh = cases.box_mesh(12, 10, 6, jitter=0.2, seed=4)
    s = cases.pressure_system(mesh, rhs_seed=5)
    pat = _pattern(oracle, mesh)
    x0 = 0.3 * np.sin(0.05 * np.arange(mesh.n_cells))
    init_np, r0_np = _dense_normalised_residual(mesh, s.diag, s.upper, s.lower, s.source, x0)
    cache = oracle.AmgCacheState()
    x = x0.copy()
    out = oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, s.diag, s.upper, s.lower, s.source, x, 0,
                           tol, rel, 500)
    assert out["initial"] == pytest.approx(init_np, rel=1e-12)
    assert out["target"] == pytest.approx(max(rel, tol / init_np) * r0_np, rel=1e-12)
    r = s.source - oracle.ldu_multiply(mesh.owner, mesh.neighbour, s.diag, s.upper, s.lower, x)
    assert np.linalg.norm(r) <= out["target"] * (1 + 1e-6)


def _anisotropic(mesh, diag, up, lo, factor):
    """The same mesh with the x-direction coefficients scaled by factor."""
    xfaces = (mesh.neighbour - mesh.owner) == 1
    up2, lo2 = up.copy(), lo.copy()
    up2[xfaces] *= factor
    lo2[xfaces] *= factor
    d2 = diag.copy() - (np.bincount(mesh.owner, -up, mesh.n_cells) + np.bincount(mesh.neighbour, -lo, mesh.n_cells))
    d2 += np.bincount(mesh.owner, -up2, mesh.n_cells) + np.bincount(mesh.neighbour, -lo2, mesh.n_cells)
    return d2, up2, lo2


def stale_sequence():
    """An isotropic solve, then the same mesh with x-coefficients 10x larger: the hierarchy of
    the first matrix is a poor preconditioner for the second (12 iterations, then 59 with the
    stale hierarchy against 2 * 12 + 2 = 26 for the rule)."""
    mesh = cases.box_mesh(24, 24, 4)
    diag, up, lo = build_poisson_ldu(mesh, True)
    d2, up2, lo2 = _anisotropic(mesh, diag, up, lo, 10.0)
    b = np.cos(0.07 * np.arange(mesh.n_cells))
    return mesh, [(diag, up, lo, b), (d2, up2, lo2, b)]


@pytest.mark.parametrize("stale_rebuild", [False, True])
def test_stale_rule_rebuilds_a_poor_hierarchy(oracle, stale_rebuild):
    mesh, seq = stale_sequence()
    pat = _pattern(oracle, mesh)
    cache = oracle.AmgCacheState(stale_rebuild=stale_rebuild)
    outs = []
    for diag, up, lo, b in seq:
        x = np.zeros(mesh.n_cells)
        outs.append(oracle.ldu_solve(cache, mesh.owner, mesh.neighbour, pat, diag, up, lo, b, x, 0, 1e-10, 0.0,
                                     2000))
    first, second = outs
    assert first["setups"] == 1
    if stale_rebuild:
        assert second["setups"] == 1 and cache.setups == 2 and cache.age == 1
        assert cache.setup_iterations < first["iterations"] * 2 + 2 < second["iterations"]
    else:
        assert second["setups"] == 0 and cache.setups == 1 and cache.age == 2
        assert second["iterations"] > 2 * first["iterations"] + 2  # the rule's trigger holds here
