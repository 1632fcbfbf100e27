"""Failure and guard paths of the HIP solvers against the oracle and, where it was built, the
reference's own compiled host solvers (oracle/_ref):

  * a zero diagonal: the Jacobi scaling uses 1.0 for it (krylov_openmp.cpp:147-157,
    krylov_cuda.cu:557-562), on the CSR one-shot path and on the LDU persistent path;
  * CG breakdown on p.q = 0 (krylov_kernels.inc:68-70);
  * BiCGStab's guards: r0.v = 0 (:133-135), omega = 0 from t.s = 0 and from t.t = 0 (:151-152),
    then rho_new = 0 / omega = 0 on the next iteration (:121-123);
  * a singular coarsest operator: the setup's LDL^T meets an exactly zero pivot and falls back to
    the pseudo-inverse (amg.cpp:185-189, 227-240), and AMG-PCG runs on that hierarchy.

Small exact cases are constructed so that every guard fires on known iterations; they are
checked against the closed-form outcome as well as against the oracle and the reference.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, CsrView, Method, SolveControl, cases

from .conftest import poisson_2d_csr

pytestmark = pytest.mark.gpu


def _blocks(block, nb):
    """Block-diagonal CSR of nb copies of a dense 2x2 block (explicit zeros kept)."""
    rp, ci, va = [0], [], []
    for k in range(nb):
        for i in range(2):
            for j in range(2):
                ci.append(2 * k + j)
                va.append(block[i][j])
            rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64)


def _run_all(oracle, rp, ci, va, b, method, tol, max_it, x0=None):
    n = len(rp) - 1
    out = {}
    x = np.zeros(n) if x0 is None else x0.copy()
    st = cfd_amd.solve(CsrView(n, len(va), rp, ci, va), b, x, method, SolveControl(tol, max_it))
    out["gpu"] = (st.iterations, st.converged, st.residual, x)
    xo = np.zeros(n) if x0 is None else x0.copy()
    so = oracle.solve(oracle.Csr(n, n, rp, ci, va), b, xo, int(method), tol, max_it)
    out["oracle"] = (so.iterations, so.converged, so.residual, xo)
    if oracle.ref_available():
        xr = np.zeros(n) if x0 is None else x0.copy()
        sr = oracle.ref_solve_host(oracle.Csr(n, n, rp, ci, va), b, xr, int(method), tol, max_it)
        out["reference"] = (sr.iterations, sr.converged, sr.residual, xr)
    return out, st


def test_cg_breakdown_on_zero_pq(oracle):
    """A = diag([[0, 1], [1, 0]] x 500), b = (1, 0, 1, 0, ...): dinv = 1 (zero diagonal), z = p = r,
    q = A p is orthogonal to p, so the first iteration breaks down with x = 0."""
    rp, ci, va = _blocks([[0.0, 1.0], [1.0, 0.0]], 500)
    b = np.tile([1.0, 0.0], 500)
    out, st = _run_all(oracle, rp, ci, va, b, Method.Cg, 1e-10, 50)
    for who, (it, conv, res, x) in out.items():
        print(f"{who}: it {it} converged {conv} res {res}")
        assert it == 1 and not conv and res == np.sqrt(500.0) and np.array_equal(x, np.zeros(1000)), who
    assert st.exit_reason == 2  # breakdown


@pytest.mark.parametrize("block,label", [([[0.0, 1.0], [1.0, 0.0]], "r0.v = 0"),
                                         ([[1.0, 1.0], [-1.0, 0.0]], "t.s = 0"),
                                         ([[1.0, 0.0], [1.0, 0.0]], "t.t = 0")])
def test_bicgstab_guards(oracle, block, label):
    """b = (1, 0, ...). r0.v = 0: the first iteration stops at alpha's guard with x = 0.
    t.s = 0: alpha = 1, s = (0, 1), t = A s = (1, 0), so omega = 0, x = (1, 0), r = s; the second
    iteration stops on rho_new = r0.r = 0 (and omega = 0). t.t = 0: s = (0, -1), t = A s = 0, so
    omega = 0 by the guard, x = (1, 0), r = s, and the second iteration stops likewise."""
    rp, ci, va = _blocks(block, 300)
    b = np.tile([1.0, 0.0], 300)
    out, st = _run_all(oracle, rp, ci, va, b, Method.BiCgStab, 1e-10, 50)
    want_x = np.zeros(600) if label == "r0.v = 0" else np.tile([1.0, 0.0], 300)
    want_it = 1 if label == "r0.v = 0" else 2
    want_res = np.sqrt(300.0)
    for who, (it, conv, res, x) in out.items():
        print(f"{label} {who}: it {it} converged {conv} res {res}")
        assert it == want_it and not conv, who
        assert res == want_res and np.array_equal(x, want_x), who
    assert st.exit_reason == 2


@pytest.mark.parametrize("method", [Method.Cg, Method.BiCgStab])
def test_zero_diagonal_jacobi_scaling_csr(oracle, method):
    """Five-point operator of tests/test_gpu.cpp with three diagonal entries set to zero: the
    iterates after a fixed number of iterations agree with the oracle and the reference.

    The nonsymmetric operator with zero diagonals makes BiCGStab's iterates sensitive to
    rounding: a 1e-15 relative perturbation of b moves the oracle's iterate by 2e-12 after two
    iterations, 1.3e-9 after three and 1.7e-4 after five, so BiCGStab is compared after three
    iterations at 1e-6 (any other reduction order, the GPU's included, differs at that level),
    while the scaling 0.0 instead of 1.0 at the zero diagonals moves the iterate by 110%."""
    rp, ci, va = poisson_2d_csr(20, 0.3 if method == Method.BiCgStab else 0.0)
    va = va.copy()
    for row in (0, 57, 399):
        for k in range(rp[row], rp[row + 1]):
            if ci[k] == row:
                va[k] = 0.0
    b = np.cos(0.05 * np.arange(400))
    its, bar = (5, 1e-12) if method == Method.Cg else (3, 1e-6)
    out, _ = _run_all(oracle, rp, ci, va, b, method, 1e-30, its)
    x_o = out["oracle"][3]
    for who, (it, conv, res, x) in out.items():
        rel = float(np.linalg.norm(x - x_o) / np.linalg.norm(x_o))
        print(f"{method.name} {who}: it {it} res {res:.6e} rel vs oracle {rel:.2e}")
        assert it == its and not conv and rel <= bar, who


def test_zero_diagonal_on_the_ldu_seam(oracle):
    """DeviceSolver::solve with a zero LDU diagonal entry: the device's inverse diagonal is 1.0
    there; Jacobi CG and BiCGStab iterates match the oracle on the same CSR matrix."""
    mesh = cases.box_mesh(12, 10, 1)
    n, nf = mesh.n_cells, mesh.n_faces
    pat = cases.csr_pattern(n, mesh.owner, mesh.neighbour)
    up = -np.ones(nf)
    lo = -np.ones(nf)
    diag = cases.accumulate_diag(n, mesh.owner, mesh.neighbour, np.ones(nf), np.ones(nf)) + 0.1
    diag[[3, 64]] = 0.0
    b = np.sin(0.2 * np.arange(n))
    a = cases.ldu_to_csr(pat, diag, up, lo)
    with cfd_amd.DeviceSolver(pat) as dev:
        for method in (Method.Cg, Method.BiCgStab):
            x = np.zeros(n)
            st = dev.solve(diag, up, lo, b, x, method, False, SolveControl(1e-30, 4))
            xo = np.zeros(n)
            so = oracle.solve(oracle.Csr(n, n, a.row_ptr, a.col, a.val), b, xo, int(method), 1e-30, 4)
            rel = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
            print(f"{method.name}: gpu it {st.iterations} oracle it {so.iterations} rel {rel:.2e}")
            # a 1e-15 relative perturbation of b moves the oracle's fourth BiCGStab iterate by
            # 1e-11 here (CG: 4e-14), so reduction orders differ at that level
            bar = 1e-12 if method == Method.Cg else 1e-9
            assert st.iterations == so.iterations == 4 and rel <= bar


def _singular_coarse_system():
    """Dirichlet-walled 2-D Laplacian on 30 x 30 cells plus one cell with no faces and a zero
    diagonal: its aggregate's row of every coarse operator is empty, so the coarsest operator
    has an exactly zero pivot."""
    rp, ci, va = poisson_2d_csr(30)
    n = 901
    rp = np.append(rp, rp[-1] + 1).astype(np.int32)
    ci = np.append(ci, n - 1).astype(np.int32)
    va = np.append(va, 0.0)
    return CsrView(n, len(va), rp, ci, va)


def test_singular_coarse_operator_uses_pseudo_inverse_and_solves(oracle):
    a = _singular_coarse_system()
    h = AmgHierarchy(a)
    assert h.n_levels > 1
    c = h.matrix(h.n_levels - 1).to_scipy().toarray()
    assert np.linalg.matrix_rank(c) == c.shape[0] - 1
    ci = h.coarse_inverse()
    pinv = np.linalg.pinv(c)
    assert np.abs(ci - pinv).max() <= 1e-10 * np.abs(pinv).max()
    b = np.cos(0.1 * np.arange(a.n))
    b[-1] = 0.0
    tol = 1e-8 * np.linalg.norm(b)
    x = np.zeros(a.n)
    st = cfd_amd.solve_amg_cg(h.view(), b, x, SolveControl(tol, 200))
    a_l = [oracle.Csr(m.n, m.n_cols or m.n, m.row_ptr, m.col, m.val)
           for m in (h.matrix(l) for l in range(h.n_levels))]
    p_l = [oracle.Csr(m.n, m.n_cols, m.row_ptr, m.col, m.val) if l + 1 < h.n_levels else None
           for l, m in ((l, h.matrix(l, 1)) for l in range(h.n_levels))]
    r_l = [oracle.Csr(m.n, m.n_cols, m.row_ptr, m.col, m.val) if l + 1 < h.n_levels else None
           for l, m in ((l, h.matrix(l, 2)) for l in range(h.n_levels))]
    xo = np.zeros(a.n)
    so = oracle.amg_cg_lists(a_l, p_l, r_l, ci, b, xo, tol, 200)
    rel = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    print(f"singular coarse: gpu it {st.iterations} oracle it {so.iterations} rel {rel:.2e}")
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 1 and rel <= 1e-10
    assert x[-1] == 0.0
    if oracle.ref_available():
        xr = np.zeros(a.n)
        sr = oracle.ref_solve_amg_cg_host(a_l, p_l, r_l, ci, b, xr, tol, 200)
        assert sr.converged and abs(st.iterations - sr.iterations) <= 1
        assert float(np.linalg.norm(x - xr) / np.linalg.norm(xr)) <= 1e-10
