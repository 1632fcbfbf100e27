"""Pins of the oracle (CPU restatement of the reference) by known answers, dense brute force,
closed forms, the invariants the reference's own tests assert, and - where it was built - the
reference's own host solver code (oracle/_ref).

Each check is chosen so that a plausible slip in the restatement (a dropped term, a sign, an
index, a transposed operand) fails it.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from cfd_amd import cases

from .conftest import build_poisson_ldu, poisson_2d_csr


def _csr_of(oracle, m):
    m = sp.csr_matrix(m)
    m.sort_indices()
    return oracle.Csr(m.shape[0], m.shape[1], m.indptr.astype(np.int32), m.indices.astype(np.int32),
                      m.data.astype(np.float64))


def _box_csr(oracle, nx, ny, nz, dirichlet=False, jitter=0.0, seed=1):
    mesh = cases.box_mesh(nx, ny, nz, jitter=jitter, seed=seed)
    diag, up, lo = build_poisson_ldu(mesh, dirichlet)
    n = mesh.n_cells
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(n, mesh.owner, mesh.neighbour)
    val = np.zeros(nnz)
    val[dpos], val[upos], val[lpos] = diag, up, lo
    return mesh, oracle.Csr(n, n, rp, ci, val)


# ---------------------------------------------------------------- LDU (test_ldu_matrix.cpp) --
def _chain(oracle):
    mesh = cases.box_mesh(10, 1, 1)
    nf = mesh.n_faces
    diag = np.bincount(mesh.owner, np.ones(nf), 10) + np.bincount(mesh.neighbour, np.ones(nf), 10)
    diag[0] += 1.0
    diag[9] += 1.0
    return mesh, diag, -np.ones(nf), -np.ones(nf)


def test_ldu_multiply_matches_dense(oracle):
    """LduMatrix.MultiplyMatchesDense (tests/test_ldu_matrix.cpp:94-127), with asymmetry."""
    mesh, diag, up, lo = _chain(oracle)
    up = up - 0.3
    x = np.sin(0.3 * np.arange(10)) + 0.1 * np.arange(10)
    dense = np.diag(diag)
    for f in range(mesh.n_faces):
        dense[mesh.owner[f], mesh.neighbour[f]] = up[f]
        dense[mesh.neighbour[f], mesh.owner[f]] = lo[f]
    y = oracle.ldu_multiply(mesh.owner, mesh.neighbour, diag, up, lo, x)
    np.testing.assert_allclose(y, dense @ x, rtol=0, atol=1e-12)


def test_normalised_residual_zero_at_solution_and_reference_pin(oracle):
    """NormalisedResidualZeroAtSolution and SetReferenceFixesValue
    (tests/test_ldu_matrix.cpp:171-197) on the pure-Neumann chain."""
    mesh, diag, up, lo = _chain(oracle)
    xe = np.sin(0.3 * np.arange(10))
    b = oracle.ldu_multiply(mesh.owner, mesh.neighbour, diag, up, lo, xe)
    assert oracle.ldu_normalised_residual(mesh.owner, mesh.neighbour, diag, up, lo, b, xe) < 1e-12
    # pure Neumann + set_reference(0, 2) => x == 2
    dn = np.bincount(mesh.owner, np.ones(9), 10) + np.bincount(mesh.neighbour, np.ones(9), 10)
    src = np.zeros(10)
    oracle.orc().orc_ldu_set_reference(oracle.dp(dn), oracle.dp(src), 0, 2.0)
    assert dn[0] == 2.0 and src[0] == 2.0  # diag doubled, source += diag * value
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(10, mesh.owner, mesh.neighbour)
    val = np.zeros(nnz)
    val[dpos], val[upos], val[lpos] = dn, up, lo
    x = np.zeros(10)
    st = oracle.solve(oracle.Csr(10, 10, rp, ci, val), src, x, 0, 1e-12, 1000)
    assert st.converged
    np.testing.assert_allclose(x, 2.0, atol=1e-6)


def test_csr_pattern_closed_form(oracle):
    """build_csr_pattern: sorted columns; for the box, row c holds c-nx*ny, c-nx, c-1, c, c+1,
    c+nx, c+nx*ny where they exist."""
    mesh = cases.box_mesh(4, 3, 2)
    n = mesh.n_cells
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(n, mesh.owner, mesh.neighbour)
    assert nnz == n + 2 * mesh.n_faces
    for c in range(n):
        i, j, k = c % 4, (c // 4) % 3, c // 12
        want = []
        if k > 0: want.append(c - 12)
        if j > 0: want.append(c - 4)
        if i > 0: want.append(c - 1)
        want.append(c)
        if i < 3: want.append(c + 1)
        if j < 2: want.append(c + 4)
        if k < 1: want.append(c + 12)
        assert list(ci[rp[c]:rp[c + 1]]) == want
        assert ci[dpos[c]] == c
    for f in range(mesh.n_faces):
        o, nb = mesh.owner[f], mesh.neighbour[f]
        assert rp[o] <= upos[f] < rp[o + 1] and ci[upos[f]] == nb
        assert rp[nb] <= lpos[f] < rp[nb + 1] and ci[lpos[f]] == o


def test_assembly_closed_form(oracle):
    """Gpu.AssemblyMatchesCpu's per-face closed form (tests/test_gpu.cpp:295-309)."""
    mesh = cases.box_mesh(7, 5, 3)
    nf, nc = mesh.n_faces, mesh.n_cells
    f = np.arange(nf)
    flux, gamma, dc = np.sin(1.3 * f), 0.1 + 0.01 * (f % 7), 2.0 + 0.1 * (f % 3)
    diag, div = np.ones(nc), np.zeros(nc)
    up, lo = oracle.assemble(mesh.owner, mesh.neighbour, flux, gamma, dc, 1.0, diag, div)
    g = gamma * dc
    assert np.array_equal(up, np.minimum(flux, 0.0) - g)
    assert np.array_equal(lo, -np.maximum(flux, 0.0) - g)
    d_ref, v_ref = np.ones(nc), np.zeros(nc)
    for k in range(nf):
        p, n = mesh.owner[k], mesh.neighbour[k]
        d_ref[p] += max(flux[k], 0.0) + g[k]
        d_ref[n] += -min(flux[k], 0.0) + g[k]
        v_ref[p] += flux[k]
        v_ref[n] -= flux[k]
    assert np.array_equal(diag, d_ref) and np.array_equal(div, v_ref)


# ------------------------------------------------------------------ aggregation (amg.cpp) --
def test_aggregation_chain_known_answer(oracle):
    """10-cell chain, upper = lower = -1, diag 2: roots 0, 3, 6, 9 (hand-derived from
    src/fv/src/amg.cpp:12-78)."""
    mesh, diag, up, lo = _chain(oracle)
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(10, mesh.owner, mesh.neighbour)
    val = np.zeros(nnz)
    val[dpos], val[upos], val[lpos] = diag, up, lo
    agg, n_agg = oracle.aggregate(oracle.Csr(10, 10, rp, ci, val), diag)
    assert n_agg == 4
    assert list(agg) == [0, 0, 1, 1, 1, 2, 2, 2, 3, 3]


def test_aggregation_3x3_known_answer(oracle):
    """3x3x1 box, unit coefficients, pure Neumann: root 0 takes {0,1,3}, root 5 takes
    {2,4,5,8}; pass 2 puts 6 with 0 (via 3) and 7 with 1 (via 4; 8 ties, strict >)."""
    mesh = cases.box_mesh(3, 3, 1)
    nf = mesh.n_faces
    diag = np.bincount(mesh.owner, np.ones(nf), 9) + np.bincount(mesh.neighbour, np.ones(nf), 9)
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(9, mesh.owner, mesh.neighbour)
    val = np.zeros(nnz)
    val[dpos], val[upos], val[lpos] = diag, -np.ones(nf), -np.ones(nf)
    agg, n_agg = oracle.aggregate(oracle.Csr(9, 9, rp, ci, val), diag)
    assert n_agg == 2
    assert list(agg) == [0, 0, 1, 0, 1, 1, 0, 1, 1]


def test_aggregation_threshold_and_isolated_nodes(oracle):
    """Weak links (|a_ij| <= theta sqrt(d_i d_j)) never aggregate: a diagonal matrix gives one
    singleton aggregate per row (pass 3); raising one link above the threshold pairs it."""
    n = 5
    m = sp.diags([np.full(n, 4.0)], [0]).tolil()
    m[1, 2] = m[2, 1] = -0.32          # 0.32 == 0.08 * sqrt(16): not strong (strict >)
    a = _csr_of(oracle, m)
    agg, n_agg = oracle.aggregate(a, np.full(n, 4.0))
    assert n_agg == n and list(agg) == list(range(n))
    # Pass 1 seeds row 1 (aggregate 0 = {1, 2}); rows 0, 3, 4 have no strong link and become
    # singletons in pass 3, numbered after the pass-1 aggregates.
    m[1, 2] = m[2, 1] = -0.33
    agg, n_agg = oracle.aggregate(_csr_of(oracle, m), np.full(n, 4.0))
    assert n_agg == 4 and list(agg) == [1, 0, 0, 2, 3]


# ------------------------------------------------- smoothed prolongation / Galerkin (dense) --
def _dense_level(a_dense, agg):
    """P = (I - omega D^-1 A) P_tent with omega = 4 / (3 rho), rho the Gershgorin bound
    max_i sum_j |a_ij| / |a_ii| (amg.cpp:150-178), computed densely."""
    n = a_dense.shape[0]
    na = agg.max() + 1
    pt = np.zeros((n, na))
    pt[np.arange(n), agg] = 1.0
    d = np.diag(a_dense)
    inv = np.where(d != 0.0, 1.0 / d, 0.0)
    rho = np.max(np.abs(a_dense).sum(axis=1) * np.abs(inv))
    omega = 4.0 / (3.0 * rho)
    p = pt - omega * (inv[:, None] * a_dense) @ pt
    return p, p.T @ a_dense @ p


@pytest.mark.parametrize("shape,jitter", [((6, 5, 1), 0.0), ((5, 4, 3), 0.2), ((8, 8, 1), 0.0)])
def test_setup_matches_dense_brute_force(oracle, shape, jitter):
    mesh, a = _box_csr(oracle, *shape, dirichlet=True, jitter=jitter, seed=3)
    h = oracle.Hierarchy(a, coarse_size=4)
    assert h.n_levels >= 2
    ad = a.to_scipy().toarray()
    for lvl in range(h.n_levels - 1):
        agg = h.aggregates(lvl)
        p_dense, ac_dense = _dense_level(ad, agg)
        p = h.matrix(lvl, 1).to_scipy().toarray()
        r = h.matrix(lvl, 2).to_scipy().toarray()
        ac = h.matrix(lvl + 1, 0).to_scipy().toarray()
        scale = np.abs(ad).max()
        np.testing.assert_allclose(p, p_dense, rtol=0, atol=1e-14)
        assert np.array_equal(r, p.T)                                   # R = P^T exactly
        np.testing.assert_allclose(ac, ac_dense, rtol=0, atol=1e-13 * scale)
        # pattern: the stored entries are exactly the structurally reachable non-zeros
        assert np.array_equal(p != 0.0, np.abs(p_dense) > 1e-300)
        assert np.array_equal(ac != 0.0, np.abs(ac_dense) > 1e-14 * scale)
        np.testing.assert_allclose(ac, ac.T, rtol=0, atol=1e-13 * scale)
        ad = ac


def test_hierarchy_invariants_64x64(oracle):
    """Amg.HierarchyIsBuilt (tests/test_amg.cpp:155-170)."""
    mesh, a = _box_csr(oracle, 64, 64, 1, dirichlet=True)
    h = oracle.Hierarchy(a)
    sizes = h.level_sizes()
    assert len(sizes) >= 3
    assert all(sizes[i] < sizes[i - 1] for i in range(1, len(sizes)))
    assert sizes[-1] <= 400
    assert 1.0 <= h.operator_complexity() <= 3.0


def test_coarse_inverse_is_inverse(oracle):
    mesh, a = _box_csr(oracle, 20, 20, 1, dirichlet=True)
    h = oracle.Hierarchy(a)
    ac = h.matrix(h.n_levels - 1).to_scipy().toarray()
    ci = h.coarse_inverse()
    np.testing.assert_allclose(ci @ ac, np.eye(ac.shape[0]), rtol=0, atol=1e-10)
    np.testing.assert_allclose(ci, np.linalg.inv(ac), rtol=1e-10, atol=1e-14 * np.abs(ci).max())


# --------------------------------------------------------------- Krylov (krylov_kernels.inc) --
def test_cg_and_bicgstab_reference_bounds(oracle):
    """Gpu.CgSolvesPoisson / BiCgStabSolvesNonSymmetric (tests/test_gpu.cpp:68-91)."""
    rp, ci, va = poisson_2d_csr(40, 0.0)
    a = oracle.Csr(1600, 1600, rp, ci, va)
    x = np.zeros(1600)
    st = oracle.solve(a, np.ones(1600), x, 0, 1e-10, 2000)
    assert st.converged and st.iterations > 10 and st.residual < 1e-10
    assert np.linalg.norm(np.ones(1600) - a.to_scipy() @ x) < 1e-9
    rp, ci, va = poisson_2d_csr(40, 0.4)
    a = oracle.Csr(1600, 1600, rp, ci, va)
    b = np.sin(0.1 * np.arange(1600))
    x = np.zeros(1600)
    st = oracle.solve(a, b, x, 1, 1e-10, 2000)
    assert st.converged and np.linalg.norm(b - a.to_scipy() @ x) < 1e-9


def test_amg_cg_fast_and_mesh_independent(oracle):
    """Gpu.AmgCgSolvesPoissonFast (tests/test_gpu.cpp:170-223)."""
    for m in (40, 80):
        rp, ci, va = poisson_2d_csr(m)
        a = oracle.Csr(m * m, m * m, rp, ci, va)
        h = oracle.Hierarchy(a)
        assert h.n_levels > 1
        b = np.ones(m * m)
        x = np.zeros(m * m)
        st = h.amg_cg(b, x, 1e-10, 500)
        assert st.converged and np.linalg.norm(b - a.to_scipy() @ x) < 1e-9
        xj = np.zeros(m * m)
        sj = oracle.solve(a, b, xj, 0, 1e-10, 5000)
        assert 3 * st.iterations < sj.iterations and st.iterations < 40


def test_sgs_pcg_reference_bounds(oracle):
    """Amg.PcgConvergesFastAndIndependentOfSize and Amg.Poisson3D (tests/test_amg.cpp:172-202)."""
    iters = []
    for shape in ((32, 32, 1), (64, 64, 1), (128, 128, 1), (16, 16, 16)):
        mesh, a = _box_csr(oracle, *shape, dirichlet=True)
        n = a.n_rows
        xe = np.sin(0.01 * np.arange(n)) + 1.0
        b = a.to_scipy() @ xe
        tol = 1e-8 * np.linalg.norm(b)
        h = oracle.Hierarchy(a)
        x = np.zeros(n)
        st = h.sgs_pcg(b, x, tol, 200)
        assert st.residual <= tol and st.iterations < 30
        assert np.abs(x - xe).max() < 1e-5
        iters.append(st.iterations)
    assert iters[2] <= iters[0] + 10


def _sgs_dense_operator(a, p, r, c):
    """M^-1 of one two-level SGS V-cycle from x = 0 (amg.cpp:192-225) formed from dense matrices:
    forward Gauss-Seidel x1 = (D + L)^-1 b, coarse correction x2 = x1 + P C^-1 R (b - A x1),
    backward Gauss-Seidel x3 = x2 + (D + U)^-1 (b - A x2); the coarse solve is exact."""
    import scipy.linalg as sl

    n = a.shape[0]
    eye = np.eye(n)
    f = sl.solve_triangular(np.tril(a), eye, lower=True)
    g = sl.solve_triangular(np.triu(a), eye, lower=False)
    k = p @ np.linalg.solve(c, r)
    t = f + k @ (eye - a @ f)
    return t + g @ (eye - a @ t), f, g, k


def test_sgs_vcycle_is_the_two_level_symmetric_gauss_seidel_operator(oracle):
    """The oracle's SGS V-cycle (orc_cycle) applied to every unit vector equals the dense
    operator of a two-level cycle with a forward pre-sweep, exact coarse solve and backward
    post-sweep, and is symmetric (<M^-1 x, y> = <x, M^-1 y>). Swapping the sweep directions or
    making both sweeps forward changes the operator by far more than the tolerance (checked)."""
    _, a = _box_csr(oracle, 24, 20, 1, dirichlet=True)
    h = oracle.Hierarchy(a)
    assert h.n_levels == 2
    ad = a.to_scipy().toarray()
    p = h.matrix(0, 1).to_scipy().toarray()
    r = h.matrix(0, 2).to_scipy().toarray()
    c = h.matrix(1).to_scipy().toarray()
    m_dense, f, g, k = _sgs_dense_operator(ad, p, r, c)
    n = a.n_rows
    m_orc = np.column_stack([h.sgs_apply(np.eye(n)[:, j]) for j in range(n)])
    scale = np.abs(m_dense).max()
    assert np.abs(m_orc - m_dense).max() <= 1e-12 * scale
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    assert abs(h.sgs_apply(x) @ y - x @ h.sgs_apply(y)) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y) * scale
    eye = np.eye(n)
    t_sw = g + k @ (eye - ad @ g)          # backward pre-sweep, forward post-sweep
    swapped = t_sw + f @ (eye - ad @ t_sw)
    t_ff = f + k @ (eye - ad @ f)          # forward both times
    forward = t_ff + f @ (eye - ad @ t_ff)
    assert np.abs(swapped - m_dense).max() > 1e-6 * scale
    assert np.abs(forward - m_dense).max() > 1e-6 * scale


def test_sgs_pcg_is_preconditioned_cg_with_that_operator(oracle):
    """orc_sgs_pcg (amg.cpp:253-285) against scipy's preconditioned CG with the dense V-cycle
    operator as preconditioner, from x = 0 and stopping at |r| <= tol: same iteration count
    (+-1) and solution to 1e-10."""
    import scipy.sparse.linalg as sla

    _, a = _box_csr(oracle, 24, 20, 1, dirichlet=True)
    h = oracle.Hierarchy(a)
    ad = a.to_scipy().toarray()
    m_dense, _, _, _ = _sgs_dense_operator(ad, h.matrix(0, 1).to_scipy().toarray(),
                                           h.matrix(0, 2).to_scipy().toarray(),
                                           h.matrix(1).to_scipy().toarray())
    n = a.n_rows
    b = np.cos(0.37 * np.arange(n))
    tol = 1e-10 * np.linalg.norm(b)
    x = np.zeros(n)
    st = h.sgs_pcg(b, x, tol, 100)
    count = []
    xs, info = sla.cg(a.to_scipy(), b, x0=np.zeros(n), rtol=0.0, atol=tol, maxiter=100,
                      M=sla.aslinearoperator(m_dense), callback=lambda _: count.append(1))
    assert info == 0 and st.converged
    assert abs(st.iterations - len(count)) <= 1
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)


def test_conservative_product_known_answers(oracle):
    """The setup's sparse product (amg.cpp:170, 176-177) adds the terms of an entry in the order
    k runs through row i of A, starting from the first product, and keeps exact zeros for the
    later prune(0.0). Hand-derived in binary64, where 1 + 1e16 = 1e16 and 1 - 1e16 = -1e16:
      C(0, 0): terms (1, 1e16, -1e16) -> (1 + 1e16) - 1e16 = 0, kept (reversed order gives 1);
      C(0, 1): terms (1e16, -1e16, 1) -> 1                       (reversed order gives 0);
      C(1, 0): terms (1, -1e16) -> -1e16;  C(1, 1): terms (1e16, 1) -> 1e16;
      C(2, :) = -1 x row 1 of B = (-1e16, 1e16);
    and a single product -1 x 0 is stored as -0.0 (an accumulator started at 0.0 would give +0.0)."""
    a = oracle.Csr(3, 3, np.array([0, 3, 5, 6], np.int32), np.array([0, 1, 2, 0, 2, 1], np.int32),
                   np.array([1.0, 1.0, 1.0, 1.0, 1.0, -1.0]))
    b = oracle.Csr(3, 2, np.array([0, 2, 4, 6], np.int32), np.array([0, 1, 0, 1, 0, 1], np.int32),
                   np.array([1.0, 1e16, 1e16, -1e16, -1e16, 1.0]))
    c = oracle.spgemm(a, b)
    assert list(c.rp) == [0, 2, 4, 6]
    assert list(c.ci) == [0, 1, 0, 1, 0, 1]
    assert c.va[0] == 0.0 and c.va[1] == 1.0
    assert c.va[2] == -1e16 and c.va[3] == 1e16
    assert c.va[4] == -1e16 and c.va[5] == 1e16
    one = np.array([0, 1], np.int32)
    zero = np.array([0], np.int32)
    z = oracle.spgemm(oracle.Csr(1, 1, one, zero, np.array([-1.0])), oracle.Csr(1, 1, one, zero, np.array([0.0])))
    assert z.nnz == 1 and z.va[0] == 0.0 and np.signbit(z.va[0])


def test_singular_neumann_with_reference(oracle):
    """Amg.SingularNeumannWithReference (tests/test_amg.cpp:254-265): x == 3."""
    mesh = cases.box_mesh(40, 40, 1)
    diag, up, lo = build_poisson_ldu(mesh, False)
    src = np.zeros(mesh.n_cells)
    diag = diag.copy()
    oracle.orc().orc_ldu_set_reference(oracle.dp(diag), oracle.dp(src), 0, 3.0)
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    val = np.zeros(nnz)
    val[dpos], val[upos], val[lpos] = diag, up, lo
    a = oracle.Csr(mesh.n_cells, mesh.n_cells, rp, ci, val)
    h = oracle.Hierarchy(a)
    x = np.zeros(mesh.n_cells)
    h.sgs_pcg(src, x, 1e-12 * max(np.linalg.norm(src), 1.0), 200)
    np.testing.assert_allclose(x, 3.0, atol=1e-6)
    x = np.zeros(mesh.n_cells)
    h.amg_cg(src, x, 1e-12 * np.linalg.norm(src), 200)
    np.testing.assert_allclose(x, 3.0, atol=1e-6)


# ------------------------------------------------ the reference's own host code (oracle/_ref) --
needs_ref = pytest.mark.skipif(
    not __import__("oracle.bind", fromlist=["ref_available"]).ref_available(),
    reason="oracle/_ref not built (reference checkout absent); golden fixtures cover this")


@needs_ref
@pytest.mark.parametrize("m", [40, 80])
def test_amg_cg_bitwise_equals_reference(oracle, m):
    """The oracle's Jacobi-V-cycle AMG-PCG is the reference's solve_amg_cg_host bit for bit
    (both serial: n <= 20000)."""
    rp, ci, va = poisson_2d_csr(m)
    a = oracle.Csr(m * m, m * m, rp, ci, va)
    h = oracle.Hierarchy(a)
    b = np.ones(m * m)
    x = np.zeros(m * m)
    st = h.amg_cg(b, x, 1e-10, 500)
    a_l, p_l, r_l = oracle.hierarchy_lists(h)
    xr = np.zeros(m * m)
    sr = oracle.ref_solve_amg_cg_host(a_l, p_l, r_l, h.coarse_inverse(), b, xr, 1e-10, 500)
    assert st.iterations == sr.iterations and st.residual == sr.residual
    assert np.array_equal(x, xr)


@needs_ref
@pytest.mark.parametrize("method,c", [(0, 0.0), (1, 0.4)])
def test_krylov_bitwise_equals_reference(oracle, method, c):
    rp, ci, va = poisson_2d_csr(40, c)
    a = oracle.Csr(1600, 1600, rp, ci, va)
    b = np.sin(0.1 * np.arange(1600)) + 0.5
    x = np.zeros(1600)
    st = oracle.solve(a, b, x, method, 1e-10, 2000)
    xr = np.zeros(1600)
    sr = oracle.ref_solve_host(a, b, xr, method, 1e-10, 2000)
    assert st.iterations == sr.iterations and st.residual == sr.residual
    assert np.array_equal(x, xr)


@needs_ref
def test_assembly_bitwise_equals_reference(oracle):
    mesh = cases.box_mesh(7, 5, 3)
    nf, nc = mesh.n_faces, mesh.n_cells
    f = np.arange(nf)
    flux, gamma, dc = np.sin(1.3 * f), 0.1 + 0.01 * (f % 7), 2.0 + 0.1 * (f % 3)
    d1, v1 = np.ones(nc), np.zeros(nc)
    u1, l1 = oracle.assemble(mesh.owner, mesh.neighbour, flux, gamma, dc, 0.7, d1, v1)
    d2, v2 = np.ones(nc), np.zeros(nc)
    u2, l2 = oracle.ref_assemble(mesh.owner, mesh.neighbour, flux, gamma, dc, 0.7, d2, v2)
    for p, q in ((u1, u2), (l1, l2), (d1, d2), (v1, v2)):
        assert np.array_equal(p, q)
