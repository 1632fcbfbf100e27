"""The caller policy on the device (cfdg_amg_cache_* / cfdg_ldu_solve, cfd_amd.AmgCache) against
the oracle's restatement of LduMatrix::solve with a pattern and an AmgCache
(src/fv/src/ldu_matrix.cpp:111-131, 287-366; stale rule :242-256), call by call: whether it
solved, whether AMG ran, the hierarchies built, the cache's age, the initial normalised residual
and the mapped target (1e-12 relative), the iterations (+-1) and the solution (1e-10 relative,
or within the residual bound where a stale hierarchy amplifies rounding differences);
and the device normalised residual (cfdg_system_normalised_residual, ldu_matrix.cpp:65-103)
against the oracle's. The cache rebuilds with the device setup (its default) unless a test asks
for the host setup."""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgCache, LinearSolverSettings, Method, Precision, cases

from .conftest import build_poisson_ldu
from .test_policy_oracle import _manufactured, stale_sequence

pytestmark = pytest.mark.gpu


def _rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


def _lambda_min(a) -> float:
    import scipy.sparse.linalg as sla

    m = a.to_scipy().tocsc()
    return float(sla.eigsh(m, k=1, sigma=0, which="LM", return_eigenvectors=False)[0])


def _run(oracle, mesh, seq, method, tol, rel, max_it, max_age=20, stale=False,
         precision=Precision.FP64, device_setup=True):
    pat = cases.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    opat = oracle.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    cache_o = oracle.AmgCacheState(max_age=max_age, stale_rebuild=stale)
    outs = []
    with AmgCache(pat, max_age=max_age, stale_rebuild=stale, device_setup=device_setup) as cache:
        for k, (diag, up, lo, b, x0) in enumerate(seq):
            x = x0.copy()
            r = cache.solve(diag, up, lo, b, x, LinearSolverSettings(method, tol, rel, max_it, precision))
            xo = x0.copy()
            o = oracle.ldu_solve(cache_o, mesh.owner, mesh.neighbour, opat, diag, up, lo, b, xo, int(method),
                                 tol, rel, max_it)
            label = f"call {k}"
            print(f"{label}: solved {r.solved}/{o['solved']} amg {r.used_amg}/{o['used_amg']} setups "
                  f"{r.setups}/{o['setups']} it {r.iterations}/{o['iterations']} initial "
                  f"{r.initial_residual:.6e}/{o['initial']:.6e} rel x {_rel(x, xo):.2e}")
            assert (r.solved, r.used_amg, r.setups) == (o["solved"], o["used_amg"], o["setups"]), label
            # absolute 1e-14: at a converged start the normalised residual is rounding itself
            assert r.initial_residual == pytest.approx(o["initial"], rel=1e-12, abs=1e-14), label
            age, levels, setups = cache.state()
            assert (age, setups) == (cache_o.age, cache_o.setups), label
            if cache_o.h is not None:
                assert levels == cache_o.h.n_levels, label
            if o["solved"]:
                assert r.target == pytest.approx(o["target"], rel=1e-12), label
                if precision == Precision.FP64:
                    assert abs(r.iterations - o["iterations"]) <= 1, label
                    if _rel(x, xo) > 1e-10:
                        # A stale hierarchy amplifies the rounding differences of the two
                        # solvers' sums (2e-10 after six solves on one hierarchy here); the two
                        # converged solutions are then held to A (x - xo) = ro - r:
                        # |x - xo| <= (|r| + |ro|) / lambda_min(A).
                        a = oracle.ldu_csr(opat, diag, up, lo)
                        res = lambda v: float(np.linalg.norm(b - oracle.spmv(a, v)))
                        bound = (res(x) + res(xo)) / (0.9 * _lambda_min(a))
                        print(f"{label}: |x - xo| {np.linalg.norm(x - xo):.3e} bound {bound:.3e}")
                        assert np.linalg.norm(x - xo) <= bound, label
            else:
                assert r.iterations == 0 and np.array_equal(x, x0), label
            outs.append((r, x))
    return outs


def test_amg_cache_reuse(oracle):
    """Amg.AmgCacheReuse (tests/test_amg.cpp:232-252) with GpuCg: five solves on one cache,
    each within 1e-6 of its manufactured solution; age 5 and one hierarchy at the end."""
    mesh = cases.box_mesh(40, 40, 1)
    diag, up, lo = build_poisson_ldu(mesh, True)
    seq, xes = [], []
    for k in range(5):
        xe = _manufactured(mesh.n_cells, 0.01 * (k + 1))
        b = oracle.ldu_multiply(mesh.owner, mesh.neighbour, diag, up, lo, xe)
        seq.append((diag, up, lo, b, np.zeros(mesh.n_cells)))
        xes.append(xe)
    outs = _run(oracle, mesh, seq, Method.Cg, 1e-12, 0.0, 200)
    for (r, x), xe in zip(outs, xes):
        assert r.used_amg and np.abs(x - xe).max() < 1e-6


@pytest.mark.parametrize("device_setup", [True, False])
def test_rebuild_every_max_age_solves(oracle, device_setup):
    """Rebuilds at age 20 with the setup on the GPU (the default) and on the host."""
    mesh = cases.box_mesh(20, 20, 2)
    diag0, up, lo = build_poisson_ldu(mesh, True)
    seq = [(diag0 + 0.01 * k, up, lo, np.cos(0.1 * np.arange(mesh.n_cells) + k), np.zeros(mesh.n_cells))
           for k in range(45)]
    outs = _run(oracle, mesh, seq, Method.Cg, 1e-10, 0.0, 500, device_setup=device_setup)
    assert [k for k, (r, _) in enumerate(outs) if r.setups] == [0, 20, 40]


def test_one_level_bicgstab_and_converged_start(oracle):
    """A one-level hierarchy runs Jacobi CG (ldu_matrix.cpp:358); GpuBiCgStab never builds a
    hierarchy; a start below the tolerance returns before any solve (:121-124)."""
    small = cases.box_mesh(15, 15, 1)
    d, u, l = build_poisson_ldu(small, True)
    b = np.cos(0.1 * np.arange(small.n_cells))
    outs = _run(oracle, small, [(d, u, l, b, np.zeros(small.n_cells))], Method.Cg, 1e-10, 0.0, 500)
    assert not outs[0][0].used_amg and outs[0][0].setups == 1
    mesh = cases.box_mesh(20, 20, 2)
    d, u, l = build_poisson_ldu(mesh, True)
    xe = _manufactured(mesh.n_cells, 0.03)
    be = oracle.ldu_multiply(mesh.owner, mesh.neighbour, d, u, l, xe)
    b = np.cos(0.1 * np.arange(mesh.n_cells))
    x0 = 0.2 * np.sin(0.01 * np.arange(mesh.n_cells))
    _run(oracle, mesh, [(d, u, l, b, x0), (d, u, l, be, xe)], Method.BiCgStab, 1e-10, 0.0, 500)
    outs = _run(oracle, mesh, [(d, u, l, be, xe)], Method.Cg, 1e-10, 0.0, 500)
    assert not outs[0][0].solved


@pytest.mark.parametrize("tol,rel", [(1e-6, 0.05), (1e-3, 1e-4)])
def test_tolerance_mapping_from_a_nonzero_guess(oracle, tol, rel):
    mesh = cases.box_mesh(24, 20, 12, jitter=0.2, seed=4)
    s = cases.pressure_system(mesh, rhs_seed=5)
    x0 = 0.3 * np.sin(0.05 * np.arange(mesh.n_cells))
    _run(oracle, mesh, [(s.diag, s.upper, s.lower, s.source, x0)] * 2, Method.Cg, tol, rel, 500)


@pytest.mark.parametrize("stale", [False, True])
def test_stale_rule(oracle, stale):
    mesh, seq = stale_sequence()
    outs = _run(oracle, mesh, [(d, u, l, b, np.zeros(mesh.n_cells)) for d, u, l, b in seq], Method.Cg,
                1e-10, 0.0, 2000, stale=stale)
    assert outs[1][0].setups == (1 if stale else 0)


def test_mixed_precision_through_the_cache(oracle):
    """The V-cycle precision of the policy's AMG-PCG: mixed FP32 meets the same mapped target."""
    mesh = cases.box_mesh(32, 30, 28, jitter=0.2, seed=9)
    s = cases.pressure_system(mesh, rhs_seed=10)
    outs = _run(oracle, mesh, [(s.diag, s.upper, s.lower, s.source, np.zeros(mesh.n_cells))], Method.Cg,
                1e-10, 0.0, 500, precision=Precision.MIXED_FP32)
    r, x = outs[0]
    res = np.linalg.norm(s.source - oracle.ldu_multiply(mesh.owner, mesh.neighbour, s.diag, s.upper, s.lower, x))
    assert r.used_amg and r.stats.converged and res <= r.target * 1.05


def test_device_normalised_residual_matches_the_oracle(oracle):
    mesh = cases.box_mesh(37, 29, 23, jitter=0.2, seed=2)
    s = cases.pressure_system(mesh, rhs_seed=3)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    rng = np.random.default_rng(7)
    with cfd_amd.DeviceSolver(pat) as dev:
        dev.assemble(s.diag, s.upper, s.lower)
        dev.set_rhs(s.source)
        for x in (np.zeros(s.n), rng.uniform(-1, 1, s.n), np.full(s.n, 2.5)):
            dev.set_x(x)
            v, res, norm = dev.normalised_residual()
            vo = oracle.ldu_normalised_residual(s.owner, s.neighbour, s.diag, s.upper, s.lower, s.source, x)
            r = s.source - oracle.ldu_multiply(s.owner, s.neighbour, s.diag, s.upper, s.lower, x)
            print(f"normalised residual {v:.15e} oracle {vo:.15e} |r|_1 {res:.6e} ({np.abs(r).sum():.6e})")
            assert v == pytest.approx(vo, rel=1e-12)
            assert res == pytest.approx(np.abs(r).sum(), rel=1e-12)
