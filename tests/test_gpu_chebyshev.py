"""Mixed-precision smoothers (cfdg_set_mixed_smoother): Chebyshev smoothers of degree 1 to 3
(degree k: k damped-Jacobi sweeps per side weighted by the reciprocals of the Chebyshev
polynomial's roots over [h / ratio, h], h the top of each level's D^-1 A spectrum) reach the
tolerance on a jittered pressure system and on batch-shaped systems, against the FP64 oracle's
solution within the tolerance-implied bound; higher degrees need no more iterations; and ratio 0
(the hierarchy's own weight) with two sweeps is the FP64 path's V-cycle in FP32, which needs the
FP64 solve's iterations (-1 to +2: flexible CG around an FP32 cycle)."""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases
from cfd_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_smoother():
    yield
    cfd_amd._apply_tuning()  # the installed table's smoother


def _system(shape, jitter, seed, lz=1.0):
    mesh = cases.box_mesh(*shape, hi=(1.0, 1.0, lz), jitter=jitter, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed + 1)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    return s, pat, a, AmgHierarchy(a, AmgSettings(strength_decay=0.5))


def _solve(s, pat, h, prec, sweeps, ratio, tol):
    L.call("cfdg_set_mixed_smoother", sweeps, ratio)  # applies to hierarchies set afterwards
    with cfd_amd.DeviceSolver(pat) as dev:
        dev.assemble(s.diag, s.upper, s.lower)
        dev.set_amg(h.view())
        dev.set_rhs(s.source)
        dev.set_x(None)
        st = dev.solve_resident(Method.Cg, True, prec, SolveControl(tol, 500))
        return st, dev.get_x()


def _lambda_min(a) -> float:
    import scipy.sparse.linalg as sla

    return float(sla.eigsh(a.to_scipy().tocsc(), k=1, sigma=0, which="LM", return_eigenvectors=False)[0])


@pytest.mark.parametrize("shape,jitter,lz", [((40, 40, 40), 0.2, 1.0), ((18, 18, 70), 0.0, 8.0)])
def test_chebyshev_degrees_converge_against_the_oracle(oracle, restore_smoother, shape, jitter, lz):
    s, pat, a, h = _system(shape, jitter, 31, lz)
    tol = 1e-8 * float(np.linalg.norm(s.source))
    oa = oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val)
    xo = np.zeros(s.n)
    so = oracle.Hierarchy(oa, decay=0.5).amg_cg(s.source, xo, tol, 500)
    lam = _lambda_min(a)
    res = lambda x: float(np.linalg.norm(s.source - oracle.spmv(oa, x)))
    its = {}
    for degree, ratio in ((1, 6.0), (2, 8.0), (3, 12.0)):
        st, x = _solve(s, pat, h, Precision.MIXED_FP32, degree, ratio, tol)
        its[degree] = st.iterations
        bound = (res(x) + res(xo)) / (0.9 * lam)
        print(f"{shape} degree {degree} ratio {ratio}: it {st.iterations} (oracle FP64 {so.iterations}), "
              f"|x - xo| {np.linalg.norm(x - xo):.2e} bound {bound:.2e}")
        assert st.converged and res(x) <= 1.05 * tol
        assert np.linalg.norm(x - xo) <= bound
    assert its[2] <= its[1] and its[3] <= its[2] + 1


def test_chebyshev_against_the_reference_weight_mode(monkeypatch, restore_smoother):
    monkeypatch.setenv("CFDG_DENSE_TAIL", "0")  # the FP32 cycle keeps every level of the FP64 one
    s, pat, _, h = _system((48, 48, 48), 0.2, 5)
    tol = 1e-8 * float(np.linalg.norm(s.source))
    st64, _ = _solve(s, pat, h, Precision.FP64, 1, 6.0, tol)       # FP64: the reference cycle
    st_ref, _ = _solve(s, pat, h, Precision.MIXED_FP32, 2, 0.0, tol)  # the same cycle in FP32
    st_cheb, _ = _solve(s, pat, h, Precision.MIXED_FP32, 2, 8.0, tol)
    print(f"FP64 {st64.iterations}, FP32 reference cycle {st_ref.iterations}, degree-2 Chebyshev "
          f"{st_cheb.iterations}")
    assert st_ref.converged and st_cheb.converged
    # flexible CG around an FP32 cycle against standard PCG around the FP64 one
    assert st64.iterations - 1 <= st_ref.iterations <= st64.iterations + 2
    assert st_cheb.iterations <= st_ref.iterations + 1
