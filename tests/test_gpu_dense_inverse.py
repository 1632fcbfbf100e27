"""Batched dense inversion on the FP64 matrix cores (cfdg_dense_inverse_batch; the coarsest-level
inverse of the AMG setup, amg.cpp:185-189, 227-240): one batch of symmetric positive definite
matrices of orders 1 to 512 (one workgroup each; 16-tile edges and their neighbours) and 513 to
2048 (the whole device, two launches per block step), and coarsest operators of benchmark
hierarchies, against numpy's LU inverse and the oracle's LDL^T inverse; orders above 2048 and a
matrix with an exactly zero pivot take the host path (the pseudo-inverse, as the reference);
solves with a device-built hierarchy against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from cfd_amd import AmgHierarchy, AmgSettings, SolveControl, cases, solve_amg_cg
from cfd_amd.amg import dense_inverse_batch

pytestmark = pytest.mark.gpu


def _spd(n, rng, cond=1e3):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    return (q * ev) @ q.T


def test_batch_of_spd_matrices_against_numpy():
    rng = np.random.default_rng(5)
    orders = [1, 2, 15, 16, 17, 31, 48, 100, 257, 400, 512, 513, 527, 944, 1636, 2048]
    mats = [_spd(n, rng) for n in orders]
    invs, on_dev = dense_inverse_batch(mats)
    for n, m, inv, dev in zip(orders, mats, invs, on_dev):
        ref = np.linalg.inv(m)
        err = np.abs(inv - ref).max() / np.abs(ref).max()
        res = np.abs(m @ inv - np.eye(n)).max()
        print(f"order {n}: device {dev}, rel. difference to numpy {err:.2e}, |A A^-1 - I| {res:.2e}")
        assert dev
        assert err <= 1e-12 and res <= 1e-11


def test_large_and_singular_matrices_take_the_host_path(oracle):
    rng = np.random.default_rng(6)
    big = _spd(2100, rng)
    sing = _spd(40, rng)
    sing[7, :] = 0.0
    sing[:, 7] = 0.0  # an exactly zero pivot: LDL^T fails, the reference takes the pseudo-inverse
    ok = _spd(33, rng)
    invs, on_dev = dense_inverse_batch([big, sing, ok])
    assert on_dev == [False, False, True]
    assert np.abs(big @ invs[0] - np.eye(2100)).max() <= 1e-11
    pinv = np.linalg.pinv(sing)
    assert np.abs(invs[1] - pinv).max() <= 1e-10 * np.abs(pinv).max()
    assert np.abs(ok @ invs[2] - np.eye(33)).max() <= 1e-11


@pytest.mark.parametrize("shape,decay", [((48, 48, 48), 0.5), ((32, 32, 32), 1.0), ((20, 20, 60), 0.5)])
def test_coarsest_operators_of_benchmark_hierarchies(oracle, shape, decay):
    """The device setup inverts its coarsest operator on the matrix cores; the inverse matches the
    oracle's LDL^T inverse to rounding and the FP64 solve with it matches the oracle's solve."""
    mesh = cases.box_mesh(*shape, hi=(1.0, 1.0, shape[2] / shape[0]), jitter=0.2, seed=sum(shape))
    s = cases.pressure_system(mesh, rhs_seed=2)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    d = AmgHierarchy(a, AmgSettings(strength_decay=decay), device=True)
    oh = oracle.Hierarchy(oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val), decay=decay)
    assert d.level_sizes() == oh.level_sizes()
    oc = oh.coarse_inverse()
    dc = d.view().coarse_inverse
    print(f"{shape}: levels {d.level_sizes()}, coarsest on device {d.coarse_on_device}, "
          f"rel. difference to the oracle {np.abs(dc - oc).max() / np.abs(oc).max():.2e}")
    assert d.coarse_on_device == (d.level_sizes()[-1] <= 2048)
    assert np.abs(dc - oc).max() <= 1e-11 * np.abs(oc).max()
    tol = 1e-8 * np.linalg.norm(s.source)
    x = np.zeros(s.n)
    st = solve_amg_cg(d.view(), s.source, x, SolveControl(tol, 500))
    xo = np.zeros(s.n)
    so = oh.amg_cg(s.source, xo, tol, 500)
    print(f"  FP64 AMG-PCG: gpu it {st.iterations} oracle it {so.iterations}, rel "
          f"{np.linalg.norm(x - xo) / np.linalg.norm(xo):.2e}")
    assert st.converged and abs(st.iterations - so.iterations) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10
