"""Smoothed-aggregation setup with the sparse products on the GPU (cfdg_amg_build_device)
against the host setup (cfdg_amg_build, pinned to the reference by tests/test_oracle_pins.py and
the golden fixtures): every level's operator, prolongation, restriction and aggregates are
bit-identical, and the coarsest inverse, computed on the matrix cores (up to order 512), equals
the host's LDL^T-based inverse to rounding (bit-identical where the host path is taken). Each
case runs with the aggregation on the host (the default) and with its first two passes as
parallel root rounds on the GPU (CFDG_DEVICE_AGGREGATION=all).
"""
from __future__ import annotations

import numpy as np
import pytest

from cfd_amd import AmgHierarchy, AmgSettings, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["host", "all"])
def aggregation(request, monkeypatch):
    if request.param == "host":
        monkeypatch.delenv("CFDG_DEVICE_AGGREGATION", raising=False)
    else:
        monkeypatch.setenv("CFDG_DEVICE_AGGREGATION", request.param)
    return request.param


def same_hierarchy(h, d):
    assert h.n_levels == d.n_levels
    for l in range(h.n_levels):
        for which in (0, 1, 2):
            if which and l == h.n_levels - 1:
                continue
            a, b = h.matrix(l, which), d.matrix(l, which)
            assert a.n == b.n and a.nnz == b.nnz, (l, which)
            assert np.array_equal(a.row_ptr, b.row_ptr), (l, which)
            assert np.array_equal(a.col, b.col), (l, which)
            assert np.array_equal(a.val, b.val), (l, which)
    hc, dc = h.view().coarse_inverse, d.view().coarse_inverse
    if d.coarse_on_device:
        # inverted on the matrix cores (block Gauss-Jordan) instead of by LDL^T solves: equal to
        # rounding, and an inverse of the same coarsest operator
        c = h.matrix(h.n_levels - 1).to_scipy().toarray()
        assert np.abs(dc - hc).max() <= 1e-11 * np.abs(hc).max()
        assert np.abs(c @ dc - np.eye(c.shape[0])).max() <= 1e-11
    else:
        assert np.array_equal(hc, dc)


@pytest.mark.parametrize("shape,jitter,decay", [((24, 20, 18), 0.2, 0.5), ((40, 40, 40), 0.0, 1.0),
                                                 ((64, 48, 30), 0.25, 0.5), ((7, 5, 3), 0.0, 0.5)])
def test_device_setup_matches_host_setup(aggregation, shape, jitter, decay):
    mesh = cases.box_mesh(*shape, hi=(1.0, 1.0, shape[2] / shape[0]), jitter=jitter, seed=sum(shape))
    s = cases.pressure_system(mesh, rhs_seed=3)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    settings = AmgSettings(strength_decay=decay)
    h = AmgHierarchy(a, settings)
    d = AmgHierarchy(a, settings, device=True)
    print(f"{shape} ({aggregation} aggregation): levels {h.level_sizes()}, host {h.setup_ms:.1f} ms, device {d.setup_ms:.1f} ms", flush=True)
    same_hierarchy(h, d)


def test_device_setup_of_a_variable_coefficient_operator(aggregation):
    mesh = cases.box_mesh(30, 30, 30, jitter=0.2, seed=9)
    rng = np.random.default_rng(2)
    kappa = np.exp(rng.uniform(-3.0, 3.0, mesh.n_cells))  # strongly varying cell coefficients
    s = cases.pressure_system(mesh, kappa=kappa, rhs_seed=4)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    settings = AmgSettings(strength_decay=0.5)
    same_hierarchy(AmgHierarchy(a, settings), AmgHierarchy(a, settings, device=True))
