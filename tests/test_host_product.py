"""CPU tests of the product's host side: the C-ABI library, the case generator and the AMG
setup (bit-exact against the oracle). None of these call a device kernel."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, CsrView, SolveControl, cases
from cfd_amd import _lib as L

from .conftest import ROOT, build_poisson_ldu, has_gpu, poisson_2d_csr


# ----------------------------------------------------------------------------- C-ABI --
def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "cfdgpu.h").read_text()
    declared = set(re.findall(r"\b(cfdg_[a-z0-9_]+)\s*\(", header))
    assert len(declared) > 30
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [name for name in sorted(declared) if not hasattr(lib, name)]
    assert not missing, missing


def test_backend_name_and_description():
    assert cfd_amd.backend_name() == "hip"
    assert cfd_amd.device_description().startswith("hip")
    assert L.lib().cfdg_version() == 1


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_no_device_is_an_error_not_a_fallback():
    rp, ci, va = poisson_2d_csr(4)
    a = CsrView(16, len(va), rp, ci, va)
    with pytest.raises(L.NoDevice):
        cfd_amd.solve(a, np.ones(16), np.zeros(16), cfd_amd.Method.Cg, SolveControl())
    pat = cases.csr_pattern(8, np.array([0, 1, 2], np.int32), np.array([1, 2, 3], np.int32))
    with pytest.raises(L.NoDevice):
        cfd_amd.DeviceSolver(pat)
    with pytest.raises(L.NoDevice):
        cfd_amd.Server()
    with pytest.raises(L.NoDevice):
        cfd_amd.DeviceFv(cfd_amd.fv_box(3, 2, 2))


def test_footprint_estimate_bounds_the_stored_layout():
    """The admission estimate covers at least every stored slot (value, FP32 copy, BF16 copy,
    column offset, LDU source) and the 15 FP64 work vectors; it is host-only."""
    mesh = cases.box_mesh(12, 10, 9, jitter=0.2, seed=3)
    s = cases.pressure_system(mesh)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    plain = cfd_amd.estimate_footprint(pat)
    assert plain >= pat.nnz * (8 + 4 + 2 + 2 + 4) + 15 * 8 * s.n
    h = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5))
    sizes = h.level_sizes()
    with_amg = cfd_amd.estimate_footprint(pat, h.view())
    assert with_amg > plain + 8 * sizes[-1] ** 2  # coarse levels plus the dense coarsest inverse


def test_invalid_arguments_are_reported():
    with pytest.raises(L.InvalidArgument):
        cfd_amd.solve_amg_cg(cfd_amd.AmgView(), np.zeros(1), np.zeros(1), SolveControl())
    bad = CsrView(3, 2, np.array([0, 1, 2, 2]), np.array([0, 5]), np.array([1.0, 1.0]))
    with pytest.raises(L.InvalidArgument):
        AmgHierarchy(bad)
    with pytest.raises(L.InvalidArgument):
        cases.box_mesh(0, 1, 1)


# ------------------------------------------------------------------------ generator --
def test_uniform_box_coefficients_closed_form():
    m = cases.box_mesh(4, 3, 2, hi=(2.0, 1.5, 1.0))
    hx, hy, hz = 0.5, 0.5, 0.5
    assert m.n_cells == 24 and m.n_faces == 3 * 3 * 2 + 4 * 2 * 2 + 4 * 3 * 1
    d = m.neighbour - m.owner
    np.testing.assert_allclose(m.coeff[d == 1], hy * hz / hx, rtol=1e-14)
    np.testing.assert_allclose(m.coeff[d == 4], hx * hz / hy, rtol=1e-14)
    np.testing.assert_allclose(m.coeff[d == 12], hx * hy / hz, rtol=1e-14)
    np.testing.assert_allclose(m.weight, 0.5, rtol=1e-13)
    np.testing.assert_allclose(m.volume, hx * hy * hz, rtol=1e-13)
    i = np.arange(24) % 4
    np.testing.assert_allclose(m.centre[:, 0], (i + 0.5) * hx, rtol=1e-13)
    # faces ordered by (owner, neighbour) with owner < neighbour
    key = m.owner.astype(np.int64) * 100 + m.neighbour
    assert np.all(np.diff(key) > 0) and np.all(m.owner < m.neighbour)


def test_jittered_box_is_a_closed_partition():
    m = cases.box_mesh(10, 9, 8, jitter=0.2, seed=17)
    assert abs(m.volume.sum() - 1.0) < 1e-12
    assert m.volume.min() > 0.3 / 720 and m.coeff.min() > 0.0
    assert 0.0 < m.weight.min() and m.weight.max() < 1.0
    # the same seed reproduces the mesh exactly, another seed does not
    m2 = cases.box_mesh(10, 9, 8, jitter=0.2, seed=17)
    m3 = cases.box_mesh(10, 9, 8, jitter=0.2, seed=18)
    assert np.array_equal(m.coeff, m2.coeff) and not np.array_equal(m.coeff, m3.coeff)


def test_mt19937_64_known_answer():
    """The 10000th output of mt19937_64 with seed 5489 is 9981545732273789042 (C++11
    [rand.predef]); draws map to (u >> 11) 2^-53."""
    u = cases.uniform(5489, 10000, 0.0, 1.0)
    assert u[-1] == (9981545732273789042 >> 11) * 2.0 ** -53


def test_pressure_system_properties():
    mesh, s = cases.config1()
    assert s.n == 32768 and s.n_faces == 95232
    off = np.bincount(s.owner, -s.upper, s.n) + np.bincount(s.neighbour, -s.lower, s.n)
    assert s.diag[0] == 2.0 * off[0]                       # set_reference doubles diag[0]
    np.testing.assert_allclose(s.diag[1:], off[1:], rtol=1e-15)
    assert abs(s.source.sum()) < 1e-10                      # zero-mean right-hand side
    np.testing.assert_allclose(-s.upper, 1.0 / 32, rtol=1e-13)  # area / distance = h


def test_csr_pattern_matches_oracle(oracle):
    mesh = cases.box_mesh(9, 7, 5, jitter=0.2, seed=2)
    p = cases.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    nnz, rp, ci, dpos, upos, lpos = oracle.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    assert p.nnz == nnz
    for mine, theirs in ((p.row_ptr, rp), (p.col, ci), (p.diag_pos, dpos), (p.upper_pos, upos),
                         (p.lower_pos, lpos)):
        assert np.array_equal(mine, theirs)


# ----------------------------------------------------------- AMG setup (bit-exact) --
def _same_hierarchy(h: AmgHierarchy, oh) -> None:
    assert h.n_levels == oh.n_levels
    for lvl in range(h.n_levels):
        for which in (0, 1, 2):
            if which and lvl + 1 == h.n_levels:
                continue
            m, o = h.matrix(lvl, which), oh.matrix(lvl, which)
            assert np.array_equal(m.row_ptr, o.rp), (lvl, which)
            assert np.array_equal(m.col, o.ci), (lvl, which)
            assert np.array_equal(m.val, o.va), (lvl, which)      # bit-exact values
        if lvl + 1 < h.n_levels:
            assert np.array_equal(h.aggregates(lvl), oh.aggregates(lvl))
    np.testing.assert_allclose(h.coarse_inverse(), oh.coarse_inverse(), rtol=1e-12,
                               atol=1e-13 * np.abs(oh.coarse_inverse()).max())


@pytest.mark.parametrize("shape,jitter,decay,dirichlet", [
    ((64, 64, 1), 0.0, 1.0, True),
    ((32, 32, 32), 0.0, 1.0, False),
    ((24, 20, 16), 0.2, 1.0, False),
    ((40, 40, 40), 0.2, 0.5, False),
])
def test_setup_bit_exact_against_oracle(oracle, shape, jitter, decay, dirichlet):
    mesh = cases.box_mesh(*shape, jitter=jitter, seed=23)
    if dirichlet:
        diag, up, lo = build_poisson_ldu(mesh, True)
    else:
        s = cases.pressure_system(mesh)
        diag, up, lo = s.diag, s.upper, s.lower
    pat = cases.csr_pattern(mesh.n_cells, mesh.owner, mesh.neighbour)
    a = cases.ldu_to_csr(pat, diag, up, lo)
    h = AmgHierarchy(a, AmgSettings(strength_decay=decay))
    oh = oracle.Hierarchy(oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val), decay=decay)
    _same_hierarchy(h, oh)


def test_setup_bit_exact_on_the_128_cube_benchmark_system(oracle):
    """The hierarchy the single-system benchmark solves on: the 128^3 jittered pressure system
    (2,097,152 cells, threshold halved per level). Every level's P, R and A_c, patterns and
    values, and every aggregate map equal the oracle's bit for bit."""
    _, p, _ = cases.config2(128)
    pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
    a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    oh = oracle.Hierarchy(oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val), decay=0.5)
    assert h.level_sizes() == [2097152, 263534, 14687, 488, 22]
    _same_hierarchy(h, oh)


def test_setup_bit_exact_on_variable_coefficients(oracle):
    """Config-3-like contrast (rho 1000 sphere, harmonic faces) on a small jittered box."""
    mesh = cases.box_mesh(30, 30, 30, jitter=0.2, seed=5)
    s = cases.pressure_system(mesh, kappa=cases.sphere_kappa(mesh, 7))
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    oh = oracle.Hierarchy(oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val), decay=0.5)
    _same_hierarchy(h, oh)


def test_setup_unsorted_input_is_sorted():
    rp, ci, va = poisson_2d_csr(24)
    n = 576
    a = CsrView(n, len(va), rp, ci, va)
    h1 = AmgHierarchy(a)
    ci2, va2 = ci.copy(), va.copy()
    for i in range(n):   # reverse every row
        s, e = rp[i], rp[i + 1]
        ci2[s:e], va2[s:e] = ci[s:e][::-1], va[s:e][::-1]
    h2 = AmgHierarchy(CsrView(n, len(va), rp, ci2, va2))
    assert h1.n_levels > 1 and h1.level_sizes() == h2.level_sizes()
    assert np.array_equal(h1.matrix(1).val, h2.matrix(1).val)


def test_layout_rules_and_built_in_choice():
    """The built-in layout rule (no device needed) and a tuned rule overriding it."""
    def layout(rp):
        lanes, csr = C.c_int(), C.c_int()
        L.call("cfdg_layout_of", len(rp) - 1, L.iptr(np.asarray(rp, np.int32)), C.byref(lanes), C.byref(csr))
        return lanes.value, csr.value

    short = np.arange(0, 7 * 1000 + 1, 7)     # 1000 rows of 7 entries: one lane per row, ELL
    long_ = np.arange(0, 31 * 1000 + 1, 31)   # 31 entries: lanes for about 4 entries each
    L.call("cfdg_set_layout_rules", None, 0)
    assert layout(short) == (1, 0)
    assert layout(long_) == (4, 0)
    rule = (L.LayoutRule * 1)(L.LayoutRule(500, 2000, 20.0, 40.0, 16, 1))
    try:
        L.call("cfdg_set_layout_rules", rule, 1)
        assert layout(long_) == (16, 1)
        assert layout(short) == (1, 0)  # outside the rule's class
    finally:
        cfd_amd._apply_tuning()  # back to the table installed at import (if any)
    bad = (L.LayoutRule * 1)(L.LayoutRule(0, 10, 0.0, 1.0, 3, 0))  # lanes not a power of two
    with pytest.raises(L.InvalidArgument):
        L.call("cfdg_set_layout_rules", bad, 1)
    with pytest.raises(L.InvalidArgument):
        L.call("cfdg_set_mixed_smoother", 1, 0.5)


def test_tuning_table_is_well_formed():
    import json

    path = ROOT / "cfd_amd" / "tuning.json"
    if not path.exists():
        pytest.skip("no tuning table")
    t = json.loads(path.read_text())
    for r in t["layout_rules"]:
        assert r["min_rows"] <= r["max_rows"] and r["min_mean"] <= r["max_mean"]
        assert r["lanes"] in (1, 2, 4, 8, 16, 32, 64) and r["csr"] in (0, 1)
    assert t["mixed_smoother"]["sweeps"] in (1, 2)


def test_singular_coarsest_operator_falls_back_to_the_pseudo_inverse():
    """A cell with no faces and a zero diagonal leaves an empty row in every coarse operator, so
    the coarsest LDL^T meets an exactly zero pivot; the setup then returns the Moore-Penrose
    pseudo-inverse (amg.cpp:185-189, 227-240: Eigen's complete orthogonal decomposition)."""
    from .test_gpu_failure_paths import _singular_coarse_system

    h = AmgHierarchy(_singular_coarse_system())
    c = h.matrix(h.n_levels - 1).to_scipy().toarray()
    assert h.n_levels == 2 and np.linalg.matrix_rank(c) == c.shape[0] - 1
    pinv = np.linalg.pinv(c)
    assert np.abs(h.coarse_inverse() - pinv).max() <= 1e-10 * np.abs(pinv).max()
