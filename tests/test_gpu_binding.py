"""The reference-side binding of the integration guide, compiled
against the reference's own cfd::gpu header (oracle/build_shim.py -> oracle/_ref/libshimtest.so)
and called through the reference's types (CsrView, AmgView, LduPattern, SolveControl,
SolveStats): the counterparts of Gpu.CgSolvesPoisson, BiCgStabSolvesNonSymmetric,
ZeroIterationsWhenConverged, AmgCgSolvesPoissonFast, PersistentSystemRepeatedSolves and
AssemblyMatchesCpu (tests/test_gpu.cpp:68-329) against the oracle, results identical to the
direct C-ABI calls, and the binding's exception mapping (krylov_cuda.cu:18-22)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, CsrView, Method, SolveControl, cases
from cfd_amd import _lib as L

from .conftest import poisson_2d_csr

SHIM = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "libshimtest.so"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not SHIM.exists(), reason="binding not built (needs the reference header)")]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("on_device", C.c_int),
                ("residual", C.c_double)]


P = C.c_void_p


def _shim():
    L.lib()  # libcfdgpu first: the binding links against it
    s = C.CDLL(str(SHIM))
    s.shim_solve.argtypes = [C.c_int, C.c_int, P, P, P, P, P, C.c_int, C.c_double, C.c_int, C.POINTER(Stats)]
    s.shim_solve_amg.argtypes = [C.c_int, P, P, P, P, P, P, C.c_int, C.c_double, P, P, C.c_double, C.c_int,
                                 C.POINTER(Stats)]
    s.shim_assemble.argtypes = [C.c_int, C.c_int, P, P, P, P, P, C.c_double, P, P, P, P]
    s.shim_system_create.argtypes = [C.c_int, C.c_int, C.c_int, P, P, P, P, P, C.POINTER(C.c_void_p)]
    s.shim_system_set_amg.argtypes = [P, C.c_int, P, P, P, P, P, P, C.c_int, C.c_double]
    s.shim_system_solve.argtypes = [P, P, P, P, P, P, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(Stats)]
    s.shim_system_destroy.argtypes = [P]
    s.shim_device_name.argtypes = [C.c_char_p, C.c_int]
    return s


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _Hierarchy:
    """Per-level arrays of a hierarchy as the driver takes them (a, p, r per level)."""

    def __init__(self, h: AmgHierarchy):
        self.keep = []
        n_levels = h.n_levels
        rows, nnz, rp, ci, va = [], [], [], [], []
        for lvl in range(n_levels):
            for which in range(3):
                if which > 0 and lvl + 1 == n_levels:
                    m = None
                else:
                    m = h.matrix(lvl, which)
                if m is None:
                    arr = (np.zeros(1, np.int32), np.zeros(1, np.int32), np.zeros(1))
                    rows.append(0)
                    nnz.append(0)
                else:
                    arr = (np.ascontiguousarray(m.row_ptr, np.int32), np.ascontiguousarray(m.col, np.int32),
                           np.ascontiguousarray(m.val, np.float64))
                    rows.append(m.n)
                    nnz.append(m.nnz)
                self.keep.append(arr)
                rp.append(arr[0].ctypes.data)
                ci.append(arr[1].ctypes.data)
                va.append(arr[2].ctypes.data)
        self.n_levels = n_levels
        self.rows = np.array(rows, np.int32)
        self.nnz = np.array(nnz, np.int32)
        self.rp = (C.c_void_p * len(rp))(*rp)
        self.ci = (C.c_void_p * len(ci))(*ci)
        self.va = (C.c_void_p * len(va))(*va)
        self.cinv = np.ascontiguousarray(h.coarse_inverse(), np.float64)

    def args(self):
        return (self.n_levels, _p(self.rows), _p(self.nnz), C.cast(self.rp, C.c_void_p),
                C.cast(self.ci, C.c_void_p), C.cast(self.va, C.c_void_p), _p(self.cinv))


def test_device_name_through_the_binding():
    s = _shim()
    buf = C.create_string_buffer(256)
    assert s.shim_device_name(buf, 256) == 0
    assert buf.value.decode().startswith("hip")


@pytest.mark.parametrize("method,conv", [(Method.Cg, 0.0), (Method.BiCgStab, 0.3)])
def test_krylov_through_the_binding(oracle, method, conv):
    """Gpu.CgSolvesPoisson / BiCgStabSolvesNonSymmetric: 40^2 five-point operator, b = 1, to
    1e-8 relative. BiCGStab's iteration count on this case follows rounding (a 1e-15 relative
    perturbation of b moves the oracle from 70 to 67 iterations), so for it the oracle bounds the
    solution only, as the reference test does; the binding equals the direct call bit for bit."""
    s = _shim()
    rp, ci, va = poisson_2d_csr(40, conv)
    n = 1600
    b = np.ones(n)
    tol = 1e-8 * np.linalg.norm(b)
    x = np.zeros(n)
    st = Stats()
    assert s.shim_solve(n, len(va), _p(rp), _p(ci), _p(va), _p(b), _p(x), int(method), tol, 2000,
                        C.byref(st)) == 0
    xd = np.zeros(n)
    sd = cfd_amd.solve(CsrView(n, len(va), rp, ci, va), b, xd, method, SolveControl(tol, 2000))
    xo = np.zeros(n)
    so = oracle.solve(oracle.Csr(n, n, rp, ci, va), b, xo, int(method), tol, 2000)
    print(f"{method.name}: binding it {st.iterations} direct {sd.iterations} oracle {so.iterations}")
    assert st.converged and st.on_device and st.iterations == sd.iterations and np.array_equal(x, xd)
    assert so.converged
    if method == Method.Cg:
        assert abs(st.iterations - so.iterations) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= (1e-10 if method == Method.Cg else 1e-6)


def test_zero_iterations_when_converged_through_the_binding():
    s = _shim()
    rp, ci, va = poisson_2d_csr(10)
    x = np.ones(100)
    b = np.add.reduceat(va, rp[:-1])
    st = Stats()
    assert s.shim_solve(100, len(va), _p(rp), _p(ci), _p(va), _p(b), _p(x), 0, 1e-12, 10, C.byref(st)) == 0
    assert st.converged and st.iterations == 0


def test_amg_cg_through_the_binding(oracle):
    """Gpu.AmgCgSolvesPoissonFast on a 32^3 jittered pressure system: the binding's
    solve_amg_cg equals the direct C-ABI solve bit for bit and the oracle within 1e-10."""
    s = _shim()
    mesh = cases.box_mesh(32, 32, 32, jitter=0.2, seed=11)
    p = cases.pressure_system(mesh, rhs_seed=12)
    pat = cases.csr_pattern(p.n, p.owner, p.neighbour)
    a = cases.ldu_to_csr(pat, p.diag, p.upper, p.lower)
    h = AmgHierarchy(a, AmgSettings(strength_decay=0.5))
    hv = _Hierarchy(h)
    tol = 1e-8 * np.linalg.norm(p.source)
    x = np.zeros(p.n)
    st = Stats()
    assert s.shim_solve_amg(*hv.args(), 2, 0.7, _p(p.source), _p(x), tol, 500, C.byref(st)) == 0
    xd = np.zeros(p.n)
    sd = cfd_amd.solve_amg_cg(h.view(), p.source, xd, SolveControl(tol, 500))
    oh = oracle.Hierarchy(oracle.Csr(a.n, a.n, a.row_ptr, a.col, a.val), decay=0.5)
    xo = np.zeros(p.n)
    so = oh.amg_cg(p.source, xo, tol, 500)
    print(f"AMG-CG: binding it {st.iterations} direct {sd.iterations} oracle {so.iterations}")
    assert st.converged and st.iterations == sd.iterations and np.array_equal(x, xd)
    assert abs(st.iterations - so.iterations) <= 1
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10


def test_persistent_system_through_the_binding(oracle):
    """Gpu.PersistentSystemRepeatedSolves (tests/test_gpu.cpp:229-268) through the binding's
    DeviceSolver functions: one hierarchy, three rounds of new diagonals; GpuCg with AMG and
    GpuBiCgStab without, as the reference's caller sets use_amg (ldu_matrix.cpp:358)."""
    s = _shim()
    mesh = cases.box_mesh(16, 16, 1)
    n, nf = mesh.n_cells, mesh.n_faces
    pat = cases.csr_pattern(n, mesh.owner, mesh.neighbour)
    h = C.c_void_p()
    assert s.shim_system_create(n, nf, pat.nnz, _p(pat.row_ptr), _p(pat.col), _p(pat.diag_pos),
                                _p(pat.upper_pos), _p(pat.lower_pos), C.byref(h)) == 0
    try:
        hier = None
        for rnd in range(3):
            diag = np.bincount(mesh.owner, np.ones(nf), n) + np.bincount(mesh.neighbour, np.ones(nf), n)
            diag = diag + 0.01 * (rnd + 1)
            up, lo = -np.ones(nf), -np.ones(nf)
            src = np.sin(0.3 * np.arange(n))
            a = cases.ldu_to_csr(pat, diag, up, lo)
            if hier is None:
                hier = AmgHierarchy(a)
                hv = _Hierarchy(hier)
                assert s.shim_system_set_amg(h, *hv.args(), 2, 0.7) == 0
            oa = oracle.Csr(n, n, a.row_ptr, a.col, a.val)
            m0 = hier.matrix(0)
            oh = oracle.Hierarchy(oracle.Csr(n, n, m0.row_ptr, m0.col, m0.val))
            for method in (Method.Cg, Method.BiCgStab):
                tol = 1e-8 * np.linalg.norm(src)
                x = np.zeros(n)
                st = Stats()
                assert s.shim_system_solve(h, _p(diag), _p(up), _p(lo), _p(src), _p(x), int(method),
                                           int(method == Method.Cg), tol, 1000, C.byref(st)) == 0
                ref = np.zeros(n)
                so = (oh.amg_cg(src, ref, tol, 1000, level0=oa) if method == Method.Cg
                      else oracle.solve(oa, src, ref, 1, tol, 1000))
                print(f"round {rnd} {method.name}: binding it {st.iterations} oracle {so.iterations}")
                assert st.converged and st.on_device and abs(st.iterations - so.iterations) <= 1
                np.testing.assert_allclose(x, ref, rtol=0, atol=1e-7 * (1 + np.abs(ref).max()))
    finally:
        s.shim_system_destroy(h)


def test_assembly_through_the_binding(oracle):
    """Gpu.AssemblyMatchesCpu (tests/test_gpu.cpp:274-329): bit-identical to the oracle."""
    s = _shim()
    mesh = cases.box_mesh(7, 5, 3)
    n, nf = mesh.n_cells, mesh.n_faces
    rng = np.random.default_rng(3)
    flux, gamma, delta = rng.uniform(-1, 1, nf), rng.uniform(0.1, 1, nf), rng.uniform(0.5, 2, nf)
    d0, div0 = rng.uniform(0, 1, n), rng.uniform(-1, 1, n)
    up, lo, dg, dv = np.zeros(nf), np.zeros(nf), d0.copy(), div0.copy()
    assert s.shim_assemble(n, nf, _p(mesh.owner), _p(mesh.neighbour), _p(flux), _p(gamma), _p(delta), 1.0,
                           _p(up), _p(lo), _p(dg), _p(dv)) == 0
    dg_o, dv_o = d0.copy(), div0.copy()
    up_o, lo_o = oracle.assemble(mesh.owner, mesh.neighbour, flux, gamma, delta, 1.0, dg_o, dv_o)
    assert np.array_equal(up, up_o) and np.array_equal(lo, lo_o)
    assert np.array_equal(dg, dg_o) and np.array_equal(dv, dv_o)


def test_binding_maps_errors_to_the_reference_exceptions():
    """Bad input is std::invalid_argument (code 1 here), as in the reference."""
    s = _shim()
    rp, ci, va = poisson_2d_csr(4)
    bad = ci.copy()
    bad[3] = 99  # column out of range
    b, x, st = np.ones(16), np.zeros(16), Stats()
    assert s.shim_solve(16, len(va), _p(rp), _p(bad), _p(va), _p(b), _p(x), 0, 1e-10, 10, C.byref(st)) == 1
    empty = (0, None, None, None, None, None, None)
    assert s.shim_solve_amg(*empty, 2, 0.7, _p(b), _p(x), 1e-10, 10, C.byref(st)) == 1
