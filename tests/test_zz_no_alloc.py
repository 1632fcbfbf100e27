"""The allocation-free CSR solve (cfdg_solve_csr_noalloc: device arrays in, work space in the
library's fixed scratch area) against the allocating one-shot solve (cfdg_solve_csr, host arrays)
on the same inputs: the same iterations, residuals and solution, bit for bit, on calls in a row
whose later systems are smaller (the scratch area keeps the previous call's contents), from a
zero and from a non-zero initial guess; oversized and malformed inputs are errors.

The caller's device buffers are allocated here through the HIP runtime the library itself links
(libamdhip64), as a caller that owns its device memory would."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cfd_amd
from cfd_amd import CsrView, Method, SolveControl, cases
from cfd_amd import _lib as L
from cfd_amd.krylov import NOALLOC_MAX_ROWS, solve_device_noalloc

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libcfdgpu links."""

    def __init__(self):
        L.lib()  # loads libcfdgpu and with it libamdhip64.so.7
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def upload(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p.value)
        assert self.rt.hipMemcpy(p.value, a.ctypes.data, a.nbytes, _H2D) == 0
        return p.value

    def download(self, ptr: int, n: int, dtype) -> np.ndarray:
        out = np.empty(n, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, _D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.fixture
def hip():
    h = _Hip()
    yield h
    h.free()


def _system(shape, kind, seed):
    mesh = cases.box_mesh(*shape, jitter=0.2, seed=seed)
    s = (cases.pressure_system(mesh, rhs_seed=seed + 1) if kind == "pressure"
         else cases.momentum_system(mesh, 1.0, seed + 1))
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    return cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), s.source


def _on_device(hip, a, b, x0):
    return dict(rp=hip.upload(np.asarray(a.row_ptr, np.int32)), col=hip.upload(np.asarray(a.col, np.int32)),
                val=hip.upload(np.asarray(a.val, np.float64)), b=hip.upload(np.asarray(b, np.float64)),
                x=hip.upload(np.asarray(x0, np.float64)))


def _solve_both(hip, a, b, x0, method, control):
    x_ref = x0.copy()
    st_ref = cfd_amd.solve(CsrView(a.n, a.nnz, a.row_ptr, a.col, a.val), b, x_ref, method, control)
    t = _on_device(hip, a, b, x0)
    st = solve_device_noalloc(a.n, a.nnz, t["rp"], t["col"], t["val"], t["b"], t["x"], method, control)
    x = hip.download(t["x"], a.n, np.float64)
    return st, x, st_ref, x_ref


def test_noalloc_matches_the_allocating_solve_on_calls_in_a_row(hip):
    runs = [((48, 40, 36), "pressure", Method.Cg, 0.0),       # 69,120 rows
            ((30, 26, 22), "momentum", Method.BiCgStab, 0.0),  # 17,160 rows: smaller than the first
            ((17, 9, 5), "pressure", Method.Cg, 0.5),          # 765 rows, non-zero initial guess
            ((12, 11, 10), "momentum", Method.BiCgStab, 0.5)]
    for i, (shape, kind, method, guess) in enumerate(runs):
        a, b = _system(shape, kind, 40 + i)
        x0 = guess * np.cos(0.01 * np.arange(a.n))
        control = SolveControl(1e-8 * float(np.linalg.norm(b)), 2000)
        st, x, st_ref, x_ref = _solve_both(hip, a, b, x0, method, control)
        print(f"call {i}: {kind} n {a.n} {method.name}: it {st.iterations} (allocating {st_ref.iterations}) "
              f"res {st.residual:.6e} true {st.true_residual:.6e} status {st.exit_reason}", flush=True)
        assert st.converged and st_ref.converged
        assert st.iterations == st_ref.iterations and st.exit_reason == st_ref.exit_reason
        assert st.residual == st_ref.residual and st.true_residual == st_ref.true_residual
        assert st.rhs_norm == st_ref.rhs_norm
        assert np.array_equal(x, x_ref)


def test_noalloc_iteration_limit_and_converged_start(hip):
    a, b = _system((20, 18, 16), "pressure", 7)
    for control in (SolveControl(1e-12 * float(np.linalg.norm(b)), 7), SolveControl(1e300, 50),
                    SolveControl(1e-8, 0)):
        for method in (Method.Cg, Method.BiCgStab):
            st, x, st_ref, x_ref = _solve_both(hip, a, b, np.zeros(a.n), method, control)
            assert (st.iterations, st.converged, st.exit_reason) == (st_ref.iterations, st_ref.converged,
                                                                     st_ref.exit_reason)
            assert st.residual == st_ref.residual and np.array_equal(x, x_ref)


def test_noalloc_rejects_oversized_and_malformed_input(hip):
    p = hip.upload(np.zeros(16))
    with pytest.raises(L.CfdgError) as err:  # rejected on the host, before any device access
        solve_device_noalloc(NOALLOC_MAX_ROWS + 1, 1, p, p, p, p, p, Method.Cg, SolveControl(1e-8, 10))
    assert err.value.status == 6 and "maximum" in str(err.value)  # CFDG_ERR_CAPACITY
    a, b = _system((6, 5, 4), "pressure", 3)
    bad = a.col.copy()
    bad[17] = a.n  # a column past the last row: the device check rejects it before any solve kernel
    t = _on_device(hip, CsrView(a.n, a.nnz, a.row_ptr, bad, a.val), b, np.zeros(a.n))
    with pytest.raises(L.InvalidArgument):
        solve_device_noalloc(a.n, a.nnz, t["rp"], t["col"], t["val"], t["b"], t["x"], Method.Cg,
                             SolveControl(1e-8, 10))
    # the next call on the same scratch area is unaffected
    a2, b2 = _system((10, 9, 8), "pressure", 5)
    control = SolveControl(1e-8 * float(np.linalg.norm(b2)), 500)
    st, x, st_ref, x_ref = _solve_both(hip, a2, b2, np.zeros(a2.n), Method.Cg, control)
    assert st.iterations == st_ref.iterations and np.array_equal(x, x_ref)
