"""Device memory pool: buffers of dropped systems are reused at once by new ones.

Systems of different sizes are created and closed between two builds of the same system; the
second build reuses ranges the others held (now holding their stale contents), and every solve
must still give bit-identical results. Releasing the cache returns only empty slabs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases
from cfd_amd import _lib as L

pytestmark = pytest.mark.gpu


def _build(ns, seed):
    mesh = cases.box_mesh(ns, ns + 1, ns + 3, jitter=0.2, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed + 1)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    return s, pat, AmgHierarchy(a, AmgSettings(strength_decay=0.5)).view()


def _solve(case, prec):
    s, pat, view = case
    with cfd_amd.DeviceSolver(pat) as dev:
        dev.assemble(s.diag, s.upper, s.lower)
        dev.set_amg(view)
        dev.set_rhs(s.source)
        out = []
        for _ in range(2):  # eager, then captured graphs
            dev.set_x(None)
            st = dev.solve_resident(Method.Cg, True, prec, SolveControl(1e-8 * np.linalg.norm(s.source), 300))
            out.append((st.iterations, st.residual, dev.get_x()))
    return out


def _release():
    freed, reserved = C.c_size_t(), C.c_size_t()
    L.call("cfdg_release_cached_memory", C.byref(freed), C.byref(reserved))
    return freed.value, reserved.value


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32])
def test_reused_memory_gives_identical_results(prec):
    target = _build(30, 5)
    others = [_build(ns, 11 + ns) for ns in (44, 18, 52, 25)]
    _release()
    first = _solve(target, prec)
    for o in others:  # grow and churn the pool with other sizes
        _solve(o, prec)
    again = _solve(target, prec)
    for (it1, r1, x1), (it2, r2, x2) in zip(first, again):
        assert it1 == it2 and r1 == r2
        assert np.array_equal(x1, x2)
    for (it1, _, x1), (it2, _, x2) in zip(first, first[1:]):
        assert it1 == it2 and np.array_equal(x1, x2)


def test_release_returns_empty_slabs():
    _release()
    _, before = _release()
    case = _build(40, 3)
    with cfd_amd.DeviceSolver(case[1]) as dev:
        dev.assemble(case[0].diag, case[0].upper, case[0].lower)
        dev.set_amg(case[2])
        _, held = _release()  # the live system's slabs stay
        assert held > 0
    freed, after = _release()
    assert after <= before
    assert freed + after >= held
