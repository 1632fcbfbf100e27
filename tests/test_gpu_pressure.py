"""Pressure-equation assembly on the GPU (cfdg_fv_pressure_equation) against the oracle's face
loops, bit for bit, and the device-resident hand-over to a solver system
(cfdg_system_assemble_pressure) against the same equation passed as host arrays.
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

import cfd_amd
from cfd_amd import (AmgHierarchy, AmgSettings, DeviceFv, Method, Precision, SolveControl, cases, fv_box,
                     synthetic_pressure_inputs)

from cfd_amd import _lib as L
from oracle.bind import FvOracle

pytestmark = pytest.mark.gpu

FIELDS = ("diag", "upper", "lower", "source", "phi_hbya", "rau_f", "hbya", "rau")


def _check(mesh, inp):
    ref = FvOracle(mesh).pressure_equation(inp)
    with DeviceFv(mesh) as dev:
        got = dev.pressure_equation(inp)
    for k in FIELDS:
        assert np.array_equal(getattr(got, k), ref[k]), k


@pytest.mark.parametrize("shape,kw", [((6, 5, 4), dict(jitter=0.25, seed=1)),
                                      ((11, 9, 7), dict(jitter=0.2, seed=2, xy_skew=0.25)),
                                      ((24, 20, 16), dict(jitter=0.2, seed=3, grading=(1.0, 1.5, 1.0)))])
@pytest.mark.parametrize("random_kinds,relax,pin", [(False, False, True), (True, True, False), (True, False, True)])
def test_pressure_equation_matches_oracle_bitwise(shape, kw, random_kinds, relax, pin):
    mesh = fv_box(*shape, **kw)
    inp = synthetic_pressure_inputs(mesh, 11, random_kinds=random_kinds, relax=relax)
    inp.pin_reference = pin
    _check(mesh, inp)


def test_pressure_equation_large_mesh_matches_oracle():
    mesh = fv_box(48, 40, 36, jitter=0.2, seed=5)
    _check(mesh, synthetic_pressure_inputs(mesh, 12, random_kinds=True, relax=True))


def test_pressure_equation_repeated_calls_and_fields():
    """Two calls on one object (different fields, relaxation switched off again) each equal the
    oracle: no state of the first call leaks into the second."""
    mesh = fv_box(9, 8, 7, jitter=0.2, seed=6)
    a = synthetic_pressure_inputs(mesh, 1, random_kinds=True, relax=True)
    b = synthetic_pressure_inputs(mesh, 2)
    orc = FvOracle(mesh)
    with DeviceFv(mesh) as dev:
        for inp in (a, b, a):
            got = dev.pressure_equation(inp)
            ref = orc.pressure_equation(inp)
            for k in FIELDS:
                assert np.array_equal(getattr(got, k), ref[k]), k


def _pattern(mesh):
    nif = mesh.n_internal_faces
    return cases.csr_pattern(mesh.n_cells, mesh.owner[:nif].astype(np.int32), mesh.neighbour.astype(np.int32))


def test_resident_pressure_solve_equals_host_arrays():
    mesh = fv_box(20, 18, 16, jitter=0.2, seed=7)
    inp = synthetic_pressure_inputs(mesh, 13)
    pat = _pattern(mesh)
    with DeviceFv(mesh) as fv:
        eq = fv.pressure_equation(inp)
        view = AmgHierarchy(cases.ldu_to_csr(pat, eq.diag, eq.upper, eq.lower), AmgSettings(strength_decay=0.5)).view()
        tol = 1e-9 * np.linalg.norm(eq.source)
        out = {}
        for resident in (True, False):
            dev = cfd_amd.DeviceSolver(pat)
            try:
                if resident:
                    dev.assemble_pressure(fv)
                else:
                    dev.assemble(eq.diag, eq.upper, eq.lower)
                    dev.set_rhs(eq.source)
                dev.set_amg(view)
                for prec in (Precision.FP64, Precision.MIXED_FP32):
                    dev.set_x(None)
                    st = dev.solve_resident(Method.Cg, True, prec, SolveControl(tol, 300))
                    assert st.converged
                    out[(resident, prec)] = (st.iterations, dev.get_x())
            finally:
                dev.close()
    for prec in (Precision.FP64, Precision.MIXED_FP32):
        assert out[(True, prec)][0] == out[(False, prec)][0]
        assert np.array_equal(out[(True, prec)][1], out[(False, prec)][1])


def test_pressure_errors():
    mesh = fv_box(5, 4, 3, jitter=0.1, seed=8)
    inp = synthetic_pressure_inputs(mesh, 14)
    other = fv_box(4, 4, 3, jitter=0.1, seed=8)
    with DeviceFv(mesh) as fv:
        dev = cfd_amd.DeviceSolver(_pattern(mesh))
        try:
            with pytest.raises(L.InvalidArgument):
                dev.assemble_pressure(fv)  # nothing assembled yet
            bad = copy.deepcopy(inp)
            bad.boundary_kind = bad.boundary_kind.copy()
            bad.boundary_kind[0] = 7
            with pytest.raises(L.InvalidArgument):
                fv.pressure_equation(bad)
            fv.pressure_equation(inp)
            dev.assemble_pressure(fv)
        finally:
            dev.close()
        dev2 = cfd_amd.DeviceSolver(_pattern(other))
        try:
            with pytest.raises(L.InvalidArgument):
                dev2.assemble_pressure(fv)  # another mesh
        finally:
            dev2.close()


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 1, 1), (1, 1, 3), (65, 1, 1)])
def test_pressure_equation_tiny_meshes(shape):
    """No internal faces, one or two of them, and a partial second slice of 64 cells."""
    mesh = fv_box(*shape, jitter=0.0, seed=1)
    _check(mesh, synthetic_pressure_inputs(mesh, 3, random_kinds=True, relax=True))
