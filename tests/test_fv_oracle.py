"""The oracle's restatement of gpu::DeviceFv, pinned: against the reference's own host path
(fixtures in tests/golden/ written by make_fv_golden.py, and oracle/_ref when it was built here),
and against what the mathematics fixes: least squares and skewness-corrected Gauss are exact for
linear fields on a mesh with planar faces (plain Gauss is not, on a skewed one), face corrections
conserve the source, limiters lie in [0, 1] and keep face extrapolations inside the neighbour
bounds, a vector gradient's columns are the scalar gradients of its components, and the scalar
corrections equal a face-major loop (the form of the reference's test, tests/test_gpu.cpp:418-446).
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from cfd_amd.fv import FvMesh, fv_box
from oracle import bind as ob

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = ["ref_fv_skewed_6x5x4", "ref_fv_jittered_5x4x6"]


def load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    g = lambda k: z[f"mesh_{k}"]
    mesh = FvMesh(int(g("cell_offsets").shape[0] - 1), int(g("neighbour").shape[0]), int(g("owner").shape[0]),
                  g("owner"), g("neighbour"), g("cell_offsets"), g("cell_faces"), g("weight"), g("sf"),
                  g("face_centre"), g("cell_centre"), g("volume"), g("correction"), g("lsq_d"), g("lsq_inverse"))
    inputs = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
    outputs = {k: z[k] for k in z.files if not k.startswith(("in_", "mesh_", "meta"))}
    return meta, mesh, inputs, outputs


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference_fixtures(name):
    meta, mesh, fi, out = load(name)
    o = ob.FvOracle(mesh)
    for s in (0, 1, 2):
        assert np.array_equal(o.gradient(fi["phi"], fi["bv"], s), out[f"grad_{s}"]), f"scheme {s}"
        assert np.array_equal(o.vector_gradient(fi["u"], fi["ub"], s), out[f"vgrad_{s}"]), f"vector scheme {s}"
    for lu in (0, 1):
        got = o.scalar_corrections(fi["phi"], fi["bv"], 1, fi["flux"], fi["gamma"], lu, fi["source"])
        assert np.array_equal(got, out[f"scorr_{lu}"])
    g = o.vector_gradient(fi["u"], fi["ub"], 1)
    for k, (lu, lim, ven, cen, tr) in enumerate(meta["momentum_cases"]):
        got = o.momentum_corrections(fi["u"], fi["ub"], g, fi["flux"], fi["gamma"], lu, lim, ven, fi["eps2"], cen,
                                     tr, fi["source3"])
        assert np.array_equal(got, out[f"mom_{k}"]), (lu, lim, ven, cen, tr)


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference checkout)")
def test_oracle_matches_reference_build_on_larger_mesh():
    mesh = fv_box(14, 11, 9, jitter=0.3, seed=5)
    o, r = ob.FvOracle(mesh), ob.FvReference(mesh)
    rng = np.random.default_rng(1)
    phi, bv = rng.standard_normal(mesh.n_cells), rng.standard_normal(mesh.n_boundary)
    u, ub = rng.standard_normal((mesh.n_cells, 3)), rng.standard_normal((mesh.n_boundary, 3))
    for s in (0, 1, 2):
        assert np.array_equal(o.gradient(phi, bv, s), r.gradient(phi, bv, s))
        assert np.array_equal(o.vector_gradient(u, ub, s), r.vector_gradient(u, ub, s))
    flux = rng.standard_normal(mesh.n_internal_faces)
    nu = rng.uniform(0.01, 0.2, mesh.n_internal_faces)
    src3 = rng.standard_normal((mesh.n_cells, 3))
    g = r.vector_gradient(u, ub, 2)
    eps2 = np.array([1e-2, 1e-3, 1e-4])
    for lim, ven in ((1, 0), (1, 1)):
        a = o.momentum_corrections(u, ub, g, flux, nu, 1, lim, ven, eps2, 0.75, 1, src3)
        b = r.momentum_corrections(flux, nu, 1, lim, ven, eps2, 0.75, 1, src3)
        assert np.array_equal(a, b)


def linear_case(mesh, a=(0.7, -1.3, 2.1), c0=0.4):
    a = np.asarray(a)
    phi = mesh.cell_centre @ a + c0
    bv = mesh.face_centre[mesh.n_internal_faces:] @ a + c0
    return a, phi, bv


@pytest.mark.parametrize("grading", [(1.0, 1.0, 1.0), (2.0, 0.6, 1.4)])
def test_linear_fields_exact_for_least_squares_and_corrected_gauss(grading):
    # the reference's check on its skewed box (tests/test_gpu.cpp:392-413)
    mesh = fv_box(6, 5, 4, xy_skew=0.25, grading=grading, shear=(0.2, 0.0, 0.1))
    o = ob.FvOracle(mesh)
    a, phi, bv = linear_case(mesh)
    err = {s: np.abs(o.gradient(phi, bv, s) - a).max() for s in (0, 1, 2)}
    assert err[1] < 1e-10 and err[2] < 1e-10
    assert err[0] > 1e-3  # plain Gauss (no skewness term) is inexact on the skewed mesh


def test_gauss_exact_for_linear_fields_on_uniform_mesh():
    mesh = fv_box(6, 5, 4)
    a, phi, bv = linear_case(mesh)
    assert np.abs(ob.FvOracle(mesh).gradient(phi, bv, 0) - a).max() < 1e-12


def test_vector_gradient_columns_are_scalar_gradients():
    mesh = fv_box(6, 5, 7, jitter=0.3, seed=9)
    o = ob.FvOracle(mesh)
    rng = np.random.default_rng(4)
    u, ub = rng.standard_normal((mesh.n_cells, 3)), rng.standard_normal((mesh.n_boundary, 3))
    for s in (0, 1, 2):
        v = o.vector_gradient(u, ub, s)
        for i in range(3):
            assert np.array_equal(v[:, 3 * i:3 * i + 3], o.gradient(u[:, i], ub[:, i], s))


def test_face_corrections_conserve_and_match_face_loop():
    mesh = fv_box(8, 6, 5, jitter=0.25, seed=2)
    o = ob.FvOracle(mesh)
    rng = np.random.default_rng(7)
    phi, bv = rng.standard_normal(mesh.n_cells), rng.standard_normal(mesh.n_boundary)
    nif = mesh.n_internal_faces
    flux, gamma = np.sin(0.7 * np.arange(nif)), 0.05 + 0.01 * (np.arange(nif) % 5)
    src = np.full(mesh.n_cells, 0.5)
    got = o.scalar_corrections(phi, bv, 1, flux, gamma, 1, src)
    assert abs((got - src).sum()) < 1e-12 * np.abs(got - src).sum()
    g = o.gradient(phi, bv, 1)
    ref = src.copy()
    own, nei = mesh.owner[:nif], mesh.neighbour
    cf, cc, w = mesh.face_centre, mesh.cell_centre, mesh.weight
    for f in range(nif):
        p, n = own[f], nei[f]
        up = p if flux[f] >= 0.0 else n
        e = -flux[f] * g[up].dot(cf[f] - cc[up]) + gamma[f] * mesh.correction[f].dot(w[f] * g[p] + (1 - w[f]) * g[n])
        ref[p] += e
        ref[n] -= e
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("venkat", [0, 1])
def test_limiter_bounds(venkat):
    mesh = fv_box(7, 6, 6, jitter=0.3, seed=6)
    o = ob.FvOracle(mesh)
    rng = np.random.default_rng(8)
    u, ub = rng.standard_normal((mesh.n_cells, 3)), rng.standard_normal((mesh.n_boundary, 3))
    grad = o.vector_gradient(u, ub, 1)
    lim = o.limiter(u, ub, grad, venkat, np.array([1e-3, 1e-3, 1e-3]))
    assert np.all(np.isfinite(lim))
    if venkat == 0:
        assert lim.min() >= 0.0 and lim.max() <= 1.0
        # Barth-Jespersen: every limited face extrapolation stays within the neighbour bounds
        nif = mesh.n_internal_faces
        for c in range(mesh.n_cells):
            faces = mesh.cell_faces[mesh.cell_offsets[c]:mesh.cell_offsets[c + 1]]
            vals = [u[c]]
            for f in faces:
                vals.append(u[mesh.neighbour[f] if mesh.owner[f] == c else mesh.owner[f]] if f < nif else ub[f - nif])
            lo, hi = np.min(vals, axis=0), np.max(vals, axis=0)
            g = grad[c].reshape(3, 3)
            for f in faces:
                x = u[c] + lim[c] * (g @ (mesh.face_centre[f] - mesh.cell_centre[c]))
                assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
    else:
        assert lim.max() <= 1.0 + 1e-12
