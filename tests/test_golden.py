"""The oracle and the HIP solvers against golden fixtures written by the reference's own host
solver code (tests/golden/make_golden.py; oracle/_ref built from the reference sources)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from cfd_amd import cases

from .conftest import poisson_2d_csr

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


@pytest.mark.parametrize("m", [40, 80])
def test_oracle_amg_cg_equals_reference_fixture(oracle, m):
    g = load(f"ref_amg_cg_poisson{m}")
    rp, ci, va = poisson_2d_csr(m)
    h = oracle.Hierarchy(oracle.Csr(m * m, m * m, rp, ci, va))
    assert h.level_sizes() == list(g["level_sizes"])
    x = np.zeros(m * m)
    st = h.amg_cg(np.ones(m * m), x, 1e-10, 500)
    assert st.iterations == int(g["iterations"]) and st.residual == float(g["residual"])
    assert np.array_equal(x, g["x"])


@pytest.mark.parametrize("method", [0, 1])
def test_oracle_krylov_equals_reference_fixture(oracle, method):
    g = load(f"ref_krylov_m{method}")
    rp, ci, va = poisson_2d_csr(40, 0.4 if method else 0.0)
    x = np.zeros(1600)
    st = oracle.solve(oracle.Csr(1600, 1600, rp, ci, va), np.sin(0.1 * np.arange(1600)) + 0.5, x,
                      method, 1e-10, 2000)
    assert st.iterations == int(g["iterations"]) and np.array_equal(x, g["x"])


def test_oracle_assembly_equals_reference_fixture(oracle):
    g = load("ref_assembly_7x5x3")
    mesh = cases.box_mesh(7, 5, 3)
    f = np.arange(mesh.n_faces)
    diag, div = np.ones(mesh.n_cells), np.zeros(mesh.n_cells)
    up, lo = oracle.assemble(mesh.owner, mesh.neighbour, np.sin(1.3 * f), 0.1 + 0.01 * (f % 7),
                             2.0 + 0.1 * (f % 3), 1.0, diag, div)
    for mine, key in ((up, "upper"), (lo, "lower"), (diag, "diag"), (div, "divergence")):
        assert np.array_equal(mine, g[key])


@pytest.mark.gpu
@pytest.mark.parametrize("m", [40, 80])
def test_gpu_amg_cg_against_reference_fixture(m):
    import cfd_amd
    from cfd_amd import AmgHierarchy, CsrView, SolveControl

    g = load(f"ref_amg_cg_poisson{m}")
    rp, ci, va = poisson_2d_csr(m)
    h = AmgHierarchy(CsrView(m * m, len(va), rp, ci, va))
    x = np.zeros(m * m)
    st = cfd_amd.solve_amg_cg(h.view(), np.ones(m * m), x, SolveControl(1e-10, 500))
    assert st.converged and abs(st.iterations - int(g["iterations"])) <= 1
    assert np.linalg.norm(x - g["x"]) <= 1e-10 * np.linalg.norm(g["x"])


@pytest.mark.gpu
def test_gpu_assembly_against_reference_fixture():
    import cfd_amd

    g = load("ref_assembly_7x5x3")
    mesh = cases.box_mesh(7, 5, 3)
    f = np.arange(mesh.n_faces)
    up, lo = np.zeros(mesh.n_faces), np.zeros(mesh.n_faces)
    diag, div = np.ones(mesh.n_cells), np.zeros(mesh.n_cells)
    cfd_amd.assemble_convection_diffusion(mesh.n_cells, mesh.n_faces, mesh.owner, mesh.neighbour,
                                          np.sin(1.3 * f), 0.1 + 0.01 * (f % 7),
                                          2.0 + 0.1 * (f % 3), 1.0, up, lo, diag, div)
    for mine, key in ((up, "upper"), (lo, "lower"), (diag, "diag"), (div, "divergence")):
        assert np.array_equal(mine, g[key])
