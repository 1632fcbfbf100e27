"""Shared test configuration: the `gpu` marker, import paths and system builders."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP path)")


def has_gpu() -> bool:
    import cfd_amd

    return cfd_amd.device_available()


@pytest.fixture(scope="session")
def oracle():
    from oracle import bind

    bind.set_threads(1)
    return bind


def poisson_2d_csr(m: int, c: float = 0.0):
    """Five-point operator of tests/test_gpu.cpp:14-58 (-laplace + c d/dx)."""
    rp, ci, va = [0], [], []
    for j in range(m):
        for i in range(m):
            def add(ii, jj, v):
                if 0 <= ii < m and 0 <= jj < m:
                    ci.append(ii + jj * m)
                    va.append(v)
            add(i, j, 4.0)
            add(i - 1, j, -1.0 - c)
            add(i + 1, j, -1.0 + c)
            add(i, j - 1, -1.0)
            add(i, j + 1, -1.0)
            rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64))


def build_poisson_ldu(mesh, dirichlet_xmin: bool):
    """build_poisson of tests/test_amg.cpp:46-72 on a uniform box (xmin face coefficient
    area / |C_f - C_P| = 2 * area / dx for a uniform box)."""
    n = mesh.n_cells
    c = mesh.coeff
    diag = np.bincount(mesh.owner, c, n) + np.bincount(mesh.neighbour, c, n)
    if dirichlet_xmin:
        dx = (mesh.hi[0] - mesh.lo[0]) / mesh.nx
        area = (mesh.hi[1] - mesh.lo[1]) / mesh.ny * (mesh.hi[2] - mesh.lo[2]) / mesh.nz
        for k in range(mesh.nz):
            for j in range(mesh.ny):
                diag[mesh.nx * (j + mesh.ny * k)] += area / (0.5 * dx)
    return diag, -c.copy(), -c.copy()
