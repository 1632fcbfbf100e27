"""Writes the golden fixtures of tests/golden/ from the reference's own host solver code.

Run in a container that has the reference checkout, after `make -C oracle` built
oracle/_ref/libcfdref.so from the reference sources (src/gpu/src/krylov_openmp.cpp,
dispatch.cpp, device_solver.cpp):

    python tests/golden/make_golden.py

Each fixture holds only inputs that the tests regenerate (described in its `meta` field) and the
reference's outputs: iteration counts, residuals and solution vectors. Hierarchies are built by
the oracle's restatement of the reference setup (the reference setup needs Eigen, which is absent).
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import bind as ob  # noqa: E402
from tests.conftest import poisson_2d_csr  # noqa: E402


def save(name, meta, **arrays):
    np.savez_compressed(HERE / f"{name}.npz", meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", name)


def main() -> int:
    if not ob.ref_available():
        print("oracle/_ref/libcfdref.so missing: run `make -C oracle` with the reference present")
        return 1
    ob.set_threads(1)
    for m in (40, 80):
        rp, ci, va = poisson_2d_csr(m)
        a = ob.Csr(m * m, m * m, rp, ci, va)
        h = ob.Hierarchy(a)
        a_l, p_l, r_l = ob.hierarchy_lists(h)
        b = np.ones(m * m)
        x = np.zeros(m * m)
        st = ob.ref_solve_amg_cg_host(a_l, p_l, r_l, h.coarse_inverse(), b, x, 1e-10, 500)
        save(f"ref_amg_cg_poisson{m}",
             {"matrix": f"5-point Poisson {m}x{m} (tests/test_gpu.cpp:14-58, c=0)",
              "rhs": "ones", "x0": "zeros", "tol": 1e-10, "max_it": 500,
              "hierarchy": "oracle setup, theta 0.08, coarse 400", "routine": "solve_amg_cg_host"},
             iterations=np.int32(st.iterations), residual=np.float64(st.residual), x=x,
             level_sizes=np.array(h.level_sizes(), np.int32))
    for method, c in ((0, 0.0), (1, 0.4)):
        rp, ci, va = poisson_2d_csr(40, c)
        a = ob.Csr(1600, 1600, rp, ci, va)
        b = np.sin(0.1 * np.arange(1600)) + 0.5
        x = np.zeros(1600)
        st = ob.ref_solve_host(a, b, x, method, 1e-10, 2000)
        save(f"ref_krylov_m{method}",
             {"matrix": f"5-point 40x40, c={c}", "rhs": "sin(0.1 i) + 0.5", "x0": "zeros",
              "tol": 1e-10, "max_it": 2000, "method": ["Cg", "BiCgStab"][method],
              "routine": "solve_host"},
             iterations=np.int32(st.iterations), residual=np.float64(st.residual), x=x)
    # Assembly on the 7x5x3 box of tests/test_gpu.cpp:274-309.
    from cfd_amd import cases

    mesh = cases.box_mesh(7, 5, 3)
    nf, nc = mesh.n_faces, mesh.n_cells
    f = np.arange(nf)
    flux, gamma, dc = np.sin(1.3 * f), 0.1 + 0.01 * (f % 7), 2.0 + 0.1 * (f % 3)
    diag, div = np.ones(nc), np.zeros(nc)
    up, lo = ob.ref_assemble(mesh.owner, mesh.neighbour, flux, gamma, dc, 1.0, diag, div)
    save("ref_assembly_7x5x3",
         {"mesh": "box 7x5x3", "flux": "sin(1.3 f)", "gamma": "0.1 + 0.01 (f % 7)",
          "delta": "2 + 0.1 (f % 3)", "convection": 1.0, "diag0": "ones",
          "routine": "assemble_convection_diffusion"},
         upper=up, lower=lo, diag=diag, divergence=div)
    return 0


if __name__ == "__main__":
    sys.exit(main())
