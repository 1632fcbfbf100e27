"""Writes the finite-volume fixtures of tests/golden/ from the reference's own gpu::DeviceFv
host path (src/gpu/src/fv_device.cpp, fv_openmp.cpp, fv_cell.inc), compiled by oracle/Makefile
into oracle/_ref/libcfdref.so. Run where the reference checkout exists, after `make -C oracle`:

    python tests/golden/make_fv_golden.py

Each fixture holds the mesh arrays (cfd_amd.fv.fv_box), the input fields and the reference's
outputs: scalar gradients (schemes 0, 1, 2), vector gradients (0, 1, 2), scalar corrections
(linear upwind off / on) and momentum corrections for every option combination.
"""
from __future__ import annotations

import itertools
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from cfd_amd.fv import fv_box  # noqa: E402
from oracle import bind as ob  # noqa: E402

MESH_FIELDS = ("owner", "neighbour", "cell_offsets", "cell_faces", "weight", "sf", "face_centre",
               "cell_centre", "volume", "correction", "lsq_d", "lsq_inverse")
MOMENTUM_CASES = list(itertools.product((0, 1), (0, 1), (0, 1), (0.0, 0.75), (0, 1)))


def fields(mesh, seed):
    rng = np.random.default_rng(seed)
    x = mesh.cell_centre
    xf = mesh.face_centre[mesh.n_internal_faces:]
    phi = np.sin(2.0 * x[:, 0]) + x[:, 1] ** 2 + x[:, 0] * x[:, 2]
    bv = np.sin(2.0 * xf[:, 0]) + xf[:, 1] ** 2 + xf[:, 0] * xf[:, 2]
    u = np.stack([np.sin(x[:, 1]) + x[:, 2], x[:, 0] * x[:, 1], np.cos(x[:, 2]) - x[:, 0]], axis=1)
    ub = np.stack([np.sin(xf[:, 1]) + xf[:, 2], xf[:, 0] * xf[:, 1], np.cos(xf[:, 2]) - xf[:, 0]], axis=1)
    f = np.arange(mesh.n_internal_faces)
    flux = np.sin(0.7 * f)
    gamma = 0.05 + 0.01 * (f % 5)
    source = rng.standard_normal(mesh.n_cells)
    source3 = rng.standard_normal((mesh.n_cells, 3))
    return dict(phi=phi, bv=bv, u=u, ub=ub, flux=flux, gamma=gamma, source=source, source3=source3,
                eps2=np.array([1e-3, 4e-3, 2e-4]))


def make(name, mesh, seed, meta):
    r = ob.FvReference(mesh)
    fl = fields(mesh, seed)
    out = {}
    for s in (0, 1, 2):
        out[f"grad_{s}"] = r.gradient(fl["phi"], fl["bv"], s)
    for s in (0, 1, 2):
        out[f"vgrad_{s}"] = r.vector_gradient(fl["u"], fl["ub"], s)
    for lu in (0, 1):
        out[f"scorr_{lu}"] = r.scalar_corrections(fl["phi"], fl["bv"], 1, fl["flux"], fl["gamma"], lu, fl["source"])
    r.vector_gradient(fl["u"], fl["ub"], 1)  # momentum corrections use the last vector gradient
    for k, (lu, lim, ven, cen, tr) in enumerate(MOMENTUM_CASES):
        out[f"mom_{k}"] = r.momentum_corrections(fl["flux"], fl["gamma"], lu, lim, ven, fl["eps2"], cen, tr,
                                                 fl["source3"])
    r.close()
    meta = dict(meta, momentum_cases=MOMENTUM_CASES, scalar_corrections_scheme=1, momentum_gradient_scheme=1,
                routine="gpu::DeviceFv host path")
    arrays = {f"mesh_{k}": getattr(mesh, k) for k in MESH_FIELDS}
    arrays.update({f"in_{k}": v for k, v in fl.items()})
    arrays.update(out)
    np.savez_compressed(HERE / f"{name}.npz", meta=np.array(json.dumps(meta)), **arrays)


def main() -> int:
    if not ob.ref_available():
        print("oracle/_ref/libcfdref.so missing: build it with `make -C oracle` where the reference exists")
        return 1
    make("ref_fv_skewed_6x5x4", fv_box(6, 5, 4, xy_skew=0.25, grading=(1.5, 0.7, 1.3)), 11,
         {"mesh": "fv_box(6, 5, 4, xy_skew=0.25, grading=(1.5, 0.7, 1.3))"})
    make("ref_fv_jittered_5x4x6", fv_box(5, 4, 6, jitter=0.25, seed=3), 12,
         {"mesh": "fv_box(5, 4, 6, jitter=0.25, seed=3)"})
    return 0


if __name__ == "__main__":
    sys.exit(main())
