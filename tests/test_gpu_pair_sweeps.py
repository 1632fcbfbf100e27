"""Level-0 FP32 post-sweep with two slices per wavefront against the one-slice kernel
(CFDG_NO_PAIR_SWEEPS): mixed FP32 and BF16 solves, alone and in a batch, with the stencil and
the per-slot column layouts (equal and unequal slice widths), are bit-identical. The members'
sizes leave a block's last slice pair incomplete or empty.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

pytestmark = pytest.mark.gpu


def _system(shape, seed):
    mesh = cases.box_mesh(*shape, jitter=0.2, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    view = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5)).view()
    return s, pat, view


def _build(s, pat, view):
    dev = cfd_amd.DeviceSolver(pat)
    dev.assemble(s.diag, s.upper, s.lower)
    dev.set_amg(view)
    dev.set_rhs(s.source)
    return dev


SHAPES = [(20, 20, 20), (9, 13, 40), (17, 11, 23)]  # 125, 74 and 68 slices of 64 rows


@pytest.mark.parametrize("stencil", [True, False])
@pytest.mark.parametrize("prec", [Precision.MIXED_FP32, Precision.MIXED_BF16])
def test_pair_sweeps_bitwise_alone_and_batched(monkeypatch, stencil, prec):
    if stencil:
        monkeypatch.delenv("CFDG_NO_STENCIL", raising=False)
    else:
        monkeypatch.setenv("CFDG_NO_STENCIL", "1")
    systems = [_system(shape, 60 + i) for i, shape in enumerate(SHAPES)]
    results = {}
    for pair in (True, False):
        if pair:
            monkeypatch.delenv("CFDG_NO_PAIR_SWEEPS", raising=False)
        else:
            monkeypatch.setenv("CFDG_NO_PAIR_SWEEPS", "1")
        devs = [_build(*t) for t in systems]
        try:
            ctls = [SolveControl(1e-9 * np.linalg.norm(t[0].source), 300) for t in systems]
            alone = []
            for d, c in zip(devs, ctls):
                d.set_x(None)
                st = d.solve_resident(Method.Cg, True, prec, c)
                assert st.converged
                alone.append((st.iterations, st.residual, d.get_x()))
            with cfd_amd.Batch(devs, Method.Cg, True, prec) as batch:
                for d in devs:
                    d.set_x(None)
                stats = batch.solve(ctls)
                batched = [(st.iterations, st.residual, d.get_x()) for st, d in zip(stats, devs)]
            results[pair] = (alone, batched)
        finally:
            for d in devs:
                d.close()
    for which in (0, 1):
        for a, b in zip(results[True][which], results[False][which]):
            assert a[0] == b[0] and a[1] == b[1]
            assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("lanes,csr", [(2, 0), (4, 1), (1, 1)])
def test_pair_prolongation_with_other_transfer_layouts(monkeypatch, lanes, csr):
    """P_0 laid out with several lanes per row or as CSR (forced by a layout rule) takes the
    paired prolongation's general path; solves stay bit-identical to the one-slice kernels."""
    import json
    from pathlib import Path

    from cfd_amd import autotune

    rule = {"min_rows": 1000, "max_rows": 20000, "min_mean": 3.0, "max_mean": 6.0, "lanes": lanes, "csr": csr}
    autotune.apply({"layout_rules": [rule]})
    try:
        systems = [_system(shape, 70 + i) for i, shape in enumerate(SHAPES[1:])]
        results = {}
        for pair in (True, False):
            if pair:
                monkeypatch.delenv("CFDG_NO_PAIR_SWEEPS", raising=False)
            else:
                monkeypatch.setenv("CFDG_NO_PAIR_SWEEPS", "1")
            devs = [_build(*t) for t in systems]
            try:
                ctls = [SolveControl(1e-9 * np.linalg.norm(t[0].source), 300) for t in systems]
                out = []
                for d, c in zip(devs, ctls):
                    d.set_x(None)
                    st = d.solve_resident(Method.Cg, True, Precision.MIXED_FP32, c)
                    assert st.converged
                    out.append((st.iterations, d.get_x()))
                with cfd_amd.Batch(devs, Method.Cg, True, Precision.MIXED_FP32) as batch:
                    for d in devs:
                        d.set_x(None)
                    stats = batch.solve(ctls)
                    out += [(st.iterations, d.get_x()) for st, d in zip(stats, devs)]
                results[pair] = out
            finally:
                for d in devs:
                    d.close()
        for a, b in zip(results[True], results[False]):
            assert a[0] == b[0]
            assert np.array_equal(a[1], b[1])
    finally:
        path = Path(cfd_amd.__file__).resolve().parent / "tuning.json"
        autotune.apply(json.loads(path.read_text()) if path.exists() else {"layout_rules": []})
