"""Layout and launch switches without a test of their own: 32-bit columns everywhere
(CFDG_NO_SHORT_COLS) and kernel-by-kernel launches instead of HIP graph replays (CFDG_NO_GRAPHS)
give the same iterations, recurrence residuals and solutions, bit for bit, as the defaults, for
FP64 and mixed FP32 solves alone and in a batch.
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

pytestmark = pytest.mark.gpu

SHAPES = [(20, 20, 20), (9, 13, 40)]


def _system(shape, seed):
    mesh = cases.box_mesh(*shape, jitter=0.2, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    view = AmgHierarchy(cases.ldu_to_csr(pat, s.diag, s.upper, s.lower), AmgSettings(strength_decay=0.5)).view()
    return s, pat, view


def _solve_all(systems, prec):
    devs = []
    try:
        for s, pat, view in systems:
            dev = cfd_amd.DeviceSolver(pat)
            devs.append(dev)
            dev.assemble(s.diag, s.upper, s.lower)
            dev.set_amg(view)
            dev.set_rhs(s.source)
        ctls = [SolveControl(1e-9 * np.linalg.norm(t[0].source), 300) for t in systems]
        out = []
        # Each system and the batch are solved twice: a runner launches kernel by kernel on its
        # first solve and replays captured graphs from its second on (unless CFDG_NO_GRAPHS).
        for d, c in zip(devs, ctls):
            for _ in range(2):
                d.set_x(None)
                st = d.solve_resident(Method.Cg, True, prec, c)
                assert st.converged
                out.append((st.iterations, st.residual, d.get_x()))
        reps = []
        with cfd_amd.Batch(devs, Method.Cg, True, prec) as batch:
            for _ in range(2):
                for d in devs:
                    d.set_x(None)
                stats = batch.solve(ctls)
                reps.append([(st.iterations, st.residual, d.get_x()) for st, d in zip(stats, devs)])
        for first, second in zip(*reps):  # per member: first batch solve, then the replayed one
            out += [first, second]
        return out
    finally:
        for d in devs:
            d.close()


@pytest.mark.parametrize("switch,stencil", [("CFDG_NO_SHORT_COLS", True), ("CFDG_NO_SHORT_COLS", False),
                                            ("CFDG_NO_GRAPHS", True)])
@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32])
def test_switch_gives_identical_solves(monkeypatch, switch, stencil, prec):
    """stencil=False lays level 0 out with per-slot columns too (CFDG_NO_STENCIL in both runs), so
    the 16-bit column offsets are compared on the finest operator as well as the coarse ones."""
    if stencil:
        monkeypatch.delenv("CFDG_NO_STENCIL", raising=False)
    else:
        monkeypatch.setenv("CFDG_NO_STENCIL", "1")
    systems = [_system(shape, 80 + i) for i, shape in enumerate(SHAPES)]
    monkeypatch.delenv(switch, raising=False)
    default = _solve_all(systems, prec)
    monkeypatch.setenv(switch, "1")
    switched = _solve_all(systems, prec)
    assert len(default) == len(switched) == 4 * len(SHAPES)
    for a, b in zip(default, switched):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2])
    for k in range(0, len(default), 2):  # eager first solve and replayed second solve agree too
        assert default[k][0] == default[k + 1][0] and np.array_equal(default[k][2], default[k + 1][2])
