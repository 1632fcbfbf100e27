"""Batched resident solves: every member of a batch gets exactly the result it gets alone.

Each system keeps its own block partials, reductions and stopping test inside a batch, so a
member's iterate, iteration count and residuals are bit-identical to a solve of that system by
itself, whatever else shares the launches (sizes, iteration counts, convergence order).
"""
from __future__ import annotations

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

pytestmark = pytest.mark.gpu


def _member(ns, seed, momentum=False):
    mesh = cases.box_mesh(ns, ns + 2, ns - 1, jitter=0.15, seed=seed)
    s = cases.momentum_system(mesh, rhs_seed=seed) if momentum else \
        cases.pressure_system(mesh, rhs_seed=seed)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    dev = cfd_amd.DeviceSolver(pat)
    dev.assemble(s.diag, s.upper, s.lower)
    if not momentum:
        a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
        dev.set_amg(AmgHierarchy(a, AmgSettings(strength_decay=0.5)).view())
    dev.set_rhs(s.source)
    return dev, s


def _alone(dev, method, amg, prec, ctl):
    dev.set_x(None)
    st = dev.solve_resident(method, amg, prec, ctl)
    return st, dev.get_x()


@pytest.mark.parametrize("method,amg,prec", [
    (Method.Cg, True, Precision.FP64),
    (Method.Cg, True, Precision.MIXED_FP32),
    (Method.Cg, True, Precision.MIXED_BF16),
    (Method.BiCgStab, False, Precision.FP64),
])
def test_batch_members_match_alone(method, amg, prec):
    momentum = method == Method.BiCgStab
    members = [_member(ns, 40 + i, momentum) for i, ns in enumerate((9, 17, 12, 21))]
    ctls = [SolveControl(tol * np.linalg.norm(s.source), 400)
            for (_, s), tol in zip(members, (1e-8, 1e-6, 1e-10, 1e-8))]
    alone = [_alone(d, method, amg, prec, c) for (d, _), c in zip(members, ctls)]
    with cfd_amd.Batch([d for d, _ in members], method, amg, prec) as batch:
        for rep in range(2):  # the second solve replays the cached graphs
            for d, _ in members:
                d.set_x(None)
            stats = batch.solve(ctls)
            for i, ((d, _), st, (st1, x1)) in enumerate(zip(members, stats, alone)):
                x = d.get_x()
                print(f"rep {rep} member {i}: n={d.n} it={st.iterations} alone={st1.iterations}")
                assert st.converged and st1.converged
                assert st.iterations == st1.iterations
                assert st.residual == st1.residual
                assert st.true_residual == st1.true_residual
                assert np.array_equal(x, x1)


def test_batch_with_max_iterations_zero_and_converged_start():
    (d0, s0), (d1, s1) = _member(10, 3), _member(14, 4)
    with cfd_amd.Batch([d0, d1], Method.Cg, True, Precision.FP64) as batch:
        d0.set_x(None)
        d1.set_x(np.zeros(d1.n))
        st = batch.solve([SolveControl(1e-30, 0), SolveControl(1e-8 * np.linalg.norm(s1.source), 300)])
        assert st[0].iterations == 0 and not st[0].converged
        assert np.array_equal(d0.get_x(), np.zeros(d0.n))
        assert st[1].converged
        x1 = d1.get_x()
        d1.set_x(x1)
        st = batch.solve([SolveControl(1e-30, 0), SolveControl(1.01 * st[1].true_residual, 300)])
        assert st[1].iterations == 0 and st[1].converged
        assert np.array_equal(d1.get_x(), x1)


def test_batch_rejects_mismatched_controls():
    from cfd_amd._lib import InvalidArgument

    (d0, _), (d1, _) = _member(8, 1), _member(9, 2)
    with cfd_amd.Batch([d0, d1]) as batch:
        with pytest.raises(InvalidArgument):
            batch.solve([SolveControl()])


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.MIXED_FP32])
def test_batch_zero_x_matches_per_member_set_x(prec):
    """Batch.zero_x (one launch for every member's initial guess) gives the solves that
    DeviceSolver.set_x(None) member by member gives, bit for bit, on repeated solves; members
    whose x was last set from the host are zeroed too."""
    members = [_member(ns, 40 + ns) for ns in (9, 14, 6)]
    devs = [d for d, _ in members]
    ctls = [SolveControl(1e-9 * np.linalg.norm(s.source), 300) for _, s in members]
    try:
        with cfd_amd.Batch(devs, Method.Cg, True, prec) as batch:
            for d in devs:
                d.set_x(None)
            ref = [(st.iterations, st.residual, d.get_x()) for st, d in zip(batch.solve(ctls), devs)]
            for rep in range(2):
                if rep == 1:
                    devs[1].set_x(np.ones(devs[1].n))  # a host-set guess, then zeroed by the batch
                batch.zero_x()
                got = [(st.iterations, st.residual, d.get_x()) for st, d in zip(batch.solve(ctls), devs)]
                for a, b in zip(ref, got):
                    assert a[0] == b[0] and a[1] == b[1]
                    assert np.array_equal(a[2], b[2])
    finally:
        for d in devs:
            d.close()
