"""Serving loop: continuous batching keeps every result equal to the request served alone.

Each request is a jittered box's pressure system with a run of time-step right-hand sides
(each the previous one plus 5% seeded noise) solved from the previous step's solution. The
same requests are served one at a time and all together through a server with fewer slots than
requests (so that solves join and leave the running set mid-way); every step's solution and
iteration count must match bit for bit, and also match the persistent-system path.
"""
from __future__ import annotations

import threading

import numpy as np
import pytest

import cfd_amd
from cfd_amd import AmgHierarchy, AmgSettings, Method, Precision, SolveControl, cases

pytestmark = pytest.mark.gpu


def _request(ns, seed, steps):
    mesh = cases.box_mesh(ns, ns + 1, ns + 3, jitter=0.2, seed=seed)
    s = cases.pressure_system(mesh, rhs_seed=seed)
    pat = cases.csr_pattern(s.n, s.owner, s.neighbour)
    a = cases.ldu_to_csr(pat, s.diag, s.upper, s.lower)
    view = AmgHierarchy(a, AmgSettings(strength_decay=0.5)).view()
    rng = np.random.default_rng(seed)
    rhs = [s.source.copy()]
    for _ in range(steps - 1):
        prev = rhs[-1]
        noise = rng.uniform(-1.0, 1.0, s.n)
        noise -= noise.mean()
        rhs.append(prev + 0.05 * np.linalg.norm(prev) / np.linalg.norm(noise) * noise)
    return dict(pattern=pat, diag=s.diag, upper=s.upper, lower=s.lower, amg=view, rhs=rhs)


REQS = [(10, 3, 3), (16, 4, 2), (7, 5, 4), (13, 6, 3), (19, 7, 2), (9, 8, 3)]


@pytest.fixture(scope="module")
def requests():
    return [_request(ns, seed, steps) for ns, seed, steps in REQS]


@pytest.mark.parametrize("prec", [Precision.MIXED_FP32, Precision.FP64])
def test_served_results_equal_alone_and_in_mix(requests, prec):
    alone = []
    with cfd_amd.Server(max_active_solves=2, precision=prec) as srv:
        for q in requests:
            alone.append(srv.wait(srv.submit(**q)))
    with cfd_amd.Server(max_active_solves=3, precision=prec) as srv:
        first = [srv.submit(**q) for q in requests[:4]]
        late = []

        def submit_rest():  # arrivals while the first ones are on the device
            for q in requests[4:]:
                late.append(srv.submit(**q))

        t = threading.Thread(target=submit_rest)
        t.start()
        mixed = [srv.wait(r) for r in first]
        t.join()
        mixed += [srv.wait(r) for r in late]
        c = srv.counters()
    assert c.completed == len(requests) and c.peak_active_solves <= 3
    for i, (a, m) in enumerate(zip(alone, mixed)):
        for k in range(a.n_steps):
            sa, sm = a.stats[k], m.stats[k]
            print(f"request {i} step {k}: alone it={sa.iterations} mix it={sm.iterations} "
                  f"rel res {sm.true_residual / sm.rhs_norm:.2e}")
            assert sa.converged and sm.converged
            assert sm.true_residual <= 1.05e-8 * sm.rhs_norm
            assert sa.iterations == sm.iterations and sa.residual == sm.residual
            assert np.array_equal(a.solutions[k], m.solutions[k])
            assert m.step_latency_ms[k] > 0.0


def test_served_steps_match_persistent_system(requests):
    q = requests[0]
    with cfd_amd.Server(max_active_solves=4, precision=Precision.MIXED_FP32) as srv:
        r = srv.wait(srv.submit(**q))
    dev = cfd_amd.DeviceSolver(q["pattern"])
    dev.assemble(q["diag"], q["upper"], q["lower"])
    dev.set_amg(q["amg"])
    dev.set_x(None)
    for k, b in enumerate(q["rhs"]):
        dev.set_rhs(b)
        tol = 1e-8 * np.linalg.norm(b)
        st = dev.solve_resident(Method.Cg, True, Precision.MIXED_FP32, SolveControl(tol, 500))
        assert st.iterations == r.stats[k].iterations
        assert np.array_equal(dev.get_x(), r.solutions[k])


def test_admission_holds_requests_beyond_the_memory_budget(requests):
    q = requests[1]
    one = cfd_amd.estimate_footprint(q["pattern"], q["amg"]) + 2 * len(q["rhs"]) * q["pattern"].n * 8
    with cfd_amd.Server(max_active_solves=8, memory_budget_bytes=int(1.5 * one)) as srv:
        rs = [srv.submit(**q) for _ in range(3)]
        out = [srv.wait(r) for r in rs]
        c = srv.counters()
    assert c.peak_active_solves == 1  # never two resident at once
    assert c.peak_resident_bytes <= int(1.5 * one)
    for r in out[1:]:
        for k in range(r.n_steps):
            assert np.array_equal(r.solutions[k], out[0].solutions[k])
    with cfd_amd.Server(max_active_solves=8, memory_budget_bytes=one // 4) as srv:
        r = srv.submit(**q)
        with pytest.raises(cfd_amd._lib.CfdgError) as e:
            srv.wait(r)
        assert e.value.status == 6  # CFDG_ERR_CAPACITY


def test_server_rejects_mismatched_hierarchy(requests):
    q = dict(requests[2])
    with cfd_amd.Server(max_active_solves=2, smooth_sweeps=3) as srv:
        with pytest.raises(cfd_amd._lib.InvalidArgument):
            srv.submit(**q)
