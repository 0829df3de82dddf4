"""Serving: requests solved through the continuous-batching loop give the same results as solo
solves, admission keeps the reserved device memory within the budget, and a failing request
reports its error without disturbing the others."""
import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, rel

pytestmark = pytest.mark.gpu


def _mix():
    return [
        G.qp(n=600, m=300, seed=101),
        G.socp(n=500, m=250, ncone=20, seed=102, adens=0.02),
        G.qp(n=150, m=60, seed=103, pdens=0.03, adens=0.05),
        G.qp(n=1100, m=500, seed=104),
        G.socp(n=300, m=120, ncone=10, seed=105, adens=0.03),
        G.qp(n=400, m=200, seed=106),
    ]


def _close(a, b):
    assert a.status == b.status == mh.IPMStatus.OPTIMAL
    assert abs(a.niter - b.niter) <= 1
    assert rel(a.pobj, b.pobj) < 1e-8 and rel(a.dobj, b.dobj) < 1e-8
    assert np.linalg.norm(a.p - b.p) <= 1e-6 * (1.0 + np.linalg.norm(b.p))
    assert np.linalg.norm(a.y - b.y) <= 1e-6 * (1.0 + np.linalg.norm(b.y))


def test_each_result_is_the_same_alone_and_in_the_mix():
    data = _mix()
    probs = [hip_problem(d) for d in data]
    solo = [mh.solve(p) for p in probs]
    server = mh.IPMServer(nslots=4)
    try:
        tickets = [server.submit(p) for p in probs]
        served = [server.wait(t) for t in tickets]
        st = server.stats()
    finally:
        server.close()
    assert st["completed"] == len(probs) and st["running"] == 0 and st["queued"] == 0
    for (res, times), ref in zip(served, solo):
        _close(res, ref)
        assert times["submitted"] <= times["admitted"] <= times["finished"]
        assert 0 < times["device_bytes"] <= times["reserved_bytes"]  # the admission bound holds


def test_admission_keeps_reservations_within_the_budget():
    data = [G.qp(n=700, m=300, seed=200 + k) for k in range(6)]
    probs = [hip_problem(d) for d in data]
    budget_gb = 0.03  # about two of these instances at a time
    server = mh.IPMServer(nslots=6, mem_budget_gb=budget_gb)
    try:
        tickets = [server.submit(p) for p in probs]
        out = [server.wait(t) for t in tickets]
        st = server.stats()
    finally:
        server.close()
    assert all(r.status == mh.IPMStatus.OPTIMAL for r, _ in out)
    assert 0 < st["peak_reserved_bytes"] <= budget_gb * 2.0 ** 30
    # admission made some requests wait for memory rather than for a slot
    waits = sorted(t["admitted"] - t["submitted"] for _, t in out)
    assert waits[-1] > waits[0]


def test_a_failing_request_reports_its_error():
    # a power cone whose exponent lies outside (0, 1): the host analysis rejects it
    bad = G.qp(n=3, m=1, seed=300, pdens=0.5, adens=1.0)
    bad_prob = mh.IPMProblem(bad["Q"], bad["B"], bad["f"], bad["g"], 0.0, [mh.PowerCone(1.5)], [3])
    good = hip_problem(G.qp(n=200, m=80, seed=301))
    server = mh.IPMServer(nslots=2)
    try:
        tb = server.submit(bad_prob)
        tg = server.submit(good)
        with pytest.raises(RuntimeError, match="exponent"):
            server.wait(tb)
        res, _ = server.wait(tg)
    finally:
        server.close()
    assert res.status == mh.IPMStatus.OPTIMAL


def test_mixed_precision_requests_alongside_fp64_ones():
    """FP32-factor requests (precision 1) and FP64 ones in one stream: each result is the same
    as the same request solved alone, and the mixed results agree with FP64 to 1e-8."""
    data = _mix()[:4]
    probs = [hip_problem(d) for d in data]
    mixed = mh.IPMSettings(precision=1)
    solo64 = [mh.solve(p) for p in probs]
    solo32, fp32_calls = [], 0
    for p in probs:
        s = mh.init(p, mixed)
        solo32.append(s.solve())
        fp32_calls += s.stats()["factor_calls_fp32"]
        s.close()
    assert fp32_calls > 0  # the mix has fronts large enough for the FP32 path
    server = mh.IPMServer(nslots=4)
    try:
        tickets = [server.submit(p, s) for p in probs for s in (mixed, None)]
        out = [server.wait(t) for t in tickets]
        st = server.stats()
    finally:
        server.close()
    served = [r for r, _ in out]
    assert all(0 < t["device_bytes"] <= t["reserved_bytes"] for _, t in out)
    assert st["completed"] == 2 * len(probs)
    for k, (r64, r32) in enumerate(zip(solo64, solo32)):
        _close(served[2 * k], r32)
        _close(served[2 * k + 1], r64)
        assert rel(r32.pobj, r64.pobj) < 1e-8 and rel(r32.dobj, r64.dobj) < 1e-8


def test_a_request_larger_than_the_budget_runs_alone():
    """A request whose device bound exceeds the whole budget is admitted once the device is
    empty, and the requests behind it wait for it instead of starving it."""
    big = hip_problem(G.qp(n=900, m=400, seed=400))
    small = [hip_problem(G.qp(n=200, m=80, seed=401 + k)) for k in range(3)]
    server = mh.IPMServer(nslots=3, mem_budget_gb=1e-4)  # 100 KiB: every request is oversized
    try:
        tickets = [server.submit(p) for p in [small[0], big, small[1], small[2]]]
        out = [server.wait(t, timeout=120) for t in tickets]
        st = server.stats()
    finally:
        server.close()
    assert all(o is not None for o in out)
    assert all(r.status == mh.IPMStatus.OPTIMAL for r, _ in out)
    assert st["completed"] == 4
    spans = sorted((t["admitted"], t["finished"]) for _, t in out)
    for (a0, f0), (a1, _) in zip(spans, spans[1:]):
        assert f0 <= a1  # one at a time


def test_a_single_oversized_request_completes():
    server = mh.IPMServer(nslots=2, mem_budget_gb=1e-4)
    try:
        t = server.submit(hip_problem(G.qp(n=300, m=120, seed=410)))
        out = server.wait(t, timeout=60)
    finally:
        server.close()
    assert out is not None and out[0].status == mh.IPMStatus.OPTIMAL
