"""The served stream as a test: 128 requests of the stream workload (60 % QP, 30 % SOCP, 10 %
chordal max-cut SDP, n log-uniform in [50, 4000]) arriving as a Poisson process and solved
through the continuous-batching loop. Each served result is compared with the same request
solved alone, and with the oracle where the oracle is quick (n ≤ 300).

Served and solo solves run the same kernels with the same summation orders (the persistent
kernels' sums do not depend on how many workgroups they get, and the extend-add adds a front's
children in a fixed order), so a served result equals the solo one bit for bit. Against the
oracle, objectives are compared to 1e-8, or to the two solves' own duality gaps where one ended
NEAR_OPTIMAL (small SOCPs whose last iterate sits just above the gap test)."""
import time

import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem, rel

pytestmark = pytest.mark.gpu

GOOD = (mh.IPMStatus.OPTIMAL, mh.IPMStatus.NEAR_OPTIMAL)


def _agree(a, b):
    assert a.status in GOOD and b.status in GOOD, (a.status, b.status)
    both = a.status == b.status == mh.IPMStatus.OPTIMAL
    scale = max(1.0, abs(b.pobj))
    gaps = (abs(a.pobj - a.dobj) + abs(b.pobj - b.dobj)) / scale
    tol = 1e-8 if both else max(1e-8, gaps + 1e-8)
    assert tol <= 1e-6
    return rel(a.pobj, b.pobj) < tol and rel(a.dobj, b.dobj) < tol


def test_served_stream_matches_solo_and_oracle():
    from oracle import ipm as O
    rng = np.random.default_rng(11)
    K = 128
    data = [G.served_request(rng, k, 4000, seed0=7000) for k in range(K)]
    probs = [hip_problem(d) for d in data]
    gaps = rng.exponential(1.0 / 400.0, K)
    server = mh.IPMServer(nslots=8)
    t0 = mh.clock()
    tickets = []
    try:
        for k in range(K):
            target = t0 + float(np.sum(gaps[:k + 1]))
            while mh.clock() < target:
                time.sleep(min(0.002, max(0.0, target - mh.clock())))
            tickets.append(server.submit(probs[k]))
        print(f"stream: {K} requests submitted", flush=True)
        served = [server.wait(t, timeout=600) for t in tickets]
        st = server.stats()
    finally:
        server.close()
    assert all(x is not None for x in served)
    assert st["completed"] == K and st["running"] == 0
    print("stream: served; solving each alone", flush=True)
    kinds = {"qp": 0, "socp": 0, "maxcut": 0}
    checked = 0
    for k, ((res, times), d) in enumerate(zip(served, data)):
        kinds[d["name"].split("_")[0]] += 1
        assert times["submitted"] <= times["admitted"] <= times["finished"]
        assert 0 < times["device_bytes"] <= times["reserved_bytes"]
        solo = mh.solve(probs[k])
        if res.status not in GOOD or solo.status not in GOOD:
            print(f"request {k} {d['name']} B {d['B'].shape}: served {res.status.name} in {res.niter} iterations "
                  f"(mu {res.mu:.3g}, pres {res.pres:.3g}, dres {res.dres:.3g}), alone {solo.status.name} in "
                  f"{solo.niter}", flush=True)
        assert res.status == solo.status and res.niter == solo.niter, (k, d["name"], res.status, solo.status)
        assert np.array_equal(res.p, solo.p) and np.array_equal(res.y, solo.y), (k, d["name"])
        assert _agree(res, solo), (k, d["name"], res.status, solo.status, res.pobj, solo.pobj)
        if d["B"].shape[1] <= 300:
            ores = O.solve(oracle_problem(d))
            assert int(ores.status) in (int(x) for x in GOOD)
            assert _agree(res, ores), (k, d["name"], res.status, ores.status, res.pobj, ores.pobj)
            checked += 1
    assert all(v > 0 for v in kinds.values()), kinds
    assert checked >= 20
