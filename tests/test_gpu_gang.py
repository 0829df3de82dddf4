"""Lockstep batching (IPMGang): several solvers advance together and every phase of their IPM
steps is one launch for all of them. Each member keeps its own decisions and its own data, and
the combined kernels compute every instance as the solo kernels do (fixed summation orders), so
each result must equal the same solver's solo result bit for bit: status, iteration count and
the iterate. Covered: small fronts only (level-batched panels across instances), one large front
per instance (one persistent factorization and one pipelined solve launch for all), mixed cone
kinds with PSD blocks, mixed precision, more solvers than slots (members joining as others
leave), a member the oracle also solves, and the serving loop with its slots as gang members."""
import sys
import threading

import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem, rel

pytestmark = pytest.mark.gpu


def _same(a, b, tag):
    assert a.status == b.status and a.niter == b.niter, (tag, a.status, b.status, a.niter, b.niter)
    assert np.array_equal(a.p, b.p) and np.array_equal(a.d, b.d) and np.array_equal(a.y, b.y), tag


def _gang_vs_solo(data, capacity, settings=None, tag=""):
    print("case", tag, "members", len(data), "capacity", capacity, flush=True)
    sys.stdout.flush()
    probs = [hip_problem(d) for d in data]
    st = settings or mh.IPMSettings()
    solo = [mh.init(p, st).solve() for p in probs]
    g = mh.IPMGang(capacity=capacity)
    solvers = [mh.init(p, st) for p in probs]
    out = g.solve(solvers)
    stats = g.stats()
    for k, (a, b) in enumerate(zip(out, solo)):
        _same(a, b, (tag, k))
    for s in solvers:
        s.close()
    g.close()
    assert stats["joins"] == len(data)
    assert stats["combined_ops"] > 0  # launches were shared
    return out, stats


def test_gang_small_fronts_bit_exact():
    data = [G.qp(n=200 + 20 * k, m=90 + 10 * k, seed=100 + k) for k in range(4)]
    _, st = _gang_vs_solo(data, 4, tag="qp small")
    assert st["combined_ops"] >= 0.5 * st["ops"]


def test_gang_large_fronts_bit_exact():
    # one front of order ≥ 512 per instance: the persistent factorization and the pipelined
    # solves of all members in one launch each
    data = [G.qp(n=900, m=400, seed=200 + k) for k in range(3)]
    _, st = _gang_vs_solo(data, 3, tag="qp large")
    assert st["factor_fronts"] >= 2 * st["factor_launches"] > 0


def test_gang_odd_leading_dimension_bit_exact():
    # fronts of order 680..735 (a column is not a whole number of cache lines, so a line holds
    # rows of two tiles) and ≥ 8 of them per combined factorization (per-XCD task queues)
    data = [G.qp(n=680 + 5 * k, m=340, seed=250 + k) for k in range(12)]
    _, st = _gang_vs_solo(data, 12, tag="odd ld")
    assert st["factor_fronts"] >= 4 * st["factor_launches"] > 0


def test_gang_mixed_cones_bit_exact():
    data = [G.qp(n=300, m=150, seed=300), G.socp(n=400, m=200, ncone=15, seed=301, adens=0.02),
            G.maxcut(nv=60, band=4, clique=12, extra=0.01, seed=302),
            G.socp(n=600, m=300, ncone=25, seed=303, adens=0.015)]
    _gang_vs_solo(data, 4, tag="mixed cones")


def test_gang_mixed_precision_bit_exact():
    data = [G.qp(n=1000, m=450, seed=400 + k) for k in range(3)]
    _gang_vs_solo(data, 3, settings=mh.IPMSettings(precision=1), tag="mixed precision")


def test_gang_members_join_as_others_leave():
    # six solvers through three slots: each finished member's slot takes the next solver
    data = [G.qp(n=150 + 40 * k, m=70 + 20 * k, seed=500 + k) for k in range(6)]
    _, st = _gang_vs_solo(data, 3, tag="joins")
    assert st["members"] == 0


def test_gang_member_matches_oracle():
    from oracle import ipm as O
    print("case gang vs oracle", flush=True)
    data = [G.qp(n=250, m=110, seed=600 + k) for k in range(3)]
    g = mh.IPMGang(capacity=3)
    solvers = [mh.init(hip_problem(d)) for d in data]
    out = g.solve(solvers)
    for d, r in zip(data, out):
        o = O.solve(oracle_problem(d))
        assert r.status == mh.IPMStatus.OPTIMAL and o.status == O.OPTIMAL
        assert rel(r.pobj, o.pobj) < 1e-8 and rel(r.dobj, o.dobj) < 1e-8
    for s in solvers:
        s.close()
    g.close()


def test_gang_solve_in_from_own_threads():
    # the per-solver entry point, each solver on a thread of the caller
    print("case solve_in threads", flush=True)
    data = [G.socp(n=300, m=140, ncone=12, seed=700 + k, adens=0.03) for k in range(4)]
    probs = [hip_problem(d) for d in data]
    solo = [mh.solve(p) for p in probs]
    g = mh.IPMGang(capacity=4)
    solvers = [mh.init(p) for p in probs]
    stat = [None] * 4

    def run(k):
        stat[k] = g.solve_in(solvers[k])

    th = [threading.Thread(target=run, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k in range(4):
        _same(solvers[k].result(), solo[k], k)
        solvers[k].close()
    g.close()


def test_served_gang_matches_solo():
    print("case served gang", flush=True)
    data = [G.qp(n=120 + 30 * k, m=60 + 10 * k, seed=800 + k) for k in range(8)] + \
           [G.socp(n=350, m=160, ncone=14, seed=810 + k, adens=0.03) for k in range(4)]
    probs = [hip_problem(d) for d in data]
    solo = [mh.solve(p) for p in probs]
    srv = mh.IPMServer(nslots=4, gang=True)
    tickets = [srv.submit(p) for p in probs]
    for k, t in enumerate(tickets):
        r, times = srv.wait(t, timeout=120.0)
        assert r.status == solo[k].status and r.niter == solo[k].niter, k
        assert np.array_equal(r.p, solo[k].p) and np.array_equal(r.y, solo[k].y), k
        assert 0 < times["device_bytes"] <= times["reserved_bytes"]
    srv.close()
