"""Solves are reproducible bit for bit: the persistent factorization subtracts each tile's updates
in a fixed column order, the pipelined triangular solves sum their contributions in a fixed block
order however the published blocks were batched, and the extend-add adds a front's children in a
fixed order (one child per parent in each launch), and data passed between the workgroups of one
launch moves with the device-coherent cache policy. So results do not depend on the run, on the
number of workgroups the persistent kernels get, on the concurrency of independent fronts, or on
other work sharing the GPU."""
import sys

import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem

pytestmark = pytest.mark.gpu

CASES = {
    "entropy200": lambda: G.exppow(nexp=200, npow=0, m=0, eps=0.0, seed=0),
    "qp1500": lambda: G.qp(n=1500, m=700, seed=3),
    "socp900": lambda: G.socp(n=900, m=450, ncone=40, seed=14, adens=0.01),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("precision", [0, 1])
def test_single_front_solves_repeat_exactly(name, precision):
    print("case", name, "precision", precision, flush=True)
    sys.stdout.flush()
    prob = hip_problem(CASES[name]())
    st = mh.IPMSettings(precision=precision)
    runs = []
    for _ in range(3):
        s = mh.init(prob, st)
        assert s.stats()["nfront"] == 1
        runs.append(s.solve())
        s.close()
    for r in runs[1:]:
        assert r.status == runs[0].status and r.niter == runs[0].niter
        assert np.array_equal(r.p, runs[0].p) and np.array_equal(r.d, runs[0].d) and np.array_equal(r.y, runs[0].y)


MULTI = {
    "maxcut_chordal": lambda: G.maxcut_chordal(nv=160, band=6, extra=2e-3, seed=5, merge=24),
    "qp_multifront": lambda: G.qp(n=900, m=400, seed=8),
}


@pytest.mark.parametrize("name", sorted(MULTI))
def test_multi_front_solves_repeat_exactly(name):
    print("case", name, flush=True)
    sys.stdout.flush()
    prob = hip_problem(MULTI[name]())
    runs = []
    for grid in (0, 0, 24):
        s = mh.init(prob, mh.IPMSettings(big_front_ld=64, dag_grid=grid))
        assert s.stats()["nfront"] > 1
        runs.append(s.solve())
        s.close()
    assert runs[0].status == mh.IPMStatus.OPTIMAL
    for r in runs[1:]:
        assert r.status == runs[0].status and r.niter == runs[0].niter
        assert np.array_equal(r.p, runs[0].p) and np.array_equal(r.d, runs[0].d) and np.array_equal(r.y, runs[0].y)


def test_small_front_solves_repeat_exactly_under_concurrency():
    """Small fronts (order < 512) take the level-batched panel kernels. Solved alone and as eight
    copies at once from eight host threads (their kernels interleaving on the device, so the
    workgroups of one panel launch start at different times), the results are the same bit for
    bit: a row tile of a panel step reads the diagonal factor only after the diagonal workgroup
    published it."""
    import threading
    print("case small fronts under concurrency", flush=True)
    sys.stdout.flush()
    data = [G.qp(n=200, m=100, seed=64 + k) for k in range(4)] + \
           [G.socp(n=300, m=140, ncone=12, seed=70 + k, adens=0.03) for k in range(4)]
    probs = [hip_problem(d) for d in data]
    solo = [mh.solve(p) for p in probs]
    for rep in range(3):
        solvers = [mh.init(p) for p in probs]
        out = [None] * len(solvers)

        def run(k):
            out[k] = solvers[k].solve()

        threads = [threading.Thread(target=run, args=(k,)) for k in range(len(solvers))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for s in solvers:
            s.close()
        for k, (a, b) in enumerate(zip(out, solo)):
            assert a.status == b.status and a.niter == b.niter, (rep, k)
            assert np.array_equal(a.p, b.p) and np.array_equal(a.y, b.y), (rep, k)


def test_served_mix_repeats_exactly_beside_other_gpu_work():
    """The first 192 requests of the served mix (QP / SOCP / chordal max-cut SDP, orders 50 to
    4000: small fronts, large fronts, PSD cones), solved once alone and once while another host
    thread keeps the GPU busy with dense 3000×3000 Cholesky factorizations (the persistent kernel
    on every CU) on the null stream. Every result repeats bit for bit. The other work shifts when
    each workgroup starts, which is what exposed a panel step's row tiles reading the diagonal
    factor through a cache that still held the line from before it was stored (about 1 solve in
    40 differed under such load)."""
    import threading
    from devmem import DeviceArray, hip
    print("case served mix beside other GPU work", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(7)
    probs = [hip_problem(G.served_request(rng, k, 4000)) for k in range(192)]
    solo = [mh.solve(p) for p in probs]
    n = 3000
    M = np.random.default_rng(1).standard_normal((n, n))
    A = M @ M.T / n + np.eye(n)
    src, work = DeviceArray(A), DeviceArray(A)
    stop = threading.Event()
    count = [0]

    def load():
        while not stop.is_set():
            assert hip().hipMemcpy(work.ptr, src.ptr, A.nbytes, 3) == 0
            mh.dense_cholesky(work.ptr, n, n)
            count[0] += 1

    bg = threading.Thread(target=load)
    bg.start()
    try:
        again = [mh.solve(p) for p in probs]
    finally:
        stop.set()
        bg.join()
        src.free()
        work.free()
    assert count[0] > 0
    bad = [k for k, (a, b) in enumerate(zip(again, solo))
           if not (a.status == b.status and a.niter == b.niter and np.array_equal(a.p, b.p) and
                   np.array_equal(a.y, b.y))]
    assert not bad, bad
