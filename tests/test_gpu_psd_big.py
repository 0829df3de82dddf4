"""PSD cones above the LDS kernels' order (d > 48) on the device: the HBM-resident pipeline of
psd_big.hip (tile products on MFMA, blocked Cholesky, one-sided Jacobi, Householder + Sturm) and
the implicit scaling block W ⊗ₛ W.

Cone by cone through the C ABI at random interior points against the oracle's restatement of
sdp.jl (LAPACK SVD and eigvalsh there, Jacobi and Sturm here: 1e-9); whole solves against closed
forms (the trace SDP attains λmin(C)), against the oracle, and the chordal max-cut conversion
against the dense max-cut SDP it decomposes."""
import numpy as np
import pytest
import scipy.sparse as sp

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem, rel
from oracle import cones as C
from test_gpu_cone_fanout import _cone_of, _dual_point, _primal_point, _print_case
from test_gpu_psd import _trace_sdp

pytestmark = pytest.mark.gpu


def _problem(orders, seed):
    K = [mh.PositiveCone(), mh.SecondOrderCone()] + [mh.SemidefiniteCone() for _ in orders]
    s = [4, 5] + [d * (d + 1) // 2 for d in orders]
    n, m = sum(s), 3
    rng = np.random.default_rng(seed)
    B = sp.csc_matrix(rng.normal(size=(m, n)))
    return mh.IPMProblem(sp.csc_matrix((n, n)), B, rng.normal(size=n), rng.normal(size=m), 0.0, K, s)


class _Big:
    def __init__(self, orders, seed, dscale=0.5):
        self.prob = _problem(orders, seed)
        self.s = mh.init(self.prob, mh.IPMSettings(scale_max_iter=0))
        lay = self.s.layout()
        self.xcol = lay["xcol"]
        self.K = [_cone_of(k, a) for k, a in zip(lay["kind"], lay["param"])]
        rng = np.random.default_rng(200 + seed)
        n = self.prob.n
        self.p, self.d, self.dp, self.dd = (np.zeros(n) for _ in range(4))
        for v, cone in enumerate(self.K):
            a, b = self.xcol[v], self.xcol[v + 1]
            self.p[a:b] = _primal_point(rng, cone, b - a)
            self.d[a:b] = _dual_point(rng, cone, b - a)
            self.dp[a:b] = rng.normal(size=b - a) * dscale * np.linalg.norm(self.p[a:b])
            self.dd[a:b] = rng.normal(size=b - a) * dscale * np.linalg.norm(self.d[a:b])
        self.caches = [C.make_cache(cone, self.xcol[v + 1] - self.xcol[v]) for v, cone in enumerate(self.K)]
        self.H, self.spsd = [], 0.0
        for v, cone in enumerate(self.K):
            a, b = self.xcol[v], self.xcol[v + 1]
            H = np.zeros((b - a, b - a))
            fl, sv = C.scale(cone, H, self.p[a:b], self.d[a:b], self.caches[v])
            assert fl
            self.H.append(H)
            self.spsd += sv

    def ranges(self):
        for v, cone in enumerate(self.K):
            yield v, cone, self.xcol[v], self.xcol[v + 1]


@pytest.fixture(scope="module", params=[(1, (49, 64)), (2, (100,)), (3, (57, 90))])
def bx(request):
    seed, orders = request.param
    _print_case(test="large PSD fan-out", seed=seed, orders=orders)
    f = _Big(orders, seed)
    yield f
    f.s.close()


def test_scale_blocks_match_oracle(bx):
    ok, spsd = bx.s.cone_scale(bx.p, bx.d)
    assert ok
    assert spsd == pytest.approx(bx.spsd, rel=1e-9)
    D = bx.s.cone_blocks()
    o = 0
    for v, cone, a, b in bx.ranges():
        nv = b - a
        if cone[0] == "pos":
            blk, ref = D[o:o + nv], np.diag(bx.H[v])
            o += nv
        else:
            blk, ref = D[o:o + nv * nv].reshape(nv, nv, order="F"), bx.H[v]
            o += nv * nv
        err = np.linalg.norm(blk - ref) / np.linalg.norm(ref)
        assert err < 1e-9, (v, cone, nv, err)
    assert o == len(D)


@pytest.mark.parametrize("smu", [0.0, 0.37])
def test_predictor_and_corrector_rhs_match_oracle(bx, smu):
    bx.s.cone_scale(bx.p, bx.d)
    r0 = bx.s.cone_corr0(smu)
    r1 = bx.s.cone_corr(bx.dp, bx.dd, smu)
    for v, cone, a, b in bx.ranges():
        ref0, ref1 = np.zeros(b - a), np.zeros(b - a)
        C.corr0(cone, ref0, bx.p[a:b], bx.d[a:b], smu, bx.caches[v])
        C.corr(cone, ref1, bx.p[a:b], bx.d[a:b], bx.dp[a:b], bx.dd[a:b], smu, bx.caches[v])
        assert np.linalg.norm(r0[a:b] - ref0) / max(1.0, np.linalg.norm(ref0)) < 1e-9, (v, cone)
        assert np.linalg.norm(r1[a:b] - ref1) / max(1.0, np.linalg.norm(ref1)) < 1e-9, (v, cone)


def test_maxsteps_match_oracle(bx):
    bx.s.cone_scale(bx.p, bx.d)
    zero = np.zeros(bx.prob.n)
    blocked = False
    for v, cone, a, b in bx.ranges():
        if cone[0] != "psd":
            continue
        dp, dd = zero.copy(), zero.copy()
        dp[a:b], dd[a:b] = bx.dp[a:b], bx.dd[a:b]
        tp, td = C.maxsteps(cone, bx.p[a:b], dp[a:b], bx.d[a:b], dd[a:b], bx.caches[v])
        ref = min(1.0, tp, td)
        blocked |= ref < 1.0
        got = bx.s.cone_maxsteps(dp, dd, 1.0)
        assert abs(got - ref) <= 1e-10 * ref, (v, got, ref)
        # a direction that keeps the cone (the Cholesky test passes): step 1
        got = bx.s.cone_maxsteps(0.1 * bx.p * (np.arange(bx.prob.n) >= a) * (np.arange(bx.prob.n) < b),
                                 zero, 1.0)
        assert got == 1.0
    assert blocked


def test_scale_flag_outside_the_cone():
    _print_case(test="large PSD flag")
    f = _Big((64,), 7)
    try:
        v, cone, a, b = next(t for t in f.ranges() if t[1][0] == "psd")
        d = C.triroot(b - a)
        p = f.p.copy()
        p[a:b] = C.svec(np.diag(np.concatenate([[-1.0], np.ones(d - 1)])))
        ok, _ = f.s.cone_scale(p, f.d)
        assert not ok
        ok, _ = f.s.cone_scale(f.p, f.d)
        assert ok
    finally:
        f.s.close()


@pytest.mark.parametrize("d,seed", [(49, 11), (64, 12), (128, 13), (200, 14)])
def test_trace_sdp_attains_lambda_min(d, seed):
    _print_case(test="trace SDP", d=d)
    data, lam = _trace_sdp(d, seed)
    res = mh.solve(hip_problem(data))
    assert res.status == mh.IPMStatus.OPTIMAL
    assert abs(res.pobj - lam) <= 1e-8 * max(1.0, abs(lam))
    assert abs(res.dobj - lam) <= 1e-8 * max(1.0, abs(lam))
    X = C.smat(res.p)
    assert np.linalg.eigvalsh(X)[0] >= -1e-10


def test_trace_sdp_matches_oracle():
    from oracle import ipm as O
    data, lam = _trace_sdp(64, 21)
    res = mh.solve(hip_problem(data))
    ores = O.solve(oracle_problem(data))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8
    assert abs(res.niter - ores.niter) <= 2


def test_chordal_maxcut_matches_oracle():
    """Cliques above 48 vertices from a chordal extension with uniformly random extra edges."""
    from oracle import ipm as O
    data = G.maxcut_chordal(nv=120, band=6, extra=0.01, seed=3, merge=64)
    assert max(len(c) for c in data["cliques"]) > 48
    _print_case(test="chordal max-cut vs oracle", n=data["B"].shape[1], m=data["B"].shape[0])
    res = mh.solve(hip_problem(data))
    ores = O.solve(oracle_problem(data))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8


def test_chordal_conversion_has_the_dense_optimum():
    """The clique decomposition and the dense SDP on one PSD cone of order nv (the large-cone
    path) reach the same optimum; every clique block of the solution is PSD."""
    nv = 60
    data = G.maxcut_chordal(nv=nv, band=4, extra=0.05, seed=1, merge=0)
    # the dense SDP: recover C from the chordal problem's objective (each entry owned once)
    Cm = np.zeros((nv, nv))
    col = 0
    seen = set()
    for c, sz in zip(data["cliques"], data["s"]):
        f = -data["f"][col:col + sz]
        d = len(c)
        k = 0
        for j in range(d):
            for i in range(j, d):
                a, b = c[i], c[j]
                key = (max(a, b), min(a, b))
                val = f[k] if i == j else f[k] / np.sqrt(2.0)
                if val != 0.0 and key not in seen:
                    seen.add(key)
                    Cm[key[0], key[1]] = Cm[key[1], key[0]] = val
                k += 1
        col += sz
    nd = nv * (nv + 1) // 2
    rows, cols = [], []
    k = 0
    for j in range(nv):
        rows.append(j)
        cols.append(k)
        k += nv - j
    Bd = sp.csc_matrix((np.ones(nv), (rows, cols)), shape=(nv, nd))
    dense = dict(Q=sp.csc_matrix((nd, nd)), B=Bd, f=-C.svec(Cm), g=np.ones(nv), K=[("psd",)], s=[nd])
    rd = mh.solve(hip_problem(dense))
    rc = mh.solve(hip_problem(data))
    assert rd.status == rc.status == mh.IPMStatus.OPTIMAL
    assert rel(rc.pobj, rd.pobj) < 1e-7
    col = 0
    for sz in data["s"]:
        assert np.linalg.eigvalsh(C.smat(rc.p[col:col + sz]))[0] >= -1e-9
        col += sz
