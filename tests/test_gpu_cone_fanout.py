"""The cone fan-outs through the C ABI (mb_cone_scale / _blocks / _corr0 / _corr / _maxsteps) at
random interior points of every cone kind, cone by cone against the oracle's restatement of
pos.jl, noc.jl, soc.jl, sdp.jl and tdc.jl (src/IPM/src/cone/).

Tolerances: nonnegative and second-order cones are closed-form arithmetic (1e-12 relative to the
block's norm). The semidefinite block goes through an SVD (one-sided Jacobi on the device, LAPACK
in the oracle) and the exponential/power blocks through a Newton solve for the dual gradient
seeded differently (1e-9). The exponential/power step comes out of a certified bisection whose
bracket is a few noise bands (65536·eps / |g′|) wide, so two correct implementations agree to
about that width (1e-8)."""
import math
import sys

import numpy as np
import pytest
import scipy.sparse as sp

import mumblebee_hip as mh
from oracle import cones as C
from oracle import tdc as T

pytestmark = pytest.mark.gpu

KIND = {0: "pos", 1: "free", 2: "soc", 3: "psd", 4: "exp", 5: "pow"}
TOL = {"pos": 1e-12, "free": 0.0, "soc": 1e-12, "psd": 1e-9, "exp": 1e-9, "pow": 1e-9}


def _print_case(**kw):
    print("case", kw, flush=True)
    sys.stdout.flush()


def _cone_of(kind, param):
    k = KIND[int(kind)]
    return ("pow", float(param)) if k == "pow" else (k,)


def _problem(seed):
    """pos 5, free 3, soc 6, psd 4×4 and 7×7, three exponential and three power cones."""
    K = [mh.PositiveCone(), mh.CofreeCone(), mh.SecondOrderCone(), mh.SemidefiniteCone(), mh.SemidefiniteCone(),
         mh.ExponentialCone(), mh.ExponentialCone(), mh.ExponentialCone(), mh.PowerCone(0.3), mh.PowerCone(0.7),
         mh.PowerCone(0.5)]
    s = [5, 3, 6, 10, 28, 3, 3, 3, 3, 3, 3]
    n, m = sum(s), 4
    rng = np.random.default_rng(seed)
    B = sp.csc_matrix(rng.normal(size=(m, n)))
    Q = sp.csc_matrix((n, n))
    return mh.IPMProblem(Q, B, rng.normal(size=n), rng.normal(size=m), 0.0, K, s)


def _primal_point(rng, cone, nv):
    k = cone[0]
    if k == "pos":
        return rng.uniform(0.3, 3.0, nv)
    if k == "free":
        return rng.normal(size=nv)
    if k == "soc":
        x = rng.normal(size=nv - 1)
        return np.concatenate([[np.linalg.norm(x) * rng.uniform(1.1, 3.0) + 0.1], x])
    if k == "psd":
        d = C.triroot(nv)
        A = rng.normal(size=(d, d))
        return C.svec(A @ A.T / d + 0.3 * np.eye(d))
    if k == "exp":
        y, z = rng.uniform(0.5, 2.0), rng.uniform(-1.0, 1.0)
        return np.array([y * math.exp(z / y) * rng.uniform(1.2, 3.0), y, z])
    al = cone[1]
    x, y = rng.uniform(0.5, 2.0, 2)
    return np.array([x, y, rng.uniform(-0.9, 0.9) * x ** al * y ** (1 - al)])


def _dual_point(rng, cone, nv):
    k = cone[0]
    if k == "free":
        return np.zeros(nv)
    if k in ("exp", "pow"):  # −f′ of a primal interior point lies in the dual interior
        s = np.zeros(3)
        T.dual_shadow(cone, s, _primal_point(rng, cone, nv))
        return s * rng.uniform(0.5, 2.0)
    return _primal_point(rng, cone, nv)  # self-dual


class _Fixture:
    def __init__(self, seed, dscale=0.5):
        self.prob = _problem(seed)
        self.s = mh.init(self.prob)
        lay = self.s.layout()
        self.xcol = lay["xcol"]
        self.K = [_cone_of(k, a) for k, a in zip(lay["kind"], lay["param"])]
        rng = np.random.default_rng(100 + seed)
        n = self.prob.n
        self.p, self.d, self.dp, self.dd = (np.zeros(n) for _ in range(4))
        for v, cone in enumerate(self.K):
            a, b = self.xcol[v], self.xcol[v + 1]
            self.p[a:b] = _primal_point(rng, cone, b - a)
            self.d[a:b] = _dual_point(rng, cone, b - a)
            self.dp[a:b] = rng.normal(size=b - a) * dscale * np.linalg.norm(self.p[a:b])
            self.dd[a:b] = rng.normal(size=b - a) * dscale * np.linalg.norm(self.d[a:b])
            if cone[0] == "free":
                self.dd[a:b] = 0.0
        # the oracle at the same point, cone by cone
        self.caches = [C.make_cache(cone, self.xcol[v + 1] - self.xcol[v]) for v, cone in enumerate(self.K)]
        self.H, self.flags, self.spsd = [], [], 0.0
        for v, cone in enumerate(self.K):
            a, b = self.xcol[v], self.xcol[v + 1]
            H = np.zeros((b - a, b - a))
            fl, sv = C.scale(cone, H, self.p[a:b], self.d[a:b], self.caches[v])
            self.H.append(H)
            self.flags.append(fl)
            self.spsd += sv

    def ranges(self):
        for v, cone in enumerate(self.K):
            yield v, cone, self.xcol[v], self.xcol[v + 1]

    def close(self):
        self.s.close()


@pytest.fixture(scope="module", params=[1, 2, 3])
def fx(request):
    _print_case(test="cone fan-out", seed=request.param)
    f = _Fixture(request.param)
    yield f
    f.close()


def test_every_kind_is_present(fx):
    assert {c[0] for c in fx.K} == set(KIND.values())


def test_scale_blocks_match_oracle(fx):
    ok, spsd = fx.s.cone_scale(fx.p, fx.d)
    assert ok and all(fx.flags)
    assert spsd == pytest.approx(fx.spsd, rel=1e-9)
    D = fx.s.cone_blocks()
    o = 0
    for v, cone, a, b in fx.ranges():
        nv = b - a
        if cone[0] == "pos":
            blk, ref = D[o:o + nv], np.diag(fx.H[v])
            o += nv
        elif cone[0] == "free":
            continue
        else:
            blk, ref = D[o:o + nv * nv].reshape(nv, nv, order="F"), fx.H[v]
            o += nv * nv
        err = np.linalg.norm(blk - ref) / np.linalg.norm(ref)
        assert err < TOL[cone[0]], (v, cone, err)
    assert o == len(D)


@pytest.mark.parametrize("smu", [0.0, 0.37])
def test_predictor_rhs_matches_oracle(fx, smu):
    fx.s.cone_scale(fx.p, fx.d)
    r = fx.s.cone_corr0(smu)
    for v, cone, a, b in fx.ranges():
        ref = np.zeros(b - a)
        C.corr0(cone, ref, fx.p[a:b], fx.d[a:b], smu, fx.caches[v])
        err = np.linalg.norm(r[a:b] - ref) / max(1.0, np.linalg.norm(ref))
        assert err < max(TOL[cone[0]], 1e-12), (v, cone, err)


@pytest.mark.parametrize("smu", [0.0, 0.37])
def test_corrector_rhs_matches_oracle(fx, smu):
    fx.s.cone_scale(fx.p, fx.d)
    r = fx.s.cone_corr(fx.dp, fx.dd, smu)
    for v, cone, a, b in fx.ranges():
        ref = np.zeros(b - a)
        C.corr(cone, ref, fx.p[a:b], fx.d[a:b], fx.dp[a:b], fx.dd[a:b], smu, fx.caches[v])
        err = np.linalg.norm(r[a:b] - ref) / max(1.0, np.linalg.norm(ref))
        assert err < max(TOL[cone[0]], 1e-12), (v, cone, err)


def test_maxsteps_cone_by_cone(fx):
    """One call per cone with the other cones' directions zero, and one with all of them."""
    fx.s.cone_scale(fx.p, fx.d)
    zero = np.zeros(fx.prob.n)
    taus = []
    for v, cone, a, b in fx.ranges():
        dp, dd = zero.copy(), zero.copy()
        dp[a:b], dd[a:b] = fx.dp[a:b], fx.dd[a:b]
        tp, td = C.maxsteps(cone, fx.p[a:b], dp[a:b], fx.d[a:b], dd[a:b], fx.caches[v])
        ref = min(1.0, tp, td)
        got = fx.s.cone_maxsteps(dp, dd, 1.0)
        tol = 1e-8 if cone[0] in ("exp", "pow") else 1e-10
        assert abs(got - ref) <= tol * ref, (v, cone, got, ref)
        taus.append((cone[0], tp, td))
    assert any(min(tp, td) < 1.0 for _, tp, td in taus)  # some step is blocked
    ref = min(min(tp, td) * (1.0 if k == "free" else 0.99) for k, tp, td in taus)
    got = fx.s.cone_maxsteps(fx.dp, fx.dd, 0.99)
    assert got == pytest.approx(ref, rel=1e-8)


def test_scale_flags_at_points_outside_the_cones():
    """The flag of scale! where p leaves a cone, as the reference sets it: the second-order,
    semidefinite and exponential factorizations fail; posscale! (pos.jl:75-84) never checks, and
    the power cone's factorization can succeed outside the cone (the oracle says which)."""
    _print_case(test="cone fan-out flags")
    f = _Fixture(4)
    try:
        for kind in ("pos", "soc", "psd", "exp", "pow"):
            v, cone, a, b = next(t for t in f.ranges() if t[1][0] == kind)
            p = f.p.copy()
            if kind == "pos":
                p[a] = -1.0
            elif kind == "soc":
                p[a] = 0.5 * np.linalg.norm(p[a + 1:b])
            elif kind == "psd":
                dd = C.triroot(b - a)
                p[a:b] = C.svec(np.diag(np.concatenate([[-1.0], np.ones(dd - 1)])))
            elif kind == "exp":
                p[a:b] = [0.1, 1.0, 0.5]  # y e^{z/y} > x
            else:
                p[a:b] = [1.0, 1.0, 2.0]  # x^α y^{1−α} < |z|
            ref, _ = C.scale(cone, np.zeros((b - a, b - a)), p[a:b], f.d[a:b], C.make_cache(cone, b - a))
            assert ref == (kind in ("pos", "pow"))
            ok, _ = f.s.cone_scale(p, f.d)
            assert ok == ref, kind
            ok, _ = f.s.cone_scale(f.p, f.d)
            assert ok
    finally:
        f.close()


def test_fanout_vectors_are_checked():
    _print_case(test="cone fan-out lengths")
    f = _Fixture(5)
    try:
        with pytest.raises(ValueError):
            f.s.cone_scale(f.p[:-1], f.d)
    finally:
        f.close()


def test_kkt_solve_matches_a_dense_saddle_point_solve():
    """initkkt! + solvekkt! through the C ABI (uzawa.jl:73-115, 265-344) on the H that scale!
    formed at an interior point: the solution of H x − Bᵀ y = f, B x = g matches a dense solve of
    the same saddle-point system (equilibration off, so the internal coordinates are the caller's
    permuted by P1 and P2)."""
    _print_case(test="kkt solve")
    K = [mh.PositiveCone(), mh.CofreeCone(), mh.SecondOrderCone(), mh.SemidefiniteCone(), mh.ExponentialCone(),
         mh.PowerCone(0.4)]
    sizes = [6, 3, 5, 10, 3, 3]
    n, m = sum(sizes), 7
    rng = np.random.default_rng(77)
    A = rng.normal(size=(n, 3))
    Q = sp.csc_matrix(0.1 * A @ A.T)
    B = sp.csc_matrix(rng.normal(size=(m, n)))
    prob = mh.IPMProblem(Q, B, rng.normal(size=n), rng.normal(size=m), 0.0, K, sizes)
    s = mh.init(prob, mh.IPMSettings(scale_max_iter=0))
    try:
        lay = s.layout()
        assert np.all(lay["rowscale"] == 1.0) and np.all(lay["colscale"] == 1.0)
        P1, P2, xcol = lay["P1"], lay["P2"], lay["xcol"]
        cones = [_cone_of(k, a) for k, a in zip(lay["kind"], lay["param"])]
        p, d = np.zeros(n), np.zeros(n)
        for v, cone in enumerate(cones):
            a, b = xcol[v], xcol[v + 1]
            p[a:b] = _primal_point(rng, cone, b - a)
            d[a:b] = _dual_point(rng, cone, b - a)
        ok, _ = s.cone_scale(p, d)
        assert ok
        D = s.cone_blocks()
        H = Q.toarray()[np.ix_(P2, P2)]
        o = 0
        for v, cone in enumerate(cones):
            a, b = xcol[v], xcol[v + 1]
            nv = b - a
            if cone[0] == "pos":
                H[np.arange(a, b), np.arange(a, b)] += D[o:o + nv]
                o += nv
            elif cone[0] != "free":
                H[a:b, a:b] += D[o:o + nv * nv].reshape(nv, nv, order="F")
                o += nv * nv
        Bi = B.toarray()[np.ix_(P1, P2)]
        okf, rho = s.kkt_factor(1.0)
        assert okf and rho == 0.0
        f, g = rng.normal(size=n), rng.normal(size=m)
        x, y, ncg, nir, status, dmin, dmax = s.kkt_solve(f, g, ftol=1e-12, gtol=1e-12)
        assert status in (mh.KKTStatus.KKT_SOLVED, mh.KKTStatus.KKT_STAGNATED) and nir >= 1
        Kd = np.block([[H, -Bi.T], [Bi, np.zeros((m, m))]])
        ref = np.linalg.solve(Kd, np.concatenate([f, g]))
        assert np.linalg.norm(x - ref[:n]) <= 1e-8 * (1 + np.linalg.norm(ref[:n]))
        assert np.linalg.norm(y - ref[n:]) <= 1e-8 * (1 + np.linalg.norm(ref[n:]))
        # a warm start from the solution stops at once
        x2, y2, _, nir2, status2, _, _ = s.kkt_solve(f, g, x0=x, y0=y, ftol=1e-9, gtol=1e-9)
        assert status2 == mh.KKTStatus.KKT_SOLVED and nir2 == 0
        assert np.linalg.norm(x2 - x) <= 1e-9 * (1 + np.linalg.norm(x))
    finally:
        s.close()
