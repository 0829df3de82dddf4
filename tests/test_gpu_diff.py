"""Differentiation of the solution map on the device: frule! and rrule! (src/IPM/src/ipm/
differential.jl:5-75, 252-334) through the C ABI, at the solver's last iterate.

What the mathematics fixes, and what is checked here:
(1) free cones only (ν = 0): the solution of an equality-constrained QP is linear in (f, g), and
    frule! must be the exact KKT solution of the seeds;
(2) on the central path (a target barrier μt > 0 solved to gap_tol 1e-12) the last iterate is the
    barrier problem's optimum, the NT scaling there is μt ∇²F(p), and frule!(Δf, Δg, Δμ) is the
    exact derivative of the map (f, g, μt) ↦ (p, y, d): central finite differences of whole solves,
    with μt perturbed too, agree to ~1e-6 for every cone kind (nonnegative, second-order, PSD of
    both device paths);
(3) rrule! is the adjoint of frule!: ⟨p̄, Δp⟩ + ⟨ȳ, Δy⟩ + ⟨d̄, Δd⟩ = ⟨f̄, Δf⟩ + ⟨ḡ, Δg⟩ + μ̄ Δμ;
(4) the device's frule! and rrule! equal the oracle's on the same central-path point.

At the optimum itself (μt = 0) the last NT scaling is taken at an iterate that is only roughly
centred. Where a cone's p and d both sit on its boundary (second-order and PSD cones with strict
complementarity), the scaling's tangential curvature, which the derivative depends on, is then off
by the centring error, and frule! is an approximation of the optimum's derivative (the oracle shows
the same 50-80 % difference from finite differences). On nonnegative cones no curvature enters,
and there the optimum's finite differences are checked. The Δμ seed at the optimum multiplies the
dual shadow d* = 1/p, which is ~1/μ on active bounds; the adjoint identity at the optimum is
therefore checked with Δμ = 0 (to 1e-5: the solves stop at the reference's forcing tolerance on a
KKT matrix of condition ~1/μ) and with Δμ ≠ 0 on the central path (to 1e-8)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem
from test_gpu_psd import _trace_sdp

pytestmark = pytest.mark.gpu

MUT = 1e-2
TIGHT = dict(gap_tol=1e-12)


def _solved(d, mu=0.0, **kw):
    s = mh.init(hip_problem(d, mu=mu), mh.IPMSettings(**kw) if kw else None)
    r = s.solve()
    assert r.status == mh.IPMStatus.OPTIMAL, r.status
    return s, r


def _rel(a, b):
    return np.linalg.norm(a - b) / max(1e-12, np.linalg.norm(b))


def test_frule_of_an_equality_qp_is_its_kkt_solve():
    rng = np.random.default_rng(51)
    n, m = 30, 9
    M = rng.standard_normal((n, n))
    Q = M @ M.T / n + np.eye(n)
    A = rng.standard_normal((m, n))
    d = dict(Q=sp.csc_matrix(Q), B=sp.csc_matrix(A), f=rng.standard_normal(n), g=rng.standard_normal(m),
             K=[("free",)], s=[n])
    s, _ = _solved(d)
    try:
        df, dg = rng.standard_normal(n), rng.standard_normal(m)
        dp, dy, dd = s.frule(df, dg, 0.0)
        K = np.block([[Q, -A.T], [A, np.zeros((m, m))]])
        ref = np.linalg.solve(K, np.concatenate([df, dg]))
        assert np.linalg.norm(dp - ref[:n]) <= 1e-8 * np.linalg.norm(ref[:n])
        assert np.linalg.norm(dy - ref[n:]) <= 1e-8 * np.linalg.norm(ref[n:])
        assert np.linalg.norm(dd) <= 1e-12  # no cone blocks: Δd = Δμ d* − (H − Q) Δp = 0
    finally:
        s.close()


def _cases():
    yield "qp", G.qp(n=80, m=30, seed=52, pdens=0.05, adens=0.1)
    yield "socp", G.socp(n=120, m=50, ncone=6, seed=53, adens=0.05)
    yield "psd-lds", _trace_sdp(10, 54)[0]
    yield "psd-big", _trace_sdp(56, 55)[0]


CASES = list(_cases())
IDS = [c[0] for c in CASES]


def _adjoint(s, rng, n, m, dmu):
    df, dg = rng.standard_normal(n), rng.standard_normal(m)
    dp, dy, dd = s.frule(df, dg, dmu)
    pb, yb, db = rng.standard_normal(n), rng.standard_normal(m), rng.standard_normal(n)
    fb, gb, mub = s.rrule(pb, yb, db)
    lhs = pb @ dp + yb @ dy + db @ dd
    rhs = fb @ df + gb @ dg + mub * dmu
    scale = np.linalg.norm(np.concatenate([pb, yb, db])) * np.linalg.norm(np.concatenate([dp, dy, dd]))
    return lhs, rhs, scale


@pytest.mark.parametrize("name,d", CASES, ids=IDS)
def test_rrule_is_the_adjoint_of_frule(name, d):
    print("adjoint", name, flush=True)
    rng = np.random.default_rng(60)
    n, m = d["B"].shape[1], d["B"].shape[0]
    s, _ = _solved(d)
    try:
        lhs, rhs, scale = _adjoint(s, rng, n, m, 0.0)
        # each solve stops at a residual of 0.1·(μ/μ0)·(1 + ‖f‖) (the reference's forcing rule)
        # on a KKT matrix whose condition grows like 1/μ: the identity holds to ~1e-6 here
        assert abs(lhs - rhs) <= 1e-5 * scale, (name, "optimum", lhs, rhs)
    finally:
        s.close()
    s, _ = _solved(d, mu=MUT, **TIGHT)
    try:
        lhs, rhs, scale = _adjoint(s, rng, n, m, 0.3 * MUT)
        assert abs(lhs - rhs) <= 1e-8 * scale, (name, "central path", lhs, rhs)
    finally:
        s.close()


@pytest.mark.parametrize("name,d", CASES, ids=IDS)
def test_frule_is_the_derivative_of_the_central_path(name, d):
    """(x(f + εΔf, g + εΔg, μt + εΔμ) − x(f − εΔf, …)) / 2ε against frule!(Δf, Δg, Δμ) for
    x = p, y, d: the solves stop at a barrier gap of 1e-12; the O(ε²) term is ~1e-6 relative."""
    print("central path", name, flush=True)
    rng = np.random.default_rng(61)
    n, m = d["B"].shape[1], d["B"].shape[0]
    df, dg, dmu = rng.standard_normal(n) * 0.1, rng.standard_normal(m) * 0.1, 0.3 * MUT
    s, _ = _solved(d, mu=MUT, **TIGHT)
    try:
        dp, dy, dd = s.frule(df, dg, dmu)
    finally:
        s.close()
    eps = 1e-4
    sols = []
    for sign in (1.0, -1.0):
        e = dict(d, f=d["f"] + sign * eps * df, g=d["g"] + sign * eps * dg)
        r = mh.solve(hip_problem(e, mu=MUT + sign * eps * dmu), mh.IPMSettings(**TIGHT))
        assert r.status == mh.IPMStatus.OPTIMAL
        sols.append(r)
    fd = [(getattr(sols[0], k) - getattr(sols[1], k)) / (2 * eps) for k in "pyd"]
    errs = [_rel(fd[0], dp), _rel(fd[1], dy), _rel(fd[2], dd)]
    # d carries the centring error of the last iterate (pᵢdᵢ = μt to the barrier gap) most visibly
    assert errs[0] <= 1e-4 and errs[1] <= 1e-4 and errs[2] <= 1e-3, (name, errs)


@pytest.mark.parametrize("name,d", CASES[:3], ids=IDS[:3])
def test_frule_and_rrule_match_the_oracle(name, d):
    from oracle import ipm as O
    print("oracle", name, flush=True)
    rng = np.random.default_rng(62)
    n, m = d["B"].shape[1], d["B"].shape[0]
    df, dg, dmu = rng.standard_normal(n), rng.standard_normal(m), 0.3 * MUT
    pb, yb, db = rng.standard_normal(n), rng.standard_normal(m), rng.standard_normal(n)
    s, _ = _solved(d, mu=MUT, **TIGHT)
    try:
        fw = s.frule(df, dg, dmu)
        rv = s.rrule(pb, yb, db)
    finally:
        s.close()
    os_ = O.Solver(oracle_problem(d, mu=MUT), O.Settings(**TIGHT))
    assert int(os_.solve().status) == int(mh.IPMStatus.OPTIMAL)
    ofw = O.frule(os_, df, dg, dmu)
    orv = O.rrule(os_, pb, yb, db)
    for a, b in zip(fw, ofw):
        assert _rel(a, b) <= 1e-6, (name, _rel(a, b))
    for a, b in zip(rv[:2], orv[:2]):
        assert _rel(a, b) <= 1e-6, (name, _rel(a, b))
    assert abs(rv[2] - orv[2]) <= 1e-6 * max(1.0, abs(orv[2]))


def test_frule_matches_finite_differences_at_the_optimum_of_a_qp():
    """Nonnegative cones: at the optimum an active bound keeps Δp = 0 and an inactive one adds
    nothing, so frule! at the last iterate is the optimum's derivative."""
    name, d = CASES[0]
    rng = np.random.default_rng(63)
    n, m = d["B"].shape[1], d["B"].shape[0]
    df, dg = rng.standard_normal(n) * 0.1, rng.standard_normal(m) * 0.1
    s, _ = _solved(d, **TIGHT)
    try:
        dp, dy, _ = s.frule(df, dg, 0.0)
    finally:
        s.close()
    eps = 1e-4
    sols = [mh.solve(hip_problem(dict(d, f=d["f"] + sg * eps * df, g=d["g"] + sg * eps * dg)),
                     mh.IPMSettings(**TIGHT)) for sg in (1.0, -1.0)]
    assert all(r.status == mh.IPMStatus.OPTIMAL for r in sols)
    assert _rel((sols[0].p - sols[1].p) / (2 * eps), dp) <= 1e-5
    assert _rel((sols[0].y - sols[1].y) / (2 * eps), dy) <= 1e-5
