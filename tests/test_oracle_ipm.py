"""Pins the oracle's IPM driver by solver-independent certificates, an exact LP solver
(HiGHS via scipy.optimize.linprog) and forced failure paths."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import linprog

from oracle import ipm as O
from oracle.problem import CSC, make_problem
from mumblebee_hip import generators as G
from helpers import kkt_certificate, oracle_problem


def _lp(seed, n=60, m=25):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=0.2, random_state=rng, data_rvs=rng.standard_normal, format="csc")
    x0 = rng.uniform(0.1, 1.0, n)
    y0 = rng.standard_normal(m)
    c = A.T @ y0 + rng.uniform(0.1, 1.0, n)
    return dict(Q=sp.csc_matrix((n, n)), B=A, f=-c, g=A @ x0, K=[("pos",)], s=[n])


@pytest.mark.parametrize("seed", range(4))
def test_lp_objective_matches_highs(seed):
    d = _lp(seed)
    r = O.solve(oracle_problem(d))
    assert r.status == O.OPTIMAL
    ref = linprog(-d["f"], A_eq=d["B"].toarray(), b_eq=d["g"], bounds=(0, None), method="highs")
    assert ref.status == 0
    # min ½pᵀQp − fᵀp with Q = 0 is min cᵀp, c = −f
    assert abs(r.pobj - ref.fun) <= 1e-7 * max(1.0, abs(ref.fun))


@pytest.mark.parametrize("gen,kw", [(G.qp, dict(n=150, m=70, seed=21)),
                                    (G.socp, dict(n=300, m=140, ncone=12, seed=22, adens=0.03))])
def test_certificates(gen, kw):
    d = gen(**kw)
    r = O.solve(oracle_problem(d))
    assert r.status == O.OPTIMAL
    rp, rd = kkt_certificate(d, r.p, r.d, r.y)
    assert rp < 1e-7 and rd < 1e-7
    # cone membership of p and d
    j = 0
    for k, n in zip(d["K"], d["s"]):
        pv, dv = r.p[j:j + n], r.d[j:j + n]
        if k[0] == "pos":
            assert pv.min() > -1e-9 and dv.min() > -1e-9
        elif k[0] == "soc":
            assert pv[0] - np.linalg.norm(pv[1:]) > -1e-8 and dv[0] - np.linalg.norm(dv[1:]) > -1e-8
        j += n
    # objective = dual objective (zero gap)
    assert abs(r.pobj - r.dobj) <= 1e-7 * max(1.0, abs(r.pobj))
    # complementarity pᵀd ≈ 0 relative to the objective scale
    assert abs(r.p @ r.d) <= 1e-6 * max(1.0, abs(r.pobj))


def test_equality_qp_closed_form():
    """A QP whose bounds are inactive at the optimum equals the closed-form KKT solution."""
    rng = np.random.default_rng(7)
    n, m = 20, 6
    M = rng.standard_normal((n, n))
    Q = M @ M.T + n * np.eye(n)
    A = rng.standard_normal((m, n))
    xs = rng.uniform(1.0, 2.0, n)
    ys = rng.standard_normal(m)
    # choose f so that (xs, ys) satisfies Q x − f − Aᵀy = 0 with x > 0 (d = 0)
    f = Q @ xs - A.T @ ys
    d = dict(Q=sp.csc_matrix(Q), B=sp.csc_matrix(A), f=f, g=A @ xs, K=[("pos",)], s=[n])
    r = O.solve(oracle_problem(d))
    assert r.status == O.OPTIMAL
    np.testing.assert_allclose(r.p, xs, rtol=1e-6)
    np.testing.assert_allclose(r.y, ys, rtol=1e-5, atol=1e-6)


def test_iteration_limit():
    d = G.qp(n=80, m=30, seed=3)
    r = O.solve(oracle_problem(d), O.Settings(max_iter=1))
    assert r.status in (O.ITERATION_LIMIT, O.NEAR_OPTIMAL)
    assert r.niter == 2


def test_primal_infeasible():
    n = 6
    A = np.ones((1, n))
    d = dict(Q=sp.csc_matrix((n, n)), B=sp.csc_matrix(A), f=-np.ones(n), g=np.array([-1.0]), K=[("pos",)], s=[n])
    r = O.solve(oracle_problem(d))
    assert r.status in (O.PRIMAL_INFEASIBLE, O.NEAR_PRIMAL_INFEASIBLE)


def test_dual_infeasible():
    n = 4
    A = np.array([[1.0, -1.0, 0.0, 0.0]])
    d = dict(Q=sp.csc_matrix((n, n)), B=sp.csc_matrix(A), f=np.array([1.0, 1.0, 0.0, 0.0]) * 1.0,
             g=np.array([0.0]), K=[("pos",)], s=[n])
    r = O.solve(oracle_problem(d))
    assert r.status in (O.DUAL_INFEASIBLE, O.NEAR_DUAL_INFEASIBLE)



def test_frule_of_an_equality_qp_is_its_kkt_solve():
    """ν = 0 (free cones): the solution is linear in (f, g), so frule! is the KKT solve of the
    seeds (differential.jl:5-75)."""
    rng = np.random.default_rng(51)
    n, m = 20, 6
    M = rng.standard_normal((n, n))
    Q = M @ M.T / n + np.eye(n)
    A = rng.standard_normal((m, n))
    d = dict(Q=sp.csc_matrix(Q), B=sp.csc_matrix(A), f=rng.standard_normal(n), g=rng.standard_normal(m),
             K=[("free",)], s=[n])
    s = O.Solver(oracle_problem(d))
    assert s.solve().status == O.OPTIMAL
    df, dg = rng.standard_normal(n), rng.standard_normal(m)
    dp, dy, dd = O.frule(s, df, dg, 0.0)
    ref = np.linalg.solve(np.block([[Q, -A.T], [A, np.zeros((m, m))]]), np.concatenate([df, dg]))
    np.testing.assert_allclose(dp, ref[:n], rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(dy, ref[n:], rtol=1e-8, atol=1e-10)
    assert np.abs(dd).max() < 1e-12


def _adjoint_gap(s, rng, n, m, dmu):
    df, dg = rng.standard_normal(n), rng.standard_normal(m)
    dp, dy, dd = O.frule(s, df, dg, dmu)
    pb, yb, db = rng.standard_normal(n), rng.standard_normal(m), rng.standard_normal(n)
    fb, gb, mub = O.rrule(s, pb, yb, db)
    lhs = pb @ dp + yb @ dy + db @ dd
    rhs = fb @ df + gb @ dg + mub * dmu
    scale = np.linalg.norm(np.concatenate([pb, yb, db])) * np.linalg.norm(np.concatenate([dp, dy, dd]))
    return abs(lhs - rhs) / scale


DIFF_CASES = [(G.qp, dict(n=60, m=20, seed=52, pdens=0.05, adens=0.1)),
              (G.socp, dict(n=90, m=40, ncone=5, seed=53, adens=0.06)),
              (G.maxcut_chordal, dict(nv=30, band=4, extra=0.0, seed=54))]


@pytest.mark.parametrize("gen,kw", DIFF_CASES, ids=["qp", "socp", "psd"])
def test_frule_is_the_derivative_of_the_central_path_and_rrule_its_adjoint(gen, kw):
    """On the central path (target barrier μt = 1e-2, barrier gap 1e-12) the NT scaling is
    μt ∇²F(p), so frule!(Δf, Δg, Δμ) is the derivative of (f, g, μt) ↦ (p, y, d): central
    finite differences of whole solves pin it for every cone kind, Δμ included. rrule! is pinned
    by the adjoint identity ⟨p̄, Δp⟩ + ⟨ȳ, Δy⟩ + ⟨d̄, Δd⟩ = ⟨f̄, Δf⟩ + ⟨ḡ, Δg⟩ + μ̄ Δμ, at the
    optimum (Δμ = 0) and on the central path."""
    d = gen(**kw)
    rng = np.random.default_rng(60)
    n, m = d["B"].shape[1], d["B"].shape[0]
    st = O.Settings(gap_tol=1e-12)
    s = O.Solver(oracle_problem(d))
    assert s.solve().status == O.OPTIMAL
    assert _adjoint_gap(s, rng, n, m, 0.0) <= 1e-5  # forcing tolerance × KKT condition ~1/μ
    mut = 1e-2
    s = O.Solver(oracle_problem(d, mu=mut), st)
    assert s.solve().status == O.OPTIMAL
    assert _adjoint_gap(s, rng, n, m, 0.3 * mut) <= 1e-8
    df, dg, dmu = rng.standard_normal(n) * 0.1, rng.standard_normal(m) * 0.1, 0.3 * mut
    dx = O.frule(s, df, dg, dmu)
    eps = 1e-4  # the O(ε²) term is ~1e-4 relative at ε = 1e-3 on the chordal max-cut
    r = [O.solve(oracle_problem(dict(d, f=d["f"] + sg * eps * df, g=d["g"] + sg * eps * dg), mu=mut + sg * eps * dmu),
                 st) for sg in (1, -1)]
    assert all(x.status == O.OPTIMAL for x in r)
    # d carries the centring error of the last iterate (pᵢdᵢ = μt to the barrier gap) most visibly
    for k, dk, tol in zip("pyd", dx, (1e-4, 1e-4, 1e-3)):
        fd = (getattr(r[0], k) - getattr(r[1], k)) / (2 * eps)
        assert np.linalg.norm(fd - dk) <= tol * np.linalg.norm(dk), k


def test_a_dropped_dual_shadow_term_fails_the_central_path_check():
    """The check above is sharp: frule! without the Δμ d* term misses the μ-derivative."""
    d = G.qp(n=60, m=20, seed=52, pdens=0.05, adens=0.1)
    mut, eps = 1e-2, 1e-3
    st = O.Settings(gap_tol=1e-12)
    s = O.Solver(oracle_problem(d, mu=mut), st)
    assert s.solve().status == O.OPTIMAL
    n, m = d["B"].shape[1], d["B"].shape[0]
    dp, _, _ = O.frule(s, np.zeros(n), np.zeros(m), 0.0)  # the planted mistake: Δμ not seeded
    r = [O.solve(oracle_problem(d, mu=mut + sg * eps * mut), st) for sg in (1, -1)]
    fd = (r[0].p - r[1].p) / (2 * eps)
    assert np.linalg.norm(fd - dp) > 0.1 * np.linalg.norm(fd)
