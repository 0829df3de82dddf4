"""Pins of the exponential / power cone oracle (src/IPM/src/cone/tdc/*.jl) by what the
mathematics fixes: an independently written barrier value differentiated numerically, the
identity fixed point, the shadow maps inverting each other, the Tunçel secant equations, the
max step against bisection on membership, and whole solves against closed-form optima."""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import ipm as O
from oracle import tdc
from oracle.problem import CSC, make_problem

ALPHAS = (0.2, 0.5, 0.77)


def barrier(cone, x):
    """f(x), written from the cone definitions (exp.jl:37-40, pow.jl:40-45), plain double."""
    x1, x2, x3 = x
    if cone[0] == "exp":
        return -math.log(x2 * math.log(x1 / x2) - x3) - math.log(x1) - math.log(x2)
    al = cone[1]
    return -math.log(x1 ** (2 * al) * x2 ** (2 - 2 * al) - x3 ** 2) - (1 - al) * math.log(x1) - al * math.log(x2)


def member(cone, x):
    x1, x2, x3 = x
    if not (x1 > 0 and x2 > 0):
        return False
    if cone[0] == "exp":
        return x2 * math.log(x1 / x2) - x3 > 0
    al = cone[1]
    return x1 ** al * x2 ** (1 - al) - abs(x3) > 0


def interior(cone, rng, margin=(0.1, 1.0)):
    x1, x2 = rng.uniform(0.5, 2.0, 2)
    if cone[0] == "exp":
        return np.array([x1, x2, x2 * math.log(x1 / x2) - rng.uniform(*margin)])
    al = cone[1]
    return np.array([x1, x2, rng.uniform(-0.8, 0.8) * x1 ** al * x2 ** (1 - al)])


CONES = [("exp",)] + [("pow", a) for a in ALPHAS]


def state(cone, p):
    c = tdc.TdCache(cone[0])
    tdc.det(cone, p, c)
    return c


def test_two_word_log_and_exp2():
    getcontext().prec = 80
    rng = np.random.default_rng(0)
    assert tdc.twolograt(3.7, 3.7) == (0.0, 0.0)
    for k in (-40, -1, 0, 5, 300):
        assert tdc.twoexp2(float(k), 0.0) == (2.0 ** k, 0.0)
    for _ in range(50):
        a, b, c = np.exp(rng.uniform(-30, 30, 3))
        # log(a/b) + log(b/c) = log(a/c) to two-word accuracy
        x = Decimal(tdc.twolograt(a, b)[0]) + Decimal(tdc.twolograt(a, b)[1])
        y = Decimal(tdc.twolograt(b, c)[0]) + Decimal(tdc.twolograt(b, c)[1])
        z = Decimal(tdc.twolograt(a, c)[0]) + Decimal(tdc.twolograt(a, c)[1])
        assert abs(x + y - z) <= Decimal(1e-29) * max(Decimal(1), abs(z))
        # 2^(log2(a/b)) = a/b
        h, l = tdc.twolog2rat(a, b)
        e = tdc.twoexp2(h, l)
        assert abs((Decimal(e[0]) + Decimal(e[1])) / (Decimal(a) / Decimal(b)) - 1) < Decimal(1e-28)


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_identity_is_the_barrier_fixed_point(cone):
    """−f′(e) = e. The power cone's e = (√(1+α), √(2−α), 0) is exact; the exponential cone's
    constants (exp.jl:15-20) are the fixed point to about six digits, which Newton on
    −f′(x) = x confirms."""
    e = tdc.identity(cone)
    g = tdc.barrgrad(cone, e, state(cone, e))
    if cone[0] == "pow":
        assert np.allclose(-g, e, rtol=0, atol=1e-14)
        return
    assert np.abs(-g - e).max() < 3e-6
    x = e.copy()
    for _ in range(20):
        c = state(cone, x)
        r = -tdc.barrgrad(cone, x, c) - x
        J = -tdc.hessian(tdc.fact(cone, x, c)[1]) - np.eye(3)
        x = x - np.linalg.solve(J, r)
    assert np.abs(-tdc.barrgrad(cone, x, state(cone, x)) - x).max() < 1e-15
    assert np.abs(x - e).max() < 3e-6


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_gradient_hessian_third_match_finite_differences(cone):
    rng = np.random.default_rng(1)
    for _ in range(5):
        p = interior(cone, rng)
        c = state(cone, p)
        g = tdc.barrgrad(cone, p, c)
        h = 1e-6
        gfd = np.array([(barrier(cone, p + h * e) - barrier(cone, p - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.allclose(g, gfd, rtol=1e-6, atol=1e-7)
        ok, L = tdc.fact(cone, p, c)
        assert ok
        H = tdc.hessian(L)

        def grad_at(x):
            return tdc.barrgrad(cone, x, state(cone, x))
        Hfd = np.column_stack([(grad_at(p + h * e) - grad_at(p - h * e)) / (2 * h) for e in np.eye(3)])
        assert np.allclose(H, Hfd, rtol=1e-6, atol=1e-6 * np.abs(H).max())
        u = rng.standard_normal(3)
        D = tdc.third(cone, p, u, c)

        def hess_at(x):
            return tdc.hessian(tdc.fact(cone, x, state(cone, x))[1])
        Dfd = (hess_at(p + h * u) - hess_at(p - h * u)) / (2 * h)
        assert np.allclose(D, Dfd, rtol=1e-5, atol=1e-5 * np.abs(D).max())
        # the pivoted factor solves what it multiplies
        b = rng.standard_normal(3)
        assert np.allclose(tdc.hessldiv(L, tdc.hessmul(L, b)), b, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_shadow_maps_invert_each_other(cone):
    rng = np.random.default_rng(2)
    for _ in range(10):
        x = interior(cone, rng, margin=(1e-3, 2.0))
        d = np.zeros(3)
        tdc.dual_shadow(cone, d, x)
        for seed in (0.0, 1.0, 0.3, 7.0):
            sp, _ = tdc.dualgrad(cone, seed, d, tdc.TdCache(cone[0]))
            assert np.allclose(sp, x, rtol=1e-9, atol=1e-12), (seed, sp, x)


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_tuncel_block_satisfies_secant_equations(cone):
    rng = np.random.default_rng(3)
    for _ in range(10):
        p, q = interior(cone, rng), interior(cone, rng)
        d = np.zeros(3)
        tdc.dual_shadow(cone, d, q)
        c = tdc.TdCache(cone[0])
        H = np.zeros((3, 3))
        ok, spsd = tdc.scale(cone, H, p, d, c)
        assert ok
        sp, _ = tdc.dualgrad(cone, 1.0, d, tdc.TdCache(cone[0]))
        assert np.allclose(H, H.T, rtol=0, atol=1e-12 * np.abs(H).max())
        assert np.linalg.eigvalsh(H).min() > 0
        assert np.allclose(H @ p, d, rtol=1e-12, atol=1e-13 * np.linalg.norm(d))
        assert np.allclose(H @ sp, c.sd, rtol=1e-11, atol=1e-12 * np.linalg.norm(c.sd))
        assert spsd == pytest.approx(float(sp @ c.sd), rel=1e-13)


def _bisect_boundary(cone, x, dx, hi):
    if member(cone, x + hi * dx):
        return hi
    lo = 0.0
    for _ in range(200):
        m = 0.5 * (lo + hi)
        if member(cone, x + m * dx):
            lo = m
        else:
            hi = m
    return lo


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_max_step_matches_bisection_on_membership(cone):
    rng = np.random.default_rng(4)
    for _ in range(20):
        p, q = interior(cone, rng), interior(cone, rng)
        d = np.zeros(3)
        tdc.dual_shadow(cone, d, q)
        c = tdc.TdCache(cone[0])
        ok, _ = tdc.scale(cone, np.zeros((3, 3)), p, d, c)
        assert ok
        dp = rng.standard_normal(3) * 2.0
        tp, _ = tdc.maxsteps(cone, p, dp, d, np.zeros(3), c)
        ref = _bisect_boundary(cone, p, dp, 1.0)
        # τ may land on the closure (an affine clamp where x₁ or x₂ reaches 0): just short of
        # it the point is interior
        assert member(cone, p + (1 - 1e-9) * tp * dp) or tp == 0.0
        assert tp == pytest.approx(ref, rel=1e-7, abs=1e-12)


def _dual_member(cone, z):
    z1, z2, z3 = z
    if cone[0] == "exp":
        return z1 > 0 and z3 < 0 and 1 + math.log(z1) - z2 / z3 - math.log(-z3) > 0
    al = cone[1]
    return z1 > 0 and z2 > 0 and (z1 / al) ** al * (z2 / (1 - al)) ** (1 - al) - abs(z3) > 0


@pytest.mark.parametrize("cone", CONES, ids=str)
def test_dual_max_step_matches_bisection(cone):
    rng = np.random.default_rng(5)
    for _ in range(20):
        p, q = interior(cone, rng), interior(cone, rng)
        d = np.zeros(3)
        tdc.dual_shadow(cone, d, q)
        c = tdc.TdCache(cone[0])
        assert tdc.scale(cone, np.zeros((3, 3)), p, d, c)[0]
        dd = rng.standard_normal(3) * 2.0
        _, td = tdc.maxsteps(cone, p, np.zeros(3), d, dd, c)
        hi = 1.0
        if _dual_member(cone, d + dd):
            ref = 1.0
        else:
            lo = 0.0
            for _ in range(200):
                m = 0.5 * (lo + hi)
                if _dual_member(cone, d + m * dd):
                    lo = m
                else:
                    hi = m
            ref = lo
        assert td == pytest.approx(ref, rel=1e-7, abs=1e-12)


# ---------------------------------------------------------------------------------------
# whole solves with closed-form optima
# ---------------------------------------------------------------------------------------

def entropy_problem(k):
    """max Σ −x_i log x_i s.t. Σ x_i = 1, as min −Σ t_i over (u_i, x_i, t_i) ∈ K_exp, u_i = 1.
    Optimum −log k at x_i = 1/k."""
    n = 3 * k
    rows, cols, vals = [], [], []
    for i in range(k):
        rows.append(i), cols.append(3 * i), vals.append(1.0)       # u_i = 1
        rows.append(k), cols.append(3 * i + 1), vals.append(1.0)   # Σ x_i = 1
    B = sp.csc_matrix((vals, (rows, cols)), shape=(k + 1, n))
    f = np.zeros(n)
    f[2::3] = 1.0  # the objective is ½pᵀQp − fᵀp
    g = np.ones(k + 1)
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=f, g=g, K=[("exp",)] * k, s=[3] * k), -math.log(k)


def geomean_problem(alphas):
    """max Σ x3_i s.t. (x1_i, x2_i, x3_i) ∈ K_pow(α_i), x1_i + x2_i = 2: optimum
    Σ (2α_i)^α_i (2(1−α_i))^(1−α_i)."""
    k = len(alphas)
    n = 3 * k
    rows, cols, vals = [], [], []
    for i in range(k):
        rows += [i, i]
        cols += [3 * i, 3 * i + 1]
        vals += [1.0, 1.0]
    B = sp.csc_matrix((vals, (rows, cols)), shape=(k, n))
    f = np.zeros(n)
    f[2::3] = 1.0  # the objective is ½pᵀQp − fᵀp
    opt = -sum((2 * a) ** a * (2 * (1 - a)) ** (1 - a) for a in alphas)
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=f, g=2.0 * np.ones(k), K=[("pow", a) for a in alphas],
                s=[3] * k), opt


def _solve(data):
    prob = make_problem(CSC.from_scipy(data["Q"]), CSC.from_scipy(data["B"]), data["f"], data["g"], 0.0,
                        data["K"], data["s"])
    return O.Solver(prob).solve()


@pytest.mark.parametrize("k", [1, 3, 10])
def test_entropy_on_the_simplex(k):
    data, opt = entropy_problem(k)
    r = _solve(data)
    assert r.status == O.OPTIMAL
    assert r.pobj == pytest.approx(opt, rel=1e-7, abs=1e-8)
    # the optimum is flat to second order: x is determined to about the square root of the gap
    assert np.allclose(r.p[1::3], 1.0 / k, rtol=1e-4)


def test_weighted_geometric_means_under_budgets():
    data, opt = geomean_problem([0.2, 0.5, 0.77, 0.35])
    r = _solve(data)
    assert r.status == O.OPTIMAL
    assert r.pobj == pytest.approx(opt, rel=1e-7)
