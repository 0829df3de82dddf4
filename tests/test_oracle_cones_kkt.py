"""Pins the oracle's cone kernels and KKT solver by invariants, closed forms and brute force."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import cones as C
from oracle.kkt import KKT_SOLVED, UzawaSolver, solvekkt
from oracle.bsm import BSM


def soc_point(rng, n, margin=0.5):
    x = rng.standard_normal(n)
    x[0] = np.linalg.norm(x[1:]) + margin + rng.random()
    return x


def in_soc(x, tol=0.0):
    return x[0] >= np.linalg.norm(x[1:]) - tol


@pytest.mark.parametrize("seed", range(5))
def test_twoprod_twosum_are_exact(seed):
    rng = np.random.default_rng(seed)
    for _ in range(200):
        a, b = rng.standard_normal(2) * 10.0 ** rng.integers(-8, 8, 2)
        p, e = C.twoprod(float(a), float(b))
        assert Fraction(p) + Fraction(e) == Fraction(float(a)) * Fraction(float(b))
        s, e = C.twosum(float(a), float(b))
        assert Fraction(s) + Fraction(e) == Fraction(float(a)) + Fraction(float(b))


def test_identities_are_barrier_fixed_points():
    # −f′(e) = e: pos f = −Σ log x → −f′ = 1/x; soc f = −log det x → −f′ = 2Jx/det x
    for cone, n in ((("pos",), 5), (("soc",), 6)):
        e = np.zeros(n)
        C.identity(e, cone)
        s = np.zeros(n)
        assert C.shadow(cone, s, e)
        np.testing.assert_allclose(s, e, rtol=1e-15)


@pytest.mark.parametrize("seed", range(6))
def test_soc_nt_scaling_maps_p_to_d(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 12))
    p, d = soc_point(rng, n), soc_point(rng, n)
    H = np.zeros((n, n))
    cache = C.SocCache(n)
    ok, spsd = C.socscale(H, cache, p, d)
    assert ok
    np.testing.assert_allclose(H @ p, d, rtol=1e-10, atol=1e-10)     # NT: H p = d
    assert np.all(np.linalg.eigvalsh(H) > 0)
    # H = f″(w) at the NT point: f″(w) = (2 (Jw)(Jw)ᵀ/det² − J/det)·2 … equivalently H⁻¹ d = p
    np.testing.assert_allclose(np.linalg.solve(H, d), p, rtol=1e-10)
    # ⟨p*, d*⟩ with p* = −f′(p), d* = −f′*(d)
    ps, ds = np.zeros(n), np.zeros(n)
    C.shadow(("soc",), ps, p)
    C.shadow(("soc",), ds, d)
    assert abs(spsd - ps @ ds) < 1e-10 * abs(ps @ ds)
    # √H applied twice is H
    x = rng.standard_normal(n)
    y = C.socroot(C.socroot(x.copy(), cache.w, cache.beta, False), cache.w, cache.beta, False)
    np.testing.assert_allclose(y, H @ x, rtol=1e-10, atol=1e-10)
    z = C.socroot(C.socroot(x.copy(), cache.w, cache.beta, True), cache.w, cache.beta, False)
    np.testing.assert_allclose(z, x, rtol=1e-10, atol=1e-10)


def test_pos_scaling_maps_p_to_d():
    rng = np.random.default_rng(1)
    p, d = rng.random(7) + 0.1, rng.random(7) + 0.1
    H = np.zeros((7, 7))
    ok, spsd = C.scale(("pos",), H, p, d, None)
    np.testing.assert_allclose(H @ p, d)
    assert abs(spsd - np.sum(1 / (p * d))) < 1e-12


def _jordan(x, y):
    return np.concatenate([[x @ y], x[0] * y[1:] + y[0] * x[1:]])


@pytest.mark.parametrize("seed", range(6))
def test_soc_corrector_closed_form(seed):
    """r = −d + (σμ e − Δp̃∘Δd̃)/p̃ in the NT-scaled space (soc.jl:282-351), checked by solving the
    Jordan system with a dense linear solve."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 9))
    p, d = soc_point(rng, n), soc_point(rng, n)
    H = np.zeros((n, n))
    cache = C.SocCache(n)
    C.socscale(H, cache, p, d)
    dp, dd = rng.standard_normal(n), rng.standard_normal(n)
    smu = 0.3
    r = np.zeros(n)
    C.corr(("soc",), r, p, d, dp, dd, smu, cache)
    # reference: W = √H; v = W p (= W⁻¹ d); solve v ∘ x = (W dp)∘(W⁻¹ dd); r = −W (x + v) + γ J p
    W = np.zeros((n, n))
    for k in range(n):
        W[:, k] = C.socroot(np.eye(n)[k].copy(), cache.w, cache.beta, False)
    Winv = np.linalg.inv(W)
    v = W @ p
    rhs = _jordan(W @ dp, Winv @ dd)
    Lv = np.zeros((n, n))
    for k in range(n):
        Lv[:, k] = _jordan(v, np.eye(n)[k])
    x = np.linalg.solve(Lv, rhs)
    detp = p[0] ** 2 - p[1:] @ p[1:]
    Jp = p.copy()
    Jp[1:] *= -1
    ref = -(W @ (x + v)) + (2 * smu / detp) * Jp
    np.testing.assert_allclose(r, ref, rtol=1e-9, atol=1e-9)
    # with Δp = Δd = 0 the corrector equals corr0
    r0 = np.zeros(n)
    C.corr0(("soc",), r0, p, d, smu, cache)
    r00 = np.zeros(n)
    C.corr(("soc",), r00, p, d, np.zeros(n), np.zeros(n), smu, cache)
    np.testing.assert_allclose(r0, r00, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("seed", range(10))
def test_maxstep_against_bisection(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 10))
    x = soc_point(rng, n, margin=0.1)
    dx = rng.standard_normal(n) * 3
    c, clo = C.socdet(x)
    tau = C.socmaxstep(x, dx, c, clo)
    lo, hi = 0.0, 1.0
    if in_soc(x + dx):
        assert tau == 1.0 or in_soc(x + tau * dx, 1e-12)
    else:
        for _ in range(80):
            mid = (lo + hi) / 2
            (lo, hi) = (mid, hi) if in_soc(x + mid * dx) else (lo, mid)
        assert abs(tau - lo) < 1e-9
    xp = rng.random(n) + 0.1
    dxp = rng.standard_normal(n)
    t = C.posmaxstep(xp, dxp)
    assert np.all(xp + t * dxp >= -1e-14)
    assert t == 1.0 or np.min(xp + t * dxp) < 1e-12


def _dense_bsm(A):
    m, n = A.shape
    val = A.reshape(-1, order="F")
    return BSM(1, 1, [0, 1], [0, n], [0, m], [0, m * n], [0], val)


@pytest.mark.parametrize("seed", range(4))
def test_solvekkt_meets_tolerances_and_matches_direct_solve(seed):
    rng = np.random.default_rng(seed)
    n, m = 30, 12
    Bd = rng.standard_normal((m, n))
    G0 = rng.standard_normal((n, n))
    H = G0 @ G0.T / n + np.eye(n)
    solver = UzawaSolver(_dense_bsm(Bd))
    ok, rho = solver.init(H, 0.05)
    assert ok and rho == 0.0
    f, g = rng.standard_normal(n), rng.standard_normal(m)
    x, y = np.zeros(n), np.zeros(m)
    ftol = gtol = 1e-10
    ncg, nir, st, dmin, dmax = solvekkt(solver, H, Bd, x, y, f, g, False, ftol, gtol, 0.5, 10, 100, 1)
    assert st == KKT_SOLVED
    assert np.linalg.norm(H @ x - Bd.T @ y - f) <= ftol
    assert np.linalg.norm(Bd @ x - g) <= gtol
    K = np.block([[H, -Bd.T], [Bd, np.zeros((m, m))]])
    sol = np.linalg.solve(K, np.concatenate([f, g]))
    np.testing.assert_allclose(np.concatenate([x, y]), sol, rtol=1e-7, atol=1e-8)


def test_factor_failure_takes_the_rho_retry():
    # δH + BᵀB indefinite → first Cholesky fails, shifted one may too: flag and ρ are reported
    n = 4
    Bd = np.zeros((1, n))
    H = -np.eye(n)
    solver = UzawaSolver(_dense_bsm(Bd))
    ok, rho = solver.init(H, 1.0)
    assert not ok and rho > 0
    H2 = np.diag([1.0, 1.0, 1.0, 1e-300]) - np.diag([0, 0, 0, 2e-300])
    ok2, rho2 = solver.init(H2, 1.0)
    assert rho2 > 0 and ok2


@pytest.mark.parametrize("seed", range(4))
def test_socroot_squared_is_the_nt_scaling(seed):
    # √H applied twice equals the scaling block H that socscale adds (soc.jl:82-110 vs
    # soc.jl:192-263), and the inverse form undoes the forward one
    rng = np.random.default_rng(seed)
    n = 7
    p, d = soc_point(rng, n), soc_point(rng, n)
    cache = C.SocCache(n)
    H = np.zeros((n, n))
    flag, _ = C.socscale(H, cache, p, d)
    assert flag
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        x = C.socroot(C.socroot(e.copy(), cache.w, cache.beta, False), cache.w, cache.beta, False)
        np.testing.assert_allclose(x, H[:, i], rtol=1e-12, atol=1e-12)
    x = rng.standard_normal(n)
    y = C.socroot(C.socroot(x.copy(), cache.w, cache.beta, False), cache.w, cache.beta, True)
    np.testing.assert_allclose(y, x, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("cone", [("pos",), ("soc",)])
def test_shadow_is_an_involution(cone):
    # −f′ maps the interior onto itself and −f′(−f′(x)) = x for −Σ log xᵢ and −log det x
    rng = np.random.default_rng(7)
    n = 6
    x = rng.random(n) + 0.1 if cone[0] == "pos" else soc_point(rng, n)
    s, t = np.zeros(n), np.zeros(n)
    assert C.shadow(cone, s, x)
    assert C.shadow(cone, t, s)
    np.testing.assert_allclose(t, x, rtol=1e-13)


def _kkt_instance(seed, n=30, m=12):
    rng = np.random.default_rng(seed)
    Bd = rng.standard_normal((m, n))
    G0 = rng.standard_normal((n, n))
    H = G0 @ G0.T / n + np.eye(n)
    solver = UzawaSolver(_dense_bsm(Bd))
    ok, rho = solver.init(H, 0.05)
    assert ok and rho == 0.0
    return solver, H, Bd, rng.standard_normal(n), rng.standard_normal(m)


def test_solvekkt_irmax_zero_skips_refinement():
    # nir ≥ irmax before the first pass: no refinement, the dual CG still runs, and the
    # status is ITMAX (uzawa.jl:384-390, 480-488)
    from oracle.kkt import KKT_ITMAX
    solver, H, Bd, f, g = _kkt_instance(1)
    x, y = np.zeros(len(f)), np.zeros(len(g))
    ncg, nir, st, _, _ = solvekkt(solver, H, Bd, x, y, f, g, False, 1e-12, 1e-12, 0.5, 0, 100)
    assert nir == 0 and ncg > 0 and st == KKT_ITMAX


def test_solvekkt_cgmax_zero_reports_itmax():
    from oracle.kkt import KKT_ITMAX
    solver, H, Bd, f, g = _kkt_instance(2)
    x, y = np.zeros(len(f)), np.zeros(len(g))
    ncg, nir, st, _, _ = solvekkt(solver, H, Bd, x, y, f, g, False, 1e-12, 1e-12, 0.5, 10, 0, 1)
    assert ncg == 0 and nir >= 1 and st == KKT_ITMAX


def test_solvekkt_stall_stops_the_second_pass():
    # stall = 1 demands fres ≤ 0·fprv, so the second pass is declared stagnated unless solved
    from oracle.kkt import KKT_STAGNATED
    solver, H, Bd, f, g = _kkt_instance(3)
    x, y = np.zeros(len(f)), np.zeros(len(g))
    ncg, nir, st, _, _ = solvekkt(solver, H, Bd, x, y, f, g, False, 0.0, 0.0, 1.0, 10, 100)
    assert nir == 2 and st == KKT_STAGNATED
