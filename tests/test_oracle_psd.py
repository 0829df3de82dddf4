"""Pins the oracle's PSD cone (sdp.jl; svec/skron: cone/utils.jl:475-664) by the NT invariant,
the Kronecker identity, closed forms and brute force, and the IPM on SDPs with known optima."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import cones as C
from oracle import ipm as O
from helpers import oracle_problem


def _spd(rng, d, shift=0.5):
    M = rng.standard_normal((d, d))
    return M @ M.T / d + shift * np.eye(d)


def _sym(rng, d):
    M = rng.standard_normal((d, d))
    return (M + M.T) / 2


@pytest.mark.parametrize("d", [1, 2, 5, 9])
def test_svec_is_an_isometry_and_smat_inverts_it(d):
    rng = np.random.default_rng(d)
    X, Y = _sym(rng, d), _sym(rng, d)
    np.testing.assert_allclose(C.svec(X) @ C.svec(Y), np.trace(X @ Y), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(C.smat(C.svec(X)), X, rtol=0, atol=1e-14)


@pytest.mark.parametrize("d", [2, 4, 7])
def test_skron_applies_the_congruence(d):
    # (A ⊗ₛ A) svec(X) = svec(A X Aᵀ) for symmetric A, X
    rng = np.random.default_rng(10 + d)
    A, X = _sym(rng, d), _sym(rng, d)
    np.testing.assert_allclose(C.skron(A) @ C.svec(X), C.svec(A @ X @ A.T), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d", [1, 3, 6, 10])
def test_psd_nt_scaling_maps_p_to_d(d):
    # the NT block H = W ⊗ₛ W maps svec(P) to svec(D) and is positive definite
    rng = np.random.default_rng(20 + d)
    P, D = _spd(rng, d), _spd(rng, d)
    n = d * (d + 1) // 2
    H = np.zeros((n, n))
    cache = C.PsdCache(d)
    ok, spsd = C.psdscale(H, cache, C.svec(P), C.svec(D))
    assert ok
    np.testing.assert_allclose(H @ C.svec(P), C.svec(D), rtol=1e-10, atol=1e-12)
    assert np.linalg.eigvalsh(H).min() > 0
    # ⟨p*, d*⟩ = tr(P⁻¹ D⁻¹)
    np.testing.assert_allclose(spsd, np.trace(np.linalg.inv(P) @ np.linalg.inv(D)), rtol=1e-10)
    # R Rᵀ = P⁻¹, Rᵀ S = I
    np.testing.assert_allclose(cache.R @ cache.R.T, np.linalg.inv(P), rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(cache.R.T @ cache.S, np.eye(d), rtol=0, atol=1e-11)


def test_psd_corr0_is_sigma_mu_p_inverse_minus_d():
    rng = np.random.default_rng(31)
    d = 6
    P, D = _spd(rng, d), _spd(rng, d)
    n = d * (d + 1) // 2
    cache = C.PsdCache(d)
    assert C.psdscale(np.zeros((n, n)), cache, C.svec(P), C.svec(D))[0]
    r = np.zeros(n)
    C.corr0(("psd",), r, C.svec(P), C.svec(D), 0.3, cache)
    np.testing.assert_allclose(r, C.svec(0.3 * np.linalg.inv(P) - D), rtol=1e-9, atol=1e-11)
    # with Δp = Δd = 0 the corrector reduces to corr0
    r2 = np.zeros(n)
    C.corr(("psd",), r2, C.svec(P), C.svec(D), np.zeros(n), np.zeros(n), 0.3, cache)
    np.testing.assert_allclose(r2, r, rtol=1e-12, atol=1e-13)


def test_psd_corrector_on_diagonal_matrices_is_the_orthant_corrector():
    # with P, D, ΔP, ΔD diagonal the PSD cone is the nonnegative orthant on the diagonal, so the
    # Mehrotra corrector must be (σμ − Δpᵢ Δdᵢ)/pᵢ − dᵢ there and zero off the diagonal
    # (pos.jl:90-120), whatever order the SVD returns the singular values in
    rng = np.random.default_rng(32)
    d = 6
    p, dd = rng.uniform(0.2, 3.0, d), rng.uniform(0.2, 3.0, d)
    dp, ddd = rng.standard_normal(d), rng.standard_normal(d)
    n = d * (d + 1) // 2
    cache = C.PsdCache(d)
    assert C.psdscale(np.zeros((n, n)), cache, C.svec(np.diag(p)), C.svec(np.diag(dd)))[0]
    r = np.zeros(n)
    C.corr(("psd",), r, C.svec(np.diag(p)), C.svec(np.diag(dd)), C.svec(np.diag(dp)), C.svec(np.diag(ddd)), 0.2, cache)
    np.testing.assert_allclose(C.smat(r), np.diag((0.2 - dp * ddd) / p - dd), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_psd_maxstep_against_bisection(seed):
    rng = np.random.default_rng(40 + seed)
    d = 5
    P = _spd(rng, d)
    dP = _sym(rng, d) * (3.0 if seed % 2 else 0.2)
    LP = np.linalg.cholesky(P)
    import scipy.linalg as sla
    F = sla.solve_triangular(LP.T, np.eye(d), lower=False)  # F Fᵀ = P⁻¹ (any such factor)
    tau = C.psdmaxstep(F, C.svec(dP), None)
    ispd = lambda t: np.linalg.eigvalsh(P + t * dP).min() > 0  # noqa: E731
    lo, hi = 0.0, 1.0
    if ispd(1.0):
        assert tau == 1.0
    else:
        for _ in range(60):
            mid = (lo + hi) / 2
            lo, hi = (mid, hi) if ispd(mid) else (lo, mid)
        assert abs(tau - lo) <= 1e-9


def test_psd_shadow_and_identity():
    rng = np.random.default_rng(50)
    d = 4
    X = _spd(rng, d)
    s1, s2 = np.zeros(10), np.zeros(10)
    assert C.shadow(("psd",), s1, C.svec(X))
    assert C.shadow(("psd",), s2, s1)
    np.testing.assert_allclose(s2, C.svec(X), rtol=1e-12)
    e = np.zeros(10)
    C.identity(e, ("psd",))
    s = np.zeros(10)
    assert C.shadow(("psd",), s, e)
    np.testing.assert_allclose(s, e, rtol=0, atol=1e-15)


def _trace_sdp(C0):
    """min tr(C X) s.t. tr(X) = 1, X ⪰ 0: optimum λmin(C)."""
    d = C0.shape[0]
    n = d * (d + 1) // 2
    B = sp.csc_matrix(C.svec(np.eye(d))[None, :])
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=-C.svec(C0), g=np.array([1.0]), K=[("psd",)], s=[n])


@pytest.mark.parametrize("d", [2, 5, 8])
def test_trace_constrained_sdp_attains_lambda_min(d):
    rng = np.random.default_rng(60 + d)
    C0 = _sym(rng, d)
    r = O.solve(oracle_problem(_trace_sdp(C0)))
    assert r.status == O.OPTIMAL
    lam = np.linalg.eigvalsh(C0)[0]
    assert abs(r.pobj - lam) <= 1e-7 * max(1.0, abs(lam))
    assert np.linalg.eigvalsh(C.smat(r.p)).min() > -1e-8
    assert np.linalg.eigvalsh(C.smat(r.d)).min() > -1e-8


@pytest.mark.parametrize("nv,band,clique", [(14, 3, 7), (22, 4, 9)])
def test_chordal_decomposition_attains_the_dense_sdp_optimum(nv, band, clique):
    # the windowed max-cut SDP (one PSD cone per window, consensus rows) attains the optimum of
    # the same SDP over one dense PSD block (chordal completion); band graphs, the same weights
    # through the same seed
    from mumblebee_hip import generators as G
    dw = G.maxcut(nv=nv, band=band, clique=clique, extra=0.0, seed=5)
    dd = G.maxcut(nv=nv, band=band, clique=nv, extra=0.0, seed=5)
    rw = O.solve(oracle_problem(dw))
    rd = O.solve(oracle_problem(dd))
    assert rw.status == rd.status == O.OPTIMAL
    assert abs(rw.pobj - rd.pobj) <= 1e-7 * max(1.0, abs(rd.pobj))
    assert rd.pobj < 0


def test_maxcut_sdp_bounds_every_cut_by_brute_force():
    # min ⟨C, X⟩ ≤ ⟨C, xxᵀ⟩ for every ±1 vector x (each cut is feasible), with equality
    # impossible to beat: compare against the best cut by enumeration on a tiny graph
    from mumblebee_hip import generators as G
    import itertools
    nv = 9
    d = G.maxcut(nv=nv, band=3, clique=nv, extra=0.0, seed=6)
    r = O.solve(oracle_problem(d))
    assert r.status == O.OPTIMAL
    Cm = C.smat(-d["f"])  # single window: f = −svec(C)
    best = min(float(np.array(x) @ Cm @ np.array(x)) for x in itertools.product([-1.0, 1.0], repeat=nv))
    assert r.pobj <= best + 1e-8
    assert r.pobj >= 1.1381 * best  # Goemans–Williamson: SDP ≥ best / 0.878… for C = −L/4
