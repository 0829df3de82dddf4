"""The reference's control paths on the device, each against the oracle on the same input
(solver/ipm.jl:502-571, 278-315, 219-238; solver/solver.jl:9-30; problem.jl:247-349):
infeasibility certificates, the iteration limit, a Newton (CG) budget of zero, the ν = 0 exact
solve, the target-μ barrier, compressed columns, warm starts inside and outside the cone, and
whole solves at the full 2000-variable QP size in FP64 and in mixed precision.

Every comparison is status for status; objectives and iterates to 1e-8 relative (the certificate
paths end on iterates that diverge by construction, so there only the status and the iteration
at which the certificate fires are compared)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem, rel
from oracle import ipm as O

pytestmark = pytest.mark.gpu


def _pair(d, settings=None, osettings=None, mu=0.0, compress=1.0, **start):
    print("control case", d.get("name", ""), dict(mu=mu, compress=compress), flush=True)
    s = mh.init(hip_problem(d, mu=mu, compress=compress), settings, **start)
    res = s.solve()
    hist = s.history()
    s.close()
    ores = O.solve(oracle_problem(d, mu=mu, compress=compress), osettings, **start)
    return res, hist, ores


def _same(res, ores, niter=1, tol=1e-8):
    assert int(res.status) == int(ores.status), (res.status, ores.status)
    assert abs(res.niter - ores.niter) <= niter, (res.niter, ores.niter)
    assert rel(res.pobj, ores.pobj) < tol and rel(res.dobj, ores.dobj) < tol, (res.pobj, ores.pobj)


@pytest.mark.parametrize("n,m,seed,qd", [(6, 1, 0, 0.0), (20, 3, 1, 0.0), (50, 5, 2, 0.0), (50, 5, 3, 1.0),
                                         (200, 20, 4, 0.0), (200, 20, 5, 0.5), (80, 8, 6, 0.1)])
def test_primal_infeasible_certificate(n, m, seed, qd):
    """B > 0 entrywise and g < 0: no p ≥ 0 has Bp = g, and y ≤ 0 certifies it (solver/ipm.jl:
    502-517). The certificate fires at the same iteration (±1) as in the oracle. (An instance whose
    columns are all identical, with f constant, is left out: there the Uzawa window sets δ from
    residuals at the rounding floor, which differ between any two implementations, and the two
    paths part at noise level before the certificate forms.)"""
    rng = np.random.default_rng(seed)
    B = rng.uniform(0.5, 1.5, (m, n)) if m > 1 else np.ones((1, n))
    d = dict(Q=sp.csc_matrix(qd * np.eye(n)), B=sp.csc_matrix(B), f=-rng.uniform(0.5, 1.5, n),
             g=-rng.uniform(0.5, 1.5, m), K=[("pos",)], s=[n], name=f"primal infeasible {n}x{m}")
    res, _, ores = _pair(d)
    assert res.status in (mh.IPMStatus.PRIMAL_INFEASIBLE, mh.IPMStatus.NEAR_PRIMAL_INFEASIBLE)
    assert int(res.status) == int(ores.status) and abs(res.niter - ores.niter) <= 1


def test_dual_infeasible_certificate():
    n = 4
    d = dict(Q=sp.csc_matrix((n, n)), B=sp.csc_matrix(np.array([[1.0, -1.0, 0.0, 0.0]])),
             f=np.array([1.0, 1.0, 0.0, 0.0]), g=np.array([0.0]), K=[("pos",)], s=[n], name="dual infeasible")
    res, _, ores = _pair(d)
    assert res.status in (mh.IPMStatus.DUAL_INFEASIBLE, mh.IPMStatus.NEAR_DUAL_INFEASIBLE)
    assert int(res.status) == int(ores.status) and res.niter == ores.niter


def test_iteration_limit_and_history():
    """max_iter: ITERATION_LIMIT after the same number of iterations as the oracle. Mid-way the
    KKT solves are inexact by design (forcing tolerance 0.1·μ/μ₁, solver/ipm.jl:111-163) and δ
    comes from residual histories at the rounding floor, so the two trajectories agree there to
    about that tolerance (1e-5 relative), not to the 1e-8 of a finished solve."""
    d = G.qp(n=300, m=120, seed=31)
    res, hist, ores = _pair(d, mh.IPMSettings(max_iter=3), O.Settings(max_iter=3))
    assert res.status in (mh.IPMStatus.ITERATION_LIMIT, mh.IPMStatus.NEAR_OPTIMAL)
    _same(res, ores, niter=0, tol=1e-5)
    # the history rows of the iterations taken agree with the oracle's
    assert len(hist) == len(ores.history)
    for k, row in enumerate(hist):
        for key in ("pres", "dres", "mu"):
            ref = ores.history[k][key]
            assert abs(row[key] - ref) <= 1e-5 * max(1.0, abs(ref)), (k, key, row[key], ref)


def test_zero_newton_budget():
    """newton_max_iter = 0: the dual CG may take no step (cg.jl:69-167), so every KKT solve ends on
    the refinement alone; the device takes the same decisions as the oracle."""
    d = G.qp(n=200, m=80, seed=32)
    res, _, ores = _pair(d, mh.IPMSettings(newton_max_iter=0), O.Settings(newton_max_iter=0))
    assert int(res.status) == int(ores.status)
    assert abs(res.niter - ores.niter) <= 1
    if res.status == mh.IPMStatus.OPTIMAL:
        assert rel(res.pobj, ores.pobj) < 1e-8


def test_free_cones_only_take_the_exact_solve():
    """ν = 0 (free cones only): solveexact! (solver/ipm.jl:278-315); the equality-constrained QP
    optimum from its KKT system."""
    rng = np.random.default_rng(33)
    n, m = 40, 12
    M = rng.standard_normal((n, n))
    Q = M @ M.T / n + np.eye(n)
    A = rng.standard_normal((m, n))
    f = rng.standard_normal(n)
    g = rng.standard_normal(m)
    d = dict(Q=sp.csc_matrix(Q), B=sp.csc_matrix(A), f=f, g=g, K=[("free",)], s=[n], name="free only")
    res, _, ores = _pair(d)
    assert res.status == mh.IPMStatus.OPTIMAL and int(ores.status) == int(res.status)
    K = np.block([[Q, -A.T], [A, np.zeros((m, m))]])
    sol = np.linalg.solve(K, np.concatenate([f, g]))
    assert np.linalg.norm(res.p - sol[:n]) <= 1e-8 * np.linalg.norm(sol[:n])
    assert rel(res.pobj, ores.pobj) < 1e-10


def test_target_barrier():
    """μ > 0: the solve stops on the barrier problem's own optimality test (solver.jl:20-30) at
    the central-path point of μt, where pᵢdᵢ = μt on the nonnegative orthant."""
    d = G.qp(n=250, m=100, seed=34)
    mut = 1e-3
    res, _, ores = _pair(d, mu=mut)
    assert res.status == mh.IPMStatus.OPTIMAL
    _same(res, ores)
    assert np.linalg.norm(res.p - ores.p) <= 1e-6 * np.linalg.norm(ores.p)
    pd = res.p * res.d
    assert np.all(np.abs(pd - mut) <= 1e-2 * mut)


def test_compressed_columns():
    """compress = 0.6: columns of similar pattern merge into one vertex (twins with the cosine
    test, compress.jl:147-265), the block-sparse problem the device solves is the oracle's."""
    rng = np.random.default_rng(35)
    n, m = 120, 50
    base = sp.random(m, n // 4, density=0.15, random_state=rng, format="csc")
    cols = []
    for j in range(n):
        c = base[:, j // 4].toarray().ravel().copy()
        if j % 4 == 3:
            c[rng.integers(0, m)] = rng.standard_normal()  # near twin
        c[c != 0] = rng.standard_normal(np.count_nonzero(c))
        cols.append(c)
    A = sp.csc_matrix(np.array(cols).T)
    x0 = rng.uniform(0.2, 1.0, n)
    y0 = rng.standard_normal(m)
    d = dict(Q=sp.csc_matrix(sp.identity(n) * 0.1), B=A, f=-(A.T @ y0 + rng.uniform(0.1, 1.0, n)), g=A @ x0,
             K=[("pos",)], s=[n], name="compress 0.6")
    e = hip_problem(d, compress=0.6).export()
    assert e["B_xcol"].size - 1 < n  # some columns were merged
    res, _, ores = _pair(d, compress=0.6)
    assert res.status == mh.IPMStatus.OPTIMAL
    _same(res, ores)


def test_warm_start_inside_the_cone():
    d = G.qp(n=200, m=90, seed=36)
    ref = O.solve(oracle_problem(d))
    p0 = np.maximum(ref.p, 0.0) + 0.05
    res, _, ores = _pair(d, p0=p0, y0=ref.y)
    assert res.status == mh.IPMStatus.OPTIMAL
    _same(res, ores)


def test_warm_start_outside_the_cone():
    """A starting point with negative entries: scale! fails at the first iterate in the reference
    (the nonnegative block's shadow is not in the cone), and the device reports the same."""
    d = G.qp(n=150, m=60, seed=37)
    p0 = np.ones(150)
    p0[::7] = -0.5
    res, _, ores = _pair(d, p0=p0)
    assert int(res.status) == int(ores.status)
    assert res.niter == ores.niter


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_qp_matches_oracle(precision):
    """The 2000-variable QP at full size (n = 2000, m = 1000): FP64 and the mixed-precision form
    (FP32 large-front factor, FP64 refinement) both against the oracle's FP64 solve."""
    d = G.qp(n=2000, m=1000, seed=1000)
    s = mh.init(hip_problem(d), mh.IPMSettings(precision=precision))
    res = s.solve()
    st = s.stats()
    s.close()
    ores = O.solve(oracle_problem(d), order="identity")
    assert res.status == mh.IPMStatus.OPTIMAL and int(ores.status) == int(res.status)
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8
    assert np.linalg.norm(res.p - ores.p) <= 1e-6 * np.linalg.norm(ores.p)
    if precision == 1:
        assert st["factor_calls_fp32"] > 0


@pytest.mark.parametrize("precision", [0, 1])
def test_full_size_socp_matches_golden(precision):
    """The SOCP workload at full size (n = 10000, m = 5000, 500 second-order cones and a
    nonnegative block) against the oracle's solve, kept as a golden fixture because the oracle
    takes minutes on it (tests/golden/make_golden.py regenerates it from the seeded generator)."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "socp10000_seed1000.npz"))
    d = G.socp(seed=1000)
    s = mh.init(hip_problem(d), mh.IPMSettings(precision=precision))
    res = s.solve()
    st = s.stats()
    s.close()
    assert int(res.status) == int(gold["status"]) == int(mh.IPMStatus.OPTIMAL)
    # FP32 factors change the refinement passes and the δ windows: the mixed form may take a few
    # more iterations to the same 1e-8 optimum
    assert abs(res.niter - int(gold["niter"])) <= (1 if precision == 0 else 3)
    assert rel(res.pobj, float(gold["pobj"])) < 1e-8 and rel(res.dobj, float(gold["dobj"])) < 1e-8
    assert res.pres < 1e-8 and res.dres < 1e-8
    # the iterate: a solve stopped at gap 1e-8 sits within about √μ of the optimal set along its
    # flat directions; the FP64 path retraces the oracle's trajectory (1e-6), the mixed path takes
    # a different one to another point of the same 1e-8-optimal set (1e-5)
    tol = 1e-6 if precision == 0 else 1e-5
    assert np.linalg.norm(res.p - gold["p"]) <= tol * np.linalg.norm(gold["p"])
    if precision == 1:
        assert st["factor_calls_fp32"] > 0
