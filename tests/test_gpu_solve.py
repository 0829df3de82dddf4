"""GPU parity tests: the HIP path against the oracle on the same seeded inputs."""
import numpy as np
import pytest

import mumblebee_hip as mh
from helpers import hip_problem, oracle_problem, rel, small_cases, kkt_certificate

pytestmark = pytest.mark.gpu


def test_mfma_f64_lane_layout():
    """v_mfma_f64_16x16x4f64: A[i][k] from lane i+16k, B[k][j] from lane j+16k, D[i][j] in
    lane j+16(i%4), element i//4 — the layout the trailing-update kernel assumes."""
    out = mh.probe_mfma_f64().reshape(64, 4)
    A = np.array([[i + 100 * k for k in range(4)] for i in range(16)], dtype=float)
    B = np.array([[j + 1000 * k for j in range(16)] for k in range(4)], dtype=float)
    D = A @ B
    exp = np.zeros((64, 4))
    for lane in range(64):
        for t in range(4):
            exp[lane, t] = D[(lane // 16) + 4 * t, lane % 16]
    if not np.array_equal(out, exp):
        print("actual lane outputs:\n", out)
    assert np.array_equal(out, exp)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_factor_matches_dense_cholesky(case):
    """initkkt! on H = Q: the supernodal factor equals LAPACK's dense Cholesky of δQ + BᵀB."""
    from oracle import ipm as O
    d = small_cases()[case]
    solver = mh.init(hip_problem(d))
    osolver = O.Solver(oracle_problem(d))
    delta = 0.37
    ok, rho = solver.kkt_factor(delta)
    M = osolver.kkt.L + delta * osolver.Qd
    # the ρ retry (uzawa.jl:97-112) is taken exactly when the unshifted factorization fails;
    # with Q = 0 (SOCP cases) BᵀB is singular, so it must be
    ook, orho = osolver.kkt.init(osolver.Qd, delta)
    assert ok and ook
    assert (rho == 0.0) == (orho == 0.0)
    if rho > 0:
        assert abs(rho - orho) <= 1e-12 * orho
        M = M + rho * np.eye(M.shape[0])
    _check_factor(np.tril(solver.dense_factor()), M)


def _check_factor(L, M):
    """Backward error ‖LLᵀ − M‖/‖M‖ at rounding level; the forward error against LAPACK only
    where M is well conditioned (the ρ-shifted BᵀB of the SOCP cases is not)."""
    n = M.shape[0]
    back = np.linalg.norm(L @ L.T - M) / np.linalg.norm(M)
    assert back < 10 * n * 2.2e-16, back
    if np.linalg.cond(M) < 1e8:
        Lref = np.linalg.cholesky(M)
        assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) < 1e-11


def _regularized(d, shift=0.5):
    """The same problem with Q + shift·I: δQ + BᵀB is then positive definite (with Q = 0 and
    m < n, BᵀB is singular and its ρ-shifted factor is rounding noise past the rank)."""
    import scipy.sparse as sp
    d = dict(d)
    d["Q"] = (d["Q"] + shift * sp.identity(d["Q"].shape[0], format="csc")).tocsc()
    return d


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_factor_persistent_kernels_match_dense_cholesky(case):
    """Every front of order ≥ 64 through the tile-DAG kernel (fronts with separators included)."""
    from oracle import ipm as O
    d = _regularized(small_cases()[case])
    solver = mh.init(hip_problem(d), mh.IPMSettings(big_front_ld=64))
    osolver = O.Solver(oracle_problem(d))
    delta = 0.37
    ok, rho = solver.kkt_factor(delta)
    ook, orho = osolver.kkt.init(osolver.Qd, delta)
    assert ok and ook and rho == 0.0 and orho == 0.0
    M = osolver.kkt.L + delta * osolver.Qd
    assert np.linalg.eigvalsh(M).min() > 1e-8 * np.linalg.norm(M, 2)
    _check_factor(np.tril(solver.dense_factor()), M)


@pytest.mark.parametrize("case", [2, 3])
def test_factor_persistent_kernels_singular_takes_rho_retry(case):
    """SOCP cases with Q = 0: BᵀB (rank ≤ m < n) is singular, so the unshifted factorization
    fails on both sides and the retry shift is ρ = eps·‖F‖ (uzawa.jl:97-112), the same value."""
    from oracle import ipm as O
    d = small_cases()[case]
    solver = mh.init(hip_problem(d), mh.IPMSettings(big_front_ld=64))
    osolver = O.Solver(oracle_problem(d))
    ok, rho = solver.kkt_factor(0.37)
    ook, orho = osolver.kkt.init(osolver.Qd, 0.37)
    assert rho > 0.0 and orho > 0.0
    assert abs(rho - orho) <= 1e-12 * orho


def test_factor_large_front():
    """A near-dense 1200-variable QP: one large root front through the persistent kernel."""
    from mumblebee_hip import generators as G
    from oracle import ipm as O
    d = G.qp(n=1200, m=600, seed=31)
    solver = mh.init(hip_problem(d))
    osolver = O.Solver(oracle_problem(d), order="mindegree")
    delta = 0.05
    ok, rho = solver.kkt_factor(delta)
    assert ok and rho == 0.0
    M = osolver.kkt.L + delta * osolver.Qd
    _check_factor(np.tril(solver.dense_factor()), M)


def test_long_sparse_rows_match_oracle():
    """A dense row of B and a variable in every constraint, among short rows: each exceeds
    SPMV_LONG × its matrix's lanes per row, so the sparse products hand it to a workgroup of its
    own (B row 0 in the m-row products, Bᵀ row 1 in the Bᵀ-tiled n-row products). The solve
    matches the oracle."""
    from mumblebee_hip import generators as G
    from oracle import ipm as O
    n, m = 600, 150
    rng = np.random.default_rng(41)
    d = G.qp(n=n, m=m, seed=41, adens=4.0 / n)
    B = d["B"].tolil()
    B[0, :] = rng.standard_normal((1, n))
    B[:, 1] = rng.standard_normal((m, 1))
    B = B.tocsc()
    B.sort_indices()
    def lanes(nnz, rows):  # lanes per row of the tiling: largest power of two ≤ the mean, ≤ 64
        g = 1
        while g < 64 and 2 * g <= nnz / rows:
            g *= 2
        return g

    Bc = B.tocsr()
    rowlen, collen = np.diff(Bc.indptr), np.diff(B.indptr)
    assert rowlen[0] > 64 * lanes(B.nnz, m) and collen[1] > 64 * lanes(B.nnz, n)
    d = dict(d, B=B, g=B @ rng.uniform(0.1, 1.0, n), name="qp_long_rows")
    res = mh.solve(hip_problem(d))
    ores = O.solve(oracle_problem(d))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_solve_persistent_kernels_match_oracle(case):
    from oracle import ipm as O
    d = small_cases()[case]
    res = mh.solve(hip_problem(d), mh.IPMSettings(big_front_ld=64))
    ores = O.solve(oracle_problem(d))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8
    assert abs(res.niter - ores.niter) <= 1


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_solve_matches_oracle(case):
    from oracle import ipm as O
    d = small_cases()[case]
    res = mh.solve(hip_problem(d))
    ores = O.solve(oracle_problem(d))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8
    assert rel(res.dobj, ores.dobj) < 1e-8
    assert res.pres < 1e-8 and res.dres < 1e-8
    assert abs(res.niter - ores.niter) <= 1
    # per-iteration trajectory (first iterations agree to rounding)
    for h, oh in list(zip(res.history, ores.history))[:4]:
        for key in ("pres", "dres", "mu"):
            assert rel(h[key], oh[key]) < 1e-6, (key, h[key], oh[key])
    rp, rd = kkt_certificate(d, res.p, res.d, res.y)
    assert rp < 1e-7 and rd < 1e-7
    sc = max(1.0, np.linalg.norm(ores.p))
    assert np.linalg.norm(res.p - ores.p) / sc < 1e-5
