"""Block-sparse problem input through the C ABI (mb_problem_create_bsm): the form every reference
example builds, IPMProblem(Q::BlockSparseMatrix, B::BlockSparseMatrix, f, g, μ, K)
(problem.jl:49-53), solved on the device against the oracle constructing the same problem.

The obstacle QP follows examples/qp/obstacle_problem.jl:120-199 (its grid, obstacle and boundary
lift, Q the 5-point stiffness with a 1×1 block per grid vertex, no equality rows); the solution is
also checked as the linear complementarity problem it is (s ≥ 0, Ks − c ≥ 0, sᵀ(Ks − c) = 0)."""
import math

import numpy as np
import pytest

import mumblebee_hip as mh
from helpers import rel
from oracle import ipm as O
from oracle.bsm import BSM as OBSM
from oracle.problem import problem_from_blocks

pytestmark = pytest.mark.gpu


def _contact_radius():
    f = lambda r: -r * r * math.log(r) - (0.25 - r * r)  # noqa: E731
    lo, hi = 1e-6, 0.5 - 1e-12
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


RSTAR = _contact_radius()
ASTAR = -RSTAR ** 2 / math.sqrt(0.25 - RSTAR ** 2)


def _obstacle(x, y):
    r = math.hypot(x, y)
    r0, beta = 0.5, 0.9
    b = r0 * beta
    tmp = math.sqrt(r0 * r0 - b * b)
    B = tmp + b * b / tmp
    C = -b / tmp
    return B + r * C if r > b else math.sqrt(r0 * r0 - r * r)


def _exact(x, y):
    r = math.hypot(x, y)
    return ASTAR * math.log(r) if r > RSTAR else math.sqrt(max(0.25 - r * r, 0.0))


def obstacle_problem(L):
    """The eliminated form of the obstacle problem on an L×L grid of [−1, 1]²: variables s = u − ψ
    ≥ 0 at the (L−2)² interior nodes, min ½ sᵀKs − (c − Kψ)ᵀs."""
    h = 2.0 / (L - 1)
    coord = lambda i: -1.0 + i * h  # noqa: E731
    idx = -np.ones((L, L), dtype=np.int64)
    k = 0
    for j in range(1, L - 1):
        for i in range(1, L - 1):
            idx[i, j] = k
            k += 1
    n = k
    psi, c = np.zeros(n), np.zeros(n)
    for j in range(1, L - 1):
        for i in range(1, L - 1):
            v = idx[i, j]
            psi[v] = _obstacle(coord(i), coord(j))
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                if idx[i + di, j + dj] < 0:
                    c[v] += _exact(coord(i + di), coord(j + dj))
    I, J, V = [], [], []
    Kpsi = np.zeros(n)
    for v in range(n):
        I.append(v), J.append(v), V.append([[4.0]])
        Kpsi[v] += 4.0 * psi[v]
    for j in range(1, L - 1):
        for i in range(1, L - 1):
            v = idx[i, j]
            for di, dj in ((1, 0), (0, 1)):
                w = idx[i + di, j + dj]
                if w < 0:
                    continue
                I += [v, w]
                J += [w, v]
                V += [[[-1.0]], [[-1.0]]]
                Kpsi[v] -= psi[w]
                Kpsi[w] -= psi[v]
    Q = mh.blocksparse(I, J, V)
    B = mh.blocksparse([], [], [], [], [1] * n)
    return Q, B, c - Kpsi, np.zeros(0), [mh.PositiveCone() for _ in range(n)]


def _oracle_blocks(A):
    return OBSM(A.nout, A.nvtx, A.xsrc, A.xcol, A.xrow, A.xblk, A.tgt, A.val)


def _cone_names(K):
    names = {0: "pos", 1: "free", 2: "soc", 3: "psd", 4: "exp"}
    return [("pow", k.param) if k.kind == 5 else (names[k.kind],) for k in K]


def _solve_both(Q, B, f, g, K):
    res = mh.solve(mh.IPMProblem.from_blocks(Q, B, f, g, 0.0, K))
    ores = O.solve(problem_from_blocks(_oracle_blocks(Q), _oracle_blocks(B), f, g, 0.0, _cone_names(K)))
    return res, ores


def test_obstacle_qp_from_blocks_matches_oracle():
    Q, B, f, g, K = obstacle_problem(22)
    res, ores = _solve_both(Q, B, f, g, K)
    assert res.status == mh.IPMStatus.OPTIMAL and int(ores.status) == int(res.status)
    assert rel(res.pobj, ores.pobj) < 1e-8
    assert np.linalg.norm(res.p - ores.p) <= 1e-6 * max(1.0, np.linalg.norm(ores.p))
    Kd = Q.toscipy()
    s = res.p
    w = Kd @ s - f
    assert s.min() > -1e-9 and w.min() > -1e-7
    assert abs(s @ w) <= 1e-6 * max(1.0, np.linalg.norm(f))


def test_random_block_problem_matches_oracle_and_the_csc_route():
    """Multi-column vertices (nonnegative blocks of 3, second-order cones of 5) with dense blocks
    of B over row blocks of several heights; the same problem through the CSC constructor reaches
    the same optimum."""
    rng = np.random.default_rng(41)
    widths = [3] * 8 + [5] * 6
    K = [mh.PositiveCone()] * 8 + [mh.SecondOrderCone()] * 6
    heights = [2, 3, 1, 4, 2, 3]
    I, J, V = [], [], []
    for u, hgt in enumerate(heights):
        for v in sorted(rng.choice(len(widths), size=4, replace=False)):
            I.append(u), J.append(int(v)), V.append(rng.standard_normal((hgt, widths[v])))
    B = mh.blocksparse(I, J, V, heights, widths)
    QI, QJ, QV = [], [], []
    for v, w in enumerate(widths):
        M = rng.standard_normal((w, w))
        QI.append(v), QJ.append(v), QV.append(M @ M.T / w + 0.1 * np.eye(w))
    Q = mh.blocksparse(QI, QJ, QV, widths, widths)
    n, m = sum(widths), sum(heights)
    x0 = np.concatenate([rng.uniform(0.5, 1.0, 3) for _ in range(8)] +
                        [np.concatenate([[3.0], 0.3 * rng.standard_normal(4)]) for _ in range(6)])
    Bd = B.toscipy()
    g = Bd @ x0
    f = rng.standard_normal(n)
    res, ores = _solve_both(Q, B, f, g, K)
    assert res.status == mh.IPMStatus.OPTIMAL and int(ores.status) == int(res.status)
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8
    assert abs(res.niter - ores.niter) <= 1
    csc = mh.solve(mh.IPMProblem(Q.toscipy(), Bd, f, g, 0.0, K, widths))
    assert csc.status == mh.IPMStatus.OPTIMAL
    assert rel(csc.pobj, res.pobj) < 1e-8


def test_kkt_factor_with_the_callers_h():
    """initkkt!(kkt, H; δ) with an H the caller holds (uzawa.jl:73-76): block-sparse over the
    solver's vertex partition in its internal coordinates. The following solvekkt! solves
    H x − Bᵀ y = f, B x = g with that H (a dense solve of the same saddle-point system); scale!
    brings the solver's own H back."""
    import scipy.sparse as sp
    rng = np.random.default_rng(43)
    K = [mh.PositiveCone(), mh.SecondOrderCone(), mh.SecondOrderCone(), mh.CofreeCone()]
    sizes = [7, 4, 6, 3]
    n, m = sum(sizes), 6
    A = rng.normal(size=(n, 2))
    Q = sp.csc_matrix(0.2 * A @ A.T)
    Bm = sp.csc_matrix(rng.normal(size=(m, n)))
    prob = mh.IPMProblem(Q, Bm, rng.normal(size=n), rng.normal(size=m), 0.0, K, sizes)
    s = mh.init(prob, mh.IPMSettings(scale_max_iter=0))
    try:
        lay = s.layout()
        P1, P2, xcol = lay["P1"], lay["P2"], lay["xcol"]
        nv = len(xcol) - 1
        widths = np.diff(xcol)
        Qi = Q.toarray()[np.ix_(P2, P2)]
        # H: Q's blocks (dense here: Q is dense) plus an SPD matrix on every diagonal block
        Hd = Qi.copy()
        for v in range(nv):
            a, b = xcol[v], xcol[v + 1]
            M = rng.normal(size=(b - a, b - a))
            Hd[a:b, a:b] += M @ M.T / (b - a) + np.eye(b - a)
        I, J, V = [], [], []
        for v in range(nv):
            for u in range(nv):
                I.append(u), J.append(v), V.append(Hd[xcol[u]:xcol[u + 1], xcol[v]:xcol[v + 1]])
        H = mh.blocksparse(I, J, V, widths, widths)
        ok, rho = s.kkt_factor_h(H, 0.7)
        assert ok and rho == 0.0
        Bi = Bm.toarray()[np.ix_(P1, P2)]
        f, g = rng.normal(size=n), rng.normal(size=m)
        x, y, ncg, nir, status, _, _ = s.kkt_solve(f, g, ftol=1e-12, gtol=1e-12)
        assert status in (mh.KKTStatus.KKT_SOLVED, mh.KKTStatus.KKT_STAGNATED)
        ref = np.linalg.solve(np.block([[Hd, -Bi.T], [Bi, np.zeros((m, m))]]), np.concatenate([f, g]))
        assert np.linalg.norm(x - ref[:n]) <= 1e-8 * (1 + np.linalg.norm(ref[:n]))
        assert np.linalg.norm(y - ref[n:]) <= 1e-8 * (1 + np.linalg.norm(ref[n:]))
        # scale! at an interior point brings the solver's own H back: Q plus the cone blocks
        p = np.zeros(n)
        for v, k in enumerate(lay["kind"]):
            a, b = xcol[v], xcol[v + 1]
            if k == 0:
                p[a:b] = 1.0
            elif k == 2:
                p[a] = 2.0
                p[a + 1:b] = 0.1
        ok, _ = s.cone_scale(p, p)
        assert ok
        D = s.cone_blocks()
        Hown = Qi.copy()
        o = 0
        for v, k in enumerate(lay["kind"]):
            a, b = xcol[v], xcol[v + 1]
            if k == 0:
                Hown[np.arange(a, b), np.arange(a, b)] += D[o:o + b - a]
                o += b - a
            elif k == 2:
                Hown[a:b, a:b] += D[o:o + (b - a) ** 2].reshape(b - a, b - a, order="F")
                o += (b - a) ** 2
        Hown[xcol[3]:, xcol[3]:] += 0.0  # the free block adds nothing
        ok, _ = s.kkt_factor(0.7)
        assert ok
        x, y, _, _, status, _, _ = s.kkt_solve(f, g, ftol=1e-12, gtol=1e-12)
        free = [v for v, k in enumerate(lay["kind"]) if k == 1]
        if not free:  # with a free block the own H is singular there; the solve is then not unique
            ref = np.linalg.solve(np.block([[Hown, -Bi.T], [Bi, np.zeros((m, m))]]), np.concatenate([f, g]))
            assert np.linalg.norm(x - ref[:n]) <= 1e-7 * (1 + np.linalg.norm(ref[:n]))
        else:
            r = Hown @ x - Bi.T @ y - f
            assert np.linalg.norm(r) <= 1e-6 * (1 + np.linalg.norm(f)) and np.linalg.norm(Bi @ x - g) <= 1e-6 * (1 + np.linalg.norm(g))
    finally:
        s.close()
