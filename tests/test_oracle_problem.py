"""Pins the oracle's problem construction (problem.jl / compress.jl restatement) by brute force
and invariants, not by re-typing its formulas."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.problem import (CSC, dropoffdz, make_problem, twins_exact, twins_of_cols, twins_tau)
from oracle.bsm import gram, selectvtxs, symselectvtxs
from mumblebee_hip import generators as G


def _pattern_cols(X):
    return [tuple(X.rowval[X.colptr[j]:X.colptr[j + 1]]) for j in range(X.n)]


def _random_pattern(rng, m, n, dens, dup=0.3):
    A = (rng.random((m, n)) < dens).astype(float)
    # duplicate some columns so that twins exist
    for j in range(n):
        if rng.random() < dup and j > 0:
            A[:, j] = A[:, rng.integers(0, j)]
    return A


@pytest.mark.parametrize("seed", range(6))
def test_twins_exact_is_the_identical_pattern_partition(seed):
    rng = np.random.default_rng(seed)
    A = _random_pattern(rng, 12, 30, 0.25)
    X = CSC.from_dense(A)
    nsv, ptr, perm = twins_of_cols(X, 1.0)
    assert sorted(perm.tolist()) == list(range(X.n))
    pats = _pattern_cols(X)
    groups = [perm[ptr[k]:ptr[k + 1]].tolist() for k in range(nsv)]
    # brute force: columns grouped iff identical pattern
    for gcols in groups:
        assert len({pats[j] for j in gcols}) == 1
    assert len({pats[g[0]] for g in groups}) == nsv


def test_twins_exact_empty_and_singleton():
    assert twins_exact(0, 0, np.zeros(1, np.int64), np.zeros(0, np.int64))[0] == 0
    nsv, ptr, perm = twins_exact(1, 2, np.array([0, 1, 1]), np.array([0]))
    assert nsv == 1 and perm.tolist() == [0]


@pytest.mark.parametrize("seed", range(4))
def test_twins_tau_groups_satisfy_cosine_threshold(seed):
    rng = np.random.default_rng(seed)
    A = _random_pattern(rng, 20, 25, 0.3, dup=0.2)
    X = CSC.from_dense(A)
    tau = 0.7
    nb, ptr, perm = twins_of_cols(X, tau)
    assert sorted(perm.tolist()) == list(range(X.n))
    cols = [set(X.rowval[X.colptr[j]:X.colptr[j + 1]].tolist()) for j in range(X.n)]
    for k in range(nb):
        g = perm[ptr[k]:ptr[k + 1]].tolist()
        rep = min(g)  # the first member opens the group
        for j in g:
            cos = len(cols[rep] & cols[j])
            assert cos * cos >= tau * tau * len(cols[rep]) * len(cols[j]) - 1e-12


def test_dropoffdz_keeps_structural_diagonal():
    Q = np.array([[0.0, 0.0, 2.0], [0.0, 1.0, 0.0], [2.0, 0.0, 0.0]])
    C = CSC.from_dense(Q)
    # add an explicit stored off-diagonal zero
    C = CSC(3, 3, [0, 1, 2, 4], [2, 1, 0, 1], [2.0, 1.0, 2.0, 0.0])
    D = dropoffdz(C)
    np.testing.assert_array_equal(D.to_dense(), np.array([[0, 0, 2], [0, 1, 0], [2, 0, 0.0]]))
    for j in range(3):
        assert j in D.rowval[D.colptr[j]:D.colptr[j + 1]].tolist()
    for j in range(3):
        for p in range(D.colptr[j], D.colptr[j + 1]):
            assert D.rowval[p] == j or D.nzval[p] != 0


@pytest.mark.parametrize("gen,kw", [(G.qp, dict(n=80, m=40, seed=3, pdens=0.05, adens=0.08)),
                                    (G.socp, dict(n=200, m=90, ncone=8, seed=4, adens=0.03))])
def test_make_problem_blocks_reproduce_the_permuted_matrices(gen, kw):
    d = gen(**kw)
    prob = make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])
    Bd = d["B"].toarray()
    Qd = d["Q"].toarray()
    np.testing.assert_array_equal(prob.B.to_dense(), Bd[np.ix_(prob.P1, prob.P2)])
    np.testing.assert_array_equal(prob.Q.to_dense(), Qd[np.ix_(prob.P2, prob.P2)])
    np.testing.assert_array_equal(prob.f, d["f"][prob.P2])
    np.testing.assert_array_equal(prob.g, d["g"][prob.P1])
    # column order POS | FREE | others, one vertex per non-orthant cone
    kinds = [k[0] for k in prob.K]
    order = {"pos": 0, "free": 1}
    assert kinds == sorted(kinds, key=lambda k: order.get(k, 2))
    # each POS vertex groups identical [Q;B] column patterns
    X = np.vstack([Qd, Bd]) != 0
    for v in range(prob.B.nvtx):
        cols = prob.P2[prob.B.xcol[v]:prob.B.xcol[v + 1]]
        if prob.K[v][0] == "pos":
            for j in cols:
                assert np.array_equal(X[:, j], X[:, cols[0]])
    # each arc's block holds the values, absent arcs are all zero
    for v in range(prob.B.nvtx):
        tg = set(prob.B.tgt[prob.B.xsrc[v]:prob.B.xsrc[v + 1]].tolist())
        for u in range(prob.B.nout):
            blk = prob.B.to_dense()[prob.B.xrow[u]:prob.B.xrow[u + 1], prob.B.xcol[v]:prob.B.xcol[v + 1]]
            if u not in tg:
                assert not blk.any()


def test_bsm_products_against_dense():
    rng = np.random.default_rng(0)
    d = G.socp(n=150, m=70, ncone=6, seed=5, adens=0.04)
    prob = make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])
    B = prob.B
    Bd = B.to_dense()
    x = rng.standard_normal(B.ncol)
    y = rng.standard_normal(B.nrow)
    np.testing.assert_allclose(B.gemv(x), Bd @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(B.gemv(y, trans=True), Bd.T @ y, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(gram(B).to_dense(), Bd.T @ Bd, rtol=1e-12, atol=1e-12)
    sel = rng.permutation(B.nvtx)
    cols = np.concatenate([np.arange(B.xcol[v], B.xcol[v + 1]) for v in sel])
    np.testing.assert_array_equal(selectvtxs(B, sel).to_dense(), Bd[:, cols])
    L = gram(B)
    Ld = L.to_dense()
    np.testing.assert_array_equal(symselectvtxs(L, sel).to_dense(), Ld[np.ix_(cols, cols)])
    assert abs(B.dot(y, x) - y @ Bd @ x) < 1e-10 * (1 + abs(y @ Bd @ x))
