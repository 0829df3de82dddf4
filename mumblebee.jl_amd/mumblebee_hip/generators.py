"""Seeded synthetic conic problems (the benchmark workloads), in the reference's problem
form (problem.jl:1-19):

    min ½ pᵀQp − fᵀp   s.t.  Bp = g,  p ∈ K

Each generator returns a dict with scipy CSC matrices Q (n×n), B (m×n), dense f, g, the
cone list K and block sizes s, ready for `IPMProblem(Q, B, f, g, 0.0, K, s)`.

Workloads:
  qp        sparse convex QP on the nonnegative orthant: Q = PᵀP + 1e-3 I with P n×n at
            0.5 % density, B m×n at 1 % density, N(0,1) values; g = B x0 with
            x0 ~ U(0.1, 1); f = −q with q ~ N(0,1). n=2000, m=1000 is the 2000-variable
            QP workload.
  socp      500 second-order cones plus a nonnegative block, n=10000, B 5000×10000 at
            0.2 % density. Cone sizes are drawn U{3..50} and then scaled by 9000/Σk (floored,
            at least 3) because 500 draws of U{3..50} average 13 250 > n; the remainder of the
            n columns is the nonnegative block. g = B x0 with x0 strictly interior. The linear
            objective is made dual-feasible (an objective drawn N(0,1) makes the problem
            unbounded with high probability when n = 2m): f = −(Bᵀy0 + s0) with y0 ~ N(0,1)
            and s0 a point of int K* built from N(0,1) draws.
"""
import math

import numpy as np
import scipy.sparse as sp


def _sprandn(rng, m, n, density):
    nnz = int(round(density * m * n))
    if nnz == 0:
        return sp.csc_matrix((m, n))
    idx = rng.choice(m * n, size=nnz, replace=False)
    rows = idx % m
    cols = idx // m
    vals = rng.standard_normal(nnz)
    return sp.csc_matrix((vals, (rows, cols)), shape=(m, n))


def qp(n=2000, m=1000, seed=0, pdens=0.005, adens=0.01):
    rng = np.random.default_rng(seed)
    P = _sprandn(rng, n, n, pdens)
    Q = (P.T @ P + 1e-3 * sp.identity(n, format="csc")).tocsc()
    A = _sprandn(rng, m, n, adens)
    x0 = rng.uniform(0.1, 1.0, n)
    g = A @ x0
    q = rng.standard_normal(n)
    Q.sort_indices()
    A.sort_indices()
    return dict(Q=Q, B=A, f=-q, g=g, K=[("pos",)], s=[n], name=f"qp_n{n}_m{m}_s{seed}")


def soc_sizes(rng, ncone, n):
    k = rng.integers(3, 51, size=ncone)
    target = min(n - max(1, n // 10), int(k.sum()))
    if k.sum() > target:
        k = np.maximum(3, np.floor(k * (target / k.sum())).astype(np.int64))
    return k


def socp(n=10000, m=5000, ncone=500, seed=0, adens=0.002):
    rng = np.random.default_rng(seed)
    ks = soc_sizes(rng, ncone, n)
    npos = n - int(ks.sum())
    assert npos >= 0
    K = []
    s = []
    if npos > 0:
        K.append(("pos",))
        s.append(npos)
    for k in ks:
        K.append(("soc",))
        s.append(int(k))
    A = _sprandn(rng, m, n, adens)
    # strictly interior x0 and s0
    x0 = np.empty(n)
    s0 = np.empty(n)
    x0[:npos] = rng.uniform(0.1, 1.0, npos)
    s0[:npos] = np.abs(rng.standard_normal(npos)) + 0.1
    j = npos
    for k in ks:
        k = int(k)
        t = rng.standard_normal(k - 1)
        x0[j + 1:j + k] = t
        x0[j] = np.linalg.norm(t) + rng.uniform(0.1, 1.0)
        t = rng.standard_normal(k - 1)
        s0[j + 1:j + k] = t
        s0[j] = np.linalg.norm(t) + abs(rng.standard_normal()) + 0.1
        j += k
    g = A @ x0
    y0 = rng.standard_normal(m)
    c = A.T @ y0 + s0
    Q = sp.csc_matrix((n, n))
    A.sort_indices()
    return dict(Q=Q, B=A, f=-c, g=g, K=K, s=s, name=f"socp_n{n}_m{m}_s{seed}")


def log_uniform_n(rng, lo=50, hi=4000):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
