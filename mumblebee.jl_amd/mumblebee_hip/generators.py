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
  maxcut    max-cut SDP relaxation min ⟨C, X⟩ s.t. diag(X) = 1, X ⪰ 0 on a graph of nv
            vertices: the band of half-width `band` plus extra random edges with probability
            `extra`, weights U(0,1), C = −Laplacian/4. X is decomposed chordally into
            overlapping windows of `clique` consecutive vertices (consecutive windows share
            `band` vertices), one PSD cone per window in svec form (lower triangle column by
            column, off-diagonal ×√2), with diag(X)=1 rows on each vertex's first window and
            consensus rows X_k[i,j] = X_{k+1}[i,j] on every overlap entry; each C entry goes to
            the first window containing it. The windows are a chordal cover of the band; extra
            edges are drawn only between vertices that share a window (uniformly random
            long-range edges make the chordal extension of a 2000-vertex band nearly dense).
  exppow    entropy on a probability simplex plus weighted geometric means under budgets:
            `nexp` exponential cones (u_i, x_i, t_i) with u_i = 1, Σ x_i = 1 and t_i ≤ x_i
            log(1/x_i); `npow` power cones (a_j, b_j, c_j) with α_j ~ U(0.2, 0.8), a_j + b_j = 2
            and |c_j| ≤ a_j^α b_j^(1−α); `m` further sparse equalities (density `dens`, N(0,1))
            over the x, a, b coordinates with g = B x0 at the strictly interior x0 (x_i = 1/nexp,
            a_j = b_j = 1). Objective: maximise Σ t_i + Σ c_j + εᵀ(x, a, b), ε ~ 0.1 N(0,1),
            bounded because x, a, b range over compact sets. With m = 0 and ε = 0 the optimum
            is log(nexp) + Σ_j (2α_j)^α_j (2(1−α_j))^(1−α_j).
            X = I is strictly feasible.
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


def maxcut(nv=2000, band=20, clique=48, extra=0.001, seed=0):
    rng = np.random.default_rng(seed)
    stride = clique - band
    starts = list(range(0, max(1, nv - band), stride))
    if starts[-1] + clique < nv:
        starts.append(nv - clique)
    wins = [(a, min(nv, a + clique)) for a in starts]
    # graph: band plus in-window extra edges, weights U(0, 1)
    W = {}
    for i in range(nv):
        for j in range(i + 1, min(nv, i + band + 1)):
            W[(j, i)] = rng.uniform(0.0, 1.0)
    for a, b in wins:
        for i in range(a, b):
            for j in range(max(i + band + 1, a), b):
                if (j, i) not in W and rng.random() < extra:
                    W[(j, i)] = rng.uniform(0.0, 1.0)
    deg = np.zeros(nv)
    for (j, i), w in W.items():
        deg[i] += w
        deg[j] += w
    Cent = {(i, i): -deg[i] / 4.0 for i in range(nv)}  # C = −L/4, L = diag(deg) − W
    for (j, i), w in W.items():
        Cent[(j, i)] = w / 4.0
    # one PSD cone per window
    sizes = [(b - a) * (b - a + 1) // 2 for a, b in wins]
    col0 = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    n = int(col0[-1])

    def pos(k, i, j):  # global column of X_k[i, j], i ≥ j (global vertex indices)
        a, b = wins[k]
        d = b - a
        ii, jj = i - a, j - a
        return int(col0[k] + jj * d - jj * (jj - 1) // 2 + (ii - jj))

    r2 = math.sqrt(2.0)
    f = np.zeros(n)
    owner = {}
    for k, (a, b) in enumerate(wins):
        for j in range(a, b):
            for i in range(j, b):
                if (i, j) in Cent and (i, j) not in owner:
                    owner[(i, j)] = k
                    f[pos(k, i, j)] = -(Cent[(i, j)] if i == j else r2 * Cent[(i, j)])
    rows, cols, vals, g = [], [], [], []
    seen = set()
    for k, (a, b) in enumerate(wins):  # diag(X) = 1 on the first window of each vertex
        for i in range(a, b):
            if i not in seen:
                seen.add(i)
                rows.append(len(g)); cols.append(pos(k, i, i)); vals.append(1.0); g.append(1.0)
    for k in range(len(wins) - 1):  # consensus on the overlap of consecutive windows
        lo, hi = wins[k + 1][0], wins[k][1]
        for j in range(lo, hi):
            for i in range(j, hi):
                r = len(g)
                rows += [r, r]; cols += [pos(k, i, j), pos(k + 1, i, j)]; vals += [1.0, -1.0]; g.append(0.0)
    B = sp.csc_matrix((vals, (rows, cols)), shape=(len(g), n))
    B.sort_indices()
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=f, g=np.array(g), K=[("psd",)] * len(wins), s=sizes,
                name=f"maxcut_nv{nv}_w{band}_c{clique}_s{seed}")


def chordal_cliques(nv, edges):
    """Maximal cliques of a chordal extension of the graph (nv vertices, edge set of (j, i) pairs)
    and a clique tree over them: minimum-degree elimination (ties to the smallest index) gives a
    perfect elimination ordering with higher neighbourhoods H(v); the vertex cliques {v} ∪ H(v)
    form a tree through the elimination-tree parent p(v) = the first-eliminated vertex of H(v),
    with H(v) ⊆ {p(v)} ∪ H(p(v)) (running intersection); each non-maximal vertex clique merges into
    a child clique that contains it. Returns (cliques as sorted vertex lists, parent clique index
    or −1)."""
    adj = [set() for _ in range(nv)]
    for j, i in edges:
        adj[i].add(j)
        adj[j].add(i)
    alive = set(range(nv))
    order, higher = [], {}
    # bucketed minimum degree: degree → set of vertices
    deg = [len(a) for a in adj]
    buckets = {}
    for v in range(nv):
        buckets.setdefault(deg[v], set()).add(v)
    while alive:
        dmin = min(k for k, b in buckets.items() if b)
        v = min(buckets[dmin])
        buckets[dmin].discard(v)
        nb = set(adj[v])
        order.append(v)
        higher[v] = nb
        alive.discard(v)
        for a in nb:
            adj[a].discard(v)
            adj[a] |= nb - {a}
            old = deg[a]
            deg[a] = len(adj[a])
            if old != deg[a]:
                buckets[old].discard(a)
                buckets.setdefault(deg[a], set()).add(a)
    pos = {v: k for k, v in enumerate(order)}
    parent = {v: (min(higher[v], key=pos.get) if higher[v] else None) for v in order}
    kids = {v: [] for v in order}
    for v in order:
        if parent[v] is not None:
            kids[parent[v]].append(v)
    rep = {}
    for v in order:  # children come first in elimination order
        r = v
        for w in kids[v]:
            if len(higher[w]) == len(higher[v]) + 1:  # {v} ∪ H(v) ⊂ {w} ∪ H(w)
                r = rep[w]
                break
        rep[v] = r
    reps = [v for v in order if rep[v] == v]
    index = {v: k for k, v in enumerate(reps)}
    cliques = [sorted({v} | higher[v]) for v in reps]
    cparent = [-1] * len(reps)
    for v in order:
        pv = parent[v]
        if pv is not None and rep[v] != rep[pv]:
            cparent[index[rep[v]]] = index[rep[pv]]
    return cliques, cparent


def merge_cliques(cliques, cparent, cap):
    """Clique-tree merging: children first (index order is a topological order), a clique joins
    its parent when their union has at most `cap` vertices; the merged tree keeps the running
    intersection property. Returns the merged (cliques, parent)."""
    sets = [set(c) for c in cliques]
    par = list(cparent)
    alive = [True] * len(sets)
    for k in range(len(sets)):
        q = par[k]
        if q < 0 or len(sets[k] | sets[q]) > cap:
            continue
        sets[q] |= sets[k]
        alive[k] = False
        for c in range(len(sets)):
            if par[c] == k:
                par[c] = q
    keep = [k for k in range(len(sets)) if alive[k]]
    index = {k: i for i, k in enumerate(keep)}
    return [sorted(sets[k]) for k in keep], [index[par[k]] if par[k] >= 0 else -1 for k in keep]


def maxcut_chordal(nv=2000, band=20, extra=1e-4, seed=0, merge=48):
    """Max-cut SDP relaxation min ⟨C, X⟩ s.t. diag(X) = 1, X ⪰ 0 on a graph of nv vertices: the
    band of half-width `band` plus extra edges drawn uniformly among all other vertex pairs, each
    with probability `extra` (a binomial count, placed uniformly); weights U(0, 1); C = −L/4. X is
    decomposed over the maximal cliques of a chordal extension (chordal_cliques), one PSD cone per
    clique in svec form, with the clique-tree consensus rows X_child[i, j] = X_parent[i, j] on
    every entry of each child's intersection with its parent, diag(X) = 1 rows on each vertex's
    first clique, and each C entry in the first clique that holds it (the chordal conversion of
    the reference's networked H₂ example, examples/sdp/networked_h2.jl:172-296). Cliques are merged
    into their parents while the union stays within `merge` vertices (0: no merging). Its optimum
    is the dense max-cut SDP's."""
    rng = np.random.default_rng(seed)
    W = {}
    for i in range(nv):
        for j in range(i + 1, min(nv, i + band + 1)):
            W[(j, i)] = rng.uniform(0.0, 1.0)
    npairs = nv * (nv - 1) // 2 - len(W)
    k = int(rng.binomial(npairs, extra)) if npairs > 0 else 0
    while k > 0:
        i, j = (int(x) for x in rng.integers(0, nv, 2))
        if i == j:
            continue
        a, b = max(i, j), min(i, j)
        if (a, b) in W:
            continue
        W[(a, b)] = rng.uniform(0.0, 1.0)
        k -= 1
    cliques, cparent = chordal_cliques(nv, W.keys())
    if merge:
        cliques, cparent = merge_cliques(cliques, cparent, merge)
    deg = np.zeros(nv)
    for (j, i), w in W.items():
        deg[i] += w
        deg[j] += w
    Cent = {(i, i): -deg[i] / 4.0 for i in range(nv)}
    for (j, i), w in W.items():
        Cent[(j, i)] = w / 4.0
    sizes = [len(c) * (len(c) + 1) // 2 for c in cliques]
    col0 = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(col0[-1])
    where = [{v: a for a, v in enumerate(c)} for c in cliques]

    def pos(k, i, j):  # global column of X_k[i, j] (vertices i, j of clique k)
        d = len(cliques[k])
        a, b = where[k][i], where[k][j]
        if a < b:
            a, b = b, a
        return int(col0[k] + b * d - b * (b - 1) // 2 + (a - b))

    r2 = math.sqrt(2.0)
    f = np.zeros(n)
    owner = set()
    for k, c in enumerate(cliques):
        for b_, j in enumerate(c):
            for i in c[b_:]:
                key = (max(i, j), min(i, j))
                if key in Cent and key not in owner:
                    owner.add(key)
                    f[pos(k, i, j)] = -(Cent[key] if i == j else r2 * Cent[key])
    rows, cols, vals, g = [], [], [], []
    seen = set()
    for k, c in enumerate(cliques):
        for i in c:
            if i not in seen:
                seen.add(i)
                rows.append(len(g)); cols.append(pos(k, i, i)); vals.append(1.0); g.append(1.0)
    for k, c in enumerate(cliques):
        q = cparent[k]
        if q < 0:
            continue
        sep = sorted(set(c) & set(cliques[q]))
        for b_, j in enumerate(sep):
            for i in sep[b_:]:
                r = len(g)
                rows += [r, r]; cols += [pos(k, i, j), pos(q, i, j)]; vals += [1.0, -1.0]; g.append(0.0)
    B = sp.csc_matrix((vals, (rows, cols)), shape=(len(g), n))
    B.sort_indices()
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=f, g=np.array(g), K=[("psd",)] * len(cliques), s=sizes,
                cliques=cliques, name=f"maxcut_chordal_nv{nv}_w{band}_x{extra}_s{seed}")


def exppow(nexp=2000, npow=500, m=499, seed=0, dens=0.005, eps=0.1, psd=0, rows="xab"):
    """Exponential and power cones (see the module docstring), optionally with one PSD block of
    order `psd` (svec columns after the power cones): objective ⟨C, X⟩ with C dense N(0,1)
    symmetrised, bounded by tr(X) = 1. rows = "xab": the m random equalities run over the x, a, b
    coordinates; rows = "all": over every column (X included) at density `dens`, with at least two
    entries per row. g = B x0 at the strictly interior x0 (u = 1, x = 1/nexp, t = 0, a = b = 1,
    c = 0, X = I/psd)."""
    rng = np.random.default_rng(seed)
    ntd = 3 * (nexp + npow)
    npsd = psd * (psd + 1) // 2
    n = ntd + npsd
    alphas = rng.uniform(0.2, 0.8, npow)
    rows_, cols, vals = [], [], []
    r = 0
    for i in range(nexp):  # u_i = 1
        rows_.append(r), cols.append(3 * i), vals.append(1.0)
        r += 1
    for i in range(nexp):  # Σ x_i = 1
        rows_.append(r), cols.append(3 * i + 1), vals.append(1.0)
    r += 1
    for j in range(npow):  # a_j + b_j = 2
        c0 = 3 * (nexp + j)
        rows_ += [r, r]
        cols += [c0, c0 + 1]
        vals += [1.0, 1.0]
        r += 1
    if psd:  # tr(X) = 1
        k = ntd
        for jj in range(psd):
            rows_.append(r), cols.append(k), vals.append(1.0)
            k += psd - jj
        r += 1
    xab = np.array([3 * i + 1 for i in range(nexp)] + [3 * (nexp + j) + k for j in range(npow) for k in (0, 1)])
    pool = xab if rows == "xab" else np.arange(n)
    for _ in range(m):  # sparse equalities; at least two entries per row
        k = max(2, rng.binomial(len(pool), dens))
        for c in rng.choice(pool, size=k, replace=False):
            rows_.append(r), cols.append(int(c)), vals.append(float(rng.standard_normal()))
        r += 1
    B = sp.csc_matrix((vals, (rows_, cols)), shape=(r, n))
    x0 = np.zeros(n)
    x0[0:3 * nexp:3] = 1.0
    x0[1:3 * nexp:3] = 1.0 / max(nexp, 1)
    x0[3 * nexp:ntd:3] = 1.0
    x0[3 * nexp + 1:ntd:3] = 1.0
    if psd:
        k = ntd
        for jj in range(psd):
            x0[k] = 1.0 / psd
            k += psd - jj
    g = B @ x0
    f = np.zeros(n)
    f[2:ntd:3] = 1.0  # t_i and c_j (objective −fᵀp)
    f[xab] = eps * rng.standard_normal(len(xab))
    K = [("exp",)] * nexp + [("pow", float(a)) for a in alphas]
    s_ = [3] * (nexp + npow)
    if psd:
        M = rng.standard_normal((psd, psd))
        Cm = (M + M.T) / 2.0
        r2 = math.sqrt(2.0)
        k = ntd
        for jj in range(psd):
            f[k] = -Cm[jj, jj]
            f[k + 1:k + psd - jj] = -r2 * Cm[jj + 1:, jj]
            k += psd - jj
        K.append(("psd",))
        s_.append(npsd)
    return dict(Q=sp.csc_matrix((n, n)), B=B, f=f, g=g, K=K, s=s_, alphas=alphas,
                name=f"exppow_e{nexp}_p{npow}_m{m}_psd{psd}_{rows}_s{seed}")


def served_request(rng, k, nmax=4000, seed0=5000):
    """Request k of the served stream: 60 % QP, 30 % SOCP, 10 % max-cut SDP (the chordal clique
    decomposition of maxcut_chordal on nv = n/25 vertices), n log-uniform in [50, nmax], the
    densities of the full-size workloads (SOCP rows scaled up for small n)."""
    n = int(log_uniform_n(rng, 50, nmax))
    u = rng.random()
    if u < 0.6:
        return qp(n=n, m=max(1, n // 2), seed=seed0 + k)
    if u < 0.9:
        return socp(n=n, m=max(1, n // 2), ncone=max(1, n // 40), seed=seed0 + k, adens=min(0.2, 20.0 / n))
    return maxcut_chordal(nv=max(12, n // 25), band=4, extra=0.01, seed=seed0 + k, merge=16)
