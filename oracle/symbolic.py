"""Symbolic analysis of the KKT matrix δH + BᵀB (oracle restatement of kkt.jl:23-53 and
utils.jl:135-145).

The vertex graph (`weightedgraph`, utils.jl:135-145) is restated exactly: vertices are cone
blocks weighted by their column count; edges are the union of the Q block pattern and the
line graph of B (two vertices touching a common row block).

The elimination ordering, elimination tree and supernode partition come, in the reference,
from CliqueTrees.jl `symbolic(weights, graph; alg=DEFAULT_ELIMINATION_ALGORITHM)` (kkt.jl:36),
an external package (compat "1.19.4", Project.toml:20) whose source is not available here.
PARITY UNPINNED against CliqueTrees. This module defines the ordering the build uses (a
weighted exact-external-degree minimum-degree ordering on the quotient graph, ties to the
smallest vertex index, followed by an elimination-tree postorder and fundamental
supernodes) and the C++ host restatement must match it bit for bit. Structural invariants
(etree parent after child, fronts partition the columns into contiguous ranges, each
separator is contained in the ancestors' columns, the factor pattern covers the pattern of
M) are checked in tests/test_host_parity.py::test_symbolic_structure_invariants, which also
checks the fronts against brute-force symbolic elimination of the permuted graph.
"""
import numpy as np


def vertex_graph(B, Q):
    """Adjacency sets of the weighted vertex graph pattern(Q) ∪ linegraph(B)."""
    N = B.nvtx
    adj = [set() for _ in range(N)]
    for v in range(Q.nvtx):
        for e in Q.arcs(v):
            u = int(Q.tgt[e])
            if u != v:
                adj[u].add(v)
                adj[v].add(u)
    rows = [[] for _ in range(B.nout)]
    for v in range(N):
        for e in B.arcs(v):
            rows[int(B.tgt[e])].append(v)
    for lst in rows:
        for a in lst:
            for b in lst:
                if a != b:
                    adj[a].add(b)
    weights = np.array([B.ncols(v) for v in range(N)], dtype=np.int64)
    return weights, adj


def mindegree(weights, adj):
    """Weighted minimum exact external degree on the quotient graph.

    Returns (order, struct): order[k] = vertex eliminated k-th; struct[v] = set of vertices
    (original labels) in the factor column block of v below v.
    """
    N = len(weights)
    A = [set(a) for a in adj]
    E = [set() for _ in range(N)]
    L = {}
    alive = [True] * N
    remaining = set(range(N))
    W = int(np.sum(weights))

    def reach(i):
        r = set(A[i])
        for e in E[i]:
            r |= L[e]
        r.discard(i)
        return r

    deg = [int(sum(weights[j] for j in reach(i))) for i in range(N)]
    order = []
    struct = [None] * N
    while remaining:
        p = min(remaining, key=lambda i: (deg[i], i))
        if deg[p] + weights[p] == W:
            # p reaches every remaining variable: once p is eliminated the rest form a
            # clique, in which minimum degree picks by (-weight, index)
            order.append(p)
            remaining.discard(p)
            struct[p] = set(remaining)
            rest = sorted(remaining, key=lambda i: (-int(weights[i]), i))
            for k, i in enumerate(rest):
                order.append(i)
                struct[i] = set(rest[k + 1:])
            break
        Lp = reach(p)
        struct[p] = set(Lp)
        Ep = E[p]
        for i in Lp:
            A[i].discard(p)
            A[i] -= Lp
            E[i] -= Ep
            E[i].add(p)
        for e in Ep:
            del L[e]
        L[p] = Lp
        alive[p] = False
        remaining.discard(p)
        W -= int(weights[p])
        order.append(p)
        for i in Lp:
            deg[i] = int(sum(weights[j] for j in reach(i)))
    return order, struct


class Symbolic:
    """Front structure in the final (postordered) vertex order.

    vperm[k]      old vertex label of new vertex k
    front_vtx     front f owns new vertices front_vtx[f]:front_vtx[f+1]
    front_sep     sorted new-vertex separator of each front
    front_parent  parent front (-1 for a root)
    xcol          column offsets of the new vertices (length N+1)
    res(f), sep(f) column ranges/lists
    """

    def __init__(self, vperm, weights_new, front_vtx, front_sep, front_parent):
        self.vperm = np.asarray(vperm, dtype=np.int64)
        self.weights = np.asarray(weights_new, dtype=np.int64)
        self.xcol = np.concatenate([[0], np.cumsum(self.weights)]).astype(np.int64)
        self.front_vtx = np.asarray(front_vtx, dtype=np.int64)
        self.front_sep = [np.asarray(s, dtype=np.int64) for s in front_sep]
        self.front_parent = np.asarray(front_parent, dtype=np.int64)
        self.nfront = len(front_sep)

    def res(self, f):
        return range(int(self.xcol[self.front_vtx[f]]), int(self.xcol[self.front_vtx[f + 1]]))

    def sep(self, f):
        cols = [np.arange(self.xcol[v], self.xcol[v + 1]) for v in self.front_sep[f]]
        return np.concatenate(cols).astype(np.int64) if cols else np.zeros(0, dtype=np.int64)

    def colperm(self, xcol_old):
        """Column permutation new → old induced by vperm."""
        return np.concatenate([np.arange(xcol_old[v], xcol_old[v + 1]) for v in self.vperm]).astype(np.int64) \
            if len(self.vperm) else np.zeros(0, dtype=np.int64)

    def flops(self):
        """Σ_f r³/3 + s·r² + s²·r (the factor flop model)."""
        tot = 0.0
        for f in range(self.nfront):
            r = len(self.res(f))
            s = len(self.sep(f))
            tot += r ** 3 / 3.0 + s * r * r + s * s * r
        return tot


def analyse(weights, adj, amalgamate=True):
    """Ordering → elimination tree → postorder → fundamental supernodes → relaxed
    amalgamation (a child front merges into its parent when the explicit zeros this adds stay
    below 80 % / 10 % / 5 % of the merged front for ≤16 / ≤48 / larger column counts, or
    always when the merged front has ≤4 columns, or when the child has ≤64 columns and its
    separator is at least half the parent's order, or when the merged front costs at most 2 %
    more flops than the two apart)."""
    N = len(weights)
    order, struct = mindegree(weights, adj)
    pos = np.empty(N, dtype=np.int64)
    pos[np.asarray(order, dtype=np.int64)] = np.arange(N)
    parent = np.full(N, -1, dtype=np.int64)
    for v in range(N):
        if struct[v]:
            parent[v] = min(struct[v], key=lambda u: pos[u])
    children = [[] for _ in range(N)]
    roots = []
    for v in sorted(range(N), key=lambda u: pos[u]):
        if parent[v] >= 0:
            children[parent[v]].append(v)
        else:
            roots.append(v)
    # postorder: roots and children in elimination order
    post = []
    for r in roots:
        stack = [(r, 0)]
        while stack:
            v, k = stack.pop()
            if k < len(children[v]):
                stack.append((v, k + 1))
                stack.append((children[v][k], 0))
            else:
                post.append(v)
    newpos = np.empty(N, dtype=np.int64)
    newpos[np.asarray(post, dtype=np.int64)] = np.arange(N)
    nparent = np.array([newpos[parent[v]] if parent[v] >= 0 else -1 for v in post], dtype=np.int64)
    nstruct = [sorted(int(newpos[u]) for u in struct[v]) for v in post]
    nchild = np.zeros(N, dtype=np.int64)
    for k in range(N):
        if nparent[k] >= 0:
            nchild[nparent[k]] += 1
    # fundamental supernodes
    front_vtx = [0]
    for k in range(N - 1):
        merge = (nparent[k] == k + 1 and nchild[k + 1] == 1
                 and len(nstruct[k]) == len(nstruct[k + 1]) + 1)
        if not merge:
            front_vtx.append(k + 1)
    if N > 0:
        front_vtx.append(N)
    nf = len(front_vtx) - 1
    vfront = np.empty(N, dtype=np.int64)
    for f in range(nf):
        vfront[front_vtx[f]:front_vtx[f + 1]] = f
    wpost = [int(weights[v]) for v in post]
    xc = np.concatenate([[0], np.cumsum(wpost)]).astype(np.int64) if N else np.zeros(1, dtype=np.int64)
    fsep, fpar, r, s = [], [], [], []
    flops = 0.0
    for f in range(nf):
        sep = nstruct[front_vtx[f + 1] - 1]
        fsep.append(sep)
        fpar.append(int(vfront[sep[0]]) if sep else -1)
        r.append(int(xc[front_vtx[f + 1]] - xc[front_vtx[f]]))
        s.append(sum(wpost[u] for u in sep))
        flops += r[f] ** 3 / 3.0 + s[f] * r[f] * r[f] + s[f] * s[f] * r[f]
    rep = list(range(nf))
    members = [[f] for f in range(nf)]
    if amalgamate:
        fch = [[] for _ in range(nf)]
        for f in range(nf):
            if fpar[f] >= 0:
                fch[fpar[f]].append(f)
        z = [0.0] * nf
        rr = list(r)
        kept = [[] for _ in range(nf)]  # children that stay separate fronts
        for f in range(nf):
            mem = []
            # worklist: original children, then the kept children of every child merged into f
            queue = list(fch[f])
            qi = 0
            while qi < len(queue):
                c = queue[qi]
                qi += 1
                rc, rf = float(rr[c]), float(rr[f])
                rt, st = rc + rf, float(s[f])
                zz = z[c] + z[f] + rc * (rf + st - float(s[c]))
                tot = rt * (rt + 1.0) / 2.0 + rt * st
                ok = (rt <= 4 or (rt <= 16 and zz <= 0.8 * tot) or (rt <= 48 and zz <= 0.1 * tot)
                      or zz <= 0.05 * tot)
                # the GPU rule: a child of at most one panel with a large separator merges
                ok = ok or (rc <= 64 and 2.0 * float(s[c]) >= rf + st)
                # and any merge that costs at most 2 % more flops than the two fronts apart
                fl = lambda a, b: a * a * a / 3.0 + b * a * a + b * b * a  # noqa: E731
                ok = ok or fl(rt, st) <= 1.02 * (fl(rc, float(s[c])) + fl(rf, st))
                if ok:
                    rr[f] = int(rt)
                    z[f] = zz
                    rep[c] = f
                    mem.extend(members[c])
                    members[c] = []
                    queue.extend(kept[c])
                else:
                    kept[f].append(c)
            mem.append(f)
            members[f] = mem

    def find(f):
        while rep[f] != f:
            f = rep[f]
        return f

    reps = [f for f in range(nf) if rep[f] == f]
    repidx = {R: k for k, R in enumerate(reps)}
    vperm, fvtx, vnew = [], [0], np.empty(N, dtype=np.int64)
    for R in reps:
        for f in members[R]:
            for k in range(front_vtx[f], front_vtx[f + 1]):
                vnew[k] = len(vperm)
                vperm.append(post[k])
        fvtx.append(len(vperm))
    front_sep = [sorted(int(vnew[u]) for u in fsep[R]) for R in reps]
    front_parent = [repidx[find(fpar[R])] if fpar[R] >= 0 else -1 for R in reps]
    wnew = np.asarray(weights)[np.asarray(vperm, dtype=np.int64)] if N else np.zeros(0, dtype=np.int64)
    sym = Symbolic(vperm, wnew, fvtx, front_sep, front_parent)
    sym.exact_flops = flops
    return sym
