"""Problem construction (oracle restatement of src/IPM/src/ipm/problem.jl and
src/BlockSparseArrays/src/compress.jl). Pure integer work plus value scatter.

Pinned by: invariants and brute force in tests/test_oracle_problem.py (every twin group
has one pattern and the groups are maximal; permutations are valid; block values equal
the CSC values; the C++ host restatement agrees bit-for-bit).

Sparse matrices are plain CSC triples (colptr, rowval, nzval), 0-based, rows sorted inside
each column, stored zeros kept (as Julia's SparseMatrixCSC keeps them).
"""
import numpy as np

from .bsm import BSM

POS, FREE, SOC, PSD, EXP, POW = "pos", "free", "soc", "psd", "exp", "pow"


class CSC:
    __slots__ = ("m", "n", "colptr", "rowval", "nzval")

    def __init__(self, m, n, colptr, rowval, nzval):
        self.m, self.n = int(m), int(n)
        self.colptr = np.asarray(colptr, dtype=np.int64)
        self.rowval = np.asarray(rowval, dtype=np.int64)
        self.nzval = np.asarray(nzval, dtype=np.float64)

    @staticmethod
    def from_scipy(A):
        A = A.tocsc(copy=True)
        A.sort_indices()
        return CSC(A.shape[0], A.shape[1], A.indptr, A.indices, A.data)

    @staticmethod
    def from_dense(A):
        m, n = A.shape
        colptr, rowval, nzval = [0], [], []
        for j in range(n):
            for i in range(m):
                if A[i, j] != 0:
                    rowval.append(i)
                    nzval.append(A[i, j])
            colptr.append(len(rowval))
        return CSC(m, n, colptr, rowval, nzval)

    def to_dense(self):
        A = np.zeros((self.m, self.n))
        for j in range(self.n):
            for p in range(self.colptr[j], self.colptr[j + 1]):
                A[self.rowval[p], j] += self.nzval[p]
        return A

    def nnz(self):
        return int(self.colptr[self.n])

    def transpose(self):
        cnt = np.zeros(self.m + 1, dtype=np.int64)
        for i in self.rowval[:self.nnz()]:
            cnt[i + 1] += 1
        colptr = np.cumsum(cnt)
        nxt = colptr[:-1].copy()
        rowval = np.empty(self.nnz(), dtype=np.int64)
        nzval = np.empty(self.nnz())
        for j in range(self.n):
            for p in range(self.colptr[j], self.colptr[j + 1]):
                i = self.rowval[p]
                q = nxt[i]
                rowval[q] = j
                nzval[q] = self.nzval[p]
                nxt[i] += 1
        return CSC(self.n, self.m, colptr, rowval, nzval)

    def select_cols(self, cols):
        colptr, rowval, nzval = [0], [], []
        for j in cols:
            a, b = self.colptr[j], self.colptr[j + 1]
            rowval.extend(self.rowval[a:b])
            nzval.extend(self.nzval[a:b])
            colptr.append(len(rowval))
        return CSC(self.m, len(cols), colptr, rowval, nzval)

    def permute(self, rowperm, colperm):
        """A[rowperm, colperm] (Julia SparseArrays.permute), rows re-sorted."""
        inv = np.empty(self.m, dtype=np.int64)
        inv[np.asarray(rowperm)] = np.arange(self.m)
        colptr, rowval, nzval = [0], [], []
        for j in colperm:
            a, b = self.colptr[j], self.colptr[j + 1]
            pairs = sorted(zip(inv[self.rowval[a:b]], self.nzval[a:b]))
            rowval.extend(int(i) for i, _ in pairs)
            nzval.extend(x for _, x in pairs)
            colptr.append(len(rowval))
        return CSC(self.m, len(colperm), colptr, rowval, nzval)


def vcat(A, B):
    colptr, rowval, nzval = [0], [], []
    for j in range(A.n):
        rowval.extend(A.rowval[A.colptr[j]:A.colptr[j + 1]])
        nzval.extend(A.nzval[A.colptr[j]:A.colptr[j + 1]])
        rowval.extend(B.rowval[B.colptr[j]:B.colptr[j + 1]] + A.m)
        nzval.extend(B.nzval[B.colptr[j]:B.colptr[j + 1]])
        colptr.append(len(rowval))
    return CSC(A.m + B.m, A.n, colptr, rowval, nzval)


def dropzeros(A):
    colptr, rowval, nzval = [0], [], []
    for j in range(A.n):
        for p in range(A.colptr[j], A.colptr[j + 1]):
            if A.nzval[p] != 0:
                rowval.append(A.rowval[p])
                nzval.append(A.nzval[p])
        colptr.append(len(rowval))
    return CSC(A.m, A.n, colptr, rowval, nzval)


def dropoffdz(Q):
    """Keep a structural diagonal, drop stored off-diagonal zeros (problem.jl:133-186)."""
    n = Q.m
    colptr, rowval, nzval = [0], [], []
    for j in range(n):
        iprv = -1
        for p in range(Q.colptr[j], Q.colptr[j + 1]):
            i = int(Q.rowval[p])
            x = float(Q.nzval[p])
            if x != 0 and iprv < j < i:
                rowval.append(j)
                nzval.append(0.0)
                iprv = j
            if x != 0 or i == j:
                rowval.append(i)
                nzval.append(x)
                iprv = i
        if iprv < j:
            rowval.append(j)
            nzval.append(0.0)
        colptr.append(len(rowval))
    return CSC(n, n, colptr, rowval, nzval)


def twins_exact(nout, nvtx, xsrc, tgt):
    """Exact supervariable partition by partition refinement (compress.jl:9-116).

    Partitions the `nout` outputs by their pattern over the `nvtx` inputs, where input j
    lists outputs tgt[xsrc[j]:xsrc[j+1]]. Restated 1-based with the same singly/doubly
    linked-list mechanics (abstract_linked_lists/*.jl) so that the output order matches.
    Returns (nsv, ptr[0..nsv], perm[0..nout-1]) 0-based.
    """
    N = nout
    var = [0] * (N + 2)
    svar = [0] * (N + 1)
    flag = [0] * (N + 1)
    card = [0] * (N + 1)
    head = [0] * (N + 1)
    prev = [0] * (N + 1)
    nxt = [0] * (N + 1)
    new = 0
    nsv = 0
    # free list: prepend!(free, 1:N)
    if N > 0:
        new = 1
        for i in range(1, N):
            var[i] = i + 1
        var[N] = 0

    def pop_free():
        nonlocal new
        i = new
        new = var[i]
        return i

    def push_free(i):
        nonlocal new
        var[i] = new
        new = i

    def dl_pushfirst(s, i):
        n = head[s]
        nxt[i] = n
        head[s] = i
        if n > 0:
            prev[n] = i

    def dl_delete(s, i):
        h = head[s]
        n = nxt[i]
        if i == h:
            head[s] = n
        else:
            p = prev[i]
            nxt[p] = n
            if n > 0:
                prev[n] = p

    if N > 0:
        s = pop_free()
        nsv += 1
        # prepend!(set(s), 1:N)
        h = head[s]
        head[s] = 1
        i = 1
        for j in range(2, N + 1):
            prev[j] = i
            nxt[i] = j
            i = j
        nxt[i] = h
        card[s] = N
        for i in range(1, N + 1):
            svar[i] = s

    for j in range(1, nvtx + 1):
        for e in range(int(xsrc[j - 1]), int(xsrc[j])):
            i = int(tgt[e]) + 1
            s = svar[i]
            if flag[s] < j:
                flag[s] = j
                if card[s] > 1:
                    dl_delete(s, i)
                    card[s] -= 1
                    var[s] = i
                    ns = pop_free()
                    svar[i] = ns
                    nsv += 1
                    dl_pushfirst(ns, i)
                    card[ns] += 1
            else:
                dl_delete(s, i)
                card[s] -= 1
                k = var[s]
                ns = svar[k]
                svar[i] = ns
                dl_pushfirst(ns, i)
                card[ns] += 1
                if head[s] == 0:
                    push_free(s)
                    nsv -= 1

    ptr = [0] * (N + 2)
    perm = [0] * (N + 1)
    t = 1
    ptr[t] = p = 1
    for s in range(1, N + 1):
        if head[s] == 0:
            continue
        i = head[s]
        while i > 0:
            svar[i] = t
            perm[p] = i
            p += 1
            i = nxt[i]
        t += 1
        ptr[t] = p
    return nsv, np.array(ptr[1:nsv + 2], dtype=np.int64) - 1, np.array(perm[1:N + 1], dtype=np.int64) - 1


def twins_tau(nout, nvtx, xrow, cadj, xcol, radj, tau):
    """Cosine-similarity grouping, tau < 1 (compress.jl:147-265). 0-based inputs:
    output i lists inputs cadj[xrow[i]:xrow[i+1]]; input j lists outputs radj[xcol[j]:xcol[j+1]]
    sorted ascending. Returns (nb, ptr, perm)."""
    assert 0 < tau <= 1
    adjmap = [0] * nout
    cosine = [0] * nout
    degree = [int(xrow[i + 1] - xrow[i]) for i in range(nout)]
    nb = 0
    for i in range(nout):
        if adjmap[i] != 0:
            continue
        nb += 1
        adjmap[i] = nb
        lst = []  # pushfirst order: iterate most-recent-first
        for e in range(int(xrow[i]), int(xrow[i + 1])):
            j = int(cadj[e])
            for f in range(int(xcol[j + 1]) - 1, int(xcol[j]) - 1, -1):
                k = int(radj[f])
                if not k > i:
                    break
                if adjmap[k] == 0:
                    if cosine[k] == 0:
                        lst.insert(0, k)
                    cosine[k] += 1
        for k in lst:
            cos = float(cosine[k])
            if cos * cos >= tau * tau * degree[i] * degree[k]:
                adjmap[k] = nb
            cosine[k] = 0
    ptr = [0] * (nb + 2)
    for i in range(nout):
        ptr[adjmap[i] + 1] += 1
    ptr[1] = 1
    for t in range(1, nb + 1):
        ptr[t + 1] += ptr[t]
    nxt = ptr[:]
    perm = [0] * (nout + 1)
    for i in range(nout):
        t = adjmap[i]
        perm[nxt[t]] = i + 1
        nxt[t] += 1
    return nb, np.array(ptr[1:nb + 2], dtype=np.int64) - 1, np.array(perm[1:], dtype=np.int64) - 1


def twins_of_cols(X, tau):
    """Group the columns of X by row pattern, as `twins(transpose(X), X, tau)` does
    (compress.jl:118-145 dispatch)."""
    Xt = X.transpose()  # rows of Xt = columns of X
    if tau == 1:
        return twins_exact(X.n, X.m, Xt.colptr, Xt.rowval)
    # A = transpose(X): nout = X.n, xrow = X.colptr / cadj = X.rowval,
    # xcol = Xt.colptr / radj = Xt.rowval
    return twins_tau(X.n, X.m, X.colptr, X.rowval, Xt.colptr, Xt.rowval, tau)


def twins_of_rows(A, tau):
    """Group the rows of A by column pattern, as `twins(A, transpose(A), tau)` does."""
    if tau == 1:
        return twins_exact(A.m, A.n, A.colptr, A.rowval)
    At = A.transpose()
    return twins_tau(A.m, A.n, At.colptr, At.rowval, A.colptr, A.rowval, tau)


def colpartition(B, Q, K, s, tau):
    """Column order POS | FREE | other cones, twin-grouping POS and FREE columns by their
    [Q; B] pattern (problem.jl:247-349). Returns (nvtx, xcol, perm, cones)."""
    ncol = B.n
    ncolP = sum(int(n) for k, n in zip(K, s) if k[0] == POS)
    ncolC = sum(int(n) for k, n in zip(K, s) if k[0] == FREE)
    work = [0] * ncol
    p = c = r = 0
    j = 0
    for k, n in zip(K, s):
        for _ in range(int(n)):
            if k[0] == POS:
                work[p] = j
                p += 1
            elif k[0] == FREE:
                work[ncolP + c] = j
                c += 1
            else:
                work[ncolP + ncolC + r] = j
                r += 1
            j += 1
    indP = work[:ncolP]
    indC = work[ncolP:ncolP + ncolC]
    XP = vcat(Q.select_cols(indP), B.select_cols(indP))
    XC = vcat(Q.select_cols(indC), B.select_cols(indC))
    nvtxP, xcolP, permP = twins_of_cols(XP, tau)
    nvtxC, xcolC, permC = twins_of_cols(XC, tau)
    nvtxR = sum(1 for k in K if k[0] not in (POS, FREE))
    nvtx = nvtxP + nvtxC + nvtxR
    perm = np.zeros(ncol, dtype=np.int64)
    xcol = np.zeros(nvtx + 1, dtype=np.int64)
    cones = []
    for jj in range(ncolP):
        perm[jj] = work[permP[jj]]
    for v in range(nvtxP):
        xcol[v] = xcolP[v]
        cones.append((POS,))
    for jj in range(ncolC):
        perm[ncolP + jj] = work[ncolP + permC[jj]]
    for v in range(nvtxC):
        xcol[nvtxP + v] = ncolP + xcolC[v]
        cones.append((FREE,))
    for jj in range(ncolP + ncolC, ncol):
        perm[jj] = work[jj]
    jj = 0
    for k, n in zip(K, s):
        if k[0] not in (POS, FREE):
            xcol[len(cones)] = ncolP + ncolC + jj
            cones.append(tuple(k))
            jj += int(n)
    xcol[nvtx] = ncol
    return nvtx, xcol, perm, cones


def colcompress(A, nvtx, xcol, colperm):
    """Pattern of A with the columns of each vertex merged (problem.jl:188-245)."""
    colptr, rowval = [0], []
    for v in range(nvtx):
        rows = set()
        for j in range(xcol[v], xcol[v + 1]):
            k = colperm[j]
            rows.update(int(i) for i in A.rowval[A.colptr[k]:A.colptr[k + 1]])
        rowval.extend(sorted(rows))
        colptr.append(len(rowval))
    return CSC(A.m, nvtx, colptr, rowval, np.ones(len(rowval)))


def compress2(A, nvtx, nout, xcol, xrow):
    """CSC → block-sparse with zero-filled dense blocks (compress.jl:359-483)."""
    row = np.zeros(A.m, dtype=np.int64)
    for u in range(nout):
        row[xrow[u]:xrow[u + 1]] = u
    xsrc, tgt, xblk = [0], [], [0]
    for v in range(nvtx):
        vn = int(xcol[v + 1] - xcol[v])
        outs = set()
        for j in range(xcol[v], xcol[v + 1]):
            outs.update(int(row[i]) for i in A.rowval[A.colptr[j]:A.colptr[j + 1]])
        for u in sorted(outs):
            tgt.append(u)
            xblk.append(xblk[-1] + int(xrow[u + 1] - xrow[u]) * vn)
        xsrc.append(len(tgt))
    val = np.zeros(xblk[-1])
    for v in range(nvtx):
        for e in range(xsrc[v], xsrc[v + 1]):
            u = tgt[e]
            nr = int(xrow[u + 1] - xrow[u])
            for jl, j in enumerate(range(xcol[v], xcol[v + 1])):
                for p in range(A.colptr[j], A.colptr[j + 1]):
                    i = int(A.rowval[p])
                    if xrow[u] <= i < xrow[u + 1]:
                        val[xblk[e] + jl * nr + (i - xrow[u])] = A.nzval[p]
    return BSM(nout, nvtx, xsrc, xcol, xrow, xblk, tgt, val)


class Problem:
    """IPMProblem restated (problem.jl:21-42): block-sparse Q, B; permuted f, g; cones;
    P1 (row permutation), P2 (column permutation) as index arrays new → old."""

    def __init__(self, Q, B, f, g, mu, K, P1, P2):
        self.Q, self.B, self.f, self.g, self.mu, self.K = Q, B, f, g, mu, K
        self.P1, self.P2 = P1, P2


def make_problem(Q, B, f, g, mu, K, s, compress=1.0):
    """IPMProblem{T,I,V}(Q::CSC, B::CSC, f, g, μ, K, s; compress) (problem.jl:84-127)."""
    m, n = B.m, B.n
    assert Q.m == n and Q.n == n and len(f) == n and len(g) == m
    assert len(s) == len(K) and sum(int(x) for x in s) == n and 0 < compress <= 1
    Q = dropoffdz(Q)
    B = dropzeros(B)
    nvtx, xcol, colperm, cones = colpartition(B, Q, K, s, compress)
    Bc = colcompress(B, nvtx, xcol, colperm)
    nout, xrow, rowperm = twins_of_rows(Bc, compress)
    Bp = B.permute(rowperm, colperm)
    Qp = Q.permute(colperm, colperm)
    Bb = compress2(Bp, nvtx, nout, xcol, xrow)
    Qb = compress2(Qp, nvtx, nvtx, xcol, xcol)
    fp = np.asarray(f, dtype=np.float64)[colperm]
    gp = np.asarray(g, dtype=np.float64)[rowperm]
    return Problem(Qb, Bb, fp, gp, float(mu), cones, np.asarray(rowperm), np.asarray(colperm))


def problem_from_blocks(Q, B, f, g, mu, K):
    """IPMProblem(Q::BlockSparseMatrix, B::BlockSparseMatrix, f, g, μ, K) (problem.jl:49-53): the
    block-sparse matrices as given, identity permutations (P1 = rows(B), P2 = cols(B))."""
    m, n = int(B.xrow[B.nout]), int(B.xcol[B.nvtx])
    return Problem(Q, B, np.asarray(f, dtype=np.float64), np.asarray(g, dtype=np.float64), float(mu), list(K),
                   np.arange(m), np.arange(n))
