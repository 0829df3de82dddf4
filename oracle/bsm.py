"""Block-sparse matrix (oracle restatement).

Restates the storage contract of src/BlockSparseArrays/src/block_sparse_matrix.jl:10-158:
a diagram U <-tgt- E -src-> V with dense column-major blocks. Vertex v owns columns
xcol[v]:xcol[v+1], out-vertex u owns rows xrow[u]:xrow[u+1], arc e (a block (u, v)) owns
values xblk[e]:xblk[e+1], arcs of a vertex are sorted by target (block_sparse_matrix.jl:703-711).
0-based here.

Pinned by: dense numpy products on random small inputs
(tests/test_oracle_problem.py::test_bsm_products_against_dense).
"""
import numpy as np


class BSM:
    __slots__ = ("nout", "nvtx", "narc", "ncol", "nrow", "nbnz",
                 "xsrc", "xcol", "xrow", "xblk", "tgt", "val")

    def __init__(self, nout, nvtx, xsrc, xcol, xrow, xblk, tgt, val):
        self.nout = int(nout)
        self.nvtx = int(nvtx)
        self.xsrc = np.asarray(xsrc, dtype=np.int64)
        self.xcol = np.asarray(xcol, dtype=np.int64)
        self.xrow = np.asarray(xrow, dtype=np.int64)
        self.xblk = np.asarray(xblk, dtype=np.int64)
        self.tgt = np.asarray(tgt, dtype=np.int64)
        self.val = np.asarray(val, dtype=np.float64)
        self.narc = int(self.xsrc[self.nvtx])
        self.ncol = int(self.xcol[self.nvtx])
        self.nrow = int(self.xrow[self.nout])
        self.nbnz = int(self.xblk[self.narc])

    # block_sparse_matrix.jl:559-619
    def ncols(self, v):
        return int(self.xcol[v + 1] - self.xcol[v])

    def nrows(self, u):
        return int(self.xrow[u + 1] - self.xrow[u])

    def colrange(self, v):
        return range(int(self.xcol[v]), int(self.xcol[v + 1]))

    def rowrange(self, u):
        return range(int(self.xrow[u]), int(self.xrow[u + 1]))

    def arcs(self, v):
        return range(int(self.xsrc[v]), int(self.xsrc[v + 1]))

    def block(self, e):
        """View of block e as an (nrows(tgt[e]) x ncols(v)) column-major matrix."""
        u = int(self.tgt[e])
        v = int(np.searchsorted(self.xsrc, e, side="right") - 1)
        return self.val[self.xblk[e]:self.xblk[e + 1]].reshape(
            (self.nrows(u), self.ncols(v)), order="F")

    def blocks(self):
        """Yield (u, v, e, view) for every arc."""
        for v in range(self.nvtx):
            nc = self.ncols(v)
            for e in self.arcs(v):
                u = int(self.tgt[e])
                blk = self.val[self.xblk[e]:self.xblk[e + 1]].reshape(
                    (self.nrows(u), nc), order="F")
                yield u, v, e, blk

    def copy(self):
        return BSM(self.nout, self.nvtx, self.xsrc.copy(), self.xcol.copy(),
                   self.xrow.copy(), self.xblk.copy(), self.tgt.copy(), self.val.copy())

    def to_dense(self):
        A = np.zeros((self.nrow, self.ncol))
        for u, v, e, blk in self.blocks():
            A[self.xrow[u]:self.xrow[u + 1], self.xcol[v]:self.xcol[v + 1]] = blk
        return A

    # block_sparse_matrix.jl:1084-1092: 2-norm of the stored values
    def norm(self):
        return float(np.linalg.norm(self.val[:self.nbnz]))

    # block_sparse_matrix.jl:720-725 → blas/gemv.jl: y ← α A x + β y  (or Aᵀ)
    def gemv(self, x, alpha=1.0, beta=0.0, y=None, trans=False):
        if trans:
            out = np.zeros(self.ncol) if y is None else y
            acc = np.zeros(self.ncol)
            for u, v, e, blk in self.blocks():
                acc[self.xcol[v]:self.xcol[v + 1]] += blk.T @ x[self.xrow[u]:self.xrow[u + 1]]
        else:
            out = np.zeros(self.nrow) if y is None else y
            acc = np.zeros(self.nrow)
            for u, v, e, blk in self.blocks():
                acc[self.xrow[u]:self.xrow[u + 1]] += blk @ x[self.xcol[v]:self.xcol[v + 1]]
        if beta == 0.0:
            out[:] = alpha * acc
        else:
            out[:] = alpha * acc + beta * out
        return out

    # block_sparse_matrix.jl:1094-1109: xᵀ A y
    def dot(self, x, y):
        s = 0.0
        for u, v, e, blk in self.blocks():
            s += float(x[self.xrow[u]:self.xrow[u + 1]] @ blk @ y[self.xcol[v]:self.xcol[v + 1]])
        return s


def selectvtxs(A, sel):
    """Column-vertex selection B[:, sel] (block_sparse_matrix.jl:1111-1169)."""
    xsrc = [0]
    xcol = [0]
    xblk = [0]
    tgt = []
    vals = []
    for av in sel:
        av = int(av)
        xcol.append(xcol[-1] + A.ncols(av))
        for e in A.arcs(av):
            tgt.append(int(A.tgt[e]))
            vals.append(A.val[A.xblk[e]:A.xblk[e + 1]])
            xblk.append(xblk[-1] + int(A.xblk[e + 1] - A.xblk[e]))
        xsrc.append(len(tgt))
    val = np.concatenate(vals) if vals else np.zeros(0)
    return BSM(A.nout, len(sel), xsrc, xcol, A.xrow.copy(), xblk, tgt, val)


def symselectvtxs(A, sel):
    """Symmetric vertex permutation A[sel, sel] of a square block-sparse matrix.

    Equals halfselectvtxs(halfselectvtxs(A, sel), sel) (block_sparse_matrix.jl:787-874,
    as used at kkt.jl:38): new vertex k is old vertex sel[k], every block keeps its
    orientation, and the arcs of each vertex are sorted by their new target.
    """
    sel = np.asarray(sel, dtype=np.int64)
    inv = np.empty_like(sel)
    inv[sel] = np.arange(len(sel))
    xcol = [0]
    for av in sel:
        xcol.append(xcol[-1] + A.ncols(int(av)))
    xsrc = [0]
    xblk = [0]
    tgt = []
    vals = []
    for av in sel:
        av = int(av)
        items = sorted((int(inv[A.tgt[e]]), e) for e in A.arcs(av))
        for nu, e in items:
            tgt.append(nu)
            vals.append(A.val[A.xblk[e]:A.xblk[e + 1]])
            xblk.append(xblk[-1] + int(A.xblk[e + 1] - A.xblk[e]))
        xsrc.append(len(tgt))
    val = np.concatenate(vals) if vals else np.zeros(0)
    return BSM(len(sel), len(sel), xsrc, xcol, xcol, xblk, tgt, val)


def gram(B):
    """L = Bᵀ B as a block-sparse matrix with both triangles (uzawa.jl:85 → `B' * B`,
    block_sparse_matrix.jl:745-751, pattern from matprod_dest :1171-1252, arcs sorted)."""
    # rows of B per out-vertex: vertices touching each out-vertex
    touch = [[] for _ in range(B.nout)]
    for v in range(B.nvtx):
        for e in B.arcs(v):
            touch[int(B.tgt[e])].append((v, e))
    xsrc = [0]
    xblk = [0]
    tgt = []
    vals = []
    for v in range(B.nvtx):
        nbr = {}
        for e in B.arcs(v):
            u = int(B.tgt[e])
            for w, f in touch[u]:
                nbr.setdefault(w, []).append((e, f))
        for w in sorted(nbr):
            blk = np.zeros((B.ncols(w), B.ncols(v)))
            for e, f in nbr[w]:
                blk += B.block(f).T @ B.block(e)
            tgt.append(w)
            vals.append(blk.reshape(-1, order="F"))
            xblk.append(xblk[-1] + blk.size)
        xsrc.append(len(tgt))
    val = np.concatenate(vals) if vals else np.zeros(0)
    return BSM(B.nvtx, B.nvtx, xsrc, B.xcol.copy(), B.xcol.copy(), xblk, tgt, val)
