"""Block-aware Ruiz equilibration (oracle restatement of src/IPM/src/scaling/scaling.jl).

Pinned by: the invariant that after the sweeps every row / vertex-block ∞-norm of the scaled
data is within tol of 1 or itmax sweeps ran (scaling.jl:143-169), and that scale/unscale
round-trips as a diagonal congruence (tests/test_oracle_scaling.py).
"""
import numpy as np


def infnorms(B, Q):
    """Row and vertex-block ∞-norms in one pass (scaling.jl:52-82)."""
    rnrm = np.zeros(B.nrow)
    vnrm = np.zeros(B.nvtx)
    for v in range(B.nvtx):
        nv = 0.0
        for e in Q.arcs(v):
            blk = Q.val[Q.xblk[e]:Q.xblk[e + 1]]
            if blk.size:
                nv = max(nv, float(np.max(np.abs(blk))))
        nc = B.ncols(v)
        for e in B.arcs(v):
            u = int(B.tgt[e])
            nr = B.nrows(u)
            blk = np.abs(B.val[B.xblk[e]:B.xblk[e + 1]].reshape((nr, nc), order="F"))
            if blk.size:
                nv = max(nv, float(blk.max()))
                r0 = int(B.xrow[u])
                rnrm[r0:r0 + nr] = np.maximum(rnrm[r0:r0 + nr], blk.max(axis=1))
        vnrm[v] = nv
    return rnrm, vnrm


def applyscaling(B, Q, yscl, vscl):
    """B ← E B D, Q ← D Q D (scaling.jl:88-110)."""
    for v in range(B.nvtx):
        sv = vscl[v]
        nc = B.ncols(v)
        for e in B.arcs(v):
            u = int(B.tgt[e])
            nr = B.nrows(u)
            r0 = int(B.xrow[u])
            blk = B.val[B.xblk[e]:B.xblk[e + 1]].reshape((nr, nc), order="F")
            blk *= (yscl[r0:r0 + nr] * sv)[:, None]
        for e in Q.arcs(v):
            Q.val[Q.xblk[e]:Q.xblk[e + 1]] *= vscl[int(Q.tgt[e])] * sv


def equilibrate(B, Q, f, g, itmax=10, tol=1e-3):
    """Mutates B, Q, f, g; returns (pscl, yscl) (scaling.jl:122-184)."""
    nrow, nvtx = B.nrow, B.nvtx
    pscl = np.ones(B.ncol)
    yscl = np.ones(nrow)
    vscl = np.ones(nvtx)
    vswp = np.ones(nvtx)
    rswp = np.ones(nrow)
    for _ in range(itmax):
        rnrm, vnrm = infnorms(B, Q)
        d = 0.0
        for v in range(nvtx):
            n = np.sqrt(vnrm[v])
            d = max(d, abs(1.0 - n))
            if n != 0:
                vswp[v] = 1.0 / n
                vscl[v] *= vswp[v]
        for i in range(nrow):
            n = np.sqrt(rnrm[i])
            d = max(d, abs(1.0 - n))
            if n != 0:
                rswp[i] = 1.0 / n
                yscl[i] *= rswp[i]
        applyscaling(B, Q, rswp, vswp)
        if d < tol:
            break
    for v in range(nvtx):
        s = vscl[v]
        for j in B.colrange(v):
            pscl[j] = s
            f[j] *= s
    g *= yscl
    return pscl, yscl
