"""mumblebee_hip — the MI355X conic IPM solve path behind Mumblebee.jl's own interface.

This module mirrors the reference's Julia API for the solve path (same names, argument meaning
and error behaviour) over the C ABI in include/mumblebee_hip.h:

    IPMProblem(Q, B, f, g, μ, K, s; compress)   src/IPM/src/ipm/problem.jl:64-127
    IPMSettings(...)                            src/IPM/src/ipm/settings/ipm.jl:1-18
    init(prob, settings; p0, d0, y0)            src/IPM/src/ipm/solver/ipm.jl:485-491
    step(solver)  (step!)                       src/IPM/src/ipm/solver/ipm.jl:577-774
    solve(prob, settings) / solve_(solver)      src/IPM/src/ipm/solver/ipm.jl:798-804
    IPMResult, IPMHistory rows, IPMStatus       result/ipm.jl, history/ipm.jl, ipm/ipm.jl:1

All numerics run in the HIP library; there is no CPU fallback. Loading fails loudly when the
library has not been built.
"""
import atexit
import ctypes
import enum
import os
import weakref
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmumblebee_hip.so")
_lib = None
_live_solvers = weakref.WeakSet()
_live_problems = weakref.WeakSet()
_live_servers = weakref.WeakSet()


@atexit.register
def _release_all():
    """Free device state while the HIP runtime is still up (before its own exit handlers)."""
    for s in list(_live_servers):
        s.close()
    for s in list(_live_solvers):
        s.close()
    for p in list(_live_problems):
        p.close()


class IPMStatus(enum.IntEnum):
    CONTINUE = 0
    OPTIMAL = 1
    NEAR_OPTIMAL = 2
    STALLED = 3
    NUMERICAL_FAILURE = 4
    ITERATION_LIMIT = 5
    PRIMAL_INFEASIBLE = 6
    DUAL_INFEASIBLE = 7
    ILL_POSED = 8
    NEAR_PRIMAL_INFEASIBLE = 9
    NEAR_DUAL_INFEASIBLE = 10
    NEAR_ILL_POSED = 11


class KKTStatus(enum.IntEnum):
    KKT_CONTINUE = 0
    KKT_SOLVED = 1
    KKT_STAGNATED = 2
    KKT_ITMAX = 3
    KKT_NUMERICAL_FAILURE = 4
    KKT_IRMAX = 5


# cones (src/IPM/src/cone/*.jl); kind codes are the C ABI's
class AbstractCone:
    kind = -1
    param = 0.0


class PositiveCone(AbstractCone):
    kind = 0


class CofreeCone(AbstractCone):
    kind = 1


class SecondOrderCone(AbstractCone):
    kind = 2


class SemidefiniteCone(AbstractCone):
    kind = 3


class ExponentialCone(AbstractCone):
    kind = 4


class PowerCone(AbstractCone):
    kind = 5

    def __init__(self, alpha):
        self.param = float(alpha)


_I64P = ctypes.POINTER(ctypes.c_int64)
_I32P = ctypes.POINTER(ctypes.c_int32)
_DP = ctypes.POINTER(ctypes.c_double)


class _Settings(ctypes.Structure):
    _fields_ = [("verbose", ctypes.c_int32), ("step_frac", ctypes.c_double), ("feas_tol", ctypes.c_double),
                ("gap_tol", ctypes.c_double), ("max_iter", ctypes.c_int32), ("near_factor", ctypes.c_double),
                ("stall_tol", ctypes.c_double), ("refine_stall_tol", ctypes.c_double),
                ("scale_max_iter", ctypes.c_int32), ("refine_max_iter", ctypes.c_int32),
                ("newton_max_iter", ctypes.c_int32), ("aug_tol", ctypes.c_double),
                ("infeas_abs", ctypes.c_double), ("infeas_rel", ctypes.c_double),
                ("amalgamate", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("big_front_ld", ctypes.c_int32), ("dag_grid", ctypes.c_int32), ("mixed_mu", ctypes.c_double)]


EXPORTS = {
    # name: (restype, argtypes)
    "mb_settings_default": (None, [ctypes.POINTER(_Settings)]),
    "mb_problem_create": (ctypes.c_void_p, [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, _DP, _I64P, _I64P, _DP,
                                            _DP, _DP, ctypes.c_double, ctypes.c_int64, _I32P, _DP, _I64P,
                                            ctypes.c_double]),
    "mb_problem_destroy": (None, [ctypes.c_void_p]),
    "mb_problem_create_bsm": (ctypes.c_void_p, [ctypes.c_int64, ctypes.c_int64, _I64P, _I64P, _I64P, _I64P, _I64P,
                                                _DP, _I64P, _I64P, _I64P, _DP, _DP, _DP, ctypes.c_double, _I32P,
                                                _DP, ctypes.c_int32]),
    "mb_problem_dims": (ctypes.c_int, [ctypes.c_void_p, _I64P]),
    "mb_problem_export": (ctypes.c_int, [ctypes.c_void_p, _I64P, _I64P, _I32P, _I64P, _I64P, _I64P, _I64P, _I64P,
                                         _DP, _I64P, _I64P, _I64P, _DP, _DP, _DP]),
    "mb_problem_symbolic_dims": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I64P]),
    "mb_problem_symbolic": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I64P, _I64P, _I64P, _I64P, _I64P,
                                           _DP]),
    "mb_solver_create": (ctypes.c_void_p, [ctypes.c_void_p, ctypes.POINTER(_Settings), _DP, _DP, _DP]),
    "mb_solver_destroy": (None, [ctypes.c_void_p]),
    "mb_solver_step": (ctypes.c_int, [ctypes.c_void_p]),
    "mb_solver_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "mb_solver_result": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP, _DP]),
    "mb_solver_niter": (ctypes.c_int64, [ctypes.c_void_p]),
    "mb_solver_history": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_solver_stats": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_solver_reinit": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP]),
    "mb_solver_set_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "mb_solver_profile": (ctypes.c_int, [ctypes.c_void_p, _DP, ctypes.c_int32]),
    "mb_solver_dag_trace": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _I64P, ctypes.c_int32]),
    "mb_server_create": (ctypes.c_void_p, [ctypes.c_int32, ctypes.c_double, ctypes.c_int32]),
    "mb_server_destroy": (None, [ctypes.c_void_p]),
    "mb_server_submit": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_Settings)]),
    "mb_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, _DP, _DP, _DP, _DP, _DP]),
    "mb_server_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "mb_server_stats": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_clock": (ctypes.c_double, []),
    "mb_device_info": (ctypes.c_int, [_DP]),
    "mb_cached_bytes": (ctypes.c_int64, []),
    "mb_release_cached_memory": (ctypes.c_int64, []),
    "mb_solver_kkt_factor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _DP]),
    "mb_solver_frule": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, ctypes.c_double, _DP, _DP, _DP]),
    "mb_solver_rrule": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP, _DP, _DP, _DP]),
    "mb_solver_kkt_factor_h": (ctypes.c_int, [ctypes.c_void_p, _I64P, _I64P, _I64P, _DP, ctypes.c_int32,
                                               ctypes.c_double, _DP]),
    "mb_solver_dense_factor": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_probe_mfma_f64": (ctypes.c_int, [_DP]),
    "mb_solver_ncone": (ctypes.c_int64, [ctypes.c_void_p]),
    "mb_solver_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int32), _DP, _DP, _DP]),
    "mb_solver_kkt_solve": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP, _DP, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_int32), _DP]),
    "mb_cone_scale": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, _DP]),
    "mb_cone_blocks_size": (ctypes.c_int64, [ctypes.c_void_p]),
    "mb_cone_blocks": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_cone_corr0": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, _DP]),
    "mb_cone_corr": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, ctypes.c_double, _DP]),
    "mb_cone_maxsteps": (ctypes.c_int, [ctypes.c_void_p, _DP, _DP, ctypes.c_double, _DP]),
    "mb_probe_mfma_f32": (ctypes.c_int, [_DP]),
    "mb_dense_cholesky": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_void_p]),
    "mb_dense_cholesky_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                                  ctypes.c_void_p]),
    "mb_server_create_gang": (ctypes.c_void_p, [ctypes.c_int32, ctypes.c_double]),
    "mb_gang_create": (ctypes.c_void_p, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "mb_gang_destroy": (None, [ctypes.c_void_p]),
    "mb_solver_solve_in": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "mb_solve_batch": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32, ctypes.c_void_p, _I32P]),
    "mb_gang_stats": (ctypes.c_int, [ctypes.c_void_p, _DP]),
    "mb_last_error": (ctypes.c_char_p, []),
}


def lib():
    """Load libmumblebee_hip.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build the HIP library first (__graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _err(what):
    msg = lib().mb_last_error()
    return RuntimeError(f"{what}: {msg.decode() if msg else 'unknown error'}")


def _p(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def _csc(A, m, n):
    import scipy.sparse as sp
    A = sp.csc_matrix(A, shape=(m, n), copy=True)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, dtype=np.int64), np.ascontiguousarray(A.indices, dtype=np.int64),
            np.ascontiguousarray(A.data, dtype=np.float64))


@dataclass
class IPMSettings:
    """IPMSettings{Float64} (settings/ipm.jl:1-18) plus this build's knobs: `amalgamate`
    (relaxed supernode amalgamation), `precision` (0: FP64 factor; 1: the large fronts factored
    in FP32 on FP32 MFMA with FP64 Uzawa/CG refinement, back to FP64 once μ < `mixed_mu`),
    `big_front_ld`, `dag_grid`."""
    verbose: int = 0
    step_frac: float = 0.99
    feas_tol: float = 1e-8
    gap_tol: float = 1e-8
    max_iter: int = 100
    near_factor: float = 1000.0
    stall_tol: float = 1e-3
    refine_stall_tol: float = 0.5
    scale_max_iter: int = 10
    refine_max_iter: int = 10
    newton_max_iter: int = 100
    aug_tol: float = 1e-7
    infeas_abs: float = 1e-8
    infeas_rel: float = 1e-8
    amalgamate: int = 1
    precision: int = 0
    big_front_ld: int = 512
    dag_grid: int = 0
    mixed_mu: float = 1e-5

    def _c(self):
        s = _Settings()
        for k in _Settings._fields_:
            setattr(s, k[0], getattr(self, k[0]))
        return s


class BlockSparse:
    """BlockSparseMatrix{Float64, Int} (block_sparse_matrix.jl:10-109) as 0-based arrays: vertex v
    owns columns xcol[v]:xcol[v+1], out vertex u rows xrow[u]:xrow[u+1], arc e of vertex v
    (xsrc[v]:xsrc[v+1]) targets out vertex tgt[e] with the column-major block
    val[xblk[e]:xblk[e+1]] of shape (rows of tgt[e], columns of v)."""

    def __init__(self, nout, nvtx, xsrc, xcol, xrow, xblk, tgt, val):
        self.nout, self.nvtx = int(nout), int(nvtx)
        self.xsrc = np.ascontiguousarray(xsrc, dtype=np.int64)
        self.xcol = np.ascontiguousarray(xcol, dtype=np.int64)
        self.xrow = np.ascontiguousarray(xrow, dtype=np.int64)
        self.xblk = np.ascontiguousarray(xblk, dtype=np.int64)
        self.tgt = np.ascontiguousarray(tgt, dtype=np.int64)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def shape(self):
        return int(self.xrow[-1]), int(self.xcol[-1])

    def toscipy(self):
        """SparseArrays.sparse(A) (block_sparse_matrix.jl:361-395): every block entry stored."""
        import scipy.sparse as sp
        rows, cols, vals = [], [], []
        for v in range(self.nvtx):
            c0, nc = int(self.xcol[v]), int(self.xcol[v + 1] - self.xcol[v])
            for e in range(int(self.xsrc[v]), int(self.xsrc[v + 1])):
                u = int(self.tgt[e])
                r0, nr = int(self.xrow[u]), int(self.xrow[u + 1] - self.xrow[u])
                blk = self.val[self.xblk[e]:self.xblk[e + 1]].reshape((nr, nc), order="F")
                for jl in range(nc):
                    for il in range(nr):
                        rows.append(r0 + il)
                        cols.append(c0 + jl)
                        vals.append(blk[il, jl])
        return sp.csc_matrix((vals, (rows, cols)), shape=self.shape)


def blocksparse(I, J, V, R=None, C=None):
    """blocksparse(I, J, V[, R, C]) (block_sparse_matrix.jl:330-359): block (I[k], J[k]) (0-based
    out-vertex and vertex ids) is the dense matrix V[k]; blocks at the same position are summed,
    as sparse(I, J, V) sums them. R / C: rows per out vertex and columns per vertex (default: from
    the blocks; a vertex or out vertex without a block then has none)."""
    I = [int(x) for x in I]
    J = [int(x) for x in J]
    V = [np.atleast_2d(np.asarray(x, dtype=np.float64)) for x in V]
    nout = len(R) if R is not None else (max(I) + 1 if I else 0)
    nvtx = len(C) if C is not None else (max(J) + 1 if J else 0)
    rsz = list(R) if R is not None else [0] * nout
    csz = list(C) if C is not None else [0] * nvtx
    if R is None or C is None:
        for i, j, b in zip(I, J, V):
            if R is None:
                rsz[i] = b.shape[0]
            if C is None:
                csz[j] = b.shape[1]
    blocks = {}
    for i, j, b in zip(I, J, V):
        if b.shape != (rsz[i], csz[j]):
            raise ValueError(f"blocksparse: block ({i}, {j}) has shape {b.shape}, expected {(rsz[i], csz[j])}")
        blocks[(j, i)] = blocks[(j, i)] + b if (j, i) in blocks else b.copy()
    xcol = np.concatenate([[0], np.cumsum(csz)]).astype(np.int64)
    xrow = np.concatenate([[0], np.cumsum(rsz)]).astype(np.int64)
    xsrc, tgt, xblk, val = [0], [], [0], []
    for v in range(nvtx):
        for u in sorted(i for (j, i) in blocks if j == v):
            b = blocks[(v, u)]
            tgt.append(u)
            val.append(b.ravel(order="F"))
            xblk.append(xblk[-1] + b.size)
        xsrc.append(len(tgt))
    return BlockSparse(nout, nvtx, xsrc, xcol, xrow, xblk, tgt,
                       np.concatenate(val) if val else np.zeros(0))


class IPMProblem:
    """IPMProblem(Q, B, f, g, μ, K, s; compress=1.0) (problem.jl:64-127).

    Q (n×n) and B (m×n) are scipy sparse matrices (stored zeros are kept); f, g dense; K a list
    of cone objects with block sizes s (Σs = n)."""

    def __init__(self, Q, B, f, g, mu, K, s, compress=1.0):
        m, n = B.shape
        self.n, self.m = int(n), int(m)
        self._Q = _csc(Q, n, n)
        self._B = _csc(B, m, n)
        self._f = np.ascontiguousarray(f, dtype=np.float64)
        self._g = np.ascontiguousarray(g, dtype=np.float64)
        if self._f.shape != (n,) or self._g.shape != (m,):
            raise ValueError("IPMProblem: f and g must have lengths n and m")
        kinds = np.array([k.kind for k in K], dtype=np.int32)
        params = np.array([k.param for k in K], dtype=np.float64)
        sizes = np.array(s, dtype=np.int64)
        self.K = list(K)
        h = lib().mb_problem_create(n, m, _p(self._Q[0], ctypes.c_int64), _p(self._Q[1], ctypes.c_int64),
                                    _p(self._Q[2], ctypes.c_double), _p(self._B[0], ctypes.c_int64),
                                    _p(self._B[1], ctypes.c_int64), _p(self._B[2], ctypes.c_double),
                                    _p(self._f, ctypes.c_double), _p(self._g, ctypes.c_double), float(mu),
                                    len(K), _p(kinds, ctypes.c_int32), _p(params, ctypes.c_double),
                                    _p(sizes, ctypes.c_int64), float(compress))
        if not h:
            raise _err("IPMProblem")
        self._h = ctypes.c_void_p(h)
        _live_problems.add(self)

    @classmethod
    def from_blocks(cls, Q, B, f, g, mu, K):
        """IPMProblem(Q::BlockSparseMatrix, B::BlockSparseMatrix, f, g, μ, K) (problem.jl:49-53): the
        block-sparse Q and B (BlockSparse, e.g. from blocksparse) as they are, one cone per vertex,
        identity row and column permutations."""
        self = cls.__new__(cls)
        m, n = B.shape
        self.n, self.m = int(n), int(m)
        if Q.nvtx != B.nvtx or not np.array_equal(Q.xcol, B.xcol) or not np.array_equal(Q.xrow, B.xcol):
            raise ValueError("IPMProblem: Q must be square over B's vertex partition")
        if len(K) != B.nvtx:
            raise ValueError("IPMProblem: one cone per vertex")
        self._f = np.ascontiguousarray(f, dtype=np.float64)
        self._g = np.ascontiguousarray(g, dtype=np.float64)
        if self._f.shape != (n,) or self._g.shape != (m,):
            raise ValueError("IPMProblem: f and g must have lengths n and m")
        self._blocks = (Q, B)
        self.K = list(K)
        kinds = np.array([k.kind for k in K], dtype=np.int32)
        params = np.array([k.param for k in K], dtype=np.float64)
        i64, dbl = ctypes.c_int64, ctypes.c_double
        h = lib().mb_problem_create_bsm(B.nvtx, B.nout, _p(B.xcol, i64), _p(B.xrow, i64), _p(B.xsrc, i64),
                                        _p(B.tgt, i64), _p(B.xblk, i64), _p(B.val, dbl), _p(Q.xsrc, i64),
                                        _p(Q.tgt, i64), _p(Q.xblk, i64), _p(Q.val, dbl), _p(self._f, dbl),
                                        _p(self._g, dbl), float(mu), _p(kinds, ctypes.c_int32),
                                        _p(params, dbl), 0)
        if not h:
            raise _err("IPMProblem")
        self._h = ctypes.c_void_p(h)
        _live_problems.add(self)
        return self

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.mb_problem_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def dims(self):
        d = np.zeros(7, dtype=np.int64)
        if lib().mb_problem_dims(self._h, _p(d, ctypes.c_int64)) != 0:
            raise _err("dims")
        return dict(nvtx=int(d[0]), nout=int(d[1]), narcB=int(d[2]), nbnzB=int(d[3]), narcQ=int(d[4]),
                    nbnzQ=int(d[5]), n=int(d[6]))

    def export(self):
        """The constructed block-sparse problem (0-based), as plain arrays."""
        d = self.dims()
        N, U = d["nvtx"], d["nout"]
        out = dict(P1=np.zeros(self.m, np.int64), P2=np.zeros(self.n, np.int64), kind=np.zeros(N, np.int32),
                   B_xsrc=np.zeros(N + 1, np.int64), B_xcol=np.zeros(N + 1, np.int64),
                   B_xrow=np.zeros(U + 1, np.int64), B_xblk=np.zeros(d["narcB"] + 1, np.int64),
                   B_tgt=np.zeros(d["narcB"], np.int64), B_val=np.zeros(d["nbnzB"]),
                   Q_xsrc=np.zeros(N + 1, np.int64), Q_xblk=np.zeros(d["narcQ"] + 1, np.int64),
                   Q_tgt=np.zeros(d["narcQ"], np.int64), Q_val=np.zeros(d["nbnzQ"]), f=np.zeros(self.n),
                   g=np.zeros(self.m))
        i = ctypes.c_int64
        r = lib().mb_problem_export(self._h, _p(out["P1"], i), _p(out["P2"], i), _p(out["kind"], ctypes.c_int32),
                                    _p(out["B_xsrc"], i), _p(out["B_xcol"], i), _p(out["B_xrow"], i),
                                    _p(out["B_xblk"], i), _p(out["B_tgt"], i), _p(out["B_val"], ctypes.c_double),
                                    _p(out["Q_xsrc"], i), _p(out["Q_xblk"], i), _p(out["Q_tgt"], i),
                                    _p(out["Q_val"], ctypes.c_double), _p(out["f"], ctypes.c_double),
                                    _p(out["g"], ctypes.c_double))
        if r != 0:
            raise _err("export")
        return out

    def symbolic(self, amalgamate=True):
        """The build's symbolic analysis (ordering, fronts) of this problem's KKT graph."""
        d = np.zeros(3, dtype=np.int64)
        if lib().mb_problem_symbolic_dims(self._h, int(amalgamate), _p(d, ctypes.c_int64)) != 0:
            raise _err("symbolic")
        N, F, S = (int(x) for x in d)
        vperm = np.zeros(N, np.int64)
        fv = np.zeros(F + 1, np.int64)
        sp = np.zeros(F + 1, np.int64)
        sep = np.zeros(max(S, 1), np.int64)
        par = np.zeros(max(F, 1), np.int64)
        fl = np.zeros(2)
        i = ctypes.c_int64
        if lib().mb_problem_symbolic(self._h, int(amalgamate), _p(vperm, i), _p(fv, i), _p(sp, i), _p(sep, i),
                                     _p(par, i), _p(fl, ctypes.c_double)) != 0:
            raise _err("symbolic")
        return dict(vperm=vperm, front_vtx=fv, sep_ptr=sp, sep=sep[:S], parent=par[:F], exact_flops=fl[0],
                    front_flops=fl[1])


HIST_FIELDS = ["pres", "dres", "pobj", "dobj", "step", "piter", "ppass", "pstat", "citer", "cpass", "cstat",
               "dmin", "dmax", "rho", "delta", "mu", "chi"]


@dataclass
class IPMResult:
    """IPMResult (result/ipm.jl:1-15)."""
    p: np.ndarray
    d: np.ndarray
    y: np.ndarray
    status: IPMStatus
    niter: int
    nsolve: int
    history: list = field(repr=False)
    mu: float
    pres: float
    dres: float
    pobj: float
    dobj: float


class IPMSolver:
    """IPMSolver (solver/ipm.jl:1-28): device-resident iterate, scaling, factor."""

    def __init__(self, prob, settings=None, p0=None, d0=None, y0=None):
        settings = settings or IPMSettings()
        self.prob = prob
        self.settings = settings
        cs = settings._c()
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (p0, d0, y0)]
        h = lib().mb_solver_create(prob._h, ctypes.byref(cs), *(_p(a, ctypes.c_double) for a in arr))
        if not h:
            raise _err("init")
        self._h = ctypes.c_void_p(h)
        _live_solvers.add(self)

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.mb_solver_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def step(self):
        r = lib().mb_solver_step(self._h)
        if r < 0:
            raise _err("step!")
        return IPMStatus(r)

    def solve(self):
        r = lib().mb_solver_solve(self._h)
        if r < 0:
            raise _err("solve!")
        return self.result()

    def history(self):
        n = int(lib().mb_solver_niter(self._h))
        rows = np.zeros((max(n, 1), 17))
        if lib().mb_solver_history(self._h, _p(rows, ctypes.c_double)) != 0:
            raise _err("history")
        out = []
        for k in range(n):
            row = dict(zip(HIST_FIELDS, rows[k].tolist()))
            for key in ("piter", "ppass", "pstat", "citer", "cpass", "cstat"):
                row[key] = int(row[key])
            out.append(row)
        return out

    def stats(self):
        st = np.zeros(19)
        if lib().mb_solver_stats(self._h, _p(st, ctypes.c_double)) != 0:
            raise _err("stats")
        return dict(exact_flops=st[0], front_flops=st[1], nfront=int(st[2]), nlevel=int(st[3]),
                    factor_ms_total=st[4], factor_calls=int(st[5]), factor_ms_last=st[6],
                    factor_calls_fp32=int(st[7]), init_host_ms=st[8], init_device_ms=st[9],
                    init_device_parts_ms=dict(stream=st[10], buffers=st[11], schedule=st[12], start=st[13]),
                    build_marks_ms=dict(fronts=st[14], union=st[15], cones=st[16], allocs=st[17], sparse=st[18]))

    def reinit(self, p0=None, d0=None, y0=None):
        """reinit!(solver; p0, d0, y0) (solver/ipm.jl:371-429)."""
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (p0, d0, y0)]
        if lib().mb_solver_reinit(self._h, *(_p(a, ctypes.c_double) for a in arr)) != 0:
            raise _err("reinit!")

    PROF_KINDS = ["factor", "update", "panel", "fsolve", "bsolve", "extend_add", "dag"]

    def set_profile(self, kinds, trace=False):
        mask = 0
        for k in kinds:
            mask |= 1 << self.PROF_KINDS.index(k)
        if trace:
            mask |= 1 << 16
        if lib().mb_solver_set_profile(self._h, mask) != 0:
            raise _err("profile")

    def dag_trace(self, front=0):
        """Task timeline of the last persistent factorization of large front `front`: an
        (ntask, 8) array of 100 MHz wall-clock stamps in schedule order (claimed, updates done,
        last input ready, done; a diagonal task also: factor panels done, diagonal-block
        inverses done, inverse complete, inverse published; 0 where a task has none)."""
        n = lib().mb_solver_dag_trace(self._h, int(front), None, 0)
        if n < 0:
            raise _err("trace")
        out = np.zeros(max(n, 0), dtype=np.int64)
        if n > 0 and lib().mb_solver_dag_trace(self._h, int(front), _p(out, ctypes.c_int64), n) < 0:
            raise _err("trace")
        return out.reshape(-1, 8)

    def profile(self):
        out = np.zeros(4 * len(self.PROF_KINDS))
        if lib().mb_solver_profile(self._h, _p(out, ctypes.c_double), len(out)) < 0:
            raise _err("profile")
        return {k: dict(ms=out[4 * i], launches=int(out[4 * i + 1]), flops=out[4 * i + 2], bytes=out[4 * i + 3])
                for i, k in enumerate(self.PROF_KINDS)}

    def kkt_factor(self, delta):
        rho = ctypes.c_double(0.0)
        r = lib().mb_solver_kkt_factor(self._h, float(delta), ctypes.byref(rho))
        if r < 0:
            raise _err("initkkt!")
        return bool(r), rho.value

    def frule(self, df, dg, dmu=0.0):
        """frule!(Δp, Δy, Δd, s, Δf, Δg, Δμ) (differential.jl:5-75) at the last iterate, in the
        problem's space: returns (Δp, Δy, Δd)."""
        n, m = self.prob.n, self.prob.m
        df = np.ascontiguousarray(df, dtype=np.float64)
        dg = np.ascontiguousarray(dg, dtype=np.float64)
        if df.shape != (n,) or dg.shape != (m,):
            raise ValueError("frule: Δf and Δg must have lengths n and m")
        dp, dy, dd = np.zeros(n), np.zeros(m), np.zeros(n)
        if lib().mb_solver_frule(self._h, _p(df, ctypes.c_double), _p(dg, ctypes.c_double), float(dmu),
                                 _p(dp, ctypes.c_double), _p(dy, ctypes.c_double), _p(dd, ctypes.c_double)) != 0:
            raise _err("frule!")
        return dp, dy, dd

    def rrule(self, dp, dy, dd):
        """rrule!(Δf, Δg, s, Δp, Δy, Δd) (differential.jl:252-334): returns (Δf, Δg, dμ)."""
        n, m = self.prob.n, self.prob.m
        dp, dy, dd = (np.ascontiguousarray(a, dtype=np.float64) for a in (dp, dy, dd))
        if dp.shape != (n,) or dd.shape != (n,) or dy.shape != (m,):
            raise ValueError("rrule: Δp, Δd must have length n and Δy length m")
        df, dg, dmu = np.zeros(n), np.zeros(m), ctypes.c_double(0.0)
        if lib().mb_solver_rrule(self._h, _p(dp, ctypes.c_double), _p(dy, ctypes.c_double), _p(dd, ctypes.c_double),
                                 _p(df, ctypes.c_double), _p(dg, ctypes.c_double), ctypes.byref(dmu)) != 0:
            raise _err("rrule!")
        return df, dg, dmu.value

    def kkt_factor_h(self, H, delta):
        """initkkt!(kkt, H; δ) (uzawa.jl:73-76) with the caller's H: a BlockSparse over the internal
        vertex partition (layout()["xcol"]) in internal coordinates; kkt_solve then uses it."""
        rho = ctypes.c_double(0.0)
        i64 = ctypes.c_int64
        r = lib().mb_solver_kkt_factor_h(self._h, _p(H.xsrc, i64), _p(H.tgt, i64), _p(H.xblk, i64),
                                         _p(H.val, ctypes.c_double), 0, float(delta), ctypes.byref(rho))
        if r < 0:
            raise _err("initkkt!")
        return bool(r), rho.value

    # -- cone fan-outs (cone.jl:146-553) on vectors in the solver's internal coordinates --

    def layout(self):
        """Internal row and column order and cones: P1 (internal → caller row), P2 (internal →
        caller column), xcol, kind, param, rowscale and colscale (the equilibration factors)."""
        n, m, N = self.prob.n, self.prob.m, int(lib().mb_solver_ncone(self._h))
        P1, P2, xcol = np.zeros(m, np.int64), np.zeros(n, np.int64), np.zeros(N + 1, np.int64)
        kind, param, rs, cs = np.zeros(N, np.int32), np.zeros(N), np.zeros(m), np.zeros(n)
        i = ctypes.c_int64
        if lib().mb_solver_layout(self._h, _p(P1, i), _p(P2, i), _p(xcol, i), _p(kind, ctypes.c_int32),
                                  _p(param, ctypes.c_double), _p(rs, ctypes.c_double),
                                  _p(cs, ctypes.c_double)) != 0:
            raise _err("layout")
        return dict(P1=P1, P2=P2, xcol=xcol, kind=kind, param=param, rowscale=rs, colscale=cs)

    def kkt_solve(self, f, g, x0=None, y0=None, ftol=1e-10, gtol=1e-10, irmin=0, use_stop=False):
        """solvekkt! (uzawa.jl:265-344) on the current H and factor: H x − Bᵀ y = f, B x = g in
        internal coordinates. Returns (x, y, ncg, nir, KKTStatus, δmin, δmax)."""
        f = self._vec(f)
        g = np.ascontiguousarray(g, dtype=np.float64)
        if g.shape != (self.prob.m,):
            raise ValueError("kkt_solve: g must have length m")
        warm = x0 is not None or y0 is not None
        x = np.array(x0, dtype=np.float64) if x0 is not None else np.zeros(self.prob.n)
        y = np.array(y0, dtype=np.float64) if y0 is not None else np.zeros(self.prob.m)
        counts, window = np.zeros(3, np.int32), np.zeros(2)
        if lib().mb_solver_kkt_solve(self._h, _p(x, ctypes.c_double), _p(y, ctypes.c_double),
                                     _p(f, ctypes.c_double), _p(g, ctypes.c_double), int(warm), float(ftol),
                                     float(gtol), int(irmin), int(bool(use_stop)), _p(counts, ctypes.c_int32),
                                     _p(window, ctypes.c_double)) != 0:
            raise _err("solvekkt!")
        return x, y, int(counts[0]), int(counts[1]), KKTStatus(int(counts[2])), window[0], window[1]

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.prob.n,):
            raise ValueError("cone fan-out vectors must have length n")
        return x

    def cone_scale(self, p, d):
        """scale! (cone.jl:146-247): sets the iterate to (p, d); returns (all flags, Σ spsd)."""
        p, d = self._vec(p), self._vec(d)
        spsd = ctypes.c_double(0.0)
        r = lib().mb_cone_scale(self._h, _p(p, ctypes.c_double), _p(d, ctypes.c_double), ctypes.byref(spsd))
        if r < 0:
            raise _err("scale!")
        return bool(r), spsd.value

    def cone_blocks(self):
        """The scaling blocks in cone order: diagonal (nonneg), nothing (free), n_v² column-major."""
        out = np.zeros(int(lib().mb_cone_blocks_size(self._h)))
        if lib().mb_cone_blocks(self._h, _p(out, ctypes.c_double)) != 0:
            raise _err("cone blocks")
        return out

    def cone_corr0(self, sigma_mu):
        """initpredictor! right-hand side (cone.jl:361-444) at the scaled iterate."""
        r = np.zeros(self.prob.n)
        if lib().mb_cone_corr0(self._h, float(sigma_mu), _p(r, ctypes.c_double)) != 0:
            raise _err("corr0!")
        return r

    def cone_corr(self, dp, dd, sigma_mu):
        """initcorrector! right-hand side (cone.jl:261-352) for directions (dp, dd)."""
        dp, dd = self._vec(dp), self._vec(dd)
        r = np.zeros(self.prob.n)
        if lib().mb_cone_corr(self._h, _p(dp, ctypes.c_double), _p(dd, ctypes.c_double), float(sigma_mu),
                              _p(r, ctypes.c_double)) != 0:
            raise _err("corr!")
        return r

    def cone_maxsteps(self, dp, dd, step_frac=1.0):
        """maxsteps (cone.jl:455-553): the largest step keeping both iterates interior."""
        dp, dd = self._vec(dp), self._vec(dd)
        st = ctypes.c_double(0.0)
        if lib().mb_cone_maxsteps(self._h, _p(dp, ctypes.c_double), _p(dd, ctypes.c_double), float(step_frac),
                                  ctypes.byref(st)) != 0:
            raise _err("maxsteps")
        return st.value

    def dense_factor(self):
        n = self.prob.n
        L = np.zeros((n, n), order="F")
        if lib().mb_solver_dense_factor(self._h, L.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) != 0:
            raise _err("dense_factor")
        return L

    def result(self):
        n, m = self.prob.n, self.prob.m
        p, d, y, info = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(8)
        if lib().mb_solver_result(self._h, _p(p, ctypes.c_double), _p(d, ctypes.c_double), _p(y, ctypes.c_double),
                                  _p(info, ctypes.c_double)) != 0:
            raise _err("result")
        return IPMResult(p, d, y, IPMStatus(int(info[7])), int(info[5]), int(info[6]), self.history(),
                         info[0], info[1], info[2], info[3], info[4])


def init(prob, settings=None, p0=None, d0=None, y0=None):
    """CommonSolve.init(prob, settings; p0, d0, y0) (solver/ipm.jl:485-491)."""
    return IPMSolver(prob, settings, p0, d0, y0)


def step(solver):
    """step!(solver) (solver/ipm.jl:577-774)."""
    return solver.step()


def solve(prob, settings=None, p0=None, d0=None, y0=None):
    """CommonSolve.solve(prob, settings; p0, d0, y0) (solver/ipm.jl:798-804)."""
    return init(prob, settings, p0, d0, y0).solve()


class IPMServer:
    """Continuous-batching request loop on one GPU (no counterpart in the reference; each request
    is a solve(prob, settings), solver/solver.jl:119-150). `nslots` slots, each a host thread with
    its own HIP stream, take requests in arrival order and admit them as slots and device memory
    (`mem_budget_gb`, default 90 % of the free memory) free up; their kernels overlap on the GPU.
    `dag_grid` workgroups per slot for the persistent factorization (0: the CUs split evenly).
    With `gang`, requests are routed by the flops of their factorization to one gang (IPMGang) of
    `nslots` members per size class (below 1e8, below 2e9, above): members advance in lockstep
    and share every launch."""

    def __init__(self, nslots=8, mem_budget_gb=0.0, dag_grid=0, gang=False):
        if gang:  # one gang per size class, nslots members each
            h = lib().mb_server_create_gang(int(nslots), float(mem_budget_gb) * 2.0 ** 30)
        else:
            h = lib().mb_server_create(int(nslots), float(mem_budget_gb) * 2.0 ** 30, int(dag_grid))
        if not h:
            raise _err("server")
        self._h = ctypes.c_void_p(h)
        self._probs = {}
        _live_servers.add(self)

    def submit(self, prob, settings=None):
        """Queue a solve; returns its ticket. The problem is kept alive until collected."""
        cs = (settings or IPMSettings())._c()
        t = lib().mb_server_submit(self._h, prob._h, ctypes.byref(cs))
        if t < 0:
            raise _err("submit")
        self._probs[int(t)] = prob
        return int(t)

    def wait(self, ticket, timeout=None):
        """Block until the ticket's solve ends (or `timeout` seconds pass: returns None). Returns
        (IPMResult without history, dict of submitted/admitted/finished clock times, device
        bytes allocated and device bytes reserved at admission)."""
        prob = self._probs[ticket]
        n, m = prob.n, prob.m
        p, d, y, info, tm = np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(8), np.zeros(5)
        r = lib().mb_server_wait(self._h, int(ticket), -1.0 if timeout is None else float(timeout),
                                 _p(p, ctypes.c_double), _p(d, ctypes.c_double), _p(y, ctypes.c_double),
                                 _p(info, ctypes.c_double), _p(tm, ctypes.c_double))
        if r == -2:
            return None
        if r < 0:
            err = _err("serve")
            self.release(ticket)
            raise err
        self.release(ticket)
        res = IPMResult(p, d, y, IPMStatus(int(info[7])), int(info[5]), int(info[6]), [],
                        info[0], info[1], info[2], info[3], info[4])
        return res, dict(submitted=tm[0], admitted=tm[1], finished=tm[2], device_bytes=tm[3], reserved_bytes=tm[4])

    def release(self, ticket):
        lib().mb_server_release(self._h, int(ticket))
        self._probs.pop(int(ticket), None)

    def stats(self):
        out = np.zeros(5)
        if lib().mb_server_stats(self._h, _p(out, ctypes.c_double)) != 0:
            raise _err("stats")
        return dict(queued=int(out[0]), running=int(out[1]), completed=int(out[2]), reserved_bytes=out[3],
                    peak_reserved_bytes=out[4])

    def close(self):
        if getattr(self, "_h", None):
            lib().mb_server_destroy(self._h)
            self._h = None
            self._probs.clear()


class IPMGang:
    """Lockstep batching of independent solves (no counterpart in the reference; each member is a
    solve!(solver), solver/solver.jl:119-150). Up to `capacity` solvers advance together, each
    taking the reference's decisions for its own instance, while every phase of their IPM steps
    is one launch for all of them; results equal the solvers' solo results bit for bit."""

    def __init__(self, capacity=8, dag_grid=0, profile=False):
        h = lib().mb_gang_create(int(capacity), int(dag_grid), 1 if profile else 0)
        if not h:
            raise _err("gang")
        self._h = ctypes.c_void_p(h)
        self.capacity = int(capacity)
        _live_servers.add(self)

    def solve_in(self, solver):
        """solve!(solver) as a member (call from one thread per solver); returns the IPMStatus."""
        r = lib().mb_solver_solve_in(solver._h, self._h)
        if r < 0:
            raise _err("solve! in gang")
        return IPMStatus(r)

    def solve(self, solvers):
        """Solve the solvers through the gang (library threads, a finished solver's slot taken by
        the next); returns their IPMResults."""
        n = len(solvers)
        hs = (ctypes.c_void_p * max(n, 1))(*[s._h.value for s in solvers])
        st = np.zeros(max(n, 1), dtype=np.int32)
        if lib().mb_solve_batch(hs, n, self._h, _p(st, ctypes.c_int32)) != 0:
            raise _err("batch solve")
        return [s.result() for s in solvers]

    def stats(self):
        out = np.zeros(18)
        if lib().mb_gang_stats(self._h, _p(out, ctypes.c_double)) != 0:
            raise _err("gang stats")
        keys = ("ticks", "launches", "ops", "combined_ops", "readbacks", "joins", "factor_launches",
                "factor_fronts", "factor_flops", "factor_ms", "dag_launches", "dag_flops", "dag_ms", "members",
                "admit_ms", "fiber_ms", "tick_ms", "wait_ms")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "_h", None):
            lib().mb_gang_destroy(self._h)
            self._h = None


def clock():
    """Seconds on the steady clock of the server's timestamps."""
    return lib().mb_clock()


def device_info():
    """Compute units, free and total bytes of the current device."""
    out = np.zeros(3)
    if lib().mb_device_info(_p(out, ctypes.c_double)) != 0:
        raise _err("device_info")
    return dict(cus=int(out[0]), free_bytes=out[1], total_bytes=out[2])


def cached_bytes():
    """Device bytes held in the library's cache of freed solver buffers."""
    return int(lib().mb_cached_bytes())


def release_cached_memory():
    """Return the cached device memory to the HIP runtime; returns the bytes released."""
    return int(lib().mb_release_cached_memory())


def dense_cholesky(A_ptr, n, lda, grid=0, stream=None):
    """Factor the dense n×n front at device address A_ptr (column-major, leading dimension lda)
    in place: its lower triangle becomes L with A = L Lᵀ. Returns True on success, False when a
    pivot was not positive. The frontal POTRF inside CliqueTrees' cholesky! (uzawa.jl:93, 111)."""
    r = lib().mb_dense_cholesky(ctypes.c_void_p(int(A_ptr)), int(n), int(lda), int(grid),
                                ctypes.c_void_p(int(stream) if stream else 0))
    if r < 0:
        raise _err("dense_cholesky")
    return r == 0


def dense_cholesky_noalloc(A_ptr, n, lda, grid=0, stream=None):
    """dense_cholesky without any device allocation (workspace: a per-GPU device array, n ≤ 8192;
    calls overlapping on one GPU race)."""
    r = lib().mb_dense_cholesky_noalloc(ctypes.c_void_p(int(A_ptr)), int(n), int(lda), int(grid),
                                        ctypes.c_void_p(int(stream) if stream else 0))
    if r < 0:
        raise _err("dense_cholesky_noalloc")
    return r == 0


def probe_mfma_f64():
    out = np.zeros(256)
    if lib().mb_probe_mfma_f64(_p(out, ctypes.c_double)) != 0:
        raise _err("probe")
    return out


def probe_mfma_f32():
    out = np.zeros(256)
    if lib().mb_probe_mfma_f32(_p(out, ctypes.c_double)) != 0:
        raise _err("probe")
    return out
