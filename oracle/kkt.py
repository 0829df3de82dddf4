"""KKT solver (oracle restatement of src/IPM/src/kkt/{kkt,uzawa,cg}.jl).

The reference factors δH + BᵀB with CliqueTrees' supernodal multifrontal Cholesky
(uzawa.jl:93, 111) — an external package whose source is not available here. The oracle
factors the same matrix densely with LAPACK (numpy.linalg.cholesky): the exact factor is
the same mathematical object, so the factor is pinned by LLᵀ = M to rounding, not by the
reference's accumulation order (which is unpinned). The Uzawa / CG control logic is restated
line by line: stopping rules, stall test, irmin/irmax, cgmax, statuses, kktwindow.

Pinned by (tests/test_oracle_cones_kkt.py): the returned (x, y) satisfy the KKT system to the
requested tolerances on random SPD inputs; against a direct dense solve of the saddle-point
system; the status/counter semantics on forced cases (irmax=0, cgmax=0, stall=1 stagnation,
failed factorization → ρ retry).
"""
import math

import numpy as np

KKT_CONTINUE, KKT_SOLVED, KKT_STAGNATED, KKT_ITMAX, KKT_NUMERICAL_FAILURE, KKT_IRMAX = range(6)
CG_CONTINUE, CG_SOLVED, CG_NUMERICAL_FAILURE, CG_ITMAX = range(4)
EPS = float(np.finfo(np.float64).eps)


class UzawaSolver:
    """UzawaSolver (uzawa.jl:3-30): factor of δH + BᵀB plus refinement workspaces."""

    def __init__(self, B, irmax=10, cgmax=100):
        self.B = B                        # oracle BSM (already permuted / scaled)
        self.Bd = B.to_dense()
        self.L = self.Bd.T @ self.Bd      # uzawa.jl:29 — L = B' * B
        self.F = None
        self.delta = 1.0
        self.hist = np.zeros(irmax + 2)
        self.cghist = np.zeros(cgmax + 1)

    # uzawa.jl:73-115
    def init(self, Hd, delta):
        self.delta = delta
        M = self.L + delta * Hd
        rho = 0.0
        try:
            F = np.linalg.cholesky(M)
            ok = bool(np.all(np.isfinite(F)))
        except np.linalg.LinAlgError:
            ok = False
        if not ok:
            rho = EPS * float(np.linalg.norm(np.tril(M)))
            M = M + rho * np.eye(M.shape[0])
            try:
                F = np.linalg.cholesky(M)
                ok = bool(np.all(np.isfinite(F)))
            except np.linalg.LinAlgError:
                ok = False
        self.F = F if ok else None
        return ok, rho

    def Lsolve(self, v):
        import scipy.linalg as sl
        return sl.solve_triangular(self.F, v, lower=True, check_finite=False)

    def Usolve(self, v):
        import scipy.linalg as sl
        return sl.solve_triangular(self.F, v, lower=True, trans="T", check_finite=False)


def mulkkt(Hd, Bd, x, y):
    """[u; v] = [H −Bᵀ; B 0][x; y] (kkt.jl:9-21)."""
    u = Hd @ x - Bd.T @ y
    v = Bd @ x
    return u, v


def kktwindow(rhist, chist, delta, nir, ncg, ftol, gtol):
    """uzawa.jl:199-227 (1-based rhist[2] → rhist[1] here)."""
    dmin = dmax = math.nan
    if nir >= 1:
        logd = math.log(delta)
        logfres = math.log(rhist[1]) if rhist[1] > 0 else -math.inf
        loggtol = math.log(gtol) if gtol > 0 else -math.inf
        logftol = math.log(ftol) if ftol > 0 else -math.inf
        i = 0
        logdmax = -math.inf
        logdmin = logd + logfres - logftol
        while logdmin > logdmax and i < ncg + 1:
            i += 1
            loggres = math.log(chist[i - 1]) if chist[i - 1] > 0 else -math.inf
            logdmax = logd + (loggtol - loggres) / i
        dmin = math.exp(logdmin) if not math.isnan(logdmin) else math.nan
        dmax = math.exp(logdmax) if not math.isnan(logdmax) else math.nan
    return dmin, dmax


def cg(stop, solver, Bd, x, y, u, atol, itmax):
    """Dual CG on B(δH+BᵀB)⁻¹Bᵀ (cg.jl:69-167). Mutates x, y, u in place."""
    hist = solver.cghist
    niter = 0
    nu2 = float(u @ u)
    nu = math.sqrt(nu2)
    hist[0] = nu
    np2 = nu2
    y[:] = 0.0
    status = CG_SOLVED
    if nu > atol and not stop(u):
        status = CG_ITMAX
        if itmax > 0:
            p = u.copy()
            status = CG_CONTINUE
            while status == CG_CONTINUE:
                niter += 1
                w = solver.Lsolve(Bd.T @ p)
                ep2 = float(w @ w)
                if ep2 <= EPS * np2:
                    status = CG_NUMERICAL_FAILURE
                else:
                    w = solver.Usolve(w)
                    alpha = nu2 / ep2
                    y += alpha * p
                    x += alpha * w
                    u -= alpha * (Bd @ w)
                    pnu2 = nu2
                    nu2 = float(u @ u)
                    nu = math.sqrt(nu2)
                    hist[niter] = nu
                    if nu <= atol or 1.0 + nu <= 1.0 or stop(u):
                        status = CG_SOLVED
                    elif niter >= itmax:
                        status = CG_ITMAX
                    else:
                        beta = nu2 / pnu2
                        p = u + beta * p
                        np2 = nu2 + beta ** 2 * np2
    return niter, status


def solvekkt(solver, Hd, Bd, x, y, f, g, warm, ftol, gtol, stall, irmax, cgmax, irmin=0,
             fstop=lambda v: False, gstop=lambda v: False):
    """Uzawa iterative refinement + dual CG (uzawa.jl:319-507). Mutates x, y.
    Returns (ncg, nir, status, dmin, dmax)."""
    delta = solver.delta
    alpha = 1.0 / delta
    nir = ncg = 0
    hist = solver.hist
    if warm:
        df, dg = mulkkt(Hd, Bd, x, y)
        df = f - df
        dg = g - dg
    else:
        df = f.copy()
        dg = g.copy()
        x[:] = 0.0
        y[:] = 0.0
    fprv = fres = float(np.linalg.norm(df))
    hist[0] = fres
    if nir >= irmin and (fres <= ftol or fstop(df)):
        irstat = KKT_SOLVED
    elif nir >= irmax:
        irstat = KKT_IRMAX
    else:
        irstat = KKT_CONTINUE
    while irstat == KKT_CONTINUE:
        nir += 1
        dx = df * delta + Bd.T @ dg
        dx = solver.Usolve(solver.Lsolve(dx))
        dg = dg - Bd @ dx
        x += dx
        y += alpha * dg
        df, dg = mulkkt(Hd, Bd, x, y)
        df = f - df
        dg = g - dg
        fres = float(np.linalg.norm(df))
        hist[nir] = fres
        if nir >= irmin and (fres <= ftol or fstop(df)):
            irstat = KKT_SOLVED
        elif nir > 1 and fres > (1 - stall) * fprv:
            irstat = KKT_STAGNATED
        elif nir >= irmax:
            irstat = KKT_IRMAX
        else:
            irstat = KKT_CONTINUE
        fprv = fres
    y -= alpha * dg
    dy = np.zeros_like(y)
    ncg, cgstat = cg(gstop, solver, Bd, x, dy, dg, gtol, cgmax)
    y += alpha * dy
    y += alpha * dg
    if cgstat == CG_NUMERICAL_FAILURE:
        status = KKT_NUMERICAL_FAILURE
    elif irstat == KKT_STAGNATED:
        status = KKT_STAGNATED
    elif irstat == KKT_IRMAX or cgstat == CG_ITMAX:
        status = KKT_ITMAX
    else:
        status = KKT_SOLVED
    dmin, dmax = kktwindow(hist, solver.cghist, delta, nir, ncg, ftol, gtol)
    return ncg, nir, status, dmin, dmax
