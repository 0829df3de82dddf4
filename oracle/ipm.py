"""IPM driver (oracle restatement of src/IPM/src/ipm/solver/{ipm,solver,utils}.jl,
history/history.jl, settings/ipm.jl).

Pinned by (tests/test_oracle_ipm.py): solver-independent optimality certificates at 1e-8
(primal/dual residuals, gap, cone membership) on seeded QPs and SOCPs; LP optima against
scipy's HiGHS; the closed-form optimum of small equality QPs; forced failure paths
(max_iter=1 → ITERATION_LIMIT, infeasible Bp=g, p≥0 → PRIMAL_INFEASIBLE, unbounded
objective → DUAL_INFEASIBLE). The KKT status semantics are pinned at the solvekkt level
(tests/test_oracle_cones_kkt.py).
"""
import math
from dataclasses import dataclass

import numpy as np

from . import cones as C
from .bsm import selectvtxs, symselectvtxs
from .kkt import (KKT_SOLVED, UzawaSolver, solvekkt)
from .scaling import equilibrate
from .symbolic import analyse, vertex_graph

(CONTINUE, OPTIMAL, NEAR_OPTIMAL, STALLED, NUMERICAL_FAILURE, ITERATION_LIMIT,
 PRIMAL_INFEASIBLE, DUAL_INFEASIBLE, ILL_POSED, NEAR_PRIMAL_INFEASIBLE,
 NEAR_DUAL_INFEASIBLE, NEAR_ILL_POSED) = range(12)
STATUS_NAMES = ["CONTINUE", "OPTIMAL", "NEAR_OPTIMAL", "STALLED", "NUMERICAL_FAILURE",
                "ITERATION_LIMIT", "PRIMAL_INFEASIBLE", "DUAL_INFEASIBLE", "ILL_POSED",
                "NEAR_PRIMAL_INFEASIBLE", "NEAR_DUAL_INFEASIBLE", "NEAR_ILL_POSED"]
FORCING_FRAC = 0.1  # solver/solver.jl:3
EPS = float(np.finfo(np.float64).eps)


@dataclass
class Settings:
    """IPMSettings defaults (settings/ipm.jl:1-18)."""
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


def scalenorm(v, s):
    """solver/utils.jl:11-19."""
    return math.sqrt(float(np.sum((v / s) ** 2)))


def isoptimal_raw(mu, mus, mut, pobj, dobj, pres, dres, nu, gap_tol, feas_tol):
    """solver/solver.jl:20-30."""
    if not (pres < feas_tol and dres < feas_tol):
        return False
    tol = gap_tol * max(1.0, min(abs(pobj), abs(dobj)))
    if mut > 0 and nu > 0:
        return nu * (mu - mut + mut * math.log1p(mus * mut - 1.0)) <= tol
    return pobj - dobj < tol


class Solver:
    """IPMSolver (solver/ipm.jl:1-28, constructor :431-478)."""

    def __init__(self, prob, settings=None, order="mindegree", p0=None, d0=None, y0=None):
        st = settings or Settings()
        self.settings = st
        B0, Q0 = prob.B, prob.Q
        n, m = B0.ncol, B0.nrow
        self.nu = sum(C.degree(prob.K[v], B0.ncols(v)) for v in range(B0.nvtx))  # utils.jl:1-9
        # symbkkt (kkt.jl:23-53)
        if order == "mindegree":
            w, adj = vertex_graph(B0, Q0)
            self.sym = analyse(w, adj)
            vperm = self.sym.vperm
        else:
            self.sym = None
            vperm = np.arange(B0.nvtx)
        colp = np.concatenate([np.arange(B0.xcol[v], B0.xcol[v + 1]) for v in vperm]).astype(np.int64)
        self.Q = symselectvtxs(Q0, vperm)
        self.B = selectvtxs(B0, vperm)
        self.f = prob.f[colp].copy()
        self.g = prob.g.copy()
        self.K = [prob.K[v] for v in vperm]
        self.P1 = prob.P1
        self.P2 = prob.P2[colp]
        self.mu_t = prob.mu
        if st.scale_max_iter > 0:
            self.pscl, self.yscl = equilibrate(self.B, self.Q, self.f, self.g, itmax=st.scale_max_iter)
        else:
            self.pscl, self.yscl = np.ones(n), np.ones(m)
        self.kkt = UzawaSolver(self.B, irmax=st.refine_max_iter, cgmax=st.newton_max_iter)
        self.Bd = self.kkt.Bd
        self.Qd = self.Q.to_dense()
        self.Hd = self.Qd.copy()                      # H = copy(Q)
        self.Hnorm = self.Q.norm()
        self.caches = [C.make_cache(self.K[v], self.B.ncols(v)) for v in range(self.B.nvtx)]
        self.nB = self.B.norm()
        self.nf = float(np.linalg.norm(self.f))
        self.ng = float(np.linalg.norm(self.g))
        self.sg = scalenorm(self.g, self.yscl)
        self.sf = scalenorm(self.f, self.pscl)
        self.p = np.zeros(n)
        self.d = np.zeros(n)
        self.y = np.zeros(m)
        self.hist = []
        self.reinit(p0, d0, y0)

    # ---------------------------------------------------------------------------------
    def vranges(self):
        B = self.B
        for v in range(B.nvtx):
            yield v, int(B.xcol[v]), int(B.xcol[v + 1])

    def reinit(self, p0=None, d0=None, y0=None):
        """solver/ipm.jl:375-429 (cold start via startingpoint!, :331-365)."""
        colp = self.P2
        if p0 is None and d0 is None:
            for v, a, b in self.vranges():
                C.identity(self.p[a:b], self.K[v])
                C.identity(self.d[a:b], self.K[v])
            z = self.B.gemv(self.p)
            w = self.Q.gemv(self.p)
            np_ = float(np.linalg.norm(self.p))
            nz = float(np.linalg.norm(z))
            nw = float(np.linalg.norm(w))
            sp = max(1.0, float(np.linalg.norm(self.g)) / nz) if nz > EPS * np_ else 1.0
            sd = max(1.0, (float(np.linalg.norm(self.f)) + sp * nw) / np_) if np_ > EPS else 1.0
            self.p *= sp
            self.d *= sd
            if y0 is None:
                self.y[:] = 0.0
            else:
                self.y[:] = np.asarray(y0)[self.P1] / self.yscl
        else:
            if p0 is not None:
                self.p[:] = np.asarray(p0)[colp]
            if d0 is not None:
                self.d[:] = np.asarray(d0)[colp]
            for v, a, b in self.vranges():
                if d0 is None:
                    C.shadow(self.K[v], self.d[a:b], self.p[a:b])
                elif p0 is None:
                    C.shadow(self.K[v], self.p[a:b], self.d[a:b], primal=True)
            self.y[:] = 0.0 if y0 is None else np.asarray(y0)[self.P1]
            self.p /= self.pscl
            self.d *= self.pscl
            self.y /= self.yscl
        for v in range(self.B.nvtx):  # initcache! (solver/ipm.jl:422-424)
            if self.K[v][0] in ("exp", "pow"):
                from . import tdc
                self.caches[v].seed = tdc.seed0(self.K[v][0])
        self.hist = []

    # ---------------------------------------------------------------------------------
    def residuals(self):
        """Δf = d + f − (Qp − Bᵀy), Δg = g − Bp (solver/ipm.jl:86-99)."""
        u = self.Qd @ self.p - self.Bd.T @ self.y
        self.df = self.d + self.f - u
        self.dg = self.g - self.Bd @ self.p

    def getaug(self):
        """solver.jl:40-52 and history.jl:9-25."""
        nH = self.Hnorm
        if self.nB == 0 or nH == 0:
            return 1.0
        if not self.hist:
            return self.settings.aug_tol * self.nB ** 2 / nH
        h = self.hist[-1]
        delta, dmin, dmax = h["delta"], h["dmin"], h["dmax"]
        if h["ppass"] > 1:
            if math.isfinite(dmin):
                delta = max(delta, dmin)
            delta *= 10
        elif math.isfinite(dmin) and math.isfinite(dmax) and dmin > 0 and dmax > 0:
            delta = math.sqrt(dmin) * math.sqrt(dmax)
        return delta

    def scale(self):
        """scale! fan-out (cone.jl:176-247): H_vv ← Q_vv + cone block."""
        flags = True
        spsd = 0.0
        self.Hd[:, :] = self.Qd
        for v, a, b in self.vranges():
            Hv = self.Hd[a:b, a:b]
            fl, sp = C.scale(self.K[v], Hv, self.p[a:b], self.d[a:b], self.caches[v])
            self.Hd[a:b, a:b] = Hv
            flags = flags and fl
            spsd += sp
        self.Hnorm = self._hnorm()
        return flags, spsd

    def _hnorm(self):
        # ‖H‖ over the stored block values (block_sparse_matrix.jl:1084-1092); entries
        # outside the stored pattern are zero, so this is the dense Frobenius norm
        return float(np.linalg.norm(self.Hd))

    def maxsteps(self, dp, dd, step_frac):
        """cone.jl:455-497."""
        step = 1.0
        for v, a, b in self.vranges():
            tp, td = C.maxsteps(self.K[v], self.p[a:b], dp[a:b], self.d[a:b], dd[a:b], self.caches[v])
            if self.K[v][0] != "free":
                tp *= step_frac
                td *= step_frac
            step = min(step, tp, td)
        return step

    def isstalled(self, window=5):
        """history.jl:27-46."""
        tol = self.settings.stall_tol
        n = len(self.hist)
        if n <= window:
            return False
        for i in range(n - window, n):
            h, hp = self.hist[i], self.hist[i - 1]
            if (h["mu"] < (1 - tol) * hp["mu"] or h["pres"] < (1 - tol) * hp["pres"]
                    or h["dres"] < (1 - tol) * hp["dres"]):
                return False
        return True

    def isprimalinfeasible(self, rtol, atol):
        """solver/ipm.jl:502-517."""
        gy = float(self.g @ self.y)
        ny = float(np.linalg.norm(self.y))
        flag = gy > atol * ny * (1 + self.ng)
        if flag:
            nQp = float(np.linalg.norm(self.Qp))
            r = self.Qp - self.d - self.Bd.T @ self.y
            flag = max(nQp, float(np.linalg.norm(r))) < rtol * gy * (1 + float(np.linalg.norm(self.d)) / ny)
        return flag

    def isdualinfeasible(self, rtol, atol):
        """solver/ipm.jl:528-540."""
        fp = float(self.f @ self.p)
        np_ = float(np.linalg.norm(self.p))
        flag = fp > atol * np_ * (1 + self.nf)
        if flag:
            flag = max(float(np.linalg.norm(self.Bd @ self.p)), float(np.linalg.norm(self.Qp))) < rtol * abs(fp)
        return flag

    def isopt(self, mu, mus, pobj, dobj, pres, dres, factor=1.0):
        st = self.settings
        return isoptimal_raw(mu, mus, self.mu_t, pobj, dobj, pres, dres, self.nu,
                             factor * st.gap_tol, factor * st.feas_tol)

    def nearstatus(self, status, mu, mus, pobj, dobj, pres, dres):
        """solver/ipm.jl:561-571."""
        st = self.settings
        f = st.near_factor
        if self.isopt(mu, mus, pobj, dobj, pres, dres, f):
            return NEAR_OPTIMAL
        if self.isprimalinfeasible(f * st.infeas_rel, st.infeas_abs):
            return NEAR_PRIMAL_INFEASIBLE
        if self.isdualinfeasible(f * st.infeas_rel, st.infeas_abs):
            return NEAR_DUAL_INFEASIBLE
        return status

    def zerofree(self, x):
        for v, a, b in self.vranges():
            if self.K[v][0] == "free":
                x[a:b] = 0.0

    def fnorm(self, df):
        return scalenorm(df, self.pscl) / (1 + self.sf)

    def gnorm(self, dg):
        return scalenorm(dg, self.yscl) / (1 + self.sg)

    # ---------------------------------------------------------------------------------
    def step(self):
        """One Mehrotra predictor-corrector iteration (solver/ipm.jl:577-774)."""
        st = self.settings
        status = CONTINUE
        row = dict(pres=math.nan, dres=math.nan, pobj=math.nan, dobj=math.nan, step=0.0,
                   piter=0, ppass=0, pstat=KKT_SOLVED, citer=0, cpass=0, cstat=KKT_SOLVED,
                   dmin=math.nan, dmax=math.nan, rho=math.nan, delta=0.0, mu=math.nan, chi=math.nan)
        mut = self.mu_t
        delta = self.getaug()
        row["delta"] = delta
        pd = float(self.p @ self.d)
        flag, spsd = self.scale()
        if not flag:
            status = NUMERICAL_FAILURE
        else:
            self.residuals()
            nu1 = max(self.nu, 1)
            mu = pd / nu1
            mus = spsd / nu1
            chi = mu * mus - 1
            pres = self.gnorm(self.dg)
            dres = self.fnorm(self.df)
            self.Qp = self.Qd @ self.p
            pQp = float(self.p @ self.Qp)
            pobj = pQp / 2 - float(self.f @ self.p)
            dobj = float(self.g @ self.y) - pQp / 2
            row.update(mu=mu, chi=chi, pres=pres, dres=dres, pobj=pobj, dobj=dobj)
            if self.isopt(mu, mus, pobj, dobj, pres, dres):
                status = OPTIMAL
            elif self.isprimalinfeasible(st.infeas_rel, st.infeas_abs):
                status = PRIMAL_INFEASIBLE
            elif self.isdualinfeasible(st.infeas_rel, st.infeas_abs):
                status = DUAL_INFEASIBLE
            elif len(self.hist) >= st.max_iter:
                status = self.nearstatus(ITERATION_LIMIT, mu, mus, pobj, dobj, pres, dres)
            elif self.isstalled():
                status = self.nearstatus(STALLED, mu, mus, pobj, dobj, pres, dres)
            elif self.nu == 0:
                ok, rho = self.kkt.init(self.Hd, delta)
                row["rho"] = rho
                if not ok:
                    status = self.nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres)
                else:
                    dpa = np.zeros_like(self.p)
                    dya = np.zeros_like(self.y)
                    res = solvekkt(self.kkt, self.Hd, self.Bd, dpa, dya, self.df, self.dg, False,
                                   EPS, EPS, st.refine_stall_tol, st.refine_max_iter,
                                   st.newton_max_iter, 0,
                                   lambda v: self.fnorm(v) <= st.feas_tol,
                                   lambda v: self.gnorm(v) <= st.feas_tol)
                    row.update(piter=res[0], ppass=res[1], pstat=res[2], dmin=res[3], dmax=res[4])
                    self.p += dpa
                    self.y += dya
            elif not mu > 0:
                status = self.nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres)
            else:
                ok, rho = self.kkt.init(self.Hd, delta)
                row["rho"] = rho
                if not ok:
                    status = self.nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres)
                else:
                    mu1 = mu if not self.hist else self.hist[0]["mu"]
                    if mut > 0:
                        tol = FORCING_FRAC * min(abs(mu - mut) / abs(mu1 - mut), 1.0)
                    else:
                        tol = FORCING_FRAC * mu / mu1
                    ftol = tol * (1 + self.nf)
                    gtol = tol * (1 + self.ng)
                    # predictor (solver/ipm.jl:111-163)
                    fr = np.zeros_like(self.p)
                    if mut > 0:
                        self.fanout_corr0(fr, mut)
                    else:
                        fr[:] = -self.d
                    fr += self.df
                    dpa = np.zeros_like(self.p)
                    dya = np.zeros_like(self.y)
                    piter, ppass, pstat, dmin, dmax = solvekkt(
                        self.kkt, self.Hd, self.Bd, dpa, dya, fr, self.dg, False, ftol, gtol,
                        st.refine_stall_tol, st.refine_max_iter, st.newton_max_iter, 1)
                    dda = self.df.copy()
                    dda = -dda - self.Bd.T @ dya + self.Qd @ dpa
                    self.zerofree(dda)
                    # corrector (solver/ipm.jl:181-272)
                    taua = self.maxsteps(dpa, dda, 1.0)
                    smu = float(np.sum((self.p + taua * dpa) * (self.d + taua * dda)))
                    mua = smu / self.nu
                    if mut > 0:
                        dmu = mu - mut
                        if dmu > 0:
                            dmua = min(max(mua - mut, 0.0), dmu)
                            smu = mut + min(max(dmua * (dmua / dmu) ** 2, 0.0), dmu)
                        else:
                            smu = mut
                    else:
                        smu = min(max(mua * (mua / mu) ** 2, 0.0), mu)
                    fr = np.zeros_like(self.p)
                    self.fanout_corr(fr, dpa, dda, smu)
                    fr += self.df
                    dp = dpa.copy()
                    dy = dya.copy()
                    citer, cpass, cstat, _, _ = solvekkt(
                        self.kkt, self.Hd, self.Bd, dp, dy, fr, self.dg, True, ftol, gtol,
                        st.refine_stall_tol, st.refine_max_iter, st.newton_max_iter, 1)
                    dd = -self.df - self.Bd.T @ dy + self.Qd @ dp
                    self.zerofree(dd)
                    row.update(piter=piter, ppass=ppass, pstat=pstat, dmin=dmin, dmax=dmax,
                               citer=citer, cpass=cpass, cstat=cstat)
                    if pstat != KKT_SOLVED or cstat != KKT_SOLVED:
                        status = self.nearstatus(NUMERICAL_FAILURE, mu, mus, pobj, dobj, pres, dres)
                    else:
                        stp = self.maxsteps(dp, dd, st.step_frac)
                        row["step"] = stp
                        self.p += stp * dp
                        self.d += stp * dd
                        self.y += stp * dy
        self.hist.append(row)
        return status

    def fanout_corr0(self, r, smu):
        for v, a, b in self.vranges():
            C.corr0(self.K[v], r[a:b], self.p[a:b], self.d[a:b], smu, self.caches[v])

    def fanout_corr(self, r, dp, dd, smu):
        for v, a, b in self.vranges():
            C.corr(self.K[v], r[a:b], self.p[a:b], self.d[a:b], dp[a:b], dd[a:b], smu, self.caches[v])

    def solve(self):
        """solve_impl! (solver/solver.jl:119-145) + result (solver/ipm.jl:30-74)."""
        status = CONTINUE
        while status == CONTINUE:
            status = self.step()
        return self.result(status)

    def result(self, status):
        pu = self.p * self.pscl
        du = self.d / self.pscl
        yu = self.y * self.yscl
        if status in (PRIMAL_INFEASIBLE, NEAR_PRIMAL_INFEASIBLE, DUAL_INFEASIBLE, NEAR_DUAL_INFEASIBLE) or not self.hist:
            mu = pres = dres = pobj = dobj = math.nan
        else:
            self.residuals()
            pQp = float(self.p @ self.Qd @ self.p)
            pobj = pQp / 2 - float(self.f @ self.p)
            dobj = float(self.g @ self.y) - pQp / 2
            mu = math.nan if self.nu == 0 else float(self.p @ self.d) / self.nu
            pres = self.gnorm(self.dg)
            dres = self.fnorm(self.df)
        p = np.empty_like(pu)
        d = np.empty_like(du)
        y = np.empty_like(yu)
        p[self.P2] = pu
        d[self.P2] = du
        y[self.P1] = yu
        niter = len(self.hist)
        nsolve = sum(h["piter"] + h["ppass"] + h["citer"] + h["cpass"] for h in self.hist)
        return Result(p, d, y, status, niter, nsolve, self.hist, mu, pres, dres, pobj, dobj)


@dataclass
class Result:
    """IPMResult (result/ipm.jl:1-15)."""
    p: np.ndarray
    d: np.ndarray
    y: np.ndarray
    status: int
    niter: int
    nsolve: int
    history: list
    mu: float
    pres: float
    dres: float
    pobj: float
    dobj: float


def _dualshadow(solver):
    """dualshadow!(d*, B, p, K, …) (cone.jl:556-566) at the solver's iterate."""
    ds = np.zeros_like(solver.p)
    for v, a, b in solver.vranges():
        C.shadow(solver.K[v], ds[a:b], solver.p[a:b])
    return ds


def _diff_setup(solver):
    """differential.jl:5-16: initkkt! at getaug's δ, the forcing tolerances from μ's history."""
    delta = solver.getaug()
    ok, _ = solver.kkt.init(solver.Hd, delta)
    if not ok:
        raise RuntimeError("initkkt! failed")
    mu0 = solver.hist[0]["mu"]
    tol = FORCING_FRAC * solver.settings.gap_tol if (mu0 == 0 or not math.isfinite(mu0)) else \
        FORCING_FRAC * solver.hist[-1]["mu"] / mu0
    return tol * (1 + solver.nf), tol * (1 + solver.ng)


def _kkt(solver, f, g, ftol, gtol):
    st = solver.settings
    x, y = np.zeros_like(f), np.zeros_like(g)
    solvekkt(solver.kkt, solver.Hd, solver.Bd, x, y, f, g, False, ftol, gtol, st.refine_stall_tol,
             st.refine_max_iter, st.newton_max_iter, 1)
    return x, y


def frule(solver, df, dg, dmu):
    """frule!(Δp, Δy, Δd, s, Δf, Δg, Δμ) (differential.jl:5-75) at the solver's last iterate, in
    the problem's space. Pinned by tests/test_oracle_ipm.py: the equality-QP closed form, the
    adjoint identity with rrule and central finite differences of whole solves."""
    ftol, gtol = _diff_setup(solver)
    ps, ys = solver.pscl, solver.yscl
    fx = np.asarray(df, dtype=np.float64)[solver.P2] * ps
    gx = np.asarray(dg, dtype=np.float64)[solver.P1] * ys
    dx = dmu * _dualshadow(solver)
    px, yx = _kkt(solver, fx + dx, gx, ftol, gtol)
    dx = dx - (solver.Hd - solver.Qd) @ px
    dp, dy, dd = (np.zeros(len(px)), np.zeros(len(yx)), np.zeros(len(px)))
    dp[solver.P2] = px * ps
    dy[solver.P1] = yx * ys
    dd[solver.P2] = dx / ps
    return dp, dy, dd


def rrule(solver, dp, dy, dd):
    """rrule!(Δf, Δg, s, Δp, Δy, Δd) (differential.jl:252-334): (Δf, Δg, dμ)."""
    ftol, gtol = _diff_setup(solver)
    ps, ys = solver.pscl, solver.yscl
    px = np.asarray(dp, dtype=np.float64)[solver.P2] * ps
    dx = np.asarray(dd, dtype=np.float64)[solver.P2] / ps
    yx = -np.asarray(dy, dtype=np.float64)[solver.P1] * ys
    px = px - (solver.Hd - solver.Qd) @ dx
    fx, gx = _kkt(solver, px, yx, ftol, gtol)
    dx = dx + fx
    df, dg = np.zeros(len(fx)), np.zeros(len(gx))
    df[solver.P2] = fx * ps
    dg[solver.P1] = -gx * ys
    return df, dg, float(dx @ _dualshadow(solver))


def solve(prob, settings=None, order="mindegree", **kw):
    return Solver(prob, settings, order=order, **kw).solve()
