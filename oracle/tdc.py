"""Exponential and power cones (oracle restatement of src/IPM/src/cone/tdc/{tdc,exp,pow,utils}.jl).

Test infrastructure only: the device path never imports this module.

The reference evaluates the determinant chains in two-word (double-double) arithmetic, with the
two-word logarithms and powers of two of DoubleFloats.jl (an external dependency, absent here;
called from utils.jl:84-121 and pow.jl:47-70). This restatement takes those two-word values from
Python's decimal module at 60 digits, rounded to the nearest double-double. That is the value
the reference's two-word routines approximate, so it is no less accurate than theirs.

Pinned by tests/test_oracle_tdc.py:
  - the two-word log / exp2 against decimal at 60 digits and against the closed forms log 1 = 0
    and 2^k exact;
  - barrier gradient, Hessian (L Lᵀ through the pivot) and third derivative against central
    finite differences of an independently written barrier value;
  - −f′(e) = e at the identity point (exact for the power cone; the exponential cone's
    constants, exp.jl:15-20, are the fixed point to about six digits);
  - the shadow primal solves −f′(p*) = d (residual checked through the gradient);
  - the Tunçel block M is symmetric positive definite and satisfies the secant equations
    M p = d and M p* = d*;
  - the max step against bisection on the cone-membership test;
  - the corrector's third-order term against finite differences of the Hessian;
  - whole solves against closed-form optima (entropy on the simplex; the weighted geometric
    mean under a budget).
"""
import math
from decimal import Decimal, getcontext

import numpy as np

from .cones import twoprod, twosum

EPS = float(np.finfo(np.float64).eps)
TINY = float(np.finfo(np.float64).tiny)
NBAND = 65536.0 * EPS  # tdc/utils.jl:6
NFK = 8.0              # tdc/utils.jl:11

CACHE = 19
C_L, C_SD, C_SEED, C_D1, C_D2, C_D3, C_D4, C_H0P, C_H0D = 0, 9, 12, 13, 14, 15, 16, 17, 18

EXP_E = (1.2909282315382298, 0.8051015526498357, -0.8278379086082098)  # exp.jl:15-20


# ---------------------------------------------------------------------------------------
# two-word values (utils.jl:28-38, 84-129; the log₂ / exp₂ kernels are DoubleFloats')
# ---------------------------------------------------------------------------------------

def _dd_of(x):
    """Nearest double-double of a Decimal."""
    h = float(x)
    return h, float(x - Decimal(h))


def twolograt(x1, x2):
    """log(x1/x2) as two words (utils.jl:114-118)."""
    getcontext().prec = 60
    return _dd_of((Decimal(x1) / Decimal(x2)).ln())


def twolog2rat(x1, x2):
    """log₂(x1/x2) as two words (utils.jl:84-111)."""
    getcontext().prec = 60
    return _dd_of((Decimal(x1) / Decimal(x2)).ln() / Decimal(2).ln())


def twoexp2(ah, al):
    """2^(ah + al) as two words (utils.jl:120-123): 2^k exactly times 2^f, f = ah + al − k."""
    getcontext().prec = 60
    k = int(round(ah))
    f = (Decimal(ah) - k) + Decimal(al)
    h, lo = _dd_of((f * Decimal(2).ln()).exp())
    return math.ldexp(h, k), math.ldexp(lo, k)


def twodiv(ah, al, bh, bl):
    """(ah + al)/(bh + bl) (utils.jl:125-129)."""
    q1 = ah / bh
    r = _fma(-q1, bh, ah) + al - q1 * bl
    return twosum(q1, r / bh)


def _fma(a, b, c):
    """a·b + c rounded once (exact product through the error-free transformation)."""
    p, e = twoprod(a, b)
    s, t = twosum(p, c)
    return s + (t + e)


def cdot3(x, y):
    s = c = 0.0
    for a, b in zip(x, y):
        pr, e = twoprod(float(a), float(b))
        s, e2 = twosum(s, pr)
        c += e + e2
    return s + c


def dot3(x, y):
    return float(x[0] * y[0] + x[1] * y[1] + x[2] * y[2])


def cross3(x, y):
    return np.array([x[1] * y[2] - x[2] * y[1], x[2] * y[0] - x[0] * y[2], x[0] * y[1] - x[1] * y[0]])


# ---------------------------------------------------------------------------------------
# pivoted factor helpers (tdc.jl:51-119): P f''(p) Pᵀ = L Lᵀ, P v = (v₃, v₁, v₂)
# ---------------------------------------------------------------------------------------

_PERM = [2, 0, 1]


def hessian(L):
    """f''(p) = Pᵀ L Lᵀ P as a dense 3×3 matrix."""
    G = L @ L.T
    H = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            H[_PERM[a], _PERM[b]] = G[a, b]
    return H


def hessmul(L, v):
    """f''(p) v = Pᵀ L (Lᵀ (P v)) (tdc.jl:76-94)."""
    pv = np.array([v[2], v[0], v[1]], dtype=np.float64)
    t = L @ (L.T @ pv)
    return np.array([t[1], t[2], t[0]])


def hessldiv(L, b):
    """f''(p)⁻¹ b by the two triangular solves through the pivot (tdc.jl:51-67)."""
    import scipy.linalg as sla
    pb = np.array([b[2], b[0], b[1]], dtype=np.float64)
    y = sla.solve_triangular(L, pb, lower=True)
    z = sla.solve_triangular(L.T, y, lower=False)
    return np.array([z[1], z[2], z[0]])


# ---------------------------------------------------------------------------------------
# exponential cone (exp.jl)
# ---------------------------------------------------------------------------------------

def expdet(x):
    """(ψ, 1/ψ, log(x₁/x₂), log(x₁/x₂) − 1), ψ = x₂ log(x₁/x₂) − x₃ (exp.jl:43-61)."""
    x1, x2, x3 = map(float, x)
    lrh, lrl = twolograt(x1, x2)
    mh, me = twosum(lrh, -1.0)
    lrm1 = mh + (me + lrl)
    ph, pe = twoprod(lrh, x2)
    pe = _fma(lrl, x2, pe)
    sh, se = twosum(ph, -x3)
    qh, ql = twosum(sh, se + pe)
    ih, il = twodiv(1.0, 0.0, qh, ql)
    return qh + ql, ih + il, lrh + lrl, lrm1


def expgrad(x, ipsi, lrm1):
    x1, x2 = float(x[0]), float(x[1])
    return np.array([-(x2 * ipsi + 1.0) / x1, -lrm1 * ipsi - 1.0 / x2, ipsi])


def expfact(x, psi, ipsi, lrm1):
    x1, x2 = float(x[0]), float(x[1])
    r1 = 1.0 + x2 * ipsi
    r2 = (psi + 2.0 * x2) / (psi + x2)
    ok = psi > 0 and r1 > 0 and r2 > 0 and math.isfinite(r1) and math.isfinite(r2)
    L = np.zeros((3, 3))
    if ok:
        s1 = math.sqrt(r1)
        L[0, 0] = ipsi
        L[1, 0] = -x2 * ipsi / x1
        L[2, 0] = -lrm1 * ipsi
        L[1, 1] = s1 / x1
        L[2, 1] = -ipsi / s1
        L[2, 2] = math.sqrt(r2) / x2
    return ok, L


def expthird(x, u, ipsi, lrm1):
    """f'''(x)[u] (exp.jl:105-141)."""
    x1, x2 = float(x[0]), float(x[1])
    u1, u2, u3 = map(float, u)
    g1, g2 = x2 / x1, lrm1
    gu = g1 * u1 + g2 * u2 - u3
    H11, H21, H22 = -x2 / x1 ** 2, 1.0 / x1, -1.0 / x2
    Hu1, Hu2 = H11 * u1 + H21 * u2, H21 * u1 + H22 * u2
    i2 = ipsi * ipsi
    i3 = i2 * ipsi
    al = -2.0 * gu * i3
    D = np.zeros((3, 3))
    D[0, 0] = al * g1 ** 2 + (2 * g1 * Hu1 + H11 * gu) * i2 - (2 * x2 * u1 / x1 ** 3 - u2 / x1 ** 2) * ipsi - 2 * u1 / x1 ** 3
    D[1, 0] = al * g1 * g2 + (g1 * Hu2 + g2 * Hu1 + H21 * gu) * i2 + u1 / x1 ** 2 * ipsi
    D[2, 0] = -al * g1 - Hu1 * i2
    D[1, 1] = al * g2 ** 2 + (2 * g2 * Hu2 + H22 * gu) * i2 - u2 / x2 ** 2 * ipsi - 2 * u2 / x2 ** 3
    D[2, 1] = -al * g2 - Hu2 * i2
    D[2, 2] = al
    D[0, 1], D[0, 2], D[1, 2] = D[1, 0], D[2, 0], D[2, 1]
    return D


def expdualgrad(seed, d):
    """Shadow primal by Halley on v + log1p(v) = R, v ∈ [R/2, R] (exp.jl:199-364).
    Returns (p*, next seed, h_dual(0))."""
    d1, d2, d3 = map(float, d)
    if d3 == 0.0:
        p2 = 1.0 / d2
        s = d2 / d1
        return np.array([s * p2, p2, -math.inf]), p2, (math.inf if d1 >= 0 and d2 >= 0 else -math.inf)
    md3 = -d3
    rat = md3 / d1
    lr = math.log(rat) if TINY <= rat < math.inf else math.log(md3) - math.log(d1)
    w = d2 / md3
    R = w - (lr - 1.0)
    if not R > 0:
        R = TINY
    lo, hi = R / 2, R
    v = 1.0 / (max(seed, TINY) * md3)
    if not lo < v < hi:
        v = R / 2 + R * R / 16 if R <= 2 else R - math.log(R)
        if not lo < v < hi:
            v = 0.7 * R
    for _ in range(60):
        fv = v + math.log1p(v) - R
        fp = 1.0 + 1.0 / (1.0 + v)
        if fv < 0:
            lo = v
        elif fv > 0:
            hi = v
        else:
            break
        fpp = -(fp - 1.0) ** 2
        vn = v - 2 * fv * fp / (2 * fp * fp - fv * fpp)
        if not lo < vn < hi:
            if vn == lo or vn == hi:
                v = vn
                break
            vn = (lo + hi) / 2
        conv = abs(vn - v) < 4 * EPS * abs(vn)
        v = vn
        if conv:
            break
    p2 = 1.0 / (md3 * v)
    s = md3 * (v + 1.0) / d1
    if not 0.0 < s < math.inf:
        s = math.exp(lr + math.log1p(v))
    return np.array([s * p2, p2, p2 * math.log(s) + 1.0 / d3]), p2, 1.0 - lr + w


def expjetprim(tau, p, dp):
    x1, x2, x3 = (float(p[i] + tau * dp[i]) for i in range(3))
    if x1 > 0 and x2 > 0:
        q = x1 / x2
        s = math.log(q) if TINY <= q < math.inf else math.log(x1) - math.log(x2)
        w = x3 / x2
        wp = (dp[2] - w * dp[1]) / x2
        return (s - w, dp[0] / x1 - dp[1] / x2 - wp,
                -(dp[0] / x1) ** 2 + (dp[1] / x2) ** 2 + 2 * wp * dp[1] / x2)
    if x1 >= 0 and x2 == 0:
        return (math.inf if x3 <= 0 else -math.inf), math.nan, math.nan
    return -math.inf, math.nan, math.nan


def expjetdual(tau, d, dd):
    z1, z2, z3 = (float(d[i] + tau * dd[i]) for i in range(3))
    if z1 > 0 and z3 < 0:
        r = z2 / z3
        rp = (dd[1] - r * dd[2]) / z3
        return (1.0 + math.log(z1) - r - math.log(-z3), dd[0] / z1 - rp - dd[2] / z3,
                -(dd[0] / z1) ** 2 + 2 * rp * dd[2] / z3 + (dd[2] / z3) ** 2)
    if z1 >= 0 and z2 >= 0 and z3 == 0:
        return math.inf, math.nan, math.nan
    return -math.inf, math.nan, math.nan


# ---------------------------------------------------------------------------------------
# power cone (pow.jl)
# ---------------------------------------------------------------------------------------

def _dd_add(a, b):
    s, e = twosum(a[0], b[0])
    t, f = twosum(a[1], b[1])
    e += t
    h = s + e
    e = e - (h - s)
    s2 = h + (e + f)
    return s2, (e + f) - (s2 - h)


def _dd_mul(a, b):
    p, e = twoprod(a[0], b[0])
    e += a[0] * b[1] + a[1] * b[0]
    h = p + e
    return h, e - (h - p)


def powdet(x, al):
    """(P, ρ, 1/det) with P = x₁^{2α} x₂^{2−2α}, det = P − x₃², ρ = P/det (pow.jl:47-70)."""
    x1, x2, x3 = map(float, x)
    a = 2.0 * al
    r = twolog2rat(x1, x2)
    g = twolog2rat(x2, 1.0)
    L2 = _dd_add(_dd_mul((a, 0.0), r), (2 * g[0], 2 * g[1]))
    Ph, Pl = twoexp2(*L2)
    qh, ql = twoprod(x3, x3)
    fh, fe = twosum(Ph, -qh)
    dh, dl = twosum(fh, fe + (Pl - ql))
    ih, il = twodiv(1.0, 0.0, dh, dl)
    rh, re = twoprod(Ph, ih)
    re = _fma(Ph, il, _fma(Pl, ih, re))
    rh, rl = twosum(rh, re)
    return Ph + Pl, rh + rl, ih + il


def powgrad(x, al, rho, idet):
    a, b = 2 * al, 2 - 2 * al
    return np.array([-(a * rho + 1 - al) / x[0], -(b * rho + al) / x[1], 2 * x[2] * idet])


def powfact(x, al, P, rho, idet):
    x1, x2, x3 = map(float, x)
    a, b = 2 * al, 2 - 2 * al
    w = rho * idet
    s = P + x3 * x3
    l1, l2 = a / x1, b / x2
    d1 = (2 * rho * a + b) / (2 * x1 ** 2)
    d2 = (2 * rho * b + a) / (2 * x2 ** 2)
    c = rho * x3 ** 2 / s
    r1 = d1 - c * l1 ** 2
    r2 = (d1 * d2 - c * (l2 ** 2 * d1 + l1 ** 2 * d2)) / r1 if r1 != 0 else math.nan
    ok = r1 > 0 and r2 > 0 and math.isfinite(r2)
    L = np.zeros((3, 3))
    if ok:
        L[0, 0] = math.sqrt(2 * s) * idet
        L[1, 0] = -2 * l1 * x3 * w / L[0, 0]
        L[2, 0] = -2 * l2 * x3 * w / L[0, 0]
        L[1, 1] = math.sqrt(r1)
        L[2, 2] = math.sqrt(r2)
        L[2, 1] = -c * l1 * l2 / L[1, 1]
    return ok, L


def powthird(x, u, al, P, rho, idet):
    """f'''(x)[u] (pow.jl:139-191)."""
    x1, x2, x3 = map(float, x)
    u1, u2, u3 = map(float, u)
    a, b = 2 * al, 2 - 2 * al
    l1, l2, m1, m2 = a / x1, b / x2, u1 / x1, u2 / x2
    f1, f2, f3 = P * l1, P * l2, -2 * x3
    fdot = f1 * u1 + f2 * u2 + f3 * u3
    f11, f22, f12 = P * l1 * (l1 - 1 / x1), P * l2 * (l2 - 1 / x2), P * l1 * l2
    fd1, fd2, fd3 = f11 * u1 + f12 * u2, f12 * u1 + f22 * u2, -2 * u3
    i2 = idet * idet
    i3 = i2 * idet
    c2, c3 = fdot * i2, 2 * fdot * i3
    K = rho * a * b * (a - 1) * (m2 - m1)
    D = np.zeros((3, 3))
    D[0, 0] = 2 * fd1 * f1 * i2 - c3 * f1 * f1 - K / (x1 * x1) + c2 * f11 - b * m1 / (x1 * x1)
    D[1, 1] = 2 * fd2 * f2 * i2 - c3 * f2 * f2 - K / (x2 * x2) + c2 * f22 - a * m2 / (x2 * x2)
    D[1, 0] = (fd2 * f1 + f2 * fd1) * i2 - c3 * f2 * f1 + K / (x1 * x2) + c2 * f12
    D[2, 2] = 2 * fd3 * f3 * i2 - c3 * f3 * f3 - 2 * c2
    D[2, 0] = (fd3 * f1 + f3 * fd1) * i2 - c3 * f3 * f1
    D[2, 1] = (fd3 * f2 + f3 * fd2) * i2 - c3 * f3 * f2
    D[0, 1], D[0, 2], D[1, 2] = D[1, 0], D[2, 0], D[2, 1]
    return D


def powdualgrad(seed, d, al):
    """Shadow primal by the log residual of 1 − u = W₀(1 + k₁u)^a(1 + k₂u)^b (pow.jl:262-447).
    Returns (p*, next seed ρ, h_dual(0))."""
    d1, d2, d3 = map(float, d)
    a, b = 2 * al, 2 - 2 * al
    if d3 == 0.0:
        return np.array([(a + 1 - al) / d1, (b + al) / d2, 0.0]), 1.0, math.inf
    k1, k2 = (1 - al) / a, al / b
    t1, t2 = a / d1, b / d2
    z1 = (d3 / 2) * t2
    z1s = z1 * z1
    z2 = t1 / t2
    z2a = z2 ** a
    W0 = z1s * z2a
    normal = all(TINY <= v < math.inf for v in (z1s, z2, z2a, W0))
    if not normal:
        lW0 = 2 * math.log(abs(d3)) - math.log(4.0) + a * (math.log(a) - math.log(d1)) + b * (math.log(b) - math.log(d2))
        W0 = math.exp(lW0)
    else:
        lW0 = math.log(W0)
    f0 = 1 - W0
    if not f0 > 0:
        f0 = TINY
    lo, hi = f0 / 2, f0
    u = 1.0 / max(seed, 1.0)
    if not lo < u < hi:
        Fpp0 = 1 - (k1 ** 2 * a + k2 ** 2 * b)
        u = -4 * lW0 / (8 - lW0 * Fpp0)
        if not lo < u < hi:
            u = 0.7 * f0
    for _ in range(60):
        r1, r2 = 1 + k1 * u, 1 + k2 * u
        Fv = lW0 + a * math.log1p(k1 * u) + b * math.log1p(k2 * u) - math.log1p(-u)
        sg = k1 * a / r1 + k2 * b / r2
        Fp = sg + 1 / (1 - u)
        if Fv < 0:
            lo = u
        elif Fv > 0:
            hi = u
        else:
            break
        sgp = -(k1 ** 2 * a / r1 ** 2 + k2 ** 2 * b / r2 ** 2)
        Fpp = sgp + 1 / (1 - u) ** 2
        halley = u - 2 * Fv * Fp / (2 * Fp * Fp - Fv * Fpp)
        pole = 1 - (1 - u) * math.exp(Fv)
        un, alt = (pole, halley) if 1 / (1 - u) > 4 * sg else (halley, pole)
        if not lo < un < hi:
            if un == lo or un == hi:
                u = un
                break
            un = alt if lo < alt < hi else (lo + hi) / 2
        conv = abs(un - u) < 4 * EPS * (1 + abs(un))
        u = un
        if conv:
            break
    rho = 1 / u
    return np.array([(a * rho + 1 - al) / d1, (b * rho + al) / d2, -2 * (rho - 1) / d3]), rho, -lW0 / 2


def powjetprim(tau, p, dp, al):
    be = 1 - al
    x1, x2, x3 = (float(p[i] + tau * dp[i]) for i in range(3))
    if x1 > 0 and x2 > 0:
        if x3 == 0:
            return math.inf, math.nan, math.nan
        return (al * math.log(x1) + be * math.log(x2) - math.log(abs(x3)),
                al * dp[0] / x1 + be * dp[1] / x2 - dp[2] / x3,
                -al * (dp[0] / x1) ** 2 - be * (dp[1] / x2) ** 2 + (dp[2] / x3) ** 2)
    if (x1 == 0 and x2 >= 0) or (x1 >= 0 and x2 == 0):
        return (0.0 if x3 == 0 else -math.inf), math.nan, math.nan
    return -math.inf, math.nan, math.nan


def powjetdual(tau, d, dd, al):
    be = 1 - al
    s1, s2, s3 = (float(d[i] + tau * dd[i]) for i in range(3))
    if s1 > 0 and s2 > 0:
        if s3 == 0:
            return math.inf, math.nan, math.nan
        return (al * math.log(s1 / al) + be * math.log(s2 / be) - math.log(abs(s3)),
                al * dd[0] / s1 + be * dd[1] / s2 - dd[2] / s3,
                -al * (dd[0] / s1) ** 2 - be * (dd[1] / s2) ** 2 + (dd[2] / s3) ** 2)
    if (s1 == 0 and s2 >= 0) or (s1 >= 0 and s2 == 0):
        return (0.0 if s3 == 0 else -math.inf), math.nan, math.nan
    return -math.inf, math.nan, math.nan


# ---------------------------------------------------------------------------------------
# cone-generic layer (tdc.jl)
# ---------------------------------------------------------------------------------------

class TdCache:
    """The 19 cache slots of tdc.jl:142-163 as named fields."""
    __slots__ = ("L", "sd", "seed", "d1", "d2", "d3", "d4", "h0p", "h0d")

    def __init__(self, kind):
        self.L = np.zeros((3, 3))
        self.sd = np.zeros(3)
        self.seed = seed0(kind)
        self.d1 = self.d2 = self.d3 = self.d4 = 0.0
        self.h0p = self.h0d = 0.0


def seed0(kind):
    """initcache! (exp.jl:27-30, pow.jl:35-38)."""
    return EXP_E[1] if kind == "exp" else 1.0


def identity(cone):
    if cone[0] == "exp":
        return np.array(EXP_E)
    al = cone[1]
    return np.array([math.sqrt(1 + al), math.sqrt(2 - al), 0.0])


def det(cone, p, c):
    """tddet! (exp.jl:374-389, pow.jl:456-470)."""
    if cone[0] == "exp":
        c.d1, c.d2, c.d3, c.d4 = expdet(p)
        c.h0p = c.d3 - p[2] / p[1]
    else:
        c.d1, c.d2, c.d3 = powdet(p, cone[1])
        with np.errstate(divide="ignore"):
            c.h0p = math.log(c.d1) / 2 - (math.log(abs(p[2])) if p[2] != 0 else -math.inf)


def barrgrad(cone, p, c):
    if cone[0] == "exp":
        return expgrad(p, c.d2, c.d4)
    return powgrad(p, cone[1], c.d2, c.d3)


def fact(cone, p, c):
    if cone[0] == "exp":
        return expfact(p, c.d1, c.d2, c.d4)
    return powfact(p, cone[1], c.d1, c.d2, c.d3)


def third(cone, p, u, c):
    if cone[0] == "exp":
        return expthird(p, u, c.d2, c.d4)
    return powthird(p, u, cone[1], c.d1, c.d2, c.d3)


def dualgrad(cone, seed, d, c):
    if cone[0] == "exp":
        sp, nxt, h0 = expdualgrad(seed, d)
    else:
        sp, nxt, h0 = powdualgrad(seed, d, cone[1])
    c.h0d = h0
    return sp, nxt


def bfgs(L, sp, sd, z, p, mu, mut):
    """t of the rank-one term (tdbfgs, tdc.jl:278-354)."""
    w = sp - mut * p
    Hw = hessmul(L, w)
    wHw = dot3(w, Hw)
    if not wHw > 0:
        return 0.0
    Hz = hessmul(L, z)
    fzz, sdz, wHz = dot3(z, Hz), dot3(sd, z), dot3(Hw, z)
    s1, e1 = twosum(fzz, -sdz ** 2 / 3)
    s2, e2 = twosum(s1, -wHz ** 2 / wHw)
    return mu * (s2 + (e1 + e2))


def scale(cone, H, p, d, c):
    """tdscale! (tdc.jl:185-270): adds the Tunçel block to H; returns (flag, spsd)."""
    det(cone, p, c)
    c.sd[:] = -barrgrad(cone, p, c)
    ok, L = fact(cone, p, c)
    if not ok:
        c.seed = 1.0
        return False, 0.0
    c.L[:, :] = L
    sp, c.seed = dualgrad(cone, c.seed, d, c)
    mu = dot3(p, d) / 3
    spsd = cdot3(sp, c.sd)
    mut = spsd / 3
    z = cross3(p, sp)
    nz, np_, nsp = (float(np.linalg.norm(v)) for v in (z, p, sp))
    if nz < EPS * (np_ * nsp + EPS):
        H += mu * hessian(L)
        return True, spsd
    z = z / nz
    t = bfgs(L, sp, c.sd, z, p, mu, mut)
    if t <= 0 or not math.isfinite(t):
        H += mu * hessian(L)
        return True, spsd
    dp = p - mu * sp
    dd = d - mu * c.sd
    H += np.outer(d, d) / (3 * mu) + np.outer(dd, dd) / cdot3(dp, dd) + t * np.outer(z, z)
    return True, spsd


def corr0(r, d, c, smu):
    """tdcorr0! (tdc.jl:414-419): σμ d* − d."""
    r[:] = smu * c.sd - d


def corr(cone, r, p, d, dp, dd, c, smu):
    """tdcorr! (tdc.jl:371-408): σμ d* − d − η, η = −½ f'''(p)[Δp, f''(p)⁻¹Δd]."""
    v = hessldiv(c.L, np.asarray(dd, dtype=np.float64))
    eta = -0.5 * third(cone, p, dp, c) @ v
    r[:] = smu * c.sd - d - eta


def nflog3(jet, hi, glo, gplo):
    """Certified step on the log-form membership function (tdc/utils.jl:24-200)."""
    tb, nb, lo = hi, NBAND, 0.0
    if not glo >= nb:
        return lo
    ghi, gphi, gpphi = jet(hi)
    if ghi >= nb:
        return hi
    whi = nb / abs(gphi) if math.isfinite(gphi) and gphi != 0 else 0.0
    st = dict(lo=lo, hi=hi, glo=glo, gplo=gplo, ghi=ghi, gphi=gphi, gpphi=gpphi, whi=whi)

    def take(m):
        g, gp, gpp = jet(m)
        if g >= nb:
            st.update(lo=m, glo=g, gplo=gp)
        else:
            st.update(hi=m, ghi=g, gphi=gp, gpphi=gpp,
                      whi=nb / abs(gp) if math.isfinite(gp) and gp != 0 else 0.0)
        return st["hi"] - st["lo"] < max(NFK * st["whi"], EPS * (1 + st["hi"]))

    for _ in range(53):
        lo, hi, glo, ghi, gplo = st["lo"], st["hi"], st["glo"], st["ghi"], st["gplo"]
        wid = hi - lo
        if math.isfinite(ghi) and ghi >= -nb:
            m = hi - 3 * max(st["whi"], EPS * (1 + hi)) / 2
        elif math.isfinite(ghi) and math.isfinite(glo):
            m = lo + glo * wid / (glo - ghi)
            best = abs((ghi - glo) / wid - gplo)
            dlo, dhi = tb - lo, tb - hi
            if dhi > 0:
                ulo, uhi = math.log(dlo), math.log(dhi)
                cc = (glo - ghi) / (ulo - uhi)
                e = abs(-cc / dlo - gplo)
                if e < best:
                    best = e
                    m2 = tb - math.exp(ulo - glo * (ulo - uhi) / (glo - ghi))
                    if lo < m2 < hi:
                        m = m2
                bb = (glo - ghi) / (1 / dhi - 1 / dlo)
                aa = glo + bb / dlo
                e = abs(-bb / (dlo * dlo) - gplo)
                if e < best and aa != 0:
                    m2 = tb - bb / aa
                    if lo < m2 < hi:
                        m = m2
        else:
            m = lo + wid / 2
        if not lo < m < hi:
            m = lo + wid / 2
        if take(m):
            break
        lo, hi, ghi, gphi, gpphi = st["lo"], st["hi"], st["ghi"], st["gphi"], st["gpphi"]
        dh = tb - hi
        m = math.nan
        if math.isfinite(gpphi) and math.isfinite(gphi) and gphi != 0 and dh > 0:
            rr = gpphi * dh / gphi
            if 0.5 < rr < 1.5:
                m = tb - dh * math.exp(ghi / (gphi * dh))
            elif 1.5 <= rr < 3:
                den = ghi - gphi * dh
                if den != 0:
                    m = tb + gphi * dh * dh / den
        if not lo < m < hi:
            den = 2 * gphi * gphi - ghi * gpphi
            m = hi - 2 * ghi * gphi / den if den != 0 else math.nan
            if not lo < m < hi:
                m = hi - ghi / gphi if gphi != 0 else math.nan
        if not lo < m < hi:
            m = lo + (hi - lo) / 2
        if take(m):
            break
        if 2 * (st["hi"] - st["lo"]) > wid:
            if take(st["lo"] + (st["hi"] - st["lo"]) / 2):
                break
    return st["lo"]


def _bound(x, dx, i0, i1, neg1):
    hi = 1.0
    if dx[i0] < 0:
        hi = min(hi, -x[i0] / dx[i0])
    if (dx[i1] > 0) if neg1 else (dx[i1] < 0):
        hi = min(hi, -x[i1] / dx[i1])
    return hi


def maxsteps(cone, p, dp, d, dd, c):
    """(τp, τd) before step_frac (tdmaxsteps, tdc.jl:437-458)."""
    p, dp, d, dd = (np.asarray(v, dtype=np.float64) for v in (p, dp, d, dd))
    if cone[0] == "exp":
        w = p[2] / p[1]
        hp0 = dp[0] / p[0] - dp[1] / p[1] - (dp[2] - w * dp[1]) / p[1]
        r = d[1] / d[2]
        hd0 = dd[0] / d[0] - (dd[1] - r * dd[2]) / d[2] - dd[2] / d[2]
        tp = nflog3(lambda t: expjetprim(t, p, dp), _bound(p, dp, 0, 1, False), c.h0p, hp0)
        td = nflog3(lambda t: expjetdual(t, d, dd), _bound(d, dd, 0, 2, True), c.h0d, hd0)
    else:
        al = cone[1]
        hp0 = al * dp[0] / p[0] + (1 - al) * dp[1] / p[1] - dp[2] / p[2]
        hd0 = al * dd[0] / d[0] + (1 - al) * dd[1] / d[1] - dd[2] / d[2]
        tp = nflog3(lambda t: powjetprim(t, p, dp, al), _bound(p, dp, 0, 1, False), c.h0p, hp0)
        td = nflog3(lambda t: powjetdual(t, d, dd, al), _bound(d, dd, 0, 1, False), c.h0d, hd0)
    return tp, td


def dual_shadow(cone, s, x):
    """s = −f′(x) (dualshadow!, tdc.jl:464-469)."""
    c = TdCache(cone[0])
    det(cone, x, c)
    s[:] = -barrgrad(cone, x, c)
    return True


def primal_shadow(cone, x, s):
    """x with −f′(x) = s (primalshadow!, tdc.jl:471-474)."""
    c = TdCache(cone[0])
    x[:], _ = dualgrad(cone, 0.0, s, c)
    return True
