"""Cone kernels (oracle restatement of src/IPM/src/cone/{pos,noc,soc}.jl and the
error-free transformations of src/IPM/src/utils.jl:27-53).

Pinned by (tests/test_oracle_cones_kkt.py):
  - NT invariant H·p = d for the scaling block added by `scale` (pos and SOC), H ≻ 0;
  - −f′(e) = e at the identity; −f′ is an involution on the interior (shadow twice);
  - √H (socroot) applied twice equals the NT block H, and its inverse form undoes it;
  - the corrector reduces to the closed-form
    Jordan-product formula; max step against brute-force bisection on membership;
  - twoprod / twosum exactness checked with fractions.Fraction.
"""
import math

import numpy as np

_SPLIT = 134217729.0  # 2^27 + 1 (Veltkamp split for binary64)


def twosum(a, b):
    """s + e = a + b exactly (utils.jl:28-32)."""
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _split(a):
    c = _SPLIT * a
    hi = c - (c - a)
    return hi, a - hi


def twoprod(a, b):
    """p + e = a·b exactly (utils.jl:35-38). The reference uses fma(a, b, -p); Dekker's
    product with Veltkamp splitting yields the same exact error term without overflow."""
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
    return p, e


def cdot(p, d):
    """Compensated dot product (utils.jl:41-53)."""
    s = c = 0.0
    for x, y in zip(p, d):
        pr, e = twoprod(float(x), float(y))
        s, e2 = twosum(s, pr)
        c += e + e2
    return s + c


def socdot(x, y):
    """Compensated J-inner product x₁y₁ − Σ xᵢyᵢ as (hi, lo) (soc.jl:48-61)."""
    s, c = twoprod(float(x[0]), float(y[0]))
    for i in range(1, len(x)):
        p, e = twoprod(float(x[i]), float(y[i]))
        s, e2 = twosum(s, -p)
        c += -e + e2
    return twosum(s, c)


def socdet(x):
    return socdot(x, x)


# ---------------------------------------------------------------------------------------
# degree / cache sizes (pos.jl:20-30, noc.jl:20-30, soc.jl:116-126)
# ---------------------------------------------------------------------------------------

def degree(cone, n):
    k = cone[0]
    if k == "pos":
        return n
    if k == "free":
        return 0
    if k == "soc":
        return 2
    if k == "psd":
        return triroot(n)
    if k in ("exp", "pow"):
        return 3
    raise ValueError(k)


def triroot(n):
    return (math.isqrt(1 + 8 * n) - 1) // 2


def identity(x, cone):
    """Barrier fixed point −f′(e) = e (pos.jl:44-55, noc.jl:264-267, soc.jl:155-167)."""
    k = cone[0]
    if k == "pos":
        x[:] = 1.0
    elif k == "free":
        x[:] = 0.0
    elif k == "soc":
        x[:] = 0.0
        x[0] = math.sqrt(2.0)
    elif k == "psd":
        x[:] = svec(np.eye(triroot(len(x))))  # sdp.jl:92-111
    elif k in ("exp", "pow"):
        from . import tdc
        x[:] = tdc.identity(cone)  # exp.jl:15-25, pow.jl:27-33
    else:
        raise NotImplementedError(k)


# ---------------------------------------------------------------------------------------
# PSD cone in svec coordinates (sdp.jl; smat/svec/skron: cone/utils.jl:475-664)
# ---------------------------------------------------------------------------------------

_R2 = math.sqrt(2.0)


def svec(M):
    """Lower triangle column by column, off-diagonal entries times √2 (utils.jl:570-597)."""
    d = M.shape[0]
    out = []
    for j in range(d):
        out.append(M[j, j])
        out.extend(_R2 * M[j + 1:, j])
    return np.array(out, dtype=np.float64)


def smat(v):
    """Inverse of svec, symmetric (utils.jl:631-664 then symmetrize!)."""
    d = triroot(len(v))
    M = np.zeros((d, d))
    k = 0
    for j in range(d):
        M[j, j] = v[k]
        M[j + 1:, j] = v[k + 1:k + d - j] / _R2
        k += d - j
    return np.tril(M) + np.tril(M, -1).T


def svec_index(d):
    """(i, j) with i ≥ j of each svec position."""
    return [(i, j) for j in range(d) for i in range(j, d)]


def skron(A):
    """H = A ⊗ₛ A in svec coordinates (utils.jl:475-513): for i ≥ j, k ≥ l,
    H[(i,j),(k,l)] = c_ij c_kl (A_ik A_jl + A_il A_jk) / 2 with c = 1 on the diagonal, √2 off it."""
    d = A.shape[0]
    idx = svec_index(d)
    I = np.array([a for a, _ in idx])
    J = np.array([b for _, b in idx])
    c = np.where(I == J, 1.0, _R2)
    H = (A[I][:, I] * A[J][:, J] + A[I][:, J] * A[J][:, I]) / 2.0
    return H * c[:, None] * c[None, :]


class PsdCache:
    """R = LP⁻ᵀU, S = LP U, s = Σ with LPᵀLD = UΣVᵀ (sdp.jl:10-30)."""
    __slots__ = ("R", "S", "s")

    def __init__(self, d):
        self.R = np.zeros((d, d))
        self.S = np.zeros((d, d))
        self.s = np.zeros(d)


def psdscale(H, cache, p, d):
    """sdpscaledynamic! (sdp.jl:212-262): Cholesky of P = smat p and D = smat d, the SVD of
    LPᵀ LD (the reference runs one-sided Jacobi, eig.jl:1-174; LAPACK's SVD gives the same
    factorization up to the order and signs of the singular triplets, which leave W and the
    cached products' uses invariant), then W = R Σ Rᵀ and H += W ⊗ₛ W."""
    try:
        LP = np.linalg.cholesky(smat(p))
        LD = np.linalg.cholesky(smat(d))
    except np.linalg.LinAlgError:
        return False, 0.0
    U, sv, _ = np.linalg.svd(LP.T @ LD)
    if not np.all(sv > 0):
        return False, 0.0
    import scipy.linalg as sla
    cache.S[:, :] = LP @ U
    cache.R[:, :] = sla.solve_triangular(LP.T, U, lower=False)
    cache.s[:] = sv
    Y = cache.R * np.sqrt(sv)[None, :]
    H += skron(Y @ Y.T)
    return True, float(np.sum(1.0 / sv ** 2))


def psdcorr(r, cache, dp, dd, smu):
    """sdpcorrdynamic! (sdp.jl:374-433): A = Rᵀ ΔP R, B = Sᵀ ΔD S, X = A B,
    Ŵ_ij = −(s_i X_ij + s_j X_ji)/(s_i + s_j) off the diagonal, σμ − s_j² − X_jj on it,
    r = svec(R Ŵ Rᵀ)."""
    R, S, sv = cache.R, cache.S, cache.s
    A = R.T @ smat(dp) @ R
    B = S.T @ smat(dd) @ S
    X = A @ B
    W = -(sv[:, None] * X + sv[None, :] * X.T) / (sv[:, None] + sv[None, :])
    W[np.diag_indices(len(sv))] = smu - sv ** 2 - np.diag(X)
    r[:] = svec(R @ W @ R.T)


def psdcorr0(r, cache, smu):
    """sdpcorr0dynamic! (sdp.jl:490-509): r = svec(R (σμ I − Σ²) Rᵀ)."""
    R, sv = cache.R, cache.s
    r[:] = svec((R * (smu - sv ** 2)[None, :]) @ R.T)


def psdmaxstep(F, dx, scale):
    """sdpmaxstepdynamic (sdp.jl:583-630): M = Fᵀ smat(Δx) F (divided by s_i s_j for the dual);
    τ = 1 if I + M is positive definite, else 1/max(1, −λmin(M)) (the reference finds λmin by
    Sturm bisection on a tridiagonal form, eig.jl:180-589; here LAPACK's eigvalsh)."""
    M = F.T @ smat(dx) @ F
    if scale is not None:
        M = M / (scale[:, None] * scale[None, :])
    try:
        np.linalg.cholesky(np.eye(M.shape[0]) + M)
        return 1.0
    except np.linalg.LinAlgError:
        pass
    lam = float(np.linalg.eigvalsh((M + M.T) / 2.0)[0])
    return 1.0 / max(1.0, -lam)


class SocCache:
    __slots__ = ("beta", "w", "pdet", "ddet", "pdetlo", "ddetlo")

    def __init__(self, n):
        self.beta = 0.0
        self.w = np.zeros(n)
        self.pdet = self.ddet = self.pdetlo = self.ddetlo = 0.0


def make_cache(cone, n):
    if cone[0] == "soc":
        return SocCache(n)
    if cone[0] == "psd":
        return PsdCache(triroot(n))
    if cone[0] in ("exp", "pow"):
        from . import tdc
        return tdc.TdCache(cone[0])
    return None


# ---------------------------------------------------------------------------------------
# scale (pos.jl:75-84, noc.jl:273-279, soc.jl:192-263)
# ---------------------------------------------------------------------------------------

def scale(cone, H, p, d, cache):
    """Add the NT scaling block to the (already Q-initialised) diagonal block H.
    Returns (flag, spsd)."""
    k = cone[0]
    if k == "pos":
        H[np.diag_indices(len(p))] += d / p
        return True, float(np.sum(1.0 / (p * d)))
    if k == "free":
        return True, 0.0
    if k == "soc":
        return socscale(H, cache, p, d)
    if k == "psd":
        return psdscale(H, cache, p, d)
    if k in ("exp", "pow"):
        from . import tdc
        return tdc.scale(cone, H, p, d, cache)
    raise NotImplementedError(k)


def socscale(H, cache, p, d):
    n = len(p)
    pdet, pdetlo = socdet(p)
    ddet, ddetlo = socdet(d)
    cache.pdet, cache.pdetlo, cache.ddet, cache.ddetlo = pdet, pdetlo, ddet, ddetlo
    flag = pdet > 0 and p[0] > 0 and ddet > 0 and d[0] > 0
    beta = spsd = 0.0
    if flag:
        pdot = cdot(p, d)
        spdet = math.sqrt(pdet)
        sddet = math.sqrt(ddet)
        kappa = math.sqrt(2.0 * (spdet * sddet + pdot))
        beta = math.sqrt(spdet / sddet)
        w = cache.w
        w[0] = (d[0] * beta + p[0] / beta) / kappa
        w[1:] = (-d[1:] * beta + p[1:] / beta) / kappa
        eta = 1.0 / beta ** 2
        # H += (2 (Jw)(Jw)ᵀ − J) / β²
        Jw = w.copy()
        Jw[1:] = -Jw[1:]
        J = np.ones(n)
        J[1:] = -1.0
        H += 2.0 * eta * np.outer(Jw, Jw)
        H[np.diag_indices(n)] -= eta * J
        spsd = 4.0 * pdot / (pdet * ddet)
    cache.beta = beta
    return flag, spsd


def socroot(x, w, beta, inverse):
    """y = √H x (inverse=False) or solve √H y = x (soc.jl:82-110), in place."""
    if not inverse:
        sigma, alpha = -1.0, 1.0 / beta
    else:
        sigma, alpha = 1.0, beta
    w1p1 = w[0] + 1.0
    wpdx = w1p1 * x[0] + sigma * float(np.dot(w[1:], x[1:]))
    wpdx /= w1p1
    x0 = alpha * (wpdx * w1p1 - x[0])
    x[1:] = alpha * (sigma * wpdx * w[1:] + x[1:])
    x[0] = x0
    return x


# ---------------------------------------------------------------------------------------
# correctors (pos.jl:107-137, noc.jl:285-306, soc.jl:282-379)
# ---------------------------------------------------------------------------------------

def corr0(cone, r, p, d, smu, cache):
    k = cone[0]
    if k == "pos":
        r[:] = smu / p - d
    elif k == "free":
        r[:] = 0.0
    elif k == "soc":
        c = cache
        eta = (1.0 / c.beta) ** 2
        wp, _ = socdot(c.w, p)
        t = 2.0 * eta * wp
        c1 = 2.0 * smu / c.pdet + eta
        r[0] = c1 * p[0] - t * c.w[0]
        r[1:] = t * c.w[1:] - c1 * p[1:]
    elif k == "psd":
        psdcorr0(r, cache, smu)
    elif k in ("exp", "pow"):
        from . import tdc
        tdc.corr0(r, d, cache, smu)
    else:
        raise NotImplementedError(k)


def corr(cone, r, p, d, dp, dd, smu, cache):
    k = cone[0]
    if k == "pos":
        r[:] = (smu - dp * dd) / p - d
    elif k == "free":
        r[:] = 0.0
    elif k == "soc":
        soccorr(r, cache.w, cache.beta, p, cache.pdet, dp, dd, smu)
    elif k == "psd":
        psdcorr(r, cache, dp, dd, smu)
    elif k in ("exp", "pow"):
        from . import tdc
        tdc.corr(cone, r, p, d, dp, dd, cache, smu)
    else:
        raise NotImplementedError(k)


def soccorr(r, w, beta, p, pdet, dp0, dd0, smu):
    n = len(p)
    dp = socroot(dp0.astype(np.float64).copy(), w, beta, False)
    dd = socroot(dd0.astype(np.float64).copy(), w, beta, True)
    # Jordan product dp ∘ dd
    r[0] = float(np.dot(dp, dd))
    r[1:] = dd[0] * dp[1:] + dp[0] * dd[1:]
    v = socroot(p.astype(np.float64).copy(), w, beta, False)
    vdet, _ = socdet(v)
    vr, _ = socdot(v, r)
    r1 = vr / vdet
    r[0] = r1
    r[1:] = (-r1 * v[1:] + r[1:]) * (1.0 / v[0])
    r += v
    socroot(r, w, beta, False)
    gamma = 2.0 * smu / pdet
    r[0] = gamma * p[0] - r[0]
    r[1:] = -gamma * p[1:] - r[1:]
    return r


# ---------------------------------------------------------------------------------------
# max steps (pos.jl:143-167, noc.jl:312-314, soc.jl:385-470)
# ---------------------------------------------------------------------------------------

def posmaxstep(x, dx):
    tau = 1.0
    neg = dx < 0
    if np.any(neg):
        tau = min(tau, float(np.min(-x[neg] / dx[neg])))
    return tau


def socmaxstep(x, dx, c, clo):
    eps = np.finfo(np.float64).eps
    a, alo = socdet(dx)
    b, blo = socdot(x, dx)
    b *= 2.0
    blo *= 2.0
    tau = 1.0
    if abs(a) > eps * max(abs(b), abs(c)):
        bb, bblo = twoprod(b, b)
        bblo = 2.0 * b * blo + bblo
        ac, aclo = twoprod(4.0 * a, c)
        aclo = 4.0 * a * clo + (4.0 * alo * c + aclo)
        s0, s0lo = twosum(bb, -ac)
        disc = s0 + ((bblo - aclo) + s0lo)
        if disc > -(eps * max(bb, abs(ac))):
            s = math.sqrt(max(disc, 0.0))
            q = -(b + math.copysign(s, b)) / 2.0
            tau1 = q / a if b >= 0 else c / q
            if tau1 > 0:
                tau = min(tau, tau1)
    elif b < 0.0:
        tau = min(tau, -c / b)
    if dx[0] < 0:
        tau = min(tau, -x[0] / dx[0])
    return tau


def maxsteps(cone, p, dp, d, dd, cache):
    k = cone[0]
    if k == "pos":
        return posmaxstep(p, dp), posmaxstep(d, dd)
    if k == "free":
        return 1.0, 1.0
    if k == "soc":
        return (socmaxstep(p, dp, cache.pdet, cache.pdetlo),
                socmaxstep(d, dd, cache.ddet, cache.ddetlo))
    if k == "psd":
        return psdmaxstep(cache.R, dp, None), psdmaxstep(cache.S, dd, cache.s)
    if k in ("exp", "pow"):
        from . import tdc
        return tdc.maxsteps(cone, p, dp, d, dd, cache)
    raise NotImplementedError(k)


# ---------------------------------------------------------------------------------------
# shadows (pos.jl:173-187, noc.jl:320-328, soc.jl:476-504)
# ---------------------------------------------------------------------------------------

def shadow(cone, s, x, primal=False):
    """s = −f′(x); returns flag (interior). For the exponential and power cones `primal`
    selects the inverse map: s with −f′(s) = x (primalshadow!, tdc.jl:471-474)."""
    k = cone[0]
    if k in ("exp", "pow"):
        from . import tdc
        return tdc.primal_shadow(cone, s, x) if primal else tdc.dual_shadow(cone, s, x)
    if k == "pos":
        s[:] = 1.0 / x
        return True
    if k == "free":
        s[:] = 0.0
        return True
    if k == "soc":
        xdet, _ = socdet(x)
        flag = xdet > 0 and x[0] > 0
        if flag:
            g = 2.0 / xdet
            s[0] = g * x[0]
            s[1:] = -g * x[1:]
        return flag
    if k == "psd":
        # sdpshadowdynamic! (sdp.jl:666-678): −f′(X) = X⁻¹ through the Cholesky factor
        try:
            Lc = np.linalg.cholesky(smat(x))
        except np.linalg.LinAlgError:
            return False
        Li = np.linalg.inv(Lc)
        s[:] = svec(Li.T @ Li)
        return True
    raise NotImplementedError(k)
