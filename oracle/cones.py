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
    else:
        raise NotImplementedError(k)


class SocCache:
    __slots__ = ("beta", "w", "pdet", "ddet", "pdetlo", "ddetlo")

    def __init__(self, n):
        self.beta = 0.0
        self.w = np.zeros(n)
        self.pdet = self.ddet = self.pdetlo = self.ddetlo = 0.0


def make_cache(cone, n):
    if cone[0] == "soc":
        return SocCache(n)
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
    raise NotImplementedError(k)


# ---------------------------------------------------------------------------------------
# shadows (pos.jl:173-187, noc.jl:320-328, soc.jl:476-504)
# ---------------------------------------------------------------------------------------

def shadow(cone, s, x):
    """s = −f′(x); returns flag (interior)."""
    k = cone[0]
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
    raise NotImplementedError(k)
