"""PSD cones on the device: whole solves against the oracle on SDPs with known structure."""
import numpy as np
import pytest
import scipy.sparse as sp

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem, rel

pytestmark = pytest.mark.gpu


def _svec_eye(d):
    v = []
    for j in range(d):
        v.append(1.0)
        v.extend([0.0] * (d - j - 1))
    return np.array(v)


def _trace_sdp(d, seed):
    """min tr(C X) s.t. tr(X) = 1, X ⪰ 0 (optimum λmin(C)); f = −svec(C)."""
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((d, d))
    Cm = (M + M.T) / 2
    r2 = np.sqrt(2.0)
    f = []
    for j in range(d):
        f.append(-Cm[j, j])
        f.extend(-r2 * Cm[j + 1:, j])
    n = d * (d + 1) // 2
    return dict(Q=sp.csc_matrix((n, n)), B=sp.csc_matrix(_svec_eye(d)[None, :]), f=np.array(f),
                g=np.array([1.0]), K=[("psd",)], s=[n]), np.linalg.eigvalsh(Cm)[0]


@pytest.mark.parametrize("d,seed", [(3, 1), (8, 2), (17, 3), (33, 4), (48, 5)])
def test_trace_sdp_attains_lambda_min(d, seed):
    data, lam = _trace_sdp(d, seed)
    res = mh.solve(hip_problem(data))
    assert res.status == mh.IPMStatus.OPTIMAL
    assert abs(res.pobj - lam) <= 1e-7 * max(1.0, abs(lam))


@pytest.mark.parametrize("kw", [dict(nv=40, band=4, clique=10, extra=0.05, seed=1),
                                dict(nv=120, band=8, clique=24, extra=0.02, seed=2),
                                dict(nv=120, band=20, clique=48, extra=0.001, seed=3)])
def test_maxcut_sdp_matches_oracle(kw):
    from oracle import ipm as O
    d = G.maxcut(**kw)
    res = mh.solve(hip_problem(d))
    ores = O.solve(oracle_problem(d))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8 and rel(res.dobj, ores.dobj) < 1e-8
    assert abs(res.niter - ores.niter) <= 2


def test_mixed_psd_soc_pos_problem():
    # an SDP block, second-order cones and a nonnegative block in one problem
    from oracle import ipm as O
    rng = np.random.default_rng(9)
    d = 6
    npsd = d * (d + 1) // 2
    K = [("psd",), ("soc",), ("soc",), ("pos",)]
    s = [npsd, 5, 7, 20]
    n = sum(s)
    m = 15
    B = sp.random(m, n, density=0.3, random_state=rng, data_rvs=rng.standard_normal, format="csc")
    x0 = np.concatenate([_svec_eye(d), [2.0, 0.3, -0.2, 0.1, 0.4], [3.0] + [0.2] * 6, np.ones(20)])
    y0 = rng.standard_normal(m)
    s0 = np.concatenate([_svec_eye(d), [2.0, 0.1, 0.1, -0.2, 0.3], [2.5] + [-0.1] * 6, np.ones(20)])
    data = dict(Q=sp.csc_matrix((n, n)), B=B, f=-(B.T @ y0 + s0), g=B @ x0, K=K, s=s)
    res = mh.solve(hip_problem(data))
    ores = O.solve(oracle_problem(data))
    assert res.status == ores.status == mh.IPMStatus.OPTIMAL
    assert rel(res.pobj, ores.pobj) < 1e-8
