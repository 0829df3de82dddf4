"""Mixed precision on the device: the large fronts factored in FP32 on FP32 MFMA, refined in
FP64 by the Uzawa/CG loop, with the factor back in FP64 once μ < mixed_mu. Each result is
compared with the plain FP64 solve of the same instance."""
import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, rel

pytestmark = pytest.mark.gpu


def test_mfma_f32_lane_layout():
    """v_mfma_f32_16x16x4f32: A[i][k] from lane i+16k, B[k][j] from lane j+16k, D[i][j] in
    lane j+16(i//4), element i%4 — the layout the FP32 factorization assumes."""
    out = mh.probe_mfma_f32().reshape(64, 4)
    A = np.array([[i + 100 * k for k in range(4)] for i in range(16)], dtype=float)
    B = np.array([[j + 1000 * k for j in range(16)] for k in range(4)], dtype=float)
    D = A @ B
    exp = np.zeros((64, 4))
    for lane in range(64):
        for t in range(4):
            exp[lane, t] = D[4 * (lane // 16) + t, lane % 16]
    if not np.array_equal(out, exp):
        print("actual lane outputs:\n", out)
    assert np.array_equal(out, exp)


CASES = [
    dict(kind="qp", n=2000, m=1000, seed=1000, big=512),    # one 2000×2000 front
    dict(kind="qp", n=600, m=300, seed=7, big=128),        # a tree with several large fronts
    dict(kind="socp", n=1500, m=700, seed=8, big=128),
]


def _data(c):
    if c["kind"] == "qp":
        return G.qp(n=c["n"], m=c["m"], seed=c["seed"])
    return G.socp(n=c["n"], m=c["m"], ncone=max(1, c["n"] // 40), seed=c["seed"], adens=0.01)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_mixed_matches_fp64(case):
    c = CASES[case]
    print("case", c, flush=True)
    d = _data(c)
    ref = mh.init(hip_problem(d), mh.IPMSettings(big_front_ld=c["big"]))
    r64 = ref.solve()
    s = mh.init(hip_problem(d), mh.IPMSettings(big_front_ld=c["big"], precision=1))
    r32 = s.solve()
    st = s.stats()
    s.close()
    ref.close()
    assert r64.status == mh.IPMStatus.OPTIMAL
    assert r32.status == mh.IPMStatus.OPTIMAL
    assert st["factor_calls_fp32"] > 0
    # objective, residuals and gap within 1e-8 relative of the FP64 solve
    assert rel(r32.pobj, r64.pobj) < 1e-8 and rel(r32.dobj, r64.dobj) < 1e-8
    assert r32.pres < 1e-8 and r32.dres < 1e-8
    print("iterations fp64", r64.niter, "mixed", r32.niter, "fp32 factors", st["factor_calls_fp32"],
          "of", st["factor_calls"])
