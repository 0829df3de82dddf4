"""Shared test helpers: build the same seeded problem for the HIP path and for the oracle."""
import numpy as np

import mumblebee_hip as mh
from mumblebee_hip import generators as G

KMAP = {"pos": mh.PositiveCone, "free": mh.CofreeCone, "soc": mh.SecondOrderCone,
        "psd": mh.SemidefiniteCone, "exp": mh.ExponentialCone}


def cones_of(K):
    out = []
    for k in K:
        if k[0] == "pow":
            out.append(mh.PowerCone(k[1]))
        else:
            out.append(KMAP[k[0]]())
    return out


def hip_problem(d, mu=0.0, compress=1.0):
    return mh.IPMProblem(d["Q"], d["B"], d["f"], d["g"], mu, cones_of(d["K"]), d["s"], compress=compress)


def oracle_problem(d, mu=0.0, compress=1.0):
    from oracle.problem import CSC, make_problem
    return make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], mu, d["K"], d["s"],
                        compress=compress)


def small_cases():
    return [
        G.qp(n=120, m=60, seed=11, pdens=0.03, adens=0.05),
        G.qp(n=300, m=150, seed=12),
        G.socp(n=400, m=200, ncone=15, seed=13, adens=0.02),
        G.socp(n=900, m=450, ncone=40, seed=14, adens=0.01),
    ]


def rel(a, b):
    return abs(a - b) / max(1.0, abs(b))


def kkt_certificate(d, p, dd, y):
    """Solver-independent residuals of the original problem: ‖Bp−g‖, ‖Qp−f−Bᵀy−d‖ (relative),
    gap and the min cone eigenvalue of p and d."""
    Q, B, f, g = d["Q"], d["B"], d["f"], d["g"]
    rp = np.linalg.norm(B @ p - g) / (1 + np.linalg.norm(g))
    rd = np.linalg.norm(Q @ p - f - B.T @ y - dd) / (1 + np.linalg.norm(f))
    return rp, rd
