"""Writes the golden fixtures of the full-size workloads that the oracle takes minutes to solve:
inputs are regenerated from the seeded generators (mumblebee_hip/generators.py), so a fixture
holds only the oracle's outputs (status, iterations, objectives, residuals, and the iterate).

    python tests/golden/make_golden.py

socp10000_seed1000.npz: the SOCP workload at full size (n = 10000, m = 5000, 500 second-order
cones and a nonnegative block), solved by the oracle (oracle/ipm.py, FP64, minimum-degree order)."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "mumblebee.jl_amd"))

from mumblebee_hip import generators as G  # noqa: E402
from oracle import ipm as O  # noqa: E402
from oracle.problem import CSC, make_problem  # noqa: E402


def socp_full():
    d = G.socp(seed=1000)
    t = time.time()
    prob = make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])
    r = O.solve(prob)
    print(f"socp n=10000: {O.STATUS_NAMES[r.status]} in {r.niter} iterations, {time.time() - t:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "socp10000_seed1000.npz"), status=r.status, niter=r.niter,
                        pobj=r.pobj, dobj=r.dobj, pres=r.pres, dres=r.dres, mu=r.mu, p=r.p, d=r.d, y=r.y)


if __name__ == "__main__":
    socp_full()
