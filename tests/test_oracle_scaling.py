"""Pins the oracle's block-aware Ruiz equilibration (scaling.jl:52-184) by its invariants."""
import numpy as np
import pytest

from oracle.problem import CSC, make_problem
from oracle.scaling import equilibrate, infnorms
from mumblebee_hip import generators as G


def _problem(kind, seed):
    d = G.qp(n=80, m=40, seed=seed, pdens=0.05, adens=0.08) if kind == "qp" else \
        G.socp(n=120, m=50, ncone=6, seed=seed, adens=0.06)
    return make_problem(CSC.from_scipy(d["Q"]), CSC.from_scipy(d["B"]), d["f"], d["g"], 0.0, d["K"], d["s"])


@pytest.mark.parametrize("kind,seed", [("qp", 1), ("qp", 2), ("socp", 3)])
def test_equilibrate_drives_block_norms_to_one(kind, seed):
    # after convergence every nonzero row ∞-norm and vertex-block ∞-norm is within a few tol of 1
    P = _problem(kind, seed)
    f, g = P.f.copy(), P.g.copy()
    equilibrate(P.B, P.Q, f, g, itmax=200, tol=1e-6)
    rn, vn = infnorms(P.B, P.Q)
    rn, vn = np.sqrt(rn[rn > 0]), np.sqrt(vn[vn > 0])
    assert np.abs(rn - 1).max() < 1e-5 and np.abs(vn - 1).max() < 1e-5


@pytest.mark.parametrize("kind,seed", [("qp", 4), ("socp", 5)])
def test_equilibrate_is_a_diagonal_congruence(kind, seed):
    # B ← E B D, Q ← D Q D, f ← D f, g ← E g with D constant on each cone block, so unscaling
    # with the returned pscl, yscl recovers the data exactly (up to rounding)
    P = _problem(kind, seed)
    B0, Q0 = P.B.to_dense(), P.Q.to_dense()
    f0, g0 = P.f.copy(), P.g.copy()
    f, g = P.f.copy(), P.g.copy()
    pscl, yscl = equilibrate(P.B, P.Q, f, g, itmax=10, tol=1e-3)
    np.testing.assert_allclose(P.B.to_dense(), yscl[:, None] * B0 * pscl[None, :], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(P.Q.to_dense(), pscl[:, None] * Q0 * pscl[None, :], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(f, pscl * f0, rtol=1e-14)
    np.testing.assert_allclose(g, yscl * g0, rtol=1e-14)
    for v in range(P.B.nvtx):
        cols = list(P.B.colrange(v))
        assert np.all(pscl[cols] == pscl[cols[0]]) if cols else True
