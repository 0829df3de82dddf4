"""CPU tests of the native host side: the C++ problem construction and symbolic analysis agree
bit for bit with the oracle; structural invariants of the fronts; the C ABI exports every
symbol the header declares."""
import ctypes
import os
import re

import numpy as np
import pytest

import mumblebee_hip as mh
from mumblebee_hip import generators as G
from helpers import hip_problem, oracle_problem
from oracle.symbolic import analyse, vertex_graph

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "mumblebee_hip.h")).read()
    names = sorted(set(re.findall(r"\b(mb_[a-z0-9_]+)\s*\(", hdr)))
    L = ctypes.CDLL(mh.LIBPATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(mh.EXPORTS), set(names) - set(mh.EXPORTS)


CASES = [(G.qp, dict(n=150, m=70, seed=1)),
         (G.qp, dict(n=300, m=100, seed=2, pdens=0.01, adens=0.02)),
         (G.socp, dict(n=500, m=240, ncone=18, seed=3, adens=0.02)),
         (G.socp, dict(n=800, m=300, ncone=30, seed=4, adens=0.008))]


@pytest.mark.parametrize("gen,kw", CASES)
def test_problem_construction_bit_exact(gen, kw):
    d = gen(**kw)
    e = hip_problem(d).export()
    o = oracle_problem(d)
    for key, ref in [("P1", o.P1), ("P2", o.P2), ("B_xsrc", o.B.xsrc), ("B_xcol", o.B.xcol), ("B_xrow", o.B.xrow),
                     ("B_xblk", o.B.xblk), ("B_tgt", o.B.tgt), ("B_val", o.B.val), ("Q_xsrc", o.Q.xsrc),
                     ("Q_xblk", o.Q.xblk), ("Q_tgt", o.Q.tgt), ("Q_val", o.Q.val), ("f", o.f), ("g", o.g)]:
        np.testing.assert_array_equal(e[key], ref, err_msg=key)
    assert [int(k) for k in e["kind"]] == [{"pos": 0, "free": 1, "soc": 2}[k[0]] for k in o.K]


def test_problem_construction_with_compression_and_free_cones():
    rng = np.random.default_rng(9)
    import scipy.sparse as sp
    n, m = 40, 15
    A = sp.random(m, n, density=0.15, random_state=rng, format="csc")
    A[:, 5] = A[:, 3]  # twins
    A[:, 30] = A[:, 31]
    d = dict(Q=sp.identity(n, format="csc") * 0.0, B=sp.csc_matrix(A), f=rng.standard_normal(n),
             g=rng.standard_normal(m), K=[("pos",), ("free",), ("soc",), ("pos",)], s=[10, 10, 5, 15])
    for tau in (1.0, 0.6):
        e = hip_problem(d, compress=tau).export()
        o = oracle_problem(d, compress=tau)
        np.testing.assert_array_equal(e["P2"], o.P2)
        np.testing.assert_array_equal(e["P1"], o.P1)
        np.testing.assert_array_equal(e["B_val"], o.B.val)
        np.testing.assert_array_equal(e["B_xcol"], o.B.xcol)


@pytest.mark.parametrize("gen,kw", CASES)
@pytest.mark.parametrize("amalg", [False, True])
def test_symbolic_bit_exact(gen, kw, amalg):
    d = gen(**kw)
    s = hip_problem(d).symbolic(amalg)
    o = oracle_problem(d)
    w, adj = vertex_graph(o.B, o.Q)
    os_ = analyse(w, adj, amalgamate=amalg)
    np.testing.assert_array_equal(s["vperm"], os_.vperm)
    np.testing.assert_array_equal(s["front_vtx"], os_.front_vtx)
    np.testing.assert_array_equal(s["parent"], os_.front_parent)
    seps = [s["sep"][s["sep_ptr"][f]:s["sep_ptr"][f + 1]].tolist() for f in range(len(s["parent"]))]
    assert seps == [list(map(int, x)) for x in os_.front_sep]
    assert s["exact_flops"] == os_.exact_flops


@pytest.mark.parametrize("gen,kw", CASES[:3])
@pytest.mark.parametrize("amalg", [False, True])
def test_symbolic_structure_invariants(gen, kw, amalg):
    """Fronts partition the vertices into contiguous ranges; parents come later; each separator
    lies in the ancestors; the fronts cover the fill of brute-force elimination in the new order."""
    d = gen(**kw)
    o = oracle_problem(d)
    w, adj = vertex_graph(o.B, o.Q)
    S = analyse(w, adj, amalgamate=amalg)
    N = len(w)
    assert sorted(S.vperm.tolist()) == list(range(N))
    fv = S.front_vtx
    assert fv[0] == 0 and fv[-1] == N and np.all(np.diff(fv) > 0)
    vfront = np.empty(N, dtype=int)
    for f in range(S.nfront):
        vfront[fv[f]:fv[f + 1]] = f
        if S.front_parent[f] >= 0:
            assert S.front_parent[f] > f
    # ancestors
    for f in range(S.nfront):
        anc = set()
        p = S.front_parent[f]
        while p >= 0:
            anc.add(p)
            p = S.front_parent[p]
        for u in S.front_sep[f]:
            assert vfront[u] in anc
            assert u >= fv[f + 1]
    # brute-force elimination of the vertex graph in the new order
    inv = np.empty(N, dtype=int)
    inv[S.vperm] = np.arange(N)
    nadj = [set(int(inv[u]) for u in adj[S.vperm[k]]) for k in range(N)]
    struct = [set() for _ in range(N)]
    for k in range(N):
        later = {u for u in nadj[k] if u > k}
        struct[k] = later
        for a in later:
            nadj[a] |= later - {a}
    for k in range(N):
        f = vfront[k]
        cover = set(range(k + 1, fv[f + 1])) | set(int(u) for u in S.front_sep[f])
        assert struct[k] <= cover


def test_block_sparse_input_is_taken_as_given():
    """mb_problem_create_bsm keeps the caller's block-sparse Q and B (identity permutations) and
    rejects inconsistent layouts with a message."""
    rng = np.random.default_rng(3)
    widths, heights = [2, 3, 1, 4], [2, 1, 3]
    B = mh.blocksparse([0, 1, 2, 0, 2], [0, 1, 1, 3, 2],
                       [rng.standard_normal((2, 2)), rng.standard_normal((1, 3)), rng.standard_normal((3, 3)),
                        rng.standard_normal((2, 4)), rng.standard_normal((3, 1))], heights, widths)
    Q = mh.blocksparse([v for v in range(4)], [v for v in range(4)], [np.eye(w) for w in widths], widths, widths)
    K = [mh.PositiveCone(), mh.SecondOrderCone(), mh.CofreeCone(), mh.PositiveCone()]
    P = mh.IPMProblem.from_blocks(Q, B, np.arange(10.0), np.arange(6.0), 0.0, K)
    e = P.export()
    for key in ("xsrc", "xcol", "xrow", "xblk", "tgt", "val"):
        assert np.array_equal(e["B_" + key], getattr(B, key)), key
    assert np.array_equal(e["Q_val"], Q.val)
    assert np.array_equal(e["P1"], np.arange(6)) and np.array_equal(e["P2"], np.arange(10))
    assert np.array_equal(e["f"], np.arange(10.0))
    assert np.allclose(B.toscipy().toarray()[:2, :2], B.val[:4].reshape(2, 2, order="F"))
    bad = mh.BlockSparse(B.nout, B.nvtx, B.xsrc, B.xcol, B.xrow, B.xblk, B.tgt[::-1].copy(), B.val)
    with pytest.raises(RuntimeError, match="not strictly increasing|nrows"):
        mh.IPMProblem.from_blocks(Q, bad, np.zeros(10), np.zeros(6), 0.0, K)
