"""mb_dense_cholesky_noalloc (the entry point that allocates no device memory) against
mb_dense_cholesky, the allocating entry point it varies: the same kernels on the same input give
the same factor bit for bit, over two calls in a row with the second one smaller (the shared
workspace keeps the first call's contents), and beside LAPACK's Cholesky."""
import numpy as np
import pytest

import mumblebee_hip as mh
from devmem import DeviceArray

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M @ M.T / n + np.eye(n)


def run(fn, A, lda):
    n = A.shape[0]
    buf = np.full((n, lda), -3.5)
    buf[:, :n] = A.T
    t = DeviceArray(buf)
    ok = fn(t.ptr, n, lda)
    out = t.numpy()
    t.free()
    return ok, out[:, :n].T.copy(), out[:, n:]


def test_noalloc_matches_the_allocating_entry_point():
    for n, lda, seed in [(1500, 1503, 1), (700, 700, 2), (65, 65, 3)]:
        print(f"no-alloc dense cholesky n={n} lda={lda}", flush=True)
        A = spd(n, seed)
        ok0, L0, pad0 = run(mh.dense_cholesky, A, lda)
        ok1, L1, pad1 = run(mh.dense_cholesky_noalloc, A, lda)
        assert ok0 and ok1
        assert np.array_equal(np.tril(L1), np.tril(L0))
        assert np.array_equal(L1, L0) and np.array_equal(pad1, pad0)
        Lref = np.linalg.cholesky(A)
        assert np.linalg.norm(np.tril(L1) - Lref) / np.linalg.norm(Lref) < 1e-12


def test_noalloc_reports_failure_then_recovers():
    A = spd(600, 4)
    A[300, 300] = -2.0
    ok, _, _ = run(mh.dense_cholesky_noalloc, A, 600)
    assert not ok
    B = spd(300, 5)  # the next, smaller call starts from cleared flags and scalars
    ok, L, _ = run(mh.dense_cholesky_noalloc, B, 300)
    assert ok
    assert np.linalg.norm(np.tril(L) - np.linalg.cholesky(B)) / np.linalg.norm(B) < 1e-12


def test_noalloc_rejects_orders_beyond_its_workspace():
    with pytest.raises(RuntimeError, match="workspace limit"):
        mh.dense_cholesky_noalloc(0, 8193, 8193)
