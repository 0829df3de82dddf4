"""GPU tests of mb_dense_cholesky: the frontal POTRF on a dense front held in HBM (caller-owned
device memory from the HIP runtime), checked against LAPACK's Cholesky (numpy) and by the backward error
‖LLᵀ − A‖/‖A‖; a matrix that is not positive definite reports failure as the reference's
issuccess(F) = false (uzawa.jl:93-97)."""
import numpy as np
import pytest

import mumblebee_hip as mh
from devmem import DeviceArray

pytestmark = pytest.mark.gpu


def spd(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return M @ M.T / n + np.eye(n)


def factor_on_device(A, lda, grid=0):
    """Column-major A with leading dimension lda in device memory (row j of the (n, lda) host
    image holds column j; the padding rows hold a sentinel), factored in place."""
    n = A.shape[0]
    buf = np.full((n, lda), 7.25)
    buf[:, :n] = A.T
    t = DeviceArray(buf)
    ok = mh.dense_cholesky(t.ptr, n, lda, grid)
    out = t.numpy()
    t.free()
    return ok, out[:, :n].T.copy(), out[:, n:]


@pytest.mark.parametrize("n,lda,grid", [(1, 1, 0), (63, 63, 0), (64, 70, 0), (65, 65, 0), (300, 301, 0),
                                        (1000, 1000, 0), (2000, 2003, 0), (777, 777, 7)])
def test_dense_cholesky_matches_lapack(n, lda, grid):
    print(f"dense cholesky n={n} lda={lda} grid={grid}", flush=True)
    A = spd(n, seed=n)
    ok, out, pad = factor_on_device(A, lda, grid)
    assert ok
    L = np.tril(out)
    Lref = np.linalg.cholesky(A)
    assert np.linalg.norm(L - Lref) / np.linalg.norm(Lref) < 1e-12
    assert np.linalg.norm(L @ L.T - A) / np.linalg.norm(A) < 10 * n * 2.2e-16
    iu = np.triu_indices(n, 1)
    assert np.array_equal(out[iu], A[iu])  # strict upper triangle untouched
    assert np.all(pad == 7.25)             # nothing written past row n of a column


def test_dense_cholesky_repeats_bit_for_bit():
    A = spd(900, seed=3)
    _, a, _ = factor_on_device(A, 900)
    _, b, _ = factor_on_device(A, 900, grid=13)
    assert np.array_equal(np.tril(a), np.tril(b))


def test_dense_cholesky_not_positive_definite():
    A = spd(700, seed=5)
    A[400, 400] = -1.0
    ok, _, _ = factor_on_device(A, 700)
    assert not ok


def test_dense_cholesky_rejects_bad_arguments():
    with pytest.raises(RuntimeError, match="lda"):
        mh.dense_cholesky(0, 10, 5)
