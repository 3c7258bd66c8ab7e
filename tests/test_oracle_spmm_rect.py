"""Pins of the oracle's rectangular and transposed CSR products (csrmv.m:27-76)
and the host-side shape refusals of the Python layer.  CPU only."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import oracle as O


def _rand_csr(nrows, ncols, density, seed, integer=False):
    rng = np.random.default_rng(seed)
    M = sp.random(nrows, ncols, density=density, format="csr", random_state=rng,
                  data_rvs=(lambda s: rng.integers(-4, 5, s).astype(float)) if integer else None)
    M.sort_indices()
    # a few empty rows and columns
    M = M.tolil()
    M[nrows // 3, :] = 0
    M[:, ncols // 2] = 0
    M = M.tocsr()
    M.eliminate_zeros()
    return O.CSR(nrows, M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data, ncols=ncols), M


@pytest.mark.parametrize("nrows,ncols", [(300, 170), (90, 411), (1, 7), (64, 64)])
def test_spmm_t_is_the_row_stable_transpose_product(nrows, ncols):
    A, M = _rand_csr(nrows, ncols, 0.05, nrows + ncols)
    X = np.random.default_rng(1).standard_normal((nrows, 5))
    Y = O.spmm_t(A, X)
    # scipy's CSR -> CSC conversion visits rows in increasing order: each row of
    # the transpose lists its entries in increasing row of A (the 'T' order)
    T = M.T.tocsr()
    for j in range(T.shape[0]):
        seg = T.indices[T.indptr[j]:T.indptr[j + 1]]
        assert np.all(np.diff(seg) > 0)
    At = O.CSR(ncols, T.indptr.astype(np.int64), T.indices.astype(np.int32), T.data, ncols=nrows)
    assert np.array_equal(Y, O.spmm(At, X))
    assert np.max(np.abs(Y - M.toarray().T @ X)) <= 1e-13 * max(1.0, np.abs(M).max() * np.abs(X).max())


def test_spmm_rect_integer_exact_and_adjoint():
    A, M = _rand_csr(200, 120, 0.08, 3, integer=True)
    rng = np.random.default_rng(2)
    X = rng.integers(-9, 10, (120, 3)).astype(float)
    Y = rng.integers(-9, 10, (200, 3)).astype(float)
    D = M.toarray()
    assert np.array_equal(O.spmm(A, X), D @ X)
    assert np.array_equal(O.spmm_t(A, Y), D.T @ Y)
    Xr = rng.standard_normal((120, 1))
    Yr = rng.standard_normal((200, 1))
    lhs = float(Yr[:, 0] @ O.spmm(A, Xr)[:, 0])
    rhs = float(O.spmm_t(A, Yr)[:, 0] @ Xr[:, 0])
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(D).sum() * 10)


def test_planted_transpose_mistake_is_caught():
    # A^T X with X read as if A were square (the untransposed product) differs
    A, M = _rand_csr(150, 150, 0.05, 9)
    X = np.random.default_rng(4).standard_normal((150, 2))
    assert not np.allclose(O.spmm_t(A, X), O.spmm(A, X))


# ---- host-side refusals of the Python layer (no GPU call is reached) --------
def _host_csr(n, ncols):
    from mixedprecisionofrr_amd import api
    rp = torch.zeros(n + 1, dtype=torch.int64)
    ci = torch.zeros(0, dtype=torch.int32)
    va = torch.zeros(0, dtype=torch.float64)
    return api.CSRMatrix(n, rp, ci, va, ncols=ncols)


def test_spmm_refuses_wrong_row_counts():
    from mixedprecisionofrr_amd import api
    A = _host_csr(10, 6)
    with pytest.raises(ValueError, match="rows"):
        api.spmm(None, A, torch.zeros(10, 2, dtype=torch.float64))
    with pytest.raises(ValueError, match="rows"):
        api.spmm_rect(None, A, torch.zeros(6, 2, dtype=torch.float64), trans=True)
    with pytest.raises(ValueError, match="Y must be"):
        api.spmm(None, A, torch.zeros(6, 2, dtype=torch.float64), Y=torch.zeros(6, 2, dtype=torch.float64))


def test_square_only_entry_points_refuse_rectangular_operators():
    from mixedprecisionofrr_amd import api
    A = _host_csr(10, 6)
    X = torch.zeros(10, 4, dtype=torch.float64)
    for call in (lambda: api.filtered_eig(None, A, X, 2),
                 lambda: api.residuals(None, A, X, [1.0] * 4, 1.0),
                 lambda: api.subspace_iter_eig(None, A, X, 1),
                 lambda: api.arnoldi_hess_eig(None, A, 2, 4, 1, torch.zeros(10))):
        with pytest.raises(ValueError, match="square"):
            call()
    S = _host_csr(10, 10)
    with pytest.raises(ValueError, match="rows"):
        api.filtered_eig(None, S, torch.zeros(9, 4, dtype=torch.float64), 2)
    with pytest.raises(ValueError, match="rows"):
        api.residuals(None, S, torch.zeros(11, 4, dtype=torch.float64), [1.0] * 4, 1.0)


def test_csr_structure_checks():
    from mixedprecisionofrr_amd import api
    rp = np.array([0, 2, 3]); ci = np.array([0, 5], np.int32)
    with pytest.raises(ValueError, match="differ"):
        api._check_csr(rp, ci, np.zeros(3), 2, 6)
    with pytest.raises(ValueError, match="row pointers"):
        api._check_csr(np.array([0, 2, 1]), np.array([0, 1], np.int32), np.zeros(2), 2, 6)
    with pytest.raises(ValueError, match="column indices"):
        api._check_csr(np.array([0, 1, 2]), np.array([0, 6], np.int32), np.zeros(2), 2, 6)
    api._check_csr(np.array([0, 1, 2]), np.array([0, 5], np.int32), np.zeros(2), 2, 6)
