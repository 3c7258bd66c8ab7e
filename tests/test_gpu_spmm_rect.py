"""Rectangular and transposed CSR products on the GPU (ofrr_spmm_csr_rect,
ofrr_csr_transpose; csrmv.m:27-76) against the oracle: FP64 bitwise for both
directions, the transpose against scipy's row-stable transpose exactly."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rand(nrows, ncols, density, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(nrows, ncols, density=density, format="csr", random_state=rng)
    M = M.tolil()
    M[nrows // 3, :] = 0
    M[:, ncols // 2] = 0
    M = M.tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    return M


def _both(ctx, M):
    from mixedprecisionofrr_amd import api
    rp, ci, va = M.indptr.astype(np.int64), M.indices.astype(np.int32), M.data
    A = api.CSRMatrix.from_host(ctx, rp, ci, va, ncols=M.shape[1])
    return A, O.CSR(M.shape[0], rp, ci, va, ncols=M.shape[1])


@pytest.mark.parametrize("shape", [(3000, 1700), (900, 4111), (513, 1)])
@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("k", [1, 16, 61])
def test_spmm_rect_fp64_bitwise_both_directions(ctx, shape, layout, k):
    from mixedprecisionofrr_amd import api
    M = _rand(*shape, 0.01, shape[0] + k)
    A, Ah = _both(ctx, M)
    rng = np.random.default_rng(k)
    for trans in (False, True):
        xr = shape[0] if trans else shape[1]
        X = rng.standard_normal((xr, k))
        Xd = torch.as_tensor(X, device="cuda:0") if layout == "row" else \
            torch.as_tensor(np.ascontiguousarray(X.T), device="cuda:0").t()
        print(f"spmm_rect shape={shape} trans={trans} layout={layout} k={k}", flush=True)
        Y = api.spmm_rect(ctx, A, Xd, trans=trans).cpu().numpy()
        ref = O.spmm_t(Ah, X) if trans else O.spmm(Ah, X)
        assert np.array_equal(Y, ref), np.abs(Y - ref).max()


def test_spmm_rect_beta_and_fp32_output(ctx):
    from mixedprecisionofrr_amd import api
    M = _rand(700, 300, 0.02, 5)
    A, Ah = _both(ctx, M)
    X = np.random.default_rng(0).standard_normal((700, 8))
    Y0 = np.random.default_rng(1).standard_normal((300, 8))
    Y = torch.as_tensor(Y0.copy(), device="cuda:0")
    print("spmm_rect alpha/beta trans", flush=True)
    api.spmm_rect(ctx, A, torch.as_tensor(X, device="cuda:0"), trans=True, alpha=2.0, beta=-0.5, Y=Y)
    ref = 2.0 * (M.toarray().T @ X) - 0.5 * Y0
    assert np.max(np.abs(Y.cpu().numpy() - ref)) < 1e-12 * np.abs(ref).max()
    Y32 = api.spmm_rect(ctx, A, torch.as_tensor(X, device="cuda:0"), trans=True, out_dtype=torch.float32)
    r32 = O.spmm_t(Ah, X)
    assert np.max(np.abs(Y32.cpu().numpy() - r32)) <= 2 ** -23 * np.abs(r32).max() * 2


def test_transpose_matches_scipy(ctx):
    from mixedprecisionofrr_amd import api
    M = _rand(2500, 777, 0.01, 11)
    A, _ = _both(ctx, M)
    print("csr_transpose 2500x777", flush=True)
    T = api.transpose(ctx, A)
    S = M.T.tocsr()
    assert (T.n, T.ncols) == (777, 2500)
    assert np.array_equal(T.row_ptr.cpu().numpy(), S.indptr)
    assert np.array_equal(T.col_idx.cpu().numpy(), S.indices)
    assert np.array_equal(T.vals.cpu().numpy(), S.data)
    # the transpose of the transpose is A again
    TT = api.transpose(ctx, T)
    assert np.array_equal(TT.row_ptr.cpu().numpy(), M.indptr)
    assert np.array_equal(TT.col_idx.cpu().numpy(), M.indices)
    assert np.array_equal(TT.vals.cpu().numpy(), M.data)


def test_transpose_of_a_stencil_is_itself(ctx):
    from mixedprecisionofrr_amd import api
    A = api.CSRMatrix.stencil(ctx, 3, 20, seed=0)
    print("csr_transpose of a symmetric stencil", flush=True)
    T = api.transpose(ctx, A)
    assert torch.equal(T.row_ptr, A.row_ptr) and torch.equal(T.col_idx, A.col_idx)
    assert torch.equal(T.vals, A.vals)


def test_transpose_reports_a_bad_column(ctx):
    from mixedprecisionofrr_amd import api
    dev = "cuda:0"
    rp = torch.tensor([0, 1, 2], dtype=torch.int64, device=dev)
    ci = torch.tensor([0, 5], dtype=torch.int32, device=dev)   # 5 >= ncols = 3
    va = torch.ones(2, dtype=torch.float64, device=dev)
    A = api.CSRMatrix(2, rp, ci, va, ncols=3)
    print("csr_transpose with an out-of-range column", flush=True)
    with pytest.raises(api.OfrrError) as e:
        api.transpose(ctx, A)
    assert e.value.status == 1000
