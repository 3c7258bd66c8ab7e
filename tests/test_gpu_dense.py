"""Dense operators on the GPU: MFMA products (gemm over dense operands,
mvops.cu:423-605 / gemm.m:27-59), filtered OFRR on a dense symmetric A
(si_eig over a dense Matrix, si_eig.hpp:24-27) and the truncated SVD by
OFRR on A^T A (si_svd / subspace_iter_svd.m; the dense low-rank-plus-noise SVD workload at a
reduced size).

References: FP64 products of the same (rounded) operands, the closed-form
Laplacian spectrum, and numpy's LAPACK SVD of the same FP32 matrix."""
import numpy as np
import pytest
import torch

from mixedprecisionofrr_amd import api
from oracle import oracle as O

pytestmark = pytest.mark.gpu
dev = "cuda:0"


def _bound(P, Q):
    return (P.double().abs() @ Q.double().abs())


@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("R,K,nc", [(1000, 300, 64), (64, 65536, 32), (4097, 4096, 128), (100, 40, 16)])
def test_dense_gemm_low(ctx, dt, R, K, nc):
    g = torch.Generator(device="cpu").manual_seed(R + K + nc)
    ldp = (K + 7) // 8 * 8
    Pf = torch.randn(R, ldp, generator=g)
    P = Pf.to(dt).to(dev)[:, :K]
    Q = torch.randn(K, nc, generator=g).to(dt).to(dev)
    C = api.dense_gemm(ctx, P, Q, torch.float32)
    torch.cuda.synchronize()
    ref = P.double() @ Q.double()
    err = (C.double() - ref).abs()
    # FP32 accumulation of exact products: |err| <= K u32 sum |p||q| (worst case)
    assert torch.all(err <= 2.0 ** -24 * K * _bound(P, Q) + 1e-30), float(err.max())
    # typical error is far below the worst case
    assert float((err / (_bound(P, Q) + 1e-30)).max()) < 1e-5


@pytest.mark.parametrize("pdt", [torch.float64, torch.float32])
@pytest.mark.parametrize("R,K,nc", [(1000, 300, 64), (64, 65536, 32), (333, 4097, 48)])
def test_dense_gemm_f64(ctx, pdt, R, K, nc):
    g = torch.Generator(device="cpu").manual_seed(7 * R + K)
    P = torch.randn(R, K, generator=g, dtype=torch.float64).to(pdt).to(dev)
    Q = torch.randn(K, nc, generator=g, dtype=torch.float64).to(dev)
    C = api.dense_gemm(ctx, P, Q, torch.float64)
    Cb = api.dense_gemm(ctx, P, Q, torch.float64, C=C.clone(), alpha=2.0, beta=-1.0)
    torch.cuda.synchronize()
    ref = P.double() @ Q
    tol = 1e-14 * _bound(P, Q)
    assert torch.all((C - ref).abs() <= tol), float((C - ref).abs().max())
    assert torch.all((Cb - ref).abs() <= 3 * tol)


def test_dense_gemm_rejects(ctx):
    P = torch.zeros(16, 10, dtype=torch.bfloat16, device=dev)
    Q = torch.zeros(10, 16, dtype=torch.float32, device=dev)
    with pytest.raises(RuntimeError):
        api.dense_gemm(ctx, P, Q)  # mixed operand types
    Q2 = torch.zeros(10, 24, dtype=torch.bfloat16, device=dev)
    with pytest.raises(RuntimeError):
        api.dense_gemm(ctx, P, Q2)  # nc not a multiple of 16


def _dense_laplacian_2d(N):
    A = O.stencil(2, N)
    D = np.zeros((A.n, A.n))
    for i in range(A.n):
        for p in range(A.row_ptr[i], A.row_ptr[i + 1]):
            D[i, A.col_idx[p]] = A.vals[p]
    return D


@pytest.mark.parametrize("bdt", [torch.float64, torch.float16, torch.bfloat16])
def test_filtered_eig_dense_laplacian(ctx, bdt):
    N = 32
    D = torch.from_numpy(_dense_laplacian_2d(N)).to(dev)
    n, k, nwant = N * N, 32, 16
    X0 = torch.empty(n, k, dtype=torch.float64, device=dev)
    api.fill_randn(ctx, X0, seed=3)
    st, th, X, res, stats = api.filtered_eig_dense(ctx, D, X0, nwant, nd=1, degree=12, maxiter=100,
                                                   basis_dtype=bdt, tol=1e-8)
    assert st == 0, stats
    exact = O.laplacian_eigs(2, N)[:nwant]
    assert np.max(np.abs(th[:nwant] - exact) / exact) < 1e-12
    assert np.max(res[:nwant]) <= 1e-8
    # independent residual of the returned block
    Xd = X.double()
    R = D @ Xd - Xd * torch.from_numpy(th).to(dev)[None, :]
    rr = (R.norm(dim=0) / (exact[0] * Xd.norm(dim=0))).cpu().numpy()
    assert np.max(rr[:nwant]) <= 2e-8


def _lowrank_matrix(M, N, decay, noise, seed):
    g = np.random.default_rng(seed)
    U, _ = np.linalg.qr(g.standard_normal((M, N)))
    W, _ = np.linalg.qr(g.standard_normal((N, N)))
    s = np.exp(-np.arange(N) / decay)
    A = (U * s[None, :]) @ W.T + noise * g.standard_normal((M, N))
    return A.astype(np.float32)


@pytest.mark.parametrize("bdt", [torch.bfloat16, torch.float64])
def test_svd_filtered_lowrank(ctx, bdt):
    M, N, k, nwant = 8192, 512, 32, 20
    A32 = _lowrank_matrix(M, N, 32.0, 1e-6, 11)
    ref = np.linalg.svd(A32.astype(np.float64), compute_uv=False)[:nwant]
    # column-major storage on the device: the transpose of a row-major N x M tensor
    Acm = torch.from_numpy(np.ascontiguousarray(A32.T)).to(dev).t()
    V0 = torch.empty(N, k, dtype=torch.float64, device=dev)
    api.fill_randn(ctx, V0, seed=5)
    st, sig, U, V, res, stats = api.svd_filtered(ctx, Acm, V0, nwant, nd=1, degree=16, maxiter=60,
                                                 basis_dtype=bdt, tol=1e-8)
    assert st == 0, stats
    assert np.max(np.abs(sig[:nwant] - ref) / ref) < 1e-9
    assert np.max(res[:nwant]) <= 1e-8
    # left vectors: A v = sigma u, unit norm
    Ad = torch.from_numpy(A32.astype(np.float64)).to(dev)
    AV = Ad @ V[:, :nwant]
    assert float((AV - U[:, :nwant] * torch.from_numpy(sig[:nwant]).to(dev)[None, :]).abs().max()) < 1e-10
    assert np.allclose(U[:, :nwant].norm(dim=0).cpu().numpy(), 1.0, atol=1e-8)


@pytest.mark.parametrize("dt,drops", [(torch.float64, (9,)), (torch.float32, (9,)), (torch.float16, (9,)),
                                      (torch.float64, (15, 16, 31, 39)),
                                      (torch.float16, (15, 16, 31, 39))])
def test_hessenberg_v3_and_steps(ctx, dt, drops):
    """C++ v3 (dropped columns still eliminated) against the oracle's variant 1,
    and the column-by-column step entry point bit for bit equal to the full
    factorizations of both variants.  The full factorization runs in panels of
    16 columns: the second case drops the last column of a panel, the first of
    the next, and the last column of all."""
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(44)
    A = rng.uniform(-1, 1, (700, 40))
    for d in drops:  # dropped at select_tol 1e-6
        A[:, d] = A[:, d - 6] + 1e-9 * rng.uniform(-1, 1, 700)
    tol = 1e-6 if dt != torch.float16 else 1e-2
    Ad = torch.as_tensor(A, device="cuda:0").to(dt)
    for variant in (0, 1):
        print(f"hessenberg variant={variant} dtype={dt} full and by steps", flush=True)
        Q, R, perm, skip = api.hessenberg(ctx, Ad, tol, variant=variant)
        Qs, Rs, prow, skips = api.hessenberg_by_steps(ctx, Ad, tol, variant=variant)
        assert torch.equal(Q, Qs), (Q.double() - Qs.double()).abs().max()
        assert torch.equal(skip, skips)
        assert np.array_equal(np.triu(R.cpu().numpy()), np.triu(Rs.cpu().numpy()))
        kept = skip.cpu().numpy() == 0
        assert np.array_equal(perm.cpu().numpy()[kept], prow.cpu().numpy()[kept] - 1)
        if dt == torch.float64:
            Qo, Ro, so, po = O.hessenberg(A, tol, variant=variant)
            assert list(np.nonzero(so == 0)[0]) == list(drops)
            assert np.array_equal(Q.cpu().numpy(), Qo)
            assert np.array_equal(R.cpu().numpy(), Ro)
