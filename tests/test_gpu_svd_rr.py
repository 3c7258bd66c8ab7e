"""Two-sided SVD Rayleigh-Ritz (svd_rr.m) and subspace iteration for the SVD
(subspace_iter_svd.m) on the GPU in FP64 against the oracle on the same inputs:
singular values within 1e-12 relative, Ritz vectors within 1e-9 (same signs:
the same Cholesky-reduced Jacobi eigensolver on both sides)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _lowrank(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    W, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.exp(-np.arange(n) / decay)
    return (U * s) @ W.T + 1e-9 * rng.standard_normal((m, n)), rng


def _colmajor(a):
    return torch.as_tensor(np.asfortranarray(a).T.copy(), device="cuda:0").t()


@pytest.mark.parametrize("ku,kv,generalized", [(10, 10, True), (12, 7, True), (16, 16, False)])
def test_svd_rr_matches_oracle(ctx, ku, kv, generalized):
    from mixedprecisionofrr_amd import api
    A, rng = _lowrank(1000, 300, 20.0, ku + kv)
    U0 = rng.standard_normal((1000, ku))
    V0 = rng.standard_normal((300, kv))
    if not generalized:
        U0, _ = np.linalg.qr(U0)
        V0, _ = np.linalg.qr(V0)
    print(f"svd_rr ku={ku} kv={kv} generalized={generalized}", flush=True)
    sig, U, V = api.svd_rr(ctx, _colmajor(A), _colmajor(U0), _colmajor(V0), generalized)
    so, Uo, Vo = O.svd_rr(A, U0, V0, generalized)
    assert np.max(np.abs(sig - so) / so[0]) < 1e-12
    assert np.max(np.abs(U.cpu().numpy() - Uo)) < 1e-9
    assert np.max(np.abs(V.cpu().numpy() - Vo)) < 1e-9


def test_subspace_iter_svd_matches_oracle_and_lapack(ctx):
    from mixedprecisionofrr_amd import api
    A, rng = _lowrank(2000, 400, 12.0, 5)
    V0 = rng.standard_normal((400, 16))
    print("subspace_iter_svd 2000 x 400, k 16", flush=True)
    sig, U, V, kept = api.subspace_iter_svd(ctx, _colmajor(A), _colmajor(V0), 25, 1)
    so, Uo, Vo, kc = O.subspace_iter_svd(A, V0, 25, 1)
    assert kept == kc == 16
    assert np.max(np.abs(sig - so) / so[0]) < 1e-12
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.max(np.abs(sig[:10] - ref[:10]) / ref[:10]) < 1e-10
    Uh, Vh = U.cpu().numpy(), V.cpu().numpy()
    r = np.linalg.norm(A @ Vh[:, :10] - Uh[:, :10] * sig[:10], axis=0) / sig[0]
    assert np.max(r) < 1e-9
