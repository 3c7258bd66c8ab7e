"""The oracle's restatement of svd_rr.m and subspace_iter_svd.m, pinned by what
the mathematics fixes (no GPU):
* exact subspaces: for U0 = U_k C1, V0 = W_k C2 of A = U diag(s) W^T the
  generalised augmented pencil returns s_1..s_k, unit U and V with A V = U S;
* the plain branch on orthonormal U, V returns the singular values of U^T A V
  (LAPACK through numpy), and sqrt(2)-scaled vectors of the augmented matrix;
* subspace iteration converges to LAPACK's top singular values;
* a planted mistake (a transposed B block) is caught."""
import numpy as np

from oracle import oracle as O


def _lowrank(m, n, decay, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    W, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.exp(-np.arange(n) / decay)
    return (U * s) @ W.T, U, s, W, rng


def test_svd_rr_exact_subspaces():
    A, U, s, W, rng = _lowrank(300, 120, 10.0, 1)
    for ku, kv in [(8, 8), (10, 14), (15, 9)]:
        U0 = U[:, :ku] @ rng.standard_normal((ku, ku))
        V0 = W[:, :kv] @ rng.standard_normal((kv, kv))
        sig, Uo, Vo = O.svd_rr(A, U0, V0, generalized=True)
        kk = min(ku, kv)
        assert sig.shape == (kk,)
        if ku == kv:  # the subspaces pair up exactly
            assert np.max(np.abs(sig - s[:kk]) / s[:kk]) < 1e-12
            assert np.max(np.abs(A @ Vo - Uo * sig)) < 1e-12
        assert np.max(np.abs(np.linalg.norm(Uo, axis=0) - 1)) < 1e-10
        assert np.max(np.abs(np.linalg.norm(Vo, axis=0) - 1)) < 1e-10
        assert np.all(np.diff(sig) <= 0)


def test_svd_rr_plain_branch_matches_lapack():
    A, U, s, W, rng = _lowrank(200, 80, 30.0, 2)
    P, _ = np.linalg.qr(rng.standard_normal((200, 12)))
    Q, _ = np.linalg.qr(rng.standard_normal((80, 12)))
    sig, Uo, Vo = O.svd_rr(A, P, Q, generalized=False)
    ref = np.linalg.svd(P.T @ A @ Q, compute_uv=False)
    assert np.max(np.abs(sig - ref) / ref[0]) < 1e-12
    # U = P Z, V = Q W with (Z; W) unit eigenvectors of [0 B; B^T 0] scaled by sqrt 2
    assert np.max(np.abs(np.linalg.norm(Uo, axis=0) - 1)) < 1e-10
    assert np.max(np.abs((Uo.T @ A @ Vo) - np.diag(sig))) < 1e-10


def test_subspace_iter_svd_converges_to_lapack():
    A, U, s, W, rng = _lowrank(400, 150, 8.0, 3)
    V0 = rng.standard_normal((150, 12))
    sig, Uo, Vo, kc = O.subspace_iter_svd(A, V0, 40, 1, generalized=True)
    assert kc == 12
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.max(np.abs(sig[:8] - ref[:8]) / ref[:8]) < 1e-12
    r = np.linalg.norm(A @ Vo[:, :8] - Uo[:, :8] * sig[:8], axis=0) / sig[0]
    assert np.max(r) < 1e-10


def test_planted_transposed_block_is_caught():
    A, U, s, W, rng = _lowrank(150, 60, 6.0, 4)
    U0 = U[:, :6] @ rng.standard_normal((6, 6))
    V0 = W[:, :6] @ rng.standard_normal((6, 6))
    good, _, _ = O.svd_rr(A, U0, V0, generalized=True)
    # the same pencil built with A^T in place of A (a transposed operand) on a
    # non-square problem gives different numbers: the check above would fail
    bad, _, _ = O.svd_rr(A[:60, :].T.copy(), V0, V0, generalized=True)
    assert np.max(np.abs(good - s[:6]) / s[:6]) < 1e-12
    assert np.max(np.abs(bad - s[:6]) / s[:6]) > 1e-3
