"""The CPU FP64 oracle pinned against the reference's known answers, published
vectors, closed forms and library routines (scipy) -- no GPU."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import oracle as O


def test_philox_published_vectors(kats):
    for c in kats["philox4x32_10"]["cases"]:
        assert list(O.philox(c["ctr"], c["key"])) == c["out"]


def test_randn_randu_laws():
    from scipy import stats
    z = O.randn(7, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert stats.kstest(z, "norm").pvalue > 1e-3
    u = O.randu(7, 100000, stream=3)
    assert u.min() >= 0 and u.max() < 1
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    # counter-based: offsets address the same stream
    assert np.array_equal(O.randn(7, 10, offset=5), z[5:15])


@pytest.mark.parametrize("dim,N", [(2, 64), (3, 16), (2, 7), (3, 5)])
def test_stencil_closed_form(dim, N, kats):
    A = O.stencil(dim, N)
    assert A.nnz == (5 * N * N - 4 * N if dim == 2 else 7 * N ** 3 - 6 * N ** 2)
    S = A.to_scipy()
    assert abs(S - S.T).max() == 0
    # same operator as the Kronecker-sum build
    e = np.ones(N)
    T = sp.diags([-e[:-1], 2 * e, -e[:-1]], [-1, 0, 1])
    I = sp.identity(N)
    K = sp.kron(I, T) + sp.kron(T, I) if dim == 2 else \
        sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)
    assert abs(S - K).max() == 0
    # sine modes are eigenvectors with the closed-form eigenvalue
    p = (3, 2, 1)[:dim]
    x = [np.sin(pi * np.pi * np.arange(1, N + 1) / (N + 1)) for pi in p]
    s = x[0][None, :] * x[1][:, None] if dim == 2 else \
        x[0][None, None, :] * x[1][None, :, None] * x[2][:, None, None]
    s = s.ravel()
    lam = sum(4 * np.sin(pi * np.pi / (2 * (N + 1))) ** 2 for pi in p)
    assert np.abs(O.spmm(A, s)[:, 0] - lam * s).max() < 1e-12
    if (dim, N) == (2, 64):
        assert np.allclose(O.laplacian_eigs(2, 64)[:24], kats["laplacian"]["lap2d_N64"]["top"], rtol=0, atol=1e-15)


def test_stencil_variable_coefficients():
    A = O.stencil(3, 6, seed=11)
    S = A.to_scipy()
    assert abs(S - S.T).max() == 0
    off = -S.data[S.data < 0]
    assert off.min() >= 0.1 - 1e-12 and off.max() <= 10 + 1e-12
    # diagonal = sum of incident face coefficients, including boundary faces
    # (interior rows sum to zero, boundary rows to their Dirichlet face coefficients)
    rowsum = np.asarray(S.sum(axis=1)).ravel()
    x = np.arange(A.n) % 6
    y = (np.arange(A.n) // 6) % 6
    z = np.arange(A.n) // 36
    interior = (x > 0) & (x < 5) & (y > 0) & (y < 5) & (z > 0) & (z < 5)
    assert np.abs(rowsum[interior]).max() < 1e-12
    assert rowsum[~interior].min() >= 0.1 - 1e-12
    assert np.all(np.diff(A.row_ptr) <= 7)
    assert np.linalg.eigvalsh(S.toarray()).min() > 0


def test_spmm_brute_force_and_exact_integers():
    A = O.stencil(3, 7, seed=3)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((A.n, 5))
    Y = O.spmm(A, X)
    assert np.abs(Y - A.to_scipy().toarray() @ X).max() < 1e-12
    L = O.stencil(2, 9)
    Xi = rng.integers(-50, 50, size=(L.n, 3)).astype(float)
    assert np.array_equal(O.spmm(L, Xi), L.to_scipy() @ Xi)


def test_iamax_reference_kats(kats):
    n = kats["iamax"]["n"]
    for c in kats["iamax"]["cases"]:
        x = (np.arange(n) - 65536) / 131072.0
        x[c["idx0"]] = c["val"]
        i, v = O.iamax(x)
        assert i == c["idx0"] + 1 and v == c["val"]
    x = np.where(np.arange(n) % 2 == 0, np.arange(n), -np.arange(n)).astype(float)
    assert O.iamax(x)[0] == n
    # ties keep the first index
    assert O.iamax(np.array([1.0, -3.0, 3.0, -3.0]))[0:2] == (2, -3.0)


def _glibc_planted(rows, cols, seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(seed)
    return [libc.rand() % rows for _ in range(cols)]


def test_matrix_iamax_reference_kat(kats):
    k = kats["matrix_iamax"]
    rows, cols = k["rows"], 200  # first 200 columns of the reference case
    planted = _glibc_planted(rows, k["cols"], k["srand"])[:cols]
    A = np.tile(((np.arange(rows) - 5000) / 10000.0)[:, None], (1, cols))
    for c in range(cols):
        A[planted[c], c] = -c - 1 if c % 2 == 0 else c + 1
    assert list(O.matrix_iamax(A)) == planted


def test_scalecols():
    A = np.array([[1.0, 0.0, 2.0], [-4.0, 0.0, -2.0], [2.0, 0.0, 1.0]])
    S = O.scalecols(A)
    assert np.array_equal(S[:, 0], A[:, 0] / -4.0)
    assert np.array_equal(S[:, 1], A[:, 1])          # zero column untouched
    assert np.array_equal(S[:, 2], A[:, 2] / 2.0)    # tie: first index (+2) wins


def test_gemm_reference_kat(kats):
    g = kats["gemm_3x2_2x4"]
    A, B, Cx = np.array(g["A"]), np.array(g["B"]), np.array(g["C"])
    assert np.array_equal(O.gemm_nn(A, B), Cx)
    assert np.array_equal(O.gemm_tn(A.T.copy(), B), Cx)
    rng = np.random.default_rng(1)
    V, W = rng.standard_normal((300, 7)), rng.standard_normal((300, 5))
    assert np.allclose(O.gemm_tn(V, W), V.T @ W, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("m,n", [(2000, 200), (500, 40)])
def test_hessenberg_invariants_and_skip(m, n, kats):
    rng = np.random.default_rng(815)
    A = rng.uniform(-1, 1, (m, n))
    Q, R, sel, perm = O.hessenberg(A, 2.0 ** -52)
    assert sel.all()
    assert np.abs(A - Q @ R).max() < kats["hessenberg"]["qr_tol"]["f64"]
    assert np.abs(Q).max() <= 1.0
    assert np.allclose(np.tril(R, -1), 0)
    for i in range(n):
        assert Q[perm[i], i] == 1.0
        assert np.all(Q[perm[:i], i] == 0.0)
        assert R[i, i] == pytest.approx(A[perm[i], i] - Q[perm[i], :i] @ R[:i, i], rel=1e-10)
    # planted linear dependence: columns 0-2 copied onto 3-5 are dropped
    A2 = A.copy()
    A2[:, 3:6] = A2[:, 0:3]
    Q2, R2, sel2, perm2 = O.hessenberg(A2, kats["hessenberg"]["dup_select_tol"]["f64"])
    skip = 1 - sel2
    assert list(np.nonzero(skip)[0]) == [3, 4, 5]


def _rtilde(R, sel):
    Rt = R.copy()
    for j in np.nonzero(sel == 0)[0]:
        Rt[j, j] = 1.0  # a dropped column keeps its unscaled residual
    return Rt


def test_hessenberg_variants_with_dropped_columns(kats):
    """Variant 0 (MATLAB / C++ v1) leaves dropped columns out of later
    eliminations; variant 1 (C++ v3, hessenberg.cu:768-779) eliminates with them,
    using their argmax row and unscaled residual.  Both satisfy A = Q R~ (R~ = R
    with a unit diagonal at dropped columns); they agree bit for bit without
    drops and differ, with nonzero R(j, i) for a dropped j, with them."""
    rng = np.random.default_rng(31)
    A = rng.uniform(-1, 1, (400, 30))
    Q0, R0, s0, p0 = O.hessenberg(A, 2.0 ** -52, variant=0)
    Q1, R1, s1, p1 = O.hessenberg(A, 2.0 ** -52, variant=1)
    assert s0.all() and np.array_equal(Q0, Q1) and np.array_equal(R0, R1)
    A2 = A.copy()
    A2[:, 7] = A2[:, 2] + 1e-9 * rng.uniform(-1, 1, 400)   # nearly dependent: dropped at 1e-6
    Q0, R0, s0, p0 = O.hessenberg(A2, 1e-6, variant=0)
    Q1, R1, s1, p1 = O.hessenberg(A2, 1e-6, variant=1)
    assert list(np.nonzero(s0 == 0)[0]) == [7] and list(np.nonzero(s1 == 0)[0]) == [7]
    for Q, R, sl in ((Q0, R0, s0), (Q1, R1, s1)):
        assert np.abs(A2 - Q @ _rtilde(R, sl)).max() < 1e-12
    assert np.all(R0[7, 8:] == 0.0) and np.any(R1[7, 8:] != 0.0)
    assert not np.array_equal(Q0[:, 8:], Q1[:, 8:])
    np.testing.assert_array_equal(Q0[:, :8], Q1[:, :8])


def test_syev_sygv_reference_kats(kats):
    n = 10
    A = np.diag(2 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    w, V = O.syev(A)
    assert np.abs(w - kats["syev_tridiag10"]["eigenvalues"]).max() < kats["syev_tridiag10"]["tol_double"]
    assert np.abs(A @ V - V * w).max() < 1e-13
    info, w2, V2 = O.sygv(A, np.diag(np.arange(1.0, 11)))
    assert info == 0
    assert np.abs(w2 - kats["sygv_tridiag10_diag"]["eigenvalues"]).max() < 1e-10
    assert np.allclose(V2.T @ np.diag(np.arange(1.0, 11)) @ V2, np.eye(n), atol=1e-12)


def test_syev_vs_scipy_random_and_degenerate():
    rng = np.random.default_rng(2)
    M = rng.standard_normal((60, 60))
    A = M + M.T
    w, V = O.syev(A)
    assert np.allclose(w, sla.eigh(A, eigvals_only=True), rtol=0, atol=1e-12)
    # exact multiplicities (the Laplacian case): eigenvalues still exact, vectors orthonormal
    Q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    d = np.repeat([1.0, 2.0, 2.0, 5.0, 5.0, 5.0], 5)
    B = (Q * d) @ Q.T
    w, V = O.syev(B)
    assert np.abs(w - np.sort(d)).max() < 1e-13
    assert np.abs(V.T @ V - np.eye(30)).max() < 1e-13
    info, _, _ = O.sygv(A, -np.eye(60))
    assert info == 1


def test_rr_generalized_vs_scipy():
    rng = np.random.default_rng(3)
    V = rng.standard_normal((400, 12))
    M = rng.standard_normal((400, 400))
    S = M + M.T
    G, H = V.T @ V, V.T @ S @ V
    info, th, Y = O.rr_generalized(G, H, 0.0)
    assert info == 0
    assert np.allclose(th, sla.eigh(H, G, eigvals_only=True)[::-1], rtol=1e-11, atol=1e-11)


def test_rr_truncated_exact_ritz_and_rank_drop():
    N = 20
    A = O.stencil(2, N)
    S = A.to_scipy().toarray()
    lam, U = np.linalg.eigh(S)
    # basis spanning 6 exact eigenvectors, mixed and rescaled -> exact Ritz values
    rng = np.random.default_rng(4)
    idx = [399, 398, 396, 390, 380, 300]
    V = U[:, idx] @ rng.standard_normal((6, 6)) * 1e3
    G, H = V.T @ V, V.T @ S @ V
    r, th, Y = O.rr_truncated(G, H, 1e-14, 6)
    assert r == 6
    assert np.allclose(th, np.sort(lam[idx])[::-1], rtol=1e-12, atol=0)
    X = V @ Y
    assert np.abs(S @ X - X * th).max() < 1e-10
    # planted rank drop: duplicated columns are truncated away
    V2 = np.concatenate([V, V[:, :2]], axis=1)
    r2, th2, _ = O.rr_truncated(V2.T @ V2, V2.T @ S @ V2, 1e-12, 6)
    assert r2 == 6
    assert np.allclose(th2, th, rtol=1e-11)


def _split_basis(S, kx, kd, seed, collinear=True):
    """A solver-like basis [X | D] for a symmetric S: X orthonormal (previous
    Ritz block), D correlated with X and, optionally, within itself."""
    rng = np.random.default_rng(seed)
    n = S.shape[0]
    X = np.linalg.qr(rng.standard_normal((n, kx)))[0]
    D = rng.standard_normal((n, kd)) + X @ rng.standard_normal((kx, kd)) * 0.7
    if collinear:  # nearly parallel columns, as the two filter snapshots give
        D[:, kd // 2:] = D[:, :kd - kd // 2] + 1e-3 * rng.standard_normal((n, kd - kd // 2))
    return np.concatenate([X, D], axis=1)


@pytest.mark.parametrize("kx,kd", [(6, 10), (16, 32), (7, 9)])
def test_rr_split_matches_textbook_pencil(kx, kd):
    """orc_rr_split against scipy's generalised eigensolver on a well-conditioned
    pencil: the top Ritz values agree, Y^T G Y = I and Y^T H Y = diag(theta)."""
    A = O.stencil(2, 14)
    S = A.to_scipy().toarray()
    V = _split_basis(S, kx, kd, kx + kd, collinear=False)
    G, H = V.T @ V, V.T @ S @ V
    nk = kx
    r, th, Y = O.rr_split(G, H, kx, 1e-12, nk)
    assert r == kx + kd
    ref = sla.eigh(H, G, eigvals_only=True)[::-1][:nk]
    assert np.abs(th - ref).max() < 1e-12 * abs(ref[0])
    assert np.abs(Y.T @ G @ Y - np.eye(nk)).max() < 1e-12
    assert np.abs(Y.T @ H @ Y - np.diag(th)).max() < 1e-12 * abs(th[0])


def test_rr_split_exact_ritz_rank_drop_and_indefinite_block():
    """Exact Ritz values from an invariant subspace; D columns inside span(X)
    are truncated away; a non-positive-definite X block returns -1."""
    N = 20
    A = O.stencil(2, N)
    S = A.to_scipy().toarray()
    lam, U = np.linalg.eigh(S)
    rng = np.random.default_rng(5)
    idx = [399, 398, 396, 390, 380, 300, 250, 200]
    X = U[:, idx[:4]] @ np.linalg.qr(rng.standard_normal((4, 4)))[0]
    D = U[:, idx] @ rng.standard_normal((8, 6))
    V = np.concatenate([X, D, X[:, :2] * 3.0], axis=1)  # two columns inside span(X)
    G, H = V.T @ V, V.T @ S @ V
    r, th, Y = O.rr_split(G, H, 4, 1e-12, 8)
    assert r == 8  # X and the 4 directions of D outside span(X)
    assert np.allclose(th, np.sort(lam[idx])[::-1], rtol=1e-11, atol=0)
    Xr = V @ Y
    assert np.abs(S @ Xr - Xr * th).max() < 1e-9
    # the whole form agrees
    r2, th2, _ = O.rr_truncated(G, H, 1e-12, 8)
    assert np.abs(th2 - th).max() < 1e-11 * th[0]
    # indefinite X block
    G2 = G.copy()
    G2[1, 1] = -1.0
    r3, _, _ = O.rr_split(G2, H, 4, 1e-12, 4)
    assert r3 == -1


@pytest.mark.parametrize("seed", [0, 1])
def test_rr_split_solver_like_basis(seed):
    """Collinear correction columns (Gram condition ~6.5e6): the split form and
    the whole form both meet the Ritz values of an orthonormal basis of the same
    span (Householder QR) within the accuracy the Gram matrix carries,
    eps cond(G) ||S||."""
    A = O.stencil(3, 10)
    S = A.to_scipy().toarray()
    V = _split_basis(S, 16, 32, seed)
    G, H = V.T @ V, V.T @ S @ V
    r, th, Y = O.rr_split(G, H, 16, 1e-12, 16)
    r2, th2, Y2 = O.rr_truncated(G, H, 1e-12, 16)
    Q = np.linalg.qr(V)[0]
    ref = np.linalg.eigvalsh(Q.T @ S @ Q)[::-1][:16]
    bound = 2.2e-16 * np.linalg.cond(G) * ref[0]
    assert r == r2 == 48
    assert np.abs(th - ref).max() < bound
    assert np.abs(th2 - ref).max() < bound
    assert np.abs(Y.T @ G @ Y - np.eye(16)).max() < 1e-9


def test_arnoldi_exact_termination():
    # v0 = sum of 3 sine modes: the Krylov space is invariant after 3 columns
    N = 16
    A = O.stencil(2, N)
    S = A.to_scipy().toarray()
    lam, U = np.linalg.eigh(S)
    idx = [255, 250, 240]
    v0 = U[:, idx] @ np.array([1.0, 0.5, 0.25])
    th, X, built = O.arnoldi_hess_eig(A, 3, 10, 11, v0, select_tol=1e-10)
    assert built == 3
    assert np.allclose(th, np.sort(lam[idx])[::-1], rtol=1e-12)


def test_arnoldi_laplacian64_converges_top():
    A = O.stencil(2, 64)
    v0 = O.randn(0, A.n)
    th, X, built = O.arnoldi_hess_eig(A, 20, 60, 1 + 60 * 4, v0)
    ex = O.laplacian_eigs(2, 64)
    assert built == 61
    assert abs(th[0] - ex[0]) / ex[0] < 1e-10
    res = O.residuals(A, X, th, ex[0])
    assert res[0] < 1e-6  # the eigenvalue converges quadratically faster than the vector


def test_subspace_iteration_laplacian():
    A = O.stencil(2, 12)
    V0 = O.randn(0, A.n * 8).reshape(A.n, 8, order="F")
    th, V = O.subspace_iter_eig(A, V0, maxiter=300, step_size=2)
    ex = O.laplacian_eigs(2, 12)
    assert abs(th[0] - ex[0]) / ex[0] < 1e-12
    assert np.all(np.isfinite(th))


@pytest.mark.parametrize("nd", [1, 2])
def test_filtered_ofrr_closed_form(nd):
    N, k, nwant = 14, 32, 20
    A = O.stencil(3, N)
    X0 = O.randn(0, A.n * k).reshape(A.n, k, order="F")
    it, th, X, res = O.filtered_ofrr(A, X0, nwant, nd=nd, degree=14, maxiter=60, tol=1e-9)
    assert it >= 0
    ex = O.laplacian_eigs(3, N)[:nwant]
    assert np.abs(th[:nwant] - ex).max() / ex[0] < 1e-12
    assert res[:nwant].max() <= 1e-9
    res2 = O.residuals(A, X, th, th[0])
    assert np.allclose(res2[:nwant], res[:nwant], rtol=1e-6, atol=1e-14)


def test_filtered_ofrr_soft_locking():
    """Soft locking: with no pair converged the solve is the unlocked one,
    bitwise; once the leading pairs converge and later sweeps filter only the
    trailing columns, the spectrum is still the closed form and the residuals
    meet the tolerance, in the same number of iterations."""
    N, k, nwant = 14, 32, 20
    A = O.stencil(3, N)
    X0 = O.randn(0, A.n * k).reshape(A.n, k, order="F")
    a = O.filtered_ofrr(A, X0, nwant, nd=2, degree=10, maxiter=3, tol=1e-30, lock=True)
    b = O.filtered_ofrr(A, X0, nwant, nd=2, degree=10, maxiter=3, tol=1e-30, lock=False)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    ex = O.laplacian_eigs(3, N)[:nwant]
    its = {}
    for lock in (True, False):
        it, th, X, res = O.filtered_ofrr(A, X0, nwant, nd=2, degree=10, maxiter=60, tol=1e-9,
                                         lock=lock)
        assert it >= 0
        assert np.abs(th[:nwant] - ex).max() / ex[0] < 1e-12
        assert res[:nwant].max() <= 1e-9
        assert np.allclose(O.residuals(A, X, th, th[0])[:nwant], res[:nwant], rtol=1e-6, atol=1e-14)
        its[lock] = it
    assert its[True] == its[False]


def test_residuals_hand_case():
    A = O.stencil(2, 10)
    lam, U = np.linalg.eigh(A.to_scipy().toarray())
    X = U[:, -3:]
    r = O.residuals(A, X, lam[-3:], lam[-1])
    assert r.max() < 1e-14
    Xp = X.copy()
    Xp[:, 0] += 1e-3 * U[:, 0]
    r = O.residuals(A, Xp, lam[-3:], lam[-1])
    expect = 1e-3 * abs(lam[0] - lam[-3]) / (lam[-1] * np.sqrt(1 + 1e-6))
    assert r[0] == pytest.approx(expect, rel=1e-6)
