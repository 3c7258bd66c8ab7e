"""GPU kernels against the CPU FP64 oracle (bitwise for FP64 integer-exact work,
rounding-model bounds for low precision).  Marked gpu."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda:0")


def _csr(ctx, Ah):
    from mixedprecisionofrr_amd import api
    return api.CSRMatrix.from_host(ctx, Ah.row_ptr, Ah.col_idx, Ah.vals)


def test_randn_matches_oracle(ctx):
    from mixedprecisionofrr_amd import api
    x = torch.empty(100003, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, x, seed=1234, stream=2, offset=17)
    ref = O.randn(1234, 100003, stream=2, offset=17)
    assert np.max(np.abs(x.cpu().numpy() - ref)) <= 1e-14 * np.max(np.abs(ref))
    u = torch.empty(5001, dtype=torch.float64, device="cuda:0")
    api.fill_randu(ctx, u, seed=9, stream=1)
    assert np.array_equal(u.cpu().numpy(), O.randu(9, 5001, stream=1))


@pytest.mark.parametrize("dim,N,seed", [(2, 64, 0), (3, 17, 0), (3, 9, 5), (2, 31, 77)])
def test_stencil_generator_matches_oracle(ctx, dim, N, seed):
    from mixedprecisionofrr_amd import api
    A = api.CSRMatrix.stencil(ctx, dim, N, seed)
    Ah = O.stencil(dim, N, seed)
    assert np.array_equal(A.row_ptr.cpu().numpy(), Ah.row_ptr)
    assert np.array_equal(A.col_idx.cpu().numpy(), Ah.col_idx)
    v = A.vals.cpu().numpy()
    if seed == 0:
        assert np.array_equal(v, Ah.vals)
    else:
        assert np.max(np.abs(v - Ah.vals) / np.abs(Ah.vals)) < 4e-16 * 8


@pytest.mark.parametrize("layout", ["row", "col"])
@pytest.mark.parametrize("k", [1, 8, 64, 61])
def test_spmm_fp64_bitwise(ctx, layout, k):
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(3, 13, seed=2)
    A = _csr(ctx, Ah)
    X = np.random.default_rng(k).standard_normal((Ah.n, k))
    Xd = _dev(X) if layout == "row" else _dev(X.T).t()
    print(f"spmm fp64 layout={layout} k={k}", flush=True)
    Y = api.spmm(ctx, A, Xd).cpu().numpy()
    ref = O.spmm(Ah, X)
    bad = np.argwhere(Y != ref)
    assert bad.shape[0] == 0, (f"{bad.shape[0]} mismatches, first {bad[:4].tolist()}, "
                               f"max |diff| {np.max(np.abs(Y - ref)):.3e}, "
                               f"max |ref| {np.max(np.abs(ref)):.3e}")


@pytest.mark.parametrize("xdt,ydt", [(torch.float32, torch.float32), (torch.float16, torch.float32),
                                     (torch.bfloat16, torch.float32), (torch.float16, torch.float16)])
def test_spmm_low_precision_rounding_bound(ctx, xdt, ydt):
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(3, 16, seed=4)
    A = _csr(ctx, Ah)
    X = np.random.default_rng(5).uniform(-1, 1, (Ah.n, 64))
    Xl = _dev(X, xdt)
    Xq = Xl.double().cpu().numpy()  # exactly the stored operand
    Y = api.spmm(ctx, A, Xl, out_dtype=ydt).double().cpu().numpy()
    ref = O.spmm(Ah, Xq)
    absA = Ah.to_scipy().copy()
    absA.data = np.abs(absA.data)
    bound = absA @ np.abs(Xq)
    u_acc = 2.0 ** -24 * 8  # FP32 accumulation over <= 7 terms
    u_out = {torch.float32: 2.0 ** -24, torch.float16: 2.0 ** -11}[ydt]
    assert np.all(np.abs(Y - ref) <= u_acc * bound + u_out * np.abs(ref) + 1e-30)


def test_spmm_fp16_values_on_laplacian_exact(ctx):
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(3, 16)
    A = _csr(ctx, Ah).with_low_precision(torch.float16)
    Xi = np.random.default_rng(6).integers(-64, 64, (Ah.n, 64)).astype(np.float64)
    Y = api.spmm(ctx, A, _dev(Xi, torch.float16), out_dtype=torch.float32, vals=A.vals_lo)
    assert np.array_equal(Y.double().cpu().numpy(), O.spmm(Ah, Xi))


def test_iamax_reference_kats(ctx, kats):
    from mixedprecisionofrr_amd import api
    n = kats["iamax"]["n"]
    dts = {"f64": torch.float64, "f32": torch.float32, "f16": torch.float16}
    for c in kats["iamax"]["cases"]:
        x = (np.arange(n) - 65536) / 131072.0
        x[c["idx0"]] = c["val"]
        i, v = api.iamax(ctx, _dev(x, dts[c["dtype"]]))
        assert i == c["idx0"] + 1 and v == c["val"], c
    x = np.where(np.arange(n) % 2 == 0, np.arange(n), -np.arange(n)).astype(np.float32)
    assert api.iamax(ctx, _dev(x, torch.float32))[0] == n
    assert api.iamax(ctx, _dev(np.array([1.0, -3.0, 3.0, -3.0])))[0:2] == (2, -3.0)
    # ties across the chunk boundary of the two-stage reduction keep the lowest index
    y = np.zeros(50000)
    y[[9000, 20000, 40000]] = [-5.0, 5.0, 5.0]
    assert api.iamax(ctx, _dev(y)) == (9001, -5.0)


def test_matrix_iamax_reference_kat(ctx, kats):
    from mixedprecisionofrr_amd import api
    k = kats["matrix_iamax"]
    rows, cols = k["rows"], k["cols"]
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(k["srand"])
    planted = [libc.rand() % rows for _ in range(cols)]
    A = np.tile(((np.arange(rows) - 5000) / 10000.0)[:, None], (1, cols))
    for c in range(cols):
        A[planted[c], c] = -c - 1 if c % 2 == 0 else c + 1
    for dt in (torch.float64, torch.float32):
        Ad = _dev(A.T, dt).t()  # column-major like the reference Matrix<T>
        assert api.matrix_iamax(ctx, Ad).cpu().tolist() == planted


def test_scalecols_matches_oracle(ctx):
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(8).standard_normal((3000, 12))
    A[:, 5] = 0.0
    A[[10, 20], 7] = [4.0, -4.0]
    Ad = _dev(A.T).t().clone()
    piv = api.scalecols(ctx, Ad)
    assert np.array_equal(Ad.cpu().numpy(), O.scalecols(A))
    assert piv.cpu().numpy()[7] == 4.0 and piv.cpu().numpy()[5] == 0.0


def test_check_special(ctx):
    from mixedprecisionofrr_amd import api
    x = torch.zeros(1000, 3, dtype=torch.float64, device="cuda:0")
    assert api.check_special(ctx, x) == 0
    x[17, 1] = float("inf")
    assert api.check_special(ctx, x) == 1006
    x[3, 2] = float("nan")
    assert api.check_special(ctx, x) == 1005


@pytest.mark.parametrize("dt,m,n", [(torch.float64, 2000, 200), (torch.float32, 2000, 200),
                                    (torch.float16, 2000, 100)])
def test_hessenberg_reference_tolerances_and_skip(ctx, kats, dt, m, n):
    from mixedprecisionofrr_amd import api
    key = {torch.float64: "f64", torch.float32: "f32", torch.float16: "f16"}[dt]
    A = np.random.default_rng(815).uniform(-1, 1, (m, n))
    Ad = _dev(A.T, dt).t()
    print(f"hessenberg {key} {m}x{n}", flush=True)
    Q, R, perm, skip = api.hessenberg(ctx, Ad)
    Qh = Q.double().cpu().numpy()
    Rh = R.cpu().numpy()
    Aq = Ad.double().cpu().numpy()
    assert api.check_special(ctx, Q) == 0 and api.check_special(ctx, R) == 0
    assert np.abs(Aq - Qh @ Rh).max() < kats["hessenberg"]["qr_tol"][key]
    assert skip.cpu().numpy().sum() == 0
    p = perm.cpu().numpy()
    assert np.all(np.abs(Qh) <= 1.0 + 1e-3)
    for i in range(n):
        assert Qh[p[i], i] == 1.0
    if dt == torch.float64:
        Qo, Ro, so, po = O.hessenberg(A, 2.0 ** -52)
        assert np.array_equal(p, po)
        assert np.array_equal(Qh, Qo) and np.array_equal(Rh, Ro)
    # planted linear dependence -> exact skip pattern
    A2 = A.copy()
    A2[:, 3:6] = A2[:, 0:3]
    Q2, R2, perm2, skip2 = api.hessenberg(ctx, _dev(A2.T, dt).t(), kats["hessenberg"]["dup_select_tol"][key])
    assert list(np.nonzero(skip2.cpu().numpy())[0]) == [3, 4, 5]
    assert list(perm2.cpu().numpy()[3:6]) == [-1, -1, -1]


@pytest.mark.parametrize("vdt,wdt,tol", [(torch.float64, torch.float64, 1e-14),
                                         (torch.float16, torch.float16, 1e-15),
                                         (torch.float64, torch.float16, 1e-14),
                                         (torch.float16, torch.float32, 1e-14)])
def test_gram_fp64_accumulation(ctx, vdt, wdt, tol):
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(9)
    V = _dev(rng.uniform(-1, 1, (100000, 64)), vdt)
    W = _dev(rng.uniform(-1, 1, (100000, 72)), wdt)
    G = api.gram(ctx, V, W).cpu().numpy()
    Vh, Wh = V.double().cpu().numpy(), W.double().cpu().numpy()
    ref = O.gemm_tn(Vh, Wh)
    scale = np.abs(Vh).T @ np.abs(Wh)
    assert np.max(np.abs(G - ref) / scale) < tol
    # bitwise reproducible (fixed-order split over rows)
    assert np.array_equal(G, api.gram(ctx, V, W).cpu().numpy())
    # the device accumulates each chunk's rows in order with FMAs (MFMA f64) and
    # sums the chunks in the same fixed order as the oracle: equal bits
    assert np.array_equal(G, ref)


def test_gram_reference_kat_and_layouts(ctx, kats):
    from mixedprecisionofrr_amd import api
    g = kats["gemm_3x2_2x4"]
    At = np.array(g["A"]).T  # A = At^T, so A B = At^T B
    Vd = _dev(At)            # 2 x 3 row-major, V^T W with V = At (2x3)?
    # gemm(transA=true): G = V^T W with V (2 x 3) = A^T, W = B (2 x 4)
    G = api.gram(ctx, _dev(np.array(g["A"]).T.copy()), _dev(np.array(g["B"])))
    assert np.array_equal(G.cpu().numpy(), np.array(g["C"]))
    Gc = api.gram(ctx, _dev(np.array(g["A"])).t().contiguous().t(), _dev(np.array(g["B"])))
    del Vd, Gc


def test_basis_update_mixed_blocks(ctx):
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(10)
    n, k = 50000, 64
    X = _dev(rng.standard_normal((n, k)))
    D = _dev(rng.uniform(-1, 1, (n, k)), torch.float16)
    C1 = _dev(rng.standard_normal((128, k)).T).t()  # column-major 128 x 64
    Y = api.basis_update(ctx, [(X, C1[:64]), (D, C1[64:])], k).cpu().numpy()
    ref = X.cpu().numpy() @ C1[:64].cpu().numpy() + D.double().cpu().numpy() @ C1[64:].cpu().numpy()
    assert np.max(np.abs(Y - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,k,kbs,dts,out", [
    (1000, 16, (16, 32), (torch.float64, torch.bfloat16), torch.float64),  # ragged last tile
    (4099, 128, (128, 128), (torch.float64, torch.float16), torch.float64),
    (777, 48, (48, 48, 48), (torch.float64, torch.float8_e4m3fn, torch.float32), torch.float32),
    (513, 64, (64, 40), (torch.float64, torch.float16), torch.float64),  # width 40: LDS kernel
    # more 16-row tiles than one pass of the persistent workgroups covers (256 CUs x 16 waves)
    (150001, 64, (64, 64, 64), (torch.float64, torch.float16, torch.float16), torch.float64),
])
def test_basis_update_shapes(ctx, n, k, kbs, dts, out):
    """Back-transform Y = sum_b V_b C_b against FP64.  Row-major blocks of widths
    that are multiples of 16 take the FP64 MFMA path (coefficients staged in LDS by
    persistent workgroups for k <= 64, read from L2 above); other layouts take the
    FMA kernel with LDS tiles."""
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(n + k)
    blocks, ref = [], 0.0
    for kb, dt in zip(kbs, dts):
        V = _dev(rng.uniform(-1, 1, (n, kb)), dt)
        Cb = _dev(rng.standard_normal((kb, k)).T).t()  # column-major kb x k
        blocks.append((V, Cb))
        ref = ref + V.double().cpu().numpy() @ Cb.cpu().numpy()
    Y = api.basis_update(ctx, blocks, k, out_dtype=out).double().cpu().numpy()
    tol = 1e-12 if out == torch.float64 else 1e-6
    assert np.max(np.abs(Y - ref)) < tol * np.max(np.abs(ref))
    # a column-major block takes the LDS kernel: same result within rounding
    Vc = blocks[0][0].t().contiguous().t()
    Y2 = api.basis_update(ctx, [(Vc, blocks[0][1])] + blocks[1:], k, out_dtype=out).double().cpu().numpy()
    assert np.max(np.abs(Y2 - ref)) < tol * np.max(np.abs(ref))


def test_syev_reference_kat_batched_and_random(ctx, kats):
    from mixedprecisionofrr_amd import api
    n = 10
    A = np.diag(2 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    w, V = api.syev_batched(ctx, _dev(A))
    assert np.abs(w.cpu().numpy()[0] - kats["syev_tridiag10"]["eigenvalues"]).max() < 1e-13
    rng = np.random.default_rng(11)
    for m in (7, 64, 96, 128, 192):
        M = rng.standard_normal((3, m, m))
        S = M + M.transpose(0, 2, 1)
        print(f"syev batched m={m}", flush=True)
        w, V = api.syev_batched(ctx, _dev(S))
        w, V = w.cpu().numpy(), V.cpu().numpy()
        for b in range(3):
            wo, _ = O.syev(S[b])
            assert np.abs(w[b] - wo).max() < 1e-12 * np.abs(wo).max()
            assert np.abs(S[b] @ V[b] - V[b] * w[b]).max() < 1e-11 * np.abs(wo).max()
            assert np.abs(V[b].T @ V[b] - np.eye(m)).max() < 1e-12


def test_syev_bitwise_vs_oracle(ctx):
    """The device eigensolver and its CPU restatement agree bit for bit
    (fixed reduction orders, explicit fused multiply-adds on both sides)."""
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(21)
    for m in (2, 3, 17, 64, 127, 128, 130, 191, 192, 193, 200, 256, 257, 300, 383, 384):
        M = rng.standard_normal((m, m))
        S = M + M.T
        w, V = api.syev_batched(ctx, _dev(S[None]))
        wo, Vo = O.syev(S)
        print(f"syev bitwise m={m}", flush=True)
        assert np.array_equal(w.cpu().numpy()[0], wo)
        assert np.array_equal(V.cpu().numpy()[0], Vo)


def test_syev_exact_multiplicities(ctx):
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(12)
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    d = np.repeat([1.0, 2.0, 5.0, 7.0], 10)
    B = (Q * d) @ Q.T
    w, V = api.syev_batched(ctx, _dev(B))
    # B is the rounded product Q diag(d) Q^T: its spectrum is d to within ~eps ||B|| sqrt(n)
    assert np.abs(w.cpu().numpy()[0] - np.sort(d)).max() < 5e-13
    assert np.abs(V.cpu().numpy()[0].T @ V.cpu().numpy()[0] - np.eye(40)).max() < 1e-13


def test_sygv_reference_kat(ctx, kats):
    from mixedprecisionofrr_amd import api
    n = 10
    A = np.diag(2 * np.ones(n)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    B = np.diag(np.arange(1.0, 11))
    w, V = api.sygv(ctx, _dev(A), _dev(B))
    assert np.abs(w.cpu().numpy() - kats["sygv_tridiag10_diag"]["eigenvalues"]).max() < 1e-12
    Vh = V.cpu().numpy()
    assert np.allclose(Vh.T @ B @ Vh, np.eye(n), atol=1e-12)
    from mixedprecisionofrr_amd.api import OfrrError
    with pytest.raises(OfrrError) as e:
        api.sygv(ctx, _dev(A), _dev(-np.eye(n)))
    assert e.value.status == 3001


@pytest.mark.parametrize("m,drop", [(64, 4), (128, 4), (192, 4), (192, 3), (127, 3)])
def test_rr_truncated_vs_oracle(ctx, m, drop):
    """Truncated RR; an odd kept rank runs the eigensolver on an odd size read
    from the device (zero-padded virtual index)."""
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(3, 12)
    S = Ah.to_scipy()
    rng = np.random.default_rng(m)
    V = rng.standard_normal((Ah.n, m))
    V[:, -drop:] = V[:, :drop] + 1e-13 * rng.standard_normal((Ah.n, drop))  # near rank drop
    G, H = V.T @ V, V.T @ (S @ V)
    r_o, th_o, Y_o = O.rr_truncated(G, H, 1e-10, 32)
    print(f"rr_truncated m={m} drop={drop}", flush=True)
    th, Y, r = api.rr_truncated(ctx, _dev(G.T).t(), _dev(H.T).t(), 1e-10, 32)
    assert r == r_o
    assert np.abs(th.cpu().numpy() - th_o).max() < 1e-12 * abs(th_o[0])
    X = V @ Y.cpu().numpy()
    res = np.linalg.norm(S @ X - X * th.cpu().numpy(), axis=0) / np.linalg.norm(X, axis=0)
    Xo = V @ Y_o
    res_o = np.linalg.norm(S @ Xo - Xo * th_o, axis=0) / np.linalg.norm(Xo, axis=0)
    assert np.all(res <= 10 * res_o + 1e-10)


def _split_basis(n, kx, kd, seed):
    """A solver-like basis [X | D]: X orthonormal, D correlated with X, its second
    half nearly parallel to its first (the two filter snapshots)."""
    rng = np.random.default_rng(seed)
    X = np.linalg.qr(rng.standard_normal((n, kx)))[0]
    D = rng.standard_normal((n, kd)) + X @ rng.standard_normal((kx, kd)) * 0.7
    h = kd // 2
    D[:, h:] = D[:, :kd - h] + 1e-4 * rng.standard_normal((n, kd - h))
    return np.concatenate([X, D], axis=1)


@pytest.mark.parametrize("kx,kd", [(64, 128), (64, 64), (32, 33), (16, 111), (128, 172), (7, 0)])
def test_rr_split_bitwise_vs_oracle(ctx, kx, kd):
    """The split Rayleigh-Ritz on the GPU runs the oracle's arithmetic: theta,
    Y and the kept rank are bitwise equal (D-block eigensolves of even and odd
    size, the LDS and global-memory eigensolvers, no D block at all)."""
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(3, 12)
    S = Ah.to_scipy()
    V = _split_basis(Ah.n, kx, kd, kx * 1000 + kd)
    G, H = V.T @ V, V.T @ (S @ V)
    nk = min(kx, 64)
    r_o, th_o, Y_o = O.rr_split(G, H, kx, 1e-12, nk)
    print(f"rr_split kx={kx} kd={kd}", flush=True)
    th, Y, r = api.rr_split(ctx, _dev(G.T).t(), _dev(H.T).t(), kx, 1e-12, nk)
    assert r == r_o
    assert np.array_equal(th.cpu().numpy(), th_o)
    assert np.array_equal(Y.cpu().numpy(), Y_o)


def test_rr_split_indefinite_block_reports_minus_one(ctx):
    """An X block that is not positive definite: rank -1, outputs untouched."""
    from mixedprecisionofrr_amd import api
    V = _split_basis(500, 8, 8, 3)
    G, H = V.T @ V, V.T @ V * 2.0
    G[2, 2] = -1.0
    print("rr_split indefinite X block", flush=True)
    th, Y, r = api.rr_split(ctx, _dev(G.T).t(), _dev(H.T).t(), 8, 1e-12, 8)
    assert r == -1
    assert np.all(np.isnan(th.cpu().numpy()))


def test_rr_generalized_vs_oracle(ctx):
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(13)
    V = rng.standard_normal((500, 30))
    M = rng.standard_normal((500, 500))
    S = M + M.T
    G, H = V.T @ V, V.T @ S @ V
    info, th_o, _ = O.rr_generalized(G, H, 0.0)
    th, Y = api.rr_generalized(ctx, _dev(G.T).t(), _dev(H.T).t(), 0.0)
    assert np.abs(th.cpu().numpy() - th_o).max() < 1e-11 * np.abs(th_o).max()


def test_residuals_vs_oracle(ctx):
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(2, 20)
    lam, U = np.linalg.eigh(Ah.to_scipy().toarray())
    X = U[:, -5:] + 1e-4 * U[:, :5]
    th = lam[-5:]
    ref = O.residuals(Ah, X, th, lam[-1])
    got = api.residuals(ctx, _csr(ctx, Ah), _dev(X.T).t(), th, lam[-1])
    assert np.allclose(got, ref, rtol=1e-12, atol=0)
    # row-major block (the layout the solvers keep): the coalesced row kernel
    got_r = api.residuals(ctx, _csr(ctx, Ah), _dev(X), th, lam[-1])
    assert np.allclose(got_r, ref, rtol=1e-12, atol=0)


def test_fp8_hardware_conversion_agrees(ctx):
    """The device FP8 conversions (gfx950's v_cvt_f32_fp8 / v_cvt_pk_fp8_f32,
    clamped to +-448) against the OCP E4M3 software definition: every code
    decodes identically and every float encodes identically (round to nearest
    even, saturating)."""
    import ctypes
    counts = (ctypes.c_int64 * 3)()
    ctx.check(ctx.lib.ofrr_fp8_check(ctx.h, counts), "fp8_check")
    print("fp8 mismatches (decode, encode <= 448, encode > 448):", list(counts), flush=True)
    assert counts[0] == 0 and counts[1] == 0 and counts[2] == 0
