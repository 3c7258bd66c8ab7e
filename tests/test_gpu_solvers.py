"""GPU solver drivers against the CPU FP64 oracle and closed-form spectra.
Marked gpu."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _csr(ctx, Ah):
    from mixedprecisionofrr_amd import api
    return api.CSRMatrix.from_host(ctx, Ah.row_ptr, Ah.col_idx, Ah.vals)


def test_arnoldi_laplacian64_matches_cpu(ctx):
    """2D 64x64 Laplacian, kdim 60, 20 largest, N(0,1) start from seed 0:
    FP64 GPU Ritz values within 1e-12 (relative) of the CPU FP64 oracle."""
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(2, 64)
    A = api.CSRMatrix.stencil(ctx, 2, 64)
    v0 = torch.empty(Ah.n, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, v0, seed=0)
    v0h = v0.cpu().numpy()
    for maxiter in (61, 1 + 60 * 3):
        print(f"arnoldi 64x64 maxiter={maxiter}", flush=True)
        th, X, built = api.arnoldi_hess_eig(ctx, A, 20, 60, maxiter, v0)
        th_o, X_o, built_o = O.arnoldi_hess_eig(Ah, 20, 60, maxiter, v0h)
        assert built == built_o == 61
        # the FP64 pipeline runs the same operations in the same order on both sides
        assert np.max(np.abs(th - th_o) / np.abs(th_o)) <= 1e-14
        res = O.residuals(Ah, X.cpu().numpy(), th, O.laplacian_eigs(2, 64)[0])
        res_o = O.residuals(Ah, X_o, th_o, O.laplacian_eigs(2, 64)[0])
        assert np.allclose(res, res_o, rtol=1e-6, atol=1e-13)


def test_arnoldi_exact_termination(ctx):
    from mixedprecisionofrr_amd import api
    N = 16
    Ah = O.stencil(2, N)
    lam, U = np.linalg.eigh(Ah.to_scipy().toarray())
    v0 = U[:, [255, 250, 240]] @ np.array([1.0, 0.5, 0.25])
    th, X, built = api.arnoldi_hess_eig(ctx, _csr(ctx, Ah), 3, 10, 11,
                                        torch.as_tensor(v0, device="cuda:0"), select_tol=1e-10)
    assert built == 3
    assert np.allclose(th, np.sort(lam[[255, 250, 240]])[::-1], rtol=1e-12)


def test_subspace_iteration_matches_cpu(ctx):
    from mixedprecisionofrr_amd import api
    Ah = O.stencil(2, 12)
    V0 = O.randn(0, Ah.n * 8).reshape(Ah.n, 8, order="F")
    th_o, V_o = O.subspace_iter_eig(Ah, V0, maxiter=40, step_size=2)
    th, V = api.subspace_iter_eig(ctx, _csr(ctx, Ah), torch.as_tensor(V0, device="cuda:0"),
                                  maxiter=40, step_size=2)
    assert np.max(np.abs(th - th_o) / np.abs(th_o)) <= 1e-14


@pytest.mark.parametrize("nd", [1, 2])
def test_filtered_fp64_matches_cpu(ctx, nd):
    """Same algorithm, same start block, fixed iterations: Ritz values agree to 1e-12."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 14, 32, 20
    Ah = O.stencil(3, N)
    X0 = O.randn(0, Ah.n * k).reshape(Ah.n, k)
    for iters in (1, 4):
        it_o, th_o, X_o, res_o = O.filtered_ofrr(Ah, np.asfortranarray(X0), nwant, nd=nd, degree=10,
                                                 maxiter=iters, tol=1e-30)
        print(f"filtered fp64 nd={nd} iters={iters}", flush=True)
        st, th, X, res, stats = api.filtered_eig(ctx, _csr(ctx, Ah), torch.as_tensor(X0, device="cuda:0"),
                                                 nwant, nd=nd, degree=10, maxiter=iters,
                                                 basis_dtype=torch.float64, tol=1e-30)
        assert st == 4001 and stats["iterations"] == iters, (st, stats)
        assert np.max(np.abs(th - th_o) / np.abs(th_o)) < 1e-12
        # individual Ritz vectors are not unique inside the Laplacian's degenerate
        # clusters; the residual of the whole block is
        fro, fro_o = np.sqrt(np.sum(res[:nwant] ** 2)), np.sqrt(np.sum(res_o[:nwant] ** 2))
        assert abs(fro - fro_o) <= 1e-6 * fro_o


def test_filtered_fp64_strip_schedule_matches_cpu(ctx):
    """64^3 grid: bandwidth 64^2 rows = 16 tiles of 256 rows, so the CSR kernels
    walk the row tiles in the strip schedule (chunks of 16 tiles, 8 strips of 2
    tiles, one per XCD).  Every row is still computed once in the same order of
    its entries: the FP64 Ritz values agree with the CPU restatement to 1e-12."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 64, 16, 8
    Ah = O.stencil(3, N)
    X0 = O.randn(2, Ah.n * k).reshape(Ah.n, k)
    it_o, th_o, X_o, res_o = O.filtered_ofrr(Ah, np.asfortranarray(X0), nwant, nd=1, degree=10,
                                             maxiter=2, tol=1e-30)
    print("filtered fp64 strip schedule N=64", flush=True)
    st, th, X, res, stats = api.filtered_eig(ctx, _csr(ctx, Ah), torch.as_tensor(X0, device="cuda:0"),
                                             nwant, nd=1, degree=10, maxiter=2,
                                             basis_dtype=torch.float64, tol=1e-30)
    assert st == 4001 and stats["iterations"] == 2, (st, stats)
    assert np.max(np.abs(th - th_o) / np.abs(th_o)) < 1e-12
    fro, fro_o = np.sqrt(np.sum(res[:nwant] ** 2)), np.sqrt(np.sum(res_o[:nwant] ** 2))
    assert abs(fro - fro_o) <= 1e-6 * fro_o


@pytest.mark.parametrize("nd,rr_whole", [(1, False), (2, False), (1, True), (2, True)])
def test_filtered_fp64_soft_locking_matches_cpu(ctx, nd, rr_whole):
    """Soft locking engaged (the leading pairs converge first and the later
    sweeps filter 16 of the 32 columns; with two correction blocks they are
    packed for the products): the FP64 path converges in the iterations of the
    CPU restatement, to the same Ritz values (1e-12) and the closed-form
    spectrum, with the split Rayleigh-Ritz (default) and the whole form."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 14, 32, 20
    Ah = O.stencil(3, N)
    X0 = O.randn(0, Ah.n * k).reshape(Ah.n, k)
    # tol is chosen away from the residuals at the deciding iterations: at 1e-9 the
    # last residuals are 9.96e-10 (split, one block) and 9.92e-10 (split, two
    # blocks) on the CPU, within rounding of tol
    tol = 2e-9
    it_o, th_o, X_o, res_o = O.filtered_ofrr(Ah, np.asfortranarray(X0), nwant, nd=nd, degree=10,
                                             maxiter=60, tol=tol, rr_whole=rr_whole)
    print(f"filtered fp64 soft locking N=14 k=32 nd={nd} rr_whole={rr_whole}", flush=True)
    st, th, X, res, stats = api.filtered_eig(ctx, _csr(ctx, Ah), torch.as_tensor(X0, device="cuda:0"),
                                             nwant, nd=nd, degree=10, maxiter=60,
                                             basis_dtype=torch.float64, tol=tol, rr_whole=rr_whole)
    assert stats["rr_fallbacks"] == 0
    assert st == 0 and stats["iterations"] == it_o, (st, stats, it_o)
    # the iteration counts can only agree by design if no deciding residual sits
    # within rounding of tol (a fragile tol would fail here, not by a count)
    assert np.all(np.abs(res_o[:nwant] - tol) > 0.01 * tol), res_o[:nwant]
    assert stats["filter_col_steps"] < stats["filter_steps"] * k, stats
    assert stats["filter_col_steps"] % 16 == 0
    assert np.max(np.abs(th - th_o)[:nwant] / np.abs(th_o[:nwant])) < 1e-12
    ex = O.laplacian_eigs(3, N)[:nwant]
    assert np.abs(th[:nwant] - ex).max() / ex[0] < 1e-12
    assert res[:nwant].max() <= tol


@pytest.mark.parametrize("bdt", [torch.float16, torch.bfloat16, torch.float8_e4m3fn, torch.float64])
def test_filtered_soft_locking_formats(ctx, bdt):
    """k = 64, basis 192, degree 120 on the 32^3 Laplacian, every growth format:
    with and without soft locking the solve reaches residual 1e-8 and the
    closed-form eigenvalues; locking filters fewer columns (FP8 refines in FP16).
    FP64 runs at tol 1e-9: at 1e-8 all 50 pairs meet tol in the same iteration
    (the third), so no sweep is ever locked."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 32, 64, 50
    tol = 1e-9 if bdt == torch.float64 else 1e-8
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=0)
    ex = O.laplacian_eigs(3, N)[:nwant]
    Ah = O.stencil(3, N)
    fp8 = bdt == torch.float8_e4m3fn
    out = {}
    for lock in (True, False):
        print(f"filtered soft locking basis={bdt} lock={lock}", flush=True)
        st, th, X, res, stats = api.filtered_eig(
            ctx, A, X0, nwant, nd=2, degree=120, maxiter=100, basis_dtype=bdt, tol=tol,
            lock=lock, refine_dtype=torch.float16 if fp8 else None, refine_tol=1e-4)
        assert st == 0, (lock, stats)
        assert np.max(np.abs(th[:nwant] - ex) / ex) < 1e-10
        assert O.residuals(Ah, X.cpu().numpy(), th, ex[0])[:nwant].max() <= 1.05 * tol
        out[lock] = stats
    assert out[False]["filter_col_steps"] == out[False]["filter_steps"] * k
    # locking engages in every format: some sweep filtered fewer than k columns
    assert out[True]["filter_col_steps"] < out[True]["filter_steps"] * k, out[True]
    assert out[True]["filter_col_steps"] < out[False]["filter_col_steps"]


@pytest.mark.parametrize("nd", [1, 2])
def test_filtered_fp64_padded_width_matches_cpu(ctx, nd):
    """A block width that is not a power of two (k = 24, stored 32 wide with zero
    columns, the Rayleigh-Ritz on the compacted matrices): the FP64 Ritz values
    of fixed iterations agree with the CPU restatement at k = 24 to 1e-12."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 14, 24, 12
    Ah = O.stencil(3, N)
    X0 = O.randn(3, Ah.n * k).reshape(Ah.n, k)
    for iters in (1, 4):
        it_o, th_o, X_o, res_o = O.filtered_ofrr(Ah, np.asfortranarray(X0), nwant, nd=nd, degree=10,
                                                 maxiter=iters, tol=1e-30)
        print(f"filtered fp64 padded k=24 nd={nd} iters={iters}", flush=True)
        st, th, X, res, stats = api.filtered_eig(ctx, _csr(ctx, Ah), torch.as_tensor(X0, device="cuda:0"),
                                                 nwant, nd=nd, degree=10, maxiter=iters,
                                                 basis_dtype=torch.float64, tol=1e-30, lock=False)
        assert st == 4001 and stats["iterations"] == iters, (st, stats)
        assert th.shape == (k,) and X.shape == (Ah.n, k)
        assert np.max(np.abs(th - th_o) / np.abs(th_o)) < 1e-12
        fro, fro_o = np.sqrt(np.sum(res[:nwant] ** 2)), np.sqrt(np.sum(res_o[:nwant] ** 2))
        assert abs(fro - fro_o) <= 1e-6 * fro_o


@pytest.mark.parametrize("bdt", [torch.float16, torch.float8_e4m3fn])
def test_filtered_padded_width_converges(ctx, bdt):
    """k = 40 (stored 64 wide), two correction blocks (basis 120), soft locking
    on: the mixed solve reaches residual 1e-8 and the closed-form spectrum."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 24, 40, 32
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=4)
    fp8 = bdt == torch.float8_e4m3fn
    print(f"filtered padded k=40 basis={bdt}", flush=True)
    st, th, X, res, stats = api.filtered_eig(
        ctx, A, X0, nwant, nd=2, degree=60, maxiter=100, basis_dtype=bdt, tol=1e-8,
        refine_dtype=torch.float64 if fp8 else None, refine_tol=1e-4)
    assert st == 0, stats
    ex = O.laplacian_eigs(3, N)[:nwant]
    assert np.max(np.abs(th[:nwant] - ex) / ex) < 1e-10
    Ah = O.stencil(3, N)
    assert O.residuals(Ah, X.cpu().numpy(), th, ex[0])[:nwant].max() <= 1.05e-8


def _box_laplacian(nx, ny, nz):
    """7-point Dirichlet Laplacian on an nx x ny x nz box, x fastest (scipy kron sums)."""
    import scipy.sparse as sp
    T = lambda m: sp.diags([-np.ones(m - 1), 2 * np.ones(m), -np.ones(m - 1)], [-1, 0, 1])
    I = sp.identity
    L = (sp.kron(I(nz), sp.kron(I(ny), T(nx))) + sp.kron(I(nz), sp.kron(T(ny), I(nx)))
         + sp.kron(T(nz), sp.kron(I(ny), I(nx))))
    L = sp.csr_matrix(L)
    L.sort_indices()
    return O.CSR(L.shape[0], L.indptr.astype(np.int64), L.indices.astype(np.int32),
                 L.data.astype(np.float64))


def test_filtered_fp64_strip_schedule_box_matches_cpu(ctx):
    """128 x 128 x 16 box, k = 64 FP64: planes of 64 tiles in 16 chunks, each XCD
    walking strips of 8 tiles (the schedule and row width of the 128^3 FP64
    path).  Same algorithm and start block as the CPU restatement: Ritz values
    agree to 1e-12."""
    from mixedprecisionofrr_amd import api
    Ah = _box_laplacian(128, 128, 16)
    k, nwant = 64, 32
    X0 = O.randn(3, Ah.n * k).reshape(Ah.n, k)
    it_o, th_o, X_o, res_o = O.filtered_ofrr(Ah, np.asfortranarray(X0), nwant, nd=1, degree=10,
                                             maxiter=1, tol=1e-30)
    print("filtered fp64 strip schedule 128x128x16 k=64", flush=True)
    st, th, X, res, stats = api.filtered_eig(ctx, _csr(ctx, Ah), torch.as_tensor(X0, device="cuda:0"),
                                             nwant, nd=1, degree=10, maxiter=1,
                                             basis_dtype=torch.float64, tol=1e-30)
    assert st == 4001 and stats["iterations"] == 1, (st, stats)
    assert np.max(np.abs(th - th_o) / np.abs(th_o)) < 1e-12
    fro, fro_o = np.sqrt(np.sum(res[:nwant] ** 2)), np.sqrt(np.sum(res_o[:nwant] ** 2))
    assert abs(fro - fro_o) <= 1e-6 * fro_o


@pytest.mark.parametrize("bdt", [torch.float16, torch.bfloat16, torch.float32, torch.float64])
def test_filtered_mixed_reaches_1e8(ctx, bdt):
    """Mixed bases reach residual 1e-8 and the closed-form eigenvalues."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 24, 64, 50
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=0)
    print(f"filtered mixed basis={bdt}", flush=True)
    st, th, X, res, stats = api.filtered_eig(ctx, A, X0, nwant, nd=2, degree=16, maxiter=80,
                                             basis_dtype=bdt, tol=1e-8)
    assert st == 0, stats
    ex = O.laplacian_eigs(3, N)[:nwant]
    assert np.max(np.abs(th[:nwant] - ex) / ex) < 1e-10
    Ah = O.stencil(3, N)
    res_o = O.residuals(Ah, X.cpu().numpy(), th, ex[0])
    assert res_o[:nwant].max() <= 1.05e-8


def test_filtered_wide_sweeps(ctx):
    """FP16 growth blocks with FP32 for the sweeps after the first that start
    above residual 1e-3 (wide_dtype / wide_tol): the wide sweeps are counted,
    the solve reaches residual 1e-8 and the closed-form spectrum, and it takes
    fewer iterations than with every sweep in FP16 (whose second sweep stalls on
    the rounding of the correction blocks; the FP64 path's count is the floor)."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 40, 64, 50
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=0)
    print("filtered wide sweeps N=40 k=64", flush=True)
    runs = {}
    for name, bdt, wide in [("f16", torch.float16, None), ("wide", torch.float16, torch.float32),
                            ("f64", torch.float64, None)]:
        runs[name] = api.filtered_eig(ctx, A, X0, nwant, nd=2, degree=120, maxiter=60,
                                      basis_dtype=bdt, tol=1e-8, wide_dtype=wide,
                                      wide_tol=1e-3 if wide is not None else 0.0)
        assert runs[name][0] == 0, (name, runs[name][4])
    st, th, X, res, stats = runs["wide"]
    assert stats["wide_sweeps"] >= 1 and runs["f16"][4]["wide_sweeps"] == 0
    ex = O.laplacian_eigs(3, N)[:nwant]
    assert np.max(np.abs(th[:nwant] - ex) / ex) < 1e-10
    res_o = O.residuals(O.stencil(3, N), X.cpu().numpy(), th, ex[0])
    assert res_o[:nwant].max() <= 1.05e-8
    its = {name: r[4]["iterations"] for name, r in runs.items()}
    assert its["f64"] <= its["wide"] < its["f16"], its


@pytest.mark.parametrize("N", [40, 64])
def test_filtered_fp32_basis_high_degree(ctx, N):
    """FP32 growth blocks at degree 120: their Gram matrix resolves directions
    down to ~1e-15 of its largest eigenvalue, so Z = A D must be FP64 (an FP32 Z
    put spurious Ritz values above the wanted ones: residual ~1 from the second
    iteration on).  The solve reaches 1e-8 and the closed-form spectrum in no
    more iterations than the FP16 one."""
    from mixedprecisionofrr_amd import api
    k, nwant = 64, 50
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=0)
    print(f"filtered fp32 basis degree 120 N={N}", flush=True)
    its = {}
    for bdt in (torch.float32, torch.float16):
        st, th, X, res, stats = api.filtered_eig(ctx, A, X0, nwant, nd=2, degree=120, maxiter=40,
                                                 basis_dtype=bdt, tol=1e-8)
        assert st == 0, (bdt, stats)
        its[bdt] = stats["iterations"]
        if bdt == torch.float32:
            ex = O.laplacian_eigs(3, N)[:nwant]
            assert np.max(np.abs(th[:nwant] - ex) / ex) < 1e-10
            res_o = O.residuals(O.stencil(3, N), X.cpu().numpy(), th, ex[0])
            assert res_o[:nwant].max() <= 1.05e-8
    assert its[torch.float32] <= its[torch.float16], its


def test_filtered_wide_format_must_be_wider(ctx):
    """The wide format must be wider than the growth format."""
    from mixedprecisionofrr_amd import api
    A = api.CSRMatrix.stencil(ctx, 3, 8).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, 16, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=0)
    with pytest.raises(Exception, match="wide format"):
        api.filtered_eig(ctx, A, X0, 8, nd=1, degree=8, maxiter=4, basis_dtype=torch.float16,
                         wide_dtype=torch.bfloat16, wide_tol=1e-3)


def test_filtered_same_alone_and_repeated(ctx):
    """Deterministic reductions: the same solve twice gives bitwise equal results."""
    from mixedprecisionofrr_amd import api
    A = api.CSRMatrix.stencil(ctx, 3, 16, seed=3).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, 32, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=1)
    r1 = api.filtered_eig(ctx, A, X0, 16, nd=1, degree=12, maxiter=40, basis_dtype=torch.float16)
    r2 = api.filtered_eig(ctx, A, X0, 16, nd=1, degree=12, maxiter=40, basis_dtype=torch.float16)
    assert r1[0] == 0
    assert np.array_equal(r1[1], r2[1]) and torch.equal(r1[2], r2[2])


@pytest.mark.parametrize("refine", [torch.float16, torch.float64])
def test_filtered_fp8_precision_ladder_reaches_1e8(ctx, refine):
    """FP8 E4M3 growth blocks with the precision ladder: the solve switches to
    the refinement format (FP16, or FP64 with the FP64 operator values) and
    reaches residual 1e-8 with eigenvalues equal to the closed form."""
    from mixedprecisionofrr_amd import api
    N, k, nwant = 40, 32, 20
    A = api.CSRMatrix.stencil(ctx, 3, N).with_low_precision(torch.float16)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=9)
    print(f"fp8 ladder -> {refine}", flush=True)
    st, th, X, res, stats = api.filtered_eig(ctx, A, X0, nwant, nd=2, degree=60, maxiter=200,
                                             basis_dtype=torch.float8_e4m3fn, tol=1e-8,
                                             refine_dtype=refine, refine_tol=1e-4)
    assert st == 0, (st, stats)
    exact = O.laplacian_eigs(3, N)[:nwant]
    assert np.max(np.abs(th[:nwant] - exact) / exact) < 1e-11
    res_o = O.residuals(O.stencil(3, N), X.cpu().numpy(), th, exact[0])
    assert np.max(res_o[:nwant]) < 2e-8


@pytest.mark.parametrize("case", ["f16_locking", "fp8_ladder", "f16_padded"])
def test_filtered_graph_sweeps_match_plain_launches(ctx, case):
    """Filter sweeps recorded as HIP graphs and replayed give bitwise the same
    solve as plain launches: every iteration replays (or re-records) a sweep
    between the solver's own eager work and host synchronisations, so any
    replay that differed would change the Ritz values, residuals or block.
    Soft locking narrows the sweeps (new shapes), the FP8 ladder switches the
    format, and a padded width compacts the Rayleigh-Ritz.  A second graph
    solve on new buffers re-records through the context's graphs of the same
    shapes and matches too."""
    from mixedprecisionofrr_amd import api
    N, k, nwant, deg, kw = 28, 64, 40, 40, {}
    bdt = torch.float16
    if case == "fp8_ladder":
        N, k, nwant, deg, bdt = 40, 32, 20, 60, torch.float8_e4m3fn
        kw = dict(refine_dtype=torch.float16, refine_tol=1e-4)
    if case == "f16_padded":
        k, nwant = 48, 36
    A = api.CSRMatrix.stencil(ctx, 3, N, seed=4).with_low_precision(torch.float32)
    X0 = torch.empty(A.n, k, dtype=torch.float64, device="cuda:0")
    api.fill_randn(ctx, X0, seed=2)
    print(f"graph sweeps vs plain: {case}", flush=True)
    args = dict(nd=2, degree=deg, maxiter=200, basis_dtype=bdt, tol=1e-8, **kw)
    plain = api.filtered_eig(ctx, A, X0, nwant, graph=False, **args)
    g1 = api.filtered_eig(ctx, A, X0, nwant, graph=True, **args)
    g2 = api.filtered_eig(ctx, A, X0.clone(), nwant, graph=True, **args)
    assert plain[0] == 0, plain[4]
    assert plain[4]["graph_replays"] == 0
    for g in (g1, g2):
        assert g[0] == plain[0] and g[4]["iterations"] == plain[4]["iterations"]
        assert np.array_equal(g[1], plain[1]) and np.array_equal(g[3], plain[3])
        assert torch.equal(g[2], plain[2])
        st = g[4]
        # one graph launch per sweep, at most one recording per sweep
        assert st["graph_replays"] == st["iterations"]
        assert st["graph_captures"] <= st["graph_replays"]
        assert st["filter_steps"] == plain[4]["filter_steps"]
    # the first graph solve records each shape; the second finds them in the
    # context (same shapes, and here the same solver buffers), so the two
    # solves together record fewer sweeps than they launch
    assert g1[4]["graph_captures"] >= 1
    assert g2[4]["graph_captures"] <= g1[4]["graph_captures"]
    assert (g1[4]["graph_captures"] + g2[4]["graph_captures"]
            < g1[4]["graph_replays"] + g2[4]["graph_replays"])
