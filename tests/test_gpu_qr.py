"""Gram-Schmidt QR on the GPU (ofrr_gs_qr / ofrr_gs_qr_step) against the oracle and
the reference's own test cases (CPP/tests/qr_test.cpp).  Marked gpu."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEY = {torch.float64: "f64", torch.float32: "f32", torch.float16: "f16", torch.bfloat16: "bf16"}


def _colmajor(a, dtype=torch.float64):
    """device tensor of a (m x n) in column-major storage (stride(0) == 1)"""
    return torch.as_tensor(np.ascontiguousarray(a.T), dtype=dtype, device="cuda:0").t()


def _loss(Q, skip):
    k = skip == 0
    Qk = Q[:, k]
    return np.abs(Qk.T @ Qk - np.eye(int(k.sum()))).max()


def _host(Q, R, skip):
    return Q.double().cpu().numpy(), R.cpu().numpy(), skip.cpu().numpy()


@pytest.mark.parametrize("variant,dt", [("mgs", torch.float64), ("mgs", torch.float32),
                                        ("mgs", torch.float16), ("cgs2", torch.float64),
                                        ("cgs2", torch.float32), ("cgs2", torch.float16),
                                        ("mgs_v2", torch.float64)])
def test_reference_qr_cases(ctx, kats, variant, dt):
    """qr_test.cpp:30-651: Q^T Q = I and A = QR within the test's tolerance, then
    columns 3..5 copied from 0..2 are the dropped ones."""
    from mixedprecisionofrr_amd import api
    q = kats["qr"]
    key = KEY[dt]
    tol = q["tol"][variant][key]
    m, n = q["shape"]
    A = np.random.default_rng(815).uniform(-1, 1, (m, n))
    Ad = _colmajor(A, dt)
    print(f"gs_qr {variant} {key} {m}x{n}", flush=True)
    Q, R, skip = api.gs_qr(ctx, Ad, variant)
    Qh, Rh, sh = _host(Q, R, skip)
    Ah = Ad.double().cpu().numpy()
    assert sh.sum() == 0
    assert _loss(Qh, sh) < tol
    assert np.abs(Ah - Qh @ Rh).max() < tol
    assert np.all(np.tril(Rh, -1) == 0)
    # planted linear dependence
    A2 = A.copy()
    A2[:, q["dup_columns"]["copied_to"]] = A2[:, q["dup_columns"]["copied_from"]]
    A2d = _colmajor(A2, dt)
    stol = q["dup_select_tol"][variant][key]
    otol = q["dup_orth_tol"]["mgs"][key] if variant == "mgs" else None
    Q2, R2, skip2 = api.gs_qr(ctx, A2d, variant, select_tol=stol, orth_tol=otol)
    Q2h, R2h, s2h = _host(Q2, R2, skip2)
    assert list(np.nonzero(s2h)[0]) == q["dup_columns"]["copied_to"]
    assert np.abs(A2d.double().cpu().numpy() - Q2h @ R2h).max() < tol
    assert _loss(Q2h, s2h) < tol


@pytest.mark.parametrize("variant,rtol", [("mgs", -1.0), ("mgs", 2 ** -0.5), ("mgs_v2", -1.0),
                                          ("cgs", -1.0), ("cgs", 2 ** -0.5), ("cgs2", -1.0)])
@pytest.mark.parametrize("m,n", [(3000, 64), (300000, 24), (257, 257)])
def test_fp64_matches_oracle(ctx, variant, rtol, m, n):
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(m + n).uniform(-1, 1, (m, n))
    print(f"gs_qr fp64 {variant} rtol={rtol} {m}x{n}", flush=True)
    Q, R, skip = api.gs_qr(ctx, _colmajor(A), variant, reorth_tol=rtol)
    Qh, Rh, sh = _host(Q, R, skip)
    Qo, Ro, so = O.gs_qr(A, variant, reorth_tol=rtol)
    assert np.array_equal(sh, so)
    # the square matrix has condition ~1.8e3: single-pass CGS amplifies a change of
    # the summation order (a 1e-16 relative perturbation of A moves its Q by 1e-12)
    tq = 1e-10 if (m == n and variant == "cgs" and rtol < 0) else 1e-12
    assert np.abs(Qh - Qo).max() < tq
    assert np.abs(Rh - Ro).max() < tq * np.abs(Ro).max()
    assert np.all(np.tril(Rh, -1) == 0)


@pytest.mark.parametrize("kappa", [1e6, 1e8])
def test_fp64_loss_of_orthogonality_classes(ctx, kappa):
    """the textbook classes the oracle is pinned to, on a graded 4000 x 30 matrix"""
    from mixedprecisionofrr_amd import api
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((4000, 30)))
    V, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    A = U @ np.diag(np.logspace(0, -np.log10(kappa), 30)) @ V.T
    u = 2.0 ** -53
    loss = {}
    for variant, rtol in [("mgs", -1.0), ("mgs", 2 ** -0.5), ("mgs_v2", -1.0), ("cgs", -1.0),
                          ("cgs", 2 ** -0.5), ("cgs2", -1.0)]:
        print(f"gs_qr graded kappa={kappa:g} {variant} rtol={rtol}", flush=True)
        Q, R, skip = api.gs_qr(ctx, _colmajor(A), variant, reorth_tol=rtol)
        Qh, Rh, sh = _host(Q, R, skip)
        assert sh.sum() == 0
        assert np.abs(A - Qh @ Rh).max() < 1e-15
        loss[(variant, rtol > 0)] = _loss(Qh, sh)
    assert loss[("cgs", False)] > min(0.1, 1e-2 * u * kappa ** 2)
    for key in [("mgs", False), ("mgs_v2", False)]:
        assert 1e-3 * u * kappa < loss[key] < 1e3 * u * kappa
    for key in [("mgs", True), ("cgs", True), ("cgs2", False)]:
        assert loss[key] < 1e-14


@pytest.mark.parametrize("variant", ["mgs", "mgs_v2", "cgs", "cgs2"])
def test_fp64_drop_conventions(ctx, variant):
    """duplicated and zero columns: skip flags, R(i,i) of dropped columns (0 for MGS,
    1 otherwise), dropped columns left out of later projections, as the oracle"""
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(7).uniform(-1, 1, (5000, 40))
    A[:, 3:6] = A[:, 0:3]
    A[:, 9] = 0.0
    print(f"gs_qr drops {variant}", flush=True)
    Q, R, skip = api.gs_qr(ctx, _colmajor(A), variant, select_tol=1e-12, orth_tol=1e-12)
    Qh, Rh, sh = _host(Q, R, skip)
    Qo, Ro, so = O.gs_qr(A, variant, select_tol=1e-12, orth_tol=1e-12,
                         reorth_tol=2 ** -0.5 if variant == "mgs" else -1.0)
    assert list(np.nonzero(sh)[0]) == [3, 4, 5, 9] == list(np.nonzero(so)[0])
    rd = 0.0 if variant == "mgs" else 1.0
    assert all(Rh[i, i] == rd for i in (3, 4, 5, 9))
    assert np.all(Rh[[3, 4, 5, 9], 10:] == 0.0)
    assert np.all(Qh[:, 9] == 0.0)
    kept = sh == 0
    assert np.abs(Qh[:, kept] - Qo[:, kept]).max() < 1e-12
    assert np.abs(A - Qh @ Rh).max() < 1e-12


@pytest.mark.parametrize("variant", ["mgs", "cgs2"])
def test_row_major_and_determinism(ctx, variant):
    """the same factorization in row-major storage, bitwise; two runs bitwise equal"""
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(21).uniform(-1, 1, (20000, 48))
    Qc, Rc, sc = api.gs_qr(ctx, _colmajor(A), variant)
    Qr, Rr, sr = api.gs_qr(ctx, torch.as_tensor(A, device="cuda:0"), variant)
    assert Qr.stride(1) == 1
    Qc2, Rc2, _ = api.gs_qr(ctx, _colmajor(A), variant)
    assert torch.equal(Qc, Qr) and torch.equal(Rc, Rr) and torch.equal(sc, sr)
    assert torch.equal(Qc, Qc2) and torch.equal(Rc, Rc2)


@pytest.mark.parametrize("variant,exact", [("mgs", False), ("cgs", True), ("cgs2", True)])
def test_column_steps_match_full(ctx, variant, exact):
    """mgs_step / cgs_step / cgs2_step over i = 0..n-1 give the full factorization
    (CGS: the same launches, bitwise)"""
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(4).uniform(-1, 1, (6000, 32))
    A[:, 7] = A[:, 2]
    Q, R, skip = api.gs_qr(ctx, _colmajor(A), variant, select_tol=1e-12)
    Qs = _colmajor(A)
    Rs = torch.zeros(32, 32, dtype=torch.float64, device="cuda:0").t()
    ss = torch.zeros(32, dtype=torch.int32, device="cuda:0")
    flags = [api.gs_qr_step(ctx, Qs, Rs, ss, i, variant, select_tol=1e-12) for i in range(32)]
    assert flags == list(skip.cpu().numpy()) and flags[7] == 1
    assert torch.equal(ss, skip)
    if exact:
        assert torch.equal(Qs, Q) and torch.equal(Rs, R)
    else:
        assert (Qs - Q).abs().max().item() < 1e-13
        assert (Rs - R).abs().max().item() < 1e-12


def test_bf16_runs(ctx):
    from mixedprecisionofrr_amd import api
    A = np.random.default_rng(8).uniform(-1, 1, (4096, 64))
    Q, R, skip = api.gs_qr(ctx, _colmajor(A, torch.bfloat16), "cgs2")
    Qh, Rh, sh = _host(Q, R, skip)
    assert sh.sum() == 0 and _loss(Qh, sh) < 0.1


def test_invalid_arguments(ctx):
    from mixedprecisionofrr_amd import api
    A = _colmajor(np.ones((10, 20)))
    with pytest.raises(api.OfrrError) as e:
        api.gs_qr(ctx, A, "mgs")                       # wide
    assert e.value.status == 1000
    B = _colmajor(np.random.default_rng(0).uniform(-1, 1, (50, 5)))
    R = torch.zeros(5, 5, dtype=torch.float64, device="cuda:0").t()
    s = torch.zeros(5, dtype=torch.int32, device="cuda:0")
    with pytest.raises(api.OfrrError) as e:
        api.gs_qr_step(ctx, B, R, s, 0, "mgs_v2")      # no column step for mgs_v2
    assert e.value.status == 1000
    with pytest.raises(api.OfrrError) as e:
        api.gs_qr_step(ctx, B, R, s, 5, "mgs")         # column out of range
    assert e.value.status == 1000
