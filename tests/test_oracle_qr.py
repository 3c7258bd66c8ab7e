"""The oracle's Gram-Schmidt QR (orc_gs_qr) pinned without the reference:

* full-rank input: every variant equals numpy's Householder QR with the signs
  fixed so that diag(R) > 0 (the QR factorization is unique);
* graded input (condition number kappa): the textbook loss of orthogonality of
  each variant -- CGS ~ u kappa^2, MGS ~ u kappa, reorthogonalized MGS/CGS and
  CGS2 ~ u -- which tells a CGS from an MGS and a pass that is skipped from one
  that is applied;
* the reference's duplicated-column case (CPP/tests/qr_test.cpp:90-120, columns
  3..5 copied from 0..2) and the drop conventions of qr.cu (R(i,i) = 0 for MGS,
  qr.cu:70,217-246; 1 for the others, qr.cu:315,459,636).
"""
import numpy as np
import pytest

from oracle import oracle as O

VARIANTS = [("mgs", -1.0), ("mgs", 2 ** -0.5), ("mgs_v2", -1.0), ("cgs", -1.0),
            ("cgs", 2 ** -0.5), ("cgs2", -1.0)]


def _loss(Q, skip):
    k = skip == 0
    Qk = Q[:, k]
    return np.abs(Qk.T @ Qk - np.eye(int(k.sum()))).max()


def _graded(m, n, kappa, seed=3):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return U @ np.diag(np.logspace(0, -np.log10(kappa), n)) @ V.T


@pytest.mark.parametrize("variant,rtol", VARIANTS)
def test_full_rank_equals_householder(variant, rtol):
    A = np.random.default_rng(11).uniform(-1, 1, (300, 40))
    Q, R, skip = O.gs_qr(A, variant, reorth_tol=rtol)
    Qh, Rh = np.linalg.qr(A)
    d = np.sign(np.diag(Rh))
    assert skip.sum() == 0
    assert np.abs(Q - Qh * d).max() < 1e-14
    assert np.abs(R - d[:, None] * Rh).max() < 1e-13
    assert np.all(np.tril(R, -1) == 0)


@pytest.mark.parametrize("kappa", [1e6, 1e8])
def test_loss_of_orthogonality_classes(kappa):
    A = _graded(400, 30, kappa)
    u = 2.0 ** -53
    loss = {}
    for variant, rtol in VARIANTS:
        Q, R, skip = O.gs_qr(A, variant, reorth_tol=rtol)
        assert skip.sum() == 0
        assert np.abs(A - Q @ R).max() < 1e-15      # backward stable, every variant
        loss[(variant, rtol > 0)] = _loss(Q, skip)
    # CGS: ~ u kappa^2 (saturates at 1)
    assert loss[("cgs", False)] > min(0.1, 1e-2 * u * kappa ** 2)
    # MGS (either order): ~ u kappa, far from both u and u kappa^2
    for key in [("mgs", False), ("mgs_v2", False)]:
        assert 1e-3 * u * kappa < loss[key] < 1e3 * u * kappa
    assert loss[("mgs", False)] == loss[("mgs_v2", False)]
    # reorthogonalized / two passes: ~ u
    for key in [("mgs", True), ("cgs", True), ("cgs2", False)]:
        assert loss[key] < 1e-14


def test_reorthogonalization_triggers_on_cancellation():
    rng = np.random.default_rng(5)
    a = rng.standard_normal(500)
    A = np.stack([a, a + 1e-9 * rng.standard_normal(500), rng.standard_normal(500)], axis=1)
    Q0, _, s0 = O.gs_qr(A, "mgs", reorth_tol=-1.0)
    Q1, R1, s1 = O.gs_qr(A, "mgs", reorth_tol=2 ** -0.5)
    assert s0.sum() == 0 and s1.sum() == 0
    assert _loss(Q0, s0) > 1e-9
    assert _loss(Q1, s1) < 1e-14
    assert np.abs(A - Q1 @ R1).max() < 1e-14


@pytest.mark.parametrize("variant", ["mgs", "mgs_v2", "cgs", "cgs2"])
def test_duplicated_columns_dropped(kats, variant):
    q = kats["qr"]
    m, n = q["shape"]
    A = np.random.default_rng(815).uniform(-1, 1, (m, n))
    A[:, q["dup_columns"]["copied_to"]] = A[:, q["dup_columns"]["copied_from"]]
    stol = q["dup_select_tol"].get(variant, q["dup_select_tol"]["cgs2"])["f64"]
    Q, R, skip = O.gs_qr(A, variant, select_tol=stol, orth_tol=stol,
                         reorth_tol=q["reorth_tol"] if variant == "mgs" else -1.0)
    assert list(np.nonzero(skip)[0]) == q["dup_columns"]["copied_to"]
    assert _loss(Q, skip) < 1e-10
    assert np.abs(A - Q @ R).max() < 1e-10
    rd = 0.0 if variant == "mgs" else 1.0
    assert all(R[i, i] == rd for i in (3, 4, 5))
    assert np.all(R[3:6, 6:] == 0.0)          # dropped columns take no part later
    assert np.abs(Q[:, 3:6]).max() < stol     # their residual stays in Q


@pytest.mark.parametrize("variant", ["mgs", "mgs_v2", "cgs", "cgs2"])
def test_zero_column(variant):
    A = np.random.default_rng(2).uniform(-1, 1, (50, 6))
    A[:, 2] = 0.0
    Q, R, skip = O.gs_qr(A, variant)
    assert list(skip) == [0, 0, 1, 0, 0, 0]
    assert np.all(Q[:, 2] == 0.0)
    rd = 0.0 if variant == "mgs" else 1.0
    col = R[:, 2].copy()
    assert col[2] == rd
    col[2] = 0.0
    assert np.all(col == 0.0)
    assert _loss(Q, skip) < 1e-14
