"""Request planning of the serving path (ofrr_request_footprint): host only."""
import ctypes

import pytest
import torch

from mixedprecisionofrr_amd import _lib, api


def _state_bytes(n, k, nd, degree, bdt):
    L = _lib.load()
    o = _lib.FilterOpts(k, 1, nd, degree, 1, bdt, 1e-8, 1e-12, 0.0, 0.0)
    b = ctypes.c_int64()
    assert L.ofrr_filtered_eig_state_bytes(n, ctypes.byref(o), ctypes.byref(b)) == 0
    return b.value


@pytest.mark.parametrize("nwant,k,nd", [(1, 16, 2), (8, 16, 2), (13, 32, 2), (16, 32, 2),
                                        (32, 48, 2), (50, 64, 2), (52, 80, 1), (64, 80, 1),
                                        (100, 128, 1), (128, 128, 1)])
def test_block_width_and_corrections(nwant, k, nd):
    """k: the smallest multiple of 16 >= 1.25 nwant (16..128, >= nwant); two
    correction blocks while the basis stays <= 192 (the LDS-resident eigensolves)."""
    kk, ndd, _ = api.request_footprint(api.request(3, 64, nwant))
    assert (kk, ndd) == (k, nd)
    assert kk * (1 + ndd) <= 256
    if kk <= 64:
        assert kk * (1 + ndd) <= 192


def test_footprint_covers_operator_block_and_state():
    for dim, N, nwant, seed, dt in [(3, 128, 64, 5, torch.float16), (2, 2000, 8, 0, torch.float16),
                                    (3, 100, 32, 0, torch.float64), (2, 700, 16, 9, torch.bfloat16)]:
        k, nd, fp = api.request_footprint(api.request(dim, N, nwant, coef_seed=seed, basis_dtype=dt))
        n = N ** dim
        nnz = ctypes.c_int64()
        _lib.load().ofrr_stencil_nnz(dim, N, ctypes.byref(nnz))
        csr = 8 * (n + 1) + 12 * nnz.value
        state = _state_bytes(n, k, nd, 120, api.dtype_code(dt))
        # FP64 Ritz, new Ritz and AX blocks alone are 3 n k 8 bytes
        assert state >= 3 * n * k * 8
        assert fp >= csr + n * k * 8 + state
        assert fp <= csr + 4 * nnz.value + n * k * 8 + state + (64 << 20) + 8 * 256


def test_footprint_grows_with_size():
    a = api.request_footprint(api.request(3, 100, 16))[2]
    b = api.request_footprint(api.request(3, 150, 16))[2]
    c = api.request_footprint(api.request(3, 150, 64))[2]
    assert a < b < c


@pytest.mark.parametrize("bad", [dict(dim=4, N=10, nwant=8), dict(dim=2, N=2, nwant=8),
                                 dict(dim=2, N=100, nwant=0), dict(dim=2, N=100, nwant=129),
                                 dict(dim=3, N=1300, nwant=8), dict(dim=2, N=4, nwant=64)])
def test_invalid_requests_are_refused(bad):
    with pytest.raises(api.OfrrError):
        api.request_footprint(api.request(**bad))


def test_bad_tolerance_refused():
    with pytest.raises(api.OfrrError):
        api.request_footprint(api.request(2, 100, 8, tol=0.0))


def test_fixed_plan_fields():
    # the configured large FP8 solve: 100 pairs, basis 300 = 128 + 128 + 44
    k, nd, fp = api.request_footprint(api.request(3, 64, 100, k=128, nd=2, mid_cols=44,
                                                  basis_dtype=torch.float8_e4m3fn,
                                                  refine_dtype=torch.float16, refine_tol=1e-4))
    assert (k, nd) == (128, 2)
    n = 64 ** 3
    assert fp >= 3 * n * 128 * 8
    # a block width that is not a power of two is stored at the next one
    k, nd, fp = api.request_footprint(api.request(3, 64, 100, k=100, nd=2))
    assert (k, nd) == (100, 2) and fp >= 3 * n * 128 * 8
    assert api.request_footprint(api.request(3, 64, 16, nd=1))[:2] == (32, 1)
    # mid_cols >= k keeps the whole mid-filter block (basis 384, the largest)
    assert api.request_footprint(api.request(3, 64, 100, k=128, nd=2, mid_cols=129))[:2] == (128, 2)


@pytest.mark.parametrize("bad", [dict(k=8), dict(k=129), dict(nd=3), dict(refine_dtype=7)])
def test_invalid_plan_fields(bad):
    args = dict(dim=3, N=64, nwant=16)
    args.update(bad)
    rdt = args.pop("refine_dtype", None)
    req = api.request(**args)
    if rdt is not None:
        req.refine, req.refine_dtype = 1, rdt
    with pytest.raises(api.OfrrError):
        api.request_footprint(req)


def test_largest_basis_is_384():
    assert api.request_footprint(api.request(3, 64, 100, k=128, nd=2))[:2] == (128, 2)
