"""ofrr_gram_noalloc (the projection product with its chunk partials in a
static device array instead of allocated workspace) against ofrr_gram through
the C-ABI: two calls in a row, the second smaller, each within FP64 rounding
of the allocating entry point (their chunkings differ, so not bit for bit);
sizes past the scratch are refused."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _gram(ctx, fn, V, W, alpha=1.0, beta=0.0, G=None):
    from mixedprecisionofrr_amd import api
    n, k1 = V.shape
    k2 = W.shape[1]
    if G is None:
        G = torch.zeros(k2, k1, dtype=torch.float64, device=V.device).t()
    st = fn(ctx.h, n, k1, k2, C.c_void_p(V.data_ptr()), api.dtype_code(V), V.stride(0), V.stride(1),
            C.c_void_p(W.data_ptr()), api.dtype_code(W), W.stride(0), W.stride(1), alpha, beta,
            C.c_void_p(G.data_ptr()), G.stride(1))
    return st, G


def _bound(V, W):
    # |V^T W - G| <= n u sum |v| |w| for FP64 accumulation in any order
    a, b = V.double().abs(), W.double().abs()
    return (a.t() @ b).cpu().numpy() * V.shape[0] * 2.0 ** -53 * 2


def test_gram_noalloc_matches_gram(ctx):
    g = torch.Generator(device="cuda:0").manual_seed(7)
    cases = [(400_000, 64, 192, torch.float16, torch.float64), (5_000, 40, 24, torch.float64, torch.float32)]
    for n, k1, k2, vdt, wdt in cases:
        print(f"gram_noalloc n={n} k1={k1} k2={k2} {vdt} x {wdt}", flush=True)
        V = torch.randn(n, k1, device="cuda:0", generator=g).to(vdt)
        W = torch.randn(n, k2, device="cuda:0", generator=g).to(wdt)
        st0, G0 = _gram(ctx, ctx.lib.ofrr_gram, V, W)
        st1, G1 = _gram(ctx, ctx.lib.ofrr_gram_noalloc, V, W)
        torch.cuda.synchronize()
        assert st0 == 0 and st1 == 0, (st0, st1)
        diff = np.abs(G1.cpu().numpy() - G0.cpu().numpy())
        assert np.all(diff <= _bound(V, W)), diff.max()
    # alpha / beta against the allocating path, column-major W of a wider buffer
    V = torch.randn(70_000, 64, device="cuda:0", generator=g, dtype=torch.float64)
    Wb = torch.randn(48, 70_000, device="cuda:0", generator=g, dtype=torch.float64)
    W = Wb[:32].t()
    Gi = torch.randn(32, 64, device="cuda:0", generator=g, dtype=torch.float64).t()
    st0, G0 = _gram(ctx, ctx.lib.ofrr_gram, V, W, 0.5, -2.0, Gi.clone())
    st1, G1 = _gram(ctx, ctx.lib.ofrr_gram_noalloc, V, W, 0.5, -2.0, Gi.clone())
    torch.cuda.synchronize()
    assert st0 == 0 and st1 == 0
    assert np.all(np.abs(G1.cpu().numpy() - G0.cpu().numpy()) <= 0.5 * _bound(V, W) + 1e-15 * 4)


def test_gram_noalloc_refuses_oversized_products(ctx):
    V = torch.zeros(1000, 1100, dtype=torch.float64, device="cuda:0")
    print("gram_noalloc 1100 x 1100 (past the scratch)", flush=True)
    st, _ = _gram(ctx, ctx.lib.ofrr_gram_noalloc, V, V)
    assert st == 2005
    assert b"256 tiles" in ctx.lib.ofrr_ctx_last_error(ctx.h)
