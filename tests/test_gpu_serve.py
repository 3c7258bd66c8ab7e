"""Serving path (ofrr_server_*): continuous batching of eigen requests on one GPU.

* each result is the same alone and in the mix: the requests are solved
  together on 4 slots, then one at a time on a 1-slot server, and the Ritz
  values and residuals must agree exactly (tolerance 0: every reduction of a
  solve has a fixed order and solves share no device memory);
* the solves overlap (more than one ran at once) and every pair converged;
* admission keeps the running footprint within the byte budget;
* eigenvalues match an FP64 dense eigensolver on the oracle's operator
  (variable coefficients) and the closed form (Laplacian).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (dim, N, nwant, coef_seed): 2-D 5-point and 3-D 7-point diffusion operators,
# variable face coefficients (seed > 0) and the Laplacian (seed 0)
MIX = [
    (2, 100, 8, 11), (3, 24, 16, 12), (2, 160, 16, 13), (3, 20, 8, 0),
    (2, 120, 32, 14), (3, 30, 8, 15), (2, 64, 16, 0), (3, 16, 32, 16),
]


def _requests(api):
    return [api.request(d, N, w, coef_seed=s, start_seed=100 + i, basis_dtype=torch.float16,
                        tol=1e-8) for i, (d, N, w, s) in enumerate(MIX)]


def _run(api, reqs, slots, budget=0, graphs=False, rr_group_us=0):
    with api.Server(0, slots=slots, budget=budget, graphs=graphs, rr_group_us=rr_group_us) as srv:
        if slots == 1:
            out = []
            for r in reqs:  # strictly one after the other
                out.append(srv.wait(srv.submit(r)))
        else:
            ids = [srv.submit(r) for r in reqs]
            out = [srv.wait(i) for i in ids]
        stats = srv.stats()
    return out, stats


@pytest.fixture(scope="module")
def api():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mixedprecisionofrr_amd import api as A
    return A


def test_serve_same_alone_and_in_mix(api):
    reqs = _requests(api)
    print("serve: mix of", len(reqs), "requests on 4 slots", flush=True)
    mix, st_mix = _run(api, reqs, 4)
    print("serve: the same requests one at a time", flush=True)
    alone, _ = _run(api, _requests(api), 1)
    for (d, N, w, s), a, b in zip(MIX, mix, alone):
        assert a["status"] == 0 and b["status"] == 0, (d, N, w, s, a["status"], b["status"])
        assert a["converged"] == w
        assert np.max(a["res"]) <= 1e-8
        assert a["iterations"] == b["iterations"]
        assert np.array_equal(a["theta"], b["theta"]), (d, N, w, s)
        assert np.array_equal(a["res"], b["res"]), (d, N, w, s)
    assert st_mix["completed"] == len(reqs)
    assert st_mix["running_peak"] > 1  # the solves overlapped
    assert st_mix["bytes_running"] == 0 and st_mix["running"] == 0


def test_serve_admission_respects_budget(api):
    reqs = _requests(api)
    fps = [api.request_footprint(r)[2] for r in reqs]
    slots = 4
    # what the slots and the grouped Rayleigh-Ritz hold for the server's life
    with api.Server(0, slots=slots) as probe:
        reserve = probe.stats()["slot_bytes"]
    assert reserve >= slots * api.server_slot_bytes(0)
    budget = reserve + int(max(fps) * 1.2)  # no two of the largest fit together
    out, st = _run(api, reqs, slots, budget=budget)
    assert st["budget"] == budget and st["slot_bytes"] == reserve
    assert st["ws_regrowths"] == 0  # the slots' reservation covered every solve
    assert st["bytes_peak"] <= budget - reserve
    assert all(o["status"] == 0 for o in out)
    # requests are admitted in arrival order
    admits = [o["t_admit"] for o in out]
    assert admits == sorted(admits)
    # a request larger than the budget is refused at submission
    with api.Server(0, slots=2, budget=min(fps) - 1) as srv:
        with pytest.raises(api.OfrrError) as e:
            srv.submit(reqs[0])
        assert e.value.status == 1002


def test_serve_eigenvalues_match_dense_fp64(api):
    reqs = [api.request(2, 40, 16, coef_seed=7, start_seed=3, tol=1e-8),
            api.request(3, 12, 8, coef_seed=0, start_seed=4, tol=1e-8)]
    out, _ = _run(api, reqs, 2)
    A = O.stencil(2, 40, seed=7)
    dense = np.zeros((A.n, A.n))
    for i in range(A.n):
        for p in range(A.row_ptr[i], A.row_ptr[i + 1]):
            dense[i, A.col_idx[p]] = A.vals[p]
    ref = np.linalg.eigvalsh(dense)[::-1][:16]
    assert out[0]["status"] == 0
    assert np.max(np.abs(out[0]["theta"] - ref) / ref) < 1e-10
    lap = O.laplacian_eigs(3, 12)[:8]
    assert out[1]["status"] == 0
    assert np.max(np.abs(out[1]["theta"] - lap) / lap) < 1e-10


def _fp8_basis300(api):
    """The large FP8 solve's configuration on its scaled twin: 3-D
    variable-coefficient diffusion (coefficient seed 5) on a 32^3 grid, the 100
    largest pairs, basis dimension 300 = a 128-column Ritz block (100 wanted +
    28 guard columns), the 128-column FP8 E4M3 correction block and the
    trailing 44 columns of the FP8 mid-filter block (per-column power-of-two
    scales, FP32 accumulation), FP64 Gram / truncated Rayleigh-Ritz /
    residuals, and the precision ladder finishing the growth blocks in FP16
    below 1e-4."""
    return api.request(3, 32, 100, coef_seed=5, start_seed=7, basis_dtype=torch.float8_e4m3fn,
                       tol=1e-8, k=128, nd=2, mid_cols=44, refine_dtype=torch.float16, refine_tol=1e-4)


def test_serve_fp8_basis300_beside_a_mix(api):
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "fp8_twin.json")))
    print("serve: FP8 basis-300 solve (32^3 twin) beside the mix on 4 slots", flush=True)
    mix, st = _run(api, _requests(api)[:4] + [_fp8_basis300(api)] + _requests(api)[4:], 4)
    print("serve: the FP8 basis-300 solve alone", flush=True)
    alone, _ = _run(api, [_fp8_basis300(api)], 1)
    b, a = mix[4], alone[0]
    print(f"  status {b['status']} k {b['k']} nd {b['nd']} iterations {b['iterations']} "
          f"max residual {np.max(b['res']):.2e}", flush=True)
    assert b["status"] == 0 and (b["k"], b["nd"]) == (128, 2)
    assert b["converged"] == 100 and np.max(b["res"]) <= 1e-8
    ref = np.array(gold["top_eigenvalues"][:100])
    assert np.max(np.abs(b["theta"] - ref) / ref[0]) < 1e-10
    # the same result alone and in the mix, bit for bit
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["res"], b["res"])
    assert st["running_peak"] > 1 and st["completed"] == 9
    for o in mix:
        assert o["status"] == 0


def test_serve_graph_sweeps_match_plain_launches(api):
    """The slots' filter sweeps run as captured graphs kept per slot and shape
    (re-recorded into the same executable graphs for the next request of a
    shape); the same stream with plain launches gives bitwise equal results."""
    reqs = _requests(api)[:6]
    print("serve: graphs on", flush=True)
    on, _ = _run(api, reqs, 2, graphs=True)
    print("serve: graphs off", flush=True)
    off, _ = _run(api, _requests(api)[:6], 2, graphs=False)
    for a, b in zip(on, off):
        assert a["status"] == 0 and b["status"] == 0
        assert a["iterations"] == b["iterations"]
        assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["res"], b["res"])


def test_serve_grouped_rr_matches_separate_and_direct(api):
    """The grouped Rayleigh-Ritz (jobs of several running solves in one launch
    per stage) gives each solve bitwise the result of its own Rayleigh-Ritz:
    the same stream with grouping off, and the direct ofrr_filtered_eig call on
    the request's operator and start block, agree exactly.  A wide window makes
    the solves' jobs meet, so groups of more than one job form."""
    reqs = [api.request(2, 60 + 10 * i, 8 + 4 * (i % 3), coef_seed=20 + i, start_seed=50 + i,
                        basis_dtype=torch.float16, tol=1e-8) for i in range(8)]
    print("serve: grouped Rayleigh-Ritz (window 2 ms)", flush=True)
    grouped, st = _run(api, reqs, 4, rr_group_us=2000)
    print("serve: separate Rayleigh-Ritz", flush=True)
    sep, st0 = _run(api, reqs, 4, rr_group_us=-1)
    assert st["rr_group_max"] > 1 and st["rr_group_jobs"] > st["rr_groups"]
    assert st0["rr_groups"] == 0
    for a, b in zip(grouped, sep):
        assert a["status"] == 0 and b["status"] == 0
        assert a["iterations"] == b["iterations"]
        assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["res"], b["res"])
    # the direct call with the server's options (degree 120, tau 1e-12, soft
    # locking, FP32 operator values for variable coefficients)
    from mixedprecisionofrr_amd import api as A
    ctx = A.Context(0)
    for r, g in zip(reqs[:3], grouped[:3]):
        k = g["k"]
        M = A.CSRMatrix.stencil(ctx, r.dim, r.N, seed=r.coef_seed).with_low_precision(torch.float32)
        X0 = torch.empty(M.n, k, dtype=torch.float64, device="cuda:0")
        A.fill_randn(ctx, X0, seed=r.start_seed)
        out = A.filtered_eig(ctx, M, X0, r.nwant, nd=g["nd"], degree=120, maxiter=200,
                             basis_dtype=torch.float16, tol=1e-8, tau=1e-12)
        assert out[0] == 0 and out[4]["iterations"] == g["iterations"]
        assert np.array_equal(out[1][:r.nwant], g["theta"]) and np.array_equal(out[3][:r.nwant], g["res"])
