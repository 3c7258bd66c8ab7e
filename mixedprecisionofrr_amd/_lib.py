"""ctypes declarations of the C-ABI in include/ofrr.h (libofrr.so, built in-tree).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libofrr.so")

F64, F32, F16, BF16, FP8 = 0, 1, 2, 3, 4

# every extern "C" entry point declared in include/ofrr.h
EXPORTS = [
    "ofrr_ctx_create", "ofrr_ctx_destroy", "ofrr_ctx_sync", "ofrr_ctx_set_stream",
    "ofrr_ctx_use_own_stream",
    "ofrr_ctx_stream", "ofrr_ctx_last_error", "ofrr_status_string", "ofrr_version",
    "ofrr_fill_randn", "ofrr_fill_randu", "ofrr_stencil_nnz", "ofrr_stencil_csr",
    "ofrr_spmm_csr", "ofrr_iamax", "ofrr_matrix_iamax", "ofrr_scalecols", "ofrr_check_special",
    "ofrr_hessenberg", "ofrr_gram", "ofrr_basis_update", "ofrr_syev_batched", "ofrr_sygv",
    "ofrr_rr_truncated", "ofrr_rr_split", "ofrr_rr_generalized", "ofrr_residuals", "ofrr_arnoldi_hess_eig",
    "ofrr_subspace_iter_eig", "ofrr_filtered_eig", "ofrr_dense_gemm", "ofrr_convert",
    "ofrr_filtered_eig_dense", "ofrr_svd_filtered", "ofrr_filtered_eig_state_bytes",
    "ofrr_request_footprint", "ofrr_server_create", "ofrr_server_submit", "ofrr_server_poll",
    "ofrr_server_wait", "ofrr_server_stats_get", "ofrr_server_destroy",
    "ofrr_csr_file_info", "ofrr_csr_file_read", "ofrr_singular_values_file", "ofrr_io_last_error",
    "ofrr_svd_rr", "ofrr_subspace_iter_svd", "ofrr_fp8_check", "ofrr_gs_qr", "ofrr_gs_qr_step",
    "ofrr_spmm_csr_rect", "ofrr_csr_transpose", "ofrr_server_slot_bytes",
    "ofrr_hessenberg_v3", "ofrr_hessenberg_step", "ofrr_gemm", "ofrr_axpb",
    "ofrr_arnoldi_hess_eig_mp", "ofrr_gram_noalloc",
]

SERVE_MAX_WANT = 128
PENDING = 1


class FilterOpts(C.Structure):
    _fields_ = [("k", C.c_int), ("nwant", C.c_int), ("nd", C.c_int), ("degree", C.c_int),
                ("maxiter", C.c_int), ("basis_dtype", C.c_int), ("tol", C.c_double),
                ("tau", C.c_double), ("lower", C.c_double), ("normA", C.c_double),
                ("refine", C.c_int), ("refine_dtype", C.c_int), ("refine_tol", C.c_double),
                ("lock", C.c_int), ("mid_cols", C.c_int), ("graph", C.c_int),
                ("rr_whole", C.c_int), ("wide_dtype", C.c_int), ("wide_tol", C.c_double)]


class FilterStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("spmm_calls", C.c_int),
                ("rank_min", C.c_int), ("filter_steps", C.c_int), ("resid_passes", C.c_int),
                ("t_total_ms", C.c_double), ("t_filter_ms", C.c_double),
                ("t_zspmm_ms", C.c_double), ("t_gram_ms", C.c_double), ("t_rr_ms", C.c_double),
                ("t_update_ms", C.c_double), ("t_resid_ms", C.c_double),
                ("filter_col_steps", C.c_int), ("filter_first_cols", C.c_int),
                ("graph_replays", C.c_int), ("graph_captures", C.c_int),
                ("rr_fallbacks", C.c_int), ("wide_sweeps", C.c_int), ("t_aux_ms", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int), ("budget", C.c_int64), ("graphs", C.c_int),
                ("rr_group_us", C.c_int)]


class EigRequest(C.Structure):
    _fields_ = [("dim", C.c_int), ("N", C.c_int64), ("coef_seed", C.c_uint64),
                ("start_seed", C.c_uint64), ("nwant", C.c_int), ("basis_dtype", C.c_int),
                ("tol", C.c_double), ("degree", C.c_int), ("maxiter", C.c_int),
                ("k", C.c_int), ("nd", C.c_int), ("refine", C.c_int), ("refine_dtype", C.c_int),
                ("refine_tol", C.c_double), ("mid_cols", C.c_int)]


class EigResult(C.Structure):
    _fields_ = [("status", C.c_int), ("k", C.c_int), ("nd", C.c_int), ("iterations", C.c_int),
                ("converged", C.c_int), ("filter_steps", C.c_int), ("n", C.c_int64),
                ("footprint", C.c_int64), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double), ("theta", C.c_double * SERVE_MAX_WANT),
                ("res", C.c_double * SERVE_MAX_WANT)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("running", C.c_int64),
                ("queued", C.c_int64), ("bytes_running", C.c_int64), ("bytes_peak", C.c_int64),
                ("running_peak", C.c_int), ("budget", C.c_int64), ("slot_bytes", C.c_int64),
                ("ws_regrowths", C.c_int64), ("rr_groups", C.c_int64), ("rr_group_jobs", C.c_int64),
                ("rr_group_max", C.c_int)]


_lib = None
vp = C.c_void_p
i64 = C.c_int64
i32 = C.c_int
dbl = C.c_double


def _decl(L):
    L.ofrr_ctx_create.argtypes = [i32, C.POINTER(vp)]
    L.ofrr_ctx_destroy.argtypes = [vp]
    L.ofrr_ctx_sync.argtypes = [vp]
    L.ofrr_ctx_set_stream.argtypes = [vp, vp]
    L.ofrr_ctx_use_own_stream.argtypes = [vp]
    L.ofrr_ctx_stream.argtypes = [vp]
    L.ofrr_ctx_stream.restype = vp
    L.ofrr_ctx_last_error.argtypes = [vp]
    L.ofrr_ctx_last_error.restype = C.c_char_p
    L.ofrr_status_string.argtypes = [i32]
    L.ofrr_status_string.restype = C.c_char_p
    L.ofrr_fill_randn.argtypes = [vp, i32, i64, vp, i64, C.c_uint64, C.c_uint32, C.c_uint64]
    L.ofrr_fill_randu.argtypes = [vp, i32, i64, vp, i64, C.c_uint64, C.c_uint32, C.c_uint64]
    L.ofrr_stencil_nnz.argtypes = [i32, i64, C.POINTER(i64)]
    L.ofrr_stencil_csr.argtypes = [vp, i32, i64, C.c_uint64, vp, vp, i32, vp]
    L.ofrr_spmm_csr.argtypes = [vp, i64, i64, vp, vp, vp, i32, i32, vp, i32, i64, i64, dbl, dbl,
                                vp, i32, i64, i64]
    L.ofrr_spmm_csr_rect.argtypes = [vp, i32, i64, i64, i64, vp, vp, vp, i32, i32, vp, i32, i64,
                                     i64, dbl, dbl, vp, i32, i64, i64]
    L.ofrr_csr_transpose.argtypes = [vp, i64, i64, i64, vp, vp, vp, i32, vp, vp, vp]
    L.ofrr_server_slot_bytes.argtypes = [i32, C.POINTER(i64)]
    L.ofrr_gemm.argtypes = [vp, i32, i32, i64, i64, i64, dbl, vp, i32, i64, vp, i32, i64, dbl, vp,
                            i32, i64, i32]
    L.ofrr_axpb.argtypes = [vp, i32, i64, i64, vp, i64, i64, dbl, dbl]
    L.ofrr_gram_noalloc.argtypes = [vp, i64, i32, i32, vp, i32, i64, i64, vp, i32, i64, i64, dbl,
                                    dbl, vp, i32]
    L.ofrr_arnoldi_hess_eig_mp.argtypes = [vp, i32, i64, i64, vp, vp, vp, i32, i32, i32, dbl, dbl, vp,
                                           C.POINTER(dbl), vp, i64, C.POINTER(i32)]
    L.ofrr_hessenberg_v3.argtypes = [vp, i32, i64, i32, vp, i64, i64, dbl, vp, i32, vp, vp]
    L.ofrr_hessenberg_step.argtypes = [vp, i32, i32, i64, i32, vp, i64, i64, i32, dbl, vp, i32,
                                       vp, vp, C.POINTER(i32)]
    L.ofrr_iamax.argtypes = [vp, i32, i64, vp, i64, vp, vp]
    L.ofrr_matrix_iamax.argtypes = [vp, i32, i64, i32, vp, i64, i64, vp]
    L.ofrr_scalecols.argtypes = [vp, i32, i64, i32, vp, i64, i64, vp]
    L.ofrr_check_special.argtypes = [vp, i32, i64, i32, vp, i64, i64]
    L.ofrr_hessenberg.argtypes = [vp, i32, i64, i32, vp, i64, i64, dbl, vp, i32, vp, vp]
    L.ofrr_gs_qr.argtypes = [vp, i32, i32, i64, i32, vp, i64, i64, dbl, dbl, dbl, vp, i32, vp]
    L.ofrr_gs_qr_step.argtypes = [vp, i32, i32, i64, i32, vp, i64, i64, i32, dbl, dbl, dbl, vp,
                                  i32, vp, C.POINTER(C.c_int32)]
    L.ofrr_gram.argtypes = [vp, i64, i32, i32, vp, i32, i64, i64, vp, i32, i64, i64, dbl, dbl,
                            vp, i32]
    L.ofrr_basis_update.argtypes = [vp, i64, i32, i32, C.POINTER(vp), C.POINTER(i32),
                                    C.POINTER(i32), C.POINTER(i64), C.POINTER(i64), C.POINTER(vp),
                                    C.POINTER(i32), vp, i32, i64, i64]
    L.ofrr_syev_batched.argtypes = [vp, i32, i32, vp, i32, i64, vp, i64, vp, i32, i64]
    L.ofrr_sygv.argtypes = [vp, i32, vp, i32, vp, i32, vp, vp, i32]
    L.ofrr_rr_truncated.argtypes = [vp, i32, vp, i32, vp, i32, dbl, i32, vp, vp, i32, vp]
    L.ofrr_rr_split.argtypes = [vp, i32, i32, vp, i32, vp, i32, dbl, i32, vp, vp, i32, vp]
    L.ofrr_rr_generalized.argtypes = [vp, i32, vp, i32, vp, i32, dbl, vp, vp, i32]
    L.ofrr_residuals.argtypes = [vp, i64, i64, vp, vp, vp, i32, vp, i64, i64,
                                 C.POINTER(dbl), dbl, C.POINTER(dbl)]
    L.ofrr_arnoldi_hess_eig.argtypes = [vp, i64, i64, vp, vp, vp, i32, i32, i32, dbl, dbl, vp,
                                        C.POINTER(dbl), vp, i64, C.POINTER(i32)]
    L.ofrr_subspace_iter_eig.argtypes = [vp, i64, i64, vp, vp, vp, i32, i32, i32, dbl, dbl, vp,
                                         i64, C.POINTER(dbl)]
    L.ofrr_filtered_eig.argtypes = [vp, i64, i64, vp, vp, vp, vp, i32, C.POINTER(FilterOpts), vp,
                                    C.POINTER(dbl), C.POINTER(dbl), C.POINTER(FilterStats)]
    L.ofrr_dense_gemm.argtypes = [vp, i64, i64, i32, vp, i32, i64, vp, i32, i64, dbl, dbl, vp, i32,
                                  i64]
    L.ofrr_convert.argtypes = [vp, i64, i64, vp, i32, i64, i64, vp, i32, i64, i64]
    L.ofrr_filtered_eig_dense.argtypes = [vp, i64, vp, i64, C.POINTER(FilterOpts), vp,
                                          C.POINTER(dbl), C.POINTER(dbl), C.POINTER(FilterStats)]
    L.ofrr_svd_filtered.argtypes = [vp, i64, i64, vp, i64, C.POINTER(FilterOpts), vp, vp,
                                    C.POINTER(dbl), C.POINTER(dbl), C.POINTER(FilterStats)]
    L.ofrr_filtered_eig_state_bytes.argtypes = [i64, C.POINTER(FilterOpts), C.POINTER(i64)]
    L.ofrr_request_footprint.argtypes = [C.POINTER(EigRequest), C.POINTER(i32), C.POINTER(i32),
                                         C.POINTER(i64)]
    L.ofrr_server_create.argtypes = [i32, C.POINTER(ServerOpts), C.POINTER(vp)]
    L.ofrr_server_submit.argtypes = [vp, C.POINTER(EigRequest), C.POINTER(i64)]
    L.ofrr_server_poll.argtypes = [vp, i64, C.POINTER(EigResult)]
    L.ofrr_server_wait.argtypes = [vp, i64, C.POINTER(EigResult)]
    L.ofrr_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
    L.ofrr_server_destroy.argtypes = [vp]
    L.ofrr_csr_file_info.argtypes = [C.c_char_p, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    L.ofrr_csr_file_read.argtypes = [C.c_char_p, vp, vp, vp]
    L.ofrr_singular_values_file.argtypes = [C.c_char_p, vp, i64, C.POINTER(i64)]
    L.ofrr_io_last_error.restype = C.c_char_p
    L.ofrr_fp8_check.argtypes = [vp, C.POINTER(i64)]
    L.ofrr_svd_rr.argtypes = [vp, i64, i64, vp, i64, i32, i32, vp, i64, vp, i64, i32, dbl,
                              C.POINTER(dbl)]
    L.ofrr_subspace_iter_svd.argtypes = [vp, i64, i64, vp, i64, i32, i32, i32, dbl, i32, dbl, vp,
                                         i64, vp, i64, vp, i64, C.POINTER(dbl), C.POINTER(i32)]


def load():
    """Load libofrr.so; raises if it is missing (never falls back)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (hipcc, gfx950); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        _decl(L)
        _lib = L
    return _lib
