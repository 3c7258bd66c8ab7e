"""Python host layer over the C-ABI (include/ofrr.h).

PyTorch only supplies device memory and the stream; every operation runs in
libofrr.so (hand-written HIP for gfx950).  Names and argument meanings follow
the reference's operator interface: the MATLAB plugins (csrmv, scalecols,
hessenberg, eig_rr, arnoldi_hess_eig, subspace_iter_eig) and the C++ free
functions (iamax, matrix_iamax, gemm, syev, sygv, check_special_values,
rr_eig/ofrr_eig).  Errors raise OfrrError carrying the MSVDStatus code.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import BF16, F16, F32, F64, FP8, FilterOpts, FilterStats

_DT = {torch.float64: F64, torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16,
       torch.float8_e4m3fn: FP8}
_TORCH = {v: k for k, v in _DT.items()}

STATUS_SUCCESS = 0
STATUS_MAX_ITER = 4001


class OfrrError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[{status}] {msg}")
        self.status = status


def dtype_code(t):
    if isinstance(t, int):
        return t
    if isinstance(t, torch.dtype):
        return _DT[t]
    return _DT[t.dtype]


def torch_dtype(code):
    return _TORCH[code]


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def colmajor(rows, cols, dtype=torch.float64, device="cuda"):
    """Empty rows x cols tensor in column-major storage (ld = rows)."""
    return torch.empty(cols, rows, dtype=dtype, device=device).t()


class Context:
    """Owns an ofrr_ctx bound to torch's current stream on `device`."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        self.device = device
        h = C.c_void_p()
        st = self.lib.ofrr_ctx_create(device, C.byref(h))
        if st != 0:
            raise OfrrError(st, "ofrr_ctx_create failed")
        self.h = h
        self.bind_stream()

    def bind_stream(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.lib.ofrr_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream))

    def check(self, st, what=""):
        if st != 0:
            msg = self.lib.ofrr_ctx_last_error(self.h).decode() or \
                self.lib.ofrr_status_string(st).decode()
            raise OfrrError(st, f"{what}: {msg}")

    def sync(self):
        self.check(self.lib.ofrr_ctx_sync(self.h), "sync")

    def close(self):
        if self.h:
            self.lib.ofrr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CSRMatrix:
    """Device CSR operator: int64 row_ptr, int32 col_idx, FP64 values, and an
    optional low-precision copy of the values for the growth SpMMs.  n rows and
    ncols columns (default n); only spmm / spmm_rect / transpose take a
    rectangular matrix, every other entry point needs a square one."""

    def __init__(self, n, row_ptr, col_idx, vals, ncols=None):
        self.n = int(n)
        self.ncols = self.n if ncols is None else int(ncols)
        self.row_ptr = row_ptr
        self.col_idx = col_idx
        self.vals = vals
        self.vals_lo = None

    @property
    def shape(self):
        return (self.n, self.ncols)

    def require_square(self, what):
        if self.n != self.ncols:
            raise ValueError(f"{what}: needs a square operator, got {self.n} x {self.ncols}")

    @property
    def nnz(self):
        return int(self.col_idx.numel())

    @staticmethod
    def stencil(ctx, dim, N, seed=0):
        nnz = C.c_int64()
        ctx.check(ctx.lib.ofrr_stencil_nnz(dim, N, C.byref(nnz)), "stencil_nnz")
        n = N ** dim
        dev = f"cuda:{ctx.device}"
        rp = torch.empty(n + 1, dtype=torch.int64, device=dev)
        ci = torch.empty(nnz.value, dtype=torch.int32, device=dev)
        va = torch.empty(nnz.value, dtype=torch.float64, device=dev)
        ctx.check(ctx.lib.ofrr_stencil_csr(ctx.h, dim, N, seed, _p(rp), _p(ci), F64, _p(va)),
                  "stencil_csr")
        return CSRMatrix(n, rp, ci, va)

    @staticmethod
    def from_files(ctx, prefix, scale=1.0, symmetrize=False, shift=0.0):
        """The binary CSR files <prefix>_{row_ptr,col_ind,values}.bin
        (MatrixReader::read_csr_matrix, CPP/io/matrix_reader.cpp:8-45) on the
        device, with the experiment's preprocessing applied on the host
        (MATLAB/+tests/+test4/test.m:31-49): values / scale, (K + K^T) / 2 when
        symmetrize, + shift I."""
        rp, ci, va, rows, cols = read_csr(prefix)
        if rows != cols and (symmetrize or shift):
            raise ValueError("from_files: symmetrize / shift need a square matrix")
        va = va / scale
        if symmetrize or shift:
            rp, ci, va = _sym_shift(rp, ci, va, rows, symmetrize, shift)
        return CSRMatrix.from_host(ctx, rp, ci, va, ncols=cols)

    @staticmethod
    def from_host(ctx, row_ptr, col_idx, vals, ncols=None):
        """Upload host CSR arrays; ncols defaults to the row count.  The
        structure is checked on the host (monotone row pointers spanning the
        entries, column indices in [0, ncols)), so a bad matrix is a ValueError
        here rather than an out-of-bounds gather on the device."""
        rp_h = np.asarray(row_ptr, np.int64)
        ci_h = np.asarray(col_idx, np.int32)
        va_h = np.asarray(vals, np.float64)
        n = len(rp_h) - 1
        nc = n if ncols is None else int(ncols)
        _check_csr(rp_h, ci_h, va_h, n, nc)
        dev = f"cuda:{ctx.device}"
        rp = torch.as_tensor(rp_h, device=dev)
        ci = torch.as_tensor(ci_h, device=dev)
        va = torch.as_tensor(va_h, device=dev)
        return CSRMatrix(n, rp, ci, va, ncols=nc)

    def with_low_precision(self, dtype=torch.float16):
        self.vals_lo = self.vals.to(dtype)
        return self


def _check_csr(rp, ci, va, n, ncols):
    if n < 0 or ncols < 0:
        raise ValueError("CSR: negative dimension")
    if len(ci) != len(va):
        raise ValueError("CSR: column indices and values differ in length")
    if rp[0] != 0 or rp[-1] != len(ci) or np.any(np.diff(rp) < 0):
        raise ValueError("CSR: row pointers must start at 0, never decrease and end at nnz")
    if len(ci) and (ci.min() < 0 or ci.max() >= ncols):
        raise ValueError(f"CSR: column indices must lie in [0, {ncols})")


def _check_rows(X, rows, what, name="X"):
    if X.shape[0] != rows:
        raise ValueError(f"{what}: {name} has {X.shape[0]} rows, the operator needs {rows}")


# ---------------------------------------------------------------------------
def spmm(ctx, A, X, out_dtype=None, alpha=1.0, beta=0.0, Y=None, vals=None):
    """Y = alpha A X + beta Y (csrmv semantics, csrmv.m:48-59) for a square or
    rectangular A: X has A.ncols rows, Y A.n rows."""
    return spmm_rect(ctx, A, X, trans=False, out_dtype=out_dtype, alpha=alpha, beta=beta, Y=Y,
                     vals=vals)


def spmm_rect(ctx, A, X, trans=False, out_dtype=None, alpha=1.0, beta=0.0, Y=None, vals=None):
    """csrmv(Acsr, trans, X) of csrmv.m:27-76: Y = alpha op(A) X + beta Y with
    op(A) = A (X: A.ncols rows, Y: A.n rows) or A^T (trans; X: A.n rows, Y:
    A.ncols rows; the 'T' branch, :61-71, bitwise in FP64)."""
    if X.dim() == 1:
        X = X.view(-1, 1)
    xr, yr = (A.n, A.ncols) if trans else (A.ncols, A.n)
    _check_rows(X, xr, "spmm")
    k = X.shape[1]
    if Y is None:
        Y = torch.zeros(yr, k, dtype=out_dtype or X.dtype, device=X.device)
    if tuple(Y.shape) != (yr, k):
        raise ValueError(f"spmm: Y must be {yr} x {k}, got {tuple(Y.shape)}")
    va = A.vals if vals is None else vals
    if not trans and A.n == A.ncols:
        ctx.check(ctx.lib.ofrr_spmm_csr(ctx.h, A.n, A.nnz, _p(A.row_ptr), _p(A.col_idx), _p(va),
                                        dtype_code(va), k, _p(X), dtype_code(X), X.stride(0),
                                        X.stride(1), alpha, beta, _p(Y), dtype_code(Y),
                                        Y.stride(0), Y.stride(1)), "spmm_csr")
    else:
        ctx.check(ctx.lib.ofrr_spmm_csr_rect(ctx.h, 1 if trans else 0, A.n, A.ncols, A.nnz,
                                             _p(A.row_ptr), _p(A.col_idx), _p(va), dtype_code(va),
                                             k, _p(X), dtype_code(X), X.stride(0), X.stride(1),
                                             alpha, beta, _p(Y), dtype_code(Y), Y.stride(0),
                                             Y.stride(1)), "spmm_csr_rect")
    return Y


def transpose(ctx, A):
    """A^T as a CSRMatrix (A.ncols x A.n), each row's entries in increasing
    row of A (ofrr_csr_transpose)."""
    dev = A.row_ptr.device
    rp = torch.empty(A.ncols + 1, dtype=torch.int64, device=dev)
    ci = torch.empty(max(1, A.nnz), dtype=torch.int32, device=dev)
    va = torch.empty(max(1, A.nnz), dtype=A.vals.dtype, device=dev)
    ctx.check(ctx.lib.ofrr_csr_transpose(ctx.h, A.n, A.ncols, A.nnz, _p(A.row_ptr),
                                         _p(A.col_idx), _p(A.vals), dtype_code(A.vals), _p(rp),
                                         _p(ci), _p(va)), "csr_transpose")
    return CSRMatrix(A.ncols, rp, ci[:A.nnz], va[:A.nnz], ncols=A.n)


def fill_randn(ctx, t, seed, stream=0, offset=0):
    ctx.check(ctx.lib.ofrr_fill_randn(ctx.h, dtype_code(t), t.numel(), _p(t), 1, seed, stream,
                                      offset), "fill_randn")
    return t


def fill_randu(ctx, t, seed, stream=0, offset=0):
    ctx.check(ctx.lib.ofrr_fill_randu(ctx.h, dtype_code(t), t.numel(), _p(t), 1, seed, stream,
                                      offset), "fill_randu")
    return t


def iamax(ctx, x):
    """(1-based index, signed value) of the first max |x_i| (mvops.hpp:98-99)."""
    idx = torch.zeros(1, dtype=torch.int32, device=x.device)
    val = torch.zeros(1, dtype=torch.float64, device=x.device)
    ctx.check(ctx.lib.ofrr_iamax(ctx.h, dtype_code(x), x.numel(), _p(x), x.stride(0) if x.dim() else 1,
                                 _p(idx), _p(val)), "iamax")
    return int(idx.item()), float(val.item())


def matrix_iamax(ctx, A):
    """0-based per-column argmax |A(:,j)| (mvops.hpp:113-114)."""
    m, n = A.shape
    idx = torch.zeros(n, dtype=torch.int32, device=A.device)
    ctx.check(ctx.lib.ofrr_matrix_iamax(ctx.h, dtype_code(A), m, n, _p(A), A.stride(0),
                                        A.stride(1), _p(idx)), "matrix_iamax")
    return idx


def scalecols(ctx, A):
    """In place: each column divided by its signed max-abs entry (scalecols.m:21-34)."""
    m, n = A.shape
    piv = torch.zeros(n, dtype=torch.float64, device=A.device)
    ctx.check(ctx.lib.ofrr_scalecols(ctx.h, dtype_code(A), m, n, _p(A), A.stride(0), A.stride(1),
                                     _p(piv)), "scalecols")
    return piv


def check_special(ctx, A):
    """0, 1005 (NaN) or 1006 (Inf) -- check_special_values (mvops.cu:1388-1599)."""
    if A.dim() == 1:
        A = A.view(-1, 1)
    m, n = A.shape
    return ctx.lib.ofrr_check_special(ctx.h, dtype_code(A), m, n, _p(A), A.stride(0), A.stride(1))


def hessenberg(ctx, A, select_tol=None, variant=0):
    """[Q, R, perm, skip] of the Hessenberg process on a copy of A: variant 0
    (hessenberg.m:55-74, C++ hessenberg / hessenberg_v2) leaves dropped columns
    out of later eliminations, variant 1 (C++ hessenberg_v3) eliminates with
    them."""
    Q = A.clone() if A.stride(0) == 1 else A.t().contiguous().t().clone()
    m, n = Q.shape
    if select_tol is None:
        select_tol = _EPS[Q.dtype]
    R = colmajor(n, n, torch.float64, Q.device)
    perm = torch.zeros(n, dtype=torch.int32, device=Q.device)
    skip = torch.zeros(n, dtype=torch.int32, device=Q.device)
    fn = ctx.lib.ofrr_hessenberg_v3 if variant == 1 else ctx.lib.ofrr_hessenberg
    ctx.check(fn(ctx.h, dtype_code(Q), m, n, _p(Q), Q.stride(0), Q.stride(1), select_tol, _p(R), n,
                 _p(perm), _p(skip)), "hessenberg")
    return Q, R, perm, skip


def hessenberg_by_steps(ctx, A, select_tol=None, variant=1):
    """The same factorization built one column at a time with
    ofrr_hessenberg_step (hessenberg_v3_step): [Q, R, prow1 (1-based argmax
    rows, dropped columns included), skip]."""
    Q = A.clone() if A.stride(0) == 1 else A.t().contiguous().t().clone()
    m, n = Q.shape
    if select_tol is None:
        select_tol = _EPS[Q.dtype]
    R = torch.zeros(n, n, dtype=torch.float64, device=Q.device).t()
    prow = torch.zeros(n, dtype=torch.int32, device=Q.device)
    skip = torch.zeros(n, dtype=torch.int32, device=Q.device)
    dropped = C.c_int32()
    for i in range(n):
        ctx.check(ctx.lib.ofrr_hessenberg_step(ctx.h, variant, dtype_code(Q), m, n, _p(Q), Q.stride(0),
                                               Q.stride(1), i, select_tol, _p(R), n, _p(prow),
                                               _p(skip), C.byref(dropped)), "hessenberg_step")
    return Q, R, prow, skip


QR_VARIANTS = {"mgs": 0, "mgs_v2": 1, "cgs": 2, "cgs2": 3}
_EPS = {torch.float64: 2.0 ** -52, torch.float32: 2.0 ** -23, torch.float16: 2.0 ** -10,
        torch.bfloat16: 2.0 ** -7}


def gs_qr(ctx, A, variant="mgs", select_tol=None, orth_tol=None, reorth_tol=None):
    """[Q, R, skip] of a Gram-Schmidt QR of A (qr.hpp:9-141): 'mgs' (mgs, reorthogonalized
    with the default reorth_tol = 1/sqrt(2); <= 0 turns it off), 'mgs_v2', 'cgs'
    (reorth_tol > 0: iterated, cgs.m:73-89; default off as in the C++ cgs) or 'cgs2'.
    Tolerances default to the output format's epsilon (type_utils.hpp:125-140).  Q keeps
    A's storage order (column- or row-major)."""
    Q = A.clone()
    m, n = Q.shape
    eps = _EPS[Q.dtype]
    select_tol = eps if select_tol is None else select_tol
    orth_tol = eps if orth_tol is None else orth_tol
    if reorth_tol is None:
        reorth_tol = 2.0 ** -0.5 if variant == "mgs" else -1.0
    R = colmajor(n, n, torch.float64, Q.device)
    skip = torch.zeros(n, dtype=torch.int32, device=Q.device)
    ctx.check(ctx.lib.ofrr_gs_qr(ctx.h, QR_VARIANTS[variant], dtype_code(Q), m, n, _p(Q),
                                 Q.stride(0), Q.stride(1), select_tol, orth_tol, reorth_tol, _p(R),
                                 n, _p(skip)), "gs_qr")
    return Q, R, skip


def gs_qr_step(ctx, Q, R, skip, i, variant="mgs", select_tol=None, orth_tol=None,
               reorth_tol=None):
    """Column step i in place on (Q, R, skip): mgs_step / cgs_step / cgs2_step
    (qr.hpp:52-141).  Columns 0..i-1 must be final.  Returns the drop flag of column i."""
    m, n = Q.shape
    eps = _EPS[Q.dtype]
    select_tol = eps if select_tol is None else select_tol
    orth_tol = eps if orth_tol is None else orth_tol
    if reorth_tol is None:
        reorth_tol = 2.0 ** -0.5 if variant == "mgs" else -1.0
    out = C.c_int32(0)
    ctx.check(ctx.lib.ofrr_gs_qr_step(ctx.h, QR_VARIANTS[variant], dtype_code(Q), m, n, _p(Q),
                                      Q.stride(0), Q.stride(1), i, select_tol, orth_tol,
                                      reorth_tol, _p(R), R.stride(1), _p(skip),
                                      C.byref(out)), "gs_qr_step")
    return int(out.value)


def gram(ctx, V, W, alpha=1.0, beta=0.0, G=None):
    """G = alpha V^T W + beta G, FP64 accumulation (eig_rr.m:49,51)."""
    n, k1 = V.shape
    k2 = W.shape[1]
    if G is None:
        G = colmajor(k1, k2, torch.float64, V.device)
        G.zero_()
    ctx.check(ctx.lib.ofrr_gram(ctx.h, n, k1, k2, _p(V), dtype_code(V), V.stride(0), V.stride(1),
                                _p(W), dtype_code(W), W.stride(0), W.stride(1), alpha, beta,
                                _p(G), G.stride(1)), "gram")
    return G


def basis_update(ctx, blocks, k, out_dtype=torch.float64, Y=None):
    """Y = sum_b V_b C_b for [(V_b, C_b)] (C_b column-major FP64)."""
    n = blocks[0][0].shape[0]
    nb = len(blocks)
    if Y is None:
        Y = torch.empty(n, k, dtype=out_dtype, device=blocks[0][0].device)
    Vp = (C.c_void_p * nb)(*[V.data_ptr() for V, _ in blocks])
    dt = (C.c_int * nb)(*[dtype_code(V) for V, _ in blocks])
    kb = (C.c_int * nb)(*[V.shape[1] for V, _ in blocks])
    rs = (C.c_int64 * nb)(*[V.stride(0) for V, _ in blocks])
    cs = (C.c_int64 * nb)(*[V.stride(1) for V, _ in blocks])
    Cp = (C.c_void_p * nb)(*[Cb.data_ptr() for _, Cb in blocks])
    ld = (C.c_int * nb)(*[Cb.stride(1) for _, Cb in blocks])
    for _, Cb in blocks:
        assert Cb.stride(0) == 1 and Cb.dtype == torch.float64
    ctx.check(ctx.lib.ofrr_basis_update(ctx.h, n, k, nb, Vp, dt, kb, rs, cs, Cp, ld, _p(Y),
                                        dtype_code(Y), Y.stride(0), Y.stride(1)), "basis_update")
    return Y


def syev_batched(ctx, A):
    """Eigenvalues ascending and vectors of each symmetric matrix in A (batch, n, n),
    read from the upper triangle in column-major storage (syev, mvops.hpp:184-197)."""
    if A.dim() == 2:
        A = A.unsqueeze(0)
    b, n, _ = A.shape
    Acm = A.transpose(1, 2).contiguous()  # element (i, j) at [j, i]: column-major per matrix
    w = torch.empty(b, n, dtype=torch.float64, device=A.device)
    V = torch.empty(b, n, n, dtype=torch.float64, device=A.device)
    ctx.check(ctx.lib.ofrr_syev_batched(ctx.h, b, n, _p(Acm), n, n * n, _p(w), n, _p(V), n, n * n),
              "syev_batched")
    return w, V.transpose(1, 2)


def sygv(ctx, A, B):
    """A x = lambda B x (itype 1), ascending, B-normalised vectors (sygv, mvops.hpp:184-197)."""
    n = A.shape[0]
    Acm = A.t().contiguous()
    Bcm = B.t().contiguous()
    w = torch.empty(n, dtype=torch.float64, device=A.device)
    V = colmajor(n, n, torch.float64, A.device)
    ctx.check(ctx.lib.ofrr_sygv(ctx.h, n, _p(Acm), n, _p(Bcm), n, _p(w), _p(V), n), "sygv")
    return w, V


def rr_truncated(ctx, G, H, tau, nkeep):
    """Truncated OFRR Rayleigh-Ritz (ofrr_eig, rr_eig.hpp:21-35): (theta desc, Y, rank)."""
    m = G.shape[0]
    Gc = G.t().contiguous().t() if G.stride(0) != 1 else G
    Hc = H.t().contiguous().t() if H.stride(0) != 1 else H
    th = torch.empty(nkeep, dtype=torch.float64, device=G.device)
    Y = colmajor(m, nkeep, torch.float64, G.device)
    rk = torch.zeros(1, dtype=torch.int32, device=G.device)
    ctx.check(ctx.lib.ofrr_rr_truncated(ctx.h, m, _p(Gc), Gc.stride(1), _p(Hc), Hc.stride(1), tau,
                                        nkeep, _p(th), _p(Y), m, _p(rk)), "rr_truncated")
    return th, Y, int(rk.item())


def rr_split(ctx, G, H, kx, tau, nkeep):
    """Split truncated Rayleigh-Ritz over a basis whose first kx columns are a
    previous Ritz block: (theta desc, Y, rank); rank -1 when G[:kx, :kx] is not
    positive definite (theta and Y are then not written)."""
    m = G.shape[0]
    Gc = G.t().contiguous().t() if G.stride(0) != 1 else G
    Hc = H.t().contiguous().t() if H.stride(0) != 1 else H
    th = torch.full((nkeep,), float("nan"), dtype=torch.float64, device=G.device)
    Y = colmajor(m, nkeep, torch.float64, G.device)
    rk = torch.zeros(1, dtype=torch.int32, device=G.device)
    ctx.check(ctx.lib.ofrr_rr_split(ctx.h, m, kx, _p(Gc), Gc.stride(1), _p(Hc), Hc.stride(1), tau,
                                    nkeep, _p(th), _p(Y), m, _p(rk)), "rr_split")
    return th, Y, int(rk.item())


def rr_generalized(ctx, G, H, gtol=0.0):
    """eig(H, G + gtol I) sorted descending (eig_rr.m:52)."""
    m = G.shape[0]
    Gc = G.t().contiguous().t() if G.stride(0) != 1 else G
    Hc = H.t().contiguous().t() if H.stride(0) != 1 else H
    th = torch.empty(m, dtype=torch.float64, device=G.device)
    Y = colmajor(m, m, torch.float64, G.device)
    ctx.check(ctx.lib.ofrr_rr_generalized(ctx.h, m, _p(Gc), Gc.stride(1), _p(Hc), Hc.stride(1),
                                          gtol, _p(th), _p(Y), m), "rr_generalized")
    return th, Y


def residuals(ctx, A, X, theta, normA):
    """||A x_j - theta_j x_j|| / (normA ||x_j||) in FP64 (test4/test.m:340-350)."""
    A.require_square("residuals")
    _check_rows(X, A.n, "residuals")
    k = X.shape[1]
    if len(theta) < k:
        raise ValueError("residuals: one Ritz value per column of X")
    th = (C.c_double * k)(*[float(t) for t in theta])
    res = (C.c_double * k)()
    ctx.check(ctx.lib.ofrr_residuals(ctx.h, A.n, A.nnz, _p(A.row_ptr), _p(A.col_idx), _p(A.vals),
                                     k, _p(X), X.stride(0), X.stride(1), th, normA, res),
              "residuals")
    return np.array(res[:])


def arnoldi_hess_eig(ctx, A, k, kdim, maxiter, v0, select_tol=2.0 ** -52, gtol=0.0,
                     basis_dtype=torch.float64):
    """arnoldi_hess_eig.m on the GPU: (theta desc, X col-major FP64, built).
    basis_dtype is precision_output (the stored Krylov basis: FP64, FP32, FP16
    or BF16; FP32 arithmetic for the 32- and 16-bit formats); the
    Rayleigh-Ritz runs in FP64."""
    A.require_square("arnoldi_hess_eig")
    if v0.numel() != A.n:
        raise ValueError(f"arnoldi_hess_eig: the start vector needs {A.n} entries")
    th = (C.c_double * k)()
    X = colmajor(A.n, k, torch.float64, v0.device)
    built = C.c_int()
    v = v0.to(torch.float64).contiguous()
    if basis_dtype == torch.float64:
        st = ctx.lib.ofrr_arnoldi_hess_eig(ctx.h, A.n, A.nnz, _p(A.row_ptr), _p(A.col_idx),
                                           _p(A.vals), k, kdim, maxiter, select_tol, gtol, _p(v), th,
                                           _p(X), A.n, C.byref(built))
    else:
        st = ctx.lib.ofrr_arnoldi_hess_eig_mp(ctx.h, dtype_code(basis_dtype), A.n, A.nnz,
                                              _p(A.row_ptr), _p(A.col_idx), _p(A.vals), k, kdim,
                                              maxiter, select_tol, gtol, _p(v), th, _p(X), A.n,
                                              C.byref(built))
    ctx.check(st, "arnoldi_hess_eig")
    return np.array(th[:]), X, built.value


def subspace_iter_eig(ctx, A, V0, maxiter, step_size=1, select_tol=2.0 ** -52, gtol=0.0):
    """subspace_iter_eig.m with hessenberg + eig_rr(generalised), FP64 on the GPU."""
    A.require_square("subspace_iter_eig")
    _check_rows(V0, A.n, "subspace_iter_eig", "V0")
    m = V0.shape[1]
    V = colmajor(A.n, m, torch.float64, V0.device)
    V.copy_(V0)
    th = (C.c_double * m)()
    ctx.check(ctx.lib.ofrr_subspace_iter_eig(ctx.h, A.n, A.nnz, _p(A.row_ptr), _p(A.col_idx),
                                             _p(A.vals), m, maxiter, step_size, select_tol, gtol,
                                             _p(V), A.n, th), "subspace_iter_eig")
    return np.array(th[:]), V


def filtered_eig(ctx, A, X0, nwant, nd=1, degree=20, maxiter=100, basis_dtype=torch.float16,
                 tol=1e-8, tau=1e-12, lower=0.0, normA=0.0, use_lo_values=True,
                 refine_dtype=None, refine_tol=0.0, lock=True, mid_cols=0, graph=False,
                 rr_whole=False, wide_dtype=None, wide_tol=0.0):
    """Filtered OFRR solve; X0 (n x k, FP64) is copied.  refine_dtype: the
    precision ladder's refinement format (None: off).  lock: soft locking (a
    sweep filters only the columns after the leading converged wanted pairs).
    graph: each filter sweep runs as a captured HIP graph, replayed while its
    shape repeats (same results as plain launches).  rr_whole: every
    Rayleigh-Ritz eigendecomposes the Gram matrix of the whole basis instead of
    the split form (X whitened by Cholesky, the D block's Schur complement).
    wide_dtype / wide_tol: every sweep after the first that starts with a wanted
    residual above wide_tol builds its blocks in wide_dtype (wider than
    basis_dtype; None / 0: off).
    Returns (status, theta, X row-major, res, stats dict)."""
    A.require_square("filtered_eig")
    _check_rows(X0, A.n, "filtered_eig", "X0")
    n, k = X0.shape
    X = X0.to(torch.float64).contiguous().clone()
    bdt = dtype_code(basis_dtype)
    o = FilterOpts(k, nwant, nd, degree, maxiter, bdt, tol, tau, lower, normA,
                   0 if refine_dtype is None else 1,
                   0 if refine_dtype is None else dtype_code(refine_dtype), refine_tol,
                   1 if lock else 0, mid_cols, 1 if graph else 0, 1 if rr_whole else 0,
                   0 if wide_dtype is None else dtype_code(wide_dtype),
                   0.0 if wide_dtype is None else float(wide_tol))
    th = (C.c_double * k)()
    res = (C.c_double * k)()
    stats = FilterStats()
    vlo, lodt = C.c_void_p(0), F64
    if use_lo_values and A.vals_lo is not None and bdt != F64:
        vlo, lodt = _p(A.vals_lo), dtype_code(A.vals_lo)
    st = ctx.lib.ofrr_filtered_eig(ctx.h, A.n, A.nnz, _p(A.row_ptr), _p(A.col_idx), _p(A.vals),
                                   vlo, lodt, C.byref(o), _p(X), th, res, C.byref(stats))
    if st not in (STATUS_SUCCESS, STATUS_MAX_ITER):
        ctx.check(st, "filtered_eig")
    sd = {f: getattr(stats, f) for f, _ in FilterStats._fields_}
    return st, np.array(th[:]), X, np.array(res[:]), sd


# ---- dense operators ---------------------------------------------------------
def dense_gemm(ctx, P, Q, out_dtype=torch.float32, C=None, alpha=1.0, beta=0.0):
    """C = alpha P Q + beta C for row-major P (R x K, unit column stride) and
    Q (K x nc); gemm.m:27-59 over dense operands."""
    R, K = P.shape
    K2, nc = Q.shape
    if K2 != K or P.stride(1) != 1 or Q.stride(1) != 1:
        raise ValueError("dense_gemm: P (R x K) and Q (K x nc) must be row-major")
    if C is None:
        C = torch.zeros(R, nc, dtype=out_dtype, device=P.device)
    ctx.check(ctx.lib.ofrr_dense_gemm(ctx.h, R, K, nc, _p(P), dtype_code(P), P.stride(0), _p(Q),
                                      dtype_code(Q), Q.stride(0), alpha, beta, _p(C),
                                      dtype_code(C), C.stride(0)), "dense_gemm")
    return C


def convert(ctx, X, Y):
    """Y[...] = X element-wise (any formats and strides of two equally shaped 2-D tensors)."""
    r, c = X.shape
    ctx.check(ctx.lib.ofrr_convert(ctx.h, r, c, _p(X), dtype_code(X), X.stride(0), X.stride(1),
                                   _p(Y), dtype_code(Y), Y.stride(0), Y.stride(1)), "convert")
    return Y


def filtered_eig_dense(ctx, A, X0, nwant, nd=1, degree=20, maxiter=100,
                       basis_dtype=torch.float16, tol=1e-8, tau=1e-12, lower=0.0, normA=0.0):
    """Filtered OFRR on a dense symmetric FP64 A (n x n, column-major storage =
    the same tensor, symmetric).  Returns (status, theta, X, res, stats dict)."""
    n, k = X0.shape
    X = X0.to(torch.float64).contiguous().clone()
    o = FilterOpts(k, nwant, nd, degree, maxiter, dtype_code(basis_dtype), tol, tau, lower, normA)
    th = (C.c_double * k)()
    res = (C.c_double * k)()
    stats = FilterStats()
    st = ctx.lib.ofrr_filtered_eig_dense(ctx.h, n, _p(A), A.stride(0), C.byref(o), _p(X), th, res,
                                         C.byref(stats))
    if st not in (STATUS_SUCCESS, STATUS_MAX_ITER):
        ctx.check(st, "filtered_eig_dense")
    sd = {f: getattr(stats, f) for f, _ in FilterStats._fields_}
    return st, np.array(th[:]), X, np.array(res[:]), sd


def svd_filtered(ctx, A, V0, nwant, nd=1, degree=20, maxiter=100, basis_dtype=torch.bfloat16,
                 tol=1e-8, tau=1e-12, want_u=True):
    """Truncated SVD of a dense FP32 A (M x N) by filtered OFRR on A^T A.
    A is given in column-major storage: pass A.t() of a row-major (N x M) tensor,
    or any tensor with A.stride(0) == 1.  Returns (status, sigma, U, V, res, stats)."""
    M, N = A.shape
    if A.stride(0) != 1 or A.dtype != torch.float32:
        raise ValueError("svd_filtered: A must be FP32 column-major (stride(0) == 1)")
    k = V0.shape[1]
    V = V0.to(torch.float64).contiguous().clone()
    U = torch.zeros(M, k, dtype=torch.float64, device=A.device) if want_u else None
    o = FilterOpts(k, nwant, nd, degree, maxiter, dtype_code(basis_dtype), tol, tau, 0.0, 0.0)
    sig = (C.c_double * k)()
    res = (C.c_double * k)()
    stats = FilterStats()
    st = ctx.lib.ofrr_svd_filtered(ctx.h, M, N, _p(A), A.stride(1), C.byref(o), _p(V),
                                   _p(U) if U is not None else None, sig, res, C.byref(stats))
    if st not in (STATUS_SUCCESS, STATUS_MAX_ITER):
        ctx.check(st, "svd_filtered")
    sd = {f: getattr(stats, f) for f, _ in FilterStats._fields_}
    return st, np.array(sig[:]), U, V, np.array(res[:]), sd


# ---- two-sided SVD (reference algorithms, FP64) --------------------------------
def svd_rr(ctx, A, U, V, generalized=True, gtol=0.0):
    """svd_rr.m on the GPU over a dense FP64 A (M x N, any tensor with
    A.stride(0) == 1, i.e. column-major).  U (M x ku), V (N x kv) are copied.
    Returns (sigma (kk), U (M x kk), V (N x kk)), kk = min(ku, kv)."""
    M, N = A.shape
    if A.stride(0) != 1 or A.dtype != torch.float64:
        raise ValueError("svd_rr: A must be FP64 column-major (stride(0) == 1)")
    Uc = colmajor(M, U.shape[1], torch.float64, A.device)
    Uc.copy_(U)
    Vc = colmajor(N, V.shape[1], torch.float64, A.device)
    Vc.copy_(V)
    kk = min(U.shape[1], V.shape[1])
    sig = (C.c_double * kk)()
    ctx.check(ctx.lib.ofrr_svd_rr(ctx.h, M, N, _p(A), A.stride(1), U.shape[1], V.shape[1], _p(Uc),
                                  M, _p(Vc), N, int(generalized), gtol, sig), "svd_rr")
    return np.array(sig[:]), Uc[:, :kk], Vc[:, :kk]


def subspace_iter_svd(ctx, A, V0, maxiter, step_size=1, select_tol=2.0 ** -52, generalized=True,
                      gtol=0.0):
    """subspace_iter_svd.m (Hessenberg orth + svd_rr) in FP64 on the GPU.
    Returns (sigma (k, NaN past the kept width), U (M x k), V (N x k), kept)."""
    M, N = A.shape
    if A.stride(0) != 1 or A.dtype != torch.float64:
        raise ValueError("subspace_iter_svd: A must be FP64 column-major (stride(0) == 1)")
    k = V0.shape[1]
    V0c = colmajor(N, k, torch.float64, A.device)
    V0c.copy_(V0)
    U = colmajor(M, k, torch.float64, A.device)
    V = colmajor(N, k, torch.float64, A.device)
    sig = (C.c_double * k)()
    kept = C.c_int()
    ctx.check(ctx.lib.ofrr_subspace_iter_svd(ctx.h, M, N, _p(A), A.stride(1), k, maxiter, step_size,
                                             select_tol, int(generalized), gtol, _p(V0c), N, _p(U),
                                             M, _p(V), N, sig, C.byref(kept)), "subspace_iter_svd")
    return np.array(sig[:]), U, V, kept.value


# ---- matrix input -----------------------------------------------------------
def read_csr(prefix):
    """Host arrays (row_ptr int64, col_idx int32, vals float64, rows, cols) of the
    binary CSR files at prefix; errors carry the reference's messages."""
    L = _lib.load()
    pb = str(prefix).encode()
    rows, cols, nnz = C.c_int64(), C.c_int64(), C.c_int64()
    st = L.ofrr_csr_file_info(pb, C.byref(rows), C.byref(cols), C.byref(nnz))
    if st != 0:
        raise OfrrError(st, L.ofrr_io_last_error().decode())
    rp = np.zeros(rows.value + 1, np.int64)
    ci = np.zeros(nnz.value, np.int32)
    va = np.zeros(nnz.value, np.float64)
    st = L.ofrr_csr_file_read(pb, rp.ctypes.data, ci.ctypes.data, va.ctypes.data)
    if st != 0:
        raise OfrrError(st, L.ofrr_io_last_error().decode())
    return rp, ci, va, rows.value, cols.value


def read_singular_values(prefix):
    L = _lib.load()
    pb = str(prefix).encode()
    cnt = C.c_int64()
    st = L.ofrr_singular_values_file(pb, None, 0, C.byref(cnt))
    if st != 0:
        raise OfrrError(st, L.ofrr_io_last_error().decode())
    out = np.zeros(cnt.value, np.float64)
    L.ofrr_singular_values_file(pb, out.ctypes.data, cnt.value, C.byref(cnt))
    return out


def _sym_shift(rp, ci, va, n, symmetrize, shift):
    """CSR of (K + K^T) / 2 (symmetrize) + shift I, rows sorted by column."""
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    r, c, v = [rows], [ci.astype(np.int64)], [va]
    if symmetrize:
        r, c, v = [rows, ci.astype(np.int64)], [ci.astype(np.int64), rows], [va / 2, va / 2]
    if shift:
        d = np.arange(n, dtype=np.int64)
        r.append(d)
        c.append(d)
        v.append(np.full(n, float(shift)))
    r, c, v = np.concatenate(r), np.concatenate(c), np.concatenate(v)
    key = r * n + c
    uk, inv = np.unique(key, return_inverse=True)
    vals = np.zeros(len(uk))
    np.add.at(vals, inv, v)
    ur, uc = uk // n, uk % n
    nrp = np.zeros(n + 1, np.int64)
    np.add.at(nrp, ur + 1, 1)
    return np.cumsum(nrp), uc.astype(np.int32), vals


# ---- serving -----------------------------------------------------------------
def request(dim, N, nwant, coef_seed=0, start_seed=0, basis_dtype=torch.float16, tol=1e-8,
            degree=0, maxiter=0, k=0, nd=0, refine_dtype=None, refine_tol=0.0, refine=None,
            mid_cols=0):
    """An eigen request: the nwant largest eigenpairs of the dim-D diffusion
    stencil on an N^dim grid (coef_seed 0: the Laplacian).  k / nd fix the
    block width and correction blocks (0: the server's plan), mid_cols the
    mid-filter block's columns kept in the basis (0: all; basis k + k +
    mid_cols with nd = 2); refine_dtype sets
    the precision ladder's refinement format (refine=False: no ladder; default:
    FP8 finishes in FP16)."""
    if refine_dtype is not None:
        rf, rdt = 1, dtype_code(refine_dtype)
    else:
        rf, rdt = (-1 if refine is False else 0), F64
    return _lib.EigRequest(dim, N, coef_seed, start_seed, nwant, dtype_code(basis_dtype), tol,
                           degree, maxiter, k, nd, rf, rdt, refine_tol, mid_cols)


def server_slot_bytes(device=0):
    """Device bytes each server slot reserves for its lifetime (its workspace);
    a server of s slots admits requests against budget - s * this."""
    L = _lib.load()
    b = C.c_int64()
    st = L.ofrr_server_slot_bytes(device, C.byref(b))
    if st != 0:
        raise OfrrError(st, "server_slot_bytes")
    return b.value


def request_footprint(req):
    """(block width k, correction blocks nd, device bytes) the server plans for
    a request; host only."""
    L = _lib.load()
    k, nd, b = C.c_int(), C.c_int(), C.c_int64()
    st = L.ofrr_request_footprint(C.byref(req), C.byref(k), C.byref(nd), C.byref(b))
    if st != 0:
        raise OfrrError(st, f"request_footprint: {L.ofrr_status_string(st).decode()}")
    return k.value, nd.value, b.value


def _result_dict(req, r):
    w = req.nwant
    return {"status": r.status, "k": r.k, "nd": r.nd, "iterations": r.iterations,
            "converged": r.converged, "filter_steps": r.filter_steps, "n": r.n,
            "footprint": r.footprint, "t_submit": r.t_submit, "t_admit": r.t_admit,
            "t_done": r.t_done, "theta": np.array(r.theta[:w]), "res": np.array(r.res[:w])}


class Server:
    """Continuous-batching eigen server on one GPU (ofrr_server_*): `slots`
    solves run at once on their own streams, admitted in arrival order against
    a device-byte budget (0: 90% of free memory)."""

    def __init__(self, device=0, slots=4, budget=0, graphs=False, rr_group_us=0):
        """graphs: filter sweeps as captured HIP graphs; rr_group_us: the grouped
        Rayleigh-Ritz (0: the jobs that meet, w > 0: also wait up to w us for
        every running solve's job, -1: off)."""
        self.lib = _lib.load()
        h = C.c_void_p()
        opts = _lib.ServerOpts(slots, budget, 1 if graphs else -1, rr_group_us)
        st = self.lib.ofrr_server_create(device, C.byref(opts), C.byref(h))
        if st != 0:
            raise OfrrError(st, f"ofrr_server_create: {self.lib.ofrr_status_string(st).decode()}")
        self.h = h
        self._reqs = {}

    def submit(self, req):
        i = C.c_int64()
        st = self.lib.ofrr_server_submit(self.h, C.byref(req), C.byref(i))
        if st != 0:
            raise OfrrError(st, f"ofrr_server_submit: {self.lib.ofrr_status_string(st).decode()}")
        self._reqs[i.value] = req
        return i.value

    def poll(self, rid):
        """The result dict when request rid is done, else None."""
        r = _lib.EigResult()
        st = self.lib.ofrr_server_poll(self.h, rid, C.byref(r))
        if st == _lib.PENDING:
            return None
        if st != 0:
            raise OfrrError(st, "ofrr_server_poll")
        return _result_dict(self._reqs.pop(rid), r)

    def wait(self, rid):
        r = _lib.EigResult()
        st = self.lib.ofrr_server_wait(self.h, rid, C.byref(r))
        if st != 0:
            raise OfrrError(st, "ofrr_server_wait")
        return _result_dict(self._reqs.pop(rid), r)

    def stats(self):
        s = _lib.ServerStats()
        self.lib.ofrr_server_stats_get(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in _lib.ServerStats._fields_}

    def close(self):
        if self.h:
            self.lib.ofrr_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
