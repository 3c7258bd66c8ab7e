"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product package
never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_ip64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class OfrrOpts(C.Structure):
    _fields_ = [("k", C.c_int), ("nwant", C.c_int), ("nd", C.c_int), ("degree", C.c_int),
                ("maxiter", C.c_int), ("tol", C.c_double), ("tau", C.c_double),
                ("lower", C.c_double), ("normA", C.c_double), ("lock", C.c_int),
                ("rr_whole", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [_u32, _u32, _u32]
        L.orc_randn.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, _dp]
        L.orc_randu.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int64, _dp]
        L.orc_stencil_nnz.argtypes = [C.c_int, C.c_int64]
        L.orc_stencil_nnz.restype = C.c_int64
        L.orc_stencil_csr.argtypes = [C.c_int, C.c_int64, C.c_uint64, _ip64, _ip32, _dp]
        L.orc_spmm_csr.argtypes = [C.c_int64, _ip64, _ip32, _dp, C.c_int, _dp, C.c_int64, _dp,
                                   C.c_int64]
        L.orc_spmm_csr_t.argtypes = [C.c_int64, C.c_int64, _ip64, _ip32, _dp, C.c_int, _dp,
                                     C.c_int64, _dp, C.c_int64]
        L.orc_scalecols.argtypes = [C.c_int64, C.c_int, _dp, C.c_int64]
        L.orc_iamax.argtypes = [C.c_int64, _dp, C.POINTER(C.c_double)]
        L.orc_matrix_iamax.argtypes = [C.c_int64, C.c_int, _dp, C.c_int64, _ip32]
        L.orc_gemm_tn.argtypes = [C.c_int64, C.c_int, C.c_int, _dp, C.c_int64, _dp, C.c_int64,
                                  _dp, C.c_int]
        L.orc_gemm_nn.argtypes = [C.c_int64, C.c_int, C.c_int, _dp, C.c_int64, _dp, C.c_int64,
                                  _dp, C.c_int64]
        L.orc_hessenberg.argtypes = [C.c_int64, C.c_int, _dp, C.c_int64, _dp, C.c_int, C.c_double,
                                     C.c_int, _ip32, _ip32]
        L.orc_gs_qr.argtypes = [C.c_int64, C.c_int, _dp, C.c_int64, _dp, C.c_int, C.c_int,
                                C.c_double, C.c_double, C.c_double, _ip32]
        L.orc_syev.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_sygv.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.orc_rr_truncated.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_int,
                                       _dp, _dp, C.c_int]
        L.orc_rr_split.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double,
                                   C.c_int, _dp, _dp, C.c_int]
        L.orc_rr_generalized.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, _dp,
                                         _dp, C.c_int]
        L.orc_arnoldi_hess_eig.argtypes = [C.c_int64, _ip64, _ip32, _dp, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, _dp, _dp, _dp, C.c_int64]
        L.orc_subspace_iter_eig.argtypes = [C.c_int64, _ip64, _ip32, _dp, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_double, _dp, C.c_int64, _dp]
        L.orc_filtered_ofrr.argtypes = [C.c_int64, _ip64, _ip32, _dp, C.POINTER(OfrrOpts), _dp,
                                        C.c_int64, _dp, _dp]
        L.orc_residuals.argtypes = [C.c_int64, _ip64, _ip32, _dp, C.c_int, _dp, C.c_int64, _dp,
                                    C.c_double, _dp]
        L.orc_svd_rr.argtypes = [C.c_int64, C.c_int64, _dp, C.c_int64, C.c_int, C.c_int, _dp,
                                 C.c_int64, _dp, C.c_int64, C.c_int, C.c_double, _dp]
        L.orc_subspace_iter_svd.argtypes = [C.c_int64, C.c_int64, _dp, C.c_int64, C.c_int, C.c_int,
                                            C.c_int, C.c_double, C.c_int, C.c_double, _dp, _dp,
                                            C.c_int64, _dp, C.c_int64, _dp]
        _lib = L
    return _lib


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def randn(seed, n, stream=0, offset=0):
    out = np.zeros(n, np.float64)
    lib().orc_randn(seed, stream, offset, n, out)
    return out


def randu(seed, n, stream=0, offset=0):
    out = np.zeros(n, np.float64)
    lib().orc_randu(seed, stream, offset, n, out)
    return out


class CSR:
    def __init__(self, n, row_ptr, col_idx, vals, ncols=None):
        self.n = n
        self.ncols = n if ncols is None else ncols
        self.row_ptr = np.ascontiguousarray(row_ptr, np.int64)
        self.col_idx = np.ascontiguousarray(col_idx, np.int32)
        self.vals = np.asfortranarray(vals, np.float64)

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.col_idx, self.row_ptr), shape=(self.n, self.ncols))


def stencil(dim, N, seed=0):
    L = lib()
    n = N ** dim
    nnz = L.orc_stencil_nnz(dim, N)
    rp = np.zeros(n + 1, np.int64)
    ci = np.zeros(nnz, np.int32)
    va = np.zeros(nnz, np.float64)
    L.orc_stencil_csr(dim, N, seed, rp, ci, va)
    return CSR(n, rp, ci, va)


def spmm(A, X):
    """A X (csrmv 'N'); X has A.ncols rows."""
    X = _f(X.reshape(A.ncols, -1))
    k = X.shape[1]
    Y = np.zeros((A.n, k), order="F")
    lib().orc_spmm_csr(A.n, A.row_ptr, A.col_idx, A.vals, k, X, A.ncols, Y, A.n)
    return Y


def spmm_t(A, X):
    """A^T X (csrmv 'T', csrmv.m:61-71); X has A.n rows."""
    X = _f(X.reshape(A.n, -1))
    k = X.shape[1]
    Y = np.zeros((A.ncols, k), order="F")
    lib().orc_spmm_csr_t(A.n, A.ncols, A.row_ptr, A.col_idx, A.vals, k, X, A.n, Y, A.ncols)
    return Y


def scalecols(A):
    A = _f(A).copy(order="F")
    lib().orc_scalecols(A.shape[0], A.shape[1], A, A.shape[0])
    return A


def iamax(x):
    x = _f(x)
    v = C.c_double()
    i = lib().orc_iamax(x.shape[0], x, C.byref(v))
    return i, v.value


def matrix_iamax(A):
    A = _f(A)
    idx = np.zeros(A.shape[1], np.int32)
    lib().orc_matrix_iamax(A.shape[0], A.shape[1], A, A.shape[0], idx)
    return idx


def gemm_tn(V, W):
    V, W = _f(V), _f(W)
    G = np.zeros((V.shape[1], W.shape[1]), order="F")
    lib().orc_gemm_tn(V.shape[0], V.shape[1], W.shape[1], V, V.shape[0], W, W.shape[0], G,
                      V.shape[1])
    return G


def gemm_nn(A, B):
    A, B = _f(A), _f(B)
    Cm = np.zeros((A.shape[0], B.shape[1]), order="F")
    lib().orc_gemm_nn(A.shape[0], B.shape[1], A.shape[1], A, A.shape[0], B, B.shape[0], Cm,
                      A.shape[0])
    return Cm


def hessenberg(A, select_tol, variant=0):
    Q = _f(A).copy(order="F")
    m, n = Q.shape
    R = np.zeros((n, n), order="F")
    sel = np.zeros(n, np.int32)
    perm = np.zeros(n, np.int32)
    lib().orc_hessenberg(m, n, Q, m, R, n, select_tol, variant, sel, perm)
    return Q, R, sel, perm


QR_VARIANTS = {"mgs": 0, "mgs_v2": 1, "cgs": 2, "cgs2": 3}


def gs_qr(A, variant, select_tol=2.0 ** -52, orth_tol=2.0 ** -52, reorth_tol=-1.0):
    """(Q, R, skip) of the Gram-Schmidt variant ('mgs', 'mgs_v2', 'cgs', 'cgs2')."""
    Q = _f(A).copy(order="F")
    m, n = Q.shape
    R = np.zeros((n, n), order="F")
    skip = np.zeros(n, np.int32)
    lib().orc_gs_qr(m, n, Q, m, R, n, QR_VARIANTS[variant], select_tol, orth_tol, reorth_tol,
                    skip)
    return Q, R, skip


def syev(A):
    A = _f(A)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order="F")
    lib().orc_syev(n, A, n, w, V, n)
    return w, V


def sygv(A, B):
    A, B = _f(A), _f(B)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order="F")
    info = lib().orc_sygv(n, A, n, B, n, w, V, n)
    return info, w, V


def rr_truncated(G, H, tau, nkeep):
    G, H = _f(G), _f(H)
    m = G.shape[0]
    th = np.zeros(nkeep)
    Y = np.zeros((m, nkeep), order="F")
    r = lib().orc_rr_truncated(m, G, m, H, m, tau, nkeep, th, Y, m)
    return r, th, Y


def rr_split(G, H, kx, tau, nkeep):
    G, H = _f(G), _f(H)
    m = G.shape[0]
    th = np.zeros(nkeep)
    Y = np.zeros((m, nkeep), order="F")
    r = lib().orc_rr_split(m, kx, G, m, H, m, tau, nkeep, th, Y, m)
    return r, th, Y


def rr_generalized(G, H, gtol):
    G, H = _f(G), _f(H)
    m = G.shape[0]
    th = np.zeros(m)
    Y = np.zeros((m, m), order="F")
    info = lib().orc_rr_generalized(m, G, m, H, m, gtol, th, Y, m)
    return info, th, Y


def arnoldi_hess_eig(A, k, kdim, maxiter, v0, select_tol=2.0 ** -52, gtol=0.0):
    th = np.zeros(k)
    X = np.zeros((A.n, k), order="F")
    built = lib().orc_arnoldi_hess_eig(A.n, A.row_ptr, A.col_idx, A.vals, k, kdim, maxiter,
                                       select_tol, gtol, _f(v0), th, X, A.n)
    return th, X, built


def subspace_iter_eig(A, V0, maxiter, step_size, select_tol=2.0 ** -52, gtol=0.0):
    V = _f(V0).copy(order="F")
    m = V.shape[1]
    th = np.zeros(m)
    lib().orc_subspace_iter_eig(A.n, A.row_ptr, A.col_idx, A.vals, m, maxiter, step_size,
                                select_tol, gtol, V, A.n, th)
    return th, V


def svd_rr(A, U, V, generalized=True, gtol=0.0):
    """svd_rr.m over a dense A: (sigma (kk), U (m x kk), V (n x kk)), kk = min(ku, kv)."""
    A = _f(A)
    U = _f(U).copy(order="F")
    V = _f(V).copy(order="F")
    m, n = A.shape
    ku, kv = U.shape[1], V.shape[1]
    sig = np.zeros(min(ku, kv))
    kk = lib().orc_svd_rr(m, n, A, m, ku, kv, U, m, V, n, int(generalized), gtol, sig)
    return sig[:kk], U[:, :kk], V[:, :kk]


def subspace_iter_svd(A, V0, maxiter, step_size=1, select_tol=2.0 ** -52, generalized=True,
                      gtol=0.0):
    """subspace_iter_svd.m with hessenberg + svd_rr: (sigma (k), U, V, kept width)."""
    A = _f(A)
    m, n = A.shape
    k = V0.shape[1]
    U = np.zeros((m, k), order="F")
    V = np.zeros((n, k), order="F")
    sig = np.zeros(k)
    kc = lib().orc_subspace_iter_svd(m, n, A, m, k, maxiter, step_size, select_tol,
                                     int(generalized), gtol, _f(V0), U, m, V, n, sig)
    return sig, U, V, kc


def filtered_ofrr(A, X0, nwant, nd=1, degree=20, maxiter=50, tol=1e-8, tau=1e-12, lower=0.0,
                  normA=0.0, lock=True, rr_whole=False):
    X = _f(X0).copy(order="F")
    k = X.shape[1]
    o = OfrrOpts(k, nwant, nd, degree, maxiter, tol, tau, lower, normA, 1 if lock else 0,
                 1 if rr_whole else 0)
    th = np.zeros(k)
    res = np.zeros(k)
    it = lib().orc_filtered_ofrr(A.n, A.row_ptr, A.col_idx, A.vals, C.byref(o), X, A.n, th, res)
    return it, th, X, res


def residuals(A, X, theta, normA):
    X = _f(X)
    k = X.shape[1]
    res = np.zeros(k)
    lib().orc_residuals(A.n, A.row_ptr, A.col_idx, A.vals, k, X, A.n,
                        np.ascontiguousarray(theta, np.float64), normA, res)
    return res


def laplacian_eigs(dim, N):
    """Closed-form spectrum of the Dirichlet stencil Laplacian, descending."""
    q = np.arange(1, N + 1)
    l1 = 4.0 * np.sin(q * np.pi / (2.0 * (N + 1))) ** 2
    if dim == 2:
        L = (l1[:, None] + l1[None, :]).ravel()
    else:
        L = (l1[:, None, None] + l1[None, :, None] + l1[None, None, :]).ravel()
    return np.sort(L)[::-1]
