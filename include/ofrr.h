/*
 * ofrr.h -- C-ABI of the MI355X (gfx950) OFRR projection pipeline.
 *
 * Every entry point returns an int status whose values are the reference's
 * MSVDStatus codes (CPP/core/utils/error_handling.hpp:24-62): 0 success,
 * 1000 invalid argument, 1003 not implemented, 1005 NaN, 1006 Inf,
 * 2005 invalid size, 4001 max iterations, 6000-6003 HIP runtime errors.
 * The message of the last failure on a context is ofrr_ctx_last_error().
 *
 * Pointers named *_dev are device pointers; everything else documented as
 * host.  Dense blocks are addressed with two strides: element (i, j) of X is
 * X[i * rs + j * cs], so column-major (rs = 1, cs = ld, the reference's
 * Matrix<T> layout, CPP/containers/matrix.cu:89-99) and row-major
 * (rs = ld, cs = 1, the layout the hot path keeps in HBM) are both accepted.
 * All work is enqueued on the context's stream; the call returns when the work
 * is enqueued unless documented as synchronous.
 *
 * The reference C++ functions each entry point replaces are cited per
 * declaration (paths relative to the reference repository root).
 */
#ifndef OFRR_H
#define OFRR_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OFRR_API __attribute__((visibility("default")))

typedef struct ofrr_ctx ofrr_ctx;

/* element types (the reference's T_I/T_O/T_COMPUTE instantiations
 * double/float/__half, CPP/containers/vector.cu:11-13, plus bf16 and OCP FP8 E4M3) */
enum ofrr_dtype { OFRR_F64 = 0, OFRR_F32 = 1, OFRR_F16 = 2, OFRR_BF16 = 3, OFRR_FP8 = 4 };

/* ---- context: replaces the CUDAHandler singleton
 *      (CPP/core/utils/cuda_handler.cu:9-45): one per host thread and device */
OFRR_API int ofrr_ctx_create(int device, ofrr_ctx **ctx);
OFRR_API int ofrr_ctx_destroy(ofrr_ctx *ctx);
OFRR_API int ofrr_ctx_sync(ofrr_ctx *ctx);
/* enqueue on an external hipStream_t; NULL selects the null (default) stream */
OFRR_API int ofrr_ctx_set_stream(ofrr_ctx *ctx, void *stream);
/* enqueue on the context's own stream (the state after ofrr_ctx_create) */
OFRR_API int ofrr_ctx_use_own_stream(ofrr_ctx *ctx);
OFRR_API void *ofrr_ctx_stream(ofrr_ctx *ctx);
OFRR_API const char *ofrr_ctx_last_error(ofrr_ctx *ctx);
OFRR_API const char *ofrr_status_string(int status);
OFRR_API int ofrr_version(void);

/* ---- random fill: replaces Vector<T>::fill_random (CPP/containers/vector.cu:188-225).
 * x[i*inc] = N(0,1) draw number (offset + i) of the Philox4x32-10 stream
 * (seed, stream), Box-Muller, rounded to dtype. */
OFRR_API int ofrr_fill_randn(ofrr_ctx *ctx, int dtype, int64_t n, void *x_dev, int64_t inc,
                             uint64_t seed, uint32_t stream, uint64_t offset);
/* x[i*inc] = U[0,1) draw number (offset + i) of stream (seed, stream) */
OFRR_API int ofrr_fill_randu(ofrr_ctx *ctx, int dtype, int64_t n, void *x_dev, int64_t inc,
                             uint64_t seed, uint32_t stream, uint64_t offset);

/* ---- operators.  CSR with 0-based int64 row_ptr and int32 col_idx.  The CSR
 * format and SpMM semantics follow MATLAB/+src/+csr/sparse2csr.m:21-50 and
 * csrmv.m:27-76 (y(i,:) = sum_p a_p x(j_p,:)); the reference has no C++ CSR path. */
/* stencil operator on an N^dim grid (dim 2: 5-point, 3: 7-point), natural
 * x-fastest ordering, Dirichlet; seed 0 = Laplacian (diag 2*dim, off -1),
 * else face coefficients 10^(2u-1), u ~ U[0,1).  Returns nnz in *nnz (host). */
OFRR_API int ofrr_stencil_nnz(int dim, int64_t N, int64_t *nnz);
OFRR_API int ofrr_stencil_csr(ofrr_ctx *ctx, int dim, int64_t N, uint64_t seed,
                              int64_t *row_ptr_dev, int32_t *col_idx_dev, int val_dtype,
                              void *vals_dev);
/* Y = alpha * A X + beta * Y, FP64 accumulation for FP64 operands, FP32
 * otherwise (beta == 0: Y is not read).  X: n x k, Y: n x k (csrmv.m:48-59). */
OFRR_API int ofrr_spmm_csr(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                           const int32_t *col_idx_dev, const void *vals_dev, int val_dtype,
                           int k, const void *X_dev, int x_dtype, int64_t x_rs, int64_t x_cs,
                           double alpha, double beta, void *Y_dev, int y_dtype, int64_t y_rs,
                           int64_t y_cs);

/* Rectangular A (nrows x ncols CSR, column indices in [0, ncols)) and the
 * transposed product, csrmv(Acsr, trans, X) of MATLAB/+src/+csr/csrmv.m:27-76:
 * trans 0: Y (nrows x k) = alpha A X + beta Y, X ncols x k ('N', :48-59);
 * trans 1: Y (ncols x k) = alpha A^T X + beta Y, X nrows x k ('T', :61-71).
 * The transposed product runs over A^T built on the device (ofrr_csr_transpose,
 * in the context's workspace, synchronous), so FP64 results equal the 'T'
 * branch's sequential scatter (rows in increasing order, fused multiply-adds)
 * bit for bit.  Accumulation as ofrr_spmm_csr. */
OFRR_API int ofrr_spmm_csr_rect(ofrr_ctx *ctx, int trans, int64_t nrows, int64_t ncols, int64_t nnz,
                                const int64_t *row_ptr_dev, const int32_t *col_idx_dev,
                                const void *vals_dev, int val_dtype, int k, const void *X_dev,
                                int x_dtype, int64_t x_rs, int64_t x_cs, double alpha, double beta,
                                void *Y_dev, int y_dtype, int64_t y_rs, int64_t y_cs);
/* A^T of a CSR A (nrows x ncols) as a CSR of ncols x nrows (row_ptrT ncols + 1,
 * col_idxT and valsT nnz, device), the entries of each row of A^T in increasing
 * row of A (a stable sort by column).  Build it once for repeated transposed
 * products (subspace_iter_svd.m:63-77 alternates A and A^T) and pass it to
 * ofrr_spmm_csr_rect with trans 0.  1000 when a column index lies outside
 * [0, ncols).  Synchronous. */
OFRR_API int ofrr_csr_transpose(ofrr_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                                const int64_t *row_ptr_dev, const int32_t *col_idx_dev,
                                const void *vals_dev, int val_dtype, int64_t *row_ptrT_dev,
                                int32_t *col_idxT_dev, void *valsT_dev);

/* ---- pivots and scaling */
/* 1-based index of the first max |x_i| and its signed value, both written to
 * device memory: iamax<T_O,T_COMPUTE>(x, result, result_val, kDEVICE)
 * (CPP/linalg/blas/mvops.hpp:98-99, mvops.cu:607-946). */
OFRR_API int ofrr_iamax(ofrr_ctx *ctx, int dtype, int64_t n, const void *x_dev, int64_t inc,
                        int32_t *idx1_dev, double *val_dev);
/* 0-based per-column argmax |A(:,j)| (mvops.hpp:113-114, mvops.cu:955-1199) */
OFRR_API int ofrr_matrix_iamax(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A_dev,
                               int64_t rs, int64_t cs, int32_t *idx0_dev);
/* divide each column by its signed max-abs entry, first index on ties, skip
 * zero columns (MATLAB/+src/+mvops/scalecols.m:21-34); the pivot values are
 * written to piv_dev (n doubles) when it is not NULL */
OFRR_API int ofrr_scalecols(ofrr_ctx *ctx, int dtype, int64_t m, int n, void *A_dev, int64_t rs,
                            int64_t cs, double *piv_dev);
/* check_special_values (mvops.cu:1388-1599): synchronous; returns 0, 1005 (NaN
 * first) or 1006 (Inf) */
OFRR_API int ofrr_check_special(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A_dev,
                                int64_t rs, int64_t cs);

/* ---- Hessenberg process (pivoted, max-abs scaled, no orthogonalization):
 * hessenberg<T_I,T_O,T_COMPUTE> / hessenberg_v3 (CPP/linalg/factorization/
 * hessenberg.hpp:22-108) with MATLAB/+src/+qr/hessenberg.m:55-74 semantics.
 * In place on Q (m x n, dtype q_dtype, arithmetic in FP64 for F64, FP32
 * otherwise).  R (n x n FP64, column-major, device) receives the upper
 * triangular factor, perm_dev the 0-based pivot rows (-1 for dropped columns),
 * skip_dev 1 for dropped columns.  variant 0 skips dropped columns in later
 * eliminations (hessenberg.m:58, hessenberg.cu:520); a dropped column keeps its
 * unscaled residual (hessenberg.cu:515-522). */
OFRR_API int ofrr_hessenberg(ofrr_ctx *ctx, int q_dtype, int64_t m, int n, void *Q_dev,
                             int64_t rs, int64_t cs, double select_tol, double *R_dev, int ldr,
                             int32_t *perm_dev, int32_t *skip_dev);

/* The C++ v3 variant (hessenberg_v3, CPP/linalg/factorization/hessenberg.cu:
 * 617-801): as ofrr_hessenberg, except that a dropped column still takes part
 * in the later eliminations, with its argmax row and its unscaled residual
 * (the v3 step eliminates against every earlier column, hessenberg.cu:768-779).
 * Without dropped columns the two give the same result bit for bit. */
OFRR_API int ofrr_hessenberg_v3(ofrr_ctx *ctx, int q_dtype, int64_t m, int n, void *Q_dev,
                                int64_t rs, int64_t cs, double select_tol, double *R_dev, int ldr,
                                int32_t *perm_dev, int32_t *skip_dev);
/* One column of the process (hessenberg_v3_step, hessenberg.hpp:93-104):
 * column i of Q against its finished columns 0..i-1 (variant 1: every one, as
 * v3; variant 0: the kept ones, as hessenberg.m), then its pivot: R(0..i, i)
 * (FP64, device), prow1_dev[i] = the 1-based argmax row (written for dropped
 * columns too; prow1_dev[0..i) must hold those of the finished columns),
 * skip_dev[i] and *skipped_host (nullable; when given, the call synchronizes).
 * Left-looking: the coefficients follow by forward substitution over the pivot
 * rows, then one pass over the rows -- the same rounded operations as the full
 * factorization, so columns stepped one at a time equal ofrr_hessenberg(_v3)
 * bit for bit. */
OFRR_API int ofrr_hessenberg_step(ofrr_ctx *ctx, int variant, int q_dtype, int64_t m, int n,
                                  void *Q_dev, int64_t rs, int64_t cs, int i, double select_tol,
                                  double *R_dev, int ldr, int32_t *prow1_dev, int32_t *skip_dev,
                                  int32_t *skipped_host);

/* ---- Gram-Schmidt QR, the orthogonalizing baselines (CPP/linalg/
 * factorization/qr.hpp:9-141, MATLAB/+src/+qr/mgs.m, cgs.m).  In place on Q
 * (m x n, m >= n, dtype q_dtype: FP64, FP32, FP16, BF16); dot products and
 * norms accumulate in FP64.  R (n x n FP64, column-major, device) receives the
 * triangular factor, skip_dev 1 for dropped columns.  A dropped column keeps its
 * unnormalized residual in Q and is left out of later projections.
 *   OFRR_QR_MGS     mgs (qr.hpp:9-28, qr.cu:12-249): sequential projections,
 *                   reorthogonalized while orth_tol <= ||v|| < reorth_tol *
 *                   ||v before the pass|| (reorth_tol <= 0: off); dropped when
 *                   the initial or final norm is below select_tol, R(i,i) = 0;
 *   OFRR_QR_MGS_V2  mgs_v2 (qr.hpp:30-50, qr.cu:251-393): right-looking MGS,
 *                   no reorthogonalization; dropped R(i,i) = 1;
 *   OFRR_QR_CGS     cgs (qr.hpp:77-98, qr.cu:395-569): one projection product
 *                   per pass; reorth_tol > 0 adds passes under the MGS
 *                   criterion (cgs.m:73-89; the C++ cgs ignores it: pass 0);
 *                   dropped R(i,i) = 1;
 *   OFRR_QR_CGS2    cgs2 (qr.hpp:110-131, qr.cu:571-754): two passes, R(:,i) =
 *                   r1 + r2, dropped R(i,i) = 1.  The second projection is
 *                   subtracted from v (the reference adds r2 to R only,
 *                   qr.cu:713-717, which leaves the CGS Q).
 * Decisions stay on the device; with reorthogonalization on, one flag per column
 * is read back (the call then synchronizes once per column). */
#define OFRR_QR_MGS 0
#define OFRR_QR_MGS_V2 1
#define OFRR_QR_CGS 2
#define OFRR_QR_CGS2 3
OFRR_API int ofrr_gs_qr(ofrr_ctx *ctx, int variant, int q_dtype, int64_t m, int n, void *Q_dev,
                        int64_t rs, int64_t cs, double select_tol, double orth_tol,
                        double reorth_tol, double *R_dev, int ldr, int32_t *skip_dev);

/* One column step: column i of Q against the final columns 0..i-1 whose drop
 * flags are in skip_dev[0..i): mgs_step (qr.hpp:52-75, qr.cu:84-249),
 * cgs_step (qr.hpp:100-107, qr.cu:489-569), cgs2_step (qr.hpp:133-141,
 * qr.cu:666-754).  Writes R(0..i, i), skip_dev[i] and *skipped_host;
 * synchronous.  OFRR_QR_MGS_V2 has no step (1000). */
OFRR_API int ofrr_gs_qr_step(ofrr_ctx *ctx, int variant, int q_dtype, int64_t m, int n,
                             void *Q_dev, int64_t rs, int64_t cs, int i, double select_tol,
                             double orth_tol, double reorth_tol, double *R_dev, int ldr,
                             int32_t *skip_dev, int32_t *skipped_host);

/* ---- projection products: G = alpha * (V^T W) + beta * G in FP64
 * accumulation with a fixed-order (deterministic) split over rows.
 * gemm(transA = true, ...) of mvops.hpp:79-82 / eig_rr.m:49,51 used as
 * V^T V and V^T (A V); V: n x k1, W: n x k2, G: k1 x k2 FP64 column-major. */
OFRR_API int ofrr_gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V_dev, int v_dtype,
                       int64_t v_rs, int64_t v_cs, const void *W_dev, int w_dtype, int64_t w_rs,
                       int64_t w_cs, double alpha, double beta, double *G_dev, int ldg);
/* ofrr_gram without device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and FP64 accumulation as ofrr_gram; the per-chunk partial sums
 * live in a static 8 MiB device array per GPU (set up by ofrr_ctx_create),
 * so it calls no hipMalloc / hipFree.  Limit: ceil(k1/64) * ceil(k2/64) <= 256
 * output tiles (k1 = k2 <= 1024), else 2005 with a message.  Its chunking
 * differs from ofrr_gram's (about 256 chunk tiles instead of 2048), so the two
 * agree to FP64 rounding, not bit for bit.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other host
 * threads) race. */
OFRR_API int ofrr_gram_noalloc(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V_dev,
                               int v_dtype, int64_t v_rs, int64_t v_cs, const void *W_dev,
                               int w_dtype, int64_t w_rs, int64_t w_cs, double alpha, double beta,
                               double *G_dev, int ldg);
/* Y = sum_b V_b C_b over nb blocks (Y n x k FP64 or FP32; V_b n x kb of any
 * dtype; C_b kb x k FP64 column-major): the back-transform V <- V U of
 * eig_rr.m:56 for a basis whose column blocks have different dtypes. */
OFRR_API int ofrr_basis_update(ofrr_ctx *ctx, int64_t n, int k, int nb, const void *const *V_dev,
                               const int *v_dtype, const int *kb, const int64_t *v_rs,
                               const int64_t *v_cs, const double *const *C_dev, const int *ldc,
                               void *Y_dev, int y_dtype, int64_t y_rs, int64_t y_cs);

/* ---- FP64 dense symmetric eigensolver and Rayleigh-Ritz on the GPU */
/* batched syev: eigenvalues ascending (w_dev, n per matrix), eigenvectors in
 * V (column-major, ldv), A read from its upper triangle (syev of
 * mvops.hpp:184-197, LAPACKE dsyev 'V','U' at mvops.cu:1617).  One
 * workgroup per matrix, parallel cyclic Jacobi. */
OFRR_API int ofrr_syev_batched(ofrr_ctx *ctx, int batch, int n, const double *A_dev, int lda,
                               int64_t strideA, double *w_dev, int64_t stridew, double *V_dev,
                               int ldv, int64_t strideV);
/* A x = lambda B x (itype 1), eigenvalues ascending, B-normalised vectors
 * (sygv of mvops.hpp:184-197, LAPACKE dsygv_work at mvops.cu:1650); returns
 * 3001 (kErrorSingularMatrix) when B is not positive definite */
OFRR_API int ofrr_sygv(ofrr_ctx *ctx, int n, const double *A_dev, int lda, const double *B_dev,
                       int ldb, double *w_dev, double *V_dev, int ldv);
/* Truncated OFRR Rayleigh-Ritz (ofrr_eig of CPP/linalg/solvers/eig/rr_eig.hpp:
 * 21-35 with select_tol = tau): S = diag(G)^{-1/2}; S G S = U L U^T; keep
 * L_i > tau * max(L); B = S U_r L_r^{-1/2}; eig(B^T H B) -> theta (nkeep,
 * descending) and Y = B Z (m x nkeep, column-major).  rank_dev: rank kept. */
OFRR_API int ofrr_rr_truncated(ofrr_ctx *ctx, int m, const double *G_dev, int ldg,
                               const double *H_dev, int ldh, double tau, int nkeep,
                               double *theta_dev, double *Y_dev, int ldy, int32_t *rank_dev);
/* The same Rayleigh-Ritz for a basis V = [X | D] whose first kx columns
 * (kx <= 128) are a previous Ritz block, in split form (the filtered solver's
 * default): G_XX = L L^T; S = G_DD - G_XD^T G_XX^{-1} G_XD scaled to unit
 * diagonal; eig(S) = U L, keep L_i > tau * max(1, max L); W = [[L^{-T}, -G_XX^{-1}
 * G_XD B_D], [0, B_D]] with B_D = s U_r L_r^{-1/2}; eig of the D block of W^T H W,
 * then of the coupled matrix -> theta (nkeep, descending), Y (m x nkeep) with
 * Y^T G Y = I.  rank_dev: kx + the kept D rank, or -1 when G_XX is not positive
 * definite (theta and Y are then not written).  No counterpart in the reference:
 * a faster route to rr_eig's pencil for the solver's bases. */
OFRR_API int ofrr_rr_split(ofrr_ctx *ctx, int m, int kx, const double *G_dev, int ldg,
                           const double *H_dev, int ldh, double tau, int nkeep, double *theta_dev,
                           double *Y_dev, int ldy, int32_t *rank_dev);
/* the reference's generalised form eig(H, G + gtol I) (eig_rr.m:52), sorted
 * descending; all m pairs */
OFRR_API int ofrr_rr_generalized(ofrr_ctx *ctx, int m, const double *G_dev, int ldg,
                                 const double *H_dev, int ldh, double gtol, double *theta_dev,
                                 double *Y_dev, int ldy);

/* ---- residuals: res[j] = ||A x_j - theta_j x_j||_2 / (normA ||x_j||_2) in
 * FP64 (MATLAB/+tests/+test4/test.m:340-350, normalised); synchronous, host res */
OFRR_API int ofrr_residuals(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                            const int32_t *col_idx_dev, const double *vals_dev, int k,
                            const double *X_dev, int64_t x_rs, int64_t x_cs,
                            const double *theta_host, double normA, double *res_host);

/* ---- drivers ------------------------------------------------------------ */
/* Arnoldi-Hessenberg eigensolver (arnoldi_eig of CPP/linalg/solvers/eig/
 * arnoldi_eig.hpp:25-28 over a CSR operator, semantics of MATLAB/+src/+krylov/
 * arnoldi_hess_eig.m:43-107) in FP64; kdim + 1 <= 384.  v0_dev: start vector (n).  theta_host
 * (k, descending) and X_dev (n x k column-major, ldx) out.  Synchronous.
 * *built (host, nullable) receives the basis columns of the last sweep. */
OFRR_API int ofrr_arnoldi_hess_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz,
                                   const int64_t *row_ptr_dev, const int32_t *col_idx_dev,
                                   const double *vals_dev, int k, int kdim, int maxiter,
                                   double select_tol, double gtol, const double *v0_dev,
                                   double *theta_host, double *X_dev, int64_t ldx, int *built);

/* The same solver with a low-precision Krylov basis (arnoldi_hess_eig.m:9-31
 * precision options): the basis stored in basis_dtype (precision_output: FP64,
 * FP32, FP16 or BF16), its arithmetic and the sweep's operator products in FP32
 * (FP64 for an FP64 basis; precision_compute / precision_A_matvec_*), the
 * Rayleigh-Ritz in FP64 (the basis widened, V^T V and V^T A V with FP64
 * accumulation: precision_eigenvalues double).  The pivot decisions stay on the
 * device: a sweep of kdim columns synchronizes once, not once per column.
 * theta_host and X_dev (FP64) as for ofrr_arnoldi_hess_eig; kdim + 1 <= 384. */
OFRR_API int ofrr_arnoldi_hess_eig_mp(ofrr_ctx *ctx, int basis_dtype, int64_t n, int64_t nnz,
                                      const int64_t *row_ptr_dev, const int32_t *col_idx_dev,
                                      const double *vals_dev, int k, int kdim, int maxiter,
                                      double select_tol, double gtol, const double *v0_dev,
                                      double *theta_host, double *X_dev, int64_t ldx, int *built);

/* Block subspace iteration (si_eig of CPP/linalg/solvers/eig/si_eig.hpp:24-27
 * over a CSR operator, semantics of MATLAB/+src/+subspace/subspace_iter_eig.m:
 * 53-68 with function_orth = hessenberg and function_rr = eig_rr generalised)
 * in FP64.  V_dev (n x m column-major, ldv) start block in / Ritz block out;
 * theta_host (m) out, sorted descending.  Synchronous. */
OFRR_API int ofrr_subspace_iter_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz,
                                    const int64_t *row_ptr_dev, const int32_t *col_idx_dev,
                                    const double *vals_dev, int m, int maxiter, int step_size,
                                    double select_tol, double gtol, double *V_dev, int64_t ldv,
                                    double *theta_host);

/* Filtered OFRR: block subspace iteration whose growth blocks are Chebyshev-
 * filtered residual corrections built in basis_dtype (FP64, FP32, FP16, BF16
 * or FP8 E4M3) with FP32 accumulation, FP64 Ritz block X, FP64 Gram /
 * projected products, FP64 truncated Rayleigh-Ritz on the GPU. */
typedef struct {
   int k;            /* block width (Ritz block columns) */
   int nwant;        /* wanted largest eigenpairs */
   int nd;           /* correction blocks in the basis (basis dim = k * (1 + nd)) */
   int degree;       /* filter degree per outer iteration */
   int maxiter;      /* outer iterations */
   int basis_dtype;  /* ofrr_dtype of the growth blocks */
   double tol;       /* ||A x - theta x|| / (||A|| ||x||) */
   double tau;       /* Gram truncation tolerance */
   double lower;     /* lower bound of the spectrum (damped interval start) */
   double normA;     /* ||A||_2, 0: use the largest Ritz value */
   int refine;       /* CSR operator only.  1: precision ladder -- once the wanted
                        residuals fall below refine_tol or have not halved in three
                        iterations, the growth blocks continue in refine_dtype (an
                        FP64 refinement also uses the FP64 operator values); 0: off */
   int refine_dtype; /* ofrr_dtype of the refinement format */
   double refine_tol;
   int lock;         /* CSR operator only.  1: soft locking -- when the leading wanted
                        pairs 0..p-1 meet tol, the sweep filters only the trailing
                        columns: width the smallest power of two >= k - p, halved from
                        k, at least 16 (the Ritz block keeps all k columns, so locked
                        pairs stay in every Rayleigh-Ritz); 0: every sweep filters k */
   int mid_cols;     /* nd = 2 only.  0: the mid-filter block enters the basis whole;
                        w > 0: only its trailing w columns do (the ones nearest the
                        block edge), so the basis dimension is k + ka + min(w, ka)
                        with ka the filtered width -- a Ritz block wider than nwant
                        (guard columns) at a given basis dimension, e.g. k = 128,
                        w = 44 for 100 pairs in a basis of 300.  Every Ritz column
                        still gets its full correction block D_d. */
   int graph;        /* CSR operator only.  1: each filter sweep (its degree - 1
                        Chebyshev steps) is recorded once as a HIP graph and replayed
                        while its shape repeats (the context keeps the graphs; a sweep
                        of the same shape on other buffers updates the graph's
                        arguments); results are the same as with plain launches */
   int rr_whole;     /* 1: every Rayleigh-Ritz eigendecomposes the Gram matrix of the
                        whole basis [X | D] (ofrr_rr_truncated).  0 (default): after
                        the first, the split form -- the Ritz block X is whitened by
                        the Cholesky factor of X^T X and only the Gram matrix of D with
                        its X component removed is eigendecomposed (size m - k), then
                        the D block of the projected matrix before the coupled one; the
                        whole form is used where X^T X is not positive definite */
   int wide_dtype;   /* CSR operator only, with wide_tol > 0: every sweep after the
                        first whose largest wanted residual exceeds wide_tol builds its
                        growth blocks in wide_dtype, a format wider than basis_dtype
                        (e.g. FP32 for FP16 blocks); the other sweeps use basis_dtype.
                        FP16 rounding of the correction stalls the second sweep, whose
                        residual block is still large against its useful part */
   double wide_tol;  /* 0: off */
} ofrr_filter_opts;

typedef struct {
   int iterations;       /* outer iterations run */
   int converged;        /* pairs meeting tol among the nwant */
   int spmm_calls;       /* operator applications on k-column blocks */
   int rank_min;         /* smallest rank kept by the truncation */
   int filter_steps;     /* Chebyshev step launches (one kernel each) */
   int resid_passes;     /* FP64 residual launches */
   double t_total_ms;    /* device time of the solve (events on the stream) */
   double t_filter_ms;   /* Chebyshev steps only */
   double t_zspmm_ms;    /* operator applied to the growth blocks for V^T A V */
   double t_gram_ms;     /* projection products on the solver's stream (all of them,
                            or G = V^T V alone when t_aux_ms > 0) */
   double t_rr_ms;       /* Rayleigh-Ritz */
   double t_update_ms;   /* back-transform */
   double t_resid_ms;    /* FP64 operator application + residual (incl. host checks) */
   int filter_col_steps; /* sum over the Chebyshev steps of the columns filtered */
   int filter_first_cols;/* sum over the sweeps of the columns of the first step */
   int graph_replays;    /* sweeps launched as one captured graph (opts.graph) */
   int graph_captures;   /* of those, sweeps recorded anew (first of a shape, or new buffers) */
   int rr_fallbacks;     /* split Rayleigh-Ritz steps redone in the whole form (X^T X
                            not positive definite) */
   int wide_sweeps;      /* sweeps built in opts.wide_dtype */
   double t_aux_ms;      /* Z = A D and the H = V^T A V products on the context's second
                            stream, run beside the Rayleigh-Ritz (overlaps t_rr_ms; 0
                            when the products all run on the solver's stream) */
} ofrr_filter_stats;

/* X_dev: n x k FP64 row-major (row stride k), start block in / Ritz block out;
 * theta_host, res_host: k values out (descending Ritz values and their
 * residuals).  vals_lo_dev: the operator values in lo_dtype for the growth
 * SpMMs (NULL: use vals_dev).  Synchronous.  Returns 0 when the nwant pairs
 * converged, 4001 otherwise. */
OFRR_API int ofrr_filtered_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *row_ptr_dev,
                               const int32_t *col_idx_dev, const double *vals_dev,
                               const void *vals_lo_dev, int lo_dtype, const ofrr_filter_opts *opt,
                               double *X_dev, double *theta_host, double *res_host,
                               ofrr_filter_stats *stats);

/* ---- dense operators -------------------------------------------------------
 * C (R x nc, row-major, row stride ldc) = alpha P Q + beta C, with P R x K
 * row-major (row stride ldp) and Q K x nc row-major (row stride ldq); nc a
 * multiple of 16, <= 128.  Types: (BF16, BF16) or (FP16, FP16) on MFMA with FP32
 * accumulation (P rows 16-byte aligned); (FP32 | FP64, FP64) on FP64 MFMA.  Any
 * output type but FP8.  A X for column-major A is P = A^T row-major: pass the
 * row-major copy; A^T Y is P = A column-major with ldp = lda.
 * Replaces gemm(transA = false / true) over dense Matrix<T> operands
 * (CPP/linalg/blas/mvops.hpp:79-82, mvops.cu:423-605; gemm.m:27-59). */
OFRR_API int ofrr_dense_gemm(ofrr_ctx *ctx, int64_t R, int64_t K, int nc, const void *P_dev,
                             int p_dtype, int64_t ldp, const void *Q_dev, int q_dtype, int64_t ldq,
                             double alpha, double beta, void *C_dev, int c_dtype, int64_t ldc);

/* C = alpha op(A) op(B) + beta C, column-major BLAS form (gemm<T_I, T_O,
 * T_COMPUTE> of CPP/linalg/blas/mvops.hpp:79-82, mvops.cu:423-605, any sizes and
 * transposes): op(A) m x k, op(B) k x n, C m x n with leading dimensions lda,
 * ldb, ldc.  A and B share a format (FP64, FP32, FP16, BF16); C any of these.
 * compute_dtype OFRR_F64 accumulates in FP64, OFRR_F32 / OFRR_F16 / OFRR_BF16 in
 * FP32 (an FP16 compute rounds once at the end).  Fixed summation order.
 * beta = 0 does not read C. */
OFRR_API int ofrr_gemm(ofrr_ctx *ctx, int transA, int transB, int64_t m, int64_t n, int64_t k,
                       double alpha, const void *A_dev, int a_dtype, int64_t lda, const void *B_dev,
                       int b_dtype, int64_t ldb, double beta, void *C_dev, int c_dtype, int64_t ldc,
                       int compute_dtype);
/* X = alpha X + shift element-wise over a rows x cols block (element (i, j) at
 * X[i * rs + j * cs]; FP64, FP32, FP16, BF16): Vector / Matrix::scale
 * (CPP/containers/vector.hpp:140, matrix.hpp:145) with shift = 0, and the map of
 * U[0,1) draws to [lo, hi) for fill_random. */
OFRR_API int ofrr_axpb(ofrr_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X_dev, int64_t rs,
                       int64_t cs, double alpha, double shift);

/* Y = X converted element-wise to another format / layout: element (i, j) is
 * X[i * xrs + j * xcs] and Y[i * yrs + j * ycs] (Matrix::cast, matrix.cu:178-187,
 * without the host round trip). */
OFRR_API int ofrr_convert(ofrr_ctx *ctx, int64_t rows, int64_t cols, const void *X_dev, int x_dtype,
                          int64_t xrs, int64_t xcs, void *Y_dev, int y_dtype, int64_t yrs,
                          int64_t ycs);

/* Filtered OFRR on a dense symmetric FP64 operator A (n x n, column-major, lda):
 * the si_eig(const Matrix<T_I>& A, ...) boundary over a dense A
 * (CPP/linalg/solvers/eig/si_eig.hpp:24-27).  Same options, outputs and status as
 * ofrr_filtered_eig; k a power of two in [16, 128]; basis FP64, FP16 or BF16 (the
 * growth products then use a 16-bit copy of A on MFMA). */
OFRR_API int ofrr_filtered_eig_dense(ofrr_ctx *ctx, int64_t n, const double *A_dev, int64_t lda,
                                     const ofrr_filter_opts *opt, double *X_dev,
                                     double *theta_host, double *res_host,
                                     ofrr_filter_stats *stats);

/* Truncated SVD of a dense M x N FP32 A (column-major, lda) by filtered OFRR on
 * the implicit A^T A: the si_svd boundary (CPP/linalg/solvers/svd/si_svd.hpp,
 * subspace_iter_svd.m).  V_dev (N x k row-major FP64): start block in, right
 * singular vectors out; U_dev (M x k row-major FP64, nullable): left singular
 * vectors A v / sigma; sigma_host descending; res_host: ||A^T A v - sigma^2 v|| /
 * (sigma_1^2 ||v||).  Basis BF16 (growth products A X, A^T Y as BF16 MFMA GEMMs
 * with FP32 accumulation) or FP64; the residual products use the FP32 A with
 * FP64 accumulation. */
OFRR_API int ofrr_svd_filtered(ofrr_ctx *ctx, int64_t M, int64_t N, const float *A_dev, int64_t lda,
                               const ofrr_filter_opts *opt, double *V_dev, double *U_dev,
                               double *sigma_host, double *res_host, ofrr_filter_stats *stats);

/* Two-sided Rayleigh-Ritz for the SVD (svd_rr.m:54-91; the rr_svd boundary,
 * CPP/linalg/solvers/svd/rr_svd.hpp) over a dense FP64 A (M x N, column-major):
 * B = U^T A V; eig([0 B; B^T 0], blkdiag(U^T U, V^T V) + gtol I) when generalized,
 * else eig([0 B; B^T 0]) (the SVD of B for orthonormal U, V); sigma_host (min(ku, kv),
 * descending); U (M x ku) and V (N x kv), column-major, are replaced in their first
 * min(ku, kv) columns by sqrt(2) U Z and sqrt(2) V W ((Z; W) the eigenvectors).
 * ku + kv <= 256.  Synchronous. */
OFRR_API int ofrr_svd_rr(ofrr_ctx *ctx, int64_t M, int64_t N, const double *A_dev, int64_t lda,
                         int ku, int kv, double *U_dev, int64_t ldu, double *V_dev, int64_t ldv,
                         int generalized, double gtol, double *sigma_host);
/* Subspace iteration for the SVD (subspace_iter_svd.m:57-77; the si_svd boundary,
 * CPP/linalg/solvers/svd/si_svd.hpp:25-28) in FP64 with function_orth = the
 * Hessenberg process (select_tol, kept columns only) and function_rr = ofrr_svd_rr:
 * V0 (N x k) start block; each iteration step_size x (U = A V, scalecols; V = A^T U,
 * scalecols), then the RR of the two Hessenberg bases.  U (M x k), V (N x k) and
 * sigma_host (k, NaN past the kept width *kept) out.  2 k <= 256.  Synchronous. */
OFRR_API int ofrr_subspace_iter_svd(ofrr_ctx *ctx, int64_t M, int64_t N, const double *A_dev,
                                    int64_t lda, int k, int maxiter, int step_size,
                                    double select_tol, int generalized, double gtol,
                                    const double *V0_dev, int64_t ldv0, double *U_dev, int64_t ldu,
                                    double *V_dev, int64_t ldv, double *sigma_host, int *kept);

/* Diagnostic (no counterpart in the reference): counts where gfx950's FP8
 * conversion instructions disagree with the OCP FP8 E4M3 definition the library
 * stores (round to nearest even, saturating at 448): counts_host[0] decodes of
 * the 256 codes, [1] encodes of floats up to 448 in magnitude, [2] above 448. */
OFRR_API int ofrr_fp8_check(ofrr_ctx *ctx, int64_t counts_host[3]);

/* ---- matrix input (host only): the binary CSR files read by
 * MatrixReader::read_csr_matrix (CPP/io/matrix_reader.cpp:8-45):
 * <prefix>_row_ptr.bin (int32, rows + 1), <prefix>_col_ind.bin (int32, 0-based),
 * <prefix>_values.bin (float64); cols = max column index + 1.  The structure is
 * validated; a missing or empty file gives 1000 and the reference's message
 * ("Failed to read row pointers from <file>", ...) in ofrr_io_last_error(). */
OFRR_API int ofrr_csr_file_info(const char *prefix, int64_t *rows, int64_t *cols, int64_t *nnz);
/* row_ptr (rows + 1, widened to int64), col_idx (nnz), vals (nnz): host arrays */
OFRR_API int ofrr_csr_file_read(const char *prefix, int64_t *row_ptr, int32_t *col_idx,
                                double *vals);
/* <prefix>_singular_values.bin (float64; MatrixReader::read_singular_values,
 * matrix_reader.cpp:47-50): *count values, the first min(cap, count) into out */
OFRR_API int ofrr_singular_values_file(const char *prefix, double *out, int64_t cap,
                                       int64_t *count);
/* message of the last failed file call on this thread */
OFRR_API const char *ofrr_io_last_error(void);

/* Device bytes ofrr_filtered_eig holds for its solver state (Ritz, residual,
 * filter and projected blocks) at n rows and the given options; the solve
 * takes them from a grow-only per-thread arena. */
OFRR_API int ofrr_filtered_eig_state_bytes(int64_t n, const ofrr_filter_opts *opt, int64_t *bytes);

/* ---- serving: continuous batching of eigen requests on one GPU ---------------
 * No counterpart in the reference (the caller of the path: a request loop that
 * feeds ofrr_filtered_eig).  A server runs `slots` solves at once, each on its
 * own host worker thread, context and non-blocking HIP stream, so the kernels of
 * different solves overlap on the device.  Requests queue in arrival order; the
 * head of the queue is admitted when a slot is free and its device footprint
 * (operator + start block + solver state, ofrr_request_footprint) fits in the
 * byte budget beside the solves already running.  A request's memory comes
 * from the device's stream-ordered pool at admission and returns to it on
 * completion.  Every reduction of a solve has a fixed order and no solve
 * shares device memory with another, so a request's result does not depend on
 * what runs beside it. */
typedef struct ofrr_server ofrr_server;

typedef struct {
   int slots;          /* concurrent solves (0: 4) */
   int64_t budget;     /* device bytes the running solves may hold together
                          (0: 90% of the free device memory at creation) */
   int graphs;         /* 1: filter sweeps as captured HIP graphs (ofrr_filter_opts.graph),
                          kept per slot and shape across requests; 0 or -1: plain launches
                          (the default: on ROCm 7.2 a replayed sweep measured slower per step
                          than its launches) */
   int rr_group_us;    /* grouped Rayleigh-Ritz: the truncated Rayleigh-Ritz jobs of the
                          running solves that meet run as one batched launch per stage on
                          the stream of the solve that formed the group (each job's result
                          is the one it gets alone).  0: the jobs already handed in when a group is formed;
                          w > 0: a group also waits up to w microseconds for every running
                          solve's job; -1: every solve runs its own */
} ofrr_server_opts;

#define OFRR_SERVE_MAX_WANT 128
typedef struct {
   int dim;            /* 2: 5-point, 3: 7-point diffusion stencil */
   int64_t N;          /* grid side; n = N^dim */
   uint64_t coef_seed; /* face coefficients 10^(2u-1) (ofrr_stencil_csr); 0: Laplacian */
   uint64_t start_seed;/* Philox seed of the N(0,1) start block */
   int nwant;          /* largest eigenpairs wanted, 1..OFRR_SERVE_MAX_WANT */
   int basis_dtype;    /* growth-block format (OFRR_F16, OFRR_BF16, OFRR_FP8, OFRR_F32, OFRR_F64) */
   double tol;         /* ||A x - theta x|| / (||A|| ||x||) */
   int degree;         /* filter degree per outer iteration (0: by size) */
   int maxiter;        /* outer iterations (0: 200) */
   int k;              /* Ritz block width, nwant..128 (0: the smallest multiple of 16
                          >= 1.25 nwant, at least 16; stored padded to a power of two) */
   int nd;             /* correction blocks, 1 or 2 (0: 2 while the basis dimension
                          k (1 + nd) stays <= 192, else 1); basis dimension <= 384 */
   int refine;         /* precision ladder: 0 default (FP8 growth finishes in FP16 below
                          1e-4), 1 use refine_dtype / refine_tol, -1 off */
   int refine_dtype;   /* refinement format when refine = 1 (OFRR_F64: the growth
                          products then also use the FP64 operator values) */
   double refine_tol;  /* switch once the wanted residuals are below it (or stall) */
   int mid_cols;       /* ofrr_filter_opts.mid_cols (0: the whole mid-filter block);
                          basis dimension k + k + mid_cols with nd = 2 */
} ofrr_eig_request;

typedef struct {
   int status;         /* ofrr_filtered_eig's status for the solve */
   int k, nd;          /* block width and correction blocks the solve used */
   int iterations, converged, filter_steps;
   int64_t n, footprint;             /* rows; device bytes held while running */
   double t_submit, t_admit, t_done; /* seconds since the server was created */
   double theta[OFRR_SERVE_MAX_WANT];/* descending Ritz values (nwant) */
   double res[OFRR_SERVE_MAX_WANT];  /* their residuals */
} ofrr_eig_result;

typedef struct {
   int64_t submitted, completed, running, queued;
   int64_t bytes_running, bytes_peak; /* footprint of the running solves, and its maximum */
   int running_peak;                  /* most solves running at once */
   int64_t budget;
   int64_t slot_bytes;                /* workspace the slots hold for the server's life
                                         (admission: bytes_running + footprint <=
                                         budget - slot_bytes) */
   int64_t ws_regrowths;              /* solves whose workspace outgrew the slot's
                                         reservation (counted into slot_bytes) */
   int64_t rr_groups, rr_group_jobs;  /* grouped Rayleigh-Ritz launches and the jobs in them */
   int rr_group_max;                  /* most jobs in one group */
} ofrr_server_stats;

/* the plan of a request (block width k, correction blocks nd) and its device
 * footprint in bytes; host only */
OFRR_API int ofrr_request_footprint(const ofrr_eig_request *req, int *k, int *nd, int64_t *bytes);
/* device bytes each slot of a server on `device` reserves for its whole life
 * (the largest workspace a solve asks for); a server of s slots admits requests
 * against budget - s * this.  Host only (reads the device's CU count). */
OFRR_API int ofrr_server_slot_bytes(int device, int64_t *bytes);
OFRR_API int ofrr_server_create(int device, const ofrr_server_opts *opt, ofrr_server **srv);
/* enqueue; *id identifies the request.  1002 when its footprint exceeds the budget */
OFRR_API int ofrr_server_submit(ofrr_server *srv, const ofrr_eig_request *req, int64_t *id);
/* 0 and *out filled when request id is done (the server then forgets it),
 * OFRR_PENDING while it is queued or running */
#define OFRR_PENDING 1
OFRR_API int ofrr_server_poll(ofrr_server *srv, int64_t id, ofrr_eig_result *out);
/* blocks until request id is done; fills *out and forgets it */
OFRR_API int ofrr_server_wait(ofrr_server *srv, int64_t id, ofrr_eig_result *out);
OFRR_API int ofrr_server_stats_get(ofrr_server *srv, ofrr_server_stats *out);
/* finishes the queued and running requests, then stops the workers */
OFRR_API int ofrr_server_destroy(ofrr_server *srv);

#ifdef __cplusplus
}
#endif
#endif
