/*
 * ofrr_oracle.h -- CPU FP64 restatement of the OFRR projection pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (mixedprecisionofrr_amd/) never links, loads or calls it.
 *
 * Each function cites the reference file:line whose semantics it restates
 * (paths relative to the reference repository root).  Every function is
 * pinned by a check in tests/test_oracle_*.py; see the table at the top of
 * ofrr_oracle.c for which check pins which function.
 *
 * Conventions: column-major dense matrices with explicit leading dimension,
 * 0-based CSR with int32 row_ptr/col_idx (int64 row_ptr for the large
 * generators), FP64 everywhere.
 */
#ifndef OFRR_ORACLE_H
#define OFRR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* out[i] = N(0,1) number (offset+i) of stream (seed, stream) */
void orc_randn(uint64_t seed, uint32_t stream, uint64_t offset, int64_t n, double *out);
/* out[i] = U[0,1) number (offset+i) of stream (seed, stream) */
void orc_randu(uint64_t seed, uint32_t stream, uint64_t offset, int64_t n, double *out);

/* ---- CSR stencil operators (natural x-fastest ordering, Dirichlet) ------- */
/* number of nonzeros of the dim-D (2 or 3) stencil on an N^D grid */
int64_t orc_stencil_nnz(int dim, int64_t N);
/* seed == 0 -> constant coefficients (the 2D 5-point / 3D 7-point Laplacian);
 * otherwise face coefficients 10^(2u-1), u ~ U[0,1) from Philox stream d. */
void orc_stencil_csr(int dim, int64_t N, uint64_t seed, int64_t *row_ptr, int32_t *col_idx,
                     double *vals);

/* ---- operator and BLAS-like kernels -------------------------------------- */
/* Y = A X (csrmv.m:48-59 semantics, FP64, fused multiply-add in row order) */
void orc_spmm_csr(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
                  int k, const double *X, int64_t ldx, double *Y, int64_t ldy);
/* Y = A^T X for an nrows x ncols CSR A: the 'T' branch of csrmv.m:61-71 --
 * Y = 0, then for each row i in increasing order and each stored entry a_ij,
 * Y(j,:) += a_ij X(i,:) (fused multiply-add); X nrows x k, Y ncols x k */
void orc_spmm_csr_t(int64_t nrows, int64_t ncols, const int64_t *row_ptr, const int32_t *col_idx,
                    const double *vals, int k, const double *X, int64_t ldx, double *Y, int64_t ldy);
/* divide each column by its signed max-abs entry (scalecols.m:26-33) */
void orc_scalecols(int64_t m, int n, double *A, int64_t lda);
/* 1-based index of the first max |x|, signed value (mvops.cu:607-946) */
int orc_iamax(int64_t n, const double *x, double *val);
/* 0-based per-column argmax |A(:,j)| (mvops.cu:955-1199) */
void orc_matrix_iamax(int64_t m, int n, const double *A, int64_t lda, int *idx);
/* G = V^T W, G is k1 x k2 (gemm transA=true, mvops.cu:423) */
void orc_gemm_tn(int64_t n, int k1, int k2, const double *V, int64_t ldv, const double *W,
                 int64_t ldw, double *G, int ldg);
/* C = A B, A m x k, B k x n (gemm transA=false) */
void orc_gemm_nn(int64_t m, int n, int k, const double *A, int64_t lda, const double *B,
                 int64_t ldb, double *C, int64_t ldc);

/* ---- Hessenberg process (hessenberg.m:55-74), in place on Q ---------------
 * Q: m x n in (the matrix A) / out (the basis Q); R: n x n (zeroed here);
 * select[i] = 1 kept, 0 dropped; perm[i] = 0-based pivot row (-1 if dropped).
 * variant 0: skip dropped columns in later eliminations (MATLAB, C++ v1)
 * variant 1: C++ v3, dropped columns still eliminated with their argmax row and
 *            unscaled residual (hessenberg.cu:768-779)
 */
void orc_hessenberg(int64_t m, int n, double *Q, int64_t ldq, double *R, int ldr,
                    double select_tol, int variant, int *select, int *perm);

/* Gram-Schmidt QR (CPP/linalg/factorization/qr.cu, MATLAB/+src/+qr/mgs.m,
 * cgs.m).  variant 0 MGS (reorth_tol > 0: reorthogonalization), 1 MGS_V2,
 * 2 CGS (reorth_tol > 0: iterated, cgs.m:73-89), 3 CGS2 (second projection
 * subtracted).  Q holds A on entry.  skip[i] = 1 for dropped columns, which keep
 * their residual in Q and are left out of later projections; R(i,i) of a dropped
 * column is 0 (MGS) or 1 (the others). */
void orc_gs_qr(int64_t m, int n, double *Q, int64_t ldq, double *R, int ldr, int variant,
               double select_tol, double orth_tol, double reorth_tol, int *skip);

/* ---- dense symmetric eigensolvers (FP64) --------------------------------- */
/* eigenvalues ascending in w, eigenvectors in V (LAPACK dsyev 'V','U' order);
 * returns the number of Jacobi sweeps run */
int orc_syev(int n, const double *A, int lda, double *w, double *V, int ldv);
/* A x = lambda B x, itype 1, B-normalised vectors (LAPACK dsygv order).
 * returns 0, or i>0 if the leading minor of order i of B is not positive */
int orc_sygv(int n, const double *A, int lda, const double *B, int ldb, double *w, double *V,
             int ldv);
/* truncated OFRR Rayleigh-Ritz: G = U L U^T, keep L_i > tau * max(L),
 * B = S U_r L_r^{-1/2} with S = diag(G)^{-1/2} scaling, eig(B^T H B) -> top nkeep.
 * theta descending, Y (m x nkeep) coefficient vectors. returns the rank kept. */
int orc_rr_truncated(int m, const double *G, int ldg, const double *H, int ldh, double tau,
                     int nkeep, double *theta, double *Y, int ldy);
/* eig_rr.m:48-57 generalised form: eig(H, G + gtol I), sorted descending */
/* Split truncated Rayleigh-Ritz over V = [X | D], X the first kx columns:
 * G_XX = L L^T, S = G_DD - G_XD^T G_XX^{-1} G_XD scaled to unit diagonal,
 * eig(S) truncated at tau max(1, max eig) -> whitened basis W; eig of the D block
 * of W^T H W first, then of the coupled matrix.  theta descending, Y (m x nkeep).
 * Returns kx + the kept D rank, or -1 (G_XX not positive definite). */
int orc_rr_split(int m, int kx, const double *G, int ldg, const double *H, int ldh, double tau,
                 int nkeep, double *theta, double *Y, int ldy);
int orc_rr_generalized(int m, const double *G, int ldg, const double *H, int ldh, double gtol,
                       double *theta, double *Y, int ldy);

/* ---- drivers -------------------------------------------------------------- */
/* arnoldi_hess_eig.m:43-107: Hessenberg-type Krylov builder + generalised RR.
 * v0: start vector (n).  Outputs the top k Ritz values (descending) and X (n x k).
 * returns the number of basis columns built in the last sweep. */
int orc_arnoldi_hess_eig(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                         const double *vals, int k, int kdim, int maxiter, double select_tol,
                         double gtol, const double *v0, double *theta, double *X, int64_t ldx);

/* subspace_iter_eig.m:53-68 with function_orth = hessenberg (select_tol) and
 * function_rr = eig_rr(generalised, gtol).  V (n x m) start block in, Ritz block out. */
void orc_subspace_iter_eig(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                           const double *vals, int m, int maxiter, int step_size,
                           double select_tol, double gtol, double *V, int64_t ldv, double *theta);

/* svd_rr.m:54-91 over a dense column-major A (m x n): the augmented pencil
 * (generalised: blkdiag(U^T U, V^T V) + gtol I, else I); sigma (kk), U, V in/out
 * (first kk = min(ku, kv) columns written, scaled by sqrt 2).  Returns kk. */
int orc_svd_rr(int64_t m, int64_t n, const double *A, int64_t lda, int ku, int kv, double *U,
               int64_t ldu, double *V, int64_t ldv, int generalized, double gtol, double *sigma);
/* subspace_iter_svd.m:57-77 with hessenberg orth and svd_rr; returns the kept width */
int orc_subspace_iter_svd(int64_t m, int64_t n, const double *A, int64_t lda, int k, int maxiter,
                          int step_size, double select_tol, int generalized, double gtol,
                          const double *V0, double *U, int64_t ldu, double *V, int64_t ldv,
                          double *sigma);

/* Filtered OFRR (block subspace iteration with a Chebyshev filter on the
 * residual correction, FP64 throughout) -- the algorithm the FP64 GPU path runs.
 * X (n x k) start block in / Ritz block out; theta (k) out descending;
 * res (k) out relative residuals; returns iterations run (negative if not converged). */
typedef struct {
   int k;            /* block width */
   int nwant;        /* wanted pairs (largest) */
   int nd;           /* correction blocks in the basis: 1 or 2 */
   int degree;       /* filter degree per outer iteration */
   int maxiter;      /* outer iterations */
   double tol;       /* residual tolerance ||Ax - theta x|| / (||A|| ||x||) */
   double tau;       /* Gram truncation tolerance */
   double lower;     /* lower bound of the spectrum */
   double normA;     /* ||A||_2 (0: use the largest Ritz value) */
   int lock;         /* 1: soft locking (a sweep filters the columns after the
                        leading converged wanted pairs only); 0: all k columns */
   int rr_whole;     /* 1: every Rayleigh-Ritz eigendecomposes the Gram matrix of the
                        whole basis (orc_rr_truncated); 0: the split form
                        (orc_rr_split, X whitened by Cholesky), the whole form only
                        where G_XX is not positive definite */
} orc_ofrr_opts;
int orc_filtered_ofrr(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                      const double *vals, const orc_ofrr_opts *opt, double *X, int64_t ldx,
                      double *theta, double *res);

/* r_j = ||A x_j - theta_j x_j||_2 / (normA ||x_j||_2)  (test4/test.m:340-350, normalised) */
void orc_residuals(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
                   int k, const double *X, int64_t ldx, const double *theta, double normA,
                   double *res);

#ifdef __cplusplus
}
#endif
#endif
