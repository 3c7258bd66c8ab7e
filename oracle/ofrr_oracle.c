/*
 * ofrr_oracle.c -- CPU FP64 restatement of the OFRR projection pipeline.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Nothing in the product links it.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   orc_philox4x32_10   Random123 published known-answer vectors (Philox4x32-10)
 *   orc_randn/randu     moments + Kolmogorov-Smirnov against the normal/uniform laws
 *   orc_stencil_csr     closed-form nnz counts (64^2 grid 20,224 / 128^3 grid 14,581,760), symmetry,
 *                       sine-mode eigenvectors A s = lambda s (closed form), scipy kron build
 *   orc_spmm_csr        brute-force dense A X, integer-valued exact results
 *   orc_spmm_csr_t      bitwise orc_spmm_csr on scipy's row-stable transpose, brute-force
 *                       dense A^T X, integer-valued exact results, <A^T y, x> = <y, A x>
 *   orc_scalecols       hand cases incl. ties and zero columns
 *   orc_iamax           reference KATs mvops_test.cpp:774-850 (index 72345 -> -8.0, 1-based)
 *   orc_matrix_iamax    reference KATs mvops_test.cpp:1151-1497 (0-based)
 *   orc_gemm_tn/nn      reference gemm KAT mvops_test.cpp:455-721 and numpy on random data
 *   orc_hessenberg      A = QR, Q(perm(i),i) = 1, Q(perm(j),i) = 0 (j<i), |Q| <= 1, and the
 *                       duplicated-column skip pattern of hessenberg_test.cpp:73-91; with
 *                       dropped columns A = Q R~ (R~ = R with a unit diagonal where dropped)
 *                       for both variants, which differ exactly there
 *   orc_syev            reference KAT mvops_test.cpp:2292-2447 (tridiag(-1,2,-1), n = 10)
 *   orc_sygv            reference KAT mvops_test.cpp:2450-2626 (B = diag(1..10))
 *   orc_rr_truncated    exact Ritz values for sine-mode bases, planted rank drop
 *   orc_rr_generalized  scipy.linalg.eigh(H, G) on random SPD pencils
 *   orc_arnoldi_hess_eig exact termination: v0 = sum of j sine modes -> closed-form eigenvalues
 *   orc_subspace_iter_eig / orc_filtered_ofrr  closed-form Laplacian spectra, residual bound;
 *                       soft locking: bitwise equal to the unlocked solve while no pair has
 *                       converged, closed-form spectrum in the unlocked iteration count
 *   orc_residuals       hand-computed residual of exact eigenpairs (zero) and perturbed ones
 *   orc_gs_qr           numpy Householder QR (sign-fixed) on full-rank input; the textbook
 *                       loss of orthogonality of each variant on graded matrices (CGS ~ u k^2,
 *                       MGS ~ u k, reorthogonalized and CGS2 ~ u); the duplicated-column skip
 *                       pattern of qr_test.cpp:90-120; zero-column and drop conventions
 */
#include "ofrr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Philox4x32-10                                                              */
/* ------------------------------------------------------------------------- */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t *hi, uint32_t *lo) {
   uint64_t p = (uint64_t)a * (uint64_t)b;
   *hi = (uint32_t)(p >> 32);
   *lo = (uint32_t)p;
}

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
   const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
   const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
   uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
   uint32_t k0 = key_in[0], k1 = key_in[1];
   for (int r = 0; r < 10; ++r) {
      uint32_t hi0, lo0, hi1, lo1;
      mulhilo32(M0, c0, &hi0, &lo0);
      mulhilo32(M1, c2, &hi1, &lo1);
      uint32_t n0 = hi1 ^ c1 ^ k0;
      uint32_t n2 = hi0 ^ c3 ^ k1;
      c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
      k0 += W0; k1 += W1;
   }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* block p of stream (seed, stream): counter (p_lo, p_hi, stream, 0), key (seed_lo, seed_hi) */
static void philox_block(uint64_t seed, uint32_t stream, uint64_t p, uint32_t out[4]) {
   uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), stream, 0u};
   uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
   orc_philox4x32_10(ctr, key, out);
}

/* 53-bit uniform in [0,1) from two words */
static inline double u53(uint32_t a, uint32_t b) {
   return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

void orc_randn(uint64_t seed, uint32_t stream, uint64_t offset, int64_t n, double *out) {
   const double two_pi = 6.283185307179586476925286766559;
   for (int64_t i = 0; i < n; ++i) {
      uint64_t e = offset + (uint64_t)i;
      uint32_t w[4];
      philox_block(seed, stream, e >> 1, w);
      double u1 = u53(w[0], w[1]);
      double u2 = u53(w[2], w[3]);
      double r = sqrt(-2.0 * log(1.0 - u1)); /* 1-u1 in (0,1] */
      double a = two_pi * u2;
      out[i] = (e & 1) ? r * sin(a) : r * cos(a);
   }
}

void orc_randu(uint64_t seed, uint32_t stream, uint64_t offset, int64_t n, double *out) {
   for (int64_t i = 0; i < n; ++i) {
      uint64_t e = offset + (uint64_t)i;
      uint32_t w[4];
      philox_block(seed, stream, e >> 1, w);
      out[i] = (e & 1) ? u53(w[2], w[3]) : u53(w[0], w[1]);
   }
}

/* ------------------------------------------------------------------------- */
/* stencil operators                                                          */
/* ------------------------------------------------------------------------- */
int64_t orc_stencil_nnz(int dim, int64_t N) {
   if (dim == 2) return 5 * N * N - 4 * N;
   return 7 * N * N * N - 6 * N * N;
}

/* coefficient of the face between cells c-1 and c along direction d (c = 0..N),
 * other coordinates (o1, o2): index c + (N+1)*(o1 + N*o2) of stream d. */
static double face_coef(uint64_t seed, int d, int64_t N, int64_t c, int64_t o1, int64_t o2) {
   if (seed == 0) return 1.0;
   uint64_t idx = (uint64_t)(c + (N + 1) * (o1 + N * o2));
   double u;
   orc_randu(seed, (uint32_t)d, idx, 1, &u);
   return exp((2.0 * u - 1.0) * 2.302585092994045684017991454684);
}

void orc_stencil_csr(int dim, int64_t N, uint64_t seed, int64_t *row_ptr, int32_t *col_idx,
                     double *vals) {
   int64_t n = (dim == 2) ? N * N : N * N * N;
   int64_t pos = 0;
   for (int64_t p = 0; p < n; ++p) {
      int64_t x = p % N, y = (p / N) % N, z = (dim == 3) ? p / (N * N) : 0;
      /* faces: lower/upper along x, y, z; the "other" coordinates per direction */
      double kxl = face_coef(seed, 0, N, x, y, z), kxu = face_coef(seed, 0, N, x + 1, y, z);
      double kyl = face_coef(seed, 1, N, y, x, z), kyu = face_coef(seed, 1, N, y + 1, x, z);
      double kzl = 0, kzu = 0;
      if (dim == 3) {
         kzl = face_coef(seed, 2, N, z, x, y);
         kzu = face_coef(seed, 2, N, z + 1, x, y);
      }
      double diag = kxl + kxu + kyl + kyu + kzl + kzu;
      row_ptr[p] = pos;
      if (dim == 3 && z > 0) { col_idx[pos] = (int32_t)(p - N * N); vals[pos++] = -kzl; }
      if (y > 0) { col_idx[pos] = (int32_t)(p - N); vals[pos++] = -kyl; }
      if (x > 0) { col_idx[pos] = (int32_t)(p - 1); vals[pos++] = -kxl; }
      col_idx[pos] = (int32_t)p; vals[pos++] = diag;
      if (x < N - 1) { col_idx[pos] = (int32_t)(p + 1); vals[pos++] = -kxu; }
      if (y < N - 1) { col_idx[pos] = (int32_t)(p + N); vals[pos++] = -kyu; }
      if (dim == 3 && z < N - 1) { col_idx[pos] = (int32_t)(p + N * N); vals[pos++] = -kzu; }
   }
   row_ptr[n] = pos;
}

/* ------------------------------------------------------------------------- */
/* kernels                                                                    */
/* ------------------------------------------------------------------------- */
void orc_spmm_csr(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
                  int k, const double *X, int64_t ldx, double *Y, int64_t ldy) {
#pragma omp parallel for schedule(static)
   for (int64_t i = 0; i < n; ++i) {
      for (int j = 0; j < k; ++j) {
         double acc = 0.0;
         for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
            acc = fma(vals[p], X[(int64_t)col_idx[p] + j * ldx], acc);
         Y[i + j * ldy] = acc;
      }
   }
}

void orc_spmm_csr_t(int64_t nrows, int64_t ncols, const int64_t *row_ptr, const int32_t *col_idx,
                    const double *vals, int k, const double *X, int64_t ldx, double *Y, int64_t ldy) {
   /* columns of X are independent: the scatter order within a column is the
    * reference's row order */
#pragma omp parallel for schedule(static)
   for (int j = 0; j < k; ++j) {
      double *y = Y + (int64_t)j * ldy;
      const double *x = X + (int64_t)j * ldx;
      for (int64_t c = 0; c < ncols; ++c) y[c] = 0.0;
      for (int64_t i = 0; i < nrows; ++i)
         for (int64_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
            y[col_idx[p]] = fma(vals[p], x[i], y[col_idx[p]]);
   }
}

int orc_iamax(int64_t n, const double *x, double *val) {
   int64_t best = 0;
   double bv = -1.0;
   for (int64_t i = 0; i < n; ++i) {
      double a = fabs(x[i]);
      if (a > bv) { bv = a; best = i; } /* strict: first index on ties */
   }
   if (val) *val = (n > 0) ? x[best] : 0.0;
   return (int)(best + 1);
}

void orc_matrix_iamax(int64_t m, int n, const double *A, int64_t lda, int *idx) {
   for (int j = 0; j < n; ++j) idx[j] = orc_iamax(m, A + (int64_t)j * lda, NULL) - 1;
}

void orc_scalecols(int64_t m, int n, double *A, int64_t lda) {
   for (int j = 0; j < n; ++j) {
      double *a = A + (int64_t)j * lda;
      double t;
      orc_iamax(m, a, &t);
      if (t != 0.0)
         for (int64_t i = 0; i < m; ++i) a[i] = a[i] / t;
   }
}

/* G = V^T W with a fixed summation order: rows split into chunks (the split the
 * device Gram kernel uses: about 2048 / (tiles of 64 x 64) chunks of a multiple
 * of 4 rows, at least 64 rows); within a chunk every entry accumulates its rows
 * in order with fused multiply-adds; the chunk partials are added in 8
 * consecutive groups (each in chunk order) and the group sums in order.  The
 * result does not depend on the thread count. */
void orc_gemm_tn(int64_t n, int k1, int k2, const double *V, int64_t ldv, const double *W,
                 int64_t ldw, double *G, int ldg) {
   enum { SLAB = 32 };
   const int64_t ti = (k1 + 63) / 64, tj = (k2 + 63) / 64;
   int64_t want = 2048 / (ti * tj);
   if (want < 1) want = 1;
   int64_t rows = (n + want - 1) / want;
   rows = ((rows + 3) / 4) * 4;
   if (rows < 64) rows = 64;
   const int64_t nch = n > 0 ? (n + rows - 1) / rows : 0;
   double *part = (double *)calloc((size_t)(nch ? nch : 1) * k1 * k2, sizeof(double));
#pragma omp parallel
   {
      double *vb = (double *)malloc(sizeof(double) * SLAB * (size_t)k1);
      double *wb = (double *)malloc(sizeof(double) * SLAB * (size_t)k2);
#pragma omp for schedule(dynamic, 1)
      for (int64_t c = 0; c < nch; ++c) {
         double *g = part + (size_t)c * k1 * k2;
         const int64_t r0 = c * rows, r1 = (r0 + rows < n) ? r0 + rows : n;
         for (int64_t rb = r0; rb < r1; rb += SLAB) {
            int nr = (int)((r1 - rb) < SLAB ? (r1 - rb) : SLAB);
            for (int a = 0; a < k1; ++a)
               for (int r = 0; r < nr; ++r) vb[r * k1 + a] = V[rb + r + (int64_t)a * ldv];
            for (int b2 = 0; b2 < k2; ++b2)
               for (int r = 0; r < nr; ++r) wb[r * k2 + b2] = W[rb + r + (int64_t)b2 * ldw];
            for (int r = 0; r < nr; ++r)
               for (int a = 0; a < k1; ++a) {
                  const double va = vb[r * k1 + a];
                  double *ga = g + (size_t)a * k2;
                  const double *wr = wb + (size_t)r * k2;
                  for (int b2 = 0; b2 < k2; ++b2) ga[b2] = fma(va, wr[b2], ga[b2]);
               }
         }
      }
      free(vb);
      free(wb);
   }
   const int64_t gsz = (nch + 7) / 8;
   for (int a = 0; a < k1; ++a)
      for (int b2 = 0; b2 < k2; ++b2) {
         double t = 0.0;
         for (int g = 0; g < 8; ++g) {
            double s = 0.0;
            const int64_t c0 = g * gsz, c1 = (c0 + gsz < nch) ? c0 + gsz : nch;
            for (int64_t c = c0; c < c1; ++c) s += part[(size_t)c * k1 * k2 + (size_t)a * k2 + b2];
            t += s;
         }
         G[a + (int64_t)b2 * ldg] = t;
      }
   free(part);
}

void orc_gemm_nn(int64_t m, int n, int k, const double *A, int64_t lda, const double *B,
                 int64_t ldb, double *C, int64_t ldc) {
#pragma omp parallel for schedule(static)
   for (int64_t i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j) {
         double acc = 0.0;
         for (int p = 0; p < k; ++p) acc = fma(A[i + p * lda], B[p + j * ldb], acc);
         C[i + j * ldc] = acc;
      }
}

/* ------------------------------------------------------------------------- */
/* Hessenberg process (hessenberg.m:55-74; C++ v1 keeps the unscaled residual */
/* in a dropped column, hessenberg.cu:515-522)                                 */
/* ------------------------------------------------------------------------- */
void orc_hessenberg(int64_t m, int n, double *Q, int64_t ldq, double *R, int ldr,
                    double select_tol, int variant, int *select, int *perm) {
   /* the argmax row of every column, dropped ones included: C++ v3 takes it
    * from iamax before the drop decision (hessenberg.cu:783-795) */
   int *prow = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) R[i + (int64_t)j * ldr] = 0.0;
   for (int i = 0; i < n; ++i) {
      double *v = Q + (int64_t)i * ldq;
      for (int j = 0; j < i; ++j) {
         /* variant 0 leaves dropped columns out (hessenberg.m:58, C++ v1
          * hessenberg.cu:520); variant 1 eliminates with every earlier column,
          * a dropped one with its unscaled residual (v3, hessenberg.cu:768-779) */
         if (!select[j] && variant == 0) continue;
         double c = v[prow[j]];
         R[j + (int64_t)i * ldr] = c;
         const double *q = Q + (int64_t)j * ldq;
         for (int64_t r = 0; r < m; ++r) v[r] = fma(-c, q[r], v[r]);
      }
      double piv;
      int k1 = orc_iamax(m, v, &piv);
      R[i + (int64_t)i * ldr] = piv;
      prow[i] = k1 - 1;
      if (fabs(piv) < select_tol) {
         select[i] = 0;
         perm[i] = -1;
      } else {
         select[i] = 1;
         perm[i] = k1 - 1;
         for (int64_t r = 0; r < m; ++r) v[r] = v[r] / piv;
      }
   }
   free(prow);
}

/* ------------------------------------------------------------------------- */
/* Gram-Schmidt QR.  MGS follows qr.cu:84-249 / mgs.m:55-92 (left-looking,     */
/* r = v^T q_j then v -= r q_j per kept column, reorthogonalization while      */
/* orth_tol <= t < reorth_tol * normv); MGS_V2 qr.cu:251-393 (right-looking);  */
/* CGS qr.cu:489-569 with the iterated passes of cgs.m:73-89; CGS2 the two     */
/* passes of qr.hpp:110-131 with the second projection subtracted.             */
/* ------------------------------------------------------------------------- */
static double o_dot(int64_t m, const double *x, const double *y) {
   double s = 0.0;
   for (int64_t r = 0; r < m; ++r) s = fma(x[r], y[r], s);
   return s;
}

void orc_gs_qr(int64_t m, int n, double *Q, int64_t ldq, double *R, int ldr, int variant,
               double select_tol, double orth_tol, double reorth_tol, int *skip) {
   for (int j = 0; j < n; ++j) {
      skip[j] = 0;
      for (int i = 0; i < n; ++i) R[i + (int64_t)j * ldr] = 0.0;
   }
#define RR(i, j) R[(i) + (int64_t)(j) * ldr]
   const double rd_skip = variant == 0 ? 0.0 : 1.0;
   if (variant == 1) {
      for (int i = 0; i < n; ++i) {
         double *v = Q + (int64_t)i * ldq;
         const double t = sqrt(o_dot(m, v, v));
         if (t < select_tol) {
            skip[i] = 1;
            RR(i, i) = 1.0;
            continue;
         }
         RR(i, i) = t;
         const double s = 1.0 / t;
         for (int64_t r = 0; r < m; ++r) v[r] = v[r] * s;
         for (int c = i + 1; c < n; ++c) {
            double *w = Q + (int64_t)c * ldq;
            const double rc = o_dot(m, v, w);
            RR(i, c) = rc;
            for (int64_t r = 0; r < m; ++r) w[r] = fma(-rc, v[r], w[r]);
         }
      }
      return;
   }
   double *rv = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
   const int reorth = reorth_tol > 0.0 && (variant == 0 || variant == 2);
   for (int i = 0; i < n; ++i) {
      double *v = Q + (int64_t)i * ldq;
      double normv = sqrt(o_dot(m, v, v));
      if (normv < select_tol) {
         skip[i] = 1;
         RR(i, i) = rd_skip;
         continue;
      }
      int pass = 0, reorths = 0;
      for (;;) {
         if (variant == 0) {
            for (int j = 0; j < i; ++j) {
               if (skip[j]) continue;
               const double *q = Q + (int64_t)j * ldq;
               const double r = o_dot(m, v, q);
               RR(j, i) += r;
               for (int64_t k = 0; k < m; ++k) v[k] = fma(-r, q[k], v[k]);
            }
         } else {
            for (int j = 0; j < i; ++j) rv[j] = skip[j] ? 0.0 : o_dot(m, Q + (int64_t)j * ldq, v);
            for (int j = 0; j < i; ++j) RR(j, i) += rv[j];
            for (int64_t k = 0; k < m; ++k) {
               double s = 0.0;
               for (int j = 0; j < i; ++j)
                  if (rv[j] != 0.0) s = fma(rv[j], Q[k + (int64_t)j * ldq], s);
               v[k] = v[k] - s;
            }
         }
         ++pass;
         if (variant == 3 && pass < 2) continue;
         const double t = sqrt(o_dot(m, v, v));
         /* at most 64 reorthogonalization passes; none for t = 0 */
         if (reorth && reorths < 64 && t > 0.0 && t >= orth_tol && t < reorth_tol * normv) {
            normv = t;
            ++reorths;
            continue;
         }
         if (t < select_tol) {
            skip[i] = 1;
            RR(i, i) = rd_skip;
         } else {
            RR(i, i) = t;
            const double s = 1.0 / t;
            for (int64_t k = 0; k < m; ++k) v[k] = v[k] * s;
         }
         break;
      }
   }
#undef RR
   free(rv);
}

/* ------------------------------------------------------------------------- */
/* symmetric eigensolver: parallel-ordered cyclic Jacobi (round-robin pairing,  */
/* Brent & Luk 1985), Demmel-Veselic threshold with an absolute floor           */
/* eps max|a_ij|, explicit fused multiply-adds.                                 */
/* Every round rotates all disjoint pairs at once; the 2x2 blocks of A are      */
/* transformed independently and the rotations are then applied to V in round   */
/* order -- the same arithmetic, in the same order, as the device solver.       */
/* ------------------------------------------------------------------------- */
static void pair_of(int nn, int r, int s, int *p, int *q) {
   int a, b;
   if (s == 0) { a = nn - 1; b = r; }
   else { a = (r + s) % (nn - 1); b = (r - s + nn - 1) % (nn - 1); }
   if (a > b) { int t = a; a = b; b = t; }
   *p = a; *q = b;
}

static int jacobi_eig(int n, double *a /* n x n symmetric, destroyed */, double *w, double *V,
                      int ldv) {
   const int nn = n + (n & 1), np = nn / 2;
   double *rc = (double *)malloc(sizeof(double) * (size_t)np);
   double *rs = (double *)malloc(sizeof(double) * (size_t)np);
   int *pp = (int *)malloc(sizeof(int) * (size_t)np), *qq = (int *)malloc(sizeof(int) * (size_t)np);
   double *Vw = (double *)malloc(sizeof(double) * (size_t)n * n);
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) Vw[i + (size_t)j * n] = (i == j) ? 1.0 : 0.0;
#define A_(i, j) a[(i) + (size_t)(j) * n]
   double amax = 0.0;
   for (size_t t = 0; t < (size_t)n * n; ++t) amax = fmax(amax, fabs(a[t]));
   const double tol_abs = 2.220446049250313e-16 * amax;
   int sweep;
   for (sweep = 0; sweep < 30; ++sweep) {
      int rotated = 0;
      for (int r = 0; r < nn - 1; ++r) {
         for (int s = 0; s < np; ++s) {
            int p, q;
            pair_of(nn, r, s, &p, &q);
            pp[s] = p; qq[s] = q;
            double c = 1.0, sn = 0.0;
            if (q < n) {
               const double apq = A_(p, q), app = A_(p, p), aqq = A_(q, q);
               if (fabs(apq) > fmax(1.1102230246251565e-16 * sqrt(fabs(app) * fabs(aqq)), tol_abs)) {
                  const double th = (aqq - app) / (2.0 * apq);
                  const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(fma(th, th, 1.0)));
                  c = 1.0 / sqrt(fma(t, t, 1.0));
                  sn = t * c;
                  rotated = 1;
               }
            }
            rc[s] = c; rs[s] = sn;
         }
         for (int I = 0; I < np; ++I)
            for (int J = I; J < np; ++J) {
               const double s1 = rs[I], s2 = rs[J];
               if (s1 == 0.0 && s2 == 0.0) continue;
               const int p1 = pp[I], q1 = qq[I], p2 = pp[J], q2 = qq[J];
               const double c1 = rc[I], c2 = rc[J];
               const int q1ok = q1 < n, q2ok = q2 < n;
               if (I == J) {
                  if (!q1ok) continue;
                  const double app = A_(p1, p1), apq = A_(p1, q1), aqq = A_(q1, q1);
                  const double m11 = fma(c1, app, -(s1 * apq)), m12 = fma(s1, app, c1 * apq);
                  const double m21 = fma(c1, apq, -(s1 * aqq)), m22 = fma(s1, apq, c1 * aqq);
                  A_(p1, p1) = fma(c1, m11, -(s1 * m21));
                  A_(q1, q1) = fma(s1, m12, c1 * m22);
                  A_(p1, q1) = A_(q1, p1) = 0.0;
               } else {
                  const double a11 = A_(p1, p2);
                  const double a12 = q2ok ? A_(p1, q2) : 0.0;
                  const double a21 = q1ok ? A_(q1, p2) : 0.0;
                  const double a22 = (q1ok && q2ok) ? A_(q1, q2) : 0.0;
                  const double m11 = fma(c2, a11, -(s2 * a12)), m12 = fma(s2, a11, c2 * a12);
                  const double m21 = fma(c2, a21, -(s2 * a22)), m22 = fma(s2, a21, c2 * a22);
                  const double b11 = fma(c1, m11, -(s1 * m21)), b12 = fma(c1, m12, -(s1 * m22));
                  const double b21 = fma(s1, m11, c1 * m21), b22 = fma(s1, m12, c1 * m22);
                  A_(p1, p2) = A_(p2, p1) = b11;
                  if (q2ok) A_(p1, q2) = A_(q2, p1) = b12;
                  if (q1ok) A_(q1, p2) = A_(p2, q1) = b21;
                  if (q1ok && q2ok) A_(q1, q2) = A_(q2, q1) = b22;
               }
            }
         for (int s = 0; s < np; ++s) {
            if (rs[s] == 0.0 || qq[s] >= n) continue;
            const double c = rc[s], sn = rs[s];
            double *vp = Vw + (size_t)pp[s] * n, *vq = Vw + (size_t)qq[s] * n;
            for (int row = 0; row < n; ++row) {
               const double x = vp[row], y = vq[row];
               vp[row] = fma(c, x, -(sn * y));
               vq[row] = fma(sn, x, c * y);
            }
         }
      }
      if (!rotated) break;
   }
   /* eigenvalues ascending (stable rank), vectors permuted accordingly */
   for (int i = 0; i < n; ++i) {
      const double di = A_(i, i);
      int rk = 0;
      for (int j = 0; j < n; ++j) {
         const double dj = A_(j, j);
         rk += (dj < di || (dj == di && j < i)) ? 1 : 0;
      }
      w[rk] = di;
      memcpy(V + (int64_t)rk * ldv, Vw + (size_t)i * n, sizeof(double) * (size_t)n);
   }
#undef A_
   free(rc); free(rs); free(pp); free(qq); free(Vw);
   return sweep;
}

int orc_syev(int n, const double *A, int lda, double *w, double *V, int ldv) {
   double *a = (double *)malloc(sizeof(double) * (size_t)n * n);
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) /* symmetrise from the upper triangle ('U') */
         a[i + j * n] = (i <= j) ? A[i + (int64_t)j * lda] : A[j + (int64_t)i * lda];
   int sweeps = jacobi_eig(n, a, w, V, ldv);
   free(a);
   return sweeps; /* number of Jacobi sweeps run */
}

/* Cholesky B = L L^T, right-looking (column j scaled, then the trailing lower
 * triangle updated with fused multiply-adds); returns 0 or the failing order */
static int cholesky(int n, double *L) {
   for (int j = 0; j < n; ++j) {
      double d = L[j + j * n];
      if (!(d > 0.0)) return j + 1;
      d = sqrt(d);
      L[j + j * n] = d;
      for (int i = j + 1; i < n; ++i) L[i + j * n] = L[i + j * n] / d;
      for (int c = j + 1; c < n; ++c)
         for (int i = c; i < n; ++i) L[i + c * n] = fma(-L[i + j * n], L[c + j * n], L[i + c * n]);
   }
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < j; ++i) L[i + j * n] = 0.0;
   return 0;
}

int orc_sygv(int n, const double *A, int lda, const double *B, int ldb, double *w, double *V,
             int ldv) {
   double *L = (double *)malloc(sizeof(double) * (size_t)n * n);
   double *C = (double *)malloc(sizeof(double) * (size_t)n * n);
   double *Y = (double *)malloc(sizeof(double) * (size_t)n * n);
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
         L[i + j * n] = (i >= j) ? B[j + (int64_t)i * ldb] : B[i + (int64_t)j * ldb];
         C[i + j * n] = (i <= j) ? A[i + (int64_t)j * lda] : A[j + (int64_t)i * lda];
      }
   int info = cholesky(n, L);
   if (info) { free(L); free(C); free(Y); return info; }
   /* C <- L^{-1} C L^{-T}: forward-substitute columns, then rows */
   for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
         double s = C[i + j * n];
         for (int p = 0; p < i; ++p) s = fma(-L[i + p * n], C[p + j * n], s);
         C[i + j * n] = s / L[i + i * n];
      }
   for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
         double s = C[i + j * n];
         for (int p = 0; p < j; ++p) s = fma(-L[j + p * n], C[i + p * n], s);
         C[i + j * n] = s / L[j + j * n];
      }
   for (int j = 0; j < n; ++j) /* exact symmetry */
      for (int i = j + 1; i < n; ++i) {
         double s = 0.5 * (C[i + j * n] + C[j + i * n]);
         C[i + j * n] = C[j + i * n] = s;
      }
   jacobi_eig(n, C, w, Y, n);
   /* V = L^{-T} Y: back-substitution */
   for (int c = 0; c < n; ++c)
      for (int i = n - 1; i >= 0; --i) {
         double s = Y[i + c * n];
         for (int p = i + 1; p < n; ++p) s = fma(-L[p + i * n], V[p + (int64_t)c * ldv], s);
         V[i + (int64_t)c * ldv] = s / L[i + i * n];
      }
   free(L); free(C); free(Y);
   return 0;
}

int orc_rr_truncated(int m, const double *G, int ldg, const double *H, int ldh, double tau,
                     int nkeep, double *theta, double *Y, int ldy) {
   double *sc = (double *)malloc(sizeof(double) * (size_t)m);
   double *Gs = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *Hs = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *U = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *lam = (double *)malloc(sizeof(double) * (size_t)m);
   for (int i = 0; i < m; ++i) {
      double d = G[i + (int64_t)i * ldg];
      sc[i] = d > 0 ? 1.0 / sqrt(d) : 0.0;
   }
   for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
         double g = 0.5 * (G[i + (int64_t)j * ldg] + G[j + (int64_t)i * ldg]);
         double h = 0.5 * (H[i + (int64_t)j * ldh] + H[j + (int64_t)i * ldh]);
         Gs[i + j * m] = g * sc[i] * sc[j];
         Hs[i + j * m] = h * sc[i] * sc[j];
      }
   jacobi_eig(m, Gs, lam, U, m);
   double lmax = lam[m - 1];
   int r = 0;
   for (int i = 0; i < m; ++i)
      if (lam[i] > tau * lmax) ++r;
   /* B (m x r) = U_r L_r^{-1/2}, the r largest */
   double *B = (double *)malloc(sizeof(double) * (size_t)m * (r ? r : 1));
   for (int c = 0; c < r; ++c) {
      int src = m - r + c;
      const double f = sqrt(lam[src]);
      for (int i = 0; i < m; ++i) B[i + c * m] = U[i + src * m] / f;
   }
   /* Hr = B^T Hs B */
   double *T = (double *)malloc(sizeof(double) * (size_t)m * (r ? r : 1));
   double *Hr = (double *)malloc(sizeof(double) * (size_t)(r ? r : 1) * (r ? r : 1));
   double *Z = (double *)malloc(sizeof(double) * (size_t)(r ? r : 1) * (r ? r : 1));
   double *th = (double *)malloc(sizeof(double) * (size_t)(r ? r : 1));
   orc_gemm_nn(m, r, m, Hs, m, B, m, T, m);
   for (int a2 = 0; a2 < r; ++a2) /* Hr = B^T T, one sequential dot per entry */
      for (int b2 = 0; b2 < r; ++b2) {
         double acc = 0.0;
         for (int p = 0; p < m; ++p) acc = fma(B[p + a2 * m], T[p + b2 * m], acc);
         Hr[a2 + b2 * r] = acc;
      }
   for (int j = 0; j < r; ++j) /* the eigensolver reads the upper triangle */
      for (int i = j + 1; i < r; ++i) Hr[i + j * r] = Hr[j + i * r];
   if (r > 0) jacobi_eig(r, Hr, th, Z, r);
   int nout = nkeep < r ? nkeep : r;
   for (int c = 0; c < nout; ++c) {
      int src = r - 1 - c; /* descending */
      theta[c] = th[src];
      for (int i = 0; i < m; ++i) {
         double s = 0.0;
         for (int p = 0; p < r; ++p) s = fma(B[i + p * m], Z[p + src * r], s);
         Y[i + (int64_t)c * ldy] = s * sc[i];
      }
   }
   for (int c = nout; c < nkeep; ++c) {
      theta[c] = NAN;
      for (int i = 0; i < m; ++i) Y[i + (int64_t)c * ldy] = 0.0;
   }
   free(sc); free(Gs); free(Hs); free(U); free(lam); free(B); free(T); free(Hr); free(Z); free(th);
   return r;
}

/* Split truncated Rayleigh-Ritz for a basis V = [X | D] whose first kx columns
 * are the previous Ritz block (X^T X ~ I, so its Gram block is well
 * conditioned).  Same pencil and truncation idea as orc_rr_truncated
 * (eig_rr.m:48-57 with the shift replaced by truncation), reached through a
 * congruence that keeps the eigendecompositions small:
 *   G_XX = L L^T (Cholesky), C = G_XX^{-1} G_XD, S = G_DD - G_XD^T C (the Gram
 *   matrix of D with its X component removed), s = diag(S)^{-1/2};
 *   eig(s S s) = U Lam, kept Lam_i > tau max(1, max Lam) -> B_D = s U_r Lam_r^{-1/2};
 *   W = [[L^{-T}, -C B_D], [0, B_D]] (W^T G W = I on the kept directions);
 *   Hr = W^T H W; eig(Hr_DD) = Q hd;  Hr2 = [[Hr_XX, Hr_XD Q], [., diag(hd)]];
 *   eig(Hr2) = Z th;  Y = W blkdiag(I, Q) Z (top nkeep, descending).
 * Only the upper triangles of S, Hr_DD and Hr2 are read by the eigensolver.
 * Every sum is a sequential fused multiply-add chain in increasing index order,
 * except the back substitutions with L^T, whose terms run in decreasing index
 * order (the order in which a right-looking substitution finalises them); the
 * device kernels run the same chains.  Returns the kept dimension kx + r,
 * or -1 when G_XX is not positive definite. */
int orc_rr_split(int m, int kx, const double *G, int ldg, const double *H, int ldh, double tau,
                 int nkeep, double *theta, double *Y, int ldy) {
   const int md = m - kx;
#define Gs_(i, j) (0.5 * (G[(i) + (int64_t)(j) * ldg] + G[(j) + (int64_t)(i) * ldg]))
#define Hs_(i, j) (0.5 * (H[(i) + (int64_t)(j) * ldh] + H[(j) + (int64_t)(i) * ldh]))
   double *L = (double *)calloc((size_t)kx * kx + 1, sizeof(double));
   for (int j = 0; j < kx; ++j)
      for (int i = j; i < kx; ++i) L[i + (size_t)j * kx] = Gs_(i, j);
   if (cholesky(kx, L)) { free(L); return -1; }
#define L_(i, j) L[(i) + (size_t)(j) * kx]
   /* C = G_XX^{-1} G_XD: forward then backward substitution per column */
   double *Cc = (double *)malloc(sizeof(double) * ((size_t)kx * md + 1));
   for (int c = 0; c < md; ++c) {
      double *x = Cc + (size_t)c * kx;
      for (int i = 0; i < kx; ++i) {
         double s2 = Gs_(i, kx + c);
         for (int p = 0; p < i; ++p) s2 = fma(-L_(i, p), x[p], s2);
         x[i] = s2 / L_(i, i);
      }
      for (int i = kx - 1; i >= 0; --i) {
         double s2 = x[i];
         for (int p = kx - 1; p > i; --p) s2 = fma(-L_(p, i), x[p], s2);
         x[i] = s2 / L_(i, i);
      }
   }
   /* S (upper) = G_DD - G_XD^T C, scaled by diag(S)^{-1/2} */
   double *S = (double *)malloc(sizeof(double) * ((size_t)md * md + 1));
   double *sd = (double *)malloc(sizeof(double) * (md + 1));
   for (int j = 0; j < md; ++j)
      for (int i = 0; i <= j; ++i) {
         double s2 = Gs_(kx + i, kx + j);
         for (int p = 0; p < kx; ++p) s2 = fma(-Gs_(p, kx + i), Cc[p + (size_t)j * kx], s2);
         S[i + (size_t)j * md] = s2;
      }
   for (int i = 0; i < md; ++i) {
      const double d = S[i + (size_t)i * md];
      sd[i] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
   }
   for (int j = 0; j < md; ++j)
      for (int i = 0; i <= j; ++i) {
         const double v = S[i + (size_t)j * md] * sd[i] * sd[j];
         S[i + (size_t)j * md] = v;
         S[j + (size_t)i * md] = v;
      }
   double *lam = (double *)malloc(sizeof(double) * (md + 1));
   double *U = (double *)malloc(sizeof(double) * ((size_t)md * md + 1));
   if (md > 0) jacobi_eig(md, S, lam, U, md);
   double lmax = (md > 0 && lam[md - 1] > 1.0) ? lam[md - 1] : 1.0;
   int r = 0;
   for (int i = 0; i < md; ++i)
      if (lam[i] > tau * lmax) ++r;
   const int mr = kx + r;
   /* W (m x mr) */
   double *W = (double *)calloc((size_t)m * mr + 1, sizeof(double));
#define W_(i, j) W[(i) + (size_t)(j) * m]
   for (int c = 0; c < kx; ++c) /* L^{-T}: solve L^T w = e_c */
      for (int i = kx - 1; i >= 0; --i) {
         double s2 = (i == c) ? 1.0 : 0.0;
         for (int p = kx - 1; p > i; --p) s2 = fma(-L_(p, i), W_(p, c), s2);
         W_(i, c) = s2 / L_(i, i);
      }
   for (int c = 0; c < r; ++c) {
      const int src = md - r + c;
      const double f = sqrt(lam[src]);
      for (int i = 0; i < md; ++i) W_(kx + i, kx + c) = U[i + (size_t)src * md] / f * sd[i];
   }
   for (int c = 0; c < r; ++c)
      for (int i = 0; i < kx; ++i) {
         double s2 = 0.0;
         for (int p = 0; p < md; ++p) s2 = fma(Cc[i + (size_t)p * kx], W_(kx + p, kx + c), s2);
         W_(i, kx + c) = -s2;
      }
   /* Hr = W^T (H W), full */
   double *T = (double *)malloc(sizeof(double) * ((size_t)m * mr + 1));
   double *Hr = (double *)malloc(sizeof(double) * ((size_t)mr * mr + 1));
   for (int c = 0; c < mr; ++c)
      for (int i = 0; i < m; ++i) {
         double s2 = 0.0;
         for (int p = 0; p < m; ++p) s2 = fma(Hs_(i, p), W_(p, c), s2);
         T[i + (size_t)c * m] = s2;
      }
   for (int b = 0; b < mr; ++b)
      for (int a = 0; a < mr; ++a) {
         double s2 = 0.0;
         for (int p = 0; p < m; ++p) s2 = fma(W_(p, a), T[p + (size_t)b * m], s2);
         Hr[a + (size_t)b * mr] = s2;
      }
   /* eig of the D block (upper triangle) */
   double *Hd = (double *)malloc(sizeof(double) * ((size_t)r * r + 1));
   double *hd = (double *)malloc(sizeof(double) * (r + 1));
   double *Q = (double *)malloc(sizeof(double) * ((size_t)r * r + 1));
   for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i) {
         const int a = i < j ? i : j, b = i < j ? j : i;
         Hd[i + (size_t)j * r] = Hr[(kx + a) + (size_t)(kx + b) * mr];
      }
   if (r > 0) jacobi_eig(r, Hd, hd, Q, r);
   /* Hr2 (full from its upper triangle) and W2 = W blkdiag(I, Q) */
   double *H2 = (double *)malloc(sizeof(double) * ((size_t)mr * mr + 1));
   for (int b = 0; b < mr; ++b)
      for (int a = 0; a <= b; ++a) {
         double v;
         if (b < kx) v = Hr[a + (size_t)b * mr];
         else if (a < kx) {
            double s2 = 0.0;
            for (int p = 0; p < r; ++p) s2 = fma(Hr[a + (size_t)(kx + p) * mr], Q[p + (size_t)(b - kx) * r], s2);
            v = s2;
         } else v = (a == b) ? hd[a - kx] : 0.0;
         H2[a + (size_t)b * mr] = v;
         H2[b + (size_t)a * mr] = v;
      }
   double *W2 = (double *)malloc(sizeof(double) * ((size_t)m * mr + 1));
   for (int c = 0; c < mr; ++c)
      for (int i = 0; i < m; ++i) {
         double v = W_(i, c);
         if (c >= kx) {
            double s2 = 0.0;
            for (int p = 0; p < r; ++p) s2 = fma(W_(i, kx + p), Q[p + (size_t)(c - kx) * r], s2);
            v = s2;
         }
         W2[i + (size_t)c * m] = v;
      }
   double *th = (double *)malloc(sizeof(double) * (mr + 1));
   double *Z = (double *)malloc(sizeof(double) * ((size_t)mr * mr + 1));
   jacobi_eig(mr, H2, th, Z, mr);
   const int nout = nkeep < mr ? nkeep : mr;
   for (int c = 0; c < nout; ++c) {
      const int src = mr - 1 - c;
      theta[c] = th[src];
      for (int i = 0; i < m; ++i) {
         double s2 = 0.0;
         for (int p = 0; p < mr; ++p) s2 = fma(W2[i + (size_t)p * m], Z[p + (size_t)src * mr], s2);
         Y[i + (int64_t)c * ldy] = s2;
      }
   }
   for (int c = nout; c < nkeep; ++c) {
      theta[c] = NAN;
      for (int i = 0; i < m; ++i) Y[i + (int64_t)c * ldy] = 0.0;
   }
#undef Gs_
#undef Hs_
#undef L_
#undef W_
   free(L); free(Cc); free(S); free(sd); free(lam); free(U); free(W); free(T); free(Hr);
   free(Hd); free(hd); free(Q); free(H2); free(W2); free(th); free(Z);
   return mr;
}

int orc_rr_generalized(int m, const double *G, int ldg, const double *H, int ldh, double gtol,
                       double *theta, double *Y, int ldy) {
   double *B = (double *)calloc((size_t)m * m + 1, sizeof(double));
   double *V = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *w = (double *)malloc(sizeof(double) * (size_t)m);
   for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) B[i + j * m] = G[i + (int64_t)j * ldg] + (i == j ? gtol : 0.0);
   int info = orc_sygv(m, H, ldh, B, m, w, V, m);
   if (!info)
      for (int c = 0; c < m; ++c) {
         theta[c] = w[m - 1 - c];
         memcpy(Y + (int64_t)c * ldy, V + (int64_t)(m - 1 - c) * m, sizeof(double) * (size_t)m);
      }
   free(B); free(V); free(w);
   return info;
}

/* ------------------------------------------------------------------------- */
/* drivers                                                                    */
/* ------------------------------------------------------------------------- */
int orc_arnoldi_hess_eig(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                         const double *vals, int k, int kdim, int maxiter, double select_tol,
                         double gtol, const double *v0, double *theta, double *X, int64_t ldx) {
   int m = kdim + 1;
   if (k > n) k = (int)n;
   double *V = (double *)calloc((size_t)n * m, sizeof(double));
   double *AV = (double *)malloc(sizeof(double) * (size_t)n * m);
   double *G = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *H = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *Y = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *th = (double *)malloc(sizeof(double) * (size_t)m);
   double *U = (double *)malloc(sizeof(double) * (size_t)n * m);
   int *perm = (int *)malloc(sizeof(int) * (size_t)m);
   memcpy(V, v0, sizeof(double) * (size_t)n);
   int tits = 1, built = 0, have = 0;
   while (tits < maxiter) {
      double piv;
      int l = orc_iamax(n, V, &piv) - 1;
      perm[0] = l;
      if (fabs(piv) < select_tol) break;
      for (int64_t r = 0; r < n; ++r) V[r] = V[r] / piv;
      int ncols = 1;
      for (int i = 0; i < kdim; ++i) {
         double *w = V + (int64_t)(i + 1) * n;
         orc_spmm_csr(n, row_ptr, col_idx, vals, 1, V + (int64_t)i * n, n, w, n);
         for (int j = 0; j <= i; ++j) {
            double c = w[perm[j]];
            const double *q = V + (int64_t)j * n;
            for (int64_t r = 0; r < n; ++r) w[r] = fma(-c, q[r], w[r]);
         }
         l = orc_iamax(n, w, &piv) - 1;
         perm[i + 1] = l;
         if (fabs(piv) < select_tol) { tits = maxiter; break; }
         for (int64_t r = 0; r < n; ++r) w[r] = w[r] / piv;
         ++ncols;
         ++tits;
      }
      /* Rayleigh-Ritz on the columns built in this sweep (eig_rr.m:48-57) */
      built = ncols;
      orc_spmm_csr(n, row_ptr, col_idx, vals, ncols, V, n, AV, n);
      orc_gemm_tn(n, ncols, ncols, V, n, AV, n, H, ncols);
      orc_gemm_tn(n, ncols, ncols, V, n, V, n, G, ncols);
      int info = orc_rr_generalized(ncols, G, ncols, H, ncols, gtol, th, Y, ncols);
      if (info) break;
      orc_gemm_nn(n, ncols, ncols, V, n, Y, ncols, U, n);
      have = ncols;
      if (tits < maxiter) {
         memcpy(V, U, sizeof(double) * (size_t)n); /* restart on the top Ritz vector */
      }
   }
   int kk = k < have ? k : have;
   for (int c = 0; c < kk; ++c) {
      theta[c] = th[c];
      memcpy(X + (int64_t)c * ldx, U + (int64_t)c * n, sizeof(double) * (size_t)n);
   }
   for (int c = kk; c < k; ++c) theta[c] = NAN;
   free(V); free(AV); free(G); free(H); free(Y); free(th); free(U); free(perm);
   return built;
}

void orc_subspace_iter_eig(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                           const double *vals, int m, int maxiter, int step_size,
                           double select_tol, double gtol, double *V, int64_t ldv, double *theta) {
   double *T = (double *)malloc(sizeof(double) * (size_t)n * m);
   double *AV = (double *)malloc(sizeof(double) * (size_t)n * m);
   double *R = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *G = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *H = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *Y = (double *)malloc(sizeof(double) * (size_t)m * m);
   int *sel = (int *)malloc(sizeof(int) * (size_t)m), *perm = (int *)malloc(sizeof(int) * (size_t)m);
   for (int j = 0; j < m; ++j) memcpy(T + (int64_t)j * n, V + (int64_t)j * ldv, sizeof(double) * (size_t)n);
   orc_scalecols(n, m, T, n);
   for (int it = 0; it < maxiter; ++it) {
      for (int s = 0; s < step_size; ++s) {
         orc_spmm_csr(n, row_ptr, col_idx, vals, m, T, n, AV, n);
         memcpy(T, AV, sizeof(double) * (size_t)n * m);
         orc_scalecols(n, m, T, n);
      }
      orc_hessenberg(n, m, T, n, R, m, select_tol, 0, sel, perm);
      /* economic Q: selected columns only (hessenberg.m:76-78) */
      int ms = 0;
      for (int j = 0; j < m; ++j)
         if (sel[j]) {
            if (ms != j) memcpy(T + (int64_t)ms * n, T + (int64_t)j * n, sizeof(double) * (size_t)n);
            ++ms;
         }
      orc_spmm_csr(n, row_ptr, col_idx, vals, ms, T, n, AV, n);
      orc_gemm_tn(n, ms, ms, T, n, AV, n, H, ms);
      orc_gemm_tn(n, ms, ms, T, n, T, n, G, ms);
      orc_rr_generalized(ms, G, ms, H, ms, gtol, theta, Y, ms);
      orc_gemm_nn(n, ms, ms, T, n, Y, ms, AV, n);
      memcpy(T, AV, sizeof(double) * (size_t)n * ms);
      for (int j = ms; j < m; ++j) { theta[j] = NAN; memset(T + (int64_t)j * n, 0, sizeof(double) * (size_t)n); }
   }
   for (int j = 0; j < m; ++j) memcpy(V + (int64_t)j * ldv, T + (int64_t)j * n, sizeof(double) * (size_t)n);
   free(T); free(AV); free(R); free(G); free(H); free(Y); free(sel); free(perm);
}

/* Two-sided Rayleigh-Ritz for the SVD (MATLAB/+src/+rr/svd_rr.m:54-91) over a
 * dense A (m x n, column-major): AV = A V, B = U^T AV (ku x kv); the augmented
 * pencil eig([0 B; B^T 0], blkdiag(U^T U, V^T V) + gtol I) (the generalised form,
 * svd_rr.m:57-76) or eig([0 B; B^T 0]) (the SVD of B, svd_rr.m:78); sorted
 * descending; sigma = the top kk = min(ku, kv) eigenvalues and, with (Z; W)
 * their eigenvectors, U <- sqrt(2) U Z, V <- sqrt(2) V W (svd_rr.m:85-91).
 * U (m x ku) and V (n x kv) in; their first kk columns out.  Returns kk. */
int orc_svd_rr(int64_t m, int64_t n, const double *A, int64_t lda, int ku, int kv, double *U,
               int64_t ldu, double *V, int64_t ldv, int generalized, double gtol, double *sigma) {
   const int K = ku + kv, kk = ku < kv ? ku : kv;
   double *AV = (double *)malloc(sizeof(double) * (size_t)m * kv);
   double *B = (double *)malloc(sizeof(double) * (size_t)ku * kv);
   double *Ae = (double *)calloc((size_t)K * K, sizeof(double));
   double *Me = (double *)calloc((size_t)K * K, sizeof(double));
   double *th = (double *)malloc(sizeof(double) * (size_t)K);
   double *Y = (double *)malloc(sizeof(double) * (size_t)K * K);
   double *Un = (double *)malloc(sizeof(double) * (size_t)m * kk);
   double *Vn = (double *)malloc(sizeof(double) * (size_t)n * kk);
   orc_gemm_nn(m, kv, n, A, lda, V, ldv, AV, m);
   orc_gemm_tn(m, ku, kv, U, ldu, AV, m, B, ku);
   for (int j = 0; j < kv; ++j)
      for (int i = 0; i < ku; ++i) {
         Ae[i + (int64_t)(ku + j) * K] = B[i + (int64_t)j * ku];
         Ae[(ku + j) + (int64_t)i * K] = B[i + (int64_t)j * ku];
      }
   if (generalized) {
      double *G = (double *)malloc(sizeof(double) * (size_t)(ku > kv ? ku : kv) * (ku > kv ? ku : kv));
      orc_gemm_tn(m, ku, ku, U, ldu, U, ldu, G, ku);
      for (int j = 0; j < ku; ++j)
         for (int i = 0; i < ku; ++i) Me[i + (int64_t)j * K] = G[i + (int64_t)j * ku];
      orc_gemm_tn(n, kv, kv, V, ldv, V, ldv, G, kv);
      for (int j = 0; j < kv; ++j)
         for (int i = 0; i < kv; ++i) Me[(ku + i) + (int64_t)(ku + j) * K] = G[i + (int64_t)j * kv];
      free(G);
   } else {
      for (int i = 0; i < K; ++i) Me[i + (int64_t)i * K] = 1.0;
      gtol = 0.0;
   }
   const int info = orc_rr_generalized(K, Me, K, Ae, K, gtol, th, Y, K);
   if (!info) {
      const double r2 = sqrt(2.0);
      /* Z = Y(0:ku, 0:kk), W = Y(ku:K, 0:kk) */
      orc_gemm_nn(m, kk, ku, U, ldu, Y, K, Un, m);
      orc_gemm_nn(n, kk, kv, V, ldv, Y + ku, K, Vn, n);
      for (int c = 0; c < kk; ++c) {
         sigma[c] = th[c];
         for (int64_t i = 0; i < m; ++i) U[i + (int64_t)c * ldu] = r2 * Un[i + (int64_t)c * m];
         for (int64_t i = 0; i < n; ++i) V[i + (int64_t)c * ldv] = r2 * Vn[i + (int64_t)c * n];
      }
   }
   free(AV); free(B); free(Ae); free(Me); free(th); free(Y); free(Un); free(Vn);
   return info ? -info : kk;
}

/* subspace_iter_svd.m:57-77 with function_orth = hessenberg (select_tol; the
 * kept columns only) and function_rr = svd_rr (generalised with gtol, or not):
 * V (n x k) start block, scaled by scalecols; each iteration step_size x (U = A V,
 * scalecols; V = A^T U, scalecols), then P = hessenberg(U), Q = hessenberg(V),
 * [U, S, V] = svd_rr(A, P, Q).  U (m x k), V (n x k) and sigma (k) out (sigma NaN
 * and zero columns past the kept width).  Returns the kept width. */
int orc_subspace_iter_svd(int64_t m, int64_t n, const double *A, int64_t lda, int k, int maxiter,
                          int step_size, double select_tol, int generalized, double gtol,
                          const double *V0, double *U, int64_t ldu, double *V, int64_t ldv,
                          double *sigma) {
   double *Tu = (double *)calloc((size_t)m * k, sizeof(double));
   double *Tv = (double *)malloc(sizeof(double) * (size_t)n * k);
   double *R = (double *)malloc(sizeof(double) * (size_t)k * k);
   double *At = (double *)malloc(sizeof(double) * (size_t)n * m);
   int *sel = (int *)malloc(sizeof(int) * (size_t)k), *perm = (int *)malloc(sizeof(int) * (size_t)k);
   for (int64_t j = 0; j < m; ++j)
      for (int64_t i = 0; i < n; ++i) At[i + j * n] = A[j + i * lda];
   for (int j = 0; j < k; ++j) memcpy(Tv + (int64_t)j * n, V0 + (int64_t)j * n, sizeof(double) * (size_t)n);
   orc_scalecols(n, k, Tv, n);
   int kc = k;
   for (int it = 0; it < maxiter && kc > 0; ++it) {
      for (int s = 0; s < step_size; ++s) {
         orc_gemm_nn(m, kc, n, A, lda, Tv, n, Tu, m);
         orc_scalecols(m, kc, Tu, m);
         orc_gemm_nn(n, kc, m, At, n, Tu, m, Tv, n);
         orc_scalecols(n, kc, Tv, n);
      }
      int ku = 0, kv = 0;
      orc_hessenberg(m, kc, Tu, m, R, k, select_tol, 0, sel, perm);
      for (int j = 0; j < kc; ++j)
         if (sel[j]) {
            if (ku != j) memcpy(Tu + (int64_t)ku * m, Tu + (int64_t)j * m, sizeof(double) * (size_t)m);
            ++ku;
         }
      orc_hessenberg(n, kc, Tv, n, R, k, select_tol, 0, sel, perm);
      for (int j = 0; j < kc; ++j)
         if (sel[j]) {
            if (kv != j) memcpy(Tv + (int64_t)kv * n, Tv + (int64_t)j * n, sizeof(double) * (size_t)n);
            ++kv;
         }
      if (ku == 0 || kv == 0) { kc = 0; break; }
      kc = orc_svd_rr(m, n, A, lda, ku, kv, Tu, m, Tv, n, generalized, gtol, sigma);
      if (kc < 0) { kc = 0; break; }
   }
   for (int j = 0; j < k; ++j) {
      if (j < kc) {
         memcpy(U + (int64_t)j * ldu, Tu + (int64_t)j * m, sizeof(double) * (size_t)m);
         memcpy(V + (int64_t)j * ldv, Tv + (int64_t)j * n, sizeof(double) * (size_t)n);
      } else {
         sigma[j] = NAN;
         memset(U + (int64_t)j * ldu, 0, sizeof(double) * (size_t)m);
         memset(V + (int64_t)j * ldv, 0, sizeof(double) * (size_t)n);
      }
   }
   free(Tu); free(Tv); free(R); free(At); free(sel); free(perm);
   return kc;
}

void orc_residuals(int64_t n, const int64_t *row_ptr, const int32_t *col_idx, const double *vals,
                   int k, const double *X, int64_t ldx, const double *theta, double normA,
                   double *res) {
   double *AX = (double *)malloc(sizeof(double) * (size_t)n * k);
   orc_spmm_csr(n, row_ptr, col_idx, vals, k, X, ldx, AX, n);
   for (int j = 0; j < k; ++j) {
      double s = 0.0, t = 0.0;
      for (int64_t i = 0; i < n; ++i) {
         double r = fma(-theta[j], X[i + j * ldx], AX[i + j * n]);
         s += r * r;
         t += X[i + j * ldx] * X[i + j * ldx];
      }
      res[j] = sqrt(s) / (normA * sqrt(t));
   }
   free(AX);
}

/* ---- filtered OFRR ------------------------------------------------------- */
static double colmaxabs(int64_t n, const double *a) {
   double m = 0.0;
   for (int64_t i = 0; i < n; ++i) m = fmax(m, fabs(a[i]));
   return m;
}

int orc_filtered_ofrr(int64_t n, const int64_t *row_ptr, const int32_t *col_idx,
                      const double *vals, const orc_ofrr_opts *o, double *X, int64_t ldx,
                      double *theta, double *res) {
   const int k = o->k, nd = o->nd, dmax = o->degree;
   const int m = k * (1 + nd);
   double *V = (double *)malloc(sizeof(double) * (size_t)n * m);  /* [X | Dd | Dmid] */
   double *W = (double *)malloc(sizeof(double) * (size_t)n * m);  /* A V */
   double *Rr = (double *)malloc(sizeof(double) * (size_t)n * k);
   double *Dp = (double *)malloc(sizeof(double) * (size_t)n * k);
   double *T = (double *)malloc(sizeof(double) * (size_t)n * k);
   double *G = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *H = (double *)malloc(sizeof(double) * (size_t)m * m);
   double *C = (double *)malloc(sizeof(double) * (size_t)m * k);
   double *zc = (double *)malloc(sizeof(double) * k), *zp = (double *)malloc(sizeof(double) * k);
   double *Lth = (double *)malloc(sizeof(double) * k);
   for (int j = 0; j < k; ++j) memcpy(V + (int64_t)j * n, X + (int64_t)j * ldx, sizeof(double) * (size_t)n);
   /* initial Rayleigh-Ritz on the start block */
   orc_spmm_csr(n, row_ptr, col_idx, vals, k, V, n, W, n);
   orc_gemm_tn(n, k, k, V, n, V, n, G, k);
   orc_gemm_tn(n, k, k, V, n, W, n, H, k);
   orc_rr_truncated(k, G, k, H, k, o->tau, k, theta, C, k);
   orc_gemm_nn(n, k, k, V, n, C, k, T, n);
   memcpy(V, T, sizeof(double) * (size_t)n * k);
   int it, converged = 0;
   for (it = 0; it < o->maxiter; ++it) {
      double *Xc = V;
      orc_spmm_csr(n, row_ptr, col_idx, vals, k, Xc, n, W, n);
      double normA = o->normA > 0 ? o->normA : theta[0];
      converged = 1;
      for (int j = 0; j < k; ++j) {
         double s = 0.0, t = 0.0;
         for (int64_t i = 0; i < n; ++i) {
            double r = fma(-theta[j], Xc[i + j * n], W[i + j * n]);
            Rr[i + j * n] = r;
            s += r * r;
            t += Xc[i + j * n] * Xc[i + j * n];
         }
         res[j] = sqrt(s) / (normA * sqrt(t));
         if (j < o->nwant && !(res[j] <= o->tol)) converged = 0;
      }
      if (converged) break;
      /* Chebyshev filter on the correction: damp [lower, theta_k] */
      double cut = theta[k - 1], lo = o->lower;
      double e = 0.5 * (cut - lo), c0 = 0.5 * (cut + lo);
      /* degree of this sweep: capped where the top-over-edge gain
       * (rho(L(theta_1)) / rho(L(theta_nwant)))^d would exceed 1e5 */
      int d = dmax;
      {
         const double l1 = (theta[0] - c0) / e, lw = (theta[o->nwant - 1] - c0) / e;
         const double r1 = l1 > 1.0 ? l1 + sqrt(l1 * l1 - 1.0) : 1.0;
         const double rw = lw > 1.0 ? lw + sqrt(lw * lw - 1.0) : 1.0;
         const double ratio = r1 / rw;
         if (ratio > 1.0) {
            const int cap = (int)floor(log(1e5) / log(ratio));
            d = cap < 2 ? 2 : (cap < dmax ? cap : dmax);
         }
      }
      const int mid = d / 2;
      /* soft locking: with the leading wanted pairs 0..pre-1 converged, the
       * sweep filters columns [c, k) only, ka = k - c halved from k while it
       * stays >= 16 and >= k - pre; the Ritz block keeps all k columns */
      int ka = k;
      if (o->lock) {
         int pre = 0;
         while (pre < o->nwant && res[pre] <= o->tol) ++pre;
         while (ka / 2 >= 16 && ka / 2 >= k - pre) ka /= 2;
      }
      const int c = k - ka, me = k + nd * ka;
      double *Dd = V + (int64_t)k * n, *Dm = V + (int64_t)(k + ka) * n;
      for (int jj = 0; jj < ka; ++jj) {
         const int j = c + jj;
         double s = colmaxabs(n, Rr + (int64_t)j * n) / e;
         double f = s > 0 ? 1.0 / s : 1.0;
         Lth[j] = (theta[j] - c0) / e;
         for (int64_t i = 0; i < n; ++i) {
            Dd[i + jj * n] = Rr[i + j * n] / e * f; /* D_1 */
            Dp[i + jj * n] = 0.0;                 /* D_0 */
         }
         zp[j] = f;            /* z_0 * f */
         zc[j] = Lth[j] * f;   /* z_1 * f */
      }
      if (nd == 2 && mid == 1) memcpy(Dm, Dd, sizeof(double) * (size_t)n * ka);
      for (int step = 1; step < d; ++step) {
         orc_spmm_csr(n, row_ptr, col_idx, vals, ka, Dd, n, T, n);
         for (int jj = 0; jj < ka; ++jj) {
            const int j = c + jj;
            double mx = 0.0;
            for (int64_t i = 0; i < n; ++i) {
               double t = (2.0 / e) * (T[i + jj * n] - c0 * Dd[i + jj * n]) - Dp[i + jj * n] +
                          (2.0 / e) * Rr[i + j * n] * zc[j];
               T[i + jj * n] = t;
               mx = fmax(mx, fabs(t));
            }
            double f = mx > 0 ? 1.0 / mx : 1.0;
            double znext = 2.0 * Lth[j] * zc[j] - zp[j];
            for (int64_t i = 0; i < n; ++i) {
               Dp[i + jj * n] = Dd[i + jj * n] * f;
               Dd[i + jj * n] = T[i + jj * n] * f;
            }
            zp[j] = zc[j] * f;
            zc[j] = znext * f;
         }
         if (nd == 2 && step + 1 == mid) memcpy(Dm, Dd, sizeof(double) * (size_t)n * ka);
      }
      /* projection products over V = [X | Dd (| Dmid)] */
      orc_spmm_csr(n, row_ptr, col_idx, vals, me - k, Dd, n, W + (int64_t)k * n, n);
      orc_gemm_tn(n, me, me, V, n, V, n, G, me);
      orc_gemm_tn(n, me, me, V, n, W, n, H, me);
      int r = o->rr_whole ? -1 : orc_rr_split(me, k, G, me, H, me, o->tau, k, theta, C, me);
      if (r < 0) r = orc_rr_truncated(me, G, me, H, me, o->tau, k, theta, C, me);
      if (r < k) { it = -(it + 1); break; }
      orc_gemm_nn(n, k, me, V, n, C, me, T, n);
      memcpy(V, T, sizeof(double) * (size_t)n * k);
   }
   if (!converged && it >= 0) {
      orc_residuals(n, row_ptr, col_idx, vals, k, V, n, theta, o->normA > 0 ? o->normA : theta[0], res);
   }
   for (int j = 0; j < k; ++j) memcpy(X + (int64_t)j * ldx, V + (int64_t)j * n, sizeof(double) * (size_t)n);
   free(V); free(W); free(Rr); free(Dp); free(T); free(G); free(H); free(C); free(zc); free(zp); free(Lth);
   if (it < 0) return it;
   return converged ? it : -(it + 1);
}
