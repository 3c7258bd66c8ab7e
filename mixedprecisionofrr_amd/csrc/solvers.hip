// Eigen-solver drivers over a CSR operator, composed from the library's
// kernels.  The reference declares these solvers as empty C++ stubs
// (CPP/linalg/solvers/eig/si_eig.cu:10-18, arnoldi_eig.cu:10-18,
// rr_eig.cu:9-34); their semantics come from the MATLAB prototype:
//   ofrr_arnoldi_hess_eig   MATLAB/+src/+krylov/arnoldi_hess_eig.m:43-107
//   ofrr_subspace_iter_eig  MATLAB/+src/+subspace/subspace_iter_eig.m:53-68
//   ofrr_filtered_eig       block subspace iteration with Chebyshev-filtered,
//                           low-precision residual corrections, FP64 Ritz block,
//                           FP64 projections and truncated Rayleigh-Ritz
//   ofrr_residuals          MATLAB/+tests/+test4/test.m:340-350 (normalised)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "internal.hpp"

namespace ofrr {

// ---- small device kernels used by the drivers ------------------------------
// per-column sums of (Ax - theta x)^2 and x^2 for strided X, partial per block
__global__ __launch_bounds__(256) void resid_cols_kernel(int64_t n, const double *__restrict__ AX,
                                                         int64_t ars, int64_t acs,
                                                         const double *__restrict__ X,
                                                         int64_t xrs, int64_t xcs,
                                                         const double *__restrict__ theta,
                                                         int64_t chunk, double *__restrict__ part) {
   __shared__ double s1[256], s2[256];
   const int col = blockIdx.y;
   const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(n, r0 + chunk);
   double a = 0.0, b = 0.0;
   const double th = theta[col];
   for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
      double x = X[r * xrs + (int64_t)col * xcs];
      double d = AX[r * ars + (int64_t)col * acs] - th * x;
      a = fma(d, d, a);
      b = fma(x, x, b);
   }
   s1[threadIdx.x] = a;
   s2[threadIdx.x] = b;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) { s1[threadIdx.x] += s1[threadIdx.x + o]; s2[threadIdx.x] += s2[threadIdx.x + o]; }
      __syncthreads();
   }
   if (threadIdx.x == 0) {
      part[(int64_t)blockIdx.x * 2 * gridDim.y + 2 * col] = s1[0];
      part[(int64_t)blockIdx.x * 2 * gridDim.y + 2 * col + 1] = s2[0];
   }
}

// the same sums for row-major X (row stride xrs) and AX (row stride k): thread
// t serves column t % k of the rows t / k, t / k + G, ... (G = 256 / k rows
// per pass, k <= 256), so a wave reads whole row segments; the G partial sums
// of a column are added in a fixed order
__global__ __launch_bounds__(256) void resid_rows_kernel(int64_t n, int k,
                                                         const double *__restrict__ AX,
                                                         const double *__restrict__ X, int64_t xrs,
                                                         const double *__restrict__ theta,
                                                         int64_t chunk, double *__restrict__ part) {
   __shared__ double s1[256], s2[256];
   const int G = 256 / k;
   const int t = threadIdx.x, c = t % k, g = t / k;
   const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(n, r0 + chunk);
   double a = 0.0, b = 0.0;
   if (g < G) {
      const double th = theta[c];
      for (int64_t r = r0 + g; r < r1; r += G) {
         const double x = X[r * xrs + c];
         const double d = AX[r * k + c] - th * x;
         a = fma(d, d, a);
         b = fma(x, x, b);
      }
   }
   s1[t] = a;
   s2[t] = b;
   __syncthreads();
   if (t < k) {
      double sa = 0.0, sb = 0.0;
      for (int q = 0; q < G; ++q) {
         sa += s1[q * k + t];
         sb += s2[q * k + t];
      }
      part[(int64_t)blockIdx.x * 2 * k + 2 * t] = sa;
      part[(int64_t)blockIdx.x * 2 * k + 2 * t + 1] = sb;
   }
}

// mirror the upper triangle of an m x m column-major matrix into its lower one
__global__ void mirror_upper_kernel(int m, double *__restrict__ G, int ldg) {
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= m * m) return;
   int i = t % m, j = t / m;
   if (i > j) G[i + (int64_t)j * ldg] = G[j + (int64_t)i * ldg];
}

// order-preserving map of a double onto uint64 (for atomicMax over any sign)
__device__ inline unsigned long long ord_bits(double v) {
   const unsigned long long u = (unsigned long long)__double_as_longlong(v);
   return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
static double ord_value(unsigned long long u) {
   const unsigned long long b = (u >> 63) ? (u & 0x7fffffffffffffffull) : ~u;
   double v;
   std::memcpy(&v, &b, 8);
   return v;
}

// upper Gershgorin bound max_i (a_ii + sum_{j != i} |a_ij|) of a symmetric CSR
// matrix (the top of the window the Chebyshev filter is normalised on) into
// out[0], and the bandwidth max |j - i| over the stored entries into out[1]
__global__ void gershgorin_upper_kernel(int64_t n, const int64_t *__restrict__ rp,
                                        const int32_t *__restrict__ ci,
                                        const double *__restrict__ va,
                                        unsigned long long *__restrict__ out) {
   double m = -INFINITY;
   unsigned long long bw = 0;
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (int64_t)gridDim.x * blockDim.x) {
      double s = 0.0;
      for (int64_t p = rp[i]; p < rp[i + 1]; ++p) {
         const int64_t j = ci[p];
         s += (j == i) ? va[p] : fabs(va[p]);
         bw = max(bw, (unsigned long long)(j > i ? j - i : i - j));
      }
      m = fmax(m, s);
   }
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      m = fmax(m, __shfl_xor(m, o, 64));
      bw = max(bw, (unsigned long long)__shfl_xor((long long)bw, o, 64));
   }
   if ((threadIdx.x & 63) == 0) {
      atomicMax(out, ord_bits(m));
      atomicMax(out + 1, bw);
   }
}

// ---- device arena for the drivers (grow-only, reused between solves) -------
struct Arena {
   char *base = nullptr;
   size_t bytes = 0, off = 0;
};
static thread_local Arena g_arena;
static thread_local int g_arena_dev = -1;
// caller-owned memory of the current solve on this thread (null: g_arena)
static thread_local char *g_ext_base = nullptr;
static thread_local size_t g_ext_off = 0;
// bytes the current solve reserved: a take past them returns null (and so does
// every later take, since the offset only grows), so checking the last pointer
// of a sequence of takes catches an undercounted reservation
static thread_local size_t g_reserved = 0;

static int arena_reserve(ofrr_ctx *ctx, size_t bytes) {
   if (ctx->state_mem) {
      // memory the caller handed over for this solve (the serving path)
      if (ctx->state_bytes < bytes)
         return set_error(ctx, kErrorOutOfMemory, "filtered_eig: caller state memory too small");
      g_ext_base = (char *)ctx->state_mem;
      g_ext_off = 0;
      g_reserved = bytes;
      return kSuccess;
   }
   g_ext_base = nullptr;
   g_reserved = bytes;
   if (g_arena.base && g_arena.bytes >= bytes && g_arena_dev == ctx->device) {
      g_arena.off = 0;
      return kSuccess;
   }
   if (g_arena.base) {
      OFRR_HIP(ctx, hipDeviceSynchronize());
      (void)hipFree(g_arena.base);
      g_arena = Arena();
   }
   OFRR_HIP(ctx, hipMalloc((void **)&g_arena.base, bytes));
   g_arena.bytes = bytes;
   g_arena.off = 0;
   g_arena_dev = ctx->device;
   return kSuccess;
}
static void *arena_take(size_t bytes) {
   size_t &off = g_ext_base ? g_ext_off : g_arena.off;
   char *base = g_ext_base ? g_ext_base : g_arena.base;
   const size_t a = (off + 255) & ~(size_t)255;
   off = a + bytes;
   return (off <= g_reserved) ? base + a : nullptr;
}

static inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_residuals(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rp, const int32_t *ci,
                   const double *va, int k, const double *X, int64_t xrs, int64_t xcs,
                   const double *theta_host, double normA, double *res_host) {
   OFRR_CHECK_CTX(ctx);
   if (n <= 0 || k <= 0 || !rp || !ci || !va || !X || !theta_host || !res_host)
      return set_error(ctx, kErrorInvalidArgument, "residuals: bad arguments");
   (void)nnz;
   const bool rowmajor = xcs == 1 && k <= 256;
   int64_t chunk = rowmajor ? 2048 : 16384;
   int nch = (int)((n + chunk - 1) / chunk);
   size_t need = al((size_t)n * k * 8) + al((size_t)k * 8) + al((size_t)nch * 2 * k * 8) + al((size_t)2 * k * 8);
   char *ws = (char *)workspace(ctx, need);
   double *hs = (double *)pinned(ctx, (size_t)2 * k * 8);
   if (!ws || !hs) return set_error(ctx, kErrorCUDAMemory, "residuals: workspace");
   double *AX = (double *)ws;
   double *th = (double *)(ws + al((size_t)n * k * 8));
   double *part = (double *)((char *)th + al((size_t)k * 8));
   double *sums = (double *)((char *)part + al((size_t)nch * 2 * k * 8));
   OFRR_HIP(ctx, hipMemcpyAsync(th, theta_host, (size_t)k * 8, hipMemcpyHostToDevice, ctx->stream));
   // A X in the layout of X (row-major X takes the row-gather SpMM)
   const int64_t ars = (xcs == 1) ? k : 1, acs = (xcs == 1) ? 1 : n;
   int st = spmm_csr(ctx, n, rp, ci, va, OFRR_F64, k, X, OFRR_F64, xrs, xcs, 1.0, 0.0, AX, OFRR_F64, ars, acs);
   if (st) return st;
   if (rowmajor)
      resid_rows_kernel<<<nch, 256, 0, ctx->stream>>>(n, k, AX, X, xrs, th, chunk, part);
   else
      resid_cols_kernel<<<dim3(nch, k), 256, 0, ctx->stream>>>(n, AX, ars, acs, X, xrs, xcs, th, chunk, part);
   OFRR_LAUNCH(ctx);
   if ((st = reduce_partials(ctx, nch, 2 * k, part, sums))) return st;
   OFRR_HIP(ctx, hipMemcpyAsync(hs, sums, (size_t)2 * k * 8, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int j = 0; j < k; ++j) res_host[j] = std::sqrt(hs[2 * j]) / (normA * std::sqrt(hs[2 * j + 1]));
   return kSuccess;
}

// ---------------------------------------------------------------------------
// Arnoldi-Hessenberg (arnoldi_hess_eig.m:43-107) with the Krylov basis stored
// in vdt (precision_output: FP64, FP32, FP16 or BF16), its arithmetic in FP64
// for FP64 and FP32 otherwise (precision_compute), the operator products of
// the sweep accumulated in that arithmetic (precision_A_matvec_*), and the
// Rayleigh-Ritz in FP64: the basis widened to FP64, A V, V^T V and V^T A V
// with FP64 accumulation, eig(V^T A V, V^T V + gtol I).  Each new column is
// one left-looking Hessenberg step against the earlier ones (the pivot-row
// forward substitution, ofrr_hessenberg_step), so a sweep of kdim columns runs
// on the device without a host round trip per column: a column whose pivot is
// below select_tol is flagged there, and one read after the sweep finds where
// the process stopped (the columns after it are never used).
static int arnoldi_impl(ofrr_ctx *ctx, int vdt, int64_t n, const int64_t *rp, const int32_t *ci,
                        const double *va, int k, int kdim, int maxiter, double select_tol, double gtol,
                        const double *v0, double *theta_host, double *X, int64_t ldx, int *built_out) {
   if (n <= 0 || k <= 0 || kdim <= 0 || kdim + 1 > 384 || !rp || !ci || !va || !v0 || !theta_host || !X || ldx < n)
      return set_error(ctx, kErrorInvalidArgument, "arnoldi_hess_eig: bad arguments");
   if (vdt != OFRR_F64 && vdt != OFRR_F32 && vdt != OFRR_F16 && vdt != OFRR_BF16)
      return set_error(ctx, kErrorBLASUnsupportedType, "arnoldi_hess_eig: basis FP64, FP32, FP16 or BF16");
   if (k > n) k = (int)n;
   const int m = kdim + 1;
   const bool wide = vdt != OFRR_F64;
   const size_t es = dtype_size(vdt);
   size_t need = al((size_t)n * m * es) + al((size_t)n * m * 8) * (wide ? 3 : 2) + al((size_t)m * m * 8) * 4 +
                 al((size_t)m * 8) + al((size_t)m * 4) * 2;
   int st = arena_reserve(ctx, need);
   if (st) return st;
   void *V = arena_take((size_t)n * m * es);
   double *V64 = wide ? (double *)arena_take((size_t)n * m * 8) : (double *)V;
   double *AV = (double *)arena_take((size_t)n * m * 8);
   double *U = (double *)arena_take((size_t)n * m * 8);
   double *G = (double *)arena_take((size_t)m * m * 8);
   double *H = (double *)arena_take((size_t)m * m * 8);
   double *Y = (double *)arena_take((size_t)m * m * 8);
   double *R = (double *)arena_take((size_t)m * m * 8);
   double *th = (double *)arena_take((size_t)m * 8);
   int32_t *prow1 = (int32_t *)arena_take((size_t)m * 4);
   int32_t *skip = (int32_t *)arena_take((size_t)m * 4);
   if (!skip) return set_error(ctx, kErrorInternalError, "arnoldi: state layout exceeds its reservation");
   double *hp = (double *)pinned(ctx, (size_t)m * 8 + (size_t)m * 4 + 64);
   if (!hp) return set_error(ctx, kErrorCUDAMemory, "arnoldi: pinned");
   int32_t *hskip = (int32_t *)(hp + m);
   const int64_t ldv = n;
   auto col = [&](int j) { return (char *)V + (size_t)j * ldv * es; };
   if ((st = convert(ctx, n, 1, v0, OFRR_F64, 1, n, V, vdt, 1, ldv))) return st;
   int tits = 1, built = 0, have = 0;
   while (tits < maxiter) {
      // first column: its pivot and scale (a tiny start vector ends the solve)
      int32_t dropped = 0;
      if ((st = ofrr_hessenberg_step(ctx, 0, vdt, n, m, V, 1, ldv, 0, select_tol, R, m, prow1, skip, &dropped)))
         return st;
      if (dropped) break;
      for (int i = 0; i < kdim; ++i) {
         // w = A v_i in the basis arithmetic, stored in the basis format
         if ((st = spmm_csr(ctx, n, rp, ci, va, OFRR_F64, 1, col(i), vdt, 1, ldv, 1.0, 0.0, col(i + 1), vdt, 1, ldv)))
            return st;
         if ((st = ofrr_hessenberg_step(ctx, 0, vdt, n, m, V, 1, ldv, i + 1, select_tol, R, m, prow1, skip, nullptr)))
            return st;
      }
      OFRR_HIP(ctx, hipMemcpyAsync(hskip, skip, (size_t)m * 4, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      int ncols = m;
      for (int j = 1; j < m; ++j)
         if (hskip[j]) { ncols = j; break; }
      if (ncols < m) tits = maxiter;  // the process stopped: this is the last Rayleigh-Ritz
      else tits += kdim;
      built = ncols;
      // Rayleigh-Ritz in FP64
      if (wide && (st = convert(ctx, n, ncols, V, vdt, 1, ldv, V64, OFRR_F64, 1, n))) return st;
      if ((st = spmm_csr(ctx, n, rp, ci, va, OFRR_F64, ncols, V64, OFRR_F64, 1, n, 1.0, 0.0, AV, OFRR_F64, 1, n))) return st;
      if ((st = gram(ctx, n, ncols, ncols, V64, OFRR_F64, 1, n, AV, OFRR_F64, 1, n, 1.0, 0.0, H, ncols))) return st;
      if ((st = gram(ctx, n, ncols, ncols, V64, OFRR_F64, 1, n, V64, OFRR_F64, 1, n, 1.0, 0.0, G, ncols))) return st;
      st = rr_generalized(ctx, ncols, G, ncols, H, ncols, gtol, th, Y, ncols);
      if (st == kErrorSingularMatrix) break;
      if (st) return st;
      {
         const void *vb[1] = {V64};
         int dt[1] = {OFRR_F64}, kb[1] = {ncols}, ldc[1] = {ncols};
         int64_t rs[1] = {1}, cs[1] = {n};
         const double *cb[1] = {Y};
         if ((st = basis_update(ctx, n, ncols, 1, vb, dt, kb, rs, cs, cb, ldc, U, OFRR_F64, 1, n))) return st;
      }
      have = ncols;
      // restart from the top Ritz vector, in the basis format
      if (tits < maxiter && (st = convert(ctx, n, 1, U, OFRR_F64, 1, n, V, vdt, 1, ldv))) return st;
   }
   int kk = std::min(k, have);
   if (kk > 0) {
      OFRR_HIP(ctx, hipMemcpyAsync(hp, th, (size_t)kk * 8, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipMemcpy2DAsync(X, (size_t)ldx * 8, U, (size_t)n * 8, (size_t)n * 8, kk,
                                     hipMemcpyDeviceToDevice, ctx->stream));
   }
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int c = 0; c < k; ++c) theta_host[c] = (c < kk) ? hp[c] : NAN;
   if (built_out) *built_out = built;
   return kSuccess;
}

int ofrr_arnoldi_hess_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rp,
                          const int32_t *ci, const double *va, int k, int kdim, int maxiter,
                          double select_tol, double gtol, const double *v0, double *theta_host,
                          double *X, int64_t ldx, int *built_out) {
   OFRR_CHECK_CTX(ctx);
   (void)nnz;
   return arnoldi_impl(ctx, OFRR_F64, n, rp, ci, va, k, kdim, maxiter, select_tol, gtol, v0, theta_host, X, ldx,
                       built_out);
}

int ofrr_arnoldi_hess_eig_mp(ofrr_ctx *ctx, int basis_dtype, int64_t n, int64_t nnz, const int64_t *rp,
                             const int32_t *ci, const double *va, int k, int kdim, int maxiter,
                             double select_tol, double gtol, const double *v0, double *theta_host,
                             double *X, int64_t ldx, int *built_out) {
   OFRR_CHECK_CTX(ctx);
   (void)nnz;
   return arnoldi_impl(ctx, basis_dtype, n, rp, ci, va, k, kdim, maxiter, select_tol, gtol, v0, theta_host, X,
                       ldx, built_out);
}

// ---------------------------------------------------------------------------
int ofrr_subspace_iter_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rp,
                           const int32_t *ci, const double *va, int m, int maxiter, int step_size,
                           double select_tol, double gtol, double *Vout, int64_t ldv,
                           double *theta_host) {
   OFRR_CHECK_CTX(ctx);
   (void)nnz;
   if (n <= 0 || m <= 0 || m > 256 || m > n || !rp || !ci || !va || !Vout || ldv < n || !theta_host)
      return set_error(ctx, kErrorInvalidArgument, "subspace_iter_eig: bad arguments");
   size_t need = al((size_t)n * m * 8) * 2 + al((size_t)m * m * 8) * 4 + al((size_t)m * 8) * 2 + al((size_t)m * 4) * 2;
   int st = arena_reserve(ctx, need);
   if (st) return st;
   double *T = (double *)arena_take((size_t)n * m * 8);
   double *AV = (double *)arena_take((size_t)n * m * 8);
   double *R = (double *)arena_take((size_t)m * m * 8);
   double *G = (double *)arena_take((size_t)m * m * 8);
   double *H = (double *)arena_take((size_t)m * m * 8);
   double *Y = (double *)arena_take((size_t)m * m * 8);
   double *th = (double *)arena_take((size_t)m * 8);
   double *piv = (double *)arena_take((size_t)m * 8);
   int32_t *perm = (int32_t *)arena_take((size_t)m * 4);
   int32_t *skip = (int32_t *)arena_take((size_t)m * 4);
   if (!skip) return set_error(ctx, kErrorInternalError, "subspace: state layout exceeds its reservation");
   int32_t *hskip = (int32_t *)pinned(ctx, (size_t)m * 4 + (size_t)m * 8);
   if (!hskip) return set_error(ctx, kErrorCUDAMemory, "subspace: pinned");
   double *hth = (double *)(hskip + m);
   OFRR_HIP(ctx, hipMemcpy2DAsync(T, (size_t)n * 8, Vout, (size_t)ldv * 8, (size_t)n * 8, m,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   if ((st = ofrr_scalecols(ctx, OFRR_F64, n, m, T, 1, n, piv))) return st;
   int ms = m;
   for (int it = 0; it < maxiter; ++it) {
      for (int s = 0; s < step_size; ++s) {
         if ((st = spmm_csr(ctx, n, rp, ci, va, OFRR_F64, ms, T, OFRR_F64, 1, n, 1.0, 0.0, AV, OFRR_F64, 1, n))) return st;
         std::swap(T, AV);
         if ((st = ofrr_scalecols(ctx, OFRR_F64, n, ms, T, 1, n, piv))) return st;
      }
      if ((st = ofrr_hessenberg(ctx, OFRR_F64, n, ms, T, 1, n, select_tol, R, m, perm, skip))) return st;
      OFRR_HIP(ctx, hipMemcpyAsync(hskip, skip, (size_t)ms * 4, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      // economic Q: kept columns only (hessenberg.m:76-78)
      int keep = 0;
      for (int j = 0; j < ms; ++j)
         if (!hskip[j]) {
            if (keep != j)
               OFRR_HIP(ctx, hipMemcpyAsync(T + (int64_t)keep * n, T + (int64_t)j * n, (size_t)n * 8,
                                            hipMemcpyDeviceToDevice, ctx->stream));
            ++keep;
         }
      ms = keep;
      if (ms == 0) return set_error(ctx, kErrorFactorizationFailed, "subspace_iter_eig: every column dropped");
      if ((st = spmm_csr(ctx, n, rp, ci, va, OFRR_F64, ms, T, OFRR_F64, 1, n, 1.0, 0.0, AV, OFRR_F64, 1, n))) return st;
      if ((st = gram(ctx, n, ms, ms, T, OFRR_F64, 1, n, AV, OFRR_F64, 1, n, 1.0, 0.0, H, ms))) return st;
      if ((st = gram(ctx, n, ms, ms, T, OFRR_F64, 1, n, T, OFRR_F64, 1, n, 1.0, 0.0, G, ms))) return st;
      if ((st = rr_generalized(ctx, ms, G, ms, H, ms, gtol, th, Y, ms))) return st;
      const void *vb[1] = {T};
      int dt[1] = {OFRR_F64}, kb[1] = {ms}, ldc[1] = {ms};
      int64_t rs[1] = {1}, cs[1] = {n};
      const double *cb[1] = {Y};
      if ((st = basis_update(ctx, n, ms, 1, vb, dt, kb, rs, cs, cb, ldc, AV, OFRR_F64, 1, n))) return st;
      std::swap(T, AV);
   }
   OFRR_HIP(ctx, hipMemcpy2DAsync(Vout, (size_t)ldv * 8, T, (size_t)n * 8, (size_t)n * 8, ms,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   if (ms < m) {
      for (int j = ms; j < m; ++j)
         OFRR_HIP(ctx, hipMemsetAsync(Vout + (int64_t)j * ldv, 0, (size_t)n * 8, ctx->stream));
   }
   OFRR_HIP(ctx, hipMemcpyAsync(hth, th, (size_t)ms * 8, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int j = 0; j < m; ++j) theta_host[j] = (j < ms) ? hth[j] : NAN;
   return kSuccess;
}

// ---- two-sided Rayleigh-Ritz for the SVD and its subspace-iteration driver
// (svd_rr.m:54-91, subspace_iter_svd.m:57-77) over a dense FP64 A (M x N,
// column-major).  A V and A^T U are tall products on the back-transform
// kernel (A^T is A read with swapped strides), U^T A V, U^T U and V^T V are
// Gram products, and the augmented pencil goes to the Cholesky-reduced FP64
// eigensolver.

}  // extern "C"

namespace ofrr {

// Ae = [0 B; B^T 0] and Me = blkdiag(Gu, Gv) (generalised) or I, K = ku + kv
__global__ void svd_aug_kernel(int ku, int kv, const double *__restrict__ B,
                               const double *__restrict__ Gu, const double *__restrict__ Gv,
                               int generalized, double *__restrict__ Ae, double *__restrict__ Me) {
   const int K = ku + kv;
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= K * K) return;
   const int i = t % K, j = t / K;
   double a = 0.0, m = 0.0;
   if (i < ku && j >= ku) a = B[i + (int64_t)(j - ku) * ku];
   else if (i >= ku && j < ku) a = B[j + (int64_t)(i - ku) * ku];
   if (generalized) {
      if (i < ku && j < ku) m = Gu[i + (int64_t)j * ku];
      else if (i >= ku && j >= ku) m = Gv[(i - ku) + (int64_t)(j - ku) * kv];
   } else {
      m = (i == j) ? 1.0 : 0.0;
   }
   Ae[t] = a;
   Me[t] = m;
}

// C(0:rows, 0:kk) = sqrt(2) Y(off:off+rows, 0:kk), Y column-major ld K
__global__ void svd_coef_kernel(int rows, int kk, int off, int K, const double *__restrict__ Y,
                                double *__restrict__ C) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= rows * kk) return;
   const int i = t % rows, c = t / rows;
   C[t] = 1.4142135623730951 * Y[(off + i) + (int64_t)c * K];
}

// the two-sided RR on device buffers: U (M x ku, ld M) and V (N x kv, ld N) are
// replaced by their first kk = min(ku, kv) Ritz vectors; sigma (kk) on the device
static int svd_rr_dev(ofrr_ctx *ctx, int64_t M, int64_t N, const double *A, int64_t lda, int ku,
                      int kv, double *U, double *V, int generalized, double gtol, double *sigma,
                      double *work_tall /* max(M, N) x max(ku, kv) */, double *small) {
   const int K = ku + kv, kk = std::min(ku, kv);
   int st;
   double *B = small, *Gu = B + (size_t)ku * kv, *Gv = Gu + (size_t)ku * ku;
   double *Ae = Gv + (size_t)kv * kv, *Me = Ae + (size_t)K * K, *Y = Me + (size_t)K * K;
   double *Cu = Y + (size_t)K * K, *Cv = Cu + (size_t)ku * kk;
   // AV = A V into work_tall (M x kv)
   {
      const void *vb[1] = {A};
      int dt[1] = {OFRR_F64}, kb[1] = {(int)N}, ldc[1] = {(int)N};
      int64_t rs[1] = {1}, cs[1] = {lda};
      const double *cb[1] = {V};
      if ((st = basis_update(ctx, M, kv, 1, vb, dt, kb, rs, cs, cb, ldc, work_tall, OFRR_F64, 1, M))) return st;
   }
   if ((st = gram(ctx, M, ku, kv, U, OFRR_F64, 1, M, work_tall, OFRR_F64, 1, M, 1.0, 0.0, B, ku))) return st;
   if (generalized) {
      if ((st = gram(ctx, M, ku, ku, U, OFRR_F64, 1, M, U, OFRR_F64, 1, M, 1.0, 0.0, Gu, ku))) return st;
      if ((st = gram(ctx, N, kv, kv, V, OFRR_F64, 1, N, V, OFRR_F64, 1, N, 1.0, 0.0, Gv, kv))) return st;
   }
   svd_aug_kernel<<<(K * K + 255) / 256, 256, 0, ctx->stream>>>(ku, kv, B, Gu, Gv, generalized, Ae, Me);
   OFRR_LAUNCH(ctx);
   if ((st = rr_generalized(ctx, K, Me, K, Ae, K, generalized ? gtol : 0.0, sigma, Y, K))) return st;
   svd_coef_kernel<<<(ku * kk + 255) / 256, 256, 0, ctx->stream>>>(ku, kk, 0, K, Y, Cu);
   svd_coef_kernel<<<(kv * kk + 255) / 256, 256, 0, ctx->stream>>>(kv, kk, ku, K, Y, Cv);
   OFRR_LAUNCH(ctx);
   // U <- U Cu (through work_tall), V <- V Cv
   {
      const void *vb[1] = {U};
      int dt[1] = {OFRR_F64}, kb[1] = {ku}, ldc[1] = {ku};
      int64_t rs[1] = {1}, cs[1] = {M};
      const double *cb[1] = {Cu};
      if ((st = basis_update(ctx, M, kk, 1, vb, dt, kb, rs, cs, cb, ldc, work_tall, OFRR_F64, 1, M))) return st;
      OFRR_HIP(ctx, hipMemcpyAsync(U, work_tall, (size_t)M * kk * 8, hipMemcpyDeviceToDevice, ctx->stream));
   }
   {
      const void *vb[1] = {V};
      int dt[1] = {OFRR_F64}, kb[1] = {kv}, ldc[1] = {kv};
      int64_t rs[1] = {1}, cs[1] = {N};
      const double *cb[1] = {Cv};
      if ((st = basis_update(ctx, N, kk, 1, vb, dt, kb, rs, cs, cb, ldc, work_tall, OFRR_F64, 1, N))) return st;
      OFRR_HIP(ctx, hipMemcpyAsync(V, work_tall, (size_t)N * kk * 8, hipMemcpyDeviceToDevice, ctx->stream));
   }
   return kSuccess;
}

static size_t svd_small_doubles(int ku, int kv) {
   const int K = ku + kv, kk = std::min(ku, kv);
   return (size_t)ku * kv + (size_t)ku * ku + (size_t)kv * kv + 3 * (size_t)K * K +
          (size_t)(ku + kv) * kk + K;
}

// keep the columns of T (rows x kc, ld rows) whose skip flag is 0; returns the kept width
static int compact_kept(ofrr_ctx *ctx, double *T, int64_t rows, int kc, const int32_t *skip_dev,
                        int32_t *hskip, int *kept) {
   OFRR_HIP(ctx, hipMemcpyAsync(hskip, skip_dev, (size_t)kc * 4, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   int keep = 0;
   for (int j = 0; j < kc; ++j)
      if (!hskip[j]) {
         if (keep != j)
            OFRR_HIP(ctx, hipMemcpyAsync(T + (int64_t)keep * rows, T + (int64_t)j * rows, (size_t)rows * 8,
                                         hipMemcpyDeviceToDevice, ctx->stream));
         ++keep;
      }
   *kept = keep;
   return kSuccess;
}

}  // namespace ofrr

extern "C" {

int ofrr_svd_rr(ofrr_ctx *ctx, int64_t M, int64_t N, const double *A, int64_t lda, int ku, int kv,
                double *U, int64_t ldu, double *V, int64_t ldv, int generalized, double gtol,
                double *sigma_host) {
   OFRR_CHECK_CTX(ctx);
   if (M <= 0 || N <= 0 || ku < 1 || kv < 1 || ku + kv > 256 || ku > M || kv > N || lda < M ||
       ldu < M || ldv < N || !A || !U || !V || !sigma_host)
      return set_error(ctx, kErrorInvalidArgument, "svd_rr: bad arguments");
   const int kk = std::min(ku, kv);
   const int64_t tall = std::max(M, N) * std::max(ku, kv);
   size_t need = al((size_t)M * ku * 8) + al((size_t)N * kv * 8) + al((size_t)tall * 8) +
                 al(svd_small_doubles(ku, kv) * 8) + al((size_t)(ku + kv) * 8);
   int st = arena_reserve(ctx, need);
   if (st) return st;
   double *Ud = (double *)arena_take((size_t)M * ku * 8);
   double *Vd = (double *)arena_take((size_t)N * kv * 8);
   double *wt = (double *)arena_take((size_t)tall * 8);
   double *small = (double *)arena_take(svd_small_doubles(ku, kv) * 8);
   double *sig = (double *)arena_take((size_t)(ku + kv) * 8);
   if (!sig) return set_error(ctx, kErrorInternalError, "svd_rr: state layout exceeds its reservation");
   OFRR_HIP(ctx, hipMemcpy2DAsync(Ud, (size_t)M * 8, U, (size_t)ldu * 8, (size_t)M * 8, ku,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   OFRR_HIP(ctx, hipMemcpy2DAsync(Vd, (size_t)N * 8, V, (size_t)ldv * 8, (size_t)N * 8, kv,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   if ((st = svd_rr_dev(ctx, M, N, A, lda, ku, kv, Ud, Vd, generalized, gtol, sig, wt, small))) return st;
   OFRR_HIP(ctx, hipMemcpy2DAsync(U, (size_t)ldu * 8, Ud, (size_t)M * 8, (size_t)M * 8, kk,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   OFRR_HIP(ctx, hipMemcpy2DAsync(V, (size_t)ldv * 8, Vd, (size_t)N * 8, (size_t)N * 8, kk,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   double *h = (double *)pinned(ctx, (size_t)kk * 8);
   if (!h) return set_error(ctx, kErrorCUDAMemory, "svd_rr: pinned");
   OFRR_HIP(ctx, hipMemcpyAsync(h, sig, (size_t)kk * 8, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int j = 0; j < kk; ++j) sigma_host[j] = h[j];
   return kSuccess;
}

int ofrr_subspace_iter_svd(ofrr_ctx *ctx, int64_t M, int64_t N, const double *A, int64_t lda, int k,
                           int maxiter, int step_size, double select_tol, int generalized,
                           double gtol, const double *V0, int64_t ldv0, double *U, int64_t ldu,
                           double *V, int64_t ldv, double *sigma_host, int *kept) {
   OFRR_CHECK_CTX(ctx);
   if (M <= 0 || N <= 0 || k < 1 || 2 * k > 256 || k > M || k > N || lda < M || ldv0 < N || ldu < M ||
       ldv < N || step_size < 1 || maxiter < 0 || !A || !V0 || !U || !V || !sigma_host)
      return set_error(ctx, kErrorInvalidArgument, "subspace_iter_svd: bad arguments");
   const int64_t tall = std::max(M, N) * k;
   size_t need = al((size_t)M * k * 8) + al((size_t)N * k * 8) + al((size_t)tall * 8) +
                 al(svd_small_doubles(k, k) * 8) + al((size_t)2 * k * 8) + al((size_t)k * k * 8) +
                 al((size_t)k * 8) + al((size_t)k * 4) * 2;
   int st = arena_reserve(ctx, need);
   if (st) return st;
   double *Tu = (double *)arena_take((size_t)M * k * 8);
   double *Tv = (double *)arena_take((size_t)N * k * 8);
   double *wt = (double *)arena_take((size_t)tall * 8);
   double *small = (double *)arena_take(svd_small_doubles(k, k) * 8);
   double *sig = (double *)arena_take((size_t)2 * k * 8);
   double *R = (double *)arena_take((size_t)k * k * 8);
   double *piv = (double *)arena_take((size_t)k * 8);
   int32_t *perm = (int32_t *)arena_take((size_t)k * 4);
   int32_t *skip = (int32_t *)arena_take((size_t)k * 4);
   if (!skip) return set_error(ctx, kErrorInternalError, "subspace_iter_svd: state layout exceeds its reservation");
   int32_t *hskip = (int32_t *)pinned(ctx, (size_t)k * 4 + (size_t)k * 8);
   if (!hskip) return set_error(ctx, kErrorCUDAMemory, "subspace_iter_svd: pinned");
   double *hsig = (double *)(hskip + k);
   OFRR_HIP(ctx, hipMemsetAsync(Tu, 0, (size_t)M * k * 8, ctx->stream));
   OFRR_HIP(ctx, hipMemcpy2DAsync(Tv, (size_t)N * 8, V0, (size_t)ldv0 * 8, (size_t)N * 8, k,
                                  hipMemcpyDeviceToDevice, ctx->stream));
   if ((st = ofrr_scalecols(ctx, OFRR_F64, N, k, Tv, 1, N, piv))) return st;
   int kc = k;
   for (int it = 0; it < maxiter && kc > 0; ++it) {
      for (int s = 0; s < step_size; ++s) {
         // Tu = A Tv, Tv = A^T Tu (A^T: element (i, j) = A[j + i lda])
         const void *va[1] = {A};
         int dt[1] = {OFRR_F64}, kbn[1] = {(int)N}, kbm[1] = {(int)M}, ldn[1] = {(int)N}, ldm[1] = {(int)M};
         int64_t rs1[1] = {1}, cs1[1] = {lda}, rs2[1] = {lda}, cs2[1] = {1};
         const double *cv[1] = {Tv}, *cu[1] = {Tu};
         if ((st = basis_update(ctx, M, kc, 1, va, dt, kbn, rs1, cs1, cv, ldn, Tu, OFRR_F64, 1, M))) return st;
         if ((st = ofrr_scalecols(ctx, OFRR_F64, M, kc, Tu, 1, M, piv))) return st;
         if ((st = basis_update(ctx, N, kc, 1, va, dt, kbm, rs2, cs2, cu, ldm, Tv, OFRR_F64, 1, N))) return st;
         if ((st = ofrr_scalecols(ctx, OFRR_F64, N, kc, Tv, 1, N, piv))) return st;
      }
      int ku = 0, kv = 0;
      if ((st = ofrr_hessenberg(ctx, OFRR_F64, M, kc, Tu, 1, M, select_tol, R, k, perm, skip))) return st;
      if ((st = compact_kept(ctx, Tu, M, kc, skip, hskip, &ku))) return st;
      if ((st = ofrr_hessenberg(ctx, OFRR_F64, N, kc, Tv, 1, N, select_tol, R, k, perm, skip))) return st;
      if ((st = compact_kept(ctx, Tv, N, kc, skip, hskip, &kv))) return st;
      if (ku == 0 || kv == 0) { kc = 0; break; }
      if ((st = svd_rr_dev(ctx, M, N, A, lda, ku, kv, Tu, Tv, generalized, gtol, sig, wt, small))) return st;
      kc = std::min(ku, kv);
   }
   if (kc > 0) {
      OFRR_HIP(ctx, hipMemcpy2DAsync(U, (size_t)ldu * 8, Tu, (size_t)M * 8, (size_t)M * 8, kc,
                                     hipMemcpyDeviceToDevice, ctx->stream));
      OFRR_HIP(ctx, hipMemcpy2DAsync(V, (size_t)ldv * 8, Tv, (size_t)N * 8, (size_t)N * 8, kc,
                                     hipMemcpyDeviceToDevice, ctx->stream));
      OFRR_HIP(ctx, hipMemcpyAsync(hsig, sig, (size_t)kc * 8, hipMemcpyDeviceToHost, ctx->stream));
   }
   for (int j = kc; j < k; ++j) {
      OFRR_HIP(ctx, hipMemsetAsync(U + (int64_t)j * ldu, 0, (size_t)M * 8, ctx->stream));
      OFRR_HIP(ctx, hipMemsetAsync(V + (int64_t)j * ldv, 0, (size_t)N * 8, ctx->stream));
   }
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int j = 0; j < k; ++j) sigma_host[j] = (j < kc) ? hsig[j] : NAN;
   if (kept) *kept = kc;
   return kc > 0 ? kSuccess : set_error(ctx, kErrorFactorizationFailed, "subspace_iter_svd: every column dropped");
}

}  // extern "C"

namespace ofrr {

constexpr double kFilterGain = 1e5;  // largest top-over-edge gain of one sweep

// The operator of the filtered solver.  The CSR operator fuses each step into
// one SpMM kernel; the dense operators (symmetric A, and A^T A for the SVD) run
// MFMA products followed by the shared epilogue kernels.
struct FilterOp {
   int64_t n = 0;
   virtual ~FilterOp() = default;
   // an upper bound of the spectrum, or <= 0 when none is cheap to get
   virtual int upper(ofrr_ctx *ctx, double *ub) = 0;
   // W = A X, FP64, row-major n x k
   virtual int apply64(ofrr_ctx *ctx, int k, const double *X, double *W) = 0;
   // W = A X, R = W - X theta quantised into Rq (scale 1 / sR), column max |R|,
   // FP64 sums [r^2, x^2] per column
   virtual int resid(ofrr_ctx *ctx, int k, const double *X, const double *th, double *W, void *Rq,
                     int bdt, const double *inv_sR, float *colmax, double *sums) = 0;
   // one Chebyshev step
   virtual int cheb(ofrr_ctx *ctx, int k, const void *Dc, const void *Dp, const void *R, int bdt,
                    const ChebArgs &args, void *Dn) = 0;
   // Z (row stride zrs) = A D in the growth arithmetic
   virtual int apply_lo(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt,
                        int64_t zrs) = 0;
   // whether the growth products may change format mid-solve (the precision
   // ladder), and the switch to the FP64 operator values for an FP64 format
   virtual bool can_refine() const { return false; }
   virtual void set_precise(bool) {}
   // whether a sweep may filter a trailing column range only (soft locking)
   virtual bool can_lock() const { return false; }
   // whether apply_lo may run on the context's second stream (its launches
   // take their stream from the context and it keeps no scratch of its own)
   virtual bool can_overlap() const { return false; }
   // the narrowest block storage the operator's products take
   virtual int min_width() const { return 1; }
   // whether a sweep's launches may be recorded as a graph, and what of the
   // operator a recorded sweep depends on (its buffers and value format)
   virtual bool can_graph() const { return false; }
   virtual void sweep_sig(std::vector<uint64_t> &) const {}
};

struct CsrOp final : FilterOp {
   const int64_t *rp;
   const int32_t *ci;
   const double *va;
   const void *vlo;
   int vlo_dt;
   int64_t bw = 0;  // bandwidth max |j - i|, found with the Gershgorin bound
   int upper(ofrr_ctx *ctx, double *ub) override {
      unsigned long long *gub = (unsigned long long *)workspace(ctx, 16);
      if (!gub) return set_error(ctx, kErrorCUDAMemory, "filtered_eig: workspace");
      OFRR_HIP(ctx, hipMemsetAsync(gub, 0, 16, ctx->stream));
      gershgorin_upper_kernel<<<1024, 256, 0, ctx->stream>>>(n, rp, ci, va, gub);
      OFRR_LAUNCH(ctx);
      unsigned long long u[2] = {0, 0};
      OFRR_HIP(ctx, hipMemcpyAsync(u, gub, 16, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      *ub = ord_value(u[0]);
      bw = (int64_t)u[1];
      return kSuccess;
   }
   int apply64(ofrr_ctx *ctx, int k, const double *X, double *W) override {
      return spmm_csr(ctx, n, rp, ci, va, OFRR_F64, k, X, OFRR_F64, k, 1, 1.0, 0.0, W, OFRR_F64, k, 1, bw);
   }
   int resid(ofrr_ctx *ctx, int k, const double *X, const double *th, double *W, void *Rq, int bdt,
             const double *inv_sR, float *colmax, double *sums) override {
      return spmm_resid(ctx, n, rp, ci, va, k, X, th, W, Rq, bdt, inv_sR, colmax, sums, bw);
   }
   bool precise = false;  // growth products on the FP64 values
   int cheb(ofrr_ctx *ctx, int k, const void *Dc, const void *Dp, const void *R, int bdt,
            const ChebArgs &args, void *Dn) override {
      return cheb_step(ctx, n, rp, ci, precise ? (const void *)va : vlo, precise ? OFRR_F64 : vlo_dt,
                       k, Dc, Dp, R, bdt, args, Dn, bw);
   }
   int apply_lo(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt, int64_t zrs) override {
      return spmm_csr(ctx, n, rp, ci, precise ? (const void *)va : vlo, precise ? OFRR_F64 : vlo_dt,
                      k, D, bdt, k, 1, 1.0, 0.0, Z, zdt, zrs, 1, bw);
   }
   bool can_refine() const override { return true; }
   void set_precise(bool p) override { precise = p; }
   bool can_lock() const override { return true; }
   bool can_overlap() const override { return true; }
   bool can_graph() const override { return true; }
   void sweep_sig(std::vector<uint64_t> &v) const override {
      v.insert(v.end(), {(uint64_t)(uintptr_t)rp, (uint64_t)(uintptr_t)ci,
                         (uint64_t)(uintptr_t)(precise ? (const void *)va : vlo),
                         (uint64_t)(precise ? OFRR_F64 : vlo_dt), (uint64_t)bw});
   }
};

// Y (n x wr 8-byte words, row-major) = words [c0, c0 + wr) of every row of X
// (row stride ws words): the trailing columns of a row-major block, compacted
__global__ void copy_cols_kernel(int64_t n, int wr, int64_t ws, int64_t c0,
                                 const uint64_t *__restrict__ X, uint64_t *__restrict__ Y) {
   const int64_t total = n * wr;
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
        t += (int64_t)gridDim.x * blockDim.x) {
      const int64_t i = t / wr;
      Y[t] = X[i * ws + c0 + (t - i * wr)];
   }
}

// Y (n x 2 wr words, row-major) = [A | B], A and B n x wr words row-major
__global__ void pack2_rows_kernel(int64_t n, int wr, const uint64_t *__restrict__ A,
                                  const uint64_t *__restrict__ B, uint64_t *__restrict__ Y) {
   const int64_t total = n * 2 * wr;
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
        t += (int64_t)gridDim.x * blockDim.x) {
      const int64_t i = t / (2 * wr);
      const int w = (int)(t - i * 2 * wr);
      Y[t] = (w < wr) ? A[i * wr + w] : B[i * wr + w - wr];
   }
}

// Y(i, j) *= s[j] for a row-major m x k block
__global__ void scale_cols_rm_kernel(int64_t m, int k, double *__restrict__ Y,
                                     const double *__restrict__ s) {
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < m * k) Y[t] *= s[t % k];
}

// device buffers owned by a dense operator for the length of one solve
struct DevBuf {
   void *p = nullptr;
   ~DevBuf() { if (p) (void)hipFree(p); }
   bool alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 16) == hipSuccess; }
};

static int64_t round8(int64_t x) { return (x + 7) / 8 * 8; }

// Dense symmetric A (FP64, column-major = row-major).  The growth products use
// a 16-bit copy of A on MFMA with FP32 accumulation (or A itself for an FP64
// basis); the residual uses A in FP64.
struct DenseSymOp final : FilterOp {
   int min_width() const override { return 16; }
   const double *A = nullptr;
   int64_t lda = 0;
   const void *Alo = nullptr;  // row-major n x n, row stride ldlo, type lo_dt (16-bit)
   int64_t ldlo = 0;
   int lo_dt = OFRR_F64;
   void *T = nullptr;          // n x k product of a growth step
   int upper(ofrr_ctx *ctx, double *ub) override { return dense_gershgorin(ctx, n, A, lda, ub); }
   int apply64(ofrr_ctx *ctx, int k, const double *X, double *W) override {
      return dense_gemm(ctx, n, n, k, A, OFRR_F64, lda, X, OFRR_F64, k, 1.0, 0.0, W, OFRR_F64, k);
   }
   int resid(ofrr_ctx *ctx, int k, const double *X, const double *th, double *W, void *Rq, int bdt,
             const double *inv_sR, float *colmax, double *sums) override {
      int st = apply64(ctx, k, X, W);
      if (st) return st;
      return resid_epilogue(ctx, n, k, X, th, W, Rq, bdt, inv_sR, colmax, sums);
   }
   int product(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt, int64_t zrs) {
      if (bdt == OFRR_F64)
         return dense_gemm(ctx, n, n, k, A, OFRR_F64, lda, D, OFRR_F64, k, 1.0, 0.0, Z, zdt, zrs);
      return dense_gemm(ctx, n, n, k, Alo, lo_dt, ldlo, D, bdt, k, 1.0, 0.0, Z, zdt, zrs);
   }
   int cheb(ofrr_ctx *ctx, int k, const void *Dc, const void *Dp, const void *R, int bdt,
            const ChebArgs &args, void *Dn) override {
      const int tdt = (bdt == OFRR_F64) ? OFRR_F64 : OFRR_F32;
      int st = product(ctx, k, Dc, bdt, T, tdt, k);
      if (st) return st;
      return cheb_combine(ctx, n, k, T, tdt, Dc, Dp, R, bdt, args, Dn);
   }
   // Z = A D for the projected matrix: the stored D (exact in its format) times
   // the FP64 A with FP64 accumulation, so V^T A V carries no growth-product
   // rounding (the error the truncated Rayleigh-Ritz would amplify along the
   // small-eigenvalue directions of the Gram matrix)
   double *D64 = nullptr;  // n x k
   int apply_lo(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt, int64_t zrs) override {
      if (bdt == OFRR_F64)
         return dense_gemm(ctx, n, n, k, A, OFRR_F64, lda, D, OFRR_F64, k, 1.0, 0.0, Z, zdt, zrs);
      int st = convert(ctx, n, k, D, bdt, k, 1, D64, OFRR_F64, k, 1);
      if (st) return st;
      return dense_gemm(ctx, n, n, k, A, OFRR_F64, lda, D64, OFRR_F64, k, 1.0, 0.0, Z, zdt, zrs);
   }
};

// A^T A for a dense M x N A (FP32, column-major): the operator of the singular
// value solver (filtered OFRR on the implicit A^T A).  A X and A^T Y of
// the growth steps are BF16 MFMA products with FP32 accumulation; the residual
// and the final vectors use the FP32 A with FP64 accumulation.
struct AtaOp final : FilterOp {
   int min_width() const override { return 16; }
   int64_t M = 0;
   const float *Acm = nullptr;  // column-major, lda: as P of A^T Y
   int64_t lda = 0;
   const float *Arm = nullptr;  // row-major copy, row stride N: as P of A X
   const void *Blo_rm = nullptr, *Blo_cm = nullptr;  // BF16 copies, row strides ldr / ldc
   int64_t ldr = 0, ldc = 0;
   double *Y64 = nullptr;       // M x k
   void *Ylo = nullptr;         // M x k BF16
   void *T = nullptr;           // N x k FP32
   int upper(ofrr_ctx *, double *ub) override {
      *ub = 0.0;  // no cheap bound of sigma_1^2: the window top follows the Ritz values
      return kSuccess;
   }
   int apply64(ofrr_ctx *ctx, int k, const double *X, double *W) override {
      int st = dense_gemm(ctx, M, n, k, Arm, OFRR_F32, n, X, OFRR_F64, k, 1.0, 0.0, Y64, OFRR_F64, k);
      if (st) return st;
      return dense_gemm(ctx, n, M, k, Acm, OFRR_F32, lda, Y64, OFRR_F64, k, 1.0, 0.0, W, OFRR_F64, k);
   }
   int resid(ofrr_ctx *ctx, int k, const double *X, const double *th, double *W, void *Rq, int bdt,
             const double *inv_sR, float *colmax, double *sums) override {
      int st = apply64(ctx, k, X, W);
      if (st) return st;
      return resid_epilogue(ctx, n, k, X, th, W, Rq, bdt, inv_sR, colmax, sums);
   }
   int product(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt, int64_t zrs) {
      int st;
      if (bdt == OFRR_F64) {
         if ((st = dense_gemm(ctx, M, n, k, Arm, OFRR_F32, n, D, OFRR_F64, k, 1.0, 0.0, Y64, OFRR_F64, k))) return st;
         return dense_gemm(ctx, n, M, k, Acm, OFRR_F32, lda, Y64, OFRR_F64, k, 1.0, 0.0, Z, zdt, zrs);
      }
      if ((st = dense_gemm(ctx, M, n, k, Blo_rm, OFRR_BF16, ldr, D, OFRR_BF16, k, 1.0, 0.0, Ylo, OFRR_BF16, k))) return st;
      return dense_gemm(ctx, n, M, k, Blo_cm, OFRR_BF16, ldc, Ylo, OFRR_BF16, k, 1.0, 0.0, Z, zdt, zrs);
   }
   int cheb(ofrr_ctx *ctx, int k, const void *Dc, const void *Dp, const void *R, int bdt,
            const ChebArgs &args, void *Dn) override {
      const int tdt = (bdt == OFRR_F64) ? OFRR_F64 : OFRR_F32;
      int st = product(ctx, k, Dc, bdt, T, tdt, k);
      if (st) return st;
      return cheb_combine(ctx, n, k, T, tdt, Dc, Dp, R, bdt, args, Dn);
   }
   // Z = A^T (A D) for the projected matrix with the FP32 A and FP64
   // accumulation (see DenseSymOp::apply_lo): BF16 products there pollute the
   // Ritz values when the Gram matrix is nearly singular
   double *D64 = nullptr;  // N x k
   int apply_lo(ofrr_ctx *ctx, int k, const void *D, int bdt, void *Z, int zdt, int64_t zrs) override {
      const void *Dq = D;
      if (bdt != OFRR_F64) {
         int st = convert(ctx, n, k, D, bdt, k, 1, D64, OFRR_F64, k, 1);
         if (st) return st;
         Dq = D64;
      }
      int st = dense_gemm(ctx, M, n, k, Arm, OFRR_F32, n, Dq, OFRR_F64, k, 1.0, 0.0, Y64, OFRR_F64, k);
      if (st) return st;
      return dense_gemm(ctx, n, M, k, Acm, OFRR_F32, lda, Y64, OFRR_F64, k, 1.0, 0.0, Z, zdt, zrs);
   }
};

// Storage width of the filter blocks: the smallest power of two >= k (the
// lane layout of the step kernels and of the fused projection products).  For
// a k that is not a power of two the blocks carry kp - k zero columns: X and W
// hold zeros there, so R, every filter block and Z do too; the Rayleigh-Ritz
// runs on the compacted Gram and projected matrices (the padded rows and
// columns are dropped before the eigensolves, where their zero diagonal would
// break the scaling) and its coefficients are scattered back with zeros for the
// padding, so the padding never mixes into the real columns.
static int filter_width(int k, int minw = 1) {
   int kp = 1;
   while (kp < k || kp < minw) kp *= 2;
   return kp;
}
// the narrowest storage the kernels take: the step kernels read 16 bytes per
// lane (8 FP16 / 16 FP8 columns), the residual kernel two FP64 columns, and
// the dense operators' MFMA products 16 columns (op_minw)
// the format of the wide sweeps (the growth format itself when they are off)
static int wide_format(const ofrr_filter_opts &o) {
   return o.wide_tol > 0.0 ? o.wide_dtype : o.basis_dtype;
}
static int filter_min_width(const ofrr_filter_opts &o, int op_minw) {
   const int rdt = o.refine ? o.refine_dtype : o.basis_dtype;
   const size_t es = std::min(std::min(dtype_size(o.basis_dtype), dtype_size(rdt)),
                              dtype_size(wide_format(o)));
   return std::max(std::max(2, (int)(16 / std::max<size_t>(1, es))), op_minw);
}

// The filtered solver's device state.  One layout serves the sizing
// (ofrr_filtered_eig_state_bytes and the serving footprint) and the solve, so
// the two cannot disagree; a take past the reservation returns null.
struct FiltBufs {
   double *X, *Xn, *W;        // n x kp FP64 row-major: Ritz block, next block, A X
   void *Rq, *Db[4], *Z;      // quantised residual and filter blocks (n x kp), Z (n x kp nd)
   double *G, *H, *C;         // projected matrices (mp x mp), coefficients (mp x kp)
   double *Gc, *Hc, *Cc;      // compacted m x m, m x k (padded widths only)
   int32_t *cmap, *cinv;      // compacted -> padded basis index (m), padded -> compacted (mp)
   double *th, *sums, *invsR, *sR, *scale[4], *sD1, *zall, *steps;
   float *cmax[3], *cmaxR;
   int32_t *rank;
};
struct Carve {
   char *base;
   size_t cap, off;
   void *take(size_t b) {
      const size_t a = al(off);
      off = a + b;
      return (base && off <= cap) ? base + a : nullptr;
   }
};
static void filtered_layout(Carve &cv, int64_t n, const ofrr_filter_opts &o, int op_minw, FiltBufs &b) {
   const int k = o.k, kp = filter_width(k, filter_min_width(o, op_minw)), nd = o.nd, d = o.degree;
   const int m = k * (1 + nd), mp = kp * (1 + nd);
   const int bdt = o.basis_dtype, rdt = o.refine ? o.refine_dtype : bdt, wdt = wide_format(o);
   // blocks sized for the largest of the growth, refinement and wide formats
   const size_t bsz = std::max(std::max(dtype_size(bdt), dtype_size(rdt)), dtype_size(wdt));
   // Z = A D in FP64 for the 4- and 8-byte formats (see set_format)
   auto zbytes = [](int dt) { return (dt == OFRR_F64 || dt == OFRR_F32) ? 8 : 4; };
   const size_t zsz = (size_t)std::max(std::max(zbytes(bdt), zbytes(rdt)), zbytes(wdt));
   const size_t nk = (size_t)n * kp, kq = (size_t)kp;
   b.X = (double *)cv.take(nk * 8);
   b.Xn = (double *)cv.take(nk * 8);
   b.W = (double *)cv.take(nk * 8);
   b.Rq = cv.take(nk * bsz);
   for (int i = 0; i < 4; ++i) b.Db[i] = cv.take(nk * bsz);
   b.Z = cv.take(nk * nd * zsz);
   b.G = (double *)cv.take((size_t)mp * mp * 8);
   b.H = (double *)cv.take((size_t)mp * mp * 8);
   b.C = (double *)cv.take((size_t)mp * kp * 8);
   b.Gc = b.Hc = b.Cc = nullptr;
   b.cmap = b.cinv = nullptr;
   if (kp != k || (nd == 2 && o.mid_cols > 0)) {
      b.Gc = (double *)cv.take((size_t)m * m * 8);
      b.Hc = (double *)cv.take((size_t)m * m * 8);
      b.Cc = (double *)cv.take((size_t)m * k * 8);
      b.cmap = (int32_t *)cv.take((size_t)m * 4);
      b.cinv = (int32_t *)cv.take((size_t)mp * 4);
   }
   b.th = (double *)cv.take(kq * 8);
   b.sums = (double *)cv.take(2 * kq * 8);
   b.invsR = (double *)cv.take(kq * 8);
   b.sR = (double *)cv.take(kq * 8);
   for (int i = 0; i < 4; ++i) b.scale[i] = (double *)cv.take(kq * 8);
   b.sD1 = (double *)cv.take(kq * 8);
   b.zall = (double *)cv.take((size_t)d * kq * 8);
   b.steps = (double *)cv.take((size_t)3 * d * 8);
   for (int i = 0; i < 3; ++i) b.cmax[i] = (float *)cv.take(kMaxSlots * kq * 4);  // partial maxima
   b.cmaxR = (float *)cv.take(kq * 4);
   b.rank = (int32_t *)cv.take(64);
}
static size_t filtered_state_bytes(int64_t n, const ofrr_filter_opts &o, int op_minw = 1) {
   Carve cv{nullptr, 0, 0};
   FiltBufs b;
   filtered_layout(cv, n, o, op_minw, b);
   return al(cv.off);
}

size_t filtered_workspace_bound(int num_cus) {
   // gram / gram_multi partials, the Rayleigh-Ritz at basis 384, the
   // back-transform's repacked coefficients (3 blocks of 128, width 128)
   const size_t basis = (size_t)(3 * 128 / 16) * 128 * 16 * sizeof(double);
   const size_t rr = std::max(rr_workspace_bytes(384), rr_split_workspace_bytes(1, 384, 128));
   return std::max(std::max(gram_workspace_bound(num_cus), rr), std::max(basis, (size_t)256));
}

// p[0 .. nwords) = 0 (4-byte words): a kernel rather than a memset, so the
// solver's clears are ordinary kernel nodes under stream capture
__global__ void zero_words_kernel(int64_t nwords, uint32_t *__restrict__ p) {
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nwords;
        t += (int64_t)gridDim.x * blockDim.x)
      p[t] = 0u;
}
static int zero_bytes(ofrr_ctx *ctx, void *p, size_t bytes) {
   const int64_t w = (int64_t)(bytes / 4);
   if (w <= 0) return kSuccess;
   const unsigned g = (unsigned)std::min<int64_t>((w + 255) / 256, 65536);
   zero_words_kernel<<<g, 256, 0, ctx->stream>>>(w, (uint32_t *)p);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// Gc (m x m) = rows and columns map[0 .. m) of G (column-major, ld ldg)
__global__ void sym_gather_kernel(int m, const int32_t *__restrict__ map, const double *__restrict__ G,
                                  int ldg, double *__restrict__ Gc) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= m * m) return;
   const int i = t % m, j = t / m;
   Gc[t] = G[map[i] + (int64_t)map[j] * ldg];
}
// C (mp x kp, ld mp) = the compacted coefficients Cc (m x k, ld m) at the
// padded rows inv^-1, zero in the padded rows and in the columns >= k
__global__ void coef_scatter_kernel(int mp, int kp, int k, int m, const int32_t *__restrict__ inv,
                                    const double *__restrict__ Cc, double *__restrict__ C) {
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= mp * kp) return;
   const int p = t % mp, j = t / mp;
   const int i = inv[p];
   C[t] = (j < k && i >= 0) ? Cc[i + (int64_t)j * m] : 0.0;
}

// 64-bit FNV-1a over a list of words: graph keys and argument signatures
static uint64_t hash_words(const std::vector<uint64_t> &v) {
   uint64_t h = 1469598103934665603ull;
   for (uint64_t w : v)
      for (int b = 0; b < 8; ++b) {
         h ^= (w >> (8 * b)) & 0xffu;
         h *= 1099511628211ull;
      }
   return h;
}

// Launch a filter sweep through the context's graph of its shape `key`:
// replayed as is when its kernel arguments (`sig`) repeat; otherwise the
// sweep is recorded again (enqueue() under stream capture) and either updates
// the executable graph of that shape in place or, for a new shape, is
// instantiated (the least recently used of kMaxSweepGraphs is dropped).
constexpr size_t kMaxSweepGraphs = 8;
static int sweep_graph_on(ofrr_ctx *ctx, uint64_t key, uint64_t sig, const std::function<int()> &enqueue,
                          int &captured) {
   captured = 0;
   ofrr_ctx::SweepGraph *hit = nullptr;
   for (auto &g : ctx->graphs)
      if (g.key == key && g.exec) { hit = &g; break; }
   if (hit && hit->sig == sig) {
      hit->used = ++ctx->graph_clock;
      OFRR_HIP(ctx, hipGraphLaunch(hit->exec, ctx->stream));
      return kSuccess;
   }
   hipGraph_t graph = nullptr;
   OFRR_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
   const int st = enqueue();
   const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
   if (st || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return st ? st : hip_status(ctx, ec, "filter sweep: stream capture");
   }
   captured = 1;
   bool updated = false;
   if (hit) {
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateError;
      if (hipGraphExecUpdate(hit->exec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
         updated = true;
      } else {
         (void)hipGetLastError();
         (void)hipGraphExecDestroy(hit->exec);
         hit->exec = nullptr;
      }
   }
   if (!updated) {
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
      if (ei != hipSuccess) {
         (void)hipGraphDestroy(graph);
         return hip_status(ctx, ei, "filter sweep: graph instantiation");
      }
      if (!hit) {
         if (ctx->graphs.size() >= kMaxSweepGraphs) {
            auto lru = ctx->graphs.begin();
            for (auto it = ctx->graphs.begin(); it != ctx->graphs.end(); ++it)
               if (it->used < lru->used) lru = it;
            if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
            ctx->graphs.erase(lru);
         }
         ctx->graphs.push_back(ofrr_ctx::SweepGraph{});
         hit = &ctx->graphs.back();
         hit->key = key;
      }
      hit->exec = ex;
   }
   (void)hipGraphDestroy(graph);
   hit->sig = sig;
   hit->used = ++ctx->graph_clock;
   OFRR_HIP(ctx, hipGraphLaunch(hit->exec, ctx->stream));
   return kSuccess;
}
// the same on the context's stream, or (null stream) on its own stream between
// a fork and a join event; enqueue() launches on ctx->stream, which points to
// the recording stream meanwhile
static int sweep_graph(ofrr_ctx *ctx, uint64_t key, uint64_t sig, const std::function<int()> &enqueue,
                       int &captured) {
   hipStream_t user = ctx->stream;
   if (user) return sweep_graph_on(ctx, key, sig, enqueue, captured);
   if (!ctx->own_stream) return set_error(ctx, kErrorInvalidOperation, "filter sweep graph: no stream");
   if (!ctx->ev_fork) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
   if (!ctx->ev_join) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
   OFRR_HIP(ctx, hipEventRecord(ctx->ev_fork, user));
   OFRR_HIP(ctx, hipStreamWaitEvent(ctx->own_stream, ctx->ev_fork, 0));
   ctx->stream = ctx->own_stream;
   const int st = sweep_graph_on(ctx, key, sig, enqueue, captured);
   ctx->stream = user;
   if (st) return st;
   OFRR_HIP(ctx, hipEventRecord(ctx->ev_join, ctx->own_stream));
   OFRR_HIP(ctx, hipStreamWaitEvent(user, ctx->ev_join, 0));
   return kSuccess;
}

static int filtered_core(ofrr_ctx *ctx, FilterOp &op, const ofrr_filter_opts *o, double *Xio,
                         double *theta_host, double *res_host, ofrr_filter_stats *stats) {
   const int64_t n = op.n;
   // kr: the Ritz block width asked for; k: its storage width (a power of two)
   const int kr = o->k, nd = o->nd, dmax = o->degree;
   int bdt = o->basis_dtype;  // the growth format of the current iteration
   const int d = dmax;        // buffer sizing: the largest degree of any sweep
   const int minw = filter_min_width(*o, op.min_width());
   if (kr < 2 || kr > 128 || nd < 1 || nd > 2 || d < 2 || o->nwant < 1 || o->nwant > kr ||
       !dtype_valid(bdt) || (int64_t)filter_width(kr, minw) > n || o->mid_cols < 0)
      return set_error(ctx, kErrorInvalidArgument,
                       "filtered_eig: need k in [2,128], nd in {1,2}, degree >= 2, 1 <= nwant <= k");
   if (o->refine && (!dtype_valid(o->refine_dtype) || !op.can_refine()))
      return set_error(ctx, kErrorInvalidArgument,
                       "filtered_eig: the refinement format is valid for a CSR operator only");
   const bool wide_on = o->wide_tol > 0.0;
   if (wide_on && (!dtype_valid(o->wide_dtype) || !op.can_refine() ||
                   dtype_size(o->wide_dtype) <= dtype_size(o->basis_dtype)))
      return set_error(ctx, kErrorInvalidArgument,
                       "filtered_eig: the wide format must be wider than the growth format (CSR operator only)");
   const int k = filter_width(kr, minw);
   // mid: the trailing columns of the mid-filter block kept in the basis (0: all)
   const int mid_keep = (nd == 2 && o->mid_cols > 0 && o->mid_cols < kr) ? o->mid_cols : 0;
   const bool padded = k != kr || mid_keep > 0;  // columns left out: compacted Rayleigh-Ritz
   const int m = k * (1 + nd);
   if (kr * nd + (mid_keep ? mid_keep : (nd == 2 ? kr : 0)) > 384)
      return set_error(ctx, kErrorInvalidArgument, "filtered_eig: basis dimension > 384");
   int zdt = (bdt == OFRR_F64 || bdt == OFRR_F32) ? OFRR_F64 : OFRR_F32;  // as set_format
   size_t zsz = dtype_size(zdt);
   // ---- buffers
   const size_t state_bytes = filtered_state_bytes(n, *o, op.min_width());
   int st = arena_reserve(ctx, state_bytes);
   if (st) return st;
   FiltBufs fb;
   {
      Carve cv{(char *)arena_take(state_bytes), state_bytes, 0};
      if (!cv.base) return set_error(ctx, kErrorInternalError, "filtered_eig: state reservation");
      filtered_layout(cv, n, *o, op.min_width(), fb);
      if (!fb.rank) return set_error(ctx, kErrorInternalError, "filtered_eig: state layout exceeds its reservation");
   }
   double *X = fb.X, *Xn = fb.Xn, *W = fb.W;
   void *Rq = fb.Rq, *Z = fb.Z;
   void *const *Db = fb.Db;
   double *G = fb.G, *H = fb.H, *C = fb.C;
   double *th = fb.th, *sums = fb.sums, *invsR = fb.invsR, *sR = fb.sR, *sD1 = fb.sD1, *zall = fb.zall;
   double *const *scale = fb.scale;
   float *const *cmax = fb.cmax;
   float *cmaxR = fb.cmaxR;
   int32_t *rank = fb.rank;
   double *sRd = sR;
   // pinned host mirrors (the compaction maps after the rank)
   const size_t hbytes = (size_t)(2 * k + k + k + d * k + 4 * k + 16) * 8 + (size_t)2 * m * 4 +
                         (size_t)3 * d * 8;
   double *hp = (double *)pinned(ctx, hbytes);
   if (!hp) return set_error(ctx, kErrorCUDAMemory, "filtered_eig: pinned");
   double *h_sums = hp, *h_th = hp + 2 * k, *h_sR = h_th + k, *h_z = h_sR + k;
   float *h_cmaxR = (float *)(h_z + (size_t)d * k);
   int32_t *h_rank = (int32_t *)(h_cmaxR + 2 * k);
   int32_t *h_cmap = (int32_t *)(hp + (size_t)(2 * k + k + k + d * k + 4 * k + 16));
   int32_t *h_cinv = h_cmap + m;
   double *h_step = (double *)(h_cinv + m);  // 3 scalars per filter step
   // Rayleigh-Ritz of the basis [X | D_d (| D_mid)] with me = k + nd ka padded
   // columns, of which the correction blocks hold kar = ka - (k - kr) real
   // ones: eigensolves on the compacted matrices when the blocks are padded
   // the truncated Rayleigh-Ritz itself, or (a serving slot) its job handed
   // to the server's grouped Rayleigh-Ritz
   // kx > 0: the split form with the first kx columns the previous Ritz block
   // h_ready: set while the projected matrix H is still being formed on the
   // second stream (the Rayleigh-Ritz waits for it before its first read of H)
   hipEvent_t h_ready = nullptr;
   auto rr_run = [&](int mm, const double *Gm, const double *Hm, double *Y, int kx) -> int {
      const RRJob job{Gm, Hm, mm, mm, mm, kr, o->tau, th, Y, mm, rank, kx, h_ready};
      if (!ctx->rr_group) return rr_group_run(ctx, 1, &job);
      return ctx->rr_group(ctx->rr_group_arg, ctx, &job);
   };
   auto rayleigh_ritz = [&](int me, int ka, int kx) -> int {
      if (!padded) return rr_run(me, G, H, C, kx);
      const int kar = ka - (k - kr), nb = (me - k) / std::max(1, ka);
      int mr = 0;
      for (int p = 0; p < me; ++p) h_cinv[p] = -1;
      for (int j = 0; j < kr; ++j) { h_cinv[j] = mr; h_cmap[mr++] = j; }
      for (int bb = 0; bb < nb; ++bb) {
         // the mid-filter block (bb = 1) keeps its trailing mid_keep real columns
         const int j0 = (bb == 1 && mid_keep) ? std::max(0, kar - mid_keep) : 0;
         for (int j = j0; j < kar; ++j) {
            const int p = k + bb * ka + j;
            h_cinv[p] = mr;
            h_cmap[mr++] = p;
         }
      }
      OFRR_HIP(ctx, hipMemcpyAsync(fb.cmap, h_cmap, (size_t)mr * 4, hipMemcpyHostToDevice, ctx->stream));
      OFRR_HIP(ctx, hipMemcpyAsync(fb.cinv, h_cinv, (size_t)me * 4, hipMemcpyHostToDevice, ctx->stream));
      if (h_ready) OFRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, h_ready, 0));
      sym_gather_kernel<<<(mr * mr + 255) / 256, 256, 0, ctx->stream>>>(mr, fb.cmap, G, me, fb.Gc);
      sym_gather_kernel<<<(mr * mr + 255) / 256, 256, 0, ctx->stream>>>(mr, fb.cmap, H, me, fb.Hc);
      OFRR_LAUNCH(ctx);
      int s2 = rr_run(mr, fb.Gc, fb.Hc, fb.Cc, kx);
      if (s2) return s2;
      coef_scatter_kernel<<<(me * k + 255) / 256, 256, 0, ctx->stream>>>(me, k, kr, mr, fb.cinv, fb.Cc, C);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   };

   hipEvent_t ev[11];
   for (int i = 0; i < 11; ++i) OFRR_HIP(ctx, hipEventCreate(&ev[i]));
   float t_acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
   int filter_steps = 0, resid_passes = 0, filter_col_steps = 0, filter_first_cols = 0;
   int graph_replays = 0, graph_captures = 0, rr_fallbacks = 0, wide_sweeps = 0;
   auto rec = [&](int i) { return hipEventRecord(ev[i], ctx->stream); };
   OFRR_HIP(ctx, rec(0));

   int spmm_calls = 0, rank_min = m;
   // ---- top of the filter window: an upper bound of the spectrum (the
   // Gershgorin bound for a CSR operator); without one, the top Ritz value
   double upper = 0.0;
   if ((st = op.upper(ctx, &upper))) return st;
   // ---- initial Rayleigh-Ritz on the start block (zero padding columns)
   if (padded) {
      if ((st = zero_bytes(ctx, X, (size_t)n * k * 8))) return st;
      if ((st = convert(ctx, n, kr, Xio, OFRR_F64, kr, 1, X, OFRR_F64, k, 1))) return st;
   } else {
      OFRR_HIP(ctx, hipMemcpyAsync(X, Xio, (size_t)n * k * 8, hipMemcpyDeviceToDevice, ctx->stream));
   }
   if ((st = zero_bytes(ctx, th, (size_t)k * 8))) return st;
   if ((st = op.apply64(ctx, k, X, W))) return st;
   ++spmm_calls;
   if ((st = gram(ctx, n, k, k, X, OFRR_F64, k, 1, X, OFRR_F64, k, 1, 1.0, 0.0, G, k))) return st;
   if ((st = gram(ctx, n, k, k, X, OFRR_F64, k, 1, W, OFRR_F64, k, 1, 1.0, 0.0, H, k))) return st;
   if ((st = rayleigh_ritz(k, k, 0))) return st;
   {
      const void *vb[1] = {X};
      int dt[1] = {OFRR_F64}, kb[1] = {k}, ldc[1] = {k};
      int64_t rs[1] = {k}, cs[1] = {1};
      const double *cb[1] = {C};
      if ((st = basis_update(ctx, n, k, 1, vb, dt, kb, rs, cs, cb, ldc, Xn, OFRR_F64, k, 1))) return st;
      std::swap(X, Xn);
   }
   OFRR_HIP(ctx, hipMemcpyAsync(h_th, th, (size_t)k * 8, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));

   int it = 0, converged = 0, nconv = 0;
   std::vector<double> Lth(k), res(k, NAN), theta(k);
   // host staging kept alive across the asynchronous uploads of an iteration
   std::vector<double> inv(k), s1(k), pred(k, 1.0);
   std::vector<float> m1((size_t)kMaxSlots * k, 0.f);  // slot 0: the maxima, the others 0
   // quantisation targets: the stored maximum of R sits near 2^tgt_resid and
   // the filter blocks near 2^tgt_step (FP8 E4M3 uses the top of its range)
   bool scaled = false;
   int tgt_resid = 0, tgt_step = 0;
   double lo_ratio = 1.0;
   auto set_format = [&](int dt) {
      bdt = dt;
      // Z = A D accumulated and kept in FP64 for FP32 blocks too: their Gram
      // matrix resolves directions down to about 1e-15 of its largest
      // eigenvalue, above the truncation, and an FP32 Z (relative error ~6e-8)
      // would put errors of order 1e-10 / lambda into those directions of H
      // (spurious Ritz values above the wanted ones); the FP16 blocks' own
      // rounding keeps their Gram spectrum above ~1e-7, where FP32 Z suffices
      zdt = (bdt == OFRR_F64 || bdt == OFRR_F32) ? OFRR_F64 : OFRR_F32;
      zsz = dtype_size(zdt);
      scaled = (bdt == OFRR_F16 || bdt == OFRR_FP8);
      tgt_resid = (bdt == OFRR_FP8) ? 8 : 0;
      tgt_step = (bdt == OFRR_FP8) ? 6 : 0;
      lo_ratio = (bdt == OFRR_FP8) ? 16.0 : 1.0 / 256.0;
   };
   set_format(bdt);
   // precision ladder: the largest wanted residual of each iteration
   std::vector<double> worst_hist;
   for (int j = 0; j < k; ++j) h_sR[j] = 1.0;
   for (int j = 0; j < k; ++j) theta[j] = h_th[j];
   // column maxima |R| of the last two iterations' final residual passes (FP16)
   std::vector<double> cm_last(k, 0.0), cm_prev(k, 0.0);
   for (it = 0;; ++it) {
      // ---- FP64 operator application + residual, quantised R
      OFRR_HIP(ctx, rec(1));
      if (bdt == OFRR_F16) {
         // predicted scale of R: the last maximum times the column's last
         // reduction factor (within [2^-6, 1]), so that the stored range stays
         // near the top of the format and R is seldom formed twice
         for (int j = 0; j < k; ++j) {
            if (!(cm_last[j] > 0.0)) continue;
            double f = cm_prev[j] > 0.0 ? cm_last[j] / cm_prev[j] : 1.0;
            f = std::min(1.0, std::max(1.0 / 64.0, f));
            h_sR[j] = std::exp2(std::ceil(std::log2(cm_last[j] * f)) - tgt_resid);
         }
      }
   resid_pass:
      for (int pass = 0; pass < 2; ++pass) {
         for (int j = 0; j < k; ++j) inv[j] = 1.0 / h_sR[j];
         OFRR_HIP(ctx, hipMemcpyAsync(invsR, inv.data(), (size_t)k * 8, hipMemcpyHostToDevice, ctx->stream));
         OFRR_HIP(ctx, hipMemcpyAsync(sRd, h_sR, (size_t)k * 8, hipMemcpyHostToDevice, ctx->stream));
         if ((st = zero_bytes(ctx, cmaxR, (size_t)k * 4))) return st;
         if ((st = op.resid(ctx, k, X, th, W, Rq, bdt, invsR, cmaxR, sums))) return st;
         ++spmm_calls;
         ++resid_passes;
         OFRR_HIP(ctx, hipMemcpyAsync(h_sums, sums, (size_t)2 * k * 8, hipMemcpyDeviceToHost, ctx->stream));
         OFRR_HIP(ctx, hipMemcpyAsync(h_cmaxR, cmaxR, (size_t)k * 4, hipMemcpyDeviceToHost, ctx->stream));
         OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
         if (!scaled) break;  // FP64 / FP32 / BF16 hold R unscaled
         // R was quantised with the predicted scale h_sR; redo once when the
         // stored range left the window of the format
         bool redo = false;
         for (int j = 0; j < k; ++j) {
            const double cmj = (double)h_cmaxR[j];
            const double ratio = cmj / h_sR[j];
            const double want = cmj > 0 ? std::exp2(std::ceil(std::log2(cmj)) - tgt_resid) : 1.0;
            if (j < kr && (ratio > 0.5 * dtype_max(bdt) || ratio < lo_ratio)) redo = true;
            pred[j] = want;
         }
         if (!redo || pass == 1) break;
         for (int j = 0; j < k; ++j) h_sR[j] = pred[j];
      }
      if (bdt == OFRR_F16)
         for (int j = 0; j < k; ++j) {
            cm_prev[j] = cm_last[j];
            cm_last[j] = (double)h_cmaxR[j];
         }
      OFRR_HIP(ctx, rec(2));
      OFRR_HIP(ctx, hipEventSynchronize(ev[2]));
      const double normA = o->normA > 0 ? o->normA : theta[0];
      nconv = 0;
      double worst = 0.0;
      for (int j = 0; j < k; ++j) {
         res[j] = std::sqrt(h_sums[2 * j]) / (normA * std::sqrt(h_sums[2 * j + 1]));
         if (j < o->nwant && res[j] <= o->tol) ++nconv;
         if (j < o->nwant) worst = std::max(worst, res[j]);
      }
      // wide sweeps: after the first, a sweep that starts above wide_tol builds
      // its blocks in the wide format, the others in the growth format (until
      // the precision ladder, if on, has switched); R is then formed again in
      // the sweep's format, with the scale measured in this pass
      if (wide_on && nconv < o->nwant && !(o->refine && bdt == o->refine_dtype)) {
         const int tf = (it >= 1 && worst > o->wide_tol) ? o->wide_dtype : o->basis_dtype;
         if (tf != bdt) {
            set_format(tf);
            op.set_precise(tf == OFRR_F64);
            for (int j = 0; j < k; ++j)
               h_sR[j] = (scaled && h_cmaxR[j] > 0.f)
                             ? std::exp2(std::ceil(std::log2((double)h_cmaxR[j])) - tgt_resid)
                             : 1.0;
            goto resid_pass;
         }
      }
      // precision ladder: continue in the refinement format once the wanted
      // residuals are below refine_tol or have not halved in three iterations;
      // R is then formed again in the new format
      if (o->refine && bdt != o->refine_dtype && nconv < o->nwant) {
         worst_hist.push_back(worst);
         const size_t h = worst_hist.size();
         const bool stalled = h >= 4 && worst > 0.5 * worst_hist[h - 4];
         if (worst <= o->refine_tol || stalled) {
            set_format(o->refine_dtype);
            op.set_precise(bdt == OFRR_F64);
            for (int j = 0; j < k; ++j) h_sR[j] = 1.0;
            goto resid_pass;
         }
      }
      {
         float ms_;
         OFRR_HIP(ctx, hipEventElapsedTime(&ms_, ev[1], ev[2]));
         t_acc[5] += ms_;
      }
      if (nconv == o->nwant) { converged = 1; break; }
      if (it >= o->maxiter) break;

      // ---- Chebyshev filter on the correction, damping [lower, theta_k].
      // D_{j+1} = (2/e)(A - c)D_j - D_{j-1} + (2/e) z_j R with z_j = T_j(L(theta)),
      // L(t) = (t - c)/e, run on Dh_j = D_j / mu_j, mu_j = T_j(L(b)), b = the
      // upper Gershgorin bound (no eigenvalue lies above it, so |Dh_j| stays
      // within a small multiple of |R| / e whatever the degree):
      // Dh_{j+1} = (2/e) q_j (A - c) Dh_j - q_{j-1} q_j Dh_{j-1} + (2/e) zeta_j q_j R
      // with q_j = mu_j / mu_{j+1} and zeta_j = z_j / mu_j (all O(1): no overflow
      // at any degree).
      const double cut = theta[kr - 1], lo = o->lower;
      const double e = 0.5 * (cut - lo), c0 = 0.5 * (cut + lo);
      for (int j = 0; j < k; ++j) Lth[j] = (theta[j] - c0) / e;
      const double top = (upper > 0.0) ? std::max(theta[0], upper) : theta[0] * (1.0 + 1e-3);
      const double Lb = std::max(1.0, (top - c0) / e);
      // Degree of this sweep.  The filter raises the top of the spectrum over
      // the wanted edge by (rho(L(theta_1)) / rho(L(theta_nwant)))^degree,
      // rho(x) = x + sqrt(x^2 - 1).  When that exceeds kFilterGain the new
      // directions drop below the rounding of the converged ones and the
      // residual stalls (a steeply decaying spectrum: A^T A of the SVD); the
      // degree is capped there.  A flat top (the Laplacian) is not affected.
      int d = dmax;
      {
         auto rho = [](double x) { return x > 1.0 ? x + std::sqrt(x * x - 1.0) : 1.0; };
         const double ratio = rho(Lth[0]) / rho(Lth[o->nwant - 1]);
         if (ratio > 1.0) {
            const int cap = (int)std::floor(std::log(kFilterGain) / std::log(ratio));
            d = std::max(2, std::min(dmax, cap));
         }
      }
      const int mid = d / 2;
      // Soft locking: the leading wanted pairs that meet tol keep their Ritz
      // vectors in X (and in every Rayleigh-Ritz) but get no correction; the
      // sweep filters columns [c, k), ka = k - c a power of two (the step
      // kernels' lane layout) and >= 16 (the back-transform's column groups)
      int ka = k;
      if (o->lock && op.can_lock()) {
         int pre = 0;
         while (pre < o->nwant && res[pre] <= o->tol) ++pre;
         while (ka / 2 >= 16 && ka / 2 >= k - pre) ka /= 2;
      }
      const int c = k - ka, me = k + nd * ka;
      std::vector<double> q(d + 1);
      q[0] = 1.0 / Lb;
      for (int s = 1; s <= d; ++s) q[s] = 1.0 / (2.0 * Lb - q[s - 1]);
      for (int jj = 0; jj < ka; ++jj) {
         const int j = c + jj;
         double zp = 1.0, zc = Lth[j] / Lb;  // zeta_0, zeta_1
         for (int s = 1; s < d; ++s) {
            h_z[(size_t)(s - 1) * ka + jj] = (2.0 / e) * zc * q[s];  // weight of R in step s
            const double zn = 2.0 * Lth[j] * zc * q[s] - zp * q[s - 1] * q[s];
            zp = zc;
            zc = zn;
         }
      }
      OFRR_HIP(ctx, hipMemcpyAsync(zall, h_z, (size_t)(d - 1) * ka * 8, hipMemcpyHostToDevice, ctx->stream));
      // Dh_1 = R / (e L(b)), stored in Rq: scale sR / (e Lb), range max|R| / (e Lb)
      for (int jj = 0; jj < ka; ++jj) s1[jj] = h_sR[c + jj] / (e * Lb);
      OFRR_HIP(ctx, hipMemcpyAsync(sD1, s1.data(), (size_t)ka * 8, hipMemcpyHostToDevice, ctx->stream));
      for (int jj = 0; jj < ka; ++jj) m1[jj] = (float)((double)h_cmaxR[c + jj] / (e * Lb));
      OFRR_HIP(ctx, hipMemcpyAsync(cmax[0], m1.data(), (size_t)kMaxSlots * k * 4, hipMemcpyHostToDevice,
                                   ctx->stream));
      OFRR_HIP(ctx, hipEventRecord(ev[3], ctx->stream));
      // the filtered columns of the quantised residual, n x ka row-major: Rq
      // itself, or (locked columns present) a compact copy in the second half
      // of Db[0], whose first half is all a step writes (n ka <= n k / 2)
      const void *Ra = Rq;
      const int es = (int)dtype_size(bdt);
      if (c > 0) Ra = (char *)Db[0] + (size_t)n * ka * es;
      // The buffers of every step (host bookkeeping): the recurrence keeps the
      // current and previous stored blocks and their scales; Db[b] has scale
      // scale[b]; D_1 lives in Ra with scale sD1.  The column maxima rotate
      // over three buffers: step s reads cmax[(s-1)%3], accumulates into
      // cmax[s%3] (cleared by step s-1, or before the sweep for s = 1) and
      // clears cmax[(s+1)%3] for step s+1.
      struct StepBufs { int out; const void *Dc, *Dp; double *sc, *sp; };
      std::vector<StepBufs> sb(d);
      const void *Dc = Ra;
      const void *Dmid = (nd == 2 && mid == 1) ? Ra : nullptr;
      int curIdx = -1, snap = -1;
      {
         const void *Dp = nullptr;
         double *sc = sD1, *sp = nullptr;
         int prevIdx = -1;
         for (int s = 1; s < d; ++s) {
            int out = -1;
            for (int b = 0; b < 4; ++b)
               if (b != curIdx && b != prevIdx && b != snap) { out = b; break; }
            sb[s] = StepBufs{out, Dc, Dp, sc, sp};
            prevIdx = curIdx;
            Dp = Dc;
            sp = sc;
            curIdx = out;
            Dc = Db[out];
            sc = scale[out];
            if (nd == 2 && s + 1 == mid) { snap = out; Dmid = Db[out]; }
         }
      }
      // step scalars in device memory (a replayed sweep reads this iteration's)
      for (int s = 1; s < d; ++s) {
         h_step[3 * (s - 1)] = (2.0 / e) * q[s];
         h_step[3 * (s - 1) + 1] = q[s - 1] * q[s];
         h_step[3 * (s - 1) + 2] = c0;
      }
      OFRR_HIP(ctx, hipMemcpyAsync(fb.steps, h_step, (size_t)(d - 1) * 3 * 8, hipMemcpyHostToDevice,
                                   ctx->stream));
      auto enqueue_sweep = [&]() -> int {
         if (c > 0) {
            const int wr = ka * es / 8;
            const unsigned g = (unsigned)std::min<int64_t>((n * wr + 255) / 256, 65536);
            copy_cols_kernel<<<g, 256, 0, ctx->stream>>>(n, wr, (int64_t)k * es / 8, (int64_t)c * es / 8,
                                                         (const uint64_t *)Rq, (uint64_t *)Ra);
            OFRR_LAUNCH(ctx);
         }
         int st2 = zero_bytes(ctx, cmax[1], (size_t)kMaxSlots * k * 4);
         if (st2) return st2;
         for (int s = 1; s < d; ++s) {
            ChebArgs ca;
            ca.alpha = h_step[3 * (s - 1)];
            ca.beta = h_step[3 * (s - 1) + 1];
            ca.c = c0;
            ca.step = fb.steps + 3 * (s - 1);
            ca.w = zall + (size_t)(s - 1) * ka;
            ca.sR = sRd + c;
            ca.sc = sb[s].sc;
            ca.sp = sb[s].sp;
            ca.maxc = cmax[(s - 1) % 3];
            ca.maxn = cmax[s % 3];
            ca.zero_next = (s + 1 < d) ? cmax[(s + 1) % 3] : nullptr;
            ca.slot_ld = k;
            ca.sn_out = scale[sb[s].out];
            ca.target_exp = tgt_step;
            // alternate steps walk the row tiles backwards, so a step starts on
            // the rows whose blocks the previous step left in the memory-side cache
            ca.reverse = s & 1;
            if ((st2 = op.cheb(ctx, ka, sb[s].Dc, sb[s].Dp, Ra, bdt, ca, Db[sb[s].out]))) return st2;
         }
         return kSuccess;
      };
      if (o->graph > 0 && op.can_graph()) {
         // the shape of the sweep (its kernels, grids and scalar arguments) and
         // every buffer it touches
         std::vector<uint64_t> kv = {(uint64_t)n, (uint64_t)k, (uint64_t)ka, (uint64_t)c, (uint64_t)d,
                                     (uint64_t)bdt, (uint64_t)tgt_step, (uint64_t)nd, (uint64_t)mid};
         std::vector<uint64_t> sv;
         op.sweep_sig(sv);
         kv.push_back(sv.size() > 3 ? sv[3] : 0);  // the operator's value format
         const void *bufs[] = {Rq, Ra, Db[0], Db[1], Db[2], Db[3], zall, sRd, sD1, fb.steps};
         for (const void *ptr : bufs) sv.push_back((uint64_t)(uintptr_t)ptr);
         for (int i = 0; i < 4; ++i) sv.push_back((uint64_t)(uintptr_t)scale[i]);
         for (int i = 0; i < 3; ++i) sv.push_back((uint64_t)(uintptr_t)cmax[i]);
         int captured = 0;
         if ((st = sweep_graph(ctx, hash_words(kv), hash_words(sv), enqueue_sweep, captured))) return st;
         ++graph_replays;
         graph_captures += captured;
      } else if ((st = enqueue_sweep())) {
         return st;
      }
      spmm_calls += d - 1;
      filter_steps += d - 1;
      if (wide_on && bdt == o->wide_dtype) ++wide_sweeps;
      filter_col_steps += ka * (d - 1);
      filter_first_cols += ka;
      OFRR_HIP(ctx, hipEventRecord(ev[4], ctx->stream));
      // ---- projection: V = [X | Dd (| Dmid)], AV = [W | Z]
      const void *Dd = Dc;
      const int64_t zs = (int64_t)k * nd;
      // Outside a serving slot (whose Rayleigh-Ritz is grouped) the products
      // split over two streams: G = V^T V first on the solver's stream, where
      // the Rayleigh-Ritz then starts (its Gram-only stages, one CU), while the
      // second stream forms Z = A D and H = V^T A V on the other CUs; the
      // Rayleigh-Ritz waits for them before it first reads H.
      const bool overlap = !ctx->rr_group && op.can_overlap();
      if (overlap) {
         if (!ctx->aux_stream) OFRR_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
         if (!ctx->ev_aux_fork) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_aux_fork, hipEventDisableTiming));
         if (!ctx->ev_aux_join) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_aux_join, hipEventDisableTiming));
      }
      auto z_products = [&]() -> int {
         int s2 = op.apply_lo(ctx, ka, Dd, bdt, Z, zdt, (int64_t)ka * nd);
         if (s2) return s2;
         ++spmm_calls;
         if (nd == 2) {
            if ((s2 = op.apply_lo(ctx, ka, Dmid, bdt, (char *)Z + (size_t)ka * zsz, zdt, (int64_t)ka * nd))) return s2;
            ++spmm_calls;
         }
         return kSuccess;
      };
      // locked columns present with two correction blocks: [D_d | D_mid] packed
      // row by row next to the Z layout [Z_d | Z_mid] (a free filter buffer:
      // neither block, nor Db[0] whose second half may hold D_mid = D_1)
      const int kd = nd * ka;
      const void *Dp2 = Dd;
      if (c > 0 && nd == 2) {
         int b = -1;
         for (int i = 1; i < 4 && b < 0; ++i)
            if (i != curIdx && i != snap) b = i;
         const int es = (int)dtype_size(bdt);
         const int wr = ka * es / 8;
         const unsigned g = (unsigned)std::min<int64_t>((n * 2 * wr + 255) / 256, 65536);
         pack2_rows_kernel<<<g, 256, 0, ctx->stream>>>(n, wr, (const uint64_t *)Dd,
                                                       (const uint64_t *)Dmid, (uint64_t *)Db[b]);
         OFRR_LAUNCH(ctx);
         Dp2 = Db[b];
      }
      // the G blocks (want & 1) and the H blocks (want & 2) of the upper block
      // triangles (column-major, ld me), then their mirror images
      auto products = [&](int want) -> int {
         // (rows operand, columns operand, destination, block row, block column)
         struct Blk { int a, b; double *dst; int br, bc; };
         int s2 = kSuccess;
         if (c == 0 || kd == k) {
            // fused launches: G = [X Dd Dm]^T [X Dd Dm]; H = [X Dd Dm]^T [W Zd Zm]
            // with the X-D blocks of H formed as W^T D (W = A X in FP64, D exact in
            // its format); with locked columns the packed correction block is as
            // wide as X and takes the place of one full-width block
            GramMultiArgs ga{};
            ga.op[0] = {X, (int64_t)k, 0};
            ga.op[1] = {W, (int64_t)k, 0};
            int nb;
            const Blk *blks;
            static const int kG = 0, kH = 1;
            Blk b12[12], b6[6];
            double *dst[2] = {G, H};
            if (c == 0) {
               ga.op[2] = {Dd, (int64_t)k, 1};
               ga.op[3] = {Dmid, (int64_t)k, 1};
               ga.op[4] = {Z, zs, 2};
               ga.op[5] = {(const char *)Z + (size_t)k * zsz, zs, 2};
               const int t12[12][5] = {{0, 0, kG, 0, 0}, {0, 2, kG, 0, 1}, {2, 2, kG, 1, 1},
                                       {0, 1, kH, 0, 0}, {1, 2, kH, 0, 1}, {2, 4, kH, 1, 1},
                                       {0, 3, kG, 0, 2}, {2, 3, kG, 1, 2}, {3, 3, kG, 2, 2},
                                       {1, 3, kH, 0, 2}, {2, 5, kH, 1, 2}, {3, 5, kH, 2, 2}};
               for (int q = 0; q < 12; ++q)
                  b12[q] = Blk{t12[q][0], t12[q][1], dst[t12[q][2]], t12[q][3], t12[q][4]};
               nb = (nd == 2) ? 12 : 6;
               blks = b12;
            } else {
               ga.op[2] = {Dp2, (int64_t)kd, 1};
               ga.op[3] = {Z, (int64_t)kd, 2};
               const int t6[6][5] = {{0, 0, kG, 0, 0}, {0, 2, kG, 0, 1}, {2, 2, kG, 1, 1},
                                     {0, 1, kH, 0, 0}, {1, 2, kH, 0, 1}, {2, 3, kH, 1, 1}};
               for (int q = 0; q < 6; ++q)
                  b6[q] = Blk{t6[q][0], t6[q][1], dst[t6[q][2]], t6[q][3], t6[q][4]};
               nb = 6;
               blks = b6;
            }
            const int nt = (k + 63) / 64;
            int t = 0;
            for (int q = 0; q < nb; ++q) {
               if (!(want & (blks[q].dst == G ? 1 : 2))) continue;
               for (int ti = 0; ti < nt; ++ti)
                  for (int tj = 0; tj < nt; ++tj, ++t) {
                     ga.ia[t] = blks[q].a;
                     ga.ib[t] = blks[q].b;
                     ga.ti[t] = ti;
                     ga.tj[t] = tj;
                     ga.dst[t] = blks[q].dst;
                     ga.ld[t] = me;
                     ga.ro[t] = blks[q].br * k;
                     ga.co[t] = blks[q].bc * k;
                  }
            }
            ga.nitems = t;
            // chunks sized for all nb blocks: G and H launched apart sum each
            // entry as their fused launch does (64 x 32 halves: ceil(k / 32) per tile row)
            ga.size_halves = nb * nt * ((k + 31) / 32);
            if ((s2 = gram_multi(ctx, n, k, bdt, zdt, ga))) return s2;
         } else {
            // the six products X^T X, X^T D, D^T D, X^T W, W^T D, D^T Z, one
            // launch each
            struct Prod { const void *a; int adt, wa; const void *b; int bdt_, wb; double *dst; int ro, co; };
            const Prod pr[6] = {{X, OFRR_F64, k, X, OFRR_F64, k, G, 0, 0},
                                {X, OFRR_F64, k, Dp2, bdt, kd, G, 0, k},
                                {Dp2, bdt, kd, Dp2, bdt, kd, G, k, k},
                                {X, OFRR_F64, k, W, OFRR_F64, k, H, 0, 0},
                                {W, OFRR_F64, k, Dp2, bdt, kd, H, 0, k},
                                {Dp2, bdt, kd, Z, zdt, kd, H, k, k}};
            for (int q = 0; q < 6; ++q) {
               const Prod &P = pr[q];
               if (!(want & (P.dst == G ? 1 : 2))) continue;
               if ((s2 = gram(ctx, n, P.wa, P.wb, P.a, P.adt, P.wa, 1, P.b, P.bdt_, P.wb, 1, 1.0, 0.0,
                              P.dst + P.ro + (size_t)P.co * me, me)))
                  return s2;
            }
         }
         if (want & 1) mirror_upper_kernel<<<(me * me + 255) / 256, 256, 0, ctx->stream>>>(me, G, me);
         if (want & 2) mirror_upper_kernel<<<(me * me + 255) / 256, 256, 0, ctx->stream>>>(me, H, me);
         OFRR_LAUNCH(ctx);
         return kSuccess;
      };
      if (!overlap) {
         if ((st = z_products())) return st;
         OFRR_HIP(ctx, hipEventRecord(ev[5], ctx->stream));
         if ((st = products(3))) return st;
      } else {
         if ((st = products(1))) return st;
         OFRR_HIP(ctx, hipEventRecord(ctx->ev_aux_fork, ctx->stream));
         OFRR_HIP(ctx, hipStreamWaitEvent(ctx->aux_stream, ctx->ev_aux_fork, 0));
         {
            OnAuxStream aux(ctx);
            OFRR_HIP(ctx, hipEventRecord(ev[8], ctx->stream));
            if ((st = z_products())) return st;
            OFRR_HIP(ctx, hipEventRecord(ev[9], ctx->stream));
            if ((st = products(2))) return st;
            OFRR_HIP(ctx, hipEventRecord(ev[10], ctx->stream));
            OFRR_HIP(ctx, hipEventRecord(ctx->ev_aux_join, ctx->stream));
         }
      }
      OFRR_HIP(ctx, hipEventRecord(ev[6], ctx->stream));
      // the split Rayleigh-Ritz (X = the first kr columns); where X^T X is not
      // positive definite it reports rank -1 and the whole form runs instead
      h_ready = overlap ? ctx->ev_aux_join : nullptr;
      for (int form = o->rr_whole ? 1 : 0; form < 2; ++form) {
         if ((st = rayleigh_ritz(me, ka, form ? 0 : kr))) return st;
         // (already waited for inside; kept explicit so that nothing later on
         // this stream can overtake the second stream's work)
         if (h_ready) OFRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, h_ready, 0));
         OFRR_HIP(ctx, hipEventRecord(ev[7], ctx->stream));
         {
            const void *vb[3] = {X, Dd, Dmid};
            int dt[3] = {OFRR_F64, bdt, bdt}, kb[3] = {k, ka, ka}, ldc[3] = {me, me, me};
            int64_t rs[3] = {k, ka, ka}, cs[3] = {1, 1, 1};
            const double *cb[3] = {C, C + k, C + k + ka};
            if ((st = basis_update(ctx, n, k, 1 + nd, vb, dt, kb, rs, cs, cb, ldc, Xn, OFRR_F64, k, 1))) return st;
         }
         OFRR_HIP(ctx, hipMemcpyAsync(h_th, th, (size_t)k * 8, hipMemcpyDeviceToHost, ctx->stream));
         OFRR_HIP(ctx, hipMemcpyAsync(h_rank, rank, 4, hipMemcpyDeviceToHost, ctx->stream));
         OFRR_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
         OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
         if (h_rank[0] >= 0) break;
         ++rr_fallbacks;
      }
      std::swap(X, Xn);
      {
         float a, b, c, dd, e2;
         OFRR_HIP(ctx, hipEventElapsedTime(&a, ev[3], ev[4]));
         if (overlap) {
            // Z and the H blocks ran on the second stream beside the Rayleigh-Ritz:
            // c is G on the solver's stream, the second stream's time goes to t_aux
            float aux;
            OFRR_HIP(ctx, hipEventElapsedTime(&b, ev[8], ev[9]));
            OFRR_HIP(ctx, hipEventElapsedTime(&c, ev[4], ev[6]));
            OFRR_HIP(ctx, hipEventElapsedTime(&aux, ev[8], ev[10]));
            t_acc[7] += aux;
         } else {
            OFRR_HIP(ctx, hipEventElapsedTime(&b, ev[4], ev[5]));
            OFRR_HIP(ctx, hipEventElapsedTime(&c, ev[5], ev[6]));
         }
         OFRR_HIP(ctx, hipEventElapsedTime(&dd, ev[6], ev[7]));
         OFRR_HIP(ctx, hipEventElapsedTime(&e2, ev[7], ev[2]));
         t_acc[0] += a;
         t_acc[6] += b;
         t_acc[1] += c;
         t_acc[2] += dd;
         t_acc[3] += e2;
      }
      rank_min = std::min(rank_min, (int)h_rank[0]);
      if (h_rank[0] < kr) {
         st = set_error(ctx, kErrorSolverDivergence, "filtered_eig: truncated rank below the block width");
         break;
      }
      for (int j = 0; j < k; ++j) theta[j] = h_th[j];
   }
   OFRR_HIP(ctx, hipEventRecord(ev[2], ctx->stream));
   if (padded) {
      const int st2 = convert(ctx, n, kr, X, OFRR_F64, k, 1, Xio, OFRR_F64, kr, 1);
      if (st2) return st2;
   } else {
      OFRR_HIP(ctx, hipMemcpyAsync(Xio, X, (size_t)n * k * 8, hipMemcpyDeviceToDevice, ctx->stream));
   }
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   float total;
   OFRR_HIP(ctx, hipEventElapsedTime(&total, ev[0], ev[2]));
   for (int i = 0; i < 11; ++i) (void)hipEventDestroy(ev[i]);
   for (int j = 0; j < kr; ++j) {
      theta_host[j] = theta[j];
      res_host[j] = res[j];
   }
   if (stats) {
      stats->iterations = it;
      stats->converged = nconv;
      stats->spmm_calls = spmm_calls;
      stats->rank_min = rank_min;
      stats->filter_steps = filter_steps;
      stats->filter_col_steps = filter_col_steps;
      stats->filter_first_cols = filter_first_cols;
      stats->graph_replays = graph_replays;
      stats->graph_captures = graph_captures;
      stats->rr_fallbacks = rr_fallbacks;
      stats->wide_sweeps = wide_sweeps;
      stats->t_aux_ms = t_acc[7];
      stats->resid_passes = resid_passes;
      stats->t_total_ms = total;
      stats->t_filter_ms = t_acc[0];
      stats->t_zspmm_ms = t_acc[6];
      stats->t_gram_ms = t_acc[1];
      stats->t_rr_ms = t_acc[2];
      stats->t_update_ms = t_acc[3];
      stats->t_resid_ms = t_acc[5];
   }
   if (st) return st;
   return converged ? kSuccess : kErrorSolverMaxIterations;
}

}  // namespace ofrr

extern "C" {

int ofrr_filtered_eig(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rp, const int32_t *ci,
                      const double *va, const void *va_lo, int lo_dt, const ofrr_filter_opts *o,
                      double *Xio, double *theta_host, double *res_host, ofrr_filter_stats *stats) {
   OFRR_CHECK_CTX(ctx);
   (void)nnz;
   if (!o || n <= 0 || !rp || !ci || !va || !Xio || !theta_host || !res_host)
      return set_error(ctx, kErrorInvalidArgument, "filtered_eig: bad arguments");
   CsrOp op;
   op.n = n;
   op.rp = rp;
   op.ci = ci;
   op.va = va;
   op.vlo = va_lo ? va_lo : (const void *)va;
   op.vlo_dt = va_lo ? lo_dt : OFRR_F64;
   return filtered_core(ctx, op, o, Xio, theta_host, res_host, stats);
}

int ofrr_filtered_eig_state_bytes(int64_t n, const ofrr_filter_opts *o, int64_t *bytes) {
   if (!o || !bytes || n <= 0 || o->k < 1 || o->nd < 1 || o->degree < 2 || !dtype_valid(o->basis_dtype))
      return kErrorInvalidArgument;
   *bytes = (int64_t)filtered_state_bytes(n, *o);
   return kSuccess;
}


int ofrr_filtered_eig_dense(ofrr_ctx *ctx, int64_t n, const double *A, int64_t lda,
                            const ofrr_filter_opts *o, double *Xio, double *theta_host,
                            double *res_host, ofrr_filter_stats *stats) {
   OFRR_CHECK_CTX(ctx);
   if (!o || n <= 0 || !A || lda < n || !Xio || !theta_host || !res_host)
      return set_error(ctx, kErrorInvalidArgument, "filtered_eig_dense: bad arguments");
   const int k = o->k, bdt = o->basis_dtype;
   if (k < 2 || k > 128)
      return set_error(ctx, kErrorInvalidArgument, "filtered_eig_dense: k must lie in [2, 128]");
   const int kp = filter_width(k, 16);  // the operator's products take the stored width
   if (bdt != OFRR_F64 && bdt != OFRR_F16 && bdt != OFRR_BF16)
      return set_error(ctx, kErrorBLASUnsupportedType, "filtered_eig_dense: basis must be FP64, FP16 or BF16");
   DenseSymOp op;
   op.n = n;
   op.A = A;
   op.lda = lda;
   DevBuf alo, tb;
   if (bdt != OFRR_F64) {
      op.lo_dt = bdt;
      op.ldlo = round8(n);
      if (!alo.alloc((size_t)n * op.ldlo * 2)) return set_error(ctx, kErrorCUDAMemory, "filtered_eig_dense: A copy");
      // symmetric: the row-major copy of A is its column-major storage
      int st = convert(ctx, n, n, A, OFRR_F64, 1, lda, alo.p, bdt, op.ldlo, 1);
      if (st) return st;
      op.Alo = alo.p;
   }
   DevBuf d64;
   if (!tb.alloc((size_t)n * kp * 8) || !d64.alloc((size_t)n * kp * 8))
      return set_error(ctx, kErrorCUDAMemory, "filtered_eig_dense: product");
   op.T = tb.p;
   op.D64 = (double *)d64.p;
   int st = filtered_core(ctx, op, o, Xio, theta_host, res_host, stats);
   (void)hipStreamSynchronize(ctx->stream);
   return st;
}

int ofrr_svd_filtered(ofrr_ctx *ctx, int64_t M, int64_t N, const float *A, int64_t lda,
                      const ofrr_filter_opts *o, double *Vio, double *U, double *sigma_host,
                      double *res_host, ofrr_filter_stats *stats) {
   OFRR_CHECK_CTX(ctx);
   if (!o || M <= 0 || N <= 0 || !A || lda < M || !Vio || !sigma_host || !res_host)
      return set_error(ctx, kErrorInvalidArgument, "svd_filtered: bad arguments");
   const int k = o->k, bdt = o->basis_dtype;
   const int kp = filter_width(k, 16);  // the operator's products take the stored width
   if (k < 2 || k > 128 || kp > N)
      return set_error(ctx, kErrorInvalidArgument, "svd_filtered: k must lie in [2, 128], its storage width <= N");
   if (bdt != OFRR_F64 && bdt != OFRR_BF16)
      return set_error(ctx, kErrorBLASUnsupportedType, "svd_filtered: basis must be BF16 or FP64");
   AtaOp op;
   op.n = N;
   op.M = M;
   op.Acm = A;
   op.lda = lda;
   DevBuf arm, brm, bcm, y64, ylo, tb, d64;
   int st;
   if (!arm.alloc((size_t)M * N * 4) || !y64.alloc((size_t)M * kp * 8) || !tb.alloc((size_t)N * kp * 8) ||
       !d64.alloc((size_t)N * kp * 8))
      return set_error(ctx, kErrorCUDAMemory, "svd_filtered: operator copies");
   op.D64 = (double *)d64.p;
   if ((st = convert(ctx, M, N, A, OFRR_F32, 1, lda, arm.p, OFRR_F32, N, 1))) return st;
   op.Arm = (const float *)arm.p;
   op.Y64 = (double *)y64.p;
   op.T = tb.p;
   if (bdt == OFRR_BF16) {
      op.ldr = round8(N);
      op.ldc = round8(M);
      if (!brm.alloc((size_t)M * op.ldr * 2) || !bcm.alloc((size_t)N * op.ldc * 2) ||
          !ylo.alloc((size_t)M * kp * 2))
         return set_error(ctx, kErrorCUDAMemory, "svd_filtered: BF16 copies");
      if ((st = convert(ctx, M, N, A, OFRR_F32, 1, lda, brm.p, OFRR_BF16, op.ldr, 1))) return st;
      if ((st = convert(ctx, M, N, A, OFRR_F32, 1, lda, bcm.p, OFRR_BF16, 1, op.ldc))) return st;
      op.Blo_rm = brm.p;
      op.Blo_cm = bcm.p;
      op.Ylo = ylo.p;
   }
   std::vector<double> th(k);
   st = filtered_core(ctx, op, o, Vio, th.data(), res_host, stats);
   if (st != kSuccess && st != kErrorSolverMaxIterations) {
      (void)hipStreamSynchronize(ctx->stream);
      return st;
   }
   for (int j = 0; j < k; ++j) sigma_host[j] = std::sqrt(std::max(0.0, th[j]));
   if (U) {
      // U = A V diag(1 / sigma), FP64 accumulation
      int st2;
      if (kp == k) {
         st2 = dense_gemm(ctx, M, N, k, op.Arm, OFRR_F32, N, Vio, OFRR_F64, k, 1.0, 0.0, U, OFRR_F64, k);
      } else {
         // the product takes the stored width: V zero-padded to kp columns, A V
         // into the M x kp scratch, its first k columns to U
         st2 = zero_bytes(ctx, op.D64, (size_t)N * kp * 8);
         if (!st2) st2 = convert(ctx, N, k, Vio, OFRR_F64, k, 1, op.D64, OFRR_F64, kp, 1);
         if (!st2)
            st2 = dense_gemm(ctx, M, N, kp, op.Arm, OFRR_F32, N, op.D64, OFRR_F64, kp, 1.0, 0.0, op.Y64, OFRR_F64, kp);
         if (!st2) st2 = convert(ctx, M, k, op.Y64, OFRR_F64, kp, 1, U, OFRR_F64, k, 1);
      }
      if (st2) return st2;
      std::vector<double> inv(k);
      for (int j = 0; j < k; ++j) inv[j] = sigma_host[j] > 0.0 ? 1.0 / sigma_host[j] : 0.0;
      double *dinv = (double *)workspace(ctx, (size_t)k * 8);
      if (!dinv) return set_error(ctx, kErrorCUDAMemory, "svd_filtered: workspace");
      OFRR_HIP(ctx, hipMemcpyAsync(dinv, inv.data(), (size_t)k * 8, hipMemcpyHostToDevice, ctx->stream));
      scale_cols_rm_kernel<<<(unsigned)((M * k + 255) / 256), 256, 0, ctx->stream>>>(M, k, U, dinv);
      OFRR_LAUNCH(ctx);
   }
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   return st;
}

}  // extern "C"
