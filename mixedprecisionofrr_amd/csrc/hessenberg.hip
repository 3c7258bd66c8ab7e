// Hessenberg process on the GPU: a pivoted, max-abs scaled, LU-on-columns
// basis builder with no orthogonalization.  Semantics of
// MATLAB/+src/+qr/hessenberg.m:55-74 (and of the C++ variants in
// CPP/linalg/factorization/hessenberg.cu:405-604): column i is eliminated
// against every earlier kept column j at its pivot row perm(j) with the
// current value R(j,i) = v(perm(j)), then divided by its signed max-abs entry
// R(i,i) (first index on ties); a column whose pivot is below select_tol is
// dropped (skip = 1) and keeps its unscaled residual.
//
// Right-looking formulation, all on the device (the reference copies the skip
// flag to the host every column, hessenberg.cu:519): per column one argmax
// (two small kernels), one kernel that records the pivot decision and gathers
// the pivot row of the trailing columns into R, and one kernel over (row block,
// group of 16 trailing columns) in which each thread owns a row: it scales the
// row's entry of column i and eliminates it from the group's columns.
#include "internal.hpp"

#include <algorithm>

namespace ofrr {

// decide keep/drop for column i and gather R(i, c) = Q(perm_i, c), c > i
// (one workgroup).  variant 0: a dropped column takes no part in later
// eliminations (R(i, c) = 0); variant 1 (C++ v3): it does, with its argmax
// row and its unscaled residual.
template <typename T>
__global__ void hess_pivot_kernel(int n, int i, const T *__restrict__ Q, int64_t rs, int64_t cs,
                                  double select_tol, double *__restrict__ R, int ldr,
                                  int32_t *__restrict__ perm, int32_t *__restrict__ skip,
                                  int variant) {
   const double piv = R[i + (int64_t)i * ldr];
   const bool drop = !(fabs(piv) >= select_tol);
   const int32_t pr = perm[i];
   __syncthreads();  // every thread has read perm[i] before it is rewritten
   if (threadIdx.x == 0) {
      skip[i] = drop ? 1 : 0;
      if (drop) perm[i] = -1;
   }
   const bool gather = !drop || variant == 1;
   for (int c = i + 1 + threadIdx.x; c < n; c += blockDim.x)
      R[i + (int64_t)c * ldr] = gather ? (double)Elem<T>::load(Q[(int64_t)pr * rs + (int64_t)c * cs]) : 0.0;
}

// Elimination of column i (not dropped) from the trailing columns of this
// workgroup's column group (kHessCols columns, blockIdx.y), one row per thread,
// with q = Q(r, i) / R(i,i) rounded to the storage format.  Every column group
// reads column i unscaled, so the scaled column is stored by the next launch:
// column group 0 stores the scaled column `prev` (the previous pivot column,
// when kept).  i < 0: that store only.
constexpr int kHessCols = 16;
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_update_kernel(int64_t m, int n, int i, int prev,
                                                          T *__restrict__ Q, int64_t rs, int64_t cs,
                                                          const double *__restrict__ R, int ldr,
                                                          const int32_t *__restrict__ skip,
                                                          int variant) {
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= m) return;
   if (blockIdx.y == 0 && prev >= 0 && !skip[prev]) {
      T *qp = Q + r * rs + (int64_t)prev * cs;
      *qp = Elem<T>::store((double)((ACC)Elem<T>::load(*qp) / (ACC)R[prev + (int64_t)prev * ldr]));
   }
   if (i < 0) return;
   const bool dropped = skip[i] != 0;
   if (dropped && variant == 0) return;
   // a kept column enters scaled (as it is stored), a dropped one (variant 1)
   // with its unscaled residual
   const ACC piv = (ACC)R[i + (int64_t)i * ldr];
   const ACC q = dropped ? (ACC)Elem<T>::load(Q[r * rs + (int64_t)i * cs])
                         : (ACC)Elem<T>::load(Elem<T>::store(
                               (double)((ACC)Elem<T>::load(Q[r * rs + (int64_t)i * cs]) / piv)));
   const int c0 = i + 1 + (int)blockIdx.y * kHessCols, c1 = min(n, c0 + kHessCols);
   for (int c = c0; c < c1; ++c) {
      T *p = Q + r * rs + (int64_t)c * cs;
      const ACC coef = (ACC)R[i + (int64_t)c * ldr];
      *p = Elem<T>::store((double)fma(-coef, q, (ACC)Elem<T>::load(*p)));
   }
}

__global__ void zero_lower_kernel(int n, double *__restrict__ R, int ldr) {
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n * n) return;
   int r = t % n, c = t / n;
   if (r > c) R[r + (int64_t)c * ldr] = 0.0;
}

template <typename T, typename ACC>
static int hessenberg_impl(ofrr_ctx *ctx, int64_t m, int n, T *Q, int64_t rs, int64_t cs,
                           double tol, double *R, int ldr, int32_t *perm, int32_t *skip,
                           int variant) {
   zero_lower_kernel<<<(n * n + 255) / 256, 256, 0, ctx->stream>>>(n, R, ldr);
   OFRR_LAUNCH(ctx);
   const int ublocks = (int)((m + 255) / 256);
   for (int i = 0; i < n; ++i) {
      // argmax of column i -> perm[i] (0-based), R(i,i) (signed)
      int st = argmax_columns(ctx, Elem<T>::id, m, 1, Q + (int64_t)i * cs, rs, 0, 0, perm + i,
                              R + i + (int64_t)i * ldr);
      if (st) return st;
      hess_pivot_kernel<T><<<1, 256, 0, ctx->stream>>>(
          n, i, Q, rs, cs, tol, R, ldr, perm, skip, variant);
      dim3 grid(ublocks, std::max(1, (n - i - 1 + kHessCols - 1) / kHessCols));
      hess_update_kernel<T, ACC><<<grid, 256, 0, ctx->stream>>>(m, n, i, i - 1, Q, rs, cs, R, ldr, skip,
                                                                variant);
      OFRR_LAUNCH(ctx);
   }
   hess_update_kernel<T, ACC><<<ublocks, 256, 0, ctx->stream>>>(m, n, -1, n - 1, Q, rs, cs, R, ldr, skip,
                                                                variant);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---- one column at a time (hessenberg_v3_step, hessenberg.cu:617-801) ------
// Column i against the finished columns 0..i-1, left-looking.  Its coefficient
// against column j is its entry at j's pivot row after the eliminations with
// 0..j-1, so all of them follow from the column's entries at the pivot rows by
// a forward substitution over the pivot rows of the finished columns (one
// workgroup, the same rounded fused multiply-adds as the full elimination);
// then every row applies the eliminations in column order in one pass.  The
// result is bit for bit that of the right-looking factorization.
template <typename T, typename ACC>
__global__ __launch_bounds__(1024) void hess_step_coef_kernel(int i, const T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, const int32_t *__restrict__ prow1,
                                                              const int32_t *__restrict__ skip, int variant,
                                                              double *__restrict__ R, int ldr) {
   extern __shared__ double val[];  // i entries: column i at the pivot rows, being eliminated
   for (int t = threadIdx.x; t < i; t += blockDim.x)
      val[t] = (double)Elem<T>::load(Q[(int64_t)(prow1[t] - 1) * rs + (int64_t)i * cs]);
   __syncthreads();
   for (int j = 0; j < i; ++j) {
      const bool use = variant == 1 || !skip[j];
      const double c = use ? val[j] : 0.0;
      if (threadIdx.x == 0) R[j + (int64_t)i * ldr] = c;
      if (use)
         for (int t = j + 1 + threadIdx.x; t < i; t += blockDim.x) {
            const ACC q = (ACC)Elem<T>::load(Q[(int64_t)(prow1[t] - 1) * rs + (int64_t)j * cs]);
            val[t] = (double)Elem<T>::load(Elem<T>::store((double)fma(-(ACC)c, q, (ACC)val[t])));
         }
      __syncthreads();
   }
}

// every row of column i: the eliminations with the columns j < i in order
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_step_apply_kernel(int64_t m, int i, T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, const int32_t *__restrict__ skip,
                                                              int variant, const double *__restrict__ R,
                                                              int ldr) {
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= m) return;
   T *p = Q + r * rs + (int64_t)i * cs;
   ACC v = (ACC)Elem<T>::load(*p);
   for (int j = 0; j < i; ++j) {
      if (variant == 0 && skip[j]) continue;
      const ACC c = (ACC)R[j + (int64_t)i * ldr];
      v = (ACC)Elem<T>::load(Elem<T>::store((double)fma(-c, (ACC)Elem<T>::load(Q[r * rs + (int64_t)j * cs]), v)));
   }
   *p = Elem<T>::store((double)v);
}

// pivot of column i (argmax written by argmax_columns to prow1[i] (1-based)
// and R(i, i)): drop or scale
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_step_scale_kernel(int64_t m, int i, T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, double select_tol,
                                                              const double *__restrict__ R, int ldr,
                                                              int32_t *__restrict__ skip) {
   const double piv = R[i + (int64_t)i * ldr];
   const bool drop = !(fabs(piv) >= select_tol);
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r == 0) skip[i] = drop ? 1 : 0;
   if (drop || r >= m) return;
   T *p = Q + r * rs + (int64_t)i * cs;
   *p = Elem<T>::store((double)((ACC)Elem<T>::load(*p) / (ACC)piv));
}

template <typename T, typename ACC>
static int hessenberg_step_impl(ofrr_ctx *ctx, int variant, int64_t m, int i, T *Q, int64_t rs, int64_t cs,
                                double tol, double *R, int ldr, int32_t *prow1, int32_t *skip) {
   if (i > 0) {
      hess_step_coef_kernel<T, ACC><<<1, 1024, (size_t)i * sizeof(double), ctx->stream>>>(
          i, Q, rs, cs, prow1, skip, variant, R, ldr);
      hess_step_apply_kernel<T, ACC><<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(
          m, i, Q, rs, cs, skip, variant, R, ldr);
      OFRR_LAUNCH(ctx);
   }
   int st = argmax_columns(ctx, Elem<T>::id, m, 1, Q + (int64_t)i * cs, rs, 0, 1, prow1 + i,
                           R + i + (int64_t)i * ldr);
   if (st) return st;
   hess_step_scale_kernel<T, ACC><<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(
       m, i, Q, rs, cs, tol, R, ldr, skip);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

static int hessenberg_variant(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q, int64_t rs,
                              int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                              int32_t *skip) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n <= 0 || !Q || !R || !perm || !skip || ldr < n)
      return set_error(ctx, kErrorInvalidArgument, "hessenberg: bad arguments");
   if (m < n) return set_error(ctx, kErrorInvalidArgument, "hessenberg: A must be a tall-thin matrix");
   switch (qdt) {
      case OFRR_F64:
         return hessenberg_impl<double, double>(ctx, m, n, (double *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_F32:
         return hessenberg_impl<float, float>(ctx, m, n, (float *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_F16:
         return hessenberg_impl<__half, float>(ctx, m, n, (__half *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_BF16:
         return hessenberg_impl<__hip_bfloat16, float>(ctx, m, n, (__hip_bfloat16 *)Q, rs, cs, select_tol, R, ldr,
                                                       perm, skip, variant);
      default:
         return set_error(ctx, kErrorBLASUnsupportedType, "hessenberg: unsupported dtype");
   }
}

extern "C" int ofrr_hessenberg(ofrr_ctx *ctx, int qdt, int64_t m, int n, void *Q, int64_t rs,
                               int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                               int32_t *skip) {
   return hessenberg_variant(ctx, 0, qdt, m, n, Q, rs, cs, select_tol, R, ldr, perm, skip);
}

extern "C" int ofrr_hessenberg_v3(ofrr_ctx *ctx, int qdt, int64_t m, int n, void *Q, int64_t rs,
                                  int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                                  int32_t *skip) {
   return hessenberg_variant(ctx, 1, qdt, m, n, Q, rs, cs, select_tol, R, ldr, perm, skip);
}

extern "C" int ofrr_hessenberg_step(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q,
                                    int64_t rs, int64_t cs, int i, double select_tol, double *R, int ldr,
                                    int32_t *prow1, int32_t *skip, int32_t *skipped_host) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n <= 0 || i < 0 || i >= n || !Q || !R || !prow1 || !skip || ldr < n ||
       (variant != 0 && variant != 1))
      return set_error(ctx, kErrorInvalidArgument, "hessenberg_step: bad arguments");
   int st;
   switch (qdt) {
      case OFRR_F64:
         st = hessenberg_step_impl<double, double>(ctx, variant, m, i, (double *)Q, rs, cs, select_tol, R, ldr,
                                                   prow1, skip);
         break;
      case OFRR_F32:
         st = hessenberg_step_impl<float, float>(ctx, variant, m, i, (float *)Q, rs, cs, select_tol, R, ldr,
                                                 prow1, skip);
         break;
      case OFRR_F16:
         st = hessenberg_step_impl<__half, float>(ctx, variant, m, i, (__half *)Q, rs, cs, select_tol, R, ldr,
                                                  prow1, skip);
         break;
      case OFRR_BF16:
         st = hessenberg_step_impl<__hip_bfloat16, float>(ctx, variant, m, i, (__hip_bfloat16 *)Q, rs, cs,
                                                          select_tol, R, ldr, prow1, skip);
         break;
      default:
         return set_error(ctx, kErrorBLASUnsupportedType, "hessenberg_step: unsupported dtype");
   }
   if (st) return st;
   if (skipped_host) {
      int32_t *h = (int32_t *)pinned(ctx, sizeof(int32_t));
      if (!h) return set_error(ctx, kErrorCUDAMemory, "hessenberg_step: pinned");
      OFRR_HIP(ctx, hipMemcpyAsync(h, skip + i, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      *skipped_host = *h;
   }
   return kSuccess;
}
