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
// (one workgroup)
template <typename T>
__global__ void hess_pivot_kernel(int n, int i, const T *__restrict__ Q, int64_t rs, int64_t cs,
                                  double select_tol, double *__restrict__ R, int ldr,
                                  int32_t *__restrict__ perm, int32_t *__restrict__ skip) {
   const double piv = R[i + (int64_t)i * ldr];
   const bool drop = !(fabs(piv) >= select_tol);
   const int32_t pr = perm[i];
   __syncthreads();  // every thread has read perm[i] before it is rewritten
   if (threadIdx.x == 0) {
      skip[i] = drop ? 1 : 0;
      if (drop) perm[i] = -1;
   }
   for (int c = i + 1 + threadIdx.x; c < n; c += blockDim.x)
      R[i + (int64_t)c * ldr] = drop ? 0.0 : (double)Elem<T>::load(Q[(int64_t)pr * rs + (int64_t)c * cs]);
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
                                                          const int32_t *__restrict__ skip) {
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= m) return;
   if (blockIdx.y == 0 && prev >= 0 && !skip[prev]) {
      T *qp = Q + r * rs + (int64_t)prev * cs;
      *qp = Elem<T>::store((double)((ACC)Elem<T>::load(*qp) / (ACC)R[prev + (int64_t)prev * ldr]));
   }
   if (i < 0 || skip[i]) return;
   const ACC piv = (ACC)R[i + (int64_t)i * ldr];
   const ACC q = (ACC)Elem<T>::load(
       Elem<T>::store((double)((ACC)Elem<T>::load(Q[r * rs + (int64_t)i * cs]) / piv)));
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
                           double tol, double *R, int ldr, int32_t *perm, int32_t *skip) {
   zero_lower_kernel<<<(n * n + 255) / 256, 256, 0, ctx->stream>>>(n, R, ldr);
   OFRR_LAUNCH(ctx);
   const int ublocks = (int)((m + 255) / 256);
   for (int i = 0; i < n; ++i) {
      // argmax of column i -> perm[i] (0-based), R(i,i) (signed)
      int st = argmax_columns(ctx, Elem<T>::id, m, 1, Q + (int64_t)i * cs, rs, 0, 0, perm + i,
                              R + i + (int64_t)i * ldr);
      if (st) return st;
      hess_pivot_kernel<T><<<1, 256, 0, ctx->stream>>>(
          n, i, Q, rs, cs, tol, R, ldr, perm, skip);
      dim3 grid(ublocks, std::max(1, (n - i - 1 + kHessCols - 1) / kHessCols));
      hess_update_kernel<T, ACC><<<grid, 256, 0, ctx->stream>>>(m, n, i, i - 1, Q, rs, cs, R, ldr, skip);
      OFRR_LAUNCH(ctx);
   }
   hess_update_kernel<T, ACC><<<ublocks, 256, 0, ctx->stream>>>(m, n, -1, n - 1, Q, rs, cs, R, ldr, skip);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

extern "C" int ofrr_hessenberg(ofrr_ctx *ctx, int qdt, int64_t m, int n, void *Q, int64_t rs,
                               int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                               int32_t *skip) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n <= 0 || !Q || !R || !perm || !skip || ldr < n)
      return set_error(ctx, kErrorInvalidArgument, "hessenberg: bad arguments");
   if (m < n) return set_error(ctx, kErrorInvalidArgument, "hessenberg: A must be a tall-thin matrix");
   switch (qdt) {
      case OFRR_F64:
         return hessenberg_impl<double, double>(ctx, m, n, (double *)Q, rs, cs, select_tol, R, ldr, perm, skip);
      case OFRR_F32:
         return hessenberg_impl<float, float>(ctx, m, n, (float *)Q, rs, cs, select_tol, R, ldr, perm, skip);
      case OFRR_F16:
         return hessenberg_impl<__half, float>(ctx, m, n, (__half *)Q, rs, cs, select_tol, R, ldr, perm, skip);
      case OFRR_BF16:
         return hessenberg_impl<__hip_bfloat16, float>(ctx, m, n, (__hip_bfloat16 *)Q, rs, cs, select_tol, R, ldr, perm, skip);
      default:
         return set_error(ctx, kErrorBLASUnsupportedType, "hessenberg: unsupported dtype");
   }
}
