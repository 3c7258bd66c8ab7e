// General dense products and element-wise updates behind the C++ adapter's
// gemm<T_I, T_O, T_COMPUTE> (CPP/linalg/blas/mvops.hpp:79-82, mvops.cu:423-605)
// and Vector / Matrix::scale (CPP/containers/vector.hpp:140, matrix.hpp:145).
//
// gemm: C = alpha op(A) op(B) + beta C in BLAS column-major form (any m, n, k,
// any transposes, leading dimensions), the operand formats FP64 / FP32 / FP16 /
// BF16, the accumulator FP64 (compute FP64) or FP32 (compute FP32 or FP16: the
// reference's gemm<__half, __half, __half> accumulates in FP16; this one
// accumulates those products in FP32 and rounds once at the end).  One
// workgroup of 256 threads owns a 64 x 64 tile of C and walks k in steps of
// 16: each step stages the 64 x 16 slice of op(A) and the 16 x 64 slice of
// op(B) in LDS (converted to the accumulator type; the global reads follow
// the contiguous direction of each operand, so both transposes load
// coalesced), then each thread accumulates its 4 x 4 outputs with fused
// multiply-adds in increasing k.  The order is fixed, so results are
// reproducible run to run.  beta = 0 does not read C.
#include "internal.hpp"

namespace ofrr {

constexpr int kGT = 64;   // C tile
constexpr int kGK = 16;   // k step

template <typename TA, typename TB, typename TC, typename ACC>
__global__ __launch_bounds__(256) void gemm_tile_kernel(int transA, int transB, int64_t m, int64_t n,
                                                        int64_t k, double alpha, const TA *__restrict__ A,
                                                        int64_t lda, const TB *__restrict__ B, int64_t ldb,
                                                        double beta, TC *__restrict__ C, int64_t ldc) {
   __shared__ ACC As[kGK][kGT + 1];
   __shared__ ACC Bs[kGK][kGT + 1];
   const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
   const int64_t i0 = (int64_t)blockIdx.x * kGT, j0 = (int64_t)blockIdx.y * kGT;
   ACC acc[4][4];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0;
   for (int64_t p0 = 0; p0 < k; p0 += kGK) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
         const int e = threadIdx.x + 256 * u;
         // op(A)(i, p): along i for A as stored, along p for A^T
         int ii, pp;
         if (!transA) { ii = e % kGT; pp = e / kGT; } else { pp = e % kGK; ii = e / kGK; }
         const int64_t gi = i0 + ii, gp = p0 + pp;
         ACC va = 0;
         if (gi < m && gp < k)
            va = (ACC)Elem<TA>::load(transA ? A[gp + gi * lda] : A[gi + gp * lda]);
         As[pp][ii] = va;
         // op(B)(p, j): along p for B as stored, along j for B^T
         int jj, qq;
         if (!transB) { qq = e % kGK; jj = e / kGK; } else { jj = e % kGT; qq = e / kGT; }
         const int64_t gj = j0 + jj, gq = p0 + qq;
         ACC vb = 0;
         if (gj < n && gq < k)
            vb = (ACC)Elem<TB>::load(transB ? B[gj + gq * ldb] : B[gq + gj * ldb]);
         Bs[qq][jj] = vb;
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < kGK; ++p) {
         ACC ra[4], rb[4];
#pragma unroll
         for (int a = 0; a < 4; ++a) ra[a] = As[p][tx + 16 * a];
#pragma unroll
         for (int b = 0; b < 4; ++b) rb[b] = Bs[p][ty + 16 * b];
#pragma unroll
         for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(ra[a], rb[b], acc[a][b]);
      }
      __syncthreads();
   }
#pragma unroll
   for (int b = 0; b < 4; ++b) {
      const int64_t j = j0 + ty + 16 * b;
      if (j >= n) continue;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
         const int64_t i = i0 + tx + 16 * a;
         if (i >= m) continue;
         TC *c = C + i + j * ldc;
         double r = alpha * (double)acc[a][b];
         if (beta != 0.0) r = fma(beta, (double)Elem<TC>::load(*c), r);
         *c = Elem<TC>::store(r);
      }
   }
}

// X = alpha X + shift over a rows x cols block with strides (element (i, j) at
// X[i * rs + j * cs])
template <typename T>
__global__ void axpb_kernel(int64_t rows, int64_t cols, T *__restrict__ X, int64_t rs, int64_t cs,
                            double alpha, double shift) {
   const int64_t total = rows * cols;
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
        t += (int64_t)gridDim.x * blockDim.x) {
      const int64_t i = t % rows, j = t / rows;
      T *p = X + i * rs + j * cs;
      *p = Elem<T>::store(fma(alpha, (double)Elem<T>::load(*p), shift));
   }
}

// the four dense formats (FP8 is a growth-block format only)
template <typename F> inline int dispatch_dense(int dt, F &&f) {
   switch (dt) {
      case OFRR_F64: return f(double{});
      case OFRR_F32: return f(float{});
      case OFRR_F16: return f(__half{});
      case OFRR_BF16: return f(__hip_bfloat16{});
      default: return kErrorBLASUnsupportedType;
   }
}

// A and B share a format (the reference's T_I), C has its own (T_O)
template <typename ACC>
static int gemm_dispatch(ofrr_ctx *ctx, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                         const void *A, int idt, int64_t lda, const void *B, int64_t ldb, double beta,
                         void *C, int cdt, int64_t ldc) {
   const dim3 grid((unsigned)((m + kGT - 1) / kGT), (unsigned)((n + kGT - 1) / kGT));
   return dispatch_dense(idt, [&](auto vi) -> int {
      using TI = decltype(vi);
      return dispatch_dense(cdt, [&](auto vc) -> int {
         using TC = decltype(vc);
         gemm_tile_kernel<TI, TI, TC, ACC><<<grid, 256, 0, ctx->stream>>>(
             ta, tb, m, n, k, alpha, (const TI *)A, lda, (const TI *)B, ldb, beta, (TC *)C, ldc);
         OFRR_LAUNCH(ctx);
         return kSuccess;
      });
   });
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_gemm(ofrr_ctx *ctx, int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
              const void *A, int a_dtype, int64_t lda, const void *B, int b_dtype, int64_t ldb,
              double beta, void *C, int c_dtype, int64_t ldc, int compute_dtype) {
   OFRR_CHECK_CTX(ctx);
   if (m < 0 || n < 0 || k < 0) return set_error(ctx, kErrorBLASInvalidSize, "gemm: negative size");
   if (m == 0 || n == 0) return kSuccess;
   if (!C || (k > 0 && (!A || !B))) return set_error(ctx, kErrorInvalidArgument, "gemm: null pointer");
   const int64_t ar = transA ? k : m, br = transB ? n : k;
   if (ldc < m || (k > 0 && (lda < std::max<int64_t>(1, ar) || ldb < std::max<int64_t>(1, br))))
      return set_error(ctx, kErrorBLASInvalidSize, "gemm: leading dimension too small");
   for (int dt : {a_dtype, b_dtype, c_dtype})
      if (dt != OFRR_F64 && dt != OFRR_F32 && dt != OFRR_F16 && dt != OFRR_BF16)
         return set_error(ctx, kErrorBLASUnsupportedType, "gemm: operand formats FP64, FP32, FP16, BF16");
   if (a_dtype != b_dtype) return set_error(ctx, kErrorBLASUnsupportedType, "gemm: A and B must share a format");
   if (compute_dtype == OFRR_F64)
      return gemm_dispatch<double>(ctx, transA ? 1 : 0, transB ? 1 : 0, m, n, k, alpha, A, a_dtype, lda, B,
                                   ldb, beta, C, c_dtype, ldc);
   if (compute_dtype == OFRR_F32 || compute_dtype == OFRR_F16 || compute_dtype == OFRR_BF16)
      return gemm_dispatch<float>(ctx, transA ? 1 : 0, transB ? 1 : 0, m, n, k, alpha, A, a_dtype, lda, B,
                                  ldb, beta, C, c_dtype, ldc);
   return set_error(ctx, kErrorBLASUnsupportedType, "gemm: compute format");
}

int ofrr_axpb(ofrr_ctx *ctx, int dtype, int64_t rows, int64_t cols, void *X, int64_t rs, int64_t cs,
              double alpha, double shift) {
   OFRR_CHECK_CTX(ctx);
   if (rows < 0 || cols < 0) return set_error(ctx, kErrorInvalidArgument, "axpb: negative size");
   if (rows == 0 || cols == 0) return kSuccess;
   if (!X) return set_error(ctx, kErrorInvalidArgument, "axpb: null pointer");
   if (dtype == OFRR_FP8 || !dtype_valid(dtype)) return set_error(ctx, kErrorBLASUnsupportedType, "axpb: dtype");
   const int64_t total = rows * cols;
   const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
   return dispatch_dense(dtype, [&](auto v) -> int {
      using T = decltype(v);
      axpb_kernel<T><<<g, 256, 0, ctx->stream>>>(rows, cols, (T *)X, rs, cs, alpha, shift);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

}  // extern "C"
