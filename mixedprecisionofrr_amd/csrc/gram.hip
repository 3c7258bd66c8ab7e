// Projection products and back-transform.
//
//   gram          G = alpha V^T W + beta G: the projections V^T V and V^T (A V)
//                 of eig_rr.m:49,51 (gemm with transA = true, mvops.cu:423-605).
//                 Tall-skinny: each workgroup owns a 64 x 64 output tile and a
//                 contiguous range of rows, stages 32-row slabs of V and W in
//                 LDS (converted to FP64), and accumulates a 4 x 4 register
//                 block per thread with FP64 fused multiply-adds.  The per-range
//                 partial tiles are summed in a fixed order by a second kernel,
//                 so the result is bitwise reproducible and independent of other
//                 work on the device.
//   basis_update  Y = sum_b V_b C_b: the back-transform V <- V U (eig_rr.m:56)
//                 for a basis whose column blocks are stored in different
//                 precisions (FP64 Ritz block, low-precision growth blocks).
//   small_gemm    C = alpha op(A) B + beta C for the k x k Rayleigh-Ritz algebra.
#include "internal.hpp"

namespace ofrr {

constexpr int GT = 64;   // output tile
constexpr int GR = 32;   // rows staged per slab

template <typename TV, typename TW>
__global__ __launch_bounds__(256) void gram_partial_kernel(
    int64_t n, int k1, int k2, const TV *__restrict__ V, int64_t vrs, int64_t vcs,
    const TW *__restrict__ W, int64_t wrs, int64_t wcs, int64_t rows_per_chunk,
    double *__restrict__ part) {
   __shared__ double sV[GR][GT + 1];
   __shared__ double sW[GR][GT + 1];
   const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
   const int i0 = blockIdx.y * GT, j0 = blockIdx.z * GT;
   const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
   const int64_t r1 = min(n, r0 + rows_per_chunk);
   double acc[4][4];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
   const bool vrow = (vcs == 1), wrow = (wcs == 1);
   for (int64_t rb = r0; rb < r1; rb += GR) {
      // stage V(rb:rb+GR, i0:i0+GT) and W(rb:rb+GR, j0:j0+GT)
      for (int t = threadIdx.x; t < GR * GT; t += 256) {
         int rr, cc;
         if (vrow) { rr = t / GT; cc = t % GT; } else { cc = t / GR; rr = t % GR; }
         int64_t row = rb + rr;
         int col = i0 + cc;
         sV[rr][cc] = (row < r1 && col < k1) ? (double)Elem<TV>::load(V[row * vrs + (int64_t)col * vcs]) : 0.0;
         if (wrow) { rr = t / GT; cc = t % GT; } else { cc = t / GR; rr = t % GR; }
         row = rb + rr;
         col = j0 + cc;
         sW[rr][cc] = (row < r1 && col < k2) ? (double)Elem<TW>::load(W[row * wrs + (int64_t)col * wcs]) : 0.0;
      }
      __syncthreads();
#pragma unroll 4
      for (int r = 0; r < GR; ++r) {
         double va[4], wb[4];
#pragma unroll
         for (int a = 0; a < 4; ++a) va[a] = sV[r][ty * 4 + a];
#pragma unroll
         for (int b = 0; b < 4; ++b) wb[b] = sW[r][tx * 4 + b];
#pragma unroll
         for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(va[a], wb[b], acc[a][b]);
      }
      __syncthreads();
   }
   // partial tile layout: part[chunk][tile_i][tile_j][GT*GT]
   const int64_t tile = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
   double *p = part + ((int64_t)blockIdx.x * gridDim.y * gridDim.z + tile) * (GT * GT);
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) p[(ty * 4 + a) * GT + tx * 4 + b] = acc[a][b];
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(int nchunks, int ntiles_i, int ntiles_j,
                                                          int k1, int k2,
                                                          const double *__restrict__ part,
                                                          double alpha, double beta,
                                                          double *__restrict__ G, int ldg) {
   const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   const int64_t per_chunk = (int64_t)ntiles_i * ntiles_j * GT * GT;
   if (e >= per_chunk) return;
   const int64_t tile = e / (GT * GT);
   const int w = (int)(e % (GT * GT));
   const int i = (int)(tile / ntiles_j) * GT + w / GT;
   const int j = (int)(tile % ntiles_j) * GT + w % GT;
   if (i >= k1 || j >= k2) return;
   double s = 0.0;
   for (int c = 0; c < nchunks; ++c) s += part[c * per_chunk + e];
   double *g = G + i + (int64_t)j * ldg;
   *g = (beta == 0.0) ? alpha * s : fma(alpha, s, beta * *g);
}

int gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
         int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha, double beta,
         double *G, int ldg) {
   if (k1 <= 0 || k2 <= 0) return kSuccess;
   const int ti = (k1 + GT - 1) / GT, tj = (k2 + GT - 1) / GT;
   // about 512 workgroups in total, each with a whole number of 32-row slabs
   int64_t want = std::max<int64_t>(1, 512 / (ti * tj));
   int64_t rows = (n + want - 1) / want;
   rows = std::max<int64_t>(GR * 4, ((rows + GR - 1) / GR) * GR);
   const int nchunks = (int)((n + rows - 1) / rows);
   const int64_t per_chunk = (int64_t)ti * tj * GT * GT;
   double *part = (double *)workspace(ctx, (size_t)nchunks * per_chunk * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "gram: workspace");
   int st = dispatch_dtype(vdt, [&](auto tv) -> int {
      using TV = decltype(tv);
      return dispatch_dtype(wdt, [&](auto tw) -> int {
         using TW = decltype(tw);
         gram_partial_kernel<TV, TW><<<dim3(nchunks, ti, tj), 256, 0, ctx->stream>>>(
             n, k1, k2, (const TV *)V, vrs, vcs, (const TW *)W, wrs, wcs, rows, part);
         OFRR_LAUNCH(ctx);
         return kSuccess;
      });
   });
   if (st) return st;
   gram_reduce_kernel<<<(unsigned)((per_chunk + 255) / 256), 256, 0, ctx->stream>>>(
       nchunks, ti, tj, k1, k2, part, alpha, beta, G, ldg);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---------------------------------------------------------------------------
// Y(rows, 0:k) = sum_b V_b(rows, :) C_b, 64 rows x 64 columns per workgroup
struct UpdBlock {
   const void *V;
   const double *C;
   int64_t rs, cs;
   int kb, ldc, dtype;
};
struct UpdArgs {
   UpdBlock b[4];
   int nb;
};

template <typename T>
__device__ inline double ld_any(const void *p, int64_t off) {
   return (double)Elem<T>::load(((const T *)p)[off]);
}
__device__ inline double load_dt(const void *p, int64_t off, int dt) {
   switch (dt) {
      case OFRR_F64: return ld_any<double>(p, off);
      case OFRR_F32: return ld_any<float>(p, off);
      case OFRR_F16: return ld_any<__half>(p, off);
      case OFRR_BF16: return ld_any<__hip_bfloat16>(p, off);
      default: return ld_any<fp8_t>(p, off);
   }
}

template <typename TY>
__global__ __launch_bounds__(256) void basis_update_kernel(int64_t n, int k, UpdArgs args,
                                                           TY *__restrict__ Y, int64_t yrs,
                                                           int64_t ycs) {
   __shared__ double sV[GT][GR + 1];  // 64 rows x 32 inner
   __shared__ double sC[GR][GT + 1];  // 32 inner x 64 cols
   const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
   const int64_t r0 = (int64_t)blockIdx.x * GT;
   const int c0 = blockIdx.y * GT;
   double acc[4][4];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
   for (int bi = 0; bi < args.nb; ++bi) {
      const UpdBlock B = args.b[bi];
      const bool vrow = (B.cs == 1);
      for (int j0 = 0; j0 < B.kb; j0 += GR) {
         for (int t = threadIdx.x; t < GT * GR; t += 256) {
            int rr, jj;
            if (vrow) { rr = t / GR; jj = t % GR; } else { jj = t / GT; rr = t % GT; }
            int64_t row = r0 + rr;
            int j = j0 + jj;
            sV[rr][jj] = (row < n && j < B.kb) ? load_dt(B.V, row * B.rs + (int64_t)j * B.cs, B.dtype) : 0.0;
            int cj = t / GT, cc = t % GT;
            int jr = j0 + cj, col = c0 + cc;
            sC[cj][cc] = (jr < B.kb && col < k) ? B.C[jr + (int64_t)col * B.ldc] : 0.0;
         }
         __syncthreads();
#pragma unroll 4
         for (int j = 0; j < GR; ++j) {
            double va[4], cb[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) va[a] = sV[ty * 4 + a][j];
#pragma unroll
            for (int b = 0; b < 4; ++b) cb[b] = sC[j][tx * 4 + b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
               for (int b = 0; b < 4; ++b) acc[a][b] = fma(va[a], cb[b], acc[a][b]);
         }
         __syncthreads();
      }
   }
#pragma unroll
   for (int a = 0; a < 4; ++a) {
      int64_t row = r0 + ty * 4 + a;
      if (row >= n) continue;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
         int col = c0 + tx * 4 + b;
         if (col < k) Y[row * yrs + (int64_t)col * ycs] = Elem<TY>::store(acc[a][b]);
      }
   }
}

int basis_update(ofrr_ctx *ctx, int64_t n, int k, int nb, const void *const *V, const int *vdt,
                 const int *kb, const int64_t *vrs, const int64_t *vcs, const double *const *C,
                 const int *ldc, void *Y, int ydt, int64_t yrs, int64_t ycs) {
   if (nb < 1 || nb > 4) return set_error(ctx, kErrorInvalidArgument, "basis_update: 1 to 4 blocks");
   UpdArgs a;
   a.nb = nb;
   for (int i = 0; i < nb; ++i) {
      if (!dtype_valid(vdt[i])) return set_error(ctx, kErrorBLASUnsupportedType, "basis_update: dtype");
      a.b[i] = UpdBlock{V[i], C[i], vrs[i], vcs[i], kb[i], ldc[i], vdt[i]};
   }
   dim3 grid((unsigned)((n + GT - 1) / GT), (unsigned)((k + GT - 1) / GT));
   if (ydt == OFRR_F64) {
      basis_update_kernel<double><<<grid, 256, 0, ctx->stream>>>(n, k, a, (double *)Y, yrs, ycs);
   } else if (ydt == OFRR_F32) {
      basis_update_kernel<float><<<grid, 256, 0, ctx->stream>>>(n, k, a, (float *)Y, yrs, ycs);
   } else {
      return set_error(ctx, kErrorBLASUnsupportedType, "basis_update: output must be FP64 or FP32");
   }
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
              int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha,
              double beta, double *G, int ldg) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || k1 < 0 || k2 < 0 || ldg < k1) return set_error(ctx, kErrorInvalidArgument, "gram: bad sizes");
   if (k1 == 0 || k2 == 0) return kSuccess;
   if (!V || !W || !G) return set_error(ctx, kErrorInvalidArgument, "gram: null pointer");
   return gram(ctx, n, k1, k2, V, vdt, vrs, vcs, W, wdt, wrs, wcs, alpha, beta, G, ldg);
}

int ofrr_basis_update(ofrr_ctx *ctx, int64_t n, int k, int nb, const void *const *V,
                      const int *vdt, const int *kb, const int64_t *vrs, const int64_t *vcs,
                      const double *const *C, const int *ldc, void *Y, int ydt, int64_t yrs,
                      int64_t ycs) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || k < 0 || !V || !vdt || !kb || !vrs || !vcs || !C || !ldc || !Y)
      return set_error(ctx, kErrorInvalidArgument, "basis_update: bad arguments");
   if (n == 0 || k == 0) return kSuccess;
   return basis_update(ctx, n, k, nb, V, vdt, kb, vrs, vcs, C, ldc, Y, ydt, yrs, ycs);
}

}  // extern "C"
