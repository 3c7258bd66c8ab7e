// Projection products and back-transform.
//
//   gram          G = alpha V^T W + beta G: the projections V^T V and V^T (A V)
//                 of eig_rr.m:49,51 (gemm with transA = true, mvops.cu:423-605).
//                 Tall-skinny: one wave per (contiguous range of rows, 64 x 64
//                 output tile) streams its rows from HBM straight into FP64
//                 MFMA accumulators (v_mfma_f64_16x16x4, operands converted
//                 to FP64 in registers).  The per-range partial tiles are
//                 summed in a fixed order by a second kernel, so the result is
//                 bitwise reproducible and independent of other work on the
//                 device.
//   basis_update  Y = sum_b V_b C_b: the back-transform V <- V U (eig_rr.m:56)
//                 for a basis whose column blocks are stored in different
//                 precisions (FP64 Ritz block, low-precision growth blocks).
//   small_gemm    C = alpha op(A) B + beta C for the k x k Rayleigh-Ritz algebra.
#include "internal.hpp"

namespace ofrr {

constexpr int GT = 64;   // output tile
constexpr int GR = 32;   // rows staged per slab (back-transform)
constexpr int kGramChunks = 2048;  // row chunks of a one-tile product
constexpr int kRedGroups = 8;      // first-level groups of the chunk sum

typedef double dbl4 __attribute__((ext_vector_type(4)));

// One wave per (row chunk, 64 x 64 output tile).  v_mfma_f64_16x16x4 takes
// A = V^T (16 columns x 4 rows) and B = W (4 rows x 16 columns): lane l holds
// V(r + l/16, i + l%16) and W(r + l/16, j + l%16), so a 16-lane group reads 16
// consecutive entries of one row.  Each of the 16 accumulator tiles keeps the
// rows of the chunk in order.  Entry (i, j) of a 16 x 16 result sits in lane
// j + 16 (i % 4), register i / 4 (layout measured on gfx950 with a one-hot probe).
template <typename TV, typename TW>
__global__ __launch_bounds__(64) void gram_mfma_kernel(
    int64_t n, int k1, int k2, const TV *__restrict__ V, int64_t vrs, int64_t vcs,
    const TW *__restrict__ W, int64_t wrs, int64_t wcs, int64_t rows_per_chunk,
    double *__restrict__ part) {
   const int lane = threadIdx.x;
   const int li = lane & 15, lk = lane >> 4;
   const int i0 = blockIdx.y * GT, j0 = blockIdx.z * GT;
   const int64_t r0 = (int64_t)blockIdx.x * rows_per_chunk;
   const int64_t r1 = min(n, r0 + rows_per_chunk);
   dbl4 acc[4][4];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
   bool cv[4], cw[4];
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      cv[t] = i0 + 16 * t + li < k1;
      cw[t] = j0 + 16 * t + li < k2;
   }
   const TV *vb = V + (int64_t)(i0 + li) * vcs;
   const TW *wb = W + (int64_t)(j0 + li) * wcs;
#pragma unroll 2
   for (int64_t r = r0; r < r1; r += 4) {
      const int64_t row = r + lk;
      const bool rok = row < r1;
      double a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
         a[t] = (rok && cv[t]) ? (double)Elem<TV>::load(vb[row * vrs + (int64_t)(16 * t) * vcs]) : 0.0;
         b[t] = (rok && cw[t]) ? (double)Elem<TW>::load(wb[row * wrs + (int64_t)(16 * t) * wcs]) : 0.0;
      }
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
         for (int tj = 0; tj < 4; ++tj)
            acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
   }
   const int64_t tile = (int64_t)blockIdx.y * gridDim.z + blockIdx.z;
   double *p = part + ((int64_t)blockIdx.x * gridDim.y * gridDim.z + tile) * (GT * GT);
#pragma unroll
   for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int tj = 0; tj < 4; ++tj)
#pragma unroll
         for (int rg = 0; rg < 4; ++rg)
            p[(16 * ti + lk + 4 * rg) * GT + 16 * tj + li] = acc[ti][tj][rg];
}

// G(i, j) = alpha * sum over chunks of the partial tiles: the chunks are summed in
// kRedGroups consecutive groups (each in chunk order), then the group sums in order
__global__ __launch_bounds__(256) void gram_reduce_kernel(int nchunks, int ntiles_i, int ntiles_j,
                                                          int k1, int k2,
                                                          const double *__restrict__ part,
                                                          double alpha, double beta,
                                                          double *__restrict__ G, int ldg) {
   __shared__ double gs[256];
   const int64_t per_chunk = (int64_t)ntiles_i * ntiles_j * GT * GT;
   const int el = threadIdx.x % 32, grp = threadIdx.x / 32;  // 32 entries x 8 groups
   const int64_t e = (int64_t)blockIdx.x * 32 + el;
   const int gsz = (nchunks + kRedGroups - 1) / kRedGroups;
   double s = 0.0;
   if (e < per_chunk) {
      const int c0 = grp * gsz, c1 = min(nchunks, c0 + gsz);
      for (int c = c0; c < c1; ++c) s += part[c * per_chunk + e];
   }
   gs[threadIdx.x] = s;
   __syncthreads();
   if (grp != 0 || e >= per_chunk) return;
   double t = 0.0;
   for (int g = 0; g < kRedGroups; ++g) t += gs[g * 32 + el];
   const int64_t tile = e / (GT * GT);
   const int w = (int)(e % (GT * GT));
   const int i = (int)(tile / ntiles_j) * GT + w / GT;
   const int j = (int)(tile % ntiles_j) * GT + w % GT;
   if (i >= k1 || j >= k2) return;
   double *g = G + i + (int64_t)j * ldg;
   *g = (beta == 0.0) ? alpha * t : fma(alpha, t, beta * *g);
}

int gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
         int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha, double beta,
         double *G, int ldg) {
   if (k1 <= 0 || k2 <= 0) return kSuccess;
   const int ti = (k1 + GT - 1) / GT, tj = (k2 + GT - 1) / GT;
   // about kGramChunks waves per tile pair, each with a multiple of 4 rows
   int64_t want = std::max<int64_t>(1, kGramChunks / (ti * tj));
   int64_t rows = (n + want - 1) / want;
   rows = std::max<int64_t>(64, ((rows + 3) / 4) * 4);
   const int nchunks = (int)((n + rows - 1) / rows);
   const int64_t per_chunk = (int64_t)ti * tj * GT * GT;
   double *part = (double *)workspace(ctx, (size_t)nchunks * per_chunk * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "gram: workspace");
   int st = dispatch_dtype(vdt, [&](auto tv) -> int {
      using TV = decltype(tv);
      return dispatch_dtype(wdt, [&](auto tw) -> int {
         using TW = decltype(tw);
         gram_mfma_kernel<TV, TW><<<dim3(nchunks, ti, tj), 64, 0, ctx->stream>>>(
             n, k1, k2, (const TV *)V, vrs, vcs, (const TW *)W, wrs, wcs, rows, part);
         OFRR_LAUNCH(ctx);
         return kSuccess;
      });
   });
   if (st) return st;
   gram_reduce_kernel<<<(unsigned)((per_chunk + 31) / 32), 256, 0, ctx->stream>>>(
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
