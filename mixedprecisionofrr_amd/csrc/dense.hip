// Dense tall-skinny products on MFMA: C = alpha P Q + beta C.
//
// P is R x K row-major (element (r, k) at P[r * ldp + k]); Q is K x nc
// row-major (nc <= 128 columns, a multiple of 16); C is R x nc row-major.
// Both products of the dense operators of this library have this form:
//   * A X        with P = A in row-major storage (R = rows of A, K = cols);
//   * A^T Y      with P = A^T row-major = A column-major (R = cols, K = rows);
//   * A X for a symmetric dense A in either storage.
// They are HBM-bound (arithmetic intensity = nc flops per element of P), so P is
// streamed once, straight from HBM into MFMA fragments with 16-byte loads, and
// the small operand Q is staged through LDS once per K chunk and shared by the
// workgroup's four waves (64 rows of P each in the low-precision kernel, whose
// Q is first transposed once per product so that the staging reads 16-byte
// rows; 16 rows each in the FP64 kernel).
//
//   low   P, Q in BF16 or FP16, FP32 accumulation: v_mfma_f32_16x16x32_{bf16,f16}
//         (lane l holds P[l%16][8(l/16)+e], Q[8(l/16)+e][l%16]; result register g
//         of lane l is C[4(l/16)+g][l%16], checked on gfx950 with a probe)
//   f64   P in FP32 or FP64, Q in FP64, FP64 accumulation: v_mfma_f64_16x16x4
//         (lane l holds one P[l%16][k] and Q[k][l%16] with k from lane group
//         l/16; register g of lane l is C[4g + l/16][l%16])
//
// A long K (A^T Y over the rows of a tall A) is split into S ranges whose
// partial products are summed in a fixed order by a second kernel, so the
// result is deterministic.  MATLAB semantics: gemm.m:27-59 (transA variants).
#include "internal.hpp"

namespace ofrr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kDMaxNc = 128;

template <typename T> struct LowMfma;
template <> struct LowMfma<__hip_bfloat16> {
   using vec = bf16x8;
   __device__ static f32x4 mma(vec a, vec b, f32x4 c) {
      return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
   }
};
template <> struct LowMfma<__half> {
   using vec = f16x8;
   __device__ static f32x4 mma(vec a, vec b, f32x4 c) {
      return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
   }
};

// 8 consecutive 16-bit elements of P starting at column k (zero past K)
template <typename T>
__device__ inline typename LowMfma<T>::vec load8(const T *__restrict__ row, int64_t k, int64_t K) {
   typename LowMfma<T>::vec v;
   if (k + 8 <= K) {
      const uint4 u = *reinterpret_cast<const uint4 *>(row + k);
      __builtin_memcpy(&v, &u, 16);
   } else {
      unsigned short s[8];
      for (int e = 0; e < 8; ++e)
         s[e] = (k + e < K) ? *reinterpret_cast<const unsigned short *>(row + k + e) : 0;
      __builtin_memcpy(&v, s, 16);
   }
   return v;
}

// ---------------------------------------------------------------------------
// Qt (nc x KT, row c = column c of Q, zero for k >= K) from Q (K x nc row-major):
// the low-precision kernel reads its B fragments (8 consecutive k of one column)
// as 16-byte rows of Qt.  Q is the small operand (K nc elements), so this costs
// one read and one write of it per product.
template <typename T>
__global__ void q_transpose_kernel(int64_t K, int nc, const T *__restrict__ Q, int64_t ldq, int64_t KT,
                                   T *__restrict__ Qt) {
   __shared__ T tile[64][65];
   const int64_t k0 = (int64_t)blockIdx.x * 64;
   const int c0 = blockIdx.y * 64;
   const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
   for (int r = ty; r < 64; r += 4) {
      const int64_t k = k0 + r;
      const int c = c0 + tx;
      tile[r][tx] = (k < K && c < nc) ? Q[k * ldq + c] : T(0);
   }
   __syncthreads();
   for (int r = ty; r < 64; r += 4) {
      const int c = c0 + r;
      const int64_t k = k0 + tx;
      if (c < nc && k < KT) Qt[(int64_t)c * KT + k] = tile[tx][r];
   }
}

// low precision: out = alpha * P Q over K range [kb, ke) (+ beta out), or the
// raw FP32 partial into part[split] when part != nullptr.  A workgroup owns
// 256 rows of P, each of its four waves 64 (four 16-row MFMA tiles), against
// all nc columns: per 32-long K step a lane loads 16 bytes of each of its four
// P rows straight from HBM (P is read once) and 16 bytes of Qt per column tile
// from the LDS chunk the workgroup staged (64 K x nc), and issues 4 ntile MFMAs.
constexpr int kTallRows = 256;
constexpr int kTallKC = 64;
template <typename T, typename TO>
__global__ __launch_bounds__(256, 2) void dense_low_kernel(
    int64_t R, int64_t K, int nc, const T *__restrict__ P, int64_t ldp, const T *__restrict__ Qt,
    int64_t KT, int64_t kper, double alpha, double beta, TO *__restrict__ C, int64_t ldc,
    float *__restrict__ part) {
   constexpr int QLD = kTallKC + 8;  // 16-byte pad: the 16 columns of a fragment read spread over banks
   __shared__ __align__(16) T qs[kDMaxNc * QLD];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int li = lane & 15, lg = lane >> 4;
   const int64_t r0 = (int64_t)blockIdx.x * kTallRows + 64 * w;
   const int64_t kb = (int64_t)blockIdx.y * kper, ke = min(K, kb + kper);
   const int ntile = nc >> 4;
   f32x4 acc[4][kDMaxNc / 16];
#pragma unroll
   for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int t = 0; t < kDMaxNc / 16; ++t) acc[rt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
   const T *prw[4];
#pragma unroll
   for (int rt = 0; rt < 4; ++rt) prw[rt] = P + min(r0 + 16 * rt + li, R - 1) * ldp;
   for (int64_t kc = kb; kc < ke; kc += kTallKC) {
      __syncthreads();  // the previous chunk is consumed
      // Qt[0 .. nc)[kc .. kc + 64) -> qs, 16 bytes per access (Qt is zero past K)
      for (int e = threadIdx.x; e < nc * (kTallKC / 8); e += 256) {
         const int c = e / (kTallKC / 8), q = e - c * (kTallKC / 8);
         *reinterpret_cast<uint4 *>(&qs[c * QLD + 8 * q]) =
             *reinterpret_cast<const uint4 *>(&Qt[(int64_t)c * KT + kc + 8 * q]);
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < kTallKC / 32; ++s) {
         const int64_t k = kc + 32 * s + 8 * lg;
         typename LowMfma<T>::vec a[4];
#pragma unroll
         for (int rt = 0; rt < 4; ++rt)
            a[rt] = (k < ke) ? load8<T>(prw[rt], k, ke) : typename LowMfma<T>::vec{};
#pragma unroll
         for (int t = 0; t < kDMaxNc / 16; ++t) {
            if (t >= ntile) break;
            typename LowMfma<T>::vec b;
            const uint4 u = *reinterpret_cast<const uint4 *>(&qs[(16 * t + li) * QLD + 32 * s + 8 * lg]);
            __builtin_memcpy(&b, &u, 16);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) acc[rt][t] = LowMfma<T>::mma(a[rt], b, acc[rt][t]);
         }
      }
   }
   // register g of lane l: row 4 lg + g, column 16 t + li (of row tile rt)
#pragma unroll
   for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int t = 0; t < kDMaxNc / 16; ++t) {
         if (t >= ntile) break;
#pragma unroll
         for (int g = 0; g < 4; ++g) {
            const int64_t r = r0 + 16 * rt + 4 * lg + g;
            if (r >= R) continue;
            const int c = 16 * t + li;
            if (part) {
               part[((int64_t)blockIdx.y * R + r) * nc + c] = acc[rt][t][g];
            } else {
               TO *o = C + r * ldc + c;
               double v = alpha * (double)acc[rt][t][g];
               if (beta != 0.0) v = fma(beta, (double)Elem<TO>::load(*o), v);
               *o = Elem<TO>::store(v);
            }
         }
      }
}

// ---------------------------------------------------------------------------
// FP64 accumulation: P in FP32 or FP64, Q in FP64.  A workgroup owns 128 rows
// of P, each wave 32 (two 16-row tiles), against all nc columns; Q is staged
// through LDS in its own row-major layout, 32 K at a time, with 16-byte loads.
// In MFMA step q of a 16-long K block s, lane group g supplies K index
// 16 s + 4 g + q on both operands, so a lane reads 4 consecutive elements of
// each of its P rows with one vector load, and the B operand of a step (16
// consecutive columns of one Q row per lane group) is a conflict-free LDS read.
constexpr int kF64Rows = 128;
constexpr int kF64KC = 32;
template <typename TP> struct P4;
template <> struct P4<float> {
   __device__ static void load(const float *p, double a[4]) {
      const float4 v = *reinterpret_cast<const float4 *>(p);
      a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
   }
};
template <> struct P4<double> {
   __device__ static void load(const double *p, double a[4]) {
      const double2 u = *reinterpret_cast<const double2 *>(p);
      const double2 v = *reinterpret_cast<const double2 *>(p + 2);
      a[0] = u.x; a[1] = u.y; a[2] = v.x; a[3] = v.y;
   }
};
template <typename TP, typename TO>
__global__ __launch_bounds__(256, 2) void dense_f64_kernel(
    int64_t R, int64_t K, int nc, const TP *__restrict__ P, int64_t ldp,
    const double *__restrict__ Q, int64_t ldq, int64_t kper, double alpha, double beta,
    TO *__restrict__ C, int64_t ldc, double *__restrict__ part, int pvec) {
   constexpr int QLD = kDMaxNc + 2;  // 16-byte aligned rows, staggered banks between rows
   __shared__ __align__(16) double qs[kF64KC * QLD];
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   const int li = lane & 15, lg = lane >> 4;
   const int64_t r0 = (int64_t)blockIdx.x * kF64Rows + 32 * w;
   const int64_t kb = (int64_t)blockIdx.y * kper, ke = min(K, kb + kper);
   const int ntile = nc >> 4;
   f64x4 acc[2][kDMaxNc / 16];
#pragma unroll
   for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int t = 0; t < kDMaxNc / 16; ++t) acc[rt][t] = f64x4{0.0, 0.0, 0.0, 0.0};
   const TP *prw[2];
#pragma unroll
   for (int rt = 0; rt < 2; ++rt) prw[rt] = P + min(r0 + 16 * rt + li, R - 1) * ldp;
   const bool qvec = (ldq % 2 == 0) && (((uintptr_t)Q & 15u) == 0);
   // 4 P elements from K index k on (zero past ke), one vector load when aligned
   auto load_p4 = [&](const TP *row, int64_t k, double a[4]) {
      if (pvec && k + 4 <= ke) {
         P4<TP>::load(row + k, a);
      } else {
#pragma unroll
         for (int q = 0; q < 4; ++q) a[q] = (k + q < ke) ? (double)row[k + q] : 0.0;
      }
   };
   for (int64_t kc = kb; kc < ke; kc += kF64KC) {
      __syncthreads();  // the previous chunk is consumed
      // Q[kc .. kc + 32)[0 .. nc) -> qs (zero past ke), two doubles per access
      for (int e = threadIdx.x; e < kF64KC * (nc / 2); e += 256) {
         const int kk = e / (nc / 2), c = 2 * (e - kk * (nc / 2));
         const int64_t k = kc + kk;
         double2 v = make_double2(0.0, 0.0);
         if (k < ke) {
            if (qvec) v = *reinterpret_cast<const double2 *>(&Q[k * ldq + c]);
            else v = make_double2(Q[k * ldq + c], Q[k * ldq + c + 1]);
         }
         *reinterpret_cast<double2 *>(&qs[kk * QLD + c]) = v;
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < kF64KC / 16; ++s) {
         const int64_t k = kc + 16 * s + 4 * lg;
         double a[2][4];
#pragma unroll
         for (int rt = 0; rt < 2; ++rt) load_p4(prw[rt], k, a[rt]);
#pragma unroll
         for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < kDMaxNc / 16; ++t) {
               if (t >= ntile) break;
               const double b = qs[(16 * s + 4 * lg + q) * QLD + 16 * t + li];
#pragma unroll
               for (int rt = 0; rt < 2; ++rt)
                  acc[rt][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt][q], b, acc[rt][t], 0, 0, 0);
            }
      }
   }
   // register g of lane l: row 4 g + lg, column 16 t + li (of row tile rt)
#pragma unroll
   for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int t = 0; t < kDMaxNc / 16; ++t) {
         if (t >= ntile) break;
#pragma unroll
         for (int g = 0; g < 4; ++g) {
            const int64_t r = r0 + 16 * rt + 4 * g + lg;
            if (r >= R) continue;
            const int c = 16 * t + li;
            if (part) {
               part[((int64_t)blockIdx.y * R + r) * nc + c] = acc[rt][t][g];
            } else {
               TO *o = C + r * ldc + c;
               double v = alpha * acc[rt][t][g];
               if (beta != 0.0) v = fma(beta, (double)Elem<TO>::load(*o), v);
               *o = Elem<TO>::store(v);
            }
         }
      }
}

// C = alpha * sum_s part[s] (+ beta C), splits summed in order
template <typename TA, typename TO>
__global__ void dense_split_reduce_kernel(int64_t R, int nc, int S, const TA *__restrict__ part,
                                          double alpha, double beta, TO *__restrict__ C,
                                          int64_t ldc) {
   const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= R * nc) return;
   double s = 0.0;
   for (int i = 0; i < S; ++i) s += (double)part[(int64_t)i * R * nc + e];
   const int64_t r = e / nc;
   const int c = (int)(e - r * nc);
   TO *o = C + r * ldc + c;
   double v = alpha * s;
   if (beta != 0.0) v = fma(beta, (double)Elem<TO>::load(*o), v);
   *o = Elem<TO>::store(v);
}

// number of K ranges: enough workgroups to fill the chip when R is short
static int64_t split_count(int64_t R, int64_t K, int64_t chunk, int64_t rows) {
   const int64_t row_blocks = (R + rows - 1) / rows;
   int64_t S = 1;
   while (row_blocks * S < 2048 && K / (S * 2) >= 4 * chunk) S *= 2;
   return S;
}

int dense_gemm(ofrr_ctx *ctx, int64_t R, int64_t K, int nc, const void *P, int pdt, int64_t ldp,
               const void *Q, int qdt, int64_t ldq, double alpha, double beta, void *C, int cdt,
               int64_t ldc) {
   if (R <= 0 || nc <= 0) return kSuccess;
   if (nc % 16 != 0 || nc > kDMaxNc)
      return set_error(ctx, kErrorInvalidArgument, "dense_gemm: nc must be a multiple of 16 and <= 128");
   if (cdt == OFRR_FP8)
      return set_error(ctx, kErrorBLASUnsupportedType, "dense_gemm: FP8 output is not supported");
   const bool low = (pdt == OFRR_BF16 && qdt == OFRR_BF16) || (pdt == OFRR_F16 && qdt == OFRR_F16);
   const bool f64 = (pdt == OFRR_F32 || pdt == OFRR_F64) && qdt == OFRR_F64;
   if (!low && !f64)
      return set_error(ctx, kErrorBLASUnsupportedType,
                       "dense_gemm: operand types must be (BF16, BF16), (FP16, FP16) or (FP32|FP64, FP64)");
   if (low && (ldp % 8 != 0 || ((uintptr_t)P & 15u) != 0))
      return set_error(ctx, kErrorInvalidArgument, "dense_gemm: 16-bit P needs 16-byte aligned rows");
   const int64_t chunk = low ? kTallKC : kF64KC;
   const int64_t rows = low ? kTallRows : kF64Rows;
   const int64_t S = split_count(R, K, chunk, rows);
   const int64_t kper = ((K + S - 1) / S + chunk - 1) / chunk * chunk;
   const int64_t Sr = (K + kper - 1) / kper;
   const int64_t rb = (R + rows - 1) / rows;
   dim3 grid((unsigned)rb, (unsigned)std::max<int64_t>(1, Sr));
   const size_t esz = low ? 4 : 8;
   // workspace: [Qt (low) | split partials]
   const int64_t KT = (K + kTallKC - 1) / kTallKC * kTallKC;
   const size_t qt_bytes = low ? ((size_t)nc * KT * 2 + 255) / 256 * 256 : 0;
   const size_t part_bytes = Sr > 1 ? (size_t)Sr * R * nc * esz : 0;
   char *ws = nullptr;
   if (qt_bytes + part_bytes) {
      ws = (char *)workspace(ctx, qt_bytes + part_bytes);
      if (!ws) return set_error(ctx, kErrorCUDAMemory, "dense_gemm: workspace");
   }
   void *part = part_bytes ? (void *)(ws + qt_bytes) : nullptr;
   void *Qt = low ? (void *)ws : nullptr;
   if (low) {
      const dim3 tg((unsigned)(KT / 64), (unsigned)((nc + 63) / 64));
      if (pdt == OFRR_BF16)
         q_transpose_kernel<__hip_bfloat16><<<tg, 256, 0, ctx->stream>>>(
             K, nc, (const __hip_bfloat16 *)Q, ldq, KT, (__hip_bfloat16 *)Qt);
      else
         q_transpose_kernel<__half><<<tg, 256, 0, ctx->stream>>>(K, nc, (const __half *)Q, ldq, KT,
                                                                  (__half *)Qt);
      OFRR_LAUNCH(ctx);
   }
   return dispatch_dtype(cdt, [&](auto to) -> int {
      using TO = decltype(to);
      if (low) {
         if (pdt == OFRR_BF16)
            dense_low_kernel<__hip_bfloat16, TO><<<grid, 256, 0, ctx->stream>>>(
                R, K, nc, (const __hip_bfloat16 *)P, ldp, (const __hip_bfloat16 *)Qt, KT, kper,
                alpha, beta, (TO *)C, ldc, (float *)part);
         else
            dense_low_kernel<__half, TO><<<grid, 256, 0, ctx->stream>>>(
                R, K, nc, (const __half *)P, ldp, (const __half *)Qt, KT, kper, alpha, beta,
                (TO *)C, ldc, (float *)part);
      } else {
         // vector loads of 4 P elements when the rows keep 16-byte (FP32) / 8-byte
         // (FP64, two 16-byte halves need 16) alignment
         const size_t pes = pdt == OFRR_F32 ? 4 : 8;
         const int pvec = (((uintptr_t)P % 16) == 0 && ((size_t)ldp * pes) % 16 == 0) ? 1 : 0;
         if (pdt == OFRR_F32)
            dense_f64_kernel<float, TO><<<grid, 256, 0, ctx->stream>>>(
                R, K, nc, (const float *)P, ldp, (const double *)Q, ldq, kper, alpha, beta,
                (TO *)C, ldc, (double *)part, pvec);
         else
            dense_f64_kernel<double, TO><<<grid, 256, 0, ctx->stream>>>(
                R, K, nc, (const double *)P, ldp, (const double *)Q, ldq, kper, alpha, beta,
                (TO *)C, ldc, (double *)part, pvec);
      }
      OFRR_LAUNCH(ctx);
      if (Sr > 1) {
         const int64_t tot = R * nc;
         if (low)
            dense_split_reduce_kernel<float, TO><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
                R, nc, (int)Sr, (const float *)part, alpha, beta, (TO *)C, ldc);
         else
            dense_split_reduce_kernel<double, TO><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
                R, nc, (int)Sr, (const double *)part, alpha, beta, (TO *)C, ldc);
         OFRR_LAUNCH(ctx);
      }
      return kSuccess;
   });
}

// element conversion between formats (dense operator copies), any strides
template <typename TI, typename TO>
__global__ void convert_kernel(int64_t rows, int64_t cols, const TI *__restrict__ X, int64_t xrs,
                               int64_t xcs, TO *__restrict__ Y, int64_t yrs, int64_t ycs) {
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= rows * cols) return;
   // consecutive threads walk the destination's contiguous direction
   int64_t i, j;
   if (ycs == 1) { i = t / cols; j = t - i * cols; } else { j = t / rows; i = t - j * rows; }
   Y[i * yrs + j * ycs] = Elem<TO>::store((double)Elem<TI>::load(X[i * xrs + j * xcs]));
}

int convert(ofrr_ctx *ctx, int64_t rows, int64_t cols, const void *X, int xdt, int64_t xrs,
            int64_t xcs, void *Y, int ydt, int64_t yrs, int64_t ycs) {
   const int64_t tot = rows * cols;
   if (tot <= 0) return kSuccess;
   return dispatch_dtype(xdt, [&](auto ti) -> int {
      using TI = decltype(ti);
      return dispatch_dtype(ydt, [&](auto to) -> int {
         using TO = decltype(to);
         convert_kernel<TI, TO><<<(unsigned)((tot + 255) / 256), 256, 0, ctx->stream>>>(
             rows, cols, (const TI *)X, xrs, xcs, (TO *)Y, yrs, ycs);
         OFRR_LAUNCH(ctx);
         return kSuccess;
      });
   });
}

// ---------------------------------------------------------------------------
// epilogues of the dense operators (the CSR kernels fuse these into the SpMM)

// R = W - X diag(theta): quantised copy (scaled by 1 / sR), column max |R|, and
// per-block FP64 sums of r^2 and x^2 (fixed order).  Thread t owns column
// t % k (k divides 256) and every (256 / k)-th row of the block.
template <typename TB>
__global__ __launch_bounds__(256) void resid_epilogue_kernel(
    int64_t n, int k, const double *__restrict__ X, const double *__restrict__ th,
    const double *__restrict__ W, TB *__restrict__ Rq, const double *__restrict__ inv_sR,
    float *__restrict__ colmax, double *__restrict__ part, int64_t rows_per_block) {
   __shared__ double sr[256], sx[256];
   __shared__ float sm[256];
   const int c = threadIdx.x % k, rstep = 256 / k;
   const int64_t r0 = (int64_t)blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
   const double t = th[c], is = inv_sR[c];
   double a = 0.0, b = 0.0;
   float m = 0.f;
   for (int64_t r = r0 + threadIdx.x / k; r < r1; r += rstep) {
      const double x = X[r * k + c];
      const double rv = fma(-t, x, W[r * k + c]);
      a = fma(rv, rv, a);
      b = fma(x, x, b);
      m = fmaxf(m, (float)fabs(rv));
      Rq[r * k + c] = Elem<TB>::store(rv * is);
   }
   sr[threadIdx.x] = a;
   sx[threadIdx.x] = b;
   sm[threadIdx.x] = m;
   __syncthreads();
   if (threadIdx.x < k) {
      double pa = 0.0, pb = 0.0;
      float pm = 0.f;
      for (int g = 0; g < rstep; ++g) {
         pa += sr[g * k + threadIdx.x];
         pb += sx[g * k + threadIdx.x];
         pm = fmaxf(pm, sm[g * k + threadIdx.x]);
      }
      part[(int64_t)blockIdx.x * 2 * k + 2 * threadIdx.x] = pa;
      part[(int64_t)blockIdx.x * 2 * k + 2 * threadIdx.x + 1] = pb;
      atomic_max_nonneg(colmax + threadIdx.x, pm);
   }
}

int resid_epilogue(ofrr_ctx *ctx, int64_t n, int k, const double *X, const double *th,
                   const double *W, void *Rq, int rdt, const double *inv_sR, float *colmax,
                   double *sums) {
   if (256 % k != 0) return set_error(ctx, kErrorInvalidArgument, "resid_epilogue: k must divide 256");
   const int64_t rows = 1024;
   const int64_t nblk = (n + rows - 1) / rows;
   double *part = (double *)workspace(ctx, (size_t)nblk * 2 * k * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "resid_epilogue: workspace");
   int st = dispatch_dtype(rdt, [&](auto tb) -> int {
      using TB = decltype(tb);
      resid_epilogue_kernel<TB><<<(unsigned)nblk, 256, 0, ctx->stream>>>(n, k, X, th, W, (TB *)Rq,
                                                                        inv_sR, colmax, part, rows);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
   if (st) return st;
   return reduce_partials(ctx, nblk, 2 * k, part, sums);
}

// power-of-two exponent e with 2^e >= 2 m > 2^(e-1)
__device__ inline int ceil_log2_2m_d(float m) {
   int e;
   const float f = frexpf(2.0f * m, &e);
   return (f == 0.5f) ? e - 1 : e;
}

// one Chebyshev step from the operator product T = A Dc (stored values):
// the same coefficients, scale rule and outputs as the CSR step kernel
template <typename TT, typename TB>
__global__ __launch_bounds__(256) void cheb_combine_kernel(
    int64_t n, int k, const TT *__restrict__ T, const TB *__restrict__ Dc,
    const TB *__restrict__ Dp, const TB *__restrict__ R, ChebArgs ca_, TB *__restrict__ Dn,
    int64_t rows_per_block) {
   using ACC = typename std::conditional<std::is_same<TB, double>::value, double, float>::type;
   __shared__ float sm[256];
   const int c = threadIdx.x % k, rstep = 256 / k;
   float mc = 0.f;
   for (int sl = 0; sl < kMaxSlots; ++sl) mc = fmaxf(mc, ca_.maxc[sl * ca_.slot_ld + c]);
   const int se = (mc > 0.f) ? ceil_log2_2m_d(mc) - ca_.target_exp : 0;
   const double alpha = ca_.step ? ca_.step[0] : ca_.alpha;
   const double beta = ca_.step ? ca_.step[1] : ca_.beta;
   const ACC ca = (ACC)ldexp(alpha * ca_.sc[c], -se);
   const ACC cg = (ACC)(Dp ? ldexp(-beta * ca_.sp[c], -se) : 0.0);
   const ACC cd = (ACC)ldexp(ca_.w[c] * ca_.sR[c], -se);
   const ACC cshift = (ACC)(ca_.step ? ca_.step[2] : ca_.c);
   if (blockIdx.x == 0 && threadIdx.x < k) {
      ca_.sn_out[c] = ldexp(1.0, se);
      if (ca_.zero_next)
         for (int sl = 0; sl < kMaxSlots; ++sl) ca_.zero_next[sl * ca_.slot_ld + c] = 0.f;
   }
   const int64_t r0 = (int64_t)blockIdx.x * rows_per_block, r1 = min(n, r0 + rows_per_block);
   float m = 0.f;
   for (int64_t r = r0 + threadIdx.x / k; r < r1; r += rstep) {
      const int64_t e = r * k + c;
      ACC t = ca * fma(-cshift, (ACC)Elem<TB>::load(Dc[e]), (ACC)T[e]);
      if (Dp) t = fma(cg, (ACC)Elem<TB>::load(Dp[e]), t);
      t = fma(cd, (ACC)Elem<TB>::load(R[e]), t);
      m = fmaxf(m, (float)fabs(t));
      Dn[e] = Elem<TB>::store((double)t);
   }
   sm[threadIdx.x] = m;
   __syncthreads();
   if (threadIdx.x < k) {
      float pm = 0.f;
      for (int g = 0; g < rstep; ++g) pm = fmaxf(pm, sm[g * k + threadIdx.x]);
      atomic_max_nonneg(ca_.maxn + (blockIdx.x & (kMaxSlots - 1)) * ca_.slot_ld + threadIdx.x,
                        ldexpf(pm, se));
   }
}

int cheb_combine(ofrr_ctx *ctx, int64_t n, int k, const void *T, int tdt, const void *Dc,
                 const void *Dp, const void *R, int bdt, const ChebArgs &args, void *Dn) {
   if (256 % k != 0) return set_error(ctx, kErrorInvalidArgument, "cheb_combine: k must divide 256");
   const int64_t rows = 1024;
   const int64_t nblk = (n + rows - 1) / rows;
   return dispatch_dtype(bdt, [&](auto tb) -> int {
      using TB = decltype(tb);
      if (tdt == OFRR_F64)
         cheb_combine_kernel<double, TB><<<(unsigned)nblk, 256, 0, ctx->stream>>>(
             n, k, (const double *)T, (const TB *)Dc, (const TB *)Dp, (const TB *)R, args, (TB *)Dn, rows);
      else if (tdt == OFRR_F32)
         cheb_combine_kernel<float, TB><<<(unsigned)nblk, 256, 0, ctx->stream>>>(
             n, k, (const float *)T, (const TB *)Dc, (const TB *)Dp, (const TB *)R, args, (TB *)Dn, rows);
      else
         return set_error(ctx, kErrorBLASUnsupportedType, "cheb_combine: product must be FP32 or FP64");
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

// upper Gershgorin bound max_i (a_ii + sum_{j != i} |a_ij|) of a dense symmetric
// matrix (row i = column i), one block per row
__global__ __launch_bounds__(256) void dense_gershgorin_kernel(int64_t n, const double *__restrict__ A,
                                                               int64_t lda, double *__restrict__ rowb) {
   __shared__ double s[256];
   const int64_t i = blockIdx.x;
   double a = 0.0;
   for (int64_t j = threadIdx.x; j < n; j += 256) {
      const double v = A[j + i * lda];
      a += (j == i) ? v : fabs(v);
   }
   s[threadIdx.x] = a;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
      __syncthreads();
   }
   if (threadIdx.x == 0) rowb[i] = s[0];
}

__global__ void max_reduce_kernel(int64_t n, const double *__restrict__ v, double *__restrict__ out) {
   __shared__ double s[256];
   double m = -INFINITY;
   for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
   s[threadIdx.x] = m;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
      __syncthreads();
   }
   if (threadIdx.x == 0) *out = s[0];
}

int dense_gershgorin(ofrr_ctx *ctx, int64_t n, const double *A, int64_t lda, double *ub_host) {
   double *rowb = (double *)workspace(ctx, (size_t)(n + 1) * sizeof(double));
   if (!rowb) return set_error(ctx, kErrorCUDAMemory, "gershgorin: workspace");
   dense_gershgorin_kernel<<<(unsigned)n, 256, 0, ctx->stream>>>(n, A, lda, rowb);
   OFRR_LAUNCH(ctx);
   max_reduce_kernel<<<1, 256, 0, ctx->stream>>>(n, rowb, rowb + n);
   OFRR_LAUNCH(ctx);
   OFRR_HIP(ctx, hipMemcpyAsync(ub_host, rowb + n, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_dense_gemm(ofrr_ctx *ctx, int64_t R, int64_t K, int nc, const void *P, int pdt,
                    int64_t ldp, const void *Q, int qdt, int64_t ldq, double alpha, double beta,
                    void *C, int cdt, int64_t ldc) {
   OFRR_CHECK_CTX(ctx);
   if (R < 0 || K < 0 || nc < 0) return set_error(ctx, kErrorInvalidArgument, "dense_gemm: negative size");
   if (R == 0 || nc == 0) return kSuccess;
   if (!P || !Q || !C || ldp < K || ldq < nc || ldc < nc)
      return set_error(ctx, kErrorInvalidArgument, "dense_gemm: bad pointers or leading dimensions");
   if (!dtype_valid(pdt) || !dtype_valid(qdt) || !dtype_valid(cdt))
      return set_error(ctx, kErrorBLASUnsupportedType, "dense_gemm: bad dtype");
   return dense_gemm(ctx, R, K, nc, P, pdt, ldp, Q, qdt, ldq, alpha, beta, C, cdt, ldc);
}

int ofrr_convert(ofrr_ctx *ctx, int64_t rows, int64_t cols, const void *X, int xdt, int64_t xrs,
                 int64_t xcs, void *Y, int ydt, int64_t yrs, int64_t ycs) {
   OFRR_CHECK_CTX(ctx);
   if (rows < 0 || cols < 0 || (!X && rows * cols > 0) || (!Y && rows * cols > 0))
      return set_error(ctx, kErrorInvalidArgument, "convert: bad arguments");
   if (!dtype_valid(xdt) || !dtype_valid(ydt)) return set_error(ctx, kErrorBLASUnsupportedType, "convert: bad dtype");
   return convert(ctx, rows, cols, X, xdt, xrs, xcs, Y, ydt, yrs, ycs);
}

}  // extern "C"
