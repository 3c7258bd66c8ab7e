// Reductions: deterministic partial sums, pivot searches (iamax with a 1-based
// index and the signed pivot value, mvops.cu:607-946; per-column argmax with
// 0-based indices, mvops.cu:955-1199), column scaling by the signed max-abs
// entry (MATLAB/+src/+mvops/scalecols.m:21-34) and the NaN/Inf scan
// (mvops.cu:1388-1599).  Ties keep the lowest index, as the reference's
// strict '>' comparison does (mvops.cu:614-621).  The reference reduces in
// two levels with a cudaMalloc/cudaFree per level (mvops.cu:824-857); here a
// fixed two-kernel split over a preallocated workspace is used, with wave64
// shuffles inside a workgroup.
#include "internal.hpp"

namespace ofrr {

// ---- deterministic partial sums -------------------------------------------
__global__ __launch_bounds__(256) void reduce_partials_kernel(int64_t nblk, int width,
                                                              const double *__restrict__ part,
                                                              double *__restrict__ out) {
   __shared__ double sh[256];
   const int j = blockIdx.x;
   double s = 0.0;
   for (int64_t b = threadIdx.x; b < nblk; b += 256) s += part[b * width + j];
   sh[threadIdx.x] = s;
   __syncthreads();
   for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
   }
   if (threadIdx.x == 0) out[j] = sh[0];
}

int reduce_partials(ofrr_ctx *ctx, int64_t nblk, int width, const double *part, double *out) {
   reduce_partials_kernel<<<width, 256, 0, ctx->stream>>>(nblk, width, part, out);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---- argmax |x| ------------------------------------------------------------
struct AMax {
   double a;   // |x|
   int64_t i;  // index (0-based), INT64_MAX when empty
   double v;   // signed value
};

__device__ inline AMax amax_better(const AMax &p, const AMax &q) {
   // larger |x| wins; equal |x| -> lower index
   if (q.a > p.a || (q.a == p.a && q.i < p.i)) return q;
   return p;
}

__device__ inline AMax amax_wave(AMax m) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      AMax q;
      q.a = __shfl_xor(m.a, o, 64);
      q.i = __shfl_xor(m.i, o, 64);
      q.v = __shfl_xor(m.v, o, 64);
      m = amax_better(m, q);
   }
   return m;
}

__device__ inline AMax amax_block(AMax m) {
   __shared__ double sa[16], sv[16];
   __shared__ int64_t si[16];
   m = amax_wave(m);
   const int w = threadIdx.x / 64, l = threadIdx.x % 64;
   if (l == 0) { sa[w] = m.a; si[w] = m.i; sv[w] = m.v; }
   __syncthreads();
   const int nw = (blockDim.x + 63) / 64;
   if (w == 0) {
      AMax q{-1.0, INT64_MAX, 0.0};
      if (l < nw) { q.a = sa[l]; q.i = si[l]; q.v = sv[l]; }
      m = amax_wave(q);
   }
   return m;  // valid in wave 0
}

// stage 1: partial argmax per (chunk, column)
template <typename T>
__global__ __launch_bounds__(256) void amax_partial_kernel(int64_t m, int ncols, const T *__restrict__ A,
                                                           int64_t rs, int64_t cs, int64_t chunk,
                                                           AMax *__restrict__ part) {
   const int col = blockIdx.y;
   const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(m, r0 + chunk);
   const T *a = A + (int64_t)col * cs;
   AMax best{-1.0, INT64_MAX, 0.0};
   for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
      double v = (double)Elem<T>::load(a[r * rs]);
      AMax q{fabs(v), r, v};
      best = amax_better(best, q);
   }
   best = amax_block(best);
   if (threadIdx.x == 0) part[(int64_t)col * gridDim.x + blockIdx.x] = best;
}

// stage 2: reduce the chunks of each column; writes 1-based or 0-based index
// and the signed value; optionally divides the column by its pivot value
template <typename T>
__global__ __launch_bounds__(256) void amax_final_kernel(int nchunks, const AMax *__restrict__ part,
                                                         int base, int32_t *__restrict__ idx,
                                                         double *__restrict__ val) {
   const int col = blockIdx.x;
   AMax best{-1.0, INT64_MAX, 0.0};
   for (int c = threadIdx.x; c < nchunks; c += blockDim.x)
      best = amax_better(best, part[(int64_t)col * nchunks + c]);
   best = amax_block(best);
   if (threadIdx.x == 0) {
      int64_t i = best.i == INT64_MAX ? 0 : best.i;
      if (idx) idx[col] = (int32_t)(i + base);
      if (val) val[col] = best.i == INT64_MAX ? 0.0 : best.v;
   }
}

template <typename T>
__global__ void divcols_kernel(int64_t m, T *__restrict__ A, int64_t rs, int64_t cs,
                               const double *__restrict__ piv) {
   const int col = blockIdx.y;
   const double t = piv[col];
   if (t == 0.0) return;
   for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m;
        r += (int64_t)gridDim.x * blockDim.x) {
      T *p = A + r * rs + (int64_t)col * cs;
      *p = Elem<T>::store((double)Elem<T>::load(*p) / t);
   }
}

int argmax_columns(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A, int64_t rs,
                        int64_t cs, int base, int32_t *idx, double *val) {
   int64_t chunk = 8192;
   int nchunks = (int)((m + chunk - 1) / chunk);
   if (nchunks < 1) nchunks = 1;
   AMax *part = (AMax *)workspace(ctx, (size_t)nchunks * n * sizeof(AMax));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "argmax workspace");
   return dispatch_dtype(dtype, [&](auto t) -> int {
      using T = decltype(t);
      amax_partial_kernel<T><<<dim3(nchunks, n), 256, 0, ctx->stream>>>(m, n, (const T *)A, rs, cs,
                                                                       chunk, part);
      amax_final_kernel<T><<<n, 256, 0, ctx->stream>>>(nchunks, part, base, idx, val);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

int iamax(ofrr_ctx *ctx, int dtype, int64_t n, const void *x, int64_t inc, int32_t *idx1,
          double *val) {
   return argmax_columns(ctx, dtype, n, 1, x, inc, 0, 1, idx1, val);
}

// ---- NaN / Inf scan ---------------------------------------------------------
template <typename T>
__global__ void special_kernel(int64_t m, int n, const T *__restrict__ A, int64_t rs, int64_t cs,
                               int *__restrict__ flag) {
   int local = 0;
   const int64_t total = m * (int64_t)n;
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
        t += (int64_t)gridDim.x * blockDim.x) {
      int64_t r = t % m, c = t / m;
      double v = (double)Elem<T>::load(A[r * rs + c * cs]);
      if (v != v) local |= 1;
      else if (isinf(v)) local |= 2;
   }
   if (local) atomicOr(flag, local);
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_iamax(ofrr_ctx *ctx, int dtype, int64_t n, const void *x, int64_t inc, int32_t *idx1,
               double *val) {
   OFRR_CHECK_CTX(ctx);
   if (n <= 0 || !x || inc < 1) return set_error(ctx, kErrorInvalidArgument, "iamax: bad arguments");
   if (n >= ((int64_t)1 << 31) - 1) return set_error(ctx, kErrorBLASInvalidSize, "iamax: n >= 2^31 - 1");
   return iamax(ctx, dtype, n, x, inc, idx1, val);
}

int ofrr_matrix_iamax(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A, int64_t rs,
                      int64_t cs, int32_t *idx0) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n < 0 || !A || !idx0) return set_error(ctx, kErrorInvalidArgument, "matrix_iamax: bad arguments");
   if (n == 0) return kSuccess;
   return argmax_columns(ctx, dtype, m, n, A, rs, cs, 0, idx0, nullptr);
}

int ofrr_scalecols(ofrr_ctx *ctx, int dtype, int64_t m, int n, void *A, int64_t rs, int64_t cs,
                   double *piv) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n < 0 || !A) return set_error(ctx, kErrorInvalidArgument, "scalecols: bad arguments");
   if (n == 0) return kSuccess;
   double *pv = piv;
   if (!pv) {
      // the pivot values live past the argmax partials in the workspace
      int64_t chunk = 8192;
      int nchunks = (int)((m + chunk - 1) / chunk);
      size_t off = ((size_t)nchunks * n * sizeof(AMax) + 255) & ~(size_t)255;
      char *ws = (char *)workspace(ctx, off + n * sizeof(double));
      if (!ws) return set_error(ctx, kErrorCUDAMemory, "scalecols workspace");
      pv = (double *)(ws + off);
   }
   int st = argmax_columns(ctx, dtype, m, n, A, rs, cs, 0, nullptr, pv);
   if (st) return st;
   int gx = (int)std::min<int64_t>((m + 255) / 256, 1024);
   return dispatch_dtype(dtype, [&](auto t) -> int {
      using T = decltype(t);
      divcols_kernel<T><<<dim3(gx, n), 256, 0, ctx->stream>>>(m, (T *)A, rs, cs, pv);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

int ofrr_check_special(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A, int64_t rs,
                       int64_t cs) {
   OFRR_CHECK_CTX(ctx);
   if (m < 0 || n < 0 || (m * n > 0 && !A)) return set_error(ctx, kErrorInvalidArgument, "check_special: bad arguments");
   if (m == 0 || n == 0) return kSuccess;
   int *flag = (int *)workspace(ctx, sizeof(int));
   int *host = (int *)pinned(ctx, sizeof(int));
   if (!flag || !host) return set_error(ctx, kErrorCUDAMemory, "check_special workspace");
   OFRR_HIP(ctx, hipMemsetAsync(flag, 0, sizeof(int), ctx->stream));
   int64_t total = m * (int64_t)n;
   int grid = (int)std::min<int64_t>((total + 255) / 256, 1024);
   int st = dispatch_dtype(dtype, [&](auto t) -> int {
      using T = decltype(t);
      special_kernel<T><<<grid, 256, 0, ctx->stream>>>(m, n, (const T *)A, rs, cs, flag);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
   if (st) return st;
   OFRR_HIP(ctx, hipMemcpyAsync(host, flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   if (*host & 1) return kErrorNaN;
   if (*host & 2) return kErrorInf;
   return kSuccess;
}

}  // extern "C"
