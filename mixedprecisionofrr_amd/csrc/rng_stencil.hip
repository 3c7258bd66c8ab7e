// Counter-based random fills (Philox4x32-10, Salmon et al., SC'11, with a
// Box-Muller transform) and the stencil-operator generators.  The reference
// fills with cuRAND (CPP/containers/vector.cu:188-225); this library uses its
// own published-method generator so that host and device draw the same
// numbers from (seed, stream, index).
#include "common.hpp"

namespace ofrr {

__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
   const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
   const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
   for (int r = 0; r < 10; ++r) {
      uint32_t lo0 = M0 * c[0], hi0 = __umulhi(M0, c[0]);
      uint32_t lo1 = M1 * c[2], hi1 = __umulhi(M1, c[2]);
      uint32_t n0 = hi1 ^ c[1] ^ k0;
      uint32_t n2 = hi0 ^ c[3] ^ k1;
      c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
      k0 += W0; k1 += W1;
   }
}

__device__ inline void philox_block(uint64_t seed, uint32_t stream, uint64_t p, uint32_t w[4]) {
   w[0] = (uint32_t)p; w[1] = (uint32_t)(p >> 32); w[2] = stream; w[3] = 0u;
   philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ inline double u53(uint32_t a, uint32_t b) {
   return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__device__ inline double randu_at(uint64_t seed, uint32_t stream, uint64_t e) {
   uint32_t w[4];
   philox_block(seed, stream, e >> 1, w);
   return (e & 1) ? u53(w[2], w[3]) : u53(w[0], w[1]);
}

template <typename T>
__global__ void randn_kernel(int64_t n, T *x, int64_t inc, uint64_t seed, uint32_t stream,
                             uint64_t offset) {
   int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   uint64_t e = offset + (uint64_t)i;
   uint32_t w[4];
   philox_block(seed, stream, e >> 1, w);
   double u1 = u53(w[0], w[1]), u2 = u53(w[2], w[3]);
   double r = sqrt(-2.0 * log(1.0 - u1));
   double a = 6.283185307179586476925286766559 * u2;
   double v = (e & 1) ? r * sin(a) : r * cos(a);
   x[i * inc] = Elem<T>::store(v);
}

template <typename T>
__global__ void randu_kernel(int64_t n, T *x, int64_t inc, uint64_t seed, uint32_t stream,
                             uint64_t offset) {
   int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   x[i * inc] = Elem<T>::store(randu_at(seed, stream, offset + (uint64_t)i));
}

// face coefficient between cells c-1 and c along direction d, other coordinates
// (o1, o2): draw c + (N+1)(o1 + N o2) of stream d, 10^(2u-1)
__device__ inline double face_coef(uint64_t seed, int d, int64_t N, int64_t c, int64_t o1,
                                   int64_t o2) {
   if (seed == 0) return 1.0;
   double u = randu_at(seed, (uint32_t)d, (uint64_t)(c + (N + 1) * (o1 + N * o2)));
   return exp((2.0 * u - 1.0) * 2.302585092994045684017991454684);
}

// entries stored before row p (closed form over the boundary rows that lack a
// neighbour), natural x-fastest ordering
__device__ inline int64_t stencil_row_start(int dim, int64_t N, int64_t p) {
   int64_t x = p % N, y = (p / N) % N;
   if (dim == 2) {
      // rows q < p missing x-: lines before (y) + (x > 0); missing x+: y
      int64_t miss = y + (x > 0 ? 1 : 0) + y;
      miss += (y > 0) ? N : x;            // y == 0 rows before p
      miss += (y == N - 1) ? x : 0;       // y == N-1 rows before p
      return 5 * p - miss;
   }
   int64_t z = p / (N * N);
   int64_t lines = y + N * z;
   int64_t miss = lines + (x > 0 ? 1 : 0) + lines;       // x- and x+ missing
   miss += z * N + ((y > 0) ? N : x);                      // y == 0
   miss += z * N + ((y == N - 1) ? x : 0);                 // y == N-1
   miss += (z > 0) ? N * N : (y * N + x);                  // z == 0
   miss += (z == N - 1) ? (y * N + x) : 0;                 // z == N-1
   return 7 * p - miss;
}

template <typename TV>
__global__ void stencil_kernel(int dim, int64_t N, uint64_t seed, int64_t n, int64_t *row_ptr,
                               int32_t *col_idx, TV *vals) {
   int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (p > n) return;
   if (p == n) {
      row_ptr[n] = (dim == 2) ? 5 * N * N - 4 * N : 7 * N * N * N - 6 * N * N;
      return;
   }
   int64_t x = p % N, y = (p / N) % N, z = (dim == 3) ? p / (N * N) : 0;
   double kxl = face_coef(seed, 0, N, x, y, z), kxu = face_coef(seed, 0, N, x + 1, y, z);
   double kyl = face_coef(seed, 1, N, y, x, z), kyu = face_coef(seed, 1, N, y + 1, x, z);
   double kzl = 0.0, kzu = 0.0;
   if (dim == 3) {
      kzl = face_coef(seed, 2, N, z, x, y);
      kzu = face_coef(seed, 2, N, z + 1, x, y);
   }
   double diag = kxl + kxu + kyl + kyu + kzl + kzu;
   int64_t pos = stencil_row_start(dim, N, p);
   row_ptr[p] = pos;
   if (dim == 3 && z > 0) { col_idx[pos] = (int32_t)(p - N * N); vals[pos++] = Elem<TV>::store(-kzl); }
   if (y > 0) { col_idx[pos] = (int32_t)(p - N); vals[pos++] = Elem<TV>::store(-kyl); }
   if (x > 0) { col_idx[pos] = (int32_t)(p - 1); vals[pos++] = Elem<TV>::store(-kxl); }
   col_idx[pos] = (int32_t)p; vals[pos++] = Elem<TV>::store(diag);
   if (x < N - 1) { col_idx[pos] = (int32_t)(p + 1); vals[pos++] = Elem<TV>::store(-kxu); }
   if (y < N - 1) { col_idx[pos] = (int32_t)(p + N); vals[pos++] = Elem<TV>::store(-kyu); }
   if (dim == 3 && z < N - 1) { col_idx[pos] = (int32_t)(p + N * N); vals[pos++] = Elem<TV>::store(-kzu); }
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_fill_randn(ofrr_ctx *ctx, int dtype, int64_t n, void *x, int64_t inc, uint64_t seed,
                    uint32_t stream, uint64_t offset) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || (n > 0 && !x) || inc < 1) return set_error(ctx, kErrorInvalidArgument, "fill_randn: bad arguments");
   if (n == 0) return kSuccess;
   int64_t grid = (n + 255) / 256;
   return dispatch_dtype(dtype, [&](auto t) -> int {
      using T = decltype(t);
      randn_kernel<T><<<grid, 256, 0, ctx->stream>>>(n, (T *)x, inc, seed, stream, offset);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

int ofrr_fill_randu(ofrr_ctx *ctx, int dtype, int64_t n, void *x, int64_t inc, uint64_t seed,
                    uint32_t stream, uint64_t offset) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || (n > 0 && !x) || inc < 1) return set_error(ctx, kErrorInvalidArgument, "fill_randu: bad arguments");
   if (n == 0) return kSuccess;
   int64_t grid = (n + 255) / 256;
   return dispatch_dtype(dtype, [&](auto t) -> int {
      using T = decltype(t);
      randu_kernel<T><<<grid, 256, 0, ctx->stream>>>(n, (T *)x, inc, seed, stream, offset);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

int ofrr_stencil_nnz(int dim, int64_t N, int64_t *nnz) {
   if (!nnz || (dim != 2 && dim != 3) || N < 1) return kErrorInvalidArgument;
   *nnz = (dim == 2) ? 5 * N * N - 4 * N : 7 * N * N * N - 6 * N * N;
   return kSuccess;
}

int ofrr_stencil_csr(ofrr_ctx *ctx, int dim, int64_t N, uint64_t seed, int64_t *row_ptr,
                     int32_t *col_idx, int val_dtype, void *vals) {
   OFRR_CHECK_CTX(ctx);
   if ((dim != 2 && dim != 3) || N < 1 || !row_ptr || !col_idx || !vals)
      return set_error(ctx, kErrorInvalidArgument, "stencil_csr: bad arguments");
   int64_t n = (dim == 2) ? N * N : N * N * N;
   if (n >= ((int64_t)1 << 31)) return set_error(ctx, kErrorBLASInvalidSize, "stencil_csr: n >= 2^31");
   int64_t grid = (n + 1 + 255) / 256;
   return dispatch_dtype(val_dtype, [&](auto t) -> int {
      using T = decltype(t);
      stencil_kernel<T><<<grid, 256, 0, ctx->stream>>>(dim, N, seed, n, row_ptr, col_idx, (T *)vals);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
}

}  // extern "C"
