// Rectangular and transposed CSR products (MATLAB/+src/+csr/csrmv.m:27-76).
//
// csrmv takes a CSR A of nrows x ncols and trans 'N' (y = A x, x has ncols
// rows, :48-59) or 'T' (y = A^T x, x has nrows rows, :61-71).  The 'T' branch
// scatters: y starts at zero and, row i by row i in increasing order, each
// stored entry adds a_ij x(i,:) to y(j,:).  So y(j,:) is a sum over the rows i
// holding column j, accumulated in increasing i.
//
// On the GPU a scatter needs atomics, whose order (and so the FP64 rounding)
// varies from run to run.  Instead the transpose is built once as a CSR of
// ncols x nrows whose row j lists the entries of column j in increasing row of
// A: a stable radix sort of the entries by column (rocPRIM, LSD, stable), the
// row of each entry and its value gathered in sorted order, and the row
// pointers read off the sorted keys.  The ordinary row-gather SpMM on that
// transpose then accumulates y(j,:) in exactly the order of the 'T' branch,
// with the same fused multiply-adds, so FP64 results are bitwise those of the
// sequential scatter.
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"

namespace ofrr {

__global__ void iota_kernel(int64_t n, int32_t *__restrict__ p) {
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
        t += (int64_t)gridDim.x * blockDim.x)
      p[t] = (int32_t)t;
}

// row_of[p] = the row holding stored entry p
__global__ void row_of_kernel(int64_t nrows, const int64_t *__restrict__ rp,
                              int32_t *__restrict__ row_of) {
   for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows;
        i += (int64_t)gridDim.x * blockDim.x)
      for (int64_t p = rp[i]; p < rp[i + 1]; ++p) row_of[p] = (int32_t)i;
}

// sorted entry q (original entry perm[q]): column index of A^T = its row in
// A, value copied as raw bytes of width VB
template <int VB>
__global__ void transpose_gather_kernel(int64_t nnz, const int32_t *__restrict__ perm,
                                        const int32_t *__restrict__ row_of,
                                        const unsigned char *__restrict__ va,
                                        int32_t *__restrict__ ciT, unsigned char *__restrict__ vaT) {
   for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
        q += (int64_t)gridDim.x * blockDim.x) {
      const int64_t p = perm[q];
      ciT[q] = row_of[p];
#pragma unroll
      for (int b = 0; b < VB; ++b) vaT[q * VB + b] = va[p * VB + b];
   }
}

// rpT[c] = first sorted position with key >= c, for c in [0, ncols]; a key
// outside [0, ncols) raises *bad and writes nothing past rpT[ncols]
__global__ void transpose_rowptr_kernel(int64_t nnz, int64_t ncols, const uint32_t *__restrict__ keys,
                                        int64_t *__restrict__ rpT, int32_t *__restrict__ bad) {
   for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= nnz;
        q += (int64_t)gridDim.x * blockDim.x) {
      const int64_t prev = (q == 0) ? -1 : (int64_t)keys[q - 1];
      int64_t cur = (q == nnz) ? ncols : (int64_t)keys[q];
      if (q < nnz && cur >= ncols) {
         *bad = 1;
         cur = ncols;
      }
      for (int64_t c = prev + 1; c <= cur && c <= ncols; ++c)
         if (c >= 0) rpT[c] = q;
   }
}

static unsigned grid_for(int64_t n) {
   return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
}

// scratch bytes of csr_transpose (device), besides its outputs
static size_t transpose_scratch(ofrr_ctx *ctx, int64_t nnz, size_t &sort_bytes) {
   sort_bytes = 0;
   if (nnz > 0)
      (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                      (const int32_t *)nullptr, (int32_t *)nullptr, (size_t)nnz, 0, 32,
                                      ctx->stream);
   auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
   return al((size_t)nnz * 4) * 4 + al(sort_bytes) + 256;
}

// A^T of a CSR A (nrows x ncols) as CSR (ncols x nrows), entries of a row of
// A^T in increasing row of A.  `scratch` holds transpose_scratch bytes.
static int csr_transpose(ofrr_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz,
                         const int64_t *rp, const int32_t *ci, const void *va, int vdt,
                         int64_t *rpT, int32_t *ciT, void *vaT, char *scratch, size_t sort_bytes) {
   const size_t vb = dtype_size(vdt);
   auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
   // keys as unsigned: a negative index sorts past every valid one and is reported
   uint32_t *keys = (uint32_t *)scratch;
   int32_t *pin = (int32_t *)(scratch + al((size_t)nnz * 4));
   int32_t *pout = (int32_t *)(scratch + 2 * al((size_t)nnz * 4));
   int32_t *row_of = (int32_t *)(scratch + 3 * al((size_t)nnz * 4));
   char *sort_tmp = scratch + 4 * al((size_t)nnz * 4);
   int32_t *bad = (int32_t *)(sort_tmp + al(sort_bytes));
   int32_t *hbad = (int32_t *)pinned(ctx, sizeof(int32_t));
   if (!hbad) return set_error(ctx, kErrorCUDAMemory, "csr_transpose: pinned");
   // a kernel clears the flag (no memset node under stream capture)
   iota_kernel<<<1, 64, 0, ctx->stream>>>(1, bad);  // bad = 0
   if (nnz > 0) {
      iota_kernel<<<grid_for(nnz), 256, 0, ctx->stream>>>(nnz, pin);
      row_of_kernel<<<grid_for(nrows), 256, 0, ctx->stream>>>(nrows, rp, row_of);
      OFRR_LAUNCH(ctx);
      size_t sb = sort_bytes;
      OFRR_HIP(ctx, rocprim::radix_sort_pairs(sort_tmp, sb, (const uint32_t *)ci, keys, pin, pout,
                                              (size_t)nnz, 0, 32, ctx->stream));
      switch (vb) {
         case 8:
            transpose_gather_kernel<8><<<grid_for(nnz), 256, 0, ctx->stream>>>(
                nnz, pout, row_of, (const unsigned char *)va, ciT, (unsigned char *)vaT);
            break;
         case 4:
            transpose_gather_kernel<4><<<grid_for(nnz), 256, 0, ctx->stream>>>(
                nnz, pout, row_of, (const unsigned char *)va, ciT, (unsigned char *)vaT);
            break;
         default:
            transpose_gather_kernel<2><<<grid_for(nnz), 256, 0, ctx->stream>>>(
                nnz, pout, row_of, (const unsigned char *)va, ciT, (unsigned char *)vaT);
            break;
      }
      OFRR_LAUNCH(ctx);
   }
   transpose_rowptr_kernel<<<grid_for(nnz + 1), 256, 0, ctx->stream>>>(nnz, ncols, keys, rpT, bad);
   OFRR_LAUNCH(ctx);
   OFRR_HIP(ctx, hipMemcpyAsync(hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   if (*hbad) return set_error(ctx, kErrorInvalidArgument, "csr_transpose: a column index lies outside [0, ncols)");
   return kSuccess;
}

static int check_csr_args(ofrr_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rp,
                          const int32_t *ci, const void *va, int vdt, const char *who) {
   if (nrows < 0 || ncols < 0 || nnz < 0)
      return set_error(ctx, kErrorInvalidArgument, std::string(who) + ": negative size");
   if (nrows >= ((int64_t)1 << 31) || ncols >= ((int64_t)1 << 31) || nnz >= ((int64_t)1 << 31))
      return set_error(ctx, kErrorInvalidArgument, std::string(who) + ": sizes must be below 2^31");
   if (!rp || (nnz > 0 && (!ci || !va)))
      return set_error(ctx, kErrorInvalidArgument, std::string(who) + ": null pointer");
   if (vdt != OFRR_F64 && vdt != OFRR_F32 && vdt != OFRR_F16)
      return set_error(ctx, kErrorBLASUnsupportedType, std::string(who) + ": values must be FP64, FP32 or FP16");
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_csr_transpose(ofrr_ctx *ctx, int64_t nrows, int64_t ncols, int64_t nnz, const int64_t *rp,
                       const int32_t *ci, const void *va, int vdt, int64_t *rpT, int32_t *ciT,
                       void *vaT) {
   OFRR_CHECK_CTX(ctx);
   int st = check_csr_args(ctx, nrows, ncols, nnz, rp, ci, va, vdt, "csr_transpose");
   if (st) return st;
   if (!rpT || (nnz > 0 && (!ciT || !vaT)))
      return set_error(ctx, kErrorInvalidArgument, "csr_transpose: null output");
   size_t sort_bytes = 0;
   const size_t need = transpose_scratch(ctx, nnz, sort_bytes);
   char *scratch = (char *)workspace(ctx, need);
   if (!scratch) return set_error(ctx, kErrorCUDAMemory, "csr_transpose: workspace");
   return csr_transpose(ctx, nrows, ncols, nnz, rp, ci, va, vdt, rpT, ciT, vaT, scratch, sort_bytes);
}

int ofrr_spmm_csr_rect(ofrr_ctx *ctx, int trans, int64_t nrows, int64_t ncols, int64_t nnz,
                       const int64_t *rp, const int32_t *ci, const void *va, int vdt, int k,
                       const void *X, int xdt, int64_t xrs, int64_t xcs, double alpha, double beta,
                       void *Y, int ydt, int64_t yrs, int64_t ycs) {
   OFRR_CHECK_CTX(ctx);
   int st = check_csr_args(ctx, nrows, ncols, nnz, rp, ci, va, vdt, "spmm_csr_rect");
   if (st) return st;
   if (trans != 0 && trans != 1) return set_error(ctx, kErrorInvalidArgument, "spmm_csr_rect: trans must be 0 or 1");
   if (k < 0) return set_error(ctx, kErrorInvalidArgument, "spmm_csr_rect: negative k");
   const int64_t ny = trans ? ncols : nrows, nx = trans ? nrows : ncols;
   if (ny == 0 || k == 0) return kSuccess;
   if (!X || !Y) return set_error(ctx, kErrorInvalidArgument, "spmm_csr_rect: null pointer");
   if (!dtype_valid(xdt) || !dtype_valid(ydt))
      return set_error(ctx, kErrorBLASUnsupportedType, "spmm_csr_rect: block dtype");
   if (nx == 0) {
      // op(A) has no columns, so A holds no entries and Y = beta Y: the product
      // over empty rows (no gathers)
      if (nnz > 0) return set_error(ctx, kErrorInvalidArgument, "spmm_csr_rect: entries in a matrix without columns");
      if (trans) return set_error(ctx, kErrorNotImplemented, "spmm_csr_rect: A^T of a matrix with no rows");
      return spmm_csr(ctx, ny, rp, ci, va, vdt, k, X, xdt, xrs, xcs, 0.0, beta, Y, ydt, yrs, ycs, 0, 1);
   }
   if (!trans) return spmm_csr(ctx, nrows, rp, ci, va, vdt, k, X, xdt, xrs, xcs, alpha, beta, Y, ydt, yrs, ycs, 0, ncols);
   // A^T: transpose into the workspace, then the row-gather product over it
   auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
   size_t sort_bytes = 0;
   const size_t scr = transpose_scratch(ctx, nnz, sort_bytes);
   const size_t vb = dtype_size(vdt);
   const size_t tbytes = al((size_t)(ncols + 1) * 8) + al((size_t)nnz * 4) + al((size_t)nnz * vb + 16);
   char *ws = (char *)workspace(ctx, tbytes + scr);
   if (!ws) return set_error(ctx, kErrorCUDAMemory, "spmm_csr_rect: workspace");
   int64_t *rpT = (int64_t *)ws;
   int32_t *ciT = (int32_t *)(ws + al((size_t)(ncols + 1) * 8));
   void *vaT = ws + al((size_t)(ncols + 1) * 8) + al((size_t)nnz * 4);
   if ((st = csr_transpose(ctx, nrows, ncols, nnz, rp, ci, va, vdt, rpT, ciT, vaT, ws + tbytes, sort_bytes)))
      return st;
   return spmm_csr(ctx, ncols, rpT, ciT, vaT, vdt, k, X, xdt, xrs, xcs, alpha, beta, Y, ydt, yrs, ycs, 0, nrows);
}

}  // extern "C"
