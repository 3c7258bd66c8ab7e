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
#include <mutex>
#include <type_traits>

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
   // 16-column groups of the tile inside k1 x k2: a narrow product (a 32- or
   // 16-column block) issues only the products of its own groups
   const int nti = min(4, (k1 - i0 + 15) / 16), ntj = min(4, (k2 - j0 + 15) / 16);
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
            if (ti < nti && tj < ntj)
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

// rows per chunk and chunk count of a product over n rows with ti x tj output
// tiles: about `chunk_tiles` (chunk, tile) waves, each chunk a multiple of 4
// rows; the partials take nchunks x ti x tj tiles of GT x GT doubles, at most
// max(chunk_tiles, ti tj) of them
static void gram_chunks(int64_t n, int ti, int tj, int64_t chunk_tiles, int64_t &rows, int &nchunks) {
   const int64_t want = std::max<int64_t>(1, chunk_tiles / ((int64_t)ti * tj));
   rows = (n + want - 1) / want;
   rows = std::max<int64_t>(64, ((rows + 3) / 4) * 4);
   nchunks = (int)std::max<int64_t>(1, (n + rows - 1) / rows);
}

// the product into `part` (nchunks x ti x tj partial tiles, gram_chunks)
static int gram_run(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
                    int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha,
                    double beta, double *G, int ldg, double *part, int64_t chunk_tiles) {
   const int ti = (k1 + GT - 1) / GT, tj = (k2 + GT - 1) / GT;
   int64_t rows;
   int nchunks;
   gram_chunks(n, ti, tj, chunk_tiles, rows, nchunks);
   const int64_t per_chunk = (int64_t)ti * tj * GT * GT;
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

int gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
         int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha, double beta,
         double *G, int ldg) {
   if (k1 <= 0 || k2 <= 0) return kSuccess;
   const int ti = (k1 + GT - 1) / GT, tj = (k2 + GT - 1) / GT;
   int64_t rows;
   int nchunks;
   gram_chunks(n, ti, tj, kGramChunks, rows, nchunks);
   const int64_t per_chunk = (int64_t)ti * tj * GT * GT;
   double *part = (double *)workspace(ctx, (size_t)nchunks * per_chunk * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "gram: workspace");
   return gram_run(ctx, n, k1, k2, V, vdt, vrs, vcs, W, wdt, wrs, wcs, alpha, beta, G, ldg, part,
                   kGramChunks);
}

// ---- the allocation-free Gram product (ofrr_gram_noalloc) -----------------
// Its chunk partials live in `scratch`, a file-scope device array of
// kScratchChunkTiles tiles of GT x GT doubles (8 MiB on every GPU the library
// runs on, held until the process ends).  Host code never passes the array's
// name to a kernel: ofrr_ctx_create looks up its device address once per
// device (gram_scratch_init, with that device current) and the entry point
// passes that address.  Every chunk writes its whole partial tile before the
// reduction reads it, so nothing needs clearing between calls.
constexpr int64_t kScratchChunkTiles = 256;
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchChunkTiles * GT * GT];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchMaxDevices] = {};

int gram_scratch_init(int device) {
   if (device < 0 || device >= kScratchMaxDevices) return kErrorInvalidArgument;
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (g_scratch_ptr[device]) return kSuccess;
   double *p = nullptr;
   if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return kErrorCUDAGeneral;
   g_scratch_ptr[device] = p;
   return kSuccess;
}

static double *gram_scratch(int device) {
   if (device < 0 || device >= kScratchMaxDevices) return nullptr;
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   return g_scratch_ptr[device];
}

// ---------------------------------------------------------------------------
// Fused projection products of one solver iteration: every block V_a^T V_b of
// the Gram matrix and of the projected matrix in one launch.  The output is cut
// into 64 x 32 tiles, one wave each; a workgroup of 12 waves takes one chunk of
// rows and up to 12 tiles.  It stages the rows of the operands its tiles read
// into LDS as FP64, a slab of R rows at a time (the next slab is loaded into
// registers while the MFMAs run on the current one), so every row
// of every operand is read from HBM once per workgroup and converted once.  The
// operands are row-major n x k blocks whose element kinds are FP64 (Ritz block,
// A X), the growth-block format TB or the product format TZ.  Per-chunk partial
// tiles are summed in a fixed order (deterministic, independent of other work).
constexpr int kGramWaves = 12;                      // tiles (waves) per workgroup
constexpr int kGramThreads = 64 * kGramWaves;
constexpr int kGramStageOps = 6;                    // operands staged per workgroup
// staging loads per thread and operand: slabs of R = 1024 / k rows.  Three
// (R = 2048 / k) need 36 prefetch registers per thread and spill 19-26 at
// 3 waves per SIMD; two measured 30.1 against 31.4 ms of mixed and 27.1
// against 41.5 ms of FP64 projection products per 128^3 solve
constexpr int kGramPasses = 2;
constexpr int kMaxHalfItems = 2 * kMaxGramItems;

// the staged operands op[0 .. N0 + N1 + N2) in kind order: N0 FP64, N1 in the
// basis format TB, N2 in the format TZ; every workgroup stages them all
struct GramLdsArgs {
   GramOperand op[kMaxGramOps];
   int nitems, R, stride, lk;
   short slot_a[kMaxHalfItems], slot_b[kMaxHalfItems], i0[kMaxHalfItems], j0[kMaxHalfItems];
};

// raw bits of one staged element as two 32-bit halves: a load of any width
// writes its registers directly and nothing touches them until the store
// (hi is read for 8-byte formats only)
struct GramRaw {
   uint32_t lo, hi;
};
template <typename T> __device__ inline void gram_raw_load(GramRaw &r, const void *p, int64_t off) {
   if constexpr (sizeof(T) == 8) {
      const uint2 v = reinterpret_cast<const uint2 *>(p)[off];
      r.lo = v.x;
      r.hi = v.y;
   } else if constexpr (sizeof(T) == 4) {
      r.lo = reinterpret_cast<const uint32_t *>(p)[off];
   } else if constexpr (sizeof(T) == 2) {
      r.lo = reinterpret_cast<const uint16_t *>(p)[off];
   } else {
      r.lo = reinterpret_cast<const uint8_t *>(p)[off];
   }
}
template <typename T> __device__ inline double gram_raw_value(const GramRaw &r) {
   if constexpr (sizeof(T) == 8)
      return __builtin_bit_cast(double, ((uint64_t)r.hi << 32) | r.lo);
   else if constexpr (sizeof(T) == 4)
      return (double)Elem<T>::load(__builtin_bit_cast(T, r.lo));
   else if constexpr (sizeof(T) == 2)
      return (double)Elem<T>::load(__builtin_bit_cast(T, (uint16_t)r.lo));
   else
      return (double)Elem<T>::load((T)(uint8_t)r.lo);
}

template <typename TB, typename TZ, int N0, int N1, int N2>
__global__ __launch_bounds__(kGramThreads) void gram_lds_kernel(int64_t n, int k, GramLdsArgs args,
                                                                int64_t rows_per_chunk, int nchunks,
                                                                int ngroups, double *__restrict__ part) {
   constexpr int NOP = N0 + N1 + N2;
   static_assert(NOP <= kGramStageOps, "staged operands");
   extern __shared__ double sbuf[];  // [NOP][R][stride]
   const int chunk = blockIdx.x / ngroups, grp = blockIdx.x % ngroups;
   const int64_t r0 = (int64_t)chunk * rows_per_chunk;
   const int64_t r1 = min(n, r0 + rows_per_chunk);
   const int R = args.R, stride = args.stride;
   const int slab = R * stride;                       // doubles per operand slab
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int li = lane & 15, lk = lane >> 4;
   const int item = grp * kGramWaves + wave;
   const bool has = item < args.nitems;
   const int sa = has ? args.slot_a[item] : 0, sb = has ? args.slot_b[item] : 0;
   const int i0 = has ? args.i0[item] : 0, j0 = has ? args.j0[item] : 0;
   bool cv[4], cw[2];
#pragma unroll
   for (int t = 0; t < 4; ++t) cv[t] = i0 + 16 * t + li < k;
#pragma unroll
   for (int t = 0; t < 2; ++t) cw[t] = j0 + 16 * t + li < k;
   dbl4 acc[4][2];
#pragma unroll
   for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};
   // staging: operand slot o, element e = tid + 768 q of its R x k slab
   // (R k <= 1024, k and R powers of two): row e >> lk, column e & (k - 1).
   // The slots' formats are compile-time, so the prefetch is a run of
   // unconditional loads from clamped (valid) addresses with nothing done to
   // their registers until store() converts them after the MFMAs; store()
   // zeroes the rows past the end of the chunk.
   const int lk2 = args.lk, rk = R << lk2;
   auto fetch = [&](int64_t rb, GramRaw pre[kGramPasses * NOP]) {
#pragma unroll
      for (int o = 0; o < NOP; ++o) {
         const GramOperand &op = args.op[o];
#pragma unroll
         for (int q = 0; q < kGramPasses; ++q) {
            const int e = min((int)threadIdx.x + q * kGramThreads, rk - 1);
            const int64_t row = min(rb + (e >> lk2), r1 - 1);
            const int64_t off = row * op.rs + (e & (k - 1));
            GramRaw &r = pre[kGramPasses * o + q];
            if (o < N0)
               gram_raw_load<double>(r, op.p, off);
            else if (o < N0 + N1)
               gram_raw_load<TB>(r, op.p, off);
            else
               gram_raw_load<TZ>(r, op.p, off);
         }
      }
   };
   auto store = [&](int64_t rb, const GramRaw pre[kGramPasses * NOP]) {
#pragma unroll
      for (int o = 0; o < NOP; ++o) {
#pragma unroll
         for (int q = 0; q < kGramPasses; ++q) {
            const int e = threadIdx.x + q * kGramThreads;
            const GramRaw &r = pre[kGramPasses * o + q];
            double v = (o < N0) ? gram_raw_value<double>(r)
                       : (o < N0 + N1) ? gram_raw_value<TB>(r) : gram_raw_value<TZ>(r);
            if (rb + (e >> lk2) >= r1) v = 0.0;
            if (e < rk) sbuf[o * slab + (e >> lk2) * stride + (e & (k - 1))] = v;
         }
      }
   };
   GramRaw pre[kGramPasses * NOP];
   fetch(r0, pre);
   store(r0, pre);
   __syncthreads();
   for (int64_t rb = r0; rb < r1; rb += R) {
      const bool more = rb + R < r1;
      if (more) fetch(rb + R, pre);  // in flight during the MFMAs below
      if (has) {
         const double *A = sbuf + sa * slab;
         const double *B = sbuf + sb * slab;
         for (int t = 0; t < R; t += 4) {
            const int rr = t + lk;
            double a[4], b[2];
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = cv[q] ? A[rr * stride + i0 + 16 * q + li] : 0.0;
#pragma unroll
            for (int q = 0; q < 2; ++q) b[q] = cw[q] ? B[rr * stride + j0 + 16 * q + li] : 0.0;
#pragma unroll
            for (int ti = 0; ti < 4; ++ti)
#pragma unroll
               for (int tj = 0; tj < 2; ++tj)
                  acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b[tj], acc[ti][tj], 0, 0, 0);
         }
      }
      __syncthreads();  // every wave is done with this slab
      if (more) store(rb + R, pre);
      __syncthreads();
   }
   if (!has) return;
   // entry (i, j) of a 16 x 16 result sits in lane j + 16 (i % 4), register i / 4
   double *p = part + ((int64_t)chunk * args.nitems + item) * (GT * GT / 2);
#pragma unroll
   for (int ti = 0; ti < 4; ++ti)
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
         for (int rg = 0; rg < 4; ++rg)
            p[(16 * ti + lk + 4 * rg) * (GT / 2) + 16 * tj + li] = acc[ti][tj][rg];
}

struct GramOutArgs {
   double *dst[kMaxHalfItems];
   int ld[kMaxHalfItems], ro[kMaxHalfItems], co[kMaxHalfItems];
   short i0[kMaxHalfItems], j0[kMaxHalfItems];
   int nitems;
};

// item's 64 x 32 tile = sum over chunks in a fixed order (kRedGroups groups of
// consecutive chunks, each in chunk order, then the group sums in order)
__global__ __launch_bounds__(256) void gram_lds_reduce_kernel(int nchunks, int k, GramOutArgs o,
                                                              const double *__restrict__ part) {
   __shared__ double gs[256];
   const int item = blockIdx.y;
   const int el = threadIdx.x % 32, grp = threadIdx.x / 32;
   const int e = blockIdx.x * 32 + el;  // entry of the 64 x 32 tile
   const int64_t stride = (int64_t)o.nitems * (GT * GT / 2);
   const int gsz = (nchunks + kRedGroups - 1) / kRedGroups;
   double s = 0.0;
   {
      const int c0 = grp * gsz, c1 = min(nchunks, c0 + gsz);
      const double *pp = part + (int64_t)item * (GT * GT / 2) + e;
      for (int c = c0; c < c1; ++c) s += pp[c * stride];
   }
   gs[threadIdx.x] = s;
   __syncthreads();
   if (grp != 0) return;
   double t = 0.0;
   for (int g = 0; g < kRedGroups; ++g) t += gs[g * 32 + el];
   const int i = o.i0[item] + e / (GT / 2), j = o.j0[item] + e % (GT / 2);
   if (i >= k || j >= k) return;
   o.dst[item][(o.ro[item] + i) + (int64_t)(o.co[item] + j) * o.ld[item]] = t;
}

// Largest workspace (bytes) gram or gram_multi asks for at any n: gram's
// partials are at most max(tiles, kGramChunks) chunk tiles of GT x GT doubles;
// gram_multi's at most max(groups, 2 num_cus) chunks of kGramWaves half tiles
size_t gram_workspace_bound(int num_cus) {
   const size_t g1 = (size_t)std::max(kGramChunks, 36) * GT * GT * sizeof(double);
   const int groups = (2 * kMaxGramItems + kGramWaves - 1) / kGramWaves;
   const size_t g2 = (size_t)std::max(2 * num_cus, groups) * kGramWaves * (GT * GT / 2) * sizeof(double);
   return std::max(g1, g2);
}

int gram_multi(ofrr_ctx *ctx, int64_t n, int k, int tb, int tz, GramMultiArgs &args) {
   if (args.nitems <= 0) return kSuccess;
   if (args.nitems > kMaxGramItems || k <= 0 || k > 128)
      return set_error(ctx, kErrorInvalidArgument, "gram_multi: items / k");
   // 64 x 32 output tiles of the 64 x 64 items (halves past column k dropped)
   GramLdsArgs la{};
   GramOutArgs oa{};
   int nh = 0;
   int ha[kMaxHalfItems], hb[kMaxHalfItems];
   for (int t = 0; t < args.nitems; ++t)
      for (int h = 0; h < 2; ++h) {
         const int j0 = args.tj[t] * GT + h * (GT / 2);
         if (j0 >= k) continue;
         ha[nh] = args.ia[t];
         hb[nh] = args.ib[t];
         la.i0[nh] = oa.i0[nh] = (short)(args.ti[t] * GT);
         la.j0[nh] = oa.j0[nh] = (short)j0;
         oa.dst[nh] = args.dst[t];
         oa.ld[nh] = args.ld[t];
         oa.ro[nh] = args.ro[t];
         oa.co[nh] = args.co[t];
         ++nh;
      }
   la.nitems = oa.nitems = nh;
   const int ngroups = (nh + kGramWaves - 1) / kGramWaves;
   // the staged operands: those the items use, in kind order
   bool used[kMaxGramOps] = {};
   for (int t = 0; t < nh; ++t) used[ha[t]] = used[hb[t]] = true;
   int pos[kMaxGramOps], cnt[3] = {0, 0, 0}, nop = 0;
   for (int kind = 0; kind < 3; ++kind)
      for (int q = 0; q < kMaxGramOps; ++q)
         if (used[q] && args.op[q].kind == kind) {
            pos[q] = nop;
            la.op[nop++] = args.op[q];
            ++cnt[kind];
         }
   for (int q = 0; q < kMaxGramOps; ++q)
      if (used[q] && (args.op[q].kind < 0 || args.op[q].kind > 2))
         return set_error(ctx, kErrorInvalidArgument, "gram_multi: operand kind");
   if (nop > kGramStageOps) return set_error(ctx, kErrorInternalError, "gram_multi: staged operands");
   for (int t = 0; t < nh; ++t) {
      la.slot_a[t] = (short)pos[ha[t]];
      la.slot_b[t] = (short)pos[hb[t]];
   }
   const int nop_max = nop;
   // R rows per slab: R k <= 1024 elements per operand, two per thread
   la.R = std::max(4, std::min(128, 1024 / k));  // a power of two: R k <= 1024 <= 2 x 768
   la.lk = 0;
   while ((1 << la.lk) < k) ++la.lk;
   if ((1 << la.lk) != k) return set_error(ctx, kErrorInvalidArgument, "gram_multi: k a power of two");
   la.stride = k + ((16 - k % 32) + 32) % 32;  // row stride = 16 mod 32 doubles: two-pass LDS reads
   const size_t lds = (size_t)nop_max * la.R * la.stride * sizeof(double);
   // about two workgroups per CU over the device (for the sizing item count)
   const int size_groups =
       ((args.size_halves > 0 ? std::max(args.size_halves, nh) : nh) + kGramWaves - 1) / kGramWaves;
   int64_t want = std::max<int64_t>(1, (2 * (int64_t)ctx->num_cus) / size_groups);
   int64_t rows = (n + want - 1) / want;
   rows = std::max<int64_t>(la.R, (rows + la.R - 1) / la.R * la.R);
   const int nchunks = (int)((n + rows - 1) / rows);
   double *part = (double *)workspace(ctx, (size_t)nchunks * nh * (GT * GT / 2) * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "gram_multi: workspace");
   const unsigned grid = (unsigned)(nchunks * ngroups);
   // operand patterns (FP64, basis format, Z format) of the filtered solver:
   // [X W | D_d D_mid | Z_d Z_mid] and [X W | D_d | Z_d] (G and H together, or
   // H alone), [X | D_d D_mid] and [X | D_d] (G alone)
   const int pattern = 100 * cnt[0] + 10 * cnt[1] + cnt[2];
   if (pattern != 222 && pattern != 211 && pattern != 120 && pattern != 110)
      return set_error(ctx, kErrorInvalidArgument, "gram_multi: operand pattern");
   int st;
   if (cnt[2] == 0) {
      // no operand in the Z format: one instantiation per basis format
      st = dispatch_dtype(tb, [&](auto b) -> int {
         using TB = decltype(b);
         auto launch = [&](auto tag) -> int {
            constexpr int P = decltype(tag)::value;
            constexpr auto kern = (P == 120) ? gram_lds_kernel<TB, TB, 1, 2, 0> : gram_lds_kernel<TB, TB, 1, 1, 0>;
            static std::once_flag once;
            static hipError_t attr = hipSuccess;
            std::call_once(once, [] {
               attr = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          160 * 1024);
            });
            OFRR_HIP(ctx, attr);
            kern<<<grid, kGramThreads, lds, ctx->stream>>>(n, k, la, rows, nchunks, ngroups, part);
            OFRR_LAUNCH(ctx);
            return kSuccess;
         };
         return pattern == 120 ? launch(std::integral_constant<int, 120>{})
                               : launch(std::integral_constant<int, 110>{});
      });
   } else st = dispatch_dtype(tb, [&](auto b) -> int {
      using TB = decltype(b);
      return dispatch_dtype(tz, [&](auto z) -> int {
         using TZ = decltype(z);
         // one instantiation (and one attribute set-up) per pattern
         auto launch = [&](auto tag) -> int {
            constexpr int P = decltype(tag)::value;
            constexpr auto kern = (P == 222) ? gram_lds_kernel<TB, TZ, 2, 2, 2> : gram_lds_kernel<TB, TZ, 2, 1, 1>;
            static std::once_flag once;
            static hipError_t attr = hipSuccess;
            std::call_once(once, [] {
               attr = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          160 * 1024);
            });
            OFRR_HIP(ctx, attr);
            kern<<<grid, kGramThreads, lds, ctx->stream>>>(n, k, la, rows, nchunks, ngroups, part);
            OFRR_LAUNCH(ctx);
            return kSuccess;
         };
         return pattern == 222 ? launch(std::integral_constant<int, 222>{})
                               : launch(std::integral_constant<int, 211>{});
      });
   });
   if (st) return st;
   gram_lds_reduce_kernel<<<dim3(GT * GT / 2 / 32, nh), 256, 0, ctx->stream>>>(nchunks, k, oa, part);
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

// Y (n x k, row-major) = sum_b V_b C_b on FP64 MFMA for row-major blocks whose
// widths are multiples of 16.  One wave owns 16 rows of Y and all k columns.
// The inner index of MFMA step s (4 per group of 16 inner columns) and lane
// group g = lane / 16 is 16 G + 4 g + s, so each lane reads 4 consecutive
// entries of its row of V_b; the coefficients are repacked beforehand
// (repack_coef_kernel) so that its 4 matching entries of C are contiguous too.
// Entry (i, j) of a 16 x 16 result sits in lane j + 16 (i % 4), register i / 4.
struct UpdMfmaArgs {
   const void *V[4];
   int64_t rs[4];
   int kb[4], dtype[4];
   int nb;
};

// Cp[((G k + col) 4 + g) 4 + s] = C_b(16 G' + 4 g + s, col) with G running over
// the 16-column groups of all blocks in order
__global__ void repack_coef_kernel(int k, int nb, UpdArgs a, double *__restrict__ Cp) {
   int G = blockIdx.x, b = 0, Gb = G;
   while (b < nb && Gb >= a.b[b].kb / 16) { Gb -= a.b[b].kb / 16; ++b; }
   if (b >= nb) return;
   for (int t = threadIdx.x; t < k * 16; t += blockDim.x) {
      const int col = t / 16, g = (t / 4) % 4, s = t % 4;
      Cp[(((int64_t)G * k + col) * 4 + g) * 4 + s] =
          a.b[b].C[(16 * Gb + 4 * g + s) + (int64_t)col * a.b[b].ldc];
   }
}

template <typename TV>
__device__ inline void upd_load4(const void *V, int64_t off, bool ok, double out[4]) {
   const TV *p = (const TV *)V + off;
#pragma unroll
   for (int s = 0; s < 4; ++s) out[s] = ok ? (double)Elem<TV>::load(p[s]) : 0.0;
}

template <typename TY, int NCT, int RT>
__global__ __launch_bounds__(64) void basis_update_mfma_kernel(int64_t n, int k, UpdMfmaArgs a,
                                                               const double *__restrict__ Cp,
                                                               TY *__restrict__ Y, int64_t yrs) {
   // one wave: RT tiles of 16 rows; every coefficient fragment loaded is used
   // by the RT row tiles (the coefficients come through L2)
   const int lane = threadIdx.x, li = lane & 15, g = lane >> 4;
   const int64_t row0 = (int64_t)blockIdx.x * 16 * RT;
   dbl4 acc[RT][NCT];
#pragma unroll
   for (int r = 0; r < RT; ++r)
#pragma unroll
      for (int c = 0; c < NCT; ++c) acc[r][c] = dbl4{0.0, 0.0, 0.0, 0.0};
   int G = 0;
   for (int b = 0; b < a.nb; ++b) {
      const int ng = a.kb[b] / 16, dt = a.dtype[b];
      for (int gb = 0; gb < ng; ++gb, ++G) {
         double av[RT][4];
#pragma unroll
         for (int r = 0; r < RT; ++r) {
            const int64_t row = row0 + 16 * r + li;
            const bool rok = row < n;
            const int64_t off = row * a.rs[b] + 16 * gb + 4 * g;
            switch (dt) {  // wave-uniform
               case OFRR_F64: upd_load4<double>(a.V[b], off, rok, av[r]); break;
               case OFRR_F32: upd_load4<float>(a.V[b], off, rok, av[r]); break;
               case OFRR_F16: upd_load4<__half>(a.V[b], off, rok, av[r]); break;
               case OFRR_BF16: upd_load4<__hip_bfloat16>(a.V[b], off, rok, av[r]); break;
               default: upd_load4<fp8_t>(a.V[b], off, rok, av[r]); break;
            }
         }
         const double *cg = Cp + ((int64_t)G * k * 4 + g) * 4;
#pragma unroll
         for (int c = 0; c < NCT; ++c) {
            const double *cb = cg + (int64_t)(16 * c + li) * 16;
            const double b0 = cb[0], b1 = cb[1], b2 = cb[2], b3 = cb[3];
#pragma unroll
            for (int r = 0; r < RT; ++r) {
               acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][0], b0, acc[r][c], 0, 0, 0);
               acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][1], b1, acc[r][c], 0, 0, 0);
               acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][2], b2, acc[r][c], 0, 0, 0);
               acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][3], b3, acc[r][c], 0, 0, 0);
            }
         }
      }
   }
#pragma unroll
   for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int c = 0; c < NCT; ++c)
#pragma unroll
         for (int t = 0; t < 4; ++t) {
            const int64_t r = row0 + 16 * rt + 4 * t + g;
            if (r < n) Y[r * yrs + 16 * c + li] = Elem<TY>::store(acc[rt][c][t]);
         }
}

// The same product with the repacked coefficients staged once per workgroup in
// LDS.  Read from L2, every wave fetched all of them for its 32 rows (about
// 98 KB against 10 KB of basis rows); here a persistent workgroup of kUpdWaves
// waves (one per CU, 4 waves per SIMD) copies them once and walks row tiles of
// 16 rows per wave (launched with RT = 1: with the coefficients in LDS, two
// tiles per wave only cost registers).
// Column col of group G sits at (G k + col) kUpdLdsStride + 4 g + s: the 2
// doubles of padding put the 16 lanes of one g on distinct banks.  The MFMA
// sequence of every output is the one of basis_update_mfma_kernel.
constexpr int kUpdWaves = 16;
constexpr int kUpdLdsStride = 18;
constexpr size_t kUpdLdsMax = 160 * 1024;
template <typename TY, int NCT, int RT>
__global__ __launch_bounds__(64 * kUpdWaves) void basis_update_lds_kernel(
    int64_t n, int k, int groups, UpdMfmaArgs a, const double *__restrict__ Cp, TY *__restrict__ Y,
    int64_t yrs) {
   extern __shared__ __align__(16) double sC[];
   const int total = groups * k * 16;  // Cp: [((G k + col) 4 + g) 4 + s], 16 doubles per column
   for (int t = 2 * (int)threadIdx.x; t < total; t += 2 * (int)blockDim.x) {
      const int col = t >> 4, w = t & 15;
      *reinterpret_cast<double2 *>(sC + col * kUpdLdsStride + w) = *reinterpret_cast<const double2 *>(Cp + t);
   }
   __syncthreads();
   const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, g = lane >> 4;
   const int64_t tiles = (n + 16 * RT - 1) / (16 * RT);
   for (int64_t tile = (int64_t)blockIdx.x * kUpdWaves + wave; tile < tiles;
        tile += (int64_t)gridDim.x * kUpdWaves) {
      const int64_t row0 = tile * 16 * RT;
      dbl4 acc[RT][NCT];
#pragma unroll
      for (int r = 0; r < RT; ++r)
#pragma unroll
         for (int c = 0; c < NCT; ++c) acc[r][c] = dbl4{0.0, 0.0, 0.0, 0.0};
      int G = 0;
      for (int b = 0; b < a.nb; ++b) {
         const int ng = a.kb[b] / 16, dt = a.dtype[b];
         for (int gb = 0; gb < ng; ++gb, ++G) {
            double av[RT][4];
#pragma unroll
            for (int r = 0; r < RT; ++r) {
               const int64_t row = row0 + 16 * r + li;
               const bool rok = row < n;
               const int64_t off = row * a.rs[b] + 16 * gb + 4 * g;
               switch (dt) {  // wave-uniform
                  case OFRR_F64: upd_load4<double>(a.V[b], off, rok, av[r]); break;
                  case OFRR_F32: upd_load4<float>(a.V[b], off, rok, av[r]); break;
                  case OFRR_F16: upd_load4<__half>(a.V[b], off, rok, av[r]); break;
                  case OFRR_BF16: upd_load4<__hip_bfloat16>(a.V[b], off, rok, av[r]); break;
                  default: upd_load4<fp8_t>(a.V[b], off, rok, av[r]); break;
               }
            }
#pragma unroll
            for (int c = 0; c < NCT; ++c) {
               const double *cb = sC + (G * k + 16 * c + li) * kUpdLdsStride + 4 * g;
               const double b0 = cb[0], b1 = cb[1], b2 = cb[2], b3 = cb[3];
#pragma unroll
               for (int r = 0; r < RT; ++r) {
                  acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][0], b0, acc[r][c], 0, 0, 0);
                  acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][1], b1, acc[r][c], 0, 0, 0);
                  acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][2], b2, acc[r][c], 0, 0, 0);
                  acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r][3], b3, acc[r][c], 0, 0, 0);
               }
            }
         }
      }
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
         for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
               const int64_t r = row0 + 16 * rt + 4 * t + g;
               if (r < n) Y[r * yrs + 16 * c + li] = Elem<TY>::store(acc[rt][c][t]);
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
   // FP64 MFMA path: row-major blocks and output, widths multiples of 16
   bool mfma = (ycs == 1) && (k % 16 == 0) && (k <= 128) && (ydt == OFRR_F64 || ydt == OFRR_F32);
   int groups = 0;
   for (int i = 0; i < nb; ++i) {
      mfma = mfma && (vcs[i] == 1) && (kb[i] % 16 == 0) && (kb[i] > 0);
      groups += kb[i] / 16;
   }
   if (mfma) {
      double *Cp = (double *)workspace(ctx, (size_t)groups * k * 16 * sizeof(double));
      if (!Cp) return set_error(ctx, kErrorCUDAMemory, "basis_update: workspace");
      repack_coef_kernel<<<groups, 256, 0, ctx->stream>>>(k, nb, a, Cp);
      OFRR_LAUNCH(ctx);
      UpdMfmaArgs m{};
      m.nb = nb;
      for (int i = 0; i < nb; ++i) {
         m.V[i] = V[i];
         m.rs[i] = vrs[i];
         m.kb[i] = kb[i];
         m.dtype[i] = vdt[i];
      }
      // two 16-row tiles per wave share each coefficient load: 13.2 -> 10.2 ms
      // of back-transforms per 128^3 solve against one tile (four tiles drop
      // to one wave per SIMD and are slower)
      constexpr int kUpdRT = 2;
      const size_t lds = (size_t)groups * k * kUpdLdsStride * sizeof(double);
      int lds_st = kSuccess;
      auto launch = [&](auto nct_tag) {
         constexpr int NCT = decltype(nct_tag)::value;
         constexpr int RT = kUpdRT;
         const int64_t tiles = (n + 16 * RT - 1) / (16 * RT);
         // (wider outputs need more accumulators than 3 waves per SIMD leave)
         if constexpr (NCT <= 4) if (lds <= kUpdLdsMax) {
            // coefficients in LDS, persistent workgroups (at most one per CU),
            // one 16-row tile per wave
            const int64_t tiles1 = (n + 15) / 16;
            const unsigned grl = (unsigned)std::max<int64_t>(
                1, std::min<int64_t>(ctx->num_cus, (tiles1 + kUpdWaves - 1) / kUpdWaves));
            auto go = [&](auto ytag) {
               using TYk = decltype(ytag);
               constexpr auto kern = basis_update_lds_kernel<TYk, NCT, 1>;
               static std::once_flag once;
               static hipError_t attr = hipSuccess;
               std::call_once(once, [] {
                  attr = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)kUpdLdsMax);
               });
               if (attr != hipSuccess) {
                  lds_st = set_error(ctx, kErrorCUDAGeneral, "basis_update: LDS size attribute");
                  return;
               }
               kern<<<grl, 64 * kUpdWaves, lds, ctx->stream>>>(n, k, groups, m, Cp, (TYk *)Y, yrs);
            };
            if (ydt == OFRR_F64) go(double{}); else go(float{});
            return;
         }
         const unsigned gr = (unsigned)tiles;
         if (ydt == OFRR_F64)
            basis_update_mfma_kernel<double, NCT, RT><<<gr, 64, 0, ctx->stream>>>(n, k, m, Cp, (double *)Y, yrs);
         else
            basis_update_mfma_kernel<float, NCT, RT><<<gr, 64, 0, ctx->stream>>>(n, k, m, Cp, (float *)Y, yrs);
      };
      switch (k / 16) {
         case 1: launch(std::integral_constant<int, 1>{}); break;
         case 2: launch(std::integral_constant<int, 2>{}); break;
         case 3: launch(std::integral_constant<int, 3>{}); break;
         case 4: launch(std::integral_constant<int, 4>{}); break;
         case 5: launch(std::integral_constant<int, 5>{}); break;
         case 6: launch(std::integral_constant<int, 6>{}); break;
         case 7: launch(std::integral_constant<int, 7>{}); break;
         default: launch(std::integral_constant<int, 8>{}); break;
      }
      if (lds_st) return lds_st;
      OFRR_LAUNCH(ctx);
      return kSuccess;
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

int ofrr_gram_noalloc(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
                      int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha,
                      double beta, double *G, int ldg) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || k1 < 0 || k2 < 0 || ldg < k1) return set_error(ctx, kErrorInvalidArgument, "gram_noalloc: bad sizes");
   if (k1 == 0 || k2 == 0) return kSuccess;
   if (!V || !W || !G) return set_error(ctx, kErrorInvalidArgument, "gram_noalloc: null pointer");
   const int64_t tiles = (int64_t)((k1 + GT - 1) / GT) * ((k2 + GT - 1) / GT);
   if (tiles > kScratchChunkTiles)
      return set_error(ctx, kErrorBLASInvalidSize,
                       "gram_noalloc: ceil(k1/64) * ceil(k2/64) exceeds the 256 tiles of its scratch");
   if (!dtype_valid(vdt) || !dtype_valid(wdt)) return set_error(ctx, kErrorBLASUnsupportedType, "gram_noalloc: dtype");
   double *part = gram_scratch(ctx->device);
   if (!part) return set_error(ctx, kErrorInternalError, "gram_noalloc: scratch address not set up for this device");
   return gram_run(ctx, n, k1, k2, V, vdt, vrs, vcs, W, wdt, wrs, wcs, alpha, beta, G, ldg, part,
                   kScratchChunkTiles);
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
