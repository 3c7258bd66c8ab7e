// CSR SpMM kernels for gfx950 (semantics of MATLAB/+src/+csr/csrmv.m:48-59:
// y(i,:) = sum over the stored entries p of row i of a_p * x(j_p,:)).
//
// Layout in HBM: the hot path keeps n x k blocks row-major (a row of k values
// is contiguous), so one operator entry gathers one contiguous row of X.  A
// group of G lanes of a wave64 serves one row, each lane loading 16 bytes
// (VEC elements) of the gathered row: for k = 64 that is 8 rows per wave in
// FP16 (128 B per row) and 2 rows per wave in FP64 (512 B per row).
// Accumulation is FP64 for FP64 operands and FP32 otherwise; entries of a row
// are accumulated with fused multiply-adds in storage order.  Each workgroup
// owns 256 consecutive rows and first copies their row pointers, column
// indices and values to LDS with coalesced loads; a row then costs one global
// latency, its (up to 8) gathers in flight together.  The block index is
// remapped so that each XCD walks its own part of the matrix: a contiguous row
// range in general, and for operators with a wide band (3D stencils) one strip
// of every plane (sched_block), so the XCD's L2 serves the +-1, +-N and +-N^2
// neighbours.  Read-once and write-once blocks use non-temporal accesses.
//
// Kernels:
//   spmm_rm      Y = alpha A X + beta Y
//   spmm_resid   W = A X, R = W - X diag(theta) -> quantised R, column max |R|,
//                per-block FP64 sums of r^2 and x^2 (residual norms)
//   cheb_step    T = a.(A Dc) + b.Dc + g.Dp + d.R, one Chebyshev-filter step;
//                the per-column coefficients and the output scale s are formed
//                in the kernel from device-side scales, T / s is stored in the
//                basis format and column max |T| is folded in with atomics
//   spmm_strided generic strides (any layout), one thread per output
#include "internal.hpp"

namespace ofrr {

template <typename T, int N> struct alignas(sizeof(T) * N) VecT { T v[N]; };

template <typename T, int N> __device__ inline VecT<T, N> vload(const T *p) {
   return *reinterpret_cast<const VecT<T, N> *>(p);
}
template <typename T, int N> __device__ inline void vstore(T *p, const VecT<T, N> &v) {
   *reinterpret_cast<VecT<T, N> *>(p) = v;
}

// streamed (read-once / write-once) 16-byte accesses bypass the cache
// hierarchy's retention so the gathered block keeps its L2 / MALL residency
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <int B> struct RawVec;
template <> struct RawVec<16> { using type = u32x4; };
template <> struct RawVec<8> { using type = u32x2; };
template <typename T, int N> __device__ inline VecT<T, N> vload_nt(const T *p) {
   using R = typename RawVec<sizeof(VecT<T, N>)>::type;
   R u = __builtin_nontemporal_load(reinterpret_cast<const R *>(p));
   VecT<T, N> r;
   __builtin_memcpy(&r, &u, sizeof(r));
   return r;
}
template <typename T, int N> __device__ inline void vstore_nt(T *p, const VecT<T, N> &v) {
   using R = typename RawVec<sizeof(VecT<T, N>)>::type;
   R u;
   __builtin_memcpy(&u, &v, sizeof(u));
   __builtin_nontemporal_store(u, reinterpret_cast<R *>(p));
}

// store an accumulator value in a block format without a detour through FP64
template <typename TB, typename ACC> __device__ inline TB store_acc(ACC v) {
   if constexpr (std::is_same<TB, __half>::value) return __float2half((float)v);
   else if constexpr (std::is_same<TB, __hip_bfloat16>::value) return __float2bfloat16((float)v);
   else if constexpr (std::is_same<TB, float>::value) return (float)v;
   else return Elem<TB>::store((double)v);
}

// power-of-two exponent e with 2^e >= 2 m > 2^(e-1), i.e. ceil(log2(2 m)), m > 0
__device__ inline int ceil_log2_2m(float m) {
   int e;
   const float f = frexpf(2.0f * m, &e);  // 2 m = f 2^e, f in [0.5, 1)
   return (f == 0.5f) ? e - 1 : e;
}

// acc[v] += sum_p a_p * X(col_p, c0 + v) over entries [b, e) of the index /
// value arrays cs / vs (LDS or global), in storage order, 8 gathers in flight;
// the padding slots of the last group of 8 add 0 * X(self) (finite).
// OVERREAD: cs / vs may be read up to 7 entries past e (padded LDS copies), so
// the index loads are unconditional and only the selects depend on the bound.
template <typename TV, typename TX, int VEC, typename ACC, bool OVERREAD, typename CI, typename VA>
__device__ inline void row_gather8(CI cs, VA vs, int64_t b, int64_t e, const TX *__restrict__ X,
                                   int64_t ldx, int64_t self, int c0, ACC acc[VEC]) {
   for (int64_t p = b; p < e; p += 8) {
      int32_t c[8];
      ACC a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
         const bool ok = p + u < e;
         if (OVERREAD) {
            const int32_t cl = cs[p + u];
            const ACC al = (ACC)Elem<TV>::load(vs[p + u]);
            c[u] = ok ? cl : (int32_t)self;
            a[u] = ok ? al : (ACC)0;
         } else {
            c[u] = ok ? cs[p + u] : (int32_t)self;
            a[u] = ok ? (ACC)Elem<TV>::load(vs[p + u]) : (ACC)0;
         }
      }
      VecT<TX, VEC> xv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) xv[u] = vload<TX, VEC>(X + (int64_t)c[u] * ldx + c0);
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
         for (int v = 0; v < VEC; ++v) acc[v] = fma(a[u], (ACC)Elem<TX>::load(xv[u].v[v]), acc[v]);
   }
}

// ---------------------------------------------------------------------------
// LDS copy of the CSR rows of one 256-row block: row offsets relative to the
// block's first entry, and the indices / values padded by 8 entries so that the
// gathers may read past a row's end (row_gather8<..., true>).
constexpr int kTileRows = 256;
constexpr int kTileCap = 8 * kTileRows;
template <typename TV, int ROWS = kTileRows> struct CsrTile {
   static constexpr int kRows = ROWS, kCap = 8 * ROWS;
   int32_t rp[ROWS + 1];
   int32_t ci[kCap + 8];
   TV va[kCap + 8];
};
// The copy runs in two phases, loads into registers and then LDS stores, over a
// fixed number of passes (workgroups of kCsrThreads threads), so a workgroup
// waits for one memory round trip.  A loop of load / wait / store per pass of
// 256 entries waited for nine of them in a row (2,056 entries at most) before
// its first row could start.
constexpr int kCsrThreads = 256;
template <typename TV, int ROWS = kTileRows> struct TileRegs {
   static constexpr int kCap = 8 * ROWS;
   static constexpr int kRp = (ROWS + 1 + kCsrThreads - 1) / kCsrThreads;
   static constexpr int kNz = (kCap + 8 + kCsrThreads - 1) / kCsrThreads;
   int64_t base, nnzb;
   bool in_lds;
   int64_t rp[kRp];
   int32_t ci[kNz];
   TV va[kNz];
};
// load phase for rows [r0, r0 + nr); in_lds is false when they hold more than
// 8 entries per row (the offsets are staged anyway; indices then come from
// ci + base, va + base).  Every address is clamped into the tile's rows.
template <typename TV, int ROWS>
__device__ inline void tile_load(TileRegs<TV, ROWS> &g, const int64_t *__restrict__ rp,
                                 const int32_t *__restrict__ ci, const TV *__restrict__ va,
                                 int64_t r0, int nr) {
   g.base = rp[r0];
   g.nnzb = rp[r0 + nr] - g.base;
   g.in_lds = g.nnzb <= TileRegs<TV, ROWS>::kCap;
   if (g.in_lds && g.nnzb > 0) {  // uniform across the workgroup
      const int last = (int)g.nnzb - 1;
#pragma unroll
      for (int u = 0; u < TileRegs<TV, ROWS>::kNz; ++u) {
         const int j = min((int)threadIdx.x + u * kCsrThreads, last);
         g.ci[u] = ci[g.base + j];
         g.va[u] = va[g.base + j];
      }
   } else {
#pragma unroll
      for (int u = 0; u < TileRegs<TV, ROWS>::kNz; ++u) {
         g.ci[u] = 0;
         g.va[u] = TV(0);
      }
   }
#pragma unroll
   for (int u = 0; u < TileRegs<TV, ROWS>::kRp; ++u)
      g.rp[u] = rp[r0 + min((int)threadIdx.x + u * kCsrThreads, nr)];
}
// store phase: row offsets relative to the first entry, indices and values
// padded by 8 zero entries
template <typename TV, int ROWS>
__device__ inline void tile_store(CsrTile<TV, ROWS> &t, const TileRegs<TV, ROWS> &g, int nr) {
#pragma unroll
   for (int u = 0; u < TileRegs<TV, ROWS>::kRp; ++u) {
      const int i = (int)threadIdx.x + u * kCsrThreads;
      if (i <= nr) t.rp[i] = (int32_t)(g.rp[u] - g.base);
   }
   if (g.in_lds) {
      const int nz = (int)g.nnzb;
#pragma unroll
      for (int u = 0; u < TileRegs<TV, ROWS>::kNz; ++u) {
         const int i = (int)threadIdx.x + u * kCsrThreads;
         if (i < nz + 8) {
            const bool ok = i < nz;
            t.ci[i] = ok ? g.ci[u] : 0;
            t.va[i] = ok ? g.va[u] : TV(0);
         }
      }
   }
}
template <typename TV, int ROWS>
__device__ inline bool stage_tile(CsrTile<TV, ROWS> &t, const int64_t *__restrict__ rp,
                                  const int32_t *__restrict__ ci, const TV *__restrict__ va,
                                  int64_t r0, int nr, int64_t &base) {
   TileRegs<TV, ROWS> g;
   tile_load(g, rp, ci, va, r0, nr);
   tile_store(t, g, nr);
   base = g.base;
   return g.in_lds;
}

// ---------------------------------------------------------------------------
template <typename TV, typename TX, typename TY, int VEC, typename ACC, bool OVERREAD, bool YVEC,
          typename CI, typename VA>
__device__ inline void spmm_rows(int64_t r0, int nr, int grp, int groups, int lane_g, int G, int k,
                                 CI cs, VA vs, const int32_t *s_rp, const TX *__restrict__ X,
                                 int64_t ldx, int64_t nx, double alpha, double beta,
                                 TY *__restrict__ Y, int64_t ldy) {
   for (int c0 = lane_g * VEC; c0 < k; c0 += G * VEC) {
      for (int lr = grp; lr < nr; lr += groups) {
         const int64_t row = r0 + lr;
         ACC acc[VEC];
#pragma unroll
         for (int v = 0; v < VEC; ++v) acc[v] = 0;
         // the padding slots gather a row of X that exists (X has nx rows)
         const int64_t self = row < nx ? row : nx - 1;
         row_gather8<TV, TX, VEC, ACC, OVERREAD>(cs, vs, s_rp[lr], s_rp[lr + 1], X, ldx, self, c0,
                                                 acc);
         TY *y = Y + row * ldy + c0;
         if (YVEC) {
            // the lane's VEC outputs are contiguous and aligned: one vector store
            VecT<TY, VEC> out;
            if (beta != 0.0) {
               const VecT<TY, VEC> yo = vload<TY, VEC>(y);
#pragma unroll
               for (int v = 0; v < VEC; ++v)
                  out.v[v] = Elem<TY>::store(alpha * (double)acc[v] + beta * (double)Elem<TY>::load(yo.v[v]));
            } else {
#pragma unroll
               for (int v = 0; v < VEC; ++v) out.v[v] = Elem<TY>::store(alpha * (double)acc[v]);
            }
            vstore<TY, VEC>(y, out);
         } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) {
               double r = alpha * (double)acc[v];
               if (beta != 0.0) r += beta * (double)Elem<TY>::load(y[v]);
               y[v] = Elem<TY>::store(r);
            }
         }
      }
   }
}

template <typename TV, typename TX, typename TY, int VEC, bool YVEC>
__global__ __launch_bounds__(256) void spmm_rm_kernel(
    int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const TV *__restrict__ va, int k, const TX *__restrict__ X, int64_t ldx, int64_t nx,
    double alpha, double beta, TY *__restrict__ Y, int64_t ldy, int lg, int64_t nblk,
    int64_t plane) {
   // FP64 accumulation for FP64 operands or an FP64 result
   using ACC = typename std::conditional<std::is_same<TX, double>::value || std::is_same<TY, double>::value,
                                         double, float>::type;
   __shared__ CsrTile<TV> tile;
   const int64_t blk = sched_block(nblk, plane);
   if (blk < 0) return;
   const int G = 1 << lg;
   const int lane_g = threadIdx.x & (G - 1), grp = threadIdx.x >> lg;
   const int GROUPS = 256 >> lg;
   const int64_t r0 = blk * kTileRows;
   const int nr = (int)min((int64_t)kTileRows, n - r0);
   int64_t base;
   const bool in_lds = stage_tile(tile, rp, ci, va, r0, nr, base);
   __syncthreads();
   if (in_lds)
      spmm_rows<TV, TX, TY, VEC, ACC, true, YVEC>(r0, nr, grp, GROUPS, lane_g, G, k, tile.ci,
                                                  tile.va, tile.rp, X, ldx, nx, alpha, beta, Y, ldy);
   else
      spmm_rows<TV, TX, TY, VEC, ACC, false, YVEC>(r0, nr, grp, GROUPS, lane_g, G, k, ci + base,
                                                   va + base, tile.rp, X, ldx, nx, alpha, beta, Y,
                                                   ldy);
}

// ---------------------------------------------------------------------------
// W = A X (FP64), R = W - X theta, Rq = quantise(R * inv_sR), colmax |R|,
// partial sums of r^2 and x^2 per block (fixed order -> deterministic)
template <typename TB, bool OVERREAD, typename CI, typename VA>
__device__ inline void resid_rows(int64_t r0, int nr, int grp, int groups, int c0, int k, CI cs,
                                  VA vs, const int32_t *s_rp, const double *__restrict__ X,
                                  const double th[2], const double isr[2],
                                  double *__restrict__ W, TB *__restrict__ Rq, double s_r[2],
                                  double s_x[2], float mx[2]) {
   constexpr int VEC = 2;
   for (int lr = grp; lr < nr; lr += groups) {
      const int64_t row = r0 + lr;
      const VecT<double, VEC> xv = vload<double, VEC>(X + row * k + c0);
      double acc[VEC] = {0.0, 0.0};
      row_gather8<double, double, VEC, double, OVERREAD>(cs, vs, s_rp[lr], s_rp[lr + 1], X, k, row,
                                                         c0, acc);
      VecT<double, VEC> wv;
      VecT<TB, VEC> rq;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
         wv.v[v] = acc[v];
         const double r = fma(-th[v], xv.v[v], acc[v]);
         s_r[v] = fma(r, r, s_r[v]);
         s_x[v] = fma(xv.v[v], xv.v[v], s_x[v]);
         mx[v] = fmaxf(mx[v], (float)fabs(r));
         rq.v[v] = Elem<TB>::store(r * isr[v]);
      }
      vstore_nt<double, VEC>(W + row * k + c0, wv);
      vstore<TB, VEC>(Rq + row * k + c0, rq);
   }
}

// W = A X (FP64), R = W - X theta, Rq = quantise(R * inv_sR), colmax |R|,
// partial sums of r^2 and x^2 per block (fixed order -> deterministic)
template <typename TB>
__global__ __launch_bounds__(256) void spmm_resid_kernel(
    int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const double *__restrict__ va, int k, const double *__restrict__ X,
    const double *__restrict__ theta, double *__restrict__ W, TB *__restrict__ Rq,
    const double *__restrict__ inv_sR, float *__restrict__ colmax, double *__restrict__ part,
    int lg, int64_t nblk, int64_t plane) {
   constexpr int VEC = 2;
   __shared__ CsrTile<double> tile;
   __shared__ double red[256][VEC * 2];
   __shared__ float redm[256][VEC];
   const int64_t blk = sched_block(nblk, plane);
   if (blk < 0) return;
   const int G = 1 << lg;
   const int GROUPS = 256 >> lg;
   const int lane_g = threadIdx.x & (G - 1), grp = threadIdx.x >> lg;
   const int c0 = lane_g * VEC;  // host guarantees k == G * VEC
   const int64_t r0 = blk * kTileRows;
   const int nr = (int)min((int64_t)kTileRows, n - r0);
   int64_t base;
   const bool in_lds = stage_tile(tile, rp, ci, va, r0, nr, base);
   double th[VEC], isr[VEC], s_r[VEC] = {0, 0}, s_x[VEC] = {0, 0};
   float mx[VEC] = {0.f, 0.f};
#pragma unroll
   for (int v = 0; v < VEC; ++v) { th[v] = theta[c0 + v]; isr[v] = inv_sR[c0 + v]; }
   __syncthreads();
   if (in_lds)
      resid_rows<TB, true>(r0, nr, grp, GROUPS, c0, k, tile.ci, tile.va, tile.rp, X, th, isr, W,
                           Rq, s_r, s_x, mx);
   else
      resid_rows<TB, false>(r0, nr, grp, GROUPS, c0, k, ci + base, va + base, tile.rp, X, th, isr,
                            W, Rq, s_r, s_x, mx);
#pragma unroll
   for (int v = 0; v < VEC; ++v) {
      red[threadIdx.x][2 * v] = s_r[v];
      red[threadIdx.x][2 * v + 1] = s_x[v];
      redm[threadIdx.x][v] = mx[v];
   }
   __syncthreads();
   // column c = lane * VEC + v; sums over the row groups in a fixed order
   for (int t = threadIdx.x; t < G * VEC * 2; t += blockDim.x) {
      const int c = t >> 1, which = t & 1, lane = c / VEC, v = c % VEC;
      double s = 0.0;
      for (int g = 0; g < GROUPS; ++g) s += red[g * G + lane][2 * v + which];
      part[blk * (int64_t)(2 * k) + t] = s;  // [col*2 + {0: r^2, 1: x^2}]
   }
   for (int t = threadIdx.x; t < G * VEC; t += blockDim.x) {
      const int lane = t / VEC, v = t % VEC;
      float m = 0.f;
      for (int g = 0; g < GROUPS; ++g) m = fmaxf(m, redm[g * G + lane][v]);
      atomic_max_nonneg(colmax + t, m);
   }
}

// ---------------------------------------------------------------------------
// one Chebyshev step on the correction block:
// T = ca.(A Dc) + cb.Dc + cg.Dp + cd.R,  Dn = quantise(T * inv_s), colmax |T|
// A block owns kChebRows consecutive rows.  Their row pointers, column indices
// and values are first copied to LDS with coalesced loads, so a row costs one
// global latency (its gathers, all in flight together) instead of the
// row-pointer -> index -> block chain; the streamed blocks (Dp, R, Dn) use
// non-temporal accesses.  A block whose rows hold more than kTileCap entries
// reads them from global memory instead (same arithmetic order).
constexpr int kChebRows = kTileRows;
constexpr int kChebMaxK = 256;

// the row loop of the Chebyshev step; HAS_DP and the index source are compile-
// time so that no branch sits between the streamed loads and the gathers
// (a join there makes the compiler wait for every outstanding load)
template <typename TV, typename TB, int VEC, typename ACC, bool HAS_DP, bool OVERREAD, typename CI,
          typename VA>
__device__ inline void cheb_rows(int64_t r0, int nr, int grp, int groups, int c0, int k, CI cs,
                                 VA vs, const int32_t *s_rp, const ACC (*s_coef)[kChebMaxK],
                                 ACC cshift, const TB *__restrict__ Dc,
                                 const TB *__restrict__ Dp, const TB *__restrict__ R,
                                 TB *__restrict__ Dn, float mx[VEC]) {
   for (int lr = grp; lr < nr; lr += groups) {
      const int64_t row = r0 + lr;
      // the streamed operands first: independent of the gathers
      const VecT<TB, VEC> rr = vload_nt<TB, VEC>(R + row * k + c0);
      VecT<TB, VEC> dp;
      if (HAS_DP) dp = vload_nt<TB, VEC>(Dp + row * k + c0);
      const VecT<TB, VEC> dc = vload<TB, VEC>(Dc + row * k + c0);
      ACC acc[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc[v] = 0;
      row_gather8<TV, TB, VEC, ACC, OVERREAD>(cs, vs, s_rp[lr], s_rp[lr + 1], Dc, k, row, c0, acc);
      asm volatile("" ::: "memory");  // keep the coefficient reads inside the loop
      const VecT<ACC, VEC> ca = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[0][c0]);
      const VecT<ACC, VEC> cd = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[2][c0]);
      VecT<ACC, VEC> cg;
      if (HAS_DP) cg = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[1][c0]);
      VecT<TB, VEC> out;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
         ACC t = ca.v[v] * fma(-cshift, (ACC)Elem<TB>::load(dc.v[v]), acc[v]);
         if (HAS_DP) t = fma(cg.v[v], (ACC)Elem<TB>::load(dp.v[v]), t);
         t = fma(cd.v[v], (ACC)Elem<TB>::load(rr.v[v]), t);
         mx[v] = fmaxf(mx[v], (float)fabs(t));
         out.v[v] = store_acc<TB>(t);
      }
      vstore_nt<TB, VEC>(Dn + row * k + c0, out);
   }
}

// The same row loop for blocks of fewer than 2^31 bytes: every address is a
// 32-bit byte offset from a kernel-argument base, so a gather costs one integer
// multiply-add and one VGPR (base + offset addressing) instead of a 64-bit
// multiply and a register pair (94 VGPRs instead of 116 for FP16 blocks: five
// waves per SIMD instead of four).  The arithmetic (order of the fused
// multiply-adds, padding slots of the last group of 8 gathering the row itself
// with weight 0) is the one of cheb_rows.
template <typename TV, typename TB, int VEC, typename ACC, bool HAS_DP, bool OVERREAD>
__device__ inline void cheb_rows32(int64_t r0, int nr, int grp, int groups, int c0, uint32_t rowb,
                                   const int32_t *cs, const TV *vs, const int32_t *s_rp,
                                   const ACC (*s_coef)[kChebMaxK], ACC cshift,
                                   const char *__restrict__ Dcb, const char *__restrict__ Dpb,
                                   const char *__restrict__ Rb, char *__restrict__ Dnb, float mx[VEC]) {
   const uint32_t c0b = (uint32_t)c0 * (uint32_t)sizeof(TB);
   for (int lr = grp; lr < nr; lr += groups) {
      const uint32_t self = (uint32_t)(r0 + lr) * rowb + c0b;
      const VecT<TB, VEC> rr = vload_nt<TB, VEC>(reinterpret_cast<const TB *>(Rb + self));
      VecT<TB, VEC> dp;
      if (HAS_DP) dp = vload_nt<TB, VEC>(reinterpret_cast<const TB *>(Dpb + self));
      const VecT<TB, VEC> dc = vload<TB, VEC>(reinterpret_cast<const TB *>(Dcb + self));
      ACC acc[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) acc[v] = 0;
      const int e = s_rp[lr + 1];
      for (int p = s_rp[lr]; p < e; p += 8) {
         uint32_t off[8];
         ACC a[8];
#pragma unroll
         for (int u = 0; u < 8; ++u) {
            const bool ok = p + u < e;
            if (OVERREAD) {  // the LDS tile is padded by 8 entries
               const int32_t cl = cs[p + u];
               const ACC al = (ACC)Elem<TV>::load(vs[p + u]);
               off[u] = ok ? (uint32_t)cl * rowb + c0b : self;
               a[u] = ok ? al : (ACC)0;
            } else {
               off[u] = ok ? (uint32_t)cs[p + u] * rowb + c0b : self;
               a[u] = ok ? (ACC)Elem<TV>::load(vs[p + u]) : (ACC)0;
            }
         }
         VecT<TB, VEC> xv[8];
#pragma unroll
         for (int u = 0; u < 8; ++u) xv[u] = vload<TB, VEC>(reinterpret_cast<const TB *>(Dcb + off[u]));
#pragma unroll
         for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int v = 0; v < VEC; ++v) acc[v] = fma(a[u], (ACC)Elem<TB>::load(xv[u].v[v]), acc[v]);
      }
      asm volatile("" ::: "memory");  // keep the coefficient reads inside the loop
      const VecT<ACC, VEC> ca = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[0][c0]);
      const VecT<ACC, VEC> cd = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[2][c0]);
      VecT<ACC, VEC> cg;
      if (HAS_DP) cg = *reinterpret_cast<const VecT<ACC, VEC> *>(&s_coef[1][c0]);
      VecT<TB, VEC> out;
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
         ACC t = ca.v[v] * fma(-cshift, (ACC)Elem<TB>::load(dc.v[v]), acc[v]);
         if (HAS_DP) t = fma(cg.v[v], (ACC)Elem<TB>::load(dp.v[v]), t);
         t = fma(cd.v[v], (ACC)Elem<TB>::load(rr.v[v]), t);
         mx[v] = fmaxf(mx[v], (float)fabs(t));
         out.v[v] = store_acc<TB>(t);
      }
      vstore_nt<TB, VEC>(reinterpret_cast<TB *>(Dnb + self), out);
   }
}

template <typename TV, typename TB, int VEC, bool HAS_DP, bool OFF32>
__global__ __launch_bounds__(256) void cheb_step_kernel(
    int64_t n, const int64_t *__restrict__ rp, const int32_t *__restrict__ ci,
    const TV *__restrict__ va, int k, const TB *__restrict__ Dc, const TB *__restrict__ Dp,
    const TB *__restrict__ R, ChebArgs ca_, TB *__restrict__ Dn, int lg, int64_t nblk, int64_t plane) {
   using ACC = typename std::conditional<std::is_same<TB, double>::value, double, float>::type;
   __shared__ CsrTile<TV, kChebRows> tile;
   __shared__ float redm[256][VEC];
   __shared__ __align__(16) ACC s_coef[3][kChebMaxK];
   __shared__ int s_se[kChebMaxK];
   const int64_t blk = sched_block(nblk, plane, ca_.reverse != 0);
   if (blk < 0) return;
   const int G = 1 << lg;
   const int GROUPS = 256 >> lg;
   const int lane_g = threadIdx.x & (G - 1), grp = threadIdx.x >> lg;
   const int c0 = lane_g * VEC;  // host guarantees k == G * VEC
   const int64_t r0 = blk * kChebRows;
   const int nr = (int)min((int64_t)kChebRows, n - r0);
   int64_t base;
   const bool in_lds = stage_tile(tile, rp, ci, va, r0, nr, base);
   // per-column coefficients from the device-side scales, with the output
   // scale 2^-se folded in (exact: a power of two):
   // T / s = (alpha sc / s)(A Dc - c Dc) + (-beta sp / s) Dp + (w sR / s) R
   // They live in LDS and are re-read per row (in registers they cost occupancy).
   const double alpha = ca_.step ? ca_.step[0] : ca_.alpha;
   const double beta = ca_.step ? ca_.step[1] : ca_.beta;
   for (int j = threadIdx.x; j < k; j += 256) {
      float m = 0.f;
#pragma unroll
      for (int sl = 0; sl < kMaxSlots; ++sl) m = fmaxf(m, ca_.maxc[sl * ca_.slot_ld + j]);
      // stored scale s = 2^(ceil(log2(2 max|T|)) - target): the stored maximum lands in [2^(t-1), 2^t]
      const int se = (m > 0.f) ? ceil_log2_2m(m) - ca_.target_exp : 0;
      s_coef[0][j] = (ACC)ldexp(alpha * ca_.sc[j], -se);
      s_coef[1][j] = (ACC)(HAS_DP ? ldexp(-beta * ca_.sp[j], -se) : 0.0);
      s_coef[2][j] = (ACC)ldexp(ca_.w[j] * ca_.sR[j], -se);
      s_se[j] = se;
      if (blk == 0) {
         ca_.sn_out[j] = ldexp(1.0, se);
         if (ca_.zero_next)
            for (int sl = 0; sl < kMaxSlots; ++sl) ca_.zero_next[sl * ca_.slot_ld + j] = 0.f;
      }
   }
   float mx[VEC];
#pragma unroll
   for (int v = 0; v < VEC; ++v) mx[v] = 0.f;
   const ACC cshift = (ACC)(ca_.step ? ca_.step[2] : ca_.c);
   __syncthreads();
   if (OFF32) {
      const uint32_t rowb = (uint32_t)(k * sizeof(TB));
      if (in_lds)
         cheb_rows32<TV, TB, VEC, ACC, HAS_DP, true>(r0, nr, grp, GROUPS, c0, rowb, tile.ci, tile.va,
                                                     tile.rp, s_coef, cshift, (const char *)Dc,
                                                     (const char *)Dp, (const char *)R, (char *)Dn, mx);
      else
         cheb_rows32<TV, TB, VEC, ACC, HAS_DP, false>(r0, nr, grp, GROUPS, c0, rowb, ci + base,
                                                      va + base, tile.rp, s_coef, cshift,
                                                      (const char *)Dc, (const char *)Dp,
                                                      (const char *)R, (char *)Dn, mx);
   } else if (in_lds)
      cheb_rows<TV, TB, VEC, ACC, HAS_DP, true>(r0, nr, grp, GROUPS, c0, k, tile.ci, tile.va,
                                                tile.rp, s_coef, cshift, Dc, Dp, R, Dn, mx);
   else  // more than kTileCap entries in this block's rows: indices from global memory
      cheb_rows<TV, TB, VEC, ACC, HAS_DP, false>(r0, nr, grp, GROUPS, c0, k, ci + base, va + base,
                                                 tile.rp, s_coef, cshift, Dc, Dp, R, Dn, mx);
   // publish the unscaled column maxima (max |T / s| * s, exact)
#pragma unroll
   for (int v = 0; v < VEC; ++v) redm[threadIdx.x][v] = ldexpf(mx[v], s_se[c0 + v]);
   __syncthreads();
   for (int t = threadIdx.x; t < G * VEC; t += blockDim.x) {
      const int lane = t / VEC, v = t % VEC;
      float m = 0.f;
      for (int g = 0; g < GROUPS; ++g) m = fmaxf(m, redm[g * G + lane][v]);
      atomic_max_nonneg(ca_.maxn + (blockIdx.x & (kMaxSlots - 1)) * ca_.slot_ld + t, m);
   }
}

// ---------------------------------------------------------------------------
template <typename TV, typename TX, typename TY>
__global__ void spmm_strided_kernel(int64_t n, const int64_t *__restrict__ rp,
                                    const int32_t *__restrict__ ci, const TV *__restrict__ va,
                                    int k, const TX *__restrict__ X, int64_t xrs, int64_t xcs,
                                    double alpha, double beta, TY *__restrict__ Y, int64_t yrs,
                                    int64_t ycs) {
   // FP64 accumulation for FP64 operands or an FP64 result
   using ACC = typename std::conditional<std::is_same<TX, double>::value || std::is_same<TY, double>::value,
                                         double, float>::type;
   int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n * (int64_t)k) return;
   int64_t row = t % n;
   int j = (int)(t / n);
   ACC acc = 0;
   for (int64_t p = rp[row]; p < rp[row + 1]; ++p)
      acc = fma((ACC)Elem<TV>::load(va[p]), (ACC)Elem<TX>::load(X[(int64_t)ci[p] * xrs + j * xcs]), acc);
   TY *y = Y + row * yrs + (int64_t)j * ycs;
   double r = alpha * (double)acc;
   if (beta != 0.0) r += beta * (double)Elem<TY>::load(*y);
   *y = Elem<TY>::store(r);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

template <typename TV, typename TX, typename TY>
static int launch_spmm(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci,
                       const TV *va, int k, const TX *X, int64_t xrs, int64_t xcs, double alpha,
                       double beta, TY *Y, int64_t yrs, int64_t ycs, int64_t bw, int64_t nx) {
   constexpr int VEC = 16 / sizeof(TX) > 0 ? 16 / sizeof(TX) : 1;
   bool fast = xcs == 1 && ycs == 1 && k % VEC == 0 && aligned16(X) &&
               (xrs * (int64_t)sizeof(TX)) % 16 == 0;
   if (fast) {
      int groups = k / VEC;
      int G = 1;
      while (G < groups && G < 64) G <<= 1;
      int64_t nblk = (n + kTileRows - 1) / kTileRows;
      const int64_t pt = plane_tiles(n, bw, kTileRows);
      dim3 grid((unsigned)sched_grid(nblk, pt));
      int lg = 0;
      while ((1 << lg) < G) ++lg;
      // a lane's VEC outputs go out as one vector store when Y's rows keep
      // that alignment
      const size_t yb = sizeof(TY) * VEC;
      const bool yvec = ((uintptr_t)Y % yb) == 0 && ((size_t)yrs * sizeof(TY)) % yb == 0;
      if (yvec)
         spmm_rm_kernel<TV, TX, TY, VEC, true><<<grid, 256, 0, ctx->stream>>>(
             n, rp, ci, va, k, X, xrs, nx, alpha, beta, Y, yrs, lg, nblk, pt);
      else
         spmm_rm_kernel<TV, TX, TY, VEC, false><<<grid, 256, 0, ctx->stream>>>(
             n, rp, ci, va, k, X, xrs, nx, alpha, beta, Y, yrs, lg, nblk, pt);
   } else {
      int64_t total = n * (int64_t)k;
      spmm_strided_kernel<TV, TX, TY><<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(
          n, rp, ci, va, k, X, xrs, xcs, alpha, beta, Y, yrs, ycs);
   }
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

int spmm_csr(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const void *va,
             int vdt, int k, const void *X, int xdt, int64_t xrs, int64_t xcs, double alpha,
             double beta, void *Y, int ydt, int64_t yrs, int64_t ycs, int64_t bw, int64_t nx) {
   if (nx < 0) nx = n;
   if (vdt != OFRR_F64 && vdt != OFRR_F32 && vdt != OFRR_F16)
      return set_error(ctx, kErrorBLASUnsupportedType, "spmm: operator values must be FP64, FP32 or FP16");
   if (ydt == OFRR_FP8)
      return set_error(ctx, kErrorBLASUnsupportedType, "spmm: FP8 output is produced by the filter step only");
   return dispatch_value_dtype(vdt, [&](auto tv) -> int {
      using TV = decltype(tv);
      return dispatch_dtype(xdt, [&](auto tx) -> int {
         using TX = decltype(tx);
         return dispatch_dtype(ydt, [&](auto ty) -> int {
            using TY = decltype(ty);
            return launch_spmm<TV, TX, TY>(ctx, n, rp, ci, (const TV *)va, k, (const TX *)X, xrs,
                                           xcs, alpha, beta, (TY *)Y, yrs, ycs, bw, nx);
         });
      });
   });
}

int spmm_resid(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const double *va,
               int k, const double *X, const double *theta, double *W, void *Rq, int rdt,
               const double *inv_sR, float *colmax, double *sums /* 2k */, int64_t bw) {
   if (k % 2 != 0 || k > 128) return set_error(ctx, kErrorInvalidArgument, "spmm_resid: k must be even and <= 128");
   int G = k / 2;
   if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16 && G != 32 && G != 64)
      return set_error(ctx, kErrorInvalidArgument, "spmm_resid: k/2 must be a power of two");
   int64_t nblk = (n + kTileRows - 1) / kTileRows;
   double *part = (double *)workspace(ctx, (size_t)nblk * 2 * k * sizeof(double));
   if (!part) return set_error(ctx, kErrorCUDAMemory, "spmm_resid: workspace");
   int st = dispatch_dtype(rdt, [&](auto tb) -> int {
      using TB = decltype(tb);
      const int64_t pt = plane_tiles(n, bw, kTileRows);
      dim3 grid((unsigned)sched_grid(nblk, pt));
      int lg = 0;
      while ((1 << lg) < G) ++lg;
      // 20 KiB of dynamic LDS the kernel does not touch hold it to three
      // workgroups per CU (56 KiB each) instead of four: the 512-byte FP64 rows
      // a CU has in flight then leave the XCD's L2 room for their +-N^2
      // neighbours (6.85 -> 6.30 ms of residual passes per 128^3 mixed solve;
      // two workgroups per CU: 8.2 ms)
      constexpr int kResidLdsPad = 20 * 1024;
      spmm_resid_kernel<TB><<<grid, 256, kResidLdsPad, ctx->stream>>>(n, rp, ci, va, k, X, theta, W, (TB *)Rq,
                                                           inv_sR, colmax, part, lg, nblk, pt);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   });
   if (st) return st;
   return reduce_partials(ctx, nblk, 2 * k, part, sums);
}

int cheb_step(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const void *va,
              int vdt, int k, const void *Dc, const void *Dp, const void *R, int bdt,
              const ChebArgs &args, void *Dn, int64_t bw) {
   if (vdt != OFRR_F64 && vdt != OFRR_F32 && vdt != OFRR_F16)
      return set_error(ctx, kErrorBLASUnsupportedType, "cheb_step: operator values must be FP64, FP32 or FP16");
   return dispatch_value_dtype(vdt, [&](auto tv) -> int {
      using TV = decltype(tv);
      return dispatch_dtype(bdt, [&](auto tb) -> int {
         using TB = decltype(tb);
         // 16-byte lane accesses; 8 bytes (8 elements) for the 1-byte format,
         // whose 16-element lanes need too many registers for 3 waves per SIMD
         constexpr int VEC = sizeof(TB) == 1 ? 8 : 16 / sizeof(TB);
         if (k % VEC != 0 || k > kChebMaxK)
            return set_error(ctx, kErrorInvalidArgument, "cheb_step: k must be a multiple of the vector width and <= 256");
         int G = k / VEC;
         if (G != 1 && G != 2 && G != 4 && G != 8 && G != 16 && G != 32 && G != 64)
            return set_error(ctx, kErrorInvalidArgument, "cheb_step: k / vector width must be a power of two <= 64");
         if (n >= ((int64_t)1 << 31)) return set_error(ctx, kErrorInvalidArgument, "cheb_step: n >= 2^31");
         int64_t nblk = (n + kChebRows - 1) / kChebRows;
         const int64_t pt = plane_tiles(n, bw, kChebRows);
         dim3 grid((unsigned)sched_grid(nblk, pt));
         int lg = 0;
         while ((1 << lg) < G) ++lg;
         // 32-bit byte offsets while every block stays below 2^31 bytes
         const bool off32 = (uint64_t)n * (uint64_t)k * sizeof(TB) < ((uint64_t)1 << 31);
         // 2-byte formats: dynamic LDS the kernel does not touch brings a
         // workgroup to about 36 KiB, four per CU instead of the five its
         // registers allow (FP16 step 195.9 -> 193.7 us; three per CU: 199.1 us)
         using ACC = typename std::conditional<std::is_same<TB, double>::value, double, float>::type;
         constexpr size_t kStatic = sizeof(CsrTile<TV, kChebRows>) + 256 * VEC * sizeof(float) +
                                    3 * kChebMaxK * sizeof(ACC) + kChebMaxK * sizeof(int);
         constexpr size_t kPad = (sizeof(TB) == 2 && kStatic < 36 * 1024) ? 36 * 1024 - kStatic : 0;
         auto go = [&](auto has_dp, auto o32) {
            cheb_step_kernel<TV, TB, VEC, decltype(has_dp)::value, decltype(o32)::value>
                <<<grid, 256, kPad, ctx->stream>>>(n, rp, ci, (const TV *)va, k, (const TB *)Dc,
                                                (const TB *)Dp, (const TB *)R, args, (TB *)Dn, lg,
                                                nblk, pt);
         };
         using T_ = std::true_type;
         using F_ = std::false_type;
         if (Dp) {
            if (off32) go(T_{}, T_{}); else go(T_{}, F_{});
         } else {
            if (off32) go(F_{}, T_{}); else go(F_{}, F_{});
         }
         OFRR_LAUNCH(ctx);
         return kSuccess;
      });
   });
}

}  // namespace ofrr

using namespace ofrr;

extern "C" int ofrr_spmm_csr(ofrr_ctx *ctx, int64_t n, int64_t nnz, const int64_t *rp,
                             const int32_t *ci, const void *va, int vdt, int k, const void *X,
                             int xdt, int64_t xrs, int64_t xcs, double alpha, double beta,
                             void *Y, int ydt, int64_t yrs, int64_t ycs) {
   OFRR_CHECK_CTX(ctx);
   if (n < 0 || k < 0 || nnz < 0) return set_error(ctx, kErrorInvalidArgument, "spmm_csr: negative size");
   if (n == 0 || k == 0) return kSuccess;
   if (!rp || !ci || !va || !X || !Y) return set_error(ctx, kErrorInvalidArgument, "spmm_csr: null pointer");
   if (xrs < 1 && xcs < 1) return set_error(ctx, kErrorInvalidArgument, "spmm_csr: bad X strides");
   return spmm_csr(ctx, n, rp, ci, va, vdt, k, X, xdt, xrs, xcs, alpha, beta, Y, ydt, yrs, ycs, 0);
}
