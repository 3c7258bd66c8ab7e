// Hessenberg process on the GPU: a pivoted, max-abs scaled, LU-on-columns
// basis builder with no orthogonalization.  Semantics of
// MATLAB/+src/+qr/hessenberg.m:55-74 (and of the C++ variants in
// CPP/linalg/factorization/hessenberg.cu:405-604): column i is eliminated
// against every earlier kept column j at its pivot row perm(j) with the
// current value R(j,i) = v(perm(j)), then divided by its signed max-abs entry
// R(i,i) (first index on ties); a column whose pivot is below select_tol is
// dropped (skip = 1) and keeps its unscaled residual.
//
// All decisions stay on the device (the reference copies the skip flag to the
// host every column, hessenberg.cu:519); the columns run in panels (below).
#include "internal.hpp"

#include <algorithm>

namespace ofrr {

// Panel-blocked formulation.  The columns are processed in panels of kPanel;
// every element Q(r, c) still receives the eliminations of the columns j < c
// in increasing j, each rounded to the storage format, with the coefficient
// R(j, c) = Q(perm_j, c) after the eliminations of the columns before j --
// the arithmetic of the right-looking process, so results are bitwise those of
// the column-at-a-time form (and of the oracle in FP64).
//   * column kernel (one launch per column i of panel [j0, j1)): every row
//     stores the scaled previous panel column, applies the eliminations of
//     the panel columns j0..i-1 to column i and takes part in its argmax; the
//     last workgroup to finish (an atomic count, no waiting) decides the pivot
//     and forms R(i, c) for the panel columns c in (i, j1) by replaying the
//     panel's eliminations at the pivot row;
//   * coefficient kernel (per panel): R(j, c) for the panel columns j and the
//     trailing columns c, by the same replay at each panel pivot row;
//   * trailing kernel (per panel): the panel's eliminations of every trailing
//     column, one row per thread, the panel's scaled columns in registers.
// The trailing matrix is read and written once per panel instead of once per
// column, and a column costs one launch instead of four.
constexpr int kPanel = 16;
constexpr int kHessCols = 16;   // trailing columns per workgroup
constexpr int kColRows = 256;   // rows per workgroup of the column kernel

struct HAMax {
   double a;   // |x|
   int64_t i;  // row (0-based), INT64_MAX when empty
   double v;   // signed value
};
__device__ inline HAMax hamax_better(const HAMax &p, const HAMax &q) {
   // larger |x| wins; equal |x| -> lower row (the argmax of the column-at-a-time form)
   if (q.a > p.a || (q.a == p.a && q.i < p.i)) return q;
   return p;
}
// block reduction (blockDim.x a multiple of 64, at most 1024); valid in thread 0
__device__ inline HAMax hamax_block(HAMax m) {
   __shared__ double sa[16], sv[16];
   __shared__ int64_t si[16];
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) {
      HAMax q;
      q.a = __shfl_xor(m.a, o, 64);
      q.i = __shfl_xor(m.i, o, 64);
      q.v = __shfl_xor(m.v, o, 64);
      m = hamax_better(m, q);
   }
   const int w = threadIdx.x / 64, l = threadIdx.x % 64;
   if (l == 0) { sa[w] = m.a; si[w] = m.i; sv[w] = m.v; }
   __syncthreads();
   if (threadIdx.x == 0)
      for (int u = 1; u < (int)(blockDim.x / 64); ++u) m = hamax_better(m, HAMax{sa[u], si[u], sv[u]});
   __syncthreads();
   return m;
}

// x rounded to the storage format, back in the accumulation type
template <typename T, typename ACC>
__device__ inline ACC hround(ACC x) {
   return (ACC)Elem<T>::load(Elem<T>::store((double)x));
}

// The pivot of a column from the argmax partials of the launch that formed
// it, decided by every workgroup alike (the same partials in the same order);
// workgroup 0 records it (R(i, i), skip, perm, prow).
struct HPivot {
   int64_t row;
   double val;
   int drop;
};
__device__ inline HPivot hess_finish(int i, int nparts, const HAMax *__restrict__ parts, double tol,
                                     double *__restrict__ R, int ldr, int32_t *__restrict__ perm,
                                     int32_t *__restrict__ skip, int32_t *__restrict__ prow) {
   __shared__ HPivot s_p;
   HAMax b{-1.0, INT64_MAX, 0.0};
   for (int u = threadIdx.x; u < nparts; u += blockDim.x) b = hamax_better(b, parts[u]);
   b = hamax_block(b);
   if (threadIdx.x == 0) {
      HPivot p;
      p.row = b.i == INT64_MAX ? 0 : b.i;
      p.val = b.i == INT64_MAX ? 0.0 : b.v;
      p.drop = !(fabs(p.val) >= tol) ? 1 : 0;
      s_p = p;
      if (blockIdx.x == 0 && blockIdx.y == 0) {
         R[i + (int64_t)i * ldr] = p.val;
         skip[i] = p.drop;
         perm[i] = p.drop ? -1 : (int32_t)p.row;
         prow[i] = (int32_t)p.row;
      }
   }
   __syncthreads();
   return s_p;
}

// Column launch i of panel [j0, j1): every workgroup first finishes column
// i - 1 (i > j0) -- its pivot from the previous launch's partials, and the
// coefficient R(i - 1, i) by replaying the panel's eliminations at the pivot
// row (workgroup 0 records R(i - 1, c) for every panel column c > i - 1) --
// then each row stores column i - 1 into Q (scaled when kept), applies the
// eliminations of the panel columns j0..i-1 to column i, and the workgroup
// writes its argmax partial of column i.  The eliminated column i goes to
// col_cur, not to Q: the replays of this launch read Q(pivot row, i) before
// the eliminations, so Q keeps it until the next launch.  No workgroup waits
// for another: what a launch needs from the previous one is read after the
// kernel boundary.
template <typename T, typename ACC>
__global__ __launch_bounds__(kColRows) void hess_col_kernel(int64_t m, int i, int j0, int j1, T *__restrict__ Q,
                                                            int64_t rs, int64_t cs, double tol,
                                                            double *__restrict__ R, int ldr,
                                                            int32_t *__restrict__ perm, int32_t *__restrict__ skip,
                                                            int32_t *__restrict__ prow,
                                                            const HAMax *__restrict__ part_prev,
                                                            HAMax *__restrict__ part_cur,
                                                            const T *__restrict__ col_prev,
                                                            T *__restrict__ col_cur, int variant) {
   __shared__ double s_row_q[kPanel], s_coef_i;
   const int64_t r = (int64_t)blockIdx.x * kColRows + threadIdx.x;
   const bool in = r < m;
   // this row's operands first: they do not depend on the pivot
   ACC qv[kPanel];
#pragma unroll
   for (int jt = 0; jt < kPanel; ++jt) {
      const int j = j0 + jt;
      qv[jt] = (in && j < i - 1) ? (ACC)Elem<T>::load(Q[r * rs + (int64_t)j * cs]) : (ACC)0;
   }
   const ACC vprev = (in && i > 0) ? (ACC)Elem<T>::load(col_prev[r]) : (ACC)0;
   ACC v = in ? (ACC)Elem<T>::load(Q[r * rs + (int64_t)i * cs]) : (ACC)0;
   double piv_p = 0.0;
   int drop_p = 1;
   if (i > j0) {
      const HPivot pv = hess_finish(i - 1, (int)gridDim.x, part_prev, tol, R, ldr, perm, skip, prow);
      piv_p = pv.val;
      drop_p = pv.drop;
      const bool gather = !pv.drop || variant == 1;
      const int np = i - 1 - j0;  // panel columns before i - 1
      if (gather)
         for (int t = threadIdx.x; t < j1 - j0; t += blockDim.x)
            s_row_q[t] = (double)Elem<T>::load(Q[pv.row * rs + (int64_t)(j0 + t) * cs]);
      __syncthreads();
      // R(i - 1, c), c in [i, j1): workgroup 0 all, the others c = i only
      const int ncoef = (blockIdx.x == 0) ? j1 - i : 1;
      for (int t = threadIdx.x; t < ncoef; t += blockDim.x) {
         const int c = i + t;
         double coef = 0.0;
         if (gather) {
            ACC val = (ACC)s_row_q[c - j0];
            for (int jt = 0; jt < np; ++jt) {
               if (variant == 0 && skip[j0 + jt]) continue;
               val = hround<T, ACC>(fma(-(ACC)R[j0 + jt + (int64_t)c * ldr], (ACC)s_row_q[jt], val));
            }
            coef = (double)val;
         }
         if (blockIdx.x == 0) R[i - 1 + (int64_t)c * ldr] = coef;
         if (t == 0) s_coef_i = coef;
      }
      __syncthreads();
   } else if (i > 0) {
      piv_p = R[i - 1 + (int64_t)(i - 1) * ldr];
      drop_p = skip[i - 1];
   }
   HAMax best{-1.0, INT64_MAX, 0.0};
   if (in) {
      // column i - 1 is stored scaled here: the launches before this one read
      // it unscaled (and scale it themselves)
      ACC qprev = vprev;
      if (i > 0) {
         if (!drop_p) qprev = hround<T, ACC>(vprev / (ACC)piv_p);
         Q[r * rs + (int64_t)(i - 1) * cs] = Elem<T>::store((double)qprev);
      }
#pragma unroll
      for (int jt = 0; jt < kPanel; ++jt) {
         const int j = j0 + jt;
         if (j < i - 1 && !(variant == 0 && skip[j]))
            v = hround<T, ACC>(fma(-(ACC)R[j + (int64_t)i * ldr], qv[jt], v));
      }
      if (i > j0 && !(variant == 0 && drop_p)) v = hround<T, ACC>(fma(-(ACC)s_coef_i, qprev, v));
      col_cur[r] = Elem<T>::store((double)v);
      best = HAMax{fabs((double)v), r, (double)v};
   }
   best = hamax_block(best);
   if (threadIdx.x == 0) part_cur[blockIdx.x] = best;
}

// entry of panel column j at row r as the eliminations use it: scaled and
// rounded when kept (stored that way once `stored`), unscaled when dropped
template <typename T, typename ACC>
__device__ inline ACC hess_q(const T *Q, int64_t rs, int64_t cs, int64_t r, int j, bool stored,
                             double piv, int drop) {
   const ACC v = (ACC)Elem<T>::load(Q[r * rs + (int64_t)j * cs]);
   if (stored || drop) return v;
   return hround<T, ACC>(v / (ACC)piv);
}

// Per panel: every workgroup finishes the last panel column (its pivot from
// the last column launch), then R(j, c) for the panel columns j and the
// trailing columns c >= j1 -- the value of column c at the pivot row of j
// after the eliminations before j.
template <typename T, typename ACC>
__global__ __launch_bounds__(64) void hess_coef_kernel(int n, int j0, int j1, const T *__restrict__ Q,
                                                       int64_t rs, int64_t cs, double tol, double *__restrict__ R,
                                                       int ldr, int32_t *__restrict__ perm,
                                                       int32_t *__restrict__ skip, int32_t *__restrict__ prow,
                                                       const HAMax *__restrict__ part_last, int nparts,
                                                       int variant) {
   const int nb = j1 - j0;
   const HPivot pl = hess_finish(j1 - 1, nparts, part_last, tol, R, ldr, perm, skip, prow);
   auto row_of = [&](int t) -> int64_t { return (t == nb - 1) ? pl.row : (int64_t)prow[j0 + t]; };
   auto drop_of = [&](int t) -> int { return (t == nb - 1) ? pl.drop : skip[j0 + t]; };
   // P(t, j) = panel column j at the pivot row of panel column t, as the
   // eliminations use it (the same for every trailing column)
   __shared__ ACC P[kPanel * kPanel];
   for (int u = threadIdx.x; u < nb * nb; u += blockDim.x) {
      const int t = u % nb, j = u / nb;
      if (j < t && !(variant == 0 && drop_of(t)))
         P[t + j * kPanel] = hess_q<T, ACC>(Q, rs, cs, row_of(t), j0 + j, true, 0.0, skip[j0 + j]);
   }
   __syncthreads();
   const int c = j1 + blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= n) return;
   ACC val[kPanel];
#pragma unroll
   for (int t = 0; t < kPanel; ++t)
      val[t] = (t < nb && !(variant == 0 && drop_of(t)))
                   ? (ACC)Elem<T>::load(Q[row_of(t) * rs + (int64_t)c * cs])
                   : (ACC)0;
#pragma unroll
   for (int jt = 0; jt < kPanel; ++jt) {
      if (jt >= nb) break;
      const int j = j0 + jt;
      const bool use = variant == 1 || !drop_of(jt);
      const ACC coef = use ? val[jt] : (ACC)0;
      R[j + (int64_t)c * ldr] = (double)coef;
      if (!use) continue;
#pragma unroll
      for (int t = jt + 1; t < kPanel; ++t) {
         if (t >= nb || (variant == 0 && drop_of(t))) continue;
         val[t] = hround<T, ACC>(fma(-coef, P[t + jt * kPanel], val[t]));
      }
   }
}

// the panel's eliminations of the trailing columns [cb, ce) (a group of
// kHessCols per workgroup, one row per thread).  The last panel column is
// still in its column buffer (the next column launch stores it into Q): the
// launch for the next panel's columns (side_in = 0) scales it from there and
// keeps a copy in `side`; the launch for the rest of the columns (side_in = 1),
// which runs beside the next panel's column launches, reads that copy.
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_trail_kernel(int64_t m, int cb, int ce, int j0, int j1,
                                                         T *__restrict__ Q, int64_t rs, int64_t cs,
                                                         const double *__restrict__ R, int ldr,
                                                         const int32_t *__restrict__ skip, int variant,
                                                         const T *__restrict__ col_last,
                                                         T *__restrict__ side, int side_in) {
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= m) return;
   const int nb = j1 - j0;
   ACC q[kPanel], qlast;
#pragma unroll
   for (int t = 0; t < kPanel; ++t)
      q[t] = (t < nb - 1) ? (ACC)Elem<T>::load(Q[r * rs + (int64_t)(j0 + t) * cs]) : (ACC)0;
   if (side_in) {
      qlast = (ACC)Elem<T>::load(side[r]);
   } else {
      const ACC v = (ACC)Elem<T>::load(col_last[r]);
      qlast = skip[j1 - 1] ? v : hround<T, ACC>(v / (ACC)R[j1 - 1 + (int64_t)(j1 - 1) * ldr]);
      if (blockIdx.y == 0) side[r] = Elem<T>::store((double)qlast);
   }
   const int c0 = cb + (int)blockIdx.y * kHessCols, c1 = min(ce, c0 + kHessCols);
   // the group's columns loaded together; each element still receives the
   // panel's eliminations in column order
   ACC v[kHessCols];
#pragma unroll
   for (int u = 0; u < kHessCols; ++u)
      v[u] = (c0 + u < c1) ? (ACC)Elem<T>::load(Q[r * rs + (int64_t)(c0 + u) * cs]) : (ACC)0;
#pragma unroll
   for (int t = 0; t < kPanel; ++t) {
      if (t >= nb || (variant == 0 && skip[j0 + t])) continue;
      const ACC qt = (t == nb - 1) ? qlast : q[t];
#pragma unroll
      for (int u = 0; u < kHessCols; ++u)
         if (c0 + u < c1) v[u] = hround<T, ACC>(fma(-(ACC)R[j0 + t + (int64_t)(c0 + u) * ldr], qt, v[u]));
   }
#pragma unroll
   for (int u = 0; u < kHessCols; ++u)
      if (c0 + u < c1) Q[r * rs + (int64_t)(c0 + u) * cs] = Elem<T>::store((double)v[u]);
}

// the last column: its pivot from the last column launch; into Q from its
// column buffer, scaled when kept
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_last_kernel(int64_t m, int i, T *__restrict__ Q, int64_t rs,
                                                        int64_t cs, double tol, double *__restrict__ R, int ldr,
                                                        int32_t *__restrict__ perm, int32_t *__restrict__ skip,
                                                        int32_t *__restrict__ prow,
                                                        const HAMax *__restrict__ part_last, int nparts,
                                                        const T *__restrict__ col_last) {
   const HPivot pv = hess_finish(i, nparts, part_last, tol, R, ldr, perm, skip, prow);
   for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += (int64_t)gridDim.x * blockDim.x) {
      const ACC v = (ACC)Elem<T>::load(col_last[r]);
      Q[r * rs + (int64_t)i * cs] = Elem<T>::store((double)(pv.drop ? v : hround<T, ACC>(v / (ACC)pv.val)));
   }
}

__global__ void zero_lower_kernel(int n, double *__restrict__ R, int ldr) {
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n * n) return;
   int r = t % n, c = t / n;
   if (r > c) R[r + (int64_t)c * ldr] = 0.0;
}

// Panel by panel on the context's stream; each panel's trailing update is
// split: the next panel's columns first (on the stream), the rest on the
// context's second stream, where it overlaps the next panel's column launches
// (which touch neither those columns nor the panel's) and is joined before the
// next panel's coefficients read them.
template <typename T, typename ACC>
static int hessenberg_impl(ofrr_ctx *ctx, int64_t m, int n, T *Q, int64_t rs, int64_t cs,
                           double tol, double *R, int ldr, int32_t *perm, int32_t *skip,
                           int variant) {
   const int cblocks = (int)((m + kColRows - 1) / kColRows);
   // scratch: two sets of argmax partials and two column buffers (consecutive
   // column launches) | the argmax row of every column | the scaled last column
   // of the current panel
   const size_t pbytes = ((size_t)cblocks * sizeof(HAMax) + 255) & ~(size_t)255;
   const size_t rbytes = ((size_t)n * sizeof(int32_t) + 255) & ~(size_t)255;
   const size_t cbytes = ((size_t)m * sizeof(T) + 255) & ~(size_t)255;
   char *scr = (char *)workspace(ctx, 2 * pbytes + rbytes + 3 * cbytes);
   if (!scr) return set_error(ctx, kErrorCUDAMemory, "hessenberg: workspace");
   HAMax *part[2] = {(HAMax *)scr, (HAMax *)(scr + pbytes)};
   int32_t *prow = (int32_t *)(scr + 2 * pbytes);
   T *colb[2] = {(T *)(scr + 2 * pbytes + rbytes), (T *)(scr + 2 * pbytes + rbytes + cbytes)};
   T *side = (T *)(scr + 2 * pbytes + rbytes + 2 * cbytes);
   if (n > 2 * kPanel) {
      if (!ctx->aux_stream) OFRR_HIP(ctx, hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
      if (!ctx->ev_aux_fork) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_aux_fork, hipEventDisableTiming));
      if (!ctx->ev_aux_join) OFRR_HIP(ctx, hipEventCreateWithFlags(&ctx->ev_aux_join, hipEventDisableTiming));
   }
   zero_lower_kernel<<<(n * n + 255) / 256, 256, 0, ctx->stream>>>(n, R, ldr);
   OFRR_LAUNCH(ctx);
   const int ublocks = (int)((m + 255) / 256);
   bool pending = false;  // a trailing update running on the second stream
   for (int j0 = 0; j0 < n; j0 += kPanel) {
      const int j1 = std::min(n, j0 + kPanel), j2 = std::min(n, j1 + kPanel);
      for (int i = j0; i < j1; ++i)
         hess_col_kernel<T, ACC><<<cblocks, kColRows, 0, ctx->stream>>>(m, i, j0, j1, Q, rs, cs, tol, R, ldr,
                                                                        perm, skip, prow, part[(i + 1) & 1],
                                                                        part[i & 1], colb[(i + 1) & 1],
                                                                        colb[i & 1], variant);
      OFRR_LAUNCH(ctx);
      if (pending) {
         OFRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->ev_aux_join, 0));
         pending = false;
      }
      if (j1 >= n) break;
      hess_coef_kernel<T, ACC><<<(n - j1 + 63) / 64, 64, 0, ctx->stream>>>(
          n, j0, j1, Q, rs, cs, tol, R, ldr, perm, skip, prow, part[(j1 - 1) & 1], cblocks, variant);
      hess_trail_kernel<T, ACC><<<dim3(ublocks, (j2 - j1 + kHessCols - 1) / kHessCols), 256, 0, ctx->stream>>>(
          m, j1, j2, j0, j1, Q, rs, cs, R, ldr, skip, variant, colb[(j1 - 1) & 1], side, 0);
      OFRR_LAUNCH(ctx);
      if (j2 < n) {
         OFRR_HIP(ctx, hipEventRecord(ctx->ev_aux_fork, ctx->stream));
         OFRR_HIP(ctx, hipStreamWaitEvent(ctx->aux_stream, ctx->ev_aux_fork, 0));
         hess_trail_kernel<T, ACC><<<dim3(ublocks, (n - j2 + kHessCols - 1) / kHessCols), 256, 0,
                                     ctx->aux_stream>>>(m, j2, n, j0, j1, Q, rs, cs, R, ldr, skip, variant,
                                                        colb[(j1 - 1) & 1], side, 1);
         OFRR_LAUNCH(ctx);
         OFRR_HIP(ctx, hipEventRecord(ctx->ev_aux_join, ctx->aux_stream));
         pending = true;
      }
   }
   hess_last_kernel<T, ACC><<<std::min(ublocks, 1024), 256, 0, ctx->stream>>>(
       m, n - 1, Q, rs, cs, tol, R, ldr, perm, skip, prow, part[(n - 1) & 1], cblocks, colb[(n - 1) & 1]);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---- one column at a time (hessenberg_v3_step, hessenberg.cu:617-801) ------
// Column i against the finished columns 0..i-1, left-looking.  Its coefficient
// against column j is its entry at j's pivot row after the eliminations with
// 0..j-1, so all of them follow from the column's entries at the pivot rows by
// a forward substitution over the pivot rows of the finished columns (one
// workgroup, the same rounded fused multiply-adds as the full elimination);
// then every row applies the eliminations in column order in one pass.  The
// result is bit for bit that of the right-looking factorization.
template <typename T, typename ACC>
__global__ __launch_bounds__(1024) void hess_step_coef_kernel(int i, const T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, const int32_t *__restrict__ prow1,
                                                              const int32_t *__restrict__ skip, int variant,
                                                              double *__restrict__ R, int ldr) {
   extern __shared__ double val[];  // i entries: column i at the pivot rows, being eliminated
   for (int t = threadIdx.x; t < i; t += blockDim.x)
      val[t] = (double)Elem<T>::load(Q[(int64_t)(prow1[t] - 1) * rs + (int64_t)i * cs]);
   __syncthreads();
   for (int j = 0; j < i; ++j) {
      const bool use = variant == 1 || !skip[j];
      const double c = use ? val[j] : 0.0;
      if (threadIdx.x == 0) R[j + (int64_t)i * ldr] = c;
      if (use)
         for (int t = j + 1 + threadIdx.x; t < i; t += blockDim.x) {
            const ACC q = (ACC)Elem<T>::load(Q[(int64_t)(prow1[t] - 1) * rs + (int64_t)j * cs]);
            val[t] = (double)Elem<T>::load(Elem<T>::store((double)fma(-(ACC)c, q, (ACC)val[t])));
         }
      __syncthreads();
   }
}

// every row of column i: the eliminations with the columns j < i in order
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_step_apply_kernel(int64_t m, int i, T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, const int32_t *__restrict__ skip,
                                                              int variant, const double *__restrict__ R,
                                                              int ldr) {
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r >= m) return;
   T *p = Q + r * rs + (int64_t)i * cs;
   ACC v = (ACC)Elem<T>::load(*p);
   for (int j = 0; j < i; ++j) {
      if (variant == 0 && skip[j]) continue;
      const ACC c = (ACC)R[j + (int64_t)i * ldr];
      v = (ACC)Elem<T>::load(Elem<T>::store((double)fma(-c, (ACC)Elem<T>::load(Q[r * rs + (int64_t)j * cs]), v)));
   }
   *p = Elem<T>::store((double)v);
}

// pivot of column i (argmax written by argmax_columns to prow1[i] (1-based)
// and R(i, i)): drop or scale
template <typename T, typename ACC>
__global__ __launch_bounds__(256) void hess_step_scale_kernel(int64_t m, int i, T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, double select_tol,
                                                              const double *__restrict__ R, int ldr,
                                                              int32_t *__restrict__ skip) {
   const double piv = R[i + (int64_t)i * ldr];
   const bool drop = !(fabs(piv) >= select_tol);
   const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (r == 0) skip[i] = drop ? 1 : 0;
   if (drop || r >= m) return;
   T *p = Q + r * rs + (int64_t)i * cs;
   *p = Elem<T>::store((double)((ACC)Elem<T>::load(*p) / (ACC)piv));
}

template <typename T, typename ACC>
static int hessenberg_step_impl(ofrr_ctx *ctx, int variant, int64_t m, int i, T *Q, int64_t rs, int64_t cs,
                                double tol, double *R, int ldr, int32_t *prow1, int32_t *skip) {
   if (i > 0) {
      hess_step_coef_kernel<T, ACC><<<1, 1024, (size_t)i * sizeof(double), ctx->stream>>>(
          i, Q, rs, cs, prow1, skip, variant, R, ldr);
      hess_step_apply_kernel<T, ACC><<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(
          m, i, Q, rs, cs, skip, variant, R, ldr);
      OFRR_LAUNCH(ctx);
   }
   int st = argmax_columns(ctx, Elem<T>::id, m, 1, Q + (int64_t)i * cs, rs, 0, 1, prow1 + i,
                           R + i + (int64_t)i * ldr);
   if (st) return st;
   hess_step_scale_kernel<T, ACC><<<(unsigned)((m + 255) / 256), 256, 0, ctx->stream>>>(
       m, i, Q, rs, cs, tol, R, ldr, skip);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

}  // namespace ofrr

using namespace ofrr;

static int hessenberg_variant(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q, int64_t rs,
                              int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                              int32_t *skip) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n <= 0 || !Q || !R || !perm || !skip || ldr < n)
      return set_error(ctx, kErrorInvalidArgument, "hessenberg: bad arguments");
   if (m < n) return set_error(ctx, kErrorInvalidArgument, "hessenberg: A must be a tall-thin matrix");
   switch (qdt) {
      case OFRR_F64:
         return hessenberg_impl<double, double>(ctx, m, n, (double *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_F32:
         return hessenberg_impl<float, float>(ctx, m, n, (float *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_F16:
         return hessenberg_impl<__half, float>(ctx, m, n, (__half *)Q, rs, cs, select_tol, R, ldr, perm, skip, variant);
      case OFRR_BF16:
         return hessenberg_impl<__hip_bfloat16, float>(ctx, m, n, (__hip_bfloat16 *)Q, rs, cs, select_tol, R, ldr,
                                                       perm, skip, variant);
      default:
         return set_error(ctx, kErrorBLASUnsupportedType, "hessenberg: unsupported dtype");
   }
}

extern "C" int ofrr_hessenberg(ofrr_ctx *ctx, int qdt, int64_t m, int n, void *Q, int64_t rs,
                               int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                               int32_t *skip) {
   return hessenberg_variant(ctx, 0, qdt, m, n, Q, rs, cs, select_tol, R, ldr, perm, skip);
}

extern "C" int ofrr_hessenberg_v3(ofrr_ctx *ctx, int qdt, int64_t m, int n, void *Q, int64_t rs,
                                  int64_t cs, double select_tol, double *R, int ldr, int32_t *perm,
                                  int32_t *skip) {
   return hessenberg_variant(ctx, 1, qdt, m, n, Q, rs, cs, select_tol, R, ldr, perm, skip);
}

extern "C" int ofrr_hessenberg_step(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q,
                                    int64_t rs, int64_t cs, int i, double select_tol, double *R, int ldr,
                                    int32_t *prow1, int32_t *skip, int32_t *skipped_host) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || n <= 0 || i < 0 || i >= n || !Q || !R || !prow1 || !skip || ldr < n ||
       (variant != 0 && variant != 1))
      return set_error(ctx, kErrorInvalidArgument, "hessenberg_step: bad arguments");
   int st;
   switch (qdt) {
      case OFRR_F64:
         st = hessenberg_step_impl<double, double>(ctx, variant, m, i, (double *)Q, rs, cs, select_tol, R, ldr,
                                                   prow1, skip);
         break;
      case OFRR_F32:
         st = hessenberg_step_impl<float, float>(ctx, variant, m, i, (float *)Q, rs, cs, select_tol, R, ldr,
                                                 prow1, skip);
         break;
      case OFRR_F16:
         st = hessenberg_step_impl<__half, float>(ctx, variant, m, i, (__half *)Q, rs, cs, select_tol, R, ldr,
                                                  prow1, skip);
         break;
      case OFRR_BF16:
         st = hessenberg_step_impl<__hip_bfloat16, float>(ctx, variant, m, i, (__hip_bfloat16 *)Q, rs, cs,
                                                          select_tol, R, ldr, prow1, skip);
         break;
      default:
         return set_error(ctx, kErrorBLASUnsupportedType, "hessenberg_step: unsupported dtype");
   }
   if (st) return st;
   if (skipped_host) {
      int32_t *h = (int32_t *)pinned(ctx, sizeof(int32_t));
      if (!h) return set_error(ctx, kErrorCUDAMemory, "hessenberg_step: pinned");
      OFRR_HIP(ctx, hipMemcpyAsync(h, skip + i, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      *skipped_host = *h;
   }
   return kSuccess;
}
