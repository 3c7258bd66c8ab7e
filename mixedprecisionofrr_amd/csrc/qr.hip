// Gram-Schmidt QR on the GPU: the orthogonalizing baselines the projection
// pipeline is compared against.  Semantics of CPP/linalg/factorization/qr.cu
// (mgs :12-81 with mgs_step :84-249, mgs_v2 :251-393, cgs :395-569,
// cgs2 :571-754) and of
// MATLAB/+src/+qr/mgs.m:55-92, cgs.m:48-95:
//   MGS     column i is projected against the kept columns j < i one at a time
//           (r = q_j^T v, v -= r q_j) and reorthogonalized the same way while
//           orth_tol <= ||v|| < reorth_tol * normv (normv: the norm before the
//           pass; reorth_tol <= 0 turns this off); a column whose initial or
//           final norm is below select_tol is dropped with R(i,i) = 0;
//   MGS_V2  right-looking MGS: column i is normalized when its turn comes
//           (dropped if ||v|| < select_tol, R(i,i) = 1) and removed from every
//           later column;
//   CGS     one product r = Q_kept^T v and v -= Q_kept r per pass (dropped on
//           the initial or final norm, R(i,i) = 1); with reorth_tol > 0 more
//           passes under the MGS criterion (cgs.m:73-89);
//   CGS2    two such passes, R(:,i) = r1 + r2 (qr.hpp:110-131; qr.cu:713-717
//           adds r2 to R without subtracting Q r2 from v, so its Q is the CGS
//           one: here the second projection is removed).
// A dropped column keeps its unnormalized residual in Q and takes no part in
// later projections (mgs.m:66, cgs.m:60; the C++ cgs also projects against
// dropped columns, qr.cu:523).
//
// Every decision (drop, accept, reorthogonalize) is taken on the device: a
// factorization is a fixed sequence of launches per column with no host round
// trip (the reference reads every norm and dot product back to the host).  A
// reorthogonalization loop reads one flag per column.  Rows are split into
// chunks of 256..4096 rows, one workgroup per chunk; every reduction over rows
// goes through per-chunk partials summed in a fixed order, so results are
// deterministic.  Dot products and norms accumulate in FP64; updates are
// computed in the format's compute type and rounded to the storage format.
//
// MGS runs right-looking: once column j is final, one launch forms q_j^T Q(:,c)
// for every later column c and one rank-1 launch removes them -- per column the
// same operations in the same order as the left-looking loop.  A sequential
// MGS pass over one column (reorthogonalization, the step entry point) is a
// chain of launches: each finishes the previous partial dot in every
// workgroup (fixed order), applies that axpy to its rows and forms the next
// partial dot.
#include "internal.hpp"

#include <algorithm>
#include <vector>

namespace ofrr {
namespace {

constexpr int kQrThreads = 256;
constexpr int kQrWaves = kQrThreads / 64;
constexpr int kQrCols = 16;  // columns per workgroup in the dot and rank-1 launches
// reorthogonalization passes per column at most (each must shrink ||v|| by
// reorth_tol, so more than a few means a degenerate column)
constexpr int kQrMaxPasses = 64;

struct QrLayout {
   int64_t chunk;  // rows per workgroup
   int nblk;       // row chunks
};

QrLayout qr_layout(int64_t m) {
   int64_t per = (m + 65535) / 65536;  // about 256 chunks, 256..4096 rows each
   per = std::min<int64_t>(std::max<int64_t>(per, 1), 16);
   QrLayout L;
   L.chunk = 256 * per;
   L.nblk = (int)((m + L.chunk - 1) / L.chunk);
   return L;
}

template <typename T>
__device__ inline T *qat(T *Q, int64_t rs, int64_t cs, int64_t r, int c) {
   return Q + r * rs + (int64_t)c * cs;
}

// sum of p[0..cnt) by one wave in a fixed order (lane 0's value on every lane)
__device__ inline double wave_sum_fixed(const double *p, int cnt) {
   const int lane = threadIdx.x & 63;
   double s = 0.0;
   for (int t = lane; t < cnt; t += 64) s += p[t];
   s = wave_sum(s);
   return __shfl(s, 0, 64);
}

// workgroup sum, valid in thread 0 (one call per kernel)
__device__ inline double block_sum(double v, double *red) {
   v = wave_sum(v);
   if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
   __syncthreads();
   double s = 0.0;
   if (threadIdx.x == 0)
      for (int i = 0; i < kQrWaves; ++i) s += red[i];
   return s;
}

__global__ void qr_init_kernel(int n, int c0, int c1, double *R, int ldr, int32_t *skip) {
   // zero R(:, c0..c1) rows 0..c1 and skip[c0..c1)
   const int64_t rows = c1, cols = c1 - c0;
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < rows * cols) R[t % rows + (c0 + t / rows) * (int64_t)ldr] = 0.0;
   if (t < cols) skip[c0 + t] = 0;
   (void)n;
}

// part[c * nblk + b] = sum over chunk b of Q(r, c)^2, c in [c0, c1)
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_colnorm_kernel(int64_t m, int64_t chunk, int nblk,
                                                                const T *__restrict__ Q, int64_t rs,
                                                                int64_t cs, int c0, int c1,
                                                                double *__restrict__ part) {
   const int b = blockIdx.x;
   const int64_t r0 = (int64_t)b * chunk, r1 = min(m, r0 + chunk);
   const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
   const int cb = c0 + blockIdx.y * kQrCols, ce = min(c1, cb + kQrCols);
   for (int c = cb + w; c < ce; c += kQrWaves) {
      double s = 0.0;
      for (int64_t r = r0 + lane; r < r1; r += 64) {
         const double x = ld_d(qat(Q, rs, cs, r, c));
         s = fma(x, x, s);
      }
      s = wave_sum(s);
      if (lane == 0) part[(int64_t)c * nblk + b] = s;
   }
}

// normv[c] = ||Q(:, c)||, pre[c] = normv[c] < select_tol (one wave per column)
__global__ void qr_norms_kernel(int n, int nblk, const double *__restrict__ part, double stol,
                                double *__restrict__ normv, int32_t *__restrict__ pre) {
   const int c = blockIdx.x * kQrWaves + (threadIdx.x >> 6);
   if (c >= n) return;
   const double t = sqrt(wave_sum_fixed(part + (int64_t)c * nblk, nblk));
   if ((threadIdx.x & 63) == 0) {
      normv[c] = t;
      pre[c] = (t < stol) ? 1 : 0;
   }
}

// part[c * nblk + b] = sum over chunk b of Q(r, c) v(r) for the columns c in
// [c0, c1) not excluded, v = Q(:, vcol), times *vscale when vscale is given
// (the right-looking pivot: the value its scale launch stores).  pcol: a column
// whose normalization is still pending (left-looking), scaled and stored here
// by the one wave that owns it.  vnorm_part[b]: partial ||v||^2 (nullable).
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_dots_kernel(
    int64_t m, int64_t chunk, int nblk, T *__restrict__ Q, int64_t rs, int64_t cs, int vcol,
    const double *vscale, int c0, int c1, const int32_t *excl, int pcol, const double *pscale,
    const int32_t *gate, double *__restrict__ part, double *__restrict__ vnorm_part) {
   using ACC = typename Elem<T>::acc;
   if (gate && *gate) return;
   extern __shared__ double vs[];
   const int b = blockIdx.x;
   const int64_t r0 = (int64_t)b * chunk;
   const int rows = (int)min(chunk, m - r0);
   const ACC s = vscale ? (ACC)*vscale : (ACC)1;
   for (int t = threadIdx.x; t < rows; t += kQrThreads) {
      double x = ld_d(qat(Q, rs, cs, r0 + t, vcol));
      if (vscale) x = (double)Elem<T>::load(Elem<T>::store((double)((ACC)x * s)));
      vs[t] = x;
   }
   __syncthreads();
   const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
   if (vnorm_part && blockIdx.y == 0 && w == 0) {
      double a = 0.0;
      for (int t = lane; t < rows; t += 64) a = fma(vs[t], vs[t], a);
      a = wave_sum(a);
      if (lane == 0) vnorm_part[b] = a;
   }
   const int cb = c0 + blockIdx.y * kQrCols, ce = min(c1, cb + kQrCols);
   for (int c = cb + w; c < ce; c += kQrWaves) {
      if (excl && excl[c]) continue;
      double a = 0.0;
      if (c == pcol) {
         const ACC ps = (ACC)*pscale;
         for (int t = lane; t < rows; t += 64) {
            T *p = qat(Q, rs, cs, r0 + t, c);
            const T y = Elem<T>::store((double)((ACC)Elem<T>::load(*p) * ps));
            *p = y;
            a = fma((double)Elem<T>::load(y), vs[t], a);
         }
      } else {
         for (int t = lane; t < rows; t += 64) a = fma(ld_d(qat(Q, rs, cs, r0 + t, c)), vs[t], a);
      }
      a = wave_sum(a);
      if (lane == 0) part[(int64_t)c * nblk + b] = a;
   }
}

// Coefficients of the columns [c0, c1): coef[c] = the fixed-order sum of
// part[c * nblk + .], 0 for an excluded column; added to R(c, rfix) (left-
// looking) or stored in R(rfix, c) (right-looking).  pre_col >= 0: the first
// pass of column pre_col, whose norm partials are in vnorm_part: if its norm is
// below select_tol the column is dropped (R(pre_col, pre_col) = rd_skip) and
// nothing else is written, else normv[pre_col] = the norm.  Workgroups from
// colblocks on scale column scol by *sscale (the right-looking pivot).
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_reduce_kernel(
    int nblk, const double *__restrict__ part, int c0, int c1, int colblocks, const int32_t *excl,
    const int32_t *gate, int pre_col, const double *vnorm_part, double stol, double rd_skip,
    int32_t *skip, double *normv, double *__restrict__ coef, double *__restrict__ R, int ldr,
    int left, int rfix, int64_t m, T *__restrict__ Q, int64_t rs, int64_t cs, int scol,
    const double *sscale) {
   using ACC = typename Elem<T>::acc;
   if (gate && *gate) return;
   if ((int)blockIdx.x >= colblocks) {
      const ACC s = (ACC)*sscale;
      const int64_t stride = (int64_t)(gridDim.x - colblocks) * kQrThreads;
      for (int64_t r = (int64_t)(blockIdx.x - colblocks) * kQrThreads + threadIdx.x; r < m; r += stride) {
         T *p = qat(Q, rs, cs, r, scol);
         *p = Elem<T>::store((double)((ACC)Elem<T>::load(*p) * s));
      }
      return;
   }
   const int lane = threadIdx.x & 63;
   if (pre_col >= 0) {
      const double t = sqrt(wave_sum_fixed(vnorm_part, nblk));
      const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
      if (t < stol) {
         if (lead) {
            skip[pre_col] = 1;
            R[pre_col + (int64_t)pre_col * ldr] = rd_skip;
         }
         return;
      }
      if (lead) normv[pre_col] = t;
   }
   const int c = c0 + (int)blockIdx.x * kQrWaves + (threadIdx.x >> 6);
   if (c >= c1) return;
   if (excl && excl[c]) {
      if (lane == 0) coef[c] = 0.0;
      return;
   }
   const double v = wave_sum_fixed(part + (int64_t)c * nblk, nblk);
   if (lane == 0) {
      coef[c] = v;
      if (left)
         R[c + (int64_t)rfix * ldr] += v;
      else
         R[rfix + (int64_t)c * ldr] = v;
   }
}

// Decision for column col from its norm partials p (one wave).  flag[0] = 1:
// reorthogonalize (normv[col] <- ||v||; never for ||v|| = 0); else final: dropped (flag[1] = 1,
// R(col,col) = rd_skip) or accepted (R(col,col) = ||v||, scale[col] = 1/||v||).
__global__ void qr_decide_kernel(int nblk, const double *__restrict__ p, int col,
                                 const int32_t *pre, int32_t *skip, double *R, int ldr,
                                 double *normv, double *scale, int32_t *flag, double stol,
                                 double otol, double rtol, double rd_skip) {
   const bool lead = threadIdx.x == 0;
   if (skip[col]) {
      if (lead) { flag[0] = 0; flag[1] = 1; }
      return;
   }
   if (pre && pre[col]) {
      if (lead) {
         skip[col] = 1;
         R[col + (int64_t)col * ldr] = rd_skip;
         scale[col] = 0.0;
         flag[0] = 0;
         flag[1] = 1;
      }
      return;
   }
   const double t = sqrt(wave_sum_fixed(p, nblk));
   if (!lead) return;
   if (rtol > 0.0 && t > 0.0 && t >= otol && t < rtol * normv[col]) {
      normv[col] = t;
      flag[0] = 1;
      flag[1] = 0;
      return;
   }
   flag[0] = 0;
   if (t < stol) {
      skip[col] = 1;
      R[col + (int64_t)col * ldr] = rd_skip;
      scale[col] = 0.0;
      flag[1] = 1;
   } else {
      R[col + (int64_t)col * ldr] = t;
      scale[col] = 1.0 / t;
      flag[1] = 0;
   }
}

// right-looking update Q(:, c) -= q_j coef[c], c in [c0, c1), q_j = Q(:, j)
// normalized; a zero coefficient (excluded column) leaves the column alone.
// next_part[b]: partial ||Q(:, c0)||^2 after the update (also when column j
// was dropped, then with no update).
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_rank1_kernel(int64_t m, int64_t chunk,
                                                              T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, int j, int c0, int c1,
                                                              const double *__restrict__ coef,
                                                              const int32_t *gate,
                                                              double *__restrict__ next_part) {
   using ACC = typename Elem<T>::acc;
   __shared__ double red[kQrWaves];
   const bool dropped = *gate != 0;
   const int b = blockIdx.x;
   const int64_t r0 = (int64_t)b * chunk, r1 = min(m, r0 + chunk);
   const int cb = c0 + blockIdx.y * kQrCols, ce = min(c1, cb + kQrCols);
   const bool has_next = blockIdx.y == 0;
   if (dropped && !has_next) return;
   double a = 0.0;
   for (int64_t r = r0 + threadIdx.x; r < r1; r += kQrThreads) {
      if (!dropped) {
         const ACC q = (ACC)Elem<T>::load(*qat(Q, rs, cs, r, j));
         for (int c = cb; c < ce; ++c) {
            const double cf = coef[c];
            if (cf == 0.0) continue;
            T *p = qat(Q, rs, cs, r, c);
            *p = Elem<T>::store((double)fma(-(ACC)cf, q, (ACC)Elem<T>::load(*p)));
         }
      }
      if (has_next) {
         const double x = ld_d(qat(Q, rs, cs, r, c0));
         a = fma(x, x, a);
      }
   }
   if (has_next) {
      const double s = block_sum(a, red);
      if (threadIdx.x == 0) next_part[b] = s;
   }
}

// Left-looking update v = Q(:, vcol) -= sum_{c < c1} coef[c] Q(:, c) in two
// launches, so that the columns spread over workgroups as in the dot launch:
// qr_gemvn_part_kernel forms S[g][r] = sum of coef[c] Q(r, c) over the columns
// of group g (kQrGemvCols columns, blockIdx.y; zero coefficients skipped), and
// qr_gemvn_apply_kernel subtracts sum_g S[g][r] (fixed order) from v and forms
// vnorm_part[b] = partial ||v||^2 of the result.
constexpr int kQrGemvCols = 32;
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_gemvn_part_kernel(int64_t m, int64_t chunk,
                                                                   const T *__restrict__ Q,
                                                                   int64_t rs, int64_t cs, int c1,
                                                                   const double *__restrict__ coef,
                                                                   const int32_t *gate,
                                                                   double *__restrict__ S) {
   if (*gate) return;
   const int64_t r0 = (int64_t)blockIdx.x * chunk, r1 = min(m, r0 + chunk);
   const int g = blockIdx.y;
   const int cb = g * kQrGemvCols, ce = min(c1, cb + kQrGemvCols);
   double *Sg = S + (int64_t)g * m;
   for (int64_t r = r0 + threadIdx.x; r < r1; r += kQrThreads) {
      double s = 0.0;
      for (int c = cb; c < ce; ++c) {
         const double cf = coef[c];
         if (cf != 0.0) s = fma(cf, ld_d(qat(Q, rs, cs, r, c)), s);
      }
      Sg[r] = s;
   }
}

template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_gemvn_apply_kernel(int64_t m, int64_t chunk,
                                                                    T *__restrict__ Q, int64_t rs,
                                                                    int64_t cs, int vcol, int ngroups,
                                                                    const double *__restrict__ S,
                                                                    const int32_t *gate,
                                                                    double *__restrict__ vnorm_part) {
   using ACC = typename Elem<T>::acc;
   if (*gate) return;
   __shared__ double red[kQrWaves];
   const int b = blockIdx.x;
   const int64_t r0 = (int64_t)b * chunk, r1 = min(m, r0 + chunk);
   double a = 0.0;
   for (int64_t r = r0 + threadIdx.x; r < r1; r += kQrThreads) {
      double s = 0.0;
      for (int g = 0; g < ngroups; ++g) s += S[(int64_t)g * m + r];
      T *p = qat(Q, rs, cs, r, vcol);
      const T y = Elem<T>::store((double)((ACC)Elem<T>::load(*p) - (ACC)s));
      *p = y;
      const double yd = (double)Elem<T>::load(y);
      a = fma(yd, yd, a);
   }
   const double s = block_sum(a, red);
   if (threadIdx.x == 0) vnorm_part[b] = s;
}

// One step of a sequential MGS pass over v = Q(:, vcol): every workgroup
// finishes the partial dot of column lp (prev, fixed order); workgroup 0 adds
// it to R(lp, vcol); v -= r q_lp on the chunk; next[b] = the partial dot with
// column ln, or partial ||v||^2 when ln < 0.
template <typename T>
__global__ __launch_bounds__(kQrThreads) void qr_chain_kernel(int64_t m, int64_t chunk, int nblk,
                                                              T *__restrict__ Q, int64_t rs,
                                                              int64_t cs, int vcol, int lp, int ln,
                                                              const double *prev, double *next,
                                                              double *R, int ldr,
                                                              const int32_t *gate) {
   using ACC = typename Elem<T>::acc;
   if (gate && *gate) return;
   __shared__ double red[kQrWaves];
   __shared__ double rsh;
   if (lp >= 0) {
      if (threadIdx.x < 64) {
         const double r = wave_sum_fixed(prev, nblk);
         if (threadIdx.x == 0) {
            rsh = r;
            if (blockIdx.x == 0) R[lp + (int64_t)vcol * ldr] += r;
         }
      }
      __syncthreads();
   }
   const ACC rc = (lp >= 0) ? (ACC)rsh : (ACC)0;
   const int b = blockIdx.x;
   const int64_t r0 = (int64_t)b * chunk, r1 = min(m, r0 + chunk);
   double a = 0.0;
   for (int64_t r = r0 + threadIdx.x; r < r1; r += kQrThreads) {
      T *p = qat(Q, rs, cs, r, vcol);
      ACC x = (ACC)Elem<T>::load(*p);
      if (lp >= 0) {
         const T y = Elem<T>::store((double)fma(-rc, (ACC)Elem<T>::load(*qat(Q, rs, cs, r, lp)), x));
         *p = y;
         x = (ACC)Elem<T>::load(y);
      }
      const double xd = (double)x;
      a = (ln >= 0) ? fma(xd, ld_d(qat(Q, rs, cs, r, ln)), a) : fma(xd, xd, a);
   }
   const double s = block_sum(a, red);
   if (threadIdx.x == 0) next[b] = s;
}

template <typename T>
struct GsQr {
   ofrr_ctx *ctx;
   int variant;
   int64_t m;
   int n;
   T *Q;
   int64_t rs, cs;
   double stol, otol, rtol, rd_skip;
   double *R;
   int ldr;
   int32_t *skip;
   QrLayout L;
   double *part, *coef, *normv, *scale, *gsum;
   int32_t *pre, *flag, *hflag;

   double *slot(int s) const { return part + (int64_t)s * L.nblk; }
   double *norm_slot() const { return slot(n); }
   double *chain_slot(int k) const { return slot(n + 1 + (k & 1)); }

   int setup() {
      L = qr_layout(m);
      const bool left = variant == OFRR_QR_CGS || variant == OFRR_QR_CGS2;
      const size_t sgroups = left ? (size_t)(n + kQrGemvCols - 1) / kQrGemvCols : 0;
      const size_t bytes = ((size_t)(n + 3) * L.nblk + 3 * (size_t)n + sgroups * (size_t)m) * 8 +
                           (size_t)n * 4 + 16;
      char *w = (char *)workspace(ctx, bytes);
      hflag = (int32_t *)pinned(ctx, 16);
      if (!w || !hflag) return set_error(ctx, kErrorCUDAMemory, "gs_qr workspace");
      part = (double *)w;
      coef = part + (size_t)(n + 3) * L.nblk;
      normv = coef + n;
      scale = normv + n;
      gsum = scale + n;
      pre = (int32_t *)(gsum + sgroups * (size_t)m);
      flag = pre + n;
      return kSuccess;
   }

   int init_columns(int c0, int c1) {
      const int64_t cnt = (int64_t)c1 * (c1 - c0);
      const int64_t total = std::max<int64_t>(cnt, c1 - c0);
      qr_init_kernel<<<(unsigned)((total + 255) / 256), 256, 0, ctx->stream>>>(n, c0, c1, R, ldr, skip);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   int colnorms(int c0, int c1) {
      dim3 grid(L.nblk, (c1 - c0 + kQrCols - 1) / kQrCols);
      qr_colnorm_kernel<T><<<grid, kQrThreads, 0, ctx->stream>>>(m, L.chunk, L.nblk, Q, rs, cs, c0, c1, part);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   int dots(int vcol, const double *vscale, int c0, int c1, const int32_t *excl, int pcol,
            const int32_t *gate, double *vnorm) {
      dim3 grid(L.nblk, std::max(1, (c1 - c0 + kQrCols - 1) / kQrCols));
      qr_dots_kernel<T><<<grid, kQrThreads, (size_t)L.chunk * 8, ctx->stream>>>(
          m, L.chunk, L.nblk, Q, rs, cs, vcol, vscale, c0, c1, excl, pcol,
          pcol >= 0 ? scale + pcol : nullptr, gate, part, vnorm);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   int reduce(int c0, int c1, const int32_t *excl, const int32_t *gate, int pre_col, int left,
              int rfix, int scol) {
      const int colblocks = std::max(1, (c1 - c0 + kQrWaves - 1) / kQrWaves);
      const int sblocks = scol >= 0 ? (int)std::min<int64_t>((m + kQrThreads - 1) / kQrThreads, 1024) : 0;
      qr_reduce_kernel<T><<<colblocks + sblocks, kQrThreads, 0, ctx->stream>>>(
          L.nblk, part, c0, c1, colblocks, excl, gate, pre_col, norm_slot(), stol, rd_skip, skip,
          normv, coef, R, ldr, left, rfix, m, Q, rs, cs, scol, scol >= 0 ? scale + scol : nullptr);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   int decide(const double *p, int col, const int32_t *prep, bool reorth) {
      qr_decide_kernel<<<1, 64, 0, ctx->stream>>>(L.nblk, p, col, prep, skip, R, ldr, normv, scale,
                                                   flag, stol, otol, reorth ? rtol : 0.0, rd_skip);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   // flag[0] (reorthogonalize), flag[1] (dropped) of the last decision
   int read_flags(int *again, int *dropped) {
      OFRR_HIP(ctx, hipMemcpyAsync(hflag, flag, 8, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      *again = hflag[0];
      *dropped = hflag[1];
      return kSuccess;
   }

   int rank1(int j) {
      dim3 grid(L.nblk, (n - j - 1 + kQrCols - 1) / kQrCols);
      qr_rank1_kernel<T><<<grid, kQrThreads, 0, ctx->stream>>>(m, L.chunk, Q, rs, cs, j, j + 1, n,
                                                               coef, skip + j, norm_slot());
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   int gemvn(int i) {
      const int ngroups = (i + kQrGemvCols - 1) / kQrGemvCols;
      if (ngroups > 0) {
         dim3 grid(L.nblk, ngroups);
         qr_gemvn_part_kernel<T><<<grid, kQrThreads, 0, ctx->stream>>>(m, L.chunk, Q, rs, cs, i,
                                                                        coef, skip + i, gsum);
         OFRR_LAUNCH(ctx);
      }
      qr_gemvn_apply_kernel<T><<<L.nblk, kQrThreads, 0, ctx->stream>>>(
          m, L.chunk, Q, rs, cs, i, ngroups, gsum, skip + i, norm_slot());
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   // one CGS pass (projection product + update) on column i; pass 0 also
   // checks the initial norm and finishes the normalization of column i-1
   int cgs_pass(int i, bool first, int pcol) {
      int st = dots(i, nullptr, 0, i, skip, pcol, first ? nullptr : skip + i,
                    first ? norm_slot() : nullptr);
      if (!st) st = reduce(0, i, skip, first ? nullptr : skip + i, first ? i : -1, 1, i, -1);
      if (!st) st = gemvn(i);
      return st;
   }

   // one sequential MGS pass of column vcol over the kept columns; the norm
   // partials of the result end in the norm slot
   int mgs_pass(int vcol, const std::vector<int> &kept) {
      const int K = (int)kept.size();
      for (int s = 0; s <= K; ++s) {
         const int lp = s > 0 ? kept[s - 1] : -1;
         const int ln = s < K ? kept[s] : -1;
         double *next = ln >= 0 ? chain_slot(s + 1) : norm_slot();
         qr_chain_kernel<T><<<L.nblk, kQrThreads, 0, ctx->stream>>>(
             m, L.chunk, L.nblk, Q, rs, cs, vcol, lp, ln, chain_slot(s), next, R, ldr, skip + vcol);
         OFRR_LAUNCH(ctx);
      }
      return kSuccess;
   }

   int scale_column(int col) {
      const int sblocks = (int)std::min<int64_t>((m + kQrThreads - 1) / kQrThreads, 1024);
      qr_reduce_kernel<T><<<sblocks, kQrThreads, 0, ctx->stream>>>(
          L.nblk, part, 0, 0, 0, nullptr, skip + col, -1, nullptr, stol, rd_skip, skip, normv, coef,
          R, ldr, 0, 0, m, Q, rs, cs, col, scale + col);
      OFRR_LAUNCH(ctx);
      return kSuccess;
   }

   // MGS (variant 0) and MGS_V2: right-looking
   int right_looking() {
      const bool v0 = variant == OFRR_QR_MGS;
      const bool reorth = v0 && rtol > 0.0;
      std::vector<int> kept;
      int st = colnorms(0, v0 ? n : 1);
      if (st) return st;
      if (v0) {
         qr_norms_kernel<<<(n + kQrWaves - 1) / kQrWaves, kQrThreads, 0, ctx->stream>>>(
             n, L.nblk, part, stol, normv, pre);
         OFRR_LAUNCH(ctx);
      }
      const int32_t *excl = v0 ? pre : nullptr;
      for (int j = 0; j < n; ++j) {
         const double *np = j == 0 ? slot(0) : norm_slot();
         for (int pass = 0;; ++pass) {
            if ((st = decide(np, j, excl, reorth && pass < kQrMaxPasses))) return st;
            if (!reorth) break;
            int again = 0, dropped = 0;
            if ((st = read_flags(&again, &dropped))) return st;
            if (!again) {
               if (!dropped) kept.push_back(j);
               break;
            }
            if ((st = mgs_pass(j, kept))) return st;
            np = norm_slot();
         }
         if (j + 1 < n && (st = dots(j, scale + j, j + 1, n, excl, -1, skip + j, nullptr))) return st;
         if ((st = reduce(j + 1, n, excl, skip + j, -1, 0, j, j))) return st;
         if (j + 1 < n && (st = rank1(j))) return st;
      }
      return kSuccess;
   }

   // CGS (variant 2, iterated with reorth_tol > 0) and CGS2: left-looking
   int left_looking() {
      const bool reorth = variant == OFRR_QR_CGS && rtol > 0.0;
      int st;
      for (int i = 0; i < n; ++i) {
         if ((st = cgs_pass(i, true, i > 0 ? i - 1 : -1))) return st;
         if (variant == OFRR_QR_CGS2 && (st = cgs_pass(i, false, -1))) return st;
         for (int pass = 0;; ++pass) {
            if ((st = decide(norm_slot(), i, nullptr, reorth && pass < kQrMaxPasses))) return st;
            if (!reorth) break;
            int again = 0, dropped = 0;
            if ((st = read_flags(&again, &dropped))) return st;
            if (!again) break;
            if ((st = cgs_pass(i, false, -1))) return st;
         }
      }
      return scale_column(n - 1);
   }

   int run() {
      int st = setup();
      if (!st) st = init_columns(0, n);
      if (st) return st;
      return (variant == OFRR_QR_MGS || variant == OFRR_QR_MGS_V2) ? right_looking() : left_looking();
   }

   // column i against the final columns 0..i-1 (the *_step entry points)
   int step(int i, int32_t *skipped) {
      int st = setup();
      if (st) return st;
      std::vector<int32_t> hs(i > 0 ? i : 1);
      if (i > 0) {
         OFRR_HIP(ctx, hipMemcpyAsync(hs.data(), skip, (size_t)i * 4, hipMemcpyDeviceToHost, ctx->stream));
         OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      }
      if ((st = init_columns(i, i + 1))) return st;
      const bool reorth = (variant == OFRR_QR_MGS || variant == OFRR_QR_CGS) && rtol > 0.0;
      std::vector<int> kept;
      for (int j = 0; j < i; ++j)
         if (!hs[j]) kept.push_back(j);
      if (variant == OFRR_QR_MGS) {
         // initial norm check, then the sequential pass
         if ((st = dots(i, nullptr, i, i, nullptr, -1, nullptr, norm_slot()))) return st;
         if ((st = reduce(i, i, nullptr, nullptr, i, 1, i, -1))) return st;
         if ((st = mgs_pass(i, kept))) return st;
      } else {
         if ((st = cgs_pass(i, true, -1))) return st;
         if (variant == OFRR_QR_CGS2 && (st = cgs_pass(i, false, -1))) return st;
      }
      for (int pass = 0;; ++pass) {
         if ((st = decide(norm_slot(), i, nullptr, reorth && pass < kQrMaxPasses))) return st;
         if (!reorth) break;
         int again = 0, dropped = 0;
         if ((st = read_flags(&again, &dropped))) return st;
         if (!again) break;
         st = variant == OFRR_QR_MGS ? mgs_pass(i, kept) : cgs_pass(i, false, -1);
         if (st) return st;
      }
      if ((st = scale_column(i))) return st;
      OFRR_HIP(ctx, hipMemcpyAsync(hflag, skip + i, 4, hipMemcpyDeviceToHost, ctx->stream));
      OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
      *skipped = hflag[0];
      return kSuccess;
   }
};

template <typename F>
int qr_dispatch(ofrr_ctx *ctx, int qdt, F &&f) {
   switch (qdt) {
      case OFRR_F64: return f(double());
      case OFRR_F32: return f(float());
      case OFRR_F16: return f(__half());
      case OFRR_BF16: return f(__hip_bfloat16());
      default: return set_error(ctx, kErrorBLASUnsupportedType, "gs_qr: unsupported dtype");
   }
}

int qr_check(ofrr_ctx *ctx, int variant, int64_t m, int n, const void *Q, int64_t rs, int64_t cs,
             const double *R, int ldr, const int32_t *skip) {
   if (variant < OFRR_QR_MGS || variant > OFRR_QR_CGS2)
      return set_error(ctx, kErrorInvalidArgument, "gs_qr: unknown variant");
   if (m <= 0 || n <= 0 || !Q || !R || !skip || ldr < n || rs <= 0 || cs <= 0)
      return set_error(ctx, kErrorInvalidArgument, "gs_qr: bad arguments");
   if (m < n) return set_error(ctx, kErrorInvalidArgument, "gs_qr: A must be a tall-thin matrix");
   return kSuccess;
}

}  // namespace
}  // namespace ofrr

using namespace ofrr;

extern "C" int ofrr_gs_qr(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q, int64_t rs,
                          int64_t cs, double select_tol, double orth_tol, double reorth_tol,
                          double *R, int ldr, int32_t *skip) {
   OFRR_CHECK_CTX(ctx);
   int st = qr_check(ctx, variant, m, n, Q, rs, cs, R, ldr, skip);
   if (st) return st;
   return qr_dispatch(ctx, qdt, [&](auto t) -> int {
      using T = decltype(t);
      GsQr<T> q{ctx, variant, m, n, (T *)Q, rs, cs, select_tol, orth_tol, reorth_tol,
                variant == OFRR_QR_MGS ? 0.0 : 1.0, R, ldr, skip};
      return q.run();
   });
}

extern "C" int ofrr_gs_qr_step(ofrr_ctx *ctx, int variant, int qdt, int64_t m, int n, void *Q,
                               int64_t rs, int64_t cs, int i, double select_tol, double orth_tol,
                               double reorth_tol, double *R, int ldr, int32_t *skip,
                               int32_t *skipped) {
   OFRR_CHECK_CTX(ctx);
   int st = qr_check(ctx, variant, m, n, Q, rs, cs, R, ldr, skip);
   if (st) return st;
   if (variant == OFRR_QR_MGS_V2)
      return set_error(ctx, kErrorInvalidArgument, "gs_qr_step: mgs_v2 has no column step");
   if (i < 0 || i >= n || !skipped) return set_error(ctx, kErrorInvalidArgument, "gs_qr_step: bad column");
   return qr_dispatch(ctx, qdt, [&](auto t) -> int {
      using T = decltype(t);
      GsQr<T> q{ctx, variant, m, n, (T *)Q, rs, cs, select_tol, orth_tol, reorth_tol,
                variant == OFRR_QR_MGS ? 0.0 : 1.0, R, ldr, skip};
      return q.step(i, skipped);
   });
}
