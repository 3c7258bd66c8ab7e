// FP64 dense symmetric eigensolver and Rayleigh-Ritz on the GPU.
//
// syev: parallel cyclic Jacobi, one workgroup per matrix (batched for the
// serving path).  A round pairs all indices with the round-robin ("circle")
// schedule; every 2x2 block (pair I, pair J) of the packed upper triangle of A
// is then transformed independently, A'_IJ = J_I^T A_IJ J_J, so one round
// needs two barriers.  A rotation is skipped when |a_pq| <= max(eps/2
// sqrt(|a_pp a_qq|), eps max|a_ij|) (the Demmel-Veselic threshold with an
// absolute floor), and the solver stops after a sweep without rotations.  The packed triangle lives in LDS for n <= 192
// (145 KiB of the 160 KiB per CU).  The rotations are logged, and a second,
// many-workgroup kernel forms V = J_1 J_2 ... row by row (rows are
// independent), so the sequential kernel touches only A.
// Replaces the host LAPACKE dsyev/dsygv of the reference
// (CPP/linalg/blas/mvops.cu:1601-1664).
//
// rr_truncated: the OFRR projection step (ofrr_eig, CPP/linalg/solvers/eig/
// rr_eig.hpp:21-35; MATLAB/+src/+rr/eig_rr.m:48-57 with the shift replaced by
// truncation): S = diag(G)^{-1/2}; eig(S G S) = U L; keep L_i > tau max L;
// B = S U_r L_r^{-1/2}; eig(B^T H B) = Z Theta; Y = B Z, descending.
// The kept rank stays on the device: the kernels after the truncation read it.
#include <mutex>

#include "internal.hpp"

namespace ofrr {

constexpr int kJacobiThreads = 1024;
constexpr int kLdsFullN = 128;    // full storage, row stride n + 1: <= 129 KiB of LDS
constexpr int kLdsMaxN = 192;     // packed upper triangle: n(n+1)/2 doubles <= 144.8 KiB
constexpr int kMaxSweeps = 30;
// 2x2 blocks per thread: ceil(np (np + 1) / 2 / 1024) with np = nn / 2 pairs
constexpr int kBlkFull = 3;    // nn <= 128
constexpr int kBlkPacked = 5;  // nn <= 192
constexpr int kBlkGlobal = 9;  // nn <= 256
constexpr int kBlkWide = 19;   // nn <= 384
constexpr int kJacobiMaxN = 384;
constexpr int kMaxPairs = kJacobiMaxN / 2;
constexpr int kGroup = 3;      // blocks whose loads are in flight together
static_assert(kBlkFull * kJacobiThreads >= (64 * 65) / 2, "blocks per thread, nn <= 128");
static_assert(kBlkPacked * kJacobiThreads >= (96 * 97) / 2, "blocks per thread, nn <= 192");
static_assert(kBlkGlobal * kJacobiThreads >= (128 * 129) / 2, "blocks per thread, nn <= 256");
static_assert(kBlkWide * kJacobiThreads >= (kMaxPairs * (kMaxPairs + 1)) / 2, "blocks per thread, nn <= 384");

// Pair s of round r in the circle schedule on nn indices (M = nn - 1): s = 0
// pairs M with r, s > 0 pairs r + s with r - s (mod M); sorted so p < q.
__device__ inline void pair_of(int M, int r, int s, int &p, int &q) {
   if (s == 0) {
      p = M;
      q = r;
   } else {
      p = r + s;
      if (p >= M) p -= M;
      q = r - s;
      if (q < 0) q += M;
   }
   if (p > q) { const int t = p; p = q; q = t; }
}

// Storage of the symmetric working matrix on nn = n + (n odd) indices; an odd
// n gets a virtual index n whose row and column hold zeros (its pairs never
// rotate and its entries are never written, so the arithmetic equals skipping
// it).  Only the upper triangle is stored and touched.  FULL lays it out in
// square storage with row stride nn + 1 (conflict-free LDS rows): entry (i, j),
// i <= j, at off(i) + j with off(i) = i (nn + 1).  Otherwise the packed upper
// triangle: (i, j), i <= j, at off(j) + i with off(j) = j (j + 1) / 2.  Both
// offsets grow with the index, so the entry of (i, j) in either order is
// min(oi, oj) + max(i, j) (full) or max(oi, oj) + min(i, j) (packed): three
// integer operations, and the same on byte-scaled arguments (atb: 8 i, 8 oi).
template <bool FULL> struct SymIdx;
template <> struct SymIdx<true> {
   int ld;
   __device__ int off(int i) const { return i * ld; }
   __device__ int at(int i, int j, int oi, int oj) const { return min(oi, oj) + max(i, j); }
   __device__ int atb(int ib, int jb, int oib, int ojb) const { return min(oib, ojb) + max(ib, jb); }
};
template <> struct SymIdx<false> {
   int ld;
   __device__ int off(int i) const { return (i * (i + 1)) >> 1; }
   __device__ int at(int i, int j, int oi, int oj) const { return max(oi, oj) + min(i, j); }
   __device__ int atb(int ib, int jb, int oib, int ojb) const { return max(oib, ojb) + min(ib, jb); }
};
// the double at byte offset ob of a
__device__ inline double &at_byte(double *a, int ob) {
   return *reinterpret_cast<double *>(reinterpret_cast<char *>(a) + ob);
}

// Rotations: A is diagonalised in place (LDS or global) and the (c, s) of every
// pair of every round is logged; the eigenvectors are formed afterwards by
// jacobi_apply_kernel.  Outputs per matrix: eigenvalues ascending (w), the rank
// of each diagonal index (rank), and the number of logged sweeps.
// Each thread owns up to NB 2x2 blocks (pair I, pair J >= I) of the upper block
// triangle for the whole solve; a round loads every owned block before it
// writes any (the blocks of a round are disjoint), so the loads of all blocks
// are in flight together.
template <bool LDS, bool FULL, int NB>
__global__ __launch_bounds__(kJacobiThreads) void jacobi_rot_kernel(
    int n_max, const int32_t *__restrict__ n_dev, const double *__restrict__ A, int lda,
    int64_t strideA, double *__restrict__ w, int64_t stridew, double2 *__restrict__ rotlog,
    int64_t strideLog, int32_t *__restrict__ rank, int32_t *__restrict__ nsweeps,
    int64_t strideInts, double *__restrict__ work, int64_t strideWork, int max_sweeps) {
   extern __shared__ double lds[];
   __shared__ double2 rcs[kMaxPairs];  // (c, s) of pair slot s in this round
   __shared__ int4 rpq[kMaxPairs];     // 8 (p, q, off(p), off(q)) of pair slot s, p < q (bytes)
   __shared__ int rotated, s_n;
   __shared__ double s_amax[16];
   const int b = blockIdx.x;
   const double *Ab = A + b * strideA;
   double2 *log = rotlog + b * strideLog;
   if (threadIdx.x == 0) s_n = n_dev ? min((int)n_dev[b], n_max) : n_max;
   __syncthreads();
   const int n = s_n;
   if (n <= 0) {
      if (threadIdx.x == 0) nsweeps[b * strideInts] = 0;
      return;
   }
   const int nn = n + (n & 1), np = nn / 2, M = nn - 1;
   double *a = LDS ? lds : work + b * strideWork;
   const SymIdx<FULL> pk{nn + 1};
   double lmax = 0.0;
   for (int j = threadIdx.x; j < nn; j += blockDim.x) {
      const int oj = pk.off(j);
      for (int i = 0; i <= j; ++i) {
         const double v = (j < n) ? Ab[i + (int64_t)j * lda] : 0.0;
         a[pk.at(i, j, pk.off(i), oj)] = v;
         lmax = fmax(lmax, fabs(v));
      }
   }
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, o, 64));
   if ((threadIdx.x & 63) == 0) s_amax[threadIdx.x >> 6] = lmax;
   __syncthreads();
   double amax = 0.0;
   for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) amax = fmax(amax, s_amax[i]);
   // rotations below eps * max|a_ij| cannot change any eigenvalue beyond the
   // accuracy the matrix carries: they end the sweep like the relative test
   const double tol_abs = 2.220446049250313e-16 * amax;
   // this thread's blocks: t = threadIdx.x + u * kJacobiThreads in I-major order
   int blk[NB];
   {
      const int nblocks = np * (np + 1) / 2;
      const float b2 = 2.0f * np + 1.0f;
#pragma unroll
      for (int u = 0; u < NB; ++u) {
         const int t = threadIdx.x + u * kJacobiThreads;
         blk[u] = -1;
         if (t < nblocks) {
            // I = largest row with I*np - I*(I-1)/2 <= t
            int I = (int)((b2 - sqrtf(b2 * b2 - 8.0f * (float)t)) * 0.5f);
            while (I > 0 && I * np - I * (I - 1) / 2 > t) --I;
            while ((I + 1) * np - (I + 1) * I / 2 <= t) ++I;
            const int J = I + (t - (I * np - I * (I - 1) / 2));
            blk[u] = (I << 16) | J;
         }
      }
   }
   int done = 0;
   // the sweeps, with the virtual-index tests compiled in only for odd sizes
   // (ODD = true is correct for every size)
   auto sweeps = [&](auto odd_tag) {
   constexpr bool ODD = decltype(odd_tag)::value;
   for (int sweep = 0; sweep < max_sweeps; ++sweep) {
      if (threadIdx.x == 0) rotated = 0;
      __syncthreads();
      for (int r = 0; r < M; ++r) {
         for (int s = threadIdx.x; s < np; s += blockDim.x) {
            int p, q;
            pair_of(M, r, s, p, q);
            const int op = pk.off(p), oq = pk.off(q);
            double c = 1.0, sn = 0.0;
            if (!ODD || q < n) {
               const double apq = a[pk.at(p, q, op, oq)], app = a[pk.at(p, p, op, op)],
                            aqq = a[pk.at(q, q, oq, oq)];
               // the rotation is formed whether or not it is taken, so its
               // divide/square-root chain overlaps the threshold's square root
               const double th = (aqq - app) / (2.0 * apq);
               const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(fma(th, th, 1.0)));
               const double cr = 1.0 / sqrt(fma(t, t, 1.0));
               if (fabs(apq) > fmax(1.1102230246251565e-16 * sqrt(fabs(app) * fabs(aqq)), tol_abs)) {
                  c = cr;
                  sn = t * cr;
                  rotated = 1;
               }
            }
            rcs[s] = make_double2(c, sn);
            rpq[s] = make_int4(8 * p, 8 * q, 8 * op, 8 * oq);
            log[((int64_t)sweep * M + r) * np + s] = make_double2(c, sn);
         }
         __syncthreads();
         // A' = J^T A J block by block over the upper block triangle:
         // A'_IJ = J_I^T A_IJ J_J (I <= J), explicit fused multiply-adds.  A
         // diagonal block (I = J) runs the same arithmetic; its a_pq is then 0.
         // groups of kGroup blocks: all loads of a group, then its stores
#pragma unroll
         for (int g = 0; g < NB; g += kGroup) {
            double v11[kGroup], v12[kGroup], v21[kGroup], v22[kGroup];
            double2 cI[kGroup], cJ[kGroup];
            int4 qI[kGroup], qJ[kGroup];
            bool act[kGroup];
#pragma unroll
            for (int e = 0; e < kGroup; ++e) {
               const int u = g + e;
               act[e] = (u < NB) && blk[u < NB ? u : 0] >= 0;
               const int bu = act[e] ? blk[u < NB ? u : 0] : 0;
               const int I = bu >> 16, J = bu & 0xffff;
               cI[e] = rcs[I];
               cJ[e] = rcs[J];
               act[e] = act[e] && (cI[e].y != 0.0 || cJ[e].y != 0.0);
               qI[e] = rpq[I];
               qJ[e] = rpq[J];
               if (act[e]) {
                  // rows p1 < q1 (pair I), columns p2 < q2 (pair J); a virtual
                  // index holds zeros, so every read is of stored memory
                  v11[e] = at_byte(a, pk.atb(qI[e].x, qJ[e].x, qI[e].z, qJ[e].z));
                  v12[e] = at_byte(a, pk.atb(qI[e].x, qJ[e].y, qI[e].z, qJ[e].w));
                  v21[e] = at_byte(a, pk.atb(qI[e].y, qJ[e].x, qI[e].w, qJ[e].z));
                  v22[e] = at_byte(a, pk.atb(qI[e].y, qJ[e].y, qI[e].w, qJ[e].w));
               }
            }
#pragma unroll
            for (int e = 0; e < kGroup; ++e) {
               if (!act[e]) continue;
               const double c1 = cI[e].x, s1 = cI[e].y, c2 = cJ[e].x, s2 = cJ[e].y;
               const int p1 = qI[e].x, q1 = qI[e].y, p2 = qJ[e].x, q2 = qJ[e].y;
               const int op1 = qI[e].z, oq1 = qI[e].w, op2 = qJ[e].z, oq2 = qJ[e].w;
               const double m11 = fma(c2, v11[e], -(s2 * v12[e])), m12 = fma(s2, v11[e], c2 * v12[e]);
               const double m21 = fma(c2, v21[e], -(s2 * v22[e])), m22 = fma(s2, v21[e], c2 * v22[e]);
               const double b11 = fma(c1, m11, -(s1 * m21));
               double b12 = fma(c1, m12, -(s1 * m22));
               double b21 = fma(s1, m11, c1 * m21);
               const double b22 = fma(s1, m12, c1 * m22);
               if (p1 == p2) b12 = b21 = 0.0;  // diagonal block
               // (byte-scaled indices) only an odd size has a virtual index
               const bool q1ok = !ODD || q1 < 8 * n, q2ok = !ODD || q2 < 8 * n;
               at_byte(a, pk.atb(p1, p2, op1, op2)) = b11;
               if (q2ok) at_byte(a, pk.atb(p1, q2, op1, oq2)) = b12;
               if (q1ok) at_byte(a, pk.atb(q1, p2, oq1, op2)) = b21;
               if (q1ok && q2ok) at_byte(a, pk.atb(q1, q2, oq1, oq2)) = b22;
            }
         }
         __syncthreads();
      }
      if (!rotated) break;
      done = sweep + 1;
      __syncthreads();
   }
   };
   if constexpr (LDS) {
      if (n & 1)
         sweeps(std::true_type{});
      else
         sweeps(std::false_type{});
   } else {  // the global-memory forms keep one copy (registers are at their limit)
      sweeps(std::true_type{});
   }
   if (threadIdx.x == 0) nsweeps[b * strideInts] = done;
   double *wb = w + b * stridew;
   int32_t *rk = rank + b * strideInts;
   for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const double di = a[pk.at(i, i, pk.off(i), pk.off(i))];
      int rr = 0;
      for (int j = 0; j < n; ++j) {
         const double dj = a[pk.at(j, j, pk.off(j), pk.off(j))];
         rr += (dj < di || (dj == di && j < i)) ? 1 : 0;
      }
      wb[rr] = di;
      rk[i] = rr;
   }
}
// V = J_1 J_2 ... (all logged rotations) applied to the identity, one row of V
// per wave.  The row lives in registers, arranged like the pairs of the circle
// schedule: lane l holds the row's entries at the two indices of pair slots
// l, l + 64, ..., l + 64 (S - 1), each slot's a-side (r + s mod M; M for s = 0)
// and b-side (r - s mod M).  A round rotates each slot's two entries, then the
// indices move to the next round's slots: every a-side entry one slot down,
// every b-side entry one slot up, with a[1] -> b[0] and b[np-1] -> a[np-1] at
// the ends and a[0] (index M) fixed.  The arithmetic per entry is the one of
// the rotation kernel's logged (c, s) on sorted (p, q), in round order.  The
// log is staged through LDS kApplyChunk rounds at a time, the next chunk loaded
// into registers while the current one is applied.  S = 2 serves n <= 256,
// S = 3 n <= 384.
constexpr int kApplyChunk = 16;
// Whole-wave lane shifts of a double (DPP wave_shl:1 / wave_shr:1 on each
// 32-bit half): register moves, where a general shuffle is an LDS round trip.
// The lane with no source (63, resp. 0) receives 0.
__device__ inline double wave_next(double v) {  // lane i <- lane i + 1
   const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, true);
   const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, true);
   return __hiloint2double(hi, lo);
}
__device__ inline double wave_prev(double v) {  // lane i <- lane i - 1
   const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, true);
   const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, true);
   return __hiloint2double(hi, lo);
}
__device__ inline double lane_value(double v, int l) {  // lane l's v, in every lane
   const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
   const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
   return __hiloint2double(hi, lo);
}
template <int S>
__global__ __launch_bounds__(64) void jacobi_apply_kernel(
    int n_max, const int32_t *__restrict__ n_dev, const double2 *__restrict__ rotlog,
    int64_t strideLog, const int32_t *__restrict__ rank, const int32_t *__restrict__ nsweeps,
    int64_t strideInts, double *__restrict__ V, int ldv, int64_t strideV) {
   constexpr int kPer = kApplyChunk * S;  // log entries per lane and chunk
   __shared__ double2 lg[kApplyChunk * 64 * S];
   __shared__ double row[128 * S];
   const int b = blockIdx.y;
   const int n = n_dev ? min((int)n_dev[b], n_max) : n_max;
   const int r0 = blockIdx.x;
   if (n <= 0 || r0 >= n) return;
   const int nn = n + (n & 1), np = nn / 2, M = nn - 1;
   const int rounds = nsweeps[b * strideInts] * M;
   const double2 *log = rotlog + b * strideLog;
   const int lane = threadIdx.x;
   bool has[S];
   int ia[S], ib[S];  // indices (a, b) of slot lane + 64 u; round 0: slot 0 = (M, 0), slot s = (s, M - s)
   double av[S], bv[S];
#pragma unroll
   for (int u = 0; u < S; ++u) {
      const int sl = lane + 64 * u;
      has[u] = sl < np;
      ia[u] = (sl == 0) ? M : sl;
      ib[u] = (sl == 0) ? 0 : M - sl;
      av[u] = (has[u] && ia[u] == r0) ? 1.0 : 0.0;
      bv[u] = (has[u] && ib[u] == r0) ? 1.0 : 0.0;
   }
   double2 pre[kPer];
   auto fetch = [&](int c0) {
      const int nr = min(kApplyChunk, rounds - c0);
      const double2 *src = log + (int64_t)c0 * np;
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
         const int i = lane + 64 * u;
         pre[u] = (i < nr * np) ? src[i] : make_double2(1.0, 0.0);
      }
   };
   auto rotate = [](double2 cs, int ia_, int ib_, double &va, double &vb) {
      if (cs.y == 0.0) return;
      // (p, q) = the sorted indices; row[p] = c vp - s vq, row[q] = s vp + c vq
      const bool aIsP = ia_ < ib_;
      const double vp = aIsP ? va : vb, vq = aIsP ? vb : va;
      const double np_ = fma(cs.x, vp, -(cs.y * vq));
      const double nq_ = fma(cs.y, vp, cs.x * vq);
      va = aIsP ? np_ : nq_;
      vb = aIsP ? nq_ : np_;
   };
   if (rounds > 0) fetch(0);
   for (int c0 = 0; c0 < rounds; c0 += kApplyChunk) {
      const int nr = min(kApplyChunk, rounds - c0);
#pragma unroll
      for (int u = 0; u < kPer; ++u) lg[lane + 64 * u] = pre[u];
      __syncthreads();
      if (c0 + kApplyChunk < rounds) fetch(c0 + kApplyChunk);  // in flight during this chunk
      // each round's (c, s) is read from LDS one round ahead, so the read is
      // off the round's dependence chain (rotate, lane shifts, selects)
      double2 csn[S];
#pragma unroll
      for (int u = 0; u < S; ++u) csn[u] = lg[lane + 64 * u];
      for (int k = 0; k < nr; ++k) {
         double2 csc[S];
#pragma unroll
         for (int u = 0; u < S; ++u) csc[u] = csn[u];
         const double2 *lr = lg + min(k + 1, nr - 1) * np;
#pragma unroll
         for (int u = 0; u < S; ++u) csn[u] = lr[lane + 64 * u];
#pragma unroll
         for (int u = 0; u < S; ++u)
            if (has[u]) rotate(csc[u], ia[u], ib[u], av[u], bv[u]);
         // move to the next round's slots
         double an[S], bp[S], af[S], bl[S];
#pragma unroll
         for (int u = 0; u < S; ++u) {
            an[u] = wave_next(av[u]);             // a[s + 1], same set (lane 63: unused)
            bp[u] = wave_prev(bv[u]);             // b[s - 1], same set (lane 0: unused)
            af[u] = (u > 0) ? lane_value(av[u], 0) : 0.0;       // a[64 u]: a[s + 1] of lane 63 of set u - 1
            bl[u] = (u + 1 < S) ? lane_value(bv[u], 63) : 0.0;  // b[64 u + 63]: b[s - 1] of lane 0 of set u + 1
         }
         const double a01 = lane_value(av[0], 1);  // a[1] for slot 0
         double na[S], nb[S];
#pragma unroll
         for (int u = 0; u < S; ++u) {
            const int sl = lane + 64 * u;
            if (sl == 0) {
               na[u] = av[u];
               nb[u] = (np > 1) ? a01 : bv[u];  // one slot: nothing moves
            } else {
               const double nxt = (lane == 63) ? ((u + 1 < S) ? af[u + 1 < S ? u + 1 : u] : av[u]) : an[u];
               na[u] = (sl == np - 1) ? bv[u] : nxt;
               nb[u] = (lane == 0) ? bl[u > 0 ? u - 1 : 0] : bp[u];
            }
         }
#pragma unroll
         for (int u = 0; u < S; ++u) {
            av[u] = na[u];
            bv[u] = nb[u];
            // indices of the next round
            if (lane + 64 * u == 0) {
               ib[u] = (ib[u] + 1 >= M) ? 0 : ib[u] + 1;
            } else {
               ia[u] = (ia[u] + 1 == M) ? 0 : ia[u] + 1;
               ib[u] = (ib[u] + 1 == M) ? 0 : ib[u] + 1;
            }
         }
      }
      __syncthreads();  // the chunk is consumed before the next one is stored
   }
   // scatter the row back to index order through LDS, then to its ranked columns
#pragma unroll
   for (int u = 0; u < S; ++u)
      if (has[u]) { row[ia[u]] = av[u]; row[ib[u]] = bv[u]; }
   __syncthreads();
   double *Vb = V + b * strideV;
   const int32_t *rk = rank + b * strideInts;
   for (int c = lane; c < n; c += 64) Vb[r0 + (int64_t)rk[c] * ldv] = row[c];
}

// per-matrix scratch (doubles): [packed A | rotation log | ranks | sweep count]
struct JacobiLayout {
   size_t packed, log, ints, total;
};
static JacobiLayout jacobi_layout(int n) {
   const int nn = n + (n & 1);
   JacobiLayout L;
   // packed triangle on nn indices (an odd n gets a zero virtual index)
   L.packed = ((size_t)nn * (nn + 1) / 2 + 1) & ~(size_t)1;  // even: the log is double2
   L.log = (size_t)kMaxSweeps * (nn - 1) * (nn / 2) * 2;
   L.ints = ((size_t)n + 2 + 1) / 2;
   L.total = (L.packed + L.log + L.ints + 1) & ~(size_t)1;
   return L;
}
static size_t jacobi_work_doubles(int n) { return jacobi_layout(n).total; }

int syev_batched(ofrr_ctx *ctx, int batch, int n, const int32_t *n_dev, const double *A, int lda,
                 int64_t strideA, double *w, int64_t stridew, double *V, int ldv,
                 int64_t strideV, double *work /* nullable: workspace */) {
   if (n <= 0 || batch <= 0) return kSuccess;
   if (n > kJacobiMaxN) return set_error(ctx, kErrorBLASInvalidSize, "syev: n > 384 not supported");
   const JacobiLayout L = jacobi_layout(n);
   const int64_t strideWork = (int64_t)L.total;
   if (!work) {
      work = (double *)workspace(ctx, (size_t)batch * strideWork * sizeof(double));
      if (!work) return set_error(ctx, kErrorCUDAMemory, "syev: workspace");
   }
   double2 *rotlog = reinterpret_cast<double2 *>(work + L.packed);
   int32_t *rank = reinterpret_cast<int32_t *>(work + L.packed + L.log);
   int32_t *nsw = rank + n;
   const int64_t strideLog = strideWork / 2, strideInts = strideWork * 2;
   // the LDS limits of the two LDS-resident variants, set once per process
   // (several host threads may run solves at once)
   static std::once_flag attr_once;
   static hipError_t attr_err = hipSuccess;
   std::call_once(attr_once, [] {
      const int full = (int)((size_t)kLdsFullN * (kLdsFullN + 1) * sizeof(double));
      const int packed = (int)((size_t)kLdsMaxN * (kLdsMaxN + 1) / 2 * sizeof(double));
      attr_err = hipFuncSetAttribute((const void *)jacobi_rot_kernel<true, true, kBlkFull>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, full);
      if (attr_err == hipSuccess)
         attr_err = hipFuncSetAttribute((const void *)jacobi_rot_kernel<true, false, kBlkPacked>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, packed);
   });
   OFRR_HIP(ctx, attr_err);
   // storage sized for nn = n + (n odd) indices (the largest of a batch)
   const int nn = n + (n & 1);
   if (nn <= kLdsFullN) {
      const size_t lds = (size_t)nn * (nn + 1) * sizeof(double);
      jacobi_rot_kernel<true, true, kBlkFull><<<batch, kJacobiThreads, lds, ctx->stream>>>(
          n, n_dev, A, lda, strideA, w, stridew, rotlog, strideLog, rank, nsw, strideInts, work,
          strideWork, kMaxSweeps);
   } else if (nn <= kLdsMaxN) {
      const size_t lds = (size_t)nn * (nn + 1) / 2 * sizeof(double);
      jacobi_rot_kernel<true, false, kBlkPacked><<<batch, kJacobiThreads, lds, ctx->stream>>>(
          n, n_dev, A, lda, strideA, w, stridew, rotlog, strideLog, rank, nsw, strideInts, work,
          strideWork, kMaxSweeps);
   } else if (nn <= 256) {
      jacobi_rot_kernel<false, false, kBlkGlobal><<<batch, kJacobiThreads, 0, ctx->stream>>>(
          n, n_dev, A, lda, strideA, w, stridew, rotlog, strideLog, rank, nsw, strideInts, work,
          strideWork, kMaxSweeps);
   } else {
      jacobi_rot_kernel<false, false, kBlkWide><<<batch, kJacobiThreads, 0, ctx->stream>>>(
          n, n_dev, A, lda, strideA, w, stridew, rotlog, strideLog, rank, nsw, strideInts, work,
          strideWork, kMaxSweeps);
   }
   OFRR_LAUNCH(ctx);
   dim3 grid(n, batch);
   if (nn <= 256)
      jacobi_apply_kernel<2><<<grid, 64, 0, ctx->stream>>>(n, n_dev, rotlog, strideLog, rank, nsw,
                                                          strideInts, V, ldv, strideV);
   else
      jacobi_apply_kernel<3><<<grid, 64, 0, ctx->stream>>>(n, n_dev, rotlog, strideLog, rank, nsw,
                                                          strideInts, V, ldv, strideV);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---------------------------------------------------------------------------
// C = alpha op(A) B + beta C, small dense FP64 (one thread per output)
__global__ void small_gemm_kernel(int transA, int M, int N, int K, const int32_t *__restrict__ kdev,
                                  double alpha, const double *__restrict__ A, int lda,
                                  const double *__restrict__ B, int ldb, double beta,
                                  double *__restrict__ Cm, int ldc) {
   int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
   if (i >= M || j >= N) return;
   int kk = kdev ? min(K, (int)*kdev) : K;
   double s = 0.0;
   for (int p = 0; p < kk; ++p) {
      double av = transA ? A[p + (int64_t)i * lda] : A[i + (int64_t)p * lda];
      s = fma(av, B[p + (int64_t)j * ldb], s);
   }
   double *c = Cm + i + (int64_t)j * ldc;
   *c = (beta == 0.0) ? alpha * s : fma(alpha, s, beta * *c);
}

static int small_gemm(ofrr_ctx *ctx, int transA, int M, int N, int K, const int32_t *kdev,
                      double alpha, const double *A, int lda, const double *B, int ldb,
                      double beta, double *Cm, int ldc) {
   small_gemm_kernel<<<dim3((M + 63) / 64, N), 64, 0, ctx->stream>>>(transA, M, N, K, kdev, alpha,
                                                                    A, lda, B, ldb, beta, Cm, ldc);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---------------------------------------------------------------------------
// truncated RR helpers
__global__ void rr_scale_kernel(int m, const double *__restrict__ G, int ldg,
                                const double *__restrict__ H, int ldh, double *__restrict__ sc,
                                double *__restrict__ Gs, double *__restrict__ Hs) {
   int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= m * m) return;
   int i = t % m, j = t / m;
   double gi = G[i + (int64_t)i * ldg], gj = G[j + (int64_t)j * ldg];
   double si = gi > 0.0 ? 1.0 / sqrt(gi) : 0.0, sj = gj > 0.0 ? 1.0 / sqrt(gj) : 0.0;
   if (j == 0) sc[i] = si;
   double g = 0.5 * (G[i + (int64_t)j * ldg] + G[j + (int64_t)i * ldg]);
   double h = 0.5 * (H[i + (int64_t)j * ldh] + H[j + (int64_t)i * ldh]);
   Gs[t] = g * si * sj;
   Hs[t] = h * si * sj;
}

// rank r of the eigenvalues lam (ascending) above tau * max; Bm = U_r L_r^{-1/2}
// (m x r, columns = the top r in ascending order)
__global__ void rr_trunc_kernel(int m, double tau, const double *__restrict__ lam,
                                const double *__restrict__ U, double *__restrict__ Bm,
                                int32_t *__restrict__ rank) {
   __shared__ int s_r;
   if (threadIdx.x == 0) {
      double lmax = lam[m - 1];
      int r = 0;
      for (int i = 0; i < m; ++i) r += (lam[i] > tau * lmax) ? 1 : 0;
      s_r = r;
      *rank = r;
   }
   __syncthreads();
   const int r = s_r;
   for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
      int i = t % m, c = t / m;
      double val = 0.0;
      if (c < r) {
         int src = m - r + c;
         val = U[i + (int64_t)src * m] / sqrt(lam[src]);
      }
      Bm[t] = val;
   }
}

// theta (descending) and Y = S Bm Z for the top nkeep
__global__ void rr_finish_kernel(int m, int nkeep, const int32_t *__restrict__ rank,
                                 const double *__restrict__ sc, const double *__restrict__ Bm,
                                 const double *__restrict__ th, const double *__restrict__ Z,
                                 double *__restrict__ theta, double *__restrict__ Y, int ldy) {
   const int r = *rank;
   const int c = blockIdx.x;
   for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double s = 0.0;
      if (c < r) {
         int src = r - 1 - c;
         for (int p = 0; p < r; ++p) s = fma(Bm[i + (int64_t)p * m], Z[p + (int64_t)src * m], s);
         s *= sc[i];
      }
      Y[i + (int64_t)c * ldy] = s;
   }
   if (threadIdx.x == 0) theta[c] = (c < r) ? th[r - 1 - c] : __builtin_nan("");
}

// workspace (bytes) rr_truncated asks for at basis dimension m
size_t rr_workspace_bytes(int m) {
   const size_t mm = (size_t)m * m;
   return (m + 6 * mm + 2 * m + 10 + jacobi_work_doubles(m)) * sizeof(double);
}

int rr_truncated(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                 double tau, int nkeep, double *theta, double *Y, int ldy, int32_t *rank) {
   if (m <= 0 || nkeep <= 0 || nkeep > m) return set_error(ctx, kErrorInvalidArgument, "rr_truncated: sizes");
   // workspace: sc(m) Gs Hs U Bm T Hr Z (m*m each) lam th (m) rank(1) + jacobi work
   const size_t mm = (size_t)m * m;
   double *base = (double *)workspace(ctx, rr_workspace_bytes(m));
   if (!base) return set_error(ctx, kErrorCUDAMemory, "rr_truncated: workspace");
   double *sc = base, *Gs = sc + m, *Hs = Gs + mm, *U = Hs + mm, *Bm = U + mm, *T = Bm + mm;
   double *Hr = T + mm, *lam = Hr + mm, *th = lam + m;
   int32_t *rk = (int32_t *)(th + m);
   double *jwork = base + ((3 * (size_t)m + 6 * mm + 8 + 1) & ~(size_t)1);  // 16-byte aligned
   double *Z = Gs;  // Gs is dead after its eigendecomposition
   int32_t *rank_out = rank ? rank : rk;
   rr_scale_kernel<<<(unsigned)((mm + 255) / 256), 256, 0, ctx->stream>>>(m, G, ldg, H, ldh, sc, Gs, Hs);
   OFRR_LAUNCH(ctx);
   int st = syev_batched(ctx, 1, m, nullptr, Gs, m, 0, lam, 0, U, m, 0, jwork);
   if (st) return st;
   rr_trunc_kernel<<<1, 256, 0, ctx->stream>>>(m, tau, lam, U, Bm, rank_out);
   OFRR_LAUNCH(ctx);
   // Hr = Bm^T Hs Bm over the first r columns (zero beyond)
   if ((st = small_gemm(ctx, 0, m, m, m, nullptr, 1.0, Hs, m, Bm, m, 0.0, T, m))) return st;
   if ((st = small_gemm(ctx, 1, m, m, m, nullptr, 1.0, Bm, m, T, m, 0.0, Hr, m))) return st;
   // eig of the leading r x r block of Hr (ld m); sizes from the device
   st = syev_batched(ctx, 1, m, rank_out, Hr, m, 0, th, 0, Z, m, 0, jwork);
   if (st) return st;
   rr_finish_kernel<<<nkeep, 256, 0, ctx->stream>>>(m, nkeep, rank_out, sc, Bm, th, Z, theta, Y, ldy);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---------------------------------------------------------------------------
// Grouped truncated Rayleigh-Ritz: the jobs of several solves (each its own G,
// H, m, tau, outputs) in one launch per stage, job = blockIdx.y (or .z).  Each
// job's work area holds the arrays of rr_truncated with leading dimension
// mx (the largest m of the group); every job runs the arithmetic of
// rr_truncated on its own m in the same order, so its results do not depend
// on the other jobs of its group.
struct RRJobArgs {
   int n, mx;
   int64_t slab;  // doubles per job area
   RRJob j[kMaxRRJobs];
};
// job area: sc (mx) | Gs Hs U Bm T Hr (mx^2 each) | lam th (mx)
__device__ inline double *rrj_area(double *base, const RRJobArgs &a, int b) { return base + b * a.slab; }

__global__ void rrg_scale_kernel(RRJobArgs a, double *base, int32_t *mdev) {
   const int b = blockIdx.y, m = a.j[b].m, mx = a.mx;
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t == 0) mdev[b] = m;
   if (t >= m * m) return;
   const int i = t % m, j = t / m;
   const double *G = a.j[b].G, *H = a.j[b].H;
   const int ldg = a.j[b].ldg, ldh = a.j[b].ldh;
   double *sc = rrj_area(base, a, b), *Gs = sc + mx, *Hs = Gs + (int64_t)mx * mx;
   double gi = G[i + (int64_t)i * ldg], gj = G[j + (int64_t)j * ldg];
   double si = gi > 0.0 ? 1.0 / sqrt(gi) : 0.0, sj = gj > 0.0 ? 1.0 / sqrt(gj) : 0.0;
   if (j == 0) sc[i] = si;
   double g = 0.5 * (G[i + (int64_t)j * ldg] + G[j + (int64_t)i * ldg]);
   double h = 0.5 * (H[i + (int64_t)j * ldh] + H[j + (int64_t)i * ldh]);
   Gs[i + (int64_t)j * mx] = g * si * sj;
   Hs[i + (int64_t)j * mx] = h * si * sj;
}

__global__ void rrg_trunc_kernel(RRJobArgs a, double *base, int32_t *rdev) {
   __shared__ int s_r;
   const int b = blockIdx.x, m = a.j[b].m, mx = a.mx;
   const int64_t mm = (int64_t)mx * mx;
   double *sc = rrj_area(base, a, b), *U = sc + mx + 2 * mm, *Bm = U + mm, *lam = sc + mx + 6 * mm;
   if (threadIdx.x == 0) {
      double lmax = lam[m - 1];
      int r = 0;
      for (int i = 0; i < m; ++i) r += (lam[i] > a.j[b].tau * lmax) ? 1 : 0;
      s_r = r;
      rdev[b] = r;
      *a.j[b].rank = r;
   }
   __syncthreads();
   const int r = s_r;
   for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
      int i = t % m, c = t / m;
      double val = 0.0;
      if (c < r) {
         int src = m - r + c;
         val = U[i + (int64_t)src * mx] / sqrt(lam[src]);
      }
      Bm[i + (int64_t)c * mx] = val;
   }
}

// stage 0: T = Hs Bm; stage 1: Hr = Bm^T T (the small_gemm arithmetic, ld mx)
__global__ void rrg_gemm_kernel(RRJobArgs a, double *base, int stage) {
   const int b = blockIdx.z, m = a.j[b].m, mx = a.mx;
   const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
   if (i >= m || j >= m) return;
   const int64_t mm = (int64_t)mx * mx;
   double *sc = rrj_area(base, a, b), *Hs = sc + mx + mm, *Bm = sc + mx + 3 * mm, *T = Bm + mm;
   double *Hr = T + mm;
   double s = 0.0;
   if (stage == 0) {
      for (int p = 0; p < m; ++p) s = fma(Hs[i + (int64_t)p * mx], Bm[p + (int64_t)j * mx], s);
      T[i + (int64_t)j * mx] = 1.0 * s;
   } else {
      for (int p = 0; p < m; ++p) s = fma(Bm[p + (int64_t)i * mx], T[p + (int64_t)j * mx], s);
      Hr[i + (int64_t)j * mx] = 1.0 * s;
   }
}

__global__ void rrg_finish_kernel(RRJobArgs a, double *base, const int32_t *rdev) {
   const int b = blockIdx.y, m = a.j[b].m, mx = a.mx, c = blockIdx.x;
   if (c >= a.j[b].nkeep) return;
   const int64_t mm = (int64_t)mx * mx;
   const double *sc = rrj_area(base, a, b), *Z = sc + mx, *Bm = sc + mx + 3 * mm;
   const double *th = sc + mx + 6 * mm + mx;
   const int r = rdev[b];
   double *Y = a.j[b].Y;
   const int ldy = a.j[b].ldy;
   for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double s = 0.0;
      if (c < r) {
         int src = r - 1 - c;
         for (int p = 0; p < r; ++p) s = fma(Bm[i + (int64_t)p * mx], Z[p + (int64_t)src * mx], s);
         s *= sc[i];
      }
      Y[i + (int64_t)c * ldy] = s;
   }
   if (threadIdx.x == 0) a.j[b].theta[c] = (c < r) ? th[r - 1 - c] : __builtin_nan("");
}

size_t rr_group_workspace_bytes(int njobs, int mx) {
   const size_t slab = ((size_t)mx + 6 * (size_t)mx * mx + 2 * (size_t)mx + 1) & ~(size_t)1;
   return ((size_t)njobs * slab + (size_t)njobs * jacobi_work_doubles(mx) + (size_t)njobs + 8) *
          sizeof(double);
}

int rr_truncated_group(ofrr_ctx *ctx, int njobs, const RRJob *jobs) {
   if (njobs <= 0 || njobs > kMaxRRJobs) return set_error(ctx, kErrorInvalidArgument, "rr group: job count");
   RRJobArgs a{};
   a.n = njobs;
   int mx = 0, kx = 0;
   for (int b = 0; b < njobs; ++b) {
      if (jobs[b].m <= 0 || jobs[b].nkeep <= 0 || jobs[b].nkeep > jobs[b].m || jobs[b].m > kJacobiMaxN)
         return set_error(ctx, kErrorInvalidArgument, "rr group: job sizes");
      a.j[b] = jobs[b];
      mx = std::max(mx, jobs[b].m);
      kx = std::max(kx, jobs[b].nkeep);
   }
   a.mx = mx;
   a.slab = (int64_t)(((size_t)mx + 6 * (size_t)mx * mx + 2 * (size_t)mx + 1) & ~(size_t)1);
   double *base = (double *)workspace(ctx, rr_group_workspace_bytes(njobs, mx));
   if (!base) return set_error(ctx, kErrorCUDAMemory, "rr group: workspace");
   double *jwork = base + (size_t)njobs * a.slab;
   int32_t *mdev = (int32_t *)(jwork + (size_t)njobs * jacobi_work_doubles(mx));
   int32_t *rdev = mdev + njobs;
   const int64_t mm = (int64_t)mx * mx;
   // per job: Gs at sc + mx, U at sc + mx + 2 mm, lam at sc + mx + 6 mm;
   // Hr at sc + mx + 5 mm, th at lam + mx, Z over Gs (dead after its eigensolve)
   double *Gs0 = base + mx, *U0 = base + mx + 2 * mm, *lam0 = base + mx + 6 * mm;
   double *Hr0 = base + mx + 5 * mm, *th0 = lam0 + mx;
   for (int b = 0; b < njobs; ++b)
      if (jobs[b].h_ready) OFRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, jobs[b].h_ready, 0));
   rrg_scale_kernel<<<dim3((unsigned)((mm + 255) / 256), njobs), 256, 0, ctx->stream>>>(a, base, mdev);
   OFRR_LAUNCH(ctx);
   int st = syev_batched(ctx, njobs, mx, mdev, Gs0, mx, a.slab, lam0, a.slab, U0, mx, a.slab, jwork);
   if (st) return st;
   rrg_trunc_kernel<<<njobs, 256, 0, ctx->stream>>>(a, base, rdev);
   OFRR_LAUNCH(ctx);
   for (int stage = 0; stage < 2; ++stage) {
      rrg_gemm_kernel<<<dim3((mx + 63) / 64, mx, njobs), 64, 0, ctx->stream>>>(a, base, stage);
      OFRR_LAUNCH(ctx);
   }
   st = syev_batched(ctx, njobs, mx, rdev, Hr0, mx, a.slab, th0, a.slab, Gs0, mx, a.slab, jwork);
   if (st) return st;
   rrg_finish_kernel<<<dim3(kx, njobs), 256, 0, ctx->stream>>>(a, base, rdev);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

// ---------------------------------------------------------------------------
// Split truncated Rayleigh-Ritz for a basis V = [X | D] whose first kx columns
// are the previous Ritz block (X^T X ~ I): the same pencil (G, H), reached by
// a congruence that keeps every eigendecomposition small.
//   G_XX = L L^T (Cholesky); C = G_XX^{-1} G_XD; S = G_DD - G_XD^T C, the Gram
//   matrix of D with its X component removed, scaled to unit diagonal by s;
//   eig(s S s) = U Lam, kept Lam_i > tau max(1, max Lam) (the X directions
//   have unit scaled eigenvalues) -> B_D = s U_r Lam_r^{-1/2};
//   W = [[L^{-T}, -C B_D], [0, B_D]] so that W^T G W = I on the kept directions;
//   Hr = W^T H W; eig(Hr_DD) = Q hd first (Hr_XX is already near diag(theta) as
//   X holds Ritz vectors); Hr2 = [[Hr_XX, Hr_XD Q], [., diag(hd)]];
//   eig(Hr2) = Z th; Y = W blkdiag(I, Q) Z for the top nkeep, descending.
// The D-block Gram eigendecomposition is of size m - kx instead of m, and the
// coupled one starts from diagonal blocks, so it needs few sweeps.  Jobs of a
// group run one launch per stage (job = blockIdx.y / .z) in their own work
// areas (leading dimension mx); each job's arithmetic is the oracle's
// orc_rr_split: every sum a sequential FMA chain in increasing index order.
// A job whose G_XX is not positive definite writes rank -1 and skips its
// eigensolves (the caller falls back to rr_truncated).
struct RRSplitArgs {
   int n, mx, kxm;
   int64_t slab;  // doubles per job area
   RRJob j[kMaxRRJobs];
};
// job area (ld mx): L (kxm x kxm) | Cc (kxm x mx) | S (S, then T, then H2) |
// Ss (Ss, then Hr, then Z) | U (U, then Q) | W | W2 (Hd, then W2) | sd lam hd th
// (mx each)
struct RRSplitArea {
   double *L, *Cc, *S, *Ss, *U, *W, *W2, *sd, *lam, *hd, *th;
};
__host__ __device__ inline RRSplitArea rrs_area(double *base, int mx, int kxm, int64_t slab, int b) {
   RRSplitArea a;
   const int64_t mm = (int64_t)mx * mx;
   a.L = base + b * slab;
   a.Cc = a.L + (int64_t)kxm * kxm;
   a.S = a.Cc + (int64_t)kxm * mx;
   a.Ss = a.S + mm;
   a.U = a.Ss + mm;
   a.W = a.U + mm;
   a.W2 = a.W + mm;
   a.sd = a.W2 + mm;
   a.lam = a.sd + mx;
   a.hd = a.lam + mx;
   a.th = a.hd + mx;
   return a;
}
__host__ __device__ inline int64_t rrs_slab(int mx, int kxm) {
   return ((int64_t)kxm * kxm + (int64_t)kxm * mx + 5 * (int64_t)mx * mx + 4 * (int64_t)mx + 1) &
          ~(int64_t)1;
}
// per-job device ints: fail | md | r | mr (each njobs long)
__device__ inline double rrs_g(const RRJob &J, int i, int j) {
   return 0.5 * (J.G[i + (int64_t)j * J.ldg] + J.G[j + (int64_t)i * J.ldg]);
}
__device__ inline double rrs_h(const RRJob &J, int i, int j) {
   return 0.5 * (J.H[i + (int64_t)j * J.ldh] + J.H[j + (int64_t)i * J.ldh]);
}

// Cholesky of G_XX in LDS (right-looking, the oracle's order); L (lower, ld
// kxm) to the job area
__global__ __launch_bounds__(1024) void rrs_chol_kernel(RRSplitArgs a, double *base, int32_t *ints) {
   extern __shared__ double Ls[];
   __shared__ int s_fail;
   const int b = blockIdx.x, nj = a.n;
   const RRJob &J = a.j[b];
   const int kx = J.kx, m = J.m, md = m - kx;
   const RRSplitArea A = rrs_area(base, a.mx, a.kxm, a.slab, b);
#define LS(i, j) Ls[(i) + (j) * kx]
   if (threadIdx.x == 0) s_fail = 0;
   for (int t = threadIdx.x; t < kx * kx; t += blockDim.x) {
      const int i = t % kx, j = t / kx;
      Ls[t] = (i >= j) ? rrs_g(J, i, j) : 0.0;
   }
   __syncthreads();
   for (int j = 0; j < kx; ++j) {
      const double d = LS(j, j);  // every thread reads it before anyone writes column j
      if (!(d > 0.0)) {
         if (threadIdx.x == 0) s_fail = 1;
         break;  // uniform: d is the same LDS value for every thread
      }
      const double sq = sqrt(d);
      __syncthreads();
      for (int i = j + threadIdx.x; i < kx; i += blockDim.x) LS(i, j) = (i == j) ? sq : LS(i, j) / sq;
      __syncthreads();
      const int w = kx - j - 1;
      for (int t = threadIdx.x; t < w * w; t += blockDim.x) {
         const int i = j + 1 + t % w, c = j + 1 + t / w;
         if (i >= c) LS(i, c) = fma(-LS(i, j), LS(c, j), LS(i, c));
      }
      __syncthreads();
   }
   __syncthreads();
   const int fail = s_fail;
   if (threadIdx.x == 0) {
      ints[b] = fail;
      ints[nj + b] = fail ? 0 : md;  // the D-block eigensolve's size
      if (fail) *J.rank = -1;
   }
   if (fail) return;
   for (int t = threadIdx.x; t < kx * kx; t += blockDim.x) {
      const int i = t % kx, j = t / kx;
      A.L[i + (int64_t)j * a.kxm] = Ls[t];
   }
#undef LS
}

// Triangular solves with L, one wave per column, right-looking with the
// column in registers (lane l holds rows l and l + 64): column c < md gives
// C(:, c) = L^{-T} L^{-1} G_XD(:, c); column md + c' gives W(0:kx, c') =
// L^{-T} e_c'.  The forward pass adds the terms of each row in increasing
// index order, the backward pass in decreasing order: the oracle's chains.
__global__ __launch_bounds__(64) void rrs_solve_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int kx = J.kx, md = J.m - kx, col = blockIdx.x;
   if (col >= md + kx) return;
   const RRSplitArea A = rrs_area(base, a.mx, a.kxm, a.slab, b);
   const int ldl = a.kxm, lane = threadIdx.x;
   const double *L = A.L;
   const bool rhs = col < md;
   double x0 = 0.0, x1 = 0.0;  // rows lane, lane + 64
   if (rhs) {
      if (lane < kx) x0 = rrs_g(J, lane, kx + col);
      if (lane + 64 < kx) x1 = rrs_g(J, lane + 64, kx + col);
      for (int i = 0; i < kx; ++i) {
         const double piv = L[i + (int64_t)i * ldl];
         double xi = (i < 64) ? __shfl(x0, i, 64) : __shfl(x1, i - 64, 64);
         xi = xi / piv;
         if (lane == (i & 63)) {
            if (i < 64) x0 = xi;
            else x1 = xi;
         }
         if (lane > i && lane < kx) x0 = fma(-L[lane + (int64_t)i * ldl], xi, x0);
         if (lane + 64 > i && lane + 64 < kx) x1 = fma(-L[lane + 64 + (int64_t)i * ldl], xi, x1);
      }
   } else {
      const int c = col - md;
      x0 = (lane == c) ? 1.0 : 0.0;
      x1 = (lane + 64 == c) ? 1.0 : 0.0;
   }
   for (int p = kx - 1; p >= 0; --p) {
      const double piv = L[p + (int64_t)p * ldl];
      double xp = (p < 64) ? __shfl(x0, p, 64) : __shfl(x1, p - 64, 64);
      xp = xp / piv;
      if (lane == (p & 63)) {
         if (p < 64) x0 = xp;
         else x1 = xp;
      }
      if (lane < p) x0 = fma(-L[p + (int64_t)lane * ldl], xp, x0);
      if (lane + 64 < p) x1 = fma(-L[p + (int64_t)(lane + 64) * ldl], xp, x1);
   }
   double *out = rhs ? A.Cc + (int64_t)col * a.kxm : A.W + (int64_t)(col - md) * a.mx;
   if (lane < kx) out[lane] = x0;
   if (lane + 64 < kx) out[lane + 64] = x1;
}

// S(i, j) = G_DD(i, j) - G_XD(:, i)^T C(:, j), upper triangle (i <= j)
__global__ void rrs_schur_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int kx = J.kx, md = J.m - kx, mx = a.mx;
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= md * md) return;
   const int i = t % md, j = t / md;
   if (i > j) return;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   const double *cj = A.Cc + (int64_t)j * a.kxm;
   double s2 = rrs_g(J, kx + i, kx + j);
   for (int p = 0; p < kx; ++p) s2 = fma(-rrs_g(J, p, kx + i), cj[p], s2);
   A.S[i + (int64_t)j * mx] = s2;
}

// sd = diag(S)^{-1/2} (0 for a non-positive entry); Ss = S(i, j) sd_i sd_j (upper)
__global__ void rrs_sscale_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int md = J.m - J.kx, mx = a.mx;
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= md * md) return;
   const int i = t % md, j = t / md;
   if (i > j) return;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   const double di = A.S[i + (int64_t)i * mx], dj = A.S[j + (int64_t)j * mx];
   const double si = di > 0.0 ? 1.0 / sqrt(di) : 0.0, sj = dj > 0.0 ? 1.0 / sqrt(dj) : 0.0;
   if (i == j) A.sd[i] = si;
   A.Ss[i + (int64_t)j * mx] = A.S[i + (int64_t)j * mx] * si * sj;
}

// kept D rank r; W(kx:, kx + c) = U(:, src) / sqrt(lam_src) * sd, W(kx:, 0:kx) = 0
__global__ void rrs_trunc_kernel(RRSplitArgs a, double *base, int32_t *ints) {
   __shared__ int s_r;
   const int b = blockIdx.x, nj = a.n;
   if (ints[b]) {
      if (threadIdx.x == 0) ints[2 * nj + b] = ints[3 * nj + b] = 0;
      return;
   }
   const RRJob &J = a.j[b];
   const int kx = J.kx, md = J.m - kx, mx = a.mx;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   if (threadIdx.x == 0) {
      const double lmax = (md > 0 && A.lam[md - 1] > 1.0) ? A.lam[md - 1] : 1.0;
      int r = 0;
      for (int i = 0; i < md; ++i) r += (A.lam[i] > J.tau * lmax) ? 1 : 0;
      s_r = r;
      ints[2 * nj + b] = r;
      ints[3 * nj + b] = kx + r;
      *J.rank = kx + r;
   }
   __syncthreads();
   const int r = s_r;
   for (int t = threadIdx.x; t < md * (kx + r); t += blockDim.x) {
      const int i = t % md, c = t / md;
      double v = 0.0;
      if (c >= kx) {
         const int src = md - r + (c - kx);
         v = A.U[i + (int64_t)src * mx] / sqrt(A.lam[src]) * A.sd[i];
      }
      A.W[kx + i + (int64_t)c * mx] = v;
   }
}

// W(0:kx, kx + c) = -C B_D(:, c)
__global__ void rrs_wxd_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y, nj = a.n;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int kx = J.kx, md = J.m - kx, mx = a.mx, r = ints[2 * nj + b];
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= kx * r) return;
   const int i = t % kx, c = t / kx;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   const double *bd = A.W + kx + (int64_t)(kx + c) * mx;
   double s2 = 0.0;
   for (int p = 0; p < md; ++p) s2 = fma(A.Cc[i + (int64_t)p * a.kxm], bd[p], s2);
   A.W[i + (int64_t)(kx + c) * mx] = -s2;
}

// stage 0: T = H W (m x mr, into S); stage 1: Hr = W^T T (mr x mr, into Ss), and
// its D block copied to W2 for the eigensolver
__global__ void rrs_gemm_kernel(RRSplitArgs a, double *base, const int32_t *ints, int stage) {
   const int b = blockIdx.z, nj = a.n;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int m = J.m, kx = J.kx, mx = a.mx, mr = ints[3 * nj + b];
   const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   double s2 = 0.0;
   if (stage == 0) {
      if (i >= m || c >= mr) return;
      const double *w = A.W + (int64_t)c * mx;
      for (int p = 0; p < m; ++p) s2 = fma(rrs_h(J, i, p), w[p], s2);
      A.S[i + (int64_t)c * mx] = s2;
   } else {
      if (i >= mr || c >= mr) return;
      const double *w = A.W + (int64_t)i * mx, *tc = A.S + (int64_t)c * mx;
      for (int p = 0; p < m; ++p) s2 = fma(w[p], tc[p], s2);
      A.Ss[i + (int64_t)c * mx] = s2;
      if (i >= kx && c >= kx) A.W2[(i - kx) + (int64_t)(c - kx) * mx] = s2;
   }
}

// Hr2 (upper, into S): [[Hr_XX, Hr_XD Q], [., diag(hd)]]; W2 = W blkdiag(I, Q)
__global__ void rrs_couple_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y, nj = a.n;
   if (ints[b]) return;
   const RRJob &J = a.j[b];
   const int m = J.m, kx = J.kx, mx = a.mx, r = ints[2 * nj + b], mr = kx + r;
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= m * mr) return;
   const int i = t % m, c = t / m;
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   const double *Hr = A.Ss, *Q = A.U;
   if (i <= c && i < mr) {
      double v;
      if (c < kx) {
         v = Hr[i + (int64_t)c * mx];
      } else if (i < kx) {
         double s2 = 0.0;
         for (int p = 0; p < r; ++p) s2 = fma(Hr[i + (int64_t)(kx + p) * mx], Q[p + (int64_t)(c - kx) * mx], s2);
         v = s2;
      } else {
         v = (i == c) ? A.hd[i - kx] : 0.0;
      }
      A.S[i + (int64_t)c * mx] = v;
   }
   double v = A.W[i + (int64_t)c * mx];
   if (c >= kx) {
      double s2 = 0.0;
      for (int p = 0; p < r; ++p) s2 = fma(A.W[i + (int64_t)(kx + p) * mx], Q[p + (int64_t)(c - kx) * mx], s2);
      v = s2;
   }
   A.W2[i + (int64_t)c * mx] = v;
}

// theta (descending) and Y = W2 Z for the top nkeep
__global__ void rrs_finish_kernel(RRSplitArgs a, double *base, const int32_t *ints) {
   const int b = blockIdx.y, nj = a.n, c = blockIdx.x;
   const RRJob &J = a.j[b];
   if (c >= J.nkeep || ints[b]) return;
   const int m = J.m, mx = a.mx, mr = ints[3 * nj + b];
   const RRSplitArea A = rrs_area(base, mx, a.kxm, a.slab, b);
   const double *Z = A.Ss;
   for (int i = threadIdx.x; i < m; i += blockDim.x) {
      double s2 = 0.0;
      if (c < mr) {
         const int src = mr - 1 - c;
         for (int p = 0; p < mr; ++p) s2 = fma(A.W2[i + (int64_t)p * mx], Z[p + (int64_t)src * mx], s2);
      }
      J.Y[i + (int64_t)c * J.ldy] = s2;
   }
   if (threadIdx.x == 0) J.theta[c] = (c < mr) ? A.th[mr - 1 - c] : __builtin_nan("");
}

size_t rr_split_workspace_bytes(int njobs, int mx, int kxm) {
   return ((size_t)njobs * rrs_slab(mx, kxm) + (size_t)njobs * jacobi_work_doubles(mx) +
           (size_t)2 * njobs + 8) * sizeof(double);
}

int rr_split_group(ofrr_ctx *ctx, int njobs, const RRJob *jobs) {
   if (njobs <= 0 || njobs > kMaxRRJobs) return set_error(ctx, kErrorInvalidArgument, "rr split: job count");
   RRSplitArgs a{};
   a.n = njobs;
   int mx = 0, kxm = 0, kx_min = 1 << 30, mdx = 0, nk = 0;
   for (int b = 0; b < njobs; ++b) {
      const RRJob &J = jobs[b];
      if (J.m <= 0 || J.m > kJacobiMaxN || J.kx <= 0 || J.kx > J.m || J.kx > 128 || J.nkeep <= 0 ||
          J.nkeep > J.m)
         return set_error(ctx, kErrorInvalidArgument, "rr split: job sizes");
      a.j[b] = J;
      mx = std::max(mx, J.m);
      kxm = std::max(kxm, J.kx);
      kx_min = std::min(kx_min, J.kx);
      mdx = std::max(mdx, J.m - J.kx);
      nk = std::max(nk, J.nkeep);
   }
   a.mx = mx;
   a.kxm = kxm;
   a.slab = rrs_slab(mx, kxm);
   double *base = (double *)workspace(ctx, rr_split_workspace_bytes(njobs, mx, kxm));
   if (!base) return set_error(ctx, kErrorCUDAMemory, "rr split: workspace");
   double *jwork = base + (size_t)njobs * a.slab;
   int32_t *ints = (int32_t *)(jwork + (size_t)njobs * jacobi_work_doubles(mx));
   const RRSplitArea A0 = rrs_area(base, mx, kxm, a.slab, 0);
   static std::once_flag attr_once;
   static hipError_t attr_err = hipSuccess;
   std::call_once(attr_once, [] {
      attr_err = hipFuncSetAttribute((const void *)rrs_chol_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 128 * 8);
   });
   OFRR_HIP(ctx, attr_err);
   rrs_chol_kernel<<<njobs, 1024, (size_t)kxm * kxm * sizeof(double), ctx->stream>>>(a, base, ints);
   OFRR_LAUNCH(ctx);
   rrs_solve_kernel<<<dim3(mdx + kxm, njobs), 64, 0, ctx->stream>>>(a, base, ints);
   OFRR_LAUNCH(ctx);
   const unsigned gmd = (unsigned)(((int64_t)mdx * mdx + 255) / 256);
   if (mdx > 0) {
      rrs_schur_kernel<<<dim3(gmd, njobs), 256, 0, ctx->stream>>>(a, base, ints);
      OFRR_LAUNCH(ctx);
      rrs_sscale_kernel<<<dim3(gmd, njobs), 256, 0, ctx->stream>>>(a, base, ints);
      OFRR_LAUNCH(ctx);
      // eig of the scaled D-block Gram matrix (size md per job, 0 for a failed job)
      int st = syev_batched(ctx, njobs, mdx, ints + njobs, A0.Ss, mx, a.slab, A0.lam, a.slab, A0.U,
                            mx, a.slab, jwork);
      if (st) return st;
   }
   rrs_trunc_kernel<<<njobs, 256, 0, ctx->stream>>>(a, base, ints);
   OFRR_LAUNCH(ctx);
   if (mdx > 0) {
      rrs_wxd_kernel<<<dim3((unsigned)(((int64_t)kxm * mdx + 255) / 256), njobs), 256, 0, ctx->stream>>>(
          a, base, ints);
      OFRR_LAUNCH(ctx);
   }
   // the first read of H (everything above needs G only)
   for (int b = 0; b < njobs; ++b)
      if (jobs[b].h_ready) OFRR_HIP(ctx, hipStreamWaitEvent(ctx->stream, jobs[b].h_ready, 0));
   for (int stage = 0; stage < 2; ++stage) {
      rrs_gemm_kernel<<<dim3((mx + 63) / 64, mx, njobs), 64, 0, ctx->stream>>>(a, base, ints, stage);
      OFRR_LAUNCH(ctx);
   }
   if (mdx > 0) {
      // eig of Hr's D block (size r per job, copied into W2)
      int st = syev_batched(ctx, njobs, mdx, ints + 2 * njobs, A0.W2, mx, a.slab, A0.hd, a.slab, A0.U,
                            mx, a.slab, jwork);
      if (st) return st;
   }
   rrs_couple_kernel<<<dim3((unsigned)(((int64_t)mx * mx + 255) / 256), njobs), 256, 0, ctx->stream>>>(
       a, base, ints);
   OFRR_LAUNCH(ctx);
   int st = syev_batched(ctx, njobs, mx, ints + 3 * njobs, A0.S, mx, a.slab, A0.th, a.slab, A0.Ss, mx,
                         a.slab, jwork);
   if (st) return st;
   rrs_finish_kernel<<<dim3(nk, njobs), 256, 0, ctx->stream>>>(a, base, ints);
   OFRR_LAUNCH(ctx);
   (void)kx_min;
   return kSuccess;
}

int rr_group_run(ofrr_ctx *ctx, int njobs, const RRJob *jobs) {
   RRJob whole[kMaxRRJobs], split[kMaxRRJobs];
   int nw = 0, ns = 0;
   for (int b = 0; b < njobs && b < kMaxRRJobs; ++b) {
      if (jobs[b].kx > 0) split[ns++] = jobs[b];
      else whole[nw++] = jobs[b];
   }
   int st = kSuccess;
   if (nw) st = rr_truncated_group(ctx, nw, whole);
   if (!st && ns) st = rr_split_group(ctx, ns, split);
   return st;
}

// ---------------------------------------------------------------------------
// generalised problem H x = lambda (G + gtol I) x via Cholesky (one workgroup)
__global__ __launch_bounds__(1024) void chol_transform_kernel(int m, const double *__restrict__ G,
                                                              int ldg, const double *__restrict__ H,
                                                              int ldh, double gtol,
                                                              double *__restrict__ L,
                                                              double *__restrict__ Cm,
                                                              int32_t *__restrict__ info) {
   __shared__ int s_info;
   if (threadIdx.x == 0) s_info = 0;
   for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
      int i = t % m, j = t / m;
      // lower triangle of B from G's upper triangle; C from H's upper triangle
      L[t] = (i >= j) ? G[j + (int64_t)i * ldg] + (i == j ? gtol : 0.0) : 0.0;
      Cm[t] = (i <= j) ? H[i + (int64_t)j * ldh] : H[j + (int64_t)i * ldh];
   }
   __syncthreads();
   // right-looking Cholesky, column by column
   for (int j = 0; j < m; ++j) {
      double d = L[j + j * m];
      if (!(d > 0.0)) {
         if (threadIdx.x == 0 && s_info == 0) s_info = j + 1;
         __syncthreads();
         break;
      }
      d = sqrt(d);
      __syncthreads();
      for (int i = j + threadIdx.x; i < m; i += blockDim.x) L[i + j * m] = (i == j) ? d : L[i + j * m] / d;
      __syncthreads();
      for (int t = threadIdx.x; t < (m - j - 1) * (m - j - 1); t += blockDim.x) {
         int i = j + 1 + t % (m - j - 1), c = j + 1 + t / (m - j - 1);
         if (i >= c) L[i + c * m] = fma(-L[i + j * m], L[c + j * m], L[i + c * m]);
      }
      __syncthreads();
   }
   if (threadIdx.x == 0) *info = s_info;
   if (s_info) return;
   // C <- L^{-1} C (columns independent), then C <- C L^{-T} (rows independent)
   for (int c = threadIdx.x; c < m; c += blockDim.x)
      for (int i = 0; i < m; ++i) {
         double s = Cm[i + c * m];
         for (int p = 0; p < i; ++p) s = fma(-L[i + p * m], Cm[p + c * m], s);
         Cm[i + c * m] = s / L[i + i * m];
      }
   __syncthreads();
   for (int r = threadIdx.x; r < m; r += blockDim.x)
      for (int j = 0; j < m; ++j) {
         double s = Cm[r + j * m];
         for (int p = 0; p < j; ++p) s = fma(-L[j + p * m], Cm[r + p * m], s);
         Cm[r + j * m] = s / L[j + j * m];
      }
   __syncthreads();
   for (int t = threadIdx.x; t < m * m; t += blockDim.x) {
      int i = t % m, j = t / m;
      if (i < j) {
         double s = 0.5 * (Cm[i + j * m] + Cm[j + i * m]);
         Cm[i + j * m] = s;
         Cm[j + i * m] = s;
      }
   }
}

// V = L^{-T} Yv (back substitution per column); optional descending reorder
__global__ void chol_back_kernel(int m, const double *__restrict__ L, const double *__restrict__ w,
                                 const double *__restrict__ Yv, int descending,
                                 double *__restrict__ wout, double *__restrict__ V, int ldv) {
   int c = blockIdx.x * blockDim.x + threadIdx.x;
   if (c >= m) return;
   int src = descending ? m - 1 - c : c;
   wout[c] = w[src];
   for (int i = m - 1; i >= 0; --i) {
      double s = Yv[i + (int64_t)src * m];
      for (int p = i + 1; p < m; ++p) s = fma(-L[p + i * m], V[p + (int64_t)c * ldv], s);
      V[i + (int64_t)c * ldv] = s / L[i + i * m];
   }
}

int sygv_impl(ofrr_ctx *ctx, int m, const double *A, int lda, const double *B, int ldb,
              double gtol, int descending, double *w, double *V, int ldv) {
   const size_t mm = (size_t)m * m;
   size_t need = (3 * mm + m + 8 + jacobi_work_doubles(m)) * sizeof(double);
   double *base = (double *)workspace(ctx, need);
   if (!base) return set_error(ctx, kErrorCUDAMemory, "sygv: workspace");
   double *L = base, *Cm = L + mm, *Yv = Cm + mm, *wl = Yv + mm;
   int32_t *info = (int32_t *)(wl + m);
   double *jwork = wl + m + 8;
   int32_t *hinfo = (int32_t *)pinned(ctx, sizeof(int32_t));
   if (!hinfo) return set_error(ctx, kErrorCUDAMemory, "sygv: pinned");
   chol_transform_kernel<<<1, 1024, 0, ctx->stream>>>(m, B, ldb, A, lda, gtol, L, Cm, info);
   OFRR_LAUNCH(ctx);
   OFRR_HIP(ctx, hipMemcpyAsync(hinfo, info, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   if (*hinfo) return set_error(ctx, kErrorSingularMatrix, "sygv: B is not positive definite");
   int st = syev_batched(ctx, 1, m, nullptr, Cm, m, 0, wl, 0, Yv, m, 0, jwork);
   if (st) return st;
   chol_back_kernel<<<(m + 63) / 64, 64, 0, ctx->stream>>>(m, L, wl, Yv, descending, w, V, ldv);
   OFRR_LAUNCH(ctx);
   return kSuccess;
}

int rr_generalized(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                   double gtol, double *theta, double *Y, int ldy) {
   return sygv_impl(ctx, m, H, ldh, G, ldg, gtol, 1, theta, Y, ldy);
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_syev_batched(ofrr_ctx *ctx, int batch, int n, const double *A, int lda, int64_t strideA,
                      double *w, int64_t stridew, double *V, int ldv, int64_t strideV) {
   OFRR_CHECK_CTX(ctx);
   if (batch < 0 || n < 0 || lda < n || ldv < n) return set_error(ctx, kErrorInvalidArgument, "syev: bad sizes");
   if (batch == 0 || n == 0) return kSuccess;
   if (!A || !w || !V) return set_error(ctx, kErrorInvalidArgument, "syev: null pointer");
   return syev_batched(ctx, batch, n, nullptr, A, lda, strideA, w, stridew, V, ldv, strideV, nullptr);
}

int ofrr_sygv(ofrr_ctx *ctx, int n, const double *A, int lda, const double *B, int ldb, double *w,
              double *V, int ldv) {
   OFRR_CHECK_CTX(ctx);
   if (n <= 0 || lda < n || ldb < n || ldv < n || !A || !B || !w || !V)
      return set_error(ctx, kErrorInvalidArgument, "sygv: bad arguments");
   if (n > kJacobiMaxN) return set_error(ctx, kErrorBLASInvalidSize, "sygv: n > 384 not supported");
   return sygv_impl(ctx, n, A, lda, B, ldb, 0.0, 0, w, V, ldv);
}

int ofrr_rr_truncated(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                      double tau, int nkeep, double *theta, double *Y, int ldy, int32_t *rank) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || m > kJacobiMaxN || ldg < m || ldh < m || ldy < m || !G || !H || !theta || !Y)
      return set_error(ctx, kErrorInvalidArgument, "rr_truncated: bad arguments");
   return rr_truncated(ctx, m, G, ldg, H, ldh, tau, nkeep, theta, Y, ldy, rank);
}

int ofrr_rr_split(ofrr_ctx *ctx, int m, int kx, const double *G, int ldg, const double *H, int ldh,
                  double tau, int nkeep, double *theta, double *Y, int ldy, int32_t *rank) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || m > kJacobiMaxN || kx <= 0 || kx > m || kx > 128 || nkeep <= 0 || nkeep > m ||
       ldg < m || ldh < m || ldy < m || !G || !H || !theta || !Y || !rank)
      return set_error(ctx, kErrorInvalidArgument, "rr_split: bad arguments");
   const RRJob job{G, H, ldg, ldh, m, nkeep, tau, theta, Y, ldy, rank, kx};
   return rr_split_group(ctx, 1, &job);
}

int ofrr_rr_generalized(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                        double gtol, double *theta, double *Y, int ldy) {
   OFRR_CHECK_CTX(ctx);
   if (m <= 0 || m > kJacobiMaxN || ldg < m || ldh < m || ldy < m || !G || !H || !theta || !Y)
      return set_error(ctx, kErrorInvalidArgument, "rr_generalized: bad arguments");
   return rr_generalized(ctx, m, G, ldg, H, ldh, gtol, theta, Y, ldy);
}

}  // extern "C"
