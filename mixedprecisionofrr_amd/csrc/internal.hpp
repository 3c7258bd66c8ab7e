// Host-side launchers shared between the library's translation units.
#pragma once
#include "common.hpp"

namespace ofrr {

// Y = alpha A X + beta Y (any layout; row-major contiguous blocks take the fast path).
// n rows of A and Y; nx rows of X (the columns of A; -1: n).
// bw: the operator's bandwidth max |col - row| when known (selects the strip
// schedule of the row tiles, sched_block), 0 for row order
int spmm_csr(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const void *va,
             int vdt, int k, const void *X, int xdt, int64_t xrs, int64_t xcs, double alpha,
             double beta, void *Y, int ydt, int64_t yrs, int64_t ycs, int64_t bw = 0,
             int64_t nx = -1);
// W = A X, R = W - X diag(theta) quantised by inv_sR into Rq (rdt), column max
// |R| folded into colmax (atomic max), sums[2j] = ||r_j||^2, sums[2j+1] = ||x_j||^2.
// X, W, Rq row-major with row stride k.
int spmm_resid(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const double *va,
               int k, const double *X, const double *theta, double *W, void *Rq, int rdt,
               const double *inv_sR, float *colmax, double *sums, int64_t bw = 0);
// one step of the normalised Chebyshev recurrence on row-major blocks of width k
// (see spmm.hip and solvers.hip):
// Dn = alpha (A Dc - c Dc) - beta Dp + w . R with stored blocks Dc = Dch sc,
// Dp = Dph sp, R = Rq sR; the output is stored as Dn / sn with
// sn = 2^(ceil(log2(2 max|Dc|)) - target_exp) written to sn_out
// The column maxima of a Chebyshev step are folded into kMaxSlots partial
// maxima per column (slot = workgroup index & 7, the workgroup's XCD in the
// round-robin dispatch), slot_ld floats apart; a reader takes the largest.
// One address per column drew every workgroup's atomic: 10 us of a 235 us step.
constexpr int kMaxSlots = 8;
struct ChebArgs {
   double alpha, beta, c;     // step scalars and the centre of the damped interval
   const double *step;        // nullable: the same three scalars in device memory (a
                              // captured sweep replays with new values)
   const double *w;           // per-column weight of the residual block (k values)
   const double *sR, *sc, *sp;// scales of R, Dc, Dp (sp may be null: no Dp)
   const float *maxc;         // max |D_j| per column (kMaxSlots partial maxima)
   float *maxn;               // max |D_{j+1}| per column (partial maxima, zeroed beforehand)
   float *zero_next;          // cleared for the next step (nullable; every slot)
   int slot_ld = 0;           // floats between the partial-maximum slots (0: one shared slot)
   double *sn_out;            // scale of the stored D_{j+1}
   int target_exp;
   int reverse = 0;           // walk the row tiles from the last back to the first
};
int cheb_step(ofrr_ctx *ctx, int64_t n, const int64_t *rp, const int32_t *ci, const void *va,
              int vdt, int k, const void *Dc, const void *Dp, const void *R, int bdt,
              const ChebArgs &args, void *Dn, int64_t bw = 0);
// G = alpha V^T W + beta G (FP64), strided operands
int gram(ofrr_ctx *ctx, int64_t n, int k1, int k2, const void *V, int vdt, int64_t vrs,
         int64_t vcs, const void *W, int wdt, int64_t wrs, int64_t wcs, double alpha, double beta,
         double *G, int ldg);
// Fused projection products: item t is tile (ti[t], tj[t]) (64 x 64) of the k x k
// block op[ia[t]]^T op[ib[t]], written to dst[t] (column-major, ld[t]) at rows
// ro[t] + ..., columns co[t] + ....  Operands are row-major n x k blocks of
// kind 0 (FP64), 1 (format tb) or 2 (format tz).  FP64 accumulation, fixed
// summation order.
constexpr int kMaxGramOps = 8, kMaxGramItems = 48;
struct GramOperand {
   const void *p;
   int64_t rs;
   int kind;
};
struct GramMultiArgs {
   GramOperand op[kMaxGramOps];
   int nitems;
   int ia[kMaxGramItems], ib[kMaxGramItems], ti[kMaxGramItems], tj[kMaxGramItems];
   double *dst[kMaxGramItems];
   int ld[kMaxGramItems], ro[kMaxGramItems], co[kMaxGramItems];
   // > 0: the row chunks are sized as for a launch of this many 64 x 32 half
   // items, so that a subset of a product set, launched on its own, sums each
   // entry in the order the whole set's launch does (0: this launch's items)
   int size_halves;
};
int gram_multi(ofrr_ctx *ctx, int64_t n, int k, int tb, int tz, GramMultiArgs &args);
// Y = sum_b V_b C_b
int basis_update(ofrr_ctx *ctx, int64_t n, int k, int nb, const void *const *V, const int *vdt,
                 const int *kb, const int64_t *vrs, const int64_t *vcs, const double *const *C,
                 const int *ldc, void *Y, int ydt, int64_t yrs, int64_t ycs);
// the device address of the allocation-free Gram product's scratch on
// `device` (which must be current), looked up once per device
int gram_scratch_init(int device);
// largest workspace of gram / gram_multi at any n (bytes)
size_t gram_workspace_bound(int num_cus);
// workspace of rr_truncated at basis dimension m (bytes)
size_t rr_workspace_bytes(int m);
// largest workspace a filtered solve over a CSR operator asks for (block width
// <= 128, basis dimension <= 384, any n): the serving slots reserve it once
size_t filtered_workspace_bound(int num_cus);
// truncated Rayleigh-Ritz (device pointers, FP64)
int rr_truncated(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                 double tau, int nkeep, double *theta, double *Y, int ldy, int32_t *rank);
int rr_generalized(ofrr_ctx *ctx, int m, const double *G, int ldg, const double *H, int ldh,
                   double gtol, double *theta, double *Y, int ldy);
// One job of a grouped truncated Rayleigh-Ritz: rr_truncated's arguments, and
// kx > 0 for the split form (the first kx basis columns are the previous Ritz
// block; rank -1 out when its Gram block is not positive definite).
constexpr int kMaxRRJobs = 16;
struct RRJob {
   const double *G, *H;
   int ldg, ldh, m, nkeep;
   double tau;
   double *theta, *Y;
   int ldy;
   int32_t *rank;
   int kx;
   // nullable: an event H is complete at; the job's stream waits for it
   // before its first read of H (the projected matrix may still be in flight
   // on another stream while the Gram-only stages run)
   hipEvent_t h_ready = nullptr;
};
// rr_truncated for each of njobs jobs (m <= 384 each) in one launch per stage on
// ctx's stream; each job's results equal rr_truncated's on its own arguments
int rr_truncated_group(ofrr_ctx *ctx, int njobs, const RRJob *jobs);
size_t rr_group_workspace_bytes(int njobs, int mx);
// the split form for jobs with kx > 0 (kx <= 128): results equal the oracle's
// orc_rr_split on each job's own arguments
int rr_split_group(ofrr_ctx *ctx, int njobs, const RRJob *jobs);
size_t rr_split_workspace_bytes(int njobs, int mx, int kxm);
// the jobs with kx = 0 through rr_truncated_group, the others through rr_split_group
int rr_group_run(ofrr_ctx *ctx, int njobs, const RRJob *jobs);
// argmax |x| (1-based index + signed value, device outputs)
int iamax(ofrr_ctx *ctx, int dtype, int64_t n, const void *x, int64_t inc, int32_t *idx1,
          double *val);

// per-column argmax |A(:,j)| with index base `base` (0 or 1) and signed value
int argmax_columns(ofrr_ctx *ctx, int dtype, int64_t m, int n, const void *A, int64_t rs,
                   int64_t cs, int base, int32_t *idx, double *val);
// C (R x nc, row-major) = alpha P Q + beta C with P R x K row-major and Q K x nc
// row-major: (BF16|FP16) x same -> FP32 accumulation, (FP32|FP64) x FP64 -> FP64
int dense_gemm(ofrr_ctx *ctx, int64_t R, int64_t K, int nc, const void *P, int pdt, int64_t ldp,
               const void *Q, int qdt, int64_t ldq, double alpha, double beta, void *C, int cdt,
               int64_t ldc);
// Y = X element-wise in another format / layout
int convert(ofrr_ctx *ctx, int64_t rows, int64_t cols, const void *X, int xdt, int64_t xrs,
            int64_t xcs, void *Y, int ydt, int64_t yrs, int64_t ycs);
// dense-operator epilogues: residual (R = W - X theta, quantised, norms) and the
// Chebyshev combination from T = A Dc
int resid_epilogue(ofrr_ctx *ctx, int64_t n, int k, const double *X, const double *th,
                   const double *W, void *Rq, int rdt, const double *inv_sR, float *colmax,
                   double *sums);
int cheb_combine(ofrr_ctx *ctx, int64_t n, int k, const void *T, int tdt, const void *Dc,
                 const void *Dp, const void *R, int bdt, const ChebArgs &args, void *Dn);
// max_i (a_ii + sum_{j != i} |a_ij|) of a dense symmetric FP64 matrix (synchronous)
int dense_gershgorin(ofrr_ctx *ctx, int64_t n, const double *A, int64_t lda, double *ub_host);
// out[j] = sum over b of part[b * width + j], fixed summation order
int reduce_partials(ofrr_ctx *ctx, int64_t nblk, int width, const double *part, double *out);

}  // namespace ofrr
