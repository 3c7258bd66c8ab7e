// Shared definitions of the gfx950 OFRR library: context, status codes,
// element-type traits and small device helpers.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ofrr.h"

// Status codes: the numeric values of the reference's MSVDStatus
// (CPP/core/utils/error_handling.hpp:24-62).
namespace ofrr {
enum Status : int {
   kSuccess = 0,
   kErrorInvalidArgument = 1000,
   kErrorInvalidOperation = 1001,
   kErrorOutOfMemory = 1002,
   kErrorNotImplemented = 1003,
   kErrorInternalError = 1004,
   kErrorNaN = 1005,
   kErrorInf = 1006,
   kErrorBLASInvalidLocation = 2000,
   kErrorBLASUnsupportedType = 2003,
   kErrorBLASInvalidSize = 2005,
   kErrorFactorizationFailed = 3000,
   kErrorSingularMatrix = 3001,
   kErrorSolverDivergence = 4000,
   kErrorSolverMaxIterations = 4001,
   kErrorCUDAGeneral = 6000,
   kErrorCUDAMemory = 6001,
   kErrorCUDALaunch = 6002,
   kErrorCUDASync = 6003,
};
}  // namespace ofrr

struct ofrr_ctx {
   int device = 0;
   hipStream_t stream = nullptr;      // stream in use
   hipStream_t own_stream = nullptr;  // stream created by the context
   int num_cus = 256;
   // grow-only device workspace for reduction partials
   void *ws = nullptr;
   size_t ws_bytes = 0;
   // pinned host staging for small synchronous results
   void *pinned = nullptr;
   size_t pinned_bytes = 0;
   // caller-owned device memory for the solver state of the next filtered
   // solve (the serving path); null: the grow-only per-thread arena
   void *state_mem = nullptr;
   size_t state_bytes = 0;
   // captured filter sweeps, one executable graph per sweep shape (size, width,
   // degree, formats); a later sweep of the same shape on other buffers (the
   // next request of a serving slot) updates the graph's kernel arguments
   // in place instead of instantiating it again
   struct SweepGraph {
      uint64_t key = 0, sig = 0, used = 0;
      hipGraphExec_t exec = nullptr;
   };
   std::vector<SweepGraph> graphs;
   uint64_t graph_clock = 0;
   // a graph cannot be recorded on the null stream: such a context replays
   // on own_stream, ordered with the caller's stream by these two events
   hipEvent_t ev_fork = nullptr, ev_join = nullptr;
   // grouped Rayleigh-Ritz (serving slots): when set, the filtered solver hands
   // its truncated Rayleigh-Ritz job to rr_group (which orders it after
   // ev_rr_ready on this stream and makes this stream wait for ev_rr_done)
   int (*rr_group)(void *arg, ofrr_ctx *slot, const void *job) = nullptr;
   void *rr_group_arg = nullptr;
   hipEvent_t ev_rr_ready = nullptr, ev_rr_done = nullptr;
   // a second stream of the context (the Hessenberg process runs each panel's
   // trailing update on it, beside the next panel's columns), ordered with
   // `stream` by events only; created on first use
   hipStream_t aux_stream = nullptr;
   hipEvent_t ev_aux_fork = nullptr, ev_aux_join = nullptr;
   // while ws_on_aux is set (work enqueued on aux_stream in place of stream),
   // workspace() hands out ws_aux, so the two streams never share scratch
   void *ws_aux = nullptr;
   size_t ws_aux_bytes = 0;
   bool ws_on_aux = false;
   std::string last_error;
};

namespace ofrr {

// While in scope, the library's operations on ctx enqueue on its second stream
// (created beforehand) and take their scratch from the second workspace.
struct OnAuxStream {
   ofrr_ctx *c;
   hipStream_t main;
   explicit OnAuxStream(ofrr_ctx *c_) : c(c_), main(c_->stream) {
      c->stream = c->aux_stream;
      c->ws_on_aux = true;
   }
   ~OnAuxStream() {
      c->stream = main;
      c->ws_on_aux = false;
   }
   OnAuxStream(const OnAuxStream &) = delete;
   OnAuxStream &operator=(const OnAuxStream &) = delete;
};

int set_error(ofrr_ctx *ctx, int status, const std::string &msg);
int hip_status(ofrr_ctx *ctx, hipError_t e, const char *what);
// device workspace of at least `bytes` (content undefined); nullptr on failure
void *workspace(ofrr_ctx *ctx, size_t bytes);
void *pinned(ofrr_ctx *ctx, size_t bytes);
size_t dtype_size(int dtype);
bool dtype_valid(int dtype);

#define OFRR_HIP(ctx, expr)                                          \
   do {                                                              \
      hipError_t e_ = (expr);                                        \
      if (e_ != hipSuccess) return ::ofrr::hip_status(ctx, e_, #expr); \
   } while (0)

#define OFRR_LAUNCH(ctx)                                                   \
   do {                                                                    \
      hipError_t e_ = hipGetLastError();                                   \
      if (e_ != hipSuccess) return ::ofrr::hip_status(ctx, e_, "launch");  \
   } while (0)

#define OFRR_CHECK_CTX(ctx) \
   do {                     \
      if (!(ctx)) return ::ofrr::kErrorInvalidArgument; \
   } while (0)

// ---------------------------------------------------------------------------
// element types
// ---------------------------------------------------------------------------
typedef uint8_t fp8_t;  // OCP FP8 E4M3 (bias 7, max 448, no infinities)

// OCP E4M3: software conversion (round to nearest even, saturate to +-448,
// NaN -> 0x7f).  Written from the format definition (OCP 8-bit Floating Point
// Specification v1.0, 2023).
__host__ __device__ inline float fp8_to_float(fp8_t b) {
   uint32_t s = (b >> 7) & 1u, e = (b >> 3) & 0xfu, m = b & 0x7u;
   float v;
   if (e == 0xf && m == 0x7) {
      v = __builtin_nanf("");
   } else if (e == 0) {
      v = (float)m * 0.001953125f;  // m * 2^-9
   } else {
      uint32_t bits = ((e + 120u) << 23) | (m << 20);  // 2^(e-7) * (1 + m/8)
      v = __builtin_bit_cast(float, bits);
   }
   return s ? -v : v;
}

__host__ __device__ inline fp8_t float_to_fp8(float f) {
   uint32_t bits = __builtin_bit_cast(uint32_t, f);
   uint32_t sign = (bits >> 24) & 0x80u;
   float a = __builtin_fabsf(f);
   if (!(a == a)) return (fp8_t)(sign | 0x7f);
   if (a >= 448.0f) return (fp8_t)(sign | 0x7e);  // saturate
   if (a < 0.0009765625f) {                        // below half the smallest subnormal (2^-10)
      return (fp8_t)sign;
   }
   uint32_t ab = __builtin_bit_cast(uint32_t, a);
   int32_t exp = (int32_t)((ab >> 23) & 0xff) - 127;
   if (exp < -6) {
      // subnormal: value = m * 2^-9, m in [0, 7]; round to nearest even
      float q = a * 512.0f;
      float r = __builtin_rintf(q);
      uint32_t m = (uint32_t)r;
      if (m >= 8) return (fp8_t)(sign | 0x08);  // rounds up to the smallest normal
      return (fp8_t)(sign | m);
   }
   // normal: keep 3 mantissa bits with round to nearest even
   uint32_t mant = ab & 0x7fffffu;
   uint32_t keep = mant >> 20, rest = mant & 0xfffffu;
   uint32_t half = 0x80000u;
   if (rest > half || (rest == half && (keep & 1u))) {
      keep += 1;
      if (keep == 8) { keep = 0; exp += 1; }
   }
   if (exp > 8 || (exp == 8 && keep > 6)) return (fp8_t)(sign | 0x7e);
   return (fp8_t)(sign | ((uint32_t)(exp + 7) << 3) | keep);
}

template <typename T> struct Elem;
template <> struct Elem<double> {
   using acc = double;
   static constexpr int id = OFRR_F64;
   __device__ static double load(double v) { return v; }
   __device__ static double store(double v) { return v; }
};
template <> struct Elem<float> {
   using acc = float;
   static constexpr int id = OFRR_F32;
   __device__ static float load(float v) { return v; }
   __device__ static float store(double v) { return (float)v; }
};
template <> struct Elem<__half> {
   using acc = float;
   static constexpr int id = OFRR_F16;
   __device__ static float load(__half v) { return __half2float(v); }
   __device__ static __half store(double v) { return __float2half((float)v); }
};
template <> struct Elem<__hip_bfloat16> {
   using acc = float;
   static constexpr int id = OFRR_BF16;
   __device__ static float load(__hip_bfloat16 v) { return __bfloat162float(v); }
   __device__ static __hip_bfloat16 store(double v) { return __float2bfloat16((float)v); }
};
// On the device the conversions are gfx950's v_cvt_f32_fp8 / v_cvt_pk_fp8_f32
// (after a clamp to +-448, the saturation of float_to_fp8); they agree with
// fp8_to_float / float_to_fp8 on every code and every finite float
// (ofrr_fp8_check).
template <> struct Elem<fp8_t> {
   using acc = float;
   static constexpr int id = OFRR_FP8;
   __device__ static float load(fp8_t v) { return __builtin_amdgcn_cvt_f32_fp8((int)v, 0); }
   __device__ static fp8_t store(double v) {
      float f = (float)v;
      f = (f == f) ? fminf(fmaxf(f, -448.0f), 448.0f) : f;  // saturate like float_to_fp8
      return (fp8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(f, 0.0f, 0, false) & 0xff);
   }
};

// largest finite magnitude of a storage type (quantisation range)
__host__ __device__ inline double dtype_max(int dt) {
   switch (dt) {
      case OFRR_F16: return 65504.0;
      case OFRR_FP8: return 448.0;
      case OFRR_BF16: return 3.0e38;
      case OFRR_F32: return 3.0e38;
      default: return 1.0e300;
   }
}

// load helper with conversion to double/float
template <typename T> __device__ inline double ld_d(const T *p) {
   return (double)Elem<T>::load(*p);
}

// block index remapped so that each of the 8 XCDs of the chip works on a
// contiguous range of logical blocks (hardware dispatches block b to XCD b % 8)
__device__ inline int64_t xcd_block(int64_t nblocks_logical) {
   int64_t b = blockIdx.x;
   int64_t per = ((int64_t)gridDim.x + 7) / 8;
   int64_t lb = (b % 8) * per + b / 8;
   return lb < nblocks_logical ? lb : -1;
}
inline int64_t xcd_grid(int64_t nblocks) { return ((nblocks + 7) / 8) * 8; }

// Strip schedule for row tiles of a banded operator whose long-range entries
// sit about one "plane" of `plane` tiles away (a 3D stencil: the +-N^2
// neighbours).  The tiles form chunks of `plane` tiles (a multiple of 8); each
// chunk is cut into 8 strips of plane / 8 tiles and XCD x takes strip x of
// every chunk, chunk after chunk.  Rows one plane apart are then processed
// plane / 8 blocks apart on the same XCD, so the second and third reads of a
// gathered row hit that XCD's L2 instead of returning to HBM (in row order they
// are `plane` blocks apart and miss).  Narrower strips (several per XCD) were
// measured slower on the 128^3 stencil: they lose more in-plane neighbours
// across strip edges than they gain in reuse distance.  plane <= 0 selects
// xcd_block.
__device__ inline int64_t sched_block(int64_t nblocks_logical, int64_t plane,
                                      bool reverse = false) {
   if (plane <= 0) return xcd_block(nblocks_logical);
   const int64_t b = blockIdx.x, j = b >> 3, S = plane >> 3;
   int64_t pl = j / S, w = j % S;
   if (reverse) {  // the same strips, walked from the last plane back to the first
      pl = (int64_t)(gridDim.x >> 3) / S - 1 - pl;
      w = S - 1 - w;
   }
   const int64_t t = pl * plane + (b & 7) * S + w;
   return t < nblocks_logical ? t : -1;
}
inline int64_t sched_grid(int64_t nblocks, int64_t plane) {
   if (plane <= 0) return xcd_grid(nblocks);
   return (nblocks + plane - 1) / plane * plane;
}
// plane size in tiles for an operator of bandwidth bw (max |col - row|): the
// strip schedule pays when a plane spans at least 16 tiles and the matrix
// holds at least two planes; 0 (row order) otherwise
inline int64_t plane_tiles(int64_t n, int64_t bw, int tile_rows) {
   const int64_t nblk = (n + tile_rows - 1) / tile_rows;
   int64_t p = (bw + tile_rows - 1) / tile_rows;
   p = (p + 7) / 8 * 8;
   return (p >= 16 && nblk >= 2 * p) ? p : 0;
}

// wave64 reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
   for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
   return v;
}

// non-negative float max through integer atomics (order independent).  Issued
// without reading the value first: a coherent read makes the workgroup wait for
// an L2 round trip before it retires (measured 245 -> 291 us per FP16 filter
// step), while the atomic itself is posted.
__device__ inline void atomic_max_nonneg(float *addr, float v) {
   atomicMax(reinterpret_cast<unsigned int *>(addr), __float_as_uint(v));
}

// dispatch a functor on a runtime dtype: f(T{})
// operator values: FP64, FP32 or FP16 only
template <typename F> inline int dispatch_value_dtype(int dt, F &&f) {
   switch (dt) {
      case OFRR_F64: return f(double{});
      case OFRR_F32: return f(float{});
      case OFRR_F16: return f(__half{});
      default: return kErrorBLASUnsupportedType;
   }
}
template <typename F> inline int dispatch_dtype(int dt, F &&f) {
   switch (dt) {
      case OFRR_F64: return f(double{});
      case OFRR_F32: return f(float{});
      case OFRR_F16: return f(__half{});
      case OFRR_BF16: return f(__hip_bfloat16{});
      case OFRR_FP8: return f(fp8_t{});
      default: return kErrorBLASUnsupportedType;
   }
}

}  // namespace ofrr
