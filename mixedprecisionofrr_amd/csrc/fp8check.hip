// Agreement of the library's device FP8 conversions (gfx950's v_cvt_f32_fp8
// and, after a clamp to +-448, v_cvt_pk_fp8_f32) with its software definition
// of OCP FP8 E4M3 (round to nearest even, saturating): all 256 codes and about
// a million floats spread over 2^-27 .. 2^9 in magnitude, both signs.
#include "internal.hpp"

namespace ofrr {

__device__ inline uint32_t mix32(uint32_t x) {
   x ^= x >> 16;
   x *= 0x7feb352dU;
   x ^= x >> 15;
   x *= 0x846ca68bU;
   x ^= x >> 16;
   return x;
}

// bad[0]: decode mismatches; bad[1]: encode mismatches with |f| <= 448;
// bad[2]: encode mismatches above 448 (saturation behaviour)
__global__ void fp8_check_kernel(int64_t nenc, unsigned long long *__restrict__ bad) {
   const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if (t < 256) {
      const float s = fp8_to_float((fp8_t)t);
      const float h = Elem<fp8_t>::load((fp8_t)t);  // the library's device decoder
      const bool same = (s != s) ? (h != h) : (__float_as_uint(s) == __float_as_uint(h));
      if (!same) atomicAdd(&bad[0], 1ull);
   }
   if (t < nenc) {
      const uint32_t x = mix32((uint32_t)t * 2654435761u + 12345u);
      const uint32_t e = 100u + (x >> 26) % 37u;  // 2^-27 .. 2^9
      const float f = __uint_as_float((x & 0x80000000u) | (e << 23) | (mix32(x) & 0x7fffffu));
      const fp8_t sw = float_to_fp8(f);
      const fp8_t hw = Elem<fp8_t>::store((double)f);  // the library's device encoder
      if (sw != hw) atomicAdd(&bad[fabsf(f) <= 448.0f ? 1 : 2], 1ull);
   }
}

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_fp8_check(ofrr_ctx *ctx, int64_t counts_host[3]) {
   OFRR_CHECK_CTX(ctx);
   if (!counts_host) return set_error(ctx, kErrorInvalidArgument, "fp8_check: null output");
   unsigned long long *bad = (unsigned long long *)workspace(ctx, 32);
   if (!bad) return set_error(ctx, kErrorCUDAMemory, "fp8_check: workspace");
   OFRR_HIP(ctx, hipMemsetAsync(bad, 0, 24, ctx->stream));
   const int64_t nenc = 1 << 20;
   fp8_check_kernel<<<(unsigned)(nenc / 256), 256, 0, ctx->stream>>>(nenc, bad);
   OFRR_LAUNCH(ctx);
   unsigned long long h[3] = {0, 0, 0};
   OFRR_HIP(ctx, hipMemcpyAsync(h, bad, 24, hipMemcpyDeviceToHost, ctx->stream));
   OFRR_HIP(ctx, hipStreamSynchronize(ctx->stream));
   for (int i = 0; i < 3; ++i) counts_host[i] = (int64_t)h[i];
   return kSuccess;
}

}  // extern "C"
