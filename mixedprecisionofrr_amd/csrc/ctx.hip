// Context, status strings and workspace of the gfx950 OFRR library.
// The context replaces the reference's CUDAHandler singleton
// (CPP/core/utils/cuda_handler.cu:9-45): it owns a stream, a grow-only device
// workspace and a pinned host staging buffer, one context per host thread.
#include "internal.hpp"

namespace ofrr {

int set_error(ofrr_ctx *ctx, int status, const std::string &msg) {
   if (ctx) ctx->last_error = msg;
   return status;
}

int hip_status(ofrr_ctx *ctx, hipError_t e, const char *what) {
   std::string m = std::string(what) + ": " + hipGetErrorString(e);
   int s = kErrorCUDAGeneral;
   if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) s = kErrorCUDAMemory;
   else if (e == hipErrorLaunchFailure || e == hipErrorInvalidConfiguration) s = kErrorCUDALaunch;
   return set_error(ctx, s, m);
}

void *workspace(ofrr_ctx *ctx, size_t bytes) {
   if (bytes == 0) bytes = 256;
   void *&ws = ctx->ws_on_aux ? ctx->ws_aux : ctx->ws;
   size_t &ws_bytes = ctx->ws_on_aux ? ctx->ws_aux_bytes : ctx->ws_bytes;
   if (ws_bytes >= bytes) return ws;
   if (ws) {
      // the old workspace may still be read by work enqueued on its stream
      if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
      (void)hipFree(ws);
      ws = nullptr;
      ws_bytes = 0;
   }
   size_t want = bytes + bytes / 4;
   if (hipMalloc(&ws, want) != hipSuccess) {
      ws = nullptr;
      return nullptr;
   }
   ws_bytes = want;
   return ws;
}

void *pinned(ofrr_ctx *ctx, size_t bytes) {
   if (ctx->pinned_bytes >= bytes) return ctx->pinned;
   if (ctx->pinned) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipHostFree(ctx->pinned);
      ctx->pinned = nullptr;
      ctx->pinned_bytes = 0;
   }
   if (hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
   ctx->pinned_bytes = bytes;
   return ctx->pinned;
}

size_t dtype_size(int dt) {
   switch (dt) {
      case OFRR_F64: return 8;
      case OFRR_F32: return 4;
      case OFRR_F16: return 2;
      case OFRR_BF16: return 2;
      case OFRR_FP8: return 1;
      default: return 0;
   }
}
bool dtype_valid(int dt) { return dtype_size(dt) != 0; }

}  // namespace ofrr

using namespace ofrr;

extern "C" {

int ofrr_version(void) { return 100; }

const char *ofrr_status_string(int s) {
   switch (s) {
      case kSuccess: return "Success";
      case kErrorInvalidArgument: return "Invalid argument";
      case kErrorInvalidOperation: return "Invalid operation";
      case kErrorOutOfMemory: return "Out of memory";
      case kErrorNotImplemented: return "Not implemented";
      case kErrorInternalError: return "Internal error";
      case kErrorNaN: return "NaN value detected";
      case kErrorInf: return "Inf value detected";
      case kErrorBLASInvalidLocation: return "Vectors in different locations";
      case kErrorBLASUnsupportedType: return "Unsupported data type combination";
      case kErrorBLASInvalidSize: return "Invalid size for operation";
      case kErrorFactorizationFailed: return "Factorization failed";
      case kErrorSingularMatrix: return "Singular matrix";
      case kErrorSolverDivergence: return "Solver divergence";
      case kErrorSolverMaxIterations: return "Solver reached maximum iterations";
      case kErrorCUDAGeneral: return "HIP runtime error";
      case kErrorCUDAMemory: return "HIP memory error";
      case kErrorCUDALaunch: return "HIP kernel launch error";
      case kErrorCUDASync: return "HIP synchronization error";
      default: return "Unknown status";
   }
}

int ofrr_ctx_create(int device, ofrr_ctx **out) {
   if (!out) return kErrorInvalidArgument;
   *out = nullptr;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return kErrorCUDAGeneral;
   if (device < 0 || device >= n) return kErrorInvalidArgument;
   if (hipSetDevice(device) != hipSuccess) return kErrorCUDAGeneral;
   // once per device, with it current: the scratch of ofrr_gram_noalloc
   if (gram_scratch_init(device) != kSuccess) return kErrorCUDAGeneral;
   ofrr_ctx *c = new ofrr_ctx();
   c->device = device;
   // a blocking stream: ordered with the null stream that frameworks default to
   if (hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault) != hipSuccess) {
      delete c;
      return kErrorCUDAGeneral;
   }
   c->stream = c->own_stream;
   hipDeviceProp_t prop;
   if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
   *out = c;
   return kSuccess;
}

int ofrr_ctx_destroy(ofrr_ctx *c) {
   if (!c) return kSuccess;
   (void)hipSetDevice(c->device);
   (void)hipStreamSynchronize(c->stream);
   for (auto &g : c->graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
   if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
   if (c->ev_join) (void)hipEventDestroy(c->ev_join);
   if (c->ev_rr_ready) (void)hipEventDestroy(c->ev_rr_ready);
   if (c->ev_rr_done) (void)hipEventDestroy(c->ev_rr_done);
   if (c->aux_stream) {
      (void)hipStreamSynchronize(c->aux_stream);
      (void)hipStreamDestroy(c->aux_stream);
   }
   if (c->ev_aux_fork) (void)hipEventDestroy(c->ev_aux_fork);
   if (c->ev_aux_join) (void)hipEventDestroy(c->ev_aux_join);
   if (c->ws) (void)hipFree(c->ws);
   if (c->ws_aux) (void)hipFree(c->ws_aux);
   if (c->pinned) (void)hipHostFree(c->pinned);
   if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
   delete c;
   return kSuccess;
}

int ofrr_ctx_sync(ofrr_ctx *c) {
   OFRR_CHECK_CTX(c);
   hipError_t e = hipStreamSynchronize(c->stream);
   if (e != hipSuccess) {
      hip_status(c, e, "hipStreamSynchronize");
      return kErrorCUDASync;
   }
   return kSuccess;
}

int ofrr_ctx_set_stream(ofrr_ctx *c, void *s) {
   OFRR_CHECK_CTX(c);
   c->stream = (hipStream_t)s;  // NULL selects the null (default) stream
   return kSuccess;
}

int ofrr_ctx_use_own_stream(ofrr_ctx *c) {
   OFRR_CHECK_CTX(c);
   c->stream = c->own_stream;
   return kSuccess;
}

void *ofrr_ctx_stream(ofrr_ctx *c) { return c ? (void *)c->stream : nullptr; }

const char *ofrr_ctx_last_error(ofrr_ctx *c) { return c ? c->last_error.c_str() : ""; }

}  // extern "C"
