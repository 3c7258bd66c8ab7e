// Serving loop: continuous batching of eigen requests on one GPU.
//
// The reference has no request loop; this is the caller of the filtered OFRR
// path (ofrr_filtered_eig), built for a stream of solves of differing n and k.
//
// A server owns `slots` workers.  Each worker is a host thread with its own
// context and non-blocking HIP stream, so the kernels of the solves running in
// different slots overlap on the device (the single-workgroup Rayleigh-Ritz
// kernels of one solve leave most CUs to the filter sweeps of the others).
// Requests wait in arrival order.  An idle worker admits the head of the queue
// when its device footprint fits in the byte budget beside the solves already
// running; a finished solve returns its memory and wakes the idle workers, so
// requests that arrived while others ran start as soon as a slot and memory
// are free.  Each request's memory (operator, start block, solver state) comes
// from a stream-ordered pool owned by the server: allocation and release are
// ordered on the worker's stream and never synchronise the other slots' work.
//
// Every reduction of a solve has a fixed order and the solves share no device
// memory, so a request's result is the same alone and beside other requests.
//
// Same-phase work of the running solves is grouped: a solve that reaches its
// truncated Rayleigh-Ritz hands the job (its Gram and projected matrices) to
// the server's Rayleigh-Ritz group, and the jobs that meet there (optionally
// within a window) run as one batched launch per stage on the stream of the
// solve that formed the group, ordered after each other job's inputs by an
// event on its slot's stream and before that slot's next work by another.
// Each job's arithmetic is the one it gets alone (rr_truncated_group, rr_split_group).
// On request (ofrr_server_opts.graphs) the filter sweeps of every slot run as
// captured HIP graphs, one per sweep shape, kept by the slot across requests.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

namespace ofrr {
namespace {

constexpr int kDefaultSlots = 4;
constexpr int kDefaultDegree = 120;
constexpr int kDefaultMaxIter = 200;

struct Plan {
   int k = 0, nd = 0, degree = 0, maxiter = 0, lo_dt = OFRR_F64;
   int refine = 0, refine_dtype = OFRR_F64, mid_cols = 0;
   double refine_tol = 0.0;
   int64_t n = 0, nnz = 0;
   size_t off_ci = 0, off_va = 0, off_lo = 0, off_x = 0, off_state = 0, total = 0;
   size_t state_bytes = 0;
};

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

// Block width (unless the request fixes it): the smallest multiple of 16 >=
// 1.25 nwant (at least 16, at most 128, and >= nwant), so the block carries a
// margin of unwanted directions beyond the wanted edge; a width that is not a
// power of two is stored padded to the next one.  Two correction blocks while
// the basis dimension k (1 + nd) stays <= 192, else one: up to 192 the
// Rayleigh-Ritz eigensolves keep their matrix in LDS.
int plan_request(const ofrr_eig_request &r, Plan &p) {
   if ((r.dim != 2 && r.dim != 3) || r.N < 4 || r.nwant < 1 || r.nwant > OFRR_SERVE_MAX_WANT ||
       !dtype_valid(r.basis_dtype) || !(r.tol > 0.0) || r.degree < 0 || r.maxiter < 0)
      return kErrorInvalidArgument;
   if (r.k < 0 || (r.k > 0 && (r.k < r.nwant || r.k > 128)) || r.nd < 0 || r.nd > 2 ||
       r.refine < -1 || r.refine > 1 || (r.refine == 1 && !dtype_valid(r.refine_dtype)) || r.mid_cols < 0)
      return kErrorInvalidArgument;
   p.n = r.dim == 2 ? r.N * r.N : r.N * r.N * r.N;
   if (p.n >= (int64_t)1 << 31) return kErrorBLASInvalidSize;
   int64_t nnz = 0;
   if (ofrr_stencil_nnz(r.dim, r.N, &nnz) != kSuccess) return kErrorInvalidArgument;
   p.nnz = nnz;
   if (r.k > 0) {
      p.k = r.k;
   } else {
      const int want = (5 * r.nwant + 3) / 4;
      p.k = std::max(r.nwant, std::min(128, std::max(16, (want + 15) / 16 * 16)));
   }
   if (p.k > p.n) return kErrorInvalidArgument;
   p.mid_cols = (r.mid_cols > 0 && r.mid_cols < p.k) ? r.mid_cols : 0;
   p.nd = r.nd > 0 ? r.nd : ((3 * p.k <= 192 || p.mid_cols) ? 2 : 1);
   if (p.k * p.nd + (p.nd == 2 && p.mid_cols ? p.mid_cols : p.k) > 384) return kErrorInvalidArgument;
   // precision ladder: FP8 corrections gain at most ~2^4 per sweep, so by
   // default an FP8 solve finishes in FP16 once the wanted residuals are below
   // 1e-4 or stall
   if (r.refine == 1) {
      p.refine = 1;
      p.refine_dtype = r.refine_dtype;
      p.refine_tol = r.refine_tol;
   } else if (r.refine == 0 && r.basis_dtype == OFRR_FP8) {
      p.refine = 1;
      p.refine_dtype = OFRR_F16;
      p.refine_tol = 1e-4;
   }
   p.degree = r.degree ? r.degree : kDefaultDegree;
   p.maxiter = r.maxiter ? r.maxiter : kDefaultMaxIter;
   // operator values of the growth products: FP64 for an FP64 basis, else a
   // copy whose rounding does not exceed FP32's (FP16 is exact for the
   // constant-coefficient Laplacian, FP32 keeps variable coefficients to 2^-24)
   if (r.basis_dtype == OFRR_F64) p.lo_dt = OFRR_F64;
   else p.lo_dt = (r.coef_seed == 0) ? OFRR_F16 : OFRR_F32;
   ofrr_filter_opts o{};
   o.k = p.k;
   o.nd = p.nd;
   o.degree = p.degree;
   o.basis_dtype = r.basis_dtype;
   o.refine = p.refine;
   o.refine_dtype = p.refine_dtype;
   o.mid_cols = p.mid_cols;
   int64_t sb = 0;
   if (ofrr_filtered_eig_state_bytes(p.n, &o, &sb) != kSuccess) return kErrorInvalidArgument;
   p.state_bytes = (size_t)sb;
   size_t off = al256((size_t)(p.n + 1) * 8);
   p.off_ci = off;
   off += al256((size_t)p.nnz * 4);
   p.off_va = off;
   off += al256((size_t)p.nnz * 8);
   p.off_lo = off;
   if (p.lo_dt != OFRR_F64) off += al256((size_t)p.nnz * dtype_size(p.lo_dt));
   p.off_x = off;
   off += al256((size_t)p.n * p.k * 8);
   p.off_state = off;
   off += al256(p.state_bytes);
   p.total = off;
   return kSuccess;
}

enum RecState { kQueued = 0, kRunning = 1, kDone = 2 };

struct Rec {
   ofrr_eig_request req{};
   Plan plan;
   ofrr_eig_result res{};
   int state = kQueued;
   int64_t footprint = 0;
};

}  // namespace
}  // namespace ofrr

using namespace ofrr;

namespace ofrr {
namespace {
// The grouped Rayleigh-Ritz of a server.  The first solve to hand in a job
// leads: it takes the jobs handed in so far (with a window, it first waits up
// to `window` or until every running solve has a job in), in arrival order,
// and launches them together; the others wait until their job is launched.
// The leader's own GPU work is still queued meanwhile: the host runs ahead of
// its stream, so a window costs the device nothing unless the job's inputs are
// ready before it ends.  Without a window (the default) a group holds the jobs
// that met while the previous group was being launched.
struct RRGroupExec {
   struct Pending {
      RRJob job;
      ofrr_ctx *slot = nullptr;
      int status = 0;
      bool launched = false;
   };
   std::chrono::microseconds window{0};
   int max_jobs = 1;
   std::atomic<int> *running = nullptr;  // solves running in the server
   std::mutex mu;
   std::condition_variable cv;
   std::deque<Pending *> pending;
   bool leader = false;
   int64_t groups = 0, group_jobs = 0;
   int group_max = 0;

   // the group runs on the leader's stream and workspace: a group of one is the
   // leader's own Rayleigh-Ritz, and groups of different leaders do not queue
   // behind each other
   int launch(ofrr_ctx *lead, std::vector<Pending *> &batch) {
      std::vector<RRJob> jobs;
      for (Pending *p : batch) {
         if (p->slot != lead) OFRR_HIP(lead, hipStreamWaitEvent(lead->stream, p->slot->ev_rr_ready, 0));
         jobs.push_back(p->job);
      }
      int st = rr_group_run(lead, (int)jobs.size(), jobs.data());
      if (st) return st;
      for (Pending *p : batch)
         if (p->slot != lead) OFRR_HIP(lead, hipEventRecord(p->slot->ev_rr_done, lead->stream));
      return kSuccess;
   }

   int submit(ofrr_ctx *slot, const RRJob &job) {
      if (!slot->ev_rr_ready) OFRR_HIP(slot, hipEventCreateWithFlags(&slot->ev_rr_ready, hipEventDisableTiming));
      if (!slot->ev_rr_done) OFRR_HIP(slot, hipEventCreateWithFlags(&slot->ev_rr_done, hipEventDisableTiming));
      OFRR_HIP(slot, hipEventRecord(slot->ev_rr_ready, slot->stream));
      Pending me;
      bool led = false;
      me.job = job;
      me.slot = slot;
      std::unique_lock<std::mutex> lk(mu);
      pending.push_back(&me);
      cv.notify_all();
      while (!me.launched) {
         if (leader) {
            cv.wait(lk, [&] { return me.launched || !leader; });
            continue;
         }
         leader = true;
         if (window.count() > 0)
            cv.wait_for(lk, window, [&] {
               return (int)pending.size() >= std::min(max_jobs, std::max(1, running->load()));
            });
         std::vector<Pending *> batch;
         while (!pending.empty() && (int)batch.size() < max_jobs) {
            batch.push_back(pending.front());
            pending.pop_front();
         }
         lk.unlock();
         led = true;
         const int st = launch(slot, batch);
         lk.lock();
         for (Pending *p : batch) {
            p->status = st;
            p->launched = true;
         }
         ++groups;
         group_jobs += (int64_t)batch.size();
         group_max = std::max(group_max, (int)batch.size());
         leader = false;
         cv.notify_all();
      }
      lk.unlock();
      if (me.status) return set_error(slot, me.status, "grouped Rayleigh-Ritz failed");
      if (!led) OFRR_HIP(slot, hipStreamWaitEvent(slot->stream, slot->ev_rr_done, 0));
      return kSuccess;
   }
   static int entry(void *arg, ofrr_ctx *slot, const void *job) {
      return static_cast<RRGroupExec *>(arg)->submit(slot, *static_cast<const RRJob *>(job));
   }
};
}  // namespace
}  // namespace ofrr

struct ofrr_server {
   int device = 0;
   int64_t budget = 0;
   hipMemPool_t pool = nullptr;
   std::mutex mu;
   std::condition_variable cv_admit, cv_done;
   std::deque<int64_t> queue;
   std::unordered_map<int64_t, Rec> recs;  // element references stay valid on rehash
   int64_t next_id = 1;
   bool stopping = false;
   int64_t submitted = 0, completed = 0, bytes_running = 0, bytes_peak = 0;
   int running = 0, running_peak = 0;
   std::chrono::steady_clock::time_point t0;
   std::vector<std::thread> workers;
   std::vector<int> worker_status;
   // device bytes the slots' contexts hold for the whole life of the server:
   // each reserves the largest workspace a solve asks for when it starts, so
   // admission counts request memory against budget - slot_bytes
   int64_t slot_bytes = 0;
   int64_t ws_regrowths = 0;  // solves whose workspace outgrew the slot's reservation
   bool graphs = false;       // filter sweeps as captured graphs, kept per slot and shape
   std::atomic<int> running_now{0};
   RRGroupExec *rr = nullptr;  // grouped Rayleigh-Ritz (null: each solve runs its own)

   double now() const {
      return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
   }
   // the head of the queue may start: its footprint fits beside the running solves
   bool head_fits() const {
      if (queue.empty()) return false;
      const Rec &r = recs.at(queue.front());
      return bytes_running + r.footprint <= budget - slot_bytes;
   }
   void run_one(ofrr_ctx *ctx, Rec &r);
   void worker(int slot);
};

void ofrr_server::run_one(ofrr_ctx *ctx, Rec &r) {
   const Plan &p = r.plan;
   const ofrr_eig_request &q = r.req;
   ofrr_eig_result &out = r.res;
   out.k = p.k;
   out.nd = p.nd;
   out.n = p.n;
   out.footprint = r.footprint;
   for (int j = 0; j < OFRR_SERVE_MAX_WANT; ++j) out.theta[j] = out.res[j] = __builtin_nan("");
   char *mem = nullptr;
   hipError_t e = hipMallocFromPoolAsync((void **)&mem, p.total, pool, ctx->stream);
   if (e != hipSuccess) {
      out.status = hip_status(ctx, e, "serve: hipMallocFromPoolAsync");
      return;
   }
   int64_t *rp = (int64_t *)mem;
   int32_t *ci = (int32_t *)(mem + p.off_ci);
   double *va = (double *)(mem + p.off_va);
   void *vlo = (p.lo_dt == OFRR_F64) ? nullptr : (void *)(mem + p.off_lo);
   double *X = (double *)(mem + p.off_x);
   int st = ofrr_stencil_csr(ctx, q.dim, q.N, q.coef_seed, rp, ci, OFRR_F64, va);
   if (!st && vlo) st = ofrr_convert(ctx, p.nnz, 1, va, OFRR_F64, 1, 1, vlo, p.lo_dt, 1, 1);
   if (!st) st = ofrr_fill_randn(ctx, OFRR_F64, p.n * p.k, X, 1, q.start_seed, 0, 0);
   if (!st) {
      ofrr_filter_opts o{};
      o.k = p.k;
      o.nwant = q.nwant;
      o.nd = p.nd;
      o.degree = p.degree;
      o.maxiter = p.maxiter;
      o.basis_dtype = q.basis_dtype;
      o.tol = q.tol;
      o.tau = 1e-12;
      o.lower = 0.0;  // the diffusion operators are positive definite
      o.normA = 0.0;  // ||A||_2 = the largest Ritz value
      o.lock = 1;     // later sweeps filter only the columns after the converged pairs
      o.refine = p.refine;
      o.refine_dtype = p.refine_dtype;
      o.refine_tol = p.refine_tol;
      o.mid_cols = p.mid_cols;
      o.graph = graphs ? 1 : 0;
      std::vector<double> th(p.k), rs(p.k);
      ofrr_filter_stats stats{};
      ctx->state_mem = mem + p.off_state;
      ctx->state_bytes = p.state_bytes;
      st = ofrr_filtered_eig(ctx, p.n, p.nnz, rp, ci, va, vlo, p.lo_dt, &o, X, th.data(), rs.data(),
                             &stats);
      ctx->state_mem = nullptr;
      ctx->state_bytes = 0;
      for (int j = 0; j < q.nwant; ++j) {
         out.theta[j] = th[j];
         out.res[j] = rs[j];
      }
      out.iterations = stats.iterations;
      out.converged = stats.converged;
      out.filter_steps = stats.filter_steps;
   }
   out.status = st;
   e = hipFreeAsync(mem, ctx->stream);
   if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
   if (e != hipSuccess && !out.status) out.status = hip_status(ctx, e, "serve: release");
}

void ofrr_server::worker(int slot) {
   ofrr_ctx *ctx = nullptr;
   hipStream_t s = nullptr;
   int st = ofrr_ctx_create(device, &ctx);
   if (!st && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) st = kErrorCUDAGeneral;
   if (!st) st = ofrr_ctx_set_stream(ctx, s);
   // the workspace every solve of this slot reuses, reserved once up front
   // (a slot that leads a grouped Rayleigh-Ritz runs the group in its workspace)
   size_t reserve = filtered_workspace_bound(ctx ? ctx->num_cus : 256);
   if (rr)
      reserve = std::max(reserve, std::max(rr_group_workspace_bytes(rr->max_jobs, 384),
                                           rr_split_workspace_bytes(rr->max_jobs, 384, 128)));
   if (!st && !workspace(ctx, reserve)) st = kErrorCUDAMemory;
   if (!st && rr) {
      ctx->rr_group = &RRGroupExec::entry;
      ctx->rr_group_arg = rr;
   }
   int64_t held = 0;
   {
      std::lock_guard<std::mutex> lk(mu);
      worker_status[slot] = st;
      if (!st) {
         held = (int64_t)ctx->ws_bytes;
         slot_bytes += held;
      }
   }
   cv_done.notify_all();
   for (;;) {
      std::unique_lock<std::mutex> lk(mu);
      cv_admit.wait(lk, [&] { return (stopping && queue.empty()) || (!st && head_fits()); });
      if (queue.empty() || st) break;
      const int64_t id = queue.front();
      queue.pop_front();
      Rec &r = recs.at(id);
      r.state = kRunning;
      r.res.t_admit = now();
      bytes_running += r.footprint;
      bytes_peak = std::max(bytes_peak, bytes_running);
      ++running;
      running_now.store(running);
      running_peak = std::max(running_peak, running);
      lk.unlock();
      cv_admit.notify_all();  // the next head may fit as well
      run_one(ctx, r);
      lk.lock();
      if ((int64_t)ctx->ws_bytes > held) {
         // the bound missed a workspace request: count the growth as held
         slot_bytes += (int64_t)ctx->ws_bytes - held;
         held = (int64_t)ctx->ws_bytes;
         ++ws_regrowths;
      }
      r.res.t_done = now();
      r.state = kDone;
      bytes_running -= r.footprint;
      --running;
      running_now.store(running);
      ++completed;
      lk.unlock();
      cv_done.notify_all();
      cv_admit.notify_all();
   }
   if (ctx) ofrr_ctx_destroy(ctx);
   if (s) (void)hipStreamDestroy(s);
}

extern "C" {

int ofrr_request_footprint(const ofrr_eig_request *req, int *k, int *nd, int64_t *bytes) {
   if (!req) return kErrorInvalidArgument;
   Plan p;
   const int st = plan_request(*req, p);
   if (st) return st;
   if (k) *k = p.k;
   if (nd) *nd = p.nd;
   if (bytes) *bytes = (int64_t)p.total;
   return kSuccess;
}

int ofrr_server_slot_bytes(int device, int64_t *bytes) {
   if (!bytes) return kErrorInvalidArgument;
   int cus = 0;
   if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
      return kErrorCUDAGeneral;
   const size_t b = filtered_workspace_bound(cus);
   *bytes = (int64_t)(b + b / 4);  // what workspace() allocates for it
   return kSuccess;
}

int ofrr_server_create(int device, const ofrr_server_opts *opt, ofrr_server **out) {
   if (!out) return kErrorInvalidArgument;
   *out = nullptr;
   const int slots = (opt && opt->slots > 0) ? opt->slots : kDefaultSlots;
   if (slots > 16) return kErrorInvalidArgument;
   int ndev = 0;
   if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return kErrorCUDAGeneral;
   if (device < 0 || device >= ndev) return kErrorInvalidArgument;
   if (hipSetDevice(device) != hipSuccess) return kErrorCUDAGeneral;
   ofrr_server *s = new ofrr_server();
   s->device = device;
   s->graphs = opt && opt->graphs > 0;
   s->t0 = std::chrono::steady_clock::now();
   if (opt && opt->budget > 0) {
      s->budget = opt->budget;
   } else {
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
         delete s;
         return kErrorCUDAGeneral;
      }
      s->budget = (int64_t)(0.9 * (double)fr);
   }
   hipMemPoolProps props{};
   props.allocType = hipMemAllocationTypePinned;
   props.handleTypes = hipMemHandleTypeNone;
   props.location.type = hipMemLocationTypeDevice;
   props.location.id = device;
   if (hipMemPoolCreate(&s->pool, &props) != hipSuccess) {
      delete s;
      return kErrorCUDAMemory;
   }
   // freed blocks stay in the pool for the next request instead of returning
   // to the driver at every synchronisation
   uint64_t keep = UINT64_MAX;
   (void)hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &keep);
   if (!(opt && opt->rr_group_us < 0) && slots > 1) {
      auto *g = new RRGroupExec();
      g->window = std::chrono::microseconds((opt && opt->rr_group_us > 0) ? opt->rr_group_us : 0);
      g->max_jobs = std::min(slots, kMaxRRJobs);
      g->running = &s->running_now;
      s->rr = g;
   }
   s->worker_status.assign(slots, -1);
   try {
      for (int i = 0; i < slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
   } catch (...) {
      {
         std::lock_guard<std::mutex> lk(s->mu);
         s->stopping = true;
      }
      s->cv_admit.notify_all();
      for (auto &t : s->workers) t.join();
      ofrr_server_destroy(s);
      return kErrorInternalError;
   }
   // every worker has its context and stream before the server is handed out
   int st = kSuccess;
   {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_done.wait(lk, [&] {
         for (int w : s->worker_status)
            if (w < 0) return false;
         return true;
      });
      for (int w : s->worker_status)
         if (w) st = w;
   }
   if (st) {
      ofrr_server_destroy(s);
      return st;
   }
   *out = s;
   return kSuccess;
}

int ofrr_server_submit(ofrr_server *s, const ofrr_eig_request *req, int64_t *id) {
   if (!s || !req || !id) return kErrorInvalidArgument;
   Rec r;
   r.req = *req;
   int st = plan_request(*req, r.plan);
   if (st) return st;
   r.footprint = (int64_t)r.plan.total;
   {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stopping) return kErrorInvalidOperation;
      if (r.footprint > s->budget - s->slot_bytes) return kErrorOutOfMemory;
      r.res.t_submit = s->now();
      *id = s->next_id++;
      s->recs.emplace(*id, r);
      s->queue.push_back(*id);
      ++s->submitted;
   }
   s->cv_admit.notify_all();
   return kSuccess;
}

int ofrr_server_poll(ofrr_server *s, int64_t id, ofrr_eig_result *out) {
   if (!s || !out) return kErrorInvalidArgument;
   std::lock_guard<std::mutex> lk(s->mu);
   auto it = s->recs.find(id);
   if (it == s->recs.end()) return kErrorInvalidArgument;
   if (it->second.state != kDone) return OFRR_PENDING;
   *out = it->second.res;
   s->recs.erase(it);
   return kSuccess;
}

int ofrr_server_wait(ofrr_server *s, int64_t id, ofrr_eig_result *out) {
   if (!s || !out) return kErrorInvalidArgument;
   std::unique_lock<std::mutex> lk(s->mu);
   auto it = s->recs.find(id);
   if (it == s->recs.end()) return kErrorInvalidArgument;
   // submissions may rehash the map while this thread waits: iterators are
   // invalidated then, references to the element are not
   const Rec &r = it->second;
   s->cv_done.wait(lk, [&] { return r.state == kDone; });
   *out = r.res;
   s->recs.erase(id);
   return kSuccess;
}

int ofrr_server_stats_get(ofrr_server *s, ofrr_server_stats *out) {
   if (!s || !out) return kErrorInvalidArgument;
   std::lock_guard<std::mutex> lk(s->mu);
   out->submitted = s->submitted;
   out->completed = s->completed;
   out->running = s->running;
   out->queued = (int64_t)s->queue.size();
   out->bytes_running = s->bytes_running;
   out->bytes_peak = s->bytes_peak;
   out->running_peak = s->running_peak;
   out->budget = s->budget;
   out->slot_bytes = s->slot_bytes;
   out->ws_regrowths = s->ws_regrowths;
   out->rr_groups = out->rr_group_jobs = 0;
   out->rr_group_max = 0;
   if (s->rr) {
      std::lock_guard<std::mutex> lg(s->rr->mu);
      out->rr_groups = s->rr->groups;
      out->rr_group_jobs = s->rr->group_jobs;
      out->rr_group_max = s->rr->group_max;
   }
   return kSuccess;
}

int ofrr_server_destroy(ofrr_server *s) {
   if (!s) return kSuccess;
   {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stopping = true;
   }
   s->cv_admit.notify_all();
   for (auto &t : s->workers)
      if (t.joinable()) t.join();
   (void)hipSetDevice(s->device);
   delete s->rr;
   if (s->pool) (void)hipMemPoolDestroy(s->pool);
   delete s;
   return kSuccess;
}

}  // extern "C"
