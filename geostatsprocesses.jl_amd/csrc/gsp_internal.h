// gsp_internal.h — shared internals of libgsp_hip.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "gsp.h"

// ----------------------------------------------------------------- errors ----
namespace gsp {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

struct Fail {
  int32_t code;
};

#define GSP_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::gsp::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                       __FILE__, __LINE__);                                        \
      throw ::gsp::Fail{e_ == hipErrorOutOfMemory ? GSP_ERR_MEMORY : GSP_ERR_DEVICE}; \
    }                                                                              \
  } while (0)

#define GSP_REQUIRE(cond, code, ...) \
  do {                               \
    if (!(cond)) {                   \
      ::gsp::set_error(__VA_ARGS__); \
      throw ::gsp::Fail{code};       \
    }                                \
  } while (0)

// Caching device allocator of one context. Freed blocks are kept by rounded
// size and handed out again; all work of a context runs in order on its one
// stream, so a block returned while a kernel still reads it is only reused by
// work queued after that kernel.
struct Pool {
  int device = 0;
  std::mutex mu;
  std::unordered_map<size_t, std::vector<void*>> cache;
  size_t cached = 0;
  static constexpr size_t kMaxCached = (size_t)32 << 30;
  static size_t round(size_t bytes) {
    if (bytes <= 256) return 256;
    if (bytes < ((size_t)1 << 20)) {
      size_t r = 512;
      while (r < bytes) r <<= 1;
      return r;
    }
    const size_t mb = (size_t)1 << 20;
    return (bytes + mb - 1) / mb * mb;
  }
  void* get(size_t rounded) {
    {
      std::lock_guard<std::mutex> lock(mu);
      auto it = cache.find(rounded);
      if (it != cache.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        cached -= rounded;
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, rounded);
    if (e == hipErrorOutOfMemory) {  // give the cache back and retry once
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, rounded);
    }
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu) failed: %s", rounded, hipGetErrorString(e));
      throw Fail{e == hipErrorOutOfMemory ? GSP_ERR_MEMORY : GSP_ERR_DEVICE};
    }
    return p;
  }
  void put(void* p, size_t rounded) {
    std::lock_guard<std::mutex> lock(mu);
    if (cached + rounded > kMaxCached) {
      (void)hipFree(p);
      return;
    }
    cache[rounded].push_back(p);
    cached += rounded;
  }
  void trim() {
    std::lock_guard<std::mutex> lock(mu);
    for (auto& kv : cache)
      for (void* p : kv.second) (void)hipFree(p);
    cache.clear();
    cached = 0;
  }
  ~Pool() { trim(); }
};

// The pool of the context whose call is running on this thread (set by
// gsp_guard); allocations made outside any call go straight to hipMalloc.
extern thread_local Pool* current_pool;

// A device allocation owned by host code (returned on destruction).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  Pool* pool = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      if (pool)
        pool->put(p, bytes);
      else
        (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
    pool = nullptr;
    bytes = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    if (current_pool) {
      bytes = Pool::round(count * sizeof(T));
      p = (T*)current_pool->get(bytes);
      pool = current_pool;
    } else {
      GSP_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    n = count;
  }
  // grow-only reuse for workspaces
  void reserve(size_t count) {
    if (count > n) alloc(count);
  }
};

}  // namespace gsp

struct gsp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) for work that runs beside the main
  // stream inside one call, forked from it with `fork`
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr;
  std::mutex mu;
  std::shared_ptr<gsp::Pool> pool;  // first to be built, last to go
  // reusable workspace
  gsp::DevBuf<double> X;        // sweep state, (N + n_data) x realizations
  gsp::DevBuf<double> stage_in;  // host-pointer staging
  gsp::DevBuf<double> stage_in2;
  gsp::DevBuf<double> stage_out;
  gsp::DevBuf<double> Zt;       // normals of one realization chunk, node-major
  gsp::DevBuf<int32_t> counters;  // [0] ticket, [1] error flag, [2..] scratch counters
  // FFTSIM transform plans by shape, reused across plans (gsp_fftsim.hip)
  std::shared_ptr<void> fft_plans;
  // kriging kernel chosen for systems of 17..24 and 25..32 neighbours, per
  // precision (FP64, FP32 + refinement): -1 until measured (gsp_seqsim.hip)
  int krige_variant[2][2] = {{-1, -1}, {-1, -1}};
  // kernel timing
  bool timing = false;
  struct Pending {
    hipEvent_t a, b;
    int slot;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_pool;
  gsp_timings totals{};
};

namespace gsp {
// Brackets one launch with events when timing is on; resolved by flush_timings
// after the stream has been synchronised.
struct TimedLaunch {
  gsp_ctx* ctx;
  int slot;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(gsp_ctx* c, int s);
  ~TimedLaunch();
};
void flush_timings(gsp_ctx* ctx);
// An event from the context's pool, and a measured interval [a, b] of work
// recorded on another stream (adds to the totals at the next flush).
hipEvent_t take_event(gsp_ctx* ctx);
void record_timing(gsp_ctx* ctx, hipEvent_t a, hipEvent_t b, int slot);
void release_timing_events(gsp_ctx* ctx);
}  // namespace gsp

// Runs `body` with the context's device current and its mutex held, mapping
// exceptions to status codes.
template <class F>
int32_t gsp_guard(gsp_ctx* ctx, F&& body) {
  if (!ctx) {
    gsp::set_error("null context");
    return GSP_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lock(ctx->mu);
  struct PoolScope {
    gsp::Pool* prev;
    explicit PoolScope(gsp::Pool* p) : prev(gsp::current_pool) { gsp::current_pool = p; }
    ~PoolScope() { gsp::current_pool = prev; }
  } scope(ctx->pool.get());
  try {
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
      gsp::set_error("hipSetDevice(%d): %s", ctx->device, hipGetErrorString(e));
      return GSP_ERR_DEVICE;
    }
    body();
    return GSP_OK;
  } catch (const gsp::Fail& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
}

// --------------------------------------------------------------- math ----
// The same IEEE operation sequence as the oracle (built with
// -ffp-contract=off on both sides), so FP64 plans can agree bit for bit.
namespace gsp {

__host__ __device__ inline double pow2i(int k) {
  return __builtin_bit_cast(double, (uint64_t)(k + 1023) << 52);
}

__host__ __device__ inline double ldexp_i(double p, int k) {
  if (k > 1023) return (p * pow2i(1023)) * pow2i(k - 1023);
  if (k < -1022) return (p * pow2i(k + 1000)) * pow2i(-1000);
  return p * pow2i(k);
}

// exp by Cody-Waite reduction (fdlibm split of ln2) and the degree-13 Taylor
// polynomial on |r| <= ln2/2, then scaling by 2^k.
__host__ __device__ inline double exp_cw(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_inf();
  if (x < -745.1332191019412) return 0.0;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double kd = __builtin_floor(x * inv_ln2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * ln2_hi) - kd * ln2_lo;
  double q = 1.0 / 6227020800.0;
  q = q * r + 1.0 / 479001600.0;
  q = q * r + 1.0 / 39916800.0;
  q = q * r + 1.0 / 3628800.0;
  q = q * r + 1.0 / 362880.0;
  q = q * r + 1.0 / 40320.0;
  q = q * r + 1.0 / 5040.0;
  q = q * r + 1.0 / 720.0;
  q = q * r + 1.0 / 120.0;
  q = q * r + 1.0 / 24.0;
  q = q * r + 1.0 / 6.0;
  q = q * r + 0.5;
  q = q * r + 1.0;
  q = q * r + 1.0;
  return ldexp_i(q, k);
}

// Variogram model on scaled coordinates: ranges already multiplied by alpha.
struct DevModel {
  int32_t family;
  int32_t ndim;
  int32_t has_rotation;
  int32_t pad_;
  double sill;
  double nugget;
  double irs[3];  // reciprocals of the scaled ranges
  double rot[9];  // row-major R
};

// Stabilising nugget of the gaussian model, added to its own nugget (in sill
// units): GeoStatsFunctions' predecessor Variography.jl adds 1e-6 "for
// numerical stability", and the reference's LUSIM testset (test/field.jl:29-56)
// factors Gaussian covariance matrices of unit grids that are singular in
// FP64 without it. The value is not fixed by any reference test.
constexpr double kGaussianNugget = 1e-6;

// C(u) = sill - gamma(u), gamma(0) = 0 (GeoStatsFunctions practical-range forms).
__host__ __device__ inline double cov_u(const DevModel& m, double u) {
  const double s = m.sill;
  double n = m.nugget;
  if (u == 0.0) return s;
  double shape;
  if (m.family == GSP_SPHERICAL) {
    shape = (u < 1.0) ? (1.5 * u - 0.5 * (u * u * u)) : 1.0;
  } else if (m.family == GSP_EXPONENTIAL) {
    shape = 1.0 - exp_cw(-3.0 * u);
  } else {
    n = n + kGaussianNugget;
    shape = 1.0 - exp_cw(-3.0 * (u * u));
  }
  double gamma = (s - n) * shape + n;
  return s - gamma;
}

// u = |diag(1/rs) R^T (a - b)|, rs the scaled ranges (multiplied by their
// reciprocals, computed once per model)
__host__ __device__ inline double lag(const DevModel& m, double a0, double a1, double a2,
                                      double b0, double b1, double b2) {
  const int nd = m.ndim;
  double dl0 = a0 - b0, dl1 = nd > 1 ? a1 - b1 : 0.0, dl2 = nd > 2 ? a2 - b2 : 0.0;
  double dp0 = dl0, dp1 = dl1, dp2 = dl2;
  if (m.has_rotation) {
    double dl[3] = {dl0, dl1, dl2}, dp[3];
    for (int e = 0; e < nd; ++e) {
      double acc = 0.0;
      for (int d = 0; d < nd; ++d) acc = acc + m.rot[d * 3 + e] * dl[d];
      dp[e] = acc;
    }
    dp0 = dp[0];
    dp1 = nd > 1 ? dp[1] : 0.0;
    dp2 = nd > 2 ? dp[2] : 0.0;
  }
  double t = dp0 * m.irs[0];
  double u2 = 0.0;
  u2 = u2 + t * t;
  if (nd > 1) {
    t = dp1 * m.irs[1];
    u2 = u2 + t * t;
  }
  if (nd > 2) {
    t = dp2 * m.irs[2];
    u2 = u2 + t * t;
  }
  return __builtin_sqrt(u2);
}

__host__ __device__ inline double cov_points(const DevModel& m, double a0, double a1, double a2,
                                             double b0, double b1, double b2) {
  return cov_u(m, lag(m, a0, a1, a2, b0, b1, b2));
}

DevModel make_dev_model(const gsp_model& m, int ndim, double alpha);
// Argument checks shared by the plan builders (number of cells of a valid grid).
int64_t validate_grid(const gsp_grid& g);
void validate_model(const gsp_model& m, int ndim);

// ------------------------------------------------------ seqsim internals ----
struct Offset {
  double d2;
  int16_t dx, dy, dz, pad_;
};

}  // namespace gsp

struct gsp_lusim_plan {
  std::shared_ptr<gsp::Pool> pool;  // declared first: outlives the buffers below
  int64_t N = 0, n_s = 0, n_d = 0;
  double mean = 0;
  gsp::DevBuf<double> L22;  // n_s x n_s, row-major lower
  gsp::DevBuf<double> d2;   // n_s
  gsp::DevBuf<int64_t> dinds, sinds;
  gsp::DevBuf<double> z1;
};

// kernels launched from gsp_seqsim.hip / gsp_lusim.hip
namespace gsp {
// surrogates: n_surrogates learned surrogates (none for an exact plan); a
// request takes its weights from the first one whose domain holds it
void seqsim_plan_build(gsp_ctx* ctx, int32_t n_req, const gsp_seqsim_problem* problems,
                       gsp_seqsim_plan* plan, const gsp_surrogate* const* surrogates,
                       int32_t n_surrogates, double tolerance);
void surrogate_train(gsp_ctx* ctx, int32_t n_plans, const gsp_seqsim_plan* const* plans,
                     const gsp_train_options* opt, gsp_surrogate* s, gsp_train_report* rep);
gsp_surrogate* surrogate_new(gsp_ctx* ctx);
void surrogate_delete(gsp_surrogate* s);
void surrogate_export(gsp_ctx* ctx, const gsp_surrogate* s, float* params, int64_t* count,
                      double* slots);
void surrogate_info(const gsp_surrogate* s, gsp_surrogate_info* out);
void seqsim_sweep_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                      const double* data_vals, double* out);
// Device memory of one sweep beyond its inputs and outputs: the node-major
// state X ((N + n_data) x RCmax doubles) and normals Zt (N x RCmax doubles);
// the R realizations are swept in chunks of at most RCmax (1..64).
struct SweepScratch {
  double* X = nullptr;
  double* Zt = nullptr;
  int RCmax = 64;
  bool timed = true;  // bracket the launches with timing events when enabled
  const double* mean = nullptr;  // per-request means replacing the plan's (multivariate sweeps)
};
// Device address of the allocation-free sweep's scratch on `device` (current),
// looked up once per device (gsp_noalloc.hip).
void noalloc_init(int device);
// The sweep on device pointers with caller-provided scratch: allocates nothing.
void seqsim_sweep_into(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                       const double* data_vals, double* out, const SweepScratch& scratch);
void seqsim_sweep_multivariate_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                   int32_t nv, const double* A, const double* mean,
                                   const double* z, const double* data_vals, double* out);
void seqsim_sweep_indicator_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                int32_t n_cat, const double* prob, const double* u,
                                const int32_t* data_cat, int32_t* out);
void lusim_factor_build(gsp_ctx* ctx, const gsp_lusim_problem* p, gsp_lusim_plan* plan);
void lusim_sample_run(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                      double* out);
void mfma_probe(gsp_ctx* ctx, int fp32, double* host_out);
void mfma_probe_bf16(gsp_ctx* ctx, double* host_out);
void copy_kernel(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes);
}  // namespace gsp
