// gsp_api.cpp — the C ABI of libgsp_hip.so (declared in include/gsp.h).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>

#include "gsp_seqsim.h"
#include "gsp_testing.h"

namespace gsp {

static thread_local std::string g_last_error;
thread_local Pool* current_pool = nullptr;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

hipEvent_t take_event(gsp_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  GSP_HIP(hipEventCreate(&e));
  return e;
}

TimedLaunch::TimedLaunch(gsp_ctx* c, int s) : ctx(c), slot(s) {
  if (!ctx || !ctx->timing) return;
  a = take_event(ctx);
  b = take_event(ctx);
  GSP_HIP(hipEventRecord(a, ctx->stream));
}

TimedLaunch::~TimedLaunch() {
  if (!a) return;
  if (hipEventRecord(b, ctx->stream) == hipSuccess) {
    ctx->pending.push_back({a, b, slot});
  } else {
    ctx->event_pool.push_back(a);
    ctx->event_pool.push_back(b);
  }
}

void record_timing(gsp_ctx* ctx, hipEvent_t a, hipEvent_t b, int slot) {
  ctx->pending.push_back({a, b, slot});
}

void flush_timings(gsp_ctx* ctx) {
  std::vector<gsp_ctx::Pending> keep;
  for (auto& p : ctx->pending) {
    float ms = 0.0f;
    hipError_t e = hipEventElapsedTime(&ms, p.a, p.b);
    if (e == hipErrorNotReady) {
      keep.push_back(p);
      continue;
    }
    if (e == hipSuccess) {
      ctx->totals.ms[p.slot] += ms;
      ctx->totals.launches[p.slot] += 1;
    }
    ctx->event_pool.push_back(p.a);
    ctx->event_pool.push_back(p.b);
  }
  ctx->pending.swap(keep);
}

void release_timing_events(gsp_ctx* ctx) {
  for (auto& p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  ctx->pending.clear();
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->event_pool.clear();
}

}  // namespace gsp

using gsp::DevBuf;
using gsp::Fail;

extern "C" {

int32_t gsp_abi_version(void) { return GSP_ABI_VERSION; }

const char* gsp_last_error(void) { return gsp::g_last_error.c_str(); }

int32_t gsp_ctx_create(int32_t device, gsp_ctx** out) {
  if (!out) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    gsp::set_error("no HIP device available: %s", hipGetErrorString(e));
    return GSP_ERR_DEVICE;
  }
  if (device < 0 || device >= count) {
    gsp::set_error("device %d out of range (0..%d)", device, count - 1);
    return GSP_ERR_ARGUMENT;
  }
  gsp_ctx* ctx = new (std::nothrow) gsp_ctx();
  if (!ctx) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
  ctx->device = device;
  ctx->pool = std::make_shared<gsp::Pool>();
  ctx->pool->device = device;
  e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    gsp::set_error("context creation failed: %s", hipGetErrorString(e));
    delete ctx;
    return GSP_ERR_DEVICE;
  }
  try {
    gsp::noalloc_init(device);  // this device's address of the allocation-free sweep's scratch
  } catch (const gsp::Fail& f) {
    (void)hipEventDestroy(ctx->fork);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return f.code;
  }
  *out = ctx;
  return GSP_OK;
}

int32_t gsp_ctx_destroy(gsp_ctx* ctx) {
  if (!ctx) return GSP_OK;
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->X.release();
    ctx->stage_in.release();
    ctx->stage_in2.release();
    ctx->stage_out.release();
    ctx->Zt.release();
    ctx->counters.release();
    gsp::release_timing_events(ctx);
    ctx->fft_plans.reset();
    (void)hipStreamSynchronize(ctx->side);
    ctx->pool->trim();
    (void)hipEventDestroy(ctx->fork);
    (void)hipStreamDestroy(ctx->side);
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return GSP_OK;
}

int32_t gsp_ctx_stream(gsp_ctx* ctx, void** stream) {
  if (!ctx || !stream) {
    gsp::set_error("null argument");
    return GSP_ERR_ARGUMENT;
  }
  *stream = (void*)ctx->stream;
  return GSP_OK;
}

int32_t gsp_ctx_synchronize(gsp_ctx* ctx) {
  return gsp_guard(ctx, [&] {
    GSP_HIP(hipStreamSynchronize(ctx->stream));
    gsp::flush_timings(ctx);
  });
}

int32_t gsp_ctx_enable_timing(gsp_ctx* ctx, int32_t on) {
  return gsp_guard(ctx, [&] { ctx->timing = on != 0; });
}

int32_t gsp_ctx_timings(gsp_ctx* ctx, gsp_timings* out, int32_t reset) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(out, GSP_ERR_ARGUMENT, "null output");
    GSP_HIP(hipStreamSynchronize(ctx->stream));
    gsp::flush_timings(ctx);
    *out = ctx->totals;
    if (reset) ctx->totals = gsp_timings{};
  });
}

// ----------------------------------------------------------- device memory --
int32_t gsp_device_alloc(gsp_ctx* ctx, int64_t bytes, void** ptr) {
  if (!ptr) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *ptr = nullptr;
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(bytes >= 0, GSP_ERR_ARGUMENT, "negative size");
    GSP_HIP(hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 1)));
  });
}

int32_t gsp_device_free(gsp_ctx* ctx, void* ptr) {
  return gsp_guard(ctx, [&] {
    if (ptr) GSP_HIP(hipFree(ptr));
  });
}

int32_t gsp_memcpy_to_device(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), GSP_ERR_ARGUMENT, "bad copy");
    if (bytes) {
      GSP_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
      GSP_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
}

int32_t gsp_memcpy_to_host(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(bytes >= 0 && (bytes == 0 || (dst && src)), GSP_ERR_ARGUMENT, "bad copy");
    if (bytes) {
      GSP_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
      GSP_HIP(hipStreamSynchronize(ctx->stream));
    }
  });
}

int32_t gsp_selftest_copy(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(dst && src && bytes >= 0 && bytes % 16 == 0, GSP_ERR_ARGUMENT, "bad copy");
    gsp::copy_kernel(ctx, dst, src, bytes);
  });
}

int32_t gsp_selftest_mfma(gsp_ctx* ctx, int32_t fp32, double* out) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(out, GSP_ERR_ARGUMENT, "null output");
    GSP_REQUIRE(fp32 >= 0 && fp32 <= 2, GSP_ERR_ARGUMENT, "unknown probe");
    if (fp32 == 2)
      gsp::mfma_probe_bf16(ctx, out);
    else
      gsp::mfma_probe(ctx, fp32, out);
  });
}

// ------------------------------------------------------------------ SEQSIM --
int32_t gsp_seqsim_plan_create(gsp_ctx* ctx, const gsp_seqsim_problem* problem,
                               gsp_seqsim_plan** out) {
  return gsp_seqsim_plan_create_batch(ctx, 1, problem, out);
}

int32_t gsp_seqsim_plan_create_batch(gsp_ctx* ctx, int32_t n, const gsp_seqsim_problem* problems,
                                     gsp_seqsim_plan** out) {
  if (!out) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  gsp_seqsim_plan* plan = new (std::nothrow) gsp_seqsim_plan();
  if (!plan) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
  int32_t rc = gsp_guard(ctx, [&] {
    plan->pool = ctx->pool;
    gsp::seqsim_plan_build(ctx, n, problems, plan, nullptr, 0, 0.0);
  });
  if (rc != GSP_OK) {
    delete plan;
    return rc;
  }
  *out = plan;
  return GSP_OK;
}

int32_t gsp_seqsim_plan_info(const gsp_seqsim_plan* plan, gsp_seqsim_info* info) {
  if (!plan || !info) {
    gsp::set_error("null argument");
    return GSP_ERR_ARGUMENT;
  }
  info->n_nodes = plan->N;
  info->n_data = plan->n_data;
  info->minneighbors = plan->kmin;
  info->maxneighbors = plan->kmax;
  info->depth = plan->depth;
  info->pad_ = 0;
  info->alpha = plan->alpha;
  info->n_prior = plan->n_prior;
  info->n_fallback = plan->n_fallback;
  info->n_learned = plan->n_learned;
  return GSP_OK;
}

int32_t gsp_seqsim_plan_create_learned(gsp_ctx* ctx, int32_t n, const gsp_seqsim_problem* problems,
                                       const gsp_surrogate* surrogate, double tolerance,
                                       gsp_seqsim_plan** out) {
  return gsp_seqsim_plan_create_learned_multi(ctx, n, problems, 1, &surrogate, tolerance, out);
}

int32_t gsp_seqsim_plan_create_learned_multi(gsp_ctx* ctx, int32_t n,
                                             const gsp_seqsim_problem* problems,
                                             int32_t n_surrogates,
                                             const gsp_surrogate* const* surrogates,
                                             double tolerance, gsp_seqsim_plan** out) {
  if (!out || !surrogates || n_surrogates < 1 || n_surrogates > 64) {
    gsp::set_error("null output pointer, or not 1..64 surrogates");
    return GSP_ERR_ARGUMENT;
  }
  for (int32_t i = 0; i < n_surrogates; ++i)
    if (!surrogates[i]) {
      gsp::set_error("null surrogate %d", i);
      return GSP_ERR_ARGUMENT;
    }
  if (!(tolerance >= 0)) {
    gsp::set_error("tolerance must be non-negative");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  gsp_seqsim_plan* plan = new (std::nothrow) gsp_seqsim_plan();
  if (!plan) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
  int32_t rc = gsp_guard(ctx, [&] {
    plan->pool = ctx->pool;
    gsp::seqsim_plan_build(ctx, n, problems, plan, surrogates, n_surrogates, tolerance);
  });
  if (rc != GSP_OK) {
    delete plan;
    return rc;
  }
  *out = plan;
  return GSP_OK;
}

int32_t gsp_surrogate_train(gsp_ctx* ctx, int32_t n_plans, const gsp_seqsim_plan* const* plans,
                            const gsp_train_options* options, gsp_surrogate** out,
                            gsp_train_report* report) {
  if (!out) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  gsp_surrogate* s = nullptr;
  int32_t rc = gsp_guard(ctx, [&] {
    s = gsp::surrogate_new(ctx);
    gsp::surrogate_train(ctx, n_plans, plans, options, s, report);
  });
  if (rc != GSP_OK) {
    if (s) gsp::surrogate_delete(s);
    return rc;
  }
  *out = s;
  return GSP_OK;
}

int32_t gsp_surrogate_free(gsp_surrogate* surrogate) {
  if (surrogate) gsp::surrogate_delete(surrogate);
  return GSP_OK;
}

int32_t gsp_surrogate_export(gsp_ctx* ctx, const gsp_surrogate* surrogate, float* params,
                             int64_t* count, double* slots) {
  if (!surrogate) {
    gsp::set_error("null surrogate");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] { gsp::surrogate_export(ctx, surrogate, params, count, slots); });
}

int32_t gsp_surrogate_get_info(const gsp_surrogate* surrogate, gsp_surrogate_info* info) {
  if (!surrogate || !info) {
    gsp::set_error("null surrogate or info pointer");
    return GSP_ERR_ARGUMENT;
  }
  gsp::surrogate_info(surrogate, info);
  return GSP_OK;
}

int32_t gsp_seqsim_plan_layout(const gsp_seqsim_plan* plan, int32_t* n_requests,
                               int64_t* node_offset, int64_t* data_offset) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  if (n_requests) *n_requests = plan->n_req;
  for (size_t q = 0; q < plan->node_off.size(); ++q) {
    if (node_offset) node_offset[q] = plan->node_off[q];
    if (data_offset) data_offset[q] = plan->data_off[q];
  }
  return GSP_OK;
}

int32_t gsp_seqsim_plan_export(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t* nfound,
                               int32_t* nbr, double* lambda, double* sigma, uint8_t* flags,
                               int32_t* level) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    const size_t N = (size_t)plan->N, K = (size_t)plan->kmax;
    hipStream_t s = ctx->stream;
    if (nfound) GSP_HIP(hipMemcpyAsync(nfound, plan->nfound.p, N * 4, hipMemcpyDeviceToHost, s));
    if (nbr) GSP_HIP(hipMemcpyAsync(nbr, plan->nbr.p, N * K * 4, hipMemcpyDeviceToHost, s));
    if (lambda) GSP_HIP(hipMemcpyAsync(lambda, plan->lam.p, N * K * 8, hipMemcpyDeviceToHost, s));
    if (sigma) GSP_HIP(hipMemcpyAsync(sigma, plan->sigma.p, N * 8, hipMemcpyDeviceToHost, s));
    if (flags) GSP_HIP(hipMemcpyAsync(flags, plan->flags.p, N, hipMemcpyDeviceToHost, s));
    if (level) GSP_HIP(hipMemcpyAsync(level, plan->level.p, N * 4, hipMemcpyDeviceToHost, s));
    GSP_HIP(hipStreamSynchronize(s));
    gsp::flush_timings(ctx);
  });
}

int32_t gsp_seqsim_sweep(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                         const double* data_vals, double* out, int32_t device_pointers) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(z && out, GSP_ERR_ARGUMENT, "null z or output buffer");
    GSP_REQUIRE(plan->n_data == 0 || data_vals, GSP_ERR_ARGUMENT, "null data values");
    const size_t nz = (size_t)R * (size_t)plan->N;
    hipStream_t s = ctx->stream;
    if (device_pointers) {
      gsp::seqsim_sweep_run(ctx, plan, R, z, data_vals, out);
    } else {
      ctx->stage_in.reserve(nz);
      ctx->stage_in2.reserve((size_t)plan->n_data + 1);
      ctx->stage_out.reserve(nz);
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, z, nz * 8, hipMemcpyHostToDevice, s));
      if (plan->n_data > 0)
        GSP_HIP(hipMemcpyAsync(ctx->stage_in2.p, data_vals, (size_t)plan->n_data * 8,
                               hipMemcpyHostToDevice, s));
      gsp::seqsim_sweep_run(ctx, plan, R, ctx->stage_in.p, ctx->stage_in2.p, ctx->stage_out.p);
      GSP_HIP(hipMemcpyAsync(out, ctx->stage_out.p, nz * 8, hipMemcpyDeviceToHost, s));
    }
    if (!device_pointers) {
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}

int32_t gsp_seqsim_sweep_multivariate(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                      int32_t nvars, const double* A, const double* mean,
                                      const double* z, const double* data_vals, double* out,
                                      int32_t device_pointers) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(A && mean && z && out, GSP_ERR_ARGUMENT, "null argument");
    GSP_REQUIRE(plan->n_data == 0 || data_vals, GSP_ERR_ARGUMENT, "null data values");
    GSP_REQUIRE(nvars >= 1 && nvars <= 4, GSP_ERR_UNSUPPORTED,
                "multivariate SEQSIM supports 1 to 4 variables");
    hipStream_t s = ctx->stream;
    const size_t nz = (size_t)R * (size_t)plan->N * (size_t)nvars;
    const size_t nd = (size_t)plan->n_data * (size_t)nvars;
    if (device_pointers) {
      gsp::seqsim_sweep_multivariate_run(ctx, plan, R, nvars, A, mean, z, data_vals, out);
    } else {
      ctx->stage_in.reserve(nz);
      ctx->stage_in2.reserve(nd + 1);
      ctx->stage_out.reserve(nz);
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, z, nz * 8, hipMemcpyHostToDevice, s));
      if (nd > 0)
        GSP_HIP(hipMemcpyAsync(ctx->stage_in2.p, data_vals, nd * 8, hipMemcpyHostToDevice, s));
      gsp::seqsim_sweep_multivariate_run(ctx, plan, R, nvars, A, mean, ctx->stage_in.p,
                                         ctx->stage_in2.p, ctx->stage_out.p);
      GSP_HIP(hipMemcpyAsync(out, ctx->stage_out.p, nz * 8, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}

int32_t gsp_seqsim_sweep_indicator(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                   int32_t n_cat, const double* prob, const double* u,
                                   const int32_t* data_cat, int32_t* out,
                                   int32_t device_pointers) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(u && out && prob, GSP_ERR_ARGUMENT, "null prob, u or output buffer");
    GSP_REQUIRE(plan->n_data == 0 || data_cat, GSP_ERR_ARGUMENT, "null data categories");
    double sum = 0.0;
    for (int c = 0; c < n_cat && c < 64; ++c) {
      GSP_REQUIRE(prob[c] >= 0.0 && prob[c] <= 1.0, GSP_ERR_ASSERTION,
                  "probabilities must be in [0, 1]");
      sum += prob[c];
    }
    GSP_REQUIRE(std::fabs(sum - 1.0) <= 1e-8, GSP_ERR_ASSERTION,
                "probabilities must sum up to one");
    const size_t n = (size_t)R * (size_t)plan->N;
    hipStream_t s = ctx->stream;
    if (device_pointers) {
      gsp::seqsim_sweep_indicator_run(ctx, plan, R, n_cat, prob, u, data_cat, out);
    } else {
      ctx->stage_in.reserve(n);
      ctx->stage_out.reserve((n + 1) / 2);  // int32 output
      gsp::DevBuf<int32_t> dcat;
      if (plan->n_data > 0) {
        dcat.alloc((size_t)plan->n_data);
        GSP_HIP(hipMemcpyAsync(dcat.p, data_cat, (size_t)plan->n_data * 4, hipMemcpyHostToDevice,
                               s));
      }
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, u, n * 8, hipMemcpyHostToDevice, s));
      int32_t* dout = (int32_t*)ctx->stage_out.p;
      gsp::seqsim_sweep_indicator_run(ctx, plan, R, n_cat, prob, ctx->stage_in.p, dcat.p, dout);
      GSP_HIP(hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}

int32_t gsp_seqsim_plan_free(gsp_seqsim_plan* plan) {
  delete plan;
  return GSP_OK;
}

// ------------------------------------------------------------------- LUSIM --
int32_t gsp_lusim_factor(gsp_ctx* ctx, const gsp_lusim_problem* problem, gsp_lusim_plan** out) {
  if (!out) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  gsp_lusim_plan* plan = new (std::nothrow) gsp_lusim_plan();
  if (!plan) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
  int32_t rc = gsp_guard(ctx, [&] {
    plan->pool = ctx->pool;
    gsp::lusim_factor_build(ctx, problem, plan);
  });
  if (rc != GSP_OK) {
    delete plan;
    return rc;
  }
  *out = plan;
  return GSP_OK;
}

int32_t gsp_lusim_sample(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                         double* out, int32_t device_pointers) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(w && out, GSP_ERR_ARGUMENT, "null w or output buffer");
    hipStream_t s = ctx->stream;
    const size_t nw = (size_t)R * (size_t)plan->n_s, no = (size_t)R * (size_t)plan->N;
    if (device_pointers) {
      gsp::lusim_sample_run(ctx, plan, R, w, out);
    } else {
      ctx->stage_in.reserve(nw);
      ctx->stage_out.reserve(no);
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, w, nw * 8, hipMemcpyHostToDevice, s));
      gsp::lusim_sample_run(ctx, plan, R, ctx->stage_in.p, ctx->stage_out.p);
      GSP_HIP(hipMemcpyAsync(out, ctx->stage_out.p, no * 8, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}

int32_t gsp_lusim_dims(const gsp_lusim_plan* plan, int64_t* n_sim, int64_t* n_data) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  if (n_sim) *n_sim = plan->n_s;
  if (n_data) *n_data = plan->n_d;
  return GSP_OK;
}

int32_t gsp_lusim_export(gsp_ctx* ctx, const gsp_lusim_plan* plan, double* L22, double* d2) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)plan->n_s;
    if (L22 && n) GSP_HIP(hipMemcpyAsync(L22, plan->L22.p, n * n * 8, hipMemcpyDeviceToHost, s));
    if (d2 && n) GSP_HIP(hipMemcpyAsync(d2, plan->d2.p, n * 8, hipMemcpyDeviceToHost, s));
    GSP_HIP(hipStreamSynchronize(s));
    gsp::flush_timings(ctx);
  });
}

int32_t gsp_lusim_plan_free(gsp_lusim_plan* plan) {
  delete plan;
  return GSP_OK;
}

}  // extern "C"
