// gsp_noalloc.hip — the SEQSIM sweep for callers that cannot allocate device
// memory: code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc. gsp_seqsim_sweep_noalloc runs the kernels of
// gsp_seqsim_sweep (device pointers) with their scratch — the node-major
// state and the transposed normals — in one fixed file-scope device array.
//
// Each GPU holds its own copy of `scratch` (64 MiB) from the first context
// created on it until the process ends. In host code the name `scratch` is a
// host-side stand-in, never a device address: its device address is looked
// up once per device by noalloc_init (with that device current) and only
// that address is passed to kernels.
#include <mutex>

#include "gsp_internal.h"
#include "gsp_seqsim.h"

namespace gsp {
namespace {
// Doubles of scratch per GPU: a sweep needs (2 N + n_data) x RC of them for
// chunks of RC realizations, so plans with 2 N + n_data <= 8,388,608 run
// (RC = 1 at that size; RC = 64 up to 65,536 words per realization).
constexpr int64_t kScratchWords = (int64_t)8 << 20;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
bool g_ready[kMaxDevices] = {};
}  // namespace
}  // namespace gsp

static __device__ __align__(256) double scratch[gsp::kScratchWords];

namespace gsp {

// Looks up this device's address of `scratch` once per device; the device
// must be current. Called by gsp_ctx_create.
void noalloc_init(int device) {
  GSP_REQUIRE(device >= 0 && device < kMaxDevices, GSP_ERR_UNSUPPORTED,
              "device %d beyond the %d devices of the allocation-free sweep", device,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ready[device]) return;
  void* p = nullptr;
  GSP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  GSP_REQUIRE(p != nullptr, GSP_ERR_DEVICE, "no device address for the sweep scratch");
  g_scratch[device] = static_cast<double*>(p);
  g_ready[device] = true;
}

namespace {
double* scratch_of(int device) {
  std::lock_guard<std::mutex> lock(g_mu);
  return (device >= 0 && device < kMaxDevices && g_ready[device]) ? g_scratch[device] : nullptr;
}
}  // namespace
}  // namespace gsp

extern "C" int32_t gsp_seqsim_sweep_noalloc(gsp_ctx* ctx, const gsp_seqsim_plan* plan,
                                            int32_t R, const double* z,
                                            const double* data_vals, double* out) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(z && out, GSP_ERR_ARGUMENT, "null z or output buffer");
    GSP_REQUIRE(plan->n_data == 0 || data_vals, GSP_ERR_ARGUMENT, "null data values");
    double* base = gsp::scratch_of(ctx->device);
    GSP_REQUIRE(base != nullptr, GSP_ERR_DEVICE,
                "the allocation-free sweep was not set up on device %d", ctx->device);
    const int64_t N = plan->N, nD = plan->n_data;
    // words per realization: X rows (N + n_data), Zt rows N, X padded to 32
    const int64_t per_real = 2 * N + nD + 32;
    const int64_t fit = gsp::kScratchWords / per_real;
    GSP_REQUIRE(fit >= 1, GSP_ERR_ARGUMENT,
                "plan too large for the allocation-free sweep: 2 N + n_data = %lld exceeds "
                "its %lld-word scratch (use gsp_seqsim_sweep)",
                (long long)(2 * N + nD), (long long)(gsp::kScratchWords - 32));
    const int RCmax = (int)std::min<int64_t>(std::min<int64_t>(fit, 64), R);
    gsp::SweepScratch w;
    w.RCmax = RCmax;
    w.X = base;
    const int64_t xw = ((N + nD) * RCmax + 31) / 32 * 32;
    w.Zt = base + xw;
    w.timed = false;  // timing events would be created on first use
    gsp::seqsim_sweep_into(ctx, plan, R, z, data_vals, out, w);
  });
}
