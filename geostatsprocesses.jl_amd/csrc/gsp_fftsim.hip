// gsp_fftsim.hip — FFTSIM (FFT moving average) on gfx950.
//
// The reference (src/simulation/field/fftsim.jl:54-156) builds the amplitude
// spectrum F = sqrt|fft(fftshift(C))| once and, for every realization, keeps
// F and takes the phase of the FFT of uniform white noise, transforms back,
// rescales to the sill and, with data, adds the kriging residual.
//
// The transforms run in rocFFT through hipFFT, real-to-complex and
// complex-to-real: the spectrum of a real field is Hermitian, so only the
// (nx/2 + 1) x ny x nz half is stored and multiplied, and the phase-times-
// amplitude product of two Hermitian spectra stays Hermitian. The element-wise
// work around the transforms is written here and fused where the data allow:
//   k_fft_cov     C at every position of the shifted array (fftshift folded
//                 into the index map; no C array before the shift)
//   k_fft_amp     F = sqrt|X|, F[0] = 0
//   k_fft_phase   P = F X/|X| for a batch of realizations, in place
//   k_sumsq / k_scale  per-realization sum of squares (fixed-order, two-level)
//                 and the factor sqrt(sill (N - 1) / sum Z^2), which also
//                 absorbs the 1/N of the inverse transform
//   k_fft_gather  raw values at the data cells (read before the in-place end)
//   k_fft_finish  s Z + mean, or zbar + s (Z - sum_j lambda_j Z[cell_j]) with
//                 data: the kriging of the unconditional values at the data
//                 cells, fused with the residual (the weights come from a
//                 data-only SEQSIM plan over the cell centroids).
#include <hipfft/hipfft.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "gsp_internal.h"
#include "gsp_seqsim.h"

#define GSP_FFT(call)                                                                       \
  do {                                                                                      \
    hipfftResult r_ = (call);                                                               \
    if (r_ != HIPFFT_SUCCESS) {                                                             \
      ::gsp::set_error("%s failed: hipFFT status %d (%s:%d)", #call, (int)r_, __FILE__,     \
                       __LINE__);                                                           \
      throw ::gsp::Fail{GSP_ERR_DEVICE};                                                    \
    }                                                                                       \
  } while (0)

namespace gsp {
// One hipFFT plan, destroyed with its last user.
struct FftHandle {
  hipfftHandle h = 0;
  ~FftHandle() {
    if (h) (void)hipfftDestroy(h);
  }
};
using FftRef = std::shared_ptr<FftHandle>;
}  // namespace gsp

struct gsp_fftsim_plan {
  std::shared_ptr<gsp::Pool> pool;  // declared first: outlives the buffers below
  int32_t nd = 0;
  int64_t dims[3] = {1, 1, 1};
  int64_t N = 0, Nh = 0;  // grid cells, half-spectrum entries
  double mean = 0, sill = 0;
  gsp::DevBuf<double> F;  // Nh amplitudes
  // transforms: batch of one, and batch of `nb` realizations (shared with
  // the context's cache, so plans of one grid shape are made once)
  int32_t nb = 1;
  gsp::FftRef fwd1, inv1, fwdb, invb;
  // conditioning
  int64_t n_cells = 0;
  gsp::DevBuf<double> zbar;     // N
  gsp::DevBuf<int64_t> cells;   // n_cells
  std::unique_ptr<gsp_seqsim_plan> kplan;  // kriging weights over the cell centroids
};

namespace gsp {
namespace {

struct FftGrid {
  int64_t n[3];
  int64_t c[3];  // zero-based index of the centre cell, dims .÷ 2 (1-based)
  double sp[3];
};

int fft_blocks(int64_t work) {
  int64_t g = (work + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 16384));
}

// Position k of fftshift(C) holds C[(k - n/2) mod n] per axis (fftshift
// rotates every axis by floor(n/2)), C[m] = C(x_centre, x_m).
__global__ __launch_bounds__(256) void k_fft_cov(const FftGrid G, const DevModel M, int64_t N,
                                                 double* __restrict__ out) {
  for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < N;
       k += (int64_t)gridDim.x * 256) {
    int64_t rest = k;
    double dl[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t kk = rest % G.n[d];
      rest /= G.n[d];
      int64_t m = kk - G.n[d] / 2;
      if (m < 0) m += G.n[d];
      dl[d] = (double)(m - G.c[d]) * G.sp[d];
    }
    out[k] = cov_points(M, dl[0], dl[1], dl[2], 0.0, 0.0, 0.0);
  }
}

__global__ __launch_bounds__(256) void k_fft_amp(int64_t Nh, const double2* __restrict__ X,
                                                 double* __restrict__ F) {
  for (int64_t h = (int64_t)blockIdx.x * 256 + threadIdx.x; h < Nh;
       h += (int64_t)gridDim.x * 256) {
    const double2 x = X[h];
    F[h] = h == 0 ? 0.0 : sqrt(hypot(x.x, x.y));
  }
}

// exp(i angle(X)) = X / |X| (1 where X = 0, angle(0) = 0), times F.
__global__ __launch_bounds__(256) void k_fft_phase(int64_t Nh, int64_t total,
                                                   const double* __restrict__ F,
                                                   double2* __restrict__ X) {
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * 256) {
    const double f = F[t % Nh];
    const double2 x = X[t];
    const double m = hypot(x.x, x.y);
    double2 p;
    if (m > 0.0) {
      p.x = f * (x.x / m);
      p.y = f * (x.y / m);
    } else {
      p.x = f;
      p.y = 0.0;
    }
    X[t] = p;
  }
}

constexpr int kSumBlocks = 256;  // partial sums per realization

// partial[r][b]: sum of Z[r][i]^2 over the b-th strided share of the cells
__global__ __launch_bounds__(256) void k_sumsq(int64_t N, const double* __restrict__ Z,
                                               double* __restrict__ partial) {
  __shared__ double s[256];
  const int r = blockIdx.y;
  const double* z = Z + (int64_t)r * N;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (int64_t)kSumBlocks * 256)
    acc = acc + z[i] * z[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)r * kSumBlocks + blockIdx.x] = s[0];
}

// scale[r] = sqrt(sill (N - 1) / S_r): Z / N has var(., mean=0) = S_r / (N^2 (N - 1))
__global__ __launch_bounds__(256) void k_scale(int64_t N, double sill,
                                               const double* __restrict__ partial,
                                               double* __restrict__ scale) {
  __shared__ double s[kSumBlocks];
  const int r = blockIdx.x;
  s[threadIdx.x] = partial[(int64_t)r * kSumBlocks + threadIdx.x];
  __syncthreads();
  for (int w = kSumBlocks / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[r] = sqrt(sill * (double)(N - 1) / s[0]);
}

__global__ __launch_bounds__(256) void k_fft_gather(int R, int64_t N, int64_t n_cells,
                                                    const int64_t* __restrict__ cells,
                                                    const double* __restrict__ Z,
                                                    double* __restrict__ zc) {
  const int64_t total = (int64_t)R * n_cells;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / n_cells, d = t - r * n_cells;
    zc[t] = Z[r * N + cells[d]];
  }
}

struct FinishArgs {
  int64_t N;
  int32_t R, K;
  double mean;
  const double* scale;
  // conditioning (zbar == nullptr: unconditional)
  const double* zbar;
  const int32_t* nfound;
  const int32_t* nbr;
  const double* lam;
  const uint8_t* flags;
  const double* zc;
  int64_t n_cells;
};

__global__ __launch_bounds__(256) void k_fft_finish(const FinishArgs A, double* __restrict__ Z) {
  const int64_t total = (int64_t)A.R * A.N;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / A.N, i = t - r * A.N;
    const double s = A.scale[r];
    const double z = Z[t];
    if (!A.zbar) {
      Z[t] = s * z + A.mean;
      continue;
    }
    if (A.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) {  // no kriging estimate: missing
      Z[t] = __builtin_nan("");
      continue;
    }
    const int n = A.nfound[i];
    const int32_t* nb = A.nbr + i * A.K;
    const double* lm = A.lam + i * A.K;
    const double* zr = A.zc + r * A.n_cells;
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = acc + lm[j] * zr[nb[j] - A.N];
    Z[t] = A.zbar[i] + s * (z - acc);
  }
}

// zbar = NaN where the data-only plan has no estimate (fewer than
// minneighbors data, or a failed factorisation): the reference's `missing`
__global__ __launch_bounds__(256) void k_mask_missing(int64_t N, const uint8_t* __restrict__ flags,
                                                      double* __restrict__ z) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * 256)
    if (flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) z[i] = __builtin_nan("");
}

// The context's transform plans, by (type, batch, dims). Making a rocFFT plan
// costs far more than running it on a mid-sized grid, so plans are kept for
// the life of the context (entries no plan uses are dropped past 32).
struct FftCache {
  std::map<std::array<int64_t, 5>, FftRef> plans;
};

FftRef get_plan(gsp_ctx* ctx, const gsp_fftsim_plan* P, hipfftType type, int batch) {
  if (!ctx->fft_plans) ctx->fft_plans = std::make_shared<FftCache>();
  FftCache& cache = *static_cast<FftCache*>(ctx->fft_plans.get());
  const std::array<int64_t, 5> key = {(int64_t)type, batch, P->dims[0], P->dims[1], P->dims[2]};
  auto it = cache.plans.find(key);
  if (it != cache.plans.end()) return it->second;
  if (cache.plans.size() >= 32)
    for (auto e = cache.plans.begin(); e != cache.plans.end();)
      e = e->second.use_count() == 1 ? cache.plans.erase(e) : std::next(e);
  int n[3];
  for (int d = 0; d < P->nd; ++d) n[d] = (int)P->dims[P->nd - 1 - d];  // slowest axis first
  auto ref = std::make_shared<FftHandle>();
  GSP_FFT(hipfftPlanMany(&ref->h, P->nd, n, nullptr, 1, 0, nullptr, 1, 0, type, batch));
  GSP_FFT(hipfftSetStream(ref->h, ctx->stream));
  cache.plans[key] = ref;
  return ref;
}

// A data-only SEQSIM problem (the k nearest data of every cell, simple
// kriging), host arrays.
gsp_seqsim_problem kriging_problem(const gsp_fftsim_problem* p, int64_t n, const double* coords,
                                   const int64_t* path, int64_t N) {
  gsp_seqsim_problem q;
  std::memset(&q, 0, sizeof(q));
  q.grid = p->grid;
  q.model = p->model;
  q.mean = p->mean;
  q.n_data = n;
  q.data_coords = coords;
  q.n_path = N;
  q.path = path;
  q.minneighbors = p->minneighbors;
  q.maxneighbors = (int32_t)std::min<int64_t>(p->maxneighbors, n);
  q.radius = p->radius;
  q.search_mode = GSP_DATA_ONLY;
  q.precision = GSP_FP64;
  q.device_pointers = 0;
  return q;
}

template <class T>
std::vector<T> to_host(gsp_ctx* ctx, const T* p, int64_t n, bool device) {
  std::vector<T> v((size_t)std::max<int64_t>(n, 1));
  if (n == 0) return v;
  if (device) {
    GSP_HIP(hipMemcpyAsync(v.data(), p, (size_t)n * sizeof(T), hipMemcpyDeviceToHost,
                           ctx->stream));
    GSP_HIP(hipStreamSynchronize(ctx->stream));
  } else {
    std::memcpy(v.data(), p, (size_t)n * sizeof(T));
  }
  return v;
}

void fftsim_build(gsp_ctx* ctx, const gsp_fftsim_problem* p, gsp_fftsim_plan* P) {
  GSP_REQUIRE(p, GSP_ERR_ARGUMENT, "null problem");
  hipStream_t s = ctx->stream;
  const gsp_grid& g = p->grid;
  const int64_t N = validate_grid(g);
  const int nd = g.ndim;
  validate_model(p->model, nd);
  GSP_REQUIRE(std::isfinite(p->mean), GSP_ERR_ARGUMENT, "invalid mean");
  for (int d = 0; d < nd; ++d)
    GSP_REQUIRE(g.dims[d] >= 2, GSP_ERR_ARGUMENT,
                "FFTSIM needs at least two cells along every axis (centre cell dims .÷ 2)");
  P->nd = nd;
  for (int d = 0; d < 3; ++d) P->dims[d] = d < nd ? g.dims[d] : 1;
  P->N = N;
  P->Nh = (P->dims[0] / 2 + 1) * P->dims[1] * P->dims[2];
  P->mean = p->mean;
  P->sill = p->model.sill;
  GSP_REQUIRE(N <= ((int64_t)1 << 31) - 1, GSP_ERR_UNSUPPORTED, "grid too large for FFTSIM");

  // transform plans: one realization, and batches within 512 MiB of spectrum
  const int64_t per = P->Nh * 16;
  P->nb = (int32_t)std::max<int64_t>(1, std::min<int64_t>(64, ((int64_t)512 << 20) / per));
  P->fwd1 = get_plan(ctx, P, HIPFFT_D2Z, 1);
  P->inv1 = get_plan(ctx, P, HIPFFT_Z2D, 1);
  if (P->nb > 1) {
    P->fwdb = get_plan(ctx, P, HIPFFT_D2Z, P->nb);
    P->invb = get_plan(ctx, P, HIPFFT_Z2D, P->nb);
  }

  // amplitude spectrum
  FftGrid G;
  for (int d = 0; d < 3; ++d) {
    G.n[d] = P->dims[d];
    G.c[d] = d < nd ? P->dims[d] / 2 - 1 : 0;
    G.sp[d] = d < nd ? g.spacing[d] : 1.0;
  }
  const DevModel M = make_dev_model(p->model, nd, 1.0);
  DevBuf<double> C;
  DevBuf<double2> X;
  C.alloc((size_t)N);
  X.alloc((size_t)P->Nh);
  P->F.alloc((size_t)P->Nh);
  {
    TimedLaunch tl(ctx, GSP_K_SPECTRAL);
    hipLaunchKernelGGL(k_fft_cov, dim3(fft_blocks(N)), dim3(256), 0, s, G, M, N, C.p);
  }
  {
    TimedLaunch tl(ctx, GSP_K_FFT);
    GSP_FFT(hipfftExecD2Z(P->fwd1->h, C.p, (hipfftDoubleComplex*)X.p));
  }
  {
    TimedLaunch tl(ctx, GSP_K_SPECTRAL);
    hipLaunchKernelGGL(k_fft_amp, dim3(fft_blocks(P->Nh)), dim3(256), 0, s, P->Nh, X.p, P->F.p);
  }

  // conditioning: kriging mean of the data, weights of the data cells
  if (p->n_data > 0 || p->n_cells > 0) {
    GSP_REQUIRE(p->n_data > 0 && p->n_cells > 0, GSP_ERR_ARGUMENT,
                "conditioning needs both the data and the cells they were placed on");
    GSP_REQUIRE(p->data_coords && p->data_vals && p->cells, GSP_ERR_ARGUMENT,
                "null conditioning arrays");
    const bool dev = p->device_pointers != 0;
    const std::vector<double> coords = to_host(ctx, p->data_coords, p->n_data * nd, dev);
    const std::vector<double> vals = to_host(ctx, p->data_vals, p->n_data, dev);
    const std::vector<int64_t> cells = to_host(ctx, p->cells, p->n_cells, dev);
    for (int64_t d = 0; d < p->n_cells; ++d)
      GSP_REQUIRE(cells[(size_t)d] >= 0 && cells[(size_t)d] < N, GSP_ERR_ARGUMENT,
                  "data cell %lld out of range", (long long)cells[(size_t)d]);
    std::vector<int64_t> path((size_t)N);
    for (int64_t i = 0; i < N; ++i) path[(size_t)i] = i;

    // zbar: the data-only plan swept once with zero normals
    gsp_seqsim_problem q1 = kriging_problem(p, p->n_data, coords.data(), path.data(), N);
    std::unique_ptr<gsp_seqsim_plan> plan1(new gsp_seqsim_plan());
    plan1->pool = ctx->pool;
    seqsim_plan_build(ctx, 1, &q1, plan1.get(), nullptr, 0, 0.0);
    DevBuf<double> zeros, dvals;
    zeros.alloc((size_t)N);
    dvals.alloc((size_t)p->n_data);
    GSP_HIP(hipMemsetAsync(zeros.p, 0, (size_t)N * 8, s));
    GSP_HIP(hipMemcpyAsync(dvals.p, vals.data(), (size_t)p->n_data * 8, hipMemcpyHostToDevice,
                           s));
    P->zbar.alloc((size_t)N);
    seqsim_sweep_run(ctx, plan1.get(), 1, zeros.p, dvals.p, P->zbar.p);
    hipLaunchKernelGGL(k_mask_missing, dim3(fft_blocks(N)), dim3(256), 0, s, N,
                       plan1->flags.p, P->zbar.p);

    // the cells' centroids as kriging data
    std::vector<double> cc((size_t)(p->n_cells * nd));
    for (int64_t d = 0; d < p->n_cells; ++d) {
      int64_t rest = cells[(size_t)d];
      for (int a = 0; a < nd; ++a) {
        const int64_t k = rest % g.dims[a];
        rest /= g.dims[a];
        cc[(size_t)(d * nd + a)] = g.origin[a] + ((double)k + 0.5) * g.spacing[a];
      }
    }
    gsp_seqsim_problem q2 = kriging_problem(p, p->n_cells, cc.data(), path.data(), N);
    P->kplan.reset(new gsp_seqsim_plan());
    P->kplan->pool = ctx->pool;
    seqsim_plan_build(ctx, 1, &q2, P->kplan.get(), nullptr, 0, 0.0);
    P->n_cells = p->n_cells;
    P->cells.alloc((size_t)p->n_cells);
    GSP_HIP(hipMemcpyAsync(P->cells.p, cells.data(), (size_t)p->n_cells * 8,
                           hipMemcpyHostToDevice, s));
  }
  GSP_HIP(hipStreamSynchronize(s));  // host staging ends with the call
  flush_timings(ctx);
}

void fftsim_sample_run(gsp_ctx* ctx, const gsp_fftsim_plan* P, int32_t R, const double* w,
                       double* out) {
  hipStream_t s = ctx->stream;
  const int64_t N = P->N, Nh = P->Nh;
  DevBuf<double2> X;
  X.alloc((size_t)(std::min<int32_t>(R, P->nb) * Nh));
  for (int32_t r0 = 0; r0 < R;) {
    const bool full = P->nb > 1 && R - r0 >= P->nb;
    const int32_t b = full ? P->nb : 1;
    {
      TimedLaunch tl(ctx, GSP_K_FFT);
      // forward real-to-complex transforms leave their input unchanged
      GSP_FFT(hipfftExecD2Z((full ? P->fwdb : P->fwd1)->h, const_cast<double*>(w + r0 * N),
                            (hipfftDoubleComplex*)X.p));
    }
    {
      TimedLaunch tl(ctx, GSP_K_SPECTRAL);
      hipLaunchKernelGGL(k_fft_phase, dim3(fft_blocks(b * Nh)), dim3(256), 0, s, Nh,
                         (int64_t)b * Nh, P->F.p, X.p);
    }
    {
      TimedLaunch tl(ctx, GSP_K_FFT);
      GSP_FFT(hipfftExecZ2D((full ? P->invb : P->inv1)->h, (hipfftDoubleComplex*)X.p, out + r0 * N));
    }
    r0 += b;
  }
  DevBuf<double> partial, scale, zc;
  partial.alloc((size_t)R * kSumBlocks);
  scale.alloc((size_t)R);
  FinishArgs A;
  std::memset(&A, 0, sizeof(A));
  A.N = N;
  A.R = R;
  A.mean = P->mean;
  A.scale = scale.p;
  {
    TimedLaunch tl(ctx, GSP_K_SPECTRAL);
    hipLaunchKernelGGL(k_sumsq, dim3(kSumBlocks, R), dim3(256), 0, s, N, out, partial.p);
    hipLaunchKernelGGL(k_scale, dim3(R), dim3(kSumBlocks), 0, s, N, P->sill, partial.p, scale.p);
    if (P->kplan) {
      zc.alloc((size_t)R * P->n_cells);
      hipLaunchKernelGGL(k_fft_gather, dim3(fft_blocks((int64_t)R * P->n_cells)), dim3(256), 0,
                         s, R, N, P->n_cells, P->cells.p, out, zc.p);
      A.K = P->kplan->kmax;
      A.zbar = P->zbar.p;
      A.nfound = P->kplan->nfound.p;
      A.nbr = P->kplan->nbr.p;
      A.lam = P->kplan->lam.p;
      A.flags = P->kplan->flags.p;
      A.zc = zc.p;
      A.n_cells = P->n_cells;
    }
    hipLaunchKernelGGL(k_fft_finish, dim3(fft_blocks((int64_t)R * N)), dim3(256), 0, s, A, out);
  }
}

}  // namespace
}  // namespace gsp

extern "C" {

int32_t gsp_fftsim_plan_create(gsp_ctx* ctx, const gsp_fftsim_problem* problem,
                               gsp_fftsim_plan** out) {
  if (!out) {
    gsp::set_error("null output pointer");
    return GSP_ERR_ARGUMENT;
  }
  *out = nullptr;
  gsp_fftsim_plan* plan = new (std::nothrow) gsp_fftsim_plan();
  if (!plan) {
    gsp::set_error("host allocation failed");
    return GSP_ERR_MEMORY;
  }
  int32_t rc = gsp_guard(ctx, [&] {
    plan->pool = ctx->pool;
    gsp::fftsim_build(ctx, problem, plan);
  });
  if (rc != GSP_OK) {
    delete plan;
    return rc;
  }
  *out = plan;
  return GSP_OK;
}

int32_t gsp_fftsim_sample(gsp_ctx* ctx, const gsp_fftsim_plan* plan, int32_t R, const double* w,
                          double* out, int32_t device_pointers) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1, GSP_ERR_ARGUMENT, "number of realizations must be positive");
    GSP_REQUIRE(w && out, GSP_ERR_ARGUMENT, "null w or output buffer");
    hipStream_t s = ctx->stream;
    const size_t n = (size_t)R * (size_t)plan->N;
    if (device_pointers) {
      gsp::fftsim_sample_run(ctx, plan, R, w, out);
    } else {
      ctx->stage_in.reserve(n);
      ctx->stage_out.reserve(n);
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, w, n * 8, hipMemcpyHostToDevice, s));
      gsp::fftsim_sample_run(ctx, plan, R, ctx->stage_in.p, ctx->stage_out.p);
      GSP_HIP(hipMemcpyAsync(out, ctx->stage_out.p, n * 8, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}

int32_t gsp_fftsim_export(gsp_ctx* ctx, const gsp_fftsim_plan* plan, double* amplitude,
                          double* zbar) {
  if (!plan) {
    gsp::set_error("null plan");
    return GSP_ERR_ARGUMENT;
  }
  return gsp_guard(ctx, [&] {
    hipStream_t s = ctx->stream;
    if (amplitude)
      GSP_HIP(hipMemcpyAsync(amplitude, plan->F.p, (size_t)plan->Nh * 8, hipMemcpyDeviceToHost,
                             s));
    if (zbar) {
      GSP_REQUIRE(plan->kplan, GSP_ERR_ARGUMENT, "unconditional plan: no kriging mean");
      GSP_HIP(hipMemcpyAsync(zbar, plan->zbar.p, (size_t)plan->N * 8, hipMemcpyDeviceToHost, s));
    }
    GSP_HIP(hipStreamSynchronize(s));
    gsp::flush_timings(ctx);
  });
}

int32_t gsp_fftsim_plan_free(gsp_fftsim_plan* plan) {
  delete plan;
  return GSP_OK;
}

}  // extern "C"
