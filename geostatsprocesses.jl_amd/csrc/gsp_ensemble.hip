// gsp_ensemble.hip — per-cell statistics across the realizations of an
// ensemble (ensembles.jl:42-52,76-85) on gfx950.
//
// The samplers leave an ensemble in HBM as R x N, realization-major, so the
// R values of one cell sit N apart and the values of neighbouring cells are
// adjacent: one thread per cell reads row after row, and every row read of a
// wavefront is one contiguous 512-byte segment (rows are loaded eight ahead
// of their accumulation, which stays in row order). All requested statistics
// come out of one pass over the R x N values: the mean's running sum,
// Welford's recurrence for the variance and the threshold counts of every
// cdf / ccdf are updated per row; quantiles need order statistics, so when
// any is requested the cell's values are also kept in LDS (R <= 128; beyond
// that the selection re-reads them from memory, where L2 holds a block's
// 64 x R values). Selection counts ranks (R^2 comparisons per group of four
// quantiles) instead of sorting: no per-thread array that would spill.
#include <cmath>

#include "gsp_internal.h"

namespace gsp {
namespace {

constexpr int kMaxStats = 32;
constexpr int kTile = 64;    // cells per block: one wavefront
constexpr int kStage = 128;  // most realizations kept in LDS (64 KB per block)
constexpr int kBatch = 8;    // rows loaded ahead of their (sequential) accumulation
constexpr int kQGroup = 4;   // quantiles resolved per rank pass

struct StatArgs {
  int64_t N;
  int32_t R, n_stats;
  int32_t any_count;  // some cdf / ccdf requested
  int32_t n_quant;    // quantiles requested
  int32_t kind[kMaxStats];
  double x[kMaxStats];  // cdf / ccdf threshold
  // quantiles, compacted: output row, 0-based lower order statistic, weight
  int32_t q_row[kMaxStats];
  int32_t q_j[kMaxStats];
  double q_g[kMaxStats];
};

template <bool STAGE>
__global__ __launch_bounds__(64) void k_ens_stats(const StatArgs A,
                                                  const double* __restrict__ reals,
                                                  double* __restrict__ out) {
  extern __shared__ double s_vals[];  // STAGE: [R][kTile]
  const int t = threadIdx.x;
  const int R = A.R;
  for (int64_t base = (int64_t)blockIdx.x * kTile; base < A.N;
       base += (int64_t)gridDim.x * kTile) {
    const int64_t i = base + t;
    if (i >= A.N) break;  // only the last tile is partial; no barrier below
    // one pass: sum (mean), Welford (var), threshold counts (cdf / ccdf),
    // rows accumulated in order
    double sum = 0.0, M = 0.0, S = 0.0;
    int cnt[kMaxStats];
#pragma unroll
    for (int q = 0; q < kMaxStats; ++q) cnt[q] = 0;
    for (int r0 = 0; r0 < R; r0 += kBatch) {
      double vb[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        vb[u] = r0 + u < R ? reals[(int64_t)(r0 + u) * A.N + i] : 0.0;
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int r = r0 + u;
        if (r >= R) break;
        const double v = vb[u];
        if constexpr (STAGE) s_vals[r * kTile + t] = v;
        if (r == 0) {
          sum = v;
          M = v;
        } else {
          sum = sum + v;
          const double nm = M + (v - M) / (double)(r + 1);
          S = S + (v - M) * (v - nm);
          M = nm;
        }
        if (A.any_count) {
#pragma unroll
          for (int q = 0; q < kMaxStats; ++q) {
            if (q >= A.n_stats) break;
            if (A.kind[q] == GSP_STAT_CDF) cnt[q] += v <= A.x[q];
            if (A.kind[q] == GSP_STAT_CCDF) cnt[q] += v > A.x[q];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kMaxStats; ++q) {
      if (q >= A.n_stats) break;
      const int kind = A.kind[q];
      if (kind == GSP_STAT_MEAN)
        out[(int64_t)q * A.N + i] = sum / (double)R;
      else if (kind == GSP_STAT_VAR)  // R = 1: 0 / 0, NaN like the reference
        out[(int64_t)q * A.N + i] = S / (double)(R - 1);
      else if (kind == GSP_STAT_CDF || kind == GSP_STAT_CCDF)
        out[(int64_t)q * A.N + i] = (double)cnt[q] / (double)R;
    }
    // quantiles: the rank of every value (#{< v}, #{<= v}) once per group of
    // quantiles; the k-th smallest (0-based) is the v with #{<} <= k < #{<=}
    auto val = [&](int r) -> double {
      if constexpr (STAGE)
        return s_vals[r * kTile + t];
      else
        return reals[(int64_t)r * A.N + i];
    };
#pragma unroll 1
    for (int q0 = 0; q0 < A.n_quant; q0 += kQGroup) {
      double lo[kQGroup], hi[kQGroup];
      bool nan = false;
#pragma unroll
      for (int u = 0; u < kQGroup; ++u) lo[u] = hi[u] = 0.0;
      for (int a = 0; a < R; ++a) {
        const double v = val(a);
        nan = nan || v != v;
        int lt = 0, le = 0;
        for (int b = 0; b < R; ++b) {
          const double w = val(b);
          lt += w < v;
          le += w <= v;
        }
#pragma unroll
        for (int u = 0; u < kQGroup; ++u) {
          if (q0 + u >= A.n_quant) break;
          const int k = A.q_j[q0 + u];
          if (lt <= k && k < le) lo[u] = v;
          if (lt <= k + 1 && k + 1 < le) hi[u] = v;
        }
      }
#pragma unroll
      for (int u = 0; u < kQGroup; ++u) {
        if (q0 + u >= A.n_quant) break;
        // a NaN among the values: NaN (the reference throws)
        const double a = nan ? __builtin_nan("") : lo[u];
        const double b = nan ? __builtin_nan("") : (R == 1 ? lo[u] : hi[u]);
        const double g = A.q_g[q0 + u];
        out[(int64_t)A.q_row[q0 + u] * A.N + i] =
            (__builtin_isfinite(a) && __builtin_isfinite(b)) ? a + g * (b - a)
                                                             : (1.0 - g) * a + g * b;
      }
    }
  }
}

void ensemble_stats_run(gsp_ctx* ctx, int32_t R, int64_t N, const double* reals, int32_t n_stats,
                        const int32_t* kinds, const double* params, double* out) {
  StatArgs A;
  std::memset(&A, 0, sizeof(A));
  A.N = N;
  A.R = R;
  A.n_stats = n_stats;
  for (int q = 0; q < n_stats; ++q) {
    const int k = kinds[q];
    GSP_REQUIRE(k >= GSP_STAT_MEAN && k <= GSP_STAT_QUANTILE, GSP_ERR_ARGUMENT,
                "unknown statistic %d", k);
    A.kind[q] = k;
    if (k == GSP_STAT_CDF || k == GSP_STAT_CCDF) {
      A.x[q] = params[q];
      A.any_count = 1;
    }
    if (k == GSP_STAT_QUANTILE) {
      const double p = params[q];
      GSP_REQUIRE(p >= 0.0 && p <= 1.0, GSP_ERR_ARGUMENT, "input probability out of [0,1] range");
      // Statistics.quantile with alpha = beta = 1: m = alpha + p (1 - alpha - beta)
      const double m = 1.0 + p * (1.0 - 1.0 - 1.0);
      const double h = std::fma((double)R, p, m);
      int64_t j = (int64_t)std::trunc(h);
      if (R > 1) j = std::min<int64_t>(std::max<int64_t>(j, 1), R - 1);
      double g = h - (double)j;
      g = std::min(std::max(g, 0.0), 1.0);
      const int n = A.n_quant++;
      A.q_row[n] = q;
      A.q_j[n] = R > 1 ? (int32_t)(j - 1) : 0;
      A.q_g[n] = R > 1 ? g : 0.0;
    }
  }
  const int64_t tiles = (N + kTile - 1) / kTile;
  const int blocks = (int)std::min<int64_t>(tiles, 1 << 20);
  TimedLaunch tl(ctx, GSP_K_STATS);
  if (A.n_quant > 0 && R <= kStage)
    hipLaunchKernelGGL(k_ens_stats<true>, dim3(blocks), dim3(kTile),
                       (size_t)R * kTile * sizeof(double), ctx->stream, A, reals, out);
  else
    hipLaunchKernelGGL(k_ens_stats<false>, dim3(blocks), dim3(kTile), 0, ctx->stream, A, reals,
                       out);
  GSP_HIP(hipGetLastError());
}

}  // namespace
}  // namespace gsp

extern "C" int32_t gsp_ensemble_stats(gsp_ctx* ctx, int32_t R, int64_t N, const double* reals,
                                      int32_t n_stats, const int32_t* kinds,
                                      const double* params, double* out,
                                      int32_t device_pointers) {
  return gsp_guard(ctx, [&] {
    GSP_REQUIRE(R >= 1 && N >= 1, GSP_ERR_ARGUMENT, "empty ensemble");
    GSP_REQUIRE(n_stats >= 1 && n_stats <= gsp::kMaxStats, GSP_ERR_ARGUMENT,
                "between 1 and %d statistics per call", gsp::kMaxStats);
    GSP_REQUIRE(reals && out && kinds && params, GSP_ERR_ARGUMENT, "null argument");
    hipStream_t s = ctx->stream;
    if (device_pointers) {
      gsp::ensemble_stats_run(ctx, R, N, reals, n_stats, kinds, params, out);
    } else {
      const size_t nin = (size_t)R * (size_t)N, nout = (size_t)n_stats * (size_t)N;
      ctx->stage_in.reserve(nin);
      ctx->stage_out.reserve(nout);
      GSP_HIP(hipMemcpyAsync(ctx->stage_in.p, reals, nin * 8, hipMemcpyHostToDevice, s));
      gsp::ensemble_stats_run(ctx, R, N, ctx->stage_in.p, n_stats, kinds, params,
                              ctx->stage_out.p);
      GSP_HIP(hipMemcpyAsync(out, ctx->stage_out.p, nout * 8, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      gsp::flush_timings(ctx);
    }
  });
}
