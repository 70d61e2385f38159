// gsp_krige.h — the one-system-per-wavefront kriging kernel k_krige<KB, P>
// (covariance assembly and a register-resident Cholesky solve, one matrix row
// per lane) and its launchers, shared by gsp_seqsim.hip (KB = 24, 32) and
// gsp_krige_48.hip, gsp_krige_64.hip, gsp_krige_wide.hip (KB = 48 and 64,
// compiled apart because their fully unrolled column loops dominate the
// build).
#pragma once

#include <algorithm>

#include "gsp_seqsim.h"
#include "gsp_wave.h"

namespace gsp {
namespace {

// Scaled coordinates of neighbour `nb` (a node of request R or a datum).
__device__ inline void neighbour_coords(const PlanArgs& A, const Req& R, int nb, double& x0,
                                        double& x1, double& x2) {
  if (nb < A.NT) {
    const int l = nb - R.node_off;
    const int nx = R.dims[0], ny = R.dims[1];
    const int c0 = l % nx, c1 = (l / nx) % ny, c2 = l / (nx * ny);
    x0 = R.ao[0] + ((double)c0 + 0.5) * R.as[0];
    x1 = R.ao[1] + ((double)c1 + 0.5) * R.as[1];
    x2 = R.ao[2] + ((double)c2 + 0.5) * R.as[2];
  } else {
    const double* p = A.data + (int64_t)(nb - A.NT) * 3;
    x0 = R.alpha * p[0];
    x1 = R.ndim > 1 ? R.alpha * p[1] : 0.0;
    x2 = R.ndim > 2 ? R.alpha * p[2] : 0.0;
  }
}

// Scaled coordinates of the centre of local node l of request R.
__device__ inline void node_coords(const Req& R, int l, double& t0, double& t1, double& t2) {
  const int nx = R.dims[0], ny = R.dims[1];
  const int c0 = l % nx, c1 = (l / nx) % ny, c2 = l / (nx * ny);
  t0 = R.ao[0] + ((double)c0 + 0.5) * R.as[0];
  t1 = R.ao[1] + ((double)c1 + 0.5) * R.as[1];
  t2 = R.ao[2] + ((double)c2 + 0.5) * R.as[2];
}

// Per-wavefront LDS of the kriging solve: neighbour coordinates, the FP64
// covariance matrix (both triangles; later the FP64 factor) and, for the FP32
// path, the FP32 factor.
template <int KB, int PREC>
struct KrigeLds {
  double x[KB][3];
  double A[KB][KB + 1];
  float F[PREC == GSP_FP32_REFINE ? KB : 1][KB + 1];
};

// Covariance assembly spread over the entries of the lower triangle (lane e
// computes entries e, e + 64, ...), stored to both triangles; lane j < n also
// returns its right-hand side c_j = C(x_j, x_target).
template <int KB, int PREC>
__device__ inline double assemble(KrigeLds<KB, PREC>& L, const DevModel& M, int n, int lane,
                                  double x0, double x1, double x2, double t0, double t1,
                                  double t2) {
  if (lane < n) {
    L.x[lane][0] = x0;
    L.x[lane][1] = x1;
    L.x[lane][2] = x2;
  }
  wave_sync();
  const int T = n * (n + 1) / 2;
  for (int e = lane; e < T; e += 64) {
    int j, m;
    tri_unpack(e, j, m);
    const double v = cov_points(M, L.x[j][0], L.x[j][1], L.x[j][2], L.x[m][0], L.x[m][1],
                                L.x[m][2]);
    L.A[j][m] = v;
    L.A[m][j] = v;
  }
  const double c = lane < n ? cov_points(M, x0, x1, x2, t0, t1, t2) : 0.0;
  wave_sync();
  return c;
}

// Simple-kriging solve, one matrix row per lane. Same operation order as the
// oracle: right-looking Cholesky whose per-element subtraction order is
// ascending, each column scaled by the reciprocal of its pivot (which the
// lane of that row keeps for both substitutions), forward substitution
// ascending, back substitution descending. Expects L.A assembled; leaves the
// factor in L.A.
template <int KB, int PREC>
__device__ inline int krige_fp64(KrigeLds<KB, PREC>& L, int n, int lane, double c, double C0,
                                 double& lam_out, double& sigma_out) {
  double a[KB];
#pragma unroll
  for (int m = 0; m < KB; ++m) a[m] = (m < n && lane < n) ? L.A[lane][m] : 0.0;
  double dinv = 0.0;  // 1 / L[lane][lane]
#pragma unroll
  for (int col = 0; col < KB; ++col) {
    if (col < n) {
      const double piv = rl(a[col], col);
      if (!(piv > 0.0)) return GSP_FLAG_FAIL;
      const double lcc = __builtin_sqrt(piv);
      const double rinv = 1.0 / lcc;
      if (lane == col) dinv = rinv;
      a[col] = lane == col ? lcc : (lane > col ? a[col] * rinv : a[col]);
#pragma unroll
      for (int m = col + 1; m < KB; ++m) {
        if (m < n) {
          const double lmc = rl(a[col], m);
          a[m] = a[m] - a[col] * lmc;
        }
      }
    }
  }
  double s = c, y = 0.0;
#pragma unroll
  for (int m = 0; m < KB; ++m) {
    if (m < n) {
      const double ym = rl(s * dinv, m);
      if (lane == m) y = ym;
      if (lane > m) s = s - a[m] * ym;
    }
  }
  wave_sync();
  if (lane < n) {
#pragma unroll
    for (int m = 0; m < KB; ++m)
      if (m <= lane) L.A[lane][m] = a[m];
  }
  wave_sync();
  double s2 = y, lam = 0.0;
  for (int m = n - 1; m >= 0; --m) {
    const double lm = rl(s2 * dinv, m);
    if (lane == m) lam = lm;
    if (lane < m) s2 = s2 - L.A[m][lane] * lm;
  }
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc = acc + rl(lam, j) * rl(c, j);
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  lam_out = lam;
  sigma_out = __builtin_sqrt(var);
  return 0;
}

// FP32 factor + one refinement step with an FP64 residual. A pivot at or below
// tau = sill * 2^-12 means the system is too ill-conditioned for one FP32
// refinement step to reach FP64-level weights (or singular), so the node is
// solved in FP64 instead, which also keeps the prior fallback identical.
template <int KB>
__device__ inline int krige_fp32r(KrigeLds<KB, GSP_FP32_REFINE>& L, int n, int lane, double c,
                                  double C0, float tau, double& lam_out, double& sigma_out) {
  float f[KB];
#pragma unroll
  for (int m = 0; m < KB; ++m) f[m] = (m < n && lane < n) ? (float)L.A[lane][m] : 0.0f;
  bool ok = true;
  float dinv = 0.0f;  // 1 / L[lane][lane]
#pragma unroll
  for (int col = 0; col < KB; ++col) {
    if (ok && col < n) {
      const float piv = rlf(f[col], col);
      if (!(piv > tau)) {
        ok = false;
      } else {
        const float lcc = __builtin_sqrtf(piv);
        const float rinv = 1.0f / lcc;
        if (lane == col) dinv = rinv;
        f[col] = lane == col ? lcc : (lane > col ? f[col] * rinv : f[col]);
#pragma unroll
        for (int m = col + 1; m < KB; ++m) {
          if (m < n) {
            const float lmc = rlf(f[col], m);
            f[m] = f[m] - f[col] * lmc;
          }
        }
      }
    }
  }
  if (!ok) return krige_fp64<KB, GSP_FP32_REFINE>(L, n, lane, c, C0, lam_out, sigma_out) |
                  GSP_FLAG_FP64_FALLBACK;
  if (lane < n) {
#pragma unroll
    for (int m = 0; m < KB; ++m)
      if (m <= lane) L.F[lane][m] = f[m];
  }
  wave_sync();
  // two FP32 solves with the same factor: lambda0 = A^-1 c, delta = A^-1 r
  float rhs = (float)c, sol = 0.0f;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    float s = rhs, y = 0.0f;
#pragma unroll
    for (int m = 0; m < KB; ++m) {
      if (m < n) {
        const float ym = rlf(s * dinv, m);
        if (lane == m) y = ym;
        if (lane > m) s = s - f[m] * ym;
      }
    }
    float s2 = y, l = 0.0f;
    for (int m = n - 1; m >= 0; --m) {
      const float lm = rlf(s2 * dinv, m);
      if (lane == m) l = lm;
      if (lane < m) s2 = s2 - L.F[m][lane] * lm;
    }
    if (pass == 0) {
      sol = l;
      double r = c;
      for (int m = 0; m < n; ++m) r = r - (lane < n ? L.A[lane][m] : 0.0) * (double)rlf(l, m);
      rhs = (float)r;
    } else {
      lam_out = (double)sol + (double)l;
    }
  }
  const double lam = lam_out;
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc = acc + rl(lam, j) * rl(c, j);
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  sigma_out = __builtin_sqrt(var);
  return 0;
}

// Wavefronts per block of the kriging kernel. LDS per wavefront grows with
// KB^2 and limits the resident wavefronts beyond KB = 24: single-wavefront
// blocks pack the most per CU there (KB = 32: eleven instead of eight).
constexpr int plan_waves(int KB) { return KB <= 24 ? 4 : 1; }

// Resident wavefronts per SIMD asked of the register allocator: five at
// KB <= 24 (96 registers instead of 122; measured 2-3 % faster at the 1M and
// 23M node workloads despite a small spill of the FP32 path), no constraint
// where LDS limits residency anyway.
constexpr int krige_waves_per_simd(int KB) { return KB <= 24 ? 5 : 1; }

// Kriging systems with more than 16 neighbours, one wavefront per node:
// covariance assembly, Cholesky solve, weights, sigma and flags, plus the
// dependency degrees.
template <int KB, int PREC>
__global__ __launch_bounds__(64 * plan_waves(KB)) __attribute__((amdgpu_waves_per_eu(krige_waves_per_simd(KB))))
void k_krige(const PlanArgs A, const int32_t* list,
                                                               int32_t count) {
  __builtin_amdgcn_s_setprio(3);  // issue ahead of the level kernels beside the solves
  constexpr int kW = plan_waves(KB);
  __shared__ KrigeLds<KB, PREC> s_krige[kW];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t wstride = (int64_t)gridDim.x * kW;
  for (int64_t idx = (int64_t)blockIdx.x * kW + w; idx < count; idx += wstride) {
    const int gi = __builtin_amdgcn_readfirstlane(list[idx]);
    const Req& R = A.reqs[__builtin_amdgcn_readfirstlane(A.node_req[gi])];
    const int n = A.nfound[gi];
    const int nb = lane < n ? A.nbr[(int64_t)gi * A.K + lane] : -1;
    int flag = 0;
    double lam = 0.0, sig = R.prior_sigma;
    if (n < R.kmin) {
      flag = GSP_FLAG_FEW;
    } else {
      double x0 = 0.0, x1 = 0.0, x2 = 0.0, t0, t1, t2;
      if (lane < n) neighbour_coords(A, R, nb, x0, x1, x2);
      node_coords(R, gi - R.node_off, t0, t1, t2);
      KrigeLds<KB, PREC>& KL = s_krige[w];
      const double c = assemble<KB, PREC>(KL, R.M, n, lane, x0, x1, x2, t0, t1, t2);
      if constexpr (PREC == GSP_FP32_REFINE)
        flag = krige_fp32r<KB>(KL, n, lane, c, R.M.sill, (float)(R.M.sill * 0x1p-12), lam, sig);
      else
        flag = krige_fp64<KB, PREC>(KL, n, lane, c, R.M.sill, lam, sig);
      wave_sync();
      if (flag & GSP_FLAG_FAIL) {
        lam = 0.0;
        sig = R.prior_sigma;
      }
    }
    if (lane < A.K) A.lam[(int64_t)gi * A.K + lane] = lane < n ? lam : 0.0;
    if (lane == 0) {
      A.sigma[gi] = sig;
      A.flags[gi] = (uint8_t)flag;
      if (flag & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) atomicAdd(&A.counters[CTR_PRIOR], 1);
      if (flag & GSP_FLAG_FAIL) atomicAdd(&A.counters[CTR_FAIL], 1);
      if (flag & GSP_FLAG_FP64_FALLBACK) atomicAdd(&A.counters[CTR_FALLBACK], 1);
    }
  }
}

template <int KB>
void launch_krige(const PlanArgs& A, int precision, const int32_t* list, int64_t count,
                  hipStream_t s) {
  if (count <= 0) return;
  constexpr int W = plan_waves(KB);
  const int blocks = (int)std::min<int64_t>((count + W - 1) / W, 65536 / W);
  if (precision == GSP_FP32_REFINE)
    hipLaunchKernelGGL((k_krige<KB, GSP_FP32_REFINE>), dim3(blocks), dim3(64 * W), 0, s, A, list,
                       (int32_t)count);
  else
    hipLaunchKernelGGL((k_krige<KB, GSP_FP64>), dim3(blocks), dim3(64 * W), 0, s, A, list,
                       (int32_t)count);
}

// The same for one precision only.
template <int KB, int PREC>
void launch_krige_prec(const PlanArgs& A, const int32_t* list, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  constexpr int W = plan_waves(KB);
  const int blocks = (int)std::min<int64_t>((count + W - 1) / W, 65536 / W);
  hipLaunchKernelGGL((k_krige<KB, PREC>), dim3(blocks), dim3(64 * W), 0, s, A, list,
                     (int32_t)count);
}

}  // namespace

// k_krige<KB, precision> for KB = 48 or 64 (gsp_krige_wide.hip).
void launch_krige_wide(int KB, const PlanArgs& A, int precision, const int32_t* list,
                       int64_t count, hipStream_t s);

}  // namespace gsp
