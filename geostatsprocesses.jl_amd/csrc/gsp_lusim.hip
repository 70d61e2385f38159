// gsp_lusim.hip — LUSIM on gfx950 (src/simulation/field/lusim.jl:38-175).
//
// Dense FP64 linear algebra on row-major lower-triangular storage:
//   k_cov_matrix   covariance assembly C(x_a, x_b) (utils.jl:50-62)
//   k_potrf64      Cholesky of one <=64 x 64 diagonal block, one wavefront,
//                  one row per lane held in registers
//   k_trsm_rows    X <- X L^-T for a <=64-wide panel, one row per thread
//   k_gemm_nt      C <- beta C + alpha A B^T on MFMA (FP64 16x16x4, or FP32
//                  16x16x4 for the mixed-precision trailing updates)
// composed into a right-looking blocked Cholesky, a blocked triangular solve,
// and the sampling product Z = d2 + W L22^T (a TRMM: only k <= j is read).
#include <algorithm>
#include <cmath>
#include <limits>

#include "gsp_internal.h"

namespace gsp {
namespace {

constexpr int NB = 64;  // panel width

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline double rl(double x, int lane) {
  uint64_t u = __builtin_bit_cast(uint64_t, x);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

struct CovArgs {
  DevModel M;
  int32_t ndim;
  int32_t dims[3];
  double org[3], sp[3];
};

__device__ inline void centroid(const CovArgs& A, int64_t id, double& x0, double& x1,
                                double& x2) {
  const int64_t nx = A.dims[0], ny = A.dims[1];
  const int64_t i0 = id % nx, i1 = (id / nx) % ny, i2 = id / (nx * ny);
  x0 = A.org[0] + ((double)i0 + 0.5) * A.sp[0];
  x1 = A.org[1] + ((double)i1 + 0.5) * A.sp[1];
  x2 = A.org[2] + ((double)i2 + 0.5) * A.sp[2];
}

// out[a * ldo + b] = C(x_rows[a], x_cols[b]) for a < M, b < N (physical units).
__global__ __launch_bounds__(256) void k_cov_matrix(const CovArgs A, const int64_t* rows,
                                                    const int64_t* cols, int64_t M, int64_t N,
                                                    double* out, int64_t ldo) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t a = blockIdx.y;
  if (b >= N || a >= M) return;
  double xa0, xa1, xa2, xb0, xb1, xb2;
  centroid(A, rows[a], xa0, xa1, xa2);
  centroid(A, cols[b], xb0, xb1, xb2);
  out[a * ldo + b] = cov_points(A.M, xa0, xa1, xa2, xb0, xb1, xb2);
}

// In-place Cholesky of the n x n (n <= 64) block at A (leading dimension lda).
// Lane j holds row j; the lower triangle is written back, the upper zeroed.
__global__ __launch_bounds__(64) void k_potrf64(double* A, int64_t lda, int n, int* err) {
  if (*err) return;
  const int lane = threadIdx.x;
  double a[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) a[m] = (lane < n && m < n) ? A[lane * lda + m] : 0.0;
#pragma unroll
  for (int col = 0; col < NB; ++col) {
    if (col < n) {
      const double piv = rl(a[col], col);
      if (!(piv > 0.0)) {
        if (lane == 0) atomicOr(err, 1);
        return;
      }
      const double lcc = __builtin_sqrt(piv);
      a[col] = lane == col ? lcc : (lane > col ? a[col] / lcc : a[col]);
#pragma unroll
      for (int m = col + 1; m < NB; ++m) {
        if (m < n) {
          const double lmc = rl(a[col], m);
          a[m] = a[m] - a[col] * lmc;
        }
      }
    }
  }
  if (lane < n) {
#pragma unroll
    for (int m = 0; m < NB; ++m)
      if (m < n) A[lane * lda + m] = m <= lane ? a[m] : 0.0;
  }
}

// X[r, 0:nb] <- X[r, 0:nb] L^-T (forward substitution per row) for M rows.
__global__ __launch_bounds__(256) void k_trsm_rows(const double* L, int64_t ldl, int nb,
                                                   double* X, int64_t ldx, int64_t M,
                                                   const int* err) {
  if (*err) return;
  __shared__ double sL[NB][NB + 1];
  for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
    const int c = t / NB, m = t - c * NB;
    sL[c][m] = (c < nb && m <= c) ? L[c * ldl + m] : 0.0;
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  double* row = X + r * ldx;
  double x[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) x[c] = c < nb ? row[c] : 0.0;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nb) {
      double s = x[c];
#pragma unroll
      for (int m = 0; m < c; ++m) s = s - x[m] * sL[c][m];
      x[c] = s / sL[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < nb) row[c] = x[c];
}

template <class T>
struct Mfma;
template <>
struct Mfma<double> {
  static constexpr bool kF64 = true;
  typedef v4d acc;
  __device__ static acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  static constexpr bool kF64 = false;
  typedef v4f acc;
  __device__ static acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// C <- beta C + alpha A B^T with A: M x K, B: N x K, C: M x N (row-major).
// 64 x 64 block tile, 4 waves each owning a 32 x 32 quarter (2 x 2 MFMA tiles).
// lower: only C[i][j] with i >= j is touched (tiles above the diagonal exit).
// triB: B is lower triangular (B[j][k] = 0 for k > j), so K stops at the tile.
template <class T>
__global__ __launch_bounds__(256) void k_gemm_nt(int64_t M, int64_t N, int64_t K,
                                                 const double* A, int64_t lda, const double* B,
                                                 int64_t ldb, double* C, int64_t ldc,
                                                 double alpha, double beta, int lower, int triB,
                                                 const int* err) {
  constexpr int BK = 16;
  if (err && *err) return;
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  if (lower && i0 + 63 < j0) return;
  __shared__ T sA[64][BK + 1];
  __shared__ T sB[64][BK + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  typename Mfma<T>::acc acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;
  int64_t kend = K;
  if (triB) kend = std::min<int64_t>(K, j0 + 64);
  for (int64_t k0 = 0; k0 < kend; k0 += BK) {
    for (int t = tid; t < 64 * BK; t += 256) {
      const int rr = t / BK, kk = t - rr * BK;
      const int64_t gi = i0 + rr, gj = j0 + rr, gk = k0 + kk;
      sA[rr][kk] = (gi < M && gk < kend) ? (T)A[gi * lda + gk] : (T)0;
      sB[rr][kk] = (gj < N && gk < kend) ? (T)B[gj * ldb + gk] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kl = kk + (lane >> 4);
      const int il = lane & 15;
      const T a0 = sA[wm + il][kl], a1 = sA[wm + 16 + il][kl];
      const T b0 = sB[wn + il][kl], b1 = sB[wn + 16 + il][kl];
      acc[0][0] = Mfma<T>::step(a0, b0, acc[0][0]);
      acc[0][1] = Mfma<T>::step(a0, b1, acc[0][1]);
      acc[1][0] = Mfma<T>::step(a1, b0, acc[1][0]);
      acc[1][1] = Mfma<T>::step(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // output row of accumulator register v: FP64 16x16x4 interleaves the
        // four lane groups (row 4v + l/16), FP32 16x16x4 gives each lane group
        // four consecutive rows (row 4(l/16) + v); the column is l % 16
        const int row = Mfma<T>::kF64 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v;
        const int64_t gi = i0 + wm + a * 16 + row;
        const int64_t gj = j0 + wn + b * 16 + (lane & 15);
        if (gi < M && gj < N && (!lower || gi >= gj)) {
          double* c = C + gi * ldc + gj;
          const double prod = alpha * (double)acc[a][b][v];
          *c = beta == 0.0 ? prod : beta * *c + prod;
        }
      }
}

__global__ void k_zero_upper(double* A, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j < n && j > i) A[i * n + j] = 0.0;
}

__global__ void k_rows_from_vector(double* Z, int64_t R, int64_t n, const double* v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R * n) Z[t] = v[t % n];
}

__global__ void k_scatter(const double* Z, int64_t R, int64_t n_s, int64_t N,
                          const int64_t* sinds, const int64_t* dinds, const double* z1,
                          int64_t n_d, double add, double* out) {
  const int64_t r = blockIdx.y;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_s + n_d;
       a += (int64_t)gridDim.x * blockDim.x) {
    if (a < n_s)
      out[r * N + sinds[a]] = Z[r * n_s + a] + add;
    else
      out[r * N + dinds[a - n_s]] = z1[a - n_s];
  }
}

// Lane-layout probe: setup 0 feeds A = lane + 1, B = 1 (D holds row sums of A),
// setup 1 feeds A = 1, B = lane + 1 (D holds column sums of B).
template <class T>
__global__ __launch_bounds__(64) void k_mfma_probe(double* out) {
  const int lane = threadIdx.x;
  for (int setup = 0; setup < 2; ++setup) {
    typename Mfma<T>::acc acc;
    for (int v = 0; v < 4; ++v) acc[v] = 0;
    const T a = setup == 0 ? (T)(lane + 1) : (T)1;
    const T b = setup == 0 ? (T)1 : (T)(lane + 1);
    acc = Mfma<T>::step(a, b, acc);
    for (int v = 0; v < 4; ++v) out[(setup * 64 + lane) * 4 + v] = (double)acc[v];
  }
}

// ------------------------------------------------------------------ host --
void gemm_nt(hipStream_t s, bool mixed, int64_t M, int64_t N, int64_t K, const double* A,
             int64_t lda, const double* B, int64_t ldb, double* C, int64_t ldc, double alpha,
             double beta, bool lower, bool triB, const int* err) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  if (mixed)
    hipLaunchKernelGGL(k_gemm_nt<float>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                       alpha, beta, (int)lower, (int)triB, err);
  else
    hipLaunchKernelGGL(k_gemm_nt<double>, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                       alpha, beta, (int)lower, (int)triB, err);
}

// Right-looking blocked Cholesky of the n x n row-major matrix A (lower).
void cholesky(hipStream_t s, double* A, int64_t n, bool mixed, int* err) {
  for (int64_t k0 = 0; k0 < n; k0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, n - k0);
    hipLaunchKernelGGL(k_potrf64, dim3(1), dim3(64), 0, s, A + k0 * n + k0, n, nb, err);
    const int64_t k1 = k0 + nb, rest = n - k1;
    if (rest <= 0) break;
    hipLaunchKernelGGL(k_trsm_rows, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s,
                       A + k0 * n + k0, n, nb, A + k1 * n + k0, n, rest, err);
    gemm_nt(s, mixed, rest, rest, nb, A + k1 * n + k0, n, A + k1 * n + k0, n, A + k1 * n + k1, n,
            -1.0, 1.0, true, false, err);
  }
}

// X <- X L^-T for X: M x n (leading dimension ldx), L: n x n lower (row-major).
void trsm_right_lt(hipStream_t s, const double* L, int64_t n, double* X, int64_t M, int64_t ldx,
                   int* err) {
  for (int64_t b0 = 0; b0 < n; b0 += NB) {
    const int nb = (int)std::min<int64_t>(NB, n - b0);
    hipLaunchKernelGGL(k_trsm_rows, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s,
                       L + b0 * n + b0, n, nb, X + b0, ldx, M, err);
    const int64_t b1 = b0 + nb;
    if (b1 < n)
      gemm_nt(s, false, M, n - b1, nb, X + b0, ldx, L + b1 * n + b0, n, X + b1, ldx, -1.0, 1.0,
              false, false, err);
  }
}

void cov_matrix(hipStream_t s, const CovArgs& A, const int64_t* rows, const int64_t* cols,
                int64_t M, int64_t N, double* out, int64_t ldo) {
  if (M <= 0 || N <= 0) return;
  for (int64_t a0 = 0; a0 < M; a0 += 65535) {
    const int64_t m = std::min<int64_t>(65535, M - a0);
    hipLaunchKernelGGL(k_cov_matrix, dim3((unsigned)((N + 255) / 256), (unsigned)m), dim3(256), 0,
                       s, A, rows + a0, cols, m, N, out + a0 * ldo, ldo);
  }
}

}  // namespace

namespace {
__global__ __launch_bounds__(256) void k_copy16(uint4* dst, const uint4* src, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}
}  // namespace

void copy_kernel(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes) {
  const int64_t n = bytes / 16;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
  if (n > 0)
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, ctx->stream, (uint4*)dst,
                       (const uint4*)src, n);
  GSP_HIP(hipGetLastError());
  GSP_HIP(hipStreamSynchronize(ctx->stream));
}

void mfma_probe(gsp_ctx* ctx, int fp32, double* host_out) {
  DevBuf<double> d;
  d.alloc(512);
  if (fp32)
    hipLaunchKernelGGL(k_mfma_probe<float>, dim3(1), dim3(64), 0, ctx->stream, d.p);
  else
    hipLaunchKernelGGL(k_mfma_probe<double>, dim3(1), dim3(64), 0, ctx->stream, d.p);
  GSP_HIP(hipGetLastError());
  GSP_HIP(hipMemcpyAsync(host_out, d.p, 512 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  GSP_HIP(hipStreamSynchronize(ctx->stream));
}

void lusim_factor_build(gsp_ctx* ctx, const gsp_lusim_problem* p, gsp_lusim_plan* plan) {
  GSP_REQUIRE(p, GSP_ERR_ARGUMENT, "null problem");
  const int64_t N = validate_grid(p->grid);
  const int nd = p->grid.ndim;
  validate_model(p->model, nd);
  GSP_REQUIRE(p->model.kind == GSP_COVARIANCE, GSP_ERR_ARGUMENT,
              "LUSIM requires a geostatistical function that is stationary, symmetric and "
              "banded. Covariances or composite functions of covariances satisfy these "
              "properties.");
  GSP_REQUIRE(p->precision == GSP_LU_FP64 || p->precision == GSP_LU_MIXED, GSP_ERR_ARGUMENT,
              "unknown LUSIM precision");
  GSP_REQUIRE(p->n_data >= 0 && p->n_data <= N, GSP_ERR_ARGUMENT, "invalid number of data cells");
  GSP_REQUIRE(p->n_data == 0 || (p->dinds && p->z1), GSP_ERR_ARGUMENT, "null data arrays");
  GSP_REQUIRE(std::isfinite(p->mean), GSP_ERR_ARGUMENT, "invalid mean");
  GSP_REQUIRE(N <= 65536, GSP_ERR_UNSUPPORTED, "LUSIM grids above 65536 cells are not supported");
  hipStream_t s = ctx->stream;
  const int64_t nD = p->n_data;
  std::vector<int64_t> dinds((size_t)nD);
  if (nD > 0) {
    if (p->device_pointers)
      GSP_HIP(hipMemcpy(dinds.data(), p->dinds, (size_t)nD * 8, hipMemcpyDeviceToHost));
    else
      std::memcpy(dinds.data(), p->dinds, (size_t)nD * 8);
  }
  for (int64_t i = 0; i < nD; ++i) {
    GSP_REQUIRE(dinds[(size_t)i] >= 0 && dinds[(size_t)i] < N, GSP_ERR_ARGUMENT,
                "data index out of range");
    GSP_REQUIRE(i == 0 || dinds[(size_t)i] > dinds[(size_t)i - 1], GSP_ERR_ARGUMENT,
                "data indices must be strictly ascending (findall order)");
  }
  std::vector<int64_t> sinds;
  sinds.reserve((size_t)(N - nD));
  {
    size_t q = 0;
    for (int64_t i = 0; i < N; ++i) {
      if (q < dinds.size() && dinds[q] == i) {
        ++q;
        continue;
      }
      sinds.push_back(i);
    }
  }
  const int64_t nS = (int64_t)sinds.size();
  plan->N = N;
  plan->n_s = nS;
  plan->n_d = nD;
  plan->mean = p->mean;
  plan->dinds.alloc((size_t)std::max<int64_t>(nD, 1));
  plan->sinds.alloc((size_t)std::max<int64_t>(nS, 1));
  plan->z1.alloc((size_t)std::max<int64_t>(nD, 1));
  plan->d2.alloc((size_t)std::max<int64_t>(nS, 1));
  plan->L22.alloc((size_t)std::max<int64_t>(nS * nS, 1));
  if (nD > 0) {
    GSP_HIP(hipMemcpyAsync(plan->dinds.p, dinds.data(), (size_t)nD * 8, hipMemcpyHostToDevice, s));
    GSP_HIP(hipMemcpyAsync(plan->z1.p, p->z1, (size_t)nD * 8,
                           p->device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           s));
  }
  if (nS > 0)
    GSP_HIP(hipMemcpyAsync(plan->sinds.p, sinds.data(), (size_t)nS * 8, hipMemcpyHostToDevice, s));

  CovArgs C;
  std::memset(&C, 0, sizeof(C));
  C.M = make_dev_model(p->model, nd, 1.0);
  C.ndim = nd;
  for (int d = 0; d < 3; ++d) {
    C.dims[d] = d < nd ? (int32_t)p->grid.dims[d] : 1;
    C.org[d] = d < nd ? p->grid.origin[d] : 0.0;
    C.sp[d] = d < nd ? p->grid.spacing[d] : 1.0;
  }
  ctx->counters.reserve(16);
  int* err = ctx->counters.p + 8;
  GSP_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  const bool mixed = p->precision == GSP_LU_MIXED;
  TimedLaunch tl(ctx, GSP_K_LU_FACTOR);

  if (nS > 0) cov_matrix(s, C, plan->sinds.p, plan->sinds.p, nS, nS, plan->L22.p, nS);
  if (nD == 0) {
    GSP_HIP(hipMemsetAsync(plan->d2.p, 0, (size_t)std::max<int64_t>(nS, 1) * 8, s));
    cholesky(s, plan->L22.p, nS, mixed, err);
  } else {
    DevBuf<double> L11, X, v;
    L11.alloc((size_t)(nD * nD));
    X.alloc((size_t)std::max<int64_t>(nS * nD, 1));
    v.alloc((size_t)nD);
    cov_matrix(s, C, plan->dinds.p, plan->dinds.p, nD, nD, L11.p, nD);
    cholesky(s, L11.p, nD, false, err);
    // X = B12^T = C21 L11^-T  (lusim.jl:99);  v = L11^-1 z1  (lusim.jl:102)
    cov_matrix(s, C, plan->sinds.p, plan->dinds.p, nS, nD, X.p, nD);
    if (nS > 0) trsm_right_lt(s, L11.p, nD, X.p, nS, nD, err);
    GSP_HIP(hipMemcpyAsync(v.p, plan->z1.p, (size_t)nD * 8, hipMemcpyDeviceToDevice, s));
    trsm_right_lt(s, L11.p, nD, v.p, 1, nD, err);
    if (nS > 0) {
      // d2 = B12^T v;  C22 - B12^T B12  (lusim.jl:102-103)
      gemm_nt(s, false, nS, 1, nD, X.p, nD, v.p, nD, plan->d2.p, 1, 1.0, 0.0, false, false, err);
      gemm_nt(s, false, nS, nS, nD, X.p, nD, X.p, nD, plan->L22.p, nS, -1.0, 1.0, true, false, err);
      cholesky(s, plan->L22.p, nS, mixed, err);
    }
    GSP_HIP(hipGetLastError());
    int e = 0;
    GSP_HIP(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
    GSP_HIP(hipStreamSynchronize(s));  // before the temporaries are freed
    GSP_REQUIRE(e == 0, GSP_ERR_POSDEF, "PosDefException: matrix is not positive definite");
  }
  if (nS > 0)
    hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)((nS + 255) / 256), (unsigned)nS), dim3(256),
                       0, s, plan->L22.p, nS);
  GSP_HIP(hipGetLastError());
  int e = 0;
  GSP_HIP(hipMemcpyAsync(&e, err, sizeof(int), hipMemcpyDeviceToHost, s));
  GSP_HIP(hipStreamSynchronize(s));
  GSP_REQUIRE(e == 0, GSP_ERR_POSDEF, "PosDefException: matrix is not positive definite");
}

void lusim_sample_run(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                      double* out) {
  hipStream_t s = ctx->stream;
  const int64_t nS = plan->n_s, nD = plan->n_d, N = plan->N;
  ctx->X.reserve((size_t)std::max<int64_t>((int64_t)R * nS, 1));
  double* Z = ctx->X.p;
  TimedLaunch tl(ctx, GSP_K_LU_SAMPLE);
  if (nS > 0) {
    hipLaunchKernelGGL(k_rows_from_vector, dim3((unsigned)((R * nS + 255) / 256)), dim3(256), 0, s,
                       Z, (int64_t)R, nS, plan->d2.p);
    gemm_nt(s, false, R, nS, nS, w, nS, plan->L22.p, nS, Z, nS, 1.0, 1.0, false, true, nullptr);
  }
  const double add = nD == 0 ? plan->mean : 0.0;  // lusim.jl:172
  hipLaunchKernelGGL(k_scatter, dim3((unsigned)std::min<int64_t>((nS + nD + 255) / 256, 4096),
                                     (unsigned)R),
                     dim3(256), 0, s, Z, (int64_t)R, nS, N, plan->sinds.p, plan->dinds.p,
                     plan->z1.p, nD, add, out);
  GSP_HIP(hipGetLastError());
}

}  // namespace gsp
