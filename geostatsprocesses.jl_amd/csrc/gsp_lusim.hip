// gsp_lusim.hip — LUSIM on gfx950 (src/simulation/field/lusim.jl:38-175).
//
// Dense FP64 linear algebra on row-major lower-triangular storage:
//   k_cov_tiles      covariance assembly C(x_a, x_b) (utils.jl:50-62), 64 x 64 tiles
//   k_diag_chol_inv  Cholesky of one <=64 x 64 diagonal tile and the inverse
//                    of its factor, in LDS, one 256-thread workgroup
//   k_panel_apply    X <- X L^-T for a <=64-wide panel: a GEMM with that
//                    inverse on FP64 MFMA, one 64-row tile per workgroup
//   k_trail          trailing update C <- C - P P^T of a panel P on MFMA
//                    (FP64 16x16x4; FP32 16x16x4 off the diagonal in the mixed
//                    form), whose first workgroup then factors and inverts the
//                    next diagonal tile (look-ahead)
//   k_gemm_nt        C <- beta C + alpha A B^T on MFMA (conditioning)
//   k_sample_part    the sampling product W L22^T (a TRMM: only k <= j is read)
//                    split into 64 x 64 tiles x 256-wide K chunks of equal work
//   k_sample_out     the chunks summed in order, + d2, scattered to the grid
// composed into a right-looking blocked Cholesky (two launches per panel), a
// blocked triangular solve, and the sampler d2 + W L22^T.
#include <algorithm>
#include <cmath>
#include <limits>

#include "gsp_internal.h"
#include "gsp_wave.h"

namespace gsp {
namespace {

constexpr int NB = 64;  // panel width

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

struct CovArgs {
  DevModel M;
  int32_t ndim;
  int32_t dims[3];
  double org[3], sp[3];
};

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

// Covariance tiles: one 64 x 64 tile of out[a * ldo + b] = C(x_rows[a],
// x_cols[b]) per block, the tile's 64 row and 64 column centroids computed
// once into LDS (cell ids below 2^31). `lower`: the packed lower triangle of
// tiles (rows == cols; diagonal tiles whole), else a Tm x Tn grid of tiles.
__global__ __launch_bounds__(256) void k_cov_tiles(const CovArgs A, const int64_t* rows,
                                                   const int64_t* cols, int64_t M, int64_t N,
                                                   double* out, int64_t ldo, int lower) {
  __shared__ double xr[64][3], xc[64][3];
  int ti, tj;
  if (lower) {
    tri_unpack((int)blockIdx.x, ti, tj);
  } else {
    ti = (int)blockIdx.y;
    tj = (int)blockIdx.x;
  }
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  const int t = threadIdx.x;
  if (t < 128) {
    const int q = t & 63;
    const bool isrow = t < 64;
    const int64_t g = isrow ? i0 + q : j0 + q;
    const bool in = isrow ? g < M : g < N;
    const int id = in ? (int)(isrow ? rows[g] : cols[g]) : 0;
    const int nx = A.dims[0], ny = A.dims[1];
    const int c0 = id % nx, c1 = (id / nx) % ny, c2 = id / (nx * ny);
    double(*x)[3] = isrow ? xr : xc;
    x[q][0] = A.org[0] + ((double)c0 + 0.5) * A.sp[0];
    x[q][1] = A.org[1] + ((double)c1 + 0.5) * A.sp[1];
    x[q][2] = A.org[2] + ((double)c2 + 0.5) * A.sp[2];
  }
  __syncthreads();
#pragma unroll 4
  for (int e = t; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const int64_t a = i0 + r, b = j0 + c;
    if (a < M && b < N)
      out[a * ldo + b] = cov_points(A.M, xr[r][0], xr[r][1], xr[r][2], xc[c][0], xc[c][1],
                                    xc[c][2]);
  }
}

// ---- diagonal tiles ---------------------------------------------------------
// LDS of one diagonal tile: the tile T (factored in place), the inverse V of
// its factor, and the partial products S of the inversion.
struct TileLds {
  double T[64][65];
  double V[64][65];
  double S[48][17];
  int bad;
};

// Cholesky of the symmetric 64 x 64 tile T (lower triangle read) in place
// (lower = L, upper = 0) and V = L^-1 (lower, upper 0), by 256 threads.
// Four sub-panels of 16 columns: lanes 0..15 of wave 0 factor and invert the
// 16 x 16 diagonal block D in registers (row per lane, values exchanged with
// readlane), all threads multiply the 16 columns below by D^-T and update the
// trailing part. V then follows by block rows: V_pp = D_p^-1 and
// V_pq = -V_pp sum_{m=q}^{p-1} L_pm V_mq. Returns false (in every thread) if
// a pivot is not positive.
__device__ __forceinline__ bool chol_inv_tile(TileLds& L) {
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0) L.bad = 0;
  for (int e = tid; e < 64 * 64; e += 256) L.V[e >> 6][e & 63] = 0.0;
  __syncthreads();
  for (int p = 0; p < 4; ++p) {
    const int c0 = 16 * p;
    if (tid < 64) {
      const int r = lane & 15;  // lanes 16..63 repeat lanes 0..15
      // values move between the lanes of a 16-lane row by DPP row broadcasts
      // (rbd<M>: lane M of this lane's row), which stay in the vector pipeline
      double a[16], rinv[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) a[m] = m <= r ? L.T[c0 + r][c0 + m] : 0.0;
      bool ok = true;
      static_for<0, 16>([&](auto cc) {
        constexpr int col = decltype(cc)::value;
        const double piv = rbd<col>(a[col]);
        ok = ok && (piv > 0.0);
        // 1 / sqrt(piv): the hardware estimate and two Newton steps (the
        // factor needs no bit parity with LAPACK)
        double y = __builtin_amdgcn_rsq(piv);
        y = y * __builtin_fma(-0.5 * piv, y * y, 1.5);
        y = y * __builtin_fma(-0.5 * piv, y * y, 1.5);
        rinv[col] = y;
        a[col] = r == col ? piv * y : (r > col ? a[col] * y : a[col]);
        static_for<col + 1, 16>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          const double lmc = rbd<m>(a[col]);
          a[m] = __builtin_fma(-a[col], lmc, a[m]);
        });
      });
      // column r of D^-1: x_i = (delta_ir - sum_{m<i} D_im x_m) / D_ii
      double x[16];
      static_for<0, 16>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        double acc = i == r ? 1.0 : 0.0;
        static_for<0, i>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          acc = __builtin_fma(-rbd<i>(a[m]), x[m], acc);
        });
        x[i] = i >= r ? acc * rinv[i] : 0.0;
      });
      if (lane < 16) {
#pragma unroll
        for (int m = 0; m < 16; ++m) L.T[c0 + r][c0 + m] = m <= r ? a[m] : 0.0;
#pragma unroll
        for (int i = 0; i < 16; ++i) L.V[c0 + i][c0 + r] = x[i];
        if (lane == 0 && !ok) L.bad = 1;
      }
    }
    __syncthreads();
    if (p == 3) break;
    // rows below the block: T[i][c0 + c] = sum_{m <= c} T[i][c0 + m] D^-1[c][m]
    const int rows = 48 - c0;
    double vals[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int e = tid + 256 * k;
      double acc = 0.0;
      if (e < rows * 16) {
        const int i = c0 + 16 + (e >> 4), c = e & 15;
        // D^-1 is lower triangular with zeros above: a fixed 16-term sum
#pragma unroll
        for (int m = 0; m < 16; ++m) acc = __builtin_fma(L.T[i][c0 + m], L.V[c0 + c][c0 + m], acc);
      }
      vals[k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int e = tid + 256 * k;
      if (e < rows * 16) L.T[c0 + 16 + (e >> 4)][c0 + (e & 15)] = vals[k];
    }
    __syncthreads();
    // trailing part: T[i][j] -= sum_m T[i][c0 + m] T[j][c0 + m], c0 + 16 <= j <= i
    const int ne = rows * (rows + 1) / 2;
    for (int e = tid; e < ne; e += 256) {
      int ii, jj;
      tri_unpack(e, ii, jj);
      const int i = c0 + 16 + ii, j = c0 + 16 + jj;
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < 16; ++m) acc = __builtin_fma(L.T[i][c0 + m], L.T[j][c0 + m], acc);
      L.T[i][j] = L.T[i][j] - acc;
    }
    __syncthreads();
  }
  for (int p = 1; p < 4; ++p) {
    for (int e = tid; e < p * 256; e += 256) {
      const int q = e >> 8, i = (e >> 4) & 15, j = e & 15;
      double acc = 0.0;
      for (int mb = q; mb < p; ++mb) {
#pragma unroll
        for (int m = 0; m < 16; ++m)
          acc = __builtin_fma(L.T[16 * p + i][16 * mb + m], L.V[16 * mb + m][16 * q + j], acc);
      }
      L.S[16 * q + i][j] = acc;
    }
    __syncthreads();
    for (int e = tid; e < p * 256; e += 256) {
      const int q = e >> 8, i = (e >> 4) & 15, j = e & 15;
      double acc = 0.0;
#pragma unroll
      for (int m = 0; m < 16; ++m) acc = __builtin_fma(L.V[16 * p + i][16 * p + m], L.S[16 * q + m][j], acc);
      L.V[16 * p + i][16 * q + j] = -acc;
    }
    __syncthreads();
  }
  return L.bad == 0;
}

// Writes the factored tile back (nb x nb, upper part zero) and its inverse
// (64 x 64); flags a non-positive pivot.
__device__ void store_tile(const TileLds& L, bool ok, double* A, int64_t lda, int nb, double* Vout,
                           int* err) {
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int i = e >> 6, j = e & 63;
    if (i < nb && j < nb) A[i * lda + j] = L.T[i][j];
    Vout[e] = L.V[i][j];
  }
  if (!ok && threadIdx.x == 0) atomicOr(err, 1);
}

// The first diagonal tile of a factorization (the later ones are factored by
// k_trail): nb x nb at A, padded with the identity to 64 x 64.
__global__ __launch_bounds__(256, 2) void k_diag_chol_inv(double* A, int64_t lda, int nb,
                                                       double* Vout, int* err) {
  if (*err) return;
  __shared__ TileLds L;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int i = e >> 6, j = e & 63;
    L.T[i][j] = (i < nb && j < nb) ? A[i * lda + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const bool ok = chol_inv_tile(L);
  store_tile(L, ok, A, lda, nb, Vout, err);
}

// acc[a][b] += sum_k sA[wm + 16a + i][k] sB[wn + 16b + j][k] over k < 64 on
// MFMA, FP64 or (T = float) FP32 16x16x4.
template <class T>
__device__ inline void tile_nt(const double (*sA)[65], const double (*sB)[65], int wm, int wn,
                               int lane, typename Mfma<T>::acc (&acc)[2][2]) {
  const int il = lane & 15, kq = lane >> 4;
#pragma unroll 4
  for (int k = 0; k < 64; k += 4) {
    const T a0 = (T)sA[wm + il][k + kq], a1 = (T)sA[wm + 16 + il][k + kq];
    const T b0 = (T)sB[wn + il][k + kq], b1 = (T)sB[wn + 16 + il][k + kq];
    acc[0][0] = Mfma<T>::step(a0, b0, acc[0][0]);
    acc[0][1] = Mfma<T>::step(a0, b1, acc[0][1]);
    acc[1][0] = Mfma<T>::step(a1, b0, acc[1][0]);
    acc[1][1] = Mfma<T>::step(a1, b1, acc[1][1]);
  }
}

// Row (within the 16 x 16 tile) of accumulator register v of this lane: FP64
// 16x16x4 interleaves the four lane groups (4v + l/16), FP32 16x16x4 gives each
// lane group four consecutive rows (4(l/16) + v); the column is l % 16.
template <class T>
__device__ inline int acc_row(int v, int lane) {
  return Mfma<T>::kF64 ? 4 * v + (lane >> 4) : 4 * (lane >> 4) + v;
}

// X <- X V^T for the rows x nb panel X (leading dimension ldx), V the 64 x 64
// inverse of the panel's diagonal factor: X L^-T, one 64-row tile per block.
__global__ __launch_bounds__(256) void k_panel_apply(double* X, int64_t ldx, int64_t rows, int nb,
                                                     const double* V, const int* err) {
  if (err && *err) return;
  __shared__ double sA[64][65];
  __shared__ double sB[64][65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int i = e >> 6, j = e & 63;
    sA[i][j] = (r0 + i < rows && j < nb) ? X[(r0 + i) * ldx + j] : 0.0;
    sB[i][j] = V[e];
  }
  __syncthreads();
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  tile_nt<double>(sA, sB, wm, wn, lane, acc);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t gi = r0 + wm + a * 16 + acc_row<double>(v, lane);
        const int gj = wn + b * 16 + (lane & 15);
        if (gi < rows && gj < nb) X[gi * ldx + gj] = acc[a][b][v];
      }
}

// Trailing update of one panel step: C <- C - P P^T on the lower tiles of the
// rest x rest matrix C (at A, leading dimension lda), P the rest x 64 panel
// below the factored diagonal tile. Block b updates tile (ti, tj), ti >= tj,
// of the packed lower triangle, block 0 being tile (0, 0): that block keeps
// its updated tile in LDS, factors and inverts it (the next diagonal tile,
// nb_next x nb_next) and writes the factor and Vnext. In the mixed form the
// off-diagonal tiles are updated on FP32 MFMA (inputs rounded to FP32, FP32
// accumulation, FP64 storage); diagonal tiles always in FP64.
__global__ __launch_bounds__(256, 2) void k_trail(double* A, int64_t lda, int64_t rest,
                                               const double* P, int mixed, int nb_next,
                                               double* Vnext, int* err) {
  if (*err) return;
  __shared__ union {
    struct {
      double a[64][65];
      double b[64][65];
    } g;
    TileLds t;
  } sm;
  int ti, tj;
  tri_unpack((int)blockIdx.x, ti, tj);
  const int64_t i0 = (int64_t)ti * 64, j0 = (int64_t)tj * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, k = e & 63;
    sm.g.a[r][k] = i0 + r < rest ? P[(i0 + r) * lda + k] : 0.0;
    sm.g.b[r][k] = j0 + r < rest ? P[(j0 + r) * lda + k] : 0.0;
  }
  __syncthreads();
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  if (mixed && ti != tj) {
    v4f acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    tile_nt<float>(sm.g.a, sm.g.b, wm, wn, lane, acc);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t gi = i0 + wm + a * 16 + acc_row<float>(v, lane);
          const int64_t gj = j0 + wn + b * 16 + (lane & 15);
          if (gi < rest && gj < rest) A[gi * lda + gj] -= (double)acc[a][b][v];
        }
    return;
  }
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  tile_nt<double>(sm.g.a, sm.g.b, wm, wn, lane, acc);
  const bool look = blockIdx.x == 0 && nb_next > 0;
  if (look) __syncthreads();  // the panel tiles are read: the LDS becomes the tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ri = wm + a * 16 + acc_row<double>(v, lane);
        const int cj = wn + b * 16 + (lane & 15);
        const int64_t gi = i0 + ri, gj = j0 + cj;
        const bool in = gi < rest && gj < rest;
        const double c = in ? A[gi * lda + gj] - acc[a][b][v] : (ri == cj ? 1.0 : 0.0);
        if (look)
          sm.t.T[ri][cj] = c;
        else if (in)
          A[gi * lda + gj] = c;
      }
  if (!look) return;
  __syncthreads();
  const bool ok = chol_inv_tile(sm.t);
  store_tile(sm.t, ok, A, lda, nb_next, Vnext, err);
}

// x <- L^-T x for one vector (back substitution by 64-blocks), L: n x n lower
// (row-major) with the inverses V of its diagonal factors; one workgroup.
// x_q = V_qq^T (x_q - sum_{m > q} L_mq^T x_m), blocks from the last to the
// first. Values of x written by earlier blocks are read with device-scope
// loads (past this compute unit's L1).
__global__ __launch_bounds__(256) void k_backsolve_lt(const double* L, int64_t n, const double* V,
                                                      double* x, const int* err) {
  if (*err) return;
  __shared__ double part[4][64];
  __shared__ double t[64];
  const int tid = threadIdx.x, j = tid & 63, g = tid >> 6;
  const int64_t P = (n + 63) / 64;
  for (int64_t q = P - 1; q >= 0; --q) {
    const int64_t c0 = q * 64;
    const int nb = (int)std::min<int64_t>(64, n - c0);
    double acc = 0.0;
    if (j < nb)
      for (int64_t i = c0 + 64 + g; i < n; i += 4)
        acc = acc + L[i * n + c0 + j] *
                        __hip_atomic_load(x + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    part[g][j] = acc;
    __syncthreads();
    if (tid < 64)
      t[j] = j < nb ? __hip_atomic_load(x + c0 + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) -
                          (((part[0][j] + part[1][j]) + part[2][j]) + part[3][j])
                    : 0.0;
    __syncthreads();
    if (tid < 64 && j < nb) {
      double sum = 0.0;
      for (int m = j; m < nb; ++m) sum = sum + V[q * 4096 + m * 64 + j] * t[m];
      __hip_atomic_store(x + c0 + j, sum, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// ---- sampling ----------------------------------------------------------------
// One work item of the sampling product: output tile (ti, tj) of W L22^T
// (realization rows 64 ti.., cell columns 64 tj..) over K columns [k0, k1).
struct SampleItem {
  int32_t ti, tj, k0, k1;
};
constexpr int kSampleKC = 256;  // K columns per work item

// part[item] = W[rows of ti, k0:k1] L22[cols of tj, k0:k1]^T (64 x 64, FP64
// MFMA). 32-wide K slabs through LDS; the next slab is read into registers
// while the current one is multiplied.
__global__ __launch_bounds__(256) void k_sample_part(const double* W, const double* L, int64_t nS,
                                                     int R, const SampleItem* items,
                                                     double* part) {
  __shared__ double sA[64][33];
  __shared__ double sB[64][33];
  const SampleItem it = items[blockIdx.x];
  const int64_t r0 = (int64_t)it.ti * 64, a0 = (int64_t)it.tj * 64;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int il = lane & 15, kq = lane >> 4;
  double ra[8], rb[8];
  auto fetch = [&](int64_t k) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u, row = e >> 5, kk = e & 31;
      const int64_t gk = k + kk;
      ra[u] = (r0 + row < R && gk < it.k1) ? W[(r0 + row) * nS + gk] : 0.0;
      rb[u] = (a0 + row < nS && gk < it.k1) ? L[(a0 + row) * nS + gk] : 0.0;
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};
  fetch(it.k0);
  for (int64_t k = it.k0; k < it.k1; k += 32) {
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = tid + 256 * u;
      sA[e >> 5][e & 31] = ra[u];
      sB[e >> 5][e & 31] = rb[u];
    }
    __syncthreads();
    if (k + 32 < it.k1) fetch(k + 32);
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const double x0 = sA[wm + il][kk + kq], x1 = sA[wm + 16 + il][kk + kq];
      const double y0 = sB[wn + il][kk + kq], y1 = sB[wn + 16 + il][kk + kq];
      acc[0][0] = Mfma<double>::step(x0, y0, acc[0][0]);
      acc[0][1] = Mfma<double>::step(x0, y1, acc[0][1]);
      acc[1][0] = Mfma<double>::step(x1, y0, acc[1][0]);
      acc[1][1] = Mfma<double>::step(x1, y1, acc[1][1]);
    }
  }
  double* P = part + (int64_t)blockIdx.x * 4096;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        P[(wm + a * 16 + acc_row<double>(v, lane)) * 64 + wn + b * 16 + (lane & 15)] = acc[a][b][v];
}

// out[r][sinds[a]] = (d2[a] + sum of the tile's chunks in K order) + add and
// out[r][dinds[d]] = z1[d] (lusim.jl:162-172); tile_first[ti * TS + tj] is the
// first work item of a tile, whose items are consecutive.
__global__ __launch_bounds__(256) void k_sample_out(const double* part, const int32_t* tile_first,
                                                    int TS, int R, int64_t nS, int64_t N,
                                                    const double* d2, const int64_t* sinds,
                                                    const int64_t* dinds, const double* z1,
                                                    int64_t nD, double add, double* out) {
  const int r = blockIdx.y;
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < nS + nD;
       a += (int64_t)gridDim.x * blockDim.x) {
    if (a < nS) {
      const int tile = (r >> 6) * TS + (int)(a >> 6);
      const int first = tile_first[tile], last = tile_first[tile + 1];
      const int off = (r & 63) * 64 + (int)(a & 63);
      double sum = 0.0;
      for (int c = first; c < last; ++c) sum = sum + part[(int64_t)c * 4096 + off];
      out[(int64_t)r * N + sinds[a]] = (d2[a] + sum) + add;
    } else {
      out[(int64_t)r * N + dinds[a - nS]] = z1[a - nS];
    }
  }
}

// y <- y + x
__global__ void k_axpy(double* y, const double* x, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = y[t] + x[t];
}

__global__ void k_zero_upper(double* A, int64_t n) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = blockIdx.y;
  if (j < n && j > i) A[i * n + j] = 0.0;
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

// Right-looking blocked Cholesky of the n x n row-major matrix A (lower),
// 64-wide panels: the first diagonal tile by k_diag_chol_inv, then per panel
// k_panel_apply (the panel below the diagonal, with the inverse of its
// diagonal factor) and k_trail (the trailing update, and the next diagonal
// tile). Vinv receives the inverse of every diagonal factor (64 x 64 each).
// With tctx, every trailing update is bracketed by timing events (slot
// GSP_K_LU_TRAIL) while the context's timing is on.
void cholesky(hipStream_t s, double* A, int64_t n, bool mixed, int* err, double* Vinv,
              gsp_ctx* tctx) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_diag_chol_inv, dim3(1), dim3(256), 0, s, A, n,
                     (int)std::min<int64_t>(NB, n), Vinv, err);
  for (int64_t k0 = 0, q = 0; k0 < n; k0 += NB, ++q) {
    const int64_t k1 = k0 + NB, rest = n - k1;
    if (rest <= 0) break;
    double* panel = A + k1 * n + k0;
    hipLaunchKernelGGL(k_panel_apply, dim3((unsigned)((rest + 63) / 64)), dim3(256), 0, s, panel,
                       n, rest, NB, Vinv + q * NB * NB, err);
    const int64_t t = (rest + 63) / 64;
    TimedLaunch tl(tctx, GSP_K_LU_TRAIL);
    hipLaunchKernelGGL(k_trail, dim3((unsigned)(t * (t + 1) / 2)), dim3(256), 0, s,
                       A + k1 * n + k1, n, rest, panel, (int)mixed,
                       (int)std::min<int64_t>(NB, rest), Vinv + (q + 1) * NB * NB, err);
  }
}

// X <- X L^-T for X: M x n (leading dimension ldx), L: n x n lower (row-major)
// with the inverses Vinv of its diagonal factors (from cholesky).
void trsm_right_lt(hipStream_t s, const double* L, int64_t n, const double* Vinv, double* X,
                   int64_t M, int64_t ldx, int* err) {
  if (M <= 0) return;
  for (int64_t b0 = 0, q = 0; b0 < n; b0 += NB, ++q) {
    const int nb = (int)std::min<int64_t>(NB, n - b0);
    hipLaunchKernelGGL(k_panel_apply, dim3((unsigned)((M + 63) / 64)), dim3(256), 0, s, X + b0,
                       ldx, M, nb, Vinv + q * NB * NB, err);
    const int64_t b1 = b0 + nb;
    if (b1 < n)
      gemm_nt(s, false, M, n - b1, nb, X + b0, ldx, L + b1 * n + b0, n, X + b1, ldx, -1.0, 1.0,
              false, false, err);
  }
}

// `lower` (rows == cols): only the tiles on and below the diagonal.
void cov_matrix(hipStream_t s, const CovArgs& A, const int64_t* rows, const int64_t* cols,
                int64_t M, int64_t N, double* out, int64_t ldo, bool lower = false) {
  if (M <= 0 || N <= 0) return;
  const int64_t Tm = (M + 63) / 64, Tn = (N + 63) / 64;
  if (lower)
    hipLaunchKernelGGL(k_cov_tiles, dim3((unsigned)(Tm * (Tm + 1) / 2)), dim3(256), 0, s, A, rows,
                       cols, M, N, out, ldo, 1);
  else
    hipLaunchKernelGGL(k_cov_tiles, dim3((unsigned)Tn, (unsigned)Tm), dim3(256), 0, s, A, rows,
                       cols, M, N, out, ldo, 0);
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
  const int64_t pS = (nS + NB - 1) / NB, pD = (nD + NB - 1) / NB;
  DevBuf<double> VS;  // inverses of the diagonal factors of L22
  VS.alloc((size_t)std::max<int64_t>(pS, 1) * NB * NB);

  if (nS > 0) cov_matrix(s, C, plan->sinds.p, plan->sinds.p, nS, nS, plan->L22.p, nS, true);
  if (nD == 0) {
    GSP_HIP(hipMemsetAsync(plan->d2.p, 0, (size_t)std::max<int64_t>(nS, 1) * 8, s));
    cholesky(s, plan->L22.p, nS, mixed, err, VS.p, ctx);
  } else {
    DevBuf<double> L11, X, v, VD, C11, C21, x;
    L11.alloc((size_t)(nD * nD));
    X.alloc((size_t)std::max<int64_t>(nS * nD, 1));
    v.alloc((size_t)nD);
    VD.alloc((size_t)pD * NB * NB);
    cov_matrix(s, C, plan->dinds.p, plan->dinds.p, nD, nD, L11.p, nD, !mixed);
    cov_matrix(s, C, plan->sinds.p, plan->dinds.p, nS, nD, X.p, nD);  // C21
    if (mixed) {
      // FP64 copies for the refinement of the conditioning solve
      C11.alloc((size_t)(nD * nD));
      C21.alloc((size_t)std::max<int64_t>(nS * nD, 1));
      x.alloc((size_t)nD);
      GSP_HIP(hipMemcpyAsync(C11.p, L11.p, (size_t)(nD * nD) * 8, hipMemcpyDeviceToDevice, s));
      if (nS > 0)
        GSP_HIP(hipMemcpyAsync(C21.p, X.p, (size_t)(nS * nD) * 8, hipMemcpyDeviceToDevice, s));
    }
    cholesky(s, L11.p, nD, mixed, err, VD.p, nullptr);
    // X = B12^T = C21 L11^-T  (lusim.jl:99);  v = L11^-1 z1  (lusim.jl:102)
    if (nS > 0) trsm_right_lt(s, L11.p, nD, VD.p, X.p, nS, nD, err);
    GSP_HIP(hipMemcpyAsync(v.p, plan->z1.p, (size_t)nD * 8, hipMemcpyDeviceToDevice, s));
    trsm_right_lt(s, L11.p, nD, VD.p, v.p, 1, nD, err);
    if (nS > 0) {
      if (!mixed) {
        // d2 = B12^T v  (lusim.jl:102)
        gemm_nt(s, false, nS, 1, nD, X.p, nD, v.p, nD, plan->d2.p, 1, 1.0, 0.0, false, false,
                err);
      } else {
        // d2 = C21 C11^-1 z1 with one FP64 refinement step of the solve:
        // x0 = L11^-T v, r = z1 - C11 x0 (FP64), x = x0 + L11^-T L11^-1 r
        hipLaunchKernelGGL(k_backsolve_lt, dim3(1), dim3(256), 0, s, L11.p, nD, VD.p, v.p, err);
        GSP_HIP(hipMemcpyAsync(x.p, v.p, (size_t)nD * 8, hipMemcpyDeviceToDevice, s));
        GSP_HIP(hipMemcpyAsync(v.p, plan->z1.p, (size_t)nD * 8, hipMemcpyDeviceToDevice, s));
        gemm_nt(s, false, 1, nD, nD, x.p, nD, C11.p, nD, v.p, nD, -1.0, 1.0, false, false, err);
        trsm_right_lt(s, L11.p, nD, VD.p, v.p, 1, nD, err);
        hipLaunchKernelGGL(k_backsolve_lt, dim3(1), dim3(256), 0, s, L11.p, nD, VD.p, v.p, err);
        hipLaunchKernelGGL(k_axpy, dim3((unsigned)((nD + 255) / 256)), dim3(256), 0, s, x.p, v.p,
                           nD);
        gemm_nt(s, false, nS, 1, nD, C21.p, nD, x.p, nD, plan->d2.p, 1, 1.0, 0.0, false, false,
                err);
      }
      // C22 - B12^T B12  (lusim.jl:103)
      gemm_nt(s, false, nS, nS, nD, X.p, nD, X.p, nD, plan->L22.p, nS, -1.0, 1.0, true, false, err);
      cholesky(s, plan->L22.p, nS, mixed, err, VS.p, ctx);
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

void lusim_sample_chunk(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                        double* out);

// Realizations in chunks of at most 8192 (the output kernel's grid rows).
void lusim_sample_run(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                      double* out) {
  TimedLaunch tl(ctx, GSP_K_LU_SAMPLE);
  for (int32_t r0 = 0; r0 < R; r0 += 8192)
    lusim_sample_chunk(ctx, plan, std::min<int32_t>(8192, R - r0), w + (int64_t)r0 * plan->n_s,
                       out + (int64_t)r0 * plan->N);
}

void lusim_sample_chunk(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R, const double* w,
                        double* out) {
  hipStream_t s = ctx->stream;
  const int64_t nS = plan->n_s, nD = plan->n_d, N = plan->N;
  // work items: every output tile (ti, tj) of W L22^T over K [0, 64 tj + 64)
  // (L22 is lower triangular) in chunks of kSampleKC columns
  const int TR = (R + 63) / 64, TS = (int)((nS + 63) / 64);
  std::vector<SampleItem> items;
  std::vector<int32_t> first((size_t)TR * TS + 1, 0);
  for (int ti = 0; ti < TR; ++ti)
    for (int tj = 0; tj < TS; ++tj) {
      first[(size_t)ti * TS + tj] = (int32_t)items.size();
      const int kend = (int)std::min<int64_t>(nS, (int64_t)tj * 64 + 64);
      for (int k0 = 0; k0 < kend; k0 += kSampleKC)
        items.push_back({ti, tj, k0, std::min(kend, k0 + kSampleKC)});
    }
  first[(size_t)TR * TS] = (int32_t)items.size();
  const double add = nD == 0 ? plan->mean : 0.0;  // lusim.jl:172
  DevBuf<SampleItem> ditems;
  DevBuf<int32_t> dfirst;
  DevBuf<double> part;
  if (!items.empty()) {
    ditems.alloc(items.size());
    dfirst.alloc(first.size());
    part.alloc(items.size() * 4096);
    GSP_HIP(hipMemcpyAsync(ditems.p, items.data(), items.size() * sizeof(SampleItem),
                           hipMemcpyHostToDevice, s));
    GSP_HIP(hipMemcpyAsync(dfirst.p, first.data(), first.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_sample_part, dim3((unsigned)items.size()), dim3(256), 0, s, w,
                       plan->L22.p, nS, R, ditems.p, part.p);
  }
  hipLaunchKernelGGL(k_sample_out,
                     dim3((unsigned)std::min<int64_t>((nS + nD + 255) / 256, 4096), (unsigned)R),
                     dim3(256), 0, s, part.p, dfirst.p, TS, R, nS, N, plan->d2.p, plan->sinds.p,
                     plan->dinds.p, plan->z1.p, nD, add, out);
  GSP_HIP(hipGetLastError());
  GSP_HIP(hipStreamSynchronize(s));  // the items and partial sums are freed on return
}

}  // namespace gsp
