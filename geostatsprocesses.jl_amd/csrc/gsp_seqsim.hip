// gsp_seqsim.hip — SEQSIM on gfx950.
//
// The reference (src/simulation/field/seqsim.jl:98-138) walks the path and, for
// every node of every realization, searches neighbours, assembles and factors a
// k x k simple-kriging system and draws one normal. The neighbour sets, weights
// and variances depend only on the path and the geometry, so this file splits
// the work into
//   k_plan      one wavefront per node: neighbour search (offset table over the
//               implicit grid + a cell hash for the data) fused with covariance
//               assembly and a register-resident Cholesky solve;
//   k_levels    dependency depth of every node (diagnostic);
//   k_sweep     the only sequential part: a dataflow sweep over the path in
//               ticket order, every wave waiting on its own neighbours' ready
//               flags, R realizations across the lanes;
//   k_transpose X[node][r] -> out[r][node].
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <limits>

#include "gsp_internal.h"

namespace gsp {
namespace {

constexpr int kWaves = 4;  // wavefronts per 256-thread block
constexpr int kSpinLimit = 1 << 22;

enum { CTR_TICKET = 0, CTR_ERROR = 1, CTR_PRIOR = 2, CTR_FALLBACK = 3, CTR_DEPTH = 4, CTR_POS_ERR = 5 };

// ------------------------------------------------------------ wave helpers --
__device__ inline double rl(double x, int lane) {
  uint64_t u = __builtin_bit_cast(uint64_t, x);
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ inline float rlf(float x, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), lane));
}

__device__ inline int rli(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }

// number of set bits of `mask` below this lane
__device__ inline int mbcnt(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline bool key_less(double da, int ia, double db, int ib) {
  return da < db || (da == db && ia < ib);
}

// ------------------------------------------------------------ plan kernel --
struct PlanArgs {
  int32_t N, ndim, n_data;
  int32_t dims[3];
  int32_t kmin, kmax, mode, precision, T;
  int32_t hd[3];
  double r2;
  double org[3], sp[3];  // physical grid
  double ao[3], as[3];   // scaled grid (alpha*origin, alpha*spacing)
  double hcs[3];         // data-hash cell size (physical)
  double alpha, prior_sigma;
  DevModel M;
  const int64_t* path;
  const int32_t* pos;
  const Offset* table;
  const double* data;  // n_data x ndim, physical
  const int32_t* hstart;
  const int32_t* hitems;
  int32_t* nfound;
  int32_t* nbr;
  double* lam;
  double* sigma;
  uint8_t* flags;
  int32_t* counters;
};

// Squared physical distance between datum `di` and the centre of node t.
__device__ inline double data_d2(const PlanArgs& A, int di, const int* t) {
  double acc = 0.0;
  for (int d = 0; d < A.ndim; ++d) {
    double c = A.org[d] + ((double)t[d] + 0.5) * A.sp[d];
    double u = A.data[(int64_t)di * A.ndim + d] - c;
    acc = acc + u * u;
  }
  return acc;
}

// Simple-kriging solve of the n-neighbour system held one row per lane
// (lane j: neighbour j at scaled coordinates x). Same operation order as the
// oracle: right-looking Cholesky whose per-element subtraction order is
// ascending, forward substitution ascending, back substitution descending.
template <int KB>
__device__ inline int krige_fp64(const DevModel& M, int n, int lane, double x0, double x1,
                                 double x2, double t0, double t1, double t2, double C0,
                                 double& lam_out, double& sigma_out) {
  double a[KB];
#pragma unroll
  for (int m = 0; m < KB; ++m) {
    if (m < n) {
      a[m] = cov_points(M, x0, x1, x2, rl(x0, m), rl(x1, m), rl(x2, m));
    } else {
      a[m] = 0.0;
    }
  }
  const double c = cov_points(M, x0, x1, x2, t0, t1, t2);
#pragma unroll
  for (int col = 0; col < KB; ++col) {
    if (col < n) {
      const double piv = rl(a[col], col);
      if (!(piv > 0.0)) return GSP_FLAG_FAIL;
      const double lcc = __builtin_sqrt(piv);
      a[col] = lane == col ? lcc : (lane > col ? a[col] / lcc : a[col]);
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
      const double ym = rl(s, m) / rl(a[m], m);
      if (lane == m) y = ym;
      if (lane > m) s = s - a[m] * ym;
    }
  }
  double s2 = y, lam = 0.0;
#pragma unroll
  for (int m = KB - 1; m >= 0; --m) {
    if (m < n) {
      const double lm = rl(s2, m) / rl(a[m], m);
      if (lane == m) lam = lm;
#pragma unroll
      for (int jj = 0; jj < m; ++jj) {
        const double v = rl(a[jj], m);
        if (lane == jj) s2 = s2 - v * lm;
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < KB; ++j)
    if (j < n) acc = acc + rl(lam, j) * rl(c, j);
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
__device__ inline int krige_fp32r(const DevModel& M, int n, int lane, double x0, double x1,
                                  double x2, double t0, double t1, double t2, double C0,
                                  double& lam_out, double& sigma_out) {
  const float tau = (float)(M.sill * 0x1p-12);
  double a[KB];
  float f[KB];
#pragma unroll
  for (int m = 0; m < KB; ++m) {
    if (m < n) {
      a[m] = cov_points(M, x0, x1, x2, rl(x0, m), rl(x1, m), rl(x2, m));
    } else {
      a[m] = 0.0;
    }
    f[m] = (float)a[m];
  }
  const double c = cov_points(M, x0, x1, x2, t0, t1, t2);
  bool ok = true;
#pragma unroll
  for (int col = 0; col < KB; ++col) {
    if (ok && col < n) {
      const float piv = rlf(f[col], col);
      if (!(piv > tau)) {
        ok = false;
      } else {
        const float lcc = __builtin_sqrtf(piv);
        f[col] = lane == col ? lcc : (lane > col ? f[col] / lcc : f[col]);
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
  if (!ok) {
    const int r = krige_fp64<KB>(M, n, lane, x0, x1, x2, t0, t1, t2, C0, lam_out, sigma_out);
    return r | GSP_FLAG_FP64_FALLBACK;
  }
  // two FP32 solves with the same factor: lambda0 = A^-1 c, delta = A^-1 r
  float rhs = (float)c, sol = 0.0f;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    float s = rhs, y = 0.0f;
#pragma unroll
    for (int m = 0; m < KB; ++m) {
      if (m < n) {
        const float ym = rlf(s, m) / rlf(f[m], m);
        if (lane == m) y = ym;
        if (lane > m) s = s - f[m] * ym;
      }
    }
    float s2 = y, l = 0.0f;
#pragma unroll
    for (int m = KB - 1; m >= 0; --m) {
      if (m < n) {
        const float lm = rlf(s2, m) / rlf(f[m], m);
        if (lane == m) l = lm;
#pragma unroll
        for (int jj = 0; jj < m; ++jj) {
          const float v = rlf(f[jj], m);
          if (lane == jj) s2 = s2 - v * lm;
        }
      }
    }
    if (pass == 0) {
      sol = l;
      double r = c;
#pragma unroll
      for (int m = 0; m < KB; ++m)
        if (m < n) r = r - a[m] * (double)rlf(l, m);
      rhs = (float)r;
    } else {
      lam_out = (double)sol + (double)l;
    }
  }
  const double lam = lam_out;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < KB; ++j)
    if (j < n) acc = acc + rl(lam, j) * rl(c, j);
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  sigma_out = __builtin_sqrt(var);
  return 0;
}

template <int KB, int PREC>
__global__ __launch_bounds__(256) void k_plan(const PlanArgs A) {
  __shared__ int s_gidx[kWaves][64];
  __shared__ double s_gd2[kWaves][64];
  __shared__ int s_didx[kWaves][128];
  __shared__ double s_dd2[kWaves][128];
  __shared__ int s_fidx[kWaves][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int nx = A.dims[0], ny = A.dims[1], nz = A.dims[2];
  const int64_t wstride = (int64_t)gridDim.x * kWaves;

  for (int64_t q = (int64_t)blockIdx.x * kWaves + w; q < A.N; q += wstride) {
    const int i = (int)A.path[q];
    const int pi = (int)q;
    int ti[3];
    ti[0] = i % nx;
    ti[1] = (i / nx) % ny;
    ti[2] = i / (nx * ny);

    // ---- grid candidates: the first kmax usable offsets in (d2, index) order
    int cnt = 0;
    for (int t0 = 0; cnt < A.kmax && t0 < A.T; t0 += 64) {
      const int t = t0 + lane;
      bool ok = false;
      int j = 0;
      double d2 = 0.0;
      if (t < A.T) {
        const Offset e = A.table[t];
        const int x = ti[0] + e.dx, y = ti[1] + e.dy, z = ti[2] + e.dz;
        ok = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
        if (ok) {
          j = x + nx * (y + ny * z);
          d2 = e.d2;
          if (A.mode == GSP_MASKED_KNN && A.pos[j] >= pi) ok = false;
        }
      }
      const uint64_t b = __ballot(ok);
      const int rank = cnt + mbcnt(b);
      if (ok && rank < A.kmax) {
        s_gidx[w][rank] = j;
        s_gd2[w][rank] = d2;
      }
      cnt += __popcll(b);
    }
    const int ng = cnt < A.kmax ? cnt : A.kmax;
    wave_sync();
    const double Dk = ng == A.kmax ? s_gd2[w][ng - 1] : A.r2;

    // ---- data candidates with d2 <= Dk, from the cell hash
    int nd = 0;
    if (A.n_data > 0) {
      int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
      const bool all = !(Dk < __builtin_inf());
      const double rho = all ? 0.0 : __builtin_sqrt(Dk) * (1.0 + 0x1p-20) + 0x1p-40;
      for (int d = 0; d < A.ndim; ++d) {
        if (all) {
          lo[d] = 0;
          hi[d] = A.hd[d] - 1;
        } else {
          const double c = A.org[d] + ((double)ti[d] + 0.5) * A.sp[d];
          double fl = __builtin_floor((c - rho - A.org[d]) / A.hcs[d]);
          double fh = __builtin_floor((c + rho - A.org[d]) / A.hcs[d]);
          fl = fl < 0.0 ? 0.0 : (fl > A.hd[d] - 1 ? A.hd[d] - 1 : fl);
          fh = fh < 0.0 ? 0.0 : (fh > A.hd[d] - 1 ? A.hd[d] - 1 : fh);
          lo[d] = (int)fl;
          hi[d] = (int)fh;
        }
      }
      for (int cz = lo[2]; cz <= hi[2]; ++cz)
        for (int cy = lo[1]; cy <= hi[1]; ++cy)
          for (int cx = lo[0]; cx <= hi[0]; ++cx) {
            const int cell = cx + A.hd[0] * (cy + A.hd[1] * cz);
            const int s = A.hstart[cell], e = A.hstart[cell + 1];
            for (int b0 = s; b0 < e; b0 += 64) {
              bool ok = false;
              int di = 0;
              double d2 = 0.0;
              if (b0 + lane < e) {
                di = A.hitems[b0 + lane];
                d2 = data_d2(A, di, ti);
                ok = d2 <= Dk && d2 <= A.r2;
              }
              const uint64_t b = __ballot(ok);
              const int add = __popcll(b);
              if (add == 0) continue;
              if (nd + add > 128) {
                // keep the best kmax entries found so far
                wave_sync();
                int keep_rank[2] = {128, 128};
                int kid[2] = {0, 0};
                double kd[2] = {0.0, 0.0};
                for (int h = 0; h < 2; ++h) {
                  const int m = lane + 64 * h;
                  if (m < nd) {
                    kd[h] = s_dd2[w][m];
                    kid[h] = s_didx[w][m];
                    int r = 0;
                    for (int m2 = 0; m2 < nd; ++m2)
                      r += key_less(s_dd2[w][m2], s_didx[w][m2], kd[h], kid[h]);
                    keep_rank[h] = r;
                  }
                }
                wave_sync();
                for (int h = 0; h < 2; ++h)
                  if (keep_rank[h] < A.kmax) {
                    s_dd2[w][keep_rank[h]] = kd[h];
                    s_didx[w][keep_rank[h]] = kid[h];
                  }
                nd = nd < A.kmax ? nd : A.kmax;
                wave_sync();
              }
              if (ok) {
                const int r = nd + mbcnt(b);
                s_didx[w][r] = A.N + di;
                s_dd2[w][r] = d2;
              }
              nd += add;
            }
          }
      wave_sync();
    }

    // ---- merge: top-kmax of the grid list (sorted) and the data list
    const int nf = ng + nd < A.kmax ? ng + nd : A.kmax;
    int my_rank[3] = {128, 128, 128};
    int my_id[3] = {0, 0, 0};
    if (lane < ng) {
      const double d = s_gd2[w][lane];
      const int id = s_gidx[w][lane];
      int r = lane;
      for (int m = 0; m < nd; ++m) r += key_less(s_dd2[w][m], s_didx[w][m], d, id);
      my_rank[0] = r;
      my_id[0] = id;
    }
    for (int h = 0; h < 2; ++h) {
      const int m = lane + 64 * h;
      if (m < nd) {
        const double d = s_dd2[w][m];
        const int id = s_didx[w][m];
        int r = 0;
        for (int g = 0; g < ng; ++g) r += key_less(s_gd2[w][g], s_gidx[w][g], d, id);
        for (int m2 = 0; m2 < nd; ++m2) r += key_less(s_dd2[w][m2], s_didx[w][m2], d, id);
        my_rank[1 + h] = r;
        my_id[1 + h] = id;
      }
    }
    wave_sync();
    for (int h = 0; h < 3; ++h)
      if (my_rank[h] < A.kmax) s_fidx[w][my_rank[h]] = my_id[h];
    wave_sync();

    // ---- mask: drop points not yet simulated (k nearest first, then mask)
    const int fi = lane < nf ? s_fidx[w][lane] : -1;
    const bool keep =
        lane < nf && (A.mode == GSP_MASKED_KNN || fi >= A.N || A.pos[fi] < pi);
    const uint64_t kb = __ballot(keep);
    const int n = __popcll(kb);
    const int kr = mbcnt(kb);
    wave_sync();
    if (keep) s_fidx[w][kr] = fi;
    wave_sync();
    const int nb = lane < n ? s_fidx[w][lane] : -1;

    // ---- kriging
    int flag = 0;
    double lam = 0.0, sig = A.prior_sigma;
    if (n < A.kmin) {
      flag = GSP_FLAG_FEW;
    } else {
      double x0 = 0.0, x1 = 0.0, x2 = 0.0;
      if (lane < n) {
        if (nb < A.N) {
          const int c0 = nb % nx, c1 = (nb / nx) % ny, c2 = nb / (nx * ny);
          x0 = A.ao[0] + ((double)c0 + 0.5) * A.as[0];
          x1 = A.ao[1] + ((double)c1 + 0.5) * A.as[1];
          x2 = A.ao[2] + ((double)c2 + 0.5) * A.as[2];
        } else {
          const double* p = A.data + (int64_t)(nb - A.N) * A.ndim;
          x0 = A.alpha * p[0];
          if (A.ndim > 1) x1 = A.alpha * p[1];
          if (A.ndim > 2) x2 = A.alpha * p[2];
        }
      }
      const double t0 = A.ao[0] + ((double)ti[0] + 0.5) * A.as[0];
      const double t1 = A.ao[1] + ((double)ti[1] + 0.5) * A.as[1];
      const double t2 = A.ao[2] + ((double)ti[2] + 0.5) * A.as[2];
      if (PREC == GSP_FP32_REFINE)
        flag = krige_fp32r<KB>(A.M, n, lane, x0, x1, x2, t0, t1, t2, A.M.sill, lam, sig);
      else
        flag = krige_fp64<KB>(A.M, n, lane, x0, x1, x2, t0, t1, t2, A.M.sill, lam, sig);
      if (flag & GSP_FLAG_FAIL) {
        lam = 0.0;
        sig = A.prior_sigma;
      }
    }
    if (lane < A.kmax) {
      A.nbr[(int64_t)i * A.kmax + lane] = nb;
      A.lam[(int64_t)i * A.kmax + lane] = lane < n ? lam : 0.0;
    }
    if (lane == 0) {
      A.sigma[i] = sig;
      A.flags[i] = (uint8_t)flag;
      A.nfound[i] = n;
      if (flag & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) atomicAdd(&A.counters[CTR_PRIOR], 1);
      if (flag & GSP_FLAG_FP64_FALLBACK) atomicAdd(&A.counters[CTR_FALLBACK], 1);
    }
  }
}

// ---------------------------------------------------------- small kernels --
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

// pos[path[q]] = q, flagging out-of-range ids and repeats.
__global__ void k_positions(const int64_t* path, int64_t N, int32_t* pos, int32_t* counters) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = path[q];
    if (i < 0 || i >= N) {
      atomicOr(&counters[CTR_POS_ERR], 1);
      continue;
    }
    const int old = atomicExch(&pos[i], (int)q);
    if (old != -1) atomicOr(&counters[CTR_POS_ERR], 2);
  }
}

__device__ inline bool wave_all(bool v) { return __ballot(v) == __ballot(true); }

// Dependency depth: level(i) = 0 without grid neighbours, else
// 1 + max level of its grid neighbours. Waves take 64 consecutive path
// positions in ticket order; a neighbour always has a smaller position, so the
// wave holding the earliest unfinished position can always complete.
__global__ __launch_bounds__(256) void k_levels(const int64_t* path, int32_t N, int32_t kmax,
                                                const int32_t* nfound, const uint8_t* flags,
                                                const int32_t* nbr, int32_t* level,
                                                int32_t* counters) {
  const int lane = threadIdx.x & 63;
  while (true) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&counters[CTR_TICKET], 64);
    base = rli(base, 0);
    if (base >= N) break;
    const int q = base + lane;
    const bool act = q < N;
    const int i = act ? (int)path[q] : 0;
    const int k = act && !(flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) ? nfound[i] : 0;
    const int32_t* nb = nbr + (int64_t)i * kmax;
    int j = 0, lv = 0, spins = 0;
    bool done = !act;
    while (true) {
      if (!done) {
        while (j < k) {
          const int a = nb[j];
          if (a < N) {
            const int la = __hip_atomic_load(&level[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (la < 0) break;
            lv = la + 1 > lv ? la + 1 : lv;
          }
          ++j;
        }
        if (j == k || ++spins > kSpinLimit) {
          if (j != k) atomicOr(&counters[CTR_ERROR], 1);
          __hip_atomic_store(&level[i], lv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          atomicMax(&counters[CTR_DEPTH], lv + 1);
          done = true;
        }
      }
      if (wave_all(done)) break;
      __builtin_amdgcn_s_sleep(2);
    }
  }
}

// ------------------------------------------------------------- the sweep --
struct SweepArgs {
  const int64_t* path;
  int32_t n_path, N, kmax;
  int32_t RC;     // realizations in this chunk (<= 64)
  int32_t shift;  // log2 of lanes per node
  int32_t r0;     // first realization of the chunk
  const int32_t* nfound;
  const uint8_t* flags;
  const int32_t* nbr;
  const double* lam;
  const double* sigma;
  double mean;
  const double* z;  // R x n_path
  double* X;        // (N + n_data) x RC
  uint32_t* ready;
  uint32_t epoch;
  int32_t* counters;
};

// One wave serves 64 >> shift consecutive path positions, 1 << shift lanes per
// node (lane = realization). A node is computed once all its grid neighbours
// carry this sweep's epoch; then its values are stored and its flag released.
// z reads are strided per lane, but consecutive tickets touch consecutive
// addresses, so each z line is fetched once and reused from L2.
__global__ __launch_bounds__(256) void k_sweep(const SweepArgs A) {
  const int lane = threadIdx.x & 63;
  const int per = 1 << A.shift;
  const int G = 64 >> A.shift;
  const int slot = lane >> A.shift;
  const int r = lane & (per - 1);
  while (true) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&A.counters[CTR_TICKET], G);
    base = rli(base, 0);
    if (base >= A.n_path) break;
    const int q = base + slot;
    const bool act = q < A.n_path;
    const bool lact = act && r < A.RC;
    const int i = act ? (int)A.path[q] : 0;
    const int k = act && !(A.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) ? A.nfound[i] : 0;
    const int32_t* nb = A.nbr + (int64_t)i * A.kmax;
    const double* lm = A.lam + (int64_t)i * A.kmax;
    int j = 0, spins = 0;
    bool done = !act;
    while (true) {
      if (!done) {
        while (j < k) {
          const int a = nb[j];
          if (a < A.N &&
              __hip_atomic_load(&A.ready[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != A.epoch)
            break;
          ++j;
        }
        const bool ready = j == k;
        if (ready || ++spins > kSpinLimit) {
          if (!ready) atomicOr(&A.counters[CTR_ERROR], 2);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          double acc = 0.0;
          for (int jj = 0; jj < k; ++jj) {
            const int a = nb[jj];
            const double xa = lact ? A.X[(int64_t)a * A.RC + r] : 0.0;
            acc = acc + lm[jj] * (xa - A.mean);
          }
          const double mu = A.mean + acc;
          if (lact) {
            const double zz = A.z[(int64_t)(A.r0 + r) * A.n_path + q];
            A.X[(int64_t)i * A.RC + r] = mu + A.sigma[i] * zz;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
          if (r == 0)
            __hip_atomic_store(&A.ready[i], A.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          done = true;
        }
      }
      if (wave_all(done)) break;
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// X rows N..N+n_data-1 hold the conditioning data for every realization.
__global__ void k_data_rows(double* X, int32_t N, int32_t n_data, int32_t RC,
                            const double* vals) {
  const int64_t total = (int64_t)n_data * RC;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    X[(int64_t)N * RC + t] = vals[t / RC];
}

// out[(r0 + r) * N + i] = X[i * RC + r] through a 64-node LDS tile.
__global__ __launch_bounds__(256) void k_transpose(const double* X, int32_t N, int32_t RC,
                                                   int32_t r0, double* out) {
  __shared__ double tile[64][65];
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  const int cnt = (int)((N - i0) < 64 ? (N - i0) : 64);
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int ii = t / RC, rr = t - ii * RC;
    tile[rr][ii] = X[(i0 + ii) * RC + rr];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int rr = t / cnt, ii = t - rr * cnt;
    out[(int64_t)(r0 + rr) * N + i0 + ii] = tile[rr][ii];
  }
}

// ------------------------------------------------------------- host side --
int grid_for(int64_t work, int per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// Offsets sorted by (squared distance, linear index delta). Within the search
// ball when there is one; otherwise far enough that every node finds kmax
// nodes (k nearest first), or the whole grid (masked search).
std::vector<Offset> build_offsets(const gsp_grid& g, double r2, int kmax, int mode) {
  const int nd = g.ndim;
  int64_t ext[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) ext[d] = g.dims[d] - 1;
  int64_t lim[3] = {0, 0, 0};
  double bound = r2;
  const bool ball = r2 < std::numeric_limits<double>::infinity();
  if (ball) {
    for (int d = 0; d < nd; ++d) {
      const double f = std::floor(std::sqrt(r2) / g.spacing[d]) + 1.0;
      lim[d] = std::min<int64_t>(ext[d], (int64_t)std::min(f, 1e9));
    }
  } else if (mode == GSP_KNN_THEN_MASK) {
    int64_t N = 1;
    for (int d = 0; d < nd; ++d) N *= g.dims[d];
    const int64_t need = std::min<int64_t>(kmax + 1, N);
    int64_t maxext = 0;
    for (int d = 0; d < nd; ++d) maxext = std::max(maxext, ext[d]);
    int64_t M = 0;
    for (;; ++M) {
      int64_t c = 1;
      for (int d = 0; d < nd; ++d) c *= std::min(M, ext[d]) + 1;
      if (c >= need || M >= maxext) break;
    }
    double acc = 0.0;
    for (int d = 0; d < nd; ++d) {
      const double t = (double)M * g.spacing[d];
      acc = acc + t * t;
    }
    bound = acc;
    for (int d = 0; d < nd; ++d) {
      const double f = std::floor(std::sqrt(bound) / g.spacing[d]) + 1.0;
      lim[d] = std::min<int64_t>(ext[d], (int64_t)f);
    }
  } else {
    for (int d = 0; d < nd; ++d) lim[d] = ext[d];
  }
  int64_t total = 1;
  for (int d = 0; d < 3; ++d) total *= 2 * lim[d] + 1;
  GSP_REQUIRE(total <= (int64_t)1 << 25, GSP_ERR_UNSUPPORTED,
              "neighbour search region too large (%lld offsets); use a search radius",
              (long long)total);
  GSP_REQUIRE(lim[0] < 32768 && lim[1] < 32768 && lim[2] < 32768, GSP_ERR_UNSUPPORTED,
              "grid too large for the offset table");
  std::vector<Offset> v;
  std::vector<int64_t> lin;
  v.reserve((size_t)total);
  for (int64_t dz = -lim[2]; dz <= lim[2]; ++dz)
    for (int64_t dy = -lim[1]; dy <= lim[1]; ++dy)
      for (int64_t dx = -lim[0]; dx <= lim[0]; ++dx) {
        const int64_t dd[3] = {dx, dy, dz};
        double acc = 0.0;
        for (int d = 0; d < nd; ++d) {
          const double t = (double)dd[d] * g.spacing[d];
          acc = acc + t * t;
        }
        if (!(acc <= bound)) continue;
        Offset o;
        o.d2 = acc;
        o.dx = (int16_t)dx;
        o.dy = (int16_t)dy;
        o.dz = (int16_t)dz;
        o.pad_ = 0;
        v.push_back(o);
      }
  const int64_t nx = g.dims[0], ny = nd > 1 ? g.dims[1] : 1;
  std::sort(v.begin(), v.end(), [&](const Offset& a, const Offset& b) {
    if (a.d2 != b.d2) return a.d2 < b.d2;
    const int64_t la = a.dx + nx * (a.dy + ny * (int64_t)a.dz);
    const int64_t lb = b.dx + nx * (b.dy + ny * (int64_t)b.dz);
    return la < lb;
  });
  return v;
}

// seqsim.jl:162-189
double scale_factor(const gsp_grid& g, const std::vector<double>& data, const gsp_model& m) {
  const int nd = g.ndim;
  double a1 = 0.0;
  for (int d = 0; d < nd; ++d) {
    const double lo = std::fabs(g.origin[d]);
    const double hi = std::fabs(g.origin[d] + (double)g.dims[d] * g.spacing[d]);
    a1 = std::max(a1, std::max(lo, hi));
  }
  double a2 = 0.0;
  for (double v : data) a2 = std::max(a2, std::fabs(v));
  double a3 = 0.0;
  for (int d = 0; d < nd; ++d) a3 = std::max(a3, m.ranges[d]);
  return 1.0 / std::max(a1, std::max(a2, a3));
}

}  // namespace

DevModel make_dev_model(const gsp_model& m, int ndim, double alpha) {
  DevModel M;
  std::memset(&M, 0, sizeof(M));
  M.family = m.family;
  M.ndim = ndim;
  M.has_rotation = m.has_rotation ? 1 : 0;
  M.sill = m.sill;
  M.nugget = m.nugget;
  for (int d = 0; d < 3; ++d) M.rs[d] = d < ndim ? alpha * m.ranges[d] : 1.0;
  for (int q = 0; q < 9; ++q) M.rot[q] = m.rotation[q];
  return M;
}

void validate_model(const gsp_model& m, int ndim) {
  GSP_REQUIRE(m.family >= GSP_SPHERICAL && m.family <= GSP_GAUSSIAN, GSP_ERR_UNSUPPORTED,
              "unknown variogram family %d", m.family);
  GSP_REQUIRE(std::isfinite(m.sill) && std::isfinite(m.nugget) && m.nugget >= 0 &&
                  m.sill >= m.nugget,
              GSP_ERR_ARGUMENT, "invalid sill/nugget (%g, %g)", m.sill, m.nugget);
  for (int d = 0; d < ndim; ++d)
    GSP_REQUIRE(std::isfinite(m.ranges[d]) && m.ranges[d] > 0, GSP_ERR_ARGUMENT,
                "invalid range %g", m.ranges[d]);
}

int64_t validate_grid(const gsp_grid& g) {
  GSP_REQUIRE(g.ndim >= 1 && g.ndim <= 3, GSP_ERR_ARGUMENT, "grid dimension must be 1..3");
  int64_t N = 1;
  for (int d = 0; d < g.ndim; ++d) {
    GSP_REQUIRE(g.dims[d] >= 1, GSP_ERR_ARGUMENT, "grid size must be positive");
    GSP_REQUIRE(std::isfinite(g.spacing[d]) && g.spacing[d] > 0 && std::isfinite(g.origin[d]),
                GSP_ERR_ARGUMENT, "invalid grid origin/spacing");
    N *= g.dims[d];
    GSP_REQUIRE(N < ((int64_t)1 << 30), GSP_ERR_UNSUPPORTED, "grid too large");
  }
  return N;
}

template <int KB>
void launch_plan(const PlanArgs& A, int precision, int blocks, hipStream_t s) {
  if (precision == GSP_FP32_REFINE)
    hipLaunchKernelGGL((k_plan<KB, GSP_FP32_REFINE>), dim3(blocks), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL((k_plan<KB, GSP_FP64>), dim3(blocks), dim3(256), 0, s, A);
}

void seqsim_plan_build(gsp_ctx* ctx, const gsp_seqsim_problem* p, gsp_seqsim_plan* plan) {
  GSP_REQUIRE(p, GSP_ERR_ARGUMENT, "null problem");
  const gsp_grid& g = p->grid;
  const int64_t N = validate_grid(g);
  const int nd = g.ndim;
  validate_model(p->model, nd);
  GSP_REQUIRE(p->n_data >= 0 && N + p->n_data < ((int64_t)1 << 30), GSP_ERR_ARGUMENT,
              "invalid number of data points");
  GSP_REQUIRE(p->n_data == 0 || p->data_coords, GSP_ERR_ARGUMENT, "null data coordinates");
  GSP_REQUIRE(p->n_path == N && p->path, GSP_ERR_UNSUPPORTED,
              "the path must visit every grid node exactly once");
  GSP_REQUIRE(p->search_mode == GSP_KNN_THEN_MASK || p->search_mode == GSP_MASKED_KNN,
              GSP_ERR_ARGUMENT, "unknown search mode");
  GSP_REQUIRE(p->precision == GSP_FP64 || p->precision == GSP_FP32_REFINE, GSP_ERR_ARGUMENT,
              "unknown precision");
  GSP_REQUIRE(std::isfinite(p->mean), GSP_ERR_ARGUMENT, "invalid mean");
  // seqsim.jl:202-208
  int32_t nmin = p->minneighbors, nmax = p->maxneighbors;
  if ((int64_t)nmax > N || nmax < 1) nmax = (int32_t)N;
  if (nmin > nmax || nmin < 1) nmin = 1;
  GSP_REQUIRE(nmax <= 64, GSP_ERR_UNSUPPORTED, "maxneighbors above 64 is not supported");

  hipStream_t s = ctx->stream;
  const int64_t nD = p->n_data;
  std::vector<double> data((size_t)(nD * nd));
  if (nD > 0) {
    if (p->device_pointers)
      GSP_HIP(hipMemcpy(data.data(), p->data_coords, data.size() * sizeof(double),
                        hipMemcpyDeviceToHost));
    else
      std::memcpy(data.data(), p->data_coords, data.size() * sizeof(double));
    for (double v : data) GSP_REQUIRE(std::isfinite(v), GSP_ERR_ARGUMENT, "non-finite data coordinate");
  }
  const double alpha = scale_factor(g, data, p->model);
  const double inf = std::numeric_limits<double>::infinity();
  const double r2 = p->radius > 0 ? p->radius * p->radius : inf;
  std::vector<Offset> table = build_offsets(g, r2, nmax, p->search_mode);

  // data hash: cells of B x B x B grid cells, clamped to the grid box
  int hd[3] = {1, 1, 1};
  double hcs[3] = {1.0, 1.0, 1.0};
  std::vector<int32_t> hstart(2, 0), hitems(1, 0);
  if (nD > 0) {
    int64_t B = (int64_t)std::llround(std::pow((double)N / (double)nD, 1.0 / nd));
    if (B < 1) B = 1;
    int64_t ncell = 1;
    for (int d = 0; d < nd; ++d) {
      hd[d] = (int)((g.dims[d] + B - 1) / B);
      hcs[d] = (double)B * g.spacing[d];
      ncell *= hd[d];
    }
    std::vector<int64_t> cell((size_t)nD);
    hstart.assign((size_t)ncell + 1, 0);
    for (int64_t i = 0; i < nD; ++i) {
      int64_t c = 0, stride = 1;
      for (int d = 0; d < nd; ++d) {
        double f = std::floor((data[(size_t)(i * nd + d)] - g.origin[d]) / hcs[d]);
        f = std::min(std::max(f, 0.0), (double)(hd[d] - 1));
        c += (int64_t)f * stride;
        stride *= hd[d];
      }
      cell[(size_t)i] = c;
      hstart[(size_t)c + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) hstart[(size_t)c + 1] += hstart[(size_t)c];
    hitems.assign((size_t)nD, 0);
    std::vector<int32_t> fill((size_t)ncell, 0);
    for (int64_t i = 0; i < nD; ++i)
      hitems[(size_t)(hstart[(size_t)cell[(size_t)i]] + fill[(size_t)cell[(size_t)i]]++)] = (int32_t)i;
  }

  plan->N = N;
  plan->n_data = nD;
  plan->ndim = nd;
  plan->kmin = nmin;
  plan->kmax = nmax;
  plan->alpha = alpha;
  plan->mean = p->mean;
  plan->path.alloc((size_t)N);
  plan->nfound.alloc((size_t)N);
  plan->nbr.alloc((size_t)(N * nmax));
  plan->lam.alloc((size_t)(N * nmax));
  plan->sigma.alloc((size_t)N);
  plan->flags.alloc((size_t)N);
  plan->level.alloc((size_t)N);
  plan->ready.alloc((size_t)N);

  DevBuf<int32_t> pos;
  pos.alloc((size_t)N);
  DevBuf<Offset> dtable;
  dtable.alloc(table.size());
  DevBuf<double> ddata;
  ddata.alloc((size_t)std::max<int64_t>(1, nD * nd));
  DevBuf<int32_t> dhstart, dhitems;
  dhstart.alloc(hstart.size());
  dhitems.alloc(hitems.size());
  ctx->counters.reserve(16);

  GSP_HIP(hipMemcpyAsync(plan->path.p, p->path, (size_t)N * sizeof(int64_t),
                         p->device_pointers ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dtable.p, table.data(), table.size() * sizeof(Offset),
                         hipMemcpyHostToDevice, s));
  if (nD > 0)
    GSP_HIP(hipMemcpyAsync(ddata.p, data.data(), data.size() * sizeof(double),
                           hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dhstart.p, hstart.data(), hstart.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dhitems.p, hitems.data(), hitems.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctx->counters.p, 0, (int64_t)16);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, pos.p, -1, N);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, plan->level.p, -1, N);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, plan->ready.p, 0u, N);
  hipLaunchKernelGGL(k_positions, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s, plan->path.p, N,
                     pos.p, ctx->counters.p);
  GSP_HIP(hipGetLastError());
  int32_t pos_err = 0;
  GSP_HIP(hipMemcpyAsync(&pos_err, ctx->counters.p + CTR_POS_ERR, sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  GSP_HIP(hipStreamSynchronize(s));
  GSP_REQUIRE(pos_err == 0, GSP_ERR_UNSUPPORTED,
              "the path must be a permutation of the grid ids 0..N-1");

  PlanArgs A;
  std::memset(&A, 0, sizeof(A));
  A.N = (int32_t)N;
  A.ndim = nd;
  A.n_data = (int32_t)nD;
  for (int d = 0; d < 3; ++d) {
    A.dims[d] = d < nd ? (int32_t)g.dims[d] : 1;
    A.org[d] = d < nd ? g.origin[d] : 0.0;
    A.sp[d] = d < nd ? g.spacing[d] : 1.0;
    A.ao[d] = d < nd ? alpha * g.origin[d] : 0.0;
    A.as[d] = d < nd ? alpha * g.spacing[d] : 0.0;
    A.hd[d] = hd[d];
    A.hcs[d] = hcs[d];
  }
  A.kmin = nmin;
  A.kmax = nmax;
  A.mode = p->search_mode;
  A.precision = p->precision;
  A.T = (int32_t)table.size();
  A.r2 = r2;
  A.alpha = alpha;
  A.prior_sigma = std::sqrt(p->model.sill);
  A.M = make_dev_model(p->model, nd, alpha);
  A.path = plan->path.p;
  A.pos = pos.p;
  A.table = dtable.p;
  A.data = ddata.p;
  A.hstart = dhstart.p;
  A.hitems = dhitems.p;
  A.nfound = plan->nfound.p;
  A.nbr = plan->nbr.p;
  A.lam = plan->lam.p;
  A.sigma = plan->sigma.p;
  A.flags = plan->flags.p;
  A.counters = ctx->counters.p;

  const int blocks = grid_for(N, kWaves, 16384);
  {
  TimedLaunch tl(ctx, GSP_K_PLAN);
  if (nmax <= 8) launch_plan<8>(A, p->precision, blocks, s);
  else if (nmax <= 16) launch_plan<16>(A, p->precision, blocks, s);
  else if (nmax <= 24) launch_plan<24>(A, p->precision, blocks, s);
  else if (nmax <= 32) launch_plan<32>(A, p->precision, blocks, s);
  else if (nmax <= 48) launch_plan<48>(A, p->precision, blocks, s);
  else launch_plan<64>(A, p->precision, blocks, s);
  GSP_HIP(hipGetLastError());
  }
  {
    TimedLaunch tl(ctx, GSP_K_LEVELS);
    hipLaunchKernelGGL(k_levels, dim3(grid_for(N, 256, 2048)), dim3(256), 0, s, plan->path.p,
                       (int32_t)N, nmax, plan->nfound.p, plan->flags.p, plan->nbr.p,
                       plan->level.p, ctx->counters.p);
    GSP_HIP(hipGetLastError());
  }
  int32_t ctr[16];
  GSP_HIP(hipMemcpyAsync(ctr, ctx->counters.p, sizeof(ctr), hipMemcpyDeviceToHost, s));
  GSP_HIP(hipStreamSynchronize(s));
  GSP_REQUIRE(ctr[CTR_ERROR] == 0, GSP_ERR_INTERNAL,
              "dependency wait exceeded its limit while ordering the plan");
  flush_timings(ctx);
  plan->n_prior = ctr[CTR_PRIOR];
  plan->n_fallback = ctr[CTR_FALLBACK];
  plan->depth = ctr[CTR_DEPTH];
}

void seqsim_sweep_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                      const double* data_vals, double* out) {
  hipStream_t s = ctx->stream;
  const int64_t N = plan->N, nD = plan->n_data;
  const int RCmax = 64;
  const int RC0 = R < RCmax ? R : RCmax;
  ctx->X.reserve((size_t)((N + nD) * RC0));
  ctx->counters.reserve(16);
  auto* mplan = const_cast<gsp_seqsim_plan*>(plan);
  for (int r0 = 0; r0 < R; r0 += RCmax) {
    const int RC = (R - r0) < RCmax ? (R - r0) : RCmax;
    int shift = 0;
    while ((1 << shift) < RC) ++shift;
    mplan->epoch += 1;
    if (mplan->epoch == 0) {  // wrapped: clear the flags
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(N, 256, 4096)), dim3(256), 0, s,
                         mplan->ready.p, 0u, N);
      mplan->epoch = 1;
    }
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctx->counters.p, 0, (int64_t)16);
    if (nD > 0)
      hipLaunchKernelGGL(k_data_rows, dim3(grid_for(nD * RC, 256, 1024)), dim3(256), 0, s,
                         ctx->X.p, (int32_t)N, (int32_t)nD, RC, data_vals);
    SweepArgs A;
    A.path = plan->path.p;
    A.n_path = (int32_t)N;
    A.N = (int32_t)N;
    A.kmax = plan->kmax;
    A.RC = RC;
    A.shift = shift;
    A.r0 = r0;
    A.nfound = plan->nfound.p;
    A.flags = plan->flags.p;
    A.nbr = plan->nbr.p;
    A.lam = plan->lam.p;
    A.sigma = plan->sigma.p;
    A.mean = plan->mean;
    A.z = z;
    A.X = ctx->X.p;
    A.ready = mplan->ready.p;
    A.epoch = mplan->epoch;
    A.counters = ctx->counters.p;
    const int nodes_per_wave = 64 >> shift;
    const int64_t waves = (N + nodes_per_wave - 1) / nodes_per_wave;
    {
      TimedLaunch tl(ctx, GSP_K_SWEEP);
      hipLaunchKernelGGL(k_sweep, dim3(grid_for(waves, kWaves, 2048)), dim3(256), 0, s, A);
    }
    {
      TimedLaunch tl(ctx, GSP_K_TRANSPOSE);
      hipLaunchKernelGGL(k_transpose, dim3((unsigned)((N + 63) / 64)), dim3(256), 0, s, ctx->X.p,
                         (int32_t)N, RC, r0, out);
    }
    GSP_HIP(hipGetLastError());
  }
}

}  // namespace gsp
