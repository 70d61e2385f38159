// gsp_seqsim.hip — SEQSIM on gfx950.
//
// The reference (src/simulation/field/seqsim.jl:98-138) walks the path and, for
// every node of every realization, searches neighbours, assembles and factors a
// k x k simple-kriging system and draws one normal. The neighbour sets, weights
// and variances depend only on the path and the geometry, so this file splits
// the work into
//   k_search    one wavefront per node: neighbour search (offset table over the
//               implicit grid + a cell hash for the data; rings of hash cells
//               when no distance bound is known, as in the data-only search);
//   k_krige*    covariance assembly and a register-resident Cholesky solve of
//               the kriging system (four systems per wavefront up to 16
//               neighbours, one above);
//   k_dep_* / k_frontier_*  dependency level of every node (Kahn's algorithm
//               on the reverse edges), giving the nodes grouped by level;
//   k_sweep_level  the only sequential part: one launch per level, every node
//               of a level independent, R realizations across the lanes; the
//               kernel boundaries order the levels (on MI355X every XCD has its
//               own L2, so per-node acquire/release inside one kernel would
//               write back and invalidate L2 for every node);
//   k_z_transpose / k_transpose  realization-major <-> node-major layouts.
//
// A plan holds one or many independent requests (grids, variograms, data,
// paths, neighbourhoods of their own) in one global index space: request p
// owns node ids [node_off_p, node_off_p + N_p), path positions over the same
// range and data ids NT + [data_off_p, data_off_p + n_data_p), where NT is
// the number of nodes of all requests. Every kernel above runs once over the
// whole batch; the requests' levels run side by side in the sweep.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <atomic>
#include <string>
#include <thread>
#include <type_traits>

#include <hipcub/hipcub.hpp>

#include "gsp_seqsim.h"
#include "gsp_wave.h"
#include "gsp_krige.h"

namespace gsp {
namespace {

constexpr int kWaves = 4;  // wavefronts per 256-thread block

// Squared physical distance between a datum (coordinates x) and the centre
// of node t.
template <class RQ>
__device__ inline double data_d2(const RQ& R, const double* x, const int* t) {
  double acc = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= R.ndim) break;
    double c = R.org[d] + ((double)t[d] + 0.5) * R.sp[d];
    double u = x[d] - c;
    acc = acc + u * u;
  }
  return acc;
}

// Neighbour search, one wavefront per path position (global order): the first
// kmax usable grid offsets of the request's sorted offset table, the data of
// the node's hash list within the k-th grid distance, their merge by
// (d2, index) and the mask. Writes nbr[i][0..K) and nfound[i].
// RING: searches without a distance bound (data only, or masked with no
// ball) visit rings of hash cells; the other searches never lack a bound
// (the offset table always yields kmax grid candidates).
// MI355X dispatches consecutive workgroups round-robin to its 8 XCDs, each
// with its own L2.
constexpr int kXcds = 8;

template <bool RING>
__global__ __launch_bounds__(256, 8) void k_search(const PlanArgs A) {
  __shared__ int s_gidx[kWaves][64];
  __shared__ double s_gd2[kWaves][64];
  __shared__ int s_didx[kWaves][128];
  __shared__ double s_dd2[kWaves][128];
  __shared__ int s_fidx[kWaves][64];
  __shared__ uint8_t s_gt[kWaves][64];
  __shared__ uint8_t s_ft[kWaves][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t wstride = (int64_t)gridDim.x * kWaves;

  int cur = -1;
  Offset e0 = {0.0, 0, 0, 0, 0};
  // RING: every wave takes a contiguous run of nodes, and carries
  // the k-th data distance of its previous node: if that node lies delta
  // away, the k nearest data of this one lie within (D_prev + delta), a bound
  // that replaces the ring search
  // blocks of one XCD take consecutive runs of nodes (when the grid divides)
  const int G = (int)gridDim.x;
  const int64_t lb = G % kXcds == 0 ? (int64_t)(blockIdx.x % kXcds) * (G / kXcds) + blockIdx.x / kXcds
                                    : (int64_t)blockIdx.x;
  int64_t q_begin = lb * kWaves + w, q_end = A.NT, q_step = wstride;
  if constexpr (RING) {
    const int64_t chunk = (A.NT + wstride - 1) / wstride;
    q_begin = q_begin * chunk;
    q_end = q_begin + chunk < A.NT ? q_begin + chunk : A.NT;
    q_step = 1;
  }
  int prev_req = -1;
  double prev_k = -1.0;
  int pti[3] = {0, 0, 0};
  for (int64_t q = q_begin; q < q_end; q += q_step) {
    // nodes in grid order (every node's search is independent of the others):
    // consecutive wavefronts then test the positions of overlapping
    // neighbourhoods, which share cache lines
    const int gi = __builtin_amdgcn_readfirstlane((int)q);
    const int p = __builtin_amdgcn_readfirstlane(const_view(A.node_req)[gi]);
    // the request, read through the scalar cache (the stores below cannot alias it)
    const auto& R = const_view(A.reqs)[p];
    if (p != cur) {  // the first 64 offsets serve every node of the request
      cur = p;
      e0 = Offset{0.0, 0, 0, 0, 0};
      if (lane < R.T) e0 = A.table[R.table_off + lane];
    }
    const int pi = __builtin_amdgcn_readfirstlane(const_view(A.pos)[gi]);  // its path position
    const int nx = R.dims[0], ny = R.dims[1], nz = R.dims[2];
    const int kmax = R.kmax;
    const int i = gi - R.node_off;
    int ti[3];
    ti[0] = i % nx;
    ti[1] = (i / nx) % ny;
    ti[2] = i / (nx * ny);

    // ---- grid candidates: the first kmax usable offsets in (d2, index) order
    int cnt = 0;
    for (int t0 = 0; cnt < kmax && t0 < R.T; t0 += 64) {
      const int t = t0 + lane;
      bool ok = false;
      int j = 0;
      double d2 = 0.0;
      if (t < R.T) {
        const Offset e = t0 == 0 ? e0 : A.table[R.table_off + t];
        const int x = ti[0] + e.dx, y = ti[1] + e.dy, z = ti[2] + e.dz;
        ok = x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
        if (ok) {
          j = R.node_off + x + nx * (y + ny * z);
          d2 = e.d2;
          if (R.mode == GSP_MASKED_KNN && A.pos[j] >= pi) ok = false;
        }
      }
      const uint64_t b = __ballot(ok);
      const int rank = cnt + mbcnt(b);
      if (ok && rank < kmax) {
        s_gidx[w][rank] = j;
        s_gt[w][rank] = (uint8_t)(t < 255 ? t : 255);
        s_gd2[w][rank] = d2;
      }
      cnt += __popcll(b);
    }
    const int ng = cnt < kmax ? cnt : kmax;
    wave_sync();
    double Dk = ng == kmax ? s_gd2[w][ng - 1] : R.r2;
    if constexpr (RING) {
      if (!(Dk < __builtin_inf()) && p == prev_req && prev_k >= 0.0) {
        double dl = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const double u = (double)(ti[d] - pti[d]) * R.sp[d];
          dl = dl + u * u;
        }
        const double b = (__builtin_sqrt(prev_k) + __builtin_sqrt(dl)) * (1.0 + 0x1p-30);
        Dk = b * b;
      }
    }

    // ---- data candidates with d2 <= Dk, from the cell hash
    int nd = 0;
    if (R.n_data > 0) {
      double lim = Dk;  // data farther than this cannot enter the top kmax
      // keep the best kmax entries of the LDS list, each at its rank
      auto trim = [&]() {
        wave_sync();
        int keep_rank[2] = {128, 128};
        int kid[2] = {0, 0};
        double kd[2] = {0.0, 0.0};
#pragma unroll
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
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (keep_rank[h] < kmax) {
            s_dd2[w][keep_rank[h]] = kd[h];
            s_didx[w][keep_rank[h]] = kid[h];
          }
        nd = nd < kmax ? nd : kmax;
        wave_sync();
        if (nd == kmax) lim = s_dd2[w][kmax - 1];
      };
      // one 64-lane pass: item `idx` of the hash lists on each valid lane;
      // appends data with d2 <= lim to the LDS list
      auto consider = [&](bool valid, int idx) {
        bool ok = false;
        int di = 0;
        double d2 = 0.0;
        if (valid) {
          di = A.hitems[idx];
          d2 = data_d2(R, A.data + (size_t)di * 3, ti);
          ok = d2 <= lim && d2 <= R.r2;
        }
        const uint64_t b = __ballot(ok);
        const int add = __popcll(b);
        if (add == 0) return;
        if (nd + add > 128) trim();  // may lower lim
        const bool put = ok && d2 <= lim;
        const uint64_t pb = __ballot(put);
        if (put) {
          const int r = nd + mbcnt(pb);
          s_didx[w][r] = A.NT + di;
          s_dd2[w][r] = d2;
        }
        nd += __popcll(pb);
      };
      auto scan_list = [&](int s0, int e0l) {
        for (int b0 = s0; b0 < e0l; b0 += 64) consider(b0 + lane < e0l, b0 + lane);
      };
      const int32_t* hs = A.hstart + R.cell_off;
      const auto hsv = const_view(hs);  // wave-uniform list bounds: scalar loads
      // Searches over many cells gather item ranges (runs of adjacent cells
      // along x) in lanes, up to 64 at a time, and scan them as one flat list:
      // few, full 64-lane passes instead of one pass per cell.
      int rs = 0, rn = 0, nr = 0;  // this lane's range (first item, count); ranges held
      auto flush = [&]() {
        if (nr == 0) return;
        int pre = 0, total = 0;
        for (int j = 0; j < nr; ++j) {
          const int nj = rli(rn, j);
          if (lane == j) pre = total;
          total += nj;
        }
        for (int b0 = 0; b0 < total; b0 += 64) {
          const int t = b0 + lane;
          int jj = 0;
          for (int j = 1; j < nr; ++j) jj = t >= rli(pre, j) ? j : jj;
          consider(t < total, __shfl(rs, jj) + (t - __shfl(pre, jj)));
        }
        nr = 0;
      };
      auto add_range = [&](int a, int b) {
        if (b > a) {
          if (lane == nr) {
            rs = a;
            rn = b - a;
          }
          if (++nr == 64) flush();
        }
      };
      if (R.hexp) {
        const int cell = ti[0] / R.hB + R.hd[0] * (ti[1] / R.hB + R.hd[1] * (ti[2] / R.hB));
        scan_list(hs[cell], hs[cell + 1]);
      } else if (!RING || Dk < __builtin_inf()) {
        // every cell overlapping the ball of radius sqrt(Dk) (all cells
        // without a bound)
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        const bool all = !(Dk < __builtin_inf());
        const double rho = all ? 0.0 : __builtin_sqrt(Dk) * (1.0 + 0x1p-20) + 0x1p-40;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          if (d >= R.ndim) break;
          if (all) {
            lo[d] = 0;
            hi[d] = R.hd[d] - 1;
            continue;
          }
          const double c = R.org[d] + ((double)ti[d] + 0.5) * R.sp[d];
          double fl = __builtin_floor((c - rho - R.org[d]) / R.hcs[d]);
          double fh = __builtin_floor((c + rho - R.org[d]) / R.hcs[d]);
          fl = fl < 0.0 ? 0.0 : (fl > R.hd[d] - 1 ? R.hd[d] - 1 : fl);
          fh = fh < 0.0 ? 0.0 : (fh > R.hd[d] - 1 ? R.hd[d] - 1 : fh);
          lo[d] = (int)fl;
          hi[d] = (int)fh;
        }
        // the cells lo[0]..hi[0] of one (cy, cz) row hold one run of items:
        // one lane per row reads its run's bounds (a single round trip for up
        // to 64 rows) and flush() scans the runs as one list
        const int nyr = hi[1] - lo[1] + 1;
        const int nrows = nyr * (hi[2] - lo[2] + 1);
        for (int r0 = 0; r0 < nrows; r0 += 64) {
          const int rr = r0 + lane;
          if (rr < nrows) {
            const int row = R.hd[0] * (lo[1] + rr % nyr + R.hd[1] * (lo[2] + rr / nyr));
            rs = hs[row + lo[0]];
            rn = hs[row + hi[0] + 1] - rs;
          }
          nr = nrows - r0 < 64 ? nrows - r0 : 64;
          flush();
        }
      } else {
        // no bound yet: rings of cells at Chebyshev distance r = 0, 1, ... from
        // the node's cell, until the kmax-th candidate is closer than every
        // cell not yet visited (data outside the grid sit in the edge cells,
        // beyond their cell's inner face, so the bound holds for them too)
        int cc[3];
        double c[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          cc[d] = ti[d] / R.hB;
          c[d] = R.org[d] + ((double)ti[d] + 0.5) * R.sp[d];
        }
        for (int r = 0;; ++r) {
          int lo[3], hi[3];
          bool whole = true;
          double gap = __builtin_inf();
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            lo[d] = cc[d] - r > 0 ? cc[d] - r : 0;
            hi[d] = cc[d] + r < R.hd[d] - 1 ? cc[d] + r : R.hd[d] - 1;
            if (lo[d] > 0) {
              whole = false;
              const double g = c[d] - (R.org[d] + (double)lo[d] * R.hcs[d]) - 0x1p-30 * R.hcs[d];
              gap = g < gap ? g : gap;
            }
            if (hi[d] < R.hd[d] - 1) {
              whole = false;
              const double g =
                  R.org[d] + (double)(hi[d] + 1) * R.hcs[d] - c[d] - 0x1p-30 * R.hcs[d];
              gap = g < gap ? g : gap;
            }
          }
          for (int cz = lo[2]; cz <= hi[2]; ++cz)
            for (int cy = lo[1]; cy <= hi[1]; ++cy) {
              const int row = R.hd[0] * (cy + R.hd[1] * cz);
              if (cz - cc[2] == r || cc[2] - cz == r || cy - cc[1] == r || cc[1] - cy == r) {
                add_range(hsv[row + lo[0]], hsv[row + hi[0] + 1]);
              } else {
                if (cc[0] - r >= 0) add_range(hsv[row + cc[0] - r], hsv[row + cc[0] - r + 1]);
                if (r > 0 && cc[0] + r <= R.hd[0] - 1)
                  add_range(hsv[row + cc[0] + r], hsv[row + cc[0] + r + 1]);
              }
            }
          if (whole) {
            flush();
            break;
          }
          if (r == 0) continue;  // ring 0 joins ring 1's pass
          flush();
          if (nd >= kmax) {
            trim();
            if (gap > 0.0 && lim < gap * gap) break;
          }
        }
      }
      if constexpr (RING) {
        // the k-th data distance of this node, for the next one
        prev_k = -1.0;
        if (nd >= kmax) {
          trim();
          prev_k = lim;
        }
      }
      wave_sync();
    }
    if constexpr (RING) {
      prev_req = p;
      pti[0] = ti[0];
      pti[1] = ti[1];
      pti[2] = ti[2];
    }

    // ---- merge: top-kmax of the grid list (sorted) and the data list
    const int nf = ng + nd < kmax ? ng + nd : kmax;
    int my_rank[3] = {128, 128, 128};
    int my_id[3] = {0, 0, 0};
    uint8_t my_t[3] = {255, 255, 255};
    if (lane < ng) {
      const double d = s_gd2[w][lane];
      const int id = s_gidx[w][lane];
      int r = lane;
      for (int m = 0; m < nd; ++m) r += key_less(s_dd2[w][m], s_didx[w][m], d, id);
      my_rank[0] = r;
      my_id[0] = id;
      my_t[0] = s_gt[w][lane];
    }
#pragma unroll
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
#pragma unroll
    for (int h = 0; h < 3; ++h)
      if (my_rank[h] < kmax) {
        s_fidx[w][my_rank[h]] = my_id[h];
        s_ft[w][my_rank[h]] = my_t[h];
      }
    wave_sync();

    // ---- mask: drop points not yet simulated (k nearest first, then mask)
    const int fi = lane < nf ? s_fidx[w][lane] : -1;
    const uint8_t ft = lane < nf ? s_ft[w][lane] : (uint8_t)255;
    const bool keep = lane < nf && (R.mode == GSP_MASKED_KNN || fi >= A.NT || A.pos[fi] < pi);
    const uint64_t kb = __ballot(keep);
    const int n = __popcll(kb);
    const int kr = mbcnt(kb);
    wave_sync();
    if (keep) {
      s_fidx[w][kr] = fi;
      s_ft[w][kr] = ft;
    }
    wave_sync();
    const int nb = lane < n ? s_fidx[w][lane] : -1;

    if (lane < A.K) A.nbr[(int64_t)gi * A.K + lane] = nb;
    if (A.nbr_slot && lane < A.K)
      A.nbr_slot[(int64_t)gi * A.K + lane] = lane < n ? s_ft[w][lane] : (uint8_t)255;
    if (lane == 0) A.nfound[gi] = n;
    if (A.outdeg) {
      // a node with fewer than kmin neighbours depends on nothing; data
      // neighbours (ids >= NT) are no dependencies
      const bool dep = n >= R.kmin && lane < n && nb < A.NT;
      if (dep) atomicAdd(&A.outdeg[nb], 1);
      const int c = __popcll(__ballot(dep));
      if (lane == 0) {
        A.indeg[gi] = c;
        A.mark[gi] = c == 0 ? 1 : 0;
      }
    }
  }
}

// ---- grid-only search, one node per lane ----------------------------------
// The search of a plan of one request without data, bounded by its offset
// table: node i's candidates are the first kmax in-grid offsets of the table
// in (d2, index delta) order, of which the not yet simulated are then dropped
// (GSP_KNN_THEN_MASK), or among which only simulated nodes count
// (GSP_MASKED_KNN): k_search's lists, in its order. The 64 lanes of a
// wavefront take 64 consecutive nodes and walk the table in step (one scalar
// read per offset), so the position reads and the dependency counts of one
// offset fall on consecutive addresses; eight offsets' position reads are in
// flight together. The dependency counts take a second walk once the node's
// count (and so whether it has dependencies) is known.
// STAGE: each lane writes its row (and slots) to LDS, and the workgroup then
// stores its 256 rows as one contiguous run (dynamic LDS of 256 K ints and
// 256 K bytes).
template <bool STAGE>
__global__ __launch_bounds__(256) void k_search_grid_lane(const PlanArgs A) {
  extern __shared__ int s_dyn[];
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = g < A.NT;
  const int i = live ? (int)g : 0;
  const auto& R = const_view(A.reqs)[0];  // node_off 0
  const int nx = R.dims[0], ny = R.dims[1], nz = R.dims[2];
  const int ti[3] = {i % nx, (i / nx) % ny, i / (nx * ny)};
  const int pi = live ? A.pos[i] : 0;
  const int kmax = R.kmax;
  const bool masked = R.mode == GSP_MASKED_KNN;
  const Offset* tab = A.table + R.table_off;
  uint8_t* s_slot = reinterpret_cast<uint8_t*>(s_dyn + 256 * A.K);
  int* nb = STAGE ? s_dyn + threadIdx.x * A.K : A.nbr + (int64_t)i * A.K;
  uint8_t* ns = !A.nbr_slot ? nullptr
                            : (STAGE ? s_slot + threadIdx.x * A.K : A.nbr_slot + (int64_t)i * A.K);
  // eight offsets at a time: their in-grid targets and the targets' positions
  // (loads in flight together), then the decisions in table order
  auto block = [&](int t0, bool want, int* jj, int* pj) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u < R.T ? t0 + u : R.T - 1;
      const auto& e = const_view(tab)[t];
      const int x = ti[0] + e.dx, y = ti[1] + e.dy, z = ti[2] + e.dz;
      const bool in = want && t0 + u < R.T && x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 &&
                      z < nz;
      jj[u] = in ? x + nx * (y + ny * z) : -1;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) pj[u] = jj[u] >= 0 ? A.pos[jj[u]] : 0;
  };
  int cnt = live ? 0 : kmax;  // candidates taken
  int n = 0;                  // candidates kept
  for (int t0 = 0; t0 < R.T; t0 += 8) {
    if (!__any(cnt < kmax)) break;
    int jj[8], pj[8];
    block(t0, cnt < kmax, jj, pj);
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (jj[u] >= 0 && cnt < kmax) {
        const bool before = pj[u] < pi;
        if (!masked || before) ++cnt;
        if (before) {
          nb[n] = jj[u];
          if (ns) ns[n] = (uint8_t)(t0 + u < 255 ? t0 + u : 255);
          ++n;
        }
      }
  }
  if (live) {
    for (int m = n; m < A.K; ++m) {
      nb[m] = -1;
      if (ns) ns[m] = (uint8_t)255;
    }
    A.nfound[i] = n;
  }
  if constexpr (STAGE) {
    __syncthreads();
    const int64_t first = (int64_t)blockIdx.x * 256;
    const int64_t rem = (int64_t)A.NT - first;
    const int words = (int)(rem < 256 ? rem : 256) * A.K;
    int* dst = A.nbr + first * A.K;
    for (int q = threadIdx.x; q < words; q += 256) dst[q] = s_dyn[q];
    if (A.nbr_slot) {
      uint8_t* dsl = A.nbr_slot + first * A.K;
      for (int q = threadIdx.x; q < words; q += 256) dsl[q] = s_slot[q];
    }
  }
  if (A.outdeg) {
    // a node with fewer than kmin neighbours depends on nothing
    const bool dep = live && n >= R.kmin;
    int c = dep ? 0 : kmax;
    for (int t0 = 0; t0 < R.T; t0 += 8) {
      if (!__any(c < kmax)) break;
      int jj[8], pj[8];
      block(t0, c < kmax, jj, pj);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (jj[u] >= 0 && c < kmax) {
          const bool before = pj[u] < pi;
          if (!masked || before) ++c;
          if (before) atomicAdd(&A.outdeg[jj[u]], 1);
        }
    }
    if (live) {
      const int d = dep ? n : 0;
      A.indeg[i] = d;
      A.mark[i] = d == 0 ? 1 : 0;
    }
  }
}

// ---- data-only search, one node per lane ----------------------------------
// The data-only search (the k nearest data of every node, no grid
// neighbours, no dependencies) of plans whose requests all search data only,
// with box hash cells and kmax <= KT. A workgroup takes a tile of 256 nodes of
// one request (8 x 8 x 4 in 3D, 16 x 16 in 2D, 256 in 1D), stages the data of
// every hash cell within rho of the tile in LDS, and each lane keeps the KT
// best (d2, index) keys of its node in registers, sorted. A node's result is
// final when its kmax-th distance (the ball radius when it has fewer) lies
// strictly inside the cells visited: no datum outside them can be closer (data
// outside the grid sit in the edge cells, which every region reaching the edge
// holds). Otherwise the workgroup visits again with rho doubled, until every
// node is final or every cell has been visited, so the result is the exact
// top kmax under the order of k_search.
constexpr int kTileNodes = 256;
constexpr int kTileItems = 512;  // data staged in LDS per pass (two per thread)
constexpr int kShells = 4;       // distance shells of the staging order

__host__ __device__ inline void tile_shape(int ndim, int* t) {
  t[0] = ndim >= 3 ? 8 : (ndim == 2 ? 16 : 256);
  t[1] = ndim >= 3 ? 8 : (ndim == 2 ? 16 : 1);
  t[2] = ndim >= 3 ? 4 : 1;
}

template <int KT>
__global__ __launch_bounds__(kTileNodes) void k_search_tile(const PlanArgs A, int32_t n_req) {
  using Scan = hipcub::BlockScan<int, kTileNodes>;
  __shared__ typename Scan::TempStorage s_scan;
  __shared__ double s_x[kTileItems][3];
  __shared__ int s_id[kTileItems];
  __shared__ int s_rs[kTileNodes];       // first item of each row run
  __shared__ int s_pre[kTileNodes + 1];  // items before each row run
  __shared__ int s_cnt[kShells];
  const int tid = threadIdx.x;
  const int tile = blockIdx.x;
  // the request of this tile: the last one whose first tile is <= tile
  int p = 0;
  {
    int lo = 0, hi = n_req - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (const_view(A.reqs)[mid].tile_off <= tile) lo = mid;
      else hi = mid - 1;
    }
    p = lo;
  }
  const auto& R = const_view(A.reqs)[p];
  int ts[3];
  tile_shape(R.ndim, ts);
  const int ntx = (R.dims[0] + ts[0] - 1) / ts[0], nty = (R.dims[1] + ts[1] - 1) / ts[1];
  const int tl = tile - R.tile_off;
  const int t0[3] = {(tl % ntx) * ts[0], ((tl / ntx) % nty) * ts[1], (tl / (ntx * nty)) * ts[2]};
  // each wavefront a compact block of nodes (4 x 4 x 4 in 3D, 8 x 8 in 2D)
  const int w = tid >> 6, l = tid & 63;
  const int ti[3] = {
      t0[0] + (R.ndim >= 3 ? (w & 1) * 4 + (l & 3) : (R.ndim == 2 ? (w & 1) * 8 + (l & 7) : tid)),
      t0[1] + (R.ndim >= 3 ? (w >> 1) * 4 + ((l >> 2) & 3) : (R.ndim == 2 ? (w >> 1) * 8 + (l >> 3) : 0)),
      t0[2] + (R.ndim >= 3 ? l >> 4 : 0)};
  const bool live = ti[0] < R.dims[0] && ti[1] < R.dims[1] && ti[2] < R.dims[2];
  // the tile's centre
  double tc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int last = (t0[d] + ts[d] < R.dims[d] ? t0[d] + ts[d] : R.dims[d]) - 1;
    tc[d] = R.org[d] + (0.5 * (double)(t0[d] + last) + 0.5) * R.sp[d];
  }
  // the node's centre, as data_d2 forms it
  double c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] = R.org[d] + ((double)ti[d] + 0.5) * R.sp[d];
  const int kmax = R.kmax;
  const int32_t* hs = A.hstart + R.cell_off;

  double kd[KT];
  int ki[KT];
  double rho = R.trho;  // > 0
  for (int pass = 0;; ++pass, rho = rho + rho) {
#pragma unroll
    for (int j = 0; j < KT; ++j) {
      kd[j] = __builtin_inf();
      ki[j] = 0x7fffffff;
    }
    // the cells overlapping the tile's node centres widened by rho (every
    // cell after 40 doublings, which ends the loop)
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= R.ndim) break;
      if (pass >= 40) {
        hi[d] = R.hd[d] - 1;
        continue;
      }
      const int last = (t0[d] + ts[d] < R.dims[d] ? t0[d] + ts[d] : R.dims[d]) - 1;
      const double cl = R.org[d] + ((double)t0[d] + 0.5) * R.sp[d];
      const double ch = R.org[d] + ((double)last + 0.5) * R.sp[d];
      double fl = __builtin_floor((cl - rho - R.org[d]) / R.hcs[d]);
      double fh = __builtin_floor((ch + rho - R.org[d]) / R.hcs[d]);
      fl = fl < 0.0 ? 0.0 : (fl > R.hd[d] - 1 ? R.hd[d] - 1 : fl);
      fh = fh < 0.0 ? 0.0 : (fh > R.hd[d] - 1 ? R.hd[d] - 1 : fh);
      lo[d] = (int)fl;
      hi[d] = (int)fh;
    }
    // shells at 0.75, 1.25 and 1.75 rho from the tile's centre
    const double sh2[3] = {0.5625 * rho * rho, 1.5625 * rho * rho, 3.0625 * rho * rho};
    // the cells lo[0]..hi[0] of one (cy, cz) row hold one run of items
    const int nyr = hi[1] - lo[1] + 1;
    const int nrows = nyr * (hi[2] - lo[2] + 1);
    for (int r0 = 0; r0 < nrows; r0 += kTileNodes) {
      const int rr = r0 + tid;
      int len = 0;
      if (rr < nrows) {
        const int row = R.hd[0] * (lo[1] + rr % nyr + R.hd[1] * (lo[2] + rr / nyr));
        const int a = hs[row + lo[0]];
        s_rs[tid] = a;
        len = hs[row + hi[0] + 1] - a;
      }
      int pre = 0, total = 0;
      Scan(s_scan).ExclusiveSum(len, pre, total);
      s_pre[tid] = pre;
      const int nr = nrows - r0 < kTileNodes ? nrows - r0 : kTileNodes;
      for (int i0 = 0; i0 < total; i0 += kTileItems) {
        __syncthreads();  // s_pre / s_rs written, the previous items consumed
        const int n = total - i0 < kTileItems ? total - i0 : kTileItems;
        // the items go to LDS nearest shell first (shells of the distance
        // to the tile's centre): the lanes' lists then fill with near data
        // and most of the far data fail the first comparison
        if (tid < kShells) s_cnt[tid] = 0;
        __syncthreads();
        int sh[2], slot[2], idr[2];
        double xr[2][3];  // this thread's two items, until their slots are known
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int u = tid + h * kTileNodes;
          sh[h] = -1;
          if (u < n) {
            const int t = i0 + u;
            int a = 0, b = nr - 1;  // the run holding item t: last with s_pre <= t
            while (a < b) {
              const int m = (a + b + 1) >> 1;
              if (s_pre[m] <= t) a = m;
              else b = m - 1;
            }
            const int di = A.hitems[s_rs[a] + (t - s_pre[a])];
            const double* x = A.data + (size_t)di * 3;
            double q = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double xd = x[d];
              xr[h][d] = xd;
              const double v = d < R.ndim ? xd - tc[d] : 0.0;
              q = q + v * v;
            }
            idr[h] = A.NT + di;
            sh[h] = (q > sh2[0]) + (q > sh2[1]) + (q > sh2[2]);
            slot[h] = atomicAdd(&s_cnt[sh[h]], 1);
          }
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (sh[h] >= 0) {
            int o = slot[h];
            for (int b = 0; b < sh[h]; ++b) o += s_cnt[b];
            s_x[o][0] = xr[h][0];
            s_x[o][1] = xr[h][1];
            s_x[o][2] = xr[h][2];
            s_id[o] = idr[h];
          }
        __syncthreads();
        if (live) {
          for (int u = 0; u < n; ++u) {
            double d2 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              if (d >= R.ndim) break;
              const double w = s_x[u][d] - c[d];
              d2 = d2 + w * w;
            }
            const int id = s_id[u];
            if (d2 <= R.r2 && key_less(d2, id, kd[KT - 1], ki[KT - 1])) {
              // insert into the sorted list: slot j takes slot j-1's key
              // when the new key precedes that one
              static_for<0, KT - 1>([&](auto m) {
                constexpr int j = KT - 1 - decltype(m)::value;
                const bool up = key_less(d2, id, kd[j - 1], ki[j - 1]);
                const bool here = key_less(d2, id, kd[j], ki[j]);
                kd[j] = up ? kd[j - 1] : (here ? d2 : kd[j]);
                ki[j] = up ? ki[j - 1] : (here ? id : ki[j]);
              });
              if (key_less(d2, id, kd[0], ki[0])) {
                kd[0] = d2;
                ki[0] = id;
              }
            }
          }
        }
      }
      __syncthreads();  // before the next rows overwrite s_rs / s_pre
    }
    // final: the kmax-th key (or the ball) closer than every unvisited cell
    int nf = 0;
    double D = R.r2;
#pragma unroll
    for (int j = 0; j < KT; ++j) {
      nf += (j < kmax && ki[j] != 0x7fffffff) ? 1 : 0;
      if (j == kmax - 1 && ki[j] != 0x7fffffff) D = kd[j];
    }
    double gap = __builtin_inf();
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d >= R.ndim) break;
      if (lo[d] > 0) {
        const double g = c[d] - (R.org[d] + (double)lo[d] * R.hcs[d]) - 0x1p-30 * R.hcs[d];
        gap = g < gap ? g : gap;
      }
      if (hi[d] < R.hd[d] - 1) {
        const double g = R.org[d] + (double)(hi[d] + 1) * R.hcs[d] - c[d] - 0x1p-30 * R.hcs[d];
        gap = g < gap ? g : gap;
      }
    }
    const bool done = !live || !(gap < __builtin_inf()) || (gap > 0.0 && D < gap * gap);
    if (__syncthreads_and(done)) {
      if (live) {
        const int i = ti[0] + R.dims[0] * (ti[1] + R.dims[1] * ti[2]);
        const int64_t gi = (int64_t)R.node_off + i;
        int* nb = A.nbr + gi * A.K;
#pragma unroll
        for (int j = 0; j < KT; ++j)
          if (j < A.K) nb[j] = j < nf ? ki[j] : -1;
        for (int j = KT; j < A.K; ++j) nb[j] = -1;
        if (A.nbr_slot)
          for (int j = 0; j < A.K; ++j) A.nbr_slot[gi * A.K + j] = (uint8_t)255;
        A.nfound[gi] = nf;
      }
      return;
    }
  }
}

// ---- four systems per wavefront (n <= 16) ---------------------------------
// Lane 16 g + r holds row r of system g. Values are exchanged inside a 16-lane
// row with DPP row broadcasts. The systems of one wavefront are padded to the
// largest size among them with sill * identity rows, which leaves the
// factorisation and both substitutions of every real row bit-identical. The
// four systems may come from different requests; each group then reads its
// variogram from LDS.
template <int PREC>
struct Krige16Lds {
  double x[4][16][3];
  // the assembled matrix (read into registers before the FP32 factor is
  // stored), the FP32 factor, and in the FP64 fallback the FP64 factor: one
  // buffer, the phases separated by wave_sync
  union {
    double A[4][16][17];
    float F[PREC == GSP_FP32_REFINE ? 4 : 1][16][17];
  };
  DevModel M[4];
};

static_assert(sizeof(DevModel) == 16 * sizeof(double), "DevModel is copied as 16 words");

// FP64 Cholesky + substitutions of the packed systems (rows in a[]). Returns
// true in the lanes of a group whose factorisation met a non-positive pivot.
template <int PREC>
__device__ inline bool packed_fp64(Krige16Lds<PREC>& L, int g, int r, int nmax, double c, double C0,
                                   double* a, double& lam_out, double& sigma_out) {
  bool fail = false;
  double dinv = 0.0;  // 1 / L[r][r]
  static_for<0, 16>([&](auto colc) {
    constexpr int col = decltype(colc)::value;
    if (col < nmax) {
      const double piv = rbd<col>(a[col]);
      fail = fail || !(piv > 0.0);
      const double lcc = __builtin_sqrt(piv);
      const double rinv = 1.0 / lcc;
      if (r == col) dinv = rinv;
      a[col] = r == col ? lcc : (r > col ? a[col] * rinv : a[col]);
      static_for<col + 1, 16>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if (m < nmax) {
          const double lmc = rbd<m>(a[col]);
          a[m] = a[m] - a[col] * lmc;
        }
      });
    }
  });
  double s = c, y = 0.0;
  static_for<0, 16>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < nmax) {
      const double ym = rbd<m>(s * dinv);
      if (r == m) y = ym;
      if (r > m) s = s - a[m] * ym;
    }
  });
  wave_sync();
  static_for<0, 16>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= r) L.A[g][r][m] = a[m];
  });
  wave_sync();
  double s2 = y, lam = 0.0;
  static_for_down<16>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < nmax) {
      const double lm = rbd<m>(s2 * dinv);
      if (r == m) lam = lm;
      if (r < m) s2 = s2 - L.A[g][m][r] * lm;
    }
  });
  double acc = 0.0;
  static_for<0, 16>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < nmax) acc = acc + rbd<j>(lam) * rbd<j>(c);
  });
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  lam_out = lam;
  sigma_out = __builtin_sqrt(var);
  return fail;
}

// One wavefront per block: its 15 KB of LDS then packs ten wavefronts per CU,
// where blocks of four fit only twice (the kernel is latency-bound, so the
// resident wavefronts set its rate).
constexpr int k16Waves = 1;

template <int PREC>
__global__ __launch_bounds__(64 * k16Waves) void k_krige16(const PlanArgs A, const int32_t* list,
                                                         int32_t count) {
  __builtin_amdgcn_s_setprio(3);  // issue ahead of the level kernels beside the solves
  __shared__ Krige16Lds<PREC> s_lds[k16Waves];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 4, r = lane & 15;
  Krige16Lds<PREC>& L = s_lds[w];
  const int64_t stride = (int64_t)gridDim.x * k16Waves * 4;
  for (int64_t base = ((int64_t)blockIdx.x * k16Waves + w) * 4; base < count; base += stride) {
    const int64_t slot = base + g;
    const bool act = slot < count;
    const int gi = act ? list[slot] : list[base];
    const int p = A.node_req[gi];
    const Req& R = A.reqs[p];
    const int p0 = rli(p, 0);
    const bool uniform = __ballot(p != p0) == 0;  // all four systems of one request
    const double C0 = R.M.sill;
    const int n = act ? A.nfound[gi] : 0;
    const int nb = (act && r < n) ? A.nbr[(int64_t)gi * A.K + r] : -1;
    const bool solve = act && n >= R.kmin;
    const int ns = solve ? n : 0;
    int nmax = rli(ns, 0);
    nmax = rli(ns, 16) > nmax ? rli(ns, 16) : nmax;
    nmax = rli(ns, 32) > nmax ? rli(ns, 32) : nmax;
    nmax = rli(ns, 48) > nmax ? rli(ns, 48) : nmax;
    double lam = 0.0, sig = R.prior_sigma;
    int flag = act && !solve ? GSP_FLAG_FEW : 0;
    if (nmax > 0) {
      double x0 = 0.0, x1 = 0.0, x2 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
      if (r < ns) {
        neighbour_coords(A, R, nb, x0, x1, x2);
        L.x[g][r][0] = x0;
        L.x[g][r][1] = x1;
        L.x[g][r][2] = x2;
      }
      if (solve) node_coords(R, gi - R.node_off, t0, t1, t2);
      if (!uniform) ((double*)&L.M[g])[r] = ((const double*)&R.M)[r];
      wave_sync();
      // covariance entries of each system's lower triangle over its 16 lanes
      const int T = ns * (ns + 1) / 2, Tmax = nmax * (nmax + 1) / 2;
      auto assemble16 = [&](const DevModel& M) {
        for (int e = r; e < Tmax; e += 16) {
          if (e < T) {
            int j, m;
            tri_unpack(e, j, m);
            const double v = cov_points(M, L.x[g][j][0], L.x[g][j][1], L.x[g][j][2],
                                        L.x[g][m][0], L.x[g][m][1], L.x[g][m][2]);
            L.A[g][j][m] = v;
            L.A[g][m][j] = v;
          }
        }
        return r < ns ? cov_points(M, x0, x1, x2, t0, t1, t2) : 0.0;
      };
      const double c = uniform ? assemble16(A.reqs[p0].M) : assemble16(L.M[g]);
      wave_sync();
      double a[16];
      static_for<0, 16>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        a[m] = (r < ns && m < ns) ? L.A[g][r][m] : (m == r ? C0 : 0.0);
      });
      bool fail;
      bool fell_back = false;
      if constexpr (PREC == GSP_FP32_REFINE) {
        const float tau = (float)(C0 * 0x1p-12);
        float f[16];
        static_for<0, 16>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          f[m] = (float)a[m];
        });
        bool f32fail = false;
        float dinv = 0.0f;  // 1 / L[r][r]
        static_for<0, 16>([&](auto colc) {
          constexpr int col = decltype(colc)::value;
          if (col < nmax) {
            const float piv = rbf<col>(f[col]);
            f32fail = f32fail || !(piv > tau);
            const float lcc = __builtin_sqrtf(piv);
            const float rinv = 1.0f / lcc;
            if (r == col) dinv = rinv;
            f[col] = r == col ? lcc : (r > col ? f[col] * rinv : f[col]);
            static_for<col + 1, 16>([&](auto mc) {
              constexpr int m = decltype(mc)::value;
              if (m < nmax) {
                const float lmc = rbf<m>(f[col]);
                f[m] = f[m] - f[col] * lmc;
              }
            });
          }
        });
        wave_sync();
        static_for<0, 16>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          if (m <= r) L.F[g][r][m] = f[m];
        });
        wave_sync();
        float rhs = (float)c, sol = 0.0f, l = 0.0f;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          float s = rhs, y = 0.0f;
          static_for<0, 16>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (m < nmax) {
              const float ym = rbf<m>(s * dinv);
              if (r == m) y = ym;
              if (r > m) s = s - f[m] * ym;
            }
          });
          float s2 = y;
          l = 0.0f;
          static_for_down<16>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (m < nmax) {
              const float lm = rbf<m>(s2 * dinv);
              if (r == m) l = lm;
              if (r < m) s2 = s2 - L.F[g][m][r] * lm;
            }
          });
          if (pass == 0) {
            sol = l;
            double res = c;
            static_for<0, 16>([&](auto mc) {
              constexpr int m = decltype(mc)::value;
              if (m < nmax) res = res - a[m] * (double)rbf<m>(l);
            });
            rhs = (float)res;
          }
        }
        lam = (double)sol + (double)l;
        double acc = 0.0;
        static_for<0, 16>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < nmax) acc = acc + rbd<j>(lam) * rbd<j>(c);
        });
        double var = C0 - acc;
        if (var < 0.0) var = 0.0;
        sig = __builtin_sqrt(var);
        fail = false;
        // systems whose FP32 factor failed are solved again in FP64
        if (__ballot(f32fail && r == 0)) {
          double lam64, sig64;
          wave_sync();
          const bool fail64 = packed_fp64<PREC>(L, g, r, nmax, c, C0, a, lam64, sig64);
          if (f32fail) {
            lam = lam64;
            sig = sig64;
            fail = fail64;
            fell_back = true;
          }
        }
      } else {
        fail = packed_fp64<PREC>(L, g, r, nmax, c, C0, a, lam, sig);
      }
      wave_sync();
      if (solve) {
        if (fell_back) flag |= GSP_FLAG_FP64_FALLBACK;
        if (fail) {
          flag |= GSP_FLAG_FAIL;
          lam = 0.0;
          sig = R.prior_sigma;
        }
      } else {
        lam = 0.0;
        sig = R.prior_sigma;
      }
    }
    if (act) {
      for (int col = r; col < A.K; col += 16) A.lam[(int64_t)gi * A.K + col] = col < n ? lam : 0.0;
    }
    if (act && r == 0) {
      A.sigma[gi] = sig;
      A.flags[gi] = (uint8_t)flag;
      if (flag & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) atomicAdd(&A.counters[CTR_PRIOR], 1);
      if (flag & GSP_FLAG_FAIL) atomicAdd(&A.counters[CTR_FAIL], 1);
      if (flag & GSP_FLAG_FP64_FALLBACK) atomicAdd(&A.counters[CTR_FALLBACK], 1);
    }
  }
}

// ---- one system per lane (n <= NB) ------------------------------------------
// Small systems need no exchange between lanes: each lane assembles, factors
// and solves its own system in registers (lower triangles packed row by row),
// with the operation order of k_krige16 restricted to the system's own rows
// (the padded rows of the packed kernel leave the real rows' arithmetic
// unchanged), so the results are bitwise those of k_krige16 and the oracle.
// k_krige16 spends most of its instructions moving operands between lanes and
// half its lanes idle on systems of at most 8 neighbours; here every
// instruction works on 64 systems.
template <int NB>
__device__ __forceinline__ constexpr int tri(int j, int m) {
  return j * (j + 1) / 2 + m;
}

// FP64 Cholesky and both substitutions on a[] (in place; packed_fp64's order).
template <int NB>
__device__ inline bool lane_fp64(double (&a)[NB * (NB + 1) / 2], const double (&c)[NB], int n,
                                 double C0, double (&lam)[NB], double& sig) {
  bool fail = false;
  double dinv[NB];
  static_for<0, NB>([&](auto colc) {
    constexpr int col = decltype(colc)::value;
    if (col < n) {
      const double piv = a[tri<NB>(col, col)];
      fail = fail || !(piv > 0.0);
      const double lcc = __builtin_sqrt(piv);
      const double rinv = 1.0 / lcc;
      dinv[col] = rinv;
      a[tri<NB>(col, col)] = lcc;
      static_for<col + 1, NB>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if (r < n) a[tri<NB>(r, col)] = a[tri<NB>(r, col)] * rinv;
      });
      static_for<col + 1, NB>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if (m < n) {
          const double lmc = a[tri<NB>(m, col)];
          static_for<m, NB>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if (r < n) a[tri<NB>(r, m)] = a[tri<NB>(r, m)] - a[tri<NB>(r, col)] * lmc;
          });
        }
      });
    }
  });
  double s[NB], y[NB];
  static_for<0, NB>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    s[r] = c[r];
    y[r] = 0.0;
  });
  static_for<0, NB>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < n) {
      y[m] = s[m] * dinv[m];
      static_for<m + 1, NB>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if (r < n) s[r] = s[r] - a[tri<NB>(r, m)] * y[m];
      });
    }
  });
  static_for<0, NB>([&](auto rc) {
    constexpr int r = decltype(rc)::value;
    lam[r] = 0.0;
  });
  static_for<0, NB>([&](auto kc) {
    constexpr int m = NB - 1 - decltype(kc)::value;
    if (m < n) {
      lam[m] = y[m] * dinv[m];
      static_for<0, m>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        y[r] = y[r] - a[tri<NB>(m, r)] * lam[m];
      });
    }
  });
  double acc = 0.0;
  static_for<0, NB>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < n) acc = acc + lam[j] * c[j];
  });
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  sig = __builtin_sqrt(var);
  return fail;
}

// Largest system solved one per lane, by precision. FP32 + refinement: 8
// (184 VGPRs); 10 fits in 252 without spilling but, holding more of the
// register file, slowed the level kernels beside it (batch of 512: step 78.6
// -> 80.4 ms). FP64: 10 (227 VGPRs). 12 spills into AGPRs in both.
constexpr int kLaneNB32 = 8;
constexpr int kLaneNB64 = 10;
// Fewest such systems for which the lane kernel runs (GSP_LANE_MIN
// overrides): each lane works through a whole system serially, so a launch of
// few waves waits on one long chain (a 10,000-node plan measured 51 us
// against 23 us in k_krige16).
constexpr int64_t kLaneMin = 32768;

template <int NB, int PREC>
__global__ __launch_bounds__(256) void k_krige_lane(const PlanArgs A, const int32_t* list,
                                                    int32_t count) {
  __builtin_amdgcn_s_setprio(3);  // issue ahead of the level kernels beside the solves
  constexpr int T = NB * (NB + 1) / 2;
  const int lane = threadIdx.x & 63;
  for (int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;; slot += (int64_t)gridDim.x * blockDim.x) {
    // wave-uniform exit: every lane of a wave leaves together
    if (!__ballot(slot < count)) break;
    const bool act = slot < count;
    const int gi = act ? list[slot] : 0;
    const int p = act ? A.node_req[gi] : 0;
    const Req& R = A.reqs[p];
    const int n = act ? A.nfound[gi] : 0;
    const bool solve = act && n >= R.kmin;
    const double C0 = R.M.sill;
    double lam[NB];
    static_for<0, NB>([&](auto jc) { lam[decltype(jc)::value] = 0.0; });
    double sig = R.prior_sigma;
    int flag = act && !solve ? GSP_FLAG_FEW : 0;
    bool fail = false, fell_back = false;
    if (solve && n > 0) {
      double x[NB][3], c[NB], a[T];
      static_for<0, NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        x[j][0] = x[j][1] = x[j][2] = 0.0;
        if (j < n)
          neighbour_coords(A, R, A.nbr[(int64_t)gi * A.K + j], x[j][0], x[j][1], x[j][2]);
      });
      double t0, t1, t2;
      node_coords(R, gi - R.node_off, t0, t1, t2);
      static_for<0, NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        c[j] = j < n ? cov_points(R.M, x[j][0], x[j][1], x[j][2], t0, t1, t2) : 0.0;
        static_for<0, j + 1>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          a[tri<NB>(j, m)] = j < n ? cov_points(R.M, x[j][0], x[j][1], x[j][2], x[m][0], x[m][1],
                                                 x[m][2])
                                   : 0.0;
        });
      });
      if constexpr (PREC == GSP_FP32_REFINE) {
        const float tau = (float)(C0 * 0x1p-12);
        float f[T], dinv[NB];
        static_for<0, T>([&](auto ec) { f[decltype(ec)::value] = (float)a[decltype(ec)::value]; });
        bool f32fail = false;
        static_for<0, NB>([&](auto colc) {
          constexpr int col = decltype(colc)::value;
          if (col < n) {
            const float piv = f[tri<NB>(col, col)];
            f32fail = f32fail || !(piv > tau);
            const float lcc = __builtin_sqrtf(piv);
            const float rinv = 1.0f / lcc;
            dinv[col] = rinv;
            f[tri<NB>(col, col)] = lcc;
            static_for<col + 1, NB>([&](auto rc) {
              constexpr int r = decltype(rc)::value;
              if (r < n) f[tri<NB>(r, col)] = f[tri<NB>(r, col)] * rinv;
            });
            static_for<col + 1, NB>([&](auto mc) {
              constexpr int m = decltype(mc)::value;
              if (m < n) {
                const float lmc = f[tri<NB>(m, col)];
                static_for<m, NB>([&](auto rc) {
                  constexpr int r = decltype(rc)::value;
                  if (r < n) f[tri<NB>(r, m)] = f[tri<NB>(r, m)] - f[tri<NB>(r, col)] * lmc;
                });
              }
            });
          }
        });
        float rhs[NB], sol[NB], l[NB];
        static_for<0, NB>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          rhs[r] = (float)c[r];
          sol[r] = 0.0f;
        });
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          float s[NB], y[NB];
          static_for<0, NB>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            s[r] = rhs[r];
            y[r] = 0.0f;
            l[r] = 0.0f;
          });
          static_for<0, NB>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (m < n) {
              y[m] = s[m] * dinv[m];
              static_for<m + 1, NB>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                if (r < n) s[r] = s[r] - f[tri<NB>(r, m)] * y[m];
              });
            }
          });
          static_for<0, NB>([&](auto kc) {
            constexpr int m = NB - 1 - decltype(kc)::value;
            if (m < n) {
              l[m] = y[m] * dinv[m];
              static_for<0, m>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                y[r] = y[r] - f[tri<NB>(m, r)] * l[m];
              });
            }
          });
          if (pass == 0) {
            // residual of the FP32 solution against the FP64 matrix (full rows)
            static_for<0, NB>([&](auto rc) {
              constexpr int r = decltype(rc)::value;
              sol[r] = l[r];
              double res = c[r];
              static_for<0, NB>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                if (m < n)
                  res = res - a[m <= r ? tri<NB>(r, m) : tri<NB>(m, r)] * (double)l[m];
              });
              rhs[r] = (float)res;
            });
          }
        }
        static_for<0, NB>([&](auto rc) {
          constexpr int r = decltype(rc)::value;
          lam[r] = r < n ? (double)sol[r] + (double)l[r] : 0.0;
        });
        double acc = 0.0;
        static_for<0, NB>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < n) acc = acc + lam[j] * c[j];
        });
        double var = C0 - acc;
        if (var < 0.0) var = 0.0;
        sig = __builtin_sqrt(var);
        // systems whose FP32 factor failed are solved again in FP64
        if (f32fail) {
          fail = lane_fp64<NB>(a, c, n, C0, lam, sig);
          fell_back = true;
        }
      } else {
        fail = lane_fp64<NB>(a, c, n, C0, lam, sig);
      }
    }
    if (solve) {
      if (fell_back) flag |= GSP_FLAG_FP64_FALLBACK;
      if (fail) {
        flag |= GSP_FLAG_FAIL;
        sig = R.prior_sigma;
      }
    } else {
      sig = R.prior_sigma;
    }
    if (act) {
      double* lr = A.lam + (int64_t)gi * A.K;
      static_for<0, NB>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if (j < A.K) lr[j] = (j < n && solve && !fail) ? lam[j] : 0.0;
      });
      for (int j = NB; j < A.K; ++j) lr[j] = 0.0;
      A.sigma[gi] = sig;
      A.flags[gi] = (uint8_t)flag;
    }
    // counters: one atomic per wavefront and counter
    const int prior = __popcll(__ballot(act && (flag & (GSP_FLAG_FEW | GSP_FLAG_FAIL))));
    const int failed = __popcll(__ballot(act && (flag & GSP_FLAG_FAIL)));
    const int fb = __popcll(__ballot(act && (flag & GSP_FLAG_FP64_FALLBACK)));
    if (lane == 0) {
      if (prior) atomicAdd(&A.counters[CTR_PRIOR], prior);
      if (failed) atomicAdd(&A.counters[CTR_FAIL], failed);
      if (fb) atomicAdd(&A.counters[CTR_FALLBACK], fb);
    }
  }
}

// ---- two systems per wavefront (16 < n <= KB, KB = 24 or 32) ---------------
// Lane 32 g + r holds row r of system g (rows r < KB take part). The Cholesky step of column col
// scales the column in every row's lane, stores it to LDS and every lane of
// the half reads the entries it needs back as LDS broadcasts: one LDS read and
// the update per matrix element, where the one-system kernel spends two
// readlanes and their wait states. Substitution values are made by the lane
// that owns them and broadcast the same way. Padding to the larger of the two
// sizes with sill * identity rows and the operation order are those of
// k_krige16 and the one-system kernels (bitwise equal results).
// Sized by the bucket's largest system KB (LDS limits the resident
// wavefronts); rows r >= KB of a half hold no system row.
template <int KB, int PREC>
struct Krige32Lds {
  double x[2][KB][3];
  union {
    double A[2][KB][KB + 1];  // assembled matrix, later the FP64 factor
    float F[2][KB][KB + 1];   // FP32 factor (the FP64 rows stay in registers)
  };
  double col[2][32];      // broadcast slots
  float colf[2][32];
  DevModel M[2];
};

// Value of x in lane 32 g + m of this lane's half (m wave-uniform).
__device__ inline double half_lane(double x, int m, int g) {
  const double v0 = rl(x, m), v1 = rl(x, 32 + m);
  return g ? v1 : v0;
}

__device__ inline float half_lane_f(float x, int m, int g) {
  const float v0 = rlf(x, m), v1 = rlf(x, 32 + m);
  return g ? v1 : v0;
}

// FP64 Cholesky and substitutions of the two systems (rows in a[]); true in
// the lanes of a system whose factorisation met a non-positive pivot.
template <int KB, int PREC>
__device__ inline bool packed32_fp64(Krige32Lds<KB, PREC>& L, int g, int r, int nmax, double c,
                                     double C0, double* a, double& lam_out, double& sigma_out) {
  bool fail = false;
  double dinv = 0.0;  // 1 / L[r][r]
  static_for<0, KB>([&](auto colc) {
    constexpr int col = decltype(colc)::value;
    if (col < nmax) {
      const double piv = half_lane(a[col], col, g);
      fail = fail || !(piv > 0.0);
      const double lcc = __builtin_sqrt(piv);
      const double rinv = 1.0 / lcc;
      if (r == col) dinv = rinv;
      a[col] = r == col ? lcc : (r > col ? a[col] * rinv : a[col]);
      L.col[g][r] = a[col];
      wave_sync();
      static_for<col + 1, KB>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        if (m < nmax) {
          const double lmc = L.col[g][m];
          a[m] = a[m] - a[col] * lmc;
        }
      });
      wave_sync();
    }
  });
  // forward: y_m = s_m / L_mm, s_r -= L_rm y_m below it
  double s = c, y = 0.0;
  static_for<0, KB>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < nmax) {
      const double ym = half_lane(s * dinv, m, g);
      if (r == m) y = ym;
      if (r > m) s = s - a[m] * ym;
    }
  });
  static_for<0, KB>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m <= r && r < KB) L.A[g][r][m] = a[m];
  });
  wave_sync();
  // back: lambda_m = s2_m / L_mm, s2_r -= L_mr lambda_m above it
  double s2 = y, lam = 0.0;
  static_for_down<KB>([&](auto mc) {
    constexpr int m = decltype(mc)::value;
    if (m < nmax) {
      const double lm = half_lane(s2 * dinv, m, g);
      if (r == m) lam = lm;
      if (r < m) s2 = s2 - L.A[g][m][r < KB ? r : 0] * lm;
    }
  });
  // sigma^2 = C0 - sum_j lambda_j c_j in ascending j
  double acc = 0.0;
  static_for<0, KB>([&](auto jc) {
    constexpr int j = decltype(jc)::value;
    if (j < nmax) acc = acc + half_lane(lam, j, g) * half_lane(c, j, g);
  });
  double var = C0 - acc;
  if (var < 0.0) var = 0.0;
  lam_out = lam;
  sigma_out = __builtin_sqrt(var);
  return fail;
}

constexpr int k32Waves = 1;

template <int KB, int PREC>
__global__ __launch_bounds__(64 * k32Waves) void k_krige32(const PlanArgs A, const int32_t* list,
                                                         int32_t count) {
  __builtin_amdgcn_s_setprio(3);  // issue ahead of the level kernels beside the solves
  __shared__ Krige32Lds<KB, PREC> s_lds[k32Waves];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 5, r = lane & 31;
  Krige32Lds<KB, PREC>& L = s_lds[w];
  const int64_t stride = (int64_t)gridDim.x * k32Waves * 2;
  for (int64_t base = ((int64_t)blockIdx.x * k32Waves + w) * 2; base < count; base += stride) {
    const int64_t slot = base + g;
    const bool act = slot < count;
    const int gi = act ? list[slot] : list[base];
    const int p = A.node_req[gi];
    const Req& R = A.reqs[p];
    const int p0 = rli(p, 0);
    const bool uniform = __ballot(p != p0) == 0;  // both systems of one request
    const double C0 = R.M.sill;
    const int n = act ? A.nfound[gi] : 0;
    const int nb = (act && r < n) ? A.nbr[(int64_t)gi * A.K + r] : -1;
    const bool solve = act && n >= R.kmin;
    const int ns = solve ? n : 0;
    const int nmax = rli(ns, 0) > rli(ns, 32) ? rli(ns, 0) : rli(ns, 32);
    double lam = 0.0, sig = R.prior_sigma;
    int flag = act && !solve ? GSP_FLAG_FEW : 0;
    if (nmax > 0) {
      double x0 = 0.0, x1 = 0.0, x2 = 0.0, t0 = 0.0, t1 = 0.0, t2 = 0.0;
      if (r < ns) {
        neighbour_coords(A, R, nb, x0, x1, x2);
        L.x[g][r][0] = x0;
        L.x[g][r][1] = x1;
        L.x[g][r][2] = x2;
      }
      if (solve) node_coords(R, gi - R.node_off, t0, t1, t2);
      if (!uniform && r < 16) ((double*)&L.M[g])[r] = ((const double*)&R.M)[r];
      wave_sync();
      // covariance entries of each system's lower triangle over its 32 lanes
      const int T = ns * (ns + 1) / 2, Tmax = nmax * (nmax + 1) / 2;
      auto assemble32 = [&](const DevModel& M) {
        for (int e = r; e < Tmax; e += 32) {
          if (e < T) {
            int j, m;
            tri_unpack(e, j, m);
            const double v = cov_points(M, L.x[g][j][0], L.x[g][j][1], L.x[g][j][2],
                                        L.x[g][m][0], L.x[g][m][1], L.x[g][m][2]);
            L.A[g][j][m] = v;
            L.A[g][m][j] = v;
          }
        }
        return r < ns ? cov_points(M, x0, x1, x2, t0, t1, t2) : 0.0;
      };
      const double c = uniform ? assemble32(A.reqs[p0].M) : assemble32(L.M[g]);
      wave_sync();
      double a[KB];
      static_for<0, KB>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        a[m] = (r < ns && m < ns) ? L.A[g][r][m] : (m == r ? C0 : 0.0);
      });
      wave_sync();
      bool fail;
      bool fell_back = false;
      if constexpr (PREC == GSP_FP32_REFINE) {
        const float tau = (float)(C0 * 0x1p-12);
        float f[KB];
        static_for<0, KB>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          f[m] = (float)a[m];
        });
        bool f32fail = false;
        float dinv = 0.0f;  // 1 / L[r][r]
        static_for<0, KB>([&](auto colc) {
          constexpr int col = decltype(colc)::value;
          if (col < nmax) {
            const float piv = __builtin_bit_cast(
                float, g ? __builtin_amdgcn_readlane(__builtin_bit_cast(int, f[col]), 32 + col)
                         : __builtin_amdgcn_readlane(__builtin_bit_cast(int, f[col]), col));
            f32fail = f32fail || !(piv > tau);
            const float lcc = __builtin_sqrtf(piv);
            const float rinv = 1.0f / lcc;
            if (r == col) dinv = rinv;
            f[col] = r == col ? lcc : (r > col ? f[col] * rinv : f[col]);
            L.colf[g][r] = f[col];
            wave_sync();
            static_for<col + 1, KB>([&](auto mc) {
              constexpr int m = decltype(mc)::value;
              if (m < nmax) {
                const float lmc = L.colf[g][m];
                f[m] = f[m] - f[col] * lmc;
              }
            });
            wave_sync();
          }
        });
        static_for<0, KB>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          if (m <= r && r < KB) L.F[g][r][m] = f[m];
        });
        wave_sync();
        float rhs = (float)c, sol = 0.0f, l = 0.0f;
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
          float sv = rhs, y = 0.0f;
          static_for<0, KB>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (m < nmax) {
              const float ym = half_lane_f(sv * dinv, m, g);
              if (r == m) y = ym;
              if (r > m) sv = sv - f[m] * ym;
            }
          });
          float s2 = y;
          l = 0.0f;
          static_for_down<KB>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if (m < nmax) {
              const float lm = half_lane_f(s2 * dinv, m, g);
              if (r == m) l = lm;
              if (r < m) s2 = s2 - L.F[g][m][r < KB ? r : 0] * lm;
            }
          });
          wave_sync();
          if (pass == 0) {
            sol = l;
            L.colf[g][r] = l;
            wave_sync();
            double res = c;
            static_for<0, KB>([&](auto mc) {
              constexpr int m = decltype(mc)::value;
              if (m < nmax) res = res - a[m] * (double)L.colf[g][m];
            });
            rhs = (float)res;
            wave_sync();
          }
        }
        lam = (double)sol + (double)l;
        L.col[g][r] = lam;
        wave_sync();
        double acc = 0.0;
        static_for<0, KB>([&](auto jc) {
          constexpr int j = decltype(jc)::value;
          if (j < nmax) acc = acc + L.col[g][j] * half_lane(c, j, g);
        });
        double var = C0 - acc;
        if (var < 0.0) var = 0.0;
        sig = __builtin_sqrt(var);
        fail = false;
        wave_sync();
        // systems whose FP32 factor failed are solved again in FP64
        if (__ballot(f32fail && r == 0)) {
          double lam64, sig64;
          wave_sync();
          const bool fail64 = packed32_fp64<KB, PREC>(L, g, r, nmax, c, C0, a, lam64, sig64);
          if (f32fail) {
            lam = lam64;
            sig = sig64;
            fail = fail64;
            fell_back = true;
          }
        }
      } else {
        fail = packed32_fp64<KB, PREC>(L, g, r, nmax, c, C0, a, lam, sig);
      }
      wave_sync();
      if (solve) {
        if (fell_back) flag |= GSP_FLAG_FP64_FALLBACK;
        if (fail) {
          flag |= GSP_FLAG_FAIL;
          lam = 0.0;
          sig = R.prior_sigma;
        }
      } else {
        lam = 0.0;
        sig = R.prior_sigma;
      }
    }
    if (act) {
      for (int col = r; col < A.K; col += 32) A.lam[(int64_t)gi * A.K + col] = col < n ? lam : 0.0;
    }
    if (act && r == 0) {
      A.sigma[gi] = sig;
      A.flags[gi] = (uint8_t)flag;
      if (flag & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) atomicAdd(&A.counters[CTR_PRIOR], 1);
      if (flag & GSP_FLAG_FAIL) atomicAdd(&A.counters[CTR_FAIL], 1);
      if (flag & GSP_FLAG_FP64_FALLBACK) atomicAdd(&A.counters[CTR_FALLBACK], 1);
    }
    wave_sync();
  }
}

// Nodes (ids[0..N), or 0..N-1 without ids) bucketed by neighbour count
// (counting sort), so packed wavefronts hold systems of one size.
__global__ __launch_bounds__(256) void k_n_hist(const int32_t* nfound, const int32_t* ids,
                                                int32_t N, int32_t* hist) {
  __shared__ int32_t h[65];
  for (int t = threadIdx.x; t < 65; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N;
       t += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[nfound[ids ? ids[t] : t]], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < 65; t += blockDim.x)
    if (h[t]) atomicAdd(&hist[t], h[t]);
}

// Each block counts its nodes per bucket in LDS, reserves one range per bucket
// with a single global atomic, then places its nodes.
constexpr int kScatterChunk = 4096;

__global__ __launch_bounds__(256) void k_n_scatter(const int32_t* nfound, const int32_t* ids,
                                                   int32_t N, int32_t* cursor, int32_t* list) {
  __shared__ int32_t h[65], base[65];
  const int64_t i0 = (int64_t)blockIdx.x * kScatterChunk;
  const int64_t i1 = i0 + kScatterChunk < N ? i0 + kScatterChunk : N;
  for (int t = threadIdx.x; t < 65; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (int64_t t = i0 + threadIdx.x; t < i1; t += blockDim.x)
    atomicAdd(&h[nfound[ids ? ids[t] : t]], 1);
  __syncthreads();
  for (int t = threadIdx.x; t < 65; t += blockDim.x) {
    base[t] = h[t] ? atomicAdd(&cursor[t], h[t]) : 0;
    h[t] = 0;
  }
  __syncthreads();
  for (int64_t t = i0 + threadIdx.x; t < i1; t += blockDim.x) {
    const int i = ids ? ids[t] : (int)t;
    const int k = nfound[i];
    list[base[k] + atomicAdd(&h[k], 1)] = i;
  }
}

// ---------------------------------------------------------- small kernels --
__global__ void k_iota_i32(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (int32_t)i;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = v;
}

// Request p's words [0, n) of src go to dst[off ..]; blockIdx.y = request.
struct Span8 {
  const uint64_t* src;
  int64_t off, n;
};

__global__ void k_gather8(const Span8* spans, uint64_t* dst) {
  const Span8 s = spans[blockIdx.y];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < s.n;
       t += (int64_t)gridDim.x * blockDim.x)
    dst[s.off + t] = s.src[t];
}

// Request p's path (local ids, in place) becomes (p << 32) | global id;
// node_req of the request's nodes is set. Out-of-range ids are flagged and
// replaced by -1.
__global__ void k_batch_path(int64_t* path, const int64_t* node_off, int32_t* node_req,
                             int32_t* counters) {
  const int p = blockIdx.y;
  const int64_t off = node_off[p], n = node_off[p + 1] - off;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = path[off + t];
    if (i < 0 || i >= n) {
      atomicOr(&counters[CTR_POS_ERR], 1);
      path[off + t] = -1;
    } else {
      path[off + t] = ((int64_t)p << 32) | (off + i);
    }
    node_req[off + t] = p;
  }
}

// pos[i] = q for the node id i in the low 32 bits of path[q], flagging
// out-of-range ids; plain stores (a repeated id leaves one of its positions).
__global__ void k_positions(const int64_t* path, int64_t N, int32_t* pos, int32_t* counters) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = path[q];
    const int64_t i = v & 0xFFFFFFFF;
    if (v < 0 || i >= N) {
      atomicOr(&counters[CTR_POS_ERR], 1);
      continue;
    }
    pos[i] = (int)q;
  }
}

// Repeats: a path position whose id's recorded position is another one holds
// an id that occurs twice (N positions, so a repeat also means a missing id).
__global__ void k_positions_check(const int64_t* path, int64_t N, const int32_t* pos,
                                  int32_t* counters) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = path[q];
    const int64_t i = v & 0xFFFFFFFF;
    if (v >= 0 && i < N && pos[i] != (int)q) atomicOr(&counters[CTR_POS_ERR], 2);
  }
}

// Dependency levels by Kahn's algorithm on the device. level(i) = 0 without
// grid neighbours, else 1 + the largest level of its grid neighbours; the
// nodes of level l are exactly the frontier after l rounds of removing
// finished nodes. Reverse edges (neighbour -> dependent) live in CSR form.
//
// A node draws from the prior, and so depends on nothing, when it has fewer
// than kmin neighbours (known after the search) or its kriging system is not
// positive definite (known after the solve). The levels are built from the
// search alone (FLAGS = false), beside the solves; a plan with failed systems
// builds them again from the flags (FLAGS = true).
struct DepArgs {
  int32_t N, K;
  const Req* reqs;
  const int32_t* node_req;
  const int32_t* nfound;
  const uint8_t* flags;
  const int32_t* nbr;
};

template <bool FLAGS>
__device__ inline int deps_of(const DepArgs& D, int i) {
  if constexpr (FLAGS) return (D.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) ? 0 : D.nfound[i];
  const int n = D.nfound[i];
  return n < D.reqs[D.node_req[i]].kmin ? 0 : n;
}

// Grid neighbours per node (indeg) and dependents per node (outdeg, atomic
// counts; data neighbours have ids >= N and are not edges), sixteen lanes per
// node, one neighbour each. mark[i] = 1 for the nodes without grid
// neighbours (ready for level 0), else 0.
template <bool FLAGS>
__global__ __launch_bounds__(256) void k_deg_count(const DepArgs D, int32_t* outdeg,
                                                   int32_t* indeg, uint8_t* mark,
                                                   unsigned long long* edges) {
  const int sub = threadIdx.x & 15;
  uint32_t total = 0;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < D.N;
       i += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    const int k = deps_of<FLAGS>(D, (int)i);
    const int32_t* nb = D.nbr + i * D.K;
    int c = 0;
    for (int j = sub; j < k; j += 16) {
      const int a = nb[j];
      if (a < D.N) {
        atomicAdd(&outdeg[a], 1);
        ++c;
      }
    }
    for (int o = 8; o > 0; o >>= 1) c += __shfl_xor(c, o, 16);
    if (sub == 0) {
      indeg[i] = c;
      mark[i] = c == 0 ? 1 : 0;
      total += (uint32_t)c;
    }
  }
  // the number of edges: one atomic per wavefront
  unsigned long long w = total;
  for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(edges, w);
}

// Reverse edges in CSR form, sixteen lanes per node, one edge each.
template <bool FLAGS>
__global__ __launch_bounds__(256) void k_dep_fill(const DepArgs D, const int32_t* rowptr,
                                                  int32_t* cursor, int32_t* deps) {
  const int sub = threadIdx.x & 15;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; i < D.N;
       i += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    const int k = deps_of<FLAGS>(D, (int)i);
    const int32_t* nb = D.nbr + i * D.K;
    for (int j = sub; j < k; j += 16) {
      const int a = nb[j];
      if (a < D.N) deps[rowptr[a] + atomicAdd(&cursor[a], 1)] = (int)i;
    }
  }
}

// Level L + 1 in three launches. Push: every node of frontier L (order[off[L]
// .. off[L] + cnt[L])) decrements the counters of its dependents, sixteen lanes
// per node, one edge each; the atomic that takes a counter to zero marks that
// dependent ready (mark 1). Count and place: the marks, sixteen per thread in
// one 16-byte load, give every block's number of ready nodes, then every
// block adds up the counts of the blocks before it and writes its ready nodes
// in id order: level L + 1, mark 2 (placed) and a slot in frontier L + 1. No
// same-address atomics, and each level lists its nodes by ascending id. L = -1
// places level 0 (nodes without grid neighbours, marked by the degree count).
__global__ __launch_bounds__(256) void k_level_push(int32_t L, const int32_t* rowptr,
                                                    const int32_t* deps, int32_t* indeg,
                                                    uint8_t* mark, const int32_t* order,
                                                    const int32_t* off, const int32_t* cnt) {
  const int32_t start = off[L], size = cnt[L];
  const int sub = threadIdx.x & 15;
  for (int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4; t < size;
       t += ((int64_t)gridDim.x * blockDim.x) >> 4) {
    const int a = order[start + t];
    const int e1 = rowptr[a + 1];
    for (int e = rowptr[a] + sub; e < e1; e += 16) {
      const int d = deps[e];
      if (atomicSub(&indeg[d], 1) == 1) mark[d] = 1;
    }
  }
}

// Every level of a small plan in one workgroup of 1024 threads, without
// launches or host round trips between levels: the dependency counters are
// bytes in LDS (four to a word; a count never drops below zero, so the
// packed decrements never borrow), thread t owns nodes t + 1024 m and keeps
// their placed bits in a register. Per level: each thread counts its ready
// nodes, a block scan gives every thread its slots in the new level, then
// the new level's nodes release their dependents. cnt[depth] = 0 ends the
// levels; writes depth to *depth.
constexpr int kSmallLevels = 32768;  // nodes up to which one workgroup does all levels

__global__ __launch_bounds__(1024) void k_levels_small(int32_t N, const int32_t* rowptr,
                                                       const int32_t* deps,
                                                       const int32_t* indeg, int32_t* level,
                                                       int32_t* order, int32_t* off,
                                                       int32_t* cnt, int32_t* depth) {
  __shared__ uint32_t s_deg[kSmallLevels / 4];
  __shared__ int32_t s_wave[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int q = tid; q < (N + 3) / 4; q += 1024) {
    uint32_t word = 0;
    for (int b = 0; b < 4; ++b)
      if (4 * q + b < N) word |= (uint32_t)indeg[4 * q + b] << (8 * b);
    s_deg[q] = word;
  }
  __syncthreads();
  uint32_t placed = 0;  // bit m: node tid + 1024 m has its level
  int32_t start = 0, size = 0;
  for (int32_t L = -1;; ++L) {
    uint32_t ready = 0;
    int c = 0;
    for (int m = 0; tid + 1024 * m < N; ++m) {
      const int i = tid + 1024 * m;
      const uint32_t d = (s_deg[i >> 2] >> (8 * (i & 3))) & 0xFFu;
      if (!((placed >> m) & 1u) && d == 0) {
        ready |= 1u << m;
        ++c;
      }
    }
    int incl = c;  // inclusive scan over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    int before = 0, total = 0;
    for (int v = 0; v < 16; ++v) {
      before += v < w ? s_wave[v] : 0;
      total += s_wave[v];
    }
    const int32_t next = start + size;
    int32_t slot = next + before + incl - c;
    for (int m = 0; tid + 1024 * m < N; ++m)
      if ((ready >> m) & 1u) {
        const int i = tid + 1024 * m;
        order[slot++] = i;
        level[i] = L + 1;
      }
    placed |= ready;
    if (tid == 0) {
      off[L + 1] = next;
      cnt[L + 1] = total;
    }
    if (total == 0) {
      if (tid == 0) *depth = L + 1;
      return;
    }
    start = next;
    size = total;
    __syncthreads();  // the new level's order[] entries visible; s_wave free again
    for (int32_t t = tid; t < size; t += 1024) {
      const int a = order[start + t];
      const int e0 = rowptr[a], e1 = rowptr[a + 1];
      // eight dependents' ids loaded together, then their decrements
      for (int e = e0; e < e1; e += 8) {
        int dj[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) dj[u] = e + u < e1 ? deps[e + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (dj[u] >= 0) atomicSub(&s_deg[dj[u] >> 2], 1u << (8 * (dj[u] & 3)));
      }
    }
    __syncthreads();
  }
}

// Ready nodes among 16 marks (bytes 0, 1 or 2: ready is the only odd value).
__device__ __forceinline__ int ready16(const uint4 m) {
  return __popc(m.x & 0x01010101u) + __popc(m.y & 0x01010101u) + __popc(m.z & 0x01010101u) +
         __popc(m.w & 0x01010101u);
}

// Each block takes a contiguous run of 16-node groups (the marks are padded
// to whole groups with zeros).
__device__ __forceinline__ void collect_chunk(int64_t n16, int64_t& g0, int64_t& g1) {
  const int64_t chunk = ((n16 + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
  g0 = (int64_t)blockIdx.x * chunk;
  g1 = g0 + chunk < n16 ? g0 + chunk : n16;
}

__device__ __forceinline__ int block_sum256(int v, int32_t* s_red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();  // s_red free
  if (lane == 0) s_red[w] = v;
  __syncthreads();
  return s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

__global__ __launch_bounds__(256) void k_level_count(int64_t n16, const uint4* mark16,
                                                     int32_t* blockcnt) {
  __shared__ int32_t s_red[4];
  int64_t g0, g1;
  collect_chunk(n16, g0, g1);
  int c = 0;
  for (int64_t g = g0 + threadIdx.x; g < g1; g += 256) c += ready16(mark16[g]);
  c = block_sum256(c, s_red);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(256) void k_level_place(int64_t n16, int32_t L, uint4* mark16,
                                                     const int32_t* blockcnt, int32_t* level,
                                                     int32_t* order, int32_t* off,
                                                     int32_t* cnt) {
  __shared__ int32_t s_red[4];
  __shared__ int32_t s_scan[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int32_t next = L < 0 ? 0 : off[L] + cnt[L];
  // ready nodes in the blocks before this one
  int pre = 0;
  for (int b = tid; b < (int)blockIdx.x; b += 256) pre += blockcnt[b];
  pre = block_sum256(pre, s_red);
  const int mine = blockcnt[blockIdx.x];
  if (tid == 0) {
    if (blockIdx.x == 0) off[L + 1] = next;
    if (blockIdx.x == gridDim.x - 1) cnt[L + 1] = pre + mine;
  }
  if (mine == 0) return;  // uniform per block
  int64_t g0, g1;
  collect_chunk(n16, g0, g1);
  int32_t running = next + pre;
  for (int64_t b0 = g0; b0 < g1; b0 += 256) {
    const int64_t g = b0 + tid;
    const uint4 m = g < g1 ? mark16[g] : make_uint4(0u, 0u, 0u, 0u);
    const int c = ready16(m);
    int incl = c;  // inclusive scan over the block
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o);
      if (lane >= o) incl += v;
    }
    __syncthreads();  // s_scan of the previous step has been read
    if (lane == 63) s_scan[w] = incl;
    __syncthreads();
    int before = 0;
    for (int v = 0; v < w; ++v) before += s_scan[v];
    if (c) {
      int32_t slot = running + before + incl - c;
      const uint32_t words[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (((words[q] >> (8 * k)) & 0xFFu) == 1u) {
            const int32_t i = (int32_t)(g * 16 + 4 * q + k);
            order[slot++] = i;
            level[i] = L + 1;
          }
      // ready (1) becomes placed (2); 0 and 2 stay
      mark16[g] = make_uint4(m.x + (m.x & 0x01010101u), m.y + (m.y & 0x01010101u),
                             m.z + (m.z & 0x01010101u), m.w + (m.w & 0x01010101u));
    }
    running += s_scan[0] + s_scan[1] + s_scan[2] + s_scan[3];
  }
}

// ------------------------------------------------------------- the sweep --
// Zt[q * RC + r] = z[(r0 + r) * n_path + q]: normals of one realization chunk
// made node-major (path order) through a 64-position LDS tile of RC rows
// (dynamic LDS, RC x 65 doubles: a small chunk leaves room for more
// workgroups per CU).
__global__ __launch_bounds__(256) void k_z_transpose(const double* z, int32_t n_path, int32_t RC,
                                                     int32_t r0, double* Zt) {
  extern __shared__ double s_ttile[];
  const int64_t q0 = (int64_t)blockIdx.x * 64;
  const int cnt = (int)((n_path - q0) < 64 ? (n_path - q0) : 64);
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int rr = t / cnt, qq = t - rr * cnt;
    s_ttile[rr * 65 + qq] = z[(int64_t)(r0 + rr) * n_path + q0 + qq];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int qq = t / RC, rr = t - qq * RC;
    Zt[(q0 + qq) * RC + rr] = s_ttile[rr * 65 + qq];
  }
}

// LDS bytes of the transposes' tiles for a chunk of RC realizations.
inline size_t transpose_lds(int32_t RC) { return (size_t)RC * 65 * sizeof(double); }

struct LevelArgs {
  const int32_t* order;  // the nodes of this level
  int32_t count;
  int32_t K, RC, shift;
  const int32_t* pos;
  const int32_t* nfound;
  const uint8_t* flags;
  const int32_t* nbr;
  const double* lam;
  const double* sigma;
  const int32_t* node_req;
  const double* mean;  // per request
  const double* Zt;    // n_path x RC
  double* X;           // (NT + n_data) x RC
};

// All nodes of one level are independent: x_i = (mean + sum_j lambda_ij
// (x_nbr(j) - mean)) + sigma_i z_i for every realization, 1 << shift lanes per
// node (lane = realization), neighbours from earlier levels.
// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2);
// block b works on chunk (b % 8) * per + b / 8 of the level, so every XCD
// sweeps one contiguous run of the level's (id-ordered) nodes and the
// neighbour rows those nodes share stay in one L2. The grid is 8 * per blocks.

__global__ __launch_bounds__(256) void k_sweep_level(const LevelArgs A) {
  const int per = gridDim.x / kXcds;
  const int64_t blk = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  const int64_t t = blk * blockDim.x + threadIdx.x;
  const int64_t slot = t >> A.shift;
  const int r = (int)(t & ((1 << A.shift) - 1));
  if (slot >= A.count || r >= A.RC) return;
  const int i = A.order[slot];
  const double mean = A.mean[A.node_req[i]];
  const int k = (A.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) ? 0 : A.nfound[i];
  const int32_t* nb = A.nbr + (int64_t)i * A.K;
  const double* lm = A.lam + (int64_t)i * A.K;
  // eight neighbours at a time: indices and weights, then their eight row
  // gathers in flight together, then the accumulation in neighbour order
  double acc = 0.0;
  for (int j0 = 0; j0 < k; j0 += 8) {
    int nj[8];
    double lj[8], xj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      nj[u] = j0 + u < k ? nb[j0 + u] : 0;
      lj[u] = j0 + u < k ? lm[j0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xj[u] = j0 + u < k ? A.X[(int64_t)nj[u] * A.RC + r] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < k) acc = acc + lj[u] * (xj[u] - mean);
  }
  const double mu = mean + acc;
  A.X[(int64_t)i * A.RC + r] = mu + A.sigma[i] * A.Zt[(int64_t)A.pos[i] * A.RC + r];
}

// Multivariate sweep of an intrinsic coregionalization A * c(h): the normals
// of each path node, nv per node in draw order (the reference's MvNormal
// draw, seqsim.jl:335-340), mixed by the Cholesky factor of A,
// w_v = sum_{u <= v} L_vu z_u (ascending u), one array per variable.
struct MvMix {
  double L[4][4];
};

__global__ void k_mv_mix(const double* z, int64_t n_path, int32_t R, int32_t nv, const MvMix M,
                         double* w) {
  const int64_t total = (int64_t)R * n_path;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    double zv[4];
    for (int u = 0; u < nv; ++u) zv[u] = z[t * nv + u];
    for (int v = 0; v < nv; ++v) {
      double acc = 0.0;
      for (int u = 0; u <= v; ++u) acc = acc + M.L[v][u] * zv[u];
      w[(int64_t)v * total + t] = acc;
    }
  }
}

// Indicator sweep (an IndicatorProcess with an intrinsic coregionalization
// A * c(h): the plan's weights lambda_j serve every category). X holds the
// 0-based categories as doubles, Zt the uniforms of the categorical draws.
constexpr int kMaxCat = 8;

struct IndicatorArgs {
  LevelArgs L;
  int32_t n_cat;
  int32_t pad_;
  double prob[kMaxCat];
};

// The first category c whose running sum q_0 + ... + q_c exceeds u (the last
// one at most): the categorical draw of a probability vector.
__device__ inline int draw_category(const double (&q)[kMaxCat], int n_cat, double u) {
  int cat = 0;
  double cp = q[0];
  bool more = cp <= u;
#pragma unroll
  for (int c = 1; c < kMaxCat; ++c)
    if (c < n_cat && more) {
      cat = c;
      cp = cp + q[c];
      more = cp <= u;
    }
  return cat;
}

// Per node and realization: p_c + sum_j lambda_j ([x_j == c] - p_c) for every
// category (neighbours in order), clamped to [0, 1], multiplied by the
// inverse of their sum (normalize(., 1)), then the categorical draw; nodes
// that draw from the prior use the process probabilities.
__global__ __launch_bounds__(256) void k_sweep_level_ind(const IndicatorArgs I) {
  const LevelArgs& A = I.L;
  const int per = gridDim.x / kXcds;
  const int64_t blk = (int64_t)(blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  const int64_t t = blk * blockDim.x + threadIdx.x;
  const int64_t slot = t >> A.shift;
  const int r = (int)(t & ((1 << A.shift) - 1));
  if (slot >= A.count || r >= A.RC) return;
  const int i = A.order[slot];
  const bool prior = (A.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL)) != 0;
  const int k = prior ? 0 : A.nfound[i];
  const int32_t* nb = A.nbr + (int64_t)i * A.K;
  const double* lm = A.lam + (int64_t)i * A.K;
  const int n_cat = I.n_cat;
  double q[kMaxCat];
#pragma unroll
  for (int c = 0; c < kMaxCat; ++c) q[c] = I.prob[c];
  for (int j0 = 0; j0 < k; j0 += 8) {
    int nj[8];
    double lj[8], xj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      nj[u] = j0 + u < k ? nb[j0 + u] : 0;
      lj[u] = j0 + u < k ? lm[j0 + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xj[u] = j0 + u < k ? A.X[(int64_t)nj[u] * A.RC + r] : 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < k) {
#pragma unroll
        for (int c = 0; c < kMaxCat; ++c)
          if (c < n_cat) q[c] = q[c] + lj[u] * ((xj[u] == (double)c ? 1.0 : 0.0) - I.prob[c]);
      }
  }
  if (!prior) {
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < kMaxCat; ++c)
      if (c < n_cat) {
        const double x = q[c];
        q[c] = x > 1.0 ? 1.0 : (x < 0.0 ? 0.0 : x);  // clamp(x, 0, 1)
        sum = c == 0 ? fabs(q[c]) : sum + fabs(q[c]);
      }
    // normalize(q, 1): multiply by the inverse norm, through a scaling when
    // the norm is below 1 / floatmax
    const double delta = 1.0 / 0x1.fffffffffffffp+1023;
    if (sum >= delta) {
      const double inv = 1.0 / sum;
#pragma unroll
      for (int c = 0; c < kMaxCat; ++c) q[c] = q[c] * inv;
    } else {
      const double ed = 0x1p-52 / delta;
      const double inv = 1.0 / (sum * ed);
#pragma unroll
      for (int c = 0; c < kMaxCat; ++c) q[c] = (q[c] * ed) * inv;
    }
  }
  const double u = A.Zt[(int64_t)A.pos[i] * A.RC + r];
  A.X[(int64_t)i * A.RC + r] = (double)draw_category(q, n_cat, u);
}

__global__ void k_i32_to_f64(const int32_t* in, double* out, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (double)in[t];
}

// X rows NT..NT+n_data-1 hold the conditioning data for every realization.
__global__ void k_data_rows(double* X, int32_t N, int32_t n_data, int32_t RC,
                            const double* vals) {
  const int64_t total = (int64_t)n_data * RC;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x)
    X[(int64_t)N * RC + t] = vals[t / RC];
}

// out[(r0 + r) * N + i] = X[i * RC + r] through a 64-node LDS tile of RC rows
// (dynamic LDS, RC x 65 doubles; int32 outputs: the indicator sweep's
// categories, held as exact doubles in X).
template <class OUT>
__global__ __launch_bounds__(256) void k_transpose(const double* X, int32_t N, int32_t RC,
                                                   int32_t r0, OUT* out) {
  extern __shared__ double s_ttile[];
  const int64_t i0 = (int64_t)blockIdx.x * 64;
  const int cnt = (int)((N - i0) < 64 ? (N - i0) : 64);
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int ii = t / RC, rr = t - ii * RC;
    s_ttile[rr * 65 + ii] = X[(i0 + ii) * RC + rr];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt * RC; t += blockDim.x) {
    const int rr = t / cnt, ii = t - rr * cnt;
    out[(int64_t)(r0 + rr) * N + i0 + ii] = (OUT)s_ttile[rr * 65 + ii];
  }
}

// ------------------------------------------------------------- host side --
// Optional phase trace (environment GSP_TRACE=1): synchronises the stream at
// each mark and prints the host-side elapsed time of the phase to stderr.
struct PhaseTrace {
  bool on;
  hipStream_t s;
  std::chrono::steady_clock::time_point t;
  explicit PhaseTrace(hipStream_t st) : on(std::getenv("GSP_TRACE") != nullptr), s(st) {
    if (on) t = std::chrono::steady_clock::now();
  }
  void mark(const char* what) {
    if (!on) return;
    (void)hipStreamSynchronize(s);
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[gsp] %-24s %9.3f ms\n", what,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

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
  if (mode == GSP_DATA_ONLY) return {};  // no grid candidates
  const int nd = g.ndim;
  int64_t ext[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) ext[d] = g.dims[d] - 1;
  int64_t lim[3] = {0, 0, 0};
  double bound = r2;
  const bool ball = r2 < std::numeric_limits<double>::infinity();
  if (mode == GSP_KNN_THEN_MASK) {
    // the kmax nearest in-grid nodes of any node lie within the Chebyshev
    // radius M at which the smallest in-grid window holds kmax + 1 nodes;
    // farther offsets can never be among the k nearest, ball or not
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
    bound = std::min(acc, r2);
    for (int d = 0; d < nd; ++d) {
      const double f = std::floor(std::sqrt(bound) / g.spacing[d]) + 1.0;
      lim[d] = std::min<int64_t>(ext[d], (int64_t)f);
    }
  } else if (ball) {
    for (int d = 0; d < nd; ++d) {
      const double f = std::floor(std::sqrt(r2) / g.spacing[d]) + 1.0;
      lim[d] = std::min<int64_t>(ext[d], (int64_t)std::min(f, 1e9));
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
double scale_factor(const gsp_grid& g, const double* data, int64_t n, const gsp_model& m) {
  const int nd = g.ndim;
  double a1 = 0.0;
  for (int d = 0; d < nd; ++d) {
    const double lo = std::fabs(g.origin[d]);
    const double hi = std::fabs(g.origin[d] + (double)g.dims[d] * g.spacing[d]);
    a1 = std::max(a1, std::max(lo, hi));
  }
  double a2 = 0.0;
  for (int64_t t = 0; t < n; ++t) a2 = std::max(a2, std::fabs(data[t]));
  double a3 = 0.0;
  for (int d = 0; d < nd; ++d) a3 = std::max(a3, m.ranges[d]);
  return 1.0 / std::max(a1, std::max(a2, a3));
}

// Host-side preparation of one request: clamped neighbour counts, scale
// factor, offset table and data hash (item ids local, hstart local).
struct HostReq {
  int64_t N = 0, nD = 0;
  int nd = 0;
  int32_t kmin = 0, kmax = 0;
  double alpha = 0, r2 = 0;
  const double* data = nullptr;  // nD x nd, host
  std::vector<Offset> table;
  int hd[3] = {1, 1, 1};
  double hcs[3] = {1.0, 1.0, 1.0};
  int hexp = 0;
  int64_t hB = 1;
  std::vector<int32_t> hstart;
  std::vector<int32_t> hitems;  // request-local data indices
};

// Data hash. Expanded mode: cells of B^ndim grid cells, each listing every
// datum within rho_max of the cell (rho_max bounds the k-th grid distance of
// any node, so one list holds all candidates of a node). Box mode (large
// rho_max, e.g. masked search in a big ball): each datum in its own cell and
// nodes visit every cell overlapping their ball.
void build_hash(const gsp_grid& g, int mode, HostReq& h) {
  const int nd = h.nd;
  const int64_t nD = h.nD;
  const double* data = h.data;
  const double inf = std::numeric_limits<double>::infinity();
  h.hstart.assign(2, 0);
  h.hitems.clear();
  if (nD == 0) return;
  const double rho_max = h.table.empty() ? 0.0 : std::sqrt(h.table.back().d2);
  const bool bounded = !h.table.empty() && (mode == GSP_KNN_THEN_MASK || h.r2 < inf);
  double smin = g.spacing[0];
  for (int d = 1; d < nd; ++d) smin = std::min(smin, g.spacing[d]);
  if (bounded) {
    h.hB = std::max<int64_t>(1, (int64_t)std::llround(rho_max / smin));
    int64_t ncell = 1, span = 1;
    for (int d = 0; d < nd; ++d) {
      h.hd[d] = (int)((g.dims[d] + h.hB - 1) / h.hB);
      h.hcs[d] = (double)h.hB * g.spacing[d];
      ncell *= h.hd[d];
      span *= std::min<int64_t>(h.hd[d], (int64_t)std::ceil(2.0 * rho_max / h.hcs[d]) + 2);
    }
    // worth it when the lists stay short (about one 64-lane pass per node)
    h.hexp = (span * nD <= ((int64_t)1 << 24) && ncell <= ((int64_t)1 << 26) &&
              span * nD <= 64 * ncell)
                 ? 1
                 : 0;
    if (h.hexp) {
      const double rho = rho_max * (1.0 + 1e-9) + 1e-12;
      std::vector<int64_t> lo((size_t)(nD * 3)), hi((size_t)(nD * 3));
      h.hstart.assign((size_t)ncell + 1, 0);
      for (int64_t i = 0; i < nD; ++i) {
        int64_t l[3] = {0, 0, 0}, u[3] = {0, 0, 0};
        for (int d = 0; d < nd; ++d) {
          const double x = data[(size_t)(i * nd + d)];
          double fl = std::floor((x - rho - g.origin[d]) / h.hcs[d]);
          double fh = std::floor((x + rho - g.origin[d]) / h.hcs[d]);
          l[d] = (int64_t)std::min(std::max(fl, 0.0), (double)(h.hd[d] - 1));
          u[d] = (int64_t)std::min(std::max(fh, 0.0), (double)(h.hd[d] - 1));
        }
        for (int d = 0; d < 3; ++d) {
          lo[(size_t)(i * 3 + d)] = l[d];
          hi[(size_t)(i * 3 + d)] = u[d];
        }
        for (int64_t cz = l[2]; cz <= u[2]; ++cz)
          for (int64_t cy = l[1]; cy <= u[1]; ++cy)
            for (int64_t cx = l[0]; cx <= u[0]; ++cx)
              h.hstart[(size_t)(cx + h.hd[0] * (cy + (int64_t)h.hd[1] * cz)) + 1]++;
      }
      for (int64_t c = 0; c < ncell; ++c) h.hstart[(size_t)c + 1] += h.hstart[(size_t)c];
      h.hitems.assign((size_t)h.hstart[(size_t)ncell], 0);
      std::vector<int32_t> fill((size_t)ncell, 0);
      for (int64_t i = 0; i < nD; ++i) {
        const int32_t it = (int32_t)i;
        for (int64_t cz = lo[(size_t)(i * 3 + 2)]; cz <= hi[(size_t)(i * 3 + 2)]; ++cz)
          for (int64_t cy = lo[(size_t)(i * 3 + 1)]; cy <= hi[(size_t)(i * 3 + 1)]; ++cy)
            for (int64_t cx = lo[(size_t)(i * 3)]; cx <= hi[(size_t)(i * 3)]; ++cx) {
              const int64_t c = cx + h.hd[0] * (cy + (int64_t)h.hd[1] * cz);
              h.hitems[(size_t)(h.hstart[(size_t)c] + fill[(size_t)c]++)] = it;
            }
      }
    }
  }
  if (!h.hexp) {
    // about one datum per cell; kmax / 2 per cell for the data-only search,
    // whose rings then rarely pass the first ring of neighbours
    const double per = mode == GSP_DATA_ONLY ? std::max(1.0, 0.5 * h.kmax) : 1.0;
    int64_t B = (int64_t)std::llround(std::pow((double)h.N * per / (double)nD, 1.0 / nd));
    if (B < 1) B = 1;
    h.hB = B;
    int64_t ncell = 1;
    for (int d = 0; d < nd; ++d) {
      h.hd[d] = (int)((g.dims[d] + B - 1) / B);
      h.hcs[d] = (double)B * g.spacing[d];
      ncell *= h.hd[d];
    }
    std::vector<int64_t> cell((size_t)nD);
    h.hstart.assign((size_t)ncell + 1, 0);
    for (int64_t i = 0; i < nD; ++i) {
      int64_t c = 0, stride = 1;
      for (int d = 0; d < nd; ++d) {
        double f = std::floor((data[(size_t)(i * nd + d)] - g.origin[d]) / h.hcs[d]);
        f = std::min(std::max(f, 0.0), (double)(h.hd[d] - 1));
        c += (int64_t)f * stride;
        stride *= h.hd[d];
      }
      cell[(size_t)i] = c;
      h.hstart[(size_t)c + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) h.hstart[(size_t)c + 1] += h.hstart[(size_t)c];
    h.hitems.assign((size_t)nD, 0);
    std::vector<int32_t> fill((size_t)ncell, 0);
    for (int64_t i = 0; i < nD; ++i)
      h.hitems[(size_t)(h.hstart[(size_t)cell[(size_t)i]] + fill[(size_t)cell[(size_t)i]]++)] =
          (int32_t)i;
  }
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
  for (int d = 0; d < 3; ++d) M.irs[d] = d < ndim ? 1.0 / (alpha * m.ranges[d]) : 1.0;
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

namespace {

// Checks one problem and clamps its neighbour counts (seqsim.jl:202-208).
void validate_problem(const gsp_seqsim_problem* p, HostReq& h) {
  GSP_REQUIRE(p, GSP_ERR_ARGUMENT, "null problem");
  const gsp_grid& g = p->grid;
  h.N = validate_grid(g);
  h.nd = g.ndim;
  validate_model(p->model, h.nd);
  GSP_REQUIRE(p->n_data >= 0 && h.N + p->n_data < ((int64_t)1 << 30), GSP_ERR_ARGUMENT,
              "invalid number of data points");
  GSP_REQUIRE(p->n_data == 0 || p->data_coords, GSP_ERR_ARGUMENT, "null data coordinates");
  GSP_REQUIRE(p->n_path == h.N && p->path, GSP_ERR_UNSUPPORTED,
              "the path must visit every grid node exactly once");
  GSP_REQUIRE(p->search_mode == GSP_KNN_THEN_MASK || p->search_mode == GSP_MASKED_KNN ||
                  p->search_mode == GSP_DATA_ONLY,
              GSP_ERR_ARGUMENT, "unknown search mode");
  GSP_REQUIRE(p->precision == GSP_FP64 || p->precision == GSP_FP32_REFINE, GSP_ERR_ARGUMENT,
              "unknown precision");
  GSP_REQUIRE(std::isfinite(p->mean), GSP_ERR_ARGUMENT, "invalid mean");
  int32_t nmin = p->minneighbors, nmax = p->maxneighbors;
  if ((int64_t)nmax > h.N || nmax < 1) nmax = (int32_t)h.N;
  if (nmin > nmax || nmin < 1) nmin = 1;
  GSP_REQUIRE(nmax <= 64, GSP_ERR_UNSUPPORTED, "maxneighbors above 64 is not supported");
  h.kmin = nmin;
  h.kmax = nmax;
  h.nD = p->n_data;
  const double inf = std::numeric_limits<double>::infinity();
  h.r2 = p->radius > 0 ? p->radius * p->radius : inf;
}

// ---- kriging of 16 < n <= 32: two kernels, the faster one measured --------
// k_krige32 (two systems per wavefront, LDS column broadcasts) and k_krige<KB>
// (one system per wavefront, readlanes) give identical results. Which is
// faster depends on the size bucket, the precision and the work beside them,
// so the first plan of a context with at least kTuneMin systems in a bucket
// times both, twice in alternation, on a slice of kTuneSample systems from
// the middle of the bucket (which is sorted by size) and keeps the faster for
// that bucket and precision (GSP_KRIGE_VARIANT=0 or 1 forces one; until
// measured, one system per wavefront).
enum { kVarPair = 0, kVarSingle = 1 };
constexpr int64_t kTuneMin = 65536, kTuneSample = 32768;

template <int KB>
void launch_variant(int variant, const PlanArgs& A, int precision, const int32_t* list,
                    int64_t count, hipStream_t s) {
  if (count <= 0) return;
  if (variant == kVarSingle) {
    launch_krige<KB>(A, precision, list, count, s);
    return;
  }
  const int blocks = grid_for((count + 1) / 2, k32Waves, 65536);
  if (precision == GSP_FP32_REFINE)
    hipLaunchKernelGGL((k_krige32<KB, GSP_FP32_REFINE>), dim3(blocks), dim3(64 * k32Waves), 0, s,
                       A, list, (int32_t)count);
  else
    hipLaunchKernelGGL((k_krige32<KB, GSP_FP64>), dim3(blocks), dim3(64 * k32Waves), 0, s, A, list,
                       (int32_t)count);
}

template <int KB>
void launch_mid(gsp_ctx* ctx, const PlanArgs& A, int precision, const int32_t* list,
                int64_t count, hipStream_t s) {
  if (count <= 0) return;
  int& v = ctx->krige_variant[precision == GSP_FP32_REFINE ? 1 : 0][KB == 24 ? 0 : 1];
  if (const char* env = std::getenv("GSP_KRIGE_VARIANT")) {
    v = std::atoi(env) == kVarSingle ? kVarSingle : kVarPair;
  } else if (v < 0 && count >= kTuneMin) {
    // the sample's outputs are written again by the full launch; its counters
    // go to a scratch buffer
    DevBuf<int32_t> ctr;
    ctr.alloc(16);
    PlanArgs T = A;
    T.counters = ctr.p;
    const int32_t* slice = list + (count - kTuneSample) / 2;
    float best[2] = {1e30f, 1e30f};
    for (int rep = 0; rep < 4; ++rep) {
      const int var = rep & 1;
      hipEvent_t e0 = take_event(ctx), e1 = take_event(ctx);
      GSP_HIP(hipEventRecord(e0, s));
      launch_variant<KB>(var, T, precision, slice, kTuneSample, s);
      GSP_HIP(hipEventRecord(e1, s));
      GSP_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      GSP_HIP(hipEventElapsedTime(&ms, e0, e1));
      best[var] = std::min(best[var], ms);
      ctx->event_pool.push_back(e0);
      ctx->event_pool.push_back(e1);
    }
    v = best[kVarPair] < best[kVarSingle] ? kVarPair : kVarSingle;
  }
  launch_variant<KB>(v < 0 ? kVarSingle : v, A, precision, list, count, s);
}

// Runs `f` for request index q; on failure in a batch of several requests the
// error message names the request.
template <class F>
void for_request(int32_t n, int32_t q, F&& f) {
  try {
    f();
  } catch (const Fail&) {
    if (n > 1) {
      const std::string msg = gsp_last_error();
      set_error("request %d: %s", q, msg.c_str());
    }
    throw;
  }
}

// Runs f(q) for every request on up to 16 host threads (host-only work: no
// HIP calls). The failure of the lowest request index is raised on the
// calling thread, with its message.
template <class F>
void parallel_requests(int32_t n, F&& f) {
  const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
  const int T = (int)std::min<unsigned>(std::min<unsigned>(hc, 16u), (unsigned)((n + 7) / 8));
  if (T <= 1) {
    for (int32_t q = 0; q < n; ++q) for_request(n, q, [&] { f(q); });
    return;
  }
  std::vector<int32_t> codes((size_t)n, 0);
  std::vector<std::string> msgs((size_t)n);
  std::atomic<int32_t> next{0};
  auto worker = [&] {
    for (;;) {
      const int32_t q = next.fetch_add(1);
      if (q >= n) break;
      try {
        f(q);
      } catch (const Fail& e) {
        codes[(size_t)q] = e.code;
        msgs[(size_t)q] = gsp_last_error();
      } catch (const std::bad_alloc&) {
        codes[(size_t)q] = GSP_ERR_MEMORY;
        msgs[(size_t)q] = "host allocation failed";
      }
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  for (int32_t q = 0; q < n; ++q)
    if (codes[(size_t)q]) {
      set_error("request %d: %s", q, msgs[(size_t)q].c_str());
      throw Fail{codes[(size_t)q]};
    }
}

}  // namespace

namespace {

// Buffers of the level computation, all made before the worker starts (the
// context's allocator hands out blocks in main-stream order only).
struct LevelBufs {
  DevBuf<int32_t> rowptr, cursor, deps, cnt, off, depth, blockcnt;
  DevBuf<uint8_t> mark;  // 0 waiting, 1 ready, 2 placed; zero padding to 16 nodes
  DevBuf<char> tmp;
  size_t tmp_bytes = 0;
};

// The ready marks: whole 16-node groups, zero beyond the last node.
void mark_alloc(LevelBufs& B, int64_t NT, hipStream_t s) {
  const size_t bytes = (size_t)((NT + 15) / 16 * 16);
  B.mark.alloc(bytes);
  GSP_HIP(hipMemsetAsync(B.mark.p, 0, bytes, s));
}

void level_bufs_alloc(LevelBufs& B, int64_t NT, int64_t n_edges, const int32_t* outdeg,
                      hipStream_t s) {
  B.rowptr.alloc((size_t)NT + 1);
  B.cursor.alloc((size_t)NT);
  B.cnt.alloc((size_t)NT + 16);
  B.off.alloc((size_t)NT + 16);
  B.depth.alloc(1);
  B.blockcnt.alloc(1024);
  if (!B.mark.p) mark_alloc(B, NT, s);  // kept when the degree count has filled it
  B.deps.alloc((size_t)std::max<int64_t>(n_edges, 1));
  B.tmp_bytes = 0;
  GSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, B.tmp_bytes, outdeg, B.rowptr.p,
                                           (int)NT + 1, s));
  B.tmp.alloc(B.tmp_bytes);
}

// Blocks of the passes over the edges (degree count, fill, push). They are
// bound by device atomics, which 512 blocks of 256 threads keep busy; more
// would only hold wave slots that the solves on the main stream need.
constexpr int kEdgeBlocks = 512;

// Kahn's algorithm on stream `st`: reverse CSR, level 0, then three launches
// per level (push, count, place); the frontier sizes are checked every kBatch
// levels. Plans of at most kSmallLevels nodes run every level in one workgroup. Sets
// plan->level, order, depth and level_off.
template <bool FLAGS>
void build_levels(gsp_seqsim_plan* plan, const DepArgs& D, const int32_t* outdeg,
                  const int32_t* indeg_in, LevelBufs& B, hipStream_t st) {
  int32_t* indeg = const_cast<int32_t*>(indeg_in);
  const int64_t NT = D.N;
  const int nb = grid_for(NT, 16, kEdgeBlocks);  // sixteen nodes per block
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(NT + 16, 256, 4096)), dim3(256), 0, st, B.cnt.p,
                     0, NT + 16);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, B.off.p, 0, (int64_t)1);
  size_t tmp_bytes = B.tmp_bytes;
  GSP_HIP(hipcub::DeviceScan::ExclusiveSum(B.tmp.p, tmp_bytes, outdeg, B.rowptr.p, (int)NT + 1,
                                           st));
  hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, B.cursor.p, 0, NT);
  hipLaunchKernelGGL(k_dep_fill<FLAGS>, dim3(nb), dim3(256), 0, st, D, B.rowptr.p, B.cursor.p,
                     B.deps.p);
  const int cb = grid_for(NT, 2048, 1024);
  int32_t depth = -1;
  std::vector<int32_t> hcnt;  // the level sizes, when read together with the depth
  if (NT <= kSmallLevels) {
    // one workgroup, one launch, one read of the depth and the first level sizes
    hipLaunchKernelGGL(k_levels_small, dim3(1), dim3(1024), 0, st, (int32_t)NT, B.rowptr.p,
                       B.deps.p, indeg, plan->level.p, plan->order.p, B.off.p, B.cnt.p,
                       B.depth.p);
    const int64_t npre = NT + 16 < 1024 ? NT + 16 : 1024;
    std::vector<int32_t> hpre((size_t)npre);
    int32_t d = 0;
    GSP_HIP(hipMemcpyAsync(&d, B.depth.p, 4, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipMemcpyAsync(hpre.data(), B.cnt.p, (size_t)npre * 4, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
    depth = d;
    if (depth >= 1 && depth <= npre) hcnt.assign(hpre.begin(), hpre.begin() + depth);
  } else {
  const int64_t n16 = (NT + 15) / 16;
  uint4* mark16 = reinterpret_cast<uint4*>(B.mark.p);
  auto place = [&](int32_t L) {
    hipLaunchKernelGGL(k_level_count, dim3(cb), dim3(256), 0, st, n16, mark16, B.blockcnt.p);
    hipLaunchKernelGGL(k_level_place, dim3(cb), dim3(256), 0, st, n16, L, mark16, B.blockcnt.p,
                       plan->level.p, plan->order.p, B.off.p, B.cnt.p);
  };
  place(-1);
  constexpr int kBatch = 16;
  const int fb = grid_for(NT, 16, kEdgeBlocks);
  std::vector<int32_t> hc(kBatch + 1);
  for (int32_t L = 0; depth < 0; L += kBatch) {
    GSP_REQUIRE(L + kBatch + 1 < NT + 16, GSP_ERR_INTERNAL, "dependency levels did not close");
    for (int b = 0; b < kBatch; ++b) {
      hipLaunchKernelGGL(k_level_push, dim3(fb), dim3(256), 0, st, L + b, B.rowptr.p, B.deps.p,
                         indeg, B.mark.p, plan->order.p, B.off.p, B.cnt.p);
      place(L + b);
    }
    GSP_HIP(hipMemcpyAsync(hc.data(), B.cnt.p + L, (kBatch + 1) * 4, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
    for (int b = 0; b <= kBatch; ++b)
      if (hc[(size_t)b] == 0) {
        depth = L + b;
        break;
      }
  }
  }
  GSP_REQUIRE(depth >= 1, GSP_ERR_INTERNAL, "empty dependency levels");
  if (hcnt.empty()) {
    hcnt.resize((size_t)depth);
    GSP_HIP(hipMemcpyAsync(hcnt.data(), B.cnt.p, (size_t)depth * 4, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
  }
  plan->depth = depth;
  plan->level_off.assign((size_t)depth + 1, 0);
  for (int32_t l = 0; l < depth; ++l)
    plan->level_off[(size_t)l + 1] = plan->level_off[(size_t)l] + hcnt[(size_t)l];
  GSP_REQUIRE(plan->level_off[(size_t)depth] == NT, GSP_ERR_INTERNAL,
              "dependency levels cover %lld of %lld nodes",
              (long long)plan->level_off[(size_t)depth], (long long)NT);
}

struct LevelJob {
  int32_t code = GSP_OK;
  std::string msg;
  hipEvent_t a = nullptr, b = nullptr;  // timing (when enabled)
};

// The worker: the levels from the search's degree counts on the side stream,
// after the fork.
void levels_on_side(gsp_ctx* ctx, gsp_seqsim_plan* plan, const DepArgs& D, int32_t* outdeg,
                    int32_t* indeg, LevelBufs& B, LevelJob& job) {
  try {
    GSP_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->side;
    GSP_HIP(hipStreamWaitEvent(st, ctx->fork, 0));
    if (job.a) GSP_HIP(hipEventRecord(job.a, st));
    build_levels<false>(plan, D, outdeg, indeg, B, st);
    if (job.b) GSP_HIP(hipEventRecord(job.b, st));
    GSP_HIP(hipStreamSynchronize(st));
  } catch (const Fail& f) {
    job.code = f.code;
    job.msg = gsp_last_error();
  } catch (...) {
    job.code = GSP_ERR_INTERNAL;
    job.msg = "dependency level worker failed";
  }
}

// A thread joined when it goes out of scope (also when the caller throws).
struct WorkerThread {
  std::thread t;
  template <class F>
  explicit WorkerThread(F&& f) : t(std::forward<F>(f)) {}
  ~WorkerThread() {
    if (t.joinable()) t.join();
  }
};

}  // namespace

void seqsim_plan_build(gsp_ctx* ctx, int32_t n_req, const gsp_seqsim_problem* problems,
                       gsp_seqsim_plan* plan, const gsp_surrogate* const* surrogates,
                       int32_t n_surrogates, double tolerance) {
  GSP_REQUIRE(problems, GSP_ERR_ARGUMENT, "null problem");
  const bool learned_plan = surrogates && n_surrogates > 0;
  GSP_REQUIRE(n_req >= 1 && n_req <= 65535, GSP_ERR_ARGUMENT,
              "number of requests must be 1..65535");
  hipStream_t s = ctx->stream;
  PhaseTrace tr(s);
  std::vector<HostReq> H((size_t)n_req);
  const int32_t precision = problems[0].precision;
  const int32_t devp = problems[0].device_pointers ? 1 : 0;
  int64_t NT = 0, nDT = 0, nC = 0;
  int32_t K = 1;
  bool ring = false;     // some request searches without a distance bound
  bool no_deps = true;   // every request searches data only: every node is level 0
  for (int32_t q = 0; q < n_req; ++q) {
    const int32_t mode = problems[q].search_mode;
    ring = ring || mode == GSP_DATA_ONLY || (mode == GSP_MASKED_KNN && !(problems[q].radius > 0));
    no_deps = no_deps && mode == GSP_DATA_ONLY;
    for_request(n_req, q, [&] {
      const gsp_seqsim_problem* p = &problems[q];
      validate_problem(p, H[(size_t)q]);
      GSP_REQUIRE(p->precision == precision, GSP_ERR_ARGUMENT,
                  "all requests of a batch must use one precision");
      GSP_REQUIRE((p->device_pointers ? 1 : 0) == devp, GSP_ERR_ARGUMENT,
                  "all requests of a batch must pass the same kind of pointers");
    });
    const HostReq& h = H[(size_t)q];
    NT += h.N;
    nDT += h.nD;
    nC += h.nD * h.nd;
    K = std::max(K, h.kmax);
    GSP_REQUIRE(NT + nDT < ((int64_t)1 << 30), GSP_ERR_UNSUPPORTED,
                "batch too large (nodes plus data must stay below 2^30)");
  }

  // per-request offsets of nodes, data and packed coordinates
  std::vector<int64_t> node_off((size_t)n_req + 1, 0), data_off((size_t)n_req + 1, 0),
      coord_off((size_t)n_req + 1, 0);
  for (int32_t q = 0; q < n_req; ++q) {
    const HostReq& h = H[(size_t)q];
    node_off[(size_t)q + 1] = node_off[(size_t)q] + h.N;
    data_off[(size_t)q + 1] = data_off[(size_t)q] + h.nD;
    coord_off[(size_t)q + 1] = coord_off[(size_t)q] + h.nD * h.nd;
  }

  ctx->counters.reserve(16);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ctx->counters.p, 0, (int64_t)16);
  DevBuf<int64_t> path, dnode_off;
  path.alloc((size_t)NT);
  dnode_off.alloc((size_t)n_req + 1);
  GSP_HIP(hipMemcpyAsync(dnode_off.p, node_off.data(), node_off.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));

  // conditioning coordinates on the host (the hash is built there)
  std::vector<double> coords((size_t)std::max<int64_t>(nC, 1));
  if (devp) {
    // one gather launch per array over all requests instead of a copy each
    std::vector<Span8> spans((size_t)n_req);
    DevBuf<Span8> dspans;
    dspans.alloc((size_t)n_req);
    const int gx = grid_for(NT / n_req + 1, 256, 1024);
    for (int32_t q = 0; q < n_req; ++q)
      spans[(size_t)q] = Span8{(const uint64_t*)problems[q].path, node_off[(size_t)q],
                               H[(size_t)q].N};
    GSP_HIP(hipMemcpyAsync(dspans.p, spans.data(), spans.size() * sizeof(Span8),
                           hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather8, dim3(gx, n_req), dim3(256), 0, s, dspans.p,
                       (uint64_t*)path.p);
    if (nC > 0) {
      std::vector<Span8> cs((size_t)n_req);
      DevBuf<Span8> cspans;
      cspans.alloc((size_t)n_req);
      DevBuf<double> dcoords;
      dcoords.alloc((size_t)nC);
      for (int32_t q = 0; q < n_req; ++q)
        cs[(size_t)q] = Span8{(const uint64_t*)problems[q].data_coords, coord_off[(size_t)q],
                              H[(size_t)q].nD * H[(size_t)q].nd};
      GSP_HIP(hipMemcpyAsync(cspans.p, cs.data(), cs.size() * sizeof(Span8),
                             hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_gather8, dim3(grid_for(nC / n_req + 1, 256, 64), n_req), dim3(256), 0,
                         s, cspans.p, (uint64_t*)dcoords.p);
      GSP_HIP(hipMemcpyAsync(coords.data(), dcoords.p, (size_t)nC * sizeof(double),
                             hipMemcpyDeviceToHost, s));
    }
    GSP_HIP(hipStreamSynchronize(s));  // the span tables leave scope
  } else {
    if (n_req == 1) {
      GSP_HIP(hipMemcpyAsync(path.p, problems[0].path, (size_t)NT * sizeof(int64_t),
                             hipMemcpyHostToDevice, s));
    } else {
      std::vector<int64_t> hp((size_t)NT);
      for (int32_t q = 0; q < n_req; ++q)
        std::memcpy(hp.data() + node_off[(size_t)q], problems[q].path,
                    (size_t)H[(size_t)q].N * sizeof(int64_t));
      GSP_HIP(hipMemcpyAsync(path.p, hp.data(), (size_t)NT * sizeof(int64_t),
                             hipMemcpyHostToDevice, s));
      GSP_HIP(hipStreamSynchronize(s));  // hp leaves scope
    }
    for (int32_t q = 0; q < n_req; ++q)
      if (H[(size_t)q].nD > 0)
        std::memcpy(coords.data() + coord_off[(size_t)q], problems[q].data_coords,
                    (size_t)(H[(size_t)q].nD * H[(size_t)q].nd) * sizeof(double));
  }
  tr.mark("gather inputs");

  // host preparation: scale factors, offset tables, data hashes
  parallel_requests(n_req, [&](int32_t q) {
    HostReq& h = H[(size_t)q];
    const gsp_seqsim_problem* p = &problems[q];
    h.data = coords.data() + coord_off[(size_t)q];
    for (int64_t t = 0; t < h.nD * h.nd; ++t)
      GSP_REQUIRE(std::isfinite(h.data[t]), GSP_ERR_ARGUMENT, "non-finite data coordinate");
    // data-only plans stand for GeoStatsModels.fitpredict (posterior mean,
    // FFTSIM conditioning), which krige in the domain's own coordinates
    h.alpha = p->search_mode == GSP_DATA_ONLY ? 1.0
                                              : scale_factor(p->grid, h.data, h.nD * h.nd,
                                                             p->model);
    h.table = build_offsets(p->grid, h.r2, h.kmax, p->search_mode);
    build_hash(p->grid, p->search_mode, h);
  });
  std::vector<int64_t> table_off((size_t)n_req + 1, 0), cell_off((size_t)n_req + 1, 0),
      item_off((size_t)n_req + 1, 0), tile_off((size_t)n_req + 1, 0);
  for (int32_t q = 0; q < n_req; ++q) {
    table_off[(size_t)q + 1] = table_off[(size_t)q] + (int64_t)H[(size_t)q].table.size();
    cell_off[(size_t)q + 1] = cell_off[(size_t)q] + (int64_t)H[(size_t)q].hstart.size();
    item_off[(size_t)q + 1] = item_off[(size_t)q] + (int64_t)H[(size_t)q].hitems.size();
    int ts[3];
    tile_shape(H[(size_t)q].nd, ts);
    int64_t tiles = 1;
    for (int d = 0; d < H[(size_t)q].nd; ++d) tiles *= (problems[q].grid.dims[d] + ts[d] - 1) / ts[d];
    tile_off[(size_t)q + 1] = tile_off[(size_t)q] + tiles;
  }
  const int64_t n_table = table_off.back(), n_cells = cell_off.back(), n_items = item_off.back();
  GSP_REQUIRE(n_items < ((int64_t)1 << 31), GSP_ERR_UNSUPPORTED, "data hash too large");
  // the data-only tile search: every request searches data only (no hash
  // cell lists expanded), kmax <= 16; GSP_TILE_SEARCH=0 keeps the wavefront
  // search
  bool tile_search = no_deps && K <= 16 && tile_off.back() < ((int64_t)1 << 31);
  for (int32_t q = 0; q < n_req; ++q) tile_search = tile_search && H[(size_t)q].hexp == 0;
  if (const char* env = std::getenv("GSP_TILE_SEARCH")) tile_search = tile_search && std::atoi(env) != 0;
  // one request, no data, a bounded search: the grid-only search by lanes
  bool grid_lane = !no_deps && !ring && n_req == 1 && nDT == 0 &&
                   problems[0].search_mode != GSP_DATA_ONLY;
  if (const char* env = std::getenv("GSP_GRID_LANE")) grid_lane = grid_lane && std::atoi(env) != 0;

  // concatenated device tables
  std::vector<Req> reqs((size_t)n_req);
  std::vector<Offset> table((size_t)std::max<int64_t>(n_table, 1));
  std::vector<int32_t> hstart((size_t)n_cells);
  std::vector<int32_t> hitems((size_t)std::max<int64_t>(n_items, 1));
  std::vector<double> data3((size_t)std::max<int64_t>(nDT * 3, 1), 0.0);
  std::vector<double> means((size_t)n_req);
  std::vector<ReqInfo> info((size_t)n_req);
  {
    parallel_requests(n_req, [&](int32_t q) {
      const int64_t t_off = table_off[(size_t)q], c_off = cell_off[(size_t)q],
                    i_off = item_off[(size_t)q];
      const HostReq& h = H[(size_t)q];
      const gsp_seqsim_problem* p = &problems[q];
      const gsp_grid& g = p->grid;
      Req& R = reqs[(size_t)q];
      std::memset(&R, 0, sizeof(R));
      R.ndim = h.nd;
      R.mode = p->search_mode;
      R.kmin = h.kmin;
      R.kmax = h.kmax;
      R.T = (int32_t)h.table.size();
      R.hexp = h.hexp;
      R.hB = (int32_t)h.hB;
      R.n_data = (int32_t)h.nD;
      R.node_off = (int32_t)node_off[(size_t)q];
      R.data_off = (int32_t)data_off[(size_t)q];
      R.table_off = t_off;
      R.cell_off = c_off;
      R.tile_off = (int32_t)tile_off[(size_t)q];
      R.r2 = h.r2;
      {
        // first staging radius of the tile search: 1.5 times the radius that
        // holds kmax data at the request's mean data density (a tile whose
        // nodes are not all settled by it visits again with twice the radius)
        double vol = 1.0, smin = g.spacing[0];
        for (int d = 0; d < h.nd; ++d) {
          vol *= (double)g.dims[d] * g.spacing[d];
          smin = std::min(smin, g.spacing[d]);
        }
        const double pi = 3.141592653589793;
        const double unit = h.nd == 1 ? 2.0 : (h.nd == 2 ? pi : 4.0 * pi / 3.0);
        const double per = vol / (double)std::max<int64_t>(h.nD, 1);
        double rho = 1.5 * std::pow((double)h.kmax * per / unit, 1.0 / h.nd);
        if (h.r2 < std::numeric_limits<double>::infinity())
          rho = std::min(rho, std::sqrt(h.r2) * 1.001);
        if (!std::isfinite(rho)) rho = vol;
        if (!(rho >= 0.5 * smin)) rho = 0.5 * smin;
        R.trho = rho;
      }
      R.alpha = h.alpha;
      R.prior_sigma = std::sqrt(p->model.sill);
      R.mean = p->mean;
      for (int d = 0; d < 3; ++d) {
        const bool in = d < h.nd;
        R.dims[d] = in ? (int32_t)g.dims[d] : 1;
        R.org[d] = in ? g.origin[d] : 0.0;
        R.sp[d] = in ? g.spacing[d] : 1.0;
        R.ao[d] = in ? h.alpha * g.origin[d] : 0.0;
        R.as[d] = in ? h.alpha * g.spacing[d] : 0.0;
        R.hd[d] = h.hd[d];
        R.hcs[d] = h.hcs[d];
      }
      R.M = make_dev_model(p->model, h.nd, h.alpha);
      ReqInfo& I = info[(size_t)q];
      I.model = p->model;
      I.ndim = h.nd;
      I.mode = p->search_mode;
      I.kmin = h.kmin;
      I.kmax = h.kmax;
      I.T = R.T;
      for (int d = 0; d < 3; ++d) I.spacing[d] = d < h.nd ? g.spacing[d] : 1.0;
      for (int t = 0; t < kSurrogateSlots && t < R.T; ++t) {
        I.table32[t][0] = h.table[(size_t)t].dx;
        I.table32[t][1] = h.table[(size_t)t].dy;
        I.table32[t][2] = h.table[(size_t)t].dz;
      }
      R.learned = 0;
      for (int32_t si = 0; learned_plan && si < n_surrogates; ++si)
        if (surrogate_accepts(surrogates[si], I)) {
          R.learned = si + 1;
          break;
        }
      means[(size_t)q] = p->mean;
      std::copy(h.table.begin(), h.table.end(), table.begin() + t_off);
      for (size_t c = 0; c < h.hstart.size(); ++c)
        hstart[(size_t)c_off + c] = (int32_t)(i_off + h.hstart[c]);
      for (size_t t = 0; t < h.hitems.size(); ++t)
        hitems[(size_t)i_off + t] = h.hitems[t] + R.data_off;
      for (int64_t i = 0; i < h.nD; ++i)
        for (int d = 0; d < h.nd; ++d)
          data3[(size_t)((data_off[(size_t)q] + i) * 3 + d)] = h.data[(size_t)(i * h.nd + d)];
    });
  }
  tr.mark("host tables");

  plan->N = NT;
  plan->n_data = nDT;
  plan->n_req = n_req;
  plan->ndim = H[0].nd;
  plan->kmin = H[0].kmin;
  for (const HostReq& h : H) plan->kmin = std::min(plan->kmin, h.kmin);
  plan->kmax = K;
  plan->alpha = H[0].alpha;
  plan->mean = problems[0].mean;
  plan->node_off = node_off;
  plan->data_off = data_off;
  plan->info = std::move(info);
  plan->nfound.alloc((size_t)NT);
  plan->nbr.alloc((size_t)(NT * K));
  plan->lam.alloc((size_t)(NT * K));
  plan->sigma.alloc((size_t)NT);
  plan->flags.alloc((size_t)NT);
  plan->level.alloc((size_t)NT);
  plan->order.alloc((size_t)NT);
  plan->pos.alloc((size_t)NT);
  plan->node_req.alloc((size_t)NT);
  plan->req_mean.alloc((size_t)n_req);
  DevBuf<int32_t>& pos = plan->pos;
  DevBuf<Req>& dreqs = plan->reqs;
  dreqs.alloc((size_t)n_req);
  DevBuf<Offset>& dtable = plan->table;
  dtable.alloc(table.size());
  DevBuf<double>& ddata = plan->data_xyz;  // kept: surrogate training reads the data offsets
  ddata.alloc(data3.size());
  DevBuf<int32_t> dhstart;
  DevBuf<int32_t> dhitems;
  dhstart.alloc(hstart.size());
  dhitems.alloc(hitems.size());
  GSP_HIP(hipMemcpyAsync(dreqs.p, reqs.data(), reqs.size() * sizeof(Req), hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(plan->req_mean.p, means.data(), means.size() * sizeof(double),
                         hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dtable.p, table.data(), table.size() * sizeof(Offset),
                         hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(ddata.p, data3.data(), data3.size() * sizeof(double),
                         hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dhstart.p, hstart.data(), hstart.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, s));
  GSP_HIP(hipMemcpyAsync(dhitems.p, hitems.data(), hitems.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, s));
  const int gn = grid_for(NT, 256, 4096);
  hipLaunchKernelGGL(k_batch_path, dim3(grid_for(NT / n_req + 1, 256, 1024), n_req), dim3(256), 0,
                     s, path.p, dnode_off.p, plan->node_req.p, ctx->counters.p);
  hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(256), 0, s, pos.p, -1, NT);
  hipLaunchKernelGGL(k_fill_i32, dim3(gn), dim3(256), 0, s, plan->level.p, -1, NT);
  hipLaunchKernelGGL(k_positions, dim3(gn), dim3(256), 0, s, path.p, NT, pos.p, ctx->counters.p);
  hipLaunchKernelGGL(k_positions_check, dim3(gn), dim3(256), 0, s, path.p, NT, pos.p,
                     ctx->counters.p);
  GSP_HIP(hipGetLastError());
  int32_t pos_err = 0;
  GSP_HIP(hipMemcpyAsync(&pos_err, ctx->counters.p + CTR_POS_ERR, sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  GSP_HIP(hipStreamSynchronize(s));  // also ends the host tables' uploads
  GSP_REQUIRE(pos_err == 0, GSP_ERR_UNSUPPORTED,
              "the path must be a permutation of the grid ids 0..N-1");

  DevBuf<int32_t> outdeg, indeg;
  if (!no_deps) {
    outdeg.alloc((size_t)NT + 1);
    indeg.alloc((size_t)NT);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(NT + 1, 256, 4096)), dim3(256), 0, s, outdeg.p,
                       0, NT + 1);
  }
  tr.mark("upload+positions");
  PlanArgs A;
  std::memset(&A, 0, sizeof(A));
  A.NT = (int32_t)NT;
  A.K = K;
  A.reqs = dreqs.p;
  A.node_req = plan->node_req.p;
  A.path = path.p;
  A.pos = pos.p;
  A.table = dtable.p;
  A.data = ddata.p;
  A.hstart = dhstart.p;
  A.hitems = dhitems.p;
  A.nfound = plan->nfound.p;
  A.nbr = plan->nbr.p;
  DevBuf<uint8_t> nbr_slot;
  if (learned_plan) nbr_slot.alloc((size_t)(NT * K));
  A.nbr_slot = learned_plan ? nbr_slot.p : nullptr;
  A.lam = plan->lam.p;
  A.sigma = plan->sigma.p;
  A.flags = plan->flags.p;
  A.counters = ctx->counters.p;
  // The reverse edges of the levels are sized by the bound sum(nfound) <=
  // NT * K when that fits 32-bit offsets: then the search counts the
  // dependency degrees itself and the main stream waits for nothing else
  // (else the exact count runs after the search, here).
  LevelBufs LB;
  const bool side_count = (int64_t)NT * K < ((int64_t)1 << 31);
  const bool search_counts = !no_deps && side_count;
  if (search_counts) {
    level_bufs_alloc(LB, NT, (int64_t)NT * K, outdeg.p, s);  // every buffer the worker uses
    A.outdeg = outdeg.p;
    A.indeg = indeg.p;
    A.mark = LB.mark.p;
  }

  {
    TimedLaunch tl(ctx, GSP_K_SEARCH);
    if (tile_search) {
      const dim3 tg((unsigned)tile_off.back());
      if (K <= 8)
        hipLaunchKernelGGL(k_search_tile<8>, tg, dim3(kTileNodes), 0, s, A, n_req);
      else if (K <= 12)
        hipLaunchKernelGGL(k_search_tile<12>, tg, dim3(kTileNodes), 0, s, A, n_req);
      else
        hipLaunchKernelGGL(k_search_tile<16>, tg, dim3(kTileNodes), 0, s, A, n_req);
    } else if (grid_lane) {
      const dim3 lg((unsigned)((NT + 255) / 256));
      if (K <= 32)
        hipLaunchKernelGGL(k_search_grid_lane<true>, lg, dim3(256), (size_t)256 * K * 5, s, A);
      else
        hipLaunchKernelGGL(k_search_grid_lane<false>, lg, dim3(256), 0, s, A);
    } else if (ring)
      hipLaunchKernelGGL(k_search<true>, dim3(grid_for(NT, kWaves, 16384)), dim3(256), 0, s, A);
    else
      hipLaunchKernelGGL(k_search<false>, dim3(grid_for(NT, kWaves, 16384)), dim3(256), 0, s, A);
    GSP_HIP(hipGetLastError());
  }
  tr.mark("search kernel");

  // dependency levels from the search, on the side stream in a worker thread
  // while this thread runs the solves on the main stream (the level kernels
  // are latency-bound, the solves compute-bound; they touch disjoint arrays)
  DepArgs D;
  D.N = (int32_t)NT;
  D.K = K;
  D.reqs = dreqs.p;
  D.node_req = plan->node_req.p;
  D.nfound = plan->nfound.p;
  D.flags = plan->flags.p;
  D.nbr = plan->nbr.p;
  DevBuf<unsigned long long> edges;
  if (!no_deps) {
    edges.alloc(1);
    GSP_HIP(hipMemsetAsync(edges.p, 0, 8, s));
  }
  if (no_deps || search_counts) {
    // no levels to build / degrees counted by the search
  } else {
    mark_alloc(LB, NT, s);
    hipLaunchKernelGGL(k_deg_count<false>, dim3(grid_for(NT, 16, 4096)), dim3(256), 0, s, D,
                       outdeg.p, indeg.p, LB.mark.p, edges.p);
    unsigned long long n_edges = 0;
    GSP_HIP(hipMemcpyAsync(&n_edges, edges.p, 8, hipMemcpyDeviceToHost, s));
    GSP_HIP(hipStreamSynchronize(s));
    GSP_REQUIRE(n_edges < (1ull << 31), GSP_ERR_UNSUPPORTED,
                "too many dependencies (%llu) for the level computation", n_edges);
    level_bufs_alloc(LB, NT, (int64_t)n_edges, outdeg.p, s);
  }
  GSP_HIP(hipEventRecord(ctx->fork, s));
  LevelJob job;
  if (ctx->timing && !no_deps) {
    job.a = take_event(ctx);
    job.b = take_event(ctx);
  }
  {
    std::unique_ptr<WorkerThread> worker;
    if (!no_deps)
      worker = std::make_unique<WorkerThread>([&] {
        levels_on_side(ctx, plan, D, outdeg.p, indeg.p, LB, job);
      });
  {
    // learned plans: the surrogate takes the nodes it can, the rest are listed
    // for the exact solver
    DevBuf<int32_t> exact_ids;
    int64_t n_exact = NT;
    if (learned_plan) {
      // each request's input row and slot covariances for the surrogate it uses
      std::vector<float> feat((size_t)n_req * kSurrogateInputs, 0.0f),
          cslot((size_t)n_req * kSurrogateSlots, 0.0f);
      std::vector<uint8_t> used((size_t)n_surrogates, 0);
      for (int32_t q = 0; q < n_req; ++q) {
        const int32_t l = reqs[(size_t)q].learned;
        if (l <= 0) continue;
        used[(size_t)(l - 1)] = 1;
        surrogate_request_rows(surrogates[l - 1], plan->info[(size_t)q],
                               feat.data() + (size_t)q * kSurrogateInputs,
                               cslot.data() + (size_t)q * kSurrogateSlots);
      }
      DevBuf<float> dfeat, dcslot;
      dfeat.alloc(feat.size());
      dcslot.alloc(cslot.size());
      GSP_HIP(hipMemcpyAsync(dfeat.p, feat.data(), feat.size() * 4, hipMemcpyHostToDevice, s));
      GSP_HIP(hipMemcpyAsync(dcslot.p, cslot.data(), cslot.size() * 4, hipMemcpyHostToDevice, s));
      exact_ids.alloc((size_t)NT);
      NodeOut O;
      O.NT = (int32_t)NT;
      O.K = K;
      O.reqs = dreqs.p;
      O.node_req = plan->node_req.p;
      O.table = dtable.p;
      O.nfound = plan->nfound.p;
      O.nbr = plan->nbr.p;
      O.nbr_slot = nbr_slot.p;
      O.data = ddata.p;
      O.lam = plan->lam.p;
      O.sigma = plan->sigma.p;
      O.flags = plan->flags.p;
      O.counters = ctx->counters.p;
      {
        TimedLaunch tl(ctx, GSP_K_LEARNED);
        // one pass per surrogate in use; the first pass also lists the nodes
        // of requests no surrogate takes
        bool listed = false;
        for (int32_t si = 0; si < n_surrogates; ++si) {
          if (!used[(size_t)si] && (listed || si + 1 < n_surrogates)) continue;
          surrogate_infer(ctx, surrogates[si], si + 1, !listed, O, tolerance, exact_ids.p, dfeat.p,
                          dcslot.p);
          listed = true;
        }
      }
      int32_t ne = 0;
      GSP_HIP(hipMemcpyAsync(&ne, ctx->counters.p + CTR_EXACT, 4, hipMemcpyDeviceToHost, s));
      GSP_HIP(hipStreamSynchronize(s));
      n_exact = ne;
      tr.mark("learned weights");
    }
    const int32_t* ids = learned_plan ? exact_ids.p : nullptr;
    // bucket the nodes by neighbour count: n <= 16 run four per wavefront,
    // larger systems in the smallest register tile that holds them
    DevBuf<int32_t> nhist, list;
    nhist.alloc(65);
    list.alloc((size_t)std::max<int64_t>(n_exact, 1));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, nhist.p, 0, (int64_t)65);
    hipLaunchKernelGGL(k_n_hist, dim3(grid_for(n_exact, 256, 1024)), dim3(256), 0, s,
                       plan->nfound.p, ids, (int32_t)n_exact, nhist.p);
    int32_t hh[65];
    GSP_HIP(hipMemcpyAsync(hh, nhist.p, sizeof(hh), hipMemcpyDeviceToHost, s));
    GSP_HIP(hipStreamSynchronize(s));
    int32_t cur[65];
    int64_t upto[65];  // nodes with at most k neighbours
    int64_t acc = 0;
    for (int k = 0; k <= 64; ++k) {
      cur[k] = (int32_t)acc;
      acc += hh[k];
      upto[k] = acc;
    }
    GSP_HIP(hipMemcpyAsync(nhist.p, cur, sizeof(cur), hipMemcpyHostToDevice, s));
    if (n_exact > 0)
      hipLaunchKernelGGL(k_n_scatter,
                         dim3((unsigned)((n_exact + kScatterChunk - 1) / kScatterChunk)), dim3(256),
                         0, s, plan->nfound.p, ids, (int32_t)n_exact, nhist.p, list.p);
    tr.mark("bucket by size");
    TimedLaunch tl(ctx, GSP_K_PLAN);
    // n <= kLaneNB32 / kLaneNB64: one system per lane (when there are enough
    // of them); the rest up to n = 16: four per wavefront
    int64_t lane_min = kLaneMin;
    if (const char* env = std::getenv("GSP_LANE_MIN")) lane_min = std::atoll(env);
    const int lane_nb = precision == GSP_FP32_REFINE ? kLaneNB32 : kLaneNB64;
    const int64_t tiny = upto[lane_nb] >= lane_min ? upto[lane_nb] : 0;
    if (tiny > 0) {
      const int blocks = grid_for(tiny, 256, 65536);
      if (precision == GSP_FP32_REFINE)
        hipLaunchKernelGGL((k_krige_lane<kLaneNB32, GSP_FP32_REFINE>), dim3(blocks), dim3(256), 0, s,
                           A, list.p, (int32_t)tiny);
      else
        hipLaunchKernelGGL((k_krige_lane<kLaneNB64, GSP_FP64>), dim3(blocks), dim3(256), 0, s, A,
                           list.p, (int32_t)tiny);
    }
    const int64_t small = upto[16] - tiny;
    if (small > 0) {
      const int blocks = grid_for((small + 3) / 4, k16Waves, 65536);
      if (precision == GSP_FP32_REFINE)
        hipLaunchKernelGGL(k_krige16<GSP_FP32_REFINE>, dim3(blocks), dim3(64 * k16Waves), 0, s, A,
                           list.p + tiny, (int32_t)small);
      else
        hipLaunchKernelGGL(k_krige16<GSP_FP64>, dim3(blocks), dim3(64 * k16Waves), 0, s, A,
                           list.p + tiny, (int32_t)small);
    }
    launch_mid<24>(ctx, A, precision, list.p + upto[16], upto[24] - upto[16], s);
    launch_mid<32>(ctx, A, precision, list.p + upto[24], upto[32] - upto[24], s);
    launch_krige_wide(48, A, precision, list.p + upto[32], upto[48] - upto[32], s);
    launch_krige_wide(64, A, precision, list.p + upto[48], upto[64] - upto[48], s);
    GSP_HIP(hipGetLastError());
  }
  tr.mark("kriging kernel");
  }  // joins the worker
  if (job.a) record_timing(ctx, job.a, job.b, GSP_K_LEVELS);
  GSP_REQUIRE(job.code == GSP_OK, job.code, "%s", job.msg.c_str());
  // one read of the counters serves the rebuild decision and the plan's
  // statistics (the rebuild does not touch them)
  int32_t ctr[16];
  GSP_HIP(hipMemcpyAsync(ctr, ctx->counters.p, sizeof(ctr), hipMemcpyDeviceToHost, s));
  GSP_HIP(hipStreamSynchronize(s));
  const int32_t n_fail = ctr[CTR_FAIL];
  if (no_deps) {
    // one level: every node, in id order
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(NT, 256, 4096)), dim3(256), 0, s,
                       plan->level.p, 0, NT);
    hipLaunchKernelGGL(k_iota_i32, dim3(grid_for(NT, 256, 4096)), dim3(256), 0, s,
                       plan->order.p, NT);
    plan->depth = 1;
    plan->level_off.assign({0, NT});
  } else if (n_fail > 0) {
    // failed systems draw from the prior: their grid neighbours are no
    // dependencies, so the levels are built again from the flags
    TimedLaunch tl(ctx, GSP_K_LEVELS);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(NT + 1, 256, 4096)), dim3(256), 0, s, outdeg.p,
                       0, NT + 1);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(NT, 256, 4096)), dim3(256), 0, s,
                       plan->level.p, -1, NT);
    hipLaunchKernelGGL(k_deg_count<true>, dim3(grid_for(NT, 16, 4096)), dim3(256), 0, s, D,
                       outdeg.p, indeg.p, LB.mark.p, edges.p);
    build_levels<true>(plan, D, outdeg.p, indeg.p, LB, s);
  }
  tr.mark("levels");
  flush_timings(ctx);
  plan->n_prior = ctr[CTR_PRIOR];
  plan->n_fallback = ctr[CTR_FALLBACK];
  plan->n_learned = ctr[CTR_LEARNED];
}

// Realizations per sweep pass (GSP_SWEEP_CHUNK overrides). Smaller passes keep
// the node-major state in the Infinity Cache, but each pass pays one launch
// per level; on the 1M-node 3D workload 64 realizations per pass measured
// fastest (2.6 ms vs 3.1 / 3.9 / 7.4 ms for 32 / 16 / 8), so the cap is off.
constexpr int64_t kSweepCacheBytes = (int64_t)1 << 40;

int sweep_chunk(int64_t rows, int32_t R) {
  (void)R;
  if (const char* env = std::getenv("GSP_SWEEP_CHUNK")) {
    const int v = std::atoi(env);
    if (v >= 1 && v <= 64) return v;
  }
  int rc = 64;
  while (rc > 8 && rows * rc * 8 > kSweepCacheBytes) rc >>= 1;
  return rc;
}

// Sweeps with at least this many levels are captured into a HIP graph and
// replayed (GSP_SWEEP_GRAPH=0 turns it off): one graph launch instead of one
// kernel launch per level, transposes included.
constexpr int32_t kGraphMinLevels = 8;

bool sweep_graphs_on() {
  const char* env = std::getenv("GSP_SWEEP_GRAPH");
  return !(env && env[0] == '0');
}

void seqsim_sweep_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                      const double* data_vals, double* out) {
  const int64_t N = plan->N, nD = plan->n_data;
  const int RCmax = sweep_chunk(N + nD, R);
  const int RC0 = R < RCmax ? R : RCmax;
  ctx->X.reserve((size_t)((N + nD) * RC0));
  ctx->Zt.reserve((size_t)(N * RC0));
  SweepScratch w;
  w.X = ctx->X.p;
  w.Zt = ctx->Zt.p;
  w.RCmax = RCmax;
  w.timed = true;
  // with timing events on, launches stay individual (they are bracketed)
  if (ctx->timing || plan->depth < kGraphMinLevels || !sweep_graphs_on()) {
    seqsim_sweep_into(ctx, plan, R, z, data_vals, out, w);
    return;
  }
  gsp_seqsim_plan::SweepKey key;
  key.R = R;
  key.RC = RCmax;
  key.z = z;
  key.vals = data_vals;
  key.out = out;
  key.X = w.X;
  key.Zt = w.Zt;
  hipStream_t s = ctx->stream;
  if (!plan->sweep_graph || !(plan->sweep_key == key) || plan->sweep_stream != s) {
    plan->drop_sweep_graph();
    w.timed = false;
    hipGraph_t graph = nullptr;
    GSP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
      seqsim_sweep_into(ctx, plan, R, z, data_vals, out, w);
    } catch (...) {
      if (hipStreamEndCapture(s, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
      throw;
    }
    GSP_HIP(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GSP_HIP(e);
    plan->sweep_graph = exec;
    plan->sweep_stream = s;
    plan->sweep_key = key;
  }
  GSP_HIP(hipGraphLaunch(plan->sweep_graph, s));
}

void seqsim_sweep_into(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R, const double* z,
                       const double* data_vals, double* out, const SweepScratch& W) {
  hipStream_t s = ctx->stream;
  gsp_ctx* tctx = W.timed ? ctx : nullptr;  // no event bracketing when untimed
  const int64_t N = plan->N, nD = plan->n_data;
  const int RCmax = W.RCmax;
  double* const Xs = W.X;
  double* const Zts = W.Zt;
  for (int r0 = 0; r0 < R; r0 += RCmax) {
    const int RC = (R - r0) < RCmax ? (R - r0) : RCmax;
    int shift = 0;
    while ((1 << shift) < RC) ++shift;
    if (nD > 0)
      hipLaunchKernelGGL(k_data_rows, dim3(grid_for(nD * RC, 256, 1024)), dim3(256), 0, s,
                         Xs, (int32_t)N, (int32_t)nD, RC, data_vals);
    {
      TimedLaunch tl(tctx, GSP_K_TRANSPOSE);
      hipLaunchKernelGGL(k_z_transpose, dim3((unsigned)((N + 63) / 64)), dim3(256),
                         transpose_lds(RC), s, z,
                         (int32_t)N, RC, r0, Zts);
    }
    LevelArgs A;
    A.K = plan->kmax;
    A.RC = RC;
    A.shift = shift;
    A.pos = plan->pos.p;
    A.nfound = plan->nfound.p;
    A.flags = plan->flags.p;
    A.nbr = plan->nbr.p;
    A.lam = plan->lam.p;
    A.sigma = plan->sigma.p;
    A.node_req = plan->node_req.p;
    A.mean = W.mean ? W.mean : plan->req_mean.p;
    A.Zt = Zts;
    A.X = Xs;
    {
      TimedLaunch tl(tctx, GSP_K_SWEEP);
      for (int32_t l = 0; l < plan->depth; ++l) {
        const int64_t off = plan->level_off[(size_t)l];
        const int64_t cnt = plan->level_off[(size_t)l + 1] - off;
        if (cnt == 0) continue;
        A.order = plan->order.p + off;
        A.count = (int32_t)cnt;
        const int64_t threads = cnt << shift;
        const int64_t per = ((threads + 255) / 256 + kXcds - 1) / kXcds;
        hipLaunchKernelGGL(k_sweep_level, dim3((unsigned)(per * kXcds)), dim3(256), 0, s, A);
      }
    }
    {
      TimedLaunch tl(tctx, GSP_K_TRANSPOSE);
      hipLaunchKernelGGL(k_transpose<double>, dim3((unsigned)((N + 63) / 64)), dim3(256),
                         transpose_lds(RC), s, Xs,
                         (int32_t)N, RC, r0, out);
    }
    GSP_HIP(hipGetLastError());
  }
}


// Cholesky factor of the nv x nv coefficient matrix (row-major), ascending
// sums: L_jj = sqrt(A_jj - sum_k L_jk^2), L_ij = (A_ij - sum_k L_ik L_jk) / L_jj.
MvMix mv_factor(const double* A, int nv) {
  MvMix M;
  std::memset(&M, 0, sizeof(M));
  for (int j = 0; j < nv; ++j) {
    for (int i = j; i < nv; ++i) {
      double s = A[i * nv + j];
      for (int k = 0; k < j; ++k) s = s - M.L[i][k] * M.L[j][k];
      if (i == j) {
        GSP_REQUIRE(s > 0.0 && std::isfinite(s), GSP_ERR_POSDEF,
                    "PosDefException: the coregionalization matrix is not positive definite");
        M.L[j][j] = std::sqrt(s);
      } else {
        M.L[i][j] = s / M.L[j][j];
      }
    }
  }
  return M;
}

void seqsim_sweep_multivariate_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                   int32_t nv, const double* A, const double* mean,
                                   const double* z, const double* data_vals, double* out) {
  GSP_REQUIRE(nv >= 1 && nv <= 4, GSP_ERR_UNSUPPORTED,
              "multivariate SEQSIM supports 1 to 4 variables");
  GSP_REQUIRE(plan->n_req == 1, GSP_ERR_UNSUPPORTED,
              "multivariate SEQSIM runs plans of one request");
  for (int i = 0; i < nv; ++i) {
    GSP_REQUIRE(std::isfinite(mean[i]), GSP_ERR_ARGUMENT, "invalid mean");
    for (int j = 0; j < nv; ++j)
      GSP_REQUIRE(A[i * nv + j] == A[j * nv + i], GSP_ERR_ARGUMENT,
                  "the coregionalization matrix must be symmetric");
  }
  const MvMix M = mv_factor(A, nv);
  hipStream_t s = ctx->stream;
  const int64_t N = plan->N, nD = plan->n_data, total = (int64_t)R * N;
  DevBuf<double> w, dmean;
  w.alloc((size_t)std::max<int64_t>(total * nv, 1));
  dmean.alloc((size_t)nv);
  GSP_HIP(hipMemcpyAsync(dmean.p, mean, (size_t)nv * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mv_mix, dim3(grid_for(total, 256, 8192)), dim3(256), 0, s, z, N, R, nv, M,
                     w.p);
  GSP_HIP(hipGetLastError());
  const int RCmax = sweep_chunk(N + nD, R);
  const int RC0 = R < RCmax ? R : RCmax;
  ctx->X.reserve((size_t)((N + nD) * RC0));
  ctx->Zt.reserve((size_t)(N * RC0));
  for (int v = 0; v < nv; ++v) {
    SweepScratch sw;
    sw.X = ctx->X.p;
    sw.Zt = ctx->Zt.p;
    sw.RCmax = RCmax;
    sw.timed = true;
    sw.mean = dmean.p + v;
    seqsim_sweep_into(ctx, plan, R, w.p + (int64_t)v * total,
                      nD > 0 ? data_vals + (int64_t)v * nD : nullptr, out + (int64_t)v * total, sw);
  }
  GSP_HIP(hipStreamSynchronize(s));  // w and the means are freed on return
}

void seqsim_sweep_indicator_run(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                int32_t n_cat, const double* prob, const double* u,
                                const int32_t* data_cat, int32_t* out) {
  GSP_REQUIRE(n_cat >= 2 && n_cat <= kMaxCat, GSP_ERR_UNSUPPORTED,
              "indicator simulation supports 2 to %d categories", kMaxCat);
  hipStream_t s = ctx->stream;
  const int64_t N = plan->N, nD = plan->n_data;
  const int RCmax = sweep_chunk(N + nD, R);
  const int RC0 = R < RCmax ? R : RCmax;
  ctx->X.reserve((size_t)((N + nD) * RC0));
  ctx->Zt.reserve((size_t)(N * RC0));
  DevBuf<double> dcat;
  if (nD > 0) {
    dcat.alloc((size_t)nD);
    hipLaunchKernelGGL(k_i32_to_f64, dim3(grid_for(nD, 256, 1024)), dim3(256), 0, s, data_cat,
                       dcat.p, nD);
  }
  IndicatorArgs I;
  std::memset(&I, 0, sizeof(I));
  I.n_cat = n_cat;
  for (int c = 0; c < n_cat; ++c) I.prob[c] = prob[c];
  for (int r0 = 0; r0 < R; r0 += RCmax) {
    const int RC = (R - r0) < RCmax ? (R - r0) : RCmax;
    int shift = 0;
    while ((1 << shift) < RC) ++shift;
    if (nD > 0)
      hipLaunchKernelGGL(k_data_rows, dim3(grid_for(nD * RC, 256, 1024)), dim3(256), 0, s,
                         ctx->X.p, (int32_t)N, (int32_t)nD, RC, dcat.p);
    {
      TimedLaunch tl(ctx, GSP_K_TRANSPOSE);
      hipLaunchKernelGGL(k_z_transpose, dim3((unsigned)((N + 63) / 64)), dim3(256),
                         transpose_lds(RC), s, u,
                         (int32_t)N, RC, r0, ctx->Zt.p);
    }
    LevelArgs& A = I.L;
    A.K = plan->kmax;
    A.RC = RC;
    A.shift = shift;
    A.pos = plan->pos.p;
    A.nfound = plan->nfound.p;
    A.flags = plan->flags.p;
    A.nbr = plan->nbr.p;
    A.lam = plan->lam.p;
    A.sigma = plan->sigma.p;
    A.node_req = plan->node_req.p;
    A.mean = plan->req_mean.p;
    A.Zt = ctx->Zt.p;
    A.X = ctx->X.p;
    {
      TimedLaunch tl(ctx, GSP_K_SWEEP);
      for (int32_t l = 0; l < plan->depth; ++l) {
        const int64_t off = plan->level_off[(size_t)l];
        const int64_t cnt = plan->level_off[(size_t)l + 1] - off;
        if (cnt == 0) continue;
        A.order = plan->order.p + off;
        A.count = (int32_t)cnt;
        const int64_t threads = cnt << shift;
        const int64_t per = ((threads + 255) / 256 + kXcds - 1) / kXcds;
        hipLaunchKernelGGL(k_sweep_level_ind, dim3((unsigned)(per * kXcds)), dim3(256), 0, s, I);
      }
    }
    {
      TimedLaunch tl(ctx, GSP_K_TRANSPOSE);
      hipLaunchKernelGGL(k_transpose<int32_t>, dim3((unsigned)((N + 63) / 64)), dim3(256),
                         transpose_lds(RC), s,
                         ctx->X.p, (int32_t)N, RC, r0, out);
    }
    GSP_HIP(hipGetLastError());
  }
}

}  // namespace gsp
