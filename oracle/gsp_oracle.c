/*
 * gsp_oracle.c — CPU oracle for GaussianProcess + SEQSIM (TEST INFRASTRUCTURE).
 *
 * This file is the checker for libgsp_hip.so, never part of it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every arithmetic step is written as an explicit sequence of IEEE double
 * operations (built with -ffp-contract=off) so that a device implementation
 * following the same order can be compared bit for bit. Conventions that the
 * reference's own tests do not fix are marked "unpinned" below.
 */
#include "gsp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ exp ----
 * exp(x) by Cody-Waite argument reduction with the fdlibm split of ln2
 * (ln2_hi has 21 trailing zero bits so k*ln2_hi is exact for |k| < 2^21) and
 * the degree-13 Taylor polynomial of exp on |r| <= ln2/2 (truncation
 * < 4e-18), evaluated by Horner, then scaled by 2^k. It exists so that the
 * device code, which uses the same steps, produces the same bits. */
static double or_pow2i(int k) {
  union { uint64_t u; double d; } v;
  v.u = (uint64_t)(k + 1023) << 52;
  return v.d;
}

static double or_ldexp(double p, int k) {
  if (k > 1023) return (p * or_pow2i(1023)) * or_pow2i(k - 1023);
  if (k < -1022) return (p * or_pow2i(k + 1000)) * or_pow2i(-1000);
  return p * or_pow2i(k);
}

double or_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return INFINITY;
  if (x < -745.1332191019412) return 0.0;
  const double inv_ln2 = 1.44269504088896338700e+00;
  const double ln2_hi = 6.93147180369123816490e-01;
  const double ln2_lo = 1.90821492927058770002e-10;
  double kd = floor(x * inv_ln2 + 0.5);
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
  return or_ldexp(q, k);
}

/* ------------------------------------------------------------ variograms ----
 * GeoStatsFunctions 0.15 (un-vendored; called through Kriging(f, mu) at
 * seqsim.jl:286 and _pairwise at utils.jl:50-62). Practical-range forms:
 *   spherical   gamma = (s-n)(1.5u - 0.5u^3) + n  for 0 < u < 1, (s-n) + n for u >= 1
 *   exponential gamma = (s-n)(1 - exp(-3u)) + n   for u > 0
 *   gaussian    gamma = (s-n')(1 - exp(-3u^2)) + n' for u > 0, n' = n + 1e-6
 * and gamma(0) = 0. Both variograms (kriging uses s - gamma) and covariances
 * (banded: the function value itself) reduce to C(u) = s - gamma(u).
 * The gaussian model carries a stabilising nugget of 1e-6 (sill units) on top
 * of its own, as GeoStatsFunctions' predecessor Variography.jl did "for
 * numerical stability": without it the Gaussian kernel matrix of a unit grid
 * with range 10 is singular in FP64, and the reference's LUSIM testset
 * (test/field.jl:29-56: GaussianCovariance(range=10) on CartesianGrid(100)
 * and (100,100), range 35 on a 10x10 grid) expects those calls to return a
 * field. The value 1e-6 and the x3 practical-range convention are unpinned
 * by the reference's tests. */
double or_cov_u(const or_model* m, double u) {
  double s = m->sill, n = m->nugget;
  if (u == 0.0) return s;
  double shape;
  switch (m->family) {
    case OR_SPHERICAL: shape = (u < 1.0) ? (1.5 * u - 0.5 * (u * u * u)) : 1.0; break;
    case OR_EXPONENTIAL: shape = 1.0 - or_exp(-3.0 * u); break;
    default:
      n = n + OR_GAUSSIAN_NUGGET;
      shape = 1.0 - or_exp(-3.0 * (u * u));
      break;
  }
  double gamma = (s - n) * shape + n;
  return s - gamma;
}

/* Normalised lag u between two points in scaled coordinates. The metric is the
 * ellipsoidal one of a MetricBall with radii `ranges` and rotation R:
 * u = |diag(1/r) R^T (a - b)| (unpinned: the rotation convention). Ranges are
 * multiplied by alpha, as GeoStatsFunctions.scale(f, alpha) does; each lag
 * component is multiplied by the reciprocal of its scaled range (operation
 * order shared with the device code). */
static double or_lag(const or_model* m, int ndim, double alpha, const double* a,
                     const double* b) {
  double dl[3], dp[3];
  for (int d = 0; d < ndim; ++d) dl[d] = a[d] - b[d];
  if (m->has_rotation) {
    for (int e = 0; e < ndim; ++e) {
      double acc = 0.0;
      for (int d = 0; d < ndim; ++d) acc = acc + m->rotation[d * 3 + e] * dl[d];
      dp[e] = acc;
    }
  } else {
    for (int d = 0; d < ndim; ++d) dp[d] = dl[d];
  }
  double u2 = 0.0;
  for (int e = 0; e < ndim; ++e) {
    double ir = 1.0 / (alpha * m->ranges[e]);
    double t = dp[e] * ir;
    u2 = u2 + t * t;
  }
  return sqrt(u2);
}

double or_cov_points(const or_model* m, int ndim, double alpha, const double* xs_a,
                     const double* xs_b) {
  return or_cov_u(m, or_lag(m, ndim, alpha, xs_a, xs_b));
}

/* ------------------------------------------------------------- scaling ----
 * seqsim.jl:162-189: alpha = 1 / max(alpha1, alpha2, alpha3) with
 * alpha1 = max |coordinate| of the grid bounding box corners,
 * alpha2 = same for the data bounding box, alpha3 = range(f) (largest range).
 * The data-only search stands for GeoStatsModels.fitpredict (the posterior
 * mean, expectation/field/gaussian.jl:21-25, and FFTSIM's conditioning,
 * fftsim.jl:99,148), which works in the domain's own coordinates: alpha = 1,
 * so a datum on a cell centroid lies at lag exactly 0 (this matters where the
 * model has a nugget, whose jump starts at any lag above 0). */
double or_scale_factor(const or_seqsim_problem* p) {
  if (p->search_mode == OR_DATA_ONLY) return 1.0;
  const or_grid* g = &p->grid;
  double a1 = 0.0;
  for (int d = 0; d < g->ndim; ++d) {
    double lo = fabs(g->origin[d]);
    double hi = fabs(g->origin[d] + (double)g->dims[d] * g->spacing[d]);
    if (lo > a1) a1 = lo;
    if (hi > a1) a1 = hi;
  }
  double a2 = 0.0;
  for (int64_t i = 0; i < p->n_data; ++i)
    for (int d = 0; d < g->ndim; ++d) {
      double v = fabs(p->data_coords[i * g->ndim + d]);
      if (v > a2) a2 = v;
    }
  double a3 = 0.0;
  for (int d = 0; d < g->ndim; ++d)
    if (p->model.ranges[d] > a3) a3 = p->model.ranges[d];
  double a = a1;
  if (a2 > a) a = a2;
  if (a3 > a) a = a3;
  return 1.0 / a;
}

/* seqsim.jl:202-208 */
void or_clamp_neighbors(int64_t nelm, int32_t* nmin, int32_t* nmax) {
  if ((int64_t)*nmax > nelm || *nmax < 1) *nmax = (int32_t)nelm;
  if (*nmin > *nmax || *nmin < 1) *nmin = 1;
}

/* ------------------------------------------------------------ geometry ---- */
static int64_t or_nelm(const or_grid* g) {
  int64_t n = 1;
  for (int d = 0; d < g->ndim; ++d) n *= g->dims[d];
  return n;
}

static void or_unravel(const or_grid* g, int64_t id, int64_t* ijk) {
  for (int d = 0; d < g->ndim; ++d) {
    ijk[d] = id % g->dims[d];
    id /= g->dims[d];
  }
}

/* Scaled centroid of augmented point a: the grid scaled by alpha keeps its
 * shape, with origin alpha*o and spacing alpha*s (Meshes Scale), so its cell
 * centroid is alpha*o + (i + 1/2)*(alpha*s); data points become alpha*x. */
static void or_scaled_coord(const or_seqsim_problem* p, double alpha, int64_t N, int64_t a,
                            double* xs) {
  const or_grid* g = &p->grid;
  if (a < N) {
    int64_t ijk[3];
    or_unravel(g, a, ijk);
    for (int d = 0; d < g->ndim; ++d)
      xs[d] = alpha * g->origin[d] + ((double)ijk[d] + 0.5) * (alpha * g->spacing[d]);
  } else {
    for (int d = 0; d < g->ndim; ++d) xs[d] = alpha * p->data_coords[(a - N) * g->ndim + d];
  }
}

/* Search metric. Squared Euclidean distance in physical units; between two
 * grid nodes it is formed from the integer index offsets, so equal offsets give
 * exactly equal distances and ties are broken only by the augmented index
 * (lower wins). Meshes orders ties by KD-tree heap order instead (unpinned). */
static double or_d2_grid(const or_grid* g, const int64_t* a, const int64_t* b) {
  double acc = 0.0;
  for (int d = 0; d < g->ndim; ++d) {
    double t = (double)(a[d] - b[d]) * g->spacing[d];
    acc = acc + t * t;
  }
  return acc;
}

static double or_d2_data(const or_grid* g, const double* x, const int64_t* node) {
  double acc = 0.0;
  for (int d = 0; d < g->ndim; ++d) {
    double c = g->origin[d] + ((double)node[d] + 0.5) * g->spacing[d];
    double t = x[d] - c;
    acc = acc + t * t;
  }
  return acc;
}

/* ------------------------------------------------------------- top-k ---- */
typedef struct {
  int32_t k, n;
  double* d2;
  int32_t* idx;
} or_topk;

static int or_key_less(double da, int32_t ia, double db, int32_t ib) {
  return da < db || (da == db && ia < ib);
}

static void or_topk_push(or_topk* t, double d2, int32_t idx) {
  if (t->n == t->k && !or_key_less(d2, idx, t->d2[t->n - 1], t->idx[t->n - 1])) return;
  int j = (t->n < t->k) ? t->n++ : t->n - 1;
  while (j > 0 && or_key_less(d2, idx, t->d2[j - 1], t->idx[j - 1])) {
    t->d2[j] = t->d2[j - 1];
    t->idx[j] = t->idx[j - 1];
    --j;
  }
  t->d2[j] = d2;
  t->idx[j] = idx;
}

/* Meshes KBallSearch / KNearestSearch with mask (semantics restated):
 *   OR_KNN_THEN_MASK: the k nearest of ALL augmented points are found first,
 *     then those outside the ball or not yet simulated are dropped (Meshes
 *     `searchdists!` filters the knn result by `dists .<= radius` and
 *     `mask[inds]`). This is the default.
 *   OR_MASKED_KNN: the k nearest among points that are simulated and in the ball.
 * Which one Meshes 0.57 uses is unpinned; both are kept. A grid node j is
 * simulated before target i iff pos[j] < pos[i]; data points always are.
 *   OR_DATA_ONLY: the k nearest data points in the ball, no grid nodes (simple
 *     kriging from the data: the posterior mean of
 *     expectation/field/gaussian.jl:21-25). */
static int or_eligible(int64_t N, const int64_t* pos, int64_t target, int64_t a) {
  return a >= N || pos[a] < pos[target];
}

static int32_t or_finish(const or_seqsim_problem* p, int64_t N, const int64_t* pos,
                         int64_t target, const or_topk* t, int32_t* out) {
  int32_t n = 0;
  for (int32_t j = 0; j < t->n; ++j) {
    if (p->search_mode == OR_KNN_THEN_MASK && !or_eligible(N, pos, target, t->idx[j])) continue;
    out[n++] = t->idx[j];
  }
  return n;
}

int32_t or_search_brute(const or_seqsim_problem* p, int32_t nmax, const int64_t* pos,
                        int64_t target, int32_t* out) {
  const or_grid* g = &p->grid;
  int64_t N = or_nelm(g);
  double r2 = p->radius > 0 ? p->radius * p->radius : INFINITY;
  double kd2[64];
  int32_t kidx[64];
  or_topk t = {nmax, 0, kd2, kidx};
  int64_t ti[3], ji[3];
  or_unravel(g, target, ti);
  for (int64_t a = 0; a < N + p->n_data; ++a) {
    double d2;
    if (a < N) {
      if (p->search_mode == OR_DATA_ONLY) continue;
      or_unravel(g, a, ji);
      d2 = or_d2_grid(g, ji, ti);
    } else {
      d2 = or_d2_data(g, p->data_coords + (a - N) * g->ndim, ti);
    }
    if (!(d2 <= r2)) continue;
    if (p->search_mode == OR_MASKED_KNN && !or_eligible(N, pos, target, a)) continue;
    or_topk_push(&t, d2, (int32_t)a);
  }
  return or_finish(p, N, pos, target, &t, out);
}

/* Cell-hash restatement of the same search: grid nodes are visited in
 * Chebyshev rings around the target; data points are binned by grid cell
 * (points outside the grid are checked every time). Ring m holds no grid node
 * closer than m*s_min and no binned datum closer than (m - 1/2)*s_min. */
typedef struct {
  int64_t N;
  int64_t* cell_start; /* N + 1 */
  int32_t* cell_data;
  int32_t* outside;
  int64_t n_outside;
} or_bucket;

void* or_bucket_create(const or_seqsim_problem* p) {
  const or_grid* g = &p->grid;
  or_bucket* b = (or_bucket*)calloc(1, sizeof(or_bucket));
  b->N = or_nelm(g);
  b->cell_start = (int64_t*)calloc((size_t)b->N + 1, sizeof(int64_t));
  int64_t* cell_of = (int64_t*)malloc(sizeof(int64_t) * (size_t)(p->n_data + 1));
  b->outside = (int32_t*)malloc(sizeof(int32_t) * (size_t)(p->n_data + 1));
  for (int64_t i = 0; i < p->n_data; ++i) {
    int64_t id = 0, stride = 1, inside = 1;
    for (int d = 0; d < g->ndim; ++d) {
      double f = floor((p->data_coords[i * g->ndim + d] - g->origin[d]) / g->spacing[d]);
      if (!(f >= 0.0 && f < (double)g->dims[d])) { inside = 0; break; }
      id += (int64_t)f * stride;
      stride *= g->dims[d];
    }
    if (inside) { cell_of[i] = id; b->cell_start[id + 1]++; }
    else { cell_of[i] = -1; b->outside[b->n_outside++] = (int32_t)i; }
  }
  for (int64_t c = 0; c < b->N; ++c) b->cell_start[c + 1] += b->cell_start[c];
  b->cell_data = (int32_t*)malloc(sizeof(int32_t) * (size_t)(p->n_data + 1));
  int64_t* fill = (int64_t*)calloc((size_t)b->N, sizeof(int64_t));
  for (int64_t i = 0; i < p->n_data; ++i)
    if (cell_of[i] >= 0) b->cell_data[b->cell_start[cell_of[i]] + fill[cell_of[i]]++] = (int32_t)i;
  free(fill);
  free(cell_of);
  return b;
}

void or_bucket_free(void* vb) {
  or_bucket* b = (or_bucket*)vb;
  if (!b) return;
  free(b->cell_start);
  free(b->cell_data);
  free(b->outside);
  free(b);
}

int32_t or_search_bucket(const or_seqsim_problem* p, void* vb, int32_t nmax,
                         const int64_t* pos, int64_t target, int32_t* out) {
  const or_grid* g = &p->grid;
  const or_bucket* b = (const or_bucket*)vb;
  int64_t N = b->N;
  int nd = g->ndim;
  double r2 = p->radius > 0 ? p->radius * p->radius : INFINITY;
  double kd2[64];
  int32_t kidx[64];
  or_topk t = {nmax, 0, kd2, kidx};
  int64_t ti[3] = {0, 0, 0}, ji[3] = {0, 0, 0};
  or_unravel(g, target, ti);
  int masked = p->search_mode == OR_MASKED_KNN;
  double smin = g->spacing[0];
  int64_t mmax = 0;
  for (int d = 0; d < nd; ++d) {
    if (g->spacing[d] < smin) smin = g->spacing[d];
    int64_t e = ti[d] > g->dims[d] - 1 - ti[d] ? ti[d] : g->dims[d] - 1 - ti[d];
    if (e > mmax) mmax = e;
  }
  for (int64_t o = 0; o < b->n_outside; ++o) {
    int32_t di = b->outside[o];
    double d2 = or_d2_data(g, p->data_coords + (int64_t)di * nd, ti);
    if (!(d2 <= r2)) continue;
    or_topk_push(&t, d2, (int32_t)(N + di));
  }
  for (int64_t m = 0; m <= mmax; ++m) {
    /* stop once ring m cannot beat the current k-th key or leaves the ball */
    if (m > 0) {
      double lb = ((double)m - 0.5) * smin;
      double lb2 = lb * lb;
      if (lb2 > r2) break;
      if (t.n == t.k && lb2 > t.d2[t.n - 1]) break;
    }
    int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < nd; ++d) { lo[d] = -m; hi[d] = m; }
    for (int64_t dz = lo[2]; dz <= hi[2]; ++dz) {
      int64_t z = ti[2] + dz;
      if (nd >= 3 && (z < 0 || z >= g->dims[2])) continue;
      for (int64_t dy = lo[1]; dy <= hi[1]; ++dy) {
        int64_t y = ti[1] + dy;
        if (nd >= 2 && (y < 0 || y >= g->dims[1])) continue;
        int64_t ady = dy < 0 ? -dy : dy, adz = dz < 0 ? -dz : dz;
        int full = (ady == m) || (adz == m);
        int64_t step = full || m == 0 ? 1 : 2 * m;
        for (int64_t dx = -m; dx <= m; dx += step) {
          int64_t x = ti[0] + dx;
          if (x < 0 || x >= g->dims[0]) continue;
          ji[0] = x; ji[1] = y; ji[2] = z;
          int64_t id = x + (nd >= 2 ? g->dims[0] * (y + (nd >= 3 ? g->dims[1] * z : 0)) : 0);
          double d2 = or_d2_grid(g, ji, ti);
          if (d2 <= r2 && p->search_mode != OR_DATA_ONLY &&
              (!masked || or_eligible(N, pos, target, id)))
            or_topk_push(&t, d2, (int32_t)id);
          for (int64_t q = b->cell_start[id]; q < b->cell_start[id + 1]; ++q) {
            int32_t di = b->cell_data[q];
            double dd2 = or_d2_data(g, p->data_coords + (int64_t)di * nd, ti);
            if (dd2 <= r2) or_topk_push(&t, dd2, (int32_t)(N + di));
          }
        }
      }
    }
  }
  return or_finish(p, N, pos, target, &t, out);
}

/* -------------------------------------------------------------- kriging ----
 * GeoStatsModels 0.13 simple kriging with known mean (Kriging(f, mu),
 * seqsim.jl:286; fit at :124, status at :127, predictprob at :303):
 *   LHS_jm = C(x_j, x_m), c0_j = C(x_j, x0), LHS = L L^T (Cholesky),
 *   lambda = LHS^-1 c0, mean = mu + sum_j lambda_j (z_j - mu),
 *   var = C(0) - sum_j lambda_j c0_j  (clamped at 0: unpinned).
 * status is false when a pivot is not positive (LAPACK potrf failure); the
 * caller then draws from the prior Normal(mu, sqrt(sill)) (seqsim.jl:127-131).
 * Operation order (fixed, shared with the device code):
 *   L_mm = sqrt(pivot), r_m = 1 / L_mm (the column is scaled by the
 *   reciprocal of its pivot, as LAPACK's unblocked dpotf2 does),
 *   L_jm = (A_jm - L_j0 L_m0 - L_j1 L_m1 - ...) * r_m
 *   forward  y_j = (c_j - L_j0 y_0 - ... - L_j,j-1 y_j-1) * r_j
 *   backward l_j = (y_j - L_n-1,j l_n-1 - ... - L_j+1,j l_j+1) * r_j
 *   var = C0 - (l_0 c_0 + l_1 c_1 + ...)
 */
#define OR_KMAX 64

static int or_chol64(int n, double* A /* n x OR_KMAX, lower */) {
  for (int m = 0; m < n; ++m) {
    double s = A[m * OR_KMAX + m];
    for (int c = 0; c < m; ++c) s = s - A[m * OR_KMAX + c] * A[m * OR_KMAX + c];
    if (!(s > 0.0)) return 0;
    double lmm = sqrt(s);
    double rinv = 1.0 / lmm;
    A[m * OR_KMAX + m] = lmm;
    for (int j = m + 1; j < n; ++j) {
      double v = A[j * OR_KMAX + m];
      for (int c = 0; c < m; ++c) v = v - A[j * OR_KMAX + c] * A[m * OR_KMAX + c];
      A[j * OR_KMAX + m] = v * rinv;
    }
  }
  return 1;
}

static void or_solve64(int n, const double* L, const double* c, double* lam) {
  double y[OR_KMAX];
  for (int j = 0; j < n; ++j) {
    double s = c[j];
    for (int m = 0; m < j; ++m) s = s - L[j * OR_KMAX + m] * y[m];
    y[j] = s * (1.0 / L[j * OR_KMAX + j]);
  }
  for (int j = n - 1; j >= 0; --j) {
    double s = y[j];
    for (int m = n - 1; m > j; --m) s = s - L[m * OR_KMAX + j] * lam[m];
    lam[j] = s * (1.0 / L[j * OR_KMAX + j]);
  }
}

/* FP32 factor; a pivot at or below tau sends the system to FP64 */
static int or_chol32(int n, float* A, float tau) {
  for (int m = 0; m < n; ++m) {
    float s = A[m * OR_KMAX + m];
    for (int c = 0; c < m; ++c) s = s - A[m * OR_KMAX + c] * A[m * OR_KMAX + c];
    if (!(s > tau)) return 0;
    float lmm = sqrtf(s);
    float rinv = 1.0f / lmm;
    A[m * OR_KMAX + m] = lmm;
    for (int j = m + 1; j < n; ++j) {
      float v = A[j * OR_KMAX + m];
      for (int c = 0; c < m; ++c) v = v - A[j * OR_KMAX + c] * A[m * OR_KMAX + c];
      A[j * OR_KMAX + m] = v * rinv;
    }
  }
  return 1;
}

static void or_solve32(int n, const float* L, const float* c, float* lam) {
  float y[OR_KMAX];
  for (int j = 0; j < n; ++j) {
    float s = c[j];
    for (int m = 0; m < j; ++m) s = s - L[j * OR_KMAX + m] * y[m];
    y[j] = s * (1.0f / L[j * OR_KMAX + j]);
  }
  for (int j = n - 1; j >= 0; --j) {
    float s = y[j];
    for (int m = n - 1; m > j; --m) s = s - L[m * OR_KMAX + j] * lam[m];
    lam[j] = s * (1.0f / L[j * OR_KMAX + j]);
  }
}

int32_t or_krige(const or_seqsim_problem* p, double alpha, int64_t target,
                 const int32_t* nbr, int32_t n, int32_t precision, double* lambda,
                 double* sigma) {
  const or_grid* g = &p->grid;
  int nd = g->ndim;
  int64_t N = or_nelm(g);
  double xs[OR_KMAX][3], x0[3];
  double A[OR_KMAX * OR_KMAX], c[OR_KMAX];
  or_scaled_coord(p, alpha, N, target, x0);
  for (int j = 0; j < n; ++j) or_scaled_coord(p, alpha, N, nbr[j], xs[j]);
  for (int j = 0; j < n; ++j) {
    for (int m = 0; m <= j; ++m) A[j * OR_KMAX + m] = or_cov_points(&p->model, nd, alpha, xs[j], xs[m]);
    c[j] = or_cov_points(&p->model, nd, alpha, xs[j], x0);
  }
  double c0 = or_cov_u(&p->model, 0.0);
  int32_t flags = 0;
  int done = 0;
  if (precision == OR_FP32_REFINE) {
    float A32[OR_KMAX * OR_KMAX], c32[OR_KMAX], l32[OR_KMAX], r32[OR_KMAX], d32[OR_KMAX];
    for (int j = 0; j < n; ++j) {
      for (int m = 0; m <= j; ++m) A32[j * OR_KMAX + m] = (float)A[j * OR_KMAX + m];
      c32[j] = (float)c[j];
    }
    /* tau = sill * 2^-12: below it one FP32 refinement step cannot reach
     * FP64-level weights, so the node is solved in FP64 */
    if (or_chol32(n, A32, (float)(p->model.sill * 0x1p-12))) {
      or_solve32(n, A32, c32, l32);
      /* one refinement step with the residual formed in FP64 */
      for (int j = 0; j < n; ++j) {
        double s = c[j];
        for (int m = 0; m < n; ++m) {
          double a = m <= j ? A[j * OR_KMAX + m] : A[m * OR_KMAX + j];
          s = s - a * (double)l32[m];
        }
        r32[j] = (float)s;
      }
      or_solve32(n, A32, r32, d32);
      for (int j = 0; j < n; ++j) lambda[j] = (double)l32[j] + (double)d32[j];
      done = 1;
    } else {
      flags |= OR_FLAG_FP64_FALLBACK;
    }
  }
  if (!done) {
    if (!or_chol64(n, A)) {
      for (int j = 0; j < n; ++j) lambda[j] = 0.0;
      *sigma = sqrt(p->model.sill);
      return flags | OR_FLAG_FAIL;
    }
    or_solve64(n, A, c, lambda);
  }
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc = acc + lambda[j] * c[j];
  double var = c0 - acc;
  if (var < 0.0) var = 0.0;
  *sigma = sqrt(var);
  return flags;
}

/* ----------------------------------------------------------------- plan ---- */
static int64_t* or_positions(const or_seqsim_problem* p, int64_t N) {
  if (p->n_path != N) return NULL;
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)N);
  for (int64_t i = 0; i < N; ++i) pos[i] = -1;
  for (int64_t q = 0; q < N; ++q) {
    int64_t i = p->path[q];
    if (i < 0 || i >= N || pos[i] >= 0) { free(pos); return NULL; }
    pos[i] = q;
  }
  return pos;
}

int32_t or_seqsim_plan(const or_seqsim_problem* p, int32_t use_brute, int32_t* nfound,
                       int32_t* nbr, double* lambda, double* sigma, uint8_t* flags,
                       int32_t* level, int32_t* kmax_out) {
  const or_grid* g = &p->grid;
  int64_t N = or_nelm(g);
  int32_t nmin = p->minneighbors, nmax = p->maxneighbors;
  or_clamp_neighbors(N, &nmin, &nmax);
  if (nmax > OR_KMAX) return -1;
  *kmax_out = nmax;
  int64_t* pos = or_positions(p, N);
  if (!pos) return -1;
  double alpha = or_scale_factor(p);
  void* bucket = use_brute ? NULL : or_bucket_create(p);
  for (int64_t q = 0; q < N; ++q) {
    int64_t i = p->path[q];
    int32_t* nb = nbr + i * nmax;
    double* lam = lambda + i * nmax;
    int32_t n = use_brute ? or_search_brute(p, nmax, pos, i, nb)
                          : or_search_bucket(p, bucket, nmax, pos, i, nb);
    for (int32_t j = n; j < nmax; ++j) { nb[j] = -1; lam[j] = 0.0; }
    nfound[i] = n;
    int32_t f;
    if (n < nmin) {
      for (int32_t j = 0; j < n; ++j) lam[j] = 0.0;
      sigma[i] = sqrt(p->model.sill);
      f = OR_FLAG_FEW;
    } else {
      f = or_krige(p, alpha, i, nb, n, p->precision, lam, sigma + i);
    }
    flags[i] = (uint8_t)f;
    /* level: 0 without grid predecessors, else 1 + max over them */
    int32_t lv = 0;
    if (!(f & (OR_FLAG_FEW | OR_FLAG_FAIL)))
      for (int32_t j = 0; j < n; ++j)
        if (nb[j] < N && level[nb[j]] + 1 > lv) lv = level[nb[j]] + 1;
    level[i] = lv;
  }
  or_bucket_free(bucket);
  free(pos);
  return 0;
}

/* ---------------------------------------------------------------- sweep ----
 * seqsim.jl:98-138 with the realization-invariant part hoisted: for each path
 * node, x_i = (mu + sum_j lambda_j (x_nbr(j) - mu)) + sigma_i z. Prior nodes
 * (flags FEW/FAIL) use no neighbours and sigma = sqrt(sill). Data values sit at
 * augmented slots N + d (ExplicitInit onto the augmented domain, seqsim.jl:267). */
static double or_draw(int32_t k, const int32_t* nb, const double* lam, double mean,
                      double sigma, const double* x, double z) {
  double acc = 0.0;
  for (int32_t j = 0; j < k; ++j) acc = acc + lam[j] * (x[nb[j]] - mean);
  double mu = mean + acc;
  return mu + sigma * z;
}

void or_seqsim_sweep(const or_seqsim_problem* p, int32_t kmax, const int32_t* nfound,
                     const int32_t* nbr, const double* lambda, const double* sigma,
                     const uint8_t* flags, int32_t R, const double* z,
                     const double* data_vals, double* out) {
  int64_t N = or_nelm(&p->grid);
  double* x = (double*)malloc(sizeof(double) * (size_t)(N + p->n_data));
  for (int32_t r = 0; r < R; ++r) {
    for (int64_t i = 0; i < N; ++i) x[i] = NAN;
    for (int64_t d = 0; d < p->n_data; ++d) x[N + d] = data_vals[d];
    for (int64_t q = 0; q < p->n_path; ++q) {
      int64_t i = p->path[q];
      int32_t k = (flags[i] & (OR_FLAG_FEW | OR_FLAG_FAIL)) ? 0 : nfound[i];
      x[i] = or_draw(k, nbr + i * kmax, lambda + i * kmax, p->mean, sigma[i], x,
                     z[(int64_t)r * p->n_path + q]);
    }
    memcpy(out + (int64_t)r * N, x, sizeof(double) * (size_t)N);
  }
  free(x);
}

/* ------------------------------------------------- reference cost model ---- */
int64_t or_seqsim_reference_cost(const or_seqsim_problem* p, int32_t R, const double* z,
                                 const double* data_vals, double* out, int32_t threads,
                                 int64_t n_nodes_limit) {
  const or_grid* g = &p->grid;
  int64_t N = or_nelm(g);
  int32_t nmin = p->minneighbors, nmax = p->maxneighbors;
  or_clamp_neighbors(N, &nmin, &nmax);
  if (nmax > OR_KMAX) return -1;
  int64_t* pos = or_positions(p, N);
  if (!pos) return -1;
  double alpha = or_scale_factor(p);
  void* bucket = or_bucket_create(p);
  int64_t nq = (n_nodes_limit > 0 && n_nodes_limit < N) ? n_nodes_limit : N;
  if (threads < 1) threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
  for (int32_t r = 0; r < R; ++r) {
    double* x = (double*)malloc(sizeof(double) * (size_t)(N + p->n_data));
    int32_t nb[OR_KMAX];
    double lam[OR_KMAX], sig;
    for (int64_t i = 0; i < N; ++i) x[i] = NAN;
    for (int64_t d = 0; d < p->n_data; ++d) x[N + d] = data_vals[d];
    for (int64_t q = 0; q < nq; ++q) {
      int64_t i = p->path[q];
      int32_t n = or_search_bucket(p, bucket, nmax, pos, i, nb);
      int32_t k = 0;
      if (n < nmin) {
        sig = sqrt(p->model.sill);
      } else {
        int32_t f = or_krige(p, alpha, i, nb, n, p->precision, lam, &sig);
        k = (f & OR_FLAG_FAIL) ? 0 : n;
      }
      x[i] = or_draw(k, nb, lam, p->mean, sig, x, z[(int64_t)r * p->n_path + q]);
    }
    memcpy(out + (int64_t)r * N, x, sizeof(double) * (size_t)N);
    free(x);
  }
  or_bucket_free(bucket);
  free(pos);
  return (int64_t)R * nq;
}
