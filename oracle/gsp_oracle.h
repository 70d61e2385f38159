/*
 * gsp_oracle.h — CPU oracle for the SEQSIM path (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker / the timed CPU restatement. The product
 * (libgsp_hip.so) never links or calls it.
 *
 * It restates, in plain FP64 C, what GeoStatsProcesses v0.13.0 does for
 * GaussianProcess + SEQSIM:
 *   src/simulation/field/seqsim.jl:50-70    preprocess (scale, search, model, init)
 *   src/simulation/field/seqsim.jl:72-156   randsingle (the sequential sweep)
 *   src/simulation/field/seqsim.jl:162-189  _scale / _scalefactor
 *   src/simulation/field/seqsim.jl:195-230  _search (nmin/nmax clamping, augmented set)
 *   src/simulation/field/seqsim.jl:277-290  _probmodel (simple kriging, prior)
 *   src/simulation/field/seqsim.jl:335-340  _draw! (one standard normal per node)
 * plus the un-vendored dependency semantics it calls (Meshes 0.57 KBallSearch /
 * KNearestSearch, GeoStatsFunctions 0.15 variograms, GeoStatsModels 0.13 simple
 * kriging, Distributions 0.25 Normal draws), restated from their published
 * algorithms; conventions that no reference test fixes are listed per function.
 */
#ifndef GSP_ORACLE_H
#define GSP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_SPHERICAL = 0, OR_EXPONENTIAL = 1, OR_GAUSSIAN = 2 };
/* stabilising nugget of the gaussian model (unpinned value, see or_cov_u) */
#define OR_GAUSSIAN_NUGGET 1e-6
enum { OR_KNN_THEN_MASK = 0, OR_MASKED_KNN = 1, OR_DATA_ONLY = 2 };
enum { OR_FP64 = 0, OR_FP32_REFINE = 1 };
enum { OR_FLAG_FEW = 1, OR_FLAG_FAIL = 2, OR_FLAG_FP64_FALLBACK = 4 };

/* Same field layout as gsp_model / gsp_grid in include/gsp.h. */
typedef struct {
  int32_t family;     /* OR_SPHERICAL | OR_EXPONENTIAL | OR_GAUSSIAN */
  int32_t kind;       /* 0 variogram, 1 covariance (banded) */
  double sill;
  double nugget;
  double ranges[3];   /* physical units, one per axis */
  double rotation[9]; /* row-major R; the anisotropic metric uses R^T * delta */
  int32_t has_rotation;
  int32_t pad_;
} or_model;

typedef struct {
  int32_t ndim;
  int32_t pad_;
  int64_t dims[3];
  double origin[3];   /* minimum corner of the grid */
  double spacing[3];
} or_grid;

typedef struct {
  or_grid grid;
  or_model model;
  double mean;
  int64_t n_data;
  const double* data_coords; /* n_data x ndim, point-major, physical units */
  int64_t n_path;
  const int64_t* path;       /* 0-based grid ids, a permutation of 0..N-1 */
  int32_t minneighbors;
  int32_t maxneighbors;
  double radius;             /* > 0: MetricBall radius; <= 0: plain k-nearest */
  int32_t search_mode;       /* OR_KNN_THEN_MASK | OR_MASKED_KNN | OR_DATA_ONLY */
  int32_t precision;         /* OR_FP64 | OR_FP32_REFINE */
  int32_t device_pointers;   /* ignored by the oracle */
  int32_t pad_;
} or_seqsim_problem;

/* --- primitives (exposed so tests can pin them) --- */
double or_exp(double x);
double or_cov_u(const or_model* m, double u);              /* C = sill - gamma(u), u = h/range */
double or_scale_factor(const or_seqsim_problem* p);         /* alpha, seqsim.jl:162-189 */
double or_cov_points(const or_model* m, int ndim, double alpha,
                     const double* xs_a, const double* xs_b); /* on scaled coordinates */
void or_clamp_neighbors(int64_t nelm, int32_t* nmin, int32_t* nmax); /* seqsim.jl:202-208 */

/* Neighbour search of grid node `target` (0-based). pos[j] = path position of grid
 * node j. Writes up to nmax augmented indices (grid ids, then N + data id) in
 * ascending (d2, index) order; returns the count. brute: exhaustive. bucket: a
 * cell-hash restatement that must return the identical list. */
int32_t or_search_brute(const or_seqsim_problem* p, int32_t nmax, const int64_t* pos,
                        int64_t target, int32_t* out);
void* or_bucket_create(const or_seqsim_problem* p);
void or_bucket_free(void* b);
int32_t or_search_bucket(const or_seqsim_problem* p, void* bucket, int32_t nmax,
                         const int64_t* pos, int64_t target, int32_t* out);

/* Simple-kriging system of `n` neighbours around `target`: lambda[n], sigma.
 * Returns flags (0 ok, OR_FLAG_FAIL when the Cholesky pivot is not positive). */
int32_t or_krige(const or_seqsim_problem* p, double alpha, int64_t target,
                 const int32_t* nbr, int32_t n, int32_t precision,
                 double* lambda, double* sigma);

/* Realization-invariant plan, indexed by grid id (domain order):
 * nfound[N], nbr[N*kmax], lambda[N*kmax], sigma[N], flags[N], level[N].
 * use_brute selects the exhaustive search. Returns 0, or -1 on bad input. */
int32_t or_seqsim_plan(const or_seqsim_problem* p, int32_t use_brute,
                       int32_t* nfound, int32_t* nbr, double* lambda, double* sigma,
                       uint8_t* flags, int32_t* level, int32_t* kmax_out);

/* The sweep: z is R x n_path (realization-major, path order); out is R x N. */
void or_seqsim_sweep(const or_seqsim_problem* p, int32_t kmax, const int32_t* nfound,
                     const int32_t* nbr, const double* lambda, const double* sigma,
                     const uint8_t* flags, int32_t R, const double* z,
                     const double* data_vals, double* out);

/* The reference's own cost structure (seqsim.jl:98-138): search + kriging solve
 * for every node of every realization. OpenMP over realizations with `threads`
 * (the restated pmap over workers, field.jl:117). Processes only the first
 * `n_nodes_limit` path nodes of each realization when > 0 (bounded timing
 * sample). Returns node-realizations processed. */
int64_t or_seqsim_reference_cost(const or_seqsim_problem* p, int32_t R, const double* z,
                                 const double* data_vals, double* out, int32_t threads,
                                 int64_t n_nodes_limit);

#ifdef __cplusplus
}
#endif
#endif
