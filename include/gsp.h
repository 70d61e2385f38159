/*
 * gsp.h — C ABI of libgsp_hip.so, the MI355X (gfx950) backend for
 * GeoStatsProcesses' GaussianProcess simulation methods SEQSIM and LUSIM
 * (and FFTSIM, the default method of unconditional grids, field.jl:159).
 *
 * The Julia host keeps `rand(process, domain, nreals; method=...)` unchanged
 * (src/simulation/field.jl:73-124). A package extension (the pattern of
 * ext/GeoStatsProcessesTuringPatternsExt.jl:15, `import GeoStatsProcesses:
 * preprocess, randsingle`) overrides the two plugin functions for
 * (GaussianProcess, SEQSIM), (GaussianProcess, LUSIM) and (GaussianProcess,
 * FFTSIM) and `ccall`s the entry points below. Plain C types only; every function returns a gsp_status.
 *
 * Conventions
 *  - Indices are 0-based; grids are column-major (x fastest), like Meshes'
 *    LinearIndices(size(grid)).
 *  - Coordinates, ranges and radii are in physical units with units stripped
 *    (the shim applies `ustrip` and restores units, seqsim.jl:86,141-143).
 *  - Buffers are host pointers unless a call says otherwise or the problem's
 *    `device_pointers` field is 1. The library copies host buffers in and out
 *    and never keeps a host pointer after returning.
 *  - Calls on one context are serialised by the context; distinct contexts
 *    may be used from different threads.
 */
#ifndef GSP_H
#define GSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSP_ABI_VERSION 1

typedef enum {
  GSP_OK = 0,
  GSP_ERR_ARGUMENT = 1,    /* Julia ArgumentError (lusim.jl:46, field.jl:99) */
  GSP_ERR_ASSERTION = 2,   /* Julia AssertionError (lusim.jl:59-60, explicit.jl:29-30) */
  GSP_ERR_POSDEF = 3,      /* LinearAlgebra.PosDefException (lusim.jl:92,98,103) */
  GSP_ERR_DEVICE = 4,      /* HIP runtime failure */
  GSP_ERR_MEMORY = 5,      /* device allocation failure */
  GSP_ERR_UNSUPPORTED = 6, /* a configuration this backend does not implement */
  GSP_ERR_INTERNAL = 7     /* a device-side consistency check failed */
} gsp_status;

typedef enum { GSP_SPHERICAL = 0, GSP_EXPONENTIAL = 1, GSP_GAUSSIAN = 2 } gsp_family;
typedef enum { GSP_VARIOGRAM = 0, GSP_COVARIANCE = 1 } gsp_kind;

/* Neighbour-search semantics of Meshes KBallSearch/KNearestSearch with a mask
 * (called at seqsim.jl:107): the k nearest of all augmented points, then drop
 * those outside the ball or not yet simulated (GSP_KNN_THEN_MASK, default), or
 * the k nearest among the simulated points in the ball (GSP_MASKED_KNN). Ties
 * in distance go to the lower augmented index. GSP_DATA_ONLY takes the k
 * nearest conditioning data in the ball and no grid nodes: every node is then
 * an independent simple-kriging estimate from the data, the posterior mean of
 * expectation/field/gaussian.jl:21-25 (GeoStatsModels.fitpredict with
 * Kriging(f, mean)); a sweep of one realization with z = 0 returns that mean
 * and the plan's sigma^2 is the kriging variance. */
typedef enum { GSP_KNN_THEN_MASK = 0, GSP_MASKED_KNN = 1, GSP_DATA_ONLY = 2 } gsp_search_mode;

/* Precision of the local kriging solves: FP64 Cholesky, or FP32 Cholesky with
 * one FP64-residual refinement step (FP64 retry if the FP32 factor fails). */
typedef enum { GSP_FP64 = 0, GSP_FP32_REFINE = 1 } gsp_precision;

/* LUSIM factorisation precision: FP64 throughout, or FP32 MFMA trailing
 * updates with FP64 diagonal panels and FP64 conditioning solves. */
typedef enum { GSP_LU_FP64 = 0, GSP_LU_MIXED = 1 } gsp_lu_precision;

/* Plan flags per node (SEQSIM). */
enum {
  GSP_FLAG_FEW = 1,          /* n < minneighbors: prior draw (seqsim.jl:109-111) */
  GSP_FLAG_FAIL = 2,         /* kriging Cholesky failed: prior (seqsim.jl:127-131) */
  GSP_FLAG_FP64_FALLBACK = 4, /* FP32 factor failed, FP64 solve used */
  GSP_FLAG_LEARNED = 8        /* weights and sigma from the learned surrogate */
};

/* A GeoStatsFunctions variogram or covariance (un-vendored, 0.15):
 * C(h) = sill - gamma(h), gamma(0) = 0, gamma = (sill - nugget) * shape(u) +
 * nugget for u > 0, u = |diag(1/ranges) R^T (x - y)|. */
typedef struct {
  int32_t family;       /* gsp_family */
  int32_t kind;         /* gsp_kind */
  double sill;
  double nugget;
  double ranges[3];     /* one per axis (only ndim used) */
  double rotation[9];   /* row-major R, used when has_rotation != 0 */
  int32_t has_rotation;
  int32_t pad_;
} gsp_model;

/* A Meshes CartesianGrid: cell centroid of index (i,j,k) is
 * origin + ((i,j,k) + 1/2) .* spacing. */
typedef struct {
  int32_t ndim;         /* 1..3 */
  int32_t pad_;
  int64_t dims[3];
  double origin[3];     /* minimum corner */
  double spacing[3];
} gsp_grid;

/* Everything `preprocess(rng, process, ::SEQSIM, init, domain, data)` needs
 * (seqsim.jl:50-70): the grid, the process function and mean, the data
 * locations (values come at sweep time), the traversal path and the SEQSIM
 * options (seqsim.jl:42-48; maxneighbors/minneighbors are clamped as in
 * seqsim.jl:202-208; neighborhood :range is passed as radius = range(f)). */
typedef struct {
  gsp_grid grid;
  gsp_model model;
  double mean;                  /* process.mean */
  int64_t n_data;
  const double* data_coords;    /* n_data x ndim, point-major */
  int64_t n_path;
  const int64_t* path;          /* traverse(domain, path): a permutation of 0..N-1 */
  int32_t minneighbors;
  int32_t maxneighbors;         /* <= 64 after clamping */
  double radius;                /* MetricBall radius; <= 0 means KNearestSearch */
  int32_t search_mode;          /* gsp_search_mode */
  int32_t precision;            /* gsp_precision */
  int32_t device_pointers;      /* 1: data_coords and path are device pointers */
  int32_t pad_;
} gsp_seqsim_problem;

/* Everything `preprocess(rng, process, ::LUSIM, init, domain, data)` needs
 * after `initialize` has placed the data (lusim.jl:53-72). */
typedef struct {
  gsp_grid grid;
  gsp_model model;              /* must be a covariance (lusim.jl:44-50) */
  double mean;
  int64_t n_data;               /* number of data cells */
  const int64_t* dinds;         /* findall(mask): ascending grid ids (lusim.jl:71) */
  const double* z1;             /* data values at dinds (lusim.jl:75) */
  int32_t precision;            /* gsp_lu_precision */
  int32_t device_pointers;      /* 1: dinds and z1 are device pointers */
} gsp_lusim_problem;

typedef struct {
  int64_t n_nodes;              /* N */
  int64_t n_data;
  int32_t minneighbors;         /* after clamping */
  int32_t maxneighbors;         /* after clamping */
  int32_t depth;                /* length of the longest dependency chain */
  int32_t pad_;
  double alpha;                 /* scaling factor of seqsim.jl:162-166 */
  int64_t n_prior;              /* nodes drawn from the prior */
  int64_t n_fallback;           /* nodes whose FP32 factor fell back to FP64 */
  int64_t n_learned;            /* nodes whose weights came from the surrogate */
} gsp_seqsim_info;

/* Per-kernel device time, measured with HIP events recorded on the context
 * stream around each launch while timing is enabled. */
enum {
  GSP_K_PLAN = 0,        /* kriging systems: assembly + Cholesky solves */
  GSP_K_LEVELS = 1,      /* dependency depth */
  GSP_K_SWEEP = 2,       /* dataflow draw sweep */
  GSP_K_TRANSPOSE = 3,   /* node-major -> realization-major output */
  GSP_K_LU_FACTOR = 4,   /* LUSIM assembly + factorisations (whole call) */
  GSP_K_LU_SAMPLE = 5,   /* LUSIM d2 + L22 W and scatter (whole call) */
  GSP_K_SEARCH = 6,      /* neighbour search */
  GSP_K_LEARNED = 7,     /* surrogate inference (learned plans) */
  GSP_K_FFT = 8,         /* FFTSIM transforms (rocFFT through hipFFT) */
  GSP_K_SPECTRAL = 9,    /* FFTSIM phase/amplitude, variance, scaling, conditioning */
  GSP_K_STATS = 10,      /* ensemble statistics */
  GSP_K_LU_TRAIL = 11,   /* LUSIM trailing updates of the Cholesky panels (per launch) */
  GSP_K_COUNT = 12
};
typedef struct {
  double ms[GSP_K_COUNT];       /* accumulated milliseconds */
  int64_t launches[GSP_K_COUNT];
} gsp_timings;

/* Training of the learned kriging-weight surrogate. */
typedef struct {
  int64_t max_samples;          /* kriging systems drawn from the plans (0: all usable) */
  int32_t epochs;
  int32_t batch;                /* samples per step, a multiple of 64 */
  double lr;                    /* peak Adam step (one-cycle: warm-up, cosine decay) */
  double var_weight;            /* weight of the log-variance term in the loss */
  uint64_t seed;
  int32_t n_slots;              /* neighbour slots S: 24, 32, or 0 for 24 unless a
                                   training request asks for more than 24 neighbours */
  int32_t data_slots;           /* 0: nodes with conditioning data among their
                                   neighbours are left to the exact solver; 4: such
                                   nodes with at most 4 data neighbours are training
                                   systems and may be learned */
} gsp_train_options;

/* The domain a trained surrogate serves (gsp_surrogate_get_info). */
typedef struct {
  int32_t family;               /* gsp_family of every request it takes */
  int32_t ndim;
  int32_t n_slots;              /* neighbour slots S (24 or 32): maxneighbors <= S */
  int32_t n_train_requests;     /* requests whose systems it was trained on */
  double sill;                  /* gaussian requests must have this sill */
  double nugget_ratio[2];       /* span of nugget / sill trained on */
  double scale[3][2];           /* span of spacing / range per axis trained on */
  int64_t n_samples;
  int32_t n_data_slots;         /* data neighbours a learned node may have (4, or 0:
                                   trained without data_slots or without such systems) */
  int32_t pad_;
} gsp_surrogate_info;

typedef struct {
  int64_t n_samples;            /* systems trained on */
  int64_t n_candidates;         /* usable systems in the plans */
  int64_t steps;
  double rms_weight_first_epoch;/* RMS weight error over the samples of an epoch */
  double rms_weight_last_epoch;
  double rms_logvar_last_epoch;
  double ms;                    /* wall time of the training loop */
} gsp_train_report;

typedef struct gsp_ctx gsp_ctx;
typedef struct gsp_seqsim_plan gsp_seqsim_plan;
typedef struct gsp_lusim_plan gsp_lusim_plan;
typedef struct gsp_surrogate gsp_surrogate;
typedef struct gsp_fftsim_plan gsp_fftsim_plan;

/* Everything `preprocess(rng, process, ::FFTSIM, init, domain, data)` needs
 * after `initialize` has placed the data (fftsim.jl:54-107). */
typedef struct {
  gsp_grid grid;                /* every side >= 2 (fftsim.jl:79-80) */
  gsp_model model;              /* stationary, one variable (fftsim.jl:60,69) */
  double mean;
  /* conditioning (n_data == 0: unconditional) */
  int64_t n_data;
  const double* data_coords;    /* n_data x ndim: the kriging mean (fftsim.jl:94-100) */
  const double* data_vals;      /* n_data */
  int64_t n_cells;
  const int64_t* cells;         /* findall(mask): ascending grid ids (fftsim.jl:104) */
  int32_t minneighbors;         /* FFTSIM options (fftsim.jl:49-53) */
  int32_t maxneighbors;
  double radius;                /* neighborhood MetricBall radius; <= 0: nothing */
  int32_t device_pointers;      /* data_coords, data_vals, cells on the device */
  int32_t pad_;
} gsp_fftsim_problem;

/* ---- context -------------------------------------------------------------
 * One per (process, device): owns a HIP stream and reusable device workspace.
 * Replaces nothing in the reference (Julia has no device state); the shim
 * creates it lazily and frees it in a finalizer. */
int32_t gsp_abi_version(void);
int32_t gsp_ctx_create(int32_t device, gsp_ctx** out);
int32_t gsp_ctx_destroy(gsp_ctx* ctx);
/* The hipStream_t the context launches on (for events / interop). */
int32_t gsp_ctx_stream(gsp_ctx* ctx, void** stream);
int32_t gsp_ctx_synchronize(gsp_ctx* ctx);
/* Thread-local message of the last failed call on this thread. */
const char* gsp_last_error(void);
/* Kernel timing (no counterpart in the reference, which has no tracing): when
 * enabled, every launch of the kernels listed above is bracketed by HIP events
 * on the context stream; gsp_ctx_timings returns the sums (and clears them
 * when reset != 0). */
int32_t gsp_ctx_enable_timing(gsp_ctx* ctx, int32_t on);
int32_t gsp_ctx_timings(gsp_ctx* ctx, gsp_timings* out, int32_t reset);

/* ---- device memory -------------------------------------------------------
 * For callers that pass device pointers (device_pointers = 1) and have no HIP
 * allocator of their own in the process. No counterpart in the reference.
 * Copies run on the context stream and return when complete. */
int32_t gsp_device_alloc(gsp_ctx* ctx, int64_t bytes, void** ptr);
int32_t gsp_device_free(gsp_ctx* ctx, void* ptr);
int32_t gsp_memcpy_to_device(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes);
int32_t gsp_memcpy_to_host(gsp_ctx* ctx, void* dst, const void* src, int64_t bytes);

/* ---- SEQSIM --------------------------------------------------------------
 * gsp_seqsim_plan_create replaces the realization-invariant part of
 * preprocess + randsingle for SEQSIM (seqsim.jl:50-70 and the search / fit /
 * status / predictprob steps of seqsim.jl:101-131): for every grid node the
 * neighbour set, simple-kriging weights and standard deviation, computed once
 * on the device (kernels: neighbour search + batched Cholesky solves). */
int32_t gsp_seqsim_plan_create(gsp_ctx* ctx, const gsp_seqsim_problem* problem,
                               gsp_seqsim_plan** out);
int32_t gsp_seqsim_plan_info(const gsp_seqsim_plan* plan, gsp_seqsim_info* info);
/* Copy the plan to host arrays indexed by grid id (domain order):
 * nfound[N], nbr[N*kmax] (augmented ids, -1 padded), lambda[N*kmax],
 * sigma[N], flags[N], level[N] (dependency depth). Any pointer may be NULL. */
int32_t gsp_seqsim_plan_export(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t* nfound,
                               int32_t* nbr, double* lambda, double* sigma, uint8_t* flags,
                               int32_t* level);
/* gsp_seqsim_sweep replaces R calls of randsingle(rng, process, ::SEQSIM, ...)
 * (seqsim.jl:72-156, driven by the pmap at field.jl:113-119): the sequential
 * draw x_i = mean + sum_j lambda_ij (x_nbr(j) - mean) + sigma_i z over the path.
 *   z:          R x n_path standard normals, realization-major in path order,
 *               i.e. exactly the order of the reference's `randn` calls
 *               (one per path node, realization after realization).
 *   data_vals:  n_data values of the conditioning data (ExplicitInit onto the
 *               augmented domain, seqsim.jl:267-271).
 *   out:        R x N, realization-major, domain order.
 * device_pointers = 1: z, data_vals and out are device pointers and the call
 * returns once the work is queued on the context stream. */
int32_t gsp_seqsim_sweep(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                         const double* z, const double* data_vals, double* out,
                         int32_t device_pointers);
int32_t gsp_seqsim_plan_free(gsp_seqsim_plan* plan);
/* gsp_seqsim_sweep_noalloc: no counterpart in the reference. It is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): gsp_seqsim_sweep with
 * device pointers (z, data_vals and out on the device, same layouts), whose
 * scratch lives in a fixed 64 MiB device array per GPU instead of the
 * context's pool. It calls no allocator. Plans with 2 N + n_data above
 * 8,388,576 are rejected with GSP_ERR_ARGUMENT; larger R runs in chunks of
 * realizations. Returns once the work is queued on the context stream. All
 * calls on one GPU share the array: calls that overlap on one GPU (two
 * contexts' streams) race. */
int32_t gsp_seqsim_sweep_noalloc(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                 const double* z, const double* data_vals, double* out);
/* gsp_seqsim_sweep_multivariate replaces R calls of randsingle(..., ::SEQSIM,
 * ...) for a GaussianProcess with nvars variables whose function is an
 * intrinsic coregionalization A * c(h) (seqsim.jl:72-156 with the MvNormal
 * prior of seqsim.jl:287 and GeoStatsModels' cokriging). With every variable
 * simulated at every node and known at every datum, simple cokriging with
 * A * c(h) has the weights lambda_j * I of the scalar system of c(h) and the
 * error covariance A * sigma^2, so the plan is the one built with c(h) and
 * each node draws x = mu_SK + sigma L z with L the Cholesky factor of A
 * (MvNormal's rand: nvars normals per node, z then L z) and mu_SK,v =
 * mean_v + sum_j lambda_j (x_j,v - mean_v); prior nodes x = mean + sigma L z.
 *   A:         nvars x nvars, row-major, symmetric positive definite (host memory)
 *   mean:      nvars process means (host memory)
 *   z:         R x n_path x nvars standard normals (draw order)
 *   data_vals: nvars x n_data, one row of data values per variable
 *   out:       nvars x R x N (variable-major; each variable R x N realization-major)
 * The plan must hold one request; 1 <= nvars <= 4. GSP_ERR_POSDEF if A is not
 * positive definite. */
int32_t gsp_seqsim_sweep_multivariate(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                      int32_t nvars, const double* A, const double* mean,
                                      const double* z, const double* data_vals, double* out,
                                      int32_t device_pointers);
/* gsp_seqsim_sweep_indicator replaces R calls of randsingle(..., ::SEQSIM, ...)
 * for an IndicatorProcess (seqsim.jl:247-251,292-310,342-347) whose function is
 * an intrinsic coregionalization A * c(h): with every category observed at
 * every neighbour, simple cokriging gives the weights lambda_j * I of the
 * plan built with c(h), so the plan is shared with the Gaussian case. Per node
 * and category p_c + sum_j lambda_j ([x_j == c] - p_c), clamped to [0, 1] and
 * scaled by the inverse of its sum (normalize(., 1)), then the categorical
 * draw: the first category whose running sum exceeds u; prior draws (fewer
 * than minneighbors, failed systems) use `prob`.
 *   prob:     n_cat a priori probabilities (2 <= n_cat <= 8)
 *   u:        R x n_path uniforms in [0, 1), one per path node, path order
 *   data_cat: n_data 0-based categories of the conditioning data
 *   out:      R x N 0-based categories, realization-major. */
int32_t gsp_seqsim_sweep_indicator(gsp_ctx* ctx, const gsp_seqsim_plan* plan, int32_t R,
                                   int32_t n_cat, const double* prob, const double* u,
                                   const int32_t* data_cat, int32_t* out,
                                   int32_t device_pointers);

/* ---- SEQSIM batches --------------------------------------------------------
 * gsp_seqsim_plan_create_batch builds one plan for n independent SEQSIM
 * requests, each with its own grid, variogram, data, path and neighbourhood:
 * the work of n calls of preprocess(rng, process, ::SEQSIM, ...) (seqsim.jl:
 * 50-70), which the reference runs one process at a time. The requests share
 * one global index space: request p owns node ids and path positions
 * [node_offset[p], node_offset[p+1]) and data ids [data_offset[p],
 * data_offset[p+1]) (see gsp_seqsim_plan_layout); every kernel runs once over
 * the whole batch. All requests must use one `precision` and one kind of
 * pointers (`device_pointers`). A plan made by gsp_seqsim_plan_create is a
 * batch of one.
 *
 * For a batch, gsp_seqsim_sweep takes z as R x N_total (request p's path
 * position q in column node_offset[p] + q), data_vals as the requests' data
 * values one after another (n_data_total), and writes out as R x N_total
 * (request p's grid id i in column node_offset[p] + i); every request draws
 * R realizations. gsp_seqsim_plan_export returns the global arrays: neighbour
 * ids are global node ids, or N_total + global data id, in rows of stride
 * info.maxneighbors (the largest maxneighbors of the batch). */
int32_t gsp_seqsim_plan_create_batch(gsp_ctx* ctx, int32_t n,
                                     const gsp_seqsim_problem* problems,
                                     gsp_seqsim_plan** out);
/* n_requests: number of requests; node_offset / data_offset: n + 1 entries
 * each (either may be NULL). */
int32_t gsp_seqsim_plan_layout(const gsp_seqsim_plan* plan, int32_t* n_requests,
                               int64_t* node_offset, int64_t* data_offset);

/* ---- learned kriging weights ----------------------------------------------
 * No counterpart in the reference: a surrogate for the fit / predictprob
 * steps of seqsim.jl:124-131. A BF16 multilayer perceptron 3S+1 -> 256 ->
 * 256 -> 256 -> S+1 (GELU; S = 24 slots: 73 -> ... -> 25, or S = 32: 97 ->
 * ... -> 33) maps a node's neighbour offsets (times the grid spacing, divided
 * by the ranges, in slots given by the rank of the offset in the search's
 * sorted offset table) and the model's nugget / sill to its S kriging weights
 * and log kriging variance. Trained with data_slots = 4 on conditional plans,
 * it also takes up to 4 conditioning data among a node's neighbours (their
 * offsets from the node over the ranges, and the distance to the nearest of
 * them: 13 more inputs) and gives their weights (4 more outputs): 86 -> ...
 * -> 29 for S = 24, 110 -> ... -> 37 for S = 32.
 *
 * gsp_surrogate_train draws training systems from exact plans: every node,
 * solved exactly with its grid neighbours among the S slot offsets (and, with
 * data_slots = 4, at most 4 data neighbours), of the requests compatible with the first usable
 * one (k nearest then mask, unrotated, maxneighbors <= S, at least S table
 * offsets): same variogram family, grid dimension and slot offsets (and sill,
 * for the gaussian family), any range, nugget and maxneighbors. The surrogate
 * then serves requests inside the spans of nugget / sill and spacing / range
 * it was trained on. */
int32_t gsp_surrogate_train(gsp_ctx* ctx, int32_t n_plans, const gsp_seqsim_plan* const* plans,
                            const gsp_train_options* options, gsp_surrogate** out,
                            gsp_train_report* report);
int32_t gsp_surrogate_free(gsp_surrogate* surrogate);
/* The surrogate's FP32 master parameters, layer by layer: W1 [256][128]
 * (inputs beyond the used ones zero), b1 [256], W2 [256][256], b2,
 * W3 [256][256], b3, W4 [64][256] (outputs beyond the used ones zero), b4
 * [64]; and its S slot offsets in grid cells (slots[32][3], rows S..31 zero).
 * Input k < 3S of a node is, when slot k / 3 holds a neighbour,
 * 8 * slots[k / 3][d] * spacing[d] / range[d] (d = k % 3), else 0; input 3S
 * is 10 * nugget / sill; inputs 3S+1+3q+d (q < 4, d < 3) are 8 * (x_q[d] -
 * c[d]) / range[d] for the q-th datum among the node's neighbours (neighbour
 * order) at x_q, c the node's centre, and input 3S+13 is the length of the
 * first of these offsets (all 0 without data neighbours). Outputs 0..S-1 are
 * the weights by slot, output S is log(sigma^2 / sill), outputs S+1+q the
 * weights of the data neighbours. Inference uses the weights rounded to BF16.
 * Either pointer may be NULL; *count receives the number of parameters. */
int32_t gsp_surrogate_export(gsp_ctx* ctx, const gsp_surrogate* surrogate, float* params,
                             int64_t* count, double* slots);
int32_t gsp_surrogate_get_info(const gsp_surrogate* surrogate, gsp_surrogate_info* info);
/* gsp_seqsim_plan_create_learned: gsp_seqsim_plan_create_batch whose nodes
 * take their weights and sigma from the surrogate where it applies (requests
 * inside its domain, see gsp_surrogate_get_info; nodes whose neighbours all
 * lie in its S slot offsets). A node is solved
 * exactly instead when the variance implied by the predicted weights,
 * sill - sum_j lambda_j C(x_j, x_0), differs from the predicted variance by
 * more than `tolerance` (relative). Learned nodes carry GSP_FLAG_LEARNED;
 * gsp_seqsim_plan_info counts them in n_learned. */
int32_t gsp_seqsim_plan_create_learned(gsp_ctx* ctx, int32_t n,
                                       const gsp_seqsim_problem* problems,
                                       const gsp_surrogate* surrogate, double tolerance,
                                       gsp_seqsim_plan** out);
/* The same with several surrogates (1..64, for example one per variogram
 * family of a mixed batch): each request takes its weights from the first
 * surrogate whose domain holds it, or is solved exactly. */
int32_t gsp_seqsim_plan_create_learned_multi(gsp_ctx* ctx, int32_t n,
                                             const gsp_seqsim_problem* problems,
                                             int32_t n_surrogates,
                                             const gsp_surrogate* const* surrogates,
                                             double tolerance, gsp_seqsim_plan** out);

/* ---- LUSIM ---------------------------------------------------------------
 * gsp_lusim_factor replaces preprocess(rng, process, ::LUSIM, ...) for one
 * variable (lusim.jl:66-107): C22, C11, C12 from the covariance (utils.jl:50-62),
 * L11 = chol(C11), B12 = L11 \ C12, d2 = B12' (L11 \ z1),
 * L22 = chol(C22 - B12' B12), or L22 = chol(C22) without data.
 * Returns GSP_ERR_ARGUMENT for a variogram (lusim.jl:44-50) and
 * GSP_ERR_POSDEF when a factorisation meets a non-positive pivot. */
int32_t gsp_lusim_factor(gsp_ctx* ctx, const gsp_lusim_problem* problem,
                         gsp_lusim_plan** out);
/* gsp_lusim_sample replaces R calls of randsingle(..., ::LUSIM, ...) / _lusim
 * (lusim.jl:112-126,145-175): z2 = d2 + L22 w, scatter z[dinds] = z1,
 * z[sinds] = z2, + mean only without data (lusim.jl:172).
 *   w:   R x n_s standard normals, realization-major (lusim.jl:160)
 *   out: R x N, realization-major, domain order. */
int32_t gsp_lusim_sample(gsp_ctx* ctx, const gsp_lusim_plan* plan, int32_t R,
                         const double* w, double* out, int32_t device_pointers);
/* n_s (number of simulated cells) and n_data of a factor. */
int32_t gsp_lusim_dims(const gsp_lusim_plan* plan, int64_t* n_sim, int64_t* n_data);
/* Copy L22 (n_s x n_s, row-major lower, upper zeroed) and d2 (n_s); NULL skips. */
int32_t gsp_lusim_export(gsp_ctx* ctx, const gsp_lusim_plan* plan, double* L22, double* d2);
int32_t gsp_lusim_plan_free(gsp_lusim_plan* plan);

/* ---- FFTSIM --------------------------------------------------------------
 * gsp_fftsim_plan_create replaces preprocess(rng, process, ::FFTSIM, ...)
 * (fftsim.jl:54-107): C = covariances between the centroid of cell dims .÷ 2
 * (1-based, column-major) and every centroid, F = sqrt.(abs.(fft(fftshift(C)))),
 * F[1] = 0; with data, the simple-kriging mean zbar of the data over the grid
 * (GeoStatsModels fitpredict with the FFTSIM options; missing -> NaN) and the
 * kriging weights of the data cells. */
int32_t gsp_fftsim_plan_create(gsp_ctx* ctx, const gsp_fftsim_problem* problem,
                               gsp_fftsim_plan** out);
/* gsp_fftsim_sample replaces R calls of randsingle(..., ::FFTSIM, ...)
 * (fftsim.jl:109-156): P = F .* exp.(im .* angle.(fft(w))), Z = real(ifft(P)),
 * Z = sqrt(sill / var(Z, mean=0)) Z + mean; with data
 * z = zbar + (Z - krige(Z[cells])).
 *   w:   R x N uniform [0, 1) white noise, realization-major, column-major grid
 *        order (fftsim.jl:124)
 *   out: R x N, realization-major. */
int32_t gsp_fftsim_sample(gsp_ctx* ctx, const gsp_fftsim_plan* plan, int32_t R,
                          const double* w, double* out, int32_t device_pointers);
/* Copy F on the half spectrum ((nx/2 + 1) x ny x nz, first axis fastest; the
 * other half is its mirror image) and zbar (N; conditional plans only). NULL
 * skips. */
int32_t gsp_fftsim_export(gsp_ctx* ctx, const gsp_fftsim_plan* plan, double* amplitude,
                          double* zbar);
int32_t gsp_fftsim_plan_free(gsp_fftsim_plan* plan);

/* ---- ensemble statistics -------------------------------------------------
 * Per-cell statistics across the R realizations of an ensemble
 * (ensembles.jl:42-52 through ereduce, :76-85), computed where the
 * realizations already are (the R x N realization-major output of the
 * samplers), several in one pass:
 *   GSP_STAT_MEAN      Statistics.mean: left-to-right sum / R
 *   GSP_STAT_VAR       Statistics.var (corrected, Welford's recurrence)
 *   GSP_STAT_CDF       count(<= param) / R
 *   GSP_STAT_CCDF      count(> param) / R
 *   GSP_STAT_QUANTILE  Statistics.quantile(vals, param), type 7: h = fma(R, p,
 *                      1 - p), j = clamp(trunc(h), 1, R - 1), g = clamp(h - j,
 *                      0, 1), a + g (b - a) of the j-th and (j+1)-th smallest;
 *                      NaN where a value is NaN (the reference throws)
 * kinds, params: n_stats entries (1..32, host memory); out: n_stats x N. */
typedef enum {
  GSP_STAT_MEAN = 0,
  GSP_STAT_VAR = 1,
  GSP_STAT_CDF = 2,
  GSP_STAT_CCDF = 3,
  GSP_STAT_QUANTILE = 4
} gsp_stat;
int32_t gsp_ensemble_stats(gsp_ctx* ctx, int32_t R, int64_t N, const double* reals,
                           int32_t n_stats, const int32_t* kinds, const double* params,
                           double* out, int32_t device_pointers);

#ifdef __cplusplus
}
#endif
#endif /* GSP_H */
