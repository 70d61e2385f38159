// gsp_seqsim.h — SEQSIM internals shared by the plan builder and the learned
// kriging-weight surrogate (gfx950 only).
#pragma once

#include "gsp_internal.h"

namespace gsp {

// One request of a plan, as the kernels see it.
struct Req {
  int32_t ndim, mode, kmin, kmax;
  int32_t dims[3], T;    // grid size; entries of the offset table
  int32_t hd[3], hexp;   // data-hash cells per axis; 1: expanded cell lists
  int32_t hB, n_data;    // hash cell width in grid cells; conditioning data
  int32_t node_off, data_off;
  int32_t learned;       // s + 1: the plan's surrogate s may give this request's weights
  int32_t tile_off;      // data-only tile search: first tile of this request
  int64_t table_off, cell_off;  // first entry of the offset table / of hstart
  double r2, alpha, prior_sigma, mean;
  double trho;           // data-only tile search: first staging radius (physical)
  double org[3], sp[3];  // physical grid
  double ao[3], as[3];   // scaled grid (alpha*origin, alpha*spacing)
  double hcs[3];         // data-hash cell size (physical)
  DevModel M;
};

// Neighbour slots of the learned surrogate: the first kSurrogateSlots entries
// of a request's offset table.
constexpr int kSurrogateSlots = 32;

// Host-side description of one request, kept with the plan.
struct ReqInfo {
  gsp_model model;
  int32_t ndim = 0, mode = 0, kmin = 0, kmax = 0, T = 0;
  double spacing[3] = {1.0, 1.0, 1.0};
  int16_t table32[kSurrogateSlots][3] = {};  // the first min(T, 32) offsets of the offset table
};

// Kernel-side view of a plan's per-node results (the outputs of the kriging
// step and the inputs of the dependency levels).
struct NodeOut {
  int32_t NT, K;
  const Req* reqs;
  const int32_t* node_req;
  const Offset* table;
  const int32_t* nfound;
  const int32_t* nbr;
  const uint8_t* nbr_slot;  // learned plans: table index of each grid neighbour, 255 otherwise
  const double* data;       // n_data_total x 3, physical coordinates of the data
  double* lam;
  double* sigma;
  uint8_t* flags;
  int32_t* counters;
};

// Kernel arguments of the plan builder (search, kriging, degree counts).
struct PlanArgs {
  int32_t NT;  // nodes of all requests; data ids start here
  int32_t K;   // row stride of nbr / lam (largest kmax)
  const Req* reqs;
  const int32_t* node_req;  // request of every node
  const int64_t* path;      // (request << 32) | global node id, in global path order
  const int32_t* pos;       // global path position of every node
  const Offset* table;
  const double* data;       // n_data_total x 3, physical
  const int32_t* hstart;    // absolute item index of every hash cell
  const int32_t* hitems;    // hash-cell lists: global data index of every entry
  int32_t* nfound;
  int32_t* nbr;
  uint8_t* nbr_slot;  // optional: table index of each grid neighbour (255: datum or >= 255)
  double* lam;
  double* sigma;
  uint8_t* flags;
  int32_t* counters;
  // optional: the dependency degrees, counted by the search as it writes each
  // node's neighbours (as the degree count of the level build does)
  int32_t* outdeg;
  int32_t* indeg;
  uint8_t* mark;
};


enum {
  CTR_PRIOR = 2,
  CTR_FALLBACK = 3,
  CTR_POS_ERR = 5,
  CTR_LEARNED = 6,
  CTR_EXACT = 7,
  CTR_FAIL = 8  // kriging systems that were not positive definite
};

// A table the kernels only read (request descriptors), viewed through the
// constant address space so that wave-uniform reads become scalar loads.
template <class T>
__device__ inline const __attribute__((address_space(4))) T* const_view(const T* p) {
  return (const __attribute__((address_space(4))) T*)p;
}

}  // namespace gsp

struct gsp_seqsim_plan {
  std::shared_ptr<gsp::Pool> pool;  // declared first: outlives the buffers below
  int64_t N = 0, n_data = 0;  // nodes / data of all requests
  int32_t n_req = 0;
  int32_t ndim = 0, kmin = 0, kmax = 0, depth = 0;  // kmax: row stride of nbr / lam
  double alpha = 0;  // scale factor of the first request
  double mean = 0;
  int64_t n_prior = 0, n_fallback = 0, n_learned = 0;
  std::vector<int64_t> node_off, data_off;  // n_req + 1 entries each
  std::vector<gsp::ReqInfo> info;           // per request
  gsp::DevBuf<gsp::Req> reqs;               // per request, as the kernels see it
  gsp::DevBuf<gsp::Offset> table;           // the requests' offset tables
  gsp::DevBuf<int32_t> node_req;            // request of every node
  gsp::DevBuf<double> req_mean;             // process mean of every request
  gsp::DevBuf<int32_t> nfound, nbr, level;
  gsp::DevBuf<double> lam, sigma;
  gsp::DevBuf<uint8_t> flags;
  gsp::DevBuf<int32_t> pos;     // path position of every node
  gsp::DevBuf<int32_t> order;   // nodes grouped by level
  gsp::DevBuf<double> data_xyz; // n_data x 3 physical data coordinates (padded to 3)
  std::vector<int64_t> level_off;  // level l holds order[level_off[l] .. level_off[l+1])
  // the last sweep captured as a HIP graph, replayed while every pointer and
  // size its kernels were captured with stays the same
  struct SweepKey {
    int32_t R = 0, RC = 0;
    const void *z = nullptr, *vals = nullptr, *out = nullptr, *X = nullptr, *Zt = nullptr;
    bool operator==(const SweepKey& o) const {
      return R == o.R && RC == o.RC && z == o.z && vals == o.vals && out == o.out && X == o.X &&
             Zt == o.Zt;
    }
  };
  mutable hipGraphExec_t sweep_graph = nullptr;
  mutable hipStream_t sweep_stream = nullptr;
  mutable SweepKey sweep_key;
  // (an executable graph still in flight is freed when it completes; the
  // plan's buffers go back to the context pool, which reuses them only for
  // work queued after it on the same stream)
  void drop_sweep_graph() const {
    if (!sweep_graph) return;
    (void)hipGraphExecDestroy(sweep_graph);
    sweep_graph = nullptr;
  }
  ~gsp_seqsim_plan() { drop_sweep_graph(); }
};

namespace gsp {

// The learned path of the plan builder (gsp_surrogate.hip): runs surrogate
// `s` on every node of the requests with learned == which, writes the
// accepted nodes' weights, sigmas and flags, and lists the other nodes of
// those requests (and, with list_plain, every node of a request with
// learned == 0) in `exact_list` (count in counters[CTR_EXACT]). `feat` holds
// each request's input row (kSurrogateInputs floats) and `cslot` the
// covariances C(slot offset) / sill of its slots (kSurrogateSlots floats).
constexpr int kSurrogateInputs = 128;
void surrogate_infer(gsp_ctx* ctx, const gsp_surrogate* s, int32_t which, bool list_plain,
                     const NodeOut& out, double tolerance, int32_t* exact_list, const float* feat,
                     const float* cslot);
// Whether request `r` lies in the surrogate's domain: its family, no
// rotation, k nearest then mask, maxneighbors <= 32, the surrogate's slot
// offsets as the start of its offset table, a nugget ratio and grid spacing in
// range units inside the spans trained on (and, for the gaussian family, the
// same sill).
bool surrogate_accepts(const gsp_surrogate* s, const ReqInfo& r);
// The input row and slot covariances of an accepted request.
void surrogate_request_rows(const gsp_surrogate* s, const ReqInfo& r, float* feat, float* cslot);
}  // namespace gsp
