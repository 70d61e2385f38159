// gsp_surrogate.hip — learned kriging-weight surrogate for SEQSIM on gfx950.
//
// A multilayer perceptron 3S+1 -> 256 -> 256 -> 256 -> S+1 (GELU), S = 24 or
// 32 neighbour slots (73 -> ... -> 25 or 97 -> ... -> 33), maps a node's
// neighbourhood to its S simple-kriging weights and log kriging variance,
// with BF16 weights and activations and FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x32_bf16). Inputs: S slots of neighbour offsets (dx, dy,
// dz times the grid spacing, divided by the ranges), zero for an absent
// neighbour, and the model's nugget / sill. A neighbour's slot is the rank of
// its offset in the request's sorted offset table, so an interior node of a
// k-nearest-then-mask search fills slots among the same S offsets and the
// input is the pattern of already simulated neighbours, scaled by the range.
// A surrogate trained on conditional plans also takes up to kDataSlots
// conditioning data among the neighbours: their offsets from the node (in
// ranges) and the distance to the nearest of them enter 3 kDataSlots + 1 more
// inputs, and their weights are kDataSlots more outputs (86 -> ... -> 29 for
// S = 24, 110 -> ... -> 37 for S = 32).
// One surrogate serves one variogram family and grid dimension over the spans
// of ranges (in grid cells) and nugget ratios it was trained on, so requests
// of a batch with different ranges and nuggets share it. Nodes whose
// neighbours do not fit (grid boundary, more data than data slots, or data at
// all for a surrogate trained without them) and nodes whose predicted
// variance disagrees with the variance implied by the predicted weights are
// solved exactly.
//
// Training runs on the device: a BF16 NT GEMM with fused bias/GELU epilogues
// (forward), GELU-derivative epilogues (backward), weight gradients as NT
// GEMMs over the batch, and Adam on FP32 master weights.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "gsp_seqsim.h"
#include "gsp_wave.h"

namespace gsp {
namespace {

constexpr int kIn = kSurrogateInputs;  // input features, 3 S + 1 used
constexpr int kHid = 256;
constexpr int kOutT = 64;    // output rows in training GEMMs, S + 1 used
constexpr int kSlotsMax = kSurrogateSlots;
// output rows at inference (S + 1 used): whole 16-row MFMA tiles
__host__ __device__ constexpr int out_rows(int slots) { return slots <= 31 ? 32 : 48; }
constexpr int kOutIMax = out_rows(kSlotsMax);
constexpr float kFeatScale = 8.0f;   // offsets in ranges x 8: inputs of order one
constexpr float kNugScale = 10.0f;   // nugget / sill x 10
// Data slots: the first kDataSlots conditioning data in a node's neighbour
// list (nearest first). Inputs 3 S + 1 .. 3 S + 3 kDataSlots: their offsets
// datum - node in ranges x kFeatScale; input 3 S + 3 kDataSlots + 1: the
// length of the first offset (the nearest datum among the neighbours); zero
// where absent. Outputs S + 1 .. S + kDataSlots: their weights.
constexpr int kDataSlots = 4;
constexpr int kDFeat = 3 * kDataSlots + 1;
static_assert(3 * kSlotsMax + 1 + kDFeat <= kIn && kSlotsMax + 1 + kDataSlots <= kOutIMax &&
                  kOutIMax <= kOutT,
              "network shapes");

// FP32 master parameters, one buffer
constexpr int64_t OW1 = 0, OB1 = OW1 + (int64_t)kHid * kIn, OW2 = OB1 + kHid,
                  OB2 = OW2 + (int64_t)kHid * kHid, OW3 = OB2 + kHid,
                  OB3 = OW3 + (int64_t)kHid * kHid, OW4 = OB3 + kHid,
                  OB4 = OW4 + (int64_t)kOutT * kHid, kParams = OB4 + kOutT;
// BF16 weights [out][in] and [in][out]
constexpr int64_t BW1 = 0, BW2 = BW1 + (int64_t)kHid * kIn, BW3 = BW2 + (int64_t)kHid * kHid,
                  BW4 = BW3 + (int64_t)kHid * kHid, kBf = BW4 + (int64_t)kOutT * kHid;
// BF16 weights in the fragment order of the inference kernel
constexpr int64_t PW1 = 0, PW2 = PW1 + (int64_t)kHid * kIn, PW3 = PW2 + (int64_t)kHid * kHid,
                  PW4 = PW3 + (int64_t)kHid * kHid, kPermMax = PW4 + (int64_t)kOutIMax * kHid;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite)
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
// the same with operands held as packed pairs (item 2w in the low half of word w)
__device__ inline f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> packed bf16 pair (round to nearest even; v_cvt_pk_bf16_f32)
__device__ inline uint32_t pack2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}

// GELU in its tanh form, 0.5 z (1 + tanh(u)) = z sigmoid(2u) with
// u = sqrt(2/pi) (z + 0.044715 z^3), i.e. z / (1 + exp(-z (c1 + c2 z^2)));
// the same function in training and inference.
constexpr float kGeluC1 = 1.5957691216057308f;   // 2 sqrt(2/pi)
constexpr float kGeluC2 = 0.0713548162726009f;   // 2 sqrt(2/pi) 0.044715
constexpr float kLog2e = 1.4426950408889634f;
__device__ inline float gelu_sig(float z) {
  // exp(-s) as 2^(-s log2 e), log2 e folded into the constants (v_exp_f32)
  const float t = z * ((-kGeluC1 * kLog2e) + (-kGeluC2 * kLog2e) * (z * z));
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(t));
}
__device__ inline float gelu(float z) { return z * gelu_sig(z); }
__device__ inline float gelu_grad(float z) {
  const float sg = gelu_sig(z);
  return sg + z * sg * (1.0f - sg) * (kGeluC1 + 3.0f * kGeluC2 * (z * z));
}

// Order of the reduction index inside a fragment of the inference kernel:
// lane group g, item i of step s holds index k. Items 0-3 are the rows
// 4g..4g+3 of output tile 2s of the previous layer and items 4-7 those of
// tile 2s + 1, so one layer's accumulators become the next layer's operands
// without leaving the registers.
__host__ __device__ inline int frag_k(int s, int g, int i) {
  return 32 * s + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// Input feature k of a node whose occupied slots are `mask`; `row` is its
// request's input row (slot offsets in ranges x kFeatScale, then the nugget
// feature, then zeros) and `dfeat` its kDFeat data features.
// DATA = false: no data features (the row is zero beyond input 3S).
template <bool DATA>
__device__ inline float feature(const float* row, uint32_t mask, int k, int slots,
                                const float* dfeat) {
  if (k < 3 * slots) return ((mask >> (k / 3)) & 1u) ? row[k] : 0.0f;
  if constexpr (DATA) {
    if (k > 3 * slots) {
      const int q = k - 3 * slots - 1;
      return q < kDFeat ? dfeat[q] : 0.0f;
    }
  }
  return row[k];
}

// Data features of node i (kDFeat floats, zero where absent) from its first
// kDataSlots data neighbours, and cd[d] = C(datum d, node) / sill. The
// offsets are taken in the kriging's scaled coordinates (alpha x physical)
// times the reciprocal scaled ranges, which is the physical offset over the
// range. Returns the node's number of data neighbours (possibly more than
// kDataSlots).
__device__ inline int node_data(const NodeOut& O, const Req& R, int i, int n, float* df,
                                float* cd) {
#pragma unroll
  for (int q = 0; q < kDFeat; ++q) df[q] = 0.0f;
#pragma unroll
  for (int d = 0; d < kDataSlots; ++d) cd[d] = 0.0f;
  const int l = i - R.node_off;
  const int nx = R.dims[0], ny = R.dims[1];
  const int c[3] = {l % nx, (l / nx) % ny, l / (nx * ny)};
  double t[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) t[e] = R.ao[e] + ((double)c[e] + 0.5) * R.as[e];
  int nd = 0;
  for (int j = 0; j < n; ++j) {
    const int nb = O.nbr[(int64_t)i * O.K + j];
    if (nb < O.NT) continue;
    if (nd < kDataSlots) {
      const double* p = O.data + (int64_t)(nb - O.NT) * 3;
      double x[3];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        x[e] = e < R.ndim ? R.alpha * p[e] : 0.0;
        df[3 * nd + e] = e < R.ndim ? (float)((x[e] - t[e]) * R.M.irs[e] * kFeatScale) : 0.0f;
      }
      cd[nd] = (float)(cov_points(R.M, x[0], x[1], x[2], t[0], t[1], t[2]) / R.M.sill);
    }
    ++nd;
  }
  if (nd > 0) df[3 * kDataSlots] = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
  return nd;
}

// Slot of grid offset (dx, dy, dz): its index among the first `slots`
// entries of the request's offset table, or -1.
__device__ inline int slot_of(const Offset* tab, int T, int dx, int dy, int dz, int slots) {
  const int n = T < slots ? T : slots;
  for (int s = 0; s < n; ++s)
    if (tab[s].dx == dx && tab[s].dy == dy && tab[s].dz == dz) return s;
  return -1;
}

// The surrogate's view of node i: its slot mask and number of data
// neighbours when every grid neighbour lies in the first `ns` table offsets
// and at most kDataSlots neighbours are data, else false.
__device__ inline bool node_slots(const NodeOut& O, const Req& R, int i, int n, int ns,
                                  uint32_t& mask, int& ndata) {
  const int nx = R.dims[0], ny = R.dims[1];
  const int li = i - R.node_off;
  const int ix = li % nx, iy = (li / nx) % ny, iz = li / (nx * ny);
  const Offset* tab = O.table + R.table_off;
  mask = 0;
  ndata = 0;
  for (int j = 0; j < n; ++j) {
    const int nb = O.nbr[(int64_t)i * O.K + j];
    if (nb >= O.NT) {
      if (++ndata > kDataSlots) return false;
      continue;
    }
    const int l = nb - R.node_off;
    const int s = slot_of(tab, R.T, l % nx - ix, (l / nx) % ny - iy, l / (nx * ny) - iz, ns);
    if (s < 0) return false;
    mask |= 1u << s;
  }
  return true;
}

// ------------------------------------------------------------- training --
enum { EPI_F32 = 0, EPI_GELU = 1, EPI_DGELU = 2 };

struct GemmArgs {
  const uint16_t* A;  // M x K, row stride lda
  const uint16_t* B;  // N x K, row stride ldb
  int64_t lda, ldb;
  int32_t M, N, K;
  float* C;  // EPI_F32: C[m][n] = acc (+ bias[n])
  int64_t ldc;
  const float* bias;
  uint16_t* H;   // EPI_GELU / EPI_DGELU: bf16 result [M][N]
  uint16_t* Ht;  // and its transpose [N][M]
  float* G;      // EPI_GELU: gelu'(z) [M][N]
  const float* Gin;  // EPI_DGELU: multiplies the product elementwise [M][N]
};

// C = A B^T on the matrix cores: 64 x 64 tiles per 256-thread block, 32 x 32
// per wavefront (2 x 2 MFMA tiles), operands straight from global memory (the
// matrices of one training step stay in L2). M, N multiples of 64, K of 32.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(const GemmArgs P) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < P.K; k0 += 32) {
    u16x8 fa[2], fb[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      fa[t] = *(const u16x8*)(P.A + (int64_t)(m0 + 16 * t + r) * P.lda + k0 + 8 * g);
      fb[t] = *(const u16x8*)(P.B + (int64_t)(n0 + 16 * t + r) * P.ldb + k0 + 8 * g);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = mfma(fa[a], fb[b], acc[a][b]);
  }
  // acc[a][b][v] = C[m0 + 16a + 4g + v][n0 + 16b + r]
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int n = n0 + 16 * b + r;
      const int mb = m0 + 16 * a + 4 * g;
      if constexpr (EPI == EPI_F32) {
        const float bb = P.bias ? P.bias[n] : 0.0f;
#pragma unroll
        for (int v = 0; v < 4; ++v) P.C[(int64_t)(mb + v) * P.ldc + n] = acc[a][b][v] + bb;
      } else {
        u16x4 t;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t e = (int64_t)(mb + v) * P.N + n;
          float y;
          if constexpr (EPI == EPI_GELU) {
            const float z = acc[a][b][v] + P.bias[n];
            y = gelu(z);
            P.G[e] = gelu_grad(z);
          } else {
            y = acc[a][b][v] * P.Gin[e];
          }
          const uint16_t h = f2bf(y);
          P.H[e] = h;
          t[v] = h;
        }
        *(u16x4*)(P.Ht + (int64_t)n * P.M + mb) = t;
      }
    }
}

// One minibatch: inputs (both layouts), targets, slot masks and data counts
// of the samples perm[off .. off + Bt).
__global__ void k_train_batch(const int32_t* perm, int64_t off, int32_t Bt, const uint32_t* masks,
                              const int32_t* srow, const uint8_t* ndata, const float* dfeat,
                              const float* Y, const float* rows, int32_t slots, uint16_t* X,
                              uint16_t* Xt, float* Yb, uint32_t* Mb, uint8_t* Nb) {
  const int64_t total = (int64_t)Bt * kIn;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / kIn), k = (int)(t % kIn);
    const int smp = perm[off + b];
    const uint32_t m = masks[smp];
    const uint16_t h = f2bf(feature<true>(rows + (int64_t)srow[smp] * kIn, m, k, slots,
                                          dfeat + (int64_t)smp * kDFeat));
    X[(int64_t)b * kIn + k] = h;
    Xt[(int64_t)k * Bt + b] = h;
    if (k < kOutT) Yb[(int64_t)b * kOutT + k] = Y[(int64_t)smp * kOutT + k];
    if (k == 0) {
      Mb[b] = m;
      Nb[b] = ndata[smp];
    }
  }
}

// Loss = mean over the batch of sum_s occupied (o_s - y_s)^2 + wv (o_S - y_S)^2
// + sum_{d < data count} (o_{S+1+d} - y_{S+1+d})^2; writes dLoss/dO in bf16
// (both layouts) and accumulates the loss terms.
__global__ void k_train_loss(const float* O, const float* Yb, const uint32_t* Mb,
                             const uint8_t* Nb, int32_t Bt, int32_t slots, float wv, uint16_t* dO,
                             uint16_t* dOt, float* acc) {
  __shared__ float sw[256], sv[256], sc[256];
  float lw = 0.0f, lv = 0.0f, cnt = 0.0f;
  const int64_t total = (int64_t)Bt * kOutT;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / kOutT), o = (int)(t % kOutT);
    float d = 0.0f;
    if (o < slots) {
      if ((Mb[b] >> o) & 1u) {
        d = O[t] - Yb[(int64_t)b * kOutT + o];
        lw += d * d;
        cnt += 1.0f;
      }
    } else if (o == slots) {
      const float e = O[t] - Yb[(int64_t)b * kOutT + o];
      lv += e * e;
      d = wv * e;
    } else if (o - slots - 1 < (int)Nb[b]) {
      d = O[t] - Yb[(int64_t)b * kOutT + o];
      lw += d * d;
      cnt += 1.0f;
    }
    const uint16_t h = f2bf(2.0f * d / (float)Bt);
    dO[t] = h;
    dOt[(int64_t)o * Bt + b] = h;
  }
  sw[threadIdx.x] = lw;
  sv[threadIdx.x] = lv;
  sc[threadIdx.x] = cnt;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      sw[threadIdx.x] += sw[threadIdx.x + h];
      sv[threadIdx.x] += sv[threadIdx.x + h];
      sc[threadIdx.x] += sc[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&acc[0], sw[0]);
    atomicAdd(&acc[1], sv[0]);
    atomicAdd(&acc[2], sc[0]);
  }
}

// Bias gradients: row sums of the transposed output gradients, one block per
// row (rows of dZ1^T, dZ2^T, dZ3^T, dO^T in that order).
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* d1t, const uint16_t* d2t,
                                                   const uint16_t* d3t, const uint16_t* d4t,
                                                   int32_t Bt, float* grad) {
  __shared__ float red[256];
  const int row = blockIdx.x;
  const uint16_t* src;
  int64_t dst;
  if (row < kHid) {
    src = d1t + (int64_t)row * Bt;
    dst = OB1 + row;
  } else if (row < 2 * kHid) {
    src = d2t + (int64_t)(row - kHid) * Bt;
    dst = OB2 + row - kHid;
  } else if (row < 3 * kHid) {
    src = d3t + (int64_t)(row - 2 * kHid) * Bt;
    dst = OB3 + row - 2 * kHid;
  } else {
    src = d4t + (int64_t)(row - 3 * kHid) * Bt;
    dst = OB4 + row - 3 * kHid;
  }
  float s = 0.0f;
  for (int b = threadIdx.x; b < Bt; b += blockDim.x) s += bf2f(src[b]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[dst] = red[0];
}

// Layer of a flat parameter index: weights give (out, in, layer's in, its
// bf16 offset); biases return false.
__device__ inline bool weight_of(int64_t p, int& row, int& col, int& nin, int64_t& bf) {
  int64_t base;
  if (p < OB1) { base = OW1; nin = kIn; bf = BW1; }
  else if (p < OW2) return false;
  else if (p < OB2) { base = OW2; nin = kHid; bf = BW2; }
  else if (p < OW3) return false;
  else if (p < OB3) { base = OW3; nin = kHid; bf = BW3; }
  else if (p < OW4) return false;
  else if (p < OB4) { base = OW4; nin = kHid; bf = BW4; }
  else return false;
  row = (int)((p - base) / nin);
  col = (int)((p - base) % nin);
  return true;
}

// Adam on the FP32 master parameters, then their BF16 copies in both layouts.
// Padding (inputs 3S+1..127, outputs S+1..63) keeps zero gradients and stays zero.
__global__ void k_adam(float* params, const float* grad, float* m, float* v, float lr, float b1,
                       float b2, float c1, float c2, float eps, uint16_t* wbf, uint16_t* wbt) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < kParams;
       p += (int64_t)gridDim.x * blockDim.x) {
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float x = params[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    params[p] = x;
    int row, col, nin;
    int64_t bf;
    if (weight_of(p, row, col, nin, bf)) {
      const uint16_t h = f2bf(x);
      const int nout = (bf == BW4) ? kOutT : kHid;
      wbf[bf + (int64_t)row * nin + col] = h;
      wbt[bf + (int64_t)col * nout + row] = h;
    }
  }
}

// BF16 copies of the master weights in the inference fragment order:
// fragment (tile rt, step s), lane l, item i holds W[16 rt + l % 16][frag_k(s, l / 16, i)].
__global__ void k_pack_inference(const float* params, uint16_t* wp, int32_t out_i) {
  const int64_t n = PW4 + (int64_t)out_i * kHid;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int64_t base, po;
    int nin;
    if (t < PW2) { base = PW1; po = OW1; nin = kIn; }
    else if (t < PW3) { base = PW2; po = OW2; nin = kHid; }
    else if (t < PW4) { base = PW3; po = OW3; nin = kHid; }
    else { base = PW4; po = OW4; nin = kHid; }
    const int64_t e = t - base;
    const int steps = nin / 32;
    const int i = (int)(e % 8), lane = (int)((e / 8) % 64);
    const int64_t frag = e / 512;
    const int s = (int)(frag % steps), rt = (int)(frag / steps);
    const int row = 16 * rt + (lane & 15), col = frag_k(s, lane >> 4, i);
    wp[t] = f2bf(params[po + (int64_t)row * nin + col]);
  }
}

// Training samples of one plan: every node of an accepted request that was
// solved exactly with at least one neighbour, all in table slots, drawn with
// probability `keep` (a hash of the node id and the seed). count != nullptr:
// counts the candidates only (count[0]; count[1]: those with data
// neighbours). A sample keeps its slot mask, the input row of its request
// (req_row; -1: request not used), its number of data neighbours and data
// features, and its targets: the weights by slot, log(sigma^2 / sill) and the
// weights of its data neighbours, kOutT floats.
struct ExtractArgs {
  NodeOut O;
  int32_t slots;
  const int32_t* req_row;
  uint32_t keep24;  // acceptance threshold on 24 hash bits
  uint64_t seed;
  int32_t max_data;  // data neighbours a sample may have
  int32_t* count;
  int32_t* cursor;
  int32_t capacity;
  uint32_t* masks;
  int32_t* srow;
  uint8_t* ndata;
  float* dfeat;  // kDFeat per sample
  float* Y;
};

__device__ inline uint32_t mix32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return (uint32_t)x;
}

__global__ void k_extract(const ExtractArgs E) {
  const NodeOut& O = E.O;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < O.NT;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int p = O.node_req[i];
    const int row = E.req_row[p];
    if (row < 0) continue;
    if (O.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL | GSP_FLAG_LEARNED)) continue;
    const int n = O.nfound[i];
    if (n < 1) continue;
    const Req& R = O.reqs[p];
    uint32_t mask;
    int nd;
    if (!node_slots(O, R, (int)i, n, E.slots, mask, nd) || nd > E.max_data) continue;
    if (E.count) {
      atomicAdd(E.count, 1);
      if (nd > 0) atomicAdd(E.count + 1, 1);
      continue;
    }
    if ((mix32(E.seed ^ (uint64_t)i * 0x9E3779B97F4A7C15ULL) & 0xFFFFFFu) >= E.keep24) continue;
    const int slot = atomicAdd(E.cursor, 1);
    if (slot >= E.capacity) continue;
    E.masks[slot] = mask;
    E.srow[slot] = row;
    E.ndata[slot] = (uint8_t)nd;
    float cd[kDataSlots];
    node_data(O, R, (int)i, n, E.dfeat + (int64_t)slot * kDFeat, cd);
    float* y = E.Y + (int64_t)slot * kOutT;
    for (int s = 0; s < kOutT; ++s) y[s] = 0.0f;
    const int nx = R.dims[0], ny = R.dims[1];
    const int li = (int)i - R.node_off;
    const int ix = li % nx, iy = (li / nx) % ny, iz = li / (nx * ny);
    int d = 0;
    for (int j = 0; j < n; ++j) {
      const int nb = O.nbr[i * O.K + j];
      if (nb >= O.NT) {
        y[E.slots + 1 + d++] = (float)O.lam[i * O.K + j];
        continue;
      }
      const int l = nb - R.node_off;
      const int s = slot_of(O.table + R.table_off, R.T, l % nx - ix, (l / nx) % ny - iy,
                            l / (nx * ny) - iz, E.slots);
      y[s] = (float)O.lam[i * O.K + j];
    }
    const double s2 = O.sigma[i] * O.sigma[i] / R.M.sill;
    y[E.slots] = (float)log(s2 > 1e-12 ? s2 : 1e-12);
  }
}

// ------------------------------------------------------------ inference --
struct InferArgs {
  NodeOut O;
  const uint16_t* wp;   // fragment-ordered BF16 weights
  const float* params;  // FP32 master (biases)
  const float* feat;    // input row of every request (kIn floats)
  const float* cslot;   // C(slot offset) / sill of every request (kSlotsMax floats)
  float tol;
  int32_t dslots;       // data neighbours a node may have (the surrogate's data slots)
  int32_t which;        // requests with learned == which are this surrogate's
  int32_t list_plain;   // also list the nodes of requests with learned == 0
  int32_t* exact_list;
};

constexpr int kInferWaves = 8;  // wavefronts per block: one pass of the weights through LDS
// 16-node column tiles per wavefront. Two (170 VGPRs, one block per CU) is
// the measured choice. One (118 VGPRs, two blocks per CU) made inference 37 %
// slower: each weight fragment read from LDS serves half the products. Three
// (256 VGPRs) made it 15 % faster but fills the CU's registers, so the level
// kernels that run beside it on the second stream stalled and the learned
// 256x256x128 step got 9 % slower.
constexpr int kInferTiles = 2;

// Weight fragments of output tiles, staged in LDS for the block's
// wavefronts: three buffers, so that the copy of tile rt + 2 runs while tile
// rt is multiplied and tile rt - 1 goes through its activation. The buffers
// are separate LDS objects, so the compiler can tell that reading one does
// not have to wait for the asynchronous copies into the others.
using FragBuf = u32x4[8][64];
struct FragLds {
  FragBuf& b0;
  FragBuf& b1;
  FragBuf& b2;
  template <int B>
  __device__ __forceinline__ FragBuf& buf() const {
    if constexpr (B == 0) return b0;
    else if constexpr (B == 1) return b1;
    else return b2;
  }
};

// Asynchronous copy of the fragments of one tile into buffer `dst`
// (global_load_lds_dwordx4: 1 KB per fragment, 16 B per lane). Every
// wavefront issues exactly one copy per tile, so that all of them wait for
// the same counts and no copy sits behind a branch: fragment w % STEPS. With
// fewer fragments than wavefronts, kInferWaves / STEPS wavefronts copy the same
// fragment and write identical bytes to the same LDS words (L2 serves the
// repeats). Completion: s_waitcnt vmcnt.
template <int STEPS>
__device__ __forceinline__ void copy_tile(FragBuf& dst, const uint16_t* wp, int tile, int lane,
                                          int w) {
  static_assert(kInferWaves % STEPS == 0, "whole wavefronts per fragment");
  const int s = w % STEPS;
  __builtin_amdgcn_global_load_lds(
      (const __attribute__((address_space(1))) void*)(wp + ((int64_t)(tile * STEPS + s) * 64 + lane) * 8),
      (__attribute__((address_space(3))) void*)&dst[s][0], 16, 0, 0);
}

// s_waitcnt with only the vector-memory count bounded: vmcnt <= n (expcnt and
// lgkmcnt at their "no wait" maxima, gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 16, "vmcnt field");
  __builtin_amdgcn_s_waitcnt(N | (0x7 << 4) | (0xF << 8));
}

// Block barrier for the fragment buffers. __syncthreads() (and any release
// fence) waits for every outstanding vector-memory operation, including the
// weight copy meant to stay in flight across the barrier. Here the caller has
// already waited (wait_vmcnt) for the copies the next step reads, and every
// LDS read of the buffer about to be refilled has been consumed by the
// products before the barrier, so the bare s_barrier suffices; the memory
// clobber keeps the compiler from moving loads or stores across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_barrier" ::: "memory"); }

// One layer of the fused inference: out^T = act(W in^T + b) for the wave's
// 16 CT nodes (CT column tiles); in/out hold the packed bf16 pairs of the
// operand fragments, word q of step s, column tile ct at [s][ct][q]. Every
// wavefront of the block runs it (block barriers inside). Step rt issues the
// bias load of tile rt + 1 and the copy of tile rt + 2 into LDS, then the
// products of tile rt and the activation of tile rt - 1 (the matrix cores
// and the vector ALU overlap). The step ends by waiting for the loads issued
// one step earlier only (loads complete in order; this wavefront's copy of
// tile rt + 2 stays in flight), so each copy has a whole step to land.
template <int STEPS, int TILES, bool ACT, int LT, int CT>
__device__ __forceinline__ void mlp_layer(const FragLds& F, const uint32_t (&in)[8][CT][4],
                                          uint32_t (&out)[8][CT][4], float (&last)[LT][CT][4],
                                          const uint16_t* wp, const float* bias, int lane, int w) {
  const int g = lane >> 4;
  f32x4 b_next = *(const f32x4*)(bias + 4 * g);
  copy_tile<STEPS>(F.template buf<0>(), wp, 0, lane, w);
  if constexpr (TILES > 1) copy_tile<STEPS>(F.template buf<1>(), wp, 1, lane, w);
  if constexpr (TILES > 1)
    wait_vmcnt<1>();
  else
    wait_vmcnt<0>();
  lds_barrier();
  f32x4 acc[2][CT];
  static_for<0, TILES + 1>([&](auto rtc) {
    constexpr int rt = decltype(rtc)::value;
    f32x4 b_cur = b_next;
    if constexpr (rt + 1 < TILES) b_next = *(const f32x4*)(bias + 16 * (rt + 1) + 4 * g);
    if constexpr (rt + 2 < TILES) copy_tile<STEPS>(F.template buf<(rt + 2) % 3>(), wp, rt + 2, lane, w);
    if constexpr (rt < TILES) {
      // the accumulators start at the bias of their output rows
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[rt & 1][ct] = b_cur;
      static_for<0, STEPS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const u32x4 a = F.template buf<rt % 3>()[s][lane];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
          acc[rt & 1][ct] = mfma(a, u32x4{in[s][ct][0], in[s][ct][1], in[s][ct][2], in[s][ct][3]},
                                 acc[rt & 1][ct]);
      });
    }
    if constexpr (rt >= 1) {
      constexpr int pt = rt - 1;
      if constexpr (ACT) {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const f32x4 z = acc[pt & 1][ct];
          uint32_t w0 = pack2(gelu(z[0]), gelu(z[1]));
          uint32_t w1 = pack2(gelu(z[2]), gelu(z[3]));
          // materialise the packed words here: the activation math is pure
          // and would otherwise be sunk to the next layer
          asm volatile("" : "+v"(w0), "+v"(w1));
          out[pt >> 1][ct][(pt & 1) * 2] = w0;
          out[pt >> 1][ct][(pt & 1) * 2 + 1] = w1;
        }
      } else {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) last[pt][ct][v] = acc[pt & 1][ct][v];
      }
    }
    if constexpr (rt < TILES) {
      // the copy of tile rt + 1 (and the bias of rt + 1) must have landed;
      // this wavefront's copy of tile rt + 2, issued above, may stay in flight
      if constexpr (rt + 2 < TILES)
        wait_vmcnt<1>();
      else
        wait_vmcnt<0>();
      lds_barrier();
    }
    __builtin_amdgcn_sched_barrier(0);
  });
}

// Lane layout of v_mfma_f32_16x16x32_bf16 as the kernels above use it:
// A[i][k] in lane i + 16 (k / 8), item k % 8; B[k][j] in lane j + 16 (k / 8);
// D[i][j] in lane j + 16 (i / 4), register i % 4. Setup 0: A[i][k] = i + 1,
// B = 1; setup 1: A = 1, B[k][j] = j + 1; setup 2: one nonzero product per
// element, at (lane group 0, item 0) of both operands.
__global__ __launch_bounds__(64) void k_mfma_probe_bf16(double* out) {
  const int lane = threadIdx.x;
  for (int setup = 0; setup < 3; ++setup) {
    u32x4 a, b;
    for (int w = 0; w < 4; ++w) {
      float a0, a1, b0, b1;
      if (setup == 0) {
        a0 = a1 = (float)(lane % 16 + 1);
        b0 = b1 = 1.0f;
      } else if (setup == 1) {
        a0 = a1 = 1.0f;
        b0 = b1 = (float)(lane % 16 + 1);
      } else {
        const bool first = lane < 16 && w == 0;
        a0 = first ? 1.0f : 0.0f;
        b0 = first ? (float)(lane % 16 + 1) : 0.0f;
        a1 = b1 = 0.0f;
      }
      a[w] = pack2(a0, a1);
      b[w] = pack2(b0, b1);
    }
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = mfma(a, b, acc);
    for (int v = 0; v < 4; ++v) out[(setup * 64 + lane) * 4 + v] = (double)acc[v];
  }
}

// CT column tiles of 16 nodes per wavefront: every weight fragment read from
// LDS serves CT products, and the two activation buffers of a layer take
// 2 x 32 CT registers per lane. DATA: the surrogate has data slots (nodes
// with data neighbours may be learned); without, the kernel carries none of
// their features, LDS or outputs.
template <int SLOTS, int CT, bool DATA>
__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const InferArgs A) {
  __builtin_amdgcn_s_setprio(3);  // issue ahead of the level kernels running beside it
  constexpr int kD = DATA ? kDataSlots : 0;
  constexpr int kOutI = out_rows(SLOTS), kUsedOut = SLOTS + 1 + kD, kNodes = 16 * CT;
  __shared__ FragBuf s_w0, s_w1, s_w2;
  const FragLds s_frag{s_w0, s_w1, s_w2};
  __shared__ uint32_t s_mask[kInferWaves][kNodes];
  __shared__ int32_t s_req[kInferWaves][kNodes];
  __shared__ int8_t s_slot[kInferWaves][kNodes][SLOTS];  // output of each neighbour's weight
  __shared__ float s_dfeat[kInferWaves][kNodes][DATA ? kDFeat : 1];
  __shared__ float s_cd[kInferWaves][kNodes][DATA ? kDataSlots : 1];
  __shared__ float s_out[kInferWaves][kNodes][kUsedOut + 1];
  __shared__ int32_t s_any;
  const NodeOut& O = A.O;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // one group of kNodes nodes per wavefront (no grid-stride loop: the weight
  // addresses must not become loop invariants held in registers)
  const int64_t base = ((int64_t)blockIdx.x * kInferWaves + w) * kNodes;
  if (tid == 0) s_any = 0;
  __syncthreads();
  // ---- slots of the wave's nodes
  int state = -1;  // -1: not this launch's node, 0: exact solve, 1: learned candidate, 2: learned
  const int64_t i = base + lane;
  int p = 0;
  int ndl = 0;  // data neighbours of a learned candidate
  if (lane < kNodes && i < O.NT) {
    p = O.node_req[i];
    const Req& R = O.reqs[p];
    const int n = O.nfound[i];
    if (R.learned == A.which || (A.list_plain && R.learned == 0)) state = 0;
    if (R.learned == A.which && n >= R.kmin && n >= 1) {
      // slots from the table indices the search recorded; data neighbours
      // (table index 255) take the data slots in list order
      uint32_t mask = 0;
      bool fits = true;
      int nd = 0;
      for (int j = 0; j < n; ++j) {
        const int t = O.nbr_slot[i * O.K + j];
        if (t < SLOTS) {
          mask |= 1u << t;
          s_slot[w][lane][j] = (int8_t)t;
        } else if (DATA && nd < A.dslots && O.nbr[i * O.K + j] >= O.NT) {
          s_slot[w][lane][j] = (int8_t)(SLOTS + 1 + nd);
          ++nd;
        } else {
          fits = false;
        }
      }
      if (fits) {
        state = 1;
        s_mask[w][lane] = mask;
        s_any = 1;
        ndl = nd;
        if constexpr (DATA)
          if (nd > 0) node_data(O, R, (int)i, n, s_dfeat[w][lane], s_cd[w][lane]);
      }
    }
  }
  if (lane < kNodes) {
    if (state != 1) s_mask[w][lane] = 0;
    if constexpr (DATA) {
      if (state != 1 || ndl == 0) {
        // no data neighbours: zero data inputs and covariances
#pragma unroll
        for (int q = 0; q < kDFeat; ++q) s_dfeat[w][lane][q] = 0.0f;
#pragma unroll
        for (int d = 0; d < kDataSlots; ++d) s_cd[w][lane][d] = 0.0f;
      }
    }
    s_req[w][lane] = p;
  }
  __syncthreads();
  if (s_any) {
    // ---- layer 1 operands from the slot masks
    const int r = lane & 15, g = lane >> 4;
    uint32_t h0[8][CT][4], h1[8][CT][4];
    float last[kOutI / 16][CT][4];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const uint32_t m = s_mask[w][16 * ct + r];
      const float* row = A.feat + (int64_t)s_req[w][16 * ct + r] * kIn;
      const float* df = DATA ? s_dfeat[w][16 * ct + r] : nullptr;
#pragma unroll
      for (int s = 0; s < kIn / 32; ++s)
#pragma unroll
        for (int it = 0; it < 4; ++it)
          h0[s][ct][it] = pack2(feature<DATA>(row, m, frag_k(s, g, 2 * it), SLOTS, df),
                                feature<DATA>(row, m, frag_k(s, g, 2 * it + 1), SLOTS, df));
    }
    mlp_layer<kIn / 32, kHid / 16, true>(s_frag, h0, h1, last, A.wp + PW1, A.params + OB1, lane, w);
    mlp_layer<kHid / 32, kHid / 16, true>(s_frag, h1, h0, last, A.wp + PW2, A.params + OB2, lane, w);
    mlp_layer<kHid / 32, kHid / 16, true>(s_frag, h0, h1, last, A.wp + PW3, A.params + OB3, lane, w);
    mlp_layer<kHid / 32, kOutI / 16, false>(s_frag, h1, h0, last, A.wp + PW4, A.params + OB4, lane, w);
#pragma unroll
    for (int rt = 0; rt < kOutI / 16; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int o = 16 * rt + 4 * g + v;
          if (o < kUsedOut) s_out[w][16 * ct + r][o] = last[rt][ct][v];
        }
    wave_sync();
    // ---- accept, or hand to the exact solver
    if (state == 1) {
      const Req& R = O.reqs[p];
      const int n = O.nfound[i];
      const uint32_t mask = s_mask[w][lane];
      const float* out = s_out[w][lane];
      const float* cs = A.cslot + (int64_t)p * kSlotsMax;
      float chk = 1.0f;
      for (int s = 0; s < SLOTS; ++s)
        if ((mask >> s) & 1u) chk -= out[s] * cs[s];
      if constexpr (DATA)
        for (int d = 0; d < kDataSlots; ++d) chk -= out[SLOTS + 1 + d] * s_cd[w][lane][d];
      const float pred = expf(out[SLOTS]);
      if (fabsf(pred - chk) <= A.tol * pred) {
        for (int j = 0; j < O.K; ++j)
          O.lam[i * O.K + j] = j < n ? (double)out[s_slot[w][lane][j]] : 0.0;
        O.sigma[i] = sqrt(R.M.sill * (double)pred);
        O.flags[i] = (uint8_t)GSP_FLAG_LEARNED;
        atomicAdd(&O.counters[CTR_LEARNED], 1);
        state = 2;
      } else {
        state = 0;
      }
    }
  }
  // ---- the rest goes to the exact solver
  const uint64_t ex = __ballot(state == 0);
  if (ex) {
    int first = 0;
    if (lane == 0) first = atomicAdd(&O.counters[CTR_EXACT], __popcll(ex));
    first = __builtin_amdgcn_readfirstlane(first);
    if (state == 0) A.exact_list[first + mbcnt(ex)] = (int32_t)i;
  }
}

int grid_for(int64_t work, int per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

}  // namespace
}  // namespace gsp

// ----------------------------------------------------------- the object --
struct gsp_surrogate {
  std::shared_ptr<gsp::Pool> pool;  // declared first: outlives the buffers below
  // the domain it serves: one family and grid dimension, the slot offsets
  // (grid cells), and the spans of nugget / sill and of spacing / range per
  // axis of the requests it was trained on
  int32_t family = 0;
  int32_t ndim = 0;
  double sill = 1.0;  // of the training variogram (gaussian models must match it)
  double nug_lo = 0.0, nug_hi = 0.0;
  double scale_lo[3] = {0, 0, 0}, scale_hi[3] = {0, 0, 0};
  int32_t n_slots = gsp::kSlotsMax;  // 24 or 32
  int32_t data_slots = 0;            // data neighbours a node may have (0: none)
  int16_t table[gsp::kSlotsMax][3] = {};
  int32_t n_train_requests = 0;
  gsp::DevBuf<float> params, adam_m, adam_v, grad;
  gsp::DevBuf<uint16_t> wbf, wbt, wperm;
  int64_t n_samples = 0;
};

namespace gsp {
namespace {

// Family, dimension, search, rotation, neighbourhood size and slot offsets
// (the part of the domain that does not depend on the trained spans).
bool surrogate_compatible(const gsp_surrogate* s, const ReqInfo& r) {
  const gsp_model& m = r.model;
  if (m.family != s->family || m.has_rotation || r.mode != GSP_KNN_THEN_MASK) return false;
  if (r.kmax > s->n_slots || r.ndim != s->ndim || !(m.sill > 0)) return false;
  // the gaussian model's stabilising nugget is absolute, so its weights
  // depend on the sill itself, not only on the nugget/sill ratio
  if (m.family == GSP_GAUSSIAN && std::fabs(m.sill - s->sill) > 1e-12 * s->sill) return false;
  for (int t = 0; t < s->n_slots && t < r.T; ++t)
    for (int d = 0; d < 3; ++d)
      if (r.table32[t][d] != s->table[t][d]) return false;
  return true;
}

double axis_scale(const ReqInfo& r, int d) { return r.spacing[d] / r.model.ranges[d]; }

bool within(double v, double lo, double hi) {
  const double eps = 1e-9 * std::max(std::fabs(lo), std::fabs(hi)) + 1e-15;
  return v >= lo - eps && v <= hi + eps;
}

}  // namespace

bool surrogate_accepts(const gsp_surrogate* s, const ReqInfo& r) {
  if (!s || !surrogate_compatible(s, r)) return false;
  if (!within(r.model.nugget / r.model.sill, s->nug_lo, s->nug_hi)) return false;
  for (int d = 0; d < r.ndim; ++d)
    if (!within(axis_scale(r, d), s->scale_lo[d], s->scale_hi[d])) return false;
  return true;
}

void surrogate_request_rows(const gsp_surrogate* s, const ReqInfo& r, float* feat, float* cslot) {
  DevModel M;
  std::memset(&M, 0, sizeof(M));
  M.family = s->family;
  M.sill = r.model.sill;
  M.nugget = r.model.nugget;
  for (int k = 0; k < kIn; ++k) feat[k] = 0.0f;
  if (cslot)
    for (int t = 0; t < kSlotsMax; ++t) cslot[t] = 0.0f;
  for (int t = 0; t < s->n_slots; ++t) {
    double u2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double c = d < r.ndim ? (double)s->table[t][d] * axis_scale(r, d) : 0.0;
      feat[3 * t + d] = (float)(c * kFeatScale);
      u2 += c * c;
    }
    if (cslot) cslot[t] = (float)(cov_u(M, std::sqrt(u2)) / r.model.sill);
  }
  feat[3 * s->n_slots] = (float)(r.model.nugget / r.model.sill * kNugScale);
}

void surrogate_infer(gsp_ctx* ctx, const gsp_surrogate* s, int32_t which, bool list_plain,
                     const NodeOut& out, double tolerance, int32_t* exact_list, const float* feat,
                     const float* cslot) {
  InferArgs A;
  A.O = out;
  A.wp = s->wperm.p;
  A.params = s->params.p;
  A.feat = feat;
  A.cslot = cslot;
  A.tol = (float)tolerance;
  A.dslots = s->data_slots;
  A.which = which;
  A.list_plain = list_plain ? 1 : 0;
  A.exact_list = exact_list;
  constexpr int kPerBlock = kInferWaves * 16 * kInferTiles;
  const int64_t blocks = (out.NT + kPerBlock - 1) / kPerBlock;
  GSP_REQUIRE(blocks < ((int64_t)1 << 31), GSP_ERR_UNSUPPORTED, "too many nodes");
  const dim3 grid((unsigned)blocks), block(64 * kInferWaves);
  const bool data = s->data_slots > 0;
  if (s->n_slots == 24 && !data)
    hipLaunchKernelGGL((k_mlp_infer<24, kInferTiles, false>), grid, block, 0, ctx->stream, A);
  else if (s->n_slots == 24)
    hipLaunchKernelGGL((k_mlp_infer<24, kInferTiles, true>), grid, block, 0, ctx->stream, A);
  else if (!data)
    hipLaunchKernelGGL((k_mlp_infer<32, kInferTiles, false>), grid, block, 0, ctx->stream, A);
  else
    hipLaunchKernelGGL((k_mlp_infer<32, kInferTiles, true>), grid, block, 0, ctx->stream, A);
  GSP_HIP(hipGetLastError());
}

// Kernel-side view of an existing plan.
NodeOut node_view(const gsp_seqsim_plan* p) {
  NodeOut O;
  O.NT = (int32_t)p->N;
  O.K = p->kmax;
  O.reqs = p->reqs.p;
  O.node_req = p->node_req.p;
  O.table = p->table.p;
  O.nfound = p->nfound.p;
  O.nbr = p->nbr.p;
  O.nbr_slot = nullptr;
  O.data = p->data_xyz.p;
  O.lam = p->lam.p;
  O.sigma = p->sigma.p;
  O.flags = p->flags.p;
  O.counters = nullptr;
  return O;
}

void surrogate_train(gsp_ctx* ctx, int32_t n_plans, const gsp_seqsim_plan* const* plans,
                     const gsp_train_options* opt, gsp_surrogate* S, gsp_train_report* rep) {
  GSP_REQUIRE(n_plans >= 1 && plans, GSP_ERR_ARGUMENT, "no training plans");
  GSP_REQUIRE(opt, GSP_ERR_ARGUMENT, "null training options");
  for (int32_t q = 0; q < n_plans; ++q)
    GSP_REQUIRE(plans[q] && !plans[q]->info.empty(), GSP_ERR_ARGUMENT, "null training plan");
  const int Bt = opt->batch;
  GSP_REQUIRE(Bt >= 64 && Bt % 64 == 0 && Bt <= (1 << 16), GSP_ERR_ARGUMENT,
              "batch must be a positive multiple of 64 (at most 65536)");
  GSP_REQUIRE(opt->epochs >= 1 && opt->lr > 0 && opt->var_weight >= 0, GSP_ERR_ARGUMENT,
              "invalid training options");
  GSP_REQUIRE(opt->n_slots == 0 || opt->n_slots == 24 || opt->n_slots == 32, GSP_ERR_ARGUMENT,
              "n_slots must be 0 (automatic), 24 or 32");
  GSP_REQUIRE(opt->data_slots == 0 || opt->data_slots == kDataSlots, GSP_ERR_ARGUMENT,
              "data_slots must be 0 or %d", kDataSlots);
  hipStream_t st = ctx->stream;
  auto usable = [](const ReqInfo& r) {
    return !r.model.has_rotation && r.mode == GSP_KNN_THEN_MASK && r.model.sill > 0;
  };
  // ---- slots: 24 when no usable request asks for more than 24 neighbours
  int32_t slots = opt->n_slots;
  if (slots == 0) {
    int32_t kmax = 0;
    for (int32_t q = 0; q < n_plans; ++q)
      for (const ReqInfo& r : plans[q]->info)
        if (usable(r) && r.kmax <= kSlotsMax) kmax = std::max(kmax, r.kmax);
    slots = kmax > 24 ? 32 : 24;
  }
  S->n_slots = slots;

  // ---- the domain the surrogate serves: the family, dimension and slot
  // offsets of the first usable request with at most `slots` neighbours and
  // `slots` grid offsets; then the spans of nugget / sill and spacing / range
  // of every compatible request
  const ReqInfo* first = nullptr;
  for (int32_t q = 0; q < n_plans && !first; ++q)
    for (const ReqInfo& r : plans[q]->info)
      if (usable(r) && r.T >= slots && r.kmax <= slots) {
        first = &r;
        break;
      }
  GSP_REQUIRE(first, GSP_ERR_UNSUPPORTED,
              "the surrogate needs a request with a k-nearest-then-mask search over at least %d "
              "grid offsets, maxneighbors <= %d and an unrotated variogram", slots, slots);
  S->family = first->model.family;
  S->sill = first->model.sill;
  S->ndim = first->ndim;
  std::memcpy(S->table, first->table32, sizeof(S->table));
  std::vector<float> rows;
  std::vector<std::vector<int32_t>> req_rows((size_t)n_plans);
  bool any = false;
  for (int32_t q = 0; q < n_plans; ++q) {
    const gsp_seqsim_plan* p = plans[q];
    std::vector<int32_t>& rr = req_rows[(size_t)q];
    rr.assign(p->info.size(), -1);
    for (size_t r = 0; r < p->info.size(); ++r) {
      const ReqInfo& I = p->info[r];
      if (!surrogate_compatible(S, I)) continue;
      const double nr = I.model.nugget / I.model.sill;
      S->nug_lo = any ? std::min(S->nug_lo, nr) : nr;
      S->nug_hi = any ? std::max(S->nug_hi, nr) : nr;
      for (int d = 0; d < I.ndim; ++d) {
        const double a = axis_scale(I, d);
        S->scale_lo[d] = any ? std::min(S->scale_lo[d], a) : a;
        S->scale_hi[d] = any ? std::max(S->scale_hi[d], a) : a;
      }
      any = true;
      rr[r] = (int32_t)(rows.size() / kIn);
      rows.resize(rows.size() + kIn);
    }
  }
  // input rows (after the spans are final: the rows do not depend on them)
  for (int32_t q = 0; q < n_plans; ++q)
    for (size_t r = 0; r < plans[q]->info.size(); ++r)
      if (req_rows[(size_t)q][r] >= 0)
        surrogate_request_rows(S, plans[q]->info[r], rows.data() + (size_t)req_rows[(size_t)q][r] * kIn,
                               nullptr);
  S->n_train_requests = (int32_t)(rows.size() / kIn);
  DevBuf<float> drows;
  drows.alloc(rows.size());
  GSP_HIP(hipMemcpyAsync(drows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, st));

  // ---- training systems
  ctx->counters.reserve(16);
  std::vector<DevBuf<int32_t>> ok((size_t)n_plans);
  int64_t candidates = 0, data_candidates = 0;
  for (int32_t q = 0; q < n_plans; ++q) {
    const gsp_seqsim_plan* p = plans[q];
    const std::vector<int32_t>& rr = req_rows[(size_t)q];
    ok[(size_t)q].alloc(rr.size());
    GSP_HIP(hipMemcpyAsync(ok[(size_t)q].p, rr.data(), rr.size() * 4, hipMemcpyHostToDevice, st));
    GSP_HIP(hipMemsetAsync(ctx->counters.p, 0, 16 * sizeof(int32_t), st));
    ExtractArgs E;
    std::memset(&E, 0, sizeof(E));
    E.O = node_view(p);
    E.slots = slots;
    E.req_row = ok[(size_t)q].p;
    E.max_data = opt->data_slots;
    E.count = ctx->counters.p;
    hipLaunchKernelGGL(k_extract, dim3(grid_for(p->N, 256, 4096)), dim3(256), 0, st, E);
    int32_t c[2] = {0, 0};
    GSP_HIP(hipMemcpyAsync(c, ctx->counters.p, 8, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
    candidates += c[0];
    data_candidates += c[1];
  }
  // data slots when asked for and the plans hold systems with data
  // neighbours: otherwise their weights stay zero and such nodes are left to
  // the exact solver
  const int dslots = opt->data_slots > 0 && data_candidates > 0 ? kDataSlots : 0;
  S->data_slots = dslots;
  GSP_REQUIRE(candidates >= Bt, GSP_ERR_ARGUMENT,
              "the plans hold %lld usable systems, fewer than one batch (%d)",
              (long long)candidates, Bt);
  const int64_t want = opt->max_samples > 0 ? std::min<int64_t>(opt->max_samples, candidates)
                                            : candidates;
  const double keep = (double)want / (double)candidates;
  const uint32_t keep24 = keep >= 1.0 ? (1u << 24) : (uint32_t)(keep * (double)(1u << 24));
  const int64_t capacity = std::min<int64_t>(candidates, want + want / 16 + 4096);
  DevBuf<uint32_t> masks;
  DevBuf<int32_t> srow;
  DevBuf<uint8_t> ndata;
  DevBuf<float> Y, dfeat;
  masks.alloc((size_t)capacity);
  srow.alloc((size_t)capacity);
  ndata.alloc((size_t)capacity);
  dfeat.alloc((size_t)(capacity * kDFeat));
  Y.alloc((size_t)(capacity * kOutT));
  GSP_HIP(hipMemsetAsync(ctx->counters.p, 0, 16 * sizeof(int32_t), st));
  for (int32_t q = 0; q < n_plans; ++q) {
    ExtractArgs E;
    std::memset(&E, 0, sizeof(E));
    E.O = node_view(plans[q]);
    E.slots = slots;
    E.req_row = ok[(size_t)q].p;
    E.keep24 = keep24;
    E.seed = opt->seed * 0x100000001B3ULL + (uint64_t)q;
    E.max_data = dslots;
    E.cursor = ctx->counters.p + 1;
    E.capacity = (int32_t)capacity;
    E.masks = masks.p;
    E.srow = srow.p;
    E.ndata = ndata.p;
    E.dfeat = dfeat.p;
    E.Y = Y.p;
    hipLaunchKernelGGL(k_extract, dim3(grid_for(plans[q]->N, 256, 4096)), dim3(256), 0, st, E);
  }
  int32_t got = 0;
  GSP_HIP(hipMemcpyAsync(&got, ctx->counters.p + 1, 4, hipMemcpyDeviceToHost, st));
  GSP_HIP(hipStreamSynchronize(st));
  const int64_t S_n = std::min<int64_t>(got, capacity);
  GSP_REQUIRE(S_n >= Bt, GSP_ERR_ARGUMENT, "too few training systems drawn (%lld)",
              (long long)S_n);
  S->n_samples = S_n;

  // ---- parameters: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)), padding zero
  std::vector<float> P0((size_t)kParams, 0.0f);
  std::mt19937_64 rng(opt->seed);
  auto fill = [&](int64_t off, int rows, int cols, int used_rows, int used_cols, int fan_in) {
    std::uniform_real_distribution<float> U(-1.0f / std::sqrt((float)fan_in),
                                            1.0f / std::sqrt((float)fan_in));
    for (int rr = 0; rr < rows; ++rr)
      for (int cc = 0; cc < cols; ++cc)
        P0[(size_t)(off + (int64_t)rr * cols + cc)] = (rr < used_rows && cc < used_cols) ? U(rng) : 0.0f;
  };
  const int used_in = 3 * slots + 1 + (dslots ? kDFeat : 0), used_out = slots + 1 + dslots;
  fill(OW1, kHid, kIn, kHid, used_in, used_in);
  fill(OB1, 1, kHid, 1, kHid, used_in);
  fill(OW2, kHid, kHid, kHid, kHid, kHid);
  fill(OB2, 1, kHid, 1, kHid, kHid);
  fill(OW3, kHid, kHid, kHid, kHid, kHid);
  fill(OB3, 1, kHid, 1, kHid, kHid);
  fill(OW4, kOutT, kHid, used_out, kHid, kHid);
  fill(OB4, 1, kOutT, 1, used_out, kHid);
  S->params.alloc((size_t)kParams);
  S->adam_m.alloc((size_t)kParams);
  S->adam_v.alloc((size_t)kParams);
  S->grad.alloc((size_t)kParams);
  S->wbf.alloc((size_t)kBf);
  S->wbt.alloc((size_t)kBf);
  S->wperm.alloc((size_t)kPermMax);
  GSP_HIP(hipMemcpyAsync(S->params.p, P0.data(), P0.size() * 4, hipMemcpyHostToDevice, st));
  GSP_HIP(hipMemsetAsync(S->adam_m.p, 0, (size_t)kParams * 4, st));
  GSP_HIP(hipMemsetAsync(S->adam_v.p, 0, (size_t)kParams * 4, st));
  GSP_HIP(hipMemsetAsync(S->grad.p, 0, (size_t)kParams * 4, st));
  // bf16 copies of the initial weights (an Adam step with zero gradient and lr 0)
  hipLaunchKernelGGL(k_adam, dim3(grid_for(kParams, 256, 1024)), dim3(256), 0, st, S->params.p,
                     S->grad.p, S->adam_m.p, S->adam_v.p, 0.0f, 0.9f, 0.999f, 1.0f, 1.0f, 1.0f,
                     S->wbf.p, S->wbt.p);

  // ---- workspace of one step
  DevBuf<uint16_t> X, Xt, H1, H1t, H2, H2t, H3, H3t, dO, dOt, dZ1, dZ1t, dZ2, dZ2t, dZ3, dZ3t;
  DevBuf<float> G1, G2, G3, Out, Yb, lossacc;
  DevBuf<uint32_t> Mb;
  DevBuf<uint8_t> Nb;
  DevBuf<int32_t> perm;
  const size_t B = (size_t)Bt;
  X.alloc(B * kIn);
  Xt.alloc(B * kIn);
  for (DevBuf<uint16_t>* b : {&H1, &H1t, &H2, &H2t, &H3, &H3t, &dZ1, &dZ1t, &dZ2, &dZ2t, &dZ3, &dZ3t})
    b->alloc(B * kHid);
  dO.alloc(B * kOutT);
  dOt.alloc(B * kOutT);
  G1.alloc(B * kHid);
  G2.alloc(B * kHid);
  G3.alloc(B * kHid);
  Out.alloc(B * kOutT);
  Yb.alloc(B * kOutT);
  Mb.alloc(B);
  Nb.alloc(B);
  lossacc.alloc(4);
  perm.alloc((size_t)S_n);

  auto gemm = [&](int epi, const uint16_t* A_, int64_t lda, const uint16_t* B_, int64_t ldb, int M_,
                  int N_, int K_, float* C, int64_t ldc, const float* bias, uint16_t* H, uint16_t* Ht,
                  float* G, const float* Gin) {
    GemmArgs P;
    P.A = A_;
    P.B = B_;
    P.lda = lda;
    P.ldb = ldb;
    P.M = M_;
    P.N = N_;
    P.K = K_;
    P.C = C;
    P.ldc = ldc;
    P.bias = bias;
    P.H = H;
    P.Ht = Ht;
    P.G = G;
    P.Gin = Gin;
    const dim3 grid((unsigned)(N_ / 64), (unsigned)(M_ / 64));
    if (epi == EPI_F32) hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, grid, dim3(256), 0, st, P);
    else if (epi == EPI_GELU) hipLaunchKernelGGL(k_gemm_nt<EPI_GELU>, grid, dim3(256), 0, st, P);
    else hipLaunchKernelGGL(k_gemm_nt<EPI_DGELU>, grid, dim3(256), 0, st, P);
  };

  const int64_t steps_per_epoch = S_n / Bt;
  const int64_t total_steps = steps_per_epoch * opt->epochs;
  const int64_t warm = std::max<int64_t>(1, total_steps / 20);
  std::vector<int32_t> order((size_t)S_n);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 shuf(opt->seed ^ 0xA5A5A5A5ULL);
  const auto t0 = std::chrono::steady_clock::now();
  float* grad = S->grad.p;
  int64_t step = 0;
  double first_w = -1, last_w = 0, last_v = 0;
  for (int ep = 0; ep < opt->epochs; ++ep) {
    std::shuffle(order.begin(), order.end(), shuf);
    GSP_HIP(hipMemcpyAsync(perm.p, order.data(), (size_t)S_n * 4, hipMemcpyHostToDevice, st));
    GSP_HIP(hipMemsetAsync(lossacc.p, 0, 16, st));
    for (int64_t b = 0; b < steps_per_epoch; ++b, ++step) {
      // one-cycle schedule: linear warm-up, cosine decay to lr / 100
      double lr;
      if (step < warm) {
        lr = opt->lr * (double)(step + 1) / (double)warm;
      } else {
        const double f = (double)(step - warm) / (double)std::max<int64_t>(1, total_steps - warm);
        lr = opt->lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * f)));
      }
      hipLaunchKernelGGL(k_train_batch, dim3(grid_for(B * kIn, 256, 1024)), dim3(256), 0, st,
                         perm.p, b * Bt, Bt, masks.p, srow.p, ndata.p, dfeat.p, Y.p, drows.p,
                         slots, X.p, Xt.p, Yb.p, Mb.p, Nb.p);
      // forward
      gemm(EPI_GELU, X.p, kIn, S->wbf.p + BW1, kIn, Bt, kHid, kIn, nullptr, 0, S->params.p + OB1,
           H1.p, H1t.p, G1.p, nullptr);
      gemm(EPI_GELU, H1.p, kHid, S->wbf.p + BW2, kHid, Bt, kHid, kHid, nullptr, 0,
           S->params.p + OB2, H2.p, H2t.p, G2.p, nullptr);
      gemm(EPI_GELU, H2.p, kHid, S->wbf.p + BW3, kHid, Bt, kHid, kHid, nullptr, 0,
           S->params.p + OB3, H3.p, H3t.p, G3.p, nullptr);
      gemm(EPI_F32, H3.p, kHid, S->wbf.p + BW4, kHid, Bt, kOutT, kHid, Out.p, kOutT,
           S->params.p + OB4, nullptr, nullptr, nullptr, nullptr);
      hipLaunchKernelGGL(k_train_loss, dim3(grid_for(B * kOutT, 256, 256)), dim3(256), 0, st, Out.p,
                         Yb.p, Mb.p, Nb.p, Bt, slots, (float)opt->var_weight, dO.p, dOt.p,
                         lossacc.p);
      // backward
      gemm(EPI_F32, dOt.p, Bt, H3t.p, Bt, kOutT, kHid, Bt, grad + OW4, kHid, nullptr, nullptr,
           nullptr, nullptr, nullptr);
      gemm(EPI_DGELU, dO.p, kOutT, S->wbt.p + BW4, kOutT, Bt, kHid, kOutT, nullptr, 0, nullptr,
           dZ3.p, dZ3t.p, nullptr, G3.p);
      gemm(EPI_F32, dZ3t.p, Bt, H2t.p, Bt, kHid, kHid, Bt, grad + OW3, kHid, nullptr, nullptr,
           nullptr, nullptr, nullptr);
      gemm(EPI_DGELU, dZ3.p, kHid, S->wbt.p + BW3, kHid, Bt, kHid, kHid, nullptr, 0, nullptr,
           dZ2.p, dZ2t.p, nullptr, G2.p);
      gemm(EPI_F32, dZ2t.p, Bt, H1t.p, Bt, kHid, kHid, Bt, grad + OW2, kHid, nullptr, nullptr,
           nullptr, nullptr, nullptr);
      gemm(EPI_DGELU, dZ2.p, kHid, S->wbt.p + BW2, kHid, Bt, kHid, kHid, nullptr, 0, nullptr,
           dZ1.p, dZ1t.p, nullptr, G1.p);
      gemm(EPI_F32, dZ1t.p, Bt, Xt.p, Bt, kHid, kIn, Bt, grad + OW1, kIn, nullptr, nullptr,
           nullptr, nullptr, nullptr);
      hipLaunchKernelGGL(k_bias_grad, dim3(3 * kHid + kOutT), dim3(256), 0, st, dZ1t.p, dZ2t.p,
                         dZ3t.p, dOt.p, Bt, grad);
      const double t = (double)(step + 1);
      hipLaunchKernelGGL(k_adam, dim3(grid_for(kParams, 256, 1024)), dim3(256), 0, st,
                         S->params.p, grad, S->adam_m.p, S->adam_v.p, (float)lr, 0.9f, 0.999f,
                         (float)(1.0 - std::pow(0.9, t)), (float)(1.0 - std::pow(0.999, t)), 1e-8f,
                         S->wbf.p, S->wbt.p);
    }
    GSP_HIP(hipGetLastError());
    float acc[4];
    GSP_HIP(hipMemcpyAsync(acc, lossacc.p, 16, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
    last_w = std::sqrt(acc[0] / std::max(acc[2], 1.0f));
    last_v = std::sqrt(acc[1] / (double)(steps_per_epoch * Bt));
    if (first_w < 0) first_w = last_w;
  }
  hipLaunchKernelGGL(k_pack_inference, dim3(grid_for(kPermMax, 256, 1024)), dim3(256), 0, st,
                     S->params.p, S->wperm.p, out_rows(slots));
  GSP_HIP(hipStreamSynchronize(st));
  if (rep) {
    rep->n_samples = S_n;
    rep->n_candidates = candidates;
    rep->steps = total_steps;
    rep->rms_weight_first_epoch = first_w;
    rep->rms_weight_last_epoch = last_w;
    rep->rms_logvar_last_epoch = last_v;
    rep->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
}

int64_t surrogate_param_count() { return kParams; }

void surrogate_export(gsp_ctx* ctx, const gsp_surrogate* s, float* params, int64_t* count,
                      double* slots) {
  if (count) *count = kParams;
  if (slots)
    for (int t = 0; t < kSlotsMax; ++t)
      for (int d = 0; d < 3; ++d) slots[3 * t + d] = t < s->n_slots ? (double)s->table[t][d] : 0.0;
  if (params) {
    GSP_HIP(hipMemcpyAsync(params, s->params.p, (size_t)kParams * 4, hipMemcpyDeviceToHost,
                           ctx->stream));
    GSP_HIP(hipStreamSynchronize(ctx->stream));
  }
}

void mfma_probe_bf16(gsp_ctx* ctx, double* host_out) {
  DevBuf<double> d;
  d.alloc(3 * 64 * 4);
  hipLaunchKernelGGL(k_mfma_probe_bf16, dim3(1), dim3(64), 0, ctx->stream, d.p);
  GSP_HIP(hipGetLastError());
  GSP_HIP(hipMemcpyAsync(host_out, d.p, 3 * 64 * 4 * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  GSP_HIP(hipStreamSynchronize(ctx->stream));
}

void surrogate_info(const gsp_surrogate* s, gsp_surrogate_info* out) {
  std::memset(out, 0, sizeof(*out));
  out->family = s->family;
  out->ndim = s->ndim;
  out->n_slots = s->n_slots;
  out->n_train_requests = s->n_train_requests;
  out->sill = s->sill;
  out->nugget_ratio[0] = s->nug_lo;
  out->nugget_ratio[1] = s->nug_hi;
  for (int d = 0; d < 3; ++d) {
    out->scale[d][0] = s->scale_lo[d];
    out->scale[d][1] = s->scale_hi[d];
  }
  out->n_samples = s->n_samples;
  out->n_data_slots = s->data_slots;
}

gsp_surrogate* surrogate_new(gsp_ctx* ctx) {
  gsp_surrogate* s = new gsp_surrogate();
  s->pool = ctx->pool;
  return s;
}

void surrogate_delete(gsp_surrogate* s) { delete s; }

}  // namespace gsp
