// gsp_surrogate.hip — learned kriging-weight surrogate for SEQSIM on gfx950.
//
// A multilayer perceptron 73 -> 256 -> 256 -> 256 -> 25 (GELU) maps a node's
// neighbourhood to its 24 simple-kriging weights and log kriging variance,
// with BF16 weights and activations and FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x32_bf16). Inputs: 24 slots of neighbour offsets (dx, dy,
// dz divided by the ranges), zero for an absent neighbour, and the distance to
// the nearest datum. A neighbour's slot is the rank of its offset in the
// request's sorted offset table, so an interior node of a k-nearest-then-mask
// search always fills slots 0..23 of the same 24 offsets and the input is the
// pattern of already simulated neighbours. Nodes whose neighbours do not fit
// (grid boundary, conditioning data among them) and nodes whose predicted
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

constexpr int kIn = 128;     // input features, 73 used
constexpr int kHid = 256;
constexpr int kOutT = 64;    // output rows in training GEMMs, 25 used
constexpr int kOutI = 32;    // output rows at inference, 25 used
constexpr int kSlots = 24;
constexpr int kUsedIn = 73;
constexpr int kUsedOut = 25;
constexpr float kFeatScale = 8.0f;   // offsets in ranges x 8: inputs of order one
constexpr float kDataFeature = 4.0f; // nearest-datum distance (ranges), capped

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
                  PW4 = PW3 + (int64_t)kHid * kHid, kPerm = PW4 + (int64_t)kOutI * kHid;

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

// Input feature k of a node whose occupied slots are `mask`.
__device__ inline float feature(const float* canonf, uint32_t mask, int k) {
  if (k < 3 * kSlots) return ((mask >> (k / 3)) & 1u) ? canonf[k] : 0.0f;
  return k == 3 * kSlots ? kDataFeature : 0.0f;
}

// Slot of grid offset (dx, dy, dz): its index among the first 24 entries of
// the request's offset table, or -1.
__device__ inline int slot_of(const Offset* tab, int T, int dx, int dy, int dz) {
  const int n = T < kSlots ? T : kSlots;
  for (int s = 0; s < n; ++s)
    if (tab[s].dx == dx && tab[s].dy == dy && tab[s].dz == dz) return s;
  return -1;
}

// The surrogate's view of node i: its slot mask (and slot of each neighbour)
// when every neighbour is a grid node of the 24 table offsets, else false.
__device__ inline bool node_slots(const NodeOut& O, const Req& R, int i, int n, uint32_t& mask,
                                  int8_t* slots) {
  const int nx = R.dims[0], ny = R.dims[1];
  const int li = i - R.node_off;
  const int ix = li % nx, iy = (li / nx) % ny, iz = li / (nx * ny);
  const Offset* tab = O.table + R.table_off;
  mask = 0;
  for (int j = 0; j < n; ++j) {
    const int nb = O.nbr[(int64_t)i * O.K + j];
    if (nb >= O.NT) return false;
    const int l = nb - R.node_off;
    const int s = slot_of(tab, R.T, l % nx - ix, (l / nx) % ny - iy, l / (nx * ny) - iz);
    if (s < 0) return false;
    mask |= 1u << s;
    if (slots) slots[j] = (int8_t)s;
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

// One minibatch: inputs (both layouts), targets and slot masks of the
// samples perm[off .. off + Bt).
__global__ void k_train_batch(const int32_t* perm, int64_t off, int32_t Bt, const uint32_t* masks,
                              const float* Y, const float* canonf, uint16_t* X, uint16_t* Xt,
                              float* Yb, uint32_t* Mb) {
  const int64_t total = (int64_t)Bt * kIn;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / kIn), k = (int)(t % kIn);
    const int smp = perm[off + b];
    const uint32_t m = masks[smp];
    const uint16_t h = f2bf(feature(canonf, m, k));
    X[(int64_t)b * kIn + k] = h;
    Xt[(int64_t)k * Bt + b] = h;
    if (k < 32) Yb[(int64_t)b * 32 + k] = Y[(int64_t)smp * 32 + k];
    if (k == 0) Mb[b] = m;
  }
}

// Loss = mean over the batch of sum_s occupied (o_s - y_s)^2 + wv (o_24 - y_24)^2;
// writes dLoss/dO in bf16 (both layouts) and accumulates the loss terms.
__global__ void k_train_loss(const float* O, const float* Yb, const uint32_t* Mb, int32_t Bt,
                             float wv, uint16_t* dO, uint16_t* dOt, float* acc) {
  __shared__ float sw[256], sv[256], sc[256];
  float lw = 0.0f, lv = 0.0f, cnt = 0.0f;
  const int64_t total = (int64_t)Bt * kOutT;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(t / kOutT), o = (int)(t % kOutT);
    float d = 0.0f;
    if (o < kSlots) {
      if ((Mb[b] >> o) & 1u) {
        d = O[t] - Yb[(int64_t)b * 32 + o];
        lw += d * d;
        cnt += 1.0f;
      }
    } else if (o == kSlots) {
      const float e = O[t] - Yb[(int64_t)b * 32 + o];
      lv += e * e;
      d = wv * e;
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
// Padding (inputs 73..127, outputs 25..63) keeps zero gradients and stays zero.
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
__global__ void k_pack_inference(const float* params, uint16_t* wp) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < kPerm;
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
// counts the candidates only.
struct ExtractArgs {
  NodeOut O;
  const uint8_t* req_ok;
  uint32_t keep24;  // acceptance threshold on 24 hash bits
  uint64_t seed;
  int32_t* count;
  int32_t* cursor;
  int32_t capacity;
  uint32_t* masks;
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
    if (!E.req_ok[p]) continue;
    if (O.flags[i] & (GSP_FLAG_FEW | GSP_FLAG_FAIL | GSP_FLAG_LEARNED)) continue;
    const int n = O.nfound[i];
    if (n < 1) continue;
    const Req& R = O.reqs[p];
    uint32_t mask;
    if (!node_slots(O, R, (int)i, n, mask, nullptr)) continue;
    if (E.count) {
      atomicAdd(E.count, 1);
      continue;
    }
    if ((mix32(E.seed ^ (uint64_t)i * 0x9E3779B97F4A7C15ULL) & 0xFFFFFFu) >= E.keep24) continue;
    const int slot = atomicAdd(E.cursor, 1);
    if (slot >= E.capacity) continue;
    E.masks[slot] = mask;
    float* y = E.Y + (int64_t)slot * 32;
    for (int s = 0; s < 32; ++s) y[s] = 0.0f;
    const int nx = R.dims[0], ny = R.dims[1];
    const int li = (int)i - R.node_off;
    const int ix = li % nx, iy = (li / nx) % ny, iz = li / (nx * ny);
    for (int j = 0; j < n; ++j) {
      const int l = O.nbr[i * O.K + j] - R.node_off;
      const int s = slot_of(O.table + R.table_off, R.T, l % nx - ix, (l / nx) % ny - iy,
                            l / (nx * ny) - iz);
      y[s] = (float)O.lam[i * O.K + j];
    }
    const double s2 = O.sigma[i] * O.sigma[i] / R.M.sill;
    y[kSlots] = (float)log(s2 > 1e-12 ? s2 : 1e-12);
  }
}

// ------------------------------------------------------------ inference --
struct InferArgs {
  NodeOut O;
  const uint16_t* wp;   // fragment-ordered BF16 weights
  const float* params;  // FP32 master (biases)
  const float* canonf;  // input features of the 24 slots (72 floats)
  const float* cslot;   // C(slot offset) / sill
  float tol;
  int32_t* exact_list;
};

constexpr int kInferWaves = 8;  // 256 nodes per pass of the weights through LDS

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
// 32 nodes (2 column tiles); in/out hold the packed bf16 pairs of the
// operand fragments, word q of step s, column tile ct at [s][ct][q]. Every
// wavefront of the block runs it (block barriers inside). Step rt issues the
// bias load of tile rt + 1 and the copy of tile rt + 2 into LDS, then the
// products of tile rt and the activation of tile rt - 1 (the matrix cores
// and the vector ALU overlap). The step ends by waiting for the loads issued
// one step earlier only (loads complete in order; this wavefront's copy of
// tile rt + 2 stays in flight), so each copy has a whole step to land.
template <int STEPS, int TILES, bool ACT>
__device__ __forceinline__ void mlp_layer(const FragLds& F, const uint32_t (&in)[8][2][4],
                                          uint32_t (&out)[8][2][4], float (&last)[2][2][4],
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
  f32x4 acc[2][2];
  static_for<0, TILES + 1>([&](auto rtc) {
    constexpr int rt = decltype(rtc)::value;
    f32x4 b_cur = b_next;
    if constexpr (rt + 1 < TILES) b_next = *(const f32x4*)(bias + 16 * (rt + 1) + 4 * g);
    if constexpr (rt + 2 < TILES) copy_tile<STEPS>(F.template buf<(rt + 2) % 3>(), wp, rt + 2, lane, w);
    if constexpr (rt < TILES) {
      // the accumulators start at the bias of their output rows
      acc[rt & 1][0] = b_cur;
      acc[rt & 1][1] = b_cur;
      static_for<0, STEPS>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        const u32x4 a = F.template buf<rt % 3>()[s][lane];
        acc[rt & 1][0] = mfma(a, u32x4{in[s][0][0], in[s][0][1], in[s][0][2], in[s][0][3]},
                              acc[rt & 1][0]);
        acc[rt & 1][1] = mfma(a, u32x4{in[s][1][0], in[s][1][1], in[s][1][2], in[s][1][3]},
                              acc[rt & 1][1]);
      });
    }
    if constexpr (rt >= 1) {
      constexpr int pt = rt - 1;
      if constexpr (ACT) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
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
        for (int ct = 0; ct < 2; ++ct)
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

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const InferArgs A) {
  __shared__ FragBuf s_w0, s_w1, s_w2;
  const FragLds s_frag{s_w0, s_w1, s_w2};
  __shared__ uint32_t s_mask[kInferWaves][32];
  __shared__ int8_t s_slot[kInferWaves][32][kSlots];
  __shared__ float s_out[kInferWaves][32][kUsedOut + 1];
  __shared__ int32_t s_any;
  const NodeOut& O = A.O;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // one group of 32 nodes per wavefront, four groups per block (no grid-stride
  // loop: the weight addresses must not become loop invariants held in registers)
  const int64_t base = ((int64_t)blockIdx.x * kInferWaves + w) * 32;
  if (tid == 0) s_any = 0;
  __syncthreads();
  // ---- slots of the wave's 32 nodes
  int state = -1;  // -1: no node, 0: exact solve, 1: learned candidate, 2: learned
  const int64_t i = base + lane;
  if (lane < 32 && i < O.NT) {
    state = 0;
    const int p = O.node_req[i];
    const Req& R = O.reqs[p];
    const int n = O.nfound[i];
    if (R.learned && n >= R.kmin && n >= 1) {
      // slots from the table indices the search recorded
      uint32_t mask = 0;
      bool fits = true;
      for (int j = 0; j < n; ++j) {
        const int t = O.nbr_slot[i * O.K + j];
        fits = fits && t < kSlots;
        mask |= t < kSlots ? 1u << t : 0u;
        s_slot[w][lane][j] = (int8_t)(t < kSlots ? t : 0);
      }
      if (fits) {
        state = 1;
        s_mask[w][lane] = mask;
        s_any = 1;
      }
    }
  }
  if (lane < 32 && state != 1) s_mask[w][lane] = 0;
  __syncthreads();
  if (s_any) {
    // ---- layer 1 operands from the slot masks
    const int r = lane & 15, g = lane >> 4;
    uint32_t h0[8][2][4], h1[8][2][4];
    float last[2][2][4];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const uint32_t m = s_mask[w][16 * ct + r];
#pragma unroll
      for (int s = 0; s < kIn / 32; ++s)
#pragma unroll
        for (int it = 0; it < 4; ++it)
          h0[s][ct][it] = pack2(feature(A.canonf, m, frag_k(s, g, 2 * it)),
                                feature(A.canonf, m, frag_k(s, g, 2 * it + 1)));
    }
    mlp_layer<kIn / 32, kHid / 16, true>(s_frag, h0, h1, last, A.wp + PW1, A.params + OB1, lane, w);
    mlp_layer<kHid / 32, kHid / 16, true>(s_frag, h1, h0, last, A.wp + PW2, A.params + OB2, lane, w);
    mlp_layer<kHid / 32, kHid / 16, true>(s_frag, h0, h1, last, A.wp + PW3, A.params + OB3, lane, w);
    mlp_layer<kHid / 32, kOutI / 16, false>(s_frag, h1, h0, last, A.wp + PW4, A.params + OB4, lane, w);
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int o = 16 * rt + 4 * g + v;
          if (o < kUsedOut) s_out[w][16 * ct + r][o] = last[rt][ct][v];
        }
    wave_sync();
    // ---- accept, or hand to the exact solver
    if (state == 1) {
      const Req& R = O.reqs[O.node_req[i]];
      const int n = O.nfound[i];
      const uint32_t mask = s_mask[w][lane];
      const float* out = s_out[w][lane];
      float chk = 1.0f;
      for (int s = 0; s < kSlots; ++s)
        if ((mask >> s) & 1u) chk -= out[s] * A.cslot[s];
      const float pred = expf(out[kSlots]);
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
  // the variogram and neighbourhood it serves
  int32_t family = 0;
  double nugget_ratio = 0.0;
  double sill = 1.0;  // of the training variogram (gaussian models must match it)
  double canon[gsp::kSlots][3] = {};  // slot offsets in ranges
  int32_t ndim = 0;
  gsp::DevBuf<float> params, adam_m, adam_v, grad;
  gsp::DevBuf<uint16_t> wbf, wbt, wperm;
  gsp::DevBuf<float> canonf, cslot;
  int64_t n_samples = 0;
};

namespace gsp {

bool surrogate_accepts(const gsp_surrogate* s, const ReqInfo& r) {
  if (!s) return false;
  const gsp_model& m = r.model;
  if (m.family != s->family || m.has_rotation || r.mode != GSP_KNN_THEN_MASK) return false;
  if (r.kmax > kSlots || r.T < kSlots || r.ndim != s->ndim || !(m.sill > 0)) return false;
  if (std::fabs(m.nugget / m.sill - s->nugget_ratio) > 1e-12) return false;
  // the gaussian model's stabilising nugget is absolute, so its weights
  // depend on the sill itself, not only on the nugget/sill ratio
  if (m.family == GSP_GAUSSIAN && std::fabs(m.sill - s->sill) > 1e-12 * s->sill) return false;
  for (int t = 0; t < kSlots; ++t)
    for (int d = 0; d < r.ndim; ++d) {
      const double c = (double)r.table24[t][d] * r.spacing[d] / m.ranges[d];
      if (std::fabs(c - s->canon[t][d]) > 1e-9 * (1.0 + std::fabs(c))) return false;
    }
  return true;
}

void surrogate_infer(gsp_ctx* ctx, const gsp_surrogate* s, const NodeOut& out, double tolerance,
                     int32_t* exact_list) {
  InferArgs A;
  A.O = out;
  A.wp = s->wperm.p;
  A.params = s->params.p;
  A.canonf = s->canonf.p;
  A.cslot = s->cslot.p;
  A.tol = (float)tolerance;
  A.exact_list = exact_list;
  const int64_t blocks = (out.NT + kInferWaves * 32 - 1) / (kInferWaves * 32);
  GSP_REQUIRE(blocks < ((int64_t)1 << 31), GSP_ERR_UNSUPPORTED, "too many nodes");
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)blocks), dim3(64 * kInferWaves), 0, ctx->stream, A);
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
  hipStream_t st = ctx->stream;

  // ---- the variogram and slot offsets the surrogate serves (first request)
  const ReqInfo& r0 = plans[0]->info[0];
  const gsp_model& m0 = r0.model;
  GSP_REQUIRE(!m0.has_rotation && r0.mode == GSP_KNN_THEN_MASK && r0.T >= kSlots && m0.sill > 0,
              GSP_ERR_UNSUPPORTED,
              "the surrogate needs a k-nearest-then-mask search over at least 24 grid offsets "
              "and an unrotated variogram");
  S->family = m0.family;
  S->nugget_ratio = m0.nugget / m0.sill;
  S->sill = m0.sill;
  S->ndim = r0.ndim;
  for (int t = 0; t < kSlots; ++t)
    for (int d = 0; d < 3; ++d)
      S->canon[t][d] = d < r0.ndim ? (double)r0.table24[t][d] * r0.spacing[d] / m0.ranges[d] : 0.0;
  std::vector<float> canonf(3 * kSlots), cslot(kSlots);
  DevModel M;
  std::memset(&M, 0, sizeof(M));
  M.family = S->family;
  M.sill = m0.sill;
  M.nugget = m0.nugget;
  for (int t = 0; t < kSlots; ++t) {
    double u2 = 0.0;
    for (int d = 0; d < 3; ++d) {
      canonf[(size_t)(3 * t + d)] = (float)(S->canon[t][d] * kFeatScale);
      u2 += S->canon[t][d] * S->canon[t][d];
    }
    cslot[(size_t)t] = (float)(cov_u(M, std::sqrt(u2)) / m0.sill);
  }
  S->canonf.alloc(canonf.size());
  S->cslot.alloc(cslot.size());
  GSP_HIP(hipMemcpyAsync(S->canonf.p, canonf.data(), canonf.size() * 4, hipMemcpyHostToDevice, st));
  GSP_HIP(hipMemcpyAsync(S->cslot.p, cslot.data(), cslot.size() * 4, hipMemcpyHostToDevice, st));

  // ---- training systems
  ctx->counters.reserve(16);
  std::vector<DevBuf<uint8_t>> ok((size_t)n_plans);
  int64_t candidates = 0;
  for (int32_t q = 0; q < n_plans; ++q) {
    const gsp_seqsim_plan* p = plans[q];
    std::vector<uint8_t> acc(p->info.size());
    for (size_t r = 0; r < acc.size(); ++r) acc[r] = surrogate_accepts(S, p->info[r]) ? 1 : 0;
    ok[(size_t)q].alloc(acc.size());
    GSP_HIP(hipMemcpyAsync(ok[(size_t)q].p, acc.data(), acc.size(), hipMemcpyHostToDevice, st));
    GSP_HIP(hipMemsetAsync(ctx->counters.p, 0, 16 * sizeof(int32_t), st));
    ExtractArgs E;
    std::memset(&E, 0, sizeof(E));
    E.O = node_view(p);
    E.req_ok = ok[(size_t)q].p;
    E.count = ctx->counters.p;
    hipLaunchKernelGGL(k_extract, dim3(grid_for(p->N, 256, 4096)), dim3(256), 0, st, E);
    int32_t c = 0;
    GSP_HIP(hipMemcpyAsync(&c, ctx->counters.p, 4, hipMemcpyDeviceToHost, st));
    GSP_HIP(hipStreamSynchronize(st));
    candidates += c;
  }
  GSP_REQUIRE(candidates >= Bt, GSP_ERR_ARGUMENT,
              "the plans hold %lld usable systems, fewer than one batch (%d)",
              (long long)candidates, Bt);
  const int64_t want = opt->max_samples > 0 ? std::min<int64_t>(opt->max_samples, candidates)
                                            : candidates;
  const double keep = (double)want / (double)candidates;
  const uint32_t keep24 = keep >= 1.0 ? (1u << 24) : (uint32_t)(keep * (double)(1u << 24));
  const int64_t capacity = std::min<int64_t>(candidates, want + want / 16 + 4096);
  DevBuf<uint32_t> masks;
  DevBuf<float> Y;
  masks.alloc((size_t)capacity);
  Y.alloc((size_t)(capacity * 32));
  GSP_HIP(hipMemsetAsync(ctx->counters.p, 0, 16 * sizeof(int32_t), st));
  for (int32_t q = 0; q < n_plans; ++q) {
    ExtractArgs E;
    std::memset(&E, 0, sizeof(E));
    E.O = node_view(plans[q]);
    E.req_ok = ok[(size_t)q].p;
    E.keep24 = keep24;
    E.seed = opt->seed * 0x100000001B3ULL + (uint64_t)q;
    E.cursor = ctx->counters.p + 1;
    E.capacity = (int32_t)capacity;
    E.masks = masks.p;
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
  fill(OW1, kHid, kIn, kHid, kUsedIn, kUsedIn);
  fill(OB1, 1, kHid, 1, kHid, kUsedIn);
  fill(OW2, kHid, kHid, kHid, kHid, kHid);
  fill(OB2, 1, kHid, 1, kHid, kHid);
  fill(OW3, kHid, kHid, kHid, kHid, kHid);
  fill(OB3, 1, kHid, 1, kHid, kHid);
  fill(OW4, kOutT, kHid, kUsedOut, kHid, kHid);
  fill(OB4, 1, kOutT, 1, kUsedOut, kHid);
  S->params.alloc((size_t)kParams);
  S->adam_m.alloc((size_t)kParams);
  S->adam_v.alloc((size_t)kParams);
  S->grad.alloc((size_t)kParams);
  S->wbf.alloc((size_t)kBf);
  S->wbt.alloc((size_t)kBf);
  S->wperm.alloc((size_t)kPerm);
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
  Yb.alloc(B * 32);
  Mb.alloc(B);
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
                         perm.p, b * Bt, Bt, masks.p, Y.p, S->canonf.p, X.p, Xt.p, Yb.p, Mb.p);
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
                         Yb.p, Mb.p, Bt, (float)opt->var_weight, dO.p, dOt.p, lossacc.p);
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
  hipLaunchKernelGGL(k_pack_inference, dim3(grid_for(kPerm, 256, 1024)), dim3(256), 0, st,
                     S->params.p, S->wperm.p);
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
    for (int t = 0; t < kSlots; ++t)
      for (int d = 0; d < 3; ++d) slots[3 * t + d] = s->canon[t][d];
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

gsp_surrogate* surrogate_new(gsp_ctx* ctx) {
  gsp_surrogate* s = new gsp_surrogate();
  s->pool = ctx->pool;
  return s;
}

void surrogate_delete(gsp_surrogate* s) { delete s; }

}  // namespace gsp
