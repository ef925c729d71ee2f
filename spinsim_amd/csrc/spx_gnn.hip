// The surrogate network of the control variate (include/spx_gnn.h):
// forward, backward and Adam on the device, every matrix product on BF16
// MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulation). Predictions run layer
// by layer through one tiled GEMM kernel; a training step runs the whole
// forward and backward pass of a block of graphs in one kernel (k_train),
// then one kernel that sums the blocks' gradients and applies Adam.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "spx_common.hpp"
#include "spx_ensemble.hpp"
#include "spx_gnn.h"

namespace spx {
namespace {

constexpr int H = 128;       // hidden width
constexpr int F = 16;        // padded input features
constexpr int NLAYER = 3;
constexpr size_t kPredictChunk = 4096;  // graphs per chunk of a bulk prediction

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}

struct GemmArgs {
  const uint16_t* a;
  long sam, sak;
  const uint16_t* b;
  long sbk, sbn;
  float* c;
  long scm, scn;
  int M, N, K;
  const float* bias;   // [N] or null
  float* preact;       // store the pre-activation (before ReLU) or null
  int relu;
  int accumulate;
};

// C = A B (+ C) (+ bias), optional ReLU; A is M x K, B is K x N, bf16,
// FP32 accumulation. A workgroup of four wavefronts owns a 64 x 64 tile of
// C (each wavefront 32 x 32, four 16 x 16 MFMA accumulators) and walks K 32
// at a time through LDS: A as [m][k], B transposed as [n][k], so every MFMA
// operand fragment is four consecutive bf16 of one LDS row. An operand whose
// K (or M / N) index is contiguous in memory is loaded eight elements (16
// bytes) per thread; other strides element by element. A prediction's
// products never split K, so it does not depend on the batch it is
// computed in.
constexpr int kBM = 64, kBN = 64, kBK = 32, kLP = kBK + 8;   // LDS row pitch (bf16)

__device__ __forceinline__ void gemm_store(const GemmArgs& g, int row, int col, float v) {
  if (g.bias) v += g.bias[col];
  float* dst = g.c + static_cast<long>(row) * g.scm + static_cast<long>(col) * g.scn;
  if (g.accumulate) v += *dst;
  if (g.preact) g.preact[static_cast<long>(row) * g.scm + static_cast<long>(col) * g.scn] = v;
  if (g.relu) v = v > 0.f ? v : 0.f;
  *dst = v;
}

// One 64 x 32 operand tile into LDS as t[r][k] (r: row of A or column of B),
// element (r, k) at src[(r0 + r) * sr + (k0 + k) * sk], zero outside
// [0, R) x [k0, k1).
__device__ __forceinline__ void load_tile(uint16_t (*t)[kLP], const uint16_t* __restrict__ src, long sr, long sk,
                                          int r0, int R, int k0, int k1) {
  const int tid = threadIdx.x;
  if (sk == 1) {
    // K contiguous: eight consecutive k of one row per thread
    const int r = tid >> 2, kc = (tid & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    const uint16_t* p = src + static_cast<long>(gr) * sr + gk;
    if (gr < R && gk + 8 <= k1 && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
      const uint4 v = *reinterpret_cast<const uint4*>(p);
      *reinterpret_cast<uint4*>(&t[r][kc]) = v;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) t[r][kc + e] = (gr < R && gk + e < k1) ? p[e] : uint16_t(0);
    }
  } else if (sr == 1) {
    // rows contiguous: eight consecutive rows of one k per thread
    const int k = tid >> 3, rc = (tid & 7) * 8;
    const int gk = k0 + k, gr = r0 + rc;
    const uint16_t* p = src + static_cast<long>(gk) * sk + gr;
    if (gk < k1 && gr + 8 <= R && (reinterpret_cast<uintptr_t>(p) & 15u) == 0) {
      const uint4 v = *reinterpret_cast<const uint4*>(p);
      const uint16_t* e8 = reinterpret_cast<const uint16_t*>(&v);
#pragma unroll
      for (int e = 0; e < 8; ++e) t[rc + e][k] = e8[e];
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) t[rc + e][k] = (gk < k1 && gr + e < R) ? p[e] : uint16_t(0);
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int i = tid * 8 + e, r = i / kBK, k = i % kBK;
      const int gr = r0 + r, gk = k0 + k;
      t[r][k] = (gr < R && gk < k1) ? src[static_cast<long>(gr) * sr + static_cast<long>(gk) * sk] : uint16_t(0);
    }
  }
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[kBM][kLP];
  __shared__ __align__(16) uint16_t Bs[kBN][kLP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bm = blockIdx.y * kBM, bn = blockIdx.x * kBN;
  const int kbeg = 0, kend = g.K;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += kBK) {
    load_tile(As, g.a, g.sam, g.sak, bm, g.M, k0, kend);
    load_tile(Bs, g.b, g.sbn, g.sbk, bn, g.N, k0, kend);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      bf16x4 af[2], bf[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x4*>(&As[wm + i * 16 + fr][kk + fk]);
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[j] = *reinterpret_cast<const bf16x4*>(&Bs[wn + j * 16 + fr][kk + fk]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = bn + wn + j * 16 + fr;
      if (col >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = bm + wm + i * 16 + fk + r;
        if (row >= g.M) continue;
        gemm_store(g, row, col, acc[i][j][r]);
      }
    }
}

__global__ void k_cast(const float* __restrict__ x, uint16_t* __restrict__ y, long n) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

// Edge weights and node features of one graph per block, one thread per
// spin (L <= 30): x = J / s_J, d_i = sum_j |x_ij|,
// A_ij = x_ij / sqrt((1 + d_i)(1 + d_j)),
// f_i = (h_i / s_h, d_i, log(1 + sum_j x_ij^2), 1, 0...) in bf16.
__global__ __launch_bounds__(32) void k_graph_inputs(const float* __restrict__ J, const float* __restrict__ h,
                                                     float* __restrict__ A, uint16_t* __restrict__ feat, int L,
                                                     float inv_sj, float inv_sh) {
  __shared__ float deg[32];
  const int g = blockIdx.x, i = threadIdx.x;
  const float* Jr = J + (static_cast<long>(g) * L + i) * L;
  float d = 0.f, s2 = 0.f;
  if (i < L) {
    for (int j = 0; j < L; ++j) {
      const float x = Jr[j] * inv_sj;
      d += fabsf(x);
      s2 += x * x;
    }
    uint16_t* f = feat + (static_cast<long>(g) * L + i) * F;
    f[0] = to_bf16(h[static_cast<long>(g) * L + i] * inv_sh);
    f[1] = to_bf16(d);
    f[2] = to_bf16(log1pf(s2));
    f[3] = to_bf16(1.f);
    for (int k = 4; k < F; ++k) f[k] = 0;
  }
  deg[i] = d;
  __syncthreads();
  if (i >= L) return;
  float* Ar = A + (static_cast<long>(g) * L + i) * L;
  for (int j = 0; j < L; ++j) Ar[j] = Jr[j] * inv_sj / sqrtf((1.f + d) * (1.f + deg[j]));
}

// [x | sum_j A_ij x_j] in bf16, rows of 2H: the message sum in FP32, rounded
// once with the node's own features.
__global__ void k_mix_cat(const float* __restrict__ J, const float* __restrict__ x, uint16_t* __restrict__ cat,
                          int L) {
  const int g = blockIdx.x, i = blockIdx.y, c = threadIdx.x;
  const float* Jr = J + (static_cast<long>(g) * L + i) * L;
  const float* xg = x + static_cast<long>(g) * L * H;
  float s = 0.f;
  for (int j = 0; j < L; ++j) s += Jr[j] * xg[j * H + c];
  uint16_t* row = cat + (static_cast<long>(g) * L + i) * 2 * H;
  row[c] = to_bf16(xg[i * H + c]);
  row[H + c] = to_bf16(s);
}

__global__ void k_pool(const float* __restrict__ x, float* __restrict__ p, int L) {
  const int g = blockIdx.x, c = threadIdx.x;
  float s = 0.f;
  for (int i = 0; i < L; ++i) s += x[(static_cast<long>(g) * L + i) * H + c];
  p[g * H + c] = s / static_cast<float>(L);
}

// ---------------------------------------------------------------------------
// One training step in two launches. k_train runs the forward pass, the loss
// and the backward pass of whole graphs: a workgroup takes a block of at most
// kTR rows (gb graphs of L spins, zero rows after them; the products cover
// the M tiles that hold the block's rows), keeps every
// activation of the block in LDS, and runs each product on MFMA from LDS
// operands, so the ~40 dependent launches of the layer-by-layer path become
// one. Its weight and bias gradients are partial sums over the block's rows,
// written to row blockIdx.x of `part` by a launch of the first kTrainWG
// blocks and added to it by the launches of the blocks after them;
// k_train_reduce sums the rows in order, so every result is independent of
// scheduling, and applies Adam. The rounding points are those of the layer-by-layer path: bf16
// operands of every product, FP32 accumulation, FP32 pre-activations,
// messages, pooling and gradients of the activations.

constexpr int kTR = 64;             // rows of a graph block
constexpr int kTW = 8;              // wavefronts of k_train (two per SIMD)
constexpr int kTT = kTW * 64;       // its threads
constexpr int kTrainWG = 128;       // blocks per launch (the partial rows)
// LDS row strides in bf16 elements: 4 (mod 8) dwords between rows, so the 16
// rows of an MFMA operand fragment start on distinct bank quadruples
constexpr int kSC = 2 * H + 8;      // [row][2H] layer inputs
constexpr int kSH = H + 8;          // [row][H]
constexpr int kSR = kTR + 8;        // [*][row]: K = rows (weight gradients)
constexpr int kSG = 16 + 8;         // [*][16]: K = features or graphs
constexpr int kSX = H + 4;          // FP32 rows [row][H]
// LDS regions (bytes)
constexpr int kR1 = kTR * kSH * 2 + H * kSR * 2;           // layer input; backward dz and dz^T
constexpr int kR2 = 2 * H * kSR * 2;                       // staged weights or the layer input transposed
constexpr int kR3 = kTR * kSX * 4;                         // activations / their gradients, FP32
constexpr int kR4 = kTR * kSX * 4;                         // features, readout, message gradients
constexpr int kRA = kTR * 30 * 4;                          // edge weights [graph][i][j], gb L L <= 64 L
constexpr int kRM = (NLAYER + 1) * 32 * 4 * 8;             // ReLU masks: a ballot per 16x16 tile and row group
constexpr int kRF = 16 * kSR * 2;                          // features transposed
constexpr int kCG = kTT / H;                               // row groups of the column passes
constexpr int kRS = kCG * H * 4;                           // their partial column sums
constexpr int kRB = (NLAYER + 2) * H * 4;                  // biases b0..b_NLAYER, b_out
constexpr int kTrainLds = kR1 + kR2 + kR3 + kR4 + kRA + kRM + kRF + kRS + kRB;
static_assert(kTR * kSC * 2 <= kR1 && 64 * kSC * 2 <= kR2 && H * kSH * 2 <= kR2, "LDS regions");
static_assert(16 * kSH * 2 + H * kSG * 2 + 16 * H * 4 + 16 * kSH * 2 + H * kSG * 2 <= kR4, "readout region");
static_assert(kTrainLds <= 160 * 1024, "LDS budget");

struct TrainArgs {
  const float* J;               // couplings [n][L][L], FP32
  const float* h;               // fields [n][L]
  const float* t;               // targets [n][P]
  const uint32_t* order;        // graph i of the batch is order[*step G + i]; null: graph i
  const int* step;
  int G, L, P, gb;              // batch graphs, spins, outputs, graphs per block
  int blk0;                     // block of workgroup 0: a launch covers blocks blk0 .. blk0 + gridDim.x - 1
  float inv_sj, inv_sh, scale;  // 1/s_J, 1/s_h, 1/(G P)
  const float* p;               // FP32 parameters (the biases are read here)
  const uint16_t* pbf;          // bf16 parameters
  const uint16_t* pbfT;         // bf16 W_l^T [2H][H] (l = 1..3) and W_out^T [H][16 ceil(P / 16)]
  long nparam;
  float* part;                  // [gridDim.x][nparam] gradient partials
  uint16_t* catg;               // [gridDim.x][NLAYER][kTR][2H] layer inputs for the backward
  float* sq;                    // [G][P] squared errors
};

// Parameter layout: W0 [H][F], b0, then W_l [H][2H], b_l for l = 1..NLAYER,
// then W_out [P][H], b_out; the transposed copies W_l^T [2H][H], then
// W_out^T [H][Pp] with Pp = 16 ceil(P / 16) (rows zero-padded).
__host__ __device__ constexpr long off_w(int l) {
  return l == 0 ? 0 : H * F + H + static_cast<long>(l - 1) * (2 * H * H + H);
}
__host__ __device__ constexpr long off_b(int l, int P) {
  return off_w(l) + (l == 0 ? H * F : l <= NLAYER ? 2 * H * H : static_cast<long>(P) * H);
}
__host__ __device__ constexpr long off_wt(int l) { return static_cast<long>(l - 1) * 2 * H * H; }

// C = A Bt^T: A [M][K] (row stride sa) and Bt [N][K] (stride sb), bf16 in
// LDS, mt M tiles and nt N tiles of 16. The work items (a group of up to MG
// M tiles, one N tile) go to the waves in turn, so that small products
// still occupy every wave; epi(i, r, row, col, v) (i: the M tile) is called
// for every element in lockstep across the wave (uniform control), v plus
// cbias[col] when a column bias is given. K = k <= KMAX, a multiple of 16,
// walked in unrolled runs of four k-steps (so the next fragments load
// while the MFMAs run).
template <int MG, int KMAX, class Epi>
__device__ __forceinline__ void tile_gemm(const uint16_t* A, int sa, const uint16_t* Bt, int sb, int mt, int nt,
                                          int k, Epi epi, const float* cbias = nullptr) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  const int items = (mt + MG - 1) / MG * nt;
  for (int it = wave; it < items; it += kTW) {
    const int tn = it % nt, m0 = it / nt * MG;
    // the column's bias in a register: read in the epilogue, it would be
    // re-read after every store (the compiler cannot tell the arrays apart)
    const float cb = cbias ? cbias[tn * 16 + fr] : 0.f;
    f32x4 acc[MG];
#pragma unroll
    for (int i = 0; i < MG; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int KC = KMAX < 64 ? KMAX : 64;   // k-steps unrolled four at a time
#pragma unroll 1
    for (int kc = 0; kc < KMAX; kc += KC) {
      if (kc >= k) break;
#pragma unroll
      for (int k0 = kc; k0 < kc + KC; k0 += 16) {
        if (k0 >= k) break;
        const bf16x4 b = *reinterpret_cast<const bf16x4*>(Bt + (tn * 16 + fr) * sb + k0 + fk);
#pragma unroll
        for (int i = 0; i < MG; ++i) {
          if (m0 + i >= mt) break;
          const bf16x4 x = *reinterpret_cast<const bf16x4*>(A + ((m0 + i) * 16 + fr) * sa + k0 + fk);
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, b, acc[i], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < MG; ++i) {
      if (m0 + i >= mt) break;
#pragma unroll
      for (int r = 0; r < 4; ++r)
        epi(m0 + i, r, (m0 + i) * 16 + fk + r, tn * 16 + fr, cbias ? acc[i][r] + cb : acc[i][r]);
    }
  }
}

// The same with B given as [K][N] (row stride sb), K = k <= KMAX a multiple
// of 16: each lane gathers its four consecutive k of one column with 16-bit
// reads (the 16 columns of a fragment are consecutive halfwords, the four k
// groups 16 banks apart at the strides used here).
template <int MG, int KMAX, class Epi>
__device__ __forceinline__ void tile_gemm_kn(const uint16_t* A, int sa, const uint16_t* Bkn, int sb, int mt, int nt,
                                             int k, Epi epi) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int fr = lane & 15, fk = (lane >> 4) * 4;
  const int items = (mt + MG - 1) / MG * nt;
  for (int it = wave; it < items; it += kTW) {
    const int tn = it % nt, m0 = it / nt * MG;
    f32x4 acc[MG];
#pragma unroll
    for (int i = 0; i < MG; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < KMAX; k0 += 16) {
      if (k0 >= k) break;
      const uint16_t* bp = Bkn + (k0 + fk) * sb + tn * 16 + fr;
      bf16x4 b;
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = static_cast<short>(bp[q * sb]);
#pragma unroll
      for (int i = 0; i < MG; ++i) {
        if (m0 + i >= mt) break;
        const bf16x4 x = *reinterpret_cast<const bf16x4*>(A + ((m0 + i) * 16 + fr) * sa + k0 + fk);
        acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, b, acc[i], 0, 0, 0);
      }
    }
#pragma unroll
    for (int i = 0; i < MG; ++i) {
      if (m0 + i >= mt) break;
#pragma unroll
      for (int r = 0; r < 4; ++r) epi(m0 + i, r, (m0 + i) * 16 + fk + r, tn * 16 + fr, acc[i][r]);
    }
  }
}

// A staging of bf16 rows from global memory into LDS in two halves: load()
// issues every 16-byte load of the thread into registers, store() writes
// them to LDS. Issued right after a barrier, the loads run under the product
// that precedes the staging's use. nrows rows of per 16-byte pieces
// (nrows * per <= 2048); rows at or beyond `rows` are zero.
constexpr int kStageIt = 2048 / kTT;   // 16-byte pieces per thread
struct Staging {
  uint4 v[kStageIt];
  __device__ __forceinline__ void load(const uint16_t* __restrict__ src, long ss, int rows, int nrows, int per) {
#pragma unroll
    for (int q = 0; q < kStageIt; ++q) {
      const int e = threadIdx.x + q * kTW * 64, r = e / per, c = (e % per) * 8;
      v[q] = make_uint4(0u, 0u, 0u, 0u);
      if (e < nrows * per && r < rows) v[q] = *reinterpret_cast<const uint4*>(src + r * ss + c);
    }
  }
  __device__ __forceinline__ void store(uint16_t* dst, int sd, int nrows, int per) const {
#pragma unroll
    for (int q = 0; q < kStageIt; ++q) {
      const int e = threadIdx.x + q * kTW * 64, r = e / per, c = (e % per) * 8;
      if (e < nrows * per) *reinterpret_cast<uint4*>(dst + r * sd + c) = v[q];
    }
  }
};

__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  return static_cast<uint32_t>(to_bf16(lo)) | (static_cast<uint32_t>(to_bf16(hi)) << 16);
}

// Elementwise passes over the block's 64 x 128 FP32 activations.
// Row-chunk mapping: kTPR threads per row, lane q of a row takes the kRV
// columns 4 kTPR k + 4 q + e (k < kRV / 4, e < 4) as v[4 k + e], so the
// row's lanes touch 16 kTPR contiguous bytes at each k. Column mapping:
// thread t takes column t mod H and the kCR rows from kCR (t / H).
constexpr int kTPR = kTT / kTR, kRV = H / kTPR, kCR = kTR / kCG;
static_assert(kRV % 4 == 0 && kCR % 8 == 0 && kTT % H == 0, "elementwise mappings");

__device__ __forceinline__ int row_col(int k, int q) { return 4 * kTPR * k + 4 * q; }

__device__ __forceinline__ void load_row(float* v, const float* row, int q) {
#pragma unroll
  for (int k = 0; k < kRV / 4; ++k) {
    const float4 x = *reinterpret_cast<const float4*>(row + row_col(k, q));
    v[4 * k] = x.x;
    v[4 * k + 1] = x.y;
    v[4 * k + 2] = x.z;
    v[4 * k + 3] = x.w;
  }
}

__device__ __forceinline__ void store_row(float* row, const float* v, int q) {
#pragma unroll
  for (int k = 0; k < kRV / 4; ++k)
    *reinterpret_cast<float4*>(row + row_col(k, q)) = make_float4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
}

__device__ __forceinline__ void store_row_bf16(uint16_t* row, const float* v, int q) {
#pragma unroll
  for (int k = 0; k < kRV / 4; ++k)
    *reinterpret_cast<uint2*>(row + row_col(k, q)) = make_uint2(pack2(v[4 * k], v[4 * k + 1]),
                                                                pack2(v[4 * k + 2], v[4 * k + 3]));
}

// kCR FP32 values as bf16 into 2 kCR bytes (16-byte aligned)
__device__ __forceinline__ void store_col_bf16(uint16_t* dst, const float* v) {
#pragma unroll
  for (int q = 0; q < kCR / 8; ++q)
    reinterpret_cast<uint4*>(dst)[q] = make_uint4(pack2(v[8 * q], v[8 * q + 1]), pack2(v[8 * q + 2], v[8 * q + 3]),
                                                  pack2(v[8 * q + 4], v[8 * q + 5]), pack2(v[8 * q + 6], v[8 * q + 7]));
}

// ReLU masks: one ballot per 16 x 16 tile of a 64 x 128 activation and row
// group r = row & 3; bit ((row & 15) >> 2) * 16 + (col & 15). The 16 bits of
// row `row` in the tile holding column col.
__device__ __forceinline__ uint32_t mask_bits16(const unsigned long long* mk, int row, int col) {
  const unsigned long long w = mk[((row >> 4) * 8 + (col >> 4)) * 4 + (row & 3)];
  return static_cast<uint32_t>(w >> (((row & 15) >> 2) * 16)) & 0xffffu;
}

__global__ __launch_bounds__(kTW * 64) void k_train(TrainArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t* R1 = reinterpret_cast<uint16_t*>(smem);
  uint16_t* R2 = reinterpret_cast<uint16_t*>(smem + kR1);
  float* R3 = reinterpret_cast<float*>(smem + kR1 + kR2);
  float* R4 = reinterpret_cast<float*>(smem + kR1 + kR2 + kR3);
  float* sA = reinterpret_cast<float*>(smem + kR1 + kR2 + kR3 + kR4);
  unsigned long long* mk = reinterpret_cast<unsigned long long*>(smem + kR1 + kR2 + kR3 + kR4 + kRA);
  uint16_t* featT = reinterpret_cast<uint16_t*>(smem + kR1 + kR2 + kR3 + kR4 + kRA + kRM);
  float* half_sum = reinterpret_cast<float*>(smem + kR1 + kR2 + kR3 + kR4 + kRA + kRM + kRF);   // [kCG][H]
  float* bias = half_sum + kCG * H;                                                              // [NLAYER + 2][H]
  // sub-regions
  uint16_t* cat = R1;                                           // [kTR][kSC]
  uint16_t* dzr = R1;                                           // [kTR][kSH]
  uint16_t* dzT = R1 + kTR * kSH;                               // [H][kSR]
  uint16_t* feat = reinterpret_cast<uint16_t*>(R4);             // [kTR][kSG]
  uint16_t* poolr = reinterpret_cast<uint16_t*>(R4);            // [16][kSH]
  uint16_t* poolT = poolr + 16 * kSH;                           // [H][kSG]
  float* dyf = reinterpret_cast<float*>(poolT + H * kSG);       // [16][H]
  uint16_t* dyr = reinterpret_cast<uint16_t*>(dyf + 16 * H);    // [16][sp]
  uint16_t* dyT = dyr + 16 * kSH;                               // [H][kSG]
  const int tid = threadIdx.x;
  const int L = a.L, P = a.P, Pt = (P + 15) / 16, sp = Pt * 16 + 8;
  // row-chunk mapping of the elementwise passes: row tr, lane rq of the row
  const int tr = tid / kTPR, rq = tid % kTPR;
  // column mapping: column cc, rows ch .. ch + kCR - 1
  const int cc = tid % H, ch = tid / H * kCR;
  float* part = a.part + static_cast<long>(blockIdx.x) * a.nparam;
  uint16_t* catg = a.catg + static_cast<long>(blockIdx.x) * NLAYER * kTR * 2 * H;
  const long base = a.order ? static_cast<long>(*a.step) * a.G : 0;
  const bool first = a.blk0 == 0;
  auto put = [&](long i, float v) {
    if (first) part[i] = v;
    else part[i] += v;
  };
  {
    const int blk = a.blk0 + static_cast<int>(blockIdx.x);
    const int g0 = blk * a.gb, ng = min(a.gb, a.G - g0), R = ng * L;
    const int mtr = (R + 15) / 16;   // M tiles holding the block's rows
    auto src_of = [&](int gl) -> long {
      const long gi = g0 + gl;
      return a.order ? static_cast<long>(a.order[base + gi]) : gi;
    };
    // dz = dx where the layer's pre-activation was positive: bf16 [row][c]
    // (row-chunk mapping) when want_rows, bf16 [c][row] and the bias
    // gradient (column mapping); dx in R3.
    auto relu_bwd = [&](const unsigned long long* ml, bool want_rows, long bias_off) {
      if (want_rows) {
        float v[kRV];
        load_row(v, R3 + tr * kSX, rq);
#pragma unroll
        for (int k = 0; k < kRV / 4; ++k) {
          const int col = row_col(k, rq);
          const uint32_t m = tr < R ? mask_bits16(ml, tr, col) >> (col & 15) : 0u;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (!((m >> e) & 1u)) v[4 * k + e] = 0.f;
        }
        store_row_bf16(dzr + tr * kSH, v, rq);
      }
      // loads first, unconditionally: no read waits on another
      float v[kCR], s = 0.f;
      uint32_t mb[kCR];
#pragma unroll
      for (int q = 0; q < kCR; ++q) {
        v[q] = R3[(ch + q) * kSX + cc];
        mb[q] = mask_bits16(ml, ch + q, cc);
      }
#pragma unroll
      for (int q = 0; q < kCR; ++q) {
        if (!(ch + q < R && ((mb[q] >> (cc & 15)) & 1u))) v[q] = 0.f;
        s += v[q];
      }
      store_col_bf16(dzT + cc * kSR + ch, v);
      half_sum[(ch / kCR) * H + cc] = s;
      __syncthreads();
      if (tid < H) {
        float b = half_sum[tid];
#pragma unroll
        for (int g = 1; g < kCG; ++g) b += half_sum[g * H + tid];
        put(bias_off + tid, b);
      }
    };
    Staging st;
    st.load(a.pbf + off_w(0), F, H, H, F / 8);    // W0 [c][f]
    for (int e = tid; e < (NLAYER + 2) * H; e += kTW * 64) {
      const int l = e / H, c = e % H;
      bias[e] = l <= NLAYER ? a.p[off_b(l, P) + c] : (c < P ? a.p[off_b(NLAYER + 1, P) + c] : 0.f);
    }
    // 1. edge weights and features (as k_graph_inputs)
    float* deg = R3;
    if (tid < kTR) {
      const int r = tid;
      float d = 0.f, s2 = 0.f;
      uint16_t f[4] = {0, 0, 0, 0};
      if (r < R) {
        const int gl = r / L, i = r % L;
        const long src = src_of(gl);
        const float* Jr = a.J + (src * L + i) * L;
        for (int j = 0; j < L; ++j) {
          const float x = Jr[j] * a.inv_sj;
          d += fabsf(x);
          s2 += x * x;
        }
        f[0] = to_bf16(a.h[src * L + i] * a.inv_sh);
        f[1] = to_bf16(d);
        f[2] = to_bf16(log1pf(s2));
        f[3] = to_bf16(1.f);
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const uint16_t v = k < 4 ? f[k] : uint16_t(0);
        feat[r * kSG + k] = v;
        featT[k * kSR + r] = v;
      }
      deg[r] = d;
    }
    __syncthreads();
    for (int e = tid; e < R * L; e += kTW * 64) {
      const int r = e / L, j = e % L, gl = r / L, i = r % L;
      const float* Jr = a.J + (src_of(gl) * L + i) * L;
      sA[e] = Jr[j] * a.inv_sj / sqrtf((1.f + deg[r]) * (1.f + deg[gl * L + j]));
    }
    st.store(R2, kSG, H, F / 8);
    __syncthreads();
    // 2. layer 0: x = relu(feat W0^T + b0)
    st.load(a.pbf + off_w(1), 2 * H, 64, 64, 2 * H / 8);
    {
      tile_gemm<4, F>(feat, kSG, R2, kSG, mtr, H / 16, F, [&](int i, int r, int row, int col, float z) {
        const unsigned long long w = __ballot(z > 0.f);
        if ((threadIdx.x & 63) == 0) mk[(i * 8 + (col >> 4)) * 4 + r] = w;
        R3[row * kSX + col] = row < R ? fmaxf(z, 0.f) : 0.f;
      }, bias);
    }
    __syncthreads();
    // 3. layers 1..NLAYER: cat = [x | A x] in bf16, x = relu(cat W_l^T + b_l)
    for (int l = 1; l <= NLAYER; ++l) {
      {
        float x[kRV], m[kRV];
        load_row(x, R3 + tr * kSX, rq);
#pragma unroll
        for (int q = 0; q < kRV; ++q) m[q] = 0.f;
        if (tr < R) {
          const int gl = tr / L, i = tr % L;
          const float* Ar = sA + (gl * L + i) * L;
          int j = 0;
          for (; j + 1 < L; j += 2) {
            const float w0 = Ar[j], w1 = Ar[j + 1];
            float y0[kRV], y1[kRV];
            load_row(y0, R3 + (gl * L + j) * kSX, rq);
            load_row(y1, R3 + (gl * L + j + 1) * kSX, rq);
#pragma unroll
            for (int q = 0; q < kRV; ++q) m[q] = (m[q] + w0 * y0[q]) + w1 * y1[q];
          }
          if (j < L) {
            const float w = Ar[j];
            float y[kRV];
            load_row(y, R3 + (gl * L + j) * kSX, rq);
#pragma unroll
            for (int q = 0; q < kRV; ++q) m[q] += w * y[q];
          }
        } else {
#pragma unroll
          for (int q = 0; q < kRV; ++q) x[q] = 0.f;
        }
        uint16_t* cg = catg + static_cast<long>(l - 1) * kTR * 2 * H + tr * 2 * H;
        store_row_bf16(cat + tr * kSC, x, rq);
        store_row_bf16(cat + tr * kSC + H, m, rq);
        store_row_bf16(cg, x, rq);
        store_row_bf16(cg + H, m, rq);
      }
      st.store(R2, kSC, 64, 2 * H / 8);
      const float* b = bias + l * H;
      for (int half = 0; half < 2; ++half) {
        __syncthreads();
        if (half == 0)
          st.load(a.pbf + off_w(l) + 64L * 2 * H, 2 * H, 64, 64, 2 * H / 8);
        else if (l < NLAYER)
          st.load(a.pbf + off_w(l + 1), 2 * H, 64, 64, 2 * H / 8);
        else
          st.load(a.pbf + off_w(NLAYER + 1), H, P, Pt * 16, H / 8);    // W_out [q][c]
        auto epi = [&](int i, int r, int row, int c0, float z) {
          const int col = half * 64 + c0;
          const unsigned long long w = __ballot(z > 0.f);
          if ((threadIdx.x & 63) == 0) mk[((l * 32) + i * 8 + (col >> 4)) * 4 + r] = w;
          R3[row * kSX + col] = row < R ? fmaxf(z, 0.f) : 0.f;
        };
        // up to 32 rows: one M tile per work item, so that every wave has one
        if (mtr > 2) tile_gemm<2, 2 * H>(cat, kSC, R2, kSC, mtr, 4, 2 * H, epi, b + half * 64);
        else tile_gemm<1, 2 * H>(cat, kSC, R2, kSC, mtr, 4, 2 * H, epi, b + half * 64);
        __syncthreads();
        if (half == 0) st.store(R2, kSC, 64, 2 * H / 8);
      }
    }
    // 4. readout, loss and its gradient; the block's targets into R1 (free
    // until the backward pass) while the pooling runs
    float* tg = reinterpret_cast<float*>(R1);                     // [16][H]
    for (int e = tid; e < ng * P; e += kTW * 64) tg[(e / P) * H + e % P] = a.t[src_of(e / P) * P + e % P];
    {
      // 16 graphs x H columns over the threads, kPC columns each
      constexpr int kPC = 16 * H / kTT;
      static_assert(kPC == 4, "one float4 of a row per thread");
      const int gl = tid / (H / kPC), c0 = tid % (H / kPC) * kPC;
      float s[kPC];
#pragma unroll
      for (int q = 0; q < kPC; ++q) s[q] = 0.f;
      if (gl < ng) {
        for (int i = 0; i < L; ++i) {
          const float4 u = *reinterpret_cast<const float4*>(R3 + (gl * L + i) * kSX + c0);
          s[0] += u.x;
          s[1] += u.y;
          s[2] += u.z;
          s[3] += u.w;
        }
      }
#pragma unroll
      for (int q = 0; q < kPC; ++q) {
        const uint16_t v = gl < ng ? to_bf16(s[q] / static_cast<float>(L)) : uint16_t(0);
        poolr[gl * kSH + c0 + q] = v;
        poolT[(c0 + q) * kSG + gl] = v;
      }
    }
    st.store(R2, kSH, Pt * 16, H / 8);
    __syncthreads();
    st.load(a.pbfT + off_wt(NLAYER + 1), Pt * 16, H, H, Pt * 2);   // W_out^T [c][q], rows padded
    {
      tile_gemm<1, H>(poolr, kSH, R2, kSH, 1, Pt, H, [&](int, int, int gl, int q, float y) {
        float dv = 0.f;
        if (gl < ng && q < P) {
          const float d = y - tg[gl * H + q];
          a.sq[static_cast<long>(g0 + gl) * P + q] = d * d;
          dv = 2.f * d * a.scale;
        }
        dyf[gl * H + q] = dv;
        const uint16_t db = to_bf16(dv);
        dyr[gl * sp + q] = db;
        dyT[q * kSG + gl] = db;
      }, bias + (NLAYER + 1) * H);
    }
    __syncthreads();
    for (int q = tid; q < P; q += kTW * 64) {
      float s = 0.f;
      for (int gl = 0; gl < ng; ++gl) s += dyf[gl * H + q];
      put(off_b(NLAYER + 1, P) + q, s);
    }
    tile_gemm<8, 16>(dyT, kSG, poolT, kSG, Pt, H / 16, 16, [&](int, int, int q, int c, float v) {
      if (q < P) put(off_w(NLAYER + 1) + static_cast<long>(q) * H + c, v);
    });
    st.store(R2, sp, H, Pt * 2);
    __syncthreads();
    st.load(catg + static_cast<long>(NLAYER - 1) * kTR * 2 * H, 2 * H, kTR, kTR, 2 * H / 8);
    tile_gemm<1, 128>(dyr, sp, R2, sp, 1, H / 16, Pt * 16, [&](int, int, int gl, int c, float v) {
      if (gl < ng) {
        const float dp = v / static_cast<float>(L);
        for (int i = 0; i < L; ++i) R3[(gl * L + i) * kSX + c] = dp;
      }
    });
    for (int e = R * kSX + tid; e < kTR * kSX; e += kTW * 64) R3[e] = 0.f;
    __syncthreads();
    // 5. layers NLAYER..1: dz, the bias and weight gradients, dx of the layer below
    for (int l = NLAYER; l >= 1; --l) {
      relu_bwd(mk + l * 32 * 4, true, off_b(l, P));
      st.store(R2, kSC, kTR, 2 * H / 8);               // the layer input, [row][j]
      __syncthreads();
      st.load(a.pbfT + off_wt(l) + static_cast<long>(H) * H, H, H, H, H / 8);
      tile_gemm_kn<8, kTR>(dzT, kSR, R2, kSC, 8, 2 * H / 16, 16 * mtr, [&](int, int, int c, int j, float v) {
        put(off_w(l) + static_cast<long>(c) * 2 * H + j, v);
      });
      // d[x | m] = dz W_l, the message half first
      __syncthreads();
      st.store(R2, kSH, H, H / 8);
      __syncthreads();
      st.load(a.pbfT + off_wt(l), H, H, H, H / 8);
      tile_gemm<4, H>(dzr, kSH, R2, kSH, mtr, H / 16, H, [&](int, int, int r, int c, float v) {
        R4[r * kSX + c] = v;
      });
      __syncthreads();
      st.store(R2, kSH, H, H / 8);
      __syncthreads();
      if (l > 1) st.load(catg + static_cast<long>(l - 2) * kTR * 2 * H, 2 * H, kTR, kTR, 2 * H / 8);
      tile_gemm<4, H>(dzr, kSH, R2, kSH, mtr, H / 16, H, [&](int, int, int r, int c, float v) {
        R3[r * kSX + c] = v;
      });
      __syncthreads();
      // dx = d[x] + A^T d[m] (A symmetric), rows beyond the block zero
      {
        float v[kRV];
        load_row(v, R3 + tr * kSX, rq);
        if (tr < R) {
          const int gl = tr / L, i = tr % L;
          const float* Ar = sA + (gl * L + i) * L;
          float s[kRV];
#pragma unroll
          for (int q = 0; q < kRV; ++q) s[q] = 0.f;
          int j = 0;
          for (; j + 1 < L; j += 2) {
            const float w0 = Ar[j], w1 = Ar[j + 1];
            float y0[kRV], y1[kRV];
            load_row(y0, R4 + (gl * L + j) * kSX, rq);
            load_row(y1, R4 + (gl * L + j + 1) * kSX, rq);
#pragma unroll
            for (int q = 0; q < kRV; ++q) s[q] = (s[q] + w0 * y0[q]) + w1 * y1[q];
          }
          if (j < L) {
            const float w = Ar[j];
            float y[kRV];
            load_row(y, R4 + (gl * L + j) * kSX, rq);
#pragma unroll
            for (int q = 0; q < kRV; ++q) s[q] += w * y[q];
          }
#pragma unroll
          for (int q = 0; q < kRV; ++q) v[q] += s[q];
        } else {
#pragma unroll
          for (int q = 0; q < kRV; ++q) v[q] = 0.f;
        }
        store_row(R3 + tr * kSX, v, rq);
      }
      __syncthreads();
    }
    // 6. layer 0
    relu_bwd(mk, false, off_b(0, P));
    tile_gemm<1, kTR>(dzT, kSR, featT, kSR, 8, 1, 16 * mtr, [&](int, int, int c, int f, float v) {
      put(off_w(0) + static_cast<long>(c) * F + f, v);
    });
    __syncthreads();
  }
}

// Sum of the workgroups' gradient partials in workgroup order, then (adam)
// Adam with bias corrections (b1t, b2t), or bt[2 step], bt[2 step + 1] when
// bt is given; writes the bf16 parameters and the transposed bf16 weights.
struct ReduceArgs {
  const float* part;
  int nwg;                      // partial rows; 0: no gradient (the bf16 copies only)
  long nparam;
  float* grad;
  float *p, *m, *v;
  int adam;
  float lr, b1t, b2t;
  const float* bt;
  const int* step;
  uint16_t *pbf, *pbfT;
  int P;
};

__global__ __launch_bounds__(256) void k_train_reduce(ReduceArgs a) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= a.nparam) return;
  float pv = a.p[i];
  if (a.nwg > 0) {
    // rows added in order, eight loads in flight at a time
    float s = 0.f;
    int b = 0;
    for (; b + 8 <= a.nwg; b += 8) {
      float x[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) x[q] = a.part[static_cast<long>(b + q) * a.nparam + i];
#pragma unroll
      for (int q = 0; q < 8; ++q) s += x[q];
    }
    for (; b < a.nwg; ++b) s += a.part[static_cast<long>(b) * a.nparam + i];
    a.grad[i] = s;
    if (a.adam) {
      float b1t = a.b1t, b2t = a.b2t;
      if (a.bt) {
        const int st = *a.step;
        b1t = a.bt[2 * st];
        b2t = a.bt[2 * st + 1];
      }
      const float mi = 0.9f * a.m[i] + 0.1f * s;
      const float vi = 0.999f * a.v[i] + 0.001f * s * s;
      a.m[i] = mi;
      a.v[i] = vi;
      const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
      pv -= a.lr * mh / (sqrtf(vh) + 1e-8f);
      a.p[i] = pv;
    }
  }
  const uint16_t pb = to_bf16(pv);
  a.pbf[i] = pb;
  const long o = i - off_w(1);
  if (o >= 0 && o < NLAYER * (2L * H * H + H)) {
    const long l = o / (2L * H * H + H), w = o % (2L * H * H + H);
    if (w < 2L * H * H) a.pbfT[off_wt(static_cast<int>(l) + 1) + (w % (2 * H)) * H + w / (2 * H)] = pb;
  }
  const long ow = i - off_w(NLAYER + 1);
  if (ow >= 0 && ow < static_cast<long>(a.P) * H)
    a.pbfT[off_wt(NLAYER + 1) + (ow % H) * ((a.P + 15) / 16 * 16) + ow / H] = pb;
}

// acc += sum(sq[0..n)) in a fixed order, then the step counter moves on
__global__ __launch_bounds__(256) void k_loss_step(const float* __restrict__ sq, long n, double* __restrict__ acc,
                                                   int* __restrict__ step) {
  __shared__ double part[256];
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) s += sq[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *acc += part[0];
    *step += 1;
  }
}

inline unsigned nb(long n, int b = 256) { return static_cast<unsigned>((n + b - 1) / b); }

}  // namespace
}  // namespace spx

using namespace spx;

struct spx_gnn {
  spx_ctx* ctx = nullptr;
  int P = 0;
  double sh = 1.0, sj = 1.0;
  // parameter layout
  size_t off_w[NLAYER + 2] = {}, off_b[NLAYER + 2] = {};
  size_t nparam = 0;
  float* p = nullptr;      // masters
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* pbf = nullptr; // bf16 copy of the parameters
  uint16_t* pbfT = nullptr; // bf16 W_l^T (l = 1..3) and W_out^T, for the backward products
  bool bf_stale = true;    // pbf / pbfT not yet written from the masters
  long step = 0;
  // activations (sized for the largest batch seen)
  long cap_rows = 0, cap_graphs = 0, cap_L = 0;
  float *Jin = nullptr, *hin = nullptr, *J = nullptr;  // raw couplings/fields (FP32), edge weights
  uint16_t* feat = nullptr;
  float *z[NLAYER + 1] = {}, *x[NLAYER + 1] = {};
  uint16_t* cat[NLAYER + 1] = {};
  float *pool = nullptr, *y = nullptr, *tgt = nullptr, *dy = nullptr, *sq = nullptr, *dpool = nullptr;
  uint16_t *poolbf = nullptr, *dybf = nullptr, *dzbf = nullptr;
  float *dx = nullptr, *dz = nullptr, *dcat = nullptr;
  // resident training set
  float *set_J = nullptr, *set_h = nullptr, *set_t = nullptr;
  uint32_t* set_idx = nullptr;
  int* d_step = nullptr;          // step of the running epoch (graph replays read it)
  float* d_bt = nullptr;          // per step (b1^t, b2^t)
  size_t bt_cap = 0;
  double* loss_acc = nullptr;
  size_t set_n = 0, set_L = 0;
  // pinned double buffers of the bulk prediction
  size_t pin_graphs = 0, pin_L = 0;
  float *pin_J[2] = {}, *pin_h[2] = {}, *pin_y[2] = {};
  hipEvent_t ev[2] = {};
  // the training kernel's gradient partials and saved layer inputs
  float* tpart = nullptr;
  uint16_t* tcat = nullptr;
  size_t tpart_cap = 0, tcat_cap = 0;
};

namespace {

size_t wsize(const spx_gnn* g, int l) {
  if (l == 0) return static_cast<size_t>(H) * F;
  if (l <= NLAYER) return static_cast<size_t>(H) * 2 * H;
  return static_cast<size_t>(g->P) * H;
}

size_t bsize(const spx_gnn* g, int l) { return l <= NLAYER ? H : g->P; }

int gemm(spx_gnn* net, const uint16_t* a, long sam, long sak, const uint16_t* b, long sbk, long sbn, float* out,
         long scm, long scn, int M, int N, int K, const float* bias, float* preact, int relu, int acc) {
  GemmArgs g{a, sam, sak, b, sbk, sbn, out, scm, scn, M, N, K, bias, preact, relu, acc};
  k_gemm<<<dim3(static_cast<unsigned>((N + kBN - 1) / kBN), static_cast<unsigned>((M + kBM - 1) / kBM)), 256, 0,
           net->ctx->stream>>>(g);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <class T>
int dalloc(T** p, long n) {
  if (*p) SPX_HIP(hipFree(*p));
  SPX_HIP(hipMalloc(p, (n > 0 ? n : 1) * sizeof(T)));
  return SPX_OK;
}

int reserve(spx_gnn* g, long G, long L) {
  const long rows = G * L;
  if (rows <= g->cap_rows && G <= g->cap_graphs && L * L * G <= g->cap_L) return SPX_OK;
  const long R = std::max(rows, g->cap_rows), GG = std::max(G, g->cap_graphs);
  const long JJ = std::max(L * L * G, g->cap_L);
  SPX_TRY(dalloc(&g->Jin, JJ));
  SPX_TRY(dalloc(&g->hin, R));
  SPX_TRY(dalloc(&g->J, JJ));
  SPX_TRY(dalloc(&g->feat, R * F));
  for (int l = 0; l <= NLAYER; ++l) {
    SPX_TRY(dalloc(&g->z[l], R * H));
    SPX_TRY(dalloc(&g->x[l], R * H));
    SPX_TRY(dalloc(&g->cat[l], R * 2 * H));
  }
  SPX_TRY(dalloc(&g->pool, GG * H));
  SPX_TRY(dalloc(&g->poolbf, GG * H));
  SPX_TRY(dalloc(&g->y, GG * g->P));
  SPX_TRY(dalloc(&g->tgt, GG * g->P));
  SPX_TRY(dalloc(&g->dy, GG * g->P));
  SPX_TRY(dalloc(&g->dybf, GG * g->P));
  SPX_TRY(dalloc(&g->sq, GG * g->P));
  SPX_TRY(dalloc(&g->dpool, GG * H));
  SPX_TRY(dalloc(&g->dx, R * H));
  SPX_TRY(dalloc(&g->dz, R * H));
  SPX_TRY(dalloc(&g->dzbf, R * H));
  SPX_TRY(dalloc(&g->dcat, R * 2 * H));
  g->cap_rows = R;
  g->cap_graphs = GG;
  g->cap_L = JJ;
  return SPX_OK;
}

// Upload the couplings (FP32), the fields and, when given, the targets of a
// batch. The backward pass takes the message transpose as A itself, so J
// must be symmetric.
int load_batch(spx_gnn* g, long G, long L, const double* J, const double* h, const float* target) {
  if (!J || !h) return fail(SPX_EINVAL, "null couplings or fields");
  for (long gi = 0; gi < G; ++gi)
    for (long i = 0; i < L; ++i)
      for (long j = 0; j < i; ++j)
        if (J[(gi * L + i) * L + j] != J[(gi * L + j) * L + i])
          return fail(SPX_EINVAL, "coupling matrix of graph " + std::to_string(gi) + " is not symmetric");
  SPX_TRY(reserve(g, G, L));
  std::vector<float> Jf(G * L * L), hf(G * L);
  for (long q = 0; q < G * L * L; ++q) Jf[q] = static_cast<float>(J[q]);
  for (long q = 0; q < G * L; ++q) hf[q] = static_cast<float>(h[q]);
  spx_ctx* c = g->ctx;
  SPX_TRY(copy_to_device(c, g->Jin, Jf.data(), Jf.size() * sizeof(float)));
  SPX_TRY(copy_to_device(c, g->hin, hf.data(), hf.size() * sizeof(float)));
  if (target) SPX_TRY(copy_to_device(c, g->tgt, target, G * g->P * sizeof(float)));
  return SPX_OK;
}

int forward(spx_gnn* g, long G, long L) {
  spx_ctx* c = g->ctx;
  const long R = G * L;
  k_graph_inputs<<<static_cast<unsigned>(G), 32, 0, c->stream>>>(g->Jin, g->hin, g->J, g->feat, static_cast<int>(L),
                                                                  static_cast<float>(1.0 / g->sj),
                                                                  static_cast<float>(1.0 / g->sh));
  k_cast<<<nb(static_cast<long>(g->nparam)), 256, 0, c->stream>>>(g->p, g->pbf, static_cast<long>(g->nparam));
  // x0 = relu(feat W0^T + b0)
  SPX_TRY(gemm(g, g->feat, F, 1, g->pbf + g->off_w[0], 1, F, g->x[0], H, 1, static_cast<int>(R), H, F,
               g->p + g->off_b[0], g->z[0], 1, 0));
  for (int l = 1; l <= NLAYER; ++l) {
    k_mix_cat<<<dim3(static_cast<unsigned>(G), static_cast<unsigned>(L)), H, 0, c->stream>>>(
        g->J, g->x[l - 1], g->cat[l], static_cast<int>(L));
    SPX_TRY(gemm(g, g->cat[l], 2 * H, 1, g->pbf + g->off_w[l], 1, 2 * H, g->x[l], H, 1, static_cast<int>(R),
                 H, 2 * H, g->p + g->off_b[l], g->z[l], 1, 0));
  }
  k_pool<<<static_cast<unsigned>(G), H, 0, c->stream>>>(g->x[NLAYER], g->pool, static_cast<int>(L));
  k_cast<<<nb(G * H), 256, 0, c->stream>>>(g->pool, g->poolbf, G * H);
  SPX_TRY(gemm(g, g->poolbf, H, 1, g->pbf + g->off_w[NLAYER + 1], 1, H, g->y, g->P, 1, static_cast<int>(G),
               g->P, H, g->p + g->off_b[NLAYER + 1], nullptr, 0, 0));
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

// Graphs per block of k_train: whole graphs up to 32 rows (one graph when it
// is larger), so that the forward products give every wave one M tile.
// Measured on 32768-graph epochs (80 -> 73 us per step at L = 10 and 12)
// against filling the 64 rows.
int graphs_per_block(long L) { return static_cast<int>(std::max<long>(1, std::min<long>(16, 32 / L))); }

ReduceArgs reduce_args(const spx_gnn* g, int nwg) {
  ReduceArgs r{};
  r.part = g->tpart;
  r.nwg = nwg;
  r.nparam = static_cast<long>(g->nparam);
  r.grad = g->grad;
  r.p = g->p;
  r.m = g->m;
  r.v = g->v;
  r.pbf = g->pbf;
  r.pbfT = g->pbfT;
  r.P = g->P;
  return r;
}

// The bf16 parameters and transposed weights from the masters, when a
// training step has not written them since the masters last changed.
int refresh_bf16(spx_gnn* g) {
  if (!g->bf_stale) return SPX_OK;
  k_train_reduce<<<nb(static_cast<long>(g->nparam)), 256, 0, g->ctx->stream>>>(reduce_args(g, 0));
  SPX_HIP(hipGetLastError());
  g->bf_stale = false;
  return SPX_OK;
}

// One forward + backward pass over a batch of G graphs of L spins (k_train),
// its gradient partials summed in workgroup order (k_train_reduce), with an
// Adam step when adam != 0 (bias corrections b1t, b2t, or read from bt at the
// device step counter when bt is given). The batch is J, h, t directly, or
// rows order[*step G + i] of them.
int train_pass(spx_gnn* g, long G, long L, const float* J, const float* h, const float* t, const uint32_t* order,
               const int* step, int adam, float lr, float b1t, float b2t, const float* bt) {
  spx_ctx* c = g->ctx;
  const int gb = graphs_per_block(L);
  const int nblk = static_cast<int>((G + gb - 1) / gb);
  const int nwg = std::min(nblk, kTrainWG);
  const size_t need_part = static_cast<size_t>(nwg) * g->nparam;
  const size_t need_cat = static_cast<size_t>(nwg) * NLAYER * kTR * 2 * H;
  if (need_part > g->tpart_cap || need_cat > g->tcat_cap) {
    SPX_HIP(hipStreamSynchronize(c->stream));
    for (void** q : {reinterpret_cast<void**>(&g->tpart), reinterpret_cast<void**>(&g->tcat)}) {
      if (*q) SPX_HIP(hipFree(*q));
      *q = nullptr;
    }
    g->tpart_cap = g->tcat_cap = 0;
    SPX_HIP(hipMalloc(&g->tpart, need_part * sizeof(float)));
    SPX_HIP(hipMalloc(&g->tcat, need_cat * sizeof(uint16_t)));
    g->tpart_cap = need_part;
    g->tcat_cap = need_cat;
  }
  SPX_TRY(refresh_bf16(g));
  TrainArgs a{};
  a.J = J;
  a.h = h;
  a.t = t;
  a.order = order;
  a.step = step;
  a.G = static_cast<int>(G);
  a.L = static_cast<int>(L);
  a.P = g->P;
  a.gb = gb;
  a.inv_sj = static_cast<float>(1.0 / g->sj);
  a.inv_sh = static_cast<float>(1.0 / g->sh);
  a.scale = 1.f / static_cast<float>(G * g->P);
  a.p = g->p;
  a.pbf = g->pbf;
  a.pbfT = g->pbfT;
  a.nparam = static_cast<long>(g->nparam);
  a.part = g->tpart;
  a.catg = g->tcat;
  a.sq = g->sq;
  for (int b0 = 0; b0 < nblk; b0 += kTrainWG) {
    a.blk0 = b0;
    k_train<<<static_cast<unsigned>(std::min(kTrainWG, nblk - b0)), kTW * 64, kTrainLds, c->stream>>>(a);
    SPX_HIP(hipGetLastError());
  }
  ReduceArgs r = reduce_args(g, nwg);
  r.adam = adam;
  r.lr = lr;
  r.b1t = b1t;
  r.b2t = b2t;
  r.bt = bt;
  r.step = step;
  k_train_reduce<<<nb(static_cast<long>(g->nparam)), 256, 0, c->stream>>>(r);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

// The batch's mean squared error from the per-point terms, summed in order.
int batch_loss(spx_gnn* g, long G, float* loss) {
  std::vector<float> sq(G * g->P);
  SPX_TRY(copy_to_host(g->ctx, sq.data(), g->sq, sq.size() * sizeof(float)));
  double s = 0.0;
  for (float v : sq) s += v;
  *loss = static_cast<float>(s / static_cast<double>(G * g->P));
  return SPX_OK;
}

int reserve_pinned(spx_gnn* g, size_t G, size_t L) {
  if (G <= g->pin_graphs && L <= g->pin_L) return SPX_OK;
  for (int b = 0; b < 2; ++b) {
    for (float** q : {&g->pin_J[b], &g->pin_h[b], &g->pin_y[b]}) {
      if (*q) SPX_HIP(hipHostFree(*q));
      *q = nullptr;
    }
    SPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->pin_J[b]), G * L * L * sizeof(float)));
    SPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->pin_h[b]), G * L * sizeof(float)));
    SPX_HIP(hipHostMalloc(reinterpret_cast<void**>(&g->pin_y[b]), G * g->P * sizeof(float)));
  }
  g->pin_graphs = G;
  g->pin_L = L;
  return SPX_OK;
}

int check_batch(const spx_gnn* g, size_t G, size_t L) {
  if (!g) return fail(SPX_EINVAL, "null network");
  if (G == 0 || L == 0 || L > kMaxSpins) return fail(SPX_EINVAL, "empty batch or too many spins");
  return SPX_OK;
}

}  // namespace

extern "C" {

int spx_gnn_create(spx_ctx* ctx, size_t outputs, uint64_t seed, double scale_h, double scale_j, spx_gnn** out) {
  if (!ctx || !out) return fail(SPX_EINVAL, "null argument");
  if (outputs == 0 || outputs > 128) return fail(SPX_EINVAL, "outputs must be in [1, 128]");
  if (!(scale_h > 0.0) || !(scale_j > 0.0)) return fail(SPX_EINVAL, "feature scales must be positive");
  auto* g = new spx_gnn();
  g->ctx = ctx;
  g->P = static_cast<int>(outputs);
  g->sh = scale_h;
  g->sj = scale_j;
  size_t off = 0;
  for (int l = 0; l <= NLAYER + 1; ++l) {
    g->off_w[l] = off;
    off += wsize(g, l);
    g->off_b[l] = off;
    off += bsize(g, l);
  }
  g->nparam = off;
  // the device kernels compute the same offsets from the layout
  for (int l = 0; l <= NLAYER + 1; ++l)
    if (static_cast<long>(g->off_w[l]) != off_w(l) || static_cast<long>(g->off_b[l]) != off_b(l, g->P)) {
      delete g;
      return fail(SPX_EINVAL, "parameter layout mismatch");
    }
  // W_out^T rows padded to a multiple of 16 outputs (zero padding)
  const long offT = off_wt(NLAYER + 1) + static_cast<long>(H) * ((g->P + 15) / 16 * 16);
  // He-uniform weights, zero biases, from a seeded host stream
  std::vector<float> p(off, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l <= NLAYER + 1; ++l) {
    const double fan_in = l == 0 ? 4.0 : l <= NLAYER ? 2.0 * H : H;
    const double lim = std::sqrt(6.0 / fan_in) * (l <= NLAYER ? 1.0 : 0.1);
    std::uniform_real_distribution<double> u(-lim, lim);
    for (size_t i = 0; i < wsize(g, l); ++i) {
      const size_t col = l == 0 ? i % F : 0;
      p[g->off_w[l] + i] = (l == 0 && col >= 4) ? 0.f : static_cast<float>(u(rng));
    }
  }
  int rc = SPX_OK;
  if (hipMalloc(&g->p, off * sizeof(float)) != hipSuccess || hipMalloc(&g->grad, off * sizeof(float)) != hipSuccess ||
      hipMalloc(&g->m, off * sizeof(float)) != hipSuccess || hipMalloc(&g->v, off * sizeof(float)) != hipSuccess ||
      hipMalloc(&g->pbf, off * sizeof(uint16_t)) != hipSuccess ||
      hipMalloc(&g->pbfT, static_cast<size_t>(offT) * sizeof(uint16_t)) != hipSuccess)
    rc = fail(SPX_ENOMEM, "hipMalloc of network parameters failed");
  else if (copy_to_device(g->ctx, g->p, p.data(), off * sizeof(float)) != SPX_OK ||
           hipMemset(g->m, 0, off * sizeof(float)) != hipSuccess ||
           hipMemset(g->v, 0, off * sizeof(float)) != hipSuccess ||
           hipMemset(g->pbfT, 0, static_cast<size_t>(offT) * sizeof(uint16_t)) != hipSuccess)
    rc = fail(SPX_EDEVICE, "initialising network parameters failed");
  for (int b = 0; b < 2 && rc == SPX_OK; ++b)
    if (hipEventCreateWithFlags(&g->ev[b], hipEventDisableTiming) != hipSuccess)
      rc = fail(SPX_EDEVICE, "hipEventCreate failed");
  if (rc != SPX_OK) {
    spx_gnn_destroy(g);
    return rc;
  }
  *out = g;
  return SPX_OK;
}

int spx_gnn_destroy(spx_gnn* g) {
  if (!g) return SPX_OK;
  for (void* q : {static_cast<void*>(g->p), static_cast<void*>(g->grad), static_cast<void*>(g->m),
                  static_cast<void*>(g->v), static_cast<void*>(g->pbf), static_cast<void*>(g->J),
                  static_cast<void*>(g->Jin), static_cast<void*>(g->hin), static_cast<void*>(g->feat),
                  static_cast<void*>(g->pool), static_cast<void*>(g->y), static_cast<void*>(g->tgt),
                  static_cast<void*>(g->dy), static_cast<void*>(g->sq), static_cast<void*>(g->dpool),
                  static_cast<void*>(g->poolbf), static_cast<void*>(g->dybf), static_cast<void*>(g->dzbf),
                  static_cast<void*>(g->dx), static_cast<void*>(g->dz), static_cast<void*>(g->dcat),
                  static_cast<void*>(g->pbfT), static_cast<void*>(g->tpart), static_cast<void*>(g->tcat),
                  static_cast<void*>(g->d_step), static_cast<void*>(g->d_bt)})
    if (q) (void)hipFree(q);
  for (int l = 0; l <= NLAYER; ++l)
    for (void* q : {static_cast<void*>(g->z[l]), static_cast<void*>(g->x[l]), static_cast<void*>(g->cat[l])})
      if (q) (void)hipFree(q);
  for (int b = 0; b < 2; ++b) {
    for (float* q : {g->pin_J[b], g->pin_h[b], g->pin_y[b]})
      if (q) (void)hipHostFree(q);
    if (g->ev[b]) (void)hipEventDestroy(g->ev[b]);
  }
  for (void* q : {static_cast<void*>(g->set_J), static_cast<void*>(g->set_h), static_cast<void*>(g->set_t),
                  static_cast<void*>(g->set_idx), static_cast<void*>(g->loss_acc)})
    if (q) (void)hipFree(q);
  delete g;
  return SPX_OK;
}

int spx_gnn_num_params(const spx_gnn* g, size_t* n) {
  if (!g || !n) return fail(SPX_EINVAL, "null argument");
  *n = g->nparam;
  return SPX_OK;
}

int spx_gnn_get_params(const spx_gnn* g, float* host) {
  if (!g || !host) return fail(SPX_EINVAL, "null argument");
  return copy_to_host(g->ctx, host, g->p, g->nparam * sizeof(float));
}

int spx_gnn_set_params(spx_gnn* g, const float* host) {
  if (!g || !host) return fail(SPX_EINVAL, "null argument");
  g->bf_stale = true;
  return copy_to_device(g->ctx, g->p, host, g->nparam * sizeof(float));
}

int spx_gnn_predict(spx_gnn* g, size_t G, size_t L, const double* J, const double* h, float* pred) {
  SPX_TRY(check_batch(g, G, L));
  if (!pred) return fail(SPX_EINVAL, "null prediction buffer");
  SPX_TRY(load_batch(g, static_cast<long>(G), static_cast<long>(L), J, h, nullptr));
  SPX_TRY(forward(g, static_cast<long>(G), static_cast<long>(L)));
  return copy_to_host(g->ctx, pred, g->y, G * g->P * sizeof(float));
}

int spx_gnn_predict_dipolar(spx_gnn* g, uint64_t seed, size_t first, size_t count, size_t nspins, double scale,
                            double min_separation, int angular, double disorder, float* pred) {
  if (!g || !pred) return fail(SPX_EINVAL, "null argument");
  const DipolarSpec s{seed, nspins, scale, min_separation, angular != 0, disorder};
  SPX_TRY(check_spec(s));
  if (count == 0) return SPX_OK;
  const size_t C = std::min(count, kPredictChunk), L = nspins, P = static_cast<size_t>(g->P);
  SPX_TRY(reserve(g, static_cast<long>(C), static_cast<long>(L)));
  SPX_TRY(reserve_pinned(g, C, L));
  hipStream_t st = g->ctx->stream;
  const size_t chunks = (count + C - 1) / C;
  // Chunk k is drawn on the host while the device runs chunk k - 1.
  for (size_t k = 0; k <= chunks; ++k) {
    if (k < chunks) {
      const int b = static_cast<int>(k & 1);
      const size_t n = std::min(C, count - k * C);
      SPX_HIP(hipEventSynchronize(g->ev[b]));
      SPX_TRY(draw_graphs(s, first + k * C, n, nullptr, nullptr, g->pin_J[b], g->pin_h[b], 0));
      SPX_HIP(hipMemcpyAsync(g->Jin, g->pin_J[b], n * L * L * sizeof(float), hipMemcpyHostToDevice, st));
      SPX_HIP(hipMemcpyAsync(g->hin, g->pin_h[b], n * L * sizeof(float), hipMemcpyHostToDevice, st));
      SPX_TRY(forward(g, static_cast<long>(n), static_cast<long>(L)));
      SPX_HIP(hipMemcpyAsync(g->pin_y[b], g->y, n * P * sizeof(float), hipMemcpyDeviceToHost, st));
      SPX_HIP(hipEventRecord(g->ev[b], st));
    }
    if (k >= 1) {
      const int b = static_cast<int>((k - 1) & 1);
      const size_t n = std::min(C, count - (k - 1) * C);
      SPX_HIP(hipEventSynchronize(g->ev[b]));
      std::memcpy(pred + (k - 1) * C * P, g->pin_y[b], n * P * sizeof(float));
    }
  }
  return SPX_OK;
}

int spx_gnn_gradient(spx_gnn* g, size_t G, size_t L, const double* J, const double* h, const float* target,
                     float* loss, float* grad) {
  SPX_TRY(check_batch(g, G, L));
  if (!target) return fail(SPX_EINVAL, "null targets");
  SPX_TRY(load_batch(g, static_cast<long>(G), static_cast<long>(L), J, h, target));
  SPX_TRY(train_pass(g, static_cast<long>(G), static_cast<long>(L), g->Jin, g->hin, g->tgt, nullptr, nullptr, 0, 0.f,
                     0.f, 0.f, nullptr));
  if (loss) SPX_TRY(batch_loss(g, static_cast<long>(G), loss));
  if (grad) SPX_TRY(copy_to_host(g->ctx, grad, g->grad, g->nparam * sizeof(float)));
  return SPX_OK;
}

int spx_gnn_train_step(spx_gnn* g, size_t G, size_t L, const double* J, const double* h, const float* target,
                       float lr, float* loss) {
  SPX_TRY(check_batch(g, G, L));
  if (!target) return fail(SPX_EINVAL, "null targets");
  SPX_TRY(load_batch(g, static_cast<long>(G), static_cast<long>(L), J, h, target));
  const long t = g->step + 1;
  const float b1t = static_cast<float>(std::pow(0.9, static_cast<double>(t)));
  const float b2t = static_cast<float>(std::pow(0.999, static_cast<double>(t)));
  SPX_TRY(train_pass(g, static_cast<long>(G), static_cast<long>(L), g->Jin, g->hin, g->tgt, nullptr, nullptr, 1, lr,
                     b1t, b2t, nullptr));
  g->step = t;
  if (loss) SPX_TRY(batch_loss(g, static_cast<long>(G), loss));
  return SPX_OK;
}

int spx_gnn_set_training_set(spx_gnn* g, size_t n, size_t L, const double* J, const double* h,
                             const float* target) {
  if (!g || !J || !h || !target) return fail(SPX_EINVAL, "null argument");
  if (n == 0 || L == 0 || L > kMaxSpins) return fail(SPX_EINVAL, "empty set or too many spins");
  for (size_t gi = 0; gi < n; ++gi)
    for (size_t i = 0; i < L; ++i)
      for (size_t j = 0; j < i; ++j)
        if (J[(gi * L + i) * L + j] != J[(gi * L + j) * L + i])
          return fail(SPX_EINVAL, "coupling matrix of graph " + std::to_string(gi) + " is not symmetric");
  for (void* q : {static_cast<void*>(g->set_J), static_cast<void*>(g->set_h), static_cast<void*>(g->set_t),
                  static_cast<void*>(g->set_idx)})
    if (q) SPX_HIP(hipFree(q));
  g->set_J = g->set_h = g->set_t = nullptr;
  g->set_idx = nullptr;
  g->set_n = 0;
  std::vector<float> Jf(n * L * L), hf(n * L);
  for (size_t q = 0; q < Jf.size(); ++q) Jf[q] = static_cast<float>(J[q]);
  for (size_t q = 0; q < hf.size(); ++q) hf[q] = static_cast<float>(h[q]);
  SPX_HIP(hipMalloc(&g->set_J, Jf.size() * sizeof(float)));
  SPX_HIP(hipMalloc(&g->set_h, hf.size() * sizeof(float)));
  SPX_HIP(hipMalloc(&g->set_t, n * g->P * sizeof(float)));
  SPX_HIP(hipMalloc(&g->set_idx, n * sizeof(uint32_t)));
  if (!g->loss_acc) SPX_HIP(hipMalloc(&g->loss_acc, sizeof(double)));
  SPX_TRY(copy_to_device(g->ctx, g->set_J, Jf.data(), Jf.size() * sizeof(float)));
  SPX_TRY(copy_to_device(g->ctx, g->set_h, hf.data(), hf.size() * sizeof(float)));
  SPX_TRY(copy_to_device(g->ctx, g->set_t, target, n * g->P * sizeof(float)));
  g->set_n = n;
  g->set_L = L;
  return SPX_OK;
}

int spx_gnn_train_epoch(spx_gnn* g, size_t batch, const uint32_t* order, float lr, float* mean_loss) {
  if (!g || !order) return fail(SPX_EINVAL, "null argument");
  if (g->set_n == 0) return fail(SPX_EINVAL, "no training set");
  if (batch == 0 || batch > g->set_n) return fail(SPX_EINVAL, "batch must be in [1, set size]");
  for (size_t i = 0; i < g->set_n; ++i)
    if (order[i] >= g->set_n) return fail(SPX_ERANGE, "order index out of range");
  const long G = static_cast<long>(batch), L = static_cast<long>(g->set_L);
  SPX_TRY(reserve(g, G, L));
  spx_ctx* c = g->ctx;
  SPX_TRY(copy_to_device(c, g->set_idx, order, g->set_n * sizeof(uint32_t)));
  SPX_HIP(hipMemsetAsync(g->loss_acc, 0, sizeof(double), c->stream));
  const size_t steps = g->set_n / batch;
  // Adam's bias corrections of every step of the epoch, as train_step
  // computes them, and the device step counter
  std::vector<float> bt(2 * steps);
  for (size_t st = 0; st < steps; ++st) {
    const double t = static_cast<double>(g->step + 1 + st);
    bt[2 * st] = static_cast<float>(std::pow(0.9, t));
    bt[2 * st + 1] = static_cast<float>(std::pow(0.999, t));
  }
  if (2 * steps > g->bt_cap) {
    SPX_HIP(hipStreamSynchronize(c->stream));
    if (g->d_bt) SPX_HIP(hipFree(g->d_bt));
    g->d_bt = nullptr;
    g->bt_cap = 0;
    SPX_HIP(hipMalloc(&g->d_bt, 2 * steps * sizeof(float)));
    g->bt_cap = 2 * steps;
  }
  if (!g->d_step) SPX_HIP(hipMalloc(&g->d_step, sizeof(int)));
  SPX_TRY(copy_to_device(c, g->d_bt, bt.data(), bt.size() * sizeof(float)));
  SPX_HIP(hipMemsetAsync(g->d_step, 0, sizeof(int), c->stream));
  auto one_step = [&]() -> int {
    SPX_TRY(train_pass(g, G, L, g->set_J, g->set_h, g->set_t, g->set_idx, g->d_step, 1, lr, 0.f, 0.f, g->d_bt));
    k_loss_step<<<1, 256, 0, c->stream>>>(g->sq, G * g->P, g->loss_acc, g->d_step);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  };
  // the first step eagerly (it sizes every buffer), the others as replays
  // of one captured step; a capture the runtime refuses leaves them eager
  SPX_TRY(one_step());
  size_t done = 1;
  if (steps > 1 && hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    const int rc = one_step();
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    hipGraphExec_t exec = nullptr;
    if (rc == SPX_OK && ec == hipSuccess && graph &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      hipError_t e = hipSuccess;
      for (; done < steps && e == hipSuccess; ++done) e = hipGraphLaunch(exec, c->stream);
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(SPX_EDEVICE, std::string("replaying a training step: ") + hipGetErrorString(e));
    } else {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
    }
  }
  for (; done < steps; ++done) SPX_TRY(one_step());
  g->step += static_cast<long>(steps);
  SPX_HIP(hipGetLastError());
  double acc = 0.0;
  SPX_TRY(copy_to_host(c, &acc, g->loss_acc, sizeof(double)));
  if (mean_loss) *mean_loss = static_cast<float>(acc / static_cast<double>(steps * batch * g->P));
  return SPX_OK;
}

}  // extern "C"
