// The surrogate network of the control variate (include/spx_gnn.h):
// forward, backward and Adam on the device. Every matrix product is one
// strided GEMM kernel on BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulation): a wavefront owns a 16x16 tile of C, four wavefronts a
// 32x32 block, K is walked 16 at a time. Strides give the transposed
// operands of the backward products without copies.
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
// C (each wavefront 32 x 32, four 16 x 16 MFMA accumulators) and walks its
// K range 32 at a time through LDS: A as [m][k], B transposed as [n][k], so
// every MFMA operand fragment is four consecutive bf16 of one LDS row. An
// operand whose K (or M / N) index is contiguous in memory is loaded eight
// elements (16 bytes) per thread; other strides element by element. When
// the output has few tiles and K is long (the weight gradients reduce over
// every row of the batch), K is split across workgroups into partial tiles
// that k_gemm_reduce sums in a fixed order, so results do not depend on
// scheduling.
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

// part: null for a single K range (the epilogue is applied here), else
// [gridDim.z][M][N] partial sums of K ranges of kchunk each.
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g, float* __restrict__ part, int kchunk) {
  __shared__ __align__(16) uint16_t As[kBM][kLP];
  __shared__ __align__(16) uint16_t Bs[kBN][kLP];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bm = blockIdx.y * kBM, bn = blockIdx.x * kBN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(g.K, kbeg + kchunk);
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
        if (part)
          part[(static_cast<long>(blockIdx.z) * g.M + row) * g.N + col] = acc[i][j][r];
        else
          gemm_store(g, row, col, acc[i][j][r]);
      }
    }
}

// Sum of the split-K partials in K-range order, then the epilogue.
__global__ __launch_bounds__(256) void k_gemm_reduce(GemmArgs g, const float* __restrict__ part, int splits) {
  const long i = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long mn = static_cast<long>(g.M) * g.N;
  if (i >= mn) return;
  float v = 0.f;
  for (int s = 0; s < splits; ++s) v += part[s * mn + i];
  gemm_store(g, static_cast<int>(i / g.N), static_cast<int>(i % g.N), v);
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

// dx = d[x | m][:, :H] + A^T d[x | m][:, H:] (A symmetric): the backward of
// k_mix_cat, reading both halves of the product's rows in place.
__global__ void k_mix_bwd(const float* __restrict__ J, const float* __restrict__ dcat, float* __restrict__ dx,
                          int L) {
  const int g = blockIdx.x, i = blockIdx.y, c = threadIdx.x;
  const float* Jr = J + (static_cast<long>(g) * L + i) * L;
  const float* dg = dcat + static_cast<long>(g) * L * 2 * H;
  float s = 0.f;
  for (int j = 0; j < L; ++j) s += Jr[j] * dg[j * 2 * H + H + c];
  dx[(static_cast<long>(g) * L + i) * H + c] = dg[i * 2 * H + c] + s;
}

__global__ void k_pool(const float* __restrict__ x, float* __restrict__ p, int L) {
  const int g = blockIdx.x, c = threadIdx.x;
  float s = 0.f;
  for (int i = 0; i < L; ++i) s += x[(static_cast<long>(g) * L + i) * H + c];
  p[g * H + c] = s / static_cast<float>(L);
}

__global__ void k_pool_bwd(const float* __restrict__ dp, float* __restrict__ dx, int L) {
  const int g = blockIdx.x, i = blockIdx.y, c = threadIdx.x;
  dx[(static_cast<long>(g) * L + i) * H + c] = dp[g * H + c] / static_cast<float>(L);
}

// d (MSE) / dy and the loss terms (summed on the host in a fixed order)
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                      uint16_t* __restrict__ dybf, float* __restrict__ sq, long n, float scale) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = y[i] - t[i];
  const float v = 2.f * d * scale;
  dy[i] = v;
  dybf[i] = to_bf16(v);
  sq[i] = d * d;
}

// dz = dx where the pre-activation is positive, in FP32 and in bf16
__global__ void k_relu_bwd(const float* __restrict__ dx, const float* __restrict__ z, float* __restrict__ dz,
                           uint16_t* __restrict__ dzbf, long n) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = z[i] > 0.f ? dx[i] : 0.f;
  dz[i] = v;
  dzbf[i] = to_bf16(v);
}

// Column sums of a rows x cols matrix: a block owns 64 columns; its 16 row
// lanes sum rows r = lane (mod 16) (coalesced 256-byte rows), then a fixed
// tree over the lanes, so the result does not depend on scheduling.
__global__ __launch_bounds__(1024) void k_colsum(const float* __restrict__ x, float* __restrict__ out, long rows,
                                                 int cols) {
  __shared__ float part[16][64];
  const int cl = threadIdx.x & 63, lane = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  float s = 0.f;
  if (c < cols) {
#pragma unroll 4
    for (long r = lane; r < rows; r += 16) s += x[r * cols + c];
  }
  part[lane][cl] = s;
  __syncthreads();
  for (int w = 8; w > 0; w >>= 1) {
    if (lane < w) part[lane][cl] += part[lane + w][cl];
    __syncthreads();
  }
  if (lane == 0 && c < cols) out[c] = part[0][cl];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1t, float b2t) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  m[i] = 0.9f * m[i] + 0.1f * gi;
  v[i] = 0.999f * v[i] + 0.001f * gi * gi;
  const float mh = m[i] / (1.f - b1t), vh = v[i] / (1.f - b2t);
  p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Minibatch rows of the resident training set (graph i of step s's batch is
// graph order[s batch + i] of the set: couplings, fields, targets), in the
// replayable form of an epoch's step (one HIP graph for every step but the
// first): the step index lives on the device, so the minibatch rows and
// Adam's bias corrections of step s are read at launch, not baked into the
// launch arguments.
__global__ void k_gather_step(const float* __restrict__ sJ, const float* __restrict__ sh,
                              const float* __restrict__ st, const uint32_t* __restrict__ order,
                              const int* __restrict__ step, int batch, int L, int P, float* __restrict__ J,
                              float* __restrict__ h, float* __restrict__ t) {
  const int i = blockIdx.x;
  const long src = order[static_cast<long>(*step) * batch + i];
  for (int q = threadIdx.x; q < L * L; q += blockDim.x) J[static_cast<long>(i) * L * L + q] = sJ[src * L * L + q];
  for (int q = threadIdx.x; q < L; q += blockDim.x) h[static_cast<long>(i) * L + q] = sh[src * L + q];
  for (int q = threadIdx.x; q < P; q += blockDim.x) t[static_cast<long>(i) * P + q] = st[src * P + q];
}

// k_adam with this step's (b1^t, b2^t) from bt[2 step], bt[2 step + 1]
__global__ void k_adam_step(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, const float* __restrict__ bt,
                            const int* __restrict__ step) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = *step;
  const float b1t = bt[2 * s], b2t = bt[2 * s + 1];
  const float gi = g[i];
  m[i] = 0.9f * m[i] + 0.1f * gi;
  v[i] = 0.999f * v[i] + 0.001f * gi * gi;
  const float mh = m[i] / (1.f - b1t), vh = v[i] / (1.f - b2t);
  p[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

__global__ void k_step_next(int* step) {
  if (threadIdx.x == 0) *step += 1;
}

// acc += sum(sq[0..n)) in a fixed order (one block, strided partial sums,
// then a tree), so the epoch loss does not depend on scheduling.
__global__ __launch_bounds__(256) void k_loss_acc(const float* __restrict__ sq, long n, double* __restrict__ acc) {
  __shared__ double part[256];
  double s = 0.0;
  for (long i = threadIdx.x; i < n; i += 256) s += sq[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *acc += part[0];
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
  // split-K partial tiles of the products
  float* gemm_part = nullptr;
  size_t gemm_part_cap = 0;
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
  spx_ctx* c = net->ctx;
  GemmArgs g{a, sam, sak, b, sbk, sbn, out, scm, scn, M, N, K, bias, preact, relu, acc};
  const int tiles = ((N + kBN - 1) / kBN) * ((M + kBM - 1) / kBM);
  // Split K when the output alone would leave most CUs idle and K is long:
  // that is the weight gradients, whose K is the batch's rows. The forward
  // products (K = 16, 128, 256) never split, so a prediction does not
  // depend on the batch it is computed in.
  int splits = 1;
  if (tiles < 128 && K >= 512) splits = std::min({32, K / (2 * kBK), std::max(1, 256 / tiles)});
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + kBK - 1) / kBK * kBK;
  splits = (K + kchunk - 1) / kchunk;
  if (splits <= 1) {
    k_gemm<<<dim3(static_cast<unsigned>((N + kBN - 1) / kBN), static_cast<unsigned>((M + kBM - 1) / kBM)), 256, 0,
             c->stream>>>(g, nullptr, K > 0 ? K : 1);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  }
  const size_t need = static_cast<size_t>(splits) * M * N;
  if (need > net->gemm_part_cap) {
    SPX_HIP(hipStreamSynchronize(c->stream));
    if (net->gemm_part) SPX_HIP(hipFree(net->gemm_part));
    net->gemm_part = nullptr;
    net->gemm_part_cap = 0;
    SPX_HIP(hipMalloc(&net->gemm_part, need * sizeof(float)));
    net->gemm_part_cap = need;
  }
  k_gemm<<<dim3(static_cast<unsigned>((N + kBN - 1) / kBN), static_cast<unsigned>((M + kBM - 1) / kBM),
                static_cast<unsigned>(splits)), 256, 0, c->stream>>>(g, net->gemm_part, kchunk);
  k_gemm_reduce<<<nb(static_cast<long>(M) * N), 256, 0, c->stream>>>(g, net->gemm_part, splits);
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

int backward(spx_gnn* g, long G, long L, float* loss) {
  spx_ctx* c = g->ctx;
  const long R = G * L;
  const int P = g->P;
  k_mse<<<nb(G * P), 256, 0, c->stream>>>(g->y, g->tgt, g->dy, g->dybf, g->sq, G * P,
                                          1.f / static_cast<float>(G * P));
  float* gr = g->grad;
  // readout
  SPX_TRY(gemm(g, g->dybf, 1, P, g->poolbf, H, 1, gr + g->off_w[NLAYER + 1], H, 1, P, H, static_cast<int>(G),
               nullptr, nullptr, 0, 0));
  k_colsum<<<nb(P, 64), 1024, 0, c->stream>>>(g->dy, gr + g->off_b[NLAYER + 1], G, P);
  SPX_TRY(gemm(g, g->dybf, P, 1, g->pbf + g->off_w[NLAYER + 1], H, 1, g->dpool, H, 1, static_cast<int>(G), H, P,
               nullptr, nullptr, 0, 0));
  k_pool_bwd<<<dim3(static_cast<unsigned>(G), static_cast<unsigned>(L)), H, 0, c->stream>>>(g->dpool, g->dx,
                                                                                           static_cast<int>(L));
  for (int l = NLAYER; l >= 0; --l) {
    k_relu_bwd<<<nb(R * H), 256, 0, c->stream>>>(g->dx, g->z[l], g->dz, g->dzbf, R * H);
    k_colsum<<<nb(H, 64), 1024, 0, c->stream>>>(g->dz, gr + g->off_b[l], R, H);
    if (l == 0) {
      SPX_TRY(gemm(g, g->dzbf, 1, H, g->feat, F, 1, gr + g->off_w[0], F, 1, H, F, static_cast<int>(R), nullptr,
                   nullptr, 0, 0));
      break;
    }
    SPX_TRY(gemm(g, g->dzbf, 1, H, g->cat[l], 2 * H, 1, gr + g->off_w[l], 2 * H, 1, H, 2 * H, static_cast<int>(R),
                 nullptr, nullptr, 0, 0));
    // d[x | m] = dz W_l; dx_{l-1} = first half + J^T (second half), J symmetric
    SPX_TRY(gemm(g, g->dzbf, H, 1, g->pbf + g->off_w[l], 2 * H, 1, g->dcat, 2 * H, 1, static_cast<int>(R), 2 * H,
                 H, nullptr, nullptr, 0, 0));
    k_mix_bwd<<<dim3(static_cast<unsigned>(G), static_cast<unsigned>(L)), H, 0, c->stream>>>(
        g->J, g->dcat, g->dx, static_cast<int>(L));
  }
  SPX_HIP(hipGetLastError());
  if (loss) {
    std::vector<float> sq(G * P);
    SPX_TRY(copy_to_host(c, sq.data(), g->sq, sq.size() * sizeof(float)));
    double s = 0.0;
    for (float v : sq) s += v;
    *loss = static_cast<float>(s / static_cast<double>(G * P));
  }
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
      hipMalloc(&g->pbf, off * sizeof(uint16_t)) != hipSuccess)
    rc = fail(SPX_ENOMEM, "hipMalloc of network parameters failed");
  else if (copy_to_device(g->ctx, g->p, p.data(), off * sizeof(float)) != SPX_OK ||
           hipMemset(g->m, 0, off * sizeof(float)) != hipSuccess ||
           hipMemset(g->v, 0, off * sizeof(float)) != hipSuccess)
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
                  static_cast<void*>(g->gemm_part), static_cast<void*>(g->d_step), static_cast<void*>(g->d_bt)})
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
  SPX_TRY(forward(g, static_cast<long>(G), static_cast<long>(L)));
  SPX_TRY(backward(g, static_cast<long>(G), static_cast<long>(L), loss));
  if (grad) SPX_TRY(copy_to_host(g->ctx, grad, g->grad, g->nparam * sizeof(float)));
  return SPX_OK;
}

int spx_gnn_train_step(spx_gnn* g, size_t G, size_t L, const double* J, const double* h, const float* target,
                       float lr, float* loss) {
  SPX_TRY(spx_gnn_gradient(g, G, L, J, h, target, loss, nullptr));
  ++g->step;
  const float b1t = static_cast<float>(std::pow(0.9, static_cast<double>(g->step)));
  const float b2t = static_cast<float>(std::pow(0.999, static_cast<double>(g->step)));
  k_adam<<<nb(static_cast<long>(g->nparam)), 256, 0, g->ctx->stream>>>(g->p, g->grad, g->m, g->v,
                                                                       static_cast<long>(g->nparam), lr, b1t, b2t);
  SPX_HIP(hipGetLastError());
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
    k_gather_step<<<static_cast<unsigned>(G), 128, 0, c->stream>>>(
        g->set_J, g->set_h, g->set_t, g->set_idx, g->d_step, static_cast<int>(G), static_cast<int>(L), g->P, g->Jin,
        g->hin, g->tgt);
    SPX_TRY(forward(g, G, L));
    SPX_TRY(backward(g, G, L, nullptr));
    k_loss_acc<<<1, 256, 0, c->stream>>>(g->sq, G * g->P, g->loss_acc);
    k_adam_step<<<nb(static_cast<long>(g->nparam)), 256, 0, c->stream>>>(
        g->p, g->grad, g->m, g->v, static_cast<long>(g->nparam), lr, g->d_bt, g->d_step);
    k_step_next<<<1, 64, 0, c->stream>>>(g->d_step);
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
