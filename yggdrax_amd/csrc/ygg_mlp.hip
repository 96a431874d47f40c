// Capacity model: an 8 -> 64 -> 64 -> 64 -> 3 ReLU MLP over per-node
// features, with bfloat16 weights and float32 accumulation, trained and run
// by these kernels. Output j is the log2 of a Poisson rate: the expected
// far-list count, near-list count (leaves) and number of pairs in the peak
// wavefront with this node as target, each divided by a per-output scale.
// Trained on the Poisson negative log-likelihood, the rates are unbiased, so
// their sums estimate the walk's stream sizes in one pass. A misprediction
// only costs a stream growth during the walk (or, for the pair queue, a
// rerun without the prediction): lists never depend on the model.
//
// Parameter layout (P = 9091 floats): W1[64][8] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[3][64] b4[3]; W[o][i] multiplies input i into output o.
#include "ygg_common.h"

// The model's arithmetic is not matched against the reference, so products
// and sums may fuse here.
#pragma clang fp contract(fast)

namespace ygg {

constexpr int kIn = 8, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH,
              kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kOut * kH, kParams = kOffB4 + kOut;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

// Round-to-nearest-even float32 -> bfloat16 (NaN kept quiet).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// ---- inference on the matrix cores --------------------------------------
// Each wavefront runs 16 samples at a time through the four layers with
// 16x16x16 bf16 MFMAs (fp32 accumulate). Operand layouts of
// v_mfma_f32_16x16x16_bf16 on wave64 (lane l, j = 0..3):
//   A[m][k]: lane l holds A[l % 16][4 (l / 16) + j]
//   B[k][n]: lane l holds B[4 (l / 16) + j][l % 16]
//   D[m][n]: lane l holds D[4 (l / 16) + j][l % 16]
// With A = a 16-output slice of W (m = output, k = input) and B = the
// activations (k = input feature, n = sample), a layer's D tile t is, lane
// for lane, the B operand of k-step t of the next layer: no data moves
// between layers. The weights are exactly bf16; each fp32 activation is
// split into bf16 hi + lo (split_bf16) and both halves are multiplied, so
// the products keep ~2^-16 relative accuracy and the outputs match fp32
// arithmetic to ~1e-5. Weight fragments live in registers for the whole
// kernel.
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
#ifndef MLP_INFER_BLOCKS
#define MLP_INFER_BLOCKS 256
#endif
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// hi = v truncated to bf16, so v - hi is exact in fp32; lo = that remainder
// rounded to nearest even. |v - hi - lo| <= 2^-9 |v - hi| < 2^-16 |v|.
__device__ __forceinline__ void split_bf16(const f32x4_t& v, bf16x4_t& hi, bf16x4_t& lo) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t u = __float_as_uint(v[j]);
    hi[j] = (short)(u >> 16);
    lo[j] = (short)f32_to_bf16(v[j] - __uint_as_float(u & 0xFFFF0000u));
  }
}

__device__ __forceinline__ f32x4_t relu4(f32x4_t v) {
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = fmaxf(v[j], 0.0f);
  return v;
}

__global__ __launch_bounds__(256) void k_mlp_infer(const uint16_t* __restrict__ wbf,
                                                   const float* __restrict__ x, int64_t n,
                                                   float* __restrict__ y) {
  const unsigned l = lane_id();
  const int r = (int)(l % 16), q = (int)(l / 16);
  auto w = [&](int off) { return (short)wbf[off]; };
  auto wf = [&](int off) { return bf16_to_f32(wbf[off]); };
  bf16x4_t a1[4], a2[4][4], a3[4][4], a4[4];
  f32x4_t b1[4], b2[4], b3[4], b4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      a1[t][j] = k < kIn ? w(kOffW1 + (16 * t + r) * kIn + k) : (short)0;
      b1[t][j] = wf(kOffB1 + 16 * t + k);
      b2[t][j] = wf(kOffB2 + 16 * t + k);
      b3[t][j] = wf(kOffB3 + 16 * t + k);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        a2[t][kk][j] = w(kOffW2 + (16 * t + r) * kH + 16 * kk + k);
        a3[t][kk][j] = w(kOffW3 + (16 * t + r) * kH + 16 * kk + k);
      }
    }
  }
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int j = 0; j < 4; ++j) a4[kk][j] = r < kOut ? w(kOffW4 + r * kH + 16 * kk + 4 * q + j) : (short)0;
#pragma unroll
  for (int j = 0; j < 4; ++j) b4[j] = 4 * q + j < kOut ? wf(kOffB4 + 4 * q + j) : 0.0f;

  // Wave-uniform loop over 16-sample tiles (every lane reaches every
  // MFMA). At one wavefront per SIMD nothing else hides a load, so the next
  // tile's features are loaded while this tile runs.
  const int64_t tiles = (n + 15) / 16;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / YGG_WAVE);
  auto load_tile = [&](int64_t tl) {
    f32x4_t v;
    const int64_t sm = tl * 16 + r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      v[j] = (sm < n && k < kIn) ? x[sm * kIn + k] : 0.0f;
    }
    return v;
  };
  int64_t tile = (int64_t)blockIdx.x * (blockDim.x / YGG_WAVE) + threadIdx.x / YGG_WAVE;
  f32x4_t in_next = load_tile(tile);
  for (; tile < tiles; tile += nw) {
    const int64_t smp = tile * 16 + r;
    const bool ok = smp < n;
    const f32x4_t in = in_next;
    if (tile + nw < tiles) in_next = load_tile(tile + nw);
    // Every activation tile is split once and feeds the four output tiles;
    // k-steps run outermost, so consecutive MFMAs write different
    // accumulators and pipeline.
    bf16x4_t hi[4], lo[4];
    f32x4_t acc[4];
    split_bf16(in, hi[0], lo[0]);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[t], hi[0], b1[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[t], lo[0], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) split_bf16(relu4(acc[kk]), hi[kk], lo[kk]);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = b2[t];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[t][kk], hi[kk], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[t][kk], lo[kk], acc[t], 0, 0, 0);
      }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) split_bf16(relu4(acc[kk]), hi[kk], lo[kk]);
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = b3[t];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[t][kk], hi[kk], acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3[t][kk], lo[kk], acc[t], 0, 0, 0);
      }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) split_bf16(relu4(acc[kk]), hi[kk], lo[kk]);
    f32x4_t out = b4;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4[kk], hi[kk], out, 0, 0, 0);
      out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a4[kk], lo[kk], out, 0, 0, 0);
    }
    if (q == 0 && ok) {
#pragma unroll
      for (int j = 0; j < kOut; ++j) y[smp * kOut + j] = out[j];
    }
  }
}

// ---- training: layer kernels over a batch (row-major activations) -------
// On the matrix cores with v_mfma_f32_16x16x4_f32 (fp32 operands, fp32
// accumulation; the sums run in a different order than a scalar loop, the
// products are the same). Operand layouts on wave64 (lane l):
//   A[m][k]: lane l holds A[l % 16][l / 16]
//   B[k][n]: lane l holds B[l / 16][l % 16]
//   D[m][n]: lane l holds D[4 (l / 16) + j][l % 16], j = 0..3
// (the D layout of the 16x16x16 bf16 form above). Operands past a matrix's
// edge are zero, so any size works with whole 16 x 16 tiles.
__device__ __forceinline__ f32x4_t mfma4(float a, float b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// y[b][o] = act(sum_i W[o][i] x[b][i] + bias[o]): one wavefront per 16
// batch rows, every 16-output tile (m = row, n = output, k = input).
__global__ __launch_bounds__(256) void k_lin_fwd(const uint16_t* __restrict__ W,
                                                 const uint16_t* __restrict__ bias,
                                                 const float* __restrict__ x, int B, int nin,
                                                 int nout, int relu, float* __restrict__ y) {
  const int lane = (int)lane_id(), r = lane % 16, q = lane / 16;
  const int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE * 16;
  if (b0 >= B) return;
  for (int n0 = 0; n0 < nout; n0 += 16) {
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < nin; k0 += 4) {
      const int64_t b = b0 + r;
      const int k = k0 + q, o = n0 + r;
      const float av = (b < B && k < nin) ? x[b * nin + k] : 0.0f;
      const float bv = (o < nout && k < nin) ? bf16_to_f32(W[o * nin + k]) : 0.0f;
      acc = mfma4(av, bv, acc);
    }
    const int o = n0 + r;
    if (o < nout) {  // every lane reaches the next tile's MFMAs
      const float bo = bf16_to_f32(bias[o]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t b = b0 + 4 * q + j;
        if (b < B) {
          const float v = acc[j] + bo;
          y[b * nout + o] = relu ? fmaxf(v, 0.0f) : v;
        }
      }
    }
  }
}

// dx[b][i] = sum_o W[o][i] dy[b][o], times relu'(x) when the input came from
// a ReLU (m = row, n = input, k = output).
__global__ __launch_bounds__(256) void k_lin_bwd_data(const uint16_t* __restrict__ W,
                                                      const float* __restrict__ dy,
                                                      const float* __restrict__ xin, int B,
                                                      int nin, int nout, int relu_in,
                                                      float* __restrict__ dx) {
  const int lane = (int)lane_id(), r = lane % 16, q = lane / 16;
  const int64_t b0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE * 16;
  if (b0 >= B) return;
  for (int n0 = 0; n0 < nin; n0 += 16) {
    f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int k0 = 0; k0 < nout; k0 += 4) {
      const int64_t b = b0 + r;
      const int o = k0 + q, i = n0 + r;
      const float av = (b < B && o < nout) ? dy[b * nout + o] : 0.0f;
      const float bv = (o < nout && i < nin) ? bf16_to_f32(W[o * nin + i]) : 0.0f;
      acc = mfma4(av, bv, acc);
    }
    const int i = n0 + r;
    if (i < nin) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t b = b0 + 4 * q + j;
        if (b < B) {
          const int64_t e = b * nin + i;
          dx[e] = (relu_in && !(xin[e] > 0.0f)) ? 0.0f : acc[j];
        }
      }
    }
  }
}

// Weight gradients, split over the batch: slice s of kGradSlice rows adds
// into partial[s][m][n] = sum_b dy[b][m] xa[b][n] with xa = [x | 1] (the
// column n = nin gives the bias gradient); m = output, n = input, k = row.
// One wavefront per (16 x 16 tile, slice); Mp, Np are nout and nin + 1
// rounded up to 16. The slices are summed in a fixed order afterwards, so
// the gradients are deterministic.
constexpr int kGradSlice = 256;

__global__ __launch_bounds__(256) void k_lin_bwd_weight(const float* __restrict__ dy,
                                                        const float* __restrict__ x, int B,
                                                        int nin, int nout, int Mp, int Np,
                                                        float* __restrict__ partial) {
  const int lane = (int)lane_id(), r = lane % 16, q = lane / 16;
  const int wave = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE);
  const int tiles_n = Np / 16, tiles = (Mp / 16) * tiles_n;
  const int slices = (B + kGradSlice - 1) / kGradSlice;
  if (wave >= tiles * slices) return;
  const int s = wave / tiles, t = wave % tiles;
  const int m0 = (t / tiles_n) * 16, n0 = (t % tiles_n) * 16;
  const int kb = s * kGradSlice, ke = min(B, kb + kGradSlice);
  f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int b0 = kb; b0 < ke; b0 += 4) {
    const int b = b0 + q;
    const int o = m0 + r, i = n0 + r;
    const float av = (b < ke && o < nout) ? dy[(int64_t)b * nout + o] : 0.0f;
    const float bv = b < ke ? (i < nin ? x[(int64_t)b * nin + i] : (i == nin ? 1.0f : 0.0f))
                            : 0.0f;
    acc = mfma4(av, bv, acc);
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
    partial[((int64_t)s * Mp + m0 + 4 * q + j) * Np + n0 + r] = acc[j];
}

// dW[o][i] and db[o] from the slices' partial sums, in slice order.
__global__ void k_lin_grad_reduce(const float* __restrict__ partial, int slices, int Mp, int Np,
                                  int nin, int nout, float* __restrict__ dW,
                                  float* __restrict__ db) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nout * (nin + 1)) return;
  const int o = e / (nin + 1), i = e % (nin + 1);
  float acc = 0.0f;
  for (int s = 0; s < slices; ++s) acc += partial[((int64_t)s * Mp + o) * Np + i];
  if (i < nin)
    dW[o * nin + i] = acc;
  else
    db[o] = acc;
}

// Masked mean squared error: dy = 2 (y - t) mask / B, loss accumulated.
// Masked Poisson NLL with rate 2^y (y clamped at 30): per element
// m (2^y - t y ln 2) / B, gradient m (2^y - t) ln 2 / B.
constexpr float kLn2 = 0.69314718055994531f;
constexpr float kMaxLogRate = 30.0f;

__global__ void k_poisson(const float* __restrict__ y, const float* __restrict__ t,
                          const float* __restrict__ mask, int B, float* __restrict__ dy,
                          float* __restrict__ loss) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (e < B * kOut) {
    float m = mask[e];
    float yc = fminf(y[e], kMaxLogRate);
    float lam = exp2f(yc);
    dy[e] = y[e] > kMaxLogRate ? 0.0f : m * (lam - t[e]) * kLn2 / (float)B;
    l = m * (lam - t[e] * yc * kLn2) / (float)B;
  }
  l = wave_sum_f(l);
  if (lane_id() == 0 && l != 0.0f) atomicAdd(loss, l);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ pbf, int n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  float np = p[i] - upd;
  p[i] = np;
  pbf[i] = f32_to_bf16(np);
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ q, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = f32_to_bf16(p[i]);
}

__global__ void k_zero_words(uint32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// ---- per-node features ------------------------------------------------------
// Features of node k of one tree whose arrays start at the given bases
// (root = node 0 of the tree; parent/children/sibling ids local).
__device__ __forceinline__ void node_features(int k, int NI, const int32_t* __restrict__ node_level,
                                              const int32_t* __restrict__ ranges,
                                              const int32_t* __restrict__ parent,
                                              const int32_t* __restrict__ left,
                                              const int32_t* __restrict__ right,
                                              const float* __restrict__ half,
                                              const float* __restrict__ radius,
                                              const float* __restrict__ maxext,
                                              const float* __restrict__ parent_far,
                                              float* __restrict__ o) {
  const float root_ext = maxext[0], root_rad = radius[0];
  const float root_vol = 8.0f * half[0] * half[1] * half[2];
  float cnt = (float)(ranges[2 * k + 1] - ranges[2 * k] + 1);
  int p = parent[k];
  float sib = 0.0f;
  if (p >= 0) {
    int sb = left[p] == k ? right[p] : left[p];
    sib = (float)(ranges[2 * sb + 1] - ranges[2 * sb] + 1);
  }
  float vol = 8.0f * half[3 * k] * half[3 * k + 1] * half[3 * k + 2];
  float pf = (parent_far && p >= 0) ? parent_far[p] : 0.0f;
  o[0] = (float)node_level[k] / 64.0f;
  o[1] = log2f(cnt) / 24.0f;
  o[2] = log2f(maxext[k] / root_ext) / 32.0f + 1.0f;
  o[3] = log2f(radius[k] / root_rad) / 32.0f + 1.0f;
  o[4] = log2f(1.0f + fmaxf(pf, 0.0f)) / 16.0f;
  o[5] = log2f(1.0f + sib) / 24.0f;
  float rel = (cnt / fmaxf(vol, 1e-30f)) / ((float)(ranges[1] + 1) / fmaxf(root_vol, 1e-30f));
  o[6] = fminf(fmaxf(log2f(fmaxf(rel, 1e-30f)) / 64.0f, -1.0f), 1.0f);
  o[7] = k >= NI ? 1.0f : 0.0f;
}

__global__ void k_mlp_features(int T, int NI, const int32_t* __restrict__ node_level,
                               const int32_t* __restrict__ ranges,
                               const int32_t* __restrict__ parent,
                               const int32_t* __restrict__ left,
                               const int32_t* __restrict__ right,
                               const float* __restrict__ half, const float* __restrict__ radius,
                               const float* __restrict__ maxext,
                               const float* __restrict__ parent_far, float* __restrict__ x) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  node_features(k, NI, node_level, ranges, parent, left, right, half, radius, maxext, parent_far,
                x + (int64_t)k * kIn);
}

// Batched: every node of every system, with its own tree's root.
__global__ void k_b_mlp_features(Seg sg, int64_t Ttot, const int32_t* __restrict__ node_level,
                                 const int32_t* __restrict__ ranges,
                                 const int32_t* __restrict__ parent,
                                 const int32_t* __restrict__ left,
                                 const int32_t* __restrict__ right,
                                 const float* __restrict__ half, const float* __restrict__ radius,
                                 const float* __restrict__ maxext,
                                 const float* __restrict__ parent_far, float* __restrict__ x) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Ttot) return;
  const int s = sg.of_node(g);
  const int64_t n0 = sg.node_off(s), ni0 = sg.ni_off(s);
  node_features((int)(g - n0), sg.L(s) - 1, node_level + n0, ranges + 2 * n0, parent + n0,
                left + ni0, right + ni0, half + 3 * n0, radius + n0, maxext + n0,
                parent_far ? parent_far + n0 : nullptr, x + g * kIn);
}

// Batched totals: output 1 (near) counts each system's leaves.
// Adds a workgroup's three partial sums into tot[0..2]: wavefront
// shuffles, then one LDS row per wavefront and three atomics per workgroup
// (same-address device atomics serialise across XCDs, so the grids of the
// totals kernels are kept at kTotalsBlocks workgroups).
constexpr int kTotalsBlocks = 128;

__device__ __forceinline__ void block_add3(double a0, double a1, double a2, double* tot) {
  __shared__ double red[256 / YGG_WAVE][3];
  for (int off = 32; off > 0; off >>= 1) {
    a0 += __shfl_xor(a0, off, 64);
    a1 += __shfl_xor(a1, off, 64);
    a2 += __shfl_xor(a2, off, 64);
  }
  const int w = threadIdx.x / YGG_WAVE;
  if (lane_id() == 0) {
    red[w][0] = a0;
    red[w][1] = a1;
    red[w][2] = a2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double v = 0;
    for (int q = 0; q < (int)(blockDim.x / YGG_WAVE); ++q) v += red[q][threadIdx.x];
    atomicAdd(&tot[threadIdx.x], v);
  }
}

__global__ __launch_bounds__(256) void k_b_mlp_totals(Seg sg, const float* __restrict__ y,
                                                      int64_t n, float s0, float s1, float s2,
                                                      double* __restrict__ tot) {
  __shared__ int32_t s_noff[kSegLdsCap];
  const bool staged = seg_stage(s_noff, sg.S, [&](int s) { return (int32_t)sg.node_off(s); });
  __syncthreads();
  double a0 = 0, a1 = 0, a2 = 0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
       g += (int64_t)gridDim.x * blockDim.x) {
    const int s = staged ? seg_find_lds(s_noff, sg.S, g) : sg.of_node(g);
    a0 += (double)(s0 * exp2f(fminf(y[3 * g], kMaxLogRate)));
    if (g - sg.node_off(s) >= sg.L(s) - 1)
      a1 += (double)(s1 * exp2f(fminf(y[3 * g + 1], kMaxLogRate)));
    a2 += (double)(s2 * exp2f(fminf(y[3 * g + 2], kMaxLogRate)));
  }
  block_add3(a0, a1, a2, tot);
}

// Sum of scale_j 2^y_j per output over [0, n) (stream totals), y row-major
// [n][3]; output 1 (near) counts leaves only.
__global__ __launch_bounds__(256) void k_mlp_totals(const float* __restrict__ y, int64_t n,
                                                    int64_t NI, float s0, float s1, float s2,
                                                    double* __restrict__ tot) {
  double a0 = 0, a1 = 0, a2 = 0;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < n;
       s += (int64_t)gridDim.x * blockDim.x) {
    a0 += (double)(s0 * exp2f(fminf(y[3 * s], kMaxLogRate)));
    if (s >= NI) a1 += (double)(s1 * exp2f(fminf(y[3 * s + 1], kMaxLogRate)));
    a2 += (double)(s2 * exp2f(fminf(y[3 * s + 2], kMaxLogRate)));
  }
  block_add3(a0, a1, a2, tot);
}

// far[k] = s0 2^y[k][0] (the first pass's far-count prediction).
__global__ void k_mlp_far_of(const float* __restrict__ y, int T, float s0, float* __restrict__ pf) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < T) pf[k] = s0 * exp2f(fminf(y[3 * k], kMaxLogRate));
}

}  // namespace ygg

using namespace ygg;

extern "C" {

int ygg_mlp_num_params(void) { return kParams; }

int ygg_mlp_forward(const uint16_t* params_bf16, const float* features, int64_t n,
                    float* outputs, void* stream) {
  if (n < 0 || !params_bf16) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  // 16-sample tiles, four wavefronts per workgroup, at most 256 workgroups:
  // one wavefront per SIMD (the kernel's occupancy), each loading the weight
  // fragments once and then looping over many tiles
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid_for((n + 15) / 16, 256 / YGG_WAVE, MLP_INFER_BLOCKS)),
                     dim3(256), 0, (hipStream_t)stream, params_bf16, features, n, outputs);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

// One training step on a batch: forward, masked MSE, backward, Adam.
// ws: float workspace of ygg_mlp_train_workspace_floats(B) entries.
int64_t ygg_mlp_train_workspace_floats(int64_t B) {
  const int64_t slices = (B + kGradSlice - 1) / kGradSlice;
  return (int64_t)B * (3 * kH + kOut) * 2 + kParams + 16 + slices * kH * (kH + 16);
}

int ygg_mlp_train_step(float* params, uint16_t* params_bf16, float* adam_m, float* adam_v,
                       int64_t step, float lr, const float* x, const float* target,
                       const float* mask, int64_t B64, float* grads, float* ws,
                       float* loss_out, void* stream) {
  if (B64 < 1 || !params || !grads || !ws) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int B = (int)B64;
  float* a1 = ws;
  float* a2 = a1 + (int64_t)B * kH;
  float* a3 = a2 + (int64_t)B * kH;
  float* yo = a3 + (int64_t)B * kH;
  float* d1 = yo + (int64_t)B * kOut;
  float* d2 = d1 + (int64_t)B * kH;
  float* d3 = d2 + (int64_t)B * kH;
  float* dyo = d3 + (int64_t)B * kH;
  float* partial = dyo + (int64_t)B * kOut;  // weight-gradient slices
  const uint16_t* W = params_bf16;
  auto g = [](int64_t n) { return dim3(grid_for(n, 256)); };
  // one wavefront per 16 batch rows
  const dim3 gr(grid_for(((int64_t)B + 15) / 16 * YGG_WAVE, 256));
  hipLaunchKernelGGL(k_lin_fwd, gr, dim3(256), 0, st, W + kOffW1, W + kOffB1, x, B, kIn, kH, 1,
                     a1);
  hipLaunchKernelGGL(k_lin_fwd, gr, dim3(256), 0, st, W + kOffW2, W + kOffB2, a1, B, kH, kH, 1,
                     a2);
  hipLaunchKernelGGL(k_lin_fwd, gr, dim3(256), 0, st, W + kOffW3, W + kOffB3, a2, B, kH, kH, 1,
                     a3);
  hipLaunchKernelGGL(k_lin_fwd, gr, dim3(256), 0, st, W + kOffW4, W + kOffB4, a3, B, kH, kOut, 0,
                     yo);
  if (loss_out) hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)loss_out, 1);
  hipLaunchKernelGGL(k_poisson, g((int64_t)B * kOut), dim3(256), 0, st, yo, target, mask, B,
                     dyo, loss_out);
  const int slices = (B + kGradSlice - 1) / kGradSlice;
  auto weight_grads = [&](const float* dy, const float* xin, int nin, int nout, float* dW,
                          float* db) {
    const int Mp = (nout + 15) / 16 * 16, Np = (nin + 1 + 15) / 16 * 16;
    const int64_t waves = (int64_t)(Mp / 16) * (Np / 16) * slices;
    hipLaunchKernelGGL(k_lin_bwd_weight, dim3(grid_for(waves * YGG_WAVE, 256)), dim3(256), 0, st,
                       dy, xin, B, nin, nout, Mp, Np, partial);
    hipLaunchKernelGGL(k_lin_grad_reduce, g((int64_t)nout * (nin + 1)), dim3(256), 0, st,
                       partial, slices, Mp, Np, nin, nout, dW, db);
  };
  // layer 4
  weight_grads(dyo, a3, kH, kOut, grads + kOffW4, grads + kOffB4);
  hipLaunchKernelGGL(k_lin_bwd_data, gr, dim3(256), 0, st, W + kOffW4, dyo, a3, B, kH, kOut, 1,
                     d3);
  // layer 3
  weight_grads(d3, a2, kH, kH, grads + kOffW3, grads + kOffB3);
  hipLaunchKernelGGL(k_lin_bwd_data, gr, dim3(256), 0, st, W + kOffW3, d3, a2, B, kH, kH, 1, d2);
  // layer 2
  weight_grads(d2, a1, kH, kH, grads + kOffW2, grads + kOffB2);
  hipLaunchKernelGGL(k_lin_bwd_data, gr, dim3(256), 0, st, W + kOffW2, d2, a1, B, kH, kH, 1, d1);
  // layer 1
  weight_grads(d1, x, kIn, kH, grads + kOffW1, grads + kOffB1);
  YGG_LAUNCH_CHECK();
  if (lr > 0.0f) {
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, g(kParams), dim3(256), 0, st, params, grads, adam_m, adam_v,
                       params_bf16, kParams, lr, b1, b2, 1e-8f, c1, c2);
    YGG_LAUNCH_CHECK();
  }
  (void)yo;
  return YGG_OK;
}

int ygg_mlp_params_to_bf16(const float* params, uint16_t* params_bf16, void* stream) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(kParams, 256)), dim3(256), 0, (hipStream_t)stream,
                     params, params_bf16, kParams);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_mlp_features(int64_t num_nodes, const int32_t* node_level, const int32_t* node_ranges,
                     const int32_t* parent, const int32_t* left, const int32_t* right,
                     const float* half_extent, const float* radius, const float* max_extent,
                     const float* parent_far, float* features, void* stream) {
  if (num_nodes < 1) return YGG_ERR_ARG;
  int T = (int)num_nodes, NI = (T - 1) / 2;
  hipLaunchKernelGGL(k_mlp_features, dim3(grid_for(T, 256)), dim3(256), 0, (hipStream_t)stream, T,
                     NI, node_level, node_ranges, parent, left, right, half_extent, radius,
                     max_extent, parent_far, features);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_mlp_totals(const float* outputs, int64_t num_nodes, float scale_far, float scale_near,
                   float scale_queue, double* totals, void* stream) {
  if (num_nodes < 1 || !outputs || !totals) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)totals, 6);
  hipLaunchKernelGGL(k_mlp_totals, dim3(grid_for(num_nodes, 256, kTotalsBlocks)), dim3(256), 0, st,
                     outputs, num_nodes, (num_nodes - 1) / 2, scale_far, scale_near, scale_queue,
                     totals);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_mlp_far_counts(const float* outputs, int64_t num_nodes, float scale_far,
                       float* far_counts, void* stream) {
  if (num_nodes < 1 || !outputs || !far_counts) return YGG_ERR_ARG;
  hipLaunchKernelGGL(k_mlp_far_of, dim3(grid_for(num_nodes, 256)), dim3(256), 0,
                     (hipStream_t)stream, outputs, (int)num_nodes, scale_far, far_counts);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

static inline Seg mlp_seg(const YggBatch* b) {
  Seg sg;
  sg.S = b->num_systems;
  sg.leaf_size = b->leaf_size;
  sg.pt = b->point_offsets;
  sg.leaf = b->leaf_offsets;
  return sg;
}

int ygg_batch_mlp_features(const YggBatch* b, const int32_t* node_level,
                           const int32_t* node_ranges, const int32_t* parent, const int32_t* left,
                           const int32_t* right, const float* half_extent, const float* radius,
                           const float* max_extent, const float* parent_far, float* features,
                           void* stream) {
  if (!b || b->num_systems < 1 || !features) return YGG_ERR_ARG;
  const int64_t T = 2 * b->num_leaves - b->num_systems;
  hipLaunchKernelGGL(k_b_mlp_features, dim3(grid_for(T, 256)), dim3(256), 0, (hipStream_t)stream,
                     mlp_seg(b), T, node_level, node_ranges, parent, left, right, half_extent,
                     radius, max_extent, parent_far, features);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_batch_mlp_totals(const YggBatch* b, const float* outputs, float scale_far,
                         float scale_near, float scale_queue, double* totals, void* stream) {
  if (!b || b->num_systems < 1 || !outputs || !totals) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int64_t T = 2 * b->num_leaves - b->num_systems;
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, st, (uint32_t*)totals, 6);
  hipLaunchKernelGGL(k_b_mlp_totals, dim3(grid_for(T, 256, kTotalsBlocks)), dim3(256), 0, st, mlp_seg(b),
                     outputs, T, scale_far, scale_near, scale_queue, totals);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
