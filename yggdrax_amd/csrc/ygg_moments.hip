// Per-node mass moments (yggdrax/tree_moments.py:686-746): the total mass
// and the centre of mass of every node's particle range in the
// Morton-sorted order.
//
// The reference takes float prefix sums (jnp.cumsum) of m and m * x and
// differences them per node. Here every node sum is accumulated in float64
// instead: fixed chunks of `chunk` particles are summed by groups of lanes,
// the chunk sums form a segment tree, and each node adds its full chunks
// from the tree and its partial chunks particle by particle. The result is
// rounded once to the output dtype. The per-particle terms are the
// reference's: m * x rounded to the product dtype (float32 for float32
// positions and masses), and the centre is moment / mass divided in the
// output dtype, 0 where the rounded mass is 0.
#include <type_traits>

#include "ygg_common.h"

namespace ygg {
namespace {

struct Acc4 {
  double m, x, y, z;
};

__device__ __forceinline__ void add_to(Acc4& a, const Acc4& b) {
  a.m += b.m;
  a.x += b.x;
  a.y += b.y;
  a.z += b.z;
}

template <typename P, typename M>
struct Term {
  // product dtype of masses[:, None] * positions
  using W = typename std::conditional<sizeof(P) == 8 || sizeof(M) == 8, double, float>::type;
  __device__ __forceinline__ static Acc4 at(const P* __restrict__ pos,
                                            const M* __restrict__ mass, int64_t p) {
    const W m = (W)mass[p];
    Acc4 t;
    t.m = (double)mass[p];
    t.x = (double)(W)(m * (W)pos[3 * p + 0]);
    t.y = (double)(W)(m * (W)pos[3 * p + 1]);
    t.z = (double)(W)(m * (W)pos[3 * p + 2]);
    return t;
  }
};

// Chunk sums into the segment tree's bottom row seg[P + c], zeros for the
// padding chunks c >= C. A group of g = 2^g_log2 lanes (g <= 8) takes one
// chunk, each lane summing every g-th particle of it, so a wavefront reads
// 64 / g chunks and reduces with log2(g) shuffle steps. The loop bound is
// uniform per block so every lane reaches the shuffles.
template <typename P, typename M>
__global__ void __launch_bounds__(256)
k_chunk_sums(const P* __restrict__ pos, const M* __restrict__ mass, int64_t n, int64_t chunk,
             int64_t C, int64_t Pw, int g_log2, Acc4* __restrict__ seg) {
  const int g = 1 << g_log2;
  const int64_t total = Pw << g_log2;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t gid = base + threadIdx.x;
    const int64_t c = gid >> g_log2;
    const int lane = (int)(gid & (g - 1));
    Acc4 acc = {0.0, 0.0, 0.0, 0.0};
    if (c < C) {
      const int64_t s = c * chunk;
      const int64_t e = min(s + chunk, n);
      for (int64_t p = s + lane; p < e; p += g) add_to(acc, Term<P, M>::at(pos, mass, p));
    }
    for (int o = g >> 1; o > 0; o >>= 1) {
      acc.m += __shfl_xor(acc.m, o, g);
      acc.x += __shfl_xor(acc.x, o, g);
      acc.y += __shfl_xor(acc.y, o, g);
      acc.z += __shfl_xor(acc.z, o, g);
    }
    if (lane == 0 && gid < total) seg[Pw + c] = acc;
  }
}

// One level of the segment tree: seg[i] = seg[2i] + seg[2i+1], i in [h, 2h).
__global__ void k_seg_level(Acc4* __restrict__ seg, int64_t h) {
  for (int64_t i = h + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 2 * h;
       i += (int64_t)gridDim.x * blockDim.x) {
    Acc4 a = seg[2 * i];
    add_to(a, seg[2 * i + 1]);
    seg[i] = a;
  }
}

// The top levels h, h/2, ..., 1 (h <= kTopLevel) by one block, a barrier
// between levels.
constexpr int64_t kTopLevel = 1024;

__global__ void __launch_bounds__(1024) k_seg_top(Acc4* __restrict__ seg, int64_t h) {
  for (int64_t hh = h; hh >= 1; hh >>= 1) {
    for (int64_t i = hh + threadIdx.x; i < 2 * hh; i += blockDim.x) {
      Acc4 a = seg[2 * i];
      add_to(a, seg[2 * i + 1]);
      seg[i] = a;
    }
    __syncthreads();
  }
}

template <typename P, typename M>
__device__ Acc4 range_sum(const P* __restrict__ pos, const M* __restrict__ mass, int64_t n,
                          int64_t chunk, int64_t Pw, const Acc4* __restrict__ seg, int64_t s,
                          int64_t e) {
  Acc4 acc = {0.0, 0.0, 0.0, 0.0};
  if (s < 0) s = 0;
  if (e > n - 1) e = n - 1;
  if (e < s) return acc;
  int64_t a = s / chunk, b = e / chunk;
  const bool full_left = s == a * chunk;
  const bool full_right = e == n - 1 || e == (b + 1) * chunk - 1;
  if (a == b && !(full_left && full_right)) {
    for (int64_t p = s; p <= e; ++p) add_to(acc, Term<P, M>::at(pos, mass, p));
    return acc;
  }
  if (!full_left) {
    for (int64_t p = s; p < (a + 1) * chunk; ++p) add_to(acc, Term<P, M>::at(pos, mass, p));
    ++a;
  }
  if (!full_right) {
    for (int64_t p = b * chunk; p <= e; ++p) add_to(acc, Term<P, M>::at(pos, mass, p));
    --b;
  }
  // full chunks [a, b] from the tree
  int64_t l = a + Pw, r = b + Pw + 1;
  while (l < r) {
    if (l & 1) add_to(acc, seg[l++]);
    if (r & 1) add_to(acc, seg[--r]);
    l >>= 1;
    r >>= 1;
  }
  return acc;
}

template <typename P, typename M>
__global__ void k_node_moments(const P* __restrict__ pos, const M* __restrict__ mass, int64_t n,
                               int64_t chunk, int64_t Pw, const Acc4* __restrict__ seg,
                               const int32_t* __restrict__ ranges, int64_t T,
                               M* __restrict__ mass_out,
                               typename Term<P, M>::W* __restrict__ center_out) {
  using W = typename Term<P, M>::W;
  for (int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; node < T;
       node += (int64_t)gridDim.x * blockDim.x) {
    const Acc4 acc = range_sum(pos, mass, n, chunk, Pw, seg, (int64_t)ranges[2 * node],
                               (int64_t)ranges[2 * node + 1]);
    const M mt = (M)acc.m;
    mass_out[node] = mt;
    const W mw = (W)mt;
    center_out[3 * node + 0] = mt == (M)0 ? (W)0 : (W)acc.x / mw;
    center_out[3 * node + 1] = mt == (M)0 ? (W)0 : (W)acc.y / mw;
    center_out[3 * node + 2] = mt == (M)0 ? (W)0 : (W)acc.z / mw;
  }
}

int64_t pow2_ceil(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

template <typename P, typename M>
int mass_moments(const P* pos, const M* mass, int64_t n, const int32_t* ranges, int64_t T,
                 int64_t chunk, M* mass_out, typename Term<P, M>::W* center_out, void* ws,
                 size_t ws_bytes, hipStream_t st) {
  const int64_t C = (n + chunk - 1) / chunk;
  const int64_t Pw = pow2_ceil(C);
  Arena a(ws, ws_bytes);
  Acc4* seg = a.take<Acc4>((size_t)(2 * Pw));
  if (!seg) return YGG_ERR_WORKSPACE;
  int g_log2 = 0;
  while ((1ll << g_log2) < chunk && g_log2 < 3) ++g_log2;
  hipLaunchKernelGGL((k_chunk_sums<P, M>), dim3(grid_for(Pw << g_log2, 256)), dim3(256), 0, st,
                     pos, mass, n, chunk, C, Pw, g_log2, seg);
  YGG_LAUNCH_CHECK();
  int64_t h = Pw >> 1;
  for (; h > kTopLevel; h >>= 1) {
    hipLaunchKernelGGL(k_seg_level, dim3(grid_for(h, 256)), dim3(256), 0, st, seg, h);
    YGG_LAUNCH_CHECK();
  }
  if (h >= 1) {
    hipLaunchKernelGGL(k_seg_top, dim3(1), dim3(1024), 0, st, seg, h);
    YGG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_node_moments<P, M>), dim3(grid_for(T, 256)), dim3(256), 0, st, pos, mass,
                     n, chunk, Pw, seg, ranges, T, mass_out, center_out);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // namespace
}  // namespace ygg

using namespace ygg;

extern "C" {

size_t ygg_tree_mass_moments_workspace_bytes(int64_t n, int64_t chunk) {
  if (n < 1 || chunk < 1) return 0;
  Arena a(nullptr, 0);
  a.take<Acc4>((size_t)(2 * pow2_ceil((n + chunk - 1) / chunk)));
  return a.off;
}

int ygg_tree_mass_moments(const void* pos_sorted, int pos_f64, const void* mass_sorted,
                          int mass_f64, int64_t n, const int32_t* node_ranges, int64_t num_nodes,
                          int64_t chunk, void* mass_out, void* center_out, void* ws,
                          size_t ws_bytes, void* stream) {
  if (n < 1 || num_nodes < 1 || chunk < 1 || !pos_sorted || !mass_sorted || !node_ranges ||
      !mass_out || !center_out)
    return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (!pos_f64 && !mass_f64)
    return mass_moments((const float*)pos_sorted, (const float*)mass_sorted, n, node_ranges,
                        num_nodes, chunk, (float*)mass_out, (float*)center_out, ws, ws_bytes,
                        st);
  if (pos_f64 && mass_f64)
    return mass_moments((const double*)pos_sorted, (const double*)mass_sorted, n, node_ranges,
                        num_nodes, chunk, (double*)mass_out, (double*)center_out, ws, ws_bytes,
                        st);
  if (pos_f64)
    return mass_moments((const double*)pos_sorted, (const float*)mass_sorted, n, node_ranges,
                        num_nodes, chunk, (float*)mass_out, (double*)center_out, ws, ws_bytes,
                        st);
  return mass_moments((const float*)pos_sorted, (const double*)mass_sorted, n, node_ranges,
                      num_nodes, chunk, (double*)mass_out, (double*)center_out, ws, ws_bytes,
                      st);
}

}  // extern "C"
