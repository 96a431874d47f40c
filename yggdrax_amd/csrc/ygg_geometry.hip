// Per-node geometry (compute_tree_geometry, _geometry_impl.py:176-338).
//
// Leaf boxes: one wave64 per leaf reduces the leaf's particles (contiguous in
// Morton order) with cross-lane min/max and writes them into the bottom row
// of a power-of-two segment tree. Two multi-level LDS reductions build the
// tree; every internal node then takes the min/max over its leaf range
// [first, last] in O(log L) segment-tree nodes. min/max are exact, so this
// gives the reference's level-synchronous result bit for bit on well-formed
// trees (every internal node's particles are exactly its leaf range);
// malformed trees (duplicate leaf codes) take the level-synchronous path.
#include "ygg_common.h"

namespace ygg {

__device__ __forceinline__ void finalize_node(int64_t k, const float* mn, const float* mx,
                                              float* center, float* half, float* radius,
                                              float* maxext) {
  float h[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float a = mn[d], b = mx[d];
    center[3 * k + d] = 0.5f * (a + b);
    float hh = 0.5f * (b - a);
    h[d] = hh < 1e-6f ? 1e-6f : hh;
    half[3 * k + d] = h[d];
  }
  float s = h[0] * h[0];
  s = s + h[1] * h[1];
  s = s + h[2] * h[2];
  radius[k] = sqrtf(s);  // correctly rounded (built with -fhip-fp32-correctly-rounded-divide-sqrt)
  maxext[k] = fmaxf(fmaxf(h[0], h[1]), h[2]);
}

struct __attribute__((aligned(16))) Box8 {
  float mn[4];
  float mx[4];
};

__device__ __forceinline__ Box8 empty_box() {
  Box8 b;
  for (int d = 0; d < 4; ++d) {
    b.mn[d] = INFINITY;
    b.mx[d] = -INFINITY;
  }
  return b;
}

__device__ __forceinline__ void merge_into(Box8& a, const Box8& b) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a.mn[d] = fminf(a.mn[d], b.mn[d]);
    a.mx[d] = fmaxf(a.mx[d], b.mx[d]);
  }
}

// One wave per segment-tree leaf slot j < P: real leaves reduce their
// particles, padding slots get the empty box.
__global__ __launch_bounds__(256) void k_seg_leaves(
    const float* __restrict__ pos, const float* __restrict__ prad,
    const int32_t* __restrict__ ranges, int T, int NI, int leaf_cap, int P,
    Box8* __restrict__ seg, float* __restrict__ center, float* __restrict__ half,
    float* __restrict__ radius, float* __restrict__ maxext) {
  const int L = T - NI;
  int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE;
  const unsigned lane = lane_id();
  if (j >= P) return;
  if (j >= L) {
    if (lane == 0) seg[P + j] = empty_box();
    return;
  }
  const int64_t node = NI + j;
  const int a = ranges[2 * node];
  int cnt = min(ranges[2 * node + 1] - a + 1, leaf_cap);
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = lane; m < cnt; m += YGG_WAVE) {
    const float* p = pos + 3 * ((int64_t)a + m);
    // a particle with a radius is bounded as the cube [p - r, p + r]
    const float r = prad ? prad[(int64_t)a + m] : 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float v = p[d];
      lo[d] = fminf(lo[d], prad ? v - r : v);
      hi[d] = fmaxf(hi[d], prad ? v + r : v);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off, 64));
    }
  }
  if (lane != 0) return;
  Box8 b;
  for (int d = 0; d < 3; ++d) {
    b.mn[d] = lo[d];
    b.mx[d] = hi[d];
  }
  b.mn[3] = 0.f;
  b.mx[3] = 0.f;
  seg[P + j] = b;
  finalize_node(node, lo, hi, center, half, radius, maxext);
}

// Reduce up to 9 segment-tree levels: children [S, 2S) -> parents down to
// level S / 512 (or the root). Block b owns children [S + 512 b, S + 512 b + 512).
__global__ __launch_bounds__(256) void k_seg_up(Box8* __restrict__ seg, int S) {
  __shared__ Box8 sb[2][256];
  const int tid = threadIdx.x;
  const int64_t c0 = (int64_t)S + (int64_t)blockIdx.x * 512 + 2 * tid;
  int n = S / 2 - blockIdx.x * 256;  // parents this block owns at the first level
  if (n > 256) n = 256;
  Box8 b = empty_box();
  if (tid < n) {
    b = seg[c0];
    merge_into(b, seg[c0 + 1]);
    seg[S / 2 + blockIdx.x * 256 + tid] = b;
  }
  sb[0][tid] = b;
  __syncthreads();
  int level = S / 4, width = n / 2, cur = 0;
  int64_t base = (int64_t)blockIdx.x * 128;
  while (width >= 1 && level >= 1) {
    if (tid < width) {  // ping-pong buffers: reads of one level never race its writes
      Box8 x = sb[cur][2 * tid];
      merge_into(x, sb[cur][2 * tid + 1]);
      seg[level + base + tid] = x;
      sb[1 - cur][tid] = x;
    }
    __syncthreads();
    cur = 1 - cur;
    level >>= 1;
    width >>= 1;
    base >>= 1;
  }
}

__global__ __launch_bounds__(256) void k_seg_query(const Box8* __restrict__ seg, int P,
                                                   const int32_t* __restrict__ ranges, int NI,
                                                   int leaf, float* __restrict__ center,
                                                   float* __restrict__ half,
                                                   float* __restrict__ radius,
                                                   float* __restrict__ maxext) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NI) return;
  int l = ranges[2 * i] / leaf + P;
  int r = ranges[2 * i + 1] / leaf + P + 1;
  Box8 acc = empty_box();
  while (l < r) {
    if (l & 1) merge_into(acc, seg[l++]);
    if (r & 1) merge_into(acc, seg[--r]);
    l >>= 1;
    r >>= 1;
  }
  finalize_node(i, acc.mn, acc.mx, center, half, radius, maxext);
}

// Batched: one wave per segment-tree leaf slot j < P over the batch's
// concatenated leaves (system of j by the leaf offsets).
// LANES lanes per leaf: 32 when leaves hold at most 32 points, so a
// wavefront reduces two leaves at once (every shuffle stays within its
// half), else 64.
template <int LANES>
__global__ __launch_bounds__(256) void k_b_seg_leaves(Seg sg, const float* __restrict__ pos,
                                                      const int32_t* __restrict__ ranges,
                                                      int64_t Ltot, int leaf_cap, int64_t P,
                                                      Box8* __restrict__ seg,
                                                      float* __restrict__ center,
                                                      float* __restrict__ half,
                                                      float* __restrict__ radius,
                                                      float* __restrict__ maxext) {
  const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LANES;
  const unsigned lane = lane_id() % LANES;
  if (j >= P) return;
  if (j >= Ltot) {
    if (lane == 0) seg[P + j] = empty_box();
    return;
  }
  const int s = sg.of_leaf(j);
  const int64_t node = sg.node_off(s) + (sg.L(s) - 1) + (j - sg.leaf_off(s));
  const int64_t a = sg.pt_off(s) + ranges[2 * node];
  const int cnt = min(ranges[2 * node + 1] - ranges[2 * node] + 1, leaf_cap);
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = lane; m < cnt; m += LANES) {
    const float* p = pos + 3 * (a + m);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float v = p[d];
      lo[d] = fminf(lo[d], v);
      hi[d] = fmaxf(hi[d], v);
    }
  }
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fminf(lo[d], __shfl_xor(lo[d], off, 64));
      hi[d] = fmaxf(hi[d], __shfl_xor(hi[d], off, 64));
    }
  }
  if (lane != 0) return;
  Box8 b;
  for (int d = 0; d < 3; ++d) {
    b.mn[d] = lo[d];
    b.mx[d] = hi[d];
  }
  b.mn[3] = 0.f;
  b.mx[3] = 0.f;
  seg[P + j] = b;
  finalize_node(node, lo, hi, center, half, radius, maxext);
}

__global__ __launch_bounds__(256) void k_b_seg_query(Seg sg, const Box8* __restrict__ seg,
                                                     int64_t P, const int32_t* __restrict__ ranges,
                                                     int64_t NItot, float* __restrict__ center,
                                                     float* __restrict__ half,
                                                     float* __restrict__ radius,
                                                     float* __restrict__ maxext) {
  const int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (gi >= NItot) return;
  const int s = sg.of_internal(gi);
  const int64_t node = sg.node_off(s) + (gi - sg.ni_off(s));
  const int64_t l0 = sg.leaf_off(s) + P;
  int64_t l = l0 + ranges[2 * node] / sg.leaf_size;
  int64_t r = l0 + ranges[2 * node + 1] / sg.leaf_size + 1;
  Box8 acc = empty_box();
  while (l < r) {
    if (l & 1) merge_into(acc, seg[l++]);
    if (r & 1) merge_into(acc, seg[--r]);
    l >>= 1;
    r >>= 1;
  }
  finalize_node(node, acc.mn, acc.mx, center, half, radius, maxext);
}

// geometry_to_level_major (_geometry_impl.py:342-433): element (l, j) of the
// padded [num_levels][max_nodes] views holds node nodes_by_level[off[l] + j]
// for j < count(l), else zeros (node index -1).
template <typename IdxT>
__global__ void k_level_major(const float* __restrict__ center, const float* __restrict__ half,
                              const float* __restrict__ radius, const float* __restrict__ maxext,
                              const int32_t* __restrict__ nbl, const int32_t* __restrict__ off,
                              int num_levels, int64_t max_nodes, float* __restrict__ oc,
                              float* __restrict__ oh, float* __restrict__ orad,
                              float* __restrict__ omax, IdxT* __restrict__ oidx) {
  const int64_t total = (int64_t)num_levels * max_nodes;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int l = (int)(e / max_nodes);
    const int64_t j = e - (int64_t)l * max_nodes;
    const int64_t cnt = (int64_t)off[l + 1] - off[l];
    if (j < cnt) {
      const int32_t node = nbl[off[l] + j];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        oc[3 * e + d] = center[3 * (int64_t)node + d];
        oh[3 * e + d] = half[3 * (int64_t)node + d];
      }
      orad[e] = radius[node];
      omax[e] = maxext[node];
      oidx[e] = (IdxT)node;
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        oc[3 * e + d] = 0.0f;
        oh[3 * e + d] = 0.0f;
      }
      orad[e] = 0.0f;
      omax[e] = 0.0f;
      oidx[e] = (IdxT)-1;
    }
  }
}

// Level-synchronous upward pass for malformed trees (_geometry_impl.py:302-329):
// nodes at `level` read their children's current boxes (phase 0 into tmp),
// then write them (phase 1), exactly as the reference's masked update.
__global__ void k_geom_level(const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                             const int32_t* __restrict__ depth, int NI, int level, int phase,
                             float* __restrict__ bmin, float* __restrict__ bmax,
                             float* __restrict__ tmin, float* __restrict__ tmax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NI || depth[i] != level) return;
  if (phase == 0) {
    int32_t l = left[i], r = right[i];
    for (int d = 0; d < 3; ++d) {
      tmin[3 * i + d] = fminf(bmin[3 * (int64_t)l + d], bmin[3 * (int64_t)r + d]);
      tmax[3 * i + d] = fmaxf(bmax[3 * (int64_t)l + d], bmax[3 * (int64_t)r + d]);
    }
  } else {
    for (int d = 0; d < 3; ++d) {
      bmin[3 * i + d] = tmin[3 * i + d];
      bmax[3 * i + d] = tmax[3 * i + d];
    }
  }
}

__global__ void k_geom_leaf_only(const float* __restrict__ pos, const float* __restrict__ prad,
                                 const int32_t* __restrict__ ranges, int T, int NI, int leaf_cap,
                                 float* __restrict__ bmin, float* __restrict__ bmax) {
  int64_t node = NI + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= T) return;
  int a = ranges[2 * node];
  int cnt = min(ranges[2 * node + 1] - a + 1, leaf_cap);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = 0; m < cnt; ++m) {
    const float r = prad ? prad[(int64_t)a + m] : 0.0f;
    for (int d = 0; d < 3; ++d) {
      float v = pos[3 * ((int64_t)a + m) + d];
      lo[d] = fminf(lo[d], prad ? v - r : v);
      hi[d] = fmaxf(hi[d], prad ? v + r : v);
    }
  }
  for (int d = 0; d < 3; ++d) {
    bmin[3 * node + d] = lo[d];
    bmax[3 * node + d] = hi[d];
  }
}

__global__ void k_geom_zero_internal(int NI, float* bmin, float* bmax) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * NI) return;
  bmin[i] = 0.0f;
  bmax[i] = 0.0f;
}

__global__ void k_geom_finalize(int T, const float* __restrict__ bmin, const float* __restrict__ bmax,
                                float* center, float* half, float* radius, float* maxext) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  float lo[3] = {bmin[3 * k], bmin[3 * k + 1], bmin[3 * k + 2]};
  float hi[3] = {bmax[3 * k], bmax[3 * k + 1], bmax[3 * k + 2]};
  finalize_node(k, lo, hi, center, half, radius, maxext);
}

}  // namespace ygg

using namespace ygg;

extern "C" {

static int64_t seg_width(int64_t num_nodes) {
  int64_t L = num_nodes - (num_nodes - 1) / 2, P = 1;
  while (P < L) P <<= 1;
  return P;
}

size_t ygg_geometry_workspace_bytes(int64_t num_nodes) {
  Arena a(nullptr, 0);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<Box8>((size_t)(2 * seg_width(num_nodes)));
  return a.off;
}

int ygg_geometry(const float* pos_sorted, const float* particle_radius,
                 const int32_t* node_ranges, const int32_t* left,
                 const int32_t* right, const int32_t* parent, const int32_t* node_level,
                 int64_t num_nodes, int64_t leaf_size, int64_t leaf_cap, int level_sync,
                 int32_t max_level,
                 float* center, float* half_extent, float* radius, float* max_extent, void* ws,
                 size_t ws_bytes, void* stream) {
  if (num_nodes < 1 || leaf_cap < 1) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int T = (int)num_nodes;
  const int NI = (T - 1) / 2;
  const int L = T - NI;
  Arena a(ws, ws_bytes);
  float* bmin = a.take<float>((size_t)T * 3);
  float* bmax = a.take<float>((size_t)T * 3);
  float* tmin = a.take<float>((size_t)T * 3);
  float* tmax = a.take<float>((size_t)T * 3);
  Box8* seg = a.take<Box8>((size_t)(2 * seg_width(T)));
  if (!bmin || !bmax || !tmin || !tmax || !seg) return YGG_ERR_WORKSPACE;
  if (leaf_size < 1) level_sync = 1;  // leaf ranges are not a fixed-count partition
  const int lc = (int)(leaf_cap > 0x7FFFFFFF ? 0x7FFFFFFF : leaf_cap);
  if (!level_sync) {
    const int P = (int)seg_width(T);
    int64_t threads = (int64_t)P * YGG_WAVE;
    hipLaunchKernelGGL(k_seg_leaves, dim3(grid_for(threads, 256)), dim3(256), 0, st, pos_sorted,
                       particle_radius, node_ranges, T, NI, lc, P, seg, center, half_extent,
                       radius, max_extent);
    for (int S = P; S > 1; S /= 512) {
      int blocks = (S / 2 + 255) / 256;
      hipLaunchKernelGGL(k_seg_up, dim3(blocks), dim3(256), 0, st, seg, S);
    }
    if (NI > 0)
      hipLaunchKernelGGL(k_seg_query, dim3(grid_for(NI, 256)), dim3(256), 0, st, seg, P,
                         node_ranges, NI, (int)leaf_size, center, half_extent, radius,
                         max_extent);
    YGG_LAUNCH_CHECK();
    return YGG_OK;
  }
  hipLaunchKernelGGL(k_geom_zero_internal, dim3(grid_for(3 * (int64_t)NI + 1, 256)), dim3(256), 0,
                     st, NI, bmin, bmax);
  hipLaunchKernelGGL(k_geom_leaf_only, dim3(grid_for(L, 256)), dim3(256), 0, st, pos_sorted,
                     particle_radius, node_ranges, T, NI, lc, bmin, bmax);
  for (int level = max_level - 1; level >= 0; --level) {
    for (int phase = 0; phase < 2; ++phase)
      hipLaunchKernelGGL(k_geom_level, dim3(grid_for(NI + 1, 256)), dim3(256), 0, st, left, right,
                         node_level, NI, level, phase, bmin, bmax, tmin, tmax);
  }
  hipLaunchKernelGGL(k_geom_finalize, dim3(grid_for(T, 256)), dim3(256), 0, st, T, bmin, bmax,
                     center, half_extent, radius, max_extent);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_geometry_level_major(const float* center, const float* half_extent, const float* radius,
                             const float* max_extent, const int32_t* nodes_by_level,
                             const int32_t* level_offsets, int32_t num_levels, int64_t max_nodes,
                             float* out_center, float* out_half_extent, float* out_radius,
                             float* out_max_extent, void* out_node_indices, int out_i64,
                             void* stream) {
  if (num_levels < 1 || max_nodes < 1 || !nodes_by_level || !level_offsets) return YGG_ERR_ARG;
  const int64_t total = (int64_t)num_levels * max_nodes;
  dim3 g(grid_for(total, 256, 16384));
  hipStream_t st = (hipStream_t)stream;
  if (out_i64)
    hipLaunchKernelGGL(k_level_major<int64_t>, g, dim3(256), 0, st, center, half_extent, radius,
                       max_extent, nodes_by_level, level_offsets, num_levels, max_nodes,
                       out_center, out_half_extent, out_radius, out_max_extent,
                       (int64_t*)out_node_indices);
  else
    hipLaunchKernelGGL(k_level_major<int32_t>, g, dim3(256), 0, st, center, half_extent, radius,
                       max_extent, nodes_by_level, level_offsets, num_levels, max_nodes,
                       out_center, out_half_extent, out_radius, out_max_extent,
                       (int32_t*)out_node_indices);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

static int64_t pow2_at_least(int64_t v) {
  int64_t P = 1;
  while (P < v) P <<= 1;
  return P;
}

size_t ygg_batch_geometry_workspace_bytes(const YggBatch* b) {
  if (!b || b->num_leaves < 1) return 0;
  Arena a(nullptr, 0);
  a.take<Box8>((size_t)(2 * pow2_at_least(b->num_leaves)));
  return a.off;
}

int ygg_batch_geometry(const YggBatch* b, const float* pos_sorted, const int32_t* node_ranges,
                       int64_t leaf_cap, float* center, float* half_extent, float* radius,
                       float* max_extent, void* ws, size_t ws_bytes, void* stream) {
  if (!b || b->num_systems < 1 || b->num_leaves < 1 || leaf_cap < 1 || b->leaf_size < 1)
    return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  Seg sg;
  sg.S = b->num_systems;
  sg.leaf_size = b->leaf_size;
  sg.pt = b->point_offsets;
  sg.leaf = b->leaf_offsets;
  const int64_t Ltot = b->num_leaves, NItot = Ltot - b->num_systems;
  const int64_t P = pow2_at_least(Ltot);
  if (P > (1ll << 30)) return YGG_ERR_UNSUPPORTED;
  Arena a(ws, ws_bytes);
  Box8* seg = a.take<Box8>((size_t)(2 * P));
  if (!seg) return YGG_ERR_WORKSPACE;
  const int lc = (int)(leaf_cap > 0x7FFFFFFF ? 0x7FFFFFFF : leaf_cap);
  if (lc <= 32)
    hipLaunchKernelGGL(k_b_seg_leaves<32>, dim3(grid_for(P * 32, 256)), dim3(256), 0, st, sg,
                       pos_sorted, node_ranges, Ltot, lc, P, seg, center, half_extent, radius,
                       max_extent);
  else
    hipLaunchKernelGGL(k_b_seg_leaves<64>, dim3(grid_for(P * YGG_WAVE, 256)), dim3(256), 0, st,
                       sg, pos_sorted, node_ranges, Ltot, lc, P, seg, center, half_extent, radius,
                       max_extent);
  for (int64_t S = P; S > 1; S /= 512) {
    int blocks = (int)((S / 2 + 255) / 256);
    hipLaunchKernelGGL(k_seg_up, dim3(blocks), dim3(256), 0, st, seg, (int)S);
  }
  if (NItot > 0)
    hipLaunchKernelGGL(k_b_seg_query, dim3(grid_for(NItot, 256)), dim3(256), 0, st, sg, seg, P,
                       node_ranges, NItot, center, half_extent, radius, max_extent);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
