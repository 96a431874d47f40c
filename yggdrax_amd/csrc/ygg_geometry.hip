// Per-node geometry (compute_tree_geometry, _geometry_impl.py:176-338).
//
// Leaf boxes: one wave64 per leaf reduces the leaf's particles (contiguous in
// Morton order) with cross-lane min/max. Internal boxes: Karras-style
// bottom-up merge; the second child to arrive at a parent (device-scope
// acq_rel counter) merges both children, so every node is finalised exactly
// once without level-by-level launches. min/max are exact, so this gives the
// reference's level-synchronous result bit for bit on well-formed trees;
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

__global__ __launch_bounds__(256) void k_geom_bottom_up(
    const float* __restrict__ pos, const int32_t* __restrict__ ranges,
    const int32_t* __restrict__ left, const int32_t* __restrict__ right,
    const int32_t* __restrict__ parent, int T, int NI, int leaf_cap, float* __restrict__ bmin,
    float* __restrict__ bmax, uint32_t* __restrict__ visit, float* __restrict__ center,
    float* __restrict__ half, float* __restrict__ radius, float* __restrict__ maxext) {
  const int L = T - NI;
  int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE;
  const unsigned lane = lane_id();
  if (wave >= L) return;
  const int64_t node = NI + wave;
  const int a = ranges[2 * node];
  int cnt = ranges[2 * node + 1] - a + 1;
  cnt = min(cnt, leaf_cap);
  float lo[3] = {INFINITY, INFINITY, INFINITY};
  float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = lane; m < cnt; m += YGG_WAVE) {
    const float* p = pos + 3 * ((int64_t)a + m);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float v = p[d];
      lo[d] = fminf(lo[d], v);
      hi[d] = fmaxf(hi[d], v);
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
  for (int d = 0; d < 3; ++d) {
    bmin[3 * node + d] = lo[d];
    bmax[3 * node + d] = hi[d];
  }
  finalize_node(node, lo, hi, center, half, radius, maxext);
  int64_t cur = node;
  for (int guard = 0; guard < 4096; ++guard) {  // depth bound; malformed trees go level-sync
    int32_t p = parent[cur];
    if (p < 0) break;
    uint32_t old = __hip_atomic_fetch_add(&visit[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) break;
    int32_t l = left[p], r = right[p];
    float nlo[3], nhi[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      nlo[d] = fminf(bmin[3 * (int64_t)l + d], bmin[3 * (int64_t)r + d]);
      nhi[d] = fmaxf(bmax[3 * (int64_t)l + d], bmax[3 * (int64_t)r + d]);
    }
    for (int d = 0; d < 3; ++d) {
      bmin[3 * (int64_t)p + d] = nlo[d];
      bmax[3 * (int64_t)p + d] = nhi[d];
    }
    finalize_node(p, nlo, nhi, center, half, radius, maxext);
    cur = p;
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

__global__ void k_geom_leaf_only(const float* __restrict__ pos, const int32_t* __restrict__ ranges,
                                 int T, int NI, int leaf_cap, float* __restrict__ bmin,
                                 float* __restrict__ bmax) {
  int64_t node = NI + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= T) return;
  int a = ranges[2 * node];
  int cnt = min(ranges[2 * node + 1] - a + 1, leaf_cap);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int m = 0; m < cnt; ++m)
    for (int d = 0; d < 3; ++d) {
      float v = pos[3 * ((int64_t)a + m) + d];
      lo[d] = fminf(lo[d], v);
      hi[d] = fmaxf(hi[d], v);
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

size_t ygg_geometry_workspace_bytes(int64_t num_nodes) {
  Arena a(nullptr, 0);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<float>((size_t)num_nodes * 3);
  a.take<uint32_t>((size_t)num_nodes);
  return a.off;
}

int ygg_geometry(const float* pos_sorted, const int32_t* node_ranges, const int32_t* left,
                 const int32_t* right, const int32_t* parent, const int32_t* node_level,
                 int64_t num_nodes, int64_t leaf_cap, int level_sync, int32_t max_level,
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
  uint32_t* visit = a.take<uint32_t>((size_t)T);
  if (!bmin || !bmax || !tmin || !tmax || !visit) return YGG_ERR_WORKSPACE;
  const int lc = (int)(leaf_cap > 0x7FFFFFFF ? 0x7FFFFFFF : leaf_cap);
  if (!level_sync) {
    int rc = fill_u32(visit, 0u, T, st);
    if (rc) return rc;
    int64_t threads = (int64_t)L * YGG_WAVE;
    hipLaunchKernelGGL(k_geom_bottom_up, dim3(grid_for(threads, 256)), dim3(256), 0, st,
                       pos_sorted, node_ranges, left, right, parent, T, NI, lc, bmin, bmax, visit,
                       center, half_extent, radius, max_extent);
    YGG_LAUNCH_CHECK();
    return YGG_OK;
  }
  hipLaunchKernelGGL(k_geom_zero_internal, dim3(grid_for(3 * (int64_t)NI + 1, 256)), dim3(256), 0,
                     st, NI, bmin, bmax);
  hipLaunchKernelGGL(k_geom_leaf_only, dim3(grid_for(L, 256)), dim3(256), 0, st, pos_sorted,
                     node_ranges, T, NI, lc, bmin, bmax);
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

}  // extern "C"
