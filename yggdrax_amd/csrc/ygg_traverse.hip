// Dual-tree MAC traversal (_dual_tree_walk_impl, _interactions_impl.py:1131-1905)
// and output shaping (_result_to_interactions / _result_to_neighbors,
// 3561-3837).
//
// Generation-synchronous walk: generation g's pair wavefront is processed by
// one launch; each wave64 evaluates 64 pairs, emits far/near entries through
// per-row atomic slot counters and appends the refined child pairs to the
// next wavefront with one atomic per wave. Capacity semantics follow the
// reference exactly: a pair-queue push at slot >= queue_cap, or a list entry
// at slot >= its row capacity, is dropped and raises the matching overflow
// flag, and the walk stops after the generation in which a flag was raised.
// Per-row counts then equal min(cumulative count, capacity) as in the
// reference. Within-row entry order is unspecified in the reference (it
// comes from an unstable argsort) and is arrival order here.
#include <hip/hip_runtime.h>

#include "ygg_common.h"
#include "ygg_walk.h"

namespace ygg {

__global__ void k_prep_nodes(const float* __restrict__ center, const float* __restrict__ radius,
                             const float* __restrict__ maxext, const int32_t* __restrict__ left,
                             const int32_t* __restrict__ right,
                             const int32_t* __restrict__ leaf_position, int T, int NI,
                             int mac_type, float scale, NodeRec* __restrict__ out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  NodeRec r;
  r.cx = center[3 * (int64_t)k];
  r.cy = center[3 * (int64_t)k + 1];
  r.cz = center[3 * (int64_t)k + 2];
  float e;
  if (mac_type == YGG_MAC_DEHNEN)
    e = scale * radius[k];
  else if (mac_type == YGG_MAC_ENGBLOM)
    e = radius[k];
  else
    e = maxext[k];
  r.ext = e;
  r.left = k < NI ? left[k] : -1;
  r.right = k < NI ? right[k] : -1;
  r.leafpos = k < NI ? -1 : leaf_position[k];
  r.pad = 0;
  out[k] = r;
}

__global__ void k_set_rank(const int32_t* __restrict__ nbl, int T, NodeRec* __restrict__ out) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < T) out[nbl[r]].pad = r;
}

// Batched: records of every system's nodes with batch-global child ids,
// leaf rows and level-major ranks; roots[s] = the system's root node.
__global__ void k_b_prep_nodes(Seg sg, const float* __restrict__ center,
                               const float* __restrict__ radius,
                               const float* __restrict__ maxext,
                               const int32_t* __restrict__ left,
                               const int32_t* __restrict__ right,
                               const int32_t* __restrict__ leaf_position, int64_t Ttot,
                               int mac_type, float scale, NodeRec* __restrict__ out,
                               int32_t* __restrict__ roots) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Ttot) return;
  const int s = sg.of_node(g);
  const int64_t n0 = sg.node_off(s);
  const int NI = sg.L(s) - 1;
  const int k = (int)(g - n0);
  NodeRec r;
  r.cx = center[3 * g];
  r.cy = center[3 * g + 1];
  r.cz = center[3 * g + 2];
  float e;
  if (mac_type == YGG_MAC_DEHNEN)
    e = scale * radius[g];
  else if (mac_type == YGG_MAC_ENGBLOM)
    e = radius[g];
  else
    e = maxext[g];
  r.ext = e;
  const int64_t ni = sg.ni_off(s) + k;
  r.left = k < NI ? (int32_t)(n0 + left[ni]) : -1;
  r.right = k < NI ? (int32_t)(n0 + right[ni]) : -1;
  r.leafpos = k < NI ? -1 : sg.leaf_off(s) + leaf_position[g];
  r.pad = 0;
  out[g] = r;
  if (k == 0) roots[s] = (int32_t)n0;
}

__global__ void k_b_set_rank(Seg sg, const int32_t* __restrict__ nbl, int64_t Ttot,
                             NodeRec* __restrict__ out) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Ttot) return;
  const int64_t n0 = sg.node_off(sg.of_node(g));
  out[n0 + nbl[g]].pad = (int32_t)g;
}

__global__ void k_walk_init(YggWalkParams p) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = i; k < p.num_nodes; k += stride) p.far_counts[k] = 0;
  for (int64_t k = i; k < p.num_leaves; k += stride) p.near_counts[k] = 0;
  for (int64_t k = i; k <= p.max_generations; k += stride) p.gen_sizes[k] = 0;
  if (i < 4) p.flags[i] = 0;
  if (i < 3) p.decisions[i] = 0;
  if (i == 0) {
    p.gen_sizes[0] = 1;
    p.wf0[0] = ((uint64_t)(uint32_t)p.root << 32) | (uint32_t)p.root;
  }
}

template <int MODE>
__device__ __forceinline__ void emit(int32_t* cnt, const int64_t* start, const int32_t* cap,
                                     int32_t fixed_cap, int32_t* entries, int32_t row,
                                     int32_t val, uint32_t* flag) {
  int32_t slot = atomicAdd(&cnt[row], 1);
  if (MODE == YGG_WALK_COUNT) return;
  if (MODE == YGG_WALK_FIXED) {
    if (slot < fixed_cap)
      entries[(int64_t)row * fixed_cap + slot] = val;
    else
      atomicOr(flag, 1u);
  } else {
    if (slot < cap[row])
      entries[start[row] + slot] = val;
    else
      atomicOr(flag, 1u);
  }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_walk_gen(YggWalkParams p, int g) {
  const uint32_t stop = atomic_load_agent(&p.flags[0]) | atomic_load_agent(&p.flags[1]) |
                        atomic_load_agent(&p.flags[2]);
  if (stop) return;
  const int64_t raw = p.gen_sizes[g];
  const int64_t size = raw < p.queue_cap ? raw : p.queue_cap;
  if (size == 0) return;
  const uint64_t* __restrict__ wf_in = (g & 1) ? p.wf1 : p.wf0;
  uint64_t* __restrict__ wf_out = (g & 1) ? p.wf0 : p.wf1;
  const NodeRec* __restrict__ nodes = (const NodeRec*)p.nodes;
  const int NI = p.num_internal;
  const float theta_sq = p.theta_sq;
  const unsigned lane = lane_id();
  unsigned n_acc = 0, n_near = 0, n_ref = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < size;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t idx = base + threadIdx.x;
    int nc = 0;
    uint32_t ca[4], cb[4];
    if (idx < size) {
      const uint64_t pr = wf_in[idx];
      const int32_t t = (int32_t)(pr >> 32), s = (int32_t)(uint32_t)pr;
      const NodeRec nt = nodes[t], ns = nodes[s];
      float dx = nt.cx - ns.cx, dy = nt.cy - ns.cy, dz = nt.cz - ns.cz;
      float d2 = dx * dx;
      d2 = d2 + dy * dy;
      d2 = d2 + dz * dz;
      const float et = nt.ext, es = ns.ext;
      const bool diff = t != s;
      bool ok;
      if (p.mac_type == YGG_MAC_ENGBLOM) {
        float rb = fmaxf(et, es), rs = fminf(et, es);
        float lhs = rb + p.theta_root * rs;
        ok = diff && d2 > 0.0f && (lhs * lhs <= theta_sq * d2);
      } else {
        float sum = et + es;
        ok = diff && d2 > 0.0f && (sum * sum <= theta_sq * d2);
      }
      const bool t_int = t < NI, s_int = s < NI;
      if (ok) {
        ++n_acc;
        if (p.collect_far) {
          emit<MODE>(p.far_counts, p.far_row_start, p.far_row_cap, p.far_cap, p.far_entries, t, s,
                     &p.flags[1]);
          emit<MODE>(p.far_counts, p.far_row_start, p.far_row_cap, p.far_cap, p.far_entries, s, t,
                     &p.flags[1]);
        }
      } else if (!t_int && !s_int && diff) {
        ++n_near;
        if (p.collect_near) {
          emit<MODE>(p.near_counts, p.near_row_start, p.near_row_cap, p.near_cap,
                     p.near_entries, nt.leafpos, s, &p.flags[2]);
          emit<MODE>(p.near_counts, p.near_row_start, p.near_row_cap, p.near_cap,
                     p.near_entries, ns.leafpos, t, &p.flags[2]);
        }
      } else {
        ++n_ref;
        const bool same = !diff;
        const bool split_t = t_int && (same || !s_int || et >= es);
        const bool split_s = s_int && (same || !t_int || es > et);
        if (split_t && split_s) {
          if (same) {
            ca[0] = nt.left;  cb[0] = nt.left;
            ca[1] = nt.left;  cb[1] = nt.right;
            ca[2] = nt.right; cb[2] = nt.right;
            nc = 3;
          } else {
            ca[0] = nt.left;  cb[0] = ns.left;
            ca[1] = nt.left;  cb[1] = ns.right;
            ca[2] = nt.right; cb[2] = ns.left;
            ca[3] = nt.right; cb[3] = ns.right;
            nc = 4;
          }
        } else if (split_t) {
          ca[0] = nt.left;  cb[0] = s;
          ca[1] = nt.right; cb[1] = s;
          nc = 2;
        } else if (split_s) {
          ca[0] = t; cb[0] = ns.left;
          ca[1] = t; cb[1] = ns.right;
          nc = 2;
        }
      }
    }
    // Append child pairs: one atomic per wave, slots in (pair, child) order.
    unsigned incl = wave_inclusive_sum((unsigned)nc);
    unsigned total = __shfl(incl, 63, 64);
    if (total) {
      uint32_t wbase = 0;
      if (lane == 63) wbase = atomicAdd(&p.gen_sizes[g + 1], total);
      wbase = __shfl(wbase, 63, 64);
      int64_t slot = (int64_t)wbase + incl - nc;
      bool over = false;
      for (int k = 0; k < nc; ++k, ++slot) {
        uint32_t a = ca[k], b = cb[k];
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        if (slot < p.queue_cap)
          wf_out[slot] = ((uint64_t)lo << 32) | hi;
        else
          over = true;
      }
      if (over) atomicOr(&p.flags[0], 1u);
    }
  }
  n_acc = wave_sum(n_acc);
  n_near = wave_sum(n_near);
  n_ref = wave_sum(n_ref);
  if (lane == 0) {
    if (n_acc) atomicAdd((unsigned long long*)&p.decisions[0], (unsigned long long)n_acc);
    if (n_near) atomicAdd((unsigned long long*)&p.decisions[1], (unsigned long long)n_near);
    if (n_ref) atomicAdd((unsigned long long*)&p.decisions[2], (unsigned long long)n_ref);
  }
}

// Row start/capacity tables for an exact-length fill whose rows are already
// laid out in output order: far rows at their level-major offsets, near rows
// at their native-order offsets.
__global__ void k_rows_from_counts(const int32_t* __restrict__ nbl,
                                   const int64_t* __restrict__ lm_off, int T,
                                   int64_t* __restrict__ start) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T) return;
  start[nbl ? nbl[r] : r] = lm_off[r];
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_far_expand(
    int T, const int32_t* __restrict__ nbl, const int64_t* __restrict__ lm_off,
    const int32_t* __restrict__ cnt, int32_t fixed_cap, const int64_t* __restrict__ row_start,
    const int32_t* __restrict__ entries, const int32_t* __restrict__ node_level,
    OutT* __restrict__ sources, OutT* __restrict__ targets, OutT* __restrict__ levels,
    OutT* __restrict__ node_offsets) {
  int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE;
  const unsigned lane = lane_id();
  if (r >= T) return;
  const int32_t node = nbl[r];
  const int64_t off = lm_off[r];
  int32_t c = cnt[node];
  if (fixed_cap > 0 && c > fixed_cap) c = fixed_cap;
  const int64_t src = row_start ? row_start[node] : (int64_t)node * fixed_cap;
  if (lane == 0) {
    node_offsets[node] = (OutT)off;
    if (r == 0) node_offsets[T] = (OutT)lm_off[T];
  }
  const OutT lv = (OutT)node_level[node];
  for (int e = lane; e < c; e += YGG_WAVE) {
    sources[off + e] = (OutT)entries[src + e];
    targets[off + e] = (OutT)node;
    levels[off + e] = lv;
  }
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_near_expand(
    int L, const int64_t* __restrict__ off_arr, const int32_t* __restrict__ cnt,
    int32_t fixed_cap, const int64_t* __restrict__ row_start, const int32_t* __restrict__ entries,
    const int32_t* __restrict__ leaf_position, int64_t maxc, OutT* __restrict__ neighbors,
    OutT* __restrict__ nlp) {
  int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / YGG_WAVE;
  const unsigned lane = lane_id();
  if (r >= L) return;
  const int64_t off = off_arr[r];
  int32_t c = cnt[r];
  if (fixed_cap > 0 && c > fixed_cap) c = fixed_cap;
  const int64_t src = row_start ? row_start[r] : r * (int64_t)fixed_cap;
  for (int e = lane; e < c; e += YGG_WAVE) neighbors[off + e] = (OutT)entries[src + e];
  if (nlp) {
    for (int64_t e = lane; e < maxc; e += YGG_WAVE)
      nlp[r * maxc + e] = e < c ? (OutT)leaf_position[entries[src + e]] : (OutT)-1;
  }
}

__global__ void k_min_cap(int32_t* __restrict__ cnt, int64_t n, int32_t cap) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && cnt[k] > cap) cnt[k] = cap;
}

__global__ void k_min_cap_arr(int32_t* __restrict__ cnt, int64_t n, const int32_t* __restrict__ cap) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && cnt[k] > cap[k]) cnt[k] = cap[k];
}

// Near-field target blocks (_build_target_blocks,
// yggdrax/_interactions_impl.py:3689-3770): every native leaf row is cut
// into ceil(count / k) blocks of k consecutive entries of its CSR row.
__global__ void k_target_blocks_per_row(int64_t L, const int64_t* __restrict__ off,
                                        int64_t k, int32_t* __restrict__ blocks) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L) return;
  const int64_t c = off[r + 1] - off[r];
  blocks[r] = (int32_t)((c + k - 1) / k);
}

// One thread per (block, slot): the block's row is the last row whose first
// block is <= b (upper bound over block_offsets[1..L]); slot j holds the
// native leaf row of the row's neighbour local * k + j, or 0 and invalid
// past the row's end.
template <typename OUT, typename IN>
__global__ __launch_bounds__(256) void k_target_blocks(
    int64_t L, const int64_t* __restrict__ off, const IN* __restrict__ nbrs,
    const int32_t* __restrict__ leaf_position, int64_t k,
    const int64_t* __restrict__ boff, int64_t B, OUT* __restrict__ block_row,
    OUT* __restrict__ block_src, uint8_t* __restrict__ valid) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * k) return;
  const int64_t b = i / k, j = i - b * k;
  int64_t lo = 0, hi = L;  // first r in [0, L) with boff[r + 1] > b
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (boff[mid + 1] > b)
      hi = mid;
    else
      lo = mid + 1;
  }
  const int64_t r = lo;
  const int64_t e = off[r] + (b - boff[r]) * k + j;
  const bool ok = e < off[r + 1];
  const int32_t src = ok ? leaf_position[(int64_t)nbrs[e]] : 0;
  const bool v = ok && src >= 0;
  block_src[i] = (OUT)(v ? src : 0);
  valid[i] = v ? 1 : 0;
  if (j == 0) block_row[b] = (OUT)r;
}

}  // namespace ygg

using namespace ygg;

extern "C" {

int ygg_walk_prepare_nodes(const float* center, const float* radius, const float* max_extent,
                           const int32_t* left, const int32_t* right,
                           const int32_t* leaf_position, const int32_t* nodes_by_level,
                           int64_t num_nodes, int mac_type, float dehnen_radius_scale,
                           void* nodes_out, void* stream) {
  if (num_nodes < 1 || !nodes_out) return YGG_ERR_ARG;
  int T = (int)num_nodes, NI = (T - 1) / 2;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_prep_nodes, dim3(grid_for(T, 256)), dim3(256), 0, st, center, radius,
                     max_extent, left, right, leaf_position, T, NI, mac_type, dehnen_radius_scale,
                     (NodeRec*)nodes_out);
  if (nodes_by_level)
    hipLaunchKernelGGL(k_set_rank, dim3(grid_for(T, 256)), dim3(256), 0, st, nodes_by_level, T,
                       (NodeRec*)nodes_out);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_batch_walk_prepare_nodes(const YggBatch* b, const float* center, const float* radius,
                                 const float* max_extent, const int32_t* left,
                                 const int32_t* right, const int32_t* leaf_position,
                                 const int32_t* nodes_by_level, int mac_type,
                                 float dehnen_radius_scale, void* nodes_out, int32_t* roots,
                                 void* stream) {
  if (!b || b->num_systems < 1 || !nodes_out || !roots || !nodes_by_level) return YGG_ERR_ARG;
  Seg sg;
  sg.S = b->num_systems;
  sg.leaf_size = b->leaf_size;
  sg.pt = b->point_offsets;
  sg.leaf = b->leaf_offsets;
  const int64_t T = 2 * b->num_leaves - b->num_systems;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_b_prep_nodes, dim3(grid_for(T, 256)), dim3(256), 0, st, sg, center, radius,
                     max_extent, left, right, leaf_position, T, mac_type, dehnen_radius_scale,
                     (NodeRec*)nodes_out, roots);
  hipLaunchKernelGGL(k_b_set_rank, dim3(grid_for(T, 256)), dim3(256), 0, st, sg, nodes_by_level,
                     T, (NodeRec*)nodes_out);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

static int launch_gen(const YggWalkParams* p, int g, hipStream_t st, int blocks) {
  switch (p->mode) {
    case YGG_WALK_COUNT:
      hipLaunchKernelGGL(k_walk_gen<YGG_WALK_COUNT>, dim3(blocks), dim3(256), 0, st, *p, g);
      break;
    case YGG_WALK_FIXED:
      hipLaunchKernelGGL(k_walk_gen<YGG_WALK_FIXED>, dim3(blocks), dim3(256), 0, st, *p, g);
      break;
    case YGG_WALK_OFFSETS:
      hipLaunchKernelGGL(k_walk_gen<YGG_WALK_OFFSETS>, dim3(blocks), dim3(256), 0, st, *p, g);
      break;
    default:
      return YGG_ERR_ARG;
  }
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

// Event pool for optional per-generation kernel timing (host-side objects).
static hipEvent_t g_events[2 * 1025];
static int g_events_ready = 0;

int ygg_walk_bfs(const YggWalkParams* p, uint32_t* host_status, int32_t* num_generations,
                 float* gen_ms, void* stream) {
  if (!p || !host_status || !num_generations) return YGG_ERR_ARG;
  if (gen_ms && p->max_generations > 1024) return YGG_ERR_ARG;
  if (gen_ms && !g_events_ready) {
    for (int k = 0; k < 2 * 1025; ++k) YGG_CHECK(hipEventCreate(&g_events[k]));
    g_events_ready = 1;
  }
  if (p->num_internal < 1 || p->queue_cap < 1 || p->max_generations < 2) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int64_t init_n = p->num_nodes > p->max_generations + 1 ? p->num_nodes : p->max_generations + 1;
  hipLaunchKernelGGL(k_walk_init, dim3(grid_for(init_n, 256, 4096)), dim3(256), 0, st, *p);
  YGG_LAUNCH_CHECK();
  const int blocks = 4096;
  int g = 0;
  int batch = 24;
  const int G = p->max_generations;
  while (true) {
    int end = g + batch < G ? g + batch : G;
    for (int k = g; k < end; ++k) {
      if (gen_ms) YGG_CHECK(hipEventRecord(g_events[2 * k], st));
      int rc = launch_gen(p, k, st, blocks);
      if (rc) return rc;
      if (gen_ms) YGG_CHECK(hipEventRecord(g_events[2 * k + 1], st));
    }
    YGG_CHECK(hipMemcpyAsync(host_status, p->gen_sizes, sizeof(uint32_t) * (G + 1),
                             hipMemcpyDeviceToHost, st));
    YGG_CHECK(hipMemcpyAsync(host_status + G + 1, p->flags, sizeof(uint32_t) * 4,
                             hipMemcpyDeviceToHost, st));
    YGG_CHECK(hipStreamSynchronize(st));
    const uint32_t* fl = host_status + G + 1;
    bool stopped = fl[0] | fl[1] | fl[2];
    // generations processed so far: first g' <= end with an empty wavefront
    int done = -1;
    for (int k = 0; k <= end; ++k)
      if (host_status[k] == 0) {
        done = k;
        break;
      }
    if (stopped || done >= 0) {
      if (done < 0) {
        // stopped on overflow: the last launched generation with a non-empty
        // wavefront is the one that raised the flag.
        done = end;
        for (int k = 0; k <= end; ++k)
          if (host_status[k] == 0) {
            done = k;
            break;
          }
      }
      *num_generations = done;
      if (gen_ms) {
        for (int k = 0; k < end; ++k) {
          float ms = 0.f;
          YGG_CHECK(hipEventElapsedTime(&ms, g_events[2 * k], g_events[2 * k + 1]));
          gen_ms[k] = ms;
        }
        for (int k = end; k < G; ++k) gen_ms[k] = 0.f;
      }
      return YGG_OK;
    }
    if (end >= G) return YGG_ERR_UNSUPPORTED;  // deeper than max_generations
    g = end;
    batch = 16;
  }
}

int ygg_exclusive_scan_i32(const int32_t* counts, const int32_t* gather, int64_t* out, int64_t n,
                           void* ws, size_t ws_bytes, void* stream) {
  return exclusive_scan_i32_to_i64(counts, gather, out, n, ws, ws_bytes, (hipStream_t)stream);
}

size_t ygg_scan_workspace_bytes(int64_t n) { return scan_workspace_bytes(n); }

int ygg_rows_from_offsets(const int32_t* order, const int64_t* offsets, int64_t n,
                          int64_t* row_start, void* stream) {
  if (n < 1) return YGG_OK;
  hipLaunchKernelGGL(k_rows_from_counts, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, order, offsets, (int)n, row_start);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_cap_counts(int32_t* counts, int64_t n, int32_t fixed_cap, const int32_t* caps,
                   void* stream) {
  if (n < 1) return YGG_OK;
  hipStream_t st = (hipStream_t)stream;
  if (caps)
    hipLaunchKernelGGL(k_min_cap_arr, dim3(grid_for(n, 256)), dim3(256), 0, st, counts, n, caps);
  else
    hipLaunchKernelGGL(k_min_cap, dim3(grid_for(n, 256)), dim3(256), 0, st, counts, n, fixed_cap);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_far_output(int64_t num_nodes, const int32_t* nodes_by_level, const int64_t* lm_offsets,
                   const int32_t* counts, int32_t fixed_cap, const int64_t* row_start,
                   const int32_t* entries, const int32_t* node_level, int out_i64,
                   void* sources, void* targets, void* target_levels, void* node_offsets,
                   void* stream) {
  if (num_nodes < 1) return YGG_ERR_ARG;
  int T = (int)num_nodes;
  dim3 grid(grid_for((int64_t)T * YGG_WAVE, 256));
  hipStream_t st = (hipStream_t)stream;
  if (out_i64)
    hipLaunchKernelGGL(k_far_expand<int64_t>, grid, dim3(256), 0, st, T, nodes_by_level,
                       lm_offsets, counts, fixed_cap, row_start, entries, node_level,
                       (int64_t*)sources, (int64_t*)targets, (int64_t*)target_levels,
                       (int64_t*)node_offsets);
  else
    hipLaunchKernelGGL(k_far_expand<int32_t>, grid, dim3(256), 0, st, T, nodes_by_level,
                       lm_offsets, counts, fixed_cap, row_start, entries, node_level,
                       (int32_t*)sources, (int32_t*)targets, (int32_t*)target_levels,
                       (int32_t*)node_offsets);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_target_block_offsets_workspace_bytes(int64_t num_leaves) {
  return align_up(sizeof(int32_t) * (size_t)(num_leaves + 1)) +
         ygg::scan_workspace_bytes(num_leaves);
}

int ygg_target_block_offsets(int64_t num_leaves, const int64_t* near_offsets, int64_t block_size,
                             int64_t* block_offsets, void* ws, size_t ws_bytes, void* stream) {
  if (num_leaves < 1 || block_size < 1 || !near_offsets || !block_offsets) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  Arena a(ws, ws_bytes);
  int32_t* blocks = a.take<int32_t>((size_t)num_leaves + 1);
  const size_t rest = ygg::scan_workspace_bytes(num_leaves);
  char* sws = a.take<char>(rest);
  if (!blocks || !sws) return YGG_ERR_WORKSPACE;
  hipLaunchKernelGGL(k_target_blocks_per_row, dim3(grid_for(num_leaves, 256)), dim3(256), 0, st,
                     num_leaves, near_offsets, block_size, blocks);
  YGG_LAUNCH_CHECK();
  return ygg::exclusive_scan_i32_to_i64(blocks, nullptr, block_offsets, num_leaves, sws, rest,
                                        st);
}

int ygg_target_blocks(int64_t num_leaves, const int64_t* near_offsets, const void* neighbors,
                      int in_i64, const int32_t* leaf_position, int64_t block_size,
                      const int64_t* block_offsets, int64_t num_blocks, int out_i64,
                      void* block_leaf_ids, void* block_source_leaf_ids, uint8_t* valid_mask,
                      void* stream) {
  if (num_leaves < 1 || block_size < 1 || num_blocks < 0) return YGG_ERR_ARG;
  if (num_blocks == 0) return YGG_OK;
  const int64_t n = num_blocks * block_size;
  dim3 grid(grid_for(n, 256));
  hipStream_t st = (hipStream_t)stream;
#define YGG_TB(OUT, IN)                                                                        \
  hipLaunchKernelGGL((k_target_blocks<OUT, IN>), grid, dim3(256), 0, st, num_leaves,           \
                     near_offsets, (const IN*)neighbors, leaf_position, block_size,            \
                     block_offsets, num_blocks, (OUT*)block_leaf_ids,                          \
                     (OUT*)block_source_leaf_ids, valid_mask)
  if (out_i64 && in_i64)
    YGG_TB(int64_t, int64_t);
  else if (out_i64)
    YGG_TB(int64_t, int32_t);
  else if (in_i64)
    YGG_TB(int32_t, int64_t);
  else
    YGG_TB(int32_t, int32_t);
#undef YGG_TB
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_near_output(int64_t num_leaves, const int64_t* offsets, const int32_t* counts,
                    int32_t fixed_cap, const int64_t* row_start, const int32_t* entries,
                    const int32_t* leaf_position, int64_t max_count, int out_i64,
                    void* neighbors, void* neighbor_leaf_positions, void* stream) {
  if (num_leaves < 1) return YGG_ERR_ARG;
  int L = (int)num_leaves;
  dim3 grid(grid_for((int64_t)L * YGG_WAVE, 256));
  hipStream_t st = (hipStream_t)stream;
  if (out_i64)
    hipLaunchKernelGGL(k_near_expand<int64_t>, grid, dim3(256), 0, st, L, offsets, counts,
                       fixed_cap, row_start, entries, leaf_position, max_count,
                       (int64_t*)neighbors, (int64_t*)neighbor_leaf_positions);
  else
    hipLaunchKernelGGL(k_near_expand<int32_t>, grid, dim3(256), 0, st, L, offsets, counts,
                       fixed_cap, row_start, entries, leaf_position, max_count,
                       (int32_t*)neighbors, (int32_t*)neighbor_leaf_positions);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
