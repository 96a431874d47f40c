// Downstream layouts of the far-field lists, the formats M2L kernels consume:
//  * dense level-major buffers (yggdrax/dense_interactions.py:35-140): for
//    slot (level, j) the target node nodes_by_level[level_offsets[level] + j],
//    its far sources padded to the longest list, the validity mask, the count
//    and the target-minus-source centre displacements;
//  * displacement classes (yggdrax/grouped_interactions.py:74-207): far pairs
//    keyed by (target level, source level, rint(delta / cell) per axis),
//    ordered lexicographically by key and, inside a class, by input position,
//    with CSR class offsets and the class of every input pair.
#include "ygg_common.h"

namespace ygg {
namespace {

// One thread per (slot, entry): consecutive threads write consecutive
// entries of one slot, so every output stream is written coalesced and the
// slot's CSR row is read once per wavefront. A CSR row reaching outside the
// num_far sources, or a source outside [0, num_nodes), is written as
// padding and raises *err (checked by the caller after the launch).
template <typename IdxT>
__global__ void k_dense_m2l(const IdxT* __restrict__ offsets, const IdxT* __restrict__ counts,
                            const IdxT* __restrict__ sources, int64_t num_nodes, int64_t num_far,
                            const int32_t* __restrict__ nbl, const int32_t* __restrict__ loff,
                            int num_levels, int64_t max_nodes, int64_t K,
                            const float* __restrict__ center, IdxT* __restrict__ out_src,
                            float* __restrict__ out_disp, uint8_t* __restrict__ out_mask,
                            IdxT* __restrict__ out_counts, int32_t* __restrict__ err) {
  const int64_t total = (int64_t)num_levels * max_nodes * K;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t slot = e / K;
    const int64_t j = e - slot * K;
    const int l = (int)(slot / max_nodes);
    const int64_t s = slot - (int64_t)l * max_nodes;
    const int64_t lc = (int64_t)loff[l + 1] - loff[l];
    const int64_t node = s < lc ? (int64_t)nbl[loff[l] + s] : -1;
    int64_t cnt = node >= 0 ? (int64_t)counts[node] : 0;
    bool bad = cnt < 0;
    if (bad) cnt = 0;
    bool valid = j < cnt;
    int64_t src = -1;
    if (valid) {
      const int64_t at = (int64_t)offsets[node] + j;
      if (at < 0 || at >= num_far) {
        bad = true;
      } else {
        src = (int64_t)sources[at];
        if (src < 0 || src >= num_nodes) bad = true;
      }
      if (bad) {
        valid = false;
        src = -1;
      }
    }
    if (bad) *err = 1;
    out_src[e] = (IdxT)src;
    out_mask[e] = valid ? 1 : 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      out_disp[3 * e + d] = valid ? center[3 * node + d] - center[3 * src + d] : 0.0f;
    if (j == 0) out_counts[slot] = (IdxT)cnt;
  }
}

// numpy's float64 -> int32 cast on x86-64 (cvttsd2si): values outside the
// int32 range, and NaN, become INT32_MIN. The argument is already integral.
__device__ __forceinline__ int32_t to_i32_like_numpy(double r) {
  return (r >= -2147483648.0 && r <= 2147483647.0) ? (int32_t)r : INT32_MIN;
}

// Per-axis cell size of a level: max(extent / max(2^level, 1), 1e-12)
// (grouped_interactions.py:34-38); 2^level and the division by it are exact.
__device__ __forceinline__ double cell_size(double extent, int level) {
  return fmax(extent / fmax(ldexp(1.0, level), 1.0), 1e-12);
}

struct GroupGeom {
  int64_t num_nodes;     // T
  const int32_t* level;  // node_level [T]
  const float* center;   // [T, 3]
  const float* bmin;     // [3]
  const float* bmax;     // [3]
};

// Key fields of pair i: target level, source level and the rounded
// displacement in target-level cells (grouped_interactions.py:139-165).
// Arithmetic is float64 as in the reference's NumPy code: the centre
// difference, the IEEE division and the round-half-even are all exact
// restatements, so the key is bit-identical.
template <typename IdxT>
__device__ __forceinline__ void group_fields(const IdxT* __restrict__ src,
                                             const IdxT* __restrict__ tgt, int64_t i,
                                             const GroupGeom& g, int32_t f[5]) {
  const int64_t t = (int64_t)tgt[i], s = (int64_t)src[i];
  const int tl = g.level[t];
  f[0] = tl;
  f[1] = g.level[s];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double ext = (double)g.bmax[a] - (double)g.bmin[a];
    const double delta = (double)g.center[3 * t + a] - (double)g.center[3 * s + a];
    f[2 + a] = to_i32_like_numpy(rint(delta / cell_size(ext, tl)));
  }
}

// Field-wise min (partials[b][0..4]) and max ([5..9]) of block b's pairs:
// wavefront shuffles, then one LDS row per wavefront. No atomics, so the
// reduction costs no cross-XCD traffic.
constexpr int kMinmaxBlocks = 1024;

// Pairs naming a node outside [0, T) are skipped and raise *err; the
// classify call then fails before any kernel reads through them.
template <typename IdxT>
__global__ void __launch_bounds__(256)
k_group_minmax(const IdxT* __restrict__ src, const IdxT* __restrict__ tgt, int64_t n,
               GroupGeom g, int32_t* __restrict__ partials, int32_t* __restrict__ err) {
  __shared__ int32_t red[256 / YGG_WAVE][10];
  int32_t lo[5], hi[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    lo[k] = INT32_MAX;
    hi[k] = INT32_MIN;
  }
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = (int64_t)tgt[i], s = (int64_t)src[i];
    if (t < 0 || t >= g.num_nodes || s < 0 || s >= g.num_nodes) {
      *err = 1;
      continue;
    }
    int32_t f[5];
    group_fields(src, tgt, i, g, f);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      lo[k] = min(lo[k], f[k]);
      hi[k] = max(hi[k], f[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    for (int o = YGG_WAVE / 2; o > 0; o >>= 1) {
      lo[k] = min(lo[k], __shfl_xor(lo[k], o));
      hi[k] = max(hi[k], __shfl_xor(hi[k], o));
    }
  }
  const int wave = threadIdx.x / YGG_WAVE;
  if ((threadIdx.x & (YGG_WAVE - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      red[wave][k] = lo[k];
      red[wave][5 + k] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    const int k = threadIdx.x;
    int32_t v = red[0][k];
    for (int w = 1; w < 256 / YGG_WAVE; ++w) v = k < 5 ? min(v, red[w][k]) : max(v, red[w][k]);
    partials[10 * blockIdx.x + k] = v;
  }
}

// Field offsets and widths of the packed key; the fields are concatenated
// most significant first (target level, source level, dx, dy, dz), each
// minus its minimum, so unsigned order of the packed value is the
// lexicographic order of the signed fields. total = sum of widths.
struct Pack {
  int32_t min[5];
  int32_t width[5];
  int32_t total;
  int32_t bad;  // a pair named a node outside [0, T)
};

__global__ void __launch_bounds__(256)
k_group_pack_params(const int32_t* __restrict__ partials, int nblocks,
                    const int32_t* __restrict__ err, Pack* __restrict__ pk) {
  __shared__ int32_t red[256 / YGG_WAVE][10];
  int32_t v[10];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    v[k] = INT32_MAX;
    v[5 + k] = INT32_MIN;
  }
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      v[k] = min(v[k], partials[10 * b + k]);
      v[5 + k] = max(v[5 + k], partials[10 * b + 5 + k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    for (int o = YGG_WAVE / 2; o > 0; o >>= 1) {
      v[k] = min(v[k], __shfl_xor(v[k], o));
      v[5 + k] = max(v[5 + k], __shfl_xor(v[5 + k], o));
    }
  }
  if ((threadIdx.x & (YGG_WAVE - 1)) == 0) {
#pragma unroll
    for (int k = 0; k < 10; ++k) red[threadIdx.x / YGG_WAVE][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int k = 0; k < 5; ++k) {
      int32_t lo = red[0][k], hi = red[0][5 + k];
      for (int w = 1; w < 256 / YGG_WAVE; ++w) {
        lo = min(lo, red[w][k]);
        hi = max(hi, red[w][5 + k]);
      }
      const uint32_t span = (uint32_t)((int64_t)hi - (int64_t)lo);
      const int width = span == 0 ? 0 : 32 - __clz((int)span);
      pk->min[k] = lo;
      pk->width[k] = width;
      total += width;
    }
    pk->total = total;
    pk->bad = *err;
  }
}

template <typename IdxT>
__global__ void k_group_pack(const IdxT* __restrict__ src, const IdxT* __restrict__ tgt,
                             int64_t n, GroupGeom g, Pack pk, uint64_t* __restrict__ lo_out,
                             uint64_t* __restrict__ hi_out, uint64_t* __restrict__ pairs) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int32_t f[5];
    group_fields(src, tgt, i, g, f);
    // node ids are below 2^31: one 8-byte record per pair for the emit gather
    pairs[i] = ((uint64_t)(uint32_t)tgt[i] << 32) | (uint64_t)(uint32_t)src[i];
    uint64_t lo = 0, hi = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const int w = pk.width[k];  // 0..32
      if (w == 0) continue;
      const uint64_t u = (uint64_t)((int64_t)f[k] - (int64_t)pk.min[k]);
      hi = (hi << w) | (lo >> (64 - w));
      lo = (lo << w) | u;
    }
    lo_out[i] = lo;
    if (hi_out) hi_out[i] = hi;
  }
}

__global__ void k_gather_keys(const uint64_t* __restrict__ v, const uint32_t* __restrict__ idx,
                              int64_t n, uint64_t* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = v[idx[i]];
}

__global__ void k_iota_u32(uint32_t* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (uint32_t)i;
}

// 1 where sorted position i starts a class: its packed key (lo_s, hi_s,
// already in sorted order) differs from position i - 1's.
__global__ void k_group_flags(const uint64_t* __restrict__ lo_s, const uint64_t* __restrict__ hi_s,
                              int64_t n, int32_t* __restrict__ flags) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = i == 0 || lo_s[i] != lo_s[i - 1] || (hi_s && hi_s[i] != hi_s[i - 1]);
}

// class_scan[i] = classes started before sorted position i, so position i
// belongs to class class_scan[i + 1] - 1 and starts it iff the two differ.
// A class head decodes its key fields from the sorted packed key.
template <typename IdxT>
__global__ void k_group_emit(const uint64_t* __restrict__ pairs, int64_t n,
                             const float* __restrict__ bmin,
                             const float* __restrict__ bmax, const Pack* __restrict__ pkp,
                             const uint64_t* __restrict__ lo_s, const uint64_t* __restrict__ hi_s,
                             const uint32_t* __restrict__ order, const int64_t* __restrict__ scan,
                             IdxT* __restrict__ csrc, IdxT* __restrict__ ctgt,
                             IdxT* __restrict__ cids, IdxT* __restrict__ coff,
                             int32_t* __restrict__ ckeys, float* __restrict__ cdisp) {
  const Pack pk = *pkp;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = order[i];
    const int64_t c = scan[i + 1] - 1;
    const uint64_t tp = pairs[p];
    csrc[i] = (IdxT)(uint32_t)tp;
    ctgt[i] = (IdxT)(uint32_t)(tp >> 32);
    cids[p] = (IdxT)c;
    if (scan[i + 1] != scan[i]) {
      coff[c] = (IdxT)i;
      uint64_t lo = lo_s[i], hi = pk.total > 64 ? hi_s[i] : 0;
      int32_t f[5];
#pragma unroll
      for (int k = 4; k >= 0; --k) {
        const int w = pk.width[k];
        uint64_t bits = 0;
        if (w > 0) {
          bits = lo & ((1ull << w) - 1ull);
          lo = (lo >> w) | (hi << (64 - w));
          hi >>= w;
        }
        f[k] = (int32_t)((int64_t)pk.min[k] + (int64_t)bits);
      }
#pragma unroll
      for (int k = 0; k < 5; ++k) ckeys[5 * c + k] = f[k];
      // class displacement = index * target-level cell, rounded to float32
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double ext = (double)bmax[a] - (double)bmin[a];
        cdisp[3 * c + a] = (float)((double)f[2 + a] * cell_size(ext, f[0]));
      }
    }
    if (i == n - 1) coff[scan[n]] = (IdxT)n;
  }
}

// The radix sort ranks at most 2^30 - 1 elements per call.
constexpr int64_t kMaxGroupedPairs = (1ll << 30) - 1;

// Workspace shared by the classify and emit calls. After classify, lo_s
// holds the packed keys' low words and hi holds the high words, both in
// sorted order, and pack the packing parameters.
struct GroupWs {
  int32_t* partials;
  int32_t* err;
  Pack* pack;
  uint64_t* lo;
  uint64_t* hi;
  uint64_t* hi_in;
  uint64_t* lo_s;
  uint64_t* pairs;
  uint32_t* order1;
  int32_t* flags;
  void* sort_ws;
  size_t sort_bytes;
  void* scan_ws;
  size_t scan_bytes;
};

static bool group_ws_layout(int64_t n, void* ws, size_t ws_bytes, GroupWs* w, size_t* bytes) {
  Arena a(ws, ws_bytes);
  w->partials = a.take<int32_t>(10 * kMinmaxBlocks);
  w->err = a.take<int32_t>(1);
  w->pack = a.take<Pack>(1);
  w->lo = a.take<uint64_t>((size_t)n);
  w->hi = a.take<uint64_t>((size_t)n);
  w->hi_in = a.take<uint64_t>((size_t)n);
  w->lo_s = a.take<uint64_t>((size_t)n);
  w->pairs = a.take<uint64_t>((size_t)n);
  w->order1 = a.take<uint32_t>((size_t)n);
  w->flags = a.take<int32_t>((size_t)n);
  w->sort_bytes = sort_workspace_bytes(n, 8, 64);
  w->sort_ws = a.take<char>(w->sort_bytes);
  w->scan_bytes = scan_workspace_bytes(n);
  w->scan_ws = a.take<char>(w->scan_bytes);
  if (bytes) *bytes = a.off;
  return w->partials && w->err && w->pack && w->lo && w->hi && w->hi_in && w->lo_s && w->pairs &&
         w->order1 &&
         w->flags && w->sort_ws && w->scan_ws;
}

template <typename IdxT>
int grouped_classify(const IdxT* src, const IdxT* tgt, int64_t n, const GroupGeom& g,
                     uint32_t* order, int64_t* scan, int64_t* num_classes, void* ws,
                     size_t ws_bytes, hipStream_t st) {
  GroupWs w;
  if (!group_ws_layout(n, ws, ws_bytes, &w, nullptr)) return YGG_ERR_WORKSPACE;
  const int mgrid = grid_for(n, 256, kMinmaxBlocks);
  int rc;
  if ((rc = fill_u32((uint32_t*)w.err, 0u, 1, st))) return rc;
  hipLaunchKernelGGL(k_group_minmax<IdxT>, dim3(mgrid), dim3(256), 0, st, src, tgt, n, g,
                     w.partials, w.err);
  YGG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_group_pack_params, dim3(1), dim3(256), 0, st, w.partials, mgrid, w.err,
                     w.pack);
  YGG_LAUNCH_CHECK();
  Pack pk;
  YGG_CHECK(hipMemcpyAsync(&pk, w.pack, sizeof(Pack), hipMemcpyDeviceToHost, st));
  YGG_CHECK(hipStreamSynchronize(st));
  if (pk.bad) return YGG_ERR_ARG;
  const int W = pk.total;
  if (W > 128) return YGG_ERR_UNSUPPORTED;
  const int pgrid = grid_for(n, 256);
  hipLaunchKernelGGL(k_group_pack<IdxT>, dim3(pgrid), dim3(256), 0, st, src, tgt, n, g, pk, w.lo,
                     W > 64 ? w.hi : nullptr, w.pairs);
  YGG_LAUNCH_CHECK();
  if (W == 0) {  // one class: identity order, all keys zero
    hipLaunchKernelGGL(k_iota_u32, dim3(pgrid), dim3(256), 0, st, order, n);
    YGG_LAUNCH_CHECK();
    if ((rc = copy_u32((uint32_t*)w.lo_s, (const uint32_t*)w.lo, 2 * n, st))) return rc;
  } else if (W <= 64) {
    if ((rc = sort_pairs_u64(w.lo, nullptr, w.lo_s, order, n, 0, W, w.sort_ws, w.sort_bytes,
                             st)))
      return rc;
  } else {
    // stable LSD over two words: the low word, then the high word carrying
    // the first order; then the low words gathered into sorted order
    if ((rc = sort_pairs_u64(w.lo, nullptr, w.lo_s, w.order1, n, 0, 64, w.sort_ws, w.sort_bytes,
                             st)))
      return rc;
    hipLaunchKernelGGL(k_gather_keys, dim3(pgrid), dim3(256), 0, st, w.hi, w.order1, n,
                       w.hi_in);
    YGG_LAUNCH_CHECK();
    if ((rc = sort_pairs_u64(w.hi_in, w.order1, w.hi, order, n, 0, W - 64, w.sort_ws,
                             w.sort_bytes, st)))
      return rc;
    hipLaunchKernelGGL(k_gather_keys, dim3(pgrid), dim3(256), 0, st, w.lo, order, n, w.lo_s);
    YGG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_group_flags, dim3(pgrid), dim3(256), 0, st, w.lo_s,
                     W > 64 ? w.hi : nullptr, n, w.flags);
  YGG_LAUNCH_CHECK();
  if ((rc = exclusive_scan_i32_to_i64(w.flags, nullptr, scan, n, w.scan_ws, w.scan_bytes, st)))
    return rc;
  int64_t total = 0;
  YGG_CHECK(hipMemcpyAsync(&total, scan + n, sizeof(total), hipMemcpyDeviceToHost, st));
  YGG_CHECK(hipStreamSynchronize(st));
  *num_classes = total;
  return YGG_OK;
}

}  // namespace
}  // namespace ygg

using namespace ygg;

extern "C" {

int ygg_dense_interactions(const void* offsets, const void* counts, const void* sources,
                           int idx_i64, int64_t num_nodes, int64_t num_far,
                           const int32_t* nodes_by_level, const int32_t* level_offsets,
                           int32_t num_levels, int64_t max_nodes, int64_t max_interactions,
                           const float* center, void* out_sources, float* out_displacements,
                           uint8_t* out_mask, void* out_counts, int32_t* error_flag,
                           void* stream) {
  if (num_levels < 1 || max_nodes < 1 || max_interactions < 1 || num_nodes < 1 ||
      num_far < 0 || !offsets || !counts || !sources || !nodes_by_level || !level_offsets ||
      !center || !out_sources || !out_displacements || !out_mask || !out_counts || !error_flag)
    return YGG_ERR_ARG;
  if ((int64_t)num_levels * max_nodes > (1ll << 40) / max_interactions) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = fill_u32((uint32_t*)error_flag, 0u, 1, st))) return rc;
  const int64_t total = (int64_t)num_levels * max_nodes * max_interactions;
  dim3 g(grid_for(total, 256));
  if (idx_i64)
    hipLaunchKernelGGL(k_dense_m2l<int64_t>, g, dim3(256), 0, st, (const int64_t*)offsets,
                       (const int64_t*)counts, (const int64_t*)sources, num_nodes, num_far,
                       nodes_by_level, level_offsets, num_levels, max_nodes, max_interactions,
                       center, (int64_t*)out_sources, out_displacements, out_mask,
                       (int64_t*)out_counts, error_flag);
  else
    hipLaunchKernelGGL(k_dense_m2l<int32_t>, g, dim3(256), 0, st, (const int32_t*)offsets,
                       (const int32_t*)counts, (const int32_t*)sources, num_nodes, num_far,
                       nodes_by_level, level_offsets, num_levels, max_nodes, max_interactions,
                       center, (int32_t*)out_sources, out_displacements, out_mask,
                       (int32_t*)out_counts, error_flag);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_grouped_workspace_bytes(int64_t num_pairs) {
  if (num_pairs < 1) return 0;
  GroupWs w;
  size_t bytes = 0;
  group_ws_layout(num_pairs, nullptr, 0, &w, &bytes);
  return bytes;
}

int ygg_grouped_classify(const void* sources, const void* targets, int64_t num_pairs,
                         int idx_i64, int64_t num_nodes, const int32_t* node_level,
                         const float* center,
                         const float* bounds_min, const float* bounds_max, uint32_t* order,
                         int64_t* class_scan, int64_t* num_classes, void* ws, size_t ws_bytes,
                         void* stream) {
  if (num_pairs < 1 || !sources || !targets || !node_level || !center || !bounds_min ||
      !bounds_max || !order || !class_scan || !num_classes)
    return YGG_ERR_ARG;
  if (num_pairs > kMaxGroupedPairs) return YGG_ERR_UNSUPPORTED;
  if (num_nodes < 1) return YGG_ERR_ARG;
  GroupGeom g{num_nodes, node_level, center, bounds_min, bounds_max};
  hipStream_t st = (hipStream_t)stream;
  if (idx_i64)
    return grouped_classify((const int64_t*)sources, (const int64_t*)targets, num_pairs, g,
                            order, class_scan, num_classes, ws, ws_bytes, st);
  return grouped_classify((const int32_t*)sources, (const int32_t*)targets, num_pairs, g, order,
                          class_scan, num_classes, ws, ws_bytes, st);
}

int ygg_grouped_emit(int64_t num_pairs, int idx_i64, const float* bounds_min,
                     const float* bounds_max, const uint32_t* order, const int64_t* class_scan,
                     void* class_sources, void* class_targets, void* class_ids,
                     void* class_offsets, int32_t* class_keys, float* class_displacements,
                     void* ws, size_t ws_bytes, void* stream) {
  if (num_pairs < 1 || !bounds_min || !bounds_max || !order ||
      !class_scan || !class_sources || !class_targets || !class_ids || !class_offsets ||
      !class_keys || !class_displacements)
    return YGG_ERR_ARG;
  if (num_pairs > kMaxGroupedPairs) return YGG_ERR_UNSUPPORTED;
  GroupWs w;
  if (!group_ws_layout(num_pairs, ws, ws_bytes, &w, nullptr)) return YGG_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(grid_for(num_pairs, 256));
  if (idx_i64)
    hipLaunchKernelGGL(k_group_emit<int64_t>, grid, dim3(256), 0, st, w.pairs, num_pairs,
                       bounds_min, bounds_max, w.pack,
                       w.lo_s, w.hi, order, class_scan, (int64_t*)class_sources,
                       (int64_t*)class_targets, (int64_t*)class_ids, (int64_t*)class_offsets,
                       class_keys, class_displacements);
  else
    hipLaunchKernelGGL(k_group_emit<int32_t>, grid, dim3(256), 0, st, w.pairs, num_pairs,
                       bounds_min, bounds_max, w.pack,
                       w.lo_s, w.hi, order, class_scan, (int32_t*)class_sources,
                       (int32_t*)class_targets, (int32_t*)class_ids, (int32_t*)class_offsets,
                       class_keys, class_displacements);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
