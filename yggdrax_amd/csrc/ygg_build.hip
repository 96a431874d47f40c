// Tree construction kernels: bounds, Morton encode (+ digit histograms),
// particle reorder, Karras radix build, node levels and explicit octree
// metadata. Integer work is exact; the float work of bounds and encode uses
// separately rounded IEEE operations (the library is built without FMA
// contraction) in the order the reference writes them.
#include "ygg_common.h"
#include "ygg_build_dev.h"

#include <mutex>

namespace ygg {

// ---- bounds (bounds.py:29-33) ----------------------------------------------
// Flat grid-stride over 3n floats with a stride that is a multiple of 3, so
// every thread stays on one axis.
__global__ __launch_bounds__(256) void k_bounds_reduce(const float* __restrict__ pos, int64_t n3,
                                                       uint32_t* __restrict__ acc) {
  __shared__ uint32_t smin[3], smax[3];
  if (threadIdx.x < 3) {
    smin[threadIdx.x] = 0xFFFFFFFFu;
    smax[threadIdx.x] = 0u;
  }
  __syncthreads();
  int64_t start = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int axis = (int)(start % 3);
  uint32_t mn = 0xFFFFFFFFu, mx = 0u;
  for (int64_t i = start; i < n3; i += stride) {
    uint32_t o = f2ord(pos[i]);
    mn = min(mn, o);
    mx = max(mx, o);
  }
  atomicMin(&smin[axis], mn);
  atomicMax(&smax[axis], mx);
  __syncthreads();
  if (threadIdx.x < 3) {
    atomicMin(&acc[threadIdx.x], smin[threadIdx.x]);
    atomicMax(&acc[3 + threadIdx.x], smax[threadIdx.x]);
  }
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

// Bounds from 16-byte aligned positions: four points (three float4 loads,
// x y z x | y z x y | z x y z) per thread and step, wave reductions, and one
// atomic per axis and block. Measured on gfx950 at 2^24 points: 46 us at
// 1024 blocks, against 75-100 us for one float per load; more blocks are
// slower (their atomics meet on six addresses).
__global__ __launch_bounds__(256) void k_bounds_reduce4(const float* __restrict__ pos, int64_t n,
                                                        uint32_t* __restrict__ acc) {
  __shared__ uint32_t s[4][6];
  const float4* p4 = reinterpret_cast<const float4*>(pos);
  const int64_t chunks = n / 4;
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < chunks;
       c += (int64_t)gridDim.x * blockDim.x) {
    const float4 a = p4[3 * c], b = p4[3 * c + 1], d = p4[3 * c + 2];
    const float v[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, d.x, d.y, d.z, d.w};
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const uint32_t o = f2ord(v[k]);
      mn[k % 3] = min(mn[k % 3], o);
      mx[k % 3] = max(mx[k % 3], o);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < (n & 3)) {  // the last n % 4 points
    const int64_t i = chunks * 4 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t o = f2ord(pos[3 * i + k]);
      mn[k] = min(mn[k], o);
      mx[k] = max(mx[k], o);
    }
  }
  const int w = threadIdx.x / YGG_WAVE;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    mn[k] = wave_min_u32(mn[k]);
    mx[k] = wave_max_u32(mx[k]);
  }
  if (lane_id() == 0)
    for (int k = 0; k < 3; ++k) {
      s[w][k] = mn[k];
      s[w][3 + k] = mx[k];
    }
  __syncthreads();
  if (threadIdx.x < 3) {
    uint32_t a = s[0][threadIdx.x], b = s[0][3 + threadIdx.x];
    for (int q = 1; q < 256 / YGG_WAVE; ++q) {
      a = min(a, s[q][threadIdx.x]);
      b = max(b, s[q][3 + threadIdx.x]);
    }
    atomicMin(&acc[threadIdx.x], a);
    atomicMax(&acc[3 + threadIdx.x], b);
  }
}

__global__ void k_bounds_init(uint32_t* acc) {
  if (threadIdx.x < 3) {
    acc[threadIdx.x] = 0xFFFFFFFFu;
    acc[3 + threadIdx.x] = 0u;
  }
}

__global__ void k_bounds_final(const uint32_t* __restrict__ acc, float* __restrict__ b6) {
  int d = threadIdx.x;
  if (d < 3) padded_bounds(ord2f(acc[d]), ord2f(acc[3 + d]), &b6[d], &b6[3 + d]);
}

// ---- Morton encode (morton.py:68-77) ----------------------------------------
__global__ __launch_bounds__(256) void k_encode(const float* __restrict__ pos, int64_t n,
                                                const float* __restrict__ b6,
                                                uint64_t* __restrict__ codes) {
  float lo[3] = {b6[0], b6[1], b6[2]};
  float w[3] = {b6[3] - b6[0], b6[4] - b6[1], b6[5] - b6[2]};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    codes[i] = encode_point(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], lo, w);
}

// Encode and count all eight 8-bit digit histograms of the codes.
__global__ __launch_bounds__(256) void k_encode_hist(const float* __restrict__ pos, int64_t n,
                                                     const float* __restrict__ b6,
                                                     uint64_t* __restrict__ codes,
                                                     uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kMortonHistWords];
  for (int i = threadIdx.x; i < kMortonHistWords; i += blockDim.x) h[i] = 0;
  __syncthreads();
  float lo[3] = {b6[0], b6[1], b6[2]};
  float w[3] = {b6[3] - b6[0], b6[4] - b6[1], b6[5] - b6[2]};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // uniform trip count per wave: every lane takes part in the ballots
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
    const int64_t i = b0 + threadIdx.x;
    const bool valid = i < n;
    uint64_t c = 0;
    if (valid) {
      c = encode_point(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], lo, w);
      codes[i] = c;
    }
#pragma unroll
    for (int p = 0; p < kMortonPasses; ++p) morton_hist_add(h, c, valid, p);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMortonHistWords; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ __launch_bounds__(256) void k_reorder(const float* __restrict__ pos,
                                                 const float* __restrict__ mass,
                                                 const uint32_t* __restrict__ perm, int64_t n,
                                                 float* __restrict__ pos_sorted,
                                                 float* __restrict__ mass_sorted,
                                                 int32_t* __restrict__ inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t p = perm[i];
    if (pos_sorted) {
      pos_sorted[3 * i] = pos[3 * (int64_t)p];
      pos_sorted[3 * i + 1] = pos[3 * (int64_t)p + 1];
      pos_sorted[3 * i + 2] = pos[3 * (int64_t)p + 2];
    }
    if (mass && mass_sorted) mass_sorted[i] = mass[p];
    if (inv) inv[p] = (int32_t)i;
  }
}

// Encode + digit histograms (as k_encode_hist) that also packs (x, y, z, m)
// into one float4 per particle, so the later reorder gathers one 16-byte
// record (one cache line) per particle instead of two.
__global__ __launch_bounds__(256) void k_encode_hist_pack(const float* __restrict__ pos,
                                                          const float* __restrict__ mass,
                                                          int64_t n, const float* __restrict__ b6,
                                                          uint64_t* __restrict__ codes,
                                                          uint32_t* __restrict__ hist,
                                                          float4* __restrict__ packed) {
  __shared__ uint32_t h[kMortonHistWords];
  for (int i = threadIdx.x; i < kMortonHistWords; i += blockDim.x) h[i] = 0;
  __syncthreads();
  float lo[3] = {b6[0], b6[1], b6[2]};
  float w[3] = {b6[3] - b6[0], b6[4] - b6[1], b6[5] - b6[2]};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < n; b0 += stride) {
    const int64_t i = b0 + threadIdx.x;
    const bool valid = i < n;
    uint64_t c = 0;
    if (valid) {
      const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
      c = encode_point(x, y, z, lo, w);
      codes[i] = c;
      packed[i] = make_float4(x, y, z, mass ? mass[i] : 0.0f);
    }
#pragma unroll
    for (int p = 0; p < kMortonPasses; ++p) morton_hist_add(h, c, valid, p);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMortonHistWords; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Sorted positions and masses: one 16-byte gather per particle. The
// inverse permutation is written by its own kernel (k_inverse_perm): the
// gather and the scattered inverse writes together took 682 us at 2^24
// clustered points on gfx950, apart 360 + 200 us.
__global__ __launch_bounds__(256) void k_reorder_packed(const float4* __restrict__ packed,
                                                        const uint32_t* __restrict__ perm,
                                                        int64_t n, float* __restrict__ pos_sorted,
                                                        float* __restrict__ mass_sorted) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = perm[i];
    const float4 r = packed[p];
    pos_sorted[3 * i] = r.x;
    pos_sorted[3 * i + 1] = r.y;
    pos_sorted[3 * i + 2] = r.z;
    if (mass_sorted) mass_sorted[i] = r.w;
  }
}

// One launch widening several int32 arrays: each item owns a contiguous
// range of blocks (block_start), found by a block-uniform search.
struct WidenArgs {
  const int32_t* src[YGG_WIDEN_MAX_ITEMS];
  int64_t* dst[YGG_WIDEN_MAX_ITEMS];
  int64_t n[YGG_WIDEN_MAX_ITEMS];
  int32_t block_start[YGG_WIDEN_MAX_ITEMS + 1];
  int32_t count;
};

__global__ __launch_bounds__(256) void k_widen_i32(WidenArgs a) {
  int k = 0;
  while (k + 1 < a.count && (int)blockIdx.x >= a.block_start[k + 1]) ++k;
  const int nb = a.block_start[k + 1] - a.block_start[k];
  const int64_t b = (int64_t)blockIdx.x - a.block_start[k];
  const int32_t* __restrict__ s = a.src[k];
  int64_t* __restrict__ d = a.dst[k];
  const int64_t n = a.n[k];
  for (int64_t i = b * 256 + threadIdx.x; i < n; i += (int64_t)nb * 256) d[i] = (int64_t)s[i];
}

template <typename InvT>
__global__ __launch_bounds__(256) void k_inverse_perm(const uint32_t* __restrict__ perm, int64_t n,
                                                      InvT* __restrict__ inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    inv[perm[i]] = (InvT)i;
}

// Reorder by scattering: the inverse permutation first (coalesced reads of
// perm), then every particle read in input order and written to its sorted
// slot, so all reads are coalesced and only writes are scattered.
__global__ __launch_bounds__(256) void k_inverse(const uint32_t* __restrict__ perm, int64_t n,
                                                 int32_t* __restrict__ inv) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    inv[perm[i]] = (int32_t)i;
}

__global__ __launch_bounds__(256) void k_scatter_sorted(const float* __restrict__ pos,
                                                        const float* __restrict__ mass,
                                                        const int32_t* __restrict__ inv,
                                                        int64_t n, float* __restrict__ pos_sorted,
                                                        float* __restrict__ mass_sorted) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = inv[j];
    pos_sorted[3 * i] = pos[3 * j];
    pos_sorted[3 * i + 1] = pos[3 * j + 1];
    pos_sorted[3 * i + 2] = pos[3 * j + 2];
    if (mass && mass_sorted) mass_sorted[i] = mass[j];
  }
}

// ---- Karras radix build (_tree_impl.py:1117-1270) ---------------------------
__global__ __launch_bounds__(256) void k_leaves(const uint64_t* __restrict__ sc, int64_t n,
                                                int64_t leaf, int L, uint64_t* __restrict__ lc,
                                                int32_t* __restrict__ ranges) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  int64_t s = (int64_t)j * leaf;
  int64_t e = min(s + leaf, n);
  lc[j] = sc[s];
  int64_t node = (int64_t)(L - 1) + j;
  ranges[2 * node] = (int32_t)s;
  ranges[2 * node + 1] = (int32_t)(e - 1);
}

__global__ __launch_bounds__(256) void k_karras(const uint64_t* __restrict__ lc, int L,
                                                int64_t n, int64_t leaf,
                                                int32_t* __restrict__ left,
                                                int32_t* __restrict__ right,
                                                uint8_t* __restrict__ lflag,
                                                uint8_t* __restrict__ rflag,
                                                int32_t* __restrict__ ranges,
                                                uint32_t* __restrict__ penc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L - 1) return;
  karras_node(lc, L, n, leaf, i, left, right, lflag, rflag, ranges, penc);
}

// scalars: [0] num_levels, [1] max_level, [2] deep, [3] malformed, [4] roots, [5] root
__global__ __launch_bounds__(256) void k_levels_walk(const uint32_t* __restrict__ penc, int T,
                                                     int NI, const int32_t* __restrict__ left,
                                                     const int32_t* __restrict__ right,
                                                     int32_t* __restrict__ parent,
                                                     int32_t* __restrict__ level,
                                                     uint32_t* __restrict__ lcount,
                                                     int32_t* __restrict__ scal) {
  // level histogram and max level reduced in LDS first: one global atomic
  // per workgroup on the max-level word instead of one per node
  __shared__ uint32_t h[64];
  __shared__ int s_maxd;
  if (threadIdx.x < 64) h[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_maxd = 0;
  __syncthreads();
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < T) {
    int32_t p = dec_parent(penc[k]);
    parent[k] = p;
    int d = 0;
    int32_t q = p;
    bool deep = false;
    while (q >= 0) {
      ++d;
      if (d > 255) {
        deep = true;
        break;
      }
      q = dec_parent(penc[q]);
    }
    level[k] = d;
    if (deep) atomicOr((unsigned*)&scal[2], 1u);
    if (d < 64) atomicAdd(&h[d], 1u);
    atomicMax(&s_maxd, d);
    if (p < 0) atomicAdd(&scal[4], 1);
    if (k < NI) {
      if (dec_parent(penc[left[k]]) != k || dec_parent(penc[right[k]]) != k)
        atomicOr((unsigned*)&scal[3], 1u);
    }
  }
  __syncthreads();
  if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&lcount[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 0 && k - (int)threadIdx.x < T) atomicMax(&scal[1], s_maxd);
}

__global__ void k_level_offsets(const uint32_t* __restrict__ lcount, int32_t* __restrict__ off,
                                int32_t* __restrict__ scal) {
  if (threadIdx.x == 0) {
    int32_t s = 0;
    off[0] = 0;
    for (int k = 0; k < 64; ++k) {
      s += (int32_t)lcount[k];
      off[k + 1] = s;
    }
    int nl = scal[1] + 1;
    scal[0] = nl < 1 ? 1 : (nl > 64 ? 64 : nl);
    scal[5] = 0;
  }
}

__global__ void k_single_leaf(int64_t n, int32_t* parent, int32_t* ranges, int32_t* level,
                              int32_t* off, int32_t* nbl, int32_t* scal) {
  if (threadIdx.x == 0) {
    parent[0] = -1;
    ranges[0] = 0;
    ranges[1] = (int32_t)(n - 1);
    level[0] = 0;
    nbl[0] = 0;
    for (int k = 0; k < 65; ++k) off[k] = 0;
    off[1] = 1;
    scal[0] = 1;
    scal[1] = 0;
    scal[4] = 1;
    scal[5] = 0;
  }
}

// Bounded pointer doubling (_tree_impl.py:1280-1299).
__global__ void k_pd_init(const int32_t* __restrict__ parent, int T, int32_t* __restrict__ sc,
                          int64_t* __restrict__ dist) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  bool root = parent[k] < 0;
  sc[k] = root ? k : parent[k];
  dist[k] = root ? 0 : 1;
}
__global__ void k_pd_round(const int32_t* __restrict__ sc, const int64_t* __restrict__ dist,
                           int T, int32_t* __restrict__ sc2, int64_t* __restrict__ dist2) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  int32_t s = sc[k];
  dist2[k] = dist[k] + dist[s];
  sc2[k] = sc[s];
}
__global__ __launch_bounds__(256) void k_pd_finish(const int64_t* __restrict__ dist, int T,
                                                   int32_t* __restrict__ level,
                                                   uint32_t* __restrict__ lcount,
                                                   int32_t* __restrict__ scal) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  int64_t d = dist[k];
  int32_t dv = d > 0x7FFFFFFF ? 0x7FFFFFFF : (int32_t)d;
  level[k] = dv;
  if (dv >= 0 && dv < 64) atomicAdd(&lcount[dv], 1u);
  atomicMax(&scal[1], dv);
}

// ---- explicit octree metadata (octree.py:168-550) ---------------------------
__global__ void k_oct_resolve(const uint64_t* __restrict__ sc, const int32_t* __restrict__ ranges,
                              int L, uint64_t* __restrict__ rc, int32_t* __restrict__ rd) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L) return;
  int64_t node = (int64_t)(L - 1) + j;
  uint64_t f = sc[ranges[2 * node]], l = sc[ranges[2 * node + 1]];
  int d = common_depth(f, l);
  rd[j] = d;
  rc[j] = prefix_code(f, d);
}

__global__ void k_oct_candidates(const uint64_t* __restrict__ rc, const int32_t* __restrict__ rd,
                                 int L, int M, uint64_t* __restrict__ addr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  uint64_t c;
  int d;
  candidate(rc, rd, L, k, &c, &d);
  addr[k] = oct_address(c, d);
}

__global__ void k_oct_newflags(const uint64_t* __restrict__ sa, int M, int32_t* __restrict__ f) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  f[k] = (k == 0 || sa[k] != sa[k - 1]) ? 1 : 0;
}

__global__ void k_oct_unique(const uint64_t* __restrict__ sa, const uint32_t* __restrict__ order,
                             const uint64_t* __restrict__ rc, const int32_t* __restrict__ rd,
                             int L, int M, const int64_t* __restrict__ uscan,
                             uint64_t* __restrict__ codes, int32_t* __restrict__ depths,
                             uint8_t* __restrict__ valid, uint64_t* __restrict__ oaddr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const int64_t nu = uscan[M];
  bool leader = (k == 0) || sa[k] != sa[k - 1];
  if (leader) {
    int u = (int)uscan[k];
    uint64_t c;
    int d;
    candidate(rc, rd, L, (int)order[k], &c, &d);
    codes[u] = c;
    depths[u] = d;
    oaddr[u] = sa[k];
  }
  bool v = k < nu;
  valid[k] = v;
  if (!v) {
    codes[k] = 0;
    depths[k] = -1;
    oaddr[k] = ~0ull;
  }
}

// scal: [0] oct_num_levels, [1] num_valid, [2] num_leaf, [3] max depth
__global__ __launch_bounds__(256) void k_oct_ranges(
    const uint64_t* __restrict__ codes, const int32_t* __restrict__ depths,
    const uint8_t* __restrict__ valid, const uint64_t* __restrict__ rc,
    const int32_t* __restrict__ node_ranges, int L, int M, int32_t* __restrict__ oranges,
    uint32_t* __restrict__ lcount, uint32_t* __restrict__ skey, int32_t* __restrict__ scal) {
  // depth histogram and max depth reduced in LDS, then one global atomic
  // per workgroup on the max-depth word
  __shared__ uint32_t h[22];
  __shared__ int s_maxd;
  if (threadIdx.x < 22) h[threadIdx.x] = 0;
  if (threadIdx.x == 0) s_maxd = INT32_MIN;
  __syncthreads();
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < M) {
    bool v = valid[k];
    int d = depths[k];
    uint64_t c = codes[k];
    uint64_t du = (uint64_t)max(d, 0);
    uint64_t bs = 63ull - 3ull * du;
    uint64_t bss = bs < 62ull ? bs : 62ull;
    uint64_t mask = (1ull << bss) - 1ull;
    uint64_t endc = v ? (c | mask) : 0ull;
    int f = fixed_round_search(rc, L, c, true);
    int l = fixed_round_search(rc, L, endc, false) - 1;
    int bound = max(L - 1, 0);
    f = min(max(f, 0), bound);
    l = min(max(l, 0), bound);
    const int64_t NI = L - 1;
    if (v) {
      oranges[2 * k] = node_ranges[2 * (NI + f)];
      oranges[2 * k + 1] = node_ranges[2 * (NI + l) + 1];
      atomicAdd(&h[min(max(d, 0), 21)], 1u);
      atomicMax(&s_maxd, d);
    } else {
      oranges[2 * k] = 0;
      oranges[2 * k + 1] = -1;
    }
    skey[k] = v ? (uint32_t)d : 22u;
  }
  __syncthreads();
  if (threadIdx.x < 22 && h[threadIdx.x]) atomicAdd(&lcount[threadIdx.x], h[threadIdx.x]);
  if (threadIdx.x == 0 && s_maxd != INT32_MIN) atomicMax(&scal[3], s_maxd);
}

__global__ void k_oct_level_offsets(const uint32_t* __restrict__ lcount,
                                    int32_t* __restrict__ off, int32_t* __restrict__ scal,
                                    const int64_t* __restrict__ uscan, int M) {
  if (threadIdx.x == 0) {
    int32_t s = 0;
    off[0] = 0;
    for (int k = 0; k < 22; ++k) {
      s += (int32_t)lcount[k];
      off[k + 1] = s;
    }
    scal[0] = max(scal[3], 0) + 1;
    scal[1] = (int32_t)uscan[M];
  }
}

__global__ void k_oct_parent(const uint64_t* __restrict__ codes, const int32_t* __restrict__ depths,
                             const uint8_t* __restrict__ valid, const uint64_t* __restrict__ oaddr,
                             int M, int32_t* __restrict__ oparent, uint32_t* __restrict__ cplus) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  int32_t par = -1;
  if (valid[k]) {
    int d = depths[k];
    uint64_t c = codes[k];
    for (int step = 1; step <= 21; ++step) {
      if (d < step) break;
      int pd = d - step;
      uint64_t pa = oct_address(prefix_code(c, pd), pd);
      int pos = fixed_round_search(oaddr, M, pa, true);
      int safe = min(pos, M - 1);
      if (oaddr[safe] == pa) {
        par = pos;
        break;
      }
    }
    if (par >= 0) atomicMax(&cplus[(int64_t)par * 8 + octant_at_depth(c, d)], (uint32_t)(k + 1));
  }
  oparent[k] = par;
}

__global__ void k_oct_children(const uint32_t* __restrict__ cplus,
                               const uint8_t* __restrict__ valid, int M,
                               int32_t* __restrict__ children, uint8_t* __restrict__ cmask,
                               int32_t* __restrict__ ccount, uint8_t* __restrict__ leaf_mask,
                               int32_t* __restrict__ leaf_flag) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  bool v = valid[k];
  int cnt = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    int32_t c = (int32_t)cplus[(int64_t)k * 8 + o] - 1;
    children[(int64_t)k * 8 + o] = c;
    bool m = v && c >= 0;
    cmask[(int64_t)k * 8 + o] = m;
    cnt += m;
  }
  ccount[k] = cnt;
  bool lm = v && cnt == 0;
  leaf_mask[k] = lm;
  leaf_flag[k] = lm;
}

__global__ void k_oct_leaf_nodes(const uint8_t* __restrict__ leaf_mask,
                                 const int64_t* __restrict__ pos, int M,
                                 int32_t* __restrict__ leaf_nodes, int32_t* __restrict__ scal) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  int64_t total = pos[M];
  if (leaf_mask[k]) leaf_nodes[pos[k]] = k;
  if (k >= total) leaf_nodes[k] = -1;
  if (k == 0) scal[2] = (int32_t)total;
}

__global__ void k_radix_to_oct(const uint64_t* __restrict__ sc, const int32_t* __restrict__ ranges,
                               const uint64_t* __restrict__ rc, const int32_t* __restrict__ rd,
                               int NI, int T, const uint64_t* __restrict__ oaddr, int M,
                               int32_t* __restrict__ r2o, uint32_t* __restrict__ leaf_key) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= T) return;
  int d;
  uint64_t c;
  if (k >= NI) {
    d = rd[k - NI];
    c = rc[k - NI];
  } else {
    uint64_t f = sc[ranges[2 * k]], l = sc[ranges[2 * k + 1]];
    d = common_depth(f, l);
    c = prefix_code(f, d);
  }
  int pos = fixed_round_search(oaddr, M, oct_address(c, d), true);
  r2o[k] = pos;
  if (k >= NI) leaf_key[k - NI] = (uint32_t)pos;
}

__global__ void k_leaf_order(const uint32_t* __restrict__ native, int NI, int L, int T,
                             int32_t* __restrict__ leaf_indices,
                             int32_t* __restrict__ leaf_position) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < T) leaf_position[k] = -1;
  (void)L;
}

__global__ void k_leaf_order2(const uint32_t* __restrict__ native, int NI, int L,
                              int32_t* __restrict__ leaf_indices,
                              int32_t* __restrict__ leaf_position) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L) return;
  int32_t node = NI + (native ? (int32_t)native[k] : k);
  leaf_indices[k] = node;
  leaf_position[node] = k;
}

}  // namespace ygg

using namespace ygg;

extern "C" {

size_t ygg_bounds_workspace_bytes(void) { return 256; }

int ygg_infer_bounds(const float* pos, int64_t n, float* bounds6, void* ws, size_t ws_bytes,
                     void* stream) {
  if (n < 1 || !pos || !bounds6) return YGG_ERR_ARG;
  if (ws_bytes < 24) return YGG_ERR_WORKSPACE;
  hipStream_t st = (hipStream_t)stream;
  uint32_t* acc = (uint32_t*)ws;
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, acc);
  if (((uintptr_t)pos & 15) == 0) {
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n / 4 + 255) / 256, 1024));
    hipLaunchKernelGGL(k_bounds_reduce4, dim3(blocks), dim3(256), 0, st, pos, n, acc);
  } else {
    int blocks = (int)std::min<int64_t>(((3 * n + 255) / 256 + 2) / 3 * 3, 3 * 512);
    if (blocks < 3) blocks = 3;
    hipLaunchKernelGGL(k_bounds_reduce, dim3(blocks), dim3(256), 0, st, pos, 3 * n, acc);
  }
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, st, acc, bounds6);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_morton_encode(const float* pos, int64_t n, const float* bounds6, uint64_t* codes,
                      void* stream) {
  if (n < 0) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipLaunchKernelGGL(k_encode, dim3(grid_for(n, 256, 8192)), dim3(256), 0, (hipStream_t)stream,
                     pos, n, bounds6, codes);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_morton_sort_workspace_bytes(int64_t n) {
  return align_up(sizeof(uint64_t) * (size_t)n) + align_up(sizeof(uint32_t) * kMortonHistWords) +
         ygg::sort_workspace_bytes(n, 8, kMortonEndBit, kMortonDigitBits);
}

int ygg_morton_sort(const float* pos, int64_t n, const float* bounds6, uint64_t* sorted_codes,
                    uint32_t* perm, void* ws, size_t ws_bytes, void* stream) {
  if (n < 1) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  Arena a(ws, ws_bytes);
  uint64_t* codes = a.take<uint64_t>((size_t)n);
  uint32_t* hist = a.take<uint32_t>(kMortonHistWords);
  size_t rest = ygg::sort_workspace_bytes(n, 8, kMortonEndBit, kMortonDigitBits);
  char* sws = a.take<char>(rest);
  if (!codes || !hist || !sws) return YGG_ERR_WORKSPACE;
  int rc = fill_u32(hist, 0u, kMortonHistWords, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_encode_hist, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, pos, n,
                     bounds6, codes, hist);
  YGG_LAUNCH_CHECK();
  return sort_pairs_u64(codes, nullptr, sorted_codes, perm, n, 0, kMortonEndBit, sws, rest, st, hist,
                        kMortonDigitBits);
}

// ---- ygg_morton_sort_noalloc: the same encode + sort with its workspace in
// a file-scope device array, for callers that cannot allocate. The array's
// name is a host-side stand-in: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device on that device's first call
// and kept, and only that address is handed to kernels.
static __device__ __align__(256) unsigned char scratch[YGG_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[YGG_NOALLOC_MAX_DEVICES];

static int scratch_for_current_device(void** out) {
  int dev = -1;
  YGG_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= YGG_NOALLOC_MAX_DEVICES) return YGG_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    YGG_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return YGG_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return YGG_OK;
}

int ygg_morton_sort_noalloc(const float* pos, int64_t n, const float* bounds6,
                            uint64_t* sorted_codes, uint32_t* perm, void* stream) {
  if (n < 1 || !pos || !bounds6 || !sorted_codes || !perm) return YGG_ERR_ARG;
  if (n > YGG_NOALLOC_MAX_POINTS) return YGG_ERR_TOO_LARGE;
  const size_t need = ygg_morton_sort_workspace_bytes(n);
  if (need > (size_t)YGG_NOALLOC_SCRATCH_BYTES) return YGG_ERR_TOO_LARGE;
  void* ws = nullptr;
  int rc = scratch_for_current_device(&ws);
  if (rc) return rc;
  return ygg_morton_sort(pos, n, bounds6, sorted_codes, perm, ws, need, stream);
}

size_t ygg_morton_sort_reorder_workspace_bytes(int64_t n) {
  return align_up(sizeof(float4) * (size_t)n) + ygg_morton_sort_workspace_bytes(n);
}

// Encode + histograms + pack, then the code sort, on `st`; returns the
// packed records' address in the workspace through *packed_out.
static int morton_sort_pack(const float* pos, const float* mass, int64_t n, const float* bounds6,
                            uint64_t* sorted_codes, uint32_t* perm, void* ws, size_t ws_bytes,
                            hipStream_t st, const float4** packed_out) {
  Arena a(ws, ws_bytes);
  float4* packed = a.take<float4>((size_t)n);
  uint64_t* codes = a.take<uint64_t>((size_t)n);
  uint32_t* hist = a.take<uint32_t>(kMortonHistWords);
  size_t rest = ygg::sort_workspace_bytes(n, 8, kMortonEndBit, kMortonDigitBits);
  char* sws = a.take<char>(rest);
  if (!packed || !codes || !hist || !sws) return YGG_ERR_WORKSPACE;
  int rc = fill_u32(hist, 0u, kMortonHistWords, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_encode_hist_pack, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, pos,
                     mass, n, bounds6, codes, hist, packed);
  YGG_LAUNCH_CHECK();
  if ((rc = sort_pairs_u64(codes, nullptr, sorted_codes, perm, n, 0, kMortonEndBit, sws, rest, st, hist,
                        kMortonDigitBits)))
    return rc;
  *packed_out = packed;
  return YGG_OK;
}

static int reorder_from_packed(const float4* packed, const uint32_t* perm, int64_t n,
                               float* pos_sorted, float* mass_sorted, void* inverse,
                               int inverse_i64, hipStream_t st) {
  hipLaunchKernelGGL(k_reorder_packed, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, packed,
                     perm, n, pos_sorted, mass_sorted);
  if (inverse && inverse_i64)
    hipLaunchKernelGGL(k_inverse_perm<int64_t>, dim3(grid_for(n, 256)), dim3(256), 0, st, perm, n,
                       (int64_t*)inverse);
  else if (inverse)
    hipLaunchKernelGGL(k_inverse_perm<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, st, perm, n,
                       (int32_t*)inverse);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_morton_sort_reorder(const float* pos, const float* mass, int64_t n, const float* bounds6,
                            uint64_t* sorted_codes, uint32_t* perm, float* pos_sorted,
                            float* mass_sorted, void* inverse, int inverse_i64, void* ws,
                            size_t ws_bytes, void* stream) {
  if (n < 1 || !pos || !pos_sorted) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const float4* packed = nullptr;
  int rc = morton_sort_pack(pos, mass, n, bounds6, sorted_codes, perm, ws, ws_bytes, st, &packed);
  if (rc) return rc;
  return reorder_from_packed(packed, perm, n, pos_sorted, mass ? mass_sorted : nullptr, inverse,
                             inverse_i64, st);
}

int ygg_morton_sort_reorder_async(const float* pos, const float* mass, int64_t n,
                                  const float* bounds6, uint64_t* sorted_codes, uint32_t* perm,
                                  float* pos_sorted, float* mass_sorted, void* inverse,
                                  int inverse_i64, void* ws, size_t ws_bytes, void* stream,
                                  void* reorder_stream) {
  if (n < 1 || !pos || !pos_sorted || !reorder_stream) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream, rst = (hipStream_t)reorder_stream;
  const float4* packed = nullptr;
  int rc = morton_sort_pack(pos, mass, n, bounds6, sorted_codes, perm, ws, ws_bytes, st, &packed);
  if (rc) return rc;
  hipEvent_t sorted;
  YGG_CHECK(hipEventCreateWithFlags(&sorted, hipEventDisableTiming));
  const hipError_t e1 = hipEventRecord(sorted, st);
  const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(rst, sorted, 0) : e1;
  (void)hipEventDestroy(sorted);  // released once the recorded work completes
  if (e2 != hipSuccess) return YGG_ERR_HIP;
  return reorder_from_packed(packed, perm, n, pos_sorted, mass ? mass_sorted : nullptr, inverse,
                             inverse_i64, rst);
}

int ygg_widen_i32_to_i64(const YggWidenItem* items, int count, void* stream) {
  if (count < 0 || count > YGG_WIDEN_MAX_ITEMS || (count > 0 && !items)) return YGG_ERR_ARG;
  WidenArgs a = {};
  int blocks = 0;
  for (int k = 0; k < count; ++k) {
    const YggWidenItem& it = items[k];
    if (it.n < 0 || (it.n > 0 && (!it.src || !it.dst))) return YGG_ERR_ARG;
    if (it.n == 0) continue;
    a.src[a.count] = it.src;
    a.dst[a.count] = it.dst;
    a.n[a.count] = it.n;
    a.block_start[a.count] = blocks;
    blocks += (int)std::min<int64_t>((it.n + 255) / 256, 2048);
    ++a.count;
  }
  if (a.count == 0) return YGG_OK;
  a.block_start[a.count] = blocks;
  hipLaunchKernelGGL(k_widen_i32, dim3(blocks), dim3(256), 0, (hipStream_t)stream, a);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_reorder(const float* pos, const float* mass, const uint32_t* perm, int64_t n,
                float* pos_sorted, float* mass_sorted, int32_t* inverse, void* stream) {
  if (n < 0) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipLaunchKernelGGL(k_reorder, dim3(grid_for(n, 256, 8192)), dim3(256), 0, (hipStream_t)stream,
                     pos, mass, perm, n, pos_sorted, mass_sorted, inverse);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_reorder_scatter(const float* pos, const float* mass, const uint32_t* perm, int64_t n,
                        float* pos_sorted, float* mass_sorted, int32_t* inverse, void* stream) {
  if (n < 0 || !inverse || !pos_sorted) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_inverse, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, perm, n, inverse);
  hipLaunchKernelGGL(k_scatter_sorted, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, pos, mass,
                     inverse, n, pos_sorted, mass_sorted);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_radix_build_workspace_bytes(int64_t n, int64_t leaf_size) {
  if (leaf_size < 1 || n < 1) return 0;
  int64_t L = (n + leaf_size - 1) / leaf_size;
  int64_t T = 2 * L - 1;
  Arena a(nullptr, 0);
  a.take<uint32_t>((size_t)T);  // parent encodings
  a.take<uint32_t>(64);         // level counts
  a.take<uint32_t>((size_t)T);  // sorted level keys
  a.take<char>(ygg::sort_workspace_bytes(T, 4, 32));
  a.take<int32_t>((size_t)T * 2);  // pointer doubling sc
  a.take<int64_t>((size_t)T * 2);  // pointer doubling dist
  return a.off;
}

int ygg_radix_build(const uint64_t* sorted_codes, int64_t n, int64_t leaf_size,
                    const YggRadixTopology* o, void* ws, size_t ws_bytes, void* stream) {
  if (n < 1 || leaf_size < 1 || !o) return YGG_ERR_ARG;
  if (n > 0x7FFFFFFFLL) return YGG_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int64_t L64 = (n + leaf_size - 1) / leaf_size;
  const int L = (int)L64;
  const int T = 2 * L - 1;
  const int NI = L - 1;
  int rc = fill_u32((uint32_t*)o->scalars, 0u, 8, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_leaves, dim3(grid_for(L, 256)), dim3(256), 0, st, sorted_codes, n,
                     leaf_size, L, o->leaf_codes, o->node_ranges);
  YGG_LAUNCH_CHECK();
  if (L == 1) {
    hipLaunchKernelGGL(k_single_leaf, dim3(1), dim3(64), 0, st, n, o->parent, o->node_ranges,
                       o->node_level, o->level_offsets, o->nodes_by_level, o->scalars);
    YGG_LAUNCH_CHECK();
    return YGG_OK;
  }
  Arena a(ws, ws_bytes);
  uint32_t* penc = a.take<uint32_t>((size_t)T);
  uint32_t* lcount = a.take<uint32_t>(64);
  uint32_t* skeys = a.take<uint32_t>((size_t)T);
  size_t sbytes = ygg::sort_workspace_bytes(T, 4, 32);
  char* sws = a.take<char>(sbytes);
  if (!penc || !lcount || !skeys || !sws) return YGG_ERR_WORKSPACE;
  if ((rc = fill_u32(penc, 0u, T, st))) return rc;
  if ((rc = fill_u32(lcount, 0u, 64, st))) return rc;
  hipLaunchKernelGGL(k_karras, dim3(grid_for(NI, 256)), dim3(256), 0, st, o->leaf_codes, L, n,
                     leaf_size, o->left_child, o->right_child, o->left_is_leaf, o->right_is_leaf,
                     o->node_ranges, penc);
  YGG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_levels_walk, dim3(grid_for(T, 256)), dim3(256), 0, st, penc, T, NI,
                     o->left_child, o->right_child, o->parent, o->node_level, lcount, o->scalars);
  YGG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, st, lcount, o->level_offsets,
                     o->scalars);
  YGG_LAUNCH_CHECK();
  // nodes_by_level = stable argsort(node_level): one 8-bit counting pass
  // (levels >= 256 take the pointer-doubling path).
  return sort_pairs_u32((const uint32_t*)o->node_level, nullptr, skeys,
                        (uint32_t*)o->nodes_by_level, T, 0, 8, sws, sbytes, st);
}

int ygg_levels_pointer_doubling(const YggRadixTopology* o, int64_t num_nodes, void* ws,
                                size_t ws_bytes, void* stream) {
  if (!o || num_nodes < 1) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int T = (int)num_nodes;
  Arena a(ws, ws_bytes);
  uint32_t* penc = a.take<uint32_t>((size_t)T);
  uint32_t* lcount = a.take<uint32_t>(64);
  uint32_t* skeys = a.take<uint32_t>((size_t)T);
  size_t sbytes = ygg::sort_workspace_bytes(T, 4, 32);
  char* sws = a.take<char>(sbytes);
  int32_t* sc = a.take<int32_t>((size_t)T * 2);
  int64_t* dist = a.take<int64_t>((size_t)T * 2);
  if (!penc || !lcount || !skeys || !sws || !sc || !dist) return YGG_ERR_WORKSPACE;
  int rc;
  if ((rc = fill_u32(lcount, 0u, 64, st))) return rc;
  if ((rc = fill_u32((uint32_t*)&o->scalars[1], 0u, 1, st))) return rc;
  int g = grid_for(T, 256);
  hipLaunchKernelGGL(k_pd_init, dim3(g), dim3(256), 0, st, o->parent, T, sc, dist);
  int bl = 0;
  for (int64_t v = T; v > 0; v >>= 1) ++bl;
  int rounds = bl + 1 > 1 ? bl + 1 : 1;
  int cur = 0;
  for (int r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(k_pd_round, dim3(g), dim3(256), 0, st, sc + cur * T, dist + cur * T, T,
                       sc + (1 - cur) * T, dist + (1 - cur) * T);
    cur = 1 - cur;
  }
  hipLaunchKernelGGL(k_pd_finish, dim3(g), dim3(256), 0, st, dist + cur * T, T, o->node_level,
                     lcount, o->scalars);
  hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, st, lcount, o->level_offsets,
                     o->scalars);
  YGG_LAUNCH_CHECK();
  return sort_pairs_u32((const uint32_t*)o->node_level, nullptr, skeys,
                        (uint32_t*)o->nodes_by_level, T, 0, 32, sws, sbytes, st);
}

size_t ygg_octree_meta_workspace_bytes(int64_t L) {
  int64_t M = 2 * L - 1 > 1 ? 2 * L - 1 : 1;
  int64_t T = 2 * L - 1;
  Arena a(nullptr, 0);
  a.take<uint64_t>((size_t)L);      // resolved codes
  a.take<int32_t>((size_t)L);       // resolved depths
  a.take<uint64_t>((size_t)M);      // candidate addresses
  a.take<uint64_t>((size_t)M);      // sorted addresses
  a.take<uint32_t>((size_t)M);      // order
  a.take<int32_t>((size_t)M);       // flags
  a.take<int64_t>((size_t)M + 1);   // scan
  a.take<uint64_t>((size_t)M);      // oct addresses
  a.take<uint32_t>((size_t)M);      // sort keys
  a.take<uint32_t>((size_t)M);      // sort tmp keys
  a.take<uint32_t>((size_t)M * 8);  // children+1
  a.take<uint32_t>(32);             // level counts
  a.take<uint32_t>((size_t)L);      // leaf keys
  a.take<uint32_t>((size_t)L);      // native order
  size_t s1 = ygg::sort_workspace_bytes(M, 8, 64);
  size_t s2 = ygg::sort_workspace_bytes(M > T ? M : T, 4, 32);
  a.take<char>(s1 > s2 ? s1 : s2);
  a.take<char>(ygg::scan_workspace_bytes(M));
  return a.off;
}

int ygg_octree_meta(const uint64_t* sc, const int32_t* node_ranges, int64_t L64,
                    const YggOctreeMeta* o, void* ws, size_t ws_bytes, void* stream) {
  if (L64 < 1 || !o) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const int L = (int)L64;
  const int M = 2 * L - 1 > 1 ? 2 * L - 1 : 1;
  const int T = 2 * L - 1;
  const int NI = L - 1;
  Arena a(ws, ws_bytes);
  uint64_t* rc = a.take<uint64_t>((size_t)L);
  int32_t* rd = a.take<int32_t>((size_t)L);
  uint64_t* caddr = a.take<uint64_t>((size_t)M);
  uint64_t* saddr = a.take<uint64_t>((size_t)M);
  uint32_t* order = a.take<uint32_t>((size_t)M);
  int32_t* flags = a.take<int32_t>((size_t)M);
  int64_t* uscan = a.take<int64_t>((size_t)M + 1);
  uint64_t* oaddr = a.take<uint64_t>((size_t)M);
  uint32_t* skey = a.take<uint32_t>((size_t)M);
  uint32_t* stmp = a.take<uint32_t>((size_t)M);
  uint32_t* cplus = a.take<uint32_t>((size_t)M * 8);
  uint32_t* lcount = a.take<uint32_t>(32);
  uint32_t* lkey = a.take<uint32_t>((size_t)L);
  uint32_t* native = a.take<uint32_t>((size_t)L);
  size_t s1 = ygg::sort_workspace_bytes(M, 8, 64);
  size_t s2 = ygg::sort_workspace_bytes(M > T ? M : T, 4, 32);
  size_t sb = s1 > s2 ? s1 : s2;
  char* sws = a.take<char>(sb);
  size_t scb = ygg::scan_workspace_bytes(M);
  char* scws = a.take<char>(scb);
  if (!rc || !rd || !caddr || !saddr || !order || !flags || !uscan || !oaddr || !skey || !stmp ||
      !cplus || !lcount || !lkey || !native || !sws || !scws)
    return YGG_ERR_WORKSPACE;
  int rcode;
  if ((rcode = fill_u32((uint32_t*)o->scalars, 0u, 4, st))) return rcode;
  if ((rcode = fill_u32(lcount, 0u, 32, st))) return rcode;
  if ((rcode = fill_u32(cplus, 0u, (int64_t)M * 8, st))) return rcode;
  const int gL = grid_for(L, 256), gM = grid_for(M, 256), gT = grid_for(T, 256);
  hipLaunchKernelGGL(k_oct_resolve, dim3(gL), dim3(256), 0, st, sc, node_ranges, L, rc, rd);
  hipLaunchKernelGGL(k_oct_candidates, dim3(gM), dim3(256), 0, st, rc, rd, L, M, caddr);
  YGG_LAUNCH_CHECK();
  if ((rcode = sort_pairs_u64(caddr, nullptr, saddr, order, M, 0, 64, sws, sb, st))) return rcode;
  hipLaunchKernelGGL(k_oct_newflags, dim3(gM), dim3(256), 0, st, saddr, M, flags);
  YGG_LAUNCH_CHECK();
  if ((rcode = exclusive_scan_i32_to_i64(flags, nullptr, uscan, M, scws, scb, st))) return rcode;
  hipLaunchKernelGGL(k_oct_unique, dim3(gM), dim3(256), 0, st, saddr, order, rc, rd, L, M, uscan,
                     o->oct_node_codes, o->oct_node_depths, o->oct_valid_mask, oaddr);
  hipLaunchKernelGGL(k_oct_ranges, dim3(gM), dim3(256), 0, st, o->oct_node_codes,
                     o->oct_node_depths, o->oct_valid_mask, rc, node_ranges, L, M,
                     o->oct_node_ranges, lcount, skey, o->scalars);
  YGG_LAUNCH_CHECK();
  if ((rcode = sort_pairs_u32(skey, nullptr, stmp, (uint32_t*)o->oct_nodes_by_level, M, 0, 8, sws,
                              sb, st)))
    return rcode;
  hipLaunchKernelGGL(k_oct_level_offsets, dim3(1), dim3(64), 0, st, lcount,
                     o->oct_level_offsets, o->scalars, uscan, M);
  hipLaunchKernelGGL(k_oct_parent, dim3(gM), dim3(256), 0, st, o->oct_node_codes,
                     o->oct_node_depths, o->oct_valid_mask, oaddr, M, o->oct_parent, cplus);
  hipLaunchKernelGGL(k_oct_children, dim3(gM), dim3(256), 0, st, cplus, o->oct_valid_mask, M,
                     o->oct_children, o->oct_child_mask, o->oct_child_counts, o->oct_leaf_mask,
                     flags);
  YGG_LAUNCH_CHECK();
  if ((rcode = exclusive_scan_i32_to_i64(flags, nullptr, uscan, M, scws, scb, st))) return rcode;
  hipLaunchKernelGGL(k_oct_leaf_nodes, dim3(gM), dim3(256), 0, st, o->oct_leaf_mask, uscan, M,
                     o->oct_leaf_nodes, o->scalars);
  hipLaunchKernelGGL(k_radix_to_oct, dim3(gT), dim3(256), 0, st, sc, node_ranges, rc, rd, NI, T,
                     oaddr, M, o->radix_node_to_oct, lkey);
  YGG_LAUNCH_CHECK();
  int end_bit = 32 - __builtin_clz((unsigned)M);  // values in [0, M]
  if ((rcode = sort_pairs_u32(lkey, nullptr, stmp, native, L, 0, end_bit, sws, sb, st)))
    return rcode;
  hipLaunchKernelGGL(k_leaf_order, dim3(gT), dim3(256), 0, st, native, NI, L, T, o->leaf_indices,
                     o->leaf_position);
  hipLaunchKernelGGL(k_leaf_order2, dim3(gL), dim3(256), 0, st, native, NI, L, o->leaf_indices,
                     o->leaf_position);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

// Radix trees keep particle-order leaves: leaf_indices[k] = L-1+k.
int ygg_radix_leaf_order(int64_t num_leaves, int32_t* leaf_indices, int32_t* leaf_position,
                         void* stream) {
  if (num_leaves < 1) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int L = (int)num_leaves, T = 2 * L - 1, NI = L - 1;
  hipLaunchKernelGGL(k_leaf_order, dim3(grid_for(T, 256)), dim3(256), 0, st, nullptr, NI, L, T,
                     leaf_indices, leaf_position);
  hipLaunchKernelGGL(k_leaf_order2, dim3(grid_for(L, 256)), dim3(256), 0, st, nullptr, NI, L,
                     leaf_indices, leaf_position);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
