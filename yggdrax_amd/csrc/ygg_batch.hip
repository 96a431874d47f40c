// Batched build: many independent particle systems in one set of launches.
//
// The systems' points are concatenated; a segment table (Seg) gives each
// system's point and leaf offsets. Every kernel here runs once over the
// whole batch and produces, for system s, exactly what the single-tree
// kernels produce for that system alone, in the slice of the concatenated
// arrays the layout assigns to it (all stored indices local to the system):
//   points  [pt[s], pt[s+1])          leaves   [leaf[s], leaf[s+1])
//   internal nodes [leaf[s]-s, ...)   nodes / octree slots [2 leaf[s]-s, ...)
//   level_offsets [65 s, 65 s + 65)   oct_level_offsets [23 s, ...)
// Sorting by (system, key) is a stable key sort followed by a stable sort on
// the system id, so ties keep input order exactly as in the single-tree sort.
#include "ygg_common.h"
#include "ygg_build_dev.h"

namespace ygg {

// ---- bounds: one block per system ------------------------------------------
__global__ __launch_bounds__(256) void k_b_bounds(Seg sg, const float* __restrict__ pos,
                                                  float* __restrict__ b6) {
  const int s = blockIdx.x;
  const int64_t p0 = sg.pt_off(s), n = sg.n(s);
  uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
  // four points per thread and step, all loaded before any is reduced, so
  // one workgroup keeps several loads in flight on a large system
  constexpr int kU = 4;
  for (int64_t i0 = threadIdx.x; i0 < n; i0 += kU * (int64_t)blockDim.x) {
    float v[kU][3];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = i0 + u * (int64_t)blockDim.x;
#pragma unroll
      for (int d = 0; d < 3; ++d) v[u][d] = i < n ? pos[3 * (p0 + i) + d] : pos[3 * p0 + d];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const uint32_t o = f2ord(v[u][d]);
        mn[d] = min(mn[d], o);
        mx[d] = max(mx[d], o);
      }
  }
  __shared__ uint32_t smn[3], smx[3];
  if (threadIdx.x < 3) {
    smn[threadIdx.x] = 0xFFFFFFFFu;
    smx[threadIdx.x] = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    atomicMin(&smn[d], mn[d]);
    atomicMax(&smx[d], mx[d]);
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    padded_bounds(ord2f(smn[d]), ord2f(smx[d]), &b6[6 * s + d], &b6[6 * s + 3 + d]);
  }
}

// ---- Morton encode with per-system bounds + the code sort's histograms ----
__global__ __launch_bounds__(256) void k_b_encode_hist(Seg sg, const float* __restrict__ pos,
                                                       int64_t N, const float* __restrict__ b6,
                                                       uint64_t* __restrict__ codes,
                                                       uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kMortonHistWords];
  __shared__ int32_t s_pt[kSegLdsCap];
  for (int i = threadIdx.x; i < kMortonHistWords; i += blockDim.x) h[i] = 0;
  const bool staged = seg_stage(s_pt, sg.S, [&](int s) { return (int32_t)sg.pt_off(s); });
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < N; b0 += stride) {
    const int64_t i = b0 + threadIdx.x;
    const bool valid = i < N;
    uint64_t c = 0;
    if (valid) {
      const int s = staged ? seg_find_lds(s_pt, sg.S, i) : sg.of_point(i);
      const float* b = b6 + 6 * s;
      float lo[3] = {b[0], b[1], b[2]};
      float w[3] = {b[3] - b[0], b[4] - b[1], b[5] - b[2]};
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

// key[i] = system of the point at sorted position i (a point index).
__global__ __launch_bounds__(256) void k_b_point_system(Seg sg, const uint32_t* __restrict__ idx,
                                                        int64_t N, uint32_t* __restrict__ key) {
  __shared__ int32_t s_pt[kSegLdsCap];
  const bool staged = seg_stage(s_pt, sg.S, [&](int s) { return (int32_t)sg.pt_off(s); });
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = idx[i];
    key[i] = (uint32_t)(staged ? seg_find_lds(s_pt, sg.S, g) : sg.of_point(g));
  }
}

__global__ void k_gather_u64(const uint64_t* __restrict__ src, const uint32_t* __restrict__ idx,
                             int64_t n, uint64_t* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[idx[i]];
}

// Gather positions and masses; local permutation and local inverse.
__global__ __launch_bounds__(256) void k_b_reorder(Seg sg, const float* __restrict__ pos,
                                                   const float* __restrict__ mass,
                                                   const uint32_t* __restrict__ perm, int64_t N,
                                                   float* __restrict__ pos_sorted,
                                                   float* __restrict__ mass_sorted,
                                                   int32_t* __restrict__ perm_local,
                                                   int32_t* __restrict__ inv_local) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t p = perm[i];
    const int64_t base = sg.pt_off(sg.of_point(i));
    if (pos_sorted) {
      pos_sorted[3 * i] = pos[3 * (int64_t)p];
      pos_sorted[3 * i + 1] = pos[3 * (int64_t)p + 1];
      pos_sorted[3 * i + 2] = pos[3 * (int64_t)p + 2];
    }
    if (mass && mass_sorted) mass_sorted[i] = mass[p];
    if (perm_local) perm_local[i] = (int32_t)(p - base);
    if (inv_local) inv_local[p] = (int32_t)(i - base);
  }
}

// ---- Karras build -------------------------------------------------------------
__global__ __launch_bounds__(256) void k_b_leaves(Seg sg, const uint64_t* __restrict__ sc,
                                                  int64_t Ltot, uint64_t* __restrict__ lc,
                                                  int32_t* __restrict__ ranges) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= Ltot) return;
  const int s = sg.of_leaf(j);
  const int64_t jl = j - sg.leaf_off(s);
  const int64_t n = sg.n(s);
  const int64_t st = jl * sg.leaf_size;
  const int64_t e = min(st + sg.leaf_size, n);
  lc[j] = sc[sg.pt_off(s) + st];
  const int64_t node = sg.node_off(s) + (sg.L(s) - 1) + jl;
  ranges[2 * node] = (int32_t)st;
  ranges[2 * node + 1] = (int32_t)(e - 1);
}

__global__ __launch_bounds__(256) void k_b_karras(Seg sg, const uint64_t* __restrict__ lc,
                                                  int64_t NItot, int32_t* __restrict__ left,
                                                  int32_t* __restrict__ right,
                                                  uint8_t* __restrict__ lflag,
                                                  uint8_t* __restrict__ rflag,
                                                  int32_t* __restrict__ ranges,
                                                  uint32_t* __restrict__ penc) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= NItot) return;
  const int s = sg.of_internal(g);
  const int ni0 = sg.ni_off(s), n0 = sg.node_off(s);
  karras_node(lc + sg.leaf_off(s), sg.L(s), sg.n(s), sg.leaf_size, (int)(g - ni0), left + ni0,
              right + ni0, lflag + ni0, rflag + ni0, ranges + 2 * (int64_t)n0, penc + n0);
}

// Levels by walking up (capped at 255), per-system level histograms and
// scalars [8 s + k]: 1 max_level, 2 deep, 3 malformed, 4 roots. A
// system's nodes are contiguous, so most wavefronts lie in one system:
// those reduce the max level and the root count over their lanes and add
// them with one atomic each, instead of one atomic per node on the same
// per-system word.
__global__ __launch_bounds__(256) void k_b_levels(Seg sg, const uint32_t* __restrict__ penc,
                                                  int64_t Ttot, const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right,
                                                  int32_t* __restrict__ parent,
                                                  int32_t* __restrict__ level,
                                                  uint32_t* __restrict__ lcount,
                                                  int32_t* __restrict__ scal) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool valid = g < Ttot;
  int s = 0, d = 0;
  bool root = false;
  if (valid) {
    s = sg.of_node(g);
    const int n0 = sg.node_off(s), ni0 = sg.ni_off(s);
    const int NI = sg.L(s) - 1;
    const int k = (int)(g - n0);
    const uint32_t* pe = penc + n0;
    const int32_t p = dec_parent(pe[k]);
    parent[g] = p;
    int32_t q = p;
    bool deep = false;
    while (q >= 0) {
      ++d;
      if (d > 255) {
        deep = true;
        break;
      }
      q = dec_parent(pe[q]);
    }
    level[g] = d;
    int32_t* sc = scal + 8 * (int64_t)s;
    if (deep) atomicOr((unsigned*)&sc[2], 1u);
    if (d < 64) atomicAdd(&lcount[64 * (int64_t)s + d], 1u);
    root = p < 0;
    if (k < NI) {
      if (dec_parent(pe[left[ni0 + k]]) != k || dec_parent(pe[right[ni0 + k]]) != k)
        atomicOr((unsigned*)&sc[3], 1u);
    }
  }
  // every lane reaches the ballots (no early return above)
  const uint64_t vmask = __ballot(valid);
  if (vmask == 0) return;
  const int lead = __ffsll((unsigned long long)vmask) - 1;
  const int s0 = __shfl(s, lead, 64);
  if (__ballot(valid && s == s0) == vmask) {
    int dm = valid ? d : 0, rc = (valid && root) ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      dm = max(dm, __shfl_xor(dm, off, 64));
      rc += __shfl_xor(rc, off, 64);
    }
    if ((int)lane_id() == lead) {
      atomicMax(&scal[8 * (int64_t)s0 + 1], dm);
      if (rc) atomicAdd(&scal[8 * (int64_t)s0 + 4], rc);
    }
  } else if (valid) {
    atomicMax(&scal[8 * (int64_t)s + 1], d);
    if (root) atomicAdd(&scal[8 * (int64_t)s + 4], 1);
  }
}

// Per system: level_offsets [65] (cumulative; a one-leaf tree keeps the
// reference's [0, 1, 0, ...]) and num_levels; level-sort keys.
__global__ void k_b_level_offsets(Seg sg, const uint32_t* __restrict__ lcount,
                                  int32_t* __restrict__ off, int32_t* __restrict__ scal) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sg.S) return;
  int32_t* o = off + 65 * (int64_t)s;
  int32_t* sc = scal + 8 * (int64_t)s;
  if (sg.L(s) == 1) {
    for (int k = 0; k < 65; ++k) o[k] = 0;
    o[1] = 1;
    sc[0] = 1;
    sc[5] = 0;
    return;
  }
  int32_t run = 0;
  o[0] = 0;
  for (int k = 0; k < 64; ++k) {
    run += (int32_t)lcount[64 * (int64_t)s + k];
    o[k + 1] = run;
  }
  int nl = sc[1] + 1;
  sc[0] = nl < 1 ? 1 : (nl > 64 ? 64 : nl);
  sc[5] = 0;
}

__global__ void k_b_level_keys(Seg sg, const int32_t* __restrict__ level, int64_t Ttot,
                               int shift, uint32_t* __restrict__ key) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Ttot;
       g += (int64_t)gridDim.x * blockDim.x)
    key[g] = ((uint32_t)sg.of_node(g) << shift) | (uint32_t)level[g];
}

// Sorted global node ids -> ids local to their system (position g lies in
// the system's node slice because the sort key leads with the system).
__global__ void k_b_localize_nodes(Seg sg, int32_t* __restrict__ v, int64_t Ttot) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < Ttot;
       g += (int64_t)gridDim.x * blockDim.x)
    v[g] -= sg.node_off(sg.of_node(g));
}

// ---- explicit octree metadata ---------------------------------------------------
__global__ void k_b_oct_resolve(Seg sg, const uint64_t* __restrict__ sc,
                                const int32_t* __restrict__ ranges, int64_t Ltot,
                                uint64_t* __restrict__ rc, int32_t* __restrict__ rd) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= Ltot) return;
  const int s = sg.of_leaf(j);
  const int64_t node = sg.node_off(s) + (sg.L(s) - 1) + (j - sg.leaf_off(s));
  const uint64_t* c = sc + sg.pt_off(s);
  const uint64_t f = c[ranges[2 * node]], l = c[ranges[2 * node + 1]];
  const int d = common_depth(f, l);
  rd[j] = d;
  rc[j] = prefix_code(f, d);
}

__global__ void k_b_oct_candidates(Seg sg, const uint64_t* __restrict__ rc,
                                   const int32_t* __restrict__ rd, int64_t Mtot,
                                   uint64_t* __restrict__ addr) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int s = sg.of_node(g);
  const int l0 = sg.leaf_off(s);
  uint64_t c;
  int d;
  candidate(rc + l0, rd + l0, sg.L(s), (int)(g - sg.node_off(s)), &c, &d);
  addr[g] = oct_address(c, d);
}

__global__ void k_b_node_system(Seg sg, const uint32_t* __restrict__ idx, int64_t n,
                                uint32_t* __restrict__ key) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    key[i] = (uint32_t)sg.of_node(idx[i]);
}

__global__ void k_b_oct_newflags(Seg sg, const uint64_t* __restrict__ sa, int64_t Mtot,
                                 int32_t* __restrict__ f) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int s = sg.of_node(g);
  f[g] = (g == sg.node_off(s) || sa[g] != sa[g - 1]) ? 1 : 0;
}

__global__ void k_b_oct_unique(Seg sg, const uint64_t* __restrict__ sa,
                               const uint32_t* __restrict__ order,
                               const uint64_t* __restrict__ rc, const int32_t* __restrict__ rd,
                               int64_t Mtot, const int64_t* __restrict__ uscan,
                               uint64_t* __restrict__ codes, int32_t* __restrict__ depths,
                               uint8_t* __restrict__ valid, uint64_t* __restrict__ oaddr) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int s = sg.of_node(g);
  const int64_t m0 = sg.node_off(s), m1 = sg.node_off(s + 1);
  const int64_t k = g - m0;
  const int64_t nu = uscan[m1] - uscan[m0];
  const bool leader = (g == m0) || sa[g] != sa[g - 1];
  if (leader) {
    const int64_t u = uscan[g] - uscan[m0];
    const int l0 = sg.leaf_off(s);
    uint64_t c;
    int d;
    candidate(rc + l0, rd + l0, sg.L(s), (int)(order[g] - m0), &c, &d);
    codes[m0 + u] = c;
    depths[m0 + u] = d;
    oaddr[m0 + u] = sa[g];
  }
  const bool v = k < nu;
  valid[g] = v;
  if (!v) {
    codes[g] = 0;
    depths[g] = -1;
    oaddr[g] = ~0ull;
  }
}

// oct scalars [4 s + k]: 0 oct_num_levels, 1 num_valid, 2 num_leaf, 3 max depth
__global__ __launch_bounds__(256) void k_b_oct_ranges(
    Seg sg, const uint64_t* __restrict__ codes, const int32_t* __restrict__ depths,
    const uint8_t* __restrict__ valid, const uint64_t* __restrict__ rc,
    const int32_t* __restrict__ node_ranges, int64_t Mtot, int32_t* __restrict__ oranges,
    uint32_t* __restrict__ lcount, uint32_t* __restrict__ skey, int32_t* __restrict__ scal) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool in = g < Mtot;
  int s = 0, d = 0;
  bool v = false;
  if (in) {
    s = sg.of_node(g);
    const int L = sg.L(s);
    v = valid[g];
    d = depths[g];
    const uint64_t c = codes[g];
    const uint64_t du = (uint64_t)max(d, 0);
    const uint64_t bs = 63ull - 3ull * du;
    const uint64_t bss = bs < 62ull ? bs : 62ull;
    const uint64_t mask = (1ull << bss) - 1ull;
    const uint64_t endc = v ? (c | mask) : 0ull;
    const uint64_t* r = rc + sg.leaf_off(s);
    int f = fixed_round_search(r, L, c, true);
    int l = fixed_round_search(r, L, endc, false) - 1;
    const int bound = max(L - 1, 0);
    f = min(max(f, 0), bound);
    l = min(max(l, 0), bound);
    const int64_t n0 = sg.node_off(s);
    const int64_t NI = L - 1;
    if (v) {
      oranges[2 * g] = node_ranges[2 * (n0 + NI + f)];
      oranges[2 * g + 1] = node_ranges[2 * (n0 + NI + l) + 1];
      atomicAdd(&lcount[22 * (int64_t)s + min(max(d, 0), 21)], 1u);
    } else {
      oranges[2 * g] = 0;
      oranges[2 * g + 1] = -1;
    }
    skey[g] = ((uint32_t)s << 5) | (v ? (uint32_t)d : 22u);
  }
  // Max depth per system: reduced over the wavefront when all its valid
  // octree nodes lie in one system (one atomic), else one per node. Every
  // lane reaches the ballots.
  const bool contrib = in && v;
  const uint64_t cmask = __ballot(contrib);
  if (cmask == 0) return;
  const int lead = __ffsll((unsigned long long)cmask) - 1;
  const int s0 = __shfl(s, lead, 64);
  if (__ballot(contrib && s == s0) == cmask) {
    int dm = contrib ? d : INT32_MIN;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) dm = max(dm, __shfl_xor(dm, off, 64));
    if ((int)lane_id() == lead) atomicMax(&scal[4 * (int64_t)s0 + 3], dm);
  } else if (contrib) {
    atomicMax(&scal[4 * (int64_t)s + 3], d);
  }
}

__global__ void k_b_oct_level_offsets(Seg sg, const uint32_t* __restrict__ lcount,
                                      int32_t* __restrict__ off, int32_t* __restrict__ scal,
                                      const int64_t* __restrict__ uscan) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= sg.S) return;
  int32_t* o = off + 23 * (int64_t)s;
  int32_t run = 0;
  o[0] = 0;
  for (int k = 0; k < 22; ++k) {
    run += (int32_t)lcount[22 * (int64_t)s + k];
    o[k + 1] = run;
  }
  int32_t* sc = scal + 4 * (int64_t)s;
  sc[0] = max(sc[3], 0) + 1;
  sc[1] = (int32_t)(uscan[sg.node_off(s + 1)] - uscan[sg.node_off(s)]);
}

__global__ void k_b_oct_parent(Seg sg, const uint64_t* __restrict__ codes,
                               const int32_t* __restrict__ depths,
                               const uint8_t* __restrict__ valid,
                               const uint64_t* __restrict__ oaddr, int64_t Mtot,
                               int32_t* __restrict__ oparent, uint32_t* __restrict__ cplus) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int s = sg.of_node(g);
  const int64_t m0 = sg.node_off(s);
  const int M = 2 * sg.L(s) - 1;
  const uint64_t* oa = oaddr + m0;
  int32_t par = -1;
  if (valid[g]) {
    const int d = depths[g];
    const uint64_t c = codes[g];
    for (int step = 1; step <= 21; ++step) {
      if (d < step) break;
      const int pd = d - step;
      const uint64_t pa = oct_address(prefix_code(c, pd), pd);
      const int pos = fixed_round_search(oa, M, pa, true);
      const int safe = min(pos, M - 1);
      if (oa[safe] == pa) {
        par = pos;
        break;
      }
    }
    if (par >= 0)
      atomicMax(&cplus[(m0 + par) * 8 + octant_at_depth(c, d)], (uint32_t)(g - m0 + 1));
  }
  oparent[g] = par;
}

__global__ void k_b_oct_children(const uint32_t* __restrict__ cplus,
                                 const uint8_t* __restrict__ valid, int64_t Mtot,
                                 int32_t* __restrict__ children, uint8_t* __restrict__ cmask,
                                 int32_t* __restrict__ ccount, uint8_t* __restrict__ leaf_mask,
                                 int32_t* __restrict__ leaf_flag) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const bool v = valid[g];
  int cnt = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    const int32_t c = (int32_t)cplus[g * 8 + o] - 1;
    children[g * 8 + o] = c;
    const bool m = v && c >= 0;
    cmask[g * 8 + o] = m;
    cnt += m;
  }
  ccount[g] = cnt;
  const bool lm = v && cnt == 0;
  leaf_mask[g] = lm;
  leaf_flag[g] = lm;
}

__global__ void k_b_oct_leaf_nodes(Seg sg, const uint8_t* __restrict__ leaf_mask,
                                   const int64_t* __restrict__ pos, int64_t Mtot,
                                   int32_t* __restrict__ leaf_nodes, int32_t* __restrict__ scal) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Mtot) return;
  const int s = sg.of_node(g);
  const int64_t m0 = sg.node_off(s), m1 = sg.node_off(s + 1);
  const int64_t k = g - m0;
  const int64_t total = pos[m1] - pos[m0];
  if (leaf_mask[g]) leaf_nodes[m0 + (pos[g] - pos[m0])] = (int32_t)k;
  if (k >= total) leaf_nodes[g] = -1;
  if (k == 0) scal[4 * (int64_t)s + 2] = (int32_t)total;
}

// radix_node_to_oct (unchecked lower bound, octree.py:526-531) and the leaf
// sort keys (system, oct slot) of the native leaf order.
__global__ void k_b_radix_to_oct(Seg sg, const uint64_t* __restrict__ sc,
                                 const int32_t* __restrict__ ranges,
                                 const uint64_t* __restrict__ rc, const int32_t* __restrict__ rd,
                                 int64_t Ttot, const uint64_t* __restrict__ oaddr,
                                 uint32_t stride, int32_t* __restrict__ r2o,
                                 uint32_t* __restrict__ leaf_key) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= Ttot) return;
  const int s = sg.of_node(g);
  const int64_t n0 = sg.node_off(s);
  const int L = sg.L(s), NI = L - 1, M = 2 * L - 1;
  const int k = (int)(g - n0);
  int d;
  uint64_t c;
  if (k >= NI) {
    d = rd[sg.leaf_off(s) + (k - NI)];
    c = rc[sg.leaf_off(s) + (k - NI)];
  } else {
    const uint64_t* cs = sc + sg.pt_off(s);
    const uint64_t f = cs[ranges[2 * g]], l = cs[ranges[2 * g + 1]];
    d = common_depth(f, l);
    c = prefix_code(f, d);
  }
  const int pos = fixed_round_search(oaddr + n0, M, oct_address(c, d), true);
  r2o[g] = pos;
  if (k >= NI) leaf_key[sg.leaf_off(s) + (k - NI)] = (uint32_t)s * stride + (uint32_t)pos;
}

__global__ void k_b_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// native: global leaf indices in (system, oct slot) order, or null for the
// radix (particle) order. leaf_indices[j] = local leaf node id of row j of
// its system; leaf_position[node] = that row (local).
__global__ void k_b_leaf_order(Seg sg, const uint32_t* __restrict__ native, int64_t Ltot,
                               int32_t* __restrict__ leaf_indices,
                               int32_t* __restrict__ leaf_position) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= Ltot) return;
  const int s = sg.of_leaf(j);
  const int64_t l0 = sg.leaf_off(s);
  const int NI = sg.L(s) - 1;
  const int64_t src = native ? (int64_t)native[j] : j;
  const int32_t node = NI + (int32_t)(src - l0);
  leaf_indices[j] = node;
  leaf_position[sg.node_off(s) + node] = (int32_t)(j - l0);
}

}  // namespace ygg

using namespace ygg;

static inline Seg make_seg(const YggBatch* b) {
  Seg sg;
  sg.S = b->num_systems;
  sg.leaf_size = b->leaf_size;
  sg.pt = b->point_offsets;
  sg.leaf = b->leaf_offsets;
  return sg;
}

static inline int bits_for(uint64_t v) {  // bits to hold values in [0, v]
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b < 1 ? 1 : b;
}

static int check_batch(const YggBatch* b) {
  if (!b || b->num_systems < 1 || b->leaf_size < 1 || !b->point_offsets || !b->leaf_offsets)
    return YGG_ERR_ARG;
  if (b->num_points < b->num_systems || b->num_leaves < b->num_systems) return YGG_ERR_ARG;
  if (b->num_points > 0x7FFFFFFFLL || 2 * b->num_leaves > 0x7FFFFFFFLL) return YGG_ERR_UNSUPPORTED;
  if (b->num_systems >= (1 << 24)) return YGG_ERR_UNSUPPORTED;
  return YGG_OK;
}

extern "C" {

int ygg_batch_infer_bounds(const YggBatch* b, const float* pos, float* bounds6, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_bounds, dim3(b->num_systems), dim3(256), 0, (hipStream_t)stream,
                     make_seg(b), pos, bounds6);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_batch_morton_sort_workspace_bytes(const YggBatch* b) {
  if (check_batch(b)) return 0;
  const int64_t N = b->num_points;
  Arena a(nullptr, 0);
  a.take<uint64_t>((size_t)N);      // codes
  a.take<uint32_t>(kMortonHistWords);        // histograms
  a.take<uint64_t>((size_t)N);      // code-sorted codes
  a.take<uint32_t>((size_t)N);      // code-sorted indices
  a.take<uint32_t>((size_t)N);      // system keys
  a.take<uint32_t>((size_t)N);      // system keys (sorted)
  a.take<char>(ygg::sort_workspace_bytes(N, 8, kMortonEndBit, kMortonDigitBits));
  return a.off;
}

int ygg_batch_morton_sort(const YggBatch* b, const float* pos, const float* bounds6,
                          uint64_t* sorted_codes, uint32_t* perm, void* ws, size_t ws_bytes,
                          void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = b->num_points;
  const Seg sg = make_seg(b);
  Arena a(ws, ws_bytes);
  uint64_t* codes = a.take<uint64_t>((size_t)N);
  uint32_t* hist = a.take<uint32_t>(kMortonHistWords);
  uint64_t* c1 = a.take<uint64_t>((size_t)N);
  uint32_t* i1 = a.take<uint32_t>((size_t)N);
  uint32_t* sk = a.take<uint32_t>((size_t)N);
  uint32_t* sk2 = a.take<uint32_t>((size_t)N);
  size_t rest = ygg::sort_workspace_bytes(N, 8, kMortonEndBit, kMortonDigitBits);
  char* sws = a.take<char>(rest);
  if (!codes || !hist || !c1 || !i1 || !sk || !sk2 || !sws) return YGG_ERR_WORKSPACE;
  if ((rc = fill_u32(hist, 0u, kMortonHistWords, st))) return rc;
  hipLaunchKernelGGL(k_b_encode_hist, dim3(grid_for(N, 256, 4096)), dim3(256), 0, st, sg, pos, N,
                     bounds6, codes, hist);
  YGG_LAUNCH_CHECK();
  if ((rc = sort_pairs_u64(codes, nullptr, c1, i1, N, 0, kMortonEndBit, sws, rest, st, hist,
                            kMortonDigitBits))) return rc;
  if (b->num_systems == 1) {
    if ((rc = copy_u32((uint32_t*)sorted_codes, (const uint32_t*)c1, 2 * N, st))) return rc;
    return copy_u32(perm, i1, N, st);
  }
  hipLaunchKernelGGL(k_b_point_system, dim3(grid_for(N, 256, 4096)), dim3(256), 0, st, sg, i1, N,
                     sk);
  YGG_LAUNCH_CHECK();
  if ((rc = sort_pairs_u32(sk, i1, sk2, perm, N, 0, bits_for(b->num_systems - 1), sws, rest, st)))
    return rc;
  hipLaunchKernelGGL(k_gather_u64, dim3(grid_for(N, 256, 8192)), dim3(256), 0, st, codes, perm, N,
                     sorted_codes);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_batch_reorder(const YggBatch* b, const float* pos, const float* mass,
                      const uint32_t* perm, float* pos_sorted, float* mass_sorted,
                      int32_t* perm_local, int32_t* inverse_local, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  hipLaunchKernelGGL(k_b_reorder, dim3(grid_for(b->num_points, 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, make_seg(b), pos, mass, perm, b->num_points, pos_sorted,
                     mass_sorted, perm_local, inverse_local);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_batch_radix_build_workspace_bytes(const YggBatch* b) {
  if (check_batch(b)) return 0;
  const int64_t T = 2 * b->num_leaves - b->num_systems;
  Arena a(nullptr, 0);
  a.take<uint32_t>((size_t)T);                      // parent encodings
  a.take<uint32_t>((size_t)64 * b->num_systems);    // level counts
  a.take<uint32_t>((size_t)T);                      // level keys
  a.take<uint32_t>((size_t)T);                      // sorted level keys
  a.take<char>(ygg::sort_workspace_bytes(T, 4, 32));
  return a.off;
}

int ygg_batch_radix_build(const YggBatch* b, const uint64_t* sorted_codes,
                          const YggRadixTopology* o, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!o) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Seg sg = make_seg(b);
  const int S = b->num_systems;
  const int64_t Ltot = b->num_leaves, T = 2 * Ltot - S, NItot = Ltot - S;
  Arena a(ws, ws_bytes);
  uint32_t* penc = a.take<uint32_t>((size_t)T);
  uint32_t* lcount = a.take<uint32_t>((size_t)64 * S);
  uint32_t* lkey = a.take<uint32_t>((size_t)T);
  uint32_t* lkey2 = a.take<uint32_t>((size_t)T);
  size_t sb = ygg::sort_workspace_bytes(T, 4, 32);
  char* sws = a.take<char>(sb);
  if (!penc || !lcount || !lkey || !lkey2 || !sws) return YGG_ERR_WORKSPACE;
  if ((rc = fill_u32((uint32_t*)o->scalars, 0u, 8 * (int64_t)S, st))) return rc;
  if ((rc = fill_u32(penc, 0u, T, st))) return rc;
  if ((rc = fill_u32(lcount, 0u, 64 * (int64_t)S, st))) return rc;
  hipLaunchKernelGGL(k_b_leaves, dim3(grid_for(Ltot, 256)), dim3(256), 0, st, sg, sorted_codes,
                     Ltot, o->leaf_codes, o->node_ranges);
  if (NItot > 0)
    hipLaunchKernelGGL(k_b_karras, dim3(grid_for(NItot, 256)), dim3(256), 0, st, sg, o->leaf_codes,
                       NItot, o->left_child, o->right_child, o->left_is_leaf, o->right_is_leaf,
                       o->node_ranges, penc);
  hipLaunchKernelGGL(k_b_levels, dim3(grid_for(T, 256)), dim3(256), 0, st, sg, penc, T,
                     o->left_child, o->right_child, o->parent, o->node_level, lcount, o->scalars);
  hipLaunchKernelGGL(k_b_level_offsets, dim3(grid_for(S, 64)), dim3(64), 0, st, sg, lcount,
                     o->level_offsets, o->scalars);
  YGG_LAUNCH_CHECK();
  // nodes_by_level: stable sort by (system, level); levels < 256 here (deeper
  // trees set scalars[2] and are rebuilt one at a time by the caller).
  hipLaunchKernelGGL(k_b_level_keys, dim3(grid_for(T, 256, 8192)), dim3(256), 0, st, sg,
                     o->node_level, T, 8, lkey);
  YGG_LAUNCH_CHECK();
  const int end_bit = 8 + bits_for(S - 1);
  if (end_bit > 32) return YGG_ERR_UNSUPPORTED;
  if ((rc = sort_pairs_u32(lkey, nullptr, lkey2, (uint32_t*)o->nodes_by_level, T, 0, end_bit, sws,
                           sb, st)))
    return rc;
  hipLaunchKernelGGL(k_b_localize_nodes, dim3(grid_for(T, 256, 8192)), dim3(256), 0, st, sg,
                     o->nodes_by_level, T);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_batch_octree_meta_workspace_bytes(const YggBatch* b) {
  if (check_batch(b)) return 0;
  const int64_t L = b->num_leaves, M = 2 * L - b->num_systems, S = b->num_systems;
  Arena a(nullptr, 0);
  a.take<uint64_t>((size_t)L);      // resolved codes
  a.take<int32_t>((size_t)L);       // resolved depths
  a.take<uint64_t>((size_t)M);      // candidate addresses
  a.take<uint64_t>((size_t)M);      // address-sorted addresses
  a.take<uint32_t>((size_t)M);      // address-sorted order
  a.take<uint32_t>((size_t)M);      // system keys
  a.take<uint32_t>((size_t)M);      // system keys (sorted)
  a.take<uint32_t>((size_t)M);      // order
  a.take<uint64_t>((size_t)M);      // sorted addresses
  a.take<int32_t>((size_t)M);       // flags
  a.take<int64_t>((size_t)M + 1);   // scan
  a.take<uint64_t>((size_t)M);      // oct addresses
  a.take<uint32_t>((size_t)M);      // level keys
  a.take<uint32_t>((size_t)M);      // level keys (sorted)
  a.take<uint32_t>((size_t)M * 8);  // children + 1
  a.take<uint32_t>((size_t)22 * S); // level counts
  a.take<uint32_t>((size_t)L);      // leaf keys
  a.take<uint32_t>((size_t)L);      // leaf keys (sorted)
  a.take<uint32_t>((size_t)L);      // native order
  size_t s1 = ygg::sort_workspace_bytes(M, 8, 64);
  size_t s2 = ygg::sort_workspace_bytes(M, 4, 32);
  a.take<char>(s1 > s2 ? s1 : s2);
  a.take<char>(ygg::scan_workspace_bytes(M));
  return a.off;
}

int ygg_batch_octree_meta(const YggBatch* b, const uint64_t* sc, const int32_t* node_ranges,
                          const YggOctreeMeta* o, void* ws, size_t ws_bytes, void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  if (!o) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Seg sg = make_seg(b);
  const int S = b->num_systems;
  const int64_t L = b->num_leaves, M = 2 * L - S, T = M;
  Arena a(ws, ws_bytes);
  uint64_t* rcod = a.take<uint64_t>((size_t)L);
  int32_t* rdep = a.take<int32_t>((size_t)L);
  uint64_t* caddr = a.take<uint64_t>((size_t)M);
  uint64_t* saddr1 = a.take<uint64_t>((size_t)M);
  uint32_t* order1 = a.take<uint32_t>((size_t)M);
  uint32_t* skey = a.take<uint32_t>((size_t)M);
  uint32_t* skey2 = a.take<uint32_t>((size_t)M);
  uint32_t* order = a.take<uint32_t>((size_t)M);
  uint64_t* saddr = a.take<uint64_t>((size_t)M);
  int32_t* flags = a.take<int32_t>((size_t)M);
  int64_t* uscan = a.take<int64_t>((size_t)M + 1);
  uint64_t* oaddr = a.take<uint64_t>((size_t)M);
  uint32_t* lvkey = a.take<uint32_t>((size_t)M);
  uint32_t* lvkey2 = a.take<uint32_t>((size_t)M);
  uint32_t* cplus = a.take<uint32_t>((size_t)M * 8);
  uint32_t* lcount = a.take<uint32_t>((size_t)22 * S);
  uint32_t* lfkey = a.take<uint32_t>((size_t)L);
  uint32_t* lfkey2 = a.take<uint32_t>((size_t)L);
  uint32_t* native = a.take<uint32_t>((size_t)L);
  size_t s1 = ygg::sort_workspace_bytes(M, 8, 64);
  size_t s2 = ygg::sort_workspace_bytes(M, 4, 32);
  size_t sb = s1 > s2 ? s1 : s2;
  char* sws = a.take<char>(sb);
  size_t scb = ygg::scan_workspace_bytes(M);
  char* scws = a.take<char>(scb);
  if (!rcod || !rdep || !caddr || !saddr1 || !order1 || !skey || !skey2 || !order || !saddr ||
      !flags || !uscan || !oaddr || !lvkey || !lvkey2 || !cplus || !lcount || !lfkey || !lfkey2 ||
      !native || !sws || !scws)
    return YGG_ERR_WORKSPACE;
  // leaf sort keys s * stride + slot must fit 32 bits
  int64_t maxM = 0;
  (void)maxM;
  const uint32_t stride = (uint32_t)b->max_leaves_per_system * 2;
  if ((uint64_t)stride * (uint64_t)S > 0xFFFFFFFFull) return YGG_ERR_UNSUPPORTED;
  if ((rc = fill_u32((uint32_t*)o->scalars, 0u, 4 * (int64_t)S, st))) return rc;
  if ((rc = fill_u32(lcount, 0u, 22 * (int64_t)S, st))) return rc;
  if ((rc = fill_u32(cplus, 0u, M * 8, st))) return rc;
  const dim3 gL(grid_for(L, 256)), gM(grid_for(M, 256)), gT(grid_for(T, 256));
  const dim3 gMs(grid_for(M, 256, 8192));
  hipLaunchKernelGGL(k_b_oct_resolve, gL, dim3(256), 0, st, sg, sc, node_ranges, L, rcod, rdep);
  hipLaunchKernelGGL(k_b_oct_candidates, gM, dim3(256), 0, st, sg, rcod, rdep, M, caddr);
  YGG_LAUNCH_CHECK();
  // stable sort by (system, address): address passes, then system passes
  if ((rc = sort_pairs_u64(caddr, nullptr, saddr1, order1, M, 0, 64, sws, sb, st))) return rc;
  if (S > 1) {
    hipLaunchKernelGGL(k_b_node_system, gMs, dim3(256), 0, st, sg, order1, M, skey);
    YGG_LAUNCH_CHECK();
    if ((rc = sort_pairs_u32(skey, order1, skey2, order, M, 0, bits_for(S - 1), sws, sb, st)))
      return rc;
    hipLaunchKernelGGL(k_gather_u64, gMs, dim3(256), 0, st, caddr, order, M, saddr);
  } else {
    if ((rc = copy_u32(order, order1, M, st))) return rc;
    if ((rc = copy_u32((uint32_t*)saddr, (const uint32_t*)saddr1, 2 * M, st))) return rc;
  }
  hipLaunchKernelGGL(k_b_oct_newflags, gM, dim3(256), 0, st, sg, saddr, M, flags);
  YGG_LAUNCH_CHECK();
  if ((rc = exclusive_scan_i32_to_i64(flags, nullptr, uscan, M, scws, scb, st))) return rc;
  hipLaunchKernelGGL(k_b_oct_unique, gM, dim3(256), 0, st, sg, saddr, order, rcod, rdep, M, uscan,
                     o->oct_node_codes, o->oct_node_depths, o->oct_valid_mask, oaddr);
  hipLaunchKernelGGL(k_b_oct_ranges, gM, dim3(256), 0, st, sg, o->oct_node_codes,
                     o->oct_node_depths, o->oct_valid_mask, rcod, node_ranges, M,
                     o->oct_node_ranges, lcount, lvkey, o->scalars);
  YGG_LAUNCH_CHECK();
  if ((rc = sort_pairs_u32(lvkey, nullptr, lvkey2, (uint32_t*)o->oct_nodes_by_level, M, 0,
                           5 + bits_for(S - 1), sws, sb, st)))
    return rc;
  hipLaunchKernelGGL(k_b_localize_nodes, gMs, dim3(256), 0, st, sg, o->oct_nodes_by_level, M);
  hipLaunchKernelGGL(k_b_oct_level_offsets, dim3(grid_for(S, 64)), dim3(64), 0, st, sg, lcount,
                     o->oct_level_offsets, o->scalars, uscan);
  hipLaunchKernelGGL(k_b_oct_parent, gM, dim3(256), 0, st, sg, o->oct_node_codes,
                     o->oct_node_depths, o->oct_valid_mask, oaddr, M, o->oct_parent, cplus);
  hipLaunchKernelGGL(k_b_oct_children, gM, dim3(256), 0, st, cplus, o->oct_valid_mask, M,
                     o->oct_children, o->oct_child_mask, o->oct_child_counts, o->oct_leaf_mask,
                     flags);
  YGG_LAUNCH_CHECK();
  if ((rc = exclusive_scan_i32_to_i64(flags, nullptr, uscan, M, scws, scb, st))) return rc;
  hipLaunchKernelGGL(k_b_oct_leaf_nodes, gM, dim3(256), 0, st, sg, o->oct_leaf_mask, uscan, M,
                     o->oct_leaf_nodes, o->scalars);
  hipLaunchKernelGGL(k_b_radix_to_oct, gT, dim3(256), 0, st, sg, sc, node_ranges, rcod, rdep, T,
                     oaddr, stride, o->radix_node_to_oct, lfkey);
  YGG_LAUNCH_CHECK();
  const int end_bit = bits_for((uint64_t)stride * (uint64_t)S - 1);
  if ((rc = sort_pairs_u32(lfkey, nullptr, lfkey2, native, L, 0, end_bit, sws, sb, st))) return rc;
  hipLaunchKernelGGL(k_b_fill_i32, dim3(grid_for(T, 256, 8192)), dim3(256), 0, st,
                     o->leaf_position, -1, T);
  hipLaunchKernelGGL(k_b_leaf_order, gL, dim3(256), 0, st, sg, native, L, o->leaf_indices,
                     o->leaf_position);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_batch_radix_leaf_order(const YggBatch* b, int32_t* leaf_indices, int32_t* leaf_position,
                               void* stream) {
  int rc = check_batch(b);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int64_t L = b->num_leaves, T = 2 * L - b->num_systems;
  hipLaunchKernelGGL(k_b_fill_i32, dim3(grid_for(T, 256, 8192)), dim3(256), 0, st, leaf_position,
                     -1, T);
  hipLaunchKernelGGL(k_b_leaf_order, dim3(grid_for(L, 256)), dim3(256), 0, st, make_seg(b),
                     nullptr, L, leaf_indices, leaf_position);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
