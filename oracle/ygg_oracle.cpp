// CPU oracle for the yggdrax tree pipeline.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU and in plain
// sequential C++, the algorithm of the reference (TobiBu/yggdrax, JAX) for
// the path bounds -> Morton encode -> stable sort -> Karras radix build ->
// levels -> explicit octree metadata -> geometry -> dual-tree walk. It is
// compiled with -O2 -ffp-contract=off (no FMA contraction, no fast math) so
// every float32 operation rounds exactly once, in the order the reference
// writes it. Only the test suite, the smoke check and the CPU-baseline leg of
// the benchmark may load it; the shipped library never does.
//
// Each function cites the reference file:line it follows.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#include <parallel/algorithm>
#endif

namespace {

inline int clz64(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

// _tree_impl.py:340-346 (_lcp_codes_only): 64 for equal codes, else clz(xor).
inline int64_t lcp_codes(uint64_t a, uint64_t b) {
  uint64_t x = a ^ b;
  return x == 0 ? 64 : clz64(x);
}

// morton.py:22-29 (_spread3_u64)
inline uint64_t spread3(uint64_t x) {
  x &= 0x1FFFFFull;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// octree.py:152-165 (_common_depth_from_codes)
inline int64_t common_depth(uint64_t a, uint64_t b) {
  uint64_t x = a ^ b;
  int64_t bits = (x == 0) ? 63 : std::max<int64_t>(clz64(x) - 1, 0);
  return bits / 3;
}

// octree.py:136-141 (_prefix_code)
inline uint64_t prefix_code(uint64_t c, int64_t d) {
  if (d <= 0) return 0;
  int64_t s = std::max<int64_t>(0, 63 - 3 * d);
  return (c >> s) << s;
}

// octree.py:144-149 (_octant_at_depth)
inline int64_t octant_at_depth(uint64_t c, int64_t d) {
  if (d <= 0) return 0;
  int64_t s = std::max<int64_t>(0, 63 - 3 * d);
  return (int64_t)((c >> s) & 7ull);
}

// octree.py:113-133 (_oct_address): (8^d - 1)/7 + (code >> (63 - 3d)),
// negative depths clamped to 0.
inline uint64_t oct_address(uint64_t c, int64_t d) {
  uint64_t du = (uint64_t)std::max<int64_t>(0, d);
  uint64_t shift = 63ull - 3ull * du;
  uint64_t pre = shift >= 64 ? 0 : (c >> shift);
  uint64_t p8 = (3ull * du) >= 64 ? 0 : (1ull << (3ull * du));
  return (p8 - 1ull) / 7ull + pre;
}

// jnp.searchsorted(a, q, side, method="scan") as implemented by jax 0.8
// (_searchsorted_via_scan): a fixed ceil(log2(n+1)) = bit_length(n) rounds of
// "mid = (low+high)//2; go_left = q <= a[mid] (left) / q < a[mid] (right);
// low, high = (low, mid) if go_left else (mid, high)", returning high. The
// exact iteration matters because the reference calls it on an unsorted
// array (octree.py:296-301).
inline int64_t jax_searchsorted(const uint64_t* a, int64_t n, uint64_t q,
                                bool left) {
  int levels = 0;
  for (int64_t v = n; v > 0; v >>= 1) ++levels;
  int64_t low = 0, high = n;
  for (int it = 0; it < levels; ++it) {
    int64_t mid = (int64_t)(((uint64_t)low + (uint64_t)high) / 2ull);
    int64_t m = mid < n ? mid : n - 1;  // XLA gathers clamp the index
    bool go_left = left ? (q <= a[m]) : (q < a[m]);
    if (go_left)
      high = mid;
    else
      low = mid;
  }
  return high;
}

template <class K>
void stable_argsort(const K* keys, int64_t n, int64_t* perm) {
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  auto less = [keys](int64_t a, int64_t b) { return keys[a] < keys[b]; };
#ifdef _OPENMP
  // a stable sort has one result, whatever the thread count
  __gnu_parallel::stable_sort(perm, perm + n, less);
#else
  std::stable_sort(perm, perm + n, less);
#endif
}

}  // namespace

extern "C" {

// bounds.py:29-33: lo = min - pad, hi = max + pad, pad = max(span*0.05f, 1e-6f)
void ygo_infer_bounds(const float* pos, int64_t n, float* lo, float* hi) {
  for (int d = 0; d < 3; ++d) {
    float mn = pos[d], mx = pos[d];
    for (int64_t i = 1; i < n; ++i) {
      float v = pos[i * 3 + d];
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    float span = mx - mn;
    float pad = span * 0.05f;
    if (pad < 1e-6f) pad = 1e-6f;
    lo[d] = mn - pad;
    hi[d] = mx + pad;
  }
}

// morton.py:68-77: norm = (p-lo)/(hi-lo), clip to [0, 1], rint(norm*2097151),
// interleave x | y<<1 | z<<2.
void ygo_morton_encode(const float* pos, int64_t n, const float* lo,
                       const float* hi, uint64_t* codes) {
  float w[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  for (int64_t i = 0; i < n; ++i) {
    uint64_t c[3];
    for (int d = 0; d < 3; ++d) {
      float v = (pos[i * 3 + d] - lo[d]) / w[d];
      v = std::max(0.0f, v);
      v = std::min(1.0f, v);
      float s = v * 2097151.0f;
      c[d] = (uint64_t)std::nearbyint(s);  // round half to even
    }
    codes[i] = spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
  }
}

// tree.py:402-404 / _tree_impl.py:401: stable ascending argsort.
void ygo_stable_argsort_u64(const uint64_t* keys, int64_t n, int64_t* perm) {
  stable_argsort(keys, n, perm);
}
void ygo_stable_argsort_i64(const int64_t* keys, int64_t n, int64_t* perm) {
  stable_argsort(keys, n, perm);
}

// _tree_impl.py:1009-1316 (_build_tree_from_leaf_partitions) for the
// fixed-count adaptive partition (_tree_impl.py:405-406):
// leaf j covers [j*leaf, min((j+1)*leaf, n)).
// Outputs (all int64 unless noted) sized for L = ceil(n/leaf) leaves:
//   parent[2L-1], left[L-1], right[L-1], lflag[L-1] (u8), rflag[L-1] (u8),
//   ranges[(2L-1)*2], node_level[2L-1], level_offsets[65],
//   nodes_by_level[2L-1], scalars[0] = num_levels, leaf_codes[L] (u64).
void ygo_radix_build(const uint64_t* sorted_codes, int64_t n, int64_t leaf,
                     int64_t* parent, int64_t* left, int64_t* right,
                     uint8_t* lflag, uint8_t* rflag, int64_t* ranges,
                     int64_t* node_level, int64_t* level_offsets,
                     int64_t* nodes_by_level, int64_t* scalars,
                     uint64_t* leaf_codes) {
  const int64_t L = (n + leaf - 1) / leaf;
  std::vector<int64_t> starts(L), ends(L);
  for (int64_t j = 0; j < L; ++j) {
    starts[j] = j * leaf;
    ends[j] = std::min(starts[j] + leaf, n);
    leaf_codes[j] = sorted_codes[starts[j]];
  }
  if (L == 1) {  // _tree_impl.py:1052-1115
    parent[0] = -1;
    ranges[0] = 0;
    ranges[1] = n - 1;
    node_level[0] = 0;
    for (int k = 0; k < 65; ++k) level_offsets[k] = 0;
    level_offsets[1] = 1;
    nodes_by_level[0] = 0;
    scalars[0] = 1;
    return;
  }
  const int64_t NI = L - 1, T = 2 * L - 1;
  const uint64_t* lc = leaf_codes;
  auto delta = [&](int64_t i, int64_t j) -> int64_t {  // 1121-1138
    if (j < 0 || j >= L) return -1;
    uint64_t ci = lc[i], cj = lc[j];
    if (ci == cj) return 64 + clz64((uint64_t)i ^ (uint64_t)j);
    return clz64(ci ^ cj);
  };
  for (int64_t i = 0; i < NI; ++i) {
    int64_t dp = delta(i, i + 1), dm = delta(i, i - 1);
    int64_t d = (dp > dm) ? 1 : ((dp < dm) ? -1 : 0);  // 1143-1145
    if (i == 0) d = 1;
    if (d == 0) d = 1;
    int64_t dmin = delta(i, i - d);
    int64_t span = 0;
    for (int k = 30; k >= 0; --k) {  // 1148-1155
      int64_t step = (int64_t)1 << k;
      int64_t cand = i + (span + step) * d;
      if (cand >= 0 && cand < L && delta(i, cand) > dmin) span += step;
    }
    int64_t e = i + span * d;
    int64_t first = std::min(i, e), last = std::max(i, e);
    ranges[i * 2 + 0] = starts[first];
    ranges[i * 2 + 1] = ends[last] - 1;
    uint64_t fc = lc[first];
    int64_t common = lcp_codes(fc, lc[last]);  // quirk: codes-only lcp
    int64_t split = first;
    for (int k = 30; k >= 0; --k) {  // 1170-1176
      int64_t mid = split + ((int64_t)1 << k);
      int64_t safe_mid = mid < last ? mid : last;
      if (mid < last && lcp_codes(fc, lc[safe_mid]) > common) split = mid;
    }
    bool lf = split == first, rf = split + 1 == last;
    lflag[i] = lf;
    rflag[i] = rf;
    left[i] = lf ? NI + first : split;
    right[i] = rf ? NI + last : split + 1;
  }
  // 1242-1251: parent.at[left].set(i) then parent.at[right].set(i),
  // restated with sequential last-writer semantics.
  for (int64_t k = 0; k < T; ++k) parent[k] = -1;
  for (int64_t i = 0; i < NI; ++i) parent[left[i]] = i;
  for (int64_t i = 0; i < NI; ++i) parent[right[i]] = i;
  for (int64_t j = 0; j < L; ++j) {
    ranges[(NI + j) * 2 + 0] = starts[j];
    ranges[(NI + j) * 2 + 1] = ends[j] - 1;
  }
  // 1272-1299: bounded pointer doubling.
  std::vector<int64_t> sc(T), dist(T), nsc(T), nd(T);
  for (int64_t k = 0; k < T; ++k) {
    bool root = parent[k] < 0;
    sc[k] = root ? k : parent[k];
    dist[k] = root ? 0 : 1;
  }
  int bl = 0;
  for (int64_t v = T; v > 0; v >>= 1) ++bl;
  int rounds = std::max(1, bl + 1);
  for (int r = 0; r < rounds; ++r) {
    for (int64_t k = 0; k < T; ++k) {
      nd[k] = dist[k] + dist[sc[k]];
      nsc[k] = sc[sc[k]];
    }
    std::swap(sc, nsc);
    std::swap(dist, nd);
  }
  int64_t maxl = 0;
  int64_t counts[64] = {0};
  for (int64_t k = 0; k < T; ++k) {
    node_level[k] = dist[k];
    maxl = std::max(maxl, dist[k]);
    if (dist[k] >= 0 && dist[k] < 64) counts[dist[k]] += 1;  // OOB dropped
  }
  level_offsets[0] = 0;
  for (int k = 0; k < 64; ++k) level_offsets[k + 1] = level_offsets[k] + counts[k];
  stable_argsort(node_level, T, nodes_by_level);
  scalars[0] = std::min<int64_t>(std::max<int64_t>(maxl + 1, 1), 64);
}

// octree.py:479-550 + 204-419: explicit octree metadata for a radix topology
// with adaptive leaves (leaf_depths all -1). M = max(1, 2L-1) slots.
// Outputs: oct_parent[M], oct_children[M*8], oct_child_counts[M],
// oct_child_mask[M*8] (u8), oct_valid_mask[M] (u8), oct_node_codes[M] (u64),
// oct_node_depths[M], oct_node_ranges[M*2], oct_nodes_by_level[M],
// oct_level_offsets[23], scalars[0] = oct_num_levels, oct_leaf_mask[M] (u8),
// oct_leaf_nodes[M], radix_node_to_oct[2L-1], radix_leaf_to_oct[L],
// resolved_codes[L] (u64), resolved_depths[L].
void ygo_octree_meta(const uint64_t* sorted_codes, const int64_t* node_ranges,
                     int64_t L, int64_t* oct_parent, int64_t* oct_children,
                     int64_t* oct_child_counts, uint8_t* oct_child_mask,
                     uint8_t* oct_valid_mask, uint64_t* oct_node_codes,
                     int64_t* oct_node_depths, int64_t* oct_node_ranges,
                     int64_t* oct_nodes_by_level, int64_t* oct_level_offsets,
                     int64_t* scalars, uint8_t* oct_leaf_mask,
                     int64_t* oct_leaf_nodes, int64_t* radix_node_to_oct,
                     int64_t* radix_leaf_to_oct, uint64_t* resolved_codes,
                     int64_t* resolved_depths) {
  const int64_t T = 2 * L - 1;
  const int64_t NI = (T - 1) / 2;
  const int64_t M = std::max<int64_t>(1, 2 * L - 1);
  const uint64_t U64MAX = ~0ull;
  std::vector<int64_t> starts(L), ends(L);
  // _resolved_leaf_cells / _resolved_leaf_partitions (octree.py:168-201)
  for (int64_t j = 0; j < L; ++j) {
    int64_t a = node_ranges[(NI + j) * 2], b = node_ranges[(NI + j) * 2 + 1];
    starts[j] = a;
    ends[j] = b + 1;
    uint64_t f = sorted_codes[a], l = sorted_codes[b];
    int64_t dep = common_depth(f, l);
    resolved_depths[j] = dep;
    resolved_codes[j] = prefix_code(f, dep);
  }
  const uint64_t* lc = resolved_codes;
  const int64_t* ld = resolved_depths;
  // Phase 1 (246-255): L leaves then L-1 LCAs of consecutive leaf cells.
  std::vector<uint64_t> all_codes(M);
  std::vector<int64_t> all_depths(M);
  for (int64_t j = 0; j < L; ++j) {
    all_codes[j] = lc[j];
    all_depths[j] = ld[j];
  }
  for (int64_t j = 0; j + 1 < L; ++j) {
    int64_t dep = common_depth(lc[j], lc[j + 1]);
    all_depths[L + j] = dep;
    all_codes[L + j] = prefix_code(lc[j], dep);
  }
  // Phase 2 (257-282): stable sort by address, dedupe, scatter-max.
  std::vector<uint64_t> addr(M);
  for (int64_t k = 0; k < M; ++k) addr[k] = oct_address(all_codes[k], all_depths[k]);
  std::vector<int64_t> order(M);
  stable_argsort(addr.data(), M, order.data());
  for (int64_t k = 0; k < M; ++k) {
    oct_node_codes[k] = 0;
    oct_node_depths[k] = -1;
  }
  int64_t uidx = -1;
  for (int64_t k = 0; k < M; ++k) {
    bool is_new = (k == 0) || addr[order[k]] != addr[order[k - 1]];
    if (is_new) ++uidx;
    oct_node_codes[uidx] = std::max(oct_node_codes[uidx], all_codes[order[k]]);
    oct_node_depths[uidx] = std::max(oct_node_depths[uidx], all_depths[order[k]]);
  }
  const int64_t num_unique = uidx + 1;
  for (int64_t k = 0; k < M; ++k) oct_valid_mask[k] = k < num_unique;
  // Phase 3 (284-313): ranges by searchsorted on the resolved leaf codes.
  for (int64_t k = 0; k < M; ++k) {
    bool valid = oct_valid_mask[k];
    uint64_t du = (uint64_t)std::max<int64_t>(oct_node_depths[k], 0);
    uint64_t bs = 63ull - 3ull * du;
    uint64_t bss = std::min<uint64_t>(bs, 62ull);
    uint64_t mask = (1ull << bss) - 1ull;
    uint64_t endc = valid ? (oct_node_codes[k] | mask) : 0ull;
    int64_t f = jax_searchsorted(lc, L, oct_node_codes[k], true);
    int64_t l = jax_searchsorted(lc, L, endc, false) - 1;
    int64_t bound = std::max<int64_t>(L - 1, 0);
    f = std::min(std::max<int64_t>(f, 0), bound);
    l = std::min(std::max<int64_t>(l, 0), bound);
    if (valid) {
      oct_node_ranges[k * 2] = starts[f];
      oct_node_ranges[k * 2 + 1] = ends[l] - 1;
    } else {
      oct_node_ranges[k * 2] = 0;
      oct_node_ranges[k * 2 + 1] = -1;
    }
  }
  // Phase 4 (315-337): levels.
  int64_t maxd = 0;
  int64_t lcnt[22] = {0};
  std::vector<int64_t> sort_depths(M);
  for (int64_t k = 0; k < M; ++k) {
    bool valid = oct_valid_mask[k];
    int64_t dep = oct_node_depths[k];
    if (valid) maxd = std::max(maxd, dep);
    int64_t cd = std::min<int64_t>(std::max<int64_t>(dep, 0), 21);
    if (valid) lcnt[cd] += 1;
    sort_depths[k] = valid ? dep : 22;
  }
  scalars[0] = maxd + 1;
  oct_level_offsets[0] = 0;
  for (int k = 0; k < 22; ++k) oct_level_offsets[k + 1] = oct_level_offsets[k] + lcnt[k];
  stable_argsort(sort_depths.data(), M, oct_nodes_by_level);
  // Phase 5 (339-373): parent = first (deepest) existing ancestor address.
  std::vector<uint64_t> oaddr(M);
  for (int64_t k = 0; k < M; ++k)
    oaddr[k] = oct_valid_mask[k] ? oct_address(oct_node_codes[k], oct_node_depths[k]) : U64MAX;
  for (int64_t k = 0; k < M; ++k) oct_parent[k] = -1;
  for (int64_t step = 1; step <= 21; ++step) {
    for (int64_t k = 0; k < M; ++k) {
      bool needs = oct_valid_mask[k] && oct_parent[k] < 0 && oct_node_depths[k] >= step;
      if (!needs) continue;
      int64_t pd = oct_node_depths[k] - step;
      uint64_t pa = pd >= 0 ? oct_address(prefix_code(oct_node_codes[k], pd), pd) : U64MAX;
      int64_t pos = jax_searchsorted(oaddr.data(), M, pa, true);
      int64_t safe = std::min(pos, M - 1);
      if (oaddr[safe] == pa) oct_parent[k] = pos;
    }
  }
  // Phase 6 (375-391): children by octant at the child's own depth,
  // scatter-max keeps the larger child id on collision.
  std::vector<int64_t> cplus(M * 8, 0);
  for (int64_t k = 0; k < M; ++k) {
    bool vc = oct_valid_mask[k] && oct_parent[k] >= 0;
    if (!vc) continue;
    int64_t o = octant_at_depth(oct_node_codes[k], oct_node_depths[k]);
    int64_t& slot = cplus[oct_parent[k] * 8 + o];
    slot = std::max(slot, k + 1);
  }
  int64_t nleaf = 0;
  for (int64_t k = 0; k < M; ++k) {
    int64_t cnt = 0;
    for (int o = 0; o < 8; ++o) {
      int64_t c = cplus[k * 8 + o] - 1;
      oct_children[k * 8 + o] = c;
      bool m = oct_valid_mask[k] && c >= 0;
      oct_child_mask[k * 8 + o] = m;
      cnt += m;
    }
    oct_child_counts[k] = cnt;
    oct_leaf_mask[k] = oct_valid_mask[k] && cnt == 0;
    if (oct_leaf_mask[k]) oct_leaf_nodes[nleaf++] = k;
  }
  for (int64_t k = nleaf; k < M; ++k) oct_leaf_nodes[k] = -1;
  // Radix-node mapping (octree.py:518-532): lower_bound, no match check.
  for (int64_t k = 0; k < T; ++k) {
    int64_t dep;
    uint64_t code;
    if (k >= NI) {
      dep = ld[k - NI];
      code = lc[k - NI];
    } else {
      uint64_t f = sorted_codes[node_ranges[k * 2]];
      uint64_t l = sorted_codes[node_ranges[k * 2 + 1]];
      dep = common_depth(f, l);
      code = prefix_code(f, dep);
    }
    radix_node_to_oct[k] = jax_searchsorted(oaddr.data(), M, oct_address(code, dep), true);
  }
  for (int64_t j = 0; j < L; ++j) radix_leaf_to_oct[j] = radix_node_to_oct[NI + j];
}

// _geometry_impl.py:176-338 (compute_tree_geometry) for positional leaf
// bounds. T = 2L-1 nodes, NI = L-1 internal. prad[n] (sorted order) or
// NULL is the per-particle radius. Outputs center[T*3],
// half[T*3], radius[T], max_extent[T]. Returns 0, or 1 if the depth pointer
// doubling did not converge (the reference would loop forever there).
int ygo_geometry(const float* pos_sorted, const float* prad, const int64_t* ranges,
                 const int64_t* left, const int64_t* right,
                 const int64_t* parent, int64_t T, int64_t leaf_cap,
                 float* center, float* half, float* radius, float* maxext) {
  const int64_t NI = (T - 1) / 2;
  std::vector<float> mn(T * 3, 0.0f), mx(T * 3, 0.0f);
  const float big = 3.4028234663852886e38f;
  for (int64_t k = NI; k < T; ++k) {  // _compute_leaf_bounds (63-127)
    int64_t a = ranges[k * 2], cnt = ranges[k * 2 + 1] - a + 1;
    for (int d = 0; d < 3; ++d) {
      float lo = big, hi = -big;
      for (int64_t m = 0; m < leaf_cap; ++m) {
        if (m >= cnt) continue;
        float v = pos_sorted[(a + m) * 3 + d];
        // 107-112: with particle_radius each point is the cube [v - r, v + r]
        float vl = prad ? v - prad[a + m] : v;
        float vh = prad ? v + prad[a + m] : v;
        lo = std::min(lo, vl);
        hi = std::max(hi, vh);
      }
      mn[k * 3 + d] = lo;
      mx[k * 3 + d] = hi;
    }
  }
  int status = 0;
  if (NI > 0) {
    // 285-300: pointer doubling to convergence.
    std::vector<int64_t> sc(T), dist(T), nsc(T), nd(T);
    for (int64_t k = 0; k < T; ++k) {
      bool root = parent[k] < 0;
      sc[k] = root ? k : parent[k];
      dist[k] = root ? 0 : 1;
    }
    bool changed = true;
    int guard = 0;
    while (changed) {
      changed = false;
      for (int64_t k = 0; k < T; ++k) {
        nd[k] = dist[k] + dist[sc[k]];
        nsc[k] = sc[sc[k]];
        if (nsc[k] != sc[k]) changed = true;
      }
      std::swap(sc, nsc);
      std::swap(dist, nd);
      if (++guard > 200) {
        status = 1;
        break;
      }
    }
    int64_t maxd = 0;
    for (int64_t k = 0; k < T; ++k) maxd = std::max(maxd, dist[k]);
    // 302-329: level-synchronous upward min/max, deepest level first.
    std::vector<float> cmn(NI * 3), cmx(NI * 3);
    for (int64_t r = 0; r < maxd; ++r) {
      int64_t level = maxd - 1 - r;
      for (int64_t i = 0; i < NI; ++i) {
        for (int d = 0; d < 3; ++d) {
          cmn[i * 3 + d] = std::min(mn[left[i] * 3 + d], mn[right[i] * 3 + d]);
          cmx[i * 3 + d] = std::max(mx[left[i] * 3 + d], mx[right[i] * 3 + d]);
        }
      }
      for (int64_t i = 0; i < NI; ++i) {
        if (dist[i] != level) continue;
        for (int d = 0; d < 3; ++d) {
          mn[i * 3 + d] = cmn[i * 3 + d];
          mx[i * 3 + d] = cmx[i * 3 + d];
        }
      }
    }
  }
  for (int64_t k = 0; k < T; ++k) {  // 331-338
    float h[3];
    for (int d = 0; d < 3; ++d) {
      center[k * 3 + d] = 0.5f * (mn[k * 3 + d] + mx[k * 3 + d]);
      float hh = 0.5f * (mx[k * 3 + d] - mn[k * 3 + d]);
      h[d] = std::max(hh, 1e-6f);
      half[k * 3 + d] = h[d];
    }
    float s = h[0] * h[0];
    s = s + h[1] * h[1];
    s = s + h[2] * h[2];
    radius[k] = std::sqrt(s);
    maxext[k] = std::max(std::max(h[0], h[1]), h[2]);
  }
  return status;
}

// Result of one fixed-capacity dual-tree walk attempt.
struct WalkResult {
  std::vector<int64_t> far_counts, near_counts;
  std::vector<std::vector<int64_t>> far_rows, near_rows;
  std::vector<int64_t> gen_sizes;
  int64_t flags[3];      // queue, far, near overflow
  int64_t decisions[3];  // accept, near, refine
};

}  // extern "C"

namespace {

// One evaluated pair's outcome inside a generation chunk.
struct Acc { int64_t t, s; };
struct ChunkOut {
  std::vector<Acc> acc, nea;          // accepted and near pairs, in pair order
  std::vector<int64_t> ct, cs;        // child pairs (min, max), in pair order
  int64_t dec[3] = {0, 0, 0};
};

// Evaluates pairs [p0, p1) of a generation (MAC, split rule 1444-1454,
// children as (min, max)).
void walk_chunk(int64_t p0, int64_t p1, const std::vector<int64_t>& wt,
                const std::vector<int64_t>& ws, int64_t NI, const int64_t* left,
                const int64_t* right, const float* center, const float* mac_ext,
                float theta_sq, int mac_type, ChunkOut& o) {
  for (int64_t p = p0; p < p1; ++p) {
    int64_t t = wt[p], s = ws[p];
    float dx = center[t * 3 + 0] - center[s * 3 + 0];
    float dy = center[t * 3 + 1] - center[s * 3 + 1];
    float dz = center[t * 3 + 2] - center[s * 3 + 2];
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    float et = mac_ext[t], es = mac_ext[s];
    bool diff = t != s;
    bool ok;
    if (mac_type == 2) {  // engblom (384-393)
      float rb = std::max(et, es), rs = std::min(et, es);
      float tt = std::sqrt(theta_sq);
      float lhs = rb + tt * rs;
      ok = diff && d2 > 0.0f && (lhs * lhs <= theta_sq * d2);
    } else {  // bh / dehnen (368-382)
      float es2 = et + es;
      ok = diff && d2 > 0.0f && (es2 * es2 <= theta_sq * d2);
    }
    bool t_int = t < NI, s_int = s < NI;
    bool near = !ok && !t_int && !s_int && diff;
    if (ok) {
      o.dec[0] += 1;
      o.acc.push_back({t, s});
      continue;
    }
    if (near) {
      o.dec[1] += 1;
      o.nea.push_back({t, s});
      continue;
    }
    o.dec[2] += 1;
    bool same = !diff;
    bool split_t = t_int && (same || !s_int || et >= es);  // 1444-1454
    bool split_s = s_int && (same || !t_int || es > et);
    int64_t c[4][2];
    int nc = 0;
    if (split_t && split_s) {
      int64_t tl = left[t], tr = right[t];
      if (same) {
        c[0][0] = tl; c[0][1] = tl;
        c[1][0] = tl; c[1][1] = tr;
        c[2][0] = tr; c[2][1] = tr;
        nc = 3;
      } else {
        int64_t sl = left[s], sr = right[s];
        c[0][0] = tl; c[0][1] = sl;
        c[1][0] = tl; c[1][1] = sr;
        c[2][0] = tr; c[2][1] = sl;
        c[3][0] = tr; c[3][1] = sr;
        nc = 4;
      }
    } else if (split_t) {
      c[0][0] = left[t]; c[0][1] = s;
      c[1][0] = right[t]; c[1][1] = s;
      nc = 2;
    } else if (split_s) {
      c[0][0] = t; c[0][1] = left[s];
      c[1][0] = t; c[1][1] = right[s];
      nc = 2;
    }
    for (int k = 0; k < nc; ++k) {
      int64_t a = c[k][0], b = c[k][1];
      if (a < 0 || b < 0) continue;
      o.ct.push_back(std::min(a, b));
      o.cs.push_back(std::max(a, b));
    }
  }
}

// Appends one generation's directed entries to their rows in the serial
// order of the reference's two scatters (1464-1566 far, 1568-1672 near):
// every forward entry (row = pair's first node) in pair order, then every
// backward entry. A row whose count reaches `cap` drops the entry and raises
// the overflow flag. Rows are owned by buckets (row % kBuckets), each bucket
// takes its entries in that same order, so the rows come out exactly as a
// sequential loop would write them, with any number of threads.
bool append_rows(const std::vector<ChunkOut>& outs, bool near, const int64_t* leaf_pos,
                 int64_t cap, std::vector<int64_t>& counts,
                 std::vector<std::vector<int64_t>>& rows) {
  constexpr int kBuckets = 64;
  const int64_t nch = (int64_t)outs.size();
  const int64_t nseg = 2 * nch;  // forward segments, then backward ones
  auto seg_list = [&](int64_t g) -> const std::vector<Acc>& {
    const ChunkOut& o = outs[g % nch];
    return near ? o.nea : o.acc;
  };
  auto row_val = [&](int64_t g, const Acc& e, int64_t& row, int64_t& val) {
    const bool fwd = g < nch;
    int64_t r = fwd ? e.t : e.s;
    val = fwd ? e.s : e.t;
    row = near ? leaf_pos[r] : r;
  };
  std::vector<int64_t> cnt((size_t)nseg * kBuckets, 0);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < nseg; ++g) {
    for (const Acc& e : seg_list(g)) {
      int64_t row, val;
      row_val(g, e, row, val);
      cnt[(size_t)g * kBuckets + (size_t)(row % kBuckets)] += 1;
    }
  }
  std::vector<int64_t> off((size_t)nseg * kBuckets), bstart(kBuckets + 1, 0);
  int64_t run = 0;
  for (int b = 0; b < kBuckets; ++b) {
    bstart[b] = run;
    for (int64_t g = 0; g < nseg; ++g) {
      off[(size_t)g * kBuckets + b] = run;
      run += cnt[(size_t)g * kBuckets + b];
    }
  }
  bstart[kBuckets] = run;
  std::vector<Acc> buf((size_t)run);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < nseg; ++g) {
    int64_t* o = &off[(size_t)g * kBuckets];
    for (const Acc& e : seg_list(g)) {
      int64_t row, val;
      row_val(g, e, row, val);
      buf[(size_t)o[row % kBuckets]++] = {row, val};
    }
  }
  int over = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : over)
  for (int b = 0; b < kBuckets; ++b) {
    for (int64_t i = bstart[b]; i < bstart[b + 1]; ++i) {
      const Acc& e = buf[(size_t)i];
      int64_t& c = counts[e.t];
      if (c < cap) {
        rows[e.t].push_back(e.s);
        ++c;
      } else {
        over = 1;
      }
    }
  }
  return over != 0;
}

}  // namespace

extern "C" {

// Threads of the OpenMP regions (the outputs do not depend on it).
void ygo_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int ygo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// _interactions_impl.py:1131-1785 (_dual_tree_walk_impl) restated as a
// generation-synchronous loop. mac_type: 0 = bh, 1 = dehnen, 2 = engblom.
// mac_ext[T] is the per-node MAC extent (_build_mac_extents, 993-1061).
// leaf_pos[T] maps a leaf node to its native leaf row (-1 for internal).
// Rows hold entries in emission order; capacities follow the reference:
// an entry whose slot reaches the cap is dropped and raises the flag, and
// the walk stops after the generation in which any flag was raised. Pairs
// of a generation are evaluated in chunks on OpenMP threads and merged in
// chunk order, so the result is the sequential one for any thread count.
void* ygo_walk_run(int64_t T, int64_t NI, const int64_t* left,
                   const int64_t* right, const float* center,
                   const float* mac_ext, const int64_t* leaf_pos, int64_t root,
                   float theta, int mac_type, int64_t queue_cap, int64_t far_cap,
                   int64_t near_cap, int collect_far, int collect_near) {
  WalkResult* r = new WalkResult();
  const int64_t L = T - NI;
  r->far_counts.assign(T, 0);
  r->near_counts.assign(L, 0);
  r->far_rows.assign(collect_far ? T : 0, {});
  r->near_rows.assign(collect_near ? L : 0, {});
  r->flags[0] = r->flags[1] = r->flags[2] = 0;
  r->decisions[0] = r->decisions[1] = r->decisions[2] = 0;
  if (NI == 0) {
    r->gen_sizes.push_back(0);
    return r;
  }
  const float th = theta;
  const float theta_sq = th * th;
  const int64_t Q = std::max<int64_t>(queue_cap, 4);
  std::vector<int64_t> wt{root}, ws{root};
  bool qo = false, fo = false, no = false;
  std::vector<ChunkOut> outs;
  while (!wt.empty() && !qo && !fo && !no) {
    r->gen_sizes.push_back((int64_t)wt.size());
    const int64_t n = (int64_t)wt.size();
    const int64_t nch = std::max<int64_t>(1, std::min<int64_t>(n / 2048, 512));
    outs.assign((size_t)nch, ChunkOut());
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < nch; ++c)
      walk_chunk(n * c / nch, n * (c + 1) / nch, wt, ws, NI, left, right, center, mac_ext,
                 theta_sq, mac_type, outs[(size_t)c]);
    int64_t pushes = 0;
    for (auto& o : outs) {
      for (int k = 0; k < 3; ++k) r->decisions[k] += o.dec[k];
      pushes += (int64_t)o.ct.size();
    }
    // the first Q child pairs in pair order enter the next generation
    std::vector<int64_t> nt, ns;
    const int64_t kept = std::min(pushes, Q);
    if (pushes > Q) qo = true;
    nt.reserve((size_t)kept);
    ns.reserve((size_t)kept);
    for (auto& o : outs) {
      for (size_t k = 0; k < o.ct.size() && (int64_t)nt.size() < kept; ++k) {
        nt.push_back(o.ct[k]);
        ns.push_back(o.cs[k]);
      }
    }
    if (collect_far && append_rows(outs, false, leaf_pos, far_cap, r->far_counts, r->far_rows))
      fo = true;
    if (collect_near &&
        append_rows(outs, true, leaf_pos, near_cap, r->near_counts, r->near_rows))
      no = true;
    wt.swap(nt);
    ws.swap(ns);
  }
  r->flags[0] = qo;
  r->flags[1] = fo;
  r->flags[2] = no;
  return r;
}

// sizes[0] = total far entries, [1] = total near entries, [2] = generations
// visited, [3..5] = overflow flags, [6..8] = decisions.
void ygo_walk_sizes(void* h, int64_t* sizes) {
  WalkResult* r = (WalkResult*)h;
  int64_t f = 0, n = 0;
  for (auto& row : r->far_rows) f += (int64_t)row.size();
  for (auto& row : r->near_rows) n += (int64_t)row.size();
  sizes[0] = f;
  sizes[1] = n;
  sizes[2] = (int64_t)r->gen_sizes.size();
  for (int k = 0; k < 3; ++k) sizes[3 + k] = r->flags[k];
  for (int k = 0; k < 3; ++k) sizes[6 + k] = r->decisions[k];
}

// Node-order CSR copies of the rows plus per-generation wavefront sizes.
void ygo_walk_copy(void* h, int64_t* far_counts, int64_t* far_entries,
                   int64_t* near_counts, int64_t* near_entries,
                   int64_t* gen_sizes) {
  WalkResult* r = (WalkResult*)h;
  std::copy(r->far_counts.begin(), r->far_counts.end(), far_counts);
  std::copy(r->near_counts.begin(), r->near_counts.end(), near_counts);
  int64_t o = 0;
  for (auto& row : r->far_rows)
    for (int64_t v : row) far_entries[o++] = v;
  o = 0;
  for (auto& row : r->near_rows)
    for (int64_t v : row) near_entries[o++] = v;
  std::copy(r->gen_sizes.begin(), r->gen_sizes.end(), gen_sizes);
}

void ygo_walk_free(void* h) { delete (WalkResult*)h; }

}  // extern "C"
