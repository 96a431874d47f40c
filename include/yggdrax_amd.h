/* C ABI of the yggdrax MI355X tree pipeline (libyggdrax_amd.so).
 *
 * Every entry point takes caller-owned device pointers, plain sizes, a
 * caller-provided device workspace (size from the matching *_workspace_bytes
 * query) and a hipStream_t passed as void*. No entry point allocates device
 * memory. All launch asynchronously on `stream`, except the few whose
 * comment says they block on it because a data-dependent size decides what
 * runs next (ygg_walk_bfs, ygg_walk_flat, ygg_grouped_classify).
 * Return value: YGG_OK or one of the YGG_ERR_* codes.
 *
 * Device indices are int32 (node ids, particle positions < 2^31); the host
 * wrappers widen them to the index dtype of the reference (int64 unless
 * YGGDRAX_INDEX_PRECISION=int32). Citations name the reference function each
 * entry point replaces (files under TobiBu/yggdrax, path:line).
 */
#ifndef YGGDRAX_AMD_H
#define YGGDRAX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YGG_OK 0
#define YGG_ERR_ARG 1
#define YGG_ERR_HIP 2
#define YGG_ERR_WORKSPACE 3
#define YGG_ERR_UNSUPPORTED 4
#define YGG_ERR_TOO_LARGE 5 /* input beyond a fixed maximum (ygg_morton_sort_noalloc) */

/* ---- primitives ---------------------------------------------------------- */

/* Stable LSD radix sort of (key, u32 value) pairs over key bits
 * [begin_bit, end_bit), digit_bits (8..11) bits per pass; digit_bits 0 takes
 * ygg_sort_default_digit_bits. The workspace query must use the same
 * digit_bits as the sort. vals_in may be NULL (values become 0..n-1).
 * keys_in must differ from keys_out. Replaces the XLA sorts behind
 * jnp.argsort(stable=True) (yggdrax/_tree_impl.py:401, yggdrax/octree.py:259,
 * yggdrax/_interactions_impl.py:1082). */
size_t ygg_sort_workspace_bytes(int64_t n, int key_bytes, int begin_bit, int end_bit,
                                int digit_bits);
int ygg_sort_default_digit_bits(int key_bytes, int begin_bit, int end_bit);
int ygg_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                       uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, int digit_bits,
                       void* ws, size_t ws_bytes, void* stream);
int ygg_sort_pairs_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                       uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, int digit_bits,
                       void* ws, size_t ws_bytes, void* stream);

/* ---- bounds and Morton codes -------------------------------------------- */

/* infer_bounds (yggdrax/bounds.py:11-33): bounds6 = {lo.x, lo.y, lo.z,
 * hi.x, hi.y, hi.z} of positions (n x 3 float32, row-major), padded by
 * max(5 % of the span, 1e-6). */
size_t ygg_bounds_workspace_bytes(void);
int ygg_infer_bounds(const float* pos, int64_t n, float* bounds6, void* ws, size_t ws_bytes,
                     void* stream);

/* morton_encode_impl (yggdrax/morton.py:43-77): 21-bit-per-axis codes. */
int ygg_morton_encode(const float* pos, int64_t n, const float* bounds6, uint64_t* codes,
                      void* stream);

/* Morton encode fused with the digit histograms, then a stable sort by code
 * (yggdrax/tree.py:399-404, yggdrax/_tree_impl.py:397-402): sorted_codes[n]
 * and perm[n] (particle index of each sorted position). */
size_t ygg_morton_sort_workspace_bytes(int64_t n);
int ygg_morton_sort(const float* pos, int64_t n, const float* bounds6, uint64_t* sorted_codes,
                    uint32_t* perm, void* ws, size_t ws_bytes, void* stream);

/* ygg_morton_sort without a workspace argument and without any allocation:
 * no counterpart in the reference. It is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Its workspace is a file-scope device
 * array of YGG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * n above YGG_NOALLOC_MAX_POINTS returns YGG_ERR_TOO_LARGE. The stream must
 * belong to the current device (at most YGG_NOALLOC_MAX_DEVICES devices).
 * Every call on one GPU shares that GPU's array, so calls that overlap on
 * one GPU (other streams or threads) race; calls ordered on one stream do
 * not. Outputs equal ygg_morton_sort's. */
#define YGG_NOALLOC_MAX_POINTS (1 << 20)
#define YGG_NOALLOC_SCRATCH_BYTES (24u << 20)
#define YGG_NOALLOC_MAX_DEVICES 64
int ygg_morton_sort_noalloc(const float* pos, int64_t n, const float* bounds6,
                            uint64_t* sorted_codes, uint32_t* perm, void* stream);

/* ygg_morton_sort plus the reordered particles in one call: the encode
 * kernel also packs (x, y, z, m) so the gather reads one 16-byte record per
 * particle. inverse (int64 if inverse_i64 else int32) may be NULL; masses
 * may be NULL (mass_sorted is then not written). */
size_t ygg_morton_sort_reorder_workspace_bytes(int64_t n);
int ygg_morton_sort_reorder(const float* pos, const float* mass, int64_t n, const float* bounds6,
                            uint64_t* sorted_codes, uint32_t* perm, float* pos_sorted,
                            float* mass_sorted, void* inverse, int inverse_i64, void* ws,
                            size_t ws_bytes, void* stream);
/* The same outputs, with the gather and the inverse on a second stream:
 * the encode and the code sort run on `stream`; `reorder_stream` waits for
 * them (a HIP event) and then writes pos_sorted, mass_sorted and inverse.
 * Work queued on `stream` after this call (the radix build, which reads
 * only the sorted codes and perm) overlaps the reorder. The caller orders
 * every reader of the reordered outputs after `reorder_stream` and keeps
 * `ws` and `perm` alive until `reorder_stream` passes this point. Same
 * workspace as ygg_morton_sort_reorder. */
int ygg_morton_sort_reorder_async(const float* pos, const float* mass, int64_t n,
                                  const float* bounds6, uint64_t* sorted_codes, uint32_t* perm,
                                  float* pos_sorted, float* mass_sorted, void* inverse,
                                  int inverse_i64, void* ws, size_t ws_bytes, void* stream,
                                  void* reorder_stream);

/* Gathers of the Morton-sorted particles (return_reordered of
 * yggdrax/_tree_impl.py:1045-1050): pos_sorted[i] = pos[perm[i]],
 * mass_sorted[i] = mass[perm[i]] (mass may be NULL), inverse[perm[i]] = i. */
int ygg_reorder(const float* pos, const float* mass, const uint32_t* perm, int64_t n,
                float* pos_sorted, float* mass_sorted, int32_t* inverse, void* stream);
/* The same outputs as ygg_reorder, computed by scattering: the inverse
 * permutation first, then each particle written to its sorted slot (reads
 * coalesced, writes scattered). inverse is required. */
int ygg_reorder_scatter(const float* pos, const float* mass, const uint32_t* perm, int64_t n,
                        float* pos_sorted, float* mass_sorted, int32_t* inverse, void* stream);

/* int32 -> int64 widening of up to YGG_WIDEN_MAX_ITEMS arrays in one launch:
 * dst[i] = src[i] for i < n of every item. It produces the public
 * topology's int64 index arrays (yggdrax/dtypes.py:12-40 index policy) from
 * the kernels' int32 arrays. Items with n == 0 are skipped; src and dst of
 * one item must not overlap. */
#define YGG_WIDEN_MAX_ITEMS 32
typedef struct YggWidenItem {
  const int32_t* src;
  int64_t* dst;
  int64_t n;
} YggWidenItem;
int ygg_widen_i32_to_i64(const YggWidenItem* items, int count, void* stream);

/* ---- radix (LBVH) topology ---------------------------------------------- */

/* Output arrays of the adaptive radix build for L = ceil(n/leaf_size) leaves
 * and T = 2L-1 nodes (internal nodes 0..L-2, leaf j is node L-1+j). */
typedef struct YggRadixTopology {
  int32_t* parent;         /* [T], -1 for the root */
  int32_t* left_child;     /* [L-1] */
  int32_t* right_child;    /* [L-1] */
  uint8_t* left_is_leaf;   /* [L-1] */
  uint8_t* right_is_leaf;  /* [L-1] */
  int32_t* node_ranges;    /* [T][2], inclusive particle range */
  int32_t* node_level;     /* [T] */
  int32_t* level_offsets;  /* [65] */
  int32_t* nodes_by_level; /* [T] */
  uint64_t* leaf_codes;    /* [L] */
  int32_t* scalars;        /* [8]: num_levels, max_level, deep_levels,
                              malformed, roots, root */
} YggRadixTopology;

/* _build_tree_from_leaf_partitions (yggdrax/_tree_impl.py:1009-1355) for the
 * fixed-count leaf partition of build_tree / build_octree's adaptive mode
 * (yggdrax/_tree_impl.py:405-406, yggdrax/tree.py:409-410). If
 * scalars[2] (deep_levels) is set on return, node levels exceeded 255 and
 * ygg_levels_pointer_doubling must be run before node_level, level_offsets
 * and nodes_by_level are valid. */
size_t ygg_radix_build_workspace_bytes(int64_t n, int64_t leaf_size);
int ygg_radix_build(const uint64_t* sorted_codes, int64_t n, int64_t leaf_size,
                    const YggRadixTopology* out, void* ws, size_t ws_bytes, void* stream);

/* Exact bounded pointer-doubling levels (yggdrax/_tree_impl.py:1272-1316)
 * for trees whose depth exceeds the fast path's walk (duplicate leaf
 * codes). Rewrites node_level, level_offsets, nodes_by_level, scalars[0..1]. */
int ygg_levels_pointer_doubling(const YggRadixTopology* topo, int64_t num_nodes, void* ws,
                                size_t ws_bytes, void* stream);

/* ---- explicit octree metadata -------------------------------------------- */

typedef struct YggOctreeMeta {
  int32_t* oct_parent;         /* [M], M = max(1, 2L-1) */
  int32_t* oct_children;       /* [M][8] */
  int32_t* oct_child_counts;   /* [M] */
  uint8_t* oct_child_mask;     /* [M][8] */
  uint8_t* oct_valid_mask;     /* [M] */
  uint64_t* oct_node_codes;    /* [M] */
  int32_t* oct_node_depths;    /* [M] */
  int32_t* oct_node_ranges;    /* [M][2] */
  int32_t* oct_nodes_by_level; /* [M] */
  int32_t* oct_level_offsets;  /* [23] */
  uint8_t* oct_leaf_mask;      /* [M] */
  int32_t* oct_leaf_nodes;     /* [M], -1 padded */
  int32_t* radix_node_to_oct;  /* [T]; radix_leaf_to_oct = tail L entries */
  int32_t* leaf_indices;       /* [L] native (octree) leaf order, node ids */
  int32_t* leaf_position;      /* [T] native row of a leaf node, else -1 */
  int32_t* scalars;            /* [4]: oct_num_levels, num_valid, num_leaf */
} YggOctreeMeta;

/* build_explicit_octree_metadata + augment_radix_topology_with_octree
 * (yggdrax/octree.py:204-419, 479-574) and the octree leaf ordering of
 * _resolve_leaf_ordering (yggdrax/_interactions_impl.py:1071-1103). */
size_t ygg_octree_meta_workspace_bytes(int64_t num_leaves);
int ygg_octree_meta(const uint64_t* sorted_codes, const int32_t* node_ranges, int64_t num_leaves,
                    const YggOctreeMeta* out, void* ws, size_t ws_bytes, void* stream);


/* ---- geometry ------------------------------------------------------------ */

/* compute_tree_geometry (yggdrax/geometry.py:18-75,
 * yggdrax/_geometry_impl.py:176-338): per-node AABB centre, half extent
 * (clamped to >= 1e-6), bounding-sphere radius and max half extent from the
 * Morton-sorted positions. leaf_size is the fixed-count leaf partition
 * (leaf j = particles [j*leaf_size, ...)); leaf_cap bounds the particles read
 * per leaf (max_leaf_size). level_sync = 0 takes each internal box as a
 * segment-tree range reduction over its leaves (valid for well-formed trees,
 * YggRadixTopology.scalars[3] == 0); level_sync = 1 runs the reference's
 * level-by-level pass using node_level and max_level. particle_radius
 * (float32 [n] in sorted order, or NULL) bounds each particle as the cube
 * [p - r, p + r] instead of the point p (the particle_radius argument,
 * yggdrax/_geometry_impl.py:107-112). */
size_t ygg_geometry_workspace_bytes(int64_t num_nodes);
int ygg_geometry(const float* pos_sorted, const float* particle_radius,
                 const int32_t* node_ranges, const int32_t* left,
                 const int32_t* right, const int32_t* parent, const int32_t* node_level,
                 int64_t num_nodes, int64_t leaf_size, int64_t leaf_cap, int level_sync,
                 int32_t max_level,
                 float* center, float* half_extent, float* radius, float* max_extent, void* ws,
                 size_t ws_bytes, void* stream);

/* Particle-order leaf table of a radix tree (_resolve_leaf_ordering,
 * yggdrax/_interactions_impl.py:1071-1103 without octree metadata). */
int ygg_radix_leaf_order(int64_t num_leaves, int32_t* leaf_indices, int32_t* leaf_position,
                         void* stream);

/* geometry_to_level_major (yggdrax/_geometry_impl.py:342-433): padded
 * [num_levels][max_nodes] level-major copies of the node geometry; slot j of
 * level l holds node nodes_by_level[level_offsets[l] + j] while j is below
 * the level's count, else zeros and node index -1. */
int ygg_geometry_level_major(const float* center, const float* half_extent, const float* radius,
                             const float* max_extent, const int32_t* nodes_by_level,
                             const int32_t* level_offsets, int32_t num_levels, int64_t max_nodes,
                             float* out_center, float* out_half_extent, float* out_radius,
                             float* out_max_extent, void* out_node_indices, int out_i64,
                             void* stream);

/* ---- downstream far-field layouts ----------------------------------------- */

/* _dense_m2l_buffers (yggdrax/dense_interactions.py:35-140): for every
 * level-major slot (l, j) and entry k < max_interactions, the k-th far source
 * of the slot's target node (or -1), its validity, and the target-minus-source
 * centre displacement (zeros when invalid); out_counts[l][j] is the target's
 * far count (0 for padding slots). offsets [num_nodes (+1)], counts
 * [num_nodes] and sources [num_far] are the level-major CSR far list in the
 * index dtype (int64 when idx_i64, else int32). Outputs: out_sources and
 * out_mask [num_levels][max_nodes][max_interactions], out_displacements
 * [...][3], out_counts [num_levels][max_nodes]. A row outside the sources or
 * a source outside [0, num_nodes) is written as padding and sets
 * *error_flag (device int32, cleared by the call) to 1. */
int ygg_dense_interactions(const void* offsets, const void* counts, const void* sources,
                           int idx_i64, int64_t num_nodes, int64_t num_far,
                           const int32_t* nodes_by_level, const int32_t* level_offsets,
                           int32_t num_levels, int64_t max_nodes, int64_t max_interactions,
                           const float* center, void* out_sources, float* out_displacements,
                           uint8_t* out_mask, void* out_counts, int32_t* error_flag,
                           void* stream);

/* build_grouped_interactions_from_pairs (yggdrax/grouped_interactions.py:
 * 74-207), split in two calls because the class count sizes the outputs.
 * ygg_grouped_classify computes each pair's class key (target level, source
 * level, rint((c_t - c_s) / cell(target level)) per axis, in float64, with
 * out-of-range values as INT32_MIN as NumPy's cast gives them), sorts the
 * pairs stably by key into `order` [num_pairs], writes class_scan
 * [num_pairs + 1] (classes started before each sorted position) and returns
 * the class count in *num_classes (host memory). It blocks on the stream.
 * num_pairs must be 1 .. 2^30 - 1; a pair naming a node outside
 * [0, num_nodes) makes it return YGG_ERR_ARG before anything reads it. */
size_t ygg_grouped_workspace_bytes(int64_t num_pairs);
int ygg_grouped_classify(const void* sources, const void* targets, int64_t num_pairs,
                         int idx_i64, int64_t num_nodes, const int32_t* node_level,
                         const float* center,
                         const float* bounds_min, const float* bounds_max, uint32_t* order,
                         int64_t* class_scan, int64_t* num_classes, void* ws, size_t ws_bytes,
                         void* stream);
/* ygg_grouped_emit writes the class-major outputs from classify's order,
 * class_scan and what classify left in `ws` (the sorted packed keys and the
 * node pairs; pass the same, unmodified workspace): class_sources and
 * class_targets [num_pairs] in sorted order, class_ids [num_pairs] in input
 * order, class_offsets [num_classes + 1], class_keys [num_classes][5] and
 * class_displacements [num_classes][3] (index * cell, float64 rounded to
 * float32). */
int ygg_grouped_emit(int64_t num_pairs, int idx_i64, const float* bounds_min,
                     const float* bounds_max, const uint32_t* order, const int64_t* class_scan,
                     void* class_sources, void* class_targets, void* class_ids,
                     void* class_offsets, int32_t* class_keys, float* class_displacements,
                     void* ws, size_t ws_bytes, void* stream);

/* ---- node mass moments ---------------------------------------------------- */

/* compute_tree_mass_moments (yggdrax/tree_moments.py:686-746): mass_out[t]
 * = sum of the masses of node t's particle range node_ranges[t] (inclusive)
 * and center_out[t] = (sum m x) / mass, 0 where that mass is 0, over
 * Morton-sorted positions [n][3] and masses [n]. The per-particle terms are
 * m * x rounded to the product dtype; the sums are accumulated in float64
 * (chunk sums of `chunk` particles in a segment tree plus the partial
 * chunks of each range) and rounded once. Dtypes: mass_out has the masses'
 * dtype, center_out is float64 if either input is (pos_f64, mass_f64 flags),
 * else float32. */
size_t ygg_tree_mass_moments_workspace_bytes(int64_t n, int64_t chunk);
int ygg_tree_mass_moments(const void* pos_sorted, int pos_f64, const void* mass_sorted,
                          int mass_f64, int64_t n, const int32_t* node_ranges, int64_t num_nodes,
                          int64_t chunk, void* mass_out, void* center_out, void* ws,
                          size_t ws_bytes, void* stream);

/* ---- dual-tree traversal -------------------------------------------------- */

#define YGG_MAC_BH 0
#define YGG_MAC_DEHNEN 1
#define YGG_MAC_ENGBLOM 2

#define YGG_WALK_COUNT 0   /* counts only (the count pass)                     */
#define YGG_WALK_FIXED 1   /* rows of fixed capacity: row r at r * cap          */
#define YGG_WALK_OFFSETS 2 /* rows at row_start[r] with capacity row_cap[r]     */

/* Node records for the walk: 32 bytes per node (centre, MAC extent of
 * _build_mac_extents, yggdrax/_interactions_impl.py:993-1061, children,
 * native leaf row, level-major rank from nodes_by_level). */
int ygg_walk_prepare_nodes(const float* center, const float* radius, const float* max_extent,
                           const int32_t* left, const int32_t* right,
                           const int32_t* leaf_position, const int32_t* nodes_by_level,
                           int64_t num_nodes, int mac_type, float dehnen_radius_scale,
                           void* nodes_out, void* stream);

typedef struct YggWalkParams {
  int32_t num_nodes, num_internal, num_leaves, root;
  int32_t mac_type, mode, collect_far, collect_near;
  float theta_sq, theta_root; /* float32(theta)^2 and its sqrt (engblom) */
  int32_t far_cap, near_cap;  /* fixed row capacities (YGG_WALK_FIXED) */
  int64_t queue_cap;          /* pair-queue (wavefront) capacity */
  int32_t max_generations, pad0;
  const void* nodes;          /* [num_nodes] records from ygg_walk_prepare_nodes */
  uint64_t* wf0;              /* [queue_cap] wavefront ping buffer */
  uint64_t* wf1;              /* [queue_cap] wavefront pong buffer */
  uint32_t* gen_sizes;        /* [max_generations + 1] raw push counts */
  uint32_t* flags;            /* [4]: queue, far, near overflow */
  uint64_t* decisions;        /* [3]: accept, near, refine */
  int32_t* far_counts;        /* [num_nodes] */
  int32_t* near_counts;       /* [num_leaves], native leaf rows */
  const int64_t* far_row_start;
  const int32_t* far_row_cap;
  int32_t* far_entries;
  const int64_t* near_row_start;
  const int32_t* near_row_cap;
  int32_t* near_entries;
} YggWalkParams;

/* One fixed-capacity attempt of _dual_tree_walk_impl
 * (yggdrax/_interactions_impl.py:1131-1785), generation by generation from
 * the pair (root, root). Blocks on `stream` between batches of generations
 * to read the wavefront sizes into host_status (pinned host memory of
 * max_generations + 5 words: gen sizes, then the 4 flags). If gen_ms is not
 * NULL (host array of max_generations floats, max_generations <= 1024), the
 * duration of every generation's kernel is measured with HIP events
 * recorded on `stream` around the launch. */
int ygg_walk_bfs(const YggWalkParams* params, uint32_t* host_status, int32_t* num_generations,
                 float* gen_ms, void* stream);

/* Exclusive prefix sum of int32 counts (optionally gathered through
 * `gather`) into int64 offsets[n + 1]. */
size_t ygg_scan_workspace_bytes(int64_t n);
int ygg_exclusive_scan_i32(const int32_t* counts, const int32_t* gather, int64_t* offsets,
                           int64_t n, void* ws, size_t ws_bytes, void* stream);

/* row_start[order[r]] = offsets[r] (order NULL: identity). */
int ygg_rows_from_offsets(const int32_t* order, const int64_t* offsets, int64_t n,
                          int64_t* row_start, void* stream);

/* counts[i] = min(counts[i], caps ? caps[i] : fixed_cap). */
int ygg_cap_counts(int32_t* counts, int64_t n, int32_t fixed_cap, const int32_t* caps,
                   void* stream);

/* Far list in level-major CSR (_result_to_interactions,
 * yggdrax/_interactions_impl.py:3561-3637): rows taken in nodes_by_level
 * order from entries (row r at row_start[r], or r * fixed_cap when
 * row_start is NULL). Writes sources/targets/target_levels[F] and
 * node_offsets[T + 1] as int64 (out_i64) or int32. */
int ygg_far_output(int64_t num_nodes, const int32_t* nodes_by_level, const int64_t* lm_offsets,
                   const int32_t* counts, int32_t fixed_cap, const int64_t* row_start,
                   const int32_t* entries, const int32_t* node_level, int out_i64,
                   void* sources, void* targets, void* target_levels, void* node_offsets,
                   void* stream);

/* Near list CSR in native leaf order plus the dense -1 padded
 * neighbor_leaf_positions[L][max_count] (_result_to_neighbors,
 * yggdrax/_interactions_impl.py:3640-3837). */
/* Near-field target blocks (_build_target_blocks in _result_to_neighbors,
 * yggdrax/_interactions_impl.py:3689-3770, enabled there by
 * YGGDRAX_NEARFIELD_TARGET_BLOCK_SIZE > 0): each native leaf row of the near
 * CSR list (near_offsets[L + 1]) is cut into ceil(count / block_size)
 * blocks of block_size consecutive neighbours. ygg_target_block_offsets
 * writes the int64 exclusive prefix of the blocks per row
 * (block_offsets[L + 1]; block_offsets[L] = number of blocks B). Then
 * ygg_target_blocks writes block_leaf_ids[B] (the row of each block),
 * block_source_leaf_ids[B * block_size] (native leaf row of each neighbour,
 * 0 past the row's end) and valid_mask[B * block_size]; neighbors and the
 * outputs are int64 (in_i64 / out_i64) or int32. */
size_t ygg_target_block_offsets_workspace_bytes(int64_t num_leaves);
int ygg_target_block_offsets(int64_t num_leaves, const int64_t* near_offsets, int64_t block_size,
                             int64_t* block_offsets, void* ws, size_t ws_bytes, void* stream);
int ygg_target_blocks(int64_t num_leaves, const int64_t* near_offsets, const void* neighbors,
                      int in_i64, const int32_t* leaf_position, int64_t block_size,
                      const int64_t* block_offsets, int64_t num_blocks, int out_i64,
                      void* block_leaf_ids, void* block_source_leaf_ids, uint8_t* valid_mask,
                      void* stream);

int ygg_near_output(int64_t num_leaves, const int64_t* offsets, const int32_t* counts,
                    int32_t fixed_cap, const int64_t* row_start, const int32_t* entries,
                    const int32_t* leaf_position, int64_t max_count, int out_i64,
                    void* neighbors, void* neighbor_leaf_positions, void* stream);

/* A pair policy given as data, evaluated on the device for both directions
 * of every pair (the PairPolicy contract of _resolve_pair_actions,
 * yggdrax/_interactions_impl.py:731-810, for the policies it can describe).
 * accept: YGG_ACCEPT_MAC (the built-in MAC of the walk), _DISTANCE
 * (different nodes with dist_sq >= accept_dist_sq), _DIFFERENT (any two
 * different nodes), _NEVER. near: YGG_NEAR_LEAF_PAIRS (two different leaves
 * not accepted, the default rule of _default_pair_actions, 694-711) or
 * _NEVER. Accepted pairs get a tag, every other pair -1: tag_metric
 * _DIST_SQ or _OPENING_RATIO ((e_t + e_s)^2 / dist_sq) gives
 * tag_values[k], k = number of thresholds <= the metric; _TARGET_MOD gives
 * target % tag_modulus; _NONE gives -1. */
#define YGG_ACCEPT_MAC 0
#define YGG_ACCEPT_DISTANCE 1
#define YGG_ACCEPT_DIFFERENT 2
#define YGG_ACCEPT_NEVER 3
#define YGG_NEAR_LEAF_PAIRS 0
#define YGG_NEAR_NEVER 1
#define YGG_TAG_NONE 0
#define YGG_TAG_DIST_SQ 1
#define YGG_TAG_OPENING_RATIO 2
#define YGG_TAG_TARGET_MOD 3
#define YGG_POLICY_MAX_THRESHOLDS 8
typedef struct YggPolicyDesc {
  int32_t accept;
  int32_t near;
  int32_t tag_metric;
  int32_t num_thresholds; /* 0 .. YGG_POLICY_MAX_THRESHOLDS */
  float accept_dist_sq;
  int32_t tag_modulus;
  float thresholds[YGG_POLICY_MAX_THRESHOLDS]; /* ascending */
  int32_t tag_values[YGG_POLICY_MAX_THRESHOLDS + 1];
} YggPolicyDesc;

/* Flat-emission walk: the same pair semantics as ygg_walk_bfs over the full
 * walk, with workgroup-aggregated appends to three streams: the next
 * wavefront, far entries (key = level-major rank of the target row, value =
 * source node) and near entries (key = native leaf row, value = neighbour
 * leaf node). emit = 0 only counts (entry totals and the exact wavefront
 * size of every generation; the accept / near / refine decision counts
 * follow as far/2, near/2 and the remaining evaluated pairs). A stream that would exceed its capacity sets its
 * flag (queue, far, near) and the walk stops after that generation.
 * host_status (pinned host memory of max_generations + 9 words) receives the
 * generation sizes and the 4 flags as for ygg_walk_bfs, then the far and
 * near totals as two 64-bit words (also left in counters[0..1]); gen_ms
 * (optional per-generation kernel times from HIP events on `stream`) is as
 * for ygg_walk_bfs. */
#define YGG_WALK_GEN_RECORD_WORDS 16
typedef struct YggFlatWalkParams {
  int32_t num_nodes, num_internal, num_leaves, root;
  int32_t mac_type, emit, max_generations, grid_blocks;
  float theta_sq, theta_root;
  int32_t gen_hint;           /* expected generation count (first launch batch), 0 = 24 */
  int32_t pairs_per_thread;   /* queue tile: pairs per thread of a chunk (4, 8 or 16) */
  int64_t queue_cap;          /* pairs per wavefront buffer */
  int64_t far_cap, near_cap;  /* entries per flat stream */
  const void* nodes;
  uint64_t* wf0;
  uint64_t* wf1;
  uint32_t* gen_sizes;        /* [max_generations + 1] */
  uint32_t* flags;            /* [4] */
  uint64_t* counters;         /* [2]: far total, near total */
  uint32_t* far_keys;
  uint32_t* far_vals;
  uint32_t* near_keys;
  uint32_t* near_vals;
  int32_t hist_gen;           /* node_hist[t] counts generation hist_gen's pairs by target */
  int32_t xcd_groups;         /* > 1: split each generation's chunks into that many contiguous
                                 ranges, one per XCD (workgroups dispatch round-robin) */
  int32_t* node_hist;         /* [num_nodes] or NULL */
  const int32_t* roots;       /* [num_roots] seed pairs (r, r) of a forest, or NULL: (root, root) */
  int64_t num_roots;
  uint64_t* spread_counters;  /* optional [2 * spread_lanes]: the counting walk (emit 0) adds its
                                 far/near totals here, one lane per group of workgroups, instead
                                 of reserving per chunk; the walk then writes their sums to
                                 counters[0..1]. NULL: totals go straight to counters. */
  int64_t spread_lanes;
  const YggPolicyDesc* policy; /* optional: accept/near rules of a pair-policy descriptor for both
                                  walks (NULL: the built-in MAC, near for two different leaves) */
  int32_t* far_tags;          /* emission walk with a policy: [far_cap] tag of each far entry */
  uint64_t* gen_fill;         /* optional, emission walk only: [max_generations *
                                 YGG_WALK_GEN_RECORD_WORDS] per-generation records (entry counts
                                 far << 32 | near, and the stream offsets the generation starts
                                 at), which make the walk resumable.
                                 Generation g writes after generations 0..g-1 and runs only if two
                                 entries per pair fit the remaining far and near capacity; if not,
                                 it writes nothing, flags[3] = g + 1, *num_generations = g, the
                                 counters hold the totals of generations 0..g-1, and a call with
                                 start_generation = g and larger streams (their first
                                 counters[0] / counters[1] entries copied over) continues the
                                 walk. NULL: overflowing entries set flags[1] / flags[2]. */
  int32_t start_generation;   /* 0: a new walk; g > 0: resume a paused walk at generation g (the
                                 other device buffers as the paused call left them) */
  int32_t pair_records;       /* emission walk without a policy: nonzero writes one record per
                                 accepted or near pair, the node ids (t, s), in place of the
                                 pair's two entries (t's row key, s) and (s's row key, t).
                                 far_cap / near_cap, the counters, the gen_fill counts and the
                                 overflow flags then count records. ygg_far_lists_sort /
                                 ygg_near_lists_sort list them (pair_records / leaf_nodes). */
  const uint32_t* gen_pairs_hint; /* optional, host memory: [max_generations] pairs expected in
                                 each generation (e.g. a previous walk's sizes on a similar
                                 tree). It sizes each generation's grid (1.5x the hinted chunks,
                                 at least 256 workgroups, at most grid_blocks) and so only the
                                 launch cost: any grid covers a generation, its workgroups
                                 stepping over the chunks. 0xFFFFFFFF for a generation, or
                                 NULL for all: grids from the bound roots * 4^g alone. */
} YggFlatWalkParams;

int ygg_walk_flat(const YggFlatWalkParams* params, uint32_t* host_status,
                  int32_t* num_generations, float* gen_ms, void* stream);
/* The same walk as exactly max_generations generation launches with no host
 * synchronisation, so a caller can capture it into a HIP graph. Generation
 * sizes, counters and overflow flags stay in device memory (gen_sizes,
 * counters, flags); a wavefront still non-empty after the last launch is
 * left in gen_sizes[max_generations] for the caller to detect. */
int ygg_walk_flat_fixed(const YggFlatWalkParams* params, void* stream);

/* offsets[r] = lower_bound(sorted_keys, r) for r in [0, rows]. */
int ygg_row_offsets(const uint32_t* sorted_keys, int64_t n, int64_t rows, int64_t* offsets,
                    void* stream);

/* Far lists from key-sorted flat entries (level-major CSR of
 * _result_to_interactions, yggdrax/_interactions_impl.py:3561-3637):
 * sources/targets/target_levels[n], counts[T] and node_offsets[T + 1] in the
 * output index type, plus int32 counts32[T]. */
int ygg_far_lists_from_sorted(int64_t num_nodes, const uint32_t* keys, const uint32_t* vals,
                              int64_t n, const int64_t* lm_offsets, const int32_t* nodes_by_level,
                              const int32_t* node_level, int out_i64, void* sources,
                              void* targets, void* target_levels, void* counts,
                              void* node_offsets, int32_t* counts32, void* stream);

/* Near lists from row-sorted flat entries (_result_to_neighbors,
 * yggdrax/_interactions_impl.py:3640-3837): neighbors[n], the dense
 * neighbor_leaf_positions[L][max_count] (may be NULL), counts[L] and
 * offsets[L + 1] in the output index type. */
int ygg_near_lists_from_sorted(int64_t num_leaves, const int64_t* row_offsets,
                               const uint32_t* vals, const int32_t* leaf_position,
                               int64_t max_count, int out_i64, void* neighbors,
                               void* neighbor_leaf_positions, void* counts, void* offsets,
                               void* stream);

/* The far and near lists grouped by row straight from the walk's streams:
 * every row's values contiguous, in row order, each row's count from the
 * grouping; row offsets (int64 lm_offsets[T + 1] / row_offsets[L + 1]) are
 * the scan of those counts, and the far targets and target levels are
 * filled row by row. Same outputs as ygg_sort_pairs_u32 + ygg_row_offsets +
 * ygg_far_lists_from_sorted / ygg_near_lists_from_sorted, except the order
 * inside a row, which is not specified (the walk emits in atomic order).
 * digit_bits 0 with end_bit <= 22 groups by bucketing: a split by the row
 * key's high digit, then per-bucket counts and a scatter by the low digit,
 * ranked with LDS atomics. digit_bits 8..11 (or end_bit > 22) uses the
 * stable radix sort with that digit width (0: its default), whose last pass
 * writes the lists and counts runs of equal keys. Keys are row keys below
 * 2^end_bit, with rows (T or L) <= 2^end_bit; keys >= rows are unused
 * capacity and are listed nowhere. Values are node indices: below
 * num_nodes for the far lists and below 2 * num_leaves for the near lists
 * (the bucketing packs them with the low digit). key_scratch[n] is scratch
 * (its contents afterwards are unspecified). keys and vals are not
 * modified. `count` (optional, device memory) holds the number of valid
 * entries of streams allocated at capacity n: only the first
 * min(n, *count) are listed, with no host read (for graph capture).
 * `row_sub` (optional, device memory, [rows]) is taken off every value of
 * row r, and off the far target of row r: ids local to a block of nodes,
 * as for the systems of a batched forest (bucketing only: with the stable
 * sort it returns YGG_ERR_UNSUPPORTED).
 * Pair records (bucketing only): with pair_records nonzero (far) or
 * leaf_nodes given (near, device memory, [num_leaves]: the node of each
 * leaf row, i.e. leaf_indices), keys and vals hold n records (t, s) of
 * node ids from the walk's pair_records mode, each listing s in t's row
 * and t in s's row. Far rows are taken in level-major order through
 * nodes_by_level, near rows in leaf order through leaf_nodes; the near
 * keys then span end_bit = the node ids' bits. `count` counts records, the
 * outputs and key_scratch hold 2n entries, the workspace is
 * ygg_*_lists_sort_workspace_bytes(2n, ...), and row_sub is indexed by
 * node id, over the whole key range (num_nodes entries for the far lists,
 * 2^end_bit for the near lists). The far targets still take row_sub[rank]:
 * a batched forest's per-system offsets are the same array, each system's
 * nodes and ranks sharing one index range. */
size_t ygg_far_lists_sort_workspace_bytes(int64_t n, int end_bit, int digit_bits);
int ygg_far_lists_sort(const uint32_t* keys, const uint32_t* vals, int64_t n,
                       const int64_t* count, int64_t num_nodes, int end_bit, int digit_bits,
                       const int32_t* nodes_by_level, const int32_t* node_level,
                       const int64_t* row_sub, int pair_records, int out_i64,
                       uint32_t* key_scratch, int64_t* lm_offsets, void* sources, void* targets,
                       void* target_levels, void* counts, void* node_offsets, int32_t* counts32,
                       void* ws, size_t ws_bytes, void* stream);
size_t ygg_near_lists_sort_workspace_bytes(int64_t n, int end_bit, int digit_bits);
int ygg_near_lists_sort(const uint32_t* keys, const uint32_t* vals, int64_t n,
                        const int64_t* count, int64_t num_leaves, int end_bit, int digit_bits,
                        const int64_t* row_sub, const int32_t* leaf_nodes, int out_i64,
                        uint32_t* key_scratch, int64_t* row_offsets, void* neighbors,
                        void* counts, void* offsets, void* ws, size_t ws_bytes, void* stream);
/* The dense -1 padded [L][max_count] neighbour leaf rows from a near list
 * (neighbors int64 if out_i64, else int32). */
int ygg_near_dense_from_list(int64_t num_leaves, const int64_t* row_offsets,
                             const void* neighbors, const int32_t* leaf_position,
                             int64_t max_count, int out_i64, void* neighbor_leaf_positions,
                             void* stream);

/* ---- learned capacity model -------------------------------------------- */

/* 8 -> 64 -> 64 -> 64 -> 3 ReLU MLP, bfloat16 weights, float32 accumulation.
 * No counterpart in the reference: it sizes the walk's streams in one pass
 * (the reference counts first or retries on overflow). Parameters are laid
 * out W1[64][8] b1 W2[64][64] b2 W3[64][64] b3 W4[3][64] b4. */
int ygg_mlp_num_params(void);
int ygg_mlp_forward(const uint16_t* params_bf16, const float* features, int64_t n,
                    float* outputs, void* stream);
int64_t ygg_mlp_train_workspace_floats(int64_t batch);
/* One Adam step (lr > 0) or gradient only (lr == 0) of the masked Poisson
 * negative log-likelihood (rate 2^y, target = count / per-output scale) on a
 * batch; grads receives dLoss/dparams (float32, parameter layout). */
int ygg_mlp_train_step(float* params, uint16_t* params_bf16, float* adam_m, float* adam_v,
                       int64_t step, float lr, const float* x, const float* target,
                       const float* mask, int64_t batch, float* grads, float* ws,
                       float* loss_out, void* stream);
int ygg_mlp_params_to_bf16(const float* params, uint16_t* params_bf16, void* stream);
/* Per-node features: level, particle count, max half extent, radius, parent
 * far count (may be NULL), sibling particle count, density, leaf flag. */
int ygg_mlp_features(int64_t num_nodes, const int32_t* node_level, const int32_t* node_ranges,
                     const int32_t* parent, const int32_t* left, const int32_t* right,
                     const float* half_extent, const float* radius, const float* max_extent,
                     const float* parent_far, float* features, void* stream);
/* totals[3] (double) = per-output sums of scale_j * 2^y_j (near: leaves only). */
int ygg_mlp_totals(const float* outputs, int64_t num_nodes, float scale_far, float scale_near,
                   float scale_queue, double* totals, void* stream);
/* far_counts[k] = scale_far * 2^y[k][0]: predicted far count of every node. */
int ygg_mlp_far_counts(const float* outputs, int64_t num_nodes, float scale_far,
                       float* far_counts, void* stream);

/* ---- batched mode: many independent systems per launch ------------------ */

/* A batch of S systems whose points are concatenated: system s owns points
 * [point_offsets[s], point_offsets[s+1]) and leaves [leaf_offsets[s],
 * leaf_offsets[s+1]) (leaf_offsets[s+1] - leaf_offsets[s] =
 * ceil(n_s / leaf_size)); both tables live in device memory. Per-system
 * outputs are laid out back to back: internal-node arrays at
 * leaf_offsets[s] - s, node and octree-slot arrays at 2 leaf_offsets[s] - s,
 * level_offsets at 65 s, oct_level_offsets at 23 s, radix scalars at 8 s,
 * octree scalars at 4 s. Every stored index is local to its system, so
 * each slice equals what the single-system entry point returns for that
 * system alone. No reference counterpart: the reference builds one tree per
 * call (yggdrax/tree.py:1482-1549); this is the batched mode of the same
 * computation. */
typedef struct YggBatch {
  int32_t num_systems, pad0;
  int64_t leaf_size;
  const int32_t* point_offsets;  /* [S + 1], device */
  const int32_t* leaf_offsets;   /* [S + 1], device */
  int64_t num_points;            /* point_offsets[S] */
  int64_t num_leaves;            /* leaf_offsets[S] */
  int64_t max_leaves_per_system;
} YggBatch;

/* Per-system infer_bounds (yggdrax/bounds.py:11-33): bounds6[6 s + k]. */
int ygg_batch_infer_bounds(const YggBatch* batch, const float* pos, float* bounds6, void* stream);
/* Morton codes with each system's bounds, stably sorted by (system, code):
 * perm holds global point indices. */
size_t ygg_batch_morton_sort_workspace_bytes(const YggBatch* batch);
int ygg_batch_morton_sort(const YggBatch* batch, const float* pos, const float* bounds6,
                          uint64_t* sorted_codes, uint32_t* perm, void* ws, size_t ws_bytes,
                          void* stream);
/* Reordered positions/masses, per-system particle_indices and inverse
 * permutation (local indices). */
int ygg_batch_reorder(const YggBatch* batch, const float* pos, const float* mass,
                      const uint32_t* perm, float* pos_sorted, float* mass_sorted,
                      int32_t* perm_local, int32_t* inverse_local, void* stream);
/* ygg_radix_build for every system (scalars [8 S]). Systems whose
 * scalars[8 s + 2] (deep levels) is set need the single-system path. */
size_t ygg_batch_radix_build_workspace_bytes(const YggBatch* batch);
int ygg_batch_radix_build(const YggBatch* batch, const uint64_t* sorted_codes,
                          const YggRadixTopology* out, void* ws, size_t ws_bytes, void* stream);
/* ygg_octree_meta for every system (scalars [4 S]). */
size_t ygg_batch_octree_meta_workspace_bytes(const YggBatch* batch);
int ygg_batch_octree_meta(const YggBatch* batch, const uint64_t* sorted_codes,
                          const int32_t* node_ranges, const YggOctreeMeta* out, void* ws,
                          size_t ws_bytes, void* stream);
/* ygg_radix_leaf_order for every system. */
int ygg_batch_radix_leaf_order(const YggBatch* batch, int32_t* leaf_indices,
                               int32_t* leaf_position, void* stream);
/* compute_tree_geometry for every (well-formed) system: segment tree over
 * all leaves of the batch, one range query per internal node. */
size_t ygg_batch_geometry_workspace_bytes(const YggBatch* batch);
int ygg_batch_geometry(const YggBatch* batch, const float* pos_sorted, const int32_t* node_ranges,
                       int64_t leaf_cap, float* center, float* half_extent, float* radius,
                       float* max_extent, void* ws, size_t ws_bytes, void* stream);
/* Walk node records of every system with batch-global child ids, leaf rows
 * (leaf_offsets[s] + row) and level-major ranks (node offset + rank), so one
 * walk over the forest (YggFlatWalkParams.roots = every system's root)
 * evaluates every system's pairs. */
int ygg_batch_walk_prepare_nodes(const YggBatch* batch, const float* center, const float* radius,
                                 const float* max_extent, const int32_t* left,
                                 const int32_t* right, const int32_t* leaf_position,
                                 const int32_t* nodes_by_level, int mac_type,
                                 float dehnen_radius_scale, void* nodes_out, int32_t* roots,
                                 void* stream);
/* Forest lists -> per-system lists. Far: sources/targets (batch-global node
 * ids, in place) become local; offsets_out[node_off(s) + s + k] (T_s + 1 per
 * system) = the system's level-major CSR offsets; far_level_offsets_out
 * [65 s + l] from the system's radix level_offsets. Near: neighbors (global
 * leaf node ids, in place) become local; near_offsets_out[leaf_offsets[s] +
 * s + j] (L_s + 1 per system). row_offsets are the sorted streams' offsets
 * over batch-global rows (level-major ranks / leaf rows). num_entries 0
 * leaves the entries as they are: lists grouped with row_sub (each row's
 * system node offset) are local already, and only the offsets are
 * written. */
int ygg_batch_far_localize(const YggBatch* batch, const int64_t* row_offsets,
                           const int32_t* nodes_by_level_global, const int32_t* level_offsets,
                           int64_t num_entries, int out_i64, void* sources, void* targets,
                           void* node_offsets_out, void* far_level_offsets_out, void* stream);
int ygg_batch_near_localize(const YggBatch* batch, const int64_t* row_offsets,
                            int64_t num_entries, int out_i64, void* neighbors,
                            void* near_offsets_out, void* stream);

/* Capacity-model features of every node of every system of a batch (each
 * with its own tree's root), and batch totals of the predicted rates. */
int ygg_batch_mlp_features(const YggBatch* batch, const int32_t* node_level,
                           const int32_t* node_ranges, const int32_t* parent, const int32_t* left,
                           const int32_t* right, const float* half_extent, const float* radius,
                           const float* max_extent, const float* parent_far, float* features,
                           void* stream);
int ygg_batch_mlp_totals(const YggBatch* batch, const float* outputs, float scale_far,
                         float scale_near, float scale_queue, double* totals, void* stream);

/* ---- custom pair policies ------------------------------------------------- */

/* Both directions' actions (0 accept, 1 near, 2 refine) and tags of one
 * generation's pairs under a policy descriptor, for ygg_policy_apply. */
int ygg_policy_descriptor_actions(const void* nodes, const uint64_t* pairs, int64_t n,
                                  int64_t num_internal, int mac_type, float theta_sq,
                                  float theta_root, const YggPolicyDesc* desc,
                                  int32_t* actions_fwd, int32_t* actions_rev, int32_t* tags_fwd,
                                  int32_t* tags_rev, void* stream);

/* Geometry of one generation's pairs for a host-evaluated pair policy
 * (PairPolicy keyword arrays of _resolve_pair_actions,
 * yggdrax/_interactions_impl.py:731-810). */
int ygg_policy_geometry(const void* nodes, const uint64_t* pairs, int64_t n, int64_t num_internal,
                        int mac_type, float theta_sq, float theta_root, int32_t* targets,
                        int32_t* sources, float* center_target, float* center_source,
                        float* dist_sq, float* extent_target, float* extent_source,
                        uint8_t* mac_ok, uint8_t* target_leaf, uint8_t* source_leaf,
                        void* stream);
/* Applies both directions' actions and tags: accept only if both accept,
 * near only if both say near, else refine; appends tagged far entries, near
 * entries and child pairs (pairs_out at *out_size, reserved atomically). */
int ygg_policy_apply(const void* nodes, const uint64_t* pairs, int64_t n, int64_t num_internal,
                     const int32_t* actions_fwd, const int32_t* actions_rev,
                     const int32_t* tags_fwd, const int32_t* tags_rev, uint64_t* pairs_out,
                     int64_t queue_cap, uint32_t* out_size, uint32_t* far_keys,
                     uint32_t* far_vals, int32_t* far_tags, int64_t far_cap,
                     uint32_t* near_keys, uint32_t* near_vals, int64_t near_cap,
                     uint64_t* counters, uint32_t* flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* YGGDRAX_AMD_H */
