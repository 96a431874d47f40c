// Shared device helpers for the yggdrax MI355X (gfx950) kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/yggdrax_amd.h"

#define YGG_WAVE 64

#define YGG_CHECK(expr)                                   \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return YGG_ERR_HIP;             \
  } while (0)

#define YGG_LAUNCH_CHECK() YGG_CHECK(hipGetLastError())

namespace ygg {

__device__ __forceinline__ int clz64(uint64_t x) { return x == 0 ? 64 : __clzll((long long)x); }

__device__ __forceinline__ uint64_t spread3(uint64_t x) {
  x &= 0x1FFFFFull;
  x = (x | (x << 32)) & 0x1F00000000FFFFull;
  x = (x | (x << 16)) & 0x1F0000FF0000FFull;
  x = (x | (x << 8)) & 0x100F00F00F00F00Full;
  x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Common octree depth of two Morton codes (floor((leading common bits - 1)/3)).
__device__ __forceinline__ int common_depth(uint64_t a, uint64_t b) {
  uint64_t x = a ^ b;
  int bits = (x == 0) ? 63 : max(clz64(x) - 1, 0);
  return bits / 3;
}

__device__ __forceinline__ uint64_t prefix_code(uint64_t c, int d) {
  if (d <= 0) return 0;
  int s = max(0, 63 - 3 * d);
  return (c >> s) << s;
}

__device__ __forceinline__ int octant_at_depth(uint64_t c, int d) {
  if (d <= 0) return 0;
  int s = max(0, 63 - 3 * d);
  return (int)((c >> s) & 7ull);
}

// (8^d - 1)/7 + (code >> (63 - 3d)): sortable key of an octree cell.
__device__ __forceinline__ uint64_t oct_address(uint64_t c, int d) {
  uint64_t du = (uint64_t)max(0, d);
  uint64_t shift = 63ull - 3ull * du;
  uint64_t pre = c >> shift;
  uint64_t p8 = 1ull << (3ull * du);
  return (p8 - 1ull) / 7ull + pre;
}

// Fixed-round binary search returning `high` after bit_length(n) halvings
// of [low, high), comparing q <= a[mid] (left) or q < a[mid] (right). On a
// sorted array this is lower/upper bound; it is also well defined on the
// unsorted resolved leaf-cell codes the octree ranges search.
__device__ __forceinline__ int fixed_round_search(const uint64_t* a, int n, uint64_t q, bool left) {
  int levels = 32 - __clz(n);
  int low = 0, high = n;
  for (int it = 0; it < levels; ++it) {
    int mid = (int)(((unsigned)low + (unsigned)high) >> 1);
    int m = mid < n ? mid : n - 1;
    uint64_t v = a[m];
    bool go_left = left ? (q <= v) : (q < v);
    if (go_left)
      high = mid;
    else
      low = mid;
  }
  return high;
}

__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ unsigned mbcnt64(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Histogram update aggregated over the wave: lanes holding the same 8-bit
// digit are found with eight ballots and the lowest of them adds the group
// size, so clustered digits cost one LDS atomic per distinct digit.
// Lanes of `peers` whose BITS-bit digit equals this lane's (BITS ballots).
template <int BITS>
__device__ __forceinline__ uint64_t wave_match_bits(unsigned d, uint64_t peers) {
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    bool bit = (d >> b) & 1u;
    uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  return peers;
}

template <int BITS>
__device__ __forceinline__ void wave_hist_add_bits(uint32_t* h, unsigned d, bool valid) {
  uint64_t peers = wave_match_bits<BITS>(d, __ballot(valid));
  if (valid && mbcnt64(peers) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
}

// Histogram add for the digit of one radix pass. Pass 0 adds with one LDS
// atomic per lane; later passes aggregate repeated digits across the wave
// with ballots first, over only the `used` low bits the key range can set
// (the top pass of a 63-bit Morton code has 7). Used for the top passes of
// Morton codes, whose digits repeat within a cluster.
template <int BITS>
__device__ __forceinline__ void pass_hist_add(uint32_t* h, unsigned d, bool valid, int pass,
                                              int used) {
  if (pass == 0) {
    if (valid) atomicAdd(&h[d], 1u);
    return;
  }
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    if (b >= used) break;
    bool bit = (d >> b) & 1u;
    uint64_t bal = __ballot(bit);
    peers &= bit ? bal : ~bal;
  }
  if (valid && mbcnt64(peers) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
}

__device__ __forceinline__ void wave_hist_add(uint32_t* h, unsigned d, bool valid) {
  wave_hist_add_bits<8>(h, d, valid);
}

// Wave-wide inclusive prefix sum of a 32-bit value (wave64).
__device__ __forceinline__ unsigned wave_inclusive_sum(unsigned v) {
  const int lane = (int)lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    unsigned o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T>
__device__ __forceinline__ T atomic_load_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void atomic_store_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

inline int grid_for(int64_t n, int threads, int max_blocks = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (int)b;
}

inline size_t align_up(size_t v, size_t a = 256) { return (v + a - 1) / a * a; }

// Bump allocator over a caller-provided workspace.
struct Arena {
  char* base;
  size_t cap;
  size_t off;
  __host__ Arena(void* p, size_t c) : base((char*)p), cap(c), off(0) {}
  template <typename T>
  __host__ T* take(size_t count) {
    size_t bytes = align_up(count * sizeof(T));
    if (base == nullptr) {  // sizing pass
      off += bytes;
      return nullptr;
    }
    if (off + bytes > cap) return nullptr;
    T* r = (T*)(base + off);
    off += bytes;
    return r;
  }
};

// ---- batches of independent trees (segments) --------------------------------
// System s owns points [pt[s], pt[s+1]) and leaves [leaf[s], leaf[s+1]) of the
// concatenated arrays; its L_s - 1 internal nodes start at leaf[s] - s and its
// 2 L_s - 1 nodes at 2 leaf[s] - s. Every index stored in a tree's own arrays
// is local to that tree, as the single-tree kernels produce it.
struct Seg {
  int S;
  int64_t leaf_size;
  const int32_t* pt;    // [S + 1]
  const int32_t* leaf;  // [S + 1]
  __device__ __forceinline__ int64_t pt_off(int s) const { return pt[s]; }
  __device__ __forceinline__ int64_t n(int s) const { return pt[s + 1] - pt[s]; }
  __device__ __forceinline__ int leaf_off(int s) const { return leaf[s]; }
  __device__ __forceinline__ int L(int s) const { return leaf[s + 1] - leaf[s]; }
  __device__ __forceinline__ int ni_off(int s) const { return leaf[s] - s; }
  __device__ __forceinline__ int node_off(int s) const { return 2 * leaf[s] - s; }
  // System of global index g in a layout with per-system offsets off(s)
  // (non-decreasing; empty systems are skipped): largest s with off(s) <= g.
  template <class F>
  __device__ __forceinline__ int find(int64_t g, F off) const {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if ((int64_t)off(mid) <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    return lo;
  }
  __device__ __forceinline__ int of_point(int64_t g) const {
    return find(g, [this](int s) { return pt_off(s); });
  }
  __device__ __forceinline__ int of_leaf(int64_t g) const {
    return find(g, [this](int s) { return leaf_off(s); });
  }
  __device__ __forceinline__ int of_internal(int64_t g) const {
    return find(g, [this](int s) { return ni_off(s); });
  }
  __device__ __forceinline__ int of_node(int64_t g) const {
    return find(g, [this](int s) { return node_off(s); });
  }
};

// Block-local copy of one per-system offset table (entries 0..S-1) for
// binary searches in LDS; kernels stage it when S <= kSegLdsCap and search
// the global table otherwise. Same result as Seg::find.
constexpr int kSegLdsCap = 4096;

template <class F>
__device__ __forceinline__ bool seg_stage(int32_t* s_tab, int S, F off) {
  if (S > kSegLdsCap) return false;
  for (int i = threadIdx.x; i < S; i += blockDim.x) s_tab[i] = off(i);
  return true;
}

__device__ __forceinline__ int seg_find_lds(const int32_t* s_tab, int S, int64_t g) {
  int lo = 0, hi = S - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int64_t)s_tab[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---- primitives implemented in ygg_sort.hip --------------------------------
// Stable LSD radix sort of (key, u32 value) pairs on key bits [begin, end)
// with digit_bits-wide digits (8..11; 0 = default_digit_bits).
// vals_in == nullptr means values are the input positions 0..n-1.
// A precomputed per-pass digit histogram (passes x 2^digit_bits, u32) may be
// passed through `hist`.
constexpr int kMaxDigitBits = 11;
int default_digit_bits(int key_bytes, int range_bits);
size_t sort_workspace_bytes(int64_t n, int key_bytes, int range_bits, int digit_bits = 0);
int sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                   uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                   size_t ws_bytes, hipStream_t st, const uint32_t* hist = nullptr,
                   int digit_bits = 0);
int sort_pairs_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                   uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                   size_t ws_bytes, hipStream_t st, int digit_bits = 0);

size_t scan_workspace_bytes(int64_t n);
// Row-key sorts whose last pass writes the far sources / near neighbours
// and adds the per-row entry counts into row_counts[rows] (zero before the
// call); key_scratch[n] is scratch for the passes before the last. n_dev
// (optional, device) shortens them to the first *n_dev entries.
size_t sort_rows_fused_workspace_bytes(int64_t n, int end_bit, int digit_bits);
int sort_far_rows_fused(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* key_scratch,
                        int64_t n, const int64_t* n_dev, int end_bit, int digit_bits, int out_i64,
                        void* sources, uint32_t* row_counts, int64_t rows, void* ws,
                        size_t ws_bytes, hipStream_t st);
int sort_near_rows_fused(const uint32_t* keys_in, const uint32_t* vals_in,
                         uint32_t* key_scratch, int64_t n, const int64_t* n_dev, int end_bit,
                         int digit_bits, int out_i64, void* neighbors, uint32_t* row_counts,
                         int64_t rows, void* ws, size_t ws_bytes, hipStream_t st);
// Row lists by bucketing (ygg_rowlists.hip): entries grouped by row key
// with no stable sort (order inside a row unspecified), for row keys of at
// most 22 bits (row_bucket_split false beyond). row_bucket_count adds each
// row's entries into row_counts[rows] (zero before the call) and fills
// `state`; after the caller's scan of the counts into row_offsets,
// row_bucket_scatter writes every value at its row's place. Keys >= rows
// are dropped. vals_scratch[entries] holds the split's values. With
// `records` the input is n pair records (a, b), each standing for the
// entries (a, b) and (b, a); the workspace is then sized for 2n entries.
struct RowBucketState {
  int64_t n;  // entries (capacity)
  const uint32_t* keys;
  const uint16_t* lo16;
  const uint32_t* vals;
  bool packed;  // vals hold value << lo_bits | low digit
  uint32_t rows;
  int lo_bits;
  int buckets;
  unsigned grid;
  const uint32_t* bstart;
  const uint32_t* tstart;
  const uint32_t* tile_bases;
  const int64_t* row_sub;  // taken off each row's values by the scatter (may be null)
};
bool row_bucket_split(int end_bit, int* hi_bits, int* lo_bits);
size_t row_bucket_workspace_bytes(int64_t n, int end_bit);
int row_bucket_count(const uint32_t* keys, const uint32_t* vals, bool records, int64_t n,
                     const int64_t* n_dev, int end_bit, int64_t rows, int64_t list_rows,
                     int64_t val_bound, uint32_t* vals_scratch, uint32_t* row_counts, void* ws,
                     size_t ws_bytes, hipStream_t st, RowBucketState* state);
int row_bucket_scatter(const RowBucketState& s, const int64_t* row_offsets,
                       const int64_t* row_sub, int out_i64, void* out, hipStream_t st);
// Exclusive prefix sum of int32 counts into int64 offsets (n+1 entries:
// out[n] = total). `gather`, if non-null, reads counts[gather[i]].
int exclusive_scan_i32_to_i64(const int32_t* counts, const int32_t* gather, int64_t* out,
                              int64_t n, void* ws, size_t ws_bytes, hipStream_t st);

int fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st);
// Device-to-device copy of n 32-bit words by a kernel (graph-capturable).
int copy_u32(uint32_t* dst, const uint32_t* src, int64_t n, hipStream_t st);

}  // namespace ygg
