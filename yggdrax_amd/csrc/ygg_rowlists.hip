// Row lists by bucketing: groups the walk's (row key, value) entries by row
// without a stable sort.
//
// The far and near lists only need every row's entries contiguous, in row
// order; the order inside a row is the walk's emission order, which is not
// specified (it follows atomic scheduling). So no pass has to keep equal
// digits in input order: each key is ranked inside its tile with one LDS
// atomic instead of the stable sort's per-bit ballot match, and tiles
// reserve their output ranges with atomics instead of an ordered look-back.
// A stable LSD sort needs that order for its second pass; this splits the
// work MSD first:
//
//   1. k_row_hist    - counts of the high digit (row >> lo_bits) of every
//                      valid key (key < rows);
//   2. k_row_tiles   - bucket starts (their exclusive scan) and the tiles
//                      each bucket is cut into (16384 entries, none spanning
//                      two buckets);
//   3. k_row_split   - one pass over the entries: each tile ranks its keys
//                      by high digit with LDS atomics, reserves its range in
//                      every bucket with one atomic and writes its entries
//                      grouped by bucket: value << lo_bits | low digit in one
//                      u32 when the values fit, else the low digit (u16)
//                      beside the value (u32); keys at or above `rows` are
//                      dropped here;
//   4. k_row_reserve - per bucket tile, the low-digit counts in LDS and one
//                      global atomic per (tile, row) present: the old value
//                      is the tile's first slot inside the row, the final
//                      values are the row counts;
//   5. the row offsets (exclusive scan of the counts, by the caller);
//   6. k_row_scatter - per bucket tile, ranks again by low digit, stages the
//                      values in row order in LDS and writes each row's run
//                      at row offset + tile slot.
//
// Row keys of at most 11 bits skip 1 and 3: the input itself is one bucket.
//
// Pair records: the input may hold n records (a, b), each standing for the
// two entries (a, b) and (b, a) (the walk writes one record per accepted or
// near pair: the two node ids). The histogram counts both keys and the
// split writes both entries, so every later step sees 2n entries; the
// split runs (one bucket) also for keys of at most 11 bits.
// Traffic per entry with packed values: 4 B (histogram) + 8 B read and 4 B
// written (split) + 4 B (reserve) + 4 B read and 4 or 8 B written
// (scatter); 32 B with int64 lists, against 36 B for two stable passes
// whose last writes the list. Pair records read 4 B per entry in the
// split (one 8-byte record per two entries): 28 B.
#include "ygg_common.h"

namespace ygg {
namespace {

// Workgroup shapes: 1024 threads of 16 entries for split and bucket tiles.
// The tile bounds the runs each digit leaves for the scattered writes (about
// 32 entries for 512 digits), and those writes cost more than anything else
// here: on gfx950 the near lists of the 2^24 clustered pipeline (3.3e8
// entries, 18-bit rows) took 3.2 ms with 16384-entry tiles against 3.9 ms
// with 8192 (512 threads, two workgroups per CU) and 6.2 ms with 4096.
// Splitting 18-bit rows 9 + 9 beat 8 + 10 and 7 + 11. Digit counters are
// sized for kRowMaxBits-bit digits.
constexpr int kSplitThreads = 1024;
constexpr int kBucketThreads = 1024;
constexpr int kRowItems = 16;
constexpr int kSplitTile = kSplitThreads * kRowItems;
constexpr int kBucketTile = kBucketThreads * kRowItems;
constexpr int kRowMaxBits = 11;
// entry counts fit 30 bits (the stable sort's limit too)
constexpr int64_t kRowMaxEntries = (1 << 30) - 1;

__device__ __forceinline__ int64_t effective_count(int64_t n, const int64_t* n_dev) {
  if (n_dev && *n_dev < n) n = *n_dev > 0 ? *n_dev : 0;
  return n;
}

// Exclusive prefix of v over the workgroup's THREADS threads; *total gets
// the sum over all threads. s_wsum (THREADS / 64 words) must not be in use
// by another scan until a barrier follows this one.
template <int THREADS>
__device__ __forceinline__ unsigned block_exclusive(unsigned v, unsigned* s_wsum,
                                                    unsigned* total) {
  constexpr int kW = THREADS / YGG_WAVE;
  const unsigned inc = wave_inclusive_sum(v);
  const int w = threadIdx.x / YGG_WAVE;
  if (lane_id() == YGG_WAVE - 1) s_wsum[w] = inc;
  __syncthreads();
  unsigned run = 0, all = 0;
#pragma unroll
  for (int q = 0; q < kW; ++q) {
    const unsigned s = s_wsum[q];
    run += q < w ? s : 0u;
    all += s;
  }
  *total = all;
  return run + inc - v;
}

// Bucket of tile t: the b with tstart[b] <= t < tstart[b + 1] (empty
// buckets own no tiles).
__device__ __forceinline__ int tile_bucket(const uint32_t* tstart, int nb, uint32_t t) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tstart[mid] <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// keys2: the second keys of pair records (counted too), or null.
__global__ __launch_bounds__(256) void k_row_hist(const uint32_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ keys2, int64_t n,
                                                  const int64_t* __restrict__ n_dev,
                                                  uint32_t rows, int lo_bits, int nb,
                                                  uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[1 << kRowMaxBits];
  for (int i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
  __syncthreads();
  n = effective_count(n, n_dev);
  // four keys per thread per step, all loaded before any is counted
  constexpr int kUnroll = 4;
  const int64_t step = (int64_t)gridDim.x * blockDim.x * kUnroll;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x * kUnroll; b0 < n; b0 += step) {
    uint32_t k[kUnroll], k2[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      const int64_t i = b0 + (int64_t)j * blockDim.x + threadIdx.x;
      k[j] = i < n ? keys[i] : rows;
      k2[j] = keys2 && i < n ? keys2[i] : rows;
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      if (k[j] < rows) atomicAdd(&h[k[j] >> lo_bits], 1u);
      if (k2[j] < rows) atomicAdd(&h[k2[j] >> lo_bits], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// One workgroup of 1024 threads: bucket starts bstart[nb + 1] and tile
// starts tstart[nb + 1] (tiles of kBucketTile entries) from the bucket
// counts (hist, or, with hist == nullptr, one bucket holding every entry).
// The counts are zeroed: the split reserves its ranges in them.
__global__ __launch_bounds__(1024) void k_row_tiles(uint32_t* __restrict__ hist, int nb,
                                                    int64_t n, const int64_t* __restrict__ n_dev,
                                                    uint32_t* __restrict__ bstart,
                                                    uint32_t* __restrict__ tstart) {
  constexpr int kPer = (1 << kRowMaxBits) / 1024;
  __shared__ unsigned s_wc[1024 / YGG_WAVE], s_wt[1024 / YGG_WAVE];
  const int t = threadIdx.x;
  unsigned c[kPer], tl[kPer], cs = 0, ts = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int b = t * kPer + j;
    c[j] = b < nb ? (hist ? hist[b] : (unsigned)effective_count(n, n_dev)) : 0u;
    if (hist && b < nb) hist[b] = 0u;  // the split's bucket cursors from here on
    tl[j] = (c[j] + kBucketTile - 1) / kBucketTile;
    cs += c[j];
    ts += tl[j];
  }
  unsigned ctot, ttot;
  unsigned ce = block_exclusive<1024>(cs, s_wc, &ctot);
  unsigned te = block_exclusive<1024>(ts, s_wt, &ttot);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int b = t * kPer + j;
    if (b < nb) {
      bstart[b] = ce;
      tstart[b] = te;
    }
    ce += c[j];
    te += tl[j];
  }
  if (t == 0) {
    bstart[nb] = ctot;
    tstart[nb] = ttot;
  }
}

// Layout of the entries the bucket kernels read: the input itself (one
// bucket), the split's low digits (u16) beside its values (u32), or one u32
// per entry packing value << lo_bits | low digit (values below
// 2^(32 - lo_bits)).
enum RowMode { kRowDirect = 0, kRowLo16 = 1, kRowPacked = 2 };

// Split by high digit (see the file comment, step 3) for at most BINS
// digits. Each tile reserves its range inside every bucket it adds to with
// one atomic on that bucket's cursor (zero before the launch), after it
// has staged its entries in LDS, so the atomics' round trip overlaps that
// work. Ranges are disjoint and fill each bucket; their order is the order
// the tiles reach the cursor, which the lists do not depend on.
// kPairs: the input is n pair records, half a tile of records per
// workgroup.
template <int THREADS, int BINS, bool kPacked, bool kPairs>
__global__ __launch_bounds__(THREADS) void k_row_split(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n,
    const int64_t* __restrict__ n_dev, uint32_t rows, int lo_bits, int nb,
    const uint32_t* __restrict__ bstart, uint32_t* __restrict__ cursor,
    uint16_t* __restrict__ lo_out, uint32_t* __restrict__ vals_out) {
  constexpr int kTile = THREADS * kRowItems;
  constexpr int kPer = BINS / THREADS;  // digits owned per thread
  // digit counts, then (owners only) each digit's output index minus its
  // first tile slot
  __shared__ uint32_t s_cnt[BINS];
  __shared__ uint32_t s_pos[BINS];  // each digit's first slot in the tile
  __shared__ uint32_t s_wsum[THREADS / YGG_WAVE];
  __shared__ uint32_t s_word[kTile];  // value, or packed value and low digit
  __shared__ uint16_t s_lo[kPacked ? 1 : kTile];
  __shared__ uint16_t s_dig[kTile];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * (kPairs ? kTile / 2 : kTile);
  n = effective_count(n, n_dev);
  if (base >= n) return;  // the same for every thread
  for (int i = tid; i < nb; i += THREADS) s_cnt[i] = 0;
  __syncthreads();
  const uint32_t lmask = (1u << lo_bits) - 1u;
  uint32_t key[kRowItems], val[kRowItems], rank[kRowItems];
  if (kPairs) {
#pragma unroll
    for (int k = 0; k < kRowItems / 2; ++k) {
      const int64_t idx = base + (int64_t)k * THREADS + tid;
      const bool ok = idx < n;
      key[2 * k] = ok ? keys[idx] : rows;
      key[2 * k + 1] = ok ? vals[idx] : rows;
    }
#pragma unroll
    for (int k = 0; k < kRowItems / 2; ++k) {
      val[2 * k] = key[2 * k + 1];
      val[2 * k + 1] = key[2 * k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < kRowItems; ++k) {
      const int64_t idx = base + (int64_t)k * THREADS + tid;
      const bool ok = idx < n;
      key[k] = ok ? keys[idx] : rows;
      val[k] = ok ? vals[idx] : 0u;
    }
  }
#pragma unroll
  for (int k = 0; k < kRowItems; ++k)
    rank[k] = key[k] < rows ? atomicAdd(&s_cnt[key[k] >> lo_bits], 1u) : 0u;
  __syncthreads();
  // thread tid owns digits [tid * kPer, tid * kPer + kPer)
  unsigned tot[kPer], mine = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    tot[j] = d < nb ? s_cnt[d] : 0u;
    mine += tot[j];
  }
  unsigned valid;
  const unsigned first = block_exclusive<THREADS>(mine, s_wsum, &valid);
  unsigned run = first;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    if (d < nb) s_pos[d] = run;
    run += tot[j];
  }
  __syncthreads();
  unsigned got[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j)
    got[j] = tot[j] ? atomicAdd(&cursor[tid * kPer + j], tot[j]) : 0u;
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    if (key[k] >= rows) continue;
    const unsigned d = key[k] >> lo_bits;
    const unsigned slot = s_pos[d] + rank[k];
    if (kPacked) {
      s_word[slot] = val[k] << lo_bits | (key[k] & lmask);
    } else {
      s_word[slot] = val[k];
      s_lo[slot] = (uint16_t)(key[k] & lmask);
    }
    s_dig[slot] = (uint16_t)d;
  }
  // The digit's output index is at least its first tile slot (the bucket
  // start counts every valid key of lower digits, this tile's included), so
  // the difference kept in s_cnt does not wrap.
  run = first;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    if (tot[j]) s_cnt[d] = bstart[d] + got[j] - run;
    run += tot[j];
  }
  __syncthreads();
  for (int p = tid; p < (int)valid; p += THREADS) {
    const int64_t g = (int64_t)s_cnt[s_dig[p]] + p;
    vals_out[g] = s_word[p];
    if (!kPacked) lo_out[g] = s_lo[p];
  }
}

// Tile t's bucket and entry range, found by thread 0 and shared.
struct RowTile {
  int bucket;
  int64_t begin, end;
};

__device__ __forceinline__ bool load_row_tile(const uint32_t* bstart, const uint32_t* tstart,
                                              int nb, RowTile* s_rt) {
  const uint32_t t = blockIdx.x;
  if (t >= tstart[nb]) return false;  // the same for every thread
  if (threadIdx.x == 0) {
    const int b = tile_bucket(tstart, nb, t);
    const int64_t begin = (int64_t)bstart[b] + (int64_t)(t - tstart[b]) * kBucketTile;
    const int64_t lim = bstart[b + 1];
    s_rt->bucket = b;
    s_rt->begin = begin;
    s_rt->end = begin + kBucketTile < lim ? begin + kBucketTile : lim;
  }
  return true;
}

// Raw words of entry idx of a bucket tile: (key, value) of the input,
// (low digit, value) of the split, or the packed word. Every load of a tile
// is issued before any word is decoded (a decode right after its load
// would make each thread wait for its loads one by one).
template <int kMode, bool kWithValue>
__device__ __forceinline__ void row_load(const uint32_t* keys, const uint16_t* lo_in,
                                         const uint32_t* vals, int64_t idx, bool ok,
                                         uint32_t& a, uint32_t& b) {
  if (kMode == kRowDirect) {
    a = ok ? keys[idx] : 0u;
    b = ok && kWithValue ? vals[idx] : 0u;
  } else if (kMode == kRowLo16) {
    a = ok ? (uint32_t)lo_in[idx] : 0u;
    b = ok && kWithValue ? vals[idx] : 0u;
  } else {
    a = ok ? vals[idx] : 0u;
    b = 0u;
  }
}

// Low digit of a loaded entry, -1 for none (past the tile, or a key at or
// above rows in the input), and its value.
template <int kMode>
__device__ __forceinline__ int row_digit(bool ok, uint32_t a, uint32_t rows, int lo_bits) {
  if (!ok) return -1;
  if (kMode == kRowDirect) return a < rows ? (int)a : -1;
  if (kMode == kRowLo16) return (int)a;
  return (int)(a & ((1u << lo_bits) - 1u));
}
template <int kMode>
__device__ __forceinline__ uint32_t row_value(uint32_t a, uint32_t b, int lo_bits) {
  return kMode == kRowPacked ? a >> lo_bits : b;
}

template <int BINS, int kMode>
__global__ __launch_bounds__(kBucketThreads) void k_row_reserve(
    const uint32_t* __restrict__ keys, const uint16_t* __restrict__ lo_in,
    const uint32_t* __restrict__ vals, uint32_t rows, int lo_bits, int nb,
    const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ tstart,
    uint32_t* __restrict__ row_counts, uint32_t* __restrict__ tile_bases) {
  __shared__ uint32_t s_cnt[BINS];
  __shared__ RowTile s_rt;
  const int nl = 1 << lo_bits;
  const int tid = threadIdx.x;
  if (!load_row_tile(bstart, tstart, nb, &s_rt)) return;
  for (int i = tid; i < nl; i += kBucketThreads) s_cnt[i] = 0;
  __syncthreads();
  const RowTile rt = s_rt;
  uint32_t ra[kRowItems], rb[kRowItems];
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int64_t idx = rt.begin + (int64_t)k * kBucketThreads + tid;
    row_load<kMode, false>(keys, lo_in, vals, idx, idx < rt.end, ra[k], rb[k]);
  }
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int64_t idx = rt.begin + (int64_t)k * kBucketThreads + tid;
    const int d = row_digit<kMode>(idx < rt.end, ra[k], rows, lo_bits);
    if (d >= 0) atomicAdd(&s_cnt[d], 1u);
  }
  __syncthreads();
  const int64_t tb = (int64_t)blockIdx.x * nl;
  for (int d = tid; d < nl; d += kBucketThreads) {
    const uint32_t c = s_cnt[d];
    if (c) tile_bases[tb + d] = atomicAdd(&row_counts[((uint32_t)rt.bucket << lo_bits) | d], c);
  }
}

template <int BINS, int kMode, typename OutT>
__global__ __launch_bounds__(kBucketThreads) void k_row_scatter(
    const uint32_t* __restrict__ keys, const uint16_t* __restrict__ lo_in,
    const uint32_t* __restrict__ vals, uint32_t rows, int lo_bits, int nb,
    const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ tstart,
    const uint32_t* __restrict__ tile_bases, const int64_t* __restrict__ row_offsets,
    const int64_t* __restrict__ row_sub, OutT* __restrict__ out) {
  constexpr int kPer = BINS / kBucketThreads;
  __shared__ uint32_t s_cnt[BINS];  // counts, then each digit's first slot
  __shared__ int64_t s_g[BINS];     // output index of a digit's first slot, minus it
  __shared__ int64_t s_sub[BINS];   // what row_sub takes off the row's values
  __shared__ uint32_t s_wsum[kBucketThreads / YGG_WAVE];
  __shared__ uint32_t s_val[kBucketTile];
  __shared__ uint16_t s_dig[kBucketTile];
  __shared__ RowTile s_rt;
  const int nl = 1 << lo_bits;
  const int tid = threadIdx.x;
  if (!load_row_tile(bstart, tstart, nb, &s_rt)) return;
  for (int i = tid; i < nl; i += kBucketThreads) s_cnt[i] = 0;
  __syncthreads();
  const RowTile rt = s_rt;
  int dig[kRowItems];
  uint32_t val[kRowItems], rank[kRowItems];
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int64_t idx = rt.begin + (int64_t)k * kBucketThreads + tid;
    uint32_t a, b;
    row_load<kMode, true>(keys, lo_in, vals, idx, idx < rt.end, a, b);
    dig[k] = (int)a;  // raw words until every load is issued
    val[k] = b;
  }
  // The owned digits' row offsets and this tile's first slots, loaded with
  // the entries (a digit absent from the tile leaves its slot unwritten and
  // unused; rows past the last read the final offset).
  const int64_t tb = (int64_t)blockIdx.x * nl;
  int64_t roff[kPer], rsub[kPer];
  uint32_t tbase[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    roff[j] = 0;
    rsub[j] = 0;
    tbase[j] = 0;
    if (d < nl) {
      const uint32_t row = ((uint32_t)rt.bucket << lo_bits) | (uint32_t)d;
      roff[j] = row_offsets[row < rows ? row : rows];
      if (row_sub) rsub[j] = row_sub[row < rows ? row : rows - 1];
      tbase[j] = tile_bases[tb + d];
    }
  }
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    const int64_t idx = rt.begin + (int64_t)k * kBucketThreads + tid;
    const uint32_t a = (uint32_t)dig[k];
    dig[k] = row_digit<kMode>(idx < rt.end, a, rows, lo_bits);
    val[k] = row_value<kMode>(a, val[k], lo_bits);
    rank[k] = dig[k] >= 0 ? atomicAdd(&s_cnt[dig[k]], 1u) : 0u;
  }
  __syncthreads();
  unsigned c[kPer], mine = 0;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    c[j] = d < nl ? s_cnt[d] : 0u;
    mine += c[j];
  }
  unsigned valid;
  unsigned run = block_exclusive<kBucketThreads>(mine, s_wsum, &valid);
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int d = tid * kPer + j;
    if (d < nl) {
      if (c[j]) {
        s_g[d] = roff[j] + (int64_t)tbase[j] - (int64_t)run;
        s_sub[d] = rsub[j];
      }
      s_cnt[d] = run;
    }
    run += c[j];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kRowItems; ++k) {
    if (dig[k] < 0) continue;
    const unsigned slot = s_cnt[dig[k]] + rank[k];
    s_val[slot] = val[k];
    s_dig[slot] = (uint16_t)dig[k];
  }
  __syncthreads();
  if (row_sub) {
    for (int p = tid; p < (int)valid; p += kBucketThreads) {
      const int d = s_dig[p];
      out[s_g[d] + p] = (OutT)((int64_t)s_val[p] - s_sub[d]);
    }
  } else {
    for (int p = tid; p < (int)valid; p += kBucketThreads)
      out[s_g[s_dig[p]] + p] = (OutT)s_val[p];
  }
}

template <int BINS, int kMode>
void launch_row_bucket(const RowBucketState& s, uint32_t* row_counts, const int64_t* row_offsets,
                       int out_i64, void* out, hipStream_t st) {
  dim3 g(s.grid), b(kBucketThreads);
  if (row_counts)
    hipLaunchKernelGGL((k_row_reserve<BINS, kMode>), g, b, 0, st, s.keys, s.lo16, s.vals, s.rows,
                       s.lo_bits, s.buckets, s.bstart, s.tstart, row_counts,
                       (uint32_t*)s.tile_bases);
  else if (out_i64)
    hipLaunchKernelGGL((k_row_scatter<BINS, kMode, int64_t>), g, b, 0, st, s.keys, s.lo16,
                       s.vals, s.rows, s.lo_bits, s.buckets, s.bstart, s.tstart, s.tile_bases,
                       row_offsets, s.row_sub, (int64_t*)out);
  else
    hipLaunchKernelGGL((k_row_scatter<BINS, kMode, int32_t>), g, b, 0, st, s.keys, s.lo16,
                       s.vals, s.rows, s.lo_bits, s.buckets, s.bstart, s.tstart, s.tile_bases,
                       row_offsets, s.row_sub, (int32_t*)out);
}

// Reserve (row_counts given) or scatter, for the state's layout.
void launch_row_bucket_any(const RowBucketState& s, uint32_t* row_counts,
                           const int64_t* row_offsets, int out_i64, void* out, hipStream_t st) {
  const int mode = s.lo16 ? kRowLo16 : (s.packed ? kRowPacked : kRowDirect);
  if (s.lo_bits <= 10) {
    if (mode == kRowDirect)
      launch_row_bucket<1024, kRowDirect>(s, row_counts, row_offsets, out_i64, out, st);
    else if (mode == kRowLo16)
      launch_row_bucket<1024, kRowLo16>(s, row_counts, row_offsets, out_i64, out, st);
    else
      launch_row_bucket<1024, kRowPacked>(s, row_counts, row_offsets, out_i64, out, st);
  } else {
    if (mode == kRowDirect)
      launch_row_bucket<2048, kRowDirect>(s, row_counts, row_offsets, out_i64, out, st);
    else if (mode == kRowLo16)
      launch_row_bucket<2048, kRowLo16>(s, row_counts, row_offsets, out_i64, out, st);
    else
      launch_row_bucket<2048, kRowPacked>(s, row_counts, row_offsets, out_i64, out, st);
  }
}

template <bool kPairs>
void launch_row_split(int nb, bool packed, unsigned tiles, const uint32_t* keys,
                      const uint32_t* vals, int64_t n, const int64_t* n_dev, uint32_t rows,
                      int lo, const uint32_t* bstart, uint32_t* cursor, uint16_t* lo16,
                      uint32_t* vals_out, hipStream_t st) {
  dim3 sg(tiles), sb(kSplitThreads);
  if (nb <= 1024 && packed)
    hipLaunchKernelGGL((k_row_split<kSplitThreads, 1024, true, kPairs>), sg, sb, 0, st, keys,
                       vals, n, n_dev, rows, lo, nb, bstart, cursor, lo16, vals_out);
  else if (nb <= 1024)
    hipLaunchKernelGGL((k_row_split<kSplitThreads, 1024, false, kPairs>), sg, sb, 0, st, keys,
                       vals, n, n_dev, rows, lo, nb, bstart, cursor, lo16, vals_out);
  else if (packed)
    hipLaunchKernelGGL((k_row_split<kSplitThreads, 2048, true, kPairs>), sg, sb, 0, st, keys,
                       vals, n, n_dev, rows, lo, nb, bstart, cursor, lo16, vals_out);
  else
    hipLaunchKernelGGL((k_row_split<kSplitThreads, 2048, false, kPairs>), sg, sb, 0, st, keys,
                       vals, n, n_dev, rows, lo, nb, bstart, cursor, lo16, vals_out);
}

}  // namespace

// (high, low) digit widths of a row-key range of end_bit bits; false when
// the range is too wide for one split and one bucket pass.
bool row_bucket_split(int end_bit, int* hi_bits, int* lo_bits) {
  if (end_bit < 1 || end_bit > 2 * kRowMaxBits) return false;
  if (end_bit <= kRowMaxBits) {
    *hi_bits = 0;
    *lo_bits = end_bit;
    return true;
  }
  *lo_bits = (end_bit + 1) / 2;
  *hi_bits = end_bit - *lo_bits;
  return true;
}

// Split widths when only list_rows of the 2^end_bit keys are rows (pair
// records are keyed by node id, and the near lists' rows, the leaves, are
// about half of the ids): the low digit spans the rows' count, so a bucket
// holds about as many rows as with dense keys, and the high digit the rest
// (at most kRowMaxBits; the low digit never wider than row_bucket_split's).
// On the 2^24 clustered near lists (19-bit node keys, 2^18 leaves) 10 + 9
// took the scatter from 2.01 to 1.67 ms against 9 + 10, which leaves half
// the buckets empty and doubles the rows of the others (the split over
// 1024 digits instead of 512 costs 0.13 ms of it back).
static void row_split_for_rows(int end_bit, int64_t list_rows, int* hi, int* lo) {
  if (*hi == 0) return;
  int rbits = 1;
  while (rbits < 62 && ((int64_t)1 << rbits) < list_rows) ++rbits;
  int l = (rbits + 1) / 2;
  if (l > *lo) l = *lo;
  if (end_bit - l > kRowMaxBits) l = end_bit - kRowMaxBits;
  *lo = l;
  *hi = end_bit - l;
}

size_t row_bucket_workspace_bytes(int64_t n, int end_bit) {
  int hi = 0, lo = 0;
  if (!row_bucket_split(end_bit, &hi, &lo)) return 0;
  // bucket tables for any split; tile slots for the widest low digit
  // (tiles << lo grows with lo, 2^hi << lo is 2^end_bit for every split)
  const int64_t nb = (int64_t)1 << kRowMaxBits;
  const int64_t tiles2 = (n + kBucketTile - 1) / kBucketTile + ((int64_t)1 << hi);
  Arena a(nullptr, 0);
  // the split's scratch, also for keys of at most 11 bits (pair records
  // are split into one bucket)
  a.take<uint16_t>((size_t)n);                          // low digits after the split
  a.take<uint32_t>((size_t)nb);                         // bucket counts, then cursors
  a.take<uint32_t>((size_t)nb + 1);                     // bucket starts
  a.take<uint32_t>((size_t)nb + 1);                     // tile starts
  a.take<uint32_t>((size_t)(tiles2 << lo));             // tiles' first slots per row
  return a.off;
}

// Steps 1-4 (the row counts, added into row_counts[rows], zero before the
// call). The split's values go to vals_scratch[entries], packed with the
// low digits when every value is below val_bound <= 2^(32 - lo_bits)
// (val_bound 0: no bound known). With `records` the input is n pair
// records (2n entries; n_dev counts records), else n entries. list_rows
// (<= rows) is how many keys are rows of the lists (it sets the split).
// The workspace is row_bucket_workspace_bytes of the entry count. `state`
// carries what the scatter needs.
int row_bucket_count(const uint32_t* keys, const uint32_t* vals, bool records, int64_t n,
                     const int64_t* n_dev, int end_bit, int64_t rows, int64_t list_rows,
                     int64_t val_bound, uint32_t* vals_scratch, uint32_t* row_counts, void* ws,
                     size_t ws_bytes, hipStream_t st, RowBucketState* state) {
  int hi = 0, lo = 0;
  const int64_t ne = records ? 2 * n : n;  // entries
  if (!row_bucket_split(end_bit, &hi, &lo) || rows < 1 || rows > ((int64_t)1 << end_bit) ||
      list_rows < 1 || list_rows > rows || n < 0 || ne > kRowMaxEntries || val_bound < 0 ||
      !state)
    return YGG_ERR_ARG;
  row_split_for_rows(end_bit, list_rows, &hi, &lo);
  const bool split = hi > 0 || records;
  const int64_t nbmax = (int64_t)1 << hi;
  const int64_t tiles1 = (ne + kSplitTile - 1) / kSplitTile;
  const int64_t tiles2 = (ne + kBucketTile - 1) / kBucketTile + nbmax;
  Arena a(ws, ws_bytes);
  uint16_t* lo16 = a.take<uint16_t>((size_t)ne);
  uint32_t* hist = a.take<uint32_t>((size_t)nbmax);
  if (!lo16 || !hist || (split && !vals_scratch)) return YGG_ERR_WORKSPACE;
  uint32_t* bstart = a.take<uint32_t>((size_t)nbmax + 1);
  uint32_t* tstart = a.take<uint32_t>((size_t)nbmax + 1);
  uint32_t* bases = a.take<uint32_t>((size_t)(tiles2 << lo));
  if (!bstart || !tstart || !bases) return YGG_ERR_WORKSPACE;
  const bool packed = split && val_bound > 0 && val_bound <= ((int64_t)1 << (32 - lo));
  // buckets the row range reaches
  const int nb = hi > 0 ? (int)(((rows - 1) >> lo) + 1) : 1;
  const uint32_t r32 = (uint32_t)rows;
  int rc;
  if (split) {
    if ((rc = fill_u32(hist, 0u, nb, st))) return rc;
    hipLaunchKernelGGL(k_row_hist, dim3(grid_for(n > 0 ? n : 1, 1024, 2048)), dim3(256), 0, st,
                       keys, records ? vals : nullptr, n, n_dev, r32, lo, nb, hist);
    YGG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_row_tiles, dim3(1), dim3(1024), 0, st, split ? hist : nullptr, nb, n,
                     n_dev, bstart, tstart);
  YGG_LAUNCH_CHECK();
  if (split && tiles1 > 0) {
    if (records)
      launch_row_split<true>(nb, packed, (unsigned)tiles1, keys, vals, n, n_dev, r32, lo,
                             bstart, hist, lo16, vals_scratch, st);
    else
      launch_row_split<false>(nb, packed, (unsigned)tiles1, keys, vals, n, n_dev, r32, lo,
                              bstart, hist, lo16, vals_scratch, st);
    YGG_LAUNCH_CHECK();
  }
  state->n = ne;
  state->keys = split ? nullptr : keys;
  state->lo16 = split && !packed ? lo16 : nullptr;
  state->vals = split ? vals_scratch : vals;
  state->packed = packed;
  state->rows = r32;
  state->lo_bits = lo;
  state->buckets = nb;
  // at most one partial tile per bucket beyond the full ones
  state->grid = (unsigned)((ne + kBucketTile - 1) / kBucketTile + nb);
  state->bstart = bstart;
  state->tstart = tstart;
  state->tile_bases = bases;
  state->row_sub = nullptr;
  launch_row_bucket_any(*state, row_counts, nullptr, 0, nullptr, st);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

// Step 6: every valid entry's value, less row_sub[row] when row_sub is
// given, as OutT, at its row's offset (row_offsets: exclusive scan of the
// counts) plus its slot in the row. With no entries there is nothing to
// write (and `out` may be null).
int row_bucket_scatter(const RowBucketState& s, const int64_t* row_offsets,
                       const int64_t* row_sub, int out_i64, void* out, hipStream_t st) {
  if (s.n == 0) return YGG_OK;
  if (!row_offsets || !out) return YGG_ERR_ARG;
  RowBucketState sub = s;
  sub.row_sub = row_sub;
  launch_row_bucket_any(sub, nullptr, row_offsets, out_i64, out, st);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // namespace ygg
