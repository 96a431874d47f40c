// Stable LSD radix sort for gfx950 wave64, digit width 8..11 bits, plus scans.
//
// Per digit pass one kernel ("one sweep"): each 256-thread workgroup takes
// the next 4096-key tile from an atomic tile counter, ranks its keys stably
// with wave64 ballot "match" masks and per-wave LDS digit counters, publishes
// its digit counts for decoupled look-back, reorders the tile in LDS by digit
// and writes every digit run out contiguously. A single histogram kernel
// counts all passes' digits up front, so a pass reads and writes each
// (key, value) pair exactly once. Wider digits mean fewer passes over the
// data but more look-back words per tile (2^bits per pass) and more LDS;
// the width is a parameter so callers (and the autotuner) can pick the
// cheapest pass count for a key range.
#include <type_traits>

#include "ygg_common.h"

namespace ygg {

constexpr int kSortThreads = 256;
constexpr int kWaves = kSortThreads / YGG_WAVE;  // 4
// Tile shape of one onesweep pass: threads per workgroup, 16 keys per
// thread. A tile leaves runs of about tile / 2^BITS keys per digit for the
// scatter, and runs shorter than 16 keys cost more than anything else here
// (9-bit digits on 4096-key tiles sort 3.3e8 pairs in 6.8 ms against
// 4.5 ms on 8192-key tiles, gfx950). 8-bit digits and 64-bit keys use
// 256 threads (4096 keys). 32-bit keys with 9- or 10-bit digits use 1024
// threads: 16384-key tiles at 4 waves per SIMD, with 16-bit per-wave
// counters, and for 10-bit digits keys and values staged through one
// buffer, so the workgroup fits 151 / 139 KB of LDS. 3.3e8 pairs with
// 18-bit keys sort in 4.3 ms against 5.37 ms with 256 threads of 32 keys
// (2 waves per SIMD); 1.5e8 pairs with 19-bit keys sort in 2.31 ms with two
// 10-bit passes against 3.23 ms with three 8-bit passes. 11-bit digits
// take 512 threads (8192 keys) to fit LDS.
// 64-bit keys with 9-bit digits (the 63-bit Morton codes: 7 passes) take
// 1024 threads of 12 keys: 12288-key tiles (runs of ~24 keys per digit) in
// 150 KB of LDS; 2^24 clustered codes sort in 1.17 ms against 1.30 ms with
// 8-bit digits on 4096-key tiles.
template <typename K, int BITS>
constexpr int sort_threads() {
  return sizeof(K) == 4 ? (BITS == 9 || BITS == 10 ? 1024 : (BITS >= 11 ? 512 : 256))
                        : (BITS == 9 ? 1024 : 256);
}
template <typename K, int BITS>
constexpr int sort_items() {
  return (sizeof(K) == 8 && BITS == 9) ? 12 : 16;
}
constexpr int kSortItems = 16;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPre = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

constexpr int passes_for(int range_bits, int bits) { return (range_bits + bits - 1) / bits; }

template <typename K, int BITS>
__global__ __launch_bounds__(256) void k_digit_hist(const K* __restrict__ keys, int64_t n,
                                                    int begin_bit, int end_bit, int passes,
                                                    uint32_t* __restrict__ hist,
                                                    const int64_t* __restrict__ n_dev) {
  constexpr int kBins = 1 << BITS;
  if (n_dev && *n_dev < n) n = *n_dev > 0 ? *n_dev : 0;
  constexpr int kMaxPasses = passes_for(8 * (int)sizeof(K), BITS);
  __shared__ uint32_t h[kMaxPasses * kBins];
  for (int i = threadIdx.x; i < passes * kBins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  // Four keys per thread per step, all loaded before any is counted, so
  // each wavefront keeps four coalesced loads in flight. Every lane adds
  // with its own LDS atomic: aggregating repeated digits with ballots first
  // cost more than the conflicts it saves (gfx950: 3.3e8 18-bit keys sort in
  // 5.37 ms against 5.95 ms, and the 2^24 clustered pipeline's row sorts
  // save 0.43 ms).
  constexpr int kUnroll = 4;
  const int64_t step = (int64_t)gridDim.x * blockDim.x * kUnroll;
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x * kUnroll; b0 < n; b0 += step) {
    K k[kUnroll];
    bool valid[kUnroll];
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      const int64_t i = b0 + (int64_t)j * blockDim.x + threadIdx.x;
      valid[j] = i < n;
      k[j] = valid[j] ? keys[i] : (K)0;
    }
#pragma unroll
    for (int j = 0; j < kUnroll; ++j) {
      for (int p = 0; p < passes; ++p) {
        const int lo = begin_bit + BITS * p;
        if (valid[j]) atomicAdd(&h[p * kBins + (unsigned)((k[j] >> lo) & (kBins - 1))], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * kBins; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
}

// In-place exclusive scan of each pass's 2^BITS-bin histogram (one block per
// pass; thread t owns bins [t * D, t * D + D)).
template <int BITS>
__global__ __launch_bounds__(256) void k_hist_scan(uint32_t* hist) {
  constexpr int kBins = 1 << BITS;
  constexpr int D = kBins / kSortThreads;
  __shared__ uint32_t wsum[kWaves];
  uint32_t* h = hist + (int64_t)blockIdx.x * kBins;
  unsigned v[D];
  unsigned t = 0;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    v[j] = h[threadIdx.x * D + j];
    t += v[j];
  }
  unsigned inc = wave_inclusive_sum(t);
  int w = threadIdx.x / YGG_WAVE;
  if (lane_id() == 63) wsum[w] = inc;
  __syncthreads();
  unsigned run = 0;
  for (int k = 0; k < w; ++k) run += wsum[k];
  run += inc - t;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    h[threadIdx.x * D + j] = run;
    run += v[j];
  }
}

// Where a pass writes its sorted pairs: key(g, k) and val(g, v) for the
// pair that lands at position g. Every pass but a fused last one writes the
// keys and values themselves.
template <typename K>
struct EmitKV {
  static constexpr bool kRowCounts = false;
  K* __restrict__ keys_out;
  uint32_t* __restrict__ vals_out;
  uint32_t* counts = nullptr;
  uint32_t rows = 0;
  __device__ __forceinline__ void key(int64_t g, K k) const { keys_out[g] = k; }
  __device__ __forceinline__ void val(int64_t g, uint32_t v) const { vals_out[g] = v; }
};

// Last pass of the far row-key sort fused with the far list's sources
// (_result_to_interactions, _interactions_impl.py:3561-3637): the key is the
// target's level-major rank, the value the source node. Writes the source
// and adds each tile's runs of equal keys into the per-row counts (row
// offsets, targets and levels follow from the counts in kernels with full
// occupancy: dependent gathers here would stall the low-occupancy sort);
// the sorted keys themselves are not written.
template <typename OutT>
struct EmitFarList {
  static constexpr bool kRowCounts = true;
  OutT* __restrict__ sources;
  uint32_t* __restrict__ counts;  // [rows], zero before the pass
  uint32_t rows;
  __device__ __forceinline__ void key(int64_t, uint32_t) const {}
  __device__ __forceinline__ void val(int64_t g, uint32_t v) const { sources[g] = (OutT)v; }
};

// Last pass of the near row-key sort fused with the near list: the key is
// the native leaf row, the value the neighbour leaf node; row counts as for
// the far list.
template <typename OutT>
struct EmitNearList {
  static constexpr bool kRowCounts = true;
  OutT* __restrict__ neighbors;
  uint32_t* __restrict__ counts;
  uint32_t rows;
  __device__ __forceinline__ void key(int64_t, uint32_t) const {}
  __device__ __forceinline__ void val(int64_t g, uint32_t v) const { neighbors[g] = (OutT)v; }
};

// Row counts of a last pass: the tile's keys sit in s_buf fully sorted (the
// pass orders by the top digit, stably, keys already ordered by the lower
// ones), so each run of equal keys [f, l] adds l + 1 - f to its row with
// two atomics: -f from its first element and l + 1 from its last (32-bit
// wrap-around cancels). Keys at or above `rows` (unused capacity) count
// nowhere.
template <typename K, typename Emit>
__device__ __forceinline__ void add_row_runs(const K* s_buf, int64_t valid, const Emit& em,
                                             int tid, int threads) {
  for (int p = tid; p < valid; p += threads) {
    const uint32_t kk = (uint32_t)s_buf[p];
    if (kk >= em.rows) continue;
    if (p == 0 || (uint32_t)s_buf[p - 1] != kk) atomicSub(&em.counts[kk], (uint32_t)p);
    if (p == valid - 1 || (uint32_t)s_buf[p + 1] != kk)
      atomicAdd(&em.counts[kk], (uint32_t)(p + 1));
  }
}

template <typename K, int BITS, int ITEMS, int THREADS, typename Emit>
__global__ __launch_bounds__(THREADS) void k_onesweep(const K* __restrict__ keys_in,
                                                  const uint32_t* __restrict__ vals_in,
                                                  Emit em, int64_t n,
                                                  int shift, const uint32_t* __restrict__ hist_scan,
                                                  uint32_t* __restrict__ lookback,
                                                  uint32_t* __restrict__ tile_counter,
                                                  const int64_t* __restrict__ n_dev) {
  constexpr int kBins = 1 << BITS;
  constexpr unsigned kMask = kBins - 1;
  // thread t < kBins / D owns digits [t * D, t * D + D)
  constexpr int D = kBins >= THREADS ? kBins / THREADS : 1;
  constexpr int kW = THREADS / YGG_WAVE;
  constexpr int kTile = THREADS * ITEMS;
  constexpr int kWaveKeys = kTile / kW;
  __shared__ uint32_t s_tile;
  // per-wave digit counts, then their exclusive prefix over waves; 16 bits
  // hold any count of a tile of up to 65535 keys
  using HistT = typename std::conditional<(THREADS > 256), uint16_t, uint32_t>::type;
  __shared__ HistT s_wave_hist[kW][kBins];
  __shared__ uint32_t s_digit_start[kBins];
  __shared__ uint32_t s_global[kBins];
  __shared__ uint32_t s_wsum[kW];
  // Keys, then values, are staged through one buffer (two scatter phases),
  // with each slot's digit kept beside it: less LDS per tile, more tiles
  // resident per CU.
  // (64-bit keys, and 32-bit keys where two buffers would not fit: for
  // 32-bit keys one buffer each is faster.)
  using DigT = typename std::conditional<(BITS <= 8), uint8_t, uint16_t>::type;
  constexpr bool kTwoPhase = sizeof(K) == 8 || (THREADS == 1024 && BITS >= 10);
  __shared__ K s_buf[kTile];
  __shared__ DigT s_dig[kTwoPhase ? kTile : 1];
  __shared__ uint32_t s_vals[kTwoPhase ? 1 : kTile];

  const int tid = threadIdx.x;
  const int w = tid / YGG_WAVE;
  const unsigned lane = lane_id();
  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < kW * kBins; i += THREADS) (&s_wave_hist[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile = s_tile;
  const int64_t base = tile * kTile;
  // A device-side count (n_dev) shortens the sort to its first *n_dev keys:
  // tiles past it exit at once (tiles are claimed in order, and a tile looks
  // back only at earlier ones).
  if (n_dev && *n_dev < n) n = *n_dev > 0 ? *n_dev : 0;
  if (base >= n) return;
  const int64_t wbase = base + (int64_t)w * kWaveKeys;

  K key[ITEMS];
  uint32_t val[ITEMS];
  uint32_t rank[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    int64_t idx = wbase + k * YGG_WAVE + lane;
    bool ok = idx < n;
    key[k] = ok ? keys_in[idx] : (K)~(K)0;
    val[k] = ok ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    unsigned d = (unsigned)((key[k] >> shift) & kMask);
    uint64_t peers = wave_match_bits<BITS>(d, ~0ull);
    unsigned below = mbcnt64(peers);
    unsigned cnt = (unsigned)__popcll(peers);
    unsigned before = s_wave_hist[w][d];
    rank[k] = before + below;
    if (below == 0) s_wave_hist[w][d] = (HistT)(before + cnt);
  }
  __syncthreads();
  // Thread `tid` owns digits [tid * D, tid * D + D).
  const bool owner = tid * D < kBins;
  unsigned tot[D];
  unsigned tsum = 0;
  uint32_t* lb = lookback + tile * kBins;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    tot[j] = 0;
    if (!owner) continue;
    const unsigned d = (unsigned)tid * D + j;
    unsigned total = 0;
#pragma unroll
    for (int q = 0; q < kW; ++q) {
      unsigned c = s_wave_hist[q][d];
      s_wave_hist[q][d] = (HistT)total;  // exclusive prefix over waves
      total += c;
    }
    tot[j] = total;
    tsum += total;
    atomic_store_agent(&lb[d], (tile == 0 ? kFlagPre : kFlagAgg) | total);
  }
  unsigned inc = wave_inclusive_sum(tsum);
  if (lane == 63) s_wsum[w] = inc;
  unsigned excl_lb[D];
#pragma unroll
  for (int j = 0; j < D; ++j) {
    const unsigned d = (unsigned)tid * D + j;
    unsigned e = 0;
    if (owner && tile > 0) {
      int64_t pt = tile - 1;
      while (true) {
        uint32_t v = atomic_load_agent(&lookback[pt * kBins + d]);
        uint32_t f = v & ~kCountMask;
        if (f == 0) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        e += v & kCountMask;
        if (f == kFlagPre) break;
        --pt;
      }
      atomic_store_agent(&lb[d], kFlagPre | (e + tot[j]));
    }
    excl_lb[j] = e;
  }
  __syncthreads();
  unsigned run = 0;
  for (int q = 0; q < w; ++q) run += s_wsum[q];
  run += inc - tsum;
#pragma unroll
  for (int j = 0; j < D; ++j) {
    if (!owner) break;
    const unsigned d = (unsigned)tid * D + j;
    s_digit_start[d] = run;
    s_global[d] = hist_scan[d] + excl_lb[j] - run;
    run += tot[j];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    unsigned d = (unsigned)((key[k] >> shift) & kMask);
    unsigned pos = s_digit_start[d] + s_wave_hist[w][d] + rank[k];
    rank[k] = pos;  // tile slot, reused for the values
    s_buf[pos] = key[k];
    if (kTwoPhase)
      s_dig[pos] = (DigT)d;
    else
      s_vals[pos] = val[k];
  }
  __syncthreads();
  const int64_t valid = (n - base) < kTile ? (n - base) : kTile;
  if constexpr (Emit::kRowCounts) add_row_runs(s_buf, valid, em, tid, THREADS);
  if (!kTwoPhase) {
    for (int p = tid; p < valid; p += THREADS) {
      const K kk = s_buf[p];
      const int64_t g = (int64_t)s_global[(unsigned)((kk >> shift) & kMask)] + p;
      em.key(g, kk);
      em.val(g, s_vals[p]);
    }
    return;
  }
  for (int p = tid; p < valid; p += THREADS) em.key((int64_t)s_global[s_dig[p]] + p, s_buf[p]);
  __syncthreads();
  uint32_t* s_vb = reinterpret_cast<uint32_t*>(s_buf);
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) s_vb[rank[k]] = val[k];
  __syncthreads();
  for (int p = tid; p < valid; p += THREADS) em.val((int64_t)s_global[s_dig[p]] + p, s_vb[p]);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_copy_u32(uint32_t* __restrict__ d, const uint32_t* __restrict__ s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

int copy_u32(uint32_t* dst, const uint32_t* src, int64_t n, hipStream_t st) {
  if (n <= 0) return YGG_OK;
  hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, dst, src, n);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return YGG_OK;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, st, p, v, n);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

static inline int64_t tile_keys(int key_bytes, int digit_bits) {
  const int threads =
      key_bytes == 4 ? (digit_bits == 9 || digit_bits == 10 ? 1024 : (digit_bits >= 11 ? 512 : 256))
                     : (digit_bits == 9 ? 1024 : 256);
  const int items = (key_bytes == 8 && digit_bits == 9) ? 12 : kSortItems;
  return (int64_t)threads * items;
}
static inline int64_t num_tiles(int64_t n, int key_bytes, int digit_bits) {
  const int64_t t = tile_keys(key_bytes, digit_bits);
  return (n + t - 1) / t;
}

// 8-bit digits: a 4096-key tile then has runs of ~16 keys per digit, so the
// scatter writes stay coalesced. 32-bit keys with 9-bit digits use
// 16384-key tiles (runs of ~32 and ~16) and pay where they save a pass, so
// 32-bit keys take the fewest passes among 8, 9 and 10 bits (the narrowest
// on ties). Measured on gfx950: 3.3e8 (u32, u32) pairs with 18-bit keys sort
// in 4.4 ms with two 9-bit passes (4.9 ms with two 10-bit passes); 1.5e8
// pairs with 19-bit keys in 2.31 ms with two 10-bit passes against 3.0 ms
// (three 9-bit) and 3.23 ms (three 8-bit). 64-bit Morton codes keep 8-bit
// digits (2^24 keys: 1.49 ms, against 1.72 ms with 9-bit and 2.8-4.8 ms
// with 10- or 11-bit digits). Every width remains available to the
// autotuner.
int default_digit_bits(int key_bytes, int range_bits) {
  if (key_bytes != 4) return 8;
  int best = 8;
  for (int b = 9; b <= 10; ++b)
    if (passes_for(range_bits, b) < passes_for(range_bits, best)) best = b;
  return best;
}

size_t sort_workspace_bytes(int64_t n, int key_bytes, int range_bits, int digit_bits) {
  if (digit_bits == 0) digit_bits = default_digit_bits(key_bytes, range_bits);
  const int passes = passes_for(range_bits, digit_bits);
  const size_t bins = (size_t)1 << digit_bits;
  Arena a(nullptr, 0);
  int64_t t = num_tiles(n, key_bytes, digit_bits);
  a.take<char>((size_t)n * key_bytes);                              // alternate keys
  a.take<uint32_t>((size_t)n);                                      // alternate values
  a.take<uint32_t>((size_t)passes * bins);                          // histograms
  a.take<uint32_t>((size_t)passes + (size_t)passes * t * bins);     // counters + look-back
  return a.off;
}

template <typename K, int BITS, typename Last = EmitKV<K>>
static int sort_pairs_bits(const K* keys_in, const uint32_t* vals_in, K* keys_out,
                           uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                           size_t ws_bytes, hipStream_t st, const uint32_t* hist_pre,
                           const Last* last = nullptr, const int64_t* n_dev = nullptr) {
  constexpr int kBins = 1 << BITS;
  constexpr int kThreads = sort_threads<K, BITS>();
  constexpr int kItems = sort_items<K, BITS>();
  const int passes = passes_for(end_bit - begin_bit, BITS);
  const int64_t tiles = num_tiles(n, (int)sizeof(K), BITS);
  if (tile_keys((int)sizeof(K), BITS) != (int64_t)kThreads * kItems) return YGG_ERR_ARG;
  Arena a(ws, ws_bytes);
  K* kalt = (K*)a.take<char>((size_t)n * sizeof(K));
  uint32_t* valt = a.take<uint32_t>((size_t)n);
  uint32_t* hist = a.take<uint32_t>((size_t)passes * kBins);
  uint32_t* ctr = a.take<uint32_t>((size_t)passes + (size_t)passes * tiles * kBins);
  if (!kalt || !valt || !hist || !ctr) return YGG_ERR_WORKSPACE;
  uint32_t* lookback = ctr + passes;
  int rc = fill_u32(ctr, 0u, (int64_t)passes + (int64_t)passes * tiles * kBins, st);
  if (rc) return rc;
  if (hist_pre) {
    rc = copy_u32(hist, hist_pre, (int64_t)passes * kBins, st);
    if (rc) return rc;
  } else {
    rc = fill_u32(hist, 0u, (int64_t)passes * kBins, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_digit_hist<K, BITS>), dim3(grid_for(n, 256, 2048)), dim3(256), 0, st,
                       keys_in, n, begin_bit, end_bit, passes, hist, n_dev);
    YGG_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_hist_scan<BITS>, dim3(passes), dim3(256), 0, st, hist);
  YGG_LAUNCH_CHECK();
  const K* src_k = keys_in;
  const uint32_t* src_v = vals_in;
  for (int p = 0; p < passes; ++p) {
    bool to_out = ((passes - 1 - p) % 2) == 0;
    K* dk = to_out ? keys_out : kalt;
    uint32_t* dv = to_out ? vals_out : valt;
    if (last && p == passes - 1) {
      hipLaunchKernelGGL((k_onesweep<K, BITS, kItems, kThreads, Last>), dim3((unsigned)tiles),
                         dim3(kThreads), 0, st, src_k, src_v, *last, n, begin_bit + BITS * p,
                         hist + p * kBins, lookback + (int64_t)p * tiles * kBins, ctr + p, n_dev);
    } else {
      hipLaunchKernelGGL((k_onesweep<K, BITS, kItems, kThreads, EmitKV<K>>), dim3((unsigned)tiles),
                         dim3(kThreads), 0, st, src_k, src_v, EmitKV<K>{dk, dv}, n,
                         begin_bit + BITS * p, hist + p * kBins,
                         lookback + (int64_t)p * tiles * kBins, ctr + p, n_dev);
    }
    YGG_LAUNCH_CHECK();
    src_k = dk;
    src_v = dv;
  }
  return YGG_OK;
}

// A u32 row-key sort [0, end_bit) whose last pass writes through `last`
// (the row lists) instead of keys and values. The passes before it ping-pong
// between the sort workspace and (keys_out, vals_tmp); keys_out finally
// holds the sorted keys. Workspace: n values, then the sort's own.
size_t sort_rows_fused_workspace_bytes(int64_t n, int end_bit, int digit_bits) {
  return align_up(sizeof(uint32_t) * (size_t)n) + sort_workspace_bytes(n, 4, end_bit, digit_bits);
}

template <typename Last>
static int sort_rows_fused(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                           int64_t n, int end_bit, int digit_bits, void* ws, size_t ws_bytes,
                           hipStream_t st, const Last& last, const int64_t* n_dev) {
  if (n <= 0) return YGG_OK;
  if (n > (int64_t)kCountMask || end_bit < 1 || end_bit > 32 || keys_in == keys_out)
    return YGG_ERR_ARG;
  if (digit_bits == 0) digit_bits = default_digit_bits(4, end_bit);
  Arena a(ws, ws_bytes);
  uint32_t* vals_tmp = a.take<uint32_t>((size_t)n);
  const size_t rest = sort_workspace_bytes(n, 4, end_bit, digit_bits);
  char* sws = a.take<char>(rest);
  if (!vals_tmp || !sws) return YGG_ERR_WORKSPACE;
  switch (digit_bits) {
    case 8:
      return sort_pairs_bits<uint32_t, 8, Last>(keys_in, vals_in, keys_out, vals_tmp, n, 0,
                                                end_bit, sws, rest, st, nullptr, &last, n_dev);
    case 9:
      return sort_pairs_bits<uint32_t, 9, Last>(keys_in, vals_in, keys_out, vals_tmp, n, 0,
                                                end_bit, sws, rest, st, nullptr, &last, n_dev);
    case 10:
      return sort_pairs_bits<uint32_t, 10, Last>(keys_in, vals_in, keys_out, vals_tmp, n, 0,
                                                 end_bit, sws, rest, st, nullptr, &last, n_dev);
    case 11:
      return sort_pairs_bits<uint32_t, 11, Last>(keys_in, vals_in, keys_out, vals_tmp, n, 0,
                                                 end_bit, sws, rest, st, nullptr, &last, n_dev);
    default:
      return YGG_ERR_ARG;
  }
}

int sort_far_rows_fused(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* key_scratch,
                        int64_t n, const int64_t* n_dev, int end_bit, int digit_bits, int out_i64,
                        void* sources, uint32_t* row_counts, int64_t rows, void* ws,
                        size_t ws_bytes, hipStream_t st) {
  if (out_i64) {
    EmitFarList<int64_t> e{(int64_t*)sources, row_counts, (uint32_t)rows};
    return sort_rows_fused(keys_in, vals_in, key_scratch, n, end_bit, digit_bits, ws, ws_bytes,
                           st, e, n_dev);
  }
  EmitFarList<int32_t> e{(int32_t*)sources, row_counts, (uint32_t)rows};
  return sort_rows_fused(keys_in, vals_in, key_scratch, n, end_bit, digit_bits, ws, ws_bytes, st,
                         e, n_dev);
}

int sort_near_rows_fused(const uint32_t* keys_in, const uint32_t* vals_in,
                         uint32_t* key_scratch, int64_t n, const int64_t* n_dev, int end_bit,
                         int digit_bits, int out_i64, void* neighbors, uint32_t* row_counts,
                         int64_t rows, void* ws, size_t ws_bytes, hipStream_t st) {
  if (out_i64) {
    EmitNearList<int64_t> e{(int64_t*)neighbors, row_counts, (uint32_t)rows};
    return sort_rows_fused(keys_in, vals_in, key_scratch, n, end_bit, digit_bits, ws, ws_bytes,
                           st, e, n_dev);
  }
  EmitNearList<int32_t> e{(int32_t*)neighbors, row_counts, (uint32_t)rows};
  return sort_rows_fused(keys_in, vals_in, key_scratch, n, end_bit, digit_bits, ws, ws_bytes, st,
                         e, n_dev);
}

template <typename K>
static int sort_pairs_impl(const K* keys_in, const uint32_t* vals_in, K* keys_out,
                           uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                           size_t ws_bytes, hipStream_t st, const uint32_t* hist_pre,
                           int digit_bits) {
  if (n <= 0) return YGG_OK;
  if (n > (int64_t)kCountMask) return YGG_ERR_ARG;
  if (end_bit <= begin_bit || keys_in == keys_out || begin_bit < 0 ||
      end_bit > 8 * (int)sizeof(K))
    return YGG_ERR_ARG;
  if (digit_bits == 0) digit_bits = default_digit_bits((int)sizeof(K), end_bit - begin_bit);
  switch (digit_bits) {
    case 8:
      return sort_pairs_bits<K, 8>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                   ws, ws_bytes, st, hist_pre);
    case 9:
      return sort_pairs_bits<K, 9>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                   ws, ws_bytes, st, hist_pre);
    case 10:
      return sort_pairs_bits<K, 10>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                    ws, ws_bytes, st, hist_pre);
    case 11:
      return sort_pairs_bits<K, 11>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                    ws, ws_bytes, st, hist_pre);
    default:
      return YGG_ERR_ARG;
  }
}

int sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                   uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                   size_t ws_bytes, hipStream_t st, const uint32_t* hist, int digit_bits) {
  return sort_pairs_impl<uint64_t>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                   ws, ws_bytes, st, hist, digit_bits);
}

int sort_pairs_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                   uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, void* ws,
                   size_t ws_bytes, hipStream_t st, int digit_bits) {
  return sort_pairs_impl<uint32_t>(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit,
                                   ws, ws_bytes, st, nullptr, digit_bits);
}

// ---- exclusive scan: int32 counts -> int64 offsets --------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t wave_inclusive_sum64(int64_t v) {
  const int lane = (int)lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ int64_t block_exclusive_scan64(int64_t v, int64_t* out_excl) {
  __shared__ int64_t wsum[kScanThreads / YGG_WAVE];
  int64_t inc = wave_inclusive_sum64(v);
  int w = threadIdx.x / YGG_WAVE;
  if (lane_id() == 63) wsum[w] = inc;
  __syncthreads();
  int64_t base = 0, total = 0;
  for (int q = 0; q < kScanThreads / YGG_WAVE; ++q) {
    if (q < w) base += wsum[q];
    total += wsum[q];
  }
  *out_excl = base + inc - v;
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(256) void k_scan_reduce(const int32_t* __restrict__ c,
                                                     const int32_t* __restrict__ gather,
                                                     int64_t n, int64_t* __restrict__ partial) {
  int64_t b = (int64_t)blockIdx.x * kScanTile;
  int64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = b + (int64_t)k * kScanThreads + threadIdx.x;
    if (i < n) s += c[gather ? gather[i] : i];
  }
  int64_t ex;
  int64_t tot = block_exclusive_scan64(s, &ex);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_partials(int64_t* partial, int64_t nb) {
  // single block: sequential chunks of 256
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    int64_t i = b0 + threadIdx.x;
    int64_t v = i < nb ? partial[i] : 0;
    int64_t ex;
    int64_t tot = block_exclusive_scan64(v, &ex);
    if (i < nb) partial[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ __launch_bounds__(256) void k_scan_apply(const int32_t* __restrict__ c,
                                                    const int32_t* __restrict__ gather,
                                                    int64_t n, const int64_t* __restrict__ partial,
                                                    int64_t* __restrict__ out) {
  int64_t b = (int64_t)blockIdx.x * kScanTile;
  // blocked: thread t owns items [b + t*items, ...)
  int64_t vals[kScanItems];
  int64_t s = 0;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = b + (int64_t)threadIdx.x * kScanItems + k;
    vals[k] = i < n ? (int64_t)c[gather ? gather[i] : i] : 0;
    s += vals[k];
  }
  int64_t ex;
  block_exclusive_scan64(s, &ex);
  int64_t run = partial[blockIdx.x] + ex;
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = b + (int64_t)threadIdx.x * kScanItems + k;
    if (i < n) out[i] = run;
    run += vals[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = partial[gridDim.x];
}

size_t scan_workspace_bytes(int64_t n) {
  return align_up(sizeof(int64_t) * (size_t)((n + kScanTile - 1) / kScanTile + 2));
}

int exclusive_scan_i32_to_i64(const int32_t* counts, const int32_t* gather, int64_t* out,
                              int64_t n, void* ws, size_t ws_bytes, hipStream_t st) {
  if (n <= 0) return fill_u32((uint32_t*)out, 0u, 2, st);
  int64_t nb = (n + kScanTile - 1) / kScanTile;
  if (ws_bytes < scan_workspace_bytes(n)) return YGG_ERR_WORKSPACE;
  int64_t* partial = (int64_t*)ws;
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanThreads), 0, st, counts, gather,
                     n, partial);
  YGG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, st, partial, nb);
  YGG_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kScanThreads), 0, st, counts, gather,
                     n, partial, out);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // namespace ygg

extern "C" {

size_t ygg_sort_workspace_bytes(int64_t n, int key_bytes, int begin_bit, int end_bit,
                                int digit_bits) {
  return ygg::sort_workspace_bytes(n, key_bytes, end_bit - begin_bit, digit_bits);
}

int ygg_sort_default_digit_bits(int key_bytes, int begin_bit, int end_bit) {
  return ygg::default_digit_bits(key_bytes, end_bit - begin_bit);
}

int ygg_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                       uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, int digit_bits,
                       void* ws, size_t ws_bytes, void* stream) {
  return ygg::sort_pairs_u64(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit, ws,
                             ws_bytes, (hipStream_t)stream, nullptr, digit_bits);
}

int ygg_sort_pairs_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                       uint32_t* vals_out, int64_t n, int begin_bit, int end_bit, int digit_bits,
                       void* ws, size_t ws_bytes, void* stream) {
  return ygg::sort_pairs_u32(keys_in, vals_in, keys_out, vals_out, n, begin_bit, end_bit, ws,
                             ws_bytes, (hipStream_t)stream, digit_bits);
}

}  // extern "C"
