// Dual-tree walk with flat, block-aggregated emission.
//
// Same pair semantics as the per-row-slot walk (MAC, accept / near / refine,
// split rule and child order of _dual_tree_walk_impl,
// _interactions_impl.py:1379-1454), but instead of one device-scope atomic
// per list entry, each 256-thread workgroup evaluates a chunk of pairs
// (pairs per thread x 256), prefix-sums its outputs in LDS and reserves
// space in three append-only streams (next wavefront, far, near) with one
// atomic each. A far entry is (level-major rank of its target row, source
// node); a near entry is (native leaf row, neighbour leaf node); with pair
// records the streams hold one (t, s) of node ids per accepted or near pair
// instead of its two entries. Rows are formed afterwards by grouping on the
// row key (the bucketing of ygg_rowlists.hip, or a stable radix sort), so
// no per-row counters are needed. Wavefront sizes per generation are exact,
// which lets the host reproduce the reference's pair-queue capacity
// decisions.
#include "ygg_common.h"
#include "ygg_walk.h"

namespace ygg {

constexpr int kFlatThreads = 256;
// Pairs per thread of one chunk (the queue tile): a chunk reserves its
// outputs with one atomic per stream, so larger tiles mean fewer
// same-address atomics per generation (they serialise across the XCDs);
// smaller tiles spread small generations over more workgroups.

struct FlatArgs {
  int32_t num_internal;
  int32_t mac_type;
  float theta_sq, theta_root;
  int64_t queue_cap;
  int64_t far_cap, near_cap;
  const NodeRec* nodes;
  uint64_t* wf0;
  uint64_t* wf1;
  uint32_t* gen_sizes;
  uint32_t* flags;                 // [0] queue, [1] far stream, [2] near stream
  unsigned long long* counters;    // [0] far total, [1] near total (emission walk: [0]
                                   // holds far << 32 | near until k_unpack_totals)
  uint32_t* far_keys;
  uint32_t* far_vals;
  uint32_t* near_keys;
  uint32_t* near_vals;
  int32_t hist_gen;     // generation whose pairs are histogrammed by target node
  int32_t* node_hist;   // [T] or null
  const int32_t* roots; // forest seeds (r, r), or null for the single root
  int64_t num_roots;
  int32_t xcd_groups;   // > 1: chunk ranges per XCD (see k_walk_flat)
  // Emission walk without tags: one record per accepted or near pair, the
  // two node ids (t, s), instead of its two entries (t's row key, s) and
  // (s's row key, t); the caps and counts are then in records.
  int32_t pair_records;
  unsigned long long* spread;  // counting walk: [2 * spread_lanes] far/near totals, or null
  int32_t spread_lanes;
  YggPolicyDesc policy;  // accept / near rules (the built-in MAC and leaf pairs by default)
  int32_t* far_tags;     // emission walk with tags: [far_cap]
  // Resumable emission walk: one record of kGenRecWords words per
  // generation, or null. Record g holds generation g's entry counts packed
  // as far << 32 | near (word 0, the reservation counter) and the far and
  // near stream offsets at which it starts (words 1 and 2, the totals of
  // generations 0..g-1, written by workgroup 0 of generation g). Generation
  // g runs only if its worst case, two entries per pair, fits the remaining
  // capacity of both streams; otherwise it writes nothing and records g + 1
  // in flags[3], and the host resumes from g with larger streams.
  unsigned long long* gen_fill;
};

// Words per generation record of a resumable walk: one 128-byte line each,
// so the reservations of generation g never share a line with the record
// of generation g - 1 that every workgroup of g reads.
constexpr int kGenRecWords = YGG_WALK_GEN_RECORD_WORDS;

__device__ __forceinline__ unsigned long long wave_incl_u64(unsigned long long v) {
  const int lane = (int)lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    unsigned long long o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

template <bool EMIT, int kFlatPer, bool TAGS = false>
__global__ __launch_bounds__(256) void k_walk_flat(FlatArgs a, int g) {
  static_assert(!TAGS || EMIT, "tags are written by the emission walk");
  constexpr int kFlatChunk = kFlatThreads * kFlatPer;
  if (atomic_load_agent(&a.flags[0]) | atomic_load_agent(&a.flags[1]) |
      atomic_load_agent(&a.flags[2]) | atomic_load_agent(&a.flags[3]))
    return;
  const int64_t raw = a.gen_sizes[g];
  const int64_t size = raw < a.queue_cap ? raw : a.queue_cap;
  // Chunks of kFlatChunk pairs. Workgroups are dispatched round-robin over the
  // XCDs (workgroup b on XCD b % X); with xcd_groups = X the chunks split
  // into X contiguous ranges, one per XCD, so each XCD's L2 sees one part of
  // the wavefront instead of every X-th chunk of all of it.
  const int64_t nchunks = (size + kFlatChunk - 1) / kFlatChunk;
  int64_t c_first = blockIdx.x, c_end = nchunks, c_step = gridDim.x;
  if (a.xcd_groups > 1 && gridDim.x % a.xcd_groups == 0) {
    const int X = a.xcd_groups;
    const int g = blockIdx.x % X;
    c_first = nchunks * g / X + blockIdx.x / X;
    c_end = nchunks * (g + 1) / X;
    c_step = gridDim.x / X;
  }
  // Resumable emission: this generation's stream offsets (the previous
  // record's offsets plus its final counts, the same in every workgroup;
  // workgroup 0 stores them in record g, also when it has no chunk) and the
  // capacity check that pauses the walk before anything is written.
  __shared__ unsigned long long s_gbase[2];
  if (EMIT && a.gen_fill) {
    if (c_first >= c_end && blockIdx.x != 0) return;
    if (threadIdx.x == 0) {
      unsigned long long bf = 0, bn = 0;
      if (g > 0) {
        const unsigned long long* r = a.gen_fill + (size_t)(g - 1) * kGenRecWords;
        const unsigned long long c = r[0];
        bf = r[1] + (c >> 32);
        bn = r[2] + (c & 0xFFFFFFFFull);
      }
      s_gbase[0] = bf;
      s_gbase[1] = bn;
      if (blockIdx.x == 0) {
        a.gen_fill[(size_t)g * kGenRecWords + 1] = bf;
        a.gen_fill[(size_t)g * kGenRecWords + 2] = bn;
      }
    }
    __syncthreads();
    if (c_first >= c_end) return;
    const int64_t worst = a.pair_records ? size : 2 * size;
    if ((int64_t)s_gbase[0] + worst > a.far_cap || (int64_t)s_gbase[1] + worst > a.near_cap) {
      if (threadIdx.x == 0) atomicMax(&a.flags[3], (uint32_t)(g + 1));
      return;
    }
  }
  if (c_first >= c_end) return;
  const uint64_t* __restrict__ wf_in = (g & 1) ? a.wf1 : a.wf0;
  uint64_t* __restrict__ wf_out = (g & 1) ? a.wf0 : a.wf1;
  const NodeRec* __restrict__ nodes = a.nodes;
  const unsigned lane = lane_id();
  const int w = threadIdx.x / YGG_WAVE;
  __shared__ unsigned long long s_wave[kFlatThreads / YGG_WAVE];
  __shared__ unsigned long long s_base[3];
  // counting walk with spread counters: this workgroup's far/near totals,
  // added to one lane at the end instead of one reservation per chunk
  const bool spread = !EMIT && a.spread != nullptr;
  unsigned long long acc_f = 0, acc_n = 0;
  // Output staging. Each wavefront's outputs of one chunk occupy one
  // contiguous range of each stream (slots follow the block scan), so they
  // are gathered in LDS and written out by consecutive lanes. A pair yields
  // at most two far or two near entries: 2 * 64 * kFlatPer slots hold both
  // (far first, then near). Child pairs (up to four per pair) go through the
  // same buffer in rounds of kStage.
  // Writes that nobody reads (a pair with no output of that kind, a child
  // outside the current round) go to two spare slots past the range, so
  // every pair writes the same number of slots and the staging has no
  // per-pair branches.
  constexpr int kStage = 2 * YGG_WAVE * kFlatPer;
  constexpr unsigned kDump = kStage;
  __shared__ uint2 s_stage[kFlatThreads / YGG_WAVE][kStage + 2];
  __shared__ int32_t s_tag[kFlatThreads / YGG_WAVE][TAGS ? 2 * YGG_WAVE * kFlatPer : 1];

  for (int64_t ci = c_first; ci < c_end; ci += c_step) {
    const int64_t chunk = ci * kFlatChunk;
    // mode: 0 none, 1 accept, 2 near, 3 both-same, 4 both-cross, 5 split t, 6 split s
    int mode[kFlatPer];
    uint32_t pt[kFlatPer], ps[kFlatPer], x0[kFlatPer], x1[kFlatPer], x2[kFlatPer],
        x3[kFlatPer];
    unsigned nc = 0, nf = 0, nn = 0;
    // Memory first, decisions after: every pair word of the chunk, then the
    // node records of a group of four pairs, are loaded before any of them
    // is used, so their latencies overlap instead of adding up pair by pair
    // (a lane past the wavefront's end loads node 0 and decides nothing).
    // Measured on the 2^24 clustered walk at 8 pairs per thread: 7.8 ms with
    // the loads issued pair by pair, 6.9 ms with groups of two when the
    // decisions branched per pair; with branch-free decisions 6.20 ms with
    // groups of two, 6.13 ms with four (151 VGPRs, still 3 waves per SIMD)
    // and 7.38 ms with eight (2 waves per SIMD).
    constexpr uint64_t kNoPair = ~0ull;
    uint64_t prw[kFlatPer];
#pragma unroll
    for (int k = 0; k < kFlatPer; ++k) {
      const int64_t idx = chunk + k * kFlatThreads + threadIdx.x;
      prw[k] = idx < size ? wf_in[idx] : kNoPair;
    }
    constexpr int kGroup = kFlatPer < 4 ? kFlatPer : 4;
#pragma unroll
    for (int k0 = 0; k0 < kFlatPer; k0 += kGroup) {
    NodeRec rt[kGroup], rs[kGroup];
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const uint64_t pr = prw[k0 + j];
      const bool valid = pr != kNoPair;
      rt[j] = nodes[valid ? (int32_t)(pr >> 32) : 0];
      rs[j] = nodes[valid ? (int32_t)(uint32_t)pr : 0];
    }
#pragma unroll
    for (int j = 0; j < kGroup; ++j) {
      const int k = k0 + j;
      mode[k] = 0;
      const bool pvalid = prw[k] != kNoPair;
      if (TAGS && !pvalid) continue;
      const uint64_t pr = prw[k];
      const int32_t t = (int32_t)(pr >> 32), s = (int32_t)(uint32_t)pr;
      const NodeRec& nt = rt[j];
      const NodeRec& ns = rs[j];
      float dx = nt.cx - ns.cx, dy = nt.cy - ns.cy, dz = nt.cz - ns.cz;
      float d2 = dx * dx;
      d2 = d2 + dy * dy;
      d2 = d2 + dz * dz;
      const float et = nt.ext, es = ns.ext;
      const bool diff = t != s;
      bool ok;
      if (a.mac_type == YGG_MAC_ENGBLOM) {
        float rb = fmaxf(et, es), rs = fminf(et, es);
        float lhs = rb + a.theta_root * rs;
        ok = diff && d2 > 0.0f && (lhs * lhs <= a.theta_sq * d2);
      } else {
        float sum = et + es;
        ok = diff && d2 > 0.0f && (sum * sum <= a.theta_sq * d2);
      }
      // internal nodes carry children (node ids of a forest are not
      // split at num_internal)
      const bool t_int = nt.left >= 0, s_int = ns.left >= 0;
      if (a.node_hist && g == a.hist_gen && pvalid) atomicAdd(&a.node_hist[t], 1);
      pt[k] = (uint32_t)t;
      ps[k] = (uint32_t)s;
      // accept / near rules (uniform per launch): the built-in MAC with
      // near leaf pairs unless a policy descriptor says otherwise
      bool acc = ok;
      if (a.policy.accept == YGG_ACCEPT_DISTANCE)
        acc = diff && d2 >= a.policy.accept_dist_sq;
      else if (a.policy.accept == YGG_ACCEPT_DIFFERENT)
        acc = diff;
      else if (a.policy.accept == YGG_ACCEPT_NEVER)
        acc = false;
      const bool near_rule = a.policy.near == YGG_NEAR_LEAF_PAIRS;
      const unsigned per_pair = !TAGS && a.pair_records ? 1u : 2u;  // stream slots
      if (!TAGS) {
        // Every field from selects, with no per-pair branch: accept (1),
        // near (2) or a split (3 both of the same node, 4 both, 5 t only,
        // 6 s only); nothing for a lane past the wavefront's end. On the
        // 2^24 clustered walk the branching form took 6.88 ms, this one
        // 6.21 ms (fewer exec-mask updates, 139 instead of 155 VGPRs).
        const bool fa = pvalid && acc;
        const bool fn = pvalid && !acc && near_rule && !t_int && !s_int && diff;
        const bool same = !diff;
        const bool split_t = t_int && (same || !s_int || et >= es);
        const bool split_s = s_int && (same || !t_int || es > et);
        const int ms = split_t && split_s ? (same ? 3 : 4) : (split_t ? 5 : (split_s ? 6 : 0));
        const int m = fa ? 1 : (fn ? 2 : (pvalid ? ms : 0));
        mode[k] = m;
        x0[k] = (uint32_t)(fa ? nt.pad : (fn ? nt.leafpos : nt.left));
        x1[k] = (uint32_t)(fa ? ns.pad : (fn ? ns.leafpos : nt.right));
        x2[k] = (uint32_t)ns.left;
        x3[k] = (uint32_t)ns.right;
        nf += fa ? per_pair : 0u;
        nn += fn ? per_pair : 0u;
        nc += m == 3 ? 3u : (m == 4 ? 4u : (m >= 5 ? 2u : 0u));
        continue;
      }
      if (acc) {
        mode[k] = 1;
        x0[k] = (uint32_t)nt.pad;  // level-major rank of t
        x1[k] = (uint32_t)ns.pad;
        if (TAGS) {  // directed tags: (t <- s) and (s <- t)
          x2[k] = (uint32_t)desc_tag(a.policy, d2, et, es, t);
          x3[k] = (uint32_t)desc_tag(a.policy, d2, es, et, s);
        }
        nf += 2;
      } else if (near_rule && !t_int && !s_int && diff) {
        mode[k] = 2;
        x0[k] = (uint32_t)nt.leafpos;
        x1[k] = (uint32_t)ns.leafpos;
        nn += 2;
      } else {
        const bool same = !diff;
        const bool split_t = t_int && (same || !s_int || et >= es);
        const bool split_s = s_int && (same || !t_int || es > et);
        x0[k] = (uint32_t)nt.left;
        x1[k] = (uint32_t)nt.right;
        x2[k] = (uint32_t)ns.left;
        x3[k] = (uint32_t)ns.right;
        if (split_t && split_s) {
          mode[k] = same ? 3 : 4;
          nc += same ? 3 : 4;
        } else if (split_t) {
          mode[k] = 5;
          nc += 2;
        } else if (split_s) {
          mode[k] = 6;
          nc += 2;
        }
      }
    }
    }
    // Block exclusive scan of (children, far, near) packed in 21-bit fields.
    const unsigned long long packed =
        (unsigned long long)nc | ((unsigned long long)nf << 21) | ((unsigned long long)nn << 42);
    const unsigned long long incl = wave_incl_u64(packed);
    if (lane == 63) s_wave[w] = incl;
    __syncthreads();
    unsigned long long wbase = 0, btot = 0;
#pragma unroll
    for (int q = 0; q < kFlatThreads / YGG_WAVE; ++q) {
      if (q < w) wbase += s_wave[q];
      btot += s_wave[q];
    }
    const unsigned long long mask = (1ull << 21) - 1;
    if (threadIdx.x == 0) {
      const unsigned long long tc = btot & mask, tf = (btot >> 21) & mask,
                               tn = (btot >> 42) & mask;
      s_base[0] = tc ? atomicAdd(&a.gen_sizes[g + 1], (uint32_t)tc) : 0;
      if (spread) {
        acc_f += tf;
        acc_n += tn;
        s_base[1] = s_base[2] = 0;
      } else if (EMIT && a.gen_fill) {
        // this generation's far (high) and near (low 32 bits) counts in one
        // reservation; they stay below 2 * queue_cap < 2^32
        const unsigned long long add = (tf << 32) | tn;
        const unsigned long long old =
            add ? atomicAdd(&a.gen_fill[(size_t)g * kGenRecWords], add) : 0ull;
        s_base[1] = s_gbase[0] + (old >> 32);
        s_base[2] = s_gbase[1] + (old & 0xFFFFFFFFull);
      } else if (EMIT) {
        // far (high 32 bits) and near (low 32 bits) running totals in one
        // reservation; a field that would pass 2^32 marks the walk as
        // overflowed (its result is discarded)
        const unsigned long long add = (tf << 32) | tn;
        const unsigned long long old = add ? atomicAdd(&a.counters[0], add) : 0ull;
        s_base[1] = old >> 32;
        s_base[2] = old & 0xFFFFFFFFull;
        if ((old >> 32) + tf > 0xFFFFFFFFull || (old & 0xFFFFFFFFull) + tn > 0xFFFFFFFFull)
          atomicOr(&a.flags[1], 1u);
      } else {
        s_base[1] = tf ? atomicAdd(&a.counters[0], tf) : 0;
        s_base[2] = tn ? atomicAdd(&a.counters[1], tn) : 0;
      }
    }
    __syncthreads();
    const unsigned long long ex = wbase + incl - packed;
    // wavefront-local offsets of this lane's far and near entries
    const unsigned long long lex = incl - packed;
    const unsigned wave_nf = (unsigned)((s_wave[w] >> 21) & mask);
    unsigned lf = (unsigned)((lex >> 21) & mask);
    unsigned ln = wave_nf + (unsigned)((lex >> 42) & mask);
    bool qover = false, fover = false, nover = false;
    const bool records = !TAGS && a.pair_records;
    const unsigned per_pair = records ? 1u : 2u;
#pragma unroll
    for (int k = 0; k < kFlatPer; ++k) {
      const int m = mode[k];
      const bool isf = m == 1, isn = m == 2;
      if (EMIT && !TAGS) {
        // entries (t's row key, s) and (s's row key, t), or the record
        // (t, s) with the second store going to the spare slot
        const unsigned dst = isf ? lf : (isn ? ln : kDump);
        s_stage[w][dst] = make_uint2(records ? pt[k] : x0[k], ps[k]);
        s_stage[w][records ? kDump + 1 : dst + 1] = make_uint2(x1[k], pt[k]);
      } else if (EMIT) {
        if (isf) {
          s_stage[w][lf] = make_uint2(x0[k], ps[k]);
          s_stage[w][lf + 1] = make_uint2(x1[k], pt[k]);
          s_tag[w][lf] = (int32_t)x2[k];
          s_tag[w][lf + 1] = (int32_t)x3[k];
        } else if (isn) {
          s_stage[w][ln] = make_uint2(x0[k], ps[k]);
          s_stage[w][ln + 1] = make_uint2(x1[k], pt[k]);
        }
      }
      lf += isf ? per_pair : 0u;
      ln += isn ? per_pair : 0u;
    }
    if (EMIT) {
      __syncthreads();  // staged entries visible to the whole wavefront
      const unsigned wave_nn = (unsigned)((s_wave[w] >> 42) & mask);
      const unsigned long long wex = wbase;  // block-exclusive prefix of this wavefront
      const int64_t fbase = (int64_t)s_base[1] + (int64_t)((wex >> 21) & mask);
      const int64_t nbase = (int64_t)s_base[2] + (int64_t)((wex >> 42) & mask);
      // The streams are read only after the walk, by the row lists: written
      // non-temporally they leave L2 and the Infinity Cache to the node
      // records every generation gathers (2^24 clustered emission walk:
      // 5.69 -> 5.56 ms).
      for (unsigned j = lane; j < wave_nf; j += YGG_WAVE) {
        const int64_t slot = fbase + j;
        if (slot < a.far_cap) {
          const uint2 e = s_stage[w][j];
          __builtin_nontemporal_store(e.x, &a.far_keys[slot]);
          __builtin_nontemporal_store(e.y, &a.far_vals[slot]);
          if (TAGS) __builtin_nontemporal_store(s_tag[w][j], &a.far_tags[slot]);
        } else {
          fover = true;
        }
      }
      for (unsigned j = lane; j < wave_nn; j += YGG_WAVE) {
        const int64_t slot = nbase + j;
        if (slot < a.near_cap) {
          const uint2 e = s_stage[w][wave_nf + j];
          __builtin_nontemporal_store(e.x, &a.near_keys[slot]);
          __builtin_nontemporal_store(e.y, &a.near_vals[slot]);
        } else {
          nover = true;
        }
      }
    }
    // Child pairs, in rounds of kStage slots of the wavefront's range; the
    // round count is the largest over the block's wavefronts so every thread
    // reaches the same barriers.
    {
      unsigned max_nc = 0;
#pragma unroll
      for (int q = 0; q < kFlatThreads / YGG_WAVE; ++q)
        max_nc = max(max_nc, (unsigned)(s_wave[q] & mask));
      const unsigned wave_nc = (unsigned)(s_wave[w] & mask);
      const int64_t cbase = (int64_t)s_base[0] + (int64_t)(wbase & mask);
      const unsigned lc0 = (unsigned)(lex & mask);
      for (unsigned r0 = 0; r0 < max_nc; r0 += kStage) {
        if (EMIT || r0 > 0) __syncthreads();  // previous use of s_stage finished
        unsigned lc = lc0;
#pragma unroll
        for (int k = 0; k < kFlatPer; ++k) {
          const int m = mode[k];
          // children (a_j, b_j): mode 3 (t,t) split, 4 cross split,
          // 5 split t only, 6 split s only
          const uint32_t a0 = m == 6 ? pt[k] : x0[k];
          const uint32_t b0 = m == 3 ? x0[k] : (m == 5 ? ps[k] : x2[k]);
          const uint32_t a1 = m == 6 ? pt[k] : (m == 5 ? x1[k] : x0[k]);
          const uint32_t b1 = m == 3 ? x1[k] : (m == 5 ? ps[k] : x3[k]);
          const uint32_t a2 = x1[k];
          const uint32_t b2 = m == 3 ? x1[k] : x2[k];
          const uint32_t a3 = x1[k], b3 = x3[k];
          const unsigned cnt = m < 3 ? 0u : (m == 3 ? 3u : (m == 4 ? 4u : 2u));
          // a slot outside this round's window (the offset wraps when
          // lc < r0), or past the pair's children, is the dump slot
          unsigned o = lc - r0;
          s_stage[w][cnt && o < (unsigned)kStage ? o : kDump] = make_uint2(min(a0, b0), max(a0, b0));
          o = lc + 1 - r0;
          s_stage[w][cnt && o < (unsigned)kStage ? o : kDump] = make_uint2(min(a1, b1), max(a1, b1));
          o = lc + 2 - r0;
          s_stage[w][cnt > 2 && o < (unsigned)kStage ? o : kDump] =
              make_uint2(min(a2, b2), max(a2, b2));
          o = lc + 3 - r0;
          s_stage[w][cnt > 3 && o < (unsigned)kStage ? o : kDump] =
              make_uint2(min(a3, b3), max(a3, b3));
          lc += cnt;
        }
        __syncthreads();
        const unsigned n_r = wave_nc > r0 ? min(wave_nc - r0, (unsigned)kStage) : 0u;
        for (unsigned j = lane; j < n_r; j += YGG_WAVE) {
          const int64_t slot = cbase + r0 + j;
          if (slot < a.queue_cap) {
            const uint2 e = s_stage[w][j];
            wf_out[slot] = ((uint64_t)e.x << 32) | e.y;
          } else {
            qover = true;
          }
        }
      }
    }
    if (qover) atomicOr(&a.flags[0], 1u);
    if (fover) atomicOr(&a.flags[1], 1u);
    if (nover) atomicOr(&a.flags[2], 1u);
    __syncthreads();  // s_wave / s_base / s_stage reuse
  }
  if (spread && threadIdx.x == 0 && (acc_f | acc_n)) {
    const int lane_id = (int)(blockIdx.x % (unsigned)a.spread_lanes);
    if (acc_f) atomicAdd(&a.spread[2 * lane_id], acc_f);
    if (acc_n) atomicAdd(&a.spread[2 * lane_id + 1], acc_n);
  }
}

// The emission walk's packed running total (far << 32 | near) in
// counters[0], unpacked into counters[0] (far) and counters[1] (near); for a
// resumable walk, the sums of the per-generation counts of generations
// [0, G) (zero for generations that have not run). One wavefront.
__global__ void k_unpack_totals(FlatArgs a, int G) {
  if (a.gen_fill) {
    unsigned long long f = 0, n = 0;
    for (int j = (int)threadIdx.x; j < G; j += YGG_WAVE) {
      const unsigned long long v = a.gen_fill[(size_t)j * kGenRecWords];
      f += v >> 32;
      n += v & 0xFFFFFFFFull;
    }
#pragma unroll
    for (int off = YGG_WAVE / 2; off > 0; off >>= 1) {
      f += __shfl_xor(f, off, YGG_WAVE);
      n += __shfl_xor(n, off, YGG_WAVE);
    }
    if (threadIdx.x == 0) {
      a.counters[0] = f;
      a.counters[1] = n;
    }
  } else if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long c = a.counters[0];
    a.counters[0] = c >> 32;
    a.counters[1] = c & 0xFFFFFFFFull;
  }
}

// Resuming a paused walk: clear the pause mark.
__global__ void k_flat_resume(FlatArgs a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) a.flags[3] = 0;
}

// counters[0..1] = the sums of the spread lanes (one workgroup).
__global__ void k_spread_sum(FlatArgs a) {
  __shared__ unsigned long long s_sum[2][256];
  unsigned long long f = 0, n = 0;
  for (int l = threadIdx.x; l < a.spread_lanes; l += blockDim.x) {
    f += a.spread[2 * l];
    n += a.spread[2 * l + 1];
  }
  s_sum[0][threadIdx.x] = f;
  s_sum[1][threadIdx.x] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tf = 0, tn = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      tf += s_sum[0][i];
      tn += s_sum[1][i];
    }
    a.counters[0] = tf;
    a.counters[1] = tn;
  }
}

__global__ void k_flat_init(FlatArgs a, int root, int G) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = i; k <= G; k += gridDim.x * blockDim.x) a.gen_sizes[k] = 0;
  if (i < 4) a.flags[i] = 0;
  if (i < 2) a.counters[i] = 0;  // [0] far total, [1] near total
  if (a.gen_fill)
    for (int k = i; k < G * kGenRecWords; k += gridDim.x * blockDim.x) a.gen_fill[k] = 0;
  if (a.spread)
    for (int k = i; k < 2 * a.spread_lanes; k += gridDim.x * blockDim.x) a.spread[k] = 0;
  if (a.roots) {
    for (int64_t k = i; k < a.num_roots; k += (int64_t)gridDim.x * blockDim.x) {
      const uint32_t r = (uint32_t)a.roots[k];
      a.wf0[k] = ((uint64_t)r << 32) | r;
    }
    if (i == 0) a.gen_sizes[0] = (uint32_t)a.num_roots;
  } else if (i == 0) {
    a.gen_sizes[0] = 1;
    a.wf0[0] = ((uint64_t)(uint32_t)root << 32) | (uint32_t)root;
  }
}

// offsets[r] = lower_bound(keys, r) for r in [0, rows], keys sorted ascending
// (the first min(n, *n_dev) keys when a device-side count is given).
__global__ void k_row_offsets(const uint32_t* __restrict__ keys, int64_t n, int rows,
                              int64_t* __restrict__ off, const int64_t* __restrict__ n_dev) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > rows) return;
  if (n_dev && *n_dev < n) n = *n_dev > 0 ? *n_dev : 0;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)r)
      lo = mid + 1;
    else
      hi = mid;
  }
  off[r] = lo;
}

template <typename OutT>
__global__ void k_far_from_sorted(const uint32_t* __restrict__ keys,
                                  const uint32_t* __restrict__ vals, int64_t n, int rows,
                                  const int32_t* __restrict__ nbl,
                                  const int32_t* __restrict__ node_level,
                                  OutT* __restrict__ sources, OutT* __restrict__ targets,
                                  OutT* __restrict__ levels) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t key = keys[i];
    if (key >= (uint32_t)rows) continue;  // unused stream capacity (filler keys sort last)
    int32_t node = nbl[key];
    sources[i] = (OutT)vals[i];
    targets[i] = (OutT)node;
    levels[i] = (OutT)node_level[node];
  }
}

template <typename OutT>
__global__ void k_far_rows(const int64_t* __restrict__ off, int T,
                           const int32_t* __restrict__ nbl, OutT* __restrict__ counts,
                           OutT* __restrict__ node_offsets, int32_t* __restrict__ counts32) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= T) return;
  int32_t node = nbl[r];
  int64_t c = off[r + 1] - off[r];
  counts[node] = (OutT)c;
  counts32[node] = (int32_t)c;
  node_offsets[node] = (OutT)off[r];
  if (r == 0) node_offsets[T] = (OutT)off[T];
}

// Near lists: entries converted by a flat grid-stride loop (rows can hold
// anything from a few to tens of thousands of entries), rows by one thread
// each.
template <typename OutT>
__global__ __launch_bounds__(256) void k_near_entries(const int64_t* __restrict__ off, int L,
                                                      const uint32_t* __restrict__ vals,
                                                      OutT* __restrict__ neighbors) {
  const int64_t total = off[L];
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    neighbors[e] = (OutT)vals[e];
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_near_rows(int L, const int64_t* __restrict__ off,
                                                   OutT* __restrict__ counts,
                                                   OutT* __restrict__ offsets_out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r > L) return;
  const int64_t o = off[r];
  offsets_out[r] = (OutT)o;
  if (r < L) counts[r] = (OutT)(off[r + 1] - o);
}

// Dense [L][maxc] neighbour leaf rows, -1 padded: one thread per element,
// contiguous over the whole matrix.
template <typename OutT>
__global__ __launch_bounds__(256) void k_near_dense(int L, const int64_t* __restrict__ off,
                                                    const uint32_t* __restrict__ vals,
                                                    const int32_t* __restrict__ leaf_position,
                                                    int64_t maxc, OutT* __restrict__ nlp) {
  const int64_t total = (int64_t)L * maxc;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / maxc, e = i - r * maxc;
    const int64_t o = off[r];
    OutT v = (OutT)-1;
    if (o + e < off[r + 1]) v = (OutT)leaf_position[vals[o + e]];
    __builtin_nontemporal_store(v, &nlp[i]);
  }
}

// Far targets and target levels row by row: one wavefront fills row r's
// range [off[r], off[r + 1]) with its node and that node's level
// (coalesced stores; no per-entry key read).
template <typename OutT>
__global__ __launch_bounds__(256) void k_far_targets_rows(const int64_t* __restrict__ off,
                                                          int rows,
                                                          const int32_t* __restrict__ nbl,
                                                          const int32_t* __restrict__ node_level,
                                                          const int64_t* __restrict__ row_sub,
                                                          OutT* __restrict__ targets,
                                                          OutT* __restrict__ levels) {
  const unsigned lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / YGG_WAVE);
  for (int64_t r = (int64_t)blockIdx.x * (blockDim.x / YGG_WAVE) + threadIdx.x / YGG_WAVE;
       r < rows; r += nw) {
    const int64_t b = off[r], e = off[r + 1];
    if (b >= e) continue;
    const int32_t node = nbl[r];
    const OutT t = (OutT)((int64_t)node - (row_sub ? row_sub[r] : 0)),
               lv = (OutT)node_level[node];
    for (int64_t i = b + lane; i < e; i += YGG_WAVE) {
      targets[i] = t;
      levels[i] = lv;
    }
  }
}

// Dense [L][maxc] neighbour leaf rows from the near list itself.
template <typename OutT>
__global__ __launch_bounds__(256) void k_near_dense_from_list(
    int L, const int64_t* __restrict__ off, const OutT* __restrict__ nbrs,
    const int32_t* __restrict__ leaf_position, int64_t maxc, OutT* __restrict__ nlp) {
  const int64_t total = (int64_t)L * maxc;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / maxc, e = i - r * maxc;
    const int64_t o = off[r];
    OutT v = (OutT)-1;
    if (o + e < off[r + 1]) v = (OutT)leaf_position[(int64_t)nbrs[o + e]];
    __builtin_nontemporal_store(v, &nlp[i]);
  }
}

// ---- batched (forest) lists -> per-system lists ---------------------------------
// Per far entry, the target's system (largest s with node_off(s) <= t) by a
// binary search over the node offsets, staged in LDS (seg_stage).
template <typename OutT>
__global__ __launch_bounds__(256) void k_b_far_localize_entries(
    Seg sg, const int64_t* __restrict__ off, int64_t rows, int64_t n,
    OutT* __restrict__ sources, OutT* __restrict__ targets) {
  __shared__ int32_t s_noff[kSegLdsCap];
  const bool staged = seg_stage(s_noff, sg.S, [&](int s) { return (int32_t)sg.node_off(s); });
  __syncthreads();
  const int64_t total = off[rows] < n ? off[rows] : n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t t = (int64_t)targets[e];
    const int64_t n0 = staged ? s_noff[seg_find_lds(s_noff, sg.S, t)]
                              : sg.node_off(sg.of_node(t));
    targets[e] = (OutT)(t - n0);
    sources[e] = (OutT)((int64_t)sources[e] - n0);
  }
}

// offsets_out[node + s] for every node (ranks r in [0, Ttot)), and the end
// offset of every system at node_off(s + 1) + s.
template <typename OutT>
__global__ void k_b_far_localize_rows(Seg sg, const int64_t* __restrict__ off,
                                      const int32_t* __restrict__ nbl_global, int64_t Ttot,
                                      OutT* __restrict__ offsets_out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r < Ttot) {
    const int64_t node = nbl_global[r];
    const int s = sg.of_node(node);
    offsets_out[node + s] = (OutT)(off[r] - off[sg.node_off(s)]);
  }
  if (r < sg.S) {
    const int s = (int)r;
    offsets_out[sg.node_off(s + 1) + s] = (OutT)(off[sg.node_off(s + 1)] - off[sg.node_off(s)]);
  }
}

template <typename OutT>
__global__ void k_b_far_level_offsets(Seg sg, const int64_t* __restrict__ off,
                                      const int32_t* __restrict__ level_offsets,
                                      OutT* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 65 * (int64_t)sg.S) return;
  const int s = (int)(i / 65);
  const int64_t n0 = sg.node_off(s);
  out[i] = (OutT)(off[n0 + level_offsets[i]] - off[n0]);
}

// One wavefront per near row (batch leaf j): the neighbours of a leaf lie
// in its own system, found once per row instead of once per entry (0.80 ms
// against 1.75 ms for 1.6e8 entries in rows of ~400 over 1024 Plummer
// systems, gfx950; far rows of ~73 entries gain nothing from this). Lane 0
// writes the row's local offset; the first threads also write each
// system's end offset.
template <typename OutT>
__global__ __launch_bounds__(256) void k_b_near_localize(Seg sg, const int64_t* __restrict__ off,
                                                         int64_t n, OutT* __restrict__ neighbors,
                                                         OutT* __restrict__ offsets_out) {
  const int64_t Ltot = sg.leaf[sg.S];
  const int64_t total = off[Ltot] < n ? off[Ltot] : n;
  const unsigned lane = lane_id();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / YGG_WAVE);
  for (int64_t j = (int64_t)blockIdx.x * (blockDim.x / YGG_WAVE) + threadIdx.x / YGG_WAVE;
       j < Ltot; j += nw) {
    const int s = sg.of_leaf(j);
    const int64_t n0 = sg.node_off(s);
    const int64_t b = off[j] < total ? off[j] : total;
    const int64_t e = off[j + 1] < total ? off[j + 1] : total;
    for (int64_t k = b + lane; k < e; k += YGG_WAVE)
      neighbors[k] = (OutT)((int64_t)neighbors[k] - n0);
    if (lane == 0) offsets_out[j + s] = (OutT)(off[j] - off[sg.leaf_off(s)]);
  }
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < sg.S;
       s += (int64_t)gridDim.x * blockDim.x)
    offsets_out[sg.leaf_off((int)s + 1) + s] =
        (OutT)(off[sg.leaf_off((int)s + 1)] - off[sg.leaf_off((int)s)]);
}

}  // namespace ygg

using namespace ygg;

extern "C" {

static void launch_flat(const FlatArgs& a, int k, bool emit, int per, int blocks,
                        hipStream_t st) {
  if (emit && a.far_tags) {
    if (per >= 16)
      hipLaunchKernelGGL((k_walk_flat<true, 16, true>), dim3(blocks), dim3(256), 0, st, a, k);
    else if (per >= 8)
      hipLaunchKernelGGL((k_walk_flat<true, 8, true>), dim3(blocks), dim3(256), 0, st, a, k);
    else
      hipLaunchKernelGGL((k_walk_flat<true, 4, true>), dim3(blocks), dim3(256), 0, st, a, k);
  } else if (emit) {
    if (per >= 16)
      hipLaunchKernelGGL((k_walk_flat<true, 16>), dim3(blocks), dim3(256), 0, st, a, k);
    else if (per >= 8)
      hipLaunchKernelGGL((k_walk_flat<true, 8>), dim3(blocks), dim3(256), 0, st, a, k);
    else
      hipLaunchKernelGGL((k_walk_flat<true, 4>), dim3(blocks), dim3(256), 0, st, a, k);
  } else {
    if (per >= 16)
      hipLaunchKernelGGL((k_walk_flat<false, 16>), dim3(blocks), dim3(256), 0, st, a, k);
    else if (per >= 8)
      hipLaunchKernelGGL((k_walk_flat<false, 8>), dim3(blocks), dim3(256), 0, st, a, k);
    else
      hipLaunchKernelGGL((k_walk_flat<false, 4>), dim3(blocks), dim3(256), 0, st, a, k);
  }
}

// Accept / near rules and tag stream of the walk: a policy descriptor's, or
// the built-in MAC with near leaf pairs and no tags; and the stream format
// (pair records only for an emission walk without a policy).
static int set_policy(const YggFlatWalkParams* p, FlatArgs* a) {
  if (p->pair_records && (!p->emit || p->policy)) return YGG_ERR_ARG;
  a->pair_records = p->pair_records ? 1 : 0;
  YggPolicyDesc d = {};
  d.accept = YGG_ACCEPT_MAC;
  d.near = YGG_NEAR_LEAF_PAIRS;
  d.tag_metric = YGG_TAG_NONE;
  if (p->policy) {
    d = *p->policy;
    if (d.num_thresholds < 0 || d.num_thresholds > YGG_POLICY_MAX_THRESHOLDS) return YGG_ERR_ARG;
  }
  a->policy = d;
  a->far_tags = (p->emit && p->policy) ? p->far_tags : nullptr;
  if (p->emit && p->policy && !p->far_tags) return YGG_ERR_ARG;
  return YGG_OK;
}

// Workgroups for generation k: generation k holds at most roots * 4^k pairs
// (a pair has at most four children) and never more than the queue, so the
// first generations launch only the workgroups their chunks can use (a
// launch's cost grows with its workgroup count even when most exit at once).
// Grid of generation k: enough workgroups for the bound roots * 4^k pairs
// (capped by the queue), at most `blocks`; with a hint of the generation's
// pairs, at most 1.5x the hinted chunks but never under kMinHintBlocks (one
// per CU); a hint of 0xFFFFFFFF means unknown. Workgroups step over the
// chunks, so any grid covers the
// generation; the hint only spares the dispatch of workgroups that would
// exit at once (about 2 ns each, 16-19 us for 8192).
constexpr int kMinHintBlocks = 256;
static int gen_blocks(int k, int64_t roots, int64_t queue_cap, int per, int blocks,
                      const uint32_t* hint) {
  const int64_t chunk = 256 * (per >= 16 ? 16 : (per >= 8 ? 8 : 4));
  int64_t bound = roots;
  for (int i = 0; i < k && bound < queue_cap; ++i) bound *= 4;
  if (bound > queue_cap) bound = queue_cap;
  int64_t nb = (bound + chunk - 1) / chunk;
  if (hint && hint[k] != 0xFFFFFFFFu) {
    int64_t nh = ((int64_t)hint[k] * 3 / 2 + chunk - 1) / chunk;
    if (nh < kMinHintBlocks) nh = kMinHintBlocks;
    if (nh < nb) nb = nh;
  }
  return nb < blocks ? (int)(nb > 0 ? nb : 1) : blocks;
}

static hipEvent_t g_flat_events[2 * 1025];
static int g_flat_events_ready = 0;

int ygg_walk_flat(const YggFlatWalkParams* p, uint32_t* host_status, int32_t* num_generations,
                  float* gen_ms, void* stream) {
  if (!p || !host_status || !num_generations) return YGG_ERR_ARG;
  if (gen_ms && p->max_generations > 1024) return YGG_ERR_ARG;
  if (gen_ms && !g_flat_events_ready) {
    for (int k = 0; k < 2 * 1025; ++k) YGG_CHECK(hipEventCreate(&g_flat_events[k]));
    g_flat_events_ready = 1;
  }
  if (p->num_internal < 1 || p->queue_cap < 4 || p->max_generations < 2) return YGG_ERR_ARG;
  if (p->emit && (!p->far_keys || !p->far_vals || !p->near_keys || !p->near_vals))
    return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  FlatArgs a;
  a.num_internal = p->num_internal;
  a.mac_type = p->mac_type;
  a.theta_sq = p->theta_sq;
  a.theta_root = p->theta_root;
  a.queue_cap = p->queue_cap;
  a.far_cap = p->far_cap;
  a.near_cap = p->near_cap;
  a.nodes = (const NodeRec*)p->nodes;
  a.wf0 = p->wf0;
  a.wf1 = p->wf1;
  a.gen_sizes = p->gen_sizes;
  a.flags = p->flags;
  a.counters = (unsigned long long*)p->counters;
  a.far_keys = p->far_keys;
  a.far_vals = p->far_vals;
  a.near_keys = p->near_keys;
  a.near_vals = p->near_vals;
  a.hist_gen = p->hist_gen;
  a.node_hist = p->node_hist;
  a.roots = p->roots;
  a.num_roots = p->roots ? p->num_roots : 1;
  a.xcd_groups = p->xcd_groups;
  if (p->roots && (p->num_roots < 1 || p->num_roots > p->queue_cap)) return YGG_ERR_ARG;
  if (p->spread_counters && (p->spread_lanes < 1 || p->spread_lanes > 4096)) return YGG_ERR_ARG;
  a.spread = (unsigned long long*)p->spread_counters;
  a.spread_lanes = (int32_t)p->spread_lanes;
  if (int rc = set_policy(p, &a)) return rc;
  const int G = p->max_generations;
  a.gen_fill = p->emit ? (unsigned long long*)p->gen_fill : nullptr;
  const int g0 = p->start_generation;
  if (g0 < 0 || g0 >= G || (g0 > 0 && !a.gen_fill)) return YGG_ERR_ARG;
  if (g0 == 0) {
    const int64_t init_n = (G + 1) > a.num_roots ? (G + 1) : a.num_roots;
    hipLaunchKernelGGL(k_flat_init, dim3(grid_for(init_n, 256, 4096)), dim3(256), 0, st, a,
                       p->root, G);
  } else {
    hipLaunchKernelGGL(k_flat_resume, dim3(1), dim3(64), 0, st, a);
  }
  YGG_LAUNCH_CHECK();
  const int blocks = p->grid_blocks > 0 ? p->grid_blocks : 2048;
  int g = g0, batch = p->gen_hint > 0 ? p->gen_hint - g0 : 24;
  if (batch < 2) batch = 2;
  while (true) {
    int end = g + batch < G ? g + batch : G;
    for (int k = g; k < end; ++k) {
      if (gen_ms) YGG_CHECK(hipEventRecord(g_flat_events[2 * k], st));
      launch_flat(a, k, p->emit != 0, p->pairs_per_thread,
                  gen_blocks(k, a.num_roots, a.queue_cap, p->pairs_per_thread, blocks,
                             p->gen_pairs_hint),
                  st);
      if (gen_ms) YGG_CHECK(hipEventRecord(g_flat_events[2 * k + 1], st));
    }
    // A resumable walk's totals are sums over its generation records, which
    // the launched generations never modify afterwards: summing them after
    // every batch lets the batch's one copy carry them to the host.
    if (a.gen_fill) hipLaunchKernelGGL(k_unpack_totals, dim3(1), dim3(64), 0, st, a, G);
    YGG_LAUNCH_CHECK();
    YGG_CHECK(hipMemcpyAsync(host_status, p->gen_sizes, sizeof(uint32_t) * (G + 1),
                             hipMemcpyDeviceToHost, st));
    YGG_CHECK(hipMemcpyAsync(host_status + G + 1, p->flags, sizeof(uint32_t) * 4,
                             hipMemcpyDeviceToHost, st));
    if (a.gen_fill)
      YGG_CHECK(hipMemcpyAsync(host_status + G + 5, a.counters, sizeof(uint64_t) * 2,
                               hipMemcpyDeviceToHost, st));
    YGG_CHECK(hipStreamSynchronize(st));
    const uint32_t* fl = host_status + G + 1;
    int done = -1;
    for (int k = 0; k <= end; ++k)
      if (host_status[k] == 0) {
        done = k;
        break;
      }
    // a paused walk ran generations [0, fl[3] - 1)
    if (fl[3]) done = (int)fl[3] - 1;
    if ((fl[0] | fl[1] | fl[2]) || done >= 0) {
      *num_generations = done >= 0 ? done : end;
      if (a.spread && !p->emit) {
        hipLaunchKernelGGL(k_spread_sum, dim3(1), dim3(256), 0, st, a);
        YGG_LAUNCH_CHECK();
      }
      if (p->emit && !a.gen_fill) {
        hipLaunchKernelGGL(k_unpack_totals, dim3(1), dim3(64), 0, st, a, G);
        YGG_LAUNCH_CHECK();
      }
      if (!a.gen_fill) {  // the totals to the host (already there for a resumable walk)
        YGG_CHECK(hipMemcpyAsync(host_status + G + 5, a.counters, sizeof(uint64_t) * 2,
                                 hipMemcpyDeviceToHost, st));
        YGG_CHECK(hipStreamSynchronize(st));
      }
      if (gen_ms) {
        for (int k = 0; k < G; ++k) gen_ms[k] = 0.f;
        for (int k = 0; k < end; ++k)
          YGG_CHECK(hipEventElapsedTime(&gen_ms[k], g_flat_events[2 * k],
                                        g_flat_events[2 * k + 1]));
      }
      return YGG_OK;
    }
    if (end >= G) return YGG_ERR_UNSUPPORTED;
    g = end;
    batch = 8;
  }
}

// The same walk with a fixed number of generation launches and no host
// synchronisation (for stream capture into a HIP graph): launches past the
// last non-empty generation exit at once; counters, generation sizes and
// overflow flags stay on the device for the caller to read afterwards.
int ygg_walk_flat_fixed(const YggFlatWalkParams* p, void* stream) {
  if (!p) return YGG_ERR_ARG;
  if (p->num_internal < 1 || p->queue_cap < 4 || p->max_generations < 2) return YGG_ERR_ARG;
  if (p->emit && (!p->far_keys || !p->far_vals || !p->near_keys || !p->near_vals))
    return YGG_ERR_ARG;
  if (p->roots && (p->num_roots < 1 || p->num_roots > p->queue_cap)) return YGG_ERR_ARG;
  if (p->gen_fill || p->start_generation) return YGG_ERR_ARG;  // no host to resume a pause
  hipStream_t st = (hipStream_t)stream;
  FlatArgs a;
  a.gen_fill = nullptr;
  a.num_internal = p->num_internal;
  a.mac_type = p->mac_type;
  a.theta_sq = p->theta_sq;
  a.theta_root = p->theta_root;
  a.queue_cap = p->queue_cap;
  a.far_cap = p->far_cap;
  a.near_cap = p->near_cap;
  a.nodes = (const NodeRec*)p->nodes;
  a.wf0 = p->wf0;
  a.wf1 = p->wf1;
  a.gen_sizes = p->gen_sizes;
  a.flags = p->flags;
  a.counters = (unsigned long long*)p->counters;
  a.far_keys = p->far_keys;
  a.far_vals = p->far_vals;
  a.near_keys = p->near_keys;
  a.near_vals = p->near_vals;
  a.hist_gen = p->hist_gen;
  a.node_hist = p->node_hist;
  a.roots = p->roots;
  a.num_roots = p->roots ? p->num_roots : 1;
  a.xcd_groups = p->xcd_groups;
  if (p->spread_counters && (p->spread_lanes < 1 || p->spread_lanes > 4096)) return YGG_ERR_ARG;
  a.spread = (unsigned long long*)p->spread_counters;
  a.spread_lanes = (int32_t)p->spread_lanes;
  if (int rc = set_policy(p, &a)) return rc;
  const int G = p->max_generations;
  const int64_t init_n = (G + 1) > a.num_roots ? (G + 1) : a.num_roots;
  hipLaunchKernelGGL(k_flat_init, dim3(grid_for(init_n, 256, 4096)), dim3(256), 0, st, a, p->root,
                     G);
  const int blocks = p->grid_blocks > 0 ? p->grid_blocks : 2048;
  for (int k = 0; k < G; ++k)
    launch_flat(a, k, p->emit != 0, p->pairs_per_thread,
                gen_blocks(k, a.num_roots, a.queue_cap, p->pairs_per_thread, blocks,
                           p->gen_pairs_hint),
                st);
  if (a.spread && !p->emit) hipLaunchKernelGGL(k_spread_sum, dim3(1), dim3(256), 0, st, a);
  if (p->emit) hipLaunchKernelGGL(k_unpack_totals, dim3(1), dim3(64), 0, st, a, G);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_row_offsets(const uint32_t* sorted_keys, int64_t n, int64_t rows, int64_t* offsets,
                    void* stream) {
  if (rows < 0 || n < 0) return YGG_ERR_ARG;
  hipLaunchKernelGGL(k_row_offsets, dim3(grid_for(rows + 1, 256)), dim3(256), 0,
                     (hipStream_t)stream, sorted_keys, n, (int)rows, offsets, nullptr);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_far_lists_from_sorted(int64_t num_nodes, const uint32_t* keys, const uint32_t* vals,
                              int64_t n, const int64_t* lm_offsets, const int32_t* nodes_by_level,
                              const int32_t* node_level, int out_i64, void* sources,
                              void* targets, void* target_levels, void* counts,
                              void* node_offsets, int32_t* counts32, void* stream) {
  if (num_nodes < 1 || n < 0) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int T = (int)num_nodes;
  dim3 ge(grid_for(n > 0 ? n : 1, 256, 16384)), gr(grid_for(T, 256));
  if (out_i64) {
    if (n > 0)
      hipLaunchKernelGGL(k_far_from_sorted<int64_t>, ge, dim3(256), 0, st, keys, vals, n, T,
                         nodes_by_level, node_level, (int64_t*)sources, (int64_t*)targets,
                         (int64_t*)target_levels);
    hipLaunchKernelGGL(k_far_rows<int64_t>, gr, dim3(256), 0, st, lm_offsets, T, nodes_by_level,
                       (int64_t*)counts, (int64_t*)node_offsets, counts32);
  } else {
    if (n > 0)
      hipLaunchKernelGGL(k_far_from_sorted<int32_t>, ge, dim3(256), 0, st, keys, vals, n, T,
                         nodes_by_level, node_level, (int32_t*)sources, (int32_t*)targets,
                         (int32_t*)target_levels);
    hipLaunchKernelGGL(k_far_rows<int32_t>, gr, dim3(256), 0, st, lm_offsets, T, nodes_by_level,
                       (int32_t*)counts, (int32_t*)node_offsets, counts32);
  }
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

// The row lists are grouped by bucketing (ygg::row_bucket_*) when
// digit_bits is 0 and the keys span at most 22 bits, else by the stable
// radix sort with that digit width (0: its default).
static bool lists_bucketed(int end_bit, int digit_bits) {
  int hi = 0, lo = 0;
  return digit_bits == 0 && ygg::row_bucket_split(end_bit, &hi, &lo);
}

static size_t lists_group_ws(int64_t n, int end_bit, int digit_bits) {
  return lists_bucketed(end_bit, digit_bits)
             ? ygg::row_bucket_workspace_bytes(n, end_bit)
             : ygg::sort_rows_fused_workspace_bytes(n, end_bit, digit_bits);
}

// Workspace of the row lists: the grouping's own, then the per-row counts
// (2^end_bit rows at most), their scan and, for the bucketing, the row
// offsets by node of pair records.
static size_t lists_sort_ws(int64_t n, int end_bit, int digit_bits, int64_t* rows_max) {
  const int64_t rmax = end_bit >= 1 && end_bit <= 31 ? ((int64_t)1 << end_bit) : 1;
  if (rows_max) *rows_max = rmax;
  size_t b = align_up(lists_group_ws(n, end_bit, digit_bits)) +
             align_up(sizeof(uint32_t) * (size_t)rmax) + align_up(ygg::scan_workspace_bytes(rmax));
  if (lists_bucketed(end_bit, digit_bits)) b += sizeof(int64_t) * (size_t)(rmax + 1);
  return b;
}

// Splits the row-list workspace and zeroes the row counts.
static int lists_sort_split(int64_t n, int end_bit, int digit_bits, int64_t rows, void* ws,
                            size_t ws_bytes, hipStream_t st, size_t* sort_bytes,
                            uint32_t** row_counts, void** scan_ws, size_t* scan_bytes,
                            int64_t** key_offsets) {
  int64_t rmax = 0;
  if (ws_bytes < lists_sort_ws(n, end_bit, digit_bits, &rmax)) return YGG_ERR_WORKSPACE;
  if (rows > rmax) return YGG_ERR_ARG;  // row keys need end_bit bits
  *sort_bytes = align_up(lists_group_ws(n, end_bit, digit_bits));
  *row_counts = (uint32_t*)((char*)ws + *sort_bytes);
  *scan_ws = (char*)ws + *sort_bytes + align_up(sizeof(uint32_t) * (size_t)rmax);
  *scan_bytes = ygg::scan_workspace_bytes(rmax);
  *key_offsets = lists_bucketed(end_bit, digit_bits)
                     ? (int64_t*)((char*)*scan_ws + align_up(*scan_bytes))
                     : nullptr;
  return ygg::fill_u32(*row_counts, 0u, rows, st);
}

// Row offsets by node key for pair records: every key first gets the end
// of the lists (off[rows]), then row r's node gets off[r]. A key that is
// no row's node (never in the walk's records) so lists its entries inside
// the output, past every row's.
__global__ __launch_bounds__(256) void k_key_offsets_fill(int64_t* __restrict__ by_key,
                                                          int64_t keys_n,
                                                          const int64_t* __restrict__ off,
                                                          int rows) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i <= keys_n) by_key[i] = off[rows];
}
__global__ __launch_bounds__(256) void k_key_offsets_rows(int64_t* __restrict__ by_key,
                                                          int64_t keys_n,
                                                          const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ row_nodes,
                                                          int rows) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const int32_t k = row_nodes[i];
  if (k >= 0 && k < keys_n) by_key[k] = off[i];
}

// The lists of the bucketing's counts: row offsets (int64 off[rows + 1],
// the scan of the counts, through row_nodes for pair records), then the
// scatter of every value (by node key for pair records).
static int lists_offsets_scatter(const ygg::RowBucketState& bs, const uint32_t* row_counts,
                                 const int32_t* row_nodes, int rows, int64_t keys_n,
                                 int64_t* off, int64_t* key_offsets, void* scan_ws,
                                 size_t scan_bytes, const int64_t* row_sub, int out_i64,
                                 void* out, hipStream_t st) {
  int rc = ygg::exclusive_scan_i32_to_i64((const int32_t*)row_counts, row_nodes, off, rows,
                                          scan_ws, scan_bytes, st);
  if (rc) return rc;
  if (!row_nodes) return ygg::row_bucket_scatter(bs, off, row_sub, out_i64, out, st);
  hipLaunchKernelGGL(k_key_offsets_fill, dim3(grid_for(keys_n + 1, 256)), dim3(256), 0, st,
                     key_offsets, keys_n, off, rows);
  hipLaunchKernelGGL(k_key_offsets_rows, dim3(grid_for(rows, 256)), dim3(256), 0, st,
                     key_offsets, keys_n, off, row_nodes, rows);
  YGG_LAUNCH_CHECK();
  return ygg::row_bucket_scatter(bs, key_offsets, row_sub, out_i64, out, st);
}

size_t ygg_far_lists_sort_workspace_bytes(int64_t n, int end_bit, int digit_bits) {
  return lists_sort_ws(n, end_bit, digit_bits, nullptr);
}

int ygg_far_lists_sort(const uint32_t* keys, const uint32_t* vals, int64_t n,
                       const int64_t* count, int64_t num_nodes, int end_bit, int digit_bits,
                       const int32_t* nodes_by_level, const int32_t* node_level,
                       const int64_t* row_sub, int pair_records, int out_i64,
                       uint32_t* key_scratch, int64_t* lm_offsets, void* sources, void* targets,
                       void* target_levels, void* counts, void* node_offsets, int32_t* counts32,
                       void* ws, size_t ws_bytes, void* stream) {
  if (num_nodes < 1 || n < 0 || !lm_offsets || !key_scratch || !ws) return YGG_ERR_ARG;
  if ((row_sub || pair_records) && !lists_bucketed(end_bit, digit_bits))
    return YGG_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int T = (int)num_nodes;
  const int64_t entries = pair_records ? 2 * n : n;
  size_t sort_bytes = 0, scan_bytes = 0;
  uint32_t* row_counts = nullptr;
  void* scan_ws = nullptr;
  int64_t* key_offsets = nullptr;
  int rc = lists_sort_split(entries, end_bit, digit_bits, T, ws, ws_bytes, st, &sort_bytes,
                            &row_counts, &scan_ws, &scan_bytes, &key_offsets);
  if (rc) return rc;
  if (lists_bucketed(end_bit, digit_bits)) {
    // pair records hold node ids: counts by node, offsets in level-major
    // order through nodes_by_level
    ygg::RowBucketState bs;
    rc = ygg::row_bucket_count(keys, vals, pair_records != 0, n, count, end_bit, T, T, T,
                               key_scratch, row_counts, ws, sort_bytes, st, &bs);
    if (rc) return rc;
    rc = lists_offsets_scatter(bs, row_counts, pair_records ? nodes_by_level : nullptr, T, T,
                               lm_offsets, key_offsets, scan_ws, scan_bytes, row_sub, out_i64,
                               sources, st);
  } else {
    rc = ygg::sort_far_rows_fused(keys, vals, key_scratch, n, count, end_bit, digit_bits, out_i64,
                                  sources, row_counts, T, ws, sort_bytes, st);
    if (rc) return rc;
    rc = ygg::exclusive_scan_i32_to_i64((const int32_t*)row_counts, nullptr, lm_offsets, T,
                                        scan_ws, scan_bytes, st);
  }
  if (rc) return rc;
  dim3 gr(grid_for(T, 256)), gw(grid_for((int64_t)T, 256 / YGG_WAVE, 16384));
  if (out_i64) {
    hipLaunchKernelGGL(k_far_targets_rows<int64_t>, gw, dim3(256), 0, st, lm_offsets, T,
                       nodes_by_level, node_level, row_sub, (int64_t*)targets,
                       (int64_t*)target_levels);
    hipLaunchKernelGGL(k_far_rows<int64_t>, gr, dim3(256), 0, st, lm_offsets, T, nodes_by_level,
                       (int64_t*)counts, (int64_t*)node_offsets, counts32);
  } else {
    hipLaunchKernelGGL(k_far_targets_rows<int32_t>, gw, dim3(256), 0, st, lm_offsets, T,
                       nodes_by_level, node_level, row_sub, (int32_t*)targets,
                       (int32_t*)target_levels);
    hipLaunchKernelGGL(k_far_rows<int32_t>, gr, dim3(256), 0, st, lm_offsets, T, nodes_by_level,
                       (int32_t*)counts, (int32_t*)node_offsets, counts32);
  }
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

size_t ygg_near_lists_sort_workspace_bytes(int64_t n, int end_bit, int digit_bits) {
  return lists_sort_ws(n, end_bit, digit_bits, nullptr);
}

int ygg_near_lists_sort(const uint32_t* keys, const uint32_t* vals, int64_t n,
                        const int64_t* count, int64_t num_leaves, int end_bit, int digit_bits,
                        const int64_t* row_sub, const int32_t* leaf_nodes, int out_i64,
                        uint32_t* key_scratch, int64_t* row_offsets, void* neighbors,
                        void* counts, void* offsets, void* ws, size_t ws_bytes, void* stream) {
  if (num_leaves < 1 || n < 0 || !row_offsets || !key_scratch || !ws) return YGG_ERR_ARG;
  if ((row_sub || leaf_nodes) && !lists_bucketed(end_bit, digit_bits))
    return YGG_ERR_UNSUPPORTED;
  hipStream_t st = (hipStream_t)stream;
  const int L = (int)num_leaves;
  const int64_t entries = leaf_nodes ? 2 * n : n;
  // pair records are keyed by node id: the whole 2^end_bit key range
  const int64_t keys_n = leaf_nodes ? ((int64_t)1 << end_bit) : L;
  if (leaf_nodes && (end_bit < 1 || end_bit > 31)) return YGG_ERR_ARG;
  size_t sort_bytes = 0, scan_bytes = 0;
  uint32_t* row_counts = nullptr;
  void* scan_ws = nullptr;
  int64_t* key_offsets = nullptr;
  int rc = lists_sort_split(entries, end_bit, digit_bits, keys_n, ws, ws_bytes, st, &sort_bytes,
                            &row_counts, &scan_ws, &scan_bytes, &key_offsets);
  if (rc) return rc;
  if (lists_bucketed(end_bit, digit_bits)) {
    ygg::RowBucketState bs;
    rc = ygg::row_bucket_count(keys, vals, leaf_nodes != nullptr, n, count, end_bit, keys_n,
                               leaf_nodes ? (L < keys_n ? L : keys_n) : L,
                               leaf_nodes ? keys_n : 2 * (int64_t)L, key_scratch, row_counts, ws,
                               sort_bytes, st, &bs);
    if (rc) return rc;
    rc = lists_offsets_scatter(bs, row_counts, leaf_nodes, L, keys_n, row_offsets, key_offsets,
                               scan_ws, scan_bytes, row_sub, out_i64, neighbors, st);
  } else {
    rc = ygg::sort_near_rows_fused(keys, vals, key_scratch, n, count, end_bit, digit_bits,
                                   out_i64, neighbors, row_counts, L, ws, sort_bytes, st);
    if (rc) return rc;
    rc = ygg::exclusive_scan_i32_to_i64((const int32_t*)row_counts, nullptr, row_offsets, L,
                                        scan_ws, scan_bytes, st);
  }
  if (rc) return rc;
  dim3 gr(grid_for((int64_t)L + 1, 256));
  if (out_i64)
    hipLaunchKernelGGL(k_near_rows<int64_t>, gr, dim3(256), 0, st, L, row_offsets,
                       (int64_t*)counts, (int64_t*)offsets);
  else
    hipLaunchKernelGGL(k_near_rows<int32_t>, gr, dim3(256), 0, st, L, row_offsets,
                       (int32_t*)counts, (int32_t*)offsets);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_near_dense_from_list(int64_t num_leaves, const int64_t* row_offsets,
                             const void* neighbors, const int32_t* leaf_position,
                             int64_t max_count, int out_i64, void* neighbor_leaf_positions,
                             void* stream) {
  if (num_leaves < 1 || max_count < 0) return YGG_ERR_ARG;
  if (max_count == 0) return YGG_OK;
  const int L = (int)num_leaves;
  dim3 gd(grid_for((int64_t)L * max_count, 256, 32768));
  hipStream_t st = (hipStream_t)stream;
  if (out_i64)
    hipLaunchKernelGGL(k_near_dense_from_list<int64_t>, gd, dim3(256), 0, st, L, row_offsets,
                       (const int64_t*)neighbors, leaf_position, max_count,
                       (int64_t*)neighbor_leaf_positions);
  else
    hipLaunchKernelGGL(k_near_dense_from_list<int32_t>, gd, dim3(256), 0, st, L, row_offsets,
                       (const int32_t*)neighbors, leaf_position, max_count,
                       (int32_t*)neighbor_leaf_positions);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_near_lists_from_sorted(int64_t num_leaves, const int64_t* row_offsets,
                               const uint32_t* vals, const int32_t* leaf_position,
                               int64_t max_count, int out_i64, void* neighbors,
                               void* neighbor_leaf_positions, void* counts, void* offsets,
                               void* stream) {
  if (num_leaves < 1) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  int L = (int)num_leaves;
  dim3 grid(grid_for(max_count > 0 ? (int64_t)L * max_count : (int64_t)L * 256, 256, 16384));
  dim3 gr(grid_for((int64_t)L + 1, 256));
  dim3 gd(grid_for(L * max_count > 0 ? (int64_t)L * max_count : 1, 256, 32768));
  if (out_i64) {
    hipLaunchKernelGGL(k_near_entries<int64_t>, grid, dim3(256), 0, st, row_offsets, L, vals,
                       (int64_t*)neighbors);
    hipLaunchKernelGGL(k_near_rows<int64_t>, gr, dim3(256), 0, st, L, row_offsets,
                       (int64_t*)counts, (int64_t*)offsets);
    if (neighbor_leaf_positions && max_count > 0)
      hipLaunchKernelGGL(k_near_dense<int64_t>, gd, dim3(256), 0, st, L, row_offsets, vals,
                         leaf_position, max_count, (int64_t*)neighbor_leaf_positions);
  } else {
    hipLaunchKernelGGL(k_near_entries<int32_t>, grid, dim3(256), 0, st, row_offsets, L, vals,
                       (int32_t*)neighbors);
    hipLaunchKernelGGL(k_near_rows<int32_t>, gr, dim3(256), 0, st, L, row_offsets,
                       (int32_t*)counts, (int32_t*)offsets);
    if (neighbor_leaf_positions && max_count > 0)
      hipLaunchKernelGGL(k_near_dense<int32_t>, gd, dim3(256), 0, st, L, row_offsets, vals,
                         leaf_position, max_count, (int32_t*)neighbor_leaf_positions);
  }
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

static inline Seg seg_of(const YggBatch* b) {
  Seg sg;
  sg.S = b->num_systems;
  sg.leaf_size = b->leaf_size;
  sg.pt = b->point_offsets;
  sg.leaf = b->leaf_offsets;
  return sg;
}

int ygg_batch_far_localize(const YggBatch* b, const int64_t* row_offsets,
                           const int32_t* nodes_by_level_global, const int32_t* level_offsets,
                           int64_t n, int out_i64, void* sources, void* targets,
                           void* node_offsets_out, void* far_level_offsets_out, void* stream) {
  if (!b || b->num_systems < 1 || n < 0) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Seg sg = seg_of(b);
  const int64_t T = 2 * b->num_leaves - b->num_systems;
  const int64_t rows = T > b->num_systems ? T : b->num_systems;
  // entries: grid-stride over 2048 workgroups, each staging the node offsets once
  dim3 ge(grid_for(n > 0 ? n : 1, 256, 2048)), gr(grid_for(rows, 256)),
      gl(grid_for(65 * (int64_t)b->num_systems, 256));
  if (out_i64) {
    if (n > 0)
      hipLaunchKernelGGL(k_b_far_localize_entries<int64_t>, ge, dim3(256), 0, st, sg,
                         row_offsets, T, n, (int64_t*)sources, (int64_t*)targets);
    hipLaunchKernelGGL(k_b_far_localize_rows<int64_t>, gr, dim3(256), 0, st, sg, row_offsets,
                       nodes_by_level_global, T, (int64_t*)node_offsets_out);
    hipLaunchKernelGGL(k_b_far_level_offsets<int64_t>, gl, dim3(256), 0, st, sg, row_offsets,
                       level_offsets, (int64_t*)far_level_offsets_out);
  } else {
    if (n > 0)
      hipLaunchKernelGGL(k_b_far_localize_entries<int32_t>, ge, dim3(256), 0, st, sg,
                         row_offsets, T, n, (int32_t*)sources, (int32_t*)targets);
    hipLaunchKernelGGL(k_b_far_localize_rows<int32_t>, gr, dim3(256), 0, st, sg, row_offsets,
                       nodes_by_level_global, T, (int32_t*)node_offsets_out);
    hipLaunchKernelGGL(k_b_far_level_offsets<int32_t>, gl, dim3(256), 0, st, sg, row_offsets,
                       level_offsets, (int32_t*)far_level_offsets_out);
  }
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_batch_near_localize(const YggBatch* b, const int64_t* row_offsets, int64_t n,
                            int out_i64, void* neighbors, void* near_offsets_out, void* stream) {
  if (!b || b->num_systems < 1 || n < 0) return YGG_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const Seg sg = seg_of(b);
  // one wavefront per leaf row; the grid also covers the system end offsets
  const int64_t rows = b->num_leaves > 1 ? b->num_leaves : 1;
  dim3 g(grid_for(rows, 256 / YGG_WAVE, 16384));
  if (out_i64)
    hipLaunchKernelGGL(k_b_near_localize<int64_t>, g, dim3(256), 0, st, sg, row_offsets, n,
                       (int64_t*)neighbors, (int64_t*)near_offsets_out);
  else
    hipLaunchKernelGGL(k_b_near_localize<int32_t>, g, dim3(256), 0, st, sg, row_offsets, n,
                       (int32_t*)neighbors, (int32_t*)near_offsets_out);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
