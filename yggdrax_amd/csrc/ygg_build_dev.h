// Device code shared by the single-tree build kernels (ygg_build.hip) and
// the batched (segmented) build kernels (ygg_batch.hip).
#pragma once

#include "ygg_common.h"

namespace ygg {

// Order-preserving float <-> uint32 mapping for atomic min/max.
__device__ __forceinline__ uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// infer_bounds padding (bounds.py:29-33): pad = max(5 % of the span, 1e-6).
__device__ __forceinline__ void padded_bounds(float mn, float mx, float* lo, float* hi) {
  float span = mx - mn;
  float pad = span * 0.05f;
  if (pad < 1e-6f) pad = 1e-6f;
  *lo = mn - pad;
  *hi = mx + pad;
}

// Digit histogram add for pass p of a 63-bit Morton code. The six low
// passes (bits 0-47, octree levels 5 and deeper) hold nearly distinct digits
// across a wavefront of points in input order, so each lane adds with its
// own LDS atomic; the two top passes repeat digits within a cluster and are
// aggregated with ballots first. Measured on gfx950 (2^24 clustered points,
// encode + pack + histograms): 156 us, against 340 us with ballots on
// passes 1-7.
// Morton-code sorts take 9-bit digits over the 63 used bits: 7 passes
// (2^24 clustered codes on gfx950: 1.17 ms, against 1.30 ms for 8 passes of
// 8 bits on 4096-key tiles and 1.28 ms for rocPRIM's radix_sort_pairs).
constexpr int kMortonDigitBits = 9;
constexpr int kMortonEndBit = 63;
constexpr int kMortonPasses = (kMortonEndBit + kMortonDigitBits - 1) / kMortonDigitBits;
constexpr int kMortonBins = 1 << kMortonDigitBits;
constexpr int kMortonHistWords = kMortonPasses * kMortonBins;

// Digit histograms of one code for every Morton sort pass: plain LDS
// atomics for the low passes, ballot aggregation for the two top passes,
// whose digits repeat within a cluster.
__device__ __forceinline__ void morton_hist_add(uint32_t* h, uint64_t c, bool valid, int p) {
  const unsigned d = (unsigned)((c >> (kMortonDigitBits * p)) & (kMortonBins - 1));
  if (p < kMortonPasses - 2) {
    if (valid) atomicAdd(&h[p * kMortonBins + d], 1u);
  } else {
    const int used = p == kMortonPasses - 1 ? kMortonEndBit - kMortonDigitBits * p
                                            : kMortonDigitBits;
    pass_hist_add<kMortonDigitBits>(&h[p * kMortonBins], d, valid, p, used);
  }
}

// Morton code of one point (morton.py:68-77).
__device__ __forceinline__ uint64_t encode_point(float x, float y, float z, const float* lo,
                                                 const float* w) {
  float p[3] = {x, y, z};
  uint64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float v = __fdiv_rn(p[d] - lo[d], w[d]);
    v = v < 0.0f ? 0.0f : v;
    v = v > 1.0f ? 1.0f : v;
    c[d] = (uint64_t)rintf(v * 2097151.0f);
  }
  return spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2);
}

__device__ __forceinline__ int delta_node(const uint64_t* lc, int L, int64_t i, int64_t j) {
  if (j < 0 || j >= L) return -1;
  uint64_t ci = lc[i], cj = lc[j];
  if (ci == cj) return 64 + clz64((uint64_t)i ^ (uint64_t)j);
  return clz64(ci ^ cj);
}

__device__ __forceinline__ int lcp_codes(uint64_t a, uint64_t b) {
  uint64_t x = a ^ b;
  return x == 0 ? 64 : clz64(x);
}

// Internal node i of a Karras tree over L leaf codes (_tree_impl.py:1117-1270):
// children, leaf flags, particle range, and the parent scatter
// parent.at[left].set(i) then parent.at[right].set(i) with sequential
// last-writer semantics: right writes beat left writes, larger i wins. All
// indices are local to the tree; the arrays are the tree's own.
__device__ __forceinline__ void karras_node(const uint64_t* __restrict__ lc, int L, int64_t n,
                                            int64_t leaf, int i, int32_t* __restrict__ left,
                                            int32_t* __restrict__ right,
                                            uint8_t* __restrict__ lflag,
                                            uint8_t* __restrict__ rflag,
                                            int32_t* __restrict__ ranges,
                                            uint32_t* __restrict__ penc) {
  const int NI = L - 1;
  int dp = delta_node(lc, L, i, (int64_t)i + 1);
  int dm = delta_node(lc, L, i, (int64_t)i - 1);
  int64_t d = (dp > dm) ? 1 : ((dp < dm) ? -1 : 0);
  if (i == 0) d = 1;
  if (d == 0) d = 1;
  int dmin = delta_node(lc, L, i, (int64_t)i - d);
  int64_t span = 0;
  for (int k = 30; k >= 0; --k) {
    int64_t step = (int64_t)1 << k;
    int64_t cand = (int64_t)i + (span + step) * d;
    if (cand >= 0 && cand < L && delta_node(lc, L, i, cand) > dmin) span += step;
  }
  int64_t e = (int64_t)i + span * d;
  int64_t first = min((int64_t)i, e), last = max((int64_t)i, e);
  ranges[2 * i] = (int32_t)(first * leaf);
  ranges[2 * i + 1] = (int32_t)(min((last + 1) * leaf, n) - 1);
  uint64_t fc = lc[first];
  int common = lcp_codes(fc, lc[last]);
  int64_t split = first;
  for (int k = 30; k >= 0; --k) {
    int64_t mid = split + ((int64_t)1 << k);
    int64_t safe_mid = mid < last ? mid : last;
    if (mid < last && lcp_codes(fc, lc[safe_mid]) > common) split = mid;
  }
  bool lf = split == first, rf = split + 1 == last;
  int32_t l = lf ? (int32_t)(NI + first) : (int32_t)split;
  int32_t r = rf ? (int32_t)(NI + last) : (int32_t)(split + 1);
  left[i] = l;
  right[i] = r;
  lflag[i] = lf;
  rflag[i] = rf;
  // parent.at[left].set(i) then parent.at[right].set(i) with sequential
  // last-writer semantics: right writes beat left writes, larger i wins.
  atomicMax(&penc[l], (uint32_t)(i + 1));
  atomicMax(&penc[r], 0x80000000u | (uint32_t)(i + 1));
}

__device__ __forceinline__ int32_t dec_parent(uint32_t e) {
  return e == 0 ? -1 : (int32_t)(e & 0x7FFFFFFFu) - 1;
}

__device__ __forceinline__ void candidate(const uint64_t* rc, const int32_t* rd, int L, int id,
                                          uint64_t* c, int* d) {
  if (id < L) {
    *c = rc[id];
    *d = rd[id];
  } else {
    int j = id - L;
    int dd = common_depth(rc[j], rc[j + 1]);
    *d = dd;
    *c = prefix_code(rc[j], dd);
  }
}

}  // namespace ygg
