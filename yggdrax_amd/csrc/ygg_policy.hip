// Custom pair-policy traversal (_resolve_pair_actions,
// _interactions_impl.py:731-810): the walk advances one generation at a
// time; k_policy_geometry exports each pair's geometry for the host-side
// policy, which is evaluated in both directions; k_policy_apply combines the
// two verdicts (accept only if both accept, near only if both say near, else
// refine), appends far entries with their directed tags, near entries and
// the refined child pairs (same split rule and child order as the MAC walk).
// A policy given as a descriptor is evaluated on the device instead of on
// the host (k_policy_descriptor).
#include "ygg_common.h"
#include "ygg_walk.h"

namespace ygg {

__global__ void k_policy_geometry(const NodeRec* __restrict__ nodes,
                                  const uint64_t* __restrict__ wf, int64_t n, int NI,
                                  int mac_type, float theta_sq, float theta_root,
                                  int32_t* __restrict__ tgt, int32_t* __restrict__ src,
                                  float* __restrict__ ct, float* __restrict__ cs,
                                  float* __restrict__ dist_sq, float* __restrict__ ext_t,
                                  float* __restrict__ ext_s, uint8_t* __restrict__ mac_ok,
                                  uint8_t* __restrict__ t_leaf, uint8_t* __restrict__ s_leaf) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t pr = wf[i];
    const int32_t t = (int32_t)(pr >> 32), s = (int32_t)(uint32_t)pr;
    const NodeRec nt = nodes[t], ns = nodes[s];
    float dx = nt.cx - ns.cx, dy = nt.cy - ns.cy, dz = nt.cz - ns.cz;
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    const bool diff = t != s;
    bool ok;
    if (mac_type == YGG_MAC_ENGBLOM) {
      float rb = fmaxf(nt.ext, ns.ext), rs = fminf(nt.ext, ns.ext);
      float lhs = rb + theta_root * rs;
      ok = diff && d2 > 0.0f && (lhs * lhs <= theta_sq * d2);
    } else {
      float sum = nt.ext + ns.ext;
      ok = diff && d2 > 0.0f && (sum * sum <= theta_sq * d2);
    }
    tgt[i] = t;
    src[i] = s;
    ct[3 * i] = nt.cx;
    ct[3 * i + 1] = nt.cy;
    ct[3 * i + 2] = nt.cz;
    cs[3 * i] = ns.cx;
    cs[3 * i + 1] = ns.cy;
    cs[3 * i + 2] = ns.cz;
    dist_sq[i] = d2;
    ext_t[i] = nt.ext;
    ext_s[i] = ns.ext;
    mac_ok[i] = ok;
    t_leaf[i] = t >= NI;
    s_leaf[i] = s >= NI;
  }
}

// One thread per pair; appends are reserved with one atomic per wave.
__global__ __launch_bounds__(256) void k_policy_apply(
    const NodeRec* __restrict__ nodes, const uint64_t* __restrict__ wf, int64_t n, int NI,
    const int32_t* __restrict__ act_f, const int32_t* __restrict__ act_r,
    const int32_t* __restrict__ tag_f, const int32_t* __restrict__ tag_r,
    uint64_t* __restrict__ wf_out, int64_t queue_cap, uint32_t* __restrict__ out_size,
    uint32_t* __restrict__ far_keys, uint32_t* __restrict__ far_vals,
    int32_t* __restrict__ far_tags, int64_t far_cap, uint32_t* __restrict__ near_keys,
    uint32_t* __restrict__ near_vals, int64_t near_cap,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ flags) {
  const unsigned lane = lane_id();
  for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < n;
       b0 += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = b0 + threadIdx.x;
    int kind = 0;  // 0 none, 1 accept, 2 near, 3 refine
    int32_t t = 0, s = 0, tf = -1, tr = -1;
    NodeRec nt, ns;
    unsigned nc = 0;
    uint32_t ca[4], cb[4];
    if (i < n) {
      const uint64_t pr = wf[i];
      t = (int32_t)(pr >> 32);
      s = (int32_t)(uint32_t)pr;
      nt = nodes[t];
      ns = nodes[s];
      const int af = act_f[i], ar = act_r[i];
      if (af == 0 && ar == 0) {
        kind = 1;
        tf = tag_f[i];
        tr = tag_r[i];
      } else if (af == 1 && ar == 1) {
        kind = 2;
      } else {
        kind = 3;
        const bool same = t == s, t_int = t < NI, s_int = s < NI;
        const bool split_t = t_int && (same || !s_int || nt.ext >= ns.ext);
        const bool split_s = s_int && (same || !t_int || ns.ext > nt.ext);
        if (split_t && split_s && same) {
          ca[0] = nt.left;  cb[0] = nt.left;
          ca[1] = nt.left;  cb[1] = nt.right;
          ca[2] = nt.right; cb[2] = nt.right;
          nc = 3;
        } else if (split_t && split_s) {
          ca[0] = nt.left;  cb[0] = ns.left;
          ca[1] = nt.left;  cb[1] = ns.right;
          ca[2] = nt.right; cb[2] = ns.left;
          ca[3] = nt.right; cb[3] = ns.right;
          nc = 4;
        } else if (split_t) {
          ca[0] = nt.left;  cb[0] = s;
          ca[1] = nt.right; cb[1] = s;
          nc = 2;
        } else if (split_s) {
          ca[0] = t; cb[0] = ns.left;
          ca[1] = t; cb[1] = ns.right;
          nc = 2;
        }
      }
    }
    const unsigned nf = kind == 1 ? 2u : 0u, nn = kind == 2 ? 2u : 0u;
    const unsigned ic = wave_inclusive_sum(nc), iff = wave_inclusive_sum(nf),
                   inn = wave_inclusive_sum(nn);
    const unsigned tc = __shfl(ic, 63, 64), tff = __shfl(iff, 63, 64), tnn = __shfl(inn, 63, 64);
    unsigned long long bc = 0, bf = 0, bn = 0;
    if (lane == 63) {
      if (tc) bc = atomicAdd(out_size, tc);
      if (tff) bf = atomicAdd(&counters[0], (unsigned long long)tff);
      if (tnn) bn = atomicAdd(&counters[1], (unsigned long long)tnn);
    }
    bc = __shfl(bc, 63, 64);
    bf = __shfl(bf, 63, 64);
    bn = __shfl(bn, 63, 64);
    if (kind == 1) {
      int64_t f = (int64_t)bf + iff - nf;
      if (f + 1 < far_cap) {
        far_keys[f] = (uint32_t)nt.pad;
        far_vals[f] = (uint32_t)s;
        far_tags[f] = tf;
        far_keys[f + 1] = (uint32_t)ns.pad;
        far_vals[f + 1] = (uint32_t)t;
        far_tags[f + 1] = tr;
      } else {
        atomicOr(&flags[1], 1u);
      }
    } else if (kind == 2) {
      int64_t q = (int64_t)bn + inn - nn;
      if (q + 1 < near_cap && nt.leafpos >= 0 && ns.leafpos >= 0) {
        near_keys[q] = (uint32_t)nt.leafpos;
        near_vals[q] = (uint32_t)s;
        near_keys[q + 1] = (uint32_t)ns.leafpos;
        near_vals[q + 1] = (uint32_t)t;
      } else {
        atomicOr(&flags[2], 1u);
      }
    } else if (nc) {
      int64_t slot = (int64_t)bc + ic - nc;
      for (unsigned j = 0; j < nc; ++j, ++slot) {
        uint32_t lo = ca[j] < cb[j] ? ca[j] : cb[j], hi = ca[j] < cb[j] ? cb[j] : ca[j];
        if (slot < queue_cap)
          wf_out[slot] = ((uint64_t)lo << 32) | hi;
        else
          atomicOr(&flags[0], 1u);
      }
    }
  }
}

// ---- pair-policy descriptors ------------------------------------------------
// A policy given as data (YggPolicyDesc) is evaluated here for both
// directions of every pair, producing the same four arrays a host policy
// callable would, for k_policy_apply. Floating-point steps follow
// k_policy_geometry (dist_sq summed x, y, z without contraction).

__global__ void k_policy_descriptor(const NodeRec* __restrict__ nodes,
                                    const uint64_t* __restrict__ wf, int64_t n, int NI,
                                    int mac_type, float theta_sq, float theta_root,
                                    YggPolicyDesc d, int32_t* __restrict__ act_f,
                                    int32_t* __restrict__ act_r, int32_t* __restrict__ tag_f,
                                    int32_t* __restrict__ tag_r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t pr = wf[i];
    const int32_t t = (int32_t)(pr >> 32), s = (int32_t)(uint32_t)pr;
    const NodeRec nt = nodes[t], ns = nodes[s];
    float dx = nt.cx - ns.cx, dy = nt.cy - ns.cy, dz = nt.cz - ns.cz;
    float d2 = dx * dx;
    d2 = d2 + dy * dy;
    d2 = d2 + dz * dz;
    const bool diff = t != s;
    bool acc;
    switch (d.accept) {
      case YGG_ACCEPT_MAC: {
        if (mac_type == YGG_MAC_ENGBLOM) {
          float rb = fmaxf(nt.ext, ns.ext), rs = fminf(nt.ext, ns.ext);
          float lhs = rb + theta_root * rs;
          acc = diff && d2 > 0.0f && (lhs * lhs <= theta_sq * d2);
        } else {
          float sum = nt.ext + ns.ext;
          acc = diff && d2 > 0.0f && (sum * sum <= theta_sq * d2);
        }
        break;
      }
      case YGG_ACCEPT_DISTANCE:
        acc = diff && d2 >= d.accept_dist_sq;
        break;
      case YGG_ACCEPT_DIFFERENT:
        acc = diff;
        break;
      default:
        acc = false;
    }
    const bool near = d.near == YGG_NEAR_LEAF_PAIRS && !acc && diff && t >= NI && s >= NI;
    const int32_t a = acc ? 0 : (near ? 1 : 2);
    act_f[i] = a;
    act_r[i] = a;
    tag_f[i] = acc ? desc_tag(d, d2, nt.ext, ns.ext, t) : -1;
    tag_r[i] = acc ? desc_tag(d, d2, ns.ext, nt.ext, s) : -1;
  }
}

}  // namespace ygg

using namespace ygg;

extern "C" {

int ygg_policy_descriptor_actions(const void* nodes, const uint64_t* pairs, int64_t n,
                                  int64_t num_internal, int mac_type, float theta_sq,
                                  float theta_root, const YggPolicyDesc* desc,
                                  int32_t* actions_fwd, int32_t* actions_rev, int32_t* tags_fwd,
                                  int32_t* tags_rev, void* stream) {
  if (n < 0 || !nodes || !desc) return YGG_ERR_ARG;
  if (desc->num_thresholds < 0 || desc->num_thresholds > YGG_POLICY_MAX_THRESHOLDS)
    return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipLaunchKernelGGL(k_policy_descriptor, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, (const NodeRec*)nodes, pairs, n, (int)num_internal,
                     mac_type, theta_sq, theta_root, *desc, actions_fwd, actions_rev, tags_fwd,
                     tags_rev);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_policy_geometry(const void* nodes, const uint64_t* pairs, int64_t n, int64_t num_internal,
                        int mac_type, float theta_sq, float theta_root, int32_t* targets,
                        int32_t* sources, float* center_target, float* center_source,
                        float* dist_sq, float* extent_target, float* extent_source,
                        uint8_t* mac_ok, uint8_t* target_leaf, uint8_t* source_leaf,
                        void* stream) {
  if (n < 0 || !nodes) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipLaunchKernelGGL(k_policy_geometry, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, (const NodeRec*)nodes, pairs, n, (int)num_internal,
                     mac_type, theta_sq, theta_root, targets, sources, center_target,
                     center_source, dist_sq, extent_target, extent_source, mac_ok, target_leaf,
                     source_leaf);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

int ygg_policy_apply(const void* nodes, const uint64_t* pairs, int64_t n, int64_t num_internal,
                     const int32_t* actions_fwd, const int32_t* actions_rev,
                     const int32_t* tags_fwd, const int32_t* tags_rev, uint64_t* pairs_out,
                     int64_t queue_cap, uint32_t* out_size, uint32_t* far_keys,
                     uint32_t* far_vals, int32_t* far_tags, int64_t far_cap,
                     uint32_t* near_keys, uint32_t* near_vals, int64_t near_cap,
                     uint64_t* counters, uint32_t* flags, void* stream) {
  if (n < 0 || !nodes) return YGG_ERR_ARG;
  if (n == 0) return YGG_OK;
  hipLaunchKernelGGL(k_policy_apply, dim3(grid_for(n, 256, 8192)), dim3(256), 0,
                     (hipStream_t)stream, (const NodeRec*)nodes, pairs, n, (int)num_internal,
                     actions_fwd, actions_rev, tags_fwd, tags_rev, pairs_out, queue_cap,
                     out_size, far_keys, far_vals, far_tags, far_cap, near_keys, near_vals,
                     near_cap, (unsigned long long*)counters, flags);
  YGG_LAUNCH_CHECK();
  return YGG_OK;
}

}  // extern "C"
