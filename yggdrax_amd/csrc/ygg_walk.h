// Traversal node record shared by the walk kernels.
#pragma once

#include <stdint.h>

namespace ygg {

// 32-byte node record: centre, MAC extent, children (-1 for leaves) and the
// native leaf row (-1 for internal nodes). Two 16-byte loads per node.
struct __attribute__((aligned(16))) NodeRec {
  float cx, cy, cz, ext;
  int32_t left, right, leafpos, pad;
};
static_assert(sizeof(NodeRec) == 32, "NodeRec must be 32 bytes");

// Tag of an accepted pair under a policy descriptor, target t with extent
// et against a source with extent es at squared distance d2.
__device__ __forceinline__ int32_t desc_tag(const YggPolicyDesc& d, float d2, float et, float es,
                                            int32_t target) {
  float m;
  switch (d.tag_metric) {
    case YGG_TAG_DIST_SQ:
      m = d2;
      break;
    case YGG_TAG_OPENING_RATIO: {
      const float sum = et + es;
      m = (sum * sum) / d2;
      break;
    }
    case YGG_TAG_TARGET_MOD:
      return d.tag_modulus > 0 ? target % d.tag_modulus : -1;
    default:
      return -1;
  }
  int k = 0;
  for (int j = 0; j < d.num_thresholds; ++j) k += (m >= d.thresholds[j]) ? 1 : 0;
  return d.tag_values[k];
}

}  // namespace ygg
