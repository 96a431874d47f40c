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

}  // namespace ygg
