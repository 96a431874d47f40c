"""yggdrax_amd: the yggdrax tree pipeline on AMD Instinct MI355X (gfx950).

Drop-in surface of TobiBu/yggdrax's hot path -- build_tree / build_octree,
compute_tree_geometry, build_interactions_and_neighbors,
DualTreeTraversalConfig and pair_policy -- backed by hand-written HIP
kernels in libyggdrax_amd.so (C ABI in include/yggdrax_amd.h). Arrays are
torch tensors on the HIP device.
"""

from .dtypes import INDEX_DTYPE, as_index
from .geometry import (LevelMajorTreeGeometry, TreeGeometry, compute_tree_geometry,
                       geometry_to_level_major)
from .interactions import (
    CompactTaggedFarPairs,
    DualTreeRetryEvent,
    DualTreeTraversalConfig,
    DualTreeWalkResult,
    NodeInteractionList,
    NodeNeighborList,
    build_compact_far_pairs,
    build_compact_far_pairs_and_leaf_neighbor_lists,
    build_interactions_and_neighbors,
    build_leaf_neighbor_lists,
    interactions_for_node,
    neighbors_for_leaf,
    set_default_dual_tree_config,
)
from .policy import PairPolicyDescriptor
from .tree import (
    OctreeTopology,
    OctreeTree,
    RadixTree,
    RadixTreeTopology,
    Tree,
    TreeBuildConfig,
    available_tree_types,
    build_octree,
    build_tree,
    infer_bounds,
    morton_encode,
    register_tree_builder,
    resolve_tree_topology,
    tree_from_particles,
)

__all__ = [
    "INDEX_DTYPE", "as_index", "TreeGeometry", "compute_tree_geometry", "DualTreeRetryEvent",
    "LevelMajorTreeGeometry", "geometry_to_level_major",
    "DualTreeTraversalConfig", "DualTreeWalkResult", "NodeInteractionList", "NodeNeighborList",
    "build_interactions_and_neighbors", "interactions_for_node", "neighbors_for_leaf",
    "CompactTaggedFarPairs", "build_compact_far_pairs",
    "build_compact_far_pairs_and_leaf_neighbor_lists", "build_leaf_neighbor_lists",
    "set_default_dual_tree_config", "PairPolicyDescriptor", "OctreeTopology", "OctreeTree", "RadixTree",
    "RadixTreeTopology", "Tree", "TreeBuildConfig", "available_tree_types", "build_octree",
    "build_tree", "infer_bounds", "morton_encode", "register_tree_builder",
    "resolve_tree_topology", "tree_from_particles",
]
