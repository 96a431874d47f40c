"""CPU oracle for the yggdrax tree pipeline (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker.
"""

from .oracle import (  # noqa: F401
    build_interactions_and_neighbors,
    build_octree,
    build_tree,
    compute_tree_geometry,
    compute_tree_mass_moments,
    densify_interactions,
    geometry_to_level_major,
    grouped_interactions_from_pairs,
    infer_bounds,
    lib,
    mac_extents,
    morton_encode,
    to_int32_like_numpy_cast,
    walk,
)
