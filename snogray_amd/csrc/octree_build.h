// Octree construction on the host, flattened for the device.
//
// Same insertion rules as the reference builder (space/octree-builder.cc):
// the root starts as the first surface's bounding cube and doubles toward
// any surface that does not fit (:227-317); a surface descends eagerly
// into the child that holds it, with exact-midpoint tie rules, and is
// forced into every overlapped child when its average extent is below a
// quarter of the node size (:323-500); each node's surface list is kept
// newest-first (:128-154).  Between root growths, the surfaces descend
// on up to 16 host threads (the tree and list order are those of
// sequential insertion; see Builder::build).  The result is flattened
// breadth-first so
// that the children of a node are consecutive (:538-650), which lets a
// device node be 16 bytes: first child, 8-bit child mask, first surface
// and surface count.
#pragma once
#include <cstdint>
#include <vector>

namespace sgr {

struct Box3 {
  float lo[3], hi[3];
};

struct FlatOctree {
  // 4 x uint32 per node: first_child, child_mask, surf_begin, surf_count
  std::vector<uint32_t> nodes;
  std::vector<uint32_t> surf_ids;
  float origin[3] = {0, 0, 0};
  float size = 0;
  uint32_t max_depth = 0;
  uint64_t num_real = 0;
  // Octree::stats (space/octree.cc:308-367)
  uint64_t num_leaf = 0, num_surfaces = 0;
  float avg_depth = 0;
};

// Build from per-surface boxes; `ids` gives the insertion order (only
// the listed surfaces are inserted).
void build_octree(const std::vector<Box3> &boxes,
                  const std::vector<uint32_t> &ids, FlatOctree &out);

}  // namespace sgr
