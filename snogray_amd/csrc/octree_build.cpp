// Host octree builder (rules cited in octree_build.h).
#include "octree_build.h"

#include <cstring>
#include <deque>

namespace sgr {
namespace {

enum : unsigned { XH = 4, YH = 2, ZH = 1 };

struct Work {
  uint32_t node;
  float x, y, z, sz;
};

#define SGR_NO_LINK 0xffffffffu

// A node's surfaces are a head-linked list through Builder::links, newest
// first, as the reference keeps them (space/octree-builder.cc:128-154).
struct BNode {
  uint32_t kid[8];
  uint32_t head, count;
  BNode() : head(SGR_NO_LINK), count(0) { std::memset(kid, 0, sizeof kid); }
};

struct Link {
  uint32_t id, next;
};

class Builder {
 public:
  std::vector<BNode> nodes;
  std::vector<Link> links;
  float org[3] = {0, 0, 0};
  float size = 0;

  void insert(uint32_t id, const Box3 &b) {
    if (nodes.empty()) {  // space/octree-builder.cc:205-218
      nodes.emplace_back();
      for (int a = 0; a < 3; a++) org[a] = b.lo[a];
      float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
      float s = ex;
      if (ey > s) s = ey;
      if (ez > s) s = ez;
      size = s;
      add_item(0, id);
      return;
    }
    // Grow the root until the box fits (:180-196 and :227-317).
    while (!(org[0] <= b.lo[0] && org[1] <= b.lo[1] && org[2] <= b.lo[2] &&
             (org[0] + size) >= b.hi[0] && (org[1] + size) >= b.hi[1] &&
             (org[2] + size) >= b.hi[2]))
      grow(b);
    descend(id, b);
  }

 private:
  std::vector<Work> work;  // descent work list, reused across inserts

  void add_item(uint32_t node, uint32_t id) {
    links.push_back(Link{id, nodes[node].head});
    nodes[node].head = (uint32_t)links.size() - 1;
    nodes[node].count++;
  }

  void grow(const Box3 &b) {
    uint32_t moved = (uint32_t)nodes.size();
    nodes.push_back(std::move(nodes[0]));
    nodes[0] = BNode();
    unsigned slot = 0;
    bool lo_side[3];
    for (int a = 0; a < 3; a++) {
      float lo_grow = org[a] - b.lo[a];
      float hi_grow = b.hi[a] - (org[a] + size);
      // Growing toward +axis keeps the old root in the low half.
      lo_side[a] = hi_grow <= lo_grow;
      if (lo_side[a]) slot |= (a == 0 ? XH : a == 1 ? YH : ZH);
    }
    nodes[0].kid[slot] = moved;
    for (int a = 0; a < 3; a++)
      if (lo_side[a]) org[a] -= size;
    size *= 2;
  }

  // Eager descent (:323-525), iterative over a work list.
  void descend(uint32_t id, const Box3 &b) {
    float avg = ((b.hi[0] - b.lo[0]) + (b.hi[1] - b.lo[1]) + (b.hi[2] - b.lo[2])) / 3;
    work.clear();
    work.push_back(Work{0, org[0], org[1], org[2], size});
    while (!work.empty()) {
      Work w = work.back();
      work.pop_back();
      float sub = w.sz / 2;
      float mid[3] = {w.x + sub, w.y + sub, w.z + sub};
      bool force = avg < w.sz / 4;
      bool in_lo[3], in_hi[3];
      for (int a = 0; a < 3; a++) {
        bool flat = b.lo[a] == b.hi[a];
        in_lo[a] = b.hi[a] < mid[a] || (b.hi[a] == mid[a] && !flat) ||
                   (force && b.lo[a] < mid[a]);
        in_hi[a] = b.lo[a] > mid[a] || (b.lo[a] == mid[a] && !flat) ||
                   (force && b.hi[a] > mid[a]);
      }
      bool placed = false;
      for (unsigned c = 0; c < 8; c++) {
        bool hx = c & XH, hy = c & YH, hz = c & ZH;
        if (!(hx ? in_hi[0] : in_lo[0]) || !(hy ? in_hi[1] : in_lo[1]) ||
            !(hz ? in_hi[2] : in_lo[2]))
          continue;
        uint32_t k = nodes[w.node].kid[c];
        if (!k) {
          k = (uint32_t)nodes.size();
          nodes.emplace_back();
          nodes[w.node].kid[c] = k;
        }
        work.push_back(Work{k, hx ? mid[0] : w.x, hy ? mid[1] : w.y,
                            hz ? mid[2] : w.z, sub});
        placed = true;
      }
      if (!placed) add_item(w.node, id);
    }
  }
};

// Octree::upd_stats (space/octree.cc:316-367): max depth and the
// child-averaged depth, accumulated in the same float order.
void subtree_stats(const FlatOctree &t, uint32_t ni, unsigned &max_depth,
                   float &avg) {
  const uint32_t *n = &t.nodes[4 * ni];
  unsigned sib_max = max_depth;
  max_depth = 0;
  float sib_avg = avg;
  avg = 0;
  unsigned nsub = 0, r = 0;
  for (unsigned c = 0; c < 8; c++)
    if (n[1] & (1u << c)) {
      nsub++;
      subtree_stats(t, n[0] + r, max_depth, avg);
      r++;
    }
  if (max_depth + 1 > sib_max) max_depth++;
  else max_depth = sib_max;
  if (nsub != 0) avg /= nsub;
  avg += 1 + sib_avg;
}

}  // namespace

void build_octree(const std::vector<Box3> &boxes,
                  const std::vector<uint32_t> &ids, FlatOctree &out) {
  Builder bld;
  bld.nodes.reserve(ids.size() + 1);
  bld.links.reserve(2 * ids.size() + 16);
  for (uint32_t id : ids) bld.insert(id, boxes[id]);
  out = FlatOctree();
  out.num_real = ids.size();
  for (int a = 0; a < 3; a++) out.origin[a] = bld.org[a];
  out.size = bld.size;
  if (bld.nodes.empty()) return;
  // Breadth-first copy (:538-650).
  std::deque<uint32_t> q;
  q.push_back(0);
  uint32_t next_free = 1;
  while (!q.empty()) {
    uint32_t from = q.front();
    q.pop_front();
    const BNode &bn = bld.nodes[from];
    uint32_t rec[4] = {0, 0, (uint32_t)out.surf_ids.size(), bn.count};
    for (uint32_t l = bn.head; l != SGR_NO_LINK; l = bld.links[l].next)
      out.surf_ids.push_back(bld.links[l].id);
    for (unsigned c = 0; c < 8; c++)
      if (bn.kid[c]) {
        if (!rec[1]) rec[0] = next_free;
        rec[1] |= 1u << c;
        next_free++;
        q.push_back(bn.kid[c]);
      }
    if (!rec[1]) out.num_leaf++;
    out.num_surfaces += rec[3];
    out.nodes.insert(out.nodes.end(), rec, rec + 4);
  }
  unsigned md = 0;
  float avg = 0;
  subtree_stats(out, 0, md, avg);
  out.max_depth = md;
  out.avg_depth = avg;
}

}  // namespace sgr
