// Host octree builder (rules cited in octree_build.h).
#include "octree_build.h"

#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <thread>

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
// kid[c] == 0: no child c (the root, id 0, is nobody's child).
struct BNode {
  uint32_t kid[8];
  uint32_t head, count;
};

struct Link {
  uint32_t id, next;
};

// Node storage: one pool per builder thread, in blocks that never move,
// so a node id (pool << 26 | index) stays valid while other threads add
// nodes to their own pools.
constexpr uint32_t kPoolShift = 26, kBlockShift = 16;
constexpr uint32_t kBlock = 1u << kBlockShift, kMaxBlocks = 1u << (kPoolShift - kBlockShift);
constexpr unsigned kMaxThreads = 16;

struct Pool {
  std::vector<std::unique_ptr<BNode[]>> blocks = std::vector<std::unique_ptr<BNode[]>>(kMaxBlocks);
  uint32_t used = 0;
};

struct Placement {
  uint32_t node, id;
};

class Builder {
 public:
  std::vector<Link> links;
  float org[3] = {0, 0, 0};
  float size = 0;
  bool empty = true;
  Pool pools[kMaxThreads];

  BNode &node(uint32_t id) {
    Pool &p = pools[id >> kPoolShift];
    const uint32_t i = id & ((1u << kPoolShift) - 1);
    return p.blocks[i >> kBlockShift][i & (kBlock - 1)];
  }

  // Inserts ids in order: the root grows sequentially, and each run of
  // surfaces that fit the current root descends in parallel.  Within a
  // run the descents are independent: where a surface goes depends only
  // on its box and the cell geometry, never on which children already
  // exist, so the tree is the one sequential insertion builds; the
  // surfaces then join their nodes' lists in insertion order.
  void build(const std::vector<Box3> &boxes, const std::vector<uint32_t> &ids) {
    const size_t n = ids.size();
    if (!n) return;
    root_from(ids[0], boxes[ids[0]]);
    unsigned hw = std::thread::hardware_concurrency();
    const unsigned nt = std::max(1u, std::min(hw ? hw : 1u, kMaxThreads));
    std::vector<std::vector<Placement>> placed(nt);
    for (size_t i = 1; i < n;) {
      const Box3 &b = boxes[ids[i]];
      while (!fits(b)) grow(b);  // space/octree-builder.cc:180-196 and :227-317
      size_t j = i + 1;
      while (j < n && fits(boxes[ids[j]])) j++;
      const size_t len = j - i;
      const unsigned use = len < 4096 ? 1u : nt;
      std::vector<std::exception_ptr> err(use);
      auto run = [&](unsigned t) {
        try {  // an exception must not leave a thread: carried to the caller
          std::vector<Work> work;
          uint32_t spare = SGR_NO_LINK;
          placed[t].clear();
          for (size_t k = i + len * t / use, e = i + len * (t + 1) / use; k < e; k++)
            descend(t, ids[k], boxes[ids[k]], work, spare, placed[t]);
        } catch (...) {
          err[t] = std::current_exception();
        }
      };
      if (use == 1) {
        run(0);
      } else {
        // A thread that cannot be started leaves its range to this one;
        // the threads already running are joined before anything here
        // can leave the scope their ranges refer to.
        std::vector<std::thread> th;
        unsigned started = 1;
        try {
          th.reserve(use - 1);
          for (; started < use; started++) th.emplace_back(run, started);
        } catch (...) {
        }
        run(0);
        for (unsigned t = started; t < use; t++) run(t);
        for (auto &x : th) x.join();
      }
      for (auto &e : err)
        if (e) std::rethrow_exception(e);
      for (unsigned t = 0; t < use; t++)
        for (const Placement &pl : placed[t]) add_item(pl.node, pl.id);
      i = j;
    }
  }

 private:
  uint32_t alloc(unsigned t) {
    Pool &p = pools[t];
    if (p.used >= (1u << kPoolShift)) throw std::length_error("octree: too many nodes");
    const uint32_t i = p.used++;
    auto &blk = p.blocks[i >> kBlockShift];
    if (!blk) blk.reset(new BNode[kBlock]);
    BNode &nd = blk[i & (kBlock - 1)];
    std::memset(nd.kid, 0, sizeof nd.kid);
    nd.head = SGR_NO_LINK;
    nd.count = 0;
    return (uint32_t)t << kPoolShift | i;
  }

  bool fits(const Box3 &b) const {
    return org[0] <= b.lo[0] && org[1] <= b.lo[1] && org[2] <= b.lo[2] &&
           (org[0] + size) >= b.hi[0] && (org[1] + size) >= b.hi[1] &&
           (org[2] + size) >= b.hi[2];
  }

  void add_item(uint32_t nd, uint32_t id) {
    links.push_back(Link{id, node(nd).head});
    node(nd).head = (uint32_t)links.size() - 1;
    node(nd).count++;
  }

  void root_from(uint32_t id, const Box3 &b) {  // space/octree-builder.cc:205-218
    alloc(0);
    for (int a = 0; a < 3; a++) org[a] = b.lo[a];
    float ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
    float s = ex;
    if (ey > s) s = ey;
    if (ez > s) s = ez;
    size = s;
    empty = false;
    add_item(0, id);
  }

  void grow(const Box3 &b) {
    const uint32_t moved = alloc(0);
    node(moved) = node(0);
    BNode &root = node(0);
    std::memset(root.kid, 0, sizeof root.kid);
    root.head = SGR_NO_LINK;
    root.count = 0;
    unsigned slot = 0;
    bool lo_side[3];
    for (int a = 0; a < 3; a++) {
      float lo_grow = org[a] - b.lo[a];
      float hi_grow = b.hi[a] - (org[a] + size);
      // Growing toward +axis keeps the old root in the low half.
      lo_side[a] = hi_grow <= lo_grow;
      if (lo_side[a]) slot |= (a == 0 ? XH : a == 1 ? YH : ZH);
    }
    root.kid[slot] = moved;
    for (int a = 0; a < 3; a++)
      if (lo_side[a]) org[a] -= size;
    size *= 2;
  }

  // Eager descent (:323-525), iterative over a work list; a missing child
  // is created by whichever thread gets there first.
  void descend(unsigned t, uint32_t id, const Box3 &b, std::vector<Work> &work,
               uint32_t &spare, std::vector<Placement> &out) {
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
        uint32_t *slot = &node(w.node).kid[c];
        uint32_t k = __atomic_load_n(slot, __ATOMIC_ACQUIRE);
        if (!k) {
          if (spare == SGR_NO_LINK) spare = alloc(t);
          uint32_t expect = 0;
          if (__atomic_compare_exchange_n(slot, &expect, spare, false, __ATOMIC_ACQ_REL,
                                          __ATOMIC_ACQUIRE)) {
            k = spare;
            spare = SGR_NO_LINK;
          } else {
            k = expect;  // another thread made it; keep ours for later
          }
        }
        work.push_back(Work{k, hx ? mid[0] : w.x, hy ? mid[1] : w.y,
                            hz ? mid[2] : w.z, sub});
        placed = true;
      }
      if (!placed) out.push_back(Placement{w.node, id});
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
  std::unique_ptr<Builder> bld(new Builder);
  bld->links.reserve(2 * ids.size() + 16);
  bld->build(boxes, ids);
  out = FlatOctree();
  out.num_real = ids.size();
  for (int a = 0; a < 3; a++) out.origin[a] = bld->org[a];
  out.size = bld->size;
  if (bld->empty) return;
  // Breadth-first copy (:538-650); ORDER is the queue, kept whole.
  size_t made = 0;
  for (const Pool &p : bld->pools) made += p.used;
  out.nodes.reserve(4 * made);
  out.surf_ids.reserve(bld->links.size());
  std::vector<uint32_t> order;
  order.reserve(made);
  order.push_back(0);
  uint32_t next_free = 1;
  for (size_t qi = 0; qi < order.size(); qi++) {
    const BNode &bn = bld->node(order[qi]);
    uint32_t rec[4] = {0, 0, (uint32_t)out.surf_ids.size(), bn.count};
    for (uint32_t l = bn.head; l != SGR_NO_LINK; l = bld->links[l].next)
      out.surf_ids.push_back(bld->links[l].id);
    for (unsigned c = 0; c < 8; c++)
      if (bn.kid[c]) {
        if (!rec[1]) rec[0] = next_free;
        rec[1] |= 1u << c;
        next_free++;
        order.push_back(bn.kid[c]);
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
