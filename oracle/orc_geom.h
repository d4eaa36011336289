// ORACLE -- test infrastructure only.  Nothing in the product links this.
//
// CPU restatement of the reference's geometry and octree code, templated on
// the coordinate type R (float reproduces the reference's fp32 build,
// double gives the FP64 oracle).  Every function cites the reference
// file:line it restates.  Compiled with -O2 -ffp-contract=off and no
// fast-math so each float operation is one IEEE operation in the written
// order.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <vector>
#include <algorithm>

namespace orc {

// util/snogmath.h:43-51 -- ternary min/max (NaN behaviour matters).
template<class T> static inline T mn(T a, T b) { return a < b ? a : b; }
template<class T> static inline T mx(T a, T b) { return a > b ? a : b; }
// util/snogmath.h:57-70
template<class T> static inline T clampv(T v, T lo, T hi) {
  v = v < hi ? v : hi;
  return v > lo ? v : lo;
}
template<class T> static inline T clamp01(T v) { return clampv(v, T(0), T(1)); }

static const double PI_D = 3.14159265358979323846;
static const float PIf = 3.14159265358979323846f;
static const float INV_PIf = 0.31830988618379067154f;

// geometry/vec.h, geometry/pos.h: one 3-tuple type serves as Vec and Pos.
template<class R> struct V3 {
  R x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(R a, R b, R c) : x(a), y(b), z(c) {}
  V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
  V3 &operator+=(const V3 &o) { x = x + o.x; y = y + o.y; z = z + o.z; return *this; }
  V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
  V3 operator-() const { return V3(-x, -y, -z); }
  V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
  // geometry/vec.h:59-62: division is multiplication by the reciprocal.
  V3 operator/(R d) const { return operator*(R(1) / d); }
  R length_squared() const { return x * x + y * y + z * z; }
  R length() const { return std::sqrt(x * x + y * y + z * z); }
  // geometry/vec.h:113-117
  V3 unit() const { R len = length(); return operator*(R(1) / len); }
  // geometry/vec.h:119-123
  V3 perpendicular() const {
    R nz = (x == z) ? z : -z;
    return V3(y * y + x * nz, -nz * nz - x * y, x * x - y * nz);
  }
  bool null() const { return x == 0 && y == 0 && z == 0; }
};

// geometry/vec.h:188-201
template<class R> static inline R dot(const V3<R> &a, const V3<R> &b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template<class R> static inline V3<R> cross(const V3<R> &a, const V3<R> &b) {
  return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x);
}
// geometry/vec.h:161-164: mirror about NORMAL.
template<class R> static inline V3<R> mirror(const V3<R> &v, const V3<R> &n) {
  return n * dot(v, n) * R(2) - v;
}
// geometry/vec.h:166-186: refraction; returns 0 on total internal reflection.
template<class R>
static inline V3<R> refraction(const V3<R> &v, const V3<R> &n, R ior_in,
                               R ior_out) {
  R ior_ratio = ior_in / ior_out;
  R c1 = -dot(v, n);
  R c2_sq = R(1) - (ior_ratio * ior_ratio) * (R(1) - c1 * c1);
  if (c2_sq < -R(1e-15)) return V3<R>(0, 0, 0);
  if (c2_sq < 0) c2_sq = 0;
  return (v * ior_ratio) + (n * (ior_ratio * c1 - std::sqrt(c2_sq)));
}
template<class R> static inline V3<R> vmin(const V3<R> &a, const V3<R> &b) {
  return V3<R>(mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z));
}
template<class R> static inline V3<R> vmax(const V3<R> &a, const V3<R> &b) {
  return V3<R>(mx(a.x, b.x), mx(a.y, b.y), mx(a.z, b.z));
}

// geometry/frame.h:32-136
template<class R> struct Frame {
  V3<R> origin, x, y, z;
  Frame() : origin(), x(1, 0, 0), y(0, 1, 0), z(0, 0, 1) {}
  Frame(const V3<R> &o, const V3<R> &a, const V3<R> &b, const V3<R> &c)
      : origin(o), x(a), y(b), z(c) {}
  // geometry/vec.h:125-137 to_basis / from_basis
  V3<R> to(const V3<R> &v) const {
    return V3<R>(dot(v, x), dot(v, y), dot(v, z));
  }
  V3<R> from(const V3<R> &v) const {
    return V3<R>(x.x * v.x + y.x * v.y + z.x * v.z,
                 x.y * v.x + y.y * v.y + z.y * v.z,
                 x.z * v.x + y.z * v.y + z.z * v.z);
  }
};

// geometry/bbox.h:27-174
template<class R> struct BBox {
  V3<R> min, max;
  BBox()
      : min(R(3.4028234663852886e38), R(3.4028234663852886e38),
            R(3.4028234663852886e38)),
        max(-R(3.4028234663852886e38), -R(3.4028234663852886e38),
            -R(3.4028234663852886e38)) {}
  explicit BBox(const V3<R> &p) : min(p), max(p) {}
  void add(const V3<R> &p) { min = vmin(min, p); max = vmax(max, p); }
  void add(const BBox &b) { min = vmin(min, b.min); max = vmax(max, b.max); }
  V3<R> extent() const { return max - min; }
  R diameter() const { return extent().length(); }
  // geometry/bbox.h:102-111
  R max_size() const {
    V3<R> e = extent();
    R sz = e.x;
    if (e.y > sz) sz = e.y;
    if (e.z > sz) sz = e.z;
    return sz;
  }
  // geometry/bbox.h:126-130
  R avg_size() const {
    V3<R> e = extent();
    return (e.x + e.y + e.z) / R(3);
  }
  // geometry/bbox.h:84-87: center = min + extent/2
  V3<R> center() const { return min + extent() / R(2); }
  R radius() const { return diameter() / R(2); }
};

// geometry/tripar-isec.h:56-118 (Moller-Trumbore).  Returns hit, t, u, v.
// Eps is coord_t(1e-15) (geometry/coords.h:45).
template<class R>
static inline bool tripar_intersects(const V3<R> &corner, const V3<R> &e1,
                                     const V3<R> &e2, bool parallelogram,
                                     const V3<R> &ro, const V3<R> &rd,
                                     R min_t, R &t, R &u, R &v) {
  V3<R> pvec = cross(rd, e2);
  R det = dot(e1, pvec);
  const R Eps = R(1e-15);
  if (det > -Eps && det < Eps) return false;
  R inv_det = R(1.0) / det;
  V3<R> tvec = ro - corner;
  u = dot(tvec, pvec) * inv_det;
  if (u < 0 || u > 1) return false;
  V3<R> qvec = cross(tvec, e1);
  v = dot(rd, qvec) * inv_det;
  if (parallelogram) {
    if (v < 0 || v > 1) return false;
  } else {
    if (v < 0 || u + v > 1) return false;
  }
  t = dot(e2, qvec) * inv_det;
  return t > min_t;
}

// geometry/quadratic-roots.h:30-70
template<class R>
static inline unsigned quadratic_roots(R a, R b, R c, R roots[2]) {
  if (a == 0) {
    if (b == 0) return 0;
    roots[0] = -c / b;
    return 1;
  }
  R disc = b * b - R(4) * a * c;
  if (disc < 0) return 0;
  R sqrt_disc = std::sqrt(disc);
  R t = -(b + (b < 0 ? -sqrt_disc : sqrt_disc)) / R(2);
  roots[0] = t / a;
  if (t == 0) return 1;
  roots[1] = c / t;
  if (roots[0] > roots[1]) std::swap(roots[0], roots[1]);
  return 2;
}

// geometry/sphere-isec.h:35-60
template<class R>
static inline bool sphere_intersects(const V3<R> &center, R radius,
                                     const V3<R> &ro, const V3<R> &rd, R min_t,
                                     R &t) {
  V3<R> off = ro - center;
  R a = rd.length_squared();
  R b = R(2) * dot(rd, off);
  R c = off.length_squared() - radius * radius;
  R roots[2];
  unsigned n = quadratic_roots(a, b, c, roots);
  for (unsigned i = 0; i < n; i++) {
    t = roots[i];
    if (t > min_t) return true;
  }
  return false;
}

// Ellipse::intersects (surface/ellipse.h:95-110): the bounding
// parallelogram's hit, kept when (u, v) lies in the inscribed ellipse.
template<class R>
static inline bool ellipse_intersects(const V3<R> &corner, const V3<R> &e1, const V3<R> &e2,
                                      const V3<R> &ro, const V3<R> &rd, R min_t, R &t, R &u,
                                      R &v) {
  if (!tripar_intersects(corner, e1, e2, true, ro, rd, min_t, t, u, v)) return false;
  const R x = u - R(0.5), y = v - R(0.5);
  return x * x + y * y <= R(0.25);
}

// cylinder_intersects (surface/cylinder.cc:114-150): the canonical
// cylinder (radius 1 about the z axis, z in [-1, 1], no caps); the first
// root beyond MIN_T whose z is inside.  The constant term subtracts the
// radius itself, not its square (both 1).
template<class R>
static inline bool cylinder_intersects(const V3<R> &ro, const V3<R> &rd, R min_t, R &t) {
  const R a = rd.x * rd.x + rd.y * rd.y;
  const R b = R(2) * (rd.x * ro.x + rd.y * ro.y);
  const R c = ro.x * ro.x + ro.y * ro.y - R(1);
  R roots[2];
  const unsigned n = quadratic_roots(a, b, c, roots);
  for (unsigned i = 0; i < n; i++) {
    t = roots[i];
    if (t > min_t) {
      const R z = ro.z + t * rd.z;
      if (z >= R(-1) && z <= R(1)) return true;
    }
  }
  return false;
}

template<class R> struct Ray {
  V3<R> origin, dir;
  R t0, t1;
};

// ---------------------------------------------------------------------------
// Octree (space/octree-builder.cc, space/octree.cc, space/octree-node.h)
// ---------------------------------------------------------------------------

struct BuildNode {  // space/octree-node.h:27-86 (builder form)
  unsigned child[8];
  unsigned head;
  BuildNode() : head(0) { for (int i = 0; i < 8; i++) child[i] = 0; }
};

// Flattened node: children are consecutive after BFS copy
// (space/octree-builder.cc:630-643); surf list is [surf_begin, +count).
struct FlatNode {
  unsigned first_child;
  unsigned mask;
  unsigned surf_begin;
  unsigned surf_count;
};

template<class R> struct OctreeBuilder {
  // space/octree-builder.cc:29-160
  std::vector<BuildNode> nodes;
  struct ListNode { unsigned surf, next; };
  std::vector<ListNode> lists;  // index 0 is the dummy "end of list"
  V3<R> origin;
  R size = 0;
  unsigned long num_real = 0;
  enum { X_LO = 0, X_HI = 4, Y_LO = 0, Y_HI = 2, Z_LO = 0, Z_HI = 1 };

  OctreeBuilder() { lists.push_back(ListNode{0, 0}); }

  void push_surf(unsigned s, unsigned &head) {  // :128-134
    unsigned nh = (unsigned)lists.size();
    lists.push_back(ListNode{s, head});
    head = nh;
  }

  void add(unsigned s, const BBox<R> &bb) {  // :47-51 + :174-219
    num_real++;
    add_top(s, bb);
  }

  void add_top(unsigned s, const BBox<R> &bb) {  // :174-219
    if (!nodes.empty()) {
      if (origin.x <= bb.min.x && origin.y <= bb.min.y &&
          origin.z <= bb.min.z && (origin.x + size) >= bb.max.x &&
          (origin.y + size) >= bb.max.y && (origin.z + size) >= bb.max.z)
        add_node(s, bb, 0, origin.x, origin.y, origin.z, size);
      else
        grow_to_include(s, bb);
    } else {
      nodes.push_back(BuildNode());
      origin = bb.min;
      size = bb.max_size();
      push_surf(s, nodes[0].head);
    }
  }

  void grow_to_include(unsigned s, const BBox<R> &bb) {  // :227-317
    unsigned old_root = (unsigned)nodes.size();
    nodes.push_back(nodes[0]);
    nodes[0] = BuildNode();
    R x_lo = origin.x - bb.min.x, x_hi = bb.max.x - (origin.x + size);
    R y_lo = origin.y - bb.min.y, y_hi = bb.max.y - (origin.y + size);
    R z_lo = origin.z - bb.min.z, z_hi = bb.max.z - (origin.z + size);
    unsigned slot = ((x_hi > x_lo) ? X_LO : X_HI) |
                    ((y_hi > y_lo) ? Y_LO : Y_HI) |
                    ((z_hi > z_lo) ? Z_LO : Z_HI);
    nodes[0].child[slot] = old_root;
    if (x_hi <= x_lo) origin.x -= size;
    if (y_hi <= y_lo) origin.y -= size;
    if (z_hi <= z_lo) origin.z -= size;
    size *= 2;
    add_top(s, bb);
  }

  // :323-500 -- eager split with exact-midpoint tie rules and forcing.
  void add_node(unsigned s, const BBox<R> &bb, unsigned ni, R x, R y, R z,
                R sz) {
    R sub = sz / 2;
    R mid_x = x + sub, mid_y = y + sub, mid_z = z + sub;
    bool add_here = true;
    bool force = bb.avg_size() < sz / 4;
    bool xlo = bb.max.x < mid_x || (bb.max.x == mid_x && bb.min.x != bb.max.x) ||
               (force && bb.min.x < mid_x);
    bool xhi = bb.min.x > mid_x || (bb.min.x == mid_x && bb.min.x != bb.max.x) ||
               (force && bb.max.x > mid_x);
    bool ylo = bb.max.y < mid_y || (bb.max.y == mid_y && bb.min.y != bb.max.y) ||
               (force && bb.min.y < mid_y);
    bool yhi = bb.min.y > mid_y || (bb.min.y == mid_y && bb.min.y != bb.max.y) ||
               (force && bb.max.y > mid_y);
    bool zlo = bb.max.z < mid_z || (bb.max.z == mid_z && bb.min.z != bb.max.z) ||
               (force && bb.min.z < mid_z);
    bool zhi = bb.min.z > mid_z || (bb.min.z == mid_z && bb.min.z != bb.max.z) ||
               (force && bb.max.z > mid_z);
    // The reference tests the conditions nested x -> y -> z, LO before HI,
    // and calls add_to_child in that order; the flags above are the same
    // predicates (they only read bb and the midpoints).
    if (xlo) {
      if (ylo) {
        if (zlo) { to_child(s, bb, ni, X_LO | Y_LO | Z_LO, x, y, z, sub); add_here = false; }
        if (zhi) { to_child(s, bb, ni, X_LO | Y_LO | Z_HI, x, y, mid_z, sub); add_here = false; }
      }
      if (yhi) {
        if (zlo) { to_child(s, bb, ni, X_LO | Y_HI | Z_LO, x, mid_y, z, sub); add_here = false; }
        if (zhi) { to_child(s, bb, ni, X_LO | Y_HI | Z_HI, x, mid_y, mid_z, sub); add_here = false; }
      }
    }
    if (xhi) {
      if (ylo) {
        if (zlo) { to_child(s, bb, ni, X_HI | Y_LO | Z_LO, mid_x, y, z, sub); add_here = false; }
        if (zhi) { to_child(s, bb, ni, X_HI | Y_LO | Z_HI, mid_x, y, mid_z, sub); add_here = false; }
      }
      if (yhi) {
        if (zlo) { to_child(s, bb, ni, X_HI | Y_HI | Z_LO, mid_x, mid_y, z, sub); add_here = false; }
        if (zhi) { to_child(s, bb, ni, X_HI | Y_HI | Z_HI, mid_x, mid_y, mid_z, sub); add_here = false; }
      }
    }
    if (add_here) push_surf(s, nodes[ni].head);
  }

  void to_child(unsigned s, const BBox<R> &bb, unsigned ni, unsigned cn, R x,
                R y, R z, R sz) {  // :510-525
    unsigned ci = nodes[ni].child[cn];
    if (!ci) {
      ci = (unsigned)nodes.size();
      nodes.push_back(BuildNode());
      nodes[ni].child[cn] = ci;
    }
    add_node(s, bb, ci, x, y, z, sz);
  }

  // :538-650 breadth-first copy; surface lists unrolled head-first
  // (i.e. reverse insertion order).
  void flatten(std::vector<FlatNode> &out, std::vector<unsigned> &surfs) const {
    out.clear();
    surfs.clear();
    std::deque<unsigned> q;
    unsigned next_free = 1;
    if (!nodes.empty()) q.push_back(0);
    while (!q.empty()) {
      unsigned from = q.front();
      q.pop_front();
      const BuildNode &fn = nodes[from];
      FlatNode tn;
      tn.first_child = 0;
      tn.mask = 0;
      tn.surf_begin = (unsigned)surfs.size();
      tn.surf_count = 0;
      for (unsigned li = fn.head; li; li = lists[li].next) {
        surfs.push_back(lists[li].surf);
        tn.surf_count++;
      }
      for (unsigned i = 0; i < 8; i++)
        if (fn.child[i]) {
          if (!tn.mask) tn.first_child = next_free;
          tn.mask |= 1u << i;
          q.push_back(fn.child[i]);
          next_free++;
        }
      out.push_back(tn);
    }
  }
};

// Octree::stats (space/octree.cc:308-367) on the flattened form.
struct OctStats {
  unsigned long num_nodes = 0, num_leaf_nodes = 0, num_surfaces = 0,
                num_dup_surfaces = 0;
  unsigned max_depth = 0;
  float avg_depth = 0;
};

static inline void oct_upd_stats(const std::vector<FlatNode> &nodes,
                                 unsigned ni, OctStats &st) {
  const FlatNode &n = nodes[ni];
  unsigned num_sub = 0;
  unsigned sib_max = st.max_depth;
  st.max_depth = 0;
  float sib_avg = st.avg_depth;
  st.avg_depth = 0;
  unsigned k = 0;
  for (unsigned i = 0; i < 8; i++)
    if (n.mask & (1u << i)) {
      num_sub++;
      oct_upd_stats(nodes, n.first_child + k, st);
      k++;
    }
  st.num_nodes++;
  if (num_sub == 0) st.num_leaf_nodes++;
  st.num_surfaces += n.surf_count;
  if (st.max_depth + 1 > sib_max) st.max_depth++;
  else st.max_depth = sib_max;
  if (num_sub != 0) st.avg_depth /= num_sub;
  st.avg_depth += 1 + sib_avg;
}

// Primitive storage for intersection.  Kinds: 0 = mesh triangle,
// 1 = tripar, 2 = sphere, 3 = instance, 4 = ellipse (its bounding
// parallelogram in v0, e1, e2), 5 = cylinder, 6 = sphere2 (local
// primitives: a canonical shape placed by l2w, w2l its inverse).
template<class R> struct Prim {
  int kind;
  bool parallelogram;
  V3<R> v0, e1, e2;  // triangle / tripar / ellipse
  V3<R> center;      // sphere
  R radius;
  R l2w[12], w2l[12];  // cylinder / sphere2
};

#ifdef ORC_MAILBOX
// Negative mailbox of one search (space/isec-cache.h:37-114, used at
// space/octree.cc:195-216): a surface that was tested and missed is not
// tested again in the same search, which cannot change the result (the
// ray extent only shrinks).  1024 direct-mapped entries keyed by the
// primitive index, cleared per search by a generation count.  Built only
// into the shipped-flags CPU baseline, whose surface-test counts then
// count the tests made, as the reference's do.
struct Mailbox {
  unsigned gen = 1;
  struct Entry { unsigned gen, id; } e[1024] = {};
};
static thread_local Mailbox tl_mailbox;
static inline void mailbox_new_generation() {
  if (++tl_mailbox.gen == 0) {  // generation wrapped: clear the table
    for (auto &e : tl_mailbox.e) e.gen = 0;
    tl_mailbox.gen = 1;
  }
}
#endif

// An affine 3 x 4 row-major transform applied to a position or a vector
// (geometry/pos.h:80-97, geometry/vec.h:81-95), and transposed to a
// vector (geometry/xform.h:295-309).
template<class R> static inline V3<R> xform_pos(const R *m, const V3<R> &p) {
  return V3<R>(p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3],
               p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7],
               p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11]);
}
template<class R> static inline V3<R> xform_vec(const R *m, const V3<R> &v) {
  return V3<R>(v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
               v.x * m[8] + v.y * m[9] + v.z * m[10]);
}
template<class R> static inline V3<R> xform_transpose(const R *m, const V3<R> &v) {
  return V3<R>(v.x * m[0] + v.y * m[4] + v.z * m[8], v.x * m[1] + v.y * m[5] + v.z * m[9],
               v.x * m[2] + v.y * m[6] + v.z * m[10]);
}

// Closest-hit / any-hit search over either the octree or a brute-force
// list.  The callback semantics follow space/space.cc:36-194.
template<class R> struct Accel {
  const std::vector<Prim<R>> *prims = nullptr;
  bool use_triv = false;
  // Two-level instancing (surface/instance.cc:114-146): surface
  // n_prims + i is instance i of a model with its own Accel.
  struct Inst { unsigned model; R l2w[12], w2l[12]; };
  const std::vector<Inst> *insts = nullptr;
  std::vector<const Accel *> models;
  unsigned n_prims = 0;
  std::vector<FlatNode> nodes;
  std::vector<unsigned> surfs;
  std::vector<unsigned> triv_order;
  V3<R> origin;
  R size = 0;

  // Surface test (mesh.cc:730-822, tripar.cc:93-130, sphere.cc:82-120,
  // ellipse.h:95-110, cylinder.cc:114-186, sphere2.cc:92-110).
  bool test(unsigned p, const Ray<R> &ray, R &t, R &u, R &v) const {
    const Prim<R> &pr = (*prims)[p];
    if (pr.kind == 2) {
      u = v = 0;
      if (sphere_intersects(pr.center, pr.radius, ray.origin, ray.dir, ray.t0, t))
        return t < ray.t1;
      return false;
    }
    if (pr.kind == 5 || pr.kind == 6) {
      // LocalXform::world_to_local of the ray: origin and direction, the
      // parametric extent kept
      u = v = 0;
      const V3<R> o = xform_pos(pr.w2l, ray.origin), d = xform_vec(pr.w2l, ray.dir);
      if (pr.kind == 5) return cylinder_intersects(o, d, ray.t0, t) && t < ray.t1;
      return sphere_intersects(V3<R>(0, 0, 0), R(1), o, d, ray.t0, t) && t < ray.t1;
    }
    if (pr.kind == 4) return ellipse_intersects(pr.v0, pr.e1, pr.e2, ray.origin, ray.dir, ray.t0, t, u, v) && t < ray.t1;
    return tripar_intersects(pr.v0, pr.e1, pr.e2, pr.parallelogram, ray.origin,
                             ray.dir, ray.t0, t, u, v) &&
           t < ray.t1;
  }

  struct Search {
    const Accel *acc;
    Ray<R> *ray;
    bool any;
    // Shadow rays through partially occluding surfaces (space/space.cc:
    // 131-194, material/material.h:168-192): for a surface the ray meets,
    // PARTIAL (PCTX, p, t) multiplies the ray's transmittance by the
    // surface's and returns whether the ray is blocked (a fully occluding
    // surface, or a transmittance below Eps); a surface it lets through
    // counts as a miss, as the reference's callback reports it.
    int (*partial)(void *, unsigned, R) = nullptr;
    void *pctx = nullptr;
    bool stop = false;
    unsigned hit = 0xffffffffu;
    unsigned hit_inst = 0xffffffffu;
    R hu = 0, hv = 0;
    unsigned octant;
    uint64_t node_calls = 0, surf_tests = 0;

    bool cb(unsigned p) {  // space/space.cc:42-53 / :103-116
      R t, u, v;
#ifdef ORC_MAILBOX
      Mailbox::Entry &me = tl_mailbox.e[p & 1023u];
      if (me.gen == tl_mailbox.gen && me.id == p) return false;
#endif
      surf_tests++;
      if (acc->insts && p >= acc->n_prims) {
        // an instance: the ray in model space (t kept), the model's own
        // search, whose node calls and tests count with this one's
        const unsigned i = p - acc->n_prims;
        const Inst &I = (*acc->insts)[i];
        Ray<R> r2;
        r2.origin = xform_pos(I.w2l, ray->origin);
        r2.dir = xform_vec(I.w2l, ray->dir);
        r2.t0 = ray->t0;
        r2.t1 = ray->t1;
        Search s2;
        s2.acc = acc->models[I.model];
        s2.ray = &r2;
        s2.any = any;
        s2.acc->search(s2);
        node_calls += s2.node_calls;
        surf_tests += s2.surf_tests;
#ifdef ORC_MAILBOX
        // the model's surfaces share the table's keys: start afresh
        mailbox_new_generation();
#endif
        if (s2.hit == 0xffffffffu) return false;
        hit = s2.hit;
        hit_inst = i;
        if (any) { stop = true; return true; }
        ray->t1 = r2.t1;
        hu = s2.hu;
        hv = s2.hv;
        return true;
      }
      if (acc->test(p, *ray, t, u, v)) {
        if (any) {
          if (!partial || partial(pctx, p, t)) {
            stop = true;
            hit = p;
            hit_inst = 0xffffffffu;
            return true;
          }
        } else {
          ray->t1 = t;
          hit = p;
          hit_inst = 0xffffffffu;
          hu = u;
          hv = v;
          return true;
        }
      }
#ifdef ORC_MAILBOX
      me.gen = tl_mailbox.gen;
      me.id = p;
#endif
      return false;
    }

    // space/octree.cc:175-300
    void visit(unsigned ni, R xmn, R xmx, R ymn, R ymx, R zmn, R zmx) {
      node_calls++;
      R min_t = mx(ray->t0, mx(xmn, mx(ymn, zmn)));
      R max_t = mn(ray->t1, mn(xmx, mn(ymx, zmx)));
      if (min_t >= max_t) return;
      const FlatNode &n = acc->nodes[ni];
      for (unsigned k = 0; k < n.surf_count; k++) {
        cb(acc->surfs[n.surf_begin + k]);
        if (stop) return;
      }
      if (n.mask) {
        R xh = (xmx - xmn) / 2, yh = (ymx - ymn) / 2, zh = (zmx - zmn) / 2;
        R xm = xmn + xh, ym = ymn + yh, zm = zmn + zh;
        for (unsigned i = 0; i < 8; i++) {
          unsigned child = (i == 3) ? 4 : (i == 4) ? 3 : i;
          unsigned real = child ^ octant;
          if (!(n.mask & (1u << real))) continue;
          unsigned ci = n.first_child +
                        (unsigned)__builtin_popcount(n.mask & ((1u << real) - 1));
          R t0 = ray->t0, t1 = ray->t1;
          if (((child & 4) ? (t1 > xm) : (t0 < xm)) &&
              ((child & 2) ? (t1 > ym) : (t0 < ym)) &&
              ((child & 1) ? (t1 > zm) : (t0 < zm))) {
            R cx = xmn + ((child & 4) ? xh : R(0));
            R cy = ymn + ((child & 2) ? yh : R(0));
            R cz = zmn + ((child & 1) ? zh : R(0));
            visit(ci, cx, cx + xh, cy, cy + yh, cz, cz + zh);
            if (stop) return;
          }
        }
      }
    }
  };

  // Octree::for_each_possible_intersector (space/octree.cc:101-162) or
  // TrivSpace (space/triv-space.h:48-58).
  void search(Search &ss) const {
    Ray<R> &ray = *ss.ray;
#ifdef ORC_MAILBOX
    mailbox_new_generation();
#endif
    if (use_triv) {
      for (unsigned p : triv_order) {
        ss.cb(p);
        if (ss.stop) return;
      }
      return;
    }
    if (nodes.empty()) return;
    ss.octant = (ray.dir.x >= 0 ? 0u : 4u) | (ray.dir.y >= 0 ? 0u : 2u) |
                (ray.dir.z >= 0 ? 0u : 1u);
    R inv_x = ray.dir.x == 0 ? R(1e9) : R(1) / ray.dir.x;
    R xmn = (origin.x - ray.origin.x) * inv_x;
    R xmx = (origin.x - ray.origin.x + size) * inv_x;
    if (xmn > xmx) std::swap(xmn, xmx);
    R inv_y = ray.dir.y == 0 ? R(1e9) : R(1) / ray.dir.y;
    R ymn = (origin.y - ray.origin.y) * inv_y;
    R ymx = (origin.y - ray.origin.y + size) * inv_y;
    if (ymn > ymx) std::swap(ymn, ymx);
    R inv_z = ray.dir.z == 0 ? R(1e9) : R(1) / ray.dir.z;
    R zmn = (origin.z - ray.origin.z) * inv_z;
    R zmx = (origin.z - ray.origin.z + size) * inv_z;
    if (zmn > zmx) std::swap(zmn, zmx);
    ss.visit(0, xmn, xmx, ymn, ymx, zmn, zmx);
  }

  // Closest hit; updates ray.t1 (space/space.cc:55-66).
  unsigned closest(Ray<R> &ray, R &u, R &v, uint64_t *nc = nullptr,
                   uint64_t *st = nullptr, unsigned *inst = nullptr) const {
    Search ss;
    ss.acc = this;
    ss.ray = &ray;
    ss.any = false;
    search(ss);
    u = ss.hu;
    v = ss.hv;
    if (nc) *nc += ss.node_calls;
    if (st) *st += ss.surf_tests;
    if (inst) *inst = ss.hit_inst;
    return ss.hit;
  }

  // Any hit (space/space.cc:182-194; every material here fully occludes,
  // material/material.h:101-112).
  bool occluded(const Ray<R> &ray_in, uint64_t *nc = nullptr, uint64_t *st = nullptr,
                int (*partial)(void *, unsigned, R) = nullptr, void *pctx = nullptr) const {
    Ray<R> ray = ray_in;
    Search ss;
    ss.acc = this;
    ss.ray = &ray;
    ss.any = true;
    ss.partial = partial;
    ss.pctx = pctx;
    search(ss);
    if (nc) *nc += ss.node_calls;
    if (st) *st += ss.surf_tests;
    return ss.stop;
  }
};

}  // namespace orc
