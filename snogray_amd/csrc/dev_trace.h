// Device ray queries: Moller-Trumbore, sphere test and the octree
// traversal, all in the reference's fp32 operation order so hits are bit
// identical to the CPU path.
#pragma once
#include "dev_common.h"

namespace sgr {

#define SGR_MISS 0xffffffffu


struct RayF {
  V3 o, d;
  float t0, t1;
};

// geometry/tripar-isec.h:56-118 with the ray-extent check of :140-150.
__device__ __forceinline__ bool tripar_hit(V3 corner, V3 e1, V3 e2, bool par, const RayF &r,
                                           float &t, float &u, float &v) {
  V3 pvec = cross(r.d, e2);
  float det = dot(e1, pvec);
  if (det > -1e-15f && det < 1e-15f) return false;
  float inv_det = 1.0f / det;
  V3 tvec = r.o - corner;
  u = dot(tvec, pvec) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  V3 qvec = cross(tvec, e1);
  v = dot(r.d, qvec) * inv_det;
  if (par) {
    if (v < 0.0f || v > 1.0f) return false;
  } else {
    if (v < 0.0f || u + v > 1.0f) return false;
  }
  t = dot(e2, qvec) * inv_det;
  return t > r.t0 && t < r.t1;
}

// geometry/quadratic-roots.h:30-70: the real roots of a t^2 + b t + cc,
// ascending; returns how many (0, 1 or 2).
__device__ __forceinline__ unsigned quad_roots(float a, float b, float cc, float &r0,
                                               float &r1) {
  if (a == 0.0f) {
    if (b == 0.0f) return 0;
    r0 = -cc / b;
    r1 = r0;
    return 1;
  }
  float disc = b * b - 4.0f * a * cc;
  if (disc < 0.0f) return 0;
  float sd = sqrtf(disc);
  float q = -(b + (b < 0.0f ? -sd : sd)) / 2.0f;
  r0 = q / a;
  if (q == 0.0f) {
    r1 = r0;
    return 1;
  }
  r1 = cc / q;
  if (r0 > r1) { float tmp = r0; r0 = r1; r1 = tmp; }
  return 2;
}

// geometry/sphere-isec.h:35-60
__device__ __forceinline__ bool sphere_hit(V3 c, float radius, V3 ro, V3 rd, float min_t,
                                           float &t) {
  V3 off = ro - c;
  float a = len2(rd);
  float b = 2.0f * dot(rd, off);
  float cc = len2(off) - radius * radius;
  float r0, r1;
  const unsigned n = quad_roots(a, b, cc, r0, r1);
  if (n >= 1 && r0 > min_t) { t = r0; return true; }
  if (n == 2 && r1 > min_t) { t = r1; return true; }
  return false;
}

// cylinder_intersects (surface/cylinder.cc:114-150): the canonical
// cylinder (radius 1 about z, z in [-1, 1], no caps); the first root
// beyond MIN_T whose z is inside.  The constant term subtracts the radius
// itself, not its square (both 1).
__device__ __forceinline__ bool cylinder_hit(V3 ro, V3 rd, float min_t, float &t) {
  const float a = rd.x * rd.x + rd.y * rd.y;
  const float b = 2.0f * (rd.x * ro.x + rd.y * ro.y);
  const float cc = ro.x * ro.x + ro.y * ro.y - 1.0f;
  float r0, r1;
  const unsigned n = quad_roots(a, b, cc, r0, r1);
  for (unsigned i = 0; i < n; i++) {
    const float ti = i ? r1 : r0;
    if (ti > min_t) {
      const float z = ro.z + ti * rd.z;
      if (z >= -1.0f && z <= 1.0f) { t = ti; return true; }
    }
  }
  return false;
}

// Ellipse::intersects (surface/ellipse.h:95-110): the bounding
// parallelogram's hit, kept when (u, v) lies in the inscribed ellipse.
__device__ __forceinline__ bool ellipse_hit(V3 corner, V3 e1, V3 e2, const RayF &r, float &t,
                                            float &u, float &v) {
  if (!tripar_hit(corner, e1, e2, true, r, t, u, v)) return false;
  const float x = u - 0.5f, y = v - 0.5f;
  return x * x + y * y <= 0.25f;
}

// A local primitive's ray in its own space (LocalXform::world_to_local:
// origin and direction transformed, the parametric extent kept).
__device__ __forceinline__ void local_ray(const float4 *lr, V3 o, V3 d, V3 &lo, V3 &ld) {
  const float m[12] = {lr[0].x, lr[0].y, lr[0].z, lr[0].w, lr[1].x, lr[1].y,
                       lr[1].z, lr[1].w, lr[2].x, lr[2].y, lr[2].z, lr[2].w};
  lo = xform_pos(m, o);
  ld = xform_vec(m, d);
}

// Surface test of a local primitive (surface/cylinder.cc:152-186,
// surface/sphere2.cc:92-110) with table rows LR.
__device__ __forceinline__ bool local_hit(const float4 *lr, uint32_t kind, const RayF &r,
                                          float &t) {
  V3 lo, ld;
  local_ray(lr, r.o, r.d, lo, ld);
  const bool h = kind == SGR_REF_CYLINDER ? cylinder_hit(lo, ld, r.t0, t)
                                          : sphere_hit(v3(0.0f, 0.0f, 0.0f), 1.0f, lo, ld, r.t0, t);
  return h && t < r.t1;
}

// IsecInfo::normal of a local primitive hit at T (surface/cylinder.cc:
// 84-90, surface/sphere2.cc:79-83): the local hit point (its x, y for a
// cylinder) by the inverse transpose, unit; ONORM gets the local normal.
__device__ __forceinline__ V3 local_onorm(const float4 *lr, const RayF &r, float t) {
  V3 lo, ld;
  local_ray(lr, r.o, r.d, lo, ld);
  const V3 ip = lo + ld * t;
  return __float_as_uint(lr[6].x) == SGR_REF_CYLINDER ? v3(ip.x, ip.y, 0.0f) : ip;
}
__device__ __forceinline__ V3 w2l_transpose(const float4 *lr, V3 v) {
  return V3{v.x * lr[0].x + v.y * lr[1].x + v.z * lr[2].x,
            v.x * lr[0].y + v.y * lr[1].y + v.z * lr[2].y,
            v.x * lr[0].z + v.y * lr[1].z + v.z * lr[2].z};
}
__device__ __forceinline__ V3 local_normal(const float4 *lr, const RayF &r, float t) {
  return unit(w2l_transpose(lr, local_onorm(lr, r, t)));
}

// One surface reference: three float4 in surface-list order.
//   triangle / tripar / ellipse: (v0, prim id bits) (e1, kind) (e2, 0)
//   sphere:            (centre, prim id bits) (radius, 0, 0, kind)
//   cylinder / sphere2: (0, prim id bits) (0, kind) (table index bits, 0)
// kind: 0 triangle, 1 parallelogram, 2 sphere, 3 instance of a model
// (record (0, n_prims + instance index) (0, kind) (0); see ref_test),
// 4 ellipse, 5 cylinder, 6 sphere2.  LOC: the local-primitive table.
// EXT: the scene may hold ellipses and local primitives (kinds 4-6); the
// code for them, inlined into every walk, would raise the register
// allocation of the walks over scenes that hold none (6 -> 5 waves per
// SIMD), so those scenes take the general walk.
template<bool EXT>
__device__ __forceinline__ bool ref_hit(const float4 *refs, const float4 *loc, uint32_t k,
                                        const RayF &r, float &t, float &u, float &v,
                                        uint32_t &prim) {
  // All three loads issued together (a sphere's third record is zeros),
  // as streaming loads: a reference is rarely read again soon by the same
  // CU, and keeping it out of the L1 leaves room for the nodes (Cornell
  // frame 121 -> 117 ms; the same hint on node loads cost 2 %).
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v *rv = (const f4v *)(refs + SGR_REF_STRIDE * k);
  const f4v av = __builtin_nontemporal_load(rv), bv = __builtin_nontemporal_load(rv + 1),
            cv = __builtin_nontemporal_load(rv + 2);
  const float4 a = make_float4(av.x, av.y, av.z, av.w), b = make_float4(bv.x, bv.y, bv.z, bv.w),
               c = make_float4(cv.x, cv.y, cv.z, cv.w);
  uint32_t kind = __float_as_uint(b.w);
  prim = __float_as_uint(a.w);
  if (!EXT) {
    if (kind != 2u) return tripar_hit(ld3(a), ld3(b), ld3(c), kind == 1u, r, t, u, v);
  } else {
    if (kind < 2u) return tripar_hit(ld3(a), ld3(b), ld3(c), kind == 1u, r, t, u, v);
    if (kind == SGR_REF_ELLIPSE) return ellipse_hit(ld3(a), ld3(b), ld3(c), r, t, u, v);
  }
  u = 0.0f;
  v = 0.0f;
  if (EXT && kind != 2u) return local_hit(loc + SGR_LOC_STRIDE * __float_as_uint(c.x), kind, r, t);
  if (sphere_hit(ld3(a), b.x, r.o, r.d, r.t0, t)) return t < r.t1;
  return false;
}

// nodes: node calls as the reference counts them (pruned children
// included); tests: primitive tests; fetch: node records loaded.
struct TraceCount {
  uint32_t nodes, tests, fetch;
  uint32_t povf;  // partial occluders met beyond SGR_PARTIAL_SEEN (multiplied undeduplicated)
};

// Child occupancy of a node in visiting order: bit j is set when the
// child visited j-th (parametric order, 3 and 4 swapped, xor the ray
// octant) exists.
__device__ __forceinline__ uint32_t visit_mask(uint32_t mask, uint32_t oct) {
  uint32_t m = mask;
  if (oct & 4u) m = ((m & 0x0fu) << 4) | ((m >> 4) & 0x0fu);
  if (oct & 2u) m = ((m & 0x33u) << 2) | ((m >> 2) & 0x33u);
  if (oct & 1u) m = ((m & 0x55u) << 1) | ((m >> 1) & 0x55u);
  uint32_t b3 = (m >> 3) & 1u, b4 = (m >> 4) & 1u;
  return (m & ~0x18u) | (b3 << 4) | (b4 << 3);
}

// Octree search (space/octree.cc:101-300) without recursion.
//
// The recursive search visits, for each node, its surfaces and then its
// children in perturbed parametric order, pruning a child by its
// parametric interval against the live ray extent (ray.t1 shrinks on
// every closer hit).  Here the current node lives in registers and a
// parent is pushed (first child, child mask and the siblings still to
// visit, parametric bounds) only when it still has children left to
// visit.  A child is
// tested against the same prune conditions, with the same ray extent,
// at the same point of the walk, before it is entered, so the surfaces
// are tested in the reference's order and the first of several equal-t
// hits wins exactly as on the CPU.  Node-call counts include the pruned
// calls, as the reference counts them.  ANY: stop at the first hit
// (space/space.cc:103-116 / :160-180).
//
// Traversal stack.  An entry is 8 bytes: the node's first child, its
// child mask, the siblings still to visit (a visit-order mask) and the
// node's depth.  Its parametric bounds are not stored: the bounds of a
// node follow from the root's by halving along the path to it, so a pop
// recomputes them from the root slabs and the lane's trail (3 bits per
// level: the child taken at each level of the current path, whose
// prefixes are the paths to every stack entry), with the same operations
// in the same order as the descent, hence the same bits.  The newest
// pushed entry's bounds are also kept (6 floats in LDS): a pop straight
// after a push, the common case, needs no recomputation.  The
// SGR_LDS_SLOTS newest entries of each lane live in LDS; older ones are
// spilled to the lane's private array when a newer one needs their slot
// and read back from there when popped.  LDS per 256-lane block: 9 x 8 B
// entries + 24 B of bounds per lane, 24 KB.
#define SGR_TRACE_BLOCK 256
#define SGR_LDS_SLOTS 9
#define SGR_TRAIL_LO 21  // levels in the first 64-bit trail word
struct LdsStack {
  uint2 *ent;    // &ent_lds[threadIdx.x] of a [SGR_LDS_SLOTS][SGR_TRACE_BLOCK] array
  float *bnd;    // &bnd_lds[threadIdx.x] of a [6][SGR_TRACE_BLOCK] array
};
#define SGR_STACK_LDS(name)                                             \
  __shared__ uint2 name##_ent[SGR_LDS_SLOTS * SGR_TRACE_BLOCK];         \
  __shared__ float name##_bnd[6 * SGR_TRACE_BLOCK];                     \
  const LdsStack name{&name##_ent[threadIdx.x], &name##_bnd[threadIdx.x]}
static_assert(SGR_MAX_DEPTH <= 2 * SGR_TRAIL_LO, "the trail holds 2 x 21 levels");

// ---------------------------------------------------------------------------
// The walk is a resumable state machine, for lane-refill traversal:
// trav_init sets up a ray (false: it misses the root cube), trav_step
// does one step (true: the ray is finished).  A step enters the pending
// node (its surfaces, in list order) unless the last step popped it,
// then picks the child to enter next or pops one stack entry.  Every
// step thus costs one node load at most and one evaluation of the
// eight children, so the lanes of a wave stay on the same instructions.
#define SGR_PARTIAL_SEEN 8  // partial occluders a shadow ray remembers (counted once each)
struct TravState {
  RayF ray;
  // Shadow rays in scenes with partial occluders (INST walks): the index
  // of the medium at the ray's origin, the transmittance so far and the
  // partial occluders already multiplied.
  float medium;
  C3 trans;
  uint32_t nseen;
  uint32_t seen[SGR_PARTIAL_SEEN];
  uint32_t hinst;     // instance of the hit (SGR_MISS: a surface of the scene itself)
  uint32_t oct, cn, first, mask, vis;
  float bx0, bx1, by0, by1, bz0, bz1;
  uint64_t tr0, tr1;  // trail: child taken at levels 0-20, 21-41
  int sp, nl, dep;    // stack size, entries in LDS, depth of the current node
  bool enter, top_ok; // top_ok: the LDS bounds are the top entry's
  uint32_t hit;
  float hu, hv;
};

// Root slabs of the ray (space/octree.cc:101-162): parametric bounds of
// the root cube per axis, ordered.
__device__ __forceinline__ void root_slabs(const SceneView &S, const RayF &ray, float &bx0,
                                           float &bx1, float &by0, float &by1, float &bz0,
                                           float &bz1) {
  float ix = ray.d.x == 0.0f ? 1e9f : 1.0f / ray.d.x;
  bx0 = (S.origin[0] - ray.o.x) * ix;
  bx1 = (S.origin[0] - ray.o.x + S.size) * ix;
  if (bx0 > bx1) { float q = bx0; bx0 = bx1; bx1 = q; }
  float iy = ray.d.y == 0.0f ? 1e9f : 1.0f / ray.d.y;
  by0 = (S.origin[1] - ray.o.y) * iy;
  by1 = (S.origin[1] - ray.o.y + S.size) * iy;
  if (by0 > by1) { float q = by0; by0 = by1; by1 = q; }
  float iz = ray.d.z == 0.0f ? 1e9f : 1.0f / ray.d.z;
  bz0 = (S.origin[2] - ray.o.z) * iz;
  bz1 = (S.origin[2] - ray.o.z + S.size) * iz;
  if (bz0 > bz1) { float q = bz0; bz0 = bz1; bz1 = q; }
}

// Bounds of child CHILD (parametric index, before the octant flip) of a
// node with bounds b: the lower or upper half per axis.
__device__ __forceinline__ void child_bounds(uint32_t child, float &bx0, float &bx1, float &by0,
                                             float &by1, float &bz0, float &bz1) {
  const float xh = (bx1 - bx0) / 2.0f, yh = (by1 - by0) / 2.0f, zh = (bz1 - bz0) / 2.0f;
  const float cx = bx0 + ((child & 4) ? xh : 0.0f);
  const float cy = by0 + ((child & 2) ? yh : 0.0f);
  const float cz = bz0 + ((child & 1) ? zh : 0.0f);
  bx0 = cx; bx1 = cx + xh;
  by0 = cy; by1 = cy + yh;
  bz0 = cz; bz1 = cz + zh;
}

__device__ __forceinline__ bool trav_init(const SceneView &S, TravState &T, TraceCount &cnt) {
  T.hit = SGR_MISS;
  T.hinst = SGR_MISS;
  T.hu = 0.0f;
  T.hv = 0.0f;
  if (S.n_nodes == 0) return false;
  const RayF &ray = T.ray;
  T.oct = (ray.d.x >= 0.0f ? 0u : 4u) | (ray.d.y >= 0.0f ? 0u : 2u) | (ray.d.z >= 0.0f ? 0u : 1u);
  float bx0, bx1, by0, by1, bz0, bz1;
  root_slabs(S, ray, bx0, bx1, by0, by1, bz0, bz1);
  cnt.nodes++;
  if (fmx(ray.t0, fmx(bx0, fmx(by0, bz0))) >= fmn(ray.t1, fmn(bx1, fmn(by1, bz1)))) return false;
  T.bx0 = bx0; T.bx1 = bx1; T.by0 = by0; T.by1 = by1; T.bz0 = bz0; T.bz1 = bz1;
  T.sp = 0;
  T.nl = 0;
  T.dep = 0;
  T.tr0 = 0;
  T.tr1 = 0;
  T.cn = 0;
  T.enter = true;
  T.top_ok = false;
  return true;
}

// Visit-order bit i <-> child index (parametric order with 3 and 4
// swapped): exchange bits 3 and 4 of an 8-bit child mask.
__device__ __forceinline__ uint32_t swap34(uint32_t m) {
  return (m & 0xe7u) | ((m >> 1) & 0x08u) | ((m << 1) & 0x10u);
}

// The children of the current node the recursive search would call
// (C1: the child-call test of space/octree.cc:269-272 holds) and enter
// (OK: the entry test of :186-188 on the child slabs of :276-291), for
// all eight children at once, as visit-order bit masks.  The operands and the
// order of the fmx / fmn chains are those of the per-child tests (child
// slab = low half [b0 + 0, (b0 + 0) + h] or high half [b0 + h,
// (b0 + h) + h], h = (b1 - b0) / 2), so the decisions are the same bit
// for bit, NaN operands included.
__device__ __forceinline__ void child_masks(const TravState &T, uint32_t &c1, uint32_t &ok) {
  const float t0 = T.ray.t0, t1 = T.ray.t1;
  const float xh = (T.bx1 - T.bx0) / 2.0f, yh = (T.by1 - T.by0) / 2.0f,
              zh = (T.bz1 - T.bz0) / 2.0f;
  const float xm = T.bx0 + xh, ym = T.by0 + yh, zm = T.bz0 + zh;
  const float xl = T.bx0 + 0.0f, yl = T.by0 + 0.0f, zl = T.bz0 + 0.0f;
  const float xe[2] = {xl + xh, xm + xh}, ye[2] = {yl + yh, ym + yh}, ze[2] = {zl + zh, zm + zh};
  const float xs[2] = {xl, xm}, ys[2] = {yl, ym}, zs[2] = {zl, zm};
  uint32_t m = 0;
#pragma unroll
  for (int yz = 0; yz < 4; yz++) {
    const float a = fmx(ys[yz >> 1], zs[yz & 1]), b = fmn(ye[yz >> 1], ze[yz & 1]);
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const bool in = !(fmx(t0, fmx(xs[x], a)) >= fmn(t1, fmn(xe[x], b)));
      m |= in ? 1u << (x * 4 + yz) : 0u;
    }
  }
  const uint32_t cx = (t0 < xm ? 0x0fu : 0u) | (t1 > xm ? 0xf0u : 0u);
  const uint32_t cy = (t0 < ym ? 0x33u : 0u) | (t1 > ym ? 0xccu : 0u);
  const uint32_t cz = (t0 < zm ? 0x55u : 0u) | (t1 > zm ? 0xaau : 0u);
  c1 = swap34(cx & cy & cz);
  ok = swap34(m);
}

// Bounds of the node at depth DEP of the current path: the root slabs
// halved along the trail's first DEP levels.
__device__ __forceinline__ void replay_bounds(const SceneView &S, TravState &T, int dep) {
  float bx0, bx1, by0, by1, bz0, bz1;
  root_slabs(S, T.ray, bx0, bx1, by0, by1, bz0, bz1);
  uint64_t t = T.tr0;
  for (int l = 0; l < dep; l++) {
    if (l == SGR_TRAIL_LO) t = T.tr1;
    child_bounds((uint32_t)t & 7u, bx0, bx1, by0, by1, bz0, bz1);
    t >>= 3;
  }
  T.bx0 = bx0; T.bx1 = bx1; T.by0 = by0; T.by1 = by1; T.bz0 = bz0; T.bz1 = bz1;
}

// Scene::occludes at a surface a shadow ray meets (space/space.cc:
// 131-194, material/material.h:168-192): returns true when the ray goes
// on -- a partially occluding surface (thin glass, stencil) multiplies the ray's
// transmittance by its own, once however many octree nodes list it (the
// reference's negative mailbox skips a surface already met; here the
// lane remembers SGR_PARTIAL_SEEN of them), unless the product falls
// below Eps (Color < Eps: component sum below 3 Eps).  A fully occluding
// surface, or a surface of an instanced model (models hold none),
// stops it.  Thin glass transmits colour x (1 - Fresnel (medium, ior) at
// the angle between the unit geometric normal and the reversed ray,
// material/thin-glass.cc:123-140); a stencil max (1 - opacity, 0).
__device__ __forceinline__ bool partial_pass(const SceneView &S, TravState &T, uint32_t prim,
                                             uint32_t inst, float t, TraceCount &cnt) {
  if (inst != SGR_MISS) return false;
  const DevMaterial &M = S.mats[S.prim_mat[prim]];
  if (!M.partial) return false;
  bool met = false;
#pragma unroll
  for (int k = 0; k < SGR_PARTIAL_SEEN; k++) met = met || ((uint32_t)k < T.nseen && T.seen[k] == prim);
  if (met) return true;
  if (T.nseen < SGR_PARTIAL_SEEN) {
#pragma unroll
    for (int k = 0; k < SGR_PARTIAL_SEEN; k++)
      if ((uint32_t)k == T.nseen) T.seen[k] = prim;
    T.nseen++;
  } else {
    cnt.povf++;
  }
  if (M.type == 8) {  // stencil: max(1 - opacity, 0) (material/stencil.cc:160-166)
    T.trans = T.trans * C3{fmaxf(1.0f - M.color[0], 0.0f), fmaxf(1.0f - M.color[1], 0.0f),
                           fmaxf(1.0f - M.color[2], 0.0f)};
    return !(csum(T.trans) < 3e-15f);
  }
  V3 n;
  if (prim < S.n_tri_like) {
    const float4 a = S.tri[3 * prim + 1], b = S.tri[3 * prim + 2];
    n = unit(cross(v3(b.x, b.y, b.z), v3(a.x, a.y, a.z)));  // e2 x e1
  } else if (prim < S.n_sph_end) {
    const float4 sp = S.sph[prim - S.n_tri_like];
    n = unit((T.ray.o + T.ray.d * t) - v3(sp.x, sp.y, sp.z));
  } else {
    n = local_normal(S.loc + SGR_LOC_STRIDE * (prim - S.n_sph_end), T.ray, t);
  }
  const float c = dot(n, -T.ray.d);
  const float refl = fresnel(T.medium, M.ior_n, 0.0f, c);
  T.trans = T.trans * (cload(M.color) * (1.0f - refl));
  return !(csum(T.trans) < 3e-15f);
}

// Stack of the walk inside an instanced model (a second LDS ring and
// spill array: the scene's walk keeps its own entries while it waits).
struct Nest {
  LdsStack L;
  uint2 *spill;
};

template<bool ANY, bool INST, bool EXT = INST>
__device__ __forceinline__ void trace_one(const SceneView &S, TravState &T, TraceCount &cnt,
                                          LdsStack L, uint2 *spill, Nest N);

// Surface reference K of the scene: a primitive, or (INST) an instance of
// a model: the ray is taken to model space by the inverse transform (the
// parametric t is kept, the direction is not renormalised) and the
// model's octree is searched (Instance::Renderable::intersect /
// intersects, surface/instance.cc:114-146); its node calls and surface
// tests count with the scene's, as they share the render's IsecStats.
template<bool ANY, bool INST, bool EXT>
__device__ __forceinline__ bool ref_test(const SceneView &S, uint32_t k, const RayF &r, float &t,
                                         float &u, float &v, uint32_t &prim, uint32_t &inst,
                                         TraceCount &cnt, Nest N) {
  if (!INST) return ref_hit<EXT>(S.refs, S.loc, k, r, t, u, v, prim);
  const float4 *rec = S.refs + SGR_REF_STRIDE * k;
  if (__float_as_uint(rec[1].w) != 3u) return ref_hit<EXT>(S.refs, S.loc, k, r, t, u, v, prim);
  const uint32_t i = __float_as_uint(rec[0].w) - S.n_prims;
  const DevInstance &I = S.insts[i];
  TravState T2;
  T2.ray.o = xform_pos(I.w2l, r.o);
  T2.ray.d = xform_vec(I.w2l, r.d);
  T2.ray.t0 = r.t0;
  T2.ray.t1 = r.t1;
  trace_one<ANY, false, EXT>(S.models[I.model], T2, cnt, N.L, N.spill, N);
  if (T2.hit == SGR_MISS) return false;
  t = T2.ray.t1;
  u = T2.hu;
  v = T2.hv;
  prim = T2.hit;
  inst = i;
  return true;
}

template<bool ANY, bool INST, bool EXT = INST>
__device__ __forceinline__ bool trav_step(const SceneView &S, TravState &T, TraceCount &cnt,
                                          LdsStack L, uint2 *spill, Nest N) {
  if (T.enter) {
    float t, u, v;
    uint32_t prim;
    cnt.fetch++;
    const uint4 nd = S.nodes[T.cn];
    for (uint32_t k = nd.z, e = nd.z + nd.w; k < e; k++) {
      cnt.tests++;
      uint32_t inst = SGR_MISS;
      if (ref_test<ANY, INST, EXT>(S, k, T.ray, t, u, v, prim, inst, cnt, N)) {
        if (ANY && INST && S.has_partial && partial_pass(S, T, prim, inst, t, cnt)) continue;
        T.hit = prim;
        T.hinst = inst;
        if (ANY) return true;
        T.ray.t1 = t;
        T.hu = u;
        T.hv = v;
      }
    }
    T.mask = nd.y;
    T.first = nd.x;
    T.vis = visit_mask(T.mask, T.oct);
  }
  // Children still to visit that the search would call (counted) and
  // enter, with the ray extent as it is now.  Siblings after the one
  // entered are tested again when their entry is popped: the extent can
  // only shrink, so one that fails now would fail then too.
  uint32_t c1, ok;
  child_masks(T, c1, ok);
  const uint32_t cand = T.vis & c1, go = cand & ok;
  cnt.nodes += (uint32_t)__popc(cand & ~ok) + (go ? 1u : 0u);
  if (go) {
    const uint32_t i = (uint32_t)__builtin_ctz(go), later = go & (go - 1u);
    if (later) {
      uint2 *slot = L.ent + (size_t)(T.sp % SGR_LDS_SLOTS) * SGR_TRACE_BLOCK;
      if (T.nl == SGR_LDS_SLOTS) {  // the slot holds entry sp - SLOTS: spill it
        spill[T.sp - SGR_LDS_SLOTS] = *slot;
      } else {
        T.nl++;
      }
      *slot = make_uint2(T.first, T.mask | (later << 8) | ((uint32_t)T.dep << 16));
      L.bnd[0] = T.bx0; L.bnd[SGR_TRACE_BLOCK] = T.bx1;
      L.bnd[2 * SGR_TRACE_BLOCK] = T.by0; L.bnd[3 * SGR_TRACE_BLOCK] = T.by1;
      L.bnd[4 * SGR_TRACE_BLOCK] = T.bz0; L.bnd[5 * SGR_TRACE_BLOCK] = T.bz1;
      T.top_ok = true;
      T.sp++;
    }
    const uint32_t child = (i == 3) ? 4u : (i == 4) ? 3u : i;
    const uint32_t real = child ^ T.oct;
    T.cn = T.first + __popc(T.mask & ((1u << real) - 1u));
    child_bounds(child, T.bx0, T.bx1, T.by0, T.by1, T.bz0, T.bz1);
    // the trail's entry for this level (stale bits of an earlier path
    // cleared)
    if (T.dep < SGR_TRAIL_LO) {
      const int sh = 3 * T.dep;
      T.tr0 = (T.tr0 & ~(7ull << sh)) | ((uint64_t)child << sh);
    } else {
      const int sh = 3 * (T.dep - SGR_TRAIL_LO);
      T.tr1 = (T.tr1 & ~(7ull << sh)) | ((uint64_t)child << sh);
    }
    T.dep++;
    T.enter = true;
    return false;
  }
  if (T.sp == 0) return true;
  T.sp--;
  uint2 a;
  if (T.nl > 0) {
    a = L.ent[(size_t)(T.sp % SGR_LDS_SLOTS) * SGR_TRACE_BLOCK];
    T.nl--;
  } else {
    a = spill[T.sp];
  }
  T.first = a.x;
  T.mask = a.y & 0xffu;
  T.vis = (a.y >> 8) & 0xffu;
  T.dep = (int)(a.y >> 16);
  if (T.top_ok) {  // the entry pushed last: its bounds are in LDS
    T.bx0 = L.bnd[0]; T.bx1 = L.bnd[SGR_TRACE_BLOCK];
    T.by0 = L.bnd[2 * SGR_TRACE_BLOCK]; T.by1 = L.bnd[3 * SGR_TRACE_BLOCK];
    T.bz0 = L.bnd[4 * SGR_TRACE_BLOCK]; T.bz1 = L.bnd[5 * SGR_TRACE_BLOCK];
  } else {
    replay_bounds(S, T, T.dep);
  }
  T.top_ok = false;
  T.enter = false;
  return false;
}

// One ray from start to end: brute force over the surface list
// (TrivSpace, space/triv-space.h:48-58) or the octree walk.
template<bool ANY, bool INST, bool EXT>
__device__ __forceinline__ void trace_one(const SceneView &S, TravState &T, TraceCount &cnt,
                                          LdsStack L, uint2 *spill, Nest N) {
  if (ANY && INST) {  // T.medium is the caller's
    T.trans = c3(1.0f);
    T.nseen = 0;
  }
  if (S.use_triv) {
    T.hit = SGR_MISS;
    T.hinst = SGR_MISS;
    T.hu = 0.0f;
    T.hv = 0.0f;
    float t, u, v;
    uint32_t prim;
    for (uint32_t q = 0; q < S.n_surfs; q++) {
      cnt.tests++;
      uint32_t inst = SGR_MISS;
      if (ref_test<ANY, INST, EXT>(S, q, T.ray, t, u, v, prim, inst, cnt, N)) {
        if (ANY && INST && S.has_partial && partial_pass(S, T, prim, inst, t, cnt)) continue;
        T.hit = prim;
        T.hinst = inst;
        if (ANY) break;
        T.ray.t1 = t;
        T.hu = u;
        T.hv = v;
      }
    }
  } else if (trav_init(S, T, cnt)) {
    while (!trav_step<ANY, INST, EXT>(S, T, cnt, L, spill, N)) {
    }
  }
}

// Plain driver: lane i of the grid walks rays i, i + G, ... to the end,
// one at a time (ray ranges by block_range order).  Cheaper per step than
// trace_refill when walks are short (shallow octrees), where the waves
// are balanced anyway, and for camera rays, whose lanes (the samples of
// one or two pixels) walk the same nodes in step, so their loads
// coalesce.
// INST: the scene may hold instances (N is the stack of the walk inside
// a model).
template<bool ANY, bool MULTI, bool INST, typename Get, typename Put>
__device__ __forceinline__ void trace_plain(const SceneView &S1, uint32_t jb, uint32_t je,
                                            Get get, Put put, TraceCount &cnt, LdsStack L,
                                            Nest N) {
  TravState T;
  uint2 spill[SGR_MAX_DEPTH];
  for (uint32_t i = jb + threadIdx.x; i < je; i += blockDim.x) {
    T.medium = 1.0f;  // a shadow ray's medium: air unless GET says otherwise
    const SceneView *g = get(i, T);
    trace_one<ANY, INST>(MULTI ? *g : S1, T, cnt, L, spill, N);
    put(i, T);
  }
}

// Lane-refill driver (scenes without instances): every lane of a wave
// walks its own ray; when at
// least SGR_REFILL lanes are idle the wave takes that many more rays with
// one atomic on a slice head (a wave-wide ballot ranks the idle lanes),
// so no lane idles for long while its wave still has rays to finish.
// 48: with the slice heads on lines of their own, batches of 32 / 40 /
// 48 / 56 gave 124.3 / 123.3 / 123.3 / 123.7 ms on the 1.31 M-triangle
// frame (64 lanes, a wave idle until all its lanes are, lost 10 %
// before).  The N rays are
// cut into 8 slices, one per XCD (blockIdx.x & 7 under round-robin
// dispatch), so neighbouring rays share an L2; a wave whose slice is
// used up moves on to the next slices in turn, so no XCD idles while
// another still has rays.  GET(i, ray) loads ray i and returns its scene
// (used when MULTI; S1 otherwise); PUT(i, state) stores the result.  The
// loop exits when all 8 slices are used up and the wave's last ray is
// done (every wave gets there: each step advances a finite walk).
#define SGR_REFILL 48
template<bool ANY, bool MULTI, typename Get, typename Put>
__device__ __forceinline__ void trace_refill(const SceneView &S1, uint32_t n, uint32_t *heads,
                                             Get get, Put put, TraceCount &cnt, LdsStack L) {
  // Blocks beyond what a short batch needs (a ray per lane, a block per
  // XCD at least) leave at once; the waves that stay take every slice.
  const uint32_t need = ((n + blockDim.x - 1u) / blockDim.x + 7u) & ~7u;
  if (blockIdx.x >= (need > 8u ? need : 8u)) return;
  TravState T;
  uint2 spill[SGR_MAX_DEPTH];
  bool active = false;
  uint32_t id = 0, tried = 0, slice = blockIdx.x & 7u;
  const SceneView *Sp = &S1;
  const uint32_t lane = __lane_id();
  for (;;) {
    const uint64_t idle = __ballot(!active);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (tried == 8u && nidle == 64u) break;
    if (tried < 8u && nidle >= SGR_REFILL) {
      const uint32_t beg = (uint32_t)((uint64_t)n * slice / 8u);
      const uint32_t len = (uint32_t)((uint64_t)n * (slice + 1u) / 8u) - beg;
      const int leader = __ffsll((unsigned long long)idle) - 1;
      uint32_t base = 0;
      if ((int)lane == leader) base = atomicAdd(&heads[slice * SGR_HEAD_STRIDE], nidle);
      base = (uint32_t)__shfl((int)base, leader);
      if (base + nidle >= len) {  // this slice is used up: next time, the next one
        tried++;
        slice = (slice + 1u) & 7u;
      }
      if (!active) {
        const uint32_t k = base + (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
        if (k < len) {
          id = beg + k;
          T.medium = 1.0f;
          const SceneView *g = get(id, T);
          if (MULTI) Sp = g;
          const SceneView &S = MULTI ? *Sp : S1;
          if (S.use_triv) {
            // brute force (TrivSpace, space/triv-space.h:48-58), in one go
            T.hit = SGR_MISS;
            T.hinst = SGR_MISS;
            T.hu = 0.0f;
            T.hv = 0.0f;
            float t, u, v;
            uint32_t prim;
            for (uint32_t q = 0; q < S.n_surfs; q++) {
              cnt.tests++;
              if (ref_hit<false>(S.refs, S.loc, q, T.ray, t, u, v, prim)) {
                T.hit = prim;
                if (ANY) break;
                T.ray.t1 = t;
                T.hu = u;
                T.hv = v;
              }
            }
            put(id, T);
          } else if (trav_init(S, T, cnt)) {
            active = true;
          } else {
            put(id, T);
          }
        }
      }
    }
    if (active && trav_step<ANY, false>(MULTI ? *Sp : S1, T, cnt, L, spill, Nest{})) {
      put(id, T);
      active = false;
    }
  }
}

// Sum a per-lane counter over the wave and add it once to *dst.
__device__ __forceinline__ void wave_count(unsigned long long *dst, uint32_t v) {
  unsigned long long x = v;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if (__lane_id() == 0 && x) atomicAdd(dst, x);
}

}  // namespace sgr
