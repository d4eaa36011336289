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

// geometry/quadratic-roots.h:30-70 + geometry/sphere-isec.h:35-60
__device__ __forceinline__ bool sphere_hit(V3 c, float radius, V3 ro, V3 rd, float min_t,
                                           float &t) {
  V3 off = ro - c;
  float a = len2(rd);
  float b = 2.0f * dot(rd, off);
  float cc = len2(off) - radius * radius;
  float r0, r1;
  unsigned n;
  if (a == 0.0f) {
    if (b == 0.0f) return false;
    r0 = -cc / b;
    n = 1;
    r1 = r0;
  } else {
    float disc = b * b - 4.0f * a * cc;
    if (disc < 0.0f) return false;
    float sd = sqrtf(disc);
    float q = -(b + (b < 0.0f ? -sd : sd)) / 2.0f;
    r0 = q / a;
    if (q == 0.0f) {
      n = 1;
      r1 = r0;
    } else {
      r1 = cc / q;
      if (r0 > r1) { float tmp = r0; r0 = r1; r1 = tmp; }
      n = 2;
    }
  }
  if (r0 > min_t) { t = r0; return true; }
  if (n == 2 && r1 > min_t) { t = r1; return true; }
  return false;
}

__device__ __forceinline__ bool prim_hit(const SceneView &S, uint32_t p, const RayF &r,
                                         float &t, float &u, float &v) {
  if (p < S.n_tri_like) {
    float4 a = S.tri[3 * p], b = S.tri[3 * p + 1], c = S.tri[3 * p + 2];
    return tripar_hit(ld3(a), ld3(b), ld3(c), a.w != 0.0f, r, t, u, v);
  }
  float4 s = S.sph[p - S.n_tri_like];
  u = 0.0f;
  v = 0.0f;
  if (sphere_hit(ld3(s), s.w, r.o, r.d, r.t0, t)) return t < r.t1;
  return false;
}

struct TraceCount {
  uint32_t nodes, tests;
};

// Octree search (space/octree.cc:101-300).  The recursion is kept as an
// explicit stack of parent frames (node | next-child cursor, parametric
// bounds); the current frame lives in registers.  Child visiting order,
// live ray.t1 pruning and surface order match the recursive search, so
// the first of several equal-t hits wins exactly as on the CPU.
// ANY: stop at the first hit (space/space.cc:103-116 / :160-180).
template<bool ANY>
__device__ uint32_t trace_ray(const SceneView &S, RayF &ray, float &hu, float &hv,
                              TraceCount &cnt) {
  uint32_t hit = SGR_MISS;
  float t, u, v;
  if (S.use_triv) {  // TrivSpace (space/triv-space.h:48-58)
    for (uint32_t k = 0; k < S.n_surfs; k++) {
      uint32_t p = S.surfs[k];
      cnt.tests++;
      if (prim_hit(S, p, ray, t, u, v)) {
        if (ANY) return p;
        ray.t1 = t;
        hit = p;
        hu = u;
        hv = v;
      }
    }
    return hit;
  }
  if (S.n_nodes == 0) return hit;
  const uint32_t oct = (ray.d.x >= 0.0f ? 0u : 4u) | (ray.d.y >= 0.0f ? 0u : 2u) |
                       (ray.d.z >= 0.0f ? 0u : 1u);
  float bx0, bx1, by0, by1, bz0, bz1;
  {
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
  uint32_t st_node[SGR_MAX_DEPTH];
  float st_b[SGR_MAX_DEPTH][6];
  int sp = 0;
  uint32_t cn = 0, cmask = 0, cfirst = 0, ci = 0;
  float xh = 0, yh = 0, zh = 0, xm = 0, ym = 0, zm = 0;
  bool enter = true;
  for (;;) {
    bool pop = false;
    if (enter) {
      cnt.nodes++;
      float min_t = fmx(ray.t0, fmx(bx0, fmx(by0, bz0)));
      float max_t = fmn(ray.t1, fmn(bx1, fmn(by1, bz1)));
      if (min_t >= max_t) {
        pop = true;
      } else {
        uint4 nd = S.nodes[cn];
        for (uint32_t k = 0; k < nd.w; k++) {
          uint32_t p = S.surfs[nd.z + k];
          cnt.tests++;
          if (prim_hit(S, p, ray, t, u, v)) {
            if (ANY) return p;
            ray.t1 = t;
            hit = p;
            hu = u;
            hv = v;
          }
        }
        cmask = nd.y;
        cfirst = nd.x;
        if (!cmask) {
          pop = true;
        } else {
          xh = (bx1 - bx0) / 2.0f;
          yh = (by1 - by0) / 2.0f;
          zh = (bz1 - bz0) / 2.0f;
          xm = bx0 + xh;
          ym = by0 + yh;
          zm = bz0 + zh;
          ci = 0;
          enter = false;
        }
      }
    }
    if (!pop) {
      // Next child in perturbed parametric order (3 and 4 swapped).
      bool down = false;
      while (ci < 8) {
        uint32_t i = ci++;
        uint32_t child = (i == 3) ? 4u : (i == 4) ? 3u : i;
        uint32_t real = child ^ oct;
        if (!(cmask & (1u << real))) continue;
        float t0 = ray.t0, t1 = ray.t1;
        if (((child & 4) ? (t1 > xm) : (t0 < xm)) && ((child & 2) ? (t1 > ym) : (t0 < ym)) &&
            ((child & 1) ? (t1 > zm) : (t0 < zm))) {
          st_node[sp] = cn | (ci << 28);
          st_b[sp][0] = bx0; st_b[sp][1] = bx1; st_b[sp][2] = by0;
          st_b[sp][3] = by1; st_b[sp][4] = bz0; st_b[sp][5] = bz1;
          sp++;
          float cx = bx0 + ((child & 4) ? xh : 0.0f);
          float cy = by0 + ((child & 2) ? yh : 0.0f);
          float cz = bz0 + ((child & 1) ? zh : 0.0f);
          bx0 = cx; bx1 = cx + xh;
          by0 = cy; by1 = cy + yh;
          bz0 = cz; bz1 = cz + zh;
          cn = cfirst + __popc(cmask & ((1u << real) - 1u));
          enter = true;
          down = true;
          break;
        }
      }
      if (down) continue;
    }
    // Return to the parent frame.
    if (sp == 0) break;
    sp--;
    uint32_t packed = st_node[sp];
    cn = packed & 0x0fffffffu;
    ci = packed >> 28;
    bx0 = st_b[sp][0]; bx1 = st_b[sp][1]; by0 = st_b[sp][2];
    by1 = st_b[sp][3]; bz0 = st_b[sp][4]; bz1 = st_b[sp][5];
    uint4 nd = S.nodes[cn];
    cmask = nd.y;
    cfirst = nd.x;
    xh = (bx1 - bx0) / 2.0f;
    yh = (by1 - by0) / 2.0f;
    zh = (bz1 - bz0) / 2.0f;
    xm = bx0 + xh;
    ym = by0 + yh;
    zm = bz0 + zh;
    enter = false;
  }
  return hit;
}

// Sum a per-lane counter over the wave and add it once to *dst.
__device__ __forceinline__ void wave_count(unsigned long long *dst, uint32_t v) {
  unsigned long long x = v;
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  if (__lane_id() == 0 && x) atomicAdd(dst, x);
}

}  // namespace sgr
