// Device-side data layout and small math for the gfx950 kernels.
//
// Arithmetic follows the reference's fp32 operation order where results
// must match it bit for bit (ray/triangle tests and octree traversal);
// the library is built with -ffp-contract=off, IEEE division and square
// root, and denormals kept, so each written operation is one IEEE
// operation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sgr {

#define SGR_REF_STRIDE 3  // float4s per surface reference on the device
// Surface-record kinds beyond triangle 0, parallelogram 1, sphere 2 and
// instance 3: an ellipse (its bounding parallelogram inline) and the
// local primitives, whose record holds an index into the scene's table
// of SGR_LOC_STRIDE float4s each (world-to-local rows, local-to-world
// rows, then the kind).
#define SGR_REF_ELLIPSE 4u
#define SGR_REF_CYLINDER 5u
#define SGR_REF_SPHERE2 6u
#define SGR_LOC_STRIDE 8
#define SGR_MAX_DEPTH 42  // octree levels a traversal can hold (checked at scene creation)
// uint32 words between the fetch heads of the lane-refill walk: each head
// has a 128-byte line to itself, so the atomics of different XCD slices
// do not queue on one line.
#define SGR_HEAD_STRIDE 32

// ---------------------------------------------------------------------------
// Scene view passed to kernels by value.
// ---------------------------------------------------------------------------
struct DevMaterial {
  int type;
  float color[3];
  float gloss[3];
  float m, ior_n, ior_k;
  int bsdf;      // material index providing the BSDF (-1 = none)
  int medium;    // 1 if the BSDF material is glass (has a medium)
  // Cook-Torrance constants (material/cook-torrance.cc:38-70)
  float diff_weight, inv_diff_weight, inv_gloss_weight;
  unsigned gloss_layer, have_layers;
  // Mirror: an underlying BSDF present: a Lambert of colour `color`, or
  // (has_under) the Lambert, Cook-Torrance or Phong material at this
  // entry's index + under_off
  int mirror_under;
  int partial;       // lets shadow rays through, times its transmittance
  // Stencil (material/stencil.cc:30-52) and mirror: the underlying BSDF
  // material at this entry's index + under_off of the table (has_under: there is
  // one), the opacity's intensity and its reciprocal (snapped to 0 / 1
  // below 0.001 / above 0.999); the opacity is in color.
  int under_off, has_under;
  float op_int, inv_op_int;
};

struct DevLight {
  // 0 = tripar area light, 1 = sphere light, 2 = envmap; surface lights
  // (light/surface-light-sampler.cc) of 3 = an ellipse, 4 = a cylinder,
  // 5 = a sphere2; 6 = a point light (prim: its row in SceneView::plights)
  int kind;
  uint32_t prim;
  float intensity[3];
};

struct DevMesh {
  float axis[3];
  float center[3];
  int left_handed;
  int has_normals;  // vertex normals in SceneView::tri_vn
};

#define SGR_ENV_GUIDE_ROW 256
#define SGR_ENV_GUIDE_COL 32

// An instance of a model (surface/instance.cc): 3 x 4 row-major affine
// transforms, local to world and its inverse.
struct DevInstance {
  uint32_t model;
  float l2w[12];
  float w2l[12];
};

struct SceneView {
  const uint4 *nodes;        // first_child, mask, surf_begin, surf_count
  const uint32_t *surfs;     // primitive ids (host export only)
  const float4 *refs;        // 3 per surface reference, surface-list order
  const float4 *tri;         // 3 per triangle-like prim: v0|flags, e1, e2
  const float4 *sph;         // centre, radius
  const float4 *loc;         // local primitives (SGR_LOC_STRIDE per primitive)
  // point lights, 4 per light: (pos, cos_half_angle), (frame x,
  // cos_half_core_angle), (frame y, 0), (frame z, 0)
  const float4 *plights;
  const float4 *tri_vn;      // 3 per mesh triangle: its vertex normals (or null: none)
  const uint32_t *prim_mat;  // material index per prim
  const uint32_t *prim_mesh; // mesh index per triangle (0xffffffff others)
  const DevMesh *meshes;
  const DevMaterial *mats;
  const DevLight *lights;
  const float4 *env_tex;     // rgb_ per texel, row 0 = top
  const float *env_row_cum;
  const float *env_col_cum;
  // Guide tables of the two searches: entry g of a table of G + 1 is the
  // lower_bound of g / G in its cumulative array (rows: one table of
  // SGR_ENV_GUIDE_ROW + 1; columns: SGR_ENV_GUIDE_COL + 1 per row).
  const uint32_t *env_row_guide;
  const uint32_t *env_col_guide;
  // Walk with lane refill (deep octrees; camera rays excepted) or one ray
  // per lane at a time.
  uint32_t walk_refill;
  // Two-level instancing: the models' own views (geometry and octree in
  // model space; materials shared with this scene) and the instances,
  // whose primitive ids in this scene's surface lists are n_prims + i.
  const SceneView *models;
  const DevInstance *insts;
  uint32_t n_insts;
  // Surfaces of partially occluding materials (thin glass) in this scene:
  // shadow rays take the general walk, which multiplies their
  // transmittance (such scenes, like instanced ones, use the INST kernels).
  uint32_t has_partial;
  // Ellipses or local primitives (cylinder, sphere2) in this scene or its
  // models: its walks take the general kernels, whose surface tests,
  // frames and light samplers cover them.
  uint32_t has_ext;
  uint32_t n_tri_like;       // triangles + tripars + ellipses
  uint32_t n_sph_end;        // n_tri_like + spheres: the first local primitive
  uint32_t n_tri;            // mesh triangles
  uint32_t n_prims;
  uint32_t n_nodes;
  uint32_t n_surfs;
  uint32_t n_lights;
  uint32_t use_triv;
  uint32_t env_w, env_h, env_hw, env_hh;
  int has_env;
  float env_frame[9];
  float origin[3];
  float size;
  float min_trace;
  float horizon;
  float bbox_lo[3], bbox_hi[3];
};

// ---------------------------------------------------------------------------
// Vector helpers (geometry/vec.h semantics: division multiplies by the
// reciprocal, unit() likewise, ternary min/max).
// ---------------------------------------------------------------------------
struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float a, float b, float c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ float vlen(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ V3 unit(V3 a) { float l = vlen(a); return a * (1.0f / l); }
__device__ __forceinline__ V3 vdiv(V3 a, float d) { return a * (1.0f / d); }
__device__ __forceinline__ float fmn(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float fmx(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float clamp01f(float v) {
  v = v < 1.0f ? v : 1.0f;
  return v > 0.0f ? v : 0.0f;
}
__device__ __forceinline__ V3 ld3(const float4 &f) { return V3{f.x, f.y, f.z}; }

// A 3 x 4 row-major affine transform M applied to a position, a vector
// (geometry/pos.h:80-97, geometry/vec.h:81-95) and, transposed, to a
// vector (geometry/xform.h:295-309, the normal transform when M is the
// inverse).
__device__ __forceinline__ V3 xform_pos(const float *m, V3 p) {
  return V3{p.x * m[0] + p.y * m[1] + p.z * m[2] + m[3],
            p.x * m[4] + p.y * m[5] + p.z * m[6] + m[7],
            p.x * m[8] + p.y * m[9] + p.z * m[10] + m[11]};
}
__device__ __forceinline__ V3 xform_vec(const float *m, V3 v) {
  return V3{v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
            v.x * m[8] + v.y * m[9] + v.z * m[10]};
}
__device__ __forceinline__ V3 xform_transpose(const float *m, V3 v) {
  return V3{v.x * m[0] + v.y * m[4] + v.z * m[8], v.x * m[1] + v.y * m[5] + v.z * m[9],
            v.x * m[2] + v.y * m[6] + v.z * m[10]};
}

struct Frame {
  V3 o, x, y, z;
  __device__ __forceinline__ V3 to(V3 v) const { return V3{dot(v, x), dot(v, y), dot(v, z)}; }
  __device__ __forceinline__ V3 from(V3 v) const {
    return V3{x.x * v.x + y.x * v.y + z.x * v.z, x.y * v.x + y.y * v.y + z.y * v.z,
              x.z * v.x + y.z * v.y + z.z * v.z};
  }
};

// ---------------------------------------------------------------------------
// Counter-based sample stream (same definition as the CPU oracle's;
// splitmix64 finaliser, cycle-walking permutation).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
#define SGR_GOLD 0x9e3779b97f4a7c15ULL
__device__ __forceinline__ uint64_t pixel_key(uint64_t seed, int px, int py) {
  uint64_t p = ((uint64_t)(uint32_t)px << 32) | (uint64_t)(uint32_t)py;
  return mix64(seed ^ mix64(p + SGR_GOLD));
}
__device__ __forceinline__ uint64_t path_key(uint64_t pk, uint32_t s) {
  return mix64(pk ^ (((uint64_t)s + 1) * 0xd1b54a32d192ed03ULL));
}
__device__ __forceinline__ uint64_t chan_key(uint64_t pk, uint32_t chan) {
  return mix64(pk + ((uint64_t)chan + 1) * 0xa0761d6478bd642fULL);
}
__device__ __forceinline__ float draw(uint64_t key, uint32_t d) {
  uint32_t r = (uint32_t)(mix64(key + ((uint64_t)d + 1) * SGR_GOLD) >> 32);
  return (float)(r >> 8) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ uint32_t permute(uint32_t i, uint32_t n, uint64_t key) {
  if (n <= 1) return 0;
  uint32_t w = n - 1;
  w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
  uint32_t mask = w;
  unsigned bits = 32 - __clz(w);
  unsigned hs = (bits + 1) / 2;
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  do {
    i ^= k0 & mask;
    i = (i * 0x9e3779b1u) & mask;
    i ^= i >> hs;
    i = (i * (k1 | 1u)) & mask;
    i ^= i >> hs;
    i = (i + (k1 >> 16)) & mask;
  } while (i >= n);
  return i;
}
struct Grid {
  uint32_t u, v, total;
};
// Stratified UV of channel CHAN, entry E (render/grid.cc:23-60 jittered
// grid, shuffled per pixel and channel).
__device__ __forceinline__ void strat_uv(uint64_t pk, uint32_t chan, uint32_t e, Grid g,
                                         float &u, float &v) {
  uint64_t ck = chan_key(pk, chan);
  uint32_t cell = permute(e, g.total, ck);
  uint32_t j = cell % g.u, i = cell / g.u;
  float us = 1.0f / (float)g.u, vs = 1.0f / (float)g.v;
  u = clamp01f((float)j * us + draw(ck, 2 * e) * us);
  v = clamp01f((float)i * vs + draw(ck, 2 * e + 1) * vs);
}

// Colours (color/color.h; comparisons by component sum, :241-264).
struct C3 {
  float r, g, b;
};
__device__ __forceinline__ C3 c3(float v) { return C3{v, v, v}; }
__device__ __forceinline__ C3 c3(float a, float b, float c) { return C3{a, b, c}; }
__device__ __forceinline__ C3 operator+(C3 a, C3 b) { return C3{a.r + b.r, a.g + b.g, a.b + b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, C3 b) { return C3{a.r * b.r, a.g * b.g, a.b * b.b}; }
__device__ __forceinline__ C3 operator*(C3 a, float s) { return C3{a.r * s, a.g * s, a.b * s}; }
__device__ __forceinline__ float csum(C3 a) { float s = a.r; s += a.g; s += a.b; return s; }
__device__ __forceinline__ bool cgt0(C3 a) { return csum(a) > 0.0f; }
__device__ __forceinline__ bool czero(C3 a) { return a.r == 0.0f && a.g == 0.0f && a.b == 0.0f; }
__device__ __forceinline__ C3 cload(const float *p) { return C3{p[0], p[1], p[2]}; }

// material/fresnel.h:62-130
__device__ __forceinline__ float fresnel_refl(float n, float k, float cosr) {
  float c = cosr;
  c = c < 1.0f ? c : 1.0f;
  c = c > -1.0f ? c : -1.0f;
  c = fabsf(c);
  float ang = acosf(c);
  float Rs, Rp;
  if (k == 0.0f) {
    float st = sinf(ang) / n;
    st = st < 1.0f ? st : 1.0f;
    st = st > -1.0f ? st : -1.0f;
    float ct = sqrtf(1.0f - st * st);
    float nc1 = n * c, nc2 = n * ct;
    float Fs = (nc1 - ct) / (nc1 + ct);
    float Fp = (c - nc2) / (c + nc2);
    Rs = Fs * Fs;
    Rp = Fp * Fp;
  } else {
    float n2k2 = n * n * k * k, n2mk2 = n * n - k * k;
    float sr = sinf(ang);
    float a0 = n2mk2 - sr * sr;
    float stn = sr * tanf(ang);
    float common = sqrtf(a0 * a0 + 4.0f * n2k2);
    float a2 = (common + a0) / 2.0f, b2 = (common - a0) / 2.0f;
    float apb = a2 + b2, a = sqrtf(a2);
    float rs1 = apb + c * c, rs2 = 2.0f * a * c;
    Rs = (rs1 - rs2) / (rs1 + rs2);
    float rp1 = apb + stn * stn, rp2 = 2.0f * a * stn;
    Rp = Rs * ((rp1 - rp2) / (rp1 + rp2));
  }
  return clamp01f((Rs + Rp) / 2.0f);
}
// Fresnel(n1, Ior(n2, k2)).reflectance
__device__ __forceinline__ float fresnel(float n1, float n2, float k2, float c) {
  return fresnel_refl(n2 / n1, k2 / n1, c);
}

}  // namespace sgr
