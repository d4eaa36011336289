// Device shading: intersection frames, BSDFs, light samplers and the
// envmap, in the reference's fp32 arithmetic.
#pragma once
#include "dev_trace.h"

namespace sgr {

#define SGR_PIf 3.14159265358979323846f
#define SGR_INV_PIf 0.31830988618379067154f
#define SGR_PI_D 3.14159265358979323846

enum : unsigned {  // material/bsdf.h:46-70
  F_REFLECTIVE = 0x10, F_TRANSMISSIVE = 0x20, F_ALL_DIRECTIONS = 0x30,
  F_SPECULAR = 0x01, F_GLOSSY = 0x02, F_DIFFUSE = 0x04, F_ALL = 0x37
};

// Intersection state (render/intersect.cc:86-136).
struct Isec {
  Frame nf, gf;
  V3 v, gn;
  bool back;
  int mat, bmat;
  float media_ior, encl_ior;
};

// Mesh::Part::Triangle::IsecInfo::make_frame (surface/mesh.cc:563-600):
// tangent from NORM x axis, else from the direction to the mesh centre,
// else any perpendicular.
__device__ __forceinline__ Frame mesh_frame(const DevMesh &M, V3 point, V3 norm) {
  V3 axis = v3(M.axis[0], M.axis[1], M.axis[2]);
  V3 s = cross(norm, axis);
  if (len2(s) < 1e-15f) {
    V3 cent = v3(M.center[0], M.center[1], M.center[2]) - point;
    s = cross(norm, cent);
    if (len2(s) < 1e-15f) {  // geometry/vec.h:119-123 perpendicular
      float nz = (norm.x == norm.z) ? norm.z : -norm.z;
      s = v3(norm.y * norm.y + norm.x * nz, -nz * nz - norm.x * norm.y,
             norm.x * norm.x - norm.y * nz);
    }
  }
  s = unit(s);
  V3 t = cross(s, norm);
  return Frame{point, s, t, norm};
}

// A local primitive's local-to-world linear part applied to a vector.
__device__ __forceinline__ V3 l2w_vec(const float4 *lr, V3 v) {
  return V3{v.x * lr[3].x + v.y * lr[3].y + v.z * lr[3].z,
            v.x * lr[4].x + v.y * lr[4].y + v.z * lr[4].z,
            v.x * lr[5].x + v.y * lr[5].y + v.z * lr[5].z};
}

// Normal frame of a local primitive hit (surface/cylinder.cc:61-80,
// surface/sphere2.cc:50-75): the normal by the inverse transpose, the
// second tangent t the local axis (cylinder) or onorm x z, x at the poles
// (sphere2), by the transform, s = normal x t.
__device__ __forceinline__ Frame local_frame(const float4 *lr, RayF ray, V3 point) {
  const V3 onorm = local_onorm(lr, ray, ray.t1);
  const V3 norm = unit(w2l_transpose(lr, onorm));
  V3 ot;
  if (__float_as_uint(lr[6].x) == SGR_REF_CYLINDER)
    ot = v3(0.0f, 0.0f, 1.0f);
  else
    ot = (fabsf(onorm.x) < 1e-15f && fabsf(onorm.y) < 1e-15f) ? v3(1.0f, 0.0f, 0.0f)
                                                               : cross(onorm, v3(0.0f, 0.0f, 1.0f));
  const V3 t = unit(l2w_vec(lr, ot));
  return Frame{point, cross(norm, t), t, norm};
}

// Hit frames: surface/mesh.cc:563-709, surface/tripar.cc:45-66,
// surface/sphere.cc make_intersect; then Intersect::finish_init.
// EXT: the scene may hold local primitives (the general kernels).
template<bool EXT>
__device__ __forceinline__ void make_isec_local(const SceneView &S, uint32_t p, const RayF &ray,
                                                float hu, float hv, float media_ior,
                                                float encl_ior, Isec &is) {
  V3 point = ray.o + ray.d * ray.t1;
  Frame nf, gf;
  if (p < S.n_tri_like) {
    V3 e1 = ld3(S.tri[3 * p + 1]), e2 = ld3(S.tri[3 * p + 2]);
    if (p < S.n_tri) {
      const DevMesh &M = S.meshes[S.prim_mesh[p]];
      V3 raw = M.left_handed ? cross(e2, e1) : cross(e1, e2);
      gf = mesh_frame(M, point, unit(raw));
      if (M.has_normals) {
        // surface/mesh.cc:620-641: shading normal interpolated from the
        // vertex normals; the geometric normal takes its side
        const V3 n0 = ld3(S.tri_vn[3 * p]), n1 = ld3(S.tri_vn[3 * p + 1]),
                 n2 = ld3(S.tri_vn[3 * p + 2]);
        V3 norm = n0 * (1.0f - hu - hv);
        norm = norm + n1 * hu;
        norm = norm + n2 * hv;
        norm = unit(norm);
        if (gf.to(norm).z < 0.0f) gf.z = -gf.z;
        nf = mesh_frame(M, point, norm);
      } else {
        nf = gf;
      }
    } else {
      // tripar (surface/tripar.cc:45-66); an ellipse's frame
      // (surface/ellipse.cc:48-75: s = radius1 / |radius1|, normal
      // unit(edge2 x edge1)) is the same bits, its radii being the stored
      // edges halved
      float e1_len = vlen(e1), inv = 1.0f / e1_len;
      V3 norm = unit(cross(e2, e1));
      V3 s = e1 * inv;
      V3 t = cross(s, norm);
      nf = Frame{point, s, t, norm};
      gf = nf;
    }
  } else if (EXT && p >= S.n_sph_end) {
    nf = local_frame(S.loc + SGR_LOC_STRIDE * (p - S.n_sph_end), ray, point);
    gf = nf;
  } else {
    float4 sp = S.sph[p - S.n_tri_like];
    V3 norm = unit(point - ld3(sp));
    V3 s = cross(norm, v3(0.0f, 0.0f, 1.0f));
    if (len2(s) < 1e-15f) {
      float nz = (norm.x == norm.z) ? norm.z : -norm.z;
      s = v3(norm.y * norm.y + norm.x * nz, -nz * nz - norm.x * norm.y,
             norm.x * norm.x - norm.y * nz);
    }
    s = unit(s);
    V3 t = cross(s, norm);
    nf = Frame{point, s, t, norm};
    gf = nf;
  }
  V3 wv = -unit(ray.d);
  is.v = nf.to(wv);
  V3 gv = gf.to(wv);
  is.back = gv.z < 0.0f;
  if (is.back) {
    is.v.z = -is.v.z;
    nf.z = -nf.z;
  }
  is.gn = nf.to(gf.z);
  is.gn.z = fabsf(is.gn.z);
  is.nf = nf;
  is.gf = gf;
  is.mat = (int)S.prim_mat[p];
  is.bmat = S.mats[is.mat].bsdf;
  is.media_ior = media_ior;
  is.encl_ior = encl_ior;
}

// The hit's Isec; INST (SGR_MISS: none) names the instance whose model
// holds primitive P: the intersection is made in model space with the
// model-space ray, then its normal frame is taken to world space --
// origin and tangents by the transform, the normal by the inverse
// transpose, each renormalised (Instance::Renderable::IsecInfo::
// make_intersect, surface/instance.cc:87-103).
template<bool EXT>
__device__ __forceinline__ void make_isec(const SceneView &S, uint32_t p, uint32_t inst,
                                          const RayF &ray, float hu, float hv, float media_ior,
                                          float encl_ior, Isec &is) {
  if (inst == 0xffffffffu) {
    make_isec_local<EXT>(S, p, ray, hu, hv, media_ior, encl_ior, is);
    return;
  }
  const DevInstance &I = S.insts[inst];
  RayF r2;
  r2.o = xform_pos(I.w2l, ray.o);
  r2.d = xform_vec(I.w2l, ray.d);
  r2.t0 = ray.t0;
  r2.t1 = ray.t1;
  make_isec_local<EXT>(S.models[I.model], p, r2, hu, hv, media_ior, encl_ior, is);
  is.nf.o = xform_pos(I.l2w, is.nf.o);
  is.nf.x = unit(xform_vec(I.l2w, is.nf.x));
  is.nf.y = unit(xform_vec(I.l2w, is.nf.y));
  is.nf.z = unit(xform_transpose(I.w2l, is.nf.z));
}

__device__ __forceinline__ float cos_geom(const Isec &is, V3 d) { return dot(d, is.gn); }
__device__ __forceinline__ float cos_v(const Isec &is, V3 d) { return fmn(dot(is.v, d), 1.0f); }

// geometry/spherical-coords.h:155-168 via material/dist.h:28-31
__device__ __forceinline__ V3 z_symm_vec(float cos_theta, float lng) {
  float sin_lat = cos_theta;
  float cos_lat = sqrtf(1.0f - sin_lat * sin_lat);
  float cos_lng = cosf(lng);
  float sin_lng = sqrtf(1.0f - cos_lng * cos_lng);
  if (lng < 0.0f || lng > (float)SGR_PI_D) sin_lng = -sin_lng;
  return v3(cos_lng * cos_lat, -sin_lng * cos_lat, sin_lat);
}
__device__ __forceinline__ V3 z_normal_symm_vec(float cos_theta, float v) {
  return z_symm_vec(cos_theta, v * 2.0f * SGR_PIf);
}

// material/ward-dist.h:43-76
__device__ __forceinline__ float ward_pdf(float m, float cos_theta) {
  if (cos_theta == 0.0f) return 0.0f;
  float inv_m2 = m == 0.0f ? 0.0f : 1.0f / (m * m);
  float ic = 1.0f / cos_theta, ic2 = ic * ic, ic3 = ic2 * ic;
  float tan2 = ic2 - 1.0f;
  return inv_m2 * SGR_INV_PIf * ic3 * expf(-tan2 * inv_m2);
}

struct BSample {
  C3 val;
  float pdf;
  V3 dir;
  unsigned flags;
};
struct BValue {
  C3 val;
  float pdf;
};

__device__ __forceinline__ BSample bsample_none() {
  BSample s;
  s.val = c3(0.0f);
  s.pdf = 0.0f;
  s.dir = v3(0.0f, 0.0f, 0.0f);
  s.flags = 0;
  return s;
}

// material/lambert.cc:38-62
__device__ __forceinline__ BSample lambert_sample(const Isec &is, C3 col, float u, float v,
                                                  unsigned flags) {
  BSample s = bsample_none();
  if ((flags & (F_REFLECTIVE | F_DIFFUSE)) == (F_REFLECTIVE | F_DIFFUSE)) {
    float ct = sqrtf(u);
    float pdf = ct * SGR_INV_PIf;
    V3 d = z_normal_symm_vec(ct, v);
    if (d.z > 0.0f && cos_geom(is, d) > 0.0f) {
      s.val = col * SGR_INV_PIf;
      s.pdf = pdf;
      s.dir = d;
      s.flags = F_REFLECTIVE | F_DIFFUSE;
    }
  }
  return s;
}

// material/cook-torrance.cc:236-276
__device__ __forceinline__ C3 ct_val(const Isec &is, const DevMaterial &M, V3 l, V3 h,
                                     unsigned flags, float ddw, float nv, float inv_4_nv,
                                     float &pdf) {
  float nl = l.z;
  C3 col = c3(0.0f);
  pdf = 0.0f;
  if (flags & F_DIFFUSE) {
    float diff_pdf = nl * SGR_INV_PIf;
    pdf = pdf + diff_pdf * ddw;
    col = col + cload(M.color) * SGR_INV_PIf;
  }
  if (flags & M.gloss_layer) {
    float nh = h.z;
    float inv_nl = (nl != 0.0f) ? (1.0f / nl) : 0.0f;
    float vh = cos_v(is, h);
    float G = fmn(2.0f * nh * ((nv > nl) ? nl : nv) / vh, 1.0f);
    float D = ward_pdf(M.m, nh);
    float gl = fresnel(is.media_ior, M.ior_n, M.ior_k, vh) * D * G * inv_4_nv * inv_nl;
    float gl_pdf = D / (4.0f * vh);
    pdf = pdf + gl_pdf * (1.0f - ddw);
    col = col + cload(M.gloss) * gl;
  }
  return col;
}

__device__ __forceinline__ BSample ct_sample(const Isec &is, const DevMaterial &M, float u,
                                             float v, unsigned desired) {
  BSample s = bsample_none();
  if (!(desired & F_REFLECTIVE)) return s;
  desired &= M.have_layers;
  if (!desired) return s;
  if (is.v.z < 0.0f) return s;
  float nv = is.v.z, inv_4_nv = (nv != 0.0f) ? 1.0f / (4.0f * nv) : 0.0f;
  float ddw, iddw, idgw;
  if (desired == (F_DIFFUSE | M.gloss_layer)) {
    ddw = M.diff_weight; iddw = M.inv_diff_weight; idgw = M.inv_gloss_weight;
  } else if (desired == F_DIFFUSE) {
    ddw = 1.0f; iddw = 1.0f; idgw = 0.0f;
  } else {
    ddw = 0.0f; iddw = 0.0f; idgw = 1.0f;
  }
  V3 l, h;
  unsigned flags = F_REFLECTIVE;
  if (u < ddw) {
    if (desired != F_DIFFUSE) u = u * iddw;
    l = z_normal_symm_vec(sqrtf(u), v);
    h = unit(is.v + l);
    flags |= F_DIFFUSE;
  } else {
    u = (u - ddw) * idgw;
    float ct = u >= 1.0f ? 0.0f : 1.0f / sqrtf(M.m * M.m * -logf(1.0f - u) + 1.0f);
    h = z_normal_symm_vec(ct, v);
    if (cos_v(is, h) < 0.0f) h = -h;
    l = h * dot(is.v, h) * 2.0f - is.v;
    flags |= M.gloss_layer;
  }
  if (l.z > 1e-15f && cos_geom(is, l) > 1e-15f) {
    float pdf;
    s.val = ct_val(is, M, l, h, desired, ddw, nv, inv_4_nv, pdf);
    s.pdf = pdf;
    s.dir = l;
    s.flags = flags;
  }
  return s;
}

__device__ __forceinline__ BValue ct_eval(const Isec &is, const DevMaterial &M, V3 d,
                                          unsigned flags) {
  BValue r;
  r.val = c3(0.0f);
  r.pdf = 0.0f;
  if ((flags & F_REFLECTIVE) && d.z > 0.0f) {
    flags &= M.have_layers;
    float ddw = (flags == (F_DIFFUSE | M.gloss_layer)) ? M.diff_weight
                : (flags == F_DIFFUSE) ? 1.0f : 0.0f;
    V3 h = unit(is.v + d);
    float nv = is.v.z, inv_4_nv = (nv != 0.0f) ? 1.0f / (4.0f * nv) : 0.0f;
    float pdf;
    r.val = ct_val(is, M, d, h, flags, ddw, nv, inv_4_nv, pdf);
    r.pdf = pdf;
  }
  return r;
}

// Phong (material/phong.cc:38-153, material/phong-dist.h:27-57): a
// diffuse layer (CosDist) chosen with probability intensity(color) and
// a Phong lobe of the half vector, pdf (e + 1) / 2pi cos^e.  val sums
// the layers FLAGS selects; a sample picks its direction whatever FLAGS
// says and reports REFLECTIVE only.
__device__ __forceinline__ float phong_dist_pdf(float e, float cos_theta) {
  return (e + 1.0f) * SGR_INV_PIf * 0.5f * powf(cos_theta, e);
}
__device__ __forceinline__ C3 phong_val(const Isec &is, const DevMaterial &M, V3 l, V3 h,
                                        unsigned flags, float &pdf) {
  const float nl = l.z;
  C3 col = c3(0.0f);
  pdf = 0.0f;
  if (flags & F_DIFFUSE) {
    const float diff_pdf = nl * SGR_INV_PIf;
    pdf = pdf + diff_pdf * M.diff_weight;
    col = col + cload(M.color) * SGR_INV_PIf;
  }
  if (flags & F_GLOSSY) {
    const float nh = h.z;
    const float vh = cos_v(is, h);
    const float spec = phong_dist_pdf(M.m, nh);
    const float spec_pdf = spec / (4.0f * vh);
    pdf = pdf + spec_pdf * (1.0f - M.diff_weight);
    col = col + cload(M.gloss) * spec;
  }
  return col;
}
__device__ __forceinline__ BSample phong_sample(const Isec &is, const DevMaterial &M, float u,
                                                float v, unsigned desired) {
  V3 l, h;
  if (u < M.diff_weight) {
    const float su = u * M.inv_diff_weight;
    l = z_normal_symm_vec(sqrtf(su), v);
    h = unit(is.v + l);
  } else {
    const float su = (u - M.diff_weight) * M.inv_gloss_weight;
    h = z_normal_symm_vec(powf(su, 1.0f / (M.m + 1.0f)), v);
    if (cos_v(is, h) < 0.0f) h = -h;
    l = h * dot(is.v, h) * 2.0f - is.v;
  }
  BSample s = bsample_none();
  s.dir = l;
  s.flags = F_REFLECTIVE;
  if (l.z > 1e-15f && cos_geom(is, l) > 1e-15f) {
    float pdf;
    s.val = phong_val(is, M, l, h, desired, pdf);
    s.pdf = pdf;
  }
  return s;
}
__device__ __forceinline__ BValue phong_eval(const Isec &is, const DevMaterial &M, V3 d,
                                             unsigned flags) {
  BValue r;
  r.val = c3(0.0f);
  r.pdf = 0.0f;
  if ((flags & F_REFLECTIVE) && d.z > 0.0f) {
    const V3 h = unit(is.v + d);
    float pdf;
    r.val = phong_val(is, M, d, h, flags, pdf);
    r.pdf = pdf;
  }
  return r;
}

// geometry/vec.h:166-186
__device__ __forceinline__ V3 refraction(V3 v, V3 n, float ior_in, float ior_out) {
  float ratio = ior_in / ior_out;
  float c1 = -dot(v, n);
  float c2 = 1.0f - (ratio * ratio) * (1.0f - c1 * c1);
  if (c2 < -1e-15f) return v3(0.0f, 0.0f, 0.0f);
  if (c2 < 0.0f) c2 = 0.0f;
  return (v * ratio) + (n * (ratio * c1 - sqrtf(c2)));
}

// material/glass.cc:48-117 (media: innermost ior / enclosing ior)
__device__ __forceinline__ BSample glass_sample(const Isec &is, const DevMaterial &M,
                                                float u, unsigned flags) {
  BSample s = bsample_none();
  if (flags & F_SPECULAR) {
    flags &= F_ALL_DIRECTIONS;
    bool entering = !is.back;
    float old_ior = entering ? is.media_ior : M.ior_n;
    float new_ior = entering ? M.ior_n : is.encl_ior;
    V3 xd = refraction(-is.v, v3(0.0f, 0.0f, 1.0f), old_ior, new_ior);
    float cx = -xd.z;
    float cr = fabsf(is.v.z);
    float xmit = (cx == 0.0f) ? 0.0f : 1.0f - fresnel_refl(old_ior / new_ior, 0.0f, cx);
    float refl = (cr == 0.0f) ? 0.0f : fresnel_refl(new_ior / old_ior, 0.0f, cr);
    if (xmit + refl != 0.0f) {
      float xp = (flags == F_TRANSMISSIVE) ? 1.0f : (flags == F_REFLECTIVE) ? 0.0f
                 : xmit / (xmit + refl);
      if (u < xp) {
        if (cx != 0.0f) xmit = xmit / cx;
        s.val = c3(xmit);
        s.pdf = xp;
        s.dir = xd;
        s.flags = F_SPECULAR | F_TRANSMISSIVE;
      } else {
        if (cr != 0.0f) refl = refl / cr;
        s.val = c3(refl);
        s.pdf = 1.0f - xp;
        s.dir = v3(0.0f, 0.0f, 1.0f) * (is.v.z * 2.0f) - is.v;
        s.flags = F_SPECULAR | F_REFLECTIVE;
      }
    }
  }
  return s;
}

// material/thin-glass.cc:30-70: Fresnel reflection at the view angle,
// the rest times the colour straight through.  The choice compares u
// with the colour of xmit / (xmit + refl) by component sums (Color <
// Color, with Color / Color giving 0 for a zero divisor), and a single
// allowed direction is honoured only when it is TRANSMISSIVE: asked for
// REFLECTIVE alone, the sample still transmits whenever u (0 in a
// recursion) is below that ratio.  Both branches report pdf 1.
__device__ __forceinline__ BSample thin_glass_sample(const Isec &is, const DevMaterial &M,
                                                     float u, unsigned flags) {
  BSample s = bsample_none();
  if (flags & F_SPECULAR) {
    flags &= F_ALL_DIRECTIONS;
    const float refl = fresnel(is.media_ior, M.ior_n, 0.0f, is.v.z);
    const C3 xmit = cload(M.color) * (1.0f - refl);
    const C3 sum = xmit + c3(refl);
    const C3 ratio = c3(sum.r == 0.0f ? 0.0f : xmit.r / sum.r, sum.g == 0.0f ? 0.0f : xmit.g / sum.g,
                        sum.b == 0.0f ? 0.0f : xmit.b / sum.b);
    s.pdf = 1.0f;
    if (flags == F_TRANSMISSIVE || csum(c3(u)) < csum(ratio)) {
      s.val = xmit;
      s.dir = -is.v;
      s.flags = F_SPECULAR | F_TRANSMISSIVE;
    } else {
      s.val = c3(refl);
      s.dir = v3(0.0f, 0.0f, 1.0f) * (is.v.z * 2.0f) - is.v;
      s.flags = F_SPECULAR | F_REFLECTIVE;
    }
  }
  return s;
}

// material/mirror.cc:54-120
__device__ __forceinline__ BSample mirror_sample(const Isec &is, const DevMaterial &M,
                                                 float u, float v, unsigned flags) {
  BSample s = bsample_none();
  if ((flags & (F_SPECULAR | F_REFLECTIVE)) == (F_SPECULAR | F_REFLECTIVE)) {
    float cr = is.v.z;
    if (cr != 0.0f) {
      float fr = fresnel(is.media_ior, M.ior_n, M.ior_k, cr);
      C3 refl = cload(M.gloss) * fr;
      refl = refl * (1.0f / cr);
      if (csum(refl) > 3e-15f && cos_geom(is, is.v) > 0.0f) {
        s.val = refl;
        s.pdf = 1.0f;
        s.dir = v3(0.0f, 0.0f, 1.0f) * (is.v.z * 2.0f) - is.v;
        s.flags = F_SPECULAR | F_REFLECTIVE;
      } else if (M.mirror_under) {
        const DevMaterial &U = (&M)[M.under_off];
        if (!M.has_under) s = lambert_sample(is, cload(M.color), u, v, flags);
        else if (U.type == 1) s = lambert_sample(is, cload(U.color), u, v, flags);
        else if (U.type == 2) s = ct_sample(is, U, u, v, flags);
        else s = phong_sample(is, U, u, v, flags);
        float f2 = fresnel(is.media_ior, M.ior_n, M.ior_k, s.dir.z);
        C3 rr = cload(M.gloss) * f2;
        s.val = s.val * C3{1.0f - rr.r, 1.0f - rr.g, 1.0f - rr.b};
      }
    }
  }
  return s;
}

// Bsdf::supports (material/lambert.cc, cook-torrance.cc:214-223,
// glass.cc:146-152, mirror.cc:111-119, phong.cc, thin-glass.cc): the
// flags of LIMIT the BSDF of M can produce samples for.
__device__ __forceinline__ unsigned bsdf_supports(const DevMaterial &M, unsigned limit) {
  const unsigned lam = (limit & F_REFLECTIVE) ? ((F_REFLECTIVE | F_DIFFUSE) & limit) : 0u;
  switch (M.type) {
    case 1: return lam;
    case 2:
      return ((limit & F_REFLECTIVE) && (limit & M.have_layers))
                 ? ((F_REFLECTIVE | M.have_layers) & limit) : 0u;
    case 3:
    case 7: return (limit & F_SPECULAR) ? ((F_TRANSMISSIVE | F_REFLECTIVE | F_SPECULAR) & limit) : 0u;
    case 4: {
      unsigned f = ((limit & F_REFLECTIVE) && (limit & F_SPECULAR)) ? (F_REFLECTIVE | F_SPECULAR) : 0u;
      if (M.mirror_under) {
        const DevMaterial &U = (&M)[M.under_off];
        if (!M.has_under || U.type == 1) f |= lam;
        else if (U.type == 2)
          f |= ((limit & F_REFLECTIVE) && (limit & U.have_layers))
                   ? ((F_REFLECTIVE | U.have_layers) & limit) : 0u;
        else
          f |= ((limit & F_REFLECTIVE) && (limit & (F_DIFFUSE | F_GLOSSY)))
                   ? ((F_REFLECTIVE | F_DIFFUSE | F_GLOSSY) & limit) : 0u;
      }
      return f;
    }
    case 6:
      return ((limit & F_REFLECTIVE) && (limit & (F_DIFFUSE | F_GLOSSY)))
                 ? ((F_REFLECTIVE | F_DIFFUSE | F_GLOSSY) & limit) : 0u;
  }
  return 0u;
}

// The BSDFs other than the stencil's.
__device__ __forceinline__ BSample bsdf_sample_base(const Isec &is, const DevMaterial &M, float u,
                                                    float v, unsigned flags) {
  switch (M.type) {
    case 1: return lambert_sample(is, cload(M.color), u, v, flags);
    case 2: return ct_sample(is, M, u, v, flags);
    case 3: return glass_sample(is, M, u, flags);
    case 4: return mirror_sample(is, M, u, v, flags);
    case 6: return phong_sample(is, M, u, v, flags);
    case 7: return thin_glass_sample(is, M, u, flags);
  }
  return bsample_none();
}

__device__ __forceinline__ BValue bsdf_eval_base(const Isec &is, const DevMaterial &M, V3 d,
                                                 unsigned flags) {
  BValue r;
  r.val = c3(0.0f);
  r.pdf = 0.0f;
  if (M.type == 1) {
    if ((flags & (F_DIFFUSE | F_REFLECTIVE)) == (F_DIFFUSE | F_REFLECTIVE) && d.z > 0.0f) {
      r.val = cload(M.color) * SGR_INV_PIf;
      r.pdf = d.z * SGR_INV_PIf;
    }
  } else if (M.type == 2) {
    r = ct_eval(is, M, d, flags);
  } else if (M.type == 6) {
    r = phong_eval(is, M, d, flags);
  }
  return r;
}

// Stencil (material/stencil.cc:55-114): the underlying BSDF with
// probability intensity(opacity) -- its value times the opacity, its pdf
// times 1 / intensity -- else straight through with value (1 - opacity)
// / cos and the remaining probability, flagged TRANSLUCENT (0x08).  Asked
// for no specular transmission, the underlying sample comes unscaled.
// M is an entry of the scene's material table (the underlying one is
// found relative to it).
__device__ __forceinline__ BSample stencil_sample(const Isec &is, const DevMaterial &M, float u,
                                                  float v, unsigned flags) {
  const DevMaterial &U = (&M)[M.under_off];
  const bool thru_ok =
      M.op_int < 1.0f && (flags & (F_TRANSMISSIVE | F_SPECULAR)) == (F_TRANSMISSIVE | F_SPECULAR);
  const bool undl_ok = M.op_int > 0.0f && M.has_under && bsdf_supports(U, flags) != 0u;
  if (!thru_ok && !undl_ok) return bsample_none();
  if (!thru_ok) return bsdf_sample_base(is, U, u, v, flags);
  if (!undl_ok || u > M.op_int) {
    const float cn = is.v.z;
    if (cn == 0.0f) return bsample_none();
    BSample s;
    const C3 o = cload(M.color);
    s.val = C3{1.0f - o.r, 1.0f - o.g, 1.0f - o.b} * (1.0f / cn);  // Color / float
    s.pdf = undl_ok ? 1.0f - M.op_int : 1.0f;
    s.dir = -is.v;
    s.flags = F_SPECULAR | F_TRANSMISSIVE | 0x08u;
    return s;
  }
  BSample s = bsdf_sample_base(is, U, u * M.inv_op_int, v, flags);
  s.val = s.val * cload(M.color);
  s.pdf = s.pdf * M.inv_op_int;
  return s;
}

__device__ __forceinline__ BSample bsdf_sample(const Isec &is, const DevMaterial &M, float u,
                                               float v, unsigned flags) {
  if (M.type == 8) return stencil_sample(is, M, u, v, flags);
  return bsdf_sample_base(is, M, u, v, flags);
}

__device__ __forceinline__ BValue bsdf_eval(const Isec &is, const DevMaterial &M, V3 d,
                                            unsigned flags) {
  if (M.type != 8) return bsdf_eval_base(is, M, d, flags);
  BValue r;
  r.val = c3(0.0f);
  r.pdf = 0.0f;
  if (M.has_under) {
    r = bsdf_eval_base(is, (&M)[M.under_off], d, flags);
    r.val = r.val * cload(M.color);
    r.pdf = r.pdf * M.inv_op_int;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Envmap (texture/spheremap.h, texture/matrix-linterp.h,
// geometry/hist-2d-dist.cc, light/envmap-light.cc)
// ---------------------------------------------------------------------------
__device__ __forceinline__ C3 env_texel(const SceneView &S, uint32_t x, uint32_t y) {
  float4 t = S.env_tex[(size_t)y * S.env_w + x];
  return C3{t.x, t.y, t.z};
}
__device__ __forceinline__ C3 env_lookup(const SceneView &S, float uu, float vv) {
  float u = uu - floorf(uu), v = vv - floorf(vv);
  float x = u * (float)S.env_w - 0.5f, y = v * (float)S.env_h - 0.5f;
  float xl = floorf(x), yl = floorf(y);
  float xhf = x - xl, yhf = y - yl;
  float xlf = 1.0f - xhf, ylf = 1.0f - yhf;
  int xi = (int)xl, yi = (int)yl;
  uint32_t xh = (uint32_t)(xi + 1), yh = (uint32_t)(yi + 1);
  if (xi < 0) xi += (int)S.env_w;
  if (yi < 0) yi += (int)S.env_h;
  if (xh >= S.env_w) xh -= S.env_w;
  if (yh >= S.env_h) yh -= S.env_h;
  uint32_t xlo = (uint32_t)xi, ylo = (uint32_t)yi;
  ylo = S.env_h - ylo - 1;
  yh = S.env_h - yh - 1;
  C3 a = env_texel(S, xlo, ylo) * (xlf * ylf);
  C3 b = env_texel(S, xlo, yh) * (xlf * yhf);
  C3 c = env_texel(S, xh, ylo) * (xhf * ylf);
  C3 d = env_texel(S, xh, yh) * (xhf * yhf);
  return ((a + b) + c) + d;
}
__device__ __forceinline__ void latlong(V3 d, float &u, float &v) {
  float lng = atan2f(-d.y, d.x);
  float z = d.z / sqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
  z = z < 1.0f ? z : 1.0f;
  z = z > -1.0f ? z : -1.0f;
  float lat = asinf(z);
  u = clamp01f(lng * SGR_INV_PIf * 0.5f + 0.5f);
  v = clamp01f(lat * SGR_INV_PIf + 0.5f);
}
__device__ __forceinline__ V3 latlong_to_vec(float u, float v) {
  float lng = (u - 0.5f) * SGR_PIf * 2.0f;
  float lat = (v - 0.5f) * SGR_PIf;
  float colat = (float)(SGR_PI_D / 2) - lat;
  return z_symm_vec(cosf(colat), lng);
}
__device__ __forceinline__ V3 env_to(const SceneView &S, V3 w) {
  return v3(w.x * S.env_frame[0] + w.y * S.env_frame[1] + w.z * S.env_frame[2],
            w.x * S.env_frame[3] + w.y * S.env_frame[4] + w.z * S.env_frame[5],
            w.x * S.env_frame[6] + w.y * S.env_frame[7] + w.z * S.env_frame[8]);
}
__device__ __forceinline__ V3 env_from(const SceneView &S, V3 l) {
  return v3(S.env_frame[0] * l.x + S.env_frame[3] * l.y + S.env_frame[6] * l.z,
            S.env_frame[1] * l.x + S.env_frame[4] * l.y + S.env_frame[7] * l.z,
            S.env_frame[2] * l.x + S.env_frame[5] * l.y + S.env_frame[8] * l.z);
}
__device__ __forceinline__ C3 env_map(const SceneView &S, V3 ld) {
  float u, v;
  latlong(ld, u, v);
  return env_lookup(S, u, v);
}
__device__ __forceinline__ float env_bin_pdf(const SceneView &S, uint32_t col, uint32_t row) {
  float rp = S.env_row_cum[row];
  if (row != 0) rp -= S.env_row_cum[row - 1];
  const float *cc = S.env_col_cum + (size_t)row * S.env_hw;
  float cp = cc[col];
  if (col != 0) cp -= cc[col - 1];
  return rp * cp * (float)(S.env_hw * S.env_hh);
}
// std::lower_bound: first index with a[i] >= val.
__device__ __forceinline__ uint32_t lower_bound_f(const float *a, uint32_t n, float val) {
  uint32_t lo = 0, cnt = n;
  while (cnt > 0) {
    uint32_t step = cnt / 2, mid = lo + step;
    if (a[mid] < val) {
      lo = mid + 1;
      cnt -= step + 1;
    } else {
      cnt = step;
    }
  }
  return lo;
}
// lower_bound of VAL in A[N] through a guide table GD of G + 1 entries:
// with g = floor(val * G), g / G <= val < (g + 1) / G, so the answer lies
// in [GD[g], GD[g + 1]] and a search of that range gives the same index.
__device__ __forceinline__ uint32_t guided_lower_bound(const float *a, uint32_t n,
                                                       const uint32_t *gd, uint32_t G, float val) {
  int g = (int)(val * (float)G);
  g = g < 0 ? 0 : (g > (int)G - 1 ? (int)G - 1 : g);
  // the float product can round across a boundary: step to the cell
  // whose bounds hold val exactly as the table was built
  if ((float)g / (float)G > val) g--;
  else if (g + 1 < (int)G && (float)(g + 1) / (float)G <= val) g++;
  uint32_t lo = gd[g], hi = gd[g + 1];
  if (g + 1 == (int)G) hi = n;
  return lo + lower_bound_f(a + lo, hi - lo, val);
}

__device__ __forceinline__ bool env_hist_sample(const SceneView &S, float pu, float pv,
                                                float &ou, float &ov, float &pdf) {
  float u = fmn(pu, 1.0f), v = fmn(pv, 1.0f);
  uint32_t row = guided_lower_bound(S.env_row_cum, S.env_hh, S.env_row_guide, SGR_ENV_GUIDE_ROW, v);
  if (row == S.env_hh) {
    pdf = 0.0f;
    ou = ov = 0.0f;
    return false;
  }
  uint32_t col = guided_lower_bound(S.env_col_cum + (size_t)row * S.env_hw, S.env_hw,
                                    S.env_col_guide + (size_t)row * (SGR_ENV_GUIDE_COL + 1),
                                    SGR_ENV_GUIDE_COL, u);
  pdf = env_bin_pdf(S, col, row);
  float cw = 1.0f / (float)S.env_hw, rh = 1.0f / (float)S.env_hh;
  ou = (float)col * cw + fmodf(pu, cw);
  ov = (float)row * rh + fmodf(pv, rh);
  return true;
}
__device__ __forceinline__ float env_hist_pdf(const SceneView &S, float u, float v) {
  int col = (int)(u * (float)S.env_hw);
  int row = (int)(v * (float)S.env_hh);
  int cmax = (int)S.env_hw - 1, rmax = (int)S.env_hh - 1;
  col = col < cmax ? col : cmax;
  col = col > 0 ? col : 0;
  row = row < rmax ? row : rmax;
  row = row > 0 ? row : 0;
  return env_bin_pdf(S, (uint32_t)col, (uint32_t)row);
}

// Scene::background (render/scene.cc:62-73)
__device__ __forceinline__ C3 background(const SceneView &S, V3 dir) {
  C3 r = c3(0.0f);
  if (S.has_env) r = r + env_map(S, env_to(S, dir));
  return r;
}

// ---------------------------------------------------------------------------
// Light samplers
// ---------------------------------------------------------------------------
struct LSample {
  C3 val;
  float pdf;
  V3 dir;
  float dist;
};

// surface/surface-sampler.cc:26-46
__device__ __forceinline__ bool angular(V3 pos, V3 normal, float apdf, V3 vp, float &dist,
                                        V3 &dir, float &pdf) {
  V3 vv = pos - vp;
  dist = vlen(vv);
  pdf = 0.0f;
  if (dist > 0.0f) {
    float inv = 1.0f / dist;
    dir = vv * inv;
    float adj = dot(-normal, dir);
    if (adj > 0.0f) {
      float dwda = adj * (inv * inv);
      pdf = apdf / dwda;
    }
  }
  return pdf > 0.0f;
}

// light/sphere-light-sampler.cc:40-48
__device__ __forceinline__ float sphere_solid_angle(V3 lc, float radius) {
  float dist = vlen(lc);
  if (dist < radius) return 4.0f * SGR_PIf;
  return 2.0f * SGR_PIf * (1.0f - cosf(asinf(radius / dist)));
}

// Tripar area sample at (u, v) (surface/tripar.cc:210-233).
__device__ __forceinline__ void tripar_area(V3 v0, V3 e1, V3 e2, bool par, float u, float v,
                                            V3 &pos, V3 &norm, float &apdf) {
  norm = cross(e2, e1);
  float area = vlen(norm);
  norm = norm * (1.0f / area);
  if (!par && u + v > 1.0f) {
    u = 1.0f - u;
    v = 1.0f - v;
  }
  if (!par) area = area * 0.5f;
  pos = (v0 + e1 * u) + e2 * v;
  apdf = 1.0f / area;
}

// disk_sample (geometry/disk-sample.h:27-71): concentric map of (pu, pv)
// to the disk of RADIUS.
__device__ __forceinline__ void disk_sample(float radius, float pu, float pv, float &dx,
                                            float &dy) {
  const float u = 2.0f * pu - 1.0f, v = 2.0f * pv - 1.0f;
  if (u == 0.0f && v == 0.0f) {
    dx = dy = 0.0f;
    return;
  }
  float r, theta;
  if (u >= -v) {
    if (u > v) {
      r = u;
      theta = (v > 0.0f) ? v / r : 8.0f + v / r;
    } else {
      r = v;
      theta = 2.0f - u / r;
    }
  } else {
    if (u <= v) {
      r = -u;
      theta = 4.0f - v / r;
    } else {
      r = -v;
      theta = 6.0f + u / r;
    }
  }
  theta = theta * (SGR_PIf / 2.0f);
  dx = radius * (r * cosf(theta));
  dy = radius * (r * sinf(theta));
}

// Ellipse::Sampler::sample (surface/ellipse.cc:218-229): a disk sample
// on the bounding parallelogram (corner A, edges E1, E2), pdf 1 / area.
__device__ __forceinline__ void ellipse_area(V3 a, V3 e1, V3 e2, float pu, float pv, V3 &pos,
                                             V3 &norm, float &apdf) {
  float dx, dy;
  disk_sample(0.5f, pu, pv, dx, dy);
  dx = dx + 0.5f;
  dy = dy + 0.5f;
  pos = (a + e1 * dx) + e2 * dy;
  const V3 n = cross(e2, e1);
  norm = unit(n);
  apdf = (4.0f * SGR_INV_PIf) / vlen(n);
}

// sphere_sample / sphere_sample_inverse (geometry/sphere-sample.h,
// geometry/cone-sample.h:27-62 with cos_half_angle -1).
__device__ __forceinline__ V3 sphere_sample(float pu, float pv) {
  const float z = -1.0f + pv * 2.0f;
  const float r = sqrtf(fabsf(1.0f - z * z));
  const float phi = pu * 2.0f * SGR_PIf;
  return v3(r * cosf(phi), r * sinf(phi), z);
}
__device__ __forceinline__ void sphere_sample_inverse(V3 d, float &pu, float &pv) {
  const float phi = atan2f(d.y, -d.x);
  pu = clamp01f(0.5f - (phi * SGR_INV_PIf * 0.5f));
  pv = clamp01f((d.z + 1.0f) / 2.0f);
}

// A local primitive's local-to-world rows applied to a position / vector.
__device__ __forceinline__ V3 l2w_pos(const float4 *lr, V3 p) {
  return V3{p.x * lr[3].x + p.y * lr[3].y + p.z * lr[3].z + lr[3].w,
            p.x * lr[4].x + p.y * lr[4].y + p.z * lr[4].z + lr[4].w,
            p.x * lr[5].x + p.y * lr[5].y + p.z * lr[5].z + lr[5].w};
}
__device__ __forceinline__ V3 w2l_pos(const float4 *lr, V3 p) {
  return V3{p.x * lr[0].x + p.y * lr[0].y + p.z * lr[0].z + lr[0].w,
            p.x * lr[1].x + p.y * lr[1].y + p.z * lr[1].z + lr[1].w,
            p.x * lr[2].x + p.y * lr[2].y + p.z * lr[2].z + lr[2].w};
}

// Surface::Sampler::sample_with_approx_area_pdf (surface/surface-sampler.h:
// 141-152) for a cylinder (CYL) or a sphere2: the area pdf from the
// parallelogram spanned by the positions at (pu + delta, pv) and
// (pu, pv + delta).
__device__ __forceinline__ V3 local_sample_pos(const float4 *lr, bool cyl, float pu, float pv) {
  if (cyl) {
    const float th = pu * 2.0f * SGR_PIf;
    return l2w_pos(lr, v3(cosf(th), sinf(th), 2.0f * pv - 1.0f));
  }
  return l2w_pos(lr, sphere_sample(pu, pv));
}
__device__ __forceinline__ float approx_area_pdf(const float4 *lr, bool cyl, float pu, float pv,
                                                 V3 &pos) {
  const float delta = 0.0001f;
  pos = local_sample_pos(lr, cyl, pu, pv);
  const V3 pdu = local_sample_pos(lr, cyl, pu + delta, pv);
  const V3 pdv = local_sample_pos(lr, cyl, pu, pv + delta);
  const float sample_area = vlen(cross(pdu - pos, pdv - pos));
  const float param_area = delta * delta;
  return sample_area == 0.0f ? 0.0f : param_area / sample_area;
}

// Cylinder::Sampler::sample_from_viewpoint (surface/cylinder.cc:239-306)
// and Sphere2::Sampler::sample_from_viewpoint (surface/sphere2.cc:
// 238-270): a sample facing VP (a cylinder's mirrored through its axis, a
// sphere2's through its centre, when it faces away), pdf doubled.
__device__ __forceinline__ void local_from_viewpoint(const float4 *lr, V3 vp, float pu, float pv,
                                                     V3 &pos, V3 &norm, float &apdf) {
  if (__float_as_uint(lr[6].x) == SGR_REF_CYLINDER) {
    const float theta = pu * 2.0f * SGR_PIf;
    norm = unit(w2l_transpose(lr, v3(cosf(theta), sinf(theta), 0.0f)));
    apdf = approx_area_pdf(lr, true, pu, pv, pos);
    if (dot(norm, pos - vp) > 0.0f) {
      V3 op = w2l_pos(lr, pos);
      op.x = -op.x;
      op.y = -op.y;
      pos = l2w_pos(lr, op);
      norm = -norm;
    }
    apdf = apdf * 2.0f;
    return;
  }
  const V3 ov = w2l_pos(lr, vp);
  V3 onorm = sphere_sample(pu, pv);
  float su = pu, sv = pv;
  if (dot(onorm, ov) < 0.0f) {
    onorm = -onorm;
    sphere_sample_inverse(onorm, su, sv);
  }
  norm = unit(w2l_transpose(lr, onorm));
  apdf = approx_area_pdf(lr, false, su, sv, pos) * 2.0f;
}

// SurfaceLightSampler::sample (light/surface-light-sampler.cc:47-64) of
// an ellipse, cylinder or sphere2 light from the frame's origin.
__device__ __forceinline__ LSample surface_light_sample(const float4 *tri,
                                                                 const float4 *loc,
                                                                 uint32_t n_sph_end, DevLight L,
                                                                 Frame nf, float u, float v) {
  LSample o;
  o.val = c3(0.0f);
  o.pdf = 0.0f;
  o.dir = v3(0.0f, 0.0f, 0.0f);
  o.dist = 0.0f;
  V3 pos, norm;
  float apdf;
  if (L.kind == 3) {
    const float4 a = tri[3 * L.prim], b = tri[3 * L.prim + 1], c = tri[3 * L.prim + 2];
    ellipse_area(ld3(a), ld3(b), ld3(c), u, v, pos, norm, apdf);
  } else {
    local_from_viewpoint(loc + SGR_LOC_STRIDE * (L.prim - n_sph_end), nf.o, u, v, pos, norm, apdf);
  }
  float dist, pdf;
  V3 dir;
  if (angular(pos, norm, apdf, nf.o, dist, dir, pdf)) {
    V3 d = nf.to(dir);
    if (d.z > 0.0f) {
      o.val = cload(L.intensity);
      o.pdf = pdf;
      o.dir = d;
      o.dist = dist;
    }
  }
  return o;
}

__device__ __forceinline__ LSample surface_light_eval(const float4 *tri,
                                                               const float4 *loc,
                                                               uint32_t n_sph_end, DevLight L,
                                                               Frame nf, V3 dir);

// PointLight::intensity (light/point-light.h:64-79): full inside the
// core cone, ramped down as d^4 across the fringe.
__device__ __forceinline__ C3 point_intensity(C3 col, float cos_half, float cos_core,
                                              float cos_dir) {
  if (cos_dir < cos_core) {
    const float d = (cos_dir - cos_half) / (cos_core - cos_half);
    col = col * (d * d * d * d);
  }
  return col;
}

// EXT: the scene may hold ellipse, cylinder, sphere2 and point lights
// (the general kernels).
template<bool EXT>
__device__ __forceinline__ LSample light_sample(const SceneView &S, const DevLight &L,
                                                const Isec &is, float u, float v,
                                                float extra) {
  LSample o;
  o.val = c3(0.0f);
  o.pdf = 0.0f;
  o.dir = v3(0.0f, 0.0f, 0.0f);
  o.dist = 0.0f;
  if (L.kind == 0) {  // light/surface-light-sampler.cc:47-64
    float4 a = S.tri[3 * L.prim], b = S.tri[3 * L.prim + 1], c = S.tri[3 * L.prim + 2];
    V3 pos, norm;
    float apdf, dist, pdf;
    V3 dir;
    tripar_area(ld3(a), ld3(b), ld3(c), a.w != 0.0f, u, v, pos, norm, apdf);
    if (angular(pos, norm, apdf, is.nf.o, dist, dir, pdf)) {
      V3 d = is.nf.to(dir);
      if (d.z > 0.0f) {
        o.val = cload(L.intensity);
        o.pdf = pdf;
        o.dir = d;
        o.dist = dist;
      }
    }
  } else if (L.kind == 1) {  // light/sphere-light-sampler.cc:53-92
    float4 sp = S.sph[L.prim - S.n_tri_like];
    float radius = sp.w;
    V3 lc = is.nf.to(ld3(sp) - is.nf.o);
    if (lc.z >= -radius) {
      float pdf = 1.0f / sphere_solid_angle(lc, radius);
      float rs = radius * sqrtf(u);
      float ang = v * 2.0f * SGR_PIf;
      float x = rs * cosf(ang), y = rs * sinf(ang);
      float z = sqrtf(fabsf(radius * radius - x * x - y * y)) * sinf(SGR_PIf * (extra - 0.5f));
      V3 sv = lc + v3(x, y, z);
      if (sv.z > 0.0f && cos_geom(is, sv) > 0.0f) {
        V3 sd = unit(sv);
        float dist;
        if (sphere_hit(v3(0.0f, 0.0f, 0.0f), radius, -lc, sd, 0.0f, dist)) {
          o.val = cload(L.intensity);
          o.pdf = pdf;
          o.dir = sd;
          o.dist = dist;
        }
      }
    }
  } else if (EXT && L.kind == 6) {  // PointLight::Sampler::sample (light/point-light.cc:79-111)
    const float4 *pl = S.plights + 4 * L.prim;
    const V3 lp = ld3(pl[0]);
    const V3 lvec = is.nf.to(lp - is.nf.o);
    if (lvec.z > 0.0f && cos_geom(is, lvec) > 0.0f) {
      const float dist = vlen(lvec);
      const float inv = 1.0f / dist;
      const float cos_dir = dot((is.nf.o - lp) * inv, ld3(pl[3]));
      if (cos_dir >= pl[0].w) {
        o.val = point_intensity(cload(L.intensity), pl[0].w, pl[1].w, cos_dir) * inv * inv;
        o.pdf = 1.0f;
        o.dir = lvec * inv;
        o.dist = dist;
      }
    }
  } else if (EXT && L.kind >= 3) {
    o = surface_light_sample(S.tri, S.loc, S.n_sph_end, L, is.nf, u, v);
  } else {  // light/envmap-light.cc:100-125
    float mu, mv, pdf;
    env_hist_sample(S, u, v, mu, mv, pdf);
    V3 ldir = latlong_to_vec(mu, mv);
    V3 wd = env_from(S, ldir);
    V3 d = is.nf.to(wd);
    if (d.z < 0.0f || cos_geom(is, d) < 0.0f) return o;
    o.val = env_map(S, ldir);
    o.pdf = pdf * (0.25f * SGR_INV_PIf);
    o.dir = d;
    o.dist = 0.0f;
  }
  return o;
}

template<bool EXT>
__device__ __forceinline__ LSample light_eval(const SceneView &S, const DevLight &L,
                                              const Isec &is, V3 dir) {
  LSample o;
  o.val = c3(0.0f);
  o.pdf = 0.0f;
  o.dir = dir;
  o.dist = 0.0f;
  if (L.kind == 0) {  // light/surface-light-sampler.cc:86-100
    float4 a = S.tri[3 * L.prim], b = S.tri[3 * L.prim + 1], c = S.tri[3 * L.prim + 2];
    bool par = a.w != 0.0f;
    V3 wdir = is.nf.from(dir);
    RayF r;
    r.o = is.nf.o;
    r.d = wdir;
    r.t0 = 0.0f;
    r.t1 = __builtin_inff();
    float t, uu, vv;
    if (tripar_hit(ld3(a), ld3(b), ld3(c), par, r, t, uu, vv)) {
      V3 pos, norm, sd;
      float apdf, dist, pdf;
      tripar_area(ld3(a), ld3(b), ld3(c), par, uu, vv, pos, norm, apdf);
      if (angular(pos, norm, apdf, is.nf.o, dist, sd, pdf)) {
        o.val = cload(L.intensity);
        o.pdf = pdf;
        o.dist = dist;
      }
    }
  } else if (L.kind == 1) {  // light/sphere-light-sampler.cc:130-152
    float4 sp = S.sph[L.prim - S.n_tri_like];
    V3 lc = is.nf.to(ld3(sp) - is.nf.o);
    if (lc.z >= -sp.w) {
      float dist;
      if (sphere_hit(v3(0.0f, 0.0f, 0.0f), sp.w, -lc, dir, 0.0f, dist)) {
        o.val = cload(L.intensity);
        o.pdf = 1.0f / sphere_solid_angle(lc, sp.w);
        o.dist = dist;
      }
    }
  } else if (EXT && L.kind == 6) {
    // PointLight::Sampler::eval (light/point-light.cc:130-133): no direction
    // meets a point
  } else if (EXT && L.kind >= 3) {
    o = surface_light_eval(S.tri, S.loc, S.n_sph_end, L, is.nf, dir);
  } else {  // light/envmap-light.cc:150-170
    V3 wd = is.nf.from(dir);
    V3 ld = env_to(S, wd);
    float mu, mv;
    latlong(ld, mu, mv);
    o.val = env_lookup(S, mu, mv);
    o.pdf = env_hist_pdf(S, mu, mv) * (0.25f * SGR_INV_PIf);
  }
  return o;
}

// SurfaceLightSampler::eval (light/surface-light-sampler.cc:86-100) of an
// ellipse, cylinder or sphere2 light: its eval_from_viewpoint from the
// frame's origin in direction DIR (frame coordinates).
__device__ __forceinline__ LSample surface_light_eval(const float4 *tri,
                                                               const float4 *loc,
                                                               uint32_t n_sph_end, DevLight L,
                                                               Frame nf, V3 dir) {
  LSample o;
  o.val = c3(0.0f);
  o.pdf = 0.0f;
  o.dir = dir;
  o.dist = 0.0f;
  {
    const V3 vp = nf.o, wdir = nf.from(dir);
    V3 pos, norm;
    float apdf = 0.0f;
    bool hit = false;
    if (L.kind == 3) {  // surface/ellipse.cc:233-246
      const float4 a = tri[3 * L.prim], b = tri[3 * L.prim + 1], c = tri[3 * L.prim + 2];
      RayF r;
      r.o = vp;
      r.d = wdir;
      r.t0 = 0.0f;
      r.t1 = __builtin_inff();
      float t, uu, vv;
      if (ellipse_hit(ld3(a), ld3(b), ld3(c), r, t, uu, vv)) {
        pos = vp + wdir * t;
        const V3 n = cross(ld3(c), ld3(b));
        norm = unit(n);
        apdf = (4.0f * SGR_INV_PIf) / vlen(n);
        hit = true;
      }
    } else {  // surface/cylinder.cc:322-347, surface/sphere2.cc:278-298
      const float4 *lr = loc + SGR_LOC_STRIDE * (L.prim - n_sph_end);
      V3 ov, od;
      local_ray(lr, vp, wdir, ov, od);
      float t;
      if (L.kind == 4) {
        if (cylinder_hit(ov, od, 0.0f, t)) {
          const V3 lp = ov + od * t;
          float pu = atan2f(lp.y, lp.x) * SGR_INV_PIf * 0.5f;
          if (pu < 0.0f) pu += 1.0f;
          const float pv = lp.z * 0.5f + 0.5f;
          local_from_viewpoint(lr, vp, clamp01f(pu), clamp01f(pv), pos, norm, apdf);
          hit = true;
        }
      } else if (sphere_hit(v3(0.0f, 0.0f, 0.0f), 1.0f, ov, od, 0.0f, t)) {
        float pu, pv;
        sphere_sample_inverse(ov + od * t, pu, pv);
        local_from_viewpoint(lr, vp, pu, pv, pos, norm, apdf);
        hit = true;
      }
    }
    float dist, pdf;
    V3 sd;
    if (hit && angular(pos, norm, apdf, vp, dist, sd, pdf)) {
      o.val = cload(L.intensity);
      o.pdf = pdf;
      o.dist = dist;
    }
  }
  return o;
}

// render/mis-sample-weight.h:45-77
__device__ __forceinline__ float mis_weight(float pdf, float other) {
  float term = 1.0f * pdf, ot = 1.0f * other;
  if (term > 1e10f) return 1.0f;
  if (ot > 1e10f) return 0.0f;
  return (term * pdf) / (term * term + ot * ot);
}

// The two terms of DirectIllum::sample_light (render/direct-illum.cc:
// 126-255): a sample of light L weighted against the BSDF's pdf, and a
// sample of the BSDF weighted against the light's, with the power
// heuristic; for a point light the light sample unweighted and no BSDF
// sample (render/direct-illum.cc:174, 202).  A term that can contribute goes to USE(c, ray): its value
// if unoccluded and the shadow ray that decides it (from the hit point
// to the light sample, or to the scene horizon for an envmap).
__device__ __forceinline__ RayF direct_shadow_ray(const SceneView &S, const Isec &is, V3 dir,
                                                  float dist) {
  const float md = dist == 0.0f ? S.horizon : dist;
  RayF r;
  r.o = is.nf.o;
  r.d = is.nf.from(dir);
  r.t0 = S.min_trace;
  r.t1 = md - S.min_trace;
  return r;
}
template<bool EXT, typename Use>
__device__ __forceinline__ void direct_light_term(const SceneView &S, const DevLight &L,
                                                  const Isec &is, const DevMaterial &M,
                                                  float lu, float lv, float extra, Use use) {
  const LSample ls = light_sample<EXT>(S, L, is, lu, lv, extra);
  if (ls.pdf > 0.0f && cgt0(ls.val)) {
    const BValue bval = bsdf_eval(is, M, ls.dir, F_ALL & ~F_SPECULAR);
    if (cgt0(bval.val)) {
      C3 c = ls.val;
      if (!(EXT && L.kind == 6)) c = c * mis_weight(ls.pdf, bval.pdf);
      c = c * bval.val;
      c = c * fabsf(ls.dir.z);
      c = c * (1.0f / ls.pdf);
      use(c, direct_shadow_ray(S, is, ls.dir, ls.dist));
    }
  }
}
template<bool EXT, typename Use>
__device__ __forceinline__ void direct_bsdf_term(const SceneView &S, const DevLight &L,
                                                 const Isec &is, const DevMaterial &M, float bu,
                                                 float bv, Use use) {
  if (EXT && L.kind == 6) return;
  const BSample bs = bsdf_sample(is, M, bu, bv, F_ALL & ~F_SPECULAR);
  if (bs.pdf > 0.0f && cgt0(bs.val)) {
    const LSample lval = light_eval<EXT>(S, L, is, bs.dir);
    if (lval.pdf > 0.0f && cgt0(lval.val)) {
      C3 c = lval.val;
      c = c * mis_weight(bs.pdf, lval.pdf);
      c = c * bs.val;
      c = c * fabsf(bs.dir.z);
      c = c * (1.0f / bs.pdf);
      use(c, direct_shadow_ray(S, is, bs.dir, lval.dist));
    }
  }
}

}  // namespace sgr
