// ORACLE -- test infrastructure only.  Nothing in the product links this.
//
// CPU restatement of the reference's render path: scene set-up, intersect
// frames, BSDFs, light samplers, the PathInteg bounce loop with
// DirectIllum MIS, eye rays and the Mitchell splat.  Templated on R
// (float = the reference's fp32 arithmetic, double = FP64 oracle).
#pragma once
#include "orc_geom.h"
#include "orc_rng.h"
#include "../include/snoggpu.h"
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

namespace orc {

// color/color.h:35-279.  Comparisons use component sums (:241-264).
template<class R> struct C3 {
  R r, g, b;
  C3() : r(0), g(0), b(0) {}
  C3(R v) : r(v), g(v), b(v) {}
  C3(R a, R c, R d) : r(a), g(c), b(d) {}
  C3 operator+(const C3 &o) const { return C3(r + o.r, g + o.g, b + o.b); }
  C3 operator-(const C3 &o) const { return C3(r - o.r, g - o.g, b - o.b); }
  C3 operator*(const C3 &o) const { return C3(r * o.r, g * o.g, b * o.b); }
  C3 operator*(R s) const { return C3(r * s, g * s, b * s); }
  // color/color.h:151-158: division multiplies by the reciprocal.
  C3 operator/(R d) const { return operator*(R(1) / d); }
  C3 &operator+=(const C3 &o) { r += o.r; g += o.g; b += o.b; return *this; }
  C3 &operator*=(const C3 &o) { r *= o.r; g *= o.g; b *= o.b; return *this; }
  C3 &operator*=(R s) { r *= s; g *= s; b *= s; return *this; }
  C3 &operator/=(R d) { return operator*=(R(1) / d); }
  R sum() const { R s = r; s += g; s += b; return s; }
  R intensity() const { return sum() / R(3); }        // :160-166
  bool gt0() const { return sum() > 0; }              // Color > 0
  bool is_zero() const { return r == 0 && g == 0 && b == 0; }  // == 0
  R min_component() const {  // :191-197
    R m = r;
    m = mn(g, m);
    m = mn(b, m);
    return m;
  }
};

enum {  // material/bsdf.h:46-70
  REFLECTIVE = 0x10, TRANSMISSIVE = 0x20, ALL_DIRECTIONS = 0x30,
  SPECULAR = 0x01, GLOSSY = 0x02, DIFFUSE = 0x04, ALL_LAYERS = 0x07,
  ALL = 0x37
};

template<class R> struct BsdfSample {
  C3<R> val; R pdf = 0; V3<R> dir; unsigned flags = 0;
};
template<class R> struct BsdfValue { C3<R> val; R pdf = 0; };

template<class R> struct Material {
  int type;
  C3<R> color, gloss;
  R m, ior_n, ior_k;
  int underlying;
};

// Intersection state (render/intersect.h:30-160).
template<class R> struct Isec {
  Frame<R> nf, gf;
  V3<R> v, geom_n;
  bool back;
  int mat;        // material providing Le (GLOW) or the BSDF
  int bsdf_mat;   // material providing the BSDF, -1 = none
  R media_ior;    // innermost medium (isec.media.medium.ior)
  R enclosing_ior;
  R cos_n(const V3<R> &d) const { return d.z; }
  R cos_geom_n(const V3<R> &d) const { return dot(d, geom_n); }
  R cos_v(const V3<R> &d) const { return mn(dot(v, d), R(1)); }  // :61-64
};

// material/fresnel.h:42-177
template<class R> struct Fresnel {
  R n, k, n2k2, n2_m_k2;
  Fresnel(R n1, R n2) : n(n2 / n1), k(0), n2k2(0), n2_m_k2(n * n) {}
  Fresnel(R n1, R ior2n, R ior2k)
      : n(ior2n / n1), k(ior2k / n1), n2k2(n * n * k * k),
        n2_m_k2(n * n - k * k) {}
  R reflectance(R cos_refl_angle) const {
    cos_refl_angle = std::fabs(clampv(cos_refl_angle, R(-1), R(1)));
    R refl_angle = std::acos(cos_refl_angle);
    R Rp, Rs;
    if (k == 0) {
      R sin_trans = clampv(std::sin(refl_angle) / n, R(-1), R(1));
      R cos_trans = std::sqrt(R(1) - sin_trans * sin_trans);
      R nc1 = n * cos_refl_angle;
      R nc2 = n * cos_trans;
      R Fs = (nc1 - cos_trans) / (nc1 + cos_trans);
      R Fp = (cos_refl_angle - nc2) / (cos_refl_angle + nc2);
      Rs = Fs * Fs;
      Rp = Fp * Fp;
    } else {
      R sin_refl = std::sin(refl_angle);
      R a0 = n2_m_k2 - sin_refl * sin_refl;
      R sin_tan = sin_refl * std::tan(refl_angle);
      R common = std::sqrt(a0 * a0 + R(4) * n2k2);
      R a2 = (common + a0) / R(2);
      R b2 = (common - a0) / R(2);
      R a2pb2 = a2 + b2;
      R a = std::sqrt(a2);
      R rs1 = a2pb2 + cos_refl_angle * cos_refl_angle;
      R rs2 = R(2) * a * cos_refl_angle;
      Rs = (rs1 - rs2) / (rs1 + rs2);
      R rp1 = a2pb2 + sin_tan * sin_tan;
      R rp2 = R(2) * a * sin_tan;
      Rp = Rs * ((rp1 - rp2) / (rp1 + rp2));
    }
    return clamp01((Rs + Rp) / R(2));
  }
};

// geometry/spherical-coords.h:155-168 / material/dist.h:28-31
template<class R>
static inline V3<R> z_normal_symm_vec(R cos_theta, R v) {
  R lng = v * R(2) * R(PIf);
  R sin_lat = cos_theta;
  R cos_lat = std::sqrt(R(1) - sin_lat * sin_lat);
  R cos_lng = std::cos(lng);
  R sin_lng = std::sqrt(R(1) - cos_lng * cos_lng);
  if (lng < 0 || lng > R(PI_D)) sin_lng = -sin_lng;
  return V3<R>(cos_lng * cos_lat, -sin_lng * cos_lat, sin_lat);
}

// material/ward-dist.h:22-81
template<class R> struct WardDist {
  R m, inv_m2;
  explicit WardDist(R m_) : m(m_), inv_m2(m_ == 0 ? R(0) : R(1) / (m_ * m_)) {}
  V3<R> sample(R u, R v) const {
    R cos_theta = u >= 1 ? R(0) : R(1) / std::sqrt(m * m * -std::log(R(1) - u) + R(1));
    return z_normal_symm_vec(cos_theta, v);
  }
  R pdf(R cos_theta) const {
    if (cos_theta == 0) return 0;
    R ic = R(1) / cos_theta;
    R ic2 = ic * ic;
    R ic3 = ic2 * ic;
    R tan2 = ic2 - R(1);
    return inv_m2 * R(INV_PIf) * ic3 * std::exp(-tan2 * inv_m2);
  }
};

// One BSDF instance at an intersection.
template<class R> struct Bsdf {
  int type = 0;  // 0 = none
  const Isec<R> *isec = nullptr;
  C3<R> color, gloss;
  // Cook-Torrance (material/cook-torrance.cc:38-70)
  R m = 0, diff_weight = 0, inv_diff_weight = 0, inv_gloss_weight = 0;
  R nv = 0, inv_4_nv = 0;
  unsigned gloss_layer = 0, have_layers = 0;
  R fr_n1 = 1, fr_n = 1, fr_k = 0;  // Fresnel(n1, Ior(n, k))
  // Glass (material/glass.cc:32-42)
  R old_ior = 1, new_ior = 1;
  bool has_under = false;  // Mirror's underlying Lambert
  // Stencil (material/stencil.cc:30-52): opacity, its intensity and
  // reciprocal (snapped to 0 / 1 near the ends), the underlying BSDF
  std::unique_ptr<Bsdf> under;
  C3<R> opacity;
  R op_int = 0, inv_op_int = 0;

  R fresnel_ct(R c) const { return Fresnel<R>(fr_n1, fr_n, fr_k).reflectance(c); }

  void init(const Material<R> &mt, const Isec<R> &is, R glass_medium_ior) {
    isec = &is;
    type = mt.type;
    if (type == SGR_MAT_LAMBERT) {
      color = mt.color;
    } else if (type == SGR_MAT_COOK_TORRANCE) {
      m = mt.m;
      color = mt.color;
      gloss = mt.gloss;
      R di = color.intensity(), gi = gloss.intensity();
      diff_weight = (di + gi) == 0 ? R(0) : di / (di + gi);
      inv_diff_weight = diff_weight == 0 ? R(0) : R(1) / diff_weight;
      inv_gloss_weight = diff_weight == 1 ? R(0) : R(1) / (R(1) - diff_weight);
      fr_n1 = is.media_ior;
      fr_n = mt.ior_n;
      fr_k = mt.ior_k;
      nv = is.cos_n(is.v);
      inv_4_nv = (nv != 0) ? R(1) / (R(4) * nv) : R(0);
      gloss_layer = m < R(0.5) ? GLOSSY : DIFFUSE;
      have_layers = (diff_weight > 0 ? DIFFUSE : 0) | (diff_weight < 1 ? gloss_layer : 0);
    } else if (type == SGR_MAT_GLASS) {
      bool entering = !is.back;
      old_ior = entering ? is.media_ior : glass_medium_ior;
      new_ior = entering ? glass_medium_ior : is.enclosing_ior;
    } else if (type == SGR_MAT_THIN_GLASS) {  // material/thin-glass.cc:30-46
      color = mt.color;
      fr_n1 = is.media_ior;
      fr_n = mt.ior_n;
    } else if (type == SGR_MAT_PHONG) {  // material/phong.cc:38-45
      color = mt.color;
      gloss = mt.gloss;
      m = mt.m;
      diff_weight = color.intensity();
      inv_diff_weight = diff_weight == 0 ? R(0) : R(1) / diff_weight;
      inv_gloss_weight = diff_weight == 1 ? R(0) : R(1) / (R(1) - diff_weight);
    } else if (type == SGR_MAT_MIRROR) {
      gloss = mt.gloss;  // reflectance
      color = mt.color;  // underlying Lambert colour
      fr_n1 = is.media_ior;
      fr_n = mt.ior_n;
      fr_k = mt.ior_k;
      // material/mirror.cc:28-33: no underlying material if the colour is
      // below Eps (Color < Color(Eps) compares component sums).
      has_under = !(color.sum() < R(3e-15));
    }
  }

  // --- Lambert (material/lambert.cc:38-62)
  BsdfSample<R> lambert_sample(R u, R v, unsigned flags, const C3<R> &col) const {
    BsdfSample<R> s;
    if ((flags & (REFLECTIVE | DIFFUSE)) == (REFLECTIVE | DIFFUSE)) {
      R cos_theta = std::sqrt(u);
      R pdf = cos_theta * R(INV_PIf);
      V3<R> dir = z_normal_symm_vec(cos_theta, v);
      if (isec->cos_n(dir) > 0 && isec->cos_geom_n(dir) > 0) {
        s.val = col * R(INV_PIf);
        s.pdf = pdf;
        s.dir = dir;
        s.flags = REFLECTIVE | DIFFUSE;
      }
    }
    return s;
  }
  BsdfValue<R> lambert_eval(const V3<R> &dir, unsigned flags) const {
    BsdfValue<R> r;
    R cn = isec->cos_n(dir);
    if ((flags & (DIFFUSE | REFLECTIVE)) == (DIFFUSE | REFLECTIVE) && cn > 0) {
      r.val = color * R(INV_PIf);
      r.pdf = cn * R(INV_PIf);
    }
    return r;
  }

  // --- Cook-Torrance (material/cook-torrance.cc:236-276)
  C3<R> ct_val(const V3<R> &l, const V3<R> &h, unsigned flags, R ddw, R &pdf) const {
    R nl = isec->cos_n(l);
    C3<R> col(0);
    pdf = 0;
    if (flags & DIFFUSE) {
      R diff = R(INV_PIf);
      R diff_pdf = nl * R(INV_PIf);
      pdf += diff_pdf * ddw;
      col += color * diff;
    }
    if (flags & gloss_layer) {
      WardDist<R> wd(m);
      R nh = isec->cos_n(h);
      R inv_nl = (nl != 0) ? (R(1) / nl) : R(0);
      R vh = isec->cos_v(h);
      R G = mn(R(2) * nh * ((nv > nl) ? nl : nv) / vh, R(1));
      R gl = fresnel_ct(vh) * wd.pdf(nh) * G * inv_4_nv * inv_nl;
      R gl_pdf = wd.pdf(nh) / (R(4) * vh);
      pdf += gl_pdf * (R(1) - ddw);
      col += gloss * gl;
    }
    return col;
  }
  BsdfSample<R> ct_sample(R u, R v, unsigned desired) const {  // :76-160
    BsdfSample<R> s;
    V3<R> l, h;
    unsigned flags = REFLECTIVE;
    if (!(desired & REFLECTIVE)) return s;
    desired &= have_layers;
    if (!desired) return s;
    if (isec->v.z < 0) return s;
    R ddw, iddw, idgw;
    if (desired == (DIFFUSE | gloss_layer)) {
      ddw = diff_weight; iddw = inv_diff_weight; idgw = inv_gloss_weight;
    } else if (desired == DIFFUSE) {
      ddw = 1; iddw = 1; idgw = 0;
    } else {
      ddw = 0; iddw = 0; idgw = 1;
    }
    if (u < ddw) {
      if (desired != DIFFUSE) u = u * iddw;
      l = z_normal_symm_vec(std::sqrt(u), v);
      h = (isec->v + l).unit();
      flags |= DIFFUSE;
    } else {
      u = (u - ddw) * idgw;
      h = WardDist<R>(m).sample(u, v);
      if (isec->cos_v(h) < 0) h = -h;
      l = mirror(isec->v, h);
      flags |= gloss_layer;
    }
    if (isec->cos_n(l) > R(float(1e-15)) && isec->cos_geom_n(l) > R(float(1e-15))) {
      R pdf;
      s.val = ct_val(l, h, desired, ddw, pdf);
      s.pdf = pdf;
      s.dir = l;
      s.flags = flags;
    }
    return s;
  }
  BsdfValue<R> ct_eval(const V3<R> &dir, unsigned flags) const {  // :165-190
    BsdfValue<R> r;
    R cn = isec->cos_n(dir);
    if ((flags & REFLECTIVE) && cn > 0) {
      flags &= have_layers;
      R ddw = (flags == (DIFFUSE | gloss_layer)) ? diff_weight
              : (flags == DIFFUSE) ? R(1) : R(0);
      V3<R> h = (isec->v + dir).unit();
      R pdf;
      r.val = ct_val(dir, h, flags, ddw, pdf);
      r.pdf = pdf;
    }
    return r;
  }

  // --- Glass (material/glass.cc:48-117)
  BsdfSample<R> glass_sample(R u, R, unsigned flags) const {
    BsdfSample<R> s;
    if (flags & SPECULAR) {
      flags &= ALL_DIRECTIONS;
      V3<R> xmit_dir = refraction(-isec->v, V3<R>(0, 0, 1), old_ior, new_ior);
      R cos_xmit = -isec->cos_n(xmit_dir);
      R cos_refl = std::fabs(isec->cos_n(isec->v));
      R xmit = (cos_xmit == 0) ? R(0)
               : R(1) - Fresnel<R>(new_ior, old_ior).reflectance(cos_xmit);
      R refl = (cos_refl == 0) ? R(0) : Fresnel<R>(old_ior, new_ior).reflectance(cos_refl);
      if (xmit + refl != 0) {
        R xp = (flags == TRANSMISSIVE) ? R(1) : (flags == REFLECTIVE) ? R(0)
               : xmit / (xmit + refl);
        if (u < xp) {
          if (cos_xmit != 0) xmit /= cos_xmit;
          s.val = C3<R>(xmit);
          s.pdf = xp;
          s.dir = xmit_dir;
          s.flags = SPECULAR | TRANSMISSIVE;
        } else {
          if (cos_refl != 0) refl /= cos_refl;
          s.val = C3<R>(refl);
          s.pdf = R(1) - xp;
          s.dir = mirror(isec->v, V3<R>(0, 0, 1));
          s.flags = SPECULAR | REFLECTIVE;
        }
      }
    }
    return s;
  }

  // --- Mirror (material/mirror.cc:54-120)
  BsdfSample<R> mirror_sample(R u, R v, unsigned flags) const {
    BsdfSample<R> s;
    if ((flags & (SPECULAR | REFLECTIVE)) == (SPECULAR | REFLECTIVE)) {
      R cos_refl = isec->cos_n(isec->v);
      if (cos_refl != 0) {
        Fresnel<R> fr(fr_n1, fr_n, fr_k);
        C3<R> refl = gloss * fr.reflectance(cos_refl) / cos_refl;
        if (refl.sum() > R(3e-15) && isec->cos_geom_n(isec->v) > 0) {
          s.val = refl;
          s.pdf = 1;
          s.dir = mirror(isec->v, V3<R>(0, 0, 1));
          s.flags = SPECULAR | REFLECTIVE;
        } else if (has_under) {  // the underlying material's BSDF, or a Lambert
          s = under ? under->sample(u, v, flags) : lambert_sample(u, v, flags, color);
          R fres_refl = fr.reflectance(isec->cos_n(s.dir));
          C3<R> rr = gloss * fres_refl;
          s.val *= (C3<R>(1) - rr);
        }
      }
    }
    return s;
  }

  // --- Phong (material/phong.cc:38-153, material/phong-dist.h:27-57)
  R phong_pdf(R c) const { return (m + R(1)) * R(INV_PIf) * R(0.5f) * std::pow(c, m); }
  C3<R> phong_val(const V3<R> &l, const V3<R> &h, R &pdf, unsigned flags) const {
    R nl = isec->cos_n(l);
    C3<R> col(0);
    pdf = 0;
    if (flags & DIFFUSE) {
      R diff = R(INV_PIf);
      R diff_pdf = nl * R(INV_PIf);
      pdf += diff_pdf * diff_weight;
      col += color * diff;
    }
    if (flags & GLOSSY) {
      R nh = isec->cos_n(h);
      R vh = isec->cos_v(h);
      R spec = phong_pdf(nh);
      R spec_pdf = spec / (R(4) * vh);
      pdf += spec_pdf * (R(1) - diff_weight);
      col += gloss * spec;
    }
    return col;
  }
  BsdfSample<R> phong_sample(R u, R v, unsigned desired) const {
    V3<R> l, h;
    if (u < diff_weight) {
      R su = u * inv_diff_weight;
      l = z_normal_symm_vec(std::sqrt(su), v);
      h = (isec->v + l).unit();
    } else {
      R su = (u - diff_weight) * inv_gloss_weight;
      h = z_normal_symm_vec(std::pow(su, R(1) / (m + R(1))), v);
      if (isec->cos_v(h) < 0) h = -h;
      l = mirror(isec->v, h);
    }
    BsdfSample<R> s;
    s.dir = l;
    s.flags = REFLECTIVE;
    if (isec->cos_n(l) > R(float(1e-15)) && isec->cos_geom_n(l) > R(float(1e-15))) {
      R pdf;
      s.val = phong_val(l, h, pdf, desired);
      s.pdf = pdf;
    }
    return s;
  }
  BsdfValue<R> phong_eval(const V3<R> &dir, unsigned flags) const {
    BsdfValue<R> r;
    if ((flags & REFLECTIVE) && isec->cos_n(dir) > 0) {
      V3<R> h = (isec->v + dir).unit();
      R pdf;
      r.val = phong_val(dir, h, pdf, flags);
      r.pdf = pdf;
    }
    return r;
  }

  // --- Thin glass (material/thin-glass.cc:30-70): Fresnel reflection at
  // the view angle; the rest, times the colour, goes straight through.
  // The choice compares u with the colour of xmit / (xmit + refl) as
  // Color < Color does (component sums), and a single allowed direction
  // is taken only when it is TRANSMISSIVE: asked for REFLECTIVE alone,
  // the sample still transmits whenever u (0 in a recursion) is below
  // that ratio.  Both branches report pdf 1.
  BsdfSample<R> thin_glass_sample(R u, R, unsigned flags) const {
    BsdfSample<R> s;
    if (flags & SPECULAR) {
      flags &= ALL_DIRECTIONS;
      R cos_xmit = isec->cos_n(isec->v);
      R refl = Fresnel<R>(fr_n1, fr_n).reflectance(cos_xmit);
      C3<R> xmit = color * (R(1) - refl);
      C3<R> sum = xmit + C3<R>(refl);  // Color / Color: 0 where the divisor is 0
      C3<R> ratio(sum.r == 0 ? R(0) : xmit.r / sum.r, sum.g == 0 ? R(0) : xmit.g / sum.g,
                  sum.b == 0 ? R(0) : xmit.b / sum.b);
      if (flags == TRANSMISSIVE || C3<R>(u).sum() < ratio.sum()) {
        s.val = xmit;
        s.pdf = 1;
        s.dir = -isec->v;
        s.flags = SPECULAR | TRANSMISSIVE;
      } else {
        s.val = C3<R>(refl);
        s.pdf = 1;
        s.dir = mirror(isec->v, V3<R>(0, 0, 1));
        s.flags = SPECULAR | REFLECTIVE;
      }
    }
    return s;
  }

  // Bsdf::supports (material/lambert.cc, cook-torrance.cc:214-223,
  // glass.cc:146-152, mirror.cc:111-119, phong.cc, thin-glass.cc): the
  // flags of LIMIT this BSDF can produce samples for.
  unsigned supports(unsigned limit) const {
    switch (type) {
      case SGR_MAT_LAMBERT:
        return (limit & REFLECTIVE) ? ((REFLECTIVE | DIFFUSE) & limit) : 0u;
      case SGR_MAT_COOK_TORRANCE:
        return ((limit & REFLECTIVE) && (limit & have_layers)) ? ((REFLECTIVE | have_layers) & limit)
                                                                : 0u;
      case SGR_MAT_GLASS:
      case SGR_MAT_THIN_GLASS:
        return (limit & SPECULAR) ? ((TRANSMISSIVE | REFLECTIVE | SPECULAR) & limit) : 0u;
      case SGR_MAT_MIRROR: {
        unsigned f = 0;
        if ((limit & REFLECTIVE) && (limit & SPECULAR)) f |= REFLECTIVE | SPECULAR;
        if (under) f |= under->supports(limit);
        else if (has_under) f |= (limit & REFLECTIVE) ? ((REFLECTIVE | DIFFUSE) & limit) : 0u;
        return f;
      }
      case SGR_MAT_PHONG:
        return ((limit & REFLECTIVE) && (limit & (DIFFUSE | GLOSSY)))
                   ? ((REFLECTIVE | DIFFUSE | GLOSSY) & limit) : 0u;
      case SGR_MAT_STENCIL: {  // material/stencil.cc:114-123 (not masked by LIMIT)
        unsigned f = 0;
        if (opacity.gt0() && under) f |= under->supports(limit);
        if (opacity.sum() < R(3)) f |= TRANSMISSIVE | SPECULAR;
        return f;
      }
    }
    return 0;
  }

  // --- Stencil (material/stencil.cc:55-114): the underlying BSDF with
  // probability intensity(opacity) (its value times the opacity, its pdf
  // times 1 / intensity), else straight through with value (1 - opacity)
  // / cos and the remaining probability, flagged TRANSLUCENT (0x08).
  BsdfSample<R> stencil_sample(R u, R v, unsigned flags) const {
    bool thru_ok = op_int < 1 && (flags & (TRANSMISSIVE | SPECULAR)) == (TRANSMISSIVE | SPECULAR);
    bool undl_ok = op_int > 0 && under && under->supports(flags) != 0;
    if (!thru_ok && !undl_ok) return BsdfSample<R>();
    if (!thru_ok) return under->sample(u, v, flags);
    if (!undl_ok || u > op_int) {
      R cn = isec->cos_n(isec->v);
      if (cn == 0) return BsdfSample<R>();
      BsdfSample<R> s;
      s.val = (C3<R>(1) - opacity) / cn;
      s.pdf = undl_ok ? R(1) - op_int : R(1);
      s.dir = -isec->v;
      s.flags = SPECULAR | TRANSMISSIVE | 0x08u;
      return s;
    }
    BsdfSample<R> s = under->sample(u * inv_op_int, v, flags);
    s.val *= opacity;
    s.pdf *= inv_op_int;
    return s;
  }

  BsdfSample<R> sample(R u, R v, unsigned flags) const {
    switch (type) {
      case SGR_MAT_LAMBERT: return lambert_sample(u, v, flags, color);
      case SGR_MAT_COOK_TORRANCE: return ct_sample(u, v, flags);
      case SGR_MAT_GLASS: return glass_sample(u, v, flags);
      case SGR_MAT_MIRROR: return mirror_sample(u, v, flags);
      case SGR_MAT_PHONG: return phong_sample(u, v, flags);
      case SGR_MAT_THIN_GLASS: return thin_glass_sample(u, v, flags);
      case SGR_MAT_STENCIL: return stencil_sample(u, v, flags);
    }
    return BsdfSample<R>();
  }
  BsdfValue<R> eval(const V3<R> &dir, unsigned flags) const {
    switch (type) {
      case SGR_MAT_LAMBERT: return lambert_eval(dir, flags);
      case SGR_MAT_COOK_TORRANCE: return ct_eval(dir, flags);
      case SGR_MAT_PHONG: return phong_eval(dir, flags);
      case SGR_MAT_STENCIL:
        if (under) {
          BsdfValue<R> r = under->eval(dir, flags);
          r.val *= opacity;
          r.pdf *= inv_op_int;
          return r;
        }
        return BsdfValue<R>();
    }
    return BsdfValue<R>();  // glass and mirror: specular only
  }
};

// Envmap light data (light/envmap-light.cc, geometry/hist-2d-dist.cc,
// texture/spheremap.{h,cc}, texture/matrix-linterp.h).
template<class R> struct Envmap {
  unsigned w = 0, h = 0;
  std::vector<float> tex;  // rgb, row 0 = top
  Frame<R> frame;
  unsigned hw = 0, hh = 0;
  std::vector<float> row_cum, col_cum;

  C3<R> texel(unsigned x, unsigned y) const {
    const float *p = &tex[3 * ((size_t)y * w + x)];
    return C3<R>(p[0], p[1], p[2]);
  }
  // texture/matrix-linterp.h:30-70 + texture/matrix-tex.tcc:51-67
  C3<R> lookup(float uu, float vv) const {
    float u = uu - std::floor(uu), v = vv - std::floor(vv);
    float x = u * float(w) - 0.5f, y = v * float(h) - 0.5f;
    float x_lo = std::floor(x), y_lo = std::floor(y);
    float x_hi_fr = x - x_lo, y_hi_fr = y - y_lo;
    float x_lo_fr = 1 - x_hi_fr, y_lo_fr = 1 - y_hi_fr;
    int xi_lo_u = int(x_lo), yi_lo_u = int(y_lo);
    unsigned xi_hi = (unsigned)(xi_lo_u + 1), yi_hi = (unsigned)(yi_lo_u + 1);
    if (xi_lo_u < 0) xi_lo_u += (int)w;
    if (yi_lo_u < 0) yi_lo_u += (int)h;
    if (xi_hi >= w) xi_hi -= w;
    if (yi_hi >= h) yi_hi -= h;
    unsigned xi_lo = (unsigned)xi_lo_u, yi_lo = (unsigned)yi_lo_u;
    yi_lo = h - yi_lo - 1;
    yi_hi = h - yi_hi - 1;
    return texel(xi_lo, yi_lo) * R(x_lo_fr * y_lo_fr) +
           texel(xi_lo, yi_hi) * R(x_lo_fr * y_hi_fr) +
           texel(xi_hi, yi_lo) * R(x_hi_fr * y_lo_fr) +
           texel(xi_hi, yi_hi) * R(x_hi_fr * y_hi_fr);
  }
  // geometry/spherical-coords.h:60-142 (z_axis_latlong)
  static void latlong(const V3<R> &d, float &u, float &v) {
    float lng = float(std::atan2(-d.y, d.x));
    float lat = float(std::asin(clampv(d.z / std::sqrt(d.x * d.x + d.y * d.y + d.z * d.z),
                                       R(-1), R(1))));
    u = clamp01(lng * INV_PIf * 0.5f + 0.5f);
    v = clamp01(lat * INV_PIf + 0.5f);
  }
  // geometry/spherical-coords.h:175-205 (z_axis_latlong_to_vec)
  static V3<R> latlong_to_vec(float u, float v) {
    float lng = (u - 0.5f) * PIf * 2;
    float lat = (v - 0.5f) * PIf;
    float colat = float(PI_D / 2) - lat;
    return z_normal_symm_vec_lng(R(std::cos(colat)), R(lng));
  }
  static V3<R> z_normal_symm_vec_lng(R cos_colat, R lng) {
    R sin_lat = cos_colat;
    R cos_lat = std::sqrt(R(1) - sin_lat * sin_lat);
    R cos_lng = std::cos(lng);
    R sin_lng = std::sqrt(R(1) - cos_lng * cos_lng);
    if (lng < 0 || lng > R(PI_D)) sin_lng = -sin_lng;
    return V3<R>(cos_lng * cos_lat, -sin_lng * cos_lat, sin_lat);
  }
  C3<R> map(const V3<R> &d) const {
    float u, v;
    latlong(d, u, v);
    return lookup(u, v);
  }

  void build() {
    // texture/spheremap.cc:35-71 light map (block average, y flipped)
    unsigned div = 1;
    while (w / div > 1024 || h / div > 1024) div++;
    hw = w / div;
    hh = h / div;
    float avg_scale = 1.f / float(div * div);
    std::vector<float> lmap((size_t)hw * hh * 3, 0.f);
    for (unsigned y = 0; y < h; y++)
      for (unsigned x = 0; x < w; x++) {
        unsigned lx = x / div;
        unsigned ly = hh - y / div - 1;
        if (lx >= hw || ly >= hh) continue;
        float *p = &lmap[3 * ((size_t)ly * hw + lx)];
        const float *t = &tex[3 * ((size_t)y * w + x)];
        p[0] = p[0] + t[0] * avg_scale;
        p[1] = p[1] + t[1] * avg_scale;
        p[2] = p[2] + t[2] * avg_scale;
      }
    // light/envmap-light.cc:70-95 histogram with cos(latitude) weight
    std::vector<float> bins((size_t)hw * hh, 0.f);
    double inc = PI_D / hh;
    double lat = -PI_D / 2 + inc / 2;
    for (unsigned r = 0; r < hh; r++) {
      double sc = std::cos(lat);
      for (unsigned c = 0; c < hw; c++) {
        const float *p = &lmap[3 * ((size_t)r * hw + c)];
        float s = p[0]; s += p[1]; s += p[2];
        float inten = s / 3.f;
        double val = double(inten) * sc;
        bins[(size_t)r * hw + c] += float(val);
      }
      lat += inc;
    }
    // geometry/hist-2d-dist.cc:52-110
    row_cum.assign(hh, 0.f);
    col_cum.assign((size_t)hw * hh, 0.f);
    double total = 0;
    for (size_t i = 0; i < bins.size(); i++) total += double(bins[i]);
    double inv_total = total == 0 ? 0 : 1 / total;
    double ns = 0;
    for (unsigned r = 0; r < hh; r++) {
      for (unsigned c = 0; c < hw; c++) ns += double(bins[(size_t)r * hw + c]) * inv_total;
      row_cum[r] = float(ns);
    }
    for (unsigned r = 0; r < hh; r++) {
      double rs = 0;
      for (unsigned c = 0; c < hw; c++) rs += double(bins[(size_t)r * hw + c]);
      double irs = rs == 0 ? 0 : 1 / rs;
      double nrs = 0;
      for (unsigned c = 0; c < hw; c++) {
        nrs += double(bins[(size_t)r * hw + c]) * irs;
        col_cum[(size_t)r * hw + c] = float(nrs);
      }
    }
  }

  float bin_pdf(unsigned col, unsigned row) const {  // hist-2d-dist.cc:245-262
    float rp = row_cum[row];
    if (row != 0) rp -= row_cum[row - 1];
    const float *cc = &col_cum[(size_t)row * hw];
    float cp = cc[col];
    if (col != 0) cp -= cc[col - 1];
    return rp * cp * float(hw * hh);
  }
  // hist-2d-dist.cc:180-233
  bool hist_sample(float pu, float pv, float &ou, float &ov, float &pdf) const {
    float u = mn(pu, 1.f), v = mn(pv, 1.f);
    unsigned row = (unsigned)(std::lower_bound(row_cum.begin(), row_cum.end(), v) - row_cum.begin());
    if (row == hh) { pdf = 0; ou = ov = 0; return false; }
    const float *cc = &col_cum[(size_t)row * hw];
    unsigned col = (unsigned)(std::lower_bound(cc, cc + hw, u) - cc);
    pdf = bin_pdf(col, row);
    float cw = 1.f / float(hw), rh = 1.f / float(hh);
    ou = float(col) * cw + std::fmod(pu, cw);
    ov = float(row) * rh + std::fmod(pv, rh);
    return true;
  }
  float hist_pdf(float u, float v) const {  // hist-2d-dist.cc:265-273
    int col = clampv(int(u * float(hw)), 0, int(hw) - 1);
    int row = clampv(int(v * float(hh)), 0, int(hh) - 1);
    return bin_pdf((unsigned)col, (unsigned)row);
  }
};

template<class R> struct LightSample { C3<R> val; R pdf = 0; V3<R> dir; R dist = 0; };
// Light::Sampler::FreeSample (light/light-sampler.h): a point and a
// direction leaving it, for shooting photons.
template<class R> struct FreeSample { C3<R> val; R pdf = 0; V3<R> pos, dir; };

template<class R> struct Light {
  // 0 = tripar area light, 1 = sphere light, 2 = envmap; surface lights
  // (light/surface-light-sampler.cc) of 3 = an ellipse, 4 = a cylinder,
  // 5 = a sphere2; 6 = a point light (light/point-light.h)
  int kind;
  unsigned prim;
  int env;
  C3<R> intensity;
  // point light: position, frame axes, cosines of the half cone and
  // half core angles (-1: all directions)
  V3<R> pos, fx, fy, fz;
  R cos_half = -1, cos_core = -1;
};

// PointLight::intensity (light/point-light.h:64-79)
template<class R>
static inline C3<R> point_intensity(const Light<R> &L, R cos_dir) {
  C3<R> intens = L.intensity;
  if (cos_dir < L.cos_core) {
    const R d = (cos_dir - L.cos_half) / (L.cos_core - L.cos_half);
    intens *= d * d * d * d;
  }
  return intens;
}

// Inverse of an affine 3 x 4 transform: linear part by cofactors in
// double, rounded to float (the scene's instances keep both).
static inline bool invert_affine(const float *m, float *inv) {
  const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9],
               i = m[10];
  const double A = e * i - f * h, B = f * g - d * i, Cc = d * h - e * g;
  const double det = a * A + b * B + c * Cc;
  if (!(std::isfinite(det) && det != 0.0)) return false;
  const double r[9] = {A / det, (c * h - b * i) / det, (b * f - c * e) / det,
                       B / det, (a * i - c * g) / det, (c * d - a * f) / det,
                       Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
  const double t[3] = {m[3], m[7], m[11]};
  for (int row = 0; row < 3; row++) {
    for (int col = 0; col < 3; col++) inv[4 * row + col] = (float)r[3 * row + col];
    inv[4 * row + 3] = (float)-(r[3 * row] * t[0] + r[3 * row + 1] * t[1] + r[3 * row + 2] * t[2]);
  }
  for (int k = 0; k < 12; k++)
    if (!std::isfinite(inv[k])) return false;
  return true;
}

template<class R> struct Scene {
  std::vector<Prim<R>> prims;
  std::vector<int> prim_mat;
  std::vector<int> prim_mesh;
  struct MeshInfo { V3<R> axis, center; bool left_handed, has_normals; };
  std::vector<V3<R>> tri_vn;  // 3 vertex normals per mesh triangle (meshes with normals)
  std::vector<MeshInfo> meshes;
  std::vector<Material<R>> mats;
  std::vector<Light<R>> lights;
  std::vector<Envmap<R>> envs;
  Accel<R> accel;
  OctreeBuilder<R> builder;
  // two-level instancing: models are scenes of their own (geometry only)
  std::vector<std::unique_ptr<Scene>> models;
  std::vector<typename Accel<R>::Inst> insts;
  BBox<R> bbox;
  R horizon = 0;
  R min_trace = R(1e-3);
  unsigned long num_real = 0;

  int bsdf_material(int m) const {
    int guard = 0;
    while (m >= 0 && mats[m].type == SGR_MAT_GLOW && guard++ < 8) m = mats[m].underlying;
    return m;
  }
  // Material::fully_occluding (material/material.h:98-100): thin glass
  // lets shadow rays through.
  bool partial_mat(int m) const {
    return m >= 0 && (mats[m].type == SGR_MAT_THIN_GLASS || mats[m].type == SGR_MAT_STENCIL);
  }
  bool has_partial = false;

  // IsecInfo::normal of a cylinder or sphere2 hit at T (surface/
  // cylinder.cc:84-90, surface/sphere2.cc:79-83): the local hit point's
  // normal (its x, y for the cylinder) by the inverse transpose, unit.
  static V3<R> local_onorm(const Prim<R> &pr, const Ray<R> &ray, R t) {
    const V3<R> o = xform_pos(pr.w2l, ray.origin), d = xform_vec(pr.w2l, ray.dir);
    const V3<R> ip = o + d * t;
    return pr.kind == 5 ? V3<R>(ip.x, ip.y, R(0)) : ip;
  }
  static V3<R> local_normal(const Prim<R> &pr, const Ray<R> &ray, R t) {
    return xform_transpose(pr.w2l, local_onorm(pr, ray, t)).unit();
  }

  // Material::transmittance of partially occluding surface P met by a
  // shadow ray at T, in a medium of index MEDIUM (material/thin-glass.cc:
  // 123-140): the colour times one minus the Fresnel reflectance at the
  // angle between the surface normal (IsecInfo::normal: the unit
  // geometric normal, surface/mesh.cc:714-718, surface/tripar.cc:78-81,
  // surface/sphere.cc:95-98) and the reversed ray direction.
  C3<R> transmittance(unsigned p, const Ray<R> &ray, R t, R medium) const {
    const Prim<R> &pr = prims[p];
    const Material<R> &m = mats[prim_mat[p]];
    if (m.type == SGR_MAT_STENCIL) {  // material/stencil.cc:160-166
      auto c = [](R x) { return x > R(0) ? x : R(0); };
      return C3<R>(c(R(1) - m.color.r), c(R(1) - m.color.g), c(R(1) - m.color.b));
    }
    V3<R> n = pr.kind == 2 ? (ray.origin + ray.dir * t - pr.center).unit()
              : (pr.kind == 5 || pr.kind == 6) ? local_normal(pr, ray, t)
                                               : cross(pr.e2, pr.e1).unit();
    R c = float(dot(n, -ray.dir));  // cos_angle returns a float
    R refl = Fresnel<R>(medium, m.ior_n).reflectance(c);
    return m.color * (R(1) - refl);
  }

  // Scene::occludes (render/scene.h:68-87 -> space/space.cc:131-194): a
  // fully occluding surface on the ray blocks it; each partially
  // occluding one multiplies TRANS by its transmittance, once however
  // many octree nodes list it (the reference's negative mailbox skips a
  // surface it already met), and blocks the ray once TRANS falls below
  // Eps (Color < Eps: the component sum below 3 Eps).
  struct Occl {
    const Scene *sc;
    const Ray<R> *ray;
    R medium;
    C3<R> trans;
    std::vector<unsigned> seen;
  };
  static int occl_partial(void *c, unsigned p, R t) {
    Occl &o = *(Occl *)c;
    if (!o.sc->partial_mat(o.sc->prim_mat[p])) return 1;
    for (unsigned q : o.seen)
      if (q == p) return 0;
    o.seen.push_back(p);
    o.trans *= o.sc->transmittance(p, *o.ray, t, o.medium);
    return o.trans.sum() < R(3e-15f) ? 1 : 0;
  }
  bool occludes(const Ray<R> &ray, R medium, C3<R> &trans, uint64_t *nc, uint64_t *st) const {
    if (!has_partial) return accel.occluded(ray, nc, st);
    Occl o{this, &ray, medium, trans, {}};
    const bool b = accel.occluded(ray, nc, st, &Scene::occl_partial, &o);
    trans = o.trans;
    return b;
  }

  // Material::get_bsdf of BSDF material BM at IS (a stencil wraps its
  // underlying material's BSDF, or is it when fully opaque:
  // material/stencil.cc:140-148).
  void make_bsdf(Bsdf<R> &b, int bm, const Isec<R> &is) const {
    const Material<R> &m = mats[bm];
    if (m.type == SGR_MAT_MIRROR && m.underlying >= 0) {
      // Mirror (ior, reflectance, underlying material) (material/mirror.h:
      // 37-42): the material's BSDF under the mirror, if it has one
      b.init(m, is, m.ior_n);
      const int ub = bsdf_material(m.underlying);
      b.has_under = ub >= 0;
      if (ub >= 0) {
        b.under.reset(new Bsdf<R>);
        make_bsdf(*b.under, ub, is);
      }
      return;
    }
    if (m.type != SGR_MAT_STENCIL) {
      b.init(m, is, m.ior_n);
      return;
    }
    const int ub = bsdf_material(m.underlying);
    if (m.color.sum() >= R(3)) {  // opac >= 1 (Color >= Color: sums)
      if (ub >= 0) make_bsdf(b, ub, is);
      return;
    }
    b.isec = &is;
    b.type = SGR_MAT_STENCIL;
    b.opacity = m.color;
    b.op_int = m.color.intensity();
    b.inv_op_int = b.op_int == 0 ? R(0) : R(1) / b.op_int;
    if (b.op_int < R(0.001f)) b.op_int = b.inv_op_int = 0;
    else if (b.op_int > R(0.999f)) b.op_int = b.inv_op_int = 1;
    if (ub >= 0) {
      b.under.reset(new Bsdf<R>);
      make_bsdf(*b.under, ub, is);
    }
  }

  // material/material.h medium(): the glass medium, through GLOW wrappers.
  R medium_ior(int m, bool &has) const {
    int b = bsdf_material(m);
    has = b >= 0 && mats[b].type == SGR_MAT_GLASS;  // a stencil has none
    return has ? mats[b].ior_n : R(1);
  }

  void build(const sgr_scene_desc &d) {
    for (unsigned i = 0; i < d.num_materials; i++) {
      const sgr_material &m = d.materials[i];
      if (m.type < SGR_MAT_LAMBERT || m.type > SGR_MAT_STENCIL)
        throw std::runtime_error("bad material type");
      mats.push_back(Material<R>{m.type, C3<R>(m.color[0], m.color[1], m.color[2]),
                                 C3<R>(m.gloss_color[0], m.gloss_color[1], m.gloss_color[2]),
                                 R(m.m), R(m.ior_n), R(m.ior_k), m.underlying});
    }
    for (unsigned i = 0; i < d.num_materials; i++)
      if (mats[i].type == SGR_MAT_STENCIL &&
          (mats[i].underlying < 0 || mats[i].underlying >= (int)d.num_materials ||
           (bsdf_material(mats[i].underlying) >= 0 &&
            mats[bsdf_material(mats[i].underlying)].type == SGR_MAT_STENCIL)))
        throw std::runtime_error("stencil: the underlying material must exist and not be a stencil");
    for (unsigned i = 0; i < d.num_materials; i++) {
      if (mats[i].type != SGR_MAT_MIRROR || mats[i].underlying < 0) continue;
      if (mats[i].underlying >= (int)d.num_materials || mats[i].underlying == (int)i)
        throw std::runtime_error("mirror: bad underlying");
      const int ub = bsdf_material(mats[i].underlying);
      if (ub >= 0 && mats[ub].type != SGR_MAT_LAMBERT && mats[ub].type != SGR_MAT_COOK_TORRANCE &&
          mats[ub].type != SGR_MAT_PHONG)
        throw std::runtime_error("mirror: the underlying material must be a lambert, cook-torrance or phong material");
    }
    std::vector<BBox<R>> pb;
    for (unsigned mi = 0; mi < d.num_meshes; mi++) {
      const sgr_mesh &m = d.meshes[mi];
      BBox<R> mb;
      for (unsigned k = 0; k < m.num_vertices; k++)
        mb.add(V3<R>(m.vertices[3 * k], m.vertices[3 * k + 1], m.vertices[3 * k + 2]));
      // geometry/pos.h:123-126 midpoint
      V3<R> cen((mb.min.x + mb.max.x) / R(2), (mb.min.y + mb.max.y) / R(2),
                (mb.min.z + mb.max.z) / R(2));
      meshes.push_back(MeshInfo{V3<R>(m.axis[0], m.axis[1], m.axis[2]), cen,
                                m.left_handed != 0, m.normals != nullptr});
      bbox.add(mb);
      for (unsigned t = 0; t < m.num_triangles; t++) {
        const uint32_t *ix = &m.triangles[3 * t];
        V3<R> a(m.vertices[3 * ix[0]], m.vertices[3 * ix[0] + 1], m.vertices[3 * ix[0] + 2]);
        V3<R> b(m.vertices[3 * ix[1]], m.vertices[3 * ix[1] + 1], m.vertices[3 * ix[1] + 2]);
        V3<R> c(m.vertices[3 * ix[2]], m.vertices[3 * ix[2] + 1], m.vertices[3 * ix[2] + 2]);
        Prim<R> p;
        p.kind = 0;
        p.parallelogram = false;
        p.v0 = a;
        p.e1 = b - a;  // surface/mesh.cc:735-736
        p.e2 = c - a;
        p.radius = 0;
        prims.push_back(p);
        prim_mat.push_back(m.tri_materials ? (int)m.tri_materials[t] : (int)m.material);
        prim_mesh.push_back((int)mi);
        if (m.normals) {  // surface/mesh.h:144-149 vertex_normals
          tri_vn.resize(3 * (prims.size() - 1));
          for (int c = 0; c < 3; c++) {
            const float *nv = &m.normals[3 * ix[c]];
            tri_vn.push_back(V3<R>(nv[0], nv[1], nv[2]));
          }
        }
        BBox<R> tb(a);
        tb.add(b);
        tb.add(c);
        pb.push_back(tb);
      }
    }
    for (unsigned i = 0; i < d.num_tripars; i++) {
      const sgr_tripar &t = d.tripars[i];
      Prim<R> p;
      p.kind = 1;
      p.parallelogram = t.parallelogram != 0;
      p.v0 = V3<R>(t.v0[0], t.v0[1], t.v0[2]);
      p.e1 = V3<R>(t.e1[0], t.e1[1], t.e1[2]);
      p.e2 = V3<R>(t.e2[0], t.e2[1], t.e2[2]);
      p.radius = 0;
      prims.push_back(p);
      prim_mat.push_back((int)t.material);
      prim_mesh.push_back(-1);
      BBox<R> tb(p.v0);  // surface/tripar.cc:140-148
      tb.add(p.v0 + p.e1);
      tb.add(p.v0 + p.e2);
      if (p.parallelogram) tb.add(p.v0 + p.e1 + p.e2);
      pb.push_back(tb);
      bbox.add(tb);
    }
    // Ellipses (surface/ellipse.h:38-45): the bounding parallelogram
    // corner = center - radius1 - radius2, edges 2 radius1, 2 radius2;
    // box of its four corners (surface/ellipse.cc:171-181).
    for (unsigned i = 0; i < d.num_ellipses; i++) {
      const sgr_ellipse &el = d.ellipses[i];
      const V3<R> c(el.center[0], el.center[1], el.center[2]);
      const V3<R> r1(el.radius1[0], el.radius1[1], el.radius1[2]);
      const V3<R> r2(el.radius2[0], el.radius2[1], el.radius2[2]);
      Prim<R> p;
      p.kind = 4;
      p.parallelogram = true;
      p.v0 = (c - r1) - r2;
      p.e1 = r1 * R(2);
      p.e2 = r2 * R(2);
      p.radius = 0;
      prims.push_back(p);
      prim_mat.push_back((int)el.material);
      prim_mesh.push_back(-1);
      BBox<R> tb(p.v0);
      tb.add(p.v0 + p.e1);
      tb.add(p.v0 + p.e2);
      tb.add(p.v0 + p.e1 + p.e2);
      pb.push_back(tb);
      bbox.add(tb);
    }
    for (unsigned i = 0; i < d.num_spheres; i++) {
      const sgr_sphere &s = d.spheres[i];
      Prim<R> p;
      p.kind = 2;
      p.parallelogram = false;
      p.center = V3<R>(s.center[0], s.center[1], s.center[2]);
      p.radius = R(s.radius);
      prims.push_back(p);
      prim_mat.push_back((int)s.material);
      prim_mesh.push_back(-1);
      BBox<R> sb;  // surface/sphere.cc bbox
      sb.min = V3<R>(p.center.x - p.radius, p.center.y - p.radius, p.center.z - p.radius);
      sb.max = V3<R>(p.center.x + p.radius, p.center.y + p.radius, p.center.z + p.radius);
      pb.push_back(sb);
      bbox.add(sb);
    }
    // Local primitives (surface/local-primitive.h): cylinder and sphere2,
    // a canonical shape placed by a transform; the box is the local cube
    // [-1,1]^3 with its corners taken to world space
    // (geometry/local-xform.cc:27-31, geometry/bbox.cc:25-42).
    for (unsigned i = 0; i < d.num_local_prims; i++) {
      const sgr_local_prim &lp = d.local_prims[i];
      if (lp.kind != SGR_LOCAL_CYLINDER && lp.kind != SGR_LOCAL_SPHERE2)
        throw std::runtime_error("local primitive: bad kind");
      Prim<R> p;
      p.kind = lp.kind == SGR_LOCAL_CYLINDER ? 5 : 6;
      p.parallelogram = false;
      p.radius = 0;
      float w2l[12];
      if (!invert_affine(lp.local_to_world, w2l))
        throw std::runtime_error("local primitive: singular transform");
      for (int k = 0; k < 12; k++) {
        p.l2w[k] = R(lp.local_to_world[k]);
        p.w2l[k] = R(w2l[k]);
      }
      prims.push_back(p);
      prim_mat.push_back((int)lp.material);
      prim_mesh.push_back(-1);
      BBox<R> lb;
      for (int c = 0; c < 8; c++) {
        const V3<R> corner((c & 4) ? R(1) : R(-1), (c & 2) ? R(1) : R(-1), (c & 1) ? R(1) : R(-1));
        const V3<R> w = xform_pos(p.l2w, corner);
        if (c == 0) lb = BBox<R>(w);
        else lb.add(w);
      }
      pb.push_back(lb);
      bbox.add(lb);
    }
    // Models and their instances (surface/model.cc:26-47,
    // surface/instance.cc:145-216): the instance's box is the model's box
    // with its 8 corners taken to world space (geometry/bbox.cc:25-42).
    const unsigned n_direct = (unsigned)prims.size();
    for (unsigned mi = 0; mi < d.num_models; mi++) {
      const sgr_model &m = d.models[mi];
      sgr_scene_desc md;
      std::memset(&md, 0, sizeof md);
      md.meshes = m.meshes; md.num_meshes = m.num_meshes;
      md.tripars = m.tripars; md.num_tripars = m.num_tripars;
      md.spheres = m.spheres; md.num_spheres = m.num_spheres;
      md.ellipses = m.ellipses; md.num_ellipses = m.num_ellipses;
      md.local_prims = m.local_prims; md.num_local_prims = m.num_local_prims;
      md.materials = d.materials; md.num_materials = d.num_materials;
      md.min_trace = d.min_trace;
      md.accel = d.accel;
      models.emplace_back(new Scene);
      models.back()->build(md);
      if (models.back()->has_partial)
        throw std::runtime_error("a model's surfaces must occlude fully (no thin glass)");
    }
    for (unsigned i = 0; i < d.num_instances; i++) {
      const sgr_instance &in = d.instances[i];
      if (in.model >= d.num_models) throw std::runtime_error("instance: bad model");
      typename Accel<R>::Inst I;
      I.model = in.model;
      float w2l[12];
      if (!invert_affine(in.local_to_world, w2l)) throw std::runtime_error("instance: singular transform");
      for (int k = 0; k < 12; k++) {
        I.l2w[k] = R(in.local_to_world[k]);
        I.w2l[k] = R(w2l[k]);
      }
      insts.push_back(I);
      const BBox<R> &mb = models[in.model]->bbox;
      BBox<R> ib;
      for (int c = 0; c < 8; c++) {
        V3<R> corner((c & 4) ? mb.max.x : mb.min.x, (c & 2) ? mb.max.y : mb.min.y,
                     (c & 1) ? mb.max.z : mb.min.z);
        V3<R> w = xform_pos(I.l2w, corner);
        if (c == 0) ib = BBox<R>(w);
        else ib.add(w);
      }
      Prim<R> p;
      p.kind = 3;
      p.parallelogram = false;
      p.radius = 0;
      prims.push_back(p);
      prim_mat.push_back(-1);
      prim_mesh.push_back(-1);
      pb.push_back(ib);
      bbox.add(ib);
    }
    accel.n_prims = n_direct;
    for (unsigned p = 0; p < n_direct; p++) has_partial = has_partial || partial_mat(prim_mat[p]);
    if (!insts.empty()) {
      accel.insts = &insts;
      for (auto &m : models) accel.models.push_back(&m->accel);
    }
    // Acceleration structure; degenerate mesh triangles are not inserted
    // (surface/mesh.cc:832-846).
    accel.prims = &prims;
    accel.use_triv = d.accel == SGR_ACCEL_TRIV;
    for (unsigned p = 0; p < prims.size(); p++) {
      if (prims[p].kind == 0) {
        V3<R> nu = cross(prims[p].e2, prims[p].e1);
        if (!(nu.length_squared() > 0)) continue;
      }
      if (accel.use_triv) accel.triv_order.push_back(p);
      else builder.add(p, pb[p]);
      num_real++;
    }
    if (!accel.use_triv) {
      builder.flatten(accel.nodes, accel.surfs);
      accel.origin = builder.origin;
      accel.size = builder.size;
    }
    horizon = bbox.diameter();  // render/scene.cc:28
    min_trace = R(d.min_trace);
    // Light samplers (render/scene.cc:34; surface/primitive.cc:48-54;
    // surface/sphere.cc add_light_samplers; light/envmap-light.cc:66-68)
    for (unsigned p = 0; p < prims.size(); p++) {
      int m = prim_mat[p];
      if (prims[p].kind != 0 && prims[p].kind != 3 && m >= 0 && mats[m].type == SGR_MAT_GLOW) {
        Light<R> L;
        L.kind = prims[p].kind == 1 ? 0 : prims[p].kind == 2 ? 1 : prims[p].kind - 1;
        L.prim = p;
        L.env = -1;
        L.intensity = mats[m].color;
        lights.push_back(L);
      }
    }
    for (unsigned e = 0; e < d.num_envmaps; e++) {
      const sgr_envmap &em = d.envmaps[e];
      Envmap<R> E;
      E.w = em.width;
      E.h = em.height;
      E.tex.assign(em.texels, em.texels + (size_t)3 * em.width * em.height);
      E.frame = Frame<R>(V3<R>(), V3<R>(em.frame[0], em.frame[1], em.frame[2]),
                         V3<R>(em.frame[3], em.frame[4], em.frame[5]),
                         V3<R>(em.frame[6], em.frame[7], em.frame[8]));
      E.build();
      envs.push_back(E);
      Light<R> L;
      L.kind = 2;
      L.prim = 0;
      L.env = (int)e;
      lights.push_back(L);
    }
    // point lights after the surface lights and the envmap
    // (surface/surface-group.cc:82-94); PointLight (light/point-light.h:
    // 31-47) with Frame (pos, dir) (geometry/frame.h:79-84) for a cone
    for (unsigned i = 0; i < d.num_point_lights; i++) {
      const sgr_point_light &pl = d.point_lights[i];
      Light<R> L;
      L.kind = 6;
      L.prim = i;
      L.env = -1;
      L.intensity = C3<R>(pl.intensity[0], pl.intensity[1], pl.intensity[2]);
      L.pos = V3<R>(pl.pos[0], pl.pos[1], pl.pos[2]);
      L.fx = V3<R>(1, 0, 0);
      L.fy = V3<R>(0, 1, 0);
      L.fz = V3<R>(0, 0, 1);
      if (pl.angle > 0) {
        const Frame<R> f = frame_of(L.pos, V3<R>(pl.dir[0], pl.dir[1], pl.dir[2]));
        L.fx = f.x;
        L.fy = f.y;
        L.fz = f.z;
        L.cos_half = std::cos(R(pl.angle) / 2);
        L.cos_core = std::cos((R(pl.angle) - R(pl.fringe_angle)) / 2);
      }
      lights.push_back(L);
    }
  }

  // Scene::background (render/scene.cc:62-73)
  C3<R> background(const V3<R> &dir) const {
    C3<R> r(0);
    for (const Light<R> &L : lights)
      if (L.kind == 2) {
        const Envmap<R> &E = envs[L.env];
        r += E.map(E.frame.to(dir));
      }
    return r;
  }

  // Frame construction for a hit (surface/mesh.cc:563-709,
  // surface/tripar.cc:45-66, surface/sphere.cc IsecInfo::make_intersect)
  // followed by Intersect::finish_init (render/intersect.cc:86-136).
  // Mesh::Part::Triangle::IsecInfo::make_frame (surface/mesh.cc:563-600)
  static Frame<R> mesh_frame(const MeshInfo &mi, const V3<R> &point, const V3<R> &norm) {
    V3<R> s = cross(norm, mi.axis);
    if (s.length_squared() < R(1e-15)) {
      V3<R> cent = mi.center - point;
      s = cross(norm, cent);
      if (s.length_squared() < R(1e-15)) s = norm.perpendicular();
    }
    s = s.unit();
    V3<R> t = cross(s, norm);
    return Frame<R>(point, s, t, norm);
  }

  // The hit of instance INST (0xffffffff: none): the model's Isec with
  // the model-space ray, then its normal frame to world space -- origin
  // and tangents by the transform, the normal by the inverse transpose,
  // each renormalised (surface/instance.cc:87-103).
  void make_isec(unsigned p, unsigned inst, const Ray<R> &ray, R hu, R hv, R media_ior,
                 R enclosing_ior, Isec<R> &is) const {
    if (inst == 0xffffffffu) {
      make_isec(p, ray, hu, hv, media_ior, enclosing_ior, is);
      return;
    }
    const typename Accel<R>::Inst &I = insts[inst];
    Ray<R> r2 = ray;
    r2.origin = xform_pos(I.w2l, ray.origin);
    r2.dir = xform_vec(I.w2l, ray.dir);
    models[I.model]->make_isec(p, r2, hu, hv, media_ior, enclosing_ior, is);
    is.nf.origin = xform_pos(I.l2w, is.nf.origin);
    is.nf.x = xform_vec(I.l2w, is.nf.x).unit();
    is.nf.y = xform_vec(I.l2w, is.nf.y).unit();
    is.nf.z = xform_transpose(I.w2l, is.nf.z).unit();
  }

  void make_isec(unsigned p, const Ray<R> &ray, R hu, R hv, R media_ior,
                 R enclosing_ior, Isec<R> &is) const {
    const Prim<R> &pr = prims[p];
    V3<R> point = ray.origin + ray.dir * ray.t1;
    Frame<R> nf, gf;
    if (pr.kind == 0) {
      const MeshInfo &mi = meshes[prim_mesh[p]];
      V3<R> raw = mi.left_handed ? cross(pr.e2, pr.e1) : cross(pr.e1, pr.e2);
      gf = mesh_frame(mi, point, raw.unit());
      if (mi.has_normals) {
        // surface/mesh.cc:620-641: interpolated shading normal; the
        // geometric normal is flipped to its side
        V3<R> norm = tri_vn[3 * p] * (R(1) - hu - hv);
        norm += tri_vn[3 * p + 1] * hu;
        norm += tri_vn[3 * p + 2] * hv;
        norm = norm.unit();
        if (gf.to(norm).z < 0) gf.z = -gf.z;
        nf = mesh_frame(mi, point, norm);
      } else {
        nf = gf;
      }
    } else if (pr.kind == 1 || pr.kind == 4) {
      // the tripar frame (surface/tripar.cc:45-66); an ellipse's
      // (surface/ellipse.cc:48-75: s = radius1 / |radius1|, normal
      // unit(edge2 x edge1)) is the same bits, its radii being the
      // stored edges halved
      R e1_len = pr.e1.length(), inv_e1_len = R(1) / e1_len;
      V3<R> norm = cross(pr.e2, pr.e1).unit();
      V3<R> s = pr.e1 * inv_e1_len;
      V3<R> t = cross(s, norm);
      nf = Frame<R>(point, s, t, norm);
      gf = nf;
    } else if (pr.kind == 5) {
      // Cylinder::IsecInfo::make_intersect (surface/cylinder.cc:61-80):
      // normal from the local hit point's x, y; t along the axis
      V3<R> norm = local_normal(pr, ray, ray.t1);
      V3<R> t = xform_vec(pr.l2w, V3<R>(0, 0, 1)).unit();
      V3<R> s = cross(norm, t);
      nf = Frame<R>(point, s, t, norm);
      gf = nf;
    } else if (pr.kind == 6) {
      // Sphere2::IsecInfo::make_intersect (surface/sphere2.cc:50-75):
      // local tangent onorm x z (x at the poles), taken to world space
      const V3<R> onorm = local_onorm(pr, ray, ray.t1);
      const R Eps = R(1e-15);
      const V3<R> ot = (std::fabs(onorm.x) < Eps && std::fabs(onorm.y) < Eps)
                           ? V3<R>(1, 0, 0)
                           : cross(onorm, V3<R>(0, 0, 1));
      V3<R> norm = xform_transpose(pr.w2l, onorm).unit();
      V3<R> t = xform_vec(pr.l2w, ot).unit();
      V3<R> s = cross(norm, t);
      nf = Frame<R>(point, s, t, norm);
      gf = nf;
    } else {
      V3<R> norm = (point - pr.center).unit();
      V3<R> s = cross(norm, V3<R>(0, 0, 1));
      if (s.length_squared() < R(1e-15)) s = norm.perpendicular();
      s = s.unit();
      V3<R> t = cross(s, norm);
      nf = Frame<R>(point, s, t, norm);
      gf = nf;
    }
    V3<R> wv = -ray.dir.unit();
    is.v = nf.to(wv);
    V3<R> gv = gf.to(wv);
    is.back = gv.z < 0;
    if (is.back) {
      is.v.z = -is.v.z;
      nf.z = -nf.z;
    }
    is.geom_n = nf.to(gf.z);
    is.geom_n.z = std::fabs(is.geom_n.z);
    is.nf = nf;
    is.gf = gf;
    is.mat = prim_mat[p];
    is.bsdf_mat = bsdf_material(is.mat);
    // A fully opaque stencil's get_bsdf is its underlying material's
    // (material/stencil.cc:140-148), which may have none.
    if (is.bsdf_mat >= 0 && mats[is.bsdf_mat].type == SGR_MAT_STENCIL &&
        mats[is.bsdf_mat].color.sum() >= R(3))
      is.bsdf_mat = bsdf_material(mats[is.bsdf_mat].underlying);
    is.media_ior = media_ior;
    is.enclosing_ior = enclosing_ior;
  }

  // Intersect::recursive_ray (render/intersect.cc:195-218)
  Ray<R> recursive_ray(const Isec<R> &is, const V3<R> &dir, R max_dist, bool shadow) const {
    Ray<R> r;
    r.origin = is.nf.origin;
    r.dir = is.nf.from(dir);
    r.t0 = min_trace;
    if (shadow) {
      if (max_dist == 0) max_dist = horizon;
      r.t1 = max_dist - min_trace;
    } else {
      r.t1 = horizon;
    }
    return r;
  }

  // Light::Sampler::sample (light/surface-light-sampler.cc:47-64,
  // light/sphere-light-sampler.cc:53-92, light/envmap-light.cc:100-125)
  LightSample<R> light_sample(const Light<R> &L, const Isec<R> &is, R u, R v,
                              R extra) const {
    LightSample<R> out;
    if (L.kind == 0) {
      const Prim<R> &pr = prims[L.prim];
      // surface/tripar.cc:210-233 area sample
      V3<R> norm = cross(pr.e2, pr.e1);
      R area = norm.length();
      norm = norm / area;
      if (!pr.parallelogram && u + v > 1) { u = R(1) - u; v = R(1) - v; }
      if (!pr.parallelogram) area *= R(0.5f);
      V3<R> pos = pr.v0 + pr.e1 * u + pr.e2 * v;
      R apdf = R(1) / area;
      R dist, pdf;
      V3<R> dir;
      if (angular(pos, norm, apdf, is.nf.origin, dist, dir, pdf)) {
        V3<R> d = is.nf.to(dir);
        if (d.z > 0) {
          out.val = L.intensity;
          out.pdf = pdf;
          out.dir = d;
          out.dist = dist;
        }
      }
    } else if (L.kind == 1) {
      const Prim<R> &pr = prims[L.prim];
      V3<R> lc = is.nf.to(pr.center - is.nf.origin);
      R radius = pr.radius;
      if (lc.z >= -radius) {
        R pdf = R(1) / sphere_solid_angle(lc, radius);
        R r_sqrt = radius * std::sqrt(u);
        R ang = v * R(2) * R(PIf);
        R x = r_sqrt * std::cos(ang);
        R y = r_sqrt * std::sin(ang);
        R z = std::sqrt(std::fabs(radius * radius - x * x - y * y)) *
              std::sin(R(PIf) * (extra - R(0.5f)));
        V3<R> sv = lc + V3<R>(x, y, z);
        if (is.cos_n(sv) > 0 && is.cos_geom_n(sv) > 0) {
          V3<R> sd = sv.unit();
          R dist;
          if (sphere_intersects(V3<R>(0, 0, 0), radius, -lc, sd, R(0), dist)) {
            out.val = L.intensity;
            out.pdf = pdf;
            out.dir = sd;
            out.dist = dist;
          }
        }
      }
    } else if (L.kind == 6) {
      // PointLight::Sampler::sample (light/point-light.cc:79-111)
      const V3<R> lvec = is.nf.to(L.pos - is.nf.origin);
      if (is.cos_n(lvec) > 0 && is.cos_geom_n(lvec) > 0) {
        const R dist = lvec.length();
        const R inv_dist = R(1) / dist;
        const R cos_dir = dot((is.nf.origin - L.pos) * inv_dist, L.fz);
        if (cos_dir >= L.cos_half) {
          out.val = point_intensity(L, cos_dir) * inv_dist * inv_dist;
          out.pdf = 1;
          out.dir = lvec * inv_dist;
          out.dist = dist;
        }
      }
    } else if (L.kind >= 3) {
      // SurfaceLightSampler::sample (light/surface-light-sampler.cc:
      // 47-64) over the surface's own sampler
      V3<R> pos, norm;
      R apdf;
      if (L.kind == 5) {
        sphere2_from_viewpoint(prims[L.prim], is.nf.origin, u, v, pos, norm, apdf);
      } else if (L.kind == 4) {
        cylinder_from_viewpoint(prims[L.prim], is.nf.origin, u, v, pos, norm, apdf);
      } else {
        ellipse_area(prims[L.prim], u, v, pos, norm, apdf);
      }
      R dist, pdf;
      V3<R> dir;
      if (angular(pos, norm, apdf, is.nf.origin, dist, dir, pdf)) {
        V3<R> d = is.nf.to(dir);
        if (d.z > 0) {
          out.val = L.intensity;
          out.pdf = pdf;
          out.dir = d;
          out.dist = dist;
        }
      }
    } else {
      const Envmap<R> &E = envs[L.env];
      float mu, mv, pdf;
      E.hist_sample(float(u), float(v), mu, mv, pdf);
      V3<R> ld = Envmap<R>::latlong_to_vec(mu, mv);
      V3<R> wd = E.frame.from(ld);
      V3<R> d = is.nf.to(wd);
      if (is.cos_n(d) < 0 || is.cos_geom_n(d) < 0) return out;
      out.val = E.map(ld);
      out.pdf = R(pdf) * R(0.25f * INV_PIf);
      out.dir = d;
      out.dist = 0;
    }
    return out;
  }

  // surface/surface-sampler.cc:26-46 AngularSample
  static bool angular(const V3<R> &pos, const V3<R> &normal, R apdf, const V3<R> &vp,
                      R &dist, V3<R> &dir, R &pdf) {
    V3<R> vv = pos - vp;
    dist = vv.length();
    pdf = 0;
    if (dist > 0) {
      R inv = R(1) / dist;
      dir = vv * inv;
      R adj = dot(-normal, dir);
      if (adj > 0) {
        R dw_dA = adj * (inv * inv);
        pdf = apdf / dw_dA;
      }
    }
    return pdf > 0;
  }

  // disk_sample (geometry/disk-sample.h:27-71): concentric map of PARAM
  // to the disk of RADIUS.
  static void disk_sample(R radius, R pu, R pv, R &dx, R &dy) {
    const R u = R(2) * pu - R(1), v = R(2) * pv - R(1);
    if (u == 0 && v == 0) {
      dx = dy = 0;
      return;
    }
    R r, theta;
    if (u >= -v) {
      if (u > v) {
        r = u;
        theta = (v > 0) ? v / r : R(8) + v / r;
      } else {
        r = v;
        theta = R(2) - u / r;
      }
    } else {
      if (u <= v) {
        r = -u;
        theta = R(4) - v / r;
      } else {
        r = -v;
        theta = R(6) + u / r;
      }
    }
    theta = theta * R(PIf / 2);
    dx = radius * (r * std::cos(theta));
    dy = radius * (r * std::sin(theta));
  }

  // Ellipse::Sampler::sample (surface/ellipse.cc:218-229): a disk sample
  // mapped onto the bounding parallelogram; pdf 1 / area.
  static void ellipse_area(const Prim<R> &pr, R pu, R pv, V3<R> &pos, V3<R> &norm, R &apdf) {
    R dx, dy;
    disk_sample(R(0.5), pu, pv, dx, dy);
    dx = dx + R(0.5);
    dy = dy + R(0.5);
    pos = pr.v0 + pr.e1 * dx + pr.e2 * dy;
    norm = cross(pr.e2, pr.e1).unit();
    apdf = R(4 * INV_PIf) / cross(pr.e2, pr.e1).length();
  }

  // sphere_sample / sphere_sample_inverse (geometry/sphere-sample.h,
  // geometry/cone-sample.h:27-62 with cos_half_angle -1).
  static V3<R> sphere_sample(R pu, R pv) {
    const R z = R(-1) + pv * R(2);
    const R r = std::sqrt(std::fabs(R(1) - z * z));
    const R phi = pu * R(2) * R(PIf);
    return V3<R>(r * std::cos(phi), r * std::sin(phi), z);
  }
  static void sphere_sample_inverse(const V3<R> &d, R &pu, R &pv) {
    const R phi = std::atan2(d.y, -d.x);
    pu = clamp01(R(0.5f) - (phi * R(INV_PIf) * R(0.5f)));
    pv = clamp01((d.z + R(1)) / R(2));
  }

  // Surface::Sampler::sample_with_approx_area_pdf (surface/surface-
  // sampler.h:141-152): the area pdf from the parallelogram spanned by
  // the positions at PARAM + (delta, 0) and PARAM + (0, delta).
  template<class F> static R approx_area_pdf(const F &f, R pu, R pv, V3<R> &pos) {
    const R delta = R(0.0001f);
    pos = f(pu, pv);
    const V3<R> pdu = f(pu + delta, pv), pdv = f(pu, pv + delta);
    const R sample_area = cross(pdu - pos, pdv - pos).length();
    const R param_area = delta * delta;
    return sample_area == 0 ? R(0) : param_area / sample_area;
  }

  // Cylinder::Sampler::sample_from_viewpoint (surface/cylinder.cc:
  // 239-306): an area sample, mirrored through the axis when its normal
  // faces away from VP, pdf doubled.
  static void cylinder_from_viewpoint(const Prim<R> &pr, const V3<R> &vp, R pu, R pv,
                                      V3<R> &pos, V3<R> &norm, R &apdf) {
    const R theta = pu * R(2) * R(PIf);
    norm = xform_transpose(pr.w2l, V3<R>(std::cos(theta), std::sin(theta), R(0))).unit();
    auto f = [&pr](R a, R b) {
      const R th = a * R(2) * R(PIf);
      return xform_pos(pr.l2w, V3<R>(std::cos(th), std::sin(th), R(2) * b - R(1)));
    };
    apdf = approx_area_pdf(f, pu, pv, pos);
    if (dot(norm, pos - vp) > 0) {
      V3<R> op = xform_pos(pr.w2l, pos);
      op.x = -op.x;
      op.y = -op.y;
      pos = xform_pos(pr.l2w, op);
      norm = -norm;
    }
    apdf = apdf * R(2);
  }

  // Sphere2::Sampler::sample_from_viewpoint (surface/sphere2.cc:
  // 238-270): a sample of the hemisphere facing VP (the other half
  // mirrored through the centre), pdf doubled.
  static void sphere2_from_viewpoint(const Prim<R> &pr, const V3<R> &vp, R pu, R pv,
                                     V3<R> &pos, V3<R> &norm, R &apdf) {
    const V3<R> ov = xform_pos(pr.w2l, vp);
    V3<R> onorm = sphere_sample(pu, pv);
    R su = pu, sv = pv;
    if (dot(onorm, ov) < 0) {
      onorm = -onorm;
      sphere_sample_inverse(onorm, su, sv);
    }
    norm = xform_transpose(pr.w2l, onorm).unit();
    auto f = [&pr](R a, R b) { return xform_pos(pr.l2w, sphere_sample(a, b)); };
    apdf = approx_area_pdf(f, su, sv, pos) * R(2);
  }

  // Frame (const Vec &z) (geometry/frame.h:64-73): x from z's
  // perpendicular, y = z x x.
  static Frame<R> frame_of(const V3<R> &org, const V3<R> &z) {
    const V3<R> x = z.perpendicular().unit();
    const V3<R> y = cross(z, x).unit();
    return Frame<R>(org, x, y, z);
  }

  // Light::Sampler::sample (param, dir_param), the "free" sample a photon
  // starts from: SurfaceLightSampler (light/surface-light-sampler.cc:
  // 72-100: an area sample, a cosine-distributed direction about its
  // normal, pdf = area pdf / pi), SphereLightSampler (light/sphere-light-
  // sampler.cc:120-155) and EnvmapLight::Sampler (light/envmap-light.cc:
  // 189-230: a direction from the intensity distribution, a start point
  // on the disk tangent to the scene's bounding sphere; its value is the
  // map looked up with the world-space direction, as the reference
  // passes it).
  FreeSample<R> light_free_sample(const Light<R> &L, R pu, R pv, R du, R dv) const {
    FreeSample<R> out;
    if (L.kind == 6) {
      // PointLight::Sampler::sample (light/point-light.cc:115-126) with
      // cone_sample / cone_sample_pdf (geometry/cone-sample.h:29-77)
      const R z = L.cos_half + dv * (R(1) - L.cos_half);
      const R r = std::sqrt(std::fabs(R(1) - z * z));
      const R phi = du * R(2) * R(PIf);
      V3<R> dir(r * std::cos(phi), r * std::sin(phi), z);
      out.pdf = R(0.5f) * R(INV_PIf) / (R(1) - L.cos_half);
      out.val = point_intensity(L, dir.z);
      if (L.cos_half != R(-1)) dir = Frame<R>(L.pos, L.fx, L.fy, L.fz).from(dir);
      out.pos = L.pos;
      out.dir = dir;
      return out;
    }
    if (L.kind == 2) {
      const Envmap<R> &E = envs[L.env];
      float mu, mv, pdf;
      E.hist_sample(float(du), float(dv), mu, mv, pdf);
      const V3<R> light_dir = Envmap<R>::latlong_to_vec(mu, mv);
      const V3<R> world_dir = E.frame.from(light_dir);
      R p = R(pdf) * R(0.25f) * R(INV_PIf);
      const V3<R> cen = bbox.center();
      const R rad = bbox.radius();
      R px, py;
      disk_sample(rad, pu, pv, px, py);  // geometry/tangent-disk-sample.h:31-51
      const Frame<R> f = frame_of(cen, world_dir);
      out.pos = cen + f.from(V3<R>(px, py, rad));
      p = p / (R(float(rad * rad)) * R(PIf));
      out.val = E.map(world_dir);
      out.pdf = p;
      out.dir = -world_dir;
      return out;
    }
    V3<R> pos, norm;
    R apdf;
    if (L.kind == 1) {
      const Prim<R> &pr = prims[L.prim];
      const V3<R> spv = sphere_sample(pu, pv);
      pos = pr.center + spv * pr.radius;
      apdf = R(1) / (pr.radius * pr.radius * R(float(4 * PI_D)));
      norm = spv;
    } else if (L.kind == 0) {
      const Prim<R> &pr = prims[L.prim];
      V3<R> n = cross(pr.e2, pr.e1);
      R area = n.length();
      norm = n / area;
      R u = pu, v = pv;
      if (!pr.parallelogram && u + v > 1) { u = R(1) - u; v = R(1) - v; }
      if (!pr.parallelogram) area *= R(0.5f);
      pos = pr.v0 + pr.e1 * u + pr.e2 * v;
      apdf = R(1) / area;
    } else if (L.kind == 3) {
      ellipse_area(prims[L.prim], pu, pv, pos, norm, apdf);
    } else if (L.kind == 4) {  // Cylinder::Sampler::sample (surface/cylinder.cc:225-233)
      const Prim<R> &pr = prims[L.prim];
      const R theta = pu * R(2) * R(PIf);
      norm = xform_transpose(pr.w2l, V3<R>(std::cos(theta), std::sin(theta), R(0))).unit();
      auto f = [&pr](R a, R b) {
        const R th = a * R(2) * R(PIf);
        return xform_pos(pr.l2w, V3<R>(std::cos(th), std::sin(th), R(2) * b - R(1)));
      };
      apdf = approx_area_pdf(f, pu, pv, pos);
    } else {  // Sphere2::Sampler::sample (surface/sphere2.cc:225-232)
      const Prim<R> &pr = prims[L.prim];
      norm = xform_transpose(pr.w2l, sphere_sample(pu, pv)).unit();
      auto f = [&pr](R a, R b) { return xform_pos(pr.l2w, sphere_sample(a, b)); };
      apdf = approx_area_pdf(f, pu, pv, pos);
    }
    // CosDist (material/cos-dist.h:24-44) about the normal
    const V3<R> d = z_normal_symm_vec(std::sqrt(du), dv);
    out.pos = pos;
    out.dir = frame_of(V3<R>(), norm).from(d);
    out.pdf = apdf * R(INV_PIf);
    out.val = L.intensity;
    return out;
  }

  // light/sphere-light-sampler.cc:40-48
  static R sphere_solid_angle(const V3<R> &lc, R radius) {
    R dist = lc.length();
    if (dist < radius) return R(4) * R(PIf);
    return R(2) * R(PIf) * (R(1) - std::cos(std::asin(radius / dist)));
  }

  // Light::Sampler::eval (light/surface-light-sampler.cc:86-100,
  // light/sphere-light-sampler.cc:130-152, light/envmap-light.cc:150-170)
  LightSample<R> light_eval(const Light<R> &L, const Isec<R> &is, const V3<R> &dir) const {
    LightSample<R> out;
    if (L.kind == 0) {
      const Prim<R> &pr = prims[L.prim];
      V3<R> wdir = is.nf.from(dir);
      R t, uu, vv;
      if (tripar_intersects(pr.v0, pr.e1, pr.e2, pr.parallelogram, is.nf.origin, wdir,
                            R(0), t, uu, vv)) {
        V3<R> norm = cross(pr.e2, pr.e1);
        R area = norm.length();
        norm = norm / area;
        if (!pr.parallelogram && uu + vv > 1) { uu = R(1) - uu; vv = R(1) - vv; }
        if (!pr.parallelogram) area *= R(0.5f);
        V3<R> pos = pr.v0 + pr.e1 * uu + pr.e2 * vv;
        R dist, pdf;
        V3<R> sd;
        if (angular(pos, norm, R(1) / area, is.nf.origin, dist, sd, pdf)) {
          out.val = L.intensity;
          out.pdf = pdf;
          out.dist = dist;
        }
      }
    } else if (L.kind == 1) {
      const Prim<R> &pr = prims[L.prim];
      V3<R> lc = is.nf.to(pr.center - is.nf.origin);
      if (lc.z >= -pr.radius) {
        R dist;
        if (sphere_intersects(V3<R>(0, 0, 0), pr.radius, -lc, dir, R(0), dist)) {
          out.val = L.intensity;
          out.pdf = R(1) / sphere_solid_angle(lc, pr.radius);
          out.dist = dist;
        }
      }
    } else if (L.kind == 6) {
      // PointLight::Sampler::eval (light/point-light.cc:130-133): nothing
    } else if (L.kind >= 3) {
      // SurfaceLightSampler::eval (light/surface-light-sampler.cc:
      // 86-100) over the surface's eval_from_viewpoint
      const Prim<R> &pr = prims[L.prim];
      const V3<R> vp = is.nf.origin, wdir = is.nf.from(dir);
      V3<R> pos, norm;
      R apdf = 0;
      bool hit = false;
      if (L.kind == 3) {  // surface/ellipse.cc:233-246
        R t, uu, vv;
        if (ellipse_intersects(pr.v0, pr.e1, pr.e2, vp, wdir, R(0), t, uu, vv)) {
          pos = vp + wdir * t;
          norm = cross(pr.e2, pr.e1).unit();
          apdf = R(4 * INV_PIf) / cross(pr.e2, pr.e1).length();
          hit = true;
        }
      } else if (L.kind == 4) {  // surface/cylinder.cc:322-347
        const V3<R> ov = xform_pos(pr.w2l, vp), od = xform_vec(pr.w2l, wdir);
        R t;
        if (cylinder_intersects(ov, od, R(0), t)) {
          const V3<R> lp = ov + od * t;
          R pu = R(std::atan2(lp.y, lp.x)) * R(INV_PIf) * R(0.5f);
          if (pu < 0) pu += R(1);
          const R pv = lp.z * R(0.5f) + R(0.5f);
          cylinder_from_viewpoint(pr, vp, clamp01(pu), clamp01(pv), pos, norm, apdf);
          hit = true;
        }
      } else {  // surface/sphere2.cc:278-298
        const V3<R> ov = xform_pos(pr.w2l, vp), od = xform_vec(pr.w2l, wdir);
        R t;
        if (sphere_intersects(V3<R>(0, 0, 0), R(1), ov, od, R(0), t)) {
          R pu, pv;
          sphere_sample_inverse(ov + od * t, pu, pv);
          sphere2_from_viewpoint(pr, vp, pu, pv, pos, norm, apdf);
          hit = true;
        }
      }
      R dist, pdf;
      V3<R> sd;
      if (hit && angular(pos, norm, apdf, vp, dist, sd, pdf)) {
        out.val = L.intensity;
        out.pdf = pdf;
        out.dist = dist;
      }
    } else {
      const Envmap<R> &E = envs[L.env];
      V3<R> wd = is.nf.from(dir);
      V3<R> ld = E.frame.to(wd);
      float mu, mv;
      Envmap<R>::latlong(ld, mu, mv);
      out.val = E.lookup(mu, mv);
      out.pdf = R(E.hist_pdf(mu, mv)) * R(0.25f * INV_PIf);
      out.dist = 0;
    }
    return out;
  }
};

}  // namespace orc
