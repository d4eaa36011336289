// ORACLE -- test infrastructure only.  Only tests/, the smoke check and the
// CPU-baseline leg of bench.py load liboracle.so; the product never does.
//
// CPU restatement of the reference's plain path-tracing render path
// (render/path-integ.cc:133-407, render/direct-illum.cc:82-255,
// render-mgr/renderer.cc:41-100, image/image-filter-conv.h:109-186) and
// its ray queries (space/space.cc:36-194, space/octree.cc:101-300).
// C entry points for ctypes; every call catches exceptions and returns a
// status, the message is copied into the caller's buffer.
#include "orc_render.h"
#include "orc_photon.h"
#include <atomic>
#include <thread>
#include <memory>
#include <cstdio>
#include <functional>

using namespace orc;

namespace {

struct Counters {
  uint64_t eye = 0, isect = 0, shadow = 0, nc_c = 0, st_c = 0, nc_s = 0,
           st_s = 0, illum = 0, paths = 0, media_ovf = 0, dint_ovf = 0;
  void add(const Counters &o) {
    eye += o.eye; isect += o.isect; shadow += o.shadow; nc_c += o.nc_c;
    st_c += o.st_c; nc_s += o.nc_s; st_s += o.st_s; illum += o.illum;
    paths += o.paths; media_ovf += o.media_ovf; dint_ovf += o.dint_ovf;
  }
};

// render/mis-sample-weight.h:45-77 (power heuristic, one sample each).
template<class R> static inline R mis_weight(R pdf, R other_pdf) {
  R term = R(1) * pdf;
  R other = R(1) * other_pdf;
  if (term > R(1e10f)) return 1;
  if (other > R(1e10f)) return 0;
  R t2 = term * term, o2 = other * other;
  return (term * pdf) / (t2 + o2);
}

// image/image-mitchell-filt.h:45-65
static inline float mitchell1(float x, float inv_radius, float b, float c) {
  x = std::fabs(2.f * x * inv_radius);
  if (x > 2.f) return 0;
  if (x > 1.f)
    return ((-b - 6 * c) * x * x * x + (6 * b + 30 * c) * x * x +
            (-12 * b - 48 * c) * x + (8 * b + 24 * c)) * (1.f / 6.f);
  return ((12 - 9 * b - 6 * c) * x * x * x + (-18 + 12 * b + 6 * c) * x * x +
          (6 - 2 * b)) * (1.f / 6.f);
}

// Sample-stream layout (see orc_rng.h).
struct Stream {
  uint32_t spp, minp, L, ds;
  GridShape g_cam, g_dir, g_cont, g_dir_rand, g_dint;
  uint32_t nsub_dir, nsub_dir_rand, nsub_dint;
  // The direct integrator's light samples: light_samples per light
  // (default 16), channels 1 + 2 l (light) and 2 + 2 l (BSDF).
  void init_direct(uint32_t light_samples) {
    g_dint = grid_shape(spp * (light_samples ? light_samples : 16u));
    nsub_dint = g_dint.total / spp;
  }
  void init(uint32_t spp_, uint32_t minp_, uint32_t L_, uint32_t ds_) {
    spp = spp_; minp = minp_; L = L_; ds = ds_;
    g_cam = grid_shape(spp);
    g_cont = grid_shape(spp);
    g_dir = grid_shape(spp * (ds ? ds : 1));
    nsub_dir = g_dir.total / spp;           // render/sample-set.h:330-345
    g_dir_rand = grid_shape(ds ? ds : 1);   // one-sample random set
    nsub_dir_rand = g_dir_rand.total;
  }
  uint32_t chan_light(uint32_t b, uint32_t l) const { return 1 + b * (2 * L + 1) + 2 * l; }
  uint32_t chan_cont(uint32_t b) const { return 1 + b * (2 * L + 1) + 2 * L; }
  static void strat(uint64_t pk, uint32_t chan, uint32_t entry, const GridShape &g,
                    float &u, float &v) {
    uint64_t ck = chan_key(pk, chan);
    uint32_t cell = permute(entry, g.total, ck);
    uint32_t j = cell % g.u_steps, i = cell / g.u_steps;
    float us = 1.0f / (float)g.u_steps, vs = 1.0f / (float)g.v_steps;
    u = clamp01((float)j * us + draw(ck, 2 * entry) * us);
    v = clamp01((float)i * vs + draw(ck, 2 * entry + 1) * vs);
  }
  static uint32_t dim(uint32_t b, uint32_t slot) { return (b << 20) | slot; }
};

// The photon maps of one scene and the photon integrator's resolved
// parameters (PhotonInteg::GlobalState, render/photon-integ.cc:36-148).
template<class R> struct PhotonMaps {
  PhotonMap<R> map[3];  // 0 direct, 1 caustic, 2 indirect
  float scale[3] = {0, 0, 0};
  uint64_t paths[3] = {0, 0, 0};
  uint64_t paths_shot = 0, overflows = 0;
  unsigned k = 50;
  R radius_sq = R(0.01f);
  unsigned fg = 16, fg_bsdf = 8, fg_photon = 8;
  bool direct_illum = true;
  unsigned direct_samples = 16;
};

struct OrcPhotonMap {
  int fp64 = 0;
  PhotonMaps<float> f;
  PhotonMaps<double> d;
  PhotonMaps<float> &get(float *) { return f; }
  PhotonMaps<double> &get(double *) { return d; }
};

// render/photon-integ.cc:36-148: the final-gathering sample split and
// whether direct photons are needed.  TARGETS: direct, caustic,
// indirect photon counts.
static void resolve_photon_params(const sgr_photon_params &p, unsigned &fg, unsigned &fg_bsdf,
                                  unsigned &fg_photon, unsigned targets[3]) {
  const unsigned U = SGR_PHOTON_UNSPEC;
  fg = p.fg_samples;
  fg_bsdf = p.fg_bsdf_samples;
  fg_photon = p.fg_photon_samples;
  if (p.final_gather) {
    if (fg == U && fg_bsdf == U && fg_photon == U) {
      fg = 16;
      fg_bsdf = 8;
      fg_photon = 8;
    } else if (fg == U) {
      if (fg_photon == U) fg_photon = 0;
      if (fg_bsdf == U) fg_bsdf = 0;
      fg = fg_bsdf + fg_photon;
    } else if (fg_bsdf == U || fg_photon == U) {
      if (fg_bsdf == U && fg_photon == U) {
        if (fg == 1) {
          fg_bsdf = fg_photon = 0;
        } else {
          fg_photon = fg / 2;
          fg_bsdf = (fg + 1) / 2;
        }
      } else if (fg_bsdf == U) {
        fg_photon = fg - fg_bsdf;  // the reference's own mix-up, kept
      } else {
        fg_bsdf = fg - fg_photon;
      }
    }
  } else {
    fg = fg_bsdf = fg_photon = 0;
  }
  targets[0] = (p.direct_illum && fg == 0) ? 0u : p.direct_photons;
  targets[1] = p.caustic_photons;
  targets[2] = p.indirect_photons;
}

// PhotonShooter::shoot (photon/photon-shooter.cc:29-191) with
// PhotonInteg::Shooter::deposit (render/photon-integ.cc:196-226).  Path
// N's light, light parameters and first BSDF parameters come from radical
// inverses of N; its later draws come from the path's own counter stream
// (key path_key(pixel_key(seed, -2, -2), N)): two BSDF parameters then,
// past 3 bounces, the roulette draw, per bounce.  A path deposits at
// most SGR_PHOTON_MAX_DEPOSITS photons (later hits are counted) and ends
// after SGR_PHOTON_MAX_BOUNCES hits, as the device's do.
template<class R>
static void shoot_photons(const Scene<R> &sc, const sgr_photon_params &pp,
                          const unsigned targets[3], PhotonMaps<R> &out) {
  std::vector<Photon<R>> sets[3];
  uint64_t num_paths[3] = {0, 0, 0};
  const unsigned nl = (unsigned)sc.lights.size();
  auto complete = [&](int i) { return sets[i].size() == targets[i]; };
  auto all_complete = [&]() { return complete(0) && complete(1) && complete(2); };
  const uint64_t pk = pixel_key(pp.seed, -2, -2);
  unsigned path = 0;
  if (nl != 0) {
    for (; !all_complete(); path++) {
      const unsigned li = unsigned(radical_inverse(path, 11) * nl);
      const Light<R> &L = sc.lights[li];
      FreeSample<R> samp = sc.light_free_sample(
          L, R(float(radical_inverse(path, 2))), R(float(radical_inverse(path, 3))),
          R(float(radical_inverse(path, 5))), R(float(radical_inverse(path, 7))));
      if (samp.val.is_zero() || samp.pdf == 0) continue;
      for (int i = 0; i < 3; i++)
        if (!complete(i)) num_paths[i]++;
      const uint64_t qk = path_key(pk, path);
      uint32_t nd = 0;
      unsigned history = 0, deposits = 0;
      R media[8];
      int depth = 1;
      media[0] = 1;
      V3<R> pos = samp.pos, dir = samp.dir;
      C3<R> power = samp.val * R(float(nl)) / samp.pdf;
      for (unsigned path_len = 0; path_len < SGR_PHOTON_MAX_BOUNCES; path_len++) {
        Ray<R> ray;
        ray.origin = pos;
        ray.dir = dir;
        ray.t0 = sc.min_trace;
        ray.t1 = sc.horizon;
        R hu, hv;
        unsigned inst = 0xffffffffu;
        const unsigned prim = sc.accel.closest(ray, hu, hv, nullptr, nullptr, &inst);
        if (prim == 0xffffffffu) break;
        Isec<R> is;
        const R enclosing = depth >= 2 ? media[depth - 2] : R(1);
        sc.make_isec(prim, inst, ray, hu, hv, media[depth - 1], enclosing, is);
        if (is.bsdf_mat < 0) break;
        Bsdf<R> bsdf;
        sc.make_bsdf(bsdf, is.bsdf_mat, is);
        // (no medium attenuates: the materials' media only refract)
        if (bsdf.supports(ALL & ~SPECULAR)) {
          if (deposits++ >= SGR_PHOTON_MAX_DEPOSITS) {
            out.overflows++;
          } else {
            const Photon<R> ph{is.nf.origin, power, -dir};
            int set = history == 0 ? 0
                      : (targets[1] != 0 && !(history & ALL_LAYERS & ~SPECULAR)) ? 1 : 2;
            if (!complete(set)) sets[set].push_back(ph);
          }
        }
        R u, v;
        if (path_len == 0) {
          u = R(float(radical_inverse(path, 13)));
          v = R(float(radical_inverse(path, 17)));
        } else {
          u = R(draw(qk, nd++));
          v = R(draw(qk, nd++));
        }
        const BsdfSample<R> bs = bsdf.sample(u, v, ALL);
        if (bs.val.is_zero() || bs.pdf == 0) break;
        if (path_len > 3) {
          if (draw(qk, nd++) < 0.5f) break;
          power /= R(0.5f);
        }
        pos = is.nf.origin;
        dir = is.nf.from(bs.dir);
        power *= bs.val * std::fabs(is.cos_n(bs.dir)) / bs.pdf;
        if (!(bs.flags & 0x08u)) history |= bs.flags;  // TRANSLUCENT samples leave none
        if (bs.flags & TRANSMISSIVE) {  // material/media.cc:75-118
          bool has;
          const R mior = sc.medium_ior(is.mat, has);
          if (has) {
            if (is.back) {
              if (depth >= 2) depth--;
            } else if (depth < 8) {
              media[depth++] = mior;
            }
          }
        }
      }
      if (path > 100000000u) break;
    }
  }
  out.paths_shot = path;
  for (int i = 0; i < 3; i++) {
    out.paths[i] = num_paths[i];
    out.scale[i] = num_paths[i] > 0 ? 1 / float(num_paths[i]) : 0.f;
    out.map[i].set_photons(sets[i]);
  }
}

template<class R> struct Tracer {
  const Scene<R> &sc;
  const sgr_render_params &P;
  Stream st;
  Counters cnt;

  const PhotonMaps<R> *pm = nullptr;  // SGR_INTEG_PHOTON
  std::vector<const Photon<R> *> found;
  Tracer(const Scene<R> &s, const sgr_render_params &p) : sc(s), P(p) {
    if (p.integrator == SGR_INTEG_PHOTON && p.photon_map)
      pm = &((OrcPhotonMap *)p.photon_map)->get((R *)nullptr);
    st.init(p.spp, p.min_path_len, (uint32_t)s.lights.size(), p.direct_samples);
    // the photon integrator's direct lighting takes its maps' direct_samples
    st.init_direct(pm ? pm->direct_samples : p.light_samples);
    if (pm) {
      const unsigned n[2] = {pm->fg_photon, pm->fg_bsdf};
      for (int i = 0; i < 2; i++)
        if (n[i]) {
          fg_grid[i] = grid_shape(p.spp * n[i]);
          fg_nsub[i] = fg_grid[i].total / p.spp;
        }
    }
  }

  // Scene::occludes: TRANS is multiplied by the partially occluding
  // surfaces the ray passes (MEDIUM: the index of the medium at its
  // origin, isec.media.medium).
  bool occluded(const Ray<R> &r, R medium, C3<R> &trans) {
    cnt.shadow++;
    return sc.occludes(r, medium, trans, &cnt.nc_s, &cnt.st_s);
  }

  // DirectIllum::sample_light (render/direct-illum.cc:126-255)
  C3<R> sample_light(const Isec<R> &is, const Bsdf<R> &bsdf, const Light<R> &L,
                     R lu, R lv, R bu, R bv, R extra, unsigned flags) {
    C3<R> rad(0);
    LightSample<R> ls = sc.light_sample(L, is, lu, lv, extra);
    if (ls.pdf > 0 && ls.val.gt0()) {
      BsdfValue<R> bval = bsdf.eval(ls.dir, flags);
      if (bval.val.gt0()) {
        Ray<R> ray = sc.recursive_ray(is, ls.dir, ls.dist, true);
        C3<R> trans(1);
        if (!occluded(ray, is.media_ior, trans)) {
          C3<R> r = ls.val * trans;
          if (L.kind != 6) r *= mis_weight(ls.pdf, bval.pdf);  // not for a point light
          r *= bval.val;
          r *= std::fabs(is.cos_n(ls.dir));
          r /= ls.pdf;
          rad += r;
        }
      }
    }
    if (L.kind == 6) return rad;  // no BSDF sample for a point light (direct-illum.cc:202)
    BsdfSample<R> bs = bsdf.sample(bu, bv, flags);
    if (bs.pdf > 0 && bs.val.gt0()) {
      LightSample<R> lval = sc.light_eval(L, is, bs.dir);
      if (lval.pdf > 0 && lval.val.gt0()) {
        Ray<R> ray = sc.recursive_ray(is, bs.dir, lval.dist, true);
        C3<R> trans(1);
        if (!occluded(ray, is.media_ior, trans)) {
          C3<R> r = lval.val * trans;
          r *= mis_weight(bs.pdf, lval.pdf);
          r *= bs.val;
          r *= std::fabs(is.cos_n(bs.dir));
          r /= bs.pdf;
          rad += r;
        }
      }
    }
    return rad;
  }

  // DirectIllum::sample_all_lights (render/direct-illum.cc:82-115)
  C3<R> direct(const Isec<R> &is, const Bsdf<R> &bsdf, uint32_t b, uint64_t pk,
               uint64_t qk, uint32_t s, unsigned flags) {
    cnt.illum++;
    C3<R> rad(0);
    bool strat = b < st.minp;
    uint32_t nsub = strat ? st.nsub_dir : st.nsub_dir_rand;
    for (uint32_t l = 0; l < st.L; l++) {
      C3<R> lrad(0);
      for (uint32_t j = 0; j < nsub; j++) {
        float lu, lv, bu, bv;
        uint32_t base = 16 + (l * nsub + j) * 8;
        if (strat) {
          Stream::strat(pk, st.chan_light(b, l), s * nsub + j, st.g_dir, lu, lv);
          Stream::strat(pk, st.chan_light(b, l) + 1, s * nsub + j, st.g_dir, bu, bv);
        } else {
          lu = draw(qk, Stream::dim(b, base + 0));
          lv = draw(qk, Stream::dim(b, base + 1));
          bu = draw(qk, Stream::dim(b, base + 2));
          bv = draw(qk, Stream::dim(b, base + 3));
        }
        float extra = draw(qk, Stream::dim(b, base + 4));
        lrad += sample_light(is, bsdf, sc.lights[l], R(lu), R(lv), R(bu), R(bv),
                             R(extra), flags);
      }
      rad += lrad / R(float(nsub));
    }
    return rad;
  }

  // PathInteg::Li (render/path-integ.cc:133-407).  Returns radiance and
  // sets alpha.
  // The media stack holds kMediaMax levels, as the device's does: a
  // transmission into a deeper medium is not pushed (counted in
  // media_ovf).  The reference's stack is unbounded
  // (material/media.cc:75-118); the two agree up to that depth.
  static constexpr int kMediaMax = 8;
  C3<R> Li(Ray<R> ray, uint64_t pk, uint64_t qk, uint32_t s, R &alpha) {
    R media[kMediaMax];
    int depth = 1;
    media[0] = 1;  // RenderContext::default_medium (ior 1)
    C3<R> radiance(0), T(1);
    R rr_ratio = 1;
    bool after_spec = false;
    alpha = 1;
    uint32_t path_len = 0;
    for (;;) {
      R hu, hv;
      cnt.isect++;
      unsigned inst = 0xffffffffu;
      unsigned prim = sc.accel.closest(ray, hu, hv, &cnt.nc_c, &cnt.st_c, &inst);
      bool include = path_len == 0 || after_spec || P.direct_samples == 0;
      if (prim == 0xffffffffu) {
        if (include) radiance += sc.background(ray.dir) * T;
        if (path_len == 0 && radiance.is_zero()) alpha = R(P.bg_alpha);
        break;
      }
      Isec<R> is;
      R enclosing = depth >= 2 ? media[depth - 2] : R(1);
      sc.make_isec(prim, inst, ray, hu, hv, media[depth - 1], enclosing, is);
      const Material<R> &hm = sc.mats[is.mat];
      if (include && hm.type == SGR_MAT_GLOW && !is.back)  // material/glow.cc:36-40
        radiance += hm.color * T;
      if (is.bsdf_mat < 0) break;
      Bsdf<R> bsdf;
      sc.make_bsdf(bsdf, is.bsdf_mat, is);
      if (P.direct_samples != 0)
        radiance += direct(is, bsdf, path_len, pk, qk, s, ALL & ~SPECULAR) * T;
      float cu, cv;
      if (path_len < st.minp) {
        Stream::strat(pk, st.chan_cont(path_len), s, st.g_cont, cu, cv);
      } else {
        cu = draw(qk, Stream::dim(path_len, 1));
        cv = draw(qk, Stream::dim(path_len, 2));
      }
      BsdfSample<R> bs = bsdf.sample(R(cu), R(cv), ALL);
      if (bs.pdf == 0 || bs.val.is_zero()) break;
      C3<R> f = bs.val * std::fabs(is.cos_n(bs.dir));
      f = f / bs.pdf;
      T *= f;
      if (path_len > st.minp) {
        R p = mn(T.intensity() * rr_ratio, R(1));
        R rr = R(draw(qk, Stream::dim(path_len, 0)));
        if (rr > p) break;
        T /= p;
        rr_ratio *= p;
      }
      if (path_len == P.max_path_len) break;
      ray = sc.recursive_ray(is, bs.dir, 0, false);
      after_spec = (bs.flags & SPECULAR) != 0;
      if (bs.flags & TRANSMISSIVE) {  // material/media.cc:75-118
        bool has;
        R mior = sc.medium_ior(is.mat, has);
        if (has) {
          if (is.back) {
            if (depth >= 2) depth--;
          } else if (depth < kMediaMax) {
            media[depth++] = mior;
          } else {
            cnt.media_ovf++;
          }
        }
      }
      path_len++;
    }
    return radiance;
  }

  // DirectInteg (render/direct-integ.h:31-67) on RecursiveInteg
  // (render/recursive-integ.cc:29-216).  At each hit: emission, direct
  // lighting of every light with light_samples samples each (the same
  // stratified samples at every depth: Lo gets the one Sample), then a
  // recursion into the perfectly specular reflection and the perfectly
  // specular transmission, each weighted by val |cos|; beyond depth 5 a
  // branch goes on with probability 1/2 and twice the weight, drawn from
  // the path's sequential stream (context.random()), which the sphere
  // light's extra draw also takes from.  The media of a branch are kept
  // as (innermost, enclosing) IORs, all that Media (isec, true) reads
  // (material/media.cc:25-52): entering pushes (material medium,
  // innermost), leaving sets both to the enclosing one (the new layer
  // keeps the old surrounding).  Radiance is accumulated top-down: a
  // branch's weight times its emission and direct light, depth first,
  // reflection before transmission; the reference sums the same terms
  // bottom-up.
  static constexpr uint32_t kDintDepth = SGR_DIRECT_MAX_DEPTH;
  static constexpr uint32_t kDintRand = 0xc0000000u;  // draw dims of the sequential stream
  uint64_t dq = 0;
  uint32_t dn = 0;
  R dint_draw() { return R(draw(dq, kDintRand + dn++)); }

  C3<R> dint_direct(const Isec<R> &is, const Bsdf<R> &bsdf, uint64_t pk, uint32_t s) {
    cnt.illum++;
    C3<R> rad(0);
    const uint32_t nsub = st.nsub_dint;
    for (uint32_t l = 0; l < st.L; l++) {
      C3<R> lrad(0);
      for (uint32_t j = 0; j < nsub; j++) {
        float lu, lv, bu, bv;
        Stream::strat(pk, 1 + 2 * l, s * nsub + j, st.g_dint, lu, lv);
        Stream::strat(pk, 2 + 2 * l, s * nsub + j, st.g_dint, bu, bv);
        R extra = sc.lights[l].kind == 1 ? dint_draw() : R(0);
        lrad += sample_light(is, bsdf, sc.lights[l], R(lu), R(lv), R(bu), R(bv), extra,
                             ALL & ~SPECULAR);
      }
      rad += lrad / R(float(nsub));
    }
    return rad;
  }

  // PhotonEval::Lo (photon/photon-eval.cc:44-118): the photons near the
  // hit, each weighted by the BSDF in its direction and a Gaussian filter
  // of its distance, times the map's scale, over pi r^2 with r the final
  // search radius.
  C3<R> lo_photon(const Isec<R> &is, const Bsdf<R> &bsdf, int set, unsigned flags) {
    const float scale = pm->scale[set];
    if (scale == 0) return C3<R>(0);
    if (!bsdf.supports(flags & ~SPECULAR)) return C3<R>(0);
    found.clear();
    const R maxd = pm->map[set].find(is.nf.origin, pm->k, pm->radius_sq, found);
    if (found.empty()) return C3<R>(0);
    GaussSqrt<R> g(maxd, R(2));
    const float gscale = 2.90898f;
    C3<R> rad(0);
    for (const Photon<R> *ph : found) {
      const BsdfValue<R> v = bsdf.eval(is.nf.to(ph->dir), flags);
      if (v.pdf != 0 && v.val.gt0()) {
        const float filt = g((ph->pos - is.nf.origin).length_squared());
        rad += v.val * ph->power * R(filt) * R(gscale);
      }
    }
    rad *= R(scale);
    rad /= maxd * R(PIf);
    return rad;
  }

  // PhotonEval::photon_dist (photon/photon-eval.cc:128-185): the
  // indirect photons near the hit binned by direction, weighted by their
  // power's intensity.
  void photon_dist(const Isec<R> &is, DirHistDist &dd) {
    dd.clear();
    found.clear();
    pm->map[2].find(is.nf.origin, pm->k, pm->radius_sq, found);
    for (const Photon<R> *ph : found) {
      R u, v;
      Scene<R>::sphere_sample_inverse(ph->dir, u, v);
      dd.add(float(u), float(v), float(ph->power.intensity()));
    }
    dd.calc();
  }

  // PhotonInteg::Lo_fgather_samp (render/photon-integ.cc:268-380): the
  // radiance the photon maps give at the surface a BSDF sample reaches,
  // through up to 3 levels of specular bounces there (sampled from the
  // path's sequential stream); past level 0 the surface's emission or the
  // background counts, times SCALE.  The media are the shading point's
  // throughout.
  C3<R> fg_samp(const Isec<R> &is, R m0, R m1, const BsdfSample<R> &bs, const C3<R> &scale,
                unsigned depth) {
    C3<R> rad(0);
    if (!(bs.val.gt0() && bs.pdf != 0)) return rad;
    const Ray<R> ray0 = sc.recursive_ray(is, bs.dir, 0, false);
    Ray<R> ray = ray0;
    cnt.isect++;
    R hu, hv;
    unsigned inst = 0xffffffffu;
    const unsigned prim = sc.accel.closest(ray, hu, hv, &cnt.nc_c, &cnt.st_c, &inst);
    if (prim == 0xffffffffu) {
      if (depth != 0) rad += sc.background(ray0.dir) * scale;
      return rad;
    }
    Isec<R> si;
    sc.make_isec(prim, inst, ray, hu, hv, m0, m1, si);
    if (si.bsdf_mat >= 0) {
      Bsdf<R> sb;
      sc.make_bsdf(sb, si.bsdf_mat, si);
      const C3<R> Li = lo_photon(si, sb, 0, ALL) + lo_photon(si, sb, 2, ALL) + lo_photon(si, sb, 1, ALL);
      const C3<R> li_to_lo = bs.val * std::fabs(is.cos_n(bs.dir)) / bs.pdf;
      rad += Li * li_to_lo;
      const unsigned spec = ALL_DIRECTIONS | SPECULAR;
      if (sb.supports(spec) && depth < 3) {
        const R u = dint_draw();
        const R v = dint_draw();
        const BsdfSample<R> rs = sb.sample(u, v, spec);
        rad += fg_samp(si, m0, m1, rs, scale, depth + 1) * li_to_lo;
      }
    }
    if (depth != 0) {
      const Material<R> &hm = sc.mats[si.mat];
      if (hm.type == SGR_MAT_GLOW && !si.back) rad += hm.color * scale;  // Intersect::Le
    }
    return rad;
  }

  // Color / Color with zero denominators giving 0 (color/color.h:294-300)
  static C3<R> cdiv(const C3<R> &a, const C3<R> &b) {
    return C3<R>(b.r == 0 ? R(0) : a.r / b.r, b.g == 0 ? R(0) : a.g / b.g,
                 b.b == 0 ? R(0) : a.b / b.b);
  }

  // PhotonInteg::Lo_fgather (render/photon-integ.cc:397-545): photon-
  // direction samples and BSDF samples, each through fg_samp, combined by
  // the power heuristic with sample counts.
  C3<R> lo_fgather(const Isec<R> &is, const Bsdf<R> &bsdf, R m0, R m1, uint64_t pk, uint32_t s,
                   bool avoid_caustics) {
    C3<R> rad(0);
    const unsigned nph = pm->fg_photon, nbs = pm->fg_bsdf;
    DirHistDist dd;
    photon_dist(is, dd);
    const unsigned bflags = bsdf.supports(ALL);
    for (unsigned i = 0; i < nph; i++) {
      float pu, pv, su, sv, spdf;
      Stream::strat(pk, fg_chan(0), s * fg_nsub[0] + i, fg_grid[0], pu, pv);
      dd.sample(pu, pv, su, sv, spdf);
      spdf *= 0.25f * INV_PIf;
      const V3<R> sdir = Scene<R>::sphere_sample(R(su), R(sv));
      const V3<R> bdir = is.nf.to(sdir);
      const BsdfValue<R> bv = bsdf.eval(bdir, ALL & ~SPECULAR);
      if (bv.pdf != 0 && spdf != 0) {
        C3<R> scale(1);
        if (avoid_caustics) {
          if (bflags & ALL_LAYERS & ~(SPECULAR | DIFFUSE))
            scale = cdiv(bsdf.eval(bdir, ALL & ~(SPECULAR | DIFFUSE)).val, bv.val);
          else
            scale = C3<R>(0);
        }
        BsdfSample<R> fake;
        fake.val = bv.val;
        fake.pdf = R(spdf);
        fake.dir = bdir;
        fake.flags = 0;
        rad += fg_samp(is, m0, m1, fake, scale, 0) *
               R(mis_sample_weight(spdf, float(nph), float(bv.pdf), float(nbs)));
      }
    }
    for (unsigned i = 0; i < nbs; i++) {
      float bu, bv;
      Stream::strat(pk, fg_chan(1), s * fg_nsub[1] + i, fg_grid[1], bu, bv);
      const BsdfSample<R> bs = bsdf.sample(R(bu), R(bv), ALL & ~SPECULAR);
      if (bs.pdf != 0) {
        R du, dv;
        Scene<R>::sphere_sample_inverse(is.nf.from(bs.dir), du, dv);
        const float ph_pdf = dd.pdf(float(du), float(dv)) * 0.25f * INV_PIf;
        const C3<R> scale = (avoid_caustics && (bs.flags & DIFFUSE)) ? C3<R>(0) : C3<R>(1);
        rad += fg_samp(is, m0, m1, bs, scale, 0) *
               R(mis_sample_weight(float(bs.pdf), float(nbs), ph_pdf, float(nph)));
      }
    }
    return rad;
  }

  // Final-gathering channels after the direct lighting's (1 + 2 l and
  // 2 + 2 l): photon-direction parameters 1 + 2 L, BSDF parameters 2 + 2 L,
  // fg_photon / fg_bsdf entries per sample on jittered grids.
  GridShape fg_grid[2];
  uint32_t fg_nsub[2] = {0, 0};
  uint32_t fg_chan(int which) const { return 1 + 2 * st.L + (uint32_t)which; }

  // PhotonInteg::Lo (render/photon-integ.cc:558-600)
  C3<R> pint_Lo(const Isec<R> &is, const Bsdf<R> &bsdf, uint64_t pk, uint32_t s, R m0, R m1) {
    const bool use_fg = pm->fg > 0;
    const bool caustics = !pm->map[1].photons.empty();
    C3<R> rad(0);
    if (pm->direct_illum) rad += dint_direct(is, bsdf, pk, s);
    else rad += lo_photon(is, bsdf, 0, ALL);
    if (caustics) rad += lo_photon(is, bsdf, 1, use_fg ? (ALL_DIRECTIONS | DIFFUSE) : ALL);
    if (use_fg) rad += lo_fgather(is, bsdf, m0, m1, pk, s, caustics);
    else rad += lo_photon(is, bsdf, 2, ALL);
    return rad;
  }

  void dint_Li(const Ray<R> &ray, R m0, R m1, uint32_t depth, C3<R> W, uint64_t pk,
               uint32_t s, C3<R> &total, R &alpha) {
    if (depth > 5) {  // render/recursive-integ.cc:115-138
      if (dint_draw() < R(0.5f)) return;
      W *= R(2);
    }
    cnt.isect++;
    R hu, hv;
    unsigned inst = 0xffffffffu;
    Ray<R> r = ray;
    unsigned prim = sc.accel.closest(r, hu, hv, &cnt.nc_c, &cnt.st_c, &inst);
    if (prim == 0xffffffffu) {
      total += sc.background(r.dir) * W;
      if (depth == 0) alpha = R(P.bg_alpha);
      return;
    }
    Isec<R> is;
    sc.make_isec(prim, inst, r, hu, hv, m0, m1, is);
    const Material<R> &hm = sc.mats[is.mat];
    if (hm.type == SGR_MAT_GLOW && !is.back) total += hm.color * W;
    if (is.bsdf_mat < 0) return;
    Bsdf<R> bsdf;
    sc.make_bsdf(bsdf, is.bsdf_mat, is);
    total += (pm ? pint_Lo(is, bsdf, pk, s, m0, m1) : dint_direct(is, bsdf, pk, s)) * W;
    BsdfSample<R> rs = bsdf.sample(R(0), R(0), SPECULAR | REFLECTIVE);
    if (rs.val.gt0()) {
      if (depth + 1 >= kDintDepth) {
        cnt.dint_ovf++;
      } else {
        C3<R> f = rs.val * std::fabs(is.cos_n(rs.dir));
        dint_Li(sc.recursive_ray(is, rs.dir, 0, false), m0, m1, depth + 1, W * f, pk, s, total,
                alpha);
      }
    }
    BsdfSample<R> xs = bsdf.sample(R(0), R(0), SPECULAR | TRANSMISSIVE);
    if (xs.val.gt0()) {
      if (depth + 1 >= kDintDepth) {
        cnt.dint_ovf++;
      } else {
        bool has;
        R mior = sc.medium_ior(is.mat, has);
        R n0 = is.back ? m1 : (has ? mior : R(1)), n1 = is.back ? m1 : m0;
        C3<R> f = xs.val * std::fabs(is.cos_n(xs.dir));
        dint_Li(sc.recursive_ray(is, xs.dir, 0, false), n0, n1, depth + 1, W * f, pk, s, total,
                alpha);
      }
    }
  }

  // RecursiveInteg::Li, top level (render/recursive-integ.cc:176-216)
  C3<R> dint(const Ray<R> &ray, uint64_t pk, uint64_t qk, uint32_t s, R &alpha) {
    dq = qk;
    dn = 0;
    alpha = 1;
    C3<R> total(0);
    dint_Li(ray, R(1), R(1), 0, C3<R>(1), pk, s, total, alpha);
    if (!total.is_zero()) alpha = 1;
    return total;
  }

  // One path: camera ray (render-mgr/renderer.cc:71-85,
  // camera/camera.h:322-389) and PathInteg::Li.  Output: sx, sy, Tint.
  void path(const sgr_camera &cam, R max_trace, int px, int py, uint32_t s,
            float out[6]) {
    uint64_t pk = pixel_key(P.seed, px, py);
    uint64_t qk = path_key(pk, s);
    float cu, cv;
    Stream::strat(pk, 0, s, st.g_cam, cu, cv);
    float cx = float(px) + cu, cy = float(py) + cv;
    float fu = cx / float(P.width), fv = (float(P.height) - cy) / float(P.height);
    R x = R(2) * R(fu) - R(1), y = R(2) * R(fv) - R(1);
    V3<R> fwd(cam.forward[0], cam.forward[1], cam.forward[2]);
    V3<R> right(cam.right[0], cam.right[1], cam.right[2]);
    V3<R> up(cam.up[0], cam.up[1], cam.up[2]);
    Ray<R> ray;
    ray.origin = V3<R>(cam.pos[0], cam.pos[1], cam.pos[2]);
    ray.dir = fwd + right * x * R(cam.tan_half_fov_x) + up * y * R(cam.tan_half_fov_y);
    ray.t0 = 0;
    ray.t1 = max_trace;
    cnt.eye++;
    cnt.paths++;
    R alpha;
    C3<R> L = (P.integrator == SGR_INTEG_DIRECT || P.integrator == SGR_INTEG_PHOTON)
                  ? dint(ray, pk, qk, s, alpha)
                  : Li(ray, pk, qk, s, alpha);
    out[0] = cx;
    out[1] = cy;
    out[2] = float(L.r * alpha);  // Tint(radiance, alpha) premultiplies
    out[3] = float(L.g * alpha);
    out[4] = float(L.b * alpha);
    out[5] = float(alpha);
  }
};

struct OrcScene {
  int fp64;
  std::unique_ptr<Scene<float>> sf;
  std::unique_ptr<Scene<double>> sd;
};

static void set_err(char *err, int n, const char *msg) {
  if (err && n > 0) {
    std::snprintf(err, (size_t)n, "%s", msg);
  }
}

static void parallel_for(size_t n, int nthreads,
                         const std::function<void(size_t, int)> &fn) {
  if (nthreads <= 1) {
    for (size_t i = 0; i < n; i++) fn(i, 0);
    return;
  }
  std::atomic<size_t> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&, t]() {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= n) break;
        fn(i, t);
      }
    });
  for (auto &x : th) x.join();
}

template<class R>
static void trace_impl(const Scene<R> &sc, const sgr_rays *rays, size_t n,
                       sgr_hits *hits, uint8_t *occ, int nthreads,
                       uint64_t *counters) {
  std::vector<Counters> cs((size_t)(nthreads > 0 ? nthreads : 1));
  size_t chunk = 4096;
  size_t nchunks = (n + chunk - 1) / chunk;
  parallel_for(nchunks, nthreads, [&](size_t c, int t) {
    size_t b = c * chunk, e = std::min(n, b + chunk);
    for (size_t i = b; i < e; i++) {
      Ray<R> r;
      r.origin = V3<R>(rays->ox[i], rays->oy[i], rays->oz[i]);
      r.dir = V3<R>(rays->dx[i], rays->dy[i], rays->dz[i]);
      r.t0 = rays->t0[i];
      r.t1 = rays->t1[i];
      if (hits) {
        R u = 0, v = 0;
        unsigned inst = 0xffffffffu;
        unsigned p = sc.accel.closest(r, u, v, &cs[t].nc_c, &cs[t].st_c, &inst);
        hits->prim[i] = p;
        hits->t[i] = p == 0xffffffffu ? 0.f : float(r.t1);
        hits->u[i] = float(u);
        hits->v[i] = float(v);
        if (hits->inst) hits->inst[i] = inst;
      }
      if (occ) {  // Scene::occludes from a point in air
        C3<R> trans(1);
        occ[i] = sc.occludes(r, R(1), trans, &cs[t].nc_s, &cs[t].st_s) ? 1 : 0;
      }
    }
  });
  if (counters) {
    Counters tot;
    for (auto &c : cs) tot.add(c);
    counters[0] = hits ? tot.nc_c : tot.nc_s;
    counters[1] = hits ? tot.st_c : tot.st_s;
  }
}

template<class R>
static void render_impl(const Scene<R> &sc, const sgr_camera &cam,
                        const sgr_render_params &P0, double *accum,
                        float *samples_out, int nthreads, sgr_stats *stats) {
  sgr_render_params P = P0;
  if (P.w == 0) { P.x0 = 0; P.y0 = 0; P.w = P.width; P.h = P.height; }
  int margin = (int)std::ceil(P.filter_radius - 0.5001f);  // image-filter-conv.h:39-40
  int X0 = (int)P.x0 - margin, Y0 = (int)P.y0 - margin;
  int PW = (int)P.w + 2 * margin, PH = (int)P.h + 2 * margin;
  // Renderer::render_packet: max_trace = (scene bbox + camera pos).diameter()
  BBox<R> bb = sc.bbox;
  bb.add(V3<R>(cam.pos[0], cam.pos[1], cam.pos[2]));
  R max_trace = bb.diameter();
  int nt = nthreads > 0 ? nthreads : 1;
  std::vector<Tracer<R>> tracers;
  for (int t = 0; t < nt; t++) tracers.emplace_back(sc, P);
  std::vector<R> acc((size_t)P.width * P.height * 5, R(0));
  float inv_r = 1.f / P.filter_radius;
  int fr = margin;
  const int band = 16;
  std::vector<float> buf;
  for (int yb = 0; yb < PH; yb += band) {
    int ye = std::min(PH, yb + band);
    size_t rows = (size_t)(ye - yb);
    buf.assign(rows * PW * P.spp * 6, 0.f);
    parallel_for(rows * PW, nt, [&](size_t i, int t) {
      int py = Y0 + yb + (int)(i / PW), px = X0 + (int)(i % PW);
      for (uint32_t s = 0; s < P.spp; s++)
        tracers[t].path(cam, max_trace, px, py, s, &buf[(i * P.spp + s) * 6]);
    });
    if (samples_out)
      std::memcpy(samples_out + (size_t)yb * PW * P.spp * 6, buf.data(),
                  buf.size() * sizeof(float));
    // ImageFilterConv::add_sample (image/image-filter-conv.h:109-186)
    for (size_t k = 0; k < rows * PW * P.spp; k++) {
      const float *o = &buf[k * 6];
      float sx = o[0], sy = o[1];
      int x = int(sx), y = int(sy);
      // valid_x / valid_y (image/image-sampled-output.h:155-156) of the
      // output, which for a limited render is the window alone
      // (snogray.lua:447-454): margin samples only reach window pixels
      for (int fy = -fr; fy <= fr; fy++) {
        int py = y + fy;
        if (py < (int)P.y0 || py >= (int)(P.y0 + P.h)) continue;
        for (int fx = -fr; fx <= fr; fx++) {
          int px = x + fx;
          if (px < (int)P.x0 || px >= (int)(P.x0 + P.w)) continue;
          float w = mitchell1(sx - (float(px) + 0.5f), inv_r, P.filter_b, P.filter_c) *
                    mitchell1(sy - (float(py) + 0.5f), inv_r, P.filter_b, P.filter_c);
          R ws[4] = {R(o[2]) * R(w), R(o[3]) * R(w), R(o[4]) * R(w), R(o[5]) * R(w)};
          R wr = R(w);
          if (w < 0) {
            R minc = ws[0];
            minc = mn(ws[1], minc);
            minc = mn(ws[2], minc);
            if (minc < R(P.neg_clamp)) {
              R cs = R(P.neg_clamp) / minc;
              wr *= cs;
              for (int q = 0; q < 4; q++) ws[q] *= cs;
            }
          }
          R *a = &acc[((size_t)py * P.width + px) * 5];
          for (int q = 0; q < 4; q++) a[q] += ws[q];
          a[4] += wr;
        }
      }
    }
  }
  if (accum)
    for (size_t i = 0; i < acc.size(); i++) accum[i] += double(acc[i]);
  if (stats) {
    Counters tot;
    for (auto &t : tracers) tot.add(t.cnt);
    std::memset(stats, 0, sizeof(*stats));
    stats->eye_rays = tot.eye;
    stats->scene_intersect_calls = tot.isect;
    stats->shadow_tests = tot.shadow;
    stats->node_visits_closest = tot.nc_c;
    stats->surf_tests_closest = tot.st_c;
    stats->node_visits_shadow = tot.nc_s;
    stats->surf_tests_shadow = tot.st_s;
    stats->illum_calls = tot.illum;
    stats->samples = (uint64_t)P.w * P.h * P.spp;
    stats->paths_traced = tot.paths;
    stats->media_overflows = tot.media_ovf;
    stats->direct_overflows = tot.dint_ovf;
  }
}

}  // namespace

extern "C" {

void *orc_scene_create(const sgr_scene_desc *d, int fp64, char *err, int errlen) {
  try {
    OrcScene *s = new OrcScene;
    s->fp64 = fp64;
    if (fp64) { s->sd.reset(new Scene<double>); s->sd->build(*d); }
    else { s->sf.reset(new Scene<float>); s->sf->build(*d); }
    return s;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

void orc_scene_destroy(void *s) { delete (OrcScene *)s; }

int orc_octree_export(void *h, uint32_t *nodes, uint64_t *nn, uint32_t *surfs,
                      uint64_t *ns, double *origin_size) {
  OrcScene *s = (OrcScene *)h;
  if (s->fp64) return 1;
  const Accel<float> &a = s->sf->accel;
  *nn = a.nodes.size();
  *ns = a.surfs.size();
  if (nodes)
    for (size_t i = 0; i < a.nodes.size(); i++) {
      nodes[4 * i] = a.nodes[i].first_child;
      nodes[4 * i + 1] = a.nodes[i].mask;
      nodes[4 * i + 2] = a.nodes[i].surf_begin;
      nodes[4 * i + 3] = a.nodes[i].surf_count;
    }
  if (surfs) std::memcpy(surfs, a.surfs.data(), a.surfs.size() * 4);
  if (origin_size) {
    origin_size[0] = a.origin.x; origin_size[1] = a.origin.y;
    origin_size[2] = a.origin.z; origin_size[3] = a.size;
  }
  return 0;
}

int orc_octree_stats(void *h, sgr_octree_stats *out) {
  OrcScene *s = (OrcScene *)h;
  if (s->fp64) return 1;
  OctStats st;
  if (!s->sf->accel.nodes.empty()) oct_upd_stats(s->sf->accel.nodes, 0, st);
  out->num_nodes = st.num_nodes;
  out->num_leaf_nodes = st.num_leaf_nodes;
  out->num_surfaces = st.num_surfaces;
  out->num_dup_surfaces = st.num_surfaces - s->sf->num_real;
  out->max_depth = st.max_depth;
  out->avg_depth = st.avg_depth;
  return 0;
}

int orc_scene_bounds(void *h, double *bmin, double *bmax, double *horizon) {
  OrcScene *s = (OrcScene *)h;
  if (s->fp64) {
    bmin[0] = s->sd->bbox.min.x; bmin[1] = s->sd->bbox.min.y; bmin[2] = s->sd->bbox.min.z;
    bmax[0] = s->sd->bbox.max.x; bmax[1] = s->sd->bbox.max.y; bmax[2] = s->sd->bbox.max.z;
    *horizon = s->sd->horizon;
  } else {
    bmin[0] = s->sf->bbox.min.x; bmin[1] = s->sf->bbox.min.y; bmin[2] = s->sf->bbox.min.z;
    bmax[0] = s->sf->bbox.max.x; bmax[1] = s->sf->bbox.max.y; bmax[2] = s->sf->bbox.max.z;
    *horizon = s->sf->horizon;
  }
  return 0;
}

int orc_trace(void *h, const sgr_rays *rays, size_t n, sgr_hits *hits,
              uint8_t *occ, int nthreads, uint64_t *counters) {
  OrcScene *s = (OrcScene *)h;
  if (s->fp64) trace_impl<double>(*s->sd, rays, n, hits, occ, nthreads, counters);
  else trace_impl<float>(*s->sf, rays, n, hits, occ, nthreads, counters);
  return 0;
}

int orc_render(void *h, const sgr_camera *cam, const sgr_render_params *p,
               double *accum, float *samples_out, int nthreads, sgr_stats *stats,
               char *err, int errlen) {
  try {
    OrcScene *s = (OrcScene *)h;
    if (p->spp == 0) throw std::runtime_error("spp must be > 0");
    if (s->fp64) render_impl<double>(*s->sd, *cam, *p, accum, samples_out, nthreads, stats);
    else render_impl<float>(*s->sf, *cam, *p, accum, samples_out, nthreads, stats);
    return 0;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return 1;
  }
}

// --- unit probes ----------------------------------------------------------

// BSDF probe: builds an intersection whose normal frame is the identity
// (so directions are given in the normal frame), view vector v (unit,
// v.z > 0 after the back-face flip), geometric normal gn, back flag and
// media IORs, then samples at n parameter pairs and evaluates n
// directions.  out_s: val rgb, pdf, dir xyz, flags (8 per sample);
// out_e: val rgb, pdf (4 per direction).
int orc_bsdf_probe(const sgr_material *mat, const double *v, const double *gn,
                   int back, double media_ior, double enclosing_ior,
                   const double *params, int n, double *out_s,
                   const double *dirs, int nd, double *out_e, int flags_s,
                   int flags_e) {
  Isec<float> is;
  is.v = V3<float>((float)v[0], (float)v[1], (float)v[2]);
  is.geom_n = V3<float>((float)gn[0], (float)gn[1], (float)gn[2]);
  is.back = back != 0;
  is.media_ior = (float)media_ior;
  is.enclosing_ior = (float)enclosing_ior;
  is.mat = 0;
  is.bsdf_mat = 0;
  Material<float> m{mat->type, C3<float>(mat->color[0], mat->color[1], mat->color[2]),
                    C3<float>(mat->gloss_color[0], mat->gloss_color[1], mat->gloss_color[2]),
                    mat->m, mat->ior_n, mat->ior_k, mat->underlying};
  Bsdf<float> b;
  b.init(m, is, m.ior_n);
  for (int i = 0; i < n; i++) {
    BsdfSample<float> s = b.sample((float)params[2 * i], (float)params[2 * i + 1],
                                   (unsigned)flags_s);
    double *o = &out_s[8 * i];
    o[0] = s.val.r; o[1] = s.val.g; o[2] = s.val.b; o[3] = s.pdf;
    o[4] = s.dir.x; o[5] = s.dir.y; o[6] = s.dir.z; o[7] = s.flags;
  }
  for (int i = 0; i < nd; i++) {
    V3<float> d((float)dirs[3 * i], (float)dirs[3 * i + 1], (float)dirs[3 * i + 2]);
    BsdfValue<float> e = b.eval(d, (unsigned)flags_e);
    double *o = &out_e[4 * i];
    o[0] = e.val.r; o[1] = e.val.g; o[2] = e.val.b; o[3] = e.pdf;
  }
  return 0;
}

double orc_fresnel(double n1, double n2, double k2, double cos_angle) {
  return Fresnel<float>((float)n1, (float)n2, (float)k2).reflectance((float)cos_angle);
}

double orc_mitchell(double x, double radius, double b, double c) {
  return mitchell1((float)x, 1.f / (float)radius, (float)b, (float)c);
}

// Envmap probe: for each param pair, sample the light histogram
// (u, v, pdf); then pdf_at(u, v) of the returned position.  Also returns
// the world direction and the map lookup.  out: u, v, pdf, pdf_at, dx,
// dy, dz, r, g, b (10 per sample).
int orc_envmap_probe(const float *texels, uint32_t w, uint32_t h,
                     const double *params, int n, double *out) {
  Envmap<double> E;
  E.w = w;
  E.h = h;
  E.tex.assign(texels, texels + (size_t)3 * w * h);
  E.build();
  for (int i = 0; i < n; i++) {
    float u, v, pdf;
    E.hist_sample((float)params[2 * i], (float)params[2 * i + 1], u, v, pdf);
    double *o = &out[10 * i];
    o[0] = u; o[1] = v; o[2] = pdf; o[3] = E.hist_pdf(u, v);
    V3<double> d = Envmap<double>::latlong_to_vec(u, v);
    o[4] = d.x; o[5] = d.y; o[6] = d.z;
    C3<double> c = E.map(d);
    o[7] = c.r; o[8] = c.g; o[9] = c.b;
  }
  return 0;
}

// Envmap lookup probe: the bilinear map value at n (u, v) positions of
// the lat-long parameter square (texture/matrix-linterp.h:30-70).
// out: r, g, b (3 per position).
int orc_envmap_lookup(const float *texels, uint32_t w, uint32_t h, const double *uv, int n,
                      double *out) {
  Envmap<float> E;
  E.w = w;
  E.h = h;
  E.tex.assign(texels, texels + (size_t)3 * w * h);
  for (int i = 0; i < n; i++) {
    C3<float> c = E.lookup((float)uv[2 * i], (float)uv[2 * i + 1]);
    out[3 * i] = c.r;
    out[3 * i + 1] = c.g;
    out[3 * i + 2] = c.b;
  }
  return 0;
}

// Sample-stream probe: stratified UV values of channel CHAN for all SPP
// samples of pixel (px, py), and NR plain draws of path S.
int orc_stream_probe(uint64_t seed, int px, int py, uint32_t chan, uint32_t spp,
                     float *uv_out, uint32_t s, uint32_t nr, float *r_out) {
  uint64_t pk = pixel_key(seed, px, py);
  GridShape g = grid_shape(spp);
  for (uint32_t i = 0; i < spp; i++) {
    float u, v;
    uint64_t ck = chan_key(pk, chan);
    uint32_t cell = permute(i, g.total, ck);
    uint32_t j = cell % g.u_steps, k = cell / g.u_steps;
    float us = 1.0f / (float)g.u_steps, vs = 1.0f / (float)g.v_steps;
    u = clamp01((float)j * us + draw(ck, 2 * i) * us);
    v = clamp01((float)k * vs + draw(ck, 2 * i + 1) * vs);
    uv_out[2 * i] = u;
    uv_out[2 * i + 1] = v;
  }
  uint64_t qk = path_key(pk, s);
  for (uint32_t d = 0; d < nr; d++) r_out[d] = draw(qk, d);
  return 0;
}

uint32_t orc_permute(uint32_t i, uint32_t n, uint64_t key) { return permute(i, n, key); }

}  // extern "C"

extern "C" {

// Intersection probe: for each of n rays (ox oy oz dx dy dz t0 t1), the
// closest hit and its Isec.  out (12 per ray): prim, instance (as
// doubles; -1 for none), t, normal-frame origin xyz, normal-frame z xyz,
// back flag, u, v.
int orc_isec_probe(void *h, const float *rays, int n, double *out) {
  OrcScene *s = (OrcScene *)h;
  if (s->fp64) return 1;
  const Scene<float> &sc = *s->sf;
  for (int i = 0; i < n; i++) {
    const float *r = rays + 8 * i;
    Ray<float> ray;
    ray.origin = V3<float>(r[0], r[1], r[2]);
    ray.dir = V3<float>(r[3], r[4], r[5]);
    ray.t0 = r[6];
    ray.t1 = r[7];
    float u, v;
    unsigned inst = 0xffffffffu;
    unsigned p = sc.accel.closest(ray, u, v, nullptr, nullptr, &inst);
    double *o = out + 12 * i;
    for (int k = 0; k < 12; k++) o[k] = 0;
    o[0] = p == 0xffffffffu ? -1.0 : (double)p;
    o[1] = inst == 0xffffffffu ? -1.0 : (double)inst;
    if (p == 0xffffffffu) continue;
    Isec<float> is;
    sc.make_isec(p, inst, ray, u, v, 1.0f, 1.0f, is);
    o[2] = ray.t1;
    o[3] = is.nf.origin.x; o[4] = is.nf.origin.y; o[5] = is.nf.origin.z;
    o[6] = is.nf.z.x; o[7] = is.nf.z.y; o[8] = is.nf.z.z;
    o[9] = is.back ? 1 : 0;
    o[10] = u; o[11] = v;
  }
  return 0;
}

}  // extern "C"

namespace {
// Light L of scene SC seen from points with normals (light_probe below).
template<class R>
int light_probe_t(const Scene<R> &sc, unsigned light, int mode, const double *q, int n,
                  double *out) {
  if (light >= sc.lights.size()) return 1;
  const Light<R> &L = sc.lights[light];
  for (int i = 0; i < n; i++) {
    const double *a = q + 8 * i;
    Isec<R> is;
    const V3<R> p = V3<R>(R(a[0]), R(a[1]), R(a[2]));
    const V3<R> nz = V3<R>(R(a[3]), R(a[4]), R(a[5])).unit();
    const V3<R> sx = nz.perpendicular().unit();
    const V3<R> ty = cross(nz, sx);
    is.nf = Frame<R>(p, sx, ty, nz);
    is.gf = is.nf;
    is.v = V3<R>(0, 0, 1);
    is.geom_n = V3<R>(0, 0, 1);
    is.back = false;
    is.mat = is.bsdf_mat = -1;
    is.media_ior = is.enclosing_ior = 1;
    double *o = out + 8 * i;
    LightSample<R> ls;
    if (mode == 0) {
      ls = sc.light_sample(L, is, R(a[6]), R(a[7]), R(0.5));
    } else {
      const V3<R> w = V3<R>(R(q[8 * n + 3 * i]), R(q[8 * n + 3 * i + 1]), R(q[8 * n + 3 * i + 2]));
      ls = sc.light_eval(L, is, is.nf.to(w));
    }
    o[0] = ls.val.r; o[1] = ls.val.g; o[2] = ls.val.b;
    o[3] = ls.pdf;
    o[4] = ls.dist;
    const V3<R> wdir = is.nf.from(ls.dir);
    o[5] = wdir.x; o[6] = wdir.y; o[7] = wdir.z;
  }
  return 0;
}
}  // namespace

extern "C" {

// Light probe: light LIGHT of the scene from n points.  q holds 8
// doubles per point (position xyz, normal xyz, then the sample
// parameter u v), followed, for mode 1, by 3 doubles per point (a world
// direction).  The point's normal frame is the normal with any tangent.
// Mode 0: Light::Sampler::sample at (u, v); mode 1: Light::Sampler::eval
// in the direction.  out (8 per point): value rgb, pdf (solid angle),
// distance, sample direction in world space (mode 0).
int orc_light_probe(void *h, unsigned light, int mode, const double *q, int n, double *out) {
  OrcScene *s = (OrcScene *)h;
  return s->fp64 ? light_probe_t(*s->sd, light, mode, q, n, out)
                 : light_probe_t(*s->sf, light, mode, q, n, out);
}

}  // extern "C"

namespace {
template<class R>
void photon_params_into(PhotonMaps<R> &m, const sgr_photon_params &p) {
  unsigned targets[3];
  resolve_photon_params(p, m.fg, m.fg_bsdf, m.fg_photon, targets);
  m.k = p.use_photons;
  m.radius_sq = R(p.photon_radius) * R(p.photon_radius);
  m.direct_illum = p.direct_illum != 0;
  m.direct_samples = p.direct_samples;
}
}  // namespace

extern "C" {

// Photon maps of an oracle scene (PhotonInteg::GlobalState): shot and
// built by the oracle.
void *orc_photon_map_create(void *h, const sgr_photon_params *p, char *err, int errlen) {
  try {
    OrcScene *s = (OrcScene *)h;
    OrcPhotonMap *m = new OrcPhotonMap;
    m->fp64 = s->fp64;
    unsigned fg, fb, fp, targets[3];
    resolve_photon_params(*p, fg, fb, fp, targets);
    if (s->fp64) {
      photon_params_into(m->d, *p);
      shoot_photons(*s->sd, *p, targets, m->d);
    } else {
      photon_params_into(m->f, *p);
      shoot_photons(*s->sf, *p, targets, m->f);
    }
    return m;
  } catch (const std::exception &e) {
    set_err(err, errlen, e.what());
    return nullptr;
  }
}

// Photon maps given as kd-ordered arrays (another builder's export):
// per set COUNTS[i] photons of pos xyz + split axis, power rgb, dir xyz,
// and PATHS[i], the set's path count.
void *orc_photon_map_import(const sgr_photon_params *p, const float *const *pos_axis,
                            const float *const *power, const float *const *dir,
                            const uint64_t *counts, const uint64_t *paths) {
  OrcPhotonMap *m = new OrcPhotonMap;
  photon_params_into(m->f, *p);
  photon_params_into(m->d, *p);
  for (int i = 0; i < 3; i++) {
    for (int w = 0; w < 2; w++) {
      PhotonMap<float> *mf = w == 0 ? &m->f.map[i] : nullptr;
      PhotonMap<double> *md = w == 1 ? &m->d.map[i] : nullptr;
      const size_t n = counts[i];
      if (mf) { mf->photons.resize(n); mf->axes.resize(n); }
      if (md) { md->photons.resize(n); md->axes.resize(n); }
      for (size_t k = 0; k < n; k++) {
        const float *a = pos_axis[i] + 4 * k, *c = power[i] + 3 * k, *d = dir[i] + 3 * k;
        unsigned ax;
        std::memcpy(&ax, &a[3], 4);
        if (mf) {
          mf->photons[k] = Photon<float>{V3<float>(a[0], a[1], a[2]), C3<float>(c[0], c[1], c[2]),
                                         V3<float>(d[0], d[1], d[2])};
          mf->axes[k] = (unsigned char)ax;
        } else {
          md->photons[k] = Photon<double>{V3<double>(a[0], a[1], a[2]),
                                          C3<double>(c[0], c[1], c[2]), V3<double>(d[0], d[1], d[2])};
          md->axes[k] = (unsigned char)ax;
        }
      }
    }
    m->f.paths[i] = m->d.paths[i] = paths[i];
    m->f.scale[i] = m->d.scale[i] = paths[i] > 0 ? 1 / float(paths[i]) : 0.f;
  }
  return m;
}

int orc_photon_map_info(void *h, sgr_photon_map_info *info) {
  OrcPhotonMap *m = (OrcPhotonMap *)h;
  std::memset(info, 0, sizeof *info);
  auto fill = [&](auto &pm) {
    for (int i = 0; i < 3; i++) {
      info->photons[i] = pm.map[i].photons.size();
      info->paths[i] = pm.paths[i];
      info->scale[i] = pm.scale[i];
    }
    info->paths_shot = pm.paths_shot;
    info->deposit_overflows = pm.overflows;
    info->fg_samples = pm.fg;
    info->fg_bsdf_samples = pm.fg_bsdf;
    info->fg_photon_samples = pm.fg_photon;
  };
  if (m->fp64) fill(m->d);
  else fill(m->f);
  return 0;
}

int orc_photon_map_export(void *h, int set, float *pos_axis, float *power, float *dir) {
  OrcPhotonMap *m = (OrcPhotonMap *)h;
  if (set < 0 || set > 2) return 1;
  auto put = [&](auto &pm) {
    const auto &mp = pm.map[set];
    for (size_t k = 0; k < mp.photons.size(); k++) {
      const auto &ph = mp.photons[k];
      const unsigned ax = mp.axes[k];
      if (pos_axis) {
        pos_axis[4 * k] = float(ph.pos.x);
        pos_axis[4 * k + 1] = float(ph.pos.y);
        pos_axis[4 * k + 2] = float(ph.pos.z);
        std::memcpy(&pos_axis[4 * k + 3], &ax, 4);
      }
      if (power) {
        power[3 * k] = float(ph.power.r);
        power[3 * k + 1] = float(ph.power.g);
        power[3 * k + 2] = float(ph.power.b);
      }
      if (dir) {
        dir[3 * k] = float(ph.dir.x);
        dir[3 * k + 1] = float(ph.dir.y);
        dir[3 * k + 2] = float(ph.dir.z);
      }
    }
  };
  if (m->fp64) put(m->d);
  else put(m->f);
  return 0;
}

void orc_photon_map_destroy(void *h) { delete (OrcPhotonMap *)h; }

}  // extern "C"

extern "C" {

double orc_radical_inverse(unsigned num, unsigned base) { return radical_inverse(num, base); }

// kd-tree probe: builds a PhotonMap of n photons (positions xyz; power
// and direction zero) and returns its heap-order positions, split axes
// and, for m query points, the photons find() keeps (kd indices, -1
// padded, k per query) and the final squared radius.
int orc_kd_probe(const float *pts, int n, const float *queries, int m, unsigned k, float r2,
                 float *kd_pos, int *kd_axis, int *found, float *final_r2) {
  std::vector<Photon<float>> v(n);
  for (int i = 0; i < n; i++) v[i] = Photon<float>{V3<float>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), C3<float>(0), V3<float>()};
  PhotonMap<float> pm;
  pm.set_photons(v);
  for (int i = 0; i < n; i++) {
    kd_pos[3 * i] = pm.photons[i].pos.x;
    kd_pos[3 * i + 1] = pm.photons[i].pos.y;
    kd_pos[3 * i + 2] = pm.photons[i].pos.z;
    kd_axis[i] = pm.axes[i];
  }
  std::vector<const Photon<float> *> heap;
  for (int q = 0; q < m; q++) {
    heap.clear();
    const V3<float> pos(queries[3 * q], queries[3 * q + 1], queries[3 * q + 2]);
    final_r2[q] = pm.find(pos, k, r2, heap);
    for (unsigned j = 0; j < k; j++)
      found[(size_t)q * k + j] = j < heap.size() ? int(heap[j] - pm.photons.data()) : -1;
  }
  return 0;
}

// Direction-histogram probe: bins n world directions with weights, then
// for m parameter pairs samples the distribution (direction xyz, pdf per
// steradian) and evaluates pdf() at the sampled direction (5 per
// sample); out_bins gets the 64 bin pdfs (per unit parameter area).
int orc_dirhist_probe(const float *dirs, const float *w, int n, const float *params, int m,
                      float *out, float *out_bins) {
  DirHistDist dd;
  dd.clear();
  for (int i = 0; i < n; i++) {
    double u, v;
    Scene<double>::sphere_sample_inverse(V3<double>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), u, v);
    float fu, fv;
    Scene<float>::sphere_sample_inverse(V3<float>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), fu, fv);
    (void)u;
    (void)v;
    dd.add(fu, fv, w[i]);
  }
  dd.calc();
  for (int b = 0; b < 64; b++) out_bins[b] = dd.bin_pdf(b % 8, b / 8);
  for (int i = 0; i < m; i++) {
    float su, sv, pdf;
    dd.sample(params[2 * i], params[2 * i + 1], su, sv, pdf);
    pdf *= 0.25f * INV_PIf;
    const V3<float> d = Scene<float>::sphere_sample(su, sv);
    float du, dv;
    Scene<float>::sphere_sample_inverse(d, du, dv);
    out[5 * i] = d.x;
    out[5 * i + 1] = d.y;
    out[5 * i + 2] = d.z;
    out[5 * i + 3] = pdf;
    out[5 * i + 4] = dd.pdf(du, dv) * 0.25f * INV_PIf;
  }
  return 0;
}

}  // extern "C"
