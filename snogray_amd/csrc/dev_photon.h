// Device photon mapping (render/photon-integ.cc, photon/*.cc): free light
// samples for shooting photons, the photon maps' kd-tree search, the
// Gaussian-filtered density estimate, the photon-direction histogram and
// final gathering, in the reference's fp32 arithmetic.
#pragma once
#include "dev_shade.h"
#include "kernels.h"

namespace sgr {

#define SGR_PHOTON_MAX_K 128   // photons per estimate the device keeps
#define SGR_PHOTON_KD_STACK 48 // kd-tree levels (2^48 photons)
#define SGR_PHOTON_GAUSS_SCALE 2.90898f  // photon/photon-eval.cc:83

// util/radical-inverse.h:27-43 (in double, as the reference computes it)
__device__ __forceinline__ double radical_inverse(uint32_t num, uint32_t base) {
  double val = 0;
  const double inv_base = 1 / double(base);
  double inv_bi = inv_base;
  while (num > 0) {
    const uint32_t d = num % base;
    val += d * inv_bi;
    num /= base;
    inv_bi *= inv_base;
  }
  return val;
}

// Vec::perpendicular (geometry/vec.h:119-129)
__device__ __forceinline__ V3 perpendicular(V3 n) {
  const float nz = (n.x == n.z) ? n.z : -n.z;
  return v3(n.y * n.y + n.x * nz, -nz * nz - n.x * n.y, n.x * n.x - n.y * nz);
}
// Frame (org, z) (geometry/frame.h:64-84)
__device__ __forceinline__ Frame frame_of(V3 org, V3 z) {
  const V3 x = unit(perpendicular(z));
  const V3 y = unit(cross(z, x));
  return Frame{org, x, y, z};
}

struct FreeS {
  C3 val;
  float pdf;
  V3 pos, dir;
};

// Light::Sampler::sample (param, dir_param): SurfaceLightSampler
// (light/surface-light-sampler.cc:72-100: an area sample, a cosine-
// distributed direction about its normal, pdf = area pdf / pi),
// SphereLightSampler (light/sphere-light-sampler.cc:120-155) and
// EnvmapLight::Sampler (light/envmap-light.cc:189-230: a direction from
// the intensity distribution, a start point on the disk tangent to the
// scene's bounding sphere, the map looked up with the world-space
// direction as the reference passes it).
template<bool EXT>
__device__ __forceinline__ FreeS light_free_sample(const SceneView &S, const DevLight &L, float pu,
                                                   float pv, float du, float dv) {
  FreeS o;
  o.val = c3(0.0f);
  o.pdf = 0.0f;
  if (EXT && L.kind == 6) {
    // PointLight::Sampler::sample (light/point-light.cc:115-126): a
    // direction by cone_sample and its pdf (geometry/cone-sample.h:29-77),
    // into the light's frame unless it radiates in all directions
    const float4 *pl = S.plights + 4 * L.prim;
    const float ch = pl[0].w;
    const float z = ch + dv * (1.0f - ch);
    const float r = sqrtf(fabsf(1.0f - z * z));
    const float phi = du * 2.0f * SGR_PIf;
    V3 d = v3(r * cosf(phi), r * sinf(phi), z);
    o.pdf = 0.5f * SGR_INV_PIf / (1.0f - ch);
    o.val = point_intensity(cload(L.intensity), ch, pl[1].w, d.z);
    if (ch != -1.0f) d = Frame{ld3(pl[0]), ld3(pl[1]), ld3(pl[2]), ld3(pl[3])}.from(d);
    o.pos = ld3(pl[0]);
    o.dir = d;
    return o;
  }
  if (L.kind == 2) {
    float mu, mv, pdf;
    env_hist_sample(S, du, dv, mu, mv, pdf);
    const V3 wd = env_from(S, latlong_to_vec(mu, mv));
    float p = pdf * 0.25f * SGR_INV_PIf;
    const V3 lo = v3(S.bbox_lo[0], S.bbox_lo[1], S.bbox_lo[2]);
    const V3 ext = v3(S.bbox_hi[0], S.bbox_hi[1], S.bbox_hi[2]) - lo;
    const V3 cen = lo + ext * 0.5f;     // geometry/bbox.h:84-87
    const float rad = vlen(ext) * 0.5f;
    float px, py;
    disk_sample(rad, pu, pv, px, py);  // geometry/tangent-disk-sample.h:31-51
    o.pos = cen + frame_of(cen, wd).from(v3(px, py, rad));
    p = p / ((rad * rad) * SGR_PIf);
    o.val = env_map(S, wd);
    o.pdf = p;
    o.dir = -wd;
    return o;
  }
  V3 pos, norm;
  float apdf;
  if (L.kind == 1) {
    const float4 sp = S.sph[L.prim - S.n_tri_like];
    const V3 spv = sphere_sample(pu, pv);
    pos = ld3(sp) + spv * sp.w;
    apdf = 1.0f / (sp.w * sp.w * (float)(4 * SGR_PI_D));
    norm = spv;
  } else if (L.kind == 0) {
    const float4 a = S.tri[3 * L.prim], b = S.tri[3 * L.prim + 1], c = S.tri[3 * L.prim + 2];
    tripar_area(ld3(a), ld3(b), ld3(c), a.w != 0.0f, pu, pv, pos, norm, apdf);
  } else if (EXT && L.kind == 3) {
    const float4 a = S.tri[3 * L.prim], b = S.tri[3 * L.prim + 1], c = S.tri[3 * L.prim + 2];
    ellipse_area(ld3(a), ld3(b), ld3(c), pu, pv, pos, norm, apdf);
  } else if (EXT) {  // Cylinder / Sphere2::Sampler::sample (surface/cylinder.cc:225-233, sphere2.cc:225-232)
    const float4 *lr = S.loc + SGR_LOC_STRIDE * (L.prim - S.n_sph_end);
    const bool cyl = L.kind == 4;
    V3 ln;
    if (cyl) {
      const float th = pu * 2.0f * SGR_PIf;
      ln = v3(cosf(th), sinf(th), 0.0f);
    } else {
      ln = sphere_sample(pu, pv);
    }
    norm = unit(w2l_transpose(lr, ln));
    apdf = approx_area_pdf(lr, cyl, pu, pv, pos);
  } else {
    return o;
  }
  const V3 d = z_normal_symm_vec(sqrtf(du), dv);  // CosDist (material/cos-dist.h:24-44)
  o.pos = pos;
  o.dir = frame_of(v3(0.0f, 0.0f, 0.0f), norm).from(d);
  o.pdf = apdf * SGR_INV_PIf;
  o.val = cload(L.intensity);
  return o;
}

// Bsdf::supports with the stencil (material/stencil.cc:114-123: the
// underlying BSDF's flags when some opacity, straight-through
// transmission when not fully opaque, the latter not masked by LIMIT).
__device__ __forceinline__ unsigned bsdf_supports_any(const DevMaterial &M, unsigned limit) {
  if (M.type != 8) return bsdf_supports(M, limit);
  const C3 op = cload(M.color);
  unsigned f = 0;
  if (cgt0(op) && M.has_under) f |= bsdf_supports((&M)[M.under_off], limit);
  if (csum(op) < 3.0f) f |= F_TRANSMISSIVE | F_SPECULAR;
  return f;
}

// The photons found near a point: a max-heap on squared distance.
struct PhotonHeap {
  float d[SGR_PHOTON_MAX_K];
  uint32_t i[SGR_PHOTON_MAX_K];
  uint32_t n;
};

__device__ __forceinline__ float axis_coord(V3 p, uint32_t a) {
  return a == 0 ? p.x : (a == 1 ? p.y : p.z);
}

// PhotonMap::find_photons (photon/photon-map.cc:214-300), without
// recursion: a node is searched after its children, first the child on
// POS's side of the split, then the other while the split plane lies
// within the search radius; only nodes with both children descend (2i +
// 2 < n), so a last node with a left child alone never has it searched,
// as in the reference.  The K nearest photons within sqrt(MAXD) are kept;
// once K are held the radius shrinks to the farthest.  Returns the final
// squared radius.
__device__ __forceinline__ float photon_find(const PhotonSetView &M, V3 pos, uint32_t k,
                                             float maxd, PhotonHeap &H) {
  H.n = 0;
  if (M.n == 0) return maxd;
  uint32_t stack[SGR_PHOTON_KD_STACK];
  int sp = 0;
  uint32_t node = 0, phase = 0;
  for (;;) {
    const float4 pa = M.pos[node];
    const uint32_t ax = __float_as_uint(pa.w);
    const bool inner = (uint64_t)node * 2 + 2 < M.n;
    if (inner && phase < 2) {
      const float sd = axis_coord(pos, ax) - axis_coord(ld3(pa), ax);
      if (phase == 0) {
        stack[sp++] = node << 2 | 1u;
        node = node * 2 + ((sd < 0.0f) ? 1u : 2u);
        continue;
      }
      if (sd * sd < maxd) {
        stack[sp++] = node << 2 | 2u;
        node = node * 2 + ((sd < 0.0f) ? 2u : 1u);
        phase = 0;
        continue;
      }
    }
    const float d2 = len2(pos - ld3(pa));
    if (d2 < maxd) {
      if (H.n == k) {  // replace the farthest, sift down
        uint32_t h = 0;
        for (;;) {
          const uint32_t l = 2 * h + 1, r = l + 1;
          uint32_t m = h;
          float dm = d2;
          if (l < H.n && H.d[l] > dm) { m = l; dm = H.d[l]; }
          if (r < H.n && H.d[r] > dm) { m = r; dm = H.d[r]; }
          if (m == h) break;
          H.d[h] = H.d[m];
          H.i[h] = H.i[m];
          h = m;
        }
        H.d[h] = d2;
        H.i[h] = node;
      } else {  // push, sift up
        uint32_t h = H.n++;
        while (h > 0) {
          const uint32_t par = (h - 1) / 2;
          if (!(H.d[par] < d2)) break;
          H.d[h] = H.d[par];
          H.i[h] = H.i[par];
          h = par;
        }
        H.d[h] = d2;
        H.i[h] = node;
      }
      if (H.n == k) maxd = H.d[0];
    }
    if (sp == 0) break;
    const uint32_t e = stack[--sp];
    node = e >> 2;
    phase = e & 3u;
  }
  return maxd;
}

// PhotonEval::Lo (photon/photon-eval.cc:44-118): each photon found near
// the hit weighted by the BSDF in its direction (FLAGS) and by the
// Gaussian filter of its distance (util/gaussian-filter.h:28-97, alpha 2
// over the final radius), times 2.90898 and the set's scale, over pi r^2.
__device__ __forceinline__ C3 photon_lo(const Isec &is, const DevMaterial &M,
                                        const PhotonSetView &PV, uint32_t k, float r2,
                                        unsigned flags, PhotonHeap &H) {
  if (PV.scale == 0.0f) return c3(0.0f);
  if (!bsdf_supports_any(M, flags & ~F_SPECULAR)) return c3(0.0f);
  const float maxd = photon_find(PV, is.nf.o, k, r2, H);
  if (H.n == 0) return c3(0.0f);
  const float alpha = 2.0f / maxd;
  const float edge = expf(-alpha * maxd);
  const float fscale = 1.0f / (1.0f - edge);
  C3 rad = c3(0.0f);
  for (uint32_t j = 0; j < H.n; j++) {
    const uint32_t q = H.i[j];
    const BValue v = bsdf_eval(is, M, is.nf.to(ld3(PV.dir[q])), flags);
    if (v.pdf != 0.0f && cgt0(v.val)) {
      const float d2 = len2(ld3(PV.pos[q]) - is.nf.o);
      const float filt = d2 >= maxd ? 0.0f : (expf(-alpha * d2) - edge) * fscale;
      const float4 pw = PV.pow[q];
      rad = rad + ((v.val * C3{pw.x, pw.y, pw.z}) * filt) * SGR_PHOTON_GAUSS_SCALE;
    }
  }
  rad = rad * PV.scale;
  return rad * (1.0f / (maxd * SGR_PIf));
}

// DirHist 8 x 8 with Hist2dDist and DirHistDist (geometry/dir-hist.h,
// geometry/hist-2d.h, geometry/hist-2d-dist.cc:38-272,
// geometry/dir-hist-dist.h): directions binned by their sphere-sample
// parameters, row and in-row cumulative sums summed in double and
// stored in float, sampling by lower_bound, pdf per steradian.
#define SGR_DH_W 8u
#define SGR_DH_H 8u
struct DirHistDist {
  float bins[SGR_DH_W * SGR_DH_H];
  float row_cum[SGR_DH_H];
  float col_cum[SGR_DH_W * SGR_DH_H];
};
__device__ __forceinline__ void dirhist_build(DirHistDist &D, const PhotonSetView &PV,
                                              const PhotonHeap &H) {
  for (uint32_t b = 0; b < SGR_DH_W * SGR_DH_H; b++) D.bins[b] = 0.0f;
  for (uint32_t j = 0; j < H.n; j++) {
    const uint32_t q = H.i[j];
    float u, v;
    sphere_sample_inverse(ld3(PV.dir[q]), u, v);
    const float4 pw = PV.pow[q];
    float s = pw.x;
    s += pw.y;
    s += pw.z;
    uint32_t col = (uint32_t)(u * (float)SGR_DH_W), row = (uint32_t)(v * (float)SGR_DH_H);
    col = col < SGR_DH_W - 1 ? col : SGR_DH_W - 1;
    row = row < SGR_DH_H - 1 ? row : SGR_DH_H - 1;
    D.bins[col + row * SGR_DH_W] += s / 3.0f;  // Color::intensity
  }
  double sum = 0;
  for (uint32_t b = 0; b < SGR_DH_W * SGR_DH_H; b++) sum += (double)D.bins[b];
  const double inv = sum == 0 ? 0 : 1 / sum;
  double norm = 0;
  for (uint32_t r = 0; r < SGR_DH_H; r++) {
    for (uint32_t c = 0; c < SGR_DH_W; c++) norm += (double)D.bins[r * SGR_DH_W + c] * inv;
    D.row_cum[r] = (float)norm;
  }
  for (uint32_t r = 0; r < SGR_DH_H; r++) {
    double rs = 0;
    for (uint32_t c = 0; c < SGR_DH_W; c++) rs += (double)D.bins[r * SGR_DH_W + c];
    const double ir = rs == 0 ? 0 : 1 / rs;
    double nr = 0;
    for (uint32_t c = 0; c < SGR_DH_W; c++) {
      nr += (double)D.bins[r * SGR_DH_W + c] * ir;
      D.col_cum[r * SGR_DH_W + c] = (float)nr;
    }
  }
}
__device__ __forceinline__ float dirhist_bin_pdf(const DirHistDist &D, uint32_t col, uint32_t row) {
  float rp = D.row_cum[row];
  if (row != 0) rp -= D.row_cum[row - 1];
  float cp = D.col_cum[row * SGR_DH_W + col];
  if (col != 0) cp -= D.col_cum[row * SGR_DH_W + col - 1];
  return rp * cp * (float)(SGR_DH_W * SGR_DH_H);
}
// Hist2dDist::sample then DirHistDist::sample: a world direction and its
// pdf per steradian (0 when every bin is zero).
__device__ __forceinline__ V3 dirhist_sample(const DirHistDist &D, float pu, float pv, float &pdf) {
  const float u = fmn(pu, 1.0f), v = fmn(pv, 1.0f);
  const uint32_t row = lower_bound_f(D.row_cum, SGR_DH_H, v);
  float ou = 0.0f, ov = 0.0f;
  pdf = 0.0f;
  if (row != SGR_DH_H) {
    const uint32_t col = lower_bound_f(D.col_cum + row * SGR_DH_W, SGR_DH_W, u);
    pdf = dirhist_bin_pdf(D, col, row);
    const float cw = 1.0f / (float)SGR_DH_W, rh = 1.0f / (float)SGR_DH_H;
    ou = (float)col * cw + fmodf(pu, cw);
    ov = (float)row * rh + fmodf(pv, rh);
  }
  pdf *= 0.25f * SGR_INV_PIf;
  return sphere_sample(ou, ov);
}
__device__ __forceinline__ float dirhist_pdf(const DirHistDist &D, V3 wdir) {
  float u, v;
  sphere_sample_inverse(wdir, u, v);
  int col = (int)(u * (float)SGR_DH_W), row = (int)(v * (float)SGR_DH_H);
  col = col < (int)SGR_DH_W - 1 ? col : (int)SGR_DH_W - 1;
  col = col > 0 ? col : 0;
  row = row < (int)SGR_DH_H - 1 ? row : (int)SGR_DH_H - 1;
  row = row > 0 ? row : 0;
  return dirhist_bin_pdf(D, (uint32_t)col, (uint32_t)row) * 0.25f * SGR_INV_PIf;
}

// render/mis-sample-weight.h:45-77 with sample counts
__device__ __forceinline__ float mis_sample_weight(float pdf, float num, float other_pdf,
                                                   float num_other) {
  const float term = num * pdf, other = num_other * other_pdf;
  if (term > 1e10f) return 1.0f;
  if (other > 1e10f) return 0.0f;
  const float t2 = term * term, o2 = other * other;
  return (term * pdf) / (t2 + o2);
}

// Color / Color, zero denominators giving 0 (color/color.h:294-300)
__device__ __forceinline__ C3 cdiv(C3 a, C3 b) {
  return C3{b.r == 0.0f ? 0.0f : a.r / b.r, b.g == 0.0f ? 0.0f : a.g / b.g,
            b.b == 0.0f ? 0.0f : a.b / b.b};
}

}  // namespace sgr
