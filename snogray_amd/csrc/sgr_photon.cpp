// Photon maps (render/photon-integ.cc:36-242, photon/photon-shooter.cc,
// photon/photon-map.cc): photon paths shot on the device in batches, the
// deposits kept in path order on the host until each set holds its target
// count, then each set put in a left-balanced kd-tree and uploaded.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "sgr_internal.h"

// photons per estimate the device keeps (dev_photon.h)
#define SGR_PHOTON_MAX_K 128

using namespace sgr;

namespace {

// render/photon-integ.cc:36-148: the final-gathering sample split, and
// the direct set's target (none with direct lighting on and no final
// gathering).  TARGETS: direct, caustic, indirect.
void resolve_params(const sgr_photon_params &p, uint32_t &fg, uint32_t &fg_bsdf,
                    uint32_t &fg_photon, uint32_t targets[3]) {
  const uint32_t U = SGR_PHOTON_UNSPEC;
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
        fg_photon = fg - fg_bsdf;  // as the reference assigns it (fg_bsdf stays unset)
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

// One photon as deposited: position + set bits, power, direction.
struct Rec {
  float4 pos, pow, dir;
};

// photon/photon-map.cc:37-92: nodes in the left child of a
// left-balanced tree of NUM nodes.
uint32_t left_child_nodes(uint32_t num) {
  uint32_t max_bottom = 1;
  while (max_bottom * 2 <= num) max_bottom += max_bottom;
  const uint32_t non_bottom = max_bottom - 1;
  const uint32_t bottom = num - non_bottom;
  const uint32_t left_max_bottom = max_bottom / 2;
  const uint32_t left_non_bottom = (non_bottom - 1) / 2;
  return left_non_bottom + std::min(bottom, left_max_bottom);
}

inline float coord(const float4 &p, uint32_t a) { return a == 0 ? p.x : a == 1 ? p.y : p.z; }

// PhotonMap::make_kdtree (photon/photon-map.cc:112-189): the range's
// median (the left-balanced split position) on the axis of its widest
// box at node TARGET, the two sides below it.  Iterative over a work list.
void make_kdtree(std::vector<Rec> &v, std::vector<Rec> &out) {
  out.resize(v.size());
  if (v.empty()) return;
  struct Job { size_t beg, end; uint32_t target; };
  std::vector<Job> work{{0, v.size(), 0}};
  while (!work.empty()) {
    const Job j = work.back();
    work.pop_back();
    size_t median = j.beg;
    uint32_t axis = 0;
    if (j.beg + 1 != j.end) {
      float lo[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
      float hi[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
      for (size_t i = j.beg; i < j.end; i++)
        for (uint32_t a = 0; a < 3; a++) {  // BBox::add (geometry/bbox.h)
          const float c = coord(v[i].pos, a);
          lo[a] = c < lo[a] ? c : lo[a];
          hi[a] = c > hi[a] ? c : hi[a];
        }
      float max_dim = 0;
      for (uint32_t a = 0; a < 3; a++) {
        const float d = hi[a] - lo[a];
        if (d > max_dim) {
          max_dim = d;
          axis = a;
        }
      }
      median += left_child_nodes((uint32_t)(j.end - j.beg));
      std::nth_element(v.begin() + j.beg, v.begin() + median, v.begin() + j.end,
                       [axis](const Rec &a, const Rec &b) { return coord(a.pos, axis) < coord(b.pos, axis); });
      if (median != j.beg) work.push_back({j.beg, median, j.target * 2 + 1});
      if (median + 1 != j.end) work.push_back({median + 1, j.end, j.target * 2 + 2});
    }
    Rec r = v[median];
    std::memcpy(&r.pos.w, &axis, 4);
    out[j.target] = r;
  }
}

}  // namespace

struct sgr_photon_map {
  sgr_scene *scene = nullptr;
  std::vector<void *> allocs;
  uint64_t bytes = 0;
  PhotonView view{};              // device sets and parameters (grids per render)
  uint32_t direct_samples = 16;
  std::vector<Rec> host[3];       // kd-ordered copies for export
  sgr_photon_map_info info{};
};

namespace {

int map_alloc(sgr_photon_map *m, size_t bytes, void **out) {
  sgr_ctx *ctx = m->scene->ctx;
  void *p = nullptr;
  if (bytes == 0) bytes = 16;
  if (hipMalloc(&p, bytes) != hipSuccess)
    return fail(SGR_ERR_NOMEM, "sgr_photon_map_create: device allocation of " +
                                   std::to_string(bytes) + " bytes failed");
  m->allocs.push_back(p);
  m->bytes += bytes;
  ctx->bytes += bytes;
  *out = p;
  return SGR_OK;
}

void map_free(sgr_photon_map *m) {
  for (void *p : m->allocs) (void)hipFree(p);
  if (m->scene) {
    m->scene->ctx->bytes -= m->bytes;
    m->scene->photon_maps--;
  }
  m->allocs.clear();
  m->bytes = 0;
}

// Device buffers of one shooting batch (freed at the end of the build).
struct Batch {
  uint2 *rec = nullptr;
  float4 *dep = nullptr;
  uint32_t *counter = nullptr;
  size_t n_rec = 0, cap = 0;
  ~Batch() {
    if (rec) (void)hipFree(rec);
    if (dep) (void)hipFree(dep);
    if (counter) (void)hipFree(counter);
  }
};

int shoot(sgr_photon_map *m, const sgr_photon_params &pp, const uint32_t targets[3],
          std::vector<Rec> sets[3]) {
  sgr_scene *sc = m->scene;
  sgr_ctx *ctx = sc->ctx;
  uint64_t num_paths[3] = {0, 0, 0};
  auto complete = [&](int i) { return sets[i].size() == targets[i]; };
  auto all_complete = [&]() { return complete(0) && complete(1) && complete(2); };
  m->info.paths_shot = 0;
  if (sc->n_lights == 0 || all_complete()) return SGR_OK;
  Batch B;
  std::vector<uint2> rec;
  std::vector<float4> dep;
  uint32_t path = 0;
  uint32_t batch = 1u << 16;
  bool done = false;
  hipEvent_t e0 = ctx->ev[0], e1 = ctx->ev[1];
  HIPCHK(hipEventRecord(e0, ctx->stream));
  while (!done) {
    const uint32_t n = batch;
    if (B.n_rec < n) {
      if (B.rec) (void)hipFree(B.rec);
      B.rec = nullptr;
      HIPCHK(hipMalloc(&B.rec, sizeof(uint2) * n));
      B.n_rec = n;
    }
    if (!B.counter) HIPCHK(hipMalloc(&B.counter, sizeof(uint32_t)));
    if (B.cap < (size_t)n * 2) {
      if (B.dep) (void)hipFree(B.dep);
      B.dep = nullptr;
      B.cap = (size_t)n * 2;
      HIPCHK(hipMalloc(&B.dep, sizeof(float4) * 3 * B.cap));
    }
    uint32_t used = 0;
    for (;;) {  // a batch whose deposits outgrow the array runs again with room for them
      HIPCHK(hipMemsetAsync(B.counter, 0, sizeof(uint32_t), ctx->stream));
      PhotonShootView V;
      V.seed = pp.seed;
      V.base = path;
      V.n = n;
      V.caustic_on = targets[1] != 0 ? 1u : 0u;
      V.rec = B.rec;
      V.dep = B.dep;
      V.cap = (uint32_t)B.cap;
      V.counter = B.counter;
      unsigned grid = (n + 255) / 256;
      grid = grid < 64u * launch_cus() ? grid : 64u * launch_cus();
      launch_photon_shoot(sc->view, V, ctx->stats, grid, ctx->stream);
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&used, B.counter, sizeof used, hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
      if (used <= B.cap) break;
      (void)hipFree(B.dep);
      B.dep = nullptr;
      B.cap = (size_t)used + used / 4;
      HIPCHK(hipMalloc(&B.dep, sizeof(float4) * 3 * B.cap));
    }
    rec.resize(n);
    dep.resize((size_t)used * 3);
    HIPCHK(hipMemcpy(rec.data(), B.rec, sizeof(uint2) * n, hipMemcpyDeviceToHost));
    if (used) HIPCHK(hipMemcpy(dep.data(), B.dep, sizeof(float4) * 3 * used, hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < n && !done; t++) {
      const uint32_t p = path + t;
      const uint2 r = rec[t];
      if (!(r.y >> 31)) continue;  // no light sample: not a path of any set
      for (int i = 0; i < 3; i++)
        if (!complete(i)) num_paths[i]++;
      const uint32_t attempts = r.y & 0x7fffffffu;
      const uint32_t stored = std::min(attempts, (uint32_t)SGR_PHOTON_MAX_DEPOSITS);
      m->info.deposit_overflows += attempts - stored;
      for (uint32_t q = 0; q < stored; q++) {
        const float4 *d = &dep[3 * ((size_t)r.x + q)];
        uint32_t set;
        std::memcpy(&set, &d[0].w, 4);
        if (set < 3 && !complete((int)set)) sets[set].push_back(Rec{d[0], d[1], d[2]});
      }
      m->info.paths_shot = (uint64_t)p + 1;
      if (all_complete() || p > 100000000u) done = true;  // photon/photon-shooter.cc:186-187
    }
    path += n;
    // photon/photon-shooter.cc:186-187 stops after a valid path past 1e8;
    // with no valid light sample at all the reference would never stop
    if (path > 100000000u + (1u << 22)) done = true;
    batch = std::min(batch * 4u, 1u << 22);
  }
  HIPCHK(hipEventRecord(e1, ctx->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  m->info.ms_shoot = ms;
  for (int i = 0; i < 3; i++) m->info.paths[i] = num_paths[i];
  return SGR_OK;
}

int create_impl(sgr_scene *sc, const sgr_photon_params *pp, sgr_photon_map **out,
                sgr_photon_map *&made) {
  if (!sc || !pp || !out) return fail(SGR_ERR_INVALID, "sgr_photon_map_create: null argument");
  if (pp->use_photons == 0 || pp->use_photons > SGR_PHOTON_MAX_K)
    return fail(SGR_ERR_LIMIT, "sgr_photon_map_create: use_photons must be 1.." +
                                   std::to_string(SGR_PHOTON_MAX_K));
  if (!(pp->photon_radius > 0.0f) || !std::isfinite(pp->photon_radius))
    return fail(SGR_ERR_INVALID, "sgr_photon_map_create: photon_radius must be positive");
  uint32_t fg, fgb, fgp, targets[3];
  resolve_params(*pp, fg, fgb, fgp, targets);
  if (fgb > 4096 || fgp > 4096)
    return fail(SGR_ERR_LIMIT, "sgr_photon_map_create: more than 4096 final-gathering samples");
  HIPCHK(hipSetDevice(sc->ctx->device));
  sgr_photon_map *m = new sgr_photon_map;
  made = m;
  m->scene = sc;
  sc->photon_maps++;
  m->direct_samples = pp->direct_samples ? pp->direct_samples : 16u;
  m->info.fg_samples = fg;
  m->info.fg_bsdf_samples = fgb;
  m->info.fg_photon_samples = fgp;
  std::vector<Rec> sets[3];
  int rc = shoot(m, *pp, targets, sets);
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  PhotonView &PV = m->view;
  for (int i = 0; i < 3; i++) {
    make_kdtree(sets[i], m->host[i]);
    sets[i].clear();
    sets[i].shrink_to_fit();
    const size_t n = m->host[i].size();
    std::vector<float4> pos(n), pw(n), dir(n);
    for (size_t k = 0; k < n; k++) {
      pos[k] = m->host[i][k].pos;
      pw[k] = m->host[i][k].pow;
      dir[k] = m->host[i][k].dir;
    }
    void *a = nullptr, *b = nullptr, *c = nullptr;
    rc = map_alloc(m, n * sizeof(float4), &a);
    if (!rc) rc = map_alloc(m, n * sizeof(float4), &b);
    if (!rc) rc = map_alloc(m, n * sizeof(float4), &c);
    if (rc) return rc;
    if (n) {
      HIPCHK(hipMemcpy(a, pos.data(), n * sizeof(float4), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(b, pw.data(), n * sizeof(float4), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c, dir.data(), n * sizeof(float4), hipMemcpyHostToDevice));
    }
    PV.set[i].pos = (const float4 *)a;
    PV.set[i].pow = (const float4 *)b;
    PV.set[i].dir = (const float4 *)c;
    PV.set[i].n = (uint32_t)n;
    const uint64_t np = m->info.paths[i];
    PV.set[i].scale = np > 0 ? 1 / float(np) : 0.0f;  // render/photon-integ.cc:234-241
    m->info.photons[i] = n;
    m->info.scale[i] = PV.set[i].scale;
  }
  m->info.ms_build =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  PV.k = pp->use_photons;
  PV.r2 = pp->photon_radius * pp->photon_radius;
  PV.fg = fg;
  PV.fg_bsdf = fgb;
  PV.fg_photon = fgp;
  PV.direct_illum = pp->direct_illum ? 1u : 0u;
  PV.caustics = m->host[1].empty() ? 0u : 1u;
  *out = m;
  made = nullptr;
  return SGR_OK;
}

}  // namespace

namespace sgr {
// The photon integrator's view of MAP for a render of SPP samples per
// pixel (final-gathering channel grids); DIRECT_SAMPLES gets the map's
// direct-lighting samples per light.
int photon_render_view(const sgr_scene *sc, const sgr_photon_map *map, uint32_t spp,
                       PhotonView &out, uint32_t &direct_samples) {
  if (!map) return fail(SGR_ERR_INVALID, "sgr_render: the photon integrator needs params->photon_map");
  if (map->scene != sc)
    return fail(SGR_ERR_INVALID, "sgr_render: the photon map belongs to another scene");
  out = map->view;
  const uint32_t n[2] = {out.fg_photon, out.fg_bsdf};
  Grid *g[2] = {&out.g_fgp, &out.g_fgb};
  uint32_t *ns[2] = {&out.nsub_fgp, &out.nsub_fgb};
  for (int i = 0; i < 2; i++) {
    *g[i] = Grid{1, 1, 1};
    *ns[i] = 0;
    if (n[i]) {
      if ((uint64_t)spp * n[i] >= (1ull << 31))
        return fail(SGR_ERR_LIMIT, "sgr_render: spp x final-gathering samples above 2^31");
      *g[i] = grid_shape(spp * n[i]);
      *ns[i] = g[i]->total / spp;
    }
  }
  direct_samples = map->direct_samples;
  return SGR_OK;
}
}  // namespace sgr

extern "C" {

void sgr_photon_params_default(sgr_photon_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->caustic_photons = 50000;
  p->direct_photons = 500000;
  p->indirect_photons = 500000;
  p->use_photons = 50;
  p->photon_radius = 0.1f;
  p->final_gather = 1;
  p->fg_samples = p->fg_bsdf_samples = p->fg_photon_samples = SGR_PHOTON_UNSPEC;
  p->direct_illum = 1;
  p->direct_samples = 16;
  p->seed = 1;
}

int sgr_photon_map_create(sgr_scene *sc, const sgr_photon_params *p, sgr_photon_map **out) {
  sgr_photon_map *made = nullptr;
  const int rc = guarded(
      "sgr_photon_map_create", [&] { return create_impl(sc, p, out, made); },
      [&] {
        if (made) {
          map_free(made);
          delete made;
          made = nullptr;
        }
      });
  if (rc && made) {
    map_free(made);
    delete made;
  }
  return rc;
}

int sgr_photon_map_info_get(const sgr_photon_map *m, sgr_photon_map_info *info) {
  if (!m || !info) return fail(SGR_ERR_INVALID, "sgr_photon_map_info_get: null argument");
  *info = m->info;
  return SGR_OK;
}

int sgr_photon_map_export(const sgr_photon_map *m, int set, float *pos_axis, float *power,
                          float *dir) {
  if (!m) return fail(SGR_ERR_INVALID, "sgr_photon_map_export: null map");
  if (set < 0 || set > 2) return fail(SGR_ERR_INVALID, "sgr_photon_map_export: set must be 0, 1 or 2");
  const std::vector<Rec> &h = m->host[set];
  for (size_t k = 0; k < h.size(); k++) {
    if (pos_axis) std::memcpy(pos_axis + 4 * k, &h[k].pos, 16);
    if (power) {
      power[3 * k] = h[k].pow.x;
      power[3 * k + 1] = h[k].pow.y;
      power[3 * k + 2] = h[k].pow.z;
    }
    if (dir) {
      dir[3 * k] = h[k].dir.x;
      dir[3 * k + 1] = h[k].dir.y;
      dir[3 * k + 2] = h[k].dir.z;
    }
  }
  return SGR_OK;
}

int sgr_photon_map_destroy(sgr_photon_map *m) {
  if (!m) return SGR_OK;
  (void)hipSetDevice(m->scene->ctx->device);
  map_free(m);
  delete m;
  return SGR_OK;
}

}  // extern "C"
