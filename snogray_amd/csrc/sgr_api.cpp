// C ABI of the path-tracing core (include/snoggpu.h): contexts, scene
// upload with the host octree build, ray batches and the wavefront render
// driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snoggpu.h"
#include "kernels.h"
#include "octree_build.h"
#include "sgr_internal.h"

using namespace sgr;

namespace sgr {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

}  // namespace sgr

namespace {

int dev_alloc(sgr_ctx *ctx, sgr_scene *sc, size_t bytes, void **out) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess)
    return fail(SGR_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  if (sc) {
    sc->allocs.push_back(*out);
    sc->bytes += bytes;
  }
  ctx->bytes += bytes;
  return SGR_OK;
}

template<class T>
int upload(sgr_ctx *ctx, sgr_scene *sc, hipStream_t st, const std::vector<T> &v, const T **out) {
  void *p = nullptr;
  int rc = dev_alloc(ctx, sc, v.size() * sizeof(T), &p);
  if (rc) return rc;
  if (!v.empty())
    HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  *out = (const T *)p;
  return SGR_OK;
}

// Envmap light distribution: light map (texture/spheremap.cc:35-71),
// cos(latitude)-weighted histogram (light/envmap-light.cc:70-95) and its
// row / column cumulative sums, built in double and stored in float
// (geometry/hist-2d-dist.cc:52-110).
void build_env_cdf(const sgr_envmap &em, std::vector<float> &row_cum,
                   std::vector<float> &col_cum, uint32_t &hw, uint32_t &hh) {
  uint32_t w = em.width, h = em.height, div = 1;
  while (w / div > 1024 || h / div > 1024) div++;
  hw = w / div;
  hh = h / div;
  const float scale = 1.f / float(div * div);
  std::vector<float> lmap((size_t)hw * hh * 3, 0.f);
  for (uint32_t y = 0; y < h; y++) {
    for (uint32_t x = 0; x < w; x++) {
      uint32_t lx = x / div, ly = hh - y / div - 1;
      if (lx >= hw || ly >= hh) continue;
      const float *t = em.texels + 3 * ((size_t)y * w + x);
      float *d = &lmap[3 * ((size_t)ly * hw + lx)];
      for (int c = 0; c < 3; c++) d[c] = d[c] + t[c] * scale;
    }
  }
  std::vector<float> bins((size_t)hw * hh, 0.f);
  const double step = 3.14159265358979323846 / hh;
  double lat = -3.14159265358979323846 / 2 + step / 2;
  for (uint32_t r = 0; r < hh; r++, lat += step) {
    double c = std::cos(lat);
    for (uint32_t k = 0; k < hw; k++) {
      const float *px = &lmap[3 * ((size_t)r * hw + k)];
      float s = px[0];
      s += px[1];
      s += px[2];
      bins[(size_t)r * hw + k] += float(double(s / 3.f) * c);
    }
  }
  double total = 0;
  for (float b : bins) total += double(b);
  const double inv_total = total == 0 ? 0 : 1 / total;
  row_cum.assign(hh, 0.f);
  col_cum.assign((size_t)hw * hh, 0.f);
  double run = 0;
  for (uint32_t r = 0; r < hh; r++) {
    for (uint32_t k = 0; k < hw; k++) run += double(bins[(size_t)r * hw + k]) * inv_total;
    row_cum[r] = float(run);
  }
  for (uint32_t r = 0; r < hh; r++) {
    double rs = 0;
    for (uint32_t k = 0; k < hw; k++) rs += double(bins[(size_t)r * hw + k]);
    const double inv = rs == 0 ? 0 : 1 / rs;
    double acc = 0;
    for (uint32_t k = 0; k < hw; k++) {
      acc += double(bins[(size_t)r * hw + k]) * inv;
      col_cum[(size_t)r * hw + k] = float(acc);
    }
  }
}

// Guide table of a cumulative array A[N]: entry g (g = 0..G) is the
// std::lower_bound of float(g) / G in A, appended to OUT.
void build_guide(const float *a, uint32_t n, uint32_t G, std::vector<uint32_t> &out) {
  for (uint32_t g = 0; g <= G; g++) {
    const float val = (float)g / (float)G;
    out.push_back((uint32_t)(std::lower_bound(a, a + n, val) - a));
  }
}

inline float fmin_t(float a, float b) { return a < b ? a : b; }
inline float fmax_t(float a, float b) { return a > b ? a : b; }

}  // namespace

namespace sgr {
// Grid shape of a channel with n samples (render/grid.cc:26-38).
Grid grid_shape(uint32_t n) {
  uint32_t up = 1;
  while (up * up < n) up++;
  uint32_t down = up;
  while (down > 1 && (2 * down - 1) * (2 * down - 1) > 4 * n) down--;
  return Grid{up, down, up * down};
}
}  // namespace sgr

// ---------------------------------------------------------------------------
// Scratch of sgr_render_noalloc: the path state of its chunks lives in
// this array, so that entry point needs no device allocation.  Each
// device holds its own copy for the life of the process; host code uses
// the per-device address resolved at context creation, never the array's
// own name (a host-side stand-in).
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[SGR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};

// With DEVICE current: its copy's address, resolved once per device.
hipError_t scratch_address(int device, void **out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}
}  // namespace

extern "C" {

const char *sgr_last_error(void) { return sgr::g_err.c_str(); }

int sgr_ctx_create(int device, sgr_ctx **out) {
  if (!out) return fail(SGR_ERR_INVALID, "sgr_ctx_create: null out");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(SGR_ERR_INVALID, "sgr_ctx_create: no such device " + std::to_string(device));
  HIPCHK(hipSetDevice(device));
  sgr_ctx *c = new sgr_ctx;
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
  sgr::set_launch_cus((unsigned)c->num_cu);
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  for (auto &ls : c->lane_stream)
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ls, hipStreamNonBlocking);
  if (e != hipSuccess) {
    for (auto &ls : c->lane_stream)
      if (ls) (void)hipStreamDestroy(ls);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(SGR_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  for (auto &ev : c->ev) (void)hipEventCreate(&ev);
  if (hipMalloc((void **)&c->stats, SGR_STATS_BYTES) != hipSuccess) {
    (void)hipGetLastError();
    for (auto &ev : c->ev)
      if (ev) (void)hipEventDestroy(ev);
    for (auto &ls : c->lane_stream) (void)hipStreamDestroy(ls);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(SGR_ERR_NOMEM, "sgr_ctx_create: stats buffer");
  }
  // The no-alloc entry point reports the failure if this did not work.
  if (scratch_address(device, &c->noalloc_scratch) != hipSuccess) {
    c->noalloc_scratch = nullptr;
    (void)hipGetLastError();
  }
  *out = c;
  return SGR_OK;
}

int sgr_ctx_destroy(sgr_ctx *ctx) {
  if (!ctx) return SGR_OK;
  if (ctx->live_scenes)
    return sgr::fail(SGR_ERR_INVALID, "sgr_ctx_destroy: " + std::to_string(ctx->live_scenes.load()) +
                                     " scene(s) still alive; destroy them first");
  (void)hipSetDevice(ctx->device);
  // every stream idle before any buffer its kernels use is freed
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &ls : ctx->lane_stream) (void)hipStreamSynchronize(ls);
  if (ctx->work) (void)hipFree(ctx->work);
  if (ctx->stats) (void)hipFree(ctx->stats);
  for (auto &ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto &ev : ctx->evpool) (void)hipEventDestroy(ev);
  for (auto &ev : ctx->syncpool) (void)hipEventDestroy(ev);
  sgr::engine_destroy(ctx);
  for (auto &ls : ctx->lane_stream) (void)hipStreamDestroy(ls);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SGR_OK;
}

int sgr_ctx_memory(sgr_ctx *ctx, uint64_t *bytes) {
  if (!ctx || !bytes) return fail(SGR_ERR_INVALID, "sgr_ctx_memory: null argument");
  *bytes = ctx->bytes;
  return SGR_OK;
}

}  // extern "C"

namespace {

// Host-side scene preparation: validation, flattened primitives, the
// octree, light samplers and the envmap distribution.
struct HostScene {
  std::vector<DevMaterial> mats;
  std::vector<float4> tri, sph, env_tex;
  // local primitives, SGR_LOC_STRIDE float4 each: the world-to-local
  // rows, the local-to-world rows, then (record kind, 0, 0, 0)
  std::vector<float4> loc;
  std::vector<float4> tri_vn;  // vertex normals of mesh triangles, if any mesh has them
  std::vector<uint32_t> prim_mat, prim_mesh, order;
  std::vector<DevMesh> meshes;
  std::vector<DevLight> lights;
  std::vector<float4> plights;  // 4 per point light (SceneView::plights)
  std::vector<float> row_cum, col_cum;
  std::vector<uint32_t> row_guide, col_guide;
  FlatOctree oct;
  SceneView V{};
  float blo[3], bhi[3];
  // two-level instancing: the models (prepared as scenes of their own)
  // and the instances placing them
  std::vector<std::unique_ptr<HostScene>> models;
  std::vector<DevInstance> insts;
};

// Inverse of the affine 3 x 4 transform M (linear part by cofactors, in
// double, then rounded): false if it is singular or not finite.
bool invert_affine(const float *m, float *inv) {
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

// Pos::transformed (geometry/pos.h:80-97) in float.
void xform_point(const float *m, const float *p, float *o) {
  for (int r = 0; r < 3; r++)
    o[r] = p[0] * m[4 * r] + p[1] * m[4 * r + 1] + p[2] * m[4 * r + 2] + m[4 * r + 3];
}

int prepare_scene(const sgr_scene_desc *d, HostScene &H) {
  if (!d) return fail(SGR_ERR_INVALID, "sgr_scene_create: null descriptor");
  if (d->num_envmaps > 1)
    return fail(SGR_ERR_INVALID, "sgr_scene_create: at most one envmap light is supported");
  if (d->accel != SGR_ACCEL_OCTREE && d->accel != SGR_ACCEL_TRIV)
    return fail(SGR_ERR_INVALID, "sgr_scene_create: unknown accel");
  const uint32_t nm = d->num_materials;
  // ---- materials ----
  std::vector<DevMaterial> &mats = H.mats;
  mats.resize(nm);
  for (uint32_t i = 0; i < nm; i++) {
    const sgr_material &m = d->materials[i];
    if (m.type < SGR_MAT_LAMBERT || m.type > SGR_MAT_STENCIL)
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) + ": unknown type");
    if (m.type == SGR_MAT_GLOW && m.underlying >= (int)nm)
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) + ": bad underlying");
    if (m.type == SGR_MAT_STENCIL && (m.underlying < 0 || m.underlying >= (int)nm))
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) +
                                       ": stencil: the underlying material must exist and not "
                                       "be a stencil");
    DevMaterial &o = mats[i];
    std::memset(&o, 0, sizeof o);
    o.type = m.type;
    for (int c = 0; c < 3; c++) {
      o.color[c] = m.color[c];
      o.gloss[c] = m.gloss_color[c];
    }
    o.m = m.m;
    o.ior_n = m.ior_n;
    o.ior_k = m.ior_k;
    if (m.type == SGR_MAT_GLASS && !(m.ior_n > 0))
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) + ": glass ior must be > 0");
    // ThinGlass (col, ior) is PARTIALLY_OCCLUDING (material/thin-glass.h:30-32)
    if (m.type == SGR_MAT_THIN_GLASS && !(m.ior_n > 0))
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) + ": thin glass ior must be > 0");
    // so is Stencil (material/stencil.h:30-36)
    o.partial = (m.type == SGR_MAT_THIN_GLASS || m.type == SGR_MAT_STENCIL) ? 1 : 0;
    if (m.type == SGR_MAT_COOK_TORRANCE) {  // material/cook-torrance.cc:38-70
      float di = ((m.color[0] + m.color[1]) + m.color[2]) / 3.0f;
      float gi = ((m.gloss_color[0] + m.gloss_color[1]) + m.gloss_color[2]) / 3.0f;
      o.diff_weight = (di + gi) == 0 ? 0.0f : di / (di + gi);
      o.inv_diff_weight = o.diff_weight == 0 ? 0.0f : 1.0f / o.diff_weight;
      o.inv_gloss_weight = o.diff_weight == 1 ? 0.0f : 1.0f / (1.0f - o.diff_weight);
      o.gloss_layer = m.m < 0.5f ? 0x02u : 0x04u;
      o.have_layers = (o.diff_weight > 0 ? 0x04u : 0u) | (o.diff_weight < 1 ? o.gloss_layer : 0u);
    }
    if (m.type == SGR_MAT_PHONG) {  // material/phong.cc:38-45
      float s = m.color[0];  // Color::intensity (color/color.h:160-166)
      s += m.color[1];
      s += m.color[2];
      o.diff_weight = s / 3.0f;
      o.inv_diff_weight = o.diff_weight == 0 ? 0.0f : 1.0f / o.diff_weight;
      o.inv_gloss_weight = o.diff_weight == 1 ? 0.0f : 1.0f / (1.0f - o.diff_weight);
    }
    if (m.type == SGR_MAT_MIRROR) {  // material/mirror.cc:28-33
      float s = m.color[0];
      s += m.color[1];
      s += m.color[2];
      o.mirror_under = !(s < 3e-15f);
    }
  }
  for (uint32_t i = 0; i < nm; i++) {  // resolve GLOW -> BSDF material
    int b = (int)i, guard = 0;
    while (b >= 0 && mats[b].type == SGR_MAT_GLOW && guard++ < 16) b = d->materials[b].underlying;
    if (b >= 0 && mats[b].type == SGR_MAT_GLOW) b = -1;
    mats[i].bsdf = b;
    mats[i].medium = (b >= 0 && mats[b].type == SGR_MAT_GLASS) ? 1 : 0;  // a stencil has none
  }
  // Stencil (material/stencil.cc:30-52, 140-148): its underlying BSDF
  // material (through glows), the opacity's intensity snapped to 0 / 1
  // below 0.001 / above 0.999; fully opaque (component sum >= 3), its
  // get_bsdf is the underlying material's, which may have none.
  // Mirror over a material (material/mirror.h:37-42): its underlying
  // BSDF material (through glows; one that has none leaves the mirror
  // without), a Lambert, Cook-Torrance or Phong; underlying -1 keeps the
  // Lambert of `color` (material/mirror.cc:22-26).
  for (uint32_t i = 0; i < nm; i++) {
    DevMaterial &o = mats[i];
    const int u = d->materials[i].underlying;
    if (o.type != SGR_MAT_MIRROR || u < 0) continue;
    if (u >= (int)nm || u == (int)i)
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) + ": mirror: bad underlying");
    const int ub = mats[u].bsdf;
    if (ub >= 0 && mats[ub].type != SGR_MAT_LAMBERT && mats[ub].type != SGR_MAT_COOK_TORRANCE &&
        mats[ub].type != SGR_MAT_PHONG)
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) +
                                       ": mirror: the underlying material must be a lambert, "
                                       "cook-torrance or phong material");
    o.mirror_under = ub >= 0 ? 1 : 0;
    o.has_under = ub >= 0 ? 1 : 0;
    o.under_off = ub >= 0 ? ub - (int)i : 0;
  }
  for (uint32_t i = 0; i < nm; i++) {
    DevMaterial &o = mats[i];
    if (o.type != SGR_MAT_STENCIL) continue;
    const int ub = mats[d->materials[i].underlying].bsdf;
    if (ub >= 0 && mats[ub].type == SGR_MAT_STENCIL)
      return fail(SGR_ERR_INVALID, "material " + std::to_string(i) +
                                       ": stencil: the underlying material must exist and not "
                                       "be a stencil");
    o.has_under = ub >= 0 ? 1 : 0;
    o.under_off = ub >= 0 ? ub - (int)i : 0;
    float s = o.color[0];  // Color::intensity (color/color.h:160-166)
    s += o.color[1];
    s += o.color[2];
    o.op_int = s / 3.0f;
    o.inv_op_int = o.op_int == 0 ? 0.0f : 1.0f / o.op_int;
    if (o.op_int < 0.001f) o.op_int = o.inv_op_int = 0.0f;
    else if (o.op_int > 0.999f) o.op_int = o.inv_op_int = 1.0f;
  }
  for (uint32_t i = 0; i < nm; i++) {
    const int b = mats[i].bsdf;
    if (b >= 0 && mats[b].type == SGR_MAT_STENCIL) {
      float s = mats[b].color[0];
      s += mats[b].color[1];
      s += mats[b].color[2];
      if (s >= 3.0f) mats[i].bsdf = mats[d->materials[b].underlying].bsdf;
    }
  }
  // ---- primitives ----
  uint64_t n_tri = 0;
  for (uint32_t i = 0; i < d->num_meshes; i++) n_tri += d->meshes[i].num_triangles;
  if ((d->num_ellipses && !d->ellipses) || (d->num_local_prims && !d->local_prims) ||
      (d->num_tripars && !d->tripars) || (d->num_spheres && !d->spheres) ||
      (d->num_meshes && !d->meshes) || (nm && !d->materials))
    return fail(SGR_ERR_INVALID, "sgr_scene_create: null array with a non-zero count");
  uint64_t n_tl = n_tri + d->num_tripars + d->num_ellipses;
  uint64_t n_sph_end = n_tl + d->num_spheres;
  uint64_t n_prims = n_sph_end + d->num_local_prims;
  if (n_prims >= (1ull << 28))
    return fail(SGR_ERR_LIMIT, "sgr_scene_create: more than 2^28 primitives");
  std::vector<float4> &tri = H.tri, &sph = H.sph;
  tri.resize((size_t)n_tl * 3);
  sph.resize(d->num_spheres);
  H.loc.assign((size_t)SGR_LOC_STRIDE * d->num_local_prims, make_float4(0.f, 0.f, 0.f, 0.f));
  std::vector<uint32_t> &prim_mat = H.prim_mat, &prim_mesh = H.prim_mesh, &order = H.order;
  prim_mat.assign(n_prims, 0);
  prim_mesh.assign(n_prims, 0xffffffffu);
  std::vector<DevMesh> &meshes = H.meshes;
  meshes.resize(d->num_meshes);
  // ---- models (their own geometry and octree, the scene's materials) ----
  if (d->num_instances && !d->instances) return fail(SGR_ERR_INVALID, "sgr_scene_create: null instances");
  if (d->num_models && !d->models) return fail(SGR_ERR_INVALID, "sgr_scene_create: null models");
  for (uint32_t mi = 0; mi < d->num_models; mi++) {
    const sgr_model &m = d->models[mi];
    sgr_scene_desc md;
    std::memset(&md, 0, sizeof md);
    md.meshes = m.meshes; md.num_meshes = m.num_meshes;
    md.tripars = m.tripars; md.num_tripars = m.num_tripars;
    md.spheres = m.spheres; md.num_spheres = m.num_spheres;
    md.ellipses = m.ellipses; md.num_ellipses = m.num_ellipses;
    md.local_prims = m.local_prims; md.num_local_prims = m.num_local_prims;
    md.materials = d->materials; md.num_materials = nm;
    md.min_trace = d->min_trace;
    md.accel = d->accel;
    H.models.emplace_back(new HostScene);
    int rc = prepare_scene(&md, *H.models.back());
    if (rc) return fail(rc, "model " + std::to_string(mi) + ": " + g_err);
    if (H.models.back()->order.empty())
      return fail(SGR_ERR_INVALID, "model " + std::to_string(mi) + " has no surfaces");
    if (H.models.back()->V.has_partial)
      return fail(SGR_ERR_INVALID, "model " + std::to_string(mi) +
                                       ": partially occluding materials (thin glass, stencil) are not "
                                       "supported in models");
  }
  const uint32_t n_inst = d->num_instances;
  if ((uint64_t)n_prims + n_inst >= (1ull << 28))
    return fail(SGR_ERR_LIMIT, "sgr_scene_create: more than 2^28 primitives and instances");
  std::vector<Box3> boxes(n_prims + n_inst);
  order.reserve(n_prims);
  float *blo = H.blo, *bhi = H.bhi;
  for (int a = 0; a < 3; a++) {
    blo[a] = 3.4028235e38f;
    bhi[a] = -3.4028235e38f;
  }
  auto grow = [&](const float *lo, const float *hi) {
    for (int a = 0; a < 3; a++) {
      blo[a] = fmin_t(blo[a], lo[a]);
      bhi[a] = fmax_t(bhi[a], hi[a]);
    }
  };
  uint32_t p = 0;
  for (uint32_t mi = 0; mi < d->num_meshes; mi++) {
    const sgr_mesh &m = d->meshes[mi];
    if (m.num_triangles && (!m.vertices || !m.triangles))
      return fail(SGR_ERR_INVALID, "mesh " + std::to_string(mi) + ": null arrays");
    float mlo[3] = {3.4028235e38f, 3.4028235e38f, 3.4028235e38f};
    float mhi[3] = {-3.4028235e38f, -3.4028235e38f, -3.4028235e38f};
    for (uint32_t k = 0; k < m.num_vertices; k++)
      for (int a = 0; a < 3; a++) {
        if (!std::isfinite(m.vertices[3 * k + a]))  // it would poison the scene box and horizon
          return fail(SGR_ERR_INVALID, "mesh " + std::to_string(mi) + ": non-finite vertex");
        mlo[a] = fmin_t(mlo[a], m.vertices[3 * k + a]);
        mhi[a] = fmax_t(mhi[a], m.vertices[3 * k + a]);
      }
    if (m.num_vertices) grow(mlo, mhi);
    DevMesh &dm = meshes[mi];
    for (int a = 0; a < 3; a++) {
      dm.axis[a] = m.axis[a];
      dm.center[a] = (mlo[a] + mhi[a]) / 2;  // geometry/pos.h:123-126
    }
    dm.left_handed = m.left_handed ? 1 : 0;
    dm.has_normals = m.normals ? 1 : 0;
    if (m.normals) {
      for (uint32_t k = 0; k < 3 * m.num_vertices; k++)
        if (!std::isfinite(m.normals[k]))
          return fail(SGR_ERR_INVALID, "mesh " + std::to_string(mi) + ": non-finite normal");
      if (H.tri_vn.empty()) H.tri_vn.assign((size_t)n_tri * 3, make_float4(0.f, 0.f, 0.f, 0.f));
    }
    for (uint32_t t = 0; t < m.num_triangles; t++, p++) {
      uint32_t ix[3] = {m.triangles[3 * t], m.triangles[3 * t + 1], m.triangles[3 * t + 2]};
      for (int c = 0; c < 3; c++)
        if (ix[c] >= m.num_vertices)
          return fail(SGR_ERR_INVALID, "mesh " + std::to_string(mi) + ": vertex index out of range");
      const float *a = m.vertices + 3 * ix[0], *b = m.vertices + 3 * ix[1], *c = m.vertices + 3 * ix[2];
      // surface/mesh.cc:735-736: corner v0, edges v1-v0, v2-v0.
      float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
      float e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
      tri[3 * p] = make_float4(a[0], a[1], a[2], 0.0f);
      tri[3 * p + 1] = make_float4(e1[0], e1[1], e1[2], 0.0f);
      tri[3 * p + 2] = make_float4(e2[0], e2[1], e2[2], 0.0f);
      if (m.normals)
        for (int c = 0; c < 3; c++) {
          const float *nv = m.normals + 3 * (size_t)ix[c];
          H.tri_vn[3 * p + c] = make_float4(nv[0], nv[1], nv[2], 0.0f);
        }
      uint32_t mat = m.tri_materials ? m.tri_materials[t] : m.material;
      if (mat >= nm) return fail(SGR_ERR_INVALID, "mesh " + std::to_string(mi) + ": bad material");
      prim_mat[p] = mat;
      prim_mesh[p] = mi;
      Box3 &bx = boxes[p];
      for (int k = 0; k < 3; k++) {
        bx.lo[k] = fmin_t(fmin_t(a[k], b[k]), c[k]);
        bx.hi[k] = fmax_t(fmax_t(a[k], b[k]), c[k]);
      }
      // surface/mesh.cc:832-846: degenerate triangles are not inserted.
      float n0 = e2[1] * e1[2] - e2[2] * e1[1], n1 = e2[2] * e1[0] - e2[0] * e1[2],
            n2 = e2[0] * e1[1] - e2[1] * e1[0];
      if (n0 * n0 + n1 * n1 + n2 * n2 > 0) order.push_back(p);
    }
  }
  for (uint32_t i = 0; i < d->num_tripars; i++, p++) {
    const sgr_tripar &t = d->tripars[i];
    if (t.material >= nm) return fail(SGR_ERR_INVALID, "tripar " + std::to_string(i) + ": bad material");
    tri[3 * p] = make_float4(t.v0[0], t.v0[1], t.v0[2], t.parallelogram ? 1.0f : 0.0f);
    tri[3 * p + 1] = make_float4(t.e1[0], t.e1[1], t.e1[2], 0.0f);
    tri[3 * p + 2] = make_float4(t.e2[0], t.e2[1], t.e2[2], 0.0f);
    prim_mat[p] = t.material;
    // surface/tripar.cc:140-148
    float c1[3], c2[3], c3[3];
    for (int k = 0; k < 3; k++) {
      c1[k] = t.v0[k] + t.e1[k];
      c2[k] = t.v0[k] + t.e2[k];
      c3[k] = (t.v0[k] + t.e1[k]) + t.e2[k];
    }
    Box3 &bx = boxes[p];
    for (int k = 0; k < 3; k++) {
      bx.lo[k] = fmin_t(fmin_t(t.v0[k], c1[k]), c2[k]);
      bx.hi[k] = fmax_t(fmax_t(t.v0[k], c1[k]), c2[k]);
      if (t.parallelogram) {
        bx.lo[k] = fmin_t(bx.lo[k], c3[k]);
        bx.hi[k] = fmax_t(bx.hi[k], c3[k]);
      }
    }
    grow(bx.lo, bx.hi);
    order.push_back(p);
  }
  // Ellipses (surface/ellipse.h:38-45): the bounding parallelogram,
  // corner center - radius1 - radius2 and edges 2 radius1, 2 radius2,
  // flagged 2 in the corner's w; box of its four corners
  // (surface/ellipse.cc:171-181).
  for (uint32_t i = 0; i < d->num_ellipses; i++, p++) {
    const sgr_ellipse &el = d->ellipses[i];
    if (el.material >= nm) return fail(SGR_ERR_INVALID, "ellipse " + std::to_string(i) + ": bad material");
    float v0[3], e1[3], e2[3];
    for (int k = 0; k < 3; k++) {
      v0[k] = (el.center[k] - el.radius1[k]) - el.radius2[k];
      e1[k] = el.radius1[k] * 2.0f;
      e2[k] = el.radius2[k] * 2.0f;
    }
    tri[3 * p] = make_float4(v0[0], v0[1], v0[2], 2.0f);
    tri[3 * p + 1] = make_float4(e1[0], e1[1], e1[2], 0.0f);
    tri[3 * p + 2] = make_float4(e2[0], e2[1], e2[2], 0.0f);
    prim_mat[p] = el.material;
    Box3 &bx = boxes[p];
    for (int k = 0; k < 3; k++) {
      const float c1 = v0[k] + e1[k], c2 = v0[k] + e2[k], c3 = (v0[k] + e1[k]) + e2[k];
      bx.lo[k] = fmin_t(fmin_t(fmin_t(v0[k], c1), c2), c3);
      bx.hi[k] = fmax_t(fmax_t(fmax_t(v0[k], c1), c2), c3);
    }
    grow(bx.lo, bx.hi);
    order.push_back(p);
  }
  for (uint32_t i = 0; i < d->num_spheres; i++, p++) {
    const sgr_sphere &s = d->spheres[i];
    if (s.material >= nm) return fail(SGR_ERR_INVALID, "sphere " + std::to_string(i) + ": bad material");
    sph[i] = make_float4(s.center[0], s.center[1], s.center[2], s.radius);
    prim_mat[p] = s.material;
    Box3 &bx = boxes[p];
    for (int k = 0; k < 3; k++) {
      bx.lo[k] = s.center[k] - s.radius;
      bx.hi[k] = s.center[k] + s.radius;
    }
    grow(bx.lo, bx.hi);
    order.push_back(p);
  }
  // ---- local primitives (surface/local-primitive.h): cylinder and
  // sphere2, a canonical shape under a transform and its inverse
  // (geometry/local-xform.cc:17-21); box: the local cube [-1,1]^3, its 8
  // corners taken to world space (geometry/local-xform.cc:27-31) ----
  for (uint32_t i = 0; i < d->num_local_prims; i++, p++) {
    const sgr_local_prim &lp = d->local_prims[i];
    if (lp.material >= nm)
      return fail(SGR_ERR_INVALID, "local primitive " + std::to_string(i) + ": bad material");
    if (lp.kind != SGR_LOCAL_CYLINDER && lp.kind != SGR_LOCAL_SPHERE2)
      return fail(SGR_ERR_INVALID, "local primitive " + std::to_string(i) + ": unknown kind");
    float w2l[12];
    if (!invert_affine(lp.local_to_world, w2l))
      return fail(SGR_ERR_INVALID, "local primitive " + std::to_string(i) + ": singular transform");
    float4 *lr = &H.loc[(size_t)SGR_LOC_STRIDE * i];
    for (int r = 0; r < 3; r++) {
      lr[r] = make_float4(w2l[4 * r], w2l[4 * r + 1], w2l[4 * r + 2], w2l[4 * r + 3]);
      const float *m = lp.local_to_world + 4 * r;
      lr[3 + r] = make_float4(m[0], m[1], m[2], m[3]);
    }
    uint32_t rk = lp.kind == SGR_LOCAL_CYLINDER ? SGR_REF_CYLINDER : SGR_REF_SPHERE2;
    std::memcpy(&lr[6].x, &rk, 4);
    prim_mat[p] = lp.material;
    Box3 &bx = boxes[p];
    for (int c = 0; c < 8; c++) {
      const float corner[3] = {(c & 4) ? 1.0f : -1.0f, (c & 2) ? 1.0f : -1.0f,
                               (c & 1) ? 1.0f : -1.0f};
      float w[3];
      xform_point(lp.local_to_world, corner, w);
      for (int k = 0; k < 3; k++) {
        bx.lo[k] = c ? fmin_t(bx.lo[k], w[k]) : w[k];
        bx.hi[k] = c ? fmax_t(bx.hi[k], w[k]) : w[k];
      }
    }
    grow(bx.lo, bx.hi);
    order.push_back(p);
  }
  // ---- instances (surface/instance.cc:145-216): the model's box, its 8
  // corners taken to world space (geometry/bbox.cc:25-42) ----
  for (uint32_t i = 0; i < n_inst; i++) {
    const sgr_instance &in = d->instances[i];
    if (in.model >= d->num_models)
      return fail(SGR_ERR_INVALID, "instance " + std::to_string(i) + ": bad model");
    DevInstance I;
    I.model = in.model;
    std::memcpy(I.l2w, in.local_to_world, sizeof I.l2w);
    if (!invert_affine(I.l2w, I.w2l))
      return fail(SGR_ERR_INVALID, "instance " + std::to_string(i) + ": singular transform");
    H.insts.push_back(I);
    const HostScene &M = *H.models[in.model];
    Box3 &bx = boxes[n_prims + i];
    for (int c = 0; c < 8; c++) {
      const float corner[3] = {(c & 4) ? M.bhi[0] : M.blo[0], (c & 2) ? M.bhi[1] : M.blo[1],
                               (c & 1) ? M.bhi[2] : M.blo[2]};
      float w[3];
      xform_point(I.l2w, corner, w);
      for (int k = 0; k < 3; k++) {
        bx.lo[k] = c ? fmin_t(bx.lo[k], w[k]) : w[k];
        bx.hi[k] = c ? fmax_t(bx.hi[k], w[k]) : w[k];
      }
    }
    grow(bx.lo, bx.hi);
    order.push_back((uint32_t)n_prims + i);
  }
  // A surface whose box is not finite would grow the octree root without
  // end (space/octree-builder.cc:180-196 doubles it until the box fits).
  for (uint32_t q : order)
    for (int k = 0; k < 3; k++)
      if (!std::isfinite(boxes[q].lo[k]) || !std::isfinite(boxes[q].hi[k]))
        return fail(SGR_ERR_INVALID, "primitive " + std::to_string(q) + ": non-finite coordinates");
  SceneView &V = H.V;
  std::memset(&V, 0, sizeof V);
  // ---- acceleration structure ----
  if (d->accel == SGR_ACCEL_OCTREE) {
    try {
      build_octree(boxes, order, H.oct);
    } catch (const std::bad_alloc &) {
      return fail(SGR_ERR_NOMEM, "sgr_scene_create: out of host memory building the octree");
    } catch (const std::exception &e) {
      return fail(SGR_ERR_LIMIT, std::string("sgr_scene_create: ") + e.what());
    }
    if (H.oct.max_depth > SGR_MAX_DEPTH)
      return fail(SGR_ERR_LIMIT, "sgr_scene_create: octree depth " +
                                     std::to_string(H.oct.max_depth) + " exceeds " +
                                     std::to_string(SGR_MAX_DEPTH));
    if (H.oct.nodes.size() / 4 >= (1ull << 28))
      return fail(SGR_ERR_LIMIT, "sgr_scene_create: more than 2^28 octree nodes");
    V.n_nodes = (uint32_t)(H.oct.nodes.size() / 4);
    for (int a = 0; a < 3; a++) V.origin[a] = H.oct.origin[a];
    V.size = H.oct.size;
    V.n_surfs = (uint32_t)H.oct.surf_ids.size();
  } else {
    H.oct.num_real = order.size();
    V.use_triv = 1;
    V.n_surfs = (uint32_t)order.size();
  }
  V.n_tri = (uint32_t)n_tri;
  V.n_tri_like = (uint32_t)n_tl;
  V.n_sph_end = (uint32_t)n_sph_end;
  V.n_prims = (uint32_t)n_prims;
  V.min_trace = d->min_trace;
  {  // render/scene.cc:28 horizon = bbox diameter
    float ex = bhi[0] - blo[0], ey = bhi[1] - blo[1], ez = bhi[2] - blo[2];
    V.horizon = (p || n_inst) ? std::sqrt(ex * ex + ey * ey + ez * ez) : 0.0f;
  }
  for (int a = 0; a < 3; a++) {
    V.bbox_lo[a] = blo[a];
    V.bbox_hi[a] = bhi[a];
  }
  // ---- lights (render/scene.cc:34, surface/primitive.cc:48-54) ----
  for (uint32_t q = (uint32_t)n_tri; q < n_prims; q++) {
    uint32_t m = prim_mat[q];
    if (mats[m].type == SGR_MAT_GLOW) {
      DevLight L;
      // tripar 0, sphere 1; surface lights of an ellipse 3, a cylinder 4,
      // a sphere2 5
      if (q < n_tl) L.kind = tri[3 * q].w == 2.0f ? 3 : 0;
      else if (q < n_sph_end) L.kind = 1;
      else L.kind = d->local_prims[q - n_sph_end].kind == SGR_LOCAL_CYLINDER ? 4 : 5;
      L.prim = q;
      for (int c = 0; c < 3; c++) L.intensity[c] = mats[m].color[c];
      H.lights.push_back(L);
    }
  }
  if (d->num_envmaps == 1) {
    const sgr_envmap &em = d->envmaps[0];
    if (!em.texels || em.width == 0 || em.height == 0)
      return fail(SGR_ERR_INVALID, "envmap: empty");
    H.env_tex.resize((size_t)em.width * em.height);
    for (size_t i = 0; i < H.env_tex.size(); i++)
      H.env_tex[i] = make_float4(em.texels[3 * i], em.texels[3 * i + 1], em.texels[3 * i + 2], 0.f);
    build_env_cdf(em, H.row_cum, H.col_cum, V.env_hw, V.env_hh);
    build_guide(H.row_cum.data(), V.env_hh, SGR_ENV_GUIDE_ROW, H.row_guide);
    for (uint32_t r = 0; r < V.env_hh; r++)
      build_guide(H.col_cum.data() + (size_t)r * V.env_hw, V.env_hw, SGR_ENV_GUIDE_COL,
                  H.col_guide);
    V.env_w = em.width;
    V.env_h = em.height;
    V.has_env = 1;
    for (int k = 0; k < 9; k++) V.env_frame[k] = em.frame[k];
    DevLight L;
    L.kind = 2;
    L.prim = 0;
    L.intensity[0] = L.intensity[1] = L.intensity[2] = 0;
    H.lights.push_back(L);
  }
  // point lights (light/point-light.h:31-47): all directions, or a cone
  // about dir in the frame Frame (pos, dir) (geometry/frame.h:79-84)
  if (d->num_point_lights && !d->point_lights)
    return fail(SGR_ERR_INVALID, "sgr_scene_create: null point_lights");
  for (uint32_t i = 0; i < d->num_point_lights; i++) {
    const sgr_point_light &pl = d->point_lights[i];
    const std::string what = "point light " + std::to_string(i);
    bool finite = std::isfinite(pl.angle) && std::isfinite(pl.fringe_angle);
    for (int k = 0; k < 3; k++)
      finite = finite && std::isfinite(pl.pos[k]) && std::isfinite(pl.intensity[k]) &&
               std::isfinite(pl.dir[k]);
    if (!finite) return fail(SGR_ERR_INVALID, what + ": not finite");
    float ch = -1.0f, cc = -1.0f;
    float fx[3] = {1, 0, 0}, fy[3] = {0, 1, 0}, fz[3] = {0, 0, 1};
    if (pl.angle > 0.0f) {
      const float *z = pl.dir;
      if (!(z[0] * z[0] + z[1] * z[1] + z[2] * z[2] > 0.0f))
        return fail(SGR_ERR_INVALID, what + ": a spotlight needs a direction");
      ch = std::cos(pl.angle / 2);
      cc = std::cos((pl.angle - pl.fringe_angle) / 2);
      // Vec::perpendicular (geometry/vec.h:122-129), then unit vectors
      const float nz = z[0] == z[2] ? z[2] : -z[2];
      float px[3] = {z[1] * z[1] + z[0] * nz, -nz * nz - z[0] * z[1], z[0] * z[0] - z[1] * nz};
      auto unit = [](float *v) {
        const float inv = 1 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) v[k] = v[k] * inv;
      };
      unit(px);
      float py[3] = {z[1] * px[2] - z[2] * px[1], z[2] * px[0] - z[0] * px[2],
                     z[0] * px[1] - z[1] * px[0]};
      unit(py);
      for (int k = 0; k < 3; k++) {
        fx[k] = px[k];
        fy[k] = py[k];
        fz[k] = z[k];
      }
    }
    H.plights.push_back(make_float4(pl.pos[0], pl.pos[1], pl.pos[2], ch));
    H.plights.push_back(make_float4(fx[0], fx[1], fx[2], cc));
    H.plights.push_back(make_float4(fy[0], fy[1], fy[2], 0.0f));
    H.plights.push_back(make_float4(fz[0], fz[1], fz[2], 0.0f));
    DevLight L;
    L.kind = 6;
    L.prim = i;
    for (int c = 0; c < 3; c++) L.intensity[c] = pl.intensity[c];
    H.lights.push_back(L);
  }
  V.n_lights = (uint32_t)H.lights.size();
  V.n_insts = n_inst;
  // the general kernels also carry the point lights' sampling
  V.has_ext = (d->num_ellipses || d->num_local_prims || d->num_point_lights) ? 1u : 0u;
  for (const auto &m : H.models) V.has_ext |= m->V.has_ext;
  V.has_partial = 0;
  for (uint32_t p = 0; p < n_prims; p++)
    if (mats[prim_mat[p]].partial) V.has_partial = 1;
  return SGR_OK;
}

}  // namespace

extern "C" {

static int octree_build_impl(const sgr_scene_desc *d, uint32_t *nodes, uint64_t *num_nodes,
                             uint32_t *surf_ids, uint64_t *num_surf_ids, float *origin_size4,
                             sgr_octree_stats *stats) {
  HostScene H;
  int rc = prepare_scene(d, H);
  if (rc) return rc;
  if (num_nodes) *num_nodes = H.oct.nodes.size() / 4;
  if (num_surf_ids) *num_surf_ids = H.oct.surf_ids.size();
  if (nodes) std::memcpy(nodes, H.oct.nodes.data(), H.oct.nodes.size() * 4);
  if (surf_ids) std::memcpy(surf_ids, H.oct.surf_ids.data(), H.oct.surf_ids.size() * 4);
  if (origin_size4) {
    for (int a = 0; a < 3; a++) origin_size4[a] = H.oct.origin[a];
    origin_size4[3] = H.oct.size;
  }
  if (stats) {
    stats->num_nodes = H.oct.nodes.size() / 4;
    stats->num_leaf_nodes = H.oct.num_leaf;
    stats->num_surfaces = H.oct.num_surfaces;
    stats->num_dup_surfaces = H.oct.num_surfaces - H.oct.num_real;
    stats->max_depth = H.oct.max_depth;
    stats->avg_depth = H.oct.avg_depth;
  }
  return SGR_OK;
}

int sgr_octree_build(const sgr_scene_desc *d, uint32_t *nodes, uint64_t *num_nodes,
                     uint32_t *surf_ids, uint64_t *num_surf_ids, float *origin_size4,
                     sgr_octree_stats *stats) {
  return sgr::guarded("sgr_octree_build", [&] {
    return octree_build_impl(d, nodes, num_nodes, surf_ids, num_surf_ids, origin_size4, stats);
  });
}

// Device copies of a prepared scene's (or model's) geometry: octree
// nodes, surface references with their geometry inline in list order (a
// node's surfaces contiguous; kind in the second record's w: 0 triangle,
// 1 parallelogram, 2 sphere, 3 instance, 4 ellipse, 5 cylinder, 6
// sphere2), primitives and meshes.
static int upload_geometry(sgr_ctx *ctx, sgr_scene *sc, hipStream_t up, const HostScene &H,
                           const FlatOctree &oct, SceneView &V) {
  std::vector<uint4> nodes4(V.n_nodes);
  for (uint32_t i = 0; i < V.n_nodes; i++)
    nodes4[i] = make_uint4(oct.nodes[4 * i], oct.nodes[4 * i + 1], oct.nodes[4 * i + 2],
                           oct.nodes[4 * i + 3]);
  const std::vector<uint32_t> &surfs = V.use_triv ? H.order : oct.surf_ids;
  std::vector<float4> refs(surfs.size() * SGR_REF_STRIDE, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  for (size_t k = 0; k < surfs.size(); k++) {
    uint32_t q = surfs[k];
    float pb, kind;
    std::memcpy(&pb, &q, 4);
    uint32_t kd = 0;
    if (q < V.n_tri_like) {
      const float4 &a = H.tri[3 * q], &b = H.tri[3 * q + 1], &c = H.tri[3 * q + 2];
      kd = a.w == 2.0f ? SGR_REF_ELLIPSE : a.w != 0.0f ? 1u : 0u;
      std::memcpy(&kind, &kd, 4);
      refs[SGR_REF_STRIDE * k] = make_float4(a.x, a.y, a.z, pb);
      refs[SGR_REF_STRIDE * k + 1] = make_float4(b.x, b.y, b.z, kind);
      refs[SGR_REF_STRIDE * k + 2] = make_float4(c.x, c.y, c.z, 0.0f);
    } else if (q >= V.n_sph_end && q < V.n_prims) {  // local primitive: its table index
      const uint32_t li = q - V.n_sph_end;
      std::memcpy(&kd, &H.loc[(size_t)SGR_LOC_STRIDE * li + 6].x, 4);
      std::memcpy(&kind, &kd, 4);
      float lb;
      std::memcpy(&lb, &li, 4);
      refs[SGR_REF_STRIDE * k] = make_float4(0.0f, 0.0f, 0.0f, pb);
      refs[SGR_REF_STRIDE * k + 1] = make_float4(0.0f, 0.0f, 0.0f, kind);
      refs[SGR_REF_STRIDE * k + 2] = make_float4(lb, 0.0f, 0.0f, 0.0f);
    } else if (q < V.n_prims) {
      const float4 &sp = H.sph[q - V.n_tri_like];
      kd = 2u;
      std::memcpy(&kind, &kd, 4);
      refs[SGR_REF_STRIDE * k] = make_float4(sp.x, sp.y, sp.z, pb);
      refs[SGR_REF_STRIDE * k + 1] = make_float4(sp.w, 0.0f, 0.0f, kind);
    } else {  // instance n_prims + i
      kd = 3u;
      std::memcpy(&kind, &kd, 4);
      refs[SGR_REF_STRIDE * k] = make_float4(0.0f, 0.0f, 0.0f, pb);
      refs[SGR_REF_STRIDE * k + 1] = make_float4(0.0f, 0.0f, 0.0f, kind);
    }
  }
  int rc = upload(ctx, sc, up, nodes4, &V.nodes);
  if (!rc) rc = upload(ctx, sc, up, refs, &V.refs);
  if (!rc) rc = upload(ctx, sc, up, H.tri, &V.tri);
  if (!rc) rc = upload(ctx, sc, up, H.sph, &V.sph);
  if (!rc && !H.loc.empty()) rc = upload(ctx, sc, up, H.loc, &V.loc);
  if (!rc && !H.tri_vn.empty()) rc = upload(ctx, sc, up, H.tri_vn, &V.tri_vn);
  if (!rc) rc = upload(ctx, sc, up, H.prim_mat, &V.prim_mat);
  if (!rc) rc = upload(ctx, sc, up, H.prim_mesh, &V.prim_mesh);
  if (!rc) rc = upload(ctx, sc, up, H.meshes, &V.meshes);
  return rc;
}

static int scene_create_impl(sgr_ctx *ctx, const sgr_scene_desc *d, sgr_scene **out,
                             sgr_scene *&made) {
  if (!ctx || !d || !out) return fail(SGR_ERR_INVALID, "sgr_scene_create: null argument");
  HostScene H;
  int rc = prepare_scene(d, H);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  sgr_scene *sc = new sgr_scene;
  sc->ctx = ctx;
  ctx->live_scenes++;
  made = sc;
  for (int a = 0; a < 3; a++) {
    sc->bbox_lo[a] = H.blo[a];
    sc->bbox_hi[a] = H.bhi[a];
  }
  sc->oct = std::move(H.oct);
  sc->n_lights = H.V.n_lights;
  SceneView &V = sc->view;
  V = H.V;
  // Lane refill pays off once walks are long enough to leave waves
  // unbalanced: on the 1.31 M-triangle soup (depth 19) it is 16 % faster,
  // on the Cornell box (depth 5) 30 % slower than one ray per lane.
  // Camera rays take the plain walk on every scene (launch_extend).
  V.walk_refill = sc->oct.max_depth >= 8 && V.n_insts == 0 ? 1u : 0u;
  // Uploads go through a stream of their own: a scene may be created by
  // one host thread while another drives renders on the context's stream.
  hipStream_t up = nullptr;
  {
    hipError_t e = hipStreamCreateWithFlags(&up, hipStreamNonBlocking);
    if (e != hipSuccess) rc = fail(SGR_ERR_HIP, std::string("upload stream: ") + hipGetErrorString(e));
  }
  if (!rc) rc = upload_geometry(ctx, sc, up, H, sc->oct, V);
  if (!rc) rc = upload(ctx, sc, up, H.mats, &V.mats);
  if (!rc && !H.models.empty()) {
    // the models' own views: geometry and octree in model space, the
    // scene's materials, no lights
    std::vector<SceneView> mv(H.models.size());
    for (size_t m = 0; m < H.models.size() && !rc; m++) {
      HostScene &Hm = *H.models[m];
      mv[m] = Hm.V;
      rc = upload_geometry(ctx, sc, up, Hm, Hm.oct, mv[m]);
      mv[m].mats = V.mats;
      mv[m].walk_refill = 0;
      mv[m].n_lights = 0;
      mv[m].has_env = 0;
    }
    if (!rc) rc = upload(ctx, sc, up, mv, &V.models);
    if (!rc) rc = upload(ctx, sc, up, H.insts, &V.insts);
  }
  if (!rc) rc = upload(ctx, sc, up, H.lights, &V.lights);
  if (!rc && !H.plights.empty()) rc = upload(ctx, sc, up, H.plights, &V.plights);
  if (!rc) rc = upload(ctx, sc, up, H.env_tex, &V.env_tex);
  if (!rc) rc = upload(ctx, sc, up, H.row_cum, &V.env_row_cum);
  if (!rc) rc = upload(ctx, sc, up, H.col_cum, &V.env_col_cum);
  if (!rc) rc = upload(ctx, sc, up, H.row_guide, &V.env_row_guide);
  if (!rc) rc = upload(ctx, sc, up, H.col_guide, &V.env_col_guide);
  if (up) {
    hipError_t e = hipStreamSynchronize(up);
    if (!rc && e != hipSuccess) rc = fail(SGR_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    (void)hipStreamDestroy(up);
  }
  made = nullptr;
  if (rc) {
    sgr_scene_destroy(sc);
    return rc;
  }
  *out = sc;
  return SGR_OK;
}

// No C++ exception leaves the C ABI: host allocation failures become
// SGR_ERR_NOMEM, and a half-made scene is destroyed.
int sgr_scene_create(sgr_ctx *ctx, const sgr_scene_desc *d, sgr_scene **out) {
  sgr_scene *made = nullptr;
  return sgr::guarded("sgr_scene_create", [&] { return scene_create_impl(ctx, d, out, made); },
                      [&] {
                        if (made) sgr_scene_destroy(made);
                      });
}

int sgr_scene_destroy(sgr_scene *sc) {
  if (!sc) return SGR_OK;
  if (sc->jobs_using)
    return fail(SGR_ERR_INVALID, "sgr_scene_destroy: " + std::to_string(sc->jobs_using) +
                                     " job(s) still refer to the scene; release them first");
  if (sc->photon_maps)
    return fail(SGR_ERR_INVALID, "sgr_scene_destroy: " + std::to_string(sc->photon_maps) +
                                     " photon map(s) belong to the scene; destroy them first");
  if (sc->ctx) {
    (void)hipSetDevice(sc->ctx->device);
    (void)hipDeviceSynchronize();  // no render of any thread still reads it
    sc->ctx->bytes -= sc->bytes;
    sc->ctx->live_scenes--;
  }
  for (void *p : sc->allocs) (void)hipFree(p);
  delete sc;
  return SGR_OK;
}

int sgr_scene_octree_stats(const sgr_scene *sc, sgr_octree_stats *out) {
  if (!sc || !out) return fail(SGR_ERR_INVALID, "sgr_scene_octree_stats: null argument");
  out->num_nodes = sc->oct.nodes.size() / 4;
  out->num_leaf_nodes = sc->oct.num_leaf;
  out->num_surfaces = sc->oct.num_surfaces;
  out->num_dup_surfaces = sc->oct.num_surfaces - sc->oct.num_real;
  out->max_depth = sc->oct.max_depth;
  out->avg_depth = sc->oct.avg_depth;
  return SGR_OK;
}

int sgr_octree_export(const sgr_scene *sc, uint32_t *nodes, uint64_t *num_nodes,
                      uint32_t *surf_ids, uint64_t *num_surf_ids, float *origin_size4) {
  if (!sc || !num_nodes || !num_surf_ids)
    return fail(SGR_ERR_INVALID, "sgr_octree_export: null argument");
  *num_nodes = sc->oct.nodes.size() / 4;
  *num_surf_ids = sc->oct.surf_ids.size();
  if (nodes) std::memcpy(nodes, sc->oct.nodes.data(), sc->oct.nodes.size() * 4);
  if (surf_ids) std::memcpy(surf_ids, sc->oct.surf_ids.data(), sc->oct.surf_ids.size() * 4);
  if (origin_size4) {
    for (int a = 0; a < 3; a++) origin_size4[a] = sc->oct.origin[a];
    origin_size4[3] = sc->oct.size;
  }
  return SGR_OK;
}

int sgr_scene_bounds(const sgr_scene *sc, float *lo, float *hi, float *horizon) {
  if (!sc) return fail(SGR_ERR_INVALID, "sgr_scene_bounds: null scene");
  for (int a = 0; a < 3; a++) {
    if (lo) lo[a] = sc->bbox_lo[a];
    if (hi) hi[a] = sc->bbox_hi[a];
  }
  if (horizon) *horizon = sc->view.horizon;
  return SGR_OK;
}

// Cylinder::xform (surface/cylinder.cc:24-36): translate (0, 0, 1), scale
// (radius, radius, |axis| / 2), to_basis (p, p x az, az) with az =
// unit(axis) and p = az.perpendicular() (geometry/vec.h:119-129, not
// normalised), translate base; each step composes on the left
// (geometry/xform.h:189-262), so the linear columns are radius p,
// radius (p x az), (|axis| / 2) az and the translation (|axis| / 2) az
// + base.
int sgr_cylinder_xform(const float base[3], const float axis[3], float radius, float out[12]) {
  if (!base || !axis || !out) return fail(SGR_ERR_INVALID, "sgr_cylinder_xform: null argument");
  const float len = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  if (!(len > 0.0f) || !std::isfinite(len))
    return fail(SGR_ERR_INVALID, "sgr_cylinder_xform: the axis must be non-zero and finite");
  const float inv = 1.0f / len;  // geometry/vec.h:113-117 unit
  const float az[3] = {axis[0] * inv, axis[1] * inv, axis[2] * inv};
  const float nz = (az[0] == az[2]) ? az[2] : -az[2];
  const float ax[3] = {az[1] * az[1] + az[0] * nz, -nz * nz - az[0] * az[1],
                       az[0] * az[0] - az[1] * nz};
  const float zy[3] = {ax[1] * az[2] - ax[2] * az[1], ax[2] * az[0] - ax[0] * az[2],
                       ax[0] * az[1] - ax[1] * az[0]};
  const float h = len / 2.0f;
  for (int r = 0; r < 3; r++) {
    out[4 * r] = ax[r] * radius;
    out[4 * r + 1] = zy[r] * radius;
    out[4 * r + 2] = az[r] * h;
    out[4 * r + 3] = az[r] * h + base[r];
  }
  return SGR_OK;
}

int sgr_scene_footprint(const sgr_scene *sc, uint64_t *bytes) {
  if (!sc || !bytes) return fail(SGR_ERR_INVALID, "sgr_scene_footprint: null argument");
  *bytes = sc->bytes;
  return SGR_OK;
}

// ---------------------------------------------------------------------------
// Ray batches
// ---------------------------------------------------------------------------
static int trace_common(sgr_scene *sc, const sgr_rays *rays, size_t n, sgr_hits *hits,
                        uint8_t *occ, sgr_stats *stats) {
  if (!sc || !rays) return fail(SGR_ERR_INVALID, "sgr_trace: null argument");
  if (n >= (1ull << 31)) return fail(SGR_ERR_LIMIT, "sgr_trace: batch too large");
  if (n == 0) return SGR_OK;
  sgr_ctx *ctx = sc->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  size_t in_bytes = n * 8 * sizeof(float);
  size_t out_bytes = hits ? n * 20 : n;
  void *buf = nullptr;
  HIPCHK(hipMalloc(&buf, in_bytes + out_bytes));
  float *f = (float *)buf;
  const float *src[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->t0, rays->t1};
  hipError_t e = hipSuccess;
  for (int k = 0; k < 8 && e == hipSuccess; k++)
    e = hipMemcpyAsync(f + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(ctx->stats, 0, SGR_STATS_BYTES, ctx->stream);
  RayBatchDev rb;
  rb.ox = f; rb.oy = f + n; rb.oz = f + 2 * n; rb.dx = f + 3 * n; rb.dy = f + 4 * n;
  rb.dz = f + 5 * n; rb.t0 = f + 6 * n; rb.t1 = f + 7 * n;
  uint32_t *ob = (uint32_t *)(f + 8 * n);
  rb.prim = ob;
  rb.t = (float *)(ob + n);
  rb.u = (float *)(ob + 2 * n);
  rb.v = (float *)(ob + 3 * n);
  rb.inst = hits ? ob + 4 * n : nullptr;
  rb.occ = (uint8_t *)ob;
  if (e == hipSuccess) {
    hipEventRecord(ctx->ev[0], ctx->stream);
    launch_trace(sc->view, rb, (uint32_t)n, hits == nullptr, ctx->stats, ctx->stream);
    e = hipGetLastError();
    hipEventRecord(ctx->ev[1], ctx->stream);
  }
  if (e == hipSuccess) {
    if (hits) {
      e = hipMemcpyAsync(hits->prim, rb.prim, n * 4, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(hits->t, rb.t, n * 4, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(hits->u, rb.u, n * 4, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipMemcpyAsync(hits->v, rb.v, n * 4, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess && hits->inst)
        e = hipMemcpyAsync(hits->inst, rb.inst, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    } else {
      e = hipMemcpyAsync(occ, rb.occ, n, hipMemcpyDeviceToHost, ctx->stream);
    }
  }
  unsigned long long st[SGR_NUM_STATS] = {0};
  if (e == hipSuccess)
    e = hipMemcpyAsync(st, ctx->stats, sizeof st, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && stats) {
    std::memset(stats, 0, sizeof *stats);
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    stats->ms_render = ms;
    if (hits) {
      stats->scene_intersect_calls = n;
      stats->node_visits_closest = st[0];
      stats->surf_tests_closest = st[1];
      stats->node_fetches_closest = st[SGR_STAT_FETCH_C];
      stats->ms_extend = ms;
      stats->extend_launches = 1;
    } else {
      stats->shadow_tests = n;
      stats->node_visits_shadow = st[0];
      stats->surf_tests_shadow = st[1];
      stats->node_fetches_shadow = st[SGR_STAT_FETCH_C];
      stats->ms_shadow = ms;
      stats->shadow_launches = 1;
    }
  }
  (void)hipFree(buf);
  if (e != hipSuccess) return fail(SGR_ERR_HIP, std::string("sgr_trace: ") + hipGetErrorString(e));
  return SGR_OK;
}

int sgr_trace_closest(sgr_scene *sc, const sgr_rays *rays, size_t n, sgr_hits *hits,
                      sgr_stats *stats) {
  if (!hits) return fail(SGR_ERR_INVALID, "sgr_trace_closest: null hits");
  return trace_common(sc, rays, n, hits, nullptr, stats);
}

int sgr_trace_occluded(sgr_scene *sc, const sgr_rays *rays, size_t n, uint8_t *occluded,
                       sgr_stats *stats) {
  if (!occluded) return fail(SGR_ERR_INVALID, "sgr_trace_occluded: null output");
  return trace_common(sc, rays, n, nullptr, occluded, stats);
}

// ---------------------------------------------------------------------------
// Rendering
// ---------------------------------------------------------------------------
}  // extern "C"

namespace sgr {

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Carve the path buffers for CAP paths with K shadow slots each out of
// one allocation.
size_t path_bytes(size_t cap, size_t K, PathBufs *P, char *base) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> char * {
    char *r = base ? base + off : nullptr;
    off += align_up(bytes);
    return r;
  };
  for (int k = 0; k < 8; k++) {
    char *r = take(cap * 4);
    if (P) P->ray[k] = (float *)r;
  }
  char *r = take(cap * 4); if (P) P->hprim = (uint32_t *)r;
  r = take(cap * 4); if (P) P->ht = (float *)r;
  r = take(cap * 4); if (P) P->hu = (float *)r;
  r = take(cap * 4); if (P) P->hv = (float *)r;
  r = take(cap * 4); if (P) P->hinst = (uint32_t *)r;
  for (int k = 0; k < 3; k++) { r = take(cap * 4); if (P) P->T[k] = (float *)r; }
  for (int k = 0; k < 3; k++) { r = take(cap * 4); if (P) P->rad[k] = (float *)r; }
  for (int k = 0; k < 3; k++) { r = take(cap * 4); if (P) P->Tp[k] = (float *)r; }
  r = take(cap * 4); if (P) P->rr = (float *)r;
  r = take(cap * 4); if (P) P->alpha = (float *)r;
  r = take(cap * 4); if (P) P->state = (uint32_t *)r;
  r = take(cap * 4 * SGR_MEDIA_MAX); if (P) P->media = (float *)r;
  size_t ns = cap * (K ? K : 1);
  for (int k = 0; k < 8; k++) { r = take(ns * 4); if (P) P->sray[k] = (float *)r; }
  for (int k = 0; k < 3; k++) { r = take(ns * 4); if (P) P->sc[k] = (float *)r; }
  r = take(ns); if (P) P->svis = (uint8_t *)r;
  r = take(ns * 4); if (P) P->smed = (float *)r;
  for (int k = 0; k < 3; k++) { r = take(ns * 4); if (P) P->strans[k] = (float *)r; }
  r = take(cap * 4); if (P) P->q[0] = (uint32_t *)r;
  r = take(cap * 4); if (P) P->q[1] = (uint32_t *)r;
  r = take(ns * 4); if (P) P->sq = (uint32_t *)r;
  r = take(cap); if (P) P->qflag = (uint8_t *)r;
  r = take(ns); if (P) P->sflag = (uint8_t *)r;
  r = take(((cap + SGR_COMPACT_ITEMS - 1) / SGR_COMPACT_ITEMS) * 2 * 4);
  if (P) P->cblk = (uint32_t *)r;
  r = take(SGR_NUM_COUNTERS * 4); if (P) P->counts = (uint32_t *)r;
  r = take(cap * 6 * 4); if (P) P->out = (float *)r;
  r = take(cap * 4); if (P) P->pseg = (uint32_t *)r;
  return off;
}

int plan_render(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p, RenderPlan &pl) {
  if (!sc || !cam || !p) return fail(SGR_ERR_INVALID, "sgr_render: null argument");
  if (p->width == 0 || p->height == 0 || p->spp == 0)
    return fail(SGR_ERR_INVALID, "sgr_render: width, height and spp must be > 0");
  if (p->max_path_len + 2 > SGR_MAX_BOUNCES)
    return fail(SGR_ERR_LIMIT, "sgr_render: max_path_len above " + std::to_string(SGR_MAX_BOUNCES - 2));
  if (!(p->filter_radius > 0.5001f))
    return fail(SGR_ERR_INVALID, "sgr_render: filter_radius must exceed 0.5");
  RenderView &R = pl.R;
  std::memset(&R, 0, sizeof R);
  uint32_t x0 = p->x0, y0 = p->y0, w = p->w, h = p->h;
  if (w == 0) { x0 = 0; y0 = 0; w = p->width; h = p->height; }
  if (h == 0 || (uint64_t)x0 + w > p->width || (uint64_t)y0 + h > p->height)
    return fail(SGR_ERR_INVALID, "sgr_render: window outside the image");
  int fr = (int)std::ceil(p->filter_radius - 0.5001f);  // image/image-filter-conv.h:39-40
  if (fr > 12)
    return fail(SGR_ERR_LIMIT, "sgr_render: filter_radius above 12.5 is not supported");
  R.width = p->width; R.height = p->height; R.spp = p->spp;
  R.minp = p->min_path_len; R.maxp = p->max_path_len; R.ds = p->direct_samples;
  R.L = sc->n_lights;
  uint32_t ds = R.ds ? R.ds : 1;
  R.g_cam = grid_shape(R.spp);
  R.g_cont = grid_shape(R.spp);
  R.g_dir = grid_shape(R.spp * ds);
  R.nsub_strat = R.g_dir.total / R.spp;      // render/sample-set.h:330-345
  R.nsub_rand = grid_shape(ds).total;
  uint32_t nsub_max = std::max(R.nsub_strat, R.nsub_rand);
  R.K = 2 * R.L * nsub_max;
  R.integ = p->integrator;
  if (R.integ != SGR_INTEG_PATH && R.integ != SGR_INTEG_DIRECT && R.integ != SGR_INTEG_PHOTON)
    return fail(SGR_ERR_INVALID, "sgr_render: unknown integrator " + std::to_string(R.integ));
  std::memset(&R.ph, 0, sizeof R.ph);
  if (R.integ == SGR_INTEG_DIRECT || R.integ == SGR_INTEG_PHOTON) {
    // render/direct-integ.h:72-77: "light_samples", 16 by default; the
    // shadow rays are traced where they are made (no shadow slots).  The
    // photon integrator's direct lighting takes its map's direct_samples
    // (render/photon-integ.cc:44-46).
    uint32_t ls = p->light_samples ? p->light_samples : 16u;
    if (R.integ == SGR_INTEG_PHOTON) {
      int rc = photon_render_view(sc, p->photon_map, R.spp, R.ph, ls);
      if (rc) return rc;
    }
    if ((uint64_t)R.spp * ls >= (1ull << 31))
      return fail(SGR_ERR_LIMIT, "sgr_render: spp x light_samples above 2^31");
    R.g_dint = grid_shape(R.spp * ls);
    R.nsub_dint = R.g_dint.total / R.spp;
    R.K = 0;
  }
  if (R.K > SGR_MAX_SLOTS)
    return fail(SGR_ERR_LIMIT, "sgr_render: lights x direct samples needs " + std::to_string(R.K) +
                                   " shadow slots per path (limit " + std::to_string(SGR_MAX_SLOTS) + ")");
  R.seed = p->seed;
  R.bg_alpha = p->bg_alpha;
  R.fb = p->filter_b; R.fc = p->filter_c;
  R.inv_radius = 1.0f / p->filter_radius;
  R.neg_clamp = -std::fabs(p->neg_clamp);
  R.fr = fr;
  R.X0 = (int)x0 - fr; R.Y0 = (int)y0 - fr;
  R.PW = (int)w + 2 * fr; R.PH = (int)h + 2 * fr;
  for (int a = 0; a < 3; a++) {
    R.cam_pos[a] = cam->pos[a]; R.fwd[a] = cam->forward[a];
    R.up[a] = cam->up[a]; R.right[a] = cam->right[a];
  }
  R.tanx = cam->tan_half_fov_x; R.tany = cam->tan_half_fov_y;
  // Renderer::render_packet: max_trace = (scene bbox + camera).diameter()
  float lo[3], hi[3];
  for (int a = 0; a < 3; a++) {
    lo[a] = fmin_t(sc->bbox_lo[a], cam->pos[a]);
    hi[a] = fmax_t(sc->bbox_hi[a], cam->pos[a]);
  }
  float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  R.max_trace = std::sqrt(ex * ex + ey * ey + ez * ez);
  pl.npix = (uint32_t)(R.PW * R.PH);
  uint64_t want = p->paths_in_flight ? p->paths_in_flight : SGR_DEFAULT_PATHS_IN_FLIGHT;
  uint64_t sb = want / pl.npix;
  if (sb < 1) sb = 1;
  if (sb > R.spp) sb = R.spp;
  if ((uint64_t)pl.npix * sb >= (1ull << 31))
    return fail(SGR_ERR_LIMIT, "sgr_render: image too large for one chunk");
  pl.sb_max = (uint32_t)sb;
  pl.flags = p->flags;
  return SGR_OK;
}

int ensure_work(sgr_ctx *ctx, size_t bytes) {
  if (ctx->work_bytes >= bytes) return SGR_OK;
  if (ctx->work) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &ls : ctx->lane_stream) (void)hipStreamSynchronize(ls);
    (void)hipFree(ctx->work);
    ctx->bytes -= ctx->work_bytes;
    ctx->work = nullptr;
    ctx->work_bytes = 0;
  }
  hipError_t e = hipMalloc(&ctx->work, bytes);
  if (e != hipSuccess) {
    ctx->work = nullptr;
    (void)hipGetLastError();  // a caller may retry smaller: leave no error behind
    return fail(SGR_ERR_NOMEM, std::string("render buffers: ") + hipGetErrorString(e));
  }
  ctx->work_bytes = bytes;
  ctx->bytes += bytes;
  return SGR_OK;
}

}  // namespace sgr

namespace {

// Runs the chunks; per chunk, optional host copy of the per-sample output.
// WORK_EXT: path buffers supplied by the caller (one chunk lane, no
// per-launch timing: nothing here allocates device memory or events).
int run_render(sgr_scene *sc, RenderPlan &pl, float *accum_dev, float *samples_host,
               sgr_stats *stats, sgr_guide *guide = nullptr, bool train = false,
               void *work_ext = nullptr) {
  sgr_ctx *ctx = sc->ctx;
  RenderView R = pl.R;
  size_t cap = (size_t)pl.npix * pl.sb_max;
  // Up to SGR_CHUNK_LANES chunk lanes (own path buffers, own stream)
  // when the frame has several chunks: chunks run concurrently, so one
  // chunk's sparse, latency-bound last bounces overlap another's full
  // ones.  Per-sample renders keep one lane (they copy each chunk out).
  // Guided renders use two: the chunks of a pair render concurrently with
  // the same weights, and the guide trains on both after the pair.
  const uint32_t nchunks = (R.spp + pl.sb_max - 1) / pl.sb_max;
  const int max_lanes = guide ? std::min(SGR_CHUNK_LANES, 2) : SGR_CHUNK_LANES;
  int lanes = (samples_host || work_ext || (pl.flags & SGR_RENDER_ONE_CHUNK_IN_FLIGHT))
                  ? 1 : (int)std::min<uint32_t>(nchunks, (uint32_t)max_lanes);
  size_t need = (path_bytes(cap, R.K, nullptr, nullptr) + 255) & ~(size_t)255;
  int rc = SGR_OK;
  if (!work_ext) {
    rc = ensure_work(ctx, need * lanes);
    if (rc && lanes > 1) {  // not enough device memory for a second lane: one
      lanes = 1;
      rc = ensure_work(ctx, need);
    }
  }
  if (rc) return rc;
  char *work = work_ext ? (char *)work_ext : (char *)ctx->work;
  PathBufs PL[SGR_CHUNK_LANES];
  hipStream_t lane_st[SGR_CHUNK_LANES];
  for (int l = 0; l < lanes; l++) {
    path_bytes(cap, R.K, &PL[l], work + need * l);
    lane_st[l] = l ? ctx->lane_stream[l - 1] : ctx->stream;
  }
  GuideView GV[2];
  for (int l = 0; guide && l < lanes; l++) {
    rc = guide_prepare(guide, sc, cap, l, GV[l]);
    if (rc) return rc;
  }
  hipEvent_t ev_trained = nullptr;  // the last training step of the guide
  HIPCHK(hipMemsetAsync(ctx->stats, 0, SGR_NUM_STATS * sizeof(unsigned long long), ctx->stream));
  unsigned grid = (unsigned)std::min<size_t>((cap + 255) / 256, (size_t)ctx->num_cu * 16);
  grid = (grid + 7) & ~7u;  // whole blocks per XCD (see block_range)
  std::vector<float> chunk_out;
  uint32_t n_ext = 0, n_occ = 0;
  // Per-launch timing: event pairs recorded on the render stream without
  // host waits, read after the final synchronisation.
  enum { K_EXT, K_SHADE, K_OCC, K_SPLAT, K_OTHER, K_N };
  double kms[K_N] = {0, 0, 0, 0, 0};
  const bool kt = stats && !work_ext && (pl.flags & SGR_RENDER_KERNEL_TIMING);
  std::vector<std::pair<int, size_t>> marks;
  size_t next_ev = 0;
  auto ev_get = [&]() -> hipEvent_t {
    if (next_ev == ctx->evpool.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      ctx->evpool.push_back(e);
    }
    return ctx->evpool[next_ev++];
  };
  size_t mark0 = 0;
  hipStream_t cur = ctx->stream;  // stream of the chunk being enqueued
  auto t_begin = [&]() {
    if (!kt) return;
    mark0 = next_ev;
    (void)hipEventRecord(ev_get(), cur);
  };
  auto t_end = [&](int kind) {
    if (!kt) return;
    (void)hipEventRecord(ev_get(), cur);
    marks.push_back({kind, mark0});
  };
  size_t next_sync = 0;
  auto sync_get = [&]() -> hipEvent_t {
    if (next_sync == ctx->syncpool.size()) {
      hipEvent_t e;
      (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
      ctx->syncpool.push_back(e);
    }
    return ctx->syncpool[next_sync++];
  };
  HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
  hipEvent_t ev_splat = nullptr;  // of the previous chunk
  if (lanes > 1) {  // the other lanes start after the stats are cleared
    hipEvent_t e = sync_get();
    HIPCHK(hipEventRecord(e, ctx->stream));
    for (int l = 1; l < lanes; l++) HIPCHK(hipStreamWaitEvent(lane_st[l], e, 0));
  }
  StepView J;
  std::memset(&J, 0, sizeof J);
  R.path_base = 0;
  R.out_base = 0;
  R.scene = 0;
  R.accum = accum_dev;
  // Short late queues go to k_finish (unguided renders): from bounce
  // SGR_FINISH_FROM on, a queue of fewer than SGR_FINISH_MAX rays.
  const bool finish = !guide && !(pl.flags & SGR_RENDER_WAVEFRONT_ONLY);
  for (int l = 0; l < lanes; l++) {
    PL[l].K = R.K ? R.K : 1;
    PL[l].partial = sc->view.has_partial;
    PL[l].fin_from = SGR_FINISH_FROM;
    PL[l].fin_max = finish ? SGR_FINISH_MAX : 0u;
  }
  uint32_t chunk = 0;
  for (uint32_t s0 = 0; s0 < R.spp; s0 += pl.sb_max, chunk++) {
    const int lane = (int)(chunk % (uint32_t)lanes);
    PathBufs &P = PL[lane];
    cur = lane_st[lane];
    R.s0 = s0;
    R.sb = std::min(pl.sb_max, R.spp - s0);
    R.npaths = pl.npix * R.sb;
    P.n = (uint32_t)cap;
    const GuideView *G = guide ? &GV[lane] : nullptr;
    if (guide) {
      // lane 1 waits for the training on the previous pair (which has also
      // read its records); lane 0 is the training stream itself
      if (lane && ev_trained) HIPCHK(hipStreamWaitEvent(cur, ev_trained, 0));
      rc = guide_begin_chunk(guide, GV[lane], R.npaths, R.seed * 0x9e3779b97f4a7c15ULL + s0 + 1,
                             cur);
      if (rc) return rc;
    }
    t_begin();
    launch_raygen(false, R, J, P, grid, cur);
    t_end(K_OTHER);
    if (R.integ == SGR_INTEG_DIRECT || R.integ == SGR_INTEG_PHOTON) {
      t_begin();
      launch_direct_integ(sc->view, R, P, ctx->stats, grid, cur);
      t_end(K_OTHER);
    }
    for (uint32_t b = 0; R.integ == SGR_INTEG_PATH && b <= R.maxp; b++) {
      if (finish && b >= SGR_FINISH_FROM) {
        t_begin();
        launch_finish(false, sc->view, R, J, P, b, ctx->stats, grid, cur);
        t_end(K_OTHER);
      }
      t_begin();
      launch_extend(false, sc->view, J, P, b, ctx->stats, grid, cur);
      t_end(K_EXT);
      if (guide) {
        t_begin();
        guide_infer(guide, *G, P, b, grid, cur);
        t_end(K_OTHER);
      }
      t_begin();
      launch_shade(false, sc->view, R, J, P, b, ctx->stats, grid, cur, G);
      t_end(K_SHADE);
      t_begin();
      launch_occlude(false, sc->view, J, P, b, ctx->stats, grid, cur);
      t_end(K_OCC);
      t_begin();
      launch_gather(false, R, J, P, b, grid, cur, G);
      t_end(K_OTHER);
      n_ext++;
      n_occ++;
    }
    if (accum_dev) {
      // splats reach the accumulators in chunk order
      if (ev_splat) HIPCHK(hipStreamWaitEvent(cur, ev_splat, 0));
      t_begin();
      launch_splat(false, R, J, P.out, grid, cur);
      t_end(K_SPLAT);
      if (lanes > 1) {
        ev_splat = sync_get();
        HIPCHK(hipEventRecord(ev_splat, cur));
      }
    }
    if (R.integ == SGR_INTEG_PATH) launch_count_queues(R.maxp, P, ctx->stats, cur);
    HIPCHK(hipGetLastError());
    if (guide && train && (lane == lanes - 1 || s0 + pl.sb_max >= R.spp)) {
      // the pair (or last chunk) is done: train on its chunks in order
      if (lane) {
        hipEvent_t e = sync_get();
        HIPCHK(hipEventRecord(e, cur));
        HIPCHK(hipStreamWaitEvent(ctx->stream, e, 0));
      }
      cur = ctx->stream;
      t_begin();
      for (int l = 0; l <= lane && !rc; l++) rc = guide_train_chunk(guide, GV[l], GV[l].rec_count);
      t_end(K_OTHER);
      if (rc) return rc;
      if (lanes > 1) {
        ev_trained = sync_get();
        HIPCHK(hipEventRecord(ev_trained, ctx->stream));
      }
    }
    if (samples_host) {
      chunk_out.resize((size_t)R.npaths * 6);
      HIPCHK(hipMemcpyAsync(chunk_out.data(), P.out, chunk_out.size() * 4, hipMemcpyDeviceToHost,
                            cur));
    }
    if (samples_host) {
      HIPCHK(hipStreamSynchronize(cur));
      for (uint32_t pix = 0; pix < pl.npix; pix++)
        std::memcpy(samples_host + ((size_t)pix * R.spp + s0) * 6, &chunk_out[(size_t)pix * R.sb * 6],
                    (size_t)R.sb * 6 * 4);
    }
  }
  for (int l = 1; l < lanes; l++) {  // join the lanes back into the render stream
    hipEvent_t e = sync_get();
    HIPCHK(hipEventRecord(e, lane_st[l]));
    HIPCHK(hipStreamWaitEvent(ctx->stream, e, 0));
  }
  HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
  if (guide) {
    rc = guide_finish(guide);
    if (rc) return rc;
  }
  unsigned long long st[SGR_NUM_STATS];
  HIPCHK(hipMemcpyAsync(st, ctx->stats, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    uint32_t w = pl.R.PW - 2 * pl.R.fr, h = pl.R.PH - 2 * pl.R.fr;
    stats->eye_rays = (uint64_t)pl.npix * R.spp;
    stats->paths_traced = (uint64_t)pl.npix * R.spp;
    stats->samples = (uint64_t)w * h * R.spp;
    stats->scene_intersect_calls = st[SGR_STAT_ISECT];
    stats->shadow_tests = st[SGR_STAT_SHADOW];
    stats->node_visits_closest = st[SGR_STAT_NODES_C];
    stats->surf_tests_closest = st[SGR_STAT_TESTS_C];
    stats->node_visits_shadow = st[SGR_STAT_NODES_S];
    stats->surf_tests_shadow = st[SGR_STAT_TESTS_S];
    stats->illum_calls = st[SGR_STAT_ILLUM];
    stats->media_overflows = st[SGR_STAT_MEDIA_OVF];
    stats->partial_overflows = st[SGR_STAT_PARTIAL_OVF];
    stats->direct_overflows = st[SGR_STAT_DINT_OVF];
    stats->node_fetches_closest = st[SGR_STAT_FETCH_C];
    stats->node_fetches_shadow = st[SGR_STAT_FETCH_S];
    for (auto &m : marks) {
      float e = 0;
      (void)hipEventElapsedTime(&e, ctx->evpool[m.second], ctx->evpool[m.second + 1]);
      kms[m.first] += e;
    }
    stats->ms_render = ms;
    stats->ms_extend = kms[K_EXT];
    stats->ms_shade = kms[K_SHADE];
    stats->ms_shadow = kms[K_OCC];
    stats->ms_splat = kms[K_SPLAT];
    stats->ms_other = kms[K_OTHER];
    stats->extend_launches = n_ext;
    stats->shadow_launches = n_occ;
  }
  return SGR_OK;
}

// sgr_render / sgr_render_guided: host accumulators go through a
// temporary device copy.
int render_accum(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p, float *accum,
                 int accum_on_device, sgr_stats *stats, sgr_guide *guide, bool train) {
  RenderPlan pl;
  int rc = plan_render(sc, cam, p, pl);
  if (rc) return rc;
  if (guide && pl.R.integ != SGR_INTEG_PATH)
    return fail(SGR_ERR_INVALID, "sgr_render_guided: the learned proposal guides the path integrator only");
  if (!accum) return fail(SGR_ERR_INVALID, "sgr_render: null accumulators");
  sgr_ctx *ctx = sc->ctx;
  HIPCHK(hipSetDevice(ctx->device));
  size_t abytes = (size_t)p->width * p->height * 5 * sizeof(float);
  float *dacc = accum;
  if (!accum_on_device) {
    HIPCHK(hipMalloc((void **)&dacc, abytes));
    hipError_t e = hipMemcpyAsync(dacc, accum, abytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(dacc);
      return fail(SGR_ERR_HIP, std::string("sgr_render: ") + hipGetErrorString(e));
    }
  }
  rc = run_render(sc, pl, dacc, nullptr, stats, guide, train);
  if (!accum_on_device) {
    if (!rc) {
      hipError_t e = hipMemcpyAsync(accum, dacc, abytes, hipMemcpyDeviceToHost, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess) rc = fail(SGR_ERR_HIP, std::string("sgr_render: ") + hipGetErrorString(e));
    }
    (void)hipFree(dacc);
  }
  return rc;
}

}  // namespace

extern "C" {

int sgr_render(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p, float *accum,
               int accum_on_device, sgr_stats *stats) {
  return sgr::guarded("sgr_render", [&] {
    return render_accum(sc, cam, p, accum, accum_on_device, stats, nullptr, false);
  });
}

int sgr_render_guided(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p,
                      sgr_guide *guide, int train, float *accum, int accum_on_device,
                      sgr_stats *stats) {
  if (!guide) return fail(SGR_ERR_INVALID, "sgr_render_guided: null guide");
  return sgr::guarded("sgr_render_guided", [&] {
    return render_accum(sc, cam, p, accum, accum_on_device, stats, guide, train != 0);
  });
}

int sgr_render_samples(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p,
                       float *out, sgr_stats *stats) {
  RenderPlan pl;
  int rc = plan_render(sc, cam, p, pl);
  if (rc) return rc;
  if (!out) return fail(SGR_ERR_INVALID, "sgr_render_samples: null output");
  HIPCHK(hipSetDevice(sc->ctx->device));
  return sgr::guarded("sgr_render_samples", [&] { return run_render(sc, pl, nullptr, out, stats); });
}

// sgr_render into device accumulators with the path state in the
// device's scratch array: chunks of as many samples per pattern pixel as
// the scratch holds (at most the plan's own chunk), one chunk in flight.
static int render_noalloc_impl(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p,
                               float *accum_dev, sgr_stats *stats) {
  if (!accum_dev) return fail(SGR_ERR_INVALID, "sgr_render_noalloc: null accumulators");
  RenderPlan pl;
  int rc = plan_render(sc, cam, p, pl);
  if (rc) return rc;
  sgr_ctx *ctx = sc->ctx;
  if (!ctx->noalloc_scratch)
    return fail(SGR_ERR_HIP, "sgr_render_noalloc: the scratch array's device address is unknown");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t K = pl.R.K;
  uint32_t sb = pl.sb_max;
  while (sb > 0 && path_bytes((size_t)pl.npix * sb, K, nullptr, nullptr) > SGR_NOALLOC_SCRATCH_BYTES)
    sb = sb > 8 ? sb - sb / 8 : sb - 1;  // shrink by eighths, then one at a time
  if (sb == 0)
    return fail(SGR_ERR_LIMIT,
                "sgr_render_noalloc: " + std::to_string(pl.npix) +
                    " pattern pixels (window plus filter margin) need " +
                    std::to_string(path_bytes(pl.npix, K, nullptr, nullptr)) +
                    " bytes of path state for one sample each; the scratch holds " +
                    std::to_string((unsigned long long)SGR_NOALLOC_SCRATCH_BYTES));
  pl.sb_max = sb;
  return run_render(sc, pl, accum_dev, nullptr, stats, nullptr, false, ctx->noalloc_scratch);
}

int sgr_render_noalloc(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p,
                       float *accum_dev, sgr_stats *stats) {
  return sgr::guarded("sgr_render_noalloc",
                      [&] { return render_noalloc_impl(sc, cam, p, accum_dev, stats); });
}

// ImageSampledOutput::set_raw_min_y (image/image-sampled-output.cc:63-126)
int sgr_finalize_image(const float *accum, uint32_t width, uint32_t height, float exposure,
                       float contrast, float *rgba) {
  if (!accum || !rgba) return fail(SGR_ERR_INVALID, "sgr_finalize_image: null argument");
  for (size_t i = 0; i < (size_t)width * height; i++) {
    const float *a = accum + 5 * i;
    float c[3] = {a[0], a[1], a[2]};
    float alpha = a[3], w = a[4];
    if (w > 0 && w != 1) {
      float iw = 1 / w;
      for (float &x : c) x *= iw;
      alpha *= iw;
    }
    alpha = alpha < 1.f ? alpha : 1.f;
    alpha = alpha > 0.f ? alpha : 0.f;
    for (float &x : c) x = x > 0.f ? x : 0.f;
    if (alpha != 0 && alpha != 1) {
      float ia = 1 / alpha;
      for (float &x : c) x *= ia;
    }
    if (exposure != 1)
      for (float &x : c) x *= exposure;
    if (contrast != 1)
      for (float &x : c) x = std::pow(x > 0.f ? x : 0.f, contrast);
    float *o = rgba + 4 * i;
    o[0] = c[0]; o[1] = c[1]; o[2] = c[2]; o[3] = alpha;
  }
  return SGR_OK;
}

}  // extern "C"
