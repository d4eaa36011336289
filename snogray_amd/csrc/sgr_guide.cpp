// Learned directional proposal: the network object, its parameter and
// gradient entry points, and the per-render hooks (inference buffers,
// training records, the training step after each chunk).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "sgr_internal.h"

struct sgr_guide {
  sgr_ctx *ctx = nullptr;
  float *w32 = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *aux = nullptr;
  uint16_t *w16 = nullptr;
  sgr::GuideTrain T{};
  float *upload = nullptr;      // host-supplied rows / records
  // inference rows and training records of each chunk lane
  float *in6[2] = {nullptr, nullptr}, *lobes[2] = {nullptr, nullptr}, *rec[2] = {nullptr, nullptr};
  size_t rows_cap[2] = {0, 0}, rec_cap[2] = {0, 0};
  uint32_t step = 0;
  std::vector<void *> allocs;
  uint64_t bytes = 0;
  sgr_guide_stats st{};
};

namespace sgr {
namespace {

int galloc(sgr_guide *g, size_t bytes, void **out) {
  hipError_t e = hipMalloc(out, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(SGR_ERR_NOMEM, std::string("guide: ") + hipGetErrorString(e));
  g->allocs.push_back(*out);
  g->bytes += bytes;
  g->ctx->bytes += bytes;
  return SGR_OK;
}

void gfree(sgr_guide *g, void *p, size_t bytes) {
  if (!p) return;
  for (auto &a : g->allocs)
    if (a == p) a = nullptr;
  (void)hipFree(p);
  g->bytes -= bytes;
  g->ctx->bytes -= bytes;
}

uint16_t host_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// He-normal initial weights from a splitmix64 stream (Box-Muller); the
// output layer is scaled by 0.1 so the first proposal is a set of broad
// lobes of equal weight.  Padding rows / columns stay zero.
void init_params(uint64_t seed, std::vector<float> &w) {
  w.assign(GD_NPARAM, 0.0f);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9e3779b97f4a7c15ULL;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() {
    double u1 = next(), u2 = next();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  auto fill = [&](int off, int nout, int nin, int nin_used, int nout_used, double scale) {
    const double sd = std::sqrt(2.0 / nin_used) * scale;
    for (int o = 0; o < nout; o++)
      for (int i = 0; i < nin; i++)
        w[off + o * nin + i] = (o < nout_used && i < nin_used) ? float(normal() * sd) : 0.0f;
  };
  fill(GD_W1, GD_H, GD_FEAT, GD_IN + GD_IN * GD_NFREQ * 2, GD_H, 1.0);
  fill(GD_W2, GD_H, GD_H, GD_H, GD_H, 1.0);
  fill(GD_W3, GD_H, GD_H, GD_H, GD_H, 1.0);
  fill(GD_W4, GD_OUTP, GD_H, GD_H, GD_OUT, 0.1);
}

int upload_params(sgr_guide *g, const std::vector<float> &w) {
  std::vector<uint16_t> h(GD_NPARAM);
  for (int i = 0; i < GD_NPARAM; i++) h[i] = host_bf16(w[i]);
  hipStream_t s = g->ctx->stream;
  HIPCHK(hipMemcpyAsync(g->w32, w.data(), GD_NPARAM * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(g->w16, h.data(), GD_NPARAM * 2, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(g->m, 0, GD_NPARAM * 4, s));
  HIPCHK(hipMemsetAsync(g->v, 0, GD_NPARAM * 4, s));
  HIPCHK(hipStreamSynchronize(s));
  g->step = 0;
  return SGR_OK;
}

// Forward + backward + weight gradients + reduction (+ Adam when lr > 0)
// over N records at REC (device).
int train_batch(sgr_guide *g, const float *rec, uint32_t n, float lr) {
  hipStream_t s = g->ctx->stream;
  GuideNet net{g->w32, g->w16};
  launch_mlp_rows(net, g->T, rec, n, s);
  launch_mlp_wgrad(g->T, n, s);
  if (lr > 0.0f) g->step++;
  launch_mlp_adam(g->T, n, g->grad, g->aux, g->w32, g->w16, g->m, g->v, lr, g->step, s);
  HIPCHK(hipGetLastError());
  return SGR_OK;
}

}  // namespace

int guide_prepare(sgr_guide *g, sgr_scene *sc, size_t cap, int lane, GuideView &V) {
  if (g->ctx != sc->ctx) return fail(SGR_ERR_INVALID, "guide and scene belong to different contexts");
  // about 256 k training records per chunk
  uint32_t stride = (uint32_t)std::max<size_t>(1, cap / (256u * 1024u));
  size_t nrec = (cap + stride - 1) / stride;
  if (cap > g->rows_cap[lane]) {
    gfree(g, g->in6[lane], g->rows_cap[lane] * GD_IN * 4);
    gfree(g, g->lobes[lane], g->rows_cap[lane] * GD_OUT * 4);
    g->in6[lane] = g->lobes[lane] = nullptr;
    g->rows_cap[lane] = 0;
    int rc = galloc(g, cap * GD_IN * 4, (void **)&g->in6[lane]);
    if (!rc) rc = galloc(g, cap * GD_OUT * 4, (void **)&g->lobes[lane]);
    if (rc) return rc;
    g->rows_cap[lane] = cap;
  }
  if (nrec > g->rec_cap[lane]) {
    gfree(g, g->rec[lane], g->rec_cap[lane] * GD_REC * 4);
    g->rec[lane] = nullptr;
    g->rec_cap[lane] = 0;
    int rc = galloc(g, nrec * GD_REC * 4, (void **)&g->rec[lane]);
    if (rc) return rc;
    g->rec_cap[lane] = nrec;
  }
  std::memset(&V, 0, sizeof V);
  V.net = GuideNet{g->w32, g->w16};
  V.in6 = g->in6[lane];
  V.lobes = g->lobes[lane];
  V.rec = g->rec[lane];
  V.rec_stride = stride;
  V.rec_count = (uint32_t)nrec;
  for (int a = 0; a < 3; a++) {
    float ext = sc->bbox_hi[a] - sc->bbox_lo[a];
    V.box_lo[a] = sc->bbox_lo[a];
    V.box_inv[a] = ext > 0.0f ? 1.0f / ext : 0.0f;
  }
  V.mix = 0.5f;
  return SGR_OK;
}

// Record count and slot key of a chunk of NPATHS paths; clears its
// records (before raygen, on the chunk's stream S).
int guide_begin_chunk(sgr_guide *g, GuideView &V, uint32_t npaths, uint64_t key, hipStream_t s) {
  V.rec_count = (npaths + V.rec_stride - 1) / V.rec_stride;
  V.rec_key = key;
  HIPCHK(hipMemsetAsync(V.rec, 0, (size_t)V.rec_count * GD_REC * 4, s));
  return SGR_OK;
}

// Network inputs and lobes of bounce B's queue rows (before shading).
void guide_infer(sgr_guide *g, const GuideView &V, const PathBufs &P, uint32_t b, unsigned grid,
                 hipStream_t s) {
  launch_guide_inputs(V, P.q[b & 1], &P.counts[SGR_CNT_Q + b], P.ray, P.ht, P.hprim, grid, s);
  launch_mlp_forward(V.net, V.in6, 0, &P.counts[SGR_CNT_Q + b], V.lobes, 1, grid, s);
}

// Adam steps over the chunk's records in minibatches.
int guide_train_chunk(sgr_guide *g, const GuideView &V, uint32_t nrec) {
  for (uint32_t off = 0; off < nrec; off += GD_BATCH) {
    uint32_t n = std::min<uint32_t>(GD_BATCH, nrec - off);
    int rc = train_batch(g, V.rec + (size_t)off * GD_REC, n, 1e-3f);
    if (rc) return rc;
    g->st.train_steps++;
    g->st.records += n;
  }
  return SGR_OK;
}

int guide_finish(sgr_guide *g) {
  float aux[2];
  HIPCHK(hipMemcpyAsync(aux, g->aux, sizeof aux, hipMemcpyDeviceToHost, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  g->st.last_loss = aux[1] > 0.0f ? aux[0] / aux[1] : 0.0;
  return SGR_OK;
}

}  // namespace sgr

using namespace sgr;

extern "C" {

static int sgr_guide_create_impl(sgr_ctx *ctx, uint64_t seed, sgr_guide **out) {
  if (!ctx || !out) return fail(SGR_ERR_INVALID, "sgr_guide_create: null argument");
  HIPCHK(hipSetDevice(ctx->device));
  sgr_guide *g = new sgr_guide;
  g->ctx = ctx;
  int rc = SGR_OK;
  auto A = [&](size_t bytes, void **p) {
    if (!rc) rc = galloc(g, bytes, p);
  };
  A(GD_NPARAM * 4, (void **)&g->w32);
  A(GD_NPARAM * 2, (void **)&g->w16);
  A(GD_NPARAM * 4, (void **)&g->m);
  A(GD_NPARAM * 4, (void **)&g->v);
  A(GD_NPARAM * 4, (void **)&g->grad);
  A(64, (void **)&g->aux);
  A((size_t)GD_BATCH * GD_FEAT * 2, (void **)&g->T.a0);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.a1);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.a2);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.a3);
  A((size_t)GD_BATCH * GD_OUTP * 2, (void **)&g->T.d4);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.d3);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.d2);
  A((size_t)GD_BATCH * GD_H * 2, (void **)&g->T.d1);
  A((size_t)(GD_BATCH / 16) * 2 * 4, (void **)&g->T.tile_loss);
  A((size_t)GD_WG_SLICES * GD_NPARAM * 4, (void **)&g->T.wpart);
  A((size_t)GD_BATCH * GD_REC * 4, (void **)&g->upload);
  if (!rc) {
    std::vector<float> w;
    init_params(seed, w);
    rc = upload_params(g, w);
  }
  if (rc) {
    std::string msg = g_err;
    sgr_guide_destroy(g);
    return fail(rc, msg);
  }
  *out = g;
  return SGR_OK;
}

int sgr_guide_create(sgr_ctx *ctx, uint64_t seed, sgr_guide **out) {
  return sgr::guarded("sgr_guide_create",
                      [&] { return sgr_guide_create_impl(ctx, seed, out); });
}

int sgr_guide_destroy(sgr_guide *g) {
  if (!g) return SGR_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  for (void *p : g->allocs)
    if (p) (void)hipFree(p);
  g->ctx->bytes -= g->bytes;
  delete g;
  return SGR_OK;
}

int sgr_guide_get_params(sgr_guide *g, float *out, uint64_t *count) {
  if (!g || !count) return fail(SGR_ERR_INVALID, "sgr_guide_get_params: null argument");
  *count = GD_NPARAM;
  if (!out) return SGR_OK;
  HIPCHK(hipSetDevice(g->ctx->device));
  HIPCHK(hipMemcpyAsync(out, g->w32, GD_NPARAM * 4, hipMemcpyDeviceToHost, g->ctx->stream));
  HIPCHK(hipStreamSynchronize(g->ctx->stream));
  return SGR_OK;
}

int sgr_guide_set_params(sgr_guide *g, const float *in, uint64_t count) {
  if (!g || !in) return fail(SGR_ERR_INVALID, "sgr_guide_set_params: null argument");
  if (count != GD_NPARAM)
    return fail(SGR_ERR_INVALID, "sgr_guide_set_params: expected " + std::to_string(GD_NPARAM) +
                                     " parameters");
  HIPCHK(hipSetDevice(g->ctx->device));
  std::vector<float> w(in, in + GD_NPARAM);
  return upload_params(g, w);
}

int sgr_guide_forward(sgr_guide *g, const float *in6, uint64_t n, float *out, int lobes) {
  if (!g || (n && (!in6 || !out))) return fail(SGR_ERR_INVALID, "sgr_guide_forward: null argument");
  if (n > 4096) return fail(SGR_ERR_LIMIT, "sgr_guide_forward: at most 4096 rows per call");
  if (n == 0) return SGR_OK;
  HIPCHK(hipSetDevice(g->ctx->device));
  hipStream_t s = g->ctx->stream;
  float *dout = g->upload + n * GD_IN;  // outputs after the inputs (fits: 4096 x 46 floats)
  HIPCHK(hipMemcpyAsync(g->upload, in6, n * GD_IN * 4, hipMemcpyHostToDevice, s));
  GuideNet net{g->w32, g->w16};
  launch_mlp_forward(net, g->upload, (uint32_t)n, nullptr, dout, lobes, 1024, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, n * GD_OUT * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SGR_OK;
}

int sgr_guide_train_batch(sgr_guide *g, const float *rec, uint64_t n, float lr, float *grad,
                          float *loss) {
  if (!g || !rec) return fail(SGR_ERR_INVALID, "sgr_guide_train_batch: null argument");
  if (n == 0 || n > GD_BATCH)
    return fail(SGR_ERR_LIMIT, "sgr_guide_train_batch: 1 to " + std::to_string(GD_BATCH) +
                                   " records per call");
  if (!(lr >= 0.0f)) return fail(SGR_ERR_INVALID, "sgr_guide_train_batch: negative learning rate");
  HIPCHK(hipSetDevice(g->ctx->device));
  hipStream_t s = g->ctx->stream;
  HIPCHK(hipMemcpyAsync(g->upload, rec, n * GD_REC * 4, hipMemcpyHostToDevice, s));
  int rc = train_batch(g, g->upload, (uint32_t)n, lr);
  if (rc) return rc;
  if (grad) HIPCHK(hipMemcpyAsync(grad, g->grad, GD_NPARAM * 4, hipMemcpyDeviceToHost, s));
  float aux[2];
  HIPCHK(hipMemcpyAsync(aux, g->aux, sizeof aux, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (loss) {
    loss[0] = aux[0];
    loss[1] = aux[1];
  }
  return SGR_OK;
}

int sgr_guide_stats_get(const sgr_guide *g, sgr_guide_stats *st) {
  if (!g || !st) return fail(SGR_ERR_INVALID, "sgr_guide_stats_get: null argument");
  *st = g->st;
  return SGR_OK;
}

}  // extern "C"
