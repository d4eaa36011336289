// Multi-job render engine: many sgr_render jobs share one sequence of
// wavefront launches per step.
//
// A step takes the next sample chunk of each admitted job that fits the
// path capacity (round-robin from where the previous step stopped, one
// chunk per job), lays their paths end to end (segments), uploads one
// table block (per-segment RenderView, per-slot SceneView, prefix offsets,
// header) and runs raygen, bounces 0..max_path_len of extend / shade /
// occlude / gather, splat and the queue count.  The kernels find a path's
// segment from the per-path segment id that raygen writes, so the launch
// sequence depends only on the bounce count; it is captured once per
// bounce count into a hipGraph and replayed with new tables.
//
// Each job keeps the chunking sgr_render would give it (its own
// paths_in_flight), and its chunks run in sample order, one per step, so
// its accumulators receive the same splats in the same order: a job's
// image is bit-identical whether rendered alone or in a batch.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>

#include "sgr_internal.h"

namespace sgr {

#define SGR_MAX_SCENE_SLOTS 64

// Table block uploaded before every step (pinned host copy, one device
// copy; the captured graphs read the device copy).
struct StepTables {
  RenderView R[SGR_MAX_SEGMENTS];
  SceneView S[SGR_MAX_SCENE_SLOTS];
  uint32_t seg_base[SGR_MAX_SEGMENTS + 1];
  uint32_t out_base[SGR_MAX_SEGMENTS + 1];
  uint32_t hdr[4];  // segments, paths, splat tiles
};

struct Job {
  uint64_t id = 0;
  sgr_scene *sc = nullptr;
  RenderPlan pl{};
  float *accum = nullptr;
  bool owns_accum = false;
  uint32_t next_s0 = 0;
  uint32_t sb = 1;  // chunk size in samples
  uint32_t chunks_done = 0, chunks_total = 0, steps = 0;
  int state = SGR_JOB_QUEUED;
  double t_submit = 0, t_start = -1, t_done = -1;
  uint64_t footprint = 0;
  uint32_t tiles_out = 0;  // splat tiles of a chunk
};

struct Engine {
  sgr_engine_config cfg{};
  // path buffers for cap paths with K shadow slots per path
  void *work = nullptr;
  size_t work_bytes = 0, cap = 0, K = 0;
  PathBufs P{};
  StepTables *host = nullptr;  // pinned
  StepTables *dev = nullptr;
  unsigned long long *stats = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // by bounce count | refill << 31 | inst << 30 | partial << 29
  std::map<uint32_t, hipGraphExec_t> graphs;
  std::map<uint64_t, Job> jobs;
  std::deque<uint64_t> queued;   // admission order
  std::vector<uint64_t> active;  // round-robin order
  size_t cursor = 0;
  sgr_scene *slot_scene[SGR_MAX_SCENE_SLOTS] = {};
  uint32_t slot_jobs[SGR_MAX_SCENE_SLOTS] = {};
  uint64_t next_id = 1;
  uint64_t admitted_bytes = 0;
  // engine-wide totals
  double ms_device = 0;
  uint64_t eye_rays = 0, samples = 0;
  uint32_t steps = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  unsigned grid(const sgr_ctx *ctx) const {
    unsigned g = cfg.grid_blocks ? cfg.grid_blocks : (unsigned)ctx->num_cu * 16;
    return (g + 7) & ~7u;  // whole blocks per XCD (block_range)
  }
};

namespace {

void drop_graphs(Engine *E) {
  for (auto &g : E->graphs) (void)hipGraphExecDestroy(g.second);
  E->graphs.clear();
}

int engine_get(sgr_ctx *ctx, Engine *&E) {
  if (!ctx) return fail(SGR_ERR_INVALID, "engine: null context");
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->engine) {
    E = ctx->engine;
    return SGR_OK;
  }
  E = new Engine;
  E->cfg.path_capacity = SGR_DEFAULT_PATHS_IN_FLIGHT;
  E->cfg.use_graph = 1;
  hipError_t e = hipHostMalloc((void **)&E->host, sizeof(StepTables), hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void **)&E->dev, sizeof(StepTables));
  if (e == hipSuccess) e = hipMalloc((void **)&E->stats, SGR_NUM_STATS * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemsetAsync(E->stats, 0, SGR_NUM_STATS * sizeof(unsigned long long), ctx->stream);
  if (e == hipSuccess) e = hipEventCreate(&E->e0);
  if (e == hipSuccess) e = hipEventCreate(&E->e1);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->engine = E;  // engine_destroy frees what was made
  if (e != hipSuccess) {
    engine_destroy(ctx);
    return fail(SGR_ERR_NOMEM, std::string("engine set-up: ") + hipGetErrorString(e));
  }
  std::memset(E->host, 0, sizeof(StepTables));
  ctx->bytes += sizeof(StepTables);
  return SGR_OK;
}

int ensure_engine_work(sgr_ctx *ctx, Engine *E, size_t cap, size_t K) {
  if (E->work && cap <= E->cap && K <= E->K) return SGR_OK;
  cap = std::max(cap, E->cap);
  K = std::max(K, E->K);
  size_t need = path_bytes(cap, K, nullptr, nullptr);
  if (E->work) {
    HIPCHK(hipStreamSynchronize(ctx->stream));
    (void)hipFree(E->work);
    ctx->bytes -= E->work_bytes;
    E->work = nullptr;
    E->work_bytes = 0;
  }
  drop_graphs(E);  // they hold the old buffer addresses
  hipError_t e = hipMalloc(&E->work, need);
  if (e != hipSuccess) return fail(SGR_ERR_NOMEM, std::string("engine path buffers: ") + hipGetErrorString(e));
  E->work_bytes = need;
  ctx->bytes += need;
  E->cap = cap;
  E->K = K;
  path_bytes(cap, K, &E->P, (char *)E->work);
  E->P.n = (uint32_t)cap;
  E->P.K = (uint32_t)K;
  E->P.fin_from = SGR_FINISH_FROM;
  E->P.fin_max = E->cfg.finish_max ? E->cfg.finish_max : SGR_FINISH_MAX;
  return SGR_OK;
}

uint64_t accum_bytes(const sgr_render_params *p) {
  return (uint64_t)p->width * p->height * 5 * sizeof(float);
}

// 16 x 16 tiles of image pixels the splat of a chunk writes (the
// window), as the splat kernel enumerates them.
uint32_t splat_tiles(const RenderView &R) {
  int ow = R.PW - 2 * R.fr, oh = R.PH - 2 * R.fr;  // the window
  if (ow <= 0 || oh <= 0) return 0;
  return (uint32_t)((ow + 15) / 16) * (uint32_t)((oh + 15) / 16);
}

void enqueue_step(sgr_ctx *ctx, Engine *E, uint32_t maxp, bool refill, bool inst,
                  bool partial, unsigned grid) {
  StepView J;
  J.R = E->dev->R;
  J.S = E->dev->S;
  J.seg_base = E->dev->seg_base;
  J.out_base = E->dev->out_base;
  J.hdr = E->dev->hdr;
  RenderView R0;
  std::memset(&R0, 0, sizeof R0);
  SceneView S0;
  std::memset(&S0, 0, sizeof S0);
  // select the walk of the multi-scene kernels: instanced scenes take the
  // plain walk with a stack for the models
  S0.walk_refill = refill && !inst ? 1u : 0u;
  S0.n_insts = inst ? 1u : 0u;
  S0.has_ext = inst ? 1u : 0u;  // and the shading of ellipses, cylinders and sphere2s
  // ... and shadow rays of steps with partial occluders the walk that
  // carries their transmittance
  S0.has_partial = partial ? 1u : 0u;
  PathBufs P = E->P;
  P.partial = partial ? 1u : 0u;
  hipStream_t st = ctx->stream;
  launch_raygen(true, R0, J, P, grid, st);
  for (uint32_t b = 0; b <= maxp; b++) {
    if (b >= SGR_FINISH_FROM) launch_finish(true, S0, R0, J, P, b, E->stats, grid, st);
    launch_extend(true, S0, J, P, b, E->stats, grid, st);
    launch_shade(true, S0, R0, J, P, b, E->stats, grid, st);
    launch_occlude(true, S0, J, P, b, E->stats, grid, st);
    launch_gather(true, R0, J, P, b, grid, st);
  }
  launch_splat(true, R0, J, P.out, grid, st);
  launch_count_queues(maxp, P, E->stats, st);
}

int slot_for(Engine *E, sgr_scene *sc) {
  for (int i = 0; i < SGR_MAX_SCENE_SLOTS; i++)
    if (E->slot_scene[i] == sc) return i;
  for (int i = 0; i < SGR_MAX_SCENE_SLOTS; i++)
    if (!E->slot_scene[i]) return i;
  return -1;
}

// Admits queued jobs in submission order while their footprints fit the
// budget and a scene slot is free; the first job that does not fit
// blocks the ones behind it (no overtaking, so no starvation).
void admit(Engine *E) {
  while (!E->queued.empty()) {
    Job &j = E->jobs[E->queued.front()];
    if (E->cfg.memory_budget && E->admitted_bytes + j.footprint > E->cfg.memory_budget) break;
    int s = slot_for(E, j.sc);
    if (s < 0) break;  // every scene slot is taken by other active jobs
    E->slot_scene[s] = j.sc;
    E->slot_jobs[s]++;
    j.pl.R.scene = (uint32_t)s;
    j.state = SGR_JOB_ACTIVE;
    E->admitted_bytes += j.footprint;
    E->active.push_back(j.id);
    E->queued.pop_front();
  }
}

void retire(Engine *E, Job &j) {
  uint32_t s = j.pl.R.scene;
  if (j.state != SGR_JOB_QUEUED && s < SGR_MAX_SCENE_SLOTS && E->slot_scene[s] == j.sc) {
    if (--E->slot_jobs[s] == 0) E->slot_scene[s] = nullptr;
  }
  if (j.state != SGR_JOB_QUEUED) E->admitted_bytes -= j.footprint;
}

}  // namespace

void engine_destroy(sgr_ctx *ctx) {
  Engine *E = ctx->engine;
  if (!E) return;
  (void)hipStreamSynchronize(ctx->stream);
  drop_graphs(E);
  for (auto &kv : E->jobs) {
    if (kv.second.owns_accum && kv.second.accum) (void)hipFree(kv.second.accum);
    if (kv.second.sc) kv.second.sc->jobs_using--;
  }
  if (E->work) {
    (void)hipFree(E->work);
    ctx->bytes -= E->work_bytes;
  }
  if (E->dev) {
    (void)hipFree(E->dev);
    ctx->bytes -= sizeof(StepTables);
  }
  if (E->host) (void)hipHostFree(E->host);
  if (E->stats) (void)hipFree(E->stats);
  if (E->e0) (void)hipEventDestroy(E->e0);
  if (E->e1) (void)hipEventDestroy(E->e1);
  delete E;
  ctx->engine = nullptr;
}

}  // namespace sgr

using namespace sgr;

extern "C" {

int sgr_engine_configure(sgr_ctx *ctx, const sgr_engine_config *cfg) {
  if (!cfg) return fail(SGR_ERR_INVALID, "sgr_engine_configure: null config");
  Engine *E;
  int rc = engine_get(ctx, E);
  if (rc) return rc;
  if (!E->jobs.empty()) return fail(SGR_ERR_INVALID, "sgr_engine_configure: engine holds jobs");
  uint64_t cap = cfg->path_capacity ? cfg->path_capacity : (SGR_DEFAULT_PATHS_IN_FLIGHT);
  if (cap >= (1ull << 31)) return fail(SGR_ERR_LIMIT, "sgr_engine_configure: path_capacity too large");
  if (cfg->grid_blocks > (1u << 20)) return fail(SGR_ERR_LIMIT, "sgr_engine_configure: grid_blocks too large");
  if (cfg->finish_max > (1u << 24)) return fail(SGR_ERR_LIMIT, "sgr_engine_configure: finish_max too large");
  HIPCHK(hipStreamSynchronize(ctx->stream));
  drop_graphs(E);
  if (E->work) {
    (void)hipFree(E->work);
    ctx->bytes -= E->work_bytes;
    E->work = nullptr;
    E->work_bytes = E->cap = E->K = 0;
  }
  E->cfg = *cfg;
  E->cfg.path_capacity = cap;
  HIPCHK(hipMemsetAsync(E->stats, 0, SGR_NUM_STATS * sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  E->ms_device = 0;
  E->eye_rays = E->samples = 0;
  E->steps = 0;
  return SGR_OK;
}

int sgr_job_footprint(const sgr_render_params *p, int caller_accum, uint64_t *bytes) {
  if (!p || !bytes) return fail(SGR_ERR_INVALID, "sgr_job_footprint: null argument");
  *bytes = caller_accum ? 0 : accum_bytes(p);
  return SGR_OK;
}

static int sgr_job_submit_impl(sgr_ctx *ctx, sgr_scene *sc, const sgr_camera *cam,
                               const sgr_render_params *p, float *accum_dev, uint64_t *job_id) {
  if (!job_id) return fail(SGR_ERR_INVALID, "sgr_job_submit: null job_id");
  if (!sc || sc->ctx != ctx) return fail(SGR_ERR_INVALID, "sgr_job_submit: scene of another context");
  Engine *E;
  int rc = engine_get(ctx, E);
  if (rc) return rc;
  Job j;
  rc = plan_render(sc, cam, p, j.pl);
  if (rc) return rc;
  if (j.pl.R.integ != SGR_INTEG_PATH)
    return fail(SGR_ERR_INVALID, "sgr_job_submit: the job engine renders with the path integrator only");
  uint64_t cap = E->cfg.path_capacity;
  if (j.pl.npix > cap)
    return fail(SGR_ERR_LIMIT, "sgr_job_submit: " + std::to_string(j.pl.npix) +
                                   " pattern pixels exceed the engine's path capacity");
  j.sb = std::min<uint32_t>(j.pl.sb_max, (uint32_t)(cap / j.pl.npix));
  j.chunks_total = (j.pl.R.spp + j.sb - 1) / j.sb;
  j.footprint = accum_dev ? 0 : accum_bytes(p);
  if (E->cfg.memory_budget && j.footprint > E->cfg.memory_budget)
    return fail(SGR_ERR_LIMIT, "sgr_job_submit: job footprint " + std::to_string(j.footprint) +
                                   " exceeds the memory budget");
  // The engine's containers take the job first: if they throw, nothing
  // has been acquired yet.  The accumulators and the scene reference come
  // after, and a failure there removes the entries again.
  j.id = E->next_id++;
  j.sc = sc;
  j.t_submit = E->now_ms();
  auto it = E->jobs.emplace(j.id, j).first;
  try {
    E->queued.push_back(j.id);
  } catch (...) {
    E->jobs.erase(it);
    throw;
  }
  Job &jj = it->second;
  if (accum_dev) {
    jj.accum = accum_dev;
  } else {
    hipError_t e = hipMalloc((void **)&jj.accum, jj.footprint);
    if (e == hipSuccess) e = hipMemsetAsync(jj.accum, 0, jj.footprint, ctx->stream);
    if (e != hipSuccess) {
      if (jj.accum) (void)hipFree(jj.accum);
      E->queued.pop_back();
      E->jobs.erase(it);
      return fail(SGR_ERR_NOMEM, std::string("sgr_job_submit: accumulators: ") + hipGetErrorString(e));
    }
    jj.owns_accum = true;
  }
  jj.pl.R.accum = jj.accum;
  jj.tiles_out = splat_tiles(jj.pl.R);
  sc->jobs_using++;
  *job_id = j.id;
  return SGR_OK;
}

int sgr_job_submit(sgr_ctx *ctx, sgr_scene *sc, const sgr_camera *cam,
                   const sgr_render_params *p, float *accum_dev, uint64_t *job_id) {
  return sgr::guarded("sgr_job_submit",
                      [&] { return sgr_job_submit_impl(ctx, sc, cam, p, accum_dev, job_id); });
}

static int sgr_engine_step_impl(sgr_ctx *ctx, sgr_step_info *info) {
  Engine *E;
  int rc = engine_get(ctx, E);
  if (rc) return rc;
  if (info) std::memset(info, 0, sizeof *info);
  admit(E);
  // Select segments: round-robin over active jobs with chunks left.
  std::vector<uint64_t> run;
  uint64_t paths = 0;
  size_t K = 1;
  uint32_t maxp = 0;
  bool refill = false;  // any scene of the step deep enough for lane refill
  bool inst = false;    // any scene of the step with instances, ellipses or local primitives
  bool partial = false; // any scene of the step with partial occluders
  size_t n = E->active.size();
  size_t stop = n;
  for (size_t k = 0; k < n && run.size() < SGR_MAX_SEGMENTS; k++) {
    size_t idx = (E->cursor + k) % n;
    Job &j = E->jobs[E->active[idx]];
    if (j.next_s0 >= j.pl.R.spp) continue;
    uint32_t sb = std::min(j.sb, j.pl.R.spp - j.next_s0);
    uint64_t np = (uint64_t)j.pl.npix * sb;
    if (paths + np > E->cfg.path_capacity) {
      if (stop == n) stop = idx;
      continue;
    }
    run.push_back(j.id);
    paths += np;
    K = std::max<size_t>(K, j.pl.R.K);
    maxp = std::max(maxp, j.pl.R.maxp);
    refill = refill || j.sc->view.walk_refill;
    inst = inst || j.sc->view.n_insts > 0 || j.sc->view.has_ext;
    partial = partial || j.sc->view.has_partial;
  }
  if (run.empty()) return SGR_OK;
  // next step starts with the first job that did not fit, else the job
  // after the last one taken
  if (stop != n) {
    E->cursor = stop;
  } else {
    auto it = std::find(E->active.begin(), E->active.end(), run.back());
    E->cursor = (size_t)(it - E->active.begin() + 1) % n;
  }
  rc = ensure_engine_work(ctx, E, E->cfg.path_capacity, K);
  if (rc) return rc;
  StepTables &T = *E->host;
  uint32_t pb = 0, ob = 0;
  for (size_t i = 0; i < run.size(); i++) {
    Job &j = E->jobs[run[i]];
    RenderView R = j.pl.R;
    R.s0 = j.next_s0;
    R.sb = std::min(j.sb, R.spp - j.next_s0);
    R.npaths = j.pl.npix * R.sb;
    R.path_base = pb;
    R.out_base = ob;
    T.R[i] = R;
    T.seg_base[i] = pb;
    T.out_base[i] = ob;
    pb += R.npaths;
    ob += j.tiles_out;
  }
  T.seg_base[run.size()] = pb;
  T.out_base[run.size()] = ob;
  for (int s = 0; s < SGR_MAX_SCENE_SLOTS; s++)
    if (E->slot_scene[s]) T.S[s] = E->slot_scene[s]->view;
  T.hdr[0] = (uint32_t)run.size();
  T.hdr[1] = pb;
  T.hdr[2] = ob;
  T.hdr[3] = 0;
  const unsigned grid = E->grid(ctx);
  double t_begin = E->now_ms();
  HIPCHK(hipMemcpyAsync(E->dev, E->host, sizeof(StepTables), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipEventRecord(E->e0, ctx->stream));
  bool graphed = false;
  if (E->cfg.use_graph) {
    const uint32_t key =
        maxp | (refill ? 1u << 31 : 0u) | (inst ? 1u << 30 : 0u) | (partial ? 1u << 29 : 0u);
    auto it = E->graphs.find(key);
    if (it == E->graphs.end()) {
      hipGraph_t g;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      enqueue_step(ctx, E, maxp, refill, inst, partial, grid);
      hipError_t le = hipGetLastError();
      hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
      if (le != hipSuccess) return fail(SGR_ERR_HIP, std::string("step capture: ") + hipGetErrorString(le));
      if (ce != hipSuccess) return fail(SGR_ERR_HIP, std::string("step capture: ") + hipGetErrorString(ce));
      hipGraphExec_t x;
      hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) return fail(SGR_ERR_HIP, std::string("step graph: ") + hipGetErrorString(ie));
      it = E->graphs.emplace(key, x).first;
    }
    HIPCHK(hipGraphLaunch(it->second, ctx->stream));
    graphed = true;
  } else {
    enqueue_step(ctx, E, maxp, refill, inst, partial, grid);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(E->e1, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, E->e0, E->e1);
  double t_end = E->now_ms();
  uint32_t finished = 0;
  for (size_t i = 0; i < run.size(); i++) {
    Job &j = E->jobs[run[i]];
    if (j.t_start < 0) j.t_start = t_begin;
    const RenderView &R = T.R[i];
    uint32_t w = R.PW - 2 * R.fr, h = R.PH - 2 * R.fr;
    E->eye_rays += R.npaths;
    E->samples += (uint64_t)w * h * R.sb;
    j.next_s0 += R.sb;
    j.chunks_done++;
    j.steps++;
    if (j.next_s0 >= j.pl.R.spp) {
      j.state = SGR_JOB_DONE;
      j.t_done = t_end;
      finished++;
    }
  }
  // done jobs leave the rotation (they keep their slot until released)
  std::vector<uint64_t> keep;
  uint64_t cur_id = E->active.empty() ? 0 : E->active[E->cursor % E->active.size()];
  for (uint64_t id : E->active)
    if (E->jobs[id].state != SGR_JOB_DONE) keep.push_back(id);
  E->active.swap(keep);
  E->cursor = 0;
  for (size_t i = 0; i < E->active.size(); i++)
    if (E->active[i] >= cur_id) {
      E->cursor = i;
      break;
    }
  E->ms_device += ms;
  E->steps++;
  if (info) {
    info->jobs = (uint32_t)run.size();
    info->jobs_finished = finished;
    info->paths = pb;
    info->ms_device = ms;
    info->max_path_len = maxp;
    info->graph = graphed;
  }
  return SGR_OK;
}

int sgr_engine_step(sgr_ctx *ctx, sgr_step_info *info) {
  return sgr::guarded("sgr_engine_step",
                      [&] { return sgr_engine_step_impl(ctx, info); });
}

int sgr_job_poll(sgr_ctx *ctx, uint64_t id, sgr_job_status *s) {
  if (!ctx || !s) return fail(SGR_ERR_INVALID, "sgr_job_poll: null argument");
  Engine *E = ctx->engine;
  auto it = E ? E->jobs.find(id) : std::map<uint64_t, Job>::iterator();
  if (!E || it == E->jobs.end()) return fail(SGR_ERR_INVALID, "sgr_job_poll: no such job");
  const Job &j = it->second;
  s->state = j.state;
  s->chunks_done = j.chunks_done;
  s->chunks_total = j.chunks_total;
  s->steps = j.steps;
  s->t_submit_ms = j.t_submit;
  s->t_start_ms = j.t_start;
  s->t_done_ms = j.t_done;
  s->footprint = j.footprint;
  return SGR_OK;
}

int sgr_job_result(sgr_ctx *ctx, uint64_t id, float *dst, int dst_on_device) {
  if (!ctx || !dst) return fail(SGR_ERR_INVALID, "sgr_job_result: null argument");
  Engine *E = ctx->engine;
  auto it = E ? E->jobs.find(id) : std::map<uint64_t, Job>::iterator();
  if (!E || it == E->jobs.end()) return fail(SGR_ERR_INVALID, "sgr_job_result: no such job");
  const Job &j = it->second;
  if (j.state != SGR_JOB_DONE) return fail(SGR_ERR_NOTREADY, "sgr_job_result: job not finished");
  HIPCHK(hipSetDevice(ctx->device));
  size_t bytes = (size_t)j.pl.R.width * j.pl.R.height * 5 * sizeof(float);
  HIPCHK(hipMemcpyAsync(dst, j.accum, bytes,
                        dst_on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return SGR_OK;
}

int sgr_job_release(sgr_ctx *ctx, uint64_t id) {
  if (!ctx) return fail(SGR_ERR_INVALID, "sgr_job_release: null context");
  Engine *E = ctx->engine;
  auto it = E ? E->jobs.find(id) : std::map<uint64_t, Job>::iterator();
  if (!E || it == E->jobs.end()) return fail(SGR_ERR_INVALID, "sgr_job_release: no such job");
  Job &j = it->second;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  retire(E, j);
  if (j.owns_accum) (void)hipFree(j.accum);
  j.sc->jobs_using--;
  E->queued.erase(std::remove(E->queued.begin(), E->queued.end(), id), E->queued.end());
  auto a = std::find(E->active.begin(), E->active.end(), id);
  if (a != E->active.end()) {
    size_t pos = (size_t)(a - E->active.begin());
    E->active.erase(a);
    if (E->cursor > pos) E->cursor--;
    if (!E->active.empty()) E->cursor %= E->active.size();
    else E->cursor = 0;
  }
  E->jobs.erase(it);
  return SGR_OK;
}

int sgr_engine_stats(sgr_ctx *ctx, sgr_stats *stats) {
  if (!stats) return fail(SGR_ERR_INVALID, "sgr_engine_stats: null argument");
  Engine *E;
  int rc = engine_get(ctx, E);
  if (rc) return rc;
  unsigned long long st[SGR_NUM_STATS];
  HIPCHK(hipMemcpyAsync(st, E->stats, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::memset(stats, 0, sizeof *stats);
  stats->eye_rays = E->eye_rays;
  stats->paths_traced = E->eye_rays;
  stats->samples = E->samples;
  stats->scene_intersect_calls = st[SGR_STAT_ISECT];
  stats->shadow_tests = st[SGR_STAT_SHADOW];
  stats->node_visits_closest = st[SGR_STAT_NODES_C];
  stats->surf_tests_closest = st[SGR_STAT_TESTS_C];
  stats->node_visits_shadow = st[SGR_STAT_NODES_S];
  stats->surf_tests_shadow = st[SGR_STAT_TESTS_S];
  stats->illum_calls = st[SGR_STAT_ILLUM];
  stats->media_overflows = st[SGR_STAT_MEDIA_OVF];
  stats->partial_overflows = st[SGR_STAT_PARTIAL_OVF];
  stats->node_fetches_closest = st[SGR_STAT_FETCH_C];
  stats->node_fetches_shadow = st[SGR_STAT_FETCH_S];
  stats->ms_render = E->ms_device;
  return SGR_OK;
}

// IDS: the jobs submitted so far; the caller releases them on every exit,
// exceptions included.
static int sgr_render_batch_impl(sgr_ctx *ctx, uint32_t n, sgr_scene *const *scenes,
                                 const sgr_camera *cams, const sgr_render_params *params,
                                 float *const *accum, int on_device, sgr_stats *stats,
                                 std::vector<uint64_t> &ids) {
  if (n && (!scenes || !cams || !params || !accum))
    return fail(SGR_ERR_INVALID, "sgr_render_batch: null argument");
  Engine *E;
  int rc = engine_get(ctx, E);
  if (rc) return rc;
  if (!E->jobs.empty()) return fail(SGR_ERR_INVALID, "sgr_render_batch: engine holds other jobs");
  sgr_stats before;
  if (stats) {
    rc = sgr_engine_stats(ctx, &before);
    if (rc) return rc;
  }
  ids.reserve(n);  // push_back after a successful submit cannot throw
  auto cleanup = [&]() {
    for (uint64_t id : ids) sgr_job_release(ctx, id);
    ids.clear();
  };
  for (uint32_t i = 0; i < n; i++) {
    if (!accum[i]) {
      cleanup();
      return fail(SGR_ERR_INVALID, "sgr_render_batch: null accumulators for job " + std::to_string(i));
    }
    uint64_t id;
    rc = sgr_job_submit(ctx, scenes[i], &cams[i], &params[i], on_device ? accum[i] : nullptr, &id);
    if (rc) {
      std::string msg = g_err;
      cleanup();
      return fail(rc, msg);
    }
    ids.push_back(id);
  }
  for (;;) {
    sgr_step_info si;
    rc = sgr_engine_step(ctx, &si);
    if (rc) {
      std::string msg = g_err;
      cleanup();
      return fail(rc, msg);
    }
    if (si.jobs == 0) break;
  }
  for (uint32_t i = 0; i < n && !on_device; i++) {
    const sgr_render_params &p = params[i];
    std::vector<float> tmp((size_t)p.width * p.height * 5);
    rc = sgr_job_result(ctx, ids[i], tmp.data(), 0);
    if (rc) {
      std::string msg = g_err;
      cleanup();
      return fail(rc, msg);
    }
    for (size_t k = 0; k < tmp.size(); k++) accum[i][k] += tmp[k];
  }
  if (stats) {
    sgr_stats after;
    rc = sgr_engine_stats(ctx, &after);
    if (rc) {
      cleanup();
      return rc;
    }
    *stats = after;
    stats->eye_rays -= before.eye_rays;
    stats->paths_traced -= before.paths_traced;
    stats->samples -= before.samples;
    stats->scene_intersect_calls -= before.scene_intersect_calls;
    stats->shadow_tests -= before.shadow_tests;
    stats->node_visits_closest -= before.node_visits_closest;
    stats->surf_tests_closest -= before.surf_tests_closest;
    stats->node_visits_shadow -= before.node_visits_shadow;
    stats->surf_tests_shadow -= before.surf_tests_shadow;
    stats->illum_calls -= before.illum_calls;
    stats->media_overflows -= before.media_overflows;
    stats->node_fetches_closest -= before.node_fetches_closest;
    stats->node_fetches_shadow -= before.node_fetches_shadow;
    stats->ms_render -= before.ms_render;
  }
  cleanup();
  return SGR_OK;
}

int sgr_render_batch(sgr_ctx *ctx, uint32_t n, sgr_scene *const *scenes, const sgr_camera *cams,
                     const sgr_render_params *params, float *const *accum, int on_device,
                     sgr_stats *stats) {
  std::vector<uint64_t> ids;
  return sgr::guarded(
      "sgr_render_batch",
      [&] {
        return sgr_render_batch_impl(ctx, n, scenes, cams, params, accum, on_device, stats, ids);
      },
      [&] {
        for (uint64_t id : ids) sgr_job_release(ctx, id);
        ids.clear();
      });
}

}  // extern "C"
