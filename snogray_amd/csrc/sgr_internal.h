// Library-internal state shared by the C-ABI sources (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/snoggpu.h"
#include "kernels.h"
#include "octree_build.h"

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return sgr::fail(SGR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace sgr {
struct Engine;
int fail(int code, const std::string &msg);

// Runs F (returning an sgr status); a C++ exception thrown inside turns
// into an error code after CLEANUP runs, so none crosses the C ABI.
template<class F, class C>
int guarded(const char *what, F &&f, C &&cleanup) {
  try {
    return f();
  } catch (const std::bad_alloc &) {
    cleanup();
    return fail(SGR_ERR_NOMEM, std::string(what) + ": out of host memory");
  } catch (const std::length_error &e) {  // a structure outgrew its index range
    cleanup();
    return fail(SGR_ERR_LIMIT, std::string(what) + ": " + e.what());
  } catch (const std::exception &e) {
    cleanup();
    return fail(SGR_ERR_INVALID, std::string(what) + ": " + e.what());
  }
}
template<class F>
int guarded(const char *what, F &&f) {
  return guarded(what, static_cast<F &&>(f), [] {});
}
}

// Chunks sgr_render keeps in flight at most (one stream and one set of
// path buffers each).
#define SGR_CHUNK_LANES 2

struct sgr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t lane_stream[SGR_CHUNK_LANES - 1] = {};  // chunk lanes of sgr_render after the first
  std::atomic<uint64_t> bytes{0};  // device bytes held (scenes + buffers)
  void *work = nullptr;        // path buffers, grown on demand
  size_t work_bytes = 0;
  unsigned long long *stats = nullptr;
  int num_cu = 256;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> evpool;  // per-launch timing events
  std::vector<hipEvent_t> syncpool;  // ordering events between the chunk lanes
  std::atomic<uint32_t> live_scenes{0};
  sgr::Engine *engine = nullptr;   // multi-job engine (sgr_jobs.cpp)
  // Device address of this device's copy of the no-alloc render scratch
  // (sgr_render_noalloc), or null when it could not be resolved.
  void *noalloc_scratch = nullptr;
};

struct sgr_scene {
  sgr_ctx *ctx = nullptr;
  sgr::SceneView view{};
  std::vector<void *> allocs;
  uint64_t bytes = 0;
  sgr::FlatOctree oct;
  float bbox_lo[3], bbox_hi[3];
  uint32_t n_lights = 0;
  uint32_t jobs_using = 0;         // jobs of the engine referring to it
  uint32_t photon_maps = 0;        // photon maps built for it (sgr_photon.cpp)
};

namespace sgr {

extern thread_local std::string g_err;
int fail(int code, const std::string &msg);

// Validated render set-up of one job (window, margin, sample layout).
struct RenderPlan {
  RenderView R;
  uint32_t flags;
  uint32_t sb_max;
  uint32_t npix;   // pattern pixels including the filter margin
};
int plan_render(sgr_scene *sc, const sgr_camera *cam, const sgr_render_params *p, RenderPlan &pl);
// Grid shape of a channel with n samples (render/grid.cc:26-38).
Grid grid_shape(uint32_t n);
// The photon integrator's view of MAP for a render of SPP samples per
// pixel; DIRECT_SAMPLES gets the map's direct-lighting samples per light.
int photon_render_view(const sgr_scene *sc, const sgr_photon_map *map, uint32_t spp,
                       PhotonView &out, uint32_t &direct_samples);
// Path buffers for CAP paths with K shadow slots, carved from BASE
// (NULL: only the size is computed).
size_t path_bytes(size_t cap, size_t K, PathBufs *P, char *base);
int ensure_work(sgr_ctx *ctx, size_t bytes);
void engine_destroy(sgr_ctx *ctx);

// Learned proposal hooks of a render (sgr_guide.cpp): buffers for CAP
// paths, clearing a chunk's training records, inference before shading
// bounce B, training after a chunk, final statistics.
// LANE: chunk lane (0 or 1) whose inference rows and records V uses.
int guide_prepare(sgr_guide *g, sgr_scene *sc, size_t cap, int lane, GuideView &V);
int guide_begin_chunk(sgr_guide *g, GuideView &V, uint32_t npaths, uint64_t key, hipStream_t s);
void guide_infer(sgr_guide *g, const GuideView &V, const PathBufs &P, uint32_t b, unsigned grid,
                 hipStream_t s);
int guide_train_chunk(sgr_guide *g, const GuideView &V, uint32_t nrec);
int guide_finish(sgr_guide *g);

}  // namespace sgr
