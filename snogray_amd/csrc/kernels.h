// Render-time buffers and kernel launchers shared by the host library and
// the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "dev_common.h"
#include "guide.h"

namespace sgr {

// Media stack levels per path.  The reference's stack is unbounded
// (material/media.cc:75-118); here a transmission into a ninth nested
// medium is not pushed (the path keeps refracting against the eighth)
// and is counted in SGR_STAT_MEDIA_OVF.  The oracle applies the same cap.
#define SGR_MEDIA_MAX 8
#define SGR_MAX_SLOTS 64      // shadow-ray slots per path (2 per light sample)
#define SGR_COMPACT_ITEMS 4096  // queue positions per compaction block (256 x 16)

// Path state word.
#define SGR_ST_LEN_MASK   0xffu
#define SGR_ST_SPEC_SHIFT 8
#define SGR_ST_DEPTH_SHIFT 12
#define SGR_ST_PENDING    (1u << 16)
#define SGR_ST_DONE       (1u << 17)
#define SGR_ST_STRAT      (1u << 18)

// Counter block: ray-queue lengths per bounce, shadow-queue lengths, then
// the fetch heads of the lane-refill traversal (per bounce, one per XCD
// slice of the queue, SGR_HEAD_STRIDE words apart) for closest-hit and
// any-hit rays.
#define SGR_MAX_BOUNCES   64
#define SGR_CNT_Q         0
#define SGR_CNT_S         (SGR_MAX_BOUNCES + 2)
#define SGR_CNT_HE        160  // a multiple of SGR_HEAD_STRIDE past the lengths
#define SGR_HEADS_PER_B   (8 * SGR_HEAD_STRIDE)
#define SGR_CNT_HS        (SGR_CNT_HE + SGR_HEADS_PER_B * SGR_MAX_BOUNCES)
#define SGR_NUM_COUNTERS  (SGR_CNT_HS + SGR_HEADS_PER_B * SGR_MAX_BOUNCES)
static_assert(SGR_CNT_HE >= 2 * SGR_MAX_BOUNCES + 4 && SGR_CNT_HE % SGR_HEAD_STRIDE == 0,
              "counter block layout");

// Device statistics slots (unsigned long long).
#define SGR_STAT_NODES_C 0
#define SGR_STAT_TESTS_C 1
#define SGR_STAT_NODES_S 2
#define SGR_STAT_TESTS_S 3
#define SGR_STAT_ILLUM   4
#define SGR_STAT_ISECT   5
#define SGR_STAT_SHADOW  6
#define SGR_STAT_MEDIA_OVF 7  // media pushes dropped at SGR_MEDIA_MAX
#define SGR_STAT_FETCH_C 8    // octree node records loaded (closest hit)
#define SGR_STAT_FETCH_S 9    // ... (any hit)
#define SGR_STAT_DINT_OVF 10  // direct integrator: branches beyond SGR_DIRECT_MAX_DEPTH
#define SGR_STAT_PARTIAL_OVF 11  // partial occluders beyond a shadow ray's memory
#define SGR_NUM_STATS    16
// The statistics buffer also holds the 8 fetch heads of a ray-batch query,
// from byte SGR_STATS_HEADS_OFF on, SGR_HEAD_STRIDE words apart.
#define SGR_STATS_HEADS_OFF 128
#define SGR_STATS_BYTES  (SGR_STATS_HEADS_OFF + 8 * SGR_HEAD_STRIDE * 4)
static_assert(SGR_NUM_STATS * 8 <= SGR_STATS_HEADS_OFF, "statistics layout");

// One photon map in kd-tree heap order (children of node i at 2i + 1,
// 2i + 2): position and split axis (bits of w), power, the direction the
// photon came from.
struct PhotonSetView {
  const float4 *pos;
  const float4 *pow;
  const float4 *dir;
  uint32_t n;
  float scale;  // 1 / paths of the set (0: the set contributes nothing)
};

// The photon integrator's maps and parameters (render/photon-integ.cc:
// 36-148): photons per estimate, search radius squared, final-gathering
// sample counts and their channels' grids.
struct PhotonView {
  PhotonSetView set[3];  // direct, caustic, indirect
  uint32_t k;
  float r2;
  uint32_t fg, fg_bsdf, fg_photon;
  uint32_t direct_illum, caustics;
  Grid g_fgp, g_fgb;
  uint32_t nsub_fgp, nsub_fgb;
};

// A batch of photon paths [base, base + n): per path, rec = (first
// deposit, deposit attempts | valid << 31); deposits are 3 float4 each
// (position + set in w's bits, power, the direction the photon came
// from), reserved in order of completion up to CAP.
struct PhotonShootView {
  uint64_t seed;
  uint32_t base, n;
  uint32_t caustic_on;   // a caustic set is wanted (else those photons are indirect)
  uint2 *rec;
  float4 *dep;
  uint32_t cap;
  uint32_t *counter;
};

struct RayBatchDev {
  const float *ox, *oy, *oz, *dx, *dy, *dz, *t0, *t1;
  uint32_t *prim;
  float *t, *u, *v;
  uint32_t *inst;   // may be null
  uint8_t *occ;
};

struct RenderView {
  uint32_t width, height, spp, minp, maxp, ds, L;
  uint32_t nsub_strat, nsub_rand, K;
  Grid g_cam, g_dir, g_cont;
  // Surface integrator (SGR_INTEG_*); the direct integrator's light
  // samples per light and per sample, and their grid (spp x samples).
  uint32_t integ, nsub_dint;
  Grid g_dint;
  PhotonView ph;  // SGR_INTEG_PHOTON
  uint64_t seed;
  float bg_alpha, fb, fc, inv_radius, neg_clamp;
  int fr;
  int X0, Y0, PW, PH;     // pattern window including the filter margin
  uint32_t s0, sb;        // this chunk renders samples [s0, s0 + sb)
  uint32_t npaths;        // PW * PH * sb
  float cam_pos[3], fwd[3], up[3], right[3];
  float tanx, tany, max_trace;
  // Multi-job steps: scene table slot, first path of this segment in the
  // step, first splat tile of this segment in the step, and the job's
  // accumulators (device).
  uint32_t scene, path_base, out_base;
  float *accum;
};

// One engine step over several jobs: per-segment render views, per-slot
// scene views, and prefix offsets of the segments' paths and splat tiles.
#define SGR_MAX_SEGMENTS 256
struct StepView {
  const RenderView *R;
  const SceneView *S;
  const uint32_t *seg_base;  // nseg + 1 path offsets
  const uint32_t *out_base;  // nseg + 1 splat-tile offsets
  const uint32_t *hdr;       // segments, paths, splat tiles (device: a
                             // captured step graph replays with new tables)
};

struct PathBufs {
  float *ray[8];      // ox oy oz dx dy dz t0 t1
  uint32_t *hprim;    // closest hit of queue position j of the current bounce
  float *ht, *hu, *hv;
  uint32_t *hinst;    // ... its instance (instanced scenes)
  float *T[3], *rad[3], *Tp[3];
  float *rr, *alpha;
  uint32_t *state;
  float *media;       // SGR_MEDIA_MAX x npaths
  float *sray[8];     // shadow rays, K per path
  float *sc[3];       // shadow contributions
  uint8_t *svis;
  // Scenes with partial occluders (partial != 0): the medium index at
  // each shadow ray's origin and the transmittance it found.
  float *smed, *strans[3];
  uint32_t partial = 0;
  uint32_t *q[2];     // ray queues (ping-pong), in raygen's pixel order
  uint32_t *sq;       // shadow queue, grouped by direction octant per block
  // Ordered compaction of the queues k_shade produces: per queue
  // position, whether the path goes on (1 or 0) and which of its shadow
  // slots got a ray (1 | direction octant << 1, or 0); per block of
  // SGR_COMPACT_ITEMS positions, the two counts (then their exclusive
  // prefixes).
  uint8_t *qflag, *sflag;
  uint32_t *cblk;
  uint32_t *counts;   // SGR_NUM_COUNTERS
  float *out;         // 6 floats per path: sx sy r g b a
  uint32_t *pseg;     // segment of each path (multi-job steps)
  uint32_t n;         // path capacity (stride of the media stack)
  uint32_t K;         // shadow slots per path
  // From bounce fin_from on, a ray queue shorter than fin_max rays is
  // finished by k_finish in one launch instead of bounce by bounce
  // (fin_max 0: never).
  uint32_t fin_from = 0, fin_max = 0;
  // Statistics slot counting dropped media pushes (set by the launchers).
  unsigned long long *media_ovf = nullptr;
};

// Bounce from which, and queue length below which, renders and engine
// steps finish a queue's paths in one launch (k_finish).
#define SGR_FINISH_FROM 1u
#define SGR_FINISH_MAX (1u << 18)

// Sets the CU count the per-kernel launch caps are derived from.
void set_launch_cus(unsigned num_cu);
unsigned launch_cus();
void launch_trace(const SceneView &S, const RayBatchDev &rb, uint32_t n, bool any,
                  unsigned long long *stats, hipStream_t st);
void launch_raygen(bool multi, const RenderView &R, const StepView &J, const PathBufs &P,
                   unsigned grid, hipStream_t st);
void launch_extend(bool multi, const SceneView &S, const StepView &J, const PathBufs &P,
                   uint32_t b, unsigned long long *stats, unsigned grid, hipStream_t st);
// G: the learned proposal's view, or NULL for plain BSDF sampling
// (single-job renders only).
void launch_direct_integ(const SceneView &S, const RenderView &R, const PathBufs &P,
                         unsigned long long *stats, unsigned grid, hipStream_t st);
void launch_photon_shoot(const SceneView &S, const PhotonShootView &V, unsigned long long *stats,
                         unsigned grid, hipStream_t st);
void launch_shade(bool multi, const SceneView &S, const RenderView &R, const StepView &J,
                  const PathBufs &P, uint32_t b, unsigned long long *stats, unsigned grid,
                  hipStream_t st, const GuideView *G = nullptr);
void launch_occlude(bool multi, const SceneView &S, const StepView &J, const PathBufs &P,
                    uint32_t b, unsigned long long *stats, unsigned grid, hipStream_t st);
void launch_gather(bool multi, const RenderView &R, const StepView &J, const PathBufs &P,
                   uint32_t b, unsigned grid, hipStream_t st, const GuideView *G = nullptr);
// Finishes the paths of the bounce-B queue when P.fin_* select it
// (unguided renders and multi-job steps).
void launch_finish(bool multi, const SceneView &S, const RenderView &R, const StepView &J,
                   const PathBufs &P, uint32_t b, unsigned long long *stats, unsigned grid,
                   hipStream_t st);
void launch_splat(bool multi, const RenderView &R, const StepView &J, const float *out,
                  unsigned grid, hipStream_t st);
void launch_count_queues(uint32_t maxp, const PathBufs &P, unsigned long long *stats,
                         hipStream_t st);

}  // namespace sgr
