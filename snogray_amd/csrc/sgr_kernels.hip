// gfx950 kernels of the path-tracing core.
//
// Ray batches: k_trace<ANY> (Space::intersect / Space::occludes,
// space/space.cc:55-66, :182-194).
//
// Wavefront path tracer (PathInteg::Li, render/path-integ.cc:133-407,
// restructured): one path per (pattern pixel, sample) of a chunk of
// samples; per bounce
//   k_extend   closest hit of every ray in the ray queue
//   k_shade    emission, direct lighting (DirectIllum::sample_light,
//              render/direct-illum.cc:126-255: one light sample and one
//              BSDF sample per light, power-heuristic MIS), BSDF
//              continuation, Russian roulette; flags the shadow rays it
//              wrote and whether the path goes on
//   k_compact_* ordered compaction of those flags into the shadow queue
//              and the next ray queue (block scan + wave ballots)
//   k_occlude  any hit of every queued shadow ray
//   k_gather   adds the unoccluded direct terms in the reference's order,
//              writes finished paths
// with k_finish taking over once a late queue is short; then k_splat
// gathers the chunk's samples into the image accumulators with the
// Mitchell filter (image/image-filter-conv.h:109-186).
#include <cstring>

#include "dev_shade.h"
#include "dev_photon.h"
#include "kernels.h"
#include "../../include/snoggpu.h"
#include "dev_guide.h"

namespace sgr {

// ---------------------------------------------------------------------------
// Ray batches
// ---------------------------------------------------------------------------
// Queue range of this block.  Hardware deals workgroups to the 8 XCDs
// round-robin; remapping so that the blocks of one XCD take one
// contiguous slice of the queue keeps neighbouring rays (neighbouring
// pixels) on the same L2.  The grid is a multiple of 8.
__device__ __forceinline__ void block_range(uint32_t n, uint32_t &beg, uint32_t &end) {
  const uint32_t G = gridDim.x;
  const uint32_t lb = (blockIdx.x & 7u) * (G >> 3) + (blockIdx.x >> 3);
  const uint32_t per = (n + G - 1) / G;
  beg = lb * per;
  end = beg + per < n ? beg + per : n;
}

// Walk the N rays of a launch: lane refill over XCD slices, or one ray
// per lane over this block's queue range.
// INST: scenes may hold instances (plain walk only; the walk inside a
// model takes a second LDS stack, declared here only for such kernels).
template<bool ANY, bool MULTI, bool REFILL, bool INST, typename Get, typename Put>
__device__ __forceinline__ void walk(const SceneView &S1, uint32_t n, uint32_t *heads, Get get,
                                     Put put, TraceCount &cnt, LdsStack L) {
  static_assert(!(REFILL && INST), "instanced scenes take the plain walk");
  if constexpr (REFILL) {
    trace_refill<ANY, MULTI>(S1, n, heads, get, put, cnt, L);
  } else {
    uint32_t jb, je;
    block_range(n, jb, je);
    if constexpr (INST) {
      SGR_STACK_LDS(stk2);
      uint2 spill2[SGR_MAX_DEPTH];
      trace_plain<ANY, MULTI, true>(S1, jb, je, get, put, cnt, L, Nest{stk2, spill2});
    } else {
      trace_plain<ANY, MULTI, false>(S1, jb, je, get, put, cnt, L, Nest{});
    }
  }
}

// Ray-batch queries with lane-refill traversal (heads[SGR_HEAD_STRIDE * i],
// i < 8, zeroed by the caller, are the fetch heads of the 8 slices of the
// batch).
template<bool ANY, bool REFILL, bool INST>
__global__ void __launch_bounds__(256) k_trace(SceneView S, RayBatchDev rb, uint32_t n,
                                               unsigned long long *stats) {
  SGR_STACK_LDS(stk);
  TraceCount cnt{0, 0, 0, 0};
  uint32_t *heads = (uint32_t *)((char *)stats + SGR_STATS_HEADS_OFF);
  walk<ANY, false, REFILL, INST>(
      S, n, heads,
      [&](uint32_t i, TravState &T) -> const SceneView * {
        T.ray.o = v3(rb.ox[i], rb.oy[i], rb.oz[i]);
        T.ray.d = v3(rb.dx[i], rb.dy[i], rb.dz[i]);
        T.ray.t0 = rb.t0[i];
        T.ray.t1 = rb.t1[i];
        return nullptr;
      },
      [&](uint32_t i, const TravState &T) {
        if (ANY) {
          rb.occ[i] = T.hit == SGR_MISS ? 0 : 1;
        } else {
          rb.prim[i] = T.hit;
          rb.t[i] = T.hit == SGR_MISS ? 0.0f : T.ray.t1;
          rb.u[i] = T.hu;
          rb.v[i] = T.hv;
          if (rb.inst) rb.inst[i] = T.hinst;
        }
      },
      cnt, stk);
  wave_count(&stats[0], cnt.nodes);
  wave_count(&stats[1], cnt.tests);
  wave_count(&stats[SGR_STAT_FETCH_C], cnt.fetch);
}

// ---------------------------------------------------------------------------
// Wavefront path tracer
// ---------------------------------------------------------------------------
// Direction octant of a ray (sign bits x, y, z as 4, 2, 1), the key the
// shadow queue is grouped by.
__device__ __forceinline__ uint32_t dir_octant(V3 d) {
  return (d.x >= 0.0f ? 0u : 4u) | (d.y >= 0.0f ? 0u : 2u) | (d.z >= 0.0f ? 0u : 1u);
}

__device__ __forceinline__ void path_ident(const RenderView &R, uint32_t p, int &px, int &py,
                                           uint32_t &s) {
  uint32_t i = p - R.path_base;
  uint32_t pix = i / R.sb;
  s = R.s0 + (i - pix * R.sb);
  px = R.X0 + (int)(pix % (uint32_t)R.PW);
  py = R.Y0 + (int)(pix / (uint32_t)R.PW);
}

__device__ __forceinline__ void load_ray(const PathBufs &P, uint32_t p, RayF &r) {
  r.o = v3(P.ray[0][p], P.ray[1][p], P.ray[2][p]);
  r.d = v3(P.ray[3][p], P.ray[4][p], P.ray[5][p]);
  r.t0 = P.ray[6][p];
  r.t1 = P.ray[7][p];
}
// Transmittance of shadow slot SLOT (scenes with partial occluders).
__device__ __forceinline__ void put_trans(const PathBufs &P, size_t slot, C3 t) {
  P.strans[0][slot] = t.r; P.strans[1][slot] = t.g; P.strans[2][slot] = t.b;
}
__device__ __forceinline__ C3 get_trans(const PathBufs &P, size_t slot) {
  return c3(P.strans[0][slot], P.strans[1][slot], P.strans[2][slot]);
}
__device__ __forceinline__ void store_ray(const PathBufs &P, uint32_t p, const RayF &r) {
  P.ray[0][p] = r.o.x; P.ray[1][p] = r.o.y; P.ray[2][p] = r.o.z;
  P.ray[3][p] = r.d.x; P.ray[4][p] = r.d.y; P.ray[5][p] = r.d.z;
  P.ray[6][p] = r.t0; P.ray[7][p] = r.t1;
}

// Index of the segment holding element i: largest k with base[k] <= i.
__device__ __forceinline__ uint32_t find_segment(const uint32_t *base, uint32_t nseg, uint32_t i) {
  uint32_t lo = 0, hi = nseg;  // base[lo] <= i < base[hi]
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (base[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Camera ray (render-mgr/renderer.cc:71-85, camera/camera.h:322-389) and
// path-state reset.  Thread 0 also resets the queue counters.
template<bool MULTI>
__global__ void __launch_bounds__(256) k_raygen(RenderView R1, StepView J, PathBufs P) {
  uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = MULTI ? J.hdr[1] : R1.npaths;
  for (uint32_t c = gid; c < SGR_NUM_COUNTERS; c += gridDim.x * blockDim.x)
    P.counts[c] = c == 0 ? total : 0u;
  for (uint32_t i = gid; i < total; i += gridDim.x * blockDim.x) {
    uint32_t seg = 0;
    if (MULTI) {
      seg = find_segment(J.seg_base, J.hdr[0], i);
      P.pseg[i] = seg;
    }
    const RenderView R = MULTI ? J.R[seg] : R1;
    int px, py;
    uint32_t s;
    path_ident(R, i, px, py, s);
    uint64_t pk = pixel_key(R.seed, px, py);
    float cu, cv;
    strat_uv(pk, 0, s, R.g_cam, cu, cv);
    float cx = (float)px + cu, cy = (float)py + cv;
    float fu = cx / (float)R.width, fv = ((float)R.height - cy) / (float)R.height;
    float x = 2.0f * fu - 1.0f, y = 2.0f * fv - 1.0f;
    V3 fwd = v3(R.fwd[0], R.fwd[1], R.fwd[2]);
    V3 right = v3(R.right[0], R.right[1], R.right[2]);
    V3 up = v3(R.up[0], R.up[1], R.up[2]);
    RayF r;
    r.o = v3(R.cam_pos[0], R.cam_pos[1], R.cam_pos[2]);
    r.d = (fwd + (right * x) * R.tanx) + (up * y) * R.tany;
    r.t0 = 0.0f;
    r.t1 = R.max_trace;
    store_ray(P, i, r);
    P.T[0][i] = 1.0f; P.T[1][i] = 1.0f; P.T[2][i] = 1.0f;
    P.rad[0][i] = 0.0f; P.rad[1][i] = 0.0f; P.rad[2][i] = 0.0f;
    P.rr[i] = 1.0f;
    P.alpha[i] = 1.0f;
    P.state[i] = 1u << SGR_ST_DEPTH_SHIFT;  // path_len 0, media depth 1
    P.media[i] = 1.0f;                      // default medium, ior 1
    P.q[0][i] = i;
    P.out[6 * (size_t)i + 0] = cx;
    P.out[6 * (size_t)i + 1] = cy;
  }
}

// Bounce B's queue of N rays goes to k_finish instead of the per-bounce
// kernels.
__device__ __forceinline__ bool finishing(const PathBufs &P, uint32_t b, uint32_t n) {
  return P.fin_max != 0 && b >= P.fin_from && n < P.fin_max;
}

template<bool MULTI, bool REFILL, bool INST>
__global__ void __launch_bounds__(256) k_extend(SceneView S1, StepView J, PathBufs P, uint32_t b,
                                                unsigned long long *stats) {
  SGR_STACK_LDS(stk);
  TraceCount cnt{0, 0, 0, 0};
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n)) return;
  const uint32_t *q = P.q[b & 1];
  walk<false, MULTI, REFILL, INST>(
      S1, n, &P.counts[SGR_CNT_HE + SGR_HEADS_PER_B * b],
      [&](uint32_t i, TravState &T) -> const SceneView * {
        const uint32_t p = q[i];
        load_ray(P, p, T.ray);
        return MULTI ? &J.S[J.R[P.pseg[p]].scene] : nullptr;
      },
      [&](uint32_t i, const TravState &T) {
        // by queue position: the lanes of a wave store neighbouring words
        // (by path the stores scatter over the sparse late queues)
        P.hprim[i] = T.hit;
        P.ht[i] = T.ray.t1;
        if (MULTI || S1.tri_vn) {  // read only to interpolate vertex normals
          P.hu[i] = T.hu;
          P.hv[i] = T.hv;
        }
        if (INST) P.hinst[i] = T.hinst;
      },
      cnt, stk);
  wave_count(&stats[SGR_STAT_NODES_C], cnt.nodes);
  wave_count(&stats[SGR_STAT_TESTS_C], cnt.tests);
  wave_count(&stats[SGR_STAT_FETCH_C], cnt.fetch);
}

template<bool MULTI, bool REFILL, bool INST>
__global__ void __launch_bounds__(256) k_occlude(SceneView S1, StepView J, PathBufs P, uint32_t b,
                                                 unsigned long long *stats) {
  SGR_STACK_LDS(stk);
  TraceCount cnt{0, 0, 0, 0};
  if (finishing(P, b, P.counts[SGR_CNT_Q + b])) return;
  const uint32_t n = P.counts[SGR_CNT_S + b];
  walk<true, MULTI, REFILL, INST>(
      S1, n, &P.counts[SGR_CNT_HS + SGR_HEADS_PER_B * b],
      [&](uint32_t i, TravState &T) -> const SceneView * {
        const uint32_t slot = P.sq[i];
        T.ray.o = v3(P.sray[0][slot], P.sray[1][slot], P.sray[2][slot]);
        T.ray.d = v3(P.sray[3][slot], P.sray[4][slot], P.sray[5][slot]);
        T.ray.t0 = P.sray[6][slot];
        T.ray.t1 = P.sray[7][slot];
        if (INST && P.partial) T.medium = P.smed[slot];
        return MULTI ? &J.S[J.R[P.pseg[slot / P.K]].scene] : nullptr;
      },
      [&](uint32_t i, const TravState &T) {
        const uint32_t slot = P.sq[i];
        P.svis[slot] = T.hit == SGR_MISS ? 1 : 0;
        if (INST && P.partial) put_trans(P, slot, T.trans);
      },
      cnt, stk);
  wave_count(&stats[SGR_STAT_NODES_S], cnt.nodes);
  wave_count(&stats[SGR_STAT_TESTS_S], cnt.tests);
  wave_count(&stats[SGR_STAT_FETCH_S], cnt.fetch);
  if (INST) wave_count(&stats[SGR_STAT_PARTIAL_OVF], cnt.povf);
}

// One path's vertex (render/path-integ.cc:133-407 for one bounce): the
// hit of path P at queue position J (P.hprim / P.ht [J]), emission,
// direct lighting (shadow rays written to the path's slots, each
// announced by EMIT(slot, octant of its direction)), the BSDF
// continuation and Russian roulette.  Returns whether the path goes on
// with the ray it stored.  J also indexes the learned proposal's lobes;
// ILLUM counts direct-lighting calls.
// Shadow slot SLOT of the path state: the ray, its contribution and (in
// scenes with partial occluders) the index of the medium it starts in.
__device__ __forceinline__ void put_shadow(const PathBufs &P, size_t slot, C3 c, const RayF &r,
                                           float medium) {
  if (P.partial) P.smed[slot] = medium;
  P.sray[0][slot] = r.o.x; P.sray[1][slot] = r.o.y; P.sray[2][slot] = r.o.z;
  P.sray[3][slot] = r.d.x; P.sray[4][slot] = r.d.y; P.sray[5][slot] = r.d.z;
  P.sray[6][slot] = r.t0; P.sray[7][slot] = r.t1;
  P.sc[0][slot] = c.r; P.sc[1][slot] = c.g; P.sc[2][slot] = c.b;
}

// EXT: the scenes may hold ellipses and local primitives.
template<bool MULTI, bool GUIDED, bool EXT, typename Emit>
__device__ __forceinline__ bool shade_path(const SceneView &S1, const RenderView &R1,
                                           const StepView &J, const PathBufs &P,
                                           const GuideView &G, uint32_t p, uint32_t j,
                                           uint32_t &illum, Emit emit) {
  const uint32_t N = P.n;
  const RenderView R = MULTI ? J.R[P.pseg[p]] : R1;
  const SceneView S = MULTI ? J.S[R.scene] : S1;
  uint32_t st = P.state[p];
  uint32_t path_len = st & SGR_ST_LEN_MASK;
  bool after_spec = (st >> SGR_ST_SPEC_SHIFT) & 1u;
  uint32_t depth = (st >> SGR_ST_DEPTH_SHIFT) & 0xfu;
  bool include = path_len == 0 || after_spec || R.ds == 0;
  C3 T = c3(P.T[0][p], P.T[1][p], P.T[2][p]);
  C3 rad = c3(P.rad[0][p], P.rad[1][p], P.rad[2][p]);
  float alpha = P.alpha[p];
  RayF ray;
  load_ray(P, p, ray);
  uint32_t prim = P.hprim[j];
  bool done = false, cont = false;
  uint32_t nst = st & ~(SGR_ST_PENDING | SGR_ST_DONE | SGR_ST_STRAT);
  float *grec = nullptr;  // this path's training record (learned proposal)
  if (GUIDED && p % G.rec_stride == 0) {
    // slots are shuffled so that every training minibatch spans the image
    grec = G.rec + (size_t)permute(p / G.rec_stride, G.rec_count, G.rec_key) * GD_REC;
    if (grec[9] == -1.0f) {
      // radiance gathered before the recorded continuation's first hit
      grec[13] = rad.r; grec[14] = rad.g; grec[15] = rad.b;
      grec[9] = -2.0f;
    }
  }
  if (prim == SGR_MISS) {
    if (include) rad = rad + background(S, ray.d) * T;
    if (path_len == 0 && czero(rad)) alpha = R.bg_alpha;
    done = true;
  } else {
    ray.t1 = P.ht[j];
    Isec is;
    float mtop = P.media[(size_t)(depth - 1) * N + p];
    float mencl = depth >= 2 ? P.media[(size_t)(depth - 2) * N + p] : 1.0f;
    // barycentrics only matter for interpolated vertex normals
    const float hu = S.tri_vn ? P.hu[j] : 0.0f, hv = S.tri_vn ? P.hv[j] : 0.0f;
    const uint32_t hinst = S.n_insts ? P.hinst[j] : SGR_MISS;
    make_isec<EXT>(S, prim, hinst, ray, hu, hv, mtop, mencl, is);
    const DevMaterial &HM = S.mats[is.mat];
    if (include && HM.type == 5 && !is.back) rad = rad + cload(HM.color) * T;
    if (is.bmat < 0) {
      done = true;
    } else {
      const DevMaterial &M = S.mats[is.bmat];
      int px, py;
      uint32_t s;
      path_ident(R, p, px, py, s);
      uint64_t pk = pixel_key(R.seed, px, py);
      uint64_t qk = path_key(pk, s);
      // ---- direct lighting (DirectIllum::sample_all_lights) ----
      if (R.ds != 0) {
        illum++;
        bool strat = path_len < R.minp;
        uint32_t nsub = strat ? R.nsub_strat : R.nsub_rand;
        for (uint32_t l = 0; l < R.L; l++) {
          const DevLight &L = S.lights[l];
          for (uint32_t jj = 0; jj < nsub; jj++) {
            float lu, lv, bu, bv;
            uint32_t base = 16 + (l * nsub + jj) * 8;
            if (strat) {
              uint32_t chan = 1 + path_len * (2 * R.L + 1) + 2 * l;
              strat_uv(pk, chan, s * nsub + jj, R.g_dir, lu, lv);
              strat_uv(pk, chan + 1, s * nsub + jj, R.g_dir, bu, bv);
            } else {
              lu = draw(qk, (path_len << 20) | (base + 0));
              lv = draw(qk, (path_len << 20) | (base + 1));
              bu = draw(qk, (path_len << 20) | (base + 2));
              bv = draw(qk, (path_len << 20) | (base + 3));
            }
            float extra = draw(qk, (path_len << 20) | (base + 4));
            size_t slot_a = (size_t)p * P.K + 2 * (l * nsub + jj), slot_b = slot_a + 1;
            P.svis[slot_a] = 0;
            P.svis[slot_b] = 0;
            // light sample, then BSDF sample (render/direct-illum.cc:147-252)
            direct_light_term<EXT>(S, L, is, M, lu, lv, extra, [&](C3 c, const RayF &r) {
              put_shadow(P, slot_a, c, r, is.media_ior);
              emit(slot_a, dir_octant(r.d));
            });
            direct_bsdf_term<EXT>(S, L, is, M, bu, bv, [&](C3 c, const RayF &r) {
              put_shadow(P, slot_b, c, r, is.media_ior);
              emit(slot_b, dir_octant(r.d));
            });
          }
        }
        nst |= SGR_ST_PENDING | (strat ? SGR_ST_STRAT : 0u);
        P.Tp[0][p] = T.r; P.Tp[1][p] = T.g; P.Tp[2][p] = T.b;
      }
      // ---- continuation (render/path-integ.cc:308-401) ----
      float cu, cv;
      if (path_len < R.minp) {
        uint32_t chan = 1 + path_len * (2 * R.L + 1) + 2 * R.L;
        strat_uv(pk, chan, s, R.g_cont, cu, cv);
      } else {
        cu = draw(qk, (path_len << 20) | 1u);
        cv = draw(qk, (path_len << 20) | 2u);
      }
      BSample bs;
      // Learned proposal: one-sample MIS of the BSDF (probability
      // G.mix) and the vMF mixture at this vertex; delta BSDFs (glass,
      // mirror) bypass it.
      const bool guided = GUIDED && (M.type == 1 || M.type == 2 || M.type == 6);  // not delta
      if (guided) {
        const float *L = G.lobes + (size_t)j * GD_OUT;
        const float usel = draw(qk, (path_len << 20) | 3u);
        if (usel < G.mix) {
          bs = bsdf_sample(is, M, cu, cv, F_ALL);
          if (bs.pdf > 0.0f) {
            const V3 wd = is.nf.from(bs.dir);
            bs.pdf = G.mix * bs.pdf + (1.0f - G.mix) * mix_pdf(L, wd.x, wd.y, wd.z);
          }
        } else {
          V3 wd;
          mix_sample(L, draw(qk, (path_len << 20) | 4u), draw(qk, (path_len << 20) | 5u),
                     draw(qk, (path_len << 20) | 6u), wd.x, wd.y, wd.z);
          const V3 ld = is.nf.to(wd);
          const BValue bv = bsdf_eval(is, M, ld, F_ALL);
          bs.val = bv.val;
          bs.dir = ld;
          bs.flags = F_REFLECTIVE | F_DIFFUSE;
          bs.pdf = czero(bv.val) ? 0.0f
                                 : G.mix * bv.pdf + (1.0f - G.mix) * mix_pdf(L, wd.x, wd.y, wd.z);
        }
      } else {
        bs = bsdf_sample(is, M, cu, cv, F_ALL);
      }
      const C3 T_vertex = T;
      if (bs.pdf == 0.0f || czero(bs.val)) {
        done = true;
      } else {
        C3 f = bs.val * fabsf(bs.dir.z);
        f = f * (1.0f / bs.pdf);
        T = T * f;
        float rrr = P.rr[p];
        if (path_len > R.minp) {
          float pc = fmn(((T.r + T.g) + T.b) / 3.0f * rrr, 1.0f);
          float rr = draw(qk, (path_len << 20) | 0u);
          if (rr > pc) {
            done = true;
          } else {
            T = T * (1.0f / pc);
            rrr = rrr * pc;
          }
        }
        if (!done && path_len == R.maxp) done = true;
        if (!done) {
          RayF nr;
          nr.o = is.nf.o;
          nr.d = is.nf.from(bs.dir);
          nr.t0 = S.min_trace;
          nr.t1 = S.horizon;
          store_ray(P, p, nr);
          P.rr[p] = rrr;
          bool spec = (bs.flags & F_SPECULAR) != 0;
          if ((bs.flags & F_TRANSMISSIVE) && HM.medium) {
            // Media::update_stack_for_transmission (material/media.cc:75-118)
            if (is.back) {
              if (depth >= 2) depth--;
            } else if (depth < SGR_MEDIA_MAX) {
              P.media[(size_t)depth * N + p] = M.ior_n;
              depth++;
            } else if (P.media_ovf) {
              atomicAdd(P.media_ovf, 1ull);
            }
          }
          nst = (nst & (SGR_ST_PENDING | SGR_ST_STRAT)) | (path_len + 1) |
                ((spec ? 1u : 0u) << SGR_ST_SPEC_SHIFT) | (depth << SGR_ST_DEPTH_SHIFT);
          cont = true;
          if (guided && grec && path_len == (p / G.rec_stride) % 3u) {
            const float *in = G.in6 + (size_t)j * GD_IN;
            for (int k = 0; k < GD_IN; k++) grec[k] = in[k];
            grec[6] = nr.d.x; grec[7] = nr.d.y; grec[8] = nr.d.z;
            grec[10] = T_vertex.r; grec[11] = T_vertex.g; grec[12] = T_vertex.b;
            grec[9] = -1.0f;
          }
        }
      }
    }
  }
  if (done) nst |= SGR_ST_DONE;
  P.state[p] = nst;
  P.T[0][p] = T.r; P.T[1][p] = T.g; P.T[2][p] = T.b;
  P.rad[0][p] = rad.r; P.rad[1][p] = rad.g; P.rad[2][p] = rad.b;
  P.alpha[p] = alpha;
  return cont;
}

template<bool MULTI, bool GUIDED, bool EXT>
__global__ void __launch_bounds__(256) k_shade(SceneView S1, RenderView R1, StepView J, PathBufs P,
                                               uint32_t b, unsigned long long *stats,
                                               GuideView G) {
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n)) return;
  const uint32_t *q = P.q[b & 1];
  const size_t K = P.K;
  uint32_t illum = 0;
  uint32_t jb, je;
  block_range(n, jb, je);
  for (uint32_t j = jb + threadIdx.x; j < je; j += blockDim.x) {
    const uint32_t p = q[j];
    uint8_t *sf = P.sflag + (size_t)j * K;
    for (size_t k = 0; k < K; k++) sf[k] = 0;
    const bool cont = shade_path<MULTI, GUIDED, EXT>(S1, R1, J, P, G, p, j, illum,
                                                [&](size_t slot, uint32_t oct) {
      sf[slot - (size_t)p * K] = (uint8_t)(1u | oct << 1);
    });
    P.qflag[j] = cont ? 1 : 0;
  }
  wave_count(&stats[SGR_STAT_ILLUM], illum);
}

// Ordered compaction of what k_shade flagged: the next ray queue (paths
// that go on) and the shadow queue (slots with a ray), both in the order
// of the queue they come from (the shadow queue up to a grouping by
// direction octant within a block).  Queues thus stay in raygen's pixel
// order bounce after bounce, so the slice of a queue each XCD takes (see
// block_range) keeps covering the same part of the image, and k_shade
// needs no atomics.  A block takes SGR_COMPACT_ITEMS consecutive
// positions, each of its 4 waves a run of 1024, 64 at a time (lane l
// position base + l, coalesced flag loads): a count pass, one block
// scanning the block counts, a scatter pass.
#define SGR_COMPACT_WAVE (SGR_COMPACT_ITEMS / 4)

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (uint32_t off = 32; off > 0; off >>= 1) v += (uint32_t)__shfl_xor((int)v, off);
  return v;
}

// Continuing paths and shadow rays among the wave's run of positions.
__device__ __forceinline__ void compact_wave_counts(const PathBufs &P, uint32_t n, uint32_t base,
                                                    uint32_t &cq, uint32_t &cs) {
  const uint32_t lane = __lane_id(), K = P.K;
  uint32_t q = 0, sh = 0;
  for (uint32_t it = 0; it < SGR_COMPACT_WAVE; it += 64) {
    const uint32_t j = base + it + lane;
    if (j < n) {
      q += P.qflag[j];
      const uint8_t *sf = P.sflag + (size_t)j * K;
      for (uint32_t k = 0; k < K; k++) sh += sf[k] != 0 ? 1u : 0u;
    }
  }
  cq = wave_sum(q);
  cs = wave_sum(sh);
}

__global__ void __launch_bounds__(256) k_compact_count(PathBufs P, uint32_t b) {
  __shared__ uint32_t part[8];
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n) || blockIdx.x * SGR_COMPACT_ITEMS >= n) return;
  const uint32_t w = threadIdx.x >> 6;
  uint32_t cq, cs;
  compact_wave_counts(P, n, blockIdx.x * SGR_COMPACT_ITEMS + w * SGR_COMPACT_WAVE, cq, cs);
  if (__lane_id() == 0) {
    part[2 * w] = cq;
    part[2 * w + 1] = cs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    P.cblk[2 * blockIdx.x] = (part[0] + part[2]) + (part[4] + part[6]);
    P.cblk[2 * blockIdx.x + 1] = (part[1] + part[3]) + (part[5] + part[7]);
  }
}

// One block: exclusive prefixes of the block counts (in place) and the
// queue lengths.
__global__ void __launch_bounds__(256) k_compact_scan(PathBufs P, uint32_t b) {
  __shared__ uint32_t lds[512];
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n)) return;
  const uint32_t t = threadIdx.x;
  const uint32_t nblk = (n + SGR_COMPACT_ITEMS - 1) / SGR_COMPACT_ITEMS;
  const uint32_t per = (nblk + 255) / 256;
  const uint32_t i0 = t * per < nblk ? t * per : nblk;
  const uint32_t i1 = i0 + per < nblk ? i0 + per : nblk;
  uint32_t cq = 0, cs = 0;
  for (uint32_t i = i0; i < i1; i++) {
    cq += P.cblk[2 * i];
    cs += P.cblk[2 * i + 1];
  }
  // inclusive scan over the 256 threads (Hillis-Steele in LDS)
  lds[t] = cq;
  lds[256 + t] = cs;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t va = t >= off ? lds[t - off] : 0u, vb = t >= off ? lds[256 + t - off] : 0u;
    __syncthreads();
    lds[t] += va;
    lds[256 + t] += vb;
    __syncthreads();
  }
  uint32_t oq = lds[t] - cq, os = lds[256 + t] - cs;
  for (uint32_t i = i0; i < i1; i++) {
    const uint32_t a = P.cblk[2 * i], c = P.cblk[2 * i + 1];
    P.cblk[2 * i] = oq;
    P.cblk[2 * i + 1] = os;
    oq += a;
    os += c;
  }
  if (t == 0) {
    P.counts[SGR_CNT_Q + b + 1] = lds[255];
    P.counts[SGR_CNT_S + b] = lds[511];
  }
}

// Scatter pass.  The next ray queue keeps the order of the queue it
// comes from: k_shade, k_gather and the learned proposal read their
// per-path records through it.  The shadow queue, read by k_occlude
// alone, is written with each block's shadow rays grouped by direction
// octant (shadow flags hold 1 | octant << 1; 0: no ray): rays of one
// octant from neighbouring pixels share the octree nodes they walk, so a
// wave of them touches fewer cache lines (1.31 M-triangle frame 158 ->
// 155 ms; the same grouping of the ray queue, for k_extend alone, gained
// no more there and lost 1.5 % on the Cornell box, its scattered ray
// loads and hit stores outweighing the shorter walks).
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__global__ void __launch_bounds__(256) k_compact_scatter(PathBufs P, uint32_t b) {
  __shared__ uint32_t part[4][9];  // per wave: rays, then shadow rays per octant
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n) || blockIdx.x * SGR_COMPACT_ITEMS >= n) return;
  const uint32_t w = threadIdx.x >> 6, lane = __lane_id(), K = P.K;
  const uint32_t base = blockIdx.x * SGR_COMPACT_ITEMS + w * SGR_COMPACT_WAVE;
  // Counts of the wave's run: continuing paths, and shadow rays per
  // octant (lane o < 8 holds octant o's).
  uint32_t cq = 0, cs = 0;
  for (uint32_t it = 0; it < SGR_COMPACT_WAVE; it += 64) {
    const uint32_t j = base + it + lane;
    const bool in = j < n;
    cq += (uint32_t)__popcll(__ballot(in && P.qflag[j]));
    const uint8_t *sf = P.sflag + (size_t)(in ? j : 0u) * K;
    for (uint32_t k = 0; k < K; k++) {
      const uint32_t fs = in ? sf[k] : 0u;
#pragma unroll
      for (uint32_t o = 0; o < 8; o++) {
        const uint32_t c = (uint32_t)__popcll(__ballot(fs == (1u | o << 1)));
        cs += lane == o ? c : 0u;
      }
    }
  }
  if (lane == 0) part[w][0] = cq;
  if (lane < 8) part[w][1 + lane] = cs;
  __syncthreads();
  // Starts: the wave's run of the ray queue, and per octant (lane o) the
  // wave's place in the block's octant-grouped shadow runs.
  uint32_t oq = P.cblk[2 * blockIdx.x], offs = P.cblk[2 * blockIdx.x + 1];
  for (uint32_t v = 0; v < 4; v++) {
    if (v < w) oq += part[v][0];
    for (uint32_t o = 0; o < 8; o++)
      if (o < lane || (o == lane && v < w)) offs += part[v][1 + o];
  }
  const uint32_t *q = P.q[b & 1];
  uint32_t *qn = P.q[(b + 1) & 1];
  for (uint32_t it = 0; it < SGR_COMPACT_WAVE; it += 64) {
    const uint32_t j = base + it + lane;
    const bool in = j < n;
    const uint32_t p = in ? q[j] : 0u;
    const bool go = in && P.qflag[j];
    const uint64_t m = __ballot(go);
    if (go) qn[oq + lane_rank(m)] = p;
    oq += (uint32_t)__popcll(m);
    const uint8_t *sf = P.sflag + (size_t)(in ? j : 0u) * K;
    for (uint32_t k = 0; k < K; k++) {
      const uint32_t fs = in ? sf[k] : 0u;
#pragma unroll
      for (uint32_t o = 0; o < 8; o++) {
        const uint64_t mo = __ballot(fs == (1u | o << 1));
        const uint32_t at = (uint32_t)__builtin_amdgcn_readlane((int)offs, (int)o);
        if (fs == (1u | o << 1)) P.sq[at + lane_rank(mo)] = (uint32_t)((size_t)p * K + k);
        offs += lane == o ? (uint32_t)__popcll(mo) : 0u;
      }
    }
  }
}

// Adds the unoccluded direct terms (render/direct-illum.cc:96-113 sums,
// render/path-integ.cc:277-306 product with the path transmittance) and
// writes finished paths: sx, sy, Tint(radiance, alpha).
template<bool MULTI, bool GUIDED>
__device__ __forceinline__ void gather_path(const RenderView &R1, const StepView &J,
                                            const PathBufs &P, const GuideView &G, uint32_t p) {
  uint32_t st = P.state[p];
  if (!(st & (SGR_ST_PENDING | SGR_ST_DONE))) return;
  const RenderView R = MULTI ? J.R[P.pseg[p]] : R1;
  C3 rad = c3(P.rad[0][p], P.rad[1][p], P.rad[2][p]);
  if (st & SGR_ST_PENDING) {
    uint32_t nsub = (st & SGR_ST_STRAT) ? R.nsub_strat : R.nsub_rand;
    float inv_n = 1.0f / (float)nsub;
    C3 D = c3(0.0f);
    for (uint32_t l = 0; l < R.L; l++) {
      C3 lrad = c3(0.0f);
      for (uint32_t jj = 0; jj < nsub; jj++) {
        size_t sa = (size_t)p * P.K + 2 * (l * nsub + jj);
        C3 r = c3(0.0f);
        C3 ca = c3(P.sc[0][sa], P.sc[1][sa], P.sc[2][sa]);
        C3 cb = c3(P.sc[0][sa + 1], P.sc[1][sa + 1], P.sc[2][sa + 1]);
        if (P.partial) {  // through partial occluders
          ca = ca * get_trans(P, sa);
          cb = cb * get_trans(P, sa + 1);
        }
        if (P.svis[sa]) r = r + ca;
        if (P.svis[sa + 1]) r = r + cb;
        lrad = lrad + r;
      }
      D = D + lrad * inv_n;
    }
    rad = rad + D * c3(P.Tp[0][p], P.Tp[1][p], P.Tp[2][p]);
    P.rad[0][p] = rad.r; P.rad[1][p] = rad.g; P.rad[2][p] = rad.b;
    st &= ~(SGR_ST_PENDING | SGR_ST_STRAT);
    P.state[p] = st;
  }
  if (st & SGR_ST_DONE) {
    if (GUIDED && p % G.rec_stride == 0) {
      // training weight: radiance the recorded continuation brought,
      // over the throughput at its vertex (= f cos L / pdf, averaged)
      float *grec = G.rec + (size_t)permute(p / G.rec_stride, G.rec_count, G.rec_key) * GD_REC;
      if (grec[9] == -2.0f) {
        const float d0 = rad.r - grec[13], d1 = rad.g - grec[14], d2 = rad.b - grec[15];
        float w = 0.0f;
        w += grec[10] > 0.0f ? d0 / grec[10] : 0.0f;
        w += grec[11] > 0.0f ? d1 / grec[11] : 0.0f;
        w += grec[12] > 0.0f ? d2 / grec[12] : 0.0f;
        grec[9] = fmaxf(w * (1.0f / 3.0f), 0.0f);
      } else if (grec[9] < 0.0f) {
        grec[9] = 0.0f;
      }
    }
    float a = P.alpha[p];
    float *o = P.out + 6 * (size_t)p;
    o[2] = rad.r * a;
    o[3] = rad.g * a;
    o[4] = rad.b * a;
    o[5] = a;
  }
}

template<bool MULTI, bool GUIDED>
__global__ void __launch_bounds__(256) k_gather(RenderView R1, StepView J, PathBufs P, uint32_t b,
                                                GuideView G) {
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (finishing(P, b, n)) return;
  const uint32_t *q = P.q[b & 1];
  uint32_t jb, je;
  block_range(n, jb, je);
  for (uint32_t j = jb + threadIdx.x; j < je; j += blockDim.x)
    gather_path<MULTI, GUIDED>(R1, J, P, G, q[j]);
}

// Finishes every path of a short bounce-B queue in one launch: each
// thread takes one path through all its remaining bounces (closest hit,
// the vertex, its shadow rays, the gather), in the same order and with
// the same per-path code as the per-bounce kernels, so the results are
// the same bit for bit.  Late bounces hold few rays whose walks are long
// chains of dependent loads; bounce by bounce, every launch would wait
// for its slowest ray, here only each path's own chain counts.  Rays and
// shadow rays traced after bounce B are counted here (the queue counts
// of those bounces stay zero).
template<bool MULTI, bool INST>
__global__ void __launch_bounds__(256) k_finish(SceneView S1, RenderView R1, StepView J,
                                                PathBufs P, uint32_t b,
                                                unsigned long long *stats) {
  const uint32_t n = P.counts[SGR_CNT_Q + b];
  if (!finishing(P, b, n)) return;
  SGR_STACK_LDS(stk);
  uint2 spill[SGR_MAX_DEPTH];
  Nest N{};
  if constexpr (INST) {  // stack of the walk inside an instanced model (LDS
                         // arrays are static: they outlive the block)
    SGR_STACK_LDS(stk2);
    N.L = stk2;
  }
  uint2 spill2[INST ? SGR_MAX_DEPTH : 1];
  N.spill = spill2;
  const uint32_t *q = P.q[b & 1];
  const GuideView G{};
  TraceCount cc{0, 0, 0, 0}, cs{0, 0, 0, 0};
  uint32_t illum = 0, rays = 0, shadow = 0;
  uint32_t jb, je;
  block_range(n, jb, je);
  for (uint32_t j = jb + threadIdx.x; j < je; j += blockDim.x) {
    const uint32_t p = q[j];
    const SceneView *Sp = MULTI ? &J.S[J.R[P.pseg[p]].scene] : &S1;
    for (;;) {
      const SceneView &S = *Sp;
      TravState T;
      load_ray(P, p, T.ray);
      trace_one<false, INST>(S, T, cc, stk, spill, N);
      P.hprim[j] = T.hit;
      P.ht[j] = T.ray.t1;
      if (S.tri_vn) {  // read only to interpolate vertex normals
        P.hu[j] = T.hu;
        P.hv[j] = T.hv;
      }
      if (INST) P.hinst[j] = T.hinst;
      uint64_t em = 0;  // shadow slots of the path written at this vertex
      const bool cont = shade_path<MULTI, false, INST>(S1, R1, J, P, G, p, j, illum,
                                                 [&](size_t slot, uint32_t) {
        em |= 1ull << (uint32_t)(slot - (size_t)p * P.K);
      });
      for (uint64_t m = em; m; m &= m - 1) {
        const size_t slot = (size_t)p * P.K + (uint32_t)__builtin_ctzll(m);
        TravState U;
        U.ray.o = v3(P.sray[0][slot], P.sray[1][slot], P.sray[2][slot]);
        U.ray.d = v3(P.sray[3][slot], P.sray[4][slot], P.sray[5][slot]);
        U.ray.t0 = P.sray[6][slot];
        U.ray.t1 = P.sray[7][slot];
        U.medium = (INST && P.partial) ? P.smed[slot] : 1.0f;
        trace_one<true, INST>(S, U, cs, stk, spill, N);
        P.svis[slot] = U.hit == SGR_MISS ? 1 : 0;
        if (INST && P.partial) put_trans(P, slot, U.trans);
        shadow++;
      }
      gather_path<MULTI, false>(R1, J, P, G, p);
      if (!cont) break;
      rays++;
    }
  }
  wave_count(&stats[SGR_STAT_NODES_C], cc.nodes);
  wave_count(&stats[SGR_STAT_TESTS_C], cc.tests);
  wave_count(&stats[SGR_STAT_NODES_S], cs.nodes);
  wave_count(&stats[SGR_STAT_TESTS_S], cs.tests);
  wave_count(&stats[SGR_STAT_FETCH_C], cc.fetch);
  wave_count(&stats[SGR_STAT_FETCH_S], cs.fetch);
  if (INST) wave_count(&stats[SGR_STAT_PARTIAL_OVF], cs.povf);
  wave_count(&stats[SGR_STAT_ILLUM], illum);
  wave_count(&stats[SGR_STAT_ISECT], rays);
  wave_count(&stats[SGR_STAT_SHADOW], shadow);
}

// Direct-lighting integrator (render/direct-integ.h:31-67 over
// render/recursive-integ.cc:29-216), one thread per path of the chunk:
// the sample's whole recursion.  At each hit: emission, direct lighting
// of every light with nsub_dint light and BSDF samples (the same
// stratified samples at every depth, as Lo gets the one Sample; shadow
// rays traced at once), then the perfectly specular reflection and
// transmission, each weighted by val |cos|.  Beyond depth 5 a branch goes
// on with probability 1/2 and twice the weight.  The roulette and the
// sphere light's extra coordinate come from the path's sequential
// stream (the reference's context.random()).  The recursion runs depth
// first, reflection before transmission, on a stack of pending
// transmissions (at most one per level); a branch keeps its weight and
// its media as (innermost, enclosing) IORs: entering a medium pushes
// (its IOR, innermost), leaving sets both to the enclosing IOR, as
// Media (isec, true) does (material/media.cc:25-52).  Radiance is
// accumulated top-down, weight times emission and direct light.
#define SGR_DINT_RAND 0xc0000000u  // draw dims of the sequential stream
struct DintBranch {
  RayF ray;
  C3 W;
  float m0, m1;  // innermost and enclosing IOR
  uint32_t depth;
};

__device__ __forceinline__ RayF dint_ray(const SceneView &S, const Isec &is, V3 dir) {
  RayF r;  // Intersect::recursive_ray (render/intersect.cc:195-218)
  r.o = is.nf.o;
  r.d = is.nf.from(dir);
  r.t0 = S.min_trace;
  r.t1 = S.horizon;
  return r;
}

// PhotonShooter::shoot's paths (photon/photon-shooter.cc:29-191), one per
// thread: path N takes light floor(radical_inverse(N, 11) x lights) and
// its free sample at (radical_inverse(N, 2), (N, 3)) / ((N, 5), (N, 7)),
// then bounces by BSDF samples at ((N, 13), (N, 17)) first and from its
// own counter stream (two draws, then past 3 bounces the roulette draw,
// per bounce).  At every hit whose BSDF has a non-specular part it
// records a deposit and the set it belongs to by the scattering so far
// (render/photon-integ.cc:196-226: none yet, direct; only specular
// scattering with a caustic set wanted, caustic; else indirect); which
// deposits the maps keep is decided in path order on the host.  Deposits
// are held per thread, then written to the batch's array at a reserved
// range.  The media stack is the path integrator's (8 levels).  A path
// ends after SGR_PHOTON_MAX_BOUNCES hits (past 3 the roulette halves its
// chance at each, so that is reached with probability 2^-252).
template<bool INST>
__global__ void __launch_bounds__(256) k_photon_shoot(SceneView S, PhotonShootView V,
                                                      unsigned long long *stats) {
  SGR_STACK_LDS(stk);
  uint2 spill[SGR_MAX_DEPTH];
  Nest N{};
  if constexpr (INST) {
    SGR_STACK_LDS(stk2);
    N.L = stk2;
  }
  uint2 spill2[INST ? SGR_MAX_DEPTH : 1];
  N.spill = spill2;
  float4 dep[3 * SGR_PHOTON_MAX_DEPOSITS];
  TraceCount cc{0, 0, 0, 0};
  uint32_t rays = 0, movf = 0;
  const uint64_t pk = pixel_key(V.seed, -2, -2);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < V.n; t += gridDim.x * blockDim.x) {
    const uint32_t path = V.base + t;
    const uint32_t li = (uint32_t)(radical_inverse(path, 11) * (double)S.n_lights);
    const DevLight L = S.lights[li];
    const FreeS fs = light_free_sample<INST>(S, L, (float)radical_inverse(path, 2),
                                             (float)radical_inverse(path, 3),
                                             (float)radical_inverse(path, 5),
                                             (float)radical_inverse(path, 7));
    if (czero(fs.val) || fs.pdf == 0.0f) {
      V.rec[t] = make_uint2(0u, 0u);
      continue;
    }
    const uint64_t qk = path_key(pk, path);
    uint32_t nd = 0, hist = 0, ndep = 0;
    float media[SGR_MEDIA_MAX];
    media[0] = 1.0f;
    uint32_t depth = 1;
    V3 pos = fs.pos, dir = fs.dir;
    C3 power = (fs.val * (float)S.n_lights) * (1.0f / fs.pdf);  // Color * float / float
    for (uint32_t plen = 0; plen < SGR_PHOTON_MAX_BOUNCES; plen++) {
      TravState T;
      T.ray.o = pos;
      T.ray.d = dir;
      T.ray.t0 = S.min_trace;
      T.ray.t1 = S.horizon;
      const RayF ray = T.ray;
      trace_one<false, INST>(S, T, cc, stk, spill, N);
      rays++;
      if (T.hit == SGR_MISS) break;
      RayF hr = ray;
      hr.t1 = T.ray.t1;
      Isec is;
      float encl = 1.0f;
#pragma unroll
      for (uint32_t q = 0; q < SGR_MEDIA_MAX; q++)
        if (q + 2 == depth) encl = media[q];
      float top = 1.0f;
#pragma unroll
      for (uint32_t q = 0; q < SGR_MEDIA_MAX; q++)
        if (q + 1 == depth) top = media[q];
      make_isec<INST>(S, T.hit, T.hinst, hr, S.tri_vn ? T.hu : 0.0f, S.tri_vn ? T.hv : 0.0f, top,
                      encl, is);
      if (is.bmat < 0) break;
      const DevMaterial &M = S.mats[is.bmat];
      if (bsdf_supports_any(M, F_ALL & ~F_SPECULAR)) {
        if (ndep < SGR_PHOTON_MAX_DEPOSITS) {
          const uint32_t set = hist == 0 ? 0u
                               : (V.caustic_on && !(hist & (F_GLOSSY | F_DIFFUSE))) ? 1u : 2u;
#pragma unroll
          for (uint32_t q = 0; q < SGR_PHOTON_MAX_DEPOSITS; q++)
            if (q == ndep) {
              dep[3 * q] = make_float4(is.nf.o.x, is.nf.o.y, is.nf.o.z, __uint_as_float(set));
              dep[3 * q + 1] = make_float4(power.r, power.g, power.b, 0.0f);
              dep[3 * q + 2] = make_float4(-dir.x, -dir.y, -dir.z, 0.0f);
            }
        }
        ndep++;
      }
      float u, v;
      if (plen == 0) {
        u = (float)radical_inverse(path, 13);
        v = (float)radical_inverse(path, 17);
      } else {
        u = draw(qk, nd++);
        v = draw(qk, nd++);
      }
      const BSample bs = bsdf_sample(is, M, u, v, F_ALL);
      if (czero(bs.val) || bs.pdf == 0.0f) break;
      if (plen > 3) {
        if (draw(qk, nd++) < 0.5f) break;
        power = power * (1.0f / 0.5f);
      }
      pos = is.nf.o;
      dir = is.nf.from(bs.dir);
      power = power * ((bs.val * fabsf(bs.dir.z)) * (1.0f / bs.pdf));
      if (!(bs.flags & 0x08u)) hist |= bs.flags;  // TRANSLUCENT samples leave none
      if ((bs.flags & F_TRANSMISSIVE) && S.mats[is.mat].medium) {  // material/media.cc:75-118
        const float mior = S.mats[is.bmat].ior_n;
        if (is.back) {
          if (depth >= 2) depth--;
        } else if (depth < SGR_MEDIA_MAX) {
#pragma unroll
          for (uint32_t q = 0; q < SGR_MEDIA_MAX; q++)
            if (q == depth) media[q] = mior;
          depth++;
        } else {
          movf++;
        }
      }
    }
    const uint32_t nst = ndep < SGR_PHOTON_MAX_DEPOSITS ? ndep : SGR_PHOTON_MAX_DEPOSITS;
    const uint32_t off = nst ? atomicAdd(V.counter, nst) : 0u;
    if (off + nst <= V.cap) {
      for (uint32_t q = 0; q < nst; q++) {
        V.dep[3 * (size_t)(off + q)] = dep[3 * q];
        V.dep[3 * (size_t)(off + q) + 1] = dep[3 * q + 1];
        V.dep[3 * (size_t)(off + q) + 2] = dep[3 * q + 2];
      }
    }
    V.rec[t] = make_uint2(off, ndep | 0x80000000u);
  }
  wave_count(&stats[SGR_STAT_NODES_C], cc.nodes);
  wave_count(&stats[SGR_STAT_TESTS_C], cc.tests);
  wave_count(&stats[SGR_STAT_FETCH_C], cc.fetch);
  wave_count(&stats[SGR_STAT_ISECT], rays);
  wave_count(&stats[SGR_STAT_MEDIA_OVF], movf);
}

// PhotonInteg::Lo_fgather_samp (render/photon-integ.cc:268-380) without
// recursion: the photon maps' radiance at the surface the sample BS from
// IS reaches, then up to 3 levels of specular bounces from there (their
// parameters from the sample's sequential stream), each level's
// radiance weighted by the Li -> Lo factors above it; past level 0 the
// surface's emission, or the background on a miss, counts times SCALE.
// The media are the shading point's throughout (M0, M1).
template<bool INST>
__device__ __forceinline__ C3 fg_samp(const SceneView &S, const RenderView &R, Isec is, BSample bs,
                                      C3 scale, float m0, float m1, uint64_t qk, uint32_t &nd,
                                      PhotonHeap &H, TraceCount &cc, uint32_t &rays, LdsStack stk,
                                      uint2 *spill, Nest N) {
  C3 rad = c3(0.0f), W = c3(1.0f);
  for (uint32_t depth = 0;; depth++) {
    if (!(cgt0(bs.val) && bs.pdf != 0.0f)) break;
    TravState T;
    T.ray = dint_ray(S, is, bs.dir);
    const RayF ray = T.ray;
    trace_one<false, INST>(S, T, cc, stk, spill, N);
    rays++;
    if (T.hit == SGR_MISS) {
      if (depth != 0) rad = rad + (background(S, ray.d) * scale) * W;
      break;
    }
    RayF hr = ray;
    hr.t1 = T.ray.t1;
    Isec si;
    make_isec<INST>(S, T.hit, T.hinst, hr, S.tri_vn ? T.hu : 0.0f, S.tri_vn ? T.hv : 0.0f, m0, m1,
                    si);
    bool more = false;
    BSample next;
    C3 lil;
    if (si.bmat >= 0) {
      const DevMaterial &SM = S.mats[si.bmat];
      const C3 Li = (photon_lo(si, SM, R.ph.set[0], R.ph.k, R.ph.r2, F_ALL, H) +
                     photon_lo(si, SM, R.ph.set[2], R.ph.k, R.ph.r2, F_ALL, H)) +
                    photon_lo(si, SM, R.ph.set[1], R.ph.k, R.ph.r2, F_ALL, H);
      lil = (bs.val * fabsf(bs.dir.z)) * (1.0f / bs.pdf);
      rad = rad + (Li * lil) * W;
      const unsigned spec = F_ALL_DIRECTIONS | F_SPECULAR;
      if (bsdf_supports_any(SM, spec) && depth < 3) {
        const float u = draw(qk, SGR_DINT_RAND + nd++);
        const float v = draw(qk, SGR_DINT_RAND + nd++);
        next = bsdf_sample(si, SM, u, v, spec);
        more = true;
      }
    }
    if (depth != 0) {
      const DevMaterial &HM = S.mats[si.mat];
      if (HM.type == 5 && !si.back) rad = rad + (cload(HM.color) * scale) * W;  // Intersect::Le
    }
    if (!more) break;
    W = W * lil;
    is = si;
    bs = next;
  }
  return rad;
}

// PhotonInteg::Lo_fgather (render/photon-integ.cc:397-545): samples of
// the nearby indirect photons' direction histogram and of the BSDF, each
// through fg_samp, weighted by the power heuristic with sample counts.
template<bool INST>
__device__ __forceinline__ C3 fg_gather(const SceneView &S, const RenderView &R, const Isec &is,
                                        const DevMaterial &M, float m0, float m1, uint64_t pk,
                                        uint32_t s, uint64_t qk, uint32_t &nd, PhotonHeap &H,
                                        TraceCount &cc, uint32_t &rays, LdsStack stk, uint2 *spill,
                                        Nest N) {
  C3 rad = c3(0.0f);
  const uint32_t nph = R.ph.fg_photon, nbs = R.ph.fg_bsdf;
  const bool avoid = R.ph.caustics != 0;
  DirHistDist D;
  photon_find(R.ph.set[2], is.nf.o, R.ph.k, R.ph.r2, H);
  dirhist_build(D, R.ph.set[2], H);
  const unsigned bflags = bsdf_supports_any(M, F_ALL);
  for (uint32_t i = 0; i < nph; i++) {
    float pu, pv, spdf;
    strat_uv(pk, 1 + 2 * R.L, s * R.ph.nsub_fgp + i, R.ph.g_fgp, pu, pv);
    const V3 bdir = is.nf.to(dirhist_sample(D, pu, pv, spdf));
    const BValue bv = bsdf_eval(is, M, bdir, F_ALL & ~F_SPECULAR);
    if (bv.pdf != 0.0f && spdf != 0.0f) {
      C3 scale = c3(1.0f);
      if (avoid) {
        if (bflags & (F_SPECULAR | F_GLOSSY | F_DIFFUSE) & ~(F_SPECULAR | F_DIFFUSE))
          scale = cdiv(bsdf_eval(is, M, bdir, F_ALL & ~(F_SPECULAR | F_DIFFUSE)).val, bv.val);
        else
          scale = c3(0.0f);
      }
      BSample fake;
      fake.val = bv.val;
      fake.pdf = spdf;
      fake.dir = bdir;
      fake.flags = 0;
      rad = rad + fg_samp<INST>(S, R, is, fake, scale, m0, m1, qk, nd, H, cc, rays, stk, spill, N) *
                      mis_sample_weight(spdf, (float)nph, bv.pdf, (float)nbs);
    }
  }
  for (uint32_t i = 0; i < nbs; i++) {
    float bu, bv;
    strat_uv(pk, 2 + 2 * R.L, s * R.ph.nsub_fgb + i, R.ph.g_fgb, bu, bv);
    const BSample bs = bsdf_sample(is, M, bu, bv, F_ALL & ~F_SPECULAR);
    if (bs.pdf != 0.0f) {
      const float ph_pdf = dirhist_pdf(D, is.nf.from(bs.dir));
      const C3 scale = (avoid && (bs.flags & F_DIFFUSE)) ? c3(0.0f) : c3(1.0f);
      rad = rad + fg_samp<INST>(S, R, is, bs, scale, m0, m1, qk, nd, H, cc, rays, stk, spill, N) *
                      mis_sample_weight(bs.pdf, (float)nbs, ph_pdf, (float)nph);
    }
  }
  return rad;
}

// PHOTON: the photon integrator (render/photon-integ.cc:558-600), whose
// Lo adds caustic photons and final gathering (or the indirect map) to
// the direct lighting (or the direct map).
template<bool INST, bool PHOTON>
__global__ void __launch_bounds__(256) k_direct_integ(SceneView S, RenderView R, PathBufs P,
                                                      unsigned long long *stats) {
  SGR_STACK_LDS(stk);
  uint2 spill[SGR_MAX_DEPTH];
  Nest N{};
  if constexpr (INST) {
    SGR_STACK_LDS(stk2);
    N.L = stk2;
  }
  uint2 spill2[INST ? SGR_MAX_DEPTH : 1];
  N.spill = spill2;
  DintBranch pend[SGR_DIRECT_MAX_DEPTH];
  PhotonHeap H;
  TraceCount cc{0, 0, 0, 0}, cs{0, 0, 0, 0};
  uint32_t illum = 0, rays = 0, shadow = 0, ovf = 0;
  const uint32_t nsub = R.nsub_dint;
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < R.npaths;
       p += gridDim.x * blockDim.x) {
    int px, py;
    uint32_t s;
    path_ident(R, p, px, py, s);
    const uint64_t pk = pixel_key(R.seed, px, py), qk = path_key(pk, s);
    uint32_t nd = 0;  // sequential draws taken
    C3 total = c3(0.0f);
    float alpha = 1.0f;
    DintBranch cur;
    load_ray(P, p, cur.ray);
    cur.W = c3(1.0f);
    cur.m0 = 1.0f;  // RenderContext::default_medium
    cur.m1 = 1.0f;
    cur.depth = 0;
    uint32_t sp = 0;
    for (;;) {
      bool child = false;  // CUR was replaced by its reflection branch
      bool live = true;
      if (cur.depth > 5) {  // render/recursive-integ.cc:115-138
        if (draw(qk, SGR_DINT_RAND + nd++) < 0.5f) live = false;
        else cur.W = cur.W * 2.0f;
      }
      if (live) {
        TravState T;
        T.ray = cur.ray;
        trace_one<false, INST>(S, T, cc, stk, spill, N);
        rays++;
        if (T.hit == SGR_MISS) {
          total = total + background(S, cur.ray.d) * cur.W;
          if (cur.depth == 0) alpha = R.bg_alpha;
        } else {
          RayF hr = cur.ray;
          hr.t1 = T.ray.t1;
          Isec is;
          make_isec<INST>(S, T.hit, T.hinst, hr, S.tri_vn ? T.hu : 0.0f, S.tri_vn ? T.hv : 0.0f,
                    cur.m0, cur.m1, is);
          const DevMaterial &HM = S.mats[is.mat];
          if (HM.type == 5 && !is.back) total = total + cload(HM.color) * cur.W;
          if (is.bmat >= 0) {
            const DevMaterial &M = S.mats[is.bmat];
            illum++;
            C3 D = c3(0.0f);
            for (uint32_t l = 0; l < ((!PHOTON || R.ph.direct_illum) ? R.L : 0u); l++) {
              const DevLight &L = S.lights[l];
              C3 lrad = c3(0.0f);
              for (uint32_t jj = 0; jj < nsub; jj++) {
                float lu, lv, bu, bv;
                strat_uv(pk, 1 + 2 * l, s * nsub + jj, R.g_dint, lu, lv);
                strat_uv(pk, 2 + 2 * l, s * nsub + jj, R.g_dint, bu, bv);
                const float extra = L.kind == 1 ? draw(qk, SGR_DINT_RAND + nd++) : 0.0f;
                C3 r = c3(0.0f);
                auto visible = [&](C3 c, const RayF &ray) {
                  TravState U;
                  U.ray = ray;
                  U.medium = cur.m0;  // the medium at the hit (isec.media)
                  trace_one<true, INST>(S, U, cs, stk, spill, N);
                  shadow++;
                  if (U.hit == SGR_MISS) r = r + ((INST && S.has_partial) ? c * U.trans : c);
                };
                direct_light_term<INST>(S, L, is, M, lu, lv, extra, visible);
                direct_bsdf_term<INST>(S, L, is, M, bu, bv, visible);
                lrad = lrad + r;
              }
              D = D + lrad * (1.0f / (float)nsub);  // Color / n: times 1 / n
            }
            if (PHOTON) {
              if (!R.ph.direct_illum) D = photon_lo(is, M, R.ph.set[0], R.ph.k, R.ph.r2, F_ALL, H);
              const bool fg = R.ph.fg > 0;
              if (R.ph.caustics)
                D = D + photon_lo(is, M, R.ph.set[1], R.ph.k, R.ph.r2,
                                  fg ? (F_ALL_DIRECTIONS | F_DIFFUSE) : F_ALL, H);
              if (fg)
                D = D + fg_gather<INST>(S, R, is, M, cur.m0, cur.m1, pk, s, qk, nd, H, cc, rays, stk,
                                        spill, N);
              else
                D = D + photon_lo(is, M, R.ph.set[2], R.ph.k, R.ph.r2, F_ALL, H);
            }
            total = total + D * cur.W;
            const BSample xs = bsdf_sample(is, M, 0.0f, 0.0f, F_SPECULAR | F_TRANSMISSIVE);
            const BSample rs = bsdf_sample(is, M, 0.0f, 0.0f, F_SPECULAR | F_REFLECTIVE);
            if (cgt0(xs.val)) {
              if (cur.depth + 1 >= SGR_DIRECT_MAX_DEPTH) {
                ovf++;
              } else {
                DintBranch &b = pend[sp++];
                b.ray = dint_ray(S, is, xs.dir);
                b.W = cur.W * (xs.val * fabsf(xs.dir.z));
                b.m0 = is.back ? cur.m1 : (HM.medium ? M.ior_n : 1.0f);
                b.m1 = is.back ? cur.m1 : cur.m0;
                b.depth = cur.depth + 1;
              }
            }
            if (cgt0(rs.val)) {
              if (cur.depth + 1 >= SGR_DIRECT_MAX_DEPTH) {
                ovf++;
              } else {
                cur.ray = dint_ray(S, is, rs.dir);
                cur.W = cur.W * (rs.val * fabsf(rs.dir.z));
                cur.depth++;
                child = true;
              }
            }
          }
        }
      }
      if (!child) {
        if (sp == 0) break;
        cur = pend[--sp];
      }
    }
    if (!czero(total)) alpha = 1.0f;  // render/recursive-integ.cc:212-213
    float *o = P.out + 6 * (size_t)p;
    o[2] = total.r * alpha;
    o[3] = total.g * alpha;
    o[4] = total.b * alpha;
    o[5] = alpha;
  }
  wave_count(&stats[SGR_STAT_NODES_C], cc.nodes);
  wave_count(&stats[SGR_STAT_TESTS_C], cc.tests);
  wave_count(&stats[SGR_STAT_NODES_S], cs.nodes);
  wave_count(&stats[SGR_STAT_TESTS_S], cs.tests);
  wave_count(&stats[SGR_STAT_FETCH_C], cc.fetch);
  wave_count(&stats[SGR_STAT_FETCH_S], cs.fetch);
  if (INST) wave_count(&stats[SGR_STAT_PARTIAL_OVF], cs.povf);
  wave_count(&stats[SGR_STAT_ILLUM], illum);
  wave_count(&stats[SGR_STAT_ISECT], rays);
  wave_count(&stats[SGR_STAT_SHADOW], shadow);
  wave_count(&stats[SGR_STAT_DINT_OVF], ovf);
}

// image/image-mitchell-filt.h:45-65
__device__ __forceinline__ float mitchell1(float x, float inv_radius, float B, float Cc) {
  x = fabsf(2.0f * x * inv_radius);
  if (x > 2.0f) return 0.0f;
  if (x > 1.0f)
    return ((-B - 6 * Cc) * x * x * x + (6 * B + 30 * Cc) * x * x + (-12 * B - 48 * Cc) * x +
            (8 * B + 24 * Cc)) * (1.0f / 6.0f);
  return ((12 - 9 * B - 6 * Cc) * x * x * x + (-18 + 12 * B + 6 * Cc) * x * x + (6 - 2 * B)) *
         (1.0f / 6.0f);
}

// Gather form of ImageFilterConv::add_sample: each image pixel sums the
// weighted samples of the pattern pixels within the filter radius (a
// sample of pattern pixel x reaches image pixels int(sx) +- radius; the
// taps that truncation adds or drops all carry weight 0).  No atomics,
// fixed summation order.  The image pixels of a segment are those of its
// window: a limited render's output is the window alone
// (snogray.lua:447-454, valid_x / valid_y of
// image/image-sampled-output.h:155-156), so the samples of the margin
// around it only reach window pixels.
//
// A workgroup takes a 16 x 16 tile of image pixels.  Per pass it stages
// the records (sx, sy, r, g, b, a) of G consecutive samples of every
// pattern pixel of the tile plus its filter halo into LDS with
// contiguous loads (a pixel's samples are adjacent in the output
// buffer), then every thread accumulates its pixel's taps from LDS:
// each record is read from memory once per tile instead of once per tap.
// Per pixel the sum runs over passes, then halo rows and columns, then
// the samples of the pass.
#define SPLAT_T 16
#define SPLAT_LDS_RECS 1600   // records staged per pass (38.4 KB)

__device__ __forceinline__ void splat_bounds(const RenderView &R, int &ox0, int &oy0, int &ow,
                                             int &oh) {
  // the window is the pattern without its filter margin (inside the image)
  ox0 = R.X0 + R.fr;
  oy0 = R.Y0 + R.fr;
  const int ox1 = R.X0 + R.PW - R.fr, oy1 = R.Y0 + R.PH - R.fr;
  ow = ox1 - ox0;
  oh = oy1 - oy0;
}

__device__ __forceinline__ uint32_t splat_tiles(const RenderView &R) {
  int ox0, oy0, ow, oh;
  splat_bounds(R, ox0, oy0, ow, oh);
  if (ow <= 0 || oh <= 0) return 0u;
  return (uint32_t)((ow + SPLAT_T - 1) / SPLAT_T) * (uint32_t)((oh + SPLAT_T - 1) / SPLAT_T);
}

template<bool MULTI>
__global__ void __launch_bounds__(256) k_splat(RenderView R1, StepView J, const float *out) {
  __shared__ float rec[6][SPLAT_LDS_RECS];
  const uint32_t total = MULTI ? J.hdr[2] : splat_tiles(R1);
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    const RenderView R = MULTI ? J.R[find_segment(J.out_base, J.hdr[0], t)] : R1;
    int ox0, oy0, ow, oh;
    splat_bounds(R, ox0, oy0, ow, oh);
    const uint32_t ti = t - (MULTI ? R.out_base : 0u);
    const uint32_t ntx = (uint32_t)(ow + SPLAT_T - 1) / SPLAT_T;
    const int tx0 = ox0 + (int)(ti % ntx) * SPLAT_T, ty0 = oy0 + (int)(ti / ntx) * SPLAT_T;
    const int fr = R.fr;
    const int RW = SPLAT_T + 2 * fr;            // staged region, pattern pixels
    const uint32_t region = (uint32_t)(RW * RW);
    uint32_t G = SPLAT_LDS_RECS / region;       // samples per pass
    if (G > R.sb) G = R.sb;
    const int lx = (int)(threadIdx.x % SPLAT_T), ly = (int)(threadIdx.x / SPLAT_T);
    const int qx = tx0 + lx, qy = ty0 + ly;
    const bool live = qx < ox0 + ow && qy < oy0 + oh;
    float acc[5] = {0, 0, 0, 0, 0};
    for (uint32_t g0 = 0; g0 < R.sb; g0 += G) {
      const uint32_t gn = R.sb - g0 < G ? R.sb - g0 : G;
      const uint32_t nrec = region * gn;
      __syncthreads();  // previous pass done with the records
      for (uint32_t e = threadIdx.x; e < nrec; e += blockDim.x) {
        const uint32_t cell = e / gn, k = e - cell * gn;
        const int sx = tx0 - fr + (int)(cell % (uint32_t)RW);
        const int sy = ty0 - fr + (int)(cell / (uint32_t)RW);
        float v[6] = {-1e30f, -1e30f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (sx >= R.X0 && sx < R.X0 + R.PW && sy >= R.Y0 && sy < R.Y0 + R.PH) {
          const uint32_t pix = (uint32_t)(sy - R.Y0) * (uint32_t)R.PW + (uint32_t)(sx - R.X0);
          const float2 *o = (const float2 *)(out + 6 * ((size_t)R.path_base +
                                                        (size_t)pix * R.sb + g0 + k));
          const float2 a = o[0], b = o[1], c = o[2];
          v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
        }
        const uint32_t slot = cell * gn + k;
        for (int f = 0; f < 6; f++) rec[f][slot] = v[f];
      }
      __syncthreads();
      if (!live) continue;
      for (int dy = -fr; dy <= fr; dy++)
        for (int dx = -fr; dx <= fr; dx++) {
          const uint32_t cell = (uint32_t)((ly + fr + dy) * RW + (lx + fr + dx));
          for (uint32_t k = 0; k < gn; k++) {
            const uint32_t slot = cell * gn + k;
            const float fx = rec[0][slot], fy = rec[1][slot];
            if (fx < -1e29f) continue;  // outside the pattern window
            const int ix = (int)fx, iy = (int)fy;
            if (ix - qx > fr || qx - ix > fr || iy - qy > fr || qy - iy > fr) continue;
            const float w = mitchell1(fx - ((float)qx + 0.5f), R.inv_radius, R.fb, R.fc) *
                            mitchell1(fy - ((float)qy + 0.5f), R.inv_radius, R.fb, R.fc);
            float ws0 = rec[2][slot] * w, ws1 = rec[3][slot] * w, ws2 = rec[4][slot] * w;
            float ws3 = rec[5][slot] * w, wr = w;
            if (w < 0.0f) {
              float mc = ws0;
              mc = ws1 < mc ? ws1 : mc;
              mc = ws2 < mc ? ws2 : mc;
              if (mc < R.neg_clamp) {
                const float cs = R.neg_clamp / mc;
                wr = wr * cs;
                ws0 = ws0 * cs; ws1 = ws1 * cs; ws2 = ws2 * cs; ws3 = ws3 * cs;
              }
            }
            acc[0] += ws0; acc[1] += ws1; acc[2] += ws2; acc[3] += ws3; acc[4] += wr;
          }
        }
    }
    if (live) {
      float *a = R.accum + 5 * ((size_t)qy * R.width + (size_t)qx);
      a[0] += acc[0]; a[1] += acc[1]; a[2] += acc[2]; a[3] += acc[3]; a[4] += acc[4];
    }
  }
}

// Adds this step's queue lengths (closest-hit and any-hit rays) to the
// device statistics; one wave.
__global__ void __launch_bounds__(64) k_count_queues(uint32_t maxp, PathBufs P,
                                                     unsigned long long *stats) {
  uint32_t b = threadIdx.x;
  uint32_t q = b <= maxp ? P.counts[SGR_CNT_Q + b] : 0u;
  uint32_t s = b <= maxp ? P.counts[SGR_CNT_S + b] : 0u;
  wave_count(&stats[SGR_STAT_ISECT], q);
  wave_count(&stats[SGR_STAT_SHADOW], s);
}

// ---------------------------------------------------------------------------
// Host-callable launchers
// ---------------------------------------------------------------------------
// Compute units of the device the launches go to.  The per-kernel caps
// below are resident workgroups per CU at each kernel's occupancy (set by
// its register use): a launch of that many blocks fills the chip once and
// each block loops over its range.  They apply whatever grid the caller
// asks for; the caller's grid sets raygen, the refill walks and the
// compaction-free kernels.
static unsigned s_num_cu = 256;
void set_launch_cus(unsigned num_cu) { s_num_cu = num_cu ? num_cu : 256u; }
unsigned launch_cus() { return s_num_cu; }

static inline unsigned grid_for(uint32_t n, unsigned cap) {
  unsigned g = (n + 255) / 256;
  if (g == 0) g = 1;
  return g < cap ? g : cap;
}

// Walks whose scenes hold instances, ellipses or local primitives, or
// (any-hit walks) partial occluders, take the INST kernels: the plain
// walk with the second stack for the models, the tests of the rarer
// surfaces and the shadow rays' transmittance.
static inline bool general_walk(const SceneView &S, bool any) {
  return S.n_insts || S.has_ext || (any && S.has_partial);
}

void launch_trace(const SceneView &S, const RayBatchDev &rb, uint32_t n, bool any,
                  unsigned long long *stats, hipStream_t st) {
  unsigned g = grid_for(n, 65535);
  g = (g + 7) & ~7u;  // whole blocks per XCD slice
  if (general_walk(S, any)) {  // plain walk, second stack for the models
    if (any) hipLaunchKernelGGL((k_trace<true, false, true>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
    else hipLaunchKernelGGL((k_trace<false, false, true>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
  } else if (S.walk_refill) {
    if (any) hipLaunchKernelGGL((k_trace<true, true, false>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
    else hipLaunchKernelGGL((k_trace<false, true, false>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
  } else {
    if (any) hipLaunchKernelGGL((k_trace<true, false, false>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
    else hipLaunchKernelGGL((k_trace<false, false, false>), dim3(g), dim3(256), 0, st, S, rb, n, stats);
  }
}

void launch_raygen(bool multi, const RenderView &R, const StepView &J, const PathBufs &P,
                   unsigned grid, hipStream_t st) {
  if (multi) hipLaunchKernelGGL(k_raygen<true>, dim3(grid), dim3(256), 0, st, R, J, P);
  else hipLaunchKernelGGL(k_raygen<false>, dim3(grid), dim3(256), 0, st, R, J, P);
}
// Grid of the plain walk on a shallow octree: at most the blocks six
// waves per SIMD keep resident (24 KB of LDS stack each, 6 per CU), so
// every block starts at once and loops over its range: more generations
// of short-lived blocks cost more than the walks (Cornell box 256x256
// frame 4.45 -> 3.40 ms).  On deep octrees the camera rays' plain walk
// keeps the caller's grid, whose later blocks
// even out the camera rays' uneven ranges (1.31 M-triangle frame 119.3 ->
// 120.7 ms with the smaller grid).
static inline unsigned plain_walk_grid(unsigned grid) {
  const unsigned cap = (6u * s_num_cu + 7u) & ~7u;
  return grid < cap ? grid : cap;
}

void launch_extend(bool multi, const SceneView &S, const StepView &J, const PathBufs &P,
                   uint32_t b, unsigned long long *stats, unsigned grid, hipStream_t st) {
  // Camera rays take the plain walk even where later bounces refill: a
  // wave holds the samples of one or two pixels, whose walks stay in step
  // and share their node and surface loads, so the refill's balancing
  // buys nothing there (the two tie on the 1.31 M-triangle frame) and
  // its atomics are saved; for bounce 1 and for shadow rays, whose
  // directions are random, the plain walk loses 3-9 %.
  const bool refill = S.walk_refill && b > 0;
  if (!S.walk_refill) grid = plain_walk_grid(grid);
  const bool gen = general_walk(S, false);  // the plain walk
  if (gen && multi)
    hipLaunchKernelGGL((k_extend<true, false, true>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (gen)
    hipLaunchKernelGGL((k_extend<false, false, true>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (multi && refill)
    hipLaunchKernelGGL((k_extend<true, true, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (multi)
    hipLaunchKernelGGL((k_extend<true, false, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (refill)
    hipLaunchKernelGGL((k_extend<false, true, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else
    hipLaunchKernelGGL((k_extend<false, false, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
}
void launch_direct_integ(const SceneView &S, const RenderView &R, const PathBufs &P,
                         unsigned long long *stats, unsigned grid, hipStream_t st) {
  const bool gen = general_walk(S, true), ph = R.integ == SGR_INTEG_PHOTON;
  if (gen && ph)
    hipLaunchKernelGGL((k_direct_integ<true, true>), dim3(grid), dim3(256), 0, st, S, R, P, stats);
  else if (ph)
    hipLaunchKernelGGL((k_direct_integ<false, true>), dim3(grid), dim3(256), 0, st, S, R, P, stats);
  else if (gen)
    hipLaunchKernelGGL((k_direct_integ<true, false>), dim3(grid), dim3(256), 0, st, S, R, P, stats);
  else
    hipLaunchKernelGGL((k_direct_integ<false, false>), dim3(grid), dim3(256), 0, st, S, R, P, stats);
}
void launch_photon_shoot(const SceneView &S, const PhotonShootView &V, unsigned long long *stats,
                         unsigned grid, hipStream_t st) {
  if (general_walk(S, false))
    hipLaunchKernelGGL((k_photon_shoot<true>), dim3(grid), dim3(256), 0, st, S, V, stats);
  else
    hipLaunchKernelGGL((k_photon_shoot<false>), dim3(grid), dim3(256), 0, st, S, V, stats);
}
void launch_shade(bool multi, const SceneView &S, const RenderView &R, const StepView &J,
                  const PathBufs &P, uint32_t b, unsigned long long *stats, unsigned grid,
                  hipStream_t st, const GuideView *G) {
  GuideView g0;
  if (!G) {
    std::memset(&g0, 0, sizeof g0);
    g0.rec_stride = 1;
  }
  const GuideView &g = G ? *G : g0;
  // This kernel holds 2 waves per SIMD: 512 blocks of 4 waves fill the
  // 256 CUs once, each wave looping over its block's range, and a launch
  // with a short queue (or none left) dispatches an eighth of the blocks
  // (grids of 4096 / 1024 / 512 / 256 blocks: Cornell 256x256 frame 4.94
  // / 4.53 / 4.42 / 4.55 ms, 1.31 M-triangle frame 121.8 / 120.8 / 119.1
  // / 127.2 ms).
  grid = grid < 2u * s_num_cu ? grid : 2u * s_num_cu;
  PathBufs Pk = P;
  Pk.media_ovf = stats + SGR_STAT_MEDIA_OVF;
  if (multi && S.has_ext)
    hipLaunchKernelGGL((k_shade<true, false, true>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  else if (multi)
    hipLaunchKernelGGL((k_shade<true, false, false>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  else if (G && S.has_ext)
    hipLaunchKernelGGL((k_shade<false, true, true>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  else if (G)
    hipLaunchKernelGGL((k_shade<false, true, false>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  else if (S.has_ext)
    hipLaunchKernelGGL((k_shade<false, false, true>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  else
    hipLaunchKernelGGL((k_shade<false, false, false>), dim3(grid), dim3(256), 0, st, S, R, J, Pk, b, stats, g);
  // the next ray queue and this bounce's shadow queue, in order
  const unsigned nblk = (P.n + SGR_COMPACT_ITEMS - 1) / SGR_COMPACT_ITEMS;
  hipLaunchKernelGGL(k_compact_count, dim3(nblk), dim3(256), 0, st, P, b);
  hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(256), 0, st, P, b);
  hipLaunchKernelGGL(k_compact_scatter, dim3(nblk), dim3(256), 0, st, P, b);
}
void launch_occlude(bool multi, const SceneView &S, const StepView &J, const PathBufs &P,
                    uint32_t b, unsigned long long *stats, unsigned grid, hipStream_t st) {
  if (!S.walk_refill) grid = plain_walk_grid(grid);
  const bool gen = general_walk(S, true);
  if (gen && multi)
    hipLaunchKernelGGL((k_occlude<true, false, true>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (gen)
    hipLaunchKernelGGL((k_occlude<false, false, true>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (multi && S.walk_refill)
    hipLaunchKernelGGL((k_occlude<true, true, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (multi)
    hipLaunchKernelGGL((k_occlude<true, false, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else if (S.walk_refill)
    hipLaunchKernelGGL((k_occlude<false, true, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
  else
    hipLaunchKernelGGL((k_occlude<false, false, false>), dim3(grid), dim3(256), 0, st, S, J, P, b, stats);
}
void launch_gather(bool multi, const RenderView &R, const StepView &J, const PathBufs &P,
                   uint32_t b, unsigned grid, hipStream_t st, const GuideView *G) {
  // 8 waves per SIMD keep 2048 blocks resident: one generation of blocks
  // (1.31 M-triangle frame 119.6 -> 119.0 ms against 4096)
  grid = grid < 8u * s_num_cu ? grid : 8u * s_num_cu;
  GuideView g0;
  if (!G) {
    std::memset(&g0, 0, sizeof g0);
    g0.rec_stride = 1;
  }
  const GuideView &g = G ? *G : g0;
  if (multi)
    hipLaunchKernelGGL((k_gather<true, false>), dim3(grid), dim3(256), 0, st, R, J, P, b, g);
  else if (G)
    hipLaunchKernelGGL((k_gather<false, true>), dim3(grid), dim3(256), 0, st, R, J, P, b, g);
  else
    hipLaunchKernelGGL((k_gather<false, false>), dim3(grid), dim3(256), 0, st, R, J, P, b, g);
}
void launch_splat(bool multi, const RenderView &R, const StepView &J, const float *out,
                  unsigned grid, hipStream_t st) {
  if (multi) hipLaunchKernelGGL(k_splat<true>, dim3(grid), dim3(256), 0, st, R, J, out);
  else hipLaunchKernelGGL(k_splat<false>, dim3(grid), dim3(256), 0, st, R, J, out);
}
void launch_finish(bool multi, const SceneView &S, const RenderView &R, const StepView &J,
                   const PathBufs &P, uint32_t b, unsigned long long *stats, unsigned grid,
                   hipStream_t st) {
  // A queue it takes has under SGR_FINISH_MAX rays: one per thread at most
  // needs no more blocks, and a launch that finds its queue too long (or
  // already finished) then dispatches a quarter of the blocks of the
  // per-bounce kernels (this kernel holds 2 waves per SIMD).
  unsigned g = SGR_FINISH_MAX / 256u;
  g = g < 4u * s_num_cu ? g : 4u * s_num_cu;
  g = grid < g ? grid : g;
  PathBufs Pk = P;
  Pk.media_ovf = stats + SGR_STAT_MEDIA_OVF;
  if (multi && general_walk(S, true))
    hipLaunchKernelGGL((k_finish<true, true>), dim3(g), dim3(256), 0, st, S, R, J, Pk, b, stats);
  else if (multi)
    hipLaunchKernelGGL((k_finish<true, false>), dim3(g), dim3(256), 0, st, S, R, J, Pk, b, stats);
  else if (general_walk(S, true))
    hipLaunchKernelGGL((k_finish<false, true>), dim3(g), dim3(256), 0, st, S, R, J, Pk, b, stats);
  else
    hipLaunchKernelGGL((k_finish<false, false>), dim3(g), dim3(256), 0, st, S, R, J, Pk, b, stats);
}

void launch_count_queues(uint32_t maxp, const PathBufs &P, unsigned long long *stats,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_count_queues, dim3(1), dim3(64), 0, st, maxp, P, stats);
}

}  // namespace sgr
