/*
 * snoggpu.h -- C ABI of the MI355X (gfx950) path-tracing core.
 *
 * Plain C types only: pointers, sizes, fixed-width integers and floats.
 * Every call returns an int status (SGR_OK == 0); on failure the message
 * is available from sgr_last_error() (thread-local, valid until the next
 * failing call on the same thread).
 *
 * The entry points replace, one for one, the reference's CPU render path:
 *
 *   sgr_scene_create      Scene::Scene + SpaceBuilderFactory::make_space
 *                         (render/scene.cc:26-44, space/space-builder.h:86-93,
 *                          space/octree-builder.cc:174-677) and the light
 *                          sampler set-up (render/scene.cc:34-44)
 *   sgr_trace_closest     Space::intersect (space/space.cc:55-66)
 *   sgr_trace_occluded    Space::occludes (space/space.cc:182-194)
 *   sgr_render            RenderMgr::render (render-mgr/render-mgr.cc:49-74)
 *                         driving Renderer::render_packet
 *                         (render-mgr/renderer.cc:41-100), PathInteg::Li
 *                         (render/path-integ.cc:133-407) and the
 *                         ImageSampledOutput filtered splat
 *                         (image/image-filter-conv.h:109-186)
 *   sgr_finalize_image    ImageSampledOutput::set_raw_min_y
 *                         (image/image-sampled-output.cc:63-126)
 *   sgr_scene_octree_stats Octree::stats (space/octree.cc:308-367)
 *
 * Threading: one sgr_ctx per host thread; calls on one context are
 * serialised by the caller.  Scenes are immutable after creation and may
 * be shared by renders on the same context.
 */
#ifndef SNOGGPU_H
#define SNOGGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGR_OK            0
#define SGR_ERR_INVALID   1   /* bad argument or unsupported scene feature */
#define SGR_ERR_HIP       2   /* HIP runtime error */
#define SGR_ERR_NOMEM     3   /* device allocation failed / over budget */
#define SGR_ERR_LIMIT     4   /* a documented size limit was exceeded */
#define SGR_ERR_NOTREADY  5   /* job not finished yet (sgr_job_poll) */

/* Material kinds (material/lambert.cc, material/cook-torrance.cc,
 * material/glass.cc, material/mirror.cc, material/glow.cc). */
#define SGR_MAT_LAMBERT        1
#define SGR_MAT_COOK_TORRANCE  2
#define SGR_MAT_GLASS          3
#define SGR_MAT_MIRROR         4
#define SGR_MAT_GLOW           5
#define SGR_MAT_PHONG          6
#define SGR_MAT_THIN_GLASS     7
#define SGR_MAT_STENCIL        8

/* Acceleration structure (render/global-render-state.cc:63-74 "accel"). */
#define SGR_ACCEL_OCTREE  0
#define SGR_ACCEL_TRIV    1   /* brute force, TrivSpace (space/triv-space.h) */

typedef struct sgr_ctx sgr_ctx;
typedef struct sgr_scene sgr_scene;
typedef struct sgr_photon_map sgr_photon_map;

/* One material.  Field meaning per kind:
 *   LAMBERT        color = reflectance
 *   COOK_TORRANCE  color = diffuse colour, gloss_color = specular colour,
 *                  m = Ward roughness, ior_n/ior_k = complex IOR
 *   GLASS          ior_n = IOR of the enclosed medium (no absorption)
 *   MIRROR         gloss_color = reflectance, ior_n/ior_k = IOR;
 *                  underlying = index of the material under the mirror
 *                  (a Lambert, Cook-Torrance or Phong, through glows;
 *                  material/mirror.h:37-42), or -1: then color = the
 *                  underlying Lambert colour (all zero: none)
 *   GLOW           color = emitted radiance; underlying = index of the
 *                  material supplying the BSDF, or -1 for none
 *   PHONG          color = diffuse colour, gloss_color = specular colour,
 *                  m = Phong exponent (material/phong.h:26-37)
 *   THIN_GLASS     color = transmitted colour, ior_n = IOR
 *                  (material/thin-glass.h:30-32, ThinGlass (col, ior)):
 *                  specular reflection and straight-through transmission,
 *                  and shadow rays pass it, multiplied by its
 *                  transmittance; not allowed in a model
 *   STENCIL        color = opacity, underlying = index of the material
 *                  supplying the BSDF (not a stencil; material/stencil.h:
 *                  30-37): the underlying BSDF with probability
 *                  intensity(opacity), else straight through; shadow rays
 *                  pass it with transmittance max(1 - opacity, 0); not
 *                  allowed in a model                                  */
typedef struct sgr_material {
    int32_t type;
    float   color[3];
    float   gloss_color[3];
    float   m;
    float   ior_n;
    float   ior_k;
    int32_t underlying;
} sgr_material;

/* A triangle mesh (surface/mesh.h).  Vertices are float xyz; triangles
 * are vertex-index triples.  tri_materials may be NULL, then every
 * triangle uses `material`.  axis is the mesh axis used for tangent
 * frames (surface/mesh.cc:563-600); left_handed selects the normal
 * convention (surface/mesh.h:267-271).  normals: per-vertex normals
 * (float xyz, num_vertices of them; surface/mesh.h:144-149) or NULL;
 * with them the shading frame follows the barycentric interpolation of
 * the three vertex normals and the geometric normal is flipped to its
 * side (surface/mesh.cc:620-641). */
typedef struct sgr_mesh {
    const float    *vertices;
    uint32_t        num_vertices;
    const uint32_t *triangles;
    uint32_t        num_triangles;
    const uint32_t *tri_materials;
    uint32_t        material;
    float           axis[3];
    int32_t         left_handed;
    const float    *normals;
} sgr_mesh;

/* Triangle or parallelogram primitive (surface/tripar.h). */
typedef struct sgr_tripar {
    float    v0[3];
    float    e1[3];
    float    e2[3];
    int32_t  parallelogram;
    uint32_t material;
} sgr_tripar;

/* Sphere primitive (surface/sphere.h). */
typedef struct sgr_sphere {
    float    center[3];
    float    radius;
    uint32_t material;
} sgr_sphere;

/* Ellipse primitive (surface/ellipse.h:38-45): centre and two radius
 * vectors, lying in the plane through center, center + radius1 and
 * center + radius2 (non-orthogonal radii give a skewed ellipse).  It is
 * the parallelogram with corner center - radius1 - radius2 and edges
 * 2 radius1, 2 radius2 cut to its inscribed ellipse; its normal is
 * unit(radius2 x radius1). */
typedef struct sgr_ellipse {
    float    center[3];
    float    radius1[3];
    float    radius2[3];
    uint32_t material;
} sgr_ellipse;

/* Primitives defined as a canonical shape placed by a transform
 * (surface/local-primitive.h): local_to_world is the 3 x 4 row-major
 * affine transform (rows x, y, z: linear part, then translation); the
 * library keeps its inverse (geometry/local-xform.cc:17-21).
 *   SGR_LOCAL_CYLINDER  radius 1 around the z axis, z from -1 to 1, no
 *                       caps (surface/cylinder.cc:114-150)
 *   SGR_LOCAL_SPHERE2   the unit sphere at the origin (surface/sphere2.h)
 * Their box is the local cube [-1,1]^3 with its 8 corners taken to world
 * space (geometry/local-xform.cc:27-31). */
#define SGR_LOCAL_CYLINDER 1
#define SGR_LOCAL_SPHERE2  2
typedef struct sgr_local_prim {
    int32_t  kind;
    float    local_to_world[12];
    uint32_t material;
} sgr_local_prim;

/* Point light (light/point-light.h:25-47, light/light.lua:48-56): radiates
 * `intensity` from `pos`.  angle <= 0: in all directions (the two-argument
 * constructor).  angle > 0: a spotlight, a cone of apex `angle` (radians)
 * about `dir` (used as given, not normalised: it is the frame's z axis)
 * whose outer `fringe_angle` ramps the intensity down as d^4 toward the
 * cone's edge.  A point light is never hit by a ray: direct lighting
 * takes its light sample without multiple-importance weighting and no
 * BSDF sample (render/direct-illum.cc:174, 202).  Scenes with point
 * lights take the general kernels. */
typedef struct sgr_point_light {
    float pos[3];
    float intensity[3];
    float angle;
    float dir[3];
    float fringe_angle;
} sgr_point_light;

/* Lat-long environment map light (light/envmap-light.cc,
 * texture/spheremap.h).  texels: RGB float, row-major, row 0 = top
 * (north).  frame: 3x3 row-major, rows are the light frame's x, y, z
 * axes in world space (identity = map z axis is world z). */
typedef struct sgr_envmap {
    const float *texels;
    uint32_t     width;
    uint32_t     height;
    float        frame[9];
} sgr_envmap;

/* A model: geometry with its own acceleration structure in its own
 * coordinate system, placed in the scene by instances (surface/model.h,
 * surface/model.cc:26-47).  Material indices refer to the scene's table.
 * A model holds no instances and supplies no light samplers (an emitting
 * surface in it glows where a path hits it). */
typedef struct sgr_model {
    const sgr_mesh     *meshes;
    uint32_t            num_meshes;
    const sgr_tripar   *tripars;
    uint32_t            num_tripars;
    const sgr_sphere   *spheres;
    uint32_t            num_spheres;
    const sgr_ellipse  *ellipses;
    uint32_t            num_ellipses;
    const sgr_local_prim *local_prims;
    uint32_t            num_local_prims;
} sgr_model;

/* An instance of a model (surface/instance.cc:145-216): local_to_world is
 * the 3 x 4 row-major affine transform (rows x, y, z: linear part, then
 * translation).  A ray is intersected with the model in model space
 * (the inverse transform, computed by the library); the shading frame is
 * mapped back with the transform (the normal with the inverse
 * transpose). */
typedef struct sgr_instance {
    uint32_t model;
    float    local_to_world[12];
} sgr_instance;

/* Whole scene.  Surfaces are inserted into the acceleration structure in
 * this order: meshes (triangle order), then tripars, then ellipses, then
 * spheres, then local primitives, then instances.  The primitive id of a
 * hit is its index in that order; for a surface of an instanced model it
 * is the index in the model's own order (meshes, tripars, ellipses,
 * spheres, local primitives) and the hit also names the instance.
 * Light samplers are created for GLOW tripars, ellipses, spheres and
 * local primitives (insertion order), then for each envmap. */
typedef struct sgr_scene_desc {
    const sgr_mesh     *meshes;
    uint32_t            num_meshes;
    const sgr_tripar   *tripars;
    uint32_t            num_tripars;
    const sgr_sphere   *spheres;
    uint32_t            num_spheres;
    const sgr_material *materials;
    uint32_t            num_materials;
    const sgr_envmap   *envmaps;
    uint32_t            num_envmaps;
    float               min_trace;   /* render/render-params.h:32-36 */
    int32_t             accel;       /* SGR_ACCEL_* */
    const sgr_model    *models;
    uint32_t            num_models;
    const sgr_instance *instances;
    uint32_t            num_instances;
    const sgr_ellipse  *ellipses;
    uint32_t            num_ellipses;
    const sgr_local_prim *local_prims;
    uint32_t            num_local_prims;
    /* lights after the surface lights and the envmap, in this order
     * (surface/surface-group.cc:82-94 adds surfaces' lights before the
     * group's own, those in the order added) */
    const sgr_point_light *point_lights;
    uint32_t            num_point_lights;
} sgr_scene_desc;

/* Rays and hits, structure-of-arrays (geometry/ray.h:84-89).  Pointers
 * are host memory unless the call says otherwise. */
typedef struct sgr_rays {
    const float *ox, *oy, *oz;
    const float *dx, *dy, *dz;
    const float *t0, *t1;
} sgr_rays;

typedef struct sgr_hits {
    uint32_t *prim;   /* 0xffffffff = no hit */
    float    *t;
    float    *u;
    float    *v;
    uint32_t *inst;   /* instance of the hit (0xffffffff: none); may be NULL */
} sgr_hits;

/* Pinhole camera (camera/camera.h:322-389): eye_vec = forward
 * + right*(2x-1)*tan_half_fov_x + up*(2y-1)*tan_half_fov_y, not
 * normalised. */
typedef struct sgr_camera {
    float pos[3];
    float forward[3];
    float up[3];
    float right[3];
    float tan_half_fov_x;
    float tan_half_fov_y;
} sgr_camera;

/* Render parameters (render/path-integ.cc:61-91, image/image-filter.cc,
 * image/image-mitchell-filt.h:31-33, image/image-filter-conv.h:35). */
typedef struct sgr_render_params {
    uint32_t width, height;        /* full image size */
    uint32_t spp;                  /* "samples" */
    uint32_t min_path_len;         /* default 3 */
    uint32_t max_path_len;         /* default 25 */
    uint32_t direct_samples;       /* default 1; 0 disables direct light */
    uint64_t seed;                 /* job seed for the counter RNG */
    float    filter_b, filter_c;   /* Mitchell B, C (1/3, 1/3) */
    float    filter_radius;        /* 2 (at most 12.5) */
    float    neg_clamp;            /* -0.1 */
    float    bg_alpha;             /* "background_alpha", 1 */
    /* Rendered window [x0, x0+w) x [y0, y0+h) of the image; the filter
     * margin (ceil(radius - 0.5001) pixels, snogray.lua:546-550) is added
     * around it by the renderer, and only the window's accumulators
     * receive samples (a limited render's output is the window,
     * snogray.lua:447-454). w == 0 means the whole image. */
    uint32_t x0, y0, w, h;
    /* Paths per chunk (0: the default below): a frame runs in chunks of
     * whole-image x floor(this / pixels) samples.  sgr_render keeps two
     * chunks in flight on two streams when a frame has several (path
     * buffers for both), so that one chunk's sparse late bounces overlap
     * the next chunk's full early ones; results do not depend on it. */
    uint32_t paths_in_flight;
    uint32_t flags;                /* SGR_RENDER_* */
    /* Surface integrator (render/render.lua "surface_integ"): the path
     * tracer (render/path-integ.cc) or direct lighting with perfectly
     * specular recursion (render/direct-integ.h:31-67 over
     * render/recursive-integ.cc:29-216), whose light sampling takes
     * light_samples samples per light (0: 16, the reference's default,
     * render/direct-integ.h:72-77).  The direct integrator runs in
     * sgr_render / sgr_render_samples / sgr_render_noalloc; the job
     * engine and guided renders refuse it. */
    uint32_t integrator;           /* SGR_INTEG_* */
    uint32_t light_samples;
    /* SGR_INTEG_PHOTON: the photon maps of this scene the render reads
     * (sgr_photon_map_create); NULL otherwise */
    const sgr_photon_map *photon_map;
} sgr_render_params;

#define SGR_INTEG_PATH    0u
#define SGR_INTEG_DIRECT  1u
/* Photon mapping (render/photon-integ.cc): the direct integrator's
 * recursion whose Lo adds, besides direct lighting, caustic photons and
 * final gathering over the photon maps params->photon_map holds.  Runs in
 * sgr_render / sgr_render_samples / sgr_render_noalloc; the job engine
 * and guided renders refuse it.  Its direct lighting takes the map's
 * direct_samples per light (light_samples is not read). */
#define SGR_INTEG_PHOTON  2u
/* recursion levels the direct integrator keeps per sample; deeper
 * specular branches (reached with probability 2^-26 or less through
 * its roulette) are dropped and counted in sgr_stats.direct_overflows */
#define SGR_DIRECT_MAX_DEPTH 32u

#define SGR_DEFAULT_PATHS_IN_FLIGHT (1u << 26)

/* sgr_render_params.flags: time every kernel launch with HIP events on
 * the render stream (no host waits) into sgr_stats.ms_* */
#define SGR_RENDER_KERNEL_TIMING 1u
/* ... run the chunks one after another on one stream (one chunk of path
 * buffers); the image is the same bit for bit */
#define SGR_RENDER_ONE_CHUNK_IN_FLIGHT 2u
/* ... run every bounce through the per-bounce kernels: without it, once
 * a late bounce's ray queue is short, one kernel carries each remaining
 * path through all its remaining bounces (same image bit for bit) */
#define SGR_RENDER_WAVEFRONT_ONLY 4u

/* Counters named after RenderStats::IsecStats (render/render-stats.h:30-52)
 * plus device timings. */
typedef struct sgr_stats {
    uint64_t eye_rays;
    uint64_t scene_intersect_calls;       /* closest-hit rays */
    uint64_t shadow_tests;                /* any-hit rays */
    uint64_t node_visits_closest;         /* octree nodes entered */
    uint64_t surf_tests_closest;          /* primitive tests */
    uint64_t node_visits_shadow;
    uint64_t surf_tests_shadow;
    uint64_t illum_calls;
    uint64_t samples;                     /* nominal W*H*spp */
    uint64_t paths_traced;                /* incl. filter margin */
    double   ms_render;                   /* device time of the render */
    double   ms_extend;                   /* closest-hit kernels */
    double   ms_shade;
    double   ms_shadow;                   /* any-hit kernels */
    double   ms_splat;
    double   ms_other;
    uint32_t extend_launches;
    uint32_t shadow_launches;
    /* transmissions into a medium nested deeper than 8 levels: the
     * device keeps 8 media per path (the reference's stack is unbounded,
     * material/media.cc:75-118) and does not push these */
    uint64_t media_overflows;
    /* octree node records loaded (node_visits_* also count the children
     * the reference calls and prunes at once, which load nothing) */
    uint64_t node_fetches_closest;
    uint64_t node_fetches_shadow;
    uint64_t direct_overflows;            /* see SGR_DIRECT_MAX_DEPTH */
    /* partially occluding surfaces (thin glass, stencil) beyond the 8 a shadow ray
     * remembers: multiplied into its transmittance without checking for
     * a repeat (each surface once is the rule, material/material.h:168) */
    uint64_t partial_overflows;
} sgr_stats;

/* Octree statistics (space/octree.h:53-67). */
typedef struct sgr_octree_stats {
    uint64_t num_nodes;
    uint64_t num_leaf_nodes;
    uint64_t num_surfaces;
    uint64_t num_dup_surfaces;
    uint32_t max_depth;
    float    avg_depth;
} sgr_octree_stats;

/* ---- context ---------------------------------------------------------- */
const char *sgr_last_error(void);
int sgr_ctx_create(int device, sgr_ctx **out);
int sgr_ctx_destroy(sgr_ctx *ctx);
/* Device memory currently held by the context (scenes + buffers). */
int sgr_ctx_memory(sgr_ctx *ctx, uint64_t *bytes);

/* ---- scenes ----------------------------------------------------------- */
int sgr_scene_create(sgr_ctx *ctx, const sgr_scene_desc *desc,
                     sgr_scene **out);
int sgr_scene_destroy(sgr_scene *scene);
int sgr_scene_octree_stats(const sgr_scene *scene, sgr_octree_stats *out);
/* Flattened octree, for parity tests: nodes are 4 x uint32 (first child,
 * child mask, first surface-list entry, surface count) in breadth-first
 * order; surf_ids are global primitive ids.  Pass NULL buffers to query
 * the sizes. */
int sgr_octree_export(const sgr_scene *scene, uint32_t *nodes,
                      uint64_t *num_nodes, uint32_t *surf_ids,
                      uint64_t *num_surf_ids, float *origin_size4);
/* Host-only octree build of a descriptor (no device needed); same
 * output as sgr_octree_export plus Octree::stats.  Any pointer may be
 * NULL.  Replaces Octree::Builder::add + make_space
 * (space/octree-builder.cc:174-677) for inspection. */
int sgr_octree_build(const sgr_scene_desc *desc, uint32_t *nodes,
                     uint64_t *num_nodes, uint32_t *surf_ids,
                     uint64_t *num_surf_ids, float *origin_size4,
                     sgr_octree_stats *stats);
/* Horizon (scene bbox diameter, render/scene.cc:28) and scene bbox. */
int sgr_scene_bounds(const sgr_scene *scene, float *bbox_min3,
                     float *bbox_max3, float *horizon);
/* Device memory footprint of the scene in bytes. */
int sgr_scene_footprint(const sgr_scene *scene, uint64_t *bytes);
/* The local_to_world transform of a cylinder from its base, axis and
 * radius (host only; Cylinder::xform, surface/cylinder.cc:24-36): the
 * canonical cylinder moved up by 1, scaled by (radius, radius,
 * |axis| / 2), turned to the basis (p, unit(axis) x ... ) built from
 * p = unit(axis).perpendicular() (geometry/vec.h:119-129, not normalised,
 * so the radius is scaled by |p| as in the reference), then moved to
 * base.  Fails on a zero axis. */
int sgr_cylinder_xform(const float base[3], const float axis[3], float radius,
                       float local_to_world[12]);

/* ---- ray batches (parity entry points) -------------------------------- */
int sgr_trace_closest(sgr_scene *scene, const sgr_rays *rays, size_t n,
                      sgr_hits *hits, sgr_stats *stats);
/* occluded[i] = 1 if some surface blocks ray i within (t0, t1). */
int sgr_trace_occluded(sgr_scene *scene, const sgr_rays *rays, size_t n,
                       uint8_t *occluded, sgr_stats *stats);

/* ---- rendering -------------------------------------------------------- */
/* Render into caller-owned accumulators: accum has 5 floats per image
 * pixel (premultiplied r, g, b, alpha, then filter weight), row-major,
 * width*height*5.  accum_on_device != 0: accum is a device pointer of the
 * context's device; else host memory.  Accumulators are ADDED to (clear
 * them first for a fresh image). */
int sgr_render(sgr_scene *scene, const sgr_camera *cam,
               const sgr_render_params *params, float *accum,
               int accum_on_device, sgr_stats *stats);
/* Per-sample output for parity tests: for every rendered path (window
 * plus margin, row-major over pixels, then sample index) writes
 * sx, sy, r, g, b, alpha (6 floats).  out must hold
 * (w+2m)*(h+2m)*spp*6 floats (host memory). */
int sgr_render_samples(sgr_scene *scene, const sgr_camera *cam,
                       const sgr_render_params *params, float *out,
                       sgr_stats *stats);
/* sgr_render that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  accum
 * is device memory of the context's device (added to).  The path state of
 * the render lives in a fixed scratch array of SGR_NOALLOC_SCRATCH_BYTES
 * per device (held until the process ends); the frame runs in chunks of
 * as many samples per pattern pixel as fit there (at most the chunk
 * params->paths_in_flight gives), one chunk in flight, so the image is the
 * one sgr_render gives with that chunk size.  A window whose pattern
 * pixels (window plus filter margin) need more than the scratch for one
 * sample each fails with SGR_ERR_LIMIT.  Every call on a device shares
 * that device's scratch: calls that overlap on one device race (one
 * context per device, calls serialised, as for every entry point).
 * SGR_RENDER_KERNEL_TIMING is ignored. */
#define SGR_NOALLOC_SCRATCH_BYTES (32u << 20)
int sgr_render_noalloc(sgr_scene *scene, const sgr_camera *cam,
                       const sgr_render_params *params, float *accum_dev,
                       sgr_stats *stats);
/* Normalise accumulators into an RGBA image (4 floats per pixel), as
 * ImageSampledOutput::set_raw_min_y does: divide by weight, clamp alpha,
 * clamp colour at 0, un-premultiply, exposure, contrast. */
int sgr_finalize_image(const float *accum, uint32_t width, uint32_t height,
                       float exposure, float contrast, float *rgba);

/* ---- photon maps (render/photon-integ.cc:36-242) -----------------------
 * Photons are shot from the scene's lights on the device
 * (PhotonShooter::shoot, photon/photon-shooter.cc:29-191): path n picks
 * light floor(radical_inverse(n, 11) * lights) and its free sample at
 * (radical_inverse(n, 2), (n, 3)) / ((n, 5), (n, 7)), scatters by BSDF
 * samples ((n, 13), (n, 17) at the first hit, then the path's own
 * counter stream seeded by `seed`; Russian roulette 1/2 past 3 bounces),
 * and deposits a photon at every hit whose BSDF has a non-specular part:
 * into the direct set before any scattering, the caustic set after
 * specular scattering only, else the indirect set, each set until its
 * target count (PhotonInteg::Shooter::deposit, render/photon-integ.cc:
 * 196-226).  The photons of each set are then put in a left-balanced
 * kd-tree (photon/photon-map.cc:22-189).  Parameter defaults are the
 * reference's (render/photon-integ.cc:36-148); a count of
 * SGR_PHOTON_UNSPEC means "not given". */
#define SGR_PHOTON_UNSPEC 0xffffffffu
typedef struct sgr_photon_params {
    uint32_t caustic_photons;     /* "caustic_photons": 50000 */
    uint32_t direct_photons;      /* "direct_photons": 500000; none are
                                     shot with direct lighting on and
                                     final gathering off */
    uint32_t indirect_photons;    /* "indirect_photons": 500000 */
    uint32_t use_photons;         /* photons per estimate, "use_photons": 50 */
    float    photon_radius;       /* search radius, "photon_radius": 0.1 */
    uint32_t final_gather;        /* "final_gather": 1 */
    uint32_t fg_samples;          /* "final_gather_samples" (UNSPEC) */
    uint32_t fg_bsdf_samples;     /* "final_gather_bsdf_samples" (UNSPEC) */
    uint32_t fg_photon_samples;   /* "final_gather_photon_samples" (UNSPEC) */
    uint32_t direct_illum;        /* "direct_illum": 1 (else the direct map) */
    uint32_t direct_samples;      /* "direct_samples": 16 per light */
    uint64_t seed;                /* stream of the photon paths */
} sgr_photon_params;

/* The parameters with the reference's defaults filled in. */
void sgr_photon_params_default(sgr_photon_params *params);

typedef struct sgr_photon_map_info {
    uint64_t photons[3];          /* direct, caustic, indirect set sizes */
    uint64_t paths[3];            /* PhotonSet::num_paths of each set */
    float    scale[3];            /* 1 / paths (0 for an empty set) */
    uint64_t paths_shot;          /* path numbers taken (light samples) */
    uint64_t deposit_overflows;   /* hits past SGR_PHOTON_MAX_DEPOSITS on
                                     one path, not deposited */
    uint32_t fg_samples, fg_bsdf_samples, fg_photon_samples;  /* resolved */
    double   ms_shoot;            /* device time of the shooting */
    double   ms_build;            /* host kd-tree build and upload */
} sgr_photon_map_info;
#define SGR_PHOTON_MAX_DEPOSITS 32u
/* hits a photon path takes at most (the reference has no bound; past 3
 * bounces its roulette ends a path with probability 1/2 at each) */
#define SGR_PHOTON_MAX_BOUNCES 256u

/* Shoot the photons of a scene and build its maps
 * (PhotonInteg::GlobalState::generate_photons). */
int sgr_photon_map_create(sgr_scene *scene, const sgr_photon_params *params,
                          sgr_photon_map **out);
int sgr_photon_map_info_get(const sgr_photon_map *map, sgr_photon_map_info *info);
/* Photons of set (0 direct, 1 caustic, 2 indirect) in kd-tree order:
 * per photon pos xyz + split axis (4 floats), power rgb, dir xyz (the
 * direction it came from); any output may be NULL. */
int sgr_photon_map_export(const sgr_photon_map *map, int set, float *pos_axis,
                          float *power, float *dir);
int sgr_photon_map_destroy(sgr_photon_map *map);

/* ---- image files (host only) ------------------------------------------
 * The formats the reference always supports (image/image-dispatch.cc:
 * 70-157): "pfm" (image/image-pfm.cc) and "rgbe" / "hdr" / "pic"
 * (Radiance, image/image-rgbe.cc).  format == NULL: from the file name's
 * extension (ImageIo::find_format). */
/* Write a finished RGBA image (4 floats per pixel, row-major, row 0 at
 * the top; sgr_finalize_image output): each pixel's alpha-scaled colour,
 * as ImageSink::write_row writes it.  Replaces PfmImageSink /
 * RgbeImageSink (image/image-pfm.cc:58-94, image/image-rgbe.cc:44-150). */
int sgr_image_write(const char *path, const char *format, const float *rgba,
                    uint32_t width, uint32_t height);
/* Read an image as RGB floats (3 per pixel, row 0 at the top).  rgb ==
 * NULL: only *width and *height are set; else they must match the file.
 * Replaces PfmImageSource / RgbeImageSource (image/image-pfm.cc:100-147,
 * image/image-rgbe.cc:156-333). */
int sgr_image_read(const char *path, const char *format, float *rgb,
                   uint32_t *width, uint32_t *height);
/* Recover the rows of a partially written output image into accumulators
 * (5 floats per pixel, width * height): each complete row, from the top,
 * becomes (r, g, b, alpha 1, weight 1); if the file ends or breaks inside
 * a row, the last SGR_RECOVER_DISCARD_ROWS rows read are dropped too.
 * *rows is the number of rows recovered (0 for a missing or unreadable
 * file); a render of the rest of the image continues at that row.
 * Replaces recover_image (image/recover-image.cc:22-124). */
#define SGR_RECOVER_DISCARD_ROWS 4
int sgr_image_recover(const char *path, const char *format, float *accum,
                      uint32_t width, uint32_t height, uint32_t *rows);

/* ---- mesh files (host only) --------------------------------------------
 * The mesh loaders of load/load.lua:405-418: PLY (load/mesh/load-ply.cc:
 * vertex x y z [nx ny nz], triangle faces; ASCII and binary), Wavefront
 * OBJ (load/mesh/obj.lua: v, vn, f with i or i//i indices, polygons
 * fanned into triangles), STL (load/mesh/stl.lua: the ASCII "solid ascii"
 * grammar, equal vertex positions merged), MSH (load/mesh/load-msh.cc:
 * vertex / triangle / texcoord / normal sections) and UG
 * (load/mesh/ug.lua: named vertices, named lambert colours, polygons).
 * format == NULL: from the extension, case-insensitive. */
typedef struct sgr_mesh_file sgr_mesh_file;
/* Parse a file; reports the vertex and triangle counts and whether it
 * carries per-vertex normals (any pointer but path and out may be NULL). */
int sgr_mesh_file_open(const char *path, const char *format, sgr_mesh_file **out,
                       uint64_t *num_vertices, uint64_t *num_triangles,
                       int *has_normals);
/* Copy the arrays (3 floats per vertex, 3 indices per triangle, 3 floats
 * per vertex normal) into caller buffers; NULL skips one. */
int sgr_mesh_file_copy(const sgr_mesh_file *mesh, float *vertices,
                       uint32_t *triangles, float *normals);
/* Handedness and parts.  *left_handed as the reference's loader leaves
 * the mesh: 0 for PLY, OBJ and MSH (they set it), 1 for STL and UG (the
 * Mesh default, surface/mesh.h:55).  Parts in order of first use: a UG
 * file has one per material (load/mesh/ug.lua:75-90), the others one (no
 * triangles: none).  tri_parts gets one part index per triangle,
 * part_colors 3 floats per part: the lambert colour of a UG "c" line, or
 * -1 for the caller's material.  NULL skips an output. */
int sgr_mesh_file_parts(const sgr_mesh_file *mesh, int *left_handed,
                        uint32_t *num_parts, uint32_t *tri_parts,
                        float *part_colors);
int sgr_mesh_file_close(sgr_mesh_file *mesh);

/* ---- job engine: many renders sharing one stream of launches ----------
 * The reference renders one image per RenderMgr::render call
 * (render-mgr/render-mgr.cc:35-63); these calls have no counterpart
 * there.  A job is one sgr_render (same parameters, same result bits:
 * each job keeps the sample chunking sgr_render would use).  An engine
 * step takes the next sample chunk of as many admitted jobs as fit the
 * path capacity and runs them through one raygen/extend/shade/occlude/
 * gather/splat sequence; a job contributes at most one chunk per step.
 * Not thread-safe: one host thread drives a context's engine. */
typedef struct sgr_engine_config {
    uint64_t path_capacity;   /* paths in flight per step (0: the
                                 default SGR_DEFAULT_PATHS_IN_FLIGHT) */
    uint64_t memory_budget;   /* bytes of job accumulators admitted at
                                 once (0: no limit) */
    uint32_t grid_blocks;     /* workgroups of the camera-ray, refill-walk
                                 and splat launches (0: 16 per CU);
                                 rounded up to a multiple of 8.  The
                                 shading, gather, finishing and plain-walk
                                 launches are capped at the workgroups
                                 their occupancy keeps resident (2, 8, 4
                                 and 6 per CU) */
    uint32_t use_graph;       /* replay each step from a captured
                                 hipGraph (one per bounce count) */
    uint32_t finish_max;      /* ray-queue length below which a late
                                 bounce's paths are finished in one launch
                                 (0: 2^18; results do not depend on it) */
} sgr_engine_config;

enum { SGR_JOB_QUEUED = 0, SGR_JOB_ACTIVE = 1, SGR_JOB_DONE = 2 };

typedef struct sgr_job_status {
    int32_t  state;           /* SGR_JOB_* */
    uint32_t chunks_done;
    uint32_t chunks_total;
    uint32_t steps;           /* engine steps that carried the job */
    double   t_submit_ms;     /* host clock, ms since engine creation */
    double   t_start_ms;      /* first step carrying the job began */
    double   t_done_ms;       /* its last step finished */
    uint64_t footprint;       /* bytes counted against memory_budget */
} sgr_job_status;

typedef struct sgr_step_info {
    uint32_t jobs;            /* jobs in the step */
    uint32_t jobs_finished;   /* jobs that completed in it */
    uint64_t paths;
    double   ms_device;       /* device time of the step */
    uint32_t max_path_len;    /* bounces launched */
    uint32_t graph;           /* 1: replayed from a captured graph */
} sgr_step_info;

/* (Re)configure the context's engine; only while it holds no jobs. */
int sgr_engine_configure(sgr_ctx *ctx, const sgr_engine_config *cfg);
/* Bytes a job with these parameters counts against memory_budget
 * (engine-owned accumulators; 0 when the caller supplies them). */
int sgr_job_footprint(const sgr_render_params *params, int caller_accum,
                      uint64_t *bytes);
/* Queue a job.  accum_dev: device accumulators (5 floats per pixel,
 * added to) or NULL for engine-owned zeroed ones.  The scene must stay
 * alive until the job is released (sgr_scene_destroy refuses before). */
int sgr_job_submit(sgr_ctx *ctx, sgr_scene *scene, const sgr_camera *cam,
                   const sgr_render_params *params, float *accum_dev,
                   uint64_t *job_id);
/* Admit queued jobs that fit the budget, then run one step.  info may be
 * NULL; info->jobs == 0 means there was nothing to run. */
int sgr_engine_step(sgr_ctx *ctx, sgr_step_info *info);
int sgr_job_poll(sgr_ctx *ctx, uint64_t job_id, sgr_job_status *status);
/* Copy a finished job's accumulators (width*height*5 floats) to host
 * (dst_on_device == 0) or device memory. */
int sgr_job_result(sgr_ctx *ctx, uint64_t job_id, float *dst, int dst_on_device);
/* Forget a job (any state); frees engine-owned accumulators. */
int sgr_job_release(sgr_ctx *ctx, uint64_t job_id);
/* Engine-wide counters since configure (same meaning as sgr_render's). */
int sgr_engine_stats(sgr_ctx *ctx, sgr_stats *stats);
/* Submit n jobs, step until all are done, copy results, release them.
 * accum[i] is device memory (on_device) or host memory; added to. */
int sgr_render_batch(sgr_ctx *ctx, uint32_t n, sgr_scene *const *scenes,
                     const sgr_camera *cams, const sgr_render_params *params,
                     float *const *accum, int on_device, sgr_stats *stats);

/* ---- learned directional proposal -------------------------------------
 * No counterpart in the reference (its continuation direction is always
 * a BSDF sample, render/path-integ.cc:308-401).  A guide is a small MLP
 * (position in the scene box and outgoing direction, 6-frequency sin/cos
 * encoding, 3 x 64 ReLU, 8 von Mises-Fisher lobes; bf16 weights, fp32
 * accumulation) trained online on the tracer's own radiance samples.
 * A guided render picks each non-delta continuation direction by
 * one-sample MIS: the BSDF with probability 0.5, else the lobes, and
 * divides by the mixture density, so the estimate stays unbiased; glass
 * and mirror vertices sample their BSDF only. */
typedef struct sgr_guide sgr_guide;

typedef struct sgr_guide_stats {
    uint64_t train_steps;     /* Adam steps taken */
    uint64_t records;         /* training records consumed */
    double   last_loss;       /* weighted mean -log q of the last step */
} sgr_guide_stats;

int sgr_guide_create(sgr_ctx *ctx, uint64_t seed, sgr_guide **out);
int sgr_guide_destroy(sgr_guide *guide);
/* fp32 parameters: count = 16624 (W[out][in] row-major then bias, for
 * 80->64, 64->64, 64->64, 64->48; padding entries are zero).  out may be
 * NULL to query the count. */
int sgr_guide_get_params(sgr_guide *guide, float *out, uint64_t *count);
/* Replace the parameters (bf16 copy refreshed, Adam state reset). */
int sgr_guide_set_params(sgr_guide *guide, const float *in, uint64_t count);
/* Forward pass of n <= 4096 rows of 6 inputs (position normalised to the
 * scene box, unit outgoing direction), host memory: 40 outputs per row,
 * raw (lobes == 0) or as lobes (mu xyz, kappa, weight) x 8. */
int sgr_guide_forward(sgr_guide *guide, const float *in6, uint64_t n, float *out,
                      int lobes);
/* One training step on n <= 16384 records of 16 floats (0-5 inputs,
 * 6-8 unit incoming direction, 9 weight >= 0; the rest ignored): the
 * gradient of sum(w * -log q) / sum(w) into grad (16624 floats, may be
 * NULL), loss[0..1] = (sum w * -log q, sum w), and an Adam step with
 * learning rate lr unless lr == 0. */
int sgr_guide_train_batch(sgr_guide *guide, const float *rec, uint64_t n, float lr,
                          float *grad, float *loss);
int sgr_guide_stats_get(const sgr_guide *guide, sgr_guide_stats *stats);
/* sgr_render with the guide; train != 0 takes Adam steps (lr 1e-3) on
 * each chunk's records before the next chunk. */
int sgr_render_guided(sgr_scene *scene, const sgr_camera *cam,
                      const sgr_render_params *params, sgr_guide *guide, int train,
                      float *accum, int accum_on_device, sgr_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* SNOGGPU_H */
