/*
 * vrgi.h — C ABI of the MI355X (gfx950) voxel path-tracing hot path.
 *
 * This ABI replaces the public Rust API of the reference's render_system
 * (pimpale/vulkan-voxel-restir-gi):
 *   Scene<K> / SceneUploader / SceneUploadedObjectHandle  (src/render_system/scene.rs)
 *   Renderer                                               (src/render_system/interactive_rendering.rs)
 *   Vertex3D / LightVertex3D / InstanceData                (src/render_system/vertex.rs)
 *   BvhNode                                                (src/render_system/bvh/mod.rs)
 *   build_bl_bvh / build_tl_bvh                            (src/render_system/bvh/build.rs)
 * Every entry point returns 0 on success or a negative VRGI_E_* code; the
 * message of the last failure on the calling thread is vrgi_last_error().
 * Plain pointers and sizes only: no torch, no HIP types.
 */
#ifndef VRGI_H
#define VRGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VRGI_ABI_VERSION 2u

/* status codes */
#define VRGI_OK 0
#define VRGI_E_INVALID (-1)   /* bad argument (null, size, unknown key, n_verts % 3 != 0) */
#define VRGI_E_DEVICE (-2)    /* HIP runtime failure (allocation, launch) */
#define VRGI_E_STATE (-3)     /* call not valid in the current state */
#define VRGI_E_LIMIT (-4)     /* size beyond a documented maximum */

/* Vertex3D (vertex.rs:3-12): 24 bytes, non-indexed, 3 per triangle. */
typedef struct vrgi_vertex3d {
    float position[3];
    uint32_t t; /* material slot = block * 6 + face (block.rs:116-119) */
    float uv[2];
} vrgi_vertex3d;

/* LightVertex3D (vertex.rs:23-28): 16 bytes, in light BL-BVH leaf order. */
typedef struct vrgi_light_vertex3d {
    float position[3];
    uint32_t bvh_node_idx;
} vrgi_light_vertex3d;

/* BvhNode (bvh/mod.rs:6-18): 40 bytes, scalar layout. left == 0xFFFFFFFF marks a
 * leaf whose right_or_prim holds the primitive id; both 0xFFFFFFFF is the dummy
 * node of an empty BVH (bvh/mod.rs:27-38). */
typedef struct vrgi_bvh_node {
    uint32_t left;
    uint32_t right_or_prim;
    float min[3];
    float power;
    float max[3];
    uint32_t parent;
} vrgi_bvh_node;

/* One material slot = the three constant texels the reference samples per slot
 * (raytrace.rs:465-478): reflectivity rgba (alpha = 1 - transmit probability),
 * emissivity rgb (scaled by 1000 in the shader), metallicity (r). */
typedef struct vrgi_material {
    float reflectivity[4];
    float emissivity[4];
    float metallicity[4];
} vrgi_material;

/* Renderer configuration. vrgi_config_reference() fills the reference's constants
 * (interactive_rendering.rs:766-767, restir_spatial_resampling.rs:313-314). */
typedef struct vrgi_config {
    uint32_t width, height;       /* render extent (Renderer::rebuild) */
    uint32_t num_bounces;         /* B; reference 3 */
    uint32_t rr_start_bounce;     /* Russian roulette from this bounce on; 0xFFFFFFFF = off (reference) */
    float rr_min_survival;        /* lower clamp of the survival probability */
    uint32_t restir_mode;         /* 0 = ReSTIR GI (reference), 1 = ReSTIR DI */
    float restir_spatial_radius;  /* neighbour jitter = radius * (u - 0.5); reference 15 */
    uint32_t di_candidates;       /* DI mode: light candidates per pixel */
    uint32_t seed;                /* salt of the counter-based per-frame seeds */
    uint32_t flags;               /* VRGI_FLAG_* */
    uint32_t ray_sort_bits;       /* 0: queue in compaction order (reference: no reordering);
                                     1..5: reorder each bounce's ray queue by direction octant
                                     and the Morton code of the origin, this many bits per axis */
} vrgi_config;

#define VRGI_FLAG_REBUILD_EVERY_FRAME 1u /* rebuild the geometry BVH on every render() */
#define VRGI_FLAG_KEEP_BOUNCES 2u        /* keep per-bounce buffers readable after render() */
#define VRGI_FLAG_SERIAL_NEE 4u          /* NEE pdf launches after all bounces on the frame stream (default:
                                           each bounce's on a second stream, beside the next bounce's trace) */
#define VRGI_FLAG_SERIAL_BUILD 8u        /* one set of BVH buffers, built on the frame stream (default: two
                                           sets; frame f + 1 builds on a second stream while frame f's rays
                                           still read the other) */
#define VRGI_FLAG_FULL_NEE_PDF 16u       /* compute the NEE pdf of every continued ray (the reference's
                                           nee_pdf buffer). Default: only where it can change the frame's
                                           outputs (the radiance the ray returns is non-zero, or the pdf
                                           might not be finite); elsewhere the buffer holds 0. The outputs
                                           are the same bits either way. */

typedef struct vrgi_camera {
    float eye[3], front[3], right[3], up[3]; /* Camera::eye_front_right_up (camera.rs:115-122) */
} vrgi_camera;

/* RenderingPreferences (camera.rs:37-56). */
typedef struct vrgi_prefs {
    uint32_t nee_type;                  /* 0 none, 1 NEE every bounce, 2 NEE at bounce 1 only */
    uint32_t debug_view;                /* 0 path tracer, 1 ReSTIR, 2 debug (postprocess.rs:39-88);
                                           3 ReSTIR tonemapped (x / (1 + x) per channel); 4 the path
                                           tracer's running mean over the frames since the camera, the
                                           scene, the extent or the view last changed, tonemapped
                                           (no reference counterpart; not in a batch) */
    uint32_t restir_spatial_iterations; /* spatial neighbours, 0..20 */
} vrgi_prefs;

/* Result of one render(). Device pointers stay valid until the next render/resize. */
typedef struct vrgi_frame_out {
    const float *radiance;  /* [H][W][3] plain path tracer, L_out at bounce 0 */
    const float *restir;    /* [H][W][3] ReSTIR final radiance */
    const uint8_t *bgra8;   /* [H][W][4] postprocessed view */
    uint64_t frame_index;   /* frame counter of this frame (ReSTIR slot = frame_index % 3) */
} vrgi_frame_out;

typedef struct vrgi_frame_stats {
    uint64_t frames;        /* frames counted since the last reset */
    uint64_t rays_closest;  /* closest-hit rays traced (live rays, all bounces) */
    uint64_t rays_light;    /* light-geometry any-hit rays (NEE pdf) */
    uint64_t bvh_nodes;     /* nodes of the current geometry BVH */
    uint32_t bvh_depth;     /* depth of the current geometry BVH */
    uint32_t pad;
} vrgi_frame_stats;

typedef struct vrgi_ctx vrgi_ctx;
typedef struct vrgi_object vrgi_object;

const char *vrgi_last_error(void);
uint32_t vrgi_abi_version(void);
void vrgi_config_reference(vrgi_config *cfg, uint32_t width, uint32_t height);

/* Scene::new + Renderer::new (scene.rs:112, interactive_rendering.rs:447).
 * materials: n_slots texture slots; texture_luminances: utils::get_texture_luminances
 * (utils.rs:223-235), one per slot, feeding light power = luminance * area. */
int vrgi_create(const vrgi_config *cfg, const vrgi_material *materials,
                const float *texture_luminances, uint32_t n_slots, int device, vrgi_ctx **out);
int vrgi_destroy(vrgi_ctx *ctx);

/* The texture atlas of Renderer::new (interactive_rendering.rs:447, 735-760): for every
 * texture slot t (the vertices' `t`, block.rs:116-119) three images, reflectivity
 * (rgb, alpha = 1 - transmit probability), emissivity (rgb) and metallicity (r), each
 * width x height RGBA8 in sRGB (R8G8B8A8_SRGB, interactive_rendering.rs:416), laid out
 * [slot][3][height][width][4]. Shading then reads them at the hit's interpolated uv
 * (raytrace.rs:437-481) with the reference's default sampler: nearest texel, clamp to
 * edge, colour channels through the sRGB transfer function, alpha / 255. n_slots must
 * cover the material table; rgba8 == NULL returns to the constant materials of
 * vrgi_create. Not with ReSTIR DI (restir_mode 1), nor while the scene is in a batch
 * (a textured scene may join one afterwards: its batched frames read the atlas too). */
int vrgi_set_textures(vrgi_ctx *ctx, uint32_t n_slots, uint32_t width, uint32_t height, const uint8_t *rgba8);
/* utils::get_texture_luminances (utils.rs:223-235) over such an atlas: per slot, the mean
 * over its emissivity image of r + g + b in 8-bit units (the texture_luminances of
 * vrgi_create). */
int vrgi_texture_luminances(uint32_t n_slots, uint32_t width, uint32_t height, const uint8_t *rgba8, float *out);

/* SceneUploader::upload_object (scene.rs:547-725). Thread-safe: concurrent calls run
 * side by side, each on its own stream and builder workspace (the reference uploads from
 * a pool of threads, chunk_manager.rs:214-258); it must not race with vrgi_destroy.
 * n_verts == 0 yields the Empty handle (*out == NULL, status 0); n_verts % 3 != 0 is
 * VRGI_E_INVALID. Builds the object's light BL-BVH (build_bl_bvh over its emissive
 * triangles). An object may be released after its context is destroyed. */
int vrgi_upload_object(vrgi_ctx *ctx, const vrgi_vertex3d *verts, size_t n_verts,
                       vrgi_object **out);
int vrgi_object_retain(vrgi_object *obj);
int vrgi_object_release(vrgi_object *obj);
/* light BL-BVH of an uploaded object: node count and host copy (test hook). */
int vrgi_object_light_bvh(vrgi_object *obj, vrgi_bvh_node *dst, size_t cap, size_t *n_nodes);

/* Scene::add_object / update_object / remove_object (scene.rs:150-225).
 * transform: InstanceData.transform, mat4x3 column-major (4 columns of xyz).
 * add: an object uploaded through another context is VRGI_E_INVALID; a replaced object
 * is kept until vrgi_dispose_old. update: unknown key is VRGI_E_INVALID (the reference
 * panics, scene.rs:215). remove: removed objects are kept until vrgi_dispose_old. */
int vrgi_scene_add(vrgi_ctx *ctx, uint32_t key, vrgi_object *obj, const float transform[12]);
int vrgi_scene_update(vrgi_ctx *ctx, uint32_t key, const float transform[12]);
int vrgi_scene_remove(vrgi_ctx *ctx, uint32_t key);
/* Scene::dispose_old_objects (scene.rs:227). */
int vrgi_dispose_old(vrgi_ctx *ctx);

/* Renderer::rebuild(extent) (interactive_rendering.rs:818). Frame buffers and ReSTIR
 * histories restart at the new extent (so does FP64 accumulation when it is on). A zero
 * width or height marks the context minimised: vrgi_render then draws nothing and
 * returns VRGI_OK with every field of its output zeroed until the next non-zero extent
 * (interactive_rendering.rs:978-981). */
int vrgi_resize(vrgi_ctx *ctx, uint32_t width, uint32_t height);

/* Renderer::render (interactive_rendering.rs:952-1669): (re)builds the BVHs when
 * the scene changed, then raygen, B x (traverse + shade + NEE), nee_pdf,
 * outgoing radiance, ReSTIR temporal / spatial / finalize, postprocess.
 * Work is queued on the context's stream; vrgi_sync waits for it. */
int vrgi_render(vrgi_ctx *ctx, const vrgi_camera *cam, const vrgi_prefs *prefs,
                vrgi_frame_out *out);
int vrgi_sync(vrgi_ctx *ctx);
int vrgi_stats(vrgi_ctx *ctx, vrgi_frame_stats *st, int reset);

/* ---- measurement (no counterpart in the reference: it only prints fps, main.rs:354-361) ----
 * VRGI_PROFILE_STAGES: every render() records HIP events around its stages on the
 * context's stream and waits for the frame to accumulate their durations.
 * VRGI_PROFILE_TOUCH: the next render() counts the distinct BVH nodes and triangles
 * each bounce touches (a counting instantiation of the trace kernel; that frame's
 * timings are not representative). */
#define VRGI_PROFILE_STAGES 1u
#define VRGI_PROFILE_TOUCH 2u
typedef struct vrgi_profile_out {
    uint64_t frames;             /* frames accumulated */
    double build_ms;             /* flatten + batched BVH builds + traversal copies */
    double trace_ms[16];         /* trace + shade kernel, per bounce */
    double nee_ms[16];           /* NEE pdf kernel, per bounce */
    double resolve_ms;           /* outgoing radiance + ReSTIR + postprocess */
    double frame_ms;             /* first to last event of the frame */
    uint64_t rays[16];           /* rays traced per bounce (summed over frames) */
    uint64_t touched_bytes[16];  /* unique node (64 B) + triangle (48 B) bytes per bounce, last TOUCH frame */
    uint64_t build_prims;        /* primitives of the last geometry build (both trees) */
    uint64_t build_leaf_depth_sum; /* sum over leaves of their depth, last build */
    uint64_t build_nodes;        /* nodes written by the last build */
    uint64_t nee_listed[16];     /* lazy NEE pdf: rays whose full pdf was computed, per bounce */
} vrgi_profile_out;
int vrgi_profile(vrgi_ctx *ctx, uint32_t flags);
int vrgi_profile_read(vrgi_ctx *ctx, vrgi_profile_out *out, int reset);

/* ---- readback / test hooks (host copies; each synchronises the stream) ---- */
enum vrgi_buffer_id {
    VRGI_BUF_GEOM_BVH = 0,       /* vrgi_bvh_node[2N-1], geometry BVH (build_bl_bvh semantics) */
    VRGI_BUF_GEOM_PRIM_ORDER,    /* uint32[N] global triangle id per leaf position */
    VRGI_BUF_LIGHT_TL_BVH,       /* vrgi_bvh_node[], light top-level BVH (build_tl_bvh) */
    VRGI_BUF_LIGHT_GEOM_BVH,     /* vrgi_bvh_node[], BVH over world-space light triangles */
    VRGI_BUF_PATH_LEN,           /* uint32[P] live bounces per pixel */
    VRGI_BUF_BOUNCE_ORIGIN,      /* float[B][P][3] origin written by bounce b (= ray origin b+1) */
    VRGI_BUF_BOUNCE_DIRECTION,   /* float[B][P][3] direction written by bounce b */
    VRGI_BUF_BOUNCE_NORMAL,      /* float[B][P][3] */
    VRGI_BUF_BOUNCE_EMISSIVITY,  /* float[B][P][3] */
    VRGI_BUF_BOUNCE_REFLECTIVITY,/* float[B][P][3] */
    VRGI_BUF_BOUNCE_MIS_WEIGHT,  /* float[B][P] */
    VRGI_BUF_BOUNCE_BSDF_PDF,    /* float[B][P] */
    VRGI_BUF_BOUNCE_NEE_PDF,     /* float[B][P] */
    VRGI_BUF_RADIANCE,           /* float[P][3] */
    VRGI_BUF_RESTIR,             /* float[P][3] */
    VRGI_BUF_TEMPORAL_RESERVOIR, /* float[P][20] slot (frame % 3) after the last frame */
    VRGI_BUF_SPATIAL_RESERVOIR,  /* float[P][20] */
    VRGI_BUF_BGRA8,              /* uint8[P][4] */
    VRGI_BUF_HIT_PRIM,           /* uint32[B][P] global triangle id hit at bounce b (0xFFFFFFFF miss) */
    VRGI_BUF_HIT_T,              /* float[B][P] hit distance at bounce b */
    VRGI_BUF_DI_LIGHT,           /* uint32[P] DI reservoir light (global light triangle id) */
    VRGI_BUF_DI_WEIGHTS,         /* float[P][4] DI reservoir (w_sum, W, M, p_hat) */
    VRGI_BUF_BOUNCE_ALBEDO,      /* float[B][P][3] material reflectivity at each hit (cache enabled) */
    VRGI_BUF_CACHE_RADIANCE,     /* float[P][3] cache radiance of the queried pixels (cache enabled) */
    VRGI_BUF_CACHE_QUERIES,      /* float[n][16] query records of the last frame (cache enabled) */
    VRGI_BUF_CACHE_TRAIN_RECORDS,/* float[n][16] training records of the last frame (cache enabled) */
    VRGI_BUF_CACHE_FALLBACK,     /* uint8[tiles] per-tile fallback flags for the next frame (cache enabled) */
    VRGI_BUF_COUNT
};
/* bytes needed for buffer `which` in the current state (0 if not available). */
int vrgi_buffer_size(vrgi_ctx *ctx, int which, size_t *bytes);
int vrgi_read_buffer(vrgi_ctx *ctx, int which, void *dst, size_t bytes);

/* ---- learned radiance cache (no counterpart in the reference, which has no learned part) ----
 * A small MLP (64 encoded inputs -> 4 x 64 ReLU -> 3, bf16 on the matrix cores, fp32
 * accumulation) over a 12-level x 2-feature hash grid, spherical harmonics (l < 4) of the
 * ray direction and of the normal, and the albedo; relative L2 loss, Adam. While enabled,
 * render() traces full training paths (train_bounces) from a random train_fraction of
 * the pixels, ends every other path at its bounce-1 vertex with the cache's reflected
 * radiance there, and then takes steps_per_frame Adam steps on the training paths'
 * vertices. (query_vertex selects that vertex: the camera ray's hit or the next one.) A record (query or training) is 16 floats: position, incoming direction,
 * normal, albedo, target radiance (training), key = pixel * 16 + bounce (uint bits). */
typedef struct vrgi_cache_config {
    uint32_t log2_table;        /* entries per grid level = 2^log2_table (18) */
    uint32_t res_min, res_max;  /* grid resolution of the coarsest and finest level (16, 2048) */
    float box_lo[3], box_hi[3]; /* region mapped to the grid; all zero: the geometry's root box at
                                   the first render after enabling */
    float train_fraction;       /* share of pixels tracing training paths (0.03) */
    uint32_t train_bounces;     /* bounces of a training path, 2..16 (8) */
    uint32_t batch;             /* records per Adam step, multiple of 16 (16384) */
    uint32_t steps_per_frame;   /* Adam steps per frame (4) */
    float lr, beta1, beta2, eps;/* Adam (0.01, 0.9, 0.99, 1e-10) */
    float loss_scale;           /* output-gradient scale, divided out before Adam (128) */
    uint32_t seed;              /* initial weights and table */
    uint32_t flags;             /* VRGI_CACHE_* */
    uint32_t query_vertex;      /* where non-training paths end in the cache: 0 the camera ray's
                                   hit, 1 the next vertex (0) */
    float fallback_bound;       /* per-tile fallback to the plain tracer: a tile whose held-out
                                   relative error |E[cache] - E[target]| / (E[target] + 0.01)
                                   exceeds this traces full paths next frame; 0 = off (0) */
    uint32_t fallback_tile;     /* tile side in pixels (16) */
    float fallback_ema;         /* weight of a frame in the per-tile means (0.1) */
    uint32_t precision;         /* VRGI_CACHE_BF16 (default) or VRGI_CACHE_FP16: weights, activations
                                   and gradients on the matrix cores, fp32 accumulation */
} vrgi_cache_config;

#define VRGI_CACHE_FREEZE 1u  /* query only: render() takes no training steps */
#define VRGI_CACHE_DYNAMIC_SCALE 2u /* dynamic loss scaling (always on in fp16): a step whose scaled
                                      gradients overflow the 16-bit format is skipped and the scale
                                      halved; the scale doubles after VRGI_CACHE_SCALE_INTERVAL good
                                      steps in a row (up to 2^24) */
#define VRGI_CACHE_SCALE_INTERVAL 1000u
#define VRGI_CACHE_BF16 0u
#define VRGI_CACHE_FP16 1u
#define VRGI_CACHE_WEIGHTS (4 * 64 * 64 + 64 * 16) /* floats: layers 1-4 [k][n] 64x64, layer 5 [k][n] 64x16 */

typedef struct vrgi_cache_info {
    uint32_t levels, feats, log2_table;
    uint32_t res[16];
    float box_lo[3], box_hi[3];
    uint64_t table_floats;      /* levels x 2^log2_table x feats */
    uint64_t steps;             /* Adam steps taken */
    float last_loss;            /* mean loss of the last frame's steps */
    float loss_scale;           /* the current loss scale */
    uint32_t precision;         /* VRGI_CACHE_BF16 / VRGI_CACHE_FP16 */
    uint64_t skipped_steps;     /* steps skipped by dynamic loss scaling (overflow) */
} vrgi_cache_info;

void vrgi_cache_config_default(vrgi_cache_config *cfg);
int vrgi_cache_enable(vrgi_ctx *ctx, const vrgi_cache_config *cfg);
int vrgi_cache_disable(vrgi_ctx *ctx);
int vrgi_cache_info_get(vrgi_ctx *ctx, vrgi_cache_info *out);
/* host copies of the fp32 parameters: weights (VRGI_CACHE_WEIGHTS), table (table_floats) */
int vrgi_cache_get(vrgi_ctx *ctx, float *weights, float *table);
/* replaces the parameters (either may be NULL) and clears the Adam state */
int vrgi_cache_set(vrgi_ctx *ctx, const float *weights, const float *table);
/* n samples of 12 floats (position, direction, normal, albedo) -> n x 3 radiance */
int vrgi_cache_infer(vrgi_ctx *ctx, uint64_t n, const float *samples, float *out);
/* gradients of the mean loss over n samples (targets n x 3), without an update */
int vrgi_cache_grads(vrgi_ctx *ctx, uint64_t n, const float *samples, const float *targets, float *grad_w,
                     float *grad_table, float *loss);
/* one Adam step on n samples; loss before the step */
int vrgi_cache_train(vrgi_ctx *ctx, uint64_t n, const float *samples, const float *targets, float *loss);

/* ---- convergence measurement (no counterpart in the reference) ----
 * While enabled, every render() adds its path-traced radiance (VRGI_BUF_RADIANCE) to an
 * FP64 running sum; vrgi_accum_read returns the mean over the frames since enabling. */
int vrgi_accum_enable(vrgi_ctx *ctx, int on);
/* number of the next frame (seeds: raytrace frame * B + bounce, ReSTIR frame salts) */
int vrgi_frame_set(vrgi_ctx *ctx, uint64_t frame);
int vrgi_accum_read(vrgi_ctx *ctx, double *mean /* P x 3 */, uint64_t *frames);

/* ---- standalone operators (the drop-in pieces, usable without a scene) ---- */

/* build_bl_bvh (build.rs:249-372) on the GPU over n triangles (tri_xyz: 9n floats),
 * prim_luminances (n), prim_index_ids (n). nodes: 2n-1 entries; prim_order: n
 * entries (prim index per leaf position); root_box[6] and root_power as the
 * function's returned (Aabb, f32). Host pointers. */
int vrgi_build_bl_bvh(vrgi_ctx *ctx, const float *tri_xyz, const float *prim_luminances,
                      const uint32_t *prim_index_ids, uint64_t n, vrgi_bvh_node *nodes,
                      uint32_t *prim_order, float root_box[6], float *root_power);

/* No counterpart in the reference: vrgi_build_bl_bvh for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and errors as vrgi_build_bl_bvh, for at most
 * VRGI_NOALLOC_MAX_PRIMS primitives (more is VRGI_E_LIMIT); it calls no hipMalloc /
 * hipFree. Its staging and builder workspace are one fixed array per GPU (about 18 MiB,
 * held until the process ends) shared by every call on that GPU: calls that overlap on
 * one GPU race. */
#define VRGI_NOALLOC_MAX_PRIMS 65536u
int vrgi_build_bl_bvh_noalloc(vrgi_ctx *ctx, const float *tri_xyz, const float *prim_luminances,
                              const uint32_t *prim_index_ids, uint64_t n, vrgi_bvh_node *nodes,
                              uint32_t *prim_order, float root_box[6], float *root_power);

/* build_tl_bvh (build.rs:374-480): m instance boxes (aabbs: 6m = min xyz, max xyz),
 * transforms (12m, column-major mat4x3), powers (m), ids (m). nodes: 2m-1. */
int vrgi_build_tl_bvh(vrgi_ctx *ctx, const float *transforms, const float *aabbs,
                      const float *powers, const uint32_t *ids, uint64_t m,
                      vrgi_bvh_node *nodes);

/* Closest hit of n rays against the current scene's geometry BVH (replaces the
 * ray query of raytrace.rs:354-388). t in [tmin, tmax]; ties broken by the lower
 * global triangle id. Outputs: hit triangle id (0xFFFFFFFF = miss), t, and the
 * barycentrics (u, v) of vertices 1 and 2. Host pointers. */
int vrgi_trace_closest(vrgi_ctx *ctx, const float *origins, const float *dirs, uint64_t n,
                       float tmin, float tmax, uint32_t *hit_prim, float *hit_t,
                       float *hit_uv);

/* chunk::gen_mesh (chunk.rs:158-292) on the GPU for a batch of chunks, in the same vertex
 * order as vrgi_gen_mesh_chunk. blocks: dims[0] x dims[1] x dims[2] block ids, x-major
 * (n_types = air); chunks: n_chunks x (cx, cy, cz). The chunks' vertices go to out one
 * chunk after another (chunk-local positions); counts[i] = vertices of chunk i. Returns
 * VRGI_E_LIMIT (counts filled, out untouched) when their total exceeds cap. */
int vrgi_mesh_chunks(vrgi_ctx *ctx, const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                     const uint8_t *translucent, uint32_t n_chunks, const uint32_t *chunks,
                     vrgi_vertex3d *out, uint64_t cap, uint64_t *counts);

/* ---- batched scenes ---------------------------------------------------------------
 * No counterpart in the reference (its Renderer draws one Scene per frame); for callers
 * that render many independent scenes or views of uneven size per step. A batch groups
 * contexts made with vrgi_create (same device, same bounce count and ReSTIR mode; no
 * learned cache, no ray sorting, no VRGI_FLAG_KEEP_BOUNCES, no FP64 accumulation). One
 * batched frame:
 *   - rebuilds, in ONE multi-tree build, the BVHs of every scene whose objects changed
 *     (every scene with VRGI_FLAG_REBUILD_EVERY_FRAME);
 *   - runs `spp` passes; each pass is exactly one vrgi_render frame of every scene (same
 *     bits: each scene's frame counter advances by one per pass), with ONE launch per
 *     stage for all scenes, replayed from a captured HIP graph after the first pass;
 *   - keeps per scene the mean of the passes' ReSTIR outputs,
 *     (((r_1 + r_2) + ...) + r_spp) * (1 / spp) in binary32.
 * Rays and pixels per block are fixed by the config or picked by timing each candidate
 * on the first batched frames (the results do not depend on them). */
typedef struct vrgi_batch vrgi_batch;
typedef struct vrgi_batch_config {
    uint32_t spp;              /* passes per batched frame, >= 1 */
    uint32_t rays_per_block;   /* 64, 128 or 256 (one LDS stack per 64); 0 = autotuned */
    uint32_t pixels_per_block; /* 64, 128 or 256; 0 = autotuned */
    uint32_t flags;            /* VRGI_BATCH_* */
} vrgi_batch_config;
#define VRGI_BATCH_NO_GRAPH 1u /* launch every pass directly (no graph capture) */
typedef struct vrgi_batch_info {
    uint32_t n_scenes, rays_per_block, pixels_per_block;
    uint32_t tuned;             /* 1 once the autotuner has fixed both sizes */
    uint32_t graph;             /* 1: passes replay captured graphs; 2: the last batched frame
                                   ran as one graph (its builds replayed, then every pass) */
    uint32_t pad;
    uint64_t triangles, light_triangles, pixels;
    float build_ms, passes_ms;  /* GPU time of the last batched frame's build and passes (graph
                                   2 with builds: build_ms -1, passes_ms the whole frame) */
    uint64_t rays;              /* closest-hit rays since the previous vrgi_batch_info_get */
} vrgi_batch_info;
void vrgi_batch_config_default(vrgi_batch_config *cfg);
int vrgi_batch_create(vrgi_ctx *const *scenes, uint32_t n_scenes, const vrgi_batch_config *cfg,
                      vrgi_batch **out);
int vrgi_batch_destroy(vrgi_batch *batch);
/* cams: one per scene; prefs: shared. Returns when the work is queued. */
int vrgi_batch_render(vrgi_batch *batch, const vrgi_camera *cams, const vrgi_prefs *prefs);
int vrgi_batch_sync(vrgi_batch *batch);
/* the spp mean of scene `scene` (its P x 3 floats) after the last batched frame */
int vrgi_batch_read_mean(vrgi_batch *batch, uint32_t scene, float *rgb);
int vrgi_batch_info_get(vrgi_batch *batch, vrgi_batch_info *info);

/* ---- device-resident voxel world -------------------------------------------------
 * game_system's chunk manager (chunk_manager.rs:173-399) on the GPU, for callers that
 * edit the world every frame. vrgi_voxels_load keeps the block grid on the device and
 * makes every 32^3 chunk an object of the scene, with the keys and transforms the CPU
 * chunk path gives: key (cx * ny + cy) * nz + cz, translation (32 cx, 32 cy, 32 cz),
 * vertices of chunk::gen_mesh (chunk.rs:158-292), the light BL-BVH and light vertices of
 * upload_object (scene.rs:547-725). vrgi_voxels_set writes voxels; the chunks whose
 * meshes can change (the voxel's chunk and, on a chunk border, the neighbour across it)
 * are re-meshed and their objects replaced at the next vrgi_render (or
 * vrgi_voxels_commit): one mesher pass and one multi-tree light build for all of them.
 * Replaced objects are released at the next commit. dims: multiples of 32; blocks:
 * x-major u16, air = n_types; translucent: one flag per block type. */
int vrgi_voxels_load(vrgi_ctx *ctx, const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                     const uint8_t *translucent);
/* writes blocks[i] at voxel (xyz[3i], xyz[3i+1], xyz[3i+2]); the n positions must be distinct */
int vrgi_voxels_set(vrgi_ctx *ctx, uint64_t n, const uint32_t *xyz, const uint16_t *blocks);
int vrgi_voxels_commit(vrgi_ctx *ctx);
/* test hook: the device grid (may be NULL), commits and chunks re-meshed so far */
int vrgi_voxels_read(vrgi_ctx *ctx, uint16_t *blocks, uint64_t *commits, uint64_t *chunks_remeshed);

/* ---- synthetic scenes (caller side: the game_system worldgen + chunk mesher) ---- */

/* Seeded voxel world of one of the benchmark workloads: kind 0 = 64x64x64
 * value-noise terrain (4 block types, 1% emissive surface voxels); kind 1 =
 * 256x128x256 terrain + 200 boxes (0.5% emissive); kind 100 + i = scene i of the
 * batched workload (each side 32..160 voxels, drawn per scene; 1% emissive).
 * blocks: u16 per voxel, index x*(dy*dz) + y*dz + z; air = n_block_types. */
int vrgi_world_dims(int kind, uint32_t dims[3], uint32_t *n_block_types);
int vrgi_world_generate(int kind, uint64_t seed, uint16_t *blocks);
/* material table of a generated world (6 slots per block type) and luminances. */
int vrgi_world_materials(int kind, uint64_t seed, vrgi_material *mats, float *lum,
                         uint32_t cap_slots, uint32_t *n_slots);
/* translucent flag per block type of a generated world (air excluded). */
int vrgi_world_translucent(int kind, uint8_t *translucent, uint32_t cap);

/* Edits of the large dynamic workload (caller side): n_steps steps of n_walkers seeded
 * random walkers (walkers: 3 x n_walkers voxel positions, updated in place); a step moves
 * walker (step % n_walkers) by one voxel along a random axis and toggles the voxel it
 * reaches (solid -> air, air -> block type 2). blocks is updated in place; the distinct
 * voxels touched (first-touch order) and their final ids go to xyz / vals (capacity
 * n_steps), their count to *n_out. */
int vrgi_world_walk(uint16_t *blocks, const uint32_t dims[3], uint32_t n_types, uint64_t seed,
                    uint32_t n_walkers, uint32_t *walkers, uint64_t n_steps, uint32_t *xyz,
                    uint16_t *vals, uint64_t *n_out);

/* chunk::gen_mesh (chunk.rs:158-292) for the 32^3 chunk at chunk coordinates
 * (cx, cy, cz) of a world of dims; voxels outside the world are air. Writes at
 * most cap vertices, returns the vertex count in *n_out (chunk-local positions). */
int vrgi_gen_mesh_chunk(const uint16_t *blocks, const uint32_t dims[3],
                        uint32_t n_block_types, const uint8_t *translucent, uint32_t cx,
                        uint32_t cy, uint32_t cz, vrgi_vertex3d *out, size_t cap,
                        size_t *n_out);

#ifdef __cplusplus
}
#endif
#endif /* VRGI_H */
