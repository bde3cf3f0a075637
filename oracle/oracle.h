// oracle.h — CPU restatement of the reference's hot path (TEST INFRASTRUCTURE).
//
// This is the checker, not the product: only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it. It restates, in plain C++ with FP32
// exact semantics (no FMA contraction, correctly rounded / and sqrt):
//   - the per-pixel RNG                raytrace.rs:108-149
//   - Aabb and the binned-SAH builder  bvh/aabb.rs:4-92, bvh/build.rs:36-480
//   - the exposed-face mesher          game_system/chunk.rs:158-292
//   - light gather / light vertices    scene.rs:509-571, utils.rs:223-235
//   - light TL-BVH + InstanceData      scene.rs:243-338
//   - raygen, raytrace, nee_pdf, outgoing_radiance, ReSTIR temporal / spatial /
//     finalize, postprocess            shader/*.rs
// The Vulkan ray queries (raytrace.rs:354-388, nee_pdf.rs:305-334) are replaced by
// the Woop-Benthin-Wald watertight ray/triangle test (JCGT 2(1), 2013) over a
// binned-SAH BVH built with build_bl_bvh semantics; ties between equal t are broken
// by the lower global triangle id.
//
// Parity status per function is listed at the top of oracle.cpp; the functions whose
// arithmetic lives outside the reference's sources (the `partition` 0.1.2 crate order,
// Vulkan barycentrics, GLSL built-ins) are "parity unpinned" against the reference.
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

// ---------- RNG (raytrace.rs:108-149) ----------
uint32_t or_murmur3_combine(uint32_t h, uint32_t k);
uint32_t or_murmur3_finalize(uint32_t h);
float or_float_construct(uint32_t m);
// deterministic sin / cos used by both sides (own polynomial, see oracle.cpp)
void or_sincos(float x, float *s, float *c);

// ---------- BVH (build.rs) ----------
// nodes are 40-byte BvhNode records (bvh/mod.rs:6-18)
int or_build_bl_bvh(uint64_t n, const float *tri_xyz, const float *lum, const uint32_t *ids,
                    void *nodes_out, uint32_t *prim_order_out, float *root_box6,
                    float *root_power, uint32_t *depth_out);
int or_build_tl_bvh(uint64_t m, const float *xforms12, const float *aabbs6, const float *powers,
                    const uint32_t *ids, void *nodes_out);
// brute-force check of find_best_plane (build.rs:45-128) on the range of the root
int or_find_best_plane(uint64_t n, const float *tri_xyz, uint32_t *dim_out, float *split_out);

// ---------- mesher (chunk.rs:158-292) ----------
int or_gen_mesh_chunk(const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                      const uint8_t *translucent, uint32_t cx, uint32_t cy, uint32_t cz,
                      void *out_verts, uint64_t cap, uint64_t *n_out);

// ---------- scene + renderer ----------
typedef struct or_scene or_scene;
or_scene *or_scene_create(const void *materials /*vrgi_material[n]*/, const float *lum,
                          uint32_t n_slots);
void or_scene_destroy(or_scene *s);
// texture atlas: n_slots x 3 images (reflectivity, emissivity, metallicity) of w x h RGBA8
// sRGB texels, [slot][3][h][w][4]; null clears it (constant materials)
int or_scene_set_textures(or_scene *s, uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8);
float or_srgb_decode(uint32_t v);
uint32_t or_texel_index(float c, uint32_t n);
void or_texture_luminances(uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8, float *out);
int or_scene_add(or_scene *s, uint32_t key, const void *verts /*vrgi_vertex3d*/, uint64_t n,
                 const float xform[12]);
int or_scene_remove(or_scene *s, uint32_t key);
int or_scene_build(or_scene *s); // flatten + geometry BVH + light BVHs
uint64_t or_scene_count(or_scene *s, int what); // 0 tris, 1 light tris, 2 instances, 3 tl nodes, 4 lgeom nodes
int or_scene_read(or_scene *s, int what, void *dst, uint64_t bytes);

int or_scene_light_bvh(or_scene *s, uint32_t key, void *dst, uint64_t cap, uint64_t *n);
int or_trace_closest(or_scene *s, uint64_t n, const float *o, const float *d, float tmin,
                     float tmax, int brute, uint32_t *hit, float *t, float *uv,
                     uint64_t *visited_nodes);
int or_trace_lights(or_scene *s, const float *o, const float *d, float tmin, float tmax,
                    uint32_t *hits, uint32_t cap, uint32_t *n_hits, float *uvs);
int or_traverse_light_bvh(or_scene *s, const float *point, const float *normal, uint32_t seed,
                          uint32_t *inst, uint32_t *prim, float *prob, float *importance);
float or_reverse_traverse_light_bvh(or_scene *s, const float *point, const float *normal,
                                    uint32_t inst, uint32_t bl_node);

// nee_pdf.rs:281-337 for n rays: origin o, shading normal nrm, direction d (the sum of
// computeNeePdf over every light triangle the ray crosses), the function frames use
int or_nee_pdf(or_scene *s, uint64_t n, const float *o, const float *nrm, const float *d, float *out);
// computeJacobianQToR (restir_spatial_resampling.rs:299-311) of two 20-word reservoirs
float or_restir_gi_jacobian(const float *q_res, const float *r_res);

// Planted mistakes, for the tests that show each pinning check fails on a plausible
// error in the restatement (0 = the faithful restatement).
enum {
    OR_MUT_NONE = 0,
    OR_MUT_NEE_AREA = 1,      // computeNeePdf: triangle area without the 1/2
    OR_MUT_NEE_DIST = 2,      // computeNeePdf: d instead of d^2
    OR_MUT_NEE_BARY = 3,      // computeNeePdf: raytrace's barycentric order (1-x-y, x, y)
    OR_MUT_NEE_COS = 4,       // computeNeePdf: no cosine
    OR_MUT_GI_UCW_NO_M = 5,   // temporal W = w_sum / p_hat (M left out)
    OR_MUT_GI_MERGE_NO_M = 6, // spatial merge weight without M_n
    OR_MUT_GI_Z_ALL = 7,      // spatial Z counts every reservoir (no p_hat > 0 filter)
    OR_MUT_GI_FINAL_COS = 8,  // finalize without the cosine
    OR_MUT_GI_JAC_SQ = 9,     // Jacobian with squared distances
    OR_MUT_GI_JAC_SWAP = 10,  // Jacobian of r -> q instead of q -> r
    OR_MUT_DI_AREA = 11,      // DI source pdf without the 1/area
    OR_MUT_DI_COSY = 12,      // DI contribution without the light's cosine
    OR_MUT_DI_UCW_NO_M = 13,  // DI W = w_sum / p_hat (M left out)
    OR_MUT_DI_SPATIAL_M = 14, // DI spatial merge adds 1 to M instead of M_n
};
void or_set_mutation(int m);
int or_get_mutation(void);

typedef struct or_config {
    uint32_t width, height, num_bounces, rr_start_bounce;
    float rr_min_survival;
    uint32_t restir_mode;
    float restir_spatial_radius;
    uint32_t di_candidates, seed, flags;
} or_config;
typedef struct or_camera { float eye[3], front[3], right[3], up[3]; } or_camera;
typedef struct or_prefs { uint32_t nee_type, debug_view, restir_spatial_iterations; } or_prefs;

// Dense reference-layout outputs of one frame. Any pointer may be NULL.
typedef struct or_frame {
    float *origin;      // [B+1][P][3]
    float *direction;   // [B+1][P][3]
    float *normal;      // [B][P][3]
    float *emissivity;  // [B][P][3]
    float *reflectivity;// [B][P][3]
    float *mis_weight;  // [B][P]
    float *bsdf_pdf;    // [B][P]
    float *nee_pdf;     // [B][P]
    float *l_out;       // [B][P][3]
    float *q_omega;     // [B][P]
    uint32_t *hit_prim; // [B][P]
    float *hit_t;       // [B][P]
    uint32_t *path_len; // [P]
    float *temporal;    // [P][20] (slot frame % 3 after the update)
    float *spatial;     // [P][20]
    float *restir;      // [P][3]
    float *debug;       // [P][3]
    uint8_t *bgra8;     // [P][4]
    uint32_t *di_light; // [P]
    float *di_weights;  // [P][4]
} or_frame;

// Learned-radiance-cache frame: a pixel traces a training path of all B bounces when
// u(finalize(combine(train_seed, pixel))) < train_fraction, otherwise it stops at bounce
// v = query_vertex and, when that bounce hit, ends with L_v = E_v + cache_rad[pixel] (the
// reflected radiance the cache returned there). Records are 16 floats: position, incoming direction,
// normal, albedo, target (training records: the reflected radiance of the vertex), key =
// pixel * 16 + bounce; both lists come out sorted by key.
typedef struct or_cache_frame {
    float train_fraction;
    uint32_t train_seed;
    uint32_t query_vertex;   // 0 or 1
    const uint8_t *fallback; // [P] in, may be null: pixels that trace full paths like training ones
    const float *cache_rad;  // [P][3] in
    float *albedo;           // [B][P][3] out: material reflectivity at each hit
    float *qrec;             // [P][16] out
    uint64_t n_qrec;
    float *trec;             // [trec_cap][16] out
    uint64_t trec_cap, n_trec;
} or_cache_frame;

typedef struct or_renderer or_renderer;
or_renderer *or_renderer_create(const or_config *cfg);
void or_renderer_destroy(or_renderer *r);
// renders frame number `frame` (ReSTIR slot frame % 3); threads: OpenMP threads (0 = default)
int or_render(or_renderer *r, or_scene *s, const or_camera *cam, const or_prefs *prefs,
              uint64_t frame, or_frame *out, int threads);
int or_render_cache(or_renderer *r, or_scene *s, const or_camera *cam, const or_prefs *prefs,
                    uint64_t frame, or_frame *out, int threads, or_cache_frame *cf);
// rows [y0, y1) only (for the bounded CPU baseline sample); no ReSTIR
int or_render_rows(or_renderer *r, or_scene *s, const or_camera *cam, const or_prefs *prefs,
                   uint64_t frame, uint32_t y0, uint32_t y1, float *radiance, int threads);
}
