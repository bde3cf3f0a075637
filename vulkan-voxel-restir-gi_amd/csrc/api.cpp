// api.cpp — host side of the library: the C ABI of include/vrgi.h over HIP.
//
// Mirrors the reference's render_system interface: SceneUploader::upload_object
// (scene.rs:547-725), Scene::{add,update,remove}_object / dispose_old_objects
// (scene.rs:150-231), Scene::get_tlas's light part (scene.rs:234-358) and
// Renderer::{rebuild, render} (interactive_rendering.rs:818-1669). All device work
// of a context is queued on one HIP stream; upload_object is thread-safe and uses
// its own stream.
#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "devcommon.h"
#include "vrgi_internal.h"
#include "vrgi_kernels.h"

using namespace vrgi;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define API_BEGIN try {
#define API_END                                                  \
    }                                                            \
    catch (const HipError &e) { return fail(VRGI_E_DEVICE, e.what()); } \
    catch (const std::bad_alloc &) { return fail(VRGI_E_DEVICE, "host allocation failed"); } \
    catch (const std::exception &e) { return fail(VRGI_E_INVALID, e.what()); }

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    // grows to n (a regrowth by 25 % more, so a scene that grows frame by frame does not
    // reallocate every frame)
    void reserve(size_t n) {
        if (n <= cap && p) return;
        const size_t want = std::max<size_t>(p ? n + n / 4 : n, 1);
        release();
        dev_alloc((void **)&p, want * sizeof(T));
        cap = want;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    ~DevBuf() { release(); }
};

struct Counters {
    unsigned long long rays_closest, rays_light, frames, pad;
};

}  // namespace

namespace {
struct VxArenaPool;  // voxel commits' arenas for reuse (below)

// The device and frame streams of a context, shared with its objects: an object may be
// released after its context is destroyed (a handle dropped late), and then finds
// `alive` false instead of a destroyed stream.
struct CtxLife {
    std::mutex mu;
    int device = 0;
    hipStream_t streams[2] = {nullptr, nullptr};  // frame stream, build stream
    bool alive = true;
};

// One upload in flight: its own stream and builder workspace, so that uploads from
// several threads run side by side (the reference uploads from a 15-thread pool,
// chunk_manager.rs:214-258).
struct UpSlot {
    hipStream_t s = nullptr;
    Builder b;
};
}  // namespace

struct vrgi_object {
    std::atomic<int> refs{1};
    std::shared_ptr<CtxLife> life;  // the uploading context's device and streams
    uint32_t n_tris = 0;
    vrgi_vertex3d *d_verts = nullptr;
    uint32_t n_light = 0;
    vrgi_bvh_node *d_light_bvh = nullptr;
    LPair *d_lpairs = nullptr;  // child records of the light BL-BVH
    vrgi_light_vertex3d *d_light_verts = nullptr;
    std::vector<vrgi_bvh_node> h_light_bvh;
    bool light_empty = true;
    float light_aabb[6] = {0, 0, 0, 0, 0, 0};
    float power = 0.0f;
    // set for a chunk object made by a voxel-world commit: the device pointers above are
    // slices of this shared allocation (freed with its last object)
    std::shared_ptr<void> arena;
};

struct vrgi_ctx {
    int device = 0;
    hipStream_t stream = nullptr, up_stream = nullptr;
    // learned-cache work of a frame that overlaps the training paths' bounces (default
    // priority, like every stream here); joined into `stream` before the frame ends
    hipStream_t side = nullptr;
    // NEE pdf of bounce b beside the trace of bounce b + 1, on stream b % VRGI_NEE_STREAMS:
    // the pdfs of different bounces are independent, so a slow one does not hold back the
    // next (the long latency-bound walks of short training-path queues)
    hipStream_t nee_s[VRGI_NEE_STREAMS] = {};
    hipEvent_t ev_nee_done[VRGI_NEE_STREAMS] = {};
    std::vector<hipEvent_t> ev_bounce;  // trace (and sort) of bounce b done
    hipEvent_t ev_query = nullptr, ev_infer = nullptr, ev_train = nullptr, ev_accum = nullptr;
    hipEvent_t ev_fstart = nullptr, ev_snap = nullptr, ev_held = nullptr;
    std::shared_ptr<CtxLife> life = std::make_shared<CtxLife>();
    std::mutex up_mu;                 // guards up_slots / up_idle
    std::vector<UpSlot *> up_slots;   // every upload slot made so far
    std::vector<UpSlot *> up_idle;    // the ones no upload holds
    bool minimized = false;           // zero extent: render() draws nothing (interactive_rendering.rs:978-981)
    // view 4 (accumulated path tracer): running sum, frames in it, and its restart state
    DevBuf<float> f_view_acc;
    uint32_t view_n = 0, view_prev = 0;
    vrgi_camera view_cam{};
    bool view_dirty = true;  // the scene or the extent changed since the sum started
    vrgi_config cfg{};
    uint32_t W = 0, H = 0, P = 0, B = 0;
    std::vector<vrgi_material> mats;
    std::vector<float> lum;
    DevBuf<vrgi_material> d_mats;
    // texture atlas (vrgi_set_textures): linear RGBA texels [slot][3][h][w]
    DevBuf<float4> d_tex;
    uint32_t tex_w = 0, tex_h = 0;
    bool tex_on = false;

    std::map<uint32_t, std::pair<vrgi_object *, std::vector<float>>> objects;
    std::vector<vrgi_object *> old;
    bool inst_dirty = true, geo_dirty = true;
    // results of the last geometry build made with readbacks, which a rebuild of the
    // unchanged geometry repeats (the builder replays that build's launch sequence): tree
    // depths, leaf-depth sum and light root box; cleared whenever the geometry changes
    struct GeoPlan {
        bool valid = false;
        uint32_t dep[2] = {0, 0};
        unsigned long long S = 0;
        vrgi_bvh_node lroot{};
    } geo_plan;
    DevBuf<uint32_t> d_check;     // replayed builds whose counts left their plan (sticky)
    uint32_t *h_check = nullptr;  // pinned copy, read at the next replay
    hipEvent_t ev_check = nullptr;
    bool check_pending = false;

    Builder geo_builder, tl_builder, api_builder;

    // instances
    std::vector<vrgi_object *> inst_objs;
    uint32_t n_inst = 0, n_tris = 0, n_ltri = 0;
    DevBuf<DevInst> d_inst;
    DevBuf<const vrgi_vertex3d *> d_vtxptr;
    DevBuf<const vrgi_light_vertex3d *> d_lvtxptr;
    DevBuf<uint32_t> d_light_off;
    DevBuf<vrgi_bvh_node> d_tl;
    DevBuf<LPair> d_tlpairs;
    uint32_t n_tl = 1;
    DevBuf<float> d_tl_prim9, d_tl_aabb, d_tl_xf, d_tl_power;
    DevBuf<uint32_t> d_tl_ids, d_tl_trees, d_tl_depth;

    // geometry + light-geometry BVHs (one batched build) and everything derived from
    // them. Two sets: while frame f's rays read one, frame f + 1 builds the other on
    // `up_stream` (unless VRGI_FLAG_SERIAL_BUILD)
    struct GeoSet {
        DevBuf<Tri> d_tri, d_ltri;
        DevBuf<float2> d_tri_uv;  // texture atlas on: uv of each triangle's 3 vertices
        DevBuf<uint32_t> d_tri_inst, d_ltri_inst, d_ltri_node;
        DevBuf<uint32_t> d_depth;
        DevBuf<vrgi_bvh_node> d_nodes;
        DevBuf<WideNode> d_gwide, d_lwide;
        DevBuf<Wide4> d_lwide4, d_gwide4;
        DevBuf<uint32_t> d_lanc, d_lanc_len;  // per light triangle: ancestors for the NEE pdf
        uint32_t geo_depth = 0, lgeo_depth = 0;
        DevBuf<uint8_t> d_lgrid;  // light grid of the lazy NEE pdf (SceneView::lgrid)
        float lg_lo[3] = {0, 0, 0}, lg_inv = 1.0f;
        uint32_t lg_dim[3] = {0, 0, 0};
        bool lg_on = false;
        hipEvent_t ev_free = nullptr;  // recorded after the last frame that read the set
        bool freed_once = false;
    } geo[2];
    uint32_t gi = 0;  // the set the next frame reads
    // canonical nodes of the last build when a batched build made it (its combined array)
    const vrgi_bvh_node *ext_gnodes = nullptr, *ext_lnodes = nullptr;
    bool in_batch = false;
    hipStream_t batch_stream = nullptr;  // the batch's stream while the scene is in one
    GeoSet &G() { return geo[gi]; }
    hipEvent_t ev_geo_ready = nullptr;
    DevBuf<float> d_prim9, d_power;  // builder inputs (one build at a time)
    DevBuf<uint32_t> d_ids, d_trees;

    // frame buffers
    DevBuf<float> f_origin, f_dir, f_normal, f_emis, f_refl, f_mis, f_pdf, f_nee, f_hit_t;
    DevBuf<uint32_t> f_hit, f_plen, f_queue, f_qcount;
    DevBuf<uint32_t> f_squeue, f_skeys, f_shist;  // Morton-ordered queues, keys, buckets
    // NEE pdf passes, one set per NEE stream: crossed-triangle counts and records per queue
    // slot, overflow list and its length
    DevBuf<uint32_t> nee_cnt[VRGI_NEE_STREAMS], nee_lid[VRGI_NEE_STREAMS], nee_over[VRGI_NEE_STREAMS],
        nee_oc[VRGI_NEE_STREAMS];
    DevBuf<float2> nee_uv[VRGI_NEE_STREAMS];
    DevBuf<float> f_radiance, f_restir, f_debug, f_temporal, f_spatial;
    DevBuf<float> f_lrun;  // [P][3] radiance of the next vertex for the lazy NEE pdf
    // lazy NEE pdf: rays listed for the full pdf, rays listed for the near-crossing walk;
    // counts per bounce (full pdf at [b], walk at [16 + b])
    DevBuf<uint32_t> f_nlist, f_mlist, f_ncount;
    DevBuf<uint8_t> f_bgra;
    DevBuf<float> di_T, di_hist, di_w;
    DevBuf<uint32_t> di_light;
    DevBuf<Counters> d_counters;
    uint64_t frame_count = 0;
    bool have_frame = false;

    // measurement
    uint32_t profile_flags = 0;
    vrgi_profile_out prof{};
    // stage events of frames f % 2 == 0 / 1; a frame's events are read at the end of the
    // next render (or by vrgi_profile_read), so that frames still overlap while profiled
    std::vector<hipEvent_t> events[2];
    bool pend[2] = {false, false}, pend_built[2] = {false, false};
    uint32_t pend_P[2] = {0, 0};
    uint32_t *h_qcount = nullptr;  // pinned, 2 x (2B + 2): queue counts (B + 2), lazy NEE list counts (B)
    DevBuf<uint32_t> touch_bits;
    DevBuf<unsigned long long> touch_count;
    DevBuf<unsigned long long> depth_sum_copy;

    // learned radiance cache
    struct Cache {
        bool on = false;
        vrgi_cache_config cfg{};
        uint32_t res[16] = {0};
        float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        bool box_ready = false;
        size_t n_table = 0;
        uint64_t steps = 0;
        uint32_t plane_cap = 0, tcap = 0, hcap = 0;
        DevBuf<float> w32, mw, vw, table, mt, vt, gw, gtable, loss;
        DevBuf<uint16_t> wt, wb, xplane, gplane;
        // the network as the frame starts: the frame's inference reads it while the
        // training steps update the live table and weights beside it
        DevBuf<float> table_snap;
        DevBuf<uint16_t> wt_snap;
        DevBuf<float> trec, qrec, crad, hrec, hout, albedo, d0;
        DevBuf<float> trec_sorted;                // the previous frame's records in Morton order
        DevBuf<uint32_t> sort_hist, sort_kr;
        // per-tile held-out error and fallback flags
        uint32_t tiles_x = 0, tiles_y = 0;
        DevBuf<float> t_acc, t_ema, hpred;
        DevBuf<uint8_t> t_fb;
        DevBuf<uint32_t> counts;  // [0] training records, [1] query records, [2] caller records
        DevBuf<float> scale_state;  // dynamic loss scaling: scale, good steps, overflow flag, skipped steps
    } cache;

    // FP64 running sum of the radiance output (vrgi_accum_*) over accum_P pixels
    bool accum_on = false;
    uint64_t accum_frames = 0;
    size_t accum_P = 0;
    DevBuf<double> accum;

    // device-resident voxel world (vrgi_voxels_*): grid, per-type data, pending edits
    struct Voxels {
        bool on = false, pending = false;
        uint32_t dims[3] = {0, 0, 0}, nc[3] = {0, 0, 0}, n_types = 0;
        DevBuf<uint16_t> blocks, edit_val;
        DevBuf<uint8_t> transl, em, dirty, mask;
        DevBuf<float> lum, prim9;
        DevBuf<uint32_t> edit_xyz, chunks, face_off, faces, lface_off, lfaces, trees, depth, tree_base, rank;
        DevBuf<uint8_t> scan_tmp;
        DevBuf<vrgi_bvh_node> roots;
        DevBuf<uint64_t> vert_off, light_off;
        Builder builder;
        std::shared_ptr<VxArenaPool> pool;
        uint64_t commits = 0, chunks_remeshed = 0;
    } vx;

    // GPU mesher workspace
    DevBuf<uint16_t> m_blocks;
    DevBuf<uint8_t> m_transl, m_mask;
    DevBuf<uint32_t> m_chunks, m_face_off, m_faces;
    DevBuf<uint64_t> m_vert_off;
    DevBuf<vrgi_vertex3d> m_out;

    // host-side scratch of the standalone operators
    DevBuf<float> a_f0, a_f1, a_f2;
    DevBuf<uint32_t> a_u0, a_u1, a_u2;
    DevBuf<vrgi_bvh_node> a_nodes;

    ~vrgi_ctx();
};

namespace {

// the scene's geometry changed: the next frame rebuilds it, with readbacks (no replay)
void mark_geo_dirty(vrgi_ctx *c) {
    c->geo_dirty = true;
    c->geo_plan.valid = false;
}

// current device of the context; a scene in a batch first waits for the batch's queued work
void set_dev(vrgi_ctx *c) {
    HIP_CHECK(hipSetDevice(c->device));
    if (c->batch_stream) HIP_CHECK(hipStreamSynchronize(c->batch_stream));
}

SceneView scene_view(vrgi_ctx *c) {
    SceneView sv{};
    sv.gwide = c->G().d_gwide.p;
    sv.gwide4 = c->G().d_gwide4.p;
    sv.tri = c->G().d_tri.p;
    sv.n_tris = c->n_tris;
    sv.root_prim = 0;
    sv.depth = c->G().geo_depth;
    sv.tri_inst = c->G().d_tri_inst.p;
    sv.inst = c->d_inst.p;
    sv.n_inst = c->n_inst;
    sv.tl = c->d_tl.p;
    sv.lwide = c->G().d_lwide.p;
    sv.lwide4 = c->G().d_lwide4.p;
    sv.ltri = c->G().d_ltri.p;
    sv.n_ltri = c->n_ltri;
    sv.lroot_prim = 0;
    sv.ldepth = c->G().lgeo_depth;
    sv.ltri_inst = c->G().d_ltri_inst.p;
    sv.ltri_node = c->G().d_ltri_node.p;
    sv.tlpairs = c->d_tlpairs.p;
    sv.lanc = c->G().d_lanc.p;
    sv.lanc_len = c->G().d_lanc_len.p;
    sv.mats = c->d_mats.p;
    sv.n_mats = (uint32_t)c->mats.size();
    if (c->tex_on) {
        sv.tex = c->d_tex.p;
        sv.tri_uv = c->G().d_tri_uv.p;
        sv.tex_w = c->tex_w;
        sv.tex_h = c->tex_h;
    }
    if (c->G().lg_on) {
        sv.lgrid = c->G().d_lgrid.p;
        for (int k = 0; k < 3; ++k) { sv.lg_lo[k] = c->G().lg_lo[k]; sv.lg_dim[k] = c->G().lg_dim[k]; }
        sv.lg_inv = c->G().lg_inv;
    }
    return sv;
}

// the light grid of the set just built, over the light-geometry root box `mn`..`mx`:
// cells of side 1 (larger when the box would need more than 2^25 cells), every cell within
// 1 + one side of a light triangle's box marked
void build_light_grid(vrgi_ctx *c, hipStream_t s, const float mn[3], const float mx[3]) {
    auto &G = c->G();
    G.lg_on = false;
    if (!c->n_ltri) return;
    double side = 1.0;
    for (int it = 0; it < 8; ++it) {
        const double m = 1.0 + side;
        double cells = 1.0;
        for (int k = 0; k < 3; ++k) cells *= std::ceil(((double)mx[k] - mn[k] + 2.0 * m + 2.0 * side) / side) + 1.0;
        if (!(cells <= (double)(1u << 25))) { side *= 1.5; continue; }
        const float margin = (float)m;
        for (int k = 0; k < 3; ++k) {
            G.lg_lo[k] = (float)(mn[k] - m - side);
            G.lg_dim[k] = (uint32_t)std::ceil(((double)mx[k] - mn[k] + 2.0 * m + 2.0 * side) / side) + 1u;
        }
        G.lg_inv = (float)(1.0 / side);
        const size_t n = (size_t)G.lg_dim[0] * G.lg_dim[1] * G.lg_dim[2];
        G.d_lgrid.reserve((n + 3) / 4 * 4);
        launch_light_grid(s, G.d_ltri.p, c->n_ltri, G.d_lgrid.p, G.lg_lo, G.lg_inv, G.lg_dim, margin);
        G.lg_on = true;
        return;
    }
}

NrcNet cache_net(const vrgi_ctx *c) {
    NrcNet n{};
    n.log2_T = c->cache.cfg.log2_table;
    std::memcpy(n.res, c->cache.res, sizeof(n.res));
    for (int k = 0; k < 3; ++k) { n.lo[k] = c->cache.lo[k]; n.hi[k] = c->cache.hi[k]; }
    n.table = c->cache.table.p;
    n.wt = c->cache.wt.p;
    n.wb = c->cache.wb.p;
    n.prec = c->cache.cfg.precision;
    return n;
}

// the cache box defaults to the geometry's root box (union of the root's two children)
void cache_box_from_scene(vrgi_ctx *c) {
    auto &C = c->cache;
    if (C.box_ready) return;
    if (c->n_tris >= 2) {
        WideNode w;
        HIP_CHECK(hipMemcpyAsync(&w, c->G().d_gwide.p, sizeof(w), hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        const float lo[3] = {std::min(w.a.x, w.b.z), std::min(w.a.y, w.b.w), std::min(w.a.z, w.c.x)};
        const float hi[3] = {std::max(w.a.w, w.c.y), std::max(w.b.x, w.c.z), std::max(w.b.y, w.c.w)};
        for (int k = 0; k < 3; ++k) { C.lo[k] = lo[k]; C.hi[k] = hi[k]; }
    } else {
        for (int k = 0; k < 3; ++k) { C.lo[k] = 0.0f; C.hi[k] = 1.0f; }
    }
    C.box_ready = true;
}

bool cache_dynamic_scale(const vrgi_cache_config &cfg) {
    return cfg.precision == VRGI_CACHE_FP16 || (cfg.flags & VRGI_CACHE_DYNAMIC_SCALE);
}

// one training step over records rec[k + i * steps] (at most `batch`): gradients only, or
// gradients + Adam + repacking of the bf16 weights
void cache_step(vrgi_ctx *c, hipStream_t s, const float *rec, const uint32_t *count, uint32_t cap, uint32_t k,
                uint32_t steps, uint32_t batch, bool apply) {
    auto &C = c->cache;
    NrcTrain tr{};
    tr.rec = rec;
    tr.count = count;
    tr.rec_cap = cap;
    tr.step_k = k;
    tr.steps = steps;
    tr.max_batch = batch;
    tr.loss = 0;
    tr.loss_scale = C.cfg.loss_scale;
    tr.xplane = C.xplane.p;
    tr.gplane = C.gplane.p;
    tr.gw = C.gw.p;
    tr.gtable = C.gtable.p;
    tr.loss_out = C.loss.p;
    if (!apply) {  // gradients only (a test hook): at the fixed configured scale
        launch_nrc_grads(s, cache_net(c), tr);
        return;
    }
    if (cache_dynamic_scale(C.cfg)) {
        tr.scale_state = C.scale_state.p;
        tr.scale_interval = VRGI_CACHE_SCALE_INTERVAL;
    }
    C.steps += 1;
    NrcAdam ad{};
    ad.lr = C.cfg.lr;
    ad.beta1 = C.cfg.beta1;
    ad.beta2 = C.cfg.beta2;
    ad.eps = C.cfg.eps;
    ad.c1 = (float)(1.0 - std::pow((double)C.cfg.beta1, (double)C.steps));
    ad.c2 = (float)(1.0 - std::pow((double)C.cfg.beta2, (double)C.steps));
    ad.inv_scale = 1.0f / C.cfg.loss_scale;
    launch_nrc_train_step(s, cache_net(c), tr, C.w32.p, C.mw.p, C.vw.p, C.table.p, C.mt.p, C.vt.p,
                          C.n_table, ad);
}

// scratch of NEE stream slot k for the enumeration / walk passes of one NEE pdf launch
#ifndef VRGI_NEE_PASSES_ALL
#define VRGI_NEE_PASSES_ALL 0
#endif
void nee_passes(vrgi_ctx *c, hipStream_t s, uint32_t k, BounceArgs &ba) {
    const size_t P = c->P;
    c->nee_cnt[k].reserve(P);
    c->nee_lid[k].reserve(P * VRGI_NEE_RECS);
    c->nee_uv[k].reserve(P * VRGI_NEE_RECS);
    c->nee_over[k].reserve(P);
    c->nee_oc[k].reserve(1);
    HIP_CHECK(hipMemsetAsync(c->nee_oc[k].p, 0, 4, s));
    ba.nee_cnt = c->nee_cnt[k].p;
    ba.nee_lid = c->nee_lid[k].p;
    ba.nee_uv = c->nee_uv[k].p;
    ba.nee_over = c->nee_over[k].p;
    ba.nee_over_count = c->nee_oc[k].p;
}

void alloc_frame(vrgi_ctx *c) {
    size_t P = c->P, B = c->B;
    c->f_origin.reserve(B * P * 3);
    c->f_dir.reserve(B * P * 3);
    c->f_normal.reserve(B * P * 3);
    c->f_emis.reserve(B * P * 3);
    c->f_refl.reserve(B * P * 3);
    c->f_mis.reserve(B * P);
    c->f_pdf.reserve(B * P);
    c->f_nee.reserve(B * P);
    c->f_hit_t.reserve(B * P);
    c->f_hit.reserve(B * P);
    c->f_plen.reserve(P);
    c->f_queue.reserve((B + 1) * P);
    if (c->cache.on) {
        c->cache.albedo.reserve(B * P * 3);
        c->cache.d0.reserve(P * 3);
        auto &C = c->cache;
        const uint32_t T = C.cfg.fallback_tile;
        C.tiles_x = (c->W + T - 1) / T;
        C.tiles_y = (c->H + T - 1) / T;
        const size_t nt = (size_t)C.tiles_x * C.tiles_y;
        C.t_acc.reserve(3 * nt);
        C.t_ema.reserve(2 * nt);
        C.t_fb.reserve(nt);
        HIP_CHECK(hipMemsetAsync(C.t_acc.p, 0, 12 * nt, c->stream));
        HIP_CHECK(hipMemsetAsync(C.t_ema.p, 0, 8 * nt, c->stream));
        HIP_CHECK(hipMemsetAsync(C.t_fb.p, 0, nt, c->stream));
        c->cache.qrec.reserve(P * 16);
        c->cache.crad.reserve(P * 3);
    }
    if (c->cfg.ray_sort_bits) {
        c->f_squeue.reserve((B + 1) * P);
        c->f_skeys.reserve(P);
        c->f_shist.reserve((size_t)8 << (3 * c->cfg.ray_sort_bits));
    }
    c->f_qcount.reserve(B + 2);
    c->f_radiance.reserve(P * 3);
    c->f_lrun.reserve(P * 3);
    c->f_view_acc.reserve(P * 3);
    c->view_dirty = true;
    c->f_nlist.reserve(P);
    c->f_mlist.reserve(P);
    c->f_ncount.reserve(32);
    c->f_restir.reserve(P * 3);
    c->f_debug.reserve(P * 3);
    c->f_temporal.reserve(3 * P * 20);
    c->f_spatial.reserve(P * 20);
    c->f_bgra.reserve(P * 4);
    HIP_CHECK(hipMemsetAsync(c->f_temporal.p, 0, 3 * P * 20 * 4, c->stream));
    HIP_CHECK(hipMemsetAsync(c->f_spatial.p, 0, P * 20 * 4, c->stream));
    if (c->cfg.restir_mode == 1) {
        c->di_T.reserve(P * 8);
        c->di_hist.reserve(2 * P * 8);
        c->di_w.reserve(P * 4);
        c->di_light.reserve(P);
        // an empty history: light = 0xFFFFFFFF in word 0 of every pixel
        HIP_CHECK(hipMemsetAsync(c->di_hist.p, 0xFF, 2 * P * 8 * 4, c->stream));
        HIP_CHECK(hipMemsetAsync(c->di_w.p, 0, P * 16, c->stream));
        HIP_CHECK(hipMemsetAsync(c->di_light.p, 0xFF, P * 4, c->stream));
    }
    c->have_frame = false;
}

// instance table, per-instance pointers, light TL-BVH (scene.rs:243-338)
void rebuild_instances(vrgi_ctx *c) {
    // instance data is read by the builds and rays of frames still in flight
    HIP_CHECK(hipStreamSynchronize(c->up_stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->inst_objs.clear();
    std::vector<DevInst> hi;
    std::vector<const vrgi_vertex3d *> vp;
    std::vector<const vrgi_light_vertex3d *> lp;
    std::vector<uint32_t> loff;
    std::vector<float> tl_aabb, tl_xf, tl_pw;
    std::vector<uint32_t> tl_ids;
    uint32_t tri_off = 0, light_off = 0;
    for (auto &kv : c->objects) {
        vrgi_object *o = kv.second.first;
        if (!o) continue;  // the Empty handle (scene.rs:156-158)
        uint32_t i = (uint32_t)c->inst_objs.size();
        c->inst_objs.push_back(o);
        DevInst d{};
        std::memcpy(d.xf, kv.second.second.data(), 48);
        d.tri_off = tri_off;
        d.tl_idx = 0;  // prim2node.get(i).unwrap_or(0)
        d.lin_id = (d.xf[0] == 1.0f && d.xf[1] == 0.0f && d.xf[2] == 0.0f && d.xf[3] == 0.0f && d.xf[4] == 1.0f &&
                    d.xf[5] == 0.0f && d.xf[6] == 0.0f && d.xf[7] == 0.0f && d.xf[8] == 1.0f) ? 1u : 0u;
        d.light_bvh = o->d_light_bvh;
        d.lpairs = o->d_lpairs;
        hi.push_back(d);
        vp.push_back(o->d_verts);
        lp.push_back(o->d_light_verts);
        loff.push_back(light_off);
        tri_off += o->n_tris;
        light_off += o->n_light;
        if (!o->light_empty) {
            tl_aabb.insert(tl_aabb.end(), o->light_aabb, o->light_aabb + 6);
            tl_xf.insert(tl_xf.end(), kv.second.second.begin(), kv.second.second.end());
            tl_pw.push_back(o->power);
            tl_ids.push_back(i);
        }
    }
    c->n_inst = (uint32_t)hi.size();
    c->n_tris = tri_off;
    c->n_ltri = light_off;
    hipStream_t s = c->stream;
    c->d_inst.reserve(hi.size());
    c->d_vtxptr.reserve(vp.size());
    c->d_lvtxptr.reserve(lp.size());
    c->d_light_off.reserve(loff.size());
    if (!hi.empty()) {
        HIP_CHECK(hipMemcpyAsync(c->d_inst.p, hi.data(), hi.size() * sizeof(DevInst), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(c->d_vtxptr.p, vp.data(), vp.size() * sizeof(void *), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(c->d_lvtxptr.p, lp.data(), lp.size() * sizeof(void *), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(c->d_light_off.p, loff.data(), loff.size() * 4, hipMemcpyHostToDevice, s));
    }
    uint32_t m = (uint32_t)tl_ids.size();
    c->d_tl.reserve(m ? 2 * m - 1 : 1);
    c->d_tlpairs.reserve(m ? 2 * m - 1 : 1);
    if (m == 0) {
        vrgi_bvh_node dummy;  // BvhNode::default() (bvh/mod.rs:27-38, scene.rs:275-277)
        dummy.left = dummy.right_or_prim = dummy.parent = VRGI_INVALID;
        std::memset(dummy.min, 0, 12);
        std::memset(dummy.max, 0, 12);
        dummy.power = 0.0f;
        HIP_CHECK(hipMemcpyAsync(c->d_tl.p, &dummy, sizeof(dummy), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipStreamSynchronize(s));
        c->n_tl = 1;
        return;
    }
    c->d_tl_aabb.reserve(6 * m);
    c->d_tl_xf.reserve(12 * m);
    c->d_tl_power.reserve(m);
    c->d_tl_ids.reserve(m);
    c->d_tl_prim9.reserve(9 * m);
    c->d_tl_trees.reserve(3);
    c->d_tl_depth.reserve(1);
    HIP_CHECK(hipMemcpyAsync(c->d_tl_aabb.p, tl_aabb.data(), tl_aabb.size() * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->d_tl_xf.p, tl_xf.data(), tl_xf.size() * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->d_tl_power.p, tl_pw.data(), m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->d_tl_ids.p, tl_ids.data(), m * 4, hipMemcpyHostToDevice, s));
    uint32_t trees[3] = {0, m, 0};
    HIP_CHECK(hipMemcpyAsync(c->d_tl_trees.p, trees, 12, hipMemcpyHostToDevice, s));
    launch_tl_inputs(s, c->d_tl_aabb.p, c->d_tl_xf.p, m, c->d_tl_prim9.p);
    if (c->tl_builder.build(s, c->d_tl_prim9.p, c->d_tl_power.p, c->d_tl_ids.p, m, 1, c->d_tl_trees.p,
                            c->d_tl_trees.p + 1, c->d_tl_trees.p + 2, trees + 2, c->d_tl.p,
                            c->d_tl_depth.p, nullptr) != 0)
        throw HipError("light top-level build exceeded its workspace");
    c->n_tl = 2 * m - 1;
    launch_tl_idx(s, c->d_tl.p, c->n_tl, c->d_inst.p);
    launch_light_pairs(s, c->d_tl.p, c->n_tl, c->d_tlpairs.p);
    HIP_CHECK(hipStreamSynchronize(s));
}

void build_geometry_on(vrgi_ctx *c, hipStream_t s, bool replay = false);

// adds a rendered frame's stage times (events of parity k) to the profile output
void harvest_frame(vrgi_ctx *c, uint32_t k) {
    if (!c->pend[k]) return;
    const size_t B = c->B;
    std::vector<hipEvent_t> &ev = c->events[k];
    HIP_CHECK(hipEventSynchronize(ev[2 + 2 * B]));
    auto el = [&](size_t a, size_t b) {
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, ev[a], ev[b]));
        return (double)ms;
    };
    vrgi_profile_out &p = c->prof;
    const uint32_t *q = c->h_qcount + (size_t)k * (2 * B + 2);
    if (c->pend_built[k]) p.build_ms += el(3 + 3 * B, 4 + 3 * B);
    for (size_t b = 0; b < B && b < 16; ++b) {
        p.trace_ms[b] += el(b == 0 ? 1 : 1 + b, 2 + b);
        if (b + 1 < B) p.nee_ms[b] += el(3 + 2 * B + b, 2 + B + b);
        p.rays[b] += b == 0 ? c->pend_P[k] : q[b];
        if (b + 1 < B) p.nee_listed[b] += q[B + 2 + b];
    }
    p.resolve_ms += el(2 + B + (B - 1), 2 + 2 * B);
    p.frame_ms += el(0, 2 + 2 * B);
    p.frames += 1;
    c->pend[k] = false;
}

// world triangles + world light triangles, geometry and light-geometry BVHs in one
// batched build, traversal copies
void rebuild_geometry(vrgi_ctx *c, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, bool replay = false) {
    const bool pipe = !(c->cfg.flags & VRGI_FLAG_SERIAL_BUILD);
    if (pipe) {
        // the other set; the frame that read it last must have ended (its buffers may
        // be reallocated below)
        c->gi ^= 1u;
        if (c->G().freed_once) HIP_CHECK(hipEventSynchronize(c->G().ev_free));
    }
    hipStream_t s = pipe ? c->up_stream : c->stream;
    if (e0) HIP_CHECK(hipEventRecord(e0, s));
    build_geometry_on(c, s, replay);
    if (e1) HIP_CHECK(hipEventRecord(e1, s));
    if (pipe) {
        HIP_CHECK(hipEventRecord(c->ev_geo_ready, s));
        HIP_CHECK(hipStreamWaitEvent(c->stream, c->ev_geo_ready, 0));
    }
}

// With `replay`, a rebuild of geometry unchanged since the last build: when that build's
// plan is at hand, the builder repeats its launch sequence and the host reads nothing back
// (depths, leaf-depth sum and light root box are the plan's), so the frame's launches are
// queued without waiting for the build.
void build_geometry_on(vrgi_ctx *c, hipStream_t s, bool replay) {
    c->ext_gnodes = c->ext_lnodes = nullptr;
    uint32_t nt = c->n_tris, nl = c->n_ltri, N = nt + nl;
    const uint32_t n_trees = (nt ? 1u : 0u) + (nl ? 1u : 0u);
    const BuildPlan &bp = c->geo_builder.plan();
    const bool rep = replay && c->geo_plan.valid && bp.valid && bp.n == N && bp.trees == n_trees;
    if (rep && c->check_pending) {  // the previous replay's check (its copy ended long ago)
        HIP_CHECK(hipEventSynchronize(c->ev_check));
        c->check_pending = false;
        if (*c->h_check)
            throw std::runtime_error("a replayed geometry build left its plan (the geometry changed unnoticed)");
    }
    c->G().d_tri.reserve(nt);
    c->G().d_tri_inst.reserve(nt);
    c->G().d_ltri.reserve(nl);
    c->G().d_ltri_inst.reserve(nl);
    c->G().d_ltri_node.reserve(nl);
    c->d_prim9.reserve((size_t)9 * N);
    c->d_power.reserve(N);
    c->d_ids.reserve(N);
    uint32_t gnodes = nt ? 2 * nt - 1 : 0, lnodes = nl ? 2 * nl - 1 : 0;
    c->G().d_nodes.reserve(gnodes + lnodes + 1);
    c->G().d_gwide.reserve(nt > 1 ? nt - 1 : 1);
    c->G().d_lwide.reserve(nl > 1 ? nl - 1 : 1);
    c->G().d_lwide4.reserve(nl > 1 ? nl - 1 : 1);
    if (VRGI_GEOM4) c->G().d_gwide4.reserve(nt > 1 ? nt - 1 : 1);
    c->d_trees.reserve(6);
    c->G().d_depth.reserve(2);
    c->G().geo_depth = c->G().lgeo_depth = 0;
    c->G().lg_on = false;
    if (N == 0) return;
    FlattenArgs f{};
    f.inst = c->d_inst.p;
    f.n_inst = c->n_inst;
    f.vtx = c->d_vtxptr.p;
    f.lvtx = c->d_lvtxptr.p;
    f.light_off = c->d_light_off.p;
    f.prim_stride = N;
    f.n_tris = nt;
    f.tri = c->G().d_tri.p;
    f.tri_inst = c->G().d_tri_inst.p;
    f.prim9 = c->d_prim9.p;
    f.power = c->d_power.p;
    f.ids = c->d_ids.p;
    f.prim_off = 0;
    if (c->tex_on) {
        c->G().d_tri_uv.reserve((size_t)3 * nt);
        f.tri_uv = c->G().d_tri_uv.p;
    }
    launch_flatten(s, f);
    FlattenArgs fl = f;
    fl.n_tris = nl;
    fl.tri = c->G().d_ltri.p;
    fl.tri_inst = c->G().d_ltri_inst.p;
    fl.light_node = c->G().d_ltri_node.p;
    fl.prim_off = nt;
    launch_flatten_lights(s, fl);
    c->G().d_lanc.reserve((size_t)nl * VRGI_LANC_MAX);
    c->G().d_lanc_len.reserve(nl);
    launch_light_ancestry(s, c->G().d_ltri_inst.p, c->G().d_ltri_node.p, c->d_inst.p, c->d_tl.p, nl, c->G().d_lanc.p,
                          c->G().d_lanc_len.p);
    uint32_t trees[6];
    uint32_t nt_trees = 0;
    uint32_t beg[2], cnt[2], base[2];
    if (nt) { beg[nt_trees] = 0; cnt[nt_trees] = nt; base[nt_trees] = 0; ++nt_trees; }
    if (nl) { beg[nt_trees] = nt; cnt[nt_trees] = nl; base[nt_trees] = gnodes; ++nt_trees; }
    for (uint32_t t = 0; t < nt_trees; ++t) { trees[t] = beg[t]; trees[2 + t] = cnt[t]; trees[4 + t] = base[t]; }
    if (!rep) HIP_CHECK(hipMemcpyAsync(c->d_trees.p, trees, 24, hipMemcpyHostToDevice, s));  // else unchanged
    if (rep && !c->h_check) {
        c->d_check.reserve(1);
        HIP_CHECK(hipMemsetAsync(c->d_check.p, 0, 4, s));
        HIP_CHECK(hipHostMalloc((void **)&c->h_check, 4, hipHostMallocDefault));
        *c->h_check = 0;
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_check, hipEventDisableTiming));
    }
    if (c->geo_builder.build(s, c->d_prim9.p, c->d_power.p, c->d_ids.p, N, nt_trees, c->d_trees.p,
                             c->d_trees.p + 2, c->d_trees.p + 4, base, c->G().d_nodes.p, c->G().d_depth.p,
                             nullptr, rep, rep ? c->d_check.p : nullptr) != 0)
        throw HipError("geometry build exceeded its workspace");
    uint32_t dep[2] = {0, 0};
    unsigned long long S = 0;
    vrgi_bvh_node lroot{};
    if (rep) {
        std::memcpy(dep, c->geo_plan.dep, sizeof(dep));
        S = c->geo_plan.S;
        lroot = c->geo_plan.lroot;
        HIP_CHECK(hipMemcpyAsync(c->h_check, c->d_check.p, 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(c->ev_check, s));
        c->check_pending = true;
    } else {
        HIP_CHECK(hipMemcpyAsync(dep, c->G().d_depth.p, nt_trees * 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipMemcpyAsync(&S, c->geo_builder.leaf_depth_sum(), 8, hipMemcpyDeviceToHost, s));
        if (nl) HIP_CHECK(hipMemcpyAsync(&lroot, c->G().d_nodes.p + gnodes, sizeof(lroot), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        std::memcpy(c->geo_plan.dep, dep, sizeof(dep));
        c->geo_plan.S = S;
        c->geo_plan.lroot = lroot;
        c->geo_plan.valid = true;
    }
    build_light_grid(c, s, lroot.min, lroot.max);
    // structural figures of the build (profile output)
    c->prof.build_leaf_depth_sum = S;
    c->prof.build_prims = (uint64_t)nt + nl;
    c->prof.build_nodes = (uint64_t)gnodes + lnodes;
    uint32_t t = 0;
    if (nt) c->G().geo_depth = dep[t++];
    if (nl) c->G().lgeo_depth = dep[t++];
    if (geom_stack_need(c->G().geo_depth) + 2 > VRGI_MAX_STACK || 3 * ((c->G().lgeo_depth + 1) / 2) + 2 > VRGI_MAX_STACK)
        throw std::runtime_error("BVH deeper than the traversal stack (geometry depth " + std::to_string(c->G().geo_depth) + ", light-geometry depth " + std::to_string(c->G().lgeo_depth) + ")");
    launch_wide(s, c->G().d_nodes.p, gnodes, c->G().d_gwide.p);
    if (VRGI_GEOM4) launch_wide4(s, c->G().d_nodes.p, gnodes, c->G().d_gwide4.p);
    launch_wide(s, c->G().d_nodes.p + gnodes, lnodes, c->G().d_lwide.p);
    launch_wide4(s, c->G().d_nodes.p + gnodes, lnodes, c->G().d_lwide4.p);
}

void free_object(vrgi_object *o) {
    if (o->life) {
        // frames (and the next frame's build) that may still read the object end first
        std::lock_guard<std::mutex> lk(o->life->mu);
        (void)hipSetDevice(o->life->device);
        if (o->life->alive)
            for (hipStream_t st : o->life->streams)
                if (st) (void)hipStreamSynchronize(st);
    }
    if (!o->arena) {
        if (o->d_verts) (void)hipFree(o->d_verts);
        if (o->d_light_bvh) (void)hipFree(o->d_light_bvh);
        if (o->d_lpairs) (void)hipFree(o->d_lpairs);
        if (o->d_light_verts) (void)hipFree(o->d_light_verts);
    }
    delete o;  // drops its share of the arena
}

void release(vrgi_object *o) {
    if (o && o->refs.fetch_sub(1) == 1) free_object(o);
}

// device memory of the chunk objects of one voxel-world commit
struct VxArena {
    DevBuf<vrgi_vertex3d> verts;
    DevBuf<float> tri9, power;
    DevBuf<uint32_t> ids;
    DevBuf<vrgi_bvh_node> nodes;
    DevBuf<LPair> lpairs;
    DevBuf<vrgi_light_vertex3d> lverts;
};
// arenas whose objects have all been released, for the next commits; shared by the
// context and the arenas' deleters (an object may outlive its context)
struct VxArenaPool {
    std::mutex mu;
    std::vector<VxArena *> free_list;
    ~VxArenaPool() {
        for (VxArena *a : free_list) delete a;
    }
};

// Re-meshes the dirty chunks of the voxel world and replaces their objects: what the
// reference's chunk manager does per dirty chunk (chunk_manager.rs:338-399: gen_mesh,
// upload_object, add_object), for all of them at once. Vertices, the emissive gather,
// the light BL-BVHs (one multi-tree build) and the light vertices in leaf order
// (scene.rs:509-524) stay on the device; each object's (box, power) comes from its
// tree's root, the one readback. The commit's arena is taken from a pool of arenas whose
// objects have all been released (their buffers only grow), so a commit per frame does
// not allocate and free device memory each time.
void voxels_commit(vrgi_ctx *c) {
    auto &V = c->vx;
    if (!V.on || !V.pending) return;
    hipStream_t s = c->stream;
    HIP_CHECK(hipStreamSynchronize(c->up_stream));
    const uint32_t nx = V.nc[0], ny = V.nc[1], nz = V.nc[2], nch = nx * ny * nz;
    std::vector<uint8_t> dirty(nch);
    HIP_CHECK(hipMemcpyAsync(dirty.data(), V.dirty.p, nch, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    V.pending = false;
    // every earlier frame has ended: objects replaced before now can go
    for (vrgi_object *o : c->old) release(o);
    c->old.clear();
    std::vector<uint32_t> coords, keys;
    for (uint32_t k = 0; k < nch; ++k)
        if (dirty[k]) {
            keys.push_back(k);
            coords.push_back(k / (ny * nz));
            coords.push_back((k / nz) % ny);
            coords.push_back(k % nz);
        }
    if (keys.empty()) return;
    HIP_CHECK(hipMemsetAsync(V.dirty.p, 0, nch, s));
    const uint32_t n = (uint32_t)keys.size();
    V.chunks.reserve(3 * (size_t)n);
    V.mask.reserve((size_t)n * 32768);
    V.face_off.reserve((size_t)n * 32768);
    V.lface_off.reserve((size_t)n * 32768);
    V.faces.reserve(n);
    V.lfaces.reserve(n);
    V.vert_off.reserve(n);
    V.light_off.reserve(n);
    HIP_CHECK(hipMemcpyAsync(V.chunks.p, coords.data(), (size_t)12 * n, hipMemcpyHostToDevice, s));
    MeshArgsHost h{};
    h.blocks = V.blocks.p;
    for (int k = 0; k < 3; ++k) h.dims[k] = V.dims[k];
    h.air = V.n_types;
    h.translucent = V.transl.p;
    h.chunks = V.chunks.p;
    h.n_chunks = n;
    h.mask = V.mask.p;
    h.face_off = V.face_off.p;
    h.faces = V.faces.p;
    launch_mesh_count(s, h);
    launch_light_scan(s, h, V.em.p, V.lface_off.p, V.lfaces.p);
    std::vector<uint32_t> faces(n), lfaces(n);
    HIP_CHECK(hipMemcpyAsync(faces.data(), V.faces.p, (size_t)4 * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(lfaces.data(), V.lfaces.p, (size_t)4 * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint64_t> voff(n), loff(n);
    std::vector<uint32_t> beg, cnt, base, tree_of(n, VRGI_INVALID);
    uint64_t nv = 0, nl = 0, nodes = 0;
    for (uint32_t i = 0; i < n; ++i) {
        voff[i] = nv;
        loff[i] = nl;
        nv += 6ull * faces[i];
        const uint64_t m = 2ull * lfaces[i];
        if (m) {
            tree_of[i] = (uint32_t)beg.size();
            beg.push_back((uint32_t)nl);
            cnt.push_back((uint32_t)m);
            base.push_back((uint32_t)nodes);
            nodes += 2 * m - 1;
        }
        nl += m;
    }
    if (nl >= (1ull << 31) || nodes >= (1ull << 32) - 1 || nv >= (1ull << 40))
        throw std::runtime_error("voxel commit above 2^31 light triangles");
    if (!V.pool) V.pool = std::make_shared<VxArenaPool>();
    VxArena *raw = nullptr;
    {
        std::lock_guard<std::mutex> lk(V.pool->mu);
        if (!V.pool->free_list.empty()) {
            raw = V.pool->free_list.back();
            V.pool->free_list.pop_back();
        }
    }
    if (!raw) raw = new VxArena();
    auto pool = V.pool;
    std::shared_ptr<VxArena> A(raw, [pool](VxArena *a) {
        std::lock_guard<std::mutex> lk(pool->mu);
        pool->free_list.push_back(a);
    });
    A->verts.reserve(nv);
    A->tri9.reserve(9 * nl);
    A->power.reserve(nl);
    A->ids.reserve(nl);
    A->nodes.reserve(nodes);
    A->lpairs.reserve(nodes);
    A->lverts.reserve(3 * nl);
    V.prim9.reserve(9 * nl);
    HIP_CHECK(hipMemcpyAsync(V.vert_off.p, voff.data(), (size_t)8 * n, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(V.light_off.p, loff.data(), (size_t)8 * n, hipMemcpyHostToDevice, s));
    h.vert_off = V.vert_off.p;
    h.out = A->verts.p;
    if (nv) launch_mesh_emit(s, h);
    launch_light_gather(s, h, V.lum.p, V.em.p, V.lface_off.p, V.light_off.p, A->tri9.p, A->power.p, V.prim9.p,
                        A->ids.p, (uint32_t)nl);
    const uint32_t nT = (uint32_t)beg.size();
    std::vector<vrgi_bvh_node> roots(nT);
    if (nT) {
        std::vector<uint32_t> tr(3 * (size_t)nT);
        for (uint32_t t = 0; t < nT; ++t) { tr[t] = beg[t]; tr[nT + t] = cnt[t]; tr[2 * nT + t] = base[t]; }
        V.trees.reserve(tr.size());
        V.depth.reserve(nT);
        V.tree_base.reserve(nT + 1);
        std::vector<uint32_t> tb(base);
        tb.push_back((uint32_t)nodes);
        HIP_CHECK(hipMemcpyAsync(V.trees.p, tr.data(), tr.size() * 4, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemcpyAsync(V.tree_base.p, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, s));
        if (V.builder.build(s, V.prim9.p, A->power.p, A->ids.p, (uint32_t)nl, nT, V.trees.p, V.trees.p + nT,
                            V.trees.p + 2 * nT, base.data(), A->nodes.p, V.depth.p, nullptr) != 0)
            throw HipError("voxel light build exceeded its workspace");
        launch_light_pairs_multi(s, A->nodes.p, V.tree_base.p, nT, (uint32_t)nodes, A->lpairs.p);
        // light vertices in leaf order (each leaf's right_or_prim is its triangle index in
        // the chunk, found among the chunk's lights, which are in triangle order) and the
        // trees' roots, on the device; only the roots come back
        V.rank.reserve(2 * (size_t)nodes);
        const size_t sb = light_verts_scratch_bytes((uint32_t)nodes);
        V.scan_tmp.reserve(sb);
        V.roots.reserve(nT);
        launch_light_verts_multi(s, A->nodes.p, (uint32_t)nodes, V.tree_base.p, V.trees.p, nT, A->ids.p, A->tri9.p,
                                 V.rank.p, V.scan_tmp.p, sb, A->lverts.p, V.roots.p);
        HIP_CHECK(hipMemcpyAsync(roots.data(), V.roots.p, nT * sizeof(vrgi_bvh_node), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
    }
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t key = keys[i];
        vrgi_object *o = nullptr;
        if (faces[i]) {
            o = new vrgi_object();
            o->life = c->life;
            o->arena = A;
            o->n_tris = 2 * faces[i];
            o->d_verts = A->verts.p + voff[i];
            const uint32_t t = tree_of[i];
            if (t != VRGI_INVALID) {
                const uint32_t m = cnt[t];
                o->n_light = m;
                o->d_light_bvh = A->nodes.p + base[t];
                o->d_lpairs = A->lpairs.p + base[t];
                o->d_light_verts = A->lverts.p + 3 * loff[i];
                // the (Aabb, power) build_bl_bvh returns (build.rs:359-371): the root's box
                // and power (a single-leaf root holds its triangle's box and power)
                const vrgi_bvh_node &r = roots[t];
                o->light_empty = false;
                for (int k = 0; k < 3; ++k) { o->light_aabb[k] = r.min[k]; o->light_aabb[3 + k] = r.max[k]; }
                o->power = r.power;
            }
        }
        auto it = c->objects.find(key);
        if (it != c->objects.end()) {
            if (it->second.first) c->old.push_back(it->second.first);
            c->objects.erase(it);
        }
        const float cx = (float)(32 * coords[3 * i]), cy = (float)(32 * coords[3 * i + 1]);
        const float cz = (float)(32 * coords[3 * i + 2]);
        c->objects[key] = {o, std::vector<float>{1, 0, 0, 0, 1, 0, 0, 0, 1, cx, cy, cz}};
    }
    V.commits += 1;
    V.chunks_remeshed += n;
    c->inst_dirty = true;
    mark_geo_dirty(c);
}

bool valid_xf(const float *xf) {
    if (!xf) return false;
    for (int k = 0; k < 12; ++k)
        if (!(xf[k] == xf[k])) return false;
    return true;
}

// ---------------------------------------------------------------- allocation-free build
// vrgi_build_bl_bvh_noalloc works in `scratch`, one array per GPU that lives until the
// process ends: the staging of its inputs and outputs (below) and the builder workspace
// of up to VRGI_NOALLOC_MAX_PRIMS primitives after them. Host code uses the per-device
// address hipGetSymbolAddress returns (noalloc_setup), never the array's name.
constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t NA_N = VRGI_NOALLOC_MAX_PRIMS;
constexpr size_t NA_TRI = 0;                                   // 9 floats per triangle
constexpr size_t NA_PRIM9 = NA_TRI + na_round(36 * NA_N);      // builder input planes
constexpr size_t NA_LUM = NA_PRIM9 + na_round(36 * NA_N);      // powers
constexpr size_t NA_IDS = NA_LUM + na_round(4 * NA_N);         // primitive ids
constexpr size_t NA_ORDER = NA_IDS + na_round(4 * NA_N);       // primitive order out
constexpr size_t NA_NODES = NA_ORDER + na_round(4 * NA_N);     // 2n - 1 nodes out
constexpr size_t NA_TREES = NA_NODES + na_round(40 * (2 * NA_N - 1));  // tree begin / count / base, depth
constexpr size_t NA_BUILDER = NA_TREES + 256;                  // builder workspace
constexpr size_t NA_BUILDER_BYTES = 8u << 20;
constexpr size_t NA_BYTES = NA_BUILDER + NA_BUILDER_BYTES;     // about 18 MiB

__device__ __align__(256) unsigned char scratch[NA_BYTES];

struct NoAllocDev {
    char *ptr = nullptr;  // this device's copy of `scratch`
    Builder builder;      // bound to ptr + NA_BUILDER
};
constexpr int NA_MAX_DEVICES = 64;
NoAllocDev g_na[NA_MAX_DEVICES];
std::mutex g_na_mu;

// once per device, with that device current (vrgi_create)
void noalloc_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_na_mu);
    NoAllocDev &d = g_na[device];
    if (d.ptr) return;
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (Builder::workspace_bytes(VRGI_NOALLOC_MAX_PRIMS, 1) > NA_BUILDER_BYTES)
        throw HipError("allocation-free scratch too small for its builder workspace");
    d.builder.bind((char *)p + NA_BUILDER, NA_BUILDER_BYTES, VRGI_NOALLOC_MAX_PRIMS, 1);
    d.ptr = (char *)p;
}

// the (Aabb, f32) build_bl_bvh returns (build.rs:359-371), from the host nodes
void bl_root(const float *tri, const float *lum, const vrgi_bvh_node *nodes, float *root_box, float *root_power) {
    if (!root_box && !root_power) return;
    float rb[6], rp;
    if (nodes[0].left == VRGI_INVALID) {
        for (int k = 0; k < 3; ++k) {
            float a = tri[k] + 0.0f, b = tri[3 + k] + 0.0f, d = tri[6 + k] + 0.0f;
            rb[k] = std::min(std::min(a, b), d);
            rb[3 + k] = std::max(std::max(a, b), d);
        }
        rp = lum[0];
    } else {
        for (int k = 0; k < 3; ++k) { rb[k] = nodes[0].min[k]; rb[3 + k] = nodes[0].max[k]; }
        rp = nodes[0].power;
    }
    if (root_box) std::memcpy(root_box, rb, 24);
    if (root_power) *root_power = rp;
}

}  // namespace

vrgi_ctx::~vrgi_ctx() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (up_stream) (void)hipStreamSynchronize(up_stream);
    for (int k = 0; k < 2; ++k)
        for (hipEvent_t e : events[k]) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; ++k)
        if (geo[k].ev_free) (void)hipEventDestroy(geo[k].ev_free);
    if (ev_geo_ready) (void)hipEventDestroy(ev_geo_ready);
    if (h_qcount) (void)hipHostFree(h_qcount);
    if (h_check) (void)hipHostFree(h_check);
    if (ev_check) (void)hipEventDestroy(ev_check);
    for (auto &kv : objects) release(kv.second.first);
    for (vrgi_object *o : old) release(o);
    {
        // objects still held by the caller no longer wait on this context's streams
        std::lock_guard<std::mutex> lk(life->mu);
        life->alive = false;
    }
    for (UpSlot *u : up_slots) {
        if (u->s) {
            (void)hipStreamSynchronize(u->s);
            (void)hipStreamDestroy(u->s);
        }
        delete u;
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (up_stream) (void)hipStreamDestroy(up_stream);
    if (side) (void)hipStreamSynchronize(side);
    if (side) (void)hipStreamDestroy(side);
    for (hipStream_t ns : nee_s)
        if (ns) {
            (void)hipStreamSynchronize(ns);
            (void)hipStreamDestroy(ns);
        }
    for (hipEvent_t e : ev_nee_done)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {ev_query, ev_infer, ev_train, ev_accum, ev_fstart, ev_snap, ev_held})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_bounce) (void)hipEventDestroy(e);
}

extern "C" {

const char *vrgi_last_error(void) { return g_err.c_str(); }
uint32_t vrgi_abi_version(void) { return VRGI_ABI_VERSION; }

void vrgi_config_reference(vrgi_config *cfg, uint32_t width, uint32_t height) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->width = width;
    cfg->height = height;
    cfg->num_bounces = 3;                // interactive_rendering.rs:767
    cfg->rr_start_bounce = 0xFFFFFFFFu;  // no Russian roulette in the reference
    cfg->rr_min_survival = 0.1f;
    cfg->restir_mode = 0;                // ReSTIR GI
    cfg->restir_spatial_radius = 15.0f;  // restir_spatial_resampling.rs:314
    cfg->di_candidates = 32;
    cfg->seed = 0;
    cfg->flags = 0;
}

int vrgi_create(const vrgi_config *cfg, const vrgi_material *materials, const float *lum,
                uint32_t n_slots, int device, vrgi_ctx **out) {
    API_BEGIN
    if (!cfg || !out || (n_slots && (!materials || !lum))) return fail(VRGI_E_INVALID, "null argument");
    if (cfg->width == 0 || cfg->height == 0 || cfg->num_bounces < 2 || cfg->num_bounces > 16)
        return fail(VRGI_E_INVALID, "extent must be non-zero and num_bounces in [2, 16]");
    if ((uint64_t)cfg->width * cfg->height > (1ull << 26)) return fail(VRGI_E_LIMIT, "extent above 2^26 pixels");
    if (cfg->restir_mode > 1) return fail(VRGI_E_INVALID, "restir_mode must be 0 (GI) or 1 (DI)");
    if (cfg->ray_sort_bits > 5) return fail(VRGI_E_INVALID, "ray_sort_bits must be in [0, 5]");
    if (cfg->restir_mode == 1 && (cfg->di_candidates == 0 || cfg->di_candidates > 1024))
        return fail(VRGI_E_INVALID, "di_candidates must be in [1, 1024]");
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(VRGI_E_INVALID, "no such HIP device");
    vrgi_ctx *c = new vrgi_ctx();
    try {
        c->device = device;
        HIP_CHECK(hipSetDevice(device));
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&c->up_stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        c->life->device = device;
        c->life->streams[0] = c->stream;
        c->life->streams[1] = c->up_stream;
        for (int k = 0; k < VRGI_NEE_STREAMS; ++k) {
            HIP_CHECK(hipStreamCreateWithFlags(&c->nee_s[k], hipStreamNonBlocking));
            HIP_CHECK(hipEventCreateWithFlags(&c->ev_nee_done[k], hipEventDisableTiming));
        }
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_geo_ready, hipEventDisableTiming));
        for (int k = 0; k < 2; ++k) HIP_CHECK(hipEventCreateWithFlags(&c->geo[k].ev_free, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_query, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_infer, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_train, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_accum, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_fstart, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_snap, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&c->ev_held, hipEventDisableTiming));
        c->cfg = *cfg;
        c->W = cfg->width;
        c->H = cfg->height;
        c->P = c->W * c->H;
        c->B = cfg->num_bounces;
        c->mats.assign(materials, materials + n_slots);
        c->lum.assign(lum, lum + n_slots);
        c->d_mats.reserve(n_slots);
        if (n_slots)
            HIP_CHECK(hipMemcpy(c->d_mats.p, materials, n_slots * sizeof(vrgi_material), hipMemcpyHostToDevice));
        c->d_counters.reserve(1);
        HIP_CHECK(hipMemset(c->d_counters.p, 0, sizeof(Counters)));
        noalloc_setup(device);
        alloc_frame(c);
        rebuild_instances(c);
        HIP_CHECK(hipStreamSynchronize(c->stream));
    } catch (...) {
        delete c;
        throw;
    }
    *out = c;
    return VRGI_OK;
    API_END
}

int vrgi_destroy(vrgi_ctx *c) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (c->in_batch) return fail(VRGI_E_STATE, "the scene is in a batch (destroy the batch first)");
    delete c;
    return VRGI_OK;
    API_END
}

int vrgi_upload_object(vrgi_ctx *c, const vrgi_vertex3d *v, size_t n, vrgi_object **out) {
    API_BEGIN
    if (!c || !out || (n && !v)) return fail(VRGI_E_INVALID, "null argument");
    *out = nullptr;
    if (n == 0) return VRGI_OK;  // SceneUploadedObjectHandle::Empty
    if (n % 3 != 0) return fail(VRGI_E_INVALID, "vertex count is not a multiple of 3 (scene.rs:553)");
    if (n / 3 > (1u << 30)) return fail(VRGI_E_LIMIT, "more than 2^30 triangles in one object");
    for (size_t i = 0; i < n; i += 3)
        if (v[i].t >= c->lum.size()) return fail(VRGI_E_INVALID, "material slot out of range");
    HIP_CHECK(hipSetDevice(c->device));
    // an idle upload slot (a new one when every slot is busy): no lock is held while the
    // upload runs, so uploads from several threads overlap on their own streams
    UpSlot *slot = nullptr;
    {
        std::lock_guard<std::mutex> lk(c->up_mu);
        if (!c->up_idle.empty()) {
            slot = c->up_idle.back();
            c->up_idle.pop_back();
        } else {
            slot = new UpSlot();
            c->up_slots.push_back(slot);
            HIP_CHECK(hipStreamCreateWithFlags(&slot->s, hipStreamNonBlocking));
        }
    }
    struct SlotBack {
        vrgi_ctx *c;
        UpSlot *u;
        ~SlotBack() {
            std::lock_guard<std::mutex> lk(c->up_mu);
            c->up_idle.push_back(u);
        }
    } slot_back{c, slot};
    hipStream_t s = slot->s;
    vrgi_object *o = new vrgi_object();
    o->life = c->life;
    try {
        o->n_tris = (uint32_t)(n / 3);
        dev_alloc((void **)&o->d_verts, n * sizeof(vrgi_vertex3d));
        HIP_CHECK(hipMemcpyAsync(o->d_verts, v, n * sizeof(vrgi_vertex3d), hipMemcpyHostToDevice, s));
        // emissive triangles (scene.rs:555-571): the caller-side gather of the reference
        std::vector<float> tri9, pw;
        std::vector<uint32_t> ids;
        for (uint32_t i = 0; i < o->n_tris; ++i) {
            float l = c->lum[v[3 * i].t];
            if (!(l > 0.0f)) continue;
            float a[3], b[3], cc[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = v[3 * i].position[k] + 0.0f;
                b[k] = v[3 * i + 1].position[k] + 0.0f;
                cc[k] = v[3 * i + 2].position[k] + 0.0f;
            }
            float e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {cc[0] - a[0], cc[1] - a[1], cc[2] - a[2]};
            float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            float area = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2.0f;
            tri9.insert(tri9.end(), a, a + 3);
            tri9.insert(tri9.end(), b, b + 3);
            tri9.insert(tri9.end(), cc, cc + 3);
            pw.push_back(l * area);
            ids.push_back(i);
        }
        uint32_t m = (uint32_t)ids.size();
        o->n_light = m;
        std::vector<vrgi_light_vertex3d> lv;  // outlives its asynchronous upload
        if (m) {
            // light BL-BVH on the GPU (build_bl_bvh, build.rs:249-372)
            DevBuf<float> dtri, dprim9, dpw;
            DevBuf<uint32_t> dids, dtrees, ddepth;
            dtri.reserve(9 * (size_t)m);
            dprim9.reserve(9 * (size_t)m);
            dpw.reserve(m);
            dids.reserve(m);
            dtrees.reserve(3);
            ddepth.reserve(1);
            dev_alloc((void **)&o->d_light_bvh, (2 * (size_t)m - 1) * sizeof(vrgi_bvh_node));
            uint32_t trees[3] = {0, m, 0};
            HIP_CHECK(hipMemcpyAsync(dtri.p, tri9.data(), tri9.size() * 4, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(dpw.p, pw.data(), m * 4, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(dids.p, ids.data(), m * 4, hipMemcpyHostToDevice, s));
            HIP_CHECK(hipMemcpyAsync(dtrees.p, trees, 12, hipMemcpyHostToDevice, s));
            launch_prims_from_tris(s, dtri.p, m, dprim9.p);
            if (slot->b.build(s, dprim9.p, dpw.p, dids.p, m, 1, dtrees.p, dtrees.p + 1, dtrees.p + 2,
                                    trees + 2, o->d_light_bvh, ddepth.p, nullptr) != 0)
                throw HipError("light build exceeded its workspace");
            dev_alloc((void **)&o->d_lpairs, (2 * (size_t)m - 1) * sizeof(LPair));
            launch_light_pairs(s, o->d_light_bvh, 2 * m - 1, o->d_lpairs);
            o->h_light_bvh.resize(2 * (size_t)m - 1);
            HIP_CHECK(hipMemcpyAsync(o->h_light_bvh.data(), o->d_light_bvh, o->h_light_bvh.size() * sizeof(vrgi_bvh_node),
                                     hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            // light vertices in leaf order (scene.rs:509-524)
            lv.reserve(3 * (size_t)m);
            for (size_t i = 0; i < o->h_light_bvh.size(); ++i) {
                const vrgi_bvh_node &nd = o->h_light_bvh[i];
                if (nd.left != VRGI_INVALID) continue;
                for (int k = 0; k < 3; ++k) {
                    vrgi_light_vertex3d x;
                    std::memcpy(x.position, v[(size_t)nd.right_or_prim * 3 + k].position, 12);
                    x.bvh_node_idx = (uint32_t)i;
                    lv.push_back(x);
                }
            }
            dev_alloc((void **)&o->d_light_verts, lv.size() * sizeof(vrgi_light_vertex3d));
            HIP_CHECK(hipMemcpyAsync(o->d_light_verts, lv.data(), lv.size() * sizeof(vrgi_light_vertex3d),
                                     hipMemcpyHostToDevice, s));
            // returned (Aabb, power) of build_bl_bvh (build.rs:359-371)
            const vrgi_bvh_node &r = o->h_light_bvh[0];
            o->light_empty = false;
            if (r.left == VRGI_INVALID) {
                for (int k = 0; k < 3; ++k) {
                    o->light_aabb[k] = std::min(std::min(tri9[k], tri9[3 + k]), tri9[6 + k]);
                    o->light_aabb[3 + k] = std::max(std::max(tri9[k], tri9[3 + k]), tri9[6 + k]);
                }
                o->power = pw[0];
            } else {
                for (int k = 0; k < 3; ++k) { o->light_aabb[k] = r.min[k]; o->light_aabb[3 + k] = r.max[k]; }
                o->power = r.power;
            }
        }
        HIP_CHECK(hipStreamSynchronize(s));
    } catch (...) {
        free_object(o);
        throw;
    }
    *out = o;
    return VRGI_OK;
    API_END
}

int vrgi_object_retain(vrgi_object *o) {
    if (!o) return fail(VRGI_E_INVALID, "null object");
    o->refs.fetch_add(1);
    return VRGI_OK;
}
int vrgi_object_release(vrgi_object *o) {
    API_BEGIN
    if (!o) return fail(VRGI_E_INVALID, "null object");
    release(o);
    return VRGI_OK;
    API_END
}
int vrgi_object_light_bvh(vrgi_object *o, vrgi_bvh_node *dst, size_t cap, size_t *n_nodes) {
    API_BEGIN
    if (!o || !n_nodes) return fail(VRGI_E_INVALID, "null argument");
    if (o->h_light_bvh.empty() && o->n_light && o->d_light_bvh) {  // an object of a voxel commit: read on demand
        HIP_CHECK(hipSetDevice(o->life->device));
        o->h_light_bvh.resize(2 * (size_t)o->n_light - 1);
        HIP_CHECK(hipMemcpy(o->h_light_bvh.data(), o->d_light_bvh, o->h_light_bvh.size() * sizeof(vrgi_bvh_node),
                            hipMemcpyDeviceToHost));
    }
    *n_nodes = o->h_light_bvh.size();
    if (dst) {
        if (cap < o->h_light_bvh.size()) return fail(VRGI_E_LIMIT, "destination too small");
        std::memcpy(dst, o->h_light_bvh.data(), o->h_light_bvh.size() * sizeof(vrgi_bvh_node));
    }
    return VRGI_OK;
    API_END
}

int vrgi_scene_add(vrgi_ctx *c, uint32_t key, vrgi_object *o, const float xf[12]) {
    API_BEGIN
    if (!c || !valid_xf(xf)) return fail(VRGI_E_INVALID, "null context or non-finite transform");
    if (o && o->life != c->life) return fail(VRGI_E_INVALID, "the object was uploaded through another context");
    auto it = c->objects.find(key);
    if (it != c->objects.end()) {
        if (it->second.first) c->old.push_back(it->second.first);
        c->objects.erase(it);
    }
    if (o) o->refs.fetch_add(1);
    c->objects[key] = {o, std::vector<float>(xf, xf + 12)};
    c->inst_dirty = true;
    mark_geo_dirty(c);
    c->view_dirty = true;
    return VRGI_OK;
    API_END
}

int vrgi_scene_update(vrgi_ctx *c, uint32_t key, const float xf[12]) {
    API_BEGIN
    if (!c || !valid_xf(xf)) return fail(VRGI_E_INVALID, "null context or non-finite transform");
    auto it = c->objects.find(key);
    if (it == c->objects.end()) return fail(VRGI_E_INVALID, "object with key does not exist (scene.rs:215)");
    if (!it->second.first) return VRGI_OK;  // Some(None): no-op
    it->second.second.assign(xf, xf + 12);
    c->inst_dirty = true;
    mark_geo_dirty(c);
    c->view_dirty = true;
    return VRGI_OK;
    API_END
}

int vrgi_scene_remove(vrgi_ctx *c, uint32_t key) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    auto it = c->objects.find(key);
    if (it == c->objects.end()) return VRGI_OK;
    if (it->second.first) {
        c->old.push_back(it->second.first);
        c->inst_dirty = true;
    mark_geo_dirty(c);
    }
    c->view_dirty = true;
    c->objects.erase(it);
    return VRGI_OK;
    API_END
}

int vrgi_dispose_old(vrgi_ctx *c) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    set_dev(c);
    // the frames in flight and the next frame's build (on the build stream) end first
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipStreamSynchronize(c->up_stream));
    for (vrgi_object *o : c->old) release(o);
    c->old.clear();
    return VRGI_OK;
    API_END
}

// IEC 61966-2-1 decode of an 8-bit sRGB channel, in binary64 rounded once
static float srgb_decode(uint8_t v) {
    const double x = (double)v / 255.0;
    return (float)(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
}

int vrgi_set_textures(vrgi_ctx *c, uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (c->in_batch) return fail(VRGI_E_STATE, "the scene is in a batch (no texture atlas there)");
    if (rgba8 && c->cfg.restir_mode == 1)
        return fail(VRGI_E_STATE, "ReSTIR DI reads the constant materials: no texture atlas with restir_mode 1");
    if (rgba8 && (n_slots < c->mats.size() || w == 0 || h == 0 || w > 4096 || h > 4096))
        return fail(VRGI_E_INVALID, "the atlas needs a texture for every material slot, 1 <= width, height <= 4096");
    if (rgba8 && (uint64_t)n_slots * 3 * w * h > (1ull << 28)) return fail(VRGI_E_LIMIT, "atlas above 2^28 texels");
    set_dev(c);
    // frames in flight may read the atlas or the uv of the current geometry
    for (hipStream_t st : {c->stream, c->up_stream, c->side}) HIP_CHECK(hipStreamSynchronize(st));
    for (hipStream_t st : c->nee_s) if (st) HIP_CHECK(hipStreamSynchronize(st));
    if (!rgba8) {
        c->tex_on = false;
    } else {
        float lut[256];
        for (int v = 0; v < 256; ++v) lut[v] = srgb_decode((uint8_t)v);
        const size_t texels = (size_t)n_slots * 3 * w * h;
        std::vector<float4> t(texels);
        for (size_t q = 0; q < texels; ++q)
            t[q] = make_float4(lut[rgba8[4 * q]], lut[rgba8[4 * q + 1]], lut[rgba8[4 * q + 2]],
                               (float)((double)rgba8[4 * q + 3] / 255.0));
        c->d_tex.reserve(texels);
        HIP_CHECK(hipMemcpy(c->d_tex.p, t.data(), texels * sizeof(float4), hipMemcpyHostToDevice));
        c->tex_w = w;
        c->tex_h = h;
        c->tex_on = true;
    }
    mark_geo_dirty(c);  // the next build writes (or stops writing) the triangles' uv
    c->view_dirty = true;
    return VRGI_OK;
    API_END
}

int vrgi_texture_luminances(uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8, float *out) {
    API_BEGIN
    if ((!rgba8 || !out) && n_slots) return fail(VRGI_E_INVALID, "null argument");
    for (uint32_t t = 0; t < n_slots; ++t) {  // utils.rs:223-235
        const uint8_t *e = rgba8 + ((size_t)t * 3 + 1) * w * h * 4;
        float l = 0.0f;
        for (size_t q = 0; q < (size_t)w * h; ++q) l += ((float)e[4 * q] + (float)e[4 * q + 1]) + (float)e[4 * q + 2];
        out[t] = l / (float)(w * h);
    }
    return VRGI_OK;
    API_END
}

int vrgi_resize(vrgi_ctx *c, uint32_t w, uint32_t h) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (c->in_batch) return fail(VRGI_E_STATE, "the scene is in a batch (its extent is fixed there)");
    if (w == 0 || h == 0) {
        // a minimised window: render() draws nothing until the next non-zero extent
        // (interactive_rendering.rs:978-981); the buffers of the last extent are kept
        c->minimized = true;
        return VRGI_OK;
    }
    if ((uint64_t)w * h > (1ull << 26)) return fail(VRGI_E_LIMIT, "extent above 2^26 pixels");
    set_dev(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->minimized = false;
    c->W = w;
    c->H = h;
    c->P = w * h;
    c->cfg.width = w;
    c->cfg.height = h;
    alloc_frame(c);
    if (c->accum_on) {
        // the FP64 sums restart at the new extent
        c->accum.reserve((size_t)c->P * 3);
        HIP_CHECK(hipMemsetAsync(c->accum.p, 0, (size_t)c->P * 24, c->stream));
        c->accum_P = c->P;
        c->accum_frames = 0;
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_render(vrgi_ctx *c, const vrgi_camera *cam, const vrgi_prefs *prefs, vrgi_frame_out *out) {
    API_BEGIN
    if (!c || !cam || !prefs) return fail(VRGI_E_INVALID, "null argument");
    if (prefs->nee_type > 2) return fail(VRGI_E_INVALID, "nee_type must be 0, 1 or 2");
    if (prefs->restir_spatial_iterations > 20) return fail(VRGI_E_INVALID, "restir_spatial_iterations above 20");
    if (prefs->debug_view > 4) return fail(VRGI_E_INVALID, "debug_view must be in [0, 4]");
    if (c->tex_on && c->cfg.restir_mode == 1)
        return fail(VRGI_E_STATE, "ReSTIR DI reads the constant materials: no texture atlas with restir_mode 1");
    if (c->minimized) {  // zero extent: no frame (interactive_rendering.rs:978-981)
        if (out) std::memset(out, 0, sizeof(*out));
        return VRGI_OK;
    }
    set_dev(c);
    hipStream_t s = c->stream;
    const uint32_t P = c->P, B = c->B;
    const bool stages = (c->profile_flags & VRGI_PROFILE_STAGES) != 0;
    const bool touch = (c->profile_flags & VRGI_PROFILE_TOUCH) != 0;
    // events: 0 start, 1 after the builds, 2+b after trace b, 2+B+b after NEE b (2+2B-1:
    // NEE joined), 2+2B end, 3+2B+b before NEE b, 3+3B / 4+3B around the build (on the
    // stream it runs on)
    const uint32_t par = (uint32_t)(c->frame_count & 1);
    const size_t n_ev = 5 + 3 * (size_t)B;
    if (stages) {
        harvest_frame(c, par);  // a frame two renders back, if it is still unread
        while (c->events[par].size() < n_ev) {
            hipEvent_t e;
            HIP_CHECK(hipEventCreate(&e));
            c->events[par].push_back(e);
        }
    }
    auto rec_on = [&](size_t k, hipStream_t st) {
        if (stages) HIP_CHECK(hipEventRecord(c->events[par][k], st));
    };
    auto rec = [&](size_t k) { rec_on(k, s); };
    // NEE pdf of bounce b needs only bounce b's outgoing rays (slot b, queue b + 1) and
    // writes only its slot of the pdf, while the trace of bounce b + 1 reads slot b and
    // writes slot b + 1: the two run side by side
    // Lazy NEE pdf: the pdfs run after all bounces, from the last bounce back, each after
    // the resolve step that gives the radiance its rays return, and only for rays where the
    // value can reach the outputs (launch_nee_pdf_lazy). Not with the learned cache: its
    // training paths return radiance from most of their vertices, so few pdfs could be
    // left out, and their passes, which otherwise run beside the next bounce, would add
    // to the end of the frame (measured: cache frame 2.8 -> 3.7 ms).
    const bool lazy_nee = prefs->nee_type != 0 && !c->cache.on && !(c->cfg.flags & VRGI_FLAG_FULL_NEE_PDF);
    const bool nee_overlap = prefs->nee_type != 0 && !lazy_nee && !(c->cfg.flags & VRGI_FLAG_SERIAL_NEE);
    while (c->ev_bounce.size() < B) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        c->ev_bounce.push_back(e);
    }
    rec(0);
    bool built = false;
    if (c->vx.pending) voxels_commit(c);
    if (c->inst_dirty) {
        rebuild_instances(c);
        c->inst_dirty = false;
        mark_geo_dirty(c);
    }
    if (c->geo_dirty || (c->cfg.flags & VRGI_FLAG_REBUILD_EVERY_FRAME)) {
        rebuild_geometry(c, stages ? c->events[par][3 + 3 * (size_t)B] : nullptr,
                         stages ? c->events[par][4 + 3 * (size_t)B] : nullptr, !c->geo_dirty);
        c->geo_dirty = false;
        built = true;
    }
    rec(1);
    const uint32_t frame = (uint32_t)c->frame_count;
    HIP_CHECK(hipMemsetAsync(c->f_qcount.p, 0, (B + 2) * 4, s));
    const bool cache_on = c->cache.on;
    // learned cache: training records of frame f go to half f % 2 (count word 0 or 3);
    // the side stream trains on the other half (frame f - 1) while this frame traces
    float *trec_cur = c->cache.trec.p + (size_t)par * c->cache.tcap * 16;
    float *trec_prev = c->cache.trec.p + (size_t)(par ^ 1) * c->cache.tcap * 16;
    uint32_t *tcnt_cur = c->cache.counts.p + (par ? 3 : 0), *tcnt_prev = c->cache.counts.p + (par ? 0 : 3);
    const bool cache_train = cache_on && !(c->cache.cfg.flags & VRGI_CACHE_FREEZE) && c->cache.cfg.steps_per_frame;
    if (cache_on) {
        cache_box_from_scene(c);
        HIP_CHECK(hipMemsetAsync(c->cache.counts.p + 1, 0, 4, s));
        HIP_CHECK(hipMemsetAsync(tcnt_cur, 0, 4, s));
    }
    // the frame's training steps on the previous frame's records, on the build stream
    // (idle once this frame's BVHs are built), after a snapshot of the network for this
    // frame's inference; they run beside the frame's rays and inference. They start from
    // the frame's start (ev_fstart) but are queued on the host after the first bounce's
    // launch, so that their ~30 launches do not delay the frame's first rays.
    if (cache_train) HIP_CHECK(hipEventRecord(c->ev_fstart, s));
    // the snapshot is queued with the first bounce (the inference waits for it), the rest
    // of the training after the second bounce's launch, so that the host's ~30 calls do
    // not hold back the launches of the frame's ray chain
    auto enqueue_snapshot = [&]() {
        auto &C = c->cache;
        hipStream_t ts = c->up_stream;
        HIP_CHECK(hipStreamWaitEvent(ts, c->ev_fstart, 0));
        HIP_CHECK(hipMemcpyAsync(C.table_snap.p, C.table.p, C.n_table * sizeof(float), hipMemcpyDeviceToDevice, ts));
        HIP_CHECK(hipMemcpyAsync(C.wt_snap.p, C.wt.p, NRC_N_WEIGHTS * sizeof(uint16_t), hipMemcpyDeviceToDevice, ts));
        HIP_CHECK(hipEventRecord(c->ev_snap, ts));
    };
    bool training_queued = false;
    auto enqueue_training = [&]() {
        auto &C = c->cache;
        hipStream_t ts = c->up_stream;
        training_queued = true;
        HIP_CHECK(hipMemsetAsync(C.loss.p, 0, 4, ts));
        C.trec_sorted.reserve((size_t)C.tcap * 16);
        C.sort_hist.reserve(NRC_SORT_BUCKETS + NRC_SORT_BUCKETS / 1024);
        C.sort_kr.reserve((size_t)2 * C.tcap);
        RecBox box;
        for (int k = 0; k < 3; ++k) { box.lo[k] = C.lo[k]; box.hi[k] = C.hi[k]; }
        launch_rec_sort(ts, trec_prev, tcnt_prev, C.tcap, box, C.sort_hist.p, C.sort_kr.p, C.trec_sorted.p);
        for (uint32_t k = 0; k < C.cfg.steps_per_frame; ++k)
            cache_step(c, ts, C.trec_sorted.p, tcnt_prev, C.tcap, k, C.cfg.steps_per_frame, C.cfg.batch, true);
        HIP_CHECK(hipEventRecord(c->ev_train, ts));
    };
    if (c->cfg.flags & VRGI_FLAG_KEEP_BOUNCES) {
        size_t n = (size_t)B * P;
        HIP_CHECK(hipMemsetAsync(c->f_origin.p, 0, n * 12, s));
        HIP_CHECK(hipMemsetAsync(c->f_dir.p, 0, n * 12, s));
        HIP_CHECK(hipMemsetAsync(c->f_normal.p, 0, n * 12, s));
        HIP_CHECK(hipMemsetAsync(c->f_emis.p, 0, n * 12, s));
        HIP_CHECK(hipMemsetAsync(c->f_refl.p, 0, n * 12, s));
        HIP_CHECK(hipMemsetAsync(c->f_mis.p, 0, n * 4, s));
        HIP_CHECK(hipMemsetAsync(c->f_pdf.p, 0, n * 4, s));
        HIP_CHECK(hipMemsetAsync(c->f_nee.p, 0, n * 4, s));
        HIP_CHECK(hipMemsetAsync(c->f_hit_t.p, 0, n * 4, s));
        HIP_CHECK(hipMemsetAsync(c->f_hit.p, 0xFF, n * 4, s));
    }
    SceneView sv = scene_view(c);
    BounceArgs ba{};
    ba.B = B; ba.W = c->W; ba.H = c->H; ba.P = P;
    ba.nee_type = prefs->nee_type;
    ba.rr_start = c->cfg.rr_start_bounce;
    ba.rr_min = c->cfg.rr_min_survival;
    ba.cam = *cam;
    ba.qcount = c->f_qcount.p;
    ba.origin = c->f_origin.p; ba.dir = c->f_dir.p; ba.normal = c->f_normal.p;
    ba.emis = c->f_emis.p; ba.refl = c->f_refl.p; ba.mis = c->f_mis.p; ba.pdf = c->f_pdf.p;
    ba.nee = c->f_nee.p; ba.hit_t = c->f_hit_t.p; ba.hit = c->f_hit.p; ba.path_len = c->f_plen.p;
    ba.counters = (unsigned long long *)c->d_counters.p;
    const uint32_t train_seed = frame_seed_host(c->cache.cfg.seed, frame, 5);
    if (cache_on) {
        ba.cache_on = 1;
        ba.train_seed = train_seed;
        ba.train_frac = c->cache.cfg.train_fraction;
        ba.cache_vertex = c->cache.cfg.query_vertex;
        ba.d0 = c->cache.d0.p;
        ba.fallback = c->cache.cfg.fallback_bound > 0.0f ? c->cache.t_fb.p : nullptr;
        ba.tile = c->cache.cfg.fallback_tile;
        ba.tiles_x = c->cache.tiles_x;
        ba.albedo = c->cache.albedo.p;
        ba.qrec = c->cache.qrec.p;
        ba.qrec_count = c->cache.counts.p + 1;
    }
    size_t touch_words = ((size_t)c->n_tris + 31) / 32;
    if (touch) {
        c->touch_bits.reserve(2 * touch_words);
        c->touch_count.reserve(2 * (size_t)B);
        HIP_CHECK(hipMemsetAsync(c->touch_count.p, 0, 16 * (size_t)B, s));
    }
    // queue of the rays of each bounce: compaction order, or Morton order when enabled
    std::vector<const uint32_t *> qin(B + 1, nullptr);
    // a cache frame querying at vertex 0: bounce 0 in two parts (BounceArgs::b0_part)
    const bool split_b0 = cache_on && c->cache.cfg.query_vertex == 0 && VRGI_CACHE_SPLIT_B0;
    const bool sort_rays = c->cfg.ray_sort_bits != 0 && c->n_tris >= 2;
    for (uint32_t b = 0; b < B; ++b) {
        ba.bounce = b;
        ba.inv_seed = frame * B + b;  // interactive_rendering.rs:1189
        ba.queue_in = qin[b];
        ba.queue_out = c->f_queue.p + (size_t)(b + 1) * P;
        if (touch) {
            HIP_CHECK(hipMemsetAsync(c->touch_bits.p, 0, 8 * touch_words, s));
            ba.touch_nodes = c->touch_bits.p;
            ba.touch_tris = c->touch_bits.p + touch_words;
        }
        // past the cache's query vertex only training paths (and fallback tiles) go on
        const uint32_t small_q = cache_on && b > c->cache.cfg.query_vertex ? VRGI_SMALL_QUEUE_GRID : 0u;
        if (b == 0 && split_b0) {
            // the full-path pixels' first bounce starts the frame's ray chain on this stream;
            // the other pixels' (which end at vertex 0 with a cache query) run beside it on
            // the side stream, followed there by the inference
            launch_full_list(s, ba, c->f_queue.p, c->f_qcount.p);
            HIP_CHECK(hipEventRecord(c->ev_query, s));
            HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_query, 0));
            BounceArgs bb = ba;
            bb.b0_part = 2;
            launch_trace_shade(c->side, sv, bb, P, 0);
            ba.b0_part = 1;
            ba.queue_in = c->f_queue.p;
            launch_trace_shade(s, sv, ba, P, VRGI_SMALL_QUEUE_GRID);
            ba.b0_part = 0;
            ba.queue_in = nullptr;
        } else {
            launch_trace_shade(s, sv, ba, P, small_q);
        }
        qin[b + 1] = ba.queue_out;
        if (cache_train && b == 0) enqueue_snapshot();  // before any wait on ev_snap
        if (cache_train && b == 1) enqueue_training();
        if (cache_on && b == c->cache.cfg.query_vertex) {
            // side stream: cache radiance of this bounce's queries, by the network as the
            // frame started (the snapshot while the training steps run)
            auto &C = c->cache;
            if (!split_b0) {  // split: the queries were written on the side stream
                HIP_CHECK(hipEventRecord(c->ev_query, s));
                HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_query, 0));
            }
            NrcNet net = cache_net(c);
            if (cache_train) {
                HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_snap, 0));
                net.table = C.table_snap.p;
                net.wt = C.wt_snap.p;
            }
            launch_nrc_infer(c->side, net, C.qrec.p, C.counts.p + 1, P, nullptr, C.crad.p);
            HIP_CHECK(hipEventRecord(c->ev_infer, c->side));
        }
        if (sort_rays && b + 1 < B) {
            SortArgs sa{};
            sa.bounce = b; sa.P = P; sa.bits = c->cfg.ray_sort_bits;
            sa.root = c->G().d_gwide.p;
            sa.origin = c->f_origin.p; sa.dir = c->f_dir.p;
            sa.count = c->f_qcount.p + b + 1;
            sa.queue_in = ba.queue_out;
            sa.queue_out = c->f_squeue.p + (size_t)(b + 1) * P;
            sa.keys = c->f_skeys.p; sa.hist = c->f_shist.p;
            launch_sort_queue(s, sa, P);
            qin[b + 1] = sa.queue_out;
        }
        if (touch) {
            launch_popcount(s, ba.touch_nodes, touch_words, c->touch_count.p + 2 * b);
            launch_popcount(s, ba.touch_tris, touch_words, c->touch_count.p + 2 * b + 1);
        }
        rec(2 + b);
        if (nee_overlap && b + 1 < B) {  // interactive_rendering.rs:1206
            hipStream_t ns = c->nee_s[b % VRGI_NEE_STREAMS];
            HIP_CHECK(hipEventRecord(c->ev_bounce[b], s));
            HIP_CHECK(hipStreamWaitEvent(ns, c->ev_bounce[b], 0));
            BounceArgs na = ba;
            na.queue_in = qin[b + 1];
            na.touch_nodes = na.touch_tris = nullptr;
            // short queues (the training paths past the cache's query vertex): the pdf in
            // passes, whose walks then run at full occupancy instead of in the tail of a
            // few waves that also hold the traversal
            const bool short_q = cache_on && b >= c->cache.cfg.query_vertex;
            if (short_q || VRGI_NEE_PASSES_ALL) nee_passes(c, ns, b % VRGI_NEE_STREAMS, na);
            rec_on(3 + 2 * (size_t)B + b, ns);
            launch_nee_pdf(ns, sv, na, P, short_q ? VRGI_SMALL_QUEUE_GRID : 0u);
            rec_on(2 + B + b, ns);
        }
    }
    if (cache_train && !training_queued) enqueue_training();
    FrameArgs fa{};
    fa.W = c->W; fa.H = c->H; fa.P = P; fa.B = B; fa.frame = frame;
    fa.slot = (uint32_t)(c->frame_count % 3);
    fa.tseed = frame_seed_host(c->cfg.seed, frame, 2);
    fa.sseed = frame_seed_host(c->cfg.seed, frame, 3);
    fa.spatial_iters = prefs->restir_spatial_iterations;
    fa.radius = c->cfg.restir_spatial_radius;
    fa.debug_view = prefs->debug_view;
    if (prefs->debug_view == 4) {
        const bool restart = c->view_n == 0 || c->view_prev != 4 || c->view_dirty ||
                             std::memcmp(&c->view_cam, cam, sizeof(vrgi_camera)) != 0;
        c->view_n = restart ? 1u : c->view_n + 1u;
        c->view_cam = *cam;
        c->view_dirty = false;
        fa.view_acc = c->f_view_acc.p;
        fa.view_n = c->view_n;
    }
    c->view_prev = prefs->debug_view;
    fa.origin = c->f_origin.p; fa.normal = c->f_normal.p; fa.emis = c->f_emis.p; fa.refl = c->f_refl.p;
    fa.mis = c->f_mis.p; fa.pdf = c->f_pdf.p; fa.nee = prefs->nee_type ? c->f_nee.p : nullptr;
    fa.path_len = c->f_plen.p;
    fa.radiance = c->f_radiance.p;
    fa.temporal = c->f_temporal.p + (size_t)fa.slot * P * 20;
    fa.spatial = c->f_spatial.p;
    fa.restir = c->f_restir.p;
    fa.debug = c->f_debug.p;
    fa.bgra8 = c->f_bgra.p;
    if (cache_on) {
        fa.cache_on = 1;
        fa.train_seed = train_seed;
        fa.train_frac = c->cache.cfg.train_fraction;
        fa.cache_vertex = c->cache.cfg.query_vertex;
        fa.fallback = c->cache.cfg.fallback_bound > 0.0f ? c->cache.t_fb.p : nullptr;
        fa.tile = c->cache.cfg.fallback_tile;
        fa.tiles_x = c->cache.tiles_x;
        fa.dir = c->f_dir.p;
        fa.d0 = c->cache.d0.p;
        fa.albedo = c->cache.albedo.p;
        fa.crad = c->cache.crad.p;
        fa.hit = c->f_hit.p;
        fa.trec = trec_cur;
        fa.trec_count = tcnt_cur;
        fa.tcap = c->cache.tcap;
    }
    ba.touch_nodes = ba.touch_tris = nullptr;
    if (nee_overlap) {
        for (uint32_t k = 0; k < VRGI_NEE_STREAMS && k + 1 < B; ++k) {
            HIP_CHECK(hipEventRecord(c->ev_nee_done[k], c->nee_s[k]));
            HIP_CHECK(hipStreamWaitEvent(s, c->ev_nee_done[k], 0));
        }
    } else if (lazy_nee) {
        HIP_CHECK(hipMemsetAsync(c->f_ncount.p, 0, 32 * 4, s));
        for (int b = (int)B - 2; b >= 0; --b) {  // interactive_rendering.rs:1206
            launch_resolve_step(s, fa, (uint32_t)b + 1u, c->f_lrun.p);
            rec(3 + 2 * (size_t)B + (size_t)b);
            BounceArgs na = ba;
            na.bounce = (uint32_t)b;
            na.queue_in = qin[b + 1];
            na.lrun = c->f_lrun.p;
            launch_nee_pdf_lazy(s, sv, na, P, c->f_nlist.p, c->f_ncount.p + b, c->f_mlist.p, c->f_ncount.p + 16 + b);
            rec(2 + B + (size_t)b);
        }
    } else {
        for (uint32_t b = 0; b + 1 < B; ++b) {  // interactive_rendering.rs:1206
            rec(3 + 2 * (size_t)B + b);
            if (prefs->nee_type != 0) {
                ba.bounce = b;
                ba.queue_in = qin[b + 1];
                const bool short_q = cache_on && b >= c->cache.cfg.query_vertex;
                BounceArgs na = ba;
                if (short_q) nee_passes(c, s, 0, na);
                launch_nee_pdf(s, sv, na, P, short_q ? VRGI_SMALL_QUEUE_GRID : 0u);
            }
            rec(2 + B + b);
        }
    }
    if (split_b0) {
        // pixels that end at vertex 0 with a cache query: after the inference on the side
        // stream, beside the training paths' bounces; the full-path pixels here
        FrameArgs f1 = fa;
        f1.part = 1;
        launch_accum_temporal(c->side, f1);
        HIP_CHECK(hipEventRecord(c->ev_infer, c->side));
        rec(2 + B + (B - 1));
        FrameArgs f2 = fa;
        f2.part = 2;
        f2.full_list = c->f_queue.p;
        f2.full_count = c->f_qcount.p;
        launch_accum_temporal(s, f2);
        HIP_CHECK(hipStreamWaitEvent(s, c->ev_infer, 0));  // every pixel's temporal reservoir
    } else {
        if (cache_on) HIP_CHECK(hipStreamWaitEvent(s, c->ev_infer, 0));  // cache radiance is ready
        rec(2 + B + (B - 1));
        launch_accum_temporal(s, fa);
    }
    if (cache_on && c->cache.cfg.fallback_bound > 0.0f) {
        // held-out error of this frame's records at the query vertex (the side stream has
        // finished its training steps, none of them on these records), per screen tile
        auto &C = c->cache;
        HIP_CHECK(hipEventRecord(c->ev_accum, s));
        HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_accum, 0));
        if (cache_train) HIP_CHECK(hipStreamWaitEvent(c->side, c->ev_train, 0));
        C.hpred.reserve((size_t)C.tcap * 3);
        launch_nrc_infer(c->side, cache_net(c), trec_cur, tcnt_cur, C.tcap, C.hpred.p, nullptr);
        launch_tile_stats(c->side, trec_cur, tcnt_cur, C.tcap, C.hpred.p, C.cfg.query_vertex, c->W,
                          C.cfg.fallback_tile, C.tiles_x, C.t_acc.p);
        launch_tile_update(c->side, C.t_acc.p, C.t_ema.p, C.t_fb.p, C.tiles_x * C.tiles_y, C.cfg.fallback_ema,
                           C.cfg.fallback_bound);
        HIP_CHECK(hipEventRecord(c->ev_held, c->side));
    }
    if (c->cfg.restir_mode == 1) {
        DiArgs da{};
        da.W = c->W; da.H = c->H; da.P = P; da.frame = frame;
        da.candidates = c->cfg.di_candidates;
        da.spatial_iters = prefs->restir_spatial_iterations;
        da.radius = c->cfg.restir_spatial_radius;
        da.dseed = frame_seed_host(c->cfg.seed, frame, 4);
        da.origin = c->f_origin.p; da.normal = c->f_normal.p; da.emis = c->f_emis.p; da.hit = c->f_hit.p;
        da.T = c->di_T.p;
        da.hist_prev = c->di_hist.p + (size_t)((frame + 1) % 2) * P * 8;
        da.hist_cur = c->di_hist.p + (size_t)(frame % 2) * P * 8;
        da.restir = c->f_restir.p;
        da.di_light = c->di_light.p;
        da.di_weights = c->di_w.p;
        launch_di(s, sv, da);
    } else {
        launch_spatial_finalize(s, fa);
    }
    launch_postprocess(s, fa);
    if (c->accum_on) {
        c->accum.reserve((size_t)P * 3);
        launch_accumulate(s, c->f_radiance.p, c->accum.p, (size_t)P * 3);
        c->accum_frames += 1;
    }
    if (cache_train) HIP_CHECK(hipStreamWaitEvent(s, c->ev_train, 0));  // join the training steps
    if (cache_on && c->cache.cfg.fallback_bound > 0.0f) HIP_CHECK(hipStreamWaitEvent(s, c->ev_held, 0));
    launch_frame_stats(s, c->f_qcount.p, B, P, (unsigned long long *)c->d_counters.p);
    // the set this frame read may be rebuilt once this point has passed
    HIP_CHECK(hipEventRecord(c->G().ev_free, s));
    c->G().freed_once = true;
    if (stages) {
        if (!c->h_qcount) HIP_CHECK(hipHostMalloc((void **)&c->h_qcount, 2 * (2 * (size_t)B + 2) * 4, hipHostMallocDefault));
        uint32_t *hq = c->h_qcount + (size_t)par * (2 * B + 2);
        HIP_CHECK(hipMemcpyAsync(hq, c->f_qcount.p, ((size_t)B + 2) * 4, hipMemcpyDeviceToHost, s));
        if (lazy_nee) HIP_CHECK(hipMemcpyAsync(hq + B + 2, c->f_ncount.p, (size_t)B * 4, hipMemcpyDeviceToHost, s));
        else std::memset(hq + B + 2, 0, (size_t)B * 4);
    }
    rec(2 + 2 * (size_t)B);
    if (stages) {
        c->pend[par] = true;
        c->pend_built[par] = built;
        c->pend_P[par] = P;
        harvest_frame(c, par ^ 1u);  // the previous frame: its work ran beside this one's build
    }
    if (touch) {
        HIP_CHECK(hipStreamSynchronize(s));
        vrgi_profile_out &p = c->prof;
        std::vector<unsigned long long> tc(2 * (size_t)B);
        HIP_CHECK(hipMemcpy(tc.data(), c->touch_count.p, 16 * (size_t)B, hipMemcpyDeviceToHost));
        for (uint32_t b = 0; b < B && b < 16; ++b)
            p.touched_bytes[b] = tc[2 * b] * VRGI_GEOM_NODE_BYTES + tc[2 * b + 1] * 48ull;
        c->profile_flags &= ~VRGI_PROFILE_TOUCH;
    }
    c->have_frame = true;
    if (out) {
        out->radiance = c->f_radiance.p;
        out->restir = c->f_restir.p;
        out->bgra8 = c->f_bgra.p;
        out->frame_index = c->frame_count;
    }
    ++c->frame_count;
    return VRGI_OK;
    API_END
}

int vrgi_sync(vrgi_ctx *c) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    set_dev(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_stats(vrgi_ctx *c, vrgi_frame_stats *st, int reset) {
    API_BEGIN
    if (!c || !st) return fail(VRGI_E_INVALID, "null argument");
    set_dev(c);
    Counters h{};
    HIP_CHECK(hipMemcpyAsync(&h, c->d_counters.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    st->frames = h.frames;
    st->rays_closest = h.rays_closest;
    st->rays_light = h.rays_light;
    st->bvh_nodes = c->n_tris ? 2ull * c->n_tris - 1 : 0;
    st->bvh_depth = c->G().geo_depth;
    st->pad = 0;
    if (reset) {
        HIP_CHECK(hipMemsetAsync(c->d_counters.p, 0, sizeof(Counters), c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    return VRGI_OK;
    API_END
}

int vrgi_profile(vrgi_ctx *c, uint32_t flags) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (flags & ~(VRGI_PROFILE_STAGES | VRGI_PROFILE_TOUCH)) return fail(VRGI_E_INVALID, "unknown profile flag");
    set_dev(c);
    harvest_frame(c, 0);
    harvest_frame(c, 1);
    c->profile_flags = flags;
    return VRGI_OK;
    API_END
}

int vrgi_profile_read(vrgi_ctx *c, vrgi_profile_out *out, int reset) {
    API_BEGIN
    if (!c || !out) return fail(VRGI_E_INVALID, "null argument");
    set_dev(c);
    harvest_frame(c, 0);
    harvest_frame(c, 1);
    *out = c->prof;
    if (reset) {
        vrgi_profile_out keep = c->prof;
        c->prof = vrgi_profile_out{};
        // structural figures of the last build and touch frame survive a reset
        c->prof.build_prims = keep.build_prims;
        c->prof.build_leaf_depth_sum = keep.build_leaf_depth_sum;
        c->prof.build_nodes = keep.build_nodes;
        std::memcpy(c->prof.touched_bytes, keep.touched_bytes, sizeof(keep.touched_bytes));
    }
    return VRGI_OK;
    API_END
}

int vrgi_buffer_size(vrgi_ctx *c, int which, size_t *bytes) {
    API_BEGIN
    if (!c || !bytes) return fail(VRGI_E_INVALID, "null argument");
    size_t P = c->P, B = c->B, nt = c->n_tris, nl = c->n_ltri;
    size_t r = 0;
    switch (which) {
        case VRGI_BUF_GEOM_BVH: r = nt ? (2 * nt - 1) * 40 : 0; break;
        case VRGI_BUF_GEOM_PRIM_ORDER: r = nt * 4; break;
        case VRGI_BUF_LIGHT_TL_BVH: r = (size_t)c->n_tl * 40; break;
        case VRGI_BUF_LIGHT_GEOM_BVH: r = nl ? (2 * nl - 1) * 40 : 0; break;
        case VRGI_BUF_PATH_LEN: r = P * 4; break;
        case VRGI_BUF_BOUNCE_ORIGIN: case VRGI_BUF_BOUNCE_DIRECTION: case VRGI_BUF_BOUNCE_NORMAL:
        case VRGI_BUF_BOUNCE_EMISSIVITY: case VRGI_BUF_BOUNCE_REFLECTIVITY: r = B * P * 12; break;
        case VRGI_BUF_BOUNCE_MIS_WEIGHT: case VRGI_BUF_BOUNCE_BSDF_PDF: case VRGI_BUF_BOUNCE_NEE_PDF:
        case VRGI_BUF_HIT_PRIM: case VRGI_BUF_HIT_T: r = B * P * 4; break;
        case VRGI_BUF_RADIANCE: case VRGI_BUF_RESTIR: r = P * 12; break;
        case VRGI_BUF_TEMPORAL_RESERVOIR: case VRGI_BUF_SPATIAL_RESERVOIR: r = P * 80; break;
        case VRGI_BUF_BGRA8: r = P * 4; break;
        case VRGI_BUF_DI_LIGHT: r = c->cfg.restir_mode == 1 ? P * 4 : 0; break;
        case VRGI_BUF_DI_WEIGHTS: r = c->cfg.restir_mode == 1 ? P * 16 : 0; break;
        case VRGI_BUF_BOUNCE_ALBEDO: r = c->cache.on ? B * P * 12 : 0; break;
        case VRGI_BUF_CACHE_RADIANCE: r = c->cache.on ? P * 12 : 0; break;
        case VRGI_BUF_CACHE_FALLBACK: r = c->cache.on ? (size_t)c->cache.tiles_x * c->cache.tiles_y : 0; break;
        case VRGI_BUF_CACHE_QUERIES:
        case VRGI_BUF_CACHE_TRAIN_RECORDS: {
            if (!c->cache.on) { r = 0; break; }
            uint32_t cnt[4] = {0, 0, 0, 0};
            HIP_CHECK(hipMemcpyAsync(cnt, c->cache.counts.p, 16, hipMemcpyDeviceToHost, c->stream));
            HIP_CHECK(hipStreamSynchronize(c->stream));
            const uint32_t tc = cnt[(c->frame_count & 1) ? 0 : 3];  // the last frame's half
            r = which == VRGI_BUF_CACHE_QUERIES ? (size_t)cnt[1] * 64 : (size_t)std::min(tc, c->cache.tcap) * 64;
            break;
        }
        default: r = 0;
    }
    *bytes = r;
    return VRGI_OK;
    API_END
}

int vrgi_read_buffer(vrgi_ctx *c, int which, void *dst, size_t bytes) {
    API_BEGIN
    if (!c || !dst) return fail(VRGI_E_INVALID, "null argument");
    size_t need = 0;
    vrgi_buffer_size(c, which, &need);
    if (need == 0) return fail(VRGI_E_STATE, "buffer not available");
    if (bytes < need) return fail(VRGI_E_LIMIT, "destination too small");
    set_dev(c);
    const void *src = nullptr;
    size_t P = c->P, nt = c->n_tris;
    switch (which) {
        case VRGI_BUF_GEOM_BVH: src = c->ext_gnodes ? c->ext_gnodes : c->G().d_nodes.p; break;
        case VRGI_BUF_GEOM_PRIM_ORDER: src = c->geo_builder.prim_order(); break;
        case VRGI_BUF_LIGHT_TL_BVH: src = c->d_tl.p; break;
        case VRGI_BUF_LIGHT_GEOM_BVH:
            src = c->ext_lnodes ? c->ext_lnodes : c->G().d_nodes.p + (nt ? 2 * nt - 1 : 0);
            break;
        case VRGI_BUF_PATH_LEN: src = c->f_plen.p; break;
        case VRGI_BUF_BOUNCE_ORIGIN: src = c->f_origin.p; break;
        case VRGI_BUF_BOUNCE_DIRECTION: src = c->f_dir.p; break;
        case VRGI_BUF_BOUNCE_NORMAL: src = c->f_normal.p; break;
        case VRGI_BUF_BOUNCE_EMISSIVITY: src = c->f_emis.p; break;
        case VRGI_BUF_BOUNCE_REFLECTIVITY: src = c->f_refl.p; break;
        case VRGI_BUF_BOUNCE_MIS_WEIGHT: src = c->f_mis.p; break;
        case VRGI_BUF_BOUNCE_BSDF_PDF: src = c->f_pdf.p; break;
        case VRGI_BUF_BOUNCE_NEE_PDF: src = c->f_nee.p; break;
        case VRGI_BUF_HIT_PRIM: src = c->f_hit.p; break;
        case VRGI_BUF_HIT_T: src = c->f_hit_t.p; break;
        case VRGI_BUF_RADIANCE: src = c->f_radiance.p; break;
        case VRGI_BUF_RESTIR: src = c->f_restir.p; break;
        case VRGI_BUF_TEMPORAL_RESERVOIR: {
            uint64_t last = c->frame_count ? c->frame_count - 1 : 0;
            src = c->f_temporal.p + (size_t)(last % 3) * P * 20;
            break;
        }
        case VRGI_BUF_SPATIAL_RESERVOIR: src = c->f_spatial.p; break;
        case VRGI_BUF_BGRA8: src = c->f_bgra.p; break;
        case VRGI_BUF_DI_LIGHT: src = c->di_light.p; break;
        case VRGI_BUF_DI_WEIGHTS: src = c->di_w.p; break;
        case VRGI_BUF_BOUNCE_ALBEDO: src = c->cache.albedo.p; break;
        case VRGI_BUF_CACHE_RADIANCE: src = c->cache.crad.p; break;
        case VRGI_BUF_CACHE_FALLBACK: src = c->cache.t_fb.p; break;
        case VRGI_BUF_CACHE_QUERIES: src = c->cache.qrec.p; break;
        case VRGI_BUF_CACHE_TRAIN_RECORDS:
            src = c->cache.trec.p + (size_t)((c->frame_count & 1) ? 0 : 1) * c->cache.tcap * 16;
            break;
        default: return fail(VRGI_E_INVALID, "unknown buffer id");
    }
    HIP_CHECK(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_build_bl_bvh(vrgi_ctx *c, const float *tri, const float *lum, const uint32_t *ids, uint64_t n,
                      vrgi_bvh_node *nodes, uint32_t *order, float root_box[6], float *root_power) {
    API_BEGIN
    if (!c || !tri || !lum || !ids || !nodes) return fail(VRGI_E_INVALID, "null argument");
    if (n == 0) return fail(VRGI_E_INVALID, "build_bl_bvh needs at least one primitive");
    if (n > (1ull << 30)) return fail(VRGI_E_LIMIT, "more than 2^30 primitives");
    set_dev(c);
    hipStream_t s = c->stream;
    uint32_t m = (uint32_t)n;
    c->a_f0.reserve(9 * (size_t)m);
    c->a_f1.reserve(9 * (size_t)m);
    c->a_f2.reserve(m);
    c->a_u0.reserve(m);
    c->a_u1.reserve(4);
    c->a_u2.reserve(m);
    c->a_nodes.reserve(2 * (size_t)m - 1);
    uint32_t trees[4] = {0, m, 0, 0};
    HIP_CHECK(hipMemcpyAsync(c->a_f0.p, tri, 9 * (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_f2.p, lum, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_u0.p, ids, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_u1.p, trees, 12, hipMemcpyHostToDevice, s));
    launch_prims_from_tris(s, c->a_f0.p, m, c->a_f1.p);
    if (c->api_builder.build(s, c->a_f1.p, c->a_f2.p, c->a_u0.p, m, 1, c->a_u1.p, c->a_u1.p + 1, c->a_u1.p + 2,
                             trees + 2, c->a_nodes.p, c->a_u1.p + 3, order ? c->a_u2.p : nullptr) != 0)
        return fail(VRGI_E_LIMIT, "build exceeded its workspace");
    HIP_CHECK(hipMemcpyAsync(nodes, c->a_nodes.p, (2 * (size_t)m - 1) * 40, hipMemcpyDeviceToHost, s));
    if (order) HIP_CHECK(hipMemcpyAsync(order, c->a_u2.p, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    bl_root(tri, lum, nodes, root_box, root_power);
    return VRGI_OK;
    API_END
}

int vrgi_build_bl_bvh_noalloc(vrgi_ctx *c, const float *tri, const float *lum, const uint32_t *ids, uint64_t n,
                              vrgi_bvh_node *nodes, uint32_t *order, float root_box[6], float *root_power) {
    API_BEGIN
    if (!c || !tri || !lum || !ids || !nodes) return fail(VRGI_E_INVALID, "null argument");
    if (n == 0) return fail(VRGI_E_INVALID, "build_bl_bvh needs at least one primitive");
    if (n > VRGI_NOALLOC_MAX_PRIMS)
        return fail(VRGI_E_LIMIT, "more than VRGI_NOALLOC_MAX_PRIMS primitives for the allocation-free build");
    if (c->device < 0 || c->device >= NA_MAX_DEVICES || !g_na[c->device].ptr)
        return fail(VRGI_E_STATE, "no allocation-free scratch on the context's device");
    set_dev(c);
    hipStream_t s = c->stream;
    NoAllocDev &d = g_na[c->device];
    const uint32_t m = (uint32_t)n;
    float *d_tri = (float *)(d.ptr + NA_TRI), *d_prim9 = (float *)(d.ptr + NA_PRIM9);
    float *d_lum = (float *)(d.ptr + NA_LUM);
    uint32_t *d_ids = (uint32_t *)(d.ptr + NA_IDS), *d_order = (uint32_t *)(d.ptr + NA_ORDER);
    uint32_t *d_trees = (uint32_t *)(d.ptr + NA_TREES);
    vrgi_bvh_node *d_nodes = (vrgi_bvh_node *)(d.ptr + NA_NODES);
    // tree 0 = [0, m) at node 0; the depth word after it starts at zero
    const uint32_t trees[4] = {0, m, 0, 0};
    HIP_CHECK(hipMemcpyAsync(d_tri, tri, 9 * (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_lum, lum, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_ids, ids, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(d_trees, trees, 16, hipMemcpyHostToDevice, s));
    launch_prims_from_tris(s, d_tri, m, d_prim9);
    if (d.builder.build(s, d_prim9, d_lum, d_ids, m, 1, d_trees, d_trees + 1, d_trees + 2, trees + 2, d_nodes,
                        d_trees + 3, order ? d_order : nullptr) != 0)
        return fail(VRGI_E_LIMIT, "build exceeded its workspace");
    HIP_CHECK(hipMemcpyAsync(nodes, d_nodes, (2 * (size_t)m - 1) * 40, hipMemcpyDeviceToHost, s));
    if (order) HIP_CHECK(hipMemcpyAsync(order, d_order, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    bl_root(tri, lum, nodes, root_box, root_power);
    return VRGI_OK;
    API_END
}

int vrgi_build_tl_bvh(vrgi_ctx *c, const float *xf, const float *aabbs, const float *pw, const uint32_t *ids,
                      uint64_t n, vrgi_bvh_node *nodes) {
    API_BEGIN
    if (!c || !xf || !aabbs || !pw || !ids || !nodes) return fail(VRGI_E_INVALID, "null argument");
    if (n == 0) return fail(VRGI_E_INVALID, "build_tl_bvh needs at least one instance");
    if (n > (1ull << 30)) return fail(VRGI_E_LIMIT, "more than 2^30 instances");
    set_dev(c);
    hipStream_t s = c->stream;
    uint32_t m = (uint32_t)n;
    c->a_f0.reserve(12 * (size_t)m);
    c->a_f1.reserve(9 * (size_t)m);
    c->a_f2.reserve(7 * (size_t)m);
    c->a_u0.reserve(m);
    c->a_u1.reserve(4);
    c->a_nodes.reserve(2 * (size_t)m - 1);
    uint32_t trees[4] = {0, m, 0, 0};
    HIP_CHECK(hipMemcpyAsync(c->a_f0.p, xf, 12 * (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_f2.p, aabbs, 6 * (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_f2.p + 6 * (size_t)m, pw, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_u0.p, ids, (size_t)m * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_u1.p, trees, 12, hipMemcpyHostToDevice, s));
    launch_tl_inputs(s, c->a_f2.p, c->a_f0.p, m, c->a_f1.p);
    if (c->api_builder.build(s, c->a_f1.p, c->a_f2.p + 6 * (size_t)m, c->a_u0.p, m, 1, c->a_u1.p, c->a_u1.p + 1,
                             c->a_u1.p + 2, trees + 2, c->a_nodes.p, c->a_u1.p + 3, nullptr) != 0)
        return fail(VRGI_E_LIMIT, "build exceeded its workspace");
    HIP_CHECK(hipMemcpyAsync(nodes, c->a_nodes.p, (2 * (size_t)m - 1) * 40, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return VRGI_OK;
    API_END
}

int vrgi_trace_closest(vrgi_ctx *c, const float *o, const float *d, uint64_t n, float tmin, float tmax,
                       uint32_t *hit, float *t, float *uv) {
    API_BEGIN
    if (!c || !o || !d || !hit || !t || !uv) return fail(VRGI_E_INVALID, "null argument");
    if (n == 0) return VRGI_OK;
    if (n > (1ull << 30)) return fail(VRGI_E_LIMIT, "more than 2^30 rays");
    set_dev(c);
    hipStream_t s = c->stream;
    if (c->inst_dirty) {
        rebuild_instances(c);
        c->inst_dirty = false;
        mark_geo_dirty(c);
    }
    if (c->geo_dirty) {
        rebuild_geometry(c);
        c->geo_dirty = false;
    }
    uint32_t m = (uint32_t)n;
    c->a_f0.reserve(3 * (size_t)m);
    c->a_f1.reserve(3 * (size_t)m);
    c->a_f2.reserve(3 * (size_t)m);
    c->a_u0.reserve(m);
    HIP_CHECK(hipMemcpyAsync(c->a_f0.p, o, 12 * (size_t)m, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->a_f1.p, d, 12 * (size_t)m, hipMemcpyHostToDevice, s));
    launch_trace_closest(s, scene_view(c), c->a_f0.p, c->a_f1.p, m, tmin, tmax, c->a_u0.p, c->a_f2.p,
                         c->a_f2.p + m);
    HIP_CHECK(hipMemcpyAsync(hit, c->a_u0.p, 4 * (size_t)m, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(t, c->a_f2.p, 4 * (size_t)m, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(uv, c->a_f2.p + m, 8 * (size_t)m, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return VRGI_OK;
    API_END
}

}  // extern "C"

// ---------------------------------------------------------------- learned radiance cache
namespace {

uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform(uint64_t &x, float lo, float hi) {
    return lo + (hi - lo) * (float)((double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0));
}

void cache_clear_state(vrgi_ctx *c) {
    auto &C = c->cache;
    hipStream_t s = c->stream;
    HIP_CHECK(hipMemsetAsync(C.mw.p, 0, NRC_N_WEIGHTS * 4, s));
    HIP_CHECK(hipMemsetAsync(C.vw.p, 0, NRC_N_WEIGHTS * 4, s));
    HIP_CHECK(hipMemsetAsync(C.gw.p, 0, NRC_N_WEIGHTS * 4, s));
    HIP_CHECK(hipMemsetAsync(C.mt.p, 0, C.n_table * 4, s));
    HIP_CHECK(hipMemsetAsync(C.vt.p, 0, C.n_table * 4, s));
    HIP_CHECK(hipMemsetAsync(C.gtable.p, 0, C.n_table * 4, s));
    HIP_CHECK(hipMemsetAsync(C.loss.p, 0, 4, s));
    C.steps = 0;
    const float st[4] = {C.cfg.loss_scale, 0.0f, 0.0f, 0.0f};  // scale, good steps, overflow flag, skipped
    C.scale_state.reserve(4);
    HIP_CHECK(hipMemcpyAsync(C.scale_state.p, st, 16, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));  // st is released on return
    launch_nrc_pack(s, C.w32.p, C.wt.p, C.wb.p, C.cfg.precision);
}

void cache_planes(vrgi_ctx *c, uint32_t batch) {
    auto &C = c->cache;
    if (batch <= C.plane_cap) return;
    C.xplane.reserve((size_t)5 * 64 * batch);
    C.gplane.reserve((size_t)5 * 64 * batch);
    C.plane_cap = batch;
}

// caller samples (n x 12) and targets (n x 3, may be null) as records in hrec; count[2] = n
void cache_upload_records(vrgi_ctx *c, uint64_t n, const float *samples, const float *targets) {
    auto &C = c->cache;
    std::vector<float> h((size_t)n * 16, 0.0f);
    for (uint64_t i = 0; i < n; ++i) {
        std::memcpy(&h[i * 16], samples + i * 12, 48);
        if (targets) std::memcpy(&h[i * 16 + 12], targets + i * 3, 12);
    }
    C.hrec.reserve((size_t)n * 16);
    uint32_t cnt = (uint32_t)n;
    HIP_CHECK(hipMemcpyAsync(C.hrec.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(C.counts.p + 2, &cnt, 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));  // h is released on return
}

}  // namespace

extern "C" {

void vrgi_cache_config_default(vrgi_cache_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->log2_table = 18;
    cfg->res_min = 16;
    cfg->res_max = 2048;
    cfg->train_fraction = 0.03f;
    cfg->train_bounces = 8;
    cfg->batch = 16384;
    cfg->steps_per_frame = 4;
    cfg->lr = 1e-2f;
    cfg->beta1 = 0.9f;
    cfg->beta2 = 0.99f;
    cfg->eps = 1e-10f;
    cfg->loss_scale = 128.0f;
    cfg->seed = 1;
    cfg->flags = 0;
    cfg->query_vertex = 0;
    cfg->fallback_bound = 0.0f;
    cfg->fallback_tile = 16;
    cfg->fallback_ema = 0.1f;
}

int vrgi_cache_enable(vrgi_ctx *c, const vrgi_cache_config *cfg) {
    API_BEGIN
    if (!c || !cfg) return fail(VRGI_E_INVALID, "null argument");
    if (cfg->log2_table < 10 || cfg->log2_table > 24) return fail(VRGI_E_INVALID, "log2_table must be in [10, 24]");
    if (cfg->res_min < 2 || cfg->res_max < cfg->res_min || cfg->res_max > (1u << 20))
        return fail(VRGI_E_INVALID, "need 2 <= res_min <= res_max <= 2^20");
    if (cfg->train_bounces < 2 || cfg->train_bounces > 16) return fail(VRGI_E_INVALID, "train_bounces must be in [2, 16]");
    if (cfg->batch < 16 || cfg->batch % 16 || cfg->batch > (1u << 20))
        return fail(VRGI_E_INVALID, "batch must be a multiple of 16 in [16, 2^20]");
    if (cfg->steps_per_frame > 64) return fail(VRGI_E_INVALID, "steps_per_frame above 64");
    if (cfg->query_vertex > 1) return fail(VRGI_E_INVALID, "query_vertex must be 0 or 1");
    if (cfg->precision > VRGI_CACHE_FP16) return fail(VRGI_E_INVALID, "precision must be VRGI_CACHE_BF16 or VRGI_CACHE_FP16");
    if (!(cfg->fallback_bound >= 0.0f) || cfg->fallback_tile == 0 || cfg->fallback_tile > 4096 ||
        !(cfg->fallback_ema > 0.0f && cfg->fallback_ema <= 1.0f))
        return fail(VRGI_E_INVALID, "fallback_bound >= 0, fallback_tile in [1, 4096], fallback_ema in (0, 1]");
    if (cfg->query_vertex + 1 >= cfg->train_bounces) return fail(VRGI_E_INVALID, "train_bounces too small");
    if (!(cfg->train_fraction >= 0.0f && cfg->train_fraction <= 1.0f))
        return fail(VRGI_E_INVALID, "train_fraction must be in [0, 1]");
    if (!(cfg->lr > 0.0f) || !(cfg->loss_scale > 0.0f)) return fail(VRGI_E_INVALID, "lr and loss_scale must be positive");
    if (c->in_batch) return fail(VRGI_E_STATE, "the scene is in a batch (no learned cache there)");
    bool box_zero = true;
    for (int k = 0; k < 3; ++k) box_zero = box_zero && cfg->box_lo[k] == 0.0f && cfg->box_hi[k] == 0.0f;
    if (!box_zero)
        for (int k = 0; k < 3; ++k)
            if (!(cfg->box_hi[k] > cfg->box_lo[k])) return fail(VRGI_E_INVALID, "box_hi must exceed box_lo");
    set_dev(c);
    auto &C = c->cache;
    C.cfg = *cfg;
    const double b = std::exp((std::log((double)cfg->res_max) - std::log((double)cfg->res_min)) / (NRC_LEVELS - 1));
    for (int l = 0; l < 16; ++l) C.res[l] = l < NRC_LEVELS ? (uint32_t)std::floor(cfg->res_min * std::pow(b, l)) : 0u;
    C.box_ready = !box_zero;
    for (int k = 0; k < 3; ++k) { C.lo[k] = cfg->box_lo[k]; C.hi[k] = cfg->box_hi[k]; }
    C.n_table = (size_t)NRC_LEVELS * ((size_t)1 << cfg->log2_table) * 2;
    C.w32.reserve(NRC_N_WEIGHTS); C.mw.reserve(NRC_N_WEIGHTS); C.vw.reserve(NRC_N_WEIGHTS); C.gw.reserve(NRC_N_WEIGHTS);
    C.table.reserve(C.n_table); C.mt.reserve(C.n_table); C.vt.reserve(C.n_table); C.gtable.reserve(C.n_table);
    C.wt.reserve(NRC_N_WEIGHTS); C.wb.reserve(NRC_N_WEIGHTS);
    C.table_snap.reserve(C.n_table); C.wt_snap.reserve(NRC_N_WEIGHTS);
    C.loss.reserve(1);
    C.counts.reserve(4);
    C.plane_cap = 0;
    cache_planes(c, cfg->batch);
    C.tcap = cfg->batch * std::max(cfg->steps_per_frame, 1u);
    C.trec.reserve((size_t)2 * C.tcap * 16);
    // initial parameters: Kaiming-uniform weights (no bias vectors), table U(-1e-4, 1e-4)
    uint64_t st = 0x5EEDull ^ ((uint64_t)cfg->seed << 32);
    std::vector<float> w(NRC_N_WEIGHTS, 0.0f), t(C.n_table);
    const float bound = std::sqrt(6.0f / 64.0f);
    for (int i = 0; i < 4 * 64 * 64; ++i) w[i] = uniform(st, -bound, bound);
    for (int k = 0; k < 64; ++k)
        for (int n = 0; n < 3; ++n) w[4 * 4096 + k * 16 + n] = uniform(st, -bound, bound);
    for (size_t i = 0; i < C.n_table; ++i) t[i] = uniform(st, -1e-4f, 1e-4f);
    HIP_CHECK(hipMemcpyAsync(C.w32.p, w.data(), w.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(C.table.p, t.data(), t.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemsetAsync(C.counts.p, 0, 16, c->stream));
    cache_clear_state(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));  // w, t are released on return
    C.on = true;
    c->B = cfg->train_bounces;
    alloc_frame(c);
    return VRGI_OK;
    API_END
}

int vrgi_cache_disable(vrgi_ctx *c) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    set_dev(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    c->cache.on = false;
    c->B = c->cfg.num_bounces;
    alloc_frame(c);
    return VRGI_OK;
    API_END
}

int vrgi_cache_info_get(vrgi_ctx *c, vrgi_cache_info *out) {
    API_BEGIN
    if (!c || !out) return fail(VRGI_E_INVALID, "null argument");
    if (!c->cache.on && !c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    set_dev(c);
    auto &C = c->cache;
    std::memset(out, 0, sizeof(*out));
    out->levels = NRC_LEVELS;
    out->feats = 2;
    out->log2_table = C.cfg.log2_table;
    std::memcpy(out->res, C.res, sizeof(out->res));
    for (int k = 0; k < 3; ++k) { out->box_lo[k] = C.lo[k]; out->box_hi[k] = C.hi[k]; }
    out->table_floats = C.n_table;
    out->steps = C.steps;
    out->precision = C.cfg.precision;
    float l = 0.0f, st[4] = {C.cfg.loss_scale, 0.0f, 0.0f, 0.0f};
    HIP_CHECK(hipMemcpyAsync(&l, C.loss.p, 4, hipMemcpyDeviceToHost, c->stream));
    if (cache_dynamic_scale(C.cfg) && C.scale_state.p)
        HIP_CHECK(hipMemcpyAsync(st, C.scale_state.p, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    out->loss_scale = st[0];
    out->skipped_steps = (uint64_t)st[3];
    out->last_loss = C.cfg.steps_per_frame ? l / (float)C.cfg.steps_per_frame : l;
    return VRGI_OK;
    API_END
}

int vrgi_cache_get(vrgi_ctx *c, float *weights, float *table) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (!c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    set_dev(c);
    if (weights) HIP_CHECK(hipMemcpyAsync(weights, c->cache.w32.p, NRC_N_WEIGHTS * 4, hipMemcpyDeviceToHost, c->stream));
    if (table) HIP_CHECK(hipMemcpyAsync(table, c->cache.table.p, c->cache.n_table * 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_cache_set(vrgi_ctx *c, const float *weights, const float *table) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (!c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    set_dev(c);
    if (weights) HIP_CHECK(hipMemcpyAsync(c->cache.w32.p, weights, NRC_N_WEIGHTS * 4, hipMemcpyHostToDevice, c->stream));
    if (table) HIP_CHECK(hipMemcpyAsync(c->cache.table.p, table, c->cache.n_table * 4, hipMemcpyHostToDevice, c->stream));
    cache_clear_state(c);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_cache_infer(vrgi_ctx *c, uint64_t n, const float *samples, float *out) {
    API_BEGIN
    if (!c || (n && (!samples || !out))) return fail(VRGI_E_INVALID, "null argument");
    if (!c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    if (n > (1u << 26)) return fail(VRGI_E_LIMIT, "more than 2^26 samples");
    if (n == 0) return VRGI_OK;
    set_dev(c);
    auto &C = c->cache;
    if (!C.box_ready) return fail(VRGI_E_STATE, "the cache box is set at the first render (or give one)");
    cache_upload_records(c, n, samples, nullptr);
    C.hout.reserve((size_t)n * 3);
    launch_nrc_infer(c->stream, cache_net(c), C.hrec.p, C.counts.p + 2, (uint32_t)n, C.hout.p, nullptr);
    HIP_CHECK(hipMemcpyAsync(out, C.hout.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return VRGI_OK;
    API_END
}

int vrgi_cache_grads(vrgi_ctx *c, uint64_t n, const float *samples, const float *targets, float *grad_w,
                     float *grad_table, float *loss) {
    API_BEGIN
    if (!c || !samples || !targets || !n) return fail(VRGI_E_INVALID, "null argument or empty batch");
    if (!c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    if (n > (1u << 20)) return fail(VRGI_E_LIMIT, "more than 2^20 samples");
    set_dev(c);
    auto &C = c->cache;
    if (!C.box_ready) return fail(VRGI_E_STATE, "the cache box is set at the first render (or give one)");
    const uint32_t mb = (uint32_t)((n + 15) / 16 * 16);
    cache_planes(c, mb);
    cache_upload_records(c, n, samples, targets);
    hipStream_t s = c->stream;
    HIP_CHECK(hipMemsetAsync(C.gw.p, 0, NRC_N_WEIGHTS * 4, s));
    HIP_CHECK(hipMemsetAsync(C.gtable.p, 0, C.n_table * 4, s));
    HIP_CHECK(hipMemsetAsync(C.loss.p, 0, 4, s));
    cache_step(c, s, C.hrec.p, C.counts.p + 2, (uint32_t)n, 0, 1, mb, false);
    std::vector<float> gw(NRC_N_WEIGHTS);
    HIP_CHECK(hipMemcpyAsync(gw.data(), C.gw.p, NRC_N_WEIGHTS * 4, hipMemcpyDeviceToHost, s));
    if (grad_table) HIP_CHECK(hipMemcpyAsync(grad_table, C.gtable.p, C.n_table * 4, hipMemcpyDeviceToHost, s));
    float l = 0.0f;
    HIP_CHECK(hipMemcpyAsync(&l, C.loss.p, 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    const float inv = 1.0f / C.cfg.loss_scale;
    if (grad_w)
        for (int i = 0; i < NRC_N_WEIGHTS; ++i) grad_w[i] = gw[i] * inv;
    if (grad_table)
        for (size_t i = 0; i < C.n_table; ++i) grad_table[i] *= inv;
    if (loss) *loss = l;
    HIP_CHECK(hipMemsetAsync(C.gw.p, 0, NRC_N_WEIGHTS * 4, s));
    HIP_CHECK(hipMemsetAsync(C.gtable.p, 0, C.n_table * 4, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return VRGI_OK;
    API_END
}

int vrgi_mesh_chunks(vrgi_ctx *c, const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                     const uint8_t *translucent, uint32_t n_chunks, const uint32_t *chunks, vrgi_vertex3d *out,
                     uint64_t cap, uint64_t *counts) {
    API_BEGIN
    if (!c || !blocks || !dims || !translucent || (n_chunks && (!chunks || !counts)))
        return fail(VRGI_E_INVALID, "null argument");
    if (n_types == 0 || n_types > 65535) return fail(VRGI_E_INVALID, "n_types must be in [1, 65535]");
    if (n_chunks == 0) return VRGI_OK;
    if (n_chunks > 65535) return fail(VRGI_E_LIMIT, "more than 65535 chunks in one call");
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    if (nvox == 0 || nvox > ((size_t)1 << 32)) return fail(VRGI_E_INVALID, "world size out of range");
    set_dev(c);
    hipStream_t s = c->stream;
    c->m_blocks.reserve(nvox);
    c->m_transl.reserve(n_types + 1);
    c->m_chunks.reserve((size_t)3 * n_chunks);
    c->m_mask.reserve((size_t)n_chunks * 32768);
    c->m_face_off.reserve((size_t)n_chunks * 32768);
    c->m_faces.reserve(n_chunks);
    c->m_vert_off.reserve(n_chunks);
    std::vector<uint8_t> tr(translucent, translucent + n_types);
    tr.push_back(0);  // air (never looked up: air is tested first)
    HIP_CHECK(hipMemcpyAsync(c->m_blocks.p, blocks, nvox * 2, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->m_transl.p, tr.data(), tr.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c->m_chunks.p, chunks, (size_t)12 * n_chunks, hipMemcpyHostToDevice, s));
    MeshArgsHost h{};
    h.blocks = c->m_blocks.p;
    for (int k = 0; k < 3; ++k) h.dims[k] = dims[k];
    h.air = n_types;
    h.translucent = c->m_transl.p;
    h.chunks = c->m_chunks.p;
    h.n_chunks = n_chunks;
    h.mask = c->m_mask.p;
    h.face_off = c->m_face_off.p;
    h.faces = c->m_faces.p;
    launch_mesh_count(s, h);
    std::vector<uint32_t> faces(n_chunks);
    HIP_CHECK(hipMemcpyAsync(faces.data(), c->m_faces.p, (size_t)4 * n_chunks, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint64_t> off(n_chunks);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_chunks; ++i) {
        off[i] = total;
        counts[i] = (uint64_t)faces[i] * 6;
        total += counts[i];
    }
    if (total > cap) return fail(VRGI_E_LIMIT, "vertex capacity too small");
    if (total == 0) return VRGI_OK;
    if (!out) return fail(VRGI_E_INVALID, "null output");
    c->m_out.reserve(total);
    HIP_CHECK(hipMemcpyAsync(c->m_vert_off.p, off.data(), (size_t)8 * n_chunks, hipMemcpyHostToDevice, s));
    h.vert_off = c->m_vert_off.p;
    h.out = c->m_out.p;
    launch_mesh_emit(s, h);
    HIP_CHECK(hipMemcpyAsync(out, c->m_out.p, total * sizeof(vrgi_vertex3d), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return VRGI_OK;
    API_END
}

int vrgi_voxels_load(vrgi_ctx *c, const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                     const uint8_t *translucent) {
    API_BEGIN
    if (!c || !blocks || !dims || !translucent) return fail(VRGI_E_INVALID, "null argument");
    for (int k = 0; k < 3; ++k)
        if (dims[k] == 0 || dims[k] % 32 || dims[k] > 32768) return fail(VRGI_E_INVALID, "dims must be multiples of 32 in [32, 32768]");
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    if (nvox > ((size_t)1 << 32)) return fail(VRGI_E_LIMIT, "more than 2^32 voxels");
    if (n_types == 0 || (size_t)n_types * 6 > c->lum.size()) return fail(VRGI_E_INVALID, "n_types * 6 exceeds the material slots");
    for (size_t i = 0; i < nvox; ++i)
        if (blocks[i] > n_types) return fail(VRGI_E_INVALID, "block id above n_types (air)");
    set_dev(c);
    auto &V = c->vx;
    HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 3; ++k) { V.dims[k] = dims[k]; V.nc[k] = dims[k] / 32; }
    V.n_types = n_types;
    const uint32_t nch = V.nc[0] * V.nc[1] * V.nc[2];
    V.blocks.reserve(nvox);
    V.transl.reserve(n_types + 1);
    V.em.reserve(n_types + 1);
    V.lum.reserve(c->lum.size());
    V.dirty.reserve(nch);
    std::vector<uint8_t> tr(translucent, translucent + n_types), em(n_types + 1, 0);
    tr.push_back(0);
    for (uint32_t b = 0; b < n_types; ++b)
        for (int f = 0; f < 6; ++f)
            if (c->lum[(size_t)b * 6 + f] > 0.0f) em[b] |= (uint8_t)(1u << f);
    HIP_CHECK(hipMemcpyAsync(V.blocks.p, blocks, nvox * 2, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(V.transl.p, tr.data(), tr.size(), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(V.em.p, em.data(), em.size(), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(V.lum.p, c->lum.data(), c->lum.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemsetAsync(V.dirty.p, 1, nch, c->stream));
    V.on = true;
    V.pending = true;
    c->view_dirty = true;
    voxels_commit(c);
    return VRGI_OK;
    API_END
}

int vrgi_voxels_set(vrgi_ctx *c, uint64_t n, const uint32_t *xyz, const uint16_t *blocks) {
    API_BEGIN
    if (!c || (n && (!xyz || !blocks))) return fail(VRGI_E_INVALID, "null argument");
    auto &V = c->vx;
    if (!V.on) return fail(VRGI_E_STATE, "no voxel world loaded");
    if (n > (1ull << 31)) return fail(VRGI_E_LIMIT, "more than 2^31 voxel writes");
    for (uint64_t i = 0; i < n; ++i) {
        if (blocks[i] > V.n_types) return fail(VRGI_E_INVALID, "block id above n_types (air)");
        if (xyz[3 * i] >= V.dims[0] || xyz[3 * i + 1] >= V.dims[1] || xyz[3 * i + 2] >= V.dims[2])
            return fail(VRGI_E_INVALID, "voxel outside the world");
    }
    if (n == 0) return VRGI_OK;
    set_dev(c);
    V.edit_xyz.reserve(3 * n);
    V.edit_val.reserve(n);
    HIP_CHECK(hipMemcpyAsync(V.edit_xyz.p, xyz, (size_t)12 * n, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipMemcpyAsync(V.edit_val.p, blocks, (size_t)2 * n, hipMemcpyHostToDevice, c->stream));
    launch_vx_set(c->stream, V.blocks.p, V.dims, n, V.edit_xyz.p, V.edit_val.p, V.dirty.p);
    V.pending = true;
    c->view_dirty = true;
    return VRGI_OK;
    API_END
}

int vrgi_voxels_commit(vrgi_ctx *c) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    if (!c->vx.on) return fail(VRGI_E_STATE, "no voxel world loaded");
    set_dev(c);
    voxels_commit(c);
    return VRGI_OK;
    API_END
}

int vrgi_voxels_read(vrgi_ctx *c, uint16_t *blocks, uint64_t *commits, uint64_t *chunks_remeshed) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    auto &V = c->vx;
    if (!V.on) return fail(VRGI_E_STATE, "no voxel world loaded");
    set_dev(c);
    if (blocks) {
        HIP_CHECK(hipMemcpyAsync(blocks, V.blocks.p, (size_t)V.dims[0] * V.dims[1] * V.dims[2] * 2,
                                 hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    if (commits) *commits = V.commits;
    if (chunks_remeshed) *chunks_remeshed = V.chunks_remeshed;
    return VRGI_OK;
    API_END
}

int vrgi_frame_set(vrgi_ctx *c, uint64_t frame) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    c->frame_count = frame;
    return VRGI_OK;
    API_END
}

int vrgi_accum_enable(vrgi_ctx *c, int on) {
    API_BEGIN
    if (!c) return fail(VRGI_E_INVALID, "null context");
    set_dev(c);
    c->accum_on = on != 0;
    c->accum_frames = 0;
    if (c->accum_on) {
        c->accum.reserve((size_t)c->P * 3);
        HIP_CHECK(hipMemsetAsync(c->accum.p, 0, (size_t)c->P * 24, c->stream));
        c->accum_P = c->P;
    }
    return VRGI_OK;
    API_END
}

int vrgi_accum_read(vrgi_ctx *c, double *mean, uint64_t *frames) {
    API_BEGIN
    if (!c || !mean) return fail(VRGI_E_INVALID, "null argument");
    if (!c->accum.p) return fail(VRGI_E_STATE, "accumulation was never enabled");
    if (c->accum_P != c->P) return fail(VRGI_E_STATE, "the extent changed since accumulation was enabled");
    set_dev(c);
    HIP_CHECK(hipMemcpyAsync(mean, c->accum.p, (size_t)c->P * 24, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    const double inv = c->accum_frames ? 1.0 / (double)c->accum_frames : 0.0;
    for (size_t i = 0; i < (size_t)c->P * 3; ++i) mean[i] *= inv;
    if (frames) *frames = c->accum_frames;
    return VRGI_OK;
    API_END
}

int vrgi_cache_train(vrgi_ctx *c, uint64_t n, const float *samples, const float *targets, float *loss) {
    API_BEGIN
    if (!c || !samples || !targets || !n) return fail(VRGI_E_INVALID, "null argument or empty batch");
    if (!c->cache.w32.p) return fail(VRGI_E_STATE, "the cache was never enabled");
    if (n > (1u << 20)) return fail(VRGI_E_LIMIT, "more than 2^20 samples");
    set_dev(c);
    auto &C = c->cache;
    if (!C.box_ready) return fail(VRGI_E_STATE, "the cache box is set at the first render (or give one)");
    const uint32_t mb = (uint32_t)((n + 15) / 16 * 16);
    cache_planes(c, mb);
    cache_upload_records(c, n, samples, targets);
    HIP_CHECK(hipMemsetAsync(C.loss.p, 0, 4, c->stream));
    cache_step(c, c->stream, C.hrec.p, C.counts.p + 2, (uint32_t)n, 0, 1, mb, true);
    float l = 0.0f;
    HIP_CHECK(hipMemcpyAsync(&l, C.loss.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return VRGI_OK;
    API_END
}

}  // extern "C"

// ==================================================================== batched scenes
struct vrgi_batch {
    int device = 0;
    std::vector<vrgi_ctx *> sc;
    vrgi_batch_config cfg{};
    uint32_t B = 0;
    hipStream_t stream = nullptr;
    Builder builder;
    DevBuf<float> prim9, power;
    DevBuf<uint32_t> ids, trees, depth;
    DevBuf<vrgi_bvh_node> nodes;
    // per-pass arguments: one device block, a ring of pinned host blocks (each reused only
    // after its copy has run)
    size_t off_sv = 0, off_tba = 0, off_nba = 0, off_fa = 0, off_da = 0, off_pass = 0, args_bytes = 0;
    DevBuf<uint8_t> d_args;
    static constexpr int RING = 4;
    uint8_t *h_args[RING] = {};
    hipEvent_t ev_args[RING] = {};
    bool ev_used[RING] = {};
    uint32_t ring = 0;
    // block prefixes of the current tiles, mean buffers
    DevBuf<uint32_t> d_blk;
    uint32_t blk_wpb = 0, blk_pb = 0, ray_blocks = 0, px_blocks = 0;
    std::vector<DevBuf<float>> mean;
    DevBuf<float *> d_mean;
    // graph of one pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint32_t> graph_key;
    // Frame graph: the replayed builds and every pass of a batched frame in one graph,
    // with the passes' arguments in spp consecutive regions of f_args (one copy per frame)
    static constexpr uint32_t FRAME_GRAPH_SPP = 8;  // larger spp: one graph per pass
    DevBuf<uint8_t> f_args;
    uint8_t *h_fargs[RING] = {};
    hipGraph_t fgraph = nullptr;
    hipGraphExec_t fexec = nullptr;
    std::vector<uint32_t> fgraph_key;
    // plan of the last batched build made with readbacks, which a rebuild of the same
    // scenes' unchanged geometry replays: the scenes (and their geometry buffer sets), the
    // trees' depths; gen counts such builds (graphs holding a replay key on it)
    bool plan_valid = false;
    std::vector<vrgi_ctx *> plan_list;
    std::vector<uint32_t> plan_gi, plan_dep;
    uint32_t plan_gen = 0;
    DevBuf<uint32_t> d_check;
    uint32_t *h_check = nullptr;
    hipEvent_t ev_check = nullptr;
    bool check_pending = false;
    bool last_frame_graph = false;  // the last batched frame ran as one graph
    // tiles: fixed, or tuned over the first batched frames
    uint32_t wpb = 0, pb = 0;
    int tune_step = 0;  // 0..2: rays per block 64/128/256; 3..5: pixels per block 64/128/256
    float tune_ms[6] = {};
    bool tuned = false;
    hipEvent_t e_b0 = nullptr, e_b1 = nullptr, e_end = nullptr;
    bool have_times = false;
    uint64_t rays_seen = 0;  // closest-hit rays counted at the last vrgi_batch_info_get
    ~vrgi_batch();
};

namespace {

const uint32_t BATCH_TILES[3] = {64, 128, 256};

// the device word of replayed builds' plan checks and its pinned copy (set up before any
// capture: no allocation may be recorded into a graph)
void batch_check_init(vrgi_batch *bt) {
    if (bt->h_check) return;
    bt->d_check.reserve(1);
    HIP_CHECK(hipMemsetAsync(bt->d_check.p, 0, 4, bt->stream));
    HIP_CHECK(hipHostMalloc((void **)&bt->h_check, 4, hipHostMallocDefault));
    *bt->h_check = 0;
    HIP_CHECK(hipEventCreateWithFlags(&bt->ev_check, hipEventDisableTiming));
}
// reads the previous replay's check (its copy is queued behind that frame)
void batch_check_read(vrgi_batch *bt) {
    if (!bt->check_pending) return;
    HIP_CHECK(hipEventSynchronize(bt->ev_check));
    bt->check_pending = false;
    if (*bt->h_check) throw std::runtime_error("a replayed batched build left its plan (a scene's geometry changed unnoticed)");
}

// whether the listed scenes' build can replay the batch's plan: the same scenes, in the
// same geometry buffer sets, none of whose geometry changed since that build
bool batch_replay_ok(const vrgi_batch *bt, const std::vector<vrgi_ctx *> &list) {
    if (!bt->plan_valid || list != bt->plan_list) return false;
    for (size_t k = 0; k < list.size(); ++k)
        if (list[k]->geo_dirty || list[k]->gi != bt->plan_gi[k]) return false;
    return true;
}

// One multi-tree build over the geometry and light-geometry trees of the listed scenes.
// replay (batch_replay_ok): the plan's build repeated with no readback, host state set
// from the plan (and, with launch unset, only that: the launches are in a frame graph).
void batch_build(vrgi_batch *bt, const std::vector<vrgi_ctx *> &list, bool replay = false, bool launch = true) {
    hipStream_t s = bt->stream;
    struct Part {
        vrgi_ctx *c;
        uint32_t off, nt, nl, gbase, lbase;
        int gt, lt;
    };
    std::vector<Part> parts;
    std::vector<uint32_t> beg, cnt, base;
    uint64_t N = 0, nodes = 0;
    for (vrgi_ctx *c : list) {
        Part p{c, (uint32_t)N, c->n_tris, c->n_ltri, 0, 0, -1, -1};
        if (p.nt) { p.gt = (int)beg.size(); beg.push_back(p.off); cnt.push_back(p.nt); base.push_back((uint32_t)nodes); p.gbase = (uint32_t)nodes; nodes += 2ull * p.nt - 1; }
        if (p.nl) { p.lt = (int)beg.size(); beg.push_back(p.off + p.nt); cnt.push_back(p.nl); base.push_back((uint32_t)nodes); p.lbase = (uint32_t)nodes; nodes += 2ull * p.nl - 1; }
        N += (uint64_t)p.nt + p.nl;
        parts.push_back(p);
        if (N >= (1ull << 31) || nodes >= (1ull << 32) - 1) throw std::runtime_error("batched build above 2^31 primitives");
    }
    const uint32_t nT = (uint32_t)beg.size();
    for (Part &p : parts) {
        vrgi_ctx *c = p.c;
        auto &G = c->G();
        G.geo_depth = G.lgeo_depth = 0;
        G.lg_on = false;  // a batched build makes no light grid
        c->ext_gnodes = c->ext_lnodes = nullptr;
    }
    if (N == 0) return;
    if (replay && bt->plan_dep.size() != nT) throw std::logic_error("batched build replay without its plan");
    bt->prim9.reserve((size_t)9 * N);
    bt->power.reserve(N);
    bt->ids.reserve(N);
    bt->nodes.reserve(nodes);
    bt->trees.reserve(3 * (size_t)nT);
    bt->depth.reserve(nT);
    for (Part &p : parts) {
        vrgi_ctx *c = p.c;
        auto &G = c->G();
        G.d_tri.reserve(p.nt);
        G.d_tri_inst.reserve(p.nt);
        G.d_ltri.reserve(p.nl);
        G.d_ltri_inst.reserve(p.nl);
        G.d_ltri_node.reserve(p.nl);
        G.d_gwide.reserve(p.nt > 1 ? p.nt - 1 : 1);
        G.d_lwide.reserve(p.nl > 1 ? p.nl - 1 : 1);
        G.d_lwide4.reserve(p.nl > 1 ? p.nl - 1 : 1);
        G.d_lanc.reserve((size_t)p.nl * VRGI_LANC_MAX);
        G.d_lanc_len.reserve(p.nl);
        if (c->tex_on) G.d_tri_uv.reserve((size_t)3 * p.nt);
        if (!launch) continue;
        FlattenArgs f{};
        f.inst = c->d_inst.p;
        f.n_inst = c->n_inst;
        f.vtx = c->d_vtxptr.p;
        f.lvtx = c->d_lvtxptr.p;
        f.light_off = c->d_light_off.p;
        f.prim_stride = (uint32_t)N;
        f.n_tris = p.nt;
        f.tri = G.d_tri.p;
        f.tri_inst = G.d_tri_inst.p;
        f.prim9 = bt->prim9.p;
        f.power = bt->power.p;
        f.ids = bt->ids.p;
        f.prim_off = p.off;
        if (c->tex_on) f.tri_uv = G.d_tri_uv.p;
        launch_flatten(s, f);
        FlattenArgs fl = f;
        fl.n_tris = p.nl;
        fl.tri = G.d_ltri.p;
        fl.tri_inst = G.d_ltri_inst.p;
        fl.light_node = G.d_ltri_node.p;
        fl.prim_off = p.off + p.nt;
        launch_flatten_lights(s, fl);
        launch_light_ancestry(s, G.d_ltri_inst.p, G.d_ltri_node.p, c->d_inst.p, c->d_tl.p, p.nl, G.d_lanc.p,
                              G.d_lanc_len.p);
    }
    std::vector<uint32_t> dep(nT);
    if (replay) {
        dep = bt->plan_dep;
        if (launch) {  // the tree table is the plan's, still on the device
            if (!bt->h_check) throw std::logic_error("batched build replay before its check was set up");
            if (bt->builder.build(s, bt->prim9.p, bt->power.p, bt->ids.p, (uint32_t)N, nT, bt->trees.p,
                                  bt->trees.p + nT, bt->trees.p + 2 * nT, base.data(), bt->nodes.p, bt->depth.p,
                                  nullptr, true, bt->d_check.p) != 0)
                throw HipError("batched build exceeded its workspace");
        }
    } else {
        std::vector<uint32_t> tr(3 * (size_t)nT);
        for (uint32_t t = 0; t < nT; ++t) { tr[t] = beg[t]; tr[nT + t] = cnt[t]; tr[2 * nT + t] = base[t]; }
        HIP_CHECK(hipMemcpyAsync(bt->trees.p, tr.data(), tr.size() * 4, hipMemcpyHostToDevice, s));
        if (bt->builder.build(s, bt->prim9.p, bt->power.p, bt->ids.p, (uint32_t)N, nT, bt->trees.p,
                              bt->trees.p + nT, bt->trees.p + 2 * nT, base.data(), bt->nodes.p, bt->depth.p,
                              nullptr) != 0)
            throw HipError("batched build exceeded its workspace");
        HIP_CHECK(hipMemcpyAsync(dep.data(), bt->depth.p, nT * 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        bt->plan_valid = bt->builder.plan().valid;
        bt->plan_list = list;
        bt->plan_gi.clear();
        for (vrgi_ctx *c : list) bt->plan_gi.push_back(c->gi);
        bt->plan_dep = dep;
        ++bt->plan_gen;
    }
    for (Part &p : parts) {
        vrgi_ctx *c = p.c;
        auto &G = c->G();
        if (p.gt >= 0) G.geo_depth = dep[p.gt];
        if (p.lt >= 0) G.lgeo_depth = dep[p.lt];
        if (geom_stack_need(G.geo_depth) + 2 > VRGI_MAX_STACK || 3 * ((G.lgeo_depth + 1) / 2) + 2 > VRGI_MAX_STACK)
            throw std::runtime_error("BVH deeper than the traversal stack (geometry depth " +
                                     std::to_string(G.geo_depth) + ", light-geometry depth " +
                                     std::to_string(G.lgeo_depth) + ")");
        if (launch) {
            if (p.nt) launch_wide(s, bt->nodes.p + p.gbase, 2 * p.nt - 1, G.d_gwide.p);
            if (p.nl) {
                launch_wide(s, bt->nodes.p + p.lbase, 2 * p.nl - 1, G.d_lwide.p);
                launch_wide4(s, bt->nodes.p + p.lbase, 2 * p.nl - 1, G.d_lwide4.p);
            }
        }
        c->ext_gnodes = bt->nodes.p + p.gbase;
        c->ext_lnodes = bt->nodes.p + p.lbase;
        c->geo_dirty = false;
    }
}

// block prefixes for `wpb` waves of rays and `pb` pixels per block
void batch_tables(vrgi_batch *bt, uint32_t wpb, uint32_t pb) {
    if (bt->blk_wpb == wpb && bt->blk_pb == pb) return;
    const uint32_t n = (uint32_t)bt->sc.size();
    std::vector<uint32_t> t(2 * ((size_t)n + 1));
    uint64_t r = 0, q = 0;
    for (uint32_t k = 0; k < n; ++k) {
        t[k] = (uint32_t)r;
        t[n + 1 + k] = (uint32_t)q;
        r += (bt->sc[k]->P + 64 * wpb - 1) / (64 * wpb);
        q += (bt->sc[k]->P + pb - 1) / pb;
    }
    if (r >= (1ull << 31) || q >= (1ull << 31)) throw std::runtime_error("batch grid too large");
    t[n] = (uint32_t)r;
    t[2 * n + 1] = (uint32_t)q;
    bt->d_blk.reserve(t.size());
    HIP_CHECK(hipStreamSynchronize(bt->stream));
    HIP_CHECK(hipMemcpy(bt->d_blk.p, t.data(), t.size() * 4, hipMemcpyHostToDevice));
    bt->blk_wpb = wpb;
    bt->blk_pb = pb;
    bt->ray_blocks = (uint32_t)r;
    bt->px_blocks = (uint32_t)q;
}

// the arguments of one pass of scene k, as vrgi_render forms them for a frame without the
// learned cache, ray sorting or profiling
void batch_fill(vrgi_batch *bt, uint8_t *h, uint32_t k, const vrgi_camera &cam, const vrgi_prefs &pr) {
    vrgi_ctx *c = bt->sc[k];
    const uint32_t P = c->P, B = c->B;
    const uint32_t frame = (uint32_t)c->frame_count;
    ((SceneView *)(h + bt->off_sv))[k] = scene_view(c);
    BounceArgs ba{};
    ba.B = B; ba.W = c->W; ba.H = c->H; ba.P = P;
    ba.nee_type = pr.nee_type;
    ba.rr_start = c->cfg.rr_start_bounce;
    ba.rr_min = c->cfg.rr_min_survival;
    ba.cam = cam;
    ba.qcount = c->f_qcount.p;
    ba.origin = c->f_origin.p; ba.dir = c->f_dir.p; ba.normal = c->f_normal.p;
    ba.emis = c->f_emis.p; ba.refl = c->f_refl.p; ba.mis = c->f_mis.p; ba.pdf = c->f_pdf.p;
    ba.nee = c->f_nee.p; ba.hit_t = c->f_hit_t.p; ba.hit = c->f_hit.p; ba.path_len = c->f_plen.p;
    ba.counters = (unsigned long long *)c->d_counters.p;
    BounceArgs *tba = (BounceArgs *)(h + bt->off_tba) + (size_t)k * B;
    BounceArgs *nba = (BounceArgs *)(h + bt->off_nba) + (size_t)k * B;
    for (uint32_t b = 0; b < B; ++b) {
        ba.bounce = b;
        ba.inv_seed = frame * B + b;  // interactive_rendering.rs:1189
        ba.queue_in = b == 0 ? nullptr : c->f_queue.p + (size_t)b * P;
        ba.queue_out = c->f_queue.p + (size_t)(b + 1) * P;
        tba[b] = ba;
        BounceArgs na = ba;  // NEE pdf of bounce b: the rays of bounce b + 1
        na.queue_in = c->f_queue.p + (size_t)(b + 1) * P;
        nba[b] = na;
    }
    FrameArgs fa{};
    fa.W = c->W; fa.H = c->H; fa.P = P; fa.B = B; fa.frame = frame;
    fa.slot = (uint32_t)(c->frame_count % 3);
    fa.tseed = frame_seed_host(c->cfg.seed, frame, 2);
    fa.sseed = frame_seed_host(c->cfg.seed, frame, 3);
    fa.spatial_iters = pr.restir_spatial_iterations;
    fa.radius = c->cfg.restir_spatial_radius;
    fa.debug_view = pr.debug_view;
    fa.origin = c->f_origin.p; fa.normal = c->f_normal.p; fa.emis = c->f_emis.p; fa.refl = c->f_refl.p;
    fa.mis = c->f_mis.p; fa.pdf = c->f_pdf.p; fa.nee = pr.nee_type ? c->f_nee.p : nullptr;
    fa.path_len = c->f_plen.p;
    fa.radiance = c->f_radiance.p;
    fa.temporal = c->f_temporal.p + (size_t)fa.slot * P * 20;
    fa.spatial = c->f_spatial.p;
    fa.restir = c->f_restir.p;
    fa.debug = c->f_debug.p;
    fa.bgra8 = c->f_bgra.p;
    ((FrameArgs *)(h + bt->off_fa))[k] = fa;
    DiArgs da{};
    if (c->cfg.restir_mode == 1) {
        da.W = c->W; da.H = c->H; da.P = P; da.frame = frame;
        da.candidates = c->cfg.di_candidates;
        da.spatial_iters = pr.restir_spatial_iterations;
        da.radius = c->cfg.restir_spatial_radius;
        da.dseed = frame_seed_host(c->cfg.seed, frame, 4);
        da.origin = c->f_origin.p; da.normal = c->f_normal.p; da.emis = c->f_emis.p; da.hit = c->f_hit.p;
        da.T = c->di_T.p;
        da.hist_prev = c->di_hist.p + (size_t)((frame + 1) % 2) * P * 8;
        da.hist_cur = c->di_hist.p + (size_t)(frame % 2) * P * 8;
        da.restir = c->f_restir.p;
        da.di_light = c->di_light.p;
        da.di_weights = c->di_w.p;
    }
    ((DiArgs *)(h + bt->off_da))[k] = da;
}

const char *batch_scene_ok(const vrgi_ctx *c, const vrgi_ctx *first) {
    if (c->device != first->device) return "batched scenes must share one device";
    if (c->B != first->B) return "batched scenes must share the bounce count";
    if (c->cfg.restir_mode != first->cfg.restir_mode) return "batched scenes must share the ReSTIR mode";
    if (c->cache.on) return "the learned cache is not available in a batch";
    if (c->cfg.ray_sort_bits) return "ray sorting is not available in a batch";
    if (c->cfg.flags & VRGI_FLAG_KEEP_BOUNCES) return "VRGI_FLAG_KEEP_BOUNCES is not available in a batch";
    if (c->accum_on) return "FP64 accumulation is not available in a batch";
    return nullptr;
}

}  // namespace

vrgi_batch::~vrgi_batch() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fexec) (void)hipGraphExecDestroy(fexec);
    if (fgraph) (void)hipGraphDestroy(fgraph);
    for (int k = 0; k < RING; ++k) {
        if (ev_args[k]) (void)hipEventDestroy(ev_args[k]);
        if (h_args[k]) (void)hipHostFree(h_args[k]);
        if (h_fargs[k]) (void)hipHostFree(h_fargs[k]);
    }
    if (h_check) (void)hipHostFree(h_check);
    if (ev_check) (void)hipEventDestroy(ev_check);
    for (hipEvent_t e : {e_b0, e_b1, e_end})
        if (e) (void)hipEventDestroy(e);
    for (vrgi_ctx *c : sc) {
        // the canonical nodes lived in this batch's build arrays; the next render rebuilds
        c->ext_gnodes = c->ext_lnodes = nullptr;
        mark_geo_dirty(c);
        c->in_batch = false;
        c->batch_stream = nullptr;
    }
    if (stream) (void)hipStreamDestroy(stream);
}

extern "C" {

void vrgi_batch_config_default(vrgi_batch_config *cfg) {
    if (!cfg) return;
    cfg->spp = 1;
    cfg->rays_per_block = 0;
    cfg->pixels_per_block = 0;
    cfg->flags = 0;
}

int vrgi_batch_create(vrgi_ctx *const *scenes, uint32_t n, const vrgi_batch_config *cfg, vrgi_batch **out) {
    API_BEGIN
    if (!scenes || !cfg || !out || n == 0) return fail(VRGI_E_INVALID, "null argument or no scenes");
    *out = nullptr;
    if (n > 65536) return fail(VRGI_E_LIMIT, "more than 65536 scenes");
    if (cfg->spp < 1 || cfg->spp > 1024) return fail(VRGI_E_INVALID, "spp must be in [1, 1024]");
    auto tile_ok = [](uint32_t t) { return t == 0 || t == 64 || t == 128 || t == 256; };
    if (!tile_ok(cfg->rays_per_block) || !tile_ok(cfg->pixels_per_block))
        return fail(VRGI_E_INVALID, "rays_per_block and pixels_per_block must be 0, 64, 128 or 256");
    if (cfg->flags & ~VRGI_BATCH_NO_GRAPH) return fail(VRGI_E_INVALID, "unknown batch flag");
    for (uint32_t k = 0; k < n; ++k) {
        if (!scenes[k]) return fail(VRGI_E_INVALID, "null scene");
        if (const char *e = batch_scene_ok(scenes[k], scenes[0])) return fail(VRGI_E_INVALID, e);
        if (scenes[k]->in_batch) return fail(VRGI_E_STATE, "a scene is already in a batch");
        for (uint32_t j = 0; j < k; ++j)
            if (scenes[j] == scenes[k]) return fail(VRGI_E_INVALID, "a scene appears twice");
    }
    vrgi_batch *bt = new vrgi_batch();
    try {
        bt->device = scenes[0]->device;
        HIP_CHECK(hipSetDevice(bt->device));
        bt->sc.assign(scenes, scenes + n);
        bt->cfg = *cfg;
        bt->B = scenes[0]->B;
        HIP_CHECK(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking));
        for (hipEvent_t *e : {&bt->e_b0, &bt->e_b1, &bt->e_end}) HIP_CHECK(hipEventCreate(e));
        auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
        size_t o = 0;
        bt->off_sv = o; o = align(o + n * sizeof(SceneView));
        bt->off_tba = o; o = align(o + (size_t)n * bt->B * sizeof(BounceArgs));
        bt->off_nba = o; o = align(o + (size_t)n * bt->B * sizeof(BounceArgs));
        bt->off_fa = o; o = align(o + n * sizeof(FrameArgs));
        bt->off_da = o; o = align(o + n * sizeof(DiArgs));
        bt->off_pass = o; o = align(o + sizeof(BatchPass));
        bt->args_bytes = o;
        bt->d_args.reserve(o);
        for (int k = 0; k < vrgi_batch::RING; ++k) {
            HIP_CHECK(hipHostMalloc((void **)&bt->h_args[k], o, hipHostMallocDefault));
            std::memset(bt->h_args[k], 0, o);
            HIP_CHECK(hipEventCreateWithFlags(&bt->ev_args[k], hipEventDisableTiming));
        }
        bt->mean.resize(n);
        std::vector<float *> mp(n);
        for (uint32_t k = 0; k < n; ++k) {
            bt->mean[k].reserve((size_t)scenes[k]->P * 3);
            mp[k] = bt->mean[k].p;
        }
        bt->d_mean.reserve(n);
        HIP_CHECK(hipMemcpy(bt->d_mean.p, mp.data(), n * sizeof(float *), hipMemcpyHostToDevice));
        bt->wpb = cfg->rays_per_block / 64;
        bt->pb = cfg->pixels_per_block;
        bt->tuned = bt->wpb && bt->pb;
        bt->tune_step = bt->wpb ? 3 : 0;  // rays per block fixed: tune pixels only
        for (vrgi_ctx *c : bt->sc) {
            c->in_batch = true;
            c->batch_stream = bt->stream;
        }
    } catch (...) {
        delete bt;
        throw;
    }
    *out = bt;
    return VRGI_OK;
    API_END
}

int vrgi_batch_destroy(vrgi_batch *bt) {
    API_BEGIN
    delete bt;
    return VRGI_OK;
    API_END
}

int vrgi_batch_render(vrgi_batch *bt, const vrgi_camera *cams, const vrgi_prefs *prefs) {
    API_BEGIN
    if (!bt || !cams || !prefs) return fail(VRGI_E_INVALID, "null argument");
    if (prefs->nee_type > 2) return fail(VRGI_E_INVALID, "nee_type must be 0, 1 or 2");
    if (prefs->debug_view > 3) return fail(VRGI_E_INVALID, "debug_view must be in [0, 3] in a batch");
    if (prefs->restir_spatial_iterations > 20) return fail(VRGI_E_INVALID, "restir_spatial_iterations above 20");
    const uint32_t n = (uint32_t)bt->sc.size();
    for (vrgi_ctx *c : bt->sc)
        if (const char *e = batch_scene_ok(c, bt->sc[0])) return fail(VRGI_E_STATE, e);
    HIP_CHECK(hipSetDevice(bt->device));
    hipStream_t s = bt->stream;
    // instance tables are rewritten in place: the previous batched frame must have ended
    for (vrgi_ctx *c : bt->sc)
        if (c->inst_dirty) {
            HIP_CHECK(hipStreamSynchronize(s));
            break;
        }
    // the scenes' own queued work (uploads, earlier single renders) ends first
    for (vrgi_ctx *c : bt->sc) {
        if (c->vx.pending) {
            HIP_CHECK(hipStreamSynchronize(s));
            voxels_commit(c);
        }
        if (c->inst_dirty) {
            rebuild_instances(c);
            c->inst_dirty = false;
            mark_geo_dirty(c);
        }
        HIP_CHECK(hipStreamSynchronize(c->up_stream));
        HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    HIP_CHECK(hipEventRecord(bt->e_b0, s));
    std::vector<vrgi_ctx *> list;
    for (vrgi_ctx *c : bt->sc)
        if (c->geo_dirty || (c->cfg.flags & VRGI_FLAG_REBUILD_EVERY_FRAME)) list.push_back(c);
    // Rebuilds of unchanged geometry replay the plan of the last batched build; once the
    // tiles are tuned, such a frame (or one without builds) runs as one graph holding the
    // builds and every pass.
    const bool replay = !list.empty() && batch_replay_ok(bt, list);
    const bool frame_graph = (list.empty() || replay) && !(bt->cfg.flags & VRGI_BATCH_NO_GRAPH) && bt->tuned &&
                             bt->cfg.spp <= vrgi_batch::FRAME_GRAPH_SPP;
    if (replay) {
        batch_check_init(bt);
        batch_check_read(bt);
    }
    if (!list.empty()) batch_build(bt, list, replay, !frame_graph);  // frame graph: host state only here
    HIP_CHECK(hipEventRecord(bt->e_b1, s));
    bt->last_frame_graph = frame_graph && !list.empty();
    uint32_t gneed = 0, lneed = 0;
    for (vrgi_ctx *c : bt->sc) {
        gneed = std::max(gneed, geom_stack_need(c->G().geo_depth));
        lneed = std::max(lneed, 3 * ((c->G().lgeo_depth + 1) / 2));
    }
    // tiles of this batched frame
    uint32_t wpb = bt->wpb, pb = bt->pb;
    const bool tuning = !bt->tuned;
    if (tuning) {
        const int st = bt->tune_step;
        wpb = st < 3 ? BATCH_TILES[st] / 64 : bt->wpb;
        pb = st < 3 ? (bt->cfg.pixels_per_block ? bt->cfg.pixels_per_block : 256u) : BATCH_TILES[st - 3];
    }
    if (gneed + 2 > 64 || lneed + 2 > 64) wpb = 1;  // a stack above 64 entries runs one wave per block
    batch_tables(bt, wpb, pb);
    BatchLaunch L{};
    uint8_t *dA = bt->d_args.p;
    L.svs = (const SceneView *)(dA + bt->off_sv);
    L.tba = (const BounceArgs *)(dA + bt->off_tba);
    L.nba = (const BounceArgs *)(dA + bt->off_nba);
    L.fas = (const FrameArgs *)(dA + bt->off_fa);
    L.das = (const DiArgs *)(dA + bt->off_da);
    L.pass = (const BatchPass *)(dA + bt->off_pass);
    L.blk_ray = bt->d_blk.p;
    L.blk_px = bt->d_blk.p + n + 1;
    L.mean = bt->d_mean.p;
    L.n = n; L.B = bt->B;
    L.ray_blocks = bt->ray_blocks; L.px_blocks = bt->px_blocks;
    L.gneed = gneed; L.lneed = lneed;
    L.wpb = wpb; L.pb = pb;
    L.nee = prefs->nee_type != 0;
    L.di = bt->sc[0]->cfg.restir_mode == 1;
    auto cls = [](uint32_t need) { return need + 2 <= 32 ? 32u : need + 2 <= 48 ? 48u : need + 2 <= 64 ? 64u : 160u; };
    const bool use_graph = !(bt->cfg.flags & VRGI_BATCH_NO_GRAPH) && !tuning;
    std::vector<uint32_t> key = {cls(gneed), cls(lneed), wpb, pb, L.ray_blocks, L.px_blocks, (uint32_t)L.nee,
                                 (uint32_t)L.di, bt->B};
    if (frame_graph) {
        const uint32_t spp = bt->cfg.spp;
        const size_t fbytes = (size_t)spp * bt->args_bytes;
        if (bt->f_args.cap < fbytes) {  // new buffers: any graph holding the old ones goes
            HIP_CHECK(hipStreamSynchronize(s));
            bt->f_args.reserve(fbytes);
            for (int k = 0; k < vrgi_batch::RING; ++k) {
                if (bt->h_fargs[k]) HIP_CHECK(hipHostFree(bt->h_fargs[k]));
                HIP_CHECK(hipHostMalloc((void **)&bt->h_fargs[k], fbytes, hipHostMallocDefault));
                std::memset(bt->h_fargs[k], 0, fbytes);
            }
            bt->fgraph_key.clear();
        }
        const uint32_t r = bt->ring;
        bt->ring = (r + 1) % vrgi_batch::RING;
        if (bt->ev_used[r]) HIP_CHECK(hipEventSynchronize(bt->ev_args[r]));
        uint8_t *h = bt->h_fargs[r];
        for (uint32_t pass = 0; pass < spp; ++pass) {
            uint8_t *hp = h + (size_t)pass * bt->args_bytes;
            for (uint32_t k = 0; k < n; ++k) batch_fill(bt, hp, k, cams[k], *prefs);
            BatchPass bp{};
            bp.first = pass == 0;
            bp.last = pass + 1 == spp;
            bp.inv_spp = 1.0f / (float)spp;
            *(BatchPass *)(hp + bt->off_pass) = bp;
            for (vrgi_ctx *c : bt->sc) {
                c->have_frame = true;
                ++c->frame_count;
            }
        }
        HIP_CHECK(hipMemcpyAsync(bt->f_args.p, h, fbytes, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipEventRecord(bt->ev_args[r], s));
        bt->ev_used[r] = true;
        std::vector<uint32_t> fkey = key;
        fkey.insert(fkey.end(), {spp, (uint32_t)list.size(), replay ? bt->plan_gen : 0u});
        if (!bt->fexec || bt->fgraph_key != fkey) {
            if (bt->fexec) HIP_CHECK(hipGraphExecDestroy(bt->fexec));
            if (bt->fgraph) HIP_CHECK(hipGraphDestroy(bt->fgraph));
            bt->fexec = nullptr;
            bt->fgraph = nullptr;
            HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            if (replay) batch_build(bt, list, true, true);
            for (uint32_t pass = 0; pass < spp; ++pass) {
                BatchLaunch Lp = L;
                uint8_t *dp = bt->f_args.p + (size_t)pass * bt->args_bytes;
                Lp.svs = (const SceneView *)(dp + bt->off_sv);
                Lp.tba = (const BounceArgs *)(dp + bt->off_tba);
                Lp.nba = (const BounceArgs *)(dp + bt->off_nba);
                Lp.fas = (const FrameArgs *)(dp + bt->off_fa);
                Lp.das = (const DiArgs *)(dp + bt->off_da);
                Lp.pass = (const BatchPass *)(dp + bt->off_pass);
                launch_batch_pass(s, Lp);
            }
            HIP_CHECK(hipStreamEndCapture(s, &bt->fgraph));
            HIP_CHECK(hipGraphInstantiate(&bt->fexec, bt->fgraph, nullptr, nullptr, 0));
            bt->fgraph_key = fkey;
        }
        HIP_CHECK(hipGraphLaunch(bt->fexec, s));
        if (replay) {
            HIP_CHECK(hipMemcpyAsync(bt->h_check, bt->d_check.p, 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipEventRecord(bt->ev_check, s));
            bt->check_pending = true;
        }
    }
    for (uint32_t pass = 0; pass < (frame_graph ? 0u : bt->cfg.spp); ++pass) {
        const uint32_t r = bt->ring;
        bt->ring = (r + 1) % vrgi_batch::RING;
        if (bt->ev_used[r]) HIP_CHECK(hipEventSynchronize(bt->ev_args[r]));
        uint8_t *h = bt->h_args[r];
        for (uint32_t k = 0; k < n; ++k) batch_fill(bt, h, k, cams[k], *prefs);
        BatchPass bp{};
        bp.first = pass == 0;
        bp.last = pass + 1 == bt->cfg.spp;
        bp.inv_spp = 1.0f / (float)bt->cfg.spp;
        *(BatchPass *)(h + bt->off_pass) = bp;
        HIP_CHECK(hipMemcpyAsync(dA, h, bt->args_bytes, hipMemcpyHostToDevice, s));
        HIP_CHECK(hipEventRecord(bt->ev_args[r], s));
        bt->ev_used[r] = true;
        if (use_graph) {
            if (!bt->exec || bt->graph_key != key) {
                if (bt->exec) HIP_CHECK(hipGraphExecDestroy(bt->exec));
                if (bt->graph) HIP_CHECK(hipGraphDestroy(bt->graph));
                bt->exec = nullptr;
                bt->graph = nullptr;
                HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                launch_batch_pass(s, L);
                HIP_CHECK(hipStreamEndCapture(s, &bt->graph));
                HIP_CHECK(hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0));
                bt->graph_key = key;
            }
            HIP_CHECK(hipGraphLaunch(bt->exec, s));
        } else {
            launch_batch_pass(s, L);
        }
        for (vrgi_ctx *c : bt->sc) {
            c->have_frame = true;
            ++c->frame_count;
        }
    }
    if (replay && !frame_graph) {
        HIP_CHECK(hipMemcpyAsync(bt->h_check, bt->d_check.p, 4, hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(bt->ev_check, s));
        bt->check_pending = true;
    }
    HIP_CHECK(hipEventRecord(bt->e_end, s));
    bt->have_times = true;
    if (tuning) {  // steps 0-2: 64 / 128 / 256 rays per block; steps 3-5: pixels per block
        HIP_CHECK(hipEventSynchronize(bt->e_end));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, bt->e_b1, bt->e_end));
        const int st = bt->tune_step;
        bt->tune_ms[st] = ms;
        if (st == 2) {
            int best = 0;
            for (int k = 1; k < 3; ++k)
                if (bt->tune_ms[k] < bt->tune_ms[best]) best = k;
            bt->wpb = BATCH_TILES[best] / 64;
            if (bt->cfg.pixels_per_block) bt->tuned = true;
        }
        if (st == 5) {
            int best = 3;
            for (int k = 4; k < 6; ++k)
                if (bt->tune_ms[k] < bt->tune_ms[best]) best = k;
            bt->pb = BATCH_TILES[best - 3];
            bt->tuned = true;
        }
        bt->tune_step = st + 1;
    }
    return VRGI_OK;
    API_END
}

int vrgi_batch_sync(vrgi_batch *bt) {
    API_BEGIN
    if (!bt) return fail(VRGI_E_INVALID, "null batch");
    HIP_CHECK(hipSetDevice(bt->device));
    HIP_CHECK(hipStreamSynchronize(bt->stream));
    return VRGI_OK;
    API_END
}

int vrgi_batch_read_mean(vrgi_batch *bt, uint32_t scene, float *rgb) {
    API_BEGIN
    if (!bt || !rgb) return fail(VRGI_E_INVALID, "null argument");
    if (scene >= bt->sc.size()) return fail(VRGI_E_INVALID, "scene index out of range");
    HIP_CHECK(hipSetDevice(bt->device));
    HIP_CHECK(hipMemcpyAsync(rgb, bt->mean[scene].p, (size_t)bt->sc[scene]->P * 12, hipMemcpyDeviceToHost, bt->stream));
    HIP_CHECK(hipStreamSynchronize(bt->stream));
    return VRGI_OK;
    API_END
}

int vrgi_batch_info_get(vrgi_batch *bt, vrgi_batch_info *info) {
    API_BEGIN
    if (!bt || !info) return fail(VRGI_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(bt->device));
    *info = vrgi_batch_info{};
    info->n_scenes = (uint32_t)bt->sc.size();
    info->rays_per_block = bt->wpb * 64;
    info->pixels_per_block = bt->pb;
    info->tuned = bt->tuned ? 1u : 0u;
    info->graph = bt->fexec && !bt->fgraph_key.empty() ? 2u : bt->exec ? 1u : 0u;
    for (vrgi_ctx *c : bt->sc) {
        info->triangles += c->n_tris;
        info->light_triangles += c->n_ltri;
        info->pixels += c->P;
    }
    HIP_CHECK(hipStreamSynchronize(bt->stream));
    if (bt->have_times) {
        HIP_CHECK(hipEventElapsedTime(&info->build_ms, bt->e_b0, bt->e_b1));
        HIP_CHECK(hipEventElapsedTime(&info->passes_ms, bt->e_b1, bt->e_end));
        if (bt->last_frame_graph) {  // the builds ran inside the frame's graph
            info->build_ms = -1.0f;
            HIP_CHECK(hipEventElapsedTime(&info->passes_ms, bt->e_b0, bt->e_end));
        }
    }
    uint64_t rays = 0;
    for (vrgi_ctx *c : bt->sc) {
        Counters k{};
        HIP_CHECK(hipMemcpy(&k, c->d_counters.p, sizeof(k), hipMemcpyDeviceToHost));
        rays += k.rays_closest;
    }
    info->rays = rays - bt->rays_seen;
    bt->rays_seen = rays;
    return VRGI_OK;
    API_END
}

}  // extern "C"
