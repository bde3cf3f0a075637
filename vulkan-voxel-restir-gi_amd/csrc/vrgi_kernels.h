// vrgi_kernels.h — kernel argument blocks, traversal primitives and launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "devcommon.h"

namespace vrgi {

#define VRGI_MAX_STACK 160  // traversal stack entries per lane (LDS); deeper trees are rejected
#ifndef VRGI_PEND_LDS
#define VRGI_PEND_LDS 1  // closest hit: pending triangles in the lane's top stack entries
#endif
#define VRGI_SMALL_QUEUE_GRID 2048  // blocks of a striding ray launch over a queue expected to be short
#ifndef VRGI_CACHE_SPLIT_B0
#define VRGI_CACHE_SPLIT_B0 1  // cache frames querying at vertex 0 split bounce 0 (BounceArgs::b0_part)
#endif
#ifndef VRGI_NEE_STREAMS
#define VRGI_NEE_STREAMS 4  // streams the NEE pdf launches of a frame rotate over
#endif
#define VRGI_NEE_RECS 8     // crossed light triangles a ray keeps for the NEE walk pass
#define VRGI_BOX_PAD 0.001f  // slack added to every box of the traversal copies (see box_hit)

// ---------------------------------------------------------------- argument blocks
struct SceneView {
    const WideNode *gwide;     // geometry BVH, traversal copy (one entry per internal node)
    const Wide4 *gwide4;       // the same, 4-ary
    const Tri *tri;            // world triangles by global id
    uint32_t n_tris;
    uint32_t root_prim;        // primitive of the root when the tree is a single leaf
    uint32_t depth;
    const uint32_t *tri_inst;  // instance of each global triangle
    const DevInst *inst;
    uint32_t n_inst;
    const vrgi_bvh_node *tl;   // light top-level BVH
    const WideNode *lwide;     // light-triangle BVH, traversal copy
    const Wide4 *lwide4;       // the same, 4-ary
    const Tri *ltri;           // world light triangles by global light id
    uint32_t n_ltri;
    uint32_t lroot_prim;
    uint32_t ldepth;
    const uint32_t *ltri_inst;  // instance of each light triangle
    const uint32_t *ltri_node;  // LightVertex3D.bvh_node_idx of each light triangle
    const LPair *tlpairs;       // child records of the light top-level BVH
    // ancestors of each light triangle, leaf to root (BL levels, then TL levels):
    // parent node index | 0x40000000 for a TL level | 0x80000000 when the child is the
    // right one; VRGI_LANC_MAX entries per light, lanc_len[lid] used (0xFFFFFFFF: more
    // than VRGI_LANC_MAX, walk the parent links instead)
    const uint32_t *lanc;
    const uint32_t *lanc_len;
    const vrgi_material *mats;
    uint32_t n_mats;
    // texture atlas (null: the constant materials): linear RGBA texels [slot][3][h][w] and
    // the uv of every triangle's three vertices
    const float4 *tex;
    const float2 *tri_uv;
    uint32_t tex_w, tex_h;
    // cells of a grid over the light triangles, set where some light triangle's box comes
    // within LGRID_MARGIN of the cell (null: no grid); cell of p = floor((p - lg_lo) * lg_inv)
    const uint8_t *lgrid;
    float lg_lo[3], lg_inv;
    uint32_t lg_dim[3];
};

struct BounceArgs {
    uint32_t bounce, B, W, H, P;
    uint32_t inv_seed;  // frame * B + bounce (interactive_rendering.rs:1189)
    uint32_t nee_type;
    uint32_t rr_start;
    float rr_min;
    vrgi_camera cam;
    const uint32_t *queue_in;
    uint32_t *queue_out;
    uint32_t *qcount;  // live rays per bounce
    // per bounce b, [B][P][3] / [B][P]: outputs of bounce b (origin/dir = ray b + 1)
    float *origin, *dir, *normal, *emis, *refl, *mis, *pdf, *nee, *hit_t;
    uint32_t *hit, *path_len;
    unsigned long long *counters;  // [0] closest-hit rays, [1] light rays, [2] frames
    uint32_t *touch_nodes, *touch_tris;  // non-null: counting launch (measurement only)
    // learned radiance cache (cache_on): a pixel traces a full training path when
    // u(finalize(combine(train_seed, px))) < train_frac, else it stops at bounce 1 and
    // queues a cache query record (16 floats) of that vertex
    uint32_t cache_on, train_seed, cache_vertex;  // cache_vertex: bounce of the query
    float train_frac;
    float *d0;              // [P][3] camera ray directions (incoming direction at vertex 0)
    const uint8_t *fallback;  // per tile: trace full paths (null: none)
    uint32_t tile, tiles_x;
    float *albedo;          // [B][P][3] material reflectivity at each hit
    float *qrec;            // [P][16] query records
    uint32_t *qrec_count;
    // bounce 0 of a cache frame in two launches, so that the training paths' bounces run
    // beside the other pixels' first hits: 0 every pixel; 1 the full-path pixels listed in
    // queue_in (qcount[0] of them, launch_full_list); 2 every other pixel
    uint32_t b0_part;
    // NEE pdf in three passes (nee_rec set): per queue slot i, the crossed light
    // triangles (at most VRGI_NEE_RECS: lid plane [k][P], barycentric pair plane [k][P]),
    // their count (VRGI_NEE_RECS + 1: more, the ray is on the overflow list)
    uint32_t *nee_cnt, *nee_lid;
    float2 *nee_uv;
    uint32_t *nee_over, *nee_over_count;
    // lazy NEE pdf (launch_nee_pdf_lazy): [P][3] outgoing radiance of the next vertex,
    // L_{b+1}, left by launch_resolve_step(b + 1)
    const float *lrun;
};

struct FlattenArgs {
    const DevInst *inst;
    uint32_t n_inst;
    const vrgi_vertex3d *const *vtx;        // per instance: object vertices
    const vrgi_light_vertex3d *const *lvtx;  // per instance: object light vertices
    const uint32_t *light_off;              // per instance: first global light id
    uint32_t n_tris;
    Tri *tri;
    uint32_t *tri_inst;
    uint32_t *light_node;
    float *prim9;         // builder planes of prim_stride floats; this list starts at prim_off
    float *power;
    uint32_t *ids;
    uint32_t prim_stride, prim_off;
    float2 *tri_uv;  // non-null: the uv of each triangle's 3 vertices (texture atlas on)
};

struct FrameArgs {
    uint32_t W, H, P, B, frame, slot;
    uint32_t tseed, sseed;
    uint32_t spatial_iters;
    float radius;
    uint32_t debug_view;  // 0 path tracer, 1 ReSTIR, 2 debug, 3 ReSTIR tonemapped, 4 accumulated path tracer
    float *view_acc;      // view 4: [P][3] running sum of the path-traced radiance
    uint32_t view_n;      // view 4: frames in the sum, this one included (1: restart)
    const float *origin, *normal, *emis, *refl, *mis, *pdf, *nee;
    const uint32_t *path_len;
    float *radiance;   // [P][3]
    float *temporal;   // [P][20], slot frame % 3
    float *spatial;    // [P][20]
    float *restir;     // [P][3]
    float *debug;      // [P][3]
    uint8_t *bgra8;    // [P][4]
    // learned radiance cache (see BounceArgs)
    uint32_t cache_on, train_seed, cache_vertex;
    float train_frac;
    const uint8_t *fallback;
    uint32_t tile, tiles_x;
    const float *dir, *d0, *albedo, *crad;  // crad: [P][3] cache radiance of queried pixels
    const uint32_t *hit;
    float *trec;                       // training records [tcap][16]
    uint32_t *trec_count;
    uint32_t tcap;
    // k_accum_temporal over part of the pixels: 0 all; 1 all but the full-path pixels;
    // 2 the full-path pixels of full_list[0 .. *full_count) (see BounceArgs::b0_part)
    uint32_t part;
    const uint32_t *full_list, *full_count;
};

// ReSTIR DI (benchmark mode, restir_mode = 1)
struct DiArgs {
    uint32_t W, H, P, frame, candidates, spatial_iters;
    float radius;
    uint32_t dseed;
    const float *origin, *normal, *emis;  // bounce-0 slices
    const uint32_t *hit;                  // bounce-0 hit ids
    float *T;                             // [P][8] reservoirs after the temporal merge
    const float *hist_prev;               // [P][8] previous frame's final reservoirs
    float *hist_cur;                      // [P][8]
    float *restir;                        // [P][3] final radiance
    uint32_t *di_light;                   // [P]
    float *di_weights;                    // [P][4] w_sum, W, M, p_hat
};

// ---------------------------------------------------------------- primitives
// builder inputs of one triangle: Aabb::from_points (aabb.rs:10-22) and the centroid
// ((v0 + v1) + v2) / 3 (build.rs:267-270); planes of n floats
__device__ __forceinline__ void write_prim9(float *prim9, size_t n, uint32_t g, const f3 w[3]) {
    float mn[3] = {w[0].x, w[0].y, w[0].z}, mx[3] = {w[0].x, w[0].y, w[0].z};
    for (int k = 0; k < 3; ++k) {
        mn[0] = gmin(mn[0], w[k].x); mn[1] = gmin(mn[1], w[k].y); mn[2] = gmin(mn[2], w[k].z);
        mx[0] = gmax(mx[0], w[k].x); mx[1] = gmax(mx[1], w[k].y); mx[2] = gmax(mx[2], w[k].z);
    }
    f3 c = ((w[0] + w[1]) + w[2]) / 3.0f;
    prim9[0 * (size_t)n + g] = mn[0]; prim9[1 * (size_t)n + g] = mn[1]; prim9[2 * (size_t)n + g] = mn[2];
    prim9[3 * (size_t)n + g] = mx[0]; prim9[4 * (size_t)n + g] = mx[1]; prim9[5 * (size_t)n + g] = mx[2];
    prim9[6 * (size_t)n + g] = c.x; prim9[7 * (size_t)n + g] = c.y; prim9[8 * (size_t)n + g] = c.z;
}

// Woop, Benthin, Wald: per-ray shear constants
struct RayPre {
    f3 o, d, inv, oi;  // oi = o * inv (slab tests as one fma per plane)
    int kx, ky, kz;
    float sx, sy, sz;
};
__device__ __forceinline__ RayPre ray_pre(f3 o, f3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    r.kx = r.kz == 2 ? 0 : r.kz + 1;
    r.ky = r.kx == 2 ? 0 : r.kx + 1;
    float dz = comp3(d, r.kz);
    if (dz < 0.0f) { int t = r.kx; r.kx = r.ky; r.ky = t; }
    r.sx = comp3(d, r.kx) / dz;
    r.sy = comp3(d, r.ky) / dz;
    r.sz = 1.0f / dz;
    r.inv = mk(d.x != 0.0f ? 1.0f / d.x : 1e30f, d.y != 0.0f ? 1.0f / d.y : 1e30f,
               d.z != 0.0f ? 1.0f / d.z : 1e30f);
    r.oi = mk(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    return r;
}

// watertight ray/triangle test; on a hit in [tmin, tmax] returns t and the weights
// (u, v) of vertices 1 and 2
__device__ __forceinline__ bool tri_hit(const RayPre &r, f3 v0, f3 v1, f3 v2, float tmin, float tmax,
                                        float *t_out, float *u_out, float *v_out) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = comp3(A, r.kz), Bkz = comp3(B, r.kz), Ckz = comp3(C, r.kz);
    float Ax = comp3(A, r.kx) - r.sx * Akz, Ay = comp3(A, r.ky) - r.sy * Akz;
    float Bx = comp3(B, r.kx) - r.sx * Bkz, By = comp3(B, r.ky) - r.sy * Bkz;
    float Cx = comp3(C, r.kx) - r.sx * Ckz, Cy = comp3(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {  // edge case: exact products in binary64
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

// conservative slab test (Ize 2013: far distance scaled by 1 + 2 gamma(3)); boxes of
// the traversal copies carry VRGI_BOX_PAD of slack so that a triangle hit reported at
// rounding distance outside its exact box is never culled. Each plane distance is one
// fma with o * inv rounded once: its absolute error, |o| * 2^-23 / |d_axis| at most, is
// far below the slack's VRGI_BOX_PAD / |d_axis| for scene coordinates below 10^4.
__device__ __forceinline__ bool box_hit_t(const RayPre &r, float mnx, float mny, float mnz, float mxx,
                                          float mxy, float mxz, float tmin, float tmax, float *tn) {
    float tx0 = __builtin_fmaf(mnx, r.inv.x, -r.oi.x), tx1 = __builtin_fmaf(mxx, r.inv.x, -r.oi.x);
    float ty0 = __builtin_fmaf(mny, r.inv.y, -r.oi.y), ty1 = __builtin_fmaf(mxy, r.inv.y, -r.oi.y);
    float tz0 = __builtin_fmaf(mnz, r.inv.z, -r.oi.z), tz1 = __builtin_fmaf(mxz, r.inv.z, -r.oi.z);
    float t0 = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float t1 = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    *tn = t0;
    return t0 <= t1 * 1.00000024f;
}
__device__ __forceinline__ bool box_hit(const RayPre &r, float mnx, float mny, float mnz, float mxx,
                                        float mxy, float mxz, float tmin, float tmax) {
    float tn;
    return box_hit_t(r, mnx, mny, mnz, mxx, mxy, mxz, tmin, tmax, &tn);
}

struct Hit {
    uint32_t prim;
    float t, u, v;
};

// closest hit over the geometry BVH; equal t resolves to the lower primitive id.
// stack: STACK * 64 LDS words of this wave, entry k of lane l at [k * 64 + l].
// COUNT (measurement only): marks every visited internal node / tested triangle in the
// bitmaps touch_nodes / touch_tris.
// The triangle tests of a wave run together: a lane that reaches a leaf keeps its
// triangle (at most 4 pending) and goes on with its internal nodes while other lanes
// still have none; once every lane that is still walking has a pending triangle, all
// lanes test theirs at once. The tests then run with most lanes active instead of one
// or two lanes per step; the result does not depend on the order of the tests.
template <int STACK, bool COUNT = false>
__device__ __forceinline__ Hit closest_hit2(const SceneView &sv, const RayPre &r, float tmin,
                                            float tmax, uint32_t *stack,
                                            uint32_t *touch_nodes = nullptr,
                                            uint32_t *touch_tris = nullptr) {
    Hit h;
    h.prim = VRGI_INVALID;
    h.t = 0.0f; h.u = 0.0f; h.v = 0.0f;
    if (sv.n_tris == 0) return h;
    float best = tmax;
    uint32_t *stk = stack + (threadIdx.x & 63);
    auto test = [&](uint32_t prim) {
        if (COUNT) atomicOr(touch_tris + (prim >> 5), 1u << (prim & 31));
        Tri t = sv.tri[prim];
        float tt, u, v;
        if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z), mk(t.v2.x, t.v2.y, t.v2.z),
                    tmin, best, &tt, &u, &v)) {
            if (tt < best || (tt == best && prim < h.prim)) {
                best = tt;
                h.prim = prim; h.t = tt; h.u = u; h.v = v;
            }
        }
    };
    if (sv.n_tris == 1) {
        test(sv.root_prim);
        return h;
    }
    uint32_t ref = 0;
    int sp = 0;
    bool walk = true;  // ref is an internal node still to visit
    int np = 0;
#if VRGI_PEND_LDS
    // pending triangles in the lane's top stack entries (STACK - 1 down), below which the
    // node stack (at most depth - 1 entries) never reaches
    uint32_t *const pend = stk + (STACK - 1) * 64;
    auto keep = [&](uint32_t prim) { pend[-64 * np++] = prim; };
#else
    uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    auto keep = [&](uint32_t prim) {  // a shift register: no indexed private array
        p3 = p2; p2 = p1; p1 = p0; p0 = prim;
        ++np;
    };
#endif
    for (;;) {
        // wave-uniform choice: test when no walking lane is without a pending triangle
        if (!__any(walk && np == 0)) {
            if (!__any(np > 0)) break;
#pragma unroll 1
            for (; np > 0; --np) {
#if VRGI_PEND_LDS
                test(pend[-64 * (np - 1)]);
#else
                test(p0);
                p0 = p1; p1 = p2; p2 = p3;
#endif
            }
            continue;
        }
        if (!walk || np > 2) continue;  // this lane waits (no node, or no room for two more)
        if (COUNT) atomicOr(touch_nodes + (ref >> 5), 1u << (ref & 31));
        WideNode wn = sv.gwide[ref];
        float tn0, tn1;
        bool h0 = box_hit_t(r, wn.a.x, wn.a.y, wn.a.z, wn.a.w, wn.b.x, wn.b.y, tmin, best, &tn0);
        bool h1 = box_hit_t(r, wn.b.z, wn.b.w, wn.c.x, wn.c.y, wn.c.z, wn.c.w, tmin, best, &tn1);
        uint32_t r0 = wn.r.x, r1 = wn.r.y;
        if (h0 && (r0 & 0x80000000u)) { keep(r0 & 0x7FFFFFFFu); h0 = false; }
        if (h1 && (r1 & 0x80000000u)) { keep(r1 & 0x7FFFFFFFu); h1 = false; }
        if (h0 && h1) {
            bool first0 = tn0 <= tn1;
            stk[(sp++) * 64] = first0 ? r1 : r0;
            ref = first0 ? r0 : r1;
        } else if (h0) {
            ref = r0;
        } else if (h1) {
            ref = r1;
        } else if (sp == 0) {
            walk = false;
        } else {
            ref = stk[(--sp) * 64];
        }
    }
    return h;
}

// the same over the 4-ary copy: a node tests four descendants; leaves are tested at
// once, hit internal descendants are visited nearest first (the result does not depend
// on the order: equal t resolves to the lower primitive id)
template <int STACK, bool COUNT = false>
__device__ __forceinline__ Hit closest_hit4(const SceneView &sv, const RayPre &r, float tmin, float tmax,
                                            uint32_t *stack, uint32_t *touch_nodes = nullptr,
                                            uint32_t *touch_tris = nullptr) {
    Hit h;
    h.prim = VRGI_INVALID;
    h.t = 0.0f; h.u = 0.0f; h.v = 0.0f;
    if (sv.n_tris == 0) return h;
    float best = tmax;
    uint32_t *stk = stack + (threadIdx.x & 63);
    auto test = [&](uint32_t prim) {
        if (COUNT) atomicOr(touch_tris + (prim >> 5), 1u << (prim & 31));
        Tri t = sv.tri[prim];
        float tt, u, v;
        if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z), mk(t.v2.x, t.v2.y, t.v2.z),
                    tmin, best, &tt, &u, &v)) {
            if (tt < best || (tt == best && prim < h.prim)) {
                best = tt;
                h.prim = prim; h.t = tt; h.u = u; h.v = v;
            }
        }
    };
    if (sv.n_tris == 1) {
        test(sv.root_prim);
        return h;
    }
    uint32_t ref = 0;
    int sp = 0;
    for (;;) {
        if (COUNT) atomicOr(touch_nodes + (ref >> 5), 1u << (ref & 31));
        const Wide4 *w = sv.gwide4 + ref;
        const float4 lx = w->lox, ly = w->loy, lz = w->loz, hx = w->hix, hy = w->hiy, hz = w->hiz;
        const uint4 rf = w->ref;
        float t0, t1, t2, t3;
        bool b0 = rf.x != VRGI_INVALID && box_hit_t(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, tmin, best, &t0);
        bool b1 = rf.y != VRGI_INVALID && box_hit_t(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, tmin, best, &t1);
        bool b2 = rf.z != VRGI_INVALID && box_hit_t(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, tmin, best, &t2);
        bool b3 = rf.w != VRGI_INVALID && box_hit_t(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, tmin, best, &t3);
        if (b0 && (rf.x & 0x80000000u)) { test(rf.x & 0x7FFFFFFFu); b0 = false; }
        if (b1 && (rf.y & 0x80000000u)) { test(rf.y & 0x7FFFFFFFu); b1 = false; }
        if (b2 && (rf.z & 0x80000000u)) { test(rf.z & 0x7FFFFFFFu); b2 = false; }
        if (b3 && (rf.w & 0x80000000u)) { test(rf.w & 0x7FFFFFFFu); b3 = false; }
        // internal hits sorted by entry distance (absent ones last)
        const float inf = __int_as_float(0x7f800000);
        float k0 = b0 ? t0 : inf, k1 = b1 ? t1 : inf, k2 = b2 ? t2 : inf, k3 = b3 ? t3 : inf;
        uint32_t r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
        auto cs = [](float &ka, uint32_t &ra, float &kb, uint32_t &rb) {
            if (kb < ka) { float tk = ka; ka = kb; kb = tk; uint32_t tr = ra; ra = rb; rb = tr; }
        };
        cs(k0, r0, k1, r1); cs(k2, r2, k3, r3); cs(k0, r0, k2, r2); cs(k1, r1, k3, r3); cs(k1, r1, k2, r2);
        const int n = (int)b0 + (int)b1 + (int)b2 + (int)b3;
        if (n == 0) {
            if (sp == 0) break;
            ref = stk[(--sp) * 64];
            continue;
        }
        if (n >= 4) stk[(sp++) * 64] = r3;
        if (n >= 3) stk[(sp++) * 64] = r2;
        if (n >= 2) stk[(sp++) * 64] = r1;
        ref = r0;
    }
    return h;
}

// closest hit over the geometry tree's traversal copy in use
#ifndef VRGI_GEOM4
#define VRGI_GEOM4 0
#endif
template <int STACK, bool COUNT = false>
__device__ __forceinline__ Hit closest_hit(const SceneView &sv, const RayPre &r, float tmin, float tmax,
                                           uint32_t *stack, uint32_t *touch_nodes = nullptr,
                                           uint32_t *touch_tris = nullptr) {
#if VRGI_GEOM4
    return closest_hit4<STACK, COUNT>(sv, r, tmin, tmax, stack, touch_nodes, touch_tris);
#else
    return closest_hit2<STACK, COUNT>(sv, r, tmin, tmax, stack, touch_nodes, touch_tris);
#endif
}
// stack entries a closest-hit walk needs for a tree of binary depth d
// LDS stack entries a closest-hit walk of a tree of depth d needs: the binary walk holds
// at most d - 1 nodes and, with VRGI_PEND_LDS, up to 4 pending triangles at the top
__host__ __device__ inline uint32_t geom_stack_need(uint32_t d) {
    return VRGI_GEOM4 ? 3 * ((d + 1) / 2) : (VRGI_PEND_LDS ? d + 3 : d);
}
// bytes of one traversal node of the geometry tree (touch counting)
#define VRGI_GEOM_NODE_BYTES (VRGI_GEOM4 ? 128ull : 64ull)

// ---------------------------------------------------------------- launchers
void launch_flatten(hipStream_t s, const FlattenArgs &f);
void launch_flatten_lights(hipStream_t s, const FlattenArgs &f);
void launch_wide(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, WideNode *wide);
void launch_wide4(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, Wide4 *wide4);
// the light grid of SceneView (cells of side 1 / inv, dim[0] x dim[1] x dim[2], x fastest):
// cleared, then every cell within `margin` of a light triangle's box set
void launch_light_grid(hipStream_t s, const Tri *ltri, uint32_t nl, uint8_t *grid, const float lo[3], float inv,
                       const uint32_t dim[3], float margin);
void launch_trace_closest(hipStream_t s, const SceneView &sv, const float *o, const float *d,
                          uint32_t n, float tmin, float tmax, uint32_t *hit, float *t, float *uv);
// grid_cap > 0: a grid of at most grid_cap blocks striding over the queue (for queues
// expected to be far shorter than max_rays). NEE pdf with ba.nee_cnt set: enumeration,
// walk and overflow passes instead of one kernel.
// the pixels that trace full paths in a cache frame (training pixels, fallback tiles), in
// list[0 .. *count): the full-path pixels of bounce 0 with b0_part = 1
void launch_full_list(hipStream_t s, const BounceArgs &ba, uint32_t *list, uint32_t *count);
void launch_trace_shade(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays,
                        uint32_t grid_cap = 0);
void launch_nee_pdf(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays,
                    uint32_t grid_cap = 0);
// The NEE pdf of bounce b computed only where its value can reach the frame's outputs:
// rays whose next vertex returns radiance (ba.lrun non-zero), or whose pdf might not be
// a finite non-negative number. The others get 0: their q differs from the reference's,
// but L_b = E_b + ((R_b * 0) * (bsdf_pdf / q)) does not. Needs launch_resolve_step(b + 1)
// first. Pass 1 writes the 0s and lists the other rays: those whose pdf matters (list),
// and those for which that depends on a light triangle crossed near the origin (mlist,
// decided by a walk over the light triangles in pass 1b, which lists them or writes 0);
// pass 2 computes the listed rays' pdf at full lane occupancy. list, mlist: max_rays
// entries each; *list_count, *mcount zeroed before.
#define VRGI_NEE_LIST_GRID 4096  // blocks of the passes striding over a list
void launch_nee_pdf_lazy(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays, uint32_t *list,
                         uint32_t *list_count, uint32_t *mlist, uint32_t *mcount);
// One bounce k of the outgoing-radiance recurrence (outgoing_radiance.rs:58-94) for every
// pixel whose path reached vertex k: lrun = E_k + ((R_k * lrun) * (bsdf_pdf_k / q_k)),
// lrun taken as 0 past the path's end; the arithmetic of the full resolve.
void launch_resolve_step(hipStream_t s, const FrameArgs &fa, uint32_t k, float *lrun);
void launch_accum_temporal(hipStream_t s, const FrameArgs &fa);
void launch_spatial_finalize(hipStream_t s, const FrameArgs &fa);
void launch_postprocess(hipStream_t s, const FrameArgs &fa);
void launch_accumulate(hipStream_t s, const float *rad, double *acc, size_t n);
void launch_prims_from_tris(hipStream_t s, const float *tri9, uint32_t n, float *prim9);
void launch_tl_inputs(hipStream_t s, const float *aabb6, const float *xf12, uint32_t n, float *prim9);
void launch_tl_idx(hipStream_t s, const vrgi_bvh_node *tl, uint32_t n_nodes, DevInst *inst);
#define VRGI_LANC_MAX 64
void launch_light_pairs(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, LPair *out);
void launch_light_ancestry(hipStream_t s, const uint32_t *ltri_inst, const uint32_t *ltri_node,
                           const DevInst *inst, const vrgi_bvh_node *tl, uint32_t n_ltri,
                           uint32_t *lanc, uint32_t *lanc_len);
void launch_frame_stats(hipStream_t s, const uint32_t *qcount, uint32_t B, uint32_t P,
                        unsigned long long *counters);
// learned radiance cache (nrc.hip): 12 grid levels x 2 features, 64 inputs, 4 x 64
// hidden, 3 outputs; weights [k][n] per layer, layer 5 padded to 16 columns
#define NRC_LEVELS 12
#define NRC_N_WEIGHTS (4 * 64 * 64 + 64 * 16)
struct NrcNet {
    uint32_t log2_T;
    uint32_t res[16];
    float lo[3], hi[3];       // box mapped to the unit cube
    const float *table;       // levels x 2^log2_T x 2, fp32
    const uint16_t *wt;       // 16-bit [n][k] per layer (forward A operand)
    const uint16_t *wb;       // 16-bit [k][n] per layer (backward A operand)
    uint32_t prec;            // 0 bf16, 1 fp16
};
// training records: 16 floats = position, direction, normal, albedo, target, key
struct NrcTrain {
    const float *rec;
    const uint32_t *count;    // records available (clamped to rec_cap)
    uint32_t rec_cap;
    uint32_t step_k, steps;   // this step uses records step_k + i * steps
    uint32_t max_batch;       // multiple of 16; row stride of the planes
    uint32_t loss;            // 0 relative L2, 1 plain L2
    float loss_scale;
    uint16_t *xplane, *gplane;  // [5][64][max_batch] 16-bit: layer inputs, output gradients
    float *gw, *gtable;         // gradient accumulators (scaled)
    float *loss_out;
    // dynamic loss scaling (null: the fixed loss_scale): scale, good steps in a row,
    // overflow flag (u32 bits), skipped steps; the scale doubles after scale_interval
    // good steps in a row
    float *scale_state;
    uint32_t scale_interval;
};
struct NrcAdam {
    float lr, beta1, beta2, eps;
    float c1, c2;             // 1 - beta^t
    float inv_scale;          // 1 / loss scale
};
void launch_nrc_infer(hipStream_t s, const NrcNet &net, const float *rec, const uint32_t *count, uint32_t max_n,
                      float *out, float *out_px);
void launch_nrc_train_step(hipStream_t s, const NrcNet &net, const NrcTrain &tr, float *w32, float *mw, float *vw,
                           float *table, float *mt, float *vt, size_t n_table, const NrcAdam &adam);
void launch_nrc_grads(hipStream_t s, const NrcNet &net, const NrcTrain &tr);
void launch_nrc_pack(hipStream_t s, const float *w32, uint16_t *wt, uint16_t *wb, uint32_t prec);
void launch_nrc_zero(hipStream_t s, float *p, size_t n);
// training records in Morton order of their position in the cache box; hist:
// NRC_SORT_BUCKETS + NRC_SORT_BUCKETS / 1024 words, key_rank: 2 words per record, out:
// cap records
static constexpr uint32_t NRC_SORT_BUCKETS = 32768;
struct RecBox {
    float lo[3], hi[3];
};
void launch_rec_sort(hipStream_t s, const float *rec, const uint32_t *count, uint32_t cap, const RecBox &box,
                     uint32_t *hist, uint32_t *key_rank, float *out);
// held-out error per screen tile: records of vertex `vertex` against the cache's output
void launch_tile_stats(hipStream_t s, const float *rec, const uint32_t *count, uint32_t cap, const float *pred,
                       uint32_t vertex, uint32_t W, uint32_t tile, uint32_t tiles_x, float *acc);
void launch_tile_update(hipStream_t s, float *acc, float *ema, uint8_t *fallback, uint32_t n_tiles, float alpha,
                        float bound);
// a pixel of a fallback tile traces a full path, like a training pixel
__device__ __forceinline__ bool in_fallback(const uint8_t *fb, uint32_t px, uint32_t W, uint32_t tile, uint32_t tx) {
    return fb && fb[((px / W) / tile) * tx + (px % W) / tile];
}
// a cache frame's pixel that traces a full path: a training pixel or one of a fallback tile
__device__ __forceinline__ bool full_path_px(uint32_t train_seed, float train_frac, const uint8_t *fb, uint32_t px,
                                             uint32_t W, uint32_t tile, uint32_t tx) {
    return mm_finalizef(mm_combine(train_seed, px)) < train_frac || in_fallback(fb, px, W, tile, tx);
}

// ---------------------------------------------------------------- batched scenes
// Many scenes in one launch per stage: each scene owns a contiguous range of blocks
// (`blk[s]` .. `blk[s + 1]`), so every wave belongs to one scene and reads that scene's
// arguments from the per-scene arrays below at a uniform address.
struct BatchPass {   // per pass of a batched frame
    uint32_t first, last;  // first / last of the frame's spp passes
    float inv_spp;
    uint32_t pad;
};
struct BatchLaunch {
    const SceneView *svs;     // [n]
    const BounceArgs *tba;    // [n][B]: the trace of bounce b
    const BounceArgs *nba;    // [n][B]: the NEE pdf of bounce b
    const FrameArgs *fas;     // [n]
    const DiArgs *das;        // [n]
    const uint32_t *blk_ray;  // [n + 1] block prefix for 64 * wpb rays (or pixels) per block
    const uint32_t *blk_px;   // [n + 1] block prefix for pb pixels per block
    float *const *mean;       // [n] spp mean of the ReSTIR output
    const BatchPass *pass;
    uint32_t n, B, ray_blocks, px_blocks;
    uint32_t gneed, lneed;    // traversal stack entries of the deepest geometry / light tree
    uint32_t wpb, pb;
    bool nee, di;
};
__device__ __forceinline__ uint32_t batch_scene(const uint32_t *__restrict__ blk, uint32_t n, uint32_t b) {
    uint32_t lo = 0, hi = n - 1;  // the last scene whose first block is <= b
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (blk[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
void launch_batch_pass(hipStream_t s, const BatchLaunch &L);
void launch_batch_pixels_gi(hipStream_t s, const BatchLaunch &L);  // accum + temporal
void launch_batch_pixels_spatial(hipStream_t s, const BatchLaunch &L);
void launch_batch_pixels_post(hipStream_t s, const BatchLaunch &L);  // postprocess + spp mean

// exposed-face meshing of 32^3 chunks (mesher.hip)
struct MeshArgsHost {
    const uint16_t *blocks;      // dims[0] x dims[1] x dims[2], x-major
    uint32_t dims[3];
    uint32_t air;
    const uint8_t *translucent;
    const uint32_t *chunks;      // n_chunks x (cx, cy, cz)
    uint32_t n_chunks;
    uint8_t *mask;               // [n_chunks][32768]
    uint32_t *face_off;          // [n_chunks][32768]
    uint32_t *faces;             // [n_chunks]
    const uint64_t *vert_off;    // [n_chunks]
    vrgi_vertex3d *out;
};
void launch_mesh_count(hipStream_t s, const MeshArgsHost &h);
// device-resident voxel edits (mesher.hip)
void launch_vx_set(hipStream_t s, uint16_t *blocks, const uint32_t dims[3], uint64_t n, const uint32_t *xyz,
                   const uint16_t *val, uint8_t *dirty);
void launch_light_scan(hipStream_t s, const MeshArgsHost &h, const uint8_t *em, uint32_t *lface_off, uint32_t *lfaces);
void launch_light_gather(hipStream_t s, const MeshArgsHost &h, const float *lum, const uint8_t *em,
                         const uint32_t *lface_off, const uint64_t *light_off, float *tri9, float *power,
                         float *prim9, uint32_t *ids, uint32_t n_light);
void launch_light_pairs_multi(hipStream_t s, const vrgi_bvh_node *nodes, const uint32_t *base, uint32_t n_trees,
                              uint32_t n_nodes, LPair *out);
void launch_mesh_emit(hipStream_t s, const MeshArgsHost &h);
// light vertices of many light BL-BVHs (tree t: nodes tree_base[t] .. tree_base[t + 1], its
// lights from tree_light[t], ids ascending within a tree) in leaf order, and each tree's
// root node; `rank` holds 2 * n_nodes words, `scratch` the scan's
// light_verts_scratch_bytes(n_nodes) bytes
void launch_light_verts_multi(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, const uint32_t *tree_base,
                              const uint32_t *tree_light, uint32_t n_trees, const uint32_t *ids, const float *tri9,
                              uint32_t *rank, void *scratch, size_t scratch_bytes, vrgi_light_vertex3d *out,
                              vrgi_bvh_node *roots);
size_t light_verts_scratch_bytes(uint32_t n_nodes);

// queue reordering by (direction octant, origin Morton code) between bounces
struct SortArgs {
    uint32_t bounce, P, bits;         // rays of bounce + 1 read their origin / dir from bounce
    const WideNode *root;             // root node of the geometry tree: scene box
    const float *origin, *dir;
    const uint32_t *count;            // device: rays in the queue
    const uint32_t *queue_in;
    uint32_t *queue_out, *keys, *hist;  // hist: 8 << (3 bits) counters
};
void launch_sort_queue(hipStream_t s, const SortArgs &a, uint32_t max_rays);
void launch_popcount(hipStream_t s, const uint32_t *bits, size_t n_words, unsigned long long *out);
void launch_di(hipStream_t s, const SceneView &sv, const DiArgs &da);

}  // namespace vrgi
