// trace.hip — wavefront path-tracing kernels for gfx950: scene flattening, traversal
// layout, closest-hit / all-hit software traversal (CDNA4 has no ray-tracing units),
// the per-bounce trace+shade kernel with light-BVH NEE, and the NEE pdf kernel.
//
// Replaces raygen.rs:84-111, raytrace.rs:390-577 and nee_pdf.rs:281-337. The hardware
// ray queries (raytrace.rs:354-388, nee_pdf.rs:305-334) become a stack traversal of a
// binned-SAH BVH with the Woop-Benthin-Wald watertight triangle test (JCGT 2013) and
// a conservative slab test (Ize, JCGT 2013); equal-t hits resolve to the lower
// primitive id, so the closest hit does not depend on the traversal order.
#include <hip/hip_runtime.h>

#include "devcommon.h"
#include "vrgi_kernels.h"

namespace vrgi {

// ---------------------------------------------------------------- flattening
// world-space triangles of every instance (instances in key order), builder inputs
__global__ void k_flatten(const FlattenArgs f) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= f.n_tris) return;
    // instance of triangle g: last instance with tri_off <= g (binary search)
    uint32_t lo = 0, hi = f.n_inst;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (f.inst[mid].tri_off <= g) lo = mid;
        else hi = mid;
    }
    const DevInst &in = f.inst[lo];
    const vrgi_vertex3d *v = f.vtx[lo] + 3 * (g - in.tri_off);
    f3 w[3];
    for (int k = 0; k < 3; ++k) {
        f3 p = xform_pt(in.xf, mk(v[k].position[0], v[k].position[1], v[k].position[2]));
        w[k] = mk(p.x + 0.0f, p.y + 0.0f, p.z + 0.0f);  // -0.0 -> +0.0
    }
    Tri t;
    t.v0 = make_float4(w[0].x, w[0].y, w[0].z, __uint_as_float(v[0].t));
    t.v1 = make_float4(w[1].x, w[1].y, w[1].z, 0.0f);
    t.v2 = make_float4(w[2].x, w[2].y, w[2].z, 0.0f);
    f.tri[g] = t;
    f.tri_inst[g] = lo;
    write_prim9(f.prim9 + f.prim_off, f.prim_stride, g, w);
    f.power[f.prim_off + g] = 0.0f;
    f.ids[f.prim_off + g] = g;
}

// world-space light triangles from the instances' light vertices
__global__ void k_flatten_lights(const FlattenArgs f) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= f.n_tris) return;
    uint32_t lo = 0, hi = f.n_inst;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (f.light_off[mid] <= g) lo = mid;
        else hi = mid;
    }
    const DevInst &in = f.inst[lo];
    uint32_t j = g - f.light_off[lo];
    const vrgi_light_vertex3d *v = f.lvtx[lo] + 3 * j;
    f3 w[3];
    for (int k = 0; k < 3; ++k) {
        f3 p = xform_pt(in.xf, mk(v[k].position[0], v[k].position[1], v[k].position[2]));
        w[k] = mk(p.x + 0.0f, p.y + 0.0f, p.z + 0.0f);
    }
    Tri t;
    t.v0 = make_float4(w[0].x, w[0].y, w[0].z, 0.0f);
    t.v1 = make_float4(w[1].x, w[1].y, w[1].z, 0.0f);
    t.v2 = make_float4(w[2].x, w[2].y, w[2].z, 0.0f);
    f.tri[g] = t;
    f.tri_inst[g] = lo;
    f.light_node[g] = v[0].bvh_node_idx;
    write_prim9(f.prim9 + f.prim_off, f.prim_stride, g, w);
    f.power[f.prim_off + g] = 0.0f;
    f.ids[f.prim_off + g] = g;
}

// traversal copy: internal node of preorder rank k (= (left - 1) / 2) -> WideNode[k]
__global__ void k_wide(const vrgi_bvh_node *nodes, uint32_t n_nodes, WideNode *wide) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    vrgi_bvh_node nd = nodes[i];
    if (nd.left == VRGI_INVALID) return;
    uint32_t rank = (nd.left - 1) / 2;
    vrgi_bvh_node c0 = nodes[nd.left], c1 = nodes[nd.right_or_prim];
    WideNode w;
    const float e = VRGI_BOX_PAD;
    w.a = make_float4(c0.min[0] - e, c0.min[1] - e, c0.min[2] - e, c0.max[0] + e);
    w.b = make_float4(c0.max[1] + e, c0.max[2] + e, c1.min[0] - e, c1.min[1] - e);
    w.c = make_float4(c1.min[2] - e, c1.max[0] + e, c1.max[1] + e, c1.max[2] + e);
    uint32_t r0 = c0.left == VRGI_INVALID ? (0x80000000u | c0.right_or_prim) : (c0.left - 1) / 2;
    uint32_t r1 = c1.left == VRGI_INVALID ? (0x80000000u | c1.right_or_prim) : (c1.left - 1) / 2;
    w.r = make_uint4(r0, r1, 0, 0);
    wide[rank] = w;
}

// ---------------------------------------------------------------- closest hit API kernel
template <int STACK>
__global__ __launch_bounds__(64) void k_trace_closest(const SceneView sv, const float *orig,
                                                      const float *dirs, uint32_t n, float tmin,
                                                      float tmax, uint32_t *hit, float *ht,
                                                      float *huv) {
    __shared__ uint32_t stack[STACK * 64];
    uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    RayPre r = ray_pre(ld3(orig + 3 * i), ld3(dirs + 3 * i));
    Hit h = closest_hit<STACK>(sv, r, tmin, tmax, stack);
    hit[i] = h.prim;
    ht[i] = h.t;
    huv[2 * i] = h.u;
    huv[2 * i + 1] = h.v;
}

// ---------------------------------------------------------------- light BVH walks
// raytrace.rs:181-208 and nee_pdf.rs:125-152: power / dist_sq * (visible corners / 8)
__device__ __forceinline__ float node_importance(f3 p, f3 n, const float *m,
                                                 const vrgi_bvh_node &nd, float eps) {
    f3 a = xform_pt(m, mk(nd.min[0], nd.min[1], nd.min[2]));
    f3 b = xform_pt(m, mk(nd.max[0], nd.max[1], nd.max[2]));
    uint32_t vis = 0;
    vis += dot3(a - p, n) >= eps;
    vis += dot3(mk(a.x, a.y, b.z) - p, n) >= eps;
    vis += dot3(mk(a.x, b.y, a.z) - p, n) >= eps;
    vis += dot3(mk(a.x, b.y, b.z) - p, n) >= eps;
    vis += dot3(mk(b.x, a.y, a.z) - p, n) >= eps;
    vis += dot3(mk(b.x, a.y, b.z) - p, n) >= eps;
    vis += dot3(mk(b.x, b.y, a.z) - p, n) >= eps;
    vis += dot3(b - p, n) >= eps;
    f3 diag = b - a;
    f3 mid = 0.5f * (a + b) - p;
    float dsq = gmax(dot3(diag, diag), dot3(mid, mid));
    return nd.power / dsq * ((float)vis / 8.0f);
}

__device__ const float IDENTITY12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};

struct Pick { bool ok; uint32_t inst, prim; float prob, importance; };

// traverseBvh, raytrace.rs:218-281
__device__ Pick traverse_light(const SceneView &sv, f3 p, f3 n, uint32_t seed) {
    const vrgi_bvh_node *root = sv.tl;
    vrgi_bvh_node node = root[0];
    Pick res{false, 0, 0, 1.0f, 0.0f};
    if (node.left == VRGI_INVALID && node.right_or_prim == VRGI_INVALID) return res;
    float prob = 1.0f, imp = 0.0f;
    const float *m = IDENTITY12;
    uint32_t inst = VRGI_INVALID;
    bool top = true;
    for (int guard = 0; guard < 4096; ++guard) {
        if (top && node.left == VRGI_INVALID) {
            inst = node.right_or_prim;
            m = sv.inst[inst].xf;
            top = false;
            root = sv.inst[inst].light_bvh;
            node = root[0];
            if (imp == 0.0f) imp = node_importance(p, n, m, node, VRGI_EPS_RT);
        }
        if (!top && node.left == VRGI_INVALID) {
            res.ok = true; res.inst = inst; res.prim = node.right_or_prim; res.prob = prob; res.importance = imp;
            return res;
        }
        vrgi_bvh_node l = root[node.left], r = root[node.right_or_prim];
        float li = node_importance(p, n, m, l, VRGI_EPS_RT);
        float ri = node_importance(p, n, m, r, VRGI_EPS_RT);
        float tot = li + ri;
        float ln = li / tot, rn = ri / tot;
        if (mm_finalizef(seed) < ln) { node = l; prob *= ln; imp = li; }
        else { node = r; prob *= rn; imp = ri; }
        seed = mm_combine(seed, 0);
    }
    return res;
}

// reverseTraverseBvh, nee_pdf.rs:154-228
__device__ float reverse_traverse(const SceneView &sv, f3 p, f3 n, uint32_t inst, uint32_t bl_idx) {
    const vrgi_bvh_node *root = sv.inst[inst].light_bvh;
    float prob = 1.0f;
    const float *m = sv.inst[inst].xf;
    bool top = false;
    vrgi_bvh_node node = root[bl_idx];
    uint32_t ni = bl_idx;
    for (int guard = 0; guard < 4096; ++guard) {
        uint32_t old = ni;
        if (node.parent == VRGI_INVALID) {
            if (top) return prob;
            top = true;
            root = sv.tl;
            old = sv.inst[inst].tl_idx;
            vrgi_bvh_node tl_leaf = root[old];
            if (tl_leaf.parent == VRGI_INVALID) return prob;
            ni = tl_leaf.parent;
            node = root[ni];
            m = IDENTITY12;
        } else {
            ni = node.parent;
            node = root[ni];
        }
        vrgi_bvh_node l = root[node.left], r = root[node.right_or_prim];
        float li = node_importance(p, n, m, l, VRGI_EPS_NEE);
        float ri = node_importance(p, n, m, r, VRGI_EPS_NEE);
        float tot = li + ri;
        float ln = li / tot, rn = ri / tot;
        prob *= (old == node.left) ? ln : rn;
    }
    return prob;
}

// computeNeePdf, nee_pdf.rs:240-279, for world light triangle lid; barycentric
// convention (x, y, 1-x-y) of pointOnTriangle (nee_pdf.rs:230-232)
__device__ float nee_pdf_one(const SceneView &sv, f3 p, f3 n, f3 w, uint32_t lid, float bx, float by) {
    Tri t = sv.ltri[lid];
    f3 t0 = mk(t.v0.x, t.v0.y, t.v0.z), t1 = mk(t.v1.x, t.v1.y, t.v1.z), t2 = mk(t.v2.x, t.v2.y, t.v2.z);
    f3 q = (bx * t0 + by * t1) + ((1.0f - bx) - by) * t2;
    float dist = len3(q - p);
    float area = 0.5f * len3(cross3(t0 - t1, t2 - t1));
    float cs = dot3(n, w);
    float ppp = dist * dist / (cs * area);
    float prim_p = reverse_traverse(sv, p, n, sv.ltri_inst[lid], sv.ltri_node[lid]);
    return prim_p * ppp;
}

// ---------------------------------------------------------------- trace + shade (one bounce)
struct IcS { f3 n, t, b; };
__device__ __forceinline__ IcS local_cs(f3 a, f3 b, f3 c) {  // raytrace.rs:319-331
    f3 e1 = b - a, e2 = c - a;
    f3 nn = cross3(e1, e2);
    f3 bt = cross3(nn, e1);
    IcS r;
    r.n = norm3(nn);
    r.t = norm3(e1);
    r.b = norm3(bt);
    return r;
}

template <int STACK, bool COUNT>
__global__ __launch_bounds__(64) void k_trace_shade(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t b = ba.bounce;
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    const uint32_t count = b == 0 ? ba.P : ba.qcount[b];
    bool active = i < count;
    uint32_t px = 0;
    f3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    if (active) {
        if (b == 0) {
            // raygen.rs:84-111 (jitter is multiplied by 0.00)
            px = i;
            uint32_t x = px % ba.W, y = px / ba.W;
            float uvx = 2.0f * (float)x / (float)ba.W - 1.0f;
            float uvy = 2.0f * (float)y / (float)ba.H - 1.0f;
            float aspect = (float)ba.W / (float)ba.H;
            f3 right = ld3(ba.cam.right), up = ld3(ba.cam.up), front = ld3(ba.cam.front);
            o = ld3(ba.cam.eye);
            d = norm3((right * (uvx + 0.0f) * aspect + up * (uvy + 0.0f)) + front);
        } else {
            px = ba.queue_in[i];
            size_t src = ((size_t)(b - 1) * ba.P + px) * 3;
            o = ld3(ba.origin + src);
            d = ld3(ba.dir + src);
        }
    }
    const size_t bi = (size_t)b * ba.P + px;
    const uint32_t seed = mm_combine(ba.inv_seed, px);  // raytrace.rs:403
    Hit h;
    h.prim = VRGI_INVALID;
    if (active) {
        RayPre r = ray_pre(o, d);
        h = closest_hit<STACK, COUNT>(sv, r, VRGI_EPS_RT, 1000.0f, stack, ba.touch_nodes, ba.touch_tris);
    }
    f3 out_o, out_d = mk(0, 0, 0), nrm = mk(0, 0, 0), emis = mk(0, 0, 0), refl = mk(0, 0, 0);
    float mis = 0.0f, pdf = 1.0f;
    if (active && h.prim == VRGI_INVALID) {  // raytrace.rs:421-430
        out_o = o + d * 5000.0f;
        float sky = dot3(d, mk(0, 1, 0)) > 0.9f ? 50.0f : 0.0f;
        emis = mk(sky, sky, sky);
    } else if (active) {
        Tri tr = sv.tri[h.prim];
        f3 t0 = mk(tr.v0.x, tr.v0.y, tr.v0.z), t1 = mk(tr.v1.x, tr.v1.y, tr.v1.z), t2 = mk(tr.v2.x, tr.v2.y, tr.v2.z);
        uint32_t slot = __float_as_uint(tr.v0.w);
        f3 b3 = mk((1.0f - h.u) - h.v, h.u, h.v);
        IcS ics = local_cs(t0, t1, t2);
        const vrgi_material mt = sv.mats[slot];
        out_o = (t0 * b3.x + t1 * b3.y) + t2 * b3.z;
        refl = mk(mt.reflectivity[0], mt.reflectivity[1], mt.reflectivity[2]);
        float alpha = mt.reflectivity[3];
        emis = (1000.0f * mk(mt.emissivity[0], mt.emissivity[1], mt.emissivity[2])) * -dot3(d, ics.n);
        float metal = mt.metallicity[0];
        float albedo_max;
        float u0 = mm_finalizef(mm_combine(seed, 0));
        if (u0 < metal) {  // mirror
            pdf = 1.0f;
            out_d = d - (2.0f * dot3(ics.n, d)) * ics.n;
            albedo_max = gmax(gmax(refl.x, refl.y), refl.z);
        } else if (u0 < metal + (1.0f - alpha)) {  // transmission
            out_d = d;
            refl = mk(1, 1, 1);
            pdf = 1.0f;
            albedo_max = 1.0f;
        } else {  // Lambert
            out_o = out_o + (VRGI_EPS_RT * 1.5f) * ics.n;
            albedo_max = gmax(gmax(refl.x, refl.y), refl.z);
            refl = refl / VRGI_PI_F;
            Pick pk{false, 0, 0, 1.0f, 0.0f};
            if (ba.nee_type == 1 || (ba.nee_type == 2 && b == 1))
                pk = traverse_light(sv, out_o, ics.n, mm_combine(seed, 2));
            if (pk.ok && pk.importance > 0.0f) mis = 0.3f;
            float u3 = mm_finalizef(mm_combine(seed, 3));
            float ua = mm_finalizef(mm_combine(seed, 4)), ub = mm_finalizef(mm_combine(seed, 5));
            if (u3 < mis) {  // light point, triangleSample (raytrace.rs:305-311)
                Tri lt = sv.tri[sv.inst[pk.inst].tri_off + pk.prim];
                f3 l0 = mk(lt.v0.x, lt.v0.y, lt.v0.z), l1 = mk(lt.v1.x, lt.v1.y, lt.v1.z), l2 = mk(lt.v2.x, lt.v2.y, lt.v2.z);
                float x = ua, y = ub;
                if (x + y > 1.0f) { x = 1.0f - x; y = 1.0f - y; }
                f3 bb = mk((1.0f - x) - y, x, y);
                f3 pt = (bb.x * l0 + bb.y * l1) + bb.z * l2;
                out_d = norm3(pt - out_o);
            } else {  // uniform hemisphere (raytrace.rs:286-291, 335-338)
                float u = 2.0f * VRGI_PI_F * ua;
                float vv = sqrtf(gmax(0.0f, 1.0f - ub * ub));
                float sn, cs;
                det_sincos(u, &sn, &cs);
                f3 hs = mk(vv * cs, ub, vv * sn);
                out_d = norm3((hs.x * ics.t + hs.y * ics.n) + hs.z * ics.b);
            }
            pdf = dot3(out_d, ics.n) / VRGI_PI_F;
        }
        if (b >= ba.rr_start) {  // Russian roulette of the benchmark workloads
            float surv = gmin(1.0f, gmax(ba.rr_min, albedo_max));
            float urr = mm_finalizef(mm_combine(seed, 6));
            if (urr >= surv) out_d = mk(0, 0, 0);
            else refl = refl / surv;
        }
        nrm = ics.n;
    }
    bool cont = false;
    if (active) {
        st3(ba.origin + bi * 3, out_o);
        st3(ba.dir + bi * 3, out_d);
        st3(ba.normal + bi * 3, nrm);
        st3(ba.emis + bi * 3, emis);
        st3(ba.refl + bi * 3, refl);
        ba.mis[bi] = mis;
        ba.pdf[bi] = pdf;
        ba.hit[bi] = h.prim;
        ba.hit_t[bi] = h.prim == VRGI_INVALID ? 0.0f : h.t;
        cont = !zero3(out_d) && b + 1 < ba.B;
        if (!cont) ba.path_len[px] = b + 1;
    }
    // compaction of the rays alive at the next bounce: ballot + one atomic per wave
    uint64_t bal = __ballot(cont);
    if (bal) {
        uint32_t lane = threadIdx.x;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(ba.qcount + b + 1, (uint32_t)__popcll(bal));
        base = __shfl(base, 0, 64);
        if (cont) ba.queue_out[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = px;
    }
}

// NEE pdf of bounce b (nee_pdf.rs:281-337) for the rays alive at bounce b+1
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_pdf(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t b = ba.bounce;  // pdf of bounce b, rays of bounce b + 1
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= ba.qcount[b + 1]) return;
    const uint32_t px = ba.queue_in[i];
    const size_t bi = (size_t)b * ba.P + px;
    float w = ba.mis[bi];
    f3 d = ld3(ba.dir + bi * 3);
    float np = 0.0f;
    const bool traced = !(len3(d) == 0.0f || w == 0.0f) && sv.n_ltri > 0;
    uint64_t bal = __ballot(traced);
    if ((threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)bal) - 1))
        atomicAdd(ba.counters + 1, (unsigned long long)__popcll(bal));
    if (traced) {
        f3 o = ld3(ba.origin + bi * 3);
        f3 n = ld3(ba.normal + bi * 3);
        RayPre r = ray_pre(o, d);
        // every light triangle crossed in [1e-4, 1000], left-to-right leaf order
        uint32_t *stk = stack + threadIdx.x;
        int sp = 0;
        uint32_t ref;
        bool have;
        if (sv.n_ltri == 1) { ref = 0x80000000u | sv.lroot_prim; have = true; }
        else { ref = 0; have = true; }
        while (have) {
            if (ref & 0x80000000u) {
                uint32_t lid = ref & 0x7FFFFFFFu;
                Tri t = sv.ltri[lid];
                float tt, u, v;
                if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z),
                            mk(t.v2.x, t.v2.y, t.v2.z), VRGI_EPS_NEE, 1000.0f, &tt, &u, &v))
                    np += nee_pdf_one(sv, o, n, d, lid, u, v);
                have = false;
            } else {
                WideNode wn = sv.lwide[ref];
                bool h0 = box_hit(r, wn.a.x, wn.a.y, wn.a.z, wn.a.w, wn.b.x, wn.b.y, VRGI_EPS_NEE, 1000.0f);
                bool h1 = box_hit(r, wn.b.z, wn.b.w, wn.c.x, wn.c.y, wn.c.z, wn.c.w, VRGI_EPS_NEE, 1000.0f);
                have = false;
                if (h1 && h0) { stk[(sp++) * 64] = wn.r.y; ref = wn.r.x; have = true; }
                else if (h0) { ref = wn.r.x; have = true; }
                else if (h1) { ref = wn.r.y; have = true; }
            }
            if (!have && sp > 0) { ref = stk[(--sp) * 64]; have = true; }
        }
    }
    ba.nee[bi] = np;
}

// ---------------------------------------------------------------- host launchers
// frame counters: closest-hit rays = P + live rays of bounces 1..B-1
__global__ void k_frame_stats(const uint32_t *qcount, uint32_t B, uint32_t P, unsigned long long *c) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    unsigned long long r = P;
    for (uint32_t b = 1; b < B; ++b) r += qcount[b];
    c[0] += r;
    c[2] += 1;
}
void launch_frame_stats(hipStream_t s, const uint32_t *qcount, uint32_t B, uint32_t P,
                        unsigned long long *counters) {
    k_frame_stats<<<1, 64, 0, s>>>(qcount, B, P, counters);
}

void launch_flatten(hipStream_t s, const FlattenArgs &f) {
    if (f.n_tris) k_flatten<<<(f.n_tris + 255) / 256, 256, 0, s>>>(f);
}
void launch_flatten_lights(hipStream_t s, const FlattenArgs &f) {
    if (f.n_tris) k_flatten_lights<<<(f.n_tris + 255) / 256, 256, 0, s>>>(f);
}
void launch_wide(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, WideNode *wide) {
    if (n_nodes > 1) k_wide<<<(n_nodes + 255) / 256, 256, 0, s>>>(nodes, n_nodes, wide);
}
void launch_trace_closest(hipStream_t s, const SceneView &sv, const float *o, const float *d,
                          uint32_t n, float tmin, float tmax, uint32_t *hit, float *t, float *uv) {
    if (!n) return;
    if (sv.depth + 2 <= 64) k_trace_closest<64><<<(n + 63) / 64, 64, 0, s>>>(sv, o, d, n, tmin, tmax, hit, t, uv);
    else k_trace_closest<VRGI_MAX_STACK><<<(n + 63) / 64, 64, 0, s>>>(sv, o, d, n, tmin, tmax, hit, t, uv);
}
void launch_trace_shade(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays) {
    if (!max_rays) return;
    bool cnt = ba.touch_nodes != nullptr;
    if (sv.depth + 2 <= 64) {
        if (cnt) k_trace_shade<64, true><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
        else k_trace_shade<64, false><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
    } else {
        if (cnt) k_trace_shade<VRGI_MAX_STACK, true><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
        else k_trace_shade<VRGI_MAX_STACK, false><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
    }
}

// number of set bits of a bitmap, added to *out
__global__ void k_popcount(const uint32_t *bits, size_t n_words, unsigned long long *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = i < n_words ? (unsigned long long)__popc(bits[i]) : 0ull;
    for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}
void launch_popcount(hipStream_t s, const uint32_t *bits, size_t n_words, unsigned long long *out) {
    if (n_words) k_popcount<<<(unsigned)((n_words + 255) / 256), 256, 0, s>>>(bits, n_words, out);
}
void launch_nee_pdf(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays) {
    if (!max_rays) return;
    if (sv.ldepth + 2 <= 64) k_nee_pdf<64><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
    else k_nee_pdf<VRGI_MAX_STACK><<<(max_rays + 63) / 64, 64, 0, s>>>(sv, ba);
}

}  // namespace vrgi
