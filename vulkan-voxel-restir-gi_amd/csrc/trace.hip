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

#include <type_traits>

#include "devcommon.h"
#include "vrgi_kernels.h"

namespace vrgi {

// ---------------------------------------------------------------- flattening
// world-space triangles of every instance (instances in key order), builder inputs
// Instance of entry g of a list that the instances fill in order from their first entries
// first(k): the last k with first(k) <= g. One thread of the block binary-searches the
// block's first entry; each thread then steps forward from there (a block's 256 entries
// span one or two instances of a world's chunks), instead of every thread searching.
template <class F>
__device__ __forceinline__ uint32_t block_instance(uint32_t g, uint32_t n_inst, F first) {
    __shared__ uint32_t s_lo;
    if (threadIdx.x == 0) {
        const uint32_t g0 = blockIdx.x * blockDim.x;
        uint32_t lo = 0, hi = n_inst;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (first(mid) <= g0) lo = mid;
            else hi = mid;
        }
        s_lo = lo;
    }
    __syncthreads();
    uint32_t lo = s_lo;
    while (lo + 1 < n_inst && first(lo + 1) <= g) ++lo;
    return lo;
}

__global__ void k_flatten(const FlattenArgs f) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    // instance of triangle g: last instance with tri_off <= g
    const uint32_t lo = block_instance(g, f.n_inst, [&](uint32_t k) { return f.inst[k].tri_off; });
    if (g >= f.n_tris) return;
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
    if (f.tri_uv)
        for (int k = 0; k < 3; ++k) f.tri_uv[3 * (size_t)g + k] = make_float2(v[k].uv[0], v[k].uv[1]);
    write_prim9(f.prim9 + f.prim_off, f.prim_stride, g, w);
    f.power[f.prim_off + g] = 0.0f;
    f.ids[f.prim_off + g] = g;
}

// world-space light triangles from the instances' light vertices
__global__ void k_flatten_lights(const FlattenArgs f) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lo = block_instance(g, f.n_inst, [&](uint32_t k) { return f.light_off[k]; });
    if (g >= f.n_tris) return;
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
// one thread per internal node rank: the children of rank r are nodes 2r + 1 and 2r + 2
// (build.rs's insert_blas_leaf_node numbering), read as one 80-byte pair
__global__ void k_wide(const vrgi_bvh_node *nodes, uint32_t n_nodes, WideNode *wide) {
    const uint32_t rank = blockIdx.x * blockDim.x + threadIdx.x;
    if (rank >= (n_nodes - 1) / 2) return;
    const vrgi_bvh_node c0 = nodes[2 * rank + 1], c1 = nodes[2 * rank + 2];
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

// 4-ary copy: every internal node of rank k gets its children, or its grandchildren in
// place of an internal child, in left-to-right order (boxes padded as in k_wide)
__global__ void k_wide4(const vrgi_bvh_node *nodes, uint32_t n_nodes, Wide4 *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const vrgi_bvh_node nd = nodes[i];
    if (nd.left == VRGI_INVALID) return;
    const float e = VRGI_BOX_PAD;
    float lo[3][4], hi[3][4];
    uint32_t ref[4];
    int k = 0;
    auto put = [&](const vrgi_bvh_node &c) {
        for (int a = 0; a < 3; ++a) { lo[a][k] = c.min[a] - e; hi[a][k] = c.max[a] + e; }
        ref[k] = c.left == VRGI_INVALID ? (0x80000000u | c.right_or_prim) : (c.left - 1) / 2;
        ++k;
    };
    const uint32_t ch[2] = {nd.left, nd.right_or_prim};
    for (int j = 0; j < 2; ++j) {
        const vrgi_bvh_node c = nodes[ch[j]];
        if (c.left == VRGI_INVALID) {
            put(c);
        } else {
            put(nodes[c.left]);
            put(nodes[c.right_or_prim]);
        }
    }
    for (; k < 4; ++k) {
        for (int a = 0; a < 3; ++a) { lo[a][k] = 1.0f; hi[a][k] = 0.0f; }
        ref[k] = VRGI_INVALID;
    }
    Wide4 w;
    w.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    w.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    w.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    w.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    w.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    w.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    w.ref = make_uint4(ref[0], ref[1], ref[2], ref[3]);
    w.pad = make_uint4(0, 0, 0, 0);
    out[(nd.left - 1) / 2] = w;
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
__device__ __forceinline__ float box_importance(f3 p, f3 n, const float *m, f3 mn, f3 mx, float power,
                                                float eps) {
    f3 a = xform_pt(m, mn);
    f3 b = xform_pt(m, mx);
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
    return power / dsq * ((float)vis / 8.0f);
}
// The same importance for a transform whose linear part is the identity: xform_pt then
// gives ((x + 0) + 0) + t_x = x + t_x exactly, up to the sign of a zero result, which no
// corner test, difference or square below can see; the instance-to-world offset t is
// added once per corner coordinate instead of the 3x3 product.
__device__ __forceinline__ float box_importance_t(f3 p, f3 n, f3 t, f3 mn, f3 mx, float power, float eps) {
    f3 a = mn + t;
    f3 b = mx + t;
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
    return power / dsq * ((float)vis / 8.0f);
}
__device__ __forceinline__ float node_importance(f3 p, f3 n, const float *m,
                                                 const vrgi_bvh_node &nd, float eps) {
    return box_importance(p, n, m, mk(nd.min[0], nd.min[1], nd.min[2]), mk(nd.max[0], nd.max[1], nd.max[2]),
                          nd.power, eps);
}
// normalised importances of both children of one record (raytrace.rs:266-271,
// nee_pdf.rs:201-206)
struct PairImp { float li, ri, ln, rn; };
__device__ __forceinline__ PairImp pair_importance(f3 p, f3 n, const float *m, const LPair &r, float eps) {
    PairImp q;
    q.li = box_importance(p, n, m, mk(r.a.x, r.a.y, r.a.z), mk(r.b.x, r.b.y, r.b.z), r.a.w, eps);
    q.ri = box_importance(p, n, m, mk(r.c.x, r.c.y, r.c.z), mk(r.d.x, r.d.y, r.d.z), r.b.w, eps);
    float tot = q.li + q.ri;
    q.ln = q.li / tot;
    q.rn = q.ri / tot;
    return q;
}
__device__ __forceinline__ PairImp pair_importance_t(f3 p, f3 n, f3 t, const LPair &r, float eps) {
    PairImp q;
    q.li = box_importance_t(p, n, t, mk(r.a.x, r.a.y, r.a.z), mk(r.b.x, r.b.y, r.b.z), r.a.w, eps);
    q.ri = box_importance_t(p, n, t, mk(r.c.x, r.c.y, r.c.z), mk(r.d.x, r.d.y, r.d.z), r.b.w, eps);
    float tot = q.li + q.ri;
    q.ln = q.li / tot;
    q.rn = q.ri / tot;
    return q;
}
__device__ __forceinline__ LPair ld_pair(const LPair *p) {
    const float4 *q = (const float4 *)p;
    LPair r;
    r.a = q[0]; r.b = q[1]; r.c = q[2]; r.d = q[3];
    return r;
}

__device__ const float IDENTITY12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};

struct Pick { bool ok; uint32_t inst, prim; float prob, importance; };

// Importances of the TL light tree's top two levels for one shading point, shared by
// the many descents that start there (ReSTIR DI's candidates). Records 0 (the root's
// children), 1 and 2 (the children of the root's left and right child, when inner)
// each hold li, ri, ln, rn and the two child references; `memo` points at the thread's
// slot in LDS, MEMO_FIELDS words with a stride of one word per record field.
static constexpr int MEMO_FIELDS = 18;
__device__ __forceinline__ void memo_store(float *memo, int rec, const PairImp &q, const LPair &r) {
    float *m = memo + rec * 6;
    m[0] = q.li; m[1] = q.ri; m[2] = q.ln; m[3] = q.rn; m[4] = r.c.w; m[5] = r.d.w;
}
// fills the memo at (p, n); false when the TL root is a leaf (the walk then reads no
// TL record, and the plain walk is used)
__device__ __forceinline__ bool light_memo(const SceneView &sv, f3 p, f3 n, float *memo) {
    const vrgi_bvh_node tl0 = sv.tl[0];
    if (tl0.left == VRGI_INVALID) return false;
    const f3 zero = mk(0.0f, 0.0f, 0.0f);
    const LPair r0 = ld_pair(sv.tlpairs);
    memo_store(memo, 0, pair_importance_t(p, n, zero, r0, VRGI_EPS_RT), r0);
    for (int side = 0; side < 2; ++side) {
        const uint32_t ref = __float_as_uint(side ? r0.d.w : r0.c.w);
        if (ref & 0x80000000u) continue;  // a TL leaf: no record below it
        const LPair r1 = ld_pair(sv.tlpairs + ref);
        memo_store(memo, 1 + side, pair_importance_t(p, n, zero, r1, VRGI_EPS_RT), r1);
    }
    return true;
}

// traverseBvh, raytrace.rs:218-281. Each level reads one child record: the same
// boxes and powers as the two child nodes, so the same arithmetic.
__device__ Pick traverse_light(const SceneView &sv, f3 p, f3 n, uint32_t seed) {
    const vrgi_bvh_node tl0 = sv.tl[0];
    Pick res{false, 0, 0, 1.0f, 0.0f};
    if (tl0.left == VRGI_INVALID && tl0.right_or_prim == VRGI_INVALID) return res;  // dummy root
    float prob = 1.0f, imp = 0.0f;
    const float *m = IDENTITY12;
    bool lin = true;               // linear part of m is the identity: offset `off` instead of m
    f3 off = mk(0.0f, 0.0f, 0.0f);
    const LPair *pairs = sv.tlpairs;
    uint32_t inst = VRGI_INVALID;
    bool top = true;
    bool leaf = tl0.left == VRGI_INVALID;
    uint32_t idx = leaf ? tl0.right_or_prim : 0u;  // leaf: primitive / instance, else node index
    for (int guard = 0; guard < 4096; ++guard) {
        if (top && leaf) {  // TL leaf: continue in the instance's BL-BVH from its root
            inst = idx;
            const DevInst &I = sv.inst[inst];
            m = I.xf;
            lin = I.lin_id != 0u;
            off = mk(I.xf[9], I.xf[10], I.xf[11]);
            pairs = I.lpairs;
            top = false;
            const vrgi_bvh_node bl0 = I.light_bvh[0];
            if (imp == 0.0f) imp = node_importance(p, n, m, bl0, VRGI_EPS_RT);
            leaf = bl0.left == VRGI_INVALID;
            idx = leaf ? bl0.right_or_prim : 0u;
        }
        if (!top && leaf) {
            res.ok = true; res.inst = inst; res.prim = idx; res.prob = prob; res.importance = imp;
            return res;
        }
        const LPair r = ld_pair(pairs + idx);
        const PairImp q = lin ? pair_importance_t(p, n, off, r, VRGI_EPS_RT) : pair_importance(p, n, m, r, VRGI_EPS_RT);
        uint32_t ref;
        if (mm_finalizef(seed) < q.ln) { ref = __float_as_uint(r.c.w); prob *= q.ln; imp = q.li; }
        else { ref = __float_as_uint(r.d.w); prob *= q.rn; imp = q.ri; }
        leaf = (ref & 0x80000000u) != 0u;
        idx = ref & 0x7FFFFFFFu;
        seed = mm_combine(seed, 0);
    }
    return res;
}

// traverse_light with `memo` holding light_memo at the same p, n: the top two TL levels
// take the stored importances, which are the values the walk would compute there. A
// separate copy of the walk keeps the plain one (the shading's NEE pick) compiled as
// before: folding the memo into it as a template switch, or leaving ReSTIR DI without a
// call to it, changed its machine code and cost the 1080p frame 0.06 ms.
__device__ Pick traverse_light_memo(const SceneView &sv, f3 p, f3 n, uint32_t seed, const float *memo) {
    const vrgi_bvh_node tl0 = sv.tl[0];
    Pick res{false, 0, 0, 1.0f, 0.0f};
    if (tl0.left == VRGI_INVALID && tl0.right_or_prim == VRGI_INVALID) return res;  // dummy root
    float prob = 1.0f, imp = 0.0f;
    const float *m = IDENTITY12;
    bool lin = true;               // linear part of m is the identity: offset `off` instead of m
    f3 off = mk(0.0f, 0.0f, 0.0f);
    const LPair *pairs = sv.tlpairs;
    uint32_t inst = VRGI_INVALID;
    bool top = true;
    bool leaf = tl0.left == VRGI_INVALID;
    uint32_t idx = leaf ? tl0.right_or_prim : 0u;  // leaf: primitive / instance, else node index
    int mrec = 0;  // memo record of the next TL level (3: none)
    for (int guard = 0; guard < 4096; ++guard) {
        if (top && leaf) {  // TL leaf: continue in the instance's BL-BVH from its root
            inst = idx;
            const DevInst &I = sv.inst[inst];
            m = I.xf;
            lin = I.lin_id != 0u;
            off = mk(I.xf[9], I.xf[10], I.xf[11]);
            pairs = I.lpairs;
            top = false;
            const vrgi_bvh_node bl0 = I.light_bvh[0];
            if (imp == 0.0f) imp = node_importance(p, n, m, bl0, VRGI_EPS_RT);
            leaf = bl0.left == VRGI_INVALID;
            idx = leaf ? bl0.right_or_prim : 0u;
        }
        if (!top && leaf) {
            res.ok = true; res.inst = inst; res.prim = idx; res.prob = prob; res.importance = imp;
            return res;
        }
        uint32_t ref;
        PairImp q;
        uint32_t ref_l, ref_r;
        if (top && mrec < 3) {
            const float *mm = memo + mrec * 6;
            q.li = mm[0]; q.ri = mm[1]; q.ln = mm[2]; q.rn = mm[3];
            ref_l = __float_as_uint(mm[4]); ref_r = __float_as_uint(mm[5]);
        } else {
            const LPair r = ld_pair(pairs + idx);
            q = lin ? pair_importance_t(p, n, off, r, VRGI_EPS_RT) : pair_importance(p, n, m, r, VRGI_EPS_RT);
            ref_l = __float_as_uint(r.c.w); ref_r = __float_as_uint(r.d.w);
        }
        const bool go_left = mm_finalizef(seed) < q.ln;
        if (go_left) { ref = ref_l; prob *= q.ln; imp = q.li; }
        else { ref = ref_r; prob *= q.rn; imp = q.ri; }
        mrec = mrec == 0 ? (go_left ? 1 : 2) : 3;
        leaf = (ref & 0x80000000u) != 0u;
        idx = ref & 0x7FFFFFFFu;
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

// reverseTraverseBvh over the precomputed ancestor list of light lid: the records of
// all levels are independent loads; the product is formed leaf to root as in the walk
// above
__device__ float reverse_traverse_anc(const SceneView &sv, f3 p, f3 n, uint32_t lid) {
    const uint32_t len = sv.lanc_len[lid];
    const uint32_t in = sv.ltri_inst[lid];
    if (len == VRGI_INVALID) return reverse_traverse(sv, p, n, in, sv.ltri_node[lid]);
    const DevInst &I = sv.inst[in];
    // BL levels: the instance's offset when its linear part is the identity (the TL levels'
    // transform is the identity: offset 0), else the full transform
    const bool lin = I.lin_id != 0u;
    const f3 off = mk(I.xf[9], I.xf[10], I.xf[11]), zero = mk(0.0f, 0.0f, 0.0f);
    const LPair *blp = I.lpairs;
    const uint32_t *anc = sv.lanc + (size_t)lid * VRGI_LANC_MAX;
    float prob = 1.0f;
#pragma unroll 2
    for (uint32_t k = 0; k < len; ++k) {
        const uint32_t e = anc[k];
        const bool tl = (e & 0x40000000u) != 0u;
        const uint32_t ni = e & 0x3FFFFFFFu;
        const LPair r = ld_pair((tl ? sv.tlpairs : blp) + ni);
        const f3 la = mk(r.a.x, r.a.y, r.a.z), lb = mk(r.b.x, r.b.y, r.b.z);
        const f3 ra = mk(r.c.x, r.c.y, r.c.z), rb = mk(r.d.x, r.d.y, r.d.z);
        float li, ri;
        if (tl || lin) {
            const f3 t = tl ? zero : off;
            li = box_importance_t(p, n, t, la, lb, r.a.w, VRGI_EPS_NEE);
            ri = box_importance_t(p, n, t, ra, rb, r.b.w, VRGI_EPS_NEE);
        } else {
            li = box_importance(p, n, I.xf, la, lb, r.a.w, VRGI_EPS_NEE);
            ri = box_importance(p, n, I.xf, ra, rb, r.b.w, VRGI_EPS_NEE);
        }
        // only the walked side's share: ln = li / tot or rn = ri / tot (nee_pdf.rs:201-215)
        prob *= ((e & 0x80000000u) ? ri : li) / (li + ri);
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
    float prim_p = reverse_traverse_anc(sv, p, n, lid);
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

// texel of normalised texture coordinate c over n texels: nearest, clamp to edge
// (the reference's default sampler, interactive_rendering.rs:744)
__device__ __forceinline__ uint32_t texel_index(float c, uint32_t n) {
    const float f = floorf(c * (float)n);
    return !(f >= 0.0f) ? 0u : (f >= (float)(n - 1) ? n - 1 : (uint32_t)f);
}

__device__ __forceinline__ bool full_path(const BounceArgs &ba, uint32_t px) {
    return full_path_px(ba.train_seed, ba.train_frac, ba.fallback, px, ba.W, ba.tile, ba.tiles_x);
}

// one wave-sized slice of the bounce's rays (all 64 lanes call it: compaction ballots)
template <int STACK, bool COUNT, bool CACHE>
__device__ __forceinline__ void trace_shade_slice(const SceneView &sv, const BounceArgs &ba, uint32_t *stack,
                                                  uint32_t i, uint32_t count) {
    const uint32_t b = ba.bounce;
    bool active = i < count;
    uint32_t px = 0;
    f3 o = mk(0, 0, 0), d = mk(0, 0, 0);
    if (CACHE && active && b == 0 && ba.b0_part == 2)  // the full-path pixels run in part 1
        active = !full_path(ba, i);
    if (active) {
        if (b == 0) {
            // raygen.rs:84-111 (jitter is multiplied by 0.00)
            px = (CACHE && ba.b0_part == 1) ? ba.queue_in[i] : i;
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
    bool train = false;
    uint32_t cap = ba.B;  // bounces this pixel's path may take
    if (CACHE) {
        train = full_path(ba, px);
        cap = train ? ba.B : ba.cache_vertex + 1u;
        if (active && b == 0) st3(ba.d0 + (size_t)px * 3, d);
    }
    f3 alb = mk(0, 0, 0);
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
        out_o = (t0 * b3.x + t1 * b3.y) + t2 * b3.z;
        // the slot's reflectivity, emissivity and metallicity: its textures at the hit's uv
        // (raytrace.rs:437-481) or its constant material
        float4 m0, m1, m2;
        if (sv.tex) {
            const float2 q0 = sv.tri_uv[3 * (size_t)h.prim], q1 = sv.tri_uv[3 * (size_t)h.prim + 1],
                         q2 = sv.tri_uv[3 * (size_t)h.prim + 2];
            const float uvx = (q0.x * b3.x + q1.x * b3.y) + q2.x * b3.z;
            const float uvy = (q0.y * b3.x + q1.y * b3.y) + q2.y * b3.z;
            const size_t at = ((size_t)slot * 3 * sv.tex_h + texel_index(uvy, sv.tex_h)) * sv.tex_w +
                              texel_index(uvx, sv.tex_w);
            const size_t img = (size_t)sv.tex_h * sv.tex_w;
            m0 = sv.tex[at];
            m1 = sv.tex[at + img];
            m2 = sv.tex[at + 2 * img];
        } else {
            const vrgi_material &mt = sv.mats[slot];
            m0 = make_float4(mt.reflectivity[0], mt.reflectivity[1], mt.reflectivity[2], mt.reflectivity[3]);
            m1 = make_float4(mt.emissivity[0], mt.emissivity[1], mt.emissivity[2], 0.0f);
            m2 = make_float4(mt.metallicity[0], 0.0f, 0.0f, 0.0f);
        }
        refl = mk(m0.x, m0.y, m0.z);
        alb = refl;
        float alpha = m0.w;
        emis = (1000.0f * mk(m1.x, m1.y, m1.z)) * -dot3(d, ics.n);
        float metal = m2.x;
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
        cont = !zero3(out_d) && b + 1 < cap;
        if (!cont) ba.path_len[px] = b + 1;
        if (CACHE) st3(ba.albedo + bi * 3, alb);
    }
    // cache query of a non-training path at its last vertex
    if (CACHE && b == ba.cache_vertex) {
        const bool q = active && !train && h.prim != VRGI_INVALID;
        const uint64_t qb = __ballot(q);
        if (qb) {
            const uint32_t lane = threadIdx.x & 63u;
            uint32_t base = 0;
            if (lane == (uint32_t)(__ffsll((unsigned long long)qb) - 1))
                base = atomicAdd(ba.qrec_count, (uint32_t)__popcll(qb));
            base = __shfl(base, __ffsll((unsigned long long)qb) - 1, 64);
            if (q) {
                float4 *r = (float4 *)(ba.qrec + (size_t)(base + (uint32_t)__popcll(qb & ((1ull << lane) - 1ull))) * 16);
                r[0] = make_float4(out_o.x, out_o.y, out_o.z, d.x);
                r[1] = make_float4(d.y, d.z, nrm.x, nrm.y);
                r[2] = make_float4(nrm.z, alb.x, alb.y, alb.z);
                r[3] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(px * 16u + b));
            }
        }
    }
    // compaction of the rays alive at the next bounce: ballot + one atomic per wave
    uint64_t bal = __ballot(cont);
    if (bal) {
        uint32_t lane = threadIdx.x & 63u;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(ba.qcount + b + 1, (uint32_t)__popcll(bal));
        base = __shfl(base, 0, 64);
        if (cont) ba.queue_out[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = px;
    }
}

// one wave per block, one ray per lane; the queue's length is read on the device
template <int STACK, bool COUNT, bool CACHE>
__global__ __launch_bounds__(64) void k_trace_shade(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t count = ba.bounce == 0 ? (CACHE && ba.b0_part == 1 ? ba.qcount[0] : ba.P) : ba.qcount[ba.bounce];
    if (blockIdx.x * 64u >= count) return;
    trace_shade_slice<STACK, COUNT, CACHE>(sv, ba, stack, blockIdx.x * 64u + threadIdx.x, count);
}

// The same with a capped grid whose blocks stride over the queue: for queues known to be
// small (the training paths past the cache's query vertex), where a grid sized for every
// pixel would be mostly blocks that exit at once
template <int STACK, bool CACHE>
__global__ __launch_bounds__(64) void k_trace_shade_gs(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t count = ba.bounce == 0 ? (CACHE && ba.b0_part == 1 ? ba.qcount[0] : ba.P) : ba.qcount[ba.bounce];
    for (uint32_t base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u)
        trace_shade_slice<STACK, false, CACHE>(sv, ba, stack, base + threadIdx.x, count);
}

// true when the ray crosses some light triangle in [VRGI_EPS_NEE, tmax] (the first one
// found ends the walk); `stk` is the lane's LDS stack column
__device__ __forceinline__ bool light_crossed(const SceneView &sv, const RayPre &r, float tmax, uint32_t *stk) {
    int sp = 0;
    uint32_t ref = sv.n_ltri == 1 ? 0x80000000u | sv.lroot_prim : 0u;
    bool have = true;
    while (have) {
        if (ref & 0x80000000u) {
            const Tri t = sv.ltri[ref & 0x7FFFFFFFu];
            float tt, u, v;
            if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z), mk(t.v2.x, t.v2.y, t.v2.z),
                        VRGI_EPS_NEE, tmax, &tt, &u, &v))
                return true;
            have = false;
        } else {
            const Wide4 *w4 = sv.lwide4 + ref;
            const float4 lx = w4->lox, ly = w4->loy, lz = w4->loz, hx = w4->hix, hy = w4->hiy, hz = w4->hiz;
            const uint4 rf = w4->ref;
            uint32_t m = 0;
            m |= (rf.x != VRGI_INVALID && box_hit(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, VRGI_EPS_NEE, tmax)) ? 1u : 0u;
            m |= (rf.y != VRGI_INVALID && box_hit(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, VRGI_EPS_NEE, tmax)) ? 2u : 0u;
            m |= (rf.z != VRGI_INVALID && box_hit(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, VRGI_EPS_NEE, tmax)) ? 4u : 0u;
            m |= (rf.w != VRGI_INVALID && box_hit(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, VRGI_EPS_NEE, tmax)) ? 8u : 0u;
            have = m != 0u;
            if (m & 8u && m != 8u) stk[(sp++) * 64] = rf.w;
            if (m & 4u && (m & 3u)) stk[(sp++) * 64] = rf.z;
            if (m & 2u && (m & 1u)) stk[(sp++) * 64] = rf.y;
            ref = (m & 1u) ? rf.x : (m & 2u) ? rf.y : (m & 4u) ? rf.z : rf.w;
        }
        if (!have && sp > 0) { ref = stk[(--sp) * 64]; have = true; }
    }
    return false;
}

// No light triangle lies within distance 1 of p: p's cell of the light grid is clear (the
// grid marks every cell within 1 + one cell side of a light triangle's box, a full cell
// more than the distance, so the rounding of the cell index cannot miss a marked cell).
// A point outside the grid is farther than that from every light triangle.
__device__ __forceinline__ bool lgrid_clear(const SceneView &sv, f3 p) {
    const float x = (p.x - sv.lg_lo[0]) * sv.lg_inv, y = (p.y - sv.lg_lo[1]) * sv.lg_inv,
                z = (p.z - sv.lg_lo[2]) * sv.lg_inv;
    if (!(x == x && y == y && z == z)) return false;  // NaN: take the walk
    if (x < 0.0f || y < 0.0f || z < 0.0f || x >= (float)sv.lg_dim[0] || y >= (float)sv.lg_dim[1] ||
        z >= (float)sv.lg_dim[2])
        return true;
    const uint32_t ix = min((uint32_t)x, sv.lg_dim[0] - 1u), iy = min((uint32_t)y, sv.lg_dim[1] - 1u),
                   iz = min((uint32_t)z, sv.lg_dim[2] - 1u);
    return sv.lgrid[((size_t)iz * sv.lg_dim[1] + iy) * sv.lg_dim[0] + ix] == 0u;
}

// Whether the NEE pdf of a ray can reach the frame's outputs (the lazy mode). It enters
// the recurrence only as L_b = E_b + ((R_b * L_{b+1}) * (bsdf_pdf / q)) with
// q = np * w + (1 - w) / (2 pi), w in (0, 0.3]: when L_{b+1} is zero (either sign) and R_b
// finite, the product is a zero whose sign does not depend on np, unless q is 0, NaN or
// infinite. q is finite and positive whenever np is a finite non-negative number, and np
// is one when the direction is above the tangent plane (cs = n . d > 0, so d^2 / (cs A) is
// positive) and every crossed light triangle lies at least 0.01 above that plane: every
// ancestor box of such a triangle then has a corner at height >= 0.01 > 1e-4 plus any
// rounding of the corner test, so the walked side's importance is positive and no level
// of the reverse walk divides 0 by 0. The crossing at parameter t is t * cs above the
// plane, so it suffices that no light triangle is crossed for t <= 0.01 / cs.
// 0: the pdf cannot reach the outputs; 1: it can; 2: only if a light triangle is
// crossed for t <= *tmax (light_crossed decides)
__device__ __forceinline__ int nee_pdf_class(const SceneView &sv, const BounceArgs &ba, size_t bi, uint32_t px,
                                             f3 o, f3 d, float *tmax) {
    const f3 ln = ld3(ba.lrun + (size_t)px * 3);
    const f3 R = ld3(ba.refl + bi * 3);
    if (!zero3(ln) || !(isfinite(R.x) && isfinite(R.y) && isfinite(R.z))) return 1;
    const float cs = dot3(ld3(ba.normal + bi * 3), d);
    if (!(cs > 0.0f)) return 1;
    *tmax = gmin(0.01f / cs, 1000.0f);
    // d is a unit vector, so a crossing for t <= tmax <= 1 lies within distance 1 of o
    if (sv.lgrid && *tmax <= 1.0f && lgrid_clear(sv, o)) return 0;
    return 2;
}

// NEE pdf of bounce b (nee_pdf.rs:281-337) for the rays alive at bounce b+1
constexpr int NEE_DEFER = 4;  // crossed light triangles kept per lane before their pdfs are summed
// i: the ray's slot in the queue of bounce b + 1 (BY_PX: its pixel)
template <int STACK, bool COUNT_RAYS = true, bool BY_PX = false>
__device__ __forceinline__ void nee_pdf_ray(const SceneView &sv, const BounceArgs &ba, uint32_t *stack, uint32_t i,
                                            uint32_t count) {
    const uint32_t b = ba.bounce;  // pdf of bounce b, rays of bounce b + 1
    if (i >= count) return;
    const uint32_t px = BY_PX ? i : ba.queue_in[i];
    const size_t bi = (size_t)b * ba.P + px;
    float w = ba.mis[bi];
    f3 d = ld3(ba.dir + bi * 3);
    float np = 0.0f;
    const bool traced = !(len3(d) == 0.0f || w == 0.0f) && sv.n_ltri > 0;
    if (COUNT_RAYS) {
        uint64_t bal = __ballot(traced);
        if ((threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)bal) - 1))
            atomicAdd(ba.counters + 1, (unsigned long long)__popcll(bal));
    }
    if (traced) {
        f3 o = ld3(ba.origin + bi * 3);
        f3 n = ld3(ba.normal + bi * 3);
        RayPre r = ray_pre(o, d);
        // every light triangle crossed in [1e-4, 1000], left-to-right leaf order
        uint32_t *stk = stack + (threadIdx.x & 63u);
        int sp = 0;
        uint32_t ref;
        bool have;
        if (sv.n_ltri == 1) { ref = 0x80000000u | sv.lroot_prim; have = true; }
        else { ref = 0; have = true; }
        // Crossed triangles are kept (at most NEE_DEFER) and their pdfs summed after the
        // walk, in the order they were met: the long reverse walks then run together for
        // all lanes of the wave instead of one lane at a time inside the traversal loop.
        uint32_t hl0 = 0, hl1 = 0, hl2 = 0, hl3 = 0;
        float hu0 = 0.0f, hu1 = 0.0f, hu2 = 0.0f, hu3 = 0.0f, hv0 = 0.0f, hv1 = 0.0f, hv2 = 0.0f, hv3 = 0.0f;
        int nh = 0;
        auto flush = [&]() {
#pragma unroll 1
            for (int j = 0; j < nh; ++j) {
                const uint32_t lid = j == 0 ? hl0 : j == 1 ? hl1 : j == 2 ? hl2 : hl3;
                const float u = j == 0 ? hu0 : j == 1 ? hu1 : j == 2 ? hu2 : hu3;
                const float v = j == 0 ? hv0 : j == 1 ? hv1 : j == 2 ? hv2 : hv3;
                np += nee_pdf_one(sv, o, n, d, lid, u, v);
            }
            nh = 0;
        };
        while (have) {
            if (ref & 0x80000000u) {
                uint32_t lid = ref & 0x7FFFFFFFu;
                Tri t = sv.ltri[lid];
                float tt, u, v;
                if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z),
                            mk(t.v2.x, t.v2.y, t.v2.z), VRGI_EPS_NEE, 1000.0f, &tt, &u, &v)) {
                    if (nh == NEE_DEFER) flush();
                    if (nh == 0) { hl0 = lid; hu0 = u; hv0 = v; }
                    else if (nh == 1) { hl1 = lid; hu1 = u; hv1 = v; }
                    else if (nh == 2) { hl2 = lid; hu2 = u; hv2 = v; }
                    else { hl3 = lid; hu3 = u; hv3 = v; }
                    ++nh;
                }
                have = false;
            } else {
                // 4-ary node: descendants in leaf order; the first one hit is next, the
                // others are pushed last-first, so leaves are still met left to right
                const Wide4 *w4 = sv.lwide4 + ref;
                const float4 lx = w4->lox, ly = w4->loy, lz = w4->loz, hx = w4->hix, hy = w4->hiy, hz = w4->hiz;
                const uint4 rf = w4->ref;
                uint32_t m = 0;
                m |= (rf.x != VRGI_INVALID && box_hit(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, VRGI_EPS_NEE, 1000.0f)) ? 1u : 0u;
                m |= (rf.y != VRGI_INVALID && box_hit(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, VRGI_EPS_NEE, 1000.0f)) ? 2u : 0u;
                m |= (rf.z != VRGI_INVALID && box_hit(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, VRGI_EPS_NEE, 1000.0f)) ? 4u : 0u;
                m |= (rf.w != VRGI_INVALID && box_hit(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, VRGI_EPS_NEE, 1000.0f)) ? 8u : 0u;
                have = m != 0u;
                if (m & 8u && m != 8u) stk[(sp++) * 64] = rf.w;
                if (m & 4u && (m & 3u)) stk[(sp++) * 64] = rf.z;
                if (m & 2u && (m & 1u)) stk[(sp++) * 64] = rf.y;
                ref = (m & 1u) ? rf.x : (m & 2u) ? rf.y : (m & 4u) ? rf.z : rf.w;
            }
            if (!have && sp > 0) { ref = stk[(--sp) * 64]; have = true; }
        }
        flush();
    }
    ba.nee[bi] = np;
}
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_pdf(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    nee_pdf_ray<STACK>(sv, ba, stack, blockIdx.x * 64u + threadIdx.x, ba.qcount[ba.bounce + 1]);
}
// slot in a list for every lane with `take`: one atomic per wave (all 64 lanes call it)
__device__ __forceinline__ void list_append(uint32_t *list, uint32_t *count, bool take, uint32_t v) {
    const uint64_t bal = __ballot(take);
    if (!bal) return;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(bal));
    base = __shfl(base, 0, 64);
    if (take) list[base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = v;
}
// Lazy NEE pdf, pass 1: per ray of the bounce's queue (slot i), 0 for the rays whose pdf
// cannot reach the outputs (or that trace none); the rays whose pdf can are appended to
// `list` for the full pdf, those for which it depends on a light triangle crossed near the
// origin to `mlist` (queue slots). No walk runs here, so the pass only streams the rays'
// state; the appends take one counter update per block of NEE_SELECT_THREADS rays, since
// updates of one address by every wave of so short a pass queue up behind each other.
constexpr int NEE_SELECT_THREADS = 1024;
__global__ __launch_bounds__(NEE_SELECT_THREADS) void k_nee_select(const SceneView sv, const BounceArgs ba,
                                                                   uint32_t *list, uint32_t *list_count,
                                                                   uint32_t *mlist, uint32_t *mcount) {
    constexpr int WAVES = NEE_SELECT_THREADS / 64;
    __shared__ uint32_t s_cnt[2][WAVES];
    __shared__ uint32_t s_base[2];
    const uint32_t b = ba.bounce;
    const uint32_t count = ba.qcount[b + 1];
    if (blockIdx.x * NEE_SELECT_THREADS >= count) return;  // uniform over the block
    const uint32_t i = blockIdx.x * NEE_SELECT_THREADS + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    int cls = 0;
    if (i < count) {
        const uint32_t px = ba.queue_in[i];
        const size_t bi = (size_t)b * ba.P + px;
        const float wm = ba.mis[bi];
        const f3 d = ld3(ba.dir + bi * 3);
        float tmax;
        if (!(len3(d) == 0.0f || wm == 0.0f) && sv.n_ltri > 0)
            cls = nee_pdf_class(sv, ba, bi, px, ld3(ba.origin + bi * 3), d, &tmax);
        if (cls == 0) ba.nee[bi] = 0.0f;
    }
    const uint64_t b1 = __ballot(cls == 1), b2 = __ballot(cls == 2);
    if (lane == 0) {
        s_cnt[0][w] = (uint32_t)__popcll(b1);
        s_cnt[1][w] = (uint32_t)__popcll(b2);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        for (int k = 0; k < WAVES; ++k) t += s_cnt[threadIdx.x][k];
        s_base[threadIdx.x] = t ? atomicAdd(threadIdx.x ? mcount : list_count, t) : 0u;
    }
    __syncthreads();
    if (cls != 0) {
        const uint32_t q = cls - 1;
        uint32_t off = s_base[q];
        for (uint32_t k = 0; k < w; ++k) off += s_cnt[q][k];
        const uint64_t m = q ? b2 : b1;
        (q ? mlist : list)[off + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
    }
}
// pass 1b: the rays of `mlist`, whose pdf matters only if a light triangle is crossed for
// t <= 0.01 / cos (a striding pass over the list): those go to `list`, the others get 0
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_near(const SceneView sv, const BounceArgs ba, const uint32_t *mlist,
                                                 const uint32_t *mcount, uint32_t *list, uint32_t *list_count) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t b = ba.bounce;
    const uint32_t count = ba.qcount[b + 1];
    const uint32_t n = min(*mcount, count);  // each queued ray is listed at most once
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t k = base + threadIdx.x;
        bool need = false;
        uint32_t i = 0;
        if (k < n) {
            i = mlist[k];
            const uint32_t px = ba.queue_in[i];
            const size_t bi = (size_t)b * ba.P + px;
            const f3 d = ld3(ba.dir + bi * 3);
            const float cs = dot3(ld3(ba.normal + bi * 3), d);
            need = light_crossed(sv, ray_pre(ld3(ba.origin + bi * 3), d), gmin(0.01f / cs, 1000.0f),
                                 stack + (threadIdx.x & 63u));
            if (!need) ba.nee[bi] = 0.0f;
        }
        list_append(list, list_count, need, i);
    }
}
// pass 2: the full pdf of the listed rays (a capped grid striding over the list), each
// ray exactly as nee_pdf_ray computes it
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_listed(const SceneView sv, const BounceArgs ba, const uint32_t *list,
                                                   const uint32_t *list_count) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t n = min(*list_count, ba.qcount[ba.bounce + 1]);  // each queued ray is listed at most once
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t k = base + threadIdx.x;
        nee_pdf_ray<STACK>(sv, ba, stack, k < n ? list[k] : ba.qcount[ba.bounce + 1], ba.qcount[ba.bounce + 1]);
    }
}
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_pdf_gs(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t count = ba.qcount[ba.bounce + 1];
    for (uint32_t base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u)
        nee_pdf_ray<STACK>(sv, ba, stack, base + threadIdx.x, count);
}

// The NEE pdf in passes: the enumeration keeps no reverse-walk state (fewer registers,
// more waves), the walks run in their own pass. Per ray the pdfs are summed in the order
// the triangles were crossed, as in nee_pdf_ray; a ray crossing more than VRGI_NEE_RECS
// goes to an overflow list and is done again whole by nee_pdf_ray. The 4-ary walk is
// written out again here rather than shared: with a callback the deferred triangles of
// nee_pdf_ray went to scratch memory, and an iterator form made it 0.7 ms per frame slower.
template <int STACK>
__device__ __forceinline__ void nee_enum_ray(const SceneView &sv, const BounceArgs &ba, uint32_t *stack, uint32_t i,
                                             uint32_t count) {
    const uint32_t b = ba.bounce;
    if (i >= count) return;
    const uint32_t px = ba.queue_in[i];
    const size_t bi = (size_t)b * ba.P + px;
    const float w = ba.mis[bi];
    const f3 d = ld3(ba.dir + bi * 3);
    const bool traced = !(len3(d) == 0.0f || w == 0.0f) && sv.n_ltri > 0;
    const uint64_t bal = __ballot(traced);
    if ((threadIdx.x & 63) == (uint32_t)(__ffsll((unsigned long long)bal) - 1))
        atomicAdd(ba.counters + 1, (unsigned long long)__popcll(bal));
    uint32_t nh = 0;
    if (traced) {
        const f3 o = ld3(ba.origin + bi * 3);
        const RayPre r = ray_pre(o, d);
        uint32_t *stk = stack + (threadIdx.x & 63u);
        int sp = 0;
        uint32_t ref = sv.n_ltri == 1 ? 0x80000000u | sv.lroot_prim : 0u;
        bool have = true;
        while (have) {
            if (ref & 0x80000000u) {
                const uint32_t lid = ref & 0x7FFFFFFFu;
                const Tri t = sv.ltri[lid];
                float tt, u, v;
                if (tri_hit(r, mk(t.v0.x, t.v0.y, t.v0.z), mk(t.v1.x, t.v1.y, t.v1.z), mk(t.v2.x, t.v2.y, t.v2.z),
                            VRGI_EPS_NEE, 1000.0f, &tt, &u, &v)) {
                    if (nh < VRGI_NEE_RECS) {
                        ba.nee_lid[(size_t)nh * ba.P + i] = lid;
                        ba.nee_uv[(size_t)nh * ba.P + i] = make_float2(u, v);
                    }
                    ++nh;
                    if (nh > VRGI_NEE_RECS) sp = 0;  // overflow: the whole ray is done again
                }
                have = false;
            } else {  // 4-ary node, descendants in leaf order (as in nee_pdf_ray)
                const Wide4 *w4 = sv.lwide4 + ref;
                const float4 lx = w4->lox, ly = w4->loy, lz = w4->loz, hx = w4->hix, hy = w4->hiy, hz = w4->hiz;
                const uint4 rf = w4->ref;
                uint32_t m = 0;
                m |= (rf.x != VRGI_INVALID && box_hit(r, lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, VRGI_EPS_NEE, 1000.0f)) ? 1u : 0u;
                m |= (rf.y != VRGI_INVALID && box_hit(r, lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, VRGI_EPS_NEE, 1000.0f)) ? 2u : 0u;
                m |= (rf.z != VRGI_INVALID && box_hit(r, lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, VRGI_EPS_NEE, 1000.0f)) ? 4u : 0u;
                m |= (rf.w != VRGI_INVALID && box_hit(r, lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, VRGI_EPS_NEE, 1000.0f)) ? 8u : 0u;
                have = m != 0u;
                if (m & 8u && m != 8u) stk[(sp++) * 64] = rf.w;
                if (m & 4u && (m & 3u)) stk[(sp++) * 64] = rf.z;
                if (m & 2u && (m & 1u)) stk[(sp++) * 64] = rf.y;
                ref = (m & 1u) ? rf.x : (m & 2u) ? rf.y : (m & 4u) ? rf.z : rf.w;
            }
            if (!have && sp > 0) { ref = stk[(--sp) * 64]; have = true; }
        }
    }
    if (nh > VRGI_NEE_RECS) {
        nh = VRGI_NEE_RECS + 1;
        ba.nee_over[atomicAdd(ba.nee_over_count, 1u)] = i;
    }
    ba.nee_cnt[i] = nh;
    if (nh == 0) ba.nee[bi] = 0.0f;
}
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_enum(const SceneView sv, const BounceArgs ba, uint32_t grid_stride) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t count = ba.qcount[ba.bounce + 1];
    if (!grid_stride) {
        nee_enum_ray<STACK>(sv, ba, stack, blockIdx.x * 64u + threadIdx.x, count);
        return;
    }
    for (uint32_t base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u)
        nee_enum_ray<STACK>(sv, ba, stack, base + threadIdx.x, count);
}
// the reverse walks of the kept triangles, summed in crossing order (nee_pdf.rs:281-337)
__device__ __forceinline__ void nee_walk_ray(const SceneView &sv, const BounceArgs &ba, uint32_t i, uint32_t count) {
    if (i >= count) return;
    const uint32_t nh = ba.nee_cnt[i];
    if (nh == 0 || nh > VRGI_NEE_RECS) return;
    const uint32_t px = ba.queue_in[i];
    const size_t bi = (size_t)ba.bounce * ba.P + px;
    const f3 o = ld3(ba.origin + bi * 3), n = ld3(ba.normal + bi * 3), d = ld3(ba.dir + bi * 3);
    float np = 0.0f;
#pragma unroll 1
    for (uint32_t j = 0; j < nh; ++j) {
        const uint32_t lid = ba.nee_lid[(size_t)j * ba.P + i];
        const float2 uv = ba.nee_uv[(size_t)j * ba.P + i];
        np += nee_pdf_one(sv, o, n, d, lid, uv.x, uv.y);
    }
    ba.nee[bi] = np;
}
__global__ __launch_bounds__(64) void k_nee_walk(const SceneView sv, const BounceArgs ba, uint32_t grid_stride) {
    const uint32_t count = ba.qcount[ba.bounce + 1];
    if (!grid_stride) {
        nee_walk_ray(sv, ba, blockIdx.x * 64u + threadIdx.x, count);
        return;
    }
    for (uint32_t base = blockIdx.x * 64u; base < count; base += gridDim.x * 64u)
        nee_walk_ray(sv, ba, base + threadIdx.x, count);
}
// the rays on the overflow list, whole (a small grid striding over the list)
template <int STACK>
__global__ __launch_bounds__(64) void k_nee_overflow(const SceneView sv, const BounceArgs ba) {
    __shared__ uint32_t stack[STACK * 64];
    const uint32_t count = ba.qcount[ba.bounce + 1];
    const uint32_t n = min(*ba.nee_over_count, count);  // each queued ray is listed at most once
    for (uint32_t base = blockIdx.x * 64u; base < n; base += gridDim.x * 64u) {
        const uint32_t k = base + threadIdx.x;
        nee_pdf_ray<STACK, false>(sv, ba, stack, k < n ? ba.nee_over[k] : count, count);
    }
}

// ---------------------------------------------------------------- ReSTIR DI (benchmark mode)
// Not in the reference (its ReSTIR is GI); assembled from its primitives: candidates
// from traverseBvh (raytrace.rs:218-281) and triangleSample (:305-311), reservoirs
// merged as updateReservoir / mergeReservoir (restir_*_resampling.rs), neighbours
// filtered by geometricallySimilar (restir_spatial_resampling.rs:277-288).
// Target p_hat = luminance(f * Le * cos_x * cos_y / d^2), f = albedo / pi,
// Le = 1000 * emissivity (raytrace.rs:477); source pdf = pick probability / area.
struct DiRes {
    uint32_t light;
    f3 y;
    float wsum, W, ph;
    uint32_t M;
};
__device__ __forceinline__ float di_lum(f3 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }

__device__ __forceinline__ f3 di_contrib(const SceneView &sv, f3 x, f3 n, f3 fr, uint32_t lg, f3 y) {
    Tri lt = sv.tri[lg];
    f3 l0 = mk(lt.v0.x, lt.v0.y, lt.v0.z), l1 = mk(lt.v1.x, lt.v1.y, lt.v1.z), l2 = mk(lt.v2.x, lt.v2.y, lt.v2.z);
    f3 nl = norm3(cross3(l1 - l0, l2 - l0));
    f3 w = y - x;
    float d2 = dot3(w, w);
    float dist = sqrtf(d2);
    f3 wn = w / dist;
    float cx = gmax(0.0f, dot3(n, wn));
    float cy = gmax(0.0f, -dot3(nl, wn));
    const vrgi_material &m = sv.mats[__float_as_uint(lt.v0.w)];
    f3 le = 1000.0f * mk(m.emissivity[0], m.emissivity[1], m.emissivity[2]);
    return (fr * le) * ((cx * cy) / d2);
}
__device__ __forceinline__ void di_update(uint32_t seed, DiRes &r, uint32_t light, f3 y, float w, float ph) {
    r.wsum += w;
    r.M += 1;
    if (float_construct(seed) < w / r.wsum) { r.light = light; r.y = y; r.ph = ph; }
}
__device__ __forceinline__ void di_store(float *p, const DiRes &r) {
    float4 *q = (float4 *)p;
    q[0] = make_float4(__uint_as_float(r.light), r.y.x, r.y.y, r.y.z);
    q[1] = make_float4(r.wsum, r.W, r.ph, __uint_as_float(r.M));
}
__device__ __forceinline__ DiRes di_load(const float *p) {
    const float4 *q = (const float4 *)p;
    float4 a = q[0], b = q[1];
    DiRes r;
    r.light = __float_as_uint(a.x); r.y = mk(a.y, a.z, a.w);
    r.wsum = b.x; r.W = b.y; r.ph = b.z; r.M = __float_as_uint(b.w);
    return r;
}
__device__ __forceinline__ f3 di_fr(const SceneView &sv, uint32_t hit) {
    const vrgi_material &m = sv.mats[__float_as_uint(sv.tri[hit].v0.w)];
    return mk(m.reflectivity[0], m.reflectivity[1], m.reflectivity[2]) / VRGI_PI_F;
}

// initial candidates + temporal merge with the previous frame (M capped at 20x)
__device__ __forceinline__ void di_initial_px(const SceneView &sv, const DiArgs &da, float *memo, uint32_t id) {
    if (id >= da.P) return;
    DiRes r;
    r.light = VRGI_INVALID; r.y = mk(0, 0, 0); r.wsum = 0.0f; r.W = 0.0f; r.ph = 0.0f; r.M = 0;
    const uint32_t hp = da.hit[id];
    if (hp != VRGI_INVALID) {
        f3 x = ld3(da.origin + (size_t)id * 3), n = ld3(da.normal + (size_t)id * 3);
        f3 fr = di_fr(sv, hp);
        const uint32_t pd = mm_combine(da.dseed, id);
        const bool shared = light_memo(sv, x, n, memo);
        for (uint32_t k = 0; k < da.candidates; ++k) {
            uint32_t sk = mm_combine(pd, k);
            Pick pk = shared ? traverse_light_memo(sv, x, n, mm_combine(sk, 0), memo)
                             : traverse_light(sv, x, n, mm_combine(sk, 0));
            float w = 0.0f, ph = 0.0f;
            uint32_t lg = VRGI_INVALID;
            f3 y = mk(0, 0, 0);
            if (pk.ok) {
                lg = sv.inst[pk.inst].tri_off + pk.prim;
                Tri lt = sv.tri[lg];
                f3 l0 = mk(lt.v0.x, lt.v0.y, lt.v0.z), l1 = mk(lt.v1.x, lt.v1.y, lt.v1.z), l2 = mk(lt.v2.x, lt.v2.y, lt.v2.z);
                float ux = mm_finalizef(mm_combine(sk, 1)), uy = mm_finalizef(mm_combine(sk, 2));
                if (ux + uy > 1.0f) { ux = 1.0f - ux; uy = 1.0f - uy; }
                f3 bb = mk((1.0f - ux) - uy, ux, uy);
                y = (bb.x * l0 + bb.y * l1) + bb.z * l2;
                float area = 0.5f * len3(cross3(l1 - l0, l2 - l0));
                float psrc = pk.prob / area;
                ph = di_lum(di_contrib(sv, x, n, fr, lg, y));
                w = psrc > 0.0f ? ph / psrc : 0.0f;
            }
            di_update(mm_finalize(mm_combine(sk, 3)), r, lg, y, w, ph);
        }
        const float *h = da.hist_prev + (size_t)id * 8;
        uint32_t hl = __float_as_uint(h[0]);
        if (hl != VRGI_INVALID && hl < sv.n_tris) {
            uint32_t hm = min(__float_as_uint(h[5]), 20u * da.candidates);
            f3 hy = mk(h[1], h[2], h[3]);
            float ph = di_lum(di_contrib(sv, x, n, fr, hl, hy));
            uint32_t m0 = r.M;
            di_update(mm_finalize(mm_combine(pd, 0x10000u)), r, hl, hy, (ph * h[4]) * (float)hm, ph);
            r.M = m0 + hm;
        }
        r.W = r.ph > 0.0f ? r.wsum / ((float)r.M * r.ph) : 0.0f;
    }
    di_store(da.T + (size_t)id * 8, r);
}
__global__ __launch_bounds__(64) void k_di_initial(const SceneView sv, const DiArgs da) {
    __shared__ float s_memo[64 * (MEMO_FIELDS + 1)];
    di_initial_px(sv, da, s_memo + threadIdx.x * (MEMO_FIELDS + 1), blockIdx.x * 64 + threadIdx.x);
}

// spatial merge, shadow ray, final radiance, history for the next frame
template <int STACK>
__device__ __forceinline__ void di_final_px(const SceneView &sv, const DiArgs &da, uint32_t *stack, uint32_t id) {
    if (id >= da.P) return;
    float *hc = da.hist_cur + (size_t)id * 8;
    const f3 E0 = ld3(da.emis + (size_t)id * 3);
    const uint32_t hp = da.hit[id];
    if (hp == VRGI_INVALID) {
        float4 *q = (float4 *)hc;
        q[0] = make_float4(__uint_as_float(VRGI_INVALID), 0.0f, 0.0f, 0.0f);
        q[1] = make_float4(0.0f, __uint_as_float(0u), 0.0f, 0.0f);
        st3(da.restir + (size_t)id * 3, E0);
        da.di_light[id] = VRGI_INVALID;
        ((float4 *)da.di_weights)[id] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const uint32_t pd = mm_combine(da.dseed, id);
    const uint32_t qx = id % da.W, qy = id / da.W;
    const float cos25 = 0.906307787036650f;
    DiRes r = di_load(da.T + (size_t)id * 8);
    f3 x = ld3(da.origin + (size_t)id * 3), n = ld3(da.normal + (size_t)id * 3);
    f3 fr = di_fr(sv, hp);
    for (uint32_t it = 0; it < da.spatial_iters; ++it) {
        uint32_t is = mm_combine(mm_combine(pd, 0x20000u), it);
        float jx = da.radius * (mm_finalizef(mm_combine(is, 0)) - 0.5f);
        float jy = da.radius * (mm_finalizef(mm_combine(is, 1)) - 0.5f);
        int nx = (int)(((float)qx + 0.5f) + jx), ny = (int)(((float)qy + 0.5f) + jy);
        nx = min(max(nx, 0), (int)da.W - 1);
        ny = min(max(ny, 0), (int)da.H - 1);
        size_t q = (size_t)ny * da.W + nx;
        if (da.hit[q] == VRGI_INVALID) continue;
        f3 nq = ld3(da.normal + q * 3);
        if (zero3(n) || zero3(nq) || dot3(n, nq) < cos25) continue;
        DiRes rq = di_load(da.T + q * 8);
        if (rq.light == VRGI_INVALID) continue;
        float ph = di_lum(di_contrib(sv, x, n, fr, rq.light, rq.y));
        uint32_t m0 = r.M;
        di_update(mm_finalize(mm_combine(is, 2)), r, rq.light, rq.y, (ph * rq.W) * (float)rq.M, ph);
        r.M = m0 + rq.M;
    }
    r.W = r.ph > 0.0f ? r.wsum / ((float)r.M * r.ph) : 0.0f;
    f3 L = E0;
    if (r.light != VRGI_INVALID && r.ph > 0.0f) {
        f3 w = r.y - x;
        float dist = sqrtf(dot3(w, w));
        RayPre rp = ray_pre(x, w / dist);
        Hit sh = closest_hit<STACK>(sv, rp, VRGI_EPS_RT, dist - 0.002f, stack);
        if (sh.prim == VRGI_INVALID) L = E0 + di_contrib(sv, x, n, fr, r.light, r.y) * r.W;
    }
    st3(da.restir + (size_t)id * 3, L);
    da.di_light[id] = r.light;
    ((float4 *)da.di_weights)[id] = make_float4(r.wsum, r.W, (float)r.M, r.ph);
    float4 *q = (float4 *)hc;
    q[0] = make_float4(__uint_as_float(r.light), r.y.x, r.y.y, r.y.z);
    q[1] = make_float4(r.W, __uint_as_float(r.M), 0.0f, 0.0f);
}
template <int STACK>
__global__ __launch_bounds__(64) void k_di_final(const SceneView sv, const DiArgs da) {
    __shared__ uint32_t stack[STACK * 64];
    di_final_px<STACK>(sv, da, stack, blockIdx.x * 64 + threadIdx.x);
}

// ---------------------------------------------------------------- batched scenes
// The per-scene kernels above, over many scenes per launch: a block finds its scene by
// binary search over the block prefix; WPB waves per block, one LDS stack per wave.
template <int STACK, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_trace_shade_b(const SceneView *__restrict__ svs,
                                                            const BounceArgs *__restrict__ bas,
                                                            const uint32_t *__restrict__ blk, uint32_t n,
                                                            uint32_t stride, uint32_t b) {
    __shared__ uint32_t stack[WPB * STACK * 64];
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    const BounceArgs &ba = bas[(size_t)s * stride + b];
    const uint32_t count = ba.bounce == 0 ? ba.P : ba.qcount[ba.bounce];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t first = ((blockIdx.x - blk[s]) * WPB + w) * 64u;
    if (first >= count) return;  // the whole wave
    trace_shade_slice<STACK, false, false>(svs[s], ba, stack + w * STACK * 64, first + (threadIdx.x & 63u), count);
}
template <int STACK, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_nee_pdf_b(const SceneView *__restrict__ svs,
                                                        const BounceArgs *__restrict__ bas,
                                                        const uint32_t *__restrict__ blk, uint32_t n,
                                                        uint32_t stride, uint32_t b) {
    __shared__ uint32_t stack[WPB * STACK * 64];
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    const BounceArgs &ba = bas[(size_t)s * stride + b];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t i = ((blockIdx.x - blk[s]) * WPB + w) * 64u + (threadIdx.x & 63u);
    nee_pdf_ray<STACK>(svs[s], ba, stack + w * STACK * 64, i, ba.qcount[ba.bounce + 1]);
}
template <int PB>
__global__ __launch_bounds__(PB) void k_di_initial_b(const SceneView *__restrict__ svs, const DiArgs *__restrict__ das,
                                                     const uint32_t *__restrict__ blk, uint32_t n) {
    __shared__ float s_memo[PB * (MEMO_FIELDS + 1)];
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    di_initial_px(svs[s], das[s], s_memo + threadIdx.x * (MEMO_FIELDS + 1), (blockIdx.x - blk[s]) * PB + threadIdx.x);
}
template <int STACK, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_di_final_b(const SceneView *__restrict__ svs,
                                                         const DiArgs *__restrict__ das,
                                                         const uint32_t *__restrict__ blk, uint32_t n) {
    __shared__ uint32_t stack[WPB * STACK * 64];
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    const uint32_t w = threadIdx.x >> 6;
    di_final_px<STACK>(svs[s], das[s], stack + w * STACK * 64, (blockIdx.x - blk[s]) * 64u * WPB + threadIdx.x);
}
// queue counts of every scene to zero (a kernel, so that the pass can be graph-captured)
__global__ void k_batch_reset(const BounceArgs *__restrict__ bas, uint32_t n, uint32_t stride, uint32_t words) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * words) return;
    bas[(size_t)(i / words) * stride].qcount[i % words] = 0u;
}
// per scene: closest-hit rays and frames (k_frame_stats)
__global__ void k_batch_stats(const BounceArgs *__restrict__ bas, uint32_t n, uint32_t stride, uint32_t B) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const BounceArgs &ba = bas[(size_t)s * stride];
    unsigned long long r = ba.P;
    for (uint32_t b = 1; b < B; ++b) r += ba.qcount[b];
    ba.counters[0] += r;
    ba.counters[2] += 1;
}

template <int STACK, int WPB>
static void batch_rays(hipStream_t s, const BatchLaunch &L) {
    for (uint32_t b = 0; b < L.B; ++b)
        k_trace_shade_b<STACK, WPB><<<L.ray_blocks, 64 * WPB, 0, s>>>(L.svs, L.tba, L.blk_ray, L.n, L.B, b);
}
template <int STACK, int WPB>
static void batch_nee(hipStream_t s, const BatchLaunch &L) {
    for (uint32_t b = 0; b + 1 < L.B; ++b)
        k_nee_pdf_b<STACK, WPB><<<L.ray_blocks, 64 * WPB, 0, s>>>(L.svs, L.nba, L.blk_ray, L.n, L.B, b);
}
template <int STACK, int WPB>
static void batch_di_final(hipStream_t s, const BatchLaunch &L) {
    k_di_final_b<STACK, WPB><<<L.ray_blocks, 64 * WPB, 0, s>>>(L.svs, L.das, L.blk_ray, L.n);
}
// STACK x WPB dispatch; an LDS stack of more than 64 entries runs one wave per block
#define VRGI_BATCH_DISPATCH(FN)                                                              \
    static void FN##_any(hipStream_t s, const BatchLaunch &L, uint32_t need) {               \
        const uint32_t w = L.wpb;                                                              \
        if (need + 2 <= 32) {                                                                  \
            if (w >= 4) FN<32, 4>(s, L); else if (w == 2) FN<32, 2>(s, L); else FN<32, 1>(s, L); \
        } else if (need + 2 <= 48) {                                                           \
            if (w >= 4) FN<48, 4>(s, L); else if (w == 2) FN<48, 2>(s, L); else FN<48, 1>(s, L); \
        } else if (need + 2 <= 64) {                                                           \
            if (w >= 4) FN<64, 4>(s, L); else if (w == 2) FN<64, 2>(s, L); else FN<64, 1>(s, L); \
        } else {                                                                               \
            FN<VRGI_MAX_STACK, 1>(s, L);                                                       \
        }                                                                                      \
    }
VRGI_BATCH_DISPATCH(batch_rays)
VRGI_BATCH_DISPATCH(batch_nee)
VRGI_BATCH_DISPATCH(batch_di_final)

// one spp pass of every scene: reset, B bounces, NEE pdfs, resolve, postprocess, stats
void launch_batch_pass(hipStream_t s, const BatchLaunch &L) {
    if (!L.n) return;
    const uint32_t words = L.B + 2;
    k_batch_reset<<<(L.n * words + 255) / 256, 256, 0, s>>>(L.tba, L.n, L.B, words);
    batch_rays_any(s, L, L.gneed);
    if (L.nee && L.B > 1) batch_nee_any(s, L, L.lneed);
    launch_batch_pixels_gi(s, L);
    if (L.di) {
        if (L.pb == 64) k_di_initial_b<64><<<L.px_blocks, 64, 0, s>>>(L.svs, L.das, L.blk_px, L.n);
        else if (L.pb == 128) k_di_initial_b<128><<<L.px_blocks, 128, 0, s>>>(L.svs, L.das, L.blk_px, L.n);
        else k_di_initial_b<256><<<L.px_blocks, 256, 0, s>>>(L.svs, L.das, L.blk_px, L.n);
        batch_di_final_any(s, L, L.gneed);
    } else {
        launch_batch_pixels_spatial(s, L);
    }
    launch_batch_pixels_post(s, L);
    k_batch_stats<<<(L.n + 63) / 64, 64, 0, s>>>(L.tba, L.n, L.B, L.B);
}

// ---------------------------------------------------------------- host launchers
// Traversal stack size per lane (LDS): a lane never holds more than depth - 1 pending
// nodes, so the smallest size above the tree depth is chosen; smaller stacks raise
// occupancy (64-lane blocks of 8 / 16 / 40 KB).
template <class F>
static void with_stack(uint32_t depth, F f) {  // LDS stack entries per lane for a walk's need
    if (depth + 2 <= 32) f(std::integral_constant<int, 32>());
    else if (depth + 2 <= 48) f(std::integral_constant<int, 48>());
    else if (depth + 2 <= 64) f(std::integral_constant<int, 64>());
    else f(std::integral_constant<int, VRGI_MAX_STACK>());
}

void launch_di(hipStream_t s, const SceneView &sv, const DiArgs &da) {
    k_di_initial<<<(da.P + 63) / 64, 64, 0, s>>>(sv, da);
    with_stack(geom_stack_need(sv.depth), [&](auto S) { k_di_final<decltype(S)::value><<<(da.P + 63) / 64, 64, 0, s>>>(sv, da); });
}
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
    if (n_nodes > 1) k_wide<<<((n_nodes - 1) / 2 + 255) / 256, 256, 0, s>>>(nodes, n_nodes, wide);
}
void launch_trace_closest(hipStream_t s, const SceneView &sv, const float *o, const float *d,
                          uint32_t n, float tmin, float tmax, uint32_t *hit, float *t, float *uv) {
    if (!n) return;
    with_stack(geom_stack_need(sv.depth), [&](auto S) {
        k_trace_closest<decltype(S)::value><<<(n + 63) / 64, 64, 0, s>>>(sv, o, d, n, tmin, tmax, hit, t, uv);
    });
}
void launch_trace_shade(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays,
                        uint32_t grid_cap) {
    if (!max_rays) return;
    bool cnt = ba.touch_nodes != nullptr;
    with_stack(geom_stack_need(sv.depth), [&](auto S) {
        constexpr int K = decltype(S)::value;
        const uint32_t grid = (max_rays + 63) / 64;
        if (grid_cap && !cnt && grid > grid_cap) {
            if (ba.cache_on) k_trace_shade_gs<K, true><<<grid_cap, 64, 0, s>>>(sv, ba);
            else k_trace_shade_gs<K, false><<<grid_cap, 64, 0, s>>>(sv, ba);
        } else if (ba.cache_on) k_trace_shade<K, false, true><<<grid, 64, 0, s>>>(sv, ba);
        else if (cnt) k_trace_shade<K, true, false><<<grid, 64, 0, s>>>(sv, ba);
        else k_trace_shade<K, false, false><<<grid, 64, 0, s>>>(sv, ba);
    });
}

// the full-path pixels of a cache frame: FULL_PER_THREAD pixels per thread and one counter
// update per block (every wave updating one address queues the updates up behind each
// other: 0.32 ms at 1080p)
constexpr uint32_t FULL_PER_THREAD = 16;
__global__ __launch_bounds__(256) void k_full_list(const BounceArgs ba, uint32_t *list, uint32_t *count) {
    __shared__ uint32_t s_wave[4];
    __shared__ uint32_t s_base;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    const uint32_t first = blockIdx.x * 256u * FULL_PER_THREAD + t;
    uint32_t take = 0;  // bit k: pixel first + 256 k
#pragma unroll
    for (uint32_t k = 0; k < FULL_PER_THREAD; ++k) {
        const uint32_t px = first + 256u * k;
        if (px < ba.P && full_path(ba, px)) take |= 1u << k;
    }
    const uint32_t own = (uint32_t)__popc(take);
    uint32_t inc = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += u;
    }
    if (lane == 63u) s_wave[wv] = inc;
    __syncthreads();
    if (t == 0) {
        const uint32_t tot = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        s_base = tot ? atomicAdd(count, tot) : 0u;
    }
    __syncthreads();
    uint32_t pos = s_base + inc - own;
    for (uint32_t k = 0; k < wv; ++k) pos += s_wave[k];
    for (uint32_t k = 0; k < FULL_PER_THREAD; ++k)
        if (take >> k & 1u) list[pos++] = first + 256u * k;
}
void launch_full_list(hipStream_t s, const BounceArgs &ba, uint32_t *list, uint32_t *count) {
    k_full_list<<<(ba.P + 256 * FULL_PER_THREAD - 1) / (256 * FULL_PER_THREAD), 256, 0, s>>>(ba, list, count);
}

__global__ __launch_bounds__(256) void k_lgrid_clear(uint32_t *g, size_t words) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (size_t)gridDim.x * blockDim.x)
        g[i] = 0u;
}
__global__ __launch_bounds__(256) void k_lgrid_mark(const Tri *ltri, uint32_t nl, uint8_t *grid, float lo0,
                                                    float lo1, float lo2, float inv, uint32_t d0, uint32_t d1,
                                                    uint32_t d2, float margin) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const Tri t = ltri[i];
    const float lo[3] = {lo0, lo1, lo2};
    const uint32_t dim[3] = {d0, d1, d2};
    const float v[3][3] = {{t.v0.x, t.v0.y, t.v0.z}, {t.v1.x, t.v1.y, t.v1.z}, {t.v2.x, t.v2.y, t.v2.z}};
    uint32_t c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        const float mn = fminf(fminf(v[0][a], v[1][a]), v[2][a]) - margin;
        const float mx = fmaxf(fmaxf(v[0][a], v[1][a]), v[2][a]) + margin;
        const float f0 = floorf((mn - lo[a]) * inv), f1 = floorf((mx - lo[a]) * inv);
        c0[a] = f0 <= 0.0f ? 0u : (uint32_t)fminf(f0, (float)(dim[a] - 1u));
        c1[a] = f1 <= 0.0f ? 0u : (uint32_t)fminf(f1, (float)(dim[a] - 1u));
    }
    for (uint32_t z = c0[2]; z <= c1[2]; ++z)
        for (uint32_t y = c0[1]; y <= c1[1]; ++y)
            for (uint32_t x = c0[0]; x <= c1[0]; ++x) grid[((size_t)z * d1 + y) * d0 + x] = 1u;
}
void launch_light_grid(hipStream_t s, const Tri *ltri, uint32_t nl, uint8_t *grid, const float lo[3], float inv,
                       const uint32_t dim[3], float margin) {
    const size_t cells = (size_t)dim[0] * dim[1] * dim[2];
    const size_t words = (cells + 3) / 4;
    k_lgrid_clear<<<(uint32_t)std::min<size_t>((words + 255) / 256, 8192), 256, 0, s>>>((uint32_t *)grid, words);
    if (nl)
        k_lgrid_mark<<<(nl + 255) / 256, 256, 0, s>>>(ltri, nl, grid, lo[0], lo[1], lo[2], inv, dim[0], dim[1], dim[2],
                                                      margin);
}
void launch_wide4(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, Wide4 *wide4) {
    if (n_nodes > 1) k_wide4<<<(n_nodes + 255) / 256, 256, 0, s>>>(nodes, n_nodes, wide4);
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
void launch_nee_pdf_lazy(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays, uint32_t *list,
                         uint32_t *list_count, uint32_t *mlist, uint32_t *mcount) {
    if (!max_rays) return;
    k_nee_select<<<(max_rays + NEE_SELECT_THREADS - 1) / NEE_SELECT_THREADS, NEE_SELECT_THREADS, 0, s>>>(
        sv, ba, list, list_count, mlist, mcount);
    with_stack(3 * ((sv.ldepth + 1) / 2), [&](auto S) {
        constexpr int K = decltype(S)::value;
        const uint32_t grid = std::min<uint32_t>((max_rays + 63) / 64, VRGI_NEE_LIST_GRID);
        k_nee_near<K><<<grid, 64, 0, s>>>(sv, ba, mlist, mcount, list, list_count);
        k_nee_listed<K><<<grid, 64, 0, s>>>(sv, ba, list, list_count);
    });
}
void launch_nee_pdf(hipStream_t s, const SceneView &sv, const BounceArgs &ba, uint32_t max_rays, uint32_t grid_cap) {
    if (!max_rays) return;
    // the 4-ary walk pushes at most 3 references per two binary levels
    with_stack(3 * ((sv.ldepth + 1) / 2), [&](auto S) {
        constexpr int K = decltype(S)::value;
        const uint32_t grid = (max_rays + 63) / 64;
        const bool gs = grid_cap && grid > grid_cap;
        if (ba.nee_cnt) {  // enumeration, walks, overflow rays
            k_nee_enum<K><<<gs ? grid_cap : grid, 64, 0, s>>>(sv, ba, gs ? 1u : 0u);
            k_nee_walk<<<gs ? grid_cap : grid, 64, 0, s>>>(sv, ba, gs ? 1u : 0u);
            k_nee_overflow<K><<<64, 64, 0, s>>>(sv, ba);
        } else if (gs) {
            k_nee_pdf_gs<K><<<grid_cap, 64, 0, s>>>(sv, ba);
        } else {
            k_nee_pdf<K><<<grid, 64, 0, s>>>(sv, ba);
        }
    });
}

}  // namespace vrgi
