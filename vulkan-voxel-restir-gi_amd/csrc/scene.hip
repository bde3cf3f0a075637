// scene.hip — builder inputs for triangle lists and instance boxes, and the
// InstanceData.light_bvh_tl_idx mapping (scene.rs:281-287).
#include <hip/hip_runtime.h>

#include "devcommon.h"
#include "vrgi_kernels.h"

namespace vrgi {

// build_bl_bvh inputs (build.rs:262-270) from 9 floats per triangle
__global__ void k_prims_from_tris(const float *tri9, uint32_t n, float *prim9) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    f3 w[3];
    for (int k = 0; k < 3; ++k) {
        const float *p = tri9 + (size_t)g * 9 + 3 * k;
        w[k] = mk(p[0] + 0.0f, p[1] + 0.0f, p[2] + 0.0f);  // -0.0 -> +0.0
    }
    write_prim9(prim9, (size_t)n, g, w);
}

// build_tl_bvh inputs (build.rs:391-400): Aabb::transform (aabb.rs:75-91) of each
// instance's light box, centroid (min + max) / 2
__global__ void k_tl_inputs(const float *aabb6, const float *xf12, uint32_t n, float *prim9) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const float *b = aabb6 + (size_t)g * 6;
    const float *m = xf12 + (size_t)g * 12;
    f3 lo = mk(b[0], b[1], b[2]), hi = mk(b[3], b[4], b[5]);
    f3 c[8] = {xform_pt(m, lo),
               xform_pt(m, mk(lo.x, lo.y, hi.z)),
               xform_pt(m, mk(lo.x, hi.y, lo.z)),
               xform_pt(m, mk(lo.x, hi.y, hi.z)),
               xform_pt(m, mk(hi.x, lo.y, lo.z)),
               xform_pt(m, mk(hi.x, lo.y, hi.z)),
               xform_pt(m, mk(hi.x, hi.y, lo.z)),
               xform_pt(m, hi)};
    float mn[3] = {c[0].x, c[0].y, c[0].z}, mx[3] = {c[0].x, c[0].y, c[0].z};
    for (int k = 0; k < 8; ++k) {
        mn[0] = gmin(mn[0], c[k].x); mn[1] = gmin(mn[1], c[k].y); mn[2] = gmin(mn[2], c[k].z);
        mx[0] = gmax(mx[0], c[k].x); mx[1] = gmax(mx[1], c[k].y); mx[2] = gmax(mx[2], c[k].z);
    }
    for (int k = 0; k < 3; ++k) {
        mn[k] = mn[k] + 0.0f;
        mx[k] = mx[k] + 0.0f;
        prim9[(size_t)k * n + g] = mn[k];
        prim9[(size_t)(3 + k) * n + g] = mx[k];
        prim9[(size_t)(6 + k) * n + g] = (mn[k] + mx[k]) / 2.0f;
    }
}

// prim2node: every TL leaf names its instance (scene.rs:281-287)
__global__ void k_tl_idx(const vrgi_bvh_node *tl, uint32_t n_nodes, DevInst *inst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    vrgi_bvh_node nd = tl[i];
    if (nd.left == VRGI_INVALID && nd.right_or_prim != VRGI_INVALID) inst[nd.right_or_prim].tl_idx = i;
}

// child records of every internal node of a light BVH
__global__ void k_light_pairs(const vrgi_bvh_node *nodes, uint32_t n_nodes, LPair *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const vrgi_bvh_node nd = nodes[i];
    if (nd.left == VRGI_INVALID) return;
    const vrgi_bvh_node l = nodes[nd.left], r = nodes[nd.right_or_prim];
    const uint32_t lref = l.left == VRGI_INVALID ? (0x80000000u | l.right_or_prim) : nd.left;
    const uint32_t rref = r.left == VRGI_INVALID ? (0x80000000u | r.right_or_prim) : nd.right_or_prim;
    LPair p;
    p.a = make_float4(l.min[0], l.min[1], l.min[2], l.power);
    p.b = make_float4(l.max[0], l.max[1], l.max[2], r.power);
    p.c = make_float4(r.min[0], r.min[1], r.min[2], __uint_as_float(lref));
    p.d = make_float4(r.max[0], r.max[1], r.max[2], __uint_as_float(rref));
    out[i] = p;
}

// ancestors of each world light triangle in the order reverseTraverseBvh visits them
// (nee_pdf.rs:154-228): up the instance's BL-BVH from the light's leaf, then up the
// TL-BVH from the instance's leaf
__global__ void k_light_ancestry(const uint32_t *ltri_inst, const uint32_t *ltri_node, const DevInst *inst,
                                 const vrgi_bvh_node *tl, uint32_t n_ltri, uint32_t *lanc, uint32_t *lanc_len) {
    uint32_t lid = blockIdx.x * blockDim.x + threadIdx.x;
    if (lid >= n_ltri) return;
    const uint32_t in = ltri_inst[lid];
    const vrgi_bvh_node *bl = inst[in].light_bvh;
    uint32_t *out = lanc + (size_t)lid * VRGI_LANC_MAX;
    uint32_t k = 0;
    bool fits = true;
    for (int level = 0; level < 2 && fits; ++level) {
        const vrgi_bvh_node *t = level == 0 ? bl : tl;
        uint32_t ni = level == 0 ? ltri_node[lid] : inst[in].tl_idx;
        uint32_t parent = t[ni].parent;
        while (parent != VRGI_INVALID) {
            if (k == VRGI_LANC_MAX) { fits = false; break; }
            const vrgi_bvh_node pn = t[parent];
            out[k++] = parent | (level ? 0x40000000u : 0u) | (pn.left == ni ? 0u : 0x80000000u);
            ni = parent;
            parent = pn.parent;
        }
    }
    lanc_len[lid] = fits ? k : VRGI_INVALID;
}

void launch_light_pairs(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, LPair *out) {
    if (n_nodes) k_light_pairs<<<(n_nodes + 255) / 256, 256, 0, s>>>(nodes, n_nodes, out);
}
void launch_light_ancestry(hipStream_t s, const uint32_t *ltri_inst, const uint32_t *ltri_node,
                           const DevInst *inst, const vrgi_bvh_node *tl, uint32_t n_ltri,
                           uint32_t *lanc, uint32_t *lanc_len) {
    if (n_ltri)
        k_light_ancestry<<<(n_ltri + 255) / 256, 256, 0, s>>>(ltri_inst, ltri_node, inst, tl, n_ltri, lanc, lanc_len);
}

void launch_prims_from_tris(hipStream_t s, const float *tri9, uint32_t n, float *prim9) {
    if (n) k_prims_from_tris<<<(n + 255) / 256, 256, 0, s>>>(tri9, n, prim9);
}
void launch_tl_inputs(hipStream_t s, const float *aabb6, const float *xf12, uint32_t n, float *prim9) {
    if (n) k_tl_inputs<<<(n + 255) / 256, 256, 0, s>>>(aabb6, xf12, n, prim9);
}
void launch_tl_idx(hipStream_t s, const vrgi_bvh_node *tl, uint32_t n_nodes, DevInst *inst) {
    if (n_nodes) k_tl_idx<<<(n_nodes + 255) / 256, 256, 0, s>>>(tl, n_nodes, inst);
}

}  // namespace vrgi
