// restir.hip — per-pixel resolve kernels for gfx950: outgoing radiance + ReSTIR GI
// temporal update (fused), spatial resampling + finalize (fused), postprocess.
//
// Replaces outgoing_radiance.rs:58-94, the six initial-sample copies of
// interactive_rendering.rs:1359-1414, restir_temporal_resampling.rs:235-301,
// restir_spatial_resampling.rs:253-406, restir_finalize.rs:121-141 and
// postprocess.rs:39-88. Per-bounce data is read only for the bounces a path reached
// (path_len); the reference's dead-ray entries are the constants of
// raytrace.rs:406-416 (origin passed through, everything else 0, bsdf_pdf 1), which
// these kernels substitute without reading them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "devcommon.h"
#include "vrgi_kernels.h"

namespace vrgi {

// reservoir words: x_v 0-2, n_v 3-5, x_s 6-8, n_s 9-11, l_o_hat 12-14, p_omega 15,
// seed 16 (u32), ucw 17, m 18 (u32), w_sum 19 (restir_temporal_resampling.rs:14-30)
struct Sample { f3 xv, nv, xs, ns, lo; float pw; uint32_t seed; };
struct Reservoir { Sample z; float ucw; uint32_t m; float wsum; };

__device__ __forceinline__ Reservoir res_load(const float *p) {
    const float4 *q = (const float4 *)p;
    float4 a = q[0], b = q[1], c = q[2], d = q[3], e = q[4];
    Reservoir r;
    r.z.xv = mk(a.x, a.y, a.z); r.z.nv = mk(a.w, b.x, b.y); r.z.xs = mk(b.z, b.w, c.x);
    r.z.ns = mk(c.y, c.z, c.w); r.z.lo = mk(d.x, d.y, d.z); r.z.pw = d.w;
    r.z.seed = __float_as_uint(e.x); r.ucw = e.y; r.m = __float_as_uint(e.z); r.wsum = e.w;
    return r;
}
__device__ __forceinline__ void res_store(float *p, const Reservoir &r) {
    float4 *q = (float4 *)p;
    q[0] = make_float4(r.z.xv.x, r.z.xv.y, r.z.xv.z, r.z.nv.x);
    q[1] = make_float4(r.z.nv.y, r.z.nv.z, r.z.xs.x, r.z.xs.y);
    q[2] = make_float4(r.z.xs.z, r.z.ns.x, r.z.ns.y, r.z.ns.z);
    q[3] = make_float4(r.z.lo.x, r.z.lo.y, r.z.lo.z, r.z.pw);
    q[4] = make_float4(__uint_as_float(r.z.seed), r.ucw, __uint_as_float(r.m), r.wsum);
}
__device__ __forceinline__ float luminance(f3 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
// updateReservoir, restir_temporal_resampling.rs:235-246
__device__ __forceinline__ void update_res(uint32_t seed, Reservoir &r, const Sample &z, float w) {
    r.wsum += w;
    r.m += 1;
    if (float_construct(seed) < w / r.wsum) r.z = z;
}

// one step of outgoing_radiance.rs:58-94: L_b from L_{b+1} (np 0 past the path's end,
// where the reference's nee_pdf is never written)
__device__ __forceinline__ f3 radiance_step(const FrameArgs &fa, size_t bi, bool next_live, f3 L, float bsp) {
    float bp = fa.pdf[bi], w = fa.mis[bi];
    float np = (fa.nee != nullptr && next_live) ? fa.nee[bi] : 0.0f;
    float q = np * w + (1.0f - w) * bsp;
    float rf = bp / q;
    f3 term = ((ld3(fa.refl + bi * 3) * L) * rf) * 1.0f;
    return ld3(fa.emis + bi * 3) + term;
}

__global__ __launch_bounds__(256) void k_resolve_step(const FrameArgs fa, uint32_t k, float *lrun) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= fa.P) return;
    const uint32_t plen = fa.path_len[id];
    if (plen <= k) return;
    const bool next_live = k + 1 < plen;
    const f3 Ln = next_live ? ld3(lrun + (size_t)id * 3) : mk(0, 0, 0);
    const float bsp = 1.0f / (2.0f * VRGI_PI_F);
    st3(lrun + (size_t)id * 3, radiance_step(fa, (size_t)k * fa.P + id, next_live, Ln, bsp));
}
void launch_resolve_step(hipStream_t s, const FrameArgs &fa, uint32_t k, float *lrun) {
    k_resolve_step<<<(fa.P + 255) / 256, 256, 0, s>>>(fa, k, lrun);
}

// outgoing_radiance (backward recurrence) + initial sample + temporal update. Called by
// every thread of the block (a block-wide record allocation when the cache is on).
__device__ __forceinline__ void accum_temporal_px(const FrameArgs &fa, uint32_t id) {
    const bool live = id < fa.P;
    const uint32_t plen = live ? fa.path_len[id] : 0u;
    const float bsp = 1.0f / (2.0f * VRGI_PI_F);
    f3 L = mk(0, 0, 0), L1 = mk(0, 0, 0);
    float q0 = bsp;
    // learned radiance cache: a queried path ends at vertex cache_vertex with the cache's
    // reflected radiance; a training path emits one record per vertex whose reflected
    // radiance is known (not the last vertex of a path cut by the bounce cap)
    bool train = false, queried = false;
    const uint32_t cv = fa.cache_vertex;
    if (fa.cache_on && live) {
        train = mm_finalizef(mm_combine(fa.train_seed, id)) < fa.train_frac ||
                in_fallback(fa.fallback, id, fa.W, fa.tile, fa.tiles_x);
        queried = !train && plen == cv + 1 && fa.hit[(size_t)cv * fa.P + id] != VRGI_INVALID;
    }
    // the recurrence from the path's last vertex to its first; with `write`, the records
    // go to slots slot, slot + 1, ... in that order (counted, not written, otherwise)
    auto walk = [&](bool write, uint32_t slot) -> uint32_t {
        uint32_t n_rec = 0;
        L = mk(0, 0, 0);
        for (int b = (int)plen - 1; b >= 0; --b) {
            const size_t bi = (size_t)b * fa.P + id;
            float bp = fa.pdf[bi], w = fa.mis[bi];
            // written only for continued rays, and only when NEE is on (nee == nullptr
            // otherwise: mis_weight is then 0 everywhere and the reference's nee_pdf is 0)
            float np = (fa.nee != nullptr && b + 1 < (int)plen) ? fa.nee[bi] : 0.0f;
            float q = np * w + (1.0f - w) * bsp;
            float rf = bp / q;
            f3 term = ((ld3(fa.refl + bi * 3) * L) * rf) * 1.0f;
            if (queried && b == (int)cv) term = ld3(fa.crad + (size_t)id * 3);
            L = ld3(fa.emis + bi * 3) + term;
            const bool rec = train && fa.hit[bi] != VRGI_INVALID &&
                             (b + 1 < (int)plen || zero3(ld3(fa.dir + bi * 3))) &&
                             isfinite(term.x) && isfinite(term.y) && isfinite(term.z);
            if (rec && write && slot + n_rec < fa.tcap) {
                const f3 p = ld3(fa.origin + bi * 3);
                const f3 dd = b == 0 ? ld3(fa.d0 + (size_t)id * 3) : ld3(fa.dir + (bi - fa.P) * 3);
                const f3 nn = ld3(fa.normal + bi * 3), aa = ld3(fa.albedo + bi * 3);
                float4 *r = (float4 *)(fa.trec + (size_t)(slot + n_rec) * 16);
                r[0] = make_float4(p.x, p.y, p.z, dd.x);
                r[1] = make_float4(dd.y, dd.z, nn.x, nn.y);
                r[2] = make_float4(nn.z, aa.x, aa.y, aa.z);
                r[3] = make_float4(term.x, term.y, term.z, __uint_as_float(id * 16u + (uint32_t)b));
            }
            n_rec += rec ? 1u : 0u;
            if (b == 1) L1 = L;
            if (b == 0) q0 = q;
        }
        return n_rec;
    };
    const uint32_t n_rec = walk(false, 0u);
    if (fa.cache_on) {  // uniform over the block
        // the block's records take one range of the record buffer (one counter update
        // per block); each thread's records are consecutive, in decreasing vertex order
        __shared__ uint32_t s_wave[16];
        __shared__ uint32_t s_base;
        const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
        uint32_t inc = n_rec;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += u;
        }
        if (lane == 63u) s_wave[wv] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t k = 0; k < nw; ++k) { const uint32_t v = s_wave[k]; s_wave[k] = t; t += v; }
            s_base = t ? atomicAdd(fa.trec_count, t) : 0u;
        }
        __syncthreads();
        if (n_rec) walk(true, s_base + s_wave[wv] + inc - n_rec);
        __syncthreads();  // s_wave / s_base may be reused by a later call in this block
    }
    if (!live) return;
    st3(fa.radiance + (size_t)id * 3, L);
    // initial sample (interactive_rendering.rs:1359-1414)
    Sample S;
    S.xv = ld3(fa.origin + (size_t)id * 3);
    S.nv = ld3(fa.normal + (size_t)id * 3);
    S.xs = plen >= 2 ? ld3(fa.origin + ((size_t)fa.P + id) * 3) : S.xv;
    S.ns = plen >= 2 ? ld3(fa.normal + ((size_t)fa.P + id) * 3) : mk(0, 0, 0);
    S.lo = L1;
    S.pw = q0;
    S.seed = 0;
    // restir_temporal_resampling.rs:256-301
    const uint32_t ps = mm_combine(fa.tseed, id);
    float *slot = fa.temporal + (size_t)id * 20;
    Reservoir R = res_load(slot);
    if (mm_finalizef(mm_combine(ps, 0)) < 0.4f) { R.wsum = 0.0f; R.m = 0; }
    float w = luminance(S.lo) / S.pw;
    update_res(mm_finalize(mm_combine(ps, 1)), R, S, w);
    float ph = luminance(R.z.lo);
    R.ucw = ph > 0.0f ? R.wsum / ((float)R.m * ph) : 0.0f;
    res_store(slot, R);
}
__global__ __launch_bounds__(256) void k_accum_temporal(const FrameArgs fa) {
    if (fa.part == 2) {  // the listed full-path pixels; blocks stride (uniformly per block)
        const uint32_t n = min(*fa.full_count, fa.P);
        for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
            const uint32_t i = base + threadIdx.x;
            accum_temporal_px(fa, i < n ? fa.full_list[i] : 0xFFFFFFFFu);
        }
        return;
    }
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (fa.part == 1 && id < fa.P &&
        full_path_px(fa.train_seed, fa.train_frac, fa.fallback, id, fa.W, fa.tile, fa.tiles_x))
        id = 0xFFFFFFFFu;  // left to part 2
    accum_temporal_px(fa, id);
}

// restir_spatial_resampling.rs:316-406 + restir_finalize.rs:121-141
__device__ __forceinline__ void spatial_finalize_px(const FrameArgs &fa, uint32_t id) {
    if (id >= fa.P) return;
    const uint32_t qx = id % fa.W, qy = id / fa.W;
    const uint32_t ps = mm_combine(fa.sseed, id);
    Reservoir Rs = res_load(fa.temporal + (size_t)id * 20);
    const Sample S = Rs.z;
    const float cos25 = 0.906307787036650f;  // cos(radians(25.0))
    uint32_t Z = luminance(Rs.z.lo) > 0.0f ? Rs.m : 0u;  // Q[0] = the unmerged reservoir
    for (uint32_t it = 0; it < fa.spatial_iters; ++it) {
        uint32_t is = mm_combine(ps, it);
        float jx = fa.radius * (mm_finalizef(mm_combine(is, 0)) - 0.5f);
        float jy = fa.radius * (mm_finalizef(mm_combine(is, 1)) - 0.5f);
        int nx = (int)(((float)qx + 0.5f) + jx), ny = (int)(((float)qy + 0.5f) + jy);
        nx = min(max(nx, 0), (int)fa.W - 1);
        ny = min(max(ny, 0), (int)fa.H - 1);
        Reservoir Rn = res_load(fa.temporal + ((size_t)ny * fa.W + nx) * 20);
        // geometricallySimilar, :277-288
        if (zero3(S.nv) || zero3(Rn.z.nv)) continue;
        if (dot3(S.nv, Rn.z.nv) < cos25) continue;
        // computeJacobianQToR, :299-311
        float J;
        if (zero3(Rn.z.ns)) {
            J = 1.0f;
        } else {
            f3 qv = Rn.z.xv - Rn.z.xs, rv = S.xv - Rn.z.xs;
            float cq = fabsf(dot3(norm3(qv), Rn.z.ns));
            float cr = fabsf(dot3(norm3(rv), Rn.z.ns));
            J = (cr / cq) * (len3(qv) / len3(rv));
        }
        J = gmin(gmax(J, 0.1f), 10.0f);
        float padj = luminance(Rn.z.lo) / J;
        uint32_t m0 = Rs.m;  // mergeReservoir, :266-275
        update_res(mm_finalize(mm_combine(is, 2)), Rs, Rn.z, (padj * Rn.ucw) * (float)Rn.m);
        Rs.m = m0 + Rn.m;
        if (luminance(Rn.z.lo) > 0.0f) Z += Rn.m;
    }
    float ph = luminance(Rs.z.lo);
    Rs.ucw = ph == 0.0f ? 0.0f : Rs.wsum / ((float)Z * ph);
    res_store(fa.spatial + (size_t)id * 20, Rs);
    // finalize
    f3 x1 = ld3(fa.origin + (size_t)id * 3);
    f3 dir = norm3(Rs.z.xs - x1);
    float bp = dot3(dir, Rs.z.nv) / VRGI_PI_F;
    f3 Lf = ld3(fa.emis + (size_t)id * 3) + ((ld3(fa.refl + (size_t)id * 3) * Rs.z.lo) * Rs.ucw) * bp;
    st3(fa.restir + (size_t)id * 3, Lf);
    float dv = Rs.ucw / 10.0f;
    st3(fa.debug + (size_t)id * 3, mk(dv, dv, dv));
}
__global__ __launch_bounds__(256) void k_spatial_finalize(const FrameArgs fa) {
    spatial_finalize_px(fa, blockIdx.x * blockDim.x + threadIdx.x);
}

// postprocess.rs:39-88 with scale 1: view select, BGRA8 (clamped; the reference leaves
// out-of-range conversion undefined)
__device__ __forceinline__ void postprocess_px(const FrameArgs &fa, uint32_t id) {
    if (id >= fa.P) return;
    const uint32_t v = fa.debug_view;
    f3 c;
    if (v == 4) {  // running mean of the path-traced radiance since the last restart
        const f3 L = ld3(fa.radiance + (size_t)id * 3);
        const f3 a = fa.view_n == 1u ? L : ld3(fa.view_acc + (size_t)id * 3) + L;
        st3(fa.view_acc + (size_t)id * 3, a);
        c = a / (float)fa.view_n;
    } else {
        const float *src = v == 0 ? fa.radiance : ((v == 1 || v == 3) ? fa.restir : fa.debug);
        c = ld3(src + (size_t)id * 3);
    }
    c = c / (float)(1 * 1);
    if (v >= 3) c = mk(c.x / (1.0f + c.x), c.y / (1.0f + c.y), c.z / (1.0f + c.z));  // Reinhard, per channel
    auto cv = [](float v) -> uint32_t {
        float x = v * 255.0f;
        if (!(x > 0.0f)) return 0u;
        if (x >= 255.0f) return 255u;
        return (uint32_t)x;
    };
    uint32_t px = cv(c.z) | (cv(c.y) << 8) | (cv(c.x) << 16) | (255u << 24);
    ((uint32_t *)fa.bgra8)[id] = px;
}
__global__ __launch_bounds__(256) void k_postprocess(const FrameArgs fa) {
    postprocess_px(fa, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------- batched scenes
template <int PB>
__global__ __launch_bounds__(PB) void k_accum_temporal_b(const FrameArgs *__restrict__ fas,
                                                         const uint32_t *__restrict__ blk, uint32_t n) {
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    accum_temporal_px(fas[s], (blockIdx.x - blk[s]) * PB + threadIdx.x);
}
template <int PB>
__global__ __launch_bounds__(PB) void k_spatial_finalize_b(const FrameArgs *__restrict__ fas,
                                                           const uint32_t *__restrict__ blk, uint32_t n) {
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    spatial_finalize_px(fas[s], (blockIdx.x - blk[s]) * PB + threadIdx.x);
}
// postprocess, then the spp mean of the ReSTIR output: (((r_1 + r_2) + ...) + r_spp) / spp
template <int PB>
__global__ __launch_bounds__(PB) void k_postprocess_b(const FrameArgs *__restrict__ fas, const uint32_t *__restrict__ blk,
                                                      uint32_t n, float *const *__restrict__ mean,
                                                      const BatchPass *__restrict__ pass) {
    const uint32_t s = batch_scene(blk, n, blockIdx.x);
    const FrameArgs &fa = fas[s];
    const uint32_t id = (blockIdx.x - blk[s]) * PB + threadIdx.x;
    postprocess_px(fa, id);
    if (id >= fa.P) return;
    float *m = mean[s] + (size_t)id * 3;
    const f3 r = ld3(fa.restir + (size_t)id * 3);
    f3 v = pass->first ? r : ld3(m) + r;
    if (pass->last) v = v * pass->inv_spp;
    st3(m, v);
}
#define VRGI_PB_LAUNCH(K, ...)                                                                   \
    do {                                                                                         \
        if (L.pb == 64) K<64><<<L.px_blocks, 64, 0, s>>>(__VA_ARGS__);                            \
        else if (L.pb == 128) K<128><<<L.px_blocks, 128, 0, s>>>(__VA_ARGS__);                    \
        else K<256><<<L.px_blocks, 256, 0, s>>>(__VA_ARGS__);                                     \
    } while (0)
void launch_batch_pixels_gi(hipStream_t s, const BatchLaunch &L) {
    VRGI_PB_LAUNCH(k_accum_temporal_b, L.fas, L.blk_px, L.n);
}
void launch_batch_pixels_spatial(hipStream_t s, const BatchLaunch &L) {
    VRGI_PB_LAUNCH(k_spatial_finalize_b, L.fas, L.blk_px, L.n);
}
void launch_batch_pixels_post(hipStream_t s, const BatchLaunch &L) {
    VRGI_PB_LAUNCH(k_postprocess_b, L.fas, L.blk_px, L.n, L.mean, L.pass);
}

void launch_accum_temporal(hipStream_t s, const FrameArgs &fa) {
    const uint32_t grid = (fa.P + 255) / 256;
    k_accum_temporal<<<fa.part == 2 ? std::min(grid, 2048u) : grid, 256, 0, s>>>(fa);
}
void launch_spatial_finalize(hipStream_t s, const FrameArgs &fa) {
    k_spatial_finalize<<<(fa.P + 255) / 256, 256, 0, s>>>(fa);
}
void launch_postprocess(hipStream_t s, const FrameArgs &fa) {
    k_postprocess<<<(fa.P + 255) / 256, 256, 0, s>>>(fa);
}

// running FP64 sum of the path-traced radiance (convergence measurements). A non-finite
// sample (the reference's 0/0 of a zero total light importance or a zero cosine in the
// NEE pdf) adds 0, as the reference's display shows it black.
__global__ __launch_bounds__(256) void k_accumulate(const float *rad, double *acc, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float v = rad[i];
        acc[i] += isfinite(v) ? (double)v : 0.0;
    }
}
void launch_accumulate(hipStream_t s, const float *rad, double *acc, size_t n) {
    if (n) k_accumulate<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(rad, acc, n);
}

}  // namespace vrgi
