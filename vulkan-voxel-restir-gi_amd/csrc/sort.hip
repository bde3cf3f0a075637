// Morton ordering of the ray queue between bounces (wavefront compaction, second half).
//
// The reference keeps one ray slot per pixel and never reorders (raytrace.rs:406-416).
// Here the queue of rays alive at bounce b+1 (pixel ids, written by the ballot
// compaction of bounce b) is reordered by a key made of the ray direction's octant
// (top 3 bits) and the Morton code of the ray origin quantised in the scene's root box
// (`bits` bits per axis). Rays that share a wave then start close together and travel
// the same way, so they visit the same BVH nodes. Every ray still reads and writes only
// its own pixel's slots and seeds, so the order changes no result bit.
//
// Counting sort with 2^(3 + 3 bits) buckets: histogram, one-block exclusive scan,
// scatter. Both the histogram and the scatter aggregate equal keys of a wave into one
// atomic (queue order follows pixel order, so neighbouring lanes often share a key).
// The order inside a bucket is not fixed; nothing depends on it.
#include "vrgi_kernels.h"

namespace vrgi {

__device__ __forceinline__ uint32_t spread3(uint32_t v) {  // bit i -> bit 3i, 10 bits max
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t ray_key(const SortArgs &a, uint32_t px) {
    const size_t src = ((size_t)a.bounce * a.P + px) * 3;
    const float ox = a.origin[src], oy = a.origin[src + 1], oz = a.origin[src + 2];
    const float dx = a.dir[src], dy = a.dir[src + 1], dz = a.dir[src + 2];
    const uint32_t oct = (dx < 0.0f ? 1u : 0u) | (dy < 0.0f ? 2u : 0u) | (dz < 0.0f ? 4u : 0u);
    const float4 ra = a.root->a, rb = a.root->b, rc = a.root->c;
    const float lo[3] = {fminf(ra.x, rb.z), fminf(ra.y, rb.w), fminf(ra.z, rc.x)};
    const float hi[3] = {fmaxf(ra.w, rc.y), fmaxf(rb.x, rc.z), fmaxf(rb.y, rc.w)};
    const float cells = (float)(1u << a.bits);
    auto q = [&](float v, int k) {
        float t = (v - lo[k]) / (hi[k] - lo[k]) * cells;
        t = t == t ? t : 0.0f;
        t = fminf(fmaxf(t, 0.0f), cells - 1.0f);
        return (uint32_t)t;
    };
    const uint32_t m = spread3(q(ox, 0)) | (spread3(q(oy, 1)) << 1) | (spread3(q(oz, 2)) << 2);
    return (oct << (3 * a.bits)) | m;
}

// one atomicAdd per distinct key of the wave; returns the lane's slot (old + rank)
__device__ __forceinline__ uint32_t wave_aggregated_add(uint32_t *ctr, uint32_t key, bool active) {
    const uint32_t lane = threadIdx.x & 63u;
    uint64_t todo = __ballot(active);
    uint32_t slot = 0;
    while (todo) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
        const uint32_t k = __shfl(key, (int)leader, 64);
        const uint64_t same = __ballot(active && key == k) & todo;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(ctr + k, (uint32_t)__popcll(same));
        base = __shfl(base, (int)leader, 64);
        if ((same >> lane) & 1ull) slot = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return slot;
}

__global__ __launch_bounds__(256) void k_sort_clear(uint32_t *hist, uint32_t nb) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) hist[i] = 0;
}

__global__ __launch_bounds__(256) void k_sort_hist(SortArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *a.count;
    const bool active = i < n;
    uint32_t key = 0;
    if (active) {
        key = ray_key(a, a.queue_in[i]);
        a.keys[i] = key;
    }
    (void)wave_aggregated_add(a.hist, key, active);
}

// exclusive scan of nb counters by one block of 1024 threads
__global__ __launch_bounds__(1024) void k_sort_scan(uint32_t *hist, uint32_t nb) {
    __shared__ uint32_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023u) / 1024u;
    const uint32_t beg = t * per, end = min(nb, beg + per);
    uint32_t sum = 0;
    for (uint32_t i = beg; i < end; ++i) sum += hist[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = beg; i < end; ++i) {
        uint32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(SortArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = *a.count;
    const bool active = i < n;
    const uint32_t key = active ? a.keys[i] : 0u;
    const uint32_t pos = wave_aggregated_add(a.hist, key, active);
    if (active) a.queue_out[pos] = a.queue_in[i];
}

void launch_sort_queue(hipStream_t s, const SortArgs &a, uint32_t max_rays) {
    const uint32_t nb = 8u << (3 * a.bits);
    const uint32_t grid = (max_rays + 255) / 256;
    k_sort_clear<<<(nb + 255) / 256, 256, 0, s>>>(a.hist, nb);
    k_sort_hist<<<grid, 256, 0, s>>>(a);
    k_sort_scan<<<1, 1024, 0, s>>>(a.hist, nb);
    k_sort_scatter<<<grid, 256, 0, s>>>(a);
}

}  // namespace vrgi
