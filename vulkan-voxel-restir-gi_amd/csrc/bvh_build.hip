// bvh_build.hip — bit-exact binned-SAH BVH builder for gfx950, batched over many trees.
//
// Semantics: build_bl_bvh / build_tl_bvh of the reference (bvh/build.rs:36-480):
// 32 bins x 3 axes, strict-< first minimum of area(L)*nL + area(R)*nR
// (build.rs:45-128), in-place `partition` (build.rs:145-149; the i-th false of
// [0,T) trades places with the i-th true of [T,n) counted from the right), the
// half split fallback (build.rs:152-157), the un-binned 2-primitive case
// (build.rs:203-230), one primitive per leaf, 40-byte BvhNode output with 1e-4
// padding on internal boxes (build.rs:300-339) and power = (0 + left) + right
// (build.rs:345-357). Node numbering: the reference allocates children in DFS
// preorder of splits, so the split node of preorder rank k owns children 2k+1 and
// 2k+2; a left child has rank k+1, a right child rank k + (left count). Nodes are
// therefore numbered independently of the order in which this builder visits them.
//
// Execution: level-synchronous "top" phase for nodes larger than SUB_MAX (one
// workgroup per CHUNK of a node's range; bins reduced in LDS then global
// atomics), then one 64-lane wave per remaining subtree ("jobs", <= SUB_MAX
// primitives) with the subtree's primitives resident in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "devcommon.h"
#include "vrgi_internal.h"

namespace vrgi {

static constexpr int BINS = 32;
#ifndef VRGI_CHUNK
#define VRGI_CHUNK 2048
#endif
static constexpr int CHUNK = VRGI_CHUNK;  // primitives per top-phase work item
#ifndef VRGI_TOP_THREADS
#define VRGI_TOP_THREADS 256
#endif
static constexpr int TOP_THREADS = VRGI_TOP_THREADS;
static constexpr int PER_THREAD = CHUNK / TOP_THREADS;  // 8
static constexpr int BIN_WORDS = 3 * BINS * 7;          // count + 6 box keys per bin
#ifndef VRGI_GRID_CAP_ITEMS
#define VRGI_GRID_CAP_ITEMS 4096
#endif
static constexpr uint32_t GRID_CAP_ITEMS = VRGI_GRID_CAP_ITEMS;  // blocks of the per-item kernels
static constexpr uint32_t GRID_CAP_NODES = 8192;         // blocks of the per-node kernels

struct OpenNode {
    uint32_t first, count, rank, node, parent, tree, depth, pad;
};
struct Item {
    uint32_t open, beg, len, pad;
};
struct SplitInfo {
    uint32_t dim;
    float split;
    uint32_t T;
    uint32_t fallback;
    int box[6];  // node box keys (union of the range's primitive boxes)
};

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t bin_index(float c, float bmin, float scale) {
    // build.rs:79-82: usize::min(31, ((c - min) * scale) as usize); `as` saturates, NaN -> 0
    float f = (c - bmin) * scale;
    if (!(f == f)) return 0;
    if (f >= 31.0f) return 31;
    if (f <= 0.0f) return 0;
    return (uint32_t)f;
}
__device__ __forceinline__ float area_keys(int mnx, int mny, int mnz, int mxx, int mxy, int mxz) {
    // aabb.rs:44-52
    float dx = funkey(mxx) - funkey(mnx), dy = funkey(mxy) - funkey(mny), dz = funkey(mxz) - funkey(mnz);
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

__device__ __forceinline__ float area6(const float b[6]) {
    // aabb.rs:44-52 on a box held as min xyz, max xyz
    float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}

__device__ __forceinline__ void write_node(vrgi_bvh_node *dst, uint32_t left, uint32_t right,
                                           const float mn[3], float power, const float mx[3],
                                           uint32_t parent) {
    vrgi_bvh_node n;
    n.left = left;
    n.right_or_prim = right;
    n.min[0] = mn[0]; n.min[1] = mn[1]; n.min[2] = mn[2];
    n.power = power;
    n.max[0] = mx[0]; n.max[1] = mx[1]; n.max[2] = mx[2];
    n.parent = parent;
    *dst = n;
}

// leaf (build.rs:311-326) for primitive p
__device__ __forceinline__ void emit_leaf(const BuildArgs &a, uint32_t gnode, uint32_t parent, uint32_t p) {
    const float4 b0 = a.box8[2 * p], b1 = a.box8[2 * p + 1];
    float mn[3] = {b0.x, b0.y, b0.z};
    float mx[3] = {b0.w, b1.x, b1.y};
    write_node(a.nodes + gnode, VRGI_INVALID, __float_as_uint(b1.w), mn, b1.z, mx, parent);
}

// 2-primitive node (build.rs:203-230) of tree-local rank r / index idx, leaves at p0, p1
__device__ __forceinline__ void emit_pair(const BuildArgs &a, uint32_t base, uint32_t idx, uint32_t r,
                                          uint32_t parent, uint32_t p0, uint32_t p1) {
    const float pad = 0.0001f;
    const float4 a0 = a.box8[2 * p0], a1 = a.box8[2 * p0 + 1];
    const float4 c0 = a.box8[2 * p1], c1 = a.box8[2 * p1 + 1];
    const float l0[6] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y}, l1[6] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y};
    float mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = gmin(l0[k], l1[k]) - pad;
        mx[k] = gmax(l0[3 + k], l1[3 + k]) + pad;
    }
    float pw = (0.0f + a1.z) + c1.z;
    write_node(a.nodes + base + idx, 2 * r + 1, 2 * r + 2, mn, pw, mx, parent);
    emit_leaf(a, base + 2 * r + 1, idx, p0);
    emit_leaf(a, base + 2 * r + 2, idx, p1);
}

// a node for the middle phase (k_mid): more than SUB_MAX and at most MID_MAX primitives
// of a tree whose leaves all have zero power (the middle phase writes power +0)
__device__ __forceinline__ bool mid_node(const BuildArgs &a, uint32_t count, uint32_t tree);

// ---------------------------------------------------------------- init
__global__ void k_init(BuildArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    if (a.power[i] != 0.0f) {  // the tree holding i has nodes of non-zero power
        uint32_t lo = 0, hi = a.n_trees - 1;  // last tree whose first primitive is <= i
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (a.tree_beg[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        if (!a.powered[lo]) a.powered[lo] = 1u;
    }
    a.perm[i] = i;
    a.pc[0 * a.n + i] = a.prim9[6 * a.n + i];
    a.pc[1 * a.n + i] = a.prim9[7 * a.n + i];
    a.pc[2 * a.n + i] = a.prim9[8 * a.n + i];
    a.box8[2 * i] = make_float4(a.prim9[0 * a.n + i], a.prim9[1 * a.n + i], a.prim9[2 * a.n + i], a.prim9[3 * a.n + i]);
    a.box8[2 * i + 1] = make_float4(a.prim9[4 * a.n + i], a.prim9[5 * a.n + i], a.power[i], __uint_as_float(a.ids[i]));
}

// one thread per tree: place the root (leaf / pair / job / open node)
__global__ void k_roots(BuildArgs a, BuildState s) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n_trees) return;
    uint32_t beg = a.tree_beg[t], n = a.tree_n[t], base = a.tree_base[t];
    a.depth[t] = 0;
    if (n == 1) {
        emit_leaf(a, base, VRGI_INVALID, beg);
    } else if (n == 2) {
        emit_pair(a, base, 0, 0, VRGI_INVALID, beg, beg + 1);
        a.depth[t] = 1;
        atomicAdd(a.split_count_sum, 2ull);
    } else {
        OpenNode o{beg, n, 0, 0, VRGI_INVALID, t, 0, 0};
        if (n <= (uint32_t)a.sub_max) {
            uint32_t j = atomicAdd(s.n_jobs, 1u);
            ((OpenNode *)s.jobs)[j] = o;
        } else if (mid_node(a, n, t)) {
            uint32_t j = atomicAdd(s.n_mid, 1u);
            ((OpenNode *)s.mid)[j] = o;
        } else {
            uint32_t j = atomicAdd(s.n_open_next, 1u);
            ((OpenNode *)s.open_next)[j] = o;
        }
    }
}

// ---------------------------------------------------------------- top phase
// items for the open nodes of this level: a block-wide scan of ceil(count / CHUNK) per
// node, then every thread writes items of any node (found by binary search over the
// inclusive scan), so a node with many chunks is not written by one thread
// It also starts the level: the open list of the previous level's children becomes
// current, the level's first entry in the level-node list is recorded and, when the level
// holds open nodes, the count of non-empty levels is raised to include it.
// One block of ITEMS_THREADS: the build runs beside the frame's ray kernels, and a block
// of 1024 threads waits there until one CU has 16 wave slots free at once.
static constexpr uint32_t ITEMS_THREADS = 256;
__global__ __launch_bounds__(ITEMS_THREADS) void k_items(BuildState s, uint32_t *level_start, uint32_t lvl,
                                                         uint32_t *levels_used) {
    constexpr uint32_t NT = ITEMS_THREADS, NW = NT / 64;
    __shared__ uint32_t s_sum[NT];
    __shared__ uint32_t s_nopen;
    __shared__ uint32_t s_wave[NW];
    if (threadIdx.x == 0) {
        const uint32_t n = *s.n_open_next;
        s_nopen = n;
        *s.n_open = n;
        *s.n_open_next = 0;
        level_start[lvl] = *s.n_level_nodes;
        if (n) atomicMax(levels_used, lvl + 1u);
    }
    __syncthreads();
    const OpenNode *open = (const OpenNode *)s.open;
    const uint32_t n_open = s_nopen;
    uint32_t carry = 0;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    for (uint32_t base = 0; base < n_open; base += NT) {
        uint32_t i = base + threadIdx.x;
        uint32_t c = i < n_open ? (open[i].count + CHUNK - 1) / CHUNK : 0;
        // inclusive scan: within each wave, then over the wave totals
        uint32_t inc = c;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += u;
        }
        if (lane == 63) s_wave[w] = inc;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t k = 0; k < w; ++k) before += s_wave[k];
        s_sum[threadIdx.x] = before + inc;
        __syncthreads();
        uint32_t excl = carry + s_sum[threadIdx.x] - c;
        if (i < n_open) {
            s.item_beg[i] = excl;
            s.item_end[i] = excl + c;
        }
        const uint32_t total = s_sum[NT - 1];
        for (uint32_t j = threadIdx.x; j < total; j += NT) {
            uint32_t lo = 0, hi = NT - 1;  // first t with s_sum[t] > j
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_sum[mid] > j) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t k = j - (lo ? s_sum[lo - 1] : 0u);
            const OpenNode on = open[base + lo];
            Item it;
            it.open = base + lo;
            it.beg = on.first + k * CHUNK;
            it.len = min((uint32_t)CHUNK, on.count - k * CHUNK);
            it.pad = 0;
            ((Item *)s.items)[carry + j] = it;
        }
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *s.n_items = carry;
}

// Grids of the per-level kernels are capped on the host (an upper bound of a level's
// nodes or items can be far above its actual count); blocks stride over the device count.
__global__ void k_clear_node_data(BuildState s) {
    const uint32_t n_open = *s.n_open;
    for (uint32_t o = blockIdx.x; o < n_open; o += gridDim.x) {
        if (s.item_end[o] - s.item_beg[o] > 1u) {  // a node of one item never reads its global bins
            for (uint32_t k = threadIdx.x; k < BIN_WORDS; k += blockDim.x) {
                uint32_t w = k % 7;
                s.bins[(size_t)o * BIN_WORDS + k] = w == 0 ? 0u : (w <= 3 ? (uint32_t)KEY_MIN_ID : (uint32_t)KEY_MAX_ID);
            }
        }
        if (threadIdx.x < 6) s.bounds[o * 6 + threadIdx.x] = threadIdx.x < 3 ? KEY_MIN_ID : KEY_MAX_ID;
    }
}

// wave-wide reductions over all 64 lanes (all lanes active at every call site): the
// device library's DPP wavefront reductions
extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);
extern "C" __device__ unsigned int __ockl_wfred_add_u32(unsigned int);
extern "C" __device__ unsigned int __ockl_wfred_min_u32(unsigned int);
__device__ __forceinline__ int wave_min(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max(int v) { return __ockl_wfred_max_i32(v); }
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return __ockl_wfred_add_u32(v); }

// inclusive scans within the two 32-lane halves of the wave: row shifts by 1, 2, 4, 8
// (16-lane rows) and the broadcast of lane 15 of rows 0 and 2 into rows 1 and 3 (DPP);
// lanes without a source take the identity `id`
template <class Op>
__device__ __forceinline__ int seg32_scan(int v, int id, Op op) {
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x111, 0xF, 0xF, false));  // row_shr:1
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x112, 0xF, 0xF, false));  // row_shr:2
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x114, 0xF, 0xF, false));  // row_shr:4
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x118, 0xF, 0xF, false));  // row_shr:8
    v = op(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    return v;
}

// centroid bounds of each open node (find_best_plane's first loop, build.rs:60-66)
__global__ __launch_bounds__(TOP_THREADS) void k_bounds(BuildArgs a, BuildState s) {
  const uint32_t n_items = *s.n_items;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    Item it = ((const Item *)s.items)[item];
    int mn[3] = {KEY_MIN_ID, KEY_MIN_ID, KEY_MIN_ID}, mx[3] = {KEY_MAX_ID, KEY_MAX_ID, KEY_MAX_ID};
    // all of the thread's (at most PER_THREAD) positions loaded at once: positions past
    // the item read its last element, which the item holds anyway
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t i = it.beg + min((uint32_t)q * TOP_THREADS + threadIdx.x, it.len - 1u);
        for (int d = 0; d < 3; ++d) {
            int key = fkey(a.pc[d * a.n + i]);
            mn[d] = min(mn[d], key);
            mx[d] = max(mx[d], key);
        }
    }
    __shared__ int red[TOP_THREADS / 64][6];
    int w = threadIdx.x / 64, l = threadIdx.x % 64;
    for (int d = 0; d < 3; ++d) {
        int a0 = wave_min(mn[d]), a1 = wave_max(mx[d]);
        if (l == 0) { red[w][d] = a0; red[w][3 + d] = a1; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int v = red[0][threadIdx.x];
        for (int k = 1; k < TOP_THREADS / 64; ++k)
            v = threadIdx.x < 3 ? min(v, red[k][threadIdx.x]) : max(v, red[k][threadIdx.x]);
        int *dst = s.bounds + it.open * 6 + threadIdx.x;
        if (threadIdx.x < 3) atomicMin(dst, v);
        else atomicMax(dst, v);
    }
    __syncthreads();  // red is written again by the next item
  }
}

// bin counts and bin boxes (build.rs:68-85)
// Consecutive primitives of a range are spatial neighbours, so many lanes of a wave fall
// in the same bin and their LDS atomics on one address serialise. The block therefore
// keeps up to BIN_REPL copies of its bins (fewer for short items, whose clearing and
// merging of the copies would cost more than the conflicts), lane l updating copy
// l % copies; copies are
// BIN_WORDS + 1 words apart, so one word of the BIN_REPL copies lies in BIN_REPL
// different banks. A wave whose live lanes all fall in one bin of an axis reduces them
// across the wave first (one LDS update). The copies are merged before the node's global
// bins are updated; every union is an exact min / max, so the order is free.
#ifndef VRGI_BIN_REPL
#define VRGI_BIN_REPL 8
#endif
static constexpr int BIN_REPL = VRGI_BIN_REPL;
static constexpr int BIN_STRIDE = BIN_WORDS + 1;
__device__ void sah_sweep_wave(const int *bins, const int *bounds, uint32_t *dim_out, float *split_out,
                               int box_out[6]);
// SAH sweep of open node o from its bins, by one full wave; the split record of the node
// returns the split axis on every lane; lane 0 also stores the plane to *split_out
__device__ __forceinline__ uint32_t store_split(BuildState &s, uint32_t o, const int *bins,
                                                float *split_out = nullptr) {
    uint32_t dim;
    float split;
    int box[6];
    sah_sweep_wave(bins, s.bounds + o * 6, &dim, &split, box);
    if ((threadIdx.x & 63u) == 0) {
        if (split_out) *split_out = split;
        SplitInfo si;
        si.dim = dim;
        si.split = split;
        si.T = 0;
        si.fallback = 0;
        for (int q = 0; q < 6; ++q) si.box[q] = box[q];
        ((SplitInfo *)s.split)[o] = si;
    }
    return dim;
}
__global__ __launch_bounds__(TOP_THREADS) void k_bin(BuildArgs a, BuildState s) {
  __shared__ int sb[BIN_REPL * BIN_STRIDE];
  const uint32_t n_items = *s.n_items;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    Item it = ((const Item *)s.items)[item];
    const int nrep = it.len >= 1536 ? BIN_REPL : (it.len >= 384 ? min(4, BIN_REPL) : min(2, BIN_REPL));
    for (int k = threadIdx.x; k < nrep * BIN_STRIDE; k += TOP_THREADS) {
        int w = (k % BIN_STRIDE) % 7;
        sb[k] = w == 0 ? 0 : (w <= 3 ? KEY_MIN_ID : KEY_MAX_ID);
    }
    float bmin[3], scale[3];
    for (int d = 0; d < 3; ++d) {
        float lo = funkey(s.bounds[it.open * 6 + d]), hi = funkey(s.bounds[it.open * 6 + 3 + d]);
        bmin[d] = lo;
        scale[d] = (float)BINS / (hi - lo);
    }
    __syncthreads();
    int *mine_copy = sb + (lane % (uint32_t)nrep) * BIN_STRIDE;
    const uint32_t rounds = (it.len + TOP_THREADS - 1) / TOP_THREADS;  // uniform over the block
    // the thread's primitive ids for all its rounds first (one latency), so that each
    // round waits only on its box gather
    uint32_t pp[PER_THREAD];
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q)
        pp[q] = a.perm[it.beg + min((uint32_t)q * TOP_THREADS + threadIdx.x, it.len - 1u)];
#pragma unroll
    for (uint32_t r = 0; r < (uint32_t)PER_THREAD; ++r) {
        if (r >= rounds) break;
        const uint32_t k = r * TOP_THREADS + threadIdx.x;
        const bool valid = k < it.len;
        int bk[6] = {KEY_MIN_ID, KEY_MIN_ID, KEY_MIN_ID, KEY_MAX_ID, KEY_MAX_ID, KEY_MAX_ID};
        uint32_t bb[3] = {0u, 0u, 0u};
        if (valid) {
            const uint32_t i = it.beg + k;
            const uint32_t p = pp[r];
            const float4 b0 = a.box8[2 * p], b1 = a.box8[2 * p + 1];
            bk[0] = fkey(b0.x); bk[1] = fkey(b0.y); bk[2] = fkey(b0.z);
            bk[3] = fkey(b0.w); bk[4] = fkey(b1.x); bk[5] = fkey(b1.y);
            for (int d = 0; d < 3; ++d) bb[d] = bin_index(a.pc[d * a.n + i], bmin[d], scale[d]);
        }
        const uint64_t live = __ballot(valid);
        for (int d = 0; d < 3; ++d) {
            const int lead = __ffsll((unsigned long long)live) - 1;
            const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)bb[d], lead < 0 ? 0 : lead);
            if (live && __ballot(valid && bb[d] != lb) == 0) {  // one bin for the whole wave (uniform)
                const int m0 = wave_min(bk[0]), m1 = wave_min(bk[1]), m2 = wave_min(bk[2]);
                const int x0 = wave_max(bk[3]), x1 = wave_max(bk[4]), x2 = wave_max(bk[5]);
                if ((int)lane == lead) {
                    int *bin = mine_copy + (d * BINS + lb) * 7;
                    atomicAdd(bin, (int)__popcll(live));
                    atomicMin(bin + 1, m0); atomicMin(bin + 2, m1); atomicMin(bin + 3, m2);
                    atomicMax(bin + 4, x0); atomicMax(bin + 5, x1); atomicMax(bin + 6, x2);
                }
            } else if (valid) {
                int *bin = mine_copy + (d * BINS + bb[d]) * 7;
                atomicAdd(bin, 1);
                atomicMin(bin + 1, bk[0]); atomicMin(bin + 2, bk[1]); atomicMin(bin + 3, bk[2]);
                atomicMax(bin + 4, bk[3]); atomicMax(bin + 5, bk[4]); atomicMax(bin + 6, bk[5]);
            }
        }
    }
    __syncthreads();
    int *g = (int *)s.bins + (size_t)it.open * BIN_WORDS;
    for (int k = threadIdx.x; k < BIN_WORDS; k += TOP_THREADS) {
        const int w = k % 7;
        int v = sb[k];
        for (int c = 1; c < nrep; ++c) {
            const int u = sb[c * BIN_STRIDE + k];
            v = w == 0 ? v + u : (w <= 3 ? min(v, u) : max(v, u));
        }
        sb[k] = v;  // copy 0 holds the merged bins (each word read and written by one thread)
    }
    __syncthreads();
    // A node of one item has its complete bins in LDS: its SAH sweep (build.rs:87-127)
    // runs here. The items of a larger node add their bins to the node's global bins,
    // swept by k_count (by k_split when VRGI_SPLIT_IN_COUNT is 0).
    const uint32_t o = it.open;
    if (s.item_end[o] - s.item_beg[o] > 1u) {
        for (int k = threadIdx.x; k < BIN_WORDS; k += TOP_THREADS) {
            const int bin = k / 7, w = k % 7;
            if (sb[bin * 7] == 0) continue;
            if (w == 0) atomicAdd(g + k, sb[k]);
            else if (w <= 3) atomicMin(g + k, sb[k]);
            else atomicMax(g + k, sb[k]);
        }
    } else if (threadIdx.x < 64) {  // wave 0, all lanes
        store_split(s, o, sb);
    }
    __syncthreads();  // sb is cleared again by the next item
  }
}

// SAH sweep of the nodes of several items, over their global bins
__global__ __launch_bounds__(64) void k_split(BuildState s) {
    const uint32_t n_open = *s.n_open;
    for (uint32_t o = blockIdx.x; o < n_open; o += gridDim.x)
        if (s.item_end[o] - s.item_beg[o] > 1u) store_split(s, o, (const int *)s.bins + (size_t)o * BIN_WORDS);
}

// SAH sweep over the 3 x 31 planes (build.rs:87-127) for one node held as bins in
// `bins` (3*32*7 ints, LDS or global). Executed by one full wave; returns the
// chosen (dim, split) and the node box keys on every lane.
// Bins accumulate in 32-lane scans: forward (lane j holds bin j, giving the left side
// of plane j) and backward (lane j holds bin 31 - j, giving the right side of plane
// 30 - j). Pass 0 scans axes 0 and 1 (one per half) both ways; pass 1 scans axis 2
// forward in the low half and backward in the high half. Each plane's cost is
// area(L) * n(L) + area(R) * n(R) as cost_function (build.rs:287-289) forms it; the
// first minimum in (axis, plane) order wins, as the strict < of the reference's loops.
struct BinScan { int c, k[6]; };
__device__ __forceinline__ BinScan bin_scan(const int *bins, int d, int b) {
    const int *bn = bins + (d * BINS + b) * 7;
    BinScan r;
    r.c = seg32_scan(bn[0], 0, [](int x, int y) { return x + y; });
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        r.k[q] = seg32_scan(bn[1 + q], KEY_MIN_ID, [](int x, int y) { return min(x, y); });
        r.k[3 + q] = seg32_scan(bn[4 + q], KEY_MAX_ID, [](int x, int y) { return max(x, y); });
    }
    return r;
}
__device__ __forceinline__ float side_cost(const BinScan &r) {
    const float a = r.c ? area_keys(r.k[0], r.k[1], r.k[2], r.k[3], r.k[4], r.k[5]) : 0.0f;
    return a * (float)(uint32_t)r.c;
}
__device__ void sah_sweep_wave(const int *bins, const int *bounds, uint32_t *dim_out,
                               float *split_out, int box_out[6]) {
    const int lane = threadIdx.x % 64;
    const int half = lane / 32, j = lane % 32;
    // per lane best candidate: cost bits (costs are >= +0, so their bits order as the
    // floats) and plane index d * 31 + b
    uint32_t best_c = 0xFFFFFFFFu, best_i = 0xFFFFFFFFu;
    auto offer = [&](float cost, uint32_t idx) {
        if (cost < 3.40282347e+38f) {
            const uint32_t cb = __float_as_uint(cost);
            if (cb < best_c || (cb == best_c && idx < best_i)) { best_c = cb; best_i = idx; }
        }
    };
    {  // pass 0: axis `half`
        const BinScan f = bin_scan(bins, half, j), r = bin_scan(bins, half, 31 - j);
        const float x = side_cost(f), y = side_cost(r);
        const float yr = __shfl(y, half * 32 + (j < 31 ? 30 - j : 0), 64);  // right side of plane j
        if (j < 31) offer(x + yr, (uint32_t)(half * 31 + j));
        // the forward prefix at lane 31 is the union of all primitive boxes
#pragma unroll
        for (int q = 0; q < 6; ++q) box_out[q] = __shfl(f.k[q], 31, 64);
    }
    {  // pass 1: axis 2, forward in the low half, backward in the high half
        const BinScan g = bin_scan(bins, 2, half == 0 ? j : 31 - j);
        const float v = side_cost(g);
        const float yr = __shfl(v, j < 31 ? 62 - j : 32, 64);
        if (half == 0 && j < 31) offer(v + yr, (uint32_t)(2 * 31 + j));
    }
    const uint32_t wc = __ockl_wfred_min_u32(best_c);
    const uint32_t wi = __ockl_wfred_min_u32(best_c == wc ? best_i : 0xFFFFFFFFu);
    if (wc == 0xFFFFFFFFu) {  // no cost below f32::MAX: initial values (0, 0.0)
        *dim_out = 0;
        *split_out = 0.0f;
    } else {
        uint32_t d = wi / 31, pl = wi % 31;
        float lo = funkey(bounds[d]), hi = funkey(bounds[3 + d]);
        float sc = (hi - lo) / (float)BINS;
        *dim_out = d;
        *split_out = lo + ((float)pl + 1.0f) * sc;
    }
}

// per-item count of `centroid[dim] < split` (the partition predicate, build.rs:148).
// With VRGI_SPLIT_IN_COUNT, the SAH sweep of a node of several items runs here, in
// wave 0 of every block that counts one of its items, over the node's merged global
// bins (complete once k_bin has ended): each such block derives the same split and
// stores the same record, and no separate split launch sits between k_bin and k_count.
#ifndef VRGI_SPLIT_IN_COUNT
#define VRGI_SPLIT_IN_COUNT 1
#endif
__global__ __launch_bounds__(TOP_THREADS) void k_count(BuildArgs a, BuildState s) {
  __shared__ uint32_t red[TOP_THREADS / 64];
  __shared__ uint32_t s_dim;
  __shared__ float s_split;
  const uint32_t n_items = *s.n_items;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    Item it = ((const Item *)s.items)[item];
    SplitInfo si;
    if (VRGI_SPLIT_IN_COUNT && s.item_end[it.open] - s.item_beg[it.open] > 1u) {
        if (threadIdx.x < 64) {
            const uint32_t dim = store_split(s, it.open, (const int *)s.bins + (size_t)it.open * BIN_WORDS, &s_split);
            if (threadIdx.x == 0) s_dim = dim;
        }
        __syncthreads();
        si.dim = s_dim;
        si.split = s_split;
    } else {
        si = ((const SplitInfo *)s.split)[it.open];
    }
    uint32_t c = 0;
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {  // all of the thread's positions loaded at once
        const uint32_t k = (uint32_t)q * TOP_THREADS + threadIdx.x;
        const float v = a.pc[si.dim * a.n + it.beg + min(k, it.len - 1u)];
        c += k < it.len && v < si.split ? 1u : 0u;
    }
    c = wave_sum(c);
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int k = 0; k < TOP_THREADS / 64; ++k) t += red[k];
        s.item_true[item] = t;
    }
    __syncthreads();  // red is written again by the next item
  }
}

// partition, pass A: every element learns its exclusive count of trues c_j within the
// node; right-region trues publish their position at slot T-1-c_j. The item's count of
// trues before it (and the node's total T, hence the half-split fallback of
// build.rs:152) come from the per-item counts. Element e of an item is read by thread
// e % 256 in row e / 256 (coalesced); its rank among the item's trues is the count of
// earlier (row, wave) groups plus the earlier lanes of its own ballot.
__global__ __launch_bounds__(TOP_THREADS) void k_part_a(BuildArgs a, BuildState s) {
  constexpr int WAVES = TOP_THREADS / 64;
  static_assert(PER_THREAD * WAVES <= 64, "one wave forms the prefix of the group counts");
  __shared__ uint32_t s_cnt[PER_THREAD * WAVES];  // trues per (row, wave), then their exclusive prefix
  __shared__ uint32_t s_off, s_T;
  const uint32_t n_items = *s.n_items;
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const Item it = ((const Item *)s.items)[item];
    const uint32_t o = it.open;
    SplitInfo *sip = (SplitInfo *)s.split + o;
    const uint32_t dim = sip->dim;
    const float split = sip->split;
    const OpenNode on = ((const OpenNode *)s.open)[o];
    if (w == 0) {
        const uint32_t ib = s.item_beg[o], ie = s.item_end[o];
        uint32_t before = 0, total = 0;
        for (uint32_t j = ib + lane; j < ie; j += 64) {
            const uint32_t v = s.item_true[j];
            total += v;
            before += j < item ? v : 0u;
        }
        before = wave_sum(before);
        total = wave_sum(total);
        if (lane == 0) {
            s_off = before;
            s_T = total;
            if (item == ib) {
                sip->T = total;
                sip->fallback = (total == 0 || total == on.count) ? 1u : 0u;  // build.rs:152
            }
        }
    }
    __syncthreads();
    const uint32_t T = s_T;
    if (T == 0 || T == on.count) {  // fallback: no partition (uniform over the block)
        __syncthreads();
        continue;
    }
    uint32_t flags = 0;
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t e = (uint32_t)q * TOP_THREADS + threadIdx.x;
        const bool p = e < it.len && a.pc[dim * a.n + it.beg + e] < split;
        flags |= (p ? 1u : 0u) << q;
        const uint64_t m = __ballot(p);
        if (lane == 0) s_cnt[q * WAVES + w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (w == 0) {  // exclusive prefix of the PER_THREAD * WAVES (<= 64) group counts
        const uint32_t v = lane < (uint32_t)(PER_THREAD * WAVES) ? s_cnt[lane] : 0u;
        uint32_t inc = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(inc, off, 64);
            if (lane >= (uint32_t)off) inc += u;
        }
        if (lane < (uint32_t)(PER_THREAD * WAVES)) s_cnt[lane] = inc - v;
    }
    __syncthreads();
    const uint32_t off = s_off;
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t e = (uint32_t)q * TOP_THREADS + threadIdx.x;
        const bool p = (flags >> q) & 1u;
        const uint64_t m = __ballot(p);
        if (e >= it.len) continue;
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        const uint32_t c = off + s_cnt[q * WAVES + w] + below;
        const uint32_t i = it.beg + e;
        const uint32_t j = i - on.first;
        s.rank_tmp[i] = c | (p ? 0x80000000u : 0u);
        if (p && j >= T) s.pair[on.first + (T - 1u - c)] = i;
    }
    __syncthreads();  // s_cnt, s_off, s_T are written again by the next item
  }
}

// partition, pass B: the r-th left-region false swaps with the r-th true from the right
__global__ __launch_bounds__(TOP_THREADS) void k_part_b(BuildArgs a, BuildState s) {
  const uint32_t n_items = *s.n_items;
  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    Item it = ((const Item *)s.items)[item];
    SplitInfo si = ((const SplitInfo *)s.split)[it.open];
    if (si.fallback) continue;
    const OpenNode on = ((const OpenNode *)s.open)[it.open];
    // the thread's ranks, then the partners of its left-region falses, are fetched for
    // all of its (at most PER_THREAD) positions at once: the swaps below then wait on one
    // dependent load each instead of three
    uint32_t part[PER_THREAD];
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t k = (uint32_t)q * TOP_THREADS + threadIdx.x;
        const uint32_t i = it.beg + k;
        part[q] = k < it.len && i - on.first < si.T ? s.rank_tmp[i] : 0x80000000u;
    }
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t k = (uint32_t)q * TOP_THREADS + threadIdx.x;
        const uint32_t j = it.beg + k - on.first;
        part[q] = (part[q] & 0x80000000u) ? 0xFFFFFFFFu : s.pair[on.first + (j - part[q])];
    }
#pragma unroll
    for (int q = 0; q < PER_THREAD; ++q) {
        const uint32_t i = it.beg + (uint32_t)q * TOP_THREADS + threadIdx.x;
        const uint32_t p = part[q];
        if (p == 0xFFFFFFFFu) continue;
        uint32_t t0 = a.perm[i];
        a.perm[i] = a.perm[p];
        a.perm[p] = t0;
        for (int d = 0; d < 3; ++d) {
            float f = a.pc[d * a.n + i];
            a.pc[d * a.n + i] = a.pc[d * a.n + p];
            a.pc[d * a.n + p] = f;
        }
    }
  }
}

// children of each open node: node records, leaves, pairs, jobs, next open nodes
// slot in a list for every lane with `take`: one atomic per wave (all 64 lanes call it)
__device__ __forceinline__ uint32_t wave_append(uint32_t *counter, bool take) {
    const uint64_t m = __ballot(take);
    if (!m) return 0;
    const uint32_t lane = threadIdx.x & 63u;
    const int lead = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)lane == lead) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, lead, 64);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long v) {
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, m, 64), hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
        v += ((unsigned long long)hi << 32) | lo;
    }
    return v;
}

// children of the open nodes: node records, leaves and pairs, and the next level's open
// nodes and subtree jobs. List order is free (ranks come from counts), so list slots,
// depth maxima and split sums take one atomic per wave.
// Only nodes of trees with some non-zero leaf power enter the level-node list of the power
// pass (every power of another tree is (0 + 0) + 0 = +0, the value written here), and the
// count of levels holding such nodes is raised to include this one.
__global__ void k_children(BuildArgs a, BuildState s, uint32_t lvl, uint32_t *power_levels) {
  const uint32_t n_open = *s.n_open;
  for (uint32_t o0 = blockIdx.x * blockDim.x; o0 < n_open; o0 += gridDim.x * blockDim.x) {  // uniform
    const uint32_t o = o0 + threadIdx.x;
    const bool live = o < n_open;
    OpenNode on{};
    uint32_t base = 0, dep = 0;
    unsigned long long splits = 0;
    uint32_t kind[2] = {0u, 0u};  // 1 = subtree job, 2 = open node
    OpenNode ch[2];
    if (live) {
        on = ((const OpenNode *)s.open)[o];
        const SplitInfo si = ((const SplitInfo *)s.split)[o];
        uint32_t nl = si.fallback ? on.count / 2 : si.T;
        uint32_t nr = on.count - nl;
        base = a.tree_base[on.tree];
        const float pad = 0.0001f;
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) { mn[k] = funkey(si.box[k]) - pad; mx[k] = funkey(si.box[3 + k]) + pad; }
        uint32_t li = 2 * on.rank + 1, ri = 2 * on.rank + 2;
        write_node(a.nodes + base + on.node, li, ri, mn, 0.0f, mx, on.parent);
        dep = on.depth + 1;
        splits = on.count;
        for (int side = 0; side < 2; ++side) {
            uint32_t cf = side ? on.first + nl : on.first;
            uint32_t cc = side ? nr : nl;
            uint32_t cr = side ? on.rank + nl : on.rank + 1;
            uint32_t ci = side ? ri : li;
            if (cc == 1) {
                emit_leaf(a, base + ci, on.node, a.perm[cf]);
            } else if (cc == 2) {
                emit_pair(a, base, ci, cr, on.node, a.perm[cf], a.perm[cf + 1]);
                dep = max(dep, on.depth + 2);
                splits += 2;
            } else {
                ch[side] = OpenNode{cf, cc, cr, ci, on.node, on.tree, on.depth + 1, 0};
                kind[side] = cc <= (uint32_t)a.sub_max ? 1u : (mid_node(a, cc, on.tree) ? 3u : 2u);
            }
        }
    }
    const bool pw = live && a.powered[on.tree] != 0u;
    const uint32_t ln = wave_append(s.n_level_nodes, pw);
    if (__ballot(pw) && (threadIdx.x & 63u) == 0u) atomicMax(power_levels, lvl + 1u);
    if (pw) {
        s.level_nodes[2 * ln] = base + on.node;
        s.level_nodes[2 * ln + 1] = base;
    }
    for (int side = 0; side < 2; ++side) {
        const uint32_t j = wave_append(s.n_jobs, kind[side] == 1u);
        if (kind[side] == 1u) ((OpenNode *)s.jobs)[j] = ch[side];
        const uint32_t q = wave_append(s.n_open_next, kind[side] == 2u);
        if (kind[side] == 2u) ((OpenNode *)s.open_next)[q] = ch[side];
        const uint32_t m = wave_append(s.n_mid, kind[side] == 3u);
        if (kind[side] == 3u) ((OpenNode *)s.mid)[m] = ch[side];
    }
    // depth per tree: one atomic per (wave, tree present in the wave)
    bool pend = live;
    for (;;) {
        const uint64_t m = __ballot(pend);
        if (!m) break;
        const int lead = __ffsll((unsigned long long)m) - 1;
        const uint32_t t = __shfl(on.tree, lead, 64);
        const bool mine = pend && on.tree == t;
        const int d = wave_max(mine ? (int)dep : 0);
        if ((int)(threadIdx.x & 63u) == lead) atomicMax(a.depth + t, (uint32_t)d);
        if (mine) pend = false;
    }
    const unsigned long long sp = wave_sum64(splits);
    if ((threadIdx.x & 63u) == 0u && sp) atomicAdd(a.split_count_sum, sp);
  }
}

// ---------------------------------------------------------------- bottom phase
// One 64-lane wave builds a whole subtree of <= SUB primitives held in LDS.
// Phase A: ranges larger than SMALL are split one at a time by the whole wave (an LDS
// stack of pending ranges). Phase B: ranges of 3..SMALL primitives are split by three
// lanes each (one per axis), 21 ranges at a time, one subtree level per pass. Phase C: power of every internal
// node of the subtree, children before parents: internal nodes of a subtree have
// consecutive preorder ranks and descendants have larger ranks, so one pass over
// decreasing ranks suffices (build.rs:345-357).
struct RankInfo {  // per internal node of the subtree, indexed by rank - job rank
    uint16_t first, nl;
    uint16_t count, pad;
    uint32_t gidx;
};

// Tuning of the bottom phase: ranges of at most VRGI_SUB_SMALL primitives are split by
// one lane each; VRGI_PLANE_UNROLL is the unroll factor of the per-plane union loop.
#ifndef VRGI_SUB_SMALL
#define VRGI_SUB_SMALL 12
#endif
#ifndef VRGI_PLANE_UNROLL
#define VRGI_PLANE_UNROLL 1
#endif
#define VRGI_PRAGMA(x) _Pragma(#x)

// SAH plane search (build.rs:45-128) of one range of at most 12 primitives along one
// axis d, by one lane. The primitives are put in bin order once (a counting sort held in
// registers: 5-bit bins and 4-bit per-bin counts in 64-bit words, prefix sums of the
// counts by one multiply since no partial sum exceeds 12). A sweep from the highest bin
// down stores the right-side cost of each occupied bin's plane in tmp[f..f+occupied)
// (descending order); a sweep from bin 0 up forms the left side and evaluates the planes
// in ascending order with a strict < (the first minimum). Each primitive's box is read
// once per sweep; the unions are exact min/max, so every cost equals the per-plane
// evaluation's bit for bit.
template <int SUB>
__device__ void axis_best_plane(const float (*box)[SUB], float *tmp, uint32_t f, uint32_t c, uint32_t d,
                                float *cost_out, float *split_out, float nb[6]) {
    for (int q = 0; q < 3; ++q) { nb[q] = box[q][f]; nb[3 + q] = box[3 + q][f]; }
    for (uint32_t k = 1; k < c; ++k)
        for (int q = 0; q < 3; ++q) { nb[q] = gmin(nb[q], box[q][f + k]); nb[3 + q] = gmax(nb[3 + q], box[3 + q][f + k]); }
    const float all_cost = area6(nb) * (float)c;
    float lo = box[6 + d][f], hi = lo;
    for (uint32_t k = 1; k < c; ++k) { lo = gmin(lo, box[6 + d][f + k]); hi = gmax(hi, box[6 + d][f + k]); }
    const float scale = (float)BINS / (hi - lo);
    uint64_t pk = 0, cnt_lo = 0, cnt_hi = 0;  // bins; per-bin counts of bins 0-15 and 16-31
    uint32_t occ = 0;
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t b = bin_index(box[6 + d][f + k], lo, scale);
        pk |= (uint64_t)b << (5 * k);
        occ |= 1u << b;
        if (b < 16) cnt_lo += 1ull << (4 * b);
        else cnt_hi += 1ull << (4 * (b - 16));
    }
    const uint64_t ones = 0x1111111111111111ull;
    const uint64_t inc_lo = cnt_lo * ones;                                // counts of bins <= b
    const uint64_t inc_hi = cnt_hi * ones + (inc_lo >> 60) * ones;
    uint64_t ord = 0, run_lo = 0, run_hi = 0;  // ord: primitive of each sorted position, 4 bits each
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t b = (uint32_t)((pk >> (5 * k)) & 31u);
        uint32_t pos;
        if (b < 16) {
            const uint32_t sh = 4 * b;
            pos = (uint32_t)((inc_lo >> sh) & 15u) - (uint32_t)((cnt_lo >> sh) & 15u) + (uint32_t)((run_lo >> sh) & 15u);
            run_lo += 1ull << sh;
        } else {
            const uint32_t sh = 4 * (b - 16);
            pos = (uint32_t)((inc_hi >> sh) & 15u) - (uint32_t)((cnt_hi >> sh) & 15u) + (uint32_t)((run_hi >> sh) & 15u);
            run_hi += 1ull << sh;
        }
        ord |= (uint64_t)k << (4 * pos);
    }
    auto kth = [&](uint32_t i) { return (uint32_t)((ord >> (4 * i)) & 15u); };
    auto bin_of = [&](uint32_t k) { return (uint32_t)((pk >> (5 * k)) & 31u); };
    float dbest = 3.40282347e+38f;
    uint32_t dplane = 0;
    if (!(occ & 1u)) { dbest = 0.0f * 0.0f + all_cost; dplane = 0; }  // left side empty
    auto add = [&](float (&u)[6], uint32_t n, uint32_t e) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float vn = box[q][e], vx = box[3 + q][e];
            u[q] = n == 0 ? vn : gmin(u[q], vn);
            u[3 + q] = n == 0 ? vx : gmax(u[3 + q], vx);
        }
    };
    {  // right sides, highest bin first: the plane at bin b has bins > b on the right
        float r[6] = {0, 0, 0, 0, 0, 0};
        uint32_t nr = 0, j = 0, prev = 0xFFFFFFFFu;
        for (int i = (int)c - 1; i >= 0; --i) {
            const uint32_t k = kth((uint32_t)i), b = bin_of(k);
            if (b != prev) { tmp[f + j++] = nr ? area6(r) * (float)nr : 0.0f * 0.0f; prev = b; }
            add(r, nr, f + k);
            ++nr;
        }
    }
    {  // left sides and plane costs, bin 0 up (no plane after bin 31)
        float l[6] = {0, 0, 0, 0, 0, 0};
        const uint32_t nocc = (uint32_t)__builtin_popcount(occ);
        uint32_t nl = 0, j = 0;
        for (uint32_t i = 0; i < c; ++i) {
            const uint32_t k = kth(i), b = bin_of(k);
            add(l, nl, f + k);
            ++nl;
            if (i + 1 == c || bin_of(kth(i + 1)) != b) {  // last primitive of bin b
                if (b < 31) {
                    const float al = area6(l) * (float)nl;
                    const float ar = tmp[f + (nocc - 1u - j)];
                    const float cost = al + ar;
                    if (cost < dbest) { dbest = cost; dplane = b; }
                }
                ++j;
            }
        }
    }
    *cost_out = dbest;
    *split_out = lo + ((float)dplane + 1.0f) * ((hi - lo) / (float)BINS);
}

// POWER: the variant for the jobs of trees with non-zero leaf power (powers in LDS, phase
// C). The other variant keeps no power, id or rank planes and shares the bin area with
// the partition and plane-search scratch: less LDS per wave, more waves per CU. Both are
// launched over the job list; each takes the jobs of its kind.
template <int SUB, int SMALL, bool POWER>
__global__ __launch_bounds__(64) void k_subtree(BuildArgs a, BuildState s) {
    constexpr int STK = SUB / (SMALL + 1) + 2;  // pending ranges are disjoint and hold > SMALL
    constexpr int SMQ = SUB / 3 + 2;            // small ranges are disjoint and hold >= 3
    constexpr int NBOX = POWER ? 11 : 9;
    static_assert(SUB <= 1024 && SMALL <= 127, "small-range queue packs 10 + 7 + 10 bits");
    static_assert(SMALL <= 12, "bins of a small range are packed 5 bits each into 64 bits");
    static_assert(3 * SUB <= BIN_WORDS, "partition and plane-search scratch share the bin area");
    __shared__ uint32_t s_perm[SUB];
    __shared__ float s_box[NBOX][SUB];  // min xyz, max xyz, centroid xyz (, power, id bits)
    // bins of phase A; after each sweep the partition's ranks and pairs; in phase B the
    // three axes' plane-search scratch
    __shared__ int s_bins[BIN_WORDS];
    __shared__ RankInfo s_rank[POWER ? SUB : 1];
    __shared__ float s_ipow[POWER ? SUB : 1];
    __shared__ uint32_t s_stack[STK][6];  // first, count, rank, node, depth, parent
    __shared__ uint4 s_q[2][SMQ];         // small ranges: first|count<<10|(rank-job rank)<<17, node, depth, parent
    __shared__ uint32_t s_nsmall;
    uint32_t *const s_pair = (uint32_t *)s_bins;
    uint32_t *const s_cj = s_pair + SUB;
    float *const s_tmp2 = (float *)(s_cj + SUB);

    const uint32_t job = blockIdx.x;
    if (job >= *s.n_jobs) return;
    const OpenNode jn = ((const OpenNode *)s.jobs)[job];
    if ((a.powered[jn.tree] != 0u) != POWER) return;  // the other variant's job
    const int lane = threadIdx.x;
    const uint32_t c = jn.count;
    const uint32_t base = a.tree_base[jn.tree];
    for (uint32_t k = lane; k < c; k += 64) {
        uint32_t p = a.perm[jn.first + k];
        s_perm[k] = p;
        const float4 b0 = a.box8[2 * p], b1 = a.box8[2 * p + 1];
        s_box[0][k] = b0.x; s_box[1][k] = b0.y; s_box[2][k] = b0.z;
        s_box[3][k] = b0.w; s_box[4][k] = b1.x; s_box[5][k] = b1.y;
        for (int q = 0; q < 3; ++q) s_box[6 + q][k] = a.pc[q * a.n + jn.first + k];
        if constexpr (POWER) { s_box[NBOX - 2][k] = b1.z; s_box[NBOX - 1][k] = b1.w; }
    }
    uint32_t sp = 0, maxdepth = jn.depth;
    uint32_t splitsA = 0, splitsB = 0;  // split sizes (sum = sum of leaf depths): wave-wide, per lane
    if (lane == 0) {
        uint32_t e[6] = {0, c, jn.rank, jn.node, jn.depth, jn.parent};
        if (c > (uint32_t)SMALL)
            for (int q = 0; q < 6; ++q) s_stack[0][q] = e[q];
        else
            s_q[0][0] = make_uint4(c << 10, jn.node, jn.depth, jn.parent);
        s_nsmall = c > (uint32_t)SMALL ? 0u : 1u;
    }
    if (c > (uint32_t)SMALL) sp = 1;
    __syncthreads();
    const float pad = 0.0001f;
    // leaves and 2-primitive nodes (build.rs:311-326, 203-230) from the LDS copy of the
    // primitives at job positions e, e + 1
    // power and id bits of the primitive at job position e (from LDS, or its record)
    auto pw_id = [&](uint32_t e) -> float2 {
        if constexpr (POWER) return make_float2(s_box[NBOX - 2][e], s_box[NBOX - 1][e]);
        const float4 b1 = a.box8[2 * s_perm[e] + 1];
        return make_float2(b1.z, b1.w);
    };
    auto lds_leaf = [&](uint32_t gnode, uint32_t parent, uint32_t e) {
        float mn[3] = {s_box[0][e], s_box[1][e], s_box[2][e]};
        float mx[3] = {s_box[3][e], s_box[4][e], s_box[5][e]};
        const float2 pi = pw_id(e);
        write_node(a.nodes + gnode, VRGI_INVALID, __float_as_uint(pi.y), mn, pi.x, mx, parent);
    };
    auto lds_pair = [&](uint32_t idx, uint32_t r, uint32_t parent, uint32_t e) {
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = gmin(s_box[k][e], s_box[k][e + 1]) - pad;
            mx[k] = gmax(s_box[3 + k][e], s_box[3 + k][e + 1]) + pad;
        }
        float pw = (0.0f + pw_id(e).x) + pw_id(e + 1).x;
        write_node(a.nodes + base + idx, 2 * r + 1, 2 * r + 2, mn, pw, mx, parent);
        lds_leaf(base + 2 * r + 1, idx, e);
        lds_leaf(base + 2 * r + 2, idx, e + 1);
    };
    // ---------------- phase A: ranges larger than SMALL, one wave per range
    while (sp > 0) {
        --sp;
        const uint32_t nf = s_stack[sp][0], nc = s_stack[sp][1], rank = s_stack[sp][2];
        const uint32_t gidx = s_stack[sp][3], dep = s_stack[sp][4], parent = s_stack[sp][5];
        int bmn[3] = {KEY_MIN_ID, KEY_MIN_ID, KEY_MIN_ID}, bmx[3] = {KEY_MAX_ID, KEY_MAX_ID, KEY_MAX_ID};
        for (uint32_t k = lane; k < nc; k += 64)
            for (int d = 0; d < 3; ++d) {
                int key = fkey(s_box[6 + d][nf + k]);
                bmn[d] = min(bmn[d], key);
                bmx[d] = max(bmx[d], key);
            }
        int bounds[6];
        for (int d = 0; d < 3; ++d) { bounds[d] = wave_min(bmn[d]); bounds[3 + d] = wave_max(bmx[d]); }
        for (int k = lane; k < BIN_WORDS; k += 64) {
            int w = k % 7;
            s_bins[k] = w == 0 ? 0 : (w <= 3 ? KEY_MIN_ID : KEY_MAX_ID);
        }
        float lo[3], sc[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = funkey(bounds[d]);
            sc[d] = (float)BINS / (funkey(bounds[3 + d]) - lo[d]);
        }
        __syncthreads();
        for (uint32_t k = lane; k < nc; k += 64) {
            uint32_t e = nf + k;
            int bk[6];
            for (int q = 0; q < 6; ++q) bk[q] = fkey(s_box[q][e]);
            for (int d = 0; d < 3; ++d) {
                uint32_t b = bin_index(s_box[6 + d][e], lo[d], sc[d]);
                int *bin = s_bins + (d * BINS + b) * 7;
                atomicAdd(bin, 1);
                atomicMin(bin + 1, bk[0]); atomicMin(bin + 2, bk[1]); atomicMin(bin + 3, bk[2]);
                atomicMax(bin + 4, bk[3]); atomicMax(bin + 5, bk[4]); atomicMax(bin + 6, bk[5]);
            }
        }
        __syncthreads();
        uint32_t dim;
        float split;
        int box[6];
        sah_sweep_wave(s_bins, bounds, &dim, &split, box);
        __syncthreads();  // the bin area becomes the partition scratch
        const uint32_t iters = (nc + 63) / 64;
        uint32_t T = 0;
        for (uint32_t itn = 0; itn < iters; ++itn) {
            uint32_t k = itn * 64 + lane;
            bool p = k < nc && s_box[6 + dim][nf + k] < split;
            T += (uint32_t)__popcll(__ballot(p));
        }
        const bool fb = (T == 0 || T == nc);
        if (!fb) {
            uint32_t run = 0;
            for (uint32_t itn = 0; itn < iters; ++itn) {
                uint32_t k = itn * 64 + lane;
                bool p = k < nc && s_box[6 + dim][nf + k] < split;
                uint64_t bal = __ballot(p);
                uint32_t cj = run + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                if (k < nc) {
                    s_cj[k] = cj | (p ? 0x80000000u : 0u);
                    if (p && k >= T) s_pair[T - 1 - cj] = k;
                }
                run += (uint32_t)__popcll(bal);
            }
            __syncthreads();
            for (uint32_t k = lane; k < T; k += 64) {
                uint32_t cj = s_cj[k];
                if (cj & 0x80000000u) continue;
                uint32_t e0 = nf + k, e1 = nf + s_pair[k - cj];
                uint32_t tp = s_perm[e0]; s_perm[e0] = s_perm[e1]; s_perm[e1] = tp;
                for (int q = 0; q < NBOX; ++q) { float f = s_box[q][e0]; s_box[q][e0] = s_box[q][e1]; s_box[q][e1] = f; }
            }
            __syncthreads();
        }
        const uint32_t nl = fb ? nc / 2 : T, nr = nc - nl;
        const uint32_t li = 2 * rank + 1, ri = 2 * rank + 2;
        if (lane == 0) {
            float mn[3], mx[3];
            for (int k = 0; k < 3; ++k) { mn[k] = funkey(box[k]) - pad; mx[k] = funkey(box[3 + k]) + pad; }
            write_node(a.nodes + base + gidx, li, ri, mn, 0.0f, mx, parent);
            if constexpr (POWER) s_rank[rank - jn.rank] = RankInfo{(uint16_t)nf, (uint16_t)nl, (uint16_t)nc, 0, gidx};
        }
        splitsA += nc;
        maxdepth = max(maxdepth, dep + 1);
        for (int side = 1; side >= 0; --side) {
            uint32_t cf = side ? nf + nl : nf;
            uint32_t cc = side ? nr : nl;
            uint32_t cr = side ? rank + nl : rank + 1;
            uint32_t ci = side ? ri : li;
            if (cc == 1) {
                if (lane == 0) lds_leaf(base + ci, gidx, cf);
            } else if (cc == 2) {
                if (lane == 0) {
                    lds_pair(ci, cr, gidx, cf);
                    if constexpr (POWER) s_rank[cr - jn.rank] = RankInfo{(uint16_t)cf, 1, 2, 0, ci};
                }
                splitsA += 2;
                maxdepth = max(maxdepth, dep + 2);
            } else {
                uint32_t e[6] = {cf, cc, cr, ci, dep + 1, gidx};
                if (cc > (uint32_t)SMALL) {
                    if (lane == 0)
                        for (int q = 0; q < 6; ++q) s_stack[sp][q] = e[q];
                    ++sp;
                } else if (lane == 0) {
                    s_q[0][s_nsmall++] = make_uint4(cf | (cc << 10) | ((cr - jn.rank) << 17), ci, dep + 1, gidx);
                }
            }
        }
        __syncthreads();
    }
    // ---------------- phase B: ranges of 3..SMALL primitives, three lanes each (one per
    // axis; the first leads), one level of the subtree per pass; the ranges of a pass are
    // disjoint, so at most SUB/3
    uint32_t cur = 0, ncur = s_nsmall;
    while (ncur > 0) {
        __syncthreads();
        if (lane == 0) s_nsmall = 0;
        __syncthreads();
        for (uint32_t w0 = 0; w0 < ncur; w0 += 21) {
            const uint32_t w = w0 + (uint32_t)lane / 3u, ax = (uint32_t)lane % 3u;
            const bool act = lane < 63 && w < ncur;
            uint4 e = make_uint4(0, 0, 0, 0);
            float acost = 0.0f, asplit = 0.0f, nb[6];
            if (act) {
                e = s_q[cur][w];
                float *tmp = ax == 0 ? (float *)s_pair : ax == 1 ? (float *)s_cj : s_tmp2;
                axis_best_plane<SUB>(s_box, tmp, e.x & 0x3ffu, (e.x >> 10) & 0x7fu, ax, &acost, &asplit, nb);
            }
            const float c1 = __shfl(acost, lane + 1, 64), c2 = __shfl(acost, lane + 2, 64);
            const float s1 = __shfl(asplit, lane + 1, 64), s2 = __shfl(asplit, lane + 2, 64);
            if (!act || ax != 0) continue;
            const uint32_t nf = e.x & 0x3ffu, nc = (e.x >> 10) & 0x7fu, rank = jn.rank + (e.x >> 17);
            const uint32_t gidx = e.y, dep = e.z, parent = e.w;
            uint32_t dim = 0;
            float split = 0.0f, best_c = 3.40282347e+38f;
            if (acost < best_c) { best_c = acost; dim = 0; split = asplit; }
            if (c1 < best_c) { best_c = c1; dim = 1; split = s1; }
            if (c2 < best_c) { best_c = c2; dim = 2; split = s2; }
            // partition (build.rs:145): the crate's l/r scan, in place
            uint32_t l = 0, r = nc - 1, T;
            for (;;) {
                while (l < nc && s_box[6 + dim][nf + l] < split) ++l;
                while (r > 0 && !(s_box[6 + dim][nf + r] < split)) --r;
                if (l >= r) { T = l; break; }
                const uint32_t e0 = nf + l, e1 = nf + r;
                const uint32_t tp = s_perm[e0]; s_perm[e0] = s_perm[e1]; s_perm[e1] = tp;
#pragma unroll
                for (int q = 0; q < NBOX; ++q) { float f = s_box[q][e0]; s_box[q][e0] = s_box[q][e1]; s_box[q][e1] = f; }
            }
            const uint32_t nl = (T == 0 || T == nc) ? nc / 2 : T, nr = nc - nl;
            const uint32_t li = 2 * rank + 1, ri = 2 * rank + 2;
            float mn[3], mx[3];
            for (int k = 0; k < 3; ++k) { mn[k] = nb[k] - pad; mx[k] = nb[3 + k] + pad; }
            write_node(a.nodes + base + gidx, li, ri, mn, 0.0f, mx, parent);
            if constexpr (POWER) s_rank[rank - jn.rank] = RankInfo{(uint16_t)nf, (uint16_t)nl, (uint16_t)nc, 0, gidx};
            splitsB += nc;
            maxdepth = max(maxdepth, dep + 1);
            for (int side = 1; side >= 0; --side) {
                const uint32_t cf = side ? nf + nl : nf;
                const uint32_t cc = side ? nr : nl;
                const uint32_t cr = side ? rank + nl : rank + 1;
                const uint32_t ci = side ? ri : li;
                if (cc == 1) {
                    lds_leaf(base + ci, gidx, cf);
                } else if (cc == 2) {
                    lds_pair(ci, cr, gidx, cf);
                    if constexpr (POWER) s_rank[cr - jn.rank] = RankInfo{(uint16_t)cf, 1, 2, 0, ci};
                    splitsB += 2;
                    maxdepth = max(maxdepth, dep + 2);
                } else {
                    const uint32_t slot = atomicAdd(&s_nsmall, 1u);
                    s_q[cur ^ 1][slot] = make_uint4(cf | (cc << 10) | ((cr - jn.rank) << 17), ci, dep + 1, gidx);
                }
            }
        }
        __syncthreads();
        ncur = s_nsmall;
        cur ^= 1;
    }
    __syncthreads();
    // ---------------- phase C: power, decreasing rank (children first); the left child of
    // rank k, when internal, is rank k + 1, the value computed one step earlier. With every
    // leaf power zero (the geometry trees) each sum is +0, the value the nodes were
    // written with, and the pass is skipped.
    bool any_power = false;
    if constexpr (POWER)
        for (uint32_t k = lane; k < c; k += 64) any_power |= s_box[NBOX - 2][k] != 0.0f;
    const bool powered = __any(any_power);
    if constexpr (POWER) if (lane == 0 && powered) {
        float prev = 0.0f;
        for (int k = (int)c - 2; k >= 0; --k) {
            const RankInfo ri = s_rank[k];
            const uint32_t nr = (uint32_t)ri.count - ri.nl;
            const float pl = ri.nl == 1 ? s_box[NBOX - 2][ri.first] : prev;
            const float pr = nr == 1 ? s_box[NBOX - 2][ri.first + ri.nl] : s_ipow[k + ri.nl];
            float p = 0.0f;
            p += pl;
            p += pr;
            s_ipow[k] = p;
            prev = p;
            a.nodes[base + ri.gidx].power = p;
        }
    }
    // depth and leaf-depth sum: reduce over lanes
    maxdepth = (uint32_t)wave_max((int)maxdepth);
    const unsigned long long splits = (unsigned long long)splitsA + wave_sum(splitsB);
    if (lane == 0) {
        atomicMax(a.depth + jn.tree, maxdepth);
        atomicAdd(a.split_count_sum, splits);
    }
    __syncthreads();
    for (uint32_t k = lane; k < c; k += 64) a.perm[jn.first + k] = s_perm[k];
}

// ---------------------------------------------------------------- host driver
// The workspace regions of a build of up to n primitives in `trees` trees, in one
// layout for both owners: hipMalloc'ed one by one (reserve) or carved, 256-byte
// aligned, from a caller's region (bind). `carve` receives (pointer slot, bytes).
template <class F>
static void workspace_layout(uint32_t n, uint32_t trees, uint32_t &open_cap, uint32_t *&perm, float *&pc,
                             float4 *&box8, uint32_t *&counters, uint32_t *&level_start, uint32_t *&powered,
                             BuildState &st, F carve) {
    open_cap = n / SUB_MAX_DEFAULT + trees + 16;
    const uint32_t item_cap = n / CHUNK + open_cap + 16;
    carve((void **)&perm, (size_t)n * 4);
    carve((void **)&pc, (size_t)n * 12);
    carve((void **)&box8, (size_t)n * 32);
    carve((void **)&st.rank_tmp, (size_t)n * 4);
    carve((void **)&st.pair, (size_t)n * 4);
    carve((void **)&st.open, (size_t)open_cap * sizeof(OpenNode));
    carve((void **)&st.open_next, (size_t)open_cap * sizeof(OpenNode));
    carve((void **)&st.jobs, (size_t)(n / 3 + trees + 16) * sizeof(OpenNode));
    carve((void **)&st.mid, (size_t)open_cap * sizeof(OpenNode));
    carve((void **)&st.items, (size_t)item_cap * sizeof(Item));
    carve((void **)&st.item_true, (size_t)item_cap * 4);
    carve((void **)&st.item_beg, (size_t)open_cap * 4);
    carve((void **)&st.item_end, (size_t)open_cap * 4);
    carve((void **)&st.bins, (size_t)open_cap * BIN_WORDS * 4);
    carve((void **)&st.bounds, (size_t)open_cap * 6 * 4);
    carve((void **)&st.split, (size_t)open_cap * sizeof(SplitInfo));
    carve((void **)&st.level_nodes, (size_t)n * 8 + 64);
    carve((void **)&counters, 64 * 4);
    carve((void **)&powered, (size_t)trees * 4);
    carve((void **)&level_start, (size_t)BUILDER_LEVEL_CAP * 4);
    st.n_open = counters + 0;
    st.n_open_next = counters + 1;
    st.n_jobs = counters + 2;
    st.n_items = counters + 3;
    st.n_level_nodes = counters + 4;
    st.n_mid = counters + 7;
}

size_t Builder::workspace_bytes(uint32_t n, uint32_t n_trees) {
    n = std::max(n, 16u);
    n_trees = std::max(n_trees, 1u);
    size_t total = 0;
    uint32_t oc;
    uint32_t *pe, *co, *ls, *pw;
    float *pc_;
    float4 *b8;
    BuildState st_;
    workspace_layout(n, n_trees, oc, pe, pc_, b8, co, ls, pw, st_, [&](void **, size_t bytes) {
        total += (bytes + 255) / 256 * 256;
    });
    return total;
}

void Builder::bind(void *ws, size_t bytes, uint32_t n_max, uint32_t trees_max) {
    free_all();
    n_max = std::max(n_max, 16u);
    trees_max = std::max(trees_max, 1u);
    if (workspace_bytes(n_max, trees_max) > bytes) throw HipError("builder workspace region too small");
    char *cur = (char *)ws;
    workspace_layout(n_max, trees_max, open_capacity, perm, pc, box8, counters, level_start, powered, st,
                     [&](void **slot, size_t b) {
                         *slot = cur;
                         cur += (b + 255) / 256 * 256;
                     });
    cap_n = n_max;
    cap_t = trees_max;
    level_cap = BUILDER_LEVEL_CAP;
    external = true;
}

void Builder::reserve(uint32_t n, uint32_t n_trees) {
    if (n <= cap_n && n_trees <= cap_t) return;
    if (external) throw HipError("build larger than the bound workspace");
    const bool regrow = cap_n != 0;  // 25 % headroom when a workspace grows again
    free_all();
    cap_n = std::max(regrow ? n + n / 4 : n, 16u);
    cap_t = std::max(regrow ? n_trees + n_trees / 4 : n_trees, 1u);
    workspace_layout(cap_n, cap_t, open_capacity, perm, pc, box8, counters, level_start, powered, st,
                     [](void **p, size_t bytes) { dev_alloc(p, bytes); });
    level_cap = BUILDER_LEVEL_CAP;
}

void Builder::free_all() {
    if (!external) {
        void *ptrs[] = {perm, pc, box8, st.rank_tmp, st.pair, st.open, st.open_next, st.jobs, st.items,
                        st.item_true, st.item_beg, st.item_end, st.bins, st.bounds,
                        st.split, st.level_nodes, counters, level_start, powered};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
    perm = nullptr; pc = nullptr; box8 = nullptr; counters = nullptr; level_start = nullptr; powered = nullptr;
    st = BuildState{};
    cap_n = cap_t = 0;
    external = false;
}

Builder::~Builder() { free_all(); }

// ---------------------------------------------------------------- middle phase
// Nodes of zero-power trees holding SUB_MAX < n <= MID_MAX primitives are built by one
// workgroup each, with the node's primitive order and centroids resident in LDS, instead
// of by the top phase's per-level kernels over global memory. Level by level, the
// workgroup splits all of the node's ranges larger than SUB_MAX at once, with the top
// phase's arithmetic per range (centroid bounds, 32 bins per axis over the primitive
// boxes, the SAH sweep, the in-place partition of build.rs:145-149, children), every
// thread covering positions of the node and each position knowing its range; ranges of
// at most SUB_MAX primitives become k_subtree jobs. A wave's 64 consecutive positions
// meet at most two ranges (a range holds more than SUB_MAX), so bounds and bins are
// reduced per range across the wave before their LDS updates.
#ifndef VRGI_MID
#define VRGI_MID 1
#endif
static constexpr int MID_MAX = CHUNK;  // a node of the middle phase is one top-phase item
static constexpr int MID_THREADS = 256;
static constexpr int MID_LIST = MID_MAX / (SUB_MAX_DEFAULT + 1) + 1;  // ranges > SUB_MAX of one level are disjoint
// ranges binned together (LDS for their bins), and the node's subtree jobs staged in LDS
// (the rest are appended to the global list one by one): with the partition scratch in
// the bin area, a workgroup holds about 49 KB of LDS, three per CU
#ifndef VRGI_MID_BATCH
#define VRGI_MID_BATCH 4
#endif
static constexpr int MID_BATCH = VRGI_MID_BATCH;
#ifndef VRGI_MID_JOBS_LDS
#define VRGI_MID_JOBS_LDS 128
#endif
static constexpr int MID_JOBS_LDS = VRGI_MID_JOBS_LDS;
static constexpr int MID_SCRATCH = MID_BATCH * BIN_WORDS > MID_MAX ? MID_BATCH * BIN_WORDS : MID_MAX;
static_assert(MID_MAX <= 32767 && MID_LIST < 255, "16-bit positions, 8-bit range ids");
static_assert(MID_MAX % MID_THREADS == 0, "every thread owns MID_MAX / MID_THREADS positions");
__device__ __forceinline__ bool mid_node(const BuildArgs &a, uint32_t count, uint32_t tree) {
    return VRGI_MID && count > (uint32_t)a.sub_max && count <= (uint32_t)MID_MAX && a.powered[tree] == 0u;
}

struct MidRange {
    uint32_t first, count, rank, node, depth, parent;  // first: position in the node
};

// min / max over the lanes whose range id equals the lane's own, for the (at most two)
// distinct ids of the wave; `apply(id, values)` runs on one lane per id
template <class F>
__device__ __forceinline__ void by_range(uint32_t rid, bool live, F apply) {
    uint64_t todo = __ballot(live);
    while (todo) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)rid, lead);
        const bool mine = live && rid == r;
        apply(r, mine, lead);
        todo &= ~__ballot(mine);
    }
}

__global__ __launch_bounds__(MID_THREADS) void k_mid(BuildArgs a, BuildState s) {
    constexpr int PER = MID_MAX / MID_THREADS;
    __shared__ uint32_t s_perm[MID_MAX];
    __shared__ float s_c[3][MID_MAX];       // centroids, permuted with s_perm
    __shared__ uint8_t s_rid[MID_MAX];      // range of each position at this level (0xFF: none)
    // bins of MID_BATCH ranges while binning and sweeping; then the partition's
    // exclusive counts of trues (| predicate bit 15) and its right-region trues of a
    // range, by rank from the right
    __shared__ int s_scratch[MID_SCRATCH];
    int (*const s_bins)[BIN_WORDS] = reinterpret_cast<int (*)[BIN_WORDS]>(s_scratch);
    uint16_t *const s_cj = reinterpret_cast<uint16_t *>(s_scratch);
    uint16_t *const s_pair = s_cj + MID_MAX;
    __shared__ int s_bnd[MID_LIST][6];      // centroid bound keys per range
    __shared__ float s_lo[MID_LIST][3], s_sc[MID_LIST][3];
    __shared__ uint32_t s_dim[MID_LIST], s_T[MID_LIST], s_P0[MID_LIST];
    __shared__ float s_split[MID_LIST];
    __shared__ int s_box[MID_LIST][6];
    __shared__ MidRange s_r[2][MID_LIST];
    __shared__ uint32_t s_n[2];
    __shared__ uint32_t s_wsum[MID_THREADS / 64];
    // the node's first MID_JOBS_LDS subtree jobs, appended to the global list at the end
    // with one counter update
    __shared__ uint4 s_job[MID_JOBS_LDS];  // first | count << 16, rank, node, parent
    __shared__ uint16_t s_jdep[MID_JOBS_LDS];
    __shared__ uint32_t s_nj, s_jbase;
    const OpenNode on = ((const OpenNode *)s.mid)[blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t base = a.tree_base[on.tree];
    const uint32_t c = on.count;
    for (uint32_t k = tid; k < MID_MAX; k += MID_THREADS) {
        if (k < c) {
            s_perm[k] = a.perm[on.first + k];
            for (int q = 0; q < 3; ++q) s_c[q][k] = a.pc[q * a.n + on.first + k];
        }
        s_rid[k] = k < c ? 0 : 0xFF;
    }
    if (tid == 0) {
        s_r[0][0] = MidRange{0u, c, on.rank, on.node, on.depth, on.parent};
        s_n[0] = 1;
        s_n[1] = 0;
        s_nj = 0;
    }
    __syncthreads();
    const float pad = 0.0001f;
    unsigned long long splits = 0;
    uint32_t maxdepth = 0;
    for (uint32_t cur = 0; s_n[cur] != 0u; cur ^= 1u) {
        const uint32_t nr = s_n[cur];
        const MidRange *R = s_r[cur];
        // ---- centroid bounds (build.rs:60-66)
        for (uint32_t k = tid; k < nr * 6; k += MID_THREADS) s_bnd[k / 6][k % 6] = (k % 6) < 3 ? KEY_MIN_ID : KEY_MAX_ID;
        __syncthreads();
        for (int i = 0; i < PER; ++i) {
            const uint32_t k = i * MID_THREADS + tid;
            const uint32_t rid = s_rid[k];
            const bool live = rid != 0xFFu;
            int key[3] = {0, 0, 0};
            if (live)
                for (int d = 0; d < 3; ++d) key[d] = fkey(s_c[d][k]);
            by_range(rid, live, [&](uint32_t r, bool mine, int lead) {
                for (int d = 0; d < 3; ++d) {
                    const int mn = wave_min(mine ? key[d] : KEY_MIN_ID), mx = wave_max(mine ? key[d] : KEY_MAX_ID);
                    if ((int)lane == lead) { atomicMin(&s_bnd[r][d], mn); atomicMax(&s_bnd[r][3 + d], mx); }
                }
            });
        }
        __syncthreads();
        if (tid < nr * 3) {
            const uint32_t r = tid / 3, d = tid % 3;
            const float lo = funkey(s_bnd[r][d]);
            s_lo[r][d] = lo;
            s_sc[r][d] = (float)BINS / (funkey(s_bnd[r][3 + d]) - lo);
        }
        // ---- bins (build.rs:68-85) and the SAH sweep (build.rs:87-127), MID_BATCH ranges at a time
        for (uint32_t rb = 0; rb < nr; rb += MID_BATCH) {
            const uint32_t nb = min((uint32_t)MID_BATCH, nr - rb);
            for (uint32_t k = tid; k < nb * BIN_WORDS; k += MID_THREADS) {
                const uint32_t w = (k % BIN_WORDS) % 7;
                s_bins[k / BIN_WORDS][k % BIN_WORDS] = w == 0 ? 0 : (w <= 3 ? KEY_MIN_ID : KEY_MAX_ID);
            }
            __syncthreads();
            // the thread's PER primitive records are fetched together (one memory latency)
            float4 rb0[PER], rb1[PER];
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const uint32_t k = i * MID_THREADS + tid;
                const uint32_t rid = s_rid[k];
                rb0[i] = rb1[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (rid != 0xFFu && rid >= rb && rid < rb + nb) {
                    const uint32_t p = s_perm[k];
                    rb0[i] = a.box8[2 * p];
                    rb1[i] = a.box8[2 * p + 1];
                }
            }
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const uint32_t k = i * MID_THREADS + tid;
                const uint32_t rid = s_rid[k];
                const bool live = rid != 0xFFu && rid >= rb && rid < rb + nb;
                int bk[6] = {KEY_MIN_ID, KEY_MIN_ID, KEY_MIN_ID, KEY_MAX_ID, KEY_MAX_ID, KEY_MAX_ID};
                uint32_t bb[3] = {0u, 0u, 0u};
                if (live) {
                    const float4 b0 = rb0[i], b1 = rb1[i];
                    bk[0] = fkey(b0.x); bk[1] = fkey(b0.y); bk[2] = fkey(b0.z);
                    bk[3] = fkey(b0.w); bk[4] = fkey(b1.x); bk[5] = fkey(b1.y);
                    for (int d = 0; d < 3; ++d) bb[d] = bin_index(s_c[d][k], s_lo[rid][d], s_sc[rid][d]);
                }
                by_range(rid, live, [&](uint32_t r, bool mine, int lead) {
                    int *bins = s_bins[r - rb];
                    for (int d = 0; d < 3; ++d) {
                        const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)bb[d], lead);
                        if (__ballot(mine && bb[d] != lb) == 0) {  // one bin for the range's lanes
                            const int cnt = (int)__popcll(__ballot(mine));  // outside the one-lane branch
                            const int m0 = wave_min(mine ? bk[0] : KEY_MIN_ID), m1 = wave_min(mine ? bk[1] : KEY_MIN_ID);
                            const int m2 = wave_min(mine ? bk[2] : KEY_MIN_ID), x0 = wave_max(mine ? bk[3] : KEY_MAX_ID);
                            const int x1 = wave_max(mine ? bk[4] : KEY_MAX_ID), x2 = wave_max(mine ? bk[5] : KEY_MAX_ID);
                            if ((int)lane == lead) {
                                int *bin = bins + (d * BINS + lb) * 7;
                                atomicAdd(bin, cnt);
                                atomicMin(bin + 1, m0); atomicMin(bin + 2, m1); atomicMin(bin + 3, m2);
                                atomicMax(bin + 4, x0); atomicMax(bin + 5, x1); atomicMax(bin + 6, x2);
                            }
                        } else if (mine) {
                            int *bin = bins + (d * BINS + bb[d]) * 7;
                            atomicAdd(bin, 1);
                            atomicMin(bin + 1, bk[0]); atomicMin(bin + 2, bk[1]); atomicMin(bin + 3, bk[2]);
                            atomicMax(bin + 4, bk[3]); atomicMax(bin + 5, bk[4]); atomicMax(bin + 6, bk[5]);
                        }
                    }
                });
            }
            __syncthreads();
            for (uint32_t r = rb + wave; r < rb + nb; r += MID_THREADS / 64) {  // one wave per range
                uint32_t dim;
                float split;
                int box[6];
                sah_sweep_wave(s_bins[r - rb], s_bnd[r], &dim, &split, box);
                if (lane == 0) {
                    s_dim[r] = dim;
                    s_split[r] = split;
                    for (int q = 0; q < 6; ++q) s_box[r][q] = box[q];
                }
            }
            __syncthreads();
        }
        // ---- partition predicate c[dim] < split (build.rs:148): exclusive count of trues
        // over the node's positions (thread t owns positions PER t .. PER t + PER - 1)
        {
            uint32_t flags = 0, cnt = 0;
            for (int i = 0; i < PER; ++i) {
                const uint32_t k = tid * PER + i;
                const uint32_t rid = s_rid[k];
                const bool p = rid != 0xFFu && s_c[s_dim[rid]][k] < s_split[rid];
                flags |= (p ? 1u : 0u) << i;
                cnt += p ? 1u : 0u;
            }
            uint32_t inc = cnt;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t u = __shfl_up(inc, off, 64);
                if (lane >= (uint32_t)off) inc += u;
            }
            if (lane == 63) s_wsum[wave] = inc;
            __syncthreads();
            uint32_t run = inc - cnt;
            for (uint32_t w = 0; w < wave; ++w) run += s_wsum[w];
            for (int i = 0; i < PER; ++i) {
                const uint32_t k = tid * PER + i;
                const bool p = (flags >> i) & 1u;
                s_cj[k] = (uint16_t)(run | (p ? 0x8000u : 0u));
                run += p ? 1u : 0u;
            }
        }
        __syncthreads();
        uint32_t total = 0;
        for (uint32_t w = 0; w < MID_THREADS / 64; ++w) total += s_wsum[w];
        if (tid < nr) {
            const uint32_t f = R[tid].first, e = f + R[tid].count;
            const uint32_t p0 = s_cj[f] & 0x7FFFu, p1 = e < c ? (s_cj[e] & 0x7FFFu) : total;
            s_P0[tid] = p0;
            s_T[tid] = p1 - p0;
        }
        __syncthreads();
        // right-region trues publish their position; then the r-th false of [0, T) trades
        // places with the r-th true of [T, n) counted from the right (ranges whose T is 0 or
        // n keep their order: the half split of build.rs:152)
        for (int i = 0; i < PER; ++i) {
            const uint32_t k = i * MID_THREADS + tid;
            const uint32_t rid = s_rid[k];
            if (rid == 0xFFu) continue;
            const MidRange &g = R[rid];
            const uint32_t T = s_T[rid];
            if (T == 0 || T == g.count) continue;
            const uint32_t cj = s_cj[k];
            const uint32_t j = k - g.first;
            if ((cj & 0x8000u) && j >= T) s_pair[g.first + (T - 1 - ((cj & 0x7FFFu) - s_P0[rid]))] = (uint16_t)j;
        }
        __syncthreads();
        for (int i = 0; i < PER; ++i) {
            const uint32_t k = i * MID_THREADS + tid;
            const uint32_t rid = s_rid[k];
            if (rid == 0xFFu) continue;
            const MidRange &g = R[rid];
            const uint32_t T = s_T[rid];
            const uint32_t j = k - g.first;
            if (T == 0 || T == g.count || j >= T) continue;
            const uint32_t cj = s_cj[k];
            if (cj & 0x8000u) continue;
            const uint32_t e1 = g.first + s_pair[g.first + (j - ((cj & 0x7FFFu) - s_P0[rid]))];
            const uint32_t tp = s_perm[k]; s_perm[k] = s_perm[e1]; s_perm[e1] = tp;
            for (int q = 0; q < 3; ++q) { const float f = s_c[q][k]; s_c[q][k] = s_c[q][e1]; s_c[q][e1] = f; }
        }
        __syncthreads();
        // ---- node records and children, one thread per range
        if (tid < nr) {
            const MidRange g = R[tid];
            const uint32_t T = s_T[tid];
            const bool fb = T == 0 || T == g.count;
            const uint32_t nl = fb ? g.count / 2 : T, nrt = g.count - nl;
            const uint32_t li = 2 * g.rank + 1, ri = 2 * g.rank + 2;
            float mn[3], mx[3];
            for (int k = 0; k < 3; ++k) { mn[k] = funkey(s_box[tid][k]) - pad; mx[k] = funkey(s_box[tid][3 + k]) + pad; }
            write_node(a.nodes + base + g.node, li, ri, mn, 0.0f, mx, g.parent);
            splits += g.count;
            maxdepth = max(maxdepth, g.depth + 1);
            for (int side = 0; side < 2; ++side) {
                const uint32_t cf = side ? g.first + nl : g.first;
                const uint32_t cc = side ? nrt : nl;
                const uint32_t cr = side ? g.rank + nl : g.rank + 1;
                const uint32_t ci = side ? ri : li;
                if (cc == 1) {
                    emit_leaf(a, base + ci, g.node, s_perm[cf]);
                } else if (cc == 2) {
                    emit_pair(a, base, ci, cr, g.node, s_perm[cf], s_perm[cf + 1]);
                    splits += 2;
                    maxdepth = max(maxdepth, g.depth + 2);
                } else if (cc <= (uint32_t)a.sub_max) {
                    const uint32_t jb = atomicAdd(&s_nj, 1u);
                    if (jb < (uint32_t)MID_JOBS_LDS) {
                        s_job[jb] = make_uint4(cf | (cc << 16), cr, ci, g.node);
                        s_jdep[jb] = (uint16_t)(g.depth + 1);
                    } else {
                        ((OpenNode *)s.jobs)[atomicAdd(s.n_jobs, 1u)] =
                            OpenNode{on.first + cf, cc, cr, ci, g.node, on.tree, g.depth + 1, 0};
                    }
                } else {
                    const uint32_t q = atomicAdd(&s_n[cur ^ 1u], 1u);
                    s_r[cur ^ 1u][q] = MidRange{cf, cc, cr, ci, g.depth + 1, g.node};
                }
            }
        }
        __syncthreads();
        // ---- range ids of the next level
        for (uint32_t k = tid; k < MID_MAX; k += MID_THREADS) s_rid[k] = 0xFF;
        if (tid == 0) s_n[cur] = 0;
        __syncthreads();
        const uint32_t nn = s_n[cur ^ 1u];
        for (uint32_t q = 0; q < nn; ++q) {
            const MidRange g = s_r[cur ^ 1u][q];
            for (uint32_t k = tid; k < g.count; k += MID_THREADS) s_rid[g.first + k] = (uint8_t)q;
        }
        __syncthreads();
    }
    // the jobs read the node's final primitive order from global memory
    for (uint32_t k = tid; k < c; k += MID_THREADS) {
        a.perm[on.first + k] = s_perm[k];
        for (int q = 0; q < 3; ++q) a.pc[q * a.n + on.first + k] = s_c[q][k];
    }
    const uint32_t nj = min(s_nj, (uint32_t)MID_JOBS_LDS);
    if (tid == 0) s_jbase = nj ? atomicAdd(s.n_jobs, nj) : 0u;
    __syncthreads();
    for (uint32_t k = tid; k < nj; k += MID_THREADS) {
        const uint4 j = s_job[k];
        ((OpenNode *)s.jobs)[s_jbase + k] =
            OpenNode{on.first + (j.x & 0xFFFFu), j.x >> 16, j.y, j.z, j.w, on.tree, s_jdep[k], 0};
    }
    if (maxdepth) atomicMax(a.depth + on.tree, maxdepth);
    if (splits) atomicAdd(a.split_count_sum, splits);
}

// start of a top-phase level: the open list of the previous level's children becomes
// current, and the level's first entry in the level-node list is recorded
__global__ void k_advance(BuildState s, uint32_t *level_start, uint32_t lvl) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *s.n_open = *s.n_open_next;
    *s.n_open_next = 0;
    level_start[lvl] = *s.n_level_nodes;
}

// power pass over the (global node, tree base) pairs of level l
__global__ void k_power_pairs(vrgi_bvh_node *nodes, const uint32_t *pairs, const uint32_t *level_start, uint32_t l) {
    const uint32_t beg = level_start[l], end = level_start[l + 1];
    for (uint32_t i = beg + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x) {
        uint32_t gi = pairs[2 * i], b = pairs[2 * i + 1];
        vrgi_bvh_node nd = nodes[gi];
        float p = 0.0f;
        p += nodes[b + nd.left].power;
        p += nodes[b + nd.right_or_prim].power;
        nodes[gi].power = p;
    }
}

// clears the counters and the per-tree power flags (a kernel, not a memset: a build
// replayed under stream capture may be recorded into a graph)
__global__ __launch_bounds__(256) void k_reset(uint32_t *counters, uint32_t *powered, uint32_t n_trees) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 64) counters[i] = 0u;
    if (i < n_trees) powered[i] = 0u;
}
// a replayed build's counts against its plan (counters after k_advance: [0] nodes open
// past the last level, [2] jobs, [5] non-empty levels, [6] power levels, [7] middle nodes)
__global__ void k_plan_check(const uint32_t *counters, BuildPlan p, uint32_t *check) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const bool bad = counters[0] != 0u || counters[2] != p.n_jobs || counters[5] != p.levels ||
                     min(p.levels, counters[6]) != p.power_levels || counters[7] != p.n_mid;
    if (bad) atomicOr(check, 1u);
}

int Builder::build(hipStream_t stream, const float *prim9, const float *power, const uint32_t *ids,
                   uint32_t n, uint32_t n_trees, const uint32_t *d_tree_beg,
                   const uint32_t *d_tree_n, const uint32_t *d_tree_base,
                   const uint32_t *h_tree_base, vrgi_bvh_node *nodes, uint32_t *d_depth,
                   uint32_t *prim_order_out, bool replay, uint32_t *check) {
    if (n == 0 || n_trees == 0) return 0;
    replay = replay && last_plan.valid && last_plan.n == n && last_plan.trees == n_trees;
    if (!replay) last_plan.valid = false;
    reserve(n, n_trees);
    BuildArgs a{};
    a.prim9 = prim9;
    a.power = power;
    a.ids = ids;
    a.n = n;
    a.n_trees = n_trees;
    a.tree_beg = d_tree_beg;
    a.tree_n = d_tree_n;
    a.tree_base = d_tree_base;
    a.nodes = nodes;
    a.depth = d_depth;
    a.perm = perm;
    a.pc = pc;
    a.box8 = box8;
    a.sub_max = SUB_MAX_DEFAULT;
    a.powered = powered;
    a.split_count_sum = (unsigned long long *)(counters + 8);
    (void)h_tree_base;
    k_reset<<<(std::max(n_trees, 64u) + 255) / 256, 256, 0, stream>>>(counters, powered, n_trees);
    k_init<<<(n + 255) / 256, 256, 0, stream>>>(a);
    k_roots<<<(n_trees + 255) / 256, 256, 0, stream>>>(a, st);
    // Top phase, level-synchronous. Level l holds at most n_trees * 2^l open nodes, so the
    // grids are bounded on the host without reading the device (and capped: blocks stride
    // over the device counts). Levels are queued in groups and the host reads the
    // counters once per group to see whether any node is still open (extra levels are
    // empty). The first group has as many levels as the previous build of the same size
    // had non-empty levels (a frame-by-frame rebuild of a similar scene then takes one
    // readback), later groups 1, 2, 4, 4, ...
    uint32_t group = (prev_n == n && prev_trees == n_trees && prev_levels) ? prev_levels : 1u;
    bool first_group = true;
    auto bound = [&](uint32_t lvl) -> uint32_t {
        uint64_t b = lvl >= 32 ? ~0ull : ((uint64_t)n_trees << lvl);
        return (uint32_t)std::min<uint64_t>(b, open_capacity);
    };
    uint32_t h_cnt[8];
    uint32_t lvl = 0;
    auto queue_level = [&]() {
        std::swap(st.open, st.open_next);
        const uint32_t nb = bound(lvl);
        k_items<<<1, ITEMS_THREADS, 0, stream>>>(st, level_start, lvl, counters + 5);
        const uint32_t node_grid = std::min(nb, GRID_CAP_NODES);
        const uint32_t item_grid = std::min(n / CHUNK + nb + 1, GRID_CAP_ITEMS);
        k_clear_node_data<<<node_grid, 256, 0, stream>>>(st);
        k_bounds<<<item_grid, TOP_THREADS, 0, stream>>>(a, st);
        k_bin<<<item_grid, TOP_THREADS, 0, stream>>>(a, st);
        if (!VRGI_SPLIT_IN_COUNT) k_split<<<node_grid, 64, 0, stream>>>(st);
        k_count<<<item_grid, TOP_THREADS, 0, stream>>>(a, st);
        k_part_a<<<item_grid, TOP_THREADS, 0, stream>>>(a, st);
        k_part_b<<<item_grid, TOP_THREADS, 0, stream>>>(a, st);
        k_children<<<std::min((nb + 127) / 128, GRID_CAP_NODES / 2), 128, 0, stream>>>(a, st, lvl, counters + 6);
    };
    if (replay) {
        // the plan's launch sequence: its levels, middle nodes and jobs, no readback
        const BuildPlan p = last_plan;
        for (lvl = 0; lvl < p.levels; ++lvl) queue_level();
        std::swap(st.open, st.open_next);
        k_advance<<<1, 64, 0, stream>>>(st, level_start, lvl);
        if (p.n_mid) k_mid<<<p.n_mid, MID_THREADS, 0, stream>>>(a, st);
        if (p.n_jobs) {
            k_subtree<SUB_MAX_DEFAULT, VRGI_SUB_SMALL, false><<<p.n_jobs, 64, 0, stream>>>(a, st);
            k_subtree<SUB_MAX_DEFAULT, VRGI_SUB_SMALL, true><<<p.n_jobs, 64, 0, stream>>>(a, st);
        }
        for (int l = (int)p.power_levels - 1; l >= 0; --l) {
            const uint32_t nb = bound((uint32_t)l);
            k_power_pairs<<<std::min((nb + 255) / 256, GRID_CAP_NODES), 256, 0, stream>>>(nodes, st.level_nodes,
                                                                                         level_start, (uint32_t)l);
        }
        if (check) k_plan_check<<<1, 64, 0, stream>>>(counters, p, check);
        if (prim_order_out && prim_order_out != perm)
            HIP_CHECK(hipMemcpyAsync(prim_order_out, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, stream));
        return 0;
    }
    for (;;) {
        for (uint32_t j = 0; j < group; ++j, ++lvl) {
            if (lvl + 2 >= level_cap) return -5;
            queue_level();
        }
        HIP_CHECK(hipMemcpyAsync(h_cnt, counters, 8 * 4, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        if (h_cnt[1] == 0) break;  // no node left open for the next level
        group = first_group ? 1u : std::min(group * 2, 4u);
        first_group = false;
    }
    prev_n = n;
    prev_trees = n_trees;
    prev_levels = h_cnt[5];
    std::swap(st.open, st.open_next);
    k_advance<<<1, 64, 0, stream>>>(st, level_start, lvl);  // level_start[lvl] = end of the list
    const uint32_t n_levels = std::min(lvl, h_cnt[6]);  // levels past h_cnt[6] held no powered node
    uint32_t n_jobs = h_cnt[2];
    if (h_cnt[7]) {  // the middle phase adds jobs: their count is read back
        k_mid<<<h_cnt[7], MID_THREADS, 0, stream>>>(a, st);
        HIP_CHECK(hipMemcpyAsync(h_cnt + 2, counters + 2, 4, hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        n_jobs = h_cnt[2];
    }
    if (n_jobs) {
        k_subtree<SUB_MAX_DEFAULT, VRGI_SUB_SMALL, false><<<n_jobs, 64, 0, stream>>>(a, st);
        k_subtree<SUB_MAX_DEFAULT, VRGI_SUB_SMALL, true><<<n_jobs, 64, 0, stream>>>(a, st);
    }
    // power of the top-phase nodes, deepest level first
    for (int l = (int)n_levels - 1; l >= 0; --l) {
        const uint32_t nb = bound((uint32_t)l);
        k_power_pairs<<<std::min((nb + 255) / 256, GRID_CAP_NODES), 256, 0, stream>>>(nodes, st.level_nodes, level_start,
                                                                                     (uint32_t)l);
    }
    if (prim_order_out && prim_order_out != perm)
        HIP_CHECK(hipMemcpyAsync(prim_order_out, perm, (size_t)n * 4, hipMemcpyDeviceToDevice, stream));
    last_plan.n = n;
    last_plan.trees = n_trees;
    last_plan.levels = h_cnt[5];
    last_plan.power_levels = n_levels;
    last_plan.n_mid = h_cnt[7];
    last_plan.n_jobs = n_jobs;
    last_plan.valid = true;
    return 0;
}

}  // namespace vrgi
