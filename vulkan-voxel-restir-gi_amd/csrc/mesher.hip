// Exposed-face meshing of voxel chunks on the GPU (chunk.rs:158-292), in the CPU
// mesher's vertex order: voxels x-major inside the 32^3 chunk (x, then y, then z), per
// voxel the faces left, right, down, up, back, front, per face the 6 table vertices.
//
// Three passes over (chunk, voxel): a 6-bit mask of exposed faces per voxel (a face is
// exposed when the neighbour is air or translucent, block.rs:129-131; outside the world
// is air); one block per chunk forms the exclusive scan of the face counts in voxel
// order; every voxel then writes its faces at the scanned position. The order is thus
// the sequential loop's order by construction.
#include <rocprim/device/device_scan.hpp>

#include "vrgi_internal.h"
#include "vrgi_kernels.h"

namespace vrgi {

// per face: 6 (corner, u, v) in the reference's push order; corner bit 2 = x, 1 = y, 0 = z
__constant__ int8_t MESH_FACE_C[6][6] = {
    {1, 2, 0, 3, 2, 1}, {6, 5, 4, 6, 7, 5}, {0, 4, 1, 4, 5, 1},
    {3, 6, 2, 3, 7, 6}, {2, 4, 0, 2, 6, 4}, {1, 5, 3, 5, 7, 3}};
__constant__ int8_t MESH_FACE_U[6][6] = {
    {0, 1, 1, 0, 1, 0}, {0, 1, 0, 0, 1, 1}, {0, 1, 0, 1, 1, 0},
    {1, 0, 1, 1, 0, 0}, {0, 1, 0, 0, 1, 1}, {1, 0, 1, 0, 0, 1}};
__constant__ int8_t MESH_FACE_V[6][6] = {
    {1, 0, 1, 0, 0, 1}, {0, 1, 1, 0, 0, 1}, {0, 0, 1, 0, 1, 1},
    {1, 0, 0, 1, 1, 0}, {0, 1, 1, 0, 0, 1}, {1, 1, 0, 1, 0, 0}};

struct MeshArgs {
    const uint16_t *blocks;
    uint32_t dx, dy, dz;
    uint32_t air;                // block id of air (= number of block types)
    const uint8_t *translucent;  // per block type
    const uint32_t *chunks;      // n_chunks x (cx, cy, cz)
    uint32_t n_chunks;
    uint8_t *mask;               // [n_chunks][32768] exposed faces
    uint32_t *face_off;          // [n_chunks][32768] exclusive scan of face counts
    uint32_t *faces;             // [n_chunks] faces of the chunk
    const uint64_t *vert_off;    // [n_chunks] first vertex of the chunk in out
    vrgi_vertex3d *out;
};

__device__ __forceinline__ uint32_t mesh_block(const MeshArgs &a, long x, long y, long z) {
    if (x < 0 || y < 0 || z < 0 || x >= (long)a.dx || y >= (long)a.dy || z >= (long)a.dz) return a.air;
    return a.blocks[((size_t)x * a.dy + (size_t)y) * a.dz + (size_t)z];
}

__global__ __launch_bounds__(256) void k_mesh_mask(MeshArgs a) {
    const uint32_t c = blockIdx.y;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;  // voxel index x * 1024 + y * 32 + z
    if (c >= a.n_chunks || v >= 32768) return;
    const long gx = (long)a.chunks[3 * c] * 32 + (v >> 10), gy = (long)a.chunks[3 * c + 1] * 32 + ((v >> 5) & 31);
    const long gz = (long)a.chunks[3 * c + 2] * 32 + (v & 31);
    const uint32_t b = mesh_block(a, gx, gy, gz);
    uint32_t m = 0;
    if (b != a.air) {
        const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            const uint32_t n = mesh_block(a, gx + nb[f][0], gy + nb[f][1], gz + nb[f][2]);
            if (n == a.air || (n < a.air && a.translucent[n])) m |= 1u << f;
        }
    }
    a.mask[(size_t)c * 32768 + v] = (uint8_t)m;
}

// one block of 1024 threads per chunk, 32 consecutive voxels per thread
__global__ __launch_bounds__(1024) void k_mesh_scan(MeshArgs a) {
    const uint32_t c = blockIdx.x;
    const uint8_t *m = a.mask + (size_t)c * 32768;
    uint32_t *off = a.face_off + (size_t)c * 32768;
    const uint32_t t = threadIdx.x, v0 = t * 32;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < 32; ++k) sum += (uint32_t)__popc(m[v0 + k]);
    __shared__ uint32_t part[1024];
    part[t] = sum;
    __syncthreads();
    for (uint32_t s = 1; s < 1024; s <<= 1) {
        const uint32_t u = t >= s ? part[t - s] : 0u;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t k = 0; k < 32; ++k) {
        off[v0 + k] = run;
        run += (uint32_t)__popc(m[v0 + k]);
    }
    if (t == 1023) a.faces[c] = part[1023];
}

__global__ __launch_bounds__(256) void k_mesh_emit(MeshArgs a) {
    const uint32_t c = blockIdx.y;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_chunks || v >= 32768) return;
    const uint32_t m = a.mask[(size_t)c * 32768 + v];
    if (!m) return;
    const uint32_t x = v >> 10, y = (v >> 5) & 31, z = v & 31;
    const long gx = (long)a.chunks[3 * c] * 32 + x, gy = (long)a.chunks[3 * c + 1] * 32 + y;
    const long gz = (long)a.chunks[3 * c + 2] * 32 + z;
    const uint32_t b = mesh_block(a, gx, gy, gz);
    vrgi_vertex3d *dst = a.out + a.vert_off[c] + (size_t)a.face_off[(size_t)c * 32768 + v] * 6;
    for (int f = 0; f < 6; ++f) {
        if (!((m >> f) & 1u)) continue;
        for (int k = 0; k < 6; ++k) {
            const int cr = MESH_FACE_C[f][k];
            vrgi_vertex3d w;
            w.position[0] = (float)x + (float)((cr >> 2) & 1);
            w.position[1] = (float)y + (float)((cr >> 1) & 1);
            w.position[2] = (float)z + (float)(cr & 1);
            w.t = b * 6u + (uint32_t)f;
            w.uv[0] = (float)MESH_FACE_U[f][k];
            w.uv[1] = (float)MESH_FACE_V[f][k];
            dst[k] = w;
        }
        dst += 6;
    }
}

// ---------------------------------------------------------------- voxel edits
// writes voxels (distinct positions) and marks the chunks whose meshes can change: the
// voxel's chunk and, on a chunk border, the neighbour across that face
__global__ void k_vx_set(uint16_t *blocks, uint32_t dx, uint32_t dy, uint32_t dz, uint64_t n, const uint32_t *xyz,
                         const uint16_t *val, uint8_t *dirty) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    if (x >= dx || y >= dy || z >= dz) return;
    blocks[((size_t)x * dy + y) * dz + z] = val[i];
    const uint32_t ny = dy / 32, nz = dz / 32;
    const uint32_t cx = x / 32, cy = y / 32, cz = z / 32;
    auto mark = [&](uint32_t a, uint32_t b, uint32_t c) { dirty[((size_t)a * ny + b) * nz + c] = 1; };
    mark(cx, cy, cz);
    if (x % 32 == 0 && cx > 0) mark(cx - 1, cy, cz);
    if (x % 32 == 31 && x + 1 < dx) mark(cx + 1, cy, cz);
    if (y % 32 == 0 && cy > 0) mark(cx, cy - 1, cz);
    if (y % 32 == 31 && y + 1 < dy) mark(cx, cy + 1, cz);
    if (z % 32 == 0 && cz > 0) mark(cx, cy, cz - 1);
    if (z % 32 == 31 && z + 1 < dz) mark(cx, cy, cz + 1);
}

// per chunk: exclusive scan over voxels of the exposed faces whose material slot emits
// (em[b] bit f: luminance of slot b * 6 + f > 0); same layout as k_mesh_scan
__global__ __launch_bounds__(1024) void k_light_scan(MeshArgs a, const uint8_t *em, uint32_t *lface_off, uint32_t *lfaces) {
    const uint32_t c = blockIdx.x;
    const uint8_t *m = a.mask + (size_t)c * 32768;
    uint32_t *off = lface_off + (size_t)c * 32768;
    const uint32_t t = threadIdx.x, v0 = t * 32;
    const long bx = (long)a.chunks[3 * c] * 32, by = (long)a.chunks[3 * c + 1] * 32, bz = (long)a.chunks[3 * c + 2] * 32;
    auto cnt = [&](uint32_t v) -> uint32_t {
        const uint32_t mv = m[v];
        if (!mv) return 0u;
        const uint32_t b = mesh_block(a, bx + (v >> 10), by + ((v >> 5) & 31), bz + (v & 31));
        return (uint32_t)__popc(mv & (b < a.air ? em[b] : 0u));
    };
    uint32_t sum = 0;
    for (uint32_t k = 0; k < 32; ++k) sum += cnt(v0 + k);
    __shared__ uint32_t part[1024];
    part[t] = sum;
    __syncthreads();
    for (uint32_t s = 1; s < 1024; s <<= 1) {
        const uint32_t u = t >= s ? part[t - s] : 0u;
        __syncthreads();
        part[t] += u;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t k = 0; k < 32; ++k) {
        off[v0 + k] = run;
        run += cnt(v0 + k);
    }
    if (t == 1023) lfaces[c] = part[1023];
}

// emissive triangles of the re-meshed chunks, in each chunk's triangle order (the
// gather of scene.rs:555-571 as upload_object does it): triangle corners + 0.0, power =
// luminance * |(b - a) x (c - a)| / 2, id = triangle index in the chunk; builder planes
// of stride n_light at the chunk's light offset
struct LightGatherArgs {
    const float *lum;
    const uint8_t *em;
    const uint32_t *lface_off;
    const uint64_t *light_off;  // per chunk: first light triangle
    float *tri9, *power, *prim9;
    uint32_t *ids;
    uint32_t n_light;
};
__global__ __launch_bounds__(256) void k_light_gather(MeshArgs a, LightGatherArgs g) {
    const uint32_t c = blockIdx.y;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.n_chunks || v >= 32768) return;
    const uint32_t m = a.mask[(size_t)c * 32768 + v];
    if (!m) return;
    const long gx = (long)a.chunks[3 * c] * 32 + (v >> 10), gy = (long)a.chunks[3 * c + 1] * 32 + ((v >> 5) & 31);
    const long gz = (long)a.chunks[3 * c + 2] * 32 + (v & 31);
    const uint32_t b = mesh_block(a, gx, gy, gz);
    const uint32_t emb = b < a.air ? g.em[b] : 0u;
    if (!(m & emb)) return;
    const vrgi_vertex3d *vx = a.out + a.vert_off[c] + (size_t)a.face_off[(size_t)c * 32768 + v] * 6;
    const uint32_t tri0 = 2u * a.face_off[(size_t)c * 32768 + v];
    uint64_t li = g.light_off[c] + 2ull * g.lface_off[(size_t)c * 32768 + v];
    uint32_t j = 0;  // exposed faces before f in this voxel
    for (int f = 0; f < 6; ++f) {
        if (!((m >> f) & 1u)) continue;
        if ((emb >> f) & 1u) {
            for (int h = 0; h < 2; ++h, ++li) {
                const vrgi_vertex3d *tv = vx + 6 * j + 3 * h;
                float p[3][3];
                for (int k = 0; k < 3; ++k)
                    for (int q = 0; q < 3; ++q) p[k][q] = tv[k].position[q] + 0.0f;
                const float e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
                const float e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
                const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                     e1[0] * e2[1] - e1[1] * e2[0]};
                const float area = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / 2.0f;
                float *t9 = g.tri9 + li * 9;
                for (int k = 0; k < 3; ++k)
                    for (int q = 0; q < 3; ++q) t9[3 * k + q] = p[k][q];
                g.power[li] = g.lum[tv[0].t] * area;
                g.ids[li] = tri0 + 2u * j + (uint32_t)h;
                f3 w[3] = {mk(p[0][0], p[0][1], p[0][2]), mk(p[1][0], p[1][1], p[1][2]), mk(p[2][0], p[2][1], p[2][2])};
                write_prim9(g.prim9, (size_t)g.n_light, (uint32_t)li, w);
            }
        }
        ++j;
    }
}

// child records of many light trees in one combined node array; tree t owns nodes
// [base[t], base[t + 1]) and its child indices are tree-local
__global__ void k_light_pairs_multi(const vrgi_bvh_node *nodes, const uint32_t *base, uint32_t n_trees, LPair *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= base[n_trees]) return;
    uint32_t lo = 0, hi = n_trees - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (base[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const vrgi_bvh_node *tn = nodes + base[lo];
    const vrgi_bvh_node nd = nodes[i];
    if (nd.left == VRGI_INVALID) return;
    const vrgi_bvh_node l = tn[nd.left], r = tn[nd.right_or_prim];
    const uint32_t lref = l.left == VRGI_INVALID ? (0x80000000u | l.right_or_prim) : nd.left;
    const uint32_t rref = r.left == VRGI_INVALID ? (0x80000000u | r.right_or_prim) : nd.right_or_prim;
    LPair p;
    p.a = make_float4(l.min[0], l.min[1], l.min[2], l.power);
    p.b = make_float4(l.max[0], l.max[1], l.max[2], r.power);
    p.c = make_float4(r.min[0], r.min[1], r.min[2], __uint_as_float(lref));
    p.d = make_float4(r.max[0], r.max[1], r.max[2], __uint_as_float(rref));
    out[i] = p;
}

void launch_vx_set(hipStream_t s, uint16_t *blocks, const uint32_t dims[3], uint64_t n, const uint32_t *xyz,
                   const uint16_t *val, uint8_t *dirty) {
    if (n) k_vx_set<<<(uint32_t)((n + 255) / 256), 256, 0, s>>>(blocks, dims[0], dims[1], dims[2], n, xyz, val, dirty);
}
static MeshArgs mesh_args(const MeshArgsHost &h) {
    MeshArgs a{};
    a.blocks = h.blocks; a.dx = h.dims[0]; a.dy = h.dims[1]; a.dz = h.dims[2];
    a.air = h.air; a.translucent = h.translucent; a.chunks = h.chunks; a.n_chunks = h.n_chunks;
    a.mask = h.mask; a.face_off = h.face_off; a.faces = h.faces; a.vert_off = h.vert_off; a.out = h.out;
    return a;
}
void launch_light_scan(hipStream_t s, const MeshArgsHost &h, const uint8_t *em, uint32_t *lface_off, uint32_t *lfaces) {
    if (h.n_chunks) k_light_scan<<<h.n_chunks, 1024, 0, s>>>(mesh_args(h), em, lface_off, lfaces);
}
void launch_light_gather(hipStream_t s, const MeshArgsHost &h, const float *lum, const uint8_t *em,
                         const uint32_t *lface_off, const uint64_t *light_off, float *tri9, float *power,
                         float *prim9, uint32_t *ids, uint32_t n_light) {
    if (!h.n_chunks || !n_light) return;
    LightGatherArgs g{lum, em, lface_off, light_off, tri9, power, prim9, ids, n_light};
    k_light_gather<<<dim3(32768 / 256, h.n_chunks), 256, 0, s>>>(mesh_args(h), g);
}
void launch_light_pairs_multi(hipStream_t s, const vrgi_bvh_node *nodes, const uint32_t *base, uint32_t n_trees,
                              uint32_t n_nodes, LPair *out) {
    if (n_trees && n_nodes) k_light_pairs_multi<<<(n_nodes + 255) / 256, 256, 0, s>>>(nodes, base, n_trees, out);
}

// light_bvh_buffer_to_light_vertexes (scene.rs:509-524) for the light BL-BVHs of a voxel
// commit: leaf flags, their exclusive scan (the rank of each leaf among its tree's leaves
// in node order), then each leaf writes the corners of its triangle (the light whose id
// is the leaf's right_or_prim, found by binary search) at its rank, with its node index
__global__ void k_leaf_flags(const vrgi_bvh_node *nodes, uint32_t n, uint32_t *flags) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) flags[g] = nodes[g].left == VRGI_INVALID ? 1u : 0u;
}
__global__ void k_light_verts_multi(const vrgi_bvh_node *nodes, uint32_t n_nodes, const uint32_t *tree_base,
                                    const uint32_t *tree_light, uint32_t n_trees, const uint32_t *ids,
                                    const float *tri9, const uint32_t *rank, vrgi_light_vertex3d *out) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_nodes) return;
    const vrgi_bvh_node nd = nodes[g];
    if (nd.left != VRGI_INVALID) return;
    uint32_t lo = 0, hi = n_trees - 1;  // last tree whose first node is <= g
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (tree_base[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t b0 = tree_base[lo], m = (tree_base[lo + 1] - b0 + 1) / 2, lb = tree_light[lo];
    uint32_t a = 0, c = m;  // first light of the tree with id >= right_or_prim
    while (a < c) {
        const uint32_t mid = (a + c) >> 1;
        if (ids[lb + mid] < nd.right_or_prim) a = mid + 1;
        else c = mid;
    }
    const float *tri = tri9 + 9 * ((size_t)lb + a);
    vrgi_light_vertex3d *o = out + 3 * ((size_t)lb + rank[g] - rank[b0]);
    for (int q = 0; q < 3; ++q) {
        o[q].position[0] = tri[3 * q];
        o[q].position[1] = tri[3 * q + 1];
        o[q].position[2] = tri[3 * q + 2];
        o[q].bvh_node_idx = g - b0;
    }
}
__global__ void k_tree_roots(const vrgi_bvh_node *nodes, const uint32_t *tree_base, uint32_t n_trees,
                             vrgi_bvh_node *roots) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_trees) roots[t] = nodes[tree_base[t]];
}
size_t light_verts_scratch_bytes(uint32_t n_nodes) {
    size_t bytes = 0;
    HIP_CHECK(rocprim::exclusive_scan(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, 0u,
                                      (size_t)n_nodes, rocprim::plus<uint32_t>()));
    return bytes;
}
void launch_light_verts_multi(hipStream_t s, const vrgi_bvh_node *nodes, uint32_t n_nodes, const uint32_t *tree_base,
                              const uint32_t *tree_light, uint32_t n_trees, const uint32_t *ids, const float *tri9,
                              uint32_t *rank, void *scratch, size_t scratch_bytes, vrgi_light_vertex3d *out,
                              vrgi_bvh_node *roots) {
    if (!n_trees || !n_nodes) return;
    const uint32_t grid = (n_nodes + 255) / 256;
    k_leaf_flags<<<grid, 256, 0, s>>>(nodes, n_nodes, rank + n_nodes);  // flags after the ranks
    size_t bytes = scratch_bytes;
    HIP_CHECK(rocprim::exclusive_scan(scratch, bytes, (const uint32_t *)(rank + n_nodes), rank, 0u,
                                      (size_t)n_nodes, rocprim::plus<uint32_t>(), s));
    k_light_verts_multi<<<grid, 256, 0, s>>>(nodes, n_nodes, tree_base, tree_light, n_trees, ids, tri9, rank, out);
    k_tree_roots<<<(n_trees + 255) / 256, 256, 0, s>>>(nodes, tree_base, n_trees, roots);
}

void launch_mesh_count(hipStream_t s, const MeshArgsHost &h) {
    MeshArgs a{};
    a.blocks = h.blocks; a.dx = h.dims[0]; a.dy = h.dims[1]; a.dz = h.dims[2];
    a.air = h.air; a.translucent = h.translucent; a.chunks = h.chunks; a.n_chunks = h.n_chunks;
    a.mask = h.mask; a.face_off = h.face_off; a.faces = h.faces;
    if (!h.n_chunks) return;
    k_mesh_mask<<<dim3(32768 / 256, h.n_chunks), 256, 0, s>>>(a);
    k_mesh_scan<<<h.n_chunks, 1024, 0, s>>>(a);
}

void launch_mesh_emit(hipStream_t s, const MeshArgsHost &h) {
    MeshArgs a{};
    a.blocks = h.blocks; a.dx = h.dims[0]; a.dy = h.dims[1]; a.dz = h.dims[2];
    a.air = h.air; a.translucent = h.translucent; a.chunks = h.chunks; a.n_chunks = h.n_chunks;
    a.mask = h.mask; a.face_off = h.face_off; a.faces = h.faces; a.vert_off = h.vert_off; a.out = h.out;
    if (!h.n_chunks) return;
    k_mesh_emit<<<dim3(32768 / 256, h.n_chunks), 256, 0, s>>>(a);
}

}  // namespace vrgi
