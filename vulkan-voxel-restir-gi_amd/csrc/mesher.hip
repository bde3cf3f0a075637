// Exposed-face meshing of voxel chunks on the GPU (chunk.rs:158-292), in the CPU
// mesher's vertex order: voxels x-major inside the 32^3 chunk (x, then y, then z), per
// voxel the faces left, right, down, up, back, front, per face the 6 table vertices.
//
// Three passes over (chunk, voxel): a 6-bit mask of exposed faces per voxel (a face is
// exposed when the neighbour is air or translucent, block.rs:129-131; outside the world
// is air); one block per chunk forms the exclusive scan of the face counts in voxel
// order; every voxel then writes its faces at the scanned position. The order is thus
// the sequential loop's order by construction.
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
