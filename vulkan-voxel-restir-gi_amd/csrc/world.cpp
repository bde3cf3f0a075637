// world.cpp — caller-side helpers: seeded voxel worlds of the benchmark workloads and
// the exposed-face mesher of game_system/chunk.rs:158-292.
//
// The mesher keeps the reference's face order (left, right, down, up, back, front),
// its two-triangle vertex order and uvs per face and its x-major voxel loop, because
// the primitive order of every BVH built downstream depends on them.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/vrgi.h"

namespace {

// splitmix64 finaliser: seeded integer hash used by every generator below
inline uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t hash4(uint64_t seed, int64_t a, int64_t b, int64_t c) {
    uint64_t h = mix64(seed);
    h = mix64(h ^ (uint64_t)a);
    h = mix64(h ^ (uint64_t)b);
    return mix64(h ^ (uint64_t)c);
}
inline double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

struct WorldSpec {
    uint32_t dx, dy, dz;
    int hmin, hmax;
    double base_period;  // lattice period of the first octave, in voxels
    int octaves;
    int n_boxes;
    double emissive_frac;
    uint32_t n_plain, n_emissive;
};

bool spec_of(int kind, WorldSpec *s) {
    if (kind == 0) {  // 64^3 terrain, heights 8-48, 1% emissive surface voxels
        *s = WorldSpec{64, 64, 64, 8, 48, 12.0, 4, 0, 0.01, 4, 32};
        return true;
    }
    if (kind == 1) {  // 256x128x256 terrain + 200 boxes, 0.5% emissive surface voxels
        *s = WorldSpec{256, 128, 256, 8, 120, 3.0, 4, 200, 0.005, 4, 32};
        return true;
    }
    if (kind == 4) {  // 512x256x512 terrain + 400 boxes, 1.1% emissive surface voxels
        *s = WorldSpec{512, 256, 512, 16, 240, 8.0, 4, 400, 0.011, 4, 32};
        return true;
    }
    if (kind >= 100 && kind < 100 + 65536) {
        // scene i = kind - 100 of the batched workload: each side 32..160 voxels (a multiple
        // of 32, drawn per scene), terrain up to 3/4 of the height, caves and slabs in
        // proportion to the footprint, 1% emissive surface voxels
        const uint64_t h = mix64(0xBA7Cull ^ (uint64_t)(kind - 100));
        const uint32_t sx = 32u * (1u + (uint32_t)(mix64(h + 1) % 5)), sy = 32u * (1u + (uint32_t)(mix64(h + 2) % 5));
        const uint32_t sz = 32u * (1u + (uint32_t)(mix64(h + 3) % 5));
        const double period = 4.0 + (double)(mix64(h + 4) % 12);
        *s = WorldSpec{sx, sy, sz, (int)(sy / 8), (int)(sy * 3 / 4), period, 4, (int)(sx * sz / 1024), 0.01, 4, 32};
        return true;
    }
    return false;
}

// 2D value noise, `octaves` octaves, lacunarity 2, gain 0.5, smoothstep interpolation
double value_noise(uint64_t seed, const WorldSpec &s, double x, double z) {
    double sum = 0.0, amp = 1.0, norm = 0.0, period = s.base_period;
    for (int o = 0; o < s.octaves; ++o) {
        double fx = x / period, fz = z / period;
        double ix = std::floor(fx), iz = std::floor(fz);
        double tx = fx - ix, tz = fz - iz;
        tx = tx * tx * (3.0 - 2.0 * tx);
        tz = tz * tz * (3.0 - 2.0 * tz);
        auto L = [&](double a, double b) { return unit(hash4(seed + 1000003ull * (uint64_t)o, (int64_t)a, (int64_t)b, 7)); };
        double v00 = L(ix, iz), v10 = L(ix + 1, iz), v01 = L(ix, iz + 1), v11 = L(ix + 1, iz + 1);
        double v = (v00 * (1 - tx) + v10 * tx) * (1 - tz) + (v01 * (1 - tx) + v11 * tx) * tz;
        sum += amp * v;
        norm += amp;
        amp *= 0.5;
        period *= 0.5;
    }
    return sum / norm;
}

}  // namespace

extern "C" {

int vrgi_world_dims(int kind, uint32_t dims[3], uint32_t *n_block_types) {
    WorldSpec s;
    if (!spec_of(kind, &s) || !dims || !n_block_types) return VRGI_E_INVALID;
    dims[0] = s.dx; dims[1] = s.dy; dims[2] = s.dz;
    *n_block_types = s.n_plain + s.n_emissive;
    return VRGI_OK;
}

int vrgi_world_generate(int kind, uint64_t seed, uint16_t *blocks) {
    WorldSpec s;
    if (!spec_of(kind, &s) || !blocks) return VRGI_E_INVALID;
    const uint16_t AIR = (uint16_t)(s.n_plain + s.n_emissive);
    auto at = [&](uint32_t x, uint32_t y, uint32_t z) -> uint16_t & {
        return blocks[((size_t)x * s.dy + y) * s.dz + z];
    };
    // heightmap terrain: top layer type 0, 3 layers of type 1, type 2 below with 6% type 3
    for (uint32_t x = 0; x < s.dx; ++x)
        for (uint32_t z = 0; z < s.dz; ++z) {
            double n = value_noise(seed, s, (double)x, (double)z);
            int h = s.hmin + (int)std::floor((double)(s.hmax - s.hmin) * n);
            h = std::min(std::max(h, 1), (int)s.dy);
            for (uint32_t y = 0; y < s.dy; ++y) {
                uint16_t b = AIR;
                if ((int)y < h) {
                    if ((int)y == h - 1) b = 0;
                    else if ((int)y >= h - 4) b = 1;
                    else b = unit(hash4(seed ^ 0x51ull, x, y, z)) < 0.06 ? 3 : 2;
                }
                at(x, y, z) = b;
            }
        }
    // boxes: even ones carve caves (air), odd ones add floating slabs / overhangs (type 3)
    for (int i = 0; i < s.n_boxes; ++i) {
        uint64_t h = hash4(seed ^ 0xB0C5ull, i, 0, 0);
        int sx = 4 + (int)(unit(mix64(h + 1)) * 21.0), sy = 4 + (int)(unit(mix64(h + 2)) * 13.0);
        int sz = 4 + (int)(unit(mix64(h + 3)) * 21.0);
        int cx = (int)(unit(mix64(h + 4)) * s.dx), cz = (int)(unit(mix64(h + 6)) * s.dz);
        int cy = (int)(unit(mix64(h + 5)) * (s.dy * 0.85));
        for (int x = cx - sx / 2; x < cx + (sx + 1) / 2; ++x)
            for (int y = cy - sy / 2; y < cy + (sy + 1) / 2; ++y)
                for (int z = cz - sz / 2; z < cz + (sz + 1) / 2; ++z) {
                    if (x < 0 || y < 1 || z < 0 || x >= (int)s.dx || y >= (int)s.dy || z >= (int)s.dz) continue;
                    at(x, y, z) = (i % 2 == 0) ? AIR : (uint16_t)3;
                }
    }
    // emissive surface voxels: solid voxels with an air 6-neighbour
    for (uint32_t x = 0; x < s.dx; ++x)
        for (uint32_t y = 0; y < s.dy; ++y)
            for (uint32_t z = 0; z < s.dz; ++z) {
                if (at(x, y, z) == AIR) continue;
                bool surf = false;
                const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
                for (auto &d : nb) {
                    long nx = (long)x + d[0], ny = (long)y + d[1], nz = (long)z + d[2];
                    if (nx < 0 || ny < 0 || nz < 0 || nx >= s.dx || ny >= s.dy || nz >= s.dz || at(nx, ny, nz) == AIR) {
                        surf = true;
                        break;
                    }
                }
                if (!surf) continue;
                uint64_t hh = hash4(seed ^ 0xE3155ull, x, y, z);
                if (unit(hh) < s.emissive_frac)
                    at(x, y, z) = (uint16_t)(s.n_plain + (mix64(hh) % s.n_emissive));
            }
    return VRGI_OK;
}

int vrgi_world_materials(int kind, uint64_t seed, vrgi_material *mats, float *lum, uint32_t cap,
                         uint32_t *n_slots) {
    WorldSpec s;
    if (!spec_of(kind, &s) || !n_slots) return VRGI_E_INVALID;
    uint32_t n_types = s.n_plain + s.n_emissive;
    *n_slots = n_types * 6;
    if (!mats || !lum) return VRGI_OK;
    if (cap < n_types * 6) return VRGI_E_LIMIT;
    for (uint32_t b = 0; b < n_types; ++b) {
        float alb[3], e = 0.0f;
        for (int k = 0; k < 3; ++k) alb[k] = (float)(0.2 + 0.6 * unit(hash4(seed ^ 0xA1BEull, b, k, 0)));
        if (b >= s.n_plain) e = (float)((1.0 + 19.0 * unit(hash4(seed ^ 0xE1Eull, b, 0, 0))) / 1000.0);
        for (int f = 0; f < 6; ++f) {
            vrgi_material &m = mats[b * 6 + f];
            std::memset(&m, 0, sizeof(m));
            m.reflectivity[0] = alb[0]; m.reflectivity[1] = alb[1]; m.reflectivity[2] = alb[2];
            m.reflectivity[3] = 1.0f;
            m.emissivity[0] = m.emissivity[1] = m.emissivity[2] = e;
            m.emissivity[3] = 1.0f;
            m.metallicity[3] = 1.0f;
            // utils.rs:223-235: mean over texels of r + g + b in u8 units
            lum[b * 6 + f] = 255.0f * (e + e + e);
        }
    }
    return VRGI_OK;
}

int vrgi_world_translucent(int kind, uint8_t *tr, uint32_t cap) {
    WorldSpec s;
    if (!spec_of(kind, &s) || !tr) return VRGI_E_INVALID;
    uint32_t n = s.n_plain + s.n_emissive;
    if (cap < n) return VRGI_E_LIMIT;
    std::memset(tr, 0, n);
    return VRGI_OK;
}

// chunk.rs:158-292
int vrgi_gen_mesh_chunk(const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                        const uint8_t *translucent, uint32_t cx, uint32_t cy, uint32_t cz,
                        vrgi_vertex3d *out, size_t cap, size_t *n_out) {
    if (!blocks || !dims || !translucent || !n_out) return VRGI_E_INVALID;
    const uint32_t air = n_types;
    // corner c of the unit voxel: bit 2 = x, bit 1 = y, bit 0 = z (v000 .. v111)
    // per face: 6 (corner, u, v) in the reference's push order
    static const struct { int c; float u, v; } FACE[6][6] = {
        {{1, 0, 1}, {2, 1, 0}, {0, 1, 1}, {3, 0, 0}, {2, 1, 0}, {1, 0, 1}},  // left   (chunk.rs:222-230)
        {{6, 0, 0}, {5, 1, 1}, {4, 0, 1}, {6, 0, 0}, {7, 1, 0}, {5, 1, 1}},  // right  (:233-241)
        {{0, 0, 0}, {4, 1, 0}, {1, 0, 1}, {4, 1, 0}, {5, 1, 1}, {1, 0, 1}},  // down   (:244-252)
        {{3, 1, 1}, {6, 0, 0}, {2, 1, 0}, {3, 1, 1}, {7, 0, 1}, {6, 0, 0}},  // up     (:255-263)
        {{2, 0, 0}, {4, 1, 1}, {0, 0, 1}, {2, 0, 0}, {6, 1, 0}, {4, 1, 1}},  // back   (:266-274)
        {{1, 1, 1}, {5, 0, 1}, {3, 1, 0}, {5, 0, 1}, {7, 0, 0}, {3, 1, 0}},  // front  (:277-285)
    };
    static const int NB[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    auto block = [&](long x, long y, long z) -> uint32_t {
        if (x < 0 || y < 0 || z < 0 || x >= (long)dims[0] || y >= (long)dims[1] || z >= (long)dims[2]) return air;
        return blocks[((size_t)x * dims[1] + (size_t)y) * dims[2] + (size_t)z];
    };
    size_t n = 0;
    for (uint32_t x = 0; x < 32; ++x)
        for (uint32_t y = 0; y < 32; ++y)
            for (uint32_t z = 0; z < 32; ++z) {
                long gx = (long)cx * 32 + x, gy = (long)cy * 32 + y, gz = (long)cz * 32 + z;
                uint32_t b = block(gx, gy, gz);
                if (b == air) continue;
                for (int f = 0; f < 6; ++f) {
                    uint32_t nbk = block(gx + NB[f][0], gy + NB[f][1], gz + NB[f][2]);
                    if (!(nbk == air || translucent[nbk])) continue;
                    for (int k = 0; k < 6; ++k) {
                        if (n < cap && out) {
                            int c = FACE[f][k].c;
                            vrgi_vertex3d &v = out[n];
                            v.position[0] = (float)x + (float)((c >> 2) & 1);
                            v.position[1] = (float)y + (float)((c >> 1) & 1);
                            v.position[2] = (float)z + (float)(c & 1);
                            v.t = b * 6 + (uint32_t)f;
                            v.uv[0] = FACE[f][k].u;
                            v.uv[1] = FACE[f][k].v;
                        }
                        ++n;
                    }
                }
            }
    *n_out = n;
    return n <= cap ? VRGI_OK : VRGI_E_LIMIT;
}

int vrgi_world_walk(uint16_t *blocks, const uint32_t dims[3], uint32_t n_types, uint64_t seed, uint32_t n_walkers,
                    uint32_t *walkers, uint64_t n_steps, uint32_t *xyz, uint16_t *vals, uint64_t *n_out) {
    if (!blocks || !dims || !walkers || !n_out || n_walkers == 0 || (n_steps && (!xyz || !vals))) return VRGI_E_INVALID;
    const uint64_t dx = dims[0], dy = dims[1], dz = dims[2], nvox = dx * dy * dz;
    if (nvox == 0 || n_types < 3) return VRGI_E_INVALID;
    std::vector<uint64_t> seen((nvox + 63) / 64, 0);
    std::vector<uint64_t> order;
    order.reserve(n_steps);
    for (uint64_t st = 0; st < n_steps; ++st) {
        uint32_t *w = walkers + 3 * (st % n_walkers);
        const uint64_t h = mix64(seed ^ mix64(st + 0x57A1Cull));
        const int axis = (int)(h % 3);
        const uint64_t lim = axis == 0 ? dx : axis == 1 ? dy : dz;
        if ((h >> 8) & 1u) { if (w[axis] + 1 < lim) w[axis] += 1; }
        else if (w[axis] > 0) w[axis] -= 1;
        const uint64_t i = ((uint64_t)w[0] * dy + w[1]) * dz + w[2];
        uint16_t &b = blocks[i];
        b = b == (uint16_t)n_types ? (uint16_t)2 : (uint16_t)n_types;  // solid <-> air
        if (!((seen[i >> 6] >> (i & 63)) & 1u)) {
            seen[i >> 6] |= 1ull << (i & 63);
            order.push_back(i);
        }
    }
    for (size_t k = 0; k < order.size(); ++k) {
        const uint64_t i = order[k];
        xyz[3 * k] = (uint32_t)(i / (dy * dz));
        xyz[3 * k + 1] = (uint32_t)((i / dz) % dy);
        xyz[3 * k + 2] = (uint32_t)(i % dz);
        vals[k] = blocks[i];
    }
    *n_out = order.size();
    return VRGI_OK;
}

}  // extern "C"
