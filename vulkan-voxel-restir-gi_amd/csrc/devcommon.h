// devcommon.h — device-side types and FP32-exact helpers shared by the gfx950 kernels.
//
// Every kernel of this library is compiled with -ffp-contract=off: each float
// expression below rounds once per operation in the order written, which is the
// order of the reference's Rust / GLSL line that is cited next to it. Division
// and sqrt are the correctly rounded HIP defaults (no fast-math anywhere).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vrgi.h"

#define VRGI_INVALID 0xFFFFFFFFu

// ---------------------------------------------------------------- RNG
// raytrace.rs:108-149 (duplicated in raygen.rs / restir_*.rs)
__device__ __forceinline__ uint32_t mm_combine(uint32_t h, uint32_t k) {
    k *= 0x1b873593u;
    h ^= k;
    h = (h << 13) | (h >> 19);
    h = h * 5u + 0xe6546b64u;
    return h;
}
__device__ __forceinline__ uint32_t mm_finalize(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ float float_construct(uint32_t m) {
    return __uint_as_float((m & 0x007FFFFFu) | 0x3F800000u) - 1.0f;
}
__device__ __forceinline__ float mm_finalizef(uint32_t h) { return float_construct(mm_finalize(h)); }
// counter-based stand-in for the reference's ThreadRng push constants
// (interactive_rendering.rs:1044,1448,1491): salt 2 temporal, 3 spatial
__host__ __device__ inline uint32_t frame_seed_host(uint32_t s, uint32_t frame, uint32_t salt) {
    auto comb = [](uint32_t h, uint32_t k) {
        k *= 0x1b873593u; h ^= k; h = (h << 13) | (h >> 19); return h * 5u + 0xe6546b64u;
    };
    uint32_t h = comb(comb(s, salt), frame);
    h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
    return h;
}

// ---------------------------------------------------------------- deterministic sin/cos
// GLSL cos/sin (raytrace.rs:290) are implementation-defined; this routine is the
// one both this library and its CPU checker evaluate, so directions agree bit for
// bit: Cody-Waite reduction by pi/2 (33-bit head + remainder) and Taylor series of
// sin / cos on [-pi/4, pi/4] to degree 17 / 16 in binary64, rounded once to binary32.
__device__ __forceinline__ void det_sincos(float xf, float *s, float *c) {
    const double x = (double)xf;
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.57079632673412561417;
    const double pio2_lo = 6.07710050650619224932e-11;
    const double q = floor(x * two_over_pi + 0.5);
    const double r = (x - q * pio2_hi) - q * pio2_lo;
    const double r2 = r * r;
    // nested Horner on the reciprocals of 2k(2k+1) (sine) and (2k-1)2k (cosine), each
    // rounded once at compile time (multiplications: no binary64 division in the loop)
    constexpr double sinv[9] = {0.0, 1.0 / 6, 1.0 / 20, 1.0 / 42, 1.0 / 72, 1.0 / 110, 1.0 / 156, 1.0 / 210, 1.0 / 272};
    constexpr double cinv[9] = {0.0, 1.0 / 2, 1.0 / 12, 1.0 / 30, 1.0 / 56, 1.0 / 90, 1.0 / 132, 1.0 / 182, 1.0 / 240};
    double ps = 1.0;
#pragma unroll
    for (int k = 8; k >= 1; --k) ps = 1.0 - r2 * sinv[k] * ps;
    const double sr = r * ps;
    double pc = 1.0;
#pragma unroll
    for (int k = 8; k >= 1; --k) pc = 1.0 - r2 * cinv[k] * pc;
    const double cr = pc;
    long qi = (long)q;
    int quad = (int)(((qi % 4) + 4) % 4);
    double sv, cv;
    if (quad == 0) { sv = sr; cv = cr; }
    else if (quad == 1) { sv = cr; cv = -sr; }
    else if (quad == 2) { sv = -sr; cv = -cr; }
    else { sv = -cr; cv = sr; }
    *s = (float)sv;
    *c = (float)cv;
}

// ---------------------------------------------------------------- vec3 (GLSL semantics)
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 norm3(f3 a) { return a / len3(a); }
__device__ __forceinline__ bool zero3(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ f3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ void st3(float *p, f3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
__device__ __forceinline__ float comp3(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// min / max with std::min / std::max semantics (first argument on ties and NaN), the
// choice of the CPU checker; GLSL leaves these cases unspecified
__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; }

#define VRGI_PI_F 3.1415926535897932384626433832795f
#define VRGI_EPS_RT 0.001f   /* raytrace.rs:13 */
#define VRGI_EPS_NEE 0.0001f /* nee_pdf.rs:15 */

// mat4x3 (column-major) times vec4(p, 1): ((m0 x + m1 y) + m2 z) + m3
__device__ __forceinline__ f3 xform_pt(const float *m, f3 p) {
    return mk(((m[0] * p.x + m[3] * p.y) + m[6] * p.z) + m[9] * 1.0f,
              ((m[1] * p.x + m[4] * p.y) + m[7] * p.z) + m[10] * 1.0f,
              ((m[2] * p.x + m[5] * p.y) + m[8] * p.z) + m[11] * 1.0f);
}

// ---------------------------------------------------------------- ordered float keys
// monotone map float -> int32 so that integer min / max are float min / max
__device__ __forceinline__ int fkey(float f) {
    int b = __float_as_int(f);
    return b >= 0 ? b : b ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float funkey(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7FFFFFFF); }
#define KEY_MIN_ID 0x7FFFFFFF
#define KEY_MAX_ID ((int)0x80000000)

// ---------------------------------------------------------------- traversal layout
// One internal node of the traversal copy: both children's boxes and references,
// 64 bytes, read with four 16-byte loads. ref: internal child -> its preorder rank;
// leaf child -> 0x80000000 | primitive id.
struct __align__(16) WideNode {
    float4 a;  // c0 min xyz, c0 max x
    float4 b;  // c0 max yz, c1 min xy
    float4 c;  // c1 min z, c1 max xyz
    uint4 r;   // c0 ref, c1 ref, -, -
};
// 4-ary traversal node: the binary node of rank k with two levels collapsed, its
// (up to) four descendants in left-to-right order. Per axis the four padded bounds, then
// the references (leaf 0x80000000 | prim, internal: rank of that binary node, empty slot
// 0xFFFFFFFF). Indexed by rank like WideNode, so every internal node has one.
struct __align__(16) Wide4 {
    float4 lox, loy, loz, hix, hiy, hiz;
    uint4 ref;
    uint4 pad;
};
// world triangle, 48 bytes: v0 (xyz, w = material slot), v1, v2 (w unused)
struct __align__(16) Tri { float4 v0, v1, v2; };

// ---------------------------------------------------------------- device scene view
// InstanceData (vertex.rs:31-38) of one instance, as the kernels read it
// Both children of one internal light-BVH node in one 64-byte record (four 16-byte
// loads), indexed like the node: boxes and powers exactly as in the children's
// BvhNode, plus a reference per child (leaf: 0x80000000 | right_or_prim, else the
// child's node index).
struct __align__(16) LPair {
    float4 a;  // left min xyz, left power
    float4 b;  // left max xyz, right power
    float4 c;  // right min xyz, left ref
    float4 d;  // right max xyz, right ref
};

struct DevInst {
    float xf[12];                     // transform (column-major mat4x3)
    uint32_t tri_off;                 // first global triangle id of the instance
    uint32_t tl_idx;                  // light_bvh_tl_idx
    uint32_t lin_id;                  // 1 when the transform's linear part is the identity
    const vrgi_bvh_node *light_bvh;   // bvh_node_buffer_addr: the object's light BL-BVH
    const LPair *lpairs;              // child records of that BL-BVH
};
