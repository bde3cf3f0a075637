// oracle.cpp — CPU restatement of the reference's hot path. TEST INFRASTRUCTURE:
// the checker that tests/, smoke() and bench.py's cpu_baseline compare the HIP
// product against; never linked into the product.
//
// Build: g++ -O2 -ffp-contract=off -fopenmp -shared -fPIC (see oracle/Makefile).
// Every float expression below is written in the evaluation order of the cited
// Rust / GLSL line, one IEEE-754 binary32 rounding per operation.
//
// Parity pinning (what checks each function; the tests named in tests/test_oracle_*.py):
//   RNG                      known answers (tests/golden/rng_known_answers.json)
//   build (A1-A6)            invariants + brute-force plane search + golden node arrays
//   partition order          "parity unpinned": `partition` 0.1.2 source absent
//   watertight hit           brute force vs BVH, watertight edge rays, closed forms
//   shading                  closed-form scenes (sky, Lambert floor, mirror)
//   light-BVH walks          probability sums, chi-square of the forward picks
//   computeNeePdf            solid-angle integral = 1, barycentric order at known hits,
//                            NEE-on mean = NEE-off mean
//   ReSTIR GI                closed-form mean with 0 / 5 neighbours, first-frame weights,
//                            Z filter (back-facing emitters), Jacobian known answers
//   ReSTIR DI                mean = Lambert's polygon formula (float64)
// Each of these checks is shown to fail with a planted mistake (OR_MUT_*, g_mut).

#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- RNG
// raytrace.rs:122-130 (murmur3_combine), :134-141 (murmur3_finalize), :108-117 (floatConstruct)
inline uint32_t combine(uint32_t h, uint32_t k) {
    k *= 0x1b873593u;
    h ^= k;
    h = (h << 13) | (h >> 19);
    h = h * 5u + 0xe6546b64u;
    return h;
}
inline uint32_t finalize(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
inline float fconstruct(uint32_t m) {
    m &= 0x007FFFFFu;
    m |= 0x3F800000u;
    float f;
    std::memcpy(&f, &m, 4);
    return f - 1.0f;
}
inline float finalizef(uint32_t h) { return fconstruct(finalize(h)); }
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

// counter-based replacement of the reference's ThreadRng push-constant seeds
// (interactive_rendering.rs:1044,1448,1491): salt 1 raygen, 2 temporal, 3 spatial, 4 DI
inline uint32_t frame_seed(uint32_t cfg_seed, uint32_t frame, uint32_t salt) {
    return finalize(combine(combine(cfg_seed, salt), frame));
}

// ---------------------------------------------------------------- deterministic sin/cos
// The GLSL built-ins cos/sin (raytrace.rs:290) are implementation-defined. Both the
// oracle and the HIP kernels use this routine so that their directions agree bit for
// bit: Cody-Waite reduction by pi/2 (pi/2 split into a 33-bit head and its remainder)
// and the Taylor series of sin / cos on [-pi/4, pi/4] to degree 17 / 16, all in
// binary64, rounded once to binary32.
inline void det_sincos(float xf, float *s, float *c) {
    const double x = (double)xf;
    const double two_over_pi = 0.63661977236758134308;
    const double pio2_hi = 1.57079632673412561417;  // pi/2 head (33 significant bits)
    const double pio2_lo = 6.07710050650619224932e-11;
    const double q = std::floor(x * two_over_pi + 0.5);
    const double r = (x - q * pio2_hi) - q * pio2_lo;
    const double r2 = r * r;
    // sin r = r (1 - r2/3! (1 - r2/(4*5) (1 - ...))): nested Horner on the reciprocals of
    // 2k(2k+1) (sine) and (2k-1)2k (cosine), each rounded once at compile time
    constexpr double sinv[9] = {0.0, 1.0 / 6, 1.0 / 20, 1.0 / 42, 1.0 / 72, 1.0 / 110, 1.0 / 156, 1.0 / 210, 1.0 / 272};
    constexpr double cinv[9] = {0.0, 1.0 / 2, 1.0 / 12, 1.0 / 30, 1.0 / 56, 1.0 / 90, 1.0 / 132, 1.0 / 182, 1.0 / 240};
    double ps = 1.0;
    for (int k = 8; k >= 1; --k) ps = 1.0 - r2 * sinv[k] * ps;
    const double sr = r * ps;
    double pc = 1.0;
    for (int k = 8; k >= 1; --k) pc = 1.0 - r2 * cinv[k] * pc;
    const double cr = pc;
    long qi = (long)q;
    int quad = (int)(((qi % 4) + 4) % 4);
    double sv, cv;
    switch (quad) {
        case 0: sv = sr; cv = cr; break;
        case 1: sv = cr; cv = -sr; break;
        case 2: sv = -sr; cv = -cr; break;
        default: sv = -cr; cv = sr; break;
    }
    *s = (float)sv;
    *c = (float)cv;
}

// ---------------------------------------------------------------- vec3 (GLSL semantics)
struct v3 { float x, y, z; };
inline v3 V(float x, float y, float z) { return v3{x, y, z}; }
inline v3 operator+(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
inline v3 operator-(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
inline v3 operator*(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
inline v3 operator*(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
inline v3 operator*(float s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
inline v3 operator/(v3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(v3 a, v3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(v3 a) { return std::sqrt(dot(a, a)); }
inline v3 normalize(v3 a) { return a / length(a); }
inline bool is_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
inline v3 reflect(v3 i, v3 n) { return i - (2.0f * dot(n, i)) * n; }
inline v3 ld3(const float *p) { return V(p[0], p[1], p[2]); }
inline void st3(float *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
inline float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// -0.0 -> +0.0 (x + 0.0f); applied to every vertex entering a build or a world-space
// triangle so that min / max never see both zero signs (whose order the reference's
// f32::min leaves unspecified)
inline v3 canon(v3 a) { return V(a.x + 0.0f, a.y + 0.0f, a.z + 0.0f); }

const float PI_F = 3.1415926535897932384626433832795f;  // M_PI as a GLSL float literal
const float EPS_RT = 0.001f;                            // raytrace.rs:13
const float EPS_NEE = 0.0001f;                          // nee_pdf.rs:15

// Planted mistake (OR_MUT_* in oracle.h), 0 = none. Set only by the tests that show a
// pinning check fails when the restatement is wrong in a plausible way.
int g_mut = 0;
inline bool mut(int m) { return g_mut == m; }

// mat4x3 (column-major, 4 columns of xyz) times vec4(p, 1)
inline v3 xform_point(const float *m, v3 p) {
    return V(((m[0] * p.x + m[3] * p.y) + m[6] * p.z) + m[9] * 1.0f,
             ((m[1] * p.x + m[4] * p.y) + m[7] * p.z) + m[10] * 1.0f,
             ((m[2] * p.x + m[5] * p.y) + m[8] * p.z) + m[11] * 1.0f);
}
const float IDENTITY12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};

// ---------------------------------------------------------------- Aabb (aabb.rs:4-92)
struct Aabb {
    bool empty = true;
    float mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
};
inline Aabb aabb_union(const Aabb &a, const Aabb &b) {  // aabb.rs:24-42 (inf / sup)
    if (a.empty) return b;
    if (b.empty) return a;
    Aabb r;
    r.empty = false;
    for (int i = 0; i < 3; ++i) {
        r.mn[i] = std::min(a.mn[i], b.mn[i]);
        r.mx[i] = std::max(a.mx[i], b.mx[i]);
    }
    return r;
}
inline float aabb_area(const Aabb &a) {  // aabb.rs:44-52
    if (a.empty) return 0.0f;
    float dx = a.mx[0] - a.mn[0], dy = a.mx[1] - a.mn[1], dz = a.mx[2] - a.mn[2];
    return 2.0f * (dx * dy + dx * dz + dy * dz);
}
inline Aabb aabb_from_points(const v3 *p, int n) {  // aabb.rs:10-22
    Aabb r;
    r.empty = false;
    r.mn[0] = r.mx[0] = p[0].x;
    r.mn[1] = r.mx[1] = p[0].y;
    r.mn[2] = r.mx[2] = p[0].z;
    for (int k = 0; k < n; ++k)
        for (int i = 0; i < 3; ++i) {
            r.mn[i] = std::min(r.mn[i], comp(p[k], i));
            r.mx[i] = std::max(r.mx[i], comp(p[k], i));
        }
    return r;
}
inline Aabb aabb_transform(const Aabb &a, const float *m) {  // aabb.rs:75-91
    if (a.empty) return a;
    v3 lo = V(a.mn[0], a.mn[1], a.mn[2]), hi = V(a.mx[0], a.mx[1], a.mx[2]);
    v3 c[8] = {xform_point(m, lo),
               xform_point(m, V(lo.x, lo.y, hi.z)),
               xform_point(m, V(lo.x, hi.y, lo.z)),
               xform_point(m, V(lo.x, hi.y, hi.z)),
               xform_point(m, V(hi.x, lo.y, lo.z)),
               xform_point(m, V(hi.x, lo.y, hi.z)),
               xform_point(m, V(hi.x, hi.y, lo.z)),
               xform_point(m, hi)};
    return aabb_from_points(c, 8);
}

// ---------------------------------------------------------------- BvhNode (bvh/mod.rs)
struct BvhNode {
    uint32_t left = 0xFFFFFFFFu;
    uint32_t right = 0xFFFFFFFFu;
    float mn[3] = {0, 0, 0};
    float power = 0.0f;
    float mx[3] = {0, 0, 0};
    uint32_t parent = 0xFFFFFFFFu;
};
static_assert(sizeof(BvhNode) == 40, "BvhNode is 40 bytes");

struct Vertex3D { float pos[3]; uint32_t t; float uv[2]; };
static_assert(sizeof(Vertex3D) == 24, "Vertex3D is 24 bytes");
struct LightVertex3D { float pos[3]; uint32_t node; };
struct Material { float refl[4], emis[4], metal[4]; };

// ---------------------------------------------------------------- binned SAH builder
// build.rs:10-34 (build node kinds)
struct BuildNode {
    Aabb box;
    bool leaf;
    size_t parent, first, count;  // leaf payload
    size_t left, right;           // internal payload
};

// build.rs:36-43
Aabb leaf_bounds(size_t first, size_t count, const std::vector<size_t> &idx,
                 const std::vector<Aabb> &boxes) {
    Aabb b;
    for (size_t i = first; i < first + count; ++i) b = aabb_union(b, boxes[idx[i]]);
    return b;
}

// Rust `f32 as usize`: NaN -> 0, negative -> 0, saturating above.
inline size_t rust_f32_as_usize(float f) {
    if (!(f == f)) return 0;
    if (f <= 0.0f) return 0;
    if (f >= 18446744073709551615.0f) return SIZE_MAX;
    return (size_t)f;
}

// build.rs:45-128
void find_best_plane(size_t first, size_t count, const std::vector<size_t> &idx,
                     const std::vector<float> &cen, const std::vector<Aabb> &boxes, size_t *dim_out,
                     float *split_out) {
    const size_t BINS = 32;
    float best_cost = 3.40282347e+38f;  // f32::MAX
    size_t best_dim = 0;
    float best_split = 0.0f;
    for (size_t d = 0; d < 3; ++d) {
        float bmin = 3.40282347e+38f, bmax = -3.40282347e+38f;  // f32::MAX, f32::MIN
        for (size_t i = first; i < first + count; ++i) {
            float c = cen[idx[i] * 3 + d];
            bmin = std::min(bmin, c);  // f32::min (no NaN inputs)
            bmax = std::max(bmax, c);
        }
        Aabb bin_box[32];
        size_t bin_cnt[32] = {0};
        float scale = (float)BINS / (bmax - bmin);
        for (size_t i = first; i < first + count; ++i) {
            size_t p = idx[i];
            size_t b = std::min(BINS - 1, rust_f32_as_usize((cen[p * 3 + d] - bmin) * scale));
            bin_cnt[b] += 1;
            bin_box[b] = aabb_union(bin_box[b], boxes[p]);
        }
        Aabb pl_l[31], pl_r[31];
        size_t pc_l[31], pc_r[31];
        Aabb al, ar;
        size_t cl = 0, cr = 0;
        for (size_t pl = 0; pl < BINS - 1; ++pl) {
            cl += bin_cnt[pl];
            pc_l[pl] = cl;
            al = aabb_union(al, bin_box[pl]);
            pl_l[pl] = al;
            cr += bin_cnt[BINS - 1 - pl];
            pc_r[BINS - 2 - pl] = cr;
            ar = aabb_union(ar, bin_box[BINS - 1 - pl]);
            pl_r[BINS - 2 - pl] = ar;
        }
        float scale2 = (bmax - bmin) / (float)BINS;
        for (size_t pl = 0; pl < BINS - 1; ++pl) {
            // cost_function, build.rs:287-289
            float cost = aabb_area(pl_l[pl]) * (float)pc_l[pl] + aabb_area(pl_r[pl]) * (float)pc_r[pl];
            if (cost < best_cost) {
                best_cost = cost;
                best_dim = d;
                best_split = bmin + ((float)pl + 1.0f) * scale2;
            }
        }
    }
    *dim_out = best_dim;
    *split_out = best_split;
}

// `partition` 0.1.2 (build.rs:145): l/r scan with swaps; returns the split index.
// Closed form: with T = #true, the i-th false in [0,T) (left to right) trades
// places with the i-th true in [T,n) (right to left). PARITY UNPINNED: restated
// from the crate's published algorithm; its source is not among the reference's files.
size_t partition_inplace(size_t *d, size_t len, const std::vector<float> &cen, size_t dim,
                         float split) {
    if (len == 0) return 0;
    auto pred = [&](size_t p) { return cen[p * 3 + dim] < split; };
    size_t l = 0, r = len - 1;
    for (;;) {
        while (l < len && pred(d[l])) ++l;
        while (r > 0 && !pred(d[r])) --r;
        if (l >= r) return l;
        std::swap(d[l], d[r]);
    }
}

size_t insert_leaf(size_t parent, size_t first, size_t count, std::vector<BuildNode> &nodes,
                   const std::vector<size_t> &idx, const std::vector<Aabb> &boxes) {
    BuildNode n;
    n.box = leaf_bounds(first, count, idx, boxes);
    n.leaf = true;
    n.parent = parent;
    n.first = first;
    n.count = count;
    n.left = n.right = 0;
    nodes.push_back(n);
    return nodes.size() - 1;
}

// build.rs:130-233 (recursion replaced by an explicit stack with the same order:
// children are inserted before either is subdivided, left subtree first)
void subdivide_all(std::vector<size_t> &idx, const std::vector<Aabb> &boxes,
                   const std::vector<float> &cen, std::vector<BuildNode> &nodes) {
    std::vector<size_t> stack{0};
    while (!stack.empty()) {
        size_t ni = stack.back();
        stack.pop_back();
        BuildNode nd = nodes[ni];
        if (!nd.leaf) continue;
        if (nd.count > 2) {
            size_t dim;
            float split;
            find_best_plane(nd.first, nd.count, idx, cen, boxes, &dim, &split);
            size_t nl = partition_inplace(&idx[nd.first], nd.count, cen, dim, split);
            size_t nr = nd.count - nl;
            if (nl == 0 || nr == 0) {  // build.rs:152-157
                nl = nd.count / 2;
                nr = nd.count - nl;
            }
            size_t l = insert_leaf(ni, nd.first, nl, nodes, idx, boxes);
            size_t r = insert_leaf(ni, nd.first + nl, nr, nodes, idx, boxes);
            nodes[ni].leaf = false;
            nodes[ni].left = l;
            nodes[ni].right = r;
            stack.push_back(r);
            stack.push_back(l);
        } else if (nd.count == 2) {  // build.rs:203-230
            size_t l = insert_leaf(ni, nd.first, 1, nodes, idx, boxes);
            size_t r = insert_leaf(ni, nd.first + 1, 1, nodes, idx, boxes);
            nodes[ni].leaf = false;
            nodes[ni].left = l;
            nodes[ni].right = r;
        }
    }
}

// shared body of build_bl_bvh (build.rs:249-372) and build_tl_bvh (build.rs:374-480)
void build_core(const std::vector<Aabb> &boxes, const std::vector<float> &cen, const float *power,
                const uint32_t *ids, std::vector<BvhNode> &out, std::vector<uint32_t> *order,
                uint32_t *depth_out) {
    size_t n = boxes.size();
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = i;
    std::vector<BuildNode> nodes;
    nodes.reserve(2 * n);
    insert_leaf(0xFFFFFFFFu, 0, n, nodes, idx, boxes);
    subdivide_all(idx, boxes, cen, nodes);
    const float pad = 0.0001f;  // build.rs:300
    out.assign(nodes.size(), BvhNode());
    for (size_t i = 0; i < nodes.size(); ++i) {
        const BuildNode &b = nodes[i];
        BvhNode &o = out[i];
        o.parent = (uint32_t)b.parent;
        if (b.leaf) {  // build.rs:311-326
            size_t p = idx[b.first];
            o.left = 0xFFFFFFFFu;
            o.right = ids[p];
            for (int k = 0; k < 3; ++k) { o.mn[k] = boxes[p].mn[k]; o.mx[k] = boxes[p].mx[k]; }
            o.power = power[p];
        } else {  // build.rs:327-337
            o.left = (uint32_t)b.left;
            o.right = (uint32_t)b.right;
            for (int k = 0; k < 3; ++k) { o.mn[k] = b.box.mn[k] - pad; o.mx[k] = b.box.mx[k] + pad; }
            o.power = 0.0f;
        }
    }
    for (size_t i = out.size(); i-- > 0;) {  // build.rs:345-357
        if (out[i].left != 0xFFFFFFFFu) {
            out[i].power += out[out[i].left].power;
            out[i].power += out[out[i].right].power;
        }
    }
    if (order) {
        order->resize(n);
        for (size_t i = 0; i < n; ++i) (*order)[i] = (uint32_t)idx[i];
    }
    if (depth_out) {
        std::vector<uint32_t> dep(out.size(), 0);
        uint32_t mx = 0;
        for (size_t i = 0; i < out.size(); ++i) {
            if (out[i].left != 0xFFFFFFFFu) {
                dep[out[i].left] = dep[i] + 1;
                dep[out[i].right] = dep[i] + 1;
            }
            mx = std::max(mx, dep[i]);
        }
        *depth_out = mx;
    }
}

// build.rs:249-372; tri: 9 floats per primitive
void build_bl(size_t n, const float *tri, const float *lum, const uint32_t *ids,
              std::vector<BvhNode> &out, std::vector<uint32_t> *order, Aabb *root_box,
              float *root_power, uint32_t *depth) {
    std::vector<Aabb> boxes(n);
    std::vector<float> cen(3 * n);
    for (size_t i = 0; i < n; ++i) {
        v3 v[3] = {canon(ld3(tri + 9 * i)), canon(ld3(tri + 9 * i + 3)), canon(ld3(tri + 9 * i + 6))};
        boxes[i] = aabb_from_points(v, 3);
        v3 c = ((v[0] + v[1]) + v[2]) / 3.0f;  // build.rs:267-270
        cen[3 * i] = c.x; cen[3 * i + 1] = c.y; cen[3 * i + 2] = c.z;
    }
    build_core(boxes, cen, lum, ids, out, order, depth);
    if (root_box) {  // build.rs:359-371
        if (out[0].left == 0xFFFFFFFFu) {
            *root_box = boxes[0];
            *root_power = lum[0];
        } else {
            root_box->empty = false;
            for (int k = 0; k < 3; ++k) { root_box->mn[k] = out[0].mn[k]; root_box->mx[k] = out[0].mx[k]; }
            *root_power = out[0].power;
        }
    }
}

// build.rs:374-480
void build_tl(size_t m, const float *xf, const std::vector<Aabb> &in_boxes, const float *pw,
              const uint32_t *ids, std::vector<BvhNode> &out) {
    std::vector<Aabb> boxes(m);
    std::vector<float> cen(3 * m);
    for (size_t i = 0; i < m; ++i) {
        boxes[i] = aabb_transform(in_boxes[i], xf + 12 * i);
        for (int k = 0; k < 3; ++k) { boxes[i].mn[k] += 0.0f; boxes[i].mx[k] += 0.0f; }
        for (int k = 0; k < 3; ++k) cen[3 * i + k] = (boxes[i].mn[k] + boxes[i].mx[k]) / 2.0f;
    }
    build_core(boxes, cen, pw, ids, out, nullptr, nullptr);
}

// ---------------------------------------------------------------- watertight intersection
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013.
struct RayPre {
    v3 o, d;
    int kx, ky, kz;
    float sx, sy, sz;
    v3 inv;
};
inline RayPre ray_pre(v3 o, v3 d) {
    RayPre r;
    r.o = o;
    r.d = d;
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    r.kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
    r.kx = (r.kz + 1) % 3;
    r.ky = (r.kx + 1) % 3;
    if (comp(d, r.kz) < 0.0f) std::swap(r.kx, r.ky);
    float dz = comp(d, r.kz);
    r.sx = comp(d, r.kx) / dz;
    r.sy = comp(d, r.ky) / dz;
    r.sz = 1.0f / dz;
    // traversal-only reciprocal; zero components get a huge finite value so that
    // (b - o) * inv is never 0 * inf
    auto rc = [](float v) { return v != 0.0f ? 1.0f / v : 1e30f; };
    r.inv = V(rc(d.x), rc(d.y), rc(d.z));
    return r;
}
// returns true and (t, u, v) (u, v = weights of vertices 1 and 2) when t in [tmin, tmax]
inline bool tri_intersect(const RayPre &r, v3 v0, v3 v1, v3 v2, float tmin, float tmax,
                          float *t_out, float *u_out, float *v_out) {
    v3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    float Akz = comp(A, r.kz), Bkz = comp(B, r.kz), Ckz = comp(C, r.kz);
    float Ax = comp(A, r.kx) - r.sx * Akz, Ay = comp(A, r.ky) - r.sy * Akz;
    float Bx = comp(B, r.kx) - r.sx * Bkz, By = comp(B, r.ky) - r.sy * Bkz;
    float Cx = comp(C, r.kx) - r.sx * Ckz, Cy = comp(C, r.ky) - r.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float Vv = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || Vv == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        Vv = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || Vv < 0.0f || W < 0.0f) && (U > 0.0f || Vv > 0.0f || W > 0.0f)) return false;
    float det = U + Vv + W;
    if (det == 0.0f) return false;
    float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
    float T = U * Az + Vv * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = Vv * rcp;
    *v_out = W * rcp;
    return true;
}
// conservative slab test (Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013)
inline bool box_hit(const RayPre &r, const float *mn, const float *mx, float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    for (int a = 0; a < 3; ++a) {
        float o = comp(r.o, a), inv = comp(r.inv, a);
        // 1e-3 of slack: a triangle hit reported at rounding distance outside its exact
        // box is never culled (the same slack as the HIP traversal copies)
        float tn = ((mn[a] - 0.001f) - o) * inv, tf = ((mx[a] + 0.001f) - o) * inv;
        if (tn > tf) std::swap(tn, tf);
        t0 = tn > t0 ? tn : t0;
        t1 = tf < t1 ? tf : t1;
    }
    return t0 <= t1 * 1.00000024f;  // 1 + 2*gamma(3)
}

// ---------------------------------------------------------------- scene
struct Object {
    std::vector<Vertex3D> verts;
    std::vector<BvhNode> light_bvh;          // empty when no emissive triangle
    std::vector<LightVertex3D> light_verts;  // 3 per light prim, leaf order
    Aabb light_aabb;
    float power = 0.0f;
};
struct Inst {
    std::shared_ptr<Object> obj;  // null for the Empty handle
    float xf[12];
};

}  // namespace

struct or_scene {
    std::vector<Material> mats;
    std::vector<float> lum;
    // texture atlas (or_scene_set_textures): linear RGBA per texel, [slot][3][h][w][4]
    std::vector<float> tex;
    uint32_t tex_w = 0, tex_h = 0, tex_slots = 0;
    std::vector<float> tri_uv;                // per global id: uv of its 3 vertices (6 floats)
    std::map<uint32_t, Inst> objects;
    // built state
    std::vector<const Inst *> inst;           // instances = Some objects in key order
    std::vector<uint32_t> inst_tri_off;       // global triangle offset per instance
    std::vector<float> tri;                   // world triangles, 9 per global id
    std::vector<uint32_t> tri_inst, tri_t;    // instance, material slot per global id
    std::vector<BvhNode> geom;                // geometry BVH (build_bl_bvh semantics)
    std::vector<uint32_t> geom_order;
    uint32_t geom_depth = 0;
    std::vector<BvhNode> tl;                  // light TL BVH (scene.rs:275-279)
    std::vector<uint32_t> inst_tl_idx;        // InstanceData.light_bvh_tl_idx
    std::vector<float> ltri;                  // world light triangles, 9 per light id
    std::vector<uint32_t> ltri_inst, ltri_j;  // instance, light prim j per light id
    std::vector<uint32_t> ltri_node;          // LightVertex3D.bvh_node_idx per light id
    std::vector<BvhNode> lgeom;               // BVH over world light triangles
    std::vector<uint32_t> lgeom_order;
    std::vector<uint32_t> lgeom_rank;         // leaf position of each light id
    uint64_t version = 0;                     // bumped by every add / remove
};

namespace {

// SceneUploader::upload_object light part (scene.rs:547-571, 616-650, 509-524)
std::shared_ptr<Object> upload(const or_scene *s, const Vertex3D *v, size_t n) {
    auto o = std::make_shared<Object>();
    o->verts.assign(v, v + n);
    std::vector<float> pw, tri;
    std::vector<uint32_t> ids;
    for (size_t i = 0; i < n / 3; ++i) {
        float lum = s->lum[v[3 * i].t];
        if (lum > 0.0f) {
            v3 a = canon(ld3(v[3 * i].pos)), b = canon(ld3(v[3 * i + 1].pos)), c = canon(ld3(v[3 * i + 2].pos));
            float area = length(cross(b - a, c - a)) / 2.0f;
            for (v3 p : {a, b, c}) { tri.push_back(p.x); tri.push_back(p.y); tri.push_back(p.z); }
            pw.push_back(lum * area);
            ids.push_back((uint32_t)i);
        }
    }
    if (!ids.empty()) {
        build_bl(ids.size(), tri.data(), pw.data(), ids.data(), o->light_bvh, nullptr,
                 &o->light_aabb, &o->power, nullptr);
        for (size_t i = 0; i < o->light_bvh.size(); ++i) {  // scene.rs:509-524
            const BvhNode &nd = o->light_bvh[i];
            if (nd.left == 0xFFFFFFFFu) {
                size_t pi = nd.right;
                for (int k = 0; k < 3; ++k) {
                    LightVertex3D lv;
                    std::memcpy(lv.pos, v[pi * 3 + k].pos, 12);
                    lv.node = (uint32_t)i;
                    o->light_verts.push_back(lv);
                }
            }
        }
    }
    return o;
}

void scene_build(or_scene *s) {
    s->inst.clear();
    for (auto &kv : s->objects)
        if (kv.second.obj) s->inst.push_back(&kv.second);
    size_t ni = s->inst.size();
    // light TL BVH + InstanceData.light_bvh_tl_idx (scene.rs:243-338)
    std::vector<float> xfs;
    std::vector<Aabb> boxes;
    std::vector<float> pws;
    std::vector<uint32_t> ids;
    for (size_t i = 0; i < ni; ++i) {
        const Object &o = *s->inst[i]->obj;
        if (o.light_aabb.empty) continue;
        xfs.insert(xfs.end(), s->inst[i]->xf, s->inst[i]->xf + 12);
        boxes.push_back(o.light_aabb);
        pws.push_back(o.power);
        ids.push_back((uint32_t)i);
    }
    if (boxes.empty()) s->tl.assign(1, BvhNode());
    else build_tl(boxes.size(), xfs.data(), boxes, pws.data(), ids.data(), s->tl);
    s->inst_tl_idx.assign(ni, 0);
    for (size_t k = 0; k < s->tl.size(); ++k)
        if (s->tl[k].left == 0xFFFFFFFFu && s->tl[k].right != 0xFFFFFFFFu)
            s->inst_tl_idx[s->tl[k].right] = (uint32_t)k;
    // world triangles in instance order
    s->tri.clear(); s->tri_inst.clear(); s->tri_t.clear(); s->tri_uv.clear(); s->inst_tri_off.assign(ni + 1, 0);
    s->ltri.clear(); s->ltri_inst.clear(); s->ltri_j.clear(); s->ltri_node.clear();
    for (size_t i = 0; i < ni; ++i) {
        const Object &o = *s->inst[i]->obj;
        const float *m = s->inst[i]->xf;
        s->inst_tri_off[i] = (uint32_t)s->tri_inst.size();
        for (size_t p = 0; p < o.verts.size() / 3; ++p) {
            for (int k = 0; k < 3; ++k) {
                v3 w = canon(xform_point(m, ld3(o.verts[3 * p + k].pos)));
                s->tri.push_back(w.x); s->tri.push_back(w.y); s->tri.push_back(w.z);
            }
            s->tri_inst.push_back((uint32_t)i);
            s->tri_t.push_back(o.verts[3 * p].t);
            for (int k = 0; k < 3; ++k) {
                s->tri_uv.push_back(o.verts[3 * p + k].uv[0]);
                s->tri_uv.push_back(o.verts[3 * p + k].uv[1]);
            }
        }
        for (size_t j = 0; j < o.light_verts.size() / 3; ++j) {
            for (int k = 0; k < 3; ++k) {
                v3 w = canon(xform_point(m, ld3(o.light_verts[3 * j + k].pos)));
                s->ltri.push_back(w.x); s->ltri.push_back(w.y); s->ltri.push_back(w.z);
            }
            s->ltri_inst.push_back((uint32_t)i);
            s->ltri_j.push_back((uint32_t)j);
            s->ltri_node.push_back(o.light_verts[3 * j].node);
        }
    }
    s->inst_tri_off[ni] = (uint32_t)s->tri_inst.size();
    size_t nt = s->tri_inst.size();
    if (nt) {
        std::vector<float> zero(nt, 0.0f);
        std::vector<uint32_t> gid(nt);
        for (size_t i = 0; i < nt; ++i) gid[i] = (uint32_t)i;
        build_bl(nt, s->tri.data(), zero.data(), gid.data(), s->geom, &s->geom_order, nullptr,
                 nullptr, &s->geom_depth);
    } else {
        s->geom.assign(1, BvhNode());
        s->geom_order.clear();
        s->geom_depth = 0;
    }
    size_t nl = s->ltri_inst.size();
    if (nl) {
        std::vector<float> zero(nl, 0.0f);
        std::vector<uint32_t> gid(nl);
        for (size_t i = 0; i < nl; ++i) gid[i] = (uint32_t)i;
        build_bl(nl, s->ltri.data(), zero.data(), gid.data(), s->lgeom, &s->lgeom_order, nullptr,
                 nullptr, nullptr);
        s->lgeom_rank.assign(nl, 0);
        for (size_t i = 0; i < nl; ++i) s->lgeom_rank[s->lgeom_order[i]] = (uint32_t)i;
    } else {
        s->lgeom.assign(1, BvhNode());
        s->lgeom_order.clear();
        s->lgeom_rank.clear();
    }
}

inline v3 tri_v(const std::vector<float> &t, size_t id, int k) { return ld3(&t[9 * id + 3 * k]); }

struct Hit {
    uint32_t prim = 0xFFFFFFFFu;
    float t = 0.0f, u = 0.0f, v = 0.0f;
};

// closest hit, tie -> lower global id (replaces raytrace.rs:354-388)
Hit closest_bvh(const or_scene *s, v3 o, v3 d, float tmin, float tmax, uint64_t *visited) {
    Hit h;
    if (s->tri_inst.empty()) return h;
    RayPre r = ray_pre(o, d);
    float best = tmax;
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t ni = stack[--sp];
        const BvhNode &nd = s->geom[ni];
        if (visited) ++*visited;
        if (!box_hit(r, nd.mn, nd.mx, tmin, best)) continue;
        if (nd.left == 0xFFFFFFFFu) {
            float t, u, v;
            if (tri_intersect(r, tri_v(s->tri, nd.right, 0), tri_v(s->tri, nd.right, 1),
                              tri_v(s->tri, nd.right, 2), tmin, best, &t, &u, &v)) {
                if (t < best || (t == best && nd.right < h.prim)) {
                    best = t;
                    h.prim = nd.right; h.t = t; h.u = u; h.v = v;
                }
            }
        } else {
            if (sp + 2 > 256) std::abort();
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
    return h;
}

Hit closest_brute(const or_scene *s, v3 o, v3 d, float tmin, float tmax) {
    Hit h;
    RayPre r = ray_pre(o, d);
    float best = tmax;
    for (size_t i = 0; i < s->tri_inst.size(); ++i) {
        float t, u, v;
        if (tri_intersect(r, tri_v(s->tri, i, 0), tri_v(s->tri, i, 1), tri_v(s->tri, i, 2), tmin,
                          best, &t, &u, &v)) {
            if (t < best || (t == best && (uint32_t)i < h.prim)) {
                best = t;
                h.prim = (uint32_t)i; h.t = t; h.u = u; h.v = v;
            }
        }
    }
    return h;
}

// every light triangle crossed in [tmin, tmax], in leaf (left-to-right) order of the
// light-geometry BVH (replaces the candidate loop of nee_pdf.rs:305-334)
struct LHit { uint32_t id; float u, v; };
void light_hits(const or_scene *s, v3 o, v3 d, float tmin, float tmax, std::vector<LHit> &out) {
    out.clear();
    if (s->ltri_inst.empty()) return;
    RayPre r = ray_pre(o, d);
    uint32_t stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        uint32_t ni = stack[--sp];
        const BvhNode &nd = s->lgeom[ni];
        if (!box_hit(r, nd.mn, nd.mx, tmin, tmax)) continue;
        if (nd.left == 0xFFFFFFFFu) {
            float t, u, v;
            if (tri_intersect(r, tri_v(s->ltri, nd.right, 0), tri_v(s->ltri, nd.right, 1),
                              tri_v(s->ltri, nd.right, 2), tmin, tmax, &t, &u, &v))
                out.push_back(LHit{nd.right, u, v});
        } else {
            if (sp + 2 > 256) std::abort();
            stack[sp++] = nd.right;
            stack[sp++] = nd.left;
        }
    }
}

// ---------------------------------------------------------------- light BVH walks
// raytrace.rs:181-208 (eps 1e-3, integer corner count) and nee_pdf.rs:125-152 (eps
// 1e-4, 0.125 increments): both equal power / dist_sq * (count / 8).
inline float node_importance(v3 p, v3 n, const float *m, const BvhNode &nd, float eps) {
    v3 a = xform_point(m, ld3(nd.mn)), b = xform_point(m, ld3(nd.mx));
    v3 c[8] = {a, V(a.x, a.y, b.z), V(a.x, b.y, a.z), V(a.x, b.y, b.z),
               V(b.x, a.y, a.z), V(b.x, a.y, b.z), V(b.x, b.y, a.z), b};
    uint32_t vis = 0;
    for (int k = 0; k < 8; ++k) vis += (dot(c[k] - p, n) >= eps) ? 1u : 0u;
    v3 diag = b - a;
    v3 mid = 0.5f * (a + b) - p;
    float dsq = std::max(dot(diag, diag), dot(mid, mid));
    return nd.power / dsq * ((float)vis / 8.0f);
}

struct Pick { bool ok; uint32_t inst, prim; float prob, importance; };

// raytrace.rs:218-281
Pick traverse_light(const or_scene *s, v3 p, v3 n, uint32_t seed) {
    const BvhNode *root = s->tl.data();
    BvhNode node = root[0];
    if (node.left == 0xFFFFFFFFu && node.right == 0xFFFFFFFFu) return Pick{false, 0, 0, 1.0f, 0.0f};
    float prob = 1.0f, imp = 0.0f;
    const float *m = IDENTITY12;
    uint32_t inst = 0xFFFFFFFFu;
    bool top = true;
    for (;;) {
        if (top && node.left == 0xFFFFFFFFu) {
            inst = node.right;
            m = s->inst[inst]->xf;
            top = false;
            root = s->inst[inst]->obj->light_bvh.data();
            node = root[0];
            if (imp == 0.0f) imp = node_importance(p, n, m, node, EPS_RT);
        }
        if (!top && node.left == 0xFFFFFFFFu) return Pick{true, inst, node.right, prob, imp};
        BvhNode l = root[node.left], r = root[node.right];
        float li = node_importance(p, n, m, l, EPS_RT);
        float ri = node_importance(p, n, m, r, EPS_RT);
        float tot = li + ri;
        float ln = li / tot, rn = ri / tot;
        if (finalizef(seed) < ln) { node = l; prob *= ln; imp = li; }
        else { node = r; prob *= rn; imp = ri; }
        seed = combine(seed, 0);
    }
}

// nee_pdf.rs:154-228
float reverse_traverse(const or_scene *s, v3 p, v3 n, uint32_t inst, uint32_t bl_idx) {
    const BvhNode *root = s->inst[inst]->obj->light_bvh.data();
    float prob = 1.0f;
    const float *m = s->inst[inst]->xf;
    bool top = false;
    BvhNode node = root[bl_idx];
    uint32_t ni = bl_idx;
    for (;;) {
        uint32_t old = ni;
        if (node.parent == 0xFFFFFFFFu) {
            if (top) return prob;
            top = true;
            root = s->tl.data();
            old = s->inst_tl_idx[inst];
            BvhNode tl_leaf = root[old];
            if (tl_leaf.parent == 0xFFFFFFFFu) return prob;
            ni = tl_leaf.parent;
            node = root[ni];
            m = IDENTITY12;
        } else {
            ni = node.parent;
            node = root[ni];
        }
        BvhNode l = root[node.left], r = root[node.right];
        float li = node_importance(p, n, m, l, EPS_NEE);
        float ri = node_importance(p, n, m, r, EPS_NEE);
        float tot = li + ri;
        float ln = li / tot, rn = ri / tot;
        prob *= (old == node.left) ? ln : rn;
    }
}

// nee_pdf.rs:240-279 (barycentric convention (x, y, 1-x-y), nee_pdf.rs:230-232) for
// world light triangle `lid` (the instance transform of the light vertices applied)
float nee_pdf_one(const or_scene *s, v3 p, v3 n, v3 w, uint32_t lid, float bx, float by) {
    v3 t0 = tri_v(s->ltri, lid, 0), t1 = tri_v(s->ltri, lid, 1), t2 = tri_v(s->ltri, lid, 2);
    v3 q = mut(OR_MUT_NEE_BARY) ? (((1.0f - bx) - by) * t0 + bx * t1) + by * t2
                                : (bx * t0 + by * t1) + ((1.0f - bx) - by) * t2;
    float dist = length(q - p);
    float area = mut(OR_MUT_NEE_AREA) ? length(cross(t0 - t1, t2 - t1))
                                      : 0.5f * length(cross(t0 - t1, t2 - t1));
    float cs = mut(OR_MUT_NEE_COS) ? 1.0f : dot(n, w);
    float ppp = mut(OR_MUT_NEE_DIST) ? dist / (cs * area) : dist * dist / (cs * area);
    float prim_p = reverse_traverse(s, p, n, s->ltri_inst[lid], s->ltri_node[lid]);
    return prim_p * ppp;
}

// nee_pdf.rs:281-337 for one ray leaving `o` (shading normal n) along d: the sum of
// computeNeePdf over every light triangle crossed in [1e-4, 1000] (nee_pdf.rs:305-334)
float nee_pdf_ray(const or_scene *s, v3 o, v3 n, v3 d, std::vector<LHit> &hits) {
    float np = 0.0f;
    light_hits(s, o, d, EPS_NEE, 1000.0f, hits);
    for (const LHit &lh : hits) np += nee_pdf_one(s, o, n, d, lh.id, lh.u, lh.v);
    return np;
}

// ---------------------------------------------------------------- per-pixel kernels
struct Cfg {
    uint32_t W, H, B, rr_start;
    float rr_min;
    uint32_t mode;
    float radius;
    uint32_t di_cand, seed, flags;
};

struct IcS { v3 n, t, b; };
// raytrace.rs:319-331
inline IcS local_cs(v3 a, v3 b, v3 c) {
    v3 e1 = b - a, e2 = c - a;
    v3 n = cross(e1, e2);
    v3 t = e1;
    v3 bt = cross(n, t);
    return IcS{normalize(n), normalize(t), normalize(bt)};
}

// R8G8B8A8_SRGB read by a nearest, clamp-to-edge sampler (interactive_rendering.rs:416,
// 744; vulkano 0.35.2 Sampler defaults): colour channels through the sRGB transfer
// function (IEC 61966-2-1) evaluated in binary64 and rounded once, alpha as a / 255
float srgb_decode(uint8_t v) {
    const double c = (double)v / 255.0;
    return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}
// texel of normalised coordinate c over n texels: floor(c * n), clamped to the edge
inline uint32_t texel_index(float c, uint32_t n) {
    const float f = std::floor(c * (float)n);
    return !(f >= 0.0f) ? 0u : (f >= (float)(n - 1) ? n - 1 : (uint32_t)f);
}
// image k (0 reflectivity, 1 emissivity, 2 metallicity) of texture slot t at uv
// (raytrace.rs:465-467: atlas entry t * 3 + k)
void tex_sample(const or_scene *s, uint32_t t, uint32_t k, const float uv[2], float out[4]) {
    const uint32_t i = texel_index(uv[0], s->tex_w), j = texel_index(uv[1], s->tex_h);
    const size_t at = (((size_t)t * 3 + k) * s->tex_h + j) * s->tex_w + i;
    std::memcpy(out, &s->tex[at * 4], 16);
}

struct BounceOut {
    v3 origin, dir, normal, emis, refl;
    v3 alb = V(0, 0, 0);  // material reflectivity at the hit (learned-cache inputs)
    float mis, pdf;
    uint32_t hit;
    float t;
};

// raytrace.rs:390-577 (+ optional Russian roulette of the benchmark configs)
BounceOut raytrace_px(const or_scene *s, const Cfg &c, uint32_t bounce, uint32_t nee_type,
                      uint32_t inv_seed, uint32_t bid, v3 origin, v3 dir) {
    BounceOut o;
    o.hit = 0xFFFFFFFFu;
    o.t = 0.0f;
    uint32_t seed = combine(inv_seed, bid);
    if (is_zero(dir)) {  // raytrace.rs:406-416
        o.origin = origin; o.dir = V(0, 0, 0); o.normal = V(0, 0, 0); o.emis = V(0, 0, 0);
        o.refl = V(0, 0, 0); o.mis = 0.0f; o.pdf = 1.0f;
        return o;
    }
    Hit h = closest_bvh(s, origin, dir, EPS_RT, 1000.0f, nullptr);
    if (h.prim == 0xFFFFFFFFu) {  // raytrace.rs:421-430
        o.origin = origin + dir * 5000.0f;
        o.dir = V(0, 0, 0);
        o.normal = V(0, 0, 0);
        float sky = dot(dir, V(0, 1, 0)) > 0.9f ? 50.0f : 0.0f;
        o.emis = V(sky, sky, sky);
        o.refl = V(0, 0, 0);
        o.mis = 0.0f;
        o.pdf = 1.0f;
        return o;
    }
    o.hit = h.prim;
    o.t = h.t;
    v3 b3 = V((1.0f - h.u) - h.v, h.u, h.v);
    v3 t0 = tri_v(s->tri, h.prim, 0), t1 = tri_v(s->tri, h.prim, 1), t2 = tri_v(s->tri, h.prim, 2);
    IcS ics = local_cs(t0, t1, t2);
    const uint32_t slot = s->tri_t[h.prim];
    v3 new_origin = (t0 * b3.x + t1 * b3.y) + t2 * b3.z;
    v3 new_dir;
    // the slot's three textures at the hit (raytrace.rs:437-481), or its constant material
    float t0v[4], t1v[4], t2v[4];
    if (!s->tex.empty()) {
        const float *q = &s->tri_uv[(size_t)h.prim * 6];
        float uv[2];
        for (int a = 0; a < 2; ++a) uv[a] = (q[a] * b3.x + q[2 + a] * b3.y) + q[4 + a] * b3.z;
        tex_sample(s, slot, 0, uv, t0v);
        tex_sample(s, slot, 1, uv, t1v);
        tex_sample(s, slot, 2, uv, t2v);
    } else {
        const Material &mt = s->mats[slot];
        std::memcpy(t0v, mt.refl, 16);
        std::memcpy(t1v, mt.emis, 16);
        std::memcpy(t2v, mt.metal, 16);
    }
    v3 refl = V(t0v[0], t0v[1], t0v[2]);
    o.alb = refl;
    float alpha = t0v[3];
    v3 emis = (1000.0f * V(t1v[0], t1v[1], t1v[2])) * -dot(dir, ics.n);
    float metal = t2v[0];
    float mis = 0.0f, pdf;
    float albedo_max;
    float u0 = finalizef(combine(seed, 0));
    if (u0 < metal) {
        pdf = 1.0f;
        new_dir = reflect(dir, ics.n);
        albedo_max = std::max(std::max(refl.x, refl.y), refl.z);
    } else if (u0 < metal + (1.0f - alpha)) {
        new_dir = dir;
        refl = V(1, 1, 1);
        pdf = 1.0f;
        albedo_max = 1.0f;
    } else {
        new_origin = new_origin + (EPS_RT * 1.5f) * ics.n;
        albedo_max = std::max(std::max(refl.x, refl.y), refl.z);
        refl = refl / PI_F;
        Pick pk{false, 0, 0, 1.0f, 0.0f};
        if (nee_type == 1 || (nee_type == 2 && bounce == 1))
            pk = traverse_light(s, new_origin, ics.n, combine(seed, 2));
        if (pk.ok && pk.importance > 0.0f) mis = 0.3f;
        float u3 = finalizef(combine(seed, 3));
        float ua = finalizef(combine(seed, 4)), ub = finalizef(combine(seed, 5));
        if (u3 < mis) {
            size_t lg = s->inst_tri_off[pk.inst] + pk.prim;  // world copy of the light triangle
            v3 l0 = tri_v(s->tri, lg, 0), l1 = tri_v(s->tri, lg, 1), l2 = tri_v(s->tri, lg, 2);
            float x = ua, y = ub;  // raytrace.rs:305-311
            if (x + y > 1.0f) { x = 1.0f - x; y = 1.0f - y; }
            v3 bb = V((1.0f - x) - y, x, y);
            v3 pt = (bb.x * l0 + bb.y * l1) + bb.z * l2;
            new_dir = normalize(pt - new_origin);
        } else {  // raytrace.rs:286-291, 335-338
            float u = 2.0f * PI_F * ua;
            float vv = std::sqrt(std::max(0.0f, 1.0f - ub * ub));
            float sn, cs;
            det_sincos(u, &sn, &cs);
            v3 hs = V(vv * cs, ub, vv * sn);
            new_dir = normalize((hs.x * ics.t + hs.y * ics.n) + hs.z * ics.b);
        }
        float cos_theta = dot(new_dir, ics.n);
        pdf = cos_theta / PI_F;
    }
    if (bounce >= c.rr_start) {  // Russian roulette (benchmark configs; not in the reference)
        float surv = std::min(1.0f, std::max(c.rr_min, albedo_max));
        float urr = finalizef(combine(seed, 6));
        if (urr >= surv) new_dir = V(0, 0, 0);
        else refl = refl / surv;
    }
    o.origin = new_origin;
    o.dir = new_dir;
    o.normal = ics.n;
    o.emis = emis;
    o.refl = refl;
    o.mis = mis;
    o.pdf = pdf;
    return o;
}

// reservoir word layout: x_v 0-2, n_v 3-5, x_s 6-8, n_s 9-11, l_o_hat 12-14, p_omega 15,
// seed 16 (u32 bits), ucw 17, m 18 (u32 bits), w_sum 19
struct Sample { v3 xv, nv, xs, ns, lo; float pw; uint32_t seed; };
struct Reservoir { Sample z; float ucw; uint32_t m; float wsum; };
inline Reservoir res_load(const float *p) {
    Reservoir r;
    r.z.xv = ld3(p); r.z.nv = ld3(p + 3); r.z.xs = ld3(p + 6); r.z.ns = ld3(p + 9);
    r.z.lo = ld3(p + 12); r.z.pw = p[15]; r.z.seed = fbits(p[16]);
    r.ucw = p[17]; r.m = fbits(p[18]); r.wsum = p[19];
    return r;
}
inline void res_store(float *p, const Reservoir &r) {
    st3(p, r.z.xv); st3(p + 3, r.z.nv); st3(p + 6, r.z.xs); st3(p + 9, r.z.ns);
    st3(p + 12, r.z.lo); p[15] = r.z.pw; p[16] = bitsf(r.z.seed);
    p[17] = r.ucw; p[18] = bitsf(r.m); p[19] = r.wsum;
}
inline float luminance(v3 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
inline void update_res(uint32_t seed, Reservoir &r, const Sample &z, float w) {
    r.wsum += w;  // restir_temporal_resampling.rs:235-246
    r.m += 1;
    if (fconstruct(seed) < w / r.wsum) r.z = z;
}
// computeJacobianQToR (restir_spatial_resampling.rs:299-311): reuse of q's sample at
// pixel r; sky samples (n_s = 0) give 1. The reference's form: cosine ratio times the
// ratio of the distances (not their squares).
inline float gi_jacobian(const Sample &q, const Sample &r) {
    if (is_zero(q.ns)) return 1.0f;
    v3 qv = q.xv - q.xs, rv = r.xv - q.xs;
    float cq = std::fabs(dot(normalize(qv), q.ns));
    float cr = std::fabs(dot(normalize(rv), q.ns));
    if (mut(OR_MUT_GI_JAC_SQ)) return (cr / cq) * (dot(qv, qv) / dot(rv, rv));
    if (mut(OR_MUT_GI_JAC_SWAP)) return (cq / cr) * (length(rv) / length(qv));
    return (cr / cq) * (length(qv) / length(rv));
}

}  // namespace

struct or_renderer {
    Cfg c;
    std::vector<float> temporal;  // [3][P][20]
    std::vector<float> di_hist;   // [2][P][8] ReSTIR DI history: light, y xyz, W, M, -, -
    // view 4 (accumulated path tracer): running sum, frames in it, and what it was taken with
    std::vector<float> view_acc;  // [P][3]
    uint32_t view_n = 0, view_prev = 0;
    float view_cam[12] = {0};
    const or_scene *view_scene = nullptr;
    uint64_t view_version = 0;
};

namespace {

// ---------------------------------------------------------------- ReSTIR DI (benchmark mode)
// Not in the reference (its ReSTIR is GI). Assembled from the reference's primitives:
// candidates from traverseBvh (raytrace.rs:218-281) + triangleSample (:305-311),
// updateReservoir / mergeReservoir (restir_*_resampling.rs), geometricallySimilar
// (restir_spatial_resampling.rs:277-288). Target p_hat = luminance of the unshadowed
// contribution f * Le * cos_x * cos_y / d^2 with f = albedo / pi and
// Le = 1000 * emissivity (raytrace.rs:477); source pdf = pick probability / area.
struct DiRes {
    uint32_t light = 0xFFFFFFFFu;
    v3 y{0, 0, 0};
    float wsum = 0.0f, W = 0.0f, ph = 0.0f;
    uint32_t M = 0;
};

inline v3 di_contrib(const or_scene *s, v3 x, v3 n, v3 fr, uint32_t lg, v3 y) {
    v3 l0 = tri_v(s->tri, lg, 0), l1 = tri_v(s->tri, lg, 1), l2 = tri_v(s->tri, lg, 2);
    v3 nl = normalize(cross(l1 - l0, l2 - l0));
    v3 w = y - x;
    float d2 = dot(w, w);
    float dist = std::sqrt(d2);
    v3 wn = w / dist;
    float cx = std::max(0.0f, dot(n, wn));
    float cy = mut(OR_MUT_DI_COSY) ? 1.0f : std::max(0.0f, -dot(nl, wn));
    const Material &m = s->mats[s->tri_t[lg]];
    v3 le = 1000.0f * V(m.emis[0], m.emis[1], m.emis[2]);
    return (fr * le) * ((cx * cy) / d2);
}

inline void di_update(uint32_t seed, DiRes &r, uint32_t light, v3 y, float w, float ph) {
    r.wsum += w;
    r.M += 1;
    if (fconstruct(seed) < w / r.wsum) { r.light = light; r.y = y; r.ph = ph; }
}

struct DiInputs {  // bounce-0 buffers of the frame ([B][P] layouts, bounce 0 first)
    const std::vector<uint32_t> &hit;
    const std::vector<float> &origin, &normal, &emis;
};

void restir_di(or_renderer *rr, const or_scene *s, const or_prefs *prefs, uint32_t frame,
               const DiInputs &f, std::vector<float> &fin, std::vector<uint32_t> &di_light,
               std::vector<float> &di_w) {
    const Cfg &c = rr->c;
    const size_t P = (size_t)c.W * c.H;
    if (rr->di_hist.size() != 2 * P * 8) rr->di_hist.assign(2 * P * 8, 0.0f);
    float *hist_prev = &rr->di_hist[(size_t)((frame + 1) % 2) * P * 8];
    float *hist_cur = &rr->di_hist[(size_t)(frame % 2) * P * 8];
    const uint32_t dseed = frame_seed(c.seed, frame, 4);
    const float cos25 = 0.906307787036650f;
    std::vector<DiRes> T(P);
    std::vector<v3> X(P), N(P), FR(P);
    std::vector<uint8_t> ok(P, 0);
    // initial candidates + temporal merge
    for (size_t id = 0; id < P; ++id) {
        uint32_t hp = f.hit[id];
        if (hp == 0xFFFFFFFFu) continue;
        ok[id] = 1;
        const Material &mt = s->mats[s->tri_t[hp]];
        v3 x = ld3(&f.origin[(1 * P + id) * 3]), n = ld3(&f.normal[id * 3]);
        v3 fr = V(mt.refl[0], mt.refl[1], mt.refl[2]) / PI_F;
        X[id] = x; N[id] = n; FR[id] = fr;
        const uint32_t pd = combine(dseed, (uint32_t)id);
        DiRes r;
        for (uint32_t k = 0; k < c.di_cand; ++k) {
            uint32_t sk = combine(pd, k);
            Pick pk = traverse_light(s, x, n, combine(sk, 0));
            float w = 0.0f, ph = 0.0f;
            uint32_t lg = 0xFFFFFFFFu;
            v3 y = V(0, 0, 0);
            if (pk.ok) {
                lg = s->inst_tri_off[pk.inst] + pk.prim;
                v3 l0 = tri_v(s->tri, lg, 0), l1 = tri_v(s->tri, lg, 1), l2 = tri_v(s->tri, lg, 2);
                float ux = finalizef(combine(sk, 1)), uy = finalizef(combine(sk, 2));
                if (ux + uy > 1.0f) { ux = 1.0f - ux; uy = 1.0f - uy; }
                v3 bb = V((1.0f - ux) - uy, ux, uy);
                y = (bb.x * l0 + bb.y * l1) + bb.z * l2;
                float area = 0.5f * length(cross(l1 - l0, l2 - l0));
                float psrc = mut(OR_MUT_DI_AREA) ? pk.prob : pk.prob / area;
                ph = luminance(di_contrib(s, x, n, fr, lg, y));
                w = psrc > 0.0f ? ph / psrc : 0.0f;
            }
            di_update(finalize(combine(sk, 3)), r, lg, y, w, ph);
        }
        // temporal: the previous frame's final reservoir of this pixel, M capped at 20x
        const float *h = hist_prev + id * 8;
        uint32_t hl = fbits(h[0]);
        if (hl != 0xFFFFFFFFu && hl < s->tri_inst.size()) {
            uint32_t hm = std::min(fbits(h[5]), 20u * c.di_cand);
            v3 hy = V(h[1], h[2], h[3]);
            float ph = luminance(di_contrib(s, x, n, fr, hl, hy));
            uint32_t m0 = r.M;
            di_update(finalize(combine(pd, 0x10000u)), r, hl, hy, (ph * h[4]) * (float)hm, ph);
            r.M = m0 + hm;
        }
        r.W = r.ph > 0.0f ? (mut(OR_MUT_DI_UCW_NO_M) ? r.wsum / r.ph : r.wsum / ((float)r.M * r.ph)) : 0.0f;
        T[id] = r;
    }
    // spatial merge, visibility, final radiance, history
    for (size_t id = 0; id < P; ++id) {
        float *hc = hist_cur + id * 8;
        v3 E0 = ld3(&f.emis[id * 3]);
        if (!ok[id]) {
            hc[0] = bitsf(0xFFFFFFFFu); hc[1] = hc[2] = hc[3] = hc[4] = 0.0f; hc[5] = bitsf(0u);
            st3(&fin[id * 3], E0);
            continue;
        }
        const uint32_t pd = combine(dseed, (uint32_t)id);
        uint32_t qx = (uint32_t)(id % c.W), qy = (uint32_t)(id / c.W);
        DiRes r = T[id];
        v3 x = X[id], n = N[id], fr = FR[id];
        for (uint32_t it = 0; it < prefs->restir_spatial_iterations; ++it) {
            uint32_t is = combine(combine(pd, 0x20000u), it);
            float jx = c.radius * (finalizef(combine(is, 0)) - 0.5f);
            float jy = c.radius * (finalizef(combine(is, 1)) - 0.5f);
            int nx = (int)(((float)qx + 0.5f) + jx), ny = (int)(((float)qy + 0.5f) + jy);
            nx = std::min(std::max(nx, 0), (int)c.W - 1);
            ny = std::min(std::max(ny, 0), (int)c.H - 1);
            size_t q = (size_t)ny * c.W + nx;
            if (!ok[q]) continue;
            if (is_zero(n) || is_zero(N[q]) || dot(n, N[q]) < cos25) continue;
            const DiRes &rq = T[q];
            if (rq.light == 0xFFFFFFFFu) continue;
            float ph = luminance(di_contrib(s, x, n, fr, rq.light, rq.y));
            uint32_t m0 = r.M;
            di_update(finalize(combine(is, 2)), r, rq.light, rq.y, (ph * rq.W) * (float)rq.M, ph);
            r.M = mut(OR_MUT_DI_SPATIAL_M) ? m0 + 1u : m0 + rq.M;
        }
        r.W = r.ph > 0.0f ? r.wsum / ((float)r.M * r.ph) : 0.0f;
        v3 L = E0;
        if (r.light != 0xFFFFFFFFu && r.ph > 0.0f) {
            v3 w = r.y - x;
            float dist = std::sqrt(dot(w, w));
            Hit sh = closest_bvh(s, x, w / dist, EPS_RT, dist - 0.002f, nullptr);
            if (sh.prim == 0xFFFFFFFFu) L = E0 + di_contrib(s, x, n, fr, r.light, r.y) * r.W;
        }
        st3(&fin[id * 3], L);
        di_light[id] = r.light;
        di_w[id * 4 + 0] = r.wsum;
        di_w[id * 4 + 1] = r.W;
        di_w[id * 4 + 2] = (float)r.M;
        di_w[id * 4 + 3] = r.ph;
        hc[0] = bitsf(r.light); hc[1] = r.y.x; hc[2] = r.y.y; hc[3] = r.y.z; hc[4] = r.W;
        hc[5] = bitsf(r.M);
    }
}

}  // namespace

namespace {

struct FrameBufs {
    size_t P;
    uint32_t B;
    std::vector<float> origin, dir, normal, emis, refl, mis, pdf, nee, lout, q, t;
    std::vector<uint32_t> hit, plen;
    std::vector<float> alb;                 // learned cache: [B][P][3]
    std::vector<std::vector<float>> qrec, trec;  // learned cache: per pixel records
    void alloc(size_t p, uint32_t b) {
        P = p; B = b;
        alb.assign(b * p * 3, 0);
        origin.assign((b + 1) * p * 3, 0); dir.assign((b + 1) * p * 3, 0);
        normal.assign(b * p * 3, 0); emis.assign(b * p * 3, 0); refl.assign(b * p * 3, 0);
        mis.assign(b * p, 0); pdf.assign(b * p, 0); nee.assign(b * p, 0);
        lout.assign(b * p * 3, 0); q.assign(b * p, 0); t.assign(b * p, 0);
        hit.assign(b * p, 0xFFFFFFFFu); plen.assign(p, 0);
    }
};

// one pixel, all bounces + nee_pdf + outgoing radiance. `id` is the pixel index in
// the frame (raygen coordinates and RNG stream); `slot` indexes the buffers of f.
void trace_pixel(const or_scene *s, const Cfg &c, uint32_t nee_type, uint32_t frame, size_t id,
                 size_t slot, const v3 &eye, const v3 &front, const v3 &right, const v3 &up,
                 FrameBufs &f, const or_cache_frame *cf = nullptr) {
    const size_t P = f.P;
    bool train = false;
    uint32_t cap = c.B;  // bounces this pixel's path may take
    if (cf) {
        train = finalizef(combine(cf->train_seed, (uint32_t)id)) < cf->train_fraction ||
                (cf->fallback && cf->fallback[id]);
        cap = train ? c.B : cf->query_vertex + 1u;
    }
    uint32_t x = (uint32_t)(id % c.W), y = (uint32_t)(id / c.W);
    // raygen.rs:80-111 (jitter is 0.00 * (...), i.e. zero)
    float uvx = 2.0f * (float)x / (float)c.W - 1.0f;
    float uvy = 2.0f * (float)y / (float)c.H - 1.0f;
    float aspect = (float)c.W / (float)c.H;
    v3 d0 = normalize((right * (uvx + 0.0f) * aspect + up * (uvy + 0.0f)) + front);
    st3(&f.origin[slot * 3], eye);
    st3(&f.dir[slot * 3], d0);
    uint32_t plen = 0;
    for (uint32_t b = 0; b < cap; ++b) {
        size_t bi = b * P + slot, bo = (b + 1) * P + slot;
        v3 o = ld3(&f.origin[bi * 3]), d = ld3(&f.dir[bi * 3]);
        if (!is_zero(d)) plen = b + 1;
        BounceOut r = raytrace_px(s, c, b, nee_type, frame * c.B + b, (uint32_t)id, o, d);
        if (cf) st3(&f.alb[bi * 3], r.alb);
        st3(&f.origin[bo * 3], r.origin);
        st3(&f.dir[bo * 3], r.dir);
        st3(&f.normal[bi * 3], r.normal);
        st3(&f.emis[bi * 3], r.emis);
        st3(&f.refl[bi * 3], r.refl);
        f.mis[bi] = r.mis;
        f.pdf[bi] = r.pdf;
        f.hit[bi] = r.hit;
        f.t[bi] = r.t;
    }
    f.plen[slot] = plen;
    // nee_pdf.rs:281-337, bounces 0..B-2 (the last bounce's nee_pdf is never written)
    std::vector<LHit> hits;
    for (uint32_t b = 0; b + 1 < cap; ++b) {
        size_t bi = b * P + slot, bo = (b + 1) * P + slot;
        v3 d = ld3(&f.dir[bo * 3]);
        float w = f.mis[bi];
        float np = 0.0f;
        if (!(length(d) == 0.0f || w == 0.0f))
            np = nee_pdf_ray(s, ld3(&f.origin[bo * 3]), ld3(&f.normal[bi * 3]), d, hits);
        f.nee[bi] = np;
    }
    // outgoing_radiance.rs:58-94
    v3 L = V(0, 0, 0);
    const float bsp = 1.0f / (2.0f * PI_F);
    const uint32_t qv = cf ? cf->query_vertex : 0u;
    const bool queried = cf && !train && plen == qv + 1 && f.hit[qv * P + slot] != 0xFFFFFFFFu;
    for (int b = (int)c.B - 1; b >= 0; --b) {
        size_t bid = (size_t)b * P + slot;
        float valid = is_zero(ld3(&f.dir[bid * 3])) ? 0.0f : 1.0f;
        float bp = f.pdf[bid], np = f.nee[bid], w = f.mis[bid];
        float q = np * w + (1.0f - w) * bsp;
        float rf = bp / q;
        v3 term = ((ld3(&f.refl[bid * 3]) * L) * rf) * valid;
        if (queried && b == (int)qv) term = ld3(cf->cache_rad + id * 3);
        L = ld3(&f.emis[bid * 3]) + term;
        st3(&f.lout[bid * 3], L);
        f.q[bid] = q;
        // training record of vertex b (its reflected radiance is known unless the path was
        // cut by the bounce cap right after it)
        if (cf && train && b < (int)plen && f.hit[bid] != 0xFFFFFFFFu &&
            (b + 1 < (int)plen || is_zero(ld3(&f.dir[(bid + P) * 3]))) &&
            std::isfinite(term.x) && std::isfinite(term.y) && std::isfinite(term.z)) {
            v3 p = ld3(&f.origin[(bid + P) * 3]), dd = ld3(&f.dir[bid * 3]);
            v3 nn = ld3(&f.normal[bid * 3]), aa = ld3(&f.alb[bid * 3]);
            uint32_t key = (uint32_t)id * 16u + (uint32_t)b;
            float kf;
            std::memcpy(&kf, &key, 4);
            f.trec[slot].insert(f.trec[slot].begin(), {p.x, p.y, p.z, dd.x, dd.y, dd.z, nn.x, nn.y, nn.z, aa.x, aa.y,
                                                       aa.z, term.x, term.y, term.z, kf});
        }
    }
    if (queried) {
        size_t b1 = qv * P + slot;
        v3 p = ld3(&f.origin[(b1 + P) * 3]), dd = ld3(&f.dir[b1 * 3]);
        v3 nn = ld3(&f.normal[b1 * 3]), aa = ld3(&f.alb[b1 * 3]);
        uint32_t key = (uint32_t)id * 16u + qv;
        float kf;
        std::memcpy(&kf, &key, 4);
        f.qrec[slot] = {p.x, p.y, p.z, dd.x, dd.y, dd.z, nn.x, nn.y, nn.z, aa.x, aa.y, aa.z, 0.0f, 0.0f, 0.0f, kf};
    }
}

}  // namespace

extern "C" {

uint32_t or_murmur3_combine(uint32_t h, uint32_t k) { return combine(h, k); }
uint32_t or_murmur3_finalize(uint32_t h) { return finalize(h); }
float or_float_construct(uint32_t m) { return fconstruct(m); }
void or_sincos(float x, float *s, float *c) { det_sincos(x, s, c); }

int or_build_bl_bvh(uint64_t n, const float *tri, const float *lum, const uint32_t *ids,
                    void *nodes_out, uint32_t *order_out, float *root_box6, float *root_power,
                    uint32_t *depth_out) {
    if (n == 0) return -1;
    std::vector<BvhNode> out;
    std::vector<uint32_t> order;
    Aabb rb;
    float rp = 0.0f;
    uint32_t depth = 0;
    build_bl(n, tri, lum, ids, out, &order, &rb, &rp, &depth);
    std::memcpy(nodes_out, out.data(), out.size() * sizeof(BvhNode));
    if (order_out) std::memcpy(order_out, order.data(), n * 4);
    if (root_box6) { for (int k = 0; k < 3; ++k) { root_box6[k] = rb.mn[k]; root_box6[3 + k] = rb.mx[k]; } }
    if (root_power) *root_power = rp;
    if (depth_out) *depth_out = depth;
    return 0;
}

int or_build_tl_bvh(uint64_t m, const float *xf, const float *aabbs, const float *pw,
                    const uint32_t *ids, void *nodes_out) {
    if (m == 0) return -1;
    std::vector<Aabb> boxes(m);
    for (size_t i = 0; i < m; ++i) {
        boxes[i].empty = false;
        for (int k = 0; k < 3; ++k) { boxes[i].mn[k] = aabbs[6 * i + k]; boxes[i].mx[k] = aabbs[6 * i + 3 + k]; }
    }
    std::vector<BvhNode> out;
    build_tl(m, xf, boxes, pw, ids, out);
    std::memcpy(nodes_out, out.data(), out.size() * sizeof(BvhNode));
    return 0;
}

int or_find_best_plane(uint64_t n, const float *tri, uint32_t *dim_out, float *split_out) {
    std::vector<Aabb> boxes(n);
    std::vector<float> cen(3 * n);
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; ++i) {
        v3 v[3] = {canon(ld3(tri + 9 * i)), canon(ld3(tri + 9 * i + 3)), canon(ld3(tri + 9 * i + 6))};
        boxes[i] = aabb_from_points(v, 3);
        v3 c = ((v[0] + v[1]) + v[2]) / 3.0f;
        cen[3 * i] = c.x; cen[3 * i + 1] = c.y; cen[3 * i + 2] = c.z;
        idx[i] = i;
    }
    size_t d;
    float sp;
    find_best_plane(0, n, idx, cen, boxes, &d, &sp);
    *dim_out = (uint32_t)d;
    *split_out = sp;
    return 0;
}

// chunk.rs:158-292; positions are chunk-local (the object isometry carries the offset)
int or_gen_mesh_chunk(const uint16_t *blocks, const uint32_t dims[3], uint32_t n_types,
                      const uint8_t *translucent, uint32_t cx, uint32_t cy, uint32_t cz,
                      void *out_verts, uint64_t cap, uint64_t *n_out) {
    const uint32_t AIR = n_types;
    auto get = [&](long x, long y, long z) -> uint32_t {
        if (x < 0 || y < 0 || z < 0 || x >= (long)dims[0] || y >= (long)dims[1] || z >= (long)dims[2])
            return AIR;
        return blocks[((size_t)x * dims[1] + (size_t)y) * dims[2] + (size_t)z];
    };
    auto transl = [&](uint32_t b) { return b == AIR || translucent[b] != 0; };  // block.rs:129-131
    Vertex3D *out = (Vertex3D *)out_verts;
    uint64_t n = 0;
    auto push = [&](const float p[3], uint32_t t, float u, float v) {
        if (n < cap) {
            Vertex3D &w = out[n];
            w.pos[0] = p[0]; w.pos[1] = p[1]; w.pos[2] = p[2];
            w.t = t; w.uv[0] = u; w.uv[1] = v;
        }
        ++n;
    };
    for (uint32_t x = 0; x < 32; ++x)
        for (uint32_t y = 0; y < 32; ++y)
            for (uint32_t z = 0; z < 32; ++z) {
                long gx = (long)cx * 32 + x, gy = (long)cy * 32 + y, gz = (long)cz * 32 + z;
                uint32_t blk = get(gx, gy, gz);
                if (blk == AIR) continue;  // completely_transparent
                float fx = (float)x, fy = (float)y, fz = (float)z;
                float v000[3] = {fx + 0.0f, fy + 0.0f, fz + 0.0f}, v100[3] = {fx + 1.0f, fy + 0.0f, fz + 0.0f};
                float v001[3] = {fx + 0.0f, fy + 0.0f, fz + 1.0f}, v101[3] = {fx + 1.0f, fy + 0.0f, fz + 1.0f};
                float v010[3] = {fx + 0.0f, fy + 1.0f, fz + 0.0f}, v110[3] = {fx + 1.0f, fy + 1.0f, fz + 0.0f};
                float v011[3] = {fx + 0.0f, fy + 1.0f, fz + 1.0f}, v111[3] = {fx + 1.0f, fy + 1.0f, fz + 1.0f};
                uint32_t base = blk * 6;  // get_material_offset, block.rs:116-119
                if (transl(get(gx - 1, gy, gz))) {  // left
                    uint32_t t = base + 0;
                    push(v001, t, 0, 1); push(v010, t, 1, 0); push(v000, t, 1, 1);
                    push(v011, t, 0, 0); push(v010, t, 1, 0); push(v001, t, 0, 1);
                }
                if (transl(get(gx + 1, gy, gz))) {  // right
                    uint32_t t = base + 1;
                    push(v110, t, 0, 0); push(v101, t, 1, 1); push(v100, t, 0, 1);
                    push(v110, t, 0, 0); push(v111, t, 1, 0); push(v101, t, 1, 1);
                }
                if (transl(get(gx, gy - 1, gz))) {  // down
                    uint32_t t = base + 2;
                    push(v000, t, 0, 0); push(v100, t, 1, 0); push(v001, t, 0, 1);
                    push(v100, t, 1, 0); push(v101, t, 1, 1); push(v001, t, 0, 1);
                }
                if (transl(get(gx, gy + 1, gz))) {  // up
                    uint32_t t = base + 3;
                    push(v011, t, 1, 1); push(v110, t, 0, 0); push(v010, t, 1, 0);
                    push(v011, t, 1, 1); push(v111, t, 0, 1); push(v110, t, 0, 0);
                }
                if (transl(get(gx, gy, gz - 1))) {  // back
                    uint32_t t = base + 4;
                    push(v010, t, 0, 0); push(v100, t, 1, 1); push(v000, t, 0, 1);
                    push(v010, t, 0, 0); push(v110, t, 1, 0); push(v100, t, 1, 1);
                }
                if (transl(get(gx, gy, gz + 1))) {  // front
                    uint32_t t = base + 5;
                    push(v001, t, 1, 1); push(v101, t, 0, 1); push(v011, t, 1, 0);
                    push(v101, t, 0, 1); push(v111, t, 0, 0); push(v011, t, 1, 0);
                }
            }
    *n_out = n;
    return n <= cap ? 0 : -4;
}

or_scene *or_scene_create(const void *materials, const float *lum, uint32_t n_slots) {
    or_scene *s = new or_scene();
    const Material *m = (const Material *)materials;
    s->mats.assign(m, m + n_slots);
    s->lum.assign(lum, lum + n_slots);
    return s;
}
void or_scene_destroy(or_scene *s) { delete s; }

int or_scene_set_textures(or_scene *s, uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8) {
    s->tex.clear();
    s->tex_w = s->tex_h = s->tex_slots = 0;
    if (!rgba8 || n_slots == 0) return 0;
    if (w == 0 || h == 0 || n_slots < s->mats.size()) return -1;
    const size_t texels = (size_t)n_slots * 3 * w * h;
    s->tex.resize(texels * 4);
    for (size_t q = 0; q < texels; ++q)
        for (int c = 0; c < 4; ++c)
            s->tex[q * 4 + c] = c < 3 ? srgb_decode(rgba8[q * 4 + c]) : (float)((double)rgba8[q * 4 + c] / 255.0);
    s->tex_w = w;
    s->tex_h = h;
    s->tex_slots = n_slots;
    return 0;
}
float or_srgb_decode(uint32_t v) { return srgb_decode((uint8_t)v); }
uint32_t or_texel_index(float c, uint32_t n) { return texel_index(c, n); }
// utils.rs:223-235: per slot, the mean over the emissivity image of r + g + b in u8 units
void or_texture_luminances(uint32_t n_slots, uint32_t w, uint32_t h, const uint8_t *rgba8, float *out) {
    for (uint32_t t = 0; t < n_slots; ++t) {
        const uint8_t *e = rgba8 + ((size_t)t * 3 + 1) * w * h * 4;
        float l = 0.0f;
        for (size_t q = 0; q < (size_t)w * h; ++q) l += ((float)e[q * 4] + (float)e[q * 4 + 1]) + (float)e[q * 4 + 2];
        out[t] = l / (float)(w * h);
    }
}

int or_scene_add(or_scene *s, uint32_t key, const void *verts, uint64_t n, const float xf[12]) {
    if (n % 3 != 0) return -1;
    Inst in;
    std::memcpy(in.xf, xf, 48);
    if (n > 0) in.obj = upload(s, (const Vertex3D *)verts, n);
    s->objects[key] = in;
    s->version += 1;
    return 0;
}
int or_scene_remove(or_scene *s, uint32_t key) {
    s->objects.erase(key);
    s->version += 1;
    return 0;
}
int or_scene_build(or_scene *s) {
    scene_build(s);
    return 0;
}

uint64_t or_scene_count(or_scene *s, int what) {
    switch (what) {
        case 0: return s->tri_inst.size();
        case 1: return s->ltri_inst.size();
        case 2: return s->inst.size();
        case 3: return s->tl.size();
        case 4: return s->lgeom.size();
        case 5: return s->geom.size();
        case 6: return s->geom_depth;
        default: return 0;
    }
}

// what: 0 geometry BVH nodes, 1 geometry prim order, 2 TL nodes, 3 light-geometry nodes,
// 4 world triangles (9 floats), 5 tri instance ids, 6 instance tl idx, 7 light tri ids
// (instance, j pairs), 8 light-geometry prim order
int or_scene_read(or_scene *s, int what, void *dst, uint64_t bytes) {
    const void *src = nullptr;
    size_t n = 0;
    std::vector<uint32_t> tmp;
    switch (what) {
        case 0: src = s->geom.data(); n = s->geom.size() * 40; break;
        case 1: src = s->geom_order.data(); n = s->geom_order.size() * 4; break;
        case 2: src = s->tl.data(); n = s->tl.size() * 40; break;
        case 3: src = s->lgeom.data(); n = s->lgeom.size() * 40; break;
        case 4: src = s->tri.data(); n = s->tri.size() * 4; break;
        case 5: src = s->tri_inst.data(); n = s->tri_inst.size() * 4; break;
        case 6: src = s->inst_tl_idx.data(); n = s->inst_tl_idx.size() * 4; break;
        case 7:
            for (size_t i = 0; i < s->ltri_inst.size(); ++i) { tmp.push_back(s->ltri_inst[i]); tmp.push_back(s->ltri_j[i]); }
            src = tmp.data(); n = tmp.size() * 4; break;
        case 8: src = s->lgeom_order.data(); n = s->lgeom_order.size() * 4; break;
        default: return -1;
    }
    if (bytes < n) return -4;
    if (n) std::memcpy(dst, src, n);
    return 0;
}

// light BL-BVH of the object stored under `key` (built at upload, scene.rs:616-621)
int or_scene_light_bvh(or_scene *s, uint32_t key, void *dst, uint64_t cap, uint64_t *n) {
    auto it = s->objects.find(key);
    if (it == s->objects.end() || !it->second.obj) { *n = 0; return 0; }
    const std::vector<BvhNode> &b = it->second.obj->light_bvh;
    *n = b.size();
    if (dst && cap >= b.size()) std::memcpy(dst, b.data(), b.size() * sizeof(BvhNode));
    return 0;
}

int or_trace_closest(or_scene *s, uint64_t n, const float *o, const float *d, float tmin,
                     float tmax, int brute, uint32_t *hit, float *t, float *uv,
                     uint64_t *visited) {
    uint64_t vis = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : vis)
    for (long i = 0; i < (long)n; ++i) {
        uint64_t vv = 0;
        Hit h = brute ? closest_brute(s, ld3(o + 3 * i), ld3(d + 3 * i), tmin, tmax)
                      : closest_bvh(s, ld3(o + 3 * i), ld3(d + 3 * i), tmin, tmax, &vv);
        vis += vv;
        hit[i] = h.prim;
        if (t) t[i] = h.t;
        if (uv) { uv[2 * i] = h.u; uv[2 * i + 1] = h.v; }
    }
    if (visited) *visited = vis;
    return 0;
}

int or_trace_lights(or_scene *s, const float *o, const float *d, float tmin, float tmax,
                    uint32_t *hits, uint32_t cap, uint32_t *n_hits, float *uvs) {
    std::vector<LHit> h;
    light_hits(s, ld3(o), ld3(d), tmin, tmax, h);
    *n_hits = (uint32_t)h.size();
    for (size_t i = 0; i < h.size() && i < cap; ++i) {
        hits[i] = h[i].id;
        if (uvs) { uvs[2 * i] = h[i].u; uvs[2 * i + 1] = h[i].v; }
    }
    return h.size() <= cap ? 0 : -4;
}

int or_traverse_light_bvh(or_scene *s, const float *p, const float *n, uint32_t seed,
                          uint32_t *inst, uint32_t *prim, float *prob, float *imp) {
    Pick pk = traverse_light(s, ld3(p), ld3(n), seed);
    *inst = pk.inst; *prim = pk.prim; *prob = pk.prob; *imp = pk.importance;
    return pk.ok ? 1 : 0;
}

float or_reverse_traverse_light_bvh(or_scene *s, const float *p, const float *n, uint32_t inst,
                                    uint32_t bl_node) {
    return reverse_traverse(s, ld3(p), ld3(n), inst, bl_node);
}

void or_set_mutation(int m) { g_mut = m; }
int or_get_mutation(void) { return g_mut; }

int or_nee_pdf(or_scene *s, uint64_t n, const float *o, const float *nrm, const float *d, float *out) {
#pragma omp parallel
    {
        std::vector<LHit> hits;
#pragma omp for schedule(static)
        for (long i = 0; i < (long)n; ++i)
            out[i] = nee_pdf_ray(s, ld3(o + 3 * i), ld3(nrm + 3 * i), ld3(d + 3 * i), hits);
    }
    return 0;
}

float or_restir_gi_jacobian(const float *q_res, const float *r_res) {
    return gi_jacobian(res_load(q_res).z, res_load(r_res).z);
}

or_renderer *or_renderer_create(const or_config *cfg) {
    or_renderer *r = new or_renderer();
    r->c = Cfg{cfg->width, cfg->height, cfg->num_bounces, cfg->rr_start_bounce,
               cfg->rr_min_survival, cfg->restir_mode, cfg->restir_spatial_radius,
               cfg->di_candidates, cfg->seed, cfg->flags};
    r->temporal.assign((size_t)3 * cfg->width * cfg->height * 20, 0.0f);
    return r;
}
void or_renderer_destroy(or_renderer *r) { delete r; }

static int render_impl(or_renderer *rr, or_scene *s, const or_camera *cam, const or_prefs *prefs,
                       uint64_t frame64, or_frame *out, int threads, or_cache_frame *cf) {
    const Cfg &c = rr->c;
    if (c.B < 2) return -1;
    size_t P = (size_t)c.W * c.H;
    uint32_t frame = (uint32_t)frame64;
    FrameBufs f;
    f.alloc(P, c.B);
    if (cf) {
        f.qrec.assign(P, {});
        f.trec.assign(P, {});
    }
    v3 eye = ld3(cam->eye), front = ld3(cam->front), right = ld3(cam->right), up = ld3(cam->up);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (long id = 0; id < (long)P; ++id)
        trace_pixel(s, c, prefs->nee_type, frame, (size_t)id, (size_t)id, eye, front, right, up, f, cf);
    if (cf) {
        if (cf->albedo) std::memcpy(cf->albedo, f.alb.data(), f.alb.size() * 4);
        cf->n_qrec = 0;
        cf->n_trec = 0;
        for (size_t id = 0; id < P; ++id) {
            if (!f.qrec[id].empty()) {
                if (cf->qrec) std::memcpy(cf->qrec + cf->n_qrec * 16, f.qrec[id].data(), 64);
                cf->n_qrec += 1;
            }
            for (size_t k = 0; k < f.trec[id].size(); k += 16) {
                if (cf->trec && cf->n_trec < cf->trec_cap) std::memcpy(cf->trec + cf->n_trec * 16, &f.trec[id][k], 64);
                cf->n_trec += 1;
            }
        }
    }

    // ReSTIR GI temporal (restir_temporal_resampling.rs:256-301) on slot frame % 3,
    // with the initial sample assembled as interactive_rendering.rs:1359-1414
    float *tr = &rr->temporal[(size_t)(frame64 % 3) * P * 20];
    uint32_t tseed = frame_seed(c.seed, frame, 2);
    for (size_t id = 0; id < P; ++id) {
        uint32_t ps = combine(tseed, (uint32_t)id);
        Sample S;
        S.xv = ld3(&f.origin[(1 * P + id) * 3]);
        S.nv = ld3(&f.normal[(0 * P + id) * 3]);
        S.xs = ld3(&f.origin[(2 * P + id) * 3]);
        S.ns = ld3(&f.normal[(1 * P + id) * 3]);
        S.lo = ld3(&f.lout[(1 * P + id) * 3]);
        S.pw = f.q[0 * P + id];
        S.seed = 0;
        Reservoir R = res_load(&tr[id * 20]);
        if (finalizef(combine(ps, 0)) < 0.4f) { R.wsum = 0.0f; R.m = 0; }
        float w = luminance(S.lo) / S.pw;
        update_res(finalize(combine(ps, 1)), R, S, w);
        float ph = luminance(R.z.lo);
        R.ucw = ph > 0.0f ? (mut(OR_MUT_GI_UCW_NO_M) ? R.wsum / ph : R.wsum / ((float)R.m * ph)) : 0.0f;
        res_store(&tr[id * 20], R);
    }
    // spatial (restir_spatial_resampling.rs:316-406)
    std::vector<float> sp(P * 20);
    uint32_t sseed = frame_seed(c.seed, frame, 3);
    const float cos25 = 0.906307787036650f;
    for (size_t id = 0; id < P; ++id) {
        uint32_t qx = (uint32_t)(id % c.W), qy = (uint32_t)(id / c.W);
        uint32_t ps = combine(sseed, (uint32_t)id);
        Reservoir Rs = res_load(&tr[id * 20]);
        Sample S = Rs.z;
        std::vector<Reservoir> Q{Rs};
        for (uint32_t it = 0; it < prefs->restir_spatial_iterations; ++it) {
            uint32_t is = combine(ps, it);
            float jx = c.radius * (finalizef(combine(is, 0)) - 0.5f);
            float jy = c.radius * (finalizef(combine(is, 1)) - 0.5f);
            int nx = (int)(((float)qx + 0.5f) + jx), ny = (int)(((float)qy + 0.5f) + jy);
            nx = std::min(std::max(nx, 0), (int)c.W - 1);
            ny = std::min(std::max(ny, 0), (int)c.H - 1);
            Reservoir Rn = res_load(&tr[((size_t)ny * c.W + nx) * 20]);
            // geometricallySimilar, :277-288
            if (is_zero(S.nv) || is_zero(Rn.z.nv)) continue;
            if (dot(S.nv, Rn.z.nv) < cos25) continue;
            float J = std::min(std::max(gi_jacobian(Rn.z, S), 0.1f), 10.0f);
            float padj = luminance(Rn.z.lo) / J;
            uint32_t m0 = Rs.m;  // mergeReservoir, :266-275
            update_res(finalize(combine(is, 2)), Rs, Rn.z,
                       mut(OR_MUT_GI_MERGE_NO_M) ? padj * Rn.ucw : (padj * Rn.ucw) * (float)Rn.m);
            Rs.m = m0 + Rn.m;
            Q.push_back(Rn);
        }
        uint32_t Z = 0;
        for (const Reservoir &r : Q)
            if (luminance(r.z.lo) > 0.0f || mut(OR_MUT_GI_Z_ALL)) Z += r.m;
        float ph = luminance(Rs.z.lo);
        Rs.ucw = ph == 0.0f ? 0.0f : Rs.wsum / ((float)Z * ph);
        res_store(&sp[id * 20], Rs);
    }
    // finalize (restir_finalize.rs:121-141) and postprocess (postprocess.rs:39-88)
    std::vector<float> fin(P * 3), dbg(P * 3);
    for (size_t id = 0; id < P; ++id) {
        Reservoir R = res_load(&sp[id * 20]);
        v3 dir = normalize(R.z.xs - ld3(&f.origin[(1 * P + id) * 3]));
        float bp = mut(OR_MUT_GI_FINAL_COS) ? 1.0f / PI_F : dot(dir, R.z.nv) / PI_F;
        v3 L = ld3(&f.emis[id * 3]) + ((ld3(&f.refl[id * 3]) * R.z.lo) * R.ucw) * bp;
        st3(&fin[id * 3], L);
        float dv = R.ucw / 10.0f;
        st3(&dbg[id * 3], V(dv, dv, dv));
    }
    std::vector<uint32_t> di_light(P, 0xFFFFFFFFu);
    std::vector<float> di_w(P * 4, 0.0f);
    if (c.mode == 1) restir_di(rr, s, prefs, frame, DiInputs{f.hit, f.origin, f.normal, f.emis}, fin, di_light, di_w);
    // view 4: running mean of the path-traced radiance over the frames since the camera,
    // the scene or the view last changed (restarted on any of them)
    const uint32_t view = prefs->debug_view;
    if (view == 4) {
        const bool reset = rr->view_n == 0 || rr->view_prev != 4 || std::memcmp(rr->view_cam, cam, 48) != 0 ||
                           rr->view_scene != s || rr->view_version != s->version;
        if (rr->view_acc.size() != P * 3) rr->view_acc.assign(P * 3, 0.0f);
        rr->view_n = reset ? 1u : rr->view_n + 1u;
        for (size_t k = 0; k < P * 3; ++k) rr->view_acc[k] = reset ? f.lout[k] : rr->view_acc[k] + f.lout[k];
        std::memcpy(rr->view_cam, cam, 48);
        rr->view_scene = s;
        rr->view_version = s->version;
    }
    rr->view_prev = view;
    if (out->bgra8) {
        for (size_t id = 0; id < P; ++id) {
            v3 col = view == 0 ? ld3(&f.lout[id * 3])
                     : (view == 1 || view == 3) ? ld3(&fin[id * 3])
                     : view == 4 ? ld3(&rr->view_acc[id * 3]) / (float)rr->view_n : ld3(&dbg[id * 3]);
            col = col / (float)(1 * 1);
            if (view >= 3) col = V(col.x / (1.0f + col.x), col.y / (1.0f + col.y), col.z / (1.0f + col.z));
            auto cv = [](float v) -> uint8_t {
                float x = v * 255.0f;
                if (!(x > 0.0f)) return 0;
                if (x >= 255.0f) return 255;
                return (uint8_t)x;
            };
            out->bgra8[id * 4 + 0] = cv(col.z);
            out->bgra8[id * 4 + 1] = cv(col.y);
            out->bgra8[id * 4 + 2] = cv(col.x);
            out->bgra8[id * 4 + 3] = 255;
        }
    }
    auto cp = [](void *dst, const void *src, size_t bytes) { if (dst) std::memcpy(dst, src, bytes); };
    cp(out->origin, f.origin.data(), f.origin.size() * 4);
    cp(out->direction, f.dir.data(), f.dir.size() * 4);
    cp(out->normal, f.normal.data(), f.normal.size() * 4);
    cp(out->emissivity, f.emis.data(), f.emis.size() * 4);
    cp(out->reflectivity, f.refl.data(), f.refl.size() * 4);
    cp(out->mis_weight, f.mis.data(), f.mis.size() * 4);
    cp(out->bsdf_pdf, f.pdf.data(), f.pdf.size() * 4);
    cp(out->nee_pdf, f.nee.data(), f.nee.size() * 4);
    cp(out->l_out, f.lout.data(), f.lout.size() * 4);
    cp(out->q_omega, f.q.data(), f.q.size() * 4);
    cp(out->hit_prim, f.hit.data(), f.hit.size() * 4);
    cp(out->hit_t, f.t.data(), f.t.size() * 4);
    cp(out->path_len, f.plen.data(), f.plen.size() * 4);
    cp(out->temporal, tr, P * 20 * 4);
    cp(out->spatial, sp.data(), sp.size() * 4);
    cp(out->restir, fin.data(), fin.size() * 4);
    cp(out->debug, dbg.data(), dbg.size() * 4);
    cp(out->di_light, di_light.data(), P * 4);
    cp(out->di_weights, di_w.data(), P * 16);
    return 0;
}

int or_render(or_renderer *rr, or_scene *s, const or_camera *cam, const or_prefs *prefs, uint64_t frame64,
              or_frame *out, int threads) {
    return render_impl(rr, s, cam, prefs, frame64, out, threads, nullptr);
}

int or_render_cache(or_renderer *rr, or_scene *s, const or_camera *cam, const or_prefs *prefs, uint64_t frame64,
                    or_frame *out, int threads, or_cache_frame *cf) {
    return render_impl(rr, s, cam, prefs, frame64, out, threads, cf);
}

int or_render_rows(or_renderer *rr, or_scene *s, const or_camera *cam, const or_prefs *prefs,
                   uint64_t frame64, uint32_t y0, uint32_t y1, float *radiance, int threads) {
    const Cfg &c = rr->c;
    v3 eye = ld3(cam->eye), front = ld3(cam->front), right = ld3(cam->right), up = ld3(cam->up);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel
    {
        FrameBufs f;  // one pixel's buffers, reused
        f.alloc(1, c.B);
#pragma omp for schedule(dynamic, 16)
        for (long id = (long)y0 * c.W; id < (long)y1 * c.W; ++id) {
            trace_pixel(s, c, prefs->nee_type, (uint32_t)frame64, (size_t)id, 0, eye, front, right,
                        up, f);
            radiance[(id - (long)y0 * c.W) * 3 + 0] = f.lout[0];
            radiance[(id - (long)y0 * c.W) * 3 + 1] = f.lout[1];
            radiance[(id - (long)y0 * c.W) * 3 + 2] = f.lout[2];
        }
    }
    return 0;
}

}  // extern "C"
