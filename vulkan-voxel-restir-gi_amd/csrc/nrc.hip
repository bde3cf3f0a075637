// Learned radiance cache: input encoding, 4 x 64 ReLU MLP on the matrix cores, relative
// L2 loss, backpropagation and Adam. Not in the reference (it has no learned part).
//
// Encoding of a sample (64 inputs): 12-level hash grid x 2 features of the position in
// the cache box (dense index while (N+1)^3 <= T, else (x ^ 2654435761 y ^ 805459861 z)
// mod T), real spherical harmonics l < 4 of the ray direction and of the normal,
// albedo, a constant 1 (the first layer's bias), 4 zeros.
//
// MLP on the matrix cores, fp32 accumulation, in bf16 or fp16 (the precision P of every
// kernel below): gfx950's v_mfma_f32_16x16x32_{bf16,f16} for the layers' products,
// v_mfma_f32_16x16x16_{bf16,f16} where the contraction is 16 long. A wave works on
// blocks of 16 samples in the transposed form Y^T = W^T X^T: the A operand is 16 output
// neurons x inputs of the weights ([n][k]), the B operand inputs x 16 samples. The
// accumulator of output tile t (lane l: neurons 4(l/16)+j, sample l%16) is, after ReLU
// and rounding to 16 bits, the B operand of the next layer for inputs 16t + 4g + j, so
// activations never leave the registers. Lane (s = l%16, g = l/16) therefore owns the
// input features 16t + 4g + j, i.e. feature groups q = 4t + g of four: groups 0-5 are
// grid levels 2q and 2q+1, 6-9 direction SH, 10-13 normal SH, 14 albedo + 1, 15 zeros.
// A 32-long step pairs tiles 2u and 2u + 1: lane group g brings inputs 32u + 4g + j and
// 32u + 16 + 4g + j (j < 4) as its 8 elements, in B and, from the same [n][k] rows, in A;
// the instruction pairs A's and B's elements by (lane group, element), so the product is
// the same sum over k whatever order the hardware gives the 32 elements.
//
// Backward: the same trick with the weights in [k][n] layout (gradient w.r.t. the
// inputs of a layer, ReLU mask from the stored bf16 activations). Layer inputs and
// output gradients are written as bf16 feature-major planes [k][sample], from which a
// split-K MFMA kernel forms dW = X G^T. Gradients carry the loss scale until Adam.
#include "vrgi_kernels.h"

namespace vrgi {

typedef short bf16x4 __attribute__((ext_vector_type(4)));  // four 16-bit values (bf16 or fp16 bits)
typedef short h16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bfv8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16v4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16v8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// the 16-bit format of the network: P = 0 bf16, P = 1 fp16 (IEEE binary16, round to
// nearest even, overflow to infinity)
template <int P> struct Prec;
template <> struct Prec<0> {
    static __device__ __forceinline__ uint16_t cvt(float f) { return f2bf(f); }
    static __device__ __forceinline__ float back(uint16_t h) { return bf2f(h); }
    static __device__ __forceinline__ f32x4 mma16(bf16x4 a, bf16x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ f32x4 mma32(h16x8 a, h16x8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfv8, a), __builtin_bit_cast(bfv8, b), c,
                                                       0, 0, 0);
    }
};
template <> struct Prec<1> {
    static __device__ __forceinline__ uint16_t cvt(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
    static __device__ __forceinline__ float back(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }
    static __device__ __forceinline__ f32x4 mma16(bf16x4 a, bf16x4 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(f16v4, a), __builtin_bit_cast(f16v4, b), c,
                                                     0, 0, 0);
    }
    static __device__ __forceinline__ f32x4 mma32(h16x8 a, h16x8 b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16v8, a), __builtin_bit_cast(f16v8, b), c,
                                                      0, 0, 0);
    }
};
template <int P>
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)Prec<P>::cvt(a); r[1] = (short)Prec<P>::cvt(b);
    r[2] = (short)Prec<P>::cvt(c); r[3] = (short)Prec<P>::cvt(d);
    return r;
}
// packs and records whether a value did not survive the rounding (overflow, NaN)
template <int P>
__device__ __forceinline__ bf16x4 pack4c(float a, float b, float c, float d, bool &bad) {
    bf16x4 r = pack4<P>(a, b, c, d);
#pragma unroll
    for (int i = 0; i < 4; ++i) bad |= !isfinite(Prec<P>::back((uint16_t)r[i]));
    return r;
}
__device__ __forceinline__ bf16x4 ld_bf4(const uint16_t *p) { return *(const bf16x4 *)p; }
__device__ __forceinline__ h16x8 cat8(bf16x4 lo, bf16x4 hi) {
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}
// the two 32-long steps of a 64-long contraction along a [.][64] row at w (lane group g)
template <int P>
__device__ __forceinline__ f32x4 dot64(const uint16_t *w, uint32_t g, const bf16x4 x[4], f32x4 acc) {
    acc = Prec<P>::mma32(cat8(ld_bf4(w + 4 * g), ld_bf4(w + 16 + 4 * g)), cat8(x[0], x[1]), acc);
    acc = Prec<P>::mma32(cat8(ld_bf4(w + 32 + 4 * g), ld_bf4(w + 48 + 4 * g)), cat8(x[2], x[3]), acc);
    return acc;
}

// -------------------------------------------------------------------- encoding
// real spherical harmonics l < 4 of a unit vector, coefficients 4k .. 4k + 3
__device__ __forceinline__ f32x4 sh_group(float x, float y, float z, uint32_t k) {
    const float c0 = 0.28209479177387814f, c1 = 0.48860251190291992f;
    const float c2a = 1.0925484305920792f, c2b = 0.31539156525252005f, c2c = 0.54627421529603959f;
    const float c3a = 0.59004358992664352f, c3b = 2.8906114426405538f, c3c = 0.45704579946446572f;
    const float c3d = 0.37317633259011540f, c3e = 1.4453057213202769f;
    const float xx = x * x, yy = y * y, zz = z * z;
    f32x4 o;
    if (k == 0) {
        o[0] = c0; o[1] = -c1 * y; o[2] = c1 * z; o[3] = -c1 * x;
    } else if (k == 1) {
        o[0] = c2a * x * y; o[1] = -c2a * y * z; o[2] = c2b * (3.0f * zz - 1.0f); o[3] = -c2a * x * z;
    } else if (k == 2) {
        o[0] = c2c * (xx - yy); o[1] = -c3a * y * (3.0f * xx - yy); o[2] = c3b * x * y * z;
        o[3] = -c3c * y * (5.0f * zz - 1.0f);
    } else {
        o[0] = c3d * z * (5.0f * zz - 3.0f); o[1] = -c3c * x * (5.0f * zz - 1.0f); o[2] = c3e * z * (xx - yy);
        o[3] = -c3a * x * (xx - 3.0f * yy);
    }
    return o;
}

struct GridCorners {
    uint32_t idx[8];
    float w[8];
};
__device__ __forceinline__ GridCorners grid_corners(const NrcNet &net, const float p[3], uint32_t level) {
    GridCorners r;
    const uint32_t N = net.res[level];
    const uint32_t T = 1u << net.log2_T;
    const bool dense = (uint64_t)(N + 1) * (N + 1) * (N + 1) <= (uint64_t)T;
    uint32_t cell[3];
    float fr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float u = (p[a] - net.lo[a]) / (net.hi[a] - net.lo[a]);
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        float s = u * (float)N;
        float f = floorf(s);
        if (f > (float)(N - 1)) f = (float)(N - 1);
        cell[a] = (uint32_t)f;
        fr[a] = s - f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t g[3];
        float w = 1.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t bit = (k >> a) & 1u;
            g[a] = cell[a] + bit;
            w *= bit ? fr[a] : 1.0f - fr[a];
        }
        uint32_t idx = dense ? g[0] + g[1] * (N + 1) + g[2] * (N + 1) * (N + 1)
                             : ((g[0] * 1u) ^ (g[1] * 2654435761u) ^ (g[2] * 805459861u)) & (T - 1u);
        r.idx[k] = level * T + idx;
        r.w[k] = w;
    }
    return r;
}
__device__ __forceinline__ float2 grid_level(const NrcNet &net, const float p[3], uint32_t level) {
    GridCorners c = grid_corners(net, p, level);
    float2 s = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float2 t = *(const float2 *)(net.table + (size_t)c.idx[k] * 2);
        s.x += c.w[k] * t.x;
        s.y += c.w[k] * t.y;
    }
    return s;
}

struct Sample {
    float p[3], d[3], n[3], a[3];
};
__device__ __forceinline__ Sample load_sample(const float *rec) {
    const float4 *q = (const float4 *)rec;
    float4 x = q[0], y = q[1], z = q[2];
    Sample s;
    s.p[0] = x.x; s.p[1] = x.y; s.p[2] = x.z;
    s.d[0] = x.w; s.d[1] = y.x; s.d[2] = y.y;
    s.n[0] = y.z; s.n[1] = y.w; s.n[2] = z.x;
    s.a[0] = z.y; s.a[1] = z.z; s.a[2] = z.w;
    return s;
}

// the 4 features of group q (see the file comment)
__device__ __forceinline__ f32x4 feature_group(const NrcNet &net, const Sample &s, uint32_t q) {
    f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
    if (q < 6) {
        float2 a = grid_level(net, s.p, 2 * q), b = grid_level(net, s.p, 2 * q + 1);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    } else if (q < 10) {
        v = sh_group(s.d[0], s.d[1], s.d[2], q - 6);
    } else if (q < 14) {
        v = sh_group(s.n[0], s.n[1], s.n[2], q - 10);
    } else if (q == 14) {
        v[0] = s.a[0]; v[1] = s.a[1]; v[2] = s.a[2]; v[3] = 1.0f;
    }
    return v;
}

// encoding of one sample as the four B operands of layer 1 (lane group g)
template <int P>
__device__ __forceinline__ void encode(const NrcNet &net, const Sample &s, uint32_t g, bool live, bf16x4 x[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 v = live ? feature_group(net, s, 4 * t + g) : f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        x[t] = pack4<P>(v[0], v[1], v[2], v[3]);
    }
}

// one hidden layer forward: h[nt] = round(relu(sum_k Wt[16 nt + row][k] x[k]))
template <int P>
__device__ __forceinline__ void layer_fwd(const uint16_t *wt, const bf16x4 x[4], bf16x4 h[4], uint32_t lane) {
    const uint32_t row = lane & 15u, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = dot64<P>(wt + (16 * nt + row) * 64, g, x, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
        h[nt] = pack4<P>(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f), fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f));
    }
}
template <int P>
__device__ __forceinline__ f32x4 layer_out(const uint16_t *wt5, const bf16x4 x[4], uint32_t lane) {
    const uint32_t row = lane & 15u, g = lane >> 4;
    return dot64<P>(wt5 + row * 64, g, x, f32x4{0.0f, 0.0f, 0.0f, 0.0f});  // lanes 0..15: outputs 0..3
}

// -------------------------------------------------------------------- inference
// The five layers' 16-bit [n][k] weights in LDS, one row of 64 per output neuron padded
// to WROW halves: rows 34 words apart put the 16 rows one ds_read_b64 touches in distinct
// bank pairs. Row r of the packed weights (l * 64 + n, the output layer's rows 256 + n)
// is row r here.
static constexpr int WROW = 68;
static constexpr int WROWS = 4 * 64 + 16;
__device__ __forceinline__ void weights_to_lds(const uint16_t *wt, uint16_t *sw) {
    for (uint32_t i = threadIdx.x; i < NRC_N_WEIGHTS / 4; i += blockDim.x) {
        const uint32_t h = 4u * i;
        *(bf16x4 *)(sw + (h / 64u) * WROW + h % 64u) = *(const bf16x4 *)(wt + h);
    }
    __syncthreads();
}

// records: 16 floats (position, direction, normal, albedo, target, key); out: 3 floats
// per record; when out_px is set, also out_px[3 * key + c] (negative outputs clamp to 0)
// `count` may exceed the buffer's capacity `max_n` (producers count the records they drop).
// A wave takes 16 * INFER_NB records per iteration as INFER_NB blocks of 16 (the lane
// layout of the file comment within each block), so that every weight operand read from
// LDS serves INFER_NB MFMAs; each output is the same two 32-long steps in the same order
// as in training.
#ifndef VRGI_INFER_NB
#define VRGI_INFER_NB 4
#endif
#ifndef VRGI_INFER_WAVES
#define VRGI_INFER_WAVES 3
#endif
static constexpr int INFER_NB = VRGI_INFER_NB;
template <int P>
__global__ __launch_bounds__(256, VRGI_INFER_WAVES) void k_nrc_infer(NrcNet net, const float *rec, const uint32_t *count,
                                                   uint32_t max_n, float *out, float *out_px) {
    __shared__ __align__(16) uint16_t sw[WROWS * WROW];
    weights_to_lds(net.wt, sw);
    const uint32_t lane = threadIdx.x & 63u, s = lane & 15u, g = lane >> 4;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t n = min(*count, max_n);
    const uint32_t n_groups = (n + 16u * INFER_NB - 1u) / (16u * INFER_NB);
    for (uint32_t grp = wave; grp < n_groups; grp += n_waves) {
        // the weights are read from LDS where they are used, not kept in registers across
        // iterations (a compiler barrier: the LDS reads stay inside the loop)
        asm volatile("" ::: "memory");
        bf16x4 x[INFER_NB][4];
#pragma unroll
        for (int j = 0; j < INFER_NB; ++j) {
            const uint32_t i = grp * (16u * INFER_NB) + 16u * j + s;
            const bool live = i < n;
            Sample sm{};
            if (live) sm = load_sample(rec + (size_t)i * 16);
            encode<P>(net, sm, g, live, x[j]);
        }
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            bf16x4 h[INFER_NB][4];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const uint16_t *w = sw + (l * 64 + 16 * nt + s) * WROW;
                const h16x8 a0 = cat8(ld_bf4(w + 4 * g), ld_bf4(w + 16 + 4 * g));
                const h16x8 a1 = cat8(ld_bf4(w + 32 + 4 * g), ld_bf4(w + 48 + 4 * g));
#pragma unroll
                for (int j = 0; j < INFER_NB; ++j) {
                    f32x4 acc = Prec<P>::mma32(a0, cat8(x[j][0], x[j][1]), f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                    acc = Prec<P>::mma32(a1, cat8(x[j][2], x[j][3]), acc);
                    h[j][nt] = pack4<P>(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f), fmaxf(acc[2], 0.0f),
                                        fmaxf(acc[3], 0.0f));
                }
            }
#pragma unroll
            for (int j = 0; j < INFER_NB; ++j)
#pragma unroll
                for (int t = 0; t < 4; ++t) x[j][t] = h[j][t];
        }
        const uint16_t *w5 = sw + (256 + s) * WROW;
        const h16x8 a0 = cat8(ld_bf4(w5 + 4 * g), ld_bf4(w5 + 16 + 4 * g));
        const h16x8 a1 = cat8(ld_bf4(w5 + 32 + 4 * g), ld_bf4(w5 + 48 + 4 * g));
#pragma unroll
        for (int j = 0; j < INFER_NB; ++j) {
            f32x4 y = Prec<P>::mma32(a0, cat8(x[j][0], x[j][1]), f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            y = Prec<P>::mma32(a1, cat8(x[j][2], x[j][3]), y);  // lanes 0..15: outputs 0..3
            const uint32_t i = grp * (16u * INFER_NB) + 16u * j + s;
            if (g == 0 && i < n) {
                if (out) {
                    out[3 * (size_t)i] = y[0];
                    out[3 * (size_t)i + 1] = y[1];
                    out[3 * (size_t)i + 2] = y[2];
                }
                if (out_px) {
                    uint32_t px = __float_as_uint(rec[(size_t)i * 16 + 15]) >> 4;
                    out_px[3 * (size_t)px] = fmaxf(y[0], 0.0f);
                    out_px[3 * (size_t)px + 1] = fmaxf(y[1], 0.0f);
                    out_px[3 * (size_t)px + 2] = fmaxf(y[2], 0.0f);
                }
            }
        }
    }
}

// -------------------------------------------------------------------- training
// step k of `steps` uses records k, k + steps, k + 2 steps, ... (at most max_batch)
__device__ __forceinline__ uint32_t batch_size(uint32_t count, uint32_t cap, uint32_t k, uint32_t steps,
                                               uint32_t max_batch) {
    count = count < cap ? count : cap;
    uint32_t b = count > k ? (count - k + steps - 1) / steps : 0u;
    return b < max_batch ? b : max_batch;
}

// Adds (v0, v1) at tbl[2 idx], tbl[2 idx + 1] for each of the 16 lanes of a row (one
// level, one corner, 16 samples). Runs of neighbouring lanes with the same idx are summed
// first (a segmented prefix sum over the row) and only a run's last lane adds the sum:
// with records in Morton order the samples of a row share the cells of the coarse levels,
// and the adds, which go past the L2 to where the XCDs' copies of a line meet, are the
// cost of the training step. idx 0xFFFFFFFF marks a lane without a sample.
__device__ __forceinline__ void seg_atomic2(float *tbl, uint32_t idx, float v0, float v1, uint32_t s) {
    const uint32_t prev = __shfl_up(idx, 1, 16), next = __shfl_down(idx, 1, 16);
    int head = (s == 0 || prev != idx) ? 1 : 0;  // set: the summed window reaches the run's head
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
        const float a0 = __shfl_up(v0, d, 16), a1 = __shfl_up(v1, d, 16);
        const int hp = __shfl_up(head, d, 16);
        if (!head && s >= (uint32_t)d) {
            v0 += a0;
            v1 += a1;
            head = hp;
        }
    }
    if ((s == 15 || next != idx) && idx != 0xFFFFFFFFu) {
        atomicAdd(tbl + (size_t)idx * 2, v0);
        atomicAdd(tbl + (size_t)idx * 2 + 1, v1);
    }
}

template <int P>
__global__ __launch_bounds__(256) void k_nrc_fwd_bwd(NrcNet net, NrcTrain tr) {
    const uint32_t lane = threadIdx.x & 63u, s = lane & 15u, g = lane >> 4;
    // dynamic loss scaling: the scale of this step, and whether a scaled gradient or the
    // output did not fit the 16-bit format (the step is then skipped)
    const float lscale = tr.scale_state ? tr.scale_state[0] : tr.loss_scale;
    bool bad = false;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t bsz = batch_size(*tr.count, tr.rec_cap, tr.step_k, tr.steps, tr.max_batch);
    const uint32_t n_blocks = (bsz + 15u) / 16u;
    const uint32_t B = tr.max_batch;  // row stride of the planes
    const float inv = 1.0f / (3.0f * (float)(bsz ? bsz : 1u));
    float loss = 0.0f;
    for (uint32_t blk = wave; blk < n_blocks; blk += n_waves) {
        const uint32_t j = blk * 16u + s;
        const bool live = j < bsz;
        const uint32_t src = tr.step_k + j * tr.steps;
        Sample sm{};
        float tgt[3] = {0.0f, 0.0f, 0.0f};
        if (live) {
            const float *r = tr.rec + (size_t)src * 16;
            sm = load_sample(r);
            tgt[0] = r[12]; tgt[1] = r[13]; tgt[2] = r[14];
        }
        bf16x4 a[5][4];  // inputs of layers 1..5
        encode<P>(net, sm, g, live, a[0]);
#pragma unroll
        for (int l = 0; l < 4; ++l) layer_fwd<P>(net.wt + l * 4096, a[l], a[l + 1], lane);
        f32x4 y = layer_out<P>(net.wt + 4 * 4096, a[4], lane);
        // output gradient (relative L2, denominator not differentiated), scaled
        float go[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (g == 0 && live) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float e = y[c] - tgt[c];
                float den = tr.loss == 0 ? y[c] * y[c] + 0.01f : 1.0f;
                loss += e * e / den * inv;
                go[c] = lscale * 2.0f * e / den * inv;
                bad |= !isfinite(y[c]);
            }
        }
        bf16x4 gcur[4];
        gcur[0] = pack4c<P>(go[0], go[1], go[2], go[3], bad);
        // save layer-5 input and output gradient planes
        {
            uint16_t *X = tr.xplane + (size_t)4 * 64 * B;
            uint16_t *G = tr.gplane + (size_t)4 * 64 * B;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(size_t)(16 * t + 4 * g + r) * B + j] = (uint16_t)a[4][t][r];
#pragma unroll
            for (int r = 0; r < 4; ++r) G[(size_t)(4 * g + r) * B + j] = (uint16_t)gcur[0][r];
        }
        // backward through layer 5: d(h4) = W5 g5, masked by h4 > 0
        bf16x4 gn[4];
        {
            const uint16_t *wb5 = net.wb + 4 * 4096;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = Prec<P>::mma16(ld_bf4(wb5 + (16 * mt + s) * 16 + 4 * g), gcur[0], acc);
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = Prec<P>::back((uint16_t)a[4][mt][r]) > 0.0f ? acc[r] : 0.0f;
                gn[mt] = pack4c<P>(v[0], v[1], v[2], v[3], bad);
            }
        }
        // hidden layers 4..1: gn = gradient w.r.t. the output of layer l + 1
        f32x4 gin[4];
#pragma unroll
        for (int l = 3; l >= 0; --l) {
            uint16_t *X = tr.xplane + (size_t)l * 64 * B;
            uint16_t *G = tr.gplane + (size_t)l * 64 * B;
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    X[(size_t)(16 * t + 4 * g + r) * B + j] = (uint16_t)a[l][t][r];
                    G[(size_t)(16 * t + 4 * g + r) * B + j] = (uint16_t)gn[t][r];
                }
            const uint16_t *wb = net.wb + l * 4096;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                gin[mt] = dot64<P>(wb + (16 * mt + s) * 64, g, gn, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
                for (int r = 0; r < 4; ++r) bad |= !isfinite(gin[mt][r]);
            }
            if (l > 0) {
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    float v[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = Prec<P>::back((uint16_t)a[l][mt][r]) > 0.0f ? gin[mt][r] : 0.0f;
                    gn[mt] = pack4c<P>(v[0], v[1], v[2], v[3], bad);
                }
            }
        }
        // grid gradient: gin[mt][r] = d(loss S)/d(input 16 mt + 4 g + r); grid inputs < 24
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const uint32_t q = 4 * mt + g;  // feature group (uniform over the lanes of a row)
            if (q < 6) {
#pragma unroll
                for (int h2 = 0; h2 < 2; ++h2) {
                    GridCorners c = grid_corners(net, sm.p, 2 * q + h2);
                    const float g0 = gin[mt][2 * h2], g1 = gin[mt][2 * h2 + 1];
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        seg_atomic2(tr.gtable, live ? c.idx[k] : 0xFFFFFFFFu, c.w[k] * g0, c.w[k] * g1, s);
                }
            }
        }
    }
    // batch loss (unscaled)
    for (int off = 32; off >= 1; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0 && loss != 0.0f) atomicAdd(tr.loss_out, loss);
    if (tr.scale_state && __ballot(bad) && lane == 0) atomicOr((uint32_t *)(tr.scale_state + 2), 1u);
}

// dW of layer `layer` = X G^T over the batch (split K): one wave per (tile, K chunk)
template <int P>
__global__ __launch_bounds__(64) void k_nrc_dw(NrcTrain tr, uint32_t chunk) {
    const uint32_t lane = threadIdx.x, s = lane & 15u, g = lane >> 4;
    const uint32_t layer = blockIdx.z;
    const uint32_t ntiles_n = layer < 4 ? 4u : 1u;
    const uint32_t tile = blockIdx.y;  // mt * 4 + nt
    const uint32_t mt = tile >> 2, nt = tile & 3u;
    if (nt >= ntiles_n) return;
    const uint32_t bsz = batch_size(*tr.count, tr.rec_cap, tr.step_k, tr.steps, tr.max_batch);
    const uint32_t padded = (bsz + 15u) & ~15u;
    const uint32_t k0 = blockIdx.x * chunk;
    if (k0 >= padded) return;
    const uint32_t k1 = min(padded, k0 + chunk);
    const uint32_t B = tr.max_batch;
    const uint16_t *X = tr.xplane + (size_t)layer * 64 * B + (size_t)(16 * mt + s) * B;
    const uint16_t *G = tr.gplane + (size_t)layer * 64 * B + (size_t)(16 * nt + s) * B;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t i = k0; i < k1; i += 16) acc = Prec<P>::mma16(ld_bf4(X + i + 4 * g), ld_bf4(G + i + 4 * g), acc);
    // D[k = 16 mt + 4 g + r][n = 16 nt + s]
    const uint32_t ncols = layer < 4 ? 64u : 16u;
    float *gw = tr.gw + (size_t)layer * 4096;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t k = 16 * mt + 4 * g + r, n = 16 * nt + s;
        if (acc[r] != 0.0f) atomicAdd(gw + (size_t)k * ncols + n, acc[r]);
    }
}

// Adam (fp32): gradients carry the loss scale; they are cleared after use. With dynamic
// loss scaling (scale_state set) the scale is the step's own, and a step whose gradients
// overflowed is skipped (gradients cleared, parameters and moments unchanged).
__global__ __launch_bounds__(256) void k_nrc_adam(float *p, float *gr, float *m, float *v, size_t n, NrcAdam a,
                                                  const float *scale_state) {
    const bool skip = scale_state && *(const uint32_t *)(scale_state + 2) != 0u;
    const float inv_scale = scale_state ? 1.0f / scale_state[0] : a.inv_scale;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gg = gr[i] * inv_scale;
        gr[i] = 0.0f;
        if (skip) continue;
        const float mm = a.beta1 * m[i] + (1.0f - a.beta1) * gg;
        const float vv = a.beta2 * v[i] + (1.0f - a.beta2) * gg * gg;
        m[i] = mm;
        v[i] = vv;
        const float mh = mm / a.c1, vh = vv / a.c2;
        p[i] -= a.lr * mh / (sqrtf(vh) + a.eps);
    }
}

// dynamic loss scaling after a step (state: scale, good steps in a row, overflow flag,
// skipped steps): halve the scale (not below 1) and count the step as skipped when it
// overflowed, double it (up to 2^24) after `interval` good steps in a row
__global__ void k_nrc_scale_update(float *st, uint32_t interval) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t *flag = (uint32_t *)(st + 2);
    float scale = st[0], good = st[1];
    if (*flag) {
        scale = fmaxf(scale * 0.5f, 1.0f);
        good = 0.0f;
        *flag = 0u;
        st[3] += 1.0f;
    } else {
        good += 1.0f;
        if (good >= (float)interval) {
            scale = fminf(scale * 2.0f, 16777216.0f);
            good = 0.0f;
        }
    }
    st[0] = scale;
    st[1] = good;
}

// fp32 master weights -> 16-bit in both matrix-core layouts
template <int P>
__global__ __launch_bounds__(256) void k_nrc_pack(const float *w, uint16_t *wt, uint16_t *wb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NRC_N_WEIGHTS) return;
    const uint32_t layer = i / 4096, r = i % 4096;
    const uint32_t ncols = layer < 4 ? 64u : 16u;
    const uint32_t k = r / ncols, n = r % ncols;  // w is [k][n]
    if (layer == 4 && k >= 64) return;
    const uint16_t h = Prec<P>::cvt(w[i]);
    wb[i] = h;                                  // [k][n]
    wt[layer * 4096 + n * 64 + k] = h;          // [n][k]
}

// clear the gradient accumulators of one training step
__global__ __launch_bounds__(256) void k_nrc_zero(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

// per-tile sums (cache luminance, target luminance, count) of the records at `vertex`
__global__ __launch_bounds__(256) void k_tile_stats(const float *rec, const uint32_t *count, uint32_t cap,
                                                    const float *pred, uint32_t vertex, uint32_t W, uint32_t tile,
                                                    uint32_t tiles_x, float *acc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = min(*count, cap);
    if (i >= n) return;
    const uint32_t key = __float_as_uint(rec[(size_t)i * 16 + 15]);
    if ((key & 15u) != vertex) return;
    const uint32_t px = key >> 4;
    const uint32_t t = ((px / W) / tile) * tiles_x + (px % W) / tile;
    const float *y = pred + (size_t)i * 3, *g = rec + (size_t)i * 16 + 12;
    const float ly = 0.2126f * fmaxf(y[0], 0.0f) + 0.7152f * fmaxf(y[1], 0.0f) + 0.0722f * fmaxf(y[2], 0.0f);
    const float lg = 0.2126f * g[0] + 0.7152f * g[1] + 0.0722f * g[2];
    atomicAdd(acc + 3 * (size_t)t, ly);
    atomicAdd(acc + 3 * (size_t)t + 1, lg);
    atomicAdd(acc + 3 * (size_t)t + 2, 1.0f);
}
// running means per tile and the fallback flags they imply; clears the sums
__global__ __launch_bounds__(256) void k_tile_update(float *acc, float *ema, uint8_t *fallback, uint32_t n_tiles,
                                                     float alpha, float bound) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const float c = acc[3 * (size_t)t + 2];
    if (c > 0.0f) {
        ema[2 * (size_t)t] = (1.0f - alpha) * ema[2 * (size_t)t] + alpha * (acc[3 * (size_t)t] / c);
        ema[2 * (size_t)t + 1] = (1.0f - alpha) * ema[2 * (size_t)t + 1] + alpha * (acc[3 * (size_t)t + 1] / c);
        const float e = fabsf(ema[2 * (size_t)t] - ema[2 * (size_t)t + 1]) / (ema[2 * (size_t)t + 1] + 0.01f);
        fallback[t] = e > bound ? 1 : 0;
    }
    acc[3 * (size_t)t] = 0.0f;
    acc[3 * (size_t)t + 1] = 0.0f;
    acc[3 * (size_t)t + 2] = 0.0f;
}
void launch_tile_stats(hipStream_t s, const float *rec, const uint32_t *count, uint32_t cap, const float *pred,
                       uint32_t vertex, uint32_t W, uint32_t tile, uint32_t tiles_x, float *acc) {
    if (cap) k_tile_stats<<<(cap + 255) / 256, 256, 0, s>>>(rec, count, cap, pred, vertex, W, tile, tiles_x, acc);
}
void launch_tile_update(hipStream_t s, float *acc, float *ema, uint8_t *fallback, uint32_t n_tiles, float alpha,
                        float bound) {
    if (n_tiles) k_tile_update<<<(n_tiles + 255) / 256, 256, 0, s>>>(acc, ema, fallback, n_tiles, alpha, bound);
}

// ---------------------------------------------------------------- record ordering
// The frame's training records in Morton order of their position (32 cells per axis of
// the cache box): a counting sort (bucket sizes with the lane's rank inside its bucket,
// one-block exclusive scan, scatter). Each training step then takes every steps-th record
// of that order, an even sample of the whole scene whose neighbouring samples share the
// coarse grid cells (see seg_atomic2).
__device__ __forceinline__ uint32_t spread3_5(uint32_t v) {  // bit i -> bit 3i, 5 bits
    v &= 31u;
    v = (v | (v << 8)) & 0x0000f00fu;
    v = (v | (v << 4)) & 0x000c30c3u;
    v = (v | (v << 2)) & 0x00249249u;
    return v;
}
__global__ __launch_bounds__(256) void k_rec_clear(uint32_t *hist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NRC_SORT_BUCKETS) hist[i] = 0u;
}
__global__ __launch_bounds__(256) void k_rec_hist(const float *rec, const uint32_t *count, uint32_t cap,
                                                  RecBox box, uint32_t *hist, uint32_t *key_rank) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = min(*count, cap);
    if (i >= n) return;
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float t = (rec[(size_t)i * 16 + a] - box.lo[a]) / (box.hi[a] - box.lo[a]) * 32.0f;
        t = t == t ? fminf(fmaxf(t, 0.0f), 31.0f) : 0.0f;
        q[a] = (uint32_t)t;
    }
    const uint32_t key = spread3_5(q[0]) | (spread3_5(q[1]) << 1) | (spread3_5(q[2]) << 2);
    const uint32_t rank = atomicAdd(hist + key, 1u);
    key_rank[2 * (size_t)i] = key;
    key_rank[2 * (size_t)i + 1] = rank;
}
// exclusive scan of the bucket sizes in place, per block of 1024 buckets (four consecutive
// buckets per thread of 256, a wave scan of the threads' sums, then the waves' totals);
// the block's total goes to hist[NRC_SORT_BUCKETS + block]. 256-thread blocks: the scan
// runs beside the frame's ray kernels, where 1024-thread blocks wait for a CU to free
// 16 wave slots at once.
__global__ __launch_bounds__(256) void k_rec_scan(uint32_t *hist) {
    __shared__ uint32_t wsum[4];
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    uint4 *p = (uint4 *)(hist + (size_t)blockIdx.x * 1024u) + t;
    const uint4 c = *p;
    const uint32_t own = c.x + c.y + c.z + c.w;
    uint32_t inc = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(inc, d, 64);
        if (lane >= (uint32_t)d) inc += u;
    }
    if (lane == 63u) wsum[wv] = inc;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t k = 0; k < wv; ++k) base += wsum[k];
    const uint32_t e = base + inc - own;  // exclusive prefix of this thread's first bucket
    *p = make_uint4(e, e + c.x, e + c.x + c.y, e + c.x + c.y + c.z);
    if (t == 255u) hist[NRC_SORT_BUCKETS + blockIdx.x] = e + own;
}
__global__ __launch_bounds__(256) void k_rec_scatter(const float *rec, const uint32_t *count, uint32_t cap,
                                                     const uint32_t *start, const uint32_t *key_rank, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = min(*count, cap);
    if (i >= n) return;
    const uint32_t key = key_rank[2 * (size_t)i];
    uint32_t dst = start[key] + key_rank[2 * (size_t)i + 1];
    for (uint32_t blk = 0; blk < (key >> 10); ++blk) dst += start[NRC_SORT_BUCKETS + blk];  // earlier blocks
    const float4 *a = (const float4 *)(rec + (size_t)i * 16);
    float4 *b = (float4 *)(out + (size_t)dst * 16);
    b[0] = a[0]; b[1] = a[1]; b[2] = a[2]; b[3] = a[3];
}
void launch_rec_sort(hipStream_t s, const float *rec, const uint32_t *count, uint32_t cap, const RecBox &box,
                     uint32_t *hist, uint32_t *key_rank, float *out) {
    if (!cap) return;
    k_rec_clear<<<NRC_SORT_BUCKETS / 256, 256, 0, s>>>(hist);
    k_rec_hist<<<(cap + 255) / 256, 256, 0, s>>>(rec, count, cap, box, hist, key_rank);
    k_rec_scan<<<NRC_SORT_BUCKETS / 1024, 256, 0, s>>>(hist);
    k_rec_scatter<<<(cap + 255) / 256, 256, 0, s>>>(rec, count, cap, hist, key_rank, out);
}

template <class F>
static void with_prec(uint32_t prec, F f) {
    if (prec == 1) f(std::integral_constant<int, 1>());
    else f(std::integral_constant<int, 0>());
}

void launch_nrc_infer(hipStream_t s, const NrcNet &net, const float *rec, const uint32_t *count, uint32_t max_n,
                      float *out, float *out_px) {
    // one wave per 16 * INFER_NB records, up to one residency round (4 blocks per CU)
    uint32_t groups = (max_n + 16 * INFER_NB - 1) / (16 * INFER_NB);
    uint32_t grid = std::min<uint32_t>((groups + 3) / 4, 1024u);
    if (!grid) return;
    with_prec(net.prec, [&](auto Pc) { k_nrc_infer<decltype(Pc)::value><<<grid, 256, 0, s>>>(net, rec, count, max_n, out, out_px); });
}

void launch_nrc_grads(hipStream_t s, const NrcNet &net, const NrcTrain &tr) {
    uint32_t blocks16 = (tr.max_batch + 15) / 16;
    uint32_t grid = std::min<uint32_t>((blocks16 + 3) / 4, 2048u);
    const uint32_t chunk = 1024;
    dim3 g((tr.max_batch + 15 + chunk - 1) / chunk, 16, 5);
    with_prec(net.prec, [&](auto Pc) {
        constexpr int P = decltype(Pc)::value;
        k_nrc_fwd_bwd<P><<<grid, 256, 0, s>>>(net, tr);
        k_nrc_dw<P><<<g, 64, 0, s>>>(tr, chunk);
    });
}

void launch_nrc_pack(hipStream_t s, const float *w32, uint16_t *wt, uint16_t *wb, uint32_t prec) {
    with_prec(prec, [&](auto Pc) {
        k_nrc_pack<decltype(Pc)::value><<<(NRC_N_WEIGHTS + 255) / 256, 256, 0, s>>>(w32, wt, wb);
    });
}

void launch_nrc_train_step(hipStream_t s, const NrcNet &net, const NrcTrain &tr, float *w32, float *mw, float *vw,
                           float *table, float *mt, float *vt, size_t n_table, const NrcAdam &adam) {
    launch_nrc_grads(s, net, tr);
    k_nrc_adam<<<(NRC_N_WEIGHTS + 255) / 256, 256, 0, s>>>(w32, tr.gw, mw, vw, NRC_N_WEIGHTS, adam, tr.scale_state);
    k_nrc_adam<<<4096, 256, 0, s>>>(table, tr.gtable, mt, vt, n_table, adam, tr.scale_state);
    if (tr.scale_state) k_nrc_scale_update<<<1, 64, 0, s>>>(tr.scale_state, tr.scale_interval);
    launch_nrc_pack(s, w32, const_cast<uint16_t *>(net.wt), const_cast<uint16_t *>(net.wb), net.prec);
}

void launch_nrc_zero(hipStream_t s, float *p, size_t n) {
    if (n) k_nrc_zero<<<(uint32_t)std::min<size_t>((n + 255) / 256, 4096), 256, 0, s>>>(p, n);
}

}  // namespace vrgi
