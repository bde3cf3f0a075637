// CPU restatement of the learned radiance cache (test infrastructure only: the tests,
// the smoke check and the CPU baseline use it as the checker; the product never calls
// it). Every number is a double; see nrc.h for the bf16 emulation mode.
//
// Encoding of one sample (64 values):
//   [0, 2L)   multiresolution hash grid (Mueller et al. 2022): the position mapped to
//             the unit cube of the scene box; per level l with resolution N_l the cell
//             floor(p N_l) (clamped to N_l - 1), trilinear weights of its 8 corners, and
//             per corner (integer coords g) the table entry
//               dense  (N_l + 1)^3 <= T: g.x + g.y (N_l + 1) + g.z (N_l + 1)^2
//               hashed otherwise:        (g.x * 1 ^ g.y * 2654435761 ^ g.z * 805459861) mod T
//   [24, 40)  real spherical harmonics l < 4 of the ray direction
//   [40, 56)  the same of the surface normal
//   [56, 59)  albedo, [59] 1 (acts as the first layer's bias), [60, 64) zero
// MLP: h1..h4 = relu(W_l^T h_{l-1}), y = W_5^T h4 (3 outputs), no bias vectors.
#include "nrc.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int IN = OR_NRC_IN, WID = OR_NRC_WIDTH, HID = OR_NRC_HIDDEN, OUTP = OR_NRC_OUT_PAD;

double bf16(double x) {  // double -> float -> bfloat16, round to nearest even
    float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

// double -> float -> IEEE binary16, round to nearest even, overflow to infinity
double fp16(double x) {
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = u & 0x80000000u, a = u & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) return (double)f;                                  // inf, NaN
    const float af = std::fabs(f);
    if (af >= 65520.0f) return sign ? -HUGE_VAL : HUGE_VAL;                  // rounds past 65504
    if (af < 6.103515625e-05f) {                                             // subnormal: steps of 2^-24
        const double q = std::nearbyint((double)af * 16777216.0) / 16777216.0;
        return sign ? -q : q;
    }
    const uint32_t lsb = 1u << 13, rem = a & (lsb - 1u);                     // 23 -> 10 mantissa bits
    uint32_t base = a & ~(lsb - 1u);
    if (rem > lsb / 2 || (rem == lsb / 2 && (base & lsb))) base += lsb;
    const uint32_t r = sign | base;
    float out;
    std::memcpy(&out, &r, 4);
    return (double)out;
}

size_t w_off(int layer) { return (size_t)layer * IN * WID; }  // layer 0..4
int w_cols(int layer) { return layer < HID ? WID : OUTP; }

struct Corner { uint32_t idx[8]; double w[8]; };

}  // namespace

struct or_nrc {
    or_nrc_cfg c;
    std::vector<double> w, table, m_w, v_w, m_t, v_t;
    uint64_t step = 0;
    size_t T() const { return (size_t)1 << c.log2_T; }
    // the rounding of the GPU's 16-bit points: 1 bf16, 2 fp16, 0 none
    double rnd(double x) const { return c.emulate_bf16 == 1 ? bf16(x) : c.emulate_bf16 == 2 ? fp16(x) : x; }
    bool emu() const { return c.emulate_bf16 != 0; }
};

extern "C" {

double or_nrc_round16(double x, uint32_t mode) { return mode == 1 ? bf16(x) : mode == 2 ? fp16(x) : x; }

uint32_t or_nrc_hash(uint32_t x, uint32_t y, uint32_t z, uint32_t log2_T) {
    uint32_t h = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
    return log2_T >= 32 ? h : (h & ((1u << log2_T) - 1u));
}

void or_nrc_sh16(const double *v, double *o) {
    const double x = v[0], y = v[1], z = v[2], pi = 3.14159265358979323846;
    const double c0 = 0.5 * std::sqrt(1.0 / pi), c1 = std::sqrt(3.0 / (4.0 * pi));
    const double c2a = 0.5 * std::sqrt(15.0 / pi), c2b = 0.25 * std::sqrt(5.0 / pi), c2c = 0.25 * std::sqrt(15.0 / pi);
    const double c3a = 0.25 * std::sqrt(35.0 / (2.0 * pi)), c3b = 0.5 * std::sqrt(105.0 / pi);
    const double c3c = 0.25 * std::sqrt(21.0 / (2.0 * pi)), c3d = 0.25 * std::sqrt(7.0 / pi);
    const double c3e = 0.25 * std::sqrt(105.0 / pi);
    const double xx = x * x, yy = y * y, zz = z * z;
    o[0] = c0;
    o[1] = -c1 * y;
    o[2] = c1 * z;
    o[3] = -c1 * x;
    o[4] = c2a * x * y;
    o[5] = -c2a * y * z;
    o[6] = c2b * (3.0 * zz - 1.0);
    o[7] = -c2a * x * z;
    o[8] = c2c * (xx - yy);
    o[9] = -c3a * y * (3.0 * xx - yy);
    o[10] = c3b * x * y * z;
    o[11] = -c3c * y * (5.0 * zz - 1.0);
    o[12] = c3d * z * (5.0 * zz - 3.0);
    o[13] = -c3c * x * (5.0 * zz - 1.0);
    o[14] = c3e * z * (xx - yy);
    o[15] = -c3a * x * (xx - 3.0 * yy);
}

}  // extern "C"

namespace {

Corner corners(const or_nrc *m, const float *p, int level) {
    Corner r;
    const uint32_t N = m->c.res[level];
    const size_t T = m->T();
    const bool dense = (double)(N + 1) * (N + 1) * (N + 1) <= (double)T;
    uint32_t cell[3];
    double fr[3];
    for (int a = 0; a < 3; ++a) {
        double u = ((double)p[a] - m->c.box_lo[a]) / ((double)m->c.box_hi[a] - m->c.box_lo[a]);
        u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
        double s = u * N;
        double f = std::floor(s);
        if (f > N - 1) f = N - 1;
        cell[a] = (uint32_t)f;
        fr[a] = s - f;
    }
    for (int k = 0; k < 8; ++k) {
        uint32_t g[3];
        double w = 1.0;
        for (int a = 0; a < 3; ++a) {
            uint32_t bit = (k >> a) & 1u;
            g[a] = cell[a] + bit;
            w *= bit ? fr[a] : 1.0 - fr[a];
        }
        uint32_t idx = dense ? g[0] + g[1] * (N + 1) + g[2] * (N + 1) * (N + 1)
                             : or_nrc_hash(g[0], g[1], g[2], m->c.log2_T);
        r.idx[k] = (uint32_t)((size_t)level * T + idx);
        r.w[k] = w;
    }
    return r;
}

void encode_one(const or_nrc *m, const float *p, const float *d, const float *n, const float *a, double *x) {
    const uint32_t F = m->c.feats;
    for (uint32_t l = 0; l < m->c.levels; ++l) {
        Corner cr = corners(m, p, (int)l);
        for (uint32_t f = 0; f < F; ++f) {
            double s = 0.0;
            for (int k = 0; k < 8; ++k) s += cr.w[k] * m->table[(size_t)cr.idx[k] * F + f];
            x[l * F + f] = s;
        }
    }
    const uint32_t g = m->c.levels * F;
    double dv[3] = {d[0], d[1], d[2]}, nv[3] = {n[0], n[1], n[2]};
    or_nrc_sh16(dv, x + g);
    or_nrc_sh16(nv, x + g + 16);
    x[g + 32] = a[0];
    x[g + 33] = a[1];
    x[g + 34] = a[2];
    x[g + 35] = 1.0;
    for (int k = (int)g + 36; k < IN; ++k) x[k] = 0.0;
}

// forward of one sample; h[0] = input, h[1..4] hidden (post-ReLU), y = output (3)
void forward_one(const or_nrc *m, const double *x, double h[HID + 1][IN], double y[3]) {
    for (int k = 0; k < IN; ++k) h[0][k] = m->rnd(x[k]);
    for (int l = 0; l < HID; ++l) {
        const double *W = m->w.data() + w_off(l);
        for (int nn = 0; nn < WID; ++nn) {
            double s = 0.0;
            for (int k = 0; k < IN; ++k) s += h[l][k] * m->rnd(W[(size_t)k * WID + nn]);
            s = s > 0.0 ? s : 0.0;
            h[l + 1][nn] = m->emu() ? m->rnd(s) : s;
        }
    }
    const double *W5 = m->w.data() + w_off(HID);
    for (int nn = 0; nn < 3; ++nn) {
        double s = 0.0;
        for (int k = 0; k < WID; ++k) s += h[HID][k] * m->rnd(W5[(size_t)k * OUTP + nn]);
        y[nn] = m->emu() ? (double)(float)s : s;
    }
}

double grads_impl(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm, const float *alb,
                  const float *target, double *gw, double *gt) {
    const size_t NW = OR_NRC_N_WEIGHTS, NT = (size_t)m->c.levels * m->T() * m->c.feats;
    std::memset(gw, 0, NW * sizeof(double));
    std::memset(gt, 0, NT * sizeof(double));
    const double S = m->c.loss_scale;
    double loss = 0.0;
    std::vector<double> x(IN);
    double h[HID + 1][IN], y[3], g[IN], gn[IN];
    for (uint64_t i = 0; i < n; ++i) {
        encode_one(m, pos + 3 * i, dir + 3 * i, nrm + 3 * i, alb + 3 * i, x.data());
        forward_one(m, x.data(), h, y);
        // output gradient (loss mean over n samples x 3 channels), scaled by S
        double go[OUTP] = {0};
        for (int c = 0; c < 3; ++c) {
            double t = target[3 * i + c], e = y[c] - t;
            double den = m->c.loss == 0 ? y[c] * y[c] + 0.01 : 1.0;
            loss += e * e / den / (3.0 * n);
            double gg = S * 2.0 * e / den / (3.0 * n);
            go[c] = m->emu() ? m->rnd(gg) : gg;
        }
        // layer 5
        double *G5 = gw + w_off(HID);
        const double *W5 = m->w.data() + w_off(HID);
        for (int k = 0; k < WID; ++k) {
            for (int c = 0; c < 3; ++c) G5[(size_t)k * OUTP + c] += h[HID][k] * go[c];
            double s = 0.0;
            for (int c = 0; c < 3; ++c) s += m->rnd(W5[(size_t)k * OUTP + c]) * go[c];
            s = h[HID][k] > 0.0 ? s : 0.0;
            g[k] = m->emu() ? m->rnd(s) : s;
        }
        // hidden layers 4..1: g = gradient w.r.t. the output of layer l (post-mask)
        for (int l = HID - 1; l >= 0; --l) {
            double *G = gw + w_off(l);
            const double *W = m->w.data() + w_off(l);
            for (int k = 0; k < IN; ++k)
                for (int nn = 0; nn < WID; ++nn) G[(size_t)k * WID + nn] += h[l][k] * g[nn];
            for (int k = 0; k < IN; ++k) {
                double s = 0.0;
                for (int nn = 0; nn < WID; ++nn) s += m->rnd(W[(size_t)k * WID + nn]) * g[nn];
                if (l > 0) {
                    s = h[l][k] > 0.0 ? s : 0.0;
                    gn[k] = m->emu() ? m->rnd(s) : s;
                } else {
                    gn[k] = m->emu() ? (double)(float)s : s;  // input gradient, fp32
                }
            }
            std::memcpy(g, gn, sizeof(g));
        }
        // grid: g holds d(loss * S)/d(input)
        const uint32_t F = m->c.feats;
        for (uint32_t l = 0; l < m->c.levels; ++l) {
            Corner cr = corners(m, pos + 3 * i, (int)l);
            for (int k = 0; k < 8; ++k)
                for (uint32_t f = 0; f < F; ++f) gt[(size_t)cr.idx[k] * F + f] += cr.w[k] * g[l * F + f];
        }
    }
    for (size_t k = 0; k < NW; ++k) gw[k] /= S;
    for (size_t k = 0; k < NT; ++k) gt[k] /= S;
    // the padding columns of layer 5 carry no gradient
    double *G5 = gw + w_off(HID);
    for (int k = 0; k < WID; ++k)
        for (int c = 3; c < OUTP; ++c) G5[(size_t)k * OUTP + c] = 0.0;
    return loss;
}

void adam(const or_nrc_cfg &c, uint64_t t, double *p, double *mm, double *vv, const double *g, size_t n) {
    const double b1 = c.beta1, b2 = c.beta2;
    const double c1 = 1.0 - std::pow(b1, (double)t), c2 = 1.0 - std::pow(b2, (double)t);
    for (size_t k = 0; k < n; ++k) {
        mm[k] = b1 * mm[k] + (1.0 - b1) * g[k];
        vv[k] = b2 * vv[k] + (1.0 - b2) * g[k] * g[k];
        double mh = mm[k] / c1, vh = vv[k] / c2;
        p[k] -= c.lr * mh / (std::sqrt(vh) + c.eps);
    }
}

}  // namespace

extern "C" {

or_nrc *or_nrc_create(const or_nrc_cfg *cfg, const float *weights, const float *table) {
    or_nrc *m = new or_nrc();
    m->c = *cfg;
    const size_t NT = (size_t)cfg->levels * m->T() * cfg->feats;
    m->w.assign(weights, weights + OR_NRC_N_WEIGHTS);
    m->table.assign(table, table + NT);
    m->m_w.assign(OR_NRC_N_WEIGHTS, 0.0);
    m->v_w.assign(OR_NRC_N_WEIGHTS, 0.0);
    m->m_t.assign(NT, 0.0);
    m->v_t.assign(NT, 0.0);
    return m;
}

void or_nrc_destroy(or_nrc *m) { delete m; }

void or_nrc_encode(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm, const float *alb,
                   double *enc) {
    for (uint64_t i = 0; i < n; ++i) encode_one(m, pos + 3 * i, dir + 3 * i, nrm + 3 * i, alb + 3 * i, enc + IN * i);
}

void or_nrc_forward(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm, const float *alb,
                    double *out) {
    std::vector<double> x(IN);
    double h[HID + 1][IN];
    for (uint64_t i = 0; i < n; ++i) {
        encode_one(m, pos + 3 * i, dir + 3 * i, nrm + 3 * i, alb + 3 * i, x.data());
        forward_one(m, x.data(), h, out + 3 * i);
    }
}

double or_nrc_grads(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm, const float *alb,
                    const float *target, double *grad_w, double *grad_table) {
    return grads_impl(m, n, pos, dir, nrm, alb, target, grad_w, grad_table);
}

double or_nrc_train_step(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm,
                         const float *alb, const float *target) {
    const size_t NT = m->table.size();
    std::vector<double> gw(OR_NRC_N_WEIGHTS), gt(NT);
    double loss = grads_impl(m, n, pos, dir, nrm, alb, target, gw.data(), gt.data());
    m->step += 1;
    adam(m->c, m->step, m->w.data(), m->m_w.data(), m->v_w.data(), gw.data(), OR_NRC_N_WEIGHTS);
    adam(m->c, m->step, m->table.data(), m->m_t.data(), m->v_t.data(), gt.data(), NT);
    return loss;
}

void or_nrc_get(const or_nrc *m, double *weights, double *table) {
    if (weights) std::memcpy(weights, m->w.data(), m->w.size() * sizeof(double));
    if (table) std::memcpy(table, m->table.data(), m->table.size() * sizeof(double));
}

void or_nrc_set(or_nrc *m, const double *weights, const double *table) {
    if (weights) std::memcpy(m->w.data(), weights, m->w.size() * sizeof(double));
    if (table) std::memcpy(m->table.data(), table, m->table.size() * sizeof(double));
}

}  // extern "C"
