// CPU restatement (test infrastructure only) of the learned radiance cache: input
// encoding, the 4 x 64 ReLU MLP, the relative L2 loss, backpropagation and Adam, in
// double precision. With emulate_bf16 set, inputs, weights, activations and output
// gradients are rounded to bfloat16 (round to nearest even) at the points where the
// GPU kernels feed them to the matrix cores, so the two can be compared closely; the
// sums stay in double. Nothing in the reference corresponds to this part (it has no
// learned component); the layout below is this project's own.
#pragma once
#include <cstddef>
#include <cstdint>

extern "C" {

#define OR_NRC_IN 64      // encoded inputs
#define OR_NRC_WIDTH 64   // hidden width
#define OR_NRC_HIDDEN 4   // hidden layers
#define OR_NRC_OUT_PAD 16 // output columns (3 used)
// weights: layers 1..4 are 64 x 64, layer 5 is 64 x 16, each [k][n] row-major (k = input)
#define OR_NRC_N_WEIGHTS (OR_NRC_HIDDEN * OR_NRC_IN * OR_NRC_WIDTH + OR_NRC_WIDTH * OR_NRC_OUT_PAD)

typedef struct or_nrc_cfg {
    uint32_t levels;       // grid levels (12)
    uint32_t feats;        // features per level (2)
    uint32_t log2_T;       // entries per level = 2^log2_T
    uint32_t res[16];      // grid resolution of each level
    float box_lo[3], box_hi[3];  // scene box mapped to the unit cube
    float lr, beta1, beta2, eps; // Adam
    float loss_scale;      // multiplies the output gradient; divided out before Adam
    uint32_t loss;         // 0: relative L2 (pred - y)^2 / (sg(pred)^2 + 0.01); 1: plain L2
    uint32_t emulate_bf16; // round at the GPU's 16-bit points: 1 bf16, 2 fp16 (IEEE binary16), 0 none
} or_nrc_cfg;

typedef struct or_nrc or_nrc;

or_nrc *or_nrc_create(const or_nrc_cfg *cfg, const float *weights, const float *table);
void or_nrc_destroy(or_nrc *m);
uint32_t or_nrc_hash(uint32_t x, uint32_t y, uint32_t z, uint32_t log2_T);
// the 16-bit rounding of emulate mode `mode` (1 bf16, 2 fp16) of a double
double or_nrc_round16(double x, uint32_t mode);
void or_nrc_sh16(const double *v, double *out16);
// n samples, each: position (3), direction (3), normal (3), albedo (3)
void or_nrc_encode(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm,
                   const float *alb, double *enc /* n x 64 */);
void or_nrc_forward(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm,
                    const float *alb, double *out /* n x 3 */);
// loss of the batch and its gradients (weights layout, dense table L x T x F)
double or_nrc_grads(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm,
                    const float *alb, const float *target, double *grad_w, double *grad_table);
// gradients, then one Adam step on weights and table; returns the loss before the step
double or_nrc_train_step(or_nrc *m, uint64_t n, const float *pos, const float *dir, const float *nrm,
                         const float *alb, const float *target);
void or_nrc_get(const or_nrc *m, double *weights, double *table);
void or_nrc_set(or_nrc *m, const double *weights, const double *table);

}  // extern "C"
