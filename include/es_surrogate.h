/*
 * es_surrogate.h — C ABI of the learned sigma(B) surrogate (libes_hip.so).
 *
 * No counterpart in the reference: the reference computes every point of a
 * sigma(B) curve with the full solver (SimulationRunner.cpp:129-145).  This
 * is a small MLP, 6 -> 128 -> 128 -> 2 with tanh, over normalized
 * (log B, T, log density, radius, tau, alpha) -> two normalized sigma targets,
 * kept as an ensemble of independently initialised members.  Weights are FP32
 * masters with BF16 copies; forward and backward run on MFMA (BF16 operands,
 * FP32 accumulation); Adam updates the masters.  Gradients are reduced in a
 * fixed order, so a run is reproducible bit for bit.
 *
 * Parameter layout of one member (ES_MLP_NPARAM floats, row-major, padded):
 *   W1[128][16]  (input columns 6..15 stay zero)      offset 0
 *   b1[128]                                           offset 2048
 *   W2[128][128] (W2[out][in])                        offset 2176
 *   b2[128]                                           offset 18560
 *   W3[16][128]  (output rows 2..15 stay zero)        offset 18688
 *   b3[16]       (entries 2..15 stay zero)            offset 20736
 * y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3.
 */
#ifndef ES_SURROGATE_H
#define ES_SURROGATE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_MLP_INPUTS 6
#define ES_MLP_OUTPUTS 2
#define ES_MLP_HIDDEN 128
#define ES_MLP_NPARAM 20752
#define ES_MLP_MAX_MEMBERS 16

typedef struct es_mlp es_mlp;

typedef struct es_mlp_desc {
    int32_t members;      /* ensemble size, 1..ES_MLP_MAX_MEMBERS */
    int32_t device;
    uint64_t seed;        /* Glorot-uniform initialisation, member m from seed + m */
} es_mlp_desc;

typedef struct es_mlp_train_desc {
    int32_t epochs;
    int32_t batch;        /* rows per optimizer step, 1..65536 */
    float lr;             /* Adam step size, cosine-annealed from lr to lr_final */
    float lr_final;
    float beta1, beta2, eps;
    float out_weight[2];  /* loss = mean over rows of sum_o out_weight[o] * (y_o - t_o)^2 */
    uint64_t shuffle_seed;
} es_mlp_train_desc;

int es_mlp_create(const es_mlp_desc* desc, es_mlp** out);
void es_mlp_destroy(es_mlp* m);
const char* es_mlp_last_error(const es_mlp* m);
int es_mlp_members(const es_mlp* m);

/* FP32 master parameters of every member: members * ES_MLP_NPARAM floats.
 * Setting them also clears the Adam moments and step count. */
int es_mlp_get_params(es_mlp* m, float* params);
int es_mlp_set_params(es_mlp* m, const float* params);

/* Training rows (normalized): x [n][6], t [n][2]; copied to the device. */
int es_mlp_set_data(es_mlp* m, const float* x, const float* t, int64_t n);

/* Loss and gradient of every member over the given training rows, without an
 * update: grads members * ES_MLP_NPARAM, loss members floats. */
int es_mlp_gradients(es_mlp* m, const int32_t* rows, int32_t nrows, const float* out_weight, float* grads,
                     float* loss);

/* Adam training over the rows set by es_mlp_set_data; loss_history (or NULL)
 * receives epochs * members mean losses. */
int es_mlp_train(es_mlp* m, const es_mlp_train_desc* desc, float* loss_history);

/* Forward pass of every member: x [n][6] -> y [members][n][2]. */
int es_mlp_predict(es_mlp* m, const float* x, int64_t n, float* y);

/* Device time in seconds of the last es_mlp_train (fwd/bwd + update kernels)
 * and the number of fwd/bwd launches it made. */
int es_mlp_last_train_time(es_mlp* m, double* seconds, int64_t* steps);

#ifdef __cplusplus
}
#endif

#endif
