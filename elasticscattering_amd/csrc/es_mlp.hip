// es_mlp.hip — the sigma(B) surrogate: an ensemble of 6 -> 128 -> 128 -> 2
// tanh MLPs trained and evaluated with hand-written MFMA kernels (gfx950).
//
// One workgroup (4 wave64) takes 64 rows of one ensemble member
// (blockIdx.y).  Every product is v_mfma_f32_16x16x16_bf16 with BF16
// operands and FP32 accumulation; the operand fragments of a lane are
//   A[i][k]: lane = i + 16 * (k / 4), element k % 4
//   B[k][j]: lane = j + 16 * (k / 4), element k % 4
//   D[i][j]: lane = j + 16 * (i / 4), element i % 4
// so a lane reads 4 consecutive k of one row of A and of one column of B:
// every operand is staged in LDS (or kept in global memory) in the layout
// that makes those 4 values contiguous — activations both row-major and
// transposed, W2 as [out][in] for the forward pass and [in][out] for the
// backward one.  Activations stay in FP32 registers for the tanh
// derivative.  The weights are BF16; the forward pass splits each
// activation (and input) into a BF16 pair hi + lo and issues two MFMAs, so
// the network is a smooth function of its inputs to ~2^-16 relative — with
// one BF16 rounding per activation its output would step by ~1e-2 as the
// field B moves along a sweep.  Weight gradients use the hi parts.
// Gradients are written per workgroup and summed in a fixed order by the
// Adam kernel, so training is reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/es_abi.h"
#include "../../include/es_surrogate.h"
#include "es_rng.h"

namespace {

constexpr int kIn = ES_MLP_INPUTS;
constexpr int kInP = 16;   // inputs padded to one MFMA k-step
constexpr int kH = ES_MLP_HIDDEN;
constexpr int kOutP = 16;  // outputs padded to one MFMA tile
constexpr int kRows = 64;  // rows per workgroup
constexpr int kThreads = 256;

constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kH * kInP;
constexpr int OFF_W2 = OFF_B1 + kH;
constexpr int OFF_B2 = OFF_W2 + kH * kH;
constexpr int OFF_W3 = OFF_B2 + kH;
constexpr int OFF_B3 = OFF_W3 + kOutP * kH;
constexpr int NPARAM = OFF_B3 + kOutP;
static_assert(NPARAM == ES_MLP_NPARAM, "parameter layout");

// BF16 weight copies of one member.
constexpr int BW1 = 0;                   // W1  [128][16]
constexpr int BW2 = BW1 + kH * kInP;     // W2  [out][in]
constexpr int BW2T = BW2 + kH * kH;      // W2T [in][out]
constexpr int BW3 = BW2T + kH * kH;      // W3  [16][128]
constexpr int BW3T = BW3 + kOutP * kH;   // W3T [128][16]
constexpr int NBF = BW3T + kH * kOutP;

// LDS row strides in BF16 elements, padded by 4 so that the 16 rows one
// fragment reads at a fixed column spread over the banks.
constexpr int S16 = kInP + 4;
constexpr int S64 = kRows + 4;
constexpr int S128 = kH + 4;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Round to nearest even (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline v4f mfma16(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline v4s ld4(const uint16_t* p) { return *reinterpret_cast<const v4s*>(p); }

__device__ inline float from_bf16(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// v = hi + lo with hi = bf16(v), lo = bf16(v - hi) (v - hi is exact in FP32).
__device__ inline void split_bf16(float v, uint16_t& hi, uint16_t& lo)
{
    hi = to_bf16(v);
    lo = to_bf16(v - from_bf16(hi));
}

struct __align__(16) MlpLds {
    uint16_t xb[kRows * S16];     // x    [row][in] (hi)
    uint16_t xl[kRows * S16];     // x    [row][in] (lo)
    uint16_t xT[kInP * S64];      // x^T  [in][row] (hi)
    uint16_t h1b[kRows * S128];   // h1   [row][unit] (hi)
    uint16_t h1l[kRows * S128];   // h1   [row][unit] (lo)
    uint16_t h1T[kH * S64];       // h1^T [unit][row] (hi)
    uint16_t h2b[kRows * S128];   // h2 [row][unit] (hi); then dz2 (each wave its own rows)
    uint16_t h2l[kRows * S128];   // h2 [row][unit] (lo)
    uint16_t h2T[kH * S64];       // h2^T (hi); then dz1^T
    uint16_t dyb[kRows * S16];    // dy   [row][out]
    uint16_t dyT[kOutP * S64];    // dy^T [out][row]
    uint16_t dz2T[kH * S64];      // dz2^T [unit][row]
    float tgt[kRows][2];
    int valid[kRows];
    float red1[4][kH], red2[4][kH], red3[4][kOutP];
    float lred[4];
};

struct MlpArgs {
    const float* x;          // [*][6] normalized inputs
    const float* t;          // [*][2] normalized targets (training)
    const int32_t* rows;     // row ids of this launch, or null: rows 0..nrows-1
    int nrows;
    const uint16_t* wb;      // [member][NBF]
    const float* params;     // [member][NPARAM] (biases are read in FP32)
    float ow0, ow1, inv_n;   // loss weights, 1 / rows
    float* grad_part;        // [member][gridDim.x][NPARAM]
    float* loss_part;        // [member][gridDim.x]
    float* y_out;            // [member][nrows][2] (forward only)
};

__device__ inline float wave_sum(float v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_mlp(MlpArgs a)
{
    __shared__ MlpLds s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int mem = blockIdx.y;
    const uint16_t* wb = a.wb + (size_t)mem * NBF;
    const float* prm = a.params + (size_t)mem * NPARAM;
    const int row0 = blockIdx.x * kRows;
    const int rw = 16 * w;  // this wave's 16 rows

    for (int e = tid; e < kRows * kInP; e += kThreads) {
        const int i = e / kInP, k = e % kInP;
        const int r = row0 + i;
        const bool ok = r < a.nrows;
        const int src = ok ? (a.rows ? a.rows[r] : r) : 0;
        uint16_t hi = 0, lo = 0;
        if (ok && k < kIn) split_bf16(a.x[(size_t)src * kIn + k], hi, lo);
        s.xb[i * S16 + k] = hi;
        s.xl[i * S16 + k] = lo;
        s.xT[k * S64 + i] = hi;
        if (k == 0) {
            s.valid[i] = ok ? 1 : 0;
            if (TRAIN) {
                s.tgt[i][0] = ok ? a.t[(size_t)src * 2] : 0.0f;
                s.tgt[i][1] = ok ? a.t[(size_t)src * 2 + 1] : 0.0f;
            }
        }
    }
    __syncthreads();

    // Layer 1: z1 = x W1^T + b1 (K = 16, one step per 16-unit tile).  The
    // weight fragments and biases of all 8 tiles are loaded first: one round
    // trip to the cache, not one per tile.
    float h1f[8][4];
    {
        v4s bw[8];
        float bb[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            bw[t] = ld4(wb + BW1 + (16 * t + lr) * kInP + 4 * lg);
            bb[t] = prm[OFF_B1 + 16 * t + lr];
        }
        const v4s av = ld4(&s.xb[(rw + lr) * S16 + 4 * lg]);
        const v4s al = ld4(&s.xl[(rw + lr) * S16 + 4 * lg]);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const v4s bv = bw[t];
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma16(av, bv, acc);
            acc = mfma16(al, bv, acc);
            const int col = 16 * t + lr;
            const float bias = bb[t];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = rw + 4 * lg + r;
                const float h = tanhf(acc[r] + bias);
                h1f[t][r] = h;
                uint16_t hb, hl;
                split_bf16(h, hb, hl);
                s.h1b[row * S128 + col] = hb;
                s.h1l[row * S128 + col] = hl;
                s.h1T[col * S64 + row] = hb;
            }
        }
    }
    __syncthreads();

    // Layer 2: z2 = h1 W2^T + b2 (K = 128); the next k-step's 8 weight
    // fragments are loaded while this one's MFMAs run.
    float h2f[8][4];
    {
        v4f acc[8];
        float bb[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            bb[t] = prm[OFF_B2 + 16 * t + lr];
        }
        v4s bw[8];
#pragma unroll
        for (int t = 0; t < 8; t++) bw[t] = ld4(wb + BW2 + (16 * t + lr) * kH + 4 * lg);
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
            v4s nx[8];
#pragma unroll
            for (int t = 0; t < 8; t++) nx[t] = ld4(wb + BW2 + (16 * t + lr) * kH + 16 * ((ks + 1) & 7) + 4 * lg);
            const v4s av = ld4(&s.h1b[(rw + lr) * S128 + 16 * ks + 4 * lg]);
            const v4s al = ld4(&s.h1l[(rw + lr) * S128 + 16 * ks + 4 * lg]);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                acc[t] = mfma16(av, bw[t], acc[t]);
                acc[t] = mfma16(al, bw[t], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 8; t++) bw[t] = nx[t];
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int col = 16 * t + lr;
            const float bias = bb[t];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = rw + 4 * lg + r;
                const float h = tanhf(acc[t][r] + bias);
                h2f[t][r] = h;
                uint16_t hb, hl;
                split_bf16(h, hb, hl);
                s.h2b[row * S128 + col] = hb;
                s.h2l[row * S128 + col] = hl;
                s.h2T[col * S64 + row] = hb;
            }
        }
    }
    __syncthreads();

    // Layer 3: y = h2 W3^T + b3 (K = 128, one 16-wide output tile).
    v4f y = {0.0f, 0.0f, 0.0f, 0.0f};
    {
        v4s bw[8];
#pragma unroll
        for (int ks = 0; ks < 8; ks++) bw[ks] = ld4(wb + BW3 + lr * kH + 16 * ks + 4 * lg);
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
            const v4s av = ld4(&s.h2b[(rw + lr) * S128 + 16 * ks + 4 * lg]);
            const v4s al = ld4(&s.h2l[(rw + lr) * S128 + 16 * ks + 4 * lg]);
            y = mfma16(av, bw[ks], y);
            y = mfma16(al, bw[ks], y);
        }
    }
    const int o = lr;
    const float b3 = prm[OFF_B3 + o];
    if (!TRAIN) {
        if (o < 2) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = row0 + rw + 4 * lg + r;
                if (row < a.nrows) a.y_out[((size_t)mem * a.nrows + row) * 2 + o] = y[r] + b3;
            }
        }
        return;
    }

    // Loss: mean over rows of sum_o ow_o (y_o - t_o)^2; dy = 2 ow_o (y_o - t_o) / rows.
    float lsum = 0.0f, db3 = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = rw + 4 * lg + r;
        float d = 0.0f;
        if (s.valid[row] && o < 2) {
            const float ow = o ? a.ow1 : a.ow0;
            const float e = (y[r] + b3) - s.tgt[row][o];
            d = 2.0f * ow * e * a.inv_n;
            lsum += ow * e * e * a.inv_n;
        }
        db3 += d;
        const uint16_t dbf = to_bf16(d);
        s.dyb[row * S16 + o] = dbf;
        s.dyT[o * S64 + row] = dbf;
    }
    lsum = wave_sum(lsum);
    if (lane == 0) s.lred[w] = lsum;
    db3 += __shfl_xor(db3, 16, 64);
    db3 += __shfl_xor(db3, 32, 64);
    if (lg == 0) s.red3[w][lr] = db3;
    __syncthreads();

    float* g = a.grad_part + ((size_t)mem * gridDim.x + blockIdx.x) * NPARAM;

    // dW3 = dy^T h2: [16 x 64] x [64 x 128]; this wave's column tiles 2w, 2w+1.
#pragma unroll
    for (int tt = 0; tt < 2; tt++) {
        const int t = 2 * w + tt;
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const v4s av = ld4(&s.dyT[lr * S64 + 16 * ks + 4 * lg]);
            const v4s bv = ld4(&s.h2T[(16 * t + lr) * S64 + 16 * ks + 4 * lg]);
            acc = mfma16(av, bv, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) g[OFF_W3 + (4 * lg + r) * kH + 16 * t + lr] = acc[r];
    }

    // dz2 = (dy W3) * (1 - h2^2): [64 x 16] x [16 x 128].
    {
        v4s bw[8];
#pragma unroll
        for (int t = 0; t < 8; t++) bw[t] = ld4(wb + BW3T + (16 * t + lr) * kOutP + 4 * lg);
        const v4s av = ld4(&s.dyb[(rw + lr) * S16 + 4 * lg]);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma16(av, bw[t], acc);
            const int col = 16 * t + lr;
            float sum = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = rw + 4 * lg + r;
                const float dz = acc[r] * (1.0f - h2f[t][r] * h2f[t][r]);
                sum += dz;
                const uint16_t zb = to_bf16(dz);
                s.h2b[row * S128 + col] = zb;  // this wave's rows only: layer 3 read them before
                s.dz2T[col * S64 + row] = zb;
            }
            sum += __shfl_xor(sum, 16, 64);
            sum += __shfl_xor(sum, 32, 64);
            if (lg == 0) s.red2[w][col] = sum;
        }
    }
    __syncthreads();

    // dW2 = dz2^T h1: [128 x 64] x [64 x 128]; this wave's row tiles 2w, 2w+1.
#pragma unroll
    for (int aa = 0; aa < 2; aa++) {
        const int mt = 2 * w + aa;
        v4f acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const v4s av = ld4(&s.dz2T[(16 * mt + lr) * S64 + 16 * ks + 4 * lg]);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const v4s bv = ld4(&s.h1T[(16 * t + lr) * S64 + 16 * ks + 4 * lg]);
                acc[t] = mfma16(av, bv, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) g[OFF_W2 + (16 * mt + 4 * lg + r) * kH + 16 * t + lr] = acc[t][r];
    }

    // dz1 = (dz2 W2) * (1 - h1^2): [64 x 128] x [128 x 128], W2^T from global
    // memory, the next k-step's fragments loaded while this one's MFMAs run.
    {
        v4f acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        v4s bw[8];
#pragma unroll
        for (int t = 0; t < 8; t++) bw[t] = ld4(wb + BW2T + (16 * t + lr) * kH + 4 * lg);
#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
            v4s nx[8];
#pragma unroll
            for (int t = 0; t < 8; t++) nx[t] = ld4(wb + BW2T + (16 * t + lr) * kH + 16 * ((ks + 1) & 7) + 4 * lg);
            const v4s av = ld4(&s.h2b[(rw + lr) * S128 + 16 * ks + 4 * lg]);
#pragma unroll
            for (int t = 0; t < 8; t++) acc[t] = mfma16(av, bw[t], acc[t]);
#pragma unroll
            for (int t = 0; t < 8; t++) bw[t] = nx[t];
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int col = 16 * t + lr;
            float sum = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = rw + 4 * lg + r;
                const float dz = acc[t][r] * (1.0f - h1f[t][r] * h1f[t][r]);
                sum += dz;
                s.h2T[col * S64 + row] = to_bf16(dz);  // h2^T was last read by dW3, before the barrier
            }
            sum += __shfl_xor(sum, 16, 64);
            sum += __shfl_xor(sum, 32, 64);
            if (lg == 0) s.red1[w][col] = sum;
        }
    }
    __syncthreads();

    // dW1 = dz1^T x: [128 x 64] x [64 x 16]; this wave's row tiles 2w, 2w+1.
#pragma unroll
    for (int aa = 0; aa < 2; aa++) {
        const int mt = 2 * w + aa;
        v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const v4s av = ld4(&s.h2T[(16 * mt + lr) * S64 + 16 * ks + 4 * lg]);
            const v4s bv = ld4(&s.xT[lr * S64 + 16 * ks + 4 * lg]);
            acc = mfma16(av, bv, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) g[OFF_W1 + (16 * mt + 4 * lg + r) * kInP + lr] = acc[r];
    }

    if (tid < kH) {
        g[OFF_B1 + tid] = ((s.red1[0][tid] + s.red1[1][tid]) + s.red1[2][tid]) + s.red1[3][tid];
        g[OFF_B2 + tid] = ((s.red2[0][tid] + s.red2[1][tid]) + s.red2[2][tid]) + s.red2[3][tid];
    }
    if (tid < kOutP) g[OFF_B3 + tid] = ((s.red3[0][tid] + s.red3[1][tid]) + s.red3[2][tid]) + s.red3[3][tid];
    if (tid == 0) a.loss_part[(size_t)mem * gridDim.x + blockIdx.x] = ((s.lred[0] + s.lred[1]) + s.lred[2]) + s.lred[3];
}

// BF16 copies of parameter i of one member.
__device__ inline void put_bf16(uint16_t* wb, int i, float v)
{
    const uint16_t b = to_bf16(v);
    if (i < OFF_B1) {
        wb[BW1 + i] = b;
    } else if (i >= OFF_W2 && i < OFF_B2) {
        const int j = i - OFF_W2, out = j / kH, in = j % kH;
        wb[BW2 + out * kH + in] = b;
        wb[BW2T + in * kH + out] = b;
    } else if (i >= OFF_W3 && i < OFF_B3) {
        const int j = i - OFF_W3, oo = j / kH, n = j % kH;
        wb[BW3 + oo * kH + n] = b;
        wb[BW3T + n * kOutP + oo] = b;
    }
}

__global__ void k_mlp_sync_bf16(const float* p, uint16_t* wb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int mem = blockIdx.y;
    if (i < NPARAM) put_bf16(wb + (size_t)mem * NBF, i, p[(size_t)mem * NPARAM + i]);
}

struct AdamArgs {
    float* p;
    float* m;
    float* v;
    uint16_t* wb;
    const float* grad_part;   // [member][nblk][NPARAM]
    const float* loss_part;   // [member][nblk]
    float* loss_acc;          // [member]
    int nblk;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
};

// Sums the workgroup gradients in block order, then one Adam step (Kingma & Ba 2015).
__global__ void k_mlp_adam(AdamArgs a)
{
    const int mem = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        float l = 0.0f;
        for (int b = 0; b < a.nblk; b++) l += a.loss_part[(size_t)mem * a.nblk + b];
        a.loss_acc[mem] += l;
    }
    if (i >= NPARAM) return;
    float gs = 0.0f;
    for (int b = 0; b < a.nblk; b++) gs += a.grad_part[((size_t)mem * a.nblk + b) * NPARAM + i];
    const size_t k = (size_t)mem * NPARAM + i;
    const float mm = a.b1 * a.m[k] + (1.0f - a.b1) * gs;
    const float vv = a.b2 * a.v[k] + (1.0f - a.b2) * gs * gs;
    a.m[k] = mm;
    a.v[k] = vv;
    const float pv = a.p[k] - a.lr * (mm * a.c1) / (sqrtf(vv * a.c2) + a.eps);
    a.p[k] = pv;
    put_bf16(a.wb + (size_t)mem * NBF, i, pv);
}

__global__ void k_fill0(float* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

thread_local std::string g_mlp_create_error;

}  // namespace

struct es_mlp {
    int device = 0;
    int members = 1;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string err;
    float* p = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wb = nullptr;
    float* x = nullptr;
    float* t = nullptr;
    int64_t n = 0;
    int32_t* rows = nullptr;
    size_t rows_cap = 0;
    float* gpart = nullptr;
    size_t gpart_cap = 0;
    float* lpart = nullptr;
    size_t lpart_cap = 0;
    float* lacc = nullptr;
    float* xin = nullptr;
    size_t xin_cap = 0;
    float* yout = nullptr;
    size_t yout_cap = 0;
    int64_t step = 0;
    double last_train_s = 0.0;
    int64_t last_steps = 0;
};

namespace {

int mfail(es_mlp* m, int code, const std::string& msg)
{
    m->err = msg;
    return code;
}

#define MLP_HIP(m, call)                                                                              \
    do {                                                                                              \
        hipError_t _e = (call);                                                                       \
        if (_e != hipSuccess) return mfail(m, ES_ERR_HIP, std::string(#call ": ") + hipGetErrorString(_e)); \
    } while (0)

template <typename T>
int grow(es_mlp* m, T*& ptr, size_t& cap, size_t count, const char* what)
{
    if (count <= cap && ptr) return ES_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    const hipError_t e = hipMalloc((void**)&ptr, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return mfail(m, ES_ERR_OOM, std::string("hipMalloc failed for ") + what);
    }
    cap = std::max<size_t>(count, 1);
    return ES_OK;
}

int upload_params(es_mlp* m, const float* params)
{
    const size_t np = (size_t)m->members * NPARAM;
    MLP_HIP(m, hipMemcpyAsync(m->p, params, np * sizeof(float), hipMemcpyHostToDevice, m->stream));
    const int g = (int)((np + 255) / 256);
    hipLaunchKernelGGL(k_fill0, dim3(g), dim3(256), 0, m->stream, m->m, (int)np);
    hipLaunchKernelGGL(k_fill0, dim3(g), dim3(256), 0, m->stream, m->v, (int)np);
    hipLaunchKernelGGL(k_mlp_sync_bf16, dim3((NPARAM + 255) / 256, m->members), dim3(256), 0, m->stream, m->p, m->wb);
    MLP_HIP(m, hipGetLastError());
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    m->step = 0;
    return ES_OK;
}

// Workgroup gradients of one batch of rows (already on the device).
int launch_grad(es_mlp* m, const int32_t* rows_dev, int nrows, const float* ow, int& nblk)
{
    nblk = (nrows + kRows - 1) / kRows;
    int rc;
    if ((rc = grow(m, m->gpart, m->gpart_cap, (size_t)m->members * nblk * NPARAM, "gradient partials"))) return rc;
    if ((rc = grow(m, m->lpart, m->lpart_cap, (size_t)m->members * nblk, "loss partials"))) return rc;
    MlpArgs a{};
    a.x = m->x;
    a.t = m->t;
    a.rows = rows_dev;
    a.nrows = nrows;
    a.wb = m->wb;
    a.params = m->p;
    a.ow0 = ow[0];
    a.ow1 = ow[1];
    a.inv_n = 1.0f / (float)nrows;
    a.grad_part = m->gpart;
    a.loss_part = m->lpart;
    hipLaunchKernelGGL(k_mlp<true>, dim3(nblk, m->members), dim3(kThreads), 0, m->stream, a);
    MLP_HIP(m, hipGetLastError());
    return ES_OK;
}

}  // namespace

extern "C" {

int es_mlp_create(const es_mlp_desc* d, es_mlp** out)
{
    if (!d || !out) {
        g_mlp_create_error = "null argument";
        return ES_ERR_INVALID_ARG;
    }
    *out = nullptr;
    if (d->members < 1 || d->members > ES_MLP_MAX_MEMBERS) {
        g_mlp_create_error = "members must be 1..16";
        return ES_ERR_INVALID_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        g_mlp_create_error = "no HIP device";
        return ES_ERR_NO_DEVICE;
    }
    if (d->device < 0 || d->device >= ndev) {
        g_mlp_create_error = "device ordinal out of range";
        return ES_ERR_INVALID_ARG;
    }
    es_mlp* m = new es_mlp();
    m->device = d->device;
    m->members = d->members;
    const size_t np = (size_t)m->members * NPARAM;
    hipError_t e = hipSetDevice(m->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&m->e0);
    if (e == hipSuccess) e = hipEventCreate(&m->e1);
    if (e == hipSuccess) e = hipMalloc((void**)&m->p, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->m, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->v, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc((void**)&m->wb, (size_t)m->members * NBF * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc((void**)&m->lacc, ES_MLP_MAX_MEMBERS * sizeof(float));
    if (e != hipSuccess) {
        g_mlp_create_error = std::string("HIP setup failed: ") + hipGetErrorString(e);
        es_mlp_destroy(m);
        return ES_ERR_HIP;
    }
    // Glorot-uniform weights (bound sqrt(6 / (fan_in + fan_out))), zero biases and padding.
    std::vector<float> p(np, 0.0f);
    const double g1 = std::sqrt(6.0 / (kIn + kH)), g2 = std::sqrt(6.0 / (kH + kH)), g3 = std::sqrt(6.0 / (kH + 2));
    for (int mem = 0; mem < m->members; mem++) {
        const uint64_t key = esr_key(d->seed + (uint64_t)mem);
        float* q = p.data() + (size_t)mem * NPARAM;
        for (int i = 0; i < NPARAM; i++) {
            const double u = 2.0 * esr_uniform(key, (uint64_t)i) - 1.0;
            if (i < OFF_B1) {
                if (i % kInP < kIn) q[i] = (float)(u * g1);
            } else if (i >= OFF_W2 && i < OFF_B2) {
                q[i] = (float)(u * g2);
            } else if (i >= OFF_W3 && i < OFF_B3) {
                if ((i - OFF_W3) / kH < 2) q[i] = (float)(u * g3);
            }
        }
    }
    const int rc = upload_params(m, p.data());
    if (rc != ES_OK) {
        g_mlp_create_error = m->err;
        es_mlp_destroy(m);
        return rc;
    }
    *out = m;
    return ES_OK;
}

void es_mlp_destroy(es_mlp* m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* bufs[] = {m->p, m->m, m->v, m->wb, m->x, m->t, m->rows, m->gpart, m->lpart, m->lacc, m->xin, m->yout};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (m->e0) (void)hipEventDestroy(m->e0);
    if (m->e1) (void)hipEventDestroy(m->e1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* es_mlp_last_error(const es_mlp* m) { return m ? m->err.c_str() : g_mlp_create_error.c_str(); }

int es_mlp_members(const es_mlp* m) { return m ? m->members : 0; }

int es_mlp_get_params(es_mlp* m, float* params)
{
    if (!m || !params) return ES_ERR_INVALID_ARG;
    MLP_HIP(m, hipSetDevice(m->device));
    MLP_HIP(m, hipMemcpyAsync(params, m->p, (size_t)m->members * NPARAM * sizeof(float), hipMemcpyDeviceToHost,
                              m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return ES_OK;
}

int es_mlp_set_params(es_mlp* m, const float* params)
{
    if (!m || !params) return ES_ERR_INVALID_ARG;
    MLP_HIP(m, hipSetDevice(m->device));
    return upload_params(m, params);
}

int es_mlp_set_data(es_mlp* m, const float* x, const float* t, int64_t n)
{
    if (!m) return ES_ERR_INVALID_ARG;
    if (!x || !t || n <= 0 || n > (1ll << 30)) return mfail(m, ES_ERR_INVALID_ARG, "need 1..2^30 rows of data");
    MLP_HIP(m, hipSetDevice(m->device));
    if (m->x) (void)hipFree(m->x);
    if (m->t) (void)hipFree(m->t);
    m->x = m->t = nullptr;
    m->n = 0;
    MLP_HIP(m, hipMalloc((void**)&m->x, (size_t)n * kIn * sizeof(float)));
    MLP_HIP(m, hipMalloc((void**)&m->t, (size_t)n * 2 * sizeof(float)));
    MLP_HIP(m, hipMemcpy(m->x, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice));
    MLP_HIP(m, hipMemcpy(m->t, t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice));
    m->n = n;
    return ES_OK;
}

int es_mlp_gradients(es_mlp* m, const int32_t* rows, int32_t nrows, const float* out_weight, float* grads,
                     float* loss)
{
    if (!m) return ES_ERR_INVALID_ARG;
    if (!rows || nrows <= 0 || !out_weight || !grads) return mfail(m, ES_ERR_INVALID_ARG, "bad arguments");
    if (!m->x) return mfail(m, ES_ERR_INVALID_ARG, "no data: call es_mlp_set_data");
    for (int i = 0; i < nrows; i++)
        if (rows[i] < 0 || rows[i] >= m->n) return mfail(m, ES_ERR_INVALID_ARG, "row id out of range");
    MLP_HIP(m, hipSetDevice(m->device));
    int rc;
    if ((rc = grow(m, m->rows, m->rows_cap, (size_t)nrows, "row ids"))) return rc;
    MLP_HIP(m, hipMemcpyAsync(m->rows, rows, (size_t)nrows * sizeof(int32_t), hipMemcpyHostToDevice, m->stream));
    int nblk = 0;
    if ((rc = launch_grad(m, m->rows, nrows, out_weight, nblk))) return rc;
    std::vector<float> gp((size_t)m->members * nblk * NPARAM), lp((size_t)m->members * nblk);
    MLP_HIP(m, hipMemcpyAsync(gp.data(), m->gpart, gp.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipMemcpyAsync(lp.data(), m->lpart, lp.size() * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    for (int mem = 0; mem < m->members; mem++) {
        for (int i = 0; i < NPARAM; i++) {
            float s = 0.0f;
            for (int b = 0; b < nblk; b++) s += gp[((size_t)mem * nblk + b) * NPARAM + i];
            grads[(size_t)mem * NPARAM + i] = s;
        }
        if (loss) {
            float l = 0.0f;
            for (int b = 0; b < nblk; b++) l += lp[(size_t)mem * nblk + b];
            loss[mem] = l;
        }
    }
    return ES_OK;
}

int es_mlp_train(es_mlp* m, const es_mlp_train_desc* d, float* loss_history)
{
    if (!m) return ES_ERR_INVALID_ARG;
    if (!d || d->epochs < 1 || d->batch < 1 || d->batch > 65536 || !(d->lr > 0) || !(d->lr_final >= 0) ||
        !(d->beta1 >= 0 && d->beta1 < 1) || !(d->beta2 >= 0 && d->beta2 < 1) || !(d->eps > 0))
        return mfail(m, ES_ERR_INVALID_ARG, "bad training settings");
    if (!m->x) return mfail(m, ES_ERR_INVALID_ARG, "no data: call es_mlp_set_data");
    MLP_HIP(m, hipSetDevice(m->device));
    const int64_t n = m->n;
    const int batch = (int)std::min<int64_t>(d->batch, n);
    const int64_t steps_per_epoch = (n + batch - 1) / batch;
    const int64_t total = (int64_t)d->epochs * steps_per_epoch;
    int rc;
    if ((rc = grow(m, m->rows, m->rows_cap, (size_t)n, "row ids"))) return rc;
    const int maxblk = (batch + kRows - 1) / kRows;
    if ((rc = grow(m, m->gpart, m->gpart_cap, (size_t)m->members * maxblk * NPARAM, "gradient partials"))) return rc;
    if ((rc = grow(m, m->lpart, m->lpart_cap, (size_t)m->members * maxblk, "loss partials"))) return rc;
    std::vector<int32_t> perm((size_t)n);
    std::vector<float> lh(m->members);
    int64_t k = 0;
    MLP_HIP(m, hipEventRecord(m->e0, m->stream));
    for (int ep = 0; ep < d->epochs; ep++) {
        // Fisher-Yates shuffle from the es_rng.h stream of (shuffle_seed + epoch).
        for (int64_t i = 0; i < n; i++) perm[(size_t)i] = (int32_t)i;
        const uint64_t key = esr_key(d->shuffle_seed + (uint64_t)ep);
        for (int64_t i = n - 1; i > 0; i--) {
            const int64_t j = (int64_t)(esr_uniform(key, (uint64_t)i) * (double)(i + 1));
            std::swap(perm[(size_t)i], perm[(size_t)std::min(j, i)]);
        }
        MLP_HIP(m, hipMemcpyAsync(m->rows, perm.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice,
                                  m->stream));
        hipLaunchKernelGGL(k_fill0, dim3(1), dim3(64), 0, m->stream, m->lacc, m->members);
        for (int64_t st = 0; st < steps_per_epoch; st++, k++) {
            const int64_t start = st * batch;
            const int nr = (int)std::min<int64_t>(batch, n - start);
            int nblk = 0;
            if ((rc = launch_grad(m, m->rows + start, nr, d->out_weight, nblk))) return rc;
            m->step++;
            AdamArgs a{};
            a.p = m->p;
            a.m = m->m;
            a.v = m->v;
            a.wb = m->wb;
            a.grad_part = m->gpart;
            a.loss_part = m->lpart;
            a.loss_acc = m->lacc;
            a.nblk = nblk;
            const double frac = total > 1 ? (double)k / (double)(total - 1) : 0.0;
            a.lr = (float)(d->lr_final + 0.5 * (d->lr - d->lr_final) * (1.0 + std::cos(3.14159265358979323846 * frac)));
            a.b1 = d->beta1;
            a.b2 = d->beta2;
            a.eps = d->eps;
            a.c1 = (float)(1.0 / (1.0 - std::pow((double)d->beta1, (double)m->step)));
            a.c2 = (float)(1.0 / (1.0 - std::pow((double)d->beta2, (double)m->step)));
            hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256, m->members), dim3(256), 0, m->stream, a);
            MLP_HIP(m, hipGetLastError());
        }
        MLP_HIP(m, hipMemcpyAsync(lh.data(), m->lacc, m->members * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MLP_HIP(m, hipStreamSynchronize(m->stream));  // perm and lh are reused next epoch
        if (loss_history)
            for (int mem = 0; mem < m->members; mem++)
                loss_history[(size_t)ep * m->members + mem] = lh[mem] / (float)steps_per_epoch;
    }
    MLP_HIP(m, hipEventRecord(m->e1, m->stream));
    MLP_HIP(m, hipEventSynchronize(m->e1));
    float ms = 0.0f;
    MLP_HIP(m, hipEventElapsedTime(&ms, m->e0, m->e1));
    m->last_train_s = ms * 1e-3;
    m->last_steps = total;
    return ES_OK;
}

int es_mlp_predict(es_mlp* m, const float* x, int64_t n, float* y)
{
    if (!m) return ES_ERR_INVALID_ARG;
    if (!x || !y || n <= 0 || n > (1ll << 30)) return mfail(m, ES_ERR_INVALID_ARG, "need 1..2^30 rows");
    MLP_HIP(m, hipSetDevice(m->device));
    int rc;
    if ((rc = grow(m, m->xin, m->xin_cap, (size_t)n * kIn, "inputs"))) return rc;
    if ((rc = grow(m, m->yout, m->yout_cap, (size_t)m->members * n * 2, "outputs"))) return rc;
    MLP_HIP(m, hipMemcpyAsync(m->xin, x, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MlpArgs a{};
    a.x = m->xin;
    a.nrows = (int)n;
    a.wb = m->wb;
    a.params = m->p;
    a.y_out = m->yout;
    hipLaunchKernelGGL(k_mlp<false>, dim3((unsigned)((n + kRows - 1) / kRows), m->members), dim3(kThreads), 0,
                       m->stream, a);
    MLP_HIP(m, hipGetLastError());
    MLP_HIP(m, hipMemcpyAsync(y, m->yout, (size_t)m->members * n * 2 * sizeof(float), hipMemcpyDeviceToHost,
                              m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return ES_OK;
}

int es_mlp_last_train_time(es_mlp* m, double* seconds, int64_t* steps)
{
    if (!m) return ES_ERR_INVALID_ARG;
    if (seconds) *seconds = m->last_train_s;
    if (steps) *steps = m->last_steps;
    return ES_OK;
}

}  // extern "C"
