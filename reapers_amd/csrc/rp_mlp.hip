// rp_mlp.hip -- forward, backward and AdamW kernels of the Krylov-schedule
// proposer network, a 12 -> 64 -> 64 -> 2 perceptron with tanh hidden
// layers, in FP32.
//
// Parameter layout (RP_MLP_NPARAMS = 5122 floats):
//   W1[64][12] @ 0, b1[64] @ 768, W2[64][64] @ 832, b2[64] @ 4928,
//   W3[2][64] @ 4992, b3[2] @ 5120.
//
// One 64-wide wavefront evaluates one sample: lane j owns hidden unit j of
// both layers, so the 64x64 layer is one LDS-broadcast dot product per lane
// (W2 is staged transposed so that the 64 lanes read 64 consecutive words)
// and the 2 outputs are wavefront reductions.  The gradient kernel keeps the
// per-lane slice of the gradient (row j of W1 and W2, b1[j], b2[j], column j
// of W3) in registers over the wave's samples, sums the 4 waves of a block in
// LDS and writes one partial row per block; a second kernel adds the block
// rows in a fixed order, so gradients are bitwise deterministic for a given
// sample count.
#include "rp_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>

namespace rp {
namespace {

constexpr int kIn = RP_MLP_IN, kHid = RP_MLP_HIDDEN, kOut = RP_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid,
              oW3 = oB2 + kHid, oB3 = oW3 + kOut * kHid, kNP = oB3 + kOut;
static_assert(kNP == RP_MLP_NPARAMS, "parameter layout");
static_assert(kHid == 64, "one lane per hidden unit of a 64-wide wavefront");

constexpr int kWaves = 4, kThreads = 64 * kWaves;
constexpr int kMaxGradBlocks = 32;

struct Shared {
    float w1t[kIn * kHid];     // w1t[i*64 + j] = W1[j][i]
    float w2t[kHid * kHid];    // w2t[k*64 + j] = W2[j][k]
    float w2[kHid * kHid];     // W2[j][k]
    float h1[kWaves][kHid];
    float dz2[kWaves][kHid];
    float x[kWaves][kIn];
};

__device__ inline float wave_sum(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline void stage_weights(Shared &sh, const float *__restrict__ p, bool need_w2)
{
    for (int e = threadIdx.x; e < kHid * kIn; e += kThreads) {
        int j = e / kIn, i = e % kIn;
        sh.w1t[i * kHid + j] = p[oW1 + e];
    }
    for (int e = threadIdx.x; e < kHid * kHid; e += kThreads) {
        int j = e / kHid, k = e % kHid;
        float w = p[oW2 + e];
        sh.w2t[k * kHid + j] = w;
        if (need_w2) sh.w2[e] = w;
    }
}

// forward of one sample by one wave; returns h1, h2 of this lane and y[2]
__device__ inline void wave_forward(Shared &sh, int wv, int lane, const float *__restrict__ p,
                                    const float *__restrict__ xs, float &h1, float &h2,
                                    float y[kOut])
{
    if (lane < kIn) sh.x[wv][lane] = xs[lane];
    __builtin_amdgcn_wave_barrier();
    float z = p[oB1 + lane];
    for (int i = 0; i < kIn; i++) z += sh.w1t[i * kHid + lane] * sh.x[wv][i];
    h1 = tanhf(z);
    sh.h1[wv][lane] = h1;
    __builtin_amdgcn_wave_barrier();
    z = p[oB2 + lane];
    for (int k = 0; k < kHid; k++) z += sh.w2t[k * kHid + lane] * sh.h1[wv][k];
    h2 = tanhf(z);
    for (int o = 0; o < kOut; o++) y[o] = wave_sum(p[oW3 + o * kHid + lane] * h2) + p[oB3 + o];
}

__global__ __launch_bounds__(kThreads) void k_mlp_forward(const float *__restrict__ p,
                                                          const float *__restrict__ X, long n,
                                                          float *__restrict__ Y)
{
    __shared__ Shared sh;
    stage_weights(sh, p, false);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (long s = (long)blockIdx.x * kWaves + wv; s < n; s += (long)gridDim.x * kWaves) {
        float h1, h2, y[kOut];
        wave_forward(sh, wv, lane, p, X + s * kIn, h1, h2, y);
        if (lane < kOut) Y[s * kOut + lane] = y[lane];
        __builtin_amdgcn_wave_barrier();
    }
}

// loss = sum_s sum_o w[s,o] (y[s,o] - t[s,o])^2 / n and its gradient
__global__ __launch_bounds__(kThreads) void k_mlp_grad(const float *__restrict__ p,
                                                       const float *__restrict__ X,
                                                       const float *__restrict__ T,
                                                       const float *__restrict__ W, long n,
                                                       float *__restrict__ part)
{
    __shared__ Shared sh;
    __shared__ float red[kWaves][kHid];
    stage_weights(sh, p, true);
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float inv_n = 1.0f / (float)n;
    float gw1[kIn], gw2[kHid], gb1 = 0, gb2 = 0, gw3[kOut], gb3 = 0, loss = 0;
    for (int i = 0; i < kIn; i++) gw1[i] = 0;
    for (int k = 0; k < kHid; k++) gw2[k] = 0;
    for (int o = 0; o < kOut; o++) gw3[o] = 0;
    float w3[kOut];
    for (int o = 0; o < kOut; o++) w3[o] = p[oW3 + o * kHid + lane];
    for (long s = (long)blockIdx.x * kWaves + wv; s < n; s += (long)gridDim.x * kWaves) {
        float h1, h2, y[kOut];
        wave_forward(sh, wv, lane, p, X + s * kIn, h1, h2, y);
        float dy[kOut], dh2 = 0;
        for (int o = 0; o < kOut; o++) {
            float w = W ? W[s * kOut + o] : 1.0f;
            float e = y[o] - T[s * kOut + o];
            loss += w * e * e;
            dy[o] = 2.0f * w * e * inv_n;
            gw3[o] += dy[o] * h2;
            dh2 += w3[o] * dy[o];
            if (lane == o) gb3 += dy[o];
        }
        float dz2 = dh2 * (1.0f - h2 * h2);
        gb2 += dz2;
        for (int k = 0; k < kHid; k++) gw2[k] += dz2 * sh.h1[wv][k];
        sh.dz2[wv][lane] = dz2;
        __builtin_amdgcn_wave_barrier();
        float dh1 = 0;
        for (int j = 0; j < kHid; j++) dh1 += sh.w2[j * kHid + lane] * sh.dz2[wv][j];
        float dz1 = dh1 * (1.0f - h1 * h1);
        gb1 += dz1;
        for (int i = 0; i < kIn; i++) gw1[i] += dz1 * sh.x[wv][i];
        __builtin_amdgcn_wave_barrier();
    }
    // sum the waves of the block: one parameter group at a time through LDS
    float *row = part + (size_t)blockIdx.x * (kNP + 1);
    auto block_sum = [&](float v, int dst_index) {
        red[wv][lane] = v;
        __syncthreads();
        if (wv == 0) {
            float a = red[0][lane];
            for (int q = 1; q < kWaves; q++) a += red[q][lane];
            row[dst_index] = a;
        }
        __syncthreads();
    };
    for (int i = 0; i < kIn; i++) block_sum(gw1[i], oW1 + lane * kIn + i);
    block_sum(gb1, oB1 + lane);
    for (int k = 0; k < kHid; k++) block_sum(gw2[k], oW2 + lane * kHid + k);
    block_sum(gb2, oB2 + lane);
    for (int o = 0; o < kOut; o++) block_sum(gw3[o], oW3 + o * kHid + lane);
    // b3 lives in lane o of each wave
    red[wv][lane] = gb3;
    __syncthreads();
    if (threadIdx.x < kOut) {
        float a = red[0][threadIdx.x];
        for (int q = 1; q < kWaves; q++) a += red[q][threadIdx.x];
        row[oB3 + threadIdx.x] = a;
    }
    __syncthreads();
    loss = __shfl(loss, 0, 64);   // every lane holds the same loss sum
    red[wv][lane] = loss;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a = red[0][0];
        for (int q = 1; q < kWaves; q++) a += red[q][0];
        row[kNP] = a;
    }
}

// out[e] = sum over blocks b (in order) of part[b][e]; the loss sums, at
// index kNP of each row, go to loss_out divided by n
__global__ void k_mlp_reduce(const float *__restrict__ part, int nb, float *__restrict__ out,
                             double *__restrict__ loss_out, long n)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e <= kNP; e += gridDim.x * blockDim.x) {
        float a = 0;
        for (int b = 0; b < nb; b++) a += part[(size_t)b * (kNP + 1) + e];
        if (e < kNP)
            out[e] = a;
        else
            loss_out[0] = (double)a / (double)n;
    }
}

__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ m, float *__restrict__ v, long np, float lr,
                           float b1, float b2, float eps, float wd, float c1, float c2)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < np;
         e += (long)gridDim.x * blockDim.x) {
        float ge = g[e];
        float me = b1 * m[e] + (1.0f - b1) * ge;
        float ve = b2 * v[e] + (1.0f - b2) * ge * ge;
        m[e] = me;
        v[e] = ve;
        float mh = me / c1, vh = ve / c2;
        p[e] = p[e] - lr * (mh / (sqrtf(vh) + eps) + wd * p[e]);
    }
}

int grad_blocks(long n) { return (int)std::min<long>(kMaxGradBlocks, (n + kWaves - 1) / kWaves); }

}  // namespace
}  // namespace rp

using namespace rp;

extern "C" {

int rp_mlp_forward(const float *params, const float *X, int64_t n, float *Y)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(n >= 0, "n must be non-negative");
    if (n == 0) return RP_OK;
    RP_REQUIRE(params && X && Y, "null argument");
    int g = (int)std::min<int64_t>(1024, (n + kWaves - 1) / kWaves);
    RP_LAUNCH(RP_KERNEL_OTHER, k_mlp_forward, dim3(g), dim3(kThreads), 0, params, X, (long)n, Y);
    return RP_OK;
}

int rp_mlp_loss_grad(const float *params, const float *X, const float *T, const float *W,
                     int64_t n, float *grad, double *loss)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(n > 0, "n must be positive");
    RP_REQUIRE(params && X && T && grad, "null argument");
    Context &c = ctx();
    const int nb = grad_blocks((long)n);
    float *part = (float *)c.work;
    double *dloss = (double *)((char *)c.work + c.work_bytes - 64 * 1024);
    static_assert(kMaxGradBlocks * (kNP + 1) * sizeof(float) + 64 * 1024 <=
                      (size_t)kRedBlocks * (2 * 64 + 1) * sizeof(double) + 64 * 1024,
                  "gradient partials fit the workspace");
    RP_LAUNCH(RP_KERNEL_OTHER, k_mlp_grad, dim3(nb), dim3(kThreads), 0, params, X, T, W,
              (long)n, part);
    RP_LAUNCH(RP_KERNEL_OTHER, k_mlp_reduce, dim3(21), dim3(256), 0, (const float *)part, nb,
              grad, dloss, (long)n);
    if (loss) {
        RP_CHECK_HIP(hipMemcpyAsync(c.pinned, dloss, sizeof(double), hipMemcpyDeviceToHost,
                                    c.stream));
        RP_CHECK_HIP(hipStreamSynchronize(c.stream));
        if (c.prof) {
            st = prof_collect();
            if (st) return st;
        }
        *loss = c.pinned[0];
    }
    return RP_OK;
}

int rp_mlp_adam(float *params, const float *grad, float *m, float *v, int64_t np, double lr,
                double beta1, double beta2, double eps, double weight_decay, int64_t step)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(np > 0 && step >= 1, "np and step must be positive");
    RP_REQUIRE(params && grad && m && v, "null argument");
    RP_REQUIRE(beta1 >= 0 && beta1 < 1 && beta2 >= 0 && beta2 < 1, "betas must be in [0, 1)");
    const double c1 = 1.0 - pow(beta1, (double)step), c2 = 1.0 - pow(beta2, (double)step);
    int g = (int)std::min<int64_t>(1024, (np + 255) / 256);
    RP_LAUNCH(RP_KERNEL_OTHER, k_mlp_adam, dim3(g), dim3(256), 0, params, grad, m, v, (long)np,
              (float)lr, (float)beta1, (float)beta2, (float)eps, (float)weight_decay, (float)c1,
              (float)c2);
    return RP_OK;
}

int rp_mlp_infer(const float *params, const float *x_host, int n, float *y_host)
{
    int st = check_init();
    if (st) return st;
    RP_REQUIRE(n >= 1 && n <= RP_MLP_MAX_INFER, "n must be in [1, RP_MLP_MAX_INFER]");
    RP_REQUIRE(params && x_host && y_host, "null argument");
    Context &c = ctx();
    // staging: inputs and outputs in the pinned buffer and in the tail of the
    // workspace (stream order keeps them apart from reduction partials)
    float *hin = (float *)c.pinned, *hout = hin + RP_MLP_MAX_INFER * kIn;
    float *din = (float *)((char *)c.work + c.work_bytes - 32 * 1024);
    float *dout = din + RP_MLP_MAX_INFER * kIn;
    // the previous user of the pinned buffer has completed: every entry point
    // that fills it synchronizes before returning
    memcpy(hin, x_host, sizeof(float) * n * kIn);
    RP_CHECK_HIP(hipMemcpyAsync(din, hin, sizeof(float) * n * kIn, hipMemcpyHostToDevice,
                                c.stream));
    RP_LAUNCH(RP_KERNEL_OTHER, k_mlp_forward, dim3((n + kWaves - 1) / kWaves), dim3(kThreads), 0,
              params, (const float *)din, (long)n, dout);
    RP_CHECK_HIP(hipMemcpyAsync(hout, dout, sizeof(float) * n * kOut, hipMemcpyDeviceToHost,
                                c.stream));
    RP_CHECK_HIP(hipStreamSynchronize(c.stream));
    if (c.prof) {
        st = prof_collect();
        if (st) return st;
    }
    memcpy(y_host, hout, sizeof(float) * n * kOut);
    return RP_OK;
}

}  // extern "C"
