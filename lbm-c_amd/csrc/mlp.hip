// Training and stand-alone evaluation of the learned relaxation-time
// closure (mlp.hpp), and the C ABI of its trainer (include/lbmc.h,
// lbmc_mlp_*).
//
// One training step over a minibatch of B samples:
//   1. mlp_train_kernel: one wave per 64 samples computes the forward pass,
//      the backward pass and its partial gradient (per-wave record);
//   2. mlp_update_kernel: one thread per parameter sums the records in wave
//      order (deterministic) and applies Adam to the FP32 master weights;
//   3. mlp_image_kernel: rebuilds the operand images (bf16 for training,
//      fp16 for evaluation) from the master weights.
// The loss is the mean squared error over the minibatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/lbmc.h"
#include "mlp.hpp"

namespace lbmc {

__device__ __forceinline__ __bf16 bf16_of(float x) { return __builtin_bit_cast(__bf16, bf16_bits(x)); }

// 16-bit bits of a weight in the evaluation image: fp16 (H) or bf16
template <bool H>
__device__ __forceinline__ unsigned short eval_bits(float x)
{
    if constexpr (H) return __builtin_bit_cast(unsigned short, (_Float16)x);
    else return bf16_bits(x);
}

// Block k < 12 of an evaluation image (32x32 products; fp16 operands when H)
template <bool H>
__device__ __forceinline__ void eval_image_block(int k, const float *p, MlpEval &ev)
{
    const int l = (int)threadIdx.x;
    const float *W1 = p + MLP_W1, *W2 = p + MLP_W2, *b2 = p + MLP_B2, *w3 = p + MLP_W3;
    if (k < 8) {
        const int m = k >> 2, s = k & 3, h = l >> 5, c = l & 31;
        unsigned short v[8];
        for (int j = 0; j < 8; j++) v[j] = eval_bits<H>(W2[(32 * m + c) * 64 + mlp_hid32(s, 8 * h + j)]);
        ev.w2[m][s][l] = __builtin_bit_cast(uint4, v);
    } else if (k < 10) {
        const int m = k - 8, h = l >> 5, c = l & 31;
        unsigned short v[4];
        for (int j = 0; j < 4; j++) v[j] = eval_bits<H>(W1[(32 * m + c) * 8 + 4 * h + j]);
        ev.w1[m][l] = __builtin_bit_cast(uint2, v);
    } else if (k == 10) {
        const int m = l >> 5, h = (l >> 4) & 1, e = l & 15;
        ev.b2[m][h][e] = b2[32 * m + mlp_row32(h, e)];
    } else if (l < 32) {
        const int h = l >> 4, m = (l >> 3) & 1, q = l & 7;
        ev.w3[h][m][q] = (unsigned)eval_bits<H>(w3[32 * m + mlp_row32(h, 2 * q)]) |
                         ((unsigned)eval_bits<H>(w3[32 * m + mlp_row32(h, 2 * q + 1)]) << 16);
    }
}

// Operand image from the flat FP32 parameters.  55 blocks of 64 lanes: one
// per operand array of training, 12 per evaluation image (bf16, fp16), and
// the scalars.
constexpr int MLP_IMAGE_BLOCKS = 55;

__global__ __launch_bounds__(64) void mlp_image_kernel(const float *p, MlpScalars sc, MlpImage *img)
{
    const int b = (int)blockIdx.x;
    const int l = (int)threadIdx.x, g = l >> 4, r = l & 15;
    const float *W1 = p + MLP_W1, *W2 = p + MLP_W2, *b2 = p + MLP_B2, *w3 = p + MLP_W3;
    if (b < 8) {
        const int h = b >> 1, c = b & 1;
        mlp_v8b v;
        for (int j = 0; j < 8; j++) v[j] = bf16_of(W2[(16 * h + r) * 64 + mlp_hid(c, g, j)]);
        img->inf.w2[h][c][l] = v;
    } else if (b < 10) {
        const int c = b - 8;
        mlp_v8b v;
        for (int j = 0; j < 8; j++) v[j] = bf16_of(w3[mlp_hid(c, g, j)]);
        img->inf.w3[c][l] = v;
    } else if (b < 14) {
        const int h = b - 10;
        mlp_v4f v;
        for (int j = 0; j < 4; j++) v[j] = b2[16 * h + 4 * g + j];
        img->inf.b2[h][l] = v;
    } else if (b < 18) {
        const int h = b - 14;
        mlp_v4s v;
        for (int j = 0; j < 4; j++) v[j] = g < 2 ? (short)bf16_bits(W1[(16 * h + r) * 8 + 4 * g + j]) : (short)0;
        img->inf.w1[h][l] = v;
    } else if (b < 26) {
        const int h = (b - 18) >> 1, c = (b - 18) & 1;
        mlp_v8b v;
        for (int j = 0; j < 8; j++) v[j] = bf16_of(W2[mlp_hid(c, g, j) * 64 + 16 * h + r]);
        img->w2t[h][c][l] = v;
    } else if (b < 30) {
        const int h = b - 26;
        mlp_v4f v;
        for (int j = 0; j < 4; j++)
            v[j] = __builtin_bit_cast(float, (unsigned)bf16_bits(w3[16 * h + 4 * g + j]) << 16);
        img->w3d[h][l] = v;
    } else if (b < 42) {
        eval_image_block<false>(b - 30, p, img->ev);
    } else if (b < 54) {
        eval_image_block<true>(b - 42, p, img->evh);
    } else if (l == 0) {
        MlpScalars s = sc;
        s.b3 = p[MLP_B3];
        img->inf.sc = s;
        img->ev.sc = s;
        img->evh.sc = s;
    }
}

// Network outputs of S samples (raw features, 7 per sample); fp16 operands
// when H
template <bool H>
__global__ __launch_bounds__(256) void mlp_forward_kernel(const MlpImage *img, const float *feat, long long S,
                                                          float *y)
{
    __shared__ MlpEval w;
    __shared__ uint4 xb[256];
    {
        const uint4 *src = (const uint4 *)eval_image<H>(img);
        uint4 *dst = (uint4 *)&w;
        for (int k = (int)threadIdx.x; k < (int)(sizeof(MlpEval) / 16); k += 256) dst[k] = src[k];
    }
    __syncthreads();
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < S;
    float f[7];
    for (int k = 0; k < 7; k++) f[k] = valid ? feat[i * 7 + k] : 0.f;
    const float out = mlp_wave_eval<H>(w, xb + (threadIdx.x & ~63u), mlp_input_eval<H>(w.sc, f));
    if (valid) y[i] = out;
}

// Per-wave LDS of the training kernel: matrices with the sample index
// contiguous (rows padded to 72 entries), for the products that contract
// over samples.
struct alignas(16) MlpTrainLds {
    __bf16 a1[64][72];   // a1[k][n]
    __bf16 dz2[64][72];  // bf16 dL/dz2 [i][n]
    __bf16 dz1[64][72];  // bf16 dL/dz1 [i][n]
    __bf16 xt[16][72];   // inputs [f][n], rows 8..15 zero
    mlp_v8b xb[64];
    float tb[64];        // targets
    float sb[64];        // 1/B for samples in the batch, 0 for padding lanes
};

// Forward and backward pass of 64 samples per block (one wave): writes the
// wave's gradient of L = (1/B) sum (y - t)^2 and its sum of (y - t)^2.
__global__ __launch_bounds__(64) void mlp_train_kernel(const MlpImage *img, const float *feat, const float *target,
                                                       const int *order, long long first, int count, float invB,
                                                       float *partial)
{
    __shared__ MlpTrainLds s;
    const MlpInfer &w = img->inf;
    const int l = (int)threadIdx.x, g = l >> 4, r = l & 15;
    const int i = (int)blockIdx.x * 64 + l;
    const bool valid = i < count;
    float f[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float t = 0.f;
    if (valid) {
        const long long id = order ? (long long)order[first + i] : first + i;
        for (int k = 0; k < 7; k++) f[k] = feat[id * 7 + k];
        t = target[id];
    }
    mlp_v8b xin = mlp_input(w.sc, f);
    if (!valid) {
        for (int k = 0; k < 8; k++) xin[k] = (__bf16)0.f;
    }
    s.xb[l] = xin;
    s.tb[l] = t;
    s.sb[l] = valid ? invB : 0.f;
    for (int k = 0; k < 8; k++) s.xt[k][l] = xin[k];
    for (int k = 8; k < 16; k++) s.xt[k][l] = (__bf16)0.f;
    wave_sync();

    const mlp_v4f zero = {0.f, 0.f, 0.f, 0.f};
    const float b3 = w.sc.b3;
    float s2[4][4], s3[4][4];
    for (int h = 0; h < 4; h++)
        for (int j = 0; j < 4; j++) s2[h][j] = s3[h][j] = 0.f;
    float loss = 0.f, gb3 = 0.f;
#pragma unroll
    for (int T = 0; T < 4; T++) {
        MlpTile o;
        mlp_tile_forward<true>(w, s.xb, T, l, o);
        const int n = 16 * T + r;
        const float d = (o.y + b3) - s.tb[n];
        const float dy = 2.f * d * s.sb[n];
        if (g == T) {
            loss += s.sb[n] > 0.f ? d * d : 0.f;
            gb3 += dy;
        }
        mlp_v8b dzop[2];
#pragma unroll
        for (int h = 0; h < 4; h++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = 16 * h + 4 * g + j;
                const float dz = img->w3d[h][l][j] * dy * gelu_grad(o.z2[h][j]);
                s2[h][j] += dz;
                s3[h][j] += (float)o.a2[h >> 1][(h & 1) * 4 + j] * dy;
                const __bf16 q = (__bf16)dz;
                dzop[h >> 1][(h & 1) * 4 + j] = q;
                s.dz2[row][n] = q;
                s.a1[row][n] = o.a1[h >> 1][(h & 1) * 4 + j];
            }
        }
#pragma unroll
        for (int h = 0; h < 4; h++) {
            mlp_v4f da = mfma_k32(img->w2t[h][0][l], dzop[0], zero);
            da = mfma_k32(img->w2t[h][1][l], dzop[1], da);
#pragma unroll
            for (int j = 0; j < 4; j++) s.dz1[16 * h + 4 * g + j][n] = (__bf16)(da[j] * gelu_grad(o.z1[h][j]));
        }
    }
    wave_sync();

    float *P = partial + (size_t)blockIdx.x * MLP_PSTRIDE;
    // dL/dW2 = dz2 a1^T and dL/dW1 = dz1 x^T: contractions over the 64 samples
#pragma unroll
    for (int hi = 0; hi < 4; hi++) {
#pragma unroll
        for (int hk = 0; hk < 4; hk++) {
            mlp_v4f acc = zero;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const mlp_v8b A = *(const mlp_v8b *)&s.dz2[16 * hi + r][32 * c + 8 * g];
                const mlp_v8b B = *(const mlp_v8b *)&s.a1[16 * hk + r][32 * c + 8 * g];
                acc = mfma_k32(A, B, acc);
            }
            for (int j = 0; j < 4; j++) P[MLP_W2 + (16 * hi + 4 * g + j) * 64 + 16 * hk + r] = acc[j];
        }
    }
#pragma unroll
    for (int hi = 0; hi < 4; hi++) {
        mlp_v4f acc = zero;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const mlp_v8b A = *(const mlp_v8b *)&s.dz1[16 * hi + r][32 * c + 8 * g];
            const mlp_v8b B = *(const mlp_v8b *)&s.xt[r][32 * c + 8 * g];
            acc = mfma_k32(A, B, acc);
        }
        if (r < 8) {
            for (int j = 0; j < 4; j++) P[MLP_W1 + (16 * hi + 4 * g + j) * 8 + r] = acc[j];
        }
    }
    // dL/db2 and dL/dw3: sums over the 16 sample columns of each lane row
#pragma unroll
    for (int h = 0; h < 4; h++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float v2 = s2[h][j], v3 = s3[h][j];
            for (int m = 1; m < 16; m <<= 1) {
                v2 += __shfl_xor(v2, m, 64);
                v3 += __shfl_xor(v3, m, 64);
            }
            if (r == 0) {
                P[MLP_B2 + 16 * h + 4 * g + j] = v2;
                P[MLP_W3 + 16 * h + 4 * g + j] = v3;
            }
        }
    }
    for (int m = 1; m < 64; m <<= 1) {
        loss += __shfl_xor(loss, m, 64);
        gb3 += __shfl_xor(gb3, m, 64);
    }
    if (l == 0) {
        P[MLP_B3] = gb3;
        P[MLP_NPARAM] = loss;
    }
}

// Sums the per-wave records in wave order; optional Adam step on the master
// weights (beta1 0.9, beta2 0.999, eps 1e-8; bc1/bc2 the bias corrections).
__global__ __launch_bounds__(256) void mlp_update_kernel(const float *partial, int nw, float *grad, double *loss_acc,
                                                         float *master, float *m1, float *m2, float lr, float bc1,
                                                         float bc2, int adam)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p > MLP_NPARAM) return;
    float gs = 0.f;
    for (int k = 0; k < nw; k++) gs += partial[(size_t)k * MLP_PSTRIDE + p];
    if (p == MLP_NPARAM) {
        if (loss_acc) *loss_acc += (double)gs;
        return;
    }
    if (grad) grad[p] = gs;
    if (adam) {
        const float a = 0.9f * m1[p] + 0.1f * gs;
        const float b = 0.999f * m2[p] + 0.001f * gs * gs;
        m1[p] = a;
        m2[p] = b;
        master[p] -= lr * (a / bc1) / (sqrtf(b / bc2) + 1e-8f);
    }
}

}  // namespace lbmc

using namespace lbmc;

struct lbmc_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
    MlpImage *img = nullptr;
    MlpScalars sc{};
    float *feat = nullptr, *target = nullptr;
    int *order = nullptr;
    long long S = 0;
    float *partial = nullptr;
    int pcap = 0;
    double *lossacc = nullptr;
    long long adam_t = 0;
    bool eval_fp16 = false;  // evaluation with fp16 operands (default bf16)
    std::string err;
};

namespace {

std::string g_mlp_error;

lbmc_status mfail(lbmc_mlp *m, lbmc_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (m) m->err = buf;
    else g_mlp_error = buf;
    return st;
}

#define MLP_HIP(m, call)                                                                         \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return mfail((m), LBMC_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

lbmc_status rebuild_image(lbmc_mlp *m)
{
    hipLaunchKernelGGL(mlp_image_kernel, dim3(MLP_IMAGE_BLOCKS), dim3(64), 0, m->stream, m->params, m->sc, m->img);
    MLP_HIP(m, hipGetLastError());
    return LBMC_OK;
}

lbmc_status ensure_partial(lbmc_mlp *m, int nw)
{
    if (nw <= m->pcap) return LBMC_OK;
    if (m->partial) MLP_HIP(m, hipFree(m->partial));
    m->partial = nullptr;
    MLP_HIP(m, hipMalloc(&m->partial, (size_t)nw * MLP_PSTRIDE * sizeof(float)));
    m->pcap = nw;
    return LBMC_OK;
}

void fill_scalars(MlpScalars &s, const lbmc_mlp_scaling *sc)
{
    memset(&s, 0, sizeof(s));
    for (int k = 0; k < 7; k++) {
        s.mu[k] = sc ? sc->mu[k] : 0.f;
        s.isd[k] = sc ? sc->isd[k] : 1.f;
    }
    s.out_a = sc ? sc->out_a : 1.f;
    s.out_b = sc ? sc->out_b : 0.f;
    s.lo = 0.5005f;
    s.hi = 2.0f;
}

// one forward/backward pass over `count` samples of the resident data
// starting at position `first` (through order when given)
lbmc_status grad_pass(lbmc_mlp *m, const int *order, long long first, int count, float *grad, int adam,
                      float lr)
{
    const int nw = (count + 63) / 64;
    lbmc_status st = ensure_partial(m, nw);
    if (st != LBMC_OK) return st;
    hipLaunchKernelGGL(mlp_train_kernel, dim3(nw), dim3(64), 0, m->stream, m->img, m->feat, m->target, order,
                       first, count, 1.f / (float)count, m->partial);
    MLP_HIP(m, hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (adam) {
        m->adam_t++;
        bc1 = (float)(1.0 - std::pow(0.9, (double)m->adam_t));
        bc2 = (float)(1.0 - std::pow(0.999, (double)m->adam_t));
    }
    hipLaunchKernelGGL(mlp_update_kernel, dim3((MLP_NPARAM + 256) / 256), dim3(256), 0, m->stream, m->partial, nw,
                       grad, m->lossacc, m->params, m->m1, m->m2, lr, bc1, bc2, adam);
    MLP_HIP(m, hipGetLastError());
    if (adam) return rebuild_image(m);
    return LBMC_OK;
}

lbmc_status upload_data(lbmc_mlp *m, const float *feat, const float *target, long long S)
{
    if (m->feat) MLP_HIP(m, hipFree(m->feat));
    if (m->target) MLP_HIP(m, hipFree(m->target));
    if (m->order) MLP_HIP(m, hipFree(m->order));
    m->feat = m->target = nullptr;
    m->order = nullptr;
    m->S = 0;
    MLP_HIP(m, hipMalloc(&m->feat, (size_t)S * 7 * sizeof(float)));
    MLP_HIP(m, hipMalloc(&m->target, (size_t)S * sizeof(float)));
    MLP_HIP(m, hipMalloc(&m->order, (size_t)S * sizeof(int)));
    MLP_HIP(m, hipMemcpy(m->feat, feat, (size_t)S * 7 * sizeof(float), hipMemcpyHostToDevice));
    if (target) MLP_HIP(m, hipMemcpy(m->target, target, (size_t)S * sizeof(float), hipMemcpyHostToDevice));
    else MLP_HIP(m, hipMemset(m->target, 0, (size_t)S * sizeof(float)));
    m->S = S;
    return LBMC_OK;
}

}  // namespace

namespace lbmc {

// Operand image of the trainer's current weights, with the fallback band
// [lo, hi], into dst (device memory of the same device), on stream s.
lbmc_status mlp_build_image(const lbmc_mlp *m, void *dst, int device, double lo, double hi, hipStream_t s)
{
    if (!m || !dst || m->device != device) return LBMC_ERR_INVALID;
    if (hipStreamSynchronize(m->stream) != hipSuccess) return LBMC_ERR_DEVICE;
    MlpScalars sc = m->sc;
    sc.lo = (float)lo;
    sc.hi = (float)hi;
    hipLaunchKernelGGL(mlp_image_kernel, dim3(MLP_IMAGE_BLOCKS), dim3(64), 0, s, m->params, sc, (MlpImage *)dst);
    if (hipGetLastError() != hipSuccess) return LBMC_ERR_DEVICE;
    return LBMC_OK;
}

size_t mlp_image_bytes() { return sizeof(MlpImage); }

bool mlp_eval_fp16(const lbmc_mlp *m) { return m && m->eval_fp16; }

}  // namespace lbmc

extern "C" {

const char *lbmc_mlp_last_error(const lbmc_mlp *m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

lbmc_status lbmc_mlp_create(int device, const float *params, const lbmc_mlp_scaling *sc, lbmc_mlp **out)
{
    if (!out || !params) return mfail(nullptr, LBMC_ERR_INVALID, "lbmc_mlp_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return mfail(nullptr, LBMC_ERR_DEVICE, "lbmc_mlp_create: no HIP device");
    if (device < 0 || device >= ndev) return mfail(nullptr, LBMC_ERR_INVALID, "lbmc_mlp_create: bad device %d", device);
    for (int k = 0; k < MLP_NPARAM; k++) {
        if (!std::isfinite(params[k])) return mfail(nullptr, LBMC_ERR_INVALID, "lbmc_mlp_create: parameter %d not finite", k);
    }
    lbmc_mlp *m = new lbmc_mlp();
    m->device = device;
    fill_scalars(m->sc, sc);
    lbmc_status st = LBMC_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (e != hipSuccess && st == LBMC_OK) st = mfail(nullptr, LBMC_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    chk(hipSetDevice(device), "hipSetDevice");
    chk(hipStreamCreate(&m->stream), "hipStreamCreate");
    chk(hipMalloc(&m->params, MLP_NPARAM * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&m->m1, MLP_NPARAM * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&m->m2, MLP_NPARAM * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&m->grad, MLP_NPARAM * sizeof(float)), "hipMalloc");
    chk(hipMalloc(&m->img, sizeof(MlpImage)), "hipMalloc");
    chk(hipMalloc(&m->lossacc, sizeof(double)), "hipMalloc");
    if (st == LBMC_OK) {
        chk(hipMemcpy(m->params, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        chk(hipMemset(m->m1, 0, MLP_NPARAM * sizeof(float)), "hipMemset");
        chk(hipMemset(m->m2, 0, MLP_NPARAM * sizeof(float)), "hipMemset");
    }
    if (st == LBMC_OK) st = rebuild_image(m);
    if (st == LBMC_OK) chk(hipStreamSynchronize(m->stream), "hipStreamSynchronize");
    if (st != LBMC_OK) {
        g_mlp_error = m->err.empty() ? g_mlp_error : m->err;
        lbmc_mlp_destroy(m);
        return st;
    }
    *out = m;
    return LBMC_OK;
}

void lbmc_mlp_destroy(lbmc_mlp *m)
{
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void *bufs[] = {m->params, m->m1, m->m2, m->grad, m->img, m->feat, m->target, m->order, m->partial, m->lossacc};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

lbmc_status lbmc_mlp_set_params(lbmc_mlp *m, const float *params)
{
    if (!m || !params) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_set_params: null argument");
    MLP_HIP(m, hipSetDevice(m->device));
    MLP_HIP(m, hipMemcpyAsync(m->params, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MLP_HIP(m, hipMemsetAsync(m->m1, 0, MLP_NPARAM * sizeof(float), m->stream));
    MLP_HIP(m, hipMemsetAsync(m->m2, 0, MLP_NPARAM * sizeof(float), m->stream));
    m->adam_t = 0;
    lbmc_status st = rebuild_image(m);
    if (st != LBMC_OK) return st;
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return LBMC_OK;
}

lbmc_status lbmc_mlp_get_params(lbmc_mlp *m, float *params)
{
    if (!m || !params) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_get_params: null argument");
    MLP_HIP(m, hipSetDevice(m->device));
    MLP_HIP(m, hipMemcpyAsync(params, m->params, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return LBMC_OK;
}

lbmc_status lbmc_mlp_set_scaling(lbmc_mlp *m, const lbmc_mlp_scaling *sc)
{
    if (!m || !sc) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_set_scaling: null argument");
    MLP_HIP(m, hipSetDevice(m->device));
    fill_scalars(m->sc, sc);
    lbmc_status st = rebuild_image(m);
    if (st != LBMC_OK) return st;
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    return LBMC_OK;
}

lbmc_status lbmc_mlp_forward(lbmc_mlp *m, const float *feat, long long S, float *y)
{
    if (!m || (S > 0 && (!feat || !y)) || S < 0) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_forward: bad argument");
    if (S == 0) return LBMC_OK;
    if (S > (1LL << 31) / 8) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_forward: too many samples");
    MLP_HIP(m, hipSetDevice(m->device));
    float *dx = nullptr, *dy = nullptr;
    MLP_HIP(m, hipMalloc(&dx, (size_t)S * 7 * sizeof(float)));
    hipError_t e = hipMalloc(&dy, (size_t)S * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dx, feat, (size_t)S * 7 * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) {
        if (m->eval_fp16)
            hipLaunchKernelGGL(mlp_forward_kernel<true>, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, m->stream,
                               m->img, dx, S, dy);
        else
            hipLaunchKernelGGL(mlp_forward_kernel<false>, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, m->stream,
                               m->img, dx, S, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, (size_t)S * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) return mfail(m, LBMC_ERR_DEVICE, "lbmc_mlp_forward: %s", hipGetErrorString(e));
    return LBMC_OK;
}

lbmc_status lbmc_mlp_set_eval_fp16(lbmc_mlp *m, int on)
{
    if (!m) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_set_eval_fp16: null model");
    m->eval_fp16 = on != 0;
    return LBMC_OK;
}

lbmc_status lbmc_mlp_set_data(lbmc_mlp *m, const float *feat, const float *target, long long S)
{
    if (!m || !feat || !target || S <= 0) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_set_data: bad argument");
    if (S >= (1LL << 31)) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_set_data: at most 2^31 - 1 samples");
    MLP_HIP(m, hipSetDevice(m->device));
    return upload_data(m, feat, target, S);
}

lbmc_status lbmc_mlp_loss_grad(lbmc_mlp *m, const float *feat, const float *target, long long S, float *grad,
                               double *loss)
{
    if (!m || !feat || !target || S <= 0) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_loss_grad: bad argument");
    if (S > (1 << 24)) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_loss_grad: at most 2^24 samples in one batch");
    MLP_HIP(m, hipSetDevice(m->device));
    lbmc_status st = upload_data(m, feat, target, S);
    if (st != LBMC_OK) return st;
    MLP_HIP(m, hipMemsetAsync(m->lossacc, 0, sizeof(double), m->stream));
    st = grad_pass(m, nullptr, 0, (int)S, m->grad, 0, 0.f);
    if (st != LBMC_OK) return st;
    double l = 0.0;
    if (grad) MLP_HIP(m, hipMemcpyAsync(grad, m->grad, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipMemcpyAsync(&l, m->lossacc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(m, hipStreamSynchronize(m->stream));
    if (loss) *loss = l / (double)S;
    return LBMC_OK;
}

lbmc_status lbmc_mlp_train(lbmc_mlp *m, int epochs, int batch, float lr, unsigned long long seed, double *epoch_loss)
{
    if (!m || epochs < 0 || batch <= 0 || !(lr > 0.f)) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_train: bad argument");
    if (m->S <= 0) return mfail(m, LBMC_ERR_STATE, "lbmc_mlp_train: no data (lbmc_mlp_set_data)");
    if (batch > (1 << 24)) return mfail(m, LBMC_ERR_INVALID, "lbmc_mlp_train: batch above 2^24");
    MLP_HIP(m, hipSetDevice(m->device));
    std::vector<int> order((size_t)m->S);
    std::mt19937_64 rng(seed);
    for (int ep = 0; ep < epochs; ep++) {
        std::iota(order.begin(), order.end(), 0);
        for (long long k = m->S - 1; k > 0; k--) {
            const long long j = (long long)(rng() % (unsigned long long)(k + 1));
            std::swap(order[(size_t)k], order[(size_t)j]);
        }
        MLP_HIP(m, hipMemcpyAsync(m->order, order.data(), (size_t)m->S * sizeof(int), hipMemcpyHostToDevice,
                                  m->stream));
        MLP_HIP(m, hipMemsetAsync(m->lossacc, 0, sizeof(double), m->stream));
        for (long long first = 0; first < m->S; first += batch) {
            const int count = (int)std::min<long long>(batch, m->S - first);
            lbmc_status st = grad_pass(m, m->order, first, count, nullptr, 1, lr);
            if (st != LBMC_OK) return st;
        }
        double l = 0.0;
        MLP_HIP(m, hipMemcpyAsync(&l, m->lossacc, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        MLP_HIP(m, hipStreamSynchronize(m->stream));
        if (epoch_loss) epoch_loss[ep] = l / (double)m->S;
    }
    return LBMC_OK;
}

}  // extern "C"
