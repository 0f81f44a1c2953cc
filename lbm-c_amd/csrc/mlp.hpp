// Learned relaxation-time closure: a 7 -> 64 -> 64 -> 1 perceptron with
// GELU activations, evaluated and trained on the matrix cores.  It takes the
// place of the reference's Smagorinsky hook (LBM-C/collision.cu:445-463,
// called from bgk_collision at :60-62): tau' = out_a * y + out_b, and the
// Smagorinsky tau is used wherever tau' leaves the band [lo, hi].
//
// Arithmetic: bf16 operands on the matrix cores, FP32 accumulation, biases
// and derivatives in FP32.  Training (mlp_tile_forward and the backward
// pass) uses 16x16 tiles (MFMA 16x16x16 bf16 for the first layer, 16x16x32
// bf16 for the others); evaluation in the closure steps (mlp_wave_eval)
// uses 32x32 tiles (32x32x8 and 32x32x16 bf16, the output layer as bf16 dot
// products).  A model may instead be evaluated with fp16 operands (the same
// tiles with f16 products: mlp_wave_eval<true>), which saves two vector
// instructions per pair of activations.  GELU(z) = z Phi(z), Phi from an
// odd degree-9 polynomial (least-squares fit of the normal CDF on
// [-3.5, 3.5], |error| <= 8.6e-4 in exact arithmetic), clamped to [0, 1].
// The polynomial runs in packed FP16 (gelu_pair, gelu_oct): x = fp16(z),
// t = x*x, four fused multiply-adds, the [0, 1] clamp folded into the last
// one; then z * Phi in FP32 and the bf16 rounding the next product needs
// (10 vector instructions per pair against 17 for the same polynomial in
// FP32), or with fp16 operands x * Phi in packed fp16, the next product's
// operand as it stands (8 per pair).  Over every finite fp16 x this Phi is
// within 1.8e-3 of the normal CDF and exactly 0 / 1 beyond -3.5 / 3.5, so
// the input needs no clamp; |GELU error| <= 5.2e-3, below the bf16 half-ulp
// of the result (2^-9 relative); the fp16 result adds half an fp16 ulp
// (2^-11 relative).  With fp16 operands a pre-activation beyond 65504 makes
// x infinite: the GELU is then inf or NaN and tau' leaves its band (the
// Smagorinsky tau is used).
//
// Training's wave mapping (64 lanes, lane l = 16 g + r): a wave takes 64
// samples as four tiles of 16 (mlp_wave_eval describes the evaluation's).
// Each layer is computed transposed, H = W X, with
// weights as the A operand and samples as the columns of B, so the MFMA
// result layout of one layer (lane holds rows 4g..4g+3 of column r) is the
// B-operand layout of the next: no data moves between lanes inside the
// network.  For the K = 32 products the hidden units of the two 16-row
// tiles 2c, 2c+1 are taken in the order hid(c, g, j) below, which is the
// order in which lane g of the previous result already holds them.  The
// output layer uses w3 on all 16 rows, so every lane receives y of column
// r; lane 16g + r keeps tile g's value, which is its own sample.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace lbmc {

constexpr int MLP_NIN = 7;
constexpr int MLP_H = 64;
// flat parameters: w1[64][8] (column 7 is the first-layer bias), w2[64][64]
// (row = output unit), b2[64], w3[64], b3
constexpr int MLP_W1 = 0, MLP_W2 = 512, MLP_B2 = 4608, MLP_W3 = 4672, MLP_B3 = 4736;
constexpr int MLP_NPARAM = 4737;
// per-wave partial gradient record: the parameters, then the squared-error sum
constexpr int MLP_PSTRIDE = 4740;

typedef short mlp_v4s __attribute__((ext_vector_type(4)));
typedef __bf16 mlp_v8b __attribute__((ext_vector_type(8)));
typedef float mlp_v4f __attribute__((ext_vector_type(4)));
typedef _Float16 mlp_v4h __attribute__((ext_vector_type(4)));
typedef _Float16 mlp_h8 __attribute__((ext_vector_type(8)));

// hidden unit held at K position 8g + j of the 32-wide chunk c
__host__ __device__ constexpr int mlp_hid(int c, int g, int j) { return 16 * (2 * c + (j >> 2)) + 4 * g + (j & 3); }

struct MlpScalars {
    float b3, out_a, out_b, lo, hi;
    float mu[7], isd[7];  // input normalisation x = (feature - mu) * isd
    float pad[13];
};

// Operand-ordered weights for the forward pass: entry [..][l] is what lane l
// feeds to the matrix cores, so one LDS read per operand and lane.
struct MlpInfer {
    mlp_v8b w2[4][2][64];  // A of layer 2, row tile h, chunk c: W2[16h + r][hid(c, g, j)]
    mlp_v8b w3[2][64];     // A of layer 3, chunk c: w3[hid(c, g, j)] on every row
    mlp_v4f b2[4][64];     // layer-2 accumulator start: b2[16h + 4g + j]
    mlp_v4s w1[4][64];     // A of layer 1, row tile h: W1[16h + r][4g + j] (zero for g >= 2)
    MlpScalars sc;
};

typedef float mlp_v16f __attribute__((ext_vector_type(16)));

// Row of accumulator element e held by lane half h (lanes 32h..32h+31) in a
// 32x32 matrix-core result: column = lane & 31.
__host__ __device__ constexpr int mlp_row32(int h, int e) { return 8 * (e >> 2) + 4 * h + (e & 3); }
// Hidden unit at K position k (lane half k >> 3, element k & 7) of K-step s
// of the evaluation's second layer: the order in which the first layer's
// result (two 32-row tiles) already sits in the lanes.
__host__ __device__ constexpr int mlp_hid32(int s, int k)
{
    return 16 * s + 8 * ((k & 7) >> 2) + 4 * (k >> 3) + (k & 3);
}

// Operand-ordered weights of the closure's evaluation (mlp_wave_eval):
// 32x32 products, lane l = 32h + c; 16-bit values, bf16 or fp16 by image
// (MlpImage::ev, evh), held as raw bits.
struct MlpEval {
    uint4 w2[2][4][64];    // A of layer 2, row tile m, K-step s: W2[32m + c][hid32(s, 8h + j)]
    uint2 w1[2][64];       // A of layer 1, row tile m: W1[32m + c][4h + j]
    mlp_v16f b2[2][2];     // layer-2 accumulator start, row tile m, lane half h: b2[32m + row32(h, e)]
    unsigned w3[2][2][8];  // pairs (low, high) of w3[32m + row32(h, 2q)], w3[32m + row32(h, 2q + 1)], [h][m][q]
    MlpScalars sc;
};

struct MlpImage {
    MlpInfer inf;
    // backward only
    mlp_v8b w2t[4][2][64];  // A of W2^T, row tile h, chunk c: W2[hid(c, g, j)][16h + r]
    mlp_v4f w3d[4][64];     // w3[16h + 4g + j] as rounded to bf16, in FP32
    // evaluation (closure steps and lbmc_mlp_forward): bf16 and fp16 operands
    MlpEval ev, evh;
};

static_assert(sizeof(MlpInfer) % 16 == 0, "MlpInfer is copied in 16-byte units");
static_assert(sizeof(MlpEval) % 16 == 0, "MlpEval is copied in 16-byte units");
static_assert(sizeof(MlpImage) % 16 == 0, "MlpImage is copied in 16-byte units");
static_assert(offsetof(MlpImage, ev) % 64 == 0 && offsetof(MlpImage, evh) % 64 == 0, "MlpEval holds 64-byte vectors");

// the evaluation image of a model: fp16 operands (H) or bf16
template <bool H>
__host__ __device__ inline const MlpEval *eval_image(const void *img)
{
    return H ? &((const MlpImage *)img)->evh : &((const MlpImage *)img)->ev;
}

// GELU coefficients: Phi(x) - 1/2 = x (c1 + t (c3 + t (c5 + t (c7 + t c9)))), t = x^2
constexpr float GELU_C1 = 0.3964735269546509f;
constexpr float GELU_C3 = -0.061638616025447845f;
constexpr float GELU_C5 = 0.007179733831435442f;
constexpr float GELU_C7 = -0.0004616279620677233f;
constexpr float GELU_C9 = 1.2130119102948811e-05f;

typedef _Float16 mlp_h2 __attribute__((ext_vector_type(2)));
typedef __bf16 mlp_b2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ mlp_h2 mlp_h2_splat(float c) { return mlp_h2{(_Float16)c, (_Float16)c}; }

// 1/2 + x P(x^2) of two fp16 inputs, every operation rounded to fp16 once
// (fused multiply-adds), before the [0, 1] clamp
__device__ __forceinline__ mlp_h2 gelu_phi_raw(mlp_h2 x)
{
    const mlp_h2 t = x * x;
    mlp_h2 p = __builtin_elementwise_fma(t, mlp_h2_splat(GELU_C9), mlp_h2_splat(GELU_C7));
    p = __builtin_elementwise_fma(t, p, mlp_h2_splat(GELU_C5));
    p = __builtin_elementwise_fma(t, p, mlp_h2_splat(GELU_C3));
    p = __builtin_elementwise_fma(t, p, mlp_h2_splat(GELU_C1));
    return __builtin_elementwise_fma(x, p, mlp_h2_splat(0.5f));
}

__device__ __forceinline__ mlp_h2 gelu_phi_clamp(mlp_h2 s)
{
    return __builtin_elementwise_min(__builtin_elementwise_max(s, mlp_h2_splat(0.f)), mlp_h2_splat(1.f));
}

// GELU of two FP32 pre-activations, rounded to bf16 (the next layer's
// operands).  z * Phi is one FP32 rounding (a mixed-precision FMA with a zero
// addend), then round to nearest even bf16.
__device__ __forceinline__ mlp_b2 gelu_pair(float z0, float z1)
{
    const mlp_h2 s = gelu_phi_clamp(gelu_phi_raw(mlp_h2{(_Float16)z0, (_Float16)z1}));
    return mlp_b2{(__bf16)__builtin_fmaf(z0, (float)s.x, 0.f), (__bf16)__builtin_fmaf(z1, (float)s.y, 0.f)};
}

// d GELU / dz of the evaluated form: Phi(x) + z Phi'(x) with x = fp16(z),
// Phi and the clamp test from the same fp16 evaluation as gelu_pair, Phi'
// (the polynomial's derivative) in FP32; Phi' is zero where Phi is clamped.
__device__ __forceinline__ float gelu_grad(float z)
{
    const _Float16 xh = (_Float16)z;
    const mlp_h2 raw = gelu_phi_raw(mlp_h2{xh, xh});
    const float sr = (float)raw.x;
    const float x = (float)xh, t = x * x;
    float dp = fmaf(t, 9.f * GELU_C9, 7.f * GELU_C7);
    dp = fmaf(t, dp, 5.f * GELU_C5);
    dp = fmaf(t, dp, 3.f * GELU_C3);
    dp = fmaf(t, dp, GELU_C1);
    const bool inside = sr > 0.f && sr < 1.f;
    return fminf(fmaxf(sr, 0.f), 1.f) + (inside ? z * dp : 0.f);
}

// round to nearest even bf16 (finite inputs)
__host__ __device__ inline unsigned short bf16_bits(float x)
{
    const unsigned u = __builtin_bit_cast(unsigned, x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ mlp_v4f mfma_k16(mlp_v4s a, mlp_v4s b, mlp_v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ mlp_v4f mfma_k32(mlp_v8b a, mlp_v8b b, mlp_v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS written by one lane and read by another of the same wave
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a sample's raw features -> normalised bf16 input row (constant-1 bias input last)
__device__ __forceinline__ mlp_v8b mlp_input(const MlpScalars &s, const float (&f)[7])
{
    mlp_v8b r;
#pragma unroll
    for (int k = 0; k < 7; k++) r[k] = (__bf16)((f[k] - s.mu[k]) * s.isd[k]);
    r[7] = (__bf16)1.f;
    return r;
}

// Forward values of one 16-sample tile, in result layout (lane 16g + r:
// hidden rows 4g..4g+3 of sample column r).
struct MlpTile {
    mlp_v4f z1[4], z2[4];  // pre-activations of hidden tiles h (kept for the backward pass)
    mlp_v8b a1[2], a2[2];  // bf16 activations as B operands of the next layer, chunk c
    float y;               // output of sample column r, without b3
};

// Forward pass of tile T.  xb holds the wave's 64 input rows.  Every lane of
// the wave must take part (the matrix-core operands span the wave).
template <bool KEEP>
__device__ __forceinline__ void mlp_tile_forward(const MlpInfer &w, const mlp_v8b *xb, int T, int l,
                                                 MlpTile &o)
{
    const int g = l >> 4, r = l & 15;
    const mlp_v4f zero = {0.f, 0.f, 0.f, 0.f};
    // The weight operands are read from LDS at their point of use: an opaque
    // lane index keeps the compiler from hoisting all of them out of a loop
    // over tiles (64 registers held for the whole loop).
    int lw = l;
    asm volatile("" : "+v"(lw));
    // layer-1 B operand: inputs 4g..4g+3 of sample 16T + r (K rows 8..15 are zero)
    const mlp_v4s *row = (const mlp_v4s *)&xb[16 * T + r];
    mlp_v4s xop = row[g & 1];
    if (g >= 2) xop = mlp_v4s{0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 4; h++) {
        const mlp_v4f z = mfma_k16(w.w1[h][lw], xop, zero);
        if constexpr (KEEP) o.z1[h] = z;
#pragma unroll
        for (int j = 0; j < 4; j += 2) {
            const mlp_b2 a = gelu_pair(z[j], z[j + 1]);
            o.a1[h >> 1][(h & 1) * 4 + j] = a.x;
            o.a1[h >> 1][(h & 1) * 4 + j + 1] = a.y;
        }
    }
#pragma unroll
    for (int h = 0; h < 4; h++) {
        mlp_v4f acc = w.b2[h][lw];
        acc = mfma_k32(w.w2[h][0][lw], o.a1[0], acc);
        acc = mfma_k32(w.w2[h][1][lw], o.a1[1], acc);
        if constexpr (KEEP) o.z2[h] = acc;
#pragma unroll
        for (int j = 0; j < 4; j += 2) {
            const mlp_b2 a = gelu_pair(acc[j], acc[j + 1]);
            o.a2[h >> 1][(h & 1) * 4 + j] = a.x;
            o.a2[h >> 1][(h & 1) * 4 + j + 1] = a.y;
        }
    }
    mlp_v4f yv = mfma_k32(w.w3[0][lw], o.a2[0], zero);
    yv = mfma_k32(w.w3[1][lw], o.a2[1], yv);
    o.y = yv[0];
}

// 32x32 products of 16-bit operands held as raw bits: fp16 (H) or bf16
template <bool H>
__device__ __forceinline__ mlp_v16f mfma_32k8(uint2 a, uint2 b)
{
    if constexpr (H)
        return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(mlp_v4h, a), __builtin_bit_cast(mlp_v4h, b),
                                                    mlp_v16f{}, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(__builtin_bit_cast(mlp_v4s, a),
                                                        __builtin_bit_cast(mlp_v4s, b), mlp_v16f{}, 0, 0, 0);
}

template <bool H>
__device__ __forceinline__ mlp_v16f mfma_32k16(uint4 a, uint4 b, mlp_v16f c)
{
    if constexpr (H)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(mlp_h8, a), __builtin_bit_cast(mlp_h8, b), c,
                                                      0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mlp_v8b, a), __builtin_bit_cast(mlp_v8b, b),
                                                       c, 0, 0, 0);
}

__device__ __forceinline__ mlp_h8 mlp_h8_splat(float c)
{
    const _Float16 h = (_Float16)c;
    return mlp_h8{h, h, h, h, h, h, h, h};
}

// GELU of z[e0 .. e0 + 7] as the next product's operand, element j =
// z[e0 + j]: the Phi of gelu_pair (four pairs whose packed operations issue
// side by side), then z * Phi rounded to bf16 as in gelu_pair, or with fp16
// operands (H) x * Phi rounded to fp16 once, x = fp16(z)
template <bool H>
__device__ __forceinline__ uint4 gelu_oct(const mlp_v16f &z, int e0)
{
    mlp_h8 x;
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = (_Float16)z[e0 + j];
    const mlp_h8 t = x * x;
    mlp_h8 p = __builtin_elementwise_fma(t, mlp_h8_splat(GELU_C9), mlp_h8_splat(GELU_C7));
    p = __builtin_elementwise_fma(t, p, mlp_h8_splat(GELU_C5));
    p = __builtin_elementwise_fma(t, p, mlp_h8_splat(GELU_C3));
    p = __builtin_elementwise_fma(t, p, mlp_h8_splat(GELU_C1));
    mlp_h8 s = __builtin_elementwise_fma(x, p, mlp_h8_splat(0.5f));
    s = __builtin_elementwise_min(__builtin_elementwise_max(s, mlp_h8_splat(0.f)), mlp_h8_splat(1.f));
    if constexpr (H) {
        return __builtin_bit_cast(uint4, x * s);
    } else {
        mlp_v8b r;
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = (__bf16)__builtin_fmaf(z[e0 + j], (float)s[j], 0.f);
        return __builtin_bit_cast(uint4, r);
    }
}

// sum over the four 16-bit pairs of a and b (fp16 or bf16), added to c in FP32
template <bool H>
__device__ __forceinline__ float dot8(uint4 a, uint4 b, float c)
{
    const unsigned av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if constexpr (H)
            c = __builtin_amdgcn_fdot2(__builtin_bit_cast(mlp_h2, av[q]), __builtin_bit_cast(mlp_h2, bv[q]), c, false);
        else
            c = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(mlp_b2, av[q]), __builtin_bit_cast(mlp_b2, bv[q]),
                                                c, false);
    }
    return c;
}

// a sample's raw features -> normalised input row of the evaluation (bf16,
// or fp16 when H; the constant-1 bias input last)
template <bool H>
__device__ __forceinline__ uint4 mlp_input_eval(const MlpScalars &s, const float (&f)[7])
{
    if constexpr (H) {
        mlp_h8 r;
#pragma unroll
        for (int k = 0; k < 7; k++) r[k] = (_Float16)((f[k] - s.mu[k]) * s.isd[k]);
        r[7] = (_Float16)1.f;
        return __builtin_bit_cast(uint4, r);
    } else {
        return __builtin_bit_cast(uint4, mlp_input(s, f));
    }
}

// Network output for the lane's own sample (input row xin), from the
// evaluation image of the operand type (bf16, or fp16 when H).  xb: the
// wave's 64-entry LDS staging area.  Called by all 64 lanes together.
//
// The wave's 64 samples are two tiles of 32 (lane l = 32h + c holds column c
// of a tile), each layer again computed transposed with the weights as A:
// layer 1 is two 32x32x8 products (K = the 7 inputs and the bias input),
// layer 2 two row tiles of four 32x32x16 products whose K order mlp_hid32 is
// the order in which the lanes hold layer 1's GELUs, and the output layer a
// dot product of each lane's 32 layer-2 GELUs with w3 (16-bit pairs, FP32
// sum) plus the other lane half's.  Against the 16x16 tiles of
// mlp_tile_forward (training) this skips the zero K rows of layer 1 and the
// 15 unused rows of the output layer's product, a quarter of the
// matrix-core work, and reads half the weight operands per sample; with
// bf16 operands only the FP32 summation order differs from training.
template <bool H>
__device__ __forceinline__ float mlp_wave_eval(const MlpEval &w, uint4 *xb, uint4 xin)
{
    const int l = (int)(threadIdx.x & 63), h = l >> 5, c = l & 31;
    xb[l] = xin;
    wave_sync();
    // opaque lane and half indices: the operand reads stay inside the tile
    // loop instead of holding ~90 registers across it
    int lw = l, hw = h;
    asm volatile("" : "+v"(lw), "+v"(hw));
    float y = 0.f;
#pragma unroll 1
    for (int t = 0; t < 2; t++) {
        // layer-1 B operand: inputs 4h..4h+3 of sample 32t + c
        const uint2 xop = ((const uint2 *)&xb[32 * t + c])[h];
        // one 32-row tile at a time, phases fenced from the scheduler: the
        // node's own state stays in registers beside the network's
        uint4 a[4];
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const mlp_v16f z = mfma_32k8<H>(w.w1[m][lw], xop);
            a[2 * m] = gelu_oct<H>(z, 0);
            a[2 * m + 1] = gelu_oct<H>(z, 8);
            __builtin_amdgcn_sched_barrier(0);
        }
        float part = 0.f;
#pragma unroll
        for (int m = 0; m < 2; m++) {
            mlp_v16f acc = w.b2[m][hw];
#pragma unroll
            for (int s = 0; s < 4; s++) acc = mfma_32k16<H>(w.w2[m][s][lw], a[s], acc);
            __builtin_amdgcn_sched_barrier(0);
            const uint4 *w3 = (const uint4 *)w.w3[hw][m];
            part = dot8<H>(gelu_oct<H>(acc, 0), w3[0], part);
            part = dot8<H>(gelu_oct<H>(acc, 8), w3[1], part);
            __builtin_amdgcn_sched_barrier(0);
        }
        // both lane halves hold partial sums of sample 32t + c
        const unsigned pb = __builtin_bit_cast(unsigned, part);
        const auto sw = __builtin_amdgcn_permlane32_swap(pb, pb, false, false);
        part = __builtin_bit_cast(float, (unsigned)sw[0]) + __builtin_bit_cast(float, (unsigned)sw[1]);
        if (h == t) y = part;
    }
    wave_sync();
    return y + w.sc.b3;
}

}  // namespace lbmc
