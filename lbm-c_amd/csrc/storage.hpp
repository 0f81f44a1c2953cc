// Population storage formats.
//
//   FP64  : double slots, double arithmetic (the reference's precision)
//   FP32  : float slots holding g_i = f_i - w_i, float arithmetic
//   FP16S : _Float16 slots holding 2^12 g_i, float arithmetic
//   BF16S : __bf16 slots holding g_i, float arithmetic
//
// The shifted formats store the deviation from the rest equilibrium so that
// the 24-, 11- or 8-bit significand is spent on the part of f that varies
// (|g| ~ 1e-3 w against |f| ~ w at low Mach number: for FP32 the rounding of a
// slot drops from ~3e-8 to ~1e-10); density is accumulated as 1 + sum g_i and
// momentum as sum e_i g_i (the rest weights sum to one and carry no momentum).
// FP16S scales g by 2^12 so that deviations down to ~1e-8 stay in the
// normal range of _Float16; |g| < 65504 / 4096 (about 15.99) fits, larger
// deviations store +-inf.
//
// The 16-bit formats round the step's stores stochastically (store_sr, the
// hardware's stochastic-rounding conversion, which draws on the high bits of
// a random word): a slot then holds its value's expectation, so increments
// below half a unit in the last place -- a weak body force adds ~1e-7 per
// call to populations of ~1e-3 -- accumulate instead of being rounded away
// call after call.  Set-up writes (initial states, set_f) round to nearest.
//
// gather_prio: the D3Q15 step raises the wave's issue priority while it
// issues its population loads.  For the 16-bit formats (twice the load
// instructions per byte of FP32) that keeps more bytes in flight: -3 to -4 %
// time per step on a 256^3 lattice.  It costs 3 % at FP32, whose step runs
// under a residency cap, and 1-3 % on D2Q9 lattices (9 loads per node), so
// neither takes it.
#pragma once

#include <hip/hip_runtime.h>

namespace lbmc {

enum StorageId { ST_FP64 = 0, ST_FP32 = 1, ST_FP16S = 2, ST_BF16S = 3 };

struct StFP64 {
    using S = double;
    using C = double;
    static constexpr bool shifted = false;
    static constexpr bool stochastic = false;
    static constexpr bool gather_prio = false;
    static constexpr int id = ST_FP64;
    __device__ static inline C load(S v) { return v; }
    __device__ static inline S store(C v) { return v; }
};

struct StFP32 {
    using S = float;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr bool stochastic = false;
    static constexpr bool gather_prio = false;
    static constexpr int id = ST_FP32;
    __device__ static inline C load(S v) { return v; }
    __device__ static inline S store(C v) { return v; }
};

struct StFP16S {
    using S = _Float16;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr bool stochastic = true;
    static constexpr bool gather_prio = true;
    static constexpr int id = ST_FP16S;
    static constexpr float kScale = 4096.f, kInv = 1.f / 4096.f;
    // convert and scale in one mixed-precision FMA (the fp16 operand read in
    // place from either half of a register); the zero addend makes a -0 slot
    // load as +0, the same in every kernel
    __device__ static inline C load(S v) { return __builtin_fmaf((float)v, kInv, 0.f); }
    __device__ static inline S store(C v) { return (_Float16)(v * kScale); }  // round to nearest even
    __device__ static inline S store_sr(C v, unsigned r)
    {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        return __builtin_amdgcn_cvt_sr_f16_f32((h2){0, 0}, v * kScale, r, false)[0];
    }
    // two stochastic-rounding stores into one 32-bit word: v0 low, v1 high
    __device__ static inline unsigned store_sr_pair(C v0, unsigned r0, C v1, unsigned r1)
    {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        h2 w = __builtin_amdgcn_cvt_sr_f16_f32((h2){0, 0}, v0 * kScale, r0, false);
        w = __builtin_amdgcn_cvt_sr_f16_f32(w, v1 * kScale, r1, true);
        return __builtin_bit_cast(unsigned, w);
    }
};

struct StBF16S {
    using S = __bf16;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr bool stochastic = true;
    static constexpr bool gather_prio = true;
    static constexpr int id = ST_BF16S;
    __device__ static inline C load(S v) { return (float)v; }
    __device__ static inline S store(C v) { return (__bf16)v; }  // round to nearest even
    __device__ static inline S store_sr(C v, unsigned r)
    {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        return __builtin_amdgcn_cvt_sr_bf16_f32((b2){0, 0}, v, r, false)[0];
    }
    // two stochastic-rounding stores into one 32-bit word: v0 low, v1 high
    __device__ static inline unsigned store_sr_pair(C v0, unsigned r0, C v1, unsigned r1)
    {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        b2 w = __builtin_amdgcn_cvt_sr_bf16_f32((b2){0, 0}, v0, r0, false);
        w = __builtin_amdgcn_cvt_sr_bf16_f32(w, v1, r1, true);
        return __builtin_bit_cast(unsigned, w);
    }
};

// Random words of the stochastic stores: a hash of (node, call seed: the
// call's shift counters), then
// one Weyl step (+ 2^32 / golden ratio) per population.  Each word is
// uniform over the node hash; within a node the directions' words are
// spread evenly over the interval (low discrepancy), so their rounding
// decisions stay unbiased one by one and the node's total rounding error
// is, if anything, smaller than with independent words.
__device__ __forceinline__ unsigned sr_seed(int n, unsigned tseed)
{
    unsigned h = (unsigned)n * 0x9E3779B1u ^ tseed * 0x85EBCA77u;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ unsigned sr_next(unsigned h) { return h + 0x9E3779B9u; }

}  // namespace lbmc
