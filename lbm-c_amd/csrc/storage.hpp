// Population storage formats.
//
//   FP64  : double slots, double arithmetic (the reference's precision)
//   FP32  : float slots holding g_i = f_i - w_i, float arithmetic
//   FP16S : _Float16 slots holding g_i = f_i - w_i, float arithmetic
//   BF16S : __bf16 slots holding g_i = f_i - w_i, float arithmetic
//
// The shifted formats store the deviation from the rest equilibrium so that
// the 24-, 11- or 8-bit significand is spent on the part of f that varies
// (|g| ~ 1e-3 w against |f| ~ w at low Mach number: for FP32 the rounding of a
// slot drops from ~3e-8 to ~1e-10); density is accumulated as 1 + sum g_i and
// momentum as sum e_i g_i (the rest weights sum to one and carry no momentum).
#pragma once

#include <hip/hip_runtime.h>

namespace lbmc {

enum StorageId { ST_FP64 = 0, ST_FP32 = 1, ST_FP16S = 2, ST_BF16S = 3 };

struct StFP64 {
    using S = double;
    using C = double;
    static constexpr bool shifted = false;
    static constexpr int id = ST_FP64;
    __device__ static inline C load(S v) { return v; }
    __device__ static inline S store(C v) { return v; }
};

struct StFP32 {
    using S = float;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr int id = ST_FP32;
    __device__ static inline C load(S v) { return v; }
    __device__ static inline S store(C v) { return v; }
};

struct StFP16S {
    using S = _Float16;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr int id = ST_FP16S;
    __device__ static inline C load(S v) { return (float)v; }
    __device__ static inline S store(C v) { return (_Float16)v; }  // round to nearest even
};

struct StBF16S {
    using S = __bf16;
    using C = float;
    static constexpr bool shifted = true;
    static constexpr int id = ST_BF16S;
    __device__ static inline C load(S v) { return (float)v; }
    __device__ static inline S store(C v) { return (__bf16)v; }  // round to nearest even
};

}  // namespace lbmc
