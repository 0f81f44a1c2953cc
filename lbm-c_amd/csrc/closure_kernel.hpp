// Fused step with the learned relaxation-time closure (mlp.hpp) in place of
// the Smagorinsky hook of bgk_collision (LBM-C/collision.cu:60-62,445-463).
//
// Same gather / update / scatter as step_kernel, with three differences:
//  * the network's weights (9.7 KB of operands) are staged once per block
//    into LDS, so blocks loop over 256-node chunks (grid-stride) instead of
//    one block per chunk;
//  * every node of a wave evaluates the network together, in uniform control
//    flow (the matrix-core operands span the wave); lanes past the last node
//    evaluate a copy of the last node and write nothing;
//  * plain fluid nodes relax with tau' = out_a * y + out_b + shift (the
//    lattice's shift, 0 by default), or with the Smagorinsky tau where tau'
//    leaves [lo, hi] (tau when c_smag = 0);
//    boundary and porous nodes take the reference-order update with the
//    Smagorinsky hook.
#pragma once

#include "mlp.hpp"
#include "node_index.hpp"
#include "step_kernel.hpp"

namespace lbmc {

// H: the network with fp16 operands (else bf16)
template <class L, class St, bool GUO, bool RECORD, bool H>
__device__ __forceinline__ void closure_update(const StepArgs<float> &a, int n, bool live, float (&v)[L::Q],
                                               const MlpEval &w, uint4 *xw)
{
    constexpr int D = L::D;
    const unsigned m = a.all_fluid ? 0u : ldg(a.mask + n);
    float F[3] = {0.f, 0.f, 0.f};
    if constexpr (GUO) {
        sfor<0, D>([&](auto dd) {
            constexpr int d = decltype(dd)::value;
            F[d] = a.force[0] ? ldg(a.force[d] + n) : a.F[d];
        });
    }
    FastState<L> st;
    fast_prepare<L, St::shifted, GUO>(v, F, st);
    float x[7];
    fast_closure_features<L, St::shifted>(st, x);
    const float y = mlp_wave_eval<H>(w, xw, mlp_input_eval<H>(w.sc, x));
    if (!live) return;
    if (m == 0) {
        float t = w.sc.out_a * y + w.sc.out_b + a.closure_shift;
        const bool in_band = t >= w.sc.lo && t <= w.sc.hi;
        if constexpr (RECORD) stg(a.tcache + n, in_band ? t : __builtin_nanf(""));
        if (!in_band) t = a.c_smag > 0.f ? fast_smag_tau<L, St::shifted>(st, a.tau, a.c_smag) : a.tau;
        fast_relax<L, St::shifted, GUO>(v, st, t, a.tau, F);
        if (a.store) {
            sfor<0, D>([&](auto dd) {
                constexpr int d = decltype(dd)::value;
                stg(a.u[d] + n, st.u[d]);
            });
            stg(a.rho + n, st.rho);
        }
    } else {
        node_update<L, St, 0, GUO, true, false>(a, n, v);
    }
}

// A call between evaluations, one node: plain fluid nodes relax with the
// recorded tau (the Smagorinsky tau of the current state where the record is
// NaN); no network.
template <class L, class St, bool GUO>
__device__ __forceinline__ void closure_reuse_node(const StepArgs<float> &a, int n, float (&v)[L::Q])
{
    if (a.all_fluid || ldg(a.mask + n) == 0) {
        constexpr int D = L::D;
        float F[3] = {0.f, 0.f, 0.f};
        if constexpr (GUO) {
            sfor<0, D>([&](auto dd2) {
                constexpr int d = decltype(dd2)::value;
                F[d] = a.force[0] ? ldg(a.force[d] + n) : a.F[d];
            });
        }
        FastState<L> st;
        fast_prepare<L, St::shifted, GUO>(v, F, st);
        float t = ldg(a.tcache + n);
        if (t != t) t = a.c_smag > 0.f ? fast_smag_tau<L, St::shifted>(st, a.tau, a.c_smag) : a.tau;
        fast_relax<L, St::shifted, GUO>(v, st, t, a.tau, F);
        if (a.store) {
            sfor<0, D>([&](auto dd2) {
                constexpr int d = decltype(dd2)::value;
                stg(a.u[d] + n, st.u[d]);
            });
            stg(a.rho + n, st.rho);
        }
    } else {
        node_update<L, St, 0, GUO, true, false>(a, n, v);
    }
}

template <class L, class St, bool GUO, bool RECORD, bool WIDE, bool H>
__global__ __launch_bounds__(kBlock, 4) void closure_step_kernel(const StepArgs<float> a)
{
    __shared__ MlpEval w;
    __shared__ uint4 xb[kBlock];
    {
        const uint4 *src = (const uint4 *)eval_image<H>(a.mlp);
        uint4 *dst = (uint4 *)&w;
        for (int k = (int)threadIdx.x; k < (int)(sizeof(MlpEval) / 16); k += kBlock) dst[k] = src[k];
    }
    __syncthreads();
    uint4 *xw = xb + (threadIdx.x & ~63u);
    // chunks in row_chunk order (the step kernels' XCD-aware order); the
    // grid is a multiple of 8 << rlog, so position p runs on XCD p mod 8
    const unsigned nchunks = tile_count(a.nnodes, kBlock);
    const unsigned npos = row_grid(nchunks, a.rlog);
    for (unsigned p = blockIdx.x; p < npos; p += gridDim.x) {
        const unsigned b = row_chunk(p, a.rlog);
        if (b >= nchunks) continue;  // block-uniform
        const unsigned n0 = tile_node(0u, b, kBlock, threadIdx.x);
        const bool live = n0 < (unsigned)a.nnodes;
        const int n = live ? (int)n0 : a.nnodes - 1;
        int dd[3][3];
        slot_offsets<L>(a, n, dd);
        float v[L::Q];
        gather_node<L, St, WIDE>(a, n, dd, v);
        closure_update<L, St, GUO, RECORD, H>(a, n, live, v, w, xw);
        if (live) scatter_node<L, St, WIDE>(a, n, dd, v);
    }
}

// A call between evaluations: closure_reuse_node at every node, one thread
// per node as in step_kernel; 4 more bytes per node (the record).
template <class L, class St, bool GUO, bool WIDE>
__global__ __launch_bounds__(kBlock) void closure_reuse_kernel(const StepArgs<float> a)
{
    const unsigned nu = row_chunk(blockIdx.x, a.rlog) * (unsigned)kBlock + threadIdx.x;
    if (nu >= (unsigned)a.nnodes) return;
    const int n = (int)nu;
    int dd[3][3];
    slot_offsets<L>(a, n, dd);
    float v[L::Q];
    gather_node<L, St, WIDE>(a, n, dd, v);
    closure_reuse_node<L, St, GUO>(a, n, v);
    scatter_node<L, St, WIDE>(a, n, dd, v);
}

}  // namespace lbmc
