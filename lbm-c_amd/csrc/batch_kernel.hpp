// Kernels of the batched step (batch.hip): members' arguments in device
// memory, a persistent grid over 256-node chunks of all members, and the
// per-member advance from call t to t + 1.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "closure_kernel.hpp"
#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmc {

// collision variant of a member: bit 2 MRT family, bit 1 Guo forcing, bit 0
// Smagorinsky (BGK family) or corrected D3Q15 M^-1 (MRT family); 8 and 9 are
// BGK-family members with the learned closure (9: with Guo forcing)
template <class C>
struct BatchMember {
    StepArgs<C> a;
    int variant;
    int store_every;    // 0: never store
    int store_count;    // (t + 1) mod store_every for the current call t
    int closure_every;  // closure members: the network runs on calls t with
    int closure_phase;  //   t mod closure_every == 0 (phase = t mod every);
                        //   a.tmode: 0 every call, 1 evaluate and record,
                        //   2 reuse the record
};

template <class L, class St>
__device__ __forceinline__ void batch_node(const StepArgs<typename St::C> &a, int variant, int n,
                                           const int (&dd)[3][3])
{
    switch (variant) {
    case 0: step_node<L, St, 0, false, false, false, false>(a, n, dd); break;
    case 1: step_node<L, St, 0, false, true, false, false>(a, n, dd); break;
    case 2: step_node<L, St, 0, true, false, false, false>(a, n, dd); break;
    case 3: step_node<L, St, 0, true, true, false, false>(a, n, dd); break;
    case 4: step_node<L, St, 1, false, false, false, false>(a, n, dd); break;
    case 6: step_node<L, St, 1, true, false, false, false>(a, n, dd); break;
    default:
        if constexpr (L::D == 3) {
            if (variant == 5) step_node<L, St, 1, false, false, true, false>(a, n, dd);
            else step_node<L, St, 1, true, false, true, false>(a, n, dd);
        }
        break;
    }
}

// A chunk is `sub` consecutive 256-node tiles of one member, starting at
// node ch.y: the block reads the chunk entry and the member's arguments once
// (a chain of dependent loads) and then walks its tiles.  members and
// chunks are read-only for the kernel's lifetime and reached through no
// other pointer (__restrict__), so their loads are scalar and stay in scalar
// registers across the population stores.
//
// A call is two kernels over disjoint sets of nodes.  batch_step_kernel
// takes every chunk except those of closure members that evaluate the
// network on this call: plain members, and closure members between
// evaluations (recorded tau, no network).  batch_eval_kernel takes the
// closure members' chunks and runs the network on those that evaluate; it
// needs the weights in LDS and the matrix-core registers, which the first
// kernel therefore does not carry.

__device__ __forceinline__ bool evaluates(int variant, int tmode) { return variant >= 8 && tmode != 2; }

// Calls per window: the members' arguments of kBatchWindow consecutive calls
// are laid out in device memory (batch_window_kernel), and the calls of a
// window launch the step kernels on their own entry, so the members' state
// moves on (batch_advance_kernel) once per window.
constexpr int kBatchWindow = 32;

// Arguments of call t + j from the member's state for call t: the shift
// counters and slot offsets, the macro-store flag and the closure mode that
// j single-call advances leave (j <= 2 kBatchWindow: the loops below run at
// most that often).
template <class C>
__device__ __forceinline__ void call_args(const BatchMember<C> &m, int j, StepArgs<C> &a)
{
    long long stride = 1;
    for (int d = 0; d < 3; d++) {
        const int L = a.L[d];
        int s = a.s[d] + j;
        while (s >= L) s -= L;
        a.s[d] = s;
        a.dp[d] = (int)(-(long long)s * stride);
        a.dpw[d] = (int)((long long)(L - s) * stride);
        a.dm[d] = (int)((long long)s * stride);
        a.dmw[d] = (int)((long long)(s - L) * stride);
        stride *= L;
    }
    if (m.store_every > 0) {
        int sc = m.store_count + j;
        while (sc >= m.store_every) sc -= m.store_every;
        a.store = sc == 0 ? 1 : 0;
    }
    if (m.closure_every > 1) {
        int ph = m.closure_phase + j;
        while (ph >= m.closure_every) ph -= m.closure_every;
        a.tmode = ph == 0 ? 1 : 2;
    }
}

template <class L, class St>
__global__ __launch_bounds__(kBlock) void batch_step_kernel(const BatchMember<typename St::C> *__restrict__ members,
                                                            const int2 *__restrict__ chunks, int nchunks, int sub)
{
    using C = typename St::C;
    constexpr bool CL = std::is_same<C, float>::value;
    for (int c = (int)blockIdx.x; c < nchunks; c += (int)gridDim.x) {
        const int2 ch = chunks[c];  // (member, first node); member -1: padding of the XCD order
        if (ch.x < 0) continue;
        const BatchMember<C> &m = members[ch.x];
        const StepArgs<C> &a = m.a;
        const int variant = m.variant;
        if (evaluates(variant, a.tmode)) continue;  // batch_eval_kernel's
        for (int j = 0; j < sub; j++) {
            const unsigned t0 = tile_node((unsigned)ch.y, (unsigned)j, kBlock, 0u);  // first node of the tile
            if (t0 >= (unsigned)a.nnodes) break;                                  // block-uniform
            if (t0 + threadIdx.x >= (unsigned)a.nnodes) continue;
            const int n0 = (int)(t0 + threadIdx.x);
            int dd[3][3];
            slot_offsets<L>(a, n0, dd);
            if constexpr (CL) {
                if (variant >= 8) {
                    // between evaluations: the recorded tau, no network
                    float v[L::Q];
                    gather_node<L, St, false>(a, n0, dd, v);
                    if (variant == 9) closure_reuse_node<L, St, true>(a, n0, v);
                    else closure_reuse_node<L, St, false>(a, n0, v);
                    scatter_node<L, St, false>(a, n0, dd, v);
                    continue;
                }
            }
            batch_node<L, St>(a, variant, n0, dd);
        }
    }
}

// Closure members' chunks (their own list) on a call where at least one of
// them evaluates (*any_eval, written by batch_window_kernel): the network
// over whole waves, lanes past the member's last node evaluating a copy of
// it and writing nothing.  The shared weights are staged into LDS once per
// block.
template <class L, class St, bool H>
__global__ __launch_bounds__(kBlock, 4) void batch_eval_kernel(const BatchMember<float> *__restrict__ members,
                                                               const int2 *__restrict__ chunks, int nchunks,
                                                               const void *__restrict__ mlp, int sub,
                                                               const int *__restrict__ any_eval)
{
    if (*any_eval == 0) return;
    __shared__ MlpEval w;
    __shared__ uint4 xb[kBlock];
    {
        const uint4 *src = (const uint4 *)eval_image<H>(mlp);
        uint4 *dst = (uint4 *)&w;
        for (int k = (int)threadIdx.x; k < (int)(sizeof(MlpEval) / 16); k += kBlock) dst[k] = src[k];
    }
    __syncthreads();
    uint4 *xw = xb + (threadIdx.x & ~63u);
    for (int c = (int)blockIdx.x; c < nchunks; c += (int)gridDim.x) {
        const int2 ch = chunks[c];
        if (ch.x < 0) continue;
        const BatchMember<float> &m = members[ch.x];
        const StepArgs<float> &a = m.a;
        const int variant = m.variant;
        if (!evaluates(variant, a.tmode)) continue;  // block-uniform
        for (int j = 0; j < sub; j++) {
            const unsigned t0 = tile_node((unsigned)ch.y, (unsigned)j, kBlock, 0u);
            if (t0 >= (unsigned)a.nnodes) break;
            const unsigned n0 = t0 + threadIdx.x;
            const bool live = n0 < (unsigned)a.nnodes;
            const int n = live ? (int)n0 : a.nnodes - 1;
            int dd[3][3];
            slot_offsets<L>(a, n, dd);
            float v[L::Q];
            gather_node<L, St, false>(a, n, dd, v);
            if (a.tmode == 1) {
                if (variant == 9) closure_update<L, St, true, true, H>(a, n, live, v, w, xw);
                else closure_update<L, St, false, true, H>(a, n, live, v, w, xw);
            } else {
                if (variant == 9) closure_update<L, St, true, false, H>(a, n, live, v, w, xw);
                else closure_update<L, St, false, false, H>(a, n, live, v, w, xw);
            }
            if (live) scatter_node<L, St, false>(a, n, dd, v);
        }
    }
}

// state: call t -> t + calls for every member (one block, a thread per
// member; calls <= kBatchWindow)
template <class C>
__global__ __launch_bounds__(256) void batch_advance_kernel(BatchMember<C> *state, int count, int calls)
{
    for (int k = (int)threadIdx.x; k < count; k += (int)blockDim.x) {
        BatchMember<C> m = state[k];
        call_args(m, calls, m.a);
        if (m.store_every > 0) {
            int sc = m.store_count + calls;
            while (sc >= m.store_every) sc -= m.store_every;
            m.store_count = sc;
        }
        if (m.closure_every > 1) {
            int ph = m.closure_phase + calls;
            while (ph >= m.closure_every) ph -= m.closure_every;
            m.closure_phase = ph;
        }
        state[k] = m;
    }
}

// window[q][k]: member k's arguments of call t + q (block q; the state
// describes call t), and any_eval[q]: whether some closure member
// evaluates the network on that call
template <class C>
__global__ __launch_bounds__(256) void batch_window_kernel(const BatchMember<C> *state, int count,
                                                           BatchMember<C> *window, int *any_eval)
{
    const int q = (int)blockIdx.x;
    int ev = 0;
    for (int k = (int)threadIdx.x; k < count; k += (int)blockDim.x) {
        BatchMember<C> m = state[k];
        if (q != 0) call_args(m, q, m.a);
        window[(size_t)q * (size_t)count + (size_t)k] = m;
        if (evaluates(m.variant, m.a.tmode)) ev = 1;
    }
    ev = __syncthreads_or(ev);
    if (threadIdx.x == 0) any_eval[q] = ev;
}

// Launch shape of the batch step: grid blocks, `sub` 256-node tiles per
// chunk, and lds: dynamic LDS bytes per block (unused by the kernel; caps
// the blocks resident per CU)
struct BatchShape {
    int grid, sub, lds;
};

// The closure members' evaluation: their chunk list (nchunks 0: no closure
// members), its grid, the shared weights, the device flag and the operand
// precision (fp16: true, else bf16)
struct BatchEval {
    const int2 *chunks;
    int nchunks, grid;
    const void *mlp;
    const int *any_eval;
    bool fp16;
};

template <class L, class St>
hipError_t launch_batch_lattice(const void *members, const int2 *chunks, int nchunks, BatchShape sh,
                                const BatchEval &ev, hipStream_t s)
{
    using C = typename St::C;
    hipLaunchKernelGGL((batch_step_kernel<L, St>), dim3(sh.grid), dim3(kBlock), sh.lds, s,
                       (const BatchMember<C> *)members, chunks, nchunks, sh.sub);
    if constexpr (std::is_same<C, float>::value) {
        if (ev.nchunks > 0) {
            if (ev.fp16)
                hipLaunchKernelGGL((batch_eval_kernel<L, St, true>), dim3(ev.grid), dim3(kBlock), 0, s,
                                   (const BatchMember<float> *)members, ev.chunks, ev.nchunks, ev.mlp, sh.sub,
                                   ev.any_eval);
            else
                hipLaunchKernelGGL((batch_eval_kernel<L, St, false>), dim3(ev.grid), dim3(kBlock), 0, s,
                                   (const BatchMember<float> *)members, ev.chunks, ev.nchunks, ev.mlp, sh.sub,
                                   ev.any_eval);
        }
    }
    return hipGetLastError();
}

template <class St>
hipError_t launch_batch_st(int lattice, const void *members, const int2 *chunks, int nchunks, BatchShape sh,
                           const BatchEval &ev, hipStream_t s)
{
    return lattice == 0 ? launch_batch_lattice<D2Q9, St>(members, chunks, nchunks, sh, ev, s)
                        : launch_batch_lattice<D3Q15, St>(members, chunks, nchunks, sh, ev, s);
}

}  // namespace lbmc
