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
    int store_every;  // 0: never store
    int store_count;  // (t + 1) mod store_every for the current call t
    int pad;
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
// CLOSURE: some members use the learned closure (one shared network, staged
// into LDS once per block); their tiles run the network over the whole
// wave, lanes past the member's last node evaluating a copy of it.
template <class L, class St, bool CLOSURE>
__global__ __launch_bounds__(kBlock) void batch_step_kernel(const BatchMember<typename St::C> *__restrict__ members,
                                                            const int2 *__restrict__ chunks, int nchunks,
                                                            const void *__restrict__ mlp, int sub)
{
    using C = typename St::C;
    constexpr bool CL = CLOSURE && std::is_same<C, float>::value;
    __shared__ typename std::conditional<CL, MlpInfer, char>::type w;
    __shared__ mlp_v8b xb[CL ? kBlock : 1];
    if constexpr (CL) {
        const uint4 *src = (const uint4 *)mlp;
        uint4 *dst = (uint4 *)&w;
        for (int k = (int)threadIdx.x; k < (int)(sizeof(MlpInfer) / 16); k += kBlock) dst[k] = src[k];
        __syncthreads();
    }
    for (int c = (int)blockIdx.x; c < nchunks; c += (int)gridDim.x) {
        const int2 ch = chunks[c];  // (member, first node)
        const BatchMember<C> &m = members[ch.x];
        const StepArgs<C> &a = m.a;
        const int variant = m.variant;
        for (int j = 0; j < sub; j++) {
            const int t0 = ch.y + j * kBlock;  // first node of the tile
            if (t0 >= a.nnodes) break;         // block-uniform
            const int n0 = t0 + (int)threadIdx.x;
            if constexpr (CL) {
                if (variant >= 8) {
                    const bool live = n0 < a.nnodes;
                    const int n = live ? n0 : a.nnodes - 1;
                    int dd[3][3];
                    slot_offsets<L>(a, n, dd);
                    float v[L::Q];
                    gather_node<L, St, false>(a, n, dd, v);
                    if (variant == 9) closure_update<L, St, true, false>(a, n, live, v, w, xb + (threadIdx.x & ~63u));
                    else closure_update<L, St, false, false>(a, n, live, v, w, xb + (threadIdx.x & ~63u));
                    if (live) scatter_node<L, St, false>(a, n, dd, v);
                    continue;
                }
            }
            if (n0 < a.nnodes) {
                int dd[3][3];
                slot_offsets<L>(a, n0, dd);
                batch_node<L, St>(a, variant, n0, dd);
            }
        }
    }
}

// call t -> t + 1 for every member (one thread each)
template <class C>
__global__ void batch_advance_kernel(BatchMember<C> *members, int count)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= count) return;
    BatchMember<C> m = members[k];
    long long stride = 1;
    for (int d = 0; d < 3; d++) {
        const int L = m.a.L[d];
        int s = m.a.s[d] + 1;
        if (s == L) s = 0;
        m.a.s[d] = s;
        m.a.dp[d] = (int)(-(long long)s * stride);
        m.a.dpw[d] = (int)((long long)(L - s) * stride);
        m.a.dm[d] = (int)((long long)s * stride);
        m.a.dmw[d] = (int)((long long)(s - L) * stride);
        stride *= L;
    }
    if (m.store_every > 0) {
        int sc = m.store_count + 1;
        if (sc == m.store_every) sc = 0;
        m.store_count = sc;
        m.a.store = sc == 0 ? 1 : 0;
    }
    members[k] = m;
}

// Launch shape of the batch step: grid blocks, `sub` 256-node tiles per
// chunk, and lds: dynamic LDS bytes per block (unused by the kernel; caps
// the blocks resident per CU)
struct BatchShape {
    int grid, sub, lds;
};

template <class L, class St>
hipError_t launch_batch_lattice(const void *members, const int2 *chunks, int nchunks, BatchShape sh,
                                const void *mlp, hipStream_t s)
{
    const auto *mm = (const BatchMember<typename St::C> *)members;
    if (mlp)
        hipLaunchKernelGGL((batch_step_kernel<L, St, true>), dim3(sh.grid), dim3(kBlock), sh.lds, s, mm, chunks,
                           nchunks, mlp, sh.sub);
    else
        hipLaunchKernelGGL((batch_step_kernel<L, St, false>), dim3(sh.grid), dim3(kBlock), sh.lds, s, mm, chunks,
                           nchunks, mlp, sh.sub);
    return hipGetLastError();
}

// mlp: the shared closure weights when some member uses the closure, else null
template <class St>
hipError_t launch_batch_st(int lattice, const void *members, const int2 *chunks, int nchunks, BatchShape sh,
                           const void *mlp, hipStream_t s)
{
    return lattice == 0 ? launch_batch_lattice<D2Q9, St>(members, chunks, nchunks, sh, mlp, s)
                        : launch_batch_lattice<D3Q15, St>(members, chunks, nchunks, sh, mlp, s);
}

}  // namespace lbmc
