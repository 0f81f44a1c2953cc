// Kernels of the batched step (batch.hip): members' arguments in device
// memory, a persistent grid over 256-node chunks of all members, and the
// per-member advance from call t to t + 1.
#pragma once

#include <hip/hip_runtime.h>

#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmc {

// collision variant of a member: bit 2 MRT family, bit 1 Guo forcing, bit 0
// Smagorinsky (BGK family) or corrected D3Q15 M^-1 (MRT family)
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

template <class L, class St>
__global__ __launch_bounds__(kBlock) void batch_step_kernel(const BatchMember<typename St::C> *members,
                                                            const int2 *chunks, int nchunks)
{
    for (int c = (int)blockIdx.x; c < nchunks; c += (int)gridDim.x) {
        const int2 ch = chunks[c];  // (member, first node)
        const BatchMember<typename St::C> &m = members[ch.x];
        const StepArgs<typename St::C> &a = m.a;
        const int variant = m.variant;
        const int n = ch.y + (int)threadIdx.x;
        if (n < a.nnodes) {
            int dd[3][3];
            slot_offsets<L>(a, n, dd);
            batch_node<L, St>(a, variant, n, dd);
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

template <class L, class St>
hipError_t launch_batch_lattice(const void *members, const int2 *chunks, int nchunks, int grid, hipStream_t s)
{
    hipLaunchKernelGGL((batch_step_kernel<L, St>), dim3(grid), dim3(kBlock), 0, s,
                       (const BatchMember<typename St::C> *)members, chunks, nchunks);
    return hipGetLastError();
}

template <class St>
hipError_t launch_batch_st(int lattice, const void *members, const int2 *chunks, int nchunks, int grid,
                           hipStream_t s)
{
    return lattice == 0 ? launch_batch_lattice<D2Q9, St>(members, chunks, nchunks, grid, s)
                                : launch_batch_lattice<D3Q15, St>(members, chunks, nchunks, grid, s);
}

}  // namespace lbmc
