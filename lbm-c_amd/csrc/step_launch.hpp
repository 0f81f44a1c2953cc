// Variant selection for the fused step: one instantiation per lattice,
// collision family, Guo forcing, Smagorinsky and D3Q15 M^-1 choice, plus the
// learned-closure step (FP32-compute storages).
#pragma once

#include <type_traits>

#include "closure_kernel.hpp"
#include "dispatch.hpp"
#include "step_kernel.hpp"

namespace lbmc {

// lds: dynamic LDS bytes per block (caps the resident blocks per CU)
template <class L, class St, int FAM, bool GUO, bool SMAG, bool MFIX>
static hipError_t launch_one(const StepArgs<typename St::C> &a, int grid, hipStream_t s, bool wide, int lds,
                             bool pairs)
{
    if constexpr (sizeof(typename St::S) == 2) {
        // half-interleaved layout: two nodes per lane on every call
        if (pairs && !wide && a.hs != 0) {
            const int hgrid = (int)row_grid((a.hs + kBlock - 1) / kBlock, a.rlog);
            hipLaunchKernelGGL((step_half_kernel<L, St, FAM, GUO, SMAG, MFIX>), dim3(hgrid), dim3(kBlock), lds, s,
                               a);
            return hipGetLastError();
        }
        // plain layout, even x shift on an even-length row: two nodes per lane
        if (pairs && !wide && a.hs == 0 && (a.L[0] & 1) == 0 && (a.s[0] & 1) == 0) {
            const int pgrid = (a.nnodes / 2 + kBlock - 1) / kBlock;
            hipLaunchKernelGGL((step_pair_kernel<L, St, FAM, GUO, SMAG, MFIX>), dim3(pgrid), dim3(kBlock), lds, s,
                               a);
            return hipGetLastError();
        }
    }
    if (wide)
        hipLaunchKernelGGL((step_kernel<L, St, FAM, GUO, SMAG, MFIX, true>), dim3(grid), dim3(kBlock), lds, s, a);
    else
        hipLaunchKernelGGL((step_kernel<L, St, FAM, GUO, SMAG, MFIX, false>), dim3(grid), dim3(kBlock), lds, s, a);
    return hipGetLastError();
}

// closure calls: tmode 0 evaluate, 1 evaluate and record tau, 2 reuse the
// record (grid: the persistent closure grid for 0/1, one block per 256 nodes
// for 2; lds: the step kernels' residency cap, applied to the reuse kernel,
// which has their memory pattern)
template <class L, class St, bool GUO, bool W, bool H>
static hipError_t launch_closure_w(const StepArgs<float> &a, int grid, hipStream_t s, int lds)
{
    if (a.tmode == 2)
        hipLaunchKernelGGL((closure_reuse_kernel<L, St, GUO, W>), dim3(grid), dim3(kBlock), lds, s, a);
    else if (a.tmode == 1)
        hipLaunchKernelGGL((closure_step_kernel<L, St, GUO, true, W, H>), dim3(grid), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((closure_step_kernel<L, St, GUO, false, W, H>), dim3(grid), dim3(kBlock), 0, s, a);
    return hipGetLastError();
}

// h: the network with fp16 operands (else bf16)
template <class L, class St, bool GUO>
static hipError_t launch_closure(const StepArgs<float> &a, int grid, hipStream_t s, bool wide, bool h, int lds)
{
    if (h)
        return wide ? launch_closure_w<L, St, GUO, true, true>(a, grid, s, lds)
                    : launch_closure_w<L, St, GUO, false, true>(a, grid, s, lds);
    return wide ? launch_closure_w<L, St, GUO, true, false>(a, grid, s, lds)
                : launch_closure_w<L, St, GUO, false, false>(a, grid, s, lds);
}

template <class L, class St>
static hipError_t launch_lattice(const KernelKey &k, const StepArgs<typename St::C> &a, int grid,
                                 hipStream_t s)
{
    if constexpr (std::is_same<typename St::C, float>::value) {
        if (k.closure && k.family == 0) {
            return k.guo ? launch_closure<L, St, true>(a, grid, s, k.wide != 0, k.closure_fp16 != 0, k.lds)
                         : launch_closure<L, St, false>(a, grid, s, k.wide != 0, k.closure_fp16 != 0, k.lds);
        }
    }
    if (k.family == 0) {
        if (k.guo) {
            return k.smag ? launch_one<L, St, 0, true, true, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0)
                          : launch_one<L, St, 0, true, false, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0);
        }
        return k.smag ? launch_one<L, St, 0, false, true, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0)
                      : launch_one<L, St, 0, false, false, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0);
    }
    if constexpr (L::D == 3) {
        if (k.mfix) {
            return k.guo ? launch_one<L, St, 1, true, false, true>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0)
                         : launch_one<L, St, 1, false, false, true>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0);
        }
    }
    return k.guo ? launch_one<L, St, 1, true, false, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0)
                 : launch_one<L, St, 1, false, false, false>(a, grid, s, k.wide != 0, k.lds, k.pairs != 0);
}

template <class St>
static hipError_t launch_step_st(const KernelKey &k, const StepArgs<typename St::C> &a, int grid,
                                 hipStream_t s)
{
    return k.lattice == 0 ? launch_lattice<D2Q9, St>(k, a, grid, s)
                          : launch_lattice<D3Q15, St>(k, a, grid, s);
}

template <class St>
static hipError_t launch_sf_st(int lattice, const StepArgs<typename St::C> &a, const int *sfnb,
                               int nsf, typename St::C *out, hipStream_t s)
{
    if (nsf <= 0) return hipSuccess;
    const int grid = (nsf + kBlock - 1) / kBlock;
    if (lattice == 0)
        hipLaunchKernelGGL((sf_gather_kernel<D2Q9, St>), dim3(grid), dim3(kBlock), 0, s, a, sfnb, nsf, out);
    else
        hipLaunchKernelGGL((sf_gather_kernel<D3Q15, St>), dim3(grid), dim3(kBlock), 0, s, a, sfnb, nsf, out);
    return hipGetLastError();
}

}  // namespace lbmc
