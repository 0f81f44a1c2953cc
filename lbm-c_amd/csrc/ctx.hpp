// The lattice context behind the C ABI (include/lbmc.h), shared by the
// single-lattice entry points (lbmc_api.cpp) and the batch (batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "../../include/lbmc.h"
#include "step_args.hpp"

struct lbmc_ctx {
    lbmc_desc desc{};
    int D = 0, Q = 0;
    int L[3] = {1, 1, 1};
    long long N = 0, npad = 0;
    int storage = 0;
    size_t sbytes = 0;  // bytes per population slot
    size_t cbytes = 0;  // bytes per compute-precision value
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    void *f = nullptr;
    uint8_t *mask = nullptr;
    uint16_t *cls = nullptr;
    void *rho = nullptr;
    void *u[3] = {nullptr, nullptr, nullptr};
    void *porosity = nullptr;
    void *force[3] = {nullptr, nullptr, nullptr};
    int *sfidx = nullptr;
    int *sfnb = nullptr;
    void *sfbuf = nullptr;
    int nsf = 0;
    double *partials = nullptr;
    int npartials = 0;
    double *scalar = nullptr;

    bool fields_set = false;
    bool porous = false;
    bool all_fluid = false;  // every node plain fluid (mask all 0): the step skips the mask
    bool offsets64 = false;  // a direction array spans 4 GiB or more: only 64-bit offsets reach it
                             //   (key.wide may also be set below that, by choice)
    long long hs = 0;        // half-interleaved 16-bit layout: N / 2 (StepArgs::hs), else 0
    unsigned rlog = 0;       // row_chunk block order of the narrow step kernels (StepArgs::rlog)
    double F[3] = {0, 0, 0};
    int ncls = 1;
    uint16_t cls_host[256] = {};
    lbmc::KernelKey key{};

    // learned closure: operand image (device), on when key.closure; with an
    // evaluation interval K > 1 the network runs on calls t % K == 0 (and the
    // first call after any change of state) and the other calls reuse the
    // recorded per-node tau
    void *mlp_img = nullptr;
    const lbmc_mlp *closure_model = nullptr;  // the model the image was built from
    double closure_lo = 0, closure_hi = 0;
    int closure_every = 1;
    double closure_shift = 0;  // added to the network's tau' on this lattice
    float *tcache = nullptr;
    bool tcache_valid = false;

    // launch shape
    int grid = 1;
    int closure_grid = 1;  // blocks of the grid-stride closure kernel
    int reuse_grid = 1;    // blocks of the closure-reuse kernel (one per 256 nodes)
    lbmc::FastDiv dx{}, dy{};

    // captured step sequence: `period` calls starting at t = phase (mod
    // period) with the given store pattern; replayed while valid
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    long long graph_period = 0, graph_phase = -1;
    int graph_store = 0;
    int use_graph = 1;

    std::string err;
};

namespace lbmc {

// Kernel arguments of call t (the reference's iterate_kernel(..., t)).
template <class C>
inline void fill_args(const lbmc_ctx *c, StepArgs<C> &a, long long t, int store)
{
    memset(&a, 0, sizeof(a));
    a.f = c->f;
    a.npad = c->npad;
    a.mask = c->mask;
    a.cls = c->cls;
    a.porosity = (const C *)c->porosity;
    for (int d = 0; d < 3; d++) {
        a.force[d] = (const C *)c->force[d];
        a.F[d] = (C)c->F[d];
        a.u[d] = (C *)c->u[d];
        a.L[d] = c->L[d];
        const long long r = t % c->L[d];
        a.s[d] = (int)(r < 0 ? r + c->L[d] : r);
    }
    a.rho = (C *)c->rho;
    a.sfbuf = (const C *)c->sfbuf;
    a.sfidx = c->sfidx;
    {
        long long stride = 1;
        for (int d = 0; d < 3; d++) {
            const long long L = c->L[d], sd = a.s[d];
            a.dp[d] = (int)(-sd * stride);
            a.dpw[d] = (int)((L - sd) * stride);
            a.dm[d] = (int)(sd * stride);
            a.dmw[d] = (int)((sd - L) * stride);
            stride *= L;
        }
    }
    a.nnodes = (int)c->N;
    a.mx = c->dx.m;
    a.lx = c->dx.l;
    a.my = c->dy.m;
    a.ly = c->dy.l;
    a.wide_rows = (c->key.wide && c->L[0] % 64 == 0) ? 1 : 0;
    a.hs = (unsigned)c->hs;
    a.rlog = c->rlog;
    a.tau = (C)c->desc.tau;
    a.c_smag = (C)c->desc.c_smag;
    for (int i = 0; i < 15; i++) a.tau_mrt[i] = (C)c->desc.tau_mrt[i];
    a.store = store ? 1 : 0;
    a.porous = c->porous ? 1 : 0;
    a.all_fluid = c->all_fluid ? 1 : 0;
    a.quirks = c->desc.quirks;
    a.mlp = c->mlp_img;
    a.tcache = c->tcache;
    a.tmode = 0;
    a.closure_shift = (C)c->closure_shift;
}

}  // namespace lbmc
