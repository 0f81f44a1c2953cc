// C ABI of the time-step library (declared in include/lbmc.h).
//
// Host responsibilities: descriptor validation, device allocation, building
// the 1-byte node mask and class table from the reference's per-node fields,
// choosing the kernel variant and launch shape, and the read-backs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbmc.h"
#include "dispatch.hpp"
#include "lattice.hpp"
#include "node_index.hpp"
#include "step_args.hpp"
#include "storage.hpp"
#include "ctx.hpp"

using namespace lbmc;


namespace {

std::string g_last_error;

lbmc_status fail(lbmc_ctx *c, lbmc_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_last_error = buf;
    return st;
}

lbmc_status hip_fail(lbmc_ctx *c, hipError_t e, const char *what)
{
    return fail(c, e == hipErrorOutOfMemory ? LBMC_ERR_NOMEM : LBMC_ERR_DEVICE, "%s: %s", what,
                hipGetErrorString(e));
}

#define HIP_TRY(c, expr, what)                      \
    do {                                            \
        hipError_t e_ = (expr);                     \
        if (e_ != hipSuccess) return hip_fail(c, e_, what); \
    } while (0)

bool micro_valid(int lattice, int code)
{
    if (code == 0) return true;
    if (lattice == LBMC_D2Q9) return code == 1 || code == 2 || (code >= 5 && code <= 8);
    return code >= 1 && code <= 12;
}

bool micro_is_sf(int lattice, int code)
{
    if (lattice == LBMC_D2Q9) return code >= 5 && code <= 8;
    return code >= 7 && code <= 12;
}

// (axis, side) of an sf code's face; side +1 = low face (copy from +1 neighbour)
void sf_face(int lattice, int code, int *axis, int *side)
{
    const int k = lattice == LBMC_D2Q9 ? code - 5 : code - 7;
    *axis = k / 2;
    *side = (k % 2 == 0) ? 1 : -1;
}

template <class T>
lbmc_status upload_as(lbmc_ctx *c, void *dst, const std::vector<double> &src)
{
    std::vector<T> tmp(src.size());
    for (size_t k = 0; k < src.size(); k++) tmp[k] = (T)src[k];
    HIP_TRY(c, hipMemcpy(dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice), "upload");
    return LBMC_OK;
}

lbmc_status upload_c(lbmc_ctx *c, void *dst, const std::vector<double> &src)
{
    return c->cbytes == 8 ? upload_as<double>(c, dst, src) : upload_as<float>(c, dst, src);
}

void free_dev(void *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

lbmc_status launch_step(lbmc_ctx *c, long long t, int store)
{
    hipError_t e = hipSuccess;
    if (c->cbytes == 8) {
        StepArgs<double> a;
        fill_args(c, a, t, store);
        if (c->nsf) {
            e = launch_sf_gather_fp64(c->desc.lattice, a, c->sfnb, c->nsf, (double *)c->sfbuf, c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "sf gather launch");
        }
        e = launch_step_fp64(c->key, a, c->grid, c->stream);
    } else {
        StepArgs<float> a;
        fill_args(c, a, t, store);
        if (c->nsf) {
            switch (c->storage) {
            case LBMC_FP32: e = launch_sf_gather_fp32(c->desc.lattice, a, c->sfnb, c->nsf, (float *)c->sfbuf, c->stream); break;
            case LBMC_FP16S: e = launch_sf_gather_fp16s(c->desc.lattice, a, c->sfnb, c->nsf, (float *)c->sfbuf, c->stream); break;
            default: e = launch_sf_gather_bf16s(c->desc.lattice, a, c->sfnb, c->nsf, (float *)c->sfbuf, c->stream); break;
            }
            if (e != hipSuccess) return hip_fail(c, e, "sf gather launch");
        }
        int grid = c->grid;
        if (c->key.closure) {
            if (c->closure_every > 1) {
                a.tmode = (!c->tcache_valid || t % c->closure_every == 0) ? 1 : 2;
                if (a.tmode == 1) c->tcache_valid = true;
            }
            grid = a.tmode != 2 ? c->closure_grid : c->reuse_grid;
        }
        switch (c->storage) {
        case LBMC_FP32: e = launch_step_fp32(c->key, a, grid, c->stream); break;
        case LBMC_FP16S: e = launch_step_fp16s(c->key, a, grid, c->stream); break;
        default: e = launch_step_bf16s(c->key, a, grid, c->stream); break;
        }
    }
    if (e != hipSuccess) return hip_fail(c, e, "step kernel launch");
    return LBMC_OK;
}

// Resident blocks per CU of the step kernels (0: no cap), the best of
// sweeps over 0/3/4/5/6/8 on 256^3 D3Q15 and 4096^2 D2Q9 lattices:
// D3Q15 FP32 3 (BGK family) or 4 (MRT family); D2Q9 FP32 4; FP16S/BF16S
// none (their loads run at raised wave priority instead, and a cap of 5 then
// costs 2.5 %); FP64 no gain either way.  Only lattices well beyond the
// 256 MB last-level cache are capped: a cache-resident lattice is
// latency-bound and wants every wave it can get.
int default_blocks_per_cu(const lbmc_ctx *c)
{
    const double pop_bytes = (double)c->sbytes * c->Q * c->npad;
    if (pop_bytes < 512.0 * (1 << 20)) return 0;
    const bool d3 = c->desc.lattice == LBMC_D3Q15;
    switch (c->storage) {
    case LBMC_FP32: return d3 ? (c->key.family == 0 ? 3 : 4) : 4;
    default: return 0;
    }
}

// Launch shape: one thread per node, 256-thread blocks.
void choose_shape(lbmc_ctx *c)
{
    c->grid = (int)((c->N + kBlock - 1) / kBlock);
    // rows of R = 2^k >= 2 chunks: min(R, 4) consecutive chunks per XCD
    // (row_chunk); groups of 4 measured best on 4096^2 D2Q9 (2, 8 and whole
    // 16-chunk rows slower), whole rows on 512^3 D3Q15 (+5 %)
    c->rlog = 0;
    const char *ro = getenv("LBMC_ROW_ORDER");
    if (c->L[0] % (2 * kBlock) == 0 && !(ro && atoi(ro) == 0)) {
        const int R = c->L[0] / kBlock;
        if ((R & (R - 1)) == 0)
            while ((1 << c->rlog) < R && c->rlog < 2) c->rlog++;
    }
    c->grid = (int)row_grid((unsigned)c->grid, c->rlog);
    c->dx = make_fastdiv(c->L[0]);
    c->dy = make_fastdiv(c->L[1]);
    // the closure kernel stages its weights once per block and loops over
    // 256-node chunks; 8192 blocks (8 per resident slot at 4 blocks per CU)
    // measured 2 % faster than 2048 or 1536 at 256^3, the same as 4096
    int cap = 8192;
    if (const char *env = getenv("LBMC_CLOSURE_GRID"))
        if (atoi(env) > 0) cap = atoi(env);
    const int chunks = (int)((c->N + kBlock - 1) / kBlock);
    c->closure_grid = std::min((int)row_grid((unsigned)chunks, c->rlog), cap);
    c->reuse_grid = (int)row_grid((unsigned)chunks, c->rlog);
}

void drop_graph(lbmc_ctx *c)
{
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    c->graph_phase = -1;
}

// the closure's recorded tau belongs to the state it was computed from
void invalidate_tau_record(lbmc_ctx *c)
{
    if (c->tcache_valid) {
        c->tcache_valid = false;
        drop_graph(c);
    }
}

long long lcm_ll(long long a, long long b)
{
    long long x = a, y = b;
    while (y) {
        const long long r = x % y;
        x = y;
        y = r;
    }
    return a / x * b;
}

// The kernel arguments of call t depend on t only through t mod L_d and,
// for the macro store, (t + 1) mod store_every.  A run of P = lcm of those
// calls therefore repeats exactly: capture it once into a graph and replay.
lbmc_status run_steps(lbmc_ctx *c, long long t0, int nsteps, int store_every)
{
    if (c->key.closure && c->closure_every > 1 && !c->tcache_valid && nsteps > 0) {
        // the first call after a change of state evaluates the network; the
        // captured calls then start from a valid record
        if (lbmc_status st = launch_step(c, t0, store_every > 0 && ((t0 + 1) % store_every) == 0)) return st;
        t0++;
        nsteps--;
    }
    auto stores = [&](long long tt) { return store_every > 0 && ((tt + 1) % store_every) == 0; };
    long long period = lcm_ll(lcm_ll(c->L[0], c->L[1]), c->L[2]);
    if (store_every > 0) period = lcm_ll(period, store_every);
    if (c->key.closure && c->closure_every > 1) period = lcm_ll(period, c->closure_every);
    const bool graphable = c->use_graph && period <= 4096 && nsteps >= 2 * period;
    long long t = t0;
    long long left = nsteps;
    if (graphable) {
        const long long phase = t0 % period;
        const int sflag = store_every > 0 ? store_every : 0;
        if (!c->graph_exec || c->graph_period != period || c->graph_phase != phase || c->graph_store != sflag) {
            drop_graph(c);
            HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "graph capture");
            lbmc_status st = LBMC_OK;
            for (long long k = 0; k < period && st == LBMC_OK; k++) {
                const long long tt = t0 + k;
                st = launch_step(c, tt, stores(tt));
            }
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (st != LBMC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            if (e != hipSuccess) return hip_fail(c, e, "graph capture");
            c->graph = g;
            HIP_TRY(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0), "graph instantiate");
            c->graph_period = period;
            c->graph_phase = phase;
            c->graph_store = sflag;
        }
        while (left >= period) {
            HIP_TRY(c, hipGraphLaunch(c->graph_exec, c->stream), "graph launch");
            t += period;
            left -= period;
        }
    }
    for (; left > 0; left--, t++)
        if (lbmc_status st = launch_step(c, t, stores(t))) return st;
    return LBMC_OK;
}

}  // namespace

extern "C" {

const char *lbmc_last_error(const lbmc_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

lbmc_status lbmc_create(const lbmc_desc *desc, lbmc_ctx **out)
{
    if (!desc || !out) return fail(nullptr, LBMC_ERR_INVALID, "lbmc_create: null argument");
    *out = nullptr;
    if (desc->lattice != LBMC_D2Q9 && desc->lattice != LBMC_D3Q15)
        return fail(nullptr, LBMC_ERR_INVALID, "unknown lattice %d", desc->lattice);
    if (desc->collision_type < LBMC_BGK || desc->collision_type > LBMC_MRTPOR)
        return fail(nullptr, LBMC_ERR_INVALID, "unknown collision type %d", desc->collision_type);
    if (desc->storage < LBMC_FP64 || desc->storage > LBMC_BF16S)
        return fail(nullptr, LBMC_ERR_INVALID, "unknown storage %d", desc->storage);
    const int D = desc->lattice == LBMC_D2Q9 ? 2 : 3;
    long long N = 1;
    for (int d = 0; d < D; d++) {
        if (desc->length[d] <= 0)
            return fail(nullptr, LBMC_ERR_INVALID, "length[%d] = %d must be positive", d, desc->length[d]);
        N *= desc->length[d];
    }
    if (N >= (1LL << 31))
        return fail(nullptr, LBMC_ERR_UNSUPPORTED, "%lld nodes: at most 2^31-1 per lattice", N);

    lbmc_ctx *c = new lbmc_ctx();
    c->desc = *desc;
    c->D = D;
    c->Q = D == 2 ? 9 : 15;
    for (int d = 0; d < 3; d++) c->L[d] = d < D ? desc->length[d] : 1;
    c->N = N;
    c->npad = (N + 63) / 64 * 64;
    c->storage = desc->storage;
    c->sbytes = desc->storage == LBMC_FP64 ? 8 : desc->storage == LBMC_FP32 ? 4 : 2;
    c->cbytes = desc->storage == LBMC_FP64 ? 8 : 4;
    c->key.lattice = desc->lattice;
    c->key.family = (desc->collision_type == LBMC_MRT || desc->collision_type == LBMC_MRTPOR) ? 1 : 0;
    c->key.guo = desc->forcing ? 1 : 0;
    c->key.smag = (c->key.family == 0 && desc->c_smag > 0) ? 1 : 0;
    c->key.mfix = (desc->lattice == LBMC_D3Q15 && (desc->quirks & LBMC_FIX_D3Q15_MINV)) ? 1 : 0;
    c->offsets64 = (unsigned long long)c->npad * c->sbytes >= (1ULL << 32);
    c->key.wide = c->offsets64 ? 1 : 0;
    // Large D3Q15 FP32 lattices with Lx % 64 == 0 take the 64-bit-offset
    // kernels below 4 GiB too: with Lx % 64 == 0 those address each wave's
    // slot row from a wave-uniform base, measured faster than the narrow
    // kernels' per-lane offsets from the array base once a direction array
    // reaches ~1 GB (640^3 +15 %, 768^3 +12 %, 1024^2 x 1023 +27 %), equal at
    // 384^3-512^3 and 3 % slower at 256^3.  Same results either way.
    if (desc->lattice == LBMC_D3Q15 && desc->storage == LBMC_FP32 && desc->length[0] % 64 == 0 &&
        (unsigned long long)c->npad * c->sbytes >= (768ULL << 20))
        c->key.wide = 1;
    // test hook: run the 64-bit-offset kernels on a small lattice
    if (const char *env = getenv("LBMC_FORCE_WIDE"))
        if (atoi(env) != 0) c->key.wide = 1;
    // 16-bit storage with an even last axis (Lz; Ly on D2Q9) below the
    // 4 GiB-per-array limit: slots k and k + N/2 share one 32-bit word, so
    // the step moves two nodes per lane with aligned dword accesses on every
    // call (step_half_kernel).  Default on D3Q15 only: D2Q9 lattices are
    // what batches hold, and the batch kernel (one node per lane) pays for
    // the slot-to-element mapping on every access (64-member D2Q9 ensemble:
    // FP16S 96 k -> 85 k, BF16S 103 k -> 78 k MLUPS) more than a single
    // D2Q9 lattice gains (0-2.5 %).  LBMC_HALF_LAYOUT=1 turns it on for
    // both lattices, =0 off for both.
    if (c->sbytes == 2 && !c->key.wide && c->L[D - 1] % 2 == 0) {
        bool half = desc->lattice == LBMC_D3Q15;
        if (const char *env = getenv("LBMC_HALF_LAYOUT")) half = atoi(env) != 0;
        if (half) c->hs = N / 2;
    }
    choose_shape(c);
    if (const char *env = getenv("LBMC_GRAPH")) c->use_graph = atoi(env) != 0;

    auto bail = [&](lbmc_status st) {
        g_last_error = c->err;
        lbmc_destroy(c);
        return st;
    };
    hipError_t e = hipSetDevice(desc->device);
    if (e != hipSuccess) return bail(hip_fail(c, e, "hipSetDevice"));
    if ((e = hipStreamCreate(&c->stream)) != hipSuccess)
        return bail(hip_fail(c, e, "hipStreamCreate"));
    {
        // Resident blocks per CU of the step kernels.  Each resident block
        // keeps 256 nodes x Q populations of reads and writes in flight;
        // at full occupancy the many concurrent streams (2Q + 1 arrays)
        // lower the HBM throughput, and a cap of a few blocks per CU
        // (reserved LDS, which the step kernels do not use) raises it.
        int bpc = default_blocks_per_cu(c);
        if (const char *env = getenv("LBMC_STEP_BLOCKS_PER_CU")) bpc = atoi(env);
        // two nodes per lane (step_half_kernel on every call in the
        // half-interleaved layout, else step_pair_kernel on calls with an even
        // x shift); LBMC_STEP_PAIRS=0 keeps one node per lane on every call
        c->key.pairs = 1;
        if (const char *env = getenv("LBMC_STEP_PAIRS")) c->key.pairs = atoi(env) != 0;
        int lds_cu = 0, lds_block = 0;
        if ((e = hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                       desc->device)) != hipSuccess ||
            (e = hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, desc->device)) !=
                hipSuccess)
            return bail(hip_fail(c, e, "device attribute"));
        const int lds = (bpc >= 1 && bpc <= 16 && lds_cu > 0) ? lds_cu / bpc - 1024 : 0;
        c->key.lds = lds <= lds_block ? lds : 0;
    }
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess)
        return bail(hip_fail(c, e, "hipEventCreate"));
    if ((e = hipMalloc(&c->f, c->sbytes * (size_t)c->Q * (size_t)c->npad)) != hipSuccess)
        return bail(hip_fail(c, e, "populations"));
    if ((e = hipMalloc((void **)&c->mask, (size_t)c->npad)) != hipSuccess)
        return bail(hip_fail(c, e, "node mask"));
    if ((e = hipMalloc((void **)&c->cls, 256 * sizeof(uint16_t))) != hipSuccess)
        return bail(hip_fail(c, e, "class table"));
    if ((e = hipMalloc(&c->rho, c->cbytes * (size_t)c->npad)) != hipSuccess)
        return bail(hip_fail(c, e, "rho"));
    for (int d = 0; d < 3; d++)
        if ((e = hipMalloc(&c->u[d], c->cbytes * (size_t)c->npad)) != hipSuccess)
            return bail(hip_fail(c, e, "u"));
    c->npartials = 1024;
    if ((e = hipMalloc((void **)&c->partials, sizeof(double) * c->npartials)) != hipSuccess ||
        (e = hipMalloc((void **)&c->scalar, sizeof(double))) != hipSuccess)
        return bail(hip_fail(c, e, "reduction scratch"));
    // zero state: plain fluid everywhere, zero macros, zero populations
    if ((e = hipMemset(c->mask, 0, (size_t)c->npad)) != hipSuccess ||
        (e = hipMemset(c->cls, 0, 256 * sizeof(uint16_t))) != hipSuccess ||
        (e = hipMemset(c->f, 0, c->sbytes * (size_t)c->Q * (size_t)c->npad)) != hipSuccess ||
        (e = hipMemset(c->rho, 0, c->cbytes * (size_t)c->npad)) != hipSuccess)
        return bail(hip_fail(c, e, "initial clear"));
    for (int d = 0; d < 3; d++)
        if ((e = hipMemset(c->u[d], 0, c->cbytes * (size_t)c->npad)) != hipSuccess)
            return bail(hip_fail(c, e, "initial clear"));
    if ((e = hipDeviceSynchronize()) != hipSuccess) return bail(hip_fail(c, e, "create sync"));
    *out = c;
    return LBMC_OK;
}

lbmc_status lbmc_destroy(lbmc_ctx *c)
{
    if (!c) return LBMC_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_dev(c->f);
    void *p;
    p = c->mask; free_dev(p); c->mask = nullptr;
    p = c->cls; free_dev(p); c->cls = nullptr;
    free_dev(c->rho);
    for (int d = 0; d < 3; d++) {
        free_dev(c->u[d]);
        free_dev(c->force[d]);
    }
    free_dev(c->porosity);
    p = c->sfidx; free_dev(p);
    p = c->sfnb; free_dev(p);
    free_dev(c->sfbuf);
    p = c->partials; free_dev(p);
    p = c->scalar; free_dev(p);
    free_dev(c->mlp_img);
    p = c->tcache; free_dev(p); c->tcache = nullptr;
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBMC_OK;
}

lbmc_status lbmc_set_fields(lbmc_ctx *c, const double *geometry, const int *macro_bc,
                            const int *micro_bc, const double *const *force, const double *const *u,
                            const double *rho)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    const long long N = c->N;
    const int lat = c->desc.lattice;
    const bool porous_type = c->desc.collision_type == LBMC_NTPOR || c->desc.collision_type == LBMC_MRTPOR;
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "sync");

    // node descriptors: solid (B >= 1, solver.cu:67-72), porous, macro, micro
    std::vector<uint16_t> descv((size_t)N);
    bool present[4096] = {};
    bool any_porous = false;
    std::vector<long long> sfnodes;
    for (long long n = 0; n < N; n++) {
        const double B = geometry ? geometry[n] : 0.0;
        uint16_t d;
        if (B >= 1) {
            d = CLS_SOLID;
        } else {
            const int mac = (c->desc.macro_bc && macro_bc) ? macro_bc[n] : 0;
            const int mic = (c->desc.micro_bc && micro_bc) ? micro_bc[n] : 0;
            if (mac < 0 || mac > 15)
                return fail(c, LBMC_ERR_INVALID, "macro BC code %d at node %lld (valid 0..15)", mac, n);
            if (!micro_valid(lat, mic))
                return fail(c, LBMC_ERR_INVALID, "micro BC code %d at node %lld is not in the %s table", mic,
                            n, lat == LBMC_D2Q9 ? "D2Q9" : "D3Q15");
            d = (uint16_t)((mac << 4) | (mic << 8));
            if (porous_type && B != 0) {
                d |= CLS_POROUS;
                any_porous = true;
            }
            if (mic && micro_is_sf(lat, mic)) sfnodes.push_back(n);
        }
        descv[(size_t)n] = d;
        present[d] = true;
    }
    int cls_of[4096];
    int k = 1;
    uint16_t table[256] = {};
    for (int d = 1; d < 4096; d++) {
        if (!present[d]) continue;
        if (k >= 256) return fail(c, LBMC_ERR_UNSUPPORTED, "more than 255 distinct node classes");
        cls_of[d] = k;
        table[k] = (uint16_t)d;
        k++;
    }
    cls_of[0] = 0;
    std::vector<uint8_t> maskv((size_t)N);
    bool all_fluid = true;
    for (long long n = 0; n < N; n++) {
        maskv[(size_t)n] = (uint8_t)cls_of[descv[(size_t)n]];
        all_fluid = all_fluid && maskv[(size_t)n] == 0;
    }
    c->all_fluid = all_fluid;
    HIP_TRY(c, hipMemcpy(c->mask, maskv.data(), (size_t)N, hipMemcpyHostToDevice), "mask upload");
    HIP_TRY(c, hipMemcpy(c->cls, table, sizeof(table), hipMemcpyHostToDevice), "class upload");
    memcpy(c->cls_host, table, sizeof(table));
    c->ncls = k;

    // porosity side array (read only at porous classes)
    free_dev(c->porosity);
    c->porous = any_porous;
    if (any_porous) {
        HIP_TRY(c, hipMalloc(&c->porosity, c->cbytes * (size_t)c->npad), "porosity");
        std::vector<double> b(geometry, geometry + N);
        if (lbmc_status st = upload_c(c, c->porosity, b)) return st;
    }

    // force: a kernel constant when uniform over the non-solid nodes
    for (int d = 0; d < 3; d++) free_dev(c->force[d]);
    for (int d = 0; d < 3; d++) c->F[d] = 0;
    if (c->desc.forcing && force) {
        bool uniform = true;
        bool seen = false;
        double F0[3] = {0, 0, 0};
        for (long long n = 0; n < N && uniform; n++) {
            if (descv[(size_t)n] & CLS_SOLID) continue;
            double v[3] = {0, 0, 0};
            for (int d = 0; d < c->D; d++) v[d] = force[d] ? force[d][n] : 0.0;
            if (!seen) {
                memcpy(F0, v, sizeof(F0));
                seen = true;
            } else {
                for (int d = 0; d < c->D; d++)
                    if (!(v[d] == F0[d]) && !(std::isnan(v[d]) && std::isnan(F0[d]))) uniform = false;
            }
        }
        if (uniform) {
            for (int d = 0; d < 3; d++) c->F[d] = F0[d];
        } else {
            for (int d = 0; d < c->D; d++) {
                HIP_TRY(c, hipMalloc(&c->force[d], c->cbytes * (size_t)c->npad), "force");
                std::vector<double> v((size_t)N, 0.0);
                if (force[d]) v.assign(force[d], force[d] + N);
                if (lbmc_status st = upload_c(c, c->force[d], v)) return st;
            }
        }
    }

    // macro arrays: prescribed values at macro-BC nodes, initial output
    {
        std::vector<double> v((size_t)N, 0.0);
        if (rho) v.assign(rho, rho + N);
        if (lbmc_status st = upload_c(c, c->rho, v)) return st;
        for (int d = 0; d < 3; d++) {
            std::fill(v.begin(), v.end(), 0.0);
            if (u && d < c->D && u[d]) v.assign(u[d], u[d] + N);
            if (lbmc_status st = upload_c(c, c->u[d], v)) return st;
        }
    }

    // sf nodes: neighbour list, gather buffer and per-node row index
    void *p;
    p = c->sfidx; free_dev(p); c->sfidx = nullptr;
    p = c->sfnb; free_dev(p); c->sfnb = nullptr;
    free_dev(c->sfbuf);
    c->nsf = (int)sfnodes.size();
    if (c->nsf) {
        std::vector<int> nb(c->nsf), idx((size_t)N, 0);
        for (int s = 0; s < c->nsf; s++) {
            const long long n = sfnodes[s];
            const int mic = (int)((descv[(size_t)n] >> 8) & 15);
            int axis, side;
            sf_face(lat, mic, &axis, &side);
            int xyz[3] = {(int)(n % c->L[0]), (int)((n / c->L[0]) % c->L[1]),
                          (int)(n / ((long long)c->L[0] * c->L[1]))};
            xyz[axis] = (xyz[axis] + side + c->L[axis]) % c->L[axis];
            nb[s] = (int)(xyz[0] + (long long)xyz[1] * c->L[0] + (long long)xyz[2] * c->L[0] * c->L[1]);
            idx[(size_t)n] = s;
        }
        HIP_TRY(c, hipMalloc((void **)&c->sfnb, sizeof(int) * c->nsf), "sf list");
        HIP_TRY(c, hipMalloc((void **)&c->sfidx, sizeof(int) * (size_t)c->npad), "sf index");
        HIP_TRY(c, hipMalloc(&c->sfbuf, c->cbytes * (size_t)c->Q * c->nsf), "sf buffer");
        HIP_TRY(c, hipMemcpy(c->sfnb, nb.data(), sizeof(int) * c->nsf, hipMemcpyHostToDevice), "sf upload");
        HIP_TRY(c, hipMemcpy(c->sfidx, idx.data(), sizeof(int) * (size_t)N, hipMemcpyHostToDevice), "sf upload");
    }
    HIP_TRY(c, hipDeviceSynchronize(), "set_fields sync");
    c->tcache_valid = false;
    drop_graph(c);
    c->fields_set = true;
    return LBMC_OK;
}

lbmc_status lbmc_get_node_types(lbmc_ctx *c, uint8_t *mask, uint16_t *table, int *nclasses)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "sync");
    if (mask) HIP_TRY(c, hipMemcpy(mask, c->mask, (size_t)c->N, hipMemcpyDeviceToHost), "mask read");
    if (table) HIP_TRY(c, hipMemcpy(table, c->cls, 256 * sizeof(uint16_t), hipMemcpyDeviceToHost), "class read");
    if (nclasses) *nclasses = c->ncls;
    return LBMC_OK;
}

lbmc_status lbmc_get_links(lbmc_ctx *c, uint16_t *links)
{
    if (!c || !links) return fail(c, LBMC_ERR_INVALID, "null argument");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    uint16_t *dl = nullptr;
    HIP_TRY(c, hipMalloc((void **)&dl, sizeof(uint16_t) * (size_t)c->npad), "links");
    hipError_t e = launch_links(c->desc.lattice, c->mask, c->cls, c->L[0], c->L[1], c->L[2], dl, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(links, dl, sizeof(uint16_t) * (size_t)c->N, hipMemcpyDeviceToHost);
    (void)hipFree(dl);
    if (e != hipSuccess) return hip_fail(c, e, "links");
    return LBMC_OK;
}

lbmc_status lbmc_init_static(lbmc_ctx *c)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    invalidate_tau_record(c);
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, launch_init_static(c->desc.lattice, c->storage, c->f, c->npad, c->N, c->stream), "init");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "init sync");
    return LBMC_OK;
}

lbmc_status lbmc_init_equilibrium(lbmc_ctx *c, const double *rho, const double *const *u)
{
    if (!c || !rho || !u) return fail(c, LBMC_ERR_INVALID, "null argument");
    invalidate_tau_record(c);
    for (int d = 0; d < c->D; d++)
        if (!u[d]) return fail(c, LBMC_ERR_INVALID, "u[%d] is null", d);
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    const size_t nb = sizeof(double) * (size_t)c->N;
    double *dr = nullptr, *du[3] = {nullptr, nullptr, nullptr};
    hipError_t e = hipMalloc((void **)&dr, nb);
    for (int d = 0; d < c->D && e == hipSuccess; d++) e = hipMalloc((void **)&du[d], nb);
    if (e == hipSuccess) e = hipMemcpy(dr, rho, nb, hipMemcpyHostToDevice);
    for (int d = 0; d < c->D && e == hipSuccess; d++) e = hipMemcpy(du[d], u[d], nb, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = launch_init_equilibrium(c->desc.lattice, c->storage, c->f, c->npad, c->N, c->hs, dr, du[0], du[1],
                                    c->D > 2 ? du[2] : du[1], c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dr);
    for (int d = 0; d < 3; d++)
        if (du[d]) (void)hipFree(du[d]);
    if (e != hipSuccess) return hip_fail(c, e, "init_equilibrium");
    return LBMC_OK;
}

lbmc_status lbmc_step(lbmc_ctx *c, long long t, int store_macros)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    if (!c->fields_set) return fail(c, LBMC_ERR_STATE, "lbmc_set_fields must come before stepping");
    if (t < 0) return fail(c, LBMC_ERR_INVALID, "negative time index");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    return launch_step(c, t, store_macros);
}

lbmc_status lbmc_run(lbmc_ctx *c, long long t0, int nsteps, int store_every)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    if (!c->fields_set) return fail(c, LBMC_ERR_STATE, "lbmc_set_fields must come before stepping");
    if (t0 < 0 || nsteps < 0) return fail(c, LBMC_ERR_INVALID, "negative time index or step count");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    return run_steps(c, t0, nsteps, store_every);
}

lbmc_status lbmc_time_steps(lbmc_ctx *c, long long t0, int nsteps, int store_every, float *ms)
{
    if (!c || !ms) return fail(c, LBMC_ERR_INVALID, "null argument");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipEventRecord(c->ev0, c->stream), "event");
    if (lbmc_status st = lbmc_run(c, t0, nsteps, store_every)) return st;
    HIP_TRY(c, hipEventRecord(c->ev1, c->stream), "event");
    HIP_TRY(c, hipEventSynchronize(c->ev1), "event sync");
    HIP_TRY(c, hipEventElapsedTime(ms, c->ev0, c->ev1), "event time");
    return LBMC_OK;
}

lbmc_status lbmc_get_macros(lbmc_ctx *c, double *rho, double *const *u)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "sync");
    const size_t N = (size_t)c->N;
    auto read = [&](const void *src, double *dst) -> lbmc_status {
        if (!dst) return LBMC_OK;
        if (c->cbytes == 8) {
            HIP_TRY(c, hipMemcpy(dst, src, 8 * N, hipMemcpyDeviceToHost), "macro read");
        } else {
            std::vector<float> tmp(N);
            HIP_TRY(c, hipMemcpy(tmp.data(), src, 4 * N, hipMemcpyDeviceToHost), "macro read");
            for (size_t k = 0; k < N; k++) dst[k] = tmp[k];
        }
        return LBMC_OK;
    };
    if (lbmc_status st = read(c->rho, rho)) return st;
    if (u)
        for (int d = 0; d < c->D; d++)
            if (lbmc_status st = read(c->u[d], u[d])) return st;
    return LBMC_OK;
}

lbmc_status lbmc_get_node(lbmc_ctx *c, long long idx, double *rho, double *u)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    if (idx < 0 || idx >= c->N) return fail(c, LBMC_ERR_INVALID, "node %lld out of range", idx);
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "sync");
    auto read1 = [&](const void *base, double *dst) -> lbmc_status {
        if (!dst) return LBMC_OK;
        if (c->cbytes == 8) {
            HIP_TRY(c, hipMemcpy(dst, (const double *)base + idx, 8, hipMemcpyDeviceToHost), "node read");
        } else {
            float v;
            HIP_TRY(c, hipMemcpy(&v, (const float *)base + idx, 4, hipMemcpyDeviceToHost), "node read");
            *dst = v;
        }
        return LBMC_OK;
    };
    if (lbmc_status st = read1(c->rho, rho)) return st;
    if (u)
        for (int d = 0; d < c->D; d++)
            if (lbmc_status st = read1(c->u[d], &u[d])) return st;
    return LBMC_OK;
}

lbmc_status lbmc_get_f(lbmc_ctx *c, double *const *f)
{
    if (!c || !f) return fail(c, LBMC_ERR_INVALID, "null argument");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    double *tmp = nullptr;
    HIP_TRY(c, hipMalloc((void **)&tmp, sizeof(double) * (size_t)c->N), "f read buffer");
    hipError_t e = hipSuccess;
    for (int i = 0; i < c->Q && e == hipSuccess; i++) {
        if (!f[i]) continue;
        e = launch_f_to_double(c->desc.lattice, c->storage, c->f, c->npad, i, 0, c->N, c->hs, tmp, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(f[i], tmp, sizeof(double) * (size_t)c->N, hipMemcpyDeviceToHost);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(c, e, "get_f");
    return LBMC_OK;
}

// Device scratch of lbmc_get_f_noalloc: one chunk of one direction's slots,
// converted to doubles.  Each GPU the library runs on holds its own copy for
// the life of the process.  Host code never passes the array's name to a
// kernel or a copy (on the host it is only a stand-in for the symbol); it uses
// the device address that hipGetSymbolAddress returns for the current device.
constexpr long long kNoallocChunk = 1LL << 20;  // doubles: 8 MiB
static __device__ __align__(256) double scratch[kNoallocChunk];

namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoallocMaxDevices] = {};

// Device address of `scratch` on the current device (set up once per device).
lbmc_status scratch_on_device(lbmc_ctx *c, double **out)
{
    const int dev = c->desc.device;
    if (dev < 0 || dev >= kNoallocMaxDevices)
        return fail(c, LBMC_ERR_UNSUPPORTED, "lbmc_get_f_noalloc: device %d beyond the %d the scratch table holds",
                    dev, kNoallocMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "lbmc_get_f_noalloc: scratch address");
        if (!p) return fail(c, LBMC_ERR_DEVICE, "lbmc_get_f_noalloc: scratch address is null");
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return LBMC_OK;
}
}  // namespace

lbmc_status lbmc_get_f_noalloc(lbmc_ctx *c, double *const *f)
{
    if (!c || !f) return fail(c, LBMC_ERR_INVALID, "null argument");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    double *buf = nullptr;
    if (lbmc_status st = scratch_on_device(c, &buf)) return st;
    const long long N = c->N;
    for (int i = 0; i < c->Q; i++) {
        if (!f[i]) continue;
        for (long long k0 = 0; k0 < N; k0 += kNoallocChunk) {
            const long long cnt = std::min(kNoallocChunk, N - k0);
            // slots k0 .. k0 + cnt - 1 of direction i, the kernel of lbmc_get_f
            HIP_TRY(c,
                    launch_f_to_double(c->desc.lattice, c->storage, c->f, c->npad, i, k0, cnt, c->hs, buf,
                                       c->stream),
                    "get_f_noalloc launch");
            HIP_TRY(c, hipStreamSynchronize(c->stream), "get_f_noalloc sync");
            HIP_TRY(c, hipMemcpy(f[i] + k0, buf, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost),
                    "get_f_noalloc read");
        }
    }
    return LBMC_OK;
}

lbmc_status lbmc_set_f(lbmc_ctx *c, const double *const *f)
{
    if (!c || !f) return fail(c, LBMC_ERR_INVALID, "null argument");
    invalidate_tau_record(c);
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    double *tmp = nullptr;
    HIP_TRY(c, hipMalloc((void **)&tmp, sizeof(double) * (size_t)c->N), "f write buffer");
    hipError_t e = hipStreamSynchronize(c->stream);
    for (int i = 0; i < c->Q && e == hipSuccess; i++) {
        if (!f[i]) continue;
        e = hipMemcpy(tmp, f[i], sizeof(double) * (size_t)c->N, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = launch_f_from_double(c->desc.lattice, c->storage, c->f, c->npad, i, c->N, c->hs, tmp, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return hip_fail(c, e, "set_f");
    return LBMC_OK;
}

lbmc_status lbmc_kinetic_energy(lbmc_ctx *c, double *energy)
{
    if (!c || !energy) return fail(c, LBMC_ERR_INVALID, "null argument");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    const int nb = (int)std::min<long long>(c->npartials, (c->N + 255) / 256);
    HIP_TRY(c, launch_energy(c->D, c->cbytes == 8, c->rho, c->u[0], c->u[1], c->u[2], c->N, c->partials,
                             nb, c->scalar, c->stream),
            "energy launch");
    HIP_TRY(c, hipMemcpyAsync(energy, c->scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream), "energy read");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "energy sync");
    return LBMC_OK;
}

lbmc_status lbmc_set_closure(lbmc_ctx *c, const lbmc_mlp *model, double tau_lo, double tau_hi)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "lbmc_set_closure: null context");
    if (!model) {
        c->key.closure = 0;
        c->key.closure_fp16 = 0;
        c->closure_model = nullptr;
        drop_graph(c);
        return LBMC_OK;
    }
    if (c->storage == LBMC_FP64)
        return fail(c, LBMC_ERR_UNSUPPORTED, "lbmc_set_closure: the closure computes in FP32; FP64 storage has none");
    if (c->key.family != 0)
        return fail(c, LBMC_ERR_UNSUPPORTED, "lbmc_set_closure: the closure replaces the BGK Smagorinsky hook (BGK/NTPOR only)");
    if (!(tau_lo > 0.5) || !(tau_hi >= tau_lo) || !std::isfinite(tau_hi))
        return fail(c, LBMC_ERR_INVALID, "lbmc_set_closure: band [%g, %g] must satisfy 0.5 < lo <= hi", tau_lo, tau_hi);
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    if (!c->mlp_img) HIP_TRY(c, hipMalloc(&c->mlp_img, mlp_image_bytes()), "closure image allocation");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "closure sync");
    const lbmc_status st = mlp_build_image(model, c->mlp_img, c->desc.device, tau_lo, tau_hi, c->stream);
    if (st != LBMC_OK) return fail(c, st, "lbmc_set_closure: model is on another device or the image build failed");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "closure image build");
    c->key.closure = 1;
    c->key.closure_fp16 = mlp_eval_fp16(model) ? 1 : 0;  // fixed until the model is set again
    c->closure_model = model;
    c->closure_lo = tau_lo;
    c->closure_hi = tau_hi;
    c->tcache_valid = false;
    drop_graph(c);
    return LBMC_OK;
}

lbmc_status lbmc_set_closure_interval(lbmc_ctx *c, int every)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "lbmc_set_closure_interval: null context");
    if (every < 1 || every > 4096) return fail(c, LBMC_ERR_INVALID, "lbmc_set_closure_interval: interval %d not in [1, 4096]", every);
    if (every > 1 && !c->tcache) {
        HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
        HIP_TRY(c, hipMalloc((void **)&c->tcache, sizeof(float) * (size_t)c->npad), "closure tau record");
    }
    c->closure_every = every;
    c->tcache_valid = false;
    drop_graph(c);
    return LBMC_OK;
}

lbmc_status lbmc_set_closure_shift(lbmc_ctx *c, double dtau)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "lbmc_set_closure_shift: null context");
    if (!std::isfinite(dtau)) return fail(c, LBMC_ERR_INVALID, "lbmc_set_closure_shift: shift not finite");
    c->closure_shift = dtau;
    invalidate_tau_record(c);
    drop_graph(c);
    return LBMC_OK;
}

lbmc_status lbmc_sync(lbmc_ctx *c)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    HIP_TRY(c, hipSetDevice(c->desc.device), "hipSetDevice");
    HIP_TRY(c, hipStreamSynchronize(c->stream), "sync");
    return LBMC_OK;
}

lbmc_status lbmc_info(const lbmc_ctx *c, long long *nodes, int *q, int *slot_bytes)
{
    if (!c) return fail(nullptr, LBMC_ERR_INVALID, "null context");
    if (nodes) *nodes = c->N;
    if (q) *q = c->Q;
    if (slot_bytes) *slot_bytes = (int)c->sbytes;
    return LBMC_OK;
}

lbmc_status lbmc_copy_bandwidth(int device, size_t nbytes, int reps, float *ms)
{
    if (!ms || nbytes == 0 || nbytes % 16 != 0 || reps <= 0)
        return fail(nullptr, LBMC_ERR_INVALID, "copy: nbytes must be a positive multiple of 16");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipMalloc(&a, nbytes);
    if (e == hipSuccess) e = hipMalloc(&b, nbytes);
    if (e == hipSuccess) e = hipMemset(a, 1, nbytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = launch_copy(a, b, nbytes, s);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy(a, b, nbytes, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return hip_fail(nullptr, e, "copy bandwidth");
    return LBMC_OK;
}

}  // extern "C"
