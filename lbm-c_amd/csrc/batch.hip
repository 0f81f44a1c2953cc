// Batched stepping of several lattices (ensembles of uneven lattices): one
// call for every member is one pass of launches.
//
// The members are ordinary contexts (lbmc_create + lbmc_set_fields + an
// initial condition).  The batch keeps, in device memory, each member's
// kernel arguments (its buffers, its collision variant, the slot offsets of
// the current call t and its closure schedule) and lists of chunks (runs of
// 256-node tiles of one member) over all members and over the closure
// members.  A call is
//   batch_step_kernel     blocks take chunks (striding over the list when
//                         the grid is smaller); a chunk belongs to one
//                         member, so its arguments and collision variant
//                         are uniform over the block; chunks of closure
//                         members that evaluate the network are skipped;
//   batch_eval_kernel     (closure members only) those chunks, with the
//                         network; returns at once when no member evaluates;
// The kernels of call t + j read the members' arguments of that call from
// entry j of a window of kBatchWindow calls in device memory (and the
// evaluation kernel its flag any_eval[j]).  At the end of a window of calls
//   batch_advance_kernel  one block moves every member's state from call t
//                         to t + calls (shift counters, slot offsets, the
//                         macro-store flag, the closure phase), and
//   batch_window_kernel   one block per entry lays out the next window and
//                         its evaluation flags,
// so no kernel takes t as an argument and the state moves on once per
// window.  Because nothing in the launches depends on t, a window of calls
// is captured once into a HIP graph and replayed for any t.  The launch
// shape of the step kernel is chosen by lbmc_batch_autotune (timed
// candidates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "batch_kernel.hpp"
#include "ctx.hpp"
#include "dispatch.hpp"
#include "node_index.hpp"


using namespace lbmc;

struct lbmc_batch {
    int count = 0;
    int lattice = 0, storage = 0, device = 0;
    bool cdouble = false;
    std::vector<lbmc_ctx *> members;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *dmembers = nullptr;  // BatchMember<C>[count]: the state of call t_next
    void *dwindow = nullptr;   // BatchMember<C>[kBatchWindow][count]: calls t_next + j
    int2 *chunks = nullptr;
    int nchunks = 0;
    int2 *cchunks = nullptr;  // the closure members' chunks
    int ncchunks = 0;
    int eval_grid = 0;        // blocks of the evaluation kernel
    int *any_eval = nullptr;  // device flags [kBatchWindow]: some closure member evaluates call t_next + j
    const void *mlp = nullptr;  // closure weights shared by the closure members (a member's image)
    long long nodes = 0;
    int grid = 1;
    int sub = 1;            // 256-node tiles per chunk
    int blocks_per_cu = 0;  // residency cap of the step kernel (0: none)
    int lds = 0;            //   as dynamic LDS bytes per block
    int lds_cu = 0;         // LDS bytes per CU
    int ncu = 256;          // compute units of the device
    long long t_next = -1;  // the call the device arguments describe
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_steps = 0;
    const lbmc_mlp *closure_model = nullptr;
    double closure_lo = 0, closure_hi = 0;
    int closure_fp16 = 0;  // the closure members' network with fp16 operands
    std::string err;
};

namespace {

std::string g_batch_error;

lbmc_status bfail(lbmc_batch *b, lbmc_status st, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    else g_batch_error = buf;
    return st;
}

#define B_HIP(b, call)                                                                              \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return bfail((b), LBMC_ERR_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

int member_variant(const lbmc_ctx *c)
{
    const KernelKey &k = c->key;
    if (k.closure) return 8 + (k.guo ? 1 : 0);
    const int bit0 = k.family == 0 ? (k.smag ? 1 : 0) : (k.mfix ? 1 : 0);
    return (k.family << 2) | ((k.guo ? 1 : 0) << 1) | bit0;
}

// Member arguments for call t0, read from the members' current state (so
// that set_fields / set_f on a member between runs is seen).
template <class C>
lbmc_status upload_members(lbmc_batch *b, long long t0, int store_every)
{
    std::vector<BatchMember<C>> hm((size_t)b->count);
    for (int k = 0; k < b->count; k++) {
        const lbmc_ctx *c = b->members[(size_t)k];
        if (c->nsf || c->offsets64 || !c->fields_set)
            return bfail(b, LBMC_ERR_UNSUPPORTED, "batch member %d: sf nodes or state not supported", k);
        if (c->key.closure && (c->closure_model != b->closure_model || c->closure_lo != b->closure_lo ||
                               c->closure_hi != b->closure_hi || c->key.closure_fp16 != b->closure_fp16))
            return bfail(b, LBMC_ERR_UNSUPPORTED, "batch member %d: closure members must share one model, band and operand precision", k);
        BatchMember<C> &m = hm[(size_t)k];
        memset(&m, 0, sizeof(m));
        const int se = store_every > 0 ? store_every : 0;
        fill_args(b->members[(size_t)k], m.a, t0, se > 0 && ((t0 + 1) % se) == 0);
        m.variant = member_variant(b->members[(size_t)k]);
        m.store_every = se;
        m.store_count = se > 0 ? (int)((t0 + 1) % se) : 0;
        // closure interval: the schedule of a single lattice (lbmc_run): the
        // network runs on calls t % K == 0 and on the first call whose
        // state has no valid record
        m.closure_every = c->key.closure ? c->closure_every : 1;
        m.closure_phase = (int)(t0 % m.closure_every);
        if (m.closure_every > 1) m.a.tmode = (!c->tcache_valid || m.closure_phase == 0) ? 1 : 2;
    }
    B_HIP(b, hipMemcpyAsync(b->dmembers, hm.data(), sizeof(BatchMember<C>) * hm.size(), hipMemcpyHostToDevice,
                            b->stream));
    hipLaunchKernelGGL(batch_window_kernel<C>, dim3(kBatchWindow), dim3(256), 0, b->stream,
                       (const BatchMember<C> *)b->dmembers, b->count, (BatchMember<C> *)b->dwindow, b->any_eval);
    B_HIP(b, hipGetLastError());
    B_HIP(b, hipStreamSynchronize(b->stream));
    b->t_next = t0;
    return LBMC_OK;
}

// call t_next + j (j < kBatchWindow) from window entry j
lbmc_status launch_one_step(lbmc_batch *b, int j)
{
    hipError_t e;
    const size_t msz = b->cdouble ? sizeof(BatchMember<double>) : sizeof(BatchMember<float>);
    const void *members = (const char *)b->dwindow + (size_t)j * (size_t)b->count * msz;
    const BatchShape sh{b->grid, b->sub, b->lds};
    const BatchEval ev{b->cchunks, b->mlp ? b->ncchunks : 0, b->eval_grid, b->mlp, b->any_eval + j,
                       b->closure_fp16 != 0};
    switch (b->storage) {
    case LBMC_FP64: e = launch_batch_fp64(b->lattice, members, b->chunks, b->nchunks, b->grid, b->sub, b->lds, b->stream); break;
    case LBMC_FP32: e = launch_batch_st<StFP32>(b->lattice, members, b->chunks, b->nchunks, sh, ev, b->stream); break;
    case LBMC_FP16S: e = launch_batch_st<StFP16S>(b->lattice, members, b->chunks, b->nchunks, sh, ev, b->stream); break;
    default: e = launch_batch_st<StBF16S>(b->lattice, members, b->chunks, b->nchunks, sh, ev, b->stream); break;
    }
    if (e != hipSuccess) return bfail(b, LBMC_ERR_DEVICE, "batch step launch: %s", hipGetErrorString(e));
    return LBMC_OK;
}

// the state and the window move on by `calls` (<= kBatchWindow) calls
template <class C>
void launch_advance(lbmc_batch *b, int calls)
{
    hipLaunchKernelGGL(batch_advance_kernel<C>, dim3(1), dim3(256), 0, b->stream, (BatchMember<C> *)b->dmembers,
                       b->count, calls);
    hipLaunchKernelGGL(batch_window_kernel<C>, dim3(kBatchWindow), dim3(256), 0, b->stream,
                       (const BatchMember<C> *)b->dmembers, b->count, (BatchMember<C> *)b->dwindow, b->any_eval);
}

lbmc_status advance(lbmc_batch *b, int calls)
{
    if (b->cdouble) launch_advance<double>(b, calls);
    else launch_advance<float>(b, calls);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, LBMC_ERR_DEVICE, "batch advance launch: %s", hipGetErrorString(e));
    return LBMC_OK;
}

void drop_batch_graph(lbmc_batch *b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

constexpr int kBatchGraphSteps = kBatchWindow;

lbmc_status upload(lbmc_batch *b, long long t0, int store_every)
{
    return b->cdouble ? upload_members<double>(b, t0, store_every) : upload_members<float>(b, t0, store_every);
}

// nsteps calls from t0; the device arguments must describe call t0
lbmc_status batch_steps(lbmc_batch *b, long long t0, int nsteps)
{
    int left = nsteps;
    if (left >= kBatchGraphSteps) {
        if (!b->graph_exec) {
            B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            lbmc_status st = LBMC_OK;
            for (int k = 0; k < kBatchGraphSteps && st == LBMC_OK; k++) st = launch_one_step(b, k);
            if (st == LBMC_OK) st = advance(b, kBatchGraphSteps);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (st != LBMC_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            if (e != hipSuccess) return bfail(b, LBMC_ERR_DEVICE, "batch graph capture: %s", hipGetErrorString(e));
            b->graph = g;
            B_HIP(b, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            b->graph_steps = kBatchGraphSteps;
        }
        while (left >= b->graph_steps) {
            B_HIP(b, hipGraphLaunch(b->graph_exec, b->stream));
            left -= b->graph_steps;
        }
    }
    // the rest of the run: calls of the current window, then one advance
    for (int j = 0; j < left; j++) {
        lbmc_status st = launch_one_step(b, j);
        if (st != LBMC_OK) return st;
    }
    if (left > 0)
        if (lbmc_status st = advance(b, left)) return st;
    b->t_next = t0 + nsteps;
    return LBMC_OK;
}

// List position p of a chunk list runs on XCD p mod 8 (blocks are dealt to
// the XCDs round robin, and the grid is either one block per position or a
// multiple of 8 that strides over the list).  Runs of G = 2^glog consecutive
// chunks (member-major order) go to positions p, p + 8, ..., one XCD: a
// cache line that two neighbouring chunks of a member share -- a chunk
// boundary is rarely line-aligned in the rotated rows of the index shift --
// is then fetched and written back by one L2 instead of two (row_chunk,
// node_index.hpp).  The list is padded to whole groups of 8 runs with
// entries of member -1, which the kernels skip.
static void xcd_order(std::vector<int2> &v, unsigned glog)
{
    if (glog == 0 || v.empty()) return;
    const unsigned n = (unsigned)v.size();
    const unsigned runs = (n + (1u << glog) - 1u) >> glog;
    const unsigned total = ((runs + 7u) / 8u * 8u) << glog;
    std::vector<int2> out(total, make_int2(-1, 0));
    for (unsigned p = 0; p < total; p++) {
        const unsigned c = row_chunk(p, glog);
        if (c < n) out[p] = v[c];
    }
    v.swap(out);
}

// Chunk lists for `sub` tiles per chunk: (member, first node) of every
// chunk, member by member, and the same for the closure members alone, in
// XCD order with runs of ~1024 nodes (measured best for single D2Q9
// lattices: 4 chunks of 256 nodes per XCD).
lbmc_status build_chunks(lbmc_batch *b, int sub)
{
    const long long cn = (long long)sub * kBlock;
    std::vector<int2> ch, cc;
    for (int k = 0; k < b->count; k++)
        for (long long n = 0; n < b->members[(size_t)k]->N; n += cn) {
            ch.push_back(make_int2(k, (int)n));
            if (b->members[(size_t)k]->key.closure) cc.push_back(make_int2(k, (int)n));
        }
    const unsigned glog = sub == 1 ? 2u : sub == 2 ? 1u : 0u;
    xcd_order(ch, glog);
    xcd_order(cc, glog);
    if (b->chunks) B_HIP(b, hipFree(b->chunks));
    if (b->cchunks) B_HIP(b, hipFree(b->cchunks));
    b->chunks = nullptr;
    b->cchunks = nullptr;
    B_HIP(b, hipMalloc((void **)&b->chunks, sizeof(int2) * ch.size()));
    B_HIP(b, hipMemcpy(b->chunks, ch.data(), sizeof(int2) * ch.size(), hipMemcpyHostToDevice));
    if (!cc.empty()) {
        B_HIP(b, hipMalloc((void **)&b->cchunks, sizeof(int2) * cc.size()));
        B_HIP(b, hipMemcpy(b->cchunks, cc.data(), sizeof(int2) * cc.size(), hipMemcpyHostToDevice));
    }
    b->nchunks = (int)ch.size();
    b->ncchunks = (int)cc.size();
    b->sub = sub;
    b->grid = std::min(b->grid, b->nchunks);
    // evaluation kernel: persistent, 4 blocks per CU (its residency)
    b->eval_grid = std::max(1, std::min(b->ncchunks, 4 * b->ncu));
    return LBMC_OK;
}

// after calls were run: every closure member with an interval now holds the
// record of its last evaluation
void mark_records(lbmc_batch *b)
{
    for (lbmc_ctx *c : b->members)
        if (c->key.closure && c->closure_every > 1) c->tcache_valid = true;
}

lbmc_status sync_members(lbmc_batch *b)
{
    for (lbmc_ctx *c : b->members) B_HIP(b, hipStreamSynchronize(c->stream));
    return LBMC_OK;
}

}  // namespace

extern "C" {

const char *lbmc_batch_last_error(const lbmc_batch *b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

lbmc_status lbmc_batch_create(lbmc_ctx *const *members, int count, lbmc_batch **out)
{
    if (!out || !members || count <= 0) return bfail(nullptr, LBMC_ERR_INVALID, "lbmc_batch_create: bad argument");
    *out = nullptr;
    const lbmc_ctx *c0 = members[0];
    if (!c0) return bfail(nullptr, LBMC_ERR_INVALID, "lbmc_batch_create: member 0 is null");
    long long nchunks = 0, nodes = 0;
    const lbmc_mlp *closure_model = nullptr;
    double closure_lo = 0, closure_hi = 0;
    int closure_fp16 = 0;
    const void *closure_img = nullptr;
    for (int k = 0; k < count; k++) {
        const lbmc_ctx *c = members[k];
        if (!c) return bfail(nullptr, LBMC_ERR_INVALID, "lbmc_batch_create: member %d is null", k);
        if (!c->fields_set) return bfail(nullptr, LBMC_ERR_STATE, "lbmc_batch_create: member %d has no fields", k);
        if (c->desc.lattice != c0->desc.lattice || c->storage != c0->storage || c->desc.device != c0->desc.device)
            return bfail(nullptr, LBMC_ERR_INVALID,
                         "lbmc_batch_create: member %d differs in lattice, storage or device from member 0", k);
        if (c->nsf) return bfail(nullptr, LBMC_ERR_UNSUPPORTED, "lbmc_batch_create: member %d has sf boundary nodes", k);
        if (c->key.closure) {
            if (closure_model && (c->closure_model != closure_model || c->closure_lo != closure_lo ||
                                  c->closure_hi != closure_hi || c->key.closure_fp16 != closure_fp16))
                return bfail(nullptr, LBMC_ERR_UNSUPPORTED,
                             "lbmc_batch_create: closure members must share one model, band and operand precision "
                             "(member %d)", k);
            closure_model = c->closure_model;
            closure_fp16 = c->key.closure_fp16;
            closure_lo = c->closure_lo;
            closure_hi = c->closure_hi;
            closure_img = c->mlp_img;
        }
        if (c->offsets64) return bfail(nullptr, LBMC_ERR_UNSUPPORTED, "lbmc_batch_create: member %d is too large", k);
        for (int j = 0; j < k; j++)
            if (members[j] == c) return bfail(nullptr, LBMC_ERR_INVALID, "lbmc_batch_create: member %d repeats %d", k, j);
        nchunks += (c->N + kBlock - 1) / kBlock;
        nodes += c->N;
    }
    if (nchunks >= (1LL << 31)) return bfail(nullptr, LBMC_ERR_UNSUPPORTED, "lbmc_batch_create: too many nodes");
    lbmc_batch *b = new lbmc_batch();
    b->count = count;
    b->lattice = c0->desc.lattice;
    b->storage = c0->storage;
    b->device = c0->desc.device;
    b->cdouble = c0->cbytes == 8;
    b->members.assign(members, members + count);
    b->nchunks = (int)nchunks;
    b->nodes = nodes;
    b->closure_model = closure_model;
    b->closure_fp16 = closure_fp16;
    b->closure_lo = closure_lo;
    b->closure_hi = closure_hi;
    b->mlp = closure_img;
    auto bail = [&](lbmc_status st) {
        g_batch_error = b->err;
        lbmc_batch_destroy(b);
        return st;
    };
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreate(&b->stream);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    const size_t msz = b->cdouble ? sizeof(BatchMember<double>) : sizeof(BatchMember<float>);
    if (e == hipSuccess) e = hipMalloc(&b->dmembers, msz * (size_t)count);
    if (e == hipSuccess) e = hipMalloc(&b->dwindow, msz * (size_t)count * kBatchWindow);
    if (e == hipSuccess) e = hipMalloc((void **)&b->any_eval, sizeof(int) * kBatchWindow);
    if (e != hipSuccess) return bail(bfail(b, LBMC_ERR_DEVICE, "lbmc_batch_create: %s", hipGetErrorString(e)));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess && prop.multiProcessorCount > 0)
        b->ncu = prop.multiProcessorCount;
    b->grid = b->nchunks;
    if (lbmc_status st = build_chunks(b, 1)) return bail(st);
    // default persistent grid: 8 blocks per CU, at most one block per chunk
    b->grid = (int)std::min<long long>(b->nchunks, 8LL * b->ncu);
    if (hipDeviceGetAttribute(&b->lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, b->device) != hipSuccess)
        b->lds_cu = 0;
    *out = b;
    return LBMC_OK;
}

lbmc_status lbmc_batch_destroy(lbmc_batch *b)
{
    if (!b) return LBMC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_batch_graph(b);
    if (b->dmembers) (void)hipFree(b->dmembers);
    if (b->dwindow) (void)hipFree(b->dwindow);
    if (b->chunks) (void)hipFree(b->chunks);
    if (b->cchunks) (void)hipFree(b->cchunks);
    if (b->any_eval) (void)hipFree(b->any_eval);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return LBMC_OK;
}

lbmc_status lbmc_batch_run(lbmc_batch *b, long long t0, int nsteps, int store_every)
{
    if (!b || nsteps < 0 || t0 < 0) return bfail(b, LBMC_ERR_INVALID, "lbmc_batch_run: bad argument");
    B_HIP(b, hipSetDevice(b->device));
    lbmc_status st = sync_members(b);
    if (st == LBMC_OK) st = upload(b, t0, store_every);
    if (st == LBMC_OK) st = batch_steps(b, t0, nsteps);
    if (st != LBMC_OK) return st;
    if (nsteps > 0) mark_records(b);
    B_HIP(b, hipStreamSynchronize(b->stream));
    return LBMC_OK;
}

lbmc_status lbmc_batch_time_steps(lbmc_batch *b, long long t0, int nsteps, int store_every, float *ms)
{
    if (!b || nsteps <= 0 || t0 < 0 || !ms) return bfail(b, LBMC_ERR_INVALID, "lbmc_batch_time_steps: bad argument");
    B_HIP(b, hipSetDevice(b->device));
    lbmc_status st = sync_members(b);
    if (st == LBMC_OK) st = upload(b, t0, store_every);  // outside the timed region
    if (st != LBMC_OK) return st;
    B_HIP(b, hipEventRecord(b->ev0, b->stream));
    st = batch_steps(b, t0, nsteps);
    if (st != LBMC_OK) return st;
    mark_records(b);
    B_HIP(b, hipEventRecord(b->ev1, b->stream));
    B_HIP(b, hipEventSynchronize(b->ev1));
    B_HIP(b, hipEventElapsedTime(ms, b->ev0, b->ev1));
    return LBMC_OK;
}

// Residency cap of the batch step kernel: at most `per` blocks per CU,
// through dynamic LDS that the kernel does not use (0: no cap).  Not with
// closure members, whose kernel stages the network in static LDS.
void set_blocks_per_cu(lbmc_batch *b, int per)
{
    b->blocks_per_cu = (per >= 3 && per <= 16 && b->lds_cu > 0 && !b->mlp) ? per : 0;
    b->lds = b->blocks_per_cu ? b->lds_cu / b->blocks_per_cu - 1024 : 0;
}

// Times candidate launch shapes over a few steps from the current state and
// keeps the fastest: chunks of 1, 2, 4 or 8 tiles with one block per chunk
// and the resident blocks per CU uncapped or capped at 4 or 6, and
// persistent grids of 2, 4, 8, 16 blocks per CU over one-tile chunks.  The
// steps advance the members' state; callers tune before the timed run.
lbmc_status lbmc_batch_autotune(lbmc_batch *b, long long t0, int steps_per_trial, int *grid, float *ms_per_step)
{
    if (!b || steps_per_trial <= 0 || t0 < 0) return bfail(b, LBMC_ERR_INVALID, "lbmc_batch_autotune: bad argument");
    B_HIP(b, hipSetDevice(b->device));
    const int ncu = b->ncu;
    struct Cand {
        int sub, grid_per_cu, cap;  // grid_per_cu 0: one block per chunk
    };
    std::vector<Cand> cand;
    for (int sub : {1, 2, 4, 8})
        for (int cap : {0, 4, 6}) cand.push_back({sub, 0, cap});
    for (int per : {2, 4, 8, 16}) cand.push_back({1, per, 0});
    float best = -1.f;
    Cand best_c{b->sub, 0, b->blocks_per_cu};
    int best_grid = b->grid;
    long long t = t0;
    for (const Cand &cd : cand) {
        set_blocks_per_cu(b, cd.cap);
        if (cd.cap && !b->blocks_per_cu) continue;  // cap not available
        drop_batch_graph(b);
        if (cd.sub != b->sub)
            if (lbmc_status st = build_chunks(b, cd.sub)) return st;
        b->grid = cd.grid_per_cu ? (int)std::min<long long>(b->nchunks, (long long)cd.grid_per_cu * ncu) : b->nchunks;
        float ms = 0.f;
        // one untimed pass (graph build), then the timed one
        lbmc_status st = lbmc_batch_time_steps(b, t, steps_per_trial, 0, &ms);
        if (st != LBMC_OK) return st;
        t += steps_per_trial;
        st = lbmc_batch_time_steps(b, t, steps_per_trial, 0, &ms);
        if (st != LBMC_OK) return st;
        t += steps_per_trial;
        if (best < 0.f || ms < best) {
            best = ms;
            best_c = cd;
            best_c.cap = b->blocks_per_cu;
            best_grid = b->grid;
        }
    }
    drop_batch_graph(b);
    if (best_c.sub != b->sub)
        if (lbmc_status st = build_chunks(b, best_c.sub)) return st;
    b->grid = std::min(best_grid, b->nchunks);
    set_blocks_per_cu(b, best_c.cap);
    if (grid) *grid = b->grid;
    if (ms_per_step) *ms_per_step = best / (float)steps_per_trial;
    return LBMC_OK;
}

lbmc_status lbmc_batch_info(const lbmc_batch *b, long long *nodes, int *grid, int *chunks, int *chunk_nodes,
                            int *blocks_per_cu)
{
    if (!b) return LBMC_ERR_INVALID;
    if (nodes) *nodes = b->nodes;
    if (grid) *grid = b->grid;
    if (chunks) *chunks = b->nchunks;
    if (chunk_nodes) *chunk_nodes = b->sub * kBlock;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    return LBMC_OK;
}

lbmc_status lbmc_batch_set_shape(lbmc_batch *b, int grid, int chunk_nodes, int blocks_per_cu)
{
    if (!b || grid <= 0 || blocks_per_cu < 0 || chunk_nodes < kBlock || chunk_nodes % kBlock ||
        chunk_nodes > 64 * kBlock)
        return bfail(b, LBMC_ERR_INVALID,
                     "lbmc_batch_set_shape: grid > 0, chunk_nodes a multiple of 256 up to 16384, blocks_per_cu >= 0");
    B_HIP(b, hipSetDevice(b->device));
    drop_batch_graph(b);
    if (chunk_nodes / kBlock != b->sub)
        if (lbmc_status st = build_chunks(b, chunk_nodes / kBlock)) return st;
    b->grid = std::min(grid, b->nchunks);
    set_blocks_per_cu(b, blocks_per_cu);
    return LBMC_OK;
}

}  // extern "C"
