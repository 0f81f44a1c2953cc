/*
 * lbmc.h -- C ABI of the MI355X lattice Boltzmann time-step library.
 *
 * The library replaces LBM-C's device time-step path and the host glue that
 * drives it:
 *   - the kernel iterate_kernel (LBM-C/solver.cu:10-136) and its launcher
 *     iterate(int t) (LBM-C/index.cu:284-313)        -> lbmc_step / lbmc_run
 *   - the constant upload cudaMemcpyToSymbol("domain_constants")
 *     (LBM-C/model_builder.cu:63-83)                  -> lbmc_create
 *   - the device arrays and pointer-of-pointer structs built by
 *     ModelBuilder (LBM-C/model_builder.cu:86-231)    -> lbmc_set_fields
 *   - load_static_IC (LBM-C/model_builder.cu:233-248) -> lbmc_init_static
 *   - the read-backs of output_macros / screen_mess
 *     (LBM-C/index.cu:196-216, :412-439)              -> lbmc_get_macros / lbmc_get_node
 *   - current_RMS's Thrust energy reduction
 *     (LBM-C/index.cu:315-370)                        -> lbmc_kinetic_energy
 *   - combi_malloc / cudasafe / Check_CUDA_Error
 *     (LBM-C/cuda_util.cu:6-27)                       -> status codes + lbmc_last_error
 *
 * Conventions
 *   - Plain C types only.  Host arrays use the reference's layout: one array
 *     per component, N = Lx*Ly(*Lz) nodes, x fastest (solver.cu:23,26).
 *   - Every function returns an lbmc_status; LBMC_OK is 0.  The reference
 *     printed and called exit(-1) on any error; a caller that wants that
 *     behaviour checks the status and does so itself.
 *   - The context owns all device memory; the caller owns host buffers.
 *   - One host thread per context.  Work is ordered on the context's stream;
 *     lbmc_get_* and lbmc_sync wait for it.
 */
#ifndef LBMC_H
#define LBMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum lbmc_status {
    LBMC_OK = 0,
    LBMC_ERR_INVALID = 1,     /* bad argument or field value */
    LBMC_ERR_NOMEM = 2,       /* device or host allocation failed */
    LBMC_ERR_DEVICE = 3,      /* HIP runtime or kernel launch error */
    LBMC_ERR_STATE = 4,       /* call out of order (e.g. step before fields) */
    LBMC_ERR_UNSUPPORTED = 5  /* valid request this build does not offer */
} lbmc_status;

/* lattice: the compile-time D2Q9 / D3Q15 switch of LBM-C/macros.cu:11-12 */
enum { LBMC_D2Q9 = 0, LBMC_D3Q15 = 1 };

/* ColType codes of the .lbmc file (LBM-C/infile_reader.cu:118-144) */
enum { LBMC_BGK = 0, LBMC_NTPOR = 1, LBMC_MRT = 2, LBMC_MRTPOR = 3 };

/* population storage: the reference is FP64 throughout
 * (LBM-C/data_types.cuh:7-11); the others are this library's additions.
 * FP16S/BF16S store f_i - w_i and compute in FP32. */
enum { LBMC_FP64 = 0, LBMC_FP32 = 1, LBMC_FP16S = 2, LBMC_BF16S = 3 };

/* Fixes for defects of the reference.  0 (the default) reproduces the
 * reference exactly. */
#define LBMC_FIX_D3Q15_MINV 0x1u /* D3Q15 M^-1 is all zeros (LBM-C/macros.cu:99-113) */
#define LBMC_FIX_D3Q15_MEQ 0x2u  /* (-7/3) is integer -2 (LBM-C/collision.cu:492-496) */
#define LBMC_FIX_D3Q15_ZH 0x4u   /* D3Q15 Zou-He leaves transverse momentum and, on
                                    face Z, mass (LBM-C/boundary_conditions/d3q15/
                                    d3q15_zh_defs.cu:4-92) */

/* The run-time constants of DomainConstant (LBM-C/data_types.cuh:23-43) that
 * the step uses, plus this library's storage choice. */
typedef struct lbmc_desc {
    int lattice;         /* LBMC_D2Q9 or LBMC_D3Q15 */
    int length[3];       /* Lx, Ly, Lz (Lz ignored for D2Q9) */
    int collision_type;  /* LBMC_BGK .. LBMC_MRTPOR */
    int forcing;         /* Guo body force on (the Force key) */
    int macro_bc;        /* per-node macro BC codes on (MacroBC key) */
    int micro_bc;        /* per-node micro BC codes on (MicroBC key) */
    double tau;          /* BGK relaxation time (Tau key) */
    double tau_mrt[15];  /* MRT relaxation rates S_i (TauMRT key) */
    double c_smag;       /* Smagorinsky constant; > 0 enables it (C_smag key) */
    int storage;         /* LBMC_FP64 .. LBMC_BF16S */
    unsigned int quirks; /* LBMC_FIX_* flags */
    int device;          /* HIP device ordinal (the reference hard-codes 0) */
} lbmc_desc;

typedef struct lbmc_ctx lbmc_ctx;

/* Allocate the populations (Q x N slots), macro arrays and node mask on the
 * device.  Replaces constant_loader + variable_size_allocator
 * (LBM-C/model_builder.cu:63-145). */
lbmc_status lbmc_create(const lbmc_desc *desc, lbmc_ctx **out);

/* Release everything the context owns (the reference never frees). */
lbmc_status lbmc_destroy(lbmc_ctx *ctx);

/* Message of the last failed call on ctx; ctx may be NULL for lbmc_create. */
const char *lbmc_last_error(const lbmc_ctx *ctx);

/* Per-node fields in the reference's host layout (variable_loader,
 * LBM-C/model_builder.cu:173-225): geometry = the "Porosity" field
 * (B >= 1 is solid), macro_bc / micro_bc = integer codes (used when the
 * descriptor enables them), force[d] = ForceX/Y/Z (when forcing), u[d] and
 * rho = VelocityX/Y/Z and Rho: the prescribed values at macro-BC nodes, which
 * are also the initial contents of the macro output arrays.  Any pointer may
 * be NULL (zeros).  Builds the 1-byte node mask and the class table.
 * Invalid codes (outside 0..15 for macro; outside the lattice's micro table
 * or a NULL slot of it, LBM-C/boundary_conditions/micro_bc.cuh:7-8,14-15)
 * return LBMC_ERR_INVALID. */
lbmc_status lbmc_set_fields(lbmc_ctx *ctx, const double *geometry, const int *macro_bc,
                            const int *micro_bc, const double *const *force,
                            const double *const *u, const double *rho);

/* The node mask (N bytes: class id per node) and the class table
 * (256 uint16 descriptors; *nclasses = number in use).  Descriptor bits:
 * 0 solid, 1 porous, 4-7 macro code, 8-11 micro code.  Class 0 is plain
 * fluid; the others are the distinct nonzero descriptors in increasing
 * order. */
lbmc_status lbmc_get_node_types(lbmc_ctx *ctx, uint8_t *mask, uint16_t *table, int *nclasses);

/* Bounceback link sets: links[n] bit i is set when node n is not solid and
 * its periodic neighbour n + e_i is solid.  Computed on the device. */
lbmc_status lbmc_get_links(lbmc_ctx *ctx, uint16_t *links);

/* f_i = w_i at every node (load_static_IC, LBM-C/model_builder.cu:233-248). */
lbmc_status lbmc_init_static(lbmc_ctx *ctx);

/* f_i = feq_i(rho, u) at every node (the BGK equilibrium of
 * LBM-C/collision.cu:53-59); an addition: the reference has no such IC. */
lbmc_status lbmc_init_equilibrium(lbmc_ctx *ctx, const double *rho, const double *const *u);

/* One call of iterate_kernel with time index t (LBM-C/index.cu:108 passes
 * i-1 for step i), storing rho/u when store_macros is nonzero. */
lbmc_status lbmc_step(lbmc_ctx *ctx, long long t, int store_macros);

/* nsteps calls with t = t0 .. t0+nsteps-1; macros are stored on the calls
 * where (t+1) % store_every == 0 (store_every <= 0: never). */
lbmc_status lbmc_run(lbmc_ctx *ctx, long long t0, int nsteps, int store_every);

/* As lbmc_run, timed with HIP events on the context's stream; *ms is the
 * device time of the nsteps calls. */
lbmc_status lbmc_time_steps(lbmc_ctx *ctx, long long t0, int nsteps, int store_every, float *ms);

/* The stored macro arrays (as output_macros copies them back). */
lbmc_status lbmc_get_macros(lbmc_ctx *ctx, double *rho, double *const *u);

/* One node's stored rho and u (screen_mess, LBM-C/index.cu:412-439). */
lbmc_status lbmc_get_node(lbmc_ctx *ctx, long long idx, double *rho, double *u);

/* Raw population slots f[i][slot] as doubles (w_i + g_i for shifted
 * storage), for slot-by-slot comparison and checkpoints. */
lbmc_status lbmc_get_f(lbmc_ctx *ctx, double *const *f);
lbmc_status lbmc_set_f(lbmc_ctx *ctx, const double *const *f);

/* lbmc_get_f without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * between graph captures, tight serving or checkpoint loops, harnesses that
 * forbid hipMalloc).  lbmc_get_f allocates an N-double conversion buffer on
 * every call; this variant converts through a fixed 8 MiB device array
 * (1,048,576 doubles, one per GPU for the life of the process), one chunk of
 * at most that many slots of one direction at a time, so lattices of any
 * size are read with the same results, bit for bit.  Calls that overlap on
 * one GPU (other threads, other contexts on the same device) share the array
 * and race: serialise them. */
lbmc_status lbmc_get_f_noalloc(lbmc_ctx *ctx, double *const *f);

/* E = sum_n 1/2 rho |u|^2 over the stored macro arrays (the energy functor
 * and thrust::reduce of LBM-C/index.cu:315-370), accumulated in FP64 in a
 * fixed order. */
lbmc_status lbmc_kinetic_energy(lbmc_ctx *ctx, double *energy);

/* Wait for all work on the context's stream. */
lbmc_status lbmc_sync(lbmc_ctx *ctx);

/* Information: node count, Q, bytes of one population slot. */
lbmc_status lbmc_info(const lbmc_ctx *ctx, long long *nodes, int *q, int *slot_bytes);

/* Plain device-to-device copy of nbytes (multiple of 16) inside a scratch
 * allocation, timed like lbmc_time_steps; the bandwidth reference for the
 * roofline and the HBM-counter calibration.  Not part of the reference. */
lbmc_status lbmc_copy_bandwidth(int device, size_t nbytes, int reps, float *ms);

/* ---- Learned relaxation-time closure --------------------------------------
 *
 * Replaces the reference's Smagorinsky hook turbulent_viscosity
 * (LBM-C/collision.cu:445-463, called by bgk_collision at :60-62) with a
 * 7 -> 64 -> 64 -> 1 GELU perceptron on the matrix cores (bf16 operands,
 * FP32 accumulation; the evaluation optionally with fp16 operands).  Inputs per node: Pi_ab / rho for ab = xx,
 * yy, zz, xy, yz, xz with Pi = sum_i e_i e_i (f_i - feq_i), then rho; each
 * normalised as (x - mu) * isd.  Output y gives tau' = out_a * y + out_b;
 * outside [tau_lo, tau_hi] the node falls back to the Smagorinsky tau
 * (c_smag of the descriptor; tau when c_smag is 0).  The reference has no
 * learned closure and no trainer: these entry points are additions.
 *
 * Parameters are one float array of LBMC_MLP_NPARAM values:
 *   w1[64][8]  first layer, column 7 is its bias (the input "1")
 *   w2[64][64] second layer, row = output unit; b2[64]
 *   w3[64], b3 output layer
 */
#define LBMC_MLP_NPARAM 4737

typedef struct lbmc_mlp lbmc_mlp;

typedef struct lbmc_mlp_scaling {
    float mu[7], isd[7];  /* input normalisation */
    float out_a, out_b;   /* tau' = out_a * y + out_b */
} lbmc_mlp_scaling;

/* Trainer / model on one device: FP32 master weights, Adam state, and the
 * operand images the kernels read (bf16, and an fp16 evaluation image).
 * sc may be null (identity). */
lbmc_status lbmc_mlp_create(int device, const float *params, const lbmc_mlp_scaling *sc, lbmc_mlp **out);
void lbmc_mlp_destroy(lbmc_mlp *m);
const char *lbmc_mlp_last_error(const lbmc_mlp *m);
lbmc_status lbmc_mlp_set_params(lbmc_mlp *m, const float *params); /* resets Adam */
lbmc_status lbmc_mlp_get_params(lbmc_mlp *m, float *params);
lbmc_status lbmc_mlp_set_scaling(lbmc_mlp *m, const lbmc_mlp_scaling *sc);

/* Network outputs y[S] for raw features feat[S][7] (host arrays), by the
 * same device routine the closure step uses. */
lbmc_status lbmc_mlp_forward(lbmc_mlp *m, const float *feat, long long S, float *y);

/* Evaluation with fp16 operands (on != 0) instead of bf16 (the default) in
 * lbmc_mlp_forward and in the closure steps of lattices that attach the
 * model afterwards (lbmc_set_closure reads it; a batch takes it from its
 * members).  Same tiles and FP32 accumulation; fp16 has 3 more mantissa bits
 * and its activations cost 8 instead of 10 vector instructions per pair;
 * pre-activations beyond 65504 send tau' out of its band (the Smagorinsky
 * fallback).  No counterpart in the reference (the closure is an addition). */
lbmc_status lbmc_mlp_set_eval_fp16(lbmc_mlp *m, int on);

/* Mean squared error of y against target[S] and its gradient (one batch,
 * S <= 2^24), by the training kernels; grad may be null. */
lbmc_status lbmc_mlp_loss_grad(lbmc_mlp *m, const float *feat, const float *target, long long S, float *grad,
                               double *loss);

/* Training data kept on the device (S < 2^31 samples). */
lbmc_status lbmc_mlp_set_data(lbmc_mlp *m, const float *feat, const float *target, long long S);

/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over shuffled minibatches of the
 * resident data; epoch_loss[e] (may be null) is the mean squared error seen
 * during epoch e.  The shuffle is seeded. */
lbmc_status lbmc_mlp_train(lbmc_mlp *m, int epochs, int batch, float lr, unsigned long long seed,
                           double *epoch_loss);

/* Turn the closure on for a lattice (BGK or NTPOR collision, FP32 / FP16S /
 * BF16S storage, same device as the model), with the model's current
 * weights, or off with model = null. */
lbmc_status lbmc_set_closure(lbmc_ctx *ctx, const lbmc_mlp *model, double tau_lo, double tau_hi);

/* Evaluate the network on calls t with t % every == 0 (and on the first call
 * after any change of state: set_fields, init_*, set_f, set_closure) and
 * relax the calls in between with each node's recorded tau (4 bytes per
 * node per call; nodes whose last evaluation fell back use the Smagorinsky
 * tau of the current state).  every = 1 (the default) evaluates every call. */
lbmc_status lbmc_set_closure_interval(lbmc_ctx *ctx, int every);

/* Shift dtau added to the network's tau' on this lattice before the band
 * check: tau' = out_a * y + out_b + dtau (0 by default).  Lets one model,
 * trained around a base tau, serve lattices of other base taus (batch
 * members with their own tau, for instance). */
lbmc_status lbmc_set_closure_shift(lbmc_ctx *ctx, double dtau);

/* ---- Batches of lattices ---------------------------------------------------
 *
 * Steps an ensemble of independent lattices of different sizes together:
 * one persistent launch per time step covers every member's nodes, and runs
 * are replayed from a captured HIP graph.  The reference steps one lattice
 * per process (LBM-C/index.cu:102-108 calling iterate, :284-313); a batch is
 * an addition.  Members are contexts made by lbmc_create and set up as usual
 * (fields and initial state); they share lattice type, storage and device,
 * may differ in size, collision type, forcing, tau and boundary codes, and
 * may not use sf boundaries or direction arrays of 4 GiB or more
 * (LBMC_ERR_UNSUPPORTED).  Members with the learned closure share one
 * model and band; each keeps its own evaluation interval, on the schedule
 * of lbmc_run.  Members must outlive the batch; their state is read at every
 * lbmc_batch_run, and lbmc_get_* on a member reads what the batch computed.
 */
typedef struct lbmc_batch lbmc_batch;

lbmc_status lbmc_batch_create(lbmc_ctx *const *members, int count, lbmc_batch **out);
lbmc_status lbmc_batch_destroy(lbmc_batch *b);
const char *lbmc_batch_last_error(const lbmc_batch *b);

/* nsteps calls t0 .. t0+nsteps-1 of every member, macros stored when
 * (t + 1) % store_every == 0 (never for store_every <= 0); synchronous. */
lbmc_status lbmc_batch_run(lbmc_batch *b, long long t0, int nsteps, int store_every);

/* The same, timed with HIP events on the batch's stream (ms total). */
lbmc_status lbmc_batch_time_steps(lbmc_batch *b, long long t0, int nsteps, int store_every, float *ms);

/* Times candidate launch shapes of the batch step kernel over
 * steps_per_trial calls each, starting at call t0 (the members advance
 * 2 * steps_per_trial calls per candidate), and keeps the fastest: chunks
 * of 256, 512, 1024 or 2048 nodes with one block per chunk and the resident
 * blocks per CU uncapped or capped at 4 or 6, and persistent grids of 2, 4,
 * 8, 16 blocks per CU over 256-node chunks. */
lbmc_status lbmc_batch_autotune(lbmc_batch *b, long long t0, int steps_per_trial, int *grid, float *ms_per_step);
/* The batch step's launch shape.  grid: blocks (at most one per chunk; a
 * smaller grid strides over the chunks); chunk_nodes: nodes of one member
 * that a block takes at a time, a multiple of 256 up to 16384;
 * blocks_per_cu: cap on the blocks resident per CU, 3..16, or 0 for none
 * (ignored when members use the learned closure). */
lbmc_status lbmc_batch_set_shape(lbmc_batch *b, int grid, int chunk_nodes, int blocks_per_cu);
/* Shape report: nodes over all members, grid, chunks (entries of the chunk
 * list: every member's chunks of chunk_nodes nodes and, for chunks of 256
 * or 512 nodes, the padding of the XCD-aware order, a few entries that the
 * kernels skip), chunk_nodes and the residency cap.  Any pointer may be
 * NULL. */
lbmc_status lbmc_batch_info(const lbmc_batch *b, long long *nodes, int *grid, int *chunks, int *chunk_nodes,
                            int *blocks_per_cu);

#ifdef __cplusplus
}
#endif
#endif
