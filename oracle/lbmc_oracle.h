/*
 * lbmc_oracle.h -- FP64 CPU restatement of LBM-C's time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links, calls or
 * falls back to it.
 *
 * Every function restates the reference's algorithm in the reference's
 * own floating-point operation order (compiled with -ffp-contract=off), so
 * that an FP64 device path written in the same order can be compared bit
 * for bit.  The per-function citations live in lbmc_oracle.c.
 */
#ifndef LBMC_ORACLE_H
#define LBMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXQ 15

/* quirk-fix flags; identical values to the product header's LBMC_FIX_* */
#define ORC_FIX_D3Q15_MINV 0x1u
#define ORC_FIX_D3Q15_MEQ 0x2u
#define ORC_FIX_D3Q15_ZH 0x4u

typedef struct orc_model orc_model;

/* lattice tables: e[3*Q] (row d, column i), w[Q], opp[Q], M[Q*Q], Minv[Q*Q].
 * Returns Q, or -1 for an unknown lattice. */
int orc_lattice_tables(int lattice, unsigned quirks, int *e, double *w, int *opp,
                       double *M, double *Minv);

orc_model *orc_create(int lattice, const int *length, int collision_type, int forcing,
                      int macro_bc, int micro_bc, double tau, const double *tau_mrt,
                      double c_smag, unsigned quirks);
void orc_destroy(orc_model *m);
int orc_set_fields(orc_model *m, const double *geometry, const int *macro_bc,
                   const int *micro_bc, double *const *force, double *const *u,
                   const double *rho);
void orc_init_static(orc_model *m);
void orc_init_equilibrium(orc_model *m, const double *rho, double *const *u);
void orc_step(orc_model *m, long long t, int store_macros);
void orc_run(orc_model *m, long long t0, int nsteps, int store_every);
void orc_get_f(const orc_model *m, double *const *f);
void orc_set_f(orc_model *m, double *const *f);
void orc_get_macros(const orc_model *m, double *rho, double *const *u);
int orc_node_types(const orc_model *m, uint8_t *mask, uint16_t *table, int *nclasses);
void orc_links(const orc_model *m, uint16_t *links);
double orc_kinetic_energy(const orc_model *m);
int orc_threads(void);
void orc_set_threads(int n);

/* single-node operators, for the known-answer tests.
 * code: the reference's collision code 0..8 (2*ColType + forcing, 8 = bounceback).
 * f[Q] is updated in place; rho/u are the node's moments (updated in place as
 * the reference's Node is); tau is in/out (Smagorinsky modifies it). */
void orc_node_collide(const orc_model *m, int code, double *f, double *rho, double *u,
                      const double *F, double B, double *tau);
/* micro BC on one node: code = the reference's micro_bc code. For sf codes the
 * neighbour's post-stream populations are passed in nb[Q]. */
void orc_node_micro(const orc_model *m, int code, double *f, double *rho, double *u,
                    const double *nb);
void orc_node_macro(const orc_model *m, int code, double *rho, double *u,
                    double prho, const double *pu);
void orc_equilibrium(const orc_model *m, double rho, const double *u, double *feq);

#ifdef __cplusplus
}
#endif
#endif
