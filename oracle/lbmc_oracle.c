/*
 * lbmc_oracle.c -- FP64 CPU restatement of LBM-C's time-step path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Never
 * linked into, called by, or used as a fallback for the product path.
 *
 * What it restates (reference paths relative to the LBM-C repository):
 *   - lattice tables ............ LBM-C/macros.cu:14-125
 *   - the step kernel ........... LBM-C/solver.cu:10-136 (one node per loop
 *                                  iteration, same operation order)
 *   - collision operators ....... LBM-C/collision.cu:23-503
 *   - macro BCs ................. LBM-C/boundary_conditions/macro_bc.cu:12-125
 *   - micro BCs ................. LBM-C/boundary_conditions/micro_bc.cuh:4-16,
 *                                  d2q9/d2q9_zh_defs.cu:4-27,
 *                                  d3q15/d3q15_zh_defs.cu:4-92,
 *                                  d2q9/d2q9_sf_defs.cu, d3q15/d3q15_sf_defs.cu
 *   - static initial condition .. LBM-C/model_builder.cu:233-248
 *   - kinetic energy ............ LBM-C/index.cu:315-370
 *
 * Deliberate deviations (each one also exists in the product library):
 *   - q_bar in the Smagorinsky model starts at zero (collision.cu:447 leaves
 *     it uninitialised, which is undefined behaviour).
 *   - "sf" copy conditions copy the neighbour's current post-stream
 *     populations of the directions entering through the face (the intent of
 *     d2q9_sf_defs.cu:4-46 / d3q15_sf_defs.cu:5-81; the reference reads
 *     unshifted slots, which under index-shift streaming hold other nodes'
 *     data and race with their writes).  Neighbours wrap periodically.
 *   - a zero density at a non-solid node is computed through, as the
 *     reference does (it produces inf/NaN there).
 * Optional fixes (quirk flags, default off = reference behaviour):
 *   ORC_FIX_D3Q15_MINV: D3Q15 M^-1 = M^T D^-1 instead of the all-zero table
 *                       that integer division yields (macros.cu:99-113).
 *   ORC_FIX_D3Q15_MEQ : -7/3 instead of the integer -2 (collision.cu:492-496).
 *   ORC_FIX_D3Q15_ZH  : Zou-He faces that close mass and all three momentum
 *                       components (d3q15_zh_defs.cu:4-92 do not).
 *
 * Build with -ffp-contract=off: the point is the reference's exact rounding.
 *
 * Every function is pinned by closed forms, invariants or brute force in
 * tests/test_oracle_*.py, and planted mistakes (ORC_MUTANT) show the pins
 * bite.  Parity unpinned against the reference itself: the sf deviation
 * above (only its intended semantics are pinned, by invariant) and the
 * summation order of the kinetic energy (the reference sums with Thrust;
 * the value agrees to rounding).
 */
#include "lbmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int lattice, dim, q;
    int e[3][ORC_MAXQ];
    double w[ORC_MAXQ];
    int opp[ORC_MAXQ];
    double M[ORC_MAXQ][ORC_MAXQ];
    double Minv[ORC_MAXQ][ORC_MAXQ];
} orc_lattice;

struct orc_model {
    orc_lattice lat;
    int L[3];
    long n;
    int collision_type, forcing, macro_on, micro_on;
    double tau, tau_mrt[ORC_MAXQ], c_smag;
    unsigned quirks;
    double *f[ORC_MAXQ];
    double *geometry, *force[3], *u[3], *rho;
    int *macro_bc, *micro_bc;
    double *pre; /* post-stream copy, only when sf nodes exist */
    int has_sf;
};

/* Node state, as data_types.cuh:45-55 */
typedef struct {
    long ixd;
    double f[ORC_MAXQ];
    double rho;
    double u[3];
    double F[3];
    double B;
    double c_smag;
    int coord[3];
} onode;

/* ------------------------------------------------------------------ */
/* lattice tables: macros.cu:14-58 (D2Q9) and :60-125 (D3Q15)          */
/* ------------------------------------------------------------------ */

static const int d2q9_e[2][9] = {{0, 1, 0, -1, 0, 1, -1, -1, 1},
                                 {0, 0, 1, 0, -1, 1, 1, -1, -1}};
static const int d2q9_opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const int d2q9_M[9][9] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2},
    {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
    {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
    {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 1, -1, 1, -1}};
/* The reference ships M^-1 for D2Q9 as a numeric table (macros.cu:31-43);
 * the values are data that must match for bitwise parity. */
static const double d2q9_Minv[9][9] = {
    {0.1111111111111111, -0.11111111111111112, 0.1111111111111111, 0, 0, 0, 0, 0, 0},
    {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, 0.16666666666666666,
     -0.16666666666666669, 0, 0, 0.25, 0},
    {0.1111111111111111, -0.027777777777777762, -0.055555555555555539, 0,
     -1.3877787807814457e-017, 0.16666666666666666, -0.16666666666666669, -0.25, 0},
    {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, -0.16666666666666666,
     0.16666666666666669, 0, 0, 0.25, 0},
    {0.1111111111111111, -0.02777777777777779, -0.055555555555555566, 0,
     1.3877787807814457e-017, -0.16666666666666666, 0.16666666666666669, -0.25, 0},
    {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, 0.16666666666666666,
     0.083333333333333329, 0.16666666666666666, 0.083333333333333329, 0, 0.25},
    {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, -0.16666666666666666,
     -0.083333333333333329, 0.16666666666666666, 0.083333333333333329, 0, -0.25},
    {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, -0.16666666666666666,
     -0.083333333333333329, -0.16666666666666666, -0.083333333333333329, 0, 0.25},
    {0.1111111111111111, 0.055555555555555552, 0.027777777777777776, 0.16666666666666666,
     0.083333333333333329, -0.16666666666666666, -0.083333333333333329, 0, -0.25}};

static const int d3q15_e[3][15] = {
    {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
    {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
    {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1}};
static const int d3q15_opp[15] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};
static const int d3q15_M[15][15] = {
    {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
    {-2, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1},
    {16, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1},
    {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
    {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1},
    {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
    {0, 0, 0, -4, 4, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
    {0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1},
    {0, 0, 0, 0, 0, -4, 4, 1, 1, 1, 1, -1, -1, -1, -1},
    {0, 2, 2, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0},
    {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1},
    {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, -1, -1, 1, 1},
    {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1},
    {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1}};

static void load_lattice(orc_lattice *lt, int lattice, unsigned quirks)
{
    memset(lt, 0, sizeof(*lt));
    lt->lattice = lattice;
    if (lattice == 0) {
        lt->dim = 2;
        lt->q = 9;
        for (int i = 0; i < 9; i++) {
            lt->e[0][i] = d2q9_e[0][i];
            lt->e[1][i] = d2q9_e[1][i];
            lt->opp[i] = d2q9_opp[i];
            for (int j = 0; j < 9; j++) {
                lt->M[i][j] = (double)d2q9_M[i][j];
                lt->Minv[i][j] = d2q9_Minv[i][j];
            }
        }
        /* LOAD_OMEGA, macros.cu:52 */
        lt->w[0] = 4.0 / 9.0;
        for (int i = 1; i < 5; i++) lt->w[i] = 1.0 / 9.0;
        for (int i = 5; i < 9; i++) lt->w[i] = 1.0 / 36.0;
    } else {
        lt->dim = 3;
        lt->q = 15;
        for (int i = 0; i < 15; i++) {
            for (int d = 0; d < 3; d++) lt->e[d][i] = d3q15_e[d][i];
            lt->opp[i] = d3q15_opp[i];
            for (int j = 0; j < 15; j++) lt->M[i][j] = (double)d3q15_M[i][j];
        }
        /* macros.cu:99-113 writes M^-1 as integer fractions (1/15, -1/9, ...),
         * every one of which is 0 in C.  With the fix flag, M^-1 = M^T D^-1
         * (M has orthogonal rows; D = diag of squared row norms), which is the
         * table those fractions spell. */
        if (quirks & ORC_FIX_D3Q15_MINV) {
            for (int k = 0; k < 15; k++) {
                long nrm = 0;
                for (int j = 0; j < 15; j++) nrm += (long)d3q15_M[k][j] * d3q15_M[k][j];
                for (int i = 0; i < 15; i++)
                    lt->Minv[i][k] = (double)d3q15_M[k][i] / (double)nrm;
            }
        }
        /* LOAD_OMEGA, macros.cu:117 */
        lt->w[0] = 2.0 / 9.0;
        for (int i = 1; i < 7; i++) lt->w[i] = 1.0 / 9.0;
        for (int i = 7; i < 15; i++) lt->w[i] = 1.0 / 72.0;
    }
}

int orc_lattice_tables(int lattice, unsigned quirks, int *e, double *w, int *opp, double *M,
                       double *Minv)
{
    if (lattice != 0 && lattice != 1) return -1;
    orc_lattice lt;
    load_lattice(&lt, lattice, quirks);
    int q = lt.q;
    for (int d = 0; d < 3; d++)
        for (int i = 0; i < q; i++)
            if (e) e[d * q + i] = lt.e[d][i];
    for (int i = 0; i < q; i++) {
        if (w) w[i] = lt.w[i];
        if (opp) opp[i] = lt.opp[i];
        for (int j = 0; j < q; j++) {
            if (M) M[i * q + j] = lt.M[i][j];
            if (Minv) Minv[i * q + j] = lt.Minv[i][j];
        }
    }
    return q;
}

/* ------------------------------------------------------------------ */
/* collision operators: collision.cu                                   */
/* ------------------------------------------------------------------ */

/* u_square, collision.cu:23-34 */
static double u_square(const orc_lattice *lt, const onode *nd)
{
    double v = 0;
    for (int d = 0; d < lt->dim; d++) v += (nd->u[d] * nd->u[d]);
    return v * 1.5;
}

/* ORC_MUTANT (default 0) plants one deliberate mistake per value, for
 * tests/test_oracle_pins_bite.py, which builds each mutant into a scratch
 * directory and checks that the pin written for that function fails on it.
 * The library built by oracle/Makefile has none. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

/* e_mul_u, collision.cu:36-47 */
static double e_mul_u(const orc_lattice *lt, const onode *nd, int i)
{
    double v = 0;
    for (int d = 0; d < lt->dim; d++) v += (lt->e[d][i] * nd->u[d]);
    return v * 3.;
}

/* the BGK equilibrium, collision.cu:55-59 */
static void equilibrium(const orc_lattice *lt, const onode *nd, double *feq)
{
    double usq = u_square(lt, nd);
    for (int i = 0; i < lt->q; i++) {
        double eu = e_mul_u(lt, nd, i);
#if ORC_MUTANT == 1
        feq[i] = nd->rho * lt->w[i] * (1.0 + eu - usq); /* second-order term dropped */
#else
        feq[i] = nd->rho * lt->w[i] * (1.0 + eu + (0.5 * eu * eu) - usq);
#endif
    }
}

/* turbulent_viscosity, collision.cu:445-463 (q_bar zero-initialised) */
static void turbulent_viscosity(const orc_lattice *lt, const onode *nd, const double *feq,
                                double *tau)
{
    double q_bar[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double q_hat = 0.;
    for (int a = 0; a < lt->dim; a++) {
        for (int b = 0; b < lt->dim; b++) {
            for (int q = 0; q < lt->q; q++)
                q_bar[a][b] = q_bar[a][b] + ((double)lt->e[a][q] * (double)lt->e[b][q] *
                                             (nd->f[q] - feq[q]));
            q_hat = q_hat + sqrt((double)2 * q_bar[a][b] * q_bar[a][b]);
        }
    }
    /* POW4(1/sqrt(3)) expands textually, so the divisor is
     * rho*1/s*1/s*1/s*1/s evaluated left to right (collision.cu:16,462). */
    double s3 = sqrt((double)3);
    double den = nd->rho * 1 / s3 * 1 / s3 * 1 / s3 * 1 / s3;
    double c = nd->c_smag;
#if ORC_MUTANT == 3
    *tau = *tau + 0.5 * (sqrt(((*tau) * (*tau)) + (sqrt((double)2) * (c * c) * (1 / den) * q_hat)) - *tau);
#else
    *tau = *tau + 0.5 * (sqrt(((*tau) * (*tau)) + (2 * sqrt((double)2) * (c * c) * (1 / den) * q_hat)) - *tau);
#endif
}

/* Guo source term, collision.cu:82-99 (uses the node's shifted u and the
 * relaxation time before any Smagorinsky update) */
static void guo_source(const orc_lattice *lt, const onode *nd, double tau, double *ft)
{
    for (int i = 0; i < lt->q; i++) {
        double Fc[3];
        for (int d = 0; d < lt->dim; d++) {
            double eu;
            if (lt->dim > 2)
                eu = (lt->e[0][i] * nd->u[0]) + (lt->e[1][i] * nd->u[1]) + (lt->e[2][i] * nd->u[2]);
            else
                eu = (lt->e[0][i] * nd->u[0]) + (lt->e[1][i] * nd->u[1]);
#if ORC_MUTANT == 2
            Fc[d] = lt->w[i] * (1 - (1 / (2 * tau))) *
                    (((lt->e[d][i] - nd->u[d]) * 3) + (lt->e[d][i] * 3 * eu)); /* 9 -> 3 */
#else
            Fc[d] = lt->w[i] * (1 - (1 / (2 * tau))) *
                    (((lt->e[d][i] - nd->u[d]) * 3) + (lt->e[d][i] * 9 * eu));
#endif
        }
        ft[i] = 0;
        for (int d = 0; d < lt->dim; d++) ft[i] += Fc[d] * nd->F[d];
    }
}

/* u <- u + F/(2 rho), collision.cu:74-78 */
static void guo_shift(const orc_lattice *lt, onode *nd)
{
#if ORC_MUTANT == 6
    for (int d = 0; d < lt->dim; d++) nd->u[d] = nd->u[d] + (1 / nd->rho) * nd->F[d]; /* F/rho */
#else
    for (int d = 0; d < lt->dim; d++) nd->u[d] = nd->u[d] + (1 / (2 * nd->rho)) * nd->F[d];
#endif
}

/* meq_d2q9 / meq_d3q15, collision.cu:465-503 */
static void moment_equilibrium(const orc_lattice *lt, unsigned quirks, const onode *nd,
                               double *meq)
{
    if (lt->dim == 2) {
        double jx = nd->rho * nd->u[0];
        double jy = nd->rho * nd->u[1];
        meq[0] = nd->rho;
#if ORC_MUTANT == 7
        meq[1] = (2 * nd->rho) + 3 * (jx * jx + jy * jy); /* sign of rho */
#else
        meq[1] = (-2 * nd->rho) + 3 * (jx * jx + jy * jy);
#endif
        meq[2] = nd->rho - 3 * (jx * jx + jy * jy);
        meq[3] = jx;
        meq[4] = -jx;
        meq[5] = jy;
        meq[6] = -jy;
        meq[7] = jx * jx - jy * jy;
        meq[8] = jx * jy;
    } else {
        double jx = nd->rho * nd->u[0];
        double jy = nd->rho * nd->u[1];
        double jz = nd->rho * nd->u[2];
        /* (-7/3) is integer division in the reference: -2 */
        double qc = (quirks & ORC_FIX_D3Q15_MEQ) ? (-7.0 / 3.0) : (double)(-2);
        meq[0] = nd->rho;
        meq[1] = (-1 * nd->rho) + (jx * jx + jy * jy + jz * jz);
        meq[2] = nd->rho - 5 * (jx * jx + jy * jy + jz * jz);
        meq[3] = jx;
        meq[4] = qc * jx;
        meq[5] = jy;
        meq[6] = qc * jy;
        meq[7] = jz;
        meq[8] = qc * jz;
        meq[9] = (2 * jx * jx - (jy * jy + jz * jz));
        meq[10] = jy * jy - jz * jz;
        meq[11] = jx * jy;
        meq[12] = jy * jz;
        meq[13] = jx * jz;
        meq[14] = 0;
    }
}

/* m = S (M f - meq), collision.cu:204-216 */
static void mrt_moments(const orc_model *m, const onode *nd, const double *meq, double *mm)
{
    const orc_lattice *lt = &m->lat;
    for (int i = 0; i < lt->q; i++) {
        mm[i] = 0;
        for (int j = 0; j < lt->q; j++) mm[i] = mm[i] + lt->M[i][j] * nd->f[j];
    }
    for (int i = 0; i < lt->q; i++) mm[i] = m->tau_mrt[i] * (mm[i] - meq[i]);
}

static double minv_row(const orc_lattice *lt, int i, const double *mm)
{
    double acc = 0;
    for (int j = 0; j < lt->q; j++) acc = acc + lt->Minv[i][j] * mm[j];
    return acc;
}

/* The nine collision functions of collision.cu:19-21, selected by code. */
static void collide(const orc_model *m, int code, onode *nd, double *tau)
{
    const orc_lattice *lt = &m->lat;
    const int Q = lt->q;
    double feq[ORC_MAXQ], ft[ORC_MAXQ], tmp[ORC_MAXQ], meq[ORC_MAXQ], mm[ORC_MAXQ];
    switch (code) {
    case 0: /* bgk_collision, collision.cu:49-67 */
        equilibrium(lt, nd, feq);
        if (nd->c_smag > 0) turbulent_viscosity(lt, nd, feq, tau);
        for (int i = 0; i < Q; i++) nd->f[i] = nd->f[i] - (1.0 / (*tau)) * (nd->f[i] - feq[i]);
        break;
    case 1: /* bgk_guo_collision, collision.cu:69-111 */
        guo_shift(lt, nd);
        guo_source(lt, nd, *tau, ft);
        equilibrium(lt, nd, feq);
        if (nd->c_smag > 0) turbulent_viscosity(lt, nd, feq, tau);
        for (int i = 0; i < Q; i++)
            nd->f[i] = nd->f[i] - (1.0 / (*tau)) * (nd->f[i] - feq[i]) + ft[i];
        break;
    case 2: /* bgk_ntpor_collision, collision.cu:113-138 */
        equilibrium(lt, nd, feq);
        if (nd->c_smag > 0) turbulent_viscosity(lt, nd, feq, tau);
        for (int i = 0; i < Q; i++) {
            double cb = (1.0 / (*tau)) * (nd->f[i] - feq[i]);
            double cs = nd->f[lt->opp[i]] - nd->f[i];
            tmp[i] = nd->f[i] - (1 - (nd->B)) * cb + (nd->B) * cs;
        }
        for (int i = 0; i < Q; i++) nd->f[i] = tmp[i];
        break;
    case 3: /* bgk_ntpor_guo_collision, collision.cu:140-190 */
        guo_shift(lt, nd);
        guo_source(lt, nd, *tau, ft);
        equilibrium(lt, nd, feq);
        if (nd->c_smag > 0) turbulent_viscosity(lt, nd, feq, tau);
        for (int i = 0; i < Q; i++) {
            double cb = (1.0 / (*tau)) * (nd->f[i] - feq[i]);
            double cs = nd->f[lt->opp[i]] - nd->f[i];
            tmp[i] = nd->f[i] - (1 - (nd->B)) * (cb) + (nd->B) * cs + (1 - (nd->B)) * ft[i];
        }
        for (int i = 0; i < Q; i++) nd->f[i] = tmp[i];
        break;
    case 4: /* mrt_collision, collision.cu:192-230 */
        moment_equilibrium(lt, m->quirks, nd, meq);
        mrt_moments(m, nd, meq, mm);
        for (int i = 0; i < Q; i++) nd->f[i] = nd->f[i] - minv_row(lt, i, mm);
        break;
    case 5: /* mrt_guo_collision, collision.cu:232-303 */
        guo_shift(lt, nd);
        guo_source(lt, nd, *tau, ft);
        moment_equilibrium(lt, m->quirks, nd, meq);
        mrt_moments(m, nd, meq, mm);
        for (int i = 0; i < Q; i++) nd->f[i] = nd->f[i] - minv_row(lt, i, mm) + ft[i];
        break;
    case 6: /* mrt_ntpor_collision, collision.cu:305-345: f is updated in place,
             * so f[opp[i]] is already post-collision when opp[i] < i */
        moment_equilibrium(lt, m->quirks, nd, meq);
        mrt_moments(m, nd, meq, mm);
        for (int i = 0; i < Q; i++) {
            double acc = minv_row(lt, i, mm);
            double cs = nd->f[lt->opp[i]] - nd->f[i];
            nd->f[i] = nd->f[i] - (1 - (nd->B)) * acc + (nd->B) * cs;
        }
        break;
    case 7: /* mrt_ntpor_guo_collision, collision.cu:347-421 (same in-place order) */
        guo_shift(lt, nd);
        guo_source(lt, nd, *tau, ft);
        moment_equilibrium(lt, m->quirks, nd, meq);
        mrt_moments(m, nd, meq, mm);
        for (int i = 0; i < Q; i++) {
            double acc = minv_row(lt, i, mm);
            double cs = nd->f[lt->opp[i]] - nd->f[i];
            nd->f[i] = nd->f[i] - (1 - (nd->B)) * acc + (1 - (nd->B)) * ft[i] + (nd->B) * cs;
        }
        break;
    default: /* bounceback, collision.cu:423-443 */
        for (int i = 0; i < Q; i++) tmp[i] = nd->f[i];
#if ORC_MUTANT == 5
        for (int i = 0; i < Q; i++) nd->f[i] = tmp[i]; /* no reflection */
#else
        for (int i = 0; i < Q; i++) nd->f[i] = tmp[lt->opp[i]];
#endif
        nd->u[0] = 0;
        nd->u[1] = 0;
        if (lt->dim > 2) nd->u[2] = 0;
        nd->rho = 0;
        break;
    }
}

/* ------------------------------------------------------------------ */
/* macro BCs: macro_bc.cu:12-125                                       */
/* ------------------------------------------------------------------ */

/* which of (rho, ux, uy, uz) each code 1..15 prescribes, in the listing order
 * of macro_conditions[] (macro_bc.cu:12-17) */
static const unsigned char macro_sets[16][4] = {
    {0, 0, 0, 0}, /* 0: none */
    {0, 1, 0, 0}, /* presc_ux */
    {0, 0, 1, 0}, /* presc_uy */
    {0, 0, 0, 1}, /* presc_uz */
    {0, 1, 1, 0}, /* presc_uy_ux */
    {0, 1, 0, 1}, /* presc_uz_ux */
    {0, 0, 1, 1}, /* presc_uz_uy */
    {0, 1, 1, 1}, /* presc_ux_uy_uz */
    {1, 0, 0, 0}, /* presc_rho */
    {1, 1, 0, 0}, /* presc_rho_ux */
    {1, 0, 1, 0}, /* presc_rho_uy */
    {1, 0, 0, 1}, /* presc_rho_uz */
    {1, 1, 1, 0}, /* presc_rho_ux_uy */
    {1, 1, 0, 1}, /* presc_rho_ux_uz */
    {1, 0, 1, 1}, /* presc_rho_uy_uz */
    {1, 1, 1, 1}, /* presc_rho_ux_uy_uz */
};

static void apply_macro(const orc_lattice *lt, int code, double *rho, double *u, double prho,
                        const double *pu)
{
    const unsigned char *s = macro_sets[code];
    if (s[0]) *rho = prho;
    for (int d = 0; d < lt->dim; d++)
        if (s[1 + d]) u[d] = pu[d];
}

/* ------------------------------------------------------------------ */
/* micro BCs                                                           */
/* ------------------------------------------------------------------ */

/* zh_pressure_x / zh_pressure_X, d2q9_zh_defs.cu:4-27 */
static void zh_d2q9(int east, onode *nd)
{
    double *f = nd->f;
    double r = nd->rho;
    if (!east) {
        nd->u[0] = 1.0 - ((1.0 / r) * (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])));
        nd->u[1] = 0.0;
        f[1] = f[3] + ((2.0 / 3.0) * r * nd->u[0]);
#if ORC_MUTANT == 4
        f[5] = f[7] + ((1.0 / 2.0) * (f[2] - f[4])) + ((1.0 / 6.0) * r * nd->u[0]); /* sign */
#else
        f[5] = f[7] - ((1.0 / 2.0) * (f[2] - f[4])) + ((1.0 / 6.0) * r * nd->u[0]);
#endif
        f[8] = f[6] + ((1.0 / 2.0) * (f[2] - f[4])) + ((1.0 / 6.0) * r * nd->u[0]);
    } else {
        nd->u[0] = -1.0 + ((1.0 / r) * (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])));
        nd->u[1] = 0.0;
        f[3] = f[1] - ((2.0 / 3.0) * r * nd->u[0]);
        f[6] = f[8] - ((1.0 / 2.0) * (f[2] - f[4])) - ((1.0 / 6.0) * r * nd->u[0]);
        f[7] = f[5] + ((1.0 / 2.0) * (f[2] - f[4])) - ((1.0 / 6.0) * r * nd->u[0]);
    }
}

/* D3Q15 Zou-He faces as the reference writes them, d3q15_zh_defs.cu:4-92.
 * Each face: normal axis, side (+1 low face, -1 high face), the summation
 * order of the in-plane and outgoing sums that give u_n, the axis unknown and
 * its partner, the two transverse differences (T1 = f[ta]-f[tb],
 * T2 = f[tc]-f[td]) and four diagonal unknowns (k, partner, c1, c2) with
 *   f_k = (f_p + side*(1/12) rho u_v) - 1/4 (c1 T1 + c2 T2).
 * v is the velocity component the face's u-term uses: the normal one except
 * on face Z, where the reference uses u[0] (d3q15_zh_defs.cu:87-91). */
typedef struct {
    int axis, side, v;
    int s0[5], s1[5];
    int kn, pn;
    int ta, tb, tc, td;
    int diag[4][4];
} zh_face;

static const zh_face zh_ref_faces[6] = {
    /* x: d3q15_zh_defs.cu:4-17 */
#if ORC_MUTANT == 9
    {0, 1, 0, {0, 3, 4, 5, 6}, {2, 14, 10, 12, 8}, 1, 2, 3, 4, 5, 6,
     {{7, 14, 1, 1}, {11, 10, -1, -1}, {9, 12, -1, 1}, {13, 8, 1, -1}}},
#else
    {0, 1, 0, {0, 3, 4, 5, 6}, {2, 14, 10, 12, 8}, 1, 2, 3, 4, 5, 6,
     {{7, 14, -1, 1}, {11, 10, -1, -1}, {9, 12, 1, 1}, {13, 8, 1, -1}}},
#endif
    /* X: :19-32 */
#if ORC_MUTANT == 11
    {0, -1, 0, {0, 3, 4, 5, 6}, {1, 7, 11, 9, 13}, 2, 1, 3, 4, 5, 6,
     {{14, 7, 1, -1}, {10, 11, 1, 1}, {12, 9, -1, -1}, {8, 13, -1, -1}}},
#else
    {0, -1, 0, {0, 3, 4, 5, 6}, {1, 7, 11, 9, 13}, 2, 1, 3, 4, 5, 6,
     {{14, 7, 1, -1}, {10, 11, 1, 1}, {12, 9, -1, 1}, {8, 13, -1, -1}}},
#endif
    /* y: :34-47 */
    {1, 1, 1, {0, 1, 2, 5, 6}, {4, 14, 10, 9, 13}, 3, 4, 1, 2, 5, 6,
     {{7, 14, -1, 1}, {11, 10, -1, -1}, {12, 9, 1, -1}, {8, 13, 1, 1}}},
    /* Y: :49-62 */
    {1, -1, 1, {0, 1, 2, 5, 6}, {3, 7, 11, 12, 8}, 4, 3, 1, 2, 5, 6,
     {{14, 7, 1, -1}, {10, 11, 1, 1}, {9, 12, -1, 1}, {13, 8, -1, -1}}},
    /* z: :64-77 */
    {2, 1, 2, {0, 1, 2, 3, 4}, {6, 14, 11, 12, 13}, 5, 6, 1, 2, 3, 4,
     {{7, 14, -1, 1}, {10, 11, 1, -1}, {9, 12, -1, -1}, {8, 13, 1, 1}}},
    /* Z: :79-92 (its rho-u terms read u[0]; mutant 10 reads u[2]) */
    {2, -1, ORC_MUTANT == 10 ? 2 : 0, {0, 1, 2, 3, 4}, {5, 7, 10, 9, 8}, 6, 5, 1, 2, 3, 4,
     {{14, 7, 1, -1}, {11, 10, -1, 1}, {12, 9, 1, 1}, {13, 8, -1, -1}}},
};

static void zh_d3q15_ref(int face, onode *nd)
{
    const zh_face *z = &zh_ref_faces[face];
    double *f = nd->f;
    double r = nd->rho;
    double s0 = f[z->s0[0]] + f[z->s0[1]] + f[z->s0[2]] + f[z->s0[3]] + f[z->s0[4]];
    double s1 = f[z->s1[0]] + f[z->s1[1]] + f[z->s1[2]] + f[z->s1[3]] + f[z->s1[4]];
    double un;
    if (z->side > 0)
        un = 1.0 - ((1.0 / r) * (s0 + 2.0 * s1));
    else
        un = -1.0 + ((1.0 / r) * (s0 + 2.0 * s1));
    for (int d = 0; d < 3; d++) nd->u[d] = 0.0;
    nd->u[z->axis] = un;
    double uv = nd->u[z->v];
    double T1 = f[z->ta] - f[z->tb];
    double T2 = f[z->tc] - f[z->td];
    if (z->side > 0)
        f[z->kn] = f[z->pn] + ((2.0 / 3.0) * r * uv);
    else
        f[z->kn] = f[z->pn] - ((2.0 / 3.0) * r * uv);
    for (int k = 0; k < 4; k++) {
        const int *dg = z->diag[k];
        double t = (dg[2] > 0 ? T1 : (-1.0) * T1) + (dg[3] > 0 ? T2 : -T2);
        double a = ((1.0 / 12.0) * r * uv);
        if (z->side > 0)
            f[dg[0]] = f[dg[1]] + a - ((1.0 / 4.0) * t);
        else
            f[dg[0]] = f[dg[1]] - a - ((1.0 / 4.0) * t);
    }
}

/* Corrected D3Q15 Zou-He: known populations give u_n from mass and normal
 * momentum; the axis unknown takes (2/3) rho u_n, the diagonals (1/12) rho u_n
 * and a transverse correction -1/4 sum_t e_kt T_t with T_t the transverse
 * momentum carried by the in-plane populations, so that both transverse
 * momentum components vanish. */
static void zh_d3q15_fixed(const orc_lattice *lt, int face, onode *nd)
{
    int axis = face / 2;
    int side = (face % 2 == 0) ? 1 : -1;
    double *f = nd->f;
    double r = nd->rho;
    double s0 = 0, s1 = 0;
    for (int i = 0; i < 15; i++) {
        int ea = lt->e[axis][i];
        if (ea == 0) s0 += f[i];
        else if (ea == -side) s1 += f[i];
    }
    double un = side > 0 ? 1.0 - (s0 + 2.0 * s1) / r : -1.0 + (s0 + 2.0 * s1) / r;
    for (int d = 0; d < 3; d++) nd->u[d] = 0.0;
    nd->u[axis] = un;
    double T[3] = {0, 0, 0};
    for (int i = 0; i < 15; i++)
        if (lt->e[axis][i] == 0)
            for (int d = 0; d < 3; d++)
                if (d != axis) T[d] += lt->e[d][i] * f[i];
    for (int i = 0; i < 15; i++) {
        if (lt->e[axis][i] != side) continue;
        int p = lt->opp[i];
        int diag = 0;
        for (int d = 0; d < 3; d++)
            if (d != axis && lt->e[d][i] != 0) diag = 1;
        if (!diag) {
            f[i] = f[p] + side * (2.0 / 3.0) * r * un;
        } else {
            double t = 0;
            for (int d = 0; d < 3; d++)
                if (d != axis) t += lt->e[d][i] * T[d];
            f[i] = f[p] + side * (1.0 / 12.0) * r * un - 0.25 * t;
        }
    }
}

/* sf: copy the entering populations of face (axis, side) from nb[] */
static void sf_copy(const orc_lattice *lt, int face, onode *nd, const double *nb)
{
    int axis = face / 2;
    int side = (face % 2 == 0) ? 1 : -1;
    for (int i = 0; i < lt->q; i++)
        if (lt->e[axis][i] == side) nd->f[i] = nb[i];
}

/* micro code -> (kind, face): kind 1 Zou-He, 2 sf; 0 invalid.
 * micro_bc.cuh:7-8 (D2Q9) and :14-15 (D3Q15). */
static int micro_kind(int lattice, int code, int *face)
{
    if (lattice == 0) {
        if (code == 1 || code == 2) { *face = code - 1; return 1; }
        if (code >= 5 && code <= 8) { *face = code - 5; return 2; }
        return 0;
    }
    if (code >= 1 && code <= 6) { *face = code - 1; return 1; }
    if (code >= 7 && code <= 12) { *face = code - 7; return 2; }
    return 0;
}

static void apply_micro(const orc_model *m, int code, onode *nd, const double *nb)
{
    int face = 0;
    int kind = micro_kind(m->lat.lattice, code, &face);
    if (kind == 1) {
        if (m->lat.lattice == 0) zh_d2q9(face, nd);
        else if (m->quirks & ORC_FIX_D3Q15_ZH) zh_d3q15_fixed(&m->lat, face, nd);
        else zh_d3q15_ref(face, nd);
    } else if (kind == 2) {
        sf_copy(&m->lat, face, nd, nb);
    }
}

/* ------------------------------------------------------------------ */
/* model                                                               */
/* ------------------------------------------------------------------ */

orc_model *orc_create(int lattice, const int *length, int collision_type, int forcing,
                      int macro_bc, int micro_bc, double tau, const double *tau_mrt,
                      double c_smag, unsigned quirks)
{
    if (lattice != 0 && lattice != 1) return NULL;
    if (collision_type < 0 || collision_type > 3) return NULL;
    orc_model *m = (orc_model *)calloc(1, sizeof(orc_model));
    if (!m) return NULL;
    load_lattice(&m->lat, lattice, quirks);
    m->n = 1;
    for (int d = 0; d < 3; d++) m->L[d] = 1;
    for (int d = 0; d < m->lat.dim; d++) {
        if (length[d] <= 0) { free(m); return NULL; }
        m->L[d] = length[d];
        m->n *= length[d];
    }
    m->collision_type = collision_type;
    m->forcing = forcing ? 1 : 0;
    m->macro_on = macro_bc ? 1 : 0;
    m->micro_on = micro_bc ? 1 : 0;
    m->tau = tau;
    for (int i = 0; i < m->lat.q; i++) m->tau_mrt[i] = tau_mrt ? tau_mrt[i] : 0.0;
    m->c_smag = c_smag;
    m->quirks = quirks;
    size_t nb = sizeof(double) * (size_t)m->n;
    for (int i = 0; i < m->lat.q; i++) m->f[i] = (double *)calloc(1, nb);
    m->geometry = (double *)calloc(1, nb);
    m->rho = (double *)calloc(1, nb);
    for (int d = 0; d < m->lat.dim; d++) {
        m->u[d] = (double *)calloc(1, nb);
        m->force[d] = (double *)calloc(1, nb);
    }
    m->macro_bc = (int *)calloc((size_t)m->n, sizeof(int));
    m->micro_bc = (int *)calloc((size_t)m->n, sizeof(int));
    return m;
}

void orc_destroy(orc_model *m)
{
    if (!m) return;
    for (int i = 0; i < m->lat.q; i++) free(m->f[i]);
    for (int d = 0; d < 3; d++) {
        free(m->u[d]);
        free(m->force[d]);
    }
    free(m->geometry);
    free(m->rho);
    free(m->macro_bc);
    free(m->micro_bc);
    free(m->pre);
    free(m);
}

int orc_set_fields(orc_model *m, const double *geometry, const int *macro_bc, const int *micro_bc,
                   double *const *force, double *const *u, const double *rho)
{
    size_t nb = sizeof(double) * (size_t)m->n;
    if (geometry) memcpy(m->geometry, geometry, nb);
    else memset(m->geometry, 0, nb);
    for (long n = 0; n < m->n; n++) {
        m->macro_bc[n] = (m->macro_on && macro_bc) ? macro_bc[n] : 0;
        m->micro_bc[n] = (m->micro_on && micro_bc) ? micro_bc[n] : 0;
        /* codes at solid nodes are never used (solver.cu:68-72) */
        if (geometry && geometry[n] >= 1) {
            m->macro_bc[n] = 0;
            m->micro_bc[n] = 0;
        }
        /* an out-of-range code indexes past the reference's function tables */
        if (m->macro_bc[n] < 0 || m->macro_bc[n] > 15) return 1;
        int face;
        if (m->micro_bc[n] != 0 && !micro_kind(m->lat.lattice, m->micro_bc[n], &face)) return 1;
    }
    for (int d = 0; d < m->lat.dim; d++) {
        if (force && force[d]) memcpy(m->force[d], force[d], nb);
        else memset(m->force[d], 0, nb);
        if (u && u[d]) memcpy(m->u[d], u[d], nb);
        else memset(m->u[d], 0, nb);
    }
    if (rho) memcpy(m->rho, rho, nb);
    else memset(m->rho, 0, nb);
    m->has_sf = 0;
    for (long n = 0; n < m->n; n++) {
        int face;
        if (m->geometry[n] >= 1) continue;
        if (m->micro_bc[n] && micro_kind(m->lat.lattice, m->micro_bc[n], &face) == 2)
            m->has_sf = 1;
    }
    free(m->pre);
    m->pre = NULL;
    if (m->has_sf) m->pre = (double *)malloc(nb * m->lat.q);
    return 0;
}

/* load_static_IC, model_builder.cu:233-248 */
void orc_init_static(orc_model *m)
{
    for (int i = 0; i < m->lat.q; i++)
        for (long n = 0; n < m->n; n++) m->f[i][n] = 1.0 * m->lat.w[i];
}

void orc_init_equilibrium(orc_model *m, const double *rho, double *const *u)
{
    for (long n = 0; n < m->n; n++) {
        onode nd;
        memset(&nd, 0, sizeof(nd));
        nd.rho = rho[n];
        for (int d = 0; d < m->lat.dim; d++) nd.u[d] = u[d][n];
        double feq[ORC_MAXQ];
        equilibrium(&m->lat, &nd, feq);
        for (int i = 0; i < m->lat.q; i++) m->f[i][n] = feq[i];
    }
}

/* the slot of population i of node coord at call t, solver.cu:77-92 */
static long slot_of(const orc_model *m, const int *coord, long long t, int i)
{
    long ix[3] = {0, 0, 0};
    for (int d = 0; d < m->lat.dim; d++) {
#if ORC_MUTANT == 8
        long long v = ((long long)coord[d] + (t * m->lat.e[d][i])) % m->L[d]; /* shift sign */
#else
        long long v = ((long long)coord[d] - (t * m->lat.e[d][i])) % m->L[d];
#endif
        if (v < 0) v = m->L[d] - (v * -1);
        ix[d] = (long)v;
    }
    return ix[0] + ix[1] * (long)m->L[0] + ix[2] * (long)m->L[0] * m->L[1];
}

static void coords_of(const orc_model *m, long n, int *c)
{
    c[0] = (int)(n % m->L[0]);
    c[1] = (int)((n / m->L[0]) % m->L[1]);
    c[2] = (int)(n / ((long)m->L[0] * m->L[1]));
}

static long neighbour(const orc_model *m, const int *c, int axis, int delta)
{
    int cc[3] = {c[0], c[1], c[2]};
    cc[axis] = (cc[axis] + delta + m->L[axis]) % m->L[axis];
    return cc[0] + cc[1] * (long)m->L[0] + cc[2] * (long)m->L[0] * m->L[1];
}

/* iterate_kernel body for one node, solver.cu:19-134 */
static void update_node(orc_model *m, long n, long long t, int store)
{
    const orc_lattice *lt = &m->lat;
    const int Q = lt->q, D = lt->dim;
    onode nd;
    memset(&nd, 0, sizeof(nd));
    coords_of(m, n, nd.coord);
    nd.ixd = n;
    nd.rho = 0;
    for (int d = 0; d < D; d++) nd.u[d] = 0;

    double tau = m->tau;
    nd.c_smag = m->c_smag;
    int macro = 0, micro = 0;
    if (m->macro_on) macro = m->macro_bc[n];
    if (m->micro_on) micro = m->micro_bc[n];
    int code = m->collision_type * 2;
    if (m->forcing) {
        code += 1;
        for (int d = 0; d < D; d++) nd.F[d] = m->force[d][n];
    }
    nd.B = m->geometry[n];
    if (nd.B >= 1) {
        code = 8;
        micro = 0;
        macro = 0;
    }

    long sl[ORC_MAXQ];
    for (int i = 0; i < Q; i++) {
        sl[i] = slot_of(m, nd.coord, t, i);
        nd.f[i] = m->f[i][sl[i]];
        nd.rho += nd.f[i];
        for (int d = 0; d < D; d++) nd.u[d] += lt->e[d][i] * nd.f[i];
    }
    for (int d = 0; d < D; d++) nd.u[d] = nd.u[d] / nd.rho;

    if (macro > 0) {
        double pu[3] = {0, 0, 0};
        for (int d = 0; d < D; d++) pu[d] = m->u[d][n];
        apply_macro(lt, macro, &nd.rho, nd.u, m->rho[n], pu);
    }
    if (micro > 0) {
        double nb[ORC_MAXQ];
        int face = 0;
        if (micro_kind(lt->lattice, micro, &face) == 2) {
            int axis = face / 2, side = (face % 2 == 0) ? 1 : -1;
            long nn = neighbour(m, nd.coord, axis, side);
            for (int i = 0; i < Q; i++) nb[i] = m->pre[(size_t)nn * Q + i];
        }
        apply_micro(m, micro, &nd, nb);
    }

    collide(m, code, &nd, &tau);

    for (int i = 0; i < Q; i++) m->f[i][sl[i]] = nd.f[i];

    if (store) {
        for (int d = 0; d < D; d++) m->u[d][n] = nd.u[d];
        m->rho[n] = nd.rho;
    }
}

void orc_step(orc_model *m, long long t, int store_macros)
{
    const long N = m->n;
    if (m->has_sf) {
        /* the post-stream populations of every node, before any update */
#pragma omp parallel for schedule(static)
        for (long n = 0; n < N; n++) {
            int c[3];
            coords_of(m, n, c);
            for (int i = 0; i < m->lat.q; i++)
                m->pre[(size_t)n * m->lat.q + i] = m->f[i][slot_of(m, c, t, i)];
        }
    }
#pragma omp parallel for schedule(static)
    for (long n = 0; n < N; n++) update_node(m, n, t, store_macros);
}

void orc_run(orc_model *m, long long t0, int nsteps, int store_every)
{
    for (int k = 0; k < nsteps; k++) {
        long long t = t0 + k;
        int store = store_every > 0 && ((t + 1) % store_every) == 0;
        orc_step(m, t, store);
    }
}

void orc_get_f(const orc_model *m, double *const *f)
{
    for (int i = 0; i < m->lat.q; i++) memcpy(f[i], m->f[i], sizeof(double) * (size_t)m->n);
}

void orc_set_f(orc_model *m, double *const *f)
{
    for (int i = 0; i < m->lat.q; i++) memcpy(m->f[i], f[i], sizeof(double) * (size_t)m->n);
}

void orc_get_macros(const orc_model *m, double *rho, double *const *u)
{
    if (rho) memcpy(rho, m->rho, sizeof(double) * (size_t)m->n);
    for (int d = 0; d < m->lat.dim; d++)
        if (u && u[d]) memcpy(u[d], m->u[d], sizeof(double) * (size_t)m->n);
}

/* Node classes.  A class descriptor is a 16-bit word:
 *   bit 0 solid (porosity >= 1: bounceback, BCs off; solver.cu:67-72)
 *   bit 1 porous (NTPOR/MRTPOR collision and porosity != 0)
 *   bits 4-7 macro code, bits 8-11 micro code.
 * Class 0 is plain fluid (descriptor 0); the other classes are the distinct
 * nonzero descriptors present, in increasing order. */
static uint16_t node_descriptor(const orc_model *m, long n)
{
    double B = m->geometry[n];
    if (B >= 1) return 1;
    uint16_t d = 0;
    if ((m->collision_type == 1 || m->collision_type == 3) && B != 0) d |= 2;
    d |= (uint16_t)((m->macro_bc[n] & 15) << 4);
    d |= (uint16_t)((m->micro_bc[n] & 15) << 8);
    return d;
}

int orc_node_types(const orc_model *m, uint8_t *mask, uint16_t *table, int *nclasses)
{
    static unsigned char present[4096];
    memset(present, 0, sizeof(present));
    for (long n = 0; n < m->n; n++) present[node_descriptor(m, n)] = 1;
    int cls_of[4096];
    int k = 1;
    if (table) table[0] = 0;
    for (int d = 1; d < 4096; d++) {
        if (!present[d]) continue;
        if (k >= 256) return 1;
        cls_of[d] = k;
        if (table) table[k] = (uint16_t)d;
        k++;
    }
    cls_of[0] = 0;
    if (nclasses) *nclasses = k;
    if (mask)
        for (long n = 0; n < m->n; n++) mask[n] = (uint8_t)cls_of[node_descriptor(m, n)];
    return 0;
}

/* links[n] bit i: node n is not solid and its neighbour n + e_i (periodic)
 * is solid. */
void orc_links(const orc_model *m, uint16_t *links)
{
    for (long n = 0; n < m->n; n++) {
        uint16_t b = 0;
        if (!(m->geometry[n] >= 1)) {
            int c[3];
            coords_of(m, n, c);
            for (int i = 1; i < m->lat.q; i++) {
                int cc[3];
                for (int d = 0; d < 3; d++) {
                    int e = d < m->lat.dim ? m->lat.e[d][i] : 0;
                    cc[d] = (c[d] + e + m->L[d]) % m->L[d];
                }
                long nn = cc[0] + cc[1] * (long)m->L[0] + cc[2] * (long)m->L[0] * m->L[1];
                if (m->geometry[nn] >= 1) b |= (uint16_t)(1u << i);
            }
        }
        links[n] = b;
    }
}

/* energy functor + thrust::reduce, index.cu:315-370 (sequential sum here) */
double orc_kinetic_energy(const orc_model *m)
{
    double s = 0;
    for (long n = 0; n < m->n; n++) {
        double q;
        if (m->lat.dim > 2)
            q = 0.5 * m->rho[n] *
                ((m->u[0][n] * m->u[0][n]) + (m->u[1][n] * m->u[1][n]) + (m->u[2][n] * m->u[2][n]));
        else
            q = 0.5 * m->rho[n] * ((m->u[0][n] * m->u[0][n]) + (m->u[1][n] * m->u[1][n]));
        s += q;
    }
    return s;
}

int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* single-node entry points for tests                                  */
/* ------------------------------------------------------------------ */

void orc_node_collide(const orc_model *m, int code, double *f, double *rho, double *u,
                      const double *F, double B, double *tau)
{
    onode nd;
    memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < m->lat.q; i++) nd.f[i] = f[i];
    nd.rho = *rho;
    for (int d = 0; d < m->lat.dim; d++) {
        nd.u[d] = u[d];
        nd.F[d] = F ? F[d] : 0.0;
    }
    nd.B = B;
    nd.c_smag = m->c_smag;
    collide(m, code, &nd, tau);
    for (int i = 0; i < m->lat.q; i++) f[i] = nd.f[i];
    *rho = nd.rho;
    for (int d = 0; d < m->lat.dim; d++) u[d] = nd.u[d];
}

void orc_node_micro(const orc_model *m, int code, double *f, double *rho, double *u,
                    const double *nb)
{
    onode nd;
    memset(&nd, 0, sizeof(nd));
    for (int i = 0; i < m->lat.q; i++) nd.f[i] = f[i];
    nd.rho = *rho;
    for (int d = 0; d < m->lat.dim; d++) nd.u[d] = u[d];
    apply_micro(m, code, &nd, nb);
    for (int i = 0; i < m->lat.q; i++) f[i] = nd.f[i];
    *rho = nd.rho;
    for (int d = 0; d < m->lat.dim; d++) u[d] = nd.u[d];
}

void orc_node_macro(const orc_model *m, int code, double *rho, double *u, double prho,
                    const double *pu)
{
    if (code < 0 || code > 15) return;
    apply_macro(&m->lat, code, rho, u, prho, pu);
}

void orc_equilibrium(const orc_model *m, double rho, const double *u, double *feq)
{
    onode nd;
    memset(&nd, 0, sizeof(nd));
    nd.rho = rho;
    for (int d = 0; d < m->lat.dim; d++) nd.u[d] = u[d];
    equilibrium(&m->lat, &nd, feq);
}
