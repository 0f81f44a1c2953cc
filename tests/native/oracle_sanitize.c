/*
 * Drives the FP64 oracle (oracle/lbmc_oracle.c) through every code path a
 * lattice can take, for a build with AddressSanitizer and
 * UndefinedBehaviorSanitizer: both lattices on odd-sized boxes, every
 * collision type with and without Guo forcing and Smagorinsky, solid and
 * porous nodes, every macro code, every valid micro code (Zou-He and sf on
 * arbitrary nodes, neighbours wrapping), the quirk fixes, invalid codes,
 * time indices past 2^31, store schedules, read-back, node classes, links,
 * kinetic energy and the single-node operators.  Exit status 0 when every
 * call returned what it should; the sanitizers abort on any memory or
 * undefined-behaviour error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/lbmc_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double urand(void)
{
    rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int failures = 0;
#define CHECK(c, ...)                                                                                        \
    do {                                                                                                     \
        if (!(c)) {                                                                                          \
            printf("FAIL: " __VA_ARGS__);                                                                    \
            printf("\n");                                                                                    \
            failures++;                                                                                      \
        }                                                                                                    \
    } while (0)

static void run_case(int lattice, int coll, int forcing, double c_smag, unsigned quirks)
{
    const int D = lattice == 0 ? 2 : 3, Q = lattice == 0 ? 9 : 15;
    const int len[3] = {lattice == 0 ? 23 : 11, lattice == 0 ? 13 : 7, lattice == 0 ? 1 : 5};
    const long long N = (long long)len[0] * len[1] * len[2];
    double tau_mrt[15];
    for (int i = 0; i < 15; i++) tau_mrt[i] = 1.0 / (0.6 + 0.1 * (i % 4));
    orc_model *m = orc_create(lattice, len, coll, forcing, 1, 1, 0.8, tau_mrt, c_smag, quirks);
    CHECK(m != NULL, "create lattice %d coll %d", lattice, coll);
    if (!m) return;

    double *geo = malloc(sizeof(double) * N), *rho = malloc(sizeof(double) * N);
    int *mac = malloc(sizeof(int) * N), *mic = malloc(sizeof(int) * N);
    double *force[3], *u[3], *f[15];
    for (int d = 0; d < 3; d++) {
        force[d] = malloc(sizeof(double) * N);
        u[d] = malloc(sizeof(double) * N);
    }
    for (int i = 0; i < 15; i++) f[i] = malloc(sizeof(double) * N);
    const int nmicro = lattice == 0 ? 9 : 13; /* codes 0..8 / 0..12 */
    for (long long n = 0; n < N; n++) {
        const double r = urand();
        geo[n] = r < 0.1 ? 1.0 : (r < 0.3 ? 0.5 * urand() : 0.0);
        mac[n] = (int)(urand() * 16.0);
        int mc = (int)(urand() * nmicro);
        if (lattice == 0 && (mc == 3 || mc == 4)) mc = 0; /* NULL slots of the D2Q9 table */
        mic[n] = urand() < 0.3 ? mc : 0;
        rho[n] = 1.0 + 0.01 * (urand() - 0.5);
        for (int d = 0; d < 3; d++) {
            force[d][n] = 1e-5 * (urand() - 0.5);
            u[d][n] = 0.02 * (urand() - 0.5);
        }
    }
    CHECK(orc_set_fields(m, geo, mac, mic, force, u, rho) == 0, "set_fields");

    /* invalid codes are refused, not dereferenced */
    const int good = mic[0];
    mic[0] = lattice == 0 ? 3 : 13;
    CHECK(orc_set_fields(m, geo, mac, mic, force, u, rho) != 0, "invalid micro code accepted");
    mic[0] = -1;
    CHECK(orc_set_fields(m, geo, mac, mic, force, u, rho) != 0, "negative micro code accepted");
    mic[0] = good;
    mac[0] = 16;
    CHECK(orc_set_fields(m, geo, mac, mic, force, u, rho) != 0, "macro code 16 accepted");
    mac[0] = 0;
    CHECK(orc_set_fields(m, geo, mac, mic, force, u, rho) == 0, "set_fields again");

    orc_init_static(m);
    orc_run(m, 0, 4, 2);
    orc_init_equilibrium(m, rho, u);
    orc_run(m, 0, 3, 1);
    orc_run(m, (1LL << 31) + 5, 3, 0);  /* time index past 2^31 */
    orc_step(m, (1LL << 33) - 1, 1);
    orc_get_f(m, f);
    orc_set_f(m, f);
    orc_get_macros(m, rho, u);
    double e = orc_kinetic_energy(m);
    CHECK(e == e, "energy is a number");

    uint8_t *mask = malloc((size_t)N);
    uint16_t table[256], *links = malloc(sizeof(uint16_t) * N);
    int ncls = 0;
    CHECK(orc_node_types(m, mask, table, &ncls) == 0 && ncls >= 1 && ncls <= 256, "node types");
    orc_links(m, links);

    /* single-node operators: every collision code, micro and macro code */
    double fn[15], nb[15], un[3] = {0.01, -0.02, 0.005}, F[3] = {1e-5, 0, -1e-5}, pu[3] = {0.1, 0.2, 0.3};
    for (int code = 0; code <= 8; code++) {
        for (int i = 0; i < Q; i++) fn[i] = 1.0 / Q + 0.001 * (urand() - 0.5);
        double r = 1.0, tau = 0.8, uu[3] = {un[0], un[1], un[2]};
        orc_node_collide(m, code, fn, &r, uu, F, 0.3, &tau);
    }
    for (int code = 0; code < nmicro; code++) {
        if (lattice == 0 && (code == 3 || code == 4)) continue;
        for (int i = 0; i < Q; i++) {
            fn[i] = 1.0 / Q + 0.001 * (urand() - 0.5);
            nb[i] = urand();
        }
        double r = 1.001, uu[3] = {0, 0, 0};
        orc_node_micro(m, code, fn, &r, uu, nb);
    }
    for (int code = 0; code <= 15; code++) {
        double r = 1.0, uu[3] = {un[0], un[1], un[2]};
        orc_node_macro(m, code, &r, uu, 1.01, pu);
    }
    double feq[15];
    orc_equilibrium(m, 1.0, un, feq);
    (void)D;

    orc_destroy(m);
    free(geo);
    free(rho);
    free(mac);
    free(mic);
    free(mask);
    free(links);
    for (int d = 0; d < 3; d++) {
        free(force[d]);
        free(u[d]);
    }
    for (int i = 0; i < 15; i++) free(f[i]);
}

int main(void)
{
    orc_set_threads(2);
    int e[45], opp[15];
    double w[15], M[225], Minv[225];
    CHECK(orc_lattice_tables(0, 0, e, w, opp, M, Minv) == 9, "D2Q9 tables");
    CHECK(orc_lattice_tables(1, 7, e, w, opp, M, Minv) == 15, "D3Q15 tables");
    CHECK(orc_lattice_tables(2, 0, e, w, opp, M, Minv) == -1, "unknown lattice");
    for (int lattice = 0; lattice < 2; lattice++)
        for (int coll = 0; coll < 4; coll++)
            for (int forcing = 0; forcing < 2; forcing++) {
                run_case(lattice, coll, forcing, coll == 0 ? 0.17 : 0.0, 0);
                if (lattice == 1) run_case(lattice, coll, forcing, 0.0, 7u);
            }
    printf("%s\n", failures ? "oracle sanitize: FAILED" : "oracle sanitize: ok");
    return failures ? 1 : 0;
}
