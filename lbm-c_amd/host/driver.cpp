// lbmc_run: the host time loop of LBM-C (main in LBM-C/index.cu:61-153) on
// the C ABI.
//
//   lbmc_run <case.lbmc> [--checkpoint <file>] [--reference-store-schedule]
//
// Per step i = 1..MaxT it calls the step with time index i-1
// (index.cu:108), then, as the reference:
//   * every FileOut steps appends VelocityX/Y/Z and Density (as selected by
//     OutputVars) to OutputFile (output_macros, index.cu:196-281): a CGNS
//     FlowSolution when the name ends in ".cgns", field-file records
//     otherwise (field_file.hpp);
//   * every ScreenMes steps prints the ScreenNode's rho, u and residual
//     (screen_mess, index.cu:412-439);
//   * every SteadyCheck steps computes the kinetic-energy residual
//     |E - E_prev| / SteadyCheck / E into a 3-slot ring (index.cu:315-410),
//     writes output and exits 1 if it is NaN, and writes output and stops
//     when all three residuals are below Tolerance (index.cu:122-141).
// The macros are stored on the steps that output, print or check them.  The
// reference's own condition (index.cu:104-106) stores on nearly every step
// because it tests `i % SteadyCheck` instead of `== 0`;
// --reference-store-schedule reproduces that schedule (same results, more
// traffic).  Steps between events run through lbmc_run, which replays a
// captured graph.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "field_file.hpp"
#include "model_builder.hpp"

using namespace lbmc_host;

namespace {

struct Driver {
    ModelBuilder &mb;
    double prev_energy = 0;
    SolutionSink sink;

    // output_macros, index.cu:196-281
    void output_macros(long long time)
    {
        const long long n = mb.domain_size;
        std::vector<double> rho(n);
        std::vector<std::vector<double>> u(3, std::vector<double>(n));
        double *up[3] = {u[0].data(), u[1].data(), u[2].data()};
        lbmc_safe(mb.ctx, lbmc_get_macros(mb.ctx, rho.data(), up), "Output: Copy from device memory failed!");
        std::vector<FieldRecord> recs;
        const char *labels[3] = {"VelocityX", "VelocityY", "VelocityZ"};
        for (int d = 0; d < mb.dim(); d++) {
            if (!mb.oc.u[d]) continue;
            FieldRecord r;
            r.name = labels[d];
            r.f64 = u[d];
            recs.push_back(std::move(r));
        }
        if (mb.oc.rho) {
            FieldRecord r;
            r.name = "Density";
            r.f64 = rho;
            recs.push_back(std::move(r));
        }
        std::string err;
        if (!mb.project.output_fname.empty() && !sink.append(time, recs, &err)) {
            fprintf(stderr, "ERROR: Output Handler: %s\n", err.c_str());
            exit(-1);
        }
        std::cout << "Output Handler: Solution @ Time = " << time << " written to "
                  << mb.project.output_fname << "\n";
    }

    // error_RMS / compute_residual, index.cu:377-410
    void compute_residual(long long time)
    {
        double e = 0;
        lbmc_safe(mb.ctx, lbmc_kinetic_energy(mb.ctx, &e), "Steady State Calculation Kernel Execution Failed!");
        const double r = (std::fabs(e - prev_energy) / mb.times.steady_check) / e;
        prev_energy = e;
        mb.dc.residual[time % 3] = r;
    }

    // screen_mess, index.cu:412-439
    void screen_mess(long long iter)
    {
        const int *c = mb.oc.screen_node;
        long long idx = c[0] + (long long)c[1] * mb.dc.length[0];
        if (mb.dim() > 2) idx += (long long)c[2] * mb.dc.length[0] * mb.dc.length[1];
        double rho = 0, u[3] = {0, 0, 0};
        lbmc_safe(mb.ctx, lbmc_get_node(mb.ctx, idx, &rho, u), "Screen: Copy from device memory failed!");
        std::cout << "time = " << iter << "; rho = " << rho << "; uX = " << u[0] << "; uY = " << u[1] << "; ";
        if (mb.dim() > 2) std::cout << "uZ = " << u[2] << "; ";
        std::cout << "resid = " << mb.dc.residual[iter % 3] << std::endl;
    }
};

bool due(int every, long long i) { return every > 0 && i % every == 0; }

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <case.lbmc> [--checkpoint <file>] [--reference-store-schedule]\n", argv[0]);
        return 2;
    }
    std::string checkpoint;
    bool ref_schedule = false;
    for (int k = 2; k < argc; k++) {
        if (!strcmp(argv[k], "--checkpoint") && k + 1 < argc) checkpoint = argv[++k];
        else if (!strcmp(argv[k], "--reference-store-schedule")) ref_schedule = true;
    }

    ModelBuilder mb(argv[1], std::cout);
    Driver drv{mb, 0.0, SolutionSink()};
    {
        // the output starts with the grid (the reference writes base, zone
        // and coordinates when it creates the CGNS file, index.cu:184)
        std::string err;
        if (!mb.project.output_fname.empty() &&
            !drv.sink.create(mb.project.output_fname, mb.dc.length[0], mb.dc.length[1],
                             mb.dim() > 2 ? mb.dc.length[2] : 1, &err)) {
            fprintf(stderr, "ERROR: Output Handler: %s\n", err.c_str());
            return -1;
        }
    }
    long long nodes = 0;
    int q = 0, sb = 0;
    lbmc_info(mb.ctx, &nodes, &q, &sb);
    std::cout << "Population storage: " << (double)nodes * q * sb / 1048576.0 << " Mb\n";
    printf("X-Length:\t\t%d\n", mb.dc.length[0]);
    printf("Y-Length:\t\t%d\n", mb.dc.length[1]);
    if (mb.dim() > 2) printf("Z-Length:\t\t%d\n", mb.dc.length[2]);
    printf("Relaxation Time (Tau):\t%f\n", mb.dc.tau);
    if (mb.oc.interactive) {
        printf("\nPress return to continue...");
        getchar();
    }

    const auto t1 = std::chrono::steady_clock::now();
    const Timing &T = mb.times;
    int stop = 0;
    long long i = 1;
    const long long t0 = mb.t_start;
    while (i <= T.max) {
        const bool event = due(T.plot, i) || due(T.screen, i) || due(T.steady_check, i);
        if (!event && !ref_schedule) {
            // run up to the step before the next event in one call
            long long next = T.max + 1;
            for (int every : {T.plot, T.screen, T.steady_check})
                if (every > 0) next = std::min(next, (i / every + 1) * (long long)every);
            const long long k = std::min(next, (long long)T.max + 1) - i;
            lbmc_safe(mb.ctx, lbmc_run(mb.ctx, t0 + i - 1, (int)k, 0), "Kernel \"iterate\" Execution Failed!");
            i += k;
            continue;
        }
        bool store = event;
        if (ref_schedule)
            store = due(T.plot, i) || (T.steady_check > 0 && i % T.steady_check) || (T.screen > 0 && i % T.screen);
        lbmc_safe(mb.ctx, lbmc_step(mb.ctx, t0 + i - 1, store ? 1 : 0), "Kernel \"iterate\" Execution Failed!");
        if (due(T.plot, i)) drv.output_macros(i);
        if (due(T.screen, i)) drv.screen_mess(i);
        if (due(T.steady_check, i)) {
            drv.compute_residual(i);
            for (int r = 0; r < 3; r++)
                if (mb.dc.residual[r] < mb.dc.tolerance) stop += 1;
            if (std::isnan(mb.dc.residual[i % 3])) {
                drv.output_macros(i);
                std::cout << "Residual is NaN at time " << i << ": stopping\n";
                return 1;
            } else if (stop == 3) {
                drv.output_macros(i);
                std::cout << "Steady state reached at time " << i << "\n";
                i++;
                break;
            }
            stop = 0;
        }
        i++;
    }
    lbmc_safe(mb.ctx, lbmc_sync(mb.ctx), "Kernel \"iterate\" Execution Failed!");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    printf("\n\nTotal Run Time: %fs\n", secs);
    if (!checkpoint.empty()) {
        // restartable with Init 2: the input fields plus f0..f<Q-1> and TimeStep
        mb.write_checkpoint(checkpoint + ".tmp", t0 + i - 1);
        FieldFile ck;
        const FieldFile &in = mb.inputs;
        std::string err;
        if (!ck.read(checkpoint + ".tmp", &err)) {
            fprintf(stderr, "ERROR: Checkpoint: %s\n", err.c_str());
            return -1;
        }
        FieldFile out;
        for (const FieldRecord &r : in.records)
            if (!ck.find(r.name)) out.records.push_back(r);
        for (const FieldRecord &r : ck.records) out.records.push_back(r);
        if (!out.write(checkpoint, &err)) {
            fprintf(stderr, "ERROR: Checkpoint: %s\n", err.c_str());
            return -1;
        }
        std::remove((checkpoint + ".tmp").c_str());
        std::cout << "Checkpoint written to " << checkpoint << " (next time index " << t0 + i - 1 << ")\n";
    }
    if (mb.oc.interactive) {
        printf("\nPress return to finish");
        getchar();
    }
    return 0;
}
