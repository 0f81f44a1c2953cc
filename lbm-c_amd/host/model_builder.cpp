// ModelBuilder on the C ABI; see model_builder.hpp.
#include "model_builder.hpp"

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "field_file.hpp"

namespace lbmc_host {

void lbmc_safe(lbmc_ctx *ctx, lbmc_status st, const char *message)
{
    if (st != LBMC_OK) {
        fprintf(stderr, "ERROR: %s : %s\n", message, lbmc_last_error(ctx));
        exit(-1);
    }
}

ModelBuilder::ModelBuilder(const std::string &input_filename, std::ostream &log)
    : fname_config_(input_filename), log_(log)
{
    constant_loader();
    variable_loader();
    log_ << "variable loader complete\n";
}

ModelBuilder::~ModelBuilder()
{
    if (ctx) lbmc_destroy(ctx);
}

// constant_loader, LBM-C/model_builder.cu:63-83: the .lbmc constants become
// the lattice descriptor (the reference copies them to __constant__ memory)
void ModelBuilder::constant_loader()
{
    InfileReader reader(fname_config_, &project, &dc, &times, &oc, log_);
    if (!reader.ok()) {
        fprintf(stderr, "ERROR: Model Builder: cannot read %s\n", fname_config_.c_str());
        exit(-1);
    }
    lbmc_desc d = {};
    d.lattice = dc.lattice;
    for (int k = 0; k < 3; k++) d.length[k] = k < dim() ? dc.length[k] : 1;
    d.collision_type = dc.collision_type;
    d.forcing = dc.forcing;
    d.macro_bc = dc.macro_bc;
    d.micro_bc = dc.micro_bc;
    d.tau = dc.tau;
    for (int i = 0; i < 15; i++) d.tau_mrt[i] = dc.tau_mrt[i];
    d.c_smag = dc.c_smag;
    d.storage = dc.storage;
    d.quirks = 0;
    d.device = 0;
    lbmc_safe(nullptr, lbmc_create(&d, &ctx), "Model Builder: lattice creation failed!");
    domain_size = 1;
    for (int k = 0; k < dim(); k++) domain_size *= dc.length[k];
}

namespace {

const FieldRecord &need(FieldSource &f, const std::string &label, const std::string &fname,
                        long long n, uint32_t type, std::ostream &log)
{
    std::string err;
    const FieldRecord *r = f.get(label, n, type, &err);
    if (!r && !err.empty()) {
        fprintf(stderr, "ERROR: Input Handler: %s\n", err.c_str());
        exit(-1);
    }
    if (!r) {
        // cgns_input_handler.cu:96-98
        log << "\nInput Handler: " << label << " not found in file \"" << fname << "\"\n";
        log.flush();
        exit(-1);
    }
    if ((long long)r->count() != n || r->type != type) {
        fprintf(stderr, "ERROR: Input Handler: %s has %zu values of type %u, expected %lld of type %u\n",
                label.c_str(), r->count(), r->type, n, type);
        exit(-1);
    }
    log << "\nInput Handler: " << label << " loaded\n";
    return *r;
}

}  // namespace

// variable_loader, LBM-C/model_builder.cu:173-231
void ModelBuilder::variable_loader()
{
    FieldSource in;
    std::string err;
    if (!in.open(project.domain_fname, &err)) {
        fprintf(stderr, "ERROR: Input Handler: %s\n", err.c_str());
        exit(-1);
    }
    const long long n = domain_size;
    const int D = dim();
    const FieldRecord &geo = need(in, "Porosity", project.domain_fname, n, 1, log_);
    const char *force_labels[3] = {"ForceX", "ForceY", "ForceZ"};
    const char *vel_labels[3] = {"VelocityX", "VelocityY", "VelocityZ"};
    const double *force[3] = {nullptr, nullptr, nullptr};
    const double *u[3] = {nullptr, nullptr, nullptr};
    const double *rho = nullptr;
    const int *micro = nullptr, *macro = nullptr;
    if (dc.forcing)
        for (int d = 0; d < D; d++) force[d] = need(in, force_labels[d], project.domain_fname, n, 1, log_).f64.data();
    if (dc.micro_bc) micro = need(in, "MicroBC", project.domain_fname, n, 2, log_).i32.data();
    if (dc.macro_bc) {
        macro = need(in, "MacroBC", project.domain_fname, n, 2, log_).i32.data();
        for (int d = 0; d < D; d++) u[d] = need(in, vel_labels[d], project.domain_fname, n, 1, log_).f64.data();
        rho = need(in, "Rho", project.domain_fname, n, 1, log_).f64.data();
    }
    lbmc_safe(ctx, lbmc_set_fields(ctx, geo.f64.data(), macro, micro, force, u, rho),
              "Model Builder: field upload failed!");

    if (dc.init_type == 0) {
        load_static_IC();
    } else if (dc.init_type == 1) {
        // addition: equilibrium from the Rho / Velocity fields
        const double *r0 = need(in, "Rho", project.domain_fname, n, 1, log_).f64.data();
        const double *u0[3] = {nullptr, nullptr, nullptr};
        for (int d = 0; d < D; d++) u0[d] = need(in, vel_labels[d], project.domain_fname, n, 1, log_).f64.data();
        lbmc_safe(ctx, lbmc_init_equilibrium(ctx, r0, u0), "Model Builder: equilibrium IC failed!");
    } else if (dc.init_type == 2) {
        // addition: restart from raw populations f0..f<Q-1> and TimeStep
        std::vector<const double *> fp(15, nullptr);
        for (int i = 0; i < q(); i++)
            fp[i] = need(in, "f" + std::to_string(i), project.domain_fname, n, 1, log_).f64.data();
        const FieldRecord &ts = need(in, "TimeStep", project.domain_fname, 1, 1, log_);
        t_start = (long long)ts.f64[0];
        lbmc_safe(ctx, lbmc_set_f(ctx, fp.data()), "Model Builder: restart upload failed!");
    }
    // other Init values leave the populations as allocated (zero), where the
    // reference leaves them uninitialised (model_builder.cu:227-230)
    inputs = in.file;
}

// load_static_IC, LBM-C/model_builder.cu:233-248
void ModelBuilder::load_static_IC()
{
    lbmc_safe(ctx, lbmc_init_static(ctx), "Model Builder: Copy to device memory failed!");
}

void ModelBuilder::write_checkpoint(const std::string &fname, long long next_t) const
{
    const long long n = domain_size;
    FieldFile out;
    std::vector<std::vector<double>> f(q(), std::vector<double>(n));
    std::vector<double *> fp(15, nullptr);
    for (int i = 0; i < q(); i++) fp[i] = f[i].data();
    lbmc_safe(ctx, lbmc_get_f(ctx, fp.data()), "Checkpoint: Copy from device memory failed!");
    for (int i = 0; i < q(); i++) {
        FieldRecord r;
        r.name = "f" + std::to_string(i);
        r.f64 = std::move(f[i]);
        out.records.push_back(std::move(r));
    }
    FieldRecord ts;
    ts.name = "TimeStep";
    ts.f64 = {(double)next_t};
    out.records.push_back(ts);
    std::string err;
    if (!out.write(fname, &err)) {
        fprintf(stderr, "ERROR: Checkpoint: %s\n", err.c_str());
        exit(-1);
    }
}

}  // namespace lbmc_host
