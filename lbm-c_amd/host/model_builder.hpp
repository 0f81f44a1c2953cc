// Model set-up on top of the C ABI: the role of LBM-C's ModelBuilder
// (LBM-C/model_builder.cu:16-281).  Reads the .lbmc constants, creates the
// device lattice, loads the per-node fields by the reference's names and sets
// the initial condition.  Errors print and exit(-1), as the reference's
// cudasafe (LBM-C/cuda_util.cu:14-17) and CGNS input handler
// (LBM-C/cgns/cgns_input_handler.cu:96-98) do.
#pragma once

#include <iosfwd>
#include <string>

#include "../../include/lbmc.h"
#include "field_file.hpp"
#include "infile_reader.hpp"

namespace lbmc_host {

// print-and-exit wrapper for C-ABI calls, as cudasafe
void lbmc_safe(lbmc_ctx *ctx, lbmc_status st, const char *message);

class ModelBuilder {
public:
    ModelBuilder(const std::string &input_filename, std::ostream &log);
    ~ModelBuilder();
    ModelBuilder(const ModelBuilder &) = delete;
    ModelBuilder &operator=(const ModelBuilder &) = delete;

    lbmc_ctx *ctx = nullptr;
    DomainConstants dc;
    Timing times;
    OutputController oc;
    ProjectStrings project;
    long long domain_size = 0;
    long long t_start = 0;  // first time index (a checkpoint restart sets it)
    FieldFile inputs;       // the per-node input fields that were loaded
    int dim() const { return dc.lattice == 0 ? 2 : 3; }
    int q() const { return dc.lattice == 0 ? 9 : 15; }

    // write the raw populations and the next time index (Init 2 reads it back)
    void write_checkpoint(const std::string &fname, long long next_t) const;

private:
    void constant_loader();
    void variable_loader();
    void load_static_IC();

    std::string fname_config_;
    std::ostream &log_;
};

}  // namespace lbmc_host
