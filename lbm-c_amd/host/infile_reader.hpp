// .lbmc configuration reader: the key/value format of LBM-C's InfileReader
// (LBM-C/infile_reader.cu:33-248).
//
// Each line is cut at the first '#'; the remainder is matched against the
// keys by substring, in the reference's order (so "TauMRT" is tested before
// "Tau", "MRTPOR" overrides "MRT"), and the values are read with stream
// extraction after the key word.  Recognised lines are echoed as
// "<key> = <value>", other lines verbatim, like the reference.
//
// Additions (keys the reference fixes at compile time, LBM-C/macros.cu:11-12):
//   Lattice D2Q9 | D3Q15     (default D3Q15, the reference's build)
//   Storage FP64 | FP32 | FP16S | BF16S   (default FP64: the reference's bits)
#pragma once

#include <iosfwd>
#include <string>

namespace lbmc_host {

// DomainConstant (LBM-C/data_types.cuh:23-43) fields read from the file
struct DomainConstants {
    int lattice = 1;  // 0 D2Q9, 1 D3Q15
    int storage = 0;  // LBMC_FP64
    double tau_mrt[15] = {};
    double tau = 0;
    int h = 0, dt = 0;
    int length[3] = {0, 0, 0};
    bool forcing = false, macro_bc = false, micro_bc = false;
    int collision_type = 0;
    int init_type = 0;
    double residual[3] = {1, 1, 1};
    double tolerance = 0;
    double c_smag = 0;
};

// Timing (LBM-C/data_types.cuh:57-63)
struct Timing {
    int max = 0, plot = 0, screen = 0, steady_check = 0;
};

// OutputController (LBM-C/data_types.cuh:65-72)
struct OutputController {
    bool u[3] = {false, false, false};
    bool rho = false, pressure = false, interactive = false;
    int screen_node[3] = {0, 0, 0};
};

// ProjectStrings (LBM-C/data_types.cuh:74-79)
struct ProjectStrings {
    std::string name, domain_fname, output_fname;
};

class InfileReader {
public:
    // Parses `fname`; echo goes to `log`.  Returns false if the file cannot
    // be opened (the reference silently read nothing; the driver reports it).
    InfileReader(const std::string &fname, ProjectStrings *project, DomainConstants *dc, Timing *timer,
                 OutputController *oc, std::ostream &log);
    bool ok() const { return ok_; }

    // one line, as process_data_field (infile_reader.cu:45-226)
    static void process_line(const std::string &line_tot, ProjectStrings *project, DomainConstants *dc,
                             Timing *timer, OutputController *oc, std::ostream &log);

private:
    bool ok_ = false;
};

}  // namespace lbmc_host
