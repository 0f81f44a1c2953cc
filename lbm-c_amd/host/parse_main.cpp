// lbmc_parse: parses a .lbmc file with InfileReader and prints the parsed
// constants as JSON (used by the CPU tests of the reader).
#include <iostream>
#include <sstream>

#include "infile_reader.hpp"

using namespace lbmc_host;

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cerr << "usage: lbmc_parse <case.lbmc>\n";
        return 2;
    }
    ProjectStrings p;
    DomainConstants dc;
    Timing t;
    OutputController oc;
    std::ostringstream echo;
    InfileReader r(argv[1], &p, &dc, &t, &oc, echo);
    if (!r.ok()) return 1;
    std::cout.precision(17);
    std::cout << "{\"ProjName\": \"" << p.name << "\", \"DomainFile\": \"" << p.domain_fname
              << "\", \"OutputFile\": \"" << p.output_fname << "\", \"lattice\": " << dc.lattice
              << ", \"storage\": " << dc.storage << ", \"tau\": " << dc.tau << ", \"tau_mrt\": [";
    for (int i = 0; i < 15; i++) std::cout << dc.tau_mrt[i] << (i < 14 ? ", " : "");
    std::cout << "], \"length\": [" << dc.length[0] << ", " << dc.length[1] << ", " << dc.length[2]
              << "], \"h\": " << dc.h << ", \"dt\": " << dc.dt << ", \"c_smag\": " << dc.c_smag
              << ", \"collision_type\": " << dc.collision_type << ", \"forcing\": " << dc.forcing
              << ", \"micro_bc\": " << dc.micro_bc << ", \"macro_bc\": " << dc.macro_bc
              << ", \"tolerance\": " << dc.tolerance << ", \"init_type\": " << dc.init_type
              << ", \"max\": " << t.max << ", \"plot\": " << t.plot << ", \"screen\": " << t.screen
              << ", \"steady_check\": " << t.steady_check << ", \"out_u\": [" << oc.u[0] << ", " << oc.u[1]
              << ", " << oc.u[2] << "], \"out_rho\": " << oc.rho << ", \"out_pressure\": " << oc.pressure
              << ", \"screen_node\": [" << oc.screen_node[0] << ", " << oc.screen_node[1] << ", "
              << oc.screen_node[2] << "], \"interactive\": " << oc.interactive << "}\n";
    std::ostringstream esc;
    std::cerr << echo.str();
    return 0;
}
