// .lbmc reader, following LBM-C/infile_reader.cu:33-248 (see the header).
#include "infile_reader.hpp"

#include <fstream>
#include <iostream>
#include <sstream>

namespace lbmc_host {

namespace {

bool has(const std::string &line, const char *key) { return line.find(key) != std::string::npos; }

}  // namespace

void InfileReader::process_line(const std::string &line_tot, ProjectStrings *project, DomainConstants *dc,
                                Timing *timer, OutputController *oc, std::ostream &log)
{
    std::string line;
    {
        std::stringstream ss(line_tot);
        std::getline(ss, line, '#');
    }
    std::string key;
    std::stringstream in(line);
    const int dim = dc->lattice == 0 ? 2 : 3;
    const int q = dc->lattice == 0 ? 9 : 15;
    if (has(line, "ProjName")) {
        in >> key >> project->name;
        log << key << " = " << project->name << "\n";
    } else if (has(line, "DomainFile")) {
        in >> key >> project->domain_fname;
        log << key << " = " << project->domain_fname << "\n";
    } else if (has(line, "OutputFile")) {
        in >> key >> project->output_fname;
        log << key << " = " << project->output_fname << "\n";
    } else if (has(line, "TauMRT")) {
        in >> key;
        for (int i = 0; i < q; i++) in >> dc->tau_mrt[i];
        for (int i = 0; i < q; i++) log << "S_" << i << " = " << dc->tau_mrt[i] << "\n";
    } else if (has(line, "Tau")) {
        in >> key >> dc->tau;
        log << key << " = " << dc->tau << "\n";
    } else if (has(line, "Lx")) {
        in >> key >> dc->length[0];
        log << key << " = " << dc->length[0] << "\n";
    } else if (has(line, "Ly")) {
        in >> key >> dc->length[1];
        log << key << " = " << dc->length[1] << "\n";
    } else if (has(line, "Lz")) {
        int tmp = 0;
        in >> key >> tmp;
        if (dim > 2) {
            dc->length[2] = tmp;
            log << key << " = " << tmp << "\n";
        }
    } else if (has(line, "DeltaX")) {
        in >> key >> dc->h;
        log << key << " = " << dc->h << "\n";
    } else if (has(line, "DeltaT")) {
        in >> key >> dc->dt;
        log << key << " = " << dc->dt << "\n";
    } else if (has(line, "C_smag")) {
        in >> key >> dc->c_smag;
        log << key << " = " << dc->c_smag << "\n";
    } else if (has(line, "ColType")) {
        std::string name;
        if (has(line, "BGK")) { name = "BGK"; dc->collision_type = 0; }
        if (has(line, "NTPOR")) { name = "NTPOR"; dc->collision_type = 1; }
        if (has(line, "MRT")) { name = "MRT"; dc->collision_type = 2; }
        if (has(line, "MRTPOR")) { name = "MRTPOR"; dc->collision_type = 3; }
        in >> key;
        log << key << " = " << name << "\n";
    } else if (has(line, "Force")) {
        in >> key >> dc->forcing;
        log << key << " = " << dc->forcing << "\n";
    } else if (has(line, "MicroBC")) {
        in >> key >> dc->micro_bc;
        log << key << " = " << dc->micro_bc << "\n";
    } else if (has(line, "MacroBC")) {
        in >> key >> dc->macro_bc;
        log << key << " = " << dc->macro_bc << "\n";
    } else if (has(line, "Tolerance")) {
        in >> key >> dc->tolerance;
        log << key << " = " << dc->tolerance << "\n";
    } else if (has(line, "Init")) {
        in >> key >> dc->init_type;
        log << key << " = " << dc->init_type << "\n";
    } else if (has(line, "MaxT")) {
        in >> key >> timer->max;
        log << key << " = " << timer->max << "\n";
    } else if (has(line, "FileOut")) {
        in >> key >> timer->plot;
        log << key << " = " << timer->plot << "\n";
    } else if (has(line, "ScreenMes")) {
        in >> key >> timer->screen;
        log << key << " = " << timer->screen << "\n";
    } else if (has(line, "SteadyCheck")) {
        in >> key >> timer->steady_check;
        log << key << " = " << timer->steady_check << "\n";
    } else if (has(line, "OutputVars")) {
        bool tmp = false;
        in >> key >> oc->u[0] >> oc->u[1] >> tmp >> oc->rho >> oc->pressure;
        if (dim > 2 && dc->length[2] > 0) oc->u[2] = tmp;
        if (oc->u[0]) log << "Output VelocityX = " << oc->u[0] << "\n";
        if (oc->u[1]) log << "Output VelocityY = " << oc->u[1] << "\n";
        if (dim > 2 && oc->u[2]) log << "Output VelocityZ = " << oc->u[2] << "\n";
        if (oc->rho) log << "Output Density = " << oc->rho << "\n";
        if (oc->pressure) log << "Output Pressure = " << oc->pressure << "\n";
    } else if (has(line, "ScreenNode")) {
        int tmp = 0;
        in >> key >> oc->screen_node[0] >> oc->screen_node[1] >> tmp;
        if (dim > 2 && dc->length[2] > 0) oc->screen_node[2] = tmp;
        log << "Screen Node X = " << oc->screen_node[0] << "\n";
        log << "Screen Node Y = " << oc->screen_node[1] << "\n";
        if (dim > 2) log << "Screen Node Z = " << oc->screen_node[2] << "\n";
    } else if (has(line, "Interactive")) {
        in >> key >> oc->interactive;
        log << key << " = " << oc->interactive << "\n";
    } else if (has(line, "Lattice")) {
        // addition: the reference selects the lattice at compile time
        std::string v;
        in >> key >> v;
        dc->lattice = (v == "D2Q9") ? 0 : 1;
        log << key << " = " << (dc->lattice == 0 ? "D2Q9" : "D3Q15") << "\n";
    } else if (has(line, "Storage")) {
        // addition: population storage precision
        std::string v;
        in >> key >> v;
        dc->storage = v == "FP32" ? 1 : v == "FP16S" ? 2 : v == "BF16S" ? 3 : 0;
        log << key << " = " << v << "\n";
    } else {
        log << line_tot << "\n";
    }
}

InfileReader::InfileReader(const std::string &fname, ProjectStrings *project, DomainConstants *dc,
                           Timing *timer, OutputController *oc, std::ostream &log)
{
    dc->c_smag = 0;  // initialise(), infile_reader.cu:20-23
    std::ifstream f(fname);
    log << "\nReading configuration data: \n\n";
    if (!f) {
        log << "Could not open " << fname << "\n";
        return;
    }
    // the lattice key may appear anywhere; it decides how many TauMRT values
    // and length components are read, so look for it first
    std::string line;
    while (std::getline(f, line)) {
        std::string body;
        std::stringstream ss(line);
        std::getline(ss, body, '#');
        if (body.find("Lattice") != std::string::npos) {
            std::stringstream in(body);
            std::string key, v;
            in >> key >> v;
            dc->lattice = (v == "D2Q9") ? 0 : 1;
        }
    }
    f.clear();
    f.seekg(0);
    while (std::getline(f, line)) process_line(line, project, dc, timer, oc, log);
    log << "\nFinished reading configuration data.\n";
    ok_ = true;
}

}  // namespace lbmc_host
