// Per-node field files.
//
// LBM-C reads its per-node inputs from zone 1 / solution 1 of a CGNS file by
// field name ("Porosity", "ForceX", "MicroBC", ...,
// LBM-C/cgns/cgns_input_handler.cu:57-103) and appends FlowSolutions to a
// CGNS output (LBM-C/cgns/cgns_output_handler.cu:264-293).  File names
// ending in ".cgns" are read and written as CGNS/HDF5 through libcgnsio.so
// (host/cgns_io.c, loaded from the executable's directory on first use);
// every other name uses this self-describing container with the same field
// names:
//
//   "LBMCFLD1"                         8-byte magic
//   uint32 nrecords
//   nrecords x { char name[32] (NUL-padded), uint32 type (1 = float64,
//                2 = int32), uint64 count, count values }
//
// little-endian, x-fastest node order (as the reference's arrays).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lbmc_host {

struct FieldRecord {
    std::string name;
    uint32_t type = 1;  // 1 float64, 2 int32
    std::vector<double> f64;
    std::vector<int32_t> i32;
    size_t count() const { return type == 1 ? f64.size() : i32.size(); }
};

class FieldFile {
public:
    bool read(const std::string &fname, std::string *err);
    bool write(const std::string &fname, std::string *err) const;
    // append records to an existing file (creating it if absent)
    static bool append(const std::string &fname, const std::vector<FieldRecord> &recs, std::string *err);

    const FieldRecord *find(const std::string &name) const;
    std::vector<FieldRecord> records;
};

// true for names ending in ".cgns"
bool is_cgns(const std::string &fname);

// Named per-node inputs from an LBMCFLD1 file or a CGNS file (zone 1,
// solution 1).  get() returns nullptr when the field is absent.
class FieldSource {
public:
    bool open(const std::string &fname, std::string *err);
    const FieldRecord *get(const std::string &label, long long n, uint32_t type, std::string *err);
    // the fields read so far (all records for an LBMCFLD1 file)
    FieldFile file;

private:
    std::string fname_;
    bool cgns_ = false;
};

// Time-series solution output: LBMCFLD1 records "<field>@<iter>" after a
// "Dims" record, or CGNS FlowSolutions "Solution @ Time = <iter>".
class SolutionSink {
public:
    bool create(const std::string &fname, int Lx, int Ly, int Lz, std::string *err);
    bool append(long long iter, const std::vector<FieldRecord> &fields, std::string *err);

private:
    std::string fname_;
    bool cgns_ = false;
};

}  // namespace lbmc_host
