// Drives the host code above the C ABI that parses input (lbm-c_amd/host:
// the .lbmc reader, the LBMCFLD1 field container, the CGNS/HDF5 reader and
// writer) for a build with AddressSanitizer and UndefinedBehaviorSanitizer:
// well-formed and malformed .lbmc lines, field files written, read back,
// appended, truncated and corrupted (bad magic, unknown record type, counts
// larger than the file), and CGNS files written, appended and read by name
// through the driver's field source and solution sink (libcgnsio.so is
// loaded from this executable's directory; the test builds a sanitized copy
// there; cgns_sanitize.c covers the CGNS library's own malformed inputs).
// Exit status 0 when every call returned what it should.
#include <unistd.h>

#include <climits>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../lbm-c_amd/host/field_file.hpp"
#include "../../lbm-c_amd/host/infile_reader.hpp"

using namespace lbmc_host;

static int failures = 0;
static void check(bool ok, const std::string &what)
{
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        failures++;
    }
}

static void reader(const std::string &dir)
{
    const char *lines[] = {
        "", "#", "# only a comment", "Tau", "Tau 0.8", "TauMRT 1 2", "TauMRT", "TauMRT 1 2 3 4 5 6 7 8 9 10 11 12 13 14 15 16",
        "Lx -5", "Ly abc", "Lz 99999999999999999999", "ColType", "ColType XYZ", "ColType MRTPOR", "OutputVars 1",
        "OutputVars 1 1 1 1 1 1 1", "ScreenNode 1", "ProjName", "DomainFile a b c", "Lattice D2Q9", "Lattice Q",
        "Storage FP16S", "Storage", "Init 2", "MaxT -1", "C_smag nan", "Tolerance 1e400", "\t\t  Tau\t0.6  # x",
        "\xff\xfe\x01 binary", "Force 1 MicroBC 1 MacroBC 1"};
    ProjectStrings p;
    DomainConstants dc;
    Timing t;
    OutputController oc;
    std::ostringstream log;
    for (const char *l : lines) InfileReader::process_line(l, &p, &dc, &t, &oc, log);
    InfileReader::process_line(std::string(100000, 'x') + " Tau 1", &p, &dc, &t, &oc, log);
    const std::string fn = dir + "/case.lbmc";
    {
        std::ofstream o(fn);
        for (const char *l : lines) o << l << "\n";
    }
    InfileReader r(fn, &p, &dc, &t, &oc, log);
    check(r.ok(), "reader opens a file");
    InfileReader missing(dir + "/no_such.lbmc", &p, &dc, &t, &oc, log);
    check(!missing.ok(), "reader reports a missing file");
}

static std::vector<char> slurp(const std::string &fn)
{
    std::ifstream i(fn, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(i)), std::istreambuf_iterator<char>());
}

static void spit(const std::string &fn, const std::vector<char> &b)
{
    std::ofstream o(fn, std::ios::binary);
    o.write(b.data(), (std::streamsize)b.size());
}

static void fields(const std::string &dir)
{
    const std::string fn = dir + "/f.lbmf";
    FieldFile ff;
    FieldRecord a, b;
    a.name = "Porosity";
    a.type = 1;
    a.f64 = {0, 1, 0.5, 0.25};
    b.name = std::string(40, 'N');  // longer than the 31-character name slot
    b.type = 2;
    b.i32 = {1, 2, 3};
    ff.records = {a, b};
    std::string err;
    check(ff.write(fn, &err), "write field file: " + err);
    FieldFile back;
    check(back.read(fn, &err) && back.records.size() == 2 && back.records[0].f64 == a.f64, "read back");
    check(FieldFile::append(fn, {a}, &err), "append");
    check(back.read(fn, &err) && back.records.size() == 3, "read appended");
    const std::vector<char> good = slurp(fn);
    // every truncation fails cleanly
    for (size_t cut = 0; cut < good.size(); cut += 3) {
        spit(dir + "/t.lbmf", std::vector<char>(good.begin(), good.begin() + (long)cut));
        check(!back.read(dir + "/t.lbmf", &err), "truncated file at " + std::to_string(cut) + " accepted");
    }
    // bad magic, unknown type, a count far beyond the file
    std::vector<char> bad = good;
    bad[0] = 'X';
    spit(dir + "/t.lbmf", bad);
    check(!back.read(dir + "/t.lbmf", &err), "bad magic accepted");
    bad = good;
    bad[12 + 32] = 7;  // type of record 0
    spit(dir + "/t.lbmf", bad);
    check(!back.read(dir + "/t.lbmf", &err), "unknown type accepted");
    bad = good;
    const uint64_t huge = 1ull << 60;
    std::memcpy(&bad[12 + 36], &huge, 8);  // count of record 0
    spit(dir + "/t.lbmf", bad);
    check(!back.read(dir + "/t.lbmf", &err), "oversized count accepted");
    const uint32_t many = 0xffffffffu;
    bad = good;
    std::memcpy(&bad[8], &many, 4);  // record count
    spit(dir + "/t.lbmf", bad);
    check(!back.read(dir + "/t.lbmf", &err), "oversized record count accepted");
    FieldSource src;
    check(src.open(fn, &err), "field source");
    check(src.get("Porosity", 4, 1, &err) != nullptr, "field source get");
    check(src.get("Absent", 4, 1, &err) == nullptr, "absent field");
    check(!src.open(dir + "/none.lbmf", &err), "missing field file");
}

static void cgns(const std::string &dir)
{
    const std::string fn = dir + "/out.cgns";
    std::string err;
    SolutionSink sink;
    check(sink.create(fn, 5, 4, 3, &err), "cgns grid: " + err);
    std::vector<FieldRecord> recs(2);
    recs[0].name = "VelocityX";
    recs[1].name = "Density";
    for (auto &r : recs) {
        r.type = 1;
        r.f64.assign(60, 0.0);
        for (size_t k = 0; k < 60; k++) r.f64[k] = 0.001 * (double)k;
    }
    check(sink.append(10, recs, &err), "cgns append: " + err);
    check(sink.append(20, recs, &err), "cgns append 2: " + err);
    FieldSource src;
    check(src.open(fn, &err), "cgns source: " + err);
    const FieldRecord *r = src.get("Density", 60, 1, &err);
    check(r && r->f64 == recs[1].f64, "cgns read back");
    check(src.get("Density", 59, 1, &err) != nullptr, "cached field");  // already read: the cached record
    check(src.get("VelocityX", 61, 1, &err) == nullptr, "wrong count accepted");
    check(src.get("Missing", 60, 2, &err) == nullptr, "missing cgns field");
    for (int k = 0; k < 70; k++) src.get("VelocityX", 60, 1, &err);

    // not an HDF5 file at all
    spit(dir + "/junk.cgns", std::vector<char>(100, 'j'));
    FieldSource junk;
    if (junk.open(dir + "/junk.cgns", &err)) check(junk.get("Density", 60, 1, &err) == nullptr, "junk cgns read");
}

int main(int argc, char **argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    reader(dir);
    fields(dir);
    cgns(dir);
    std::printf("%s\n", failures ? "host sanitize: FAILED" : "host sanitize: ok");
    return failures ? 1 : 0;
}
