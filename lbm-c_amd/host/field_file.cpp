// LBMCFLD1 field container (format in field_file.hpp).
#include "field_file.hpp"

#include <dlfcn.h>
#include <unistd.h>

#include <climits>
#include <cstdio>
#include <cstring>

namespace lbmc_host {

namespace {

const char kMagic[8] = {'L', 'B', 'M', 'C', 'F', 'L', 'D', '1'};

// bytes between the read position and the end of the file
uint64_t bytes_left(FILE *fp)
{
    const long pos = ftell(fp);
    if (pos < 0 || fseek(fp, 0, SEEK_END) != 0) return 0;
    const long end = ftell(fp);
    if (end < pos || fseek(fp, pos, SEEK_SET) != 0) return 0;
    return (uint64_t)(end - pos);
}

bool read_records(FILE *fp, std::vector<FieldRecord> *out, std::string *err)
{
    char magic[8];
    uint32_t n = 0;
    if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, kMagic, 8) != 0) {
        if (err) *err = "not an LBMCFLD1 file";
        return false;
    }
    if (fread(&n, 4, 1, fp) != 1) {
        if (err) *err = "truncated header";
        return false;
    }
    for (uint32_t r = 0; r < n; r++) {
        char name[33] = {};
        FieldRecord rec;
        uint64_t count = 0;
        if (fread(name, 1, 32, fp) != 32 || fread(&rec.type, 4, 1, fp) != 1 || fread(&count, 8, 1, fp) != 1) {
            if (err) *err = "truncated record header";
            return false;
        }
        rec.name = name;
        // the count comes from the file: check it against what the file
        // still holds before allocating for it
        const uint64_t esize = rec.type == 1 ? 8 : 4;
        if ((rec.type == 1 || rec.type == 2) && count > bytes_left(fp) / esize) {
            if (err) *err = "truncated record " + rec.name;
            return false;
        }
        if (rec.type == 1) {
            rec.f64.resize(count);
            if (fread(rec.f64.data(), 8, count, fp) != count) {
                if (err) *err = "truncated record " + rec.name;
                return false;
            }
        } else if (rec.type == 2) {
            rec.i32.resize(count);
            if (fread(rec.i32.data(), 4, count, fp) != count) {
                if (err) *err = "truncated record " + rec.name;
                return false;
            }
        } else {
            if (err) *err = "unknown record type in " + rec.name;
            return false;
        }
        out->push_back(std::move(rec));
    }
    return true;
}

bool write_records(FILE *fp, const std::vector<FieldRecord> &recs)
{
    const uint32_t n = (uint32_t)recs.size();
    if (fwrite(kMagic, 1, 8, fp) != 8 || fwrite(&n, 4, 1, fp) != 1) return false;
    for (const FieldRecord &rec : recs) {
        char name[32] = {};
        strncpy(name, rec.name.c_str(), 31);
        const uint64_t count = rec.count();
        if (fwrite(name, 1, 32, fp) != 32 || fwrite(&rec.type, 4, 1, fp) != 1 || fwrite(&count, 8, 1, fp) != 1)
            return false;
        if (rec.type == 1 && fwrite(rec.f64.data(), 8, count, fp) != count) return false;
        if (rec.type == 2 && fwrite(rec.i32.data(), 4, count, fp) != count) return false;
    }
    return true;
}

}  // namespace

bool FieldFile::read(const std::string &fname, std::string *err)
{
    FILE *fp = fopen(fname.c_str(), "rb");
    if (!fp) {
        if (err) *err = "cannot open " + fname;
        return false;
    }
    records.clear();
    const bool ok = read_records(fp, &records, err);
    fclose(fp);
    return ok;
}

bool FieldFile::write(const std::string &fname, std::string *err) const
{
    FILE *fp = fopen(fname.c_str(), "wb");
    if (!fp) {
        if (err) *err = "cannot create " + fname;
        return false;
    }
    const bool ok = write_records(fp, records);
    fclose(fp);
    if (!ok && err) *err = "write failed: " + fname;
    return ok;
}

bool FieldFile::append(const std::string &fname, const std::vector<FieldRecord> &recs, std::string *err)
{
    FieldFile f;
    FILE *probe = fopen(fname.c_str(), "rb");
    if (probe) {
        fclose(probe);
        if (!f.read(fname, err)) return false;
    }
    for (const FieldRecord &r : recs) f.records.push_back(r);
    return f.write(fname, err);
}

const FieldRecord *FieldFile::find(const std::string &name) const
{
    for (const FieldRecord &r : records)
        if (r.name == name) return &r;
    return nullptr;
}

bool is_cgns(const std::string &fname)
{
    return fname.size() >= 5 && fname.compare(fname.size() - 5, 5, ".cgns") == 0;
}

namespace {

// the C entry points of libcgnsio.so (host/cgns_io.c)
struct CgnsIO {
    int (*write_grid)(const char *, int, int, int);
    int (*append_solution)(const char *, int, int, const double *const *, const char *const *);
    int (*read_field)(const char *, const char *, double *, long long);
    int (*read_field_int)(const char *, const char *, int *, long long);
    const char *(*last_error)(void);
    void (*quiet)(void);
};

const CgnsIO *cgns_io(std::string *err)
{
    static CgnsIO io;
    static int state = 0;  // 0 not tried, 1 loaded, -1 failed
    static std::string why;
    if (state == 0) {
        char exe[PATH_MAX];
        const ssize_t len = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
        std::string dir = ".";
        if (len > 0) {
            exe[len] = 0;
            dir = exe;
            dir = dir.substr(0, dir.find_last_of('/'));
        }
        const std::string path = dir + "/libcgnsio.so";
        void *h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            why = "CGNS support unavailable (" + path + ": " + dlerror() + ")";
            state = -1;
        } else {
            io.write_grid = (int (*)(const char *, int, int, int))dlsym(h, "cgio_write_grid");
            io.append_solution = (int (*)(const char *, int, int, const double *const *, const char *const *))dlsym(
                h, "cgio_append_solution");
            io.read_field = (int (*)(const char *, const char *, double *, long long))dlsym(h, "cgio_read_field");
            io.read_field_int = (int (*)(const char *, const char *, int *, long long))dlsym(h, "cgio_read_field_int");
            io.last_error = (const char *(*)(void))dlsym(h, "cgio_last_error");
            io.quiet = (void (*)(void))dlsym(h, "cgio_quiet");
            if (!io.write_grid || !io.append_solution || !io.read_field || !io.read_field_int || !io.last_error ||
                !io.quiet) {
                why = "CGNS support unavailable (" + path + " lacks cgio_* symbols)";
                state = -1;
            } else {
                io.quiet();
                state = 1;
            }
        }
    }
    if (state < 0) {
        if (err) *err = why;
        return nullptr;
    }
    return &io;
}

}  // namespace

bool FieldSource::open(const std::string &fname, std::string *err)
{
    fname_ = fname;
    cgns_ = is_cgns(fname);
    if (!cgns_) return file.read(fname, err);
    if (!cgns_io(err)) return false;
    file.records.reserve(64);  // get() hands out references into this vector
    if (FILE *fp = fopen(fname.c_str(), "rb")) {
        fclose(fp);
        return true;
    }
    if (err) *err = "cannot open " + fname;
    return false;
}

const FieldRecord *FieldSource::get(const std::string &label, long long n, uint32_t type, std::string *err)
{
    if (const FieldRecord *r = file.find(label)) return r;
    if (!cgns_) return nullptr;
    const CgnsIO *io = cgns_io(err);
    if (!io) return nullptr;
    FieldRecord rec;
    rec.name = label;
    rec.type = type;
    int rc;
    if (type == 2) {
        rec.i32.resize((size_t)n);
        rc = io->read_field_int(fname_.c_str(), label.c_str(), rec.i32.data(), n);
    } else {
        rec.f64.resize((size_t)n);
        rc = io->read_field(fname_.c_str(), label.c_str(), rec.f64.data(), n);
    }
    if (rc != 0) {
        if (rc < 0 && err) *err = io->last_error();
        return nullptr;
    }
    if (file.records.size() == file.records.capacity()) {
        if (err) *err = "too many CGNS fields requested";
        return nullptr;
    }
    file.records.push_back(std::move(rec));
    return &file.records.back();
}

bool SolutionSink::create(const std::string &fname, int Lx, int Ly, int Lz, std::string *err)
{
    fname_ = fname;
    cgns_ = is_cgns(fname);
    if (cgns_) {
        const CgnsIO *io = cgns_io(err);
        if (!io) return false;
        if (io->write_grid(fname.c_str(), Lx, Ly, Lz) != 0) {
            if (err) *err = io->last_error();
            return false;
        }
        return true;
    }
    FieldFile out;
    FieldRecord dims;
    dims.name = "Dims";
    dims.type = 2;
    dims.i32 = {Lx, Ly, Lz};
    out.records.push_back(dims);
    return out.write(fname, err);
}

bool SolutionSink::append(long long iter, const std::vector<FieldRecord> &fields, std::string *err)
{
    if (cgns_) {
        const CgnsIO *io = cgns_io(err);
        if (!io) return false;
        std::vector<const double *> data;
        std::vector<const char *> labels;
        for (const FieldRecord &r : fields) {
            data.push_back(r.f64.data());
            labels.push_back(r.name.c_str());
        }
        if (io->append_solution(fname_.c_str(), (int)iter, (int)fields.size(), data.data(), labels.data()) != 0) {
            if (err) *err = io->last_error();
            return false;
        }
        return true;
    }
    std::vector<FieldRecord> recs = fields;
    for (FieldRecord &r : recs) r.name += "@" + std::to_string(iter);
    return FieldFile::append(fname_, recs, err);
}

}  // namespace lbmc_host
