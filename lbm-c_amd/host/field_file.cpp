// LBMCFLD1 field container (format in field_file.hpp).
#include "field_file.hpp"

#include <cstdio>
#include <cstring>

namespace lbmc_host {

namespace {

const char kMagic[8] = {'L', 'B', 'M', 'C', 'F', 'L', 'D', '1'};

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

}  // namespace lbmc_host
