/*
 * Drives the CGNS/HDF5 reader and writer (lbm-c_amd/host/cgns_io.c) for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer: a grid and
 * solutions written and read back, bad arguments, a missing file, a file
 * that is not HDF5, a field of the wrong size, and zone-size datasets of the
 * wrong rank and shape, which appending must refuse.  Linked with
 * cgns_io.c and libhdf5.  Exit status 0 when every call returned what it
 * should.
 */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>

int cgio_write_grid(const char *fname, int Lx, int Ly, int Lz);
int cgio_append_solution(const char *fname, int iter, int nfields, const double *const *data,
                         const char *const *labels);
int cgio_read_field(const char *fname, const char *label, double *out, long long n);
int cgio_read_field_int(const char *fname, const char *label, int *out, long long n);
const char *cgio_last_error(void);
void cgio_quiet(void);

static int failures = 0;
#define CHECK(c, what)                                                                                      \
    do {                                                                                                    \
        if (!(c)) {                                                                                         \
            printf("FAIL: %s (%s)\n", what, cgio_last_error());                                             \
            failures++;                                                                                     \
        }                                                                                                   \
    } while (0)

/* replace the zone's " data" (its size array) by an int32 dataset of the
 * given rank and extents */
static void replace_zone_size(const char *fn, int rank, const hsize_t *dims)
{
    hid_t f = H5Fopen(fn, H5F_ACC_RDWR, H5P_DEFAULT);
    hid_t zone = H5Gopen2(f, "/Base/LBM-C Output Data", H5P_DEFAULT);
    H5Ldelete(zone, " data", H5P_DEFAULT);
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t ds = H5Dcreate2(zone, " data", H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int v[64];
    for (int k = 0; k < 64; k++) v[k] = 4;
    H5Dwrite(ds, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
    H5Dclose(ds);
    H5Sclose(sp);
    H5Gclose(zone);
    H5Fclose(f);
}

int main(int argc, char **argv)
{
    char fn[4096], junk[4096];
    const char *dir = argc > 1 ? argv[1] : ".";
    snprintf(fn, sizeof fn, "%s/c.cgns", dir);
    snprintf(junk, sizeof junk, "%s/junk.cgns", dir);
    cgio_quiet();
    double a[24], b[24], back[24];
    for (int k = 0; k < 24; k++) {
        a[k] = 0.5 * k;
        b[k] = -1.0 * k;
    }
    const double *data[2] = {a, b};
    const char *labels[2] = {"Density", "VelocityX"};
    CHECK(cgio_write_grid(fn, 4, 3, 2) == 0, "write grid");
    CHECK(cgio_append_solution(fn, 5, 2, data, labels) == 0, "append");
    CHECK(cgio_append_solution(fn, 5, 2, data, labels) == 0, "append the same iteration again");
    CHECK(cgio_append_solution(fn, 6, 0, NULL, NULL) == 0, "append an empty solution");
    CHECK(cgio_read_field(fn, "Density", back, 24) == 0 && memcmp(back, a, sizeof a) == 0, "read back");
    int iv[24];
    CHECK(cgio_read_field_int(fn, "VelocityX", iv, 24) == 0 && iv[3] == -3, "read as int");
    CHECK(cgio_read_field(fn, "Density", back, 23) < 0, "wrong count refused");
    CHECK(cgio_read_field(fn, "Absent", back, 24) == 1, "absent field");
    CHECK(cgio_append_solution(fn, 7, 1, NULL, labels) < 0, "null data refused");
    CHECK(cgio_append_solution(NULL, 7, 1, data, labels) < 0, "null name refused");
    CHECK(cgio_read_field("/nonexistent/x.cgns", "Density", back, 24) < 0, "missing file");
    FILE *j = fopen(junk, "wb");
    for (int k = 0; k < 200; k++) fputc('j', j);
    fclose(j);
    CHECK(cgio_read_field(junk, "Density", back, 24) < 0, "not an HDF5 file");
    CHECK(cgio_append_solution(junk, 1, 2, data, labels) < 0, "append to a non-HDF5 file");
    CHECK(cgio_write_grid(fn, 4, 3, 1) == 0, "2-D grid");
    CHECK(cgio_append_solution(fn, 1, 2, data, labels) == 0, "2-D append");
    const hsize_t d3[3] = {3, 3, 3}, d1[1] = {9}, dbad[2] = {2, 3}, dwide[2] = {3, 7};
    replace_zone_size(fn, 3, d3);
    CHECK(cgio_append_solution(fn, 2, 2, data, labels) < 0, "rank-3 zone size refused");
    replace_zone_size(fn, 1, d1);
    CHECK(cgio_append_solution(fn, 2, 2, data, labels) < 0, "rank-1 zone size refused");
    replace_zone_size(fn, 2, dbad);
    CHECK(cgio_append_solution(fn, 2, 2, data, labels) < 0, "2 x 3 zone size refused");
    replace_zone_size(fn, 2, dwide);
    CHECK(cgio_append_solution(fn, 2, 2, data, labels) < 0, "3 x 7 zone size refused");
    printf("%s\n", failures ? "cgns sanitize: FAILED" : "cgns sanitize: ok");
    return failures ? 1 : 0;
}
