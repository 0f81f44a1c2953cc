/*
 * CGNS files in the HDF5 mapping, written and read with the HDF5 C library
 * (no libcgns): the structured-zone output of LBM-C's CGNSOutputHandler
 * (LBM-C/cgns/cgns_output_handler.cu:62-205,264-293) and the field input of
 * its CGNSInputHandler (LBM-C/cgns/cgns_input_handler.cu:57-103).
 *
 * Node mapping: every CGNS node is an HDF5 group (links tracked in creation
 * order) with the attributes "name" and "label" (33-byte strings), "type"
 * (3-byte string: MT, I4, R4, R8, C1) and "flags" (int32[1] = 1); node data
 * is the dataset " data" with the CGNS (Fortran-order) dimensions reversed.
 * The root group is the "HDF5 MotherNode" with the " format" and
 * " hdf5version" datasets and the CGNSLibraryVersion node.
 *
 * Written tree (2D: no CoordinateZ, cell dimension 2):
 *   Base (CGNSBase_t, I4 {cell_dim, phys_dim})
 *     LBM-C Output Data (Zone_t, I4 {vertex sizes, cell sizes, 0...})
 *       ZoneType (C1 "Structured")
 *       GridCoordinates / CoordinateX, CoordinateY[, CoordinateZ] (R4, vertex i, j, k)
 *       Solution @ Time = <iter> (FlowSolution_t) / GridLocation (C1 "CellCenter"),
 *                                                   <fields> (R8, cells)
 *       ZoneIterativeData_t (ZoneIterativeData_t) / FlowSolutionPointers (C1, 32 x n)
 *     TimeIterValues (BaseIterativeData_t, I4 n) / IterationValues (I4, n)
 *     SimulationType (C1 "TimeAccurate")
 *
 * Every function returns 0 on success and -1 on failure (cgio_last_error).
 * Plain C: the library links only libc and libhdf5.
 */
#define _POSIX_C_SOURCE 200809L
#include <hdf5.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NAME_LEN 33

static char g_err[512];

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

const char *cgio_last_error(void) { return g_err; }

static int put_str_attr(hid_t obj, const char *attr, const char *value, size_t len)
{
    char buf[NAME_LEN];
    memset(buf, 0, sizeof(buf));
    strncpy(buf, value, len - 1);
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, len);
    hid_t s = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(obj, attr, t, s, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, t, buf);
    if (a >= 0) H5Aclose(a);
    H5Sclose(s);
    H5Tclose(t);
    return e < 0 ? -1 : 0;
}

static int get_str_attr(hid_t obj, const char *attr, char *out, size_t len)
{
    memset(out, 0, len);
    if (H5Aexists(obj, attr) <= 0) return -1;
    hid_t a = H5Aopen(obj, attr, H5P_DEFAULT);
    if (a < 0) return -1;
    hid_t ft = H5Aget_type(a);
    size_t fl = H5Tget_size(ft);
    hid_t t = H5Tcopy(H5T_C_S1);
    H5Tset_size(t, fl);
    char *tmp = (char *)calloc(fl + 1, 1);
    herr_t e = H5Aread(a, t, tmp);
    strncpy(out, tmp, len - 1);
    free(tmp);
    H5Tclose(t);
    H5Tclose(ft);
    H5Aclose(a);
    return e < 0 ? -1 : 0;
}

static hid_t group_props(void)
{
    hid_t p = H5Pcreate(H5P_GROUP_CREATE);
    H5Pset_link_creation_order(p, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    return p;
}

static int set_node_attrs(hid_t g, const char *name, const char *label, const char *type)
{
    int flags = 1;
    hsize_t one = 1;
    if (put_str_attr(g, "name", name, NAME_LEN) || put_str_attr(g, "label", label, NAME_LEN) ||
        put_str_attr(g, "type", type, 3))
        return -1;
    hid_t s = H5Screate_simple(1, &one, NULL);
    hid_t a = H5Acreate2(g, "flags", H5T_NATIVE_INT32, s, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = a < 0 ? -1 : H5Awrite(a, H5T_NATIVE_INT32, &flags);
    if (a >= 0) H5Aclose(a);
    H5Sclose(s);
    return e < 0 ? -1 : 0;
}

/* data of a node: CGNS dims (Fortran order) -> HDF5 dims reversed */
static int write_data(hid_t g, hid_t memtype, hid_t filetype, int ndim, const long long *cgdims, const void *buf)
{
    hsize_t dims[8];
    for (int k = 0; k < ndim; k++) dims[k] = (hsize_t)cgdims[ndim - 1 - k];
    hid_t s = H5Screate_simple(ndim, dims, NULL);
    hid_t d = H5Dcreate2(g, " data", filetype, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t e = d < 0 ? -1 : H5Dwrite(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
    if (d >= 0) H5Dclose(d);
    H5Sclose(s);
    return e < 0 ? -1 : 0;
}

/* a node with optional data; returns the open group (caller closes) */
static hid_t make_node(hid_t parent, const char *name, const char *label, const char *type, hid_t memtype,
                       hid_t filetype, int ndim, const long long *cgdims, const void *buf)
{
    hid_t p = group_props();
    hid_t g = H5Gcreate2(parent, name, H5P_DEFAULT, p, H5P_DEFAULT);
    H5Pclose(p);
    if (g < 0) return -1;
    if (set_node_attrs(g, name, label, type) ||
        (buf && write_data(g, memtype, filetype, ndim, cgdims, buf))) {
        H5Gclose(g);
        return -1;
    }
    return g;
}

static int leaf(hid_t parent, const char *name, const char *label, const char *type, hid_t memtype, hid_t filetype,
                int ndim, const long long *cgdims, const void *buf)
{
    hid_t g = make_node(parent, name, label, type, memtype, filetype, ndim, cgdims, buf);
    if (g < 0) return -1;
    H5Gclose(g);
    return 0;
}

static int leaf_str(hid_t parent, const char *name, const char *label, const char *value)
{
    long long n = (long long)strlen(value);
    return leaf(parent, name, label, "C1", H5T_NATIVE_CHAR, H5T_NATIVE_CHAR, 1, &n, value);
}

int cgio_write_grid(const char *fname, int Lx, int Ly, int Lz)
{
    if (!fname || Lx <= 0 || Ly <= 0 || Lz <= 0) return fail("cgio_write_grid: bad argument");
    const int three = Lz > 1;
    const int dim = three ? 3 : 2;
    hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
    H5Pset_link_creation_order(fcpl, H5P_CRT_ORDER_TRACKED | H5P_CRT_ORDER_INDEXED);
    hid_t f = H5Fcreate(fname, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
    H5Pclose(fcpl);
    if (f < 0) return fail("cgio_write_grid: cannot create %s", fname);
    int rc = 0;
    hid_t root = H5Gopen2(f, "/", H5P_DEFAULT);
    rc |= set_node_attrs(root, "HDF5 MotherNode", "Root Node of HDF5 File", "MT");
    {
        const char fmt[] = "IEEE_LITTLE_32";
        char ver[NAME_LEN];
        unsigned maj, min, rel;
        H5get_libversion(&maj, &min, &rel);
        memset(ver, 0, sizeof(ver));
        snprintf(ver, sizeof(ver), "HDF5 Version %u.%u.%u", maj, min, rel);
        hsize_t n1 = sizeof(fmt) - 1, n2 = NAME_LEN;
        hid_t s1 = H5Screate_simple(1, &n1, NULL), s2 = H5Screate_simple(1, &n2, NULL);
        hid_t d1 = H5Dcreate2(root, " format", H5T_NATIVE_CHAR, s1, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        hid_t d2 = H5Dcreate2(root, " hdf5version", H5T_NATIVE_CHAR, s2, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        rc |= (d1 < 0 || H5Dwrite(d1, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, fmt) < 0);
        rc |= (d2 < 0 || H5Dwrite(d2, H5T_NATIVE_CHAR, H5S_ALL, H5S_ALL, H5P_DEFAULT, ver) < 0);
        if (d1 >= 0) H5Dclose(d1);
        if (d2 >= 0) H5Dclose(d2);
        H5Sclose(s1);
        H5Sclose(s2);
    }
    {
        const float version = 3.1f;
        const long long one = 1;
        rc |= leaf(root, "CGNSLibraryVersion", "CGNSLibraryVersion_t", "R4", H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, 1,
                   &one, &version);
    }
    const int bdims[2] = {dim, dim};
    const long long two = 2;
    hid_t base = make_node(root, "Base", "CGNSBase_t", "I4", H5T_NATIVE_INT32, H5T_STD_I32LE, 1, &two, bdims);
    if (base < 0) rc = -1;
    hid_t zone = -1;
    if (base >= 0) {
        /* isize[3][dim]: vertex sizes, cell sizes, boundary vertex sizes (0) */
        int isize[9];
        const int L[3] = {Lx, Ly, Lz};
        for (int d = 0; d < dim; d++) {
            isize[d] = L[d] + 1;
            isize[dim + d] = L[d];
            isize[2 * dim + d] = 0;
        }
        const long long zd[2] = {dim, 3};
        zone = make_node(base, "LBM-C Output Data", "Zone_t", "I4", H5T_NATIVE_INT32, H5T_STD_I32LE, 2, zd, isize);
        if (zone < 0) rc = -1;
    }
    if (zone >= 0) {
        rc |= leaf_str(zone, "ZoneType", "ZoneType_t", "Structured");
        hid_t gc = make_node(zone, "GridCoordinates", "GridCoordinates_t", "MT", 0, 0, 0, NULL, NULL);
        if (gc < 0) rc = -1;
        else {
            const long long ni = Lx + 1, nj = Ly + 1, nk = three ? Lz + 1 : 1;
            const long long nv = ni * nj * nk;
            float *c = (float *)malloc(sizeof(float) * (size_t)nv);
            const char *names[3] = {"CoordinateX", "CoordinateY", "CoordinateZ"};
            const long long vd[3] = {ni, nj, nk};
            for (int a = 0; a < dim && c; a++) {
                for (long long k = 0; k < nk; k++)
                    for (long long j = 0; j < nj; j++)
                        for (long long i = 0; i < ni; i++)
                            c[(k * nj + j) * ni + i] = (float)(a == 0 ? i : (a == 1 ? j : k));
                rc |= leaf(gc, names[a], "DataArray_t", "R4", H5T_NATIVE_FLOAT, H5T_IEEE_F32LE, dim, vd, c);
            }
            if (!c) rc = -1;
            free(c);
            H5Gclose(gc);
        }
        H5Gclose(zone);
    }
    if (base >= 0) H5Gclose(base);
    H5Gclose(root);
    H5Fclose(f);
    return rc ? fail("cgio_write_grid: HDF5 error writing %s", fname) : 0;
}

static int read_dims(hid_t zone, int *dim, int *cells)
{
    hid_t d = H5Dopen2(zone, " data", H5P_DEFAULT);
    if (d < 0) return -1;
    hid_t s = H5Dget_space(d);
    hsize_t dims[2] = {0, 0};
    /* the zone's size array is 3 x dim (reversed: dims = {3, dim}); check the
     * rank before reading the extent into a two-entry array */
    int nd = H5Sget_simple_extent_ndims(s);
    if (nd == 2) nd = H5Sget_simple_extent_dims(s, dims, NULL);
    int isize[9];
    int rc = (nd != 2 || dims[0] != 3 || dims[1] < 2 || dims[1] > 3) ? -1 : 0;
    if (!rc) rc = H5Dread(d, H5T_NATIVE_INT32, H5S_ALL, H5S_ALL, H5P_DEFAULT, isize) < 0 ? -1 : 0;
    if (!rc) {
        *dim = (int)dims[1];
        for (int k = 0; k < *dim; k++) cells[k] = isize[*dim + k];
    }
    H5Sclose(s);
    H5Dclose(d);
    return rc;
}

int cgio_append_solution(const char *fname, int iter, int nfields, const double *const *data,
                         const char *const *labels)
{
    if (!fname || nfields < 0 || (nfields && (!data || !labels))) return fail("cgio_append_solution: bad argument");
    hid_t f = H5Fopen(fname, H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) return fail("cgio_append_solution: cannot open %s", fname);
    int rc = 0;
    hid_t base = H5Gopen2(f, "/Base", H5P_DEFAULT);
    hid_t zone = base >= 0 ? H5Gopen2(base, "LBM-C Output Data", H5P_DEFAULT) : -1;
    int dim = 0, cells[3] = {1, 1, 1};
    if (zone < 0 || read_dims(zone, &dim, cells)) {
        if (zone >= 0) H5Gclose(zone);
        if (base >= 0) H5Gclose(base);
        H5Fclose(f);
        return fail("cgio_append_solution: %s has no LBM-C zone", fname);
    }
    char sname[NAME_LEN];
    snprintf(sname, sizeof(sname), "Solution @ Time = %d", iter);
    if (H5Lexists(zone, sname, H5P_DEFAULT) > 0) H5Ldelete(zone, sname, H5P_DEFAULT);
    hid_t sol = make_node(zone, sname, "FlowSolution_t", "MT", 0, 0, 0, NULL, NULL);
    if (sol < 0) rc = -1;
    else {
        rc |= leaf_str(sol, "GridLocation", "GridLocation_t", "CellCenter");
        const long long cd[3] = {cells[0], cells[1], cells[2]};
        for (int n = 0; n < nfields; n++)
            rc |= leaf(sol, labels[n], "DataArray_t", "R8", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, dim, cd, data[n]);
        H5Gclose(sol);
    }
    /* the iterative data lists every solution node of the zone, in creation order */
    int nsol = 0;
    int *iters = NULL;
    char *names = NULL;
    {
        H5G_info_t info;
        H5Gget_info(zone, &info);
        iters = (int *)calloc(info.nlinks + 1, sizeof(int));
        names = (char *)calloc((info.nlinks + 1) * 32, 1);
        for (hsize_t k = 0; k < info.nlinks && iters && names; k++) {
            char lname[256];
            if (H5Lget_name_by_idx(zone, ".", H5_INDEX_CRT_ORDER, H5_ITER_INC, k, lname, sizeof(lname), H5P_DEFAULT) < 0)
                continue;
            hid_t g = H5Gopen2(zone, lname, H5P_DEFAULT);
            if (g < 0) continue;
            char lab[NAME_LEN];
            if (get_str_attr(g, "label", lab, sizeof(lab)) == 0 && strcmp(lab, "FlowSolution_t") == 0) {
                int it = 0;
                if (sscanf(lname, "Solution @ Time = %d", &it) == 1) {
                    iters[nsol] = it;
                    memcpy(names + 32 * nsol, lname, strnlen(lname, 32));
                    nsol++;
                }
            }
            H5Gclose(g);
        }
    }
    if (H5Lexists(base, "TimeIterValues", H5P_DEFAULT) > 0) H5Ldelete(base, "TimeIterValues", H5P_DEFAULT);
    if (H5Lexists(base, "SimulationType", H5P_DEFAULT) > 0) H5Ldelete(base, "SimulationType", H5P_DEFAULT);
    if (H5Lexists(zone, "ZoneIterativeData_t", H5P_DEFAULT) > 0) H5Ldelete(zone, "ZoneIterativeData_t", H5P_DEFAULT);
    {
        const long long one = 1, n = nsol;
        hid_t bi = make_node(base, "TimeIterValues", "BaseIterativeData_t", "I4", H5T_NATIVE_INT32, H5T_STD_I32LE, 1,
                             &one, &nsol);
        if (bi < 0) rc = -1;
        else {
            rc |= leaf(bi, "IterationValues", "DataArray_t", "I4", H5T_NATIVE_INT32, H5T_STD_I32LE, 1, &n, iters);
            H5Gclose(bi);
        }
        hid_t zi = make_node(zone, "ZoneIterativeData_t", "ZoneIterativeData_t", "MT", 0, 0, 0, NULL, NULL);
        if (zi < 0) rc = -1;
        else {
            const long long pd[2] = {32, n};
            rc |= leaf(zi, "FlowSolutionPointers", "DataArray_t", "C1", H5T_NATIVE_CHAR, H5T_NATIVE_CHAR, 2, pd, names);
            H5Gclose(zi);
        }
        rc |= leaf_str(base, "SimulationType", "SimulationType_t", "TimeAccurate");
    }
    free(iters);
    free(names);
    H5Gclose(zone);
    H5Gclose(base);
    H5Fclose(f);
    return rc ? fail("cgio_append_solution: HDF5 error writing %s", fname) : 0;
}

/* first child (creation order) of g whose "label" attribute is `label`;
 * with `name` set, the child must also carry that "name" */
static hid_t find_child(hid_t g, const char *label, const char *name)
{
    H5G_info_t info;
    if (H5Gget_info(g, &info) < 0) return -1;
    H5_index_t idx = H5_INDEX_CRT_ORDER;
    char lname[256];
    if (info.nlinks > 0 &&
        H5Lget_name_by_idx(g, ".", idx, H5_ITER_INC, 0, lname, sizeof(lname), H5P_DEFAULT) < 0)
        idx = H5_INDEX_NAME; /* file written without creation-order tracking */
    for (hsize_t k = 0; k < info.nlinks; k++) {
        if (H5Lget_name_by_idx(g, ".", idx, H5_ITER_INC, k, lname, sizeof(lname), H5P_DEFAULT) < 0) continue;
        H5O_info_t oi;
        if (H5Oget_info_by_name(g, lname, &oi, H5P_DEFAULT) < 0 || oi.type != H5O_TYPE_GROUP) continue;
        hid_t c = H5Gopen2(g, lname, H5P_DEFAULT);
        if (c < 0) continue;
        char lab[NAME_LEN], nm[NAME_LEN];
        int ok = get_str_attr(c, "label", lab, sizeof(lab)) == 0 && strcmp(lab, label) == 0;
        if (ok && name) ok = get_str_attr(c, "name", nm, sizeof(nm)) == 0 && strcmp(nm, name) == 0;
        if (ok) return c;
        H5Gclose(c);
    }
    return -1;
}

/* Field `label` of base 1 / zone 1 / solution 1 into out[n] (converted to
 * double or int32 by HDF5).  Returns 1 when the field does not exist. */
static int read_field(const char *fname, const char *label, void *out, long long n, hid_t memtype)
{
    hid_t f = H5Fopen(fname, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cgio_read_field: cannot open %s", fname);
    hid_t root = H5Gopen2(f, "/", H5P_DEFAULT);
    hid_t base = find_child(root, "CGNSBase_t", NULL);
    hid_t zone = base >= 0 ? find_child(base, "Zone_t", NULL) : -1;
    hid_t sol = zone >= 0 ? find_child(zone, "FlowSolution_t", NULL) : -1;
    hid_t arr = sol >= 0 ? find_child(sol, "DataArray_t", label) : -1;
    int rc;
    if (sol < 0) rc = fail("cgio_read_field: %s has no base/zone/solution", fname);
    else if (arr < 0) rc = 1;
    else {
        hid_t d = H5Dopen2(arr, " data", H5P_DEFAULT);
        hid_t s = d >= 0 ? H5Dget_space(d) : -1;
        const long long count = s >= 0 ? (long long)H5Sget_simple_extent_npoints(s) : -1;
        if (count != n) rc = fail("cgio_read_field: %s has %lld values, expected %lld", label, count, n);
        else rc = H5Dread(d, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0 ? fail("cgio_read_field: read error") : 0;
        if (s >= 0) H5Sclose(s);
        if (d >= 0) H5Dclose(d);
    }
    if (arr >= 0) H5Gclose(arr);
    if (sol >= 0) H5Gclose(sol);
    if (zone >= 0) H5Gclose(zone);
    if (base >= 0) H5Gclose(base);
    H5Gclose(root);
    H5Fclose(f);
    return rc;
}

int cgio_read_field(const char *fname, const char *label, double *out, long long n)
{
    return read_field(fname, label, out, n, H5T_NATIVE_DOUBLE);
}

int cgio_read_field_int(const char *fname, const char *label, int *out, long long n)
{
    return read_field(fname, label, out, n, H5T_NATIVE_INT32);
}

/* Quiet HDF5's automatic error stack printing (errors come back as codes). */
void cgio_quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }
