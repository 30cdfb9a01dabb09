/*
 * Minimal HDF5 reader/writer for the file contract of the afterglow code:
 *   - initial-data file read by EjectaID2::_load_id_file (ejecta/ejecta_id.h:222-276):
 *     2-D [nshells x nlayers] double datasets r, mom, theta, ctheta, ek, mass;
 *   - light-curve file written by Ejecta::saveDenseLightCurves (ejecta/ejecta.h:511-573):
 *     groups "shell=i layer=j" holding "fluxdens" [n_pairs], top-level "times",
 *     "freqs" [n_pairs] and scalar double attributes on the file;
 *   - generic 1-D double dataset and double attribute readers (used to read the
 *     dynamics / spectrum / light-curve files of either implementation).
 * Plain C over the HDF5 C API; every entry point returns <0 on error.
 */
#include <hdf5.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

static int write_1d(hid_t loc, const char *name, long n, const double *data);

static int write_2d(hid_t file, const char *name, int n0, int n1, const double *data) {
    hsize_t dims[2] = {(hsize_t)n0, (hsize_t)n1};
    hid_t sp = H5Screate_simple(2, dims, NULL);
    hid_t ds = H5Dcreate2(file, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) { H5Sclose(sp); return -1; }
    herr_t st = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds); H5Sclose(sp);
    return st < 0 ? -1 : 0;
}

static int write_1d(hid_t loc, const char *name, long n, const double *data) {
    hsize_t dims[1] = {(hsize_t)n};
    hid_t sp = H5Screate_simple(1, dims, NULL);
    hid_t ds = H5Dcreate2(loc, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) { H5Sclose(sp); return -1; }
    herr_t st = H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(ds); H5Sclose(sp);
    return st < 0 ? -1 : 0;
}

static int write_attr(hid_t loc, const char *name, double v) {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t at = H5Acreate2(loc, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT);
    if (at < 0) { H5Sclose(sp); return -1; }
    herr_t st = H5Awrite(at, H5T_NATIVE_DOUBLE, &v);
    H5Aclose(at); H5Sclose(sp);
    return st < 0 ? -1 : 0;
}

/* Initial data.  one_d != 0: 1-D [nlayers] datasets (use_1d_id = yes layout,
 * ejecta_id.h:202-220, with zero "ye" and "s"); else [nshells x nlayers]. */
int h5io_write_id(const char *path, int one_d, int nshells, int nlayers,
                  const double *r, const double *mom, const double *theta,
                  const double *ctheta, const double *ek, const double *mass,
                  double theta_core, double theta_wing) {
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = 0;
    if (one_d) {
        double *z = (double *)calloc((size_t)nlayers, sizeof(double));
        rc |= write_1d(f, "r", nlayers, r);
        rc |= write_1d(f, "theta", nlayers, theta);
        rc |= write_1d(f, "ctheta", nlayers, ctheta);
        rc |= write_1d(f, "mom", nlayers, mom);
        rc |= write_1d(f, "mass", nlayers, mass);
        rc |= write_1d(f, "ek", nlayers, ek);
        rc |= write_1d(f, "ye", nlayers, z);
        rc |= write_1d(f, "s", nlayers, z);
        free(z);
    } else {
        rc |= write_2d(f, "r", nshells, nlayers, r);
        rc |= write_2d(f, "mom", nshells, nlayers, mom);
        rc |= write_2d(f, "theta", nshells, nlayers, theta);
        rc |= write_2d(f, "ctheta", nshells, nlayers, ctheta);
        rc |= write_2d(f, "ek", nshells, nlayers, ek);
        rc |= write_2d(f, "mass", nshells, nlayers, mass);
    }
    rc |= write_attr(f, "theta_core", theta_core);
    rc |= write_attr(f, "theta_wing", theta_wing);
    H5Fclose(f);
    return rc ? -2 : 0;
}

/* Number of elements of a dataset ("group/name" paths allowed), <0 if absent. */
long h5io_dataset_len(const char *path, const char *dset) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
    if (ds < 0) { H5Fclose(f); return -2; }
    hid_t sp = H5Dget_space(ds);
    hssize_t n = H5Sget_simple_extent_npoints(sp);
    H5Sclose(sp); H5Dclose(ds); H5Fclose(f);
    return (long)n;
}

/* Rank of dataset `dset` and its first `maxdims` extents (dims), or < 0 on error. */
int h5io_dataset_dims(const char *path, const char *dset, long *dims, int maxdims) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
    if (ds < 0) { H5Fclose(f); return -2; }
    hid_t sp = H5Dget_space(ds);
    hsize_t d[8];
    int rank = H5Sget_simple_extent_dims(sp, d, NULL);
    if (rank > 8) rank = -3;
    for (int i = 0; i < rank && i < maxdims; i++) dims[i] = (long)d[i];
    H5Sclose(sp); H5Dclose(ds); H5Fclose(f);
    return rank;
}

int h5io_read_dataset(const char *path, const char *dset, double *out, long n) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t ds = H5Dopen2(f, dset, H5P_DEFAULT);
    if (ds < 0) { H5Fclose(f); return -2; }
    hid_t sp = H5Dget_space(ds);
    hssize_t m = H5Sget_simple_extent_npoints(sp);
    int rc = 0;
    if ((long)m != n) rc = -3;
    else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) rc = -4;
    H5Sclose(sp); H5Dclose(ds); H5Fclose(f);
    return rc;
}

/* Scalar double attribute on an object ("/" for the file). */
int h5io_read_attr(const char *path, const char *obj, const char *name, double *out) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t at = H5Aopen_by_name(f, obj, name, H5P_DEFAULT, H5P_DEFAULT);
    if (at < 0) { H5Fclose(f); return -2; }
    int rc = H5Aread(at, H5T_NATIVE_DOUBLE, out) < 0 ? -3 : 0;
    H5Aclose(at); H5Fclose(f);
    return rc;
}

/* Light-curve file: flux is [nlayers x npairs] (shell 0). */
int h5io_write_lc(const char *path, int nshells, int nlayers, long npairs,
                  const double *flux, const double *times, const double *freqs,
                  int nattrs, const char *const *keys, const double *vals) {
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = 0;
    char name[96];
    for (int ish = 0; ish < nshells; ish++)
        for (int il = 0; il < nlayers; il++) {
            snprintf(name, sizeof name, "shell=%d layer=%d", ish, il);
            hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
            if (g < 0) { rc = -2; continue; }
            rc |= write_1d(g, "fluxdens", npairs, flux + ((long)ish * nlayers + il) * npairs);
            H5Gclose(g);
        }
    rc |= write_1d(f, "times", npairs, times);
    rc |= write_1d(f, "freqs", npairs, freqs);
    rc |= write_attr(f, "nshells", (double)nshells);
    rc |= write_attr(f, "nlayers", (double)nlayers);
    for (int i = 0; i < nattrs; i++)
        if (H5Aexists(f, keys[i]) <= 0) rc |= write_attr(f, keys[i], vals[i]);
    H5Fclose(f);
    return rc ? -3 : 0;
}

/* ---- incremental writer: open a file, add 1-D datasets and scalar attributes to groups
 * (created on first use; "" or "/" is the root group), close.  Used for the light-curve,
 * dynamics and spectrum files in the reference's layout (src/ejecta/ejecta.h:165-222, 354-595). */
static hid_t group_of(hid_t f, const char *group) {
    if (!group || !group[0] || (group[0] == '/' && !group[1])) return H5Gopen2(f, "/", H5P_DEFAULT);
    if (H5Lexists(f, group, H5P_DEFAULT) > 0) return H5Gopen2(f, group, H5P_DEFAULT);
    return H5Gcreate2(f, group, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
}

long h5io_create(const char *path) {
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    return (long)f;
}

int h5io_add_dataset(long file, const char *group, const char *name, long n, const double *data) {
    hid_t g = group_of((hid_t)file, group);
    if (g < 0) return -1;
    int rc = write_1d(g, name, n, data);
    H5Gclose(g);
    return rc;
}

/* [n0][n1] row-major dataset in a group (created on first use) */
int h5io_add_dataset2d(long file, const char *group, const char *name, int n0, int n1, const double *data) {
    hid_t g = group_of((hid_t)file, group);
    if (g < 0) return -1;
    int rc = write_2d(g, name, n0, n1, data);
    H5Gclose(g);
    return rc;
}

int h5io_add_attr(long file, const char *group, const char *name, double v) {
    hid_t g = group_of((hid_t)file, group);
    if (g < 0) return -1;
    int rc = 0;
    if (H5Aexists(g, name) <= 0) rc = write_attr(g, name, v);
    H5Gclose(g);
    return rc;
}

int h5io_close(long file) { return H5Fclose((hid_t)file) < 0 ? -1 : 0; }

/* number of attributes on an object (for layout checks) and the name of attribute i */
int h5io_attr_count(const char *path, const char *obj) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    H5O_info_t info;
    int n = -2;
    if (H5Oget_info_by_name(f, obj, &info, H5P_DEFAULT) >= 0) n = (int)info.num_attrs;
    H5Fclose(f);
    return n;
}

int h5io_attr_name(const char *path, const char *obj, int i, char *buf, int len) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t o = H5Oopen(f, obj, H5P_DEFAULT);
    int rc = -2;
    if (o >= 0) {
        hid_t a = H5Aopen_by_idx(o, ".", H5_INDEX_NAME, H5_ITER_INC, (hsize_t)i, H5P_DEFAULT, H5P_DEFAULT);
        if (a >= 0) { rc = H5Aget_name(a, (size_t)len, buf) >= 0 ? 0 : -3; H5Aclose(a); }
        H5Oclose(o);
    }
    H5Fclose(f);
    return rc;
}

/* number of links (datasets, groups) in a group and the name of link i (by name order) */
int h5io_link_count(const char *path, const char *group) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    hid_t g = H5Gopen2(f, group, H5P_DEFAULT);
    int n = -2;
    if (g >= 0) {
        H5G_info_t info;
        if (H5Gget_info(g, &info) >= 0) n = (int)info.nlinks;
        H5Gclose(g);
    }
    H5Fclose(f);
    return n;
}

int h5io_link_name(const char *path, const char *group, int i, char *buf, int len) {
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return -1;
    int rc = H5Lget_name_by_idx(f, group, H5_INDEX_NAME, H5_ITER_INC, (hsize_t)i, buf, (size_t)len, H5P_DEFAULT) >= 0
                 ? 0 : -2;
    H5Fclose(f);
    return rc;
}
