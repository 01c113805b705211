/* Run-file writer/reader on the HDF5 1.10 C library (host code, C99).
 *
 * Layout of a run file: five chunked datasets with an unlimited record axis,
 * as RunData.create_file / save_data write them (main/data_management.py:
 * 13-52).  Chunks hold one record each along axis 0; a state record is split
 * along its leading axes into chunks of at most kChunkBytes, so a record is
 * written and read as whole chunks (no read-modify-write of partial chunks,
 * and the chunk cache is bypassed). */
#include <hdf5.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "vlasov1d2v_io.h"

enum { VL_OK_ = 0, VL_E_ARG_ = -1, VL_E_NULL_ = -2 };

static __thread char g_err[512];
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static const hsize_t kChunkBytes = 16u << 20;   /* state chunks: <= 16 MiB */
static const hsize_t kSeriesChunk = 1024;       /* time / energy chunks (values) */

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *vl_io_last_error_string(void) { return g_err; }

static void quiet_errors(void) {
    static int done = 0;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, NULL, NULL);   /* errors are returned, not printed */
        done = 1;
    }
}

/* chunk = dims with axis 0 set to 1, halving the leading axes until the chunk
 * is at most kChunkBytes */
static void chunk_dims(int rank, const hsize_t *dims, size_t esize, hsize_t *chunk) {
    hsize_t bytes = esize;
    chunk[0] = 1;
    for (int d = 1; d < rank; ++d) {
        chunk[d] = dims[d] > 0 ? dims[d] : 1;
        bytes *= chunk[d];
    }
    for (int d = 1; d < rank && bytes > kChunkBytes; ++d)
        while (chunk[d] > 1 && bytes > kChunkBytes) {
            bytes /= chunk[d];
            chunk[d] = (chunk[d] + 1) / 2;
            bytes *= chunk[d];
        }
}

static int make_dataset(hid_t file, const char *name, int rank, const hsize_t *dims, const hsize_t *maxdims,
                        const hsize_t *chunk, hid_t ftype, hid_t mtype, const void *data) {
    int rc = VL_E_IO;
    hid_t space = -1, dcpl = -1, ds = -1;
    space = H5Screate_simple(rank, dims, maxdims);
    if (space < 0) { fail(VL_E_IO, "H5Screate_simple failed for '%s'", name); goto done; }
    dcpl = H5Pcreate(H5P_DATASET_CREATE);
    if (dcpl < 0 || H5Pset_chunk(dcpl, rank, chunk) < 0) { fail(VL_E_IO, "chunk layout of '%s'", name); goto done; }
    ds = H5Dcreate2(file, name, ftype, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
    if (ds < 0) { fail(VL_E_IO, "H5Dcreate2 failed for '%s'", name); goto done; }
    if (H5Dwrite(ds, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) {
        fail(VL_E_IO, "H5Dwrite failed for '%s'", name);
        goto done;
    }
    rc = VL_OK_;
done:
    if (ds >= 0) H5Dclose(ds);
    if (dcpl >= 0) H5Pclose(dcpl);
    if (space >= 0) H5Sclose(space);
    return rc;
}

int vl_run_data_create(const char *path, const int64_t shape[6], const float *pdf, int64_t rows, int64_t cols,
                       const float *potential, const float *density) {
    if (!path || !shape || !pdf || !potential || !density) return fail(VL_E_NULL_, "vl_run_data_create: null pointer");
    for (int i = 0; i < 6; ++i)
        if (shape[i] < 1) return fail(VL_E_ARG_, "vl_run_data_create: shape[%d] = %lld", i, (long long)shape[i]);
    if (rows < 1 || cols < 1 || rows > shape[0] || cols > shape[1])
        return fail(VL_E_ARG_, "vl_run_data_create: field shape (%lld, %lld) does not fit maxshape (%lld, %lld)",
                    (long long)rows, (long long)cols, (long long)shape[0], (long long)shape[1]);
    pthread_mutex_lock(&g_mu);
    quiet_errors();
    int rc = VL_E_IO;
    hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) {
        fail(VL_E_IO, "cannot create '%s'", path);
        pthread_mutex_unlock(&g_mu);
        return VL_E_IO;
    }
    {
        hsize_t d[7], m[7], c[7];
        d[0] = 1;
        m[0] = H5S_UNLIMITED;
        for (int i = 0; i < 6; ++i) d[i + 1] = m[i + 1] = (hsize_t)shape[i];
        chunk_dims(7, d, sizeof(float), c);
        if (make_dataset(file, "pdf", 7, d, m, c, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, pdf)) goto done;
    }
    {
        const hsize_t d[3] = {1, (hsize_t)rows, (hsize_t)cols};
        const hsize_t m[3] = {H5S_UNLIMITED, (hsize_t)shape[0], (hsize_t)shape[1]};
        hsize_t c[3];
        chunk_dims(3, d, sizeof(float), c);
        if (make_dataset(file, "potential", 3, d, m, c, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, potential)) goto done;
        if (make_dataset(file, "density", 3, d, m, c, H5T_IEEE_F32LE, H5T_NATIVE_FLOAT, density)) goto done;
    }
    {
        const hsize_t d[1] = {1}, m[1] = {H5S_UNLIMITED}, c[1] = {kSeriesChunk};
        const double zero = 0.0;
        if (make_dataset(file, "time", 1, d, m, c, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &zero)) goto done;
        if (make_dataset(file, "energy", 1, d, m, c, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, &zero)) goto done;
    }
    rc = VL_OK_;
done:
    if (H5Fclose(file) < 0 && rc == VL_OK_) rc = fail(VL_E_IO, "closing '%s' failed", path);
    pthread_mutex_unlock(&g_mu);
    return rc;
}

/* grows dataset `name` by one record and writes `data` into it */
/* VL_OK_ when a record of dataset `name` holds `elems` values */
static int check_record(hid_t file, const char *name, int64_t elems) {
    int rc = VL_E_IO;
    hsize_t dims[8];
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT), sp = -1;
    if (ds < 0) { fail(VL_E_IO, "no dataset '%s'", name); goto done; }
    sp = H5Dget_space(ds);
    const int rank = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    if (rank < 1 || rank > 8 || H5Sget_simple_extent_dims(sp, dims, NULL) < 0) {
        fail(VL_E_IO, "bad dataspace of '%s'", name);
        goto done;
    }
    int64_t rec = 1;
    for (int i = 1; i < rank; ++i) rec *= (int64_t)dims[i];
    if (rec != elems) {
        rc = fail(VL_E_ARG_, "a record of '%s' holds %lld values, the buffer %lld", name, (long long)rec,
                  (long long)elems);
        goto done;
    }
    rc = VL_OK_;
done:
    if (sp >= 0) H5Sclose(sp);
    if (ds >= 0) H5Dclose(ds);
    return rc;
}

/* elems: values in the caller's buffer, checked against the record shape */
static int append_record(hid_t file, const char *name, hid_t mtype, const void *data, int64_t elems,
                         hsize_t expect_len) {
    int rc = VL_E_IO;
    hid_t ds = -1, fspace = -1, mspace = -1;
    hsize_t dims[8], start[8], count[8];
    ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) { fail(VL_E_IO, "no dataset '%s'", name); goto done; }
    fspace = H5Dget_space(ds);
    const int rank = fspace < 0 ? -1 : H5Sget_simple_extent_ndims(fspace);
    if (rank < 1 || rank > 8 || H5Sget_simple_extent_dims(fspace, dims, NULL) < 0) {
        fail(VL_E_IO, "bad dataspace of '%s'", name);
        goto done;
    }
    if (dims[0] != expect_len) {
        rc = fail(VL_E_ARG_, "dataset '%s' has %llu records, 'pdf' has %llu", name, (unsigned long long)dims[0],
                  (unsigned long long)expect_len);
        goto done;
    }
    {
        int64_t rec = 1;
        for (int i = 1; i < rank; ++i) rec *= (int64_t)dims[i];
        if (rec != elems) {
            rc = fail(VL_E_ARG_, "a record of '%s' holds %lld values, the buffer %lld", name, (long long)rec,
                      (long long)elems);
            goto done;
        }
    }
    H5Sclose(fspace);
    fspace = -1;
    dims[0] += 1;
    if (H5Dset_extent(ds, dims) < 0) { fail(VL_E_IO, "cannot extend '%s'", name); goto done; }
    fspace = H5Dget_space(ds);
    for (int i = 0; i < rank; ++i) {
        start[i] = 0;
        count[i] = dims[i];
    }
    start[0] = dims[0] - 1;
    count[0] = 1;
    if (fspace < 0 || H5Sselect_hyperslab(fspace, H5S_SELECT_SET, start, NULL, count, NULL) < 0) {
        fail(VL_E_IO, "cannot select the new record of '%s'", name);
        goto done;
    }
    mspace = H5Screate_simple(rank, count, NULL);
    if (mspace < 0 || H5Dwrite(ds, mtype, mspace, fspace, H5P_DEFAULT, data) < 0) {
        fail(VL_E_IO, "H5Dwrite failed for '%s'", name);
        goto done;
    }
    rc = VL_OK_;
done:
    if (mspace >= 0) H5Sclose(mspace);
    if (fspace >= 0) H5Sclose(fspace);
    if (ds >= 0) H5Dclose(ds);
    return rc;
}

static int record_count(hid_t file, const char *name, hsize_t *len) {
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return fail(VL_E_IO, "no dataset '%s'", name);
    hid_t sp = H5Dget_space(ds);
    hsize_t dims[8];
    const int rank = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    int rc = VL_OK_;
    if (rank < 1 || rank > 8 || H5Sget_simple_extent_dims(sp, dims, NULL) < 0)
        rc = fail(VL_E_IO, "bad dataspace of '%s'", name);
    else
        *len = dims[0];
    if (sp >= 0) H5Sclose(sp);
    H5Dclose(ds);
    return rc;
}

int vl_run_data_append(const char *path, const float *pdf, int64_t pdf_elems, const float *potential,
                       int64_t potential_elems, const float *density, int64_t density_elems, double time,
                       double energy) {
    if (!path || !pdf || !potential || !density) return fail(VL_E_NULL_, "vl_run_data_append: null pointer");
    pthread_mutex_lock(&g_mu);
    quiet_errors();
    hid_t file = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
    if (file < 0) {
        fail(VL_E_IO, "cannot open '%s' for appending", path);
        pthread_mutex_unlock(&g_mu);
        return VL_E_IO;
    }
    hsize_t len = 0;
    int rc = record_count(file, "pdf", &len);
    /* every size first, so that a mismatch leaves the file unchanged */
    if (!rc) rc = check_record(file, "pdf", pdf_elems);
    if (!rc) rc = check_record(file, "potential", potential_elems);
    if (!rc) rc = check_record(file, "density", density_elems);
    if (!rc) rc = append_record(file, "pdf", H5T_NATIVE_FLOAT, pdf, pdf_elems, len);
    if (!rc) rc = append_record(file, "potential", H5T_NATIVE_FLOAT, potential, potential_elems, len);
    if (!rc) rc = append_record(file, "density", H5T_NATIVE_FLOAT, density, density_elems, len);
    if (!rc) rc = append_record(file, "time", H5T_NATIVE_DOUBLE, &time, 1, len);
    if (!rc) rc = append_record(file, "energy", H5T_NATIVE_DOUBLE, &energy, 1, len);
    if (H5Fclose(file) < 0 && !rc) rc = fail(VL_E_IO, "closing '%s' failed", path);
    pthread_mutex_unlock(&g_mu);
    return rc;
}

int vl_read_data_dims(const char *path, const char *name, int *rank, int64_t dims[8], int64_t maxdims[8]) {
    if (!path || !name || !rank || !dims || !maxdims) return fail(VL_E_NULL_, "vl_read_data_dims: null pointer");
    pthread_mutex_lock(&g_mu);
    quiet_errors();
    int rc = VL_E_IO;
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), ds = -1, sp = -1;
    if (file < 0) { fail(VL_E_IO, "cannot open '%s'", path); goto done; }
    ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) { fail(VL_E_IO, "no dataset '%s' in '%s'", name, path); goto done; }
    sp = H5Dget_space(ds);
    hsize_t d[8], m[8];
    const int r = sp < 0 ? -1 : H5Sget_simple_extent_ndims(sp);
    if (r < 1 || r > 8 || H5Sget_simple_extent_dims(sp, d, m) < 0) {
        fail(VL_E_IO, "bad dataspace of '%s'", name);
        goto done;
    }
    *rank = r;
    for (int i = 0; i < r; ++i) {
        dims[i] = (int64_t)d[i];
        maxdims[i] = m[i] == H5S_UNLIMITED ? -1 : (int64_t)m[i];
    }
    rc = VL_OK_;
done:
    if (sp >= 0) H5Sclose(sp);
    if (ds >= 0) H5Dclose(ds);
    if (file >= 0) H5Fclose(file);
    pthread_mutex_unlock(&g_mu);
    return rc;
}

int vl_read_data(const char *path, const char *name, int64_t start, int64_t count, int as_double, void *out) {
    if (!path || !name || (!out && count > 0)) return fail(VL_E_NULL_, "vl_read_data: null pointer");
    pthread_mutex_lock(&g_mu);
    quiet_errors();
    int rc = VL_E_IO;
    hid_t file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), ds = -1, fsp = -1, msp = -1;
    hsize_t dims[8], st[8], cn[8];
    if (file < 0) { fail(VL_E_IO, "cannot open '%s'", path); goto done; }
    ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) { fail(VL_E_IO, "no dataset '%s' in '%s'", name, path); goto done; }
    fsp = H5Dget_space(ds);
    const int rank = fsp < 0 ? -1 : H5Sget_simple_extent_ndims(fsp);
    if (rank < 1 || rank > 8 || H5Sget_simple_extent_dims(fsp, dims, NULL) < 0) {
        fail(VL_E_IO, "bad dataspace of '%s'", name);
        goto done;
    }
    if (start < 0 || count < 0 || (hsize_t)(start + count) > dims[0]) {
        rc = fail(VL_E_ARG_, "vl_read_data: records [%lld, %lld) outside 0..%llu of '%s'", (long long)start,
                  (long long)(start + count), (unsigned long long)dims[0], name);
        goto done;
    }
    if (count == 0) {
        rc = VL_OK_;
        goto done;
    }
    for (int i = 0; i < rank; ++i) {
        st[i] = 0;
        cn[i] = dims[i];
    }
    st[0] = (hsize_t)start;
    cn[0] = (hsize_t)count;
    if (H5Sselect_hyperslab(fsp, H5S_SELECT_SET, st, NULL, cn, NULL) < 0) {
        fail(VL_E_IO, "cannot select records of '%s'", name);
        goto done;
    }
    msp = H5Screate_simple(rank, cn, NULL);
    if (msp < 0 || H5Dread(ds, as_double ? H5T_NATIVE_DOUBLE : H5T_NATIVE_FLOAT, msp, fsp, H5P_DEFAULT, out) < 0) {
        fail(VL_E_IO, "H5Dread failed for '%s'", name);
        goto done;
    }
    rc = VL_OK_;
done:
    if (msp >= 0) H5Sclose(msp);
    if (fsp >= 0) H5Sclose(fsp);
    if (ds >= 0) H5Dclose(ds);
    if (file >= 0) H5Fclose(file);
    pthread_mutex_unlock(&g_mu);
    return rc;
}
