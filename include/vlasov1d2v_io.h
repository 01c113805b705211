/*
 * vlasov1d2v_io.h -- C ABI of the run-file writer/reader: HDF5 files with the
 * datasets of the reference's RunData / ReadData (main/data_management.py:5-71).
 *
 * A run file holds five datasets whose first axis counts records (write-outs)
 * and is unlimited:
 *   "pdf"       float32 (T, NX+2, n, NU+2, n, NV+2, n)  the ghosted state
 *   "potential" float32 (T, NX, n), maxshape (inf, NX+2, n)
 *   "density"   float32 (T, NX, n), maxshape (inf, NX+2, n)
 *   "time"      float64 (T,)
 *   "energy"    float64 (T,)
 * exactly as RunData.create_file lays them out (main/data_management.py:13-35),
 * except that "time" and "energy" are float64: the reference creates them from
 * the Python list [0], which makes them int64 and truncates every time and
 * energy written later.  The reader converts whatever type a file holds (so
 * files written by the reference read back too).
 *
 * Host pointers only (the device side is vl_export_f32 / vl_import_f32 in
 * vlasov1d2v.h).  Every function returns int: 0 = ok, negative = error
 * (VL_E_ARG / VL_E_NULL as in vlasov1d2v.h, VL_E_IO for a failed HDF5 call);
 * vl_io_last_error_string() describes the most recent failure on the calling
 * thread.  Calls are serialised inside the library, so a writer thread may
 * append while another thread reads a different file.
 *
 * Built from vlasov1d2v_amd/csrc/vl_io.c against the HDF5 1.10 C library;
 * the stepping library does not depend on it.
 */
#ifndef VLASOV1D2V_IO_H
#define VLASOV1D2V_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VL_E_IO (-7) /* an HDF5 call failed (missing file, bad dataset, disk) */

const char *vl_io_last_error_string(void);

/* Creates (truncates) the run file with one record: the state pdf (shape
 * `shape`, C order), potential and density ((rows, cols) each, rows <=
 * shape[0], cols <= shape[1]), time 0 and energy 0.
 * Replaces RunData.create_file (main/data_management.py:13-35). */
int vl_run_data_create(const char *path, const int64_t shape[6], const float *pdf, int64_t rows, int64_t cols,
                       const float *potential, const float *density);

/* Appends one record to every dataset (each grows by one along axis 0; the
 * five must have the same length).  The buffers have the record shapes the
 * file was created with; *_elems are their value counts, and a count that
 * differs from the record size is VL_E_ARG (nothing is written).
 * Replaces RunData.save_data (main/data_management.py:37-52). */
int vl_run_data_append(const char *path, const float *pdf, int64_t pdf_elems, const float *potential,
                       int64_t potential_elems, const float *density, int64_t density_elems, double time,
                       double energy);

/* Rank, current and maximum dims of dataset `name` (maxdims -1 = unlimited). */
int vl_read_data_dims(const char *path, const char *name, int *rank, int64_t dims[8], int64_t maxdims[8]);

/* Reads records [start, start + count) of dataset `name` into out (C order,
 * count x the record shape), converted to float64 (as_double != 0) or
 * float32.  Replaces ReadData.read_data / read_specific_time /
 * read_potential_only (main/data_management.py:60-84), which read whole
 * datasets, one record or a record range. */
int vl_read_data(const char *path, const char *name, int64_t start, int64_t count, int as_double, void *out);

#ifdef __cplusplus
}
#endif
#endif
