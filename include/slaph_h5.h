/*
 * slaph_h5.h -- HDF5 correlator files of the MI355X sLapH contraction path.
 *
 * The layout of the reference's WriteHDF5Correlator (include/h5-wrapper.hpp:
 * 106-200, src/h5-wrapper.cpp:18-26): one file per diagram,
 * <output_path>/<diagram><filename_ending> with filename_ending
 * "_cnfg%04d.h5" (src/global_data_build_IO_names.cpp:66-68), one 1-D dataset
 * of Lt elements per correlator, element type the compound
 * {re: f64 at offset 0, im: f64 at offset 8}.  An existing file is opened
 * read-write and its datasets are kept: a dataset that already exists is not
 * written again (h5-wrapper.hpp:145-150).
 *
 * A library of its own (libslaph_h5.so) so that libslaph_amd.so does not
 * depend on libhdf5.  Every function returns 0 on success and a negative
 * code on failure (slh5_last_error() then holds the message).
 */
#ifndef SLAPH_H5_H
#define SLAPH_H5_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slh5_file slh5_file;

/* WriteHDF5Correlator(output_path, diagram, filename_ending, comp_type):
 * creates output_path (with parents) when missing, then opens
 * <output_path>/<diagram><filename_ending> read-write, or creates it. */
int slh5_open(const char *output_path, const char *diagram, const char *filename_ending, slh5_file **out);

/* WriteHDF5Correlator::write(corr, dataset_name): n complex values (re, im
 * interleaved).  Returns 1 without writing when the dataset exists. */
int slh5_write(slh5_file *f, const char *dataset, const double *values, int n);

void slh5_close(slh5_file *f);

/* "_cnfg%04d.h5" of configuration `config` (create_correlator_file_name,
 * src/global_data_build_IO_names.cpp:66-68) into buf; returns the bytes
 * needed including the terminating NUL. */
int slh5_filename_ending(int config, char *buf, int len);

const char *slh5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SLAPH_H5_H */
