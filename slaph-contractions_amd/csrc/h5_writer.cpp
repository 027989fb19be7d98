// HDF5 correlator files (declared in include/slaph_h5.h): the file layout of
// the reference's WriteHDF5Correlator (include/h5-wrapper.hpp:106-200) written
// through the HDF5 C API.
#include "../../include/slaph_h5.h"

#include <hdf5.h>

#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>

struct slh5_file {
  hid_t file = -1;
  hid_t ctype = -1;  // {re: f64 @0, im: f64 @8}
};

namespace {

thread_local std::string g_err;

int fail(std::string const &msg) {
  g_err = msg;
  return -1;
}

// create_folder_for_hdf5_file (h5-wrapper.hpp:160-170): every missing
// directory of the path, like boost::filesystem::create_directories.
bool make_dirs(std::string const &path) {
  if (path.empty()) return true;
  std::string cur;
  size_t at = 0;
  while (at <= path.size()) {
    size_t next = path.find('/', at);
    if (next == std::string::npos) next = path.size();
    cur = path.substr(0, next);
    if (!cur.empty()) {
      struct stat st;
      if (stat(cur.c_str(), &st) != 0) {
        if (mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST) return false;
      } else if (!S_ISDIR(st.st_mode)) {
        return false;
      }
    }
    at = next + 1;
  }
  return true;
}

}  // namespace

extern "C" {

int slh5_open(const char *output_path, const char *diagram, const char *ending, slh5_file **out) {
  if (!output_path || !diagram || !ending || !out) return fail("slh5_open: null argument");
  *out = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors come back as codes
  std::string dir(output_path);
  if (!make_dirs(dir)) return fail("cannot create directory " + dir);
  std::string name = dir + "/" + diagram + ending;
  slh5_file *f = new slh5_file;
  struct stat st;
  if (stat(name.c_str(), &st) == 0)
    f->file = H5Fopen(name.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
  else
    f->file = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f->file < 0) {
    delete f;
    return fail("cannot open or create " + name);
  }
  // make_comp_type<Complex> (src/h5-wrapper.cpp:18-26)
  f->ctype = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
  if (f->ctype < 0 || H5Tinsert(f->ctype, "re", 0, H5T_NATIVE_DOUBLE) < 0 ||
      H5Tinsert(f->ctype, "im", sizeof(double), H5T_NATIVE_DOUBLE) < 0) {
    slh5_close(f);
    return fail("cannot build the complex compound type");
  }
  *out = f;
  return 0;
}

int slh5_write(slh5_file *f, const char *dataset, const double *values, int n) {
  if (!f || !dataset || (n > 0 && !values) || n < 0) return fail("slh5_write: bad argument");
  const htri_t exists = H5Lexists(f->file, dataset, H5P_DEFAULT);
  if (exists > 0) {
    std::printf("Not writing %s because it exists.\n", dataset);
    return 1;
  }
  hsize_t dim = hsize_t(n);
  hid_t space = H5Screate_simple(1, &dim, nullptr);
  if (space < 0) return fail(std::string("cannot create the dataspace of ") + dataset);
  hid_t dset = H5Dcreate2(f->file, dataset, f->ctype, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  H5Sclose(space);
  if (dset < 0) return fail(std::string("cannot create dataset ") + dataset);
  const herr_t w = n > 0 ? H5Dwrite(dset, f->ctype, H5S_ALL, H5S_ALL, H5P_DEFAULT, values) : 0;
  H5Dclose(dset);
  if (w < 0) return fail(std::string("cannot write dataset ") + dataset);
  return 0;
}

void slh5_close(slh5_file *f) {
  if (!f) return;
  if (f->ctype >= 0) H5Tclose(f->ctype);
  if (f->file >= 0) H5Fclose(f->file);
  delete f;
}

int slh5_filename_ending(int config, char *buf, int len) {
  char tmp[64];
  const int n = std::snprintf(tmp, sizeof tmp, "_cnfg%04d.h5", config);
  if (buf && len > 0) {
    std::strncpy(buf, tmp, size_t(len) - 1);
    buf[len - 1] = 0;
  }
  return n + 1;
}

const char *slh5_last_error(void) { return g_err.c_str(); }

}  // extern "C"
