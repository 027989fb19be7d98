/*
 * slaph_h5.hpp -- WriteHDF5Correlator of the reference (include/h5-wrapper.hpp:
 * 106-200) over the C ABI of slaph_h5.h (library libslaph_h5.so):
 *
 *   WriteHDF5Correlator file(output_path, "C2c", "_cnfg1000.h5");
 *   file.write(values, dataset_name);   // keeps an existing dataset
 *
 * The reference's comp_type argument is fixed to make_comp_type<Complex>
 * ({re, im} doubles), the only type its callers pass.
 */
#ifndef SLAPH_H5_HPP
#define SLAPH_H5_HPP

#include "slaph_h5.h"

#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

namespace slaph_amd {

class WriteHDF5Correlator {
 public:
  WriteHDF5Correlator(std::string const &output_path, std::string const &diagram, std::string const &output_filename) {
    if (slh5_open(output_path.c_str(), diagram.c_str(), output_filename.c_str(), &f_) != 0)
      throw std::runtime_error(slh5_last_error());
  }
  ~WriteHDF5Correlator() { slh5_close(f_); }
  WriteHDF5Correlator(WriteHDF5Correlator const &) = delete;
  WriteHDF5Correlator &operator=(WriteHDF5Correlator const &) = delete;

  void write(std::vector<std::complex<double>> const &corr, std::string const &dataset_name) {
    if (slh5_write(f_, dataset_name.c_str(), reinterpret_cast<const double *>(corr.data()), int(corr.size())) < 0)
      throw std::runtime_error(slh5_last_error());
  }

 private:
  slh5_file *f_ = nullptr;
};

// create_correlator_file_name (src/global_data_build_IO_names.cpp:66-68)
inline std::string correlator_filename_ending(int config) {
  char buf[64];
  slh5_filename_ending(config, buf, sizeof buf);
  return buf;
}

}  // namespace slaph_amd

#endif  // SLAPH_H5_HPP
