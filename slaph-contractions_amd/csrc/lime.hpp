// Gauge configurations in lime / ILDG files (reference:
// GaugeField::read_gauge_field and read_lime_gauge_field_doubleprec_timeslices,
// src/GaugeField.cpp:762-928, and map_timeslice_to_eigen, :111-149).
#pragma once

#include <complex>
#include <functional>
#include <string>

namespace slaph {

// Reads the first "ildg-binary-data" record of the lime file `path` (an FP64
// su3gauge field on Lx x Ly x Lz x Lt) one timeslice at a time and hands the
// spatial links of each t in [slice_i, slice_f] to sink(t, links), in the
// library's layout: links[(site * 3 + mu) * 9 + c * 3 + r] = element (r, c) of
// U_mu at (x, y, z, t), site = (x Ly + y) Lz + z, mu = x, y, z (the file's
// U_t is not used).  For Lx = Ly = Lz this is the reference's tslices[t]
// after its reader and timeslice map.
//
// Errors, as runtime_error: an unreadable file, no binary record, a record of
// the wrong size ("Single precision read!" for a 32-bit one, as the
// reference), an "ildg-format" record whose lattice or precision differs
// from the arguments, and a "scidac-checksum" record that does not match the
// binary data (checked over the whole record, every timeslice read).
void read_ildg_gauge(std::string const &path, int Lx, int Ly, int Lz, int Lt, int slice_i, int slice_f,
                     std::function<void(int, const std::complex<double> *)> const &sink);

}  // namespace slaph
