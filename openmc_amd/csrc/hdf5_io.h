// Incident-neutron nuclear data in the reference's HDF5 layout (layout and
// citations in hdf5_io.cpp).
#pragma once

#include <cstddef>
#include <string>

#include "host_model.h"

namespace omc {

// Whether a library path names an HDF5 file (.h5 / .hdf5).
bool is_hdf5_path(const std::string& path);

// Reads nuclide `name` from an HDF5 file at the temperature nearest to
// `temperature` [K] (an error beyond `tolerance` K): energy grid, kT, awr,
// and the derived total, absorption and elastic columns.
void read_nuclide_hdf5(HostNuclide& nuc, const std::string& path, const std::string& name, double temperature,
  double tolerance);

// Writes a nuclide with elastic scattering (MT 2) and capture (MT 102) at
// one temperature in the same layout.
void write_nuclide_hdf5(const std::string& path, const std::string& name, double awr, double kT, int temperature,
  const double* energy, const double* elastic, const double* capture, size_t n);

} // namespace omc
