// Incident-neutron nuclear data in the reference's HDF5 layout (layout and
// citations in hdf5_io.cpp).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "host_model.h"

namespace omc {

// Whether libhdf5 can be loaded.
bool hdf5_available();

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

// Writes an HDF5 file with the conventions of the reference's
// hdf5_interface.cpp, for the statepoint.  Groups stay open until the
// writer is destroyed.
class H5Writer {
public:
  enum Kind { BOOL, INT, INT64, SIZE, DOUBLE };
  explicit H5Writer(const std::string& path);
  ~H5Writer();
  H5Writer(const H5Writer&) = delete;
  H5Writer& operator=(const H5Writer&) = delete;
  int64_t root() const { return file_; }
  int64_t group(int64_t parent, const std::string& name);
  // scalar (array = false) or one-dimensional attribute of n values
  void attr(int64_t obj, const std::string& name, Kind k, const void* v, size_t n = 1, bool array = false);
  // dataset of the given shape (empty: scalar)
  void data(int64_t obj, const std::string& name, Kind k, const void* v, const std::vector<size_t>& shape = {});
  void attr_str(int64_t obj, const std::string& name, const std::string& s);
  void data_str(int64_t obj, const std::string& name, const std::string& s);
  void data_strs(int64_t obj, const std::string& name, const std::vector<std::string>& v);
  // the source-bank dataset of n_sites sites, created with the reference's
  // site type and not written
  void source_bank(int64_t obj, const std::string& name, int64_t n_sites);

private:
  int64_t file_ = -1;
  std::vector<int64_t> groups_;
};

// Reads a dataset (attribute empty) or an attribute of `object` ("" or "."
// for the file) into buf: kind 0 as doubles, 1 as int64, 2 as the raw bytes
// of its fixed-length strings.  Returns the number of values (bytes for
// kind 2); reads only when buf holds that many (cap).
size_t h5_read(const std::string& path, const std::string& object, const std::string& attribute, int kind,
  void* buf, size_t cap);

} // namespace omc
