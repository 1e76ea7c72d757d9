// Incident-neutron nuclear data in the reference's HDF5 layout.
//
// The layout is the one openmc.data.IncidentNeutron.export_to_hdf5 writes and
// Nuclide::Nuclide / Reaction::Reaction read (openmc/data/neutron.py:374-413,
// openmc/data/reaction.py:910-936, src/nuclide.cpp:52-300,
// src/reaction.cpp:25-60):
//
//   /            attrs filetype = "data_neutron", version = [3, 0]
//   /<name>      attrs Z, A, metastable, atomic_weight_ratio
//     kTs/<T>K                  scalar kT [eV]
//     energy/<T>K               energy grid [eV]
//     reactions/reaction_<mt>   attrs mt, label, Q_value, center_of_mass, redundant
//       <T>K/xs                 cross section, attr threshold_idx (0-based
//                               index of its first energy point)
//
// The same file writes the statepoint's HDF5 objects (H5Writer) with the
// conventions of src/hdf5_interface.cpp.
//
// libhdf5 is opened at run time (dlopen), so the library loads and runs on
// .nuc data where no HDF5 is installed; only reading or writing an .h5 file
// needs it.  Its location: $OPENMC_HDF5_LIBRARY, else libhdf5.so on the
// loader's path, else /opt/conda/lib/libhdf5.so.
#include "hdf5_io.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace omc {

namespace {

using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t H5P_DEFAULT_ = 0;
constexpr unsigned H5F_ACC_RDONLY_ = 0x0000u;
constexpr unsigned H5F_ACC_TRUNC_ = 0x0002u;
constexpr int H5_INDEX_NAME_ = 0;
constexpr int H5_ITER_INC_ = 0;

// The few entry points of the HDF5 C API (1.10) this file uses.
struct H5 {
  void* so = nullptr;
  herr_t (*open)() = nullptr;
  hid_t (*fopen)(const char*, unsigned, hid_t) = nullptr;
  hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
  herr_t (*fclose)(hid_t) = nullptr;
  hid_t (*gopen)(hid_t, const char*, hid_t) = nullptr;
  hid_t (*gcreate)(hid_t, const char*, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*gclose)(hid_t) = nullptr;
  hid_t (*dopen)(hid_t, const char*, hid_t) = nullptr;
  hid_t (*dcreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*) = nullptr;
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
  herr_t (*dclose)(hid_t) = nullptr;
  hid_t (*dget_space)(hid_t) = nullptr;
  int64_t (*sget_npoints)(hid_t) = nullptr;
  hid_t (*screate)(int) = nullptr;
  hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
  herr_t (*sclose)(hid_t) = nullptr;
  hid_t (*aopen)(hid_t, const char*, hid_t) = nullptr;
  herr_t (*aread)(hid_t, hid_t, void*) = nullptr;
  hid_t (*acreate)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t) = nullptr;
  herr_t (*awrite)(hid_t, hid_t, const void*) = nullptr;
  herr_t (*aclose)(hid_t) = nullptr;
  htri_t (*aexists)(hid_t, const char*) = nullptr;
  htri_t (*lexists)(hid_t, const char*, hid_t) = nullptr;
  herr_t (*gget_num)(hid_t, hsize_t*) = nullptr; // H5Gget_num_objs (1.8 API, kept in 1.10)
  int64_t (*lget_name_by_idx)(hid_t, const char*, int, int, hsize_t, char*, size_t, hid_t) = nullptr;
  hid_t (*tcopy)(hid_t) = nullptr;
  herr_t (*tset_size)(hid_t, size_t) = nullptr;
  herr_t (*tclose)(hid_t) = nullptr;
  herr_t (*eset_auto)(hid_t, void*, void*) = nullptr;
  hid_t (*tcreate)(int, size_t) = nullptr;
  herr_t (*tinsert)(hid_t, const char*, size_t, hid_t) = nullptr;
  hid_t (*dget_type)(hid_t) = nullptr;
  size_t (*tget_size)(hid_t) = nullptr;
  int (*tget_class)(hid_t) = nullptr;
  hid_t (*aopen_by_name)(hid_t, const char*, const char*, hid_t, hid_t) = nullptr;
  hid_t (*aget_space)(hid_t) = nullptr;
  hid_t (*aget_type)(hid_t) = nullptr;
  hid_t t_double = -1, t_int = -1, t_int64 = -1, t_str = -1;
  hid_t t_int8 = -1, t_ulong = -1, t_i64 = -1;
};

template<class F>
void bind(void* so, F& f, const char* name)
{
  f = reinterpret_cast<F>(dlsym(so, name));
  if (!f)
    throw std::runtime_error(std::string("libhdf5 has no symbol ") + name);
}

hid_t global_type(void* so, const char* name)
{
  void* p = dlsym(so, name);
  if (!p)
    throw std::runtime_error(std::string("libhdf5 has no symbol ") + name);
  return *static_cast<hid_t*>(p);
}

H5& h5()
{
  static H5 h;
  if (h.so)
    return h;
  std::vector<std::string> names;
  if (const char* env = std::getenv("OPENMC_HDF5_LIBRARY"))
    names.push_back(env);
  names.push_back("libhdf5.so");
  names.push_back("/opt/conda/lib/libhdf5.so");
  void* so = nullptr;
  for (const auto& n : names) {
    so = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (so)
      break;
  }
  if (!so)
    throw std::runtime_error("HDF5 nuclear data need libhdf5 (set OPENMC_HDF5_LIBRARY to its path)");
  H5 t;
  bind(so, t.open, "H5open");
  bind(so, t.fopen, "H5Fopen");
  bind(so, t.fcreate, "H5Fcreate");
  bind(so, t.fclose, "H5Fclose");
  bind(so, t.gopen, "H5Gopen2");
  bind(so, t.gcreate, "H5Gcreate2");
  bind(so, t.gclose, "H5Gclose");
  bind(so, t.dopen, "H5Dopen2");
  bind(so, t.dcreate, "H5Dcreate2");
  bind(so, t.dread, "H5Dread");
  bind(so, t.dwrite, "H5Dwrite");
  bind(so, t.dclose, "H5Dclose");
  bind(so, t.dget_space, "H5Dget_space");
  bind(so, t.sget_npoints, "H5Sget_simple_extent_npoints");
  bind(so, t.screate, "H5Screate");
  bind(so, t.screate_simple, "H5Screate_simple");
  bind(so, t.sclose, "H5Sclose");
  bind(so, t.aopen, "H5Aopen");
  bind(so, t.aread, "H5Aread");
  bind(so, t.acreate, "H5Acreate2");
  bind(so, t.awrite, "H5Awrite");
  bind(so, t.aclose, "H5Aclose");
  bind(so, t.aexists, "H5Aexists");
  bind(so, t.lexists, "H5Lexists");
  bind(so, t.gget_num, "H5Gget_num_objs");
  bind(so, t.lget_name_by_idx, "H5Lget_name_by_idx");
  bind(so, t.tcopy, "H5Tcopy");
  bind(so, t.tset_size, "H5Tset_size");
  bind(so, t.tclose, "H5Tclose");
  bind(so, t.eset_auto, "H5Eset_auto2");
  bind(so, t.tcreate, "H5Tcreate");
  bind(so, t.tinsert, "H5Tinsert");
  bind(so, t.dget_type, "H5Dget_type");
  bind(so, t.tget_size, "H5Tget_size");
  bind(so, t.tget_class, "H5Tget_class");
  bind(so, t.aopen_by_name, "H5Aopen_by_name");
  bind(so, t.aget_space, "H5Aget_space");
  bind(so, t.aget_type, "H5Aget_type");
  if (t.open() < 0)
    throw std::runtime_error("H5open failed");
  // the predefined types are globals, valid after H5open
  t.t_double = global_type(so, "H5T_NATIVE_DOUBLE_g");
  t.t_int = global_type(so, "H5T_NATIVE_INT_g");
  t.t_int64 = global_type(so, "H5T_NATIVE_LLONG_g");
  t.t_str = global_type(so, "H5T_C_S1_g");
  t.t_int8 = global_type(so, "H5T_NATIVE_INT8_g");
  t.t_ulong = global_type(so, "H5T_NATIVE_ULONG_g");
  t.t_i64 = global_type(so, "H5T_NATIVE_INT64_g");
  t.eset_auto(0, nullptr, nullptr); // errors come back as return codes, not printed
  t.so = so;
  h = t;
  return h;
}

void check(int64_t r, const std::string& what)
{
  if (r < 0)
    throw std::runtime_error("HDF5: " + what);
}

// Closes an HDF5 object when it goes out of scope.
struct Obj {
  hid_t id;
  herr_t (*close)(hid_t);
  Obj(hid_t i, herr_t (*c)(hid_t), const std::string& what) : id(i), close(c) { check(i, what); }
  ~Obj()
  {
    if (id >= 0)
      close(id);
  }
  Obj(const Obj&) = delete;
  Obj& operator=(const Obj&) = delete;
};

template<class T>
T read_attr(hid_t obj, const char* name, hid_t type)
{
  H5& h = h5();
  Obj a(h.aopen(obj, name, H5P_DEFAULT_), h.aclose, std::string("attribute ") + name);
  T v {};
  check(h.aread(a.id, type, &v), std::string("reading attribute ") + name);
  return v;
}

std::vector<double> read_vector(hid_t group, const char* name)
{
  H5& h = h5();
  Obj d(h.dopen(group, name, H5P_DEFAULT_), h.dclose, std::string("dataset ") + name);
  Obj sp(h.dget_space(d.id), h.sclose, "dataspace");
  const int64_t n = h.sget_npoints(sp.id);
  check(n, "dataspace size");
  std::vector<double> v(static_cast<size_t>(n));
  if (n > 0)
    check(h.dread(d.id, h.t_double, 0, 0, H5P_DEFAULT_, v.data()), std::string("reading ") + name);
  return v;
}

std::vector<std::string> link_names(hid_t group)
{
  H5& h = h5();
  hsize_t n = 0;
  check(h.gget_num(group, &n), "group size");
  std::vector<std::string> out;
  for (hsize_t k = 0; k < n; ++k) {
    char buf[256];
    const int64_t len = h.lget_name_by_idx(group, ".", H5_INDEX_NAME_, H5_ITER_INC_, k, buf, sizeof(buf),
      H5P_DEFAULT_);
    check(len, "link name");
    out.emplace_back(buf, static_cast<size_t>(std::min<int64_t>(len, 255)));
  }
  return out;
}

void write_attr(hid_t obj, const char* name, hid_t type, const void* v)
{
  H5& h = h5();
  Obj sp(h.screate(0 /* H5S_SCALAR */), h.sclose, "scalar dataspace");
  Obj a(h.acreate(obj, name, type, sp.id, H5P_DEFAULT_, H5P_DEFAULT_), h.aclose, std::string("attribute ") + name);
  check(h.awrite(a.id, type, v), std::string("writing attribute ") + name);
}

void write_str_attr(hid_t obj, const char* name, const std::string& s)
{
  H5& h = h5();
  Obj t(h.tcopy(h.t_str), h.tclose, "string type");
  check(h.tset_size(t.id, std::max<size_t>(s.size(), 1)), "string size");
  write_attr(obj, name, t.id, s.c_str());
}

void write_vector(hid_t group, const char* name, const double* v, size_t n)
{
  H5& h = h5();
  const hsize_t dims[1] = {static_cast<hsize_t>(n)};
  Obj sp(h.screate_simple(1, dims, nullptr), h.sclose, "dataspace");
  Obj d(h.dcreate(group, name, h.t_double, sp.id, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.dclose,
    std::string("dataset ") + name);
  check(h.dwrite(d.id, h.t_double, 0, 0, H5P_DEFAULT_, v), std::string("writing ") + name);
}

} // namespace

bool hdf5_available()
{
  try {
    h5();
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

bool is_hdf5_path(const std::string& path)
{
  const auto n = path.size();
  return (n > 3 && path.compare(n - 3, 3, ".h5") == 0) || (n > 5 && path.compare(n - 5, 5, ".hdf5") == 0);
}

void read_nuclide_hdf5(HostNuclide& nuc, const std::string& path, const std::string& name, double temperature,
  double tolerance)
{
  H5& h = h5();
  Obj f(h.fopen(path.c_str(), H5F_ACC_RDONLY_, H5P_DEFAULT_), h.fclose, "opening " + path);
  Obj g(h.gopen(f.id, name.c_str(), H5P_DEFAULT_), h.gclose, "nuclide " + name + " in " + path);
  nuc.awr = read_attr<double>(g.id, "atomic_weight_ratio", h.t_double);

  // the temperature nearest to the requested one, within the tolerance
  // (TemperatureMethod::NEAREST, src/nuclide.cpp:84-95, 133-166)
  Obj kt(h.gopen(g.id, "kTs", H5P_DEFAULT_), h.gclose, "kTs of " + name);
  std::string best;
  double best_kT = 0.0, best_dT = INFINITY;
  for (const auto& ds : link_names(kt.id)) {
    const std::vector<double> v = read_vector(kt.id, ds.c_str());
    if (v.size() != 1)
      continue;
    const double T = std::round(v[0] / K_BOLTZMANN);
    if (std::fabs(T - temperature) < best_dT) {
      best_dT = std::fabs(T - temperature);
      best = ds;
      best_kT = v[0];
    }
  }
  if (best.empty() || best_dT >= tolerance)
    throw std::runtime_error("Nuclear data library does not contain cross sections for " + name + " at or near " +
                             std::to_string(temperature) + " K");
  nuc.kT = best_kT;
  {
    Obj eg(h.gopen(g.id, "energy", H5P_DEFAULT_), h.gclose, "energy of " + name);
    nuc.energy = read_vector(eg.id, best.c_str());
  }
  const size_t n = nuc.energy.size();
  if (n < 2)
    throw std::runtime_error("nuclide grid too short: " + name + " in " + path);

  // Nuclide::create_derived (src/nuclide.cpp:359-425): TOTAL sums the
  // non-redundant reactions in file order, ABSORPTION the disappearance
  // ones (MT 102-117); the elastic column is reaction MT 2
  nuc.total.assign(n, 0.0);
  nuc.absorption.assign(n, 0.0);
  nuc.elastic.assign(n, 0.0);
  bool have_elastic = false;
  Obj rg(h.gopen(g.id, "reactions", H5P_DEFAULT_), h.gclose, "reactions of " + name);
  for (const auto& rn : link_names(rg.id)) {
    if (rn.rfind("reaction_", 0) != 0)
      continue;
    Obj rx(h.gopen(rg.id, rn.c_str(), H5P_DEFAULT_), h.gclose, rn);
    const int mt = read_attr<int>(rx.id, "mt", h.t_int);
    const int redundant = h.aexists(rx.id, "redundant") > 0 ? read_attr<int>(rx.id, "redundant", h.t_int) : 0;
    if (redundant)
      continue;
    const bool disappearance = mt >= 102 && mt <= 117;
    if (mt != 2 && !disappearance)
      throw std::runtime_error("nuclide " + name + ": reaction MT=" + std::to_string(mt) +
                               " is outside the GPU path (elastic scattering and capture only)");
    Obj tg(h.gopen(rx.id, best.c_str(), H5P_DEFAULT_), h.gclose, rn + "/" + best);
    Obj ds(h.dopen(tg.id, "xs", H5P_DEFAULT_), h.dclose, rn + "/" + best + "/xs");
    const int thr = read_attr<int>(ds.id, "threshold_idx", h.t_int);
    const std::vector<double> xs = read_vector(tg.id, "xs");
    if (thr < 0 || static_cast<size_t>(thr) + xs.size() > n)
      throw std::runtime_error("nuclide " + name + ": " + rn + " does not fit the energy grid");
    for (size_t k = 0; k < xs.size(); ++k)
      nuc.total[thr + k] += xs[k];
    if (disappearance)
      for (size_t k = 0; k < xs.size(); ++k)
        nuc.absorption[thr + k] += xs[k];
    if (mt == 2) {
      if (thr != 0)
        throw std::runtime_error("nuclide " + name + ": elastic scattering must start at the first grid point");
      nuc.elastic = xs;
      have_elastic = true;
    }
  }
  if (!have_elastic)
    throw std::runtime_error("nuclide " + name + " has no elastic scattering (MT 2)");
}

void write_nuclide_hdf5(const std::string& path, const std::string& name, double awr, double kT, int temperature,
  const double* energy, const double* elastic, const double* capture, size_t n)
{
  H5& h = h5();
  Obj f(h.fcreate(path.c_str(), H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_), h.fclose, "creating " + path);
  write_str_attr(f.id, "filetype", "data_neutron");
  {
    const hsize_t dims[1] = {2};
    const int version[2] = {3, 0};
    Obj sp(h.screate_simple(1, dims, nullptr), h.sclose, "dataspace");
    Obj a(h.acreate(f.id, "version", h.t_int, sp.id, H5P_DEFAULT_, H5P_DEFAULT_), h.aclose, "version");
    check(h.awrite(a.id, h.t_int, version), "writing version");
  }
  Obj g(h.gcreate(f.id, name.c_str(), H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, "group " + name);
  const int zero = 0;
  write_attr(g.id, "Z", h.t_int, &zero);
  write_attr(g.id, "A", h.t_int, &zero);
  write_attr(g.id, "metastable", h.t_int, &zero);
  write_attr(g.id, "atomic_weight_ratio", h.t_double, &awr);
  const std::string tname = std::to_string(temperature) + "K";
  {
    Obj kt(h.gcreate(g.id, "kTs", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, "kTs");
    Obj sp(h.screate(0), h.sclose, "scalar dataspace");
    Obj d(h.dcreate(kt.id, tname.c_str(), h.t_double, sp.id, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.dclose,
      "kT dataset");
    check(h.dwrite(d.id, h.t_double, 0, 0, H5P_DEFAULT_, &kT), "writing kT");
  }
  {
    Obj eg(h.gcreate(g.id, "energy", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, "energy");
    write_vector(eg.id, tname.c_str(), energy, n);
  }
  Obj rg(h.gcreate(g.id, "reactions", H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, "reactions");
  const struct {
    int mt;
    const char* group;
    const char* label;
    int cm;
    const double* xs;
  } rxs[2] = {{2, "reaction_002", "(n,elastic)", 1, elastic}, {102, "reaction_102", "(n,gamma)", 0, capture}};
  for (const auto& r : rxs) {
    Obj rx(h.gcreate(rg.id, r.group, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, r.group);
    write_attr(rx.id, "mt", h.t_int, &r.mt);
    write_str_attr(rx.id, "label", r.label);
    const double q = 0.0;
    write_attr(rx.id, "Q_value", h.t_double, &q);
    write_attr(rx.id, "center_of_mass", h.t_int, &r.cm);
    write_attr(rx.id, "redundant", h.t_int, &zero);
    Obj tg(h.gcreate(rx.id, tname.c_str(), H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.gclose, tname);
    write_vector(tg.id, "xs", r.xs, n);
    Obj ds(h.dopen(tg.id, "xs", H5P_DEFAULT_), h.dclose, "xs");
    write_attr(ds.id, "threshold_idx", h.t_int, &zero);
  }
}

// ---------------------------------------------------------------------------
// Writer for the reference's output files (hdf5_interface.cpp conventions:
// bool as INT8, size_t as ULONG, int64_t as INT64, a string as a fixed-length
// C string of its own length, an array of strings padded to the longest + 1,
// and no object at all for an empty string)

H5Writer::H5Writer(const std::string& path)
{
  H5& h = h5();
  file_ = h.fcreate(path.c_str(), H5F_ACC_TRUNC_, H5P_DEFAULT_, H5P_DEFAULT_);
  check(file_, "creating " + path);
}

H5Writer::~H5Writer()
{
  H5& h = h5();
  for (auto it = groups_.rbegin(); it != groups_.rend(); ++it)
    h.gclose(*it);
  if (file_ >= 0)
    h.fclose(file_);
}

int64_t H5Writer::group(int64_t parent, const std::string& name)
{
  H5& h = h5();
  const hid_t g = h.gcreate(parent, name.c_str(), H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_);
  check(g, "creating group " + name);
  groups_.push_back(g);
  return g;
}

namespace {

void write_any(hid_t obj, const std::string& name, hid_t type, int ndim, const hsize_t* dims, const void* data,
  bool attribute)
{
  H5& h = h5();
  Obj sp(ndim > 0 ? h.screate_simple(ndim, dims, nullptr) : h.screate(0 /* H5S_SCALAR */), h.sclose, "dataspace");
  if (attribute) {
    Obj a(h.acreate(obj, name.c_str(), type, sp.id, H5P_DEFAULT_, H5P_DEFAULT_), h.aclose, "attribute " + name);
    check(h.awrite(a.id, type, data), "writing attribute " + name);
  } else {
    Obj d(h.dcreate(obj, name.c_str(), type, sp.id, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.dclose,
      "dataset " + name);
    check(h.dwrite(d.id, type, 0, 0, H5P_DEFAULT_, data), "writing " + name);
  }
}

hid_t h5type(H5Writer::Kind k)
{
  H5& h = h5();
  switch (k) {
  case H5Writer::BOOL: return h.t_int8;
  case H5Writer::INT: return h.t_int;
  case H5Writer::INT64: return h.t_i64;
  case H5Writer::SIZE: return h.t_ulong;
  default: return h.t_double;
  }
}

} // namespace

void H5Writer::attr(int64_t obj, const std::string& name, Kind k, const void* v, size_t n, bool array)
{
  const hsize_t dims[1] = {static_cast<hsize_t>(n)};
  write_any(obj, name, h5type(k), array ? 1 : 0, dims, v, true);
}

void H5Writer::data(int64_t obj, const std::string& name, Kind k, const void* v, const std::vector<size_t>& shape)
{
  std::vector<hsize_t> dims(shape.begin(), shape.end());
  write_any(obj, name, h5type(k), static_cast<int>(dims.size()), dims.data(), v, false);
}

void H5Writer::attr_str(int64_t obj, const std::string& name, const std::string& s)
{
  if (s.empty())
    return;
  H5& h = h5();
  Obj t(h.tcopy(h.t_str), h.tclose, "string type");
  check(h.tset_size(t.id, s.size()), "string size");
  write_any(obj, name, t.id, 0, nullptr, s.c_str(), true);
}

void H5Writer::data_str(int64_t obj, const std::string& name, const std::string& s)
{
  if (s.empty())
    return;
  H5& h = h5();
  Obj t(h.tcopy(h.t_str), h.tclose, "string type");
  check(h.tset_size(t.id, s.size()), "string size");
  write_any(obj, name, t.id, 0, nullptr, s.c_str(), false);
}

void H5Writer::data_strs(int64_t obj, const std::string& name, const std::vector<std::string>& v)
{
  size_t m = 1;
  for (const auto& s : v)
    m = std::max(m, s.size() + 1);
  std::vector<char> buf(v.size() * m, '\0');
  for (size_t i = 0; i < v.size(); ++i)
    std::copy(v[i].begin(), v[i].end(), buf.begin() + static_cast<std::ptrdiff_t>(i * m));
  H5& h = h5();
  Obj t(h.tcopy(h.t_str), h.tclose, "string type");
  check(h.tset_size(t.id, m), "string size");
  const hsize_t dims[1] = {static_cast<hsize_t>(v.size())};
  write_any(obj, name, t.id, 1, dims, buf.data(), false);
}

void H5Writer::source_bank(int64_t obj, const std::string& name, int64_t n_sites)
{
  // the file layout of a source site (h5banktype(false), src/state_point.cpp:557-589):
  // r and u as {x, y, z} doubles, E, time, wgt, then delayed_group, surf_id, particle
  H5& h = h5();
  constexpr int COMPOUND = 6; // H5T_COMPOUND
  Obj pos(h.tcreate(COMPOUND, 24), h.tclose, "position type");
  check(h.tinsert(pos.id, "x", 0, h.t_double), "position type");
  check(h.tinsert(pos.id, "y", 8, h.t_double), "position type");
  check(h.tinsert(pos.id, "z", 16, h.t_double), "position type");
  Obj site(h.tcreate(COMPOUND, 2 * 24 + 3 * 8 + 3 * 4), h.tclose, "source site type");
  check(h.tinsert(site.id, "r", 0, pos.id), "source site type");
  check(h.tinsert(site.id, "u", 24, pos.id), "source site type");
  check(h.tinsert(site.id, "E", 48, h.t_double), "source site type");
  check(h.tinsert(site.id, "time", 56, h.t_double), "source site type");
  check(h.tinsert(site.id, "wgt", 64, h.t_double), "source site type");
  check(h.tinsert(site.id, "delayed_group", 72, h.t_int), "source site type");
  check(h.tinsert(site.id, "surf_id", 76, h.t_int), "source site type");
  check(h.tinsert(site.id, "particle", 80, h.t_int), "source site type");
  const hsize_t dims[1] = {static_cast<hsize_t>(n_sites)};
  Obj sp(h.screate_simple(1, dims, nullptr), h.sclose, "dataspace");
  // created and left unwritten: its sites read back as the fill value (zeros)
  Obj d(h.dcreate(obj, name.c_str(), site.id, sp.id, H5P_DEFAULT_, H5P_DEFAULT_, H5P_DEFAULT_), h.dclose,
    "dataset " + name);
}

size_t h5_read(const std::string& path, const std::string& object, const std::string& attribute, int kind,
  void* buf, size_t cap)
{
  H5& h = h5();
  Obj f(h.fopen(path.c_str(), H5F_ACC_RDONLY_, H5P_DEFAULT_), h.fclose, "opening " + path);
  const bool is_attr = !attribute.empty();
  hid_t id, space, type;
  herr_t (*close)(hid_t);
  if (is_attr) {
    id = h.aopen_by_name(f.id, object.empty() ? "." : object.c_str(), attribute.c_str(), H5P_DEFAULT_,
      H5P_DEFAULT_);
    check(id, "attribute " + attribute + " of " + object);
    close = h.aclose;
  } else {
    id = h.dopen(f.id, object.c_str(), H5P_DEFAULT_);
    check(id, "dataset " + object);
    close = h.dclose;
  }
  Obj o(id, close, object);
  Obj sp(is_attr ? h.aget_space(id) : h.dget_space(id), h.sclose, "dataspace");
  const int64_t n = h.sget_npoints(sp.id);
  check(n, "dataspace size");
  Obj t(is_attr ? h.aget_type(id) : h.dget_type(id), h.tclose, "type");
  constexpr int STRING = 3; // H5T_STRING
  if (kind == 2) {
    if (h.tget_class(t.id) != STRING)
      throw std::runtime_error(object + " " + attribute + " is not a string");
    const size_t size = h.tget_size(t.id);
    const size_t total = static_cast<size_t>(n) * size;
    if (buf && total <= cap && total > 0)
      check(is_attr ? h.aread(id, t.id, buf) : h.dread(id, t.id, 0, 0, H5P_DEFAULT_, buf), "reading " + object);
    return total;
  }
  const size_t count = static_cast<size_t>(n);
  if (buf && count <= cap && count > 0) {
    if (h.tget_class(t.id) == STRING)
      throw std::runtime_error(object + " " + attribute + " is a string");
    const hid_t mt = kind == 0 ? h.t_double : h.t_i64;
    check(is_attr ? h.aread(id, mt, buf) : h.dread(id, mt, 0, 0, H5P_DEFAULT_, buf), "reading " + object);
  }
  return count;
}

} // namespace omc
