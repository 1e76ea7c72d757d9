// The model-query part of OpenMC's C API: cells, materials, nuclides,
// meshes, filters, tallies and weight windows as openmc.lib reads them
// (openmc/lib/{cell,material,nuclide,mesh,filter,tally,weight_windows}.py),
// answered from the host copy of the model that openmc_init read.
//
// openmc.lib sets argtypes on about 190 functions when it is imported, so the
// library exports every one of them.  Getters return the loaded model's
// values with the reference's error codes and messages.  Functions that would
// change the model after it has been uploaded to the GPU (new cells, new
// filters, other scores, new meshes), or that belong to subsystems outside
// the fixed-source transport path (volumes, plotting, CMFD, restart, DAGMC,
// eigenvalue), return OPENMC_E_UNASSIGNED with a message that names what is
// not supported.  Filter getters of filter types the path has no use for
// (cell, material, Legendre, ...) behave as the reference does when the
// filter at that index is of another type: OPENMC_E_INVALID_TYPE.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "capi_internal.h"
#include "openmc_amd/capi.h"

// Settings globals bound by openmc/lib/settings.py:25-40 and :62.  Outside
// the GPU path they only report the reference's defaults
// (src/settings.cpp:60-120); weight_windows_on reflects the model.
extern "C" {
bool cmfd_run = false;
bool entropy_on = false;
int32_t max_lost_particles = 10;
bool need_depletion_rx = false;
bool output_summary = true;
double rel_max_lost_particles = 1.0e-6;
bool restart_run = false;
bool weight_windows_on = false;
char* path_statepoint_c = nullptr;
}

namespace omc {
namespace {

// Per-model tables the getters hand out pointers into.  They are rebuilt
// whenever openmc_init loads another model, and stay valid until then, as the
// reference's borrowed pointers do.
struct ApiTables {
  const HostModel* model = nullptr;
  long long generation = -1;
  size_t n_tallies = 0;
  // filters: the model's own list, or the concatenation of a problem set's
  std::vector<const HostFilter*> filters;
  std::vector<int> filter_part;
  std::vector<std::pair<int, const HostMesh*>> meshes; // (id, mesh)
  std::vector<int> mesh_part;
  std::vector<std::vector<int32_t>> tally_filters;
  std::vector<int> score_flux = {-1};      // SCORE_FLUX (include/openmc/constants.h)
  std::vector<int> nuclide_total = {-1};   // "total"
  std::vector<std::vector<int>> mat_nuclides;
  std::vector<std::vector<double>> mat_densities;
  std::vector<double> mat_volume;
  std::vector<bool> mat_depletable;
  std::vector<int32_t> cell_fill; // material, universe or lattice index per cell
  std::vector<int32_t> cell_instances;
  std::vector<std::string> nuclide_names;
  std::vector<std::vector<int>> mesh_dims;
  std::vector<std::vector<double>> mesh_ll, mesh_ur, mesh_width;
  std::vector<int> ww_mesh_index; // per weight-window set
};

ApiTables T;
std::string statepoint_path;

int fail(int code, const std::string& msg)
{
  api_set_errmsg(msg);
  return code;
}

int unsupported(const char* what)
{
  return fail(OPENMC_E_UNASSIGNED,
    std::string(what) + " is outside the MI355X fixed-source transport path and is not supported");
}

// Cell instance counts: how many times each universe occurs below the root
// universe, counting lattice elements but not a lattice's outer universe, as
// the reference counts them (src/geometry_aux.cpp:451-495).
int32_t universe_count(const HostModel& m, int u, int target, std::unordered_map<int, int32_t>& memo)
{
  if (u == target)
    return 1;
  auto it = memo.find(u);
  if (it != memo.end())
    return it->second;
  int32_t count = 0;
  for (int k = m.universes[u].cell_begin; k < m.universes[u].cell_end; ++k) {
    const DevCell& c = m.cells[m.univ_cells[k]];
    if (c.fill_type == FILL_UNIVERSE) {
      count += universe_count(m, c.fill, target, memo);
    } else if (c.fill_type == FILL_LATTICE) {
      const DevLattice& L = m.lattices[c.fill];
      const int n = L.n[0] * L.n[1] * (L.is_3d ? L.n[2] : 1);
      for (int e = 0; e < n; ++e) {
        int v = m.lat_univ[L.univ_offset + e];
        if (v != C_NONE)
          count += universe_count(m, v, target, memo);
      }
    }
  }
  memo[u] = count;
  return count;
}

void rebuild_tables(const HostModel* m);

// The loaded model with its tables current, or null (also when the tables
// cannot be built: the message is then set).  Never throws: every C entry
// point below calls it.
const HostModel* model_or_null()
{
  const HostModel* m = api_model();
  if (api_generation() != T.generation || m != T.model || (m && T.n_tallies != m->tallies.size())) {
    try {
      rebuild_tables(m);
    } catch (const std::exception& e) {
      T = ApiTables();
      api_set_errmsg(std::string("model tables: ") + e.what());
      return nullptr;
    }
  }
  return m;
}

void rebuild_tables(const HostModel* m)
{
  {
    T = ApiTables();
    T.model = m;
    T.generation = api_generation();
    if (!m)
      return;
    T.n_tallies = m->tallies.size();
    // meshes and filters, per problem for a problem set
    auto add_part = [&](int part, const std::vector<std::pair<int, HostMesh>>& meshes,
                      const std::vector<HostFilter>& filters) {
      const int mesh0 = static_cast<int>(T.meshes.size());
      for (const auto& e : meshes) {
        T.meshes.emplace_back(e.first, &e.second);
        T.mesh_part.push_back(part);
      }
      for (const HostFilter& f : filters) {
        T.filters.push_back(&f);
        T.filter_part.push_back(part);
      }
      (void)mesh0;
    };
    if (m->parts.size() > 1) {
      for (size_t k = 0; k < m->parts.size(); ++k)
        add_part(static_cast<int>(k), m->parts[k].meshes, m->parts[k].filters);
    } else {
      add_part(0, m->meshes, m->filters);
    }
    for (size_t i = 0; i < T.meshes.size(); ++i) {
      const HostMesh& me = *T.meshes[i].second;
      T.mesh_dims.emplace_back(me.shape, me.shape + me.n_dim);
      T.mesh_ll.emplace_back(me.ll, me.ll + me.n_dim);
      T.mesh_ur.emplace_back(me.ur, me.ur + me.n_dim);
      T.mesh_width.emplace_back(me.width, me.width + me.n_dim);
    }
    for (const HostTally& t : m->tallies) {
      std::vector<int32_t> idx;
      for (int fid : t.filter_ids) {
        for (size_t i = 0; i < T.filters.size(); ++i) {
          if (T.filters[i]->id == fid && T.filter_part[i] == t.part) {
            idx.push_back(static_cast<int32_t>(i));
            break;
          }
        }
      }
      T.tally_filters.push_back(idx);
    }
    const size_t n_mat = m->mat_ids.size();
    for (size_t i = 0; i < n_mat; ++i) {
      std::vector<int> nuc;
      std::vector<double> dens;
      for (int e = m->mat_off[i]; e < m->mat_off[i + 1]; ++e) {
        nuc.push_back(m->mat_nuc[e]);
        dens.push_back(m->mat_dens[e]);
      }
      T.mat_nuclides.push_back(nuc);
      T.mat_densities.push_back(dens);
    }
    T.mat_volume.assign(n_mat, -1.0);
    T.mat_depletable.assign(n_mat, false);
    for (size_t i = 0; i < n_mat && i < m->mat_depletable.size(); ++i)
      T.mat_depletable[i] = m->mat_depletable[i];
    for (const DevCell& c : m->cells)
      T.cell_fill.push_back(c.fill_type == FILL_MATERIAL ? c.material : c.fill);
    std::unordered_map<int, std::unordered_map<int, int32_t>> memo;
    for (const DevCell& c : m->cells) {
      int root = m->root_universe;
      if (c.part < static_cast<int>(m->parts.size()))
        root = m->parts[c.part].root_universe;
      T.cell_instances.push_back(universe_count(*m, root, c.universe, memo[c.universe]));
    }
    T.nuclide_names.resize(m->nuclides.size());
    for (const auto& e : m->nuclide_map)
      if (e.second >= 0 && e.second < static_cast<int>(T.nuclide_names.size()))
        T.nuclide_names[e.second] = e.first;
    for (size_t k = 0; k < m->nuclides.size(); ++k)
      if (T.nuclide_names[k].empty())
        T.nuclide_names[k] = m->nuclides[k].name;
    for (const HostWWSet& ws : m->ww.sets) {
      int idx = -1;
      for (size_t i = 0; i < T.meshes.size(); ++i)
        if (T.meshes[i].first == ws.mesh_id) {
          idx = static_cast<int>(i);
          break;
        }
      T.ww_mesh_index.push_back(idx);
    }
  }
}

const HostModel* need_model()
{
  const HostModel* m = model_or_null();
  if (!m)
    throw std::runtime_error("OpenMC has not been initialized.");
  return m;
}

template<class F>
int api(F f)
{
  try {
    return f();
  } catch (const std::exception& e) {
    return fail(OPENMC_E_ALLOCATE, e.what());
  }
}

bool bad_index(int64_t index, size_t size)
{
  return index < 0 || static_cast<size_t>(index) >= size;
}

#define CHECK_INDEX(idx, size, what)                                                                                  \
  if (bad_index((idx), (size)))                                                                                       \
    return fail(OPENMC_E_OUT_OF_BOUNDS, "Index in " what " array is out of bounds.");

int check_filter(int32_t index)
{
  const HostModel* m = model_or_null();
  if (!m || bad_index(index, T.filters.size()))
    return fail(OPENMC_E_OUT_OF_BOUNDS, "Filter index is out of bounds.");
  return 0;
}

// The reference's answer when a typed filter getter meets a filter of
// another type (e.g. src/tallies/filter_cell.cpp, filter_legendre.cpp)
int wrong_filter(int32_t index, const char* kind)
{
  if (int err = check_filter(index))
    return err;
  return fail(OPENMC_E_INVALID_TYPE, std::string("Tried to get/set ") + kind + " on a non-" + kind + " filter.");
}

int check_mesh(int32_t index)
{
  const HostModel* m = model_or_null();
  if (!m || bad_index(index, T.meshes.size()))
    return fail(OPENMC_E_OUT_OF_BOUNDS, "Index in meshes vector is out of bounds.");
  return 0;
}

int check_ww(int32_t index)
{
  const HostModel* m = model_or_null();
  if (!m || bad_index(index, m->ww.sets.size()))
    return fail(OPENMC_E_OUT_OF_BOUNDS, "Index in weight windows vector is out of bounds.");
  return 0;
}

const HostWWSet& ww_set(int32_t index)
{
  return model_or_null()->ww.sets[static_cast<size_t>(index)];
}

double material_density_gpcc(const HostModel& m, int index)
{
  // Material::normalize_density (src/material.cpp:457-466)
  double rho = 0.0;
  for (int e = m.mat_off[index]; e < m.mat_off[index + 1]; ++e)
    rho += m.mat_dens[e] * m.nuclides[m.mat_nuc[e]].awr * MASS_NEUTRON / N_AVOGADRO;
  return rho;
}

// Radial Zernike polynomial R_n^m(rho) from its explicit sum
// sum_k (-1)^k (n-k)! / (k! ((n+m)/2-k)! ((n-m)/2-k)!) rho^(n-2k)
double zernike_radial(int n, int m, double rho)
{
  m = std::abs(m);
  if ((n - m) % 2)
    return 0.0;
  double sum = 0.0;
  for (int k = 0; k <= (n - m) / 2; ++k) {
    double c = (k % 2) ? -1.0 : 1.0;
    c *= std::tgamma(n - k + 1.0) /
         (std::tgamma(k + 1.0) * std::tgamma((n + m) / 2 - k + 1.0) * std::tgamma((n - m) / 2 - k + 1.0));
    sum += c * std::pow(rho, n - 2 * k);
  }
  return sum;
}

} // namespace

void api_note_statepoint(const std::string& path)
{
  statepoint_path = path;
  path_statepoint_c = statepoint_path.empty() ? nullptr : &statepoint_path[0];
}

} // namespace omc

using namespace omc;

extern "C" {

// ---------------------------------------------------------------------------
// Sizes (src/cell.cpp:1820, src/material.cpp:1599, src/mesh.cpp:3900,
// src/nuclide.cpp:1105, src/tallies/filter.cpp:62)
// ---------------------------------------------------------------------------
int cells_size(void)
{
  const HostModel* m = model_or_null();
  return m ? static_cast<int>(m->cells.size()) : 0;
}

size_t n_materials(void)
{
  const HostModel* m = model_or_null();
  return m ? m->mat_ids.size() : 0;
}

int n_meshes(void)
{
  return model_or_null() ? static_cast<int>(T.meshes.size()) : 0;
}

size_t nuclides_size(void)
{
  const HostModel* m = model_or_null();
  return m ? m->nuclides.size() : 0;
}

size_t tally_filters_size(void)
{
  return model_or_null() ? T.filters.size() : 0;
}

bool openmc_master(void)
{
  return true; // one process: src/message_passing.cpp:16-19
}

// ---------------------------------------------------------------------------
// Settings (src/settings.cpp:1285-1324)
// ---------------------------------------------------------------------------
int openmc_get_n_batches(int* n_batches, bool get_max_batches)
{
  (void)get_max_batches; // no triggers: n_max_batches == n_batches
  const HostModel* m = model_or_null();
  *n_batches = m ? m->n_batches : 0;
  return 0;
}

int openmc_set_n_batches(int32_t n_batches, bool set_max_batches, bool add_statepoint_batch)
{
  (void)set_max_batches;
  return api([&]() -> int {
    HostModel* m = const_cast<HostModel*>(need_model());
    if (m->n_inactive >= n_batches)
      return fail(OPENMC_E_INVALID_ARGUMENT, "Number of active batches must be greater than zero.");
    m->n_batches = n_batches;
    if (add_statepoint_batch &&
        std::find(m->statepoint_batches.begin(), m->statepoint_batches.end(), n_batches) ==
          m->statepoint_batches.end())
      m->statepoint_batches.push_back(n_batches);
    return 0;
  });
}

// ---------------------------------------------------------------------------
// Nuclides (src/nuclide.cpp:1055-1110)
// ---------------------------------------------------------------------------
int openmc_get_nuclide_index(const char* name, int* index)
{
  const HostModel* m = model_or_null();
  if (m && name) {
    auto it = m->nuclide_map.find(name);
    if (it != m->nuclide_map.end()) {
      *index = it->second;
      return 0;
    }
  }
  return fail(OPENMC_E_DATA, std::string("No nuclide named '") + (name ? name : "") + "' has been loaded.");
}

int openmc_nuclide_name(int index, const char** name)
{
  const HostModel* m = model_or_null();
  if (!m || bad_index(index, m->nuclides.size()))
    return fail(OPENMC_E_OUT_OF_BOUNDS, "Index in nuclides vector is out of bounds.");
  *name = T.nuclide_names[index].c_str();
  return 0;
}

int openmc_load_nuclide(const char* name, const double* temps, int n)
{
  (void)temps;
  (void)n;
  const HostModel* m = model_or_null();
  if (m && name && m->nuclide_map.count(name))
    return 0; // already loaded (src/nuclide.cpp:1055-1062)
  return unsupported("loading a nuclide after openmc_init (the data are uploaded to the GPU once)");
}

int openmc_nuclide_collapse_rate(int index, int MT, double temperature, const double* energy, const double* flux,
  int n, double* xs)
{
  (void)index, (void)MT, (void)temperature, (void)energy, (void)flux, (void)n, (void)xs;
  return unsupported("openmc_nuclide_collapse_rate (depletion)");
}

// ---------------------------------------------------------------------------
// Materials (src/material.cpp:1290-1600)
// ---------------------------------------------------------------------------
int openmc_get_material_index(int32_t id, int32_t* index)
{
  const HostModel* m = model_or_null();
  if (m) {
    for (size_t i = 0; i < m->mat_ids.size(); ++i) {
      if (m->mat_ids[i] == id) {
        *index = static_cast<int32_t>(i);
        return 0;
      }
    }
  }
  return fail(OPENMC_E_INVALID_ID, "No material exists with ID=" + std::to_string(id) + ".");
}

int openmc_material_get_id(int32_t index, int32_t* id)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *id = m->mat_ids[index];
  return 0;
}

int openmc_material_set_id(int32_t index, int32_t id)
{
  HostModel* m = const_cast<HostModel*>(model_or_null());
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  for (size_t i = 0; i < m->mat_ids.size(); ++i)
    if (m->mat_ids[i] == id && static_cast<int32_t>(i) != index)
      return fail(OPENMC_E_INVALID_ID, "Two or more materials use the same unique ID: " + std::to_string(id));
  m->mat_ids[index] = id;
  return 0;
}

int openmc_material_get_name(int32_t index, const char** name)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *name = index < static_cast<int32_t>(m->mat_names.size()) ? m->mat_names[index].c_str() : "";
  return 0;
}

int openmc_material_set_name(int32_t index, const char* name)
{
  HostModel* m = const_cast<HostModel*>(model_or_null());
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  m->mat_names.resize(m->mat_ids.size());
  m->mat_names[index] = name ? name : "";
  return 0;
}

int openmc_material_get_densities(int32_t index, const int** nuclides, const double** densities, int* n)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *nuclides = T.mat_nuclides[index].data();
  *densities = T.mat_densities[index].data();
  *n = static_cast<int>(T.mat_nuclides[index].size());
  return 0;
}

int openmc_material_get_density(int32_t index, double* density)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *density = material_density_gpcc(*m, index);
  return 0;
}

int openmc_material_get_temperature(int32_t index, double* temperature)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *temperature = m->temperature_default; // no per-material temperatures on the path
  return 0;
}

int openmc_material_get_volume(int32_t index, double* volume)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  if (T.mat_volume[index] < 0.0)
    return fail(OPENMC_E_UNASSIGNED, "Volume for material with ID=" + std::to_string(m->mat_ids[index]) +
                                       " not set.");
  *volume = T.mat_volume[index];
  return 0;
}

int openmc_material_set_volume(int32_t index, double volume)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  if (volume < 0.0)
    return fail(OPENMC_E_INVALID_ARGUMENT, "Volume must be non-negative");
  T.mat_volume[index] = volume;
  return 0;
}

int openmc_material_get_depletable(int32_t index, bool* depletable)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *depletable = T.mat_depletable[index];
  return 0;
}

int openmc_material_set_depletable(int32_t index, bool depletable)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  T.mat_depletable[index] = depletable;
  return 0;
}

int openmc_material_get_fissionable(int32_t index, bool* fissionable)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->mat_ids.size() : 0, "materials");
  *fissionable = false; // the path's nuclides carry no fission data
  return 0;
}

int openmc_material_set_density(int32_t index, double density, const char* units)
{
  (void)index, (void)density, (void)units;
  return unsupported("changing a material's density after openmc_init");
}

int openmc_material_set_densities(int32_t index, int n, const char** name, const double* density)
{
  (void)index, (void)n, (void)name, (void)density;
  return unsupported("changing a material's nuclides after openmc_init");
}

int openmc_material_add_nuclide(int32_t index, const char* name, double density)
{
  (void)index, (void)name, (void)density;
  return unsupported("adding a nuclide to a material after openmc_init");
}

int openmc_extend_materials(int32_t n, int32_t* index_start, int32_t* index_end)
{
  (void)n, (void)index_start, (void)index_end;
  return unsupported("creating materials after openmc_init");
}

// ---------------------------------------------------------------------------
// Cells (src/cell.cpp:1150-1830)
// ---------------------------------------------------------------------------
int openmc_get_cell_index(int32_t id, int32_t* index)
{
  const HostModel* m = model_or_null();
  if (m) {
    for (size_t i = 0; i < m->cells.size(); ++i) {
      if (m->cells[i].id == id) {
        *index = static_cast<int32_t>(i);
        return 0;
      }
    }
  }
  return fail(OPENMC_E_INVALID_ID, "No cell exists with ID=" + std::to_string(id) + ".");
}

int openmc_cell_get_id(int32_t index, int32_t* id)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *id = m->cells[index].id;
  return 0;
}

int openmc_cell_set_id(int32_t index, int32_t id)
{
  (void)index, (void)id;
  return unsupported("changing a cell's ID after openmc_init (the GPU holds the cell table)");
}

int openmc_cell_get_name(int32_t index, const char** name)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *name = index < static_cast<int32_t>(m->cell_names.size()) ? m->cell_names[index].c_str() : "";
  return 0;
}

int openmc_cell_set_name(int32_t index, const char* name)
{
  HostModel* m = const_cast<HostModel*>(model_or_null());
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  m->cell_names.resize(m->cells.size());
  m->cell_names[index] = name ? name : "";
  return 0;
}

int openmc_cell_get_fill(int32_t index, int* type, int32_t** indices, int32_t* n)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *type = m->cells[index].fill_type; // Fill {MATERIAL, UNIVERSE, LATTICE} (include/openmc/cell.h:29)
  *indices = &T.cell_fill[index];
  *n = 1;
  return 0;
}

int openmc_cell_get_num_instances(int32_t index, int32_t* num_instances)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *num_instances = T.cell_instances[index];
  return 0;
}

int openmc_cell_get_temperature(int32_t index, const int32_t* instance, double* T_out)
{
  (void)instance;
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *T_out = m->temperature_default;
  return 0;
}

int openmc_cell_get_density(int32_t index, const int32_t* instance, double* rho)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  const DevCell& c = m->cells[index];
  if (c.fill_type != FILL_MATERIAL)
    return fail(OPENMC_E_UNASSIGNED, "Cell " + std::to_string(c.id) + ", instance " +
                                       std::to_string(instance ? *instance : -1) +
                                       " is not filled with a material.");
  *rho = c.material == MATERIAL_VOID ? 0.0 : material_density_gpcc(*m, c.material);
  return 0;
}

int openmc_cell_get_translation(int32_t index, double xyz[])
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  xyz[0] = xyz[1] = xyz[2] = 0.0; // translations are outside the path (the reader rejects them)
  return 0;
}

int openmc_cell_get_rotation(int32_t index, double rot[], size_t* n)
{
  (void)rot;
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->cells.size() : 0, "cells");
  *n = 0;
  return 0;
}

int openmc_cell_set_fill(int32_t index, int type, int32_t n, const int32_t* indices)
{
  (void)index, (void)type, (void)n, (void)indices;
  return unsupported("changing a cell's fill after openmc_init");
}

int openmc_cell_set_temperature(int32_t index, double T_in, const int32_t* instance, bool set_contained)
{
  (void)index, (void)T_in, (void)instance, (void)set_contained;
  return unsupported("changing a cell's temperature (one temperature per nuclide on the path)");
}

int openmc_cell_set_density(int32_t index, double rho, const int32_t* instance, bool set_contained)
{
  (void)index, (void)rho, (void)instance, (void)set_contained;
  return unsupported("changing a cell's density after openmc_init");
}

int openmc_cell_set_translation(int32_t index, const double xyz[])
{
  (void)index, (void)xyz;
  return unsupported("cell translations");
}

int openmc_cell_set_rotation(int32_t index, const double rot[], size_t rot_len)
{
  (void)index, (void)rot, (void)rot_len;
  return unsupported("cell rotations");
}

int openmc_cell_bounding_box(const int32_t index, double* llc, double* urc)
{
  (void)index, (void)llc, (void)urc;
  return unsupported("openmc_cell_bounding_box");
}

int openmc_global_bounding_box(double* llc, double* urc)
{
  (void)llc, (void)urc;
  return unsupported("openmc_global_bounding_box");
}

int openmc_find_cell(const double* xyz, int32_t* index, int32_t* instance)
{
  (void)xyz, (void)index, (void)instance;
  return unsupported("openmc_find_cell (cell searches run on the GPU inside the event loop)");
}

int openmc_extend_cells(int32_t n, int32_t* index_start, int32_t* index_end)
{
  (void)n, (void)index_start, (void)index_end;
  return unsupported("creating cells after openmc_init");
}

int openmc_dagmc_universe_get_cell_ids(int32_t univ_id, int32_t* ids, size_t* n)
{
  (void)univ_id, (void)ids, (void)n;
  return unsupported("DAGMC");
}

int openmc_dagmc_universe_get_num_cells(int32_t univ_id, size_t* n)
{
  (void)univ_id, (void)n;
  return unsupported("DAGMC");
}

// ---------------------------------------------------------------------------
// Meshes (src/mesh.cpp:2250-2700, 3900)
// ---------------------------------------------------------------------------
int openmc_get_mesh_index(int32_t id, int32_t* index)
{
  if (model_or_null()) {
    for (size_t i = 0; i < T.meshes.size(); ++i) {
      if (T.meshes[i].first == id) {
        *index = static_cast<int32_t>(i);
        return 0;
      }
    }
  }
  return fail(OPENMC_E_INVALID_ID, "No mesh exists with ID=" + std::to_string(id) + ".");
}

int openmc_mesh_get_id(int32_t index, int32_t* id)
{
  if (int err = check_mesh(index))
    return err;
  *id = T.meshes[index].first;
  return 0;
}

int openmc_mesh_set_id(int32_t index, int32_t id)
{
  if (int err = check_mesh(index))
    return err;
  T.meshes[index].first = id;
  return 0;
}

int openmc_mesh_get_type(int32_t index, char* type)
{
  if (int err = check_mesh(index))
    return err;
  std::strcpy(type, "regular");
  return 0;
}

int openmc_mesh_get_n_elements(int32_t index, size_t* n)
{
  if (int err = check_mesh(index))
    return err;
  *n = static_cast<size_t>(T.meshes[index].second->n_bins());
  return 0;
}

int openmc_mesh_get_volumes(int32_t index, double* volumes)
{
  if (int err = check_mesh(index))
    return err;
  const HostMesh& me = *T.meshes[index].second;
  double v = 1.0;
  for (int k = 0; k < me.n_dim; ++k)
    v *= me.width[k];
  std::fill(volumes, volumes + me.n_bins(), v);
  return 0;
}

int openmc_mesh_bounding_box(int32_t index, double* ll, double* ur)
{
  if (int err = check_mesh(index))
    return err;
  const HostMesh& me = *T.meshes[index].second;
  for (int k = 0; k < 3; ++k) {
    ll[k] = k < me.n_dim ? me.ll[k] : -INFTY;
    ur[k] = k < me.n_dim ? me.ur[k] : INFTY;
  }
  return 0;
}

int openmc_regular_mesh_get_dimension(int32_t index, int** dims, int* n)
{
  if (int err = check_mesh(index))
    return err;
  *dims = T.mesh_dims[index].data();
  *n = static_cast<int>(T.mesh_dims[index].size());
  return 0;
}

int openmc_regular_mesh_get_params(int32_t index, double** ll, double** ur, double** width, int* n)
{
  if (int err = check_mesh(index))
    return err;
  *ll = T.mesh_ll[index].data();
  *ur = T.mesh_ur[index].data();
  *width = T.mesh_width[index].data();
  *n = static_cast<int>(T.mesh_ll[index].size());
  return 0;
}

int openmc_regular_mesh_set_dimension(int32_t index, int n, const int* dims)
{
  (void)index, (void)n, (void)dims;
  return unsupported("changing a mesh after openmc_init (the tally is laid out on the GPU)");
}

int openmc_regular_mesh_set_params(int32_t index, int n, const double* ll, const double* ur, const double* width)
{
  (void)index, (void)n, (void)ll, (void)ur, (void)width;
  return unsupported("changing a mesh after openmc_init (the tally is laid out on the GPU)");
}

static int not_this_mesh(int32_t index, const char* kind)
{
  if (int err = check_mesh(index))
    return err;
  return fail(OPENMC_E_INVALID_TYPE, std::string("This function is only valid for ") + kind + " meshes.");
}

int openmc_rectilinear_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "rectilinear");
}

int openmc_rectilinear_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "rectilinear");
}

int openmc_cylindrical_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "cylindrical");
}

int openmc_cylindrical_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "cylindrical");
}

int openmc_spherical_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "spherical");
}

int openmc_spherical_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz)
{
  (void)grid_x, (void)nx, (void)grid_y, (void)ny, (void)grid_z, (void)nz;
  return not_this_mesh(index, "spherical");
}

int openmc_mesh_get_plot_bins(int32_t index, const void* origin, const void* width, int basis, int* pixels,
  int32_t* data)
{
  (void)index, (void)origin, (void)width, (void)basis, (void)pixels, (void)data;
  return unsupported("openmc_mesh_get_plot_bins (plotting)");
}

int openmc_mesh_material_volumes(int32_t index, int nx, int ny, int nz, int max_mats, int32_t* materials,
  double* volumes)
{
  (void)index, (void)nx, (void)ny, (void)nz, (void)max_mats, (void)materials, (void)volumes;
  return unsupported("openmc_mesh_material_volumes (volume calculation)");
}

int openmc_extend_meshes(int32_t n, const char* type, int32_t* index_start, int32_t* index_end)
{
  (void)n, (void)type, (void)index_start, (void)index_end;
  return unsupported("creating meshes after openmc_init");
}

// ---------------------------------------------------------------------------
// Filters (src/tallies/filter.cpp:200-260, filter_mesh.cpp, filter_energy.cpp)
// ---------------------------------------------------------------------------
int openmc_get_filter_index(int32_t id, int32_t* index)
{
  if (model_or_null()) {
    for (size_t i = 0; i < T.filters.size(); ++i) {
      if (T.filters[i]->id == id) {
        *index = static_cast<int32_t>(i);
        return 0;
      }
    }
  }
  return fail(OPENMC_E_INVALID_ID, "Could not find filter with ID " + std::to_string(id) + ".");
}

int openmc_filter_get_id(int32_t index, int32_t* id)
{
  if (int err = check_filter(index))
    return err;
  *id = T.filters[index]->id;
  return 0;
}

int openmc_filter_set_id(int32_t index, int32_t id)
{
  if (int err = check_filter(index))
    return err;
  const_cast<HostFilter*>(T.filters[index])->id = id;
  return 0;
}

int openmc_filter_get_type(int32_t index, char* type)
{
  if (int err = check_filter(index))
    return err;
  std::strcpy(type, T.filters[index]->type.c_str());
  return 0;
}

int openmc_filter_get_num_bins(int32_t index, int* n_bins)
{
  if (int err = check_filter(index))
    return err;
  *n_bins = T.filters[index]->n_bins;
  return 0;
}

int openmc_new_filter(const char* type, int32_t* index)
{
  (void)type, (void)index;
  return unsupported("creating filters after openmc_init (the tally is laid out on the GPU)");
}

int openmc_mesh_filter_get_mesh(int32_t index, int32_t* index_mesh)
{
  if (int err = check_filter(index))
    return err;
  const HostFilter& f = *T.filters[index];
  if (f.type != "mesh")
    return fail(OPENMC_E_INVALID_TYPE, "Tried to get mesh on a non-mesh filter.");
  for (size_t i = 0; i < T.meshes.size(); ++i) {
    if (T.meshes[i].first == f.mesh_id && T.mesh_part[i] == T.filter_part[index]) {
      *index_mesh = static_cast<int32_t>(i);
      return 0;
    }
  }
  return fail(OPENMC_E_INVALID_ID, "mesh filter refers to an unknown mesh");
}

int openmc_mesh_filter_set_mesh(int32_t index, int32_t index_mesh)
{
  (void)index_mesh;
  if (int err = check_filter(index))
    return err;
  return unsupported("changing a mesh filter's mesh after openmc_init");
}

int openmc_mesh_filter_get_translation(int32_t index, double translation[3])
{
  if (int err = check_filter(index))
    return err;
  if (T.filters[index]->type != "mesh")
    return fail(OPENMC_E_INVALID_TYPE, "Tried to get a translation from a non-mesh-based filter.");
  translation[0] = translation[1] = translation[2] = 0.0;
  return 0;
}

int openmc_mesh_filter_set_translation(int32_t index, double translation[3])
{
  if (int err = check_filter(index))
    return err;
  if (translation[0] != 0.0 || translation[1] != 0.0 || translation[2] != 0.0)
    return unsupported("mesh filter translations");
  return 0;
}

int openmc_mesh_filter_get_rotation(int32_t index, double rot[], size_t* n)
{
  (void)rot;
  if (int err = check_filter(index))
    return err;
  if (T.filters[index]->type != "mesh")
    return fail(OPENMC_E_INVALID_TYPE, "Tried to get a rotation from a non-mesh-based filter.");
  *n = 0;
  return 0;
}

int openmc_mesh_filter_set_rotation(int32_t index, const double rot[], size_t rot_len)
{
  (void)rot;
  if (int err = check_filter(index))
    return err;
  if (rot_len != 0)
    return unsupported("mesh filter rotations");
  return 0;
}

int openmc_energy_filter_get_bins(int32_t index, const double** energies, size_t* n)
{
  if (int err = check_filter(index))
    return err;
  const HostFilter& f = *T.filters[index];
  if (f.type != "energy")
    return fail(OPENMC_E_INVALID_TYPE, "Tried to get energy bins on a non-energy filter.");
  *energies = f.bins.data();
  *n = f.bins.size();
  return 0;
}

int openmc_energy_filter_set_bins(int32_t index, size_t n, const double* energies)
{
  (void)n, (void)energies;
  if (int err = check_filter(index))
    return err;
  return unsupported("changing energy filter bins after openmc_init");
}

// Filter types the path does not have: the reference answers with
// OPENMC_E_INVALID_TYPE when the filter at the index is of another type.
int openmc_cell_filter_get_bins(int32_t index, const int32_t** cells, int32_t* n)
{
  (void)cells, (void)n;
  return wrong_filter(index, "cells");
}
int openmc_material_filter_get_bins(int32_t index, const int32_t** bins, size_t* n)
{
  (void)bins, (void)n;
  return wrong_filter(index, "material");
}
int openmc_material_filter_set_bins(int32_t index, size_t n, const int32_t* bins)
{
  (void)bins, (void)n;
  return wrong_filter(index, "material");
}
int openmc_particle_filter_get_bins(int32_t index, int bins[])
{
  (void)bins;
  return wrong_filter(index, "particle");
}
int openmc_energyfunc_filter_get_energy(int32_t index, size_t* n, const double** energy)
{
  (void)n, (void)energy;
  return wrong_filter(index, "energyfunction");
}
int openmc_energyfunc_filter_get_y(int32_t index, size_t* n, const double** y)
{
  (void)n, (void)y;
  return wrong_filter(index, "energyfunction");
}
int openmc_energyfunc_filter_set_data(int32_t index, size_t n, const double* energies, const double* y)
{
  (void)n, (void)energies, (void)y;
  return wrong_filter(index, "energyfunction");
}
int openmc_energyfunc_filter_get_interpolation(int32_t index, int* interp)
{
  (void)interp;
  return wrong_filter(index, "energyfunction");
}
int openmc_energyfunc_filter_set_interpolation(int32_t index, const char* interp)
{
  (void)interp;
  return wrong_filter(index, "energyfunction");
}
int openmc_legendre_filter_get_order(int32_t index, int* order)
{
  (void)order;
  return wrong_filter(index, "Legendre");
}
int openmc_legendre_filter_set_order(int32_t index, int order)
{
  (void)order;
  return wrong_filter(index, "Legendre");
}
int openmc_spatial_legendre_filter_get_order(int32_t index, int* order)
{
  (void)order;
  return wrong_filter(index, "spatial Legendre");
}
int openmc_spatial_legendre_filter_set_order(int32_t index, int order)
{
  (void)order;
  return wrong_filter(index, "spatial Legendre");
}
int openmc_sphharm_filter_get_order(int32_t index, int* order)
{
  (void)order;
  return wrong_filter(index, "spherical harmonics");
}
int openmc_sphharm_filter_set_order(int32_t index, int order)
{
  (void)order;
  return wrong_filter(index, "spherical harmonics");
}
int openmc_zernike_filter_get_order(int32_t index, int* order)
{
  (void)order;
  return wrong_filter(index, "Zernike");
}
int openmc_zernike_filter_set_order(int32_t index, int order)
{
  (void)order;
  return wrong_filter(index, "Zernike");
}
int openmc_meshborn_filter_get_mesh(int32_t index, int32_t* index_mesh)
{
  (void)index_mesh;
  return wrong_filter(index, "meshborn");
}
int openmc_meshborn_filter_set_mesh(int32_t index, int32_t index_mesh)
{
  (void)index_mesh;
  return wrong_filter(index, "meshborn");
}
int openmc_meshborn_filter_get_translation(int32_t index, double translation[3])
{
  (void)translation;
  return wrong_filter(index, "meshborn");
}
int openmc_meshborn_filter_set_translation(int32_t index, double translation[3])
{
  (void)translation;
  return wrong_filter(index, "meshborn");
}
int openmc_meshsurface_filter_get_mesh(int32_t index, int32_t* index_mesh)
{
  (void)index_mesh;
  return wrong_filter(index, "meshsurface");
}
int openmc_meshsurface_filter_set_mesh(int32_t index, int32_t index_mesh)
{
  (void)index_mesh;
  return wrong_filter(index, "meshsurface");
}
int openmc_meshsurface_filter_get_translation(int32_t index, double translation[3])
{
  (void)translation;
  return wrong_filter(index, "meshsurface");
}
int openmc_meshsurface_filter_set_translation(int32_t index, double translation[3])
{
  (void)translation;
  return wrong_filter(index, "meshsurface");
}

// ---------------------------------------------------------------------------
// Tallies (src/tallies/tally.cpp:1300-1590)
// ---------------------------------------------------------------------------
int openmc_tally_get_filters(int32_t index, const int32_t** indices, size_t* n)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *indices = T.tally_filters[index].data();
  *n = T.tally_filters[index].size();
  return 0;
}

int openmc_tally_get_scores(int32_t index, int** scores, int* n)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *scores = T.score_flux.data();
  *n = 1;
  return 0;
}

int openmc_tally_get_nuclides(int32_t index, int** nuclides, int* n)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *nuclides = T.nuclide_total.data();
  *n = 1;
  return 0;
}

int openmc_tally_get_type(int32_t index, int32_t* type)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *type = 0; // TallyType::VOLUME (include/openmc/constants.h:297)
  return 0;
}

int openmc_tally_get_writable(int32_t index, bool* writable)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *writable = true;
  return 0;
}

int openmc_tally_get_multiply_density(int32_t index, bool* value)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  *value = true;
  return 0;
}

int openmc_tally_reset(int32_t index)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  return api([&]() -> int {
    api_tally_reset(index);
    return 0;
  });
}

int openmc_tally_set_id(int32_t index, int32_t id)
{
  HostModel* m = const_cast<HostModel*>(model_or_null());
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  for (size_t i = 0; i < m->tallies.size(); ++i)
    if (m->tallies[i].id == id && static_cast<int32_t>(i) != index)
      return fail(OPENMC_E_INVALID_ID, "Two or more tallies use the same unique ID: " + std::to_string(id));
  m->tallies[index].id = id;
  return 0;
}

int openmc_tally_set_active(int32_t index, bool active)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (!active)
    return unsupported("deactivating a tally (the GPU scores every tally in every active batch)");
  return 0;
}

int openmc_tally_set_writable(int32_t index, bool writable)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  (void)writable;
  return 0;
}

int openmc_tally_set_multiply_density(int32_t index, bool value)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (!value)
    return unsupported("flux tallies without the density multiplier");
  return 0;
}

int openmc_tally_set_estimator(int32_t index, const char* estimator)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (!estimator || std::strcmp(estimator, "tracklength") != 0)
    return unsupported("estimators other than tracklength");
  return 0;
}

int openmc_tally_set_type(int32_t index, const char* type)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (!type || std::strcmp(type, "volume") != 0)
    return unsupported("surface and pulse-height tallies");
  return 0;
}

int openmc_tally_set_filters(int32_t index, size_t n, const int32_t* indices)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  const auto& cur = T.tally_filters[index];
  if (n == cur.size() && std::equal(cur.begin(), cur.end(), indices))
    return 0;
  return unsupported("changing a tally's filters after openmc_init (the tally is laid out on the GPU)");
}

int openmc_tally_set_scores(int32_t index, int n, const char** scores)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (n == 1 && scores && scores[0] && std::strcmp(scores[0], "flux") == 0)
    return 0;
  return unsupported("scores other than flux");
}

int openmc_tally_set_nuclides(int32_t index, int n, const char** nuclides)
{
  const HostModel* m = model_or_null();
  CHECK_INDEX(index, m ? m->tallies.size() : 0, "tallies");
  if (n == 0 || (n == 1 && nuclides && nuclides[0] && std::strcmp(nuclides[0], "total") == 0))
    return 0;
  return unsupported("per-nuclide tallies");
}

int openmc_extend_tallies(int32_t n, int32_t* index_start, int32_t* index_end)
{
  (void)n, (void)index_start, (void)index_end;
  return unsupported("creating tallies after openmc_init (the tally is laid out on the GPU)");
}

int openmc_remove_tally(int32_t index)
{
  (void)index;
  return unsupported("removing tallies after openmc_init");
}

// ---------------------------------------------------------------------------
// Weight windows (src/weight_windows.cpp:1000-1400): the model's neutron
// sets in definition order
// ---------------------------------------------------------------------------
size_t openmc_weight_windows_size(void)
{
  const HostModel* m = model_or_null();
  return m ? m->ww.sets.size() : 0;
}

int openmc_get_weight_windows_index(int32_t id, int32_t* idx)
{
  const HostModel* m = model_or_null();
  if (m) {
    for (size_t i = 0; i < m->ww.sets.size(); ++i) {
      if (m->ww.sets[i].id == id) {
        *idx = static_cast<int32_t>(i);
        return 0;
      }
    }
  }
  return fail(OPENMC_E_INVALID_ID, "No weight windows exist with ID=" + std::to_string(id) + ".");
}

int openmc_weight_windows_get_id(int32_t index, int32_t* id)
{
  if (int err = check_ww(index))
    return err;
  *id = ww_set(index).id;
  return 0;
}

int openmc_weight_windows_get_mesh(int32_t index, int32_t* mesh_idx)
{
  if (int err = check_ww(index))
    return err;
  *mesh_idx = T.ww_mesh_index[static_cast<size_t>(index)];
  return 0;
}

int openmc_weight_windows_get_energy_bounds(int32_t index, const double** e_bounds, size_t* e_bounds_size)
{
  if (int err = check_ww(index))
    return err;
  *e_bounds = ww_set(index).ebounds.data();
  *e_bounds_size = ww_set(index).ebounds.size();
  return 0;
}

int openmc_weight_windows_get_particle(int32_t index, int* particle)
{
  if (int err = check_ww(index))
    return err;
  *particle = 0; // ParticleType::neutron
  return 0;
}

int openmc_weight_windows_get_bounds(int32_t index, const double** lower_bounds, const double** upper_bounds,
  size_t* size)
{
  if (int err = check_ww(index))
    return err;
  *lower_bounds = ww_set(index).lower.data();
  *upper_bounds = ww_set(index).upper.data();
  *size = ww_set(index).lower.size();
  return 0;
}

int openmc_weight_windows_get_survival_ratio(int32_t index, double* ratio)
{
  if (int err = check_ww(index))
    return err;
  *ratio = ww_set(index).survival_ratio;
  return 0;
}

int openmc_weight_windows_get_max_lower_bound_ratio(int32_t index, double* lb_ratio)
{
  if (int err = check_ww(index))
    return err;
  *lb_ratio = ww_set(index).max_lb_ratio;
  return 0;
}

int openmc_weight_windows_get_weight_cutoff(int32_t index, double* cutoff)
{
  if (int err = check_ww(index))
    return err;
  *cutoff = ww_set(index).weight_cutoff;
  return 0;
}

int openmc_weight_windows_get_max_split(int32_t index, int* max_split)
{
  if (int err = check_ww(index))
    return err;
  *max_split = ww_set(index).max_split;
  return 0;
}

#define WW_SETTER(name, ...)                                                                                          \
  int name(int32_t index, __VA_ARGS__)                                                                                \
  {                                                                                                                   \
    if (int err = check_ww(index))                                                                                    \
      return err;                                                                                                     \
    return unsupported("changing weight windows after openmc_init (they are uploaded to the GPU)");                  \
  }

WW_SETTER(openmc_weight_windows_set_id, int32_t)
WW_SETTER(openmc_weight_windows_set_mesh, int32_t)
WW_SETTER(openmc_weight_windows_set_energy_bounds, double*, size_t)
WW_SETTER(openmc_weight_windows_set_particle, int)
WW_SETTER(openmc_weight_windows_set_bounds, const double*, const double*, size_t)
WW_SETTER(openmc_weight_windows_set_survival_ratio, double)
WW_SETTER(openmc_weight_windows_set_max_lower_bound_ratio, double)
WW_SETTER(openmc_weight_windows_set_weight_cutoff, double)
WW_SETTER(openmc_weight_windows_set_max_split, int)
#undef WW_SETTER

int openmc_weight_windows_update_magic(int32_t ww_idx, int32_t tally_idx, const char* value, double threshold,
  double ratio)
{
  (void)ww_idx, (void)tally_idx, (void)value, (void)threshold, (void)ratio;
  return unsupported("the MAGIC weight-window update");
}

int openmc_extend_weight_windows(int32_t n, int32_t* index_start, int32_t* index_end)
{
  (void)n, (void)index_start, (void)index_end;
  return unsupported("creating weight windows after openmc_init");
}

int openmc_weight_windows_export(const char* filename)
{
  (void)filename;
  return unsupported("openmc_weight_windows_export");
}

int openmc_weight_windows_import(const char* filename)
{
  (void)filename;
  return unsupported("openmc_weight_windows_import");
}

// ---------------------------------------------------------------------------
// Subsystems outside the fixed-source transport path
// ---------------------------------------------------------------------------
int openmc_calculate_volumes(void)
{
  return unsupported("openmc_calculate_volumes (stochastic volume calculation)");
}

int openmc_plot_geometry(void)
{
  return unsupported("openmc_plot_geometry (plotting)");
}

int openmc_id_map(const void* slice, int32_t* data_out)
{
  (void)slice, (void)data_out;
  return unsupported("openmc_id_map (plotting)");
}

int openmc_property_map(const void* slice, double* data_out)
{
  (void)slice, (void)data_out;
  return unsupported("openmc_property_map (plotting)");
}

int openmc_properties_export(const char* filename)
{
  (void)filename;
  return unsupported("openmc_properties_export");
}

int openmc_properties_import(const char* filename)
{
  (void)filename;
  return unsupported("openmc_properties_import");
}

int openmc_get_keff(double k_combined[])
{
  (void)k_combined;
  return unsupported("openmc_get_keff (eigenvalue mode; this is a fixed-source run)");
}

int openmc_source_bank(void** ptr, int64_t* n)
{
  (void)ptr, (void)n;
  return unsupported("openmc_source_bank (fixed-source particles are sampled on the GPU and kept in no bank)");
}

int openmc_sample_external_source(size_t n, uint64_t* seed, void* sites)
{
  (void)n, (void)seed, (void)sites;
  return unsupported("openmc_sample_external_source (sources are sampled on the GPU inside each batch)");
}

int openmc_statepoint_load(const char* filename)
{
  (void)filename;
  return unsupported("openmc_statepoint_load (restart runs)");
}

void openmc_cmfd_reweight(const bool feedback, const double* cmfd_src)
{
  (void)feedback, (void)cmfd_src;
  unsupported("CMFD");
}

void openmc_initialize_mesh_egrid(const int meshtally_id, const int* cmfd_indices, const double norm)
{
  (void)meshtally_id, (void)cmfd_indices, (void)norm;
  unsupported("CMFD");
}

void openmc_initialize_linsolver(const int* indptr, int len_indptr, const int* indices, int n_elements, int dim,
  double spectral, const int* map, bool use_all_threads)
{
  (void)indptr, (void)len_indptr, (void)indices, (void)n_elements, (void)dim, (void)spectral, (void)map,
    (void)use_all_threads;
  unsupported("CMFD");
}

int openmc_run_linsolver(const double* A_data, const double* b, double* x, double tol)
{
  (void)A_data, (void)b, (void)x, (void)tol;
  return unsupported("CMFD");
}

// Zernike polynomials for openmc.lib.math (src/math_functions.cpp:648-750):
// zn[] holds (n, m) in the order (0,0), (1,-1), (1,1), (2,-2), (2,0), (2,2), ...
// with R_n^|m|(rho) times sin(|m| phi) for m < 0, cos(m phi) for m > 0.
void calc_zn(int n, double rho, double phi, double zn[])
{
  int i = 0;
  for (int p = 0; p <= n; ++p) {
    for (int q = -p; q <= p; q += 2) {
      const double r = zernike_radial(p, q, rho);
      if (q < 0)
        zn[i] = r * std::sin(-q * phi);
      else if (q == 0)
        zn[i] = r;
      else
        zn[i] = r * std::cos(q * phi);
      ++i;
    }
  }
}

// Even orders of the m = 0 radial polynomials: zn_rad[k] = R_{2k}^0(rho)
void calc_zn_rad(int n, double rho, double zn_rad[])
{
  for (int p = 0; p <= n; p += 2)
    zn_rad[p / 2] = zernike_radial(p, 0, rho);
}

} // extern "C"
