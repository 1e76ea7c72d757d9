// Host-side model: the result of openmc_init reading model.xml, the
// cross-section library and the nuclide files, flattened into the arrays the
// device copies of DevProblem point at.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "dev_types.h"

namespace omc {

struct HostNuclide {
  std::string name;
  std::string path; // the data file it was read from
  double awr = 0;
  double kT = 0;
  std::vector<double> energy, total, absorption, elastic;
  // Set when a problem of a problem set names a nuclide the set has already
  // loaded from the same file: the index of that nuclide in the pool, whose
  // data this entry does not hold.
  int pooled = -1;
};

struct HostTally {
  int id = 0;
  std::string name;
  bool has_mesh = false;
  int mesh_id = 0;
  int n_dim = 3;
  int shape[3] = {1, 1, 1};
  double ll[3] = {0, 0, 0}, ur[3] = {0, 0, 0}, width[3] = {0, 0, 0};
  std::vector<double> ebins;
  std::vector<int> filter_ids;
  std::vector<std::string> filter_types;
  std::vector<int> filter_nbins;
  int64_t n_filter_bins = 1;
  std::string estimator = "tracklength";
  int64_t off = 0; // first bin in the concatenated VALUE array of all tallies
  int part = 0;    // problem of a problem set that owns the tally
};

// A RegularMesh (src/mesh.cpp: RegularMesh::set_grid)
struct HostMesh {
  int n_dim = 3;
  int shape[3] = {1, 1, 1};
  double ll[3] = {0, 0, 0}, ur[3] = {0, 0, 0}, width[3] = {0, 0, 0};
  int64_t n_bins() const { return static_cast<int64_t>(shape[0]) * shape[1] * shape[2]; }
};

// A tally filter as defined in the model (all of them, in definition order,
// for the statepoint: src/tallies/filter.cpp, filter_mesh.cpp, filter_energy.cpp)
struct HostFilter {
  int id = 0;
  std::string type; // "mesh" or "energy"
  int mesh_id = 0;  // mesh filter
  std::vector<double> bins; // energy filter
  int n_bins = 0;
};

// One set of mesh weight windows for neutrons (src/weight_windows.cpp:165-229
// reads the element)
struct HostWWSet {
  int id = 0;
  int mesh_id = 0;
  HostMesh mesh;
  std::vector<double> ebounds;      // energy bounds (default [0, inf))
  std::vector<double> lower, upper; // [n_energy][n_mesh_bins]
  double survival_ratio = 3.0;
  double max_lb_ratio = 1.0;
  int max_split = 10;
  double weight_cutoff = 1.0e-38;
};

// Weight windows: every neutron set in definition order (a particle uses the
// first set that has a valid window for it, src/weight_windows.cpp:70-75) and
// the switches (src/settings.cpp:1203-1260)
struct HostWW {
  bool on = false;
  std::vector<HostWWSet> sets;
  bool ckpt_collision = true;
  bool ckpt_surface = false;
  int64_t max_history_splits = 10000000;
};

// One problem of a problem set: several independent models whose particles
// share one event loop.  Everything that differs between the problems and is
// not reached through a cell (materials, surfaces, universes and lattices are
// concatenated, so a cell index names its problem) lives here.
struct HostPart {
  std::string dir;
  int64_t n_particles = 0;
  int64_t first = 0; // first population index of the problem in a batch
  int root_universe = 0;
  int tally = -1; // index into HostModel::tallies
  int mat_begin = 0, surf_begin = 0, cell_begin = 0;
  int n_nuclides = 0;
  double energy_min = 0, energy_max = 0; // of the problem's own nuclides
  DevSource src {};
  std::vector<double> disc_x, disc_p, disc_prob;
  std::vector<int> disc_alias, domain_ids;
  // what the problem's statepoint lists
  std::vector<std::pair<int, HostMesh>> meshes;
  std::vector<HostFilter> filters;
  std::vector<int> statepoint_batches;
  std::string path_output;
};

struct HostModel {
  HostWW ww;
  // settings
  int64_t n_particles = 0;
  int n_batches = 0;
  int n_inactive = 0;
  uint64_t seed = 1;
  uint64_t stride = 152917ULL;
  bool event_based = false;
  int64_t max_in_flight = 0; // 0 = let the GPU size it
  int n_log_bins = 8000;
  int64_t max_events = 1000000;
  double free_gas_threshold = 400.0;
  double energy_cutoff = 0.0;
  double temperature_default = 293.6;  // K (src/settings.cpp), for HDF5 data
  double temperature_tolerance = 10.0; // K
  std::string run_mode = "fixed source";

  std::vector<HostNuclide> nuclides;
  std::unordered_map<std::string, int> nuclide_map;

  std::vector<int> mat_ids;
  std::vector<std::string> mat_names; // the name attribute ("" when absent)
  std::vector<bool> mat_depletable;   // the depletable attribute
  std::vector<int> mat_off;
  std::vector<int> mat_nuc;
  std::vector<double> mat_dens;

  std::vector<int> surf_ids;
  std::vector<DevSurface> surfaces;
  std::vector<DevCell> cells;
  std::vector<std::string> cell_names;
  std::vector<int> tokens;
  std::vector<int> univ_ids;
  std::vector<DevUniverse> universes;
  std::vector<int> univ_cells;
  std::vector<int> lat_ids;
  std::vector<DevLattice> lattices;
  std::vector<int> lat_univ;
  int root_universe = 0;
  int n_levels = 1;

  // source
  DevSource src {};
  std::vector<double> disc_x, disc_p, disc_prob;
  std::vector<int> disc_alias, domain_ids;

  std::vector<HostTally> tallies;
  // every mesh of the model in definition order (settings before tallies)
  std::vector<std::pair<int, HostMesh>> meshes;
  std::vector<HostFilter> filters;
  // <state_point batches> (default: the last batch) and <output><path>
  // (src/settings.cpp:823-843, 1050-1057)
  std::vector<int> statepoint_batches;
  std::string path_output;

  double energy_min = 0, energy_max = 0, log_spacing = 0;
  std::vector<int> grid_index; // [n_nuc * (n_log_bins + 1)]

  // problem set (empty for a single model)
  std::vector<HostPart> parts;
  std::vector<int> mat_nuc_local; // problem-local nuclide index of each mat_nuc entry
};

// Reads <cwd>/model.xml (or the separate settings/materials/geometry/tallies
// files) exactly as openmc_init's read_model_xml does (src/initialize.cpp:360-470).
// With a pool, nuclides the pool already holds from the same file are not
// read again (their entries get `pooled` set and no data).
void read_model(HostModel& m, const std::string& dir, const HostModel* pool = nullptr);

// Reads several independent models (one directory each) into one problem
// set: materials, surfaces, cells, universes, lattices and tallies are
// concatenated with their indices shifted, and each nuclide data file is
// held once.  The problems must agree on what is global to a run (batches,
// seed and stride, the event limit, free-gas threshold, energy cutoff, the
// energy range of their data); weight windows are not supported in a set.
void read_problem_set(HostModel& out, const std::vector<std::string>& dirs);

// initialize_data + Nuclide::init_grid (src/simulation.cpp:682-761, src/nuclide.cpp:502-533)
void initialize_data(HostModel& m);

} // namespace omc
