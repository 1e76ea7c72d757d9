/* C-ABI of libopenmc.so (MI355X build): the subset of OpenMC's C API that the
 * event-based fixed-source transport path and openmc.lib's Python bindings
 * use, plus a few build-specific entry points (prefixed omc_) for measurement
 * and parity checks.  Plain C types only.  Every openmc_* function returns 0 on
 * success or a negative OPENMC_E_* code, with the message in openmc_err_msg,
 * exactly as the reference does (include/openmc/error.h:20-33,
 * src/error.cpp:20-33).
 */
#ifndef OPENMC_AMD_CAPI_H
#define OPENMC_AMD_CAPI_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (src/error.cpp:20-30) */
extern int OPENMC_E_UNASSIGNED;
extern int OPENMC_E_ALLOCATE;
extern int OPENMC_E_OUT_OF_BOUNDS;
extern int OPENMC_E_INVALID_SIZE;
extern int OPENMC_E_INVALID_ARGUMENT;
extern int OPENMC_E_INVALID_TYPE;
extern int OPENMC_E_INVALID_ID;
extern int OPENMC_E_GEOMETRY;
extern int OPENMC_E_DATA;
extern int OPENMC_E_PHYSICS;
extern int OPENMC_E_WARNING;
extern char openmc_err_msg[256];

/* Globals bound by openmc/lib/settings.py:25-40 and openmc/lib/core.py:157 */
extern bool event_based;      /* include/openmc/settings.h:71-72 */
extern int64_t n_particles;
extern int32_t n_inactive;
extern int32_t gen_per_batch;
extern int current_batch;
extern int32_t n_realizations;
extern int run_mode;
extern int verbosity;
extern bool run_CE;
extern int n_coord_levels;
extern bool DAGMC_ENABLED;
extern bool LIBMESH_ENABLED;
extern bool UWUW_ENABLED;
/* openmc/lib/settings.py:27-40, 62: the reference's defaults
 * (src/settings.cpp); weight_windows_on reflects the model and
 * path_statepoint_c names the last statepoint written. */
extern bool cmfd_run;
extern bool entropy_on;
extern int32_t max_lost_particles;
extern bool need_depletion_rx;
extern bool output_summary;
extern double rel_max_lost_particles;
extern bool restart_run;
extern bool weight_windows_on;
extern char* path_statepoint_c;

/* include/openmc/capi.h:83 -- reads model.xml from the cwd (or argv's path);
 * "-e" selects event-based transport; the GPU path runs either way. */
int openmc_init(int argc, char* argv[], const void* intracomm);
/* include/openmc/capi.h:142 -- simulation_init + all batches + finalize */
int openmc_run(void);
/* include/openmc/capi.h:148-149 */
int openmc_simulation_init(void);
int openmc_simulation_finalize(void);
/* include/openmc/capi.h:123 -- status 0 = continue, 1 = done (src/simulation.cpp:270-278) */
int openmc_next_batch(int* status);
/* include/openmc/capi.h:62, 140, 82, 141 */
int openmc_finalize(void);
int openmc_reset(void);
int openmc_hard_reset(void);
int openmc_reset_timers(void);

/* Tallies: include/openmc/capi.h:79, 81, 164-174; src/tallies/tally.cpp:1554-1587.
 * openmc_tally_results hands out a borrowed pointer into the host mirror of
 * the [n_filter_bins][n_scores][3] (VALUE, SUM, SUM_SQ) array, current after
 * every openmc_next_batch and valid until the next batch or finalize. */
int openmc_get_tally_index(int32_t id, int32_t* index);
int openmc_tally_get_id(int32_t index, int32_t* id);
int openmc_tally_get_n_realizations(int32_t index, int32_t* n);
int openmc_tally_get_estimator(int32_t index, int* estimator);
int openmc_tally_get_active(int32_t index, bool* active);
int openmc_tally_results(int32_t index, double** results, size_t* shape);
int openmc_global_tallies(double** ptr);
size_t tallies_size(void);

/* RNG: include/openmc/random_lcg.h:92-113 */
int64_t openmc_get_seed(void);
void openmc_set_seed(int64_t new_seed);
uint64_t openmc_get_stride(void);
void openmc_set_stride(uint64_t new_stride);

/* The rest of include/openmc/capi.h that the modules of openmc/lib bind at import
 * (implemented in capi_model.cpp).  Getters answer from the model loaded by
 * openmc_init with the reference's error codes; functions that would change
 * the model after its upload to the GPU, or that belong to subsystems outside
 * the fixed-source path (volumes, plotting, CMFD, restart, DAGMC,
 * eigenvalue), return OPENMC_E_UNASSIGNED with a message.  Typed getters of
 * filter types the path does not have return OPENMC_E_INVALID_TYPE, as the
 * reference does for a filter of another type. */
/* sizes: src/cell.cpp:1820, src/material.cpp:1599, src/mesh.cpp:3900,
 * src/nuclide.cpp:1105, src/tallies/filter.cpp:62 */
int cells_size(void);
size_t n_materials(void);
int n_meshes(void);
size_t nuclides_size(void);
size_t tally_filters_size(void);
bool openmc_master(void);                                   /* src/message_passing.cpp:16 */
int openmc_get_n_batches(int* n_batches, bool get_max_batches); /* src/settings.cpp:1319 */
int openmc_set_n_batches(int32_t n_batches, bool set_max_batches, bool add_statepoint_batch); /* :1285 */
/* nuclides: include/openmc/capi.h:92,108,118; src/nuclide.cpp:1205 */
int openmc_get_nuclide_index(const char* name, int* index);
int openmc_nuclide_name(int index, const char** name);
int openmc_load_nuclide(const char* name, const double* temps, int n);
int openmc_nuclide_collapse_rate(int index, int MT, double temperature, const double* energy, const double* flux,
  int n, double* xs);
/* materials: include/openmc/capi.h:88,110-127; src/material.cpp:1443 */
int openmc_get_material_index(int32_t id, int32_t* index);
int openmc_material_get_id(int32_t index, int32_t* id);
int openmc_material_set_id(int32_t index, int32_t id);
int openmc_material_get_name(int32_t index, const char** name);
int openmc_material_set_name(int32_t index, const char* name);
int openmc_material_get_densities(int32_t index, const int** nuclides, const double** densities, int* n);
int openmc_material_get_density(int32_t index, double* density);
int openmc_material_get_temperature(int32_t index, double* temperature);
int openmc_material_get_volume(int32_t index, double* volume);
int openmc_material_set_volume(int32_t index, double volume);
int openmc_material_get_depletable(int32_t index, bool* depletable);
int openmc_material_set_depletable(int32_t index, bool depletable);
int openmc_material_get_fissionable(int32_t index, bool* fissionable);
int openmc_material_set_density(int32_t index, double density, const char* units);
int openmc_material_set_densities(int32_t index, int n, const char** name, const double* density);
int openmc_material_add_nuclide(int32_t index, const char* name, double density);
int openmc_extend_materials(int32_t n, int32_t* index_start, int32_t* index_end);
/* cells: include/openmc/capi.h:13-35,62-65,73 */
int openmc_get_cell_index(int32_t id, int32_t* index);
int openmc_cell_get_id(int32_t index, int32_t* id);
int openmc_cell_set_id(int32_t index, int32_t id);
int openmc_cell_get_name(int32_t index, const char** name);
int openmc_cell_set_name(int32_t index, const char* name);
int openmc_cell_get_fill(int32_t index, int* type, int32_t** indices, int32_t* n);
int openmc_cell_get_num_instances(int32_t index, int32_t* num_instances);
int openmc_cell_get_temperature(int32_t index, const int32_t* instance, double* T);
int openmc_cell_get_density(int32_t index, const int32_t* instance, double* rho);
int openmc_cell_get_translation(int32_t index, double xyz[]);
int openmc_cell_get_rotation(int32_t index, double rot[], size_t* n);
int openmc_cell_set_fill(int32_t index, int type, int32_t n, const int32_t* indices);
int openmc_cell_set_temperature(int32_t index, double T, const int32_t* instance, bool set_contained);
int openmc_cell_set_density(int32_t index, double rho, const int32_t* instance, bool set_contained);
int openmc_cell_set_translation(int32_t index, const double xyz[]);
int openmc_cell_set_rotation(int32_t index, const double rot[], size_t rot_len);
int openmc_cell_bounding_box(const int32_t index, double* llc, double* urc);
int openmc_global_bounding_box(double* llc, double* urc);
int openmc_find_cell(const double* xyz, int32_t* index, int32_t* instance);
int openmc_extend_cells(int32_t n, int32_t* index_start, int32_t* index_end);
int openmc_dagmc_universe_get_cell_ids(int32_t univ_id, int32_t* ids, size_t* n);
int openmc_dagmc_universe_get_num_cells(int32_t univ_id, size_t* n);
/* meshes: include/openmc/capi.h:89,130-150; src/mesh.cpp:2285,2410,2450,2670-2700 */
int openmc_get_mesh_index(int32_t id, int32_t* index);
int openmc_mesh_get_id(int32_t index, int32_t* id);
int openmc_mesh_set_id(int32_t index, int32_t id);
int openmc_mesh_get_type(int32_t index, char* type);
int openmc_mesh_get_n_elements(int32_t index, size_t* n);
int openmc_mesh_get_volumes(int32_t index, double* volumes);
int openmc_mesh_bounding_box(int32_t index, double* ll, double* ur);
int openmc_regular_mesh_get_dimension(int32_t index, int** dims, int* n);
int openmc_regular_mesh_get_params(int32_t index, double** ll, double** ur, double** width, int* n);
int openmc_regular_mesh_set_dimension(int32_t index, int n, const int* dims);
int openmc_regular_mesh_set_params(int32_t index, int n, const double* ll, const double* ur, const double* width);
int openmc_rectilinear_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz);
int openmc_rectilinear_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz);
int openmc_cylindrical_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz);
int openmc_cylindrical_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz);
int openmc_spherical_mesh_get_grid(int32_t index, double** grid_x, int* nx, double** grid_y, int* ny,
  double** grid_z, int* nz);
int openmc_spherical_mesh_set_grid(int32_t index, const double* grid_x, const int nx, const double* grid_y,
  const int ny, const double* grid_z, const int nz);
int openmc_mesh_get_plot_bins(int32_t index, const void* origin, const void* width, int basis, int* pixels,
  int32_t* data);
int openmc_mesh_material_volumes(int32_t index, int nx, int ny, int nz, int max_mats, int32_t* materials,
  double* volumes);
int openmc_extend_meshes(int32_t n, const char* type, int32_t* index_start, int32_t* index_end);
/* filters: include/openmc/capi.h:36-47,56-59,86,103-106,128-134,151-159,
 * 244-249; src/tallies/filter_mesh.cpp:253-275, filter_meshborn.cpp:37-55,
 * filter_meshsurface.cpp:103-109 */
int openmc_get_filter_index(int32_t id, int32_t* index);
int openmc_filter_get_id(int32_t index, int32_t* id);
int openmc_filter_set_id(int32_t index, int32_t id);
int openmc_filter_get_type(int32_t index, char* type);
int openmc_filter_get_num_bins(int32_t index, int* n_bins);
int openmc_new_filter(const char* type, int32_t* index);
int openmc_mesh_filter_get_mesh(int32_t index, int32_t* index_mesh);
int openmc_mesh_filter_set_mesh(int32_t index, int32_t index_mesh);
int openmc_mesh_filter_get_translation(int32_t index, double translation[3]);
int openmc_mesh_filter_set_translation(int32_t index, double translation[3]);
int openmc_mesh_filter_get_rotation(int32_t index, double rot[], size_t* n);
int openmc_mesh_filter_set_rotation(int32_t index, const double rot[], size_t rot_len);
int openmc_energy_filter_get_bins(int32_t index, const double** energies, size_t* n);
int openmc_energy_filter_set_bins(int32_t index, size_t n, const double* energies);
int openmc_cell_filter_get_bins(int32_t index, const int32_t** cells, int32_t* n);
int openmc_material_filter_get_bins(int32_t index, const int32_t** bins, size_t* n);
int openmc_material_filter_set_bins(int32_t index, size_t n, const int32_t* bins);
int openmc_particle_filter_get_bins(int32_t index, int bins[]);
int openmc_energyfunc_filter_get_energy(int32_t index, size_t* n, const double** energy);
int openmc_energyfunc_filter_get_y(int32_t index, size_t* n, const double** y);
int openmc_energyfunc_filter_set_data(int32_t index, size_t n, const double* energies, const double* y);
int openmc_energyfunc_filter_get_interpolation(int32_t index, int* interp);
int openmc_energyfunc_filter_set_interpolation(int32_t index, const char* interp);
int openmc_legendre_filter_get_order(int32_t index, int* order);
int openmc_legendre_filter_set_order(int32_t index, int order);
int openmc_spatial_legendre_filter_get_order(int32_t index, int* order);
int openmc_spatial_legendre_filter_set_order(int32_t index, int order);
int openmc_sphharm_filter_get_order(int32_t index, int* order);
int openmc_sphharm_filter_set_order(int32_t index, int order);
int openmc_zernike_filter_get_order(int32_t index, int* order);
int openmc_zernike_filter_set_order(int32_t index, int order);
int openmc_meshborn_filter_get_mesh(int32_t index, int32_t* index_mesh);
int openmc_meshborn_filter_set_mesh(int32_t index, int32_t index_mesh);
int openmc_meshborn_filter_get_translation(int32_t index, double translation[3]);
int openmc_meshborn_filter_set_translation(int32_t index, double translation[3]);
int openmc_meshsurface_filter_get_mesh(int32_t index, int32_t* index_mesh);
int openmc_meshsurface_filter_set_mesh(int32_t index, int32_t index_mesh);
int openmc_meshsurface_filter_get_translation(int32_t index, double translation[3]);
int openmc_meshsurface_filter_set_translation(int32_t index, double translation[3]);
/* tallies: include/openmc/capi.h:52,133,164-185 */
int openmc_tally_get_filters(int32_t index, const int32_t** indices, size_t* n);
int openmc_tally_get_scores(int32_t index, int** scores, int* n);
int openmc_tally_get_nuclides(int32_t index, int** nuclides, int* n);
int openmc_tally_get_type(int32_t index, int32_t* type);
int openmc_tally_get_writable(int32_t index, bool* writable);
int openmc_tally_get_multiply_density(int32_t index, bool* value);
int openmc_tally_reset(int32_t index);
int openmc_tally_set_id(int32_t index, int32_t id);
int openmc_tally_set_active(int32_t index, bool active);
int openmc_tally_set_writable(int32_t index, bool writable);
int openmc_tally_set_multiply_density(int32_t index, bool value);
int openmc_tally_set_estimator(int32_t index, const char* estimator);
int openmc_tally_set_type(int32_t index, const char* type);
int openmc_tally_set_filters(int32_t index, size_t n, const int32_t* indices);
int openmc_tally_set_scores(int32_t index, int n, const char** scores);
int openmc_tally_set_nuclides(int32_t index, int n, const char** nuclides);
int openmc_extend_tallies(int32_t n, int32_t* index_start, int32_t* index_end);
int openmc_remove_tally(int32_t index);
/* weight windows: include/openmc/capi.h:186-236 (the neutron sets in order) */
size_t openmc_weight_windows_size(void);
int openmc_get_weight_windows_index(int32_t id, int32_t* idx);
int openmc_weight_windows_get_id(int32_t index, int32_t* id);
int openmc_weight_windows_get_mesh(int32_t index, int32_t* mesh_idx);
int openmc_weight_windows_get_energy_bounds(int32_t index, const double** e_bounds, size_t* e_bounds_size);
int openmc_weight_windows_get_particle(int32_t index, int* particle);
int openmc_weight_windows_get_bounds(int32_t index, const double** lower_bounds, const double** upper_bounds,
  size_t* size);
int openmc_weight_windows_get_survival_ratio(int32_t index, double* ratio);
int openmc_weight_windows_get_max_lower_bound_ratio(int32_t index, double* lb_ratio);
int openmc_weight_windows_get_weight_cutoff(int32_t index, double* cutoff);
int openmc_weight_windows_get_max_split(int32_t index, int* max_split);
int openmc_weight_windows_set_id(int32_t index, int32_t id);
int openmc_weight_windows_set_mesh(int32_t index, int32_t mesh_idx);
int openmc_weight_windows_set_energy_bounds(int32_t index, double* e_bounds, size_t e_bounds_size);
int openmc_weight_windows_set_particle(int32_t index, int particle);
int openmc_weight_windows_set_bounds(int32_t index, const double* lower_bounds, const double* upper_bounds,
  size_t size);
int openmc_weight_windows_set_survival_ratio(int32_t index, double ratio);
int openmc_weight_windows_set_max_lower_bound_ratio(int32_t index, double lb_ratio);
int openmc_weight_windows_set_weight_cutoff(int32_t index, double cutoff);
int openmc_weight_windows_set_max_split(int32_t index, int max_split);
int openmc_weight_windows_update_magic(int32_t ww_idx, int32_t tally_idx, const char* value, double threshold,
  double ratio);
int openmc_extend_weight_windows(int32_t n, int32_t* index_start, int32_t* index_end);
int openmc_weight_windows_export(const char* filename);
int openmc_weight_windows_import(const char* filename);
/* outside the path: include/openmc/capi.h:12,71,72,79,137-139,143,147,158,
 * 250-276 */
int openmc_calculate_volumes(void);
int openmc_plot_geometry(void);
int openmc_id_map(const void* slice, int32_t* data_out);
int openmc_property_map(const void* slice, double* data_out);
int openmc_properties_export(const char* filename);
int openmc_properties_import(const char* filename);
int openmc_get_keff(double k_combined[]);
int openmc_source_bank(void** ptr, int64_t* n);
int openmc_sample_external_source(size_t n, uint64_t* seed, void* sites);
int openmc_statepoint_load(const char* filename);
void openmc_cmfd_reweight(const bool feedback, const double* cmfd_src);
void openmc_initialize_mesh_egrid(const int meshtally_id, const int* cmfd_indices, const double norm);
void openmc_initialize_linsolver(const int* indptr, int len_indptr, const int* indices, int n_elements, int dim,
  double spectral, const int* map, bool use_all_threads);
int openmc_run_linsolver(const double* A_data, const double* b, double* x, double tol);
/* Zernike polynomials (include/openmc/math_functions.h:94,117) */
void calc_zn(int n, double rho, double phi, double zn[]);
void calc_zn_rad(int n, double rho, double zn_rad[]);

/* Build-specific (no counterpart in the reference) ---------------------- */

/* Options: "trace_hash" (per-particle history hash, 0/1), "learned_index"
 * (0/1), "max_in_flight" (particles), "event_timing" (per-kernel HIP event
 * timers, 0/1), "cycles_per_check" (event cycles between host checks),
 * "use_graph" (capture cycles in a hipGraph, 0/1).  Launch shape (none of
 * them changes a history): "blocks" (workgroups of the event kernels, 0 =
 * from the capacity), "block" (threads per workgroup: 64, 128 or 256),
 * "sort" (0/1), "sort_bits" (energy buckets per binade = 2^sort_bits),
 * "sort_every" (full material+energy sort every n cycles, material-only in
 * between), "sort_blocks", "xs_lanes", "xs_group", "bank_cap", "block",
 * "wave_search" (wavefront-cooperative grid search, 0/1), "tally_copies"
 * (1/2/4/8 copies of each tally's VALUE, one per XCD), "recompute_micro"
 * (collisions repeat the nuclide lookups instead of reading per-particle
 * micro-total rows that calculate-xs would write: 1, 0, or -1 = when the
 * largest material has at most 64 nuclides), "fuse_advance" (advance in the
 * calculate-xs launch: 1, 0, or -1 = when collisions repeat the lookups),
 * "progress_s" (a progress line on
 * stderr every this many seconds of a batch, 0 = none). */
int omc_set_option(const char* name, double value);
/* Run statistics of the simulation so far (see omc_stat_names()). */
int omc_get_stats(double* out, int n);
const char* omc_stat_names(void);
/* Per-particle history hashes of the last batch: n_particles values. */
int omc_particle_hashes(const uint64_t** ptr, int64_t* n);
/* Energy grid data after initialization: energy_min, energy_max, log_spacing. */
int omc_energy_grid(double* out3);
int omc_n_nuclides(void);
const int* omc_grid_index(int i_nuc);
/* Batched lookups on the GPU: one nuclide, n energies -> i_grid[n] and
 * (total, absorption)[n]; returns the number of search probes in *probes. */
int omc_lookup_batch(int i_nuc, const double* E, int64_t n, int learned, int* i_grid, double* xs, uint64_t* probes);
/* omc_lookup_batch for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  E[n], i_grid[n] and xs[n][2] are DEVICE pointers the
 * caller owns; the search-probe partial sums use a file-scope device array of
 * 16385 uint64 per GPU (128 KB, held until the process ends), so n is at most
 * 4194304 (16384 workgroups of 256); larger n returns OPENMC_E_INVALID_SIZE.
 * No hipMalloc/hipFree and no allocating library call.  stream: a hipStream_t
 * (null = the library's stream).  probes not null: the call waits for the
 * stream and stores the probe total; null: it returns without waiting.  Every
 * value a call reads from the array was written earlier in the same call, so
 * nothing needs clearing; calls that overlap on one GPU share the array and
 * race. */
int omc_lookup_batch_noalloc(int i_nuc, const double* E, int64_t n, int learned, int* i_grid, double* xs,
  uint64_t* probes, void* stream);
/* Macroscopic (total, absorption) of material index `mat` at n energies; the
 * GPU time of the kernel in *ms. */
int omc_macro_batch(int mat, const double* E, int64_t n, int learned, int repeat, double* macro, double* ms);
/* Reads dir's model as openmc_init does, without a GPU: the model getters
 * (cells, materials, meshes, filters, tallies, weight windows) answer from
 * it and transport calls return OPENMC_E_ALLOCATE; openmc_finalize releases
 * it.  No reference counterpart (model inspection where no GPU is present). */
int omc_load_model_host(const char* dir);
/* Host-only model check (no GPU needed): reads dir/model.xml and its data as
 * openmc_init does and runs initialize_data; writes counts, energy grid data
 * and normalized atom densities to out[n_out] and the first n_gi entries of
 * the concatenated log-grid tables to grid_index_out.  Returns the number of
 * values available in out (>= 0) or an error code. */
int omc_check_model(const char* dir, double* out, int n_out, int* grid_index_out, int64_t n_gi);
/* Nuclear data in the reference's HDF5 layout (openmc.data.IncidentNeutron
 * export_to_hdf5 / Nuclide::Nuclide): cross_sections.xml may name .h5 files,
 * read through libhdf5 opened at run time ($OPENMC_HDF5_LIBRARY, else
 * libhdf5.so, else /opt/conda/lib/libhdf5.so).  The writer stores elastic
 * scattering (MT 2) and capture (MT 102) at one temperature [K]; the reader
 * returns the grid size and fills energy/total/absorption/elastic (n_max
 * points) and out2 = {awr, kT}.  No GPU needed. */
int omc_write_nuclide_hdf5(const char* path, const char* name, double awr, double kT, int temperature,
  const double* energy, const double* elastic, const double* capture, int64_t n);
int omc_read_nuclide(const char* path, const char* name, double temperature, int64_t n_max, double* energy,
  double* total, double* absorption, double* elastic, double* out2);
/* Statepoints in the reference's layout (openmc_statepoint_write,
 * include/openmc/capi.h:161; src/state_point.cpp:36-360): written by
 * openmc_next_batch at the <state_point batches> (default the last batch),
 * without the source bank, as a fixed-source run does.  filename null: the
 * default <output path>statepoint.<batch>.h5; write_source null: true (a
 * fixed-source run keeps no source bank: the dataset is created with one
 * zero site per particle).  omc_problem_statepoint_write writes one problem
 * of a problem set (no reference counterpart).  omc_h5_read reads a dataset
 * (attribute null) or attribute back: kind 0 doubles, 1 int64, 2 raw bytes of
 * fixed-length strings; *n = values (bytes) available, buf filled when cap
 * holds them. */
int openmc_statepoint_write(const char* filename, bool* write_source);
bool openmc_is_statepoint_batch(void);
int omc_problem_statepoint_write(int problem, const char* filename, bool* write_source);
int omc_h5_read(const char* path, const char* object, const char* attribute, int kind, void* buf, size_t cap,
  size_t* n);
/* Streaming copy of n_bytes, `reps` times: calibrates HBM byte counters and
 * the achievable copy bandwidth; *ms = time per copy. */
int omc_copy_kernel(int64_t n_bytes, int reps, double* ms);
/* Elementwise primitives on the GPU for bitwise host/device checks: fn 0 log,
 * 1 sin, 2 cos, 3 sqrt, 4 x/y, 5 prn(seed = bits of x), 6 u64->f64 of bits of x. */
int omc_math_batch(int fn, const double* x, const double* y, int64_t n, double* out);
/* Problem sets: several independent fixed-source models (one directory
 * each, read like openmc_init reads one) whose particles share one event loop
 * (one population per batch, sorted and advanced together).  Each problem
 * keeps its own particle ids, seeds, source, root universe, tally and
 * leakage, so its histories and tallies are those of the problem run alone.
 * The problems must share batches, seed, stride, the event limit and the
 * energy range of their data; weight windows are not supported in a set.
 * After omc_init_problems, openmc_run / openmc_next_batch run every problem;
 * openmc_tally_results(index) reads any problem's tally, and
 * omc_problem_info gives a problem's tally index. */
int omc_init_problems(int n, const char* const* dirs);
int omc_n_problems(void);
int omc_problem_info(int problem, int64_t* n_particles, int64_t* first, int32_t* tally_index);
int omc_problem_global_tallies(int problem, double** ptr);
/* Host-only reading of a problem set (no GPU needed), for tests: counts and
 * per-problem offsets (layout in capi.cpp). */
int omc_check_problem_set(int n, const char* const* dirs, double* out, int n_out);
/* Learned index: fit the per-nuclide piecewise-linear models on the GPU with
 * n_seg segments and `epochs` gradient steps; a nuclide whose recorded index
 * error exceeds the bound (option "learned_err_bound", default and maximum
 * 64) anywhere is fitted again with twice as many segments until it is within
 * it.  *max_err gets the largest recorded error bound over all nuclides. */
int omc_learned_build(int n_seg, int epochs, int* max_err);
/* Loss and gradient of one nuclide's model at given FP32 parameters
 * (params[n_seg][2]); loss[n_seg], grad[n_seg][2].  For gradient checks. */
int omc_learned_loss_grad(int i_nuc, int n_seg, const float* params, double* loss, double* grad);
/* Training summary: base n_seg, max error bound, mean error bound, loss before
 * the first gradient step, after the last one, after recentring. */
int omc_learned_stats(double* out6);
/* More of it: training rounds, the bound enforced, total segments, fewest and
 * most segments of a nuclide. */
int omc_learned_stats2(double* out5);
/* The trained model of one nuclide: returns its segment count n; with params
 * and err not null fills params[n][2] and err[n]. */
int omc_learned_model(int i_nuc, float* params, int* err);

#ifdef __cplusplus
}
#endif

#endif
