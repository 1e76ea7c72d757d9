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
/* Macroscopic (total, absorption) of material index `mat` at n energies; the
 * GPU time of the kernel in *ms. */
int omc_macro_batch(int mat, const double* E, int64_t n, int learned, int repeat, double* macro, double* ms);
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
 * n_seg segments and `epochs` gradient steps; *max_err gets the largest
 * recorded error bound over all nuclides. */
int omc_learned_build(int n_seg, int epochs, int* max_err);
/* Loss and gradient of one nuclide's model at given FP32 parameters
 * (params[n_seg][2]); loss[n_seg], grad[n_seg][2].  For gradient checks. */
int omc_learned_loss_grad(int i_nuc, int n_seg, const float* params, double* loss, double* grad);
/* Training summary: n_seg, max error bound, mean error bound, loss before the
 * first gradient step, after the last one, after recentring. */
int omc_learned_stats(double* out6);
/* The trained model of one nuclide: params[n_seg][2], err[n_seg]; returns n_seg. */
int omc_learned_model(int i_nuc, float* params, int* err);

#ifdef __cplusplus
}
#endif

#endif
