// Device-resident problem description and particle state.
//
// Everything a transport kernel reads is reached from one DevProblem passed by
// value as a kernel argument; all pointers are device pointers into HBM.
// Particle state is indexed by slot p in [0, cap): one 128-byte hot record
// per particle, one 64-byte record per deeper coordinate level, and the
// per-nuclide microscopic totals of the particle's material, contiguous per
// particle.  Event kernels visit particles in queue order, which is not slot
// order, so each particle's state is kept in as few cache lines as possible.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace omc {

constexpr int MAX_LEVELS = 4;
constexpr int SORT_MAX_BINS = 32768; // (material + 1, energy bucket) keys: a 128 KB LDS histogram
                                     // (gfx950 has 160 KB of LDS per CU)
constexpr int SORT_MAX_BLOCKS = 256; // workgroups of the sort kernels

// Constants of the reference (include/openmc/constants.h:49-113)
constexpr double TINY_BIT = 1e-8;
constexpr double FP_PRECISION = 1e-14;
constexpr double FP_REL_PRECISION = 1e-5;
constexpr double FP_COINCIDENT = 1e-12;
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double INFTY = 1.7976931348623157e308;
constexpr double MASS_NEUTRON = 1.00866491595;
constexpr double MASS_NEUTRON_EV = 939.56542052e6;
constexpr double C_LIGHT = 2.99792458e10;
constexpr double N_AVOGADRO = 0.602214076;
constexpr double K_BOLTZMANN = 8.617333262e-5;
constexpr int C_NONE = -1;
constexpr int MATERIAL_VOID = -1;
constexpr int SURFACE_NONE = 0;

enum SurfType : int { S_XPLANE = 1, S_YPLANE, S_ZPLANE, S_PLANE, S_ZCYL, S_XCYL, S_YCYL, S_SPHERE };
enum BcType : int { BC_TRANSMIT = 0, BC_VACUUM = 1, BC_REFLECT = 2 };
enum FillType : int { FILL_MATERIAL = 0, FILL_UNIVERSE = 1, FILL_LATTICE = 2 };

struct DevSurface {
  int type;
  int bc;
  double c[4];
};

// One token of a cell's region with its surface inlined, so a region is
// walked with one independent load per token (no token -> surface chain).
struct DevTok {
  DevSurface s;
  int token; // signed surface index + 1 (the region token)
  int pad;
};
static_assert(sizeof(DevTok) == 48, "three 16-byte loads per token");

struct DevCell {
  int universe;
  int fill_type;
  int fill;
  int material;
  int tok_begin;
  int tok_end;
  int id;
  int part; // problem of a problem set (0 for a single model)
};

struct DevUniverse {
  int cell_begin; // into univ_cells
  int cell_end;
};

struct DevLattice {
  double ll[3];
  double pitch[3];
  int n[3];
  int is_3d;
  int outer;
  int univ_offset; // into lat_univ
};

struct DevNuclide {
  int64_t off; // first grid point in the concatenated arrays
  int n;       // grid points
  int pad;
  double awr;
  double kT;
  double e_first; // E[0] and E[n-1] of the grid (the below/above-grid tests)
  double e_last;
};

// One nuclide of one material (indexed like mat_nuc): everything a lookup
// needs before its first grid access, in one 64-byte record, so a wavefront
// working on one material reads its nuclides' metadata with one uniform load.
struct alignas(64) DevMatNuc {
  int64_t off; // first grid point of the nuclide
  int n;       // grid points
  int nuc;     // nuclide index
  double e_first;
  double e_last;
  double dens;  // atom density (material order)
  double u_lo;  // learned index: log(E_first / E_min) (0 until built)
  double inv_w; // learned index: seg_n / (log(E_last / E_min) - u_lo)
  int seg_off;  // learned index: first segment of the nuclide in DevLearned::rec
  int seg_n;    // learned index: segments of the nuclide
};
static_assert(sizeof(DevMatNuc) == 64, "one 64-byte record per material nuclide");

// One segment of a learned index as the lookup reads it (one 16-byte load).
struct alignas(16) LSeg {
  float a;  // slope
  float b;  // intercept
  int err;  // error bound of the segment
  int pad;
};

// Learned grid index: piecewise-linear model of i(log E) per nuclide.
// Nuclide n owns segments [seg_off[n], seg_off[n] + seg_n[n]); seg_n[n] is
// the base count or a power-of-two multiple of it, as many as keep every
// recorded error bound within the limit the build enforced.
struct DevLearned {
  const LSeg* rec;     // [all segments] the lookup's copy of seg and err
  const float* seg;    // [all segments][2] slope, intercept (i ~= slope*f + icpt)
  const int* err;      // [all segments] max |i_true - i_pred| over the segment's grid points
  const double* u_lo;  // [n_nuc] log(E_first / E_min) of each nuclide grid
  const double* inv_w; // [n_nuc] seg_n / (log(E_last / E_min) - u_lo)
  const int* seg_off;  // [n_nuc]
  const int* seg_n;    // [n_nuc]
  int n_seg;           // base segment count
  int enabled;
};

struct DevSource {
  int space_type; // 0 point, 1 box
  int angle_type; // 0 isotropic, 1 monodirectional
  int energy_type; // 0 discrete, 1 watt
  int n_discrete;
  double p0[3], p1[3];
  double u_ref[3];
  double watt_a, watt_b;
  const double* disc_x;
  const double* disc_prob;
  const int* disc_alias;
  const int* domain_ids;
  int n_domain;
  int pad;
};

struct DevTally {
  int enabled;
  int n_dim;
  int shape[3];
  int n_energy; // energy bins (0 = no energy filter)
  double ll[3];
  double width[3];
  const double* ebins; // n_energy + 1 edges
  double* value;       // [n_copies][n_bins] VALUE accumulator for the current batch
  int64_t n_bins;
  int n_copies;        // 1, or one per XCD (a power of two)
  int has_mesh;        // filters [mesh] / [mesh, energy]; else [] or [energy]
};

// One set of mesh weight windows (WeightWindows, src/weight_windows.cpp:375-411)
struct DevWWSet {
  int n_dim;
  int shape[3];
  double ll[3];
  double width[3];
  int n_e; // energy groups (bounds n_e + 1)
  int max_split;
  int64_t n_mesh;
  const double* ebounds;
  const double* lower; // [n_e][n_mesh]
  const double* upper;
  double survival_ratio;
  double max_lb_ratio;
  double weight_cutoff;
};

// Mesh weight windows for neutrons (src/weight_windows.cpp:56-146): a
// particle takes the window of the first set that has a valid one for it.
struct DevWW {
  int on;
  int ckpt_collision;
  int ckpt_surface;
  int n_sets;
  int64_t max_history_splits;
  const DevWWSet* sets;
};

// One problem of a problem set: several independent models whose particles
// share one event loop.  A cell index names its problem (DevCell::part);
// what is not reached through a cell is here.
struct DevPart {
  DevSource src;
  DevTally tally;      // value points at the problem's slice (n_bins 0 without a tally)
  int64_t first;       // first population index of the problem in a batch
  int64_t n_particles; // per batch
  int root_universe;
  int mat_begin;  // first material, surface and cell of the problem (history
  int surf_begin; // hashes use problem-local indices so they equal the hashes
  int cell_begin; // of the problem run alone)
};

struct DevProblem {
  DevWW ww;
  // cross sections
  const DevNuclide* nuc;
  const double* E;   // concatenated energy grids (the searches read these)
  const double* B;   // {E, total, absorption} per grid point: the bracket of
                     // interval i is 48 contiguous bytes from B + 3 i
  const double* EL;  // elastic per grid point
  const int* grid_index; // [n_nuc][n_log_bins + 1]
  int n_nuclides;
  int n_log_bins;
  double energy_min, energy_max, log_spacing;
  DevLearned learned;

  // materials
  const int* mat_off; // [n_mat + 1]
  const int* mat_nuc;
  const double* mat_dens;
  const DevMatNuc* matnuc; // [mat_off[n_mat]]
  int n_materials;
  int max_mat_nuc;

  // geometry
  const DevSurface* surf;
  const DevCell* cell;
  const int* tokens;
  const DevTok* ctok; // tokens with their surfaces, same indexing as tokens
  const DevUniverse* univ;
  const int* univ_cells;
  const DevLattice* lat;
  const int* lat_univ;
  int root_universe;
  int n_levels;
  int n_surfaces;
  int pad_geo;

  DevSource src;
  DevTally tally;

  // problem set: n_parts > 1 switches the kernels to per-problem sources,
  // root universes, tallies and leakage (a single model has n_parts == 1)
  const DevPart* part;
  const int* mat_nuc_local; // problem-local nuclide index of each mat_nuc entry
  int n_parts;
  int pad_parts;

  // settings
  int64_t n_particles; // per batch (id arithmetic)
  int64_t max_events;
  uint64_t master_seed;
  uint64_t stride;
  double free_gas_threshold;
  double energy_cutoff;
};

// Counters accumulated by the kernels (device, int64)
enum Counter : int {
  CNT_LOOKUPS = 0,
  CNT_COLLISIONS,
  CNT_SEGMENTS,
  CNT_LOST,
  CNT_XS_EVENTS,
  CNT_ADV_EVENTS,
  CNT_SURF_EVENTS,
  CNT_COLL_EVENTS,
  CNT_SEARCH_PROBES,
  // the part of the counts above done by the history-based tail
  CNT_HIST_XS_EVENTS,
  CNT_HIST_LOOKUPS, // calculate-xs lookups only
  CNT_HIST_ADV_EVENTS,
  CNT_HIST_SEGMENTS,
  CNT_HIST_SURF_EVENTS,
  CNT_HIST_COLL_EVENTS,
  // scatters whose interpolated elastic cross section did not exceed the
  // sampled cutoff (src/physics.cpp:690-735 would leave the elastic branch)
  CNT_SCATTER_FALLTHROUGH,
  CNT_N
};

// Queue counts (device, int32), each written by one partition pass:
//   QC_XS    live particles in q_xs (the calculate-xs and advance queue)
//   QC_STAGE entries staged for the next calculate-xs partition
//   QC_SURF  surface crossings at the front of q_sc
//   QC_COLL  collisions after them (QC_COLL == QC_SURF + 1 is relied on)
//   QC_LIVE  particles still alive after a history-based tail launch
enum QueueCount : int { QC_XS = 0, QC_STAGE, QC_SURF, QC_COLL, QC_LIVE, QC_N };

// Counters and the leakage sum are spread over this many slots (a wavefront
// adds into slot wave_id % CNT_SLOTS) so no address sees more than a few
// atomics per launch; the host sums the slots.
constexpr int CNT_SLOTS = 2048;
// In a problem set each problem's leakage has this many slots of its own.
constexpr int LEAK_PART_SLOTS = 16;

// Hot state of one particle: exactly one 128-byte L2 cache line, so an event
// kernel touching a particle in any queue order moves one line, not one line
// per field.  The direction is shared by every coordinate level (no rotated
// fills on this path: find_cell_inner copies it down unchanged,
// src/geometry.cpp:196-197, and collisions copy it down, src/particle.cpp:396-406).
struct alignas(128) PHot {
  // (xs_t, xs_a are stored together as one 16-byte pair: keep them adjacent
  // at a 16-byte aligned offset)
  double rx, ry, rz; // level-0 position
  double ux, uy, uz; // direction
  double E;
  double wgt;
  double xs_t; // macroscopic total
  double xs_a; // macroscopic absorption
  double time;
  uint64_t seed; // tracking stream
  int cell0;     // level-0 cell (its universe is the root universe)
  int material;
  int material_last;
  int n_coord;
  int surface;   // signed surface index + 1
  int b_surface; // boundary of the last advance
  int b_info;    // coord level | lattice translation << 8
  int n_event;
};
static_assert(sizeof(PHot) == 128, "one cache line per particle");
static_assert(offsetof(PHot, xs_a) == offsetof(PHot, xs_t) + 8 && offsetof(PHot, xs_t) % 16 == 0,
  "xs_t, xs_a form one aligned 16-byte pair");

// Coordinate level >= 1 of one particle (64 bytes).
struct alignas(64) PLevel {
  double rx, ry, rz;
  int cell, univ, lat;
  int lix, liy, liz;
  int pad[4];
};
static_assert(sizeof(PLevel) == 64, "half a cache line per level");

// Weight-window state of one particle's history
// (include/openmc/particle_data.h:507-508, 563-564) and its bank count.
struct alignas(32) PWW {
  double wgt_born;
  double wgt_ww_born;
  double ww_factor;
  int n_split;
  int n_bank;
};
static_assert(sizeof(PWW) == 32, "32-byte weight-window record");

// A secondary bank site (the SourceSite fields Particle::split sets,
// src/particle.cpp:95-112).
struct alignas(16) PSite {
  double rx, ry, rz;
  double ux, uy, uz;
  double E, time, wgt;
  int surface;
  int pad;
};
static_assert(sizeof(PSite) == 80, "80-byte bank site");

struct DevParticles {
  int64_t cap;
  PWW* ww;       // [cap] (weight windows on)
  PSite* bank;   // [cap][bank_cap] LIFO per particle (weight windows on)
  int bank_cap;
  int* status;   // [1] bit 0: a secondary bank overflowed
  double* total_weight; // [CNT_SLOTS] per-batch source weight partial sums
  PHot* hot;      // [cap]
  PLevel* lev;    // [cap][MAX_LEVELS - 1]
  double* micro_t; // [cap][max_mat_nuc]: running sums of density x micro total, for nuclide sampling
  int max_nn;
  int* lost;
  uint64_t* hash;
  double* leak; // [CNT_SLOTS] per-batch leakage partial sums ([n_parts][LEAK_PART_SLOTS] in a problem set)
  // Queues.  No kernel appends to a shared counter: an event kernel writes
  // its result for input i at position i of a staging array (the slot and a
  // key, -1 for a particle that is done) and a partition pass compacts the
  // staging array by key into the next queue.
  int* q_xs;      // [cap] live particles for calculate-xs and advance, sorted by key
  int* adv_key;   // [cap] advance outcome of q_xs[i]: 0 surface, 1 collision, -1 done
  int* q_sc;      // [cap] surface crossings, then collisions
  int* stage_p;   // [cap] slots after surface crossing / collision
  int* stage_key; // [cap] their calculate-xs sort keys, -1 when done
  // partition scratch: sort_count is a [workgroup][nb] matrix of per-
  // workgroup key counts (SORT_MAX_BLOCKS x sort_nb ints), sort_total the
  // key totals then key starts
  int* sort_count;
  int* sort_total;
  int sort_nb;    // calculate-xs keys: 0 (void), 1 + material * sort_eb + energy bucket
  int sort_shift; // energy bucket = (bits(E) >> sort_shift) - sort_base
  int sort_base;
  int sort_eb;    // energy buckets per material
  int* qcount;        // [QC_N]
  unsigned long long* counters; // [CNT_N][CNT_SLOTS]
  int trace_hash;
  int use_learned;
  int wave_search; // plain lookups by wavefront-cooperative search (else per lane)
  int recompute_micro; // collisions look the micro totals up again (no micro_t rows)
};

} // namespace omc
