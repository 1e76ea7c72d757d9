// libopenmc.so: the C-ABI driver of the MI355X transport path.
//
// openmc_init reads the model (host_model.cpp), uploads every array the
// kernels read into HBM once, and openmc_next_batch runs one batch entirely
// on the GPU: source sampling, event cycles until the particle population is
// empty, then tally accumulation.  Only queue counts (and, at the end of a
// batch, the leakage scalar) cross PCIe during a batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <stdexcept>
#include <string>
#include <unistd.h>
#include <vector>

#include "capi_internal.h"
#include "hdf5_io.h"
#include "host_model.h"
#include "learned.h"
#include "openmc_amd/capi.h"
#include "transport.h"

// ---------------------------------------------------------------------------
// Exported globals (names as bound by openmc/lib)
// ---------------------------------------------------------------------------
extern "C" {
int OPENMC_E_UNASSIGNED = -1;
int OPENMC_E_ALLOCATE = -2;
int OPENMC_E_OUT_OF_BOUNDS = -3;
int OPENMC_E_INVALID_SIZE = -4;
int OPENMC_E_INVALID_ARGUMENT = -5;
int OPENMC_E_INVALID_TYPE = -6;
int OPENMC_E_INVALID_ID = -7;
int OPENMC_E_GEOMETRY = -8;
int OPENMC_E_DATA = -9;
int OPENMC_E_PHYSICS = -10;
int OPENMC_E_WARNING = 1;
char openmc_err_msg[256] = {0};

bool event_based = false;
int64_t n_particles = 0;
int32_t n_inactive = 0;
int32_t gen_per_batch = 1;
int current_batch = 0;
int32_t n_realizations = 0;
int run_mode = 1;
int verbosity = 7;
bool run_CE = true;
int n_coord_levels = 0;
bool DAGMC_ENABLED = false;
bool LIBMESH_ENABLED = false;
bool UWUW_ENABLED = false;
}

namespace omc {

namespace {

void set_errmsg(const std::string& s)
{
  std::snprintf(openmc_err_msg, sizeof(openmc_err_msg), "%s", s.c_str());
}

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// recompute_micro's automatic choice: up to this many nuclides in the
// largest material
constexpr int RECOMPUTE_MAX_NUC = 64;
constexpr int N_GLOBAL = 4; // K_COLLISION, K_ABSORPTION, K_TRACKLENGTH, LEAKAGE
constexpr int GT_LEAKAGE = 3;

struct KernelTimer {
  std::vector<hipEvent_t> start, stop;
  std::vector<int> kind;
  double ms[KERNEL_KINDS] = {0};
  long long launches[KERNEL_KINDS] = {0};
};

struct Sim {
  HostModel m;
  bool initialized = false;
  bool host_only = false; // omc_load_model_host: the model without a GPU (queries only)
  bool sim_initialized = false;
  DevProblem P {};
  DevParticles S {};
  hipStream_t stream = nullptr;
  std::vector<void*> dev_allocs;
  std::vector<void*> part_allocs;
  int64_t cap = 0;
  int blocks = 0;
  double* d_results = nullptr;
  std::vector<double> h_results; // [all tallies' bins][3]
  int64_t total_bins = 0;
  double* d_value = nullptr; // VALUE of every tally: [n_copies][bins] at HostTally::off * n_copies
  void* value_alloc = nullptr;
  int tally_copies = 1; // copies of each tally's VALUE (up to one per XCD; 1 measured fastest)
  double progress_s = 0; // seconds between progress lines on stderr during a batch (0: none)
  int history_below = 65536; // finish a batch history-based once fewer particles than this are live (0: never)
  int history_steps = 2048;  // event cycles per particle per history-based launch
  long long n_hist_launches = 0;
  int recompute_micro = -1; // collisions repeat the lookups instead of reading micro_t rows (-1: by material size)
  bool wave_search = false; // wavefront-cooperative grid searches (plain lookups; measured 4% slower)
  std::vector<DevPart> h_parts;
  void* parts_alloc = nullptr;
  bool results_dirty = false;
  int n_real = 0;
  std::vector<int> tally_real_base; // n_real when each tally was last reset on its own
  std::vector<double> gt = std::vector<double>(3 * N_GLOBAL, 0.0);
  std::vector<std::vector<double>> gt_part; // global tallies of each problem of a problem set
  double total_weight = 0;
  int* h_qcount = nullptr; // pinned
  std::vector<unsigned long long> h_counter_slots;
  std::vector<double> h_leak_slots;
  std::vector<uint64_t> h_hashes;
  // options
  bool trace_hash = false;
  bool use_learned = false;
  int learned_err_bound = LEARNED_MAX_ERR; // largest index error a learned segment may record
  bool event_timing = false;
  bool use_graph = true;
  int cycles_per_check = 16;
  int xs_lanes_opt = 0; // 0 = from the largest material
  size_t lds_per_block = 65536; // the device's LDS per workgroup (set at init)
  bool sort = true;
  int sort_blocks_opt = 0; // 0 = one workgroup per CU
  int blocks_opt = 0;      // 0 = from the capacity
  int xs_group_opt = -1; // -1: 2 (a lane's interleaved lookups: with 2 the calculate-xs
                       // kernels keep four waves per SIMD; measured faster than 4
                       // with either collision mode)
  int bank_cap = 32; // secondary bank sites per particle (weight windows)
  int sort_bits = -1; // mantissa bits of E in the calculate-xs sort key (-1: 5, or 8 for materials of more
                      // than 64 nuclides, whose 300-lookup events repay a finer order)
  int sort_bits_used = 5;
  int block_opt = 256;  // threads per workgroup of the event kernels
  int sort_every = 1;   // full (material, energy) sort every this many cycles
  int fuse_advance = -1; // advance in the calculate-xs launch (-1: when collisions repeat the lookups)
  LaunchConfig lc {1, 1, 0, 1, 4, 256, 1, 0};
  int64_t max_in_flight_opt = 0;
  std::vector<int64_t> nuc_off; // first grid point of each nuclide
  void* matnuc_alloc = nullptr;
  // graph
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  bool graph_valid = false;
  // stats
  double t_transport_ms = 0, t_batch_wall_s = 0;
  int active_batches = 0;
  // wall-clock timers of the statepoint's runtime group [s]
  std::string path_input; // the model directory given to openmc_init
  std::chrono::steady_clock::time_point t_start = std::chrono::steady_clock::now();
  double t_init_s = 0, t_read_s = 0, t_sim_s = 0, t_active_s = 0, t_tallies_s = 0, t_statepoint_s = 0;
  bool warned_statepoint = false;
  unsigned long long counters[CNT_N] = {0};
  long long n_cycles = 0, n_checks = 0;
  KernelTimer kt;
  Learned learned;
};

Sim* g = nullptr;
long long g_generation = 0; // bumped whenever a model is loaded or released

template<class T>
T* dev_upload(Sim& s, const std::vector<T>& v)
{
  if (v.empty())
    return nullptr;
  void* p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * v.size()), "hipMalloc");
  hip_check(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "hipMemcpy");
  s.dev_allocs.push_back(p);
  return static_cast<T*>(p);
}

template<class T>
T* dev_alloc(std::vector<void*>& list, size_t n)
{
  void* p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc");
  hip_check(hipMemset(p, 0, sizeof(T) * std::max<size_t>(n, 1)), "hipMemset");
  list.push_back(p);
  return static_cast<T*>(p);
}

void free_list(std::vector<void*>& list)
{
  for (void* p : list)
    (void)hipFree(p);
  list.clear();
}

void destroy_graph(Sim& s)
{
  if (s.graph_exec)
    (void)hipGraphExecDestroy(s.graph_exec);
  if (s.graph)
    (void)hipGraphDestroy(s.graph);
  s.graph_exec = nullptr;
  s.graph = nullptr;
  s.graph_valid = false;
}

// Per-(material, nuclide) lookup records, with the learned index's scale
// once it is built.
void upload_matnuc(Sim& s)
{
  const HostModel& m = s.m;
  std::vector<DevMatNuc> rec(m.mat_nuc.size());
  const bool lrn = s.learned.h_ulo.size() == m.nuclides.size() && !m.nuclides.empty();
  for (size_t j = 0; j < m.mat_nuc.size(); ++j) {
    const int k = m.mat_nuc[j];
    const auto& nuc = m.nuclides[k];
    DevMatNuc& r = rec[j];
    r.off = s.nuc_off[k];
    r.n = static_cast<int>(nuc.energy.size());
    r.nuc = k;
    r.e_first = nuc.energy.empty() ? 0.0 : nuc.energy.front();
    r.e_last = nuc.energy.empty() ? 0.0 : nuc.energy.back();
    r.dens = m.mat_dens[j];
    r.u_lo = lrn ? s.learned.h_ulo[k] : 0.0;
    r.inv_w = lrn ? s.learned.h_invw[k] : 0.0;
    r.seg_off = lrn ? s.learned.h_segoff[k] : 0;
    r.seg_n = lrn ? s.learned.h_segn[k] : 1;
  }
  if (s.matnuc_alloc) {
    (void)hipFree(s.matnuc_alloc);
    s.matnuc_alloc = nullptr;
  }
  if (!rec.empty()) {
    hip_check(hipMalloc(&s.matnuc_alloc, sizeof(DevMatNuc) * rec.size()), "hipMalloc");
    hip_check(hipMemcpy(s.matnuc_alloc, rec.data(), sizeof(DevMatNuc) * rec.size(), hipMemcpyHostToDevice),
      "hipMemcpy");
  }
  s.P.matnuc = static_cast<const DevMatNuc*>(s.matnuc_alloc);
}

// (Re)allocates the tallies' VALUE arrays with n_copies copies of their bins
// each (one per XCD: a wavefront adds into its XCD's copy) and points the
// device tally records at them.
void set_tally_copies(Sim& s, int n_copies)
{
  if (s.value_alloc) {
    (void)hipFree(s.value_alloc);
    s.value_alloc = nullptr;
  }
  s.d_value = nullptr;
  s.tally_copies = n_copies;
  if (s.total_bins > 0) {
    hip_check(hipMalloc(&s.value_alloc, sizeof(double) * s.total_bins * n_copies), "hipMalloc");
    hip_check(hipMemset(s.value_alloc, 0, sizeof(double) * s.total_bins * n_copies), "hipMemset");
    s.d_value = static_cast<double*>(s.value_alloc);
  }
  for (size_t k = 0; k < s.h_parts.size(); ++k) {
    const int t = s.m.parts[k].tally;
    if (t >= 0) {
      s.h_parts[k].tally.value = s.d_value + s.m.tallies[t].off * n_copies;
      s.h_parts[k].tally.n_copies = n_copies;
    }
  }
  const int enabled = s.P.tally.enabled;
  s.P.tally = s.h_parts[0].tally;
  s.P.tally.enabled = enabled;
  if (s.parts_alloc)
    (void)hipFree(s.parts_alloc);
  hip_check(hipMalloc(&s.parts_alloc, sizeof(DevPart) * s.h_parts.size()), "hipMalloc");
  hip_check(hipMemcpy(s.parts_alloc, s.h_parts.data(), sizeof(DevPart) * s.h_parts.size(), hipMemcpyHostToDevice),
    "hipMemcpy");
  s.P.part = static_cast<const DevPart*>(s.parts_alloc);
  s.graph_valid = false;
}

void upload_problem(Sim& s)
{
  HostModel& m = s.m;
  DevProblem& P = s.P;
  std::memset(&P, 0, sizeof(P));

  std::vector<DevNuclide> nucs;
  s.nuc_off.clear();
  std::vector<double> E, EL, B;
  int64_t off = 0;
  for (const auto& n : m.nuclides) {
    DevNuclide d {};
    d.off = off;
    d.n = static_cast<int>(n.energy.size());
    d.awr = n.awr;
    d.kT = n.kT;
    d.e_first = n.energy.empty() ? 0.0 : n.energy.front();
    d.e_last = n.energy.empty() ? 0.0 : n.energy.back();
    nucs.push_back(d);
    s.nuc_off.push_back(off);
    off += d.n;
  }
  E.reserve(off);
  EL.reserve(off);
  B.reserve(3 * static_cast<size_t>(off));
  for (const auto& n : m.nuclides) {
    E.insert(E.end(), n.energy.begin(), n.energy.end());
    EL.insert(EL.end(), n.elastic.begin(), n.elastic.end());
    for (size_t k = 0; k < n.energy.size(); ++k) {
      B.push_back(n.energy[k]);
      B.push_back(n.total[k]);
      B.push_back(n.absorption[k]);
    }
  }
  P.nuc = dev_upload(s, nucs);
  P.E = dev_upload(s, E);
  P.B = dev_upload(s, B);
  P.EL = dev_upload(s, EL);
  P.grid_index = dev_upload(s, m.grid_index);
  P.n_nuclides = static_cast<int>(m.nuclides.size());
  P.n_log_bins = m.n_log_bins;
  P.energy_min = m.energy_min;
  P.energy_max = m.energy_max;
  P.log_spacing = m.log_spacing;

  P.mat_off = dev_upload(s, m.mat_off);
  P.mat_nuc = dev_upload(s, m.mat_nuc);
  P.mat_dens = dev_upload(s, m.mat_dens);
  P.n_materials = static_cast<int>(m.mat_ids.size());
  int mx = 0;
  for (size_t k = 0; k + 1 < m.mat_off.size(); ++k)
    mx = std::max(mx, m.mat_off[k + 1] - m.mat_off[k]);
  P.max_mat_nuc = mx;
  upload_matnuc(s);

  P.surf = dev_upload(s, m.surfaces);
  P.cell = dev_upload(s, m.cells);
  P.tokens = dev_upload(s, m.tokens);
  {
    // region tokens with their surfaces inlined (index k matches tokens[k])
    std::vector<DevTok> ctok(m.tokens.size());
    for (size_t k = 0; k < m.tokens.size(); ++k) {
      const int t = m.tokens[k];
      const size_t i = static_cast<size_t>(std::abs(t)) - 1;
      if (t == 0 || i >= m.surfaces.size())
        throw std::runtime_error("region token " + std::to_string(t) + " names no surface");
      ctok[k].s = m.surfaces[i];
      ctok[k].token = t;
      ctok[k].pad = 0;
    }
    P.ctok = dev_upload(s, ctok);
  }
  P.univ = dev_upload(s, m.universes);
  P.univ_cells = dev_upload(s, m.univ_cells);
  P.lat = dev_upload(s, m.lattices);
  P.lat_univ = dev_upload(s, m.lat_univ);
  P.n_levels = m.n_levels;
  P.n_surfaces = static_cast<int>(m.surfaces.size());

  // A single model is a problem set of one problem.
  if (m.parts.empty()) {
    HostPart part;
    part.n_particles = m.n_particles;
    part.first = 0;
    part.root_universe = m.root_universe;
    part.tally = m.tallies.empty() ? -1 : 0;
    part.n_nuclides = static_cast<int>(m.nuclides.size());
    part.src = m.src;
    part.disc_x = m.disc_x;
    part.disc_p = m.disc_p;
    part.disc_prob = m.disc_prob;
    part.disc_alias = m.disc_alias;
    part.domain_ids = m.domain_ids;
    part.dir = s.path_input;
    part.meshes = m.meshes;
    part.filters = m.filters;
    part.statepoint_batches = m.statepoint_batches;
    part.path_output = m.path_output;
    m.parts.push_back(std::move(part));
  }

  // Tallies: one VALUE array for all of them, each at its own offset
  int64_t total_bins = 0;
  for (HostTally& t : m.tallies) {
    t.off = total_bins;
    total_bins += t.n_filter_bins;
  }
  if (total_bins > 0) {
    s.d_results = dev_alloc<double>(s.dev_allocs, 3 * total_bins);
    s.h_results.assign(3 * total_bins, 0.0);
  }
  auto dev_tally = [&](const HostTally& t) {
    DevTally T {};
    T.enabled = 0; // switched on for active batches
    T.n_dim = t.n_dim;
    for (int k = 0; k < 3; ++k) {
      T.shape[k] = t.shape[k];
      T.ll[k] = t.ll[k];
      T.width[k] = t.width[k];
    }
    T.n_energy = t.ebins.empty() ? 0 : static_cast<int>(t.ebins.size()) - 1;
    T.ebins = dev_upload(s, t.ebins);
    T.n_bins = t.n_filter_bins;
    T.has_mesh = t.has_mesh ? 1 : 0;
    T.value = nullptr; // set by set_tally_copies
    T.n_copies = 1;
    return T;
  };

  std::vector<DevPart> parts;
  for (const HostPart& hp : m.parts) {
    DevPart d {};
    d.src = hp.src;
    d.src.disc_x = dev_upload(s, hp.disc_x);
    d.src.disc_prob = dev_upload(s, hp.disc_prob);
    d.src.disc_alias = dev_upload(s, hp.disc_alias);
    d.src.domain_ids = dev_upload(s, hp.domain_ids);
    d.src.n_domain = static_cast<int>(hp.domain_ids.size());
    if (hp.tally >= 0)
      d.tally = dev_tally(m.tallies[hp.tally]);
    d.first = hp.first;
    d.n_particles = hp.n_particles;
    d.root_universe = hp.root_universe;
    d.mat_begin = hp.mat_begin;
    d.surf_begin = hp.surf_begin;
    d.cell_begin = hp.cell_begin;
    parts.push_back(d);
  }
  P.src = parts[0].src;
  P.root_universe = parts[0].root_universe;
  P.n_parts = static_cast<int>(parts.size());
  P.mat_nuc_local = P.n_parts > 1 ? dev_upload(s, m.mat_nuc_local) : nullptr;
  s.total_bins = total_bins;
  s.h_parts = parts;
  set_tally_copies(s, s.tally_copies);

  if (m.ww.on) {
    const HostWW& w = m.ww;
    DevWW& d = P.ww;
    d.on = 1;
    d.ckpt_collision = w.ckpt_collision ? 1 : 0;
    d.ckpt_surface = w.ckpt_surface ? 1 : 0;
    d.max_history_splits = w.max_history_splits;
    std::vector<DevWWSet> sets;
    for (const HostWWSet& ws : w.sets) {
      DevWWSet e {};
      e.n_dim = ws.mesh.n_dim;
      for (int k = 0; k < 3; ++k) {
        e.shape[k] = ws.mesh.shape[k];
        e.ll[k] = ws.mesh.ll[k];
        e.width[k] = ws.mesh.width[k];
      }
      e.n_e = static_cast<int>(ws.ebounds.size()) - 1;
      e.max_split = ws.max_split;
      e.n_mesh = ws.mesh.n_bins();
      e.ebounds = dev_upload(s, ws.ebounds);
      e.lower = dev_upload(s, ws.lower);
      e.upper = dev_upload(s, ws.upper);
      e.survival_ratio = ws.survival_ratio;
      e.max_lb_ratio = ws.max_lb_ratio;
      e.weight_cutoff = ws.weight_cutoff;
      sets.push_back(e);
    }
    d.n_sets = static_cast<int>(sets.size());
    d.sets = dev_upload(s, sets);
  }

  P.n_particles = m.n_particles;
  P.max_events = m.max_events;
  P.master_seed = m.seed;
  P.stride = m.stride;
  P.free_gas_threshold = m.free_gas_threshold;
  P.energy_cutoff = m.energy_cutoff;
  P.learned.enabled = 0;
}

// Whether collisions look the micro totals up again (recompute_micro; -1:
// when the largest material has at most RECOMPUTE_MAX_NUC nuclides).
bool recompute_on(const Sim& s)
{
  return s.recompute_micro < 0 ? s.P.max_mat_nuc <= RECOMPUTE_MAX_NUC : s.recompute_micro != 0;
}

void update_launch_config(Sim& s)
{
  // one lane per particle unless asked otherwise: measured fastest on the
  // 20-nuclide materials of the lattice problem (the lanes of a particle
  // group cost workgroup barriers and an LDS staging area)
  int lanes = s.xs_lanes_opt > 0 ? s.xs_lanes_opt : 1;
  s.lc.blocks = s.blocks_opt > 0 ? s.blocks_opt : s.blocks;
  s.lc.xs_lanes = lanes;
  s.lc.sort = s.sort ? 1 : 0;
  s.lc.sort_blocks = std::min({s.lc.blocks, s.sort_blocks_opt > 0 ? s.sort_blocks_opt : 256, SORT_MAX_BLOCKS});
  // auto: 2 nuclides a lane when calculate-xs writes no rows (1% faster than
  // 4 on the lattice problem, fused or not), else 4
  s.lc.xs_group = s.xs_group_opt > 0 ? s.xs_group_opt : 2;
  s.lc.block = s.block_opt;
  s.lc.sort_every = s.sort_every;
  // auto: fused when the collisions repeat the lookups (then calculate-xs
  // writes no rows and the fused kernel measured 4% faster; with rows, 6% slower)
  s.lc.fuse_advance = s.fuse_advance < 0 ? (recompute_on(s) ? 1 : 0) : (s.fuse_advance ? 1 : 0);
}

// Leakage partial sums: CNT_SLOTS for a single model, LEAK_PART_SLOTS per
// problem in a problem set.
size_t leak_slots(const Sim& s)
{
  return s.P.n_parts > 1 ? static_cast<size_t>(s.P.n_parts) * LEAK_PART_SLOTS : CNT_SLOTS;
}

void free_particles(Sim& s)
{
  destroy_graph(s);
  free_list(s.part_allocs);
  s.cap = 0;
}

void alloc_particles(Sim& s, int64_t cap)
{
  free_particles(s);
  DevParticles& S = s.S;
  std::vector<void*>& L = s.part_allocs;
  S.cap = cap;
  // levels >= 1 only; a one-level model keeps one unused record per particle
  const int64_t deep = std::max(1, s.P.n_levels - 1);
  S.hot = dev_alloc<PHot>(L, cap);
  S.lev = dev_alloc<PLevel>(L, static_cast<size_t>(deep * cap));
  // rows padded to a multiple of four values: stores and loads of the row
  // are 16-byte pairs
  // row stride: a multiple of eight values (64 bytes), so the 64-byte pieces
  // calculate-xs writes at once are aligned to them
  S.max_nn = (std::max(1, s.P.max_mat_nuc) + 7) & ~7;
  S.micro_t = dev_alloc<double>(L, static_cast<size_t>(S.max_nn) * cap);
  S.lost = dev_alloc<int>(L, cap);
  S.hash = dev_alloc<uint64_t>(L, cap);
  S.leak = dev_alloc<double>(L, leak_slots(s));
  S.total_weight = dev_alloc<double>(L, CNT_SLOTS);
  S.status = dev_alloc<int>(L, 1);
  S.bank_cap = s.bank_cap;
  if (s.P.ww.on) {
    S.ww = dev_alloc<PWW>(L, cap);
    S.bank = dev_alloc<PSite>(L, static_cast<size_t>(cap) * s.bank_cap);
  } else {
    S.ww = nullptr;
    S.bank = nullptr;
  }
  S.q_xs = dev_alloc<int>(L, cap);
  S.adv_key = dev_alloc<int>(L, cap);
  S.q_sc = dev_alloc<int>(L, cap);
  S.stage_p = dev_alloc<int>(L, cap);
  S.stage_key = dev_alloc<int>(L, cap);
  S.qcount = dev_alloc<int>(L, QC_N);
  S.counters = dev_alloc<unsigned long long>(L, static_cast<size_t>(CNT_N) * CNT_SLOTS);
  s.h_counter_slots.assign(static_cast<size_t>(CNT_N) * CNT_SLOTS, 0);
  s.h_leak_slots.assign(std::max<size_t>(CNT_SLOTS, leak_slots(s)), 0.0);
  // calculate-xs keys: 0 for void, else 1 + material * eb + energy bucket;
  // the bucket is the exponent and top sort_bits mantissa bits of E, offset
  // by those of energy_min, so eb = (binades of [energy_min, energy_max])
  // x 2^sort_bits; fewer bits when the materials times buckets exceed the
  // LDS histogram
  const int n_mat = std::max(1, s.P.n_materials);
  uint64_t emin_bits, emax_bits;
  std::memcpy(&emin_bits, &s.m.energy_min, 8);
  std::memcpy(&emax_bits, &s.m.energy_max, 8);
  const int binades = std::max(1, static_cast<int>((emax_bits >> 52) - (emin_bits >> 52)) + 1);
  int bits = s.sort_bits >= 0 ? s.sort_bits : (s.P.max_mat_nuc > RECOMPUTE_MAX_NUC ? 8 : 5);
  while (bits > 0 && 1 + static_cast<int64_t>(n_mat) * (binades << bits) > SORT_MAX_BINS)
    --bits;
  int shift = 52 - bits;
  int64_t eb = static_cast<int64_t>(binades) << bits;
  while (1 + n_mat * eb > SORT_MAX_BINS && eb > 1) {
    // more materials than the histogram holds at one bucket per binade:
    // buckets of several binades
    ++shift;
    eb = (eb + 1) / 2;
  }
  S.sort_shift = shift;
  S.sort_eb = static_cast<int>(eb);
  S.sort_base = static_cast<int>(emin_bits >> shift);
  S.sort_nb = static_cast<int>(std::min<int64_t>(1 + n_mat * eb, SORT_MAX_BINS));
  s.sort_bits_used = shift <= 52 ? 52 - shift : 0;
  // partition scratch, shared by the calculate-xs partition (sort_nb keys)
  // and the advance split (two keys)
  const int rows = std::max(S.sort_nb, 2);
  S.sort_count = dev_alloc<int>(L, static_cast<size_t>(rows) * SORT_MAX_BLOCKS);
  S.sort_total = dev_alloc<int>(L, rows);
  s.cap = cap;
  // launch shape: one particle per thread up to 16384 workgroups (64 per CU;
  // measured against 4096, 8192 and 32768 on the lattice problem: +3.6%,
  // +0.8% and +1.5%)
  int64_t b = (cap + 255) / 256;
  s.blocks = static_cast<int>(std::min<int64_t>(std::max<int64_t>(b, 1), 16384));
  update_launch_config(s);
}

int64_t bytes_per_particle(const Sim& s)
{
  const int64_t deep = std::max(1, s.P.n_levels - 1);
  // hot record, deeper levels, nuclide totals, lost flag, hash, five queue
  // and staging arrays, and the weight-window record and bank when on
  const int64_t ww = s.P.ww.on ? static_cast<int64_t>(sizeof(PWW)) + s.bank_cap * sizeof(PSite) : 0;
  return static_cast<int64_t>(sizeof(PHot)) + deep * static_cast<int64_t>(sizeof(PLevel)) +
         ((std::max(1, s.P.max_mat_nuc) + 7) & ~7) * 8 + 4 + 8 + 5 * 4 + ww;
}

void choose_capacity(Sim& s)
{
  int64_t want = s.m.n_particles;
  if (s.max_in_flight_opt > 0)
    want = std::min(want, s.max_in_flight_opt);
  else if (s.m.max_in_flight > 0 && s.m.max_in_flight < want)
    want = s.m.max_in_flight;
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
  int64_t fit = static_cast<int64_t>(free_b / 2) / bytes_per_particle(s);
  want = std::min<int64_t>(want, std::max<int64_t>(fit, 1));
  want = std::min<int64_t>(want, static_cast<int64_t>(1) << 30);
  if (want != s.cap)
    alloc_particles(s, want);
}

void record_kernel(Sim& s, int kind, bool begin)
{
  KernelTimer& kt = s.kt;
  hipEvent_t e;
  hip_check(hipEventCreate(&e), "hipEventCreate");
  hip_check(hipEventRecord(e, s.stream), "hipEventRecord");
  if (begin) {
    kt.start.push_back(e);
    kt.kind.push_back(kind);
  } else {
    kt.stop.push_back(e);
  }
}

void flush_kernel_timers(Sim& s)
{
  KernelTimer& kt = s.kt;
  for (size_t k = 0; k < kt.stop.size(); ++k) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, kt.start[k], kt.stop[k]), "hipEventElapsedTime");
    kt.ms[kt.kind[k]] += ms;
    kt.launches[kt.kind[k]] += 1;
  }
  for (auto e : kt.start)
    (void)hipEventDestroy(e);
  for (auto e : kt.stop)
    (void)hipEventDestroy(e);
  kt.start.clear();
  kt.stop.clear();
  kt.kind.clear();
}

void run_cycles(Sim& s, int n_cycles)
{
  if (s.event_timing) {
    for (int c = 0; c < n_cycles; ++c) {
      for (int k = KERNEL_SORT; k <= KERNEL_COLLISION; ++k) {
        record_kernel(s, k, true);
        launch_kernel(k, s.P, s.S, s.lc, s.stream, c);
        record_kernel(s, k, false);
      }
    }
  } else if (s.use_graph) {
    if (!s.graph_valid) {
      destroy_graph(s);
      hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      for (int c = 0; c < s.cycles_per_check; ++c)
        launch_cycle(s.P, s.S, s.lc, s.stream, c);
      hip_check(hipStreamEndCapture(s.stream, &s.graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&s.graph_exec, s.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      s.graph_valid = true;
    }
    hip_check(hipGraphLaunch(s.graph_exec, s.stream), "hipGraphLaunch");
  } else {
    for (int c = 0; c < n_cycles; ++c)
      launch_cycle(s.P, s.S, s.lc, s.stream, c);
  }
  s.n_cycles += n_cycles;
}

// The tail of a batch, history-based: launches of k_history until no
// particle is left alive (each launch runs every live particle for at most
// history_steps cycles, so launches stay short).
void run_history_tail(Sim& s, std::chrono::steady_clock::time_point& last_note)
{
  const int n_staged = s.h_qcount[QC_STAGE];
  long long rounds = 0;
  while (true) {
    s.h_qcount[QC_LIVE] = 0;
    hip_check(hipMemcpyAsync(s.S.qcount + QC_LIVE, s.h_qcount + QC_LIVE, sizeof(int), hipMemcpyHostToDevice,
                s.stream),
      "hipMemcpyAsync");
    if (s.event_timing)
      record_kernel(s, KERNEL_HISTORY, true);
    launch_history(s.P, s.S, n_staged, s.history_steps, s.stream);
    if (s.event_timing)
      record_kernel(s, KERNEL_HISTORY, false);
    hip_check(hipMemcpyAsync(s.h_qcount + QC_LIVE, s.S.qcount + QC_LIVE, sizeof(int), hipMemcpyDeviceToHost,
                s.stream),
      "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    s.n_hist_launches++;
    ++rounds;
    if (s.event_timing)
      flush_kernel_timers(s);
    if (s.h_qcount[QC_LIVE] == 0)
      break;
    // every launch retires or advances each live particle by history_steps
    // cycles; a history that outlives max_particle_events cycles (without
    // weight windows) means a failure to retire it
    if (!s.P.ww.on && rounds * static_cast<long long>(s.history_steps) > s.m.max_events + 16)
      throw std::runtime_error("history-based tail did not finish within max_particle_events cycles");
    if (s.progress_s > 0) {
      auto now = std::chrono::steady_clock::now();
      if (std::chrono::duration<double>(now - last_note).count() >= s.progress_s) {
        std::fprintf(stderr, "omc: batch %d: history-based tail, %lld launches, %d particles alive\n",
          current_batch, rounds, s.h_qcount[QC_LIVE]);
        std::fflush(stderr);
        last_note = now;
      }
    }
  }
}

void transport_batch(Sim& s)
{
  HostModel& m = s.m;
  const bool active = current_batch > m.n_inactive;
  bool tally_on = active && !m.tallies.empty();
  if (s.P.tally.enabled != (tally_on ? 1 : 0)) {
    s.P.tally.enabled = tally_on ? 1 : 0;
    s.graph_valid = false;
  }
  s.S.trace_hash = s.trace_hash ? 1 : 0;
  s.S.use_learned = (s.use_learned && s.P.learned.enabled) ? 1 : 0;
  s.S.wave_search = s.wave_search ? 1 : 0;
  // auto: look the micro totals up again in the collision when materials are
  // small (measured faster with 20 nuclides a material, slower with 300)
  s.S.recompute_micro = recompute_on(s) ? 1 : 0;
  if (s.trace_hash)
    s.h_hashes.assign(m.n_particles, 0);
  hip_check(hipMemsetAsync(s.S.leak, 0, sizeof(double) * leak_slots(s), s.stream), "hipMemsetAsync");
  hip_check(hipMemsetAsync(s.S.total_weight, 0, sizeof(double) * CNT_SLOTS, s.stream), "hipMemsetAsync");
  hip_check(hipMemsetAsync(s.S.status, 0, sizeof(int), s.stream), "hipMemsetAsync");

  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  auto w0 = std::chrono::steady_clock::now();
  hip_check(hipEventRecord(e0, s.stream), "hipEventRecord");

  int64_t remaining = m.n_particles;
  int64_t offset = 0;
  while (remaining > 0) {
    int64_t n = std::min(remaining, s.cap);
    if (s.event_timing)
      record_kernel(s, KERNEL_INIT, true);
    launch_init(s.P, s.S, n, offset, current_batch, s.stream);
    if (s.event_timing)
      record_kernel(s, KERNEL_INIT, false);
    // k_init staged the n source particles (key -1 for any it lost)
    s.h_qcount[QC_XS] = 0;
    s.h_qcount[QC_STAGE] = static_cast<int>(n);
    s.h_qcount[QC_SURF] = 0;
    s.h_qcount[QC_COLL] = 0;
    hip_check(hipMemcpyAsync(s.S.qcount, s.h_qcount, sizeof(int) * QC_N, hipMemcpyHostToDevice, s.stream),
      "hipMemcpyAsync");
    // A batch ends when no particle is left after the advance of a cycle
    // (nothing to stage for the next one).  Every history ends within
    // max_particle_events events (one per cycle), so more cycles than that
    // means a kernel failed to retire particles.
    // Without weight windows a slot holds one history, so more cycles than
    // max_particle_events means a kernel failed to retire particles.  With
    // them a slot runs its whole split tree (the secondaries are revived in
    // the same slot), which has no such bound.
    long long cycles = 0;
    auto last_note = std::chrono::steady_clock::now();
    while (true) {
      if (!s.P.ww.on && cycles > m.max_events + 16)
        throw std::runtime_error("event loop did not drain within max_particle_events cycles");
      run_cycles(s, s.cycles_per_check);
      cycles += s.cycles_per_check;
      hip_check(hipMemcpyAsync(s.h_qcount, s.S.qcount, sizeof(int) * QC_N, hipMemcpyDeviceToHost, s.stream),
        "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
      s.n_checks++;
      if (s.event_timing)
        flush_kernel_timers(s);
      if (s.h_qcount[QC_STAGE] == 0)
        break;
      // few particles left: finish them history-based (the staged entries
      // of the last cycle are exactly the live particles, each about to
      // start its next calculate-xs event)
      if (s.h_qcount[QC_XS] < s.history_below) {
        run_history_tail(s, last_note);
        break;
      }
      if (s.progress_s > 0) {
        auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(now - last_note).count() >= s.progress_s) {
          std::fprintf(stderr, "omc: batch %d: %lld cycles, %d particles in flight\n", current_batch, cycles,
            s.h_qcount[QC_STAGE]);
          std::fflush(stderr);
          last_note = now;
        }
      }
    }
    if (s.trace_hash)
      hip_check(hipMemcpy(s.h_hashes.data() + offset, s.S.hash, sizeof(uint64_t) * n, hipMemcpyDeviceToHost),
        "hipMemcpy");
    remaining -= n;
    offset += n;
  }
  hip_check(hipEventRecord(e1, s.stream), "hipEventRecord");
  hip_check(hipMemcpyAsync(s.h_leak_slots.data(), s.S.leak, sizeof(double) * leak_slots(s), hipMemcpyDeviceToHost,
              s.stream),
    "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  const bool set = s.P.n_parts > 1;
  std::vector<double> leak_part(m.parts.size(), 0.0);
  double leak = 0.0;
  if (set) {
    for (size_t k = 0; k < m.parts.size(); ++k)
      for (int j = 0; j < LEAK_PART_SLOTS; ++j)
        leak_part[k] += s.h_leak_slots[k * LEAK_PART_SLOTS + j];
  } else {
    for (size_t j = 0; j < CNT_SLOTS; ++j)
      leak += s.h_leak_slots[j];
  }
  // simulation::total_weight: the source weight after the birth weight
  // windows (src/simulation.cpp:637-639)
  hip_check(hipMemcpy(s.h_leak_slots.data(), s.S.total_weight, sizeof(double) * CNT_SLOTS, hipMemcpyDeviceToHost),
    "hipMemcpy");
  double total_weight = 0.0;
  for (size_t j = 0; j < CNT_SLOTS; ++j)
    total_weight += s.h_leak_slots[j];
  int status = 0;
  hip_check(hipMemcpy(&status, s.S.status, sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
  if (status & 1)
    throw std::runtime_error("a secondary bank overflowed its " + std::to_string(s.bank_cap) +
                             " sites (weight-window splitting); raise the option bank_cap");
  auto w1 = std::chrono::steady_clock::now();
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s.t_sim_s += std::chrono::duration<double>(w1 - w0).count();
  if (active) {
    s.t_transport_ms += ms;
    s.t_batch_wall_s += std::chrono::duration<double>(w1 - w0).count();
    s.t_active_s += std::chrono::duration<double>(w1 - w0).count();
    s.active_batches++;
  }
  const auto wt = std::chrono::steady_clock::now();

  s.total_weight = total_weight;

  // accumulate_tallies (src/tallies/tally.cpp:1059-1104), per problem in a
  // problem set (each has its own source weight: no weight windows in a set,
  // so it is the problem's particle count)
  if (active) {
    n_realizations += 1;
    if (set) {
      if (s.gt_part.size() != m.parts.size())
        s.gt_part.assign(m.parts.size(), std::vector<double>(3 * N_GLOBAL, 0.0));
      for (size_t k = 0; k < m.parts.size(); ++k) {
        double val = leak_part[k] / static_cast<double>(m.parts[k].n_particles);
        s.gt_part[k][3 * GT_LEAKAGE + 1] += val;
        s.gt_part[k][3 * GT_LEAKAGE + 2] += val * val;
      }
      s.gt = s.gt_part[0];
    } else {
      double val = leak / s.total_weight;
      s.gt[3 * GT_LEAKAGE + 1] += val;
      s.gt[3 * GT_LEAKAGE + 2] += val * val;
    }
    if (!m.tallies.empty()) {
      double total_source = 1.0;
      for (const HostTally& t : m.tallies) {
        const double n_part = static_cast<double>(set ? m.parts[t.part].n_particles : m.n_particles);
        double norm = total_source / (n_part * 1);
        launch_accumulate(s.d_value + t.off * s.tally_copies, s.d_results + 3 * t.off, t.n_filter_bins,
          s.tally_copies, norm, s.stream);
      }
      s.n_real++;
      s.results_dirty = true;
    }
  }
  s.t_tallies_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - wt).count();
  hip_check(hipMemcpyAsync(s.h_counter_slots.data(), s.S.counters,
              sizeof(unsigned long long) * CNT_N * CNT_SLOTS, hipMemcpyDeviceToHost, s.stream),
    "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
  hip_check(hipGetLastError(), "kernel launch");
  for (int c = 0; c < CNT_N; ++c) {
    unsigned long long sum = 0;
    for (int k = 0; k < CNT_SLOTS; ++k)
      sum += s.h_counter_slots[static_cast<size_t>(c) * CNT_SLOTS + k];
    s.counters[c] = sum;
  }
}

void sync_results(Sim& s)
{
  if (!s.results_dirty || s.h_results.empty())
    return;
  hip_check(hipMemcpy(s.h_results.data(), s.d_results, sizeof(double) * s.h_results.size(), hipMemcpyDeviceToHost),
    "hipMemcpy");
  s.results_dirty = false;
}

// The statepoint of problem `part` (the model itself when there is one), in
// the reference's layout: openmc_statepoint_write (src/state_point.cpp:36-360)
// with meshes_to_hdf5 / Mesh::to_hdf5 / RegularMesh::to_hdf5_inner
// (src/mesh.cpp:578-598, 1412-1418, 3874-3892), Filter::to_statepoint and
// the mesh and energy filters' own (src/tallies/filter_mesh.cpp:72-82,
// filter_energy.cpp:83-87), write_tally_results (src/hdf5_interface.cpp:689-708)
// and the source bank's site type (src/state_point.cpp:557-589).  A fixed-source
// run keeps no source bank, so when asked for it the dataset is created with
// one site per particle and left unwritten, as the reference leaves it.
void write_statepoint(Sim& s, int part, const std::string& filename, bool write_source)
{
  const auto t0 = std::chrono::steady_clock::now();
  const HostModel& m = s.m;
  const HostPart& hp = m.parts.at(static_cast<size_t>(part));
  sync_results(s);
  H5Writer w(filename);
  api_note_statepoint(filename);
  const int64_t f = w.root();
  w.attr_str(f, "filetype", "statepoint");
  const int version[2] = {18, 1};     // VERSION_STATEPOINT (include/openmc/constants.h:28)
  const int openmc_ver[3] = {0, 15, 3}; // the reference release whose layout this is
  w.attr(f, "version", H5Writer::INT, version, 2, true);
  w.attr(f, "openmc_version", H5Writer::INT, openmc_ver, 3, true);
  {
    char ts[64];
    std::time_t now = std::time(nullptr);
    std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    w.attr_str(f, "date_and_time", ts);
  }
  w.attr_str(f, "path", hp.dir);
  const int64_t seed = static_cast<int64_t>(s.P.master_seed);
  const unsigned long stride = static_cast<unsigned long>(s.P.stride);
  w.data(f, "seed", H5Writer::INT64, &seed);
  w.data(f, "stride", H5Writer::SIZE, &stride);
  w.data_str(f, "energy_mode", "continuous-energy");
  w.data_str(f, "run_mode", "fixed source");
  const int8_t no = 0, src_flag = write_source ? 1 : 0;
  w.attr(f, "photon_transport", H5Writer::BOOL, &no);
  const int64_t np = hp.n_particles;
  const int nb = m.n_batches, cb = current_batch;
  w.data(f, "n_particles", H5Writer::INT64, &np);
  w.data(f, "n_batches", H5Writer::INT, &nb);
  w.data(f, "current_batch", H5Writer::INT, &cb);
  w.attr(f, "source_present", H5Writer::BOOL, &src_flag);

  const int64_t tg = w.group(f, "tallies");
  {
    const int64_t mg = w.group(tg, "meshes");
    const int n_meshes = static_cast<int>(hp.meshes.size());
    w.attr(mg, "n_meshes", H5Writer::INT, &n_meshes);
    if (n_meshes > 0) {
      std::vector<int> ids;
      for (const auto& e : hp.meshes) {
        const HostMesh& me = e.second;
        const int64_t g1 = w.group(mg, "mesh " + std::to_string(e.first));
        w.data_str(g1, "type", "regular");
        w.attr(g1, "id", H5Writer::INT, &e.first);
        const size_t nd = static_cast<size_t>(me.n_dim);
        w.data(g1, "dimension", H5Writer::INT, me.shape, {nd});
        w.data(g1, "lower_left", H5Writer::DOUBLE, me.ll, {nd});
        w.data(g1, "upper_right", H5Writer::DOUBLE, me.ur, {nd});
        w.data(g1, "width", H5Writer::DOUBLE, me.width, {nd});
        ids.push_back(e.first);
      }
      w.attr(mg, "ids", H5Writer::INT, ids.data(), ids.size(), true);
    }
  }
  {
    const int64_t fg = w.group(tg, "filters");
    const unsigned long n_filters = hp.filters.size();
    w.attr(fg, "n_filters", H5Writer::SIZE, &n_filters);
    if (n_filters > 0) {
      std::vector<int> ids;
      for (const HostFilter& fl : hp.filters)
        ids.push_back(fl.id);
      w.attr(fg, "ids", H5Writer::INT, ids.data(), ids.size(), true);
      for (const HostFilter& fl : hp.filters) {
        const int64_t g1 = w.group(fg, "filter " + std::to_string(fl.id));
        w.data_str(g1, "type", fl.type);
        w.data(g1, "n_bins", H5Writer::INT, &fl.n_bins);
        if (fl.type == "mesh")
          w.data(g1, "bins", H5Writer::INT, &fl.mesh_id);
        else
          w.data(g1, "bins", H5Writer::DOUBLE, fl.bins.data(), {fl.bins.size()});
      }
    }
  }
  const HostTally* t = hp.tally >= 0 ? &m.tallies[static_cast<size_t>(hp.tally)] : nullptr;
  const unsigned long n_tallies = t ? 1 : 0;
  w.attr(tg, "n_tallies", H5Writer::SIZE, &n_tallies);
  int64_t tally_group = -1;
  if (t) {
    w.attr(tg, "ids", H5Writer::INT, &t->id, 1, true);
    tally_group = w.group(tg, "tally " + std::to_string(t->id));
    w.data_str(tally_group, "name", t->name);
    const int zero = 0, one = 1;
    w.attr(tally_group, "internal", H5Writer::INT, &zero);
    w.attr(tally_group, "multiply_density", H5Writer::INT, &one);
    w.attr(tally_group, "higher_moments", H5Writer::INT, &zero);
    w.data_str(tally_group, "estimator", t->estimator);
    const int n_real = s.n_real;
    w.data(tally_group, "n_realizations", H5Writer::INT, &n_real);
    const unsigned long n_tf = t->filter_ids.size();
    w.data(tally_group, "n_filters", H5Writer::SIZE, &n_tf);
    if (n_tf > 0)
      w.data(tally_group, "filters", H5Writer::INT, t->filter_ids.data(), {t->filter_ids.size()});
    w.data_strs(tally_group, "nuclides", {"total"});
    const unsigned long n_scores = 1;
    w.data(tally_group, "n_score_bins", H5Writer::SIZE, &n_scores);
    w.data_strs(tally_group, "score_bins", {"flux"});
  }
  // global tallies [N_GLOBAL][3] and the results' sum and sum of squares
  const std::vector<double>& gt = m.parts.size() > 1 && !s.gt_part.empty() ? s.gt_part[static_cast<size_t>(part)]
                                                                            : s.gt;
  w.data(f, "global_tallies", H5Writer::DOUBLE, gt.data(), {static_cast<size_t>(N_GLOBAL), 3});
  const bool tallies_active = current_batch > m.n_inactive && !m.tallies.empty();
  const int present = tallies_active ? 1 : 0;
  w.attr(f, "tallies_present", H5Writer::INT, &present);
  if (t && tallies_active) {
    const size_t nbins = static_cast<size_t>(t->n_filter_bins);
    std::vector<double> r(2 * nbins);
    const double* src = s.h_results.data() + 3 * t->off;
    for (size_t b = 0; b < nbins; ++b) {
      r[2 * b] = src[3 * b + 1];
      r[2 * b + 1] = src[3 * b + 2];
    }
    w.data(tally_group, "results", H5Writer::DOUBLE, r.data(), {nbins, 1, 2});
  }
  const int n_real_global = n_realizations;
  w.data(f, "n_realizations", H5Writer::INT, &n_real_global);

  const int64_t rg = w.group(f, "runtime");
  const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - s.t_start).count();
  const struct {
    const char* name;
    double v;
  } rt[] = {{"total initialization", s.t_init_s}, {"reading cross sections", s.t_read_s},
    {"simulation", s.t_sim_s}, {"transport", s.t_sim_s}, {"active batches", s.t_active_s},
    {"accumulating tallies", s.t_tallies_s}, {"total", total}, {"writing statepoints", s.t_statepoint_s}};
  for (const auto& e : rt)
    w.data(rg, e.name, H5Writer::DOUBLE, &e.v);
  if (write_source)
    w.source_bank(f, "source_bank", hp.n_particles);
  s.t_statepoint_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The default statepoint name: <output path>statepoint.<batch, zero-padded to
// the width of n_batches>.h5 (src/state_point.cpp:42-52); the output path of
// a problem of a problem set is relative to its directory.
std::string statepoint_name(const Sim& s, int part)
{
  const HostPart& hp = s.m.parts.at(static_cast<size_t>(part));
  const int width = static_cast<int>(std::to_string(s.m.n_batches).size());
  char num[32];
  std::snprintf(num, sizeof(num), "%0*d", width, current_batch);
  std::string dir = hp.path_output;
  if (s.m.parts.size() > 1 && (dir.empty() || dir[0] != '/'))
    dir = (hp.dir.empty() ? std::string(".") : hp.dir) + "/" + dir;
  return dir + "statepoint." + num + ".h5";
}

// Statepoints due at this batch (finalize_batch, src/simulation.cpp:432-444):
// fixed source, so without the source bank.  Without libhdf5 they are
// skipped with one warning rather than failing the run.
void write_due_statepoints(Sim& s)
{
  for (size_t k = 0; k < s.m.parts.size(); ++k) {
    const auto& b = s.m.parts[k].statepoint_batches;
    if (std::find(b.begin(), b.end(), current_batch) == b.end())
      continue;
    if (!hdf5_available()) {
      if (!s.warned_statepoint)
        std::fprintf(stderr, "omc: warning: no libhdf5, statepoints are not written\n");
      s.warned_statepoint = true;
      return;
    }
    write_statepoint(s, static_cast<int>(k), statepoint_name(s, static_cast<int>(k)), false);
  }
}

// Device set-up shared by openmc_init and omc_init_problems.
void finish_init(Sim& s)
{
  initialize_data(s.m);
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
    throw std::runtime_error("no GPU visible: the MI355X transport path needs a HIP device");
  hip_check(hipSetDevice(0), "hipSetDevice");
  hipDeviceProp_t prop {};
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.sharedMemPerBlock > 0)
    s.lds_per_block = prop.sharedMemPerBlock;
  hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_qcount), sizeof(int) * QC_N, hipHostMallocDefault),
    "hipHostMalloc");
  upload_problem(s);
  if (s.m.event_based)
    event_based = true;
  weight_windows_on = s.m.ww.on;
  n_particles = s.m.n_particles;
  n_inactive = s.m.n_inactive;
  n_coord_levels = s.m.n_levels;
  current_batch = 0;
  n_realizations = 0;
  s.initialized = true;
}

template<class F>
int guarded(F f)
{
  try {
    return f();
  } catch (const HipError& e) {
    set_errmsg(e.what());
    return OPENMC_E_ALLOCATE;
  } catch (const std::exception& e) {
    set_errmsg(e.what());
    return OPENMC_E_INVALID_ARGUMENT;
  }
}

} // namespace

Learned& learned_state()
{
  return g->learned;
}
DevProblem& problem_state()
{
  return g->P;
}
hipStream_t stream_state()
{
  return g->stream;
}
const HostModel& model_state()
{
  return g->m;
}

long long api_generation()
{
  return g_generation;
}

HostModel* api_model()
{
  return g && (g->initialized || g->host_only) ? &g->m : nullptr;
}

bool api_simulation_running()
{
  return g && g->sim_initialized;
}

void api_set_errmsg(const std::string& msg)
{
  set_errmsg(msg);
}

void api_tally_reset(int index)
{
  if (!g || !g->initialized || index < 0 || index >= static_cast<int>(g->m.tallies.size()))
    throw std::runtime_error("Index in tallies array is out of bounds.");
  if (g->h_results.empty())
    return;
  sync_results(*g);
  const HostTally& t = g->m.tallies[index];
  const size_t off = 3 * static_cast<size_t>(t.off), n = 3 * static_cast<size_t>(t.n_filter_bins);
  hip_check(hipMemset(g->d_results + off, 0, sizeof(double) * n), "hipMemset");
  std::fill(g->h_results.begin() + off, g->h_results.begin() + off + n, 0.0);
  g->tally_real_base.resize(g->m.tallies.size(), 0);
  g->tally_real_base[index] = g->n_real;
}

} // namespace omc

using namespace omc;

extern "C" {

int openmc_init(int argc, char* argv[], const void* intracomm)
{
  (void)intracomm;
  return guarded([&]() -> int {
    if (g && g->initialized)
      openmc_finalize();
    std::string dir = ".";
    bool dir_given = false;
    for (int i = 1; i < argc; ++i) {
      std::string a = argv[i] ? argv[i] : "";
      if (a == "-e" || a == "--event")
        event_based = true;
      else if (a == "-s" || a == "--threads")
        ++i; // host threads: the GPU path ignores them
      else if (!a.empty() && a[0] != '-') {
        dir = a;
        dir_given = true;
      }
    }
    delete g;
    g = new Sim();
    ++g_generation;
    g->path_input = dir_given ? dir : "";
    const auto t0 = std::chrono::steady_clock::now();
    read_model(g->m, dir);
    g->t_read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finish_init(*g);
    g->t_init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  });
}

// The model of a directory read as openmc_init reads it, with no GPU: the
// model getters answer from it, transport calls fail.  No reference
// counterpart; for inspecting a model (and the tests) where no GPU is present.
int omc_load_model_host(const char* dir)
{
  return guarded([&]() -> int {
    if (g && g->initialized)
      openmc_finalize();
    delete g;
    g = new Sim();
    ++g_generation;
    g->host_only = true;
    g->path_input = dir ? dir : "";
    try {
      read_model(g->m, dir ? dir : ".");
      initialize_data(g->m);
    } catch (...) {
      delete g;
      g = nullptr;
      throw;
    }
    n_particles = g->m.n_particles;
    n_inactive = g->m.n_inactive;
    n_coord_levels = g->m.n_levels;
    weight_windows_on = g->m.ww.on;
    current_batch = 0;
    n_realizations = 0;
    return 0;
  });
}

// A problem set: several independent models (one directory each) whose
// particles share one event loop.  No reference counterpart (the reference
// runs one model per process); the callers are batch drivers that run many
// small fixed-source problems.
int omc_init_problems(int n, const char* const* dirs)
{
  return guarded([&]() -> int {
    if (g && g->initialized)
      openmc_finalize();
    if (n < 1 || !dirs)
      throw std::runtime_error("a problem set needs at least one directory");
    std::vector<std::string> d;
    for (int k = 0; k < n; ++k)
      d.push_back(dirs[k] ? dirs[k] : ".");
    delete g;
    g = new Sim();
    ++g_generation;
    const auto t0 = std::chrono::steady_clock::now();
    read_problem_set(g->m, d);
    g->t_read_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    finish_init(*g);
    g->t_init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
  });
}


int openmc_simulation_init(void)
{
  if (g && g->host_only) {
    set_errmsg("the model was read without a GPU (omc_load_model_host); transport needs openmc_init on a GPU");
    return OPENMC_E_ALLOCATE;
  }
  if (!g || !g->initialized) {
    set_errmsg("OpenMC has not been initialized.");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    if (g->sim_initialized)
      return 0;
    if (g->P.n_parts > 1) {
      if (n_particles != g->m.n_particles)
        throw std::runtime_error("the particle counts of a problem set come from its problems' settings");
    } else {
      g->m.n_particles = n_particles;
      g->m.parts[0].n_particles = n_particles;
      g->P.n_particles = n_particles;
    }
    choose_capacity(*g);
    // src/simulation.cpp:120-126: batch counter back to zero and
    // openmc_reset(), so a second run under one openmc_init starts from
    // empty SUM/SUM_SQ and zero realizations
    current_batch = 0;
    int err = openmc_reset();
    if (err)
      return err;
    g->sim_initialized = true;
    return 0;
  });
}

int openmc_next_batch(int* status)
{
  if (!g || !g->sim_initialized) {
    set_errmsg("Simulation has not been initialized yet.");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    if (current_batch >= g->m.n_batches) {
      if (status)
        *status = 1;
      return 0;
    }
    ++current_batch;
    transport_batch(*g);
    write_due_statepoints(*g);
    if (status)
      *status = current_batch >= g->m.n_batches ? 1 : 0;
    return 0;
  });
}

int openmc_simulation_finalize(void)
{
  if (!g || !g->sim_initialized)
    return 0;
  return guarded([&]() -> int {
    sync_results(*g);
    g->sim_initialized = false;
    if (verbosity >= 6 && g->t_transport_ms > 0) {
      double rate = static_cast<double>(g->m.n_particles) * g->active_batches / (g->t_transport_ms * 1e-3);
      std::printf(" Calculation Rate (active)         = %g particles/second\n", rate);
    }
    return 0;
  });
}

int openmc_run(void)
{
  int err = openmc_simulation_init();
  if (err)
    return err;
  int status = 0;
  while (status == 0) {
    err = openmc_next_batch(&status);
    if (err)
      return err;
  }
  return openmc_simulation_finalize();
}

int openmc_reset(void)
{
  if (!g)
    return 0;
  return guarded([&]() -> int {
    if (!g->h_results.empty()) {
      hip_check(hipMemset(g->d_results, 0, sizeof(double) * g->h_results.size()), "hipMemset");
      std::fill(g->h_results.begin(), g->h_results.end(), 0.0);
    }
    std::fill(g->gt.begin(), g->gt.end(), 0.0);
    for (auto& v : g->gt_part)
      std::fill(v.begin(), v.end(), 0.0);
    g->n_real = 0;
    g->tally_real_base.assign(g->m.tallies.size(), 0);
    n_realizations = 0;
    g->results_dirty = false;
    return 0;
  });
}

int openmc_hard_reset(void)
{
  int err = openmc_reset();
  if (g) {
    current_batch = 0;
    g->m.seed = 1;
    g->m.stride = 152917ULL;
    g->P.master_seed = 1;
    g->P.stride = 152917ULL;
    g->graph_valid = false;
  }
  return err;
}

int openmc_reset_timers(void)
{
  if (g) {
    g->t_transport_ms = 0;
    g->t_batch_wall_s = 0;
    g->active_batches = 0;
    for (int k = 0; k < KERNEL_KINDS; ++k) {
      g->kt.ms[k] = 0;
      g->kt.launches[k] = 0;
    }
  }
  return 0;
}

int openmc_finalize(void)
{
  if (!g)
    return 0;
  if (g->initialized) {
    (void)hipStreamSynchronize(g->stream);
    free_particles(*g);
    free_learned(g->learned);
    if (g->matnuc_alloc)
      (void)hipFree(g->matnuc_alloc);
    if (g->value_alloc)
      (void)hipFree(g->value_alloc);
    if (g->parts_alloc)
      (void)hipFree(g->parts_alloc);
    free_list(g->dev_allocs);
    if (g->h_qcount)
      (void)hipHostFree(g->h_qcount);
    if (g->stream)
      (void)hipStreamDestroy(g->stream);
  }
  delete g;
  g = nullptr;
  ++g_generation;
  current_batch = 0;
  n_realizations = 0;
  return 0;
}

int openmc_get_tally_index(int32_t id, int32_t* index)
{
  if (!g)
    return OPENMC_E_ALLOCATE;
  for (size_t k = 0; k < g->m.tallies.size(); ++k) {
    if (g->m.tallies[k].id == id) {
      *index = static_cast<int32_t>(k);
      return 0;
    }
  }
  set_errmsg("No tally exists with ID=" + std::to_string(id) + ".");
  return OPENMC_E_INVALID_ID;
}

int openmc_tally_get_id(int32_t index, int32_t* id)
{
  if (!g || index < 0 || index >= static_cast<int32_t>(g->m.tallies.size())) {
    set_errmsg("Index in tallies array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  *id = g->m.tallies[index].id;
  return 0;
}

int openmc_tally_get_n_realizations(int32_t index, int32_t* n)
{
  if (!g || index < 0 || index >= static_cast<int32_t>(g->m.tallies.size())) {
    set_errmsg("Index in tallies array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  *n = g->n_real - (index < static_cast<int32_t>(g->tally_real_base.size()) ? g->tally_real_base[index] : 0);
  return 0;
}

int openmc_tally_get_estimator(int32_t index, int* estimator)
{
  if (!g || index < 0 || index >= static_cast<int32_t>(g->m.tallies.size())) {
    set_errmsg("Index in tallies array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  // TallyEstimator {ANALOG, TRACKLENGTH, COLLISION} (include/openmc/constants.h:299)
  *estimator = 1;
  return 0;
}

int openmc_tally_get_active(int32_t index, bool* active)
{
  if (!g || index < 0 || index >= static_cast<int32_t>(g->m.tallies.size())) {
    set_errmsg("Index in tallies array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  *active = current_batch > g->m.n_inactive;
  return 0;
}

int openmc_tally_results(int32_t index, double** results, size_t* shape)
{
  if (!g || index < 0 || index >= static_cast<int32_t>(g->m.tallies.size())) {
    set_errmsg("Index in tallies array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  if (g->h_results.empty()) {
    set_errmsg("Tally results have not been allocated yet.");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    sync_results(*g);
    *results = g->h_results.data() + 3 * g->m.tallies[index].off;
    shape[0] = static_cast<size_t>(g->m.tallies[index].n_filter_bins);
    shape[1] = 1;
    shape[2] = 3;
    return 0;
  });
}

int openmc_global_tallies(double** ptr)
{
  if (!g)
    return OPENMC_E_ALLOCATE;
  *ptr = g->gt.data();
  return 0;
}

size_t tallies_size(void)
{
  return g ? g->m.tallies.size() : 0;
}

int omc_n_problems(void)
{
  return g && g->initialized ? static_cast<int>(g->m.parts.size()) : 0;
}

int omc_problem_info(int problem, int64_t* n_particles_out, int64_t* first, int32_t* tally_index)
{
  if (!g || !g->initialized || problem < 0 || problem >= static_cast<int>(g->m.parts.size())) {
    set_errmsg("Index in problems array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  const HostPart& p = g->m.parts[problem];
  if (n_particles_out)
    *n_particles_out = p.n_particles;
  if (first)
    *first = p.first;
  if (tally_index)
    *tally_index = p.tally;
  return 0;
}

int omc_problem_global_tallies(int problem, double** ptr)
{
  if (!g || !g->initialized || problem < 0 || problem >= static_cast<int>(g->m.parts.size())) {
    set_errmsg("Index in problems array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  if (g->P.n_parts <= 1) {
    *ptr = g->gt.data();
    return 0;
  }
  if (g->gt_part.size() != g->m.parts.size())
    g->gt_part.assign(g->m.parts.size(), std::vector<double>(3 * N_GLOBAL, 0.0));
  *ptr = g->gt_part[problem].data();
  return 0;
}

int64_t openmc_get_seed(void)
{
  return g ? static_cast<int64_t>(g->P.master_seed) : 1;
}

void openmc_set_seed(int64_t new_seed)
{
  if (g) {
    g->m.seed = static_cast<uint64_t>(new_seed);
    g->P.master_seed = static_cast<uint64_t>(new_seed);
    g->graph_valid = false;
  }
}

uint64_t openmc_get_stride(void)
{
  return g ? g->P.stride : 152917ULL;
}

void openmc_set_stride(uint64_t new_stride)
{
  if (g) {
    g->m.stride = new_stride;
    g->P.stride = new_stride;
    g->graph_valid = false;
  }
}

int omc_set_option(const char* name, double value)
{
  if (!g) {
    set_errmsg("OpenMC has not been initialized.");
    return OPENMC_E_ALLOCATE;
  }
  std::string n = name ? name : "";
  g->graph_valid = false;
  if (n == "trace_hash")
    g->trace_hash = value != 0;
  else if (n == "learned_index")
    g->use_learned = value != 0;
  else if (n == "learned_err_bound") {
    if (value < 0 || value > LEARNED_MAX_ERR) {
      set_errmsg("learned_err_bound must be in [0, " + std::to_string(LEARNED_MAX_ERR) + "]");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    g->learned_err_bound = static_cast<int>(value);
  }
  else if (n == "event_timing")
    g->event_timing = value != 0;
  else if (n == "use_graph")
    g->use_graph = value != 0;
  else if (n == "cycles_per_check")
    g->cycles_per_check = std::max(1, static_cast<int>(value));
  else if (n == "xs_lanes") {
    int v = static_cast<int>(value);
    if (v != 0 && (v < 1 || v > 64 || (v & (v - 1)) != 0)) {
      set_errmsg("xs_lanes must be 0 or a power of two in [1, 64]");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    // several lanes per particle stage the material's micro cross sections
    // in LDS: (256 / lanes) particles x nuclides x 20 bytes per workgroup
    const size_t need = v > 1 ? static_cast<size_t>(256 / v) * std::max(1, g->P.max_mat_nuc) * 20 : 0;
    if (need > g->lds_per_block) {
      set_errmsg("xs_lanes " + std::to_string(v) + " needs " + std::to_string(need) + " bytes of LDS per workgroup "
                 "for materials of " + std::to_string(g->P.max_mat_nuc) + " nuclides; the device has " +
                 std::to_string(g->lds_per_block));
      return OPENMC_E_INVALID_ARGUMENT;
    }
    g->xs_lanes_opt = v;
    update_launch_config(*g);
  } else if (n == "sort") {
    g->sort = value != 0;
    update_launch_config(*g);
  } else if (n == "bank_cap" || n == "sort_bits") {
    int v = static_cast<int>(value);
    if (n == "bank_cap" && (v < 1 || v > 4096)) {
      set_errmsg("bank_cap must be in [1, 4096]");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    if (n == "sort_bits" && (v < -1 || v > 8)) {
      set_errmsg("sort_bits must be in [0, 8] (-1: by the material size)");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    (n == "bank_cap" ? g->bank_cap : g->sort_bits) = v;
    if (g->sim_initialized)
      return guarded([&]() -> int {
        alloc_particles(*g, g->cap);
        return 0;
      });
  } else if (n == "xs_group") {
    int v = static_cast<int>(value);
    if (v != -1 && v != 1 && v != 2 && v != 4 && v != 8) {
      set_errmsg("xs_group must be 1, 2, 4 or 8 (or -1: the default, 2)");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    g->xs_group_opt = v;
    update_launch_config(*g);
  } else if (n == "block") {
    int v = static_cast<int>(value);
    if (v != 64 && v != 128 && v != 256) {
      set_errmsg("block must be 64, 128 or 256 (threads per workgroup of the event kernels; their launch bound "
                 "is 256 so that they keep up to 256 VGPRs)");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    g->block_opt = v;
    update_launch_config(*g);
  } else if (n == "history_below" || n == "history_steps") {
    int v = static_cast<int>(value);
    if (v < 0 || (n == "history_steps" && v < 1)) {
      set_errmsg(n + " out of range");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    (n == "history_below" ? g->history_below : g->history_steps) = v;
  } else if (n == "progress_s") {
    g->progress_s = value;
  } else if (n == "fuse_advance") {
    g->fuse_advance = value < 0 ? -1 : (value != 0 ? 1 : 0);
    update_launch_config(*g);
  } else if (n == "recompute_micro") {
    g->recompute_micro = value < 0 ? -1 : (value != 0 ? 1 : 0);
  } else if (n == "wave_search") {
    g->wave_search = value != 0;
  } else if (n == "tally_copies") {
    int v = static_cast<int>(value);
    if (v != 1 && v != 2 && v != 4 && v != 8) {
      set_errmsg("tally_copies must be 1, 2, 4 or 8");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    if (g->n_real > 0 || current_batch > 0) {
      set_errmsg("tally_copies can only be changed before the first batch");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    return guarded([&]() -> int {
      set_tally_copies(*g, v);
      return 0;
    });
  } else if (n == "sort_every") {
    int v = static_cast<int>(value);
    if (v < 1 || v > 1024) {
      set_errmsg("sort_every must be in [1, 1024]");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    g->sort_every = v;
    update_launch_config(*g);
  } else if (n == "sort_blocks" || n == "blocks") {
    int v = static_cast<int>(value);
    if (v < 0 || v > 65536) {
      set_errmsg(n + " must be in [0, 65536]");
      return OPENMC_E_INVALID_ARGUMENT;
    }
    (n == "blocks" ? g->blocks_opt : g->sort_blocks_opt) = v;
    update_launch_config(*g);
  } else if (n == "max_in_flight") {
    g->max_in_flight_opt = static_cast<int64_t>(value);
    if (g->sim_initialized)
      return guarded([&]() -> int {
        choose_capacity(*g);
        return 0;
      });
  } else {
    set_errmsg("unknown option '" + n + "'");
    return OPENMC_E_INVALID_ARGUMENT;
  }
  return 0;
}

static const char* k_stat_names =
  "transport_ms,active_batches,lookups,collisions,segments,lost,xs_events,adv_events,surf_events,"
  "coll_events,search_probes,cycles,checks,capacity,t_sort_ms,t_xs_ms,t_adv_ms,t_split_ms,t_surf_ms,"
  "t_coll_ms,t_init_ms,t_hist_ms,n_sort,n_xs,n_adv,n_split,n_surf,n_coll,n_init,n_hist,wall_s,blocks,xs_lanes,"
  "sort,xs_group,block,sort_every,cycles_per_check,sort_bits,sort_blocks,history_below,history_steps,"
  "hist_launches,hist_xs_events,hist_lookups,hist_adv_events,hist_segments,hist_surf_events,hist_coll_events,"
  "fused_advance,recompute_micro,scatter_fallthrough";

const char* omc_stat_names(void)
{
  return k_stat_names;
}

int omc_get_stats(double* out, int n)
{
  if (!g)
    return OPENMC_E_ALLOCATE;
  std::vector<double> v;
  v.push_back(g->t_transport_ms);
  v.push_back(g->active_batches);
  v.push_back(static_cast<double>(g->counters[CNT_LOOKUPS]));
  v.push_back(static_cast<double>(g->counters[CNT_COLLISIONS]));
  v.push_back(static_cast<double>(g->counters[CNT_SEGMENTS]));
  v.push_back(static_cast<double>(g->counters[CNT_LOST]));
  v.push_back(static_cast<double>(g->counters[CNT_XS_EVENTS]));
  v.push_back(static_cast<double>(g->counters[CNT_ADV_EVENTS]));
  v.push_back(static_cast<double>(g->counters[CNT_SURF_EVENTS]));
  v.push_back(static_cast<double>(g->counters[CNT_COLL_EVENTS]));
  v.push_back(static_cast<double>(g->counters[CNT_SEARCH_PROBES]));
  v.push_back(static_cast<double>(g->n_cycles));
  v.push_back(static_cast<double>(g->n_checks));
  v.push_back(static_cast<double>(g->cap));
  for (int k = 0; k < KERNEL_KINDS; ++k)
    v.push_back(g->kt.ms[k]);
  for (int k = 0; k < KERNEL_KINDS; ++k)
    v.push_back(static_cast<double>(g->kt.launches[k]));
  v.push_back(g->t_batch_wall_s);
  v.push_back(g->lc.blocks);
  v.push_back(g->lc.xs_lanes);
  v.push_back(g->lc.sort);
  v.push_back(g->lc.xs_group);
  v.push_back(g->lc.block);
  v.push_back(g->lc.sort_every);
  v.push_back(g->cycles_per_check);
  v.push_back(g->sort_bits_used);
  v.push_back(g->lc.sort_blocks);
  v.push_back(g->history_below);
  v.push_back(g->history_steps);
  v.push_back(static_cast<double>(g->n_hist_launches));
  for (int c = CNT_HIST_XS_EVENTS; c <= CNT_HIST_COLL_EVENTS; ++c)
    v.push_back(static_cast<double>(g->counters[c]));
  v.push_back(g->lc.fuse_advance && g->lc.xs_lanes == 1 && (g->lc.xs_group == 4 || g->lc.xs_group == 2) ? 1.0 : 0.0);
  v.push_back(g->S.recompute_micro);
  v.push_back(static_cast<double>(g->counters[CNT_SCATTER_FALLTHROUGH]));
  for (int k = 0; k < n && k < static_cast<int>(v.size()); ++k)
    out[k] = v[k];
  return static_cast<int>(v.size());
}

int omc_particle_hashes(const uint64_t** ptr, int64_t* n)
{
  if (!g)
    return OPENMC_E_ALLOCATE;
  *ptr = g->h_hashes.data();
  *n = static_cast<int64_t>(g->h_hashes.size());
  return 0;
}

int omc_energy_grid(double* out3)
{
  if (!g)
    return OPENMC_E_ALLOCATE;
  out3[0] = g->m.energy_min;
  out3[1] = g->m.energy_max;
  out3[2] = g->m.log_spacing;
  return 0;
}

int omc_n_nuclides(void)
{
  return g ? static_cast<int>(g->m.nuclides.size()) : 0;
}

const int* omc_grid_index(int i_nuc)
{
  if (!g || i_nuc < 0 || i_nuc >= static_cast<int>(g->m.nuclides.size()))
    return nullptr;
  return g->m.grid_index.data() + static_cast<size_t>(i_nuc) * (g->m.n_log_bins + 1);
}

int omc_lookup_batch(int i_nuc, const double* E, int64_t n, int learned, int* i_grid, double* xs, uint64_t* probes)
{
  if (!g || !g->initialized)
    return OPENMC_E_ALLOCATE;
  if (i_nuc < 0 || i_nuc >= g->P.n_nuclides) {
    set_errmsg("nuclide index out of bounds");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  if (learned && !g->P.learned.enabled) {
    set_errmsg("learned index has not been built");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    std::vector<void*> tmp;
    double* dE = dev_alloc<double>(tmp, n);
    int* dI = dev_alloc<int>(tmp, n);
    double* dX = dev_alloc<double>(tmp, 2 * n);
    unsigned long long* dP = dev_alloc<unsigned long long>(tmp, 1);
    hip_check(hipMemcpy(dE, E, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    launch_lookup_batch(g->P, i_nuc, dE, n, learned, dI, dX, dP, g->stream);
    hip_check(hipStreamSynchronize(g->stream), "lookup kernel");
    hip_check(hipMemcpy(i_grid, dI, sizeof(int) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    hip_check(hipMemcpy(xs, dX, sizeof(double) * 2 * n, hipMemcpyDeviceToHost), "hipMemcpy");
    unsigned long long pr = 0;
    hip_check(hipMemcpy(&pr, dP, sizeof(pr), hipMemcpyDeviceToHost), "hipMemcpy");
    if (probes)
      *probes = pr;
    free_list(tmp);
    return 0;
  });
}

// omc_lookup_batch without allocation: E, i_grid and xs are device pointers
// the caller owns; the probe partial sums go to the library's file-scope
// device array (xs_lookup.hip), so n is at most NOALLOC_MAX_LOOKUPS.  stream
// null means the library's stream; with probes null the call does not wait.
int omc_lookup_batch_noalloc(int i_nuc, const double* E, int64_t n, int learned, int* i_grid, double* xs,
  uint64_t* probes, void* stream)
{
  if (!g || !g->initialized)
    return OPENMC_E_ALLOCATE;
  if (i_nuc < 0 || i_nuc >= g->P.n_nuclides) {
    set_errmsg("nuclide index out of bounds");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  if (learned && !g->P.learned.enabled) {
    set_errmsg("learned index has not been built");
    return OPENMC_E_ALLOCATE;
  }
  if (n < 0 || n > NOALLOC_MAX_LOOKUPS) {
    set_errmsg("omc_lookup_batch_noalloc: n = " + std::to_string(n) + " exceeds the scratch limit of " +
               std::to_string(NOALLOC_MAX_LOOKUPS) + " lookups per call");
    return OPENMC_E_INVALID_SIZE;
  }
  if (n > 0 && (!E || !i_grid || !xs)) {
    set_errmsg("omc_lookup_batch_noalloc: null device pointer");
    return OPENMC_E_INVALID_ARGUMENT;
  }
  return guarded([&]() -> int {
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : g->stream;
    unsigned long long pr = 0;
    launch_lookup_noalloc(g->P, i_nuc, E, n, learned, i_grid, xs, probes ? &pr : nullptr, st);
    if (probes)
      *probes = pr;
    return 0;
  });
}

int omc_macro_batch(int mat, const double* E, int64_t n, int learned, int repeat, double* macro, double* ms)
{
  if (!g || !g->initialized)
    return OPENMC_E_ALLOCATE;
  if (mat < 0 || mat >= g->P.n_materials) {
    set_errmsg("material index out of bounds");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  if (learned && !g->P.learned.enabled) {
    set_errmsg("learned index has not been built");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    std::vector<void*> tmp;
    double* dE = dev_alloc<double>(tmp, n);
    double* dM = dev_alloc<double>(tmp, 2 * n);
    unsigned long long* dP = dev_alloc<unsigned long long>(tmp, 1);
    hip_check(hipMemcpy(dE, E, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    const int group = g->xs_group_opt > 0 ? g->xs_group_opt : 2;
    launch_macro_batch(g->P, mat, dE, n, learned, dM, dP, g->stream, group); // warm-up
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    int reps = std::max(1, repeat);
    hip_check(hipEventRecord(e0, g->stream), "hipEventRecord");
    for (int r = 0; r < reps; ++r)
      launch_macro_batch(g->P, mat, dE, n, learned, dM, nullptr, g->stream, group);
    hip_check(hipEventRecord(e1, g->stream), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "macro kernel");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    if (ms)
      *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hip_check(hipMemcpy(macro, dM, sizeof(double) * 2 * n, hipMemcpyDeviceToHost), "hipMemcpy");
    free_list(tmp);
    return 0;
  });
}

int omc_check_model(const char* dir, double* out, int n_out, int* grid_index_out, int64_t n_gi)
{
  return guarded([&]() -> int {
    HostModel m;
    read_model(m, dir ? dir : ".");
    initialize_data(m);
    std::vector<double> v = {static_cast<double>(m.nuclides.size()), static_cast<double>(m.mat_ids.size()),
      static_cast<double>(m.cells.size()), static_cast<double>(m.surfaces.size()),
      static_cast<double>(m.lattices.size()), static_cast<double>(m.n_levels), m.energy_min, m.energy_max,
      m.log_spacing, static_cast<double>(m.n_particles), static_cast<double>(m.n_batches),
      m.tallies.empty() ? 0.0 : static_cast<double>(m.tallies[0].n_filter_bins),
      static_cast<double>(m.root_universe)};
    for (double d : m.mat_dens)
      v.push_back(d);
    for (int k = 0; k < n_out && k < static_cast<int>(v.size()); ++k)
      out[k] = v[k];
    if (grid_index_out)
      for (int64_t k = 0; k < n_gi && k < static_cast<int64_t>(m.grid_index.size()); ++k)
        grid_index_out[k] = m.grid_index[k];
    return static_cast<int>(v.size());
  });
}

// Host-side reading of a problem set, without a GPU (for tests):
// out = {nuclides, materials, cells, surfaces, universes, lattices, levels,
//        total particles, tallies, energy_min bits as double, energy_max}
// followed by 8 values per problem {first, n_particles, root universe,
// mat_begin, surf_begin, cell_begin, tally index, the tally's first bin}.
int omc_check_problem_set(int n, const char* const* dirs, double* out, int n_out)
{
  return guarded([&]() -> int {
    std::vector<std::string> d;
    for (int k = 0; k < n; ++k)
      d.push_back(dirs[k] ? dirs[k] : ".");
    HostModel m;
    read_problem_set(m, d);
    initialize_data(m);
    int64_t off = 0;
    for (HostTally& t : m.tallies) {
      t.off = off;
      off += t.n_filter_bins;
    }
    std::vector<double> v = {static_cast<double>(m.nuclides.size()), static_cast<double>(m.mat_ids.size()),
      static_cast<double>(m.cells.size()), static_cast<double>(m.surfaces.size()),
      static_cast<double>(m.universes.size()), static_cast<double>(m.lattices.size()),
      static_cast<double>(m.n_levels), static_cast<double>(m.n_particles), static_cast<double>(m.tallies.size()),
      m.energy_min, m.energy_max};
    for (const HostPart& p : m.parts) {
      v.push_back(static_cast<double>(p.first));
      v.push_back(static_cast<double>(p.n_particles));
      v.push_back(p.root_universe);
      v.push_back(p.mat_begin);
      v.push_back(p.surf_begin);
      v.push_back(p.cell_begin);
      v.push_back(p.tally);
      v.push_back(p.tally >= 0 ? static_cast<double>(m.tallies[p.tally].off) : -1.0);
    }
    for (int k = 0; k < n_out && k < static_cast<int>(v.size()); ++k)
      out[k] = v[k];
    return static_cast<int>(v.size());
  });
}

// Writes one nuclide (elastic scattering and capture at one temperature) in
// the reference's HDF5 layout (openmc.data.IncidentNeutron.export_to_hdf5).
int omc_write_nuclide_hdf5(const char* path, const char* name, double awr, double kT, int temperature,
  const double* energy, const double* elastic, const double* capture, int64_t n)
{
  return guarded([&]() -> int {
    if (!path || !name || n < 2)
      throw std::runtime_error("omc_write_nuclide_hdf5: need a path, a name and two or more grid points");
    write_nuclide_hdf5(path, name, awr, kT, temperature, energy, elastic, capture, static_cast<size_t>(n));
    return 0;
  });
}

// Host-only read of one nuclide's data file (.nuc or HDF5) as the library
// reads it: energy grid and the derived total, absorption and elastic
// columns (up to n_max points); out2 = {awr, kT}.  Returns the grid size.
int omc_read_nuclide(const char* path, const char* name, double temperature, int64_t n_max, double* energy,
  double* total, double* absorption, double* elastic, double* out2)
{
  return guarded([&]() -> int {
    HostNuclide nuc;
    const std::string p = path ? path : "";
    if (is_hdf5_path(p))
      read_nuclide_hdf5(nuc, p, name ? name : "", temperature, 10.0);
    else
      throw std::runtime_error("omc_read_nuclide reads HDF5 files");
    const int64_t n = static_cast<int64_t>(nuc.energy.size());
    for (int64_t k = 0; k < n && k < n_max; ++k) {
      energy[k] = nuc.energy[k];
      total[k] = nuc.total[k];
      absorption[k] = nuc.absorption[k];
      elastic[k] = nuc.elastic[k];
    }
    out2[0] = nuc.awr;
    out2[1] = nuc.kT;
    return static_cast<int>(n);
  });
}

// openmc_statepoint_write (include/openmc/capi.h:161, src/state_point.cpp:36-360):
// a null filename gives the default name of the current batch, a null
// write_source means true.
int openmc_statepoint_write(const char* filename, bool* write_source)
{
  if (!g || !g->initialized) {
    set_errmsg("OpenMC has not been initialized.");
    return OPENMC_E_ALLOCATE;
  }
  return guarded([&]() -> int {
    const std::string name = filename ? std::string(filename) : statepoint_name(*g, 0);
    const auto dot = name.find_last_of('.');
    if (dot == std::string::npos || name.substr(dot + 1) != "h5")
      std::fprintf(stderr, "omc: warning: openmc_statepoint_write was passed a file extension differing from .h5, "
                           "but an hdf5 file will be written.\n");
    write_statepoint(*g, 0, name, write_source ? *write_source : true);
    return 0;
  });
}

// openmc_is_statepoint_batch (include/openmc/capi.h:84, src/simulation.cpp)
bool openmc_is_statepoint_batch(void)
{
  if (!g || !g->initialized || g->m.parts.empty())
    return false;
  const auto& b = g->m.parts[0].statepoint_batches;
  return std::find(b.begin(), b.end(), current_batch) != b.end();
}

// The statepoint of one problem of a problem set (its own tally, global
// tallies and meshes); filename null: the default name in the problem's
// directory.  No reference counterpart (the reference runs one model per
// process).
int omc_problem_statepoint_write(int problem, const char* filename, bool* write_source)
{
  if (!g || !g->initialized || problem < 0 || problem >= static_cast<int>(g->m.parts.size())) {
    set_errmsg("Index in problems array is out of bounds.");
    return OPENMC_E_OUT_OF_BOUNDS;
  }
  return guarded([&]() -> int {
    const std::string name = filename ? std::string(filename) : statepoint_name(*g, problem);
    write_statepoint(*g, problem, name, write_source ? *write_source : true);
    return 0;
  });
}

// Reads a dataset (attribute null or empty) or an attribute of an object of
// an HDF5 file: kind 0 doubles, 1 int64, 2 the raw bytes of fixed-length
// strings.  *n = the number of values (bytes for kind 2); buf is filled only
// when cap holds them.  For reading statepoints back without h5py.
int omc_h5_read(const char* path, const char* object, const char* attribute, int kind, void* buf, size_t cap,
  size_t* n)
{
  return guarded([&]() -> int {
    if (!path || kind < 0 || kind > 2)
      throw std::runtime_error("omc_h5_read: need a path and kind 0, 1 or 2");
    const size_t r = h5_read(path, object ? object : "", attribute ? attribute : "", kind, buf, cap);
    if (n)
      *n = r;
    return 0;
  });
}

int omc_copy_kernel(int64_t n_bytes, int reps, double* ms)
{
  return guarded([&]() -> int {
    std::vector<void*> tmp;
    char* a = dev_alloc<char>(tmp, n_bytes);
    char* b = dev_alloc<char>(tmp, n_bytes);
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    launch_copy(a, b, n_bytes, nullptr);
    hip_check(hipEventRecord(e0, nullptr), "hipEventRecord");
    for (int r = 0; r < std::max(1, reps); ++r)
      launch_copy(a, b, n_bytes, nullptr);
    hip_check(hipEventRecord(e1, nullptr), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "copy kernel");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
    if (ms)
      *ms = t / std::max(1, reps);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    free_list(tmp);
    return 0;
  });
}

int omc_math_batch(int fn, const double* x, const double* y, int64_t n, double* out)
{
  return guarded([&]() -> int {
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count < 1)
      throw std::runtime_error("no GPU visible");
    std::vector<void*> tmp;
    double* dx = dev_alloc<double>(tmp, n);
    double* dy = dev_alloc<double>(tmp, n);
    double* dout = dev_alloc<double>(tmp, n);
    hip_check(hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    if (y)
      hip_check(hipMemcpy(dy, y, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    launch_math_batch(fn, dx, dy, n, dout, nullptr);
    hip_check(hipDeviceSynchronize(), "math kernel");
    hip_check(hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
    free_list(tmp);
    return 0;
  });
}

int omc_learned_build(int n_seg, int epochs, int* max_err)
{
  if (!g || !g->initialized)
    return OPENMC_E_ALLOCATE;
  return guarded([&]() -> int {
    int err = build_learned_index(g->learned, g->m, g->P, n_seg, epochs, g->stream, g->learned_err_bound);
    upload_matnuc(*g);
    if (max_err)
      *max_err = err;
    g->graph_valid = false;
    return 0;
  });
}

} // extern "C"
