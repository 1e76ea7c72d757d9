// Event kernels of the GPU transport loop (gfx950).
//
// One batch runs as: k_init (source sampling, src/simulation.cpp:573-649),
// then cycles of
//     partition(staged -> q_xs) -> k_xs -> k_advance
//       -> partition(q_xs -> q_sc) -> k_surface -> k_collision
// until no particle is alive.  The calculate-xs queue doubles as the advance
// queue (every particle that computes cross sections advances next,
// src/event.cpp:87-101).  No kernel appends to a shared queue counter: each
// writes its result for input i at position i of a staging array (the
// particle's slot and a key, -1 when its history ended), and a partition pass
// (per-workgroup key counts, a column scan, a key scan, an LDS-ranked
// scatter) compacts the staging array by key into the next queue: by outcome
// (surface crossing, collision) after advance, by (material, energy bucket)
// for calculate-xs -- the order the comment at src/event.cpp:77-85 asks for,
// so that neighbouring lanes read neighbouring parts of the same energy
// grids.  Histories do not depend on the processing order.
//
// Particle state is one 128-byte record per particle (PHot) plus one 64-byte
// record per deeper coordinate level, so visiting particles in queue order
// costs one cache line per particle rather than one per field.  Queue sizes
// live in device memory, written by the partition passes, so a cycle needs no
// host round trip and a run of cycles is captured in one hipGraph.
//
// All kernels are templated on NL, the number of coordinate levels of the
// model, so that per-level state stays in registers.
#include <hip/hip_runtime.h>

#include "device_physics.h"
#include "transport.h"

namespace omc {

constexpr int BLOCK = 256;

__device__ __forceinline__ int lane_id()
{
  return __lane_id();
}

// Key of a particle about to calculate cross sections, for the counting sort
// of the calculate-xs queue: 0 for a particle in void, else
// 1 + material * sort_eb + energy bucket, the bucket being the exponent and
// leading mantissa bits of E offset by those of the grid minimum (sort_eb
// buckets span the binades of the data's energy range).  The kernel that
// appends the particle stores it beside the slot.
__device__ __forceinline__ int sort_key(const DevParticles& S, int material, double E)
{
  if (material < 0)
    return 0;
  const uint64_t bits = static_cast<uint64_t>(__double_as_longlong(E));
  int eb = static_cast<int>(bits >> S.sort_shift) - S.sort_base;
  eb = eb < 0 ? 0 : (eb >= S.sort_eb ? S.sort_eb - 1 : eb);
  int key = 1 + material * S.sort_eb + eb;
  return key < S.sort_nb ? key : S.sort_nb - 1;
}

__device__ __forceinline__ int wave_slot()
{
  return static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / 64u) & (CNT_SLOTS - 1);
}

// Wavefront sum of a counter, added into this wavefront's slot of the
// counter's CNT_SLOTS slots.  Every lane must call it.
__device__ __forceinline__ void wave_count(const DevParticles& S, int counter, unsigned long long v)
{
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  if (lane_id() == 0 && v != 0ull)
    atomicAdd(&S.counters[counter * CNT_SLOTS + wave_slot()], v);
}

template<int NL>
__device__ __forceinline__ PLevel* level_rec(const DevParticles& S, int64_t p, int l)
{
  return S.lev + p * (NL > 1 ? NL - 1 : 1) + (l - 1);
}

// First material, surface and cell index of a particle's problem: history
// hashes record problem-local indices, so a problem's hashes in a problem set
// equal its hashes when run alone (all 0 for a single model).
struct PartBase {
  int mat, surf, cell;
};
__device__ __forceinline__ PartBase part_base(const DevProblem& P, int cell0)
{
  if (P.n_parts <= 1 || cell0 < 0)
    return PartBase {0, 0, 0};
  const DevPart& q = P.part[P.cell[cell0].part];
  return PartBase {q.mat_begin, q.surf_begin, q.cell_begin};
}

// Geometry state from a particle's records (h already loaded by the caller).
template<int NL>
__device__ __forceinline__ void load_geo(const DevProblem& P, const DevParticles& S, int64_t p, const PHot& h,
  Geo<NL>& g)
{
  g.rx[0] = h.rx;
  g.ry[0] = h.ry;
  g.rz[0] = h.rz;
  g.ux = h.ux;
  g.uy = h.uy;
  g.uz = h.uz;
  g.cell[0] = h.cell0;
  // level 0 is always the root universe (of the particle's problem in a problem set)
  g.univ[0] = h.cell0 == C_NONE ? C_NONE : (P.n_parts > 1 ? P.cell[h.cell0].universe : P.root_universe);
  g.lat[0] = C_NONE;
  g.lix[0] = g.liy[0] = g.liz[0] = 0;
#pragma unroll
  for (int l = 1; l < NL; ++l) {
    const PLevel L = *level_rec<NL>(S, p, l);
    g.rx[l] = L.rx;
    g.ry[l] = L.ry;
    g.rz[l] = L.rz;
    g.cell[l] = L.cell;
    g.univ[l] = L.univ;
    g.lat[l] = L.lat;
    g.lix[l] = L.lix;
    g.liy[l] = L.liy;
    g.liz[l] = L.liz;
  }
  g.n_coord = h.n_coord;
  g.surface = h.surface;
  g.material = h.material;
  g.material_last = h.material_last;
  g.lost = 0;
}

// Write the geometry fields of a particle back (levels >= 1 in full).
template<int NL>
__device__ __forceinline__ void store_geo(const DevParticles& S, int64_t p, const Geo<NL>& g)
{
  PHot* h = S.hot + p;
  h->rx = g.rx[0];
  h->ry = g.ry[0];
  h->rz = g.rz[0];
  h->ux = g.ux;
  h->uy = g.uy;
  h->uz = g.uz;
  h->cell0 = g.cell[0];
  h->material = g.material;
  h->material_last = g.material_last;
  h->n_coord = g.n_coord;
  h->surface = g.surface;
#pragma unroll
  for (int l = 1; l < NL; ++l) {
    PLevel L;
    L.rx = g.rx[l];
    L.ry = g.ry[l];
    L.rz = g.rz[l];
    L.cell = g.cell[l];
    L.univ = g.univ[l];
    L.lat = g.lat[l];
    L.lix = g.lix[l];
    L.liy = g.liy[l];
    L.liz = g.liz[l];
    L.pad[0] = L.pad[1] = L.pad[2] = L.pad[3] = 0;
    *level_rec<NL>(S, p, l) = L;
  }
}

template<int NL>
__device__ __forceinline__ void geo_clear(Geo<NL>& g, V3 r, V3 u)
{
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    g.set_r(l, V3 {0, 0, 0});
    g.reset(l);
  }
  g.set_r(0, r);
  g.set_u(u);
  g.n_coord = 1;
  g.surface = SURFACE_NONE;
  g.material = C_NONE;
  g.material_last = C_NONE;
  g.lost = 0;
}

//==============================================================================
// Source sampling (src/source.cpp:226-271, 355-417; src/simulation.cpp:573-649)
//==============================================================================

template<int NL>
__device__ __forceinline__ bool satisfies_spatial_constraints(const DevProblem& P, const DevSource& src, int root,
  V3 r)
{
  Geo<NL> g;
  geo_clear(g, r, V3 {0.0, 0.0, 1.0});
  g.univ[0] = root;
  if (!exhaustive_find_cell(P, g))
    return false;
  bool accepted = true;
  if (src.n_domain > 0) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      if (i >= g.n_coord)
        break;
      const int id = P.cell[g.cell[i]].id;
      accepted = false;
      for (int k = 0; k < src.n_domain; ++k)
        accepted = accepted || (src.domain_ids[k] == id);
      if (accepted)
        break;
    }
  }
  return accepted;
}

// maxwell_spectrum / watt_spectrum (src/random_dist.cpp:20-39)
__device__ __forceinline__ double watt_spectrum(double a, double b, uint64_t* seed)
{
  double r1 = prn(seed);
  double r2 = prn(seed);
  double r3 = prn(seed);
  double c = detm::cos(PI / 2. * r3);
  double w = -a * (detm::log(r1) + detm::log(r2) * c * c);
  return w + 0.25 * a * a * b + uniform_distribution(-1., 1., seed) * sqrt(a * a * b * w);
}

constexpr int MAX_SOURCE_TRIES = 1 << 20;

// Problem of population index gi in a problem set: the last problem whose
// first index is <= gi (problems are laid out in population order).
__device__ __forceinline__ int part_of_index(const DevProblem& P, int64_t gi)
{
  int lo = 0, hi = P.n_parts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (P.part[mid].first <= gi)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Source particles for population indices g_base + [0, n) of batch `batch`.
// A particle's id is its index within its own problem plus the batches
// before (src/simulation.cpp:581-583, 613-616), so a problem of a problem set
// samples exactly the histories it samples when run alone.
template<int NL, bool MULTI>
__global__ void __launch_bounds__(BLOCK) k_init(DevProblem P, DevParticles S, int64_t n, int64_t g_base,
  int64_t batch)
{
  unsigned long long n_lost = 0;
  double total_w = 0.0; // source weight after the birth weight windows
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t gi = g_base + i;
    int part = 0;
    int64_t id;
    if (MULTI) {
      part = part_of_index(P, gi);
      id = (batch - 1) * P.part[part].n_particles + (gi - P.part[part].first) + 1;
    } else {
      id = (batch - 1) * P.n_particles + gi + 1;
    }
    const DevSource& src = MULTI ? P.part[part].src : P.src;
    const int root = MULTI ? P.part[part].root_universe : P.root_universe;
    uint64_t seed = future_seed(static_cast<uint64_t>(id) * P.stride, P.master_seed + 1);
    V3 r = {0, 0, 0};
    bool lost = false;
    int tries = 0;
    bool accepted = false;
    while (!accepted) {
      if (src.space_type == 0) {
        r = {src.p0[0], src.p0[1], src.p0[2]};
      } else {
        double x1 = prn(&seed);
        double x2 = prn(&seed);
        double x3 = prn(&seed);
        V3 ll = {src.p0[0], src.p0[1], src.p0[2]};
        V3 ur = {src.p1[0], src.p1[1], src.p1[2]};
        V3 ext = v_sub(ur, ll);
        r = v_add(ll, V3 {x1 * ext.x, x2 * ext.y, x3 * ext.z});
      }
      accepted = satisfies_spatial_constraints<NL>(P, src, root, r);
      if (++tries >= MAX_SOURCE_TRIES && !accepted) {
        lost = true;
        break;
      }
    }
    V3 u;
    if (src.angle_type == 0) {
      // isotropic_direction (src/distribution_multi.cpp:106-112)
      double phi = uniform_distribution(0., 2.0 * PI, &seed);
      double mu = uniform_distribution(-1., 1., &seed);
      double s, c;
      detm::sincos(phi, &s, &c);
      u = {mu, sqrt(1.0 - mu * mu) * c, sqrt(1.0 - mu * mu) * s};
    } else {
      u = {src.u_ref[0], src.u_ref[1], src.u_ref[2]};
    }
    double E = 0.0;
    tries = 0;
    while (true) {
      if (src.energy_type == 0) {
        int nd = src.n_discrete;
        int k = 0;
        if (nd > 1) {
          size_t uu = static_cast<size_t>(prn(&seed) * nd);
          if (prn(&seed) < src.disc_prob[uu])
            k = static_cast<int>(uu);
          else
            k = src.disc_alias[uu];
        }
        E = src.disc_x[k];
      } else {
        E = watt_spectrum(src.watt_a, src.watt_b, &seed);
      }
      if (E < P.energy_max && (E > 0.0 && E < INFTY))
        break;
      if (++tries >= MAX_SOURCE_TRIES) {
        lost = true;
        break;
      }
    }
    // Particle::from_source (src/particle.cpp:114-158) + initialize_history.
    // The cell search that the first calculate_xs event performs for a new
    // particle (src/particle.cpp:180-196) is done here instead: it reads only
    // the position and direction set above, draws no random number, and
    // leaves the same coordinate levels, material and material_last.
    Geo<NL> g;
    geo_clear(g, r, u);
    g.univ[0] = root;
    if (!lost) {
      if (!exhaustive_find_cell(P, g) || g.lost)
        lost = true;
    }
    PHot h;
    h.rx = g.rx[0];
    h.ry = g.ry[0];
    h.rz = g.rz[0];
    h.ux = u.x;
    h.uy = u.y;
    h.uz = u.z;
    h.E = E;
    h.xs_t = 0.0;
    h.xs_a = 0.0;
    h.time = 0.0;
    h.seed = future_seed(static_cast<uint64_t>(id) * P.stride, P.master_seed + 0);
    // initialize_history: the birth weight window (src/simulation.cpp:599-639)
    double wgt = 1.0;
    uint64_t hh = 0xcbf29ce484222325ULL ^ static_cast<uint64_t>(id);
    if (P.ww.on) {
      PWW st;
      st.wgt_born = wgt;
      st.wgt_ww_born = -1.0;
      st.ww_factor = 0.0;
      st.n_split = 0;
      st.n_bank = 0;
      const SplitSink sink {S.bank + i * S.bank_cap, S.bank_cap, S.status};
      apply_weight_windows(P.ww, st, wgt, r, u, E, 0.0, SURFACE_NONE, &h.seed, sink);
      S.ww[i] = st;
      if (S.trace_hash) {
        hmix(hh, dbits(wgt));
        hmix(hh, static_cast<uint64_t>(st.n_bank));
      }
    }
    total_w += wgt;
    h.wgt = lost ? 0.0 : wgt;
    h.cell0 = g.cell[0];
    h.material = g.material;
    h.material_last = g.material_last;
    h.n_coord = g.n_coord;
    h.surface = SURFACE_NONE;
    h.b_surface = SURFACE_NONE;
    h.b_info = 0;
    h.n_event = 0;
    S.hot[i] = h;
    store_geo<NL>(S, i, g);
    S.lost[i] = lost ? 1 : 0;
    n_lost += lost ? 1 : 0;
    S.hash[i] = hh;
    S.stage_p[i] = static_cast<int>(i);
    S.stage_key[i] = h.wgt == 0.0 ? -1 : sort_key(S, g.material, E);
  }
  wave_count(S, CNT_LOST, n_lost);
  for (int o = 32; o > 0; o >>= 1)
    total_w += __shfl_xor(total_w, o);
  if (lane_id() == 0 && total_w != 0.0)
    unsafeAtomicAdd(&S.total_weight[wave_slot()], total_w);
}

//==============================================================================
// Partition passes: counting sort of a staging array by key
//==============================================================================

// A partition reads n = *n_in (slot, key) pairs, drops those with key < 0,
// and writes the others to `out` grouped by key in increasing key order (the
// order inside a key is free: no history depends on processing order).  It
// writes the number kept to *n_out and, if bin_tot is set, each key's count.
// The calculate-xs partition keys by (material, energy bucket) -- the order
// the comment at src/event.cpp:77-85 asks for -- so that neighbouring lanes
// search neighbouring parts of the same energy grids (one key when sorting
// is off: a plain compaction); the advance partition keys by outcome
// (0 surface crossing, 1 collision).
struct Part {
  const int* val;
  const int* key;
  const int* n_in;
  int* out;
  int* n_out;
  int* bin_tot;
  int nb;
  int div; // 0: keys as they are; else the material part of a calculate-xs key, 1 + (key - 1) / div
};

// A partition of n entries uses nw(n) = clamp(ceil(n / 2048), 1, max_blocks)
// workgroups (the launch has max_blocks; the others return at once), so
// short queues late in a batch scan a small count matrix.  Workgroup b owns
// the contiguous chunk [b * chunk, min(n, (b + 1) * chunk)).
constexpr int SORT_MIN_PER_BLOCK = 2048;
constexpr int PART_UNROLL = 8; // queue entries a thread loads before using them

__device__ __forceinline__ int sort_workgroups(int n, int max_blocks)
{
  int w = (n + SORT_MIN_PER_BLOCK - 1) / SORT_MIN_PER_BLOCK;
  return w < 1 ? 1 : (w > max_blocks ? max_blocks : w);
}

__device__ __forceinline__ int part_key(const Part& A, int i)
{
  const int k = A.key[i];
  return k < 0 ? -1 : (A.nb == 1 ? 0 : (A.div == 0 || k == 0 ? k : 1 + (k - 1) / A.div));
}

// Per-workgroup key counts of its chunk into row b of the count matrix
// sort_count[b][key] (row length nb: a workgroup writes one contiguous row).
__global__ void __launch_bounds__(BLOCK) k_part_count(DevParticles S, Part A)
{
  __shared__ int sh_cnt[SORT_MAX_BINS];
  const int n = *A.n_in;
  const int nw = sort_workgroups(n, gridDim.x);
  if (static_cast<int>(blockIdx.x) >= nw)
    return;
  for (int k = threadIdx.x; k < A.nb; k += blockDim.x)
    sh_cnt[k] = 0;
  __syncthreads();
  const int chunk = (n + nw - 1) / nw;
  const int lo = blockIdx.x * chunk;
  const int hi = min(n, lo + chunk);
  // PART_UNROLL keys loaded before their LDS atomics, so a thread keeps that
  // many loads in flight instead of one
  const int stride = blockDim.x;
  int i = lo + threadIdx.x;
  for (; i + (PART_UNROLL - 1) * stride < hi; i += PART_UNROLL * stride) {
    int k[PART_UNROLL];
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u)
      k[u] = part_key(A, i + u * stride);
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u)
      if (k[u] >= 0)
        atomicAdd(&sh_cnt[k[u]], 1);
  }
  for (; i < hi; i += stride) {
    const int k = part_key(A, i);
    if (k >= 0)
      atomicAdd(&sh_cnt[k], 1);
  }
  __syncthreads();
  int* row = S.sort_count + static_cast<int64_t>(blockIdx.x) * A.nb;
  for (int k = threadIdx.x; k < A.nb; k += blockDim.x)
    row[k] = sh_cnt[k];
}

// One thread per key, 64 keys per workgroup: exclusive scan of the key's
// column over the workgroup rows (sort_count[b][key] becomes workgroup b's
// offset inside the key) and the key total into sort_total[key].  Each step
// of the scan reads and writes 64 consecutive keys of one row.
__global__ void __launch_bounds__(64) k_part_cols(DevParticles S, Part A, int max_blocks)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= A.nb)
    return;
  const int nw = sort_workgroups(*A.n_in, max_blocks);
  int* col = S.sort_count + k;
  int run = 0;
  int b = 0;
  for (; b + 4 <= nw; b += 4) {
    int v[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      v[j] = col[static_cast<int64_t>(b + j) * A.nb];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      col[static_cast<int64_t>(b + j) * A.nb] = run;
      run += v[j];
    }
  }
  for (; b < nw; ++b) {
    const int v = col[static_cast<int64_t>(b) * A.nb];
    col[static_cast<int64_t>(b) * A.nb] = run;
    run += v;
  }
  S.sort_total[k] = run;
}

// One workgroup: exclusive scan of the key totals into key starts; writes
// the number kept and the optional per-key counts.
__global__ void __launch_bounds__(1024) k_part_scan(DevParticles S, Part A)
{
  __shared__ int sh_part[1024];
  const int nb = A.nb;
  const int per = (nb + blockDim.x - 1) / blockDim.x;
  const int b0 = threadIdx.x * per;
  int local = 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < nb)
      local += S.sort_total[b0 + k];
  sh_part[threadIdx.x] = local;
  __syncthreads();
  // Hillis-Steele inclusive scan of the per-thread partial sums
  for (int o = 1; o < static_cast<int>(blockDim.x); o <<= 1) {
    int v = threadIdx.x >= o ? sh_part[threadIdx.x - o] : 0;
    __syncthreads();
    sh_part[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sh_part[threadIdx.x] - local;
  for (int k = 0; k < per; ++k) {
    if (b0 + k < nb) {
      const int c = S.sort_total[b0 + k];
      S.sort_total[b0 + k] = run;
      if (A.bin_tot)
        A.bin_tot[b0 + k] = c;
      run += c;
    }
  }
  if (threadIdx.x == blockDim.x - 1)
    *A.n_out = sh_part[threadIdx.x];
}

// Each workgroup places its chunk: ranks inside a key come from LDS atomics.
__global__ void __launch_bounds__(BLOCK) k_part_scatter(DevParticles S, Part A)
{
  __shared__ int sh_off[SORT_MAX_BINS];
  const int n = *A.n_in;
  const int nw = sort_workgroups(n, gridDim.x);
  if (static_cast<int>(blockIdx.x) >= nw)
    return;
  for (int k = threadIdx.x; k < A.nb; k += blockDim.x)
    sh_off[k] = S.sort_total[k] + S.sort_count[static_cast<int64_t>(blockIdx.x) * A.nb + k];
  __syncthreads();
  const int chunk = (n + nw - 1) / nw;
  const int lo = blockIdx.x * chunk;
  const int hi = min(n, lo + chunk);
  const int stride = blockDim.x;
  int i = lo + threadIdx.x;
  for (; i + (PART_UNROLL - 1) * stride < hi; i += PART_UNROLL * stride) {
    int k[PART_UNROLL], v[PART_UNROLL];
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
      k[u] = part_key(A, i + u * stride);
      v[u] = A.val[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < PART_UNROLL; ++u) {
      if (k[u] >= 0) {
        const int pos = atomicAdd(&sh_off[k[u]], 1);
        A.out[pos] = v[u];
      }
    }
  }
  for (; i < hi; i += stride) {
    const int k = part_key(A, i);
    if (k >= 0) {
      const int pos = atomicAdd(&sh_off[k], 1);
      A.out[pos] = A.val[i];
    }
  }
}

//==============================================================================
// calculate_xs event (src/particle.cpp:160-230, src/material.cpp:829-901)
//
// LPP lanes cooperate on one particle: lane `sub` of the group looks up
// nuclides sub, sub + LPP, ... of the material, writes N_j * sigma_j to LDS,
// and the group's first lane forms the macroscopic sums in material order
// exactly as Material::calculate_neutron_xs does (product, then add).
//==============================================================================

template<int NL, int LPP>
__global__ void __launch_bounds__(BLOCK) k_xs(DevProblem P, DevParticles S, const int* q)
{
  extern __shared__ double sh_xs[];
  const int n = S.qcount[QC_XS];
  constexpr int PPB = BLOCK / LPP; // particles per block pass
  const int slot = threadIdx.x / LPP;
  const int sub = threadIdx.x % LPP;
  const int lane = lane_id();
  const int leader = (lane / LPP) * LPP;
  const int max_nn = P.max_mat_nuc;
  double* prod = sh_xs + static_cast<int64_t>(slot) * max_nn * 2;
  int* igs = reinterpret_cast<int*>(sh_xs + static_cast<int64_t>(PPB) * max_nn * 2) + slot * max_nn;
  const bool learned = S.use_learned != 0;
  const bool trace = S.trace_hash != 0;
  unsigned long long n_lookups = 0, n_events = 0, n_probes = 0;

  for (int base = blockIdx.x * PPB; base < n; base += gridDim.x * PPB) {
    const int i = base + slot;
    int p = 0;
    int material = MATERIAL_VOID;
    int need = 0;
    double E = 0.0;
    if (i < n && sub == 0) {
      p = q[i];
      n_events++;
      const PHot* h = S.hot + p;
      material = h->material;
      // every particle reaching this event has a cell: source particles are
      // located in k_init and crossings relocate in k_surface
      if (h->wgt != 0.0) {
        if (material != MATERIAL_VOID) {
          need = material != h->material_last ? 1 : 0;
          E = h->E;
        } else {
          S.hot[p].xs_t = 0.0;
          S.hot[p].xs_a = 0.0;
        }
      }
    }
    if (LPP > 1) {
      p = __shfl(p, leader);
      material = __shfl(material, leader);
      need = __shfl(need, leader);
      E = __shfl(E, leader);
    }
    int nn = 0;
    if (need) {
      const double lx = detm::log(E / P.energy_min);
      int i_log = static_cast<int>(lx / P.log_spacing);
      i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
      const int j0 = P.mat_off[material];
      nn = P.mat_off[material + 1] - j0;
      int probes = 0;
      for (int jj = sub; jj < nn; jj += LPP) {
        const int j = j0 + jj;
        Micro m = nuclide_xs(P, P.mat_nuc[j], E, i_log, lx, learned, probes);
        const double dens = P.mat_dens[j] * 1.0;
        prod[2 * jj] = dens * m.total;
        prod[2 * jj + 1] = dens * m.absorption;
        igs[jj] = m.i_grid;
        n_lookups++;
      }
      n_probes += probes;
    }
    __syncthreads();
    if (need && sub == 0) {
      double xs_t = 0.0;
      double xs_a = 0.0;
      double* row = S.micro_t + static_cast<int64_t>(p) * S.max_nn;
      for (int jj = 0; jj < nn; ++jj) {
        xs_t += prod[2 * jj];
        xs_a += prod[2 * jj + 1];
        if (!S.recompute_micro)
          row[jj] = xs_t; // running sums (see store_row)
      }
      S.hot[p].xs_t = xs_t;
      S.hot[p].xs_a = xs_a;
      if (trace) {
        uint64_t h = S.hash[p];
        hmix(h, 0x11);
        hmix(h, static_cast<uint64_t>(material - part_base(P, S.hot[p].cell0).mat));
        for (int jj = 0; jj < nn; ++jj)
          hmix(h, static_cast<uint64_t>(static_cast<uint32_t>(igs[jj])));
        S.hash[p] = h;
      }
    }
    __syncthreads();
  }
  wave_count(S, CNT_LOOKUPS, n_lookups);
  wave_count(S, CNT_XS_EVENTS, n_events);
  wave_count(S, CNT_SEARCH_PROBES, n_probes);
}

// Values a calculate-xs lane collects (in LDS) before it writes them to its
// particle's micro_t row: a 64-byte piece of a line at once.
constexpr int ROW_CHUNK = 8;

// kk (<= G) values of a group into the particle's micro_t row, two per
// 16-byte store: the group starts at a multiple of G in a row padded to a
// multiple of four values, so the pairs are 16-byte aligned.  The row holds
// the running sums of atom density x micro total in material order (the
// macroscopic total's partial sums), which the collision searches for the
// sampled nuclide.
template<int G>
__device__ __forceinline__ void store_row(double* dst, const double* v, int kk)
{
  if (G == 1) {
    dst[0] = v[0];
    return;
  }
#pragma unroll
  for (int k = 0; k < G; k += 2) {
    if (k + 1 < kk)
      *reinterpret_cast<double2*>(dst + k) = double2 {v[k], v[k + 1]};
    else if (k < kk)
      dst[k] = v[k];
  }
}

// Macroscopic cross sections of one particle (material given), nuclides
// looked up G at a time; sums in material order.  Returns the lookups done.
// The material must be the same in every active lane of the wavefront.
template<int G, bool LEARNED, bool STAGE = false>
__device__ __forceinline__ int particle_xs(const DevProblem& P, const DevParticles& S, int p, PHot* h,
  int material, double E, bool trace, int& probes)
{
  const double lx = detm::log(E / P.energy_min);
  int i_log = static_cast<int>(lx / P.log_spacing);
  i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
  const int j0 = uniform_load(P.mat_off + material);
  const int nn = uniform_load(P.mat_off + material + 1) - j0;
  double* micro = S.micro_t + static_cast<int64_t>(p) * S.max_nn;
  uint64_t hh = 0;
  if (trace) {
    hh = S.hash[p];
    hmix(hh, 0x11);
    hmix(hh, static_cast<uint64_t>(material - part_base(P, h->cell0).mat));
  }
  double xs_t = 0.0;
  double xs_a = 0.0;
  // STAGE: the row is collected ROW_CHUNK values at a time in this lane's
  // column of LDS and written as back-to-back 16-byte stores that fill a
  // 64-byte piece of a line at once.  Written a group at a time, the pieces
  // of a line reached L2 far apart and went to HBM in parts (measured: 5.9 KB
  // of HBM traffic per calculate-xs event on the 300-nuclide material,
  // against 2.8 KB collected).
  [[maybe_unused]] double* stage = nullptr;
  if constexpr (STAGE) {
    __shared__ double row_stage[ROW_CHUNK * BLOCK];
    stage = row_stage + threadIdx.x; // value t of the chunk at stage[t * BLOCK]
  }
  for (int jj = 0; jj < nn; jj += G) {
    const int kk = nn - jj < G ? nn - jj : G;
    Micro m[G];
    nuclide_xs_group<G, LEARNED>(P, P.matnuc + j0 + jj, kk, E, i_log, lx, probes, m);
    double run[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      if (k < kk) {
        const double dens = uniform_load(&P.matnuc[j0 + jj + k].dens) * 1.0;
        xs_t += dens * m[k].total;
        xs_a += dens * m[k].absorption;
        run[k] = xs_t;
        if (trace)
          hmix(hh, static_cast<uint64_t>(static_cast<uint32_t>(m[k].i_grid)));
      }
    }
    if (!S.recompute_micro) {
      if constexpr (STAGE && G < ROW_CHUNK) {
        const int base = jj & (ROW_CHUNK - 1); // G divides ROW_CHUNK
#pragma unroll
        for (int k = 0; k < G; ++k)
          if (k < kk)
            stage[(base + k) * BLOCK] = run[k];
        const int filled = base + kk;
        if (filled == ROW_CHUNK || jj + kk == nn) {
          double* dst = micro + (jj - base);
#pragma unroll
          for (int t = 0; t < ROW_CHUNK; t += 2) {
            if (t + 1 < filled)
              *reinterpret_cast<double2*>(dst + t) = double2 {stage[t * BLOCK], stage[(t + 1) * BLOCK]};
            else if (t < filled)
              dst[t] = stage[t * BLOCK];
          }
        }
      } else {
        store_row<G>(micro + jj, run, kk);
      }
    }
  }
  // xs_t and xs_a are adjacent and 16-byte aligned in the hot record: one store
  *reinterpret_cast<double2*>(&h->xs_t) = double2 {xs_t, xs_a};
  if (trace)
    S.hash[p] = hh;
  return nn;
}

// particle_xs with the wavefront-cooperative searches (nuclide_xs_wave):
// every lane of the wavefront calls it; lanes with act == false (another
// material, or no particle) take part in the cross-lane work and store
// nothing.  Returns the lookups done by this lane.
template<int G>
__device__ __forceinline__ int particle_xs_wave(const DevProblem& P, const DevParticles& S, int p, PHot* h,
  int material, double E_in, bool act, bool trace, int& probes)
{
  const double E = act ? E_in : P.energy_min;
  const double lx = detm::log(E / P.energy_min);
  int i_log = static_cast<int>(lx / P.log_spacing);
  i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
  // the log-grid range of the participating lanes
  const int il_lo = wave_min_int(act ? i_log : P.n_log_bins);
  const int il_hi = wave_max_int(act ? i_log : -1);
  const int j0 = uniform_load(P.mat_off + material);
  const int nn = uniform_load(P.mat_off + material + 1) - j0;
  double* micro = S.micro_t + static_cast<int64_t>(p) * S.max_nn;
  uint64_t hh = 0;
  if (trace && act) {
    hh = S.hash[p];
    hmix(hh, 0x11);
    hmix(hh, static_cast<uint64_t>(material - part_base(P, h->cell0).mat));
  }
  double xs_t = 0.0;
  double xs_a = 0.0;
  for (int jj = 0; jj < nn; jj += G) {
    const int kk = nn - jj < G ? nn - jj : G;
    Micro m[G];
    nuclide_xs_wave<G>(P, P.matnuc + j0 + jj, kk, E, i_log, il_lo, il_hi, act, probes, m);
    double run[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      if (k < kk) {
        const double dens = uniform_load(&P.matnuc[j0 + jj + k].dens) * 1.0;
        xs_t += dens * m[k].total;
        xs_a += dens * m[k].absorption;
        run[k] = xs_t;
        if (trace)
          hmix(hh, static_cast<uint64_t>(static_cast<uint32_t>(m[k].i_grid)));
      }
    }
    if (act && !S.recompute_micro)
      store_row<G>(micro + jj, run, kk);
  }
  if (act) {
    *reinterpret_cast<double2*>(&h->xs_t) = double2 {xs_t, xs_a};
    if (trace)
      S.hash[p] = hh;
  }
  return act ? nn : 0;
}


//==============================================================================
// advance event (src/particle.cpp:232-286, src/event.cpp:107-128)
//==============================================================================

// Event counts of one lane, summed per wavefront into the counter slots.
struct EvCount {
  unsigned long long events = 0, lost = 0, segs = 0, lookups = 0, probes = 0, fallthrough = 0;
};

// Particle::event_advance for slot p (src/particle.cpp:232-286): distance to
// boundary vs collision, move, time, track-length tallies.  Returns the
// outcome: 0 surface crossing, 1 collision, -1 done.
template<int NL, bool MULTI>
__device__ __forceinline__ int advance_one(const DevProblem& P, const DevParticles& S, int p, EvCount& c)
{
  const PHot h = S.hot[p];
  double wgt = h.wgt;
  if (wgt == 0.0)
    return -1;
  c.events++;
  Geo<NL> g;
  load_geo<NL>(P, S, p, h, g);
  const double E = h.E;
  const double xs_t = h.xs_t;
  uint64_t seed = h.seed;
  Boundary b = distance_to_boundary(P, g);
  double d_coll;
  if (xs_t == 0.0)
    d_coll = __builtin_inf();
  else
    d_coll = -detm::log(prn(&seed)) / xs_t;
  double speed = C_LIGHT * sqrt(E * (E + 2 * MASS_NEUTRON_EV)) / (E + MASS_NEUTRON_EV);
  double distance = b.distance;
  if (d_coll < distance)
    distance = d_coll;
  if (INFTY < distance)
    distance = INFTY;
  const V3 r0 = g.r(0);
  // move_distance (src/particle_data.cpp:86-91): r += distance * u on every level
  const double mx = g.ux * distance, my = g.uy * distance, mz = g.uz * distance;
  const V3 r1 = {g.rx[0] + mx, g.ry[0] + my, g.rz[0] + mz};
  PHot* hp = S.hot + p;
  hp->rx = r1.x;
  hp->ry = r1.y;
  hp->rz = r1.z;
#pragma unroll
  for (int l = 1; l < NL; ++l) {
    if (l < g.n_coord) {
      PLevel* L = level_rec<NL>(S, p, l);
      L->rx = g.rx[l] + mx;
      L->ry = g.ry[l] + my;
      L->rz = g.rz[l] + mz;
    }
  }
  double dt = distance / speed;
  hp->time = h.time + dt;
  if (P.tally.enabled) {
    if (MULTI) {
      const DevTally& T = P.part[h.cell0 >= 0 ? P.cell[h.cell0].part : 0].tally;
      if (T.n_bins > 0)
        c.segs += score_tracklength(T, r0, r1, g.u(), E, wgt, distance);
    } else {
      c.segs += score_tracklength(P.tally, r0, r1, g.u(), E, wgt, distance);
    }
  }
  if (distance == INFTY)
    wgt = 0.0;
  if (g.lost) {
    wgt = 0.0;
    S.lost[p] = 1;
    c.lost++;
  }
  hp->seed = seed;
  hp->b_surface = b.surface;
  hp->b_info = b.coord_level | ((b.trans[0] + 1) << 8) | ((b.trans[1] + 1) << 10) | ((b.trans[2] + 1) << 12);
  hp->wgt = wgt;
  if (S.trace_hash) {
    uint64_t hh = S.hash[p];
    hmix(hh, 0x22);
    hmix(hh, dbits(distance));
    hmix(hh, d_coll > b.distance ? 2 : 1);
    S.hash[p] = hh;
  }
  if (wgt == 0.0)
    return -1;
  return d_coll > b.distance ? 0 : 1;
}

template<int NL, bool MULTI>
__global__ void __launch_bounds__(BLOCK) k_advance(DevProblem P, DevParticles S, const int* q)
{
  const int n = S.qcount[QC_XS];
  EvCount c;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    S.adv_key[i] = advance_one<NL, MULTI>(P, S, q[i], c);
  wave_count(S, CNT_ADV_EVENTS, c.events);
  wave_count(S, CNT_SEGMENTS, c.segs);
  wave_count(S, CNT_LOST, c.lost);
}

// One lane per particle: the lane looks up the material's nuclides G at a
// time (their searches interleaved) and forms the sums in material order as
// it goes, so no LDS is needed and occupancy is set by registers alone.
// A wavefront handles one material at a time (the queue is sorted by
// material, so almost always in one pass): the material is then
// wavefront-uniform and its nuclide list, grid offsets and sizes are read
// with scalar loads into scalar registers.
//
// With ADV != 0 the same kernel then advances each lane's particle
// (advance_one; ADV 2 for a problem set): the particle's records are still in
// the caches, and the cycle has one launch less.
template<int NL, int G, bool LEARNED, int ADV, bool WAVE>
__device__ __forceinline__ void xs_lane_body(const DevProblem& P, const DevParticles& S, const int* q)
{
  const int n = S.qcount[QC_XS];
  const bool trace = S.trace_hash != 0;
  unsigned long long n_lookups = 0, n_events = 0, n_probes = 0;
  EvCount ca;
  // every lane of a wavefront runs the same number of iterations, so the
  // wavefront votes below see all of its lanes that hold a particle
  const int stride = gridDim.x * blockDim.x;
  const int first = blockIdx.x * blockDim.x + (threadIdx.x & ~63);
  for (int base = first; base < n; base += stride) {
    const int i = base + lane_id();
    bool need = false;
    int material = MATERIAL_VOID;
    int p = 0;
    PHot* h = nullptr;
    double E = 0.0;
    if (i < n) {
      p = q[i];
      n_events++;
      h = S.hot + p;
      material = h->material;
      if (h->wgt != 0.0) {
        if (material == MATERIAL_VOID) {
          h->xs_t = 0.0;
          h->xs_a = 0.0;
        } else if (material != h->material_last) {
          need = true;
          E = h->E;
        }
      }
    }
    // one pass per distinct material among the lanes that need cross
    // sections (one pass for almost every wavefront of a sorted queue)
    unsigned long long todo = __ballot(need);
    while (todo != 0ull) {
      const int leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
      const int mat = __builtin_amdgcn_readfirstlane(__shfl(material, leader));
      const bool mine = need && material == mat;
      if (!WAVE) {
        if (mine) {
          int probes = 0;
          n_lookups += particle_xs<G, LEARNED, ADV == 0>(P, S, p, h, mat, E, trace, probes);
          n_probes += probes;
        }
      } else {
        int probes = 0;
        n_lookups += particle_xs_wave<G>(P, S, p, h, mat, E, mine, trace, probes);
        n_probes += probes;
      }
      todo &= ~__ballot(mine);
    }
    if (ADV != 0 && i < n)
      S.adv_key[i] = advance_one<NL, ADV == 2>(P, S, p, ca);
  }
  wave_count(S, CNT_LOOKUPS, n_lookups);
  wave_count(S, CNT_XS_EVENTS, n_events);
  wave_count(S, CNT_SEARCH_PROBES, n_probes);
  if (ADV != 0) {
    wave_count(S, CNT_ADV_EVENTS, ca.events);
    wave_count(S, CNT_SEGMENTS, ca.segs);
    wave_count(S, CNT_LOST, ca.lost);
  }
}

template<int NL, int G, bool LEARNED, bool WAVE>
__global__ void __launch_bounds__(BLOCK) k_xs_lane(DevProblem P, DevParticles S, const int* q)
{
  xs_lane_body<NL, G, LEARNED, 0, WAVE>(P, S, q);
}

// The fused kernel is held to four wavefronts per SIMD (128 VGPRs; two of
// its values then live in scratch): measured faster than three waves with
// no scratch, its loads wait on the caches and more waves hide more of it.
template<int NL, bool LEARNED, int ADV, int G>
__global__ void __launch_bounds__(BLOCK, 4) k_xs_adv(DevProblem P, DevParticles S, const int* q)
{
  xs_lane_body<NL, G, LEARNED, ADV, false>(P, S, q);
}

//==============================================================================
// surface-crossing event (src/particle.cpp:288-329, 549-710)
//==============================================================================

// Particle::event_revive_from_secondary (src/particle.cpp:417-473) for a dead
// particle with banked secondaries: it continues as the last one it banked
// (Particle::from_source, src/particle.cpp:114-158) with the same RNG streams
// and n_event = 0.  The cell search of its next calculate_xs event is done
// here (it uses only r, u and the surface and draws no random number).
// Writes the particle's records; returns its new weight (0 if lost).
template<int NL>
__device__ __forceinline__ double revive_from_bank(const DevProblem& P, const DevParticles& S, int p, PWW& st,
  int& material, double& E, unsigned long long& n_lost)
{
  const PSite s = S.bank[static_cast<int64_t>(p) * S.bank_cap + (--st.n_bank)];
  Geo<NL> g;
  geo_clear(g, V3 {s.rx, s.ry, s.rz}, V3 {s.ux, s.uy, s.uz});
  g.surface = s.surface;
  double wgt = s.wgt;
  if (!exhaustive_find_cell(P, g) || g.lost) {
    wgt = 0.0;
    S.lost[p] = 1;
    n_lost++;
  }
  store_geo<NL>(S, p, g);
  PHot* hp = S.hot + p;
  hp->E = s.E;
  hp->wgt = wgt;
  hp->time = s.time;
  hp->n_event = 0;
  if (S.trace_hash) {
    uint64_t hh = S.hash[p];
    hmix(hh, 0x66);
    hmix(hh, dbits(s.wgt));
    hmix(hh, dbits(s.E));
    S.hash[p] = hh;
  }
  material = g.material;
  E = s.E;
  return wgt;
}

// Particle::event_cross_surface for slot p (src/particle.cpp:288-329):
// lattice crossing, boundary conditions, cell search, then the revive
// epilogue.  Leakage goes into `leak` (a single model) or the problem's
// slots (a problem set).  Returns whether the particle (or the secondary it
// was revived as) is alive; key is its calculate-xs sort key.
template<int NL, bool WW>
__device__ __forceinline__ bool surface_one(const DevProblem& P, const DevParticles& S, int p, EvCount& c,
  double& leak, int& key)
{
  c.events++;
  const PHot h = S.hot[p];
  Geo<NL> g;
  load_geo<NL>(P, S, p, h, g);
  double wgt = h.wgt;
  const int cell_last0 = g.cell[0];
  const int info = h.b_info;
  int trans[3] = {((info >> 8) & 3) - 1, ((info >> 10) & 3) - 1, ((info >> 12) & 3) - 1};
  g.surface = h.b_surface;
  g.n_coord = info & 0xff;
  if (g.n_coord < 1 || g.n_coord > NL) {
    g.n_coord = 1;
    g.lost = 1;
  }
  uint64_t hh = 0;
  const bool trace = S.trace_hash != 0;
  PartBase base {0, 0, 0};
  if (trace) {
    base = part_base(P, h.cell0);
    const int s_local = g.surface > 0 ? g.surface - base.surf : (g.surface < 0 ? g.surface + base.surf : 0);
    hh = S.hash[p];
    hmix(hh, 0x33);
    hmix(hh, static_cast<uint64_t>(static_cast<int64_t>(s_local)));
    hmix(hh, static_cast<uint64_t>((trans[0] + 1) | ((trans[1] + 1) << 2) | ((trans[2] + 1) << 4)));
  }
  if (g.lost) {
  } else if (trans[0] != 0 || trans[1] != 0 || trans[2] != 0) {
    cross_lattice(P, g, trans);
  } else if (g.surface == SURFACE_NONE || abs(g.surface) > P.n_surfaces) {
    g.lost = 1;
  } else {
    const DevSurface surf = P.surf[abs(g.surface) - 1];
    if (surf.bc == BC_VACUUM) {
      // cross_vacuum_bc (src/particle.cpp:631-655)
      if (P.n_parts > 1)
        unsafeAtomicAdd(&S.leak[(h.cell0 >= 0 ? P.cell[h.cell0].part : 0) * LEAK_PART_SLOTS +
                                (wave_slot() & (LEAK_PART_SLOTS - 1))],
          wgt);
      else
        leak += wgt;
      wgt = 0.0;
    } else if (surf.bc == BC_REFLECT) {
      // ReflectiveBC::handle_particle + cross_reflective_bc (src/boundary_condition.cpp:37-46,
      // src/particle.cpp:657-710)
      V3 u = reflect(g.u(), surf_normal(surf, g.r(0)));
      u = v_div(u, v_norm(u));
      if (g.n_coord != 1) {
        g.lost = 1;
      } else {
        g.set_u(u);
        g.cell[0] = cell_last0;
        g.surface = -g.surface;
        g.n_coord = 1;
        if (!neighbor_list_find_cell(P, g))
          g.lost = 1;
      }
    } else {
      // Particle::cross_surface (src/particle.cpp:597-628)
      if (!neighbor_list_find_cell(P, g)) {
        g.n_coord = 1;
        if (!exhaustive_find_cell(P, g)) {
          g.surface = SURFACE_NONE;
          g.n_coord = 1;
          g.set_r(0, v_add(g.r(0), v_scale(g.u(), TINY_BIT)));
          if (!exhaustive_find_cell(P, g))
            g.lost = 1;
        }
      }
    }
  }
  if (g.lost) {
    wgt = 0.0;
    S.lost[p] = 1;
    c.lost++;
  }
  PWW st;
  if (WW) {
    st = S.ww[p];
    // Particle::event_cross_surface: weight windows after a surface
    // (not lattice) crossing when that checkpoint is on (src/particle.cpp:320-322)
    const bool lattice = trans[0] != 0 || trans[1] != 0 || trans[2] != 0;
    if (P.ww.ckpt_surface && !lattice) {
      uint64_t seed = h.seed;
      const SplitSink sink {S.bank + static_cast<int64_t>(p) * S.bank_cap, S.bank_cap, S.status};
      apply_weight_windows(P.ww, st, wgt, g.r(0), g.u(), h.E, h.time, g.surface, &seed, sink);
      S.hot[p].seed = seed;
    }
  }
  if (trace) {
    hmix(hh, dbits(wgt));
    hmix(hh, static_cast<uint64_t>(static_cast<int64_t>(wgt != 0.0 ? lowest_cell(g) - base.cell : -1)));
    if (WW)
      hmix(hh, static_cast<uint64_t>(st.n_bank));
    S.hash[p] = hh;
  }
  // event_revive_from_secondary (src/particle.cpp:417-473)
  int n_event = h.n_event + 1;
  if (n_event == P.max_events)
    wgt = 0.0;
  store_geo<NL>(S, p, g);
  S.hot[p].n_event = n_event;
  S.hot[p].wgt = wgt;
  int mat = g.material;
  double E = h.E;
  if (WW) {
    if (wgt == 0.0 && st.n_bank > 0)
      wgt = revive_from_bank<NL>(P, S, p, st, mat, E, c.lost);
    S.ww[p] = st;
  }
  key = sort_key(S, mat, E);
  return wgt != 0.0;
}

template<int NL, bool WW>
__global__ void __launch_bounds__(BLOCK) k_surface(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_SURF];
  EvCount c;
  double leak = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = S.q_sc[i];
    int key = 0;
    const bool push = surface_one<NL, WW>(P, S, p, c, leak, key);
    S.stage_p[i] = p;
    S.stage_key[i] = push ? key : -1;
  }
  wave_count(S, CNT_SURF_EVENTS, c.events);
  wave_count(S, CNT_LOST, c.lost);
  for (int o = 32; o > 0; o >>= 1)
    leak += __shfl_xor(leak, o);
  if (lane_id() == 0 && leak != 0.0)
    unsafeAtomicAdd(&S.leak[wave_slot()], leak);
}

//==============================================================================
// collision event (src/particle.cpp:331-415, src/physics.cpp:43-168, 494-1041)
//==============================================================================

// sample_cxs_target_velocity (src/physics.cpp:990-1041)
__device__ __forceinline__ V3 sample_cxs_target_velocity(double awr, double E, V3 u, double kT, uint64_t* seed)
{
  double beta_vn = sqrt(awr * E / kT);
  double alpha = 1.0 / (1.0 + sqrt(PI) * beta_vn / 2.0);
  double beta_vt_sq = 0.0;
  double mu = 0.0;
  // The rejection loop accepts with probability > 1/2 per try for finite
  // inputs; the bound only guarantees termination on non-finite ones.
  for (int tries = 0; tries < (1 << 20); ++tries) {
    double r1 = prn(seed);
    double r2 = prn(seed);
    if (prn(seed) < alpha) {
      beta_vt_sq = -detm::log(r1 * r2);
    } else {
      double c = detm::cos(PI / 2.0 * prn(seed));
      beta_vt_sq = -detm::log(r1) - detm::log(r2) * c * c;
    }
    double beta_vt = sqrt(beta_vt_sq);
    mu = uniform_distribution(-1., 1., seed);
    double accept_prob =
      sqrt(beta_vn * beta_vn + beta_vt_sq - 2 * beta_vn * beta_vt * mu) / (beta_vn + beta_vt);
    if (prn(seed) < accept_prob)
      break;
  }
  double vt = sqrt(beta_vt_sq * kT / awr);
  return v_scale(rotate_angle(u, mu, seed), vt);
}

// Particle::event_collide for slot p (src/particle.cpp:331-415): sample the
// nuclide from the particle's micro_t row, analog absorption, elastic
// scattering with free-gas target motion, weight windows, then the revive
// epilogue.  Returns whether the particle (or its revived secondary) is
// alive; key is its calculate-xs sort key.  RECOMP: the micro totals are
// looked up again instead of read from the row (no rows are written then).
template<int NL, bool WW, bool RECOMP>
__device__ __forceinline__ bool collision_one(const DevProblem& P, const DevParticles& S, int p, EvCount& c,
  int& key)
{
  c.events++;
  const PHot h = S.hot[p];
  double E = h.E;
  double wgt = h.wgt;
  uint64_t seed = h.seed;
  const int material = h.material;
  V3 u = {h.ux, h.uy, h.uz};
  // sample_nuclide (src/physics.cpp:494-518)
  double cutoff = prn(&seed) * h.xs_t;
  const int j0 = P.mat_off[material];
  const int j1 = P.mat_off[material + 1];
  const double lx = detm::log(E / P.energy_min);
  int i_log = static_cast<int>(lx / P.log_spacing);
  i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
  double prob = 0.0;
  int i_nuclide = -1;
  int j_nuclide = -1; // its entry in the material's list
  Micro m {};
  int probes = 0;
  if constexpr (RECOMP) {
    // the micro totals again, CG nuclides at a time (their searches
    // interleaved), until the sum passes the cutoff: the same lookups as
    // calculate-xs, so the same values.  The lanes of one material at a time
    // (one pass for almost every wavefront of the sorted collision queue), so
    // that the material's nuclide records are the same in every active lane
    // and are read through the scalar cache.
    constexpr int CG = 2; // measured: 2 beats 1 and 4 (4 costs 195 VGPRs)
    const bool learned = S.use_learned != 0;
    unsigned long long todo = __ballot(1);
    while (todo != 0ull) {
      const int leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
      const int mat = __builtin_amdgcn_readfirstlane(__shfl(material, leader));
      const bool mine = material == mat;
      if (mine) {
        const int u0 = uniform_load(P.mat_off + mat);
        const int un = uniform_load(P.mat_off + mat + 1) - u0;
        for (int jj = 0; jj < un && i_nuclide < 0; jj += CG) {
          const int kk = un - jj < CG ? un - jj : CG;
          Micro mg[CG];
          int pr = 0;
          if (learned)
            nuclide_xs_group<CG, true>(P, P.matnuc + u0 + jj, kk, E, i_log, lx, pr, mg);
          else
            nuclide_xs_group<CG, false>(P, P.matnuc + u0 + jj, kk, E, i_log, lx, pr, mg);
          probes += pr;
#pragma unroll
          for (int k = 0; k < CG; ++k) {
            if (k < kk && i_nuclide < 0) {
              const int j = u0 + jj + k;
              const double dens = uniform_load(&P.matnuc[j].dens) * 1.0;
              prob += dens * mg[k].total;
              if (prob >= cutoff) {
                i_nuclide = uniform_load(&P.matnuc[j].nuc);
                j_nuclide = j;
                m = mg[k];
              }
            }
          }
        }
      }
      todo &= ~__ballot(mine);
    }
  } else {
    // the row holds the running sums of density x micro total in material
    // order, the very values the reference's loop accumulates
    // (src/physics.cpp:494-518): the sampled nuclide is the first whose sum
    // reaches the cutoff, found by a binary search (the sums never decrease)
    const double* row = S.micro_t + static_cast<int64_t>(p) * S.max_nn;
    int lo = 0, cnt = j1 - j0;
    while (cnt > 0) {
      const int step = cnt >> 1;
      if (row[lo + step] < cutoff) {
        lo += step + 1;
        cnt -= step + 1;
      } else {
        cnt = step;
      }
    }
    if (lo < j1 - j0) {
      i_nuclide = P.mat_nuc[j0 + lo];
      j_nuclide = j0 + lo;
    }
    // micro cross sections of the sampled nuclide at this energy
    if (i_nuclide >= 0)
      m = nuclide_xs(P, i_nuclide, E, i_log, lx, S.use_learned != 0, probes);
  }
  bool absorbed = false;
  if (i_nuclide < 0) {
    wgt = 0.0;
    S.lost[p] = 1;
    c.lost++;
  } else {
    c.lookups++;
    c.probes += probes;
    // absorption (src/physics.cpp:638-672), analog
    if (m.absorption > 0.0) {
      if (m.absorption > prn(&seed) * m.total) {
        wgt = 0.0;
        absorbed = true;
      }
    }
    if (!absorbed) {
      // scatter (src/physics.cpp:674-752): the data hold elastic scattering
      // only.  Nuclide::calculate_elastic_xs (src/nuclide.cpp:578-590) at the
      // sampled nuclide's grid index and interpolation factor decides the
      // branch; with sigma_t = sigma_el + sigma_c in the data it exceeds the
      // cutoff except for a cutoff within rounding of sigma_t - sigma_a, a
      // case the reference sends to its inelastic loop with no inelastic
      // reactions -- counted here and scattered elastically.
      const double cut = prn(&seed) * (m.total - m.absorption);
      const DevNuclide nd = P.nuc[i_nuclide];
      const double elastic = (1.0 - m.f) * P.EL[nd.off + m.i_grid] + m.f * P.EL[nd.off + m.i_grid + 1];
      if (!(elastic > cut))
        c.fallthrough++;
      const double awr = nd.awr;
      const double kT = nd.kT;
      // elastic_scatter (src/physics.cpp:754-818)
      double vel = sqrt(E);
      V3 v_n = v_scale(u, vel);
      V3 v_t = {0.0, 0.0, 0.0};
      // sample_target_velocity (src/physics.cpp:836-988), non-resonant nuclide
      if (!(E >= P.free_gas_threshold * kT && awr > 1.0))
        v_t = sample_cxs_target_velocity(awr, E, u, kT, &seed);
      V3 v_cm = v_div(v_add(v_n, v_scale(v_t, awr)), awr + 1.0);
      v_n = v_sub(v_n, v_cm);
      vel = v_norm(v_n);
      double mu_cm = uniform_distribution(-1., 1., &seed);
      V3 u_cm = v_div(v_n, vel);
      v_n = v_scale(rotate_angle(u_cm, mu_cm, &seed), vel);
      v_n = v_add(v_n, v_cm);
      E = v_dot(v_n, v_n);
      vel = sqrt(E);
      u = v_div(v_n, vel);
    }
  }
  PWW st;
  if (WW) {
    st = S.ww[p];
    // collision(): weight windows after the reaction (src/physics.cpp:64-65)
    if (P.ww.ckpt_collision) {
      const SplitSink sink {S.bank + static_cast<int64_t>(p) * S.bank_cap, S.bank_cap, S.status};
      apply_weight_windows(P.ww, st, wgt, V3 {h.rx, h.ry, h.rz}, u, E, h.time, SURFACE_NONE, &seed, sink);
    }
  }
  if (S.trace_hash) {
    uint64_t hh = S.hash[p];
    hmix(hh, 0x44);
    hmix(hh, static_cast<uint64_t>(static_cast<int64_t>(P.n_parts > 1 && j_nuclide >= 0 ?
                                                            P.mat_nuc_local[j_nuclide] : i_nuclide)));
    hmix(hh, absorbed ? 101 : 2);
    hmix(hh, dbits(E));
    hmix(hh, dbits(u.x));
    hmix(hh, dbits(u.y));
    hmix(hh, dbits(u.z));
    if (WW) {
      hmix(hh, dbits(wgt));
      hmix(hh, static_cast<uint64_t>(st.n_bank));
    }
    S.hash[p] = hh;
  }
  if (E < P.energy_cutoff)
    wgt = 0.0;
  // Particle::event_collide epilogue (src/particle.cpp:388-406); the
  // direction of the lower levels is the shared one
  int n_event = h.n_event + 1;
  if (n_event == P.max_events)
    wgt = 0.0;
  PHot* hp = S.hot + p;
  hp->surface = SURFACE_NONE;
  hp->material_last = C_NONE;
  hp->ux = u.x;
  hp->uy = u.y;
  hp->uz = u.z;
  hp->n_event = n_event;
  hp->E = E;
  hp->wgt = wgt;
  hp->seed = seed;
  int mat = material;
  if (WW) {
    if (wgt == 0.0 && st.n_bank > 0)
      wgt = revive_from_bank<NL>(P, S, p, st, mat, E, c.lost);
    S.ww[p] = st;
  }
  key = sort_key(S, mat, E);
  return wgt != 0.0;
}

// Held to four wavefronts per SIMD (128 VGPRs), which it fits without
// scratch; measured faster than the three its registers allowed otherwise.
template<int NL, bool WW, bool RECOMP>
__global__ void __launch_bounds__(BLOCK, 4) k_collision(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_COLL];
  const int n_surf = S.qcount[QC_SURF]; // collisions follow the crossings in q_sc
  EvCount c;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = S.q_sc[n_surf + i];
    int key = 0;
    const bool push = collision_one<NL, WW, RECOMP>(P, S, p, c, key);
    S.stage_p[n_surf + i] = p;
    S.stage_key[n_surf + i] = push ? key : -1;
  }
  wave_count(S, CNT_COLL_EVENTS, c.events);
  wave_count(S, CNT_COLLISIONS, c.events);
  wave_count(S, CNT_LOST, c.lost);
  wave_count(S, CNT_LOOKUPS, c.lookups);
  wave_count(S, CNT_SEARCH_PROBES, c.probes);
  wave_count(S, CNT_SCATTER_FALLTHROUGH, c.fallthrough);
}

//==============================================================================
// History-based tail: one lane per remaining particle
//==============================================================================

// Particle::event_calculate_xs for slot p by one lane on its own (the
// material need not be the same across the wavefront): the material's
// nuclides one at a time, sums in material order, as k_xs_lane computes them.
template<bool LEARNED>
__device__ __forceinline__ void xs_one(const DevProblem& P, const DevParticles& S, int p, EvCount& c)
{
  PHot* h = S.hot + p;
  c.events++;
  const int material = h->material;
  if (h->wgt == 0.0)
    return;
  if (material == MATERIAL_VOID) {
    *reinterpret_cast<double2*>(&h->xs_t) = double2 {0.0, 0.0};
    return;
  }
  if (material == h->material_last)
    return;
  const double E = h->E;
  const double lx = detm::log(E / P.energy_min);
  int i_log = static_cast<int>(lx / P.log_spacing);
  i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
  const int j0 = P.mat_off[material];
  const int nn = P.mat_off[material + 1] - j0;
  double* micro = S.micro_t + static_cast<int64_t>(p) * S.max_nn;
  const bool trace = S.trace_hash != 0;
  uint64_t hh = 0;
  if (trace) {
    hh = S.hash[p];
    hmix(hh, 0x11);
    hmix(hh, static_cast<uint64_t>(material - part_base(P, h->cell0).mat));
  }
  double xs_t = 0.0;
  double xs_a = 0.0;
  int probes = 0;
  // HG nuclides at a time, their searches interleaved: the lane is alone on
  // its particle, so its own memory-level parallelism is what hides latency
  constexpr int HG = 4;
  for (int jj = 0; jj < nn; jj += HG) {
    const int kk = nn - jj < HG ? nn - jj : HG;
    Micro m[HG];
    nuclide_xs_group<HG, LEARNED, false>(P, P.matnuc + j0 + jj, kk, E, i_log, lx, probes, m);
    double run[HG];
#pragma unroll
    for (int k = 0; k < HG; ++k) {
      if (k < kk) {
        const double dens = P.matnuc[j0 + jj + k].dens * 1.0;
        xs_t += dens * m[k].total;
        xs_a += dens * m[k].absorption;
        run[k] = xs_t;
        if (trace)
          hmix(hh, static_cast<uint64_t>(static_cast<uint32_t>(m[k].i_grid)));
      }
    }
    if (!S.recompute_micro)
      store_row<HG>(micro + jj, run, kk);
  }
  *reinterpret_cast<double2*>(&h->xs_t) = double2 {xs_t, xs_a};
  if (trace)
    S.hash[p] = hh;
  c.lookups += nn;
  c.probes += probes;
}

// The particles staged at the end of an event cycle (stage_p / stage_key,
// QC_STAGE entries, key -1 for done), each run by one lane through whole
// cycles -- calculate-xs, advance, then crossing or collision, and its
// revived secondaries -- for at most max_steps cycles.  The event functions
// are the event kernels' own, so the histories are the same.  A particle
// still alive keeps its entry (with its new key) for the next launch; the
// count of those is added to QC_LIVE.  Used for the tail of a batch, where
// few particles are left and a cycle of event kernels costs its launches.
template<int NL, bool WW, bool MULTI, bool LEARNED>
__global__ void __launch_bounds__(BLOCK) k_history(DevProblem P, DevParticles S, int max_steps)
{
  const int n = S.qcount[QC_STAGE];
  EvCount cx, ca, cs, cc;
  double leak = 0.0;
  unsigned long long live = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int key = S.stage_key[i];
    if (key < 0)
      continue;
    const int p = S.stage_p[i];
    bool alive = true;
    for (int step = 0; step < max_steps; ++step) {
      xs_one<LEARNED>(P, S, p, cx);
      const int outcome = advance_one<NL, MULTI>(P, S, p, ca);
      if (outcome < 0) {
        alive = false;
        break;
      }
      if (outcome == 0)
        alive = surface_one<NL, WW>(P, S, p, cs, leak, key);
      else if (S.recompute_micro)
        alive = collision_one<NL, WW, true>(P, S, p, cc, key);
      else
        alive = collision_one<NL, WW, false>(P, S, p, cc, key);
      if (!alive)
        break;
    }
    S.stage_key[i] = alive ? key : -1;
    live += alive ? 1 : 0;
  }
  wave_count(S, CNT_XS_EVENTS, cx.events);
  wave_count(S, CNT_ADV_EVENTS, ca.events);
  wave_count(S, CNT_SURF_EVENTS, cs.events);
  wave_count(S, CNT_COLL_EVENTS, cc.events);
  wave_count(S, CNT_COLLISIONS, cc.events);
  wave_count(S, CNT_SEGMENTS, ca.segs);
  wave_count(S, CNT_LOST, ca.lost + cs.lost + cc.lost);
  wave_count(S, CNT_LOOKUPS, cx.lookups + cc.lookups);
  wave_count(S, CNT_SEARCH_PROBES, cx.probes + cc.probes);
  wave_count(S, CNT_HIST_XS_EVENTS, cx.events);
  wave_count(S, CNT_HIST_LOOKUPS, cx.lookups);
  wave_count(S, CNT_HIST_ADV_EVENTS, ca.events);
  wave_count(S, CNT_HIST_SEGMENTS, ca.segs);
  wave_count(S, CNT_HIST_SURF_EVENTS, cs.events);
  wave_count(S, CNT_HIST_COLL_EVENTS, cc.events);
  wave_count(S, CNT_SCATTER_FALLTHROUGH, cc.fallthrough);
  for (int o = 32; o > 0; o >>= 1) {
    leak += __shfl_xor(leak, o);
    live += __shfl_xor(live, o);
  }
  if (lane_id() == 0) {
    if (leak != 0.0)
      unsafeAtomicAdd(&S.leak[wave_slot()], leak);
    if (live != 0)
      atomicAdd(&S.qcount[QC_LIVE], static_cast<int>(live));
  }
}

//==============================================================================
// Tally accumulation (src/tallies/tally.cpp:821-878)
//==============================================================================

// Tally::accumulate (src/tallies/tally.cpp:821-878) for one tally whose
// VALUE has n_copies copies of its bins: the copies are summed in copy
// order, then SUM += VALUE * norm and SUM_SQ += (VALUE * norm)^2, and VALUE
// is cleared.
__global__ void __launch_bounds__(BLOCK) k_accumulate(double* value, double* results, int64_t n_bins, int n_copies,
  double norm)
{
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_bins;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double sum = value[i];
    value[i] = 0.0;
    for (int c = 1; c < n_copies; ++c) {
      sum += value[c * n_bins + i];
      value[c * n_bins + i] = 0.0;
    }
    double val = sum * norm;
    results[3 * i + 1] += val;
    results[3 * i + 2] += val * val;
  }
}

//==============================================================================
// Host launch wrappers
//==============================================================================

namespace {

template<int NL>
void init_nl(const DevProblem& P, const DevParticles& S, int64_t n, int64_t g_base, int64_t batch, hipStream_t st)
{
  int blocks = static_cast<int>((n + BLOCK - 1) / BLOCK);
  if (P.n_parts > 1)
    hipLaunchKernelGGL((k_init<NL, true>), dim3(blocks < 1 ? 1 : blocks), dim3(BLOCK), 0, st, P, S, n, g_base, batch);
  else
    hipLaunchKernelGGL((k_init<NL, false>), dim3(blocks < 1 ? 1 : blocks), dim3(BLOCK), 0, st, P, S, n, g_base,
      batch);
}

template<int NL, int LPP>
void xs_nl(const DevProblem& P, const DevParticles& S, const int* q, int blocks, hipStream_t st)
{
  size_t lds = static_cast<size_t>(BLOCK / LPP) * (P.max_mat_nuc > 0 ? P.max_mat_nuc : 1) *
               (2 * sizeof(double) + sizeof(int));
  // the work per pass is BLOCK / LPP particles; scale the grid to keep the
  // same particle coverage per launch
  int b = blocks * (LPP < 8 ? LPP : 8);
  hipLaunchKernelGGL((k_xs<NL, LPP>), dim3(b), dim3(BLOCK), lds, st, P, S, q);
}

template<int NL, bool LEARNED, bool WAVE>
void xs_lane_l(const DevProblem& P, const DevParticles& S, const int* q, int blocks, int tpb, int group,
  hipStream_t st)
{
  switch (group) {
  case 1: hipLaunchKernelGGL((k_xs_lane<NL, 1, LEARNED, WAVE>), dim3(blocks), dim3(tpb), 0, st, P, S, q); break;
  case 2: hipLaunchKernelGGL((k_xs_lane<NL, 2, LEARNED, WAVE>), dim3(blocks), dim3(tpb), 0, st, P, S, q); break;
  case 8: hipLaunchKernelGGL((k_xs_lane<NL, 8, LEARNED, WAVE>), dim3(blocks), dim3(tpb), 0, st, P, S, q); break;
  default: hipLaunchKernelGGL((k_xs_lane<NL, 4, LEARNED, WAVE>), dim3(blocks), dim3(tpb), 0, st, P, S, q); break;
  }
}

// calculate-xs and advance in one launch (xs_group 4, one lane per particle)
template<int NL, int G>
void xs_adv_g(const DevProblem& P, const DevParticles& S, const int* q, int blocks, int tpb, hipStream_t st)
{
  if (P.n_parts > 1) {
    if (S.use_learned)
      hipLaunchKernelGGL((k_xs_adv<NL, true, 2, G>), dim3(blocks), dim3(tpb), 0, st, P, S, q);
    else
      hipLaunchKernelGGL((k_xs_adv<NL, false, 2, G>), dim3(blocks), dim3(tpb), 0, st, P, S, q);
  } else {
    if (S.use_learned)
      hipLaunchKernelGGL((k_xs_adv<NL, true, 1, G>), dim3(blocks), dim3(tpb), 0, st, P, S, q);
    else
      hipLaunchKernelGGL((k_xs_adv<NL, false, 1, G>), dim3(blocks), dim3(tpb), 0, st, P, S, q);
  }
}

template<int NL>
void xs_adv(const DevProblem& P, const DevParticles& S, const int* q, int blocks, int tpb, int G, hipStream_t st)
{
  if (G == 2)
    xs_adv_g<NL, 2>(P, S, q, blocks, tpb, st);
  else
    xs_adv_g<NL, 4>(P, S, q, blocks, tpb, st);
}

template<int NL>
void xs_lane(const DevProblem& P, const DevParticles& S, const int* q, int blocks, int tpb, int group,
  hipStream_t st)
{
  if (S.use_learned)
    xs_lane_l<NL, true, false>(P, S, q, blocks, tpb, group, st);
  else if (S.wave_search)
    xs_lane_l<NL, false, true>(P, S, q, blocks, tpb, group, st);
  else
    xs_lane_l<NL, false, false>(P, S, q, blocks, tpb, group, st);
}

template<int NL>
void xs_lpp(const DevProblem& P, const DevParticles& S, const int* q, const LaunchConfig& lc, hipStream_t st)
{
  const int blocks = lc.blocks;
  const int tpb = lc.block == 64 || lc.block == 128 ? lc.block : BLOCK;
  switch (lc.xs_lanes) {
  case 1: xs_lane<NL>(P, S, q, blocks, tpb, lc.xs_group, st); break;
  case 2: xs_nl<NL, 2>(P, S, q, blocks, st); break;
  case 4: xs_nl<NL, 4>(P, S, q, blocks, st); break;
  case 8: xs_nl<NL, 8>(P, S, q, blocks, st); break;
  case 16: xs_nl<NL, 16>(P, S, q, blocks, st); break;
  case 32: xs_nl<NL, 32>(P, S, q, blocks, st); break;
  default: xs_nl<NL, 64>(P, S, q, blocks, st); break;
  }
}

void launch_partition(const DevParticles& S, const Part& A, const LaunchConfig& lc, hipStream_t st)
{
  const int sb = lc.sort_blocks < 1 ? 1 : (lc.sort_blocks > SORT_MAX_BLOCKS ? SORT_MAX_BLOCKS : lc.sort_blocks);
  const int col_blocks = (A.nb + 63) / 64;
  hipLaunchKernelGGL(k_part_count, dim3(sb), dim3(BLOCK), 0, st, S, A);
  hipLaunchKernelGGL(k_part_cols, dim3(col_blocks), dim3(64), 0, st, S, A, sb);
  hipLaunchKernelGGL(k_part_scan, dim3(1), dim3(1024), 0, st, S, A);
  hipLaunchKernelGGL(k_part_scatter, dim3(sb), dim3(BLOCK), 0, st, S, A);
}

bool fused_advance(const LaunchConfig& lc)
{
  return lc.fuse_advance && lc.xs_lanes == 1 && (lc.xs_group == 4 || lc.xs_group == 2);
}

template<int NL>
void kernel_nl(int which, const DevProblem& P, const DevParticles& S, const LaunchConfig& lc, hipStream_t st,
  int cycle)
{
  const int tpb = lc.block == 64 || lc.block == 128 ? lc.block : BLOCK;
  switch (which) {
  case KERNEL_SORT: {
    // staged survivors -> q_xs, sorted by (material, energy bucket) key, by
    // material only on the cycles between full sorts, or only compacted
    const bool full = lc.sort_every <= 1 || cycle % lc.sort_every == 0;
    Part A {S.stage_p, S.stage_key, S.qcount + QC_STAGE, S.q_xs, S.qcount + QC_XS, nullptr,
      !lc.sort ? 1 : (full ? S.sort_nb : 1 + (S.sort_nb - 1) / S.sort_eb), full ? 0 : S.sort_eb};
    launch_partition(S, A, lc, st);
    break;
  }
  case KERNEL_XS:
    if (fused_advance(lc))
      xs_adv<NL>(P, S, S.q_xs, lc.blocks, tpb, lc.xs_group, st);
    else
      xs_lpp<NL>(P, S, S.q_xs, lc, st);
    break;
  case KERNEL_ADVANCE:
    if (fused_advance(lc))
      break; // done by the calculate-xs launch
    if (P.n_parts > 1)
      hipLaunchKernelGGL((k_advance<NL, true>), dim3(lc.blocks), dim3(tpb), 0, st, P, S, S.q_xs);
    else
      hipLaunchKernelGGL((k_advance<NL, false>), dim3(lc.blocks), dim3(tpb), 0, st, P, S, S.q_xs);
    break;
  case KERNEL_SPLIT: {
    // advance outcomes -> q_sc = [surface crossings | collisions]; the two
    // counts land in QC_SURF and QC_COLL, their sum in QC_STAGE
    Part A {S.q_xs, S.adv_key, S.qcount + QC_XS, S.q_sc, S.qcount + QC_STAGE, S.qcount + QC_SURF, 2, 0};
    launch_partition(S, A, lc, st);
    break;
  }
  case KERNEL_SURFACE:
    if (P.ww.on)
      hipLaunchKernelGGL((k_surface<NL, true>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    else
      hipLaunchKernelGGL((k_surface<NL, false>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    break;
  default:
    if (P.ww.on && S.recompute_micro)
      hipLaunchKernelGGL((k_collision<NL, true, true>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    else if (P.ww.on)
      hipLaunchKernelGGL((k_collision<NL, true, false>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    else if (S.recompute_micro)
      hipLaunchKernelGGL((k_collision<NL, false, true>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    else
      hipLaunchKernelGGL((k_collision<NL, false, false>), dim3(lc.blocks), dim3(tpb), 0, st, P, S);
    break;
  }
}

} // namespace

void launch_init(const DevProblem& P, const DevParticles& S, int64_t n, int64_t g_base, int64_t batch,
  hipStream_t st)
{
  switch (P.n_levels) {
  case 1: init_nl<1>(P, S, n, g_base, batch, st); break;
  case 2: init_nl<2>(P, S, n, g_base, batch, st); break;
  case 3: init_nl<3>(P, S, n, g_base, batch, st); break;
  default: init_nl<4>(P, S, n, g_base, batch, st); break;
  }
}

void launch_kernel(int which, const DevProblem& P, const DevParticles& S, const LaunchConfig& lc, hipStream_t st,
  int cycle)
{
  switch (P.n_levels) {
  case 1: kernel_nl<1>(which, P, S, lc, st, cycle); break;
  case 2: kernel_nl<2>(which, P, S, lc, st, cycle); break;
  case 3: kernel_nl<3>(which, P, S, lc, st, cycle); break;
  default: kernel_nl<4>(which, P, S, lc, st, cycle); break;
  }
}

void launch_cycle(const DevProblem& P, const DevParticles& S, const LaunchConfig& lc, hipStream_t st, int cycle)
{
  for (int k = KERNEL_SORT; k <= KERNEL_COLLISION; ++k)
    launch_kernel(k, P, S, lc, st, cycle);
}

namespace {
template<int NL, bool WW, bool MULTI>
void history_m(const DevProblem& P, const DevParticles& S, int n, int max_steps, hipStream_t st)
{
  const int blocks = (n + BLOCK - 1) / BLOCK;
  if (S.use_learned)
    hipLaunchKernelGGL((k_history<NL, WW, MULTI, true>), dim3(blocks < 1 ? 1 : blocks), dim3(BLOCK), 0, st, P, S,
      max_steps);
  else
    hipLaunchKernelGGL((k_history<NL, WW, MULTI, false>), dim3(blocks < 1 ? 1 : blocks), dim3(BLOCK), 0, st, P, S,
      max_steps);
}

template<int NL>
void history_nl(const DevProblem& P, const DevParticles& S, int n, int max_steps, hipStream_t st)
{
  if (P.ww.on)
    history_m<NL, true, false>(P, S, n, max_steps, st);
  else if (P.n_parts > 1)
    history_m<NL, false, true>(P, S, n, max_steps, st);
  else
    history_m<NL, false, false>(P, S, n, max_steps, st);
}
} // namespace

void launch_history(const DevProblem& P, const DevParticles& S, int n_staged, int max_steps, hipStream_t st)
{
  switch (P.n_levels) {
  case 1: history_nl<1>(P, S, n_staged, max_steps, st); break;
  case 2: history_nl<2>(P, S, n_staged, max_steps, st); break;
  case 3: history_nl<3>(P, S, n_staged, max_steps, st); break;
  default: history_nl<4>(P, S, n_staged, max_steps, st); break;
  }
}

void launch_accumulate(double* value, double* results, int64_t n_bins, int n_copies, double norm, hipStream_t st)
{
  int64_t b = (n_bins + BLOCK - 1) / BLOCK;
  int blocks = static_cast<int>(b > 4096 ? 4096 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(BLOCK), 0, st, value, results, n_bins, n_copies, norm);
}

} // namespace omc
