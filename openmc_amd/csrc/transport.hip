// Event kernels of the GPU transport loop (gfx950).
//
// One batch runs as: k_init (source sampling, src/simulation.cpp:573-649),
// then cycles of  k_xs -> k_advance -> k_surface -> k_collision  until no
// particle is alive.  The calculate-xs queue doubles as the advance queue
// (every particle that computes cross sections advances next,
// src/event.cpp:87-101); advance splits it into the surface-crossing and
// collision queues; both of those refill the calculate-xs queue with the
// survivors.  Appends are compacted per wavefront: a ballot of the lanes that
// append, one atomicAdd per wavefront for the base slot, and each lane's rank
// among the set lanes (popcount of the lower mask bits) as its offset.
//
// Queue sizes live in device memory, so a cycle needs no host round trip and
// a run of cycles can be captured in one hipGraph.  Each counter is reset by
// a kernel that neither reads nor appends to it:
//   k_xs       reads XS,   sets ADV = XS, resets SURF and COLL
//   k_advance  reads ADV,  resets XS, appends SURF / COLL
//   k_surface  reads SURF, appends XS
//   k_collision reads COLL, appends XS
#include <hip/hip_runtime.h>

#include "device_physics.h"
#include "transport.h"

namespace omc {

constexpr int BLOCK = 256;

__device__ __forceinline__ int lane_id()
{
  return __lane_id();
}

// Wavefront-compacted append of `val` to queue `q` for lanes with pred set.
// Must be reached by every lane of the wavefront.
__device__ __forceinline__ void wave_push(int* q, int* cnt, bool pred, int val)
{
  unsigned long long mask = __ballot(pred);
  if (mask == 0ull)
    return;
  int lane = lane_id();
  int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  int base = 0;
  if (lane == leader)
    base = atomicAdd(cnt, __popcll(mask));
  base = __shfl(base, leader);
  if (pred) {
    unsigned long long below = lane == 0 ? 0ull : (mask & ((1ull << lane) - 1ull));
    q[base + __popcll(below)] = val;
  }
}

// Wavefront sum then one atomic per wavefront.  Every lane must call it.
__device__ __forceinline__ void wave_count(unsigned long long* ctr, unsigned long long v)
{
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  if (lane_id() == 0 && v != 0ull)
    atomicAdd(ctr, v);
}

__device__ inline void load_geo(const DevParticles& S, int nl, int64_t p, Geo& g)
{
  const int64_t cap = S.cap;
  for (int i = 0; i < MAX_LEVELS; ++i) {
    if (i < nl) {
      int64_t k = i * cap + p;
      g.r[i] = {S.rx[k], S.ry[k], S.rz[k]};
      g.u[i] = {S.ux[k], S.uy[k], S.uz[k]};
      g.cell[i] = S.cell[k];
      g.univ[i] = S.univ[k];
      g.lat[i] = S.lat[k];
      g.li[i][0] = S.lix[k];
      g.li[i][1] = S.liy[k];
      g.li[i][2] = S.liz[k];
    } else {
      g.r[i] = {0, 0, 0};
      g.u[i] = {0, 0, 0};
      coord_reset(g, i);
    }
  }
  g.n_coord = S.n_coord[p];
  g.surface = S.surface[p];
  g.material = S.material[p];
  g.material_last = S.material_last[p];
  g.lost = 0;
}

__device__ inline void store_geo(const DevParticles& S, int nl, int64_t p, const Geo& g)
{
  const int64_t cap = S.cap;
  for (int i = 0; i < MAX_LEVELS; ++i) {
    if (i < nl) {
      int64_t k = i * cap + p;
      S.rx[k] = g.r[i].x;
      S.ry[k] = g.r[i].y;
      S.rz[k] = g.r[i].z;
      S.ux[k] = g.u[i].x;
      S.uy[k] = g.u[i].y;
      S.uz[k] = g.u[i].z;
      S.cell[k] = g.cell[i];
      S.univ[k] = g.univ[i];
      S.lat[k] = g.lat[i];
      S.lix[k] = g.li[i][0];
      S.liy[k] = g.li[i][1];
      S.liz[k] = g.li[i][2];
    }
  }
  S.n_coord[p] = g.n_coord;
  S.surface[p] = g.surface;
  S.material[p] = g.material;
  S.material_last[p] = g.material_last;
}

//==============================================================================
// Source sampling (src/source.cpp:226-271, 355-417; src/simulation.cpp:573-649)
//==============================================================================

__device__ inline bool satisfies_spatial_constraints(const DevProblem& P, V3 r)
{
  Geo g;
  for (int i = 0; i < MAX_LEVELS; ++i) {
    g.r[i] = {0, 0, 0};
    g.u[i] = {0, 0, 0};
    coord_reset(g, i);
  }
  g.r[0] = r;
  g.u[0] = {0.0, 0.0, 1.0};
  g.n_coord = 1;
  g.surface = SURFACE_NONE;
  g.material = C_NONE;
  g.material_last = C_NONE;
  g.lost = 0;
  if (!exhaustive_find_cell(P, g))
    return false;
  bool accepted = true;
  if (P.src.n_domain > 0) {
    for (int i = 0; i < g.n_coord; i++) {
      int id = P.cell[g.cell[i]].id;
      accepted = false;
      for (int k = 0; k < P.src.n_domain; ++k)
        accepted = accepted || (P.src.domain_ids[k] == id);
      if (accepted)
        break;
    }
  }
  return accepted;
}

// maxwell_spectrum / watt_spectrum (src/random_dist.cpp:20-39)
__device__ inline double watt_spectrum(double a, double b, uint64_t* seed)
{
  double r1 = prn(seed);
  double r2 = prn(seed);
  double r3 = prn(seed);
  double c = detm::cos(PI / 2. * r3);
  double w = -a * (detm::log(r1) + detm::log(r2) * c * c);
  return w + 0.25 * a * a * b + uniform_distribution(-1., 1., seed) * sqrt(a * a * b * w);
}

constexpr int MAX_SOURCE_TRIES = 1 << 20;

__global__ void __launch_bounds__(BLOCK) k_init(DevProblem P, DevParticles S, int64_t n, int64_t id_base)
{
  const int nl = P.n_levels;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t id = id_base + i + 1;
    uint64_t seed = future_seed(static_cast<uint64_t>(id) * P.stride, P.master_seed + 1);
    V3 r = {0, 0, 0};
    bool lost = false;
    int tries = 0;
    bool accepted = false;
    while (!accepted) {
      if (P.src.space_type == 0) {
        r = {P.src.p0[0], P.src.p0[1], P.src.p0[2]};
      } else {
        double x1 = prn(&seed);
        double x2 = prn(&seed);
        double x3 = prn(&seed);
        V3 ll = {P.src.p0[0], P.src.p0[1], P.src.p0[2]};
        V3 ur = {P.src.p1[0], P.src.p1[1], P.src.p1[2]};
        V3 ext = v_sub(ur, ll);
        r = v_add(ll, V3 {x1 * ext.x, x2 * ext.y, x3 * ext.z});
      }
      accepted = satisfies_spatial_constraints(P, r);
      if (++tries >= MAX_SOURCE_TRIES && !accepted) {
        lost = true;
        break;
      }
    }
    V3 u;
    if (P.src.angle_type == 0) {
      // isotropic_direction (src/distribution_multi.cpp:106-112)
      double phi = uniform_distribution(0., 2.0 * PI, &seed);
      double mu = uniform_distribution(-1., 1., &seed);
      double s, c;
      detm::sincos(phi, &s, &c);
      u = {mu, sqrt(1.0 - mu * mu) * c, sqrt(1.0 - mu * mu) * s};
    } else {
      u = {P.src.u_ref[0], P.src.u_ref[1], P.src.u_ref[2]};
    }
    double E = 0.0;
    tries = 0;
    while (true) {
      if (P.src.energy_type == 0) {
        int nd = P.src.n_discrete;
        int k = 0;
        if (nd > 1) {
          size_t uu = static_cast<size_t>(prn(&seed) * nd);
          if (prn(&seed) < P.src.disc_prob[uu])
            k = static_cast<int>(uu);
          else
            k = P.src.disc_alias[uu];
        }
        E = P.src.disc_x[k];
      } else {
        E = watt_spectrum(P.src.watt_a, P.src.watt_b, &seed);
      }
      if (E < P.energy_max && (E > 0.0 && E < INFTY))
        break;
      if (++tries >= MAX_SOURCE_TRIES) {
        lost = true;
        break;
      }
    }
    // Particle::from_source (src/particle.cpp:114-158) + initialize_history
    const int64_t cap = S.cap;
    for (int l = 0; l < nl; ++l) {
      int64_t k = l * cap + i;
      S.cell[k] = C_NONE;
      S.univ[k] = C_NONE;
      S.lat[k] = C_NONE;
      S.lix[k] = 0;
      S.liy[k] = 0;
      S.liz[k] = 0;
    }
    S.rx[i] = r.x;
    S.ry[i] = r.y;
    S.rz[i] = r.z;
    S.ux[i] = u.x;
    S.uy[i] = u.y;
    S.uz[i] = u.z;
    S.n_coord[i] = 1;
    S.E[i] = E;
    S.wgt[i] = lost ? 0.0 : 1.0;
    S.time[i] = 0.0;
    S.seed[i] = future_seed(static_cast<uint64_t>(id) * P.stride, P.master_seed + 0);
    S.surface[i] = SURFACE_NONE;
    S.material[i] = C_NONE;
    S.material_last[i] = C_NONE;
    S.n_event[i] = 0;
    S.lost[i] = lost ? 1 : 0;
    S.hash[i] = 0xcbf29ce484222325ULL ^ static_cast<uint64_t>(id);
    S.q_xs[i] = static_cast<int>(i);
  }
}

//==============================================================================
// calculate_xs event (src/particle.cpp:160-230, src/material.cpp:829-901)
//==============================================================================

__global__ void __launch_bounds__(BLOCK) k_xs(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_XS];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S.qcount[QC_ADV] = n;
    S.qcount[QC_SURF] = 0;
    S.qcount[QC_COLL] = 0;
  }
  const int nl = P.n_levels;
  const int64_t cap = S.cap;
  unsigned long long n_lookups = 0, n_events = 0, n_lost = 0, n_probes = 0;
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    if (i < n) {
      const int p = S.q_xs[i];
      n_events++;
      int nc = S.n_coord[p];
      int material = S.material[p];
      int material_last = S.material_last[p];
      bool alive = true;
      if (S.cell[(nc - 1) * cap + p] == C_NONE) {
        Geo g;
        load_geo(S, nl, p, g);
        bool found = exhaustive_find_cell(P, g);
        if (!found || g.lost) {
          S.wgt[p] = 0.0;
          S.lost[p] = 1;
          n_lost++;
          alive = false;
        }
        store_geo(S, nl, p, g);
        material = g.material;
        material_last = g.material_last;
      }
      if (alive) {
        if (material != MATERIAL_VOID) {
          if (material != material_last) {
            const double E = S.E[p];
            const double lx = detm::log(E / P.energy_min);
            int i_log = static_cast<int>(lx / P.log_spacing);
            i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
            const int j0 = P.mat_off[material];
            const int j1 = P.mat_off[material + 1];
            double xs_t = 0.0;
            double xs_a = 0.0;
            uint64_t h = S.hash[p];
            const bool trace = S.trace_hash != 0;
            if (trace) {
              hmix(h, 0x11);
              hmix(h, static_cast<uint64_t>(material));
            }
            int probes = 0;
            const bool learned = S.use_learned != 0;
            for (int j = j0; j < j1; ++j) {
              Micro m = nuclide_xs(P, P.mat_nuc[j], E, i_log, lx, learned, probes);
              S.micro_t[static_cast<int64_t>(j - j0) * cap + p] = m.total;
              const double dens = P.mat_dens[j] * 1.0;
              xs_t += dens * m.total;
              xs_a += dens * m.absorption;
              if (trace)
                hmix(h, static_cast<uint64_t>(static_cast<uint32_t>(m.i_grid)));
            }
            n_lookups += static_cast<unsigned long long>(j1 - j0);
            n_probes += static_cast<unsigned long long>(probes);
            S.xs_t[p] = xs_t;
            S.xs_a[p] = xs_a;
            if (trace)
              S.hash[p] = h;
          }
        } else {
          S.xs_t[p] = 0.0;
          S.xs_a[p] = 0.0;
        }
      }
    }
  }
  wave_count(&S.counters[CNT_LOOKUPS], n_lookups);
  wave_count(&S.counters[CNT_XS_EVENTS], n_events);
  wave_count(&S.counters[CNT_LOST], n_lost);
  wave_count(&S.counters[CNT_SEARCH_PROBES], n_probes);
}

//==============================================================================
// advance event (src/particle.cpp:232-286, src/event.cpp:107-128)
//==============================================================================

__global__ void __launch_bounds__(BLOCK) k_advance(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_ADV];
  if (blockIdx.x == 0 && threadIdx.x == 0)
    S.qcount[QC_XS] = 0;
  const int nl = P.n_levels;
  unsigned long long n_events = 0, n_segs = 0, n_lost = 0;
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    bool to_surf = false, to_coll = false;
    int p = 0;
    if (i < n) {
      p = S.q_xs[i];
      double wgt = S.wgt[p];
      if (wgt != 0.0) {
        n_events++;
        Geo g;
        load_geo(S, nl, p, g);
        const double E = S.E[p];
        const double xs_t = S.xs_t[p];
        uint64_t seed = S.seed[p];
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
        V3 r0 = g.r[0];
        for (int j = 0; j < g.n_coord; ++j)
          g.r[j] = v_add(g.r[j], v_scale(g.u[j], distance));
        double dt = distance / speed;
        S.time[p] = S.time[p] + dt;
        if (P.tally.enabled)
          n_segs += score_tracklength(P.tally, r0, g.r[0], g.u[0], E, wgt, distance);
        if (distance == INFTY)
          wgt = 0.0;
        if (g.lost) {
          wgt = 0.0;
          S.lost[p] = 1;
          n_lost++;
        }
        const int64_t cap = S.cap;
        for (int j = 0; j < g.n_coord; ++j) {
          int64_t k = j * cap + p;
          S.rx[k] = g.r[j].x;
          S.ry[k] = g.r[j].y;
          S.rz[k] = g.r[j].z;
        }
        S.seed[p] = seed;
        S.b_surface[p] = b.surface;
        S.b_info[p] = b.coord_level | ((b.trans[0] + 1) << 8) | ((b.trans[1] + 1) << 10) |
                      ((b.trans[2] + 1) << 12);
        S.wgt[p] = wgt;
        if (S.trace_hash) {
          uint64_t h = S.hash[p];
          hmix(h, 0x22);
          hmix(h, dbits(distance));
          hmix(h, d_coll > b.distance ? 2 : 1);
          S.hash[p] = h;
        }
        if (wgt != 0.0) {
          to_surf = d_coll > b.distance;
          to_coll = !to_surf;
        }
      }
    }
    wave_push(S.q_surf, &S.qcount[QC_SURF], to_surf, p);
    wave_push(S.q_coll, &S.qcount[QC_COLL], to_coll, p);
  }
  wave_count(&S.counters[CNT_ADV_EVENTS], n_events);
  wave_count(&S.counters[CNT_SEGMENTS], n_segs);
  wave_count(&S.counters[CNT_LOST], n_lost);
}

//==============================================================================
// surface-crossing event (src/particle.cpp:288-329, 549-710)
//==============================================================================

__global__ void __launch_bounds__(BLOCK) k_surface(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_SURF];
  const int nl = P.n_levels;
  unsigned long long n_events = 0, n_lost = 0;
  double leak = 0.0;
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    bool push = false;
    int p = 0;
    if (i < n) {
      p = S.q_surf[i];
      n_events++;
      Geo g;
      load_geo(S, nl, p, g);
      double wgt = S.wgt[p];
      const int cell_last0 = g.cell[0];
      const int info = S.b_info[p];
      int trans[3] = {((info >> 8) & 3) - 1, ((info >> 10) & 3) - 1, ((info >> 12) & 3) - 1};
      g.surface = S.b_surface[p];
      g.n_coord = info & 0xff;
      uint64_t h = 0;
      const bool trace = S.trace_hash != 0;
      if (trace) {
        h = S.hash[p];
        hmix(h, 0x33);
        hmix(h, static_cast<uint64_t>(static_cast<int64_t>(g.surface)));
        hmix(h, static_cast<uint64_t>((trans[0] + 1) | ((trans[1] + 1) << 2) | ((trans[2] + 1) << 4)));
      }
      if (trans[0] != 0 || trans[1] != 0 || trans[2] != 0) {
        cross_lattice(P, g, trans);
      } else if (g.surface == SURFACE_NONE || abs(g.surface) > P.n_surfaces) {
        g.lost = 1;
      } else {
        const DevSurface surf = P.surf[abs(g.surface) - 1];
        if (surf.bc == BC_VACUUM) {
          // cross_vacuum_bc (src/particle.cpp:631-655)
          leak += wgt;
          wgt = 0.0;
        } else if (surf.bc == BC_REFLECT) {
          // ReflectiveBC::handle_particle + cross_reflective_bc (src/boundary_condition.cpp:37-46,
          // src/particle.cpp:657-710)
          V3 u = reflect(g.u[0], surf_normal(surf, g.r[0]));
          u = v_div(u, v_norm(u));
          if (g.n_coord != 1) {
            g.lost = 1;
          } else {
            g.u[0] = u;
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
              g.r[0] = v_add(g.r[0], v_scale(g.u[0], TINY_BIT));
              if (!exhaustive_find_cell(P, g))
                g.lost = 1;
            }
          }
        }
      }
      if (g.lost) {
        wgt = 0.0;
        S.lost[p] = 1;
        n_lost++;
      }
      if (trace) {
        hmix(h, dbits(wgt));
        hmix(h, static_cast<uint64_t>(static_cast<int64_t>(wgt != 0.0 ? g.cell[g.n_coord - 1] : -1)));
        S.hash[p] = h;
      }
      // event_revive_from_secondary (src/particle.cpp:417-473)
      int n_event = S.n_event[p] + 1;
      if (n_event == P.max_events)
        wgt = 0.0;
      S.n_event[p] = n_event;
      store_geo(S, nl, p, g);
      S.wgt[p] = wgt;
      push = wgt != 0.0;
    }
    wave_push(S.q_xs, &S.qcount[QC_XS], push, p);
  }
  wave_count(&S.counters[CNT_SURF_EVENTS], n_events);
  wave_count(&S.counters[CNT_LOST], n_lost);
  // leakage: wavefront sum, one atomic per wavefront
  for (int o = 32; o > 0; o >>= 1)
    leak += __shfl_xor(leak, o);
  if (lane_id() == 0 && leak != 0.0)
    unsafeAtomicAdd(S.leak, leak);
}

//==============================================================================
// collision event (src/particle.cpp:331-415, src/physics.cpp:43-168, 494-1041)
//==============================================================================

// sample_cxs_target_velocity (src/physics.cpp:990-1041)
__device__ inline V3 sample_cxs_target_velocity(double awr, double E, V3 u, double kT, uint64_t* seed)
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

__global__ void __launch_bounds__(BLOCK) k_collision(DevProblem P, DevParticles S)
{
  const int n = S.qcount[QC_COLL];
  const int64_t cap = S.cap;
  unsigned long long n_events = 0, n_lost = 0, n_lookups = 0, n_probes = 0;
  for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const int i = base + threadIdx.x;
    bool push = false;
    int p = 0;
    if (i < n) {
      p = S.q_coll[i];
      n_events++;
      double E = S.E[p];
      double wgt = S.wgt[p];
      uint64_t seed = S.seed[p];
      const int material = S.material[p];
      const int nc = S.n_coord[p];
      V3 u = {S.ux[p], S.uy[p], S.uz[p]};
      // sample_nuclide (src/physics.cpp:494-518)
      double cutoff = prn(&seed) * S.xs_t[p];
      const int j0 = P.mat_off[material];
      const int j1 = P.mat_off[material + 1];
      double prob = 0.0;
      int i_nuclide = -1;
      for (int j = j0; j < j1; ++j) {
        const double dens = P.mat_dens[j] * 1.0;
        prob += dens * S.micro_t[static_cast<int64_t>(j - j0) * cap + p];
        if (prob >= cutoff) {
          i_nuclide = P.mat_nuc[j];
          break;
        }
      }
      bool absorbed = false;
      if (i_nuclide < 0) {
        wgt = 0.0;
        S.lost[p] = 1;
        n_lost++;
      } else {
        // micro cross sections of the sampled nuclide at this energy
        const double lx = detm::log(E / P.energy_min);
        int i_log = static_cast<int>(lx / P.log_spacing);
        i_log = i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
        int probes = 0;
        Micro m = nuclide_xs(P, i_nuclide, E, i_log, lx, S.use_learned != 0, probes);
        n_lookups++;
        n_probes += probes;
        // absorption (src/physics.cpp:638-672), analog
        if (m.absorption > 0.0) {
          if (m.absorption > prn(&seed) * m.total) {
            wgt = 0.0;
            absorbed = true;
          }
        }
        if (!absorbed) {
          // scatter (src/physics.cpp:674-752): the data hold elastic scattering only
          double cut = prn(&seed) * (m.total - m.absorption);
          (void)cut;
          const DevNuclide nd = P.nuc[i_nuclide];
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
      if (S.trace_hash) {
        uint64_t h = S.hash[p];
        hmix(h, 0x44);
        hmix(h, static_cast<uint64_t>(static_cast<int64_t>(i_nuclide)));
        hmix(h, absorbed ? 101 : 2);
        hmix(h, dbits(E));
        hmix(h, dbits(u.x));
        hmix(h, dbits(u.y));
        hmix(h, dbits(u.z));
        S.hash[p] = h;
      }
      if (E < P.energy_cutoff)
        wgt = 0.0;
      // Particle::event_collide epilogue (src/particle.cpp:388-406)
      S.surface[p] = SURFACE_NONE;
      S.material_last[p] = C_NONE;
      for (int j = 0; j < nc; ++j) {
        int64_t k = j * cap + p;
        S.ux[k] = u.x;
        S.uy[k] = u.y;
        S.uz[k] = u.z;
      }
      int n_event = S.n_event[p] + 1;
      if (n_event == P.max_events)
        wgt = 0.0;
      S.n_event[p] = n_event;
      S.E[p] = E;
      S.wgt[p] = wgt;
      S.seed[p] = seed;
      push = wgt != 0.0;
    }
    wave_push(S.q_xs, &S.qcount[QC_XS], push, p);
  }
  wave_count(&S.counters[CNT_COLL_EVENTS], n_events);
  wave_count(&S.counters[CNT_COLLISIONS], n_events);
  wave_count(&S.counters[CNT_LOST], n_lost);
  wave_count(&S.counters[CNT_LOOKUPS], n_lookups);
  wave_count(&S.counters[CNT_SEARCH_PROBES], n_probes);
}

//==============================================================================
// Tally accumulation (src/tallies/tally.cpp:821-878)
//==============================================================================

__global__ void __launch_bounds__(BLOCK) k_accumulate(double* value, double* results, int64_t n_bins, double norm)
{
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_bins;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double val = value[i] * norm;
    value[i] = 0.0;
    results[3 * i + 1] += val;
    results[3 * i + 2] += val * val;
  }
}

//==============================================================================
// Host launch wrappers
//==============================================================================

void launch_init(const DevProblem& P, const DevParticles& S, int64_t n, int64_t id_base, hipStream_t st)
{
  int blocks = static_cast<int>((n + BLOCK - 1) / BLOCK);
  if (blocks < 1)
    blocks = 1;
  hipLaunchKernelGGL(k_init, dim3(blocks), dim3(BLOCK), 0, st, P, S, n, id_base);
}

void launch_cycle(const DevProblem& P, const DevParticles& S, int blocks, hipStream_t st)
{
  hipLaunchKernelGGL(k_xs, dim3(blocks), dim3(BLOCK), 0, st, P, S);
  hipLaunchKernelGGL(k_advance, dim3(blocks), dim3(BLOCK), 0, st, P, S);
  hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(BLOCK), 0, st, P, S);
  hipLaunchKernelGGL(k_collision, dim3(blocks), dim3(BLOCK), 0, st, P, S);
}

void launch_kernel(int which, const DevProblem& P, const DevParticles& S, int blocks, hipStream_t st)
{
  switch (which) {
  case 0: hipLaunchKernelGGL(k_xs, dim3(blocks), dim3(BLOCK), 0, st, P, S); break;
  case 1: hipLaunchKernelGGL(k_advance, dim3(blocks), dim3(BLOCK), 0, st, P, S); break;
  case 2: hipLaunchKernelGGL(k_surface, dim3(blocks), dim3(BLOCK), 0, st, P, S); break;
  default: hipLaunchKernelGGL(k_collision, dim3(blocks), dim3(BLOCK), 0, st, P, S); break;
  }
}

void launch_accumulate(double* value, double* results, int64_t n_bins, double norm, hipStream_t st)
{
  int64_t b = (n_bins + BLOCK - 1) / BLOCK;
  int blocks = static_cast<int>(b > 4096 ? 4096 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(BLOCK), 0, st, value, results, n_bins, norm);
}

} // namespace omc
