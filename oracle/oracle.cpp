// ORACLE -- test infrastructure only.
//
// A CPU restatement of OpenMC's event-based / history-based continuous-energy
// fixed-source neutron transport loop, written function by function in the
// reference's operation order so that every history (RNG draws, energy-grid
// indices, positions, directions, energies) is reproduced bit for bit.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
// only as the checker.  The product (libopenmc.so, HIP kernels) never links it.
//
// The one shared header is include/openmc_amd/detmath.h (log/sin/cos), which
// stands in for glibc's libm; it is pinned separately against glibc and a
// 50-digit decimal evaluation (tests/test_detmath.py).  Building with
// -DORACLE_GLIBC_MATH swaps in glibc for measuring how often the two differ.
//
// Scope: configs with planes / z-cylinders / rectangular lattices, vacuum and
// reflective boundaries, nuclides with elastic (MT2) and capture (MT102) only,
// NEAREST temperature with one temperature, point/box sources with isotropic
// or monodirectional angles and discrete/Watt energies, and RegularMesh
// track-length flux tallies (optionally with an energy filter).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "openmc_amd/detmath.h"

namespace oracle {

//==============================================================================
// Constants -- include/openmc/constants.h:49-113, include/openmc/random_lcg.h
//==============================================================================

constexpr double TINY_BIT = 1e-8;
constexpr double FP_PRECISION = 1e-14;
constexpr double FP_REL_PRECISION = 1e-5;
constexpr double FP_COINCIDENT = 1e-12;
constexpr double PI = 3.141592653589793238462643383279502884L;
constexpr double INFTY = 1.7976931348623157e308; // numeric_limits<double>::max
constexpr double MASS_NEUTRON = 1.00866491595;
constexpr double MASS_NEUTRON_EV = 939.56542052e6;
constexpr double C_LIGHT = 2.99792458e10;
constexpr double N_AVOGADRO = 0.602214076;
constexpr double K_BOLTZMANN = 8.617333262e-5;
constexpr int C_NONE = -1;
constexpr int MATERIAL_VOID = -1;
constexpr int SURFACE_NONE = 0;
constexpr double CACHE_INVALID = -1.0;

constexpr int N_STREAMS = 4;
constexpr int STREAM_TRACKING = 0;
constexpr int STREAM_SOURCE = 1;
constexpr uint64_t PRN_MULT = 6364136223846793005ULL;
constexpr uint64_t PRN_ADD = 1442695040888963407ULL;

//==============================================================================
// libm stand-in
//==============================================================================

#ifdef ORACLE_GLIBC_MATH
inline double m_log(double x) { return std::log(x); }
inline double m_sin(double x) { return std::sin(x); }
inline double m_cos(double x) { return std::cos(x); }
#else
inline double m_log(double x) { return detm::log(x); }
inline double m_sin(double x) { return detm::sin(x); }
inline double m_cos(double x) { return detm::cos(x); }
#endif

//==============================================================================
// RNG -- src/random_lcg.cpp:32-120
//==============================================================================

struct Rng {
  uint64_t master_seed = 1;
  uint64_t stride = 152917ULL;
};

// src/random_lcg.cpp:32-44
inline double prn(uint64_t* seed)
{
  *seed = (PRN_MULT * (*seed) + PRN_ADD);
  uint64_t word =
    ((*seed >> ((*seed >> 59u) + 5u)) ^ *seed) * 12605985483714917081ull;
  uint64_t result = (word >> 43u) ^ word;
  return std::ldexp(static_cast<double>(result), -64);
}

// src/random_lcg.cpp:91-120 (Brown 1994 skip-ahead)
inline uint64_t future_seed(uint64_t n, uint64_t seed)
{
  uint64_t g = PRN_MULT;
  uint64_t c = PRN_ADD;
  uint64_t g_new = 1;
  uint64_t c_new = 0;
  while (n > 0) {
    if (n & 1) {
      g_new *= g;
      c_new = c_new * g + c;
    }
    c *= (g + 1);
    g *= g;
    n >>= 1;
  }
  return g_new * seed + c_new;
}

// src/random_lcg.cpp:60-76
inline uint64_t init_seed(const Rng& rng, int64_t id, int offset)
{
  return future_seed(static_cast<uint64_t>(id) * rng.stride, rng.master_seed + offset);
}

// src/random_dist.cpp:10-13
inline double uniform_distribution(double a, double b, uint64_t* seed)
{
  return a + (b - a) * prn(seed);
}

// src/random_dist.cpp:20-32
inline double maxwell_spectrum(double T, uint64_t* seed)
{
  double r1 = prn(seed);
  double r2 = prn(seed);
  double r3 = prn(seed);
  double c = m_cos(PI / 2. * r3);
  return -T * (m_log(r1) + m_log(r2) * c * c);
}

// src/random_dist.cpp:34-39
inline double watt_spectrum(double a, double b, uint64_t* seed)
{
  double w = maxwell_spectrum(a, seed);
  return w + 0.25 * a * a * b +
         uniform_distribution(-1., 1., seed) * std::sqrt(a * a * b * w);
}

//==============================================================================
// Position -- include/openmc/position.h, src/position.cpp
//==============================================================================

struct Pos {
  double x = 0., y = 0., z = 0.;
  double& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double dot(Pos o) const { return x * o.x + y * o.y + z * o.z; }
  double norm() const { return std::sqrt(x * x + y * y + z * z); }
};
inline Pos operator+(Pos a, Pos b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline Pos operator-(Pos a, Pos b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// a * Position multiplies each component by a (position.h operator*(double, Position))
inline Pos operator*(double a, Pos b) { return {b.x * a, b.y * a, b.z * a}; }
inline Pos operator*(Pos b, double a) { return {b.x * a, b.y * a, b.z * a}; }
inline Pos operator*(Pos a, Pos b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline Pos operator/(Pos a, double v) { return {a.x / v, a.y / v, a.z / v}; }
// position.h:207-213
inline Pos reflect(Pos u, Pos n)
{
  const double projection = n.dot(u);
  const double magnitude = n.dot(n);
  double s = 2.0 * projection / magnitude;
  n = n * s;
  return u - n;
}

// src/math_functions.cpp:763-792
inline Pos rotate_angle(Pos u, double mu, const double* phi, uint64_t* seed)
{
  double phi_;
  if (phi != nullptr) {
    phi_ = *phi;
  } else {
    phi_ = 2.0 * PI * prn(seed);
  }
  double sinphi = m_sin(phi_);
  double cosphi = m_cos(phi_);
  double a = std::sqrt(std::fmax(0., 1. - mu * mu));
  double b = std::sqrt(std::fmax(0., 1. - u.z * u.z));
  if (b > 1e-10) {
    return {mu * u.x + a * (u.x * u.z * cosphi - u.y * sinphi) / b,
      mu * u.y + a * (u.y * u.z * cosphi + u.x * sinphi) / b,
      mu * u.z - a * b * cosphi};
  } else {
    b = std::sqrt(1. - u.y * u.y);
    return {mu * u.x + a * (-u.x * u.y * sinphi + u.z * cosphi) / b,
      mu * u.y + a * b * sinphi,
      mu * u.z - a * (u.y * u.z * sinphi + u.x * cosphi) / b};
  }
}

// src/distribution_multi.cpp:106-112
inline Pos isotropic_direction(uint64_t* seed)
{
  double phi = uniform_distribution(0., 2.0 * PI, seed);
  double mu = uniform_distribution(-1., 1., seed);
  return {mu, std::sqrt(1.0 - mu * mu) * m_cos(phi),
    std::sqrt(1.0 - mu * mu) * m_sin(phi)};
}

//==============================================================================
// History hash (test-side protocol shared with the GPU trace)
//==============================================================================

inline uint64_t as_bits(double x)
{
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}
inline void mix(uint64_t& h, uint64_t v)
{
  h = (h ^ v) * 0x100000001b3ULL;
}

//==============================================================================
// Data model
//==============================================================================

enum SurfType { S_XPLANE = 1, S_YPLANE, S_ZPLANE, S_PLANE, S_ZCYL, S_XCYL, S_YCYL, S_SPHERE };
enum BcType { BC_TRANSMIT = 0, BC_VACUUM = 1, BC_REFLECT = 2 };

struct Surface {
  int id;
  int type;
  int bc;
  double c[4];

  double evaluate(Pos r) const
  {
    switch (type) {
    case S_XPLANE: return r.x - c[0];
    case S_YPLANE: return r.y - c[0];
    case S_ZPLANE: return r.z - c[0];
    case S_PLANE: return c[0] * r.x + c[1] * r.y + c[2] * r.z - c[3];
    case S_ZCYL: { // src/surface.cpp:394-400
      const double r1 = r.x - c[0];
      const double r2 = r.y - c[1];
      return r1 * r1 + r2 * r2 - c[2] * c[2];
    }
    case S_XCYL: {
      const double r1 = r.y - c[0];
      const double r2 = r.z - c[1];
      return r1 * r1 + r2 * r2 - c[2] * c[2];
    }
    case S_YCYL: {
      const double r1 = r.x - c[0];
      const double r2 = r.z - c[1];
      return r1 * r1 + r2 * r2 - c[2] * c[2];
    }
    case S_SPHERE: {
      const double x = r.x - c[0];
      const double y = r.y - c[1];
      const double z = r.z - c[2];
      return x * x + y * y + z * z - c[3] * c[3];
    }
    }
    return 0.0;
  }

  // src/surface.cpp:208-218
  static double plane_distance(double ri, double ui, bool coincident, double offset)
  {
    const double f = offset - ri;
    if (coincident || std::abs(f) < FP_COINCIDENT || ui == 0.0)
      return INFTY;
    const double d = f / ui;
    if (d < 0.0)
      return INFTY;
    return d;
  }

  // src/surface.cpp:403-442 (axis_aligned_cylinder_distance)
  static double cyl_distance(double u1, double r2p, double u2, double r3p, double u3,
    bool coincident, double off1, double off2, double radius)
  {
    const double a = 1.0 - u1 * u1;
    if (a == 0.0)
      return INFTY;
    const double r2 = r2p - off1;
    const double r3 = r3p - off2;
    const double k = r2 * u2 + r3 * u3;
    const double cc = r2 * r2 + r3 * r3 - radius * radius;
    const double quad = k * k - a * cc;
    if (quad < 0.0) {
      return INFTY;
    } else if (coincident || std::abs(cc) < FP_COINCIDENT) {
      if (k >= 0.0) {
        return INFTY;
      } else {
        return (-k + std::sqrt(quad)) / a;
      }
    } else if (cc < 0.0) {
      return (-k + std::sqrt(quad)) / a;
    } else {
      const double d = (-k - std::sqrt(quad)) / a;
      if (d < 0.0)
        return INFTY;
      return d;
    }
  }

  double distance(Pos r, Pos u, bool coincident) const
  {
    switch (type) {
    case S_XPLANE: return plane_distance(r.x, u.x, coincident, c[0]);
    case S_YPLANE: return plane_distance(r.y, u.y, coincident, c[0]);
    case S_ZPLANE: return plane_distance(r.z, u.z, coincident, c[0]);
    case S_PLANE: { // src/surface.cpp:354-366
      const double f = c[0] * r.x + c[1] * r.y + c[2] * r.z - c[3];
      const double projection = c[0] * u.x + c[1] * u.y + c[2] * u.z;
      if (coincident || std::abs(f) < FP_COINCIDENT || projection == 0.0) {
        return INFTY;
      } else {
        const double d = -f / projection;
        if (d < 0.0)
          return INFTY;
        return d;
      }
    }
    case S_ZCYL:
      return cyl_distance(u.z, r.x, u.x, r.y, u.y, coincident, c[0], c[1], c[2]);
    case S_XCYL:
      return cyl_distance(u.x, r.y, u.y, r.z, u.z, coincident, c[0], c[1], c[2]);
    case S_YCYL:
      return cyl_distance(u.y, r.x, u.x, r.z, u.z, coincident, c[0], c[1], c[2]);
    case S_SPHERE: { // src/surface.cpp SurfaceSphere::distance
      const double x = r.x - c[0];
      const double y = r.y - c[1];
      const double z = r.z - c[2];
      const double k = x * u.x + y * u.y + z * u.z;
      const double cc = x * x + y * y + z * z - c[3] * c[3];
      const double quad = k * k - cc;
      if (quad < 0.0) {
        return INFTY;
      } else if (coincident || std::abs(cc) < FP_COINCIDENT) {
        if (k >= 0.0) {
          return INFTY;
        } else {
          return -k + std::sqrt(quad);
        }
      } else if (cc < 0.0) {
        return -k + std::sqrt(quad);
      } else {
        const double d = -k - std::sqrt(quad);
        if (d < 0.0)
          return INFTY;
        return d;
      }
    }
    }
    return INFTY;
  }

  Pos normal(Pos r) const
  {
    switch (type) {
    case S_XPLANE: return {1., 0., 0.};
    case S_YPLANE: return {0., 1., 0.};
    case S_ZPLANE: return {0., 0., 1.};
    case S_PLANE: return {c[0], c[1], c[2]};
    case S_ZCYL: return {2.0 * (r.x - c[0]), 2.0 * (r.y - c[1]), 0.0};
    case S_XCYL: return {0.0, 2.0 * (r.y - c[0]), 2.0 * (r.z - c[1])};
    case S_YCYL: return {2.0 * (r.x - c[0]), 0.0, 2.0 * (r.z - c[1])};
    case S_SPHERE: return {2.0 * (r.x - c[0]), 2.0 * (r.y - c[1]), 2.0 * (r.z - c[2])};
    }
    return {0., 0., 0.};
  }

  // src/surface.cpp:116-130
  bool sense(Pos r, Pos u) const
  {
    const double f = evaluate(r);
    if (std::abs(f) < FP_COINCIDENT) {
      return u.dot(normal(r)) > 0.0;
    }
    return f > 0.0;
  }
};

enum FillType { FILL_MATERIAL = 0, FILL_UNIVERSE = 1, FILL_LATTICE = 2 };

struct Cell {
  int id;
  int universe;          // index
  int fill_type;
  int fill;              // universe or lattice index
  int material;          // material index or MATERIAL_VOID
  std::vector<int> region; // signed (surface index + 1)
};

struct Universe {
  int id;
  std::vector<int> cells;
};

struct Lattice {
  int id;
  int n[3];
  bool is_3d;
  Pos ll, pitch;
  int outer; // universe index or -1
  std::vector<int> universes; // flat, src/lattice.cpp:233-244 ordering

  // src/lattice.cpp:243-248
  bool valid(const int* i) const
  {
    return (i[0] >= 0) && (i[0] < n[0]) && (i[1] >= 0) && (i[1] < n[1]) &&
           (i[2] >= 0) && (i[2] < n[2]);
  }
  // src/lattice.cpp:344-348
  int flat(const int* i) const { return n[0] * n[1] * i[2] + n[0] * i[1] + i[0]; }

  // src/lattice.cpp:252-306
  double distance(Pos r, Pos u, int trans[3]) const
  {
    double x = r.x, y = r.y, z = r.z;
    double x0 = std::copysign(0.5 * pitch.x, u.x);
    double y0 = std::copysign(0.5 * pitch.y, u.y);
    double d = INFTY;
    trans[0] = trans[1] = trans[2] = 0;
    if ((std::abs(x - x0) > FP_PRECISION) && u.x != 0) {
      d = (x0 - x) / u.x;
      trans[0] = u.x > 0 ? 1 : -1;
      trans[1] = trans[2] = 0;
    }
    if ((std::abs(y - y0) > FP_PRECISION) && u.y != 0) {
      double this_d = (y0 - y) / u.y;
      if (this_d < d) {
        d = this_d;
        trans[0] = 0;
        trans[1] = u.y > 0 ? 1 : -1;
        trans[2] = 0;
      }
    }
    if (is_3d) {
      double z0 = std::copysign(0.5 * pitch.z, u.z);
      if ((std::abs(z - z0) > FP_PRECISION) && u.z != 0) {
        double this_d = (z0 - z) / u.z;
        if (this_d < d) {
          d = this_d;
          trans[0] = trans[1] = 0;
          trans[2] = u.z > 0 ? 1 : -1;
        }
      }
    }
    return d;
  }

  // src/lattice.cpp:310-342
  void get_indices(Pos r, Pos u, int* res) const
  {
    for (int k = 0; k < 3; ++k) {
      if (k == 2 && !is_3d) {
        res[2] = 0;
        break;
      }
      double i_ = (r[k] - ll[k]) / pitch[k];
      long i_close = detm::lround(i_);
      if (std::abs(i_ - static_cast<double>(i_close)) < FP_COINCIDENT) {
        res[k] = (u[k] > 0) ? static_cast<int>(i_close) : static_cast<int>(i_close - 1);
      } else {
        res[k] = static_cast<int>(std::floor(i_));
      }
    }
  }

  // src/lattice.cpp:352-361
  Pos local_position(Pos r, const int* i) const
  {
    r.x -= (ll.x + (i[0] + 0.5) * pitch.x);
    r.y -= (ll.y + (i[1] + 0.5) * pitch.y);
    if (is_3d)
      r.z -= (ll.z + (i[2] + 0.5) * pitch.z);
    return r;
  }
};

struct Nuclide {
  std::string name;
  double awr;
  double kT;
  std::vector<double> energy;
  std::vector<double> total;      // xs_(:, XS_TOTAL)
  std::vector<double> absorption; // xs_(:, XS_ABSORPTION)
  std::vector<double> elastic;    // reactions_[0]->xs_
  std::vector<int> grid_index;
};

struct Material {
  int id;
  std::vector<int> nuclide;
  std::vector<double> atom_density;
  double density;
};

struct Mesh {
  int n_dim;
  int shape[3];
  double ll[3], ur[3], width[3];
};

struct Problem {
  // settings
  int64_t n_particles = 0;
  int n_batches = 1;
  int n_inactive = 0;
  int n_log_bins = 8000;
  int64_t max_events = 1000000;
  int64_t max_in_flight = 100000;
  double free_gas_threshold = 400.0;
  double energy_cutoff = 0.0;
  Rng rng;

  std::vector<Nuclide> nuclides;
  std::vector<Material> materials;
  std::vector<Surface> surfaces;
  std::vector<Cell> cells;
  std::vector<Universe> universes;
  std::vector<Lattice> lattices;
  int root_universe = 0;
  int n_coord_levels = 1;

  // source
  int space_type = 0; // 0 point, 1 box
  Pos space_p0, space_p1;
  int angle_type = 0; // 0 isotropic, 1 monodirectional
  Pos angle_u;
  int energy_type = 0; // 0 discrete, 1 watt
  std::vector<double> energy_x, energy_p;
  std::vector<double> alias_prob; std::vector<int> alias_idx;
  double watt_a = 0, watt_b = 0;
  std::vector<int> domain_cell_ids; // cell-domain constraint

  // tally
  bool has_tally = false;
  bool has_mesh = false; // tally filters [mesh] / [mesh, energy]; else [] or [energy]
  Mesh mesh;
  std::vector<double> energy_bins; // empty -> no energy filter
  int64_t n_tally_bins = 0;

  // mesh weight windows for neutrons (src/weight_windows.cpp): the sets in
  // definition order
  bool ww_on = false;
  struct WWSet {
    Mesh mesh;
    std::vector<double> ebounds;
    std::vector<double> lower, upper; // [n_energy][n_mesh_bins]
    int64_t n_mesh = 0;
    double survival_ratio = 3.0;
    double max_lb_ratio = 1.0;
    int max_split = 10;
    double weight_cutoff = 1.0e-38;
  };
  std::vector<WWSet> ww_sets;
  bool ww_ckpt_collision = true; // src/settings.cpp:86-87
  bool ww_ckpt_surface = false;
  int64_t max_history_splits = 10000000; // src/settings.cpp:121

  double energy_min = 0, energy_max = 0, log_spacing = 0;
};

//==============================================================================
// Particle (AoS, as include/openmc/particle_data.h)
//==============================================================================

struct LocalCoord {
  Pos r, u;
  int cell = C_NONE, universe = C_NONE, lattice = C_NONE;
  int lattice_index[3] = {0, 0, 0};
  void reset()
  {
    cell = universe = lattice = C_NONE;
    lattice_index[0] = lattice_index[1] = lattice_index[2] = 0;
  }
};

struct MicroXS {
  double total, absorption, elastic;
  int index_grid;
  double interp_factor;
  double last_E = 0.0;
};

struct Boundary {
  double distance = INFINITY;
  int surface = SURFACE_NONE;
  int coord_level = 0;
  int lattice_translation[3] = {0, 0, 0};
};

struct Counters {
  int64_t lookups = 0, collisions = 0, events = 0, segments = 0, lost = 0;
  int64_t xs_events = 0, advance_events = 0, surface_events = 0, collision_events = 0;
  int64_t fallthrough = 0;  // scatters whose elastic cross section did not exceed the cutoff
  double tracklength = 0.0; // sum of wgt * distance over all advances
};

// A secondary bank site (SourceSite fields set by Particle::split,
// src/particle.cpp:95-112)
struct BankSite {
  Pos r, u;
  double E = 0, time = 0, wgt = 0;
  int surface = SURFACE_NONE; // signed surface index + 1
};

struct Particle {
  // weight-window state (include/openmc/particle_data.h:507-508, 563-564)
  double wgt_born = 1.0, wgt_ww_born = -1.0, ww_factor = 0.0;
  int n_split = 0;
  std::vector<BankSite> bank; // secondary bank, LIFO
  std::vector<LocalCoord> coord;
  std::vector<int> cell_last;
  int n_coord = 1;
  int n_coord_last = 1;
  double E = 0, E_last = 0, wgt = 1, wgt_last = 1, time = 0, mu = 0;
  Pos r_last;
  uint64_t seeds[N_STREAMS];
  int stream = STREAM_TRACKING;
  int surface = SURFACE_NONE;
  int material = C_NONE, material_last = C_NONE;
  double macro_total = 0, macro_absorption = 0;
  std::vector<MicroXS> micro;
  Boundary boundary;
  double collision_distance = 0;
  int64_t n_event = 0;
  int n_collision = 0;
  int64_t id = 0;
  double leakage = 0;
  uint64_t hash = 0;
  uint64_t dhash = 0; // the discrete part of the history only (no floating-point bits)
  bool lost = false;

  Pos& r() { return coord[0].r; }
  Pos& u() { return coord[0].u; }
  LocalCoord& lowest() { return coord[n_coord - 1]; }
  uint64_t* current_seed() { return &seeds[stream]; }
  bool alive() const { return wgt != 0.0; }
};

//==============================================================================
// Simulation state
//==============================================================================

struct Sim {
  Problem P;
  std::vector<double> tally_value; // VALUE column per bin
  std::vector<double> results;     // [bins][3] VALUE, SUM, SUM_SQ
  int n_realizations = 0;
  double leakage_batch = 0, total_weight = 0;
  double leak_sum = 0, leak_sum_sq = 0;
  std::vector<uint64_t> hashes; // last batch, by particle index
  std::vector<uint64_t> dhashes; // discrete-only hashes of the same histories
  Counters cnt;
  int current_batch = 0;
  bool trace_hash = true;
  int n_threads = 1;
};

// Forward declarations
bool find_cell_inner(const Problem& P, Particle& p, bool neighbor_search);
bool exhaustive_find_cell(const Problem& P, Particle& p);

//==============================================================================
// Initialization -- src/simulation.cpp:682-761, src/nuclide.cpp:502-533
//==============================================================================

void initialize_data(Problem& P)
{
  P.energy_max = INFTY;
  P.energy_min = 0.0;
  for (const auto& nuc : P.nuclides) {
    P.energy_min = std::max(P.energy_min, nuc.energy.front());
    P.energy_max = std::min(P.energy_max, nuc.energy.back());
  }
  int M = P.n_log_bins;
  // Nuclide::init_grid (src/nuclide.cpp:502-533)
  double spacing = m_log(P.energy_max / P.energy_min) / M;
  // xt::linspace(0, M*spacing, M+1): start + k*step, step = (stop-start)/(num-1)
  // (xtensor's generator; its exact element formula is not pinned by any
  // reference test: k*step with step = stop/M is assumed)
  double stop = M * spacing;
  double step = (stop - 0.0) / static_cast<double>(M);
  for (auto& nuc : P.nuclides) {
    nuc.grid_index.assign(M + 1, 0);
    int j = 0;
    for (int k = 0; k <= M; ++k) {
      double umesh_k = 0.0 + step * static_cast<double>(k);
      while (m_log(nuc.energy[j + 1] / P.energy_min) <= umesh_k) {
        if (j + 2 == static_cast<int>(nuc.energy.size()))
          break;
        ++j;
      }
      nuc.grid_index[k] = j;
    }
  }
  P.log_spacing = m_log(P.energy_max / P.energy_min) / P.n_log_bins;
}

//==============================================================================
// Cross sections -- src/nuclide.cpp:616-839, src/material.cpp:829-901
//==============================================================================

// Nuclide::calculate_xs (NEAREST temperature, one temperature, no S(a,b)/URR/WMP)
void nuclide_calculate_xs(const Problem& P, int i_nuc, int i_log_union, Particle& p,
  Counters& cnt)
{
  const Nuclide& nuc = P.nuclides[i_nuc];
  MicroXS& micro = p.micro[i_nuc];
  micro.elastic = CACHE_INVALID;
  const auto& E = nuc.energy;
  int i_grid;
  if (p.E < E.front()) {
    i_grid = 0;
  } else if (p.E > E.back()) {
    i_grid = static_cast<int>(E.size()) - 2;
  } else {
    int i_low = nuc.grid_index[i_log_union];
    int i_high = nuc.grid_index[i_log_union + 1] + 1;
    // lower_bound_index (include/openmc/search.h:13-20)
    int off;
    if (E[i_low] == p.E) {
      off = 0;
    } else {
      off = static_cast<int>(std::lower_bound(&E[i_low], &E[i_high], p.E) - &E[i_low]) - 1;
    }
    i_grid = i_low + off;
  }
  // (the i_grid + 2 < size guard only excludes a read past the end that the
  // reference would make for a duplicated last grid point)
  if (i_grid + 2 < static_cast<int>(E.size()) && E[i_grid] == E[i_grid + 1])
    ++i_grid;
  double f = (p.E - E[i_grid]) / (E[i_grid + 1] - E[i_grid]);
  micro.index_grid = i_grid;
  micro.interp_factor = f;
  micro.total = (1.0 - f) * nuc.total[i_grid] + f * nuc.total[i_grid + 1];
  micro.absorption = (1.0 - f) * nuc.absorption[i_grid] + f * nuc.absorption[i_grid + 1];
  micro.last_E = p.E;
  cnt.lookups++;
}

// Nuclide::calculate_elastic_xs (src/nuclide.cpp:578-590)
void calculate_elastic_xs(const Problem& P, int i_nuc, Particle& p)
{
  MicroXS& micro = p.micro[i_nuc];
  int i_grid = micro.index_grid;
  double f = micro.interp_factor;
  const auto& xs = P.nuclides[i_nuc].elastic;
  micro.elastic = (1.0 - f) * xs[i_grid] + f * xs[i_grid + 1];
}

// Material::calculate_xs / calculate_neutron_xs (src/material.cpp:814-901)
void material_calculate_xs(const Problem& P, Particle& p, Counters& cnt, bool hash)
{
  p.macro_total = 0.0;
  p.macro_absorption = 0.0;
  const Material& mat = P.materials[p.material];
  int i_grid = static_cast<int>(m_log(p.E / P.energy_min) / P.log_spacing);
  // clamp to the table: only E at or above energy_max (or below energy_min)
  // would index past it, which the reference leaves undefined
  i_grid = i_grid < 0 ? 0 : (i_grid >= P.n_log_bins ? P.n_log_bins - 1 : i_grid);
  if (hash) {
    mix(p.hash, 0x11);
    mix(p.dhash, 0x11);
    mix(p.hash, static_cast<uint64_t>(p.material));
    mix(p.dhash, static_cast<uint64_t>(p.material));
  }
  for (size_t i = 0; i < mat.nuclide.size(); ++i) {
    int i_nuclide = mat.nuclide[i];
    // Particle::update_neutron_xs (src/particle.cpp:859-878)
    MicroXS& micro = p.micro[i_nuclide];
    if (p.E != micro.last_E) {
      nuclide_calculate_xs(P, i_nuclide, i_grid, p, cnt);
    }
    if (hash)
      mix(p.hash, static_cast<uint64_t>(static_cast<uint32_t>(micro.index_grid)));
      mix(p.dhash, static_cast<uint64_t>(static_cast<uint32_t>(micro.index_grid)));
    double atom_density = mat.atom_density[i] * 1.0;
    p.macro_total += atom_density * micro.total;
    p.macro_absorption += atom_density * micro.absorption;
  }
}

//==============================================================================
// Geometry -- src/geometry.cpp, src/cell.cpp, src/universe.cpp
//==============================================================================

// Region::contains_simple (src/cell.cpp:951-970)
bool cell_contains(const Problem& P, const Cell& c, Pos r, Pos u, int on_surface)
{
  for (int token : c.region) {
    if (token == on_surface) {
    } else if (-token == on_surface) {
      return false;
    } else {
      bool sense = P.surfaces[std::abs(token) - 1].sense(r, u);
      if (sense != (token > 0))
        return false;
    }
  }
  return true;
}

// Region::distance (src/cell.cpp:910-936)
std::pair<double, int> cell_distance(const Problem& P, const Cell& c, Pos r, Pos u, int on_surface)
{
  double min_dist = INFTY;
  int i_surf = 2147483647;
  for (int token : c.region) {
    bool coincident = std::abs(token) == std::abs(on_surface);
    double d = P.surfaces[std::abs(token) - 1].distance(r, u, coincident);
    if (d < min_dist) {
      if (min_dist - d >= FP_PRECISION * min_dist) {
        min_dist = d;
        i_surf = -token;
      }
    }
  }
  return {min_dist, i_surf};
}

// Universe::find_cell (src/universe.cpp:40-60)
bool universe_find_cell(const Problem& P, Particle& p)
{
  int i_univ = p.lowest().universe;
  const Universe& univ = P.universes[i_univ];
  Pos r = p.lowest().r;
  Pos u = p.lowest().u;
  for (int i_cell : univ.cells) {
    if (P.cells[i_cell].universe != i_univ)
      continue;
    if (cell_contains(P, P.cells[i_cell], r, u, p.surface)) {
      p.lowest().cell = i_cell;
      return true;
    }
  }
  return false;
}

// find_cell_inner (src/geometry.cpp:102-252).  The neighbor-list search of the
// reference visits a subset of the same universe's cells; for a universe whose
// cells partition space (every config here) the first containing cell is the
// only containing cell, so a search over the whole universe finds the same cell.
bool find_cell_inner(const Problem& P, Particle& p, bool)
{
  bool found = false;
  for (;; ++p.n_coord) {
    found = universe_find_cell(P, p);
    if (!found)
      return found;
    int i_cell = p.lowest().cell;
    const Cell& c = P.cells[i_cell];
    if (c.fill_type == FILL_MATERIAL) {
      p.material_last = p.material;
      p.material = c.material;
      return true;
    } else if (c.fill_type == FILL_UNIVERSE) {
      LocalCoord& coord = p.coord[p.n_coord];
      coord.universe = c.fill;
      coord.r = p.lowest().r;
      coord.u = p.lowest().u;
    } else {
      const Lattice& lat = P.lattices[c.fill];
      LocalCoord& coord = p.coord[p.n_coord];
      coord.r = p.lowest().r;
      coord.u = p.lowest().u;
      lat.get_indices(coord.r, coord.u, coord.lattice_index);
      coord.r = lat.local_position(coord.r, coord.lattice_index);
      coord.lattice = c.fill;
      if (lat.valid(coord.lattice_index)) {
        coord.universe = lat.universes[lat.flat(coord.lattice_index)];
      } else {
        if (lat.outer != C_NONE) {
          coord.universe = lat.outer;
        } else {
          p.wgt = 0.0;
          p.lost = true;
          return false;
        }
      }
    }
  }
}

// neighbor_list_find_cell (src/geometry.cpp:256-282)
bool neighbor_list_find_cell(const Problem& P, Particle& p)
{
  for (int i = p.n_coord; i < P.n_coord_levels; i++)
    p.coord[i].reset();
  return find_cell_inner(P, p, true);
}

// exhaustive_find_cell (src/geometry.cpp:284-297)
bool exhaustive_find_cell(const Problem& P, Particle& p)
{
  int i_universe = p.lowest().universe;
  if (i_universe == C_NONE) {
    p.coord[0].universe = P.root_universe;
    p.n_coord = 1;
  }
  for (int i = p.n_coord; i < P.n_coord_levels; i++)
    p.coord[i].reset();
  return find_cell_inner(P, p, false);
}

// distance_to_boundary (src/geometry.cpp:361-453)
Boundary distance_to_boundary(const Problem& P, Particle& p)
{
  Boundary info;
  double d_lat = INFINITY;
  double d_surf = INFINITY;
  int level_surf_cross;
  int level_lat_trans[3] = {0, 0, 0};
  for (int i = 0; i < p.n_coord; i++) {
    const LocalCoord& coord = p.coord[i];
    Pos r = coord.r;
    Pos u = coord.u;
    const Cell& c = P.cells[coord.cell];
    auto sd = cell_distance(P, c, r, u, p.surface);
    d_surf = sd.first;
    level_surf_cross = sd.second;
    if (coord.lattice != C_NONE) {
      const Lattice& lat = P.lattices[coord.lattice];
      d_lat = lat.distance(r, u, level_lat_trans);
      if (d_lat < 0) {
        p.wgt = 0.0;
        p.lost = true;
      }
    }
    double& d = info.distance;
    if (d_surf < d_lat - FP_COINCIDENT) {
      if (d == INFINITY || (d - d_surf) / d >= FP_REL_PRECISION) {
        d = d_surf;
        // every cell here is simple (intersection-only)
        info.surface = level_surf_cross;
        info.lattice_translation[0] = 0;
        info.lattice_translation[1] = 0;
        info.lattice_translation[2] = 0;
        info.coord_level = i + 1;
      }
    } else {
      if (d == INFINITY || (d - d_lat) / d >= FP_REL_PRECISION) {
        d = d_lat;
        info.surface = SURFACE_NONE;
        info.lattice_translation[0] = level_lat_trans[0];
        info.lattice_translation[1] = level_lat_trans[1];
        info.lattice_translation[2] = level_lat_trans[2];
        info.coord_level = i + 1;
      }
    }
  }
  return info;
}

// cross_lattice (src/geometry.cpp:301-357)
void cross_lattice(const Problem& P, Particle& p, const Boundary& b)
{
  LocalCoord& coord = p.lowest();
  const Lattice& lat = P.lattices[coord.lattice];
  coord.lattice_index[0] += b.lattice_translation[0];
  coord.lattice_index[1] += b.lattice_translation[1];
  coord.lattice_index[2] += b.lattice_translation[2];
  const LocalCoord& upper = p.coord[p.n_coord - 2];
  Pos r = upper.r; // no translation / rotation in these models
  coord.r = lat.local_position(r, coord.lattice_index);
  if (!lat.valid(coord.lattice_index)) {
    p.n_coord = 1;
    bool found = exhaustive_find_cell(P, p);
    if (!found) {
      p.wgt = 0.0;
      p.lost = true;
    }
  } else {
    p.lowest().universe = lat.universes[lat.flat(coord.lattice_index)];
    bool found = exhaustive_find_cell(P, p);
    if (!found) {
      p.n_coord = 1;
      found = exhaustive_find_cell(P, p);
      if (!found) {
        p.wgt = 0.0;
        p.lost = true;
      }
    }
  }
}

//==============================================================================
// Tally -- src/mesh.cpp:1000-1146, src/tallies/tally_scoring.cpp:2416-2541
//==============================================================================

struct MeshDistance {
  int next_index = -1;
  bool max_surface = true;
  double distance = INFTY;
};

// RegularMesh::distance_to_grid_boundary (src/mesh.cpp:1350-1369)
MeshDistance mesh_distance(const Mesh& m, const int* ijk, int i, Pos r0, Pos u, double l)
{
  (void)l;
  MeshDistance d;
  d.next_index = ijk[i];
  if (std::abs(u[i]) < FP_PRECISION)
    return d;
  d.max_surface = (u[i] > 0);
  if (d.max_surface && (ijk[i] <= m.shape[i])) {
    d.next_index++;
    d.distance = ((m.ll[i] + ijk[i] * m.width[i]) - r0[i]) / u[i];
  } else if (!d.max_surface && (ijk[i] >= 1)) {
    d.next_index--;
    d.distance = ((m.ll[i] + (ijk[i] - 1) * m.width[i]) - r0[i]) / u[i];
  }
  return d;
}

// StructuredMesh::get_indices + RegularMesh::get_index_in_direction (src/mesh.cpp:874-887, 1328-1331)
void mesh_indices(const Mesh& m, Pos r, int* ijk, bool& in_mesh)
{
  in_mesh = true;
  for (int i = 0; i < m.n_dim; ++i) {
    ijk[i] = static_cast<int>(std::ceil((r[i] - m.ll[i]) / m.width[i]));
    if (ijk[i] < 1 || ijk[i] > m.shape[i])
      in_mesh = false;
  }
}

int mesh_bin(const Mesh& m, const int* ijk)
{
  switch (m.n_dim) {
  case 1: return ijk[0] - 1;
  case 2: return (ijk[1] - 1) * m.shape[0] + ijk[0] - 1;
  default: return ((ijk[2] - 1) * m.shape[1] + (ijk[1] - 1)) * m.shape[0] + ijk[0] - 1;
  }
}

// StructuredMesh::raytrace_mesh with the bins_crossed aggregator (src/mesh.cpp:1000-1146)
template<class F>
void raytrace_mesh(const Mesh& m, Pos r0, Pos r1, Pos u, F track)
{
  double total_distance = (r1 - r0).norm();
  if (total_distance == 0.0)
    return;
  Pos global_r = r0;
  Pos local_r = r0;
  const int n = m.n_dim;
  bool in_mesh;
  double traveled_distance = 0.0;
  int ijk[3] = {0, 0, 0};
  mesh_indices(m, global_r + TINY_BIT * u, ijk, in_mesh);
  if (total_distance < 2 * TINY_BIT) {
    if (in_mesh)
      track(ijk, 1.0);
    return;
  }
  MeshDistance distances[3];
  for (int k = 0; k < n; ++k)
    distances[k] = mesh_distance(m, ijk, k, local_r, u, 0.0);
  while (true) {
    if (in_mesh) {
      // std::min_element: first minimum over the 3 entries (entries past
      // n_dim keep the default INFTY distance)
      int k = 0;
      for (int kk = 1; kk < 3; ++kk)
        if (distances[kk].distance < distances[k].distance)
          k = kk;
      track(ijk, (std::min(distances[k].distance, total_distance) - traveled_distance) /
                   total_distance);
      traveled_distance = distances[k].distance;
      if (traveled_distance >= total_distance)
        return;
      ijk[k] = distances[k].next_index;
      distances[k] = mesh_distance(m, ijk, k, local_r, u, traveled_distance);
      in_mesh = ((ijk[k] >= 1) && (ijk[k] <= m.shape[k]));
    } else {
      int k_max = -1;
      for (int k = 0; k < n; ++k) {
        if ((ijk[k] < 1 || ijk[k] > m.shape[k]) &&
            (distances[k].distance > traveled_distance)) {
          traveled_distance = distances[k].distance;
          k_max = k;
        }
      }
      if (k_max == -1)
        traveled_distance += TINY_BIT;
      if (traveled_distance >= total_distance)
        return;
      mesh_indices(m, global_r + (traveled_distance + TINY_BIT) * u, ijk, in_mesh);
      for (int k = 0; k < n; ++k)
        distances[k] = mesh_distance(m, ijk, k, local_r, u, traveled_distance);
    }
  }
}

// score_tracklength_tally (src/tallies/tally_scoring.cpp:2534-2541, 2416-2490)
void score_tracklength_tally(Sim& S, Particle& p, double distance, Counters& cnt)
{
  const Problem& P = S.P;
  if (!P.has_tally)
    return;
  double flux = p.wgt * distance;
  // EnergyFilter::get_all_bins (src/tallies/filter_energy.cpp:59-81)
  int e_bin = 0;
  int n_e = 1;
  if (!P.energy_bins.empty()) {
    double E = p.E_last;
    if (!(E >= P.energy_bins.front() && E <= P.energy_bins.back()))
      return;
    if (P.energy_bins[0] == E) {
      e_bin = 0;
    } else {
      e_bin = static_cast<int>(std::lower_bound(P.energy_bins.begin(), P.energy_bins.end(), E) -
                               P.energy_bins.begin()) - 1;
    }
    n_e = static_cast<int>(P.energy_bins.size()) - 1;
  }
  if (!P.has_mesh) {
    // filters [] or [energy]: one bin per energy group, filter weight 1
    double score = flux * 1.0;
#pragma omp atomic
    S.tally_value[e_bin] += score;
    cnt.segments++;
    return;
  }
  // filters = [mesh, energy]; strides: energy 1, mesh n_e (src/tallies/tally.cpp:503-516)
  raytrace_mesh(P.mesh, p.r_last, p.r(), p.u(), [&](const int* ijk, double w) {
    int64_t bin = static_cast<int64_t>(mesh_bin(P.mesh, ijk)) * n_e + e_bin;
    double weight = 1.0;
    weight *= w;
    weight *= 1.0;
    double score = flux * weight;
#pragma omp atomic
    S.tally_value[bin] += score;
    cnt.segments++;
  });
}

//==============================================================================
// Source -- src/source.cpp:177-417, 627-660
//==============================================================================

//==============================================================================
// Weight windows -- src/weight_windows.cpp:56-146, 375-411;
// src/physics_common.cpp:12-19; src/particle.cpp:95-112
//==============================================================================

struct WeightWindow {
  double lower = -1.0; // -1: invalid
  double upper = 1.0;
  double max_lb_ratio = 1.0;
  double survival = 0.5;
  double cutoff = 1.0e-38;
  int max_split = 10;
  bool valid() const { return lower >= 0.0; }
  void scale(double f)
  {
    lower *= f;
    upper *= f;
    survival *= f;
  }
};

// lower_bound_index (include/openmc/search.h:13-20)
inline int lower_bound_index(const std::vector<double>& v, double x)
{
  if (v[0] == x)
    return 0;
  return static_cast<int>(std::lower_bound(v.begin(), v.end(), x) - v.begin()) - 1;
}

// WeightWindows::get_weight_window (src/weight_windows.cpp:375-411) of one set
WeightWindow set_weight_window(const Problem::WWSet& W, Particle& p)
{
  WeightWindow ww;
  int ijk[3] = {0, 0, 0};
  bool in_mesh;
  mesh_indices(W.mesh, p.r(), ijk, in_mesh);
  if (!in_mesh)
    return ww;
  const double E = p.E;
  if (E < W.ebounds.front() || E > W.ebounds.back())
    return ww;
  const int e_bin = lower_bound_index(W.ebounds, E);
  const int64_t k = static_cast<int64_t>(e_bin) * W.n_mesh + mesh_bin(W.mesh, ijk);
  ww.lower = W.lower[k];
  ww.upper = W.upper[k];
  ww.survival = ww.lower * W.survival_ratio;
  ww.max_lb_ratio = W.max_lb_ratio;
  ww.max_split = W.max_split;
  ww.cutoff = W.weight_cutoff;
  return ww;
}

// apply_weight_windows' search over the sets (src/weight_windows.cpp:70-75)
WeightWindow get_weight_window(const Problem& P, Particle& p)
{
  WeightWindow ww;
  for (const Problem::WWSet& W : P.ww_sets) {
    ww = set_weight_window(W, p);
    if (ww.valid())
      break;
  }
  return ww;
}

void russian_roulette(Particle& p, double weight_survive)
{
  if (weight_survive * prn(p.current_seed()) < p.wgt)
    p.wgt = weight_survive;
  else
    p.wgt = 0.0;
}

void split(Particle& p, double wgt)
{
  BankSite s;
  s.r = p.r();
  s.u = p.u();
  s.E = p.E;
  s.time = p.time;
  s.wgt = wgt;
  s.surface = p.surface;
  p.bank.push_back(s);
}

void apply_weight_windows(const Problem& P, Particle& p)
{
  if (!P.ww_on)
    return;
  if (p.E <= 0 || !p.alive())
    return;
  WeightWindow ww = get_weight_window(P, p);
  if (p.wgt_ww_born == -1.0)
    p.wgt_ww_born = ww.valid() ? (ww.lower + ww.upper) / 2 : 1.0;
  if (!ww.valid())
    return;
  ww.scale(p.wgt_born / p.wgt_ww_born);
  const double weight = p.wgt;
  if (p.wgt < ww.cutoff) {
    p.wgt = 0.0;
    return;
  }
  if (p.ww_factor == 0.0 && ww.max_lb_ratio > 1.0 && p.wgt > ww.lower * ww.max_lb_ratio)
    p.ww_factor = p.wgt / (ww.lower * ww.max_lb_ratio);
  if (p.ww_factor > 1.0)
    ww.scale(p.ww_factor);
  if (weight > ww.upper) {
    if (p.n_split >= P.max_history_splits)
      return;
    double n_split = std::ceil(weight / ww.upper);
    double max_split = ww.max_split;
    n_split = std::min(n_split, max_split);
    p.n_split = static_cast<int>(p.n_split + n_split);
    const int i_split = static_cast<int>(std::round(n_split));
    for (int l = 0; l < i_split - 1; l++)
      split(p, weight / n_split);
    p.wgt = weight / n_split;
  } else if (weight <= ww.lower) {
    const double weight_survive = std::min(weight * ww.max_split, ww.survival);
    russian_roulette(p, weight_survive);
  }
}

struct SourceSite {
  Pos r, u;
  double E = 0, time = 0, wgt = 1.0;
};

// Source::satisfies_spatial_constraints (src/source.cpp:226-271), cell domains
bool satisfies_spatial_constraints(const Problem& P, Pos r)
{
  Particle g;
  g.coord.resize(P.n_coord_levels);
  g.cell_last.resize(P.n_coord_levels);
  g.coord[0].r = r;
  g.coord[0].u = {0.0, 0.0, 1.0};
  bool found = exhaustive_find_cell(P, g);
  if (!found)
    return false;
  bool accepted = true;
  if (!P.domain_cell_ids.empty()) {
    for (int i = 0; i < g.n_coord; i++) {
      int id = P.cells[g.coord[i].cell].id;
      accepted = std::find(P.domain_cell_ids.begin(), P.domain_cell_ids.end(), id) !=
                 P.domain_cell_ids.end();
      if (accepted)
        break;
    }
  }
  return accepted;
}

// IndependentSource::sample (src/source.cpp:355-417)
SourceSite sample_source(const Problem& P, uint64_t* seed)
{
  SourceSite site;
  bool accepted = false;
  while (!accepted) {
    if (P.space_type == 0) {
      site.r = P.space_p0; // SpatialPoint::sample
    } else {
      // SpatialBox::sample (src/distribution_spatial.cpp:363-367)
      double x1 = prn(seed);
      double x2 = prn(seed);
      double x3 = prn(seed);
      Pos xi {x1, x2, x3};
      site.r = P.space_p0 + xi * (P.space_p1 - P.space_p0);
    }
    accepted = satisfies_spatial_constraints(P, site.r);
  }
  if (P.angle_type == 0) {
    site.u = isotropic_direction(seed);
  } else {
    site.u = P.angle_u;
  }
  while (true) {
    if (P.energy_type == 0) {
      // Discrete::sample -> DiscreteIndex::sample (src/distribution.cpp:87-101)
      size_t n = P.energy_x.size();
      size_t k = 0;
      if (n > 1) {
        size_t uu = static_cast<size_t>(prn(seed) * n);
        if (prn(seed) < P.alias_prob[uu])
          k = uu;
        else
          k = P.alias_idx[uu];
      }
      site.E = P.energy_x[k];
    } else {
      site.E = watt_spectrum(P.watt_a, P.watt_b, seed);
    }
    // Source::satisfies_energy_constraints defaults to (0, INFTY)
    if (site.E < P.energy_max && (site.E > 0.0 && site.E < INFTY))
      break;
  }
  site.time = 0.0;
  return site;
}

//==============================================================================
// Events -- src/particle.cpp, src/physics.cpp
//==============================================================================

// initialize_history (src/simulation.cpp:573-649), fixed source, one rank
void initialize_history(Sim& S, Particle& p, int64_t index_source)
{
  const Problem& P = S.P;
  int64_t id = (static_cast<int64_t>(S.current_batch) - 1) * P.n_particles + index_source;
  uint64_t seed = init_seed(P.rng, id, STREAM_SOURCE);
  SourceSite site = sample_source(P, &seed);
  // Particle::from_source (src/particle.cpp:114-158)
  p.coord.assign(P.n_coord_levels, LocalCoord());
  p.cell_last.assign(P.n_coord_levels, C_NONE);
  p.n_coord = 1;
  p.surface = SURFACE_NONE;
  p.material = C_NONE;
  p.material_last = C_NONE;
  p.n_collision = 0;
  p.wgt = site.wgt;
  p.wgt_last = site.wgt;
  p.coord[0].r = site.r;
  p.coord[0].u = site.u;
  p.r_last = site.r;
  p.E = site.E;
  p.E_last = site.E;
  p.time = site.time;
  p.id = index_source;
  p.n_event = 0;
  p.leakage = 0;
  p.lost = false;
  p.n_split = 0;
  p.ww_factor = 0.0;
  p.wgt_born = p.wgt;
  p.bank.clear();
  for (int i = 0; i < N_STREAMS; i++)
    p.seeds[i] = future_seed(static_cast<uint64_t>(id) * P.rng.stride, P.rng.master_seed + i);
  p.stream = STREAM_TRACKING;
  // the particle's initial weight window (src/simulation.cpp:628-630)
  p.wgt_ww_born = -1.0;
  apply_weight_windows(P, p);
#pragma omp atomic
  S.total_weight += p.wgt;
  // invalidate_neutron_xs
  p.micro.assign(P.nuclides.size(), MicroXS {0, 0, 0, 0, 0, 0.0});
  p.hash = 0xcbf29ce484222325ULL ^ static_cast<uint64_t>(id);
  p.dhash = p.hash;
  if (S.trace_hash && P.ww_on) {
    mix(p.hash, as_bits(p.wgt));
    mix(p.hash, static_cast<uint64_t>(p.bank.size()));
    mix(p.dhash, static_cast<uint64_t>(p.bank.size()));
  }
}

// Particle::event_calculate_xs (src/particle.cpp:160-230)
void event_calculate_xs(Sim& S, Particle& p, Counters& cnt)
{
  const Problem& P = S.P;
  cnt.xs_events++;
  p.stream = STREAM_TRACKING;
  p.wgt_last = p.wgt;
  p.E_last = p.E;
  p.r_last = p.r();
  if (p.lowest().cell == C_NONE) {
    if (!exhaustive_find_cell(P, p)) {
      p.wgt = 0.0;
      p.lost = true;
      return;
    }
    for (int j = 0; j < p.n_coord; ++j)
      p.cell_last[j] = p.coord[j].cell;
    p.n_coord_last = p.n_coord;
  }
  if (p.material != MATERIAL_VOID) {
    if (p.material != p.material_last) {
      material_calculate_xs(P, p, cnt, S.trace_hash);
    }
  } else {
    p.macro_total = 0.0;
    p.macro_absorption = 0.0;
  }
}

// Particle::event_advance (src/particle.cpp:232-286)
void event_advance(Sim& S, Particle& p, Counters& cnt)
{
  const Problem& P = S.P;
  cnt.advance_events++;
  p.boundary = distance_to_boundary(P, p);
  if (p.macro_total == 0.0) {
    p.collision_distance = INFINITY;
  } else {
    p.collision_distance = -m_log(prn(p.current_seed())) / p.macro_total;
  }
  double speed = C_LIGHT * std::sqrt(p.E * (p.E + 2 * MASS_NEUTRON_EV)) / (p.E + MASS_NEUTRON_EV);
  double distance_cutoff = INFTY;
  double distance = std::min({p.boundary.distance, p.collision_distance, distance_cutoff});
  for (int j = 0; j < p.n_coord; ++j) {
    Pos du = distance * p.coord[j].u;
    p.coord[j].r = p.coord[j].r + du;
  }
  double dt = distance / speed;
  p.time += dt;
  score_tracklength_tally(S, p, distance, cnt);
  cnt.tracklength += p.wgt * distance;
  if (distance == distance_cutoff)
    p.wgt = 0.0;
  if (S.trace_hash) {
    mix(p.hash, 0x22);
    mix(p.dhash, 0x22);
    mix(p.hash, as_bits(distance));
    mix(p.hash, p.collision_distance > p.boundary.distance ? 2 : 1);
    mix(p.dhash, p.collision_distance > p.boundary.distance ? 2 : 1);
  }
}

// Particle::cross_vacuum_bc / cross_reflective_bc (src/particle.cpp:631-710)
void cross_surface(const Problem& P, Particle& p, const Surface& surf)
{
  if (surf.bc == BC_VACUUM) {
    p.leakage += p.wgt;
    p.wgt = 0.0;
    return;
  }
  if (surf.bc == BC_REFLECT) {
    // ReflectiveBC::handle_particle (src/boundary_condition.cpp:37-46)
    Pos u = reflect(p.u(), surf.normal(p.r()));
    u = u / u.norm();
    if (p.n_coord != 1) {
      p.wgt = 0.0;
      p.lost = true;
      return;
    }
    p.coord[0].u = u;
    p.coord[0].cell = p.cell_last[0];
    p.surface = -p.surface;
    p.n_coord = 1;
    if (!neighbor_list_find_cell(P, p)) {
      p.wgt = 0.0;
      p.lost = true;
    }
    return;
  }
  // Particle::cross_surface (src/particle.cpp:549-629)
  if (neighbor_list_find_cell(P, p))
    return;
  p.n_coord = 1;
  bool found = exhaustive_find_cell(P, p);
  if (!found) {
    p.surface = SURFACE_NONE;
    p.n_coord = 1;
    p.coord[0].r = p.coord[0].r + TINY_BIT * p.coord[0].u;
    if (!exhaustive_find_cell(P, p)) {
      p.wgt = 0.0;
      p.lost = true;
    }
  }
}

// Particle::event_revive_from_secondary (src/particle.cpp:417-473): a dead
// particle continues as the last secondary it banked (Particle::from_source,
// src/particle.cpp:114-158); the RNG streams carry on.
void event_revive(const Sim& S, Particle& p)
{
  const Problem& P = S.P;
  ++p.n_event;
  if (p.n_event == P.max_events)
    p.wgt = 0.0;
  if (p.alive() || p.bank.empty())
    return;
  const BankSite s = p.bank.back();
  p.bank.pop_back();
  for (auto& c : p.coord)
    c.reset();
  p.n_coord = 1;
  p.surface = s.surface;
  p.material = C_NONE;
  p.n_collision = 0;
  p.wgt = s.wgt;
  p.wgt_last = s.wgt;
  p.coord[0].r = s.r;
  p.coord[0].u = s.u;
  p.r_last = s.r;
  p.E = s.E;
  p.E_last = s.E;
  p.time = s.time;
  p.n_event = 0;
  if (S.trace_hash) {
    mix(p.hash, 0x66);
    mix(p.dhash, 0x66);
    mix(p.hash, as_bits(p.wgt));
    mix(p.hash, as_bits(p.E));
  }
}

// Particle::event_cross_surface (src/particle.cpp:288-329)
void event_cross_surface(Sim& S, Particle& p, Counters& cnt)
{
  const Problem& P = S.P;
  cnt.surface_events++;
  for (int j = 0; j < p.n_coord; ++j)
    p.cell_last[j] = p.coord[j].cell;
  p.n_coord_last = p.n_coord;
  p.surface = p.boundary.surface;
  p.n_coord = p.boundary.coord_level;
  const int* t = p.boundary.lattice_translation;
  if (S.trace_hash) {
    mix(p.hash, 0x33);
    mix(p.dhash, 0x33);
    mix(p.hash, static_cast<uint64_t>(static_cast<int64_t>(p.surface)));
    mix(p.dhash, static_cast<uint64_t>(static_cast<int64_t>(p.surface)));
    mix(p.hash, static_cast<uint64_t>((t[0] + 1) | ((t[1] + 1) << 2) | ((t[2] + 1) << 4)));
    mix(p.dhash, static_cast<uint64_t>((t[0] + 1) | ((t[1] + 1) << 2) | ((t[2] + 1) << 4)));
  }
  if (t[0] != 0 || t[1] != 0 || t[2] != 0) {
    cross_lattice(P, p, p.boundary);
  } else {
    const Surface& surf = P.surfaces[std::abs(p.surface) - 1];
    cross_surface(P, p, surf);
    if (P.ww_ckpt_surface)
      apply_weight_windows(P, p);
  }
  if (S.trace_hash) {
    mix(p.hash, as_bits(p.wgt));
    mix(p.hash, static_cast<uint64_t>(static_cast<int64_t>(p.alive() ? p.lowest().cell : -1)));
    mix(p.dhash, static_cast<uint64_t>(static_cast<int64_t>(p.alive() ? p.lowest().cell : -1)));
    if (P.ww_on)
      mix(p.hash, static_cast<uint64_t>(p.bank.size()));
      mix(p.dhash, static_cast<uint64_t>(p.bank.size()));
  }
}

// sample_cxs_target_velocity (src/physics.cpp:990-1041)
Pos sample_cxs_target_velocity(double awr, double E, Pos u, double kT, uint64_t* seed)
{
  double beta_vn = std::sqrt(awr * E / kT);
  double alpha = 1.0 / (1.0 + std::sqrt(PI) * beta_vn / 2.0);
  double beta_vt_sq;
  double mu;
  while (true) {
    double r1 = prn(seed);
    double r2 = prn(seed);
    if (prn(seed) < alpha) {
      beta_vt_sq = -m_log(r1 * r2);
    } else {
      double c = m_cos(PI / 2.0 * prn(seed));
      beta_vt_sq = -m_log(r1) - m_log(r2) * c * c;
    }
    double beta_vt = std::sqrt(beta_vt_sq);
    mu = uniform_distribution(-1., 1., seed);
    double accept_prob =
      std::sqrt(beta_vn * beta_vn + beta_vt_sq - 2 * beta_vn * beta_vt * mu) /
      (beta_vn + beta_vt);
    if (prn(seed) < accept_prob)
      break;
  }
  double vt = std::sqrt(beta_vt_sq * kT / awr);
  return vt * rotate_angle(u, mu, nullptr, seed);
}

// sample_target_velocity (src/physics.cpp:836-988), non-resonant nuclides
Pos sample_target_velocity(const Problem& P, const Nuclide& nuc, double E, Pos u, double kT,
  uint64_t* seed)
{
  if (E >= P.free_gas_threshold * kT && nuc.awr > 1.0)
    return {};
  return sample_cxs_target_velocity(nuc.awr, E, u, kT, seed);
}

// elastic_scatter (src/physics.cpp:754-818), isotropic in CM (no angle data)
void elastic_scatter(const Problem& P, int i_nuc, double kT, Particle& p)
{
  const Nuclide& nuc = P.nuclides[i_nuc];
  double vel = std::sqrt(p.E);
  double awr = nuc.awr;
  Pos v_n = vel * p.u();
  Pos v_t = sample_target_velocity(P, nuc, p.E, p.u(), kT, p.current_seed());
  Pos v_cm = (v_n + awr * v_t) / (awr + 1.0);
  v_n = v_n - v_cm;
  vel = v_n.norm();
  double mu_cm = uniform_distribution(-1., 1., p.current_seed());
  Pos u_cm = v_n / vel;
  v_n = vel * rotate_angle(u_cm, mu_cm, nullptr, p.current_seed());
  v_n = v_n + v_cm;
  p.E = v_n.dot(v_n);
  vel = std::sqrt(p.E);
  p.mu = p.u().dot(v_n) / vel;
  p.coord[0].u = v_n / vel;
  if (std::abs(p.mu) > 1.0)
    p.mu = std::copysign(1.0, p.mu);
}

// Particle::event_collide + collision + sample_neutron_reaction
// (src/particle.cpp:331-415, src/physics.cpp:43-168, 494-518, 638-752)
void event_collide(Sim& S, Particle& p, Counters& cnt)
{
  const Problem& P = S.P;
  cnt.collision_events++;
  p.surface = SURFACE_NONE;
  ++p.n_collision;
  cnt.collisions++;
  // sample_nuclide
  double cutoff = prn(p.current_seed()) * p.macro_total;
  const Material& mat = P.materials[p.material];
  int n = static_cast<int>(mat.nuclide.size());
  double prob = 0.0;
  int i_nuclide = -1;
  for (int i = 0; i < n; ++i) {
    int in = mat.nuclide[i];
    double atom_density = mat.atom_density[i] * 1.0;
    prob += atom_density * p.micro[in].total;
    if (prob >= cutoff) {
      i_nuclide = in;
      break;
    }
  }
  if (i_nuclide < 0)
    throw std::runtime_error("Did not sample any nuclide during collision.");
  MicroXS& micro = p.micro[i_nuclide];
  bool absorbed = false;
  if (micro.absorption > 0.0) {
    // absorption (src/physics.cpp:638-672), analog
    if (micro.absorption > prn(p.current_seed()) * micro.total) {
      p.wgt = 0.0;
      absorbed = true;
    }
  }
  if (!absorbed) {
    // scatter (src/physics.cpp:674-752)
    double cut = prn(p.current_seed()) * (micro.total - micro.absorption);
    if (micro.elastic == CACHE_INVALID)
      calculate_elastic_xs(P, i_nuclide, p);
    double pr = micro.elastic - 0.0;
    // The data hold no inelastic reactions.  sigma_el exceeds the cutoff
    // except for a cutoff within rounding of sigma_t - sigma_a, where the
    // reference would take its inelastic loop with no inelastic reaction:
    // counted, and scattered elastically.
    if (!(pr > cut))
      cnt.fallthrough++;
    double kT = P.nuclides[i_nuclide].kT;
    elastic_scatter(P, i_nuclide, kT, p);
  }
  // collision(): weight windows after the reaction (src/physics.cpp:64-65)
  if (P.ww_ckpt_collision)
    apply_weight_windows(P, p);
  if (S.trace_hash) {
    mix(p.hash, 0x44);
    mix(p.dhash, 0x44);
    mix(p.hash, static_cast<uint64_t>(i_nuclide));
    mix(p.dhash, static_cast<uint64_t>(i_nuclide));
    mix(p.hash, absorbed ? 101 : 2);
    mix(p.dhash, absorbed ? 101 : 2);
    mix(p.hash, as_bits(p.E));
    mix(p.hash, as_bits(p.coord[0].u.x));
    mix(p.hash, as_bits(p.coord[0].u.y));
    mix(p.hash, as_bits(p.coord[0].u.z));
    if (P.ww_on) {
      mix(p.hash, as_bits(p.wgt));
      mix(p.hash, static_cast<uint64_t>(p.bank.size()));
      mix(p.dhash, static_cast<uint64_t>(p.bank.size()));
    }
  }
  if (p.E < P.energy_cutoff)
    p.wgt = 0.0;
  p.material_last = C_NONE;
  for (int j = 0; j < p.n_coord - 1; ++j)
    p.coord[j + 1].u = p.coord[j].u;
}

// Particle::event_death (src/particle.cpp:475-512)
void event_death(Sim& S, Particle& p)
{
#pragma omp atomic
  S.leakage_batch += p.leakage;
  p.leakage = 0.0;
}

//==============================================================================
// Transport loops -- src/simulation.cpp:804-883, src/event.cpp
//==============================================================================

void transport_history_based(Sim& S, int64_t n)
{
  std::vector<uint64_t>& hashes = S.hashes;
#pragma omp parallel num_threads(S.n_threads)
  {
    Counters c;
    Particle p;
#pragma omp for schedule(dynamic, 64)
    for (int64_t i_work = 1; i_work <= n; ++i_work) {
      initialize_history(S, p, i_work);
      while (p.alive()) {
        event_calculate_xs(S, p, c);
        if (p.alive())
          event_advance(S, p, c);
        if (p.alive()) {
          if (p.collision_distance > p.boundary.distance) {
            event_cross_surface(S, p, c);
          } else if (p.alive()) {
            event_collide(S, p, c);
          }
        }
        event_revive(S, p);
      }
      event_death(S, p);
      if (p.lost) {
#pragma omp atomic
        S.cnt.lost++;
      }
      hashes[i_work - 1] = p.hash;
      S.dhashes[i_work - 1] = p.dhash;
    }
#pragma omp critical
    {
      S.cnt.lookups += c.lookups;
      S.cnt.collisions += c.collisions;
      S.cnt.fallthrough += c.fallthrough;
      S.cnt.segments += c.segments;
      S.cnt.xs_events += c.xs_events;
      S.cnt.advance_events += c.advance_events;
      S.cnt.surface_events += c.surface_events;
      S.cnt.collision_events += c.collision_events;
      S.cnt.tracklength += c.tracklength;
    }
  }
}

// Event queues: longest queue first (src/simulation.cpp:852-874), one
// sub-iteration per max_particles_in_flight.
void transport_event_based(Sim& S, int64_t n_total)
{
  const Problem& P = S.P;
  int64_t remaining = n_total;
  int64_t source_offset = 0;
  while (remaining > 0) {
    int64_t n = std::min(remaining, P.max_in_flight);
    std::vector<Particle> particles(n);
    std::vector<int64_t> q_xs, q_adv, q_surf, q_coll;
    q_xs.reserve(n);
    q_adv.reserve(n);
    q_surf.reserve(n);
    q_coll.reserve(n);
#pragma omp parallel for num_threads(S.n_threads)
    for (int64_t i = 0; i < n; i++)
      initialize_history(S, particles[i], source_offset + i + 1);
    for (int64_t i = 0; i < n; i++)
      q_xs.push_back(i);
    Counters c;
    while (true) {
      size_t mx = std::max({q_xs.size(), q_adv.size(), q_surf.size(), q_coll.size()});
      if (mx == 0)
        break;
      if (mx == q_xs.size()) {
        // process_calculate_xs_events (src/event.cpp:73-105)
        for (int64_t idx : q_xs)
          event_calculate_xs(S, particles[idx], c);
        q_adv.insert(q_adv.end(), q_xs.begin(), q_xs.end());
        q_xs.clear();
      } else if (mx == q_adv.size()) {
        // process_advance_particle_events (src/event.cpp:107-128)
        for (int64_t idx : q_adv) {
          Particle& p = particles[idx];
          // a particle lost while locating its cell has no cell to advance from
          if (!p.alive())
            continue;
          event_advance(S, p, c);
          if (!p.alive())
            continue;
          if (p.collision_distance > p.boundary.distance)
            q_surf.push_back(idx);
          else
            q_coll.push_back(idx);
        }
        q_adv.clear();
      } else if (mx == q_surf.size()) {
        // process_surface_crossing_events (src/event.cpp:130-147)
        for (int64_t idx : q_surf) {
          Particle& p = particles[idx];
          event_cross_surface(S, p, c);
          event_revive(S, p);
          if (p.alive())
            q_xs.push_back(idx);
        }
        q_surf.clear();
      } else {
        // process_collision_events (src/event.cpp:149-166)
        for (int64_t idx : q_coll) {
          Particle& p = particles[idx];
          event_collide(S, p, c);
          event_revive(S, p);
          if (p.alive())
            q_xs.push_back(idx);
        }
        q_coll.clear();
      }
    }
    for (int64_t i = 0; i < n; i++) {
      event_death(S, particles[i]);
      if (particles[i].lost)
        S.cnt.lost++;
      S.hashes[source_offset + i] = particles[i].hash;
      S.dhashes[source_offset + i] = particles[i].dhash;
    }
    S.cnt.lookups += c.lookups;
    S.cnt.collisions += c.collisions;
    S.cnt.fallthrough += c.fallthrough;
    S.cnt.segments += c.segments;
    S.cnt.xs_events += c.xs_events;
    S.cnt.advance_events += c.advance_events;
    S.cnt.surface_events += c.surface_events;
    S.cnt.collision_events += c.collision_events;
    S.cnt.tracklength += c.tracklength;
    remaining -= n;
    source_offset += n;
  }
}

//==============================================================================
// Batch driver -- src/simulation.cpp:232-504, src/tallies/tally.cpp:821-1104
//==============================================================================

void run_batch(Sim& S, int mode, int64_t n_particles)
{
  S.current_batch++;
  S.leakage_batch = 0;
  S.total_weight = 0;
  std::fill(S.tally_value.begin(), S.tally_value.end(), 0.0);
  S.hashes.assign(n_particles, 0);
  S.dhashes.assign(n_particles, 0);
  if (mode == 0)
    transport_history_based(S, n_particles);
  else
    transport_event_based(S, n_particles);
  // accumulate_tallies: global leakage / total_weight, Tally::accumulate
  if (S.current_batch > S.P.n_inactive) {
    double val = S.leakage_batch / S.total_weight;
    S.leak_sum += val;
    S.leak_sum_sq += val * val;
    S.n_realizations++;
    double total_source = 1.0;
    double norm = total_source / (static_cast<double>(S.P.n_particles) * 1);
    for (int64_t i = 0; i < S.P.n_tally_bins; ++i) {
      double v = S.tally_value[i] * norm;
      S.results[3 * i + 1] += v;
      S.results[3 * i + 2] += v * v;
    }
  }
}

//==============================================================================
// Problem file reader (whitespace-separated tokens, written by oracle/problem.py)
//==============================================================================

struct Tok {
  std::istringstream in;
  std::string next()
  {
    std::string s;
    if (!(in >> s))
      throw std::runtime_error("unexpected end of problem file");
    return s;
  }
  double d() { return std::stod(next()); }
  long long i() { return std::stoll(next()); }
  void expect(const char* w)
  {
    std::string s = next();
    if (s != w)
      throw std::runtime_error(std::string("expected '") + w + "', got '" + s + "'");
  }
};

void read_nuclide_file(Nuclide& nuc, const std::string& path)
{
  std::ifstream f(path, std::ios::binary);
  if (!f)
    throw std::runtime_error("cannot open nuclide file " + path);
  char magic[8];
  f.read(magic, 8);
  if (std::memcmp(magic, "OMCNUC1", 7) != 0)
    throw std::runtime_error("bad nuclide file " + path);
  int32_t nlen;
  f.read(reinterpret_cast<char*>(&nlen), 4);
  nuc.name.resize(nlen);
  f.read(&nuc.name[0], nlen);
  f.read(reinterpret_cast<char*>(&nuc.awr), 8);
  int32_t nt;
  f.read(reinterpret_cast<char*>(&nt), 4);
  std::vector<double> kts(nt);
  f.read(reinterpret_cast<char*>(kts.data()), 8 * nt);
  nuc.kT = kts[0];
  int32_t n;
  f.read(reinterpret_cast<char*>(&n), 4);
  nuc.energy.resize(n);
  std::vector<double> el(n), cap(n);
  f.read(reinterpret_cast<char*>(nuc.energy.data()), 8 * n);
  f.read(reinterpret_cast<char*>(el.data()), 8 * n);
  f.read(reinterpret_cast<char*>(cap.data()), 8 * n);
  if (!f)
    throw std::runtime_error("truncated nuclide file " + path);
  // Nuclide::create_derived (src/nuclide.cpp:359-500): total = 0 + MT2 + MT102,
  // absorption = 0 + MT102 (disappearance), in reaction order.
  nuc.total.assign(n, 0.0);
  nuc.absorption.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    nuc.total[k] += el[k];
  }
  for (int k = 0; k < n; ++k) {
    nuc.total[k] += cap[k];
    nuc.absorption[k] += cap[k];
  }
  nuc.elastic = el;
}

// Material::normalize_density (src/material.cpp:417-467)
void normalize_density(Problem& P, Material& m, int units, double value)
{
  // units: 0 sum, 1 atom/b-cm, 2 atom/cc, 3 g/cc, 4 kg/m3
  double sum = 0.0;
  for (double a : m.atom_density)
    sum += a;
  if (units == 0)
    m.density = sum;
  else if (units == 1)
    m.density = value;
  else if (units == 2)
    m.density = 1.0e-24 * value;
  else if (units == 3)
    m.density = -value;
  else
    m.density = -1.0e-3 * value;
  bool percent_in_atom = (m.atom_density[0] >= 0.0);
  bool density_in_atom = (m.density >= 0.0);
  for (size_t i = 0; i < m.nuclide.size(); ++i) {
    double awr = P.nuclides[m.nuclide[i]].awr;
    if (!percent_in_atom)
      m.atom_density[i] = -m.atom_density[i] / awr;
  }
  double s = 0.0;
  for (double a : m.atom_density)
    s += a;
  for (double& a : m.atom_density)
    a /= s;
  if (!density_in_atom) {
    double sum_percent = 0.0;
    for (size_t i = 0; i < m.nuclide.size(); ++i)
      sum_percent += m.atom_density[i] * P.nuclides[m.nuclide[i]].awr;
    sum_percent = 1.0 / sum_percent;
    m.density = -m.density * N_AVOGADRO / MASS_NEUTRON * sum_percent;
  }
  for (double& a : m.atom_density)
    a *= m.density;
}

// DiscreteIndex::init_alias (src/distribution.cpp:40-80)
void init_alias(Problem& P)
{
  size_t n = P.energy_p.size();
  std::vector<double> p = P.energy_p;
  double s = 0.0; // std::accumulate, left to right
  for (double v : p)
    s += v;
  for (double& v : p)
    v /= s;
  P.alias_prob.assign(n, 0.0);
  P.alias_idx.assign(n, 0);
  std::vector<size_t> large, small;
  for (size_t i = 0; i < n; i++) {
    P.alias_prob[i] = n * p[i];
    if (P.alias_prob[i] > 1.0)
      large.push_back(i);
    else
      small.push_back(i);
  }
  while (!large.empty() && !small.empty()) {
    size_t j = small.back();
    size_t k = large.back();
    small.pop_back();
    P.alias_prob[k] += P.alias_prob[j] - 1.0;
    P.alias_idx[j] = static_cast<int>(k);
    if (P.alias_prob[k] < 1.0) {
      small.push_back(k);
      large.pop_back();
    }
  }
}

void load_problem(Problem& P, const std::string& path)
{
  std::ifstream f(path);
  if (!f)
    throw std::runtime_error("cannot open problem file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  Tok t;
  t.in.str(ss.str());
  std::string dir = path.substr(0, path.find_last_of('/') + 1);

  t.expect("oracle_problem");
  t.expect("1");
  t.expect("settings");
  P.n_particles = t.i();
  P.n_batches = static_cast<int>(t.i());
  P.n_inactive = static_cast<int>(t.i());
  P.rng.master_seed = static_cast<uint64_t>(t.i());
  P.n_log_bins = static_cast<int>(t.i());
  P.max_events = t.i();
  P.max_in_flight = t.i();

  t.expect("nuclides");
  int nn = static_cast<int>(t.i());
  P.nuclides.resize(nn);
  for (int k = 0; k < nn; ++k) {
    std::string name = t.next();
    std::string file = t.next();
    read_nuclide_file(P.nuclides[k], file[0] == '/' ? file : dir + file);
  }

  t.expect("materials");
  int nm = static_cast<int>(t.i());
  std::unordered_map<int, int> mat_map;
  P.materials.resize(nm);
  for (int k = 0; k < nm; ++k) {
    Material& m = P.materials[k];
    m.id = static_cast<int>(t.i());
    mat_map[m.id] = k;
    int units = static_cast<int>(t.i());
    double value = t.d();
    int n = static_cast<int>(t.i());
    for (int j = 0; j < n; ++j) {
      m.nuclide.push_back(static_cast<int>(t.i()));
      m.atom_density.push_back(t.d());
    }
    normalize_density(P, m, units, value);
  }

  t.expect("surfaces");
  int ns = static_cast<int>(t.i());
  std::unordered_map<int, int> surf_map;
  P.surfaces.resize(ns);
  for (int k = 0; k < ns; ++k) {
    Surface& s = P.surfaces[k];
    s.id = static_cast<int>(t.i());
    surf_map[s.id] = k;
    s.type = static_cast<int>(t.i());
    s.bc = static_cast<int>(t.i());
    for (int j = 0; j < 4; ++j)
      s.c[j] = t.d();
  }

  t.expect("universes");
  int nu = static_cast<int>(t.i());
  std::unordered_map<int, int> univ_map;
  P.universes.resize(nu);
  for (int k = 0; k < nu; ++k) {
    P.universes[k].id = static_cast<int>(t.i());
    univ_map[P.universes[k].id] = k;
  }

  t.expect("lattices");
  int nl = static_cast<int>(t.i());
  std::unordered_map<int, int> lat_map;
  P.lattices.resize(nl);
  std::vector<std::vector<int>> lat_uids(nl);
  std::vector<int> lat_outer(nl);
  for (int k = 0; k < nl; ++k) {
    Lattice& L = P.lattices[k];
    L.id = static_cast<int>(t.i());
    lat_map[L.id] = k;
    L.n[0] = static_cast<int>(t.i());
    L.n[1] = static_cast<int>(t.i());
    L.n[2] = static_cast<int>(t.i());
    L.is_3d = t.i() != 0;
    L.ll = {t.d(), t.d(), t.d()};
    L.pitch = {t.d(), t.d(), t.d()};
    lat_outer[k] = static_cast<int>(t.i());
    int cnt = L.n[0] * L.n[1] * L.n[2];
    lat_uids[k].resize(cnt);
    for (int j = 0; j < cnt; ++j)
      lat_uids[k][j] = static_cast<int>(t.i());
  }

  t.expect("cells");
  int nc = static_cast<int>(t.i());
  P.cells.resize(nc);
  for (int k = 0; k < nc; ++k) {
    Cell& c = P.cells[k];
    c.id = static_cast<int>(t.i());
    int uid = static_cast<int>(t.i());
    c.universe = univ_map.at(uid);
    c.fill_type = static_cast<int>(t.i());
    int fid = static_cast<int>(t.i());
    if (c.fill_type == FILL_MATERIAL) {
      c.material = fid < 0 ? MATERIAL_VOID : mat_map.at(fid);
      c.fill = C_NONE;
    } else if (c.fill_type == FILL_UNIVERSE) {
      c.fill = univ_map.at(fid);
      c.material = C_NONE;
    } else {
      c.fill = lat_map.at(fid);
      c.material = C_NONE;
    }
    int ntok = static_cast<int>(t.i());
    for (int j = 0; j < ntok; ++j) {
      int tok = static_cast<int>(t.i());
      int idx = surf_map.at(std::abs(tok)) + 1;
      c.region.push_back(tok > 0 ? idx : -idx);
    }
    P.universes[c.universe].cells.push_back(k);
  }
  // RectLattice universes in reference order (src/lattice.cpp:233-244):
  // XML lists rows from top (max y) to bottom.
  for (int k = 0; k < nl; ++k) {
    Lattice& L = P.lattices[k];
    L.universes.assign(L.n[0] * L.n[1] * L.n[2], C_NONE);
    for (int iz = 0; iz < L.n[2]; iz++)
      for (int iy = L.n[1] - 1; iy > -1; iy--)
        for (int ix = 0; ix < L.n[0]; ix++) {
          int indx1 = L.n[0] * L.n[1] * iz + L.n[0] * (L.n[1] - iy - 1) + ix;
          int indx2 = L.n[0] * L.n[1] * iz + L.n[0] * iy + ix;
          L.universes[indx1] = univ_map.at(lat_uids[k][indx2]);
        }
    L.outer = lat_outer[k] < 0 ? C_NONE : univ_map.at(lat_outer[k]);
  }

  t.expect("root");
  P.root_universe = univ_map.at(static_cast<int>(t.i()));
  P.n_coord_levels = static_cast<int>(t.i());

  t.expect("source");
  P.space_type = static_cast<int>(t.i());
  P.space_p0 = {t.d(), t.d(), t.d()};
  P.space_p1 = {t.d(), t.d(), t.d()};
  P.angle_type = static_cast<int>(t.i());
  P.angle_u = {t.d(), t.d(), t.d()};
  P.energy_type = static_cast<int>(t.i());
  if (P.energy_type == 0) {
    int n = static_cast<int>(t.i());
    for (int j = 0; j < n; ++j) {
      P.energy_x.push_back(t.d());
      P.energy_p.push_back(t.d());
    }
    init_alias(P);
  } else {
    P.watt_a = t.d();
    P.watt_b = t.d();
  }
  int nd = static_cast<int>(t.i());
  for (int j = 0; j < nd; ++j)
    P.domain_cell_ids.push_back(static_cast<int>(t.i()));

  t.expect("tally");
  P.has_tally = t.i() != 0;
  if (P.has_tally) {
    Mesh& m = P.mesh;
    m.n_dim = static_cast<int>(t.i());
    for (int j = 0; j < 3; ++j)
      m.shape[j] = static_cast<int>(t.i());
    for (int j = 0; j < 3; ++j)
      m.ll[j] = t.d();
    for (int j = 0; j < 3; ++j)
      m.ur[j] = t.d();
    for (int j = 0; j < m.n_dim; ++j)
      m.width[j] = (m.ur[j] - m.ll[j]) / static_cast<double>(m.shape[j]);
    P.has_mesh = m.n_dim > 0;
    int ne = static_cast<int>(t.i());
    for (int j = 0; j < ne; ++j)
      P.energy_bins.push_back(t.d());
    int64_t nb = 1;
    for (int j = 0; j < m.n_dim; ++j)
      nb *= m.shape[j];
    P.n_tally_bins = nb * (P.energy_bins.empty() ? 1 : static_cast<int64_t>(P.energy_bins.size()) - 1);
  }
  t.expect("ww");
  const int n_sets = static_cast<int>(t.i());
  P.ww_on = n_sets > 0;
  for (int k = 0; k < n_sets; ++k) {
    Problem::WWSet W;
    Mesh& m = W.mesh;
    m.n_dim = static_cast<int>(t.i());
    for (int j = 0; j < 3; ++j)
      m.shape[j] = static_cast<int>(t.i());
    for (int j = 0; j < 3; ++j)
      m.ll[j] = t.d();
    for (int j = 0; j < 3; ++j)
      m.ur[j] = t.d();
    for (int j = 0; j < m.n_dim; ++j)
      m.width[j] = (m.ur[j] - m.ll[j]) / static_cast<double>(m.shape[j]);
    W.n_mesh = 1;
    for (int j = 0; j < m.n_dim; ++j)
      W.n_mesh *= m.shape[j];
    const int ne = static_cast<int>(t.i());
    for (int j = 0; j < ne; ++j)
      W.ebounds.push_back(t.d());
    const int64_t nbnd = t.i();
    if (ne < 2 || nbnd != W.n_mesh * (ne - 1))
      throw std::runtime_error("weight window bounds do not match the mesh and energy bounds");
    W.lower.resize(nbnd);
    W.upper.resize(nbnd);
    for (int64_t j = 0; j < nbnd; ++j)
      W.lower[j] = t.d();
    for (int64_t j = 0; j < nbnd; ++j)
      W.upper[j] = t.d();
    W.survival_ratio = t.d();
    W.max_lb_ratio = t.d();
    W.max_split = static_cast<int>(t.i());
    W.weight_cutoff = t.d();
    P.ww_sets.push_back(std::move(W));
  }
  if (P.ww_on) {
    P.ww_ckpt_collision = t.i() != 0;
    P.ww_ckpt_surface = t.i() != 0;
    P.max_history_splits = t.i();
  } else {
    P.ww_ckpt_collision = false;
    P.ww_ckpt_surface = false;
  }
  t.expect("end");
}

} // namespace oracle

//==============================================================================
// C API for tests / bench cpu_baseline (ctypes)
//==============================================================================

using namespace oracle;

extern "C" {

static char g_err[512];

const char* oracle_last_error() { return g_err; }

void* oracle_load(const char* path)
{
  try {
    auto* S = new Sim();
    load_problem(S->P, path);
    initialize_data(S->P);
    S->tally_value.assign(S->P.n_tally_bins, 0.0);
    S->results.assign(3 * S->P.n_tally_bins, 0.0);
    return S;
  } catch (const std::exception& e) {
    std::snprintf(g_err, sizeof(g_err), "%s", e.what());
    return nullptr;
  }
}

void oracle_free(void* h) { delete static_cast<Sim*>(h); }

// mode 0 = history-based, 1 = event-based.  Runs `n_batches` batches of
// `n_particles` (<= 0 means the problem's own values).
int oracle_run(void* h, int mode, int n_threads, int64_t n_particles, int n_batches, int trace_hash)
{
  Sim& S = *static_cast<Sim*>(h);
  try {
    if (n_particles > 0)
      S.P.n_particles = n_particles;
    if (n_batches <= 0)
      n_batches = S.P.n_batches;
    S.n_threads = n_threads > 0 ? n_threads : 1;
    S.trace_hash = trace_hash != 0;
    for (int b = 0; b < n_batches; ++b)
      run_batch(S, mode, S.P.n_particles);
  } catch (const std::exception& e) {
    std::snprintf(g_err, sizeof(g_err), "%s", e.what());
    return -1;
  }
  return 0;
}

int64_t oracle_n_tally_bins(void* h) { return static_cast<Sim*>(h)->P.n_tally_bins; }
const double* oracle_results(void* h) { return static_cast<Sim*>(h)->results.data(); }
const double* oracle_batch_values(void* h) { return static_cast<Sim*>(h)->tally_value.data(); }
const uint64_t* oracle_hashes(void* h) { return static_cast<Sim*>(h)->hashes.data(); }
const uint64_t* oracle_dhashes(void* h) { return static_cast<Sim*>(h)->dhashes.data(); }
int oracle_n_realizations(void* h) { return static_cast<Sim*>(h)->n_realizations; }

void oracle_counters(void* h, int64_t* out)
{
  Counters& c = static_cast<Sim*>(h)->cnt;
  out[0] = c.lookups;
  out[1] = c.collisions;
  out[2] = c.segments;
  out[3] = c.lost;
  out[4] = c.xs_events;
  out[5] = c.advance_events;
  out[6] = c.surface_events;
  out[7] = c.collision_events;
  out[8] = c.fallthrough;
}

void oracle_leakage(void* h, double* out)
{
  Sim& S = *static_cast<Sim*>(h);
  out[0] = S.leak_sum;
  out[1] = S.leak_sum_sq;
  out[2] = S.leakage_batch;
  out[3] = S.total_weight;
  out[4] = S.cnt.tracklength;
}

void oracle_energy_grid(void* h, double* out)
{
  Sim& S = *static_cast<Sim*>(h);
  out[0] = S.P.energy_min;
  out[1] = S.P.energy_max;
  out[2] = S.P.log_spacing;
}

const int* oracle_grid_index(void* h, int i_nuc) { return static_cast<Sim*>(h)->P.nuclides[i_nuc].grid_index.data(); }

// One Nuclide::calculate_xs lookup at energy E: writes i_grid, f, total, absorption, elastic.
void oracle_lookup(void* h, int i_nuc, double E, int* i_grid, double* out)
{
  Sim& S = *static_cast<Sim*>(h);
  Particle p;
  p.micro.assign(S.P.nuclides.size(), MicroXS {0, 0, 0, 0, 0, 0.0});
  p.E = E;
  Counters c;
  int i_log = static_cast<int>(m_log(E / S.P.energy_min) / S.P.log_spacing);
  i_log = i_log < 0 ? 0 : (i_log >= S.P.n_log_bins ? S.P.n_log_bins - 1 : i_log);
  nuclide_calculate_xs(S.P, i_nuc, i_log, p, c);
  calculate_elastic_xs(S.P, i_nuc, p);
  *i_grid = p.micro[i_nuc].index_grid;
  out[0] = p.micro[i_nuc].interp_factor;
  out[1] = p.micro[i_nuc].total;
  out[2] = p.micro[i_nuc].absorption;
  out[3] = p.micro[i_nuc].elastic;
}

// Unit-level restatements used by the known-answer tests.
// apply_weight_windows on one particle of the loaded problem:
//   in  = {x, y, z, u, v, w, E, wgt, wgt_born, wgt_ww_born, ww_factor, n_split}
//   out = {wgt, wgt_ww_born, ww_factor, n_split, n_bank}
// bank_w receives the weights of the banked copies (at most bank_cap).
int oracle_ww_apply(void* h, const double* in, uint64_t* seed, double* out, double* bank_w, int bank_cap)
{
  const Problem& P = static_cast<Sim*>(h)->P;
  Particle p;
  p.coord.resize(1);
  p.coord[0].r = {in[0], in[1], in[2]};
  p.coord[0].u = {in[3], in[4], in[5]};
  p.E = in[6];
  p.wgt = in[7];
  p.wgt_born = in[8];
  p.wgt_ww_born = in[9];
  p.ww_factor = in[10];
  p.n_split = static_cast<int>(in[11]);
  p.stream = STREAM_TRACKING;
  p.seeds[STREAM_TRACKING] = *seed;
  apply_weight_windows(P, p);
  *seed = p.seeds[STREAM_TRACKING];
  out[0] = p.wgt;
  out[1] = p.wgt_ww_born;
  out[2] = p.ww_factor;
  out[3] = p.n_split;
  out[4] = static_cast<double>(p.bank.size());
  for (size_t k = 0; k < p.bank.size() && static_cast<int>(k) < bank_cap; ++k)
    bank_w[k] = p.bank[k].wgt;
  return static_cast<int>(p.bank.size());
}

double oracle_prn(uint64_t* seed) { return prn(seed); }
uint64_t oracle_future_seed(uint64_t n, uint64_t seed) { return future_seed(n, seed); }
double oracle_maxwell(double T, uint64_t* seed) { return maxwell_spectrum(T, seed); }
double oracle_watt(double a, double b, uint64_t* seed) { return watt_spectrum(a, b, seed); }
void oracle_rotate_angle(const double* u, double mu, const double* phi, uint64_t* seed, double* out)
{
  Pos r = rotate_angle({u[0], u[1], u[2]}, mu, phi, seed);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}
// Elementary pieces of the restatement, for the pins in tests/test_oracle.py
// (each checked there against what the mathematics fixes).

// Surface::distance (src/surface.cpp:208-218, 354-366, 400-442) of one
// surface of the given type and coefficients.
double oracle_surface_distance(int type, const double* c, const double* r, const double* u, int coincident)
{
  Surface s;
  s.id = 1;
  s.type = type;
  s.bc = BC_TRANSMIT;
  for (int k = 0; k < 4; ++k)
    s.c[k] = c[k];
  return s.distance({r[0], r[1], r[2]}, {u[0], u[1], u[2]}, coincident != 0);
}

// RectLattice::distance and get_indices (src/lattice.cpp:252-342) of lattice
// `lat` of a loaded problem; r is in the lattice's own (local) frame for the
// distance and in the containing universe's frame for the indices.
double oracle_lattice_distance(void* h, int lat, const double* r, const double* u, int* trans)
{
  const Problem& P = static_cast<Sim*>(h)->P;
  return P.lattices.at(lat).distance({r[0], r[1], r[2]}, {u[0], u[1], u[2]}, trans);
}

void oracle_lattice_indices(void* h, int lat, const double* r, const double* u, int* ijk)
{
  const Problem& P = static_cast<Sim*>(h)->P;
  P.lattices.at(lat).get_indices({r[0], r[1], r[2]}, {u[0], u[1], u[2]}, ijk);
}

// A particle at r going along u: find_cell, then distance_to_boundary.
// out_i = {found, n_coord, cell of each level (4), lattice index x/y/z of the
// deepest level, surface (signed index + 1), coord level, translation x/y/z};
// returns the distance.
double oracle_locate(void* h, const double* r, const double* u, int* out_i)
{
  const Problem& P = static_cast<Sim*>(h)->P;
  Particle p;
  p.coord.resize(P.n_coord_levels);
  p.cell_last.resize(P.n_coord_levels);
  p.coord[0].r = {r[0], r[1], r[2]};
  p.coord[0].u = {u[0], u[1], u[2]};
  bool found = exhaustive_find_cell(P, p);
  for (int k = 0; k < 15; ++k)
    out_i[k] = -1;
  out_i[0] = found ? 1 : 0;
  if (!found)
    return -1.0;
  out_i[1] = p.n_coord;
  for (int k = 0; k < p.n_coord && k < 4; ++k)
    out_i[2 + k] = p.coord[k].cell;
  const LocalCoord& lo = p.coord[p.n_coord - 1];
  for (int k = 0; k < 3; ++k)
    out_i[6 + k] = lo.lattice == C_NONE ? -1 : lo.lattice_index[k];
  Boundary b = distance_to_boundary(P, p);
  out_i[9] = b.surface;
  out_i[10] = b.coord_level;
  for (int k = 0; k < 3; ++k)
    out_i[11 + k] = b.lattice_translation[k];
  return b.distance;
}

// sample_cxs_target_velocity (src/physics.cpp:990-1041): out3 = target velocity
void oracle_target_velocity(double awr, double E, const double* u, double kT, uint64_t* seed, double* out3)
{
  Pos v = sample_cxs_target_velocity(awr, E, {u[0], u[1], u[2]}, kT, seed);
  out3[0] = v.x;
  out3[1] = v.y;
  out3[2] = v.z;
}

// elastic_scatter (src/physics.cpp:754-818) off one nuclide of atomic weight
// ratio awr at temperature kT: out4 = {E', u'}
void oracle_elastic(double awr, double kT, double free_gas_threshold, double E, const double* u, uint64_t* seed,
  double* out4)
{
  Problem P;
  P.free_gas_threshold = free_gas_threshold;
  Nuclide nuc;
  nuc.awr = awr;
  nuc.kT = kT;
  P.nuclides.push_back(nuc);
  Particle p;
  p.coord.resize(1);
  p.E = E;
  p.coord[0].u = {u[0], u[1], u[2]};
  p.seeds[STREAM_TRACKING] = *seed;
  p.stream = STREAM_TRACKING;
  elastic_scatter(P, 0, kT, p);
  *seed = p.seeds[STREAM_TRACKING];
  out4[0] = p.E;
  out4[1] = p.coord[0].u.x;
  out4[2] = p.coord[0].u.y;
  out4[3] = p.coord[0].u.z;
}

// raytrace_mesh over the problem's tally mesh for the segment r0 -> r1 along
// u: the crossed bins and their track fractions in order; returns how many
// there are (at most cap are written).
int oracle_raytrace(void* h, const double* r0, const double* r1, const double* u, int* bins, double* fracs,
  int cap)
{
  const Problem& P = static_cast<Sim*>(h)->P;
  int n = 0;
  raytrace_mesh(P.mesh, {r0[0], r0[1], r0[2]}, {r1[0], r1[1], r1[2]}, {u[0], u[1], u[2]},
    [&](const int* ijk, double w) {
      if (n < cap) {
        bins[n] = mesh_bin(P.mesh, ijk);
        fracs[n] = w;
      }
      ++n;
    });
  return n;
}

// Source::satisfies_spatial_constraints (src/source.cpp:226-271)
int oracle_in_source_domain(void* h, const double* r)
{
  return satisfies_spatial_constraints(static_cast<Sim*>(h)->P, {r[0], r[1], r[2]}) ? 1 : 0;
}

// IndependentSource::sample (src/source.cpp:355-417): out7 = {r, u, E}
void oracle_sample_source(void* h, uint64_t* seed, double* out7)
{
  SourceSite site = sample_source(static_cast<Sim*>(h)->P, seed);
  out7[0] = site.r.x;
  out7[1] = site.r.y;
  out7[2] = site.r.z;
  out7[3] = site.u.x;
  out7[4] = site.u.y;
  out7[5] = site.u.z;
  out7[6] = site.E;
}

double oracle_log(double x) { return m_log(x); }
double oracle_sin(double x) { return m_sin(x); }
double oracle_cos(double x) { return m_cos(x); }

} // extern "C"
