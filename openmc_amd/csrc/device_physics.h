// Device-side physics of the event kernels: RNG, CSG geometry, cross-section
// lookup, collision kinematics and the mesh ray-march.
//
// Each function restates the reference function cited beside it in the same
// floating-point operation order (compiled with -ffp-contract=off), so that a
// particle's history on the GPU is bit-identical to the reference algorithm.
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.h"
#include "openmc_amd/detmath.h"

namespace omc {

//==============================================================================
// RNG (src/random_lcg.cpp:32-120)
//==============================================================================

__device__ __forceinline__ double prn(uint64_t* seed)
{
  *seed = (6364136223846793005ULL * (*seed) + 1442695040888963407ULL);
  uint64_t word = ((*seed >> ((*seed >> 59u) + 5u)) ^ *seed) * 12605985483714917081ull;
  uint64_t result = (word >> 43u) ^ word;
  return __builtin_ldexp(static_cast<double>(result), -64);
}

__host__ __device__ __forceinline__ uint64_t future_seed(uint64_t n, uint64_t seed)
{
  uint64_t g = 6364136223846793005ULL;
  uint64_t c = 1442695040888963407ULL;
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

__device__ __forceinline__ double uniform_distribution(double a, double b, uint64_t* seed)
{
  return a + (b - a) * prn(seed);
}

//==============================================================================
// 3-vectors with the reference's Position semantics (include/openmc/position.h)
//==============================================================================

struct V3 {
  double x, y, z;
};

__device__ __forceinline__ V3 v_add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 v_sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
// double * Position: each component times the scalar
__device__ __forceinline__ V3 v_scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 v_div(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ double v_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double v_norm(V3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ double v_get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// src/math_functions.cpp:763-792
__device__ __forceinline__ V3 rotate_angle(V3 u, double mu, uint64_t* seed)
{
  double phi_ = 2.0 * PI * prn(seed);
  double sinphi, cosphi;
  detm::sincos(phi_, &sinphi, &cosphi);
  double a = sqrt(fmax(0., 1. - mu * mu));
  double b = sqrt(fmax(0., 1. - u.z * u.z));
  if (b > 1e-10) {
    return {mu * u.x + a * (u.x * u.z * cosphi - u.y * sinphi) / b,
      mu * u.y + a * (u.y * u.z * cosphi + u.x * sinphi) / b, mu * u.z - a * b * cosphi};
  } else {
    b = sqrt(1. - u.y * u.y);
    return {mu * u.x + a * (-u.x * u.y * sinphi + u.z * cosphi) / b, mu * u.y + a * b * sinphi,
      mu * u.z - a * (u.y * u.z * sinphi + u.x * cosphi) / b};
  }
}

//==============================================================================
// Geometry state of one particle held in registers: NL coordinate levels,
// one scalar array per component so that every array is small enough to live
// in VGPRs (no scratch) once the level loops are unrolled.
//==============================================================================

// The direction is one vector for all levels: the reference copies it down
// unchanged when descending (src/geometry.cpp:196-197, 216-217) and after
// collisions (src/particle.cpp:396-406), and no cell on this path is rotated.
template<int NL>
struct Geo {
  double rx[NL], ry[NL], rz[NL];
  double ux, uy, uz;
  int cell[NL], univ[NL], lat[NL];
  int lix[NL], liy[NL], liz[NL];
  int n_coord;
  int surface;
  int material;
  int material_last;
  int lost;

  __device__ __forceinline__ V3 r(int i) const { return {rx[i], ry[i], rz[i]}; }
  __device__ __forceinline__ V3 u() const { return {ux, uy, uz}; }
  __device__ __forceinline__ void set_r(int i, V3 v)
  {
    rx[i] = v.x;
    ry[i] = v.y;
    rz[i] = v.z;
  }
  __device__ __forceinline__ void set_u(V3 v)
  {
    ux = v.x;
    uy = v.y;
    uz = v.z;
  }
  __device__ __forceinline__ void reset(int i)
  {
    cell[i] = C_NONE;
    univ[i] = C_NONE;
    lat[i] = C_NONE;
    lix[i] = 0;
    liy[i] = 0;
    liz[i] = 0;
  }
};

//==============================================================================
// Surfaces (src/surface.cpp)
//==============================================================================

__device__ __forceinline__ double surf_evaluate(const DevSurface& s, V3 r)
{
  switch (s.type) {
  case S_XPLANE: return r.x - s.c[0];
  case S_YPLANE: return r.y - s.c[0];
  case S_ZPLANE: return r.z - s.c[0];
  case S_PLANE: return s.c[0] * r.x + s.c[1] * r.y + s.c[2] * r.z - s.c[3];
  case S_ZCYL: {
    const double r1 = r.x - s.c[0];
    const double r2 = r.y - s.c[1];
    return r1 * r1 + r2 * r2 - s.c[2] * s.c[2];
  }
  case S_XCYL: {
    const double r1 = r.y - s.c[0];
    const double r2 = r.z - s.c[1];
    return r1 * r1 + r2 * r2 - s.c[2] * s.c[2];
  }
  case S_YCYL: {
    const double r1 = r.x - s.c[0];
    const double r2 = r.z - s.c[1];
    return r1 * r1 + r2 * r2 - s.c[2] * s.c[2];
  }
  default: {
    const double x = r.x - s.c[0];
    const double y = r.y - s.c[1];
    const double z = r.z - s.c[2];
    return x * x + y * y + z * z - s.c[3] * s.c[3];
  }
  }
}

// axis_aligned_plane_distance (src/surface.cpp:208-218)
__device__ __forceinline__ double plane_distance(double ri, double ui, bool coincident, double offset)
{
  const double f = offset - ri;
  if (coincident || fabs(f) < FP_COINCIDENT || ui == 0.0)
    return INFTY;
  const double d = f / ui;
  if (d < 0.0)
    return INFTY;
  return d;
}

// axis_aligned_cylinder_distance (src/surface.cpp:403-442)
__device__ __forceinline__ double cyl_distance(double u1, double r2p, double u2, double r3p, double u3,
  bool coincident, double off1, double off2, double radius)
{
  const double a = 1.0 - u1 * u1;
  if (a == 0.0)
    return INFTY;
  const double r2 = r2p - off1;
  const double r3 = r3p - off2;
  const double k = r2 * u2 + r3 * u3;
  const double c = r2 * r2 + r3 * r3 - radius * radius;
  const double quad = k * k - a * c;
  if (quad < 0.0) {
    return INFTY;
  } else if (coincident || fabs(c) < FP_COINCIDENT) {
    if (k >= 0.0)
      return INFTY;
    return (-k + sqrt(quad)) / a;
  } else if (c < 0.0) {
    return (-k + sqrt(quad)) / a;
  } else {
    const double d = (-k - sqrt(quad)) / a;
    if (d < 0.0)
      return INFTY;
    return d;
  }
}

__device__ __forceinline__ double surf_distance(const DevSurface& s, V3 r, V3 u, bool coincident)
{
  switch (s.type) {
  case S_XPLANE: return plane_distance(r.x, u.x, coincident, s.c[0]);
  case S_YPLANE: return plane_distance(r.y, u.y, coincident, s.c[0]);
  case S_ZPLANE: return plane_distance(r.z, u.z, coincident, s.c[0]);
  case S_PLANE: {
    const double f = s.c[0] * r.x + s.c[1] * r.y + s.c[2] * r.z - s.c[3];
    const double projection = s.c[0] * u.x + s.c[1] * u.y + s.c[2] * u.z;
    if (coincident || fabs(f) < FP_COINCIDENT || projection == 0.0)
      return INFTY;
    const double d = -f / projection;
    if (d < 0.0)
      return INFTY;
    return d;
  }
  case S_ZCYL: return cyl_distance(u.z, r.x, u.x, r.y, u.y, coincident, s.c[0], s.c[1], s.c[2]);
  case S_XCYL: return cyl_distance(u.x, r.y, u.y, r.z, u.z, coincident, s.c[0], s.c[1], s.c[2]);
  case S_YCYL: return cyl_distance(u.y, r.x, u.x, r.z, u.z, coincident, s.c[0], s.c[1], s.c[2]);
  default: {
    const double x = r.x - s.c[0];
    const double y = r.y - s.c[1];
    const double z = r.z - s.c[2];
    const double k = x * u.x + y * u.y + z * u.z;
    const double c = x * x + y * y + z * z - s.c[3] * s.c[3];
    const double quad = k * k - c;
    if (quad < 0.0) {
      return INFTY;
    } else if (coincident || fabs(c) < FP_COINCIDENT) {
      if (k >= 0.0)
        return INFTY;
      return -k + sqrt(quad);
    } else if (c < 0.0) {
      return -k + sqrt(quad);
    } else {
      const double d = -k - sqrt(quad);
      if (d < 0.0)
        return INFTY;
      return d;
    }
  }
  }
}

__device__ __forceinline__ V3 surf_normal(const DevSurface& s, V3 r)
{
  switch (s.type) {
  case S_XPLANE: return {1., 0., 0.};
  case S_YPLANE: return {0., 1., 0.};
  case S_ZPLANE: return {0., 0., 1.};
  case S_PLANE: return {s.c[0], s.c[1], s.c[2]};
  case S_ZCYL: return {2.0 * (r.x - s.c[0]), 2.0 * (r.y - s.c[1]), 0.0};
  case S_XCYL: return {0.0, 2.0 * (r.y - s.c[0]), 2.0 * (r.z - s.c[1])};
  case S_YCYL: return {2.0 * (r.x - s.c[0]), 0.0, 2.0 * (r.z - s.c[1])};
  default: return {2.0 * (r.x - s.c[0]), 2.0 * (r.y - s.c[1]), 2.0 * (r.z - s.c[2])};
  }
}

// Surface::sense (src/surface.cpp:116-130)
__device__ __forceinline__ bool surf_sense(const DevSurface& s, V3 r, V3 u)
{
  const double f = surf_evaluate(s, r);
  if (fabs(f) < FP_COINCIDENT)
    return v_dot(u, surf_normal(s, r)) > 0.0;
  return f > 0.0;
}

//==============================================================================
// Cells, universes, lattices (src/cell.cpp, src/universe.cpp, src/lattice.cpp)
//==============================================================================

// Region::contains_simple (src/cell.cpp:951-970)
__device__ __forceinline__ bool cell_contains(const DevProblem& P, const DevCell& c, V3 r, V3 u, int on_surface)
{
  for (int k = c.tok_begin; k < c.tok_end; ++k) {
    const DevTok t = P.ctok[k];
    const int token = t.token;
    if (token == on_surface) {
    } else if (-token == on_surface) {
      return false;
    } else {
      bool sense = surf_sense(t.s, r, u);
      if (sense != (token > 0))
        return false;
    }
  }
  return true;
}

// Region::distance (src/cell.cpp:910-936)
__device__ __forceinline__ double cell_distance(const DevProblem& P, const DevCell& c, V3 r, V3 u, int on_surface,
  int* i_surf)
{
  double min_dist = INFTY;
  int isurf = 2147483647;
  for (int k = c.tok_begin; k < c.tok_end; ++k) {
    const DevTok t = P.ctok[k];
    const int token = t.token;
    bool coincident = abs(token) == abs(on_surface);
    double d = surf_distance(t.s, r, u, coincident);
    if (d < min_dist) {
      if (min_dist - d >= FP_PRECISION * min_dist) {
        min_dist = d;
        isurf = -token;
      }
    }
  }
  *i_surf = isurf;
  return min_dist;
}

// RectLattice::distance (src/lattice.cpp:252-306)
__device__ __forceinline__ double lattice_distance(const DevLattice& L, V3 r, V3 u, int* trans)
{
  double x0 = copysign(0.5 * L.pitch[0], u.x);
  double y0 = copysign(0.5 * L.pitch[1], u.y);
  double d = INFTY;
  trans[0] = trans[1] = trans[2] = 0;
  if ((fabs(r.x - x0) > FP_PRECISION) && u.x != 0) {
    d = (x0 - r.x) / u.x;
    trans[0] = u.x > 0 ? 1 : -1;
    trans[1] = 0;
    trans[2] = 0;
  }
  if ((fabs(r.y - y0) > FP_PRECISION) && u.y != 0) {
    double this_d = (y0 - r.y) / u.y;
    if (this_d < d) {
      d = this_d;
      trans[0] = 0;
      trans[1] = u.y > 0 ? 1 : -1;
      trans[2] = 0;
    }
  }
  if (L.is_3d) {
    double z0 = copysign(0.5 * L.pitch[2], u.z);
    if ((fabs(r.z - z0) > FP_PRECISION) && u.z != 0) {
      double this_d = (z0 - r.z) / u.z;
      if (this_d < d) {
        d = this_d;
        trans[0] = 0;
        trans[1] = 0;
        trans[2] = u.z > 0 ? 1 : -1;
      }
    }
  }
  return d;
}

// RectLattice::get_indices (src/lattice.cpp:310-342)
__device__ __forceinline__ void lattice_indices(const DevLattice& L, V3 r, V3 u, int* res)
{
  const double rr[3] = {r.x, r.y, r.z};
  const double uu[3] = {u.x, u.y, u.z};
  res[2] = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (k == 2 && !L.is_3d)
      break;
    double i_ = (rr[k] - L.ll[k]) / L.pitch[k];
    long i_close = detm::lround(i_);
    if (fabs(i_ - static_cast<double>(i_close)) < FP_COINCIDENT) {
      res[k] = (uu[k] > 0) ? static_cast<int>(i_close) : static_cast<int>(i_close - 1);
    } else {
      res[k] = static_cast<int>(floor(i_));
    }
  }
}

// RectLattice::get_local_position (src/lattice.cpp:352-361)
__device__ __forceinline__ V3 lattice_local(const DevLattice& L, V3 r, const int* i)
{
  r.x -= (L.ll[0] + (i[0] + 0.5) * L.pitch[0]);
  r.y -= (L.ll[1] + (i[1] + 0.5) * L.pitch[1]);
  if (L.is_3d)
    r.z -= (L.ll[2] + (i[2] + 0.5) * L.pitch[2]);
  return r;
}

__device__ __forceinline__ bool lattice_valid(const DevLattice& L, const int* i)
{
  return (i[0] >= 0) && (i[0] < L.n[0]) && (i[1] >= 0) && (i[1] < L.n[1]) && (i[2] >= 0) &&
         (i[2] < L.n[2]);
}

__device__ __forceinline__ int lattice_universe(const DevProblem& P, const DevLattice& L, const int* i)
{
  return P.lat_univ[L.univ_offset + L.n[0] * L.n[1] * i[2] + L.n[0] * i[1] + i[0]];
}

// Selects of the lowest coordinate level with static indices (the level
// arrays stay in registers).
template<int NL>
__device__ __forceinline__ int lowest_cell(const Geo<NL>& g)
{
  int c = g.cell[0];
#pragma unroll
  for (int l = 1; l < NL; ++l)
    if (l == g.n_coord - 1)
      c = g.cell[l];
  return c;
}

template<int NL>
__device__ __forceinline__ int lowest_univ(const Geo<NL>& g)
{
  int u = g.univ[0];
#pragma unroll
  for (int l = 1; l < NL; ++l)
    if (l == g.n_coord - 1)
      u = g.univ[l];
  return u;
}

// Universe::find_cell (src/universe.cpp:40-60) at coordinate level lvl
template<int NL>
__device__ __forceinline__ bool universe_find_cell(const DevProblem& P, Geo<NL>& g, const int lvl)
{
  const DevUniverse U = P.univ[g.univ[lvl]];
  const V3 r = g.r(lvl);
  const V3 u = g.u();
  for (int k = U.cell_begin; k < U.cell_end; ++k) {
    const int i_cell = P.univ_cells[k];
    if (cell_contains(P, P.cell[i_cell], r, u, g.surface)) {
      g.cell[lvl] = i_cell;
      return true;
    }
  }
  return false;
}

// find_cell_inner (src/geometry.cpp:102-252), searching the whole universe at
// each level (the neighbor lists of the reference visit a subset of the same
// universe and, for universes whose cells partition space, find the same cell)
template<int NL>
__device__ __forceinline__ bool find_cell_inner(const DevProblem& P, Geo<NL>& g)
{
  // Iteration `lvl` handles coordinate level lvl = n_coord - 1 (levels above
  // the starting n_coord are skipped), so every array index is static.
#pragma unroll
  for (int lvl = 0; lvl < NL; ++lvl) {
    if (lvl < g.n_coord - 1)
      continue;
    if (!universe_find_cell(P, g, lvl))
      return false;
    const DevCell c = P.cell[g.cell[lvl]];
    if (c.fill_type == FILL_MATERIAL) {
      g.material_last = g.material;
      g.material = c.material;
      return true;
    }
    if (lvl + 1 >= NL)
      return false;
    const int nl = lvl + 1;
    const V3 r = g.r(lvl);
    const V3 u = g.u();
    if (c.fill_type == FILL_UNIVERSE) {
      g.univ[nl] = c.fill;
      g.set_r(nl, r);
    } else {
      const DevLattice L = P.lat[c.fill];
      int idx[3];
      lattice_indices(L, r, u, idx);
      g.lix[nl] = idx[0];
      g.liy[nl] = idx[1];
      g.liz[nl] = idx[2];
      g.set_r(nl, lattice_local(L, r, idx));
      g.lat[nl] = c.fill;
      if (lattice_valid(L, idx)) {
        g.univ[nl] = lattice_universe(P, L, idx);
      } else if (L.outer != C_NONE) {
        g.univ[nl] = L.outer;
      } else {
        g.lost = 1;
        return false;
      }
    }
    ++g.n_coord;
  }
  return false;
}

// exhaustive_find_cell (src/geometry.cpp:284-297)
template<int NL>
__device__ __forceinline__ bool exhaustive_find_cell(const DevProblem& P, Geo<NL>& g)
{
  if (lowest_univ(g) == C_NONE) {
    g.univ[0] = P.root_universe;
    g.n_coord = 1;
  }
#pragma unroll
  for (int i = 1; i < NL; ++i)
    if (i >= g.n_coord)
      g.reset(i);
  return find_cell_inner(P, g);
}

// neighbor_list_find_cell (src/geometry.cpp:256-282)
template<int NL>
__device__ __forceinline__ bool neighbor_list_find_cell(const DevProblem& P, Geo<NL>& g)
{
#pragma unroll
  for (int i = 1; i < NL; ++i)
    if (i >= g.n_coord)
      g.reset(i);
  return find_cell_inner(P, g);
}

struct Boundary {
  double distance;
  int surface;
  int coord_level;
  int trans[3];
};

// distance_to_boundary (src/geometry.cpp:361-453)
template<int NL>
__device__ __forceinline__ Boundary distance_to_boundary(const DevProblem& P, Geo<NL>& g)
{
  Boundary info;
  info.distance = __builtin_inf();
  info.surface = SURFACE_NONE;
  info.coord_level = 0;
  info.trans[0] = info.trans[1] = info.trans[2] = 0;
  double d_lat = __builtin_inf();
  int level_lat_trans[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < NL; i++) {
    if (i >= g.n_coord)
      break;
    const V3 r = g.r(i);
    const V3 u = g.u();
    int level_surf_cross;
    double d_surf = cell_distance(P, P.cell[g.cell[i]], r, u, g.surface, &level_surf_cross);
    if (g.lat[i] != C_NONE) {
      d_lat = lattice_distance(P.lat[g.lat[i]], r, u, level_lat_trans);
      if (d_lat < 0)
        g.lost = 1;
    }
    double& d = info.distance;
    if (d_surf < d_lat - FP_COINCIDENT) {
      if (d == __builtin_inf() || (d - d_surf) / d >= FP_REL_PRECISION) {
        d = d_surf;
        info.surface = level_surf_cross;
        info.trans[0] = info.trans[1] = info.trans[2] = 0;
        info.coord_level = i + 1;
      }
    } else {
      if (d == __builtin_inf() || (d - d_lat) / d >= FP_REL_PRECISION) {
        d = d_lat;
        info.surface = SURFACE_NONE;
        info.trans[0] = level_lat_trans[0];
        info.trans[1] = level_lat_trans[1];
        info.trans[2] = level_lat_trans[2];
        info.coord_level = i + 1;
      }
    }
  }
  return info;
}

// cross_lattice (src/geometry.cpp:301-357)
template<int NL>
__device__ __forceinline__ void cross_lattice(const DevProblem& P, Geo<NL>& g, const int* trans)
{
  bool valid = false;
  bool done = false;
#pragma unroll
  for (int lvl = 1; lvl < NL; ++lvl) {
    if (lvl != g.n_coord - 1)
      continue;
    if (g.lat[lvl] == C_NONE)
      break;
    done = true;
    const DevLattice L = P.lat[g.lat[lvl]];
    int idx[3] = {g.lix[lvl] + trans[0], g.liy[lvl] + trans[1], g.liz[lvl] + trans[2]};
    g.lix[lvl] = idx[0];
    g.liy[lvl] = idx[1];
    g.liz[lvl] = idx[2];
    g.set_r(lvl, lattice_local(L, g.r(lvl - 1), idx));
    valid = lattice_valid(L, idx);
    if (valid)
      g.univ[lvl] = lattice_universe(P, L, idx);
  }
  if (!done) {
    // a lattice translation selected at a level without a lattice; the
    // reference would index its lattice array with -1 here
    g.lost = 1;
    return;
  }
  if (!valid) {
    g.n_coord = 1;
    if (!exhaustive_find_cell(P, g))
      g.lost = 1;
  } else {
    if (!exhaustive_find_cell(P, g)) {
      g.n_coord = 1;
      if (!exhaustive_find_cell(P, g))
        g.lost = 1;
    }
  }
}

// Position::reflect (include/openmc/position.h:207-213)
__device__ __forceinline__ V3 reflect(V3 u, V3 n)
{
  const double projection = v_dot(n, u);
  const double magnitude = v_dot(n, n);
  double s = 2.0 * projection / magnitude;
  n = v_scale(n, s);
  return v_sub(u, n);
}

//==============================================================================
// Cross sections (src/nuclide.cpp:616-839, src/material.cpp:829-901)
//==============================================================================

// Bucket bounds of the log-grid hash, clamped to the table.
__device__ __forceinline__ int log_union_index(const DevProblem& P, double E)
{
  int i = static_cast<int>(detm::log(E / P.energy_min) / P.log_spacing);
  return i;
}

// std::lower_bound over E[first, first + count): first index with !(E[i] < v)
__device__ __forceinline__ int lower_bound_idx(const double* E, int first, int count, double v, int& probes)
{
  while (count > 0) {
    int step = count >> 1;
    int it = first + step;
    ++probes;
    if (E[it] < v) {
      first = it + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return first;
}

// Learned bounded search.  The nuclide's piecewise-linear model maps
// x = log(E/E_min) to a predicted interval index j; the recorded bound err of
// the segment (the largest |pred - lower bound| a query of the segment can
// see, evaluated with this same FP32 arithmetic) guarantees that the global
// lower_bound of E lies in [j - err - 1, j + err + 3), which is searched
// exactly.  The predicted interval is tried first: when E[j] < E < E[j + 1]
// holds strictly, the lower bound is j + 1 and no search runs.  Returns the
// global lower_bound index (first i with E[i] >= E).
__device__ __forceinline__ int learned_lower_bound(const DevProblem& P, int i_nuc, const double* Eg, int n,
  double E, double lx, int& probes)
{
  const DevLearned& Lm = P.learned;
  double uu = (lx - Lm.u_lo[i_nuc]) * Lm.inv_w[i_nuc];
  const int ns = Lm.seg_n[i_nuc];
  int s = uu < 0.0 ? 0 : static_cast<int>(uu);
  s = s >= ns ? ns - 1 : s;
  const int64_t k = static_cast<int64_t>(Lm.seg_off[i_nuc]) + s;
  const float2 sg = reinterpret_cast<const float2*>(Lm.seg)[k];
  const int err = Lm.err[k];
  float fr = static_cast<float>(uu - static_cast<double>(s));
  float pred = sg.x * fr + sg.y;
  int j = static_cast<int>(pred);
  const int js = j < 0 ? 0 : (j > n - 2 ? n - 2 : j);
  if (Eg[js] < E && E < Eg[js + 1])
    return js + 1;
  int lo = j - err - 1;
  int hi = j + err + 3;
  lo = lo < 0 ? 0 : (lo > n ? n : lo);
  hi = hi > n ? n : (hi < lo ? lo : hi);
  return lower_bound_idx(Eg, lo, hi - lo, E, probes);
}

struct Micro {
  int i_grid;
  double f;
  double total;
  double absorption;
};

// One grid point's {E, total, absorption} record, and the bracket of
// interval g (points g and g + 1, global indices): 48 contiguous bytes, three
// 16-byte loads touching one or two lines, where separate energy and cross
// section arrays touch two or more.
struct PointRec {
  double e, t, a;
};
struct Bracket {
  double e0, t0, a0, e1, t1, a1;
};
__device__ __forceinline__ PointRec load_point(const DevProblem& P, int64_t g)
{
  const double* b = P.B + 3 * g;
  return PointRec {b[0], b[1], b[2]};
}
__device__ __forceinline__ Bracket load_bracket(const DevProblem& P, int64_t g)
{
  const double* b = P.B + 3 * g;
  return Bracket {b[0], b[1], b[2], b[3], b[4], b[5]};
}

// Nuclide::calculate_xs, NEAREST temperature with one temperature
__device__ __forceinline__ Micro nuclide_xs(const DevProblem& P, int i_nuc, double E, int i_log, double logE,
  bool learned, int& probes)
{
  const DevNuclide nd = P.nuc[i_nuc];
  const double* Eg = P.E + nd.off;
  const int n = nd.n;
  int i_grid;
  if (E < Eg[0]) {
    i_grid = 0;
  } else if (E > Eg[n - 1]) {
    i_grid = n - 2;
  } else {
    const int* gi = P.grid_index + static_cast<int64_t>(i_nuc) * (P.n_log_bins + 1);
    if (!learned) {
      int i_low = gi[i_log];
      int i_high = gi[i_log + 1] + 1;
      // lower_bound_index (include/openmc/search.h:13-20)
      if (Eg[i_low] == E) {
        i_grid = i_low;
      } else {
        i_grid = lower_bound_idx(Eg, i_low, i_high - i_low, E, probes) - 1;
      }
    } else {
      int lb = learned_lower_bound(P, i_nuc, Eg, n, E, logE, probes);
      // The bucket search equals lower_bound - 1 except when E equals the
      // first point of its bucket (search.h:17-18), which needs E to be a grid
      // energy; only then is the bucket start read.
      if (lb < n && Eg[lb] == E) {
        int il = gi[i_log];
        i_grid = (Eg[il] == E) ? il : lb - 1;
      } else {
        i_grid = lb - 1;
      }
    }
  }
  // keep every read inside the grid (a no-op whenever the search is correct,
  // which the parity tests check index for index)
  i_grid = i_grid < 0 ? 0 : (i_grid > n - 2 ? n - 2 : i_grid);
  // check for rare case where two energy points are the same
  Bracket b = load_bracket(P, nd.off + i_grid);
  if (i_grid + 2 < n && b.e0 == b.e1) {
    ++i_grid;
    b = load_bracket(P, nd.off + i_grid);
  }
  double f = (E - b.e0) / (b.e1 - b.e0);
  Micro m;
  m.i_grid = i_grid;
  m.f = f;
  m.total = (1.0 - f) * b.t0 + f * b.t1;
  m.absorption = (1.0 - f) * b.a0 + f * b.a1;
  return m;
}

// K lookups of one energy in K nuclides, interleaved: every phase issues the
// K independent loads before it uses any of them, and the K bounded searches
// run in lockstep, so one lane keeps up to K cache misses in flight instead of
// one.  With the learned index the brackets of the predicted interval are read
// first: when they enclose E strictly (E[j] < E < E[j+1]) that interval is
// the answer and no search runs, which costs one model read and the three
// bracket reads, the fewest loads any exact method needs.  (A final step
// that reads several points at once instead of probing was measured slower:
// the kernel is bound by vector-memory instruction issue.)
// Each result equals nuclide_xs for the same nuclide:
//  - the bucket search's special case (E equal to the first point of its
//    bucket gives that point, include/openmc/search.h:17-18) is applied after
//    the search: lower_bound over [i_low, i_high) returns i_low exactly when
//    E[i_low] >= E, so E[i_low] == E implies the lower bound is i_low, and
//    that point only needs reading when the lower bound is i_low;
//  - the duplicate-point bump reads the next point only when it applies.
// Entries k >= kk are not computed (their outputs are left unchanged).
// mn (the material's nuclide records) and kk must be the same in every
// active lane of the wavefront: the nuclide indices, grid offsets and sizes
// are then held in scalar registers (readfirstlane of values every active
// lane has loaded alike).
__device__ __forceinline__ int wave_uniform(int x)
{
  return __builtin_amdgcn_readfirstlane(x);
}

__device__ __forceinline__ int64_t wave_uniform(int64_t x)
{
  const uint64_t u = static_cast<uint64_t>(x);
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(u & 0xffffffffu)));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32)));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

__device__ __forceinline__ double wave_uniform(double x)
{
  return __longlong_as_double(wave_uniform(static_cast<int64_t>(__double_as_longlong(x))));
}

// A read-only value at an address that is the same in every active lane of
// the wavefront (the records of the wavefront's material): read through the
// scalar data cache into scalar registers (a constant-address-space load from
// the first active lane's address), so the vector-memory pipeline carries
// only the per-lane gathers.  Only for data that no kernel writes (the
// nuclear data and material records, uploaded at set-up).
template<class T>
__device__ __forceinline__ T uniform_load(const T* p)
{
  const auto a = wave_uniform(static_cast<int64_t>(reinterpret_cast<uintptr_t>(p)));
  return *reinterpret_cast<const __attribute__((address_space(4))) T*>(static_cast<uintptr_t>(a));
}

// The value at p: each lane's own where the lanes need not agree (UNI false:
// one particle per lane, any material), the wavefront's scalar copy where
// they do.
template<bool UNI, class T>
__device__ __forceinline__ T uld(const T* p)
{
  if constexpr (UNI)
    return uniform_load(p);
  else
    return *p;
}

// UNI (default) requires the wavefront-uniform material above; with UNI
// false each lane reads its own material's records, for one particle per lane.
template<int K, bool LEARNED, bool UNI = true>
__device__ __forceinline__ void nuclide_xs_group(const DevProblem& P, const DevMatNuc* mn, int kk, double E,
  int i_log, double logE, int& probes, Micro* out)
{
  const double* Eg[K];
  int64_t off[K];
  int n[K], fst[K], cnt[K], lo0[K], ig[K], nuc[K];
  double e_first[K], e_last[K];
  bool search[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const DevMatNuc* r = mn + (k < kk ? k : 0);
    nuc[k] = uld<UNI>(&r->nuc);
    off[k] = uld<UNI>(&r->off);
    n[k] = uld<UNI>(&r->n);
    e_first[k] = uld<UNI>(&r->e_first);
    e_last[k] = uld<UNI>(&r->e_last);
    Eg[k] = P.E + off[k];
  }
  // final bracket values; with the learned index they are first read
  // speculatively at the predicted interval
  double e0[K], e1[K];
  double2 x0[K], x1[K];
  bool hit[K];
  // bracket by the log-grid bucket (or the learned window); all K nuclides'
  // loads are issued before any is used
  int lo[K], hi[K];
  if (!LEARNED) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int* gi = P.grid_index + static_cast<int64_t>(nuc[k]) * (P.n_log_bins + 1);
      lo[k] = gi[i_log];
      hi[k] = gi[i_log + 1] + 1;
      hit[k] = false;
    }
  } else {
    const DevLearned& Lm = P.learned;
    LSeg sg[K];
    double uu[K];
    int seg[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const DevMatNuc* r = mn + (k < kk ? k : 0);
      uu[k] = (logE - uld<UNI>(&r->u_lo)) * uld<UNI>(&r->inv_w);
      const int ns = uld<UNI>(&r->seg_n);
      int sgi = uu[k] < 0.0 ? 0 : static_cast<int>(uu[k]);
      seg[k] = sgi >= ns ? ns - 1 : sgi;
      sg[k] = Lm.rec[uld<UNI>(&r->seg_off) + seg[k]];
    }
    int js[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float fr = static_cast<float>(uu[k] - static_cast<double>(seg[k]));
      float pred = sg[k].a * fr + sg[k].b;
      int j = static_cast<int>(pred);
      int l = j - sg[k].err - 1;
      int h = j + sg[k].err + 3;
      l = l < 0 ? 0 : (l > n[k] ? n[k] : l);
      h = h > n[k] ? n[k] : (h < l ? l : h);
      lo[k] = l;
      hi[k] = h;
      js[k] = j < 0 ? 0 : (j > n[k] - 2 ? n[k] - 2 : j);
    }
    // speculative brackets at the predicted interval js
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Bracket b = load_bracket(P, off[k] + js[k]);
      e0[k] = b.e0;
      e1[k] = b.e1;
      x0[k] = double2 {b.t0, b.a0};
      x1[k] = double2 {b.t1, b.a1};
    }
    // E[js] < E < E[js + 1] makes the lower bound js + 1 with E not a grid
    // point (so neither the bucket-start case nor the duplicate bump applies):
    // the interval is js
#pragma unroll
    for (int k = 0; k < K; ++k) {
      hit[k] = e0[k] < E && E < e1[k];
      ig[k] = js[k];
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool below = E < e_first[k];
    const bool above = !below && E > e_last[k];
    search[k] = k < kk && !below && !above && !hit[k];
    if (!hit[k])
      ig[k] = below ? 0 : n[k] - 2;
    lo0[k] = lo[k];
    fst[k] = lo[k];
    cnt[k] = search[k] ? hi[k] - lo[k] : 0;
  }
  // lower_bound of E in [fst, fst + cnt), the K searches in lockstep
  while (true) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < K; ++k)
      any = any || cnt[k] > 0;
    if (!any)
      break;
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
      v[k] = cnt[k] > 0 ? Eg[k][fst[k] + (cnt[k] >> 1)] : 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (cnt[k] > 0) {
        ++probes;
        const int step = cnt[k] >> 1;
        if (v[k] < E) {
          fst[k] += step + 1;
          cnt[k] -= step + 1;
        } else {
          cnt[k] = step;
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!search[k])
      continue;
    const int lb = fst[k];
    if (!LEARNED) {
      ig[k] = lb - 1;
      if (lb == lo0[k] && Eg[k][lb] == E)
        ig[k] = lb;
    } else {
      ig[k] = lb - 1;
      if (lb < n[k] && Eg[k][lb] == E) {
        const int il = P.grid_index[static_cast<int64_t>(nuc[k]) * (P.n_log_bins + 1) + i_log];
        if (Eg[k][il] == E)
          ig[k] = il;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (hit[k])
      continue;
    int i = ig[k];
    i = i < 0 ? 0 : (i > n[k] - 2 ? n[k] - 2 : i);
    ig[k] = i;
    const Bracket b = load_bracket(P, off[k] + i);
    e0[k] = b.e0;
    e1[k] = b.e1;
    x0[k] = double2 {b.t0, b.a0};
    x1[k] = double2 {b.t1, b.a1};
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= kk)
      continue;
    int i = ig[k];
    if (!hit[k] && i + 2 < n[k] && e0[k] == e1[k]) {
      ++i;
      const PointRec q = load_point(P, off[k] + i + 1);
      e0[k] = e1[k];
      e1[k] = q.e;
      x0[k] = x1[k];
      x1[k] = double2 {q.t, q.a};
    }
    const double f = (E - e0[k]) / (e1[k] - e0[k]);
    out[k].i_grid = i;
    out[k].f = f;
    out[k].total = (1.0 - f) * x0[k].x + f * x1[k].x;
    out[k].absorption = (1.0 - f) * x0[k].y + f * x1[k].y;
  }
}

// Minimum and maximum of an int over the lanes of a wavefront (every lane
// must call them); the result is wavefront-uniform.
__device__ __forceinline__ int wave_min_int(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_xor(x, o);
    x = y < x ? y : x;
  }
  return wave_uniform(x);
}

__device__ __forceinline__ int wave_max_int(int x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int y = __shfl_xor(x, o);
    x = y > x ? y : x;
  }
  return wave_uniform(x);
}

constexpr int WAVE_WINDOW = 64; // grid points one wavefront load covers (one per lane)

// Nuclide::calculate_xs for K nuclides of the wavefront's material, the
// searches done by the wavefront together.  Every lane of the wavefront must
// call it: lanes with act == false take part in the loads and cross-lane
// reads only.
//
// The lanes of a sorted calculate-xs queue hold nearby energies of one
// material, so for each nuclide their log-grid buckets [lo, hi) lie in a
// short window [lo_w, hi_w) of the grid (lo_w and hi_w are the bucket bounds
// at the smallest and largest log-grid index of the wavefront: the bucket
// table is non-decreasing).  When the window has at most 64 points, lane j
// loads E[lo_w + j] (one coalesced load for the wavefront) and each lane runs
// lower_bound over exactly its own bucket [lo, hi) on those registers,
// reading other lanes' values with cross-lane moves.  The index is therefore
// the one the search in memory gives (include/openmc/search.h:13-20 with the
// bucket-start case, src/nuclide.cpp:732-742), and the bracket energies
// E[i], E[i+1] come from the same registers.  A nuclide whose window is wider
// takes the per-lane search in memory.
template<int K>
__device__ __forceinline__ void nuclide_xs_wave(const DevProblem& P, const DevMatNuc* mn, int kk, double E,
  int i_log, int il_lo, int il_hi, bool act, int& probes, Micro* out)
{
  const int lane = static_cast<int>(__lane_id());
  const double* Eg[K];
  int64_t off[K];
  int n[K], lo[K], hi[K], lo_w[K], ig[K], first[K], cnt[K];
  bool coop[K], in[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const DevMatNuc* r = mn + (k < kk ? k : 0);
    const int nuc = uniform_load(&r->nuc);
    off[k] = uniform_load(&r->off);
    n[k] = uniform_load(&r->n);
    const double e_first = uniform_load(&r->e_first);
    const double e_last = uniform_load(&r->e_last);
    Eg[k] = P.E + off[k];
    const int* gi = P.grid_index + static_cast<int64_t>(nuc) * (P.n_log_bins + 1);
    lo_w[k] = wave_uniform(gi[il_lo]);
    const int hi_w = wave_uniform(gi[il_hi + 1]) + 1;
    coop[k] = hi_w - lo_w[k] <= WAVE_WINDOW;
    lo[k] = gi[i_log];
    hi[k] = gi[i_log + 1] + 1;
    const bool below = E < e_first;
    const bool above = !below && E > e_last;
    in[k] = act && k < kk && !below && !above;
    ig[k] = below ? 0 : n[k] - 2;
  }
  // the window loads, one coalesced load per nuclide
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (coop[k]) {
      int j = lo_w[k] + lane;
      j = j > n[k] - 1 ? n[k] - 1 : j;
      v[k] = Eg[k][j];
    } else {
      v[k] = 0.0;
    }
  }
  // lower_bound over [lo, hi): on the window registers (coop) or in memory,
  // all K searches in lockstep; the loop runs while any lane of the
  // wavefront still searches, so the cross-lane reads see every lane
#pragma unroll
  for (int k = 0; k < K; ++k) {
    first[k] = in[k] ? (coop[k] ? lo[k] - lo_w[k] : lo[k]) : 0;
    cnt[k] = in[k] ? hi[k] - lo[k] : 0;
    if (in[k])
      ++probes;
  }
  while (true) {
    bool mine = false;
#pragma unroll
    for (int k = 0; k < K; ++k)
      mine = mine || cnt[k] > 0;
    if (__ballot(mine) == 0ull)
      break;
    double vm[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int mid = first[k] + (cnt[k] >> 1);
      if (coop[k])
        vm[k] = __shfl(v[k], mid & (WAVE_WINDOW - 1));
      else
        vm[k] = cnt[k] > 0 ? Eg[k][mid] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (cnt[k] > 0) {
        if (!coop[k])
          ++probes;
        const int step = cnt[k] >> 1;
        if (vm[k] < E) {
          first[k] += step + 1;
          cnt[k] -= step + 1;
        } else {
          cnt[k] = step;
        }
      }
    }
  }
  // index and bracket energies
  double e0[K], e1[K];
  bool have[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int lb = coop[k] ? lo_w[k] + first[k] : first[k];
    // E[lb] for the bucket-start case
    double e_lb;
    if (coop[k])
      e_lb = __shfl(v[k], first[k] & (WAVE_WINDOW - 1));
    else
      e_lb = in[k] && lb == lo[k] ? Eg[k][lb] : 0.0;
    if (in[k])
      ig[k] = lb == lo[k] && e_lb == E ? lb : lb - 1;
    int i = ig[k];
    i = i < 0 ? 0 : (i > n[k] - 2 ? n[k] - 2 : i);
    ig[k] = i;
    const int r0 = i - lo_w[k];
    have[k] = coop[k] && r0 >= 0 && r0 + 1 < WAVE_WINDOW;
    if (coop[k]) {
      e0[k] = __shfl(v[k], r0 & (WAVE_WINDOW - 1));
      e1[k] = __shfl(v[k], (r0 + 1) & (WAVE_WINDOW - 1));
    }
  }
  double2 x0[K], x1[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const Bracket b = load_bracket(P, off[k] + ig[k]);
    if (!have[k]) {
      e0[k] = b.e0;
      e1[k] = b.e1;
    }
    x0[k] = double2 {b.t0, b.a0};
    x1[k] = double2 {b.t1, b.a1};
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= kk)
      continue;
    int i = ig[k];
    // duplicate grid point: the upper interval (src/nuclide.cpp:740-742)
    if (i + 2 < n[k] && e0[k] == e1[k]) {
      ++i;
      const PointRec q = load_point(P, off[k] + i + 1);
      e0[k] = e1[k];
      e1[k] = q.e;
      x0[k] = x1[k];
      x1[k] = double2 {q.t, q.a};
    }
    const double f = (E - e0[k]) / (e1[k] - e0[k]);
    out[k].i_grid = i;
    out[k].f = f;
    out[k].total = (1.0 - f) * x0[k].x + f * x1[k].x;
    out[k].absorption = (1.0 - f) * x0[k].y + f * x1[k].y;
  }
}

//==============================================================================
// Tally: RegularMesh ray-march (src/mesh.cpp:1000-1146, 1328-1369)
//==============================================================================

struct MeshDist {
  int next_index;
  bool max_surface;
  double distance;
};

__device__ __forceinline__ MeshDist mesh_distance(const DevTally& T, const int* ijk, int i, V3 r0, V3 u)
{
  MeshDist d;
  d.next_index = ijk[i];
  d.max_surface = true;
  d.distance = INFTY;
  double ui = v_get(u, i);
  if (fabs(ui) < FP_PRECISION)
    return d;
  d.max_surface = (ui > 0);
  if (d.max_surface && (ijk[i] <= T.shape[i])) {
    d.next_index++;
    d.distance = ((T.ll[i] + ijk[i] * T.width[i]) - v_get(r0, i)) / ui;
  } else if (!d.max_surface && (ijk[i] >= 1)) {
    d.next_index--;
    d.distance = ((T.ll[i] + (ijk[i] - 1) * T.width[i]) - v_get(r0, i)) / ui;
  }
  return d;
}

// RegularMesh::get_indices / get_bin_from_indices (src/mesh.cpp:874-900,
// 1328-1331) for any regular mesh description (tally or weight windows)
template<class M>
__device__ __forceinline__ void mesh_indices(const M& T, V3 r, int* ijk, bool& in_mesh)
{
  in_mesh = true;
  for (int i = 0; i < T.n_dim; ++i) {
    ijk[i] = static_cast<int>(ceil((v_get(r, i) - T.ll[i]) / T.width[i]));
    if (ijk[i] < 1 || ijk[i] > T.shape[i])
      in_mesh = false;
  }
}

template<class M>
__device__ __forceinline__ int64_t mesh_bin(const M& T, const int* ijk)
{
  switch (T.n_dim) {
  case 1: return ijk[0] - 1;
  case 2: return static_cast<int64_t>(ijk[1] - 1) * T.shape[0] + ijk[0] - 1;
  default:
    return (static_cast<int64_t>(ijk[2] - 1) * T.shape[1] + (ijk[1] - 1)) * T.shape[0] + ijk[0] - 1;
  }
}

// Index of the XCD (accelerator complex die) the wavefront runs on.
__device__ __forceinline__ int xcd_id()
{
  unsigned x;
  asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
  return static_cast<int>(x);
}

// The VALUE array has n_copies copies of the tally's bins; a wavefront adds
// into the copy of its XCD, so the lines a copy's atomics touch live in one
// XCD's L2 (k_accumulate sums the copies in copy order at the end of a batch).
__device__ __forceinline__ void tally_add(const DevTally& T, int64_t bin, double score)
{
  double* v = T.value;
  if (T.n_copies > 1)
    v += static_cast<int64_t>(xcd_id() & (T.n_copies - 1)) * T.n_bins;
  unsafeAtomicAdd(&v[bin], score);
}

// score_tracklength_tally with [mesh], [mesh, energy], [energy] or no filters
// (src/tallies/tally_scoring.cpp:2416-2541); returns segments scored
__device__ __forceinline__ int score_tracklength(const DevTally& T, V3 r0, V3 r1, V3 u, double E, double wgt,
  double distance)
{
  double flux = wgt * distance;
  int n_e = 1;
  int e_bin = 0;
  if (T.n_energy > 0) {
    const double* eb = T.ebins;
    int nb = T.n_energy + 1;
    if (!(E >= eb[0] && E <= eb[nb - 1]))
      return 0;
    if (eb[0] == E) {
      e_bin = 0;
    } else {
      int dummy = 0;
      e_bin = lower_bound_idx(eb, 0, nb, E, dummy) - 1;
    }
    n_e = T.n_energy;
  }
  if (!T.has_mesh) {
    // filters [] or [energy]: one bin per energy group, filter weight 1
    tally_add(T, e_bin, flux * 1.0);
    return 1;
  }
  int segs = 0;
  double total_distance = v_norm(v_sub(r1, r0));
  if (total_distance == 0.0)
    return 0;
  bool in_mesh;
  double traveled = 0.0;
  int ijk[3] = {0, 0, 0};
  mesh_indices(T, v_add(r0, v_scale(u, TINY_BIT)), ijk, in_mesh);
  if (total_distance < 2 * TINY_BIT) {
    if (in_mesh) {
      double score = flux * (1.0 * 1.0);
      tally_add(T, mesh_bin(T, ijk) * n_e + e_bin, score);
      segs++;
    }
    return segs;
  }
  MeshDist dist[3];
  for (int k = 0; k < 3; ++k) {
    dist[k].next_index = -1;
    dist[k].max_surface = true;
    dist[k].distance = INFTY;
  }
  for (int k = 0; k < T.n_dim; ++k)
    dist[k] = mesh_distance(T, ijk, k, r0, u);
  // a straight track crosses each mesh plane at most once; the cap only
  // guarantees termination on non-finite input
  const int max_steps = 2 * (T.shape[0] + T.shape[1] + T.shape[2]) + 16;
  for (int step = 0; step < max_steps; ++step) {
    if (in_mesh) {
      int k = 0;
      if (dist[1].distance < dist[k].distance)
        k = 1;
      if (dist[2].distance < dist[k].distance)
        k = 2;
      double w = (fmin(dist[k].distance, total_distance) - traveled) / total_distance;
      double weight = 1.0;
      weight *= w;
      weight *= 1.0;
      tally_add(T, mesh_bin(T, ijk) * n_e + e_bin, flux * weight);
      segs++;
      traveled = dist[k].distance;
      if (traveled >= total_distance)
        return segs;
      ijk[k] = dist[k].next_index;
      dist[k] = mesh_distance(T, ijk, k, r0, u);
      in_mesh = ((ijk[k] >= 1) && (ijk[k] <= T.shape[k]));
    } else {
      int k_max = -1;
      for (int k = 0; k < T.n_dim; ++k) {
        if ((ijk[k] < 1 || ijk[k] > T.shape[k]) && (dist[k].distance > traveled)) {
          traveled = dist[k].distance;
          k_max = k;
        }
      }
      if (k_max == -1)
        traveled += TINY_BIT;
      if (traveled >= total_distance)
        return segs;
      mesh_indices(T, v_add(r0, v_scale(u, traveled + TINY_BIT)), ijk, in_mesh);
      for (int k = 0; k < T.n_dim; ++k)
        dist[k] = mesh_distance(T, ijk, k, r0, u);
    }
  }
  return segs;
}

__device__ __forceinline__ void hmix(uint64_t& h, uint64_t v)
{
  h = (h ^ v) * 0x100000001b3ULL;
}

__device__ __forceinline__ uint64_t dbits(double x)
{
  return static_cast<uint64_t>(__double_as_longlong(x));
}

//==============================================================================
// Weight windows (src/weight_windows.cpp:56-146, 375-411;
// src/physics_common.cpp:12-19; src/particle.cpp:95-112)
//==============================================================================

struct WeightWindow {
  double lower;
  double upper;
  double max_lb_ratio;
  double survival;
  double cutoff;
  int max_split;
  __device__ __forceinline__ bool valid() const { return lower >= 0.0; }
  __device__ __forceinline__ void scale(double f)
  {
    lower *= f;
    upper *= f;
    survival *= f;
  }
};

// WeightWindows::get_weight_window of one set for a neutron at r with energy E
// (src/weight_windows.cpp:375-411); invalid (lower -1) outside its mesh or
// energy range
__device__ __forceinline__ WeightWindow set_weight_window(const DevWWSet& W, V3 r, double E)
{
  WeightWindow ww;
  ww.lower = -1.0;
  ww.upper = 1.0;
  ww.max_lb_ratio = 1.0;
  ww.survival = 0.5;
  ww.cutoff = 1.0e-38;
  ww.max_split = 10;
  int ijk[3] = {0, 0, 0};
  bool in_mesh;
  mesh_indices(W, r, ijk, in_mesh);
  if (!in_mesh)
    return ww;
  if (E < W.ebounds[0] || E > W.ebounds[W.n_e])
    return ww;
  // lower_bound_index over the energy bounds (include/openmc/search.h:13-20)
  int e_bin = 0;
  if (W.ebounds[0] != E) {
    int probes = 0;
    e_bin = lower_bound_idx(W.ebounds, 0, W.n_e + 1, E, probes) - 1;
  }
  const int64_t k = static_cast<int64_t>(e_bin) * W.n_mesh + mesh_bin(W, ijk);
  ww.lower = W.lower[k];
  ww.upper = W.upper[k];
  ww.survival = ww.lower * W.survival_ratio;
  ww.max_lb_ratio = W.max_lb_ratio;
  ww.max_split = W.max_split;
  ww.cutoff = W.weight_cutoff;
  return ww;
}

// The first set's valid window, else the last set's (invalid) answer
// (apply_weight_windows, src/weight_windows.cpp:70-75)
__device__ __forceinline__ WeightWindow get_weight_window(const DevWW& W, V3 r, double E)
{
  WeightWindow ww;
  ww.lower = -1.0;
  ww.upper = 1.0;
  ww.max_lb_ratio = 1.0;
  ww.survival = 0.5;
  ww.cutoff = 1.0e-38;
  ww.max_split = 10;
  for (int k = 0; k < W.n_sets; ++k) {
    ww = set_weight_window(W.sets[k], r, E);
    if (ww.valid())
      break;
  }
  return ww;
}

// Where a split particle's secondaries go: the next free sites of its LIFO
// bank (capacity bank_cap; overflow sets bit 0 of *status and the sites
// beyond the capacity are dropped, which the host reports as an error).
struct SplitSink {
  PSite* bank;
  int cap;
  int* status;
};

// apply_weight_windows for a live neutron with weight wgt at (r, u, E, time,
// surface); may split into the bank or play Russian roulette (tracking seed).
__device__ __forceinline__ void apply_weight_windows(const DevWW& W, PWW& st, double& wgt, V3 r, V3 u, double E,
  double time, int surface, uint64_t* seed, const SplitSink& sink)
{
  if (E <= 0 || wgt == 0.0)
    return;
  WeightWindow ww = get_weight_window(W, r, E);
  if (st.wgt_ww_born == -1.0)
    st.wgt_ww_born = ww.valid() ? (ww.lower + ww.upper) / 2 : 1.0;
  if (!ww.valid())
    return;
  ww.scale(st.wgt_born / st.wgt_ww_born);
  const double weight = wgt;
  if (wgt < ww.cutoff) {
    wgt = 0.0;
    return;
  }
  if (st.ww_factor == 0.0 && ww.max_lb_ratio > 1.0 && wgt > ww.lower * ww.max_lb_ratio)
    st.ww_factor = wgt / (ww.lower * ww.max_lb_ratio);
  if (st.ww_factor > 1.0)
    ww.scale(st.ww_factor);
  if (weight > ww.upper) {
    if (st.n_split >= W.max_history_splits)
      return;
    double n_split = ceil(weight / ww.upper);
    const double max_split = ww.max_split;
    n_split = fmin(n_split, max_split);
    st.n_split = static_cast<int>(st.n_split + n_split);
    const int i_split = static_cast<int>(round(n_split));
    const double w_each = weight / n_split;
    for (int l = 0; l < i_split - 1; l++) {
      if (st.n_bank >= sink.cap) {
        atomicOr(sink.status, 1);
        continue;
      }
      PSite s;
      s.rx = r.x;
      s.ry = r.y;
      s.rz = r.z;
      s.ux = u.x;
      s.uy = u.y;
      s.uz = u.z;
      s.E = E;
      s.time = time;
      s.wgt = w_each;
      s.surface = surface;
      s.pad = 0;
      sink.bank[st.n_bank++] = s;
    }
    wgt = weight / n_split;
  } else if (weight <= ww.lower) {
    // russian_roulette
    const double weight_survive = fmin(weight * ww.max_split, ww.survival);
    if (weight_survive * prn(seed) < wgt)
      wgt = weight_survive;
    else
      wgt = 0.0;
  }
}

} // namespace omc
