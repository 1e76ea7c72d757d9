// Learned energy-grid index (no counterpart in the reference; replaces the
// log-grid bucket table of Nuclide::init_grid, src/nuclide.cpp:502-533, and
// the bucket search of Nuclide::calculate_xs, src/nuclide.cpp:729-738, while
// keeping the grid index bit-identical).
//
// Model: for nuclide n the query coordinate is x = log(E / E_min) (the same
// value the log-grid hash computes), mapped to uu = (x - u_lo[n]) * inv_w[n]
// with segment s = clamp(floor(uu)) and local coordinate f = uu - s in [0,1).
// Segment s predicts the grid index as  pred = a_s * f + b_s  in FP32.
//
// Training (all on the GPU, one workgroup per (segment, nuclide)):
//   k_lforward   loss_s = sum_i (pred(f_i) - i)^2 over the segment's grid points
//   k_lbackward  d loss_s / d(a_s, b_s) = sum_i 2 r_i (f_i, 1)
//   k_ladam      Adam step on the mean gradient
//   k_lerror     pick the segment's line among the trained one (recentred)
//                and lines through its sample points, by the expected search
//                cost of the lookup's speculative bracket, and record err_s,
//                the largest |pred - lower bound| any query of the segment can
//                have, from the sub-interval ends (f = 0, each point, f = 1)
//                with the FP32 inference formula.
// The lookup then searches E[j - err - 1, j + err + 3) exactly (j = (int)pred),
// which contains the global lower_bound of any E that maps to segment s.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>

#include "device_physics.h"
#include "learned.h"
#include "openmc_amd/capi.h"

namespace omc {

namespace {

void lcheck(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template<class T>
T* lalloc(std::vector<void*>& list, size_t n)
{
  void* p = nullptr;
  lcheck(hipMalloc(&p, sizeof(T) * (n ? n : 1)), "hipMalloc");
  lcheck(hipMemset(p, 0, sizeof(T) * (n ? n : 1)), "hipMemset");
  list.push_back(p);
  return static_cast<T*>(p);
}

} // namespace

constexpr int LW = 64; // one wavefront per (segment, nuclide)

__device__ __forceinline__ int seg_of(double uu, int n_seg)
{
  int s = uu < 0.0 ? 0 : static_cast<int>(uu);
  return s >= n_seg ? n_seg - 1 : s;
}

// Every kernel below works on the nuclides list[0, gridDim.y) of one training
// round: blockIdx.y is the position y in that list, per-round arrays are
// indexed by y, per-nuclide ones (u_lo, inv_w, grid offsets) by list[y].

// uu_i for every grid point of the listed nuclides: grid (x over points, y)
__global__ void k_lprep(DevProblem P, const int* list, const double* u_lo, const double* inv_w, double* uu)
{
  const int n = list[blockIdx.y];
  const DevNuclide nd = P.nuc[n];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd.n; i += gridDim.x * blockDim.x) {
    double lx = detm::log(P.E[nd.off + i] / P.energy_min);
    uu[nd.off + i] = (lx - u_lo[n]) * inv_w[n];
  }
}

// seg_start[n][s] = first point whose segment >= s, s in [0, n_seg]
__global__ void k_lsegbounds(DevProblem P, const int* list, const double* uu, int n_seg, int* seg_start)
{
  const int n = list[blockIdx.y];
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= n_seg; s += gridDim.x * blockDim.x) {
    int lo = 0, cnt = nd.n;
    while (cnt > 0) {
      int step = cnt >> 1;
      int it = lo + step;
      if (seg_of(u[it], n_seg) < s) {
        lo = it + 1;
        cnt -= step + 1;
      } else {
        cnt = step;
      }
    }
    seg_start[static_cast<int64_t>(blockIdx.y) * (n_seg + 1) + s] = (s == n_seg) ? nd.n : lo;
  }
}

__device__ __forceinline__ double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ float predict(float2 ab, double uu, int s)
{
  float f = static_cast<float>(uu - static_cast<double>(s));
  return ab.x * f + ab.y;
}

// Initial parameters: the chord through the segment's first and last points.
__global__ void k_linit(DevProblem P, const int* list, const double* uu, const int* seg_start, int n_seg,
  float2* params)
{
  const int n = list[blockIdx.y];
  const int y = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg)
    return;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(y) * (n_seg + 1) + s;
  int a0 = seg_start[k], a1 = seg_start[k + 1];
  float2 ab;
  if (a1 - a0 >= 2) {
    double f0 = u[a0] - s, f1 = u[a1 - 1] - s;
    double slope = (f1 > f0) ? (static_cast<double>(a1 - 1 - a0) / (f1 - f0)) : 0.0;
    ab.x = static_cast<float>(slope);
    ab.y = static_cast<float>(a0 - slope * f0);
  } else {
    ab.x = 0.0f;
    ab.y = static_cast<float>(a0 > 0 ? a0 - 1 : 0);
  }
  params[static_cast<int64_t>(y) * n_seg + s] = ab;
}

__global__ void __launch_bounds__(LW) k_lforward(DevProblem P, const int* list, const double* uu,
  const int* seg_start, int n_seg, const float2* params, double* loss)
{
  const int n = list[blockIdx.y];
  const int y = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(y) * (n_seg + 1) + s;
  const int a0 = seg_start[k], a1 = seg_start[k + 1];
  const float2 ab = params[static_cast<int64_t>(y) * n_seg + s];
  double acc = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += LW) {
    double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    acc += r * r;
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0)
    loss[static_cast<int64_t>(y) * n_seg + s] = acc;
}

__global__ void __launch_bounds__(LW) k_lbackward(DevProblem P, const int* list, const double* uu,
  const int* seg_start, int n_seg, const float2* params, double2* grad)
{
  const int n = list[blockIdx.y];
  const int y = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(y) * (n_seg + 1) + s;
  const int a0 = seg_start[k], a1 = seg_start[k + 1];
  const float2 ab = params[static_cast<int64_t>(y) * n_seg + s];
  double ga = 0.0, gb = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += LW) {
    double f = static_cast<double>(static_cast<float>(u[i] - static_cast<double>(s)));
    double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    ga += 2.0 * r * f;
    gb += 2.0 * r;
  }
  ga = wave_sum(ga);
  gb = wave_sum(gb);
  if (threadIdx.x == 0)
    grad[static_cast<int64_t>(y) * n_seg + s] = double2 {ga, gb};
}

// Adam on the per-point mean gradient
__global__ void k_ladam(const int* seg_start, int n_seg, const double2* grad, float2* params, double2* m1,
  double2* m2, int step, double lr)
{
  const int y = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg)
    return;
  const int64_t k = static_cast<int64_t>(y) * (n_seg + 1) + s;
  const int cnt = seg_start[k + 1] - seg_start[k];
  if (cnt < 2)
    return;
  const int64_t j = static_cast<int64_t>(y) * n_seg + s;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-12;
  double2 g = grad[j];
  g.x /= cnt;
  g.y /= cnt;
  double2 a = m1[j], v = m2[j];
  a.x = b1 * a.x + (1 - b1) * g.x;
  a.y = b1 * a.y + (1 - b1) * g.y;
  v.x = b2 * v.x + (1 - b2) * g.x * g.x;
  v.y = b2 * v.y + (1 - b2) * g.y * g.y;
  m1[j] = a;
  m2[j] = v;
  double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
  float2 p = params[j];
  p.x = static_cast<float>(p.x - lr * (a.x / c1) / (sqrt(v.x / c2) + eps));
  p.y = static_cast<float>(p.y - lr * (a.y / c1) / (sqrt(v.y / c2) + eps));
  params[j] = p;
}

// The largest |pred - lower bound| a query of segment s (f in [0, 1)) can
// see.  Its points a0..a1-1 sit at f_a0 <= ... <= f_{a1-1}; the lower bound
// lb (first point >= E) is k on the sub-interval (f_{k-1}, f_k] (f_{a0-1}
// read as 0, f_{a1} as 1).  The FP32 prediction is monotone in f, so the
// extremes are at sub-interval ends:  pred(0) - a0,  pred(f_i) - i and
// pred(f_i) - (i + 1) for each point i, and pred(1) - a1.  Every lane of the
// wavefront returns the same value.
__device__ double seg_bound(float2 ab, const double* u, int a0, int a1, int s)
{
  const double us = static_cast<double>(s);
  double amax = 0.0;
  for (int i = a0 + static_cast<int>(threadIdx.x); i < a1; i += LW) {
    const double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    amax = fmax(amax, fmax(fabs(r), fabs(r - 1.0)));
  }
  const double r0 = static_cast<double>(predict(ab, us, s)) - static_cast<double>(a0);
  const double r1 = static_cast<double>(predict(ab, us + 1.0, s)) - static_cast<double>(a1);
  amax = fmax(amax, fmax(fabs(r0), fabs(r1)));
  for (int o = 32; o > 0; o >>= 1)
    amax = fmax(amax, __shfl_xor(amax, o));
  return amax;
}

// The fraction of segment s's queries (taken uniform in f over [0, 1)) for
// which the lookup's speculative bracket hits with line ab: a query between
// points i - 1 and i (lower bound i) hits when pred(f) lies in [i - 1, i).
// The prediction is taken in FP64 here; the value only ranks candidate lines,
// the error bound is computed exactly by seg_bound.  Every lane of the
// wavefront returns the same value.
__device__ double seg_hits(float2 ab, const double* u, int a0, int a1, int s)
{
  const double a = static_cast<double>(ab.x), b = static_cast<double>(ab.y);
  const double us = static_cast<double>(s);
  double h = 0.0;
  for (int i = a0 + static_cast<int>(threadIdx.x); i <= a1; i += LW) {
    double lo = (i == a0) ? 0.0 : fmin(fmax(u[i - 1] - us, 0.0), 1.0);
    double hi = (i == a1) ? 1.0 : fmin(fmax(u[i] - us, 0.0), 1.0);
    const double di = static_cast<double>(i);
    if (a > 0.0) {
      lo = fmax(lo, (di - 1.0 - b) / a);
      hi = fmin(hi, (di - b) / a);
    } else if (a < 0.0) {
      lo = fmax(lo, (di - b) / a);
      hi = fmin(hi, (di - 1.0 - b) / a);
    } else if (!(b >= di - 1.0 && b < di)) {
      continue;
    }
    h += fmax(hi - lo, 0.0);
  }
  return wave_sum(h);
}

constexpr int LSAMPLES = 16; // sample points per segment for candidate lines

// Choose the segment's line and record its error bound.  Candidates:
//  - the trained line, recentred on its points' residual range
//    pred(f_i) - i (so that a query between points i and i + 1 predicts
//    inside [i, i + 1) and the speculative bracket hits);
//  - the trained slope through each of LSAMPLES + 1 sample points, and the
//    chord through each pair of neighbouring sample points: where a cluster
//    of points that no segment can split (the points of one narrow
//    resonance) sits in the segment, these follow the grid on one side of it
//    exactly instead of missing it by half the cluster everywhere;
//  - the flat line through the middle of the segment's index range and the
//    end-to-end line, whose bounds are the smallest where the cluster is
//    large.
// Among the candidates whose bound is within cap, the one with the smallest
// expected search cost (miss fraction x the probes of its window) is kept;
// with none within cap, the one with the smallest bound (the build then
// doubles the nuclide's segments).
__global__ void __launch_bounds__(LW) k_lerror(DevProblem P, const int* list, const double* uu,
  const int* seg_start, int n_seg, int cap, float2* params, int* err)
{
  const int n = list[blockIdx.y];
  const int y = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(y) * (n_seg + 1) + s;
  const int a0 = seg_start[k], a1 = seg_start[k + 1];
  const int64_t j = static_cast<int64_t>(y) * n_seg + s;
  const double us = static_cast<double>(s);
  float2 trained = params[j];
  {
    double rmin = 1e300, rmax = -1e300;
    for (int i = a0 + threadIdx.x; i < a1; i += LW) {
      const double r = static_cast<double>(predict(trained, u[i], s)) - static_cast<double>(i);
      rmin = fmin(rmin, r);
      rmax = fmax(rmax, r);
    }
    for (int o = 32; o > 0; o >>= 1) {
      rmin = fmin(rmin, __shfl_xor(rmin, o));
      rmax = fmax(rmax, __shfl_xor(rmax, o));
    }
    if (rmax >= rmin)
      trained.y = static_cast<float>(static_cast<double>(trained.y) - 0.5 * (rmax + rmin));
  }
  const int n_pts = a1 - a0;
  const int n_cand = 3 + (n_pts >= 2 ? 2 * LSAMPLES + 1 : 0);
  float2 best_ab = trained;
  double best_bound = 1e300, best_cost = 1e300;
  bool best_ok = false;
  for (int c = 0; c < n_cand; ++c) {
    float2 ab;
    if (c == 0) {
      ab = trained;
    } else if (c == 1) {
      ab = float2 {0.0f, static_cast<float>(0.5 * (static_cast<double>(a0) + static_cast<double>(a1)))};
    } else if (c == 2) {
      ab = float2 {static_cast<float>(a1 - a0), static_cast<float>(a0)};
    } else if (c < 3 + LSAMPLES + 1) {
      // trained slope through sample point t
      const int t = c - 3;
      const int p = a0 + static_cast<int>((static_cast<int64_t>(n_pts - 1) * t) / LSAMPLES);
      const double fp = u[p] - us;
      ab = float2 {trained.x,
                   static_cast<float>(static_cast<double>(p) - static_cast<double>(trained.x) * fp)};
    } else {
      // chord through sample points t and t + 1
      const int t = c - (3 + LSAMPLES + 1);
      const int p = a0 + static_cast<int>((static_cast<int64_t>(n_pts - 1) * t) / LSAMPLES);
      const int q = a0 + static_cast<int>((static_cast<int64_t>(n_pts - 1) * (t + 1)) / LSAMPLES);
      const double fp = u[p] - us, fq = u[q] - us;
      if (!(q > p && fq > fp))
        continue;
      const double slope = static_cast<double>(q - p) / (fq - fp);
      ab = float2 {static_cast<float>(slope), static_cast<float>(static_cast<double>(p) - slope * fp)};
    }
    const double bound = seg_bound(ab, u, a0, a1, s);
    const bool ok = ceil(bound) <= static_cast<double>(cap);
    const double cost = ok ? (1.0 - seg_hits(ab, u, a0, a1, s)) * (log2(2.0 * ceil(bound) + 4.0) + 2.0) : 0.0;
    bool take;
    if (ok != best_ok)
      take = ok;
    else if (ok)
      take = cost < best_cost || (cost == best_cost && bound < best_bound);
    else
      take = bound < best_bound;
    if (take) {
      best_ok = ok;
      best_cost = cost;
      best_bound = bound;
      best_ab = ab;
    }
  }
  if (threadIdx.x == 0) {
    params[j] = best_ab;
    err[j] = static_cast<int>(ceil(best_bound));
  }
}

void free_learned(Learned& L)
{
  for (void* p : L.allocs)
    (void)hipFree(p);
  L.allocs.clear();
}

namespace {

// One training round: the nuclides `list`, each with n_seg segments.
struct Round {
  std::vector<void*> tmp;
  int* d_list = nullptr;
  double* d_ulo = nullptr;  // [n_nuc]
  double* d_invw = nullptr; // [n_nuc]
  double* uu = nullptr;     // [all grid points]
  int* seg_start = nullptr; // [list][n_seg + 1]
  double* loss = nullptr;   // [list][n_seg]
  double2* grad = nullptr;  // [list][n_seg]
  int n_list = 0;
  int n_seg = 0;
  std::vector<double> u_lo, inv_w; // [n_nuc], set for the listed nuclides
  ~Round()
  {
    for (void* p : tmp)
      (void)hipFree(p);
  }
};

// Segment scale of every listed nuclide (segment s covers
// u in [u_lo + s / inv_w, u_lo + (s + 1) / inv_w)), its grid points' uu and
// the first point of each segment.
void prepare(const DevProblem& P, const HostModel& m, const std::vector<int>& list, int n_seg, Round& r,
  hipStream_t st)
{
  const int nn = static_cast<int>(m.nuclides.size());
  r.n_list = static_cast<int>(list.size());
  r.n_seg = n_seg;
  r.u_lo.assign(nn, 0.0);
  r.inv_w.assign(nn, 0.0);
  int64_t total = 0;
  int maxn = 0;
  for (int n = 0; n < nn; ++n)
    total += static_cast<int64_t>(m.nuclides[n].energy.size());
  for (int n : list) {
    const auto& E = m.nuclides[n].energy;
    double x0 = detm::log(E.front() / m.energy_min);
    double x1 = detm::log(E.back() / m.energy_min);
    r.u_lo[n] = x0;
    r.inv_w[n] = static_cast<double>(n_seg) / (x1 - x0);
    maxn = std::max<int>(maxn, static_cast<int>(E.size()));
  }
  r.d_list = lalloc<int>(r.tmp, list.size());
  r.d_ulo = lalloc<double>(r.tmp, nn);
  r.d_invw = lalloc<double>(r.tmp, nn);
  lcheck(hipMemcpy(r.d_list, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(r.d_ulo, r.u_lo.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(r.d_invw, r.inv_w.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  r.uu = lalloc<double>(r.tmp, total);
  r.seg_start = lalloc<int>(r.tmp, list.size() * (n_seg + 1));
  r.loss = lalloc<double>(r.tmp, list.size() * n_seg);
  r.grad = lalloc<double2>(r.tmp, list.size() * n_seg);
  const int bx = std::min(1024, (maxn + 255) / 256);
  hipLaunchKernelGGL(k_lprep, dim3(bx, r.n_list), dim3(256), 0, st, P, r.d_list, r.d_ulo, r.d_invw, r.uu);
  const int sx = (n_seg + 1 + 255) / 256;
  hipLaunchKernelGGL(k_lsegbounds, dim3(sx, r.n_list), dim3(256), 0, st, P, r.d_list, r.uu, n_seg, r.seg_start);
  lcheck(hipGetLastError(), "learned prep");
}

// Per-listed-nuclide sum of the segment losses.
std::vector<double> round_loss(const DevProblem& P, Round& r, const float2* params, hipStream_t st)
{
  hipLaunchKernelGGL(k_lforward, dim3(r.n_seg, r.n_list), dim3(LW), 0, st, P, r.d_list, r.uu, r.seg_start, r.n_seg,
    params, r.loss);
  lcheck(hipGetLastError(), "learned forward");
  std::vector<double> seg(static_cast<size_t>(r.n_list) * r.n_seg);
  lcheck(hipMemcpyAsync(seg.data(), r.loss, sizeof(double) * seg.size(), hipMemcpyDeviceToHost, st), "hipMemcpy");
  lcheck(hipStreamSynchronize(st), "learned forward");
  std::vector<double> out(r.n_list, 0.0);
  for (int y = 0; y < r.n_list; ++y)
    for (int s = 0; s < r.n_seg; ++s)
      out[y] += seg[static_cast<size_t>(y) * r.n_seg + s];
  return out;
}

} // namespace

int build_learned_index(Learned& L, const HostModel& m, DevProblem& P, int n_seg, int epochs, hipStream_t st,
  int max_err_bound)
{
  if (n_seg < 1 || n_seg > LEARNED_MAX_SEG)
    throw std::runtime_error("n_seg out of range");
  if (max_err_bound < 0)
    throw std::runtime_error("the learned index's error bound must be >= 0");
  free_learned(L);
  const int nn = static_cast<int>(m.nuclides.size());
  std::vector<std::vector<float>> seg_n(nn);
  std::vector<std::vector<int>> err_n(nn);
  L.h_ulo.assign(nn, 0.0);
  L.h_invw.assign(nn, 0.0);
  L.h_segn.assign(nn, 0);
  L.loss_init = L.loss_trained = L.last_loss = 0.0;
  L.rounds = 0;
  std::vector<int> todo(nn);
  for (int n = 0; n < nn; ++n)
    todo[n] = n;
  // Train every nuclide with n_seg segments; a nuclide whose recorded bound
  // exceeds max_err_bound anywhere is trained again with twice as many.
  for (int ns = n_seg; !todo.empty(); ns *= 2) {
    if (ns > LEARNED_MAX_SEG)
      throw std::runtime_error("learned index: nuclide " + m.nuclides[todo[0]].name + " needs more than " +
                               std::to_string(LEARNED_MAX_SEG) + " segments to bound its index error by " +
                               std::to_string(max_err_bound));
    ++L.rounds;
    Round r;
    prepare(P, m, todo, ns, r, st);
    const size_t np = static_cast<size_t>(r.n_list) * ns;
    float2* params = lalloc<float2>(r.tmp, np);
    int* err = lalloc<int>(r.tmp, np);
    double2* m1 = lalloc<double2>(r.tmp, np);
    double2* m2 = lalloc<double2>(r.tmp, np);
    const int sx = (ns + 255) / 256;
    hipLaunchKernelGGL(k_linit, dim3(sx, r.n_list), dim3(256), 0, st, P, r.d_list, r.uu, r.seg_start, ns, params);
    const std::vector<double> l0 = round_loss(P, r, params, st);
    for (int e = 1; e <= epochs; ++e) {
      hipLaunchKernelGGL(k_lbackward, dim3(ns, r.n_list), dim3(LW), 0, st, P, r.d_list, r.uu, r.seg_start, ns,
        params, r.grad);
      double lr = 0.5 * std::pow(0.98, e);
      hipLaunchKernelGGL(k_ladam, dim3(sx, r.n_list), dim3(256), 0, st, r.seg_start, ns, r.grad, params, m1, m2, e,
        lr);
    }
    const std::vector<double> l1 = round_loss(P, r, params, st);
    hipLaunchKernelGGL(k_lerror, dim3(ns, r.n_list), dim3(LW), 0, st, P, r.d_list, r.uu, r.seg_start, ns,
      max_err_bound, params, err);
    lcheck(hipGetLastError(), "learned training");
    const std::vector<double> l2 = round_loss(P, r, params, st);
    std::vector<float> h_par(2 * np);
    std::vector<int> h_err(np);
    lcheck(hipMemcpy(h_par.data(), params, sizeof(float) * h_par.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    lcheck(hipMemcpy(h_err.data(), err, sizeof(int) * h_err.size(), hipMemcpyDeviceToHost), "hipMemcpy");
    std::vector<int> next;
    for (int y = 0; y < r.n_list; ++y) {
      const int n = todo[y];
      const int* e0 = h_err.data() + static_cast<size_t>(y) * ns;
      const int worst = *std::max_element(e0, e0 + ns);
      if (worst > max_err_bound) {
        next.push_back(n);
        continue;
      }
      seg_n[n].assign(h_par.begin() + 2 * static_cast<size_t>(y) * ns, h_par.begin() + 2 * static_cast<size_t>(y + 1) * ns);
      err_n[n].assign(e0, e0 + ns);
      L.h_ulo[n] = r.u_lo[n];
      L.h_invw[n] = r.inv_w[n];
      L.h_segn[n] = ns;
      L.loss_init += l0[y];
      L.loss_trained += l1[y];
      L.last_loss += l2[y];
    }
    todo.swap(next);
  }
  // concatenate the nuclides' segments
  L.h_segoff.assign(nn, 0);
  int64_t total = 0;
  for (int n = 0; n < nn; ++n) {
    if (total > INT32_MAX - L.h_segn[n])
      throw std::runtime_error("learned index: too many segments in total");
    L.h_segoff[n] = static_cast<int>(total);
    total += L.h_segn[n];
  }
  L.h_seg.assign(2 * static_cast<size_t>(total), 0.0f);
  L.h_err.assign(static_cast<size_t>(total), 0);
  std::vector<LSeg> rec(static_cast<size_t>(total));
  for (int n = 0; n < nn; ++n) {
    const size_t o = static_cast<size_t>(L.h_segoff[n]);
    std::copy(seg_n[n].begin(), seg_n[n].end(), L.h_seg.begin() + 2 * o);
    std::copy(err_n[n].begin(), err_n[n].end(), L.h_err.begin() + o);
  }
  L.max_err = 0;
  for (int v : L.h_err)
    L.max_err = std::max(L.max_err, v);
  for (size_t k = 0; k < rec.size(); ++k) {
    // the lookup's packed copy: slope, intercept and error bound in one record
    rec[k].a = L.h_seg[2 * k];
    rec[k].b = L.h_seg[2 * k + 1];
    rec[k].err = L.h_err[k];
    rec[k].pad = 0;
  }
  LSeg* d_rec = lalloc<LSeg>(L.allocs, rec.size());
  float* d_seg = lalloc<float>(L.allocs, L.h_seg.size());
  int* d_err = lalloc<int>(L.allocs, L.h_err.size());
  double* d_ulo = lalloc<double>(L.allocs, nn);
  double* d_invw = lalloc<double>(L.allocs, nn);
  int* d_segoff = lalloc<int>(L.allocs, nn);
  int* d_segn = lalloc<int>(L.allocs, nn);
  lcheck(hipMemcpy(d_rec, rec.data(), sizeof(LSeg) * rec.size(), hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_seg, L.h_seg.data(), sizeof(float) * L.h_seg.size(), hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_err, L.h_err.data(), sizeof(int) * L.h_err.size(), hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_ulo, L.h_ulo.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_invw, L.h_invw.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_segoff, L.h_segoff.data(), sizeof(int) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_segn, L.h_segn.data(), sizeof(int) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  L.n_seg = n_seg;
  L.err_bound = max_err_bound;
  P.learned.rec = d_rec;
  P.learned.seg = d_seg;
  P.learned.err = d_err;
  P.learned.u_lo = d_ulo;
  P.learned.inv_w = d_invw;
  P.learned.seg_off = d_segoff;
  P.learned.seg_n = d_segn;
  P.learned.n_seg = n_seg;
  P.learned.enabled = 1;
  return L.max_err;
}

int learned_loss_and_grad(const DevProblem& P, int i_nuc, int n_seg, const float* params_in, double* loss,
  double* grad, hipStream_t st)
{
  const HostModel& m = model_state();
  const int nn = static_cast<int>(m.nuclides.size());
  if (i_nuc < 0 || i_nuc >= nn)
    throw std::runtime_error("nuclide index out of bounds");
  if (n_seg < 1 || n_seg > LEARNED_MAX_SEG)
    throw std::runtime_error("n_seg out of range");
  Round r;
  prepare(P, m, std::vector<int> {i_nuc}, n_seg, r, st);
  float2* params = lalloc<float2>(r.tmp, n_seg);
  lcheck(hipMemcpy(params, params_in, sizeof(float2) * n_seg, hipMemcpyHostToDevice), "hipMemcpy");
  hipLaunchKernelGGL(k_lforward, dim3(n_seg, 1), dim3(LW), 0, st, P, r.d_list, r.uu, r.seg_start, n_seg, params,
    r.loss);
  hipLaunchKernelGGL(k_lbackward, dim3(n_seg, 1), dim3(LW), 0, st, P, r.d_list, r.uu, r.seg_start, n_seg, params,
    r.grad);
  lcheck(hipStreamSynchronize(st), "learned loss/grad");
  lcheck(hipMemcpy(loss, r.loss, sizeof(double) * n_seg, hipMemcpyDeviceToHost), "hipMemcpy");
  lcheck(hipMemcpy(grad, r.grad, sizeof(double2) * n_seg, hipMemcpyDeviceToHost), "hipMemcpy");
  return 0;
}

} // namespace omc

extern "C" {

// Loss and gradient of one nuclide's model at given parameters (FP32 pairs
// a_s, b_s for n_seg segments): loss[n_seg], grad[n_seg][2].  For the gradient
// checks against the CPU restatement of the same formulas.
int omc_learned_loss_grad(int i_nuc, int n_seg, const float* params, double* loss, double* grad)
{
  try {
    return omc::learned_loss_and_grad(omc::problem_state(), i_nuc, n_seg, params, loss, grad, omc::stream_state());
  } catch (const std::exception& e) {
    std::snprintf(openmc_err_msg, sizeof(openmc_err_msg), "%s", e.what());
    return -5;
  }
}

// Training summary: base n_seg, max error bound, mean error bound, loss
// before the first gradient step, after the last one, after recentring, the
// number of training rounds, the error bound asked for, and the total,
// smallest and largest segment counts over the nuclides.
int omc_learned_stats(double* out6)
{
  const omc::Learned& L = omc::learned_state();
  if (L.n_seg == 0)
    return -3;
  double mean = 0.0;
  for (int v : L.h_err)
    mean += v;
  mean /= L.h_err.empty() ? 1.0 : static_cast<double>(L.h_err.size());
  out6[0] = L.n_seg;
  out6[1] = L.max_err;
  out6[2] = mean;
  out6[3] = L.loss_init;
  out6[4] = L.loss_trained;
  out6[5] = L.last_loss;
  return 0;
}

int omc_learned_stats2(double* out5)
{
  const omc::Learned& L = omc::learned_state();
  if (L.n_seg == 0)
    return -3;
  out5[0] = L.rounds;
  out5[1] = L.err_bound;
  out5[2] = static_cast<double>(L.h_err.size());
  out5[3] = L.h_segn.empty() ? 0 : *std::min_element(L.h_segn.begin(), L.h_segn.end());
  out5[4] = L.h_segn.empty() ? 0 : *std::max_element(L.h_segn.begin(), L.h_segn.end());
  return 0;
}

// The trained model of one nuclide: its segment count is returned, and with
// params and err not null, params[n][2] and err[n] are filled (n = that count).
int omc_learned_model(int i_nuc, float* params, int* err)
{
  const omc::Learned& L = omc::learned_state();
  const int nn = static_cast<int>(L.h_segn.size());
  if (L.n_seg == 0 || i_nuc < 0 || i_nuc >= nn)
    return -3;
  const int ns = L.h_segn[i_nuc];
  const size_t o = static_cast<size_t>(L.h_segoff[i_nuc]);
  if (params)
    std::memcpy(params, L.h_seg.data() + 2 * o, sizeof(float) * 2 * ns);
  if (err)
    std::memcpy(err, L.h_err.data() + o, sizeof(int) * ns);
  return ns;
}

} // extern "C"
