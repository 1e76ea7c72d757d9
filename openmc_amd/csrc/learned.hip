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
//   k_lerror     recentre b_s on the residual range, then record
//                err_s = ceil(max_i |pred(f_i) - i|) over the segment's points
//                and one neighbour on each side, with the FP32 inference formula.
// The lookup then searches E[j - err - 1, j + err + 3) exactly (j = (int)pred),
// which contains the global lower_bound of any E that maps to segment s.
#include <hip/hip_runtime.h>

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

// uu_i for every grid point of every nuclide: grid (x over points, y = nuclide)
__global__ void k_lprep(DevProblem P, const double* u_lo, const double* inv_w, double* uu)
{
  const int n = blockIdx.y;
  const DevNuclide nd = P.nuc[n];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nd.n; i += gridDim.x * blockDim.x) {
    double lx = detm::log(P.E[nd.off + i] / P.energy_min);
    uu[nd.off + i] = (lx - u_lo[n]) * inv_w[n];
  }
}

// seg_start[n][s] = first point whose segment >= s, s in [0, n_seg]
__global__ void k_lsegbounds(DevProblem P, const double* uu, int n_seg, int* seg_start)
{
  const int n = blockIdx.y;
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
    seg_start[static_cast<int64_t>(n) * (n_seg + 1) + s] = (s == n_seg) ? nd.n : lo;
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
__global__ void k_linit(DevProblem P, const double* uu, const int* seg_start, int n_seg, float2* params)
{
  const int n = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg)
    return;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(n) * (n_seg + 1) + s;
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
  params[static_cast<int64_t>(n) * n_seg + s] = ab;
}

__global__ void __launch_bounds__(LW) k_lforward(DevProblem P, const double* uu, const int* seg_start, int n_seg,
  const float2* params, double* loss)
{
  const int n = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(n) * (n_seg + 1) + s;
  const int a0 = seg_start[k], a1 = seg_start[k + 1];
  const float2 ab = params[static_cast<int64_t>(n) * n_seg + s];
  double acc = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += LW) {
    double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    acc += r * r;
  }
  acc = wave_sum(acc);
  if (threadIdx.x == 0)
    loss[static_cast<int64_t>(n) * n_seg + s] = acc;
}

__global__ void __launch_bounds__(LW) k_lbackward(DevProblem P, const double* uu, const int* seg_start, int n_seg,
  const float2* params, double2* grad)
{
  const int n = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(n) * (n_seg + 1) + s;
  const int a0 = seg_start[k], a1 = seg_start[k + 1];
  const float2 ab = params[static_cast<int64_t>(n) * n_seg + s];
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
    grad[static_cast<int64_t>(n) * n_seg + s] = double2 {ga, gb};
}

// Adam on the per-point mean gradient
__global__ void k_ladam(DevProblem P, const int* seg_start, int n_seg, const double2* grad, float2* params,
  double2* m1, double2* m2, int step, double lr)
{
  const int n = blockIdx.y;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_seg)
    return;
  const int64_t k = static_cast<int64_t>(n) * (n_seg + 1) + s;
  const int cnt = seg_start[k + 1] - seg_start[k];
  if (cnt < 2)
    return;
  const int64_t j = static_cast<int64_t>(n) * n_seg + s;
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

// Recentre and record the error bound over points [a0 - 1, a1] of segment s.
__global__ void __launch_bounds__(LW) k_lerror(DevProblem P, const double* uu, const int* seg_start, int n_seg,
  float2* params, int* err)
{
  const int n = blockIdx.y;
  const int s = blockIdx.x;
  const DevNuclide nd = P.nuc[n];
  const double* u = uu + nd.off;
  const int64_t k = static_cast<int64_t>(n) * (n_seg + 1) + s;
  int a0 = seg_start[k] - 1, a1 = seg_start[k + 1] + 1;
  a0 = a0 < 0 ? 0 : a0;
  a1 = a1 > nd.n ? nd.n : a1;
  const int64_t j = static_cast<int64_t>(n) * n_seg + s;
  float2 ab = params[j];
  double rmin = 1e300, rmax = -1e300;
  for (int i = a0 + threadIdx.x; i < a1; i += LW) {
    double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    rmin = fmin(rmin, r);
    rmax = fmax(rmax, r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    rmin = fmin(rmin, __shfl_xor(rmin, o));
    rmax = fmax(rmax, __shfl_xor(rmax, o));
  }
  if (rmax >= rmin)
    ab.y = static_cast<float>(static_cast<double>(ab.y) - 0.5 * (rmax + rmin));
  double amax = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += LW) {
    double r = static_cast<double>(predict(ab, u[i], s)) - static_cast<double>(i);
    amax = fmax(amax, fabs(r));
  }
  for (int o = 32; o > 0; o >>= 1)
    amax = fmax(amax, __shfl_xor(amax, o));
  if (threadIdx.x == 0) {
    params[j] = ab;
    err[j] = static_cast<int>(ceil(amax));
  }
}

void free_learned(Learned& L)
{
  for (void* p : L.allocs)
    (void)hipFree(p);
  L.allocs.clear();
}

namespace {

struct Work {
  std::vector<void*> tmp;
  double* uu;
  int* seg_start;
  double* loss;
  double2* grad;
};

void prepare(const DevProblem& P, const HostModel& m, int n_seg, std::vector<double>& u_lo, std::vector<double>& inv_w,
  double*& d_ulo, double*& d_invw, Work& w, std::vector<void*>& keep, hipStream_t st)
{
  const int nn = static_cast<int>(m.nuclides.size());
  u_lo.resize(nn);
  inv_w.resize(nn);
  int64_t total = 0;
  int maxn = 0;
  for (int n = 0; n < nn; ++n) {
    const auto& E = m.nuclides[n].energy;
    double x0 = detm::log(E.front() / m.energy_min);
    double x1 = detm::log(E.back() / m.energy_min);
    u_lo[n] = x0;
    inv_w[n] = static_cast<double>(n_seg) / (x1 - x0);
    total += static_cast<int64_t>(E.size());
    maxn = std::max<int>(maxn, static_cast<int>(E.size()));
  }
  d_ulo = lalloc<double>(keep, nn);
  d_invw = lalloc<double>(keep, nn);
  lcheck(hipMemcpy(d_ulo, u_lo.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  lcheck(hipMemcpy(d_invw, inv_w.data(), sizeof(double) * nn, hipMemcpyHostToDevice), "hipMemcpy");
  w.uu = lalloc<double>(w.tmp, total);
  w.seg_start = lalloc<int>(w.tmp, static_cast<size_t>(nn) * (n_seg + 1));
  w.loss = lalloc<double>(w.tmp, static_cast<size_t>(nn) * n_seg);
  w.grad = lalloc<double2>(w.tmp, static_cast<size_t>(nn) * n_seg);
  int bx = std::min(1024, (maxn + 255) / 256);
  hipLaunchKernelGGL(k_lprep, dim3(bx, nn), dim3(256), 0, st, P, d_ulo, d_invw, w.uu);
  int sx = (n_seg + 1 + 255) / 256;
  hipLaunchKernelGGL(k_lsegbounds, dim3(sx, nn), dim3(256), 0, st, P, w.uu, n_seg, w.seg_start);
  lcheck(hipGetLastError(), "learned prep");
}

} // namespace

int build_learned_index(Learned& L, const HostModel& m, DevProblem& P, int n_seg, int epochs, hipStream_t st)
{
  if (n_seg < 1 || n_seg > 1 << 20)
    throw std::runtime_error("n_seg out of range");
  free_learned(L);
  const int nn = static_cast<int>(m.nuclides.size());
  std::vector<double> u_lo, inv_w;
  double *d_ulo = nullptr, *d_invw = nullptr;
  Work w;
  prepare(P, m, n_seg, u_lo, inv_w, d_ulo, d_invw, w, L.allocs, st);
  float2* params = lalloc<float2>(L.allocs, static_cast<size_t>(nn) * n_seg);
  int* err = lalloc<int>(L.allocs, static_cast<size_t>(nn) * n_seg);
  double2* m1 = lalloc<double2>(w.tmp, static_cast<size_t>(nn) * n_seg);
  double2* m2 = lalloc<double2>(w.tmp, static_cast<size_t>(nn) * n_seg);
  int sx = (n_seg + 255) / 256;
  std::vector<double> loss(static_cast<size_t>(nn) * n_seg);
  auto total_loss = [&]() {
    hipLaunchKernelGGL(k_lforward, dim3(n_seg, nn), dim3(LW), 0, st, P, w.uu, w.seg_start, n_seg, params, w.loss);
    lcheck(hipGetLastError(), "learned forward");
    lcheck(hipMemcpyAsync(loss.data(), w.loss, sizeof(double) * loss.size(), hipMemcpyDeviceToHost, st),
      "hipMemcpy");
    lcheck(hipStreamSynchronize(st), "learned forward");
    double t = 0.0;
    for (double v : loss)
      t += v;
    return t;
  };
  hipLaunchKernelGGL(k_linit, dim3(sx, nn), dim3(256), 0, st, P, w.uu, w.seg_start, n_seg, params);
  L.loss_init = total_loss();
  for (int e = 1; e <= epochs; ++e) {
    hipLaunchKernelGGL(k_lbackward, dim3(n_seg, nn), dim3(LW), 0, st, P, w.uu, w.seg_start, n_seg, params, w.grad);
    double lr = 0.5 * std::pow(0.98, e);
    hipLaunchKernelGGL(k_ladam, dim3(sx, nn), dim3(256), 0, st, P, w.seg_start, n_seg, w.grad, params, m1, m2, e, lr);
  }
  L.loss_trained = total_loss();
  hipLaunchKernelGGL(k_lerror, dim3(n_seg, nn), dim3(LW), 0, st, P, w.uu, w.seg_start, n_seg, params, err);
  lcheck(hipGetLastError(), "learned training");
  L.last_loss = total_loss();
  L.h_seg.resize(static_cast<size_t>(nn) * n_seg * 2);
  L.h_err.resize(static_cast<size_t>(nn) * n_seg);
  lcheck(hipMemcpy(L.h_seg.data(), params, sizeof(float) * L.h_seg.size(), hipMemcpyDeviceToHost), "hipMemcpy");
  lcheck(hipMemcpy(L.h_err.data(), err, sizeof(int) * L.h_err.size(), hipMemcpyDeviceToHost), "hipMemcpy");
  for (void* p : w.tmp)
    (void)hipFree(p);
  L.max_err = 0;
  for (int v : L.h_err)
    L.max_err = std::max(L.max_err, v);
  L.n_seg = n_seg;
  L.h_ulo = u_lo;
  L.h_invw = inv_w;
  {
    // the lookup's packed copy: slope, intercept and error bound in one record
    std::vector<LSeg> rec(L.h_err.size());
    for (size_t k = 0; k < rec.size(); ++k) {
      rec[k].a = L.h_seg[2 * k];
      rec[k].b = L.h_seg[2 * k + 1];
      rec[k].err = L.h_err[k];
      rec[k].pad = 0;
    }
    LSeg* d_rec = lalloc<LSeg>(L.allocs, rec.size());
    lcheck(hipMemcpy(d_rec, rec.data(), sizeof(LSeg) * rec.size(), hipMemcpyHostToDevice), "hipMemcpy");
    P.learned.rec = d_rec;
  }
  P.learned.seg = reinterpret_cast<const float*>(params);
  P.learned.err = err;
  P.learned.u_lo = d_ulo;
  P.learned.inv_w = d_invw;
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
  std::vector<double> u_lo, inv_w;
  double *d_ulo = nullptr, *d_invw = nullptr;
  Work w;
  std::vector<void*> keep;
  prepare(P, m, n_seg, u_lo, inv_w, d_ulo, d_invw, w, keep, st);
  float2* params = lalloc<float2>(w.tmp, static_cast<size_t>(nn) * n_seg);
  lcheck(hipMemcpy(params + static_cast<size_t>(i_nuc) * n_seg, params_in, sizeof(float2) * n_seg,
           hipMemcpyHostToDevice),
    "hipMemcpy");
  hipLaunchKernelGGL(k_lforward, dim3(n_seg, nn), dim3(LW), 0, st, P, w.uu, w.seg_start, n_seg, params, w.loss);
  hipLaunchKernelGGL(k_lbackward, dim3(n_seg, nn), dim3(LW), 0, st, P, w.uu, w.seg_start, n_seg, params, w.grad);
  lcheck(hipStreamSynchronize(st), "learned loss/grad");
  lcheck(hipMemcpy(loss, w.loss + static_cast<size_t>(i_nuc) * n_seg, sizeof(double) * n_seg, hipMemcpyDeviceToHost),
    "hipMemcpy");
  lcheck(hipMemcpy(grad, w.grad + static_cast<size_t>(i_nuc) * n_seg, sizeof(double2) * n_seg, hipMemcpyDeviceToHost),
    "hipMemcpy");
  for (void* p : w.tmp)
    (void)hipFree(p);
  for (void* p : keep)
    (void)hipFree(p);
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

// Training summary: n_seg, max error bound, mean error bound, loss before the
// first gradient step, after the last one, and after recentring.
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

// The trained model of one nuclide: params[n_seg][2] and err[n_seg].
int omc_learned_model(int i_nuc, float* params, int* err)
{
  const omc::Learned& L = omc::learned_state();
  const int nn = omc::problem_state().n_nuclides;
  if (L.n_seg == 0 || i_nuc < 0 || i_nuc >= nn)
    return -3;
  std::memcpy(params, L.h_seg.data() + static_cast<size_t>(i_nuc) * L.n_seg * 2, sizeof(float) * 2 * L.n_seg);
  std::memcpy(err, L.h_err.data() + static_cast<size_t>(i_nuc) * L.n_seg, sizeof(int) * L.n_seg);
  return L.n_seg;
}

} // extern "C"
