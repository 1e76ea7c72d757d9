// Batched cross-section lookups outside the transport loop.
//
// k_lookup_batch: one nuclide, many energies -> grid index and (total,
// absorption), i.e. Nuclide::calculate_xs (src/nuclide.cpp:616-839) applied
// to an array; used by the parity tests and the lookup-rate measurement.
// k_macro_batch: one material, many energies -> macroscopic (total,
// absorption), i.e. Material::calculate_neutron_xs (src/material.cpp:829-901).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "device_physics.h"
#include "transport.h"

namespace omc {

constexpr int LB_BLOCK = 256;

__device__ __forceinline__ int clamp_log_index(const DevProblem& P, double lx)
{
  int i_log = static_cast<int>(lx / P.log_spacing);
  return i_log < 0 ? 0 : (i_log >= P.n_log_bins ? P.n_log_bins - 1 : i_log);
}

__global__ void __launch_bounds__(LB_BLOCK) k_lookup_batch(DevProblem P, int i_nuc, const double* E, int64_t n,
  int learned, int* i_grid, double* xs, unsigned long long* probes_out)
{
  unsigned long long probes_sum = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double e = E[i];
    double lx = detm::log(e / P.energy_min);
    int i_log = clamp_log_index(P, lx);
    int probes = 0;
    Micro m = nuclide_xs(P, i_nuc, e, i_log, lx, learned != 0, probes);
    i_grid[i] = m.i_grid;
    xs[2 * i] = m.total;
    xs[2 * i + 1] = m.absorption;
    probes_sum += probes;
  }
  for (int o = 32; o > 0; o >>= 1)
    probes_sum += __shfl_xor(probes_sum, o);
  if (probes_out && __lane_id() == 0 && probes_sum)
    atomicAdd(probes_out, probes_sum);
}

// Material::calculate_neutron_xs over an array of energies with the
// calculate-xs kernel's lookup code: every wavefront works on the one
// material, so the nuclide records are wavefront-uniform, and each lane
// interleaves G nuclide lookups (nuclide_xs_group); sums in material order.
template<int G, bool LEARNED>
__global__ void __launch_bounds__(LB_BLOCK) k_macro_batch(DevProblem P, int material, const double* E, int64_t n,
  double* macro, unsigned long long* probes_out)
{
  unsigned long long probes_sum = 0;
  const int j0 = uniform_load(P.mat_off + material);
  const int nn = uniform_load(P.mat_off + material + 1) - j0;
  for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x; i0 < n;
       i0 += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t i = i0 + threadIdx.x;
    // lanes past the end repeat the last energy and store nothing, so the
    // whole wavefront runs the group lookups together
    const double e = E[i < n ? i : n - 1];
    const double lx = detm::log(e / P.energy_min);
    const int i_log = clamp_log_index(P, lx);
    int probes = 0;
    double xs_t = 0.0, xs_a = 0.0;
    for (int jj = 0; jj < nn; jj += G) {
      const int kk = nn - jj < G ? nn - jj : G;
      Micro m[G];
      nuclide_xs_group<G, LEARNED>(P, P.matnuc + j0 + jj, kk, e, i_log, lx, probes, m);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        if (k < kk) {
          const double dens = uniform_load(&P.matnuc[j0 + jj + k].dens) * 1.0;
          xs_t += dens * m[k].total;
          xs_a += dens * m[k].absorption;
        }
      }
    }
    if (i < n) {
      *reinterpret_cast<double2*>(macro + 2 * i) = double2 {xs_t, xs_a};
      probes_sum += probes;
    }
  }
  for (int o = 32; o > 0; o >>= 1)
    probes_sum += __shfl_xor(probes_sum, o);
  if (probes_out && __lane_id() == 0 && probes_sum)
    atomicAdd(probes_out, probes_sum);
}

// Elementwise primitives of the histories, for bitwise host/device checks:
// 0 log, 1 sin, 2 cos, 3 sqrt, 4 x / y, 5 prn(seed = bits of x), 6 u64 -> f64 of bits of x
__global__ void k_math_batch(int fn, const double* x, const double* y, int64_t n, double* out)
{
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double v = x[i];
    double r;
    switch (fn) {
    case 0: r = detm::log(v); break;
    case 1: r = detm::sin(v); break;
    case 2: r = detm::cos(v); break;
    case 3: r = sqrt(v); break;
    case 4: r = v / y[i]; break;
    case 5: {
      uint64_t sd = static_cast<uint64_t>(__double_as_longlong(v));
      r = prn(&sd);
      break;
    }
    default: r = static_cast<double>(static_cast<uint64_t>(__double_as_longlong(v))); break;
    }
    out[i] = r;
  }
}

// Plain streaming copy: the calibration of HBM byte counters (a copy of N
// bytes reads N and writes N).
__global__ void __launch_bounds__(LB_BLOCK) k_copy(const double4* src, double4* dst, int64_t n)
{
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

//==============================================================================
// The allocation-free lookup batch: device-resident inputs and outputs from
// the caller, and per-workgroup partial sums of the search probes in a
// file-scope device array instead of a counter the library allocates.
//==============================================================================

// scratch[0, NOALLOC_MAX_BLOCKS): one probe partial per workgroup;
// scratch[NOALLOC_MAX_BLOCKS]: their total.  Every call on a device shares
// that device's copy.
__device__ __align__(256) unsigned long long scratch[NOALLOC_MAX_BLOCKS + 1];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
unsigned long long* scratch_ptr[NOALLOC_MAX_DEVICES] = {nullptr};
} // namespace

// One energy per thread (grid = ceil(n / LB_BLOCK) <= NOALLOC_MAX_BLOCKS);
// workgroup b writes the sum of its threads' probes to partial[b].
__global__ void __launch_bounds__(LB_BLOCK) k_lookup_partial(DevProblem P, int i_nuc, const double* E, int64_t n,
  int learned, int* i_grid, double* xs, unsigned long long* partial)
{
  __shared__ unsigned long long wsum[LB_BLOCK / 64];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * LB_BLOCK + threadIdx.x;
  unsigned long long pr = 0;
  if (i < n) {
    const double e = E[i];
    const double lx = detm::log(e / P.energy_min);
    const int i_log = clamp_log_index(P, lx);
    int probes = 0;
    const Micro m = nuclide_xs(P, i_nuc, e, i_log, lx, learned != 0, probes);
    i_grid[i] = m.i_grid;
    xs[2 * i] = m.total;
    xs[2 * i + 1] = m.absorption;
    pr = static_cast<unsigned long long>(probes);
  }
  for (int o = 32; o > 0; o >>= 1)
    pr += __shfl_xor(pr, o);
  if (__lane_id() == 0)
    wsum[threadIdx.x >> 6] = pr;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < LB_BLOCK / 64; ++w)
      t += wsum[w];
    partial[blockIdx.x] = t;
  }
}

// total = sum of partial[0, n_blocks), in a fixed order: one workgroup,
// thread t adds partials t, t + LB_BLOCK, ...; then the workgroup's sums in
// thread order.
__global__ void __launch_bounds__(LB_BLOCK) k_sum_partials(const unsigned long long* partial, int n_blocks,
  unsigned long long* total)
{
  __shared__ unsigned long long part[LB_BLOCK];
  unsigned long long t = 0;
  for (int b = threadIdx.x; b < n_blocks; b += LB_BLOCK)
    t += partial[b];
  part[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sum = 0;
    for (int k = 0; k < LB_BLOCK; ++k)
      sum += part[k];
    *total = sum;
  }
}

unsigned long long* noalloc_scratch()
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    throw std::runtime_error("no current HIP device for the scratch array");
  if (!scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    scratch_ptr[dev] = static_cast<unsigned long long*>(ptr);
  }
  return scratch_ptr[dev];
}

void launch_lookup_noalloc(const DevProblem& P, int i_nuc, const double* E, int64_t n, int learned, int* i_grid,
  double* xs, unsigned long long* probes_host, hipStream_t st)
{
  if (n < 0 || n > NOALLOC_MAX_LOOKUPS)
    throw std::runtime_error("omc_lookup_batch_noalloc: n = " + std::to_string(n) + " is outside [0, " +
                             std::to_string(NOALLOC_MAX_LOOKUPS) + "] (" + std::to_string(NOALLOC_MAX_BLOCKS) +
                             " workgroups of " + std::to_string(LB_BLOCK) + " lookups in the scratch array)");
  unsigned long long* sp = noalloc_scratch();
  const int blocks = static_cast<int>((n + LB_BLOCK - 1) / LB_BLOCK);
  if (blocks > 0)
    hipLaunchKernelGGL(k_lookup_partial, dim3(blocks), dim3(LB_BLOCK), 0, st, P, i_nuc, E, n, learned, i_grid, xs,
      sp);
  // every partial the sum reads was written by this call's lookup kernel
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(LB_BLOCK), 0, st, sp, blocks, sp + NOALLOC_MAX_BLOCKS);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw std::runtime_error(std::string("lookup kernels: ") + hipGetErrorString(e));
  if (probes_host) {
    const hipError_t c = hipMemcpyAsync(probes_host, sp + NOALLOC_MAX_BLOCKS, sizeof(unsigned long long),
      hipMemcpyDeviceToHost, st);
    const hipError_t w = c == hipSuccess ? hipStreamSynchronize(st) : c;
    if (w != hipSuccess)
      throw std::runtime_error(std::string("lookup probes: ") + hipGetErrorString(w));
  }
}

void launch_copy(const void* src, void* dst, int64_t n_bytes, hipStream_t st)
{
  int64_t n = n_bytes / static_cast<int64_t>(sizeof(double4));
  hipLaunchKernelGGL(k_copy, dim3(8192), dim3(LB_BLOCK), 0, st, static_cast<const double4*>(src),
    static_cast<double4*>(dst), n);
}

void launch_math_batch(int fn, const double* x, const double* y, int64_t n, double* out, hipStream_t st)
{
  int64_t b = (n + LB_BLOCK - 1) / LB_BLOCK;
  int blocks = static_cast<int>(b > 65536 ? 65536 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(k_math_batch, dim3(blocks), dim3(LB_BLOCK), 0, st, fn, x, y, n, out);
}

void launch_lookup_batch(const DevProblem& P, int i_nuc, const double* E, int64_t n, int learned, int* i_grid,
  double* xs, unsigned long long* probes, hipStream_t st)
{
  int64_t b = (n + LB_BLOCK - 1) / LB_BLOCK;
  int blocks = static_cast<int>(b > 65536 ? 65536 : (b < 1 ? 1 : b));
  hipLaunchKernelGGL(k_lookup_batch, dim3(blocks), dim3(LB_BLOCK), 0, st, P, i_nuc, E, n, learned, i_grid, xs, probes);
}

void launch_macro_batch(const DevProblem& P, int material, const double* E, int64_t n, int learned, double* macro,
  unsigned long long* probes, hipStream_t st, int group)
{
  int64_t b = (n + LB_BLOCK - 1) / LB_BLOCK;
  int blocks = static_cast<int>(b > 65536 ? 65536 : (b < 1 ? 1 : b));
#define OMC_MACRO(G, L)                                                                                               \
  hipLaunchKernelGGL((k_macro_batch<G, L>), dim3(blocks), dim3(LB_BLOCK), 0, st, P, material, E, n, macro, probes)
  if (n <= 0)
    return;
  const bool l = learned != 0;
  if (group == 1) {
    if (l) OMC_MACRO(1, true); else OMC_MACRO(1, false);
  } else if (group == 2) {
    if (l) OMC_MACRO(2, true); else OMC_MACRO(2, false);
  } else if (group == 8) {
    if (l) OMC_MACRO(8, true); else OMC_MACRO(8, false);
  } else {
    if (l) OMC_MACRO(4, true); else OMC_MACRO(4, false);
  }
#undef OMC_MACRO
}

} // namespace omc
