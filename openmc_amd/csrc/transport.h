// Host-callable launchers of the event kernels (transport.hip, xs_lookup.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.h"

namespace omc {

// The steps of one cycle, in launch order (KERNEL_SORT .. KERNEL_COLLISION),
// then source sampling.  SORT and SPLIT are partition passes (four kernels).
enum KernelKind : int {
  KERNEL_SORT = 0,      // staged survivors -> calculate-xs queue
  KERNEL_XS = 1,
  KERNEL_ADVANCE = 2,
  KERNEL_SPLIT = 3,     // advance outcomes -> surface / collision ranges
  KERNEL_SURFACE = 4,
  KERNEL_COLLISION = 5,
  KERNEL_INIT = 6,
  KERNEL_HISTORY = 7,   // history-based tail launches
  KERNEL_KINDS = 8
};

struct LaunchConfig {
  int blocks;   // workgroups of the grid-stride event kernels
  int xs_lanes; // lanes per particle in the calculate-xs kernel (1..64, power of two)
  int sort;     // order the calculate-xs queue by key (else only compact it)
  int sort_blocks; // maximum workgroups of the partition count and scatter kernels
  int xs_group;    // nuclide lookups interleaved per lane (1, 2, 4 or 8) when xs_lanes == 1
  int block;       // threads per workgroup of the event kernels (64, 128 or 256)
  int sort_every;  // order by (material, energy) every this many cycles, by material in between
  int fuse_advance; // advance in the calculate-xs launch (xs_lanes 1, xs_group 4)
};

// Source particles for population indices g_base + [0, n) of batch `batch`
// (1-based) into slots [0, n).
void launch_init(const DevProblem& P, const DevParticles& S, int64_t n, int64_t g_base, int64_t batch,
  hipStream_t st);
void launch_cycle(const DevProblem& P, const DevParticles& S, const LaunchConfig& lc, hipStream_t st, int cycle = 0);
// `cycle` is the index of the event cycle the kernel belongs to (it selects
// the calculate-xs partition's key under sort_every).
void launch_kernel(int which, const DevProblem& P, const DevParticles& S, const LaunchConfig& lc, hipStream_t st,
  int cycle = 0);
// History-based processing of the n_staged entries staged at the end of an
// event cycle: each live particle runs up to max_steps whole cycles in one
// lane; the number still alive afterwards is added to qcount[QC_LIVE].
void launch_history(const DevProblem& P, const DevParticles& S, int n_staged, int max_steps, hipStream_t st);

// value holds n_copies copies of the tally's n_bins bins (summed in copy order)
void launch_accumulate(double* value, double* results, int64_t n_bins, int n_copies, double norm, hipStream_t st);

// Standalone batched lookups (xs_lookup.hip): E[n] -> i_grid[n], xs[n][2]
void launch_lookup_batch(const DevProblem& P, int i_nuc, const double* E, int64_t n, int learned, int* i_grid,
  double* xs, unsigned long long* probes, hipStream_t st);
// Macroscopic (total, absorption) of one material at n energies with the
// calculate-xs lookup code, `group` (1, 2, 4 or 8) nuclide lookups
// interleaved per lane.
void launch_macro_batch(const DevProblem& P, int material, const double* E, int64_t n, int learned,
  double* macro, unsigned long long* probes, hipStream_t st, int group = 4);
void launch_copy(const void* src, void* dst, int64_t n_bytes, hipStream_t st);

// The allocation-free variant of launch_lookup_batch (xs_lookup.hip): E,
// i_grid and xs are device pointers of the caller; the probe partial sums
// live in the file-scope device array `scratch` (NOALLOC_MAX_BLOCKS + 1
// values), so at most NOALLOC_MAX_LOOKUPS energies per call.  With
// probes_host not null the call waits for the stream and returns the total.
constexpr int NOALLOC_MAX_BLOCKS = 16384;
constexpr int64_t NOALLOC_MAX_LOOKUPS = static_cast<int64_t>(NOALLOC_MAX_BLOCKS) * 256;
void launch_lookup_noalloc(const DevProblem& P, int i_nuc, const double* E, int64_t n, int learned, int* i_grid,
  double* xs, unsigned long long* probes_host, hipStream_t st);
void launch_math_batch(int fn, const double* x, const double* y, int64_t n, double* out, hipStream_t st);

} // namespace omc
