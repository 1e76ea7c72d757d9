// Host-callable launchers of the event kernels (transport.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dev_types.h"

namespace omc {

void launch_init(const DevProblem& P, const DevParticles& S, int64_t n, int64_t id_base, hipStream_t st);
void launch_cycle(const DevProblem& P, const DevParticles& S, int blocks, hipStream_t st);
void launch_kernel(int which, const DevProblem& P, const DevParticles& S, int blocks, hipStream_t st);
void launch_accumulate(double* value, double* results, int64_t n_bins, double norm, hipStream_t st);

// Standalone batched lookups (xs_lookup.hip): E[n] -> i_grid[n], xs[n][2]
void launch_lookup_batch(const DevProblem& P, int i_nuc, const double* E, int64_t n, int learned, int* i_grid,
  double* xs, unsigned long long* probes, hipStream_t st);
void launch_macro_batch(const DevProblem& P, int material, const double* E, int64_t n, int learned,
  double* macro, unsigned long long* probes, hipStream_t st);

void launch_copy(const void* src, void* dst, int64_t n_bytes, hipStream_t st);
void launch_math_batch(int fn, const double* x, const double* y, int64_t n, double* out, hipStream_t st);

} // namespace omc
