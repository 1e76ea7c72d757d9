// Learned energy-grid index: per-nuclide piecewise-linear models of the grid
// index as a function of x = log(E / E_min), trained on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "dev_types.h"
#include "host_model.h"

namespace omc {

struct Learned {
  std::vector<void*> allocs;
  int n_seg = 0;
  int max_err = 0;
  double last_loss = 0.0;   // sum of squared residuals after recentring
  double loss_init = 0.0;   // before the first gradient step
  double loss_trained = 0.0; // after the last gradient step
  std::vector<float> h_seg; // [n_nuc][n_seg][2]
  std::vector<int> h_err;   // [n_nuc][n_seg]
  std::vector<double> h_ulo;  // [n_nuc] log(E_first / E_min)
  std::vector<double> h_invw; // [n_nuc] n_seg / (log(E_last / E_min) - u_lo)
};

// Fit all nuclides' models (forward/backward/Adam kernels, `epochs` steps),
// record per-segment error bounds, and point P.learned at the result.
// Returns the largest error bound.
int build_learned_index(Learned& L, const HostModel& m, DevProblem& P, int n_seg, int epochs, hipStream_t st);
void free_learned(Learned& L);

// Training kernels exposed for the gradient checks:
// loss[s] = sum_i (a_s * f_i + b_s - i)^2 over the points of segment s,
// grad[s] = d loss / d (a_s, b_s), for one nuclide with given parameters.
int learned_loss_and_grad(const DevProblem& P, int i_nuc, int n_seg, const float* params, double* loss,
  double* grad, hipStream_t st);

Learned& learned_state();
DevProblem& problem_state();
hipStream_t stream_state();
const HostModel& model_state();

} // namespace omc
