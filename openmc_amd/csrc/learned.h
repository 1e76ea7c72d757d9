// Learned energy-grid index: per-nuclide piecewise-linear models of the grid
// index as a function of x = log(E / E_min), trained on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "dev_types.h"
#include "host_model.h"

namespace omc {

// Largest recorded index error the lookup accepts by default: a nuclide whose
// model exceeds it anywhere is trained again with twice the segments.
constexpr int LEARNED_MAX_ERR = 64;
// Segments of one nuclide at most.
constexpr int LEARNED_MAX_SEG = 1 << 20;

struct Learned {
  std::vector<void*> allocs;
  int n_seg = 0;   // base segment count (every nuclide has at least this many)
  int max_err = 0;
  int err_bound = LEARNED_MAX_ERR; // the bound the build enforced
  int rounds = 0;  // training rounds (doublings + 1)
  double last_loss = 0.0;   // sum of squared residuals after recentring
  double loss_init = 0.0;   // before the first gradient step
  double loss_trained = 0.0; // after the last gradient step
  std::vector<float> h_seg; // [segments of all nuclides][2], nuclide n from h_segoff[n]
  std::vector<int> h_err;   // [segments of all nuclides]
  std::vector<int> h_segoff; // [n_nuc] first segment of each nuclide
  std::vector<int> h_segn;   // [n_nuc] segments of each nuclide
  std::vector<double> h_ulo;  // [n_nuc] log(E_first / E_min)
  std::vector<double> h_invw; // [n_nuc] segments / (log(E_last / E_min) - u_lo)
};

// Fit all nuclides' models (forward/backward/Adam kernels, `epochs` steps)
// with n_seg segments each, record per-segment error bounds, and train again
// with twice the segments every nuclide whose bound exceeds max_err_bound,
// until all are within it; then point P.learned at the result.  Returns the
// largest error bound (<= max_err_bound).
int build_learned_index(Learned& L, const HostModel& m, DevProblem& P, int n_seg, int epochs, hipStream_t st,
  int max_err_bound = LEARNED_MAX_ERR);
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
