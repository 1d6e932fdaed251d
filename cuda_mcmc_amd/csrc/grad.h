// Launchers of the chi^2 gradient / Gauss-Newton kernels (grad.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "fitdesc.h"

namespace mhre {

// chi2[n] (and grad[n][ldp] when grad != nullptr) of the 9-parameter model
// for the parameter vectors P[n][ldp] against one fit's packed data fd.
// ring != 0: the fit's closures lie inside the ring window (FitDesc::ring_ok),
// so the single-pass ring form is used.
hipError_t launch_eht_grad9(const FitDesc* fd, const float* P, int n, int ldp, double* chi2, float* grad,
                            int ring, hipStream_t s);
// chi2[n], J^T r [n][9] and the upper triangle of J^T J [n][45] (row-major)
// of the weighted residuals r.
hipError_t launch_eht_gn9(const FitDesc* fd, const float* P, int n, int ldp, double* chi2, double* jtr,
                          double* jtj, hipStream_t s);

// Residuals r [rows][neq] and Jacobian J [rows][neq][n] of quadratic systems
// x^T Q_e x + b_e . x - dat_e for parameter rows X [rows][n].
hipError_t launch_quad_jac(const float* Q, const float* b, const float* dat, const int* sys, const float* X,
                           int rows, int neq, int n, float* r, float* J, hipStream_t s);

}  // namespace mhre
