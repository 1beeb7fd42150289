// Small dense FP64 algebra on the host (K x K, column-major): the reference keeps this on the
// host through LAPACK (src/core/fe_solver.f90:41-74 dpotrf/dpotrs/dpotri, dsyev for the PSD clip
// in src/core/fe_pipeline.f90); K <= a few hundred, so it stays on the host here too.
#pragma once

#include <vector>

namespace fehip {
namespace dense {

// Upper Cholesky of the n x n SPD matrix A (col-major), solve A x = b and form A^{-1}.
// Returns 0 or the order of the first non-positive leading minor (LAPACK info convention).
int chol_solve_invert(const std::vector<double>& A, const std::vector<double>& b, int n, std::vector<double>& x,
                      std::vector<double>& Ainv);

// Symmetric eigen-decomposition by cyclic Jacobi rotations: A = V diag(w) V^T.
void sym_eig(const std::vector<double>& A, int n, std::vector<double>& w, std::vector<double>& V);

// Clip negative eigenvalues of the symmetric matrix to zero and rebuild it; returns 1 if any was clipped.
int psd_clip(std::vector<double>& A, int n);

}  // namespace dense
}  // namespace fehip
