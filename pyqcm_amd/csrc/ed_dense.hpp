// ed_dense.hpp -- host eigen-solvers for the small projected problems.
#pragma once

#include <cmath>
#include <vector>

namespace ed {

// sqrt(x^2 + y^2) without std::hypot's scaling when the squares cannot
// overflow or underflow (the common case here; std::hypot dominated the QL
// and the band reduction's rotations)
inline double fast_hypot(double x, double y) {
  const double ax = std::fabs(x), ay = std::fabs(y);
  const double m = ax > ay ? ax : ay;
  if (m < 1e150 && m > 1e-150) return std::sqrt(x * x + y * y);
  return std::hypot(x, y);
}

// implicit QL on a symmetric tridiagonal matrix (d diagonal, e[i] couples i and
// i+1); the rotations are accumulated into the nz rows of Z (row-major nz x n).
void ql_rows(int n, double *d, double *e, double *Z, int nz);

// k-th smallest eigenvalue (k = 0 lowest) of a symmetric band matrix
// (column-major n x n, half-bandwidth b), by bisection on LDL^T inertia.
double band_kth_eigenvalue(int n, int b, const double *A, int k);

// lowest eigenvalue of the leading n x n block of a symmetric band matrix held
// column-major with leading dimension lda (half-bandwidth b), by the same
// bisection.  hint: the lowest eigenvalue of a leading block of it (an upper
// bound, by Cauchy interlacing), or +inf; width: how far below hint the value
// is expected.  The bracket [hint - width, hint] is checked by inertia counts
// and widened until it holds the eigenvalue; the result is the same as from
// the full Gershgorin bracket, to the bisection's final width.
double band_lowest_eigenvalue(int n, int b, const double *A, int lda, double hint, double width);

// all eigenvalues (ascending) of a symmetric band matrix and the first p
// components of every eigenvector: R is p x n row-major, R[k*n + c] = U[k, c].
void band_eig_first_rows(int n, int b, const double *A, int p, std::vector<double> &w, std::vector<double> &R);

}  // namespace ed
