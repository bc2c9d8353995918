// ed_dense.hpp -- host eigen-solvers for the small projected problems.
#pragma once

#include <vector>

namespace ed {

// implicit QL on a symmetric tridiagonal matrix (d diagonal, e[i] couples i and
// i+1); the rotations are accumulated into the nz rows of Z (row-major nz x n).
void ql_rows(int n, double *d, double *e, double *Z, int nz);

// k-th smallest eigenvalue (k = 0 lowest) of a symmetric band matrix
// (column-major n x n, half-bandwidth b), by bisection on LDL^T inertia.
double band_kth_eigenvalue(int n, int b, const double *A, int k);

// all eigenvalues (ascending) of a symmetric band matrix and the first p
// components of every eigenvector: R is p x n row-major, R[k*n + c] = U[k, c].
void band_eig_first_rows(int n, int b, const double *A, int p, std::vector<double> &w, std::vector<double> &R);

}  // namespace ed
