// ed_dense.cpp -- small dense/banded symmetric eigenproblems on the host (FP64).
//
// The band-Lanczos projected matrix is banded (half-bandwidth = number of
// starting vectors).  Its convergence test needs only the lowest eigenvalue:
// bisection on Sylvester inertia from a banded LDL^T (O(n b^2) per count).
// The final Lehmann representation needs every eigenvalue but only the first p
// components of each eigenvector: band -> tridiagonal reduction by Givens
// rotations with bulge chasing (O(n^2 b)), tracking the first p rows of the
// accumulated transform, then implicit QL accumulating into those p rows.
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "ed_dense.hpp"
#include "ed_model.hpp"

namespace ed {

// number of eigenvalues < sigma of the symmetric band matrix A (column-major
// n x n, half-bandwidth b): inertia of LDL^T of A - sigma I.
static int band_count_below(int n, int b, const double *A, double sigma, std::vector<double> &Lw,
                            std::vector<double> &D, double eps_piv, int lda = -1) {
  if (lda < 0) lda = n;
  // L stored as n x (b) : Lw[i * b + (i - k - 1)] = l_{i k} for k in [i-b, i-1]
  int neg = 0;
  for (int i = 0; i < n; ++i) {
    const int k0 = std::max(0, i - b);
    // l_{i k} for k in [k0, i-1]
    for (int k = k0; k < i; ++k) {
      double s = A[i + (size_t)k * lda];
      const int m0 = std::max(k0, k - b);
      for (int m = m0; m < k; ++m) s -= Lw[(size_t)i * b + (i - m - 1)] * Lw[(size_t)k * b + (k - m - 1)] * D[m];
      Lw[(size_t)i * b + (i - k - 1)] = s / D[k];
    }
    double d = A[i + (size_t)i * lda] - sigma;
    for (int k = k0; k < i; ++k) {
      const double l = Lw[(size_t)i * b + (i - k - 1)];
      d -= l * l * D[k];
    }
    if (std::fabs(d) < eps_piv) d = -eps_piv;
    D[i] = d;
    if (d < 0.0) ++neg;
  }
  return neg;
}

double band_kth_eigenvalue(int n, int b, const double *A, int k) {
  if (b < 1) b = 1;
  if (b > n - 1) b = std::max(1, n - 1);
  double lo = std::numeric_limits<double>::max(), hi = -lo, scale = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    for (int j = std::max(0, i - b); j <= std::min(n - 1, i + b); ++j)
      if (j != i) r += std::fabs(A[i + (size_t)j * n]);
    const double a = A[i + (size_t)i * n];
    lo = std::min(lo, a - r);
    hi = std::max(hi, a + r);
    scale = std::max(scale, std::fabs(a) + r);
  }
  std::vector<double> Lw((size_t)n * b, 0.0), D(n, 0.0);
  const double eps_piv = 1e-300 + 4.0 * std::numeric_limits<double>::epsilon() * scale * 1e-3;
  // bisection: smallest sigma with count(sigma) >= k + 1
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (band_count_below(n, b, A, mid, Lw, D, eps_piv) >= k + 1) hi = mid;
    else lo = mid;
    if (hi - lo <= 2.0 * std::numeric_limits<double>::epsilon() * std::max(std::fabs(lo), std::fabs(hi))) break;
  }
  return 0.5 * (lo + hi);
}

double band_lowest_eigenvalue(int n, int b, const double *A, int lda, double hint, double width) {
  if (b < 1) b = 1;
  if (b > n - 1) b = std::max(1, n - 1);
  double glo = std::numeric_limits<double>::max(), ghi = -glo, scale = 0.0;
  for (int i = 0; i < n; ++i) {
    double r = 0.0;
    for (int j = std::max(0, i - b); j <= std::min(n - 1, i + b); ++j)
      if (j != i) r += std::fabs(A[i + (size_t)j * lda]);
    const double a = A[i + (size_t)i * lda];
    glo = std::min(glo, a - r);
    ghi = std::max(ghi, a + r);
    scale = std::max(scale, std::fabs(a) + r);
  }
  std::vector<double> Lw((size_t)n * b, 0.0), D(n, 0.0);
  const double eps = std::numeric_limits<double>::epsilon();
  const double eps_piv = 1e-300 + 4.0 * eps * scale * 1e-3;
  auto count = [&](double s) { return band_count_below(n, b, A, s, Lw, D, eps_piv, lda); };
  double lo = glo, hi = ghi;
  if (std::isfinite(hint) && hint > glo && hint < ghi) {
    // upper end: the hint, nudged up, when it bounds the lowest eigenvalue
    const double h = hint + 8.0 * eps * scale;
    if (count(h) >= 1) {
      hi = h;
      double w = std::max(width, 1e3 * eps * scale);
      while (true) {
        const double l = hi - w;
        if (l <= glo) break;
        if (count(l) == 0) {
          lo = l;
          break;
        }
        hi = l;  // the eigenvalue lies below l: a tighter upper end
        w *= 16.0;
      }
    }
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count(mid) >= 1) hi = mid;
    else lo = mid;
    if (hi - lo <= 2.0 * eps * std::max(std::fabs(lo), std::fabs(hi))) break;
  }
  return 0.5 * (lo + hi);
}

void band_eig_first_rows(int n, int b, const double *Ain, int p, std::vector<double> &w, std::vector<double> &R) {
  // The reduction keeps every nonzero within distance b + 1 of the diagonal
  // (the bulge), so the symmetric matrix is held as its lower band of width
  // W = b + 2: Lb[i * W + d] = M[i][i - d].  The whole working set then stays
  // in cache (a dense n x n copy made every rotation a strided row walk).
  const int W = b + 2;
  std::vector<double> Lb((size_t)n * W, 0.0);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d <= std::min(b, i); ++d) Lb[(size_t)i * W + d] = Ain[i + (size_t)(i - d) * n];
  auto at = [&](int r, int c) -> double & {
    return r >= c ? Lb[(size_t)r * W + (r - c)] : Lb[(size_t)c * W + (c - r)];
  };
  // R: p x n (row-major) = first p rows of the accumulated transform
  R.assign((size_t)p * n, 0.0);
  for (int i = 0; i < std::min(p, n); ++i) R[(size_t)i * n + i] = 1.0;
  // two-sided rotation of rows / columns (pp, pp + 1)
  auto apply_rot = [&](int pp, double c, double s) {
    const int qq = pp + 1;
    const int j0 = std::max(0, qq - W + 1), j1 = std::min(n, pp + W);
    for (int j = j0; j < j1; ++j) {
      if (j == pp || j == qq) continue;
      double &ap = at(pp, j);
      double &aq = at(qq, j);
      const double x = ap, y = aq;
      ap = c * x + s * y;
      aq = -s * x + c * y;
    }
    double &app = at(pp, pp), &aqq = at(qq, qq), &apq = at(qq, pp);
    const double xpp = app, xqq = aqq, xpq = apq;
    app = c * c * xpp + 2 * c * s * xpq + s * s * xqq;
    aqq = s * s * xpp - 2 * c * s * xpq + c * c * xqq;
    apq = (c * c - s * s) * xpq + c * s * (xqq - xpp);
    for (int k = 0; k < p; ++k) {
      double &rp = R[(size_t)k * n + pp];
      double &rq = R[(size_t)k * n + qq];
      const double x = rp, y = rq;
      rp = c * x + s * y;
      rq = -s * x + c * y;
    }
  };
  if (b > 1) {
    for (int k = 0; k + 2 < n; ++k) {
      for (int l = std::min(b, n - 1 - k); l >= 2; --l) {
        int i = k + l;
        // annihilate M[i][k] with rows (i-1, i)
        double x = at(i - 1, k), y = at(i, k);
        if (y == 0.0) continue;
        double r = fast_hypot(x, y), c = x / r, s = y / r;
        apply_rot(i - 1, c, s);
        at(i, k) = 0.0;
        // chase the bulge created at (i + b, i - 1)
        int j = i;
        while (j + b < n) {
          const int row = j + b, col = j - 1;
          x = at(row - 1, col);
          y = at(row, col);
          if (y == 0.0) break;
          r = fast_hypot(x, y);
          c = x / r;
          s = y / r;
          apply_rot(row - 1, c, s);
          at(row, col) = 0.0;
          j = row;
        }
      }
    }
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);
  ql_rows(n, d.data(), e.data(), R.data(), p);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  w.resize(n);
  std::vector<double> R2((size_t)p * n);
  for (int c = 0; c < n; ++c) {
    w[c] = d[idx[c]];
    for (int k = 0; k < p; ++k) R2[(size_t)k * n + c] = R[(size_t)k * n + idx[c]];
  }
  R.swap(R2);
}

}  // namespace ed
