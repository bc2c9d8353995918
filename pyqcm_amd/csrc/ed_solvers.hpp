// ed_solvers.hpp -- Krylov drivers (host control, device vectors).
#pragma once

#include <vector>

#include "ed_model.hpp"

namespace ed {

struct LanczosGS {
  double E0 = 0.0;
  int niter = 0;
  std::vector<double> alpha, beta;
};

// LanczosEigenvalue + LanczosEigenvector (src_ed/Hamiltonian/Lanczos.hpp:41-143).
// d_start: normalized start vector (device); d_gs receives the Ritz vector.
// proj: optional projection on a point-group representation, applied to every
// new Krylov vector and to the eigenvector (the start must already be projected)
LanczosGS lanczos_gs(Hamiltonian &H, const double *d_start, double *d_gs, const GroupProjector *proj = nullptr);

// LanczosGreen (src_ed/Hamiltonian/Lanczos.hpp:156-183); d_psi normalized.
// max_steps < 0: max_iter_CF; otherwise at most max_steps coefficients (the first
// ones of the unbounded run)
void lanczos_green(Hamiltonian &H, const double *d_psi, std::vector<double> &a, std::vector<double> &b,
                   long max_steps = -1);

// bandLanczos + build_Q_matrix (Lanczos.hpp:214-395, Hamiltonian_base.hpp:230-254).
// phi: p device vectors (unnormalized).  Returns eigenvalues (ascending) and
// Q (p x M column-major) = P[:, :p] U[:p, :].
void band_lanczos(Hamiltonian &H, const std::vector<const double *> &phi, std::vector<double> &evals,
                  std::vector<double> &Q, int M_in = -1);
// the same in two parts: the Krylov steps on the device, and the host-only
// eigen-decomposition of the projected matrix, which may run on another
// thread while the device works on something else
struct BandLanczosRun {
  int M = 0, p = 0, nd = 0;
  std::vector<double> T, P;  // projected matrix (M x M), P (p x min(p, M)), column-major
  void finish(std::vector<double> &evals, std::vector<double> &Q) const;
};
BandLanczosRun band_lanczos_run(Hamiltonian &H, const std::vector<const double *> &phi, int M_in = -1);

// Davidson-Liu (src_ed/Hamiltonian/Davidson.hpp:38-230); evecs: K device buffers of dim.
int davidson(Hamiltonian &H, int K, double accuracy, std::vector<double> &evals, std::vector<DevBuf<double>> &evecs);

// blockLanczos (qr: upper-triangular B, src_ed/Hamiltonian/Lanczos.hpp:426-572) or
// blockLanczosSVD (polar form, symmetric positive semi-definite B, :603-779); q:
// p orthonormal device vectors.  A, B: floors of p x p column-major.  Returns
// the number of floors.
int block_lanczos(Hamiltonian &H, const std::vector<const double *> &q, int M0, std::vector<double> &A,
                  std::vector<double> &B, bool qr);

// <a, b> on device
double ddot(size_t n, const double *a, const double *b);
// per-thread device buffer for per-block partial sums (16 x partials_cap())

}  // namespace ed
