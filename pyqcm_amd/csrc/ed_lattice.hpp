// ed_lattice.hpp -- sums over a wavevector grid of the superlattice for the
// lattice quantities built from a cluster Green function: the CPT Green
// function averaged over k (lattice_model_instance::projected_Green_function,
// src_qcm/CPT.cpp:283-343) and the integrand of the self-energy functional on
// a (frequency, k) grid (lattice_model_instance::Potthoff_functional and
// SEF_integrand_k, src_qcm/lattice_model_instance.cpp:503-584, :624-650).
#pragma once

#include <vector>

#include "ed_common.hpp"

namespace ed {

// Green-function dimensions handled by the wave-per-matrix kernels
constexpr int kLatticeMaxDim = 8;
constexpr int kLatticeMaxNbr = 64;

// V(k) = V0 + sum_e val_e exp(2 pi i k . R[nbr_e]) at position (row_e, col_e),
// k in reduced coordinates of the superlattice (lattice_model_instance::set_V,
// src_qcm/CPT.cpp:22-60: V = t(k) - t'); matrices column-major complex
struct Superlattice {
  int d = 0;
  std::vector<double> V0;   // 2 d^2
  std::vector<int> row, col, nbr;
  std::vector<double> val;  // 2 per element
  std::vector<double> R;    // 3 per neighbor vector
};

// wavevector grid: `dim` axes (0..3) of nk[a] points at (i + offset) / nk[a]
struct KGrid {
  int dim = 0;
  int nk[3] = {1, 1, 1};
  double offset = 0.0;
  long count() const;
};

// device copy of a Superlattice: the elements grouped by matrix position
struct SuperlatticeDev {
  int d = 0, nnbr = 0;
  DevBuf<double> V0, val, R;
  DevBuf<int> ptr, nbr;  // ptr: d^2 + 1 over positions (column-major)
  explicit SuperlatticeDev(const Superlattice &S);
};

// per frequency iz: (1/nk) sum_k log|det(V(k) G_iz - 1)| (out: nz doubles)
void lattice_logdet(const SuperlatticeDev &S, const KGrid &K, int nz, const double *dG, double *out);
// per frequency iz: (1/nk) sum_k G_iz (1 - V(k) G_iz)^{-1} (out: nz d x d
// complex, column-major)
void lattice_cpt_average(const SuperlatticeDev &S, const KGrid &K, int nz, const double *dG, double *out);

}  // namespace ed
