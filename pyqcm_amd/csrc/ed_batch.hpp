// ed_batch.hpp -- batched solves of many small sectors (parameter points of a
// VCA / CDMFT sweep): one workgroup runs a whole Lanczos ground state, or a
// whole LanczosGreen continued fraction, for one work item.
#pragma once

#include <vector>

#include "ed_common.hpp"

namespace ed {

struct Instance;

// device view of one factorized sector: per-spin merged one-body tables (CSR
// with values, both triangles) and the full diagonal, label order
// I = i_up + dim_up * i_dn
struct BatchSector {
  uint32_t du = 0, dd = 0;
  const uint32_t *up_ptr = nullptr, *up_col = nullptr, *dn_ptr = nullptr, *dn_col = nullptr;
  const double *up_val = nullptr, *dn_val = nullptr, *diag = nullptr;
};

// Lanczos ground state of one sector (LanczosEigenvalue + LanczosEigenvector,
// src_ed/Hamiltonian/Lanczos.hpp:41-143): start vector normal_draw(seed, .)
// normalized, Ritz stop |z_last beta| <= accur, eigenvector by replaying the
// recurrence
struct BatchGSItem {
  BatchSector S;
  uint64_t seed = 0;
  double *ws = nullptr;   // 3 dim
  double *gs = nullptr;   // dim: normalized ground state (output)
  double *ab = nullptr;   // 2 (max_iter + 2): alpha, then beta
  double *tri = nullptr;  // 3 (max_iter + 2): eigenvector of T and LDL^T scratch
  double *out = nullptr;  // E0, iterations, status (0 ok, 1 not converged)
};

// LanczosGreen (src_ed/Hamiltonian/Lanczos.hpp:156-183) from
// psi = c_o1 |gs> (+ c_o2 |gs>), built in the kernel from half-basis maps
struct BatchCFItem {
  BatchSector S;              // target sector
  const double *gs = nullptr; // source state
  uint32_t du_s = 0;          // source up dimension
  int spin = 0, norb = 1;
  const int32_t *map[2] = {nullptr, nullptr};
  const int8_t *sgn[2] = {nullptr, nullptr};
  const int8_t *upar = nullptr;  // spin-down orbitals: (-1)^(up occupation)
  double *ws = nullptr;       // 3 dim
  double *a = nullptr;        // max_cf
  double *b = nullptr;        // max_cf + 1 (b[0] = 1)
  double *out = nullptr;      // n (-1: psi vanishes), ||psi||^2
};

void batch_gs(const BatchGSItem *d_items, int n, int block, int max_iter, double accur, hipStream_t s);
void batch_cf(const BatchCFItem *d_items, int n, int block, int max_cf, double acc, hipStream_t s);

// solves the ground state (and with want_gf the Green function) of every
// instance; factorizable instances using Lanczos + continued fractions run
// batched, the others through their own solve
void solve_batch(const std::vector<Instance *> &insts, bool want_gf);

}  // namespace ed
