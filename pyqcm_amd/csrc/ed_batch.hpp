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
// recurrence.  ws and tri are used only by items that do not fit in LDS.
struct BatchGSItem {
  BatchSector S;
  uint64_t seed = 0;
  double *ws = nullptr;   // 2 dim
  double *gs = nullptr;   // dim: normalized ground state (output)
  double *tri = nullptr;  // kTriArrays (max_iter + 2): alpha, beta, beta^2, z, pivots
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
  double *ws = nullptr;       // 2 dim (items that do not fit in LDS)
  double *a = nullptr;        // max_cf
  double *b = nullptr;        // max_cf + 1 (b[0] = 1)
  double *out = nullptr;      // n (-1: psi vanishes), ||psi||^2
};

// tridiagonal work arrays of a ground-state item, each max_iter + 2 long
constexpr int kTriArrays = 6;
// largest LDS footprint of an item run from LDS
constexpr size_t kBatchLdsCap = 64 * 1024;

// LDS bytes of one item held entirely in LDS: nvec Lanczos vectors and ntri
// tridiagonal doubles, then the sector's diagonal and one-body tables (nu, nd
// entries); the kernels carve their LDS in this order
__host__ __device__ inline size_t batch_lds_bytes(int nvec, size_t ntri, uint32_t du, uint32_t dd, uint32_t nu,
                                                  uint32_t nd) {
  const size_t dim = (size_t)du * dd;
  return 8 * ((size_t)nvec * dim + ntri + dim + nu + nd) + 4 * ((size_t)du + 1 + nu + dd + 1 + nd);
}
constexpr int kGsVectors = 3;  // two Lanczos vectors, ground state
constexpr int kCfVectors = 2;  // two Lanczos vectors

// lds > 0: every item runs from LDS with that many dynamic LDS bytes;
// lds == 0: vectors in global memory.  block: 64 or 256 threads per item
void batch_gs(const BatchGSItem *d_items, int n, int block, size_t lds, int max_iter, double accur, hipStream_t s);
void batch_cf(const BatchCFItem *d_items, int n, int block, size_t lds, int max_cf, double acc, hipStream_t s);

// writes item k's continued fraction (out[2k] = n steps, out[2k + 1] =
// ||psi||^2), transformed with the state's energy e0[k], weight w[k] and
// electron / hole pm[k], into a continued-fraction store at off[fid[k]]
void batch_cf_store(int nitems, int max_cf, const double *out, const double *a, const double *b, const int *fid,
                    const int *off, const double *e0, const double *w, const int8_t *pm, double *A, double *B,
                    hipStream_t s);

// solves the ground state (and with want_gf the Green function) of every
// instance; factorizable instances using Lanczos + continued fractions run
// batched, the others through their own solve
void solve_batch(const std::vector<Instance *> &insts, bool want_gf);

}  // namespace ed
