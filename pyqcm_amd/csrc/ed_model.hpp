// ed_model.hpp -- sectors, bases, models, operators, Hamiltonian plans (host side).
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <limits>
#include <vector>

#include "ed_common.hpp"
#include "ed_kernels.hpp"

namespace ed {

using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// sector (src_ed/sector.hpp:23-147)
// ---------------------------------------------------------------------------
struct Sector {
  static constexpr int odd = 9999;
  static constexpr int even = 10000;
  int N = 0, S = 0, irrep = 0;
  Sector() = default;
  Sector(int n, int s, int r) : N(n), S(s), irrep(r) {}
  static Sector parse(const std::string &str);
  std::string name() const;
  int Nup() const;
  int Ndw() const;
  bool operator<(const Sector &o) const;  // sector.hpp:150-163 ordering
  bool operator==(const Sector &o) const { return N == o.N && S == o.S && irrep == o.irrep; }
};

// ---------------------------------------------------------------------------
// half basis (src_ed/ED_basis.cpp:115-135), cached per (L, N)
// ---------------------------------------------------------------------------
struct HalfBasis {
  int L = 0, N = 0;
  std::vector<uint32_t> bin;
  uint32_t dim = 0;
  uint32_t index(uint32_t b) const;  // lower_bound; returns dim if absent
};
std::shared_ptr<const HalfBasis> half_basis(int L, int N);

// ---------------------------------------------------------------------------
// model and operators (src_ed/model.hpp, src_ed/Operators/*)
// ---------------------------------------------------------------------------
struct Element {
  int r, c;
  double v;
  double vi = 0.0;  // imaginary part (operators defined with complex elements)
};

enum OpKind { OP_ONE_BODY = 0, OP_INTERACTION = 1 };

struct Operator {
  std::string name;
  int kind = OP_ONE_BODY;
  std::vector<Element> el;   // Hermitian-completed (one-body) / ordered (interaction)
  bool spin_flip = false;    // one-body elements coupling up and down
  bool spin_asym = false;    // up and down parts differ (mixing up_down)
  // defined with complex elements (one_body_operator<Complex>): an instance
  // with a nonzero coefficient on it works in a complex Hilbert space
  bool is_complex = false;
  // the imaginary parts as an operator of their own (elements v = Im, an
  // antisymmetric one-body matrix); null when every element is real
  std::shared_ptr<Operator> imag;
};

// ---------------------------------------------------------------------------
// Abelian point group of a cluster (symmetry_group, src_ed/symmetry_group.cpp:16-330):
// an element maps orbital i -> perm[i] for the first np orbitals (the sites with
// bath_irrep, all orbitals otherwise) and multiplies an occupied orbital i >= np
// by exp(2 pi i phase[i] / g).  Elements are products of the generators'
// powers, numbered with the first generator's power the fastest digit;
// irrep r has characters chi_r(k) = exp(2 pi i sum_j r_j k_j / n_j) (n_j the
// order of generator j, r_j / k_j the digits of r / k).  Only groups with real
// characters (generators of order 1 or 2) are solved here.
// ---------------------------------------------------------------------------
struct PointGroup {
  int g = 1, L = 0, ns = 0, np = 0;
  bool bath_irrep = false;
  std::vector<int> order;                 // order of each generator
  std::vector<std::vector<int>> e;        // elements (length L: permutation, then phases)
  std::vector<std::vector<double>> chi;   // chi[r][k] (real)
  // symmetric orbitals: rows of S (n_sites x n_sites, real), irrep by irrep;
  // irrep_dim[r] orbitals in irrep r, starting at row irrep_off[r]
  std::vector<double> S;
  std::vector<int> irrep_dim, irrep_off;
  static PointGroup build(int n_sites, int L, const std::vector<std::vector<int>> &gens, bool bath_irrep);
  // image of a one-spin occupation word under element k and its sign
  // (fermionic reordering x orbital phases)
  uint32_t apply(int k, uint32_t b, int &sign) const;
};

// P_r = (1 / g) sum_k chi_r(k) U(e_k) on the factorized sector (Nup, Ndw): the
// many-body action of an element is a signed permutation of each spin's half
// states (site permutation, fermionic reordering sign, orbital phases)
struct GroupProjector {
  size_t dim = 0;
  ProjArgs args;
  DevBuf<int32_t> inv_up, inv_dn;
  DevBuf<int8_t> sg_up, sg_dn;
  GroupProjector(const PointGroup &G, int irrep, int Nup, int Ndw);
  void apply(const double *x, double *y) const;
};

struct Model {
  std::string name;
  int n_sites = 0, n_bath = 0, L = 0;
  std::vector<std::string> order;            // definition order
  std::map<std::string, Operator> ops;
  bool closed = false;
  PointGroup group;                          // trivial (g = 1) without generators
};

Model &get_model(const std::string &name);
void new_model(const std::string &name, int n_sites, int n_bath,
               const std::vector<std::vector<int>> &generators = {}, bool bath_irrep = false);
void new_operator(const std::string &model, const std::string &name, const std::string &type,
                  const std::vector<Element> &el, bool is_complex = false);

// (operator, coefficient) pairs of an instance
struct TermList {
  int L = 0;
  std::vector<std::pair<const Operator *, double>> terms;
};

// ---------------------------------------------------------------------------
// Hamiltonian plans: Hamiltonian<T>::mult_add replacement on device
// (src_ed/Hamiltonian/Hamiltonian_base.hpp:43-82)
// ---------------------------------------------------------------------------
struct Hamiltonian {
  Sector sec;
  int L = 0;
  int Nup = 0, Ndw = 0;
  size_t dim = 0;
  char format = 'G';
  int variant = 0;
  TermList terms;           // the instance terms this plan realizes
  virtual ~Hamiltonian();
  // y = ax H x + beta y on device; returns the number of per-block partials of
  // <y_new, x> written to a partials slot (take_partials) when want_dot
  virtual int apply(const double *x, double *y, double ax, double beta, bool want_dot) = 0;
  // the same product with the partials written to part (device memory, when
  // not null) and nothing allocated: no autotuning (the tuned schedule once
  // tuned, else the default one), larger partial sets in aux (aux_cap doubles)
  virtual int apply_fixed(const double *x, double *y, double ax, double beta, double *part, double *aux,
                          size_t aux_cap) = 0;
  // settle the product's schedule (autotuning where the plan has it) before
  // apply_fixed is queued or captured
  virtual void prepare() {}
  // diagonal of H on device
  virtual void diag(double *d) = 0;
  double dot_result(int nb);  // sum of the partials of the last apply (deterministic order)
 protected:
  // slot for the partials of the apply being launched (an unread earlier one is discarded)
  double *take_partials();
  double *pending = nullptr;
  int pending_nb = 0;
};

// a complex Hilbert space is used when a term with a nonzero coefficient was
// defined with complex elements, or when force_complex is set: the plan then
// acts on split vectors [Re | Im] of twice the sector dimension
bool terms_complex(const TermList &T);
std::unique_ptr<Hamiltonian> make_hamiltonian(const TermList &T, const Sector &sec, char format,
                                              bool force_complex = false);

// Explicit host form of the factorized Hamiltonian of one sector: per-spin
// merged symmetric one-body tables (CSR with values; both triangles) and the
// full diagonal (dim_up * dim_dn, label order).  Used by the batched solver.
struct FactorizedHost {
  uint32_t du = 0, dd = 0;
  std::vector<uint32_t> up_ptr, up_col, dn_ptr, dn_col;
  std::vector<double> up_val, dn_val, diag;
};
bool terms_factorizable(const TermList &T);
FactorizedHost factorized_tables(const TermList &T, const Sector &sec);
// time the candidate factorized schedules (fused; two-pass x slab widths 16/32/64
// x streaming on/off) and keep the fastest; returns ms per candidate
std::vector<double> hamiltonian_autotune(Hamiltonian &H, int &best);

// host strict-lower CSR pattern (Hamiltonian_CSR.hpp:60-81) for parity checks
void csr_lower(const TermList &T, int Nup, int Ndw, std::vector<int32_t> &Iptr, std::vector<int32_t> &J,
               std::vector<double> &v, std::vector<double> &diag);

// ---------------------------------------------------------------------------
// dense host helpers
// ---------------------------------------------------------------------------
// symmetric tridiagonal: lowest eigenvalue and the last components of all
// eigenvectors (for the Ritz residual); full eigenvector of the lowest state.
// lowest eigenvalue and last eigenvector component of a symmetric tridiagonal
// matrix; guess (e.g. the previous Ritz value) seeds a Rayleigh-quotient
// iteration, with bisection as the fallback
void tridiag_lowest(int n, const double *d, const double *e, double &ev0, double &zlast0,
                    double guess = std::numeric_limits<double>::quiet_NaN());
void tridiag_eigvec0(int n, const double *d, const double *e, double &ev0, std::vector<double> &z0,
                     double guess = std::numeric_limits<double>::quiet_NaN());
// dense symmetric eigensystem, A column-major n x n, eigenvalues ascending
void symeig(int n, const double *A, double *w, double *V /* may be null */);
// dense Hermitian eigensystem (cyclic Jacobi), A column-major n x n,
// eigenvalues ascending, eigenvectors in the columns of V (may be null)
void hermeig(int n, const cplx *A, double *w, cplx *V);

// Per-block partials of a reduction go to a slot of pinned host memory
// (partials_acquire: up to 16 rows of partials_cap() doubles); sum_partials
// waits for them, sums each row in a fixed order and releases the slot.
// Other (device) pointers are folded by a kernel.
double *partials_acquire();
double sum_partials(const double *dpart, int nb, int m = 1, std::vector<double> *out = nullptr);
// the same with row r at dpart + r * stride
double sum_partials_rows(const double *dpart, int nb, int m, size_t stride, std::vector<double> *out);

}  // namespace ed
