// ed_instance.hpp -- model instances: sectors, low-energy states, Green-function
// representations (model_instance<T>, src_ed/model_instance.hpp).
#pragma once

#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "ed_model.hpp"

namespace ed {

// Green-function representation of one state and one spin
// continued fractions packed on the device, already transformed: fraction
// f's a at A[off[f], off[f+1]) and its b_0 .. b_{n-1} at the same positions
// of B (the evaluation never reads b_n)
struct CfStore {
  DevBuf<double> A, B;
  DevBuf<int> off;  // nfrac + 1, device
  int nfrac = 0;
};

struct GFRep {
  char kind = 0;  // 'L' Lehmann (Q-matrix), 'F' continued fractions, 'M' matrix continued fractions
  int n = 0;      // GF dimension (cluster orbitals)
  // Lehmann: poles e[M], residues V (n x M, column-major), weight folded in;
  // in a complex Hilbert space Vi holds their imaginary parts (empty
  // otherwise) and G(r, c) = sum_i V(c, i) conj(V(r, i)) / (z - e_i)
  std::vector<double> e, V, Vi;
  // continued fractions per ordered pair (o1 >= o2) for electrons / holes; a, b
  // already transformed (continued_fraction ctor, src_ed/continued_fraction.cpp:33-42)
  std::vector<std::vector<double>> ea, eb, ha, hb;
  // complex Hilbert space: for each pair a > b (index a (a - 1) / 2 + b) the
  // fractions of c_b - i c_a (electrons, c^dagger) and c_b + i c_a (holes),
  // as the reference's build_cf for o1 < o2; empty otherwise
  std::vector<std::vector<double>> eai, ebi, hai, hbi;
  // or, when `store` is set, fractions frac0 + 2 q (electron) and frac0 + 2 q + 1
  // (hole) of that device store, with ea .. hb empty
  std::shared_ptr<const CfStore> store;
  int frac0 = 0;
  // matrix continued fractions: floors A, B (p x p), weight W (p x n)
  struct MCF { int p = 0, nfl = 0; std::vector<double> A, B, W; };
  MCF me, mh;
  // with combine_mcf: one fraction for G_e + G_h^T (mcf_set::build_combined,
  // src_ed/mcf_set.hpp:79-90), evaluated in place of the two
  MCF comb;
};

struct State {
  Sector sec;
  double energy = 0.0, weight = 0.0;
  DevBuf<double> psi;
  GFRep gf[2];
};

struct Instance {
  int label = 0;
  Model *model = nullptr;
  std::map<std::string, double> value;
  std::set<Sector> targets;
  int mixing = 0;  // 0 normal, 4 up_down
  char format = 'G';
  char gf_method = 'L';
  bool gs_solved = false, gf_solved = false;
  double GS_energy = 0.0, E0 = 0.0;
  std::vector<std::unique_ptr<State>> states;
  std::vector<std::tuple<std::string, double, double>> averages;
  bool averages_done = false;
  // Hamiltonians of the Green function's target sectors, built on host
  // threads while the ground state is solved (band-Lanczos method, one target
  // sector); taken by the Green-function build
  std::map<Sector, std::future<std::unique_ptr<Hamiltonian>>> gf_hams;
  std::mutex gf_hams_mu;
  std::unique_ptr<Hamiltonian> take_gf_hamiltonian(const Sector &tgt);

  TermList terms() const;
  // complex Hilbert space: a term defined with complex elements has a nonzero
  // coefficient (ED::new_model_instance, src_ed/qcm_ED.cpp:145-157); state
  // vectors are then split [Re | Im], twice the sector dimension
  bool complex_hs() const;
  // fails with ED_ERR_UNSUPPORTED for a complex Hilbert space
  void need_real(const char *what) const;
  std::pair<double, std::string> ground_state_solve();
  // weights, filters and E0 of the candidate states in 'states' (each with its
  // sector, energy and normalized vector); marks the ground state solved
  void finalize_states();
  // ground-state averages and variances of every instance operator, computed
  // on first use (the reference fills them at the end of low_energy_states,
  // src_ed/model_instance.hpp:341-344; the values are the same)
  const std::vector<std::tuple<std::string, double, double>> &cluster_averages();
  void green_function_solve();
  // nz frequencies -> G (nz blocks n x n complex column-major, interleaved)
  void green_function(int nz, const double *z, bool spin_down, double *out);
  void self_energy(int nz, const double *z, bool spin_down, double *out);
  void hybridization(int nz, const double *z, bool spin_down, double *out);
  std::string gs_string() const;
  int gf_dim() const { return model->n_sites; }
  // spin block of the Green function for spin_down (checks the mixing)
  int gf_spin(bool spin_down) const;
  // one-body matrix of the spin block (L x L, row-major) from the instance's terms
  std::vector<double> one_body_matrix(bool spin_down) const;  // L x L, row-major, full matrix
  // the same with the imaginary parts (Hermitian)
  std::vector<cplx> one_body_matrix_c(bool spin_down) const;
  double sef_bath() const;
  // frequency integral of the Green function of the spin block, n x n
  // column-major (real Hamiltonians: real and symmetric)
  std::vector<cplx> gf_average(bool spin_down);
  // electrons per site from the diagonal of gf_average
  double gf_density();
  // trace of the self-energy at infinite frequency, both spins
  double tr_sigma_inf();
  // Lehmann representation of the response of one instance operator, one
  // part per low-energy state: poles e (excitation energies), residues v^2
  struct ChiPart {
    double weight = 0.0;
    std::vector<double> e, v;
  };
  std::vector<ChiPart> susceptibility_parts(const std::string &op);
  // set_hopping_matrix (src_ed/model_instance.hpp:424-474) for the spin block
  // of spin_down: cluster tc (ns x ns), bath tb (nb x nb), cluster-bath tcb
  // (ns x nb), column-major; with `diag` the bath block is diagonalized when it
  // is not diagonal already (eigenvalues ascending, tcb -> tcb S)
  struct Hopping {
    std::vector<double> tc, tb, tcb;
    // imaginary parts (complex Hilbert space; empty otherwise)
    std::vector<double> tci, tbi, tcbi;
  };
  Hopping hopping_blocks(bool spin_down, bool diag) const;
  // sum over pairs of low-energy states in the same sector of
  // |<psi_1|psi_2>|^2 w_1 w_2 (model_instance<T>::fidelity, :1261-1273)
  double fidelity(Instance &other);
  // reduced density matrix of the orbitals `sites` in the first low-energy
  // state (density_matrix_mixed, src_ed/model_instance.hpp:1547-1622): rho
  // (dA x dA column-major) over the subsystem states in order of first
  // appearance along the basis, and those states with orbital sites[i] moved
  // to bit i (up) and bit 32 + i (down)
  void density_matrix(const std::vector<int> &sites, std::vector<double> &rho, std::vector<uint64_t> &basis);
  // print_wavefunction (src_ed/model_instance.hpp:1473-1486, state<T>::write_wavefunction,
  // src_ed/state.hpp:121-131): per state a header line, then, for sectors of at
  // most max_dim_print states, |psi_i|^2, psi_i and the basis state
  std::string print_wavefunction();
};

// continued fractions of state `state`'s spin block (GF_method 'F'), fraction
// f = 2 q + eh (q: pair o1 >= o2 row by row, eh 0 electron / 1 hole), as
// evaluated: a_n and b_n (b_0 the weighted norm); lengths in len
void cf_export(Instance &I, int spin, int state, std::vector<std::vector<double>> &a,
               std::vector<std::vector<double>> &b);
// one state of energy E0 in sector `sec` whose spin block `spin` is the given
// set of continued fractions (same layout); the instance is then solved (no
// wavefunction: quantities that need one return ED_ERR_INVALID)
void cf_import(Instance &I, double E0, const Sector &sec, int spin, const std::vector<std::vector<double>> &a,
               const std::vector<std::vector<double>> &b);

// continued fractions of the ground state's spin block for the given orbital
// pairs only (o1 == o2: c_o1 |Omega>; else c_o1 |Omega> + c_o2 |Omega>), electron
// and hole, at most cap levels each (the first levels of the full fraction):
// fraction 2 k + eh for pair k, transformed as the instance's own
void cf_pairs(Instance &I, int spin, const std::vector<std::pair<int, int>> &pairs, int cap,
              std::vector<std::vector<double>> &a, std::vector<std::vector<double>> &b);

// the combined fraction of two matrix continued fractions (block size p, both
// non-empty) for n orbitals; empty when it cannot be formed
GFRep::MCF combine_mcf(const GFRep::MCF &e, const GFRep::MCF &h, int n, bool qr);

Instance &get_instance(int label);
void new_instance(const std::string &model, const std::map<std::string, double> &values, const std::string &sectors,
                  int label);
void erase_instance(int label);

// c_a / c^dagger_a (pm = -1 / +1) of spin-orbital a on a factorized sector vector
// (device); returns the target sector.  y must hold the target dimension.
Sector cdag_apply(int L, int pm, int a, const Sector &src, const double *x, double z, double *y);
// the half-basis maps of c_a (pm = -1) / c^dagger_a (pm = +1) from sector src:
// for the changed spin, map[t] = source half index (or -1) and sgn[t] the
// Jordan-Wigner sign of target half state t; for a down orbital, upar[u] =
// (-1)^popcount(up state u) (all up orbitals precede the down ones)
struct CdagMaps {
  Sector tgt;
  int spin = 0;
  uint32_t du_t = 0, dd_t = 0, du_s = 0, dd_s = 0;
  std::vector<int32_t> map;
  std::vector<int8_t> sgn, upar;
};
CdagMaps cdag_maps(int L, int pm, int a, const Sector &src);
// Green functions of many instances on one grid: out holds, in instance
// order, nz blocks of d x d complex (column-major) per instance; every
// continued-fraction representation is evaluated in a single launch
void green_function_batch(const std::vector<Instance *> &insts, int nz, const double *z, bool spin_down,
                          double *out);
size_t sector_dim(int L, const Sector &s);
// solved instances to / from HDF5 files (ed_hdf5.cpp; the HDF5 C library is
// loaded at run time)
std::string hdf5_version();
// the reference's Lanczos start vector (random() of src_util/vector_num.cpp:114-119):
// n N(0,1) draws of std::normal_distribution over std::default_random_engine(seed), normalized;
// one distribution object for the process, as there (its kept second draw carries over)
void random_start(size_t n, long seed, std::vector<double> &x);
void write_instance_hdf5(const std::string &filename, int label, const std::string &group);
void read_instance_hdf5(const std::string &filename, int label, const std::string &group);
bool sector_valid(int L, const Sector &s, int g = 1);
Sector shift_sector(const Sector &s, int pm, int spin);

}  // namespace ed
