/*
 * pyqcm_ed.h -- C-ABI of the MI355X exact-diagonalization impurity solver.
 *
 * Plain pointers and sizes only.  Every entry point returns a status code
 * (ED_OK on success); ed_last_error() gives the message of the last failure
 * on the calling thread.  No C++ exception crosses this boundary.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the pyqcm source tree).
 *
 * Conventions shared with the reference:
 *   - spin-orbital indices are 0-based here (src_python/qcm_ED_wrap.hpp:200-204
 *     converts the 1-based indices users type before calling ED::new_operator);
 *   - a < L is spin up orbital a, a >= L spin down orbital a - L, L = n_sites + n_bath;
 *   - sector strings "R<irrep>:N<N>:S<S>" (src_ed/sector.hpp:54-106);
 *   - complex outputs are interleaved (re, im) doubles; matrices are column-major
 *     (src_python/qcm_nb.hpp:27-41 copies matrix<T>::v, which is column-major).
 */
#ifndef PYQCM_ED_H
#define PYQCM_ED_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the failure paths where the reference calls qcm_ED_throw) */
enum {
  ED_OK = 0,
  ED_ERR_INVALID = 1,        /* bad argument, unknown name, malformed elements */
  ED_ERR_SECTOR = 2,         /* invalid sector string / sector (model_instance_base.cpp:99-124) */
  ED_ERR_NOT_CONVERGED = 3,  /* Lanczos / Davidson iteration limit (Lanczos.hpp:92, Davidson.hpp:228) */
  ED_ERR_QNORM = 4,          /* Q-matrix unitarity violated (Q_matrix.hpp:236-260) */
  ED_ERR_OOM = 5,            /* device or host allocation failed */
  ED_ERR_HIP = 6,            /* HIP runtime error */
  ED_ERR_NOT_FOUND = 7,      /* unknown model / instance label */
  ED_ERR_UNSUPPORTED = 8,    /* feature this library does not implement (see each entry point) */
  ED_ERR_INTERNAL = 9,
  ED_ERR_DEGENERATE = 10,    /* band Lanczos degenerate gap with no_degenerate_BL (Lanczos.hpp:368-370) */
  ED_ERR_CANCELLED = 11      /* ed_request_cancel was called (check_signals, src_python/common_Py.cpp:29-33) */
};

const char *ed_last_error(void);
/* Interrupt: the running (or next) solve stops at its next Lanczos / band-Lanczos /
   Davidson step or Green-function start vector with ED_ERR_CANCELLED, and the request
   is consumed.  Replaces check_signals (src_python/common_Py.cpp:29-33, polled in
   src_ed/Hamiltonian/Lanczos.hpp:66, :123, :172, :269, :473; model_instance.hpp:717,
   :855).  Async-signal-safe and callable from any thread. */
int ed_request_cancel(void);
int ed_version(char *buf, int cap);
int ed_device_count(int *n);

/* ---- global parameters: pyqcm.set_global_parameter (pyqcm/__init__.py:2259-2273,
        src_python/qcm_wrap.hpp:292-324; registry src_util/global_parameter.cpp:44-90) ---- */
int ed_set_global_bool(const char *name, int value);
int ed_set_global_int(const char *name, long value);
int ed_set_global_double(const char *name, double value);
int ed_set_global_char(const char *name, char value);
int ed_get_global_double(const char *name, double *value);
int ed_get_global_int(const char *name, long *value);

/* ---- model / operator / instance: namespace ED (src_ed/qcm_ED.hpp:34-339) ---- */
/* ED::new_model (src_ed/qcm_ED.cpp:43-59, symmetry_group.cpp:16-330): n_gen generators of
   n_sites + n_bath ints each, 0-based orbital images (with bath_irrep, the site images
   followed by the bath orbitals' phases, exp(2 pi i phase / g)), as the Python binding
   passes them (src_python/qcm_ED_wrap.hpp:145-160).  Groups with real characters only
   (generators of order <= 2; others ED_ERR_UNSUPPORTED).  States of sector R{r}:N:S are
   found in the full (N, S) sector projected on irrep r; Green functions are those of the
   full sectors, returned in the site basis or, with blocks, in the symmetric-orbital basis. */
int ed_new_model(const char *name, int n_sites, int n_bath, const int *generators, int n_gen, int bath_irrep);
/* ED::new_operator, real elements (src_ed/qcm_ED.cpp:74-90); type "one-body" or "interaction". */
int ed_new_operator(const char *model, const char *name, const char *type, int n,
                    const int *r, const int *c, const double *v);
/* ED::new_operator, complex elements (src_ed/qcm_ED.cpp:95-106; one_body_operator<Complex>,
   one_body_operator.hpp:59-76): type "one-body" only (others ED_ERR_UNSUPPORTED); v holds n
   complex values as (re, im) pairs, elements r <= c, the lower triangle their conjugates
   (an imaginary diagonal element is ED_ERR_INVALID: not Hermitian).  An instance with a
   nonzero value on such an operator works in a complex Hilbert space (qcm_ED.cpp:145-157):
   ground state (Lanczos, or Davidson for n_states > 1), cluster averages, Green function by
   band Lanczos (GF_method 'L') or continued fractions ('F'), Q-matrix, self-energy,
   hybridization, Potthoff functional; point groups, GF_method 'M', susceptibilities,
   fidelity, density matrices, wavefunction printing,
   continued-fraction export / import, batched solves and the HDF5 format return
   ED_ERR_UNSUPPORTED there.  Hamiltonian handles of such an instance act on split vectors
   [Re | Im] of twice the sector dimension. */
int ed_new_operator_complex(const char *model, const char *name, const char *type, int n,
                            const int *r, const int *c, const double *v);
/* ED::new_model_instance (src_ed/qcm_ED.cpp:125-158); sectors is slash-separated. */
int ed_new_model_instance(const char *model, int n_values, const char *const *names, const double *values,
                          const char *sectors, int label);
/* ED::erase_model_instance (src_ed/qcm_ED.cpp:68-71) */
int ed_erase_model_instance(int label);
/* ED::ground_state_solve (src_ed/qcm_ED.cpp:179-188) -> (E0, "sector:...") */
int ed_ground_state_solve(int label, double *E0, char *sector, int cap);
/* ED::Green_function_solve (src_ed/model_instance.hpp:517-555) */
int ed_green_function_solve(int label);
/* ED::Green_function_dimension */
int ed_green_function_dimension(int label, int *d);
/* ED::Green_function (src_ed/qcm_ED.cpp:230-236): nz frequencies (re,im pairs);
   out receives nz blocks of d x d complex, column-major. */
int ed_green_function(int label, int nz, const double *z, int spin_down, int blocks, double *out);
/* ED::Green_function for many instances on one grid (no reference counterpart: the
   reference evaluates one instance per call); out holds, in label order, nz blocks of
   d x d complex per instance (d = ed_green_function_dimension of that instance) */
int ed_green_function_batch(const int *labels, int n, int nz, const double *z, int spin_down, double *out);
/* ED::self_energy (src_ed/model_instance.hpp:683-700) */
int ed_self_energy(int label, int nz, const double *z, int spin_down, double *out);
/* ED::hybridization_function (src_ed/model_instance.hpp:1004-1018) */
int ed_hybridization_function(int label, int nz, const double *z, int spin_down, double *out);
/* ED::Potthoff_functional contribution E0 (src_ed/qcm_ED.cpp:330-335, bath part excluded) */
int ed_ground_state_energy(int label, double *E0);
/* ED::cluster_averages (src_ed/qcm_ED.cpp:192-204): for each instance operator
   (same order as ed_instance_parameters), the ground-state average and variance. */
int ed_cluster_averages(int label, int cap, double *ave, double *var, int *n);
/* ED::qmatrix (src_python/qcm_ED_wrap.hpp:339-349): the Lehmann representation of the
   first state's Green function (band-Lanczos method); e[M], v[L x M] complex column-major
   ((re, im) pairs), G(r, c) = sum_i v(c, i) conj(v(r, i)) / (z - e_i). */
int ed_qmatrix(int label, int spin_down, int cap, double *e, double *v, int *L, int *M);
/* number of low-energy states and their (energy, weight, sector) */
int ed_states(int label, int cap, double *energy, double *weight, int *n);
int ed_mixing(int label, int *mixing);

/* ---- batched solve of many instances (parameter points of a VCA / CDMFT sweep) ----
   Ground states of n instances (flags & 1: also their Green functions).  The
   reference solves one instance per call (ED::ground_state_solve,
   src_ed/qcm_ED.cpp; model_instance::low_energy_states,
   src_ed/model_instance.hpp:245-417); instances whose operators factorize by
   spin, with Ground_state_method 'L', n_states <= 1 and GF_method 'F' or 'L',
   run as one launch per stage, the others through their own solve. */
int ed_solve_batch(const int *labels, int n, int flags);

/* ---- inner plug-in: Hamiltonian<T> (src_ed/Hamiltonian/Hamiltonian_base.hpp:43-82),
        factory create_hamiltonian (src_ed/model_instance.hpp:192-238) ----
   format: 'G' (GPU factorized, CSR fallback), 'F' factorized, 'S'/'E' CSR, 'N' on-the-fly. */
int ed_hamiltonian_create(int label, const char *sector, char format, int64_t *handle);
int ed_hamiltonian_destroy(int64_t handle);
int ed_hamiltonian_dim(int64_t handle, int64_t *dim);
/* y += H x on host arrays (Hamiltonian::mult_add) */
int ed_hamiltonian_mult_add(int64_t handle, const double *x, double *y);
/* y = H x + beta y on device arrays, on the calling thread's stream (y not read when beta == 0); complete on return */
int ed_hamiltonian_apply_device(int64_t handle, const double *dx, double *dy, double beta);
/* Allocation-free variant of ed_hamiltonian_apply_device (no counterpart in the
   reference): y = H x + beta y on device arrays, and *d_dot = <x, y_new> written
   to device memory (the Lanczos alpha when beta == 0).  For callers that cannot
   allocate (HIP graph capture, tight serving loops): it calls no hipMalloc /
   hipFree; the per-block partials go to a fixed device scratch of 3.2 MB per
   device (8 192 partials + 393 216 two-pass slab partials), and a product needing
   more fails with ED_ERR_INVALID.  It never autotunes (the tuned schedule once
   ed_hamiltonian_autotune or ed_hamiltonian_apply_device has tuned the handle,
   else the default schedule).  Calls that overlap on one device race on the
   scratch. */
int ed_hamiltonian_apply_dot_noalloc(int64_t handle, const double *dx, double *dy, double beta, double *d_dot);
/* Hamiltonian::diag */
int ed_hamiltonian_diag(int64_t handle, double *d);
/* strict-lower CSR pattern of the 'S' format (Hamiltonian_CSR.hpp:60-81): Iptr[dim+1], J[nnz]
   (ED_ERR_UNSUPPORTED for the split plan of a complex Hilbert space) */
int ed_hamiltonian_csr_pattern(int64_t handle, int64_t *nnz, int32_t *Iptr, int32_t *J, double *v);
/* time nrep device applications y = H x (HIP events on the calling thread's stream) */
int ed_hamiltonian_time(int64_t handle, int nrep, double *ms_per_apply, int variant);
/* kernel selection knob for the factorized format: 0 auto (autotuned), 1 fused
   row kernel, 2 two-pass (row up-part + column-slab down-part) */
int ed_hamiltonian_set_variant(int64_t handle, int variant);
/* autotuner of the factorized schedule (no reference counterpart: the reference
   has one serial loop): times fused and two-pass x slab width {8,16,32} x
   streaming stores {off,on}; ms[k] per candidate in that order, best = chosen index */
int ed_hamiltonian_autotune(int64_t handle, int cap, double *ms, int *n, int *best);

/* ---- bases (src_ed/ED_basis.cpp:115-179) and sectors (src_ed/sector.hpp) ---- */
int ed_half_basis(int L, int N, uint32_t *out, int64_t cap, int64_t *dim);
int ed_sector_parse(const char *s, int *N, int *S, int *irrep);
/* the Lanczos start vector of the reference (random(), src_util/vector_num.cpp:114-119):
   n draws of std::normal_distribution<double>(0, 1) over std::default_random_engine(seed),
   normalized.  One distribution object serves the process, as the reference's normal_dis
   (src_ed/Hamiltonian/Lanczos.cpp:7): after an odd n its kept second draw starts the next
   vector.  The ground-state solve uses it when host_random_start is true (host only, no
   device call) */
int ed_random_start(int64_t n, int64_t seed, double *out);

/* ---- Krylov drivers on a Hamiltonian handle (host vectors in/out) ----
   Output capacity: `cap` counts coefficients (a: cap, b: cap + 1), poles (evals: cap,
   Q: p x cap) or floors (A, B: cap p x p blocks).  When the result is longer, nothing
   is written, the size needed is stored in *n / *M / *nfl and ED_ERR_INVALID is
   returned, so the caller can retry with that capacity. */
/* LanczosEigenvalue + LanczosEigenvector (src_ed/Hamiltonian/Lanczos.hpp:41-143) */
int ed_lanczos_gs(int64_t h, const double *start, double *E0, double *gs, int *niter);
/* LanczosGreen (src_ed/Hamiltonian/Lanczos.hpp:156-183): a[n], b[n+1] */
int ed_lanczos_green(int64_t h, const double *psi, int cap, double *a, double *b, int *n);
/* bandLanczos + build_Q_matrix (Lanczos.hpp:214-395, Hamiltonian_base.hpp:230-254) */
int ed_band_lanczos(int64_t h, int p, const double *phi, int cap, double *evals, double *Q, int *M);
/* Davidson (src_ed/Hamiltonian/Davidson.hpp:38-230) */
int ed_davidson(int64_t h, int K, double *evals, double *evecs, int *niter);
/* blockLanczos (src_ed/Hamiltonian/Lanczos.hpp:426-572), QR variant */
int ed_block_lanczos(int64_t h, int p, const double *q0, int M0, int cap, double *A, double *B, int *nfl);
/* lowest eigenpair of the symmetric tridiagonal (d[n], e[n-1]); z[n] may be NULL;
   guess (NaN: none) seeds a Rayleigh-quotient iteration, e.g. the previous Ritz value
   (EigensystemTridiagonal, src_ed/Hamiltonian/Lanczos.cpp:25-52, as used by
   LanczosEigenvalue / LanczosEigenvector for the Ritz pair) */
int ed_tridiag_lowest(int n, const double *d, const double *e, double guess, double *ev0, double *z);
/* c_a / c^dagger_a between factorized sectors (src_ed/model.cpp:155-182) */
int ed_create_or_destroy(int L, int pm, int a, int Nup, int Ndw, const double *x, double z, double *y);

/* ---- Green-function kernels on frequency grids ---- */
/* continued_fraction::evaluate (src_ed/continued_fraction.cpp:52-57): ncf fractions with
   offsets off[ncf+1] into a[], b[]; out[nz * ncf] complex, fraction-major per z. */
int ed_cf_eval(int ncf, const int *off, const double *a, const double *b, int nz, const double *z, double *out);
/* Q_matrix::Green_function + VDVH (src_ed/Q_matrix.hpp:134-151, src_util/VDVH_kernel.cpp:28-48) */
int ed_vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G);
/* matrix_continued_fraction::evaluate (src_ed/matrix_continued_fraction.hpp:139-182) */
int ed_mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q, int nz,
                const double *z, double *G);

/* ---- lattice sums over the superlattice Brillouin zone (CPT, VCA, CDMFT host) ----
   One spin block of the superlattice one-body part of a lattice model:
   V(k) = V0 + sum_e val_e exp(2 pi i k . R[nbr_e]) at (row_e, col_e), k in reduced
   coordinates of the superlattice, i.e. V = t(k) - t' as lattice_model_instance::set_V
   builds it (src_qcm/CPT.cpp:22-60; V0 = H - Hc, the elements are the operators'
   IGF_elem times their values).  Matrices are d x d complex, column-major, (re, im)
   pairs; d is the instance's Green-function dimension (at most 8). */
typedef struct {
  int d;
  const double *V0;  /* d x d complex */
  int ne;
  const int *row, *col, *nbr; /* ne each, 0-based */
  const double *val;          /* ne complex */
  int nnbr;                   /* at most 64 */
  const double *R;            /* nnbr x 3 */
} ed_superlattice;
/* wavevector grid: `dim` (0..3) axes of nk[a] points at (i + offset) / nk[a]: offset 0.5
   is projected_Green_function's grid (src_qcm/CPT.cpp:283-343), 0 is k_integral_grid's
   (src_util/integrate.cpp:177-235); dim 0 is the single point k = 0 */
typedef struct {
  int dim;
  int nk[3];
  double offset;
} ed_kgrid;
/* lattice_model_instance::projected_Green_function (src_qcm/CPT.cpp:283-343) for one
   cluster: for each of the nz frequencies, (1/nk) sum_k G (1 - V(k) G)^{-1} with G the
   instance's Green function; out as ed_green_function. */
int ed_cpt_projected_gf(int label, int nz, const double *z, int spin_down, const ed_superlattice *V,
                        const ed_kgrid *K, double *out);
/* the (frequency, k) grid path of lattice_model_instance::Potthoff_functional
   (src_qcm/lattice_model_instance.cpp:503-584) for one cluster: *integral = sum_n
   weights[n] (1/nk) sum_k I(i w[n], k), I = log|det(V(k) G - 1)| summed over the spin
   blocks (SEF_integrand_k, :624-650: twice the up block without spin mixing; Vdn is
   required with up/down mixing and ignored otherwise).  per_freq (may be NULL)
   receives the nw k-averages. */
int ed_sef_grid(int label, int nw, const double *w, const double *weights, const ed_superlattice *Vup,
                const ed_superlattice *Vdn, const ed_kgrid *K, double *integral, double *per_freq);
/* ED::Potthoff_functional (src_ed/qcm_ED.cpp:330-334): E0 plus the bath term */
int ed_potthoff_functional(int label, double *omega);

/* ---- one-body data, operators, overlaps ---- */
/* parametersC (src_python/qcm_ED_wrap.hpp:263-273): the instance's parameters in name
   order (names[k * name_len], values[k], k < min(*n, cap)) and its model's name. */
int ed_instance_parameters(int label, int cap, int name_len, char *names, double *values, int *n, char *model,
                           int model_cap);
/* ED::complex_HS (src_ed/qcm_ED.hpp:94): 1 when the instance works in a complex Hilbert
   space (a nonzero value on an operator defined with ed_new_operator_complex) */
int ed_complex_hs(int label, int *is_complex);
/* ED::hopping_matrix / hopping_matrix_full (src_ed/qcm_ED.cpp:269-284;
   model_instance.cpp:13-16, model_instance.hpp:1491-1522): the one-body matrix of the
   spin block; full = 0: the cluster block (n_sites); full = 1: cluster and bath
   (n_sites + n_bath), with diag = 1 the bath block diagonalized (set_hopping_matrix,
   model_instance.hpp:424-474: eigenvalues ascending unless the block is diagonal
   already; eigenvector signs are the library's own).  out: d x d complex column-major. */
int ed_hopping_matrix(int label, int spin_down, int diag, int full, double *out);
/* ED::interactions (model_instance.hpp:1527-1540): (r, c, U v) of every element of every
   interaction operator of the instance (instance parameters in name order); *n = count,
   the first cap written. */
int ed_interactions(int label, int cap, int *r, int *c, double *v, int *n);
/* ED::hybridization (model_instance.hpp:1455-1466): the diagonalized bath energies
   e[n_bath] and the cluster-bath matrix tcb (n_sites x n_bath complex column-major). */
int ed_hybridization_lehmann(int label, int spin_down, double *e, double *tcb);
/* ED::matrix_elements (src_ed/qcm_ED.cpp:442-450): the operator's type ("one_body" or
   "interaction") and its stored elements (0-based; one-body ones Hermitian-completed),
   values complex: v[2 k], v[2 k + 1] = (re, im) of element k. */
int ed_matrix_elements(const char *model, const char *op, char *type, int type_cap, int cap, int *r, int *c,
                       double *v, int *n);
/* ED::density_matrix (src_ed/qcm_ED.cpp:205-216; density_matrix_mixed,
   model_instance.hpp:1547-1622): reduced density matrix of the orbitals sites[0..nsites)
   in the first low-energy state; *dim_a = subsystem dimension; when cap >= *dim_a, rho
   receives dim_a x dim_a complex (column-major) and basis the subsystem states (orbital
   sites[i] at bit i for up, 32 + i for down), in order of first appearance along the basis.
   (For factorized Hamiltonian formats the reference returns an empty matrix; here the
   same result is returned for every format.) */
int ed_density_matrix(int label, const int *sites, int nsites, int cap, double *rho, uint64_t *basis, int *dim_a);
/* ED::print_wavefunction (src_ed/qcm_ED.cpp:431-440): the low-energy states as text
   (per state "state", then sector, energy and weight; for sectors of at most
   max_dim_print states one line per basis state: |psi|^2, psi, |up bits,down bits>).
   *len = length; the text and its terminating 0 are written when cap > *len. */
int ed_print_wavefunction(int label, char *buf, int cap, int *len);
/* ED::fidelity (src_ed/qcm_ED.cpp:360-368; model_instance.hpp:1261-1273): sum over pairs of
   low-energy states of the two instances in the same sector of |<psi_1|psi_2>|^2 w_1 w_2. */
int ed_fidelity(int label1, int label2, double *f);

/* ---- solved instances in HDF5 files (the reference's layout) ---- */
/* ED::write_instance_hdf5 (src_ed/qcm_ED.cpp:373-389; model_instance.hpp:1351-1375;
   state.hpp:52-72): the solved instance (Green function solved first) into `group`
   of `filename` (created if absent; an existing group of that name is replaced). */
int ed_write_instance_hdf5(const char *filename, int label, const char *group);
/* ED::read_instance_hdf5 (src_ed/qcm_ED.cpp:392-402; model_instance.hpp:1382-1420):
   the instance takes the file's low-energy states and Green functions and counts as
   solved (the format holds no wavefunction: cluster averages and susceptibilities of
   such an instance return ED_ERR_INVALID). */
int ed_read_instance_hdf5(const char *filename, int label, const char *group);
/* version of the HDF5 C library found at run time (ED_ERR_UNSUPPORTED if none) */
int ed_hdf5_version(char *buf, int cap);

/* ---- frequency integrals and responses of one instance ---- */
/* ED::Green_function_average (src_ed/qcm_ED.cpp:240-248; model_instance.hpp:641-652):
   the frequency-integrated Green function of the spin block, <c^dagger_b c_a>; out
   receives d x d complex, column-major, as one block of ed_green_function. */
int ed_green_function_average(int label, int spin_down, double *out);
/* ED::Green_function_density (src_ed/qcm_ED.cpp:251-257; model_instance.hpp:660-675):
   electrons per site from the trace of the average. */
int ed_green_function_density(int label, double *density);
/* ED::tr_sigma_inf (src_ed/qcm_ED.cpp:345-348; model_instance.hpp:1278-1312): the
   trace of the self-energy extrapolated to infinite frequency, both spins. */
int ed_tr_sigma_inf(int label, double *value);
/* ED::susceptibility (src_ed/qcm_ED.cpp:307-315; model_instance.hpp:1022-1058): the
   dynamic response sum_n |<n|O|Omega>|^2 / (w - E_n + E_Omega), weighted over the
   low-energy states, of instance operator `op` at nw complex frequencies (re, im
   pairs); chi receives nw complex values. */
int ed_susceptibility(int label, const char *op, int nw, const double *w, double *chi);
/* ED::susceptibility_poles (src_ed/qcm_ED.cpp:319-326; model_instance.hpp:1066-1108):
   the same response as *n (pole, residue) pairs, written to poles[2k], poles[2k+1]
   for k < cap: poles above 1e-8, residues above 1e-6, poles closer than 1e-4 merged. */
int ed_susceptibility_poles(int label, const char *op, int cap, double *poles, int *n);

/* Continued fractions of low-energy state `state` of a solved instance
   (GF_method 'F'; continued_fraction_set, src_ed/continued_fraction_set.cpp:12-98):
   fraction f = 2 q + eh, q the orbital pair o1 >= o2 row by row (q = o1 (o1 + 1) / 2
   + o2), eh 0 for the electron and 1 for the hole part, as evaluated (a_n shifted
   by the state energy, b_n squared, b_0 the weighted norm).  a[f * cap + n] and
   b[f * cap + n] for n < min(len[f], cap), zero beyond; len[f] the full length;
   *nfrac = n_sites (n_sites + 1).  Serves the learned surrogate's training targets. */
int ed_continued_fractions(int label, int spin_down, int state, int cap, double *a, double *b, int *len,
                           int *nfrac);
/* The same fractions for a list of orbital pairs only (o1[k] == o2[k]: c_o1 |Omega>;
   else c_o1 |Omega> + c_o2 |Omega>, as build_cf, src_ed/model_instance.cpp:35-100),
   from the ground state, with at most cap levels each (the first levels of the
   unbounded fraction; LanczosGreen, Lanczos.hpp:156-183): fraction 2 k + eh for
   pair k, a and b as above ([2 npairs][cap]).  Any GF_method; the instance's own
   Green function is not built. */
int ed_continued_fractions_pairs(int label, int spin_down, int npairs, const int *o1, const int *o2, int cap,
                                 double *a, double *b, int *len);
/* The inverse (no reference counterpart; the HDF5 reader does the same from a
   file, model_instance.hpp:1351-1420): the instance becomes solved, with one
   state of energy E0 in `sector` whose spin block holds these fractions (same
   layout, nfrac = n_sites (n_sites + 1)); call once per spin block.  Used to
   evaluate surrogate predictions through the Green-function, CPT and Potthoff
   entry points; quantities that need a wavefunction then return ED_ERR_INVALID. */
int ed_set_continued_fractions(int label, double E0, const char *sector, int spin_down, int nfrac, int cap,
                               const double *a, const double *b, const int *len);

/* ED::get_combined_mcf (src_ed/qcm_ED.hpp:319-339, qcm_ED.cpp:457-537): the combined
   matrix continued fraction (G_e + G_h^T in one fraction, built when combine_mcf is
   set, GF_method 'M') of the first low-energy state: W (p x d) and nfl floors A_j,
   B_j (p x p), all column-major, A / B at floor offsets j p p; *p and *nfl report the
   sizes (nothing written when nfl > cap). */
int ed_combined_mcf(int label, int spin_down, int cap, double *W, double *A, double *B, int *p, int *nfl);

/* ---- device memory helpers for callers that keep vectors resident ---- */
int ed_device_alloc(int64_t bytes, void **ptr);
int ed_device_free(void *ptr, int64_t bytes);
int ed_device_copy(void *dst, const void *src, int64_t bytes, int kind /* 0 H2D, 1 D2H, 2 D2D */);
int ed_device_sync(void);
/* device blocks the library frees are kept for reuse (a solve's Krylov basis
   included, up to ~235 GB on the 4x4 sector); this returns them to the driver,
   e.g. before another library in the process allocates */
int ed_release_cache(void);

/* ---- learned surrogate (no counterpart in the reference: pyqcm has no learned
        model).  A GELU MLP, n_in -> depth hidden layers of `width` -> n_out,
        trained and evaluated on the device: bf16 MFMA GEMMs with fp32
        accumulation, fp32 master weights and Adam.  It maps cluster parameters
        to E0 and truncated continued-fraction coefficients; the VCA / CDMFT
        drivers use its predictions as candidates that the exact solver
        verifies.  Inputs x are row-major n x n_in, outputs n x n_out, in
        physical units (the scaling below is applied inside).  Weight order:
        for each layer, W (out x in, row-major) then b (out). ---- */
int ed_surrogate_create(int n_in, int n_out, int width, int depth, int64_t seed, int64_t *handle);
int ed_surrogate_destroy(int64_t handle);
int ed_surrogate_n_params(int64_t handle, int64_t *n);
/* x' = (x - x_mean) / x_std, y' = (y - y_mean) / y_std; the loss is
   sum_o loss_weight[o] (y'_o - t'_o)^2 averaged over samples.  Null pointers
   leave a part unchanged. */
int ed_surrogate_set_scaling(int64_t handle, const double *x_mean, const double *x_std, const double *y_mean,
                             const double *y_std, const double *loss_weight);
int ed_surrogate_set_optimizer(int64_t handle, double lr, double beta1, double beta2, double eps);
int ed_surrogate_get_weights(int64_t handle, double *w);
/* sets the fp32 master weights (and their bf16 copies) and restarts Adam */
int ed_surrogate_set_weights(int64_t handle, const double *w);
int ed_surrogate_predict(int64_t handle, int n, const double *x, double *y);
/* loss of one batch (normalized units) and its gradient in the weight order
   (no update) */
int ed_surrogate_loss_grad(int64_t handle, int n, const double *x, const double *y, double *loss, double *grad);
/* `epochs` passes over the n samples in shuffled batches of `batch` (full
   batches only; shuffles from `seed`), one Adam step per batch; epoch_loss
   receives the mean training loss of each epoch; ms_per_step the device time
   of one step (gather, forward, backward, Adam) */
int ed_surrogate_train(int64_t handle, int n, const double *x, const double *y, int epochs, int batch,
                       int64_t seed, double *epoch_loss, double *ms_per_step);
/* device time of the forward pass and of forward + backward (weights
   unchanged) on `batch` resident rows, averaged over nrep */
int ed_surrogate_time(int64_t handle, int batch, int nrep, double *ms_forward, double *ms_forward_backward);

#ifdef __cplusplus
}
#endif
#endif
