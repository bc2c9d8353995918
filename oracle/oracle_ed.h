/*
 * oracle_ed.h -- CPU FP64 restatement of pyqcm's exact-diagonalization path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under oracle/ is part of the product:
 * only tests/, __graft_entry__.smoke() and bench.py (its cpu_baseline leg)
 * may load this library, and only for reference values / the CPU baseline.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Conventions (reference src_ed/binary_state.hpp:60-63, src_ed/ED_basis.cpp:142-179):
 *   - spin-orbital a in [0, 2L): a < L is spin-up orbital a, a >= L is spin-down a-L;
 *   - half basis = ascending integers with popcount N (ED_basis.cpp:115-124);
 *   - factorized label I = i_up + dim_up * i_dn (ED_basis.cpp:157-165);
 *   - one-body elements are Hermitian-completed (one_body_operator.hpp:67-76);
 *   - interaction elements (r,c,v) mean v n_r n_c (interaction_operator.hpp:43-48).
 */
#ifndef ORACLE_ED_H
#define ORACLE_ED_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ONE_BODY = 0, ORC_INTERACTION = 1 };

/* Flat model description: term t has elements [off[t], off[t+1]) of (r, c, v),
   0-based spin-orbital indices exactly as ED::new_operator receives them. */
typedef struct {
  int L;                /* orbitals per spin (cluster sites + bath) */
  int n_terms;
  const int *type;      /* ORC_ONE_BODY / ORC_INTERACTION */
  const double *value;  /* coefficient of each term in H */
  const int *off;       /* n_terms + 1 offsets */
  const int *r;
  const int *c;
  const double *v;
} orc_model;

/* ---- bases and sectors ---- */
size_t orc_half_basis(int L, int N, uint32_t *out);
uint32_t orc_half_index(const uint32_t *bin, uint32_t dim, uint32_t b);
int orc_sector_parse(const char *s, int *N, int *S, int *irrep);
int orc_sector_name(int N, int S, int irrep, char *buf, int cap);

/* ---- per-term half tables (HS_half_operator) ---- */
/* Counts (pairs, diagonal entries) of term t for spin (0 up, 1 down) with N electrons. */
int orc_half_table(const orc_model *m, int t, int spin, int N,
                   size_t *n_pairs, size_t *n_diag,
                   uint32_t *pI, uint32_t *pJ, double *pv, int *pgroup,
                   uint32_t *dI, double *dv);

/* ---- Hamiltonian-vector products: y += H x ---- */
int orc_hv_factorized(const orc_model *m, int Nup, int Ndw, const double *x, double *y);
int orc_hv_factorized_omp(const orc_model *m, int Nup, int Ndw, const double *x, double *y, int nrep);
int orc_hv_factorized_omp_timed(const orc_model *m, int Nup, int Ndw, const double *x, double *y, int nrep,
                                double *ms);
int orc_diag(const orc_model *m, int Nup, int Ndw, double *d);
int orc_csr_build(const orc_model *m, int Nup, int Ndw, size_t *nnz,
                  int32_t *Iptr, int32_t *J, double *v, double *diag);
int orc_hv_csr(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v,
               const double *diag, const double *x, double *y);
int orc_hv_csr_full_omp(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v,
                        const double *diag, const double *x, double *y, int nrep);
int orc_hv_csr_full_omp_timed(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v,
                              const double *diag, const double *x, double *y, int nrep, double *ms);
int orc_hv_otf(const orc_model *m, int Nup, int Ndw, const double *x, double *y);

/* ---- c / c^dagger between sectors (model::create_or_destroy, factorized) ---- */
int orc_create_or_destroy(int L, int pm, int a, int Nup, int Ndw,
                          const double *x, double z, double *y);

/* ---- dense helpers ---- */
int orc_tridiag_eig(int n, const double *d, const double *e, double *w, double *z);
int orc_symeig(int n, const double *A, double *w, double *V);

/* ---- Krylov solvers (all FP64) ---- */
int orc_lanczos_gs(const orc_model *m, int Nup, int Ndw, const double *start,
                   int max_iter, double accur, double *E0, double *gs, int *niter,
                   double *alpha, double *beta);
int orc_lanczos_green(const orc_model *m, int Nup, int Ndw, const double *psi,
                      double accur_cf, int max_iter_cf, double *a, double *b, int *n);
double orc_cf_eval_1(int n, const double *a, const double *b, double zr, double zi, double *gi);
int orc_cf_eval(int n, const double *a, const double *b, int nz, const double *z, double *g);
int orc_band_lanczos(const orc_model *m, int Nup, int Ndw, int p, const double *phi,
                     int M_in, int max_iter_bl, double accur_deflation, double accur_bl,
                     int *M_out, double *evals, double *Q /* p x M, column-major */);
int orc_qmatrix_streamline(int L, int M, const double *e, const double *v,
                           double wtol, double vtol, double *e2, double *v2);
double orc_qmatrix_norm_error(int L, int M, const double *v);
int orc_vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G);
int orc_davidson(const orc_model *m, int Nup, int Ndw, int K, double accuracy, int max_iter,
                 double *evals, double *evecs, int *niter);
int orc_block_lanczos(const orc_model *m, int Nup, int Ndw, int p, const double *q0,
                      int M0, int max_iter_bl, double accur_deflation, double accur_bl,
                      int *n_floors, double *A, double *B);
int orc_block_lanczos_v(const orc_model *m, int Nup, int Ndw, int p, const double *q0, int M0, int max_iter_bl,
                        double accur_deflation, double accur_bl, int qr, int *n_floors, double *A, double *B);
int orc_mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q,
                 int nz, const double *z, double *G);

/* ---- whole ED pipeline on one instance (ground state + GF representation) ---- */
int orc_openmp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
