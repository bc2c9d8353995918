// ed_kernels.hpp -- host-side launchers of the HIP kernels (ed_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ed {

// One spin's merged, symmetric one-body operator on a half basis (both
// triangles of HS_half_operator, src_ed/Operators/HS_half_operator.hpp:17-70,
// summed over terms with their coefficients).  Entries are packed as
// col | (value_id << colbits) when the value table is small enough,
// otherwise val[] holds one double per entry.
// Sliced-ELL form of a half operator: slices of 64 lanes, lane l of slice s
// holding state row[64 s + l] (or 64 s + l when row is null; dim marks an
// empty lane).  A slice is padded to its longest row, rounded up to whole
// lane groups of 4 entries; groups are stored group-major, lane-minor, so a
// wave's group load is one contiguous run.  Padding entries use column dim,
// a zero slot after the LDS-staged row.
struct SellTable {
  uint32_t nslice = 0;
  const uint32_t *off = nullptr;   // nslice + 1 entry offsets
  const uint32_t *row = nullptr;   // nslice * 64 states, or null (identity)
  const uint32_t *ent = nullptr;   // packed col | value id << colbits (packed), else col
  const double *val = nullptr;     // explicit values (packed == 0)
  const uint16_t *e16 = nullptr;   // 16-bit entries (mode16)
};

struct DevHalf {
  uint32_t dim = 0;
  const uint32_t *ptr = nullptr;   // dim + 1
  const uint32_t *ent = nullptr;   // packed entries
  const double *val = nullptr;     // explicit values (packed == 0) or value table (packed == 1)
  const double *diag = nullptr;    // one-body + same-spin interaction diagonal, per half state
  const uint32_t *states = nullptr;  // half-basis bit patterns
  int packed = 1;
  int colbits = 0;
  uint32_t nval = 0;               // size of the value table (packed == 1)
  // the same operator as sliced-ELL tables (see SellTable): [0] in state
  // order (the fused kernel's down part needs lane u on column u), [1] with
  // the states sorted by row length, which removes most of the padding
  SellTable sell[2];
  // single-magnitude operators (every value is +-t0, dim < 2^15): 16-bit
  // entries col | sign << 15 in SellTable::e16
  int mode16 = 0;
  double t0 = 0.0;
  // the states in breadth-first order of the operator's hopping graph
  const uint32_t *order = nullptr;
};

// slab configuration of the two-pass down kernel (autotuned per plan)
struct SlabConfig {
  int sw = 32;        // columns per slab (16, 32, 64; 4, 8, 16 with lds)
  int nt = 0;         // non-temporal y / own-x accesses (keep L2 for the slab)
  int nxcd = 8;       // 8: slabs partitioned over XCDs (block b -> XCD b % 8); 1: one shared sweep
  int lds = 0;        // 1: each slab held in LDS (dim_dn x sw doubles <= kDownLdsMax)
};
// largest LDS slab of the down pass (two workgroups per CU)
constexpr size_t kDownLdsMax = 64 * 1024;

// Factorized H on X[i_dn][i_up]: H = D + (1 x K_up) + (K_dn x 1).
struct FactArgs {
  DevHalf up, dn;
  int diag_mode = 0;      // 0 streamed full diagonal, 1 per-site cross U_i, 2 uniform cross U
  const double *dg = nullptr;  // full diagonal (diag_mode 0)
  double Uc[32];          // cross (up-down same-site) coefficients (diag_mode 1, 2)
  int block = 0;          // row-kernel block size: 0 auto (256 below 2048 up states, else 1024), 256, 1024
  int row_order = 0;      // fused row kernel: 0 rows in label order, 1 in dn.order (neighbour rows reused sooner)
};

// HV_TWO_PASS_LDS: the two-pass schedule with LDS-resident slabs (the widest
// of 16, 8, 4 columns that fits); an error where no slab fits
enum HvVariant { HV_AUTO = 0, HV_FUSED = 1, HV_TWO_PASS = 2, HV_TWO_PASS_LDS = 3 };
// aux / aux_cap: device array for the two-pass schedule's per-slab-block
// partials when they exceed partials_cap() (nullptr: a library-owned array,
// allocated on first need; otherwise more than aux_cap of them is an error)
int hv_factorized_cfg(const FactArgs &a, const double *x, double *y, double ax, double beta, double *dot_partials,
                      int variant, const SlabConfig &sc, hipStream_t s, double *aux = nullptr, size_t aux_cap = 0);

// y = ax H x + beta y (y not read when beta == 0); when dot_partials != nullptr,
// per-block partial sums of <y_new, x> are written there (nblocks returned).
int hv_factorized(const FactArgs &a, const double *x, double *y, double ax, double beta, double *dot_partials,
                  int variant, hipStream_t s);
int hv_factorized_max_partials(const FactArgs &a, int variant);
void fact_full_diag(const FactArgs &a, double *d, hipStream_t s);

// full symmetric CSR (row pointers 64-bit safe)
struct CsrArgs {
  uint64_t n = 0;
  const int64_t *ptr = nullptr;
  const uint32_t *col = nullptr;
  const double *val = nullptr;
  const double *diag = nullptr;
};
int hv_csr(const CsrArgs &a, const double *x, double *y, double ax, double beta, double *dot_partials,
           hipStream_t s);
int hv_csr_max_partials(const CsrArgs &a);

// on-the-fly (no stored matrix): one-body hop list + interaction diagonal
struct OtfHop { int r, c; double v; };   // y[I] += v * sign * x[J], J = I with electron moved r -> c
struct OtfArgs {
  int L = 0;
  uint32_t du = 0, dd = 0;
  const uint32_t *up_states = nullptr, *dn_states = nullptr;
  int nhop = 0;
  const OtfHop *hops = nullptr;    // device array
  const double *dg = nullptr;      // full diagonal
  const uint32_t *binom = nullptr; // (L+1) x (L+1) binomials for ranking
};
int hv_otf(const OtfArgs &a, const double *x, double *y, double ax, double beta, double *dot_partials,
           hipStream_t s);
int hv_otf_max_partials(const OtfArgs &a);

// ---- vector kernels (FP64, wave64 reductions, per-block partials) ----
int partials_cap();

// Fixed device scratch of the allocation-free entry points: a file-scope
// device array (one copy per device, kept until the process ends) whose
// address is looked up once per device.  part holds partials_cap() per-block
// partials, aux the two-pass schedule's larger partial set (aux_cap doubles).
struct NoAllocScratch {
  double *part = nullptr;
  double *aux = nullptr;
  size_t aux_cap = 0;
};
NoAllocScratch noalloc_scratch();
// out[0] = sum of part[0 .. nb) in a fixed order (one block; out on the device)
void fold_partials_device(const double *part, int nb, double *out, hipStream_t s);
// out[k * nb + block] = partial of <a, b_k>, k < m (m <= 16). returns nb
int multi_dot(size_t n, const double *a, int m, const double *const *b, double *partials, hipStream_t s);
// x *= xs (written back when xs != 1), then y_k += c_k x for k < m; if
// norm_partials, partials of ||y_0||^2 after the update. returns nb
int multi_axpy_src(size_t n, double *x, int m, double *const *y, const double *c, double *norm_partials,
                   hipStream_t s, double xs = 1.0);
// One band-Lanczos step's vector updates in one pass (no deflated vectors):
//   x *= xs;  a_k += ca_k x (k < na);  t = st t + sum_b cb_b w_b + cx x;
// and the dots the next step needs with c = a_0 (t when na == 0), as partials
// (row r at part + r * nb): <c, c>, <c, a_k> (1 <= k < na), <c, t> (na > 0),
// <c, w_b> (b < nw), <c, x>.  na, nw <= kBandStepMax; returns nb (0 when the
// counts or the vector block size are outside what the kernel is built for)
// Device-driven steps (fold given): the kernel derives the step's scalars
// itself from partials on the device, in every workgroup: the dots of v_j
// (rows of dnb: <v_j, v_j>, the na ahead, the nw behind from row rb) and
// <H v_j, v_j> (hnb partials); step j's window starts at k0 (behind) with pc
// vectors per band.  Workgroup 0 writes the host's record rec[7]: |v_j|,
// zz[0 .. 3] (the ahead coefficients), t(j, j), deflated; |v_j| < accur
// (a deflation) sets *stop, and a set *stop makes later steps no-ops.
constexpr int kBandStepMax = 4;
struct BandFold {
  const double *dpart = nullptr;
  int dnb = 0, rb = 0;
  const double *hpart = nullptr;
  int hnb = 0;
  int j = 0, pc = 0, k0 = 0;
  double accur = 0.0;
  double *rec = nullptr;
  int *stop = nullptr;
};
int band_step(size_t n, double *x, double xs, int na, double *const *a, const double *ca, double *t, double st,
              int nw, const double *const *w, const double *cb, double cx, double *part, hipStream_t s,
              const BandFold *fold = nullptr);
// t = st * t + sum_k c_k b_k ; optional partials of ||t||^2 and <t, w>. returns nb
int combine(size_t n, double *t, double st, int m, const double *const *b, const double *c, double *norm_partials,
            const double *w, double *dot_partials, hipStream_t s);
void scale(size_t n, double *x, double a, hipStream_t s);
void fill(size_t n, double *x, double v, hipStream_t s);
// out[r] = sum_{i < nb} part[r * stride + i], r < m, in one block (fixed
// order); out is host-mapped pinned memory; then *flag = seq (release,
// system scope) for the host to poll
void fold_rows_flag(const double *part, int nb, int m, size_t stride, double *out, uint64_t *flag, uint64_t seq,
                    hipStream_t s);
// x[i] ~ N(0, 1), deterministic in (seed, i)
void random_normal(size_t n, uint64_t seed, double *x, hipStream_t s);
void set_unit(double *x, size_t i, hipStream_t s);
// r_i /= (mu - d_i + tiny)
void precond(size_t n, double *r, const double *d, double mu, double tiny, hipStream_t s);
// p += c1 r ; g += c2 p
void replay_step(size_t n, double *p, const double *r, double *g, double c1, double c2, hipStream_t s);

// c_a / c^dagger_a gather between factorized sectors: y[t] += z * sgn[t_half] * x[src(t)]
struct CdagArgs {
  int spin;                 // 0: up orbital changes, 1: down orbital
  uint32_t du_t, dd_t;      // target half dims
  uint32_t du_s;            // source up dim
  const int32_t *map;       // target half index -> source half index or -1 (changed spin)
  const int8_t *sgn;        // +-1 per target half index of the changed spin (same-spin bits below a)
  const int8_t *upar;       // +-1 per up index (parity of up electrons) when spin == 1, else null
  double z;
};
void cdag_gather(const CdagArgs &a, const double *x, double *y, hipStream_t s);

// ---- device-driven Lanczos steps (no host decision inside a step) ----
// The recurrence's scalars stay on the device: a[k] at coef[k], b[k] at
// coef[cap + k], with k = *base + j for step j of a block (j is a launch
// constant, so a captured block replays unchanged; lz_next moves *base on).
// Vectors are kept unnormalized (v_k = U_k / n_k, n_k = |U_k| = b[k-1]); step k
// is W = H C (C = U_k, dot partials in part), then one pass (lz_step) that
// folds a[k] and n_k, forms U_{k+1} over W from W, C and P = U_{k-1}, copies
// v_k to slots[k] when that entry of the device slot table is not null, and
// leaves the norm partials of U_{k+1} in part + (1 + (k+1) % 2) partials_cap();
// lz_next folds the last step's b[k] and moves *base on.
struct LanczosDev {
  size_t n = 0;
  int cap = 0;
  double *coef = nullptr;   // 2 cap
  int *base = nullptr;
  double *part = nullptr;   // 3 partials_cap()
  double *const *slots = nullptr;  // device table of nslots vector pointers (or null)
  int nslots = 0;
};
void lz_step(const LanczosDev &L, double *W, const double *C, const double *P, int j, int dnb, hipStream_t s);
void lz_next(const LanczosDev &L, int steps, hipStream_t s);

// projection on an irreducible representation of the cluster's point group, in
// a factorized sector: y[J] = sum_k c_k s_k(J) x[inv_k(J)] with inv_k(J) the
// state that element k maps to J and s_k its sign (per spin: J = j_up + du j_dn)
constexpr int kMaxGroup = 64;
struct ProjArgs {
  int g = 1;
  uint32_t du = 0, dd = 0;
  const int32_t *inv_up = nullptr, *inv_dn = nullptr;  // g x du, g x dd
  const int8_t *sg_up = nullptr, *sg_dn = nullptr;
  double c[kMaxGroup];
};
void group_project(const ProjArgs &a, const double *x, double *y, hipStream_t s);

// ---- Green-function kernels ----
// continued fractions: out[iz * ncf + f] = G_f(z_iz)
void cf_eval(int ncf, const int *off, const double *a, const double *b, int nz, const double *z, double *out,
             hipStream_t s);
// G matrices of groups of states from their evaluated fractions (vals, as
// cf_eval writes them): group g owns states [pbeg[g], pbeg[g+1]); state p's
// fractions start at frac0[p], two per orbital pair q = o1 (o1 + 1) / 2 + o2
// (o1 >= o2), electron then hole.  G_aa = F_aa, G_ab = (F_ab - F_aa - F_bb) / 2
// with F = electron + hole, summed over the group's states, written at
// G + at[g] (doubles) as nz column-major n[g] x n[g] complex matrices
// cx[g] != 0: group g's representations also hold, after their n (n + 1) / 2
// pairs, the fractions of the pairs' imaginary combinations (a > b, index
// a (a - 1) / 2 + b), assembled as continued_fraction_set::Green_function does
// for a complex Hilbert space
void cf_assemble(int ngroups, const int *n, const uint64_t *at, const int *pbeg, const int *frac0, const int *cx,
                 int ncf, int nz, const double *vals, double *G, hipStream_t s);
// G[b + a L] (nz blocks) = sum_i V(a, i) conj(V(b, i)) / (z - e_i), V real or
// with imaginary parts Vi
void vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G, hipStream_t s,
          const double *Vi = nullptr);
// matrix continued fraction per z (p <= 16)
void mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q, int nz, const double *z,
              double *G, hipStream_t s);

}  // namespace ed
