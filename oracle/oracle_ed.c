/*
 * oracle_ed.c -- CPU FP64 restatement of pyqcm's exact-diagonalization path.
 *
 * TEST INFRASTRUCTURE ONLY (reference values for the tests, and the CPU baseline).  The product
 * (pyqcm_amd/) never links, loads or calls anything here.
 *
 * Written from the reference's behaviour; each function names the
 * reference file:line it restates.  Dense eigensolvers are textbook
 * Householder tridiagonalisation + implicit QL (EISPACK tred2/tql2 method),
 * standing in for LAPACK dstev / Eigen SelfAdjointEigenSolver; they are
 * pinned by residual/orthogonality tests, not by matching library output.
 */
#include "oracle_ed.h"

#include <complex.h>
#undef I
#define CI _Complex_I
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SMALL_VALUE 1.0e-12 /* src_util/vector_num.hpp:23 */
#define TINY_DAV 1e-14      /* src_ed/Hamiltonian/Davidson.hpp:15 */

/* ------------------------------------------------------------------------ */
/* prepared model: Hermitian-completed elements per term                     */
/* ------------------------------------------------------------------------ */
typedef struct { int r, c; double v; } oelem;
typedef struct { int type; double value; int n; oelem *e; } oterm;
typedef struct { int L; int nt; oterm *t; } pmodel;

/* one_body_operator ctor (src_ed/Operators/one_body_operator.hpp:59-76):
   r<c -> (r,c,v),(c,r,v); r==c -> (r,r,v); r>c dropped.
   interaction_operator ctor (src_ed/Operators/interaction_operator.hpp:43-48):
   (min, max, v). */
static pmodel *prepare(const orc_model *m) {
  pmodel *p = (pmodel *)calloc(1, sizeof(pmodel));
  p->L = m->L;
  p->nt = m->n_terms;
  p->t = (oterm *)calloc((size_t)m->n_terms, sizeof(oterm));
  for (int t = 0; t < m->n_terms; ++t) {
    int n0 = m->off[t], n1 = m->off[t + 1];
    oterm *T = &p->t[t];
    T->type = m->type[t];
    T->value = m->value[t];
    T->e = (oelem *)malloc(sizeof(oelem) * (size_t)(2 * (n1 - n0) + 1));
    T->n = 0;
    for (int k = n0; k < n1; ++k) {
      int r = m->r[k], c = m->c[k];
      double v = m->v[k];
      if (T->type == ORC_ONE_BODY) {
        if (r < c) {
          T->e[T->n++] = (oelem){r, c, v};
          T->e[T->n++] = (oelem){c, r, v};
        } else if (r == c) {
          T->e[T->n++] = (oelem){r, r, v};
        }
      } else {
        if (r <= c) T->e[T->n++] = (oelem){r, c, v};
        else T->e[T->n++] = (oelem){c, r, v};
      }
    }
  }
  return p;
}

static void release(pmodel *p) {
  if (!p) return;
  for (int t = 0; t < p->nt; ++t) free(p->t[t].e);
  free(p->t);
  free(p);
}

static inline int popc64(uint64_t x) { return __builtin_popcountll(x); }

/* binary_state::mask (src_ed/binary_state.hpp:60-63) */
static inline uint64_t bmask(int i, int n) {
  return (i >= n) ? ((uint64_t)1 << (i - n + 32)) : ((uint64_t)1 << i);
}

/* binary_state::one_body + interphase (src_ed/binary_state.cpp:23-28,
   src_ed/binary_state.hpp:42-47): move the bit at i to j; returns the sign,
   0 when Pauli-forbidden. */
static inline int one_body(uint64_t *b, uint64_t i, uint64_t j) {
  if ((*b & j) || !(*b & i)) return 0;
  *b = *b - i + j;
  if (i == j) return 1;
  uint64_t inside = (i > j) ? (i - (j << 1)) : (j - (i << 1));
  return (popc64(inside & *b) & 1) ? -1 : 1;
}

/* ------------------------------------------------------------------------ */
/* bases (src_ed/ED_basis.cpp:115-135)                                       */
/* ------------------------------------------------------------------------ */
size_t orc_half_basis(int L, int N, uint32_t *out) {
  size_t n = 0;
  if (L < 0 || L > 32 || N < 0 || N > L) return 0;
  uint64_t maxb = (uint64_t)1 << L;
  for (uint64_t b = 0; b < maxb; ++b) {
    if (popc64(b) != N) continue;
    if (out) out[n] = (uint32_t)b;
    ++n;
  }
  return n;
}

uint32_t orc_half_index(const uint32_t *bin, uint32_t dim, uint32_t b) {
  uint32_t lo = 0, hi = dim;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (bin[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  if (lo < dim && bin[lo] == b) return lo;
  return dim;
}

typedef struct { int L, Nup, Ndw; uint32_t du, dd; uint32_t *up, *dn; size_t dim; } fbasis;

static int fb_make(fbasis *B, int L, int Nup, int Ndw) {
  memset(B, 0, sizeof(*B));
  B->L = L; B->Nup = Nup; B->Ndw = Ndw;
  if (Nup < 0 || Ndw < 0 || Nup > L || Ndw > L) return -1;
  B->du = (uint32_t)orc_half_basis(L, Nup, NULL);
  B->dd = (uint32_t)orc_half_basis(L, Ndw, NULL);
  B->up = (uint32_t *)malloc(sizeof(uint32_t) * (B->du ? B->du : 1));
  B->dn = (uint32_t *)malloc(sizeof(uint32_t) * (B->dd ? B->dd : 1));
  orc_half_basis(L, Nup, B->up);
  orc_half_basis(L, Ndw, B->dn);
  B->dim = (size_t)B->du * B->dd;
  return 0;
}
static void fb_free(fbasis *B) { free(B->up); free(B->dn); }

/* mixed-basis word for label I (g = 1): up bits low, down bits high
   (src_ed/ED_basis.cpp:316-359; sorted order equals the factorized label). */
static inline uint64_t fb_word(const fbasis *B, size_t I) {
  return (uint64_t)B->up[I % B->du] | ((uint64_t)B->dn[I / B->du] << 32);
}
static inline size_t fb_index(const fbasis *B, uint64_t w) {
  uint32_t iu = orc_half_index(B->up, B->du, (uint32_t)w);
  uint32_t id = orc_half_index(B->dn, B->dd, (uint32_t)(w >> 32));
  if (iu == B->du || id == B->dd) return B->dim;
  return (size_t)iu + (size_t)B->du * id;
}

/* ------------------------------------------------------------------------ */
/* sectors (src_ed/sector.hpp:54-106)                                        */
/* ------------------------------------------------------------------------ */
#define SEC_ODD 9999
#define SEC_EVEN 10000
int orc_sector_parse(const char *s, int *N, int *S, int *irrep) {
  int n = SEC_EVEN, sp = SEC_EVEN, r = 0;
  const char *p;
  if ((p = strchr(s, 'N'))) { if (sscanf(p, "N%d", &n) != 1) return -1; }
  if ((p = strchr(s, 'S'))) { if (sscanf(p, "S%d", &sp) != 1) return -1; }
  if ((p = strchr(s, 'R'))) { if (sscanf(p, "R%d", &r) != 1) return -1; }
  if (strchr(s, 'O')) n = sp = SEC_ODD;
  if (sp == SEC_EVEN && n != SEC_EVEN && (n % 2)) sp = SEC_ODD;
  if (n == SEC_EVEN && sp != SEC_EVEN && (sp % 2)) n = SEC_ODD;
  *N = n; *S = sp; *irrep = r;
  return 0;
}
int orc_sector_name(int N, int S, int irrep, char *buf, int cap) {
  char tmp[64];
  int k = 0;
  if (S == SEC_ODD && N == SEC_ODD) k += snprintf(tmp + k, sizeof(tmp) - k, "O");
  k += snprintf(tmp + k, sizeof(tmp) - k, "R%d", irrep);
  if (N != SEC_EVEN && N != SEC_ODD) k += snprintf(tmp + k, sizeof(tmp) - k, ":N%d", N);
  if (S != SEC_EVEN && S != SEC_ODD) k += snprintf(tmp + k, sizeof(tmp) - k, ":S%d", S);
  if (k + 1 > cap) return -1;
  memcpy(buf, tmp, (size_t)k + 1);
  return k;
}

/* ------------------------------------------------------------------------ */
/* HS_half_operator (src_ed/Operators/HS_half_operator.hpp:17-70)           */
/* ------------------------------------------------------------------------ */
typedef struct { double v; size_t n, cap; uint32_t *I, *J; } vgroup;
typedef struct { uint32_t r; double v; } dentry;
typedef struct { int ng; vgroup *g; size_t nd, dcap; dentry *d; uint32_t dim; } halfop;

static void hop_insert(halfop *H, uint32_t I, uint32_t J, double z) {
  if (fabs(z) < SMALL_VALUE) return;
  if (J > I) return;
  if (J == I) {
    if (H->nd == H->dcap) { H->dcap = H->dcap ? 2 * H->dcap : 64; H->d = (dentry *)realloc(H->d, H->dcap * sizeof(dentry)); }
    H->d[H->nd++] = (dentry){I, z};
    return;
  }
  int i;
  for (i = 0; i < H->ng; ++i) if (H->g[i].v == z) break;
  if (i == H->ng) {
    H->g = (vgroup *)realloc(H->g, sizeof(vgroup) * (size_t)(H->ng + 1));
    memset(&H->g[H->ng], 0, sizeof(vgroup));
    H->g[H->ng].v = z;
    H->ng++;
  }
  vgroup *G = &H->g[i];
  if (G->n == G->cap) {
    G->cap = G->cap ? 2 * G->cap : 256;
    G->I = (uint32_t *)realloc(G->I, G->cap * sizeof(uint32_t));
    G->J = (uint32_t *)realloc(G->J, G->cap * sizeof(uint32_t));
  }
  G->I[G->n] = I; G->J[G->n] = J; G->n++;
}

static int cmp_pair(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x < y) ? -1 : (x > y);
}
static int cmp_dentry(const void *a, const void *b) {
  const dentry *x = (const dentry *)a, *y = (const dentry *)b;
  return (x->r < y->r) ? -1 : (x->r > y->r);
}

static halfop *halfop_build(const oterm *T, int spin, const uint32_t *bin, uint32_t dim, int L) {
  halfop *H = (halfop *)calloc(1, sizeof(halfop));
  H->dim = dim;
  int off = spin ? L : 0;
  for (uint32_t I = 0; I < dim; ++I) {
    uint64_t ss = bin[I];
    for (int k = 0; k < T->n; ++k) {
      const oelem *x = &T->e[k];
      if ((x->r >= L && !spin) || (x->r < L && spin)) continue;
      if (x->r != x->c) {
        uint64_t s2 = ss;
        int ph = one_body(&s2, bmask(x->c - off, L), bmask(x->r - off, L));
        if (!ph) continue;
        uint32_t J = orc_half_index(bin, dim, (uint32_t)s2);
        if (J == dim) continue;
        hop_insert(H, I, J, ph * x->v);
      } else if (ss & bmask(x->r - off, L)) {
        hop_insert(H, I, I, x->v);
      }
    }
  }
  /* sort_elements(): pairs lexicographically (I,J); diagonal by row */
  for (int g = 0; g < H->ng; ++g) {
    vgroup *G = &H->g[g];
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (G->n ? G->n : 1));
    for (size_t k = 0; k < G->n; ++k) tmp[k] = ((uint64_t)G->I[k] << 32) | G->J[k];
    qsort(tmp, G->n, sizeof(uint64_t), cmp_pair);
    for (size_t k = 0; k < G->n; ++k) { G->I[k] = (uint32_t)(tmp[k] >> 32); G->J[k] = (uint32_t)tmp[k]; }
    free(tmp);
  }
  /* std::sort on diagonal_matrix_element compares rows only; equal rows do
     not occur within one term (one diagonal element per orbital per state is
     accumulated separately), stable order is kept by qsort on (r) here. */
  qsort(H->d, H->nd, sizeof(dentry), cmp_dentry);
  return H;
}

static void halfop_free(halfop *H) {
  if (!H) return;
  for (int g = 0; g < H->ng; ++g) { free(H->g[g].I); free(H->g[g].J); }
  free(H->g); free(H->d); free(H);
}

int orc_half_table(const orc_model *m, int t, int spin, int N, size_t *n_pairs, size_t *n_diag,
                   uint32_t *pI, uint32_t *pJ, double *pv, int *pgroup, uint32_t *dI, double *dv) {
  if (t < 0 || t >= m->n_terms) return -1;
  if (m->type[t] != ORC_ONE_BODY) return -2;
  pmodel *P = prepare(m);
  size_t dim = orc_half_basis(m->L, N, NULL);
  uint32_t *bin = (uint32_t *)malloc(sizeof(uint32_t) * (dim ? dim : 1));
  orc_half_basis(m->L, N, bin);
  halfop *H = halfop_build(&P->t[t], spin, bin, (uint32_t)dim, m->L);
  size_t np = 0;
  for (int g = 0; g < H->ng; ++g) {
    for (size_t k = 0; k < H->g[g].n; ++k) {
      if (pI) { pI[np] = H->g[g].I[k]; pJ[np] = H->g[g].J[k]; pv[np] = H->g[g].v; pgroup[np] = g; }
      ++np;
    }
  }
  if (dI) for (size_t k = 0; k < H->nd; ++k) { dI[k] = H->d[k].r; dv[k] = H->d[k].v; }
  *n_pairs = np;
  *n_diag = H->nd;
  halfop_free(H); free(bin); release(P);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* interaction diagonal (src_ed/Operators/HS_interaction_operator.hpp:32-56) */
/* The state word here is the physical one (up bits low, down bits high),    */
/* which equals the reference for up/down-symmetric density-density terms.   */
/* ------------------------------------------------------------------------ */
static double inter_elem(const oterm *T, uint64_t w, int L) {
  double val = 0.0;
  for (int k = 0; k < T->n; ++k) {
    uint64_t mk = bmask(T->e[k].r, L) + bmask(T->e[k].c, L);
    if ((w & mk) == mk) val += T->e[k].v;
  }
  return val;
}

static int factorizable(const pmodel *P) {
  for (int t = 0; t < P->nt; ++t) {
    if (P->t[t].type != ORC_ONE_BODY) continue;
    for (int k = 0; k < P->t[t].n; ++k) {
      int r = P->t[t].e[k].r, c = P->t[t].e[k].c;
      if ((r < P->L) != (c < P->L)) return 0;
    }
  }
  return 1;
}

/* ------------------------------------------------------------------------ */
/* factorized H.x, reference order                                           */
/* (src_ed/Hamiltonian/Hamiltonian_Factorized.hpp:51-58,                     */
/*  src_ed/Operators/HS_factorized_operator.hpp:57-84,                       */
/*  src_ed/Operators/HS_interaction_operator.hpp:62-71)                      */
/* ------------------------------------------------------------------------ */
int orc_hv_factorized(const orc_model *m, int Nup, int Ndw, const double *x, double *y) {
  pmodel *P = prepare(m);
  if (!factorizable(P)) { release(P); return -3; }
  fbasis B;
  if (fb_make(&B, m->L, Nup, Ndw)) { release(P); return -1; }
  size_t du = B.du, dd = B.dd;
  for (int t = 0; t < P->nt; ++t) {
    const oterm *T = &P->t[t];
    double z = T->value;
    if (T->type == ORC_ONE_BODY) {
      halfop *up = halfop_build(T, 0, B.up, B.du, m->L);
      halfop *dn = halfop_build(T, 1, B.dn, B.dd, m->L);
      for (int g = 0; g < dn->ng; ++g) {
        double z2 = dn->g[g].v * z;
        for (size_t k = 0; k < dn->g[g].n; ++k) {
          size_t a = dn->g[g].I[k], b = dn->g[g].J[k];
          for (size_t i = 0; i < du; ++i) y[a * du + i] += z2 * x[b * du + i];
        }
        for (size_t k = 0; k < dn->g[g].n; ++k) {
          size_t a = dn->g[g].I[k], b = dn->g[g].J[k];
          for (size_t i = 0; i < du; ++i) y[b * du + i] += z2 * x[a * du + i];
        }
      }
      for (size_t k = 0; k < dn->nd; ++k) {
        size_t r = dn->d[k].r;
        for (size_t i = 0; i < du; ++i) y[r * du + i] += z * dn->d[k].v * x[r * du + i];
      }
      for (size_t i = 0; i < dd; ++i) {
        for (int g = 0; g < up->ng; ++g) {
          double z2 = up->g[g].v * z;
          for (size_t k = 0; k < up->g[g].n; ++k) y[up->g[g].I[k] + du * i] += z2 * x[up->g[g].J[k] + du * i];
          for (size_t k = 0; k < up->g[g].n; ++k) y[up->g[g].J[k] + du * i] += z2 * x[up->g[g].I[k] + du * i];
        }
        for (size_t k = 0; k < up->nd; ++k) y[up->d[k].r + du * i] += z * up->d[k].v * x[up->d[k].r + du * i];
      }
      halfop_free(up);
      halfop_free(dn);
    } else {
      for (size_t I = 0; I < B.dim; ++I) y[I] += x[I] * inter_elem(T, fb_word(&B, I), m->L) * z;
    }
  }
  fb_free(&B);
  release(P);
  return 0;
}

/* full diagonal of H (Hamiltonian_Factorized::diag plus the one-body
   diagonals that HS_factorized_operator::diag leaves out, see
   src_ed/Operators/HS_factorized_operator.hpp:31) */
int orc_diag(const orc_model *m, int Nup, int Ndw, double *d) {
  pmodel *P = prepare(m);
  fbasis B;
  if (fb_make(&B, m->L, Nup, Ndw)) { release(P); return -1; }
  for (size_t I = 0; I < B.dim; ++I) {
    uint64_t w = fb_word(&B, I);
    double s = 0.0;
    for (int t = 0; t < P->nt; ++t) {
      const oterm *T = &P->t[t];
      if (T->type == ORC_ONE_BODY) {
        double a = 0.0;
        for (int k = 0; k < T->n; ++k)
          if (T->e[k].r == T->e[k].c && (w & bmask(T->e[k].r, m->L))) a += T->e[k].v;
        s += T->value * a;
      } else {
        s += T->value * inter_elem(T, w, m->L);
      }
    }
    d[I] = s;
  }
  fb_free(&B);
  release(P);
  return 0;
}

/* merged, symmetric one-body tables for the parallel CPU baseline */
typedef struct { uint32_t n; uint32_t *ptr; uint32_t *col; double *val; double *diag; } symcsr;

static int cmp_u32d(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return (x < y) ? -1 : (x > y);
}

static void build_sym(const pmodel *P, int spin, const uint32_t *bin, uint32_t dim, symcsr *S) {
  /* collect (row, col, value*term) for both triangles, then merge */
  size_t cap = 1024, n = 0;
  uint64_t *key = (uint64_t *)malloc(cap * sizeof(uint64_t));
  double *val = (double *)malloc(cap * sizeof(double));
  S->n = dim;
  S->diag = (double *)calloc(dim ? dim : 1, sizeof(double));
  for (int t = 0; t < P->nt; ++t) {
    const oterm *T = &P->t[t];
    if (T->type != ORC_ONE_BODY) continue;
    halfop *H = halfop_build(T, spin, bin, dim, P->L);
    for (int g = 0; g < H->ng; ++g) {
      for (size_t k = 0; k < H->g[g].n; ++k) {
        if (n + 2 > cap) { cap *= 2; key = (uint64_t *)realloc(key, cap * 8); val = (double *)realloc(val, cap * 8); }
        uint32_t a = H->g[g].I[k], b = H->g[g].J[k];
        key[n] = ((uint64_t)a << 32) | b; val[n++] = H->g[g].v * T->value;
        key[n] = ((uint64_t)b << 32) | a; val[n++] = H->g[g].v * T->value;
      }
    }
    for (size_t k = 0; k < H->nd; ++k) S->diag[H->d[k].r] += T->value * H->d[k].v;
    halfop_free(H);
  }
  /* sort by key carrying values: pack index */
  uint64_t *idx = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t) * 2);
  for (size_t k = 0; k < n; ++k) { idx[2 * k] = key[k]; memcpy(&idx[2 * k + 1], &val[k], 8); }
  qsort(idx, n, 16, cmp_u32d);
  S->ptr = (uint32_t *)calloc((size_t)dim + 1, sizeof(uint32_t));
  S->col = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  S->val = (double *)malloc((n ? n : 1) * sizeof(double));
  size_t m = 0;
  for (size_t k = 0; k < n; ++k) {
    uint64_t kk = idx[2 * k];
    double vv; memcpy(&vv, &idx[2 * k + 1], 8);
    /* merge equal keys */
    if (k > 0 && idx[2 * (k - 1)] == kk) { S->val[m - 1] += vv; continue; }
    S->col[m] = (uint32_t)kk; S->val[m] = vv; S->ptr[(kk >> 32) + 1]++; ++m;
  }
  for (uint32_t r = 0; r < dim; ++r) S->ptr[r + 1] += S->ptr[r];
  free(idx); free(key); free(val);
}
static void free_sym(symcsr *S) { free(S->ptr); free(S->col); free(S->val); free(S->diag); }

/* OpenMP-parallel factorized H.x over the down index (CPU baseline variant (i)):
   same operator, rows of X[i_dn][i_up] processed in parallel, diagonal
   precomputed as in HS_interaction_operator.  nrep repetitions, y += H x each. */
int orc_hv_factorized_omp(const orc_model *m, int Nup, int Ndw, const double *x, double *y, int nrep) {
  return orc_hv_factorized_omp_timed(m, Nup, Ndw, x, y, nrep, NULL);
}

/* the same, with the wall time of each product in ms[rep] (tables built first,
   outside the timed products) */
int orc_hv_factorized_omp_timed(const orc_model *m, int Nup, int Ndw, const double *x, double *y, int nrep,
                                double *ms) {
  pmodel *P = prepare(m);
  if (!factorizable(P)) { release(P); return -3; }
  fbasis B;
  if (fb_make(&B, m->L, Nup, Ndw)) { release(P); return -1; }
  symcsr KU, KD;
  build_sym(P, 0, B.up, B.du, &KU);
  build_sym(P, 1, B.dn, B.dd, &KD);
  const size_t du = B.du, dd = B.dd;
  double *dg = (double *)malloc(sizeof(double) * (B.dim ? B.dim : 1));
#pragma omp parallel for schedule(static)
  for (size_t d = 0; d < dd; ++d) {
    for (size_t u = 0; u < du; ++u) {
      uint64_t w = (uint64_t)B.up[u] | ((uint64_t)B.dn[d] << 32);
      double s = KU.diag[u] + KD.diag[d];
      for (int t = 0; t < P->nt; ++t)
        if (P->t[t].type == ORC_INTERACTION) s += P->t[t].value * inter_elem(&P->t[t], w, m->L);
      dg[d * du + u] = s;
    }
  }
  for (int rep = 0; rep < nrep; ++rep) {
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t d = 0; d < dd; ++d) {
      double *yr = y + d * du;
      const double *xr = x + d * du;
      const double *gr = dg + d * du;
      for (size_t u = 0; u < du; ++u) yr[u] += gr[u] * xr[u];
      for (uint32_t k = KD.ptr[d]; k < KD.ptr[d + 1]; ++k) {
        const double v = KD.val[k];
        const double *xs = x + (size_t)KD.col[k] * du;
        for (size_t u = 0; u < du; ++u) yr[u] += v * xs[u];
      }
      for (size_t u = 0; u < du; ++u) {
        double acc = 0.0;
        for (uint32_t k = KU.ptr[u]; k < KU.ptr[u + 1]; ++k) acc += KU.val[k] * xr[KU.col[k]];
        yr[u] += acc;
      }
    }
    if (ms) ms[rep] = 1e3 * (omp_get_wtime() - t0);
  }
  free(dg);
  free_sym(&KU); free_sym(&KD);
  fb_free(&B);
  release(P);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* CSR format (src_ed/Hamiltonian/Hamiltonian_CSR.hpp:38-81,                 */
/* src_ed/Operators/HS_one_body_operator.hpp:20-36,                          */
/* src_ed/Operators/HS_nondiagonal_operator.hpp:58-76, :152-159):            */
/* strict lower triangle, row = out state > col = in state, map order (r,c)  */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t key; double v; } kv;
static int cmp_kv(const void *a, const void *b) {
  const kv *x = (const kv *)a, *y = (const kv *)b;
  return (x->key < y->key) ? -1 : (x->key > y->key);
}

int orc_csr_build(const orc_model *m, int Nup, int Ndw, size_t *nnz, int32_t *Iptr, int32_t *J,
                  double *v, double *diag) {
  pmodel *P = prepare(m);
  fbasis B;
  if (fb_make(&B, m->L, Nup, Ndw)) { release(P); return -1; }
  if (B.dim >= 0x7fffffff) { fb_free(&B); release(P); return -4; }
  size_t cap = 1024, n = 0;
  kv *E = (kv *)malloc(cap * sizeof(kv));
  if (diag) memset(diag, 0, sizeof(double) * B.dim);
  for (int t = 0; t < P->nt; ++t) {
    const oterm *T = &P->t[t];
    double z = T->value;
    for (size_t I = 0; I < B.dim; ++I) {
      uint64_t ss = fb_word(&B, I);
      if (T->type == ORC_INTERACTION) {
        if (diag) diag[I] += z * inter_elem(T, ss, m->L);
        continue;
      }
      for (int k = 0; k < T->n; ++k) {
        const oelem *x = &T->e[k];
        if (x->r != x->c) {
          uint64_t s2 = ss;
          int ph = one_body(&s2, bmask(x->c, m->L), bmask(x->r, m->L));
          if (!ph) continue;
          size_t Jo = fb_index(&B, s2);
          if (Jo == B.dim) continue;
          double X = ph * x->v;
          if (fabs(X) < SMALL_VALUE) continue;
          if (I > Jo || I == Jo) continue; /* insert(J_out, I_in): keep I_in < J_out */
          if (n == cap) { cap *= 2; E = (kv *)realloc(E, cap * sizeof(kv)); }
          E[n].key = ((uint64_t)Jo << 32) | (uint64_t)I;
          E[n].v = X * z;
          ++n;
        } else if (ss & bmask(x->r, m->L)) {
          if (diag && fabs(x->v) >= SMALL_VALUE) diag[I] += z * x->v;
        }
      }
    }
  }
  qsort(E, n, sizeof(kv), cmp_kv);
  size_t mm = 0;
  for (size_t k = 0; k < n; ++k) {
    if (mm > 0 && E[mm - 1].key == E[k].key) { E[mm - 1].v += E[k].v; continue; }
    E[mm++] = E[k];
  }
  *nnz = mm;
  if (Iptr) {
    size_t row = 0, count = 0;
    Iptr[0] = 0;
    size_t ip = 1;
    for (size_t k = 0; k < mm; ++k) {
      size_t r = (size_t)(E[k].key >> 32);
      if (r != row) {
        for (size_t i = row; i < r; ++i) Iptr[ip++] = (int32_t)count;
        row = r;
      }
      J[k] = (int32_t)(E[k].key & 0xffffffffu);
      v[k] = E[k].v;
      ++count;
    }
    Iptr[ip++] = (int32_t)count;
    /* remaining rows (after the last nonzero row) are absent in the
       reference's Iptr; pad so that Iptr has dim+1 entries */
    while (ip < B.dim + 1) Iptr[ip++] = (int32_t)count;
  }
  free(E);
  fb_free(&B);
  release(P);
  return 0;
}

/* CSR_hermitian::apply (src_ed/Hamiltonian/CSR_hermitian.hpp:27-48) */
int orc_hv_csr(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v, const double *diag,
               const double *x, double *y) {
  for (size_t i = 0; i < dim; ++i) y[i] += diag[i] * x[i];
  for (size_t i = 0; i < dim; ++i)
    for (int32_t j = Iptr[i]; j < Iptr[i + 1]; ++j) y[i] += x[J[j]] * v[j];
  for (size_t i = 0; i < dim; ++i)
    for (int32_t j = Iptr[i]; j < Iptr[i + 1]; ++j) y[J[j]] += x[i] * v[j];
  return 0;
}

/* row-parallel SpMV over the full symmetric CSR (CPU baseline variant (ii),
   the behaviour of the default Eigen RowMajor format, Hamiltonian_Eigen.hpp:75-82) */
int orc_hv_csr_full_omp(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v,
                        const double *diag, const double *x, double *y, int nrep) {
  return orc_hv_csr_full_omp_timed(dim, Iptr, J, v, diag, x, y, nrep, NULL);
}

/* the same, with the wall time of each product in ms[rep] (the full rows are
   expanded first, outside the timed products) */
int orc_hv_csr_full_omp_timed(size_t dim, const int32_t *Iptr, const int32_t *J, const double *v,
                              const double *diag, const double *x, double *y, int nrep, double *ms) {
  /* expand lower CSR into full CSR */
  size_t nnz = (size_t)Iptr[dim];
  int32_t *cnt = (int32_t *)calloc(dim + 1, sizeof(int32_t));
  for (size_t i = 0; i < dim; ++i) {
    cnt[i + 1] += Iptr[i + 1] - Iptr[i];
    for (int32_t j = Iptr[i]; j < Iptr[i + 1]; ++j) cnt[J[j] + 1]++;
  }
  for (size_t i = 0; i < dim; ++i) cnt[i + 1] += cnt[i];
  int32_t *fc = (int32_t *)malloc(sizeof(int32_t) * (2 * nnz + 1));
  double *fv = (double *)malloc(sizeof(double) * (2 * nnz + 1));
  int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (dim + 1));
  memcpy(pos, cnt, sizeof(int32_t) * (dim + 1));
  for (size_t i = 0; i < dim; ++i)
    for (int32_t j = Iptr[i]; j < Iptr[i + 1]; ++j) {
      fc[pos[i]] = J[j]; fv[pos[i]++] = v[j];
      fc[pos[J[j]]] = (int32_t)i; fv[pos[J[j]]++] = v[j];
    }
  for (int rep = 0; rep < nrep; ++rep) {
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < dim; ++i) {
      double acc = diag[i] * x[i];
      for (int32_t j = cnt[i]; j < cnt[i + 1]; ++j) acc += fv[j] * x[fc[j]];
      y[i] += acc;
    }
    if (ms) ms[rep] = 1e3 * (omp_get_wtime() - t0);
  }
  free(cnt); free(fc); free(fv); free(pos);
  return 0;
}

/* on-the-fly H.x (src_ed/Hamiltonian/Hamiltonian_OnTheFly.hpp:33-60,
   one_body_operator.hpp:473-494 real push form, interaction_operator.hpp:141-149) */
int orc_hv_otf(const orc_model *m, int Nup, int Ndw, const double *x, double *y) {
  pmodel *P = prepare(m);
  fbasis B;
  if (fb_make(&B, m->L, Nup, Ndw)) { release(P); return -1; }
  for (int t = 0; t < P->nt; ++t) {
    const oterm *T = &P->t[t];
    double z = T->value;
    for (size_t I = 0; I < B.dim; ++I) {
      uint64_t ss = fb_word(&B, I);
      for (int k = 0; k < T->n; ++k) {
        const oelem *E = &T->e[k];
        if (T->type == ORC_INTERACTION) {
          uint64_t mk = bmask(E->r, m->L) + bmask(E->c, m->L);
          if ((ss & mk) == mk) y[I] += z * E->v * x[I];
          continue;
        }
        if (E->r != E->c) {
          uint64_t s2 = ss;
          int ph = one_body(&s2, bmask(E->c, m->L), bmask(E->r, m->L));
          if (!ph) continue;
          size_t Jo = fb_index(&B, s2);
          if (Jo == B.dim) continue;
          y[Jo] += z * ph * E->v * x[I];
        } else if (ss & bmask(E->r, m->L)) {
          y[I] += z * E->v * x[I];
        }
      }
    }
  }
  fb_free(&B);
  release(P);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* c_a, c^dagger_a between factorized sectors                                */
/* (src_ed/model.cpp:155-182, src_ed/ED_basis.cpp:79-110)                    */
/* pm=-1: y += z c_a x, x in (Nup,Ndw).  pm=+1: y += z c^dagger_a x.         */
/* ------------------------------------------------------------------------ */
int orc_create_or_destroy(int L, int pm, int a, int Nup, int Ndw, const double *x, double z, double *y) {
  int spin = (a >= L);
  int tNup = Nup + ((!spin) ? pm : 0), tNdw = Ndw + (spin ? pm : 0);
  fbasis Bx, Bt;
  if (fb_make(&Bx, L, Nup, Ndw)) return -1;
  if (fb_make(&Bt, L, tNup, tNdw)) { fb_free(&Bx); return -1; }
  /* Destroy(a, I, From, To) */
  const fbasis *From = (pm == 1) ? &Bt : &Bx;
  const fbasis *To = (pm == 1) ? &Bx : &Bt;
  for (size_t I = 0; I < From->dim; ++I) {
    uint32_t lu = (uint32_t)(I % From->du), ld = (uint32_t)(I / From->du);
    uint32_t su = From->up[lu], sd = From->dn[ld];
    uint32_t lp;
    size_t Jt;
    int ph;
    if (spin) {
      uint32_t mk = 1u << (a - L);
      if (!(sd & mk)) continue;
      ph = (__builtin_popcount(sd & (mk - 1)) + __builtin_popcount(su)) % 2;
      lp = orc_half_index(To->dn, To->dd, sd ^ mk);
      if (lp == To->dd) continue;
      Jt = (size_t)lp * To->du + lu;
    } else {
      uint32_t mk = 1u << a;
      if (!(su & mk)) continue;
      ph = __builtin_popcount(su & (mk - 1)) % 2;
      lp = orc_half_index(To->up, To->du, su ^ mk);
      if (lp == To->du) continue;
      Jt = (size_t)ld * To->du + lp;
    }
    double sg = 1.0 - 2.0 * ph;
    if (pm == 1) y[I] += z * sg * x[Jt];
    else y[Jt] += z * sg * x[I];
  }
  fb_free(&Bx); fb_free(&Bt);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* dense symmetric eigensolvers (Householder + implicit QL)                  */
/* ------------------------------------------------------------------------ */
static double hypot2(double a, double b) { return hypot(a, b); }

/* implicit QL on a symmetric tridiagonal matrix; z (n x n, row-major
   z[k*n+i]) accumulates the rotations when non-NULL. d: diagonal in/out,
   e: subdiagonal with e[i] coupling i and i+1 (e[n-1] ignored). */
static int tql(int n, double *d, double *e, double *z) {
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) return -1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot2(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = hypot2(f, g);
          e[i + 1] = r;
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
          s = f / r; c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (z) {
            for (int k = 0; k < n; ++k) {
              double fz = z[(size_t)k * n + i + 1];
              z[(size_t)k * n + i + 1] = s * z[(size_t)k * n + i] + c * fz;
              z[(size_t)k * n + i] = c * z[(size_t)k * n + i] - s * fz;
            }
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p; e[l] = g; e[m] = 0.0;
      }
    } while (m != l);
  }
  return 0;
}

/* sort eigenvalues ascending, permuting the columns of z (row-major n x n) */
static void eig_sort(int n, double *d, double *z) {
  for (int i = 0; i < n - 1; ++i) {
    int k = i;
    for (int j = i + 1; j < n; ++j) if (d[j] < d[k]) k = j;
    if (k != i) {
      double t = d[i]; d[i] = d[k]; d[k] = t;
      if (z) for (int r = 0; r < n; ++r) { double q = z[(size_t)r * n + i]; z[(size_t)r * n + i] = z[(size_t)r * n + k]; z[(size_t)r * n + k] = q; }
    }
  }
}

/* symmetric tridiagonal eigensystem (stands in for LAPACK dstev,
   src_ed/Hamiltonian/Lanczos.cpp:25-52).  z (optional) is n x n column-major:
   z[i + k n] = component i of eigenvector k. */
int orc_tridiag_eig(int n, const double *d, const double *e, double *w, double *z) {
  double *dd = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  double *ee = (double *)malloc(sizeof(double) * (size_t)(n + 1));
  memcpy(dd, d, sizeof(double) * (size_t)n);
  for (int i = 0; i < n - 1; ++i) ee[i] = e[i];
  ee[n > 0 ? n - 1 : 0] = 0.0;
  double *zr = NULL;
  if (z) {
    zr = (double *)calloc((size_t)n * n, sizeof(double));
    for (int i = 0; i < n; ++i) zr[(size_t)i * n + i] = 1.0;
  }
  int info = tql(n, dd, ee, zr);
  eig_sort(n, dd, zr);
  memcpy(w, dd, sizeof(double) * (size_t)n);
  if (z) for (int i = 0; i < n; ++i) for (int k = 0; k < n; ++k) z[i + (size_t)k * n] = zr[(size_t)i * n + k];
  free(dd); free(ee); free(zr);
  return info;
}

/* Householder reduction of a symmetric matrix a (row-major n x n, lower
   triangle used) to tridiagonal form; with vec != 0 a is replaced by the
   orthogonal transform.  d: diagonal, e: e[i] couples i-1 and i. */
static void householder(int n, double *a, double *d, double *e, int vec) {
#define A(i, j) a[(size_t)(i) * n + (j)]
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
        double f = A(i, l);
        double g = (f >= 0.0) ? -sqrt(h) : sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          if (vec) A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    if (vec) {
      int l = i - 1;
      if (d[i] != 0.0) {
        for (int j = 0; j <= l; ++j) {
          double g = 0.0;
          for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
          for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
        }
      }
      d[i] = A(i, i);
      A(i, i) = 1.0;
      for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
    } else {
      d[i] = A(i, i);
    }
  }
#undef A
}

/* dense symmetric eigensystem (stands in for Eigen SelfAdjointEigenSolver,
   src_util/matrix.cpp:56-141).  A: n x n column-major; V (optional) column-
   major eigenvectors, eigenvalues ascending. */
int orc_symeig(int n, const double *A, double *w, double *V) {
  if (n <= 0) return 0;
  double *a = (double *)malloc(sizeof(double) * (size_t)n * n);
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = A[i + (size_t)j * n];
  double *d = (double *)malloc(sizeof(double) * (size_t)n);
  double *e = (double *)malloc(sizeof(double) * (size_t)n);
  householder(n, a, d, e, V != NULL);
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  int info = tql(n, d, e, V ? a : NULL);
  eig_sort(n, d, V ? a : NULL);
  memcpy(w, d, sizeof(double) * (size_t)n);
  if (V) for (int i = 0; i < n; ++i) for (int k = 0; k < n; ++k) V[i + (size_t)k * n] = a[(size_t)i * n + k];
  free(a); free(d); free(e);
  return info;
}

/* ------------------------------------------------------------------------ */
/* vector helpers                                                            */
/* ------------------------------------------------------------------------ */
static double vdot(size_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
static void vaxpy(size_t n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; ++i) y[i] += a * x[i];
}
static void vscal(size_t n, double a, double *x) {
#pragma omp parallel for schedule(static) if (n > 65536)
  for (size_t i = 0; i < n; ++i) x[i] *= a;
}

/* The Krylov drivers use the OpenMP factorized H.x with the full
   precomputed diagonal; the apply object is built once per driver call. */
typedef struct { fbasis B; symcsr KU, KD; double *dg; } happly;

static int happly_make(happly *H, const orc_model *m, int Nup, int Ndw) {
  pmodel *P = prepare(m);
  if (!factorizable(P)) { release(P); return -3; }
  if (fb_make(&H->B, m->L, Nup, Ndw)) { release(P); return -1; }
  build_sym(P, 0, H->B.up, H->B.du, &H->KU);
  build_sym(P, 1, H->B.dn, H->B.dd, &H->KD);
  size_t du = H->B.du, dd = H->B.dd;
  H->dg = (double *)malloc(sizeof(double) * (H->B.dim ? H->B.dim : 1));
#pragma omp parallel for schedule(static)
  for (size_t d = 0; d < dd; ++d)
    for (size_t u = 0; u < du; ++u) {
      uint64_t w = (uint64_t)H->B.up[u] | ((uint64_t)H->B.dn[d] << 32);
      double s = H->KU.diag[u] + H->KD.diag[d];
      for (int t = 0; t < P->nt; ++t)
        if (P->t[t].type == ORC_INTERACTION) s += P->t[t].value * inter_elem(&P->t[t], w, m->L);
      H->dg[d * du + u] = s;
    }
  release(P);
  return 0;
}
static void happly_free(happly *H) { free_sym(&H->KU); free_sym(&H->KD); free(H->dg); fb_free(&H->B); }

/* y += H x */
static void happly_run(const happly *H, const double *x, double *y) {
  const size_t du = H->B.du, dd = H->B.dd;
#pragma omp parallel for schedule(dynamic, 4) if (dd * du > 4096)
  for (size_t d = 0; d < dd; ++d) {
    double *yr = y + d * du;
    const double *xr = x + d * du;
    const double *gr = H->dg + d * du;
    for (size_t u = 0; u < du; ++u) yr[u] += gr[u] * xr[u];
    for (uint32_t k = H->KD.ptr[d]; k < H->KD.ptr[d + 1]; ++k) {
      const double v = H->KD.val[k];
      const double *xs = x + (size_t)H->KD.col[k] * du;
      for (size_t u = 0; u < du; ++u) yr[u] += v * xs[u];
    }
    for (size_t u = 0; u < du; ++u) {
      double acc = 0.0;
      for (uint32_t k = H->KU.ptr[u]; k < H->KU.ptr[u + 1]; ++k) acc += H->KU.val[k] * xr[H->KU.col[k]];
      yr[u] += acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Lanczos ground state (src_ed/Hamiltonian/Lanczos.hpp:41-143)              */
/* start must be normalized; gs receives the (unnormalized) Ritz vector.     */
/* alpha/beta (optional) receive max_iter entries at most.                   */
/* ------------------------------------------------------------------------ */
int orc_lanczos_gs(const orc_model *m, int Nup, int Ndw, const double *start, int max_iter, double accur,
                   double *E0, double *gs, int *niter, double *alpha_out, double *beta_out) {
  happly H;
  int rc = happly_make(&H, m, Nup, Ndw);
  if (rc) return rc;
  size_t n = H.B.dim;
  double *r = (double *)malloc(sizeof(double) * n), *q = (double *)calloc(n, sizeof(double));
  memcpy(r, start, sizeof(double) * n);
  double *al = (double *)malloc(sizeof(double) * (size_t)(max_iter + 2));
  double *be = (double *)malloc(sizeof(double) * (size_t)(max_iter + 2));
  double *w = (double *)malloc(sizeof(double) * (size_t)(max_iter + 2));
  double *z = NULL;
  int k = 0;
  double ritz = 2.0 * accur, e0 = 0.0;
  while (ritz > accur) {
    if (k) { vscal(n, 1.0 / be[k - 1], r); vscal(n, -be[k - 1], q); }
    happly_run(&H, r, q);
    double *t = r; r = q; q = t;
    al[k] = vdot(n, q, r);
    vaxpy(n, -al[k], q, r);
    be[k] = sqrt(vdot(n, r, r));
    ++k;
    free(z);
    z = (double *)malloc(sizeof(double) * (size_t)k * k);
    orc_tridiag_eig(k, al, be, w, z);
    e0 = w[0];
    ritz = fabs(z[k - 1] * be[k - 1]);
    if (k > max_iter) { free(z); free(r); free(q); free(al); free(be); free(w); happly_free(&H); return -10; }
  }
  /* eigenvector pass (Lanczos.hpp:109-143) */
  memcpy(r, start, sizeof(double) * n);
  memset(q, 0, sizeof(double) * n);
  for (size_t i = 0; i < n; ++i) gs[i] = start[i] * z[0];
  for (int j = 1; j < k; ++j) {
    happly_run(&H, r, q);
    vaxpy(n, -al[j - 1], r, q);
    for (size_t i = 0; i < n; ++i) {
      double tmp = r[i];
      r[i] = q[i] / be[j - 1];
      q[i] = -be[j - 1] * tmp;
    }
    vaxpy(n, z[j], r, gs);
  }
  *E0 = e0;
  *niter = k;
  if (alpha_out) memcpy(alpha_out, al, sizeof(double) * (size_t)k);
  if (beta_out) memcpy(beta_out, be, sizeof(double) * (size_t)k);
  free(z); free(r); free(q); free(al); free(be); free(w);
  happly_free(&H);
  return 0;
}

/* LanczosGreen (src_ed/Hamiltonian/Lanczos.hpp:156-183); psi normalized.
   a receives n entries, b receives n+1 entries (b[0] = 1). */
int orc_lanczos_green(const orc_model *m, int Nup, int Ndw, const double *psi0, double accur_cf, int max_iter_cf,
                      double *a, double *b, int *n_out) {
  happly H;
  int rc = happly_make(&H, m, Nup, Ndw);
  if (rc) return rc;
  size_t n = H.B.dim;
  double *psi = (double *)malloc(sizeof(double) * n), *q = (double *)calloc(n, sizeof(double));
  memcpy(psi, psi0, sizeof(double) * n);
  int na = 0;
  b[0] = 1.0;
  double prod = 1.0;
  while (prod > accur_cf && fabs(b[na]) > accur_cf && na < max_iter_cf) {
    if (na) {
      double beta = b[na], zi = 1.0 / beta;
      for (size_t i = 0; i < n; ++i) { double t = psi[i]; psi[i] = q[i] * zi; q[i] = -beta * t; }
    }
    happly_run(&H, psi, q);
    a[na] = vdot(n, psi, q);
    vaxpy(n, -a[na], psi, q);
    double zz = sqrt(vdot(n, q, q));
    b[na + 1] = zz;
    prod *= fabs(zz);
    ++na;
  }
  *n_out = na;
  free(psi); free(q);
  happly_free(&H);
  return 0;
}

/* continued_fraction::evaluate (src_ed/continued_fraction.cpp:52-57), on
   already-transformed coefficients (a shifted, b squared, b[0] = weight). */
double orc_cf_eval_1(int n, const double *a, const double *b, double zr, double zi, double *gi) {
  double complex z = zr + CI * zi, G = 0.0;
  for (int i = n - 1; i >= 0; --i) G = b[i] / (z - a[i] - G);
  *gi = cimag(G);
  return creal(G);
}
int orc_cf_eval(int n, const double *a, const double *b, int nz, const double *z, double *g) {
  for (int k = 0; k < nz; ++k) g[2 * k] = orc_cf_eval_1(n, a, b, z[2 * k], z[2 * k + 1], &g[2 * k + 1]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* band Lanczos (src_ed/Hamiltonian/Lanczos.hpp:214-395) and build_Q_matrix  */
/* (src_ed/Hamiltonian/Hamiltonian_base.hpp:230-254).  phi: p vectors of      */
/* length dim (vector k at phi + k*dim).  On return M_out poles, evals        */
/* ascending, Q (p x M_out column-major) = P[:, :p] U[:p, :].                 */
/* M_in < 0 selects the reference's M = 14 p ln(dim).                        */
/* ------------------------------------------------------------------------ */
int orc_band_lanczos(const orc_model *m, int Nup, int Ndw, int p, const double *phi, int M_in, int max_iter_bl,
                     double accur_deflation, double accur_bl, int *M_out, double *evals, double *Q) {
  happly H;
  int rc = happly_make(&H, m, Nup, Ndw);
  if (rc) return rc;
  size_t dim = H.B.dim;
  int M0 = (M_in < 0) ? (int)(14 * p * log(1.0 * dim)) : M_in;
  int pc = p;
  if (M0 > max_iter_bl) M0 = max_iter_bl;
  int M0p = (M0 / p + 1) * p;
  M0 = M0p - pc;
  double **v = (double **)calloc((size_t)M0p + 1, sizeof(double *));
  int *Idef = (int *)calloc((size_t)M0p + 1, sizeof(int));
  int *def = (int *)calloc((size_t)M0 + 1, sizeof(int));
  double *t = (double *)calloc((size_t)M0 * M0, sizeof(double)); /* column-major M0 x M0 */
  double *P = (double *)calloc((size_t)p * M0, sizeof(double));   /* column-major p x M0 */
#define T_(i, j) t[(size_t)(i) + (size_t)(j) * M0]
#define P_(i, j) P[(size_t)(i) + (size_t)(j) * p]
  for (int i = 0; i < p; ++i) {
    v[i] = (double *)malloc(sizeof(double) * dim);
    memcpy(v[i], phi + (size_t)i * dim, sizeof(double) * dim);
  }
  int nd = 0, j;
  double ev_old = 1e12;
  for (j = 0; j < M0; ++j) {
    double z = sqrt(vdot(dim, v[j], v[j]));
    if (z < accur_deflation) {
      if (j >= pc) { Idef[nd] = j - pc; def[j - pc] = 1; nd++; }
      pc--;
      if (pc == 0) break;
      free(v[j]); v[j] = NULL;
      for (int k = j; k < j + pc; ++k) v[k] = v[k + 1];
      v[j + pc] = NULL;
      j--;
      continue;
    }
    vscal(dim, 1.0 / z, v[j]);
    if (j >= pc) T_(j, j - pc) = z;
    for (int k = 0; k < p; ++k) P_(k, j) = vdot(dim, phi + (size_t)k * dim, v[j]);
    int jm = j + pc;
    for (int k = j + 1; k < jm; ++k) {
      double zz = vdot(dim, v[j], v[k]);
      vaxpy(dim, -zz, v[j], v[k]);
      if (k >= pc) T_(j, k - pc) = zz;
    }
    v[j + pc] = (double *)calloc(dim, sizeof(double));
    happly_run(&H, v[j], v[j + pc]);
    int k0 = 0;
    if (j > pc) {
      k0 = j - pc;
      if (def[k0 - 1] == 0) { free(v[k0 - 1]); v[k0 - 1] = NULL; }
    }
    for (int k = k0; k < j; ++k) {
      T_(k, j) = T_(j, k);
      vaxpy(dim, -T_(k, j), v[k], v[j + pc]);
    }
    for (int k = 0; k < nd; ++k) {
      double w = vdot(dim, v[Idef[k]], v[j + pc]);
      vaxpy(dim, -w, v[Idef[k]], v[j + pc]);
      T_(Idef[k], j) = w;
    }
    double w = vdot(dim, v[j], v[j + pc]);
    vaxpy(dim, -w, v[j], v[j + pc]);
    T_(j, j) = w;
    for (int k = 0; k < nd; ++k) T_(j, Idef[k]) = T_(Idef[k], j);
    if (j >= 4 * p && j % p == 0) {
      double *Hp = (double *)malloc(sizeof(double) * (size_t)j * j);
      for (int c = 0; c < j; ++c) for (int r = 0; r < j; ++r) Hp[r + (size_t)c * j] = T_(r, c);
      double *ev = (double *)malloc(sizeof(double) * (size_t)j);
      orc_symeig(j, Hp, ev, NULL);
      double test = fabs(ev[0] - ev_old);
      ev_old = ev[0];
      free(Hp); free(ev);
      if (test < accur_bl) break;
    }
  }
  int M = j;
  *M_out = M;
  if (M > 0) {
    double *Hp = (double *)malloc(sizeof(double) * (size_t)M * M);
    double *U = (double *)malloc(sizeof(double) * (size_t)M * M);
    for (int c = 0; c < M; ++c) for (int r = 0; r < M; ++r) Hp[r + (size_t)c * M] = T_(r, c);
    orc_symeig(M, Hp, evals, U);
    int kmax = (p < M) ? p : M;
    for (int c = 0; c < M; ++c)
      for (int i = 0; i < p; ++i) {
        double s = 0.0;
        for (int k = 0; k < kmax; ++k) s += P_(i, k) * U[k + (size_t)c * M];
        Q[i + (size_t)c * p] = s;
      }
    free(Hp); free(U);
  }
#undef T_
#undef P_
  for (int i = 0; i <= M0p; ++i) free(v[i]);
  free(v); free(Idef); free(def); free(t); free(P);
  happly_free(&H);
  return 0;
}

/* Q_matrix::streamline / streamlineQ (src_ed/Q_matrix.hpp:181-223, :286-316).
   v: L x M column-major. Returns the new M. */
int orc_qmatrix_streamline(int L, int M, const double *e, const double *v, double wtol, double vtol,
                           double *e2, double *v2) {
  if (vtol < 1e-12) {
    memcpy(e2, e, sizeof(double) * (size_t)M);
    memcpy(v2, v, sizeof(double) * (size_t)M * L);
    return M;
  }
  int i2 = 0;
  for (int i = 0; i < M; ++i) {
    int j;
    for (j = i + 1; j < M; ++j) if (fabs(e[j] - e[i]) > wtol) break;
    int n = j - i;
    if (n == 1) {
      double s = 0.0;
      for (int a = 0; a < L; ++a) s += v[a + (size_t)i * L] * v[a + (size_t)i * L];
      if (!(s < vtol)) {
        e2[i2] = e[i];
        memcpy(&v2[(size_t)i2 * L], &v[(size_t)i * L], sizeof(double) * (size_t)L);
        i2++;
      }
    } else {
      double *QQ = (double *)calloc((size_t)L * L, sizeof(double));
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += v[a + (size_t)(i + k) * L] * v[b + (size_t)(i + k) * L];
          QQ[a + (size_t)b * L] = s;
        }
      double *d = (double *)malloc(sizeof(double) * (size_t)L);
      double *U = (double *)malloc(sizeof(double) * (size_t)L * L);
      orc_symeig(L, QQ, d, U);
      for (int c = 0; c < L; ++c) {
        if (d[c] > vtol) {
          e2[i2] = e[i];
          double sq = sqrt(d[c]);
          for (int a = 0; a < L; ++a) v2[a + (size_t)i2 * L] = U[a + (size_t)c * L] * sq;
          i2++;
        }
      }
      free(QQ); free(d); free(U);
    }
    i += n - 1;
  }
  return i2;
}

/* Q_matrix::check_norm (src_ed/Q_matrix.hpp:236-260): summed squared
   deviation of Q Q^T from the identity (the quantity compared to accur_Q_matrix). */
double orc_qmatrix_norm_error(int L, int M, const double *v) {
  double tot = 0.0;
  for (int i = 0; i < L; ++i) {
    double z = 0.0;
    for (int k = 0; k < M; ++k) z += v[i + (size_t)k * L] * v[i + (size_t)k * L];
    tot += (z - 1.0) * (z - 1.0);
    for (int j = 0; j < i; ++j) {
      double zz = 0.0;
      for (int k = 0; k < M; ++k) zz += v[i + (size_t)k * L] * v[j + (size_t)k * L];
      tot += 2.0 * zz * zz;
    }
  }
  return tot;
}

/* VDVH_naive, real V (src_util/VDVH_kernel.cpp:28-37) with the u_i of
   Q_matrix::Green_function (src_ed/Q_matrix.hpp:134-151).
   G: nz blocks of L x L complex (interleaved re,im), column-major, overwritten. */
int orc_vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G) {
  for (int q = 0; q < nz; ++q) {
    double complex zz = z[2 * q] + CI * z[2 * q + 1];
    double complex *Gq = (double complex *)(G + (size_t)q * 2 * L * L);
    for (int k = 0; k < L * L; ++k) Gq[k] = 0.0;
    for (int i = 0; i < M; ++i) {
      double complex tmp = zz - e[i];
      double complex u = conj(tmp) / (creal(tmp) * creal(tmp) + cimag(tmp) * cimag(tmp));
      for (int a = 0; a < L; ++a) {
        double complex vai = V[a + (size_t)i * L] * u;
        for (int b = 0; b < L; ++b) Gq[b + a * L] += vai * V[b + (size_t)i * L];
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Davidson-Liu (src_ed/Hamiltonian/Davidson.hpp:38-230)                     */
/* evecs: K vectors of length dim.                                           */
/* ------------------------------------------------------------------------ */
static int gram_schmidt(double **b, int nb, size_t dim, double accur) {
  /* src_util/vector_num.hpp:217-235 with k = 0; returns the new count */
  for (int i = 0; i < nb; ++i) {
    double z = sqrt(vdot(dim, b[i], b[i]));
    if (z < accur) {
      double *t = b[i]; b[i] = b[nb - 1]; b[nb - 1] = t;
      free(b[nb - 1]); b[nb - 1] = NULL;
      nb--; i--;
      continue;
    }
    vscal(dim, 1.0 / z, b[i]);
    for (int j = i + 1; j < nb; ++j) {
      double zz = vdot(dim, b[i], b[j]);
      vaxpy(dim, -zz, b[i], b[j]);
    }
  }
  return nb;
}

int orc_davidson(const orc_model *m, int Nup, int Ndw, int K, double accuracy, int max_iter, double *evals,
                 double *evecs, int *niter) {
  happly H;
  int rc = happly_make(&H, m, Nup, Ndw);
  if (rc) return rc;
  size_t dim = H.B.dim;
  if ((size_t)K > dim) K = (int)dim;
  /* diagonal of H: the full one (one-body diagonals included) */
  double *d = H.dg;
  /* find_lowest (src_util/vector_num.hpp:244-267) */
  size_t *low = (size_t *)malloc(sizeof(size_t) * (size_t)K);
  double *minv = (double *)malloc(sizeof(double) * (size_t)K);
  for (int i = 0; i < K; ++i) minv[i] = 1.0e12 + 1.0 * i;
  for (size_t i = 0; i < dim; ++i) {
    int j;
    for (j = K - 1; j >= 0; j--) if (d[i] > minv[j]) break;
    j++;
    if (j < K) {
      for (int k = K - 1; k > j; k--) { minv[k] = minv[k - 1]; low[k] = low[k - 1]; }
      minv[j] = d[i]; low[j] = i;
    }
  }
  int cap = 3 * K + 2 * K + 2;
  double **b = (double **)calloc((size_t)cap, sizeof(double *));
  double **h = (double **)calloc((size_t)cap, sizeof(double *));
  double **x = (double **)calloc((size_t)K, sizeof(double *));
  int nb = K;
  for (int k = 0; k < K; ++k) {
    b[k] = (double *)calloc(dim, sizeof(double));
    h[k] = (double *)calloc(dim, sizeof(double));
    x[k] = (double *)calloc(dim, sizeof(double));
    b[k][low[k]] = 1.0;
    happly_run(&H, b[k], h[k]);
  }
  double *bhb = (double *)calloc((size_t)cap * cap, sizeof(double));
  for (int i = 0; i < nb; ++i) for (int j = 0; j < nb; ++j) bhb[i + (size_t)j * nb] = h[j][low[i]];
  double *res = (double *)malloc(sizeof(double) * (size_t)K);
  for (int k = 0; k < K; ++k) evals[k] = 0.0;
  int iter, nh = nb, status = 0;
  for (iter = 0; iter < max_iter; ++iter) {
    double *mu = (double *)malloc(sizeof(double) * (size_t)nb);
    double *a = (double *)malloc(sizeof(double) * (size_t)nb * nb);
    orc_symeig(nb, bhb, mu, a);
    for (int k = 0; k < K; ++k) {
      memset(x[k], 0, sizeof(double) * dim);
      for (int q = 0; q < nb; ++q) vaxpy(dim, a[q + (size_t)k * nb], b[q], x[k]);
    }
    double **r = (double **)calloc((size_t)K, sizeof(double *));
    int Q = 0;
    int *rk = (int *)malloc(sizeof(int) * (size_t)K);
    for (int k = 0; k < K; ++k) {
      double *rt = (double *)calloc(dim, sizeof(double));
      for (int q = 0; q < nb; ++q) vaxpy(dim, a[q + (size_t)k * nb], h[q], rt);
      vaxpy(dim, -mu[k], x[k], rt);
      res[k] = sqrt(vdot(dim, rt, rt));
      if (res[k] > accuracy) { r[Q] = rt; rk[Q] = k; Q++; }
      else free(rt);
    }
    if (Q == 0) { free(mu); free(a); free(r); free(rk); break; }
    for (int i = nb; i < K + 2 * Q; ++i) b[i] = (double *)calloc(dim, sizeof(double));
    if (nb < K + 2 * Q) nb = K + 2 * Q;
    /* push unconverged current b[k] to positions K..K+Q-1 (map order = k ascending) */
    for (int c = 0; c < Q; ++c) { double *t = b[rk[c]]; b[rk[c]] = b[K + c]; b[K + c] = t; }
    for (int k = 0; k < K; ++k) { double *t = b[k]; b[k] = x[k]; x[k] = t; }
    for (int c = 0; c < Q; ++c) {
      int k = rk[c];
      for (size_t i = 0; i < dim; ++i) r[c][i] /= (-d[i] + mu[k] + TINY_DAV);
      free(b[K + Q + c]);
      b[K + Q + c] = r[c];
    }
    for (int i = K + 2 * Q; i < nb; ++i) { free(b[i]); b[i] = NULL; }
    nb = K + 2 * Q;
    nb = gram_schmidt(b, nb, dim, TINY_DAV);
    free(r); free(rk);
    if (nb == K) { free(mu); free(a); break; }
    for (int i = 0; i < nh; ++i) { free(h[i]); h[i] = NULL; }
    for (int k = 0; k < nb; ++k) { h[k] = (double *)calloc(dim, sizeof(double)); happly_run(&H, b[k], h[k]); }
    nh = nb;
    for (int i = 0; i < nb; ++i) {
      bhb[i + (size_t)i * nb] = vdot(dim, b[i], h[i]);
      for (int k = 0; k < i; ++k) {
        double zz = vdot(dim, b[i], h[k]);
        bhb[i + (size_t)k * nb] = zz;
        bhb[k + (size_t)i * nb] = zz;
      }
    }
    for (int k = 0; k < K; ++k) evals[k] = mu[k];
    free(mu); free(a);
  }
  if (iter == max_iter) status = -11;
  for (int k = 0; k < K; ++k) memcpy(evecs + (size_t)k * dim, x[k], sizeof(double) * dim);
  *niter = iter;
  for (int i = 0; i < cap; ++i) { free(b[i]); free(h[i]); }
  for (int k = 0; k < K; ++k) free(x[k]);
  free(b); free(h); free(x); free(bhb); free(res); free(low); free(minv);
  happly_free(&H);
  return status;
}

/* ------------------------------------------------------------------------ */
/* block Lanczos, QR variant (src_ed/Hamiltonian/Lanczos.hpp:426-572).       */
/* q0: p orthonormal vectors.  A, B: up to M0 floors of p x p column-major.  */
/* ------------------------------------------------------------------------ */
int orc_block_lanczos(const orc_model *m, int Nup, int Ndw, int p, const double *q0, int M0, int max_iter_bl,
                      double accur_deflation, double accur_bl, int *n_floors, double *A, double *B) {
  return orc_block_lanczos_v(m, Nup, Ndw, p, q0, M0, max_iter_bl, accur_deflation, accur_bl, 1, n_floors, A, B);
}

/* qr = 1: blockLanczos (Lanczos.hpp:426-572), B_j upper triangular from
   modified Gram-Schmidt; qr = 0: blockLanczosSVD (Lanczos.hpp:603-779), the
   polar form: G = Z^T Z = V D V^T, B_j = V D^(1/2) V^T, Q_{j+1} = Z V D^(-1/2) V^T,
   deflation when some D_m < accur_deflation^2, checked before the window moves
   and before the convergence test (which then follows the move). */
int orc_block_lanczos_v(const orc_model *m, int Nup, int Ndw, int p, const double *q0, int M0, int max_iter_bl,
                        double accur_deflation, double accur_bl, int qr, int *n_floors, double *A, double *B) {
  happly H;
  int rc = happly_make(&H, m, Nup, Ndw);
  if (rc) return rc;
  size_t dim = H.B.dim;
  if (M0 > max_iter_bl) M0 = max_iter_bl;
  double *Qp = (double *)calloc((size_t)p * dim, sizeof(double));
  double *Qc = (double *)malloc(sizeof(double) * (size_t)p * dim);
  double *Z = (double *)calloc((size_t)p * dim, sizeof(double));
  memcpy(Qc, q0, sizeof(double) * (size_t)p * dim);
#define VEC(X, k) ((X) + (size_t)(k) * dim)
  for (int l = 0; l < p; ++l) {
    for (int k = 0; k < l; ++k) { double z = vdot(dim, VEC(Qc, k), VEC(Qc, l)); vaxpy(dim, -z, VEC(Qc, k), VEC(Qc, l)); }
    double nrm = sqrt(vdot(dim, VEC(Qc, l), VEC(Qc, l)));
    if (nrm < accur_deflation) { free(Qp); free(Qc); free(Z); happly_free(&H); return -12; }
    vscal(dim, 1.0 / nrm, VEC(Qc, l));
  }
  double ev_old = 1e12;
  int j;
  int pp = p * p;
  for (j = 0; j < M0; ++j) {
    memset(Z, 0, sizeof(double) * (size_t)p * dim);
    for (int k = 0; k < p; ++k) happly_run(&H, VEC(Qc, k), VEC(Z, k));
    double *Aj = A + (size_t)j * pp;
    for (int l = 0; l < p; ++l)
      for (int k = 0; k <= l; ++k) {
        double z = vdot(dim, VEC(Qc, k), VEC(Z, l));
        Aj[k + l * p] = z; Aj[l + k * p] = z;
      }
    for (int l = 0; l < p; ++l) for (int k = 0; k < p; ++k) vaxpy(dim, -Aj[k + l * p], VEC(Qc, k), VEC(Z, l));
    if (j > 0) {
      const double *Bp = B + (size_t)(j - 1) * pp;
      for (int l = 0; l < p; ++l) for (int k = 0; k < p; ++k) vaxpy(dim, -Bp[l + k * p], VEC(Qp, k), VEC(Z, l));
    }
    double *Bj = B + (size_t)j * pp;
    memset(Bj, 0, sizeof(double) * (size_t)pp);
    int singular = 0;
    if (qr) {
      for (int l = 0; l < p; ++l) {
        for (int k = 0; k < l; ++k) {
          double z = vdot(dim, VEC(Z, k), VEC(Z, l));
          Bj[k + l * p] = z;
          vaxpy(dim, -z, VEC(Z, k), VEC(Z, l));
        }
        double nrm = sqrt(vdot(dim, VEC(Z, l), VEC(Z, l)));
        if (nrm < accur_deflation) {
          Bj[l + l * p] = 0.0;
          for (int ll = l; ll < p; ++ll) memset(VEC(Z, ll), 0, sizeof(double) * dim);
          singular = 1;
          break;
        }
        Bj[l + l * p] = nrm;
        vscal(dim, 1.0 / nrm, VEC(Z, l));
      }
    } else {
      double *G = (double *)malloc(sizeof(double) * (size_t)pp);
      double *D = (double *)malloc(sizeof(double) * (size_t)p);
      double *V = (double *)malloc(sizeof(double) * (size_t)pp);
      for (int l = 0; l < p; ++l)
        for (int k = 0; k <= l; ++k) G[k + l * p] = G[l + k * p] = vdot(dim, VEC(Z, k), VEC(Z, l));
      orc_symeig(p, G, D, V);
      for (int q = 0; q < p; ++q)
        if (D[q] < accur_deflation * accur_deflation) singular = 1;
      for (int r = 0; r < p; ++r)
        for (int c = 0; c < p; ++c) {
          double t = 0.0;
          for (int q = 0; q < p; ++q) t += V[r + q * p] * sqrt(D[q] > 0.0 ? D[q] : 0.0) * V[c + q * p];
          Bj[r + c * p] = t;
        }
      if (!singular) {
        /* Q_{j+1} = Z Binv, written over Q_{j-1}, then the window moves */
        memset(Qp, 0, sizeof(double) * (size_t)p * dim);
        for (int l = 0; l < p; ++l)
          for (int k = 0; k < p; ++k) {
            double t = 0.0;
            for (int q = 0; q < p; ++q) t += V[k + q * p] * V[l + q * p] / sqrt(D[q]);
            vaxpy(dim, t, VEC(Z, k), VEC(Qp, l));
          }
        double *tmp = Qp; Qp = Qc; Qc = tmp;
      }
      free(G); free(D); free(V);
      if (singular) break;
    }
    int conv = 0;
    if ((j + 1) >= 2 * p && (j + 1) % p == 0) {
      int sz = (j + 1) * p;
      double *T = (double *)calloc((size_t)sz * sz, sizeof(double));
      for (int jj = 0; jj <= j; ++jj)
        for (int r = 0; r < p; ++r)
          for (int c = 0; c < p; ++c) T[(jj * p + r) + (size_t)(jj * p + c) * sz] = A[(size_t)jj * pp + r + c * p];
      for (int jj = 0; jj < j; ++jj)
        for (int r = 0; r < p; ++r)
          for (int c = 0; c < p; ++c) {
            double bv = B[(size_t)jj * pp + r + c * p];
            T[((jj + 1) * p + r) + (size_t)(jj * p + c) * sz] = bv;
            T[(jj * p + c) + (size_t)((jj + 1) * p + r) * sz] = bv;
          }
      double *ev = (double *)malloc(sizeof(double) * (size_t)sz);
      orc_symeig(sz, T, ev, NULL);
      double delta = fabs(ev[0] - ev_old);
      ev_old = ev[0];
      free(T); free(ev);
      if (delta < accur_bl) conv = 1;
    }
    if (conv) break;
    if (singular) break;
    if (qr) {
      memcpy(Qp, Qc, sizeof(double) * (size_t)p * dim);
      memcpy(Qc, Z, sizeof(double) * (size_t)p * dim);
    }
  }
#undef VEC
  *n_floors = (j < M0) ? j + 1 : M0;
  free(Qp); free(Qc); free(Z);
  happly_free(&H);
  return 0;
}

/* matrix_continued_fraction::evaluate (src_ed/matrix_continued_fraction.hpp:139-182)
   A, B: nfl floors p x p (already energy-shifted), W: p x q.  G: nz blocks q x q. */
static int cinv(int n, double complex *a) {
  /* Gauss-Jordan with partial pivoting, column-major in place */
  int *piv = (int *)malloc(sizeof(int) * (size_t)n);
  double complex *inv = (double complex *)calloc((size_t)n * n, sizeof(double complex));
  for (int i = 0; i < n; ++i) inv[i + i * n] = 1.0;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    double best = cabs(a[c + c * n]);
    for (int r = c + 1; r < n; ++r) if (cabs(a[r + c * n]) > best) { best = cabs(a[r + c * n]); pr = r; }
    if (best == 0.0) { free(piv); free(inv); return -1; }
    if (pr != c)
      for (int k = 0; k < n; ++k) {
        double complex t = a[c + k * n]; a[c + k * n] = a[pr + k * n]; a[pr + k * n] = t;
        t = inv[c + k * n]; inv[c + k * n] = inv[pr + k * n]; inv[pr + k * n] = t;
      }
    double complex d = 1.0 / a[c + c * n];
    for (int k = 0; k < n; ++k) { a[c + k * n] *= d; inv[c + k * n] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double complex f = a[r + c * n];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) { a[r + k * n] -= f * a[c + k * n]; inv[r + k * n] -= f * inv[c + k * n]; }
    }
  }
  memcpy(a, inv, sizeof(double complex) * (size_t)n * n);
  free(piv); free(inv);
  return 0;
}

int orc_mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q, int nz, const double *z,
                 double *G) {
  double complex *F = (double complex *)malloc(sizeof(double complex) * (size_t)p * p);
  double complex *D = (double complex *)malloc(sizeof(double complex) * (size_t)p * p);
  double complex *tmp = (double complex *)malloc(sizeof(double complex) * (size_t)p * p);
  for (int iz = 0; iz < nz; ++iz) {
    double complex zz = z[2 * iz] + CI * z[2 * iz + 1];
    for (int k = 0; k < p * p; ++k) F[k] = 0.0;
    for (int j = nfl - 1; j >= 0; --j) {
      const double *Aj = A + (size_t)j * p * p;
      const double *Bj = B + (size_t)j * p * p;
      for (int k = 0; k < p * p; ++k) D[k] = -Aj[k];
      for (int k = 0; k < p; ++k) D[k + k * p] += zz;
      if (j < nfl - 1) {
        /* tmp = F Bj ; D -= Bj^T tmp */
        for (int c = 0; c < p; ++c)
          for (int r = 0; r < p; ++r) {
            double complex s = 0.0;
            for (int k = 0; k < p; ++k) s += F[r + k * p] * Bj[k + c * p];
            tmp[r + c * p] = s;
          }
        for (int c = 0; c < p; ++c)
          for (int r = 0; r < p; ++r) {
            double complex s = 0.0;
            for (int k = 0; k < p; ++k) s += Bj[k + r * p] * tmp[k + c * p];
            D[r + c * p] -= s;
          }
      }
      if (cinv(p, D)) { free(F); free(D); free(tmp); return -1; }
      memcpy(F, D, sizeof(double complex) * (size_t)p * p);
    }
    double complex *Gz = (double complex *)(G + (size_t)iz * 2 * q * q);
    for (int c = 0; c < q; ++c)
      for (int r = 0; r < q; ++r) {
        double complex s = 0.0;
        for (int k = 0; k < p; ++k)
          for (int l = 0; l < p; ++l) s += W[k + r * p] * F[k + l * p] * W[l + c * p];
        Gz[r + c * q] = s;
      }
  }
  free(F); free(D); free(tmp);
  return 0;
}

int orc_openmp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
