// ed_batch.hip -- one workgroup per work item: a whole Lanczos ground state or
// a whole LanczosGreen continued fraction of one small sector, so that a
// sweep over hundreds of parameter points x sectors is a single launch
// instead of hundreds of thousands of tiny ones.  An item whose sector fits
// (tables, diagonal, Lanczos vectors and tridiagonal arrays within
// kBatchLdsCap) runs entirely from LDS; larger items keep their vectors in
// global memory.  Every reduction is a workgroup reduction; the projected
// tridiagonal problem is solved on the device by wave 0.
#include "ed_batch.hpp"

#include <cfloat>

#include "ed_rng.hpp"

namespace ed {
namespace {

constexpr int kW = 64;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kW);
  return v;
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kW));
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kW));
  return v;
}

// workgroup sum, the result in every thread
template <int BS>
__device__ __forceinline__ double bsum(double v, double *sh) {
  v = wsum(v);
  if (BS == kW) return v;
  const int lane = threadIdx.x & (kW - 1), w = threadIdx.x / kW;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < BS / kW; ++i) r += sh[i];
  return r;
}

// 1/x: the hardware estimate refined by one Newton step (relative error a
// few ulp); used only where a sign depends on it (Sturm counts)
__device__ __forceinline__ double frcp(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

// i_dn of label I = i_up + du * i_dn, from a double-precision reciprocal of
// du corrected to the exact quotient
__device__ __forceinline__ uint32_t label_dn(uint32_t I, uint32_t du, double rdu) {
  uint32_t q = (uint32_t)((double)I * rdu);
  const uint64_t qd = (uint64_t)q * du;
  if (qd > I) --q;
  else if (I - qd >= du) ++q;
  return q;
}

// walks the labels I = t, t + BS, ... of one thread with their (i_dn, i_up)
struct LabelWalk {
  uint32_t I, id, iu, sd, su, du;
  __device__ __forceinline__ LabelWalk(uint32_t t, uint32_t step, uint32_t du_, double rdu) : I(t), du(du_) {
    id = label_dn(t, du, rdu);
    iu = t - id * du;
    sd = label_dn(step, du, rdu);
    su = step - sd * du;
  }
  __device__ __forceinline__ void next(uint32_t step) {
    I += step;
    iu += su;
    id += sd;
    if (iu >= du) {
      iu -= du;
      ++id;
    }
  }
};

// (H x)[I] for one label I = iu + du * id
__device__ __forceinline__ double hx(const BatchSector &S, const double *x, uint32_t I, uint32_t id, uint32_t iu) {
  double acc = S.diag[I] * x[I];
  const double *xr = x + (size_t)S.du * id;
  for (uint32_t e = S.up_ptr[iu]; e < S.up_ptr[iu + 1]; ++e) acc += S.up_val[e] * xr[S.up_col[e]];
  const double *xc = x + iu;
  for (uint32_t e = S.dn_ptr[id]; e < S.dn_ptr[id + 1]; ++e) acc += S.dn_val[e] * xc[(size_t)S.du * S.dn_col[e]];
  return acc;
}

// copies a sector's diagonal and one-body tables to LDS (doubles at d, words
// at u, in the order batch_lds_bytes counts them) and returns the LDS view;
// the caller synchronizes before use
template <int BS>
__device__ BatchSector stage_sector(const BatchSector &G, double *d, uint32_t *u) {
  const uint32_t dim = G.du * G.dd, nu = G.up_ptr[G.du], nd = G.dn_ptr[G.dd];
  double *diag = d, *uval = d + dim, *dval = uval + nu;
  uint32_t *uptr = u, *ucol = uptr + G.du + 1, *dptr = ucol + nu, *dcol = dptr + G.dd + 1;
  for (uint32_t i = threadIdx.x; i < dim; i += BS) diag[i] = G.diag[i];
  for (uint32_t i = threadIdx.x; i < nu; i += BS) {
    uval[i] = G.up_val[i];
    ucol[i] = G.up_col[i];
  }
  for (uint32_t i = threadIdx.x; i < nd; i += BS) {
    dval[i] = G.dn_val[i];
    dcol[i] = G.dn_col[i];
  }
  for (uint32_t i = threadIdx.x; i <= G.du; i += BS) uptr[i] = G.up_ptr[i];
  for (uint32_t i = threadIdx.x; i <= G.dd; i += BS) dptr[i] = G.dn_ptr[i];
  BatchSector L;
  L.du = G.du;
  L.dd = G.dd;
  L.diag = diag;
  L.up_val = uval;
  L.dn_val = dval;
  L.up_ptr = uptr;
  L.up_col = ucol;
  L.dn_ptr = dptr;
  L.dn_col = dcol;
  return L;
}

// One Lanczos step on (v, vp), in place: vp <- w = H v - bprev vp - alpha v,
// alpha = <Hv - bprev vp, v>, beta = ||w|| (H v reads only v, and vp[I] only
// at I, so w can overwrite vp).  Returns alpha and beta in every thread; the
// caller scales vp by 1 / beta and swaps the two names.
template <int BS>
__device__ __forceinline__ void lanczos_step(const BatchSector &S, double rdu, uint32_t dim, const double *v,
                                             double *vp, double bprev, double *sh, double &alpha, double &beta) {
  double dot = 0.0;
  for (LabelWalk L(threadIdx.x, BS, S.du, rdu); L.I < dim; L.next(BS)) {
    const double t = hx(S, v, L.I, L.id, L.iu) - bprev * vp[L.I];
    vp[L.I] = t;
    dot += t * v[L.I];
  }
  alpha = bsum<BS>(dot, sh);
  double n2 = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double t = vp[I] - alpha * v[I];
    vp[I] = t;
    n2 += t * t;
  }
  beta = sqrt(bsum<BS>(n2, sh));
}

// number of eigenvalues of the tridiagonal (a, b^2 = bb) below x
__device__ __forceinline__ int sturm_below(int n, const double *a, const double *bb, double x) {
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = -1e-300;
    q = fma(-bb[i - 1], frcp(q), a[i] - x);
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Lowest eigenpair of the tridiagonal (a[0..n), b[0..n-1), bb = b^2), run by
// the 64 lanes of wave 0.
//  * Bracket [lo, hi] of the lowest eigenvalue by Sturm counts: Gershgorin
//    below; above, `*hi_io` when it is a number (an upper bound from the
//    previous, one row shorter tridiagonal: the lowest Ritz value never
//    increases as the tridiagonal grows), probed first on a geometric grid
//    so that a nearly converged value is bracketed in one round; then
//    multisection (64 counts per round, the interval shrinks 65x) until its
//    width is below 1e-8 of the norm.
//  * Lane 0: two sweeps of inverse iteration from (1, ..., 1) with the shift
//    sigma = lo, below the spectrum, so that the unpivoted LDL^T of T - sigma
//    is positive definite; the recurrences carry their last value in a
//    register.  The eigenvector's contamination shrinks as
//    ((lambda - sigma) / gap)^2, and the Rayleigh quotient z^T T z gives the
//    eigenvalue to rounding.
// z[0..n) receives the normalized eigenvector; lane 0 writes *lam (the
// Rayleigh quotient), *zlast and *hi_io (an upper bound of the eigenvalue
// for the next call).  D and l are n-long scratch arrays.
__device__ __forceinline__ void tri_lowest_wave(int n, const double *a, const double *b, const double *bb,
                                                double *hi_io, double *z, double *D, double *l, double *lam,
                                                double *zlast) {
  const int lane = threadIdx.x;
  if (n == 1) {
    // a one-step space: the eigenpair is a_0 and the unit vector (a zero
    // a_0 would otherwise leave no scale for the shifted factorization)
    if (lane == 0) {
      z[0] = 1.0;
      *lam = a[0];
      *zlast = 1.0;
      *hi_io = a[0] + 4.0 * DBL_EPSILON * fabs(a[0]) + 1e-300;
    }
    return;
  }
  double lo = 1e300, hi = -1e300, nrm = 0.0;
  for (int i = lane; i < n; i += kW) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
    nrm = fmax(nrm, fabs(a[i]) + r);
  }
  lo = wmin(lo);
  hi = wmax(hi);
  nrm = wmax(nrm);
  const double eps = DBL_EPSILON;
  hi += 1e-300 + 4.0 * eps * nrm;
  lo -= 1e-300 + 4.0 * eps * nrm;
  const double prev = *hi_io;
  if (prev == prev && prev < hi && prev > lo) {
    hi = prev;
    // x_l = hi - (hi - lo) 2^(-0.9 l): distances from the full width down to ~1e-17 of it
    const double x = hi - (hi - lo) * exp2(-0.9 * (double)lane);
    const unsigned long long mask = __ballot(sturm_below(n, a, bb, x) >= 1);
    if (mask == 0ull) {
      lo = __shfl(x, kW - 1, kW);
    } else {
      const int f = __ffsll((long long)mask) - 1;
      hi = __shfl(x, f, kW);
      if (f > 0) lo = __shfl(x, f - 1, kW);
    }
  }
  for (int round = 0; round < 64; ++round) {
    if (hi - lo <= 1e-8 * nrm) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / (double)(kW + 1);
    const bool below = sturm_below(n, a, bb, x) >= 1;
    const unsigned long long mask = __ballot(below);
    const double x63 = __shfl(x, kW - 1, kW);
    if (mask == 0ull) {
      lo = x63;
    } else {
      const int f = __ffsll((long long)mask) - 1;
      const double xf = __shfl(x, f, kW);
      const double xl = __shfl(x, f > 0 ? f - 1 : 0, kW);
      hi = xf;
      if (f > 0) lo = xl;
    }
  }
  if (lane == 0) {
    // D holds the reciprocal pivots of T - sigma
    double sigma = lo - 4.0 * eps * nrm - 1e-300, delta = 4.0 * (double)(n > 16 ? n : 16) * eps * fmax(nrm, 1e-300);
    for (int attempt = 0; attempt < 40; ++attempt) {
      bool ok = true;
      double p = a[0] - sigma;
      for (int i = 0; i < n - 1; ++i) {
        if (!(p > 0.0)) {
          ok = false;
          break;
        }
        const double ip = 1.0 / p;
        D[i] = ip;
        l[i] = b[i] * ip;
        p = a[i + 1] - sigma - l[i] * b[i];
      }
      if (ok && p > 0.0) {
        D[n - 1] = 1.0 / p;
        break;
      }
      sigma -= delta;
      delta *= 8.0;
    }
    double scale = 1.0;
    for (int sweep = 0; sweep < 2; ++sweep) {
      // forward: y_i = x_i - l_{i-1} y_{i-1}, stored as y_i / d_i
      double y = 0.0;
      for (int i = 0; i < n; ++i) {
        const double xi = sweep == 0 ? 1.0 : scale * z[i];
        y = i > 0 ? fma(-l[i - 1], y, xi) : xi;
        z[i] = y * D[i];
      }
      // backward: x_i = y_i / d_i - l_i x_{i+1}
      double x = z[n - 1], ss = x * x;
      for (int i = n - 2; i >= 0; --i) {
        x = fma(-l[i], x, z[i]);
        z[i] = x;
        ss = fma(x, x, ss);
      }
      scale = 1.0 / sqrt(ss);
    }
    // normalize; Rayleigh quotient z^T T z
    double rq = 0.0, zp = 0.0;
    for (int i = 0; i < n; ++i) {
      const double zi = z[i] * scale;
      z[i] = zi;
      rq = fma(a[i] * zi, zi, rq);
      if (i > 0) rq = fma(2.0 * b[i - 1] * zp, zi, rq);
      zp = zi;
    }
    *lam = rq;
    *zlast = zp;
    *hi_io = fmin(hi, rq + 4.0 * eps * nrm);
  }
}

template <int BS, bool LDS>
__global__ __launch_bounds__(BS) void k_batch_gs(const BatchGSItem *items, int max_iter, double accur) {
  extern __shared__ double smem[];
  __shared__ double sh[BS / kW];
  __shared__ double s_lam, s_zl, s_hi;
  const BatchGSItem &it = items[blockIdx.x];
  const BatchSector G = it.S;
  const uint32_t dim = G.du * G.dd;
  const double rdu = 1.0 / (double)G.du;
  const int M = max_iter + 2;
  double *v, *vp, *gs, *tri;
  BatchSector S;
  if constexpr (LDS) {
    v = smem;
    vp = v + dim;
    gs = vp + dim;
    tri = gs + dim;
    const uint32_t nu = G.up_ptr[G.du], nd = G.dn_ptr[G.dd];
    double *d = tri + (size_t)kTriArrays * M;
    S = stage_sector<BS>(G, d, reinterpret_cast<uint32_t *>(d + dim + nu + nd));
  } else {
    v = it.ws;
    vp = v + dim;
    gs = it.gs;
    tri = it.tri;
    S = G;
  }
  double *al = tri, *be = al + M, *bb = be + M, *z = bb + M, *D = z + M, *l = D + M;
  // start vector (the same draws as the single-sector path)
  double nn = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double r = normal_draw(it.seed, I);
    v[I] = r;
    vp[I] = 0.0;
    nn += r * r;
  }
  const double inv0 = 1.0 / sqrt(bsum<BS>(nn, sh));
  for (uint32_t I = threadIdx.x; I < dim; I += BS) v[I] *= inv0;
  if (threadIdx.x == 0) s_hi = __builtin_nan("");
  __syncthreads();
  int k = 0, status = 0;
  double bprev = 0.0;
  for (;;) {
    double alpha, beta;
    lanczos_step<BS>(S, rdu, dim, v, vp, bprev, sh, alpha, beta);
    if (threadIdx.x == 0) {
      al[k] = alpha;
      be[k] = beta;
      bb[k] = beta * beta;
    }
    __syncthreads();
    ++k;
    if (threadIdx.x < kW) tri_lowest_wave(k, al, be, bb, &s_hi, z, D, l, &s_lam, &s_zl);
    __syncthreads();
    // a zero beta: the Krylov space is exhausted (one-state sectors end here)
    if (!(beta > 0.0) || fabs(s_zl * beta) <= accur) break;
    if (k > max_iter) {
      status = 1;
      break;
    }
    const double ib = 1.0 / beta;
    for (uint32_t I = threadIdx.x; I < dim; I += BS) vp[I] *= ib;
    double *t = v;
    v = vp;
    vp = t;
    bprev = beta;
    __syncthreads();
  }
  const double lam = s_lam;
  // eigenvector: replay the recurrence, gs = sum_j z_j v_j
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double x0 = normal_draw(it.seed, I) * inv0;
    v[I] = x0;
    vp[I] = 0.0;
    gs[I] = z[0] * x0;
  }
  __syncthreads();
  for (int j = 1; j < k; ++j) {
    const double bp = j >= 2 ? be[j - 2] : 0.0, aj = al[j - 1], ib = 1.0 / be[j - 1], zj = z[j];
    for (LabelWalk L(threadIdx.x, BS, S.du, rdu); L.I < dim; L.next(BS)) {
      const double w = (hx(S, v, L.I, L.id, L.iu) - bp * vp[L.I] - aj * v[L.I]) * ib;
      vp[L.I] = w;
      gs[L.I] += zj * w;
    }
    double *t = v;
    v = vp;
    vp = t;
    __syncthreads();
  }
  double g2 = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) g2 += gs[I] * gs[I];
  const double ig = 1.0 / sqrt(bsum<BS>(g2, sh));
  for (uint32_t I = threadIdx.x; I < dim; I += BS) it.gs[I] = gs[I] * ig;
  if (threadIdx.x == 0) {
    it.out[0] = lam;
    it.out[1] = (double)k;
    it.out[2] = (double)status;
  }
}

template <int BS, bool LDS>
__global__ __launch_bounds__(BS) void k_batch_cf(const BatchCFItem *items, int max_cf, double acc) {
  extern __shared__ double smem[];
  __shared__ double sh[BS / kW];
  // item fields in registers (the two-orbital arrays are read by name, not
  // by a run-time index, so nothing is spilled to scratch)
  const BatchCFItem &itg = items[blockIdx.x];
  const BatchSector G = itg.S;
  const double *gs = itg.gs;
  const uint32_t du_s = itg.du_s;
  const int spin = itg.spin, norb = itg.norb;
  const int32_t *map0 = itg.map[0], *map1 = itg.map[1];
  const int8_t *sgn0 = itg.sgn[0], *sgn1 = itg.sgn[1], *upar = itg.upar;
  double *av = itg.a, *bv = itg.b, *out = itg.out;
  const uint32_t dim = G.du * G.dd;
  const double rdu = 1.0 / (double)G.du;
  double *v, *vp;
  BatchSector S;
  if constexpr (LDS) {
    v = smem;
    vp = v + dim;
    const uint32_t nu = G.up_ptr[G.du], nd = G.dn_ptr[G.dd];
    double *d = vp + dim;
    S = stage_sector<BS>(G, d, reinterpret_cast<uint32_t *>(d + dim + nu + nd));
  } else {
    v = itg.ws;
    vp = v + dim;
    S = G;
  }
  // psi = c_o1 |gs> (+ c_o2 |gs>), gathered over the target sector
  double nn = 0.0;
  for (LabelWalk L(threadIdx.x, BS, G.du, rdu); L.I < dim; L.next(BS)) {
    const uint32_t id = L.id, iu = L.iu;
    double val = 0.0;
    if (spin == 0) {
      const int32_t s0 = map0[iu];
      if (s0 >= 0) val += (double)sgn0[iu] * gs[(size_t)s0 + (size_t)du_s * id];
      if (norb == 2) {
        const int32_t s1 = map1[iu];
        if (s1 >= 0) val += (double)sgn1[iu] * gs[(size_t)s1 + (size_t)du_s * id];
      }
    } else {
      const double pu = (double)upar[iu];
      const int32_t s0 = map0[id];
      if (s0 >= 0) val += (double)sgn0[id] * pu * gs[iu + (size_t)du_s * s0];
      if (norb == 2) {
        const int32_t s1 = map1[id];
        if (s1 >= 0) val += (double)sgn1[id] * pu * gs[iu + (size_t)du_s * s1];
      }
    }
    v[L.I] = val;
    vp[L.I] = 0.0;
    nn += val * val;
  }
  nn = bsum<BS>(nn, sh);
  if (!(nn > 1e-14)) {
    if (threadIdx.x == 0) {
      out[0] = -1.0;
      out[1] = nn;
    }
    return;
  }
  const double inv = 1.0 / sqrt(nn);
  for (uint32_t I = threadIdx.x; I < dim; I += BS) v[I] *= inv;
  if (threadIdx.x == 0) bv[0] = 1.0;
  __syncthreads();
  int n = 0;
  double bprev = 0.0, prod = 1.0, blast = 1.0;
  while (prod > acc && fabs(blast) > acc && n < max_cf) {
    double alpha, beta;
    lanczos_step<BS>(S, rdu, dim, v, vp, bprev, sh, alpha, beta);
    if (threadIdx.x == 0) {
      av[n] = alpha;
      bv[n + 1] = beta;
    }
    prod *= fabs(beta);
    blast = beta;
    ++n;
    if (!(prod > acc && fabs(blast) > acc && n < max_cf)) break;
    const double ib = 1.0 / beta;
    for (uint32_t I = threadIdx.x; I < dim; I += BS) vp[I] *= ib;
    double *t = v;
    v = vp;
    vp = t;
    bprev = beta;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = (double)n;
    out[1] = nn;
  }
}

// writes item k's fraction, transformed as the continued_fraction ctor
// (src_ed/continued_fraction.cpp:33-42) does, at off[fid[k]] of the store:
// a_i - E0 (electrons, pm > 0) or E0 - a_i (holes) in A; b_0 = ||psi||^2 w
// and b_i^2 (0 < i < n) in B
__global__ void k_cf_store(int nitems, int max_cf, const double *out, const double *a, const double *b,
                           const int *fid, const int *off, const double *e0, const double *w, const int8_t *pm,
                           double *A, double *B) {
  const int k = blockIdx.x;
  if (k >= nitems) return;
  const int n = (int)out[2 * (size_t)k];
  if (n <= 0) return;
  const double *ak = a + (size_t)max_cf * k, *bk = b + (size_t)(max_cf + 1) * k;
  const size_t o = (size_t)off[fid[k]];
  const double E = e0[k], b0 = out[2 * (size_t)k + 1] * w[k];
  const bool el = pm[k] > 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    A[o + i] = el ? ak[i] - E : E - ak[i];
    B[o + i] = i == 0 ? b0 : bk[i] * bk[i];
  }
}

}  // namespace

void batch_gs(const BatchGSItem *d_items, int n, int block, size_t lds, int max_iter, double accur, hipStream_t s) {
  if (n <= 0) return;
  if (lds > kBatchLdsCap) fail(ED_ERR_INVALID, "batch_gs: LDS request above the cap");
  if (lds > 0 && block == 256)
    hipLaunchKernelGGL((k_batch_gs<256, true>), dim3(n), dim3(256), lds, s, d_items, max_iter, accur);
  else if (lds > 0)
    hipLaunchKernelGGL((k_batch_gs<64, true>), dim3(n), dim3(64), lds, s, d_items, max_iter, accur);
  else if (block == 64)
    hipLaunchKernelGGL((k_batch_gs<64, false>), dim3(n), dim3(64), 0, s, d_items, max_iter, accur);
  else
    hipLaunchKernelGGL((k_batch_gs<256, false>), dim3(n), dim3(256), 0, s, d_items, max_iter, accur);
  ED_HIP(hipGetLastError());
}

void batch_cf(const BatchCFItem *d_items, int n, int block, size_t lds, int max_cf, double acc, hipStream_t s) {
  if (n <= 0) return;
  if (lds > kBatchLdsCap) fail(ED_ERR_INVALID, "batch_cf: LDS request above the cap");
  if (lds > 0 && block == 256)
    hipLaunchKernelGGL((k_batch_cf<256, true>), dim3(n), dim3(256), lds, s, d_items, max_cf, acc);
  else if (lds > 0)
    hipLaunchKernelGGL((k_batch_cf<64, true>), dim3(n), dim3(64), lds, s, d_items, max_cf, acc);
  else if (block == 64)
    hipLaunchKernelGGL((k_batch_cf<64, false>), dim3(n), dim3(64), 0, s, d_items, max_cf, acc);
  else
    hipLaunchKernelGGL((k_batch_cf<256, false>), dim3(n), dim3(256), 0, s, d_items, max_cf, acc);
  ED_HIP(hipGetLastError());
}

void batch_cf_store(int nitems, int max_cf, const double *out, const double *a, const double *b, const int *fid,
                    const int *off, const double *e0, const double *w, const int8_t *pm, double *A, double *B,
                    hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_cf_store, dim3(nitems), dim3(64), 0, s, nitems, max_cf, out, a, b, fid, off, e0, w, pm, A, B);
  ED_HIP(hipGetLastError());
}

}  // namespace ed
