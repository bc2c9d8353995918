// ed_batch.hip -- one workgroup per work item: a whole Lanczos ground state or
// a whole LanczosGreen continued fraction of one small sector, so that a
// sweep over hundreds of parameter points x sectors is a single launch
// instead of hundreds of thousands of tiny ones.  Vectors live in global
// memory (they stay in L1/L2 at these sizes); every reduction is a
// workgroup reduction; the projected tridiagonal problem is solved on the
// device by wave 0.
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

// (H x)[I] for one label
__device__ __forceinline__ double hx(const BatchSector &S, const double *x, uint32_t I) {
  const uint32_t id = I / S.du, iu = I - id * S.du;
  double acc = S.diag[I] * x[I];
  for (uint32_t e = S.up_ptr[iu]; e < S.up_ptr[iu + 1]; ++e) acc += S.up_val[e] * x[S.up_col[e] + S.du * id];
  for (uint32_t e = S.dn_ptr[id]; e < S.dn_ptr[id + 1]; ++e)
    acc += S.dn_val[e] * x[iu + (size_t)S.du * S.dn_col[e]];
  return acc;
}

// One Lanczos step on (v, vp): w = H v - bprev vp - alpha v, alpha = <Hv - bprev vp, v>,
// beta = ||w||.  Returns alpha and beta in every thread.
template <int BS>
__device__ __forceinline__ void lanczos_step(const BatchSector &S, uint32_t dim, const double *v, const double *vp,
                                             double *w, double bprev, double *sh, double &alpha, double &beta) {
  double dot = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double t = hx(S, v, I) - bprev * vp[I];
    w[I] = t;
    dot += t * v[I];
  }
  alpha = bsum<BS>(dot, sh);
  double n2 = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double t = w[I] - alpha * v[I];
    w[I] = t;
    n2 += t * t;
  }
  beta = sqrt(bsum<BS>(n2, sh));
}

__device__ int sturm_below(int n, const double *a, const double *b, double x) {
  int cnt = 0;
  double q = a[0] - x;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = -1e-300;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Lowest eigenpair of the tridiagonal (a[0..n), b[0..n-1)), run by the 64
// lanes of wave 0: Gershgorin interval, then multisection (64 Sturm counts
// per round, the interval shrinks 65x), then inverse iteration on lane 0 with
// T - (lambda - delta), positive definite, so the unpivoted LDL^T is stable.
// z[0..n) receives the normalized eigenvector; *lam and *zlast are written by
// lane 0.
__device__ void tri_lowest_wave(int n, const double *a, const double *b, double *z, double *scratch, double *lam,
                                double *zlast) {
  const int lane = threadIdx.x;
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
  for (int round = 0; round < 64; ++round) {
    if (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + 1e-3 * eps * nrm) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / (double)(kW + 1);
    const bool below = sturm_below(n, a, b, x) >= 1;
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
  const double lm = 0.5 * (lo + hi);
  if (lane == 0) {
    double *D = scratch, *l = scratch + n;
    double delta = 4.0 * (double)(n > 16 ? n : 16) * eps * fmax(nrm, 1e-300);
    for (int attempt = 0; attempt < 40; ++attempt) {
      const double sigma = lm - delta;
      bool ok = true;
      D[0] = a[0] - sigma;
      for (int i = 0; i < n - 1; ++i) {
        if (!(D[i] > 0.0)) { ok = false; break; }
        l[i] = b[i] / D[i];
        D[i + 1] = a[i + 1] - sigma - l[i] * b[i];
      }
      if (ok && D[n - 1] > 0.0) break;
      delta *= 8.0;
    }
    for (int i = 0; i < n; ++i) z[i] = 1.0;
    for (int sweep = 0; sweep < 3; ++sweep) {
      for (int i = 1; i < n; ++i) z[i] -= l[i - 1] * z[i - 1];
      for (int i = 0; i < n; ++i) z[i] /= D[i];
      for (int i = n - 2; i >= 0; --i) z[i] -= l[i] * z[i + 1];
      double m = 0.0, s = 0.0;
      for (int i = 0; i < n; ++i) m = fmax(m, fabs(z[i]));
      for (int i = 0; i < n; ++i) { z[i] /= m; s += z[i] * z[i]; }
      s = 1.0 / sqrt(s);
      for (int i = 0; i < n; ++i) z[i] *= s;
    }
    *lam = lm;
    *zlast = z[n - 1];
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch_gs(const BatchGSItem *items, int max_iter, double accur) {
  __shared__ double sh[BS / kW];
  __shared__ double s_lam, s_zl;
  const BatchGSItem &it = items[blockIdx.x];
  const BatchSector S = it.S;
  const uint32_t dim = S.du * S.dd;
  double *v = it.ws, *vp = it.ws + dim, *w = it.ws + 2 * (size_t)dim;
  double *al = it.ab, *be = it.ab + (max_iter + 2);
  double *z = it.tri, *scr = it.tri + (max_iter + 2);
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
  __syncthreads();
  int k = 0, status = 0;
  double bprev = 0.0, lam = 0.0;
  for (;;) {
    double alpha, beta;
    lanczos_step<BS>(S, dim, v, vp, w, bprev, sh, alpha, beta);
    if (threadIdx.x == 0) {
      al[k] = alpha;
      be[k] = beta;
    }
    __syncthreads();
    ++k;
    if (threadIdx.x < kW) tri_lowest_wave(k, al, be, z, scr, &s_lam, &s_zl);
    __syncthreads();
    lam = s_lam;
    if (fabs(s_zl * beta) <= accur) break;
    if (k > max_iter) {
      status = 1;
      break;
    }
    const double ib = 1.0 / beta;
    for (uint32_t I = threadIdx.x; I < dim; I += BS) {
      vp[I] = v[I];
      v[I] = w[I] * ib;
    }
    bprev = beta;
    __syncthreads();
  }
  // eigenvector: replay the recurrence, gs = sum_j z_j v_j
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const double x0 = normal_draw(it.seed, I) * inv0;
    v[I] = x0;
    vp[I] = 0.0;
    it.gs[I] = z[0] * x0;
  }
  __syncthreads();
  for (int j = 1; j < k; ++j) {
    const double bp = j >= 2 ? be[j - 2] : 0.0, aj = al[j - 1], ib = 1.0 / be[j - 1], zj = z[j];
    for (uint32_t I = threadIdx.x; I < dim; I += BS) w[I] = (hx(S, v, I) - bp * vp[I] - aj * v[I]) * ib;
    __syncthreads();
    for (uint32_t I = threadIdx.x; I < dim; I += BS) {
      vp[I] = v[I];
      v[I] = w[I];
      it.gs[I] += zj * w[I];
    }
    __syncthreads();
  }
  double g2 = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) g2 += it.gs[I] * it.gs[I];
  const double ig = 1.0 / sqrt(bsum<BS>(g2, sh));
  for (uint32_t I = threadIdx.x; I < dim; I += BS) it.gs[I] *= ig;
  if (threadIdx.x == 0) {
    it.out[0] = lam;
    it.out[1] = (double)k;
    it.out[2] = (double)status;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_batch_cf(const BatchCFItem *items, int max_cf, double acc) {
  __shared__ double sh[BS / kW];
  // item fields in registers (the two-orbital arrays are read by name, not
  // by a run-time index, so nothing is spilled to scratch)
  const BatchCFItem &itg = items[blockIdx.x];
  const BatchSector S = itg.S;
  const double *gs = itg.gs;
  const uint32_t du_s = itg.du_s;
  const int spin = itg.spin, norb = itg.norb;
  const int32_t *map0 = itg.map[0], *map1 = itg.map[1];
  const int8_t *sgn0 = itg.sgn[0], *sgn1 = itg.sgn[1], *upar = itg.upar;
  double *ws = itg.ws, *av = itg.a, *bv = itg.b, *out = itg.out;
  const uint32_t dim = S.du * S.dd;
  double *v = ws, *vp = ws + dim, *w = ws + 2 * (size_t)dim;
  // psi = c_o1 |gs> (+ c_o2 |gs>), gathered over the target sector
  double nn = 0.0;
  for (uint32_t I = threadIdx.x; I < dim; I += BS) {
    const uint32_t id = I / S.du, iu = I - id * S.du;
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
    v[I] = val;
    vp[I] = 0.0;
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
    lanczos_step<BS>(S, dim, v, vp, w, bprev, sh, alpha, beta);
    if (threadIdx.x == 0) {
      av[n] = alpha;
      bv[n + 1] = beta;
    }
    prod *= fabs(beta);
    blast = beta;
    ++n;
    if (!(prod > acc && fabs(blast) > acc && n < max_cf)) break;
    const double ib = 1.0 / beta;
    for (uint32_t I = threadIdx.x; I < dim; I += BS) {
      vp[I] = v[I];
      v[I] = w[I] * ib;
    }
    bprev = beta;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = (double)n;
    out[1] = nn;
  }
}

}  // namespace

void batch_gs(const BatchGSItem *d_items, int n, int block, int max_iter, double accur, hipStream_t s) {
  if (n <= 0) return;
  if (block == 64)
    hipLaunchKernelGGL((k_batch_gs<64>), dim3(n), dim3(64), 0, s, d_items, max_iter, accur);
  else
    hipLaunchKernelGGL((k_batch_gs<256>), dim3(n), dim3(256), 0, s, d_items, max_iter, accur);
  ED_HIP(hipGetLastError());
}

void batch_cf(const BatchCFItem *d_items, int n, int block, int max_cf, double acc, hipStream_t s) {
  if (n <= 0) return;
  if (block == 64)
    hipLaunchKernelGGL((k_batch_cf<64>), dim3(n), dim3(64), 0, s, d_items, max_cf, acc);
  else
    hipLaunchKernelGGL((k_batch_cf<256>), dim3(n), dim3(256), 0, s, d_items, max_cf, acc);
  ED_HIP(hipGetLastError());
}

}  // namespace ed
