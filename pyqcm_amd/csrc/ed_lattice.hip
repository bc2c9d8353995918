// ed_lattice.hip -- (frequency, wavevector) grid sums for CPT and the
// self-energy functional.  One wave per (frequency, chunk of k points), which
// works on 64 / G k points at a time: each group of G lanes (G the smallest
// power of two >= d^2, at least 16) holds one column-major d x d matrix in
// LDS, lane e < d^2 of the group element (e mod d, e / d).  Per k point the
// group builds V(k) from the Fourier elements, forms V G, and factors it with
// partial pivoting: LU for log|det(V G - 1)|, Gauss-Jordan for
// (1 - V G)^{-1}.  Each wave accumulates its chunk in registers; the chunks
// are summed in a fixed order by a second kernel, so results do not depend
// on scheduling.
#include "ed_lattice.hpp"

#include <cmath>

namespace ed {

long KGrid::count() const {
  long n = 1;
  for (int a = 0; a < dim; ++a) n *= nk[a];
  return n;
}

SuperlatticeDev::SuperlatticeDev(const Superlattice &S) : d(S.d), nnbr((int)S.R.size() / 3) {
  const int dd = d * d;
  std::vector<int> p(dd + 1, 0), nb(S.nbr.size());
  std::vector<double> v(2 * S.nbr.size());
  for (size_t e = 0; e < S.nbr.size(); ++e) ++p[S.row[e] + d * S.col[e] + 1];
  for (int q = 0; q < dd; ++q) p[q + 1] += p[q];
  std::vector<int> fill(p.begin(), p.end() - 1);
  for (size_t e = 0; e < S.nbr.size(); ++e) {
    const int at = fill[S.row[e] + d * S.col[e]]++;
    nb[at] = S.nbr[e];
    v[2 * at] = S.val[2 * e];
    v[2 * at + 1] = S.val[2 * e + 1];
  }
  V0.upload(S.V0);
  ptr.upload(p);
  nbr.upload(nb.empty() ? std::vector<int>{0} : nb);
  val.upload(v.empty() ? std::vector<double>{0.0, 0.0} : v);
  R.upload(S.R.empty() ? std::vector<double>{0.0, 0.0, 0.0} : S.R);
}

namespace {

struct Arg {
  int d, nnbr, dim, nk0, nk1, nk2;
  double off;
  long nk, kchunk;
  const double *V0, *val, *R;
  const int *ptr, *nbr;
};

__device__ __forceinline__ void cmul(double ar, double ai, double br, double bi, double &r, double &i) {
  r = ar * br - ai * bi;
  i = ar * bi + ai * br;
}

// lanes per matrix: the smallest power of two >= d^2 (at least 16); matrices per wave
__host__ __device__ constexpr int lattice_group(int d) {
  return d * d <= 16 ? 16 : (d * d <= 32 ? 32 : 64);
}

// index of the largest value over aligned groups of G lanes (ties: the smaller index)
template <int G>
__device__ __forceinline__ int group_argmax(double v, int who) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) {
    const double v2 = __shfl_xor(v, o, 64);
    const int w2 = __shfl_xor(who, o, 64);
    if (v2 > v || (v2 == v && w2 < who)) {
      v = v2;
      who = w2;
    }
  }
  return who;
}

// One wave per (frequency, chunk of k points).  The wave holds 64 / G
// matrices at once (G lanes each, lane e < d^2 of a group holding element
// (e mod d, e / d)), i.e. that many k points of the chunk per pass.
// MODE 0: part[iz][chunk] = sum over the chunk of log|det(V G - 1)|
// MODE 1: part[iz][chunk][e] = sum over the chunk of (G (1 - V G)^{-1})_e
template <int MODE, int G>
__global__ __launch_bounds__(64) void k_lattice(Arg A, const double *Gz, double *part) {
  constexpr int NQ = 64 / G;  // matrices per wave
  __shared__ double sG[2 * kLatticeMaxDim * kLatticeMaxDim];
  __shared__ double sV[2 * 64], sM[2 * 64], sB[2 * 64];
  __shared__ double sph[2 * NQ * kLatticeMaxNbr];
  const int iz = blockIdx.y, chunk = blockIdx.x, lane = threadIdx.x, d = A.d, dd = d * d;
  const int q = lane / G, e = lane - q * G;
  const bool act = e < dd;
  const int i = act ? e % d : 0, j = act ? e / d : 0;
  const int base = q * G;  // this group's offset in sV / sM / sB
  const long nchunk = gridDim.x;
  for (int t = lane; t < dd; t += 64) {
    sG[2 * t] = Gz[2 * ((size_t)iz * dd + t)];
    sG[2 * t + 1] = Gz[2 * ((size_t)iz * dd + t) + 1];
  }
  double acc = 0.0, accr = 0.0, acci = 0.0;
  const long k0 = (long)chunk * A.kchunk, k1 = k0 + A.kchunk < A.nk ? k0 + A.kchunk : A.nk;
  for (long kb = k0; kb < k1; kb += NQ) {
    // the wavevectors of the NQ groups and their phases exp(2 pi i k . R)
    for (int t = lane; t < NQ * A.nnbr; t += 64) {
      const int g = t / A.nnbr, n = t - g * A.nnbr;
      const long kk = kb + g < k1 ? kb + g : k1 - 1;
      const long a0 = kk % A.nk0, a1 = (kk / A.nk0) % A.nk1, a2 = kk / ((long)A.nk0 * A.nk1);
      const double kx = A.dim >= 1 ? ((double)a0 + A.off) / A.nk0 : 0.0;
      const double ky = A.dim >= 2 ? ((double)a1 + A.off) / A.nk1 : 0.0;
      const double kz = A.dim >= 3 ? ((double)a2 + A.off) / A.nk2 : 0.0;
      const double ang = 2.0 * M_PI * (kx * A.R[3 * n] + ky * A.R[3 * n + 1] + kz * A.R[3 * n + 2]);
      double sn, cs;
      sincos(ang, &sn, &cs);
      sph[2 * t] = cs;
      sph[2 * t + 1] = sn;
    }
    __syncthreads();
    const bool valid = kb + q < k1;
    if (act) {
      double vr = A.V0[2 * e], vi = A.V0[2 * e + 1];
      for (int r = A.ptr[e]; r < A.ptr[e + 1]; ++r) {
        const int ph = q * A.nnbr + A.nbr[r];
        double tr, ti;
        cmul(A.val[2 * r], A.val[2 * r + 1], sph[2 * ph], sph[2 * ph + 1], tr, ti);
        vr += tr;
        vi += ti;
      }
      sV[2 * (base + e)] = vr;
      sV[2 * (base + e) + 1] = vi;
    }
    __syncthreads();
    if (act) {
      double pr = 0.0, pi = 0.0;
      for (int l = 0; l < d; ++l) {
        double tr, ti;
        cmul(sV[2 * (base + i + l * d)], sV[2 * (base + i + l * d) + 1], sG[2 * (l + j * d)], sG[2 * (l + j * d) + 1],
             tr, ti);
        pr += tr;
        pi += ti;
      }
      const double one = i == j ? 1.0 : 0.0;
      sM[2 * (base + e)] = MODE == 0 ? pr - one : one - pr;
      sM[2 * (base + e) + 1] = MODE == 0 ? pi : -pi;
      if (MODE == 1) {
        sB[2 * (base + e)] = one;
        sB[2 * (base + e) + 1] = 0.0;
      }
    }
    __syncthreads();
    double logdet = 0.0;
    for (int c = 0; c < d; ++c) {
      // pivot row of this group's matrix: the largest |M(r, c)|, r >= c
      double mag = -1.0;
      if (act && j == c && i >= c)
        mag = sM[2 * (base + e)] * sM[2 * (base + e)] + sM[2 * (base + e) + 1] * sM[2 * (base + e) + 1];
      const int p = group_argmax<G>(mag, act ? i : d);
      // swap rows c and p (nothing to do where p == c)
      double mr = 0.0, mi = 0.0, br = 0.0, bi = 0.0;
      const bool sw = act && p != c && (i == c || i == p);
      if (sw) {
        const int o = base + (i == c ? p : c) + j * d;
        mr = sM[2 * o];
        mi = sM[2 * o + 1];
        if (MODE == 1) {
          br = sB[2 * o];
          bi = sB[2 * o + 1];
        }
      }
      __syncthreads();
      if (sw) {
        sM[2 * (base + e)] = mr;
        sM[2 * (base + e) + 1] = mi;
        if (MODE == 1) {
          sB[2 * (base + e)] = br;
          sB[2 * (base + e) + 1] = bi;
        }
      }
      __syncthreads();
      const int pc = base + c + c * d;
      const double qr = sM[2 * pc], qi = sM[2 * pc + 1], q2 = qr * qr + qi * qi;
      logdet += 0.5 * log(q2);
      // 1 / pivot
      const double ir = qr / q2, ii = -qi / q2;
      if (MODE == 0) {
        // LU: rows below the pivot, columns right of it
        double nr = 0.0, ni = 0.0;
        const bool up = act && i > c && j > c;
        if (up) {
          double fr, fi, tr, ti;
          cmul(sM[2 * (base + i + c * d)], sM[2 * (base + i + c * d) + 1], ir, ii, fr, fi);
          cmul(fr, fi, sM[2 * (base + c + j * d)], sM[2 * (base + c + j * d) + 1], tr, ti);
          nr = sM[2 * (base + e)] - tr;
          ni = sM[2 * (base + e) + 1] - ti;
        }
        __syncthreads();
        if (up) {
          sM[2 * (base + e)] = nr;
          sM[2 * (base + e) + 1] = ni;
        }
        __syncthreads();
      } else {
        // Gauss-Jordan on [M | B]: row c scaled by 1 / pivot, eliminated from every other row
        double nmr = 0.0, nmi = 0.0, nbr = 0.0, nbi = 0.0;
        if (act) {
          double cr, ci, dr, di;
          cmul(sM[2 * (base + c + j * d)], sM[2 * (base + c + j * d) + 1], ir, ii, cr, ci);
          cmul(sB[2 * (base + c + j * d)], sB[2 * (base + c + j * d) + 1], ir, ii, dr, di);
          if (i == c) {
            nmr = cr;
            nmi = ci;
            nbr = dr;
            nbi = di;
          } else {
            const double fr = sM[2 * (base + i + c * d)], fi = sM[2 * (base + i + c * d) + 1];
            double tr, ti;
            cmul(fr, fi, cr, ci, tr, ti);
            nmr = sM[2 * (base + e)] - tr;
            nmi = sM[2 * (base + e) + 1] - ti;
            cmul(fr, fi, dr, di, tr, ti);
            nbr = sB[2 * (base + e)] - tr;
            nbi = sB[2 * (base + e) + 1] - ti;
          }
        }
        __syncthreads();
        if (act) {
          sM[2 * (base + e)] = nmr;
          sM[2 * (base + e) + 1] = nmi;
          sB[2 * (base + e)] = nbr;
          sB[2 * (base + e) + 1] = nbi;
        }
        __syncthreads();
      }
    }
    if (MODE == 0) {
      if (valid) acc += logdet;
    } else if (act && valid) {
      // G (1 - V G)^{-1}
      double xr = 0.0, xi = 0.0;
      for (int l = 0; l < d; ++l) {
        double tr, ti;
        cmul(sG[2 * (i + l * d)], sG[2 * (i + l * d) + 1], sB[2 * (base + l + j * d)], sB[2 * (base + l + j * d) + 1],
             tr, ti);
        xr += tr;
        xi += ti;
      }
      accr += xr;
      acci += xi;
    }
    __syncthreads();
  }
  // the groups' sums, in group order
  if (MODE == 0) {
    double s = 0.0;
#pragma unroll
    for (int g = 0; g < NQ; ++g) s += __shfl(acc, g * G, 64);
    if (lane == 0) part[(size_t)iz * nchunk + chunk] = s;
  } else {
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int g = 0; g < NQ; ++g) {
      sr += __shfl(accr, g * G + e, 64);
      si += __shfl(acci, g * G + e, 64);
    }
    if (q == 0 && act) {
      const size_t o = ((size_t)iz * nchunk + chunk) * dd + e;
      part[2 * o] = sr;
      part[2 * o + 1] = si;
    }
  }
}

// out[iz][w] = scale * sum over chunks (in order) of part[iz][chunk][w]
__global__ void k_lattice_reduce(int nz, int nchunk, int width, double scale, const double *part, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz * width) return;
  const int iz = t / width, w = t - iz * width;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((size_t)iz * nchunk + c) * width + w];
  out[t] = s * scale;
}

template <int MODE>
void run(const SuperlatticeDev &S, const KGrid &K, int nz, const double *dG, double *out) {
  if (S.d < 1 || S.d > kLatticeMaxDim)
    fail(ED_ERR_UNSUPPORTED, "lattice sums: Green-function dimension " + std::to_string(S.d) + " outside 1.." +
                                 std::to_string(kLatticeMaxDim));
  if (S.nnbr > kLatticeMaxNbr) fail(ED_ERR_UNSUPPORTED, "lattice sums: more than 64 superlattice neighbors");
  if (K.dim < 0 || K.dim > 3) fail(ED_ERR_INVALID, "lattice sums: wavevector grid dimension outside 0..3");
  for (int a = 0; a < K.dim; ++a)
    if (K.nk[a] < 1) fail(ED_ERR_INVALID, "lattice sums: empty wavevector grid");
  if (nz <= 0) return;
  Arg A;
  A.d = S.d;
  A.nnbr = S.nnbr;
  A.dim = K.dim;
  A.nk0 = K.dim >= 1 ? K.nk[0] : 1;
  A.nk1 = K.dim >= 2 ? K.nk[1] : 1;
  A.nk2 = K.dim >= 3 ? K.nk[2] : 1;
  A.off = K.offset;
  A.nk = K.count();
  // enough waves to fill the device: about 4096 per launch
  long nchunk = std::max<long>(1, std::min<long>(A.nk, (4096 + nz - 1) / nz));
  A.kchunk = (A.nk + nchunk - 1) / nchunk;
  nchunk = (A.nk + A.kchunk - 1) / A.kchunk;
  if (nz > 65535) fail(ED_ERR_INVALID, "lattice sums: more than 65535 frequencies per call");
  A.V0 = S.V0.p;
  A.val = S.val.p;
  A.R = S.R.p;
  A.ptr = S.ptr.p;
  A.nbr = S.nbr.p;
  const int width = MODE == 0 ? 1 : 2 * S.d * S.d;
  DevBuf<double> part((size_t)nz * nchunk * width), dout((size_t)nz * width);
  const int G = lattice_group(S.d);
  if (G == 16)
    hipLaunchKernelGGL((k_lattice<MODE, 16>), dim3((unsigned)nchunk, (unsigned)nz), dim3(64), 0, stream(), A, dG, part.p);
  else if (G == 32)
    hipLaunchKernelGGL((k_lattice<MODE, 32>), dim3((unsigned)nchunk, (unsigned)nz), dim3(64), 0, stream(), A, dG, part.p);
  else
    hipLaunchKernelGGL((k_lattice<MODE, 64>), dim3((unsigned)nchunk, (unsigned)nz), dim3(64), 0, stream(), A, dG, part.p);
  ED_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lattice_reduce, dim3((nz * width + 255) / 256), dim3(256), 0, stream(), nz, (int)nchunk, width,
                     1.0 / (double)A.nk, part.p, dout.p);
  ED_HIP(hipGetLastError());
  dout.download(out, (size_t)nz * width);
}

}  // namespace

void lattice_logdet(const SuperlatticeDev &S, const KGrid &K, int nz, const double *dG, double *out) {
  run<0>(S, K, nz, dG, out);
}

void lattice_cpt_average(const SuperlatticeDev &S, const KGrid &K, int nz, const double *dG, double *out) {
  run<1>(S, K, nz, dG, out);
}

}  // namespace ed
