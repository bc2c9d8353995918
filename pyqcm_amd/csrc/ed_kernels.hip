// ed_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the ED path.
//
// All arithmetic is FP64.  H.v is HBM/cache-bandwidth bound (~3 flop/B), so no
// MFMA is used: the kernels stream FP64 vectors with coalesced accesses, stage
// rows of the factorized vector in LDS, and reduce dot products with wave64
// shuffles + LDS, writing deterministic per-block partial sums.
#include <algorithm>

#include "ed_common.hpp"
#include "ed_kernels.hpp"
#include "ed_rng.hpp"

namespace ed {
namespace {

constexpr int kWave = 64;
constexpr int kPartialsCap = 8192;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum; result valid in thread 0.  sh must hold BS/64 doubles.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < kWave) {
    r = (threadIdx.x < BS / kWave) ? sh[threadIdx.x] : 0.0;
    r = wave_sum(r);
  }
  return r;
}

__device__ __forceinline__ double half_entry(const DevHalf &h, const double *vtab, uint32_t e, uint32_t &col) {
  const uint32_t w = h.ent[e];
  if (h.packed) {
    col = w & ((1u << h.colbits) - 1u);
    return vtab[w >> h.colbits];
  }
  col = w;
  return h.val[e];
}

__device__ __forceinline__ double diag_value(const FactArgs &a, const double *Ush, size_t I, uint32_t u,
                                             uint32_t d, uint32_t sd) {
  if (a.diag_mode == 0) return a.dg[I];
  const uint32_t su = a.up.states[u];
  double s = a.up.diag[u] + a.dn.diag[d];
  uint32_t m = su & sd;
  if (a.diag_mode == 2) return s + Ush[0] * (double)__popc(m);
  while (m) {
    const int i = __ffs(m) - 1;
    s += Ush[i];
    m &= m - 1u;
  }
  return s;
}

struct FusedParams {
  FactArgs a;
  const double *x;
  double *y;
  double ax;        // y = ax * H x + beta * y
  double beta;
  double *part;
  uint32_t R;       // rows of X per block iteration
  int ntab_up, ntab_dn;  // value-table sizes staged in LDS (packed mode)
};

constexpr int kVtabMax = 256;

// Row kernel: each block owns R consecutive rows of X[i_dn][i_up]; the rows are
// staged in LDS so the up-spin part (sparse K_up applied inside a row) gathers
// from LDS; the down-spin part (K_dn couples whole rows) reads coalesced
// segments of the neighbour rows through L2/MALL.  DOWN = false gives the
// first pass of the two-pass variant (diagonal + up part only).
template <int BS, bool ONE_ROW, bool DOWN, bool STAGE>
__global__ __launch_bounds__(BS) void k_hv_rows(FusedParams P) {
  extern __shared__ double smem[];
  __shared__ double Ush[32];
  __shared__ double vtu[kVtabMax];
  __shared__ double vtd[kVtabMax];
  __shared__ double red[BS / kWave];
  const FactArgs &a = P.a;
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  const uint32_t R = ONE_ROW ? 1u : P.R;
  if (threadIdx.x < 32) Ush[threadIdx.x] = a.Uc[threadIdx.x];
  for (int i = threadIdx.x; i < P.ntab_up; i += BS) vtu[i] = a.up.val[i];
  for (int i = threadIdx.x; i < P.ntab_dn; i += BS) vtd[i] = a.dn.val[i];
  const double *vtab_u = P.ntab_up ? vtu : a.up.val;
  const double *vtab_d = P.ntab_dn ? vtd : a.dn.val;
  double dot = 0.0;
  for (uint32_t d0 = blockIdx.x * R; d0 < dd; d0 += gridDim.x * R) {
    const uint32_t nr = min(R, dd - d0);
    const size_t base = (size_t)d0 * du;
    const uint32_t n = nr * du;
    if (STAGE) {
      __syncthreads();
      for (uint32_t f = threadIdx.x; f < n; f += BS) smem[f] = P.x[base + f];
      __syncthreads();
    } else {
      __syncthreads();  // value tables staged
    }
    for (uint32_t f = threadIdx.x; f < n; f += BS) {
      const uint32_t r = ONE_ROW ? 0u : f / du;
      const uint32_t u = f - r * du;
      const uint32_t d = d0 + r;
      const uint32_t sd = (a.diag_mode != 0) ? a.dn.states[d] : 0u;
      const double xv = STAGE ? smem[f] : P.x[base + f];
      const size_t I = base + f;
      double acc = diag_value(a, Ush, I, u, d, sd) * xv;
      const double *xr = STAGE ? smem + (size_t)r * du : P.x + base + (size_t)r * du;
      const uint32_t e1 = a.up.ptr[u + 1];
      for (uint32_t e = a.up.ptr[u]; e < e1; ++e) {
        uint32_t c;
        const double v = half_entry(a.up, vtab_u, e, c);
        acc += v * xr[c];
      }
      if (DOWN) {
        const uint32_t g1 = a.dn.ptr[d + 1];
        for (uint32_t e = a.dn.ptr[d]; e < g1; ++e) {
          uint32_t c;
          const double v = half_entry(a.dn, vtab_d, e, c);
          acc += v * P.x[(size_t)c * du + u];
        }
      }
      const double out = (P.beta != 0.0) ? P.ax * acc + P.beta * P.y[I] : P.ax * acc;
      P.y[I] = out;
      if (DOWN) dot += out * xv;
    }
  }
  if (DOWN && P.part) {
    const double s = block_sum<BS>(dot, red);
    if (threadIdx.x == 0) P.part[blockIdx.x] = s;
  }
}

constexpr int kMaxNb = 128;  // K_dn entries of one row staged in LDS

// Row kernel for one row of X per block iteration (rows up to ~15K states).
// Every thread owns KMAX elements of the row (u = tid + k * BS) and keeps their
// accumulators in registers.  The up part gathers from the LDS-staged row; the
// down part runs neighbour-major: for each K_dn entry (c, v) of the row, staged
// in LDS, every thread issues KMAX independent coalesced loads of row c, so
// each wave keeps many HBM/MALL requests in flight instead of one dependent
// chain per element.
template <int BS, int KMAX, bool DOWN>
__global__ __launch_bounds__(BS) void k_hv_rowk(FusedParams P) {
  extern __shared__ double smem[];
  __shared__ double Ush[32];
  __shared__ double vtu[kVtabMax];
  __shared__ double vtd[kVtabMax];
  __shared__ double red[BS / kWave];
  __shared__ uint32_t ncol[kMaxNb];
  __shared__ double nval[kMaxNb];
  const FactArgs &a = P.a;
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  if (threadIdx.x < 32) Ush[threadIdx.x] = a.Uc[threadIdx.x];
  for (int i = threadIdx.x; i < P.ntab_up; i += BS) vtu[i] = a.up.val[i];
  for (int i = threadIdx.x; i < P.ntab_dn; i += BS) vtd[i] = a.dn.val[i];
  const double *vtab_u = P.ntab_up ? vtu : a.up.val;
  const double *vtab_d = P.ntab_dn ? vtd : a.dn.val;
  double dot = 0.0;
  for (uint32_t d = blockIdx.x; d < dd; d += gridDim.x) {
    const size_t base = (size_t)d * du;
    const uint32_t g0 = a.dn.ptr[d], nnb = a.dn.ptr[d + 1] - g0;
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < du; f += BS) smem[f] = P.x[base + f];
    if (DOWN) {
      for (uint32_t j = threadIdx.x; j < min(nnb, (uint32_t)kMaxNb); j += BS) {
        uint32_t c;
        const double v = half_entry(a.dn, vtab_d, g0 + j, c);
        ncol[j] = c;
        nval[j] = v;
      }
    }
    __syncthreads();
    const uint32_t sd = (a.diag_mode != 0) ? a.dn.states[d] : 0u;
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t u = threadIdx.x + k * BS;
      acc[k] = 0.0;
      if (u < du) {
        double s = diag_value(a, Ush, base + u, u, d, sd) * smem[u];
        const uint32_t e1 = a.up.ptr[u + 1];
        uint32_t e = a.up.ptr[u];
        for (; e + 4 <= e1; e += 4) {
          uint32_t c0, c1, c2, c3;
          const double v0 = half_entry(a.up, vtab_u, e, c0);
          const double v1 = half_entry(a.up, vtab_u, e + 1, c1);
          const double v2 = half_entry(a.up, vtab_u, e + 2, c2);
          const double v3 = half_entry(a.up, vtab_u, e + 3, c3);
          s += v0 * smem[c0] + v1 * smem[c1] + v2 * smem[c2] + v3 * smem[c3];
        }
        for (; e < e1; ++e) {
          uint32_t c;
          const double v = half_entry(a.up, vtab_u, e, c);
          s += v * smem[c];
        }
        acc[k] = s;
      }
    }
    if (DOWN) {
      uint32_t j = 0;
      for (; j + 2 <= nnb; j += 2) {
        uint32_t c0, c1;
        double v0, v1;
        if (j + 1 < (uint32_t)kMaxNb) {
          c0 = ncol[j]; v0 = nval[j]; c1 = ncol[j + 1]; v1 = nval[j + 1];
        } else {
          v0 = half_entry(a.dn, vtab_d, g0 + j, c0);
          v1 = half_entry(a.dn, vtab_d, g0 + j + 1, c1);
        }
        const double *x0 = P.x + (size_t)c0 * du;
        const double *x1 = P.x + (size_t)c1 * du;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
          const uint32_t u = threadIdx.x + k * BS;
          if (u < du) acc[k] += v0 * x0[u] + v1 * x1[u];
        }
      }
      for (; j < nnb; ++j) {
        uint32_t c0;
        double v0;
        if (j < (uint32_t)kMaxNb) { c0 = ncol[j]; v0 = nval[j]; }
        else v0 = half_entry(a.dn, vtab_d, g0 + j, c0);
        const double *x0 = P.x + (size_t)c0 * du;
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
          const uint32_t u = threadIdx.x + k * BS;
          if (u < du) acc[k] += v0 * x0[u];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const uint32_t u = threadIdx.x + k * BS;
      if (u < du) {
        const size_t I = base + u;
        const double out = (P.beta != 0.0) ? P.ax * acc[k] + P.beta * P.y[I] : P.ax * acc[k];
        P.y[I] = out;
        if (DOWN) dot += out * smem[u];
      }
    }
  }
  if (DOWN && P.part) {
    const double s = block_sum<BS>(dot, red);
    if (threadIdx.x == 0) P.part[blockIdx.x] = s;
  }
}

// x value of a 16-bit sign-only entry (bit 15 = sign, bits 0-14 = column),
// the sign applied by flipping the sign bit of the double
__device__ __forceinline__ double sell16_term(const double *smem, uint32_t e) {
  const double v = smem[e & 0x7fffu];
  return __longlong_as_double(__double_as_longlong(v) ^ ((long long)(e & 0x8000u) << 48));
}

// Up part of one sliced-ELL slice for one lane: sum_j value_j * smem[col_j].
// Entries come in lane groups of 4 words (8-byte loads for 16-bit entries,
// 16-byte loads for packed 32-bit ones); up to eight group loads are issued
// before the first LDS gather, so a wave keeps a whole slice's table reads
// in flight at once.
template <int TM>
__device__ __forceinline__ double sell_slice_sum(const DevHalf &h, const SellTable &T, const double *vtab,
                                                 const double *smem, size_t o0, size_t o1, uint32_t lane) {
  constexpr uint32_t G = 4, QB = 8;
  if (TM == 2) {
    const uint32_t nq = (uint32_t)((o1 - o0) / (kWave * G));
    const uint2 *t = reinterpret_cast<const uint2 *>(T.e16 + o0) + lane;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t q = 0; q < nq; q += QB) {
      uint2 g[QB] = {};
#pragma unroll
      for (uint32_t k = 0; k < QB; ++k)
        if (q + k < nq) g[k] = t[(size_t)(q + k) * kWave];
#pragma unroll
      for (uint32_t k = 0; k < QB; ++k) {
        if (q + k < nq) {
          s0 += sell16_term(smem, g[k].x & 0xffffu);
          s1 += sell16_term(smem, g[k].x >> 16);
          s0 += sell16_term(smem, g[k].y & 0xffffu);
          s1 += sell16_term(smem, g[k].y >> 16);
        }
      }
    }
    return (s0 + s1) * h.t0;
  } else if (TM == 0) {
    const uint32_t nq = (uint32_t)((o1 - o0) / (kWave * G));
    const uint4 *t = reinterpret_cast<const uint4 *>(T.ent + o0) + lane;
    const uint32_t mask = (1u << h.colbits) - 1u;
    double s0 = 0.0, s1 = 0.0;
    for (uint32_t q = 0; q < nq; q += QB / 2) {
      uint4 g[QB / 2] = {};
#pragma unroll
      for (uint32_t k = 0; k < QB / 2; ++k)
        if (q + k < nq) g[k] = t[(size_t)(q + k) * kWave];
#pragma unroll
      for (uint32_t k = 0; k < QB / 2; ++k) {
        if (q + k < nq) {
          s0 += vtab[g[k].x >> h.colbits] * smem[g[k].x & mask];
          s1 += vtab[g[k].y >> h.colbits] * smem[g[k].y & mask];
          s0 += vtab[g[k].z >> h.colbits] * smem[g[k].z & mask];
          s1 += vtab[g[k].w >> h.colbits] * smem[g[k].w & mask];
        }
      }
    }
    return s0 + s1;
  } else {
    double s = 0.0;
    for (size_t e = o0 + lane; e < o1; e += kWave) s += T.val[e] * smem[T.ent[e]];
    return s;
  }
}

// Row kernel with the up-spin operator in sliced-ELL form: each wave takes
// slices of 64 consecutive up states and sums their entries against the
// LDS-staged row (sell_slice_sum); loop bounds are wave-uniform.  TM is the
// table mode: 0 packed entry + value table, 1 explicit values, 2 16-bit
// sign-only.  The down part (DOWN) reads the K_dn neighbour rows (staged
// list) for the same 64 states with four independent loads in flight.
template <int BS, bool DOWN, int TM>
__global__ __launch_bounds__(BS) void k_hv_rows_sell(FusedParams P) {
  extern __shared__ double smem[];
  __shared__ double Ush[32];
  __shared__ double vtu[kVtabMax];
  __shared__ double vtd[kVtabMax];
  __shared__ double red[BS / kWave];
  __shared__ uint32_t ncol[kMaxNb];
  __shared__ double nval[kMaxNb];
  const FactArgs &a = P.a;
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  if (threadIdx.x < 32) Ush[threadIdx.x] = a.Uc[threadIdx.x];
  for (int i = threadIdx.x; i < P.ntab_up; i += BS) vtu[i] = a.up.val[i];
  for (int i = threadIdx.x; i < P.ntab_dn; i += BS) vtd[i] = a.dn.val[i];
  const double *vtab_u = P.ntab_up ? vtu : a.up.val;
  const double *vtab_d = P.ntab_dn ? vtd : a.dn.val;
  const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  constexpr uint32_t NW = BS / kWave;
  const SellTable &T = a.up.sell[DOWN ? 0 : 1];
  double dot = 0.0;
  for (uint32_t di = blockIdx.x; di < dd; di += gridDim.x) {
    const uint32_t d = a.row_order ? a.dn.order[di] : di;
    const size_t base = (size_t)d * du;
    const uint32_t g0 = a.dn.ptr[d], nnb = a.dn.ptr[d + 1] - g0;
    const uint32_t nst = min(nnb, (uint32_t)kMaxNb);
    __syncthreads();
    for (uint32_t f = threadIdx.x; f < du; f += BS) smem[f] = P.x[base + f];
    if (threadIdx.x == 0) smem[du] = 0.0;  // target of the sliced-ELL padding
    if (DOWN) {
      for (uint32_t j = threadIdx.x; j < nst; j += BS) {
        uint32_t c;
        const double v = half_entry(a.dn, vtab_d, g0 + j, c);
        ncol[j] = c;
        nval[j] = v;
      }
    }
    __syncthreads();
    const uint32_t sd = (a.diag_mode != 0) ? a.dn.states[d] : 0u;
    for (uint32_t s = wave; s < T.nslice; s += NW) {
      const uint32_t u = T.row ? T.row[s * kWave + lane] : s * kWave + lane;
      const bool ok = u < du;
      const uint32_t uu = ok ? u : du - 1;
      const size_t o0 = T.off[s], o1 = T.off[s + 1];
      double acc = diag_value(a, Ush, base + uu, uu, d, sd) * smem[uu];
      acc += sell_slice_sum<TM>(a.up, T, vtab_u, smem, o0, o1, lane);
      if (DOWN) {
        // staged neighbours: eight (then four) independent row reads in flight
        uint32_t j = 0;
        for (; j + 8 <= nst; j += 8) {
          double xv[8];
#pragma unroll
          for (int m = 0; m < 8; ++m) xv[m] = P.x[(size_t)ncol[j + m] * du + uu];
#pragma unroll
          for (int m = 0; m < 8; ++m) acc += nval[j + m] * xv[m];
        }
        for (; j + 4 <= nst; j += 4) {
          double xv[4];
#pragma unroll
          for (int m = 0; m < 4; ++m) xv[m] = P.x[(size_t)ncol[j + m] * du + uu];
#pragma unroll
          for (int m = 0; m < 4; ++m) acc += nval[j + m] * xv[m];
        }
        for (; j < nst; ++j) acc += nval[j] * P.x[(size_t)ncol[j] * du + uu];
        for (; j < nnb; ++j) {  // rows with more than kMaxNb neighbours
          uint32_t c0;
          const double v0 = half_entry(a.dn, vtab_d, g0 + j, c0);
          acc += v0 * P.x[(size_t)c0 * du + uu];
        }
      }
      if (ok) {
        const size_t I = base + u;
        const double out = (P.beta != 0.0) ? P.ax * acc + P.beta * P.y[I] : P.ax * acc;
        P.y[I] = out;
        if (DOWN) dot += out * smem[u];
      }
    }
  }
  if (DOWN && P.part) {
    const double s = block_sum<BS>(dot, red);
    if (threadIdx.x == 0) P.part[blockIdx.x] = s;
  }
}

struct SlabParams {
  FactArgs a;
  const double *x;
  double *y;
  double ax;
  double *part;
  uint32_t nslab, nchunk, chunk_rows;
  int nxcd;
  int ntab_dn;
};

// Second pass of the two-pass variant: y += (K_dn x 1) x over column slabs.
// A slab of SW columns of every row (dim_dn * SW * 8 bytes) is what the down
// part re-reads ~nnz(K_dn)/row times; blocks are mapped so that all chunks of a
// slab run on the same XCD (round-robin dispatch: block b -> XCD b % nxcd), which
// keeps the slab resident in that XCD's L2.
constexpr int kSlabEnt = 2048;  // K_dn entries of one row chunk staged in LDS
constexpr int kSlabRows = 64;   // rows per chunk

// Second pass of the two-pass variant: y += ax (K_dn x 1) x on column slabs.
// One block per (slab, 64-row chunk), numbered so that block b runs on XCD
// b % nxcd (round-robin dispatch) and each XCD receives all chunks of one slab
// before the next: the slab (dim_dn x SW doubles) stays resident in that XCD's
// L2 while each x element is re-read ~nnz(K_dn) times.  The chunk's K_dn rows
// are staged in LDS; neighbours are loaded four at a time.  NT streams the
// once-touched y and own-x accesses past L2.
template <int SW, int RL, bool NT>
__global__ __launch_bounds__(SW *RL) void k_hv_down_slab(SlabParams P) {
  __shared__ double vtd[kVtabMax];
  __shared__ double red[(SW * RL) / kWave];
  __shared__ uint32_t lptr[kSlabRows + 1];
  __shared__ uint32_t lcol[kSlabEnt];
  __shared__ double lval[kSlabEnt];
  const FactArgs &a = P.a;
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  for (int i = threadIdx.x; i < P.ntab_dn; i += SW * RL) vtd[i] = a.dn.val[i];
  const double *vtab_d = P.ntab_dn ? vtd : a.dn.val;
  const uint32_t b = blockIdx.x;
  const uint32_t xcd = b % P.nxcd, k = b / P.nxcd;
  const uint32_t slab = xcd + P.nxcd * (k / P.nchunk);
  const uint32_t chunk = k % P.nchunk;
  double dot = 0.0;
  const uint32_t r0 = min(dd, chunk * kSlabRows);
  const uint32_t r1 = min(dd, r0 + kSlabRows);
  const uint32_t g0 = a.dn.ptr[r0], g1 = a.dn.ptr[r1];
  const bool staged = (g1 - g0) <= (uint32_t)kSlabEnt;
  if (staged) {
    for (uint32_t i = threadIdx.x; i <= r1 - r0; i += SW * RL) lptr[i] = a.dn.ptr[r0 + i] - g0;
    for (uint32_t e = threadIdx.x; e < g1 - g0; e += SW * RL) {
      uint32_t c;
      lval[e] = half_entry(a.dn, vtab_d, g0 + e, c);
      lcol[e] = c;
    }
  }
  __syncthreads();
  const uint32_t u = slab * SW + (threadIdx.x % SW);
  const uint32_t rl = threadIdx.x / SW;
  if (slab < P.nslab && u < du) {
    auto finish = [&](uint32_t d, double acc) {
      const size_t I = (size_t)d * du + u;
      double out, xo;
      if (NT) {
        out = __builtin_nontemporal_load(P.y + I) + P.ax * acc;
        __builtin_nontemporal_store(out, P.y + I);
        xo = __builtin_nontemporal_load(P.x + I);
      } else {
        out = P.y[I] + P.ax * acc;
        P.y[I] = out;
        xo = P.x[I];
      }
      dot += out * xo;
    };
    if (staged) {
      // two rows per iteration: up to eight independent loads in flight
      for (uint32_t d = r0 + rl; d < r1; d += 2 * RL) {
        const uint32_t dB = d + RL;
        const bool hasB = dB < r1;
        uint32_t eA = lptr[d - r0], eA1 = lptr[d - r0 + 1];
        uint32_t eB = hasB ? lptr[dB - r0] : 0u, eB1 = hasB ? lptr[dB - r0 + 1] : 0u;
        double accA = 0.0, accB = 0.0;
        while (eA + 4 <= eA1 && eB + 4 <= eB1) {
          const double a0 = P.x[(size_t)lcol[eA] * du + u], a1 = P.x[(size_t)lcol[eA + 1] * du + u];
          const double a2 = P.x[(size_t)lcol[eA + 2] * du + u], a3 = P.x[(size_t)lcol[eA + 3] * du + u];
          const double b0 = P.x[(size_t)lcol[eB] * du + u], b1 = P.x[(size_t)lcol[eB + 1] * du + u];
          const double b2 = P.x[(size_t)lcol[eB + 2] * du + u], b3 = P.x[(size_t)lcol[eB + 3] * du + u];
          accA += lval[eA] * a0 + lval[eA + 1] * a1 + lval[eA + 2] * a2 + lval[eA + 3] * a3;
          accB += lval[eB] * b0 + lval[eB + 1] * b1 + lval[eB + 2] * b2 + lval[eB + 3] * b3;
          eA += 4;
          eB += 4;
        }
        for (; eA + 4 <= eA1; eA += 4) {
          const double a0 = P.x[(size_t)lcol[eA] * du + u], a1 = P.x[(size_t)lcol[eA + 1] * du + u];
          const double a2 = P.x[(size_t)lcol[eA + 2] * du + u], a3 = P.x[(size_t)lcol[eA + 3] * du + u];
          accA += lval[eA] * a0 + lval[eA + 1] * a1 + lval[eA + 2] * a2 + lval[eA + 3] * a3;
        }
        for (; eA < eA1; ++eA) accA += lval[eA] * P.x[(size_t)lcol[eA] * du + u];
        for (; eB + 4 <= eB1; eB += 4) {
          const double b0 = P.x[(size_t)lcol[eB] * du + u], b1 = P.x[(size_t)lcol[eB + 1] * du + u];
          const double b2 = P.x[(size_t)lcol[eB + 2] * du + u], b3 = P.x[(size_t)lcol[eB + 3] * du + u];
          accB += lval[eB] * b0 + lval[eB + 1] * b1 + lval[eB + 2] * b2 + lval[eB + 3] * b3;
        }
        for (; eB < eB1; ++eB) accB += lval[eB] * P.x[(size_t)lcol[eB] * du + u];
        finish(d, accA);
        if (hasB) finish(dB, accB);
      }
    } else {
      for (uint32_t d = r0 + rl; d < r1; d += RL) {
        double acc = 0.0;
        const uint32_t e1 = a.dn.ptr[d + 1];
        for (uint32_t e = a.dn.ptr[d]; e < e1; ++e) {
          uint32_t c;
          const double v = half_entry(a.dn, vtab_d, e, c);
          acc += v * P.x[(size_t)c * du + u];
        }
        finish(d, acc);
      }
    }
  }
  if (P.part) {
    const double s = block_sum<SW * RL>(dot, red);
    if (threadIdx.x == 0) P.part[blockIdx.x] = s;
  }
}

// Down pass of the two-pass schedule for sectors whose column slab fits in
// LDS (dim_dn x W doubles): workgroup b stages columns [b W, b W + W) of every
// row of x once, then y[d][u] += ax sum_e K_dn[d, c_e] x[c_e][u] for all rows d
// from LDS (W lanes per row: consecutive columns, conflict-free), with the
// partials of <y, x>.  Each x element is read from memory once instead of once
// per neighbour row (the fused kernel's ~17 re-reads, served by the Infinity
// Cache on the small sectors where this schedule is chosen).
template <int W>
__global__ __launch_bounds__(256) void k_hv_down_lds(SlabParams P) {
  extern __shared__ double sx[];
  __shared__ double vtd[kVtabMax];
  __shared__ double red[256 / kWave];
  const FactArgs &a = P.a;
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  for (int i = threadIdx.x; i < P.ntab_dn; i += 256) vtd[i] = a.dn.val[i];
  const double *vtab_d = P.ntab_dn ? vtd : a.dn.val;
  const uint32_t c0 = blockIdx.x * W;
  const uint32_t n = dd * W;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const uint32_t d = i / W, c = i - d * W;
    sx[i] = (c0 + c < du) ? P.x[(size_t)d * du + c0 + c] : 0.0;
  }
  __syncthreads();
  double dot = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) {
    const uint32_t d = i / W, c = i - d * W;
    if (c0 + c >= du) continue;
    double acc = 0.0;
    const uint32_t e1 = a.dn.ptr[d + 1];
    for (uint32_t e = a.dn.ptr[d]; e < e1; ++e) {
      uint32_t col;
      const double v = half_entry(a.dn, vtab_d, e, col);
      acc += v * sx[col * W + c];
    }
    const size_t I = (size_t)d * du + c0 + c;
    const double out = P.y[I] + P.ax * acc;
    P.y[I] = out;
    dot += out * sx[i];
  }
  if (P.part) {
    const double s = block_sum<256>(dot, red);
    if (threadIdx.x == 0) P.part[blockIdx.x] = s;
  }
}

// partial sums -> kPartialsCap partials (deterministic, fixed order)
template <int BS>
__global__ __launch_bounds__(BS) void k_fold_partials(const double *in, size_t n, double *out) {
  __shared__ double red[BS / kWave];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) s += in[i];
  s = block_sum<BS>(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// m rows of nb partials -> m sums in one block (fixed order), written to
// host-mapped memory; then a system-scope fence and the sequence number
// 'seq' in *flag, which the host polls instead of synchronising the stream
template <int BS>
__global__ __launch_bounds__(BS) void k_fold_rows_flag(const double *in, int nb, int m, size_t stride, double *out,
                                                       uint64_t *flag, uint64_t seq) {
  // wave w folds rows w, w + BS/64, ...: lanes take strided partials, then a
  // shuffle reduction (fixed order, no block barrier per row)
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int r = w; r < m; r += BS / kWave) {
    const double *p = in + (size_t)r * stride;
    double s = 0.0;
    for (int i = lane; i < nb; i += kWave) s += p[i];
    s = wave_sum(s);
    if (lane == 0) out[r] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_fact_diag(FactArgs a, double *d) {
  __shared__ double Ush[32];
  if (threadIdx.x < 32) Ush[threadIdx.x] = a.Uc[threadIdx.x];
  __syncthreads();
  const uint32_t du = a.up.dim;
  const size_t n = (size_t)du * a.dn.dim;
  for (size_t I = (size_t)blockIdx.x * BS + threadIdx.x; I < n; I += (size_t)gridDim.x * BS) {
    const uint32_t dn = (uint32_t)(I / du), u = (uint32_t)(I - (size_t)dn * du);
    d[I] = diag_value(a, Ush, I, u, dn, a.diag_mode ? a.dn.states[dn] : 0u);
  }
}

// ---------------------------------------------------------------------------
// CSR SpMV (full symmetric matrix), G lanes per row.
// ---------------------------------------------------------------------------
template <int BS, int G>
__global__ __launch_bounds__(BS) void k_hv_csr(CsrArgs A, const double *x, double *y, double ax, double beta,
                                               double *part) {
  __shared__ double red[BS / kWave];
  const uint64_t rows_per_block = BS / G;
  double dot = 0.0;
  const int g = threadIdx.x % G;
  for (uint64_t row = (uint64_t)blockIdx.x * rows_per_block + threadIdx.x / G; row < A.n;
       row += (uint64_t)gridDim.x * rows_per_block) {
    double acc = 0.0;
    const int64_t e1 = A.ptr[row + 1];
    for (int64_t e = A.ptr[row] + g; e < e1; e += G) acc += A.val[e] * x[A.col[e]];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    if (g == 0) {
      acc += A.diag[row] * x[row];
      const double out = (beta != 0.0) ? ax * acc + beta * y[row] : ax * acc;
      y[row] = out;
      dot += out * x[row];
    }
  }
  if (part) {
    const double s = block_sum<BS>(dot, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// On-the-fly H.x (pull form): matrix elements regenerated from the bit
// patterns; target index by combinatorial ranking (binomials in LDS).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t rank_half(uint32_t s, const uint32_t *binom, int L) {
  uint32_t r = 0;
  int k = 1;
  while (s) {
    const int p = __ffs(s) - 1;
    r += binom[p * (L + 1) + k];
    ++k;
    s &= s - 1u;
  }
  return r;
}

template <int BS>
__global__ __launch_bounds__(BS) void k_hv_otf(OtfArgs A, const double *x, double *y, double ax, double beta,
                                               double *part) {
  __shared__ uint32_t bn[33 * 33];
  __shared__ double red[BS / kWave];
  const int L = A.L;
  for (int i = threadIdx.x; i < (L + 1) * (L + 1); i += BS) bn[i] = A.binom[i];
  __syncthreads();
  const size_t n = (size_t)A.du * A.dd;
  double dot = 0.0;
  for (size_t I = (size_t)blockIdx.x * BS + threadIdx.x; I < n; I += (size_t)gridDim.x * BS) {
    const uint32_t id = (uint32_t)(I / A.du), iu = (uint32_t)(I - (size_t)id * A.du);
    const uint32_t su = A.up_states[iu], sd = A.dn_states[id];
    double acc = A.dg[I] * x[I];
    for (int h = 0; h < A.nhop; ++h) {
      const OtfHop hp = A.hops[h];
      const bool down = hp.r >= L;
      const uint32_t s = down ? sd : su;
      const int r = down ? hp.r - L : hp.r, c = down ? hp.c - L : hp.c;
      if (!((s >> r) & 1u) || ((s >> c) & 1u)) continue;
      const uint32_t s2 = s ^ (1u << r) ^ (1u << c);
      const int lo = min(r, c), hi = max(r, c);
      const uint32_t inside = (hi - lo > 1) ? (((1u << hi) - 1u) & ~((2u << lo) - 1u)) : 0u;
      const double sg = (__popc(s & inside) & 1) ? -1.0 : 1.0;
      const uint32_t rk = rank_half(s2, bn, L);
      const size_t J = down ? (size_t)rk * A.du + iu : (size_t)id * A.du + rk;
      acc += hp.v * sg * x[J];
    }
    const double out = (beta != 0.0) ? ax * acc + beta * y[I] : ax * acc;
    y[I] = out;
    dot += out * x[I];
  }
  if (part) {
    const double s = block_sum<BS>(dot, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
constexpr int kMaxVec = 16;
struct PtrPack { const double *p[kMaxVec]; };
struct MPtrPack { double *p[kMaxVec]; };
struct CoefPack { double c[kMaxVec]; };

// The vector count M is a template parameter: with a run-time count the
// guarded loads of the unrolled loop were each followed by a wait, so a
// thread's loads ran one after another at full memory latency.
template <int BS, int M>
__global__ __launch_bounds__(BS) void k_multi_dot(size_t n, const double *a, PtrPack b, double *part) {
  __shared__ double red[M > 0 ? M : 1][BS / kWave];
  double acc[M > 0 ? M : 1];
#pragma unroll
  for (int k = 0; k < M; ++k) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
    const double av = a[i];
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] += av * b.p[k][i];
  }
  // one shuffle reduction per vector, one barrier per block
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const double s = wave_sum(acc[k]);
    if (lane == 0) red[k][w] = s;
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < BS / kWave; ++q) s += red[threadIdx.x][q];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

template <int BS, int M>
__global__ __launch_bounds__(BS) void k_multi_axpy_src(size_t n, double *x, MPtrPack y, CoefPack c, double *part,
                                                       double xs) {
  __shared__ double red[BS / kWave];
  double nn = 0.0;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
    double xv = x[i];
    if (xs != 1.0) {
      xv *= xs;
      x[i] = xv;
    }
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double v = y.p[k][i] + c.c[k] * xv;
      y.p[k][i] = v;
      if (k == 0) nn += v * v;
    }
  }
  if (part) {
    const double s = block_sum<BS>(nn, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
  }
}

template <int BS, int M>
__global__ __launch_bounds__(BS) void k_combine(size_t n, double *t, double st, PtrPack b, CoefPack c,
                                                double *npart, const double *w, double *dpart) {
  __shared__ double red[BS / kWave];
  double nn = 0.0, dd = 0.0;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
    double v = (st == 0.0) ? 0.0 : st * t[i];
#pragma unroll
    for (int k = 0; k < M; ++k) v += c.c[k] * b.p[k][i];
    t[i] = v;
    nn += v * v;
    if (w) dd += v * w[i];
  }
  if (npart) {
    const double s = block_sum<BS>(nn, red);
    if (threadIdx.x == 0) npart[blockIdx.x] = s;
  }
  if (dpart) {
    const double s = block_sum<BS>(dd, red);
    if (threadIdx.x == 0) dpart[blockIdx.x] = s;
  }
}

// One band-Lanczos step's vector work in one pass: x scaled, the NA vectors
// ahead updated from it, t combined from the NB vectors behind and x (the same
// operations, in the same order, as k_multi_axpy_src followed by k_combine),
// and the dots of the next step's vector c (a_0, or t when NA == 0) with the
// vectors of its window, so that step needs no separate multi-dot pass.
template <int BS, int NA, int NB>
__global__ __launch_bounds__(BS) void k_band_step(size_t n, double *x, double xs, MPtrPack a, CoefPack ca, double *t,
                                                  double st, PtrPack w, CoefPack cb, double cx, double *part,
                                                  BandFold F) {
  constexpr int ND = (NA > 0 ? 1 + NA : 1) + NB + 1;
  __shared__ double red[ND][BS / kWave];
  __shared__ double dv[NA + NB + 2];
  __shared__ double sc[3 + NA + NB + 1];
  if (F.stop && *F.stop) return;  // a deflation found by an earlier step
  if (F.dpart) {
    // device-driven step: every workgroup folds the dots of v_j (rows of
    // F.dnb partials: <v_j, v_j>, the NA ahead, the NB behind from row F.rb)
    // and <H v_j, v_j> (F.hnb partials), one wave per row in a fixed order,
    // so all of them derive the same scalars; workgroup 0 stores the record
    const int lane = threadIdx.x & (kWave - 1), wq = threadIdx.x / kWave;
    if (wq < NA + NB + 2) {
      const double *row;
      int nb = F.dnb;
      if (wq == 0) row = F.dpart;
      else if (wq <= NA) row = F.dpart + (size_t)wq * F.dnb;
      else if (wq <= NA + NB) row = F.dpart + (size_t)(F.rb + wq - 1 - NA) * F.dnb;
      else {
        row = F.hpart;
        nb = F.hnb;
      }
      double sm = 0.0;
      for (int i = lane; i < nb; i += kWave) sm += row[i];
      sm = wave_sum(sm);
      if (lane == 0) dv[wq] = sm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const double z = sqrt(fmax(dv[0], 0.0));
      if (z < F.accur) {
        sc[3 + NA + NB] = 1.0;
        if (blockIdx.x == 0) {
          *F.stop = 1;
          F.rec[6] = 1.0;
        }
      } else {
        const double iz = 1.0 / z;
        double zz[NA > 0 ? NA : 1];
#pragma unroll
        for (int k = 0; k < NA; ++k) zz[k] = dv[1 + k] * iz;
        double wdot = dv[1 + NA + NB] * iz * iz;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const int k = F.k0 + b;
          const double tb = (k == F.j - F.pc) ? z : zz[k + F.pc - F.j - 1];
          wdot -= tb * dv[1 + NA + b] * iz;
          sc[3 + NA + b] = -tb;
        }
#pragma unroll
        for (int k = 0; k < NA; ++k) sc[3 + k] = -zz[k];
        sc[0] = iz;
        sc[1] = iz;
        sc[2] = -wdot;
        sc[3 + NA + NB] = 0.0;
        if (blockIdx.x == 0) {
          F.rec[0] = z;
          for (int k = 0; k < kBandStepMax; ++k) F.rec[1 + k] = k < NA ? zz[k] : 0.0;
          F.rec[5] = wdot;
          F.rec[6] = 0.0;
        }
      }
    }
    __syncthreads();
    if (sc[3 + NA + NB] != 0.0) return;
    xs = sc[0];
    st = sc[1];
    cx = sc[2];
#pragma unroll
    for (int k = 0; k < NA; ++k) ca.c[k] = sc[3 + k];
#pragma unroll
    for (int b = 0; b < NB; ++b) cb.c[b] = sc[3 + NA + b];
  }
  double acc[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) acc[k] = 0.0;
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
    const double xv = x[i] * xs;
    x[i] = xv;
    double av[NA > 0 ? NA : 1];
#pragma unroll
    for (int k = 0; k < NA; ++k) {
      av[k] = a.p[k][i] + ca.c[k] * xv;
      a.p[k][i] = av[k];
    }
    double wv[NB > 0 ? NB : 1];
    double tv = st * t[i];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      wv[b] = w.p[b][i];
      tv += cb.c[b] * wv[b];
    }
    tv += cx * xv;
    t[i] = tv;
    const double c = NA > 0 ? av[0] : tv;
    acc[0] += c * c;
    int r = 1;
#pragma unroll
    for (int k = 1; k < NA; ++k) acc[r++] += c * av[k];
    if (NA > 0) acc[r++] += c * tv;
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[r++] += c * wv[b];
    acc[r] += c * xv;
  }
  const int lane = threadIdx.x & (kWave - 1), wv_ = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    const double sm = wave_sum(acc[k]);
    if (lane == 0) red[k][wv_] = sm;
  }
  __syncthreads();
  if (threadIdx.x < ND) {
    double sm = 0.0;
#pragma unroll
    for (int q = 0; q < BS / kWave; ++q) sm += red[threadIdx.x][q];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = sm;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_scale(size_t n, double *x, double a) {
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) x[i] *= a;
}
template <int BS>
__global__ __launch_bounds__(BS) void k_fill(size_t n, double *x, double v) {
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) x[i] = v;
}
__global__ void k_set1(double *x, size_t i) { x[i] = 1.0; }

// N(0,1) start vectors (normal_draw, ed_rng.hpp)
template <int BS>
__global__ __launch_bounds__(BS) void k_random_normal(size_t n, uint64_t seed, double *x) {
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) x[i] = normal_draw(seed, i);
}

// Davidson correction (src_ed/Hamiltonian/Davidson.hpp:181-188): r_i /= (mu - d_i + tiny)
template <int BS>
__global__ __launch_bounds__(BS) void k_precond(size_t n, double *r, const double *d, double mu, double tiny) {
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS)
    r[i] /= (-d[i] + mu + tiny);
}

// Lanczos eigenvector replay step: p = p + c1 r ; g = g + c2 p (new p)
template <int BS>
__global__ __launch_bounds__(BS) void k_replay(size_t n, double *p, const double *r, double *g, double c1, double c2) {
  for (size_t i = (size_t)blockIdx.x * BS + threadIdx.x; i < n; i += (size_t)gridDim.x * BS) {
    const double v = p[i] + c1 * r[i];
    p[i] = v;
    g[i] += c2 * v;
  }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_cdag(CdagArgs A, const double *x, double *y) {
  const size_t n = (size_t)A.du_t * A.dd_t;
  for (size_t t = (size_t)blockIdx.x * BS + threadIdx.x; t < n; t += (size_t)gridDim.x * BS) {
    const uint32_t id = (uint32_t)(t / A.du_t), iu = (uint32_t)(t - (size_t)id * A.du_t);
    if (A.spin == 0) {
      const int32_t s = A.map[iu];
      if (s < 0) continue;
      y[t] += A.z * (double)A.sgn[iu] * x[(size_t)s + (size_t)A.du_s * id];
    } else {
      const int32_t s = A.map[id];
      if (s < 0) continue;
      y[t] += A.z * (double)A.sgn[id] * (double)A.upar[iu] * x[(size_t)iu + (size_t)A.du_s * s];
    }
  }
}

// ---------------------------------------------------------------------------
// Green-function kernels (complex FP64, interleaved)
// ---------------------------------------------------------------------------
struct cplx { double re, im; };
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm for a / b
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return {(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  const double r = b.re / b.im, den = b.re * r + b.im;
  return {(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}

// one thread per (fraction, frequency), frequency fastest: the lanes of a
// wave share the fraction, hence the trip count and the (broadcast)
// coefficient loads.  G <- b_i / (z - a_i - G), innermost level first.
__global__ void k_cf_eval(int ncf, const int *off, const double *a, const double *b, int nz, const double *z,
                          double *out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)ncf * nz) return;
  const int iz = (int)(t % nz), f = (int)(t / nz);
  const double zr = z[2 * iz], zi = z[2 * iz + 1];
  double gr = 0.0, gi = 0.0;
  for (int i = off[f + 1] - 1; i >= off[f]; --i) {
    const double dr = zr - a[i] - gr, di = zi - gi, den = dr * dr + di * di;
    // 1 / den: hardware estimate, one Newton step
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    const double s = b[i] * r;
    gr = dr * s;
    gi = -di * s;
  }
  const cplx G{gr, gi};
  out[2 * ((size_t)iz * ncf + f)] = G.re;
  out[2 * ((size_t)iz * ncf + f) + 1] = G.im;
}

// G[b + a L] = sum_i V(a, i) conj(V(b, i)) / (z - e_i); Vi (imaginary parts
// of V) may be null
__global__ void k_vdvh(int L, int M, const double *e, const double *V, const double *Vi, int nz, const double *z,
                       double *G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int LL = L * L;
  if (t >= nz * LL) return;
  const int iz = t / LL, ab = t % LL;
  const int bb = ab % L, aa = ab / L;  // G[b + a L]
  const double zr = z[2 * iz], zi = z[2 * iz + 1];
  double sr = 0.0, si = 0.0;
  if (!Vi) {
    for (int i = 0; i < M; ++i) {
      const double tr = zr - e[i], ti = zi;
      const double inv = 1.0 / (tr * tr + ti * ti);
      const double w = V[aa + (size_t)i * L] * V[bb + (size_t)i * L];
      sr += w * tr * inv;
      si -= w * ti * inv;
    }
  } else {
    for (int i = 0; i < M; ++i) {
      const double tr = zr - e[i], ti = zi;
      const double inv = 1.0 / (tr * tr + ti * ti);
      const double ar = V[aa + (size_t)i * L], ai = Vi[aa + (size_t)i * L];
      const double br = V[bb + (size_t)i * L], bi = Vi[bb + (size_t)i * L];
      const double wr = ar * br + ai * bi, wi = ai * br - ar * bi;  // V(a) conj(V(b))
      sr += (wr * tr + wi * ti) * inv;
      si += (wi * tr - wr * ti) * inv;
    }
  }
  G[2 * ((size_t)iz * LL + ab)] = sr;
  G[2 * ((size_t)iz * LL + ab) + 1] = si;
}

constexpr int kMcfMax = 8;
// Matrix continued fraction on a frequency grid.  One 64-lane workgroup holds
// 64 / p^2 frequencies; each frequency's p x p complex matrices (F, D, T and
// the inverse) live in LDS, one lane per entry (r, c).  Level j (from the
// last): D = z - A_j - B_j^T F B_j, F = D^-1 by Gauss-Jordan with partial
// pivoting (first largest modulus, as a serial sweep picks it).
__global__ __launch_bounds__(64) void k_mcf_eval(int p, int nfl, const double *A, const double *B, const double *W,
                                                 int q, int nz, const double *z, double *G) {
  constexpr int M = kMcfMax * kMcfMax;
  __shared__ cplx sF[M], sD[M], sT[M], sI[M];
  __shared__ double sA[M], sB[M];
  __shared__ int sfail[M];
  const int pp = p * p;
  const int per = 64 / pp;                  // frequencies per workgroup
  const int g = threadIdx.x / pp;           // this lane's frequency slot
  const int e = threadIdx.x - g * pp;       // entry r + c p
  const int r = e % p, c = e / p;
  const bool lane_on = g < per;
  const int iz = blockIdx.x * per + g;
  const bool live = lane_on && iz < nz;
  const int o = g * pp;                     // slot's offset in the LDS matrices
  double zr = 0.0, zi = 0.0;
  if (live) {
    zr = z[2 * iz];
    zi = z[2 * iz + 1];
  }
  if (lane_on) sF[o + e] = {0.0, 0.0};
  if (threadIdx.x < per) sfail[threadIdx.x] = 0;
  for (int j = nfl - 1; j >= 0; --j) {
    __syncthreads();
    for (int k = threadIdx.x; k < pp; k += 64) {
      sA[k] = A[(size_t)j * pp + k];
      sB[k] = B[(size_t)j * pp + k];
    }
    __syncthreads();
    cplx d{0.0, 0.0};
    if (lane_on) {
      d = {-sA[e], 0.0};
      if (r == c) { d.re += zr; d.im += zi; }
    }
    if (j < nfl - 1) {
      if (lane_on) {
        cplx t{0.0, 0.0};
        for (int k = 0; k < p; ++k) {
          const cplx f = sF[o + r + k * p];
          const double bb = sB[k + c * p];
          t.re += f.re * bb;
          t.im += f.im * bb;
        }
        sT[o + e] = t;
      }
      __syncthreads();
      if (lane_on) {
        cplx t{0.0, 0.0};
        for (int k = 0; k < p; ++k) {
          const double bb = sB[k + r * p];
          const cplx u = sT[o + k + c * p];
          t.re += bb * u.re;
          t.im += bb * u.im;
        }
        d.re -= t.re;
        d.im -= t.im;
      }
    }
    // Gauss-Jordan: sD -> identity, sI -> D^-1
    if (lane_on) {
      sD[o + e] = d;
      sI[o + e] = {r == c ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    for (int k = 0; k < p; ++k) {
      int pr = k;
      double best = 0.0;
      if (lane_on) {
        best = hypot(sD[o + k + k * p].re, sD[o + k + k * p].im);
        for (int rr = k + 1; rr < p; ++rr) {
          const double m = hypot(sD[o + rr + k * p].re, sD[o + rr + k * p].im);
          if (m > best) { best = m; pr = rr; }
        }
        if (best == 0.0 && e == 0) sfail[g] = 1;
      }
      // row exchange k <-> pr, then row k scaled by 1 / D[k][k]
      cplx va{0.0, 0.0}, vi{0.0, 0.0};
      if (lane_on && best != 0.0) {
        const int src = (r == k) ? pr : (r == pr ? k : r);
        va = sD[o + src + c * p];
        vi = sI[o + src + c * p];
      }
      __syncthreads();
      if (lane_on && best != 0.0) {
        if (r == k) {
          const cplx piv = sD[o + pr + k * p];  // not yet overwritten: row pr is written below only by lanes r == pr
          const cplx dinv = cdiv(cplx{1.0, 0.0}, piv);
          va = {va.re * dinv.re - va.im * dinv.im, va.re * dinv.im + va.im * dinv.re};
          vi = {vi.re * dinv.re - vi.im * dinv.im, vi.re * dinv.im + vi.im * dinv.re};
        }
      }
      __syncthreads();
      if (lane_on && best != 0.0) {
        sD[o + e] = va;
        sI[o + e] = vi;
      }
      __syncthreads();
      // elimination of column k from the other rows
      cplx f{0.0, 0.0}, u{0.0, 0.0}, w{0.0, 0.0};
      if (lane_on && best != 0.0 && r != k) {
        f = sD[o + r + k * p];
        u = sD[o + k + c * p];
        w = sI[o + k + c * p];
      }
      __syncthreads();
      if (lane_on && best != 0.0 && r != k && (f.re != 0.0 || f.im != 0.0)) {
        sD[o + e].re -= f.re * u.re - f.im * u.im;
        sD[o + e].im -= f.re * u.im + f.im * u.re;
        sI[o + e].re -= f.re * w.re - f.im * w.im;
        sI[o + e].im -= f.re * w.im + f.im * w.re;
      }
      __syncthreads();
    }
    if (lane_on) sF[o + e] = sI[o + e];
  }
  __syncthreads();
  if (!live || r >= q || c >= q) return;
  double *out = G + 2 * ((size_t)iz * q * q + r + c * q);
  if (sfail[g]) {
    out[0] = NAN;
    out[1] = NAN;
    return;
  }
  double sr = 0.0, si = 0.0;
  for (int k = 0; k < p; ++k)
    for (int l = 0; l < p; ++l) {
      const double ww = W[k + r * p] * W[l + c * p];
      sr += ww * sF[o + k + l * p].re;
      si += ww * sF[o + k + l * p].im;
    }
  out[0] = sr;
  out[1] = si;
}

int g_ncu = 0;
int num_cu() {
  if (!g_ncu) {
    int dev = 0;
    ED_HIP(hipGetDevice(&dev));
    ED_HIP(hipDeviceGetAttribute(&g_ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (g_ncu <= 0) g_ncu = 256;
  }
  return g_ncu;
}

// blocks per CU of the Krylov vector kernels (multi-dot, axpy, combine)
int reduction_blocks_per_cu() {
  const long v = params().geti("reduction_blocks_per_cu");
  return (int)std::max(1L, std::min(v, 64L));
}

int grid_for(size_t n, int bs, int per_cu = 8) {
  size_t g = (n + bs - 1) / bs;
  size_t cap = (size_t)num_cu() * per_cu;
  if (g > cap) g = cap;
  if (g > (size_t)kPartialsCap) g = kPartialsCap;
  return g < 1 ? 1 : (int)g;
}

size_t g_lds_max = 0;
// largest dynamic LDS a block may request (160 KB per CU on gfx950, minus the
// static tables of the row kernel)
size_t lds_row_budget() {
  if (!g_lds_max) {
    int dev = 0, v = 0;
    ED_HIP(hipGetDevice(&dev));
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
      (void)hipGetLastError();
      v = 64 * 1024;
    }
    g_lds_max = (size_t)v;
  }
  const size_t reserve = 8 * 1024;
  return g_lds_max > reserve ? g_lds_max - reserve : 0;
}

uint32_t rows_per_block(uint32_t du) {
  if ((size_t)(du + 1) * 8 > lds_row_budget()) return 0;  // row does not fit: no LDS staging
  uint32_t R = 1;
  while (R * du < 512 && (size_t)(2 * R) * du * 8 <= 32 * 1024) R *= 2;
  return R;
}

template <typename K>
void allow_lds(K kern, size_t bytes) {
  if (bytes > 64 * 1024)
    ED_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)bytes));
}

}  // namespace

int partials_cap() { return kPartialsCap; }

// ---------------------------------------------------------------------------
// Fixed scratch of the allocation-free entry points (ed_*_noalloc): the
// per-block partials of one product (kPartialsCap) followed by the two-pass
// schedule's per-slab-block partials (kNoAllocAux; at most 8 XCDs x 202 slabs
// x 202 row chunks = 326 432 on the 12 870 x 12 870 sector, with 8-column
// slabs).  3.2 MB per device, held until the process ends.  Every launch overwrites
// the entries it later reads, so nothing is cleared between calls; calls that
// overlap on one device race on it.  In host code the array's name is only a
// stand-in: the device address comes from hipGetSymbolAddress, once per device.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocAux = 393216;
constexpr int kMaxNoAllocDevices = 64;
static __device__ __align__(256) double scratch[kPartialsCap + kNoAllocAux];

NoAllocScratch noalloc_scratch() {
  static double *base[kMaxNoAllocDevices] = {};
  static std::mutex mu;
  int dev = 0;
  ED_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxNoAllocDevices) fail(ED_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!base[dev]) {
    void *p = nullptr;
    ED_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(ED_ERR_HIP, "no device address for the fixed scratch");
    base[dev] = static_cast<double *>(p);
  }
  NoAllocScratch S;
  S.part = base[dev];
  S.aux = base[dev] + kPartialsCap;
  S.aux_cap = kNoAllocAux;
  return S;
}

namespace {
// one block: out[0] = sum of in[0 .. n), lanes strided then a fixed-order tree
template <int BS>
__global__ __launch_bounds__(BS) void k_fold_one(const double *in, int n, double *out) {
  __shared__ double red[BS / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += BS) s += in[i];
  s = block_sum<BS>(s, red);
  if (threadIdx.x == 0) out[0] = s;
}
}  // namespace

void fold_partials_device(const double *part, int nb, double *out, hipStream_t s) {
  hipLaunchKernelGGL((k_fold_one<256>), dim3(1), dim3(256), 0, s, part, nb, out);
  ED_HIP(hipGetLastError());
}

static FusedParams make_fused(const FactArgs &a, const double *x, double *y, double ax, double beta, double *part,
                              uint32_t R) {
  FusedParams P;
  P.a = a;
  P.x = x;
  P.y = y;
  P.ax = ax;
  P.beta = beta;
  P.part = part;
  P.R = R;
  // value tables up to kVtabMax entries are staged in LDS, larger ones are read from global
  P.ntab_up = (a.up.packed && a.up.nval <= (uint32_t)kVtabMax) ? (int)a.up.nval : 0;
  P.ntab_dn = (a.dn.packed && a.dn.nval <= (uint32_t)kVtabMax) ? (int)a.dn.nval : 0;
  return P;
}

static int fused_grid(const FactArgs &a, uint32_t R) {
  const uint32_t nb = (a.dn.dim + R - 1) / R;
  return (int)std::min<uint32_t>(nb, (uint32_t)kPartialsCap);
}

int hv_factorized_max_partials(const FactArgs &a, int variant) { return kPartialsCap; }

int hv_factorized_cfg(const FactArgs &a, const double *x, double *y, double ax, double beta, double *part, int variant,
                      const SlabConfig &sc_in, hipStream_t s, double *aux, size_t aux_cap) {
  const uint32_t du = a.up.dim, dd = a.dn.dim;
  if ((size_t)du * dd == 0) return 0;
  SlabConfig sc = sc_in;
  if (variant == HV_TWO_PASS_LDS) {
    sc.lds = 1;
    sc.sw = 0;
    for (int w : {16, 8, 4})
      if (!sc.sw && (size_t)dd * w * sizeof(double) <= kDownLdsMax) sc.sw = w;
    if (!sc.sw) fail(ED_ERR_INVALID, "hv_variant 3: a slab of this sector's rows does not fit in LDS");
    variant = HV_TWO_PASS;
  }
  uint32_t R = rows_per_block(du);
  const bool stage = R > 0;
  if (!stage) R = 1;
  if (variant == HV_AUTO) variant = ((size_t)du * dd >= ((size_t)1 << 26)) ? HV_TWO_PASS : HV_FUSED;
  const size_t lds = stage ? ((size_t)R * du + 1) * sizeof(double) : 0;
  constexpr int BS = 1024;
  const bool down_here = variant == HV_FUSED;
  {
    FusedParams P = make_fused(a, x, y, ax, beta, down_here ? part : nullptr, R);
    const int g = fused_grid(a, R);
#define ED_LAUNCH_ROWS(BSZ, ONE, DOWN, STG)                                   \
  do {                                                                          \
    auto kern = k_hv_rows<BSZ, ONE, DOWN, STG>;                                 \
    if (STG) allow_lds(kern, lds);                                              \
    hipLaunchKernelGGL(kern, dim3(g), dim3(BSZ), STG ? lds : 0, s, P);          \
  } while (0)
#define ED_LAUNCH_ROWK(KM, DOWN)                                               \
  do {                                                                          \
    auto kern = k_hv_rowk<BS, KM, DOWN>;                                        \
    allow_lds(kern, lds);                                                       \
    hipLaunchKernelGGL(kern, dim3(g), dim3(BS), lds, s, P);                     \
  } while (0)
    const uint32_t kneed = (du + BS - 1) / BS;
    if (stage && R == 1 && a.up.sell[0].nslice > 0 && a.up.sell[1].nslice > 0) {
      const int tm = a.up.mode16 ? 2 : (a.up.packed ? 0 : 1);
      // short rows: 4-wave blocks, so that all rows of a small sector are resident at once
      const bool small = a.block == 256 || (a.block == 0 && du < 2048);
#define ED_SELL(BSZ, DOWN, TMV)                                               \
  do {                                                                        \
    auto kern = k_hv_rows_sell<BSZ, DOWN, TMV>;                               \
    allow_lds(kern, lds);                                                     \
    hipLaunchKernelGGL(kern, dim3(g), dim3(BSZ), lds, s, P);                  \
  } while (0)
#define ED_SELL_TM(BSZ, DOWN)                                                 \
  do {                                                                        \
    if (tm == 2) ED_SELL(BSZ, DOWN, 2);                                       \
    else if (tm == 1) ED_SELL(BSZ, DOWN, 1);                                  \
    else ED_SELL(BSZ, DOWN, 0);                                               \
  } while (0)
      if (small) {
        if (down_here) ED_SELL_TM(256, true); else ED_SELL_TM(256, false);
      } else {
        if (down_here) ED_SELL_TM(BS, true); else ED_SELL_TM(BS, false);
      }
#undef ED_SELL_TM
#undef ED_SELL
    } else if (stage && R == 1 && kneed <= 16) {
      if (down_here) {
        if (kneed <= 1) ED_LAUNCH_ROWK(1, true);
        else if (kneed <= 2) ED_LAUNCH_ROWK(2, true);
        else if (kneed <= 4) ED_LAUNCH_ROWK(4, true);
        else if (kneed <= 8) ED_LAUNCH_ROWK(8, true);
        else ED_LAUNCH_ROWK(16, true);
      } else {
        if (kneed <= 1) ED_LAUNCH_ROWK(1, false);
        else if (kneed <= 2) ED_LAUNCH_ROWK(2, false);
        else if (kneed <= 4) ED_LAUNCH_ROWK(4, false);
        else if (kneed <= 8) ED_LAUNCH_ROWK(8, false);
        else ED_LAUNCH_ROWK(16, false);
      }
    } else if (down_here) {
      if (!stage) ED_LAUNCH_ROWS(BS, true, true, false);
      else if (R == 1) ED_LAUNCH_ROWS(BS, true, true, true);
      else ED_LAUNCH_ROWS(256, false, true, true);
    } else {
      if (!stage) ED_LAUNCH_ROWS(BS, true, false, false);
      else if (R == 1) ED_LAUNCH_ROWS(BS, true, false, true);
      else ED_LAUNCH_ROWS(256, false, false, true);
    }
#undef ED_LAUNCH_ROWS
#undef ED_LAUNCH_ROWK
    ED_HIP(hipGetLastError());
    if (down_here) return g;
  }
  // second pass of the two-pass variant: column slabs (down part)
  SlabParams Q;
  Q.a = a;
  Q.x = x;
  Q.y = y;
  Q.ax = ax;
  Q.ntab_dn = (a.dn.packed && a.dn.nval <= (uint32_t)kVtabMax) ? (int)a.dn.nval : 0;
  if (sc.lds) {
    // slabs resident in LDS (small sectors)
    const int W = (sc.sw == 4 || sc.sw == 8 || sc.sw == 16) ? sc.sw : 8;
    const size_t bytes = (size_t)dd * W * sizeof(double);
    const size_t nb = (du + W - 1) / W;
    if (bytes > kDownLdsMax || nb > (size_t)kPartialsCap)
      fail(ED_ERR_INTERNAL, "LDS slab schedule chosen for a sector it does not fit");
    Q.part = part;
#define ED_DLDS(WV)                                                              \
  do {                                                                           \
    auto kern = k_hv_down_lds<WV>;                                               \
    allow_lds(kern, bytes);                                                      \
    hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(256), bytes, s, Q);         \
  } while (0)
    if (W == 4) ED_DLDS(4);
    else if (W == 16) ED_DLDS(16);
    else ED_DLDS(8);
#undef ED_DLDS
    ED_HIP(hipGetLastError());
    return (int)nb;
  }
  Q.nxcd = (sc.nxcd == 1) ? 1 : 8;
  const int SW = (sc.sw == 8 || sc.sw == 16 || sc.sw == 64) ? sc.sw : 32;
  Q.nslab = (du + SW - 1) / SW;
  Q.chunk_rows = kSlabRows;
  Q.nchunk = (dd + kSlabRows - 1) / kSlabRows;
  Q.ntab_dn = (a.dn.packed && a.dn.nval <= (uint32_t)kVtabMax) ? (int)a.dn.nval : 0;
  const size_t slabs_per_xcd = (Q.nslab + Q.nxcd - 1) / Q.nxcd;
  const size_t nb = (size_t)Q.nxcd * slabs_per_xcd * Q.nchunk;
  // per-block partials go to a scratch array folded back to <= kPartialsCap
  // (one array per host thread, queued on that thread's stream), or to the
  // caller's aux array
  thread_local DevBuf<double> big;
  double *wide = nullptr;
  if (part && nb > (size_t)kPartialsCap) {
    if (aux) {
      if (nb > aux_cap)
        fail(ED_ERR_INVALID, "two-pass product: " + std::to_string(nb) + " block partials exceed the fixed scratch of " +
                                 std::to_string(aux_cap));
      wide = aux;
    } else {
      if (big.n < nb) big.alloc(nb);
      wide = big.p;
    }
  }
  Q.part = part ? (wide ? wide : part) : nullptr;
#define ED_SLAB(SWV, NTV) hipLaunchKernelGGL((k_hv_down_slab<SWV, 256 / SWV, NTV>), dim3((unsigned)nb), dim3(256), 0, s, Q)
  if (SW == 8) { if (sc.nt) ED_SLAB(8, true); else ED_SLAB(8, false); }
  else if (SW == 16) { if (sc.nt) ED_SLAB(16, true); else ED_SLAB(16, false); }
  else if (SW == 64) { if (sc.nt) ED_SLAB(64, true); else ED_SLAB(64, false); }
  else { if (sc.nt) ED_SLAB(32, true); else ED_SLAB(32, false); }
#undef ED_SLAB
  ED_HIP(hipGetLastError());
  if (part && nb > (size_t)kPartialsCap) {
    hipLaunchKernelGGL((k_fold_partials<256>), dim3(kPartialsCap), dim3(256), 0, s, (const double *)wide, nb, part);
    ED_HIP(hipGetLastError());
    return kPartialsCap;
  }
  return (int)nb;
}

int hv_factorized(const FactArgs &a, const double *x, double *y, double ax, double beta, double *part, int variant,
                  hipStream_t s) {
  return hv_factorized_cfg(a, x, y, ax, beta, part, variant, SlabConfig(), s);
}

void fact_full_diag(const FactArgs &a, double *d, hipStream_t s) {
  const size_t n = (size_t)a.up.dim * a.dn.dim;
  if (!n) return;
  hipLaunchKernelGGL((k_fact_diag<256>), dim3(grid_for(n, 256, 16)), dim3(256), 0, s, a, d);
  ED_HIP(hipGetLastError());
}

int hv_csr_max_partials(const CsrArgs &) { return kPartialsCap; }
int hv_csr(const CsrArgs &A, const double *x, double *y, double ax, double beta, double *part, hipStream_t s) {
  if (!A.n) return 0;
  constexpr int BS = 256, G = 8;
  const int g = grid_for(A.n * G, BS, 16);
  hipLaunchKernelGGL((k_hv_csr<BS, G>), dim3(g), dim3(BS), 0, s, A, x, y, ax, beta, part);
  ED_HIP(hipGetLastError());
  return g;
}

int hv_otf_max_partials(const OtfArgs &) { return kPartialsCap; }
int hv_otf(const OtfArgs &A, const double *x, double *y, double ax, double beta, double *part, hipStream_t s) {
  const size_t n = (size_t)A.du * A.dd;
  if (!n) return 0;
  constexpr int BS = 256;
  const int g = grid_for(n, BS, 16);
  hipLaunchKernelGGL((k_hv_otf<BS>), dim3(g), dim3(BS), 0, s, A, x, y, ax, beta, part);
  ED_HIP(hipGetLastError());
  return g;
}

// case list 0 .. kMaxVec of the vector-count templates
#define ED_VEC_COUNTS(X) \
  X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
static_assert(kMaxVec == 16, "ED_VEC_COUNTS lists 0 .. kMaxVec");

// threads per workgroup of the Krylov vector kernels (256, 512 or 1024)
static int vector_block() {
  const long v = params().geti("vector_block");
  return v >= 1024 ? 1024 : (v >= 512 ? 512 : 256);
}

template <int BS>
static int multi_dot_bs(size_t n, const double *a, int m, const PtrPack &pk, double *part, hipStream_t s) {
  const int g = grid_for(n, BS, reduction_blocks_per_cu());
  switch (m) {
#define ED_MD(M) \
  case M: hipLaunchKernelGGL((k_multi_dot<BS, M>), dim3(g), dim3(BS), 0, s, n, a, pk, part); break;
    ED_VEC_COUNTS(ED_MD)
#undef ED_MD
    default: fail(ED_ERR_INTERNAL, "multi_dot: bad vector count");
  }
  ED_HIP(hipGetLastError());
  return g;
}

int multi_dot(size_t n, const double *a, int m, const double *const *b, double *part, hipStream_t s) {
  if (m > kMaxVec) fail(ED_ERR_INTERNAL, "multi_dot: too many vectors");
  PtrPack pk{};
  for (int k = 0; k < m; ++k) pk.p[k] = b[k];
  switch (vector_block()) {
    case 1024: return multi_dot_bs<1024>(n, a, m, pk, part, s);
    case 512: return multi_dot_bs<512>(n, a, m, pk, part, s);
    default: return multi_dot_bs<256>(n, a, m, pk, part, s);
  }
}

template <int BS>
static int multi_axpy_bs(size_t n, double *x, int m, const MPtrPack &pk, const CoefPack &cp, double *part,
                         hipStream_t s, double xs) {
  const int g = grid_for(n, BS, reduction_blocks_per_cu());
  switch (m) {
#define ED_MA(M) \
  case M: hipLaunchKernelGGL((k_multi_axpy_src<BS, M>), dim3(g), dim3(BS), 0, s, n, x, pk, cp, part, xs); break;
    ED_VEC_COUNTS(ED_MA)
#undef ED_MA
    default: fail(ED_ERR_INTERNAL, "multi_axpy: bad vector count");
  }
  ED_HIP(hipGetLastError());
  return g;
}

int multi_axpy_src(size_t n, double *x, int m, double *const *y, const double *c, double *part, hipStream_t s,
                   double xs) {
  if (m > kMaxVec) fail(ED_ERR_INTERNAL, "multi_axpy: too many vectors");
  MPtrPack pk{};
  CoefPack cp{};
  for (int k = 0; k < m; ++k) { pk.p[k] = y[k]; cp.c[k] = c[k]; }
  switch (vector_block()) {
    case 1024: return multi_axpy_bs<1024>(n, x, m, pk, cp, part, s, xs);
    case 512: return multi_axpy_bs<512>(n, x, m, pk, cp, part, s, xs);
    default: return multi_axpy_bs<256>(n, x, m, pk, cp, part, s, xs);
  }
}

int band_step(size_t n, double *x, double xs, int na, double *const *a, const double *ca, double *t, double st,
              int nw, const double *const *w, const double *cb, double cx, double *part, hipStream_t s,
              const BandFold *fold) {
  const BandFold F = fold ? *fold : BandFold();
  if (fold && (fold->stop == nullptr || fold->rec == nullptr)) fail(ED_ERR_INTERNAL, "band_step: fold without outputs");
  constexpr int BS = 1024;
  if (na < 0 || na > kBandStepMax || nw < 0 || nw > kBandStepMax || vector_block() != BS || n == 0) return 0;
  MPtrPack ap{};
  CoefPack cap{}, cbp{};
  PtrPack wp{};
  for (int k = 0; k < na; ++k) { ap.p[k] = a[k]; cap.c[k] = ca ? ca[k] : 0.0; }
  for (int k = 0; k < nw; ++k) { wp.p[k] = w[k]; cbp.c[k] = cb ? cb[k] : 0.0; }
  const int g = grid_for(n, BS, reduction_blocks_per_cu());
#define ED_BS(NA, NB)                                                                                         \
  if (na == NA && nw == NB) {                                                                                \
    hipLaunchKernelGGL((k_band_step<BS, NA, NB>), dim3(g), dim3(BS), 0, s, n, x, xs, ap, cap, t, st, wp, cbp, \
                       cx, part, F);                                                                         \
    ED_HIP(hipGetLastError());                                                                               \
    return g;                                                                                                \
  }
#define ED_BS_ROW(NA) ED_BS(NA, 0) ED_BS(NA, 1) ED_BS(NA, 2) ED_BS(NA, 3) ED_BS(NA, 4)
  ED_BS_ROW(0) ED_BS_ROW(1) ED_BS_ROW(2) ED_BS_ROW(3) ED_BS_ROW(4)
#undef ED_BS_ROW
#undef ED_BS
  return 0;
}

template <int BS>
static int combine_bs(size_t n, double *t, double st, int m, const PtrPack &pk, const CoefPack &cp, double *npart,
                      const double *w, double *dpart, hipStream_t s) {
  const int g = grid_for(n, BS, reduction_blocks_per_cu());
  switch (m) {
#define ED_CB(M) \
  case M: hipLaunchKernelGGL((k_combine<BS, M>), dim3(g), dim3(BS), 0, s, n, t, st, pk, cp, npart, w, dpart); break;
    ED_VEC_COUNTS(ED_CB)
#undef ED_CB
    default: fail(ED_ERR_INTERNAL, "combine: bad vector count");
  }
  ED_HIP(hipGetLastError());
  return g;
}

int combine(size_t n, double *t, double st, int m, const double *const *b, const double *c, double *npart,
            const double *w, double *dpart, hipStream_t s) {
  if (m > kMaxVec) fail(ED_ERR_INTERNAL, "combine: too many vectors");
  PtrPack pk{};
  CoefPack cp{};
  for (int k = 0; k < m; ++k) { pk.p[k] = b[k]; cp.c[k] = c[k]; }
  switch (vector_block()) {
    case 1024: return combine_bs<1024>(n, t, st, m, pk, cp, npart, w, dpart, s);
    case 512: return combine_bs<512>(n, t, st, m, pk, cp, npart, w, dpart, s);
    default: return combine_bs<256>(n, t, st, m, pk, cp, npart, w, dpart, s);
  }
}

void scale(size_t n, double *x, double a, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_scale<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, n, x, a);
  ED_HIP(hipGetLastError());
}
void fill(size_t n, double *x, double v, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_fill<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, n, x, v);
  ED_HIP(hipGetLastError());
}
void fold_rows_flag(const double *part, int nb, int m, size_t stride, double *out, uint64_t *flag, uint64_t seq,
                    hipStream_t s) {
  hipLaunchKernelGGL((k_fold_rows_flag<1024>), dim3(1), dim3(1024), 0, s, part, nb, m, stride, out, flag, seq);
  ED_HIP(hipGetLastError());
}
void random_normal(size_t n, uint64_t seed, double *x, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_random_normal<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, n, seed, x);
  ED_HIP(hipGetLastError());
}
void replay_step(size_t n, double *p, const double *r, double *g, double c1, double c2, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_replay<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, n, p, r, g, c1, c2);
  ED_HIP(hipGetLastError());
}

void precond(size_t n, double *r, const double *d, double mu, double tiny, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL((k_precond<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, n, r, d, mu, tiny);
  ED_HIP(hipGetLastError());
}

void set_unit(double *x, size_t i, hipStream_t s) {
  hipLaunchKernelGGL(k_set1, dim3(1), dim3(1), 0, s, x, i);
  ED_HIP(hipGetLastError());
}

namespace {
// device-driven Lanczos step kernels; step k = *base + j.  Each workgroup
// folds the previous kernel's per-block partials itself (the same values in
// the same order in every workgroup, so every workgroup sees the same scalar)
// and workgroup 0 stores the scalar; the grids stay at one workgroup per CU so
// that the partial sets stay short.
constexpr int kLzBlock = 1024;
__device__ __forceinline__ double lz_fold(const double *part, int nb, double *sh) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += kLzBlock) s += part[i];
  s = block_sum<kLzBlock>(s, sh);
  if (threadIdx.x == 0) sh[kLzBlock / kWave] = s;
  __syncthreads();
  return sh[kLzBlock / kWave];
}
// One device-driven step on unnormalized vectors (v_k = U_k / n_k with
// n_k = |U_k| = b[k-1], n_0 = 1): Y = H U_k and its partials of <Y, U_k> come
// from the product; then a[k] = <Y, U_k> / n_k^2, b[k-1] = n_k (folded from the
// previous step's norm partials), and over Y
//   U_{k+1} = (Y - a[k] U_k) / n_k - (n_k / n_{k-1}) U_{k-1},
// with v_k = U_k / n_k copied to slots[k] when that entry exists and the norm
// partials of U_{k+1} written; the two norm-partial arrays alternate with the
// parity of k, so a step never overwrites the partials it folds.  n_k = 0 (an
// exhausted Krylov space) leaves non-finite values, never read: b[k-1] = 0
// ends the recurrence on the host.
__global__ __launch_bounds__(kLzBlock) void k_lz_step(size_t n, double *Y, const double *U, const double *Up,
                                                      double *coef, int cap, const int *base, int j,
                                                      const double *dpart, int dnb, double *npart, int nnb,
                                                      double *const *slots, int nslots) {
  __shared__ double red[kLzBlock / kWave + 1];
  const int k = *base + j;
  const double d = lz_fold(dpart, dnb, red);
  double nk = 1.0, nkm = 1.0;
  if (k > 0) {
    nk = sqrt(lz_fold(npart + (size_t)(k & 1) * kPartialsCap, nnb, red));
    if (k > 1) nkm = coef[cap + k - 2];
  }
  const double inv = 1.0 / nk, a = d * inv * inv, cp = k > 0 ? nk / nkm : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coef[k] = a;
    if (k > 0) coef[cap + k - 1] = nk;
  }
  double *dst = (slots && k > 0 && k < nslots) ? slots[k] : nullptr;
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * kLzBlock;
  // (the first step has no previous vector: its buffer is not read)
  for (size_t i = (size_t)blockIdx.x * kLzBlock + threadIdx.x; i < n; i += stride) {
    const double u = U[i];
    double w = (Y[i] - a * u) * inv;
    if (k > 0) w -= cp * Up[i];
    Y[i] = w;
    s += w * w;
    if (dst) dst[i] = u * inv;
  }
  s = block_sum<kLzBlock>(s, red);
  if (threadIdx.x == 0) npart[(size_t)((k + 1) & 1) * kPartialsCap + blockIdx.x] = s;
}
// end of a block of steps: b[k] = |U_{k+1}| for its last step k (the next
// step would fold the same value), then *base moves on
__global__ __launch_bounds__(kLzBlock) void k_lz_next(int *base, int steps, const double *npart, int nnb,
                                                      double *coef, int cap) {
  __shared__ double red[kLzBlock / kWave + 1];
  const int k = *base + steps - 1;
  const double nk1 = sqrt(lz_fold(npart + (size_t)((k + 1) & 1) * kPartialsCap, nnb, red));
  if (threadIdx.x == 0) {
    coef[cap + k] = nk1;
    *base += steps;
  }
}

__global__ __launch_bounds__(256) void k_group_project(ProjArgs A, const double *x, double *y) {
  const size_t n = (size_t)A.du * A.dd;
  for (size_t J = (size_t)blockIdx.x * 256 + threadIdx.x; J < n; J += (size_t)gridDim.x * 256) {
    const uint32_t ju = (uint32_t)(J % A.du), jd = (uint32_t)(J / A.du);
    double acc = 0.0;
    for (int k = 0; k < A.g; ++k) {
      const size_t ou = (size_t)k * A.du + ju, od = (size_t)k * A.dd + jd;
      const size_t I = (size_t)A.inv_up[ou] + (size_t)A.du * (size_t)A.inv_dn[od];
      acc += A.c[k] * (double)(A.sg_up[ou] * A.sg_dn[od]) * x[I];
    }
    y[J] = acc;
  }
}
}  // namespace

static int lz_grid(size_t n) {
  const size_t g = (n + kLzBlock - 1) / kLzBlock;
  return (int)std::max<size_t>(1, std::min<size_t>(g, (size_t)num_cu()));
}
void lz_step(const LanczosDev &L, double *W, const double *C, const double *P, int j, int dnb, hipStream_t s) {
  const int g = lz_grid(L.n);
  hipLaunchKernelGGL(k_lz_step, dim3(g), dim3(kLzBlock), 0, s, L.n, W, C, P, L.coef, L.cap, (const int *)L.base, j,
                     (const double *)L.part, dnb, L.part + kPartialsCap, g, L.slots, L.nslots);
  ED_HIP(hipGetLastError());
}
void lz_next(const LanczosDev &L, int steps, hipStream_t s) {
  hipLaunchKernelGGL(k_lz_next, dim3(1), dim3(kLzBlock), 0, s, L.base, steps,
                     (const double *)(L.part + kPartialsCap), lz_grid(L.n), L.coef, L.cap);
  ED_HIP(hipGetLastError());
}

void group_project(const ProjArgs &A, const double *x, double *y, hipStream_t s) {
  const size_t n = (size_t)A.du * A.dd;
  if (!n) return;
  hipLaunchKernelGGL(k_group_project, dim3(grid_for(n, 256, 8)), dim3(256), 0, s, A, x, y);
  ED_HIP(hipGetLastError());
}

void cdag_gather(const CdagArgs &A, const double *x, double *y, hipStream_t s) {
  const size_t n = (size_t)A.du_t * A.dd_t;
  if (!n) return;
  hipLaunchKernelGGL((k_cdag<256>), dim3(grid_for(n, 256, 8)), dim3(256), 0, s, A, x, y);
  ED_HIP(hipGetLastError());
}

// one thread per (group, frequency, matrix element)
__global__ void k_cf_assemble(int ngroups, const int *nn, const uint64_t *at, const int *pbeg, const int *frac0,
                              const int *cx, int ncf, int nz, const double *vals, double *G) {
  for (int g = blockIdx.y; g < ngroups; g += gridDim.y) {
    const int n = nn[g];
    const bool cplx = cx && cx[g];
    const int np = n * (n + 1) / 2;
    const size_t nsq = (size_t)n * n, m = nsq * nz;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += (size_t)gridDim.x * blockDim.x) {
      const int iz = (int)(t / nsq), e = (int)(t - (size_t)iz * nsq);
      const int o1 = e % n, o2 = e / n;
      const int a = o1 > o2 ? o1 : o2, b = o1 > o2 ? o2 : o1;
      const int qab = a * (a + 1) / 2 + b, qaa = a * (a + 1) / 2 + a, qbb = b * (b + 1) / 2 + b;
      const double *row = vals + 2 * (size_t)iz * ncf;
      double gr = 0.0, gi = 0.0;
      for (int p = pbeg[g]; p < pbeg[g + 1]; ++p) {
        const double *f = row + 2 * (size_t)frac0[p];
        const double ar = f[4 * qaa] + f[4 * qaa + 2], ai = f[4 * qaa + 1] + f[4 * qaa + 3];
        if (a == b) {
          gr += ar;
          gi += ai;
        } else if (!cplx) {
          const double br = f[4 * qbb] + f[4 * qbb + 2], bi = f[4 * qbb + 1] + f[4 * qbb + 3];
          const double cr = f[4 * qab] + f[4 * qab + 2], ci = f[4 * qab + 1] + f[4 * qab + 3];
          gr += 0.5 * (cr - ar - br);
          gi += 0.5 * (ci - ai - bi);
        } else {
          // G(b, a) = (t1 + t2 - (1 + i) s) / 2, G(a, b) = (t1 - t2 - (1 - i) s) / 2
          // with t1 the (a, b) pair's fraction, t2 = i times its imaginary
          // combination's, s = G_aa + G_bb; here (o1, o2) is (row, column)
          const int qi = np + (a * (a - 1) / 2 + b);
          const double br = f[4 * qbb] + f[4 * qbb + 2], bi = f[4 * qbb + 1] + f[4 * qbb + 3];
          const double t1r = f[4 * qab] + f[4 * qab + 2], t1i = f[4 * qab + 1] + f[4 * qab + 3];
          const double vr = f[4 * qi] + f[4 * qi + 2], vi = f[4 * qi + 1] + f[4 * qi + 3];
          const double t2r = -vi, t2i = vr;
          const double sr = ar + br, si = ai + bi;
          if (o1 < o2) {  // row b, column a
            gr += 0.5 * (t1r + t2r - (sr - si));
            gi += 0.5 * (t1i + t2i - (si + sr));
          } else {        // row a, column b
            gr += 0.5 * (t1r - t2r - (sr + si));
            gi += 0.5 * (t1i - t2i - (si - sr));
          }
        }
      }
      G[at[g] + 2 * t] = gr;
      G[at[g] + 2 * t + 1] = gi;
    }
  }
}

void cf_eval(int ncf, const int *off, const double *a, const double *b, int nz, const double *z, double *out,
             hipStream_t s) {
  const size_t n = (size_t)ncf * nz;
  if (!n) return;
  if ((n + 255) / 256 > 0x7fffffffull) fail(ED_ERR_INVALID, "cf_eval: too many fractions x frequencies");
  hipLaunchKernelGGL(k_cf_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ncf, off, a, b, nz, z, out);
  ED_HIP(hipGetLastError());
}

void cf_assemble(int ngroups, const int *n, const uint64_t *at, const int *pbeg, const int *frac0, const int *cx,
                 int ncf, int nz, const double *vals, double *G, hipStream_t s) {
  if (ngroups <= 0 || nz <= 0) return;
  const dim3 grid(64, (unsigned)(ngroups < 65535 ? ngroups : 65535));
  hipLaunchKernelGGL(k_cf_assemble, grid, dim3(256), 0, s, ngroups, n, at, pbeg, frac0, cx, ncf, nz, vals, G);
  ED_HIP(hipGetLastError());
}

void vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G, hipStream_t s,
          const double *Vi) {
  const int n = nz * L * L;
  if (!n) return;
  hipLaunchKernelGGL(k_vdvh, dim3((n + 255) / 256), dim3(256), 0, s, L, M, e, V, Vi, nz, z, G);
  ED_HIP(hipGetLastError());
}

void mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q, int nz, const double *z,
              double *G, hipStream_t s) {
  if (p > kMcfMax) fail(ED_ERR_UNSUPPORTED, "matrix continued fraction with block size > 8 on device");
  if (!nz) return;
  if (p < 1 || q < 0 || q > p) fail(ED_ERR_INVALID, "matrix continued fraction: bad block sizes");
  const int per = 64 / (p * p);
  hipLaunchKernelGGL(k_mcf_eval, dim3((nz + per - 1) / per), dim3(64), 0, s, p, nfl, A, B, W, q, nz, z, G);
  ED_HIP(hipGetLastError());
}

}  // namespace ed
