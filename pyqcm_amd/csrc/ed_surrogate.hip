// ed_surrogate.hip -- kernels of the learned surrogate (gfx950).
//
// The MLP's layers are GEMMs on the matrix cores: bf16 operands staged in LDS,
// v_mfma_f32_16x16x32_bf16 with fp32 accumulation, and the bias / GELU /
// GELU-derivative work fused into the GEMM epilogue.  Master weights, Adam
// moments, pre-activations and gradients are fp32; activations and the
// GEMM-side gradients are bf16.  Every reduction has a fixed order.
#include <algorithm>
#include <cmath>

#include "ed_surrogate.hpp"

namespace ed {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 32;
constexpr int kLdsPitch = kBK + 8;  // bf16 per LDS row: 80 B, keeps the 16-byte fragment reads aligned

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// One 64 x 64 tile of C per block of 4 waves (2 x 2); each wave owns 32 x 32 =
// 2 x 2 MFMA tiles of 16 x 16, accumulated over k in steps of 32.
// Fragment layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies row
// l % 16 of A and column l % 16 of B, each with the 8 consecutive k values
// 8 * (l / 16) .. + 7; it receives D rows 4 * (l / 16) .. + 3 of column l % 16.
template <int EPI, bool AKC, bool BKC>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16_t As[kBM * kLdsPitch];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[kBN * kLdsPitch];
  const int t = threadIdx.x, l = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  const int nkt = (g.K + kBK - 1) / kBK;
  const int per = (nkt + g.ksplit - 1) / g.ksplit;
  const int kt0 = blockIdx.z * per, kt1 = min(nkt, kt0 + per);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = {};
  for (int kt = kt0; kt < kt1; ++kt) {
    const int k0 = kt * kBK;
    // stage A (64 x 32) and B (32 x 64): 256 chunks of 8 bf16 each
    if (AKC) {
      const int r = t >> 2, kc = (t & 3) * 8, m = m0 + r, k = k0 + kc;
      bf16x8 v = zero8;
      if (m < g.M && k < g.Kpad) v = *(const bf16x8 *)(g.A + (long)m * g.sam + k);
      *(bf16x8 *)&As[r * kLdsPitch + kc] = v;
    } else {
      const int kk = t >> 3, mc = (t & 7) * 8, k = k0 + kk, m = m0 + mc;
      bf16x8 v = zero8;
      if (k < g.K && m < g.Mpad) v = *(const bf16x8 *)(g.A + (long)k * g.sak + m);
#pragma unroll
      for (int j = 0; j < 8; ++j) As[(mc + j) * kLdsPitch + kk] = v[j];
    }
    if (BKC) {
      const int r = t >> 2, kc = (t & 3) * 8, n = n0 + r, k = k0 + kc;
      bf16x8 v = zero8;
      if (n < g.N && k < g.Kpad) v = *(const bf16x8 *)(g.B + (long)n * g.sbn + k);
      *(bf16x8 *)&Bs[r * kLdsPitch + kc] = v;
    } else {
      const int kk = t >> 3, nc = (t & 7) * 8, k = k0 + kk, n = n0 + nc;
      bf16x8 v = zero8;
      if (k < g.K && n < g.Npad) v = *(const bf16x8 *)(g.B + (long)k * g.sbk + n);
#pragma unroll
      for (int j = 0; j < 8; ++j) Bs[(nc + j) * kLdsPitch + kk] = v[j];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *(const bf16x8 *)&As[(wm * 32 + i * 16 + (l & 15)) * kLdsPitch + 8 * (l >> 4)];
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *(const bf16x8 *)&Bs[(wn * 32 + j * 16 + (l & 15)) * kLdsPitch + 8 * (l >> 4)];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  float *Cf = g.Cf;
  if (EPI == EPI_STORE) Cf += (long)blockIdx.z * g.split_stride;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (l & 15);
      if (n >= g.N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (l >> 4) + r;
        if (m >= g.M) continue;
        const float v = acc[i][j][r];
        if (EPI == EPI_BIAS_GELU) {
          const float zv = v + g.bias[n];
          Cf[(long)m * g.ldc + n] = zv;
          g.Cb[(long)m * g.ldcb + n] = (bf16_t)gelu(zv);
        } else if (EPI == EPI_BIAS) {
          Cf[(long)m * g.ldc + n] = v + g.bias[n];
        } else if (EPI == EPI_GELU_GRAD) {
          g.Cb[(long)m * g.ldcb + n] = (bf16_t)(v * gelu_grad(g.Z[(long)m * g.ldz + n]));
        } else {
          Cf[(long)m * g.ldc + n] = v;
        }
      }
    }
}

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_to_bf16(size_t n, const float *in, bf16_t *out) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = (bf16_t)in[i];
}

__global__ __launch_bounds__(256) void k_loss_grad(int B, int n_out, int ld, const float *y, const float *t,
                                                   const float *w, bf16_t *dy, float *part) {
  __shared__ float red[4];
  float s = 0.f;
  const size_t n = (size_t)B * ld;
  const float scale = 2.f / (float)B;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int o = (int)(i % ld);
    float d = 0.f;
    if (o < n_out) {
      const float r = y[i] - t[i];
      s += w[o] * r * r;
      d = scale * w[o] * r;
    }
    dy[i] = (bf16_t)d;
  }
  s = wave_sum_f(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// column sums in two fixed-order stages: block (x, r) sums rows [r rc, (r+1) rc)
// of 64 columns (4 row groups) into part[r N + n]; k_colsum_fold adds the R
// row chunks in order
__global__ __launch_bounds__(256) void k_colsum(int M, int N, int ld, int rc, const bf16_t *dz, float *part) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, rg = threadIdx.x >> 6, n = blockIdx.x * 64 + c;
  const int m0 = blockIdx.y * rc, m1 = min(M, m0 + rc);
  float s = 0.f;
  if (n < N)
    for (int m = m0 + rg; m < m1; m += 4) s += (float)dz[(long)m * ld + n];
  red[rg][c] = s;
  __syncthreads();
  if (rg == 0 && n < N) part[(long)blockIdx.y * N + n] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}
// 64 columns per workgroup, 4 row groups over the R chunks, fixed order
__global__ __launch_bounds__(256) void k_colsum_fold(int R, int N, const float *part, float *db) {
  __shared__ float red[4][64];
  const int c = threadIdx.x & 63, rg = threadIdx.x >> 6, n = blockIdx.x * 64 + c;
  float s = 0.f;
  if (n < N)
    for (int r = rg; r < R; r += 4) s += part[(long)r * N + n];
  red[rg][c] = s;
  __syncthreads();
  if (rg == 0 && n < N) db[n] = (red[0][c] + red[1][c]) + (red[2][c] + red[3][c]);
}

// only the R x C logical entries are written: the padding of out keeps its zeros
__global__ __launch_bounds__(256) void k_split_reduce(int S, int R, int C, int ld, long stride, const float *part,
                                                      float *out) {
  const size_t n = (size_t)R * C;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const size_t e = (i / C) * ld + i % C;
    float s = 0.f;
    for (int z = 0; z < S; ++z) s += part[(long)z * stride + e];
    out[e] = s;
  }
}

__global__ __launch_bounds__(256) void k_adam(size_t n, float *p, const float *g, float *m, float *v, bf16_t *pb,
                                              float lr, float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = (bf16_t)pi;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(int n, int ld, const int *idx, const T *src, T *dst) {
  const size_t tot = (size_t)n * ld;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < tot; i += (size_t)gridDim.x * 256) {
    const size_t r = i / ld, c = i % ld;
    dst[i] = src[(size_t)idx[r] * ld + c];
  }
}

int grid1(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096)); }

}  // namespace

void mlp_gemm(const GemmArgs &g, int epi, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return;
  const bool akc = g.sak == 1, bkc = g.sbk == 1;
  if (!akc && g.sam != 1) fail(ED_ERR_INTERNAL, "mlp_gemm: A has no contiguous axis");
  if (!bkc && g.sbn != 1) fail(ED_ERR_INTERNAL, "mlp_gemm: B has no contiguous axis");
  if (g.Kpad % 8 || (!akc && g.Mpad % 8) || (!bkc && g.Npad % 8) || g.Kpad < g.K)
    fail(ED_ERR_INTERNAL, "mlp_gemm: padded extents must be multiples of 8");
  if ((akc && g.sam % 8) || (!akc && g.sak % 8) || (bkc && g.sbn % 8) || (!bkc && g.sbk % 8) ||
      (reinterpret_cast<uintptr_t>(g.A) & 15) || (reinterpret_cast<uintptr_t>(g.B) & 15))
    fail(ED_ERR_INTERNAL, "mlp_gemm: operands must be 16-byte aligned rows");
  if (epi != EPI_STORE && g.ksplit != 1) fail(ED_ERR_INTERNAL, "mlp_gemm: split k only for plain stores");
  const dim3 grid((g.M + kBM - 1) / kBM, (g.N + kBN - 1) / kBN, std::max(1, g.ksplit));
#define ED_GEMM(E)                                                                                      \
  do {                                                                                                  \
    if (akc && bkc) hipLaunchKernelGGL((k_gemm<E, true, true>), grid, dim3(256), 0, s, g);              \
    else if (akc) hipLaunchKernelGGL((k_gemm<E, true, false>), grid, dim3(256), 0, s, g);               \
    else if (bkc) hipLaunchKernelGGL((k_gemm<E, false, true>), grid, dim3(256), 0, s, g);               \
    else hipLaunchKernelGGL((k_gemm<E, false, false>), grid, dim3(256), 0, s, g);                       \
  } while (0)
  switch (epi) {
    case EPI_BIAS_GELU: ED_GEMM(EPI_BIAS_GELU); break;
    case EPI_BIAS: ED_GEMM(EPI_BIAS); break;
    case EPI_GELU_GRAD: ED_GEMM(EPI_GELU_GRAD); break;
    default: ED_GEMM(EPI_STORE); break;
  }
#undef ED_GEMM
  ED_HIP(hipGetLastError());
}

void mlp_to_bf16(size_t n, const float *in, bf16_t *out, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1(n)), dim3(256), 0, s, n, in, out);
  ED_HIP(hipGetLastError());
}

int mlp_loss_grad(int B, int n_out, int ld, const float *y, const float *t, const float *w, bf16_t *dy,
                  float *partials, hipStream_t s) {
  const int g = std::min(grid1((size_t)B * ld), 1024);
  hipLaunchKernelGGL(k_loss_grad, dim3(g), dim3(256), 0, s, B, n_out, ld, y, t, w, dy, partials);
  ED_HIP(hipGetLastError());
  return g;
}

int mlp_colsum_chunks(int M) { return std::max(1, std::min(64, (M + 63) / 64)); }

void mlp_colsum(int M, int N, int ld, const bf16_t *dz, float *db, float *part, hipStream_t s) {
  if (N <= 0) return;
  const int R = mlp_colsum_chunks(M);
  const int rc = (M + R - 1) / R;
  hipLaunchKernelGGL(k_colsum, dim3((N + 63) / 64, R), dim3(256), 0, s, M, N, ld, rc, dz, part);
  ED_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_colsum_fold, dim3((N + 63) / 64), dim3(256), 0, s, R, N, (const float *)part, db);
  ED_HIP(hipGetLastError());
}

void mlp_split_reduce(int S, int R, int C, int ld, long stride, const float *part, float *out, hipStream_t s) {
  if (R <= 0 || C <= 0) return;
  hipLaunchKernelGGL(k_split_reduce, dim3(grid1((size_t)R * C)), dim3(256), 0, s, S, R, C, ld, stride, part, out);
  ED_HIP(hipGetLastError());
}

void mlp_adam(size_t n, float *p, const float *g, float *m, float *v, bf16_t *pb, float lr, float b1, float b2,
              float eps, float c1, float c2, hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, s, n, p, g, m, v, pb, lr, b1, b2, eps, c1, c2);
  ED_HIP(hipGetLastError());
}

void mlp_gather_bf16(int n, int ld, const int *idx, const bf16_t *src, bf16_t *dst, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather<bf16_t>, dim3(grid1((size_t)n * ld)), dim3(256), 0, s, n, ld, idx, src, dst);
  ED_HIP(hipGetLastError());
}

void mlp_gather_f32(int n, int ld, const int *idx, const float *src, float *dst, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather<float>, dim3(grid1((size_t)n * ld)), dim3(256), 0, s, n, ld, idx, src, dst);
  ED_HIP(hipGetLastError());
}

double mlp_fold(const float *partials, int n) {
  std::vector<double> w(partials, partials + std::max(n, 0));
  size_t k = w.size();
  while (k > 1) {
    const size_t half = (k + 1) / 2;
    for (size_t i = 0; i + half < k; ++i) w[i] += w[i + half];
    k = half;
  }
  return k ? w[0] : 0.0;
}

}  // namespace ed
