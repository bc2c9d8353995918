// ed_surrogate.hpp -- learned surrogate of the impurity solver: a GELU MLP
// (cluster parameters -> E0 and truncated continued fractions) trained and
// evaluated on the device with bf16 MFMA GEMMs and fp32 accumulation.
#pragma once

#include <cstdint>
#include <vector>

#include "ed_common.hpp"

namespace ed {

// ---- kernels (ed_surrogate.hip) ----
typedef __bf16 bf16_t;

enum GemmEpi {
  EPI_BIAS_GELU = 0,  // Cf = acc + bias (pre-activation, fp32), Cb = bf16(gelu(Cf))
  EPI_BIAS = 1,       // Cf = acc + bias
  EPI_GELU_GRAD = 2,  // Cb = bf16(acc * gelu'(Z)), Z the fp32 pre-activation of the same shape
  EPI_STORE = 3       // Cf (+ z * split_stride for split z) = acc
};

// C(m, n) = sum_k A(m, k) B(k, n); A(m, k) = A[m * sam + k * sak] with sak == 1
// or sam == 1, B(k, n) = B[k * sbk + n * sbn] with sbk == 1 or sbn == 1.  The
// contiguous extents of A and B are padded to multiples of 8 elements (zeros
// beyond the logical size), and the pointers 16-byte aligned.
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  const bf16_t *A = nullptr;
  long sam = 0, sak = 0;
  const bf16_t *B = nullptr;
  long sbk = 0, sbn = 0;
  int Kpad = 0;             // padded extent of the k axis in memory (>= K, multiple of 8)
  int Mpad = 0, Npad = 0;   // padded extents of m / n when those axes are contiguous
  const float *bias = nullptr;
  float *Cf = nullptr;
  long ldc = 0;
  bf16_t *Cb = nullptr;
  long ldcb = 0;
  const float *Z = nullptr;
  long ldz = 0;
  int ksplit = 1;           // split of the k range over blockIdx.z (EPI_STORE)
  long split_stride = 0;
};
void mlp_gemm(const GemmArgs &g, int epi, hipStream_t s);

// out[i] = bf16(in[i])
void mlp_to_bf16(size_t n, const float *in, bf16_t *out, hipStream_t s);
// loss partials and output gradient: r = y - t on the first n_out columns,
// dy = bf16(2 w_o r / B) (zero in the padding), per-block partials of sum w_o r^2
int mlp_loss_grad(int B, int n_out, int ld, const float *y, const float *t, const float *w, bf16_t *dy,
                  float *partials, hipStream_t s);
// db[n] = sum_m dz[m * ld + n], n < N (fixed order, through part: at least
// mlp_colsum_chunks(M) x N floats of scratch)
int mlp_colsum_chunks(int M);
void mlp_colsum(int M, int N, int ld, const bf16_t *dz, float *db, float *part, hipStream_t s);
// out[r * ld + c] = sum_z part[z * stride + r * ld + c] for r < R, c < C (fixed
// order); the padding columns of out are left as they are
void mlp_split_reduce(int S, int R, int C, int ld, long stride, const float *part, float *out, hipStream_t s);
// Adam on fp32 masters p with gradient g; writes the bf16 copy pb (when not null)
void mlp_adam(size_t n, float *p, const float *g, float *m, float *v, bf16_t *pb, float lr, float b1, float b2,
              float eps, float c1, float c2, hipStream_t s);
// dst rows r < n: dst[r * ld + c] = src[idx[r] * ld + c]
void mlp_gather_bf16(int n, int ld, const int *idx, const bf16_t *src, bf16_t *dst, hipStream_t s);
void mlp_gather_f32(int n, int ld, const int *idx, const float *src, float *dst, hipStream_t s);
// fold of float partials into one double (host): deterministic order
double mlp_fold(const float *partials, int n);

// ---- host model (ed_surrogate.cpp) ----
struct Surrogate {
  struct Layer {
    int in = 0, out = 0, in_pad = 0, out_pad = 0;
    DevBuf<float> W, b, gW, gb, mW, vW, mb, vb;
    DevBuf<bf16_t> Wb;
  };
  int n_in = 0, n_out = 0, width = 0, depth = 0;
  std::vector<Layer> layers;
  std::vector<double> x_mean, x_std, y_mean, y_std, loss_w;
  DevBuf<float> dloss_w;
  long adam_t = 0;
  double lr = 1e-3, beta1 = 0.9, beta2 = 0.999, eps = 1e-8;
  // work buffers, sized for the current batch
  int cap_B = 0;
  DevBuf<bf16_t> xb, act_b, dz_b, dz2_b;  // activations (depth x B x width), output gradients
  DevBuf<float> z, y, t, part, split;
  int split_S = 0;

  Surrogate(int n_in, int n_out, int width, int depth, uint64_t seed);
  size_t n_params() const;
  void set_scaling(const double *xm, const double *xs, const double *ym, const double *ys, const double *w);
  void get_weights(double *w) const;
  void set_weights(const double *w);
  // device forward on normalized bf16 inputs (B rows, in_pad columns) -> y (fp32, normalized)
  void forward(int B, const bf16_t *x);
  // gradients (into gW / gb) of B rows already in xb / t after forward(B, xb):
  // the loss partials go to partials (returns their number)
  int backward(int B, float *partials);
  // the same, returning the mean weighted loss (waits for the device)
  double loss_and_grad(int B);
  void adam_step();
  // host helpers
  void upload_inputs(int n, const double *x, DevBuf<bf16_t> &dst) const;
  void upload_targets(int n, const double *y, DevBuf<float> &dst) const;
  void ensure_batch(int B);
  void predict(int n, const double *x, double *y);
  double loss_grad_host(int n, const double *x, const double *y, double *grad);
  // epochs over the n samples in shuffled mini-batches; returns the mean
  // training loss of each epoch (normalized units)
  std::vector<double> train(int n, const double *x, const double *y, int epochs, int batch, uint64_t seed,
                            double *ms_per_step);
  // device time of forward and forward + backward on `batch` resident rows
  void time(int batch, int nrep, double *ms_fwd, double *ms_fb);
};

}  // namespace ed
