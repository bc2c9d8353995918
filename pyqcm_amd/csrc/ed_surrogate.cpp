// ed_surrogate.cpp -- host side of the learned surrogate: parameters, scaling,
// the training loop (Adam on fp32 masters, bf16 copies for the GEMMs) and
// batched inference.  No reference counterpart: pyqcm has no learned model;
// the surrogate proposes VCA / CDMFT candidates that the exact solver verifies.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "ed_surrogate.hpp"

namespace ed {

namespace {
int round8(int n) { return (n + 7) / 8 * 8; }

// deterministic standard normals (Box-Muller on a 64-bit Mersenne twister), so
// that an initialization depends on the seed only
struct Normal {
  std::mt19937_64 g;
  bool have = false;
  double keep = 0.0;
  explicit Normal(uint64_t seed) : g(seed) {}
  double u() { return ((g() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double operator()() {
    if (have) {
      have = false;
      return keep;
    }
    const double r = std::sqrt(-2.0 * std::log(u())), th = 6.283185307179586 * u();
    keep = r * std::sin(th);
    have = true;
    return r * std::cos(th);
  }
};

void zero(float *p, size_t n) {
  if (n) ED_HIP(hipMemsetAsync(p, 0, n * sizeof(float), stream()));
}
void zero(bf16_t *p, size_t n) {
  if (n) ED_HIP(hipMemsetAsync(p, 0, n * sizeof(bf16_t), stream()));
}
}  // namespace

Surrogate::Surrogate(int ni, int no, int w, int d, uint64_t seed) : n_in(ni), n_out(no), width(w), depth(d) {
  if (ni < 1 || no < 1 || d < 1 || w < 8 || w % 8) fail(ED_ERR_INVALID, "surrogate: bad layer sizes (width a multiple of 8)");
  Normal nrm(seed);
  layers.resize(d + 1);
  for (int l = 0; l <= d; ++l) {
    Layer &L = layers[l];
    L.in = l == 0 ? ni : w;
    L.out = l == d ? no : w;
    L.in_pad = round8(L.in);
    L.out_pad = round8(L.out);
    const size_t nw = (size_t)L.out_pad * L.in_pad;
    std::vector<float> W(nw, 0.f), b(L.out_pad, 0.f);
    const double sc = 1.0 / std::sqrt((double)L.in);
    for (int o = 0; o < L.out; ++o)
      for (int i = 0; i < L.in; ++i) W[(size_t)o * L.in_pad + i] = (float)(sc * nrm());
    L.W.upload(W);
    L.b.upload(b);
    for (DevBuf<float> *B : {&L.gW, &L.mW, &L.vW}) {
      B->alloc(nw);
      zero(B->p, nw);
    }
    for (DevBuf<float> *B : {&L.gb, &L.mb, &L.vb}) {
      B->alloc(L.out_pad);
      zero(B->p, L.out_pad);
    }
    L.Wb.alloc(nw);
    mlp_to_bf16(nw, L.W.p, L.Wb.p, stream());
  }
  x_mean.assign(ni, 0.0);
  x_std.assign(ni, 1.0);
  y_mean.assign(no, 0.0);
  y_std.assign(no, 1.0);
  loss_w.assign(no, 1.0);
  std::vector<float> lw(layers[d].out_pad, 0.f);
  for (int o = 0; o < no; ++o) lw[o] = 1.f;
  dloss_w.upload(lw);
  sync();
}

size_t Surrogate::n_params() const {
  size_t n = 0;
  for (const Layer &L : layers) n += (size_t)L.out * L.in + L.out;
  return n;
}

void Surrogate::set_scaling(const double *xm, const double *xs, const double *ym, const double *ys, const double *w) {
  for (int i = 0; i < n_in; ++i) {
    if (xs && !(xs[i] > 0.0)) fail(ED_ERR_INVALID, "surrogate: input scales must be positive");
    if (xm) x_mean[i] = xm[i];
    if (xs) x_std[i] = xs[i];
  }
  for (int o = 0; o < n_out; ++o) {
    if (ys && !(ys[o] > 0.0)) fail(ED_ERR_INVALID, "surrogate: output scales must be positive");
    if (w && !(w[o] >= 0.0)) fail(ED_ERR_INVALID, "surrogate: loss weights must be non-negative");
    if (ym) y_mean[o] = ym[o];
    if (ys) y_std[o] = ys[o];
    if (w) loss_w[o] = w[o];
  }
  std::vector<float> lw(layers.back().out_pad, 0.f);
  for (int o = 0; o < n_out; ++o) lw[o] = (float)loss_w[o];
  dloss_w.upload(lw);
  sync();
}

void Surrogate::get_weights(double *w) const {
  size_t k = 0;
  for (const Layer &L : layers) {
    std::vector<float> W = L.W.to_host(), b = L.b.to_host();
    for (int o = 0; o < L.out; ++o)
      for (int i = 0; i < L.in; ++i) w[k++] = W[(size_t)o * L.in_pad + i];
    for (int o = 0; o < L.out; ++o) w[k++] = b[o];
  }
}

void Surrogate::set_weights(const double *w) {
  size_t k = 0;
  for (Layer &L : layers) {
    std::vector<float> W((size_t)L.out_pad * L.in_pad, 0.f), b(L.out_pad, 0.f);
    for (int o = 0; o < L.out; ++o)
      for (int i = 0; i < L.in; ++i) W[(size_t)o * L.in_pad + i] = (float)w[k++];
    for (int o = 0; o < L.out; ++o) b[o] = (float)w[k++];
    L.W.upload(W);
    L.b.upload(b);
    mlp_to_bf16(W.size(), L.W.p, L.Wb.p, stream());
    // a new starting point: the optimizer restarts
    zero(L.mW.p, L.mW.n);
    zero(L.vW.p, L.vW.n);
    zero(L.mb.p, L.mb.n);
    zero(L.vb.p, L.vb.n);
  }
  adam_t = 0;
  sync();
}

void Surrogate::ensure_batch(int B) {
  if (B <= cap_B) return;
  const Layer &L0 = layers[0], &LN = layers.back();
  xb.alloc((size_t)B * L0.in_pad);
  zero(xb.p, xb.n);
  act_b.alloc((size_t)depth * B * width);
  z.alloc((size_t)depth * B * width);
  y.alloc((size_t)B * LN.out_pad);
  t.alloc((size_t)B * LN.out_pad);
  zero(t.p, t.n);
  dz_b.alloc((size_t)B * std::max(width, LN.out_pad));
  zero(dz_b.p, dz_b.n);
  dz2_b.alloc((size_t)B * width);
  part.alloc(1024);
  // split-k partials of the weight gradients: S slabs of the largest layer
  split_S = std::max(1, std::min(8, B / 256));
  size_t wmax = 0, omax = 0;
  for (const Layer &L : layers) {
    wmax = std::max(wmax, (size_t)L.out_pad * L.in_pad);
    omax = std::max(omax, (size_t)L.out_pad);
  }
  split.alloc(std::max((size_t)split_S * wmax, (size_t)mlp_colsum_chunks(B) * omax));
  cap_B = B;
}

void Surrogate::forward(int B, const bf16_t *x) {
  for (int l = 0; l <= depth; ++l) {
    const Layer &L = layers[l];
    GemmArgs g;
    g.M = B;
    g.N = L.out;
    g.K = L.in;
    g.A = l == 0 ? x : act_b.p + (size_t)(l - 1) * B * width;
    g.sam = L.in_pad;
    g.sak = 1;
    g.Kpad = L.in_pad;
    g.B = L.Wb.p;
    g.sbk = 1;
    g.sbn = L.in_pad;
    g.bias = L.b.p;
    if (l < depth) {
      g.Cf = z.p + (size_t)l * B * width;
      g.ldc = width;
      g.Cb = act_b.p + (size_t)l * B * width;
      g.ldcb = width;
      mlp_gemm(g, EPI_BIAS_GELU, stream());
    } else {
      g.Cf = y.p;
      g.ldc = L.out_pad;
      mlp_gemm(g, EPI_BIAS, stream());
    }
  }
}

int Surrogate::backward(int B, float *partials) {
  const Layer &LN = layers.back();
  const int nb = mlp_loss_grad(B, n_out, LN.out_pad, y.p, t.p, dloss_w.p, dz_b.p, partials, stream());
  // backward: dZ of the output layer is in dz_b (row pitch out_pad)
  bf16_t *dcur = dz_b.p, *dnext = dz2_b.p;
  int ldcur = LN.out_pad;
  for (int l = depth; l >= 0; --l) {
    Layer &L = layers[l];
    const bf16_t *Aprev = l == 0 ? xb.p : act_b.p + (size_t)(l - 1) * B * width;
    // dW = dZ^T A_prev (split over the batch, fixed-order reduction)
    GemmArgs g;
    g.M = L.out;
    g.N = L.in;
    g.K = B;
    g.Kpad = round8(B);
    g.A = dcur;
    g.sam = 1;
    g.sak = ldcur;
    g.Mpad = ldcur;
    g.B = Aprev;
    g.sbk = L.in_pad;
    g.sbn = 1;
    g.Npad = L.in_pad;
    g.Cf = split.p;
    g.ldc = L.in_pad;
    g.ksplit = split_S;
    g.split_stride = (long)L.out_pad * L.in_pad;
    mlp_gemm(g, EPI_STORE, stream());
    mlp_split_reduce(split_S, L.out, L.in, L.in_pad, g.split_stride, split.p, L.gW.p, stream());
    // the split partials are folded: their buffer holds the column-sum chunks
    mlp_colsum(B, L.out, ldcur, dcur, L.gb.p, split.p, stream());
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l) * gelu'(Z_{l-1})
    GemmArgs h;
    h.M = B;
    h.N = L.in;
    h.K = L.out;
    h.A = dcur;
    h.sam = ldcur;
    h.sak = 1;
    h.Kpad = ldcur;
    h.B = L.Wb.p;
    h.sbk = L.in_pad;
    h.sbn = 1;
    h.Npad = L.in_pad;
    h.Z = z.p + (size_t)(l - 1) * B * width;
    h.ldz = width;
    h.Cb = dnext;
    h.ldcb = width;
    mlp_gemm(h, EPI_GELU_GRAD, stream());
    // the output layer's gradient is no longer needed: the two buffers alternate
    std::swap(dcur, dnext);
    ldcur = width;
  }
  return nb;
}

double Surrogate::loss_and_grad(int B) {
  const int nb = backward(B, part.p);
  std::vector<float> hp(nb);
  part.download(hp.data(), nb);
  return mlp_fold(hp.data(), nb) / B;
}

void Surrogate::adam_step() {
  ++adam_t;
  const float c1 = (float)(1.0 - std::pow(beta1, (double)adam_t)), c2 = (float)(1.0 - std::pow(beta2, (double)adam_t));
  for (Layer &L : layers) {
    mlp_adam(L.W.n, L.W.p, L.gW.p, L.mW.p, L.vW.p, L.Wb.p, (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2,
             stream());
    mlp_adam(L.b.n, L.b.p, L.gb.p, L.mb.p, L.vb.p, nullptr, (float)lr, (float)beta1, (float)beta2, (float)eps, c1,
             c2, stream());
  }
}

void Surrogate::upload_inputs(int n, const double *x, DevBuf<bf16_t> &dst) const {
  const int ip = layers[0].in_pad;
  std::vector<float> f((size_t)n * ip, 0.f);
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < n_in; ++i) f[(size_t)r * ip + i] = (float)((x[(size_t)r * n_in + i] - x_mean[i]) / x_std[i]);
  DevBuf<float> tmp;
  tmp.upload(f);
  if (dst.n < f.size()) dst.alloc(f.size());
  mlp_to_bf16(f.size(), tmp.p, dst.p, stream());
  sync();
}

void Surrogate::upload_targets(int n, const double *yv, DevBuf<float> &dst) const {
  const int op = layers.back().out_pad;
  std::vector<float> f((size_t)n * op, 0.f);
  for (int r = 0; r < n; ++r)
    for (int o = 0; o < n_out; ++o) f[(size_t)r * op + o] = (float)((yv[(size_t)r * n_out + o] - y_mean[o]) / y_std[o]);
  if (dst.n < f.size()) dst.alloc(f.size());
  dst.upload(f);
  sync();
}

void Surrogate::predict(int n, const double *x, double *yout) {
  const int chunk = 16384, op = layers.back().out_pad;
  for (int r0 = 0; r0 < n; r0 += chunk) {
    const int B = std::min(chunk, n - r0);
    ensure_batch(B);
    upload_inputs(B, x + (size_t)r0 * n_in, xb);
    forward(B, xb.p);
    std::vector<float> h((size_t)B * op);
    y.download(h.data(), h.size());
    for (int r = 0; r < B; ++r)
      for (int o = 0; o < n_out; ++o)
        yout[(size_t)(r0 + r) * n_out + o] = (double)h[(size_t)r * op + o] * y_std[o] + y_mean[o];
  }
}

double Surrogate::loss_grad_host(int n, const double *x, const double *yv, double *grad) {
  if (n < 1) fail(ED_ERR_INVALID, "surrogate: empty batch");
  ensure_batch(n);
  upload_inputs(n, x, xb);
  upload_targets(n, yv, t);
  forward(n, xb.p);
  const double loss = loss_and_grad(n);
  if (grad) {
    size_t k = 0;
    for (const Layer &L : layers) {
      std::vector<float> W = L.gW.to_host(), b = L.gb.to_host();
      for (int o = 0; o < L.out; ++o)
        for (int i = 0; i < L.in; ++i) grad[k++] = W[(size_t)o * L.in_pad + i];
      for (int o = 0; o < L.out; ++o) grad[k++] = b[o];
    }
  }
  return loss;
}

std::vector<double> Surrogate::train(int n, const double *x, const double *yv, int epochs, int batch, uint64_t seed,
                                     double *ms_per_step) {
  if (n < 1 || epochs < 0 || batch < 1) fail(ED_ERR_INVALID, "surrogate: bad training arguments");
  batch = std::min(batch, n);
  const int ip = layers[0].in_pad, op = layers.back().out_pad;
  DevBuf<bf16_t> X;
  DevBuf<float> T;
  upload_inputs(n, x, X);
  upload_targets(n, yv, T);
  ensure_batch(batch);
  const int steps = n / batch;  // full batches only: every step sees the same batch size
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::mt19937_64 g(seed);
  DevBuf<int> didx(n);
  DevBuf<float> eparts((size_t)std::max(steps, 1) * 1024);
  std::vector<int> nbs(std::max(steps, 1));
  std::vector<float> hp;
  std::vector<double> hist;
  double ms_total = 0.0;
  long nsteps = 0;
  hipEvent_t e0, e1;
  ED_HIP(hipEventCreate(&e0));
  ED_HIP(hipEventCreate(&e1));
  for (int ep = 0; ep < epochs; ++ep) {
    check_cancel();
    for (int i = n - 1; i > 0; --i) std::swap(idx[i], idx[(size_t)(g() % (uint64_t)(i + 1))]);
    didx.upload(idx);
    ED_HIP(hipEventRecord(e0, stream()));
    for (int s = 0; s < steps; ++s) {
      const int *bi = didx.p + (size_t)s * batch;
      mlp_gather_bf16(batch, ip, bi, X.p, xb.p, stream());
      mlp_gather_f32(batch, op, bi, T.p, t.p, stream());
      forward(batch, xb.p);
      nbs[s] = backward(batch, eparts.p + (size_t)s * 1024);
      adam_step();
    }
    ED_HIP(hipEventRecord(e1, stream()));
    ED_HIP(hipEventSynchronize(e1));
    // the epoch's losses, read once the epoch has run
    hp.resize((size_t)std::max(steps, 1) * 1024);
    if (steps) eparts.download(hp.data(), (size_t)steps * 1024);
    double sum = 0.0;
    for (int s = 0; s < steps; ++s) sum += mlp_fold(hp.data() + (size_t)s * 1024, nbs[s]) / batch;
    float ms = 0.f;
    ED_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms_total += ms;
    nsteps += steps;
    hist.push_back(steps ? sum / steps : 0.0);
  }
  ED_HIP(hipEventDestroy(e0));
  ED_HIP(hipEventDestroy(e1));
  if (ms_per_step) *ms_per_step = nsteps ? ms_total / nsteps : 0.0;
  return hist;
}

void Surrogate::time(int batch, int nrep, double *ms_fwd, double *ms_fb) {
  ensure_batch(batch);
  // inputs: a fixed pattern of normalized values; targets zero
  std::vector<double> x((size_t)batch * n_in);
  for (size_t i = 0; i < x.size(); ++i) x[i] = x_mean[i % n_in] + x_std[i % n_in] * std::sin(0.37 * (double)i);
  upload_inputs(batch, x.data(), xb);
  zero(t.p, (size_t)batch * layers.back().out_pad);
  hipEvent_t e0, e1, e2;
  ED_HIP(hipEventCreate(&e0));
  ED_HIP(hipEventCreate(&e1));
  ED_HIP(hipEventCreate(&e2));
  forward(batch, xb.p);
  backward(batch, part.p);
  ED_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < nrep; ++r) forward(batch, xb.p);
  ED_HIP(hipEventRecord(e1, stream()));
  for (int r = 0; r < nrep; ++r) {
    forward(batch, xb.p);
    backward(batch, part.p);
  }
  ED_HIP(hipEventRecord(e2, stream()));
  ED_HIP(hipEventSynchronize(e2));
  float a = 0.f, b = 0.f;
  ED_HIP(hipEventElapsedTime(&a, e0, e1));
  ED_HIP(hipEventElapsedTime(&b, e1, e2));
  *ms_fwd = a / nrep;
  *ms_fb = b / nrep;
  ED_HIP(hipEventDestroy(e0));
  ED_HIP(hipEventDestroy(e1));
  ED_HIP(hipEventDestroy(e2));
}

}  // namespace ed
