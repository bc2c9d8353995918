// ed_solvers.cpp -- Krylov drivers: host control flow, device vectors.
//
// Every vector operation runs on the calling thread's stream; the host only sees the
// scalars of the recurrences (alpha, beta, Gram coefficients), obtained from
// deterministic per-block partial sums.  All recurrences are FP64.
#include "ed_solvers.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>

#include "ed_dense.hpp"

namespace ed {

namespace {

void d2d(double *dst, const double *src, size_t n) {
  if (n) ED_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
}

// dots of a with each b (chunks of 16 vectors per kernel)
std::vector<double> dots(size_t n, const double *a, const std::vector<const double *> &b) {
  std::vector<double> out(b.size(), 0.0);
  for (size_t s = 0; s < b.size(); s += 16) {
    const int m = (int)std::min<size_t>(16, b.size() - s);
    double *pp = partials_acquire();
    const int nb = multi_dot(n, a, m, b.data() + s, pp, stream());
    std::vector<double> r;
    sum_partials(pp, nb, m, &r);
    for (int k = 0; k < m; ++k) out[s + k] = r[k];
  }
  return out;
}

// t = st t + sum c_k b_k (chunks); returns ||t||^2 if want_norm
double comb(size_t n, double *t, double st, const std::vector<const double *> &b, const std::vector<double> &c,
            bool want_norm, const double *w = nullptr, double *dot_w = nullptr) {
  // partials: the norm in row 0 and the dot in row 1 (stride partials_cap),
  // folded together by one reduction
  auto run = [&](double stv, int m, const double *const *bb, const double *cc, bool wn, bool wd) -> double {
    if (!wn && !wd) {
      combine(n, t, stv, m, bb, cc, nullptr, nullptr, nullptr, stream());
      return 0.0;
    }
    double *np = partials_acquire(), *dp = np + partials_cap();
    const int nb = combine(n, t, stv, m, bb, cc, wn ? np : nullptr, wd ? w : nullptr, wd ? dp : nullptr, stream());
    if (!wn) {
      *dot_w = sum_partials(dp, nb);
      return 0.0;
    }
    if (!wd) return sum_partials(np, nb);
    std::vector<double> r;
    sum_partials_rows(np, nb, 2, partials_cap(), &r);
    *dot_w = r[1];
    return r[0];
  };
  if (b.empty()) {
    if (!want_norm && !dot_w) {
      if (st != 1.0) scale(n, t, st, stream());
      return 0.0;
    }
    return run(st, 0, nullptr, nullptr, want_norm, dot_w != nullptr);
  }
  double nn = 0.0;
  for (size_t s = 0; s < b.size(); s += 16) {
    const int m = (int)std::min<size_t>(16, b.size() - s);
    const bool last = s + 16 >= b.size();
    nn = run(s == 0 ? st : 1.0, m, b.data() + s, c.data() + s, last && want_norm, last && dot_w != nullptr);
  }
  return nn;
}

// ---------------------------------------------------------------------------
// Device-driven Lanczos recurrence (kernels lz_*) over unnormalized vectors
// U_k = b[k-1] v_k: the scalars of a step are folded, used and stored on the
// device, and the normalization of U_{k+1} is folded into the next step's
// pass, so a step is one product and one vector pass and a block of steps
// runs without a host wait; the block is queued once under stream capture and
// replayed as a hipGraph.  The host reads the new coefficients once per launch
// and applies the caller's stop rule to them in step order, so the recurrence
// it keeps is the one a per-step check would keep (the steps queued past the
// stop are discarded).  Steps rotate over three buffers (step k: C = R[k % 3],
// P = R[(k + 2) % 3], W = R[(k + 1) % 3]), hence blocks of a multiple of 3
// steps.  With slots, step k also copies v_k into a vector of its own (the
// Krylov basis kept in HBM), allocated by the host before the block that
// needs it, within a byte budget.
// ---------------------------------------------------------------------------
class LanczosDeviceRun {
 public:
  LanczosDeviceRun(Hamiltonian &H, const double *d_start, int max_steps, bool keep_basis, double budget)
      : H(H), n(H.dim) {
    Params &P = params();
    K = (int)P.geti("lanczos_block");
    graph = P.getb("lanczos_graph");
    cap = ((max_steps + K - 1) / K + 1) * K;
    for (auto &r : R) r.alloc(n);
    coef.alloc(2 * (size_t)cap);
    part.alloc(3 * (size_t)partials_cap());
    base.alloc(1);
    d2d(R[0].p, d_start, n);
    const int zero = 0;
    h2d_staged(base.p, &zero, sizeof(int));
    D.n = n;
    D.cap = cap;
    D.coef = coef.p;
    D.base = base.p;
    D.part = part.p;
    if (keep_basis) {
      slot_budget = budget;
      table.alloc((size_t)cap + 1);
      host_table.assign((size_t)cap + 1, nullptr);
      basis.emplace_back(n);
      d2d(basis[0].p, d_start, n);
      host_table[0] = basis[0].p;
      D.slots = table.p;
      D.nslots = cap + 1;
    }
    H.prepare();
    sync();
  }
  ~LanczosDeviceRun() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
  int block() const { return K; }
  int queued() const { return done; }
  // true while every queued step also stored its vector
  bool basis_complete() const { return !basis.empty() && !basis_full; }
  const double *vec(int k) const { return basis[(size_t)k].p; }
  // queue nblk blocks (fewer near the coefficient capacity), wait, and append
  // the new coefficients to a and b; false once the capacity is used up
  bool run(int nblk, std::vector<double> &a, std::vector<double> &b) {
    nblk = std::min(nblk, (cap - done) / K);
    if (nblk <= 0) return false;
    grow_basis(done + nblk * K);
    for (int r = 0; r < nblk; ++r) {
      if (!graph) {
        enqueue();
      } else if (exec) {
        ED_HIP(hipGraphLaunch(exec, stream()));
      } else if (done + r * K == 0) {
        enqueue();  // the first block runs eagerly: lazy set-up happens outside the capture
      } else {
        capture();
        ED_HIP(hipGraphLaunch(exec, stream()));
      }
    }
    std::vector<double> h(2 * (size_t)nblk * K);
    const size_t m = (size_t)nblk * K;
    ED_HIP(hipMemcpyAsync(h.data(), coef.p + done, m * sizeof(double), hipMemcpyDeviceToHost, stream()));
    ED_HIP(hipMemcpyAsync(h.data() + m, coef.p + cap + done, m * sizeof(double), hipMemcpyDeviceToHost, stream()));
    sync();
    a.insert(a.end(), h.begin(), h.begin() + m);
    b.insert(b.end(), h.begin() + m, h.end());
    done += (int)m;
    return true;
  }

 private:
  void enqueue() {
    for (int j = 0; j < K; ++j) {
      double *C = R[j % 3].p, *Pv = R[(j + 2) % 3].p, *W = R[(j + 1) % 3].p;
      const int nb = H.apply_fixed(C, W, 1.0, 0.0, part.p, nullptr, 0);
      lz_step(D, W, C, Pv, j, nb, stream());
    }
    lz_next(D, K, stream());
  }
  void capture() {
    hipGraph_t g = nullptr;
    ED_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(stream(), &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    ED_HIP(hipStreamEndCapture(stream(), &g));
    const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    ED_HIP(e);
  }
  // vectors for slots 1 .. upto (step k writes slot k) within the budget
  void grow_basis(int upto) {
    if (basis.empty() || basis_full) return;
    const size_t first = basis.size();
    while ((int)basis.size() <= upto) {
      if ((double)(basis.size() + 1) * n * 8.0 > slot_budget) {
        basis_full = true;
        break;
      }
      basis.emplace_back(n);
      host_table[basis.size() - 1] = basis.back().p;
    }
    if (basis.size() > first)
      h2d_staged(table.p + first, host_table.data() + first, (basis.size() - first) * sizeof(double *));
  }
  Hamiltonian &H;
  size_t n;
  int K = 6, cap = 0, done = 0;
  bool graph = true, basis_full = false;
  double slot_budget = 0.0;
  DevBuf<double> R[3], coef, part;
  DevBuf<int> base;
  DevBuf<double *> table;
  std::vector<double *> host_table;
  std::vector<DevBuf<double>> basis;
  LanczosDev D;
  hipGraphExec_t exec = nullptr;
};

// whether the device-driven recurrence applies (a block length that is a
// positive multiple of 3, a non-empty space)
bool device_lanczos(size_t n) {
  const long K = params().geti("lanczos_block");
  if (K == 0 || n == 0) return false;
  if (K < 0 || K % 3 != 0 || K > 96) fail(ED_ERR_INVALID, "lanczos_block must be 0 or a multiple of 3 up to 96");
  return true;
}

// blocks to queue before the next host check: the steps the last six
// log-decrements of a decreasing quantity predict to reach the target
int predict_blocks(const std::vector<double> &logq, double target, int K) {
  const size_t m = logq.size();
  if (m < 7) return 2;
  const double rate = (logq[m - 1] - logq[m - 7]) / 6.0;
  if (!(rate < 0.0)) return 2;
  const double need = (target - logq[m - 1]) / rate;
  if (!(need > 0.0)) return 1;
  return (int)std::min(8.0, std::max(1.0, std::ceil(need / K)));
}

}  // namespace

double ddot(size_t n, const double *a, const double *b) {
  const double *bb[1] = {b};
  double *pp = partials_acquire();
  const int nb = multi_dot(n, a, 1, bb, pp, stream());
  return sum_partials(pp, nb);
}

// ---------------------------------------------------------------------------
// Lanczos ground state.  Vectors are kept unnormalized: buffer A holds
// beta_{k-1} v_k (norm bA), buffer B holds the previous one; the H.v kernel
// folds the scalings (y = ax H x + b y) so no separate scaling pass runs.
// With store_lanczos_basis the Krylov vectors stay resident in HBM and the
// eigenvector is one combination pass; otherwise (or when the HBM budget is
// exceeded) the recurrence is replayed as in the reference.
// ---------------------------------------------------------------------------
// the same recurrence, device-driven: the stop rule (Ritz residual of the
// lowest eigenvalue) is applied to the coefficients in step order; the
// eigenvector is combined from the kept basis, or replayed when the basis did
// not fit the HBM budget
static LanczosGS lanczos_gs_device(Hamiltonian &H, const double *d_start, double *d_gs, bool store,
                                   double budget) {
  Params &P = params();
  const int max_iter = (int)P.geti("max_iter_lanczos");
  const double accur = P.getd("accur_lanczos");
  const size_t n = H.dim;
  LanczosDeviceRun run(H, d_start, max_iter + 1, store, budget);
  LanczosGS R;
  std::vector<double> a, b, logr;
  int nblk = std::max(1, 24 / run.block());
  bool conv = false;
  while (!conv) {
    check_cancel();
    const size_t k0 = a.size();
    if (!run.run(nblk, a, b))
      fail(ED_ERR_NOT_CONVERGED, "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations");
    for (size_t k = k0; k < a.size(); ++k) {
      R.alpha.push_back(a[k]);
      R.beta.push_back(b[k]);
      double e0, zl;
      tridiag_lowest((int)R.alpha.size(), R.alpha.data(), R.beta.data(), e0, zl,
                     R.alpha.size() > 1 ? R.E0 : std::numeric_limits<double>::quiet_NaN());
      R.E0 = e0;
      const double ritz = std::fabs(zl * b[k]);
      logr.push_back(std::log(std::max(ritz, 1e-300)));
      if ((int)R.alpha.size() > max_iter)
        fail(ED_ERR_NOT_CONVERGED, "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations");
      if (!(ritz > accur)) {
        conv = true;
        break;
      }
    }
    nblk = predict_blocks(logr, std::log(accur), run.block());
  }
  const int kk = (int)R.alpha.size();
  R.niter = kk;
  if (verbose()) vlog("device Lanczos: %d steps kept of %d queued", kk, run.queued());
  std::vector<double> z;
  double e0;
  tridiag_eigvec0(kk, R.alpha.data(), R.beta.data(), e0, z);
  if (run.basis_complete()) {
    std::vector<const double *> src;
    for (int j = 0; j < kk; ++j) src.push_back(run.vec(j));
    comb(n, d_gs, 0.0, src, z, false);
  } else {
    // LanczosEigenvector replay (Lanczos.hpp:109-143); buffer A holds
    // beta_{j-2} v_{j-1} (norm bA), B the vector before it (norm bB)
    DevBuf<double> A(n), B(n);
    double bA = 1.0, bB = 1.0;
    d2d(A.p, d_start, n);
    d2d(d_gs, d_start, n);
    scale(n, d_gs, z[0], stream());
    for (int j = 1; j < kk; ++j) {
      check_cancel();
      const double cb = (j > 1) ? -R.beta[j - 2] / bB : 0.0;
      H.apply(A.p, B.p, 1.0 / bA, cb, false);
      replay_step(n, B.p, A.p, d_gs, -R.alpha[j - 1] / bA, z[j] / R.beta[j - 1], stream());
      std::swap(A.p, B.p);
      bB = bA;
      bA = R.beta[j - 1];
    }
  }
  sync();
  return R;
}

LanczosGS lanczos_gs(Hamiltonian &H, const double *d_start, double *d_gs, const GroupProjector *proj) {
  Params &P = params();
  const int max_iter = (int)P.geti("max_iter_lanczos");
  const double accur = P.getd("accur_lanczos");
  const size_t n = H.dim;
  const double gb = P.getd("lanczos_basis_budget_GB");
  const double budget = gb >= 0.0 ? gb * 1e9 : std::max(0.0, 0.9 * (double)dev_available_bytes() - 4e9);
  const bool reorth = P.getb("lanczos_reorth");
  // a projected recurrence keeps its basis (the replay would have to project too)
  bool store = P.getb("store_lanczos_basis") || reorth || proj;
  if (!reorth && !proj && device_lanczos(n)) return lanczos_gs_device(H, d_start, d_gs, store, budget);
  LanczosGS R;
  std::vector<DevBuf<double>> basis;  // basis[k] = s_k v_k
  std::vector<double> sc;             // s_k
  DevBuf<double> A(n), B(n);
  d2d(A.p, d_start, n);
  double bA = 1.0, bB = 1.0;
  if (store) {
    basis.emplace_back(n);
    d2d(basis[0].p, d_start, n);
    sc.push_back(1.0);
  }
  double ritz = 2.0 * accur;
  while (ritz > accur) {
    check_cancel();
    const int k = (int)R.alpha.size();
    double alpha, beta;
    if (store) {
      // w = H v_k - beta_{k-1} v_{k-1} - alpha v_k into a fresh buffer
      if ((double)(basis.size() + 1) * n * 8.0 > budget) {
        if (proj)
          fail(ED_ERR_OOM, "a point-group sector keeps the Krylov basis in HBM: " + std::to_string(basis.size() + 1) +
                               " vectors exceed lanczos_basis_budget_GB");
        if (reorth)
          fail(ED_ERR_OOM, "lanczos_reorth keeps the Krylov basis in HBM: " + std::to_string(basis.size() + 1) +
                               " vectors exceed lanczos_basis_budget_GB");
        // HBM budget reached: continue with two buffers and replay at the end
        store = false;
        d2d(A.p, basis[k].p, n);
        bA = sc[k];
        if (k > 0) { d2d(B.p, basis[k - 1].p, n); bB = sc[k - 1]; }
        basis.clear();
        sc.clear();
      }
    }
    if (store) {
      DevBuf<double> W(n);
      const int nb = H.apply(basis[k].p, W.p, 1.0 / sc[k], 0.0, true);
      alpha = H.dot_result(nb) / sc[k];
      std::vector<const double *> src;
      std::vector<double> cf;
      src.push_back(basis[k].p);
      cf.push_back(-alpha / sc[k]);
      if (k > 0) { src.push_back(basis[k - 1].p); cf.push_back(-R.beta[k - 1] / sc[k - 1]); }
      double w2 = comb(n, W.p, 1.0, src, cf, true);
      if (reorth) {
        // full FP64 reorthogonalisation against the stored basis (classical
        // Gram-Schmidt; a second pass when the first removed more than half
        // of the norm squared, the "twice is enough" rule)
        std::vector<const double *> all;
        for (int j = 0; j <= k; ++j) all.push_back(basis[j].p);
        for (int pass = 0; pass < 2; ++pass) {
          std::vector<double> c = dots(n, W.p, all);
          std::vector<double> cfr(all.size());
          for (size_t j = 0; j < all.size(); ++j) cfr[j] = -c[j] / (sc[j] * sc[j]);
          const double w2n = comb(n, W.p, 1.0, all, cfr, true);
          const bool again = w2n < 0.5 * w2;
          w2 = w2n;
          if (!again) break;
        }
      }
      if (proj) {
        // back onto the representation: rounding leaks weight into the others,
        // which the recurrence would otherwise amplify
        DevBuf<double> Tp(n);
        proj->apply(W.p, Tp.p);
        W = std::move(Tp);
        w2 = ddot(n, W.p, W.p);
      }
      beta = std::sqrt(w2);
      basis.push_back(std::move(W));
      sc.push_back(beta);
    } else {
      const double cb = (k > 0) ? -R.beta[k - 1] / bB : 0.0;
      const int nb = H.apply(A.p, B.p, 1.0 / bA, cb, true);
      alpha = H.dot_result(nb) / bA;
      beta = std::sqrt(comb(n, B.p, 1.0, {A.p}, {-alpha / bA}, true));
      std::swap(A.p, B.p);
      bB = bA;
      bA = beta;
    }
    R.alpha.push_back(alpha);
    R.beta.push_back(beta);
    double e0, zl;
    tridiag_lowest((int)R.alpha.size(), R.alpha.data(), R.beta.data(), e0, zl,
                   R.alpha.size() > 1 ? R.E0 : std::numeric_limits<double>::quiet_NaN());
    R.E0 = e0;
    ritz = std::fabs(zl * beta);
    if ((int)R.alpha.size() > max_iter)
      fail(ED_ERR_NOT_CONVERGED, "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations");
  }
  const int kk = (int)R.alpha.size();
  R.niter = kk;
  std::vector<double> z;
  double e0;
  tridiag_eigvec0(kk, R.alpha.data(), R.beta.data(), e0, z);
  if (store) {
    std::vector<const double *> src;
    std::vector<double> cf;
    for (int j = 0; j < kk; ++j) {
      src.push_back(basis[j].p);
      cf.push_back(z[j] / sc[j]);
    }
    comb(n, d_gs, 0.0, src, cf, false);
    if (proj) {
      DevBuf<double> Tp(n);
      proj->apply(d_gs, Tp.p);
      d2d(d_gs, Tp.p, n);
    }
  } else {
    // LanczosEigenvector replay (Lanczos.hpp:109-143)
    d2d(A.p, d_start, n);
    bA = 1.0;
    bB = 1.0;
    d2d(d_gs, d_start, n);
    scale(n, d_gs, z[0], stream());
    for (int j = 1; j < kk; ++j) {
      check_cancel();
      const double cb = (j > 1) ? -R.beta[j - 2] / bB : 0.0;
      H.apply(A.p, B.p, 1.0 / bA, cb, false);
      replay_step(n, B.p, A.p, d_gs, -R.alpha[j - 1] / bA, z[j] / R.beta[j - 1], stream());
      std::swap(A.p, B.p);
      bB = bA;
      bA = R.beta[j - 1];
    }
  }
  sync();
  return R;
}

// ---------------------------------------------------------------------------
// LanczosGreen: same recurrence, stops on the products of beta.
// ---------------------------------------------------------------------------
void lanczos_green(Hamiltonian &H, const double *d_psi, std::vector<double> &a, std::vector<double> &b,
                   long max_steps) {
  Params &P = params();
  const double acc = P.getd("accur_continued_fraction");
  const size_t max_cf = max_steps < 0 ? (size_t)P.geti("max_iter_CF") : (size_t)max_steps;
  const size_t n = H.dim;
  // with lanczos_reorth, every new vector is reorthogonalised (classical
  // Gram-Schmidt, twice) against the normalized Krylov vectors kept in HBM:
  // the fraction is then that of the exact Krylov process, down to the
  // exhaustion of the space (b_n -> 0 when accur_continued_fraction allows it)
  const bool reorth = P.getb("lanczos_reorth");
  a.clear();
  b.assign(1, 1.0);
  if (!reorth && max_cf > 0 && device_lanczos(n)) {
    // device-driven: the stop rule applied to the coefficients in step order
    LanczosDeviceRun run(H, d_psi, (int)std::min<size_t>(max_cf, 1u << 20), false, 0.0);
    std::vector<double> na, nbv, logp;
    double prod = 1.0;
    int nblk = std::max(1, 12 / run.block());
    bool stop = false;
    while (!stop) {
      check_cancel();
      const size_t k0 = na.size();
      if (!run.run(nblk, na, nbv)) break;
      for (size_t k = k0; k < na.size(); ++k) {
        if (!(prod > acc && std::fabs(b.back()) > acc && a.size() < max_cf)) {
          stop = true;
          break;
        }
        a.push_back(na[k]);
        b.push_back(nbv[k]);
        prod *= std::fabs(nbv[k]);
        logp.push_back(std::log(std::max(std::min(prod, std::fabs(nbv[k])), 1e-300)));
      }
      if (!(prod > acc && std::fabs(b.back()) > acc && a.size() < max_cf)) stop = true;
      nblk = predict_blocks(logp, std::log(acc), run.block());
    }
    return;
  }
  DevBuf<double> A(n), B(n);
  d2d(A.p, d_psi, n);
  std::vector<DevBuf<double>> basis;
  if (reorth) {
    basis.emplace_back(n);
    d2d(basis[0].p, d_psi, n);
  }
  double bA = 1.0, bB = 1.0, prod = 1.0;
  while (prod > acc && std::fabs(b.back()) > acc && a.size() < max_cf) {
    check_cancel();
    const size_t k = a.size();
    const double cb = k ? -b.back() / bB : 0.0;
    const int nb = H.apply(A.p, B.p, 1.0 / bA, cb, true);
    const double alpha = H.dot_result(nb) / bA;
    double w2 = comb(n, B.p, 1.0, {A.p}, {-alpha / bA}, true);
    if (reorth) {
      std::vector<const double *> all;
      for (const auto &v : basis) all.push_back(v.p);
      for (int pass = 0; pass < 2; ++pass) {
        std::vector<double> c = dots(n, B.p, all);
        for (auto &x : c) x = -x;
        w2 = comb(n, B.p, 1.0, all, c, true);
      }
    }
    const double beta = std::sqrt(w2);
    a.push_back(alpha);
    b.push_back(beta);
    prod *= std::fabs(beta);
    std::swap(A.p, B.p);
    bB = bA;
    bA = beta;
    if (reorth && beta > 0.0) {
      basis.emplace_back(n);
      d2d(basis.back().p, A.p, n);
      scale(n, basis.back().p, 1.0 / beta, stream());
    }
  }
}

// ---------------------------------------------------------------------------
// band Lanczos (Freund), restated from src_ed/Hamiltonian/Lanczos.hpp:214-395
// ---------------------------------------------------------------------------
BandLanczosRun band_lanczos_run(Hamiltonian &H, const std::vector<const double *> &phi, int M_in) {
  Params &Pm = params();
  const size_t dim = H.dim;
  const int p = (int)phi.size();
  const double accur_deflation = Pm.getd("accur_deflation");
  const double accur_bl = std::max(0.0, Pm.getd("accur_band_lanczos"));
  const int max_iter_bl = (int)Pm.geti("max_iter_BL");
  const double min_gap = Pm.getd("band_lanczos_minimum_gap");
  const bool no_deg = Pm.getb("no_degenerate_BL");
  if (dim == 0 || p == 0) return BandLanczosRun{};
  int M0 = (M_in < 0) ? (int)(14 * p * std::log(1.0 * dim)) : M_in;
  // 14 p log(dim) is 0 for a one-state sector: room for the p start vectors
  // and the step that deflates them
  if (M_in < 0) M0 = std::max(M0, p + 1);
  int pc = p;
  if (M0 > max_iter_bl) M0 = max_iter_bl;
  const int M0p = (M0 / p + 1) * p;
  M0 = M0p - pc;
  std::vector<std::unique_ptr<DevBuf<double>>> v((size_t)M0p + 1);
  std::vector<int> Idef((size_t)M0p + 1, 0), def((size_t)M0 + 1, 0);
  std::vector<double> t((size_t)M0 * M0, 0.0), Pmat((size_t)p * M0, 0.0);
  auto T = [&](int i, int j) -> double & { return t[(size_t)i + (size_t)j * M0]; };
  for (int i = 0; i < p; ++i) {
    v[i].reset(new DevBuf<double>(dim));
    d2d(v[i]->p, phi[i], dim);
  }
  int nd = 0, j;
  double ev_old = 1e12, ev_change = 0.0;
  // dots of the current v_j computed by the previous step's fused update
  // (band_step): partials slot, block count, rows, and the layout of the rows
  double *pre = nullptr;
  int pre_nb = 0, pre_m = 0, pre_na = 0, pre_nw = 0, pre_skip = 0;
  // H v_j queued ahead by the previous step (pre_h), with its partials count
  bool pre_h = false;
  int pre_hnb = 0;
  // a queued product or fused dots are drained before the vectors go away
  auto drain = [&]() {
    if (pre_h) H.dot_result(pre_hnb);
    if (pre) sum_partials(pre, pre_nb, pre_m, nullptr);
    pre_h = false;
    pre = nullptr;
  };
  // step (12): convergence of the lowest eigenvalue of the projected matrix,
  // checked after step j (every p steps)
  auto converged_at = [&](int j) -> bool {
    if (!(j >= 4 * p && j % p == 0)) return false;
    double e0, e1 = 0.0;
    if (nd == 0 && !no_deg) {
      // the leading j x j block of t in place; the previous check's value
      // bounds this one from above (interlacing) and the last change sizes
      // the bracket below it
      e0 = band_lowest_eigenvalue(j, p, t.data(), M0, ev_old < 1e11 ? ev_old : HUGE_VAL, 2.0 * ev_change);
    } else if (nd == 0) {
      std::vector<double> Hp((size_t)j * j);
      for (int c = 0; c < j; ++c)
        for (int r = 0; r < j; ++r) Hp[(size_t)r + (size_t)c * j] = T(r, c);
      e0 = band_kth_eigenvalue(j, p, Hp.data(), 0);
      e1 = band_kth_eigenvalue(j, p, Hp.data(), 1);
    } else {
      std::vector<double> Hp((size_t)j * j);
      for (int c = 0; c < j; ++c)
        for (int r = 0; r < j; ++r) Hp[(size_t)r + (size_t)c * j] = T(r, c);
      std::vector<double> ev(j);
      symeig(j, Hp.data(), ev.data(), nullptr);
      e0 = ev[0];
      e1 = j > 1 ? ev[1] : ev[0];
    }
    const double test = std::fabs(e0 - ev_old);
    if (ev_old < 1e11) ev_change = test;
    ev_old = e0;
    if (no_deg && e1 - e0 < min_gap) {
      drain();
      fail(ED_ERR_DEGENERATE, "band Lanczos: the gap between the first two eigenvalues is smaller than " +
                                  std::to_string(min_gap));
    }
    return test < accur_bl;
  };
  // Device-driven steps (no deflated vectors, windows of at most
  // kBandStepMax vectors, no P-matrix column): each fused step derives its own
  // scalars on the device from the dots the previous one left and the
  // product's partials (band_step with a BandFold), so blocks of
  // `band_lanczos_block` steps are queued without a host wait.  The host reads
  // a block's records (pinned, one event per block) while the next block runs,
  // fills t and applies the convergence check in step order.  A deflation
  // found on the device makes the remaining launches no-ops and the host
  // redoes that step; a convergence ends the run at the same step as the
  // host-driven loop (a block queued past it is discarded).
  const int dev_block = (int)Pm.geti("band_lanczos_block");
  const bool dev_on = dev_block >= 2 && Pm.geti("vector_block") == 1024;
  auto dev_ok = [&](int jj) {
    return dev_on && nd == 0 && jj >= p && pc - 1 <= kBandStepMax && pc <= kBandStepMax && jj + 1 < M0;
  };
  DevBuf<double> dring, hring, rring;
  DevBuf<int> dstop;
  double *rhost = nullptr;  // pinned: two blocks of records
  hipEvent_t rev[2] = {nullptr, nullptr};
  const size_t drow = (size_t)(2 * kBandStepMax + 2) * (size_t)partials_cap();
  const size_t hcap = (size_t)partials_cap();
  bool dev_converged = false;
  struct DevCleanup {
    double *&h;
    hipEvent_t *ev;
    ~DevCleanup() {
      if (h) (void)hipHostFree(h);
      for (int q = 0; q < 2; ++q)
        if (ev[q]) (void)hipEventDestroy(ev[q]);
    }
  } dev_cleanup{rhost, rev};
  // runs device blocks from step j0 (dots of v_j0 in the pinned slot `pre`);
  // returns the step the host loop continues with (its dots not available)
  auto run_device = [&](int j0) -> int {
    if (!dring.p) {
      dring.alloc(2 * drow);
      hring.alloc(2 * hcap);
      rring.alloc((size_t)dev_block * 8);
      dstop.alloc(1);
      ED_HIP(hipHostMalloc((void **)&rhost, 2 * (size_t)dev_block * 8 * sizeof(double), hipHostMallocDefault));
      for (int q = 0; q < 2; ++q) ED_HIP(hipEventCreateWithFlags(&rev[q], hipEventDisableTiming));
    }
    // the incoming dots: the pinned slot the host-driven step left
    const double *din = pre;
    int din_nb = pre_nb, din_rb = 1 + pre_na + pre_skip;
    double *pinned = pre;
    const int pinned_nb = pre_nb, pinned_m = pre_m;
    pre = nullptr;
    ED_HIP(hipMemsetAsync(dstop.p, 0, sizeof(int), stream()));
    int nstep = 0;  // steps queued so far (dots / partials slots alternate)
    auto launch_block = [&](int jb, int K, int parity) {
      for (int s2 = 0; s2 < K; ++s2, ++nstep) {
        const int js = jb + s2, jm2 = js + pc, k02 = (js > pc) ? js - pc : 0;
        const int na = jm2 - js - 1, nw = js - k02;
        v[js + pc].reset(new DevBuf<double>(dim));
        double *hp = hring.p + (size_t)(nstep & 1) * hcap;
        const int hnb2 = H.apply_fixed(v[js]->p, v[js + pc]->p, 1.0, 0.0, hp, nullptr, 0);
        std::vector<double *> ys;
        std::vector<const double *> ws;
        for (int k = js + 1; k < jm2; ++k) ys.push_back(v[k]->p);
        for (int k = k02; k < js; ++k) ws.push_back(v[k]->p);
        BandFold F;
        F.dpart = din;
        F.dnb = din_nb;
        F.rb = din_rb;
        F.hpart = hp;
        F.hnb = hnb2;
        F.j = js;
        F.pc = pc;
        F.k0 = k02;
        F.accur = accur_deflation;
        F.rec = rring.p + (size_t)s2 * 8;
        F.stop = dstop.p;
        double *dout = dring.p + (size_t)(nstep & 1) * drow;
        const int nb = band_step(dim, v[js]->p, 0.0, na, ys.data(), nullptr, v[js + pc]->p, 0.0, nw, ws.data(),
                                 nullptr, 0.0, dout, stream(), &F);
        if (nb <= 0) {
          sync();
          fail(ED_ERR_INTERNAL, "band Lanczos: device step outside the fused kernel's range");
        }
        const int k0n = (js + 1 > pc) ? js + 1 - pc : 0;
        din = dout;
        din_nb = nb;
        din_rb = 1 + na + (k0n - k02);
      }
      ED_HIP(hipMemcpyAsync(rhost + (size_t)parity * dev_block * 8, rring.p, (size_t)K * 8 * sizeof(double),
                            hipMemcpyDeviceToHost, stream()));
      ED_HIP(hipEventRecord(rev[parity], stream()));
    };
    int jb = j0, K = std::min(dev_block, M0 - jb), parity = 0;
    launch_block(jb, K, parity);
    while (true) {
      const int jn = jb + K;
      const bool next = dev_ok(jn);
      const int Kn = next ? std::min(dev_block, M0 - jn) : 0;
      if (next) launch_block(jn, Kn, parity ^ 1);
      ED_HIP(hipEventSynchronize(rev[parity]));
      if (pinned) {
        sum_partials(pinned, pinned_nb, pinned_m, nullptr);  // release the slot the first step read
        pinned = nullptr;
      }
      const double *recs = rhost + (size_t)parity * dev_block * 8;
      for (int s2 = 0; s2 < K; ++s2) {
        const int js = jb + s2, jm2 = js + pc, k02 = (js > pc) ? js - pc : 0;
        const double *r = recs + (size_t)s2 * 8;
        if (r[6] != 0.0) {  // deflation: the host redoes step js
          sync();
          return js;
        }
        if (js >= pc) T(js, js - pc) = r[0];
        for (int k = js + 1; k < jm2; ++k)
          if (k >= pc) T(js, k - pc) = r[1 + (k - js - 1)];
        for (int k = k02; k < js; ++k) T(k, js) = T(js, k);
        T(js, js) = r[5];
        if (js > pc && def[k02 - 1] == 0) v[k02 - 1].reset();
        bool conv;
        try {
          conv = converged_at(js);
        } catch (...) {
          sync();
          throw;
        }
        if (conv) {
          sync();
          dev_converged = true;
          return js;
        }
      }
      if (!next) return jn;
      jb = jn;
      K = Kn;
      parity ^= 1;
      try {
        check_cancel();
      } catch (...) {
        sync();
        throw;
      }
    }
  };
  for (j = 0; j < M0; ++j) {
    try {
      check_cancel();
    } catch (...) {
      drain();
      throw;
    }
    if (pre && !pre_h && dev_ok(j)) {
      j = run_device(j);
      if (dev_converged) break;
      if (j >= M0) break;
    }
    // one multi-dot per step: ||v_j||^2, <v_j, phi_k> (P matrix), the step (6)
    // overlaps <v_j, v_k> (k = j+1 .. j+pc-1) and <v_j, v_k> for the vectors
    // step (8) subtracts (k0 <= k < j), all against the unnormalized v_j
    const int jm = j + pc;
    const int k0 = (j > pc) ? j - pc : 0;
    // the P matrix enters Q only through its first p columns (build_Q_matrix,
    // Hamiltonian_base.hpp:244, the product over phi.size() columns), so
    // <v_j, phi_k> is computed for j < p only
    const int np = (j < p) ? p : 0;
    std::vector<double> dv;
    int hnb;
    if (pre) {
      // the dots came with the previous step's update (and H v_j was queued
      // there too, unless that step ended the range): wait once for both
      if (pre_h) {
        hnb = pre_hnb;
        pre_h = false;
      } else {
        v[j + pc].reset(new DevBuf<double>(dim));
        hnb = H.apply(v[j]->p, v[j + pc]->p, 1.0, 0.0, true);
      }
      std::vector<double> raw;
      sum_partials(pre, pre_nb, pre_m, &raw);
      pre = nullptr;
      // raw: <c,c>, the window ahead (pre_na), the vectors behind (pre_nw), <c, v_{j-1}>
      dv.push_back(raw[0]);
      for (int r = 1; r <= pre_na; ++r) dv.push_back(raw[r]);
      for (int r = 1 + pre_na + pre_skip; r <= pre_na + pre_nw + 1; ++r) dv.push_back(raw[r]);
      if ((int)dv.size() != 1 + np + (jm - j - 1) + (j - k0)) {
        H.dot_result(hnb);
        fail(ED_ERR_INTERNAL, "band Lanczos: fused dots do not match the step's window");
      }
    } else {
      std::vector<const double *> db{v[j]->p};
      for (int k = 0; k < np; ++k) db.push_back(phi[k]);
      for (int k = j + 1; k < jm; ++k) db.push_back(v[k]->p);
      for (int k = k0; k < j; ++k) db.push_back(v[k]->p);
      // multi-dot and H v_j (on the unnormalized v_j) are queued together; the
      // host then waits once for both results
      std::vector<const double *> db1(db.begin(), db.begin() + std::min<size_t>(db.size(), 16));
      double *pdot = partials_acquire();
      const int dnb = multi_dot(dim, v[j]->p, (int)db1.size(), db1.data(), pdot, stream());
      v[j + pc].reset(new DevBuf<double>(dim));
      hnb = H.apply(v[j]->p, v[j + pc]->p, 1.0, 0.0, true);
      sum_partials(pdot, dnb, (int)db1.size(), &dv);
      if (db.size() > db1.size()) {  // more than 16 vectors (p > 5): remaining dots
        std::vector<const double *> rest(db.begin() + 16, db.end());
        const std::vector<double> dr = dots(dim, v[j]->p, rest);
        dv.insert(dv.end(), dr.begin(), dr.end());
      }
    }
    const double hd_raw = H.dot_result(hnb);  // <H v_j, v_j> before normalization
    const double z = std::sqrt(std::max(dv[0], 0.0));
    if (verbose() && dim < 4096) vlog("band Lanczos step j %d pc %d: |v_j| %.3e, <Hv,v> %.6e", j, pc, z, hd_raw);
    if (z < accur_deflation) {
      v[j + pc].reset();
      if (j >= pc) {
        Idef[nd] = j - pc;
        def[j - pc] = 1;
        nd++;
      }
      pc--;
      if (pc == 0) break;
      v[j].reset();
      for (int k = j; k < j + pc; ++k) v[k] = std::move(v[k + 1]);
      v[j + pc].reset();
      j--;
      continue;
    }
    const double iz = 1.0 / z;
    if (j >= pc) T(j, j - pc) = z;
    for (int k = 0; k < np; ++k) Pmat[(size_t)k + (size_t)j * p] = dv[1 + k] * iz;
    std::vector<double *> ys;
    std::vector<double> cs;
    for (int k = j + 1; k < jm; ++k) {
      const double zz = dv[1 + np + (k - j - 1)] * iz;
      ys.push_back(v[k]->p);
      cs.push_back(-zz);
      if (k >= pc) T(j, k - pc) = zz;
    }
    // step (7): v[j+pc] = H v_j for the normalized v_j: the product above
    // times 1/z, applied in the combination below
    const double hd = hd_raw * iz * iz;
    // step (8): v[j+pc] -= sum_{k0 <= k < j} t(k, j) v_k
    std::vector<const double *> src;
    std::vector<double> cf;
    double wdot = hd;
    for (int k = k0; k < j; ++k) {
      T(k, j) = T(j, k);
      src.push_back(v[k]->p);
      cf.push_back(-T(k, j));
      wdot -= T(k, j) * dv[1 + np + (jm - j - 1) + (k - k0)] * iz;
    }
    // steps (6) - (8) in one pass, with the next step's dots (no deflated
    // vectors, windows of at most kBandStepMax vectors); the next step uses
    // the dots when it needs no P-matrix column
    if (nd == 0 && (int)ys.size() <= kBandStepMax && (int)src.size() <= kBandStepMax) {
      double *slot = partials_acquire();
      const int nb = band_step(dim, v[j]->p, iz, (int)ys.size(), ys.data(), cs.data(), v[j + pc]->p, iz,
                               (int)src.size(), src.data(), cf.data(), -wdot, slot, stream());
      if (nb > 0) {
        const int na = (int)ys.size(), nw = (int)src.size();
        const int m = (na > 0 ? 1 + na : 1) + nw + 1;
        if (j + 1 >= p) {
          pre = slot;
          pre_nb = nb;
          pre_m = m;
          pre_na = na;
          pre_nw = nw;
          pre_skip = ((j + 1 > pc) ? j + 1 - pc : 0) - k0;
          // queue the next step's product now, so that the host's
          // convergence check below overlaps it (wasted if this step ends
          // the run; drained then), unless the next steps run device-driven
          if (j + 1 < M0 && !dev_ok(j + 1)) {
            v[j + 1 + pc].reset(new DevBuf<double>(dim));
            pre_hnb = H.apply(v[j + 1]->p, v[j + 1 + pc]->p, 1.0, 0.0, true);
            pre_h = true;
          }
        } else {
          sum_partials(slot, nb, m, nullptr);  // release the slot
        }
        if (j > pc && def[k0 - 1] == 0) v[k0 - 1].reset();
        T(j, j) = wdot;
        goto converge;
      }
      sum_partials(slot, 0, 0, nullptr);
    }
    if (jm > j + 1) {
      // v_j /= z in the same pass (only the first chunk scales)
      for (size_t s0 = 0; s0 < ys.size(); s0 += 16) {
        const int m = (int)std::min<size_t>(16, ys.size() - s0);
        multi_axpy_src(dim, v[j]->p, m, ys.data() + s0, cs.data() + s0, nullptr, stream(), s0 == 0 ? iz : 1.0);
      }
    } else {
      scale(dim, v[j]->p, iz, stream());
    }
    if (j > pc && def[k0 - 1] == 0) v[k0 - 1].reset();
    if (nd == 0) {
      // t(j, j) = <v[j+pc], v_j> after step (8), from the dots above; the
      // scaling, the subtraction and the projection on v_j are one pass
      src.push_back(v[j]->p);
      cf.push_back(-wdot);
      comb(dim, v[j + pc]->p, iz, src, cf, false);
    } else {
      comb(dim, v[j + pc]->p, iz, src, cf, false, v[j]->p, &wdot);
      // step (9): deflated vectors (classical Gram-Schmidt over the mutually orthogonal set)
      std::vector<const double *> dvv;
      for (int k = 0; k < nd; ++k) dvv.push_back(v[Idef[k]]->p);
      std::vector<double> w = dots(dim, v[j + pc]->p, dvv);
      std::vector<double> c2(nd);
      for (int k = 0; k < nd; ++k) {
        c2[k] = -w[k];
        T(Idef[k], j) = w[k];
      }
      comb(dim, v[j + pc]->p, 1.0, dvv, c2, false, v[j]->p, &wdot);
      comb(dim, v[j + pc]->p, 1.0, {v[j]->p}, {-wdot}, false);
    }
    T(j, j) = wdot;
    for (int k = 0; k < nd; ++k) T(j, Idef[k]) = T(Idef[k], j);
  converge:
    if (converged_at(j)) break;
  }
  drain();  // the product and dots of a step not taken
  const int M = j;
  vlog("band Lanczos: %d steps, %d deflations, dim %zu", M, nd, dim);
  BandLanczosRun out;
  out.M = M;
  out.p = p;
  out.nd = nd;
  if (M == 0) return out;
  out.T.resize((size_t)M * M);
  for (int c = 0; c < M; ++c)
    for (int r = 0; r < M; ++r) out.T[(size_t)r + (size_t)c * M] = T(r, c);
  out.P.assign(Pmat.begin(), Pmat.begin() + (size_t)p * std::min(p, M));
  return out;
}

// build_Q_matrix (Hamiltonian_base.hpp:230-254): eigenvalues of the projected
// matrix and Q = P U restricted to the first p rows of U
void BandLanczosRun::finish(std::vector<double> &evals, std::vector<double> &Q) const {
  evals.clear();
  Q.clear();
  if (M == 0) return;
  const int kmax = std::min(p, M);
  Q.assign((size_t)p * M, 0.0);
  if (nd == 0) {
    std::vector<double> Rr;
    band_eig_first_rows(M, p, T.data(), kmax, evals, Rr);
    for (int c = 0; c < M; ++c)
      for (int i = 0; i < p; ++i) {
        double s = 0.0;
        for (int k = 0; k < kmax; ++k) s += P[(size_t)i + (size_t)k * p] * Rr[(size_t)k * M + c];
        Q[(size_t)i + (size_t)c * p] = s;
      }
  } else {
    evals.resize(M);
    std::vector<double> U((size_t)M * M);
    symeig(M, T.data(), evals.data(), U.data());
    for (int c = 0; c < M; ++c)
      for (int i = 0; i < p; ++i) {
        double s = 0.0;
        for (int k = 0; k < kmax; ++k) s += P[(size_t)i + (size_t)k * p] * U[(size_t)k + (size_t)c * M];
        Q[(size_t)i + (size_t)c * p] = s;
      }
  }
}

void band_lanczos(Hamiltonian &H, const std::vector<const double *> &phi, std::vector<double> &evals,
                  std::vector<double> &Q, int M_in) {
  band_lanczos_run(H, phi, M_in).finish(evals, Q);
}

// ---------------------------------------------------------------------------
// Davidson-Liu (src_ed/Hamiltonian/Davidson.hpp:38-230)
// ---------------------------------------------------------------------------
static int gram_schmidt(std::vector<std::unique_ptr<DevBuf<double>>> &b, size_t dim, double accur) {
  // src_util/vector_num.hpp:217-235 with k = 0
  for (size_t i = 0; i < b.size(); ++i) {
    const double z = std::sqrt(ddot(dim, b[i]->p, b[i]->p));
    if (z < accur) {
      std::swap(b[i], b.back());
      b.pop_back();
      --i;
      continue;
    }
    scale(dim, b[i]->p, 1.0 / z, stream());
    if (i + 1 < b.size()) {
      std::vector<const double *> rest;
      for (size_t j2 = i + 1; j2 < b.size(); ++j2) rest.push_back(b[j2]->p);
      std::vector<double> zz = dots(dim, b[i]->p, rest);
      std::vector<double *> ys;
      std::vector<double> cs;
      for (size_t j2 = i + 1; j2 < b.size(); ++j2) {
        ys.push_back(b[j2]->p);
        cs.push_back(-zz[j2 - i - 1]);
      }
      for (size_t s0 = 0; s0 < ys.size(); s0 += 16) {
        const int m = (int)std::min<size_t>(16, ys.size() - s0);
        multi_axpy_src(dim, b[i]->p, m, ys.data() + s0, cs.data() + s0, nullptr, stream());
      }
    }
  }
  return (int)b.size();
}

int davidson(Hamiltonian &H, int K, double accuracy, std::vector<double> &evals, std::vector<DevBuf<double>> &x) {
  const size_t dim = H.dim;
  const int max_iter = (int)params().geti("max_iter_lanczos");
  const double TINY = 1e-14;
  if ((size_t)K > dim) K = (int)dim;
  DevBuf<double> d(dim);
  H.diag(d.p);
  std::vector<double> dh = d.to_host();
  // find_lowest (src_util/vector_num.hpp:244-267)
  std::vector<size_t> low(K);
  std::vector<double> minv(K);
  for (int i = 0; i < K; ++i) minv[i] = 1.0e12 + 1.0 * i;
  for (size_t i = 0; i < dim; ++i) {
    int jj;
    for (jj = K - 1; jj >= 0; jj--)
      if (dh[i] > minv[jj]) break;
    jj++;
    if (jj < K) {
      for (int k = K - 1; k > jj; k--) { minv[k] = minv[k - 1]; low[k] = low[k - 1]; }
      minv[jj] = dh[i];
      low[jj] = i;
    }
  }
  std::vector<std::unique_ptr<DevBuf<double>>> b, h;
  x.clear();
  for (int k = 0; k < K; ++k) {
    b.emplace_back(new DevBuf<double>(dim));
    h.emplace_back(new DevBuf<double>(dim));
    x.emplace_back(dim);
    fill(dim, b[k]->p, 0.0, stream());
    set_unit(b[k]->p, low[k], stream());
    H.apply(b[k]->p, h[k]->p, 1.0, 0.0, false);
  }
  int nb = K;
  std::vector<double> bhb((size_t)nb * nb);
  {
    // bhb(i, j) = h_j[low_i]
    for (int jj = 0; jj < nb; ++jj) {
      std::vector<double> hv = h[jj]->to_host();
      for (int i = 0; i < nb; ++i) bhb[(size_t)i + (size_t)jj * nb] = hv[low[i]];
    }
  }
  evals.assign(K, 0.0);
  int iter;
  for (iter = 0; iter < max_iter; ++iter) {
    check_cancel();
    std::vector<double> mu(nb), a((size_t)nb * nb);
    symeig(nb, bhb.data(), mu.data(), a.data());
    std::vector<const double *> bp, hp;
    for (int q = 0; q < nb; ++q) { bp.push_back(b[q]->p); hp.push_back(h[q]->p); }
    for (int k = 0; k < K; ++k) {
      std::vector<double> cf(nb);
      for (int q = 0; q < nb; ++q) cf[q] = a[(size_t)q + (size_t)k * nb];
      comb(dim, x[k].p, 0.0, bp, cf, false);
    }
    std::vector<std::unique_ptr<DevBuf<double>>> r;
    std::vector<int> rk;
    for (int k = 0; k < K; ++k) {
      std::unique_ptr<DevBuf<double>> rt(new DevBuf<double>(dim));
      std::vector<const double *> src = hp;
      std::vector<double> cf(nb);
      for (int q = 0; q < nb; ++q) cf[q] = a[(size_t)q + (size_t)k * nb];
      src.push_back(x[k].p);
      cf.push_back(-mu[k]);
      const double res = std::sqrt(comb(dim, rt->p, 0.0, src, cf, true));
      if (res > accuracy) { r.push_back(std::move(rt)); rk.push_back(k); }
    }
    const int Qn = (int)r.size();
    if (Qn == 0) {
      for (int k = 0; k < K; ++k) evals[k] = mu[k];
      break;
    }
    while ((int)b.size() < K + 2 * Qn) b.emplace_back(new DevBuf<double>(dim));
    for (int c = 0; c < Qn; ++c) std::swap(b[rk[c]], b[K + c]);
    for (int k = 0; k < K; ++k) {
      // b[k] <- x[k] (the Ritz vectors); x[k] keeps a buffer for the next round
      std::swap(b[k]->p, x[k].p);
      std::swap(b[k]->n, x[k].n);
    }
    for (int c = 0; c < Qn; ++c) {
      precond(dim, r[c]->p, d.p, mu[rk[c]], TINY, stream());
      b[K + Qn + c] = std::move(r[c]);
    }
    b.resize(K + 2 * Qn);
    nb = gram_schmidt(b, dim, TINY);
    if (nb == K) {
      for (int k = 0; k < K; ++k) evals[k] = mu[k];
      break;
    }
    h.clear();
    for (int k = 0; k < nb; ++k) {
      h.emplace_back(new DevBuf<double>(dim));
      H.apply(b[k]->p, h[k]->p, 1.0, 0.0, false);
    }
    bhb.assign((size_t)nb * nb, 0.0);
    for (int i = 0; i < nb; ++i) {
      std::vector<const double *> hh;
      for (int k = 0; k <= i; ++k) hh.push_back(h[k]->p);
      std::vector<double> zz = dots(dim, b[i]->p, hh);
      for (int k = 0; k <= i; ++k) {
        bhb[(size_t)i + (size_t)k * nb] = zz[k];
        bhb[(size_t)k + (size_t)i * nb] = zz[k];
      }
    }
    for (int k = 0; k < K; ++k) evals[k] = mu[k];
  }
  if (iter == max_iter) fail(ED_ERR_NOT_CONVERGED, "Davidson method failed to converge!");
  sync();
  return iter;
}

// ---------------------------------------------------------------------------
// block Lanczos, QR variant (src_ed/Hamiltonian/Lanczos.hpp:426-572)
// ---------------------------------------------------------------------------
int block_lanczos(Hamiltonian &H, const std::vector<const double *> &q0, int M0, std::vector<double> &A,
                  std::vector<double> &B, bool qr) {
  Params &Pm = params();
  const size_t dim = H.dim;
  const int p = (int)q0.size();
  const double accur_deflation = Pm.getd("accur_deflation");
  const double accur_bl = std::max(0.0, Pm.getd("accur_band_lanczos"));
  const int max_iter_bl = (int)Pm.geti("max_iter_BL");
  if (M0 > max_iter_bl) M0 = max_iter_bl;
  A.clear();
  B.clear();
  const int pp = p * p;
  std::vector<DevBuf<double>> Qp, Qc, Z;
  for (int k = 0; k < p; ++k) {
    Qp.emplace_back(dim);
    Qc.emplace_back(dim);
    Z.emplace_back(dim);
    d2d(Qc[k].p, q0[k], dim);
  }
  // orthonormalize the starting block (MGS)
  for (int l = 0; l < p; ++l) {
    for (int k = 0; k < l; ++k) {
      const double z = ddot(dim, Qc[k].p, Qc[l].p);
      comb(dim, Qc[l].p, 1.0, {Qc[k].p}, {-z}, false);
    }
    const double nrm = std::sqrt(ddot(dim, Qc[l].p, Qc[l].p));
    if (nrm < accur_deflation)
      fail(ED_ERR_INVALID, "blockLanczos: starting block is rank-deficient at column " + std::to_string(l));
    scale(dim, Qc[l].p, 1.0 / nrm, stream());
  }
  double ev_old = 1e12;
  int j;
  for (j = 0; j < M0; ++j) {
    check_cancel();
    for (int k = 0; k < p; ++k) H.apply(Qc[k].p, Z[k].p, 1.0, 0.0, false);
    std::vector<double> Aj(pp);
    for (int l = 0; l < p; ++l) {
      std::vector<const double *> qs;
      for (int k = 0; k <= l; ++k) qs.push_back(Qc[k].p);
      std::vector<double> zz = dots(dim, Z[l].p, qs);
      for (int k = 0; k <= l; ++k) {
        Aj[k + l * p] = zz[k];
        Aj[l + k * p] = zz[k];
      }
    }
    A.insert(A.end(), Aj.begin(), Aj.end());
    for (int l = 0; l < p; ++l) {
      std::vector<const double *> src;
      std::vector<double> cf;
      for (int k = 0; k < p; ++k) { src.push_back(Qc[k].p); cf.push_back(-Aj[k + l * p]); }
      if (j > 0) {
        const double *Bp = B.data() + (size_t)(j - 1) * pp;
        for (int k = 0; k < p; ++k) { src.push_back(Qp[k].p); cf.push_back(-Bp[l + k * p]); }
      }
      comb(dim, Z[l].p, 1.0, src, cf, false);
    }
    std::vector<double> Bj(pp, 0.0);
    bool singular = false;
    if (qr) {
      // Z = Q_{j+1} B_j by modified Gram-Schmidt (B_j upper triangular)
      for (int l = 0; l < p; ++l) {
        for (int k = 0; k < l; ++k) {
          const double z = ddot(dim, Z[k].p, Z[l].p);
          Bj[k + l * p] = z;
          comb(dim, Z[l].p, 1.0, {Z[k].p}, {-z}, false);
        }
        const double nrm = std::sqrt(ddot(dim, Z[l].p, Z[l].p));
        if (nrm < accur_deflation) {
          Bj[l + l * p] = 0.0;
          for (int ll = l; ll < p; ++ll) fill(dim, Z[ll].p, 0.0, stream());
          singular = true;
          break;
        }
        Bj[l + l * p] = nrm;
        scale(dim, Z[l].p, 1.0 / nrm, stream());
      }
      B.insert(B.end(), Bj.begin(), Bj.end());
    } else {
      // polar form (blockLanczosSVD, Lanczos.hpp:603-779): Gram matrix
      // G = Z^T Z = V D V^T, B_j = V D^(1/2) V^T (symmetric positive
      // semi-definite), Q_{j+1} = Z V D^(-1/2) V^T; deflation when a singular
      // value D^(1/2) falls below accur_deflation
      std::vector<double> G(pp), D(p), V(pp);
      for (int l = 0; l < p; ++l) {
        std::vector<const double *> zs;
        for (int k = 0; k <= l; ++k) zs.push_back(Z[k].p);
        std::vector<double> gg = dots(dim, Z[l].p, zs);
        for (int k = 0; k <= l; ++k) G[k + l * p] = G[l + k * p] = gg[k];
      }
      symeig(p, G.data(), D.data(), V.data());
      for (int m = 0; m < p; ++m)
        if (D[m] < accur_deflation * accur_deflation) singular = true;
      for (int r = 0; r < p; ++r)
        for (int c = 0; c < p; ++c) {
          double t = 0.0;
          for (int m = 0; m < p; ++m) t += V[r + m * p] * std::sqrt(std::max(D[m], 0.0)) * V[c + m * p];
          Bj[r + c * p] = t;
        }
      B.insert(B.end(), Bj.begin(), Bj.end());
      if (singular) { ++j; break; }
      // Q_{j+1}[l] = sum_k Binv(k, l) Z[k] into the buffers of Q_{j-1}
      // (no longer needed), which then become Q_cur
      for (int l = 0; l < p; ++l) {
        std::vector<const double *> src;
        std::vector<double> cf;
        for (int k = 0; k < p; ++k) {
          double t = 0.0;
          for (int m = 0; m < p; ++m) t += V[k + m * p] * V[l + m * p] / std::sqrt(D[m]);
          src.push_back(Z[k].p);
          cf.push_back(t);
        }
        comb(dim, Qp[l].p, 0.0, src, cf, false);
      }
      for (int k = 0; k < p; ++k) std::swap(Qp[k].p, Qc[k].p);
    }
    bool conv = false;
    if ((j + 1) >= 2 * p && (j + 1) % p == 0) {
      const int sz = (j + 1) * p;
      std::vector<double> Tm((size_t)sz * sz, 0.0);
      for (int jj = 0; jj <= j; ++jj)
        for (int r = 0; r < p; ++r)
          for (int c = 0; c < p; ++c) Tm[(jj * p + r) + (size_t)(jj * p + c) * sz] = A[(size_t)jj * pp + r + c * p];
      for (int jj = 0; jj < j; ++jj)
        for (int r = 0; r < p; ++r)
          for (int c = 0; c < p; ++c) {
            const double bv = B[(size_t)jj * pp + r + c * p];
            Tm[((jj + 1) * p + r) + (size_t)(jj * p + c) * sz] = bv;
            Tm[(jj * p + c) + (size_t)((jj + 1) * p + r) * sz] = bv;
          }
      const double e0 = band_kth_eigenvalue(sz, 2 * p, Tm.data(), 0);
      const double delta = std::fabs(e0 - ev_old);
      ev_old = e0;
      if (delta < accur_bl) conv = true;
    }
    if (conv || singular) { ++j; break; }
    if (qr)
      for (int k = 0; k < p; ++k) {
        std::swap(Qp[k].p, Qc[k].p);
        std::swap(Qc[k].p, Z[k].p);
      }
  }
  sync();
  return (int)(A.size() / pp);
}

}  // namespace ed
