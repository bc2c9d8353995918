// ed_capi.cpp -- the extern "C" boundary.  Every entry point converts library
// exceptions into status codes; no exception crosses the C-ABI.
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <string>

#include "../../include/pyqcm_ed.h"
#include "ed_common.hpp"
#include "ed_batch.hpp"
#include "ed_instance.hpp"
#include "ed_surrogate.hpp"
#include "ed_lattice.hpp"
#include "ed_model.hpp"
#include "ed_solvers.hpp"

using namespace ed;

namespace {
thread_local std::string g_last_error;
// Concurrency.  Calls that define or remove models, instances, global
// parameters, handles or device memory take the library lock exclusively;
// calls on one instance label take it shared plus that label's own lock, so
// solves and evaluations of different labels (the reference's OpenMP loop
// over clusters, src_qcm/lattice_model_instance.cpp:107-113, :160) run
// concurrently, each queued on its host thread's stream.  Every call returns
// with its device work complete, so the next call on a label may come from
// any thread.  ed_request_cancel and ed_last_error take no lock.
std::shared_mutex g_api_mu;
std::mutex g_label_mu;
std::map<int, std::shared_ptr<std::mutex>> g_label_locks;

std::shared_ptr<std::mutex> label_lock(int label) {
  std::lock_guard<std::mutex> l(g_label_mu);
  auto &m = g_label_locks[label];
  if (!m) m = std::make_shared<std::mutex>();
  return m;
}

template <typename F>
int run(F &&f) {
  int rc = ED_OK;
  try {
    f();
  } catch (const ed::Error &e) {
    g_last_error = e.what();
    rc = e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    rc = ED_ERR_OOM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    rc = ED_ERR_INTERNAL;
  } catch (...) {
    g_last_error = "unknown error";
    rc = ED_ERR_INTERNAL;
  }
  if (stream_started() && hipStreamSynchronize(stream()) != hipSuccess && rc == ED_OK) {
    g_last_error = std::string("HIP error: ") + hipGetErrorString(hipGetLastError());
    rc = ED_ERR_HIP;
  }
  return rc;
}

template <typename F>
int guard(F &&f) {
  std::unique_lock<std::shared_mutex> lock(g_api_mu);
  return run(f);
}

template <typename F>
int guard_label(int label, F &&f) {
  std::shared_lock<std::shared_mutex> lock(g_api_mu);
  const std::shared_ptr<std::mutex> m = label_lock(label);
  std::lock_guard<std::mutex> ll(*m);
  return run(f);
}

std::mutex g_h_mu;
std::map<int64_t, std::unique_ptr<Hamiltonian>> g_hams;
std::atomic<int64_t> g_next_h{1};

std::map<int64_t, std::unique_ptr<Surrogate>> g_surr;

Surrogate &get_surr(int64_t h) {
  std::lock_guard<std::mutex> l(g_h_mu);
  auto it = g_surr.find(h);
  if (it == g_surr.end()) fail(ED_ERR_NOT_FOUND, "unknown surrogate handle " + std::to_string(h));
  return *it->second;
}

Hamiltonian &get_h(int64_t h) {
  std::lock_guard<std::mutex> l(g_h_mu);
  auto it = g_hams.find(h);
  if (it == g_hams.end()) fail(ED_ERR_NOT_FOUND, "unknown Hamiltonian handle " + std::to_string(h));
  return *it->second;
}

void need(bool c, const char *msg) {
  if (!c) fail(ED_ERR_INVALID, msg);
}
}  // namespace

extern "C" {

const char *ed_last_error(void) { return g_last_error.c_str(); }

int ed_request_cancel(void) {
  request_cancel();
  return ED_OK;
}

int ed_version(char *buf, int cap) {
  return guard([&] {
    const char *v = "pyqcm_amd ED 0.1 (gfx950, FP64)";
    need(buf && cap > (int)std::strlen(v), "buffer too small");
    std::strcpy(buf, v);
  });
}

int ed_device_count(int *n) {
  return guard([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) { (void)hipGetLastError(); c = 0; }
    *n = c;
  });
}

// ---- global parameters ----
int ed_set_global_bool(const char *name, int value) {
  return guard([&] {
    Params &P = params();
    std::lock_guard<std::mutex> l(P.mu);
    need(P.b.count(name), "unknown boolean global parameter");
    P.b[name] = value != 0;
  });
}
int ed_set_global_int(const char *name, long value) {
  return guard([&] {
    Params &P = params();
    std::lock_guard<std::mutex> l(P.mu);
    need(P.i.count(name), "unknown integer global parameter");
    P.i[name] = value;
  });
}
int ed_set_global_double(const char *name, double value) {
  return guard([&] {
    Params &P = params();
    std::lock_guard<std::mutex> l(P.mu);
    need(P.d.count(name), "unknown real global parameter");
    P.d[name] = value;
  });
}
int ed_set_global_char(const char *name, char value) {
  return guard([&] {
    Params &P = params();
    std::lock_guard<std::mutex> l(P.mu);
    need(P.c.count(name), "unknown char global parameter");
    P.c[name] = value;
  });
}
int ed_get_global_double(const char *name, double *value) {
  return guard([&] { *value = params().getd(name); });
}
int ed_get_global_int(const char *name, long *value) {
  return guard([&] { *value = params().geti(name); });
}

// ---- model / operator / instance ----
int ed_new_model(const char *name, int n_sites, int n_bath, const int *generators, int n_gen, int bath_irrep) {
  return guard([&] {
    need(n_gen >= 0 && (n_gen == 0 || generators), "ed_new_model: bad generators");
    const int L = n_sites + n_bath;
    std::vector<std::vector<int>> gens(n_gen);
    for (int j = 0; j < n_gen; ++j) gens[j].assign(generators + (size_t)j * L, generators + (size_t)(j + 1) * L);
    new_model(name, n_sites, n_bath, gens, bath_irrep != 0);
  });
}

int ed_new_operator(const char *model, const char *name, const char *type, int n, const int *r, const int *c,
                    const double *v) {
  return guard([&] {
    need(n >= 0, "negative element count");
    std::vector<Element> el(n);
    for (int k = 0; k < n; ++k) el[k] = {r[k], c[k], v[k]};
    new_operator(model, name, type, el);
  });
}

int ed_new_operator_complex(const char *model, const char *name, const char *type, int n, const int *r, const int *c,
                            const double *v) {
  return guard([&] {
    need(n >= 0, "negative element count");
    need(n == 0 || (r && c && v), "ed_new_operator_complex: null element arrays");
    std::vector<Element> el(n);
    for (int k = 0; k < n; ++k) el[k] = {r[k], c[k], v[2 * k], v[2 * k + 1]};
    new_operator(model, name, type, el, true);
  });
}

int ed_new_model_instance(const char *model, int n_values, const char *const *names, const double *values,
                          const char *sectors, int label) {
  return guard([&] {
    std::map<std::string, double> m;
    for (int k = 0; k < n_values; ++k) m[names[k]] = values[k];
    new_instance(model, m, sectors, label);
  });
}

int ed_erase_model_instance(int label) {
  return guard([&] {
    erase_instance(label);
    std::lock_guard<std::mutex> l(g_label_mu);  // no call holds the label's lock now
    g_label_locks.erase(label);
  });
}

int ed_ground_state_solve(int label, double *E0, char *sector, int cap) {
  return guard_label(label, [&] {
    auto r = get_instance(label).ground_state_solve();
    if (E0) *E0 = r.first;
    if (sector && cap > 0) {
      std::strncpy(sector, r.second.c_str(), (size_t)cap - 1);
      sector[cap - 1] = 0;
    }
  });
}

int ed_ground_state_energy(int label, double *E0) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    I.ground_state_solve();
    *E0 = I.E0;
  });
}

int ed_green_function_solve(int label) {
  return guard_label(label, [&] { get_instance(label).green_function_solve(); });
}

int ed_green_function_dimension(int label, int *d) {
  return guard_label(label, [&] { *d = get_instance(label).gf_dim(); });
}

int ed_green_function(int label, int nz, const double *z, int spin_down, int blocks, double *out) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    I.green_function(nz, z, spin_down != 0, out);
    const PointGroup &G = I.model->group;
    if (!blocks || G.g == 1) return;  // trivial group: block form == site form
    // the symmetric-orbital basis (Green_function with blocks, qcm_ED.cpp:220-240):
    // G_sym = S G S^T, the blocks of each representation on the diagonal; the
    // couplings between representations vanish by symmetry and are set to 0
    const int d = I.gf_dim();
    std::vector<int> rep(d);
    for (int r = 0; r < G.g; ++r)
      for (int a = 0; a < G.irrep_dim[r]; ++a) rep[G.irrep_off[r] + a] = r;
    std::vector<cplx> Gs((size_t)d * d), T((size_t)d * d);
    for (int iz = 0; iz < nz; ++iz) {
      cplx *Gz = reinterpret_cast<cplx *>(out) + (size_t)iz * d * d;  // column-major d x d
      for (int a = 0; a < d; ++a)
        for (int j = 0; j < d; ++j) {
          cplx t = 0.0;
          for (int i = 0; i < d; ++i) t += G.S[(size_t)a * d + i] * Gz[i + (size_t)j * d];
          T[a + (size_t)j * d] = t;
        }
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) {
          cplx t = 0.0;
          if (rep[a] == rep[b])
            for (int j = 0; j < d; ++j) t += T[a + (size_t)j * d] * G.S[(size_t)b * d + j];
          Gs[a + (size_t)b * d] = t;
        }
      std::copy(Gs.begin(), Gs.end(), Gz);
    }
  });
}

int ed_green_function_batch(const int *labels, int n, int nz, const double *z, int spin_down, double *out) {
  return guard([&] {
    if (n < 0 || (n > 0 && !labels)) fail(ED_ERR_INVALID, "ed_green_function_batch: bad arguments");
    std::vector<Instance *> insts;
    for (int k = 0; k < n; ++k) insts.push_back(&get_instance(labels[k]));
    green_function_batch(insts, nz, z, spin_down != 0, out);
  });
}

int ed_self_energy(int label, int nz, const double *z, int spin_down, double *out) {
  return guard_label(label, [&] { get_instance(label).self_energy(nz, z, spin_down != 0, out); });
}

int ed_hybridization_function(int label, int nz, const double *z, int spin_down, double *out) {
  return guard_label(label, [&] { get_instance(label).hybridization(nz, z, spin_down != 0, out); });
}

int ed_cluster_averages(int label, int cap, double *ave, double *var, int *n) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    const auto &A = I.cluster_averages();
    *n = (int)A.size();
    for (int k = 0; k < *n && k < cap; ++k) {
      ave[k] = std::get<1>(A[k]);
      var[k] = std::get<2>(A[k]);
    }
  });
}

int ed_qmatrix(int label, int spin_down, int cap, double *e, double *v, int *L, int *M) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    I.green_function_solve();
    need(!I.states.empty(), "no state");
    const GFRep &G = I.states[0]->gf[(spin_down && (I.mixing & 4)) ? 1 : 0];
    if (G.kind != 'L') fail(ED_ERR_INVALID, "qmatrix needs GF_method 'L'");
    *L = G.n;
    *M = (int)G.e.size();
    if (*M <= cap) {
      std::memcpy(e, G.e.data(), sizeof(double) * G.e.size());
      for (size_t k = 0; k < G.V.size(); ++k) {
        v[2 * k] = G.V[k];
        v[2 * k + 1] = G.Vi.empty() ? 0.0 : G.Vi[k];
      }
    }
  });
}

int ed_states(int label, int cap, double *energy, double *weight, int *n) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    I.ground_state_solve();
    *n = (int)I.states.size();
    for (int k = 0; k < *n && k < cap; ++k) {
      energy[k] = I.states[k]->energy;
      weight[k] = I.states[k]->weight;
    }
  });
}

int ed_mixing(int label, int *mixing) {
  return guard_label(label, [&] { *mixing = get_instance(label).mixing; });
}

int ed_solve_batch(const int *labels, int n, int flags) {
  return guard([&] {
    if (n < 0 || (n > 0 && !labels)) fail(ED_ERR_INVALID, "ed_solve_batch: bad arguments");
    std::vector<Instance *> insts;
    for (int k = 0; k < n; ++k) insts.push_back(&get_instance(labels[k]));
    solve_batch(insts, (flags & 1) != 0);
  });
}

// ---- lattice sums ----
static Superlattice superlattice_from(const ed_superlattice *V) {
  if (!V || V->d <= 0 || !V->V0 || V->ne < 0 || V->nnbr < 0 || (V->ne > 0 && (!V->row || !V->col || !V->nbr || !V->val)) ||
      (V->nnbr > 0 && !V->R))
    fail(ED_ERR_INVALID, "superlattice: bad arguments");
  Superlattice S;
  S.d = V->d;
  S.V0.assign(V->V0, V->V0 + 2 * (size_t)V->d * V->d);
  for (int e = 0; e < V->ne; ++e) {
    if (V->row[e] < 0 || V->row[e] >= V->d || V->col[e] < 0 || V->col[e] >= V->d || V->nbr[e] < 0 ||
        V->nbr[e] >= V->nnbr)
      fail(ED_ERR_INVALID, "superlattice: element " + std::to_string(e) + " out of range");
    S.row.push_back(V->row[e]);
    S.col.push_back(V->col[e]);
    S.nbr.push_back(V->nbr[e]);
    S.val.push_back(V->val[2 * e]);
    S.val.push_back(V->val[2 * e + 1]);
  }
  S.R.assign(V->R, V->R + 3 * (size_t)V->nnbr);
  return S;
}

static KGrid kgrid_from(const ed_kgrid *K) {
  if (!K) fail(ED_ERR_INVALID, "wavevector grid: null");
  KGrid G;
  G.dim = K->dim;
  for (int a = 0; a < 3; ++a) G.nk[a] = K->nk[a];
  G.offset = K->offset;
  return G;
}

int ed_cpt_projected_gf(int label, int nz, const double *z, int spin_down, const ed_superlattice *V,
                        const ed_kgrid *K, double *out) {
  return guard_label(label, [&] {
    if (nz < 0 || (nz > 0 && (!z || !out))) fail(ED_ERR_INVALID, "ed_cpt_projected_gf: bad arguments");
    Instance &I = get_instance(label);
    const Superlattice S = superlattice_from(V);
    const KGrid G = kgrid_from(K);
    const int d = I.gf_dim();
    if (S.d != d) fail(ED_ERR_INVALID, "superlattice dimension differs from the Green-function dimension");
    std::vector<double> g((size_t)nz * d * d * 2);
    I.green_function(nz, z, spin_down != 0, g.data());
    DevBuf<double> dg;
    dg.upload(g);
    SuperlatticeDev SD(S);
    lattice_cpt_average(SD, G, nz, dg.p, out);
  });
}

int ed_sef_grid(int label, int nw, const double *w, const double *weights, const ed_superlattice *Vup,
                const ed_superlattice *Vdn, const ed_kgrid *K, double *integral, double *per_freq) {
  return guard_label(label, [&] {
    if (nw < 0 || (nw > 0 && (!w || !weights)) || !integral) fail(ED_ERR_INVALID, "ed_sef_grid: bad arguments");
    Instance &I = get_instance(label);
    if (!I.gf_solved) I.green_function_solve();
    const bool two = (I.mixing & 4) != 0;
    if (two && !Vdn) fail(ED_ERR_INVALID, "ed_sef_grid: the spin-down superlattice is required with up/down mixing");
    const KGrid G = kgrid_from(K);
    const int d = I.gf_dim();
    std::vector<double> z(2 * (size_t)nw), acc(nw, 0.0), part(nw);
    for (int n = 0; n < nw; ++n) {
      z[2 * n] = 0.0;
      z[2 * n + 1] = w[n];
    }
    for (int spin = 0; spin < (two ? 2 : 1); ++spin) {
      const Superlattice S = superlattice_from(spin ? Vdn : Vup);
      if (S.d != d) fail(ED_ERR_INVALID, "superlattice dimension differs from the Green-function dimension");
      std::vector<double> g((size_t)nw * d * d * 2);
      I.green_function(nw, z.data(), spin == 1, g.data());
      DevBuf<double> dg;
      dg.upload(g);
      SuperlatticeDev SD(S);
      lattice_logdet(SD, G, nw, dg.p, part.data());
      for (int n = 0; n < nw; ++n) acc[n] += two ? part[n] : 2.0 * part[n];
    }
    double s = 0.0;
    for (int n = 0; n < nw; ++n) s += weights[n] * acc[n];
    *integral = s;
    if (per_freq) std::copy(acc.begin(), acc.end(), per_freq);
  });
}

int ed_potthoff_functional(int label, double *omega) {
  return guard_label(label, [&] {
    if (!omega) fail(ED_ERR_INVALID, "ed_potthoff_functional: null output");
    Instance &I = get_instance(label);
    I.ground_state_solve();
    *omega = I.E0 + I.sef_bath();
  });
}

int ed_instance_parameters(int label, int cap, int name_len, char *names, double *values, int *n, char *model,
                           int model_cap) {
  return guard_label(label, [&] {
    need(n != nullptr && (cap == 0 || (names && values && name_len > 0)), "ed_instance_parameters: bad arguments");
    Instance &I = get_instance(label);
    *n = (int)I.value.size();
    int k = 0;
    for (const auto &kv : I.value) {
      if (k < cap) {
        need((int)kv.first.size() < name_len, "ed_instance_parameters: name buffer too small");
        std::memcpy(names + (size_t)k * name_len, kv.first.c_str(), kv.first.size() + 1);
        values[k] = kv.second;
      }
      ++k;
    }
    if (model) {
      need(model_cap > (int)I.model->name.size(), "ed_instance_parameters: model buffer too small");
      std::strcpy(model, I.model->name.c_str());
    }
  });
}

int ed_complex_hs(int label, int *is_complex) {
  return guard_label(label, [&] {
    need(is_complex != nullptr, "ed_complex_hs: null output");
    *is_complex = get_instance(label).complex_hs() ? 1 : 0;
  });
}

int ed_hopping_matrix(int label, int spin_down, int diag, int full, double *out) {
  return guard_label(label, [&] {
    need(out != nullptr, "ed_hopping_matrix: null output");
    Instance &I = get_instance(label);
    const int ns = I.model->n_sites, nb = I.model->n_bath;
    const Instance::Hopping H = I.hopping_blocks(spin_down != 0, full && diag);
    const int d = full ? ns + nb : ns;
    const bool cx = !H.tci.empty();
    auto im = [cx](const std::vector<double> &x, size_t k) { return cx ? x[k] : 0.0; };
    std::vector<cplx> m((size_t)d * d, 0.0);
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < ns; ++b) {
        const size_t k = a + (size_t)b * ns;
        m[a + (size_t)b * d] = cplx(H.tc[k], im(H.tci, k));
      }
    if (full) {
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
          const size_t k = i + (size_t)j * nb;
          m[ns + i + (size_t)(ns + j) * d] = cplx(H.tb[k], im(H.tbi, k));
        }
      for (int a = 0; a < ns; ++a)
        for (int i = 0; i < nb; ++i) {
          const size_t k = a + (size_t)i * ns;
          const cplx t(H.tcb[k], im(H.tcbi, k));
          m[a + (size_t)(ns + i) * d] = t;
          m[ns + i + (size_t)a * d] = std::conj(t);
        }
    }
    for (size_t k = 0; k < m.size(); ++k) {
      out[2 * k] = m[k].real();
      out[2 * k + 1] = m[k].imag();
    }
  });
}

int ed_interactions(int label, int cap, int *r, int *c, double *v, int *n) {
  return guard_label(label, [&] {
    need(n != nullptr && (cap == 0 || (r && c && v)), "ed_interactions: bad output arrays");
    Instance &I = get_instance(label);
    int k = 0;
    for (const auto &kv : I.value) {
      const Operator &op = I.model->ops.at(kv.first);
      if (op.kind != OP_INTERACTION) continue;
      for (const auto &x : op.el) {
        if (k < cap) {
          r[k] = x.r;
          c[k] = x.c;
          v[k] = x.v * kv.second;
        }
        ++k;
      }
    }
    *n = k;
  });
}

int ed_hybridization_lehmann(int label, int spin_down, double *e, double *tcb) {
  return guard_label(label, [&] {
    Instance &I = get_instance(label);
    const int ns = I.model->n_sites, nb = I.model->n_bath;
    if (nb == 0) fail(ED_ERR_INVALID, "System has no bath. Hybridization not defined.");
    need(e && tcb, "ed_hybridization_lehmann: null output");
    const Instance::Hopping H = I.hopping_blocks(spin_down != 0, true);
    for (int i = 0; i < nb; ++i) e[i] = H.tb[i + (size_t)i * nb];
    for (size_t k = 0; k < (size_t)ns * nb; ++k) {
      tcb[2 * k] = H.tcb[k];
      tcb[2 * k + 1] = H.tcbi.empty() ? 0.0 : H.tcbi[k];
    }
  });
}

int ed_matrix_elements(const char *model, const char *op, char *type, int type_cap, int cap, int *r, int *c,
                       double *v, int *n) {
  return guard([&] {
    need(model && op && n && (cap == 0 || (r && c && v)), "ed_matrix_elements: bad arguments");
    Model &M = get_model(model);
    auto it = M.ops.find(op);
    if (it == M.ops.end())
      fail(ED_ERR_INVALID, std::string("operator ") + op + " is not defined in model . Check spelling.");
    const std::string t = it->second.kind == OP_INTERACTION ? "interaction" : "one_body";
    if (type) {
      need(type_cap > (int)t.size(), "type buffer too small");
      std::strcpy(type, t.c_str());
    }
    const auto &el = it->second.el;
    *n = (int)el.size();
    for (int k = 0; k < *n && k < cap; ++k) {
      r[k] = el[k].r;
      c[k] = el[k].c;
      v[2 * k] = el[k].v;
      v[2 * k + 1] = el[k].vi;
    }
  });
}

int ed_density_matrix(int label, const int *sites, int nsites, int cap, double *rho, uint64_t *basis, int *dim_a) {
  return guard_label(label, [&] {
    need(nsites >= 0 && (nsites == 0 || sites) && dim_a, "ed_density_matrix: bad arguments");
    std::vector<double> r;
    std::vector<uint64_t> b;
    get_instance(label).density_matrix(std::vector<int>(sites, sites + nsites), r, b);
    *dim_a = (int)b.size();
    if ((int)b.size() <= cap) {
      need(rho && basis, "ed_density_matrix: null output");
      for (size_t k = 0; k < r.size(); ++k) {
        rho[2 * k] = r[k];
        rho[2 * k + 1] = 0.0;
      }
      std::memcpy(basis, b.data(), b.size() * sizeof(uint64_t));
    }
  });
}

int ed_print_wavefunction(int label, char *buf, int cap, int *len) {
  return guard_label(label, [&] {
    need(len != nullptr && (cap == 0 || buf), "ed_print_wavefunction: bad arguments");
    const std::string s = get_instance(label).print_wavefunction();
    *len = (int)s.size();
    if ((int)s.size() < cap) std::memcpy(buf, s.c_str(), s.size() + 1);
  });
}

int ed_fidelity(int label1, int label2, double *f) {
  return guard([&] {
    need(f != nullptr, "ed_fidelity: null output");
    *f = get_instance(label1).fidelity(get_instance(label2));
  });
}

int ed_write_instance_hdf5(const char *filename, int label, const char *group) {
  return guard([&] {
    need(filename && group, "ed_write_instance_hdf5: null argument");
    write_instance_hdf5(filename, label, group);
  });
}

int ed_read_instance_hdf5(const char *filename, int label, const char *group) {
  return guard([&] {
    need(filename && group, "ed_read_instance_hdf5: null argument");
    read_instance_hdf5(filename, label, group);
  });
}

int ed_hdf5_version(char *buf, int cap) {
  return guard([&] {
    const std::string v = hdf5_version();
    need(buf && cap > (int)v.size(), "buffer too small");
    std::strcpy(buf, v.c_str());
  });
}

int ed_green_function_average(int label, int spin_down, double *out) {
  return guard_label(label, [&] {
    need(out != nullptr, "ed_green_function_average: null output");
    Instance &I = get_instance(label);
    const std::vector<cplx> G = I.gf_average(spin_down != 0);
    for (size_t k = 0; k < G.size(); ++k) {
      out[2 * k] = G[k].real();
      out[2 * k + 1] = G[k].imag();
    }
  });
}

int ed_green_function_density(int label, double *density) {
  return guard_label(label, [&] {
    need(density != nullptr, "ed_green_function_density: null output");
    *density = get_instance(label).gf_density();
  });
}

int ed_tr_sigma_inf(int label, double *value) {
  return guard_label(label, [&] {
    need(value != nullptr, "ed_tr_sigma_inf: null output");
    *value = get_instance(label).tr_sigma_inf();
  });
}

int ed_susceptibility(int label, const char *op, int nw, const double *w, double *chi) {
  return guard_label(label, [&] {
    need(op != nullptr, "ed_susceptibility: null operator name");
    need(nw >= 0 && (nw == 0 || (w && chi)), "ed_susceptibility: bad frequency arrays");
    const auto parts = get_instance(label).susceptibility_parts(op);
    // all states' poles in one Lehmann sum, residues weight * v^2, evaluated
    // for every frequency in one launch (the 1 x 1 case of the Green-function
    // Lehmann kernel)
    std::vector<double> e, v;
    for (const auto &c : parts) {
      const double sw = std::sqrt(c.weight);
      for (size_t i = 0; i < c.e.size(); ++i) {
        e.push_back(c.e[i]);
        v.push_back(c.v[i] * sw);
      }
    }
    if (nw == 0) return;
    if (e.empty()) {
      std::memset(chi, 0, sizeof(double) * 2 * (size_t)nw);
      return;
    }
    DevBuf<double> de, dv, dz, dchi((size_t)2 * nw);
    de.upload(e);
    dv.upload(v);
    dz.upload(w, (size_t)2 * nw);
    vdvh(1, (int)e.size(), de.p, dv.p, nw, dz.p, dchi.p, stream());
    dchi.download(chi, (size_t)2 * nw);
  });
}

int ed_susceptibility_poles(int label, const char *op, int cap, double *poles, int *n) {
  return guard_label(label, [&] {
    need(op != nullptr && n != nullptr, "ed_susceptibility_poles: null argument");
    need(cap >= 0 && (cap == 0 || poles), "ed_susceptibility_poles: bad output array");
    const auto parts = get_instance(label).susceptibility_parts(op);
    // poles above 1e-8 with residue weight v^2 above 1e-6, in the order found;
    // a pole within 1e-4 of one already kept adds its residue to that one
    std::vector<std::pair<double, double>> chi, kept;
    for (const auto &c : parts)
      for (size_t i = 0; i < c.e.size(); ++i) {
        if (c.e[i] < 1e-8) continue;
        const double r = c.v[i] * c.v[i] * c.weight;
        if (r > 1e-6) chi.push_back({c.e[i], r});
      }
    for (const auto &x : chi) {
      size_t j = 0;
      for (; j < kept.size(); ++j)
        if (std::fabs(kept[j].first - x.first) < 1e-4) {
          kept[j].second += x.second;
          break;
        }
      if (j == kept.size()) kept.push_back(x);
    }
    *n = (int)kept.size();
    for (int k = 0; k < *n && k < cap; ++k) {
      poles[2 * k] = kept[k].first;
      poles[2 * k + 1] = kept[k].second;
    }
  });
}

// ---- Hamiltonian handles ----
int ed_hamiltonian_create(int label, const char *sector, char format, int64_t *handle) {
  return guard([&] {
    Instance &I = get_instance(label);
    Sector s = Sector::parse(sector);
    if (!sector_valid(I.model->L, s, I.model->group.g))
      fail(ED_ERR_SECTOR, std::string("sector ") + sector + " is not valid");
    auto H = make_hamiltonian(I.terms(), s, format ? format : I.format);
    const int64_t id = g_next_h++;
    std::lock_guard<std::mutex> l(g_h_mu);
    g_hams[id] = std::move(H);
    *handle = id;
  });
}

int ed_hamiltonian_destroy(int64_t handle) {
  return guard([&] {
    std::lock_guard<std::mutex> l(g_h_mu);
    g_hams.erase(handle);
  });
}

int ed_hamiltonian_dim(int64_t handle, int64_t *dim) {
  return guard([&] { *dim = (int64_t)get_h(handle).dim; });
}

int ed_hamiltonian_mult_add(int64_t handle, const double *x, double *y) {
  return guard([&] {
    Hamiltonian &H = get_h(handle);
    DevBuf<double> dx, dy;
    dx.upload(x, H.dim);
    dy.upload(y, H.dim);
    H.apply(dx.p, dy.p, 1.0, 1.0, false);
    dy.download(y, H.dim);
  });
}

int ed_hamiltonian_apply_device(int64_t handle, const double *dx, double *dy, double beta) {
  return guard([&] {
    Hamiltonian &H = get_h(handle);
    H.apply(dx, dy, 1.0, beta, false);
    sync();
  });
}

int ed_hamiltonian_apply_dot_noalloc(int64_t handle, const double *dx, double *dy, double beta, double *d_dot) {
  return guard([&] {
    need(d_dot != nullptr, "ed_hamiltonian_apply_dot_noalloc: null dot output");
    Hamiltonian &H = get_h(handle);
    need(H.dim == 0 || (dx && dy), "ed_hamiltonian_apply_dot_noalloc: null vector");
    const NoAllocScratch S = noalloc_scratch();
    const int nb = H.apply_fixed(dx, dy, 1.0, beta, S.part, S.aux, S.aux_cap);
    fold_partials_device(S.part, nb, d_dot, stream());
    sync();
  });
}

int ed_hamiltonian_diag(int64_t handle, double *d) {
  return guard([&] {
    Hamiltonian &H = get_h(handle);
    DevBuf<double> dd(H.dim);
    H.diag(dd.p);
    dd.download(d, H.dim);
  });
}

int ed_hamiltonian_csr_pattern(int64_t handle, int64_t *nnz, int32_t *Iptr, int32_t *J, double *v) {
  return guard([&] {
    Hamiltonian &H = get_h(handle);
    if (terms_complex(H.terms))
      fail(ED_ERR_UNSUPPORTED, "the CSR pattern of a complex Hilbert space's Hamiltonian is not provided");
    std::vector<int32_t> ip, jj;
    std::vector<double> vv, dg;
    csr_lower(H.terms, H.Nup, H.Ndw, ip, jj, vv, dg);
    const int64_t cap = *nnz;
    *nnz = (int64_t)jj.size();
    if (Iptr && J && cap >= (int64_t)jj.size()) {
      std::memcpy(Iptr, ip.data(), sizeof(int32_t) * ip.size());
      std::memcpy(J, jj.data(), sizeof(int32_t) * jj.size());
      if (v) std::memcpy(v, vv.data(), sizeof(double) * vv.size());
    }
  });
}

int ed_hamiltonian_autotune(int64_t handle, int cap, double *ms, int *n, int *best) {
  return guard([&] {
    std::vector<double> t = hamiltonian_autotune(get_h(handle), *best);
    *n = (int)t.size();
    for (int k = 0; k < *n && k < cap; ++k) ms[k] = t[k];
  });
}

int ed_hamiltonian_set_variant(int64_t handle, int variant) {
  return guard([&] {
    need(variant >= 0 && variant <= 3, "ed_hamiltonian_set_variant: variant 0 .. 3");
    get_h(handle).variant = variant;
  });
}

int ed_hamiltonian_time(int64_t handle, int nrep, double *ms, int variant) {
  return guard([&] {
    Hamiltonian &H = get_h(handle);
    const int old = H.variant;
    H.variant = variant;
    DevBuf<double> x(H.dim), y(H.dim);
    fill(H.dim, x.p, 1.0 / std::sqrt((double)H.dim), stream());
    fill(H.dim, y.p, 0.0, stream());
    for (int w = 0; w < 2; ++w) H.apply(x.p, y.p, 1.0, 0.0, false);
    hipEvent_t e0, e1;
    ED_HIP(hipEventCreate(&e0));
    ED_HIP(hipEventCreate(&e1));
    ED_HIP(hipEventRecord(e0, stream()));
    for (int r = 0; r < nrep; ++r) H.apply(x.p, y.p, 1.0, 0.0, false);
    ED_HIP(hipEventRecord(e1, stream()));
    ED_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    ED_HIP(hipEventElapsedTime(&t, e0, e1));
    ED_HIP(hipEventDestroy(e0));
    ED_HIP(hipEventDestroy(e1));
    *ms = (double)t / std::max(1, nrep);
    H.variant = old;
  });
}

// ---- bases and sectors ----
int ed_half_basis(int L, int N, uint32_t *out, int64_t cap, int64_t *dim) {
  return guard([&] {
    auto B = half_basis(L, N);
    *dim = B->dim;
    if (out && cap >= (int64_t)B->dim) std::memcpy(out, B->bin.data(), sizeof(uint32_t) * B->dim);
  });
}

int ed_random_start(int64_t n, int64_t seed, double *out) {
  return guard([&] {
    need(n >= 0 && (n == 0 || out != nullptr), "ed_random_start: null output");
    std::vector<double> x;
    random_start((size_t)n, (long)seed, x);
    if (n > 0) std::memcpy(out, x.data(), sizeof(double) * (size_t)n);
  });
}

int ed_sector_parse(const char *s, int *N, int *S, int *irrep) {
  return guard([&] {
    Sector x = Sector::parse(s);
    *N = x.N;
    *S = x.S;
    *irrep = x.irrep;
  });
}

// ---- Krylov drivers on handles (host vectors) ----
int ed_lanczos_gs(int64_t h, const double *start, double *E0, double *gs, int *niter) {
  return guard([&] {
    Hamiltonian &H = get_h(h);
    DevBuf<double> d0, dg(H.dim);
    d0.upload(start, H.dim);
    LanczosGS R = lanczos_gs(H, d0.p, dg.p);
    *E0 = R.E0;
    if (niter) *niter = R.niter;
    if (gs) dg.download(gs, H.dim);
  });
}

int ed_lanczos_green(int64_t h, const double *psi, int cap, double *a, double *b, int *n) {
  return guard([&] {
    Hamiltonian &H = get_h(h);
    DevBuf<double> d0;
    d0.upload(psi, H.dim);
    std::vector<double> av, bv;
    lanczos_green(H, d0.p, av, bv);
    *n = (int)av.size();
    if ((int)av.size() > cap)
      fail(ED_ERR_INVALID, "ed_lanczos_green: " + std::to_string(av.size()) + " coefficients exceed cap " +
                               std::to_string(cap) + " (*n holds the size needed)");
    std::memcpy(a, av.data(), sizeof(double) * av.size());
    std::memcpy(b, bv.data(), sizeof(double) * bv.size());
  });
}

int ed_band_lanczos(int64_t h, int p, const double *phi, int cap, double *evals, double *Q, int *M) {
  return guard([&] {
    Hamiltonian &H = get_h(h);
    std::vector<DevBuf<double>> bufs;
    std::vector<const double *> ptrs;
    for (int k = 0; k < p; ++k) {
      bufs.emplace_back();
      bufs.back().upload(phi + (size_t)k * H.dim, H.dim);
    }
    for (auto &b : bufs) ptrs.push_back(b.p);
    std::vector<double> e, q;
    band_lanczos(H, ptrs, e, q);
    *M = (int)e.size();
    if (*M > cap)
      fail(ED_ERR_INVALID, "ed_band_lanczos: " + std::to_string(e.size()) + " poles exceed cap " +
                               std::to_string(cap) + " (*M holds the size needed)");
    std::memcpy(evals, e.data(), sizeof(double) * e.size());
    std::memcpy(Q, q.data(), sizeof(double) * q.size());
  });
}

int ed_davidson(int64_t h, int K, double *evals, double *evecs, int *niter) {
  return guard([&] {
    Hamiltonian &H = get_h(h);
    std::vector<double> ev;
    std::vector<DevBuf<double>> vecs;
    const int it = davidson(H, K, params().getd("accur_Davidson"), ev, vecs);
    for (int k = 0; k < (int)ev.size(); ++k) {
      evals[k] = ev[k];
      if (evecs) vecs[k].download(evecs + (size_t)k * H.dim, H.dim);
    }
    if (niter) *niter = it;
  });
}

int ed_block_lanczos(int64_t h, int p, const double *q0, int M0, int cap, double *A, double *B, int *nfl) {
  return guard([&] {
    Hamiltonian &H = get_h(h);
    std::vector<DevBuf<double>> bufs;
    std::vector<const double *> ptrs;
    for (int k = 0; k < p; ++k) {
      bufs.emplace_back();
      bufs.back().upload(q0 + (size_t)k * H.dim, H.dim);
    }
    for (auto &b : bufs) ptrs.push_back(b.p);
    std::vector<double> Av, Bv;
    const int n = block_lanczos(H, ptrs, M0, Av, Bv, params().getb("block_Lanczos_QR"));
    *nfl = n;
    if (n > cap)
      fail(ED_ERR_INVALID, "ed_block_lanczos: " + std::to_string(n) + " floors exceed cap " + std::to_string(cap) +
                               " (*nfl holds the size needed)");
    std::memcpy(A, Av.data(), sizeof(double) * Av.size());
    std::memcpy(B, Bv.data(), sizeof(double) * Bv.size());
  });
}

int ed_create_or_destroy(int L, int pm, int a, int Nup, int Ndw, const double *x, double z, double *y) {
  return guard([&] {
    need(pm == 1 || pm == -1, "pm must be +1 or -1");
    need(a >= 0 && a < 2 * L, "orbital out of range");
    Sector src(Nup + Ndw, Nup - Ndw, 0);
    const Sector tgt = shift_sector(src, pm, a >= L);
    const size_t ds = sector_dim(L, src), dt = sector_dim(L, tgt);
    DevBuf<double> dx, dy;
    dx.upload(x, ds);
    dy.upload(y, dt);
    if (dt && ds) cdag_apply(L, pm, a, src, dx.p, z, dy.p);
    dy.download(y, dt);
  });
}

int ed_tridiag_lowest(int n, const double *d, const double *e, double guess, double *ev0, double *z) {
  return guard([&] {
    if (n <= 0 || !d || (n > 1 && !e) || !ev0) fail(ED_ERR_INVALID, "ed_tridiag_lowest: bad arguments");
    std::vector<double> zz;
    double ev;
    tridiag_eigvec0(n, d, e, ev, zz, guess);
    *ev0 = ev;
    if (z) std::copy(zz.begin(), zz.end(), z);
  });
}

// ---- GF kernels ----
int ed_cf_eval(int ncf, const int *off, const double *a, const double *b, int nz, const double *z, double *out) {
  return guard([&] {
    DevBuf<int> doff;
    DevBuf<double> da, db, dz, dout((size_t)2 * ncf * nz + 1);
    doff.upload(off, (size_t)ncf + 1);
    const size_t na = off[ncf];
    std::vector<double> av(a, a + na), bv(b, b + na);
    da.upload(av.empty() ? std::vector<double>{0.0} : av);
    db.upload(bv.empty() ? std::vector<double>{0.0} : bv);
    dz.upload(z, (size_t)2 * nz);
    cf_eval(ncf, doff.p, da.p, db.p, nz, dz.p, dout.p, stream());
    dout.download(out, (size_t)2 * ncf * nz);
  });
}

int ed_vdvh(int L, int M, const double *e, const double *V, int nz, const double *z, double *G) {
  return guard([&] {
    DevBuf<double> de, dV, dz, dG((size_t)2 * nz * L * L + 1);
    de.upload(e, M);
    dV.upload(V, (size_t)L * M);
    dz.upload(z, (size_t)2 * nz);
    vdvh(L, M, de.p, dV.p, nz, dz.p, dG.p, stream());
    dG.download(G, (size_t)2 * nz * L * L);
  });
}

int ed_mcf_eval(int p, int nfl, const double *A, const double *B, const double *W, int q, int nz, const double *z,
                double *G) {
  return guard([&] {
    DevBuf<double> dA, dB, dW, dz, dG((size_t)2 * nz * q * q + 1);
    dA.upload(A, (size_t)nfl * p * p);
    dB.upload(B, (size_t)nfl * p * p);
    dW.upload(W, (size_t)p * q);
    dz.upload(z, (size_t)2 * nz);
    mcf_eval(p, nfl, dA.p, dB.p, dW.p, q, nz, dz.p, dG.p, stream());
    dG.download(G, (size_t)2 * nz * q * q);
  });
}

// ---- device memory ----
int ed_device_alloc(int64_t bytes, void **ptr) {
  return guard([&] { *ptr = dev_alloc((size_t)bytes); });
}
int ed_device_free(void *ptr, int64_t bytes) {
  return guard([&] { dev_free(ptr, (size_t)bytes); });
}
int ed_device_copy(void *dst, const void *src, int64_t bytes, int kind) {
  return guard([&] {
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    ED_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, stream()));
    sync();
  });
}
int ed_device_sync(void) {
  return guard([&] { sync(); });
}
int ed_release_cache(void) {
  return guard([&] { dev_release_cache(); });
}

int ed_continued_fractions(int label, int spin_down, int state, int cap, double *a, double *b, int *len,
                           int *nfrac) {
  return guard_label(label, [&] {
    need(cap >= 0 && nfrac != nullptr, "ed_continued_fractions: bad arguments");
    Instance &I = get_instance(label);
    std::vector<std::vector<double>> av, bv;
    cf_export(I, spin_down ? 1 : 0, state, av, bv);
    *nfrac = (int)av.size();
    if (!a || !b || !len) return;
    for (size_t f = 0; f < av.size(); ++f) {
      len[f] = (int)av[f].size();
      for (int k = 0; k < cap; ++k) {
        a[f * cap + k] = k < (int)av[f].size() ? av[f][k] : 0.0;
        b[f * cap + k] = k < (int)bv[f].size() ? bv[f][k] : 0.0;
      }
    }
  });
}

int ed_continued_fractions_pairs(int label, int spin_down, int npairs, const int *o1, const int *o2, int cap,
                                 double *a, double *b, int *len) {
  return guard_label(label, [&] {
    need(npairs >= 0 && cap >= 1 && (npairs == 0 || (o1 && o2 && a && b && len)),
         "ed_continued_fractions_pairs: bad arguments");
    Instance &I = get_instance(label);
    std::vector<std::pair<int, int>> pr;
    for (int k = 0; k < npairs; ++k) pr.push_back({o1[k], o2[k]});
    std::vector<std::vector<double>> av, bv;
    cf_pairs(I, spin_down ? 1 : 0, pr, cap, av, bv);
    for (size_t f = 0; f < av.size(); ++f) {
      len[f] = (int)av[f].size();
      for (int k = 0; k < cap; ++k) {
        a[f * cap + k] = k < (int)av[f].size() ? av[f][k] : 0.0;
        b[f * cap + k] = k < (int)bv[f].size() ? bv[f][k] : 0.0;
      }
    }
  });
}

int ed_set_continued_fractions(int label, double E0, const char *sector, int spin_down, int nfrac, int cap,
                               const double *a, const double *b, const int *len) {
  return guard([&] {
    need(sector && a && b && len && cap >= 0 && nfrac >= 0, "ed_set_continued_fractions: bad arguments");
    Instance &I = get_instance(label);
    std::vector<std::vector<double>> av(nfrac), bv(nfrac);
    for (int f = 0; f < nfrac; ++f) {
      need(len[f] >= 0 && len[f] <= cap, "ed_set_continued_fractions: a length exceeds cap");
      av[f].assign(a + (size_t)f * cap, a + (size_t)f * cap + len[f]);
      bv[f].assign(b + (size_t)f * cap, b + (size_t)f * cap + len[f]);
    }
    cf_import(I, E0, Sector::parse(sector), spin_down ? 1 : 0, av, bv);
  });
}

int ed_combined_mcf(int label, int spin_down, int cap, double *W, double *A, double *B, int *p, int *nfl) {
  return guard_label(label, [&] {
    need(p && nfl && cap >= 0, "ed_combined_mcf: bad arguments");
    Instance &I = get_instance(label);
    if (I.gf_method != 'M')
      fail(ED_ERR_INVALID, "get_combined_mcf: the instance at label " + std::to_string(label) +
                               " does not use the MCF format (GF_method must be 'M').");
    if (!params().getb("combine_mcf"))
      fail(ED_ERR_INVALID, "get_combined_mcf: the combined MCF is not built unless the global parameter "
                           "'combine_mcf' is set to True.");
    if (!I.gf_solved) I.green_function_solve();
    if (I.states.empty()) fail(ED_ERR_INVALID, "get_combined_mcf: no valid ground state found.");
    const GFRep::MCF &m = I.states[0]->gf[I.gf_spin(spin_down != 0)].comb;
    if (m.nfl == 0) fail(ED_ERR_INVALID, "get_combined_mcf: combined MCF is empty.");
    *p = m.p;
    *nfl = m.nfl;
    if (m.nfl > cap || !W || !A || !B) return;
    std::copy(m.W.begin(), m.W.end(), W);
    std::copy(m.A.begin(), m.A.end(), A);
    std::copy(m.B.begin(), m.B.end(), B);
  });
}

// ---- learned surrogate ----
int ed_surrogate_create(int n_in, int n_out, int width, int depth, int64_t seed, int64_t *handle) {
  return guard([&] {
    need(handle != nullptr, "ed_surrogate_create: null handle");
    auto S = std::make_unique<Surrogate>(n_in, n_out, width, depth, (uint64_t)seed);
    const int64_t h = g_next_h.fetch_add(1);
    std::lock_guard<std::mutex> l(g_h_mu);
    g_surr[h] = std::move(S);
    *handle = h;
  });
}
int ed_surrogate_destroy(int64_t handle) {
  return guard([&] {
    std::lock_guard<std::mutex> l(g_h_mu);
    if (!g_surr.erase(handle)) fail(ED_ERR_NOT_FOUND, "unknown surrogate handle " + std::to_string(handle));
  });
}
int ed_surrogate_n_params(int64_t handle, int64_t *n) {
  return guard([&] {
    need(n != nullptr, "ed_surrogate_n_params: null output");
    *n = (int64_t)get_surr(handle).n_params();
  });
}
int ed_surrogate_set_scaling(int64_t handle, const double *x_mean, const double *x_std, const double *y_mean,
                             const double *y_std, const double *loss_weight) {
  return guard([&] { get_surr(handle).set_scaling(x_mean, x_std, y_mean, y_std, loss_weight); });
}
int ed_surrogate_set_optimizer(int64_t handle, double lr, double beta1, double beta2, double eps) {
  return guard([&] {
    need(lr > 0.0 && beta1 >= 0.0 && beta1 < 1.0 && beta2 >= 0.0 && beta2 < 1.0 && eps > 0.0,
         "ed_surrogate_set_optimizer: bad hyper-parameters");
    Surrogate &S = get_surr(handle);
    S.lr = lr;
    S.beta1 = beta1;
    S.beta2 = beta2;
    S.eps = eps;
  });
}
int ed_surrogate_get_weights(int64_t handle, double *w) {
  return guard([&] {
    need(w != nullptr, "ed_surrogate_get_weights: null output");
    get_surr(handle).get_weights(w);
  });
}
int ed_surrogate_set_weights(int64_t handle, const double *w) {
  return guard([&] {
    need(w != nullptr, "ed_surrogate_set_weights: null input");
    get_surr(handle).set_weights(w);
  });
}
int ed_surrogate_predict(int64_t handle, int n, const double *x, double *y) {
  return guard([&] {
    need(n >= 0 && (n == 0 || (x && y)), "ed_surrogate_predict: bad arguments");
    if (n) get_surr(handle).predict(n, x, y);
  });
}
int ed_surrogate_loss_grad(int64_t handle, int n, const double *x, const double *y, double *loss, double *grad) {
  return guard([&] {
    need(n >= 1 && x && y && loss, "ed_surrogate_loss_grad: bad arguments");
    *loss = get_surr(handle).loss_grad_host(n, x, y, grad);
  });
}
int ed_surrogate_train(int64_t handle, int n, const double *x, const double *y, int epochs, int batch,
                       int64_t seed, double *epoch_loss, double *ms_per_step) {
  return guard([&] {
    need(n >= 1 && x && y && epochs >= 0 && batch >= 1, "ed_surrogate_train: bad arguments");
    const std::vector<double> h = get_surr(handle).train(n, x, y, epochs, batch, (uint64_t)seed, ms_per_step);
    if (epoch_loss) std::copy(h.begin(), h.end(), epoch_loss);
  });
}
int ed_surrogate_time(int64_t handle, int batch, int nrep, double *ms_forward, double *ms_forward_backward) {
  return guard([&] {
    need(batch >= 1 && nrep >= 1 && ms_forward && ms_forward_backward, "ed_surrogate_time: bad arguments");
    get_surr(handle).time(batch, nrep, ms_forward, ms_forward_backward);
  });
}

}  // extern "C"
