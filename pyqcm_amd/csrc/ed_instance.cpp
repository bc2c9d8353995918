// ed_instance.cpp -- low-energy states, Green-function build and evaluation.
#include "ed_instance.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <iomanip>
#include <mutex>
#include <random>
#include <sstream>

#include "ed_dense.hpp"
#include "ed_solvers.hpp"

namespace ed {

namespace {
std::mutex g_inst_mu;
std::map<int, std::unique_ptr<Instance>> g_instances;
constexpr double kSmall = 1.0e-12;  // SMALL_VALUE (src_util/vector_num.hpp:23)
constexpr double kMinGap = 1.0e-6;  // MIN_GAP (src_ed/model.hpp:19)
}  // namespace

Instance &get_instance(int label) {
  std::lock_guard<std::mutex> l(g_inst_mu);
  auto it = g_instances.find(label);
  if (it == g_instances.end())
    fail(ED_ERR_NOT_FOUND, "The cluster instance label " + std::to_string(label) + " is out of range.");
  return *it->second;
}

void erase_instance(int label) {
  std::lock_guard<std::mutex> l(g_inst_mu);
  g_instances.erase(label);
}

// src_ed/symmetry_group.cpp:515-533; g: order of the model's point group
bool sector_valid(int L, const Sector &s, int g) {
  if (s.N < Sector::odd) {
    if (s.N < 0 || s.N > 2 * L) return false;
    if (s.S < Sector::odd) {
      if (s.S > s.N || s.S < -s.N || ((s.N + s.S) % 2)) return false;
    }
  }
  if (s.S < Sector::odd && (s.S > L || s.S < -L)) return false;
  return s.irrep >= 0 && s.irrep < g;
}

// src_ed/symmetry_group.cpp:540-558.  The Green function is built here in the
// full target sector (N +- 1, S +- 1) whatever the representation of the
// state: the target carries no irrep label (0).
Sector shift_sector(const Sector &s0, int pm, int spin) {
  Sector s = s0;
  s.irrep = 0;
  if (s.N == Sector::odd) s.N = Sector::even;
  else if (s.N == Sector::even) s.N = Sector::odd;
  else s.N += pm;
  if (s.S == Sector::odd) s.S = Sector::even;
  else if (s.S == Sector::even) s.S = Sector::odd;
  else s.S += spin ? -pm : pm;
  return s;
}

size_t sector_dim(int L, const Sector &s) {
  const int nu = s.Nup(), nd = s.Ndw();
  if (nu < 0 || nd < 0 || nu > L || nd > L) return 0;
  return (size_t)half_basis(L, nu)->dim * half_basis(L, nd)->dim;
}

// ED::new_model_instance + model_instance_base ctor (src_ed/qcm_ED.cpp:125-158,
// src_ed/model_instance_base.cpp:14-130), trivial point group only.
void new_instance(const std::string &model, const std::map<std::string, double> &values, const std::string &sectors,
                  int label) {
  Model &M = get_model(model);
  M.closed = true;
  auto I = std::make_unique<Instance>();
  I->label = label;
  I->model = &M;
  for (const auto &kv : values)
    if (M.ops.count(kv.first)) I->value[kv.first] = kv.second;
  if (I->value.empty())
    fail(ED_ERR_INVALID, "ED model instance (label = " + std::to_string(label) + ") for model " + model +
                             " has no parameters!");
  Params &P = params();
  char gf = P.getc("GF_method");
  if (gf == 'L' && P.getb("continued_fraction")) gf = 'F';
  I->gf_method = gf;
  I->format = P.getc("Hamiltonian_format");
  bool asym = false;
  for (const auto &kv : I->value) {
    const Operator &op = M.ops.at(kv.first);
    if (op.spin_flip && kv.second != 0.0)
      fail(ED_ERR_UNSUPPORTED, "operator " + op.name + " mixes spins: spin-flip mixing is not supported");
    if (op.spin_asym && kv.second != 0.0) asym = true;
  }
  I->mixing = asym ? 4 : 0;
  if (M.group.g > 1) {
    // every operator in use must commute with the group (the reference builds
    // its Hamiltonian in the symmetric basis on that assumption)
    const PointGroup &G = M.group;
    auto image = [&](int a, int k, int &sg) {
      const int spin = a / M.L, o = a % M.L;
      int o2 = o, ph = 0;
      if (o < G.np) o2 = G.e[k][o];
      else ph = G.e[k][o];
      sg = std::cos(2.0 * M_PI * (double)(ph % G.g) / G.g) < 0.0 ? -1 : 1;
      return o2 + spin * M.L;
    };
    for (const auto &kv : I->value) {
      if (kv.second == 0.0) continue;
      const Operator &op = M.ops.at(kv.first);
      std::map<std::pair<int, int>, double> el;
      for (const auto &x : op.el) el[{x.r, x.c}] += x.v;
      for (int k = 1; k < G.g; ++k)
        for (const auto &x : op.el) {
          int s1 = 1, s2 = 1;
          const int r2 = image(x.r, k, s1), c2 = image(x.c, k, s2);
          const double vx = el.at({x.r, x.c});
          const double want = vx * (op.kind == OP_ONE_BODY ? s1 * s2 : 1);
          auto it = el.find({r2, c2});
          if (it == el.end() && op.kind != OP_ONE_BODY) it = el.find({c2, r2});
          const double got = it == el.end() ? 0.0 : it->second;
          if (std::fabs(got - want) > 1e-10 * std::max(1.0, std::fabs(vx)))
            fail(ED_ERR_INVALID, "operator " + op.name + " is not invariant under the cluster's point group");
        }
    }
  }
  std::stringstream ss(sectors);
  std::string item;
  while (std::getline(ss, item, '/')) {
    if (item.empty()) continue;
    Sector s = Sector::parse(item);
    if (!sector_valid(M.L, s, M.group.g)) fail(ED_ERR_SECTOR, "sector " + item + " is not valid");
    if (s.N >= Sector::odd || s.S >= Sector::odd)
      fail(ED_ERR_UNSUPPORTED, "sector string " + s.name() + " defines a non conserved quantity, not supported");
    if (s.S && I->mixing == 0) I->mixing = 4;
    I->targets.insert(s);
  }
  if (I->targets.empty()) fail(ED_ERR_SECTOR, "no target sectors were specified for model " + model);
  std::lock_guard<std::mutex> l(g_inst_mu);
  g_instances[label] = std::move(I);
}

TermList Instance::terms() const {
  TermList T;
  T.L = model->L;
  for (const auto &kv : value) T.terms.push_back({&model->ops.at(kv.first), kv.second});
  return T;
}

bool Instance::complex_hs() const { return terms_complex(terms()); }

std::unique_ptr<Hamiltonian> Instance::take_gf_hamiltonian(const Sector &tgt) {
  std::future<std::unique_ptr<Hamiltonian>> f;
  {
    std::lock_guard<std::mutex> lock(gf_hams_mu);
    auto it = gf_hams.find(tgt);
    if (it == gf_hams.end()) return nullptr;
    f = std::move(it->second);
    gf_hams.erase(it);
  }
  return f.get();
}

void Instance::need_real(const char *what) const {
  if (complex_hs())
    fail(ED_ERR_UNSUPPORTED, std::string(what) + " in a complex Hilbert space is not supported (instance " +
                                 std::to_string(label) + ")");
}

std::string Instance::gs_string() const {
  std::map<Sector, double> ws;
  for (const auto &s : states) ws[s->sec] += s->weight;
  std::ostringstream o;
  bool first = true;
  for (const auto &x : ws) {
    if (!first) o << '/';
    first = false;
    o << x.first.name() << ':' << std::setprecision(3) << x.second;
  }
  return o.str();
}

// ---------------------------------------------------------------------------
// c_a / c^dagger_a on factorized vectors (src_ed/model.cpp:155-182,
// src_ed/ED_basis.cpp:79-110): gather over the target sector.
// ---------------------------------------------------------------------------
CdagMaps cdag_maps(int L, int pm, int a, const Sector &src) {
  CdagMaps C;
  C.spin = a >= L ? 1 : 0;
  const int orb = a - C.spin * L;
  C.tgt = shift_sector(src, pm, C.spin);
  auto BUs = half_basis(L, src.Nup()), BDs = half_basis(L, src.Ndw());
  auto BUt = half_basis(L, C.tgt.Nup()), BDt = half_basis(L, C.tgt.Ndw());
  C.du_t = BUt->dim;
  C.dd_t = BDt->dim;
  C.du_s = BUs->dim;
  C.dd_s = BDs->dim;
  const HalfBasis &Bt = C.spin ? *BDt : *BUt;   // changed spin, target
  const HalfBasis &Bs = C.spin ? *BDs : *BUs;   // changed spin, source
  C.map.assign(Bt.dim, -1);
  C.sgn.assign(Bt.dim, 1);
  const uint32_t m = 1u << orb;
  for (uint32_t t = 0; t < Bt.dim; ++t) {
    const uint32_t st = Bt.bin[t];
    uint32_t ss;
    if (pm == -1) {  // target lacks orb, source has it
      if (st & m) continue;
      ss = st | m;
    } else {  // target has orb, source lacks it
      if (!(st & m)) continue;
      ss = st & ~m;
    }
    const uint32_t k2 = Bs.index(ss);
    if (k2 == Bs.dim) continue;
    C.map[t] = (int32_t)k2;
    C.sgn[t] = (__builtin_popcount(st & (m - 1u)) & 1) ? -1 : 1;
  }
  if (C.spin) {
    C.upar.resize(BUt->dim);
    for (uint32_t u = 0; u < BUt->dim; ++u) C.upar[u] = (__builtin_popcount(BUt->bin[u]) & 1) ? -1 : 1;
  }
  return C;
}

Sector cdag_apply(int L, int pm, int a, const Sector &src, const double *x, double z, double *y) {
  const CdagMaps C = cdag_maps(L, pm, a, src);
  const int spin = C.spin;
  const std::vector<int32_t> &map = C.map;
  const std::vector<int8_t> &sgn = C.sgn, &upar = C.upar;
  const Sector tgt = C.tgt;
  DevBuf<int32_t> dmap;
  DevBuf<int8_t> dsgn, dupar;
  dmap.upload(map.empty() ? std::vector<int32_t>{-1} : map);
  dsgn.upload(sgn.empty() ? std::vector<int8_t>{1} : sgn);
  if (spin) dupar.upload(upar.empty() ? std::vector<int8_t>{1} : upar);
  CdagArgs A;
  A.spin = spin;
  A.du_t = C.du_t;
  A.dd_t = C.dd_t;
  A.du_s = C.du_s;
  A.map = dmap.p;
  A.sgn = dsgn.p;
  A.upar = spin ? dupar.p : nullptr;
  A.z = z;
  cdag_gather(A, x, y, stream());
  sync();
  return tgt;
}

// ---------------------------------------------------------------------------
// ground state (model_instance<T>::low_energy_states, src_ed/model_instance.hpp:245-347;
// Hamiltonian<T>::states, src_ed/Hamiltonian/Hamiltonian_base.hpp:115-221;
// compute_weights, src_ed/model_instance.hpp:356-417)
// ---------------------------------------------------------------------------
// one N(0, 1) distribution for the process, as the reference's normal_dis
// (src_ed/Hamiltonian/Lanczos.cpp:7): the polar method makes its draws in pairs
// and keeps the second, so after a vector of odd length the next vector starts
// with that kept draw (from the previous seed's stream)
static std::normal_distribution<double> g_normal(0.0, 1.0);
static std::mutex g_normal_mu;

void random_start(size_t n, long seed, std::vector<double> &x) {
  // random() of src_util/vector_num.cpp:114-119: default_random_engine(seed),
  // N(0,1) draws, scaled by 1 / ||x|| (normalize, src_util/vector_num.hpp:139-145)
  std::default_random_engine gen((unsigned)seed);
  x.resize(n);
  std::lock_guard<std::mutex> lock(g_normal_mu);
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) {
    x[i] = g_normal(gen);
    s += x[i] * x[i];
  }
  const double f = 1.0 / std::sqrt(s);
  for (auto &v : x) v *= f;
}

// Distinct complex states among eigenpairs of the split plan [[A, -B], [B, A]]
// (ascending; split vectors [x_r | x_i] of 2 n doubles, unit norm): each
// vector is orthogonalized, in the complex inner product
// <u|w> = u_r.w_r + u_i.w_i + i (u_r.w_i - u_i.w_r), against the states kept
// so far; i psi, the partner of a kept psi, leaves nothing and is dropped, a
// new state leaves a vector of norm near 1 and is kept (normalized), until k
// states are kept.
static void complex_distinct(size_t n, const std::vector<double> &ev2, std::vector<DevBuf<double>> &v2, int k,
                             std::vector<double> &ev, std::vector<DevBuf<double>> &vecs) {
  for (size_t i = 0; i < ev2.size() && (int)vecs.size() < k; ++i) {
    double *w = v2[i].p;
    for (const auto &u : vecs) {
      const double cr = ddot(n, u.p, w) + ddot(n, u.p + n, w + n);
      const double ci = ddot(n, u.p, w + n) - ddot(n, u.p + n, w);
      // w -= (cr + i ci) u
      const double *s1[2] = {u.p, u.p + n}, *s2[2] = {u.p + n, u.p};
      const double c1[2] = {-cr, ci}, c2[2] = {-cr, -ci};
      combine(n, w, 1.0, 2, s1, c1, nullptr, nullptr, nullptr, stream());
      combine(n, w + n, 1.0, 2, s2, c2, nullptr, nullptr, nullptr, stream());
    }
    const double nn = ddot(2 * n, w, w);
    if (!(nn > 0.5)) continue;
    scale(2 * n, w, 1.0 / std::sqrt(nn), stream());
    ev.push_back(ev2[i]);
    vecs.push_back(std::move(v2[i]));
  }
  if ((int)vecs.size() < k) fail(ED_ERR_NOT_CONVERGED, "Davidson: fewer distinct complex states than requested");
}

std::pair<double, std::string> Instance::ground_state_solve() {
  if (gs_solved) return {GS_energy, gs_string()};
  Params &P = params();
  const double temperature = P.getd("temperature");
  double max_gap = -std::log(P.getd("minimum_weight")) * temperature;
  if (max_gap < kMinGap) max_gap = kMinGap;
  const long seed = P.geti("seed");
  long n_states = P.geti("n_states");
  char method = P.getc("Ground_state_method");
  const double acc_dav = P.getd("accur_Davidson");
  TermList T = terms();
  const bool cx = terms_complex(T);
  if (cx && model->group.g > 1)
    fail(ED_ERR_UNSUPPORTED, "complex operators with a point group are not supported");
  states.clear();
  GS_energy = 1.0e12;
  for (const auto &s : targets) {
    check_cancel();
    // vectors of a complex Hilbert space are split [Re | Im]
    const size_t dim = sector_dim(model->L, s) * (cx ? 2 : 1);
    if (dim == 0) continue;
    auto H = make_hamiltonian(T, s, format, cx);
    // the band-Lanczos Green function of this sector's ground state needs the
    // N +- 1 sectors (up spin): their tables are built and uploaded on host
    // threads, each with its own stream, while the device runs the Lanczos
    if (gf_method == 'L' && targets.size() == 1 && !cx && model->group.g == 1 && P.geti("gf_streams") > 1) {
      std::lock_guard<std::mutex> lock(gf_hams_mu);
      gf_hams.clear();
      for (int pm = -1; pm <= 1; pm += 2) {
        const Sector tgt = shift_sector(s, pm, 0);
        // small sectors only: a large one's tables take longer than nothing
        // here, and its schedule timing would run beside the Lanczos
        if (!sector_valid(model->L, tgt) || sector_dim(model->L, tgt) == 0 ||
            sector_dim(model->L, tgt) > ((size_t)1 << 24))
          continue;
        const char fmt = format;
        gf_hams[tgt] = std::async(std::launch::async, [T, tgt, fmt]() {
          auto Ht = make_hamiltonian(T, tgt, fmt);
          Ht->prepare();
          sync();
          return Ht;
        });
      }
    }
    long ns = std::max<long>(1, std::min<long>(n_states, (long)dim));
    char meth = (method == 'L' && ns > 1) ? 'D' : method;
    // a point group: the sector's states are those of the full (N, S) sector in
    // representation s.irrep
    std::unique_ptr<GroupProjector> proj;
    if (model->group.g > 1) {
      proj = std::make_unique<GroupProjector>(model->group, s.irrep, s.Nup(), s.Ndw());
      if (meth == 'D')
        fail(ED_ERR_UNSUPPORTED, "several states per sector (Davidson) with a point group are not supported");
    }
    std::vector<double> ev;
    std::vector<DevBuf<double>> vecs;
    vlog("ground state: sector %s dim %zu, Hamiltonian ready", s.name().c_str(), dim);
    if (meth == 'D' && cx) {
      // each level of the split form appears twice (psi and i psi): twice
      // the states, then the distinct complex ones
      std::vector<double> ev2;
      std::vector<DevBuf<double>> v2;
      davidson(*H, 2 * (int)ns, acc_dav, ev2, v2);
      complex_distinct(dim / 2, ev2, v2, (int)ns, ev, vecs);
    } else if (meth == 'D') {
      davidson(*H, (int)ns, acc_dav, ev, vecs);
      if (verbose()) {
        std::ostringstream os;
        os << std::setprecision(15);
        for (double e : ev) os << " " << e;
        vlog("Davidson: %ld states, eigenvalues%s", ns, os.str().c_str());
      }
    } else if (meth == 'L') {
      DevBuf<double> d0(dim);
      if (P.getb("host_random_start")) {
        // the reference's own stream (libstdc++ default_random_engine)
        std::vector<double> x0;
        random_start(dim, seed, x0);
        d0.upload(x0);
      } else {
        random_normal(dim, (uint64_t)seed, d0.p, stream());
        scale(dim, d0.p, 1.0 / std::sqrt(ddot(dim, d0.p, d0.p)), stream());
      }
      if (proj) {
        DevBuf<double> t(dim);
        proj->apply(d0.p, t.p);
        const double nrm = std::sqrt(ddot(dim, t.p, t.p));
        if (!(nrm > 1e-8)) {
          vlog("sector %s holds no state of its representation", s.name().c_str());
          continue;
        }
        scale(dim, t.p, 1.0 / nrm, stream());
        d0 = std::move(t);
      }
      vlog("random start vector ready");
      vecs.emplace_back(dim);
      LanczosGS R = lanczos_gs(*H, d0.p, vecs[0].p, proj.get());
      ev.push_back(R.E0);
      vlog("Lanczos: %d iterations, E0 = %.15g", R.niter, R.E0);
    } else {
      fail(ED_ERR_INVALID, std::string("Unknown method in Ground_state_method: ") + meth);
    }
    if (ev[0] < GS_energy) GS_energy = ev[0];
    for (size_t i = 0; i < vecs.size(); ++i) {
      if (ev[i] - GS_energy > max_gap) continue;
      auto st = std::make_unique<State>();
      st->sec = s;
      st->energy = ev[i];
      // normalize the state (the reference keeps the Lanczos Ritz vector as is)
      const double nn = std::sqrt(ddot(dim, vecs[i].p, vecs[i].p));
      scale(dim, vecs[i].p, 1.0 / nn, stream());
      st->psi = std::move(vecs[i]);
      states.push_back(std::move(st));
    }
  }
  finalize_states();
  return {GS_energy, gs_string()};
}

// compute_weights (src_ed/model_instance.hpp:356-417) and the minimum-weight
// filter over the candidate states, then the instance's E0 and mixing
void Instance::finalize_states() {
  Params &P = params();
  const double temperature = P.getd("temperature");
  double max_gap = -std::log(P.getd("minimum_weight")) * temperature;
  if (max_gap < kMinGap) max_gap = kMinGap;
  if (states.empty()) fail(ED_ERR_SECTOR, "instance " + std::to_string(label) + " has no low-energy state!");
  GS_energy = 1.0e12;
  for (auto &x : states) GS_energy = std::min(GS_energy, x->energy);
  // compute_weights
  if (temperature < 1.0e-8) {
    double e0 = 1e12;
    for (auto &x : states) e0 = std::min(e0, x->energy);
    std::vector<std::unique_ptr<State>> keep;
    for (auto &x : states)
      if (x->energy - e0 <= kSmall) keep.push_back(std::move(x));
    states.swap(keep);
    for (auto &x : states) x->weight = 1.0 / states.size();
  } else {
    std::vector<std::unique_ptr<State>> keep;
    for (auto &x : states)
      if (x->energy - GS_energy <= max_gap) keep.push_back(std::move(x));
    states.swap(keep);
    double Z = 0.0;
    for (auto &x : states) {
      x->weight = std::exp(-(x->energy - GS_energy) / temperature);
      Z += x->weight;
    }
    for (auto &x : states) x->weight /= Z;
  }
  // minimum_weight filter and renormalization
  const double minw = P.getd("minimum_weight");
  {
    std::vector<std::unique_ptr<State>> keep;
    double tot = 0.0;
    for (auto &x : states)
      if (x->weight >= minw) { tot += x->weight; keep.push_back(std::move(x)); }
    states.swap(keep);
    for (auto &x : states) x->weight /= tot;
  }
  for (auto &x : states)
    if (x->sec.S && mixing == 0) mixing = 4;
  GS_energy = 1.0e18;
  E0 = 0.0;
  for (auto &x : states) {
    E0 += x->weight * x->energy;
    GS_energy = std::min(GS_energy, x->energy);
  }
  averages.clear();
  averages_done = false;
  gs_solved = true;
}

const std::vector<std::tuple<std::string, double, double>> &Instance::cluster_averages() {
  if (!gs_solved) ground_state_solve();
  if (averages_done) return averages;
  // <O> and <O^2> - <O>^2 per operator (Hermitian_operator.hpp:67-83)
  for (auto &x : states)
    if (!x->psi.p) fail(ED_ERR_INVALID, "instance " + std::to_string(label) + " holds no ground-state vector (read from a file)");
  averages.clear();
  const bool cx = complex_hs();
  for (const auto &kv : value) {
    double ave = 0.0, var = 0.0;
    TermList one;
    one.L = model->L;
    one.terms.push_back({&model->ops.at(kv.first), 1.0});
    for (auto &x : states) {
      // <psi|O|psi> is the quadratic form of the split operator (real for a
      // Hermitian O)
      auto Ho = make_hamiltonian(one, x->sec, format, cx);
      DevBuf<double> y(Ho->dim);
      const int nb = Ho->apply(x->psi.p, y.p, 1.0, 0.0, true);
      const double e = Ho->dot_result(nb);
      const double e2 = ddot(Ho->dim, y.p, y.p);
      ave += x->weight * e;
      var += x->weight * e2;
    }
    averages.emplace_back(kv.first, ave, var - ave * ave);
  }
  averages_done = true;
  return averages;
}

// ---------------------------------------------------------------------------
// Green-function build
// ---------------------------------------------------------------------------
static void streamline(int L, std::vector<double> &e, std::vector<double> &V) {
  // Q_matrix::streamline (src_ed/Q_matrix.hpp:181-223, :286-316)
  Params &P = params();
  const double wtol = P.getd("Qmatrix_wtol"), vtol = P.getd("Qmatrix_vtol");
  if (vtol < 1e-12) return;
  const int M = (int)e.size();
  std::vector<double> e2, v2;
  for (int i = 0; i < M; ++i) {
    int j;
    for (j = i + 1; j < M; ++j)
      if (std::fabs(e[j] - e[i]) > wtol) break;
    const int n = j - i;
    if (n == 1) {
      double s = 0.0;
      for (int a = 0; a < L; ++a) s += V[a + (size_t)i * L] * V[a + (size_t)i * L];
      if (!(s < vtol)) {
        e2.push_back(e[i]);
        v2.insert(v2.end(), V.begin() + (size_t)i * L, V.begin() + (size_t)(i + 1) * L);
      }
    } else {
      std::vector<double> QQ((size_t)L * L, 0.0), d(L), U((size_t)L * L);
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += V[a + (size_t)(i + k) * L] * V[b + (size_t)(i + k) * L];
          QQ[a + (size_t)b * L] = s;
        }
      symeig(L, QQ.data(), d.data(), U.data());
      for (int c = 0; c < L; ++c)
        if (d[c] > vtol) {
          e2.push_back(e[i]);
          const double sq = std::sqrt(d[c]);
          for (int a = 0; a < L; ++a) v2.push_back(U[a + (size_t)c * L] * sq);
        }
    }
    i += n - 1;
  }
  e.swap(e2);
  V.swap(v2);
}

static double qnorm_error(int L, const std::vector<double> &V) {
  // Q_matrix::check_norm (src_ed/Q_matrix.hpp:236-260)
  const size_t M = V.size() / std::max(L, 1);
  double tot = 0.0;
  for (int i = 0; i < L; ++i) {
    double z = 0.0;
    for (size_t k = 0; k < M; ++k) z += V[i + k * L] * V[i + k * L];
    tot += (z - 1.0) * (z - 1.0);
    for (int j = 0; j < i; ++j) {
      double zz = 0.0;
      for (size_t k = 0; k < M; ++k) zz += V[i + k * L] * V[j + k * L];
      tot += 2.0 * zz * zz;
    }
  }
  return tot;
}

// ---------------------------------------------------------------------------
// combined matrix continued fraction (combine_via_lanczos,
// src_ed/matrix_continued_fraction.hpp:443-505): block Lanczos, in the QR or
// the polar form, on the block-tridiagonal operator T_e (+) T_h of the electron
// and hole fractions (dimension (M_e + M_h) p, on the host), from the block
// whose level-0 components are the two weights; G = W^T F_0 W then gives
// G_e + G_h^T in one fraction.  Both fractions need the same block size, as
// there (a deflated start block leaves them separate here).
// ---------------------------------------------------------------------------
namespace {
// y = T x for the block-tridiagonal matrix of one fraction: diagonal A_j,
// T(j+1, j) = B_j, T(j, j+1) = B_j^T (p x p column-major floors)
void mcf_apply(const GFRep::MCF &m, const double *x, double *y) {
  const int p = m.p, pp = p * p;
  for (int j = 0; j < m.nfl; ++j) {
    double *yj = y + (size_t)j * p;
    const double *Aj = m.A.data() + (size_t)j * pp;
    for (int r = 0; r < p; ++r) {
      double t = 0.0;
      for (int c = 0; c < p; ++c) t += Aj[r + c * p] * x[(size_t)j * p + c];
      if (j > 0) {
        const double *Bm = m.B.data() + (size_t)(j - 1) * pp;
        for (int c = 0; c < p; ++c) t += Bm[r + c * p] * x[(size_t)(j - 1) * p + c];
      }
      if (j + 1 < m.nfl) {
        const double *Bj = m.B.data() + (size_t)j * pp;
        for (int c = 0; c < p; ++c) t += Bj[c + r * p] * x[(size_t)(j + 1) * p + c];
      }
      yj[r] = t;
    }
  }
}

double hdot(size_t n, const double *a, const double *b) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// the block Lanczos recurrence of block_lanczos() on a host operator
int host_block_lanczos(size_t N, const std::function<void(const double *, double *)> &op,
                       std::vector<std::vector<double>> q, int M0, bool qr, std::vector<double> &A,
                       std::vector<double> &B) {
  Params &Pm = params();
  const double accur_deflation = Pm.getd("accur_deflation");
  const double accur_bl = std::max(0.0, Pm.getd("accur_band_lanczos"));
  M0 = std::min(M0, (int)Pm.geti("max_iter_BL"));
  const int p = (int)q.size(), pp = p * p;
  A.clear();
  B.clear();
  for (int l = 0; l < p; ++l) {
    for (int k = 0; k < l; ++k) {
      const double z = hdot(N, q[k].data(), q[l].data());
      for (size_t i = 0; i < N; ++i) q[l][i] -= z * q[k][i];
    }
    const double nrm = std::sqrt(hdot(N, q[l].data(), q[l].data()));
    if (nrm < accur_deflation) fail(ED_ERR_INVALID, "combined MCF: starting block is rank-deficient");
    for (auto &v : q[l]) v /= nrm;
  }
  std::vector<std::vector<double>> Qp(p, std::vector<double>(N, 0.0)), Z(p, std::vector<double>(N));
  double ev_old = 1e12;
  int j;
  for (j = 0; j < M0; ++j) {
    for (int k = 0; k < p; ++k) op(q[k].data(), Z[k].data());
    std::vector<double> Aj(pp);
    for (int l = 0; l < p; ++l)
      for (int k = 0; k <= l; ++k) Aj[k + l * p] = Aj[l + k * p] = hdot(N, q[k].data(), Z[l].data());
    A.insert(A.end(), Aj.begin(), Aj.end());
    for (int l = 0; l < p; ++l)
      for (int k = 0; k < p; ++k) {
        const double c = Aj[k + l * p];
        const double d = j > 0 ? B[(size_t)(j - 1) * pp + l + k * p] : 0.0;
        for (size_t i = 0; i < N; ++i) Z[l][i] -= c * q[k][i] + d * Qp[k][i];
      }
    std::vector<double> Bj(pp, 0.0);
    bool singular = false;
    if (qr) {
      for (int l = 0; l < p && !singular; ++l) {
        for (int k = 0; k < l; ++k) {
          const double z = hdot(N, Z[k].data(), Z[l].data());
          Bj[k + l * p] = z;
          for (size_t i = 0; i < N; ++i) Z[l][i] -= z * Z[k][i];
        }
        const double nrm = std::sqrt(hdot(N, Z[l].data(), Z[l].data()));
        if (nrm < accur_deflation) {
          for (int ll = l; ll < p; ++ll) std::fill(Z[ll].begin(), Z[ll].end(), 0.0);
          singular = true;
          break;
        }
        Bj[l + l * p] = nrm;
        for (auto &v : Z[l]) v /= nrm;
      }
      B.insert(B.end(), Bj.begin(), Bj.end());
    } else {
      std::vector<double> G(pp), D(p), V(pp);
      for (int l = 0; l < p; ++l)
        for (int k = 0; k <= l; ++k) G[k + l * p] = G[l + k * p] = hdot(N, Z[k].data(), Z[l].data());
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
      std::vector<std::vector<double>> Qn(p, std::vector<double>(N, 0.0));
      for (int l = 0; l < p; ++l)
        for (int k = 0; k < p; ++k) {
          double t = 0.0;
          for (int m = 0; m < p; ++m) t += V[k + m * p] * V[l + m * p] / std::sqrt(D[m]);
          for (size_t i = 0; i < N; ++i) Qn[l][i] += t * Z[k][i];
        }
      Qp.swap(q);
      q.swap(Qn);
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
      conv = std::fabs(e0 - ev_old) < accur_bl;
      ev_old = e0;
    }
    if (conv || singular) { ++j; break; }
    if (qr) {
      Qp.swap(q);
      q.swap(Z);
      Z.assign(p, std::vector<double>(N, 0.0));
    }
  }
  return (int)(A.size() / pp);
}
}  // namespace

GFRep::MCF combine_mcf(const GFRep::MCF &e, const GFRep::MCF &h, int n, bool qr) {
  GFRep::MCF out;
  if (e.nfl == 0 || h.nfl == 0 || e.p != h.p) return out;
  const int p = e.p;
  const size_t N = (size_t)(e.nfl + h.nfl) * p, Ne = (size_t)e.nfl * p;
  // start block: column i of W_e at the electron part's level 0, of W_h at the hole part's
  std::vector<std::vector<double>> phi(n, std::vector<double>(N, 0.0));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < p; ++k) {
      phi[i][k] = e.W[k + (size_t)i * p];
      phi[i][Ne + k] = h.W[k + (size_t)i * p];
    }
  // weight: the upper-triangular factor of the start block (modified Gram-Schmidt)
  std::vector<double> W((size_t)n * n, 0.0);
  std::vector<std::vector<double>> q(phi);
  int nd = n;
  for (int l = 0; l < n; ++l) {
    for (int k = 0; k < l; ++k) {
      const double z = hdot(N, q[k].data(), q[l].data());
      W[k + (size_t)l * n] = z;
      for (size_t i = 0; i < N; ++i) q[l][i] -= z * q[k][i];
    }
    const double nrm = std::sqrt(hdot(N, q[l].data(), q[l].data()));
    if (nrm < params().getd("accur_deflation")) { nd = l; break; }
    W[l + (size_t)l * n] = nrm;
    for (auto &v : q[l]) v /= nrm;
  }
  if (nd < n) return out;  // dependent weights: the two fractions stay separate
  auto op = [&](const double *x, double *y) {
    mcf_apply(e, x, y);
    mcf_apply(h, x + Ne, y + Ne);
  };
  out.p = n;
  out.nfl = host_block_lanczos(N, op, phi, e.nfl + h.nfl, qr, out.A, out.B);
  out.W = W;
  return out;
}

// complex residues: Q_matrix::streamline and check_norm for HilbertField =
// Complex (src_ed/Q_matrix.hpp:181-223, :236-260, streamlineQ :286-316); within
// a group of poles closer than Qmatrix_wtol, sum_k V_k V_k^H is diagonalized
// and its eigenvectors above Qmatrix_vtol, scaled by the root of their
// eigenvalue, replace the group's residues
static void streamline_c(int L, std::vector<double> &e, std::vector<cplx> &V) {
  Params &P = params();
  const double wtol = P.getd("Qmatrix_wtol"), vtol = P.getd("Qmatrix_vtol");
  if (vtol < 1e-12) return;
  const int M = (int)e.size();
  std::vector<double> e2;
  std::vector<cplx> v2;
  for (int i = 0; i < M; ++i) {
    int j;
    for (j = i + 1; j < M; ++j)
      if (std::fabs(e[j] - e[i]) > wtol) break;
    const int n = j - i;
    if (n == 1) {
      double s = 0.0;
      for (int a = 0; a < L; ++a) s += std::norm(V[a + (size_t)i * L]);
      if (!(s < vtol)) {
        e2.push_back(e[i]);
        v2.insert(v2.end(), V.begin() + (size_t)i * L, V.begin() + (size_t)(i + 1) * L);
      }
    } else {
      std::vector<cplx> QQ((size_t)L * L, 0.0), U((size_t)L * L);
      std::vector<double> d(L);
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b) {
          cplx s = 0.0;
          for (int k = 0; k < n; ++k) s += V[a + (size_t)(i + k) * L] * std::conj(V[b + (size_t)(i + k) * L]);
          QQ[a + (size_t)b * L] = s;
        }
      hermeig(L, QQ.data(), d.data(), U.data());
      for (int c = 0; c < L; ++c)
        if (d[c] > vtol) {
          e2.push_back(e[i]);
          const double sq = std::sqrt(d[c]);
          for (int a = 0; a < L; ++a) v2.push_back(U[a + (size_t)c * L] * sq);
        }
    }
    i += n - 1;
  }
  e.swap(e2);
  V.swap(v2);
}

static double qnorm_error_c(int L, const std::vector<cplx> &V) {
  const size_t M = V.size() / std::max(L, 1);
  double tot = 0.0;
  for (int i = 0; i < L; ++i) {
    double z = 0.0;
    for (size_t k = 0; k < M; ++k) z += std::norm(V[i + k * L]);
    tot += (z - 1.0) * (z - 1.0);
    for (int j = 0; j < i; ++j) {
      cplx zz = 0.0;
      for (size_t k = 0; k < M; ++k) zz += V[i + k * L] * std::conj(V[j + k * L]);
      tot += 2.0 * std::norm(zz);
    }
  }
  return tot;
}

// Green function of a state of a complex Hilbert space (GF_method 'L' only;
// build_qmatrix, src_ed/model_instance.hpp:711-818, with HilbertField = Complex).
// On split vectors the plan is the real symmetric [[A, -B], [B, A]], which
// commutes with K: [x_r | x_i] -> [-x_i | x_r] (multiplication by i).  The band
// Lanczos starts from the 2 ns vectors [phi_r | phi_i] and K [phi_r | phi_i]
// (phi_a = c^{pm}_a |Omega>), so its Krylov space is closed under K and each
// level of H appears twice among its Ritz values.  For a Ritz vector w with
// projections W on the start vectors, the complex vector it represents gives
// <phi_a|n> = W_a + i W_{ns+a}; over the two copies of a level these sum to
// twice the complex residue, hence the factor 1/sqrt(2), and streamlining
// merges the copies.  As there, the electron part stores conj(<phi_a|n>) and
// the hole part <chi_a|m>, so that G(r, c) = sum_i V(c, i) conj(V(r, i)) / (z - e_i).
// Continued fractions of a state of a complex Hilbert space (build_cf,
// src_ed/model_instance.cpp:101-177): for o1 >= o2 the fraction of
// c_o1 (+ c_o2) |Omega>, and for o1 > o2 that of c_o2 -+ i c_o1 |Omega> (- for
// electrons, + for holes), each a LanczosGreen on the split plan from the
// split vector (multiplication by i is [x_r | x_i] -> [-x_i | x_r]).  On split
// vectors <w, (z - M)^-1 w> sums both copies of every level: it equals the
// complex <psi|(z - H)^-1|psi>, so each fraction is the reference's.
static void build_cf_complex(Instance &I, State &st, int spin) {
  const int L = I.model->L, ns = I.model->n_sites;
  const TermList T = I.terms();
  const size_t ds = sector_dim(L, st.sec);
  GFRep &G = st.gf[spin];
  G = GFRep();
  G.kind = 'F';
  G.n = ns;
  const int np = ns * (ns + 1) / 2, ni = ns * (ns - 1) / 2;
  G.ea.assign(np, {});
  G.eb.assign(np, {});
  G.ha.assign(np, {});
  G.hb.assign(np, {});
  G.eai.assign(ni, {});
  G.ebi.assign(ni, {});
  G.hai.assign(ni, {});
  G.hbi.assign(ni, {});
  for (int pm = -1; pm <= 1; pm += 2) {
    check_cancel();
    Sector tgt = shift_sector(st.sec, pm, spin);
    if (!sector_valid(L, tgt)) continue;
    const size_t dt = sector_dim(L, tgt);
    if (dt == 0) continue;
    auto H = make_hamiltonian(T, tgt, I.format, true);
    std::vector<DevBuf<double>> phi;
    for (int i = 0; i < ns; ++i) {
      phi.emplace_back(2 * dt);
      fill(2 * dt, phi.back().p, 0.0, stream());
      cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p, 1.0, phi.back().p);
      cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p + ds, 1.0, phi.back().p + dt);
    }
    DevBuf<double> psi(2 * dt);
    // the fraction of the normalized psi, transformed as the reference's
    // continued_fraction constructor does (src_ed/continued_fraction.cpp:33-42)
    auto fraction = [&](std::vector<double> &a, std::vector<double> &b) {
      const double nrm = ddot(2 * dt, psi.p, psi.p);
      if (!(nrm > 1e-14)) return;
      scale(2 * dt, psi.p, 1.0 / std::sqrt(nrm), stream());
      lanczos_green(*H, psi.p, a, b);
      for (auto &x : b) x *= x;
      b[0] = nrm * st.weight;
      if (pm == 1) for (auto &x : a) x -= st.energy;
      else for (auto &x : a) x = st.energy - x;
    };
    int q = 0;
    for (int o1 = 0; o1 < ns; ++o1)
      for (int o2 = 0; o2 <= o1; ++o2, ++q) {
        std::vector<const double *> src{phi[o1].p};
        std::vector<double> cf{1.0};
        if (o2 != o1) {
          src.push_back(phi[o2].p);
          cf.push_back(1.0);
        }
        combine(2 * dt, psi.p, 0.0, (int)src.size(), src.data(), cf.data(), nullptr, nullptr, nullptr, stream());
        fraction(pm == 1 ? G.ea[q] : G.ha[q], pm == 1 ? G.eb[q] : G.hb[q]);
        if (o2 == o1) continue;
        // c_o2 -+ i c_o1: [b_r + s a_i | b_i - s a_r] with s = +1 (electrons), -1 (holes)
        const double sg = pm == 1 ? 1.0 : -1.0;
        const double *pa = phi[o1].p, *pb = phi[o2].p;
        std::vector<const double *> s1{pb, pa + dt}, s2{pb + dt, pa};
        std::vector<double> c1{1.0, sg}, c2{1.0, -sg};
        combine(dt, psi.p, 0.0, 2, s1.data(), c1.data(), nullptr, nullptr, nullptr, stream());
        combine(dt, psi.p + dt, 0.0, 2, s2.data(), c2.data(), nullptr, nullptr, nullptr, stream());
        const int qi = o1 * (o1 - 1) / 2 + o2;
        fraction(pm == 1 ? G.eai[qi] : G.hai[qi], pm == 1 ? G.ebi[qi] : G.hbi[qi]);
      }
  }
}

static void build_gf_complex(Instance &I, State &st, int spin) {
  if (I.gf_method == 'F') {
    build_cf_complex(I, st, spin);
    return;
  }
  if (I.gf_method != 'L')
    fail(ED_ERR_UNSUPPORTED, std::string("GF_method ") + I.gf_method +
                                 " in a complex Hilbert space is not supported (use L or F)");
  const int L = I.model->L, ns = I.model->n_sites;
  const TermList T = I.terms();
  const size_t ds = sector_dim(L, st.sec);
  GFRep &G = st.gf[spin];
  G = GFRep();
  G.kind = 'L';
  G.n = ns;
  std::vector<double> ee, eh;
  std::vector<cplx> Ve, Vh;
  const double E = st.energy, r2 = 1.0 / std::sqrt(2.0);
  for (int pm = -1; pm <= 1; pm += 2) {
    check_cancel();
    Sector tgt = shift_sector(st.sec, pm, spin);
    if (!sector_valid(L, tgt)) continue;
    const size_t dt = sector_dim(L, tgt);
    if (dt == 0) continue;
    auto H = make_hamiltonian(T, tgt, I.format, true);
    std::vector<DevBuf<double>> phi;
    for (int i = 0; i < ns; ++i) {
      phi.emplace_back(2 * dt);
      fill(2 * dt, phi.back().p, 0.0, stream());
      cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p, 1.0, phi.back().p);
      cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p + ds, 1.0, phi.back().p + dt);
    }
    for (int i = 0; i < ns; ++i) {
      // K phi_i = [-phi_i,im | phi_i,re]
      phi.emplace_back(2 * dt);
      double *k = phi.back().p;
      const double *x = phi[i].p;
      ED_HIP(hipMemcpyAsync(k, x + dt, dt * sizeof(double), hipMemcpyDeviceToDevice, stream()));
      ED_HIP(hipMemcpyAsync(k + dt, x, dt * sizeof(double), hipMemcpyDeviceToDevice, stream()));
      scale(dt, k, -1.0, stream());
    }
    std::vector<const double *> ph;
    for (auto &b : phi) ph.push_back(b.p);
    vlog("GF %s (pm %+d, spin %d, complex): dim %zu, %d start vectors ready", tgt.name().c_str(), pm, spin, dt,
         2 * ns);
    BandLanczosRun run = band_lanczos_run(*H, ph);
    std::vector<double> e, W;
    run.finish(e, W);
    const int p = 2 * ns, M = (int)e.size();
    std::vector<cplx> V((size_t)ns * M);
    for (int n = 0; n < M; ++n)
      for (int a = 0; a < ns; ++a) {
        const double re = W[a + (size_t)n * p], im = W[ns + a + (size_t)n * p];
        V[a + (size_t)n * ns] = r2 * cplx(re, pm == 1 ? -im : im);
      }
    for (auto &x : e) x -= E;
    if (pm == -1)
      for (auto &x : e) x = -x;
    streamline_c(ns, e, V);
    vlog("band Lanczos (complex): %d Ritz values, %zu poles", M, e.size());
    if (pm == -1) { eh.swap(e); Vh.swap(V); }
    else { ee.swap(e); Ve.swap(V); }
  }
  std::vector<cplx> V = Vh;
  V.insert(V.end(), Ve.begin(), Ve.end());
  G.e = eh;
  G.e.insert(G.e.end(), ee.begin(), ee.end());
  const double err = qnorm_error_c(ns, V);
  if (err > params().getd("accur_Q_matrix"))
    fail(ED_ERR_QNORM, "Q-matrix does not satisfy unitary condition by " + std::to_string(err));
  const double sw = std::sqrt(st.weight);
  G.V.resize(V.size());
  G.Vi.resize(V.size());
  for (size_t k = 0; k < V.size(); ++k) {
    G.V[k] = sw * V[k].real();
    G.Vi[k] = sw * V[k].imag();
  }
}

static void build_gf(Instance &I, State &st, int spin) {
  if (I.complex_hs()) {
    build_gf_complex(I, st, spin);
    return;
  }
  const int L = I.model->L, ns = I.model->n_sites;
  const TermList T = I.terms();
  const size_t dim0 = sector_dim(L, st.sec);
  GFRep &G = st.gf[spin];
  G = GFRep();
  G.kind = I.gf_method;
  G.n = ns;
  const int np = ns * (ns + 1) / 2;
  if (G.kind == 'F') { G.ea.assign(np, {}); G.eb.assign(np, {}); G.ha.assign(np, {}); G.hb.assign(np, {}); }
  std::vector<double> eh, Vh, ee, Ve;  // Lehmann parts (hole, electron)
  // poles and residues of one part from its band-Lanczos run (host only):
  // shifted by -E, negated for holes, streamlined
  using Part = std::pair<std::vector<double>, std::vector<double>>;
  const double E = st.energy;
  auto lehmann = [E, ns](const BandLanczosRun &r, int pm) -> Part {
    Part out;
    r.finish(out.first, out.second);
    vlog("band Lanczos: %zu poles", out.first.size());
    for (auto &x : out.first) x -= E;
    if (pm == -1)
      for (auto &x : out.first) x = -x;
    streamline(ns, out.first, out.second);
    return out;
  };
  if (G.kind == 'L') {
    // band Lanczos of one part (pm = -1 holes, +1 electrons) from the ns
    // vectors c^{pm}_i |Omega>; an empty run when the target sector is
    // invalid or empty
    auto run_part = [&](int pm) -> BandLanczosRun {
      check_cancel();
      Sector tgt = shift_sector(st.sec, pm, spin);
      if (!sector_valid(L, tgt)) return BandLanczosRun{};
      const size_t dt = sector_dim(L, tgt);
      if (dt == 0) return BandLanczosRun{};
      vlog("GF %s: start", tgt.name().c_str());
      std::unique_ptr<Hamiltonian> H = spin == 0 ? I.take_gf_hamiltonian(tgt) : nullptr;
      if (!H) H = make_hamiltonian(T, tgt, I.format);
      std::vector<DevBuf<double>> phi;
      std::vector<const double *> ph;
      for (int i = 0; i < ns; ++i) {
        phi.emplace_back(dt);
        fill(dt, phi.back().p, 0.0, stream());
        cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p, 1.0, phi.back().p);
        ph.push_back(phi.back().p);
      }
      vlog("GF %s (pm %+d, spin %d): dim %zu, %d start vectors ready", tgt.name().c_str(), pm, spin, dt, ns);
      return band_lanczos_run(*H, ph);
    };
    Part h, e;
    if (params().geti("gf_streams") > 1) {
      // the two parts on two host threads, each queuing on its own stream
      // (the same runs as one after the other); Omega is complete before the
      // second stream reads it, and the worker's stream is idle before its
      // buffers and Omega may be released
      sync();
      std::future<Part> fe = std::async(std::launch::async, [&]() {
        Part p = lehmann(run_part(1), 1);
        sync();
        return p;
      });
      try {
        h = lehmann(run_part(-1), -1);
      } catch (...) {
        fe.wait();
        throw;
      }
      e = fe.get();
    } else {
      // the hole part's eigen-decomposition on a host thread while the
      // device runs the electron part
      std::future<Part> fh = std::async(std::launch::async, [lehmann, r = run_part(-1)]() { return lehmann(r, -1); });
      try {
        e = lehmann(run_part(1), 1);
      } catch (...) {
        fh.wait();
        throw;
      }
      h = fh.get();
    }
    eh.swap(h.first);
    Vh.swap(h.second);
    ee.swap(e.first);
    Ve.swap(e.second);
  }
  for (int pm = -1; pm <= 1 && G.kind != 'L'; pm += 2) {
    check_cancel();
    Sector tgt = shift_sector(st.sec, pm, spin);
    if (!sector_valid(L, tgt)) continue;
    const size_t dt = sector_dim(L, tgt);
    if (dt == 0) continue;
    vlog("GF %s: start", tgt.name().c_str());
    auto H = make_hamiltonian(T, tgt, I.format);
    // phi_i = c^{pm}_{i, spin} |Omega>
    std::vector<DevBuf<double>> phi;
    for (int i = 0; i < ns; ++i) {
      phi.emplace_back(dt);
      fill(dt, phi.back().p, 0.0, stream());
      cdag_apply(L, pm, i + spin * L, st.sec, st.psi.p, 1.0, phi.back().p);
    }
    (void)dim0;
    if (G.kind == 'F') {
      // build_cf (src_ed/model_instance.cpp:35-100)
      DevBuf<double> psi(dt);
      int idx = 0;
      for (int o1 = 0; o1 < ns; ++o1)
        for (int o2 = 0; o2 <= o1; ++o2, ++idx) {
          std::vector<const double *> src{phi[o1].p};
          std::vector<double> cf{1.0};
          if (o2 != o1) { src.push_back(phi[o2].p); cf.push_back(1.0); }
          // psi = phi_o1 (+ phi_o2), norm^2
          double *pp = partials_acquire();
          const int nb = combine(dt, psi.p, 0.0, (int)src.size(), src.data(), cf.data(), pp, nullptr, nullptr,
                                 stream());
          const double nrm = sum_partials(pp, nb);
          if (!(nrm > 1e-14)) continue;
          scale(dt, psi.p, 1.0 / std::sqrt(nrm), stream());
          std::vector<double> a, b;
          lanczos_green(*H, psi.p, a, b);
          for (auto &x : b) x *= x;
          b[0] = nrm * st.weight;
          if (pm == 1) for (auto &x : a) x -= st.energy;
          else for (auto &x : a) x = st.energy - x;
          if (pm == 1) { G.ea[idx] = a; G.eb[idx] = b; }
          else { G.ha[idx] = a; G.hb[idx] = b; }
        }
    } else if (G.kind == 'M') {
      // build_mcf (src_ed/model_instance.hpp:847-946), QR block Lanczos
      const int p = ns;
      std::vector<double> W((size_t)p * p, 0.0);
      int pa = p;
      for (int l = 0; l < p; ++l) {
        for (int k = 0; k < l; ++k) {
          const double z = ddot(dt, phi[k].p, phi[l].p);
          W[k + (size_t)l * p] = z;
          std::vector<const double *> s1{phi[k].p};
          std::vector<double> c1{-z};
          combine(dt, phi[l].p, 1.0, 1, s1.data(), c1.data(), nullptr, nullptr, nullptr, stream());
        }
        const double nrm = std::sqrt(ddot(dt, phi[l].p, phi[l].p));
        if (nrm < params().getd("accur_deflation")) { pa = l; break; }
        W[l + (size_t)l * p] = nrm;
        scale(dt, phi[l].p, 1.0 / nrm, stream());
      }
      if (pa == 0) continue;
      for (auto &w : W) w *= std::sqrt(st.weight);
      GFRep::MCF mc;
      mc.p = pa;
      mc.W.assign((size_t)pa * p, 0.0);
      for (int k = 0; k < pa; ++k)
        for (int l = 0; l < p; ++l) mc.W[k + (size_t)l * pa] = W[k + (size_t)l * p];
      std::vector<const double *> q;
      for (int k = 0; k < pa; ++k) q.push_back(phi[k].p);
      // at least one floor (14 p log(dt) is 0 for a one-state sector)
      const int M0 = std::max(1, (int)(14 * pa * std::log(1.0 * dt)));
      mc.nfl = block_lanczos(*H, q, M0, mc.A, mc.B, params().getb("block_Lanczos_QR"));
      const int pp = pa * pa;
      for (int j = 0; j < mc.nfl; ++j) {
        double *Aj = mc.A.data() + (size_t)j * pp;
        if (pm == 1) {
          for (int k = 0; k < pa; ++k) Aj[k + k * pa] -= st.energy;
        } else {
          for (int k = 0; k < pp; ++k) Aj[k] = -Aj[k];
          for (int k = 0; k < pa; ++k) Aj[k + k * pa] += st.energy;
        }
      }
      if (pm == 1) G.me = mc;
      else G.mh = mc;
      if (pm == 1 && params().getb("combine_mcf"))
        G.comb = combine_mcf(G.me, G.mh, ns, params().getb("block_Lanczos_QR"));
    } else {
      fail(ED_ERR_INVALID, std::string("GF_method must be L, F or M (got ") + G.kind + ")");
    }
  }
  if (G.kind == 'L') {
    G.e = eh;
    G.e.insert(G.e.end(), ee.begin(), ee.end());
    G.V = Vh;
    G.V.insert(G.V.end(), Ve.begin(), Ve.end());
    const double err = qnorm_error(ns, G.V);
    if (err > params().getd("accur_Q_matrix"))
      fail(ED_ERR_QNORM, "Q-matrix does not satisfy unitary condition by " + std::to_string(err));
    const double sw = std::sqrt(st.weight);
    for (auto &v : G.V) v *= sw;
  }
}

void Instance::green_function_solve() {
  if (gf_solved) return;
  vlog("Green function solve: instance %d", label);
  if (!gs_solved) ground_state_solve();
  for (auto &s : states) {
    build_gf(*this, *s, 0);
    if (mixing & 4) build_gf(*this, *s, 1);
  }
  gf_solved = true;
}

// ---------------------------------------------------------------------------
// continued fractions out of / into an instance (the learned surrogate's
// targets and predictions)
// ---------------------------------------------------------------------------
void cf_export(Instance &I, int spin, int state, std::vector<std::vector<double>> &a,
               std::vector<std::vector<double>> &b) {
  I.need_real("continued-fraction export");
  if (!I.gf_solved) I.green_function_solve();
  if (state < 0 || state >= (int)I.states.size()) fail(ED_ERR_INVALID, "cf_export: no such low-energy state");
  const GFRep &G = I.states[state]->gf[I.gf_spin(spin == 1)];
  if (G.kind != 'F') fail(ED_ERR_UNSUPPORTED, "continued fractions exist for GF_method 'F' only");
  const int np = G.n * (G.n + 1) / 2;
  a.assign(2 * np, {});
  b.assign(2 * np, {});
  if (G.store) {
    const CfStore &st = *G.store;
    sync();  // the store is written on this thread's stream
    std::vector<int> off(st.nfrac + 1);
    ED_HIP(hipMemcpy(off.data(), st.off.p, off.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int f = 0; f < 2 * np; ++f) {
      const int g = G.frac0 + f;
      const int n = off[g + 1] - off[g];
      a[f].resize(n);
      b[f].resize(n);
      if (n) {
        ED_HIP(hipMemcpy(a[f].data(), st.A.p + off[g], n * sizeof(double), hipMemcpyDeviceToHost));
        ED_HIP(hipMemcpy(b[f].data(), st.B.p + off[g], n * sizeof(double), hipMemcpyDeviceToHost));
      }
    }
    return;
  }
  for (int q = 0; q < np; ++q)
    for (int eh = 0; eh < 2; ++eh) {
      const auto &av = eh ? G.ha[q] : G.ea[q];
      const auto &bv = eh ? G.hb[q] : G.eb[q];
      a[2 * q + eh] = av;
      b[2 * q + eh].assign(bv.begin(), bv.begin() + std::min(bv.size(), av.size()));
    }
}

void cf_pairs(Instance &I, int spin, const std::vector<std::pair<int, int>> &pairs, int cap,
              std::vector<std::vector<double>> &a, std::vector<std::vector<double>> &b) {
  I.need_real("continued fractions of orbital pairs");
  if (!I.gs_solved) I.ground_state_solve();
  if (I.states.empty() || !I.states[0]->psi.p) fail(ED_ERR_INVALID, "cf_pairs needs the ground-state vector");
  const int L = I.model->L, ns = I.model->n_sites;
  for (const auto &pr : pairs)
    if (pr.first < 0 || pr.first >= ns || pr.second < 0 || pr.second >= ns)
      fail(ED_ERR_INVALID, "cf_pairs: orbital outside the cluster");
  const int sp = I.gf_spin(spin == 1);
  State &st = *I.states[0];
  const TermList T = I.terms();
  a.assign(2 * pairs.size(), {});
  b.assign(2 * pairs.size(), {});
  for (int pm = -1; pm <= 1; pm += 2) {
    check_cancel();
    const Sector tgt = shift_sector(st.sec, pm, sp);
    if (!sector_valid(L, tgt)) continue;
    const size_t dt = sector_dim(L, tgt);
    if (dt == 0) continue;
    auto H = make_hamiltonian(T, tgt, I.format);
    std::map<int, DevBuf<double>> phi;
    for (const auto &pr : pairs)
      for (int o : {pr.first, pr.second})
        if (!phi.count(o)) {
          DevBuf<double> v(dt);
          fill(dt, v.p, 0.0, stream());
          cdag_apply(L, pm, o + sp * L, st.sec, st.psi.p, 1.0, v.p);
          phi.emplace(o, std::move(v));
        }
    DevBuf<double> psi(dt);
    for (size_t k = 0; k < pairs.size(); ++k) {
      const int o1 = pairs[k].first, o2 = pairs[k].second;
      std::vector<const double *> src{phi.at(o1).p};
      std::vector<double> cf{1.0};
      if (o2 != o1) { src.push_back(phi.at(o2).p); cf.push_back(1.0); }
      double *pp = partials_acquire();
      const int nb = combine(dt, psi.p, 0.0, (int)src.size(), src.data(), cf.data(), pp, nullptr, nullptr, stream());
      const double nrm = sum_partials(pp, nb);
      if (!(nrm > 1e-14)) continue;
      scale(dt, psi.p, 1.0 / std::sqrt(nrm), stream());
      std::vector<double> av, bv;
      lanczos_green(*H, psi.p, av, bv, cap);
      for (auto &x : bv) x *= x;
      bv[0] = nrm * st.weight;
      if (pm == 1) for (auto &x : av) x -= st.energy;
      else for (auto &x : av) x = st.energy - x;
      bv.resize(av.size());
      const int f = 2 * (int)k + (pm == 1 ? 0 : 1);
      a[f] = std::move(av);
      b[f] = std::move(bv);
    }
  }
}

void cf_import(Instance &I, double E0, const Sector &sec, int spin, const std::vector<std::vector<double>> &a,
               const std::vector<std::vector<double>> &b) {
  I.need_real("continued-fraction import");
  const int n = I.gf_dim(), np = n * (n + 1) / 2;
  if ((int)a.size() != 2 * np || (int)b.size() != 2 * np)
    fail(ED_ERR_INVALID, "cf_import: expected " + std::to_string(2 * np) + " fractions");
  if (!sector_valid(I.model->L, sec)) fail(ED_ERR_SECTOR, "cf_import: invalid sector " + sec.name());
  const int sp = I.gf_spin(spin == 1);
  const bool fresh = I.states.size() != 1 || I.states[0]->sec.name() != sec.name() || I.states[0]->energy != E0 ||
                     I.gf_method != 'F' || !I.gf_solved;
  if (fresh) {
    auto st = std::make_unique<State>();
    st->sec = sec;
    st->energy = E0;
    st->weight = 1.0;
    I.states.clear();
    I.states.push_back(std::move(st));
  }
  GFRep G;
  G.kind = 'F';
  G.n = n;
  G.ea.assign(np, {});
  G.eb.assign(np, {});
  G.ha.assign(np, {});
  G.hb.assign(np, {});
  for (int q = 0; q < np; ++q)
    for (int eh = 0; eh < 2; ++eh) {
      const auto &av = a[2 * q + eh], &bv = b[2 * q + eh];
      if (av.size() != bv.size()) fail(ED_ERR_INVALID, "cf_import: a and b of one fraction differ in length");
      (eh ? G.ha[q] : G.ea[q]) = av;
      (eh ? G.hb[q] : G.eb[q]) = bv;
    }
  I.states[0]->gf[sp] = std::move(G);
  I.gf_method = 'F';
  I.GS_energy = E0;
  I.E0 = E0;
  I.averages.clear();
  I.averages_done = false;
  I.gs_solved = true;
  I.gf_solved = true;
}

// ---------------------------------------------------------------------------
// Green-function evaluation on frequency grids (device kernels)
// ---------------------------------------------------------------------------
// continued_fraction_set::Green_function (src_ed/continued_fraction_set.cpp:63-98):
// h holds the values of ncf fractions per frequency (fraction-major per z);
// the representation's fractions start at frac0, two per orbital pair
// (electron, hole), pairs o1 >= o2 in order.  G_aa = F_aa, G_ab =
// (F_ab - G_aa - G_bb) / 2, added to acc (nz blocks n x n, column-major).
// the fractions of one representation, appended to a packed list
static void pack_cf(const GFRep &G, int n, std::vector<int> &off, std::vector<double> &A, std::vector<double> &B) {
  const int np = n * (n + 1) / 2;
  auto put = [&](const std::vector<double> &a, const std::vector<double> &b) {
    A.insert(A.end(), a.begin(), a.end());
    B.insert(B.end(), b.begin(), b.begin() + a.size());
    off.push_back((int)A.size());
  };
  for (int q = 0; q < np; ++q) {
    put(G.ea[q], G.eb[q]);
    put(G.ha[q], G.hb[q]);
  }
  // complex Hilbert space: the pairs' imaginary combinations follow
  for (size_t q = 0; q < G.eai.size(); ++q) {
    put(G.eai[q], G.ebi[q]);
    put(G.hai[q], G.hbi[q]);
  }
}

// continued-fraction representations summed into one G block of n x n per
// frequency, at offset `at` (doubles) of the output; cx: the representations
// hold the imaginary combinations too (complex Hilbert space)
struct CfGroup {
  std::vector<const GFRep *> reps;
  int n = 0;
  size_t at = 0;
  int cx = 0;
};

// the groups' fractions evaluated in one launch and assembled on the device
// in a second one, into dG (device memory); host-held fractions are packed
// and uploaded first, store-held ones are evaluated where they lie (one
// evaluation per store over the range of fractions the groups use)
static void eval_cf_groups(const std::vector<CfGroup> &groups, int nz, const double *dz, double *dG) {
  std::map<const CfStore *, std::vector<const CfGroup *>> by_store;
  for (const CfGroup &g : groups) {
    const CfStore *st = g.reps.empty() ? nullptr : g.reps[0]->store.get();
    for (const GFRep *r : g.reps)
      if (r->store.get() != st) fail(ED_ERR_INTERNAL, "continued fractions of one Green function held apart");
    by_store[st].push_back(&g);
  }
  for (auto &kv : by_store) {
    const CfStore *st = kv.first;
    std::vector<int> nn, pbeg{0}, frac0, cxg;
    std::vector<uint64_t> at;
    DevBuf<int> doff;
    DevBuf<double> dA, dB;
    const int *off = nullptr;
    const double *A = nullptr, *B = nullptr;
    int ncf = 0;
    if (st) {
      int fb = st->nfrac, fe = 0;
      for (const CfGroup *g : kv.second)
        for (const GFRep *r : g->reps) {
          fb = std::min(fb, r->frac0);
          fe = std::max(fe, r->frac0 + g->n * (g->n + 1));
        }
      if (fe <= fb) fb = fe = 0;
      for (const CfGroup *g : kv.second) {
        for (const GFRep *r : g->reps) frac0.push_back(r->frac0 - fb);
        pbeg.push_back((int)frac0.size());
        nn.push_back(g->n);
        at.push_back(g->at);
        cxg.push_back(0);
      }
      ncf = fe - fb;
      off = st->off.p + fb;
      A = st->A.p;
      B = st->B.p;
    } else {
      size_t na = 0, nf = 0;
      for (const CfGroup *g : kv.second)
        for (const GFRep *r : g->reps)
        {
          for (size_t q = 0; q < r->ea.size(); ++q) {
            na += r->ea[q].size() + r->ha[q].size();
            nf += 2;
          }
          for (size_t q = 0; q < r->eai.size(); ++q) {
            na += r->eai[q].size() + r->hai[q].size();
            nf += 2;
          }
        }
      std::vector<int> hoff;
      std::vector<double> hA, hB;
      hoff.reserve(nf + 1);
      hoff.push_back(0);
      hA.reserve(na);
      hB.reserve(na);
      for (const CfGroup *g : kv.second) {
        for (const GFRep *r : g->reps) {
          frac0.push_back((int)hoff.size() - 1);
          pack_cf(*r, g->n, hoff, hA, hB);
        }
        pbeg.push_back((int)frac0.size());
        nn.push_back(g->n);
        at.push_back(g->at);
        cxg.push_back(g->cx);
      }
      ncf = (int)hoff.size() - 1;
      doff.upload(hoff);
      dA.upload(hA);
      dB.upload(hB);
      off = doff.p;
      A = dA.p;
      B = dB.p;
    }
    DevBuf<int> dn, dpb, dfr, dcx;
    DevBuf<uint64_t> dat;
    DevBuf<double> dV((size_t)nz * ncf * 2);
    dn.upload(nn);
    dpb.upload(pbeg);
    dfr.upload(frac0);
    dat.upload(at);
    dcx.upload(cxg);
    cf_eval(ncf, off, A, B, nz, dz, dV.p, stream());
    cf_assemble((int)kv.second.size(), dn.p, dat.p, dpb.p, dfr.p, dcx.p, ncf, nz, dV.p, dG, stream());
  }
}

int Instance::gf_spin(bool spin_down) const {
  if (spin_down && !(mixing & 4) && mixing != 0)
    fail(ED_ERR_INVALID, "spin_down=True impossible with Hilbert space mixing " + std::to_string(mixing));
  return (spin_down && (mixing & 4)) ? 1 : 0;
}

void green_function_batch(const std::vector<Instance *> &insts, int nz, const double *z, bool spin_down,
                          double *out) {
  std::vector<size_t> base;
  size_t tot = 0;
  for (Instance *I : insts) {
    if (!I->gf_solved) I->green_function_solve();
    base.push_back(tot);
    tot += (size_t)nz * I->gf_dim() * I->gf_dim() * 2;
  }
  // instances whose states are all continued fractions: evaluated and
  // assembled together on the device, copied out once
  std::vector<CfGroup> groups;
  std::vector<size_t> other;
  for (size_t k = 0; k < insts.size(); ++k) {
    Instance &I = *insts[k];
    const int spin = I.gf_spin(spin_down);
    CfGroup g;
    g.n = I.gf_dim();
    g.at = base[k];
    bool all_cf = true;
    for (auto &s : I.states) {
      all_cf = all_cf && s->gf[spin].kind == 'F';
      if (!s->gf[spin].eai.empty()) g.cx = 1;
      g.reps.push_back(&s->gf[spin]);
    }
    if (all_cf) groups.push_back(std::move(g));
    else other.push_back(k);
  }
  if (!groups.empty()) {
    DevBuf<double> dz, dG(tot);
    dz.upload(z, (size_t)2 * nz);
    eval_cf_groups(groups, nz, dz.p, dG.p);
    vlog("gf batch: %zu instances evaluated", groups.size());
    dG.download(out, tot);
  }
  // the rest (written after the copy, which leaves their ranges undefined)
  for (size_t k : other) insts[k]->green_function(nz, z, spin_down, out + base[k]);
  vlog("gf batch: assembled");
}

void Instance::green_function(int nz, const double *z, bool spin_down, double *out) {
  if (!gf_solved) green_function_solve();
  const int spin = gf_spin(spin_down);
  const int n = gf_dim();
  std::vector<double> acc((size_t)nz * n * n * 2, 0.0);
  DevBuf<double> dz;
  dz.upload(z, (size_t)2 * nz);
  for (auto &s : states) {
    const GFRep &G = s->gf[spin];
    if (G.kind == 'L') {
      const int M = (int)G.e.size();
      if (!M) continue;
      DevBuf<double> de, dV, dVi, dG((size_t)nz * n * n * 2);
      de.upload(G.e);
      dV.upload(G.V);
      if (!G.Vi.empty()) dVi.upload(G.Vi);
      vdvh(n, M, de.p, dV.p, nz, dz.p, dG.p, stream(), G.Vi.empty() ? nullptr : dVi.p);
      std::vector<double> h = dG.to_host();
      for (size_t k = 0; k < h.size(); ++k) acc[k] += h[k];
    } else if (G.kind == 'F') {
      CfGroup g;
      g.n = n;
      g.cx = G.eai.empty() ? 0 : 1;
      g.reps.push_back(&G);
      DevBuf<double> dG((size_t)nz * n * n * 2);
      eval_cf_groups({g}, nz, dz.p, dG.p);
      std::vector<double> h = dG.to_host();
      for (size_t k = 0; k < h.size(); ++k) acc[k] += h[k];
    } else if (G.kind == 'M') {
      for (int eh = 0; eh < 3; ++eh) {
        // the combined fraction when it exists, otherwise electron and hole
        const GFRep::MCF &mc = eh == 2 ? G.comb : (eh ? G.mh : G.me);
        if (mc.nfl == 0 || (G.comb.nfl > 0) != (eh == 2)) continue;
        DevBuf<double> dA, dB, dW, dG((size_t)nz * n * n * 2);
        dA.upload(mc.A);
        dB.upload(mc.B);
        dW.upload(mc.W);
        mcf_eval(mc.p, mc.nfl, dA.p, dB.p, dW.p, n, nz, dz.p, dG.p, stream());
        std::vector<double> h = dG.to_host();
        for (size_t k = 0; k < h.size(); ++k) acc[k] += h[k];
      }
    }
  }
  std::memcpy(out, acc.data(), acc.size() * sizeof(double));
}

std::vector<double> Instance::one_body_matrix(bool spin_down) const {
  const int L = model->L, off = spin_down ? L : 0;
  std::vector<double> h((size_t)L * L, 0.0);
  for (const auto &kv : value) {
    const Operator &op = model->ops.at(kv.first);
    if (op.kind != OP_ONE_BODY) continue;
    for (const auto &x : op.el) {
      if ((x.r >= L) != spin_down || (x.c >= L) != spin_down) continue;
      // the operator holds both triangles (Hermitian completion at
      // new_operator, src_ed/Operators/one_body_operator.hpp:67-76)
      h[(size_t)(x.r - off) * L + (x.c - off)] += kv.second * x.v;
    }
  }
  return h;
}

std::vector<cplx> Instance::one_body_matrix_c(bool spin_down) const {
  const int L = model->L, off = spin_down ? L : 0;
  std::vector<cplx> h((size_t)L * L, 0.0);
  for (const auto &kv : value) {
    const Operator &op = model->ops.at(kv.first);
    if (op.kind != OP_ONE_BODY) continue;
    for (const auto &x : op.el) {
      if ((x.r >= L) != spin_down || (x.c >= L) != spin_down) continue;
      h[(size_t)(x.r - off) * L + (x.c - off)] += kv.second * cplx(x.v, x.vi);
    }
  }
  return h;
}

// bath part of ED::Potthoff_functional (model_instance::set_hopping_matrix,
// src_ed/model_instance.hpp:447-473): minus the negative eigenvalues of the
// bath one-body block, both spins (twice the up block without spin mixing)
double Instance::sef_bath() const {
  const int L = model->L, ns = model->n_sites, nb = model->n_bath;
  if (!nb) return 0.0;
  double s = 0.0;
  const bool cx = complex_hs();
  for (int spin = 0; spin < ((mixing & 4) ? 2 : 1); ++spin) {
    std::vector<double> eb(nb);
    if (cx) {
      const std::vector<cplx> h = one_body_matrix_c(spin == 1);
      std::vector<cplx> tb((size_t)nb * nb);
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
      hermeig(nb, tb.data(), eb.data(), nullptr);
    } else {
      const std::vector<double> h = one_body_matrix(spin == 1);
      std::vector<double> tb((size_t)nb * nb);
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
      symeig(nb, tb.data(), eb.data(), nullptr);
    }
    for (int i = 0; i < nb; ++i)
      if (eb[i] < 0.0) s -= eb[i];
  }
  return (mixing & 4) ? s : 2.0 * s;
}

// hybridization_function (src_ed/model_instance.hpp:1004-1018): bath block
// diagonalized, Gamma_ab = sum_i tcb(a,i) tcb(b,i) / (w - eps_i)
void Instance::hybridization(int nz, const double *z, bool spin_down, double *out) {
  const int L = model->L, ns = model->n_sites, nb = model->n_bath;
  const bool sd = spin_down && (mixing & 4);
  std::memset(out, 0, sizeof(double) * 2 * (size_t)nz * ns * ns);
  if (!nb) return;
  if (complex_hs()) {
    // Gamma = tcb (z - tb)^-1 tcb^H with tb = S diag(e) S^H (Hermitian):
    // Gamma_ab = sum_i T_ai conj(T_bi) / (z - e_i), T = tcb S
    const std::vector<cplx> h = one_body_matrix_c(sd);
    std::vector<cplx> tb((size_t)nb * nb), S((size_t)nb * nb), T((size_t)ns * nb, 0.0);
    std::vector<double> eb(nb);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
    hermeig(nb, tb.data(), eb.data(), S.data());
    for (int a = 0; a < ns; ++a)
      for (int i = 0; i < nb; ++i) {
        cplx t = 0.0;
        for (int k = 0; k < nb; ++k) t += h[(size_t)a * L + ns + k] * S[k + (size_t)i * nb];
        T[a + (size_t)i * ns] = t;
      }
    for (int iz = 0; iz < nz; ++iz) {
      const cplx w(z[2 * iz], z[2 * iz + 1]);
      double *g = out + (size_t)iz * ns * ns * 2;
      for (int i = 0; i < nb; ++i) {
        const cplx f = 1.0 / (w - eb[i]);
        for (int a = 0; a < ns; ++a)
          for (int b = 0; b < ns; ++b) {
            const cplx v = T[a + (size_t)i * ns] * std::conj(T[b + (size_t)i * ns]) * f;
            g[2 * (a + (size_t)b * ns)] += v.real();
            g[2 * (a + (size_t)b * ns) + 1] += v.imag();
          }
      }
    }
    return;
  }
  std::vector<double> h = one_body_matrix(sd);
  std::vector<double> tb((size_t)nb * nb), eb(nb), Sb((size_t)nb * nb);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
  symeig(nb, tb.data(), eb.data(), Sb.data());
  std::vector<double> tcb((size_t)ns * nb, 0.0);  // tcb' = tcb Sb
  for (int a = 0; a < ns; ++a)
    for (int i = 0; i < nb; ++i) {
      double s = 0.0;
      for (int k = 0; k < nb; ++k) s += h[(size_t)a * L + ns + k] * Sb[k + (size_t)i * nb];
      tcb[a + (size_t)i * ns] = s;
    }
  for (int iz = 0; iz < nz; ++iz) {
    const cplx w(z[2 * iz], z[2 * iz + 1]);
    double *g = out + (size_t)iz * ns * ns * 2;
    for (int i = 0; i < nb; ++i) {
      const cplx f = 1.0 / (w - eb[i]);
      for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b) {
          const cplx v = tcb[a + (size_t)i * ns] * tcb[b + (size_t)i * ns] * f;
          g[2 * (a + (size_t)b * ns)] += v.real();
          g[2 * (a + (size_t)b * ns) + 1] += v.imag();
        }
    }
  }
}

static void cinv_host(int n, std::vector<cplx> &a) {
  std::vector<cplx> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i + (size_t)i * n] = 1.0;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[r + (size_t)c * n]) > std::abs(a[pr + (size_t)c * n])) pr = r;
    if (std::abs(a[pr + (size_t)c * n]) == 0.0) fail(ED_ERR_INTERNAL, "singular Green function matrix");
    if (pr != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c + (size_t)k * n], a[pr + (size_t)k * n]);
        std::swap(inv[c + (size_t)k * n], inv[pr + (size_t)k * n]);
      }
    const cplx d = 1.0 / a[c + (size_t)c * n];
    for (int k = 0; k < n; ++k) { a[c + (size_t)k * n] *= d; inv[c + (size_t)k * n] *= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const cplx f = a[r + (size_t)c * n];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r + (size_t)k * n] -= f * a[c + (size_t)k * n];
        inv[r + (size_t)k * n] -= f * inv[c + (size_t)k * n];
      }
    }
  }
  a.swap(inv);
}

// self_energy (src_ed/model_instance.hpp:683-700): Sigma = z - tc - Gamma - G^{-1}
void Instance::self_energy(int nz, const double *z, bool spin_down, double *out) {
  const int n = gf_dim(), L = model->L;
  std::vector<double> G((size_t)nz * n * n * 2), Gam((size_t)nz * n * n * 2);
  green_function(nz, z, spin_down, G.data());
  hybridization(nz, z, spin_down, Gam.data());
  const bool sd = spin_down && (mixing & 4);
  const std::vector<cplx> h = one_body_matrix_c(sd);
  for (int iz = 0; iz < nz; ++iz) {
    std::vector<cplx> g((size_t)n * n);
    for (size_t k = 0; k < (size_t)n * n; ++k)
      g[k] = cplx(G[2 * ((size_t)iz * n * n + k)], G[2 * ((size_t)iz * n * n + k) + 1]);
    cinv_host(n, g);
    const cplx w(z[2 * iz], z[2 * iz + 1]);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        const size_t k = a + (size_t)b * n;
        cplx s = -g[k] - h[(size_t)a * L + b] - cplx(Gam[2 * ((size_t)iz * n * n + k)], Gam[2 * ((size_t)iz * n * n + k) + 1]);
        if (a == b) s += w;
        out[2 * ((size_t)iz * n * n + k)] = s.real();
        out[2 * ((size_t)iz * n * n + k) + 1] = s.imag();
      }
  }
}

Instance::Hopping Instance::hopping_blocks(bool spin_down, bool diag) const {
  const int L = model->L, ns = model->n_sites, nb = model->n_bath;
  if (complex_hs()) {
    // complex blocks; a non-diagonal bath block is diagonalized as a
    // Hermitian matrix (tcb -> tcb S)
    const std::vector<cplx> h = one_body_matrix_c(spin_down && (mixing & 4));
    std::vector<cplx> tc((size_t)ns * ns), tb((size_t)nb * nb), tcb((size_t)ns * nb);
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < ns; ++b) tc[a + (size_t)b * ns] = h[(size_t)a * L + b];
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
    for (int a = 0; a < ns; ++a)
      for (int i = 0; i < nb; ++i) tcb[a + (size_t)i * ns] = h[(size_t)a * L + ns + i];
    bool is_diag = true;
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        if (i != j && tb[i + (size_t)j * nb] != 0.0) is_diag = false;
    if (diag && nb > 0 && !is_diag) {
      std::vector<double> eb(nb);
      std::vector<cplx> S((size_t)nb * nb), t((size_t)ns * nb, 0.0);
      hermeig(nb, tb.data(), eb.data(), S.data());
      for (int a = 0; a < ns; ++a)
        for (int i = 0; i < nb; ++i) {
          cplx x = 0.0;
          for (int k = 0; k < nb; ++k) x += tcb[a + (size_t)k * ns] * S[k + (size_t)i * nb];
          t[a + (size_t)i * ns] = x;
        }
      tcb.swap(t);
      tb.assign((size_t)nb * nb, 0.0);
      for (int i = 0; i < nb; ++i) tb[i + (size_t)i * nb] = eb[i];
    }
    Hopping H;
    auto split = [](const std::vector<cplx> &m, std::vector<double> &re, std::vector<double> &im) {
      re.resize(m.size());
      im.resize(m.size());
      for (size_t k = 0; k < m.size(); ++k) {
        re[k] = m[k].real();
        im[k] = m[k].imag();
      }
    };
    split(tc, H.tc, H.tci);
    split(tb, H.tb, H.tbi);
    split(tcb, H.tcb, H.tcbi);
    return H;
  }
  const std::vector<double> h = one_body_matrix(spin_down && (mixing & 4));  // row-major L x L
  Hopping H;
  H.tc.assign((size_t)ns * ns, 0.0);
  H.tb.assign((size_t)nb * nb, 0.0);
  H.tcb.assign((size_t)ns * nb, 0.0);
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b) H.tc[a + (size_t)b * ns] = h[(size_t)a * L + b];
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) H.tb[i + (size_t)j * nb] = h[(size_t)(ns + i) * L + ns + j];
  for (int a = 0; a < ns; ++a)
    for (int i = 0; i < nb; ++i) H.tcb[a + (size_t)i * ns] = h[(size_t)a * L + ns + i];
  if (!diag || nb == 0) return H;
  bool is_diag = true;
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j)
      if (i != j && H.tb[i + (size_t)j * nb] != 0.0) is_diag = false;
  if (is_diag) return H;
  std::vector<double> eb(nb), S((size_t)nb * nb);
  symeig(nb, H.tb.data(), eb.data(), S.data());
  std::vector<double> t((size_t)ns * nb, 0.0);
  for (int a = 0; a < ns; ++a)
    for (int i = 0; i < nb; ++i) {
      double s = 0.0;
      for (int k = 0; k < nb; ++k) s += H.tcb[a + (size_t)k * ns] * S[k + (size_t)i * nb];
      t[a + (size_t)i * ns] = s;
    }
  H.tcb.swap(t);
  H.tb.assign((size_t)nb * nb, 0.0);
  for (int i = 0; i < nb; ++i) H.tb[i + (size_t)i * nb] = eb[i];
  return H;
}

void Instance::density_matrix(const std::vector<int> &sites, std::vector<double> &rho, std::vector<uint64_t> &basis) {
  need_real("the reduced density matrix");
  if (!gs_solved) ground_state_solve();
  if (states.empty() || !states[0]->psi.p)
    fail(ED_ERR_INVALID, "the density matrix needs the ground-state vector (not read from a file)");
  const int L = model->L;
  uint64_t mA = 0;
  for (int s : sites) {
    if (s < 0 || s >= L) fail(ED_ERR_INVALID, "density_matrix: orbital " + std::to_string(s) + " out of range");
    mA |= uint64_t(1) << s;
  }
  mA |= mA << 32;
  const State &st = *states[0];
  const auto BU = half_basis(L, st.sec.Nup()), BD = half_basis(L, st.sec.Ndw());
  const std::vector<double> psi = st.psi.to_host();
  // the basis in label order (i_up fastest): the words up | down << 32 ascend,
  // as the reference's mixed basis does for a trivial group
  std::map<uint64_t, int> ia, ib;
  std::vector<uint64_t> binA;
  std::vector<std::vector<std::pair<int, size_t>>> groups;  // per B: (A index, label)
  size_t I = 0;
  for (uint32_t d = 0; d < BD->dim; ++d)
    for (uint32_t u = 0; u < BU->dim; ++u, ++I) {
      const uint64_t w = (uint64_t)BU->bin[u] | ((uint64_t)BD->bin[d] << 32);
      const uint64_t A = w & mA, B = w ^ A;
      auto a = ia.find(A);
      if (a == ia.end()) {
        a = ia.emplace(A, (int)binA.size()).first;
        binA.push_back(A);
      }
      auto b = ib.find(B);
      if (b == ib.end()) {
        b = ib.emplace(B, (int)groups.size()).first;
        groups.emplace_back();
      }
      groups[b->second].push_back({a->second, I});
    }
  const size_t dA = binA.size();
  rho.assign(dA * dA, 0.0);
  for (const auto &g : groups)
    for (const auto &x : g)
      for (const auto &y : g) rho[x.first + (size_t)y.first * dA] += psi[x.second] * psi[y.second];
  // collapse (src_ed/binary_state.cpp:162-179)
  basis.assign(dA, 0);
  for (size_t k = 0; k < dA; ++k)
    for (size_t i = 0; i < sites.size(); ++i) {
      if (binA[k] >> sites[i] & 1u) basis[k] |= uint64_t(1) << i;
      if (binA[k] >> (sites[i] + 32) & 1u) basis[k] |= uint64_t(1) << (i + 32);
    }
}

// PrintBinaryDouble / BinaryToString (src_ed/binary_state.cpp:81-113): the low
// 32 bits of w, then the high 32 bits, m characters each, least significant first
static void print_binary_double(std::ostream &o, uint64_t w, int m) {
  o << '|';
  for (int i = 0; i < m; ++i) o << ((w >> i) & 1u ? '1' : '0');
  o << ',';
  for (int i = 0; i < m; ++i) o << ((w >> (32 + i)) & 1u ? '1' : '0');
  o << '>';
}

std::string Instance::print_wavefunction() {
  need_real("print_wavefunction");
  if (!gs_solved) ground_state_solve();
  const int L = model->L;
  const long max_print = params().geti("max_dim_print");
  std::ostringstream o;
  for (auto &s : states) {
    o << "state\n" << s->sec.name() << '\t' << s->energy << '\t' << s->weight << std::endl;
    const auto BU = half_basis(L, s->sec.Nup()), BD = half_basis(L, s->sec.Ndw());
    const size_t dim = (size_t)BU->dim * BD->dim;
    if ((long)dim > max_print) continue;
    if (!s->psi.p) fail(ED_ERR_INVALID, "instance " + std::to_string(label) + " holds no ground-state vector (read from a file)");
    const std::vector<double> psi = s->psi.to_host();
    for (size_t I = 0; I < dim; ++I) {
      const uint64_t up = BU->bin[I % BU->dim], dn = BD->bin[I / BU->dim];
      // the factorized basis ('F') prints its words with the down spins low
      // (ED_factorized_basis::bin, src_ed/ED_basis.cpp:157-160); the other
      // formats use the mixed basis, up spins low
      const uint64_t w = format == 'F' ? (dn | (up << 32)) : (up | (dn << 32));
      o << std::fabs(psi[I]) * std::fabs(psi[I]) << '\t' << psi[I] << '\t';
      print_binary_double(o, w, L);
      o << '\n';
    }
  }
  return o.str();
}

double Instance::fidelity(Instance &other) {
  need_real("fidelity");
  other.need_real("fidelity");
  if (!gs_solved) ground_state_solve();
  if (!other.gs_solved) other.ground_state_solve();
  double z = 0.0;
  for (auto &s1 : states)
    for (auto &s2 : other.states) {
      if (!(s1->sec == s2->sec)) continue;
      if (!s1->psi.p || !s2->psi.p) fail(ED_ERR_INVALID, "fidelity needs the ground-state vectors (not read from a file)");
      if (s1->psi.n != s2->psi.n || model->L != other.model->L)
        fail(ED_ERR_INVALID, "fidelity: the two instances belong to models of different sizes");
      const double zz = ddot(s1->psi.n, s1->psi.p, s2->psi.p);
      z += zz * zz * s1->weight * s2->weight;
    }
  return z;
}

// ---------------------------------------------------------------------------
// Frequency-integrated Green function, self-energy at infinity, response
// ---------------------------------------------------------------------------
// b_0 of the hole fraction of every orbital pair q (o1 >= o2, row by row):
// weight * <Omega| (c^dagger_o1 + c^dagger_o2)(c_o1 + c_o2) |Omega>, zero for a
// pair whose start vector vanished (no fraction was built)
static std::vector<double> hole_b0(const GFRep &R, int n) {
  const int np = n * (n + 1) / 2;
  std::vector<double> b0(np, 0.0);
  if (R.store) {
    const CfStore &S = *R.store;
    if (R.frac0 + 2 * np > S.nfrac) fail(ED_ERR_INTERNAL, "continued-fraction store shorter than its Green function");
    std::vector<int> off(2 * np + 1);
    d2h_pageable(off.data(), S.off.p + R.frac0, off.size() * sizeof(int));
    for (int q = 0; q < np; ++q) {
      const int f = 2 * q + 1;  // hole fraction of pair q
      if (off[f + 1] > off[f]) d2h_pageable(&b0[q], S.B.p + off[f], sizeof(double));
    }
  } else {
    for (int q = 0; q < np && q < (int)R.hb.size(); ++q)
      if (!R.hb[q].empty()) b0[q] = R.hb[q][0];
  }
  return b0;
}

// Green_function_average (src_ed/model_instance.hpp:641-652): the integrated
// Green function of each state added into one matrix, G(b, a) += ... as in
//  * Q_matrix::integrated_Green_function (src_ed/Q_matrix.hpp:155-168): the
//    poles below zero, sum_i v_a v_b (the weight is in v);
//  * continued_fraction_set::integrated_Green_function
//    (src_ed/continued_fraction_set.cpp:104-125): the hole fractions' b_0 on
//    the diagonal, and (b_0(o1, o2) - G_o1o1 - G_o2o2) / 2 off it, where G is the
//    running sum over the states so far (with several states this subtracts
//    the diagonals of the earlier states as well, as the reference does);
//  * mcf_set::integrated_Green_function (src_ed/mcf_set.hpp:146-158): the
//    hole weights' W_h^T W_h.
std::vector<cplx> Instance::gf_average(bool spin_down) {
  if (!gf_solved) green_function_solve();
  const int spin = gf_spin(spin_down);
  const int n = gf_dim();
  std::vector<cplx> G((size_t)n * n, 0.0);
  auto at = [&](int a, int b) -> cplx & { return G[a + (size_t)b * n]; };
  for (auto &s : states) {
    const GFRep &R = s->gf[spin];
    if (R.kind == 'L') {
      // G(b, a) += V(a, i) conj(V(b, i)) over the poles below zero
      for (size_t i = 0; i < R.e.size(); ++i) {
        if (R.e[i] >= 0.0) continue;
        const double *v = R.V.data() + i * n;
        const double *vi = R.Vi.empty() ? nullptr : R.Vi.data() + i * n;
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            at(b, a) += vi ? cplx(v[a], vi[a]) * cplx(v[b], -vi[b]) : cplx(v[a] * v[b], 0.0);
      }
    } else if (R.kind == 'F') {
      const std::vector<double> b0 = hole_b0(R, n);
      auto q = [](int o1, int o2) { return o1 * (o1 + 1) / 2 + o2; };
      for (int o1 = 0; o1 < n; ++o1) at(o1, o1) += b0[q(o1, o1)];
      for (int o1 = 0; o1 < n; ++o1)
        for (int o2 = 0; o2 < o1; ++o2) {
          const cplx t = 0.5 * (b0[q(o1, o2)] - at(o1, o1) - at(o2, o2));
          at(o1, o2) += t;
          at(o2, o1) += t;
        }
    } else if (R.kind == 'M') {
      const GFRep::MCF &mc = R.mh;
      if (mc.nfl == 0 || mc.p == 0) continue;
      // W is p x n column-major
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
          double t = 0.0;
          for (int k = 0; k < mc.p; ++k) t += mc.W[k + (size_t)a * mc.p] * mc.W[k + (size_t)b * mc.p];
          at(a, b) += t;
        }
    }
  }
  return G;
}

// Green_function_density (src_ed/model_instance.hpp:660-675): the trace of the
// average (both spin blocks) per site, twice the up block without spin mixing
double Instance::gf_density() {
  double nG = 0.0;
  const int n = gf_dim();
  for (int sd = 0; sd < ((mixing & 4) ? 2 : 1); ++sd) {
    const std::vector<cplx> G = gf_average(sd == 1);
    for (int i = 0; i < n; ++i) nG += G[i + (size_t)i * n].real();
  }
  nG /= model->n_sites;
  return mixing == 0 ? 2.0 * nG : nG;
}

// tr_sigma_inf (src_ed/model_instance.hpp:1278-1312): Tr Sigma(i / x) at x =
// 1e-5, 2e-5, 3e-5 (both spin blocks; twice the up block without spin mixing),
// extrapolated to x = 0 by the polynomial through the three points
// (polynomial_fit, src_ed/model_instance.cpp:184-222, Neville's scheme for
// that polynomial; written here in Lagrange form)
double Instance::tr_sigma_inf() {
  constexpr int nw = 3;
  const int n = gf_dim();
  double x[nw], z[2 * nw], S[nw] = {0.0, 0.0, 0.0};
  for (int i = 0; i < nw; ++i) {
    x[i] = (i + 1) * 1.0e-5;
    z[2 * i] = 0.0;
    z[2 * i + 1] = 1.0 / x[i];
  }
  std::vector<double> sig((size_t)nw * n * n * 2);
  for (int sd = 0; sd < ((mixing & 4) ? 2 : 1); ++sd) {
    self_energy(nw, z, sd == 1, sig.data());
    for (int i = 0; i < nw; ++i)
      for (int j = 0; j < n; ++j) S[i] += sig[2 * ((size_t)i * n * n + j + (size_t)j * n)];
  }
  if (mixing == 0)
    for (double &s : S) s *= 2.0;
  double y = 0.0;
  for (int i = 0; i < nw; ++i) {
    double l = 1.0;
    for (int j = 0; j < nw; ++j)
      if (j != i) l *= (0.0 - x[j]) / (x[i] - x[j]);
    y += S[i] * l;
  }
  return y;
}

// susceptibility / susceptibility_poles (src_ed/model_instance.hpp:1022-1058,
// :1066-1108): for each target sector (in sector order) and each low-energy
// state |Omega> in it, the band-Lanczos Lehmann representation (one start
// vector, build_Q_matrix, src_ed/Hamiltonian/Hamiltonian_base.hpp:229-254) of
// the Hamiltonian of that sector from O |Omega>, poles shifted by -E_Omega.
// O is the operator as defined (its instance value is not applied).
std::vector<Instance::ChiPart> Instance::susceptibility_parts(const std::string &op) {
  need_real("susceptibility");
  if (!model->ops.count(op)) fail(ED_ERR_INVALID, "Operator " + op + " is not defined in the model.");
  if (!value.count(op))
    fail(ED_ERR_INVALID, "operator " + op + " is not activated in instance " + std::to_string(label));
  if (!gs_solved) ground_state_solve();
  for (auto &s : states)
    if (!s->psi.p) fail(ED_ERR_INVALID, "instance " + std::to_string(label) + " holds no ground-state vector (read from a file)");
  TermList one;
  one.L = model->L;
  one.terms.push_back({&model->ops.at(op), 1.0});
  const TermList T = terms();
  std::vector<ChiPart> parts;
  for (const Sector &sec : targets) {
    bool any = false;
    for (auto &s : states) any = any || s->sec == sec;
    if (!any) continue;
    auto H = make_hamiltonian(T, sec, format);
    if (H->dim == 0) continue;
    auto Ho = make_hamiltonian(one, sec, format);
    for (auto &s : states) {
      if (!(s->sec == sec)) continue;
      DevBuf<double> b(H->dim);
      Ho->apply(s->psi.p, b.p, 1.0, 0.0, false);
      ChiPart c;
      c.weight = s->weight;
      std::vector<const double *> ph{b.p};
      band_lanczos(*H, ph, c.e, c.v);
      for (double &e : c.e) e -= s->energy;
      vlog("susceptibility of %s in %s: %zu poles", op.c_str(), sec.name().c_str(), c.e.size());
      parts.push_back(std::move(c));
    }
  }
  return parts;
}

}  // namespace ed
