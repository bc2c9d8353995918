// ed_model.cpp -- sectors, bases, model registry, Hamiltonian plans, host dense helpers.
#include "ed_model.hpp"
#include "ed_dense.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <limits>
#include <map>
#include <numeric>
#include <unordered_map>

namespace ed {

// ---------------------------------------------------------------------------
// Sector (src_ed/sector.hpp:54-106, ordering :150-163)
// ---------------------------------------------------------------------------
Sector Sector::parse(const std::string &str) {
  Sector s;
  bool valid = true;
  s.N = even;
  s.S = even;
  s.irrep = 0;
  auto p = str.find('N');
  if (p != std::string::npos && sscanf(str.c_str() + p, "N%d", &s.N) != 1) valid = false;
  p = str.find('S');
  if (p != std::string::npos && sscanf(str.c_str() + p, "S%d", &s.S) != 1) valid = false;
  p = str.find('R');
  if (p != std::string::npos && sscanf(str.c_str() + p, "R%d", &s.irrep) != 1) valid = false;
  if (!valid) fail(ED_ERR_SECTOR, "sector string " + str + " does not conform to standard!");
  if (str.find('O') != std::string::npos) s.N = s.S = odd;
  if (s.S == even && s.N != even && (s.N % 2)) s.S = odd;
  if (s.N == even && s.S != even && (s.S % 2)) s.N = odd;
  return s;
}

std::string Sector::name() const {
  std::string r;
  if (S == odd && N == odd) r += "O";
  r += "R" + std::to_string(irrep);
  if (N != even && N != odd) r += ":N" + std::to_string(N);
  if (S != even && S != odd) r += ":S" + std::to_string(S);
  return r;
}

int Sector::Nup() const {
  if (N == even || S == even) return even;
  if (N == odd || S == odd) return odd;
  return (N + S) / 2;
}
int Sector::Ndw() const {
  if (N == even || S == even) return even;
  if (N == odd || S == odd) return odd;
  return (N - S) / 2;
}
bool Sector::operator<(const Sector &y) const {
  if (S != y.S) return S < y.S;
  if (N != y.N) return N < y.N;
  return irrep < y.irrep;
}

// ---------------------------------------------------------------------------
// half basis
// ---------------------------------------------------------------------------
uint32_t HalfBasis::index(uint32_t b) const {
  auto it = std::lower_bound(bin.begin(), bin.end(), b);
  uint32_t I = (uint32_t)(it - bin.begin());
  if (I < dim && bin[I] == b) return I;
  return dim;
}

std::shared_ptr<const HalfBasis> half_basis(int L, int N) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, std::shared_ptr<const HalfBasis>> cache;
  std::lock_guard<std::mutex> l(mu);
  auto key = std::make_pair(L, N);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  auto B = std::make_shared<HalfBasis>();
  B->L = L;
  B->N = N;
  if (L >= 0 && L <= 32 && N >= 0 && N <= L) {
    // ascending integers with popcount N: Gosper's hack enumerates them in order
    if (N == 0) {
      B->bin.push_back(0u);
    } else {
      uint64_t v = (N == 64) ? ~0ull : ((1ull << N) - 1ull);
      const uint64_t lim = 1ull << L;
      while (v < lim) {
        B->bin.push_back((uint32_t)v);
        const uint64_t t = v | (v - 1);
        v = (t + 1) | (((~t & -~t) - 1) >> (__builtin_ctzll(v) + 1));
      }
    }
  }
  B->dim = (uint32_t)B->bin.size();
  cache[key] = B;
  return B;
}

// ---------------------------------------------------------------------------
// models and operators
// ---------------------------------------------------------------------------
namespace {
std::mutex g_model_mu;
std::map<std::string, std::unique_ptr<Model>> g_models;
}  // namespace

Model &get_model(const std::string &name) {
  std::lock_guard<std::mutex> l(g_model_mu);
  auto it = g_models.find(name);
  if (it == g_models.end()) fail(ED_ERR_NOT_FOUND, "The model " + name + " is not defined. Check spelling.");
  return *it->second;
}

// ED::new_model (src_ed/qcm_ED.cpp:43-59)
void new_model(const std::string &name, int n_sites, int n_bath, const std::vector<std::vector<int>> &generators,
               bool bath_irrep) {
  std::lock_guard<std::mutex> l(g_model_mu);
  if (g_models.count(name)) return;  // reference: warning, keeps the existing model
  if (n_bath < 0 || n_bath > 32) fail(ED_ERR_INVALID, "number of bath orbitals exceeds limits or negative!");
  if (n_sites < 1 || n_sites > 1024) fail(ED_ERR_INVALID, "number of cluster sites exceeds limits or negative!");
  if (n_sites + n_bath > 32) fail(ED_ERR_INVALID, "more than 32 orbitals per spin (ED_basis.hpp:23)");
  auto M = std::make_unique<Model>();
  M->name = name;
  M->n_sites = n_sites;
  M->n_bath = n_bath;
  M->L = n_sites + n_bath;
  M->group = PointGroup::build(n_sites, M->L, generators, bath_irrep);
  g_models[name] = std::move(M);
}

// ---------------------------------------------------------------------------
// point group (src_ed/symmetry_group.cpp:16-330, index.hpp)
// ---------------------------------------------------------------------------
PointGroup PointGroup::build(int n_sites, int L, const std::vector<std::vector<int>> &gens, bool bath_irrep) {
  PointGroup G;
  G.L = L;
  G.ns = n_sites;
  G.bath_irrep = bath_irrep && !gens.empty();
  G.np = G.bath_irrep ? n_sites : L;
  std::vector<std::vector<int>> gen = gens;
  if (gen.empty()) {
    std::vector<int> id(L, 0);
    for (int i = 0; i < G.np; ++i) id[i] = i;
    gen.push_back(id);
  }
  if (gen.size() > 8) fail(ED_ERR_INVALID, "Number of group generators exceeds the limit of 8");
  auto identity = [&]() {
    std::vector<int> v(L, 0);
    for (int i = 0; i < G.np; ++i) v[i] = i;
    return v;
  };
  auto is_identity_perm = [&](const std::vector<int> &v) {
    for (int i = 0; i < G.np; ++i)
      if (v[i] != i) return false;
    return true;
  };
  // product x.y: the permutation part composes, phases add (modulo g, once known)
  auto product = [&](const std::vector<int> &x, const std::vector<int> &y, int g) {
    std::vector<int> v(L, 0);
    for (int i = 0; i < G.np; ++i) v[i] = x[y[i]];
    for (int i = G.np; i < L; ++i) v[i] = g ? (x[i] + y[i]) % g : x[i] + y[i];
    return v;
  };
  for (size_t j = 0; j < gen.size(); ++j) {
    const auto &v = gen[j];
    if ((int)v.size() != L)
      fail(ED_ERR_INVALID, "generator " + std::to_string(j + 1) + " should have " + std::to_string(L) + " elements");
    std::vector<bool> seen(G.np, false);
    for (int i = 0; i < G.np; ++i) {
      if (v[i] < 0 || v[i] >= G.np || seen[v[i]])
        fail(ED_ERR_INVALID, "generator " + std::to_string(j + 1) + " is not valid!");
      seen[v[i]] = true;
    }
    for (int i = G.np; i < L; ++i)
      if (v[i] < 0) fail(ED_ERR_INVALID, "generator " + std::to_string(j + 1) + " is not valid!");
  }
  for (size_t i = 0; i < gen.size(); ++i)
    for (size_t j = 0; j < i; ++j) {
      const auto a = product(gen[i], gen[j], 0), b = product(gen[j], gen[i], 0);
      for (int k = 0; k < G.np; ++k)
        if (a[k] != b[k])
          fail(ED_ERR_INVALID, "generators " + std::to_string(i + 1) + " and " + std::to_string(j + 1) +
                                   " do not commute!");
    }
  G.g = 1;
  for (const auto &v : gen) {
    std::vector<int> p = identity();
    int n = 1;
    while (true) {
      p = product(p, v, 0);
      if (is_identity_perm(p)) break;
      if (++n > 64) fail(ED_ERR_INVALID, "generator of order above 64");
    }
    G.order.push_back(n);
    G.g *= n;
  }
  for (int n : G.order)
    if (n > 2)
      fail(ED_ERR_UNSUPPORTED, "point groups with complex characters (a generator of order " + std::to_string(n) +
                                   ") need a complex Hilbert space, which this solver does not provide");
  for (int i = 0; i < L; ++i)
    for (const auto &v : gen)
      if (i >= G.np && v[i] >= G.g) fail(ED_ERR_INVALID, "generator phase outside 0..g-1");
  const int ng = (int)gen.size();
  auto digits = [&](int k) {
    std::vector<int> d(ng);
    for (int j = 0; j < ng; ++j) {
      d[j] = k % G.order[j];
      k /= G.order[j];
    }
    return d;
  };
  for (int k = 0; k < G.g; ++k) {
    std::vector<int> p = identity();
    const auto d = digits(k);
    for (int j = 0; j < ng; ++j)
      for (int m = 0; m < d[j]; ++m) p = product(p, gen[j], G.g);
    if (k > 0 && is_identity_perm(p)) {
      bool trivial_phase = true;
      for (int i = G.np; i < L; ++i)
        if (p[i] != 0) trivial_phase = false;
      if (trivial_phase) fail(ED_ERR_INVALID, "The identity is generated as a non trivial element");
    }
    G.e.push_back(p);
  }
  G.chi.assign(G.g, std::vector<double>(G.g, 0.0));
  for (int r = 0; r < G.g; ++r) {
    const auto dr = digits(r);
    for (int k = 0; k < G.g; ++k) {
      const auto dk = digits(k);
      long ph = 0;
      for (int j = 0; j < ng; ++j) ph += (long)dr[j] * dk[j] * (G.g / G.order[j]);
      G.chi[r][k] = std::cos(2.0 * M_PI * (double)(ph % G.g) / G.g);
    }
  }
  // symmetric orbitals: per irrep, per orbit representative, sum_k chi_r(k) |e_k(site)>
  G.S.clear();
  G.irrep_dim.assign(G.g, 0);
  G.irrep_off.assign(G.g, 0);
  int nrow = 0;
  for (int r = 0; r < G.g; ++r) {
    G.irrep_off[r] = nrow;
    std::vector<bool> done(n_sites, false);
    for (int site = 0; site < n_sites; ++site) {
      if (done[site]) continue;
      std::vector<double> orb(n_sites, 0.0);
      for (int k = 0; k < G.g; ++k) {
        const int t = G.e[k][site] < n_sites ? G.e[k][site] : site;
        done[t] = true;
        orb[t] += G.chi[r][k];
      }
      double nn = 0.0;
      for (double x : orb) nn += x * x;
      if (nn < 1e-12) continue;
      for (double &x : orb) x /= std::sqrt(nn);
      G.S.insert(G.S.end(), orb.begin(), orb.end());
      ++G.irrep_dim[r];
      ++nrow;
    }
  }
  if (nrow != n_sites)
    fail(ED_ERR_INVALID, "Problem with construction of symmetric orbitals: " + std::to_string(nrow) +
                             " found instead of " + std::to_string(n_sites));
  return G;
}

GroupProjector::GroupProjector(const PointGroup &G, int irrep, int Nup, int Ndw) {
  if (irrep < 0 || irrep >= G.g) fail(ED_ERR_SECTOR, "irreducible representation outside the point group");
  if (G.g > kMaxGroup) fail(ED_ERR_UNSUPPORTED, "point group larger than 64 elements");
  auto BU = half_basis(G.L, Nup), BD = half_basis(G.L, Ndw);
  dim = (size_t)BU->dim * BD->dim;
  args.g = G.g;
  args.du = BU->dim;
  args.dd = BD->dim;
  for (int k = 0; k < G.g; ++k) args.c[k] = G.chi[irrep][k] / G.g;
  auto maps = [&](const HalfBasis &B, DevBuf<int32_t> &inv, DevBuf<int8_t> &sg) {
    std::vector<int32_t> iv((size_t)G.g * B.dim, -1);
    std::vector<int8_t> sv((size_t)G.g * B.dim, 0);
    for (int k = 0; k < G.g; ++k)
      for (uint32_t i = 0; i < B.dim; ++i) {
        int sign = 1;
        const uint32_t j = B.index(G.apply(k, B.bin[i], sign));
        if (j >= B.dim) fail(ED_ERR_INTERNAL, "point-group image outside the half basis");
        iv[(size_t)k * B.dim + j] = (int32_t)i;
        sv[(size_t)k * B.dim + j] = (int8_t)sign;
      }
    inv.upload(iv.empty() ? std::vector<int32_t>{0} : iv);
    sg.upload(sv.empty() ? std::vector<int8_t>{0} : sv);
  };
  maps(*BU, inv_up, sg_up);
  maps(*BD, inv_dn, sg_dn);
  args.inv_up = inv_up.p;
  args.inv_dn = inv_dn.p;
  args.sg_up = sg_up.p;
  args.sg_dn = sg_dn.p;
}

void GroupProjector::apply(const double *x, double *y) const { group_project(args, x, y, stream()); }

uint32_t PointGroup::apply(int k, uint32_t b, int &sign) const {
  const std::vector<int> &v = e[k];
  uint32_t out = 0;
  int inv = 0, ph = 0;
  int placed[32];
  int n = 0;
  for (int i = 0; i < L; ++i) {
    if (!((b >> i) & 1u)) continue;
    if (i < np) {
      out |= 1u << v[i];
      for (int m = 0; m < n; ++m)
        if (placed[m] > v[i]) ++inv;
      placed[n++] = v[i];
    } else {
      out |= 1u << i;
      ph += v[i];
    }
  }
  // exp(2 pi i ph / g) is real (+-1) for the groups solved here
  const double c = std::cos(2.0 * M_PI * (double)(ph % g) / g);
  sign = ((inv & 1) ? -1 : 1) * (c < 0.0 ? -1 : 1);
  return out;
}

// ED::new_operator (src_ed/qcm_ED.cpp:74-90, complex elements :95-106) with the
// element handling of the one_body_operator (one_body_operator.hpp:59-76,
// check_Hermiticity :179-188) and interaction_operator (interaction_operator.hpp:37-57)
// constructors.
void new_operator(const std::string &model, const std::string &name, const std::string &type,
                  const std::vector<Element> &el, bool is_complex) {
  if (el.empty()) return;
  Model &M = get_model(model);
  if (M.closed) return;  // reference: warning, model closed for modifications
  const int L = M.L;
  Operator op;
  op.name = name;
  op.is_complex = is_complex;
  for (const auto &x : el)
    if (x.r < 0 || x.c < 0 || x.r >= 2 * L || x.c >= 2 * L)
      fail(ED_ERR_INVALID, "element (" + std::to_string(x.r) + "," + std::to_string(x.c) + ") of operator " + name +
                               " is out of range for " + std::to_string(2 * L) + " spin-orbitals");
  if (is_complex && type != "one-body")
    fail(ED_ERR_UNSUPPORTED, "type of operator " + name + " (" + type +
                                 ") with complex elements is not supported by this solver");
  if (type == "one-body") {
    op.kind = OP_ONE_BODY;
    // the lower triangle is the conjugate of the upper one
    for (const auto &x : el) {
      if (x.r < x.c) {
        op.el.push_back({x.r, x.c, x.v, x.vi});
        op.el.push_back({x.c, x.r, x.v, -x.vi});
      } else if (x.r == x.c) {
        if (std::fabs(x.vi) > 1e-12) fail(ED_ERR_INVALID, "operator " + name + " is not Hermitian");
        op.el.push_back({x.r, x.r, x.v, 0.0});
      }
    }
    for (const auto &x : op.el)
      if ((x.r < L) != (x.c < L)) op.spin_flip = true;
  } else if (type == "interaction") {
    op.kind = OP_INTERACTION;
    for (const auto &x : el) op.el.push_back({std::min(x.r, x.c), std::max(x.r, x.c), x.v});
  } else {
    fail(ED_ERR_UNSUPPORTED, "type of operator " + name + " (" + type + ") is not supported by this solver");
  }
  // spin symmetry (one_body_operator.hpp:187-205): up and down parts equal?
  {
    std::map<std::pair<int, int>, cplx> m;
    for (const auto &x : op.el) m[{x.r, x.c}] += cplx(x.v, x.vi);
    for (const auto &kv : m) {
      int r = kv.first.first, c = kv.first.second;
      r = (r < L) ? r + L : r - L;
      c = (c < L) ? c + L : c - L;
      auto f = m.find({std::min(r, c), std::max(r, c)});
      auto g = m.find({r, c});
      const bool ok = (op.kind == OP_INTERACTION) ? (f != m.end() && f->second == kv.second)
                                                   : (g != m.end() && g->second == kv.second);
      if (!ok) { op.spin_asym = true; break; }
    }
  }
  // imaginary parts: an antisymmetric one-body operator of their own, applied
  // to the other half of the split [Re | Im] vectors
  bool any_im = false;
  for (const auto &x : op.el) any_im = any_im || x.vi != 0.0;
  if (any_im) {
    auto im = std::make_shared<Operator>();
    im->name = name + " (imaginary part)";
    im->kind = OP_ONE_BODY;
    im->spin_flip = op.spin_flip;
    im->spin_asym = op.spin_asym;
    for (const auto &x : op.el)
      if (x.vi != 0.0) im->el.push_back({x.r, x.c, x.vi});
    op.imag = im;
  }
  std::lock_guard<std::mutex> l(g_model_mu);
  if (!M.ops.count(name)) M.order.push_back(name);
  M.ops[name] = std::move(op);
}

// ---------------------------------------------------------------------------
// helpers on bit states (src_ed/binary_state.cpp:23-28, .hpp:42-47)
// ---------------------------------------------------------------------------
static inline int hop_sign(uint32_t s, int from, int to) {
  // s must contain 'from' and not 'to'; sign of moving the electron
  const int lo = std::min(from, to), hi = std::max(from, to);
  if (hi - lo <= 1) return 1;
  const uint32_t inside = (uint32_t)(((1ull << hi) - 1ull) & ~((2ull << lo) - 1ull));
  return (__builtin_popcount(s & inside) & 1) ? -1 : 1;
}

// Reductions.  A kernel that reduces writes one partial per block (and per
// row) into a slot of pinned, host-mapped memory.  Every slot entry holds a
// sentinel (a signalling-NaN bit pattern that FP64 arithmetic never
// produces) until the device overwrites it; the host polls until no entry of
// the rows it expects is the sentinel, sums each row in a fixed order
// (pairwise), restores the sentinels and releases the slot.  No fold kernel
// and no stream synchronisation is needed (poll_reductions; otherwise the
// stream is synchronised first).  Each 8-byte store is atomic, so no fence
// is needed on the device side.
namespace {
constexpr uint64_t kSentinel = 0x7ff4dead5eed0001ull;
constexpr int kSlots = 8;
struct PartialArena {
  size_t slot_len = 0;  // doubles per slot
  double *h = nullptr, *d = nullptr;
  bool busy[kSlots] = {};
  int next = 0;
  PartialArena() {
    slot_len = (size_t)partials_cap() * 16;
    ED_HIP(hipHostMalloc((void **)&h, kSlots * slot_len * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    ED_HIP(hipHostGetDevicePointer((void **)&d, h, 0));
    uint64_t *u = reinterpret_cast<uint64_t *>(h);
    for (size_t i = 0; i < kSlots * slot_len; ++i) u[i] = kSentinel;
  }
  int slot_of(const double *p) const {
    if (p < d || p >= d + kSlots * slot_len) return -1;
    return (int)((size_t)(p - d) / slot_len);
  }
};
// One arena per host thread (a thread polls only the reductions it queued on
// its own stream), taken from a process-wide pool on first use and handed
// back when the thread ends, so the pinned blocks are allocated once per
// concurrent thread, not once per thread.
std::mutex g_arena_mu;
std::vector<PartialArena *> *g_arena_pool = new std::vector<PartialArena *>;  // never destroyed
struct ThreadArena {
  PartialArena *a = nullptr;
  ~ThreadArena() {
    if (!a) return;
    bool dirty = false;
    for (int s = 0; s < kSlots; ++s) dirty = dirty || a->busy[s];
    // an arena with reductions abandoned by an error path may still be
    // written by queued kernels: it stays out of the pool
    if (dirty) return;
    std::lock_guard<std::mutex> l(g_arena_mu);
    g_arena_pool->push_back(a);
  }
};
thread_local ThreadArena t_arena;
PartialArena &arena() {
  if (!t_arena.a) {
    {
      std::lock_guard<std::mutex> l(g_arena_mu);
      if (!g_arena_pool->empty()) {
        t_arena.a = g_arena_pool->back();
        g_arena_pool->pop_back();
      }
    }
    if (!t_arena.a) t_arena.a = new PartialArena;
  }
  return *t_arena.a;
}
// slot of p in the calling thread's arena, -1 if p is elsewhere (or the
// thread has no arena yet: the lookup never allocates one)
int arena_slot_of(const double *p) { return t_arena.a ? t_arena.a->slot_of(p) : -1; }

// the device writes these entries while the host polls them: each read is an
// atomic acquire load, so the compiler re-reads memory on every poll and the
// row is read after the sentinel test that saw it complete
inline uint64_t load_entry(const double *p) {
  return __atomic_load_n(reinterpret_cast<const uint64_t *>(p), __ATOMIC_ACQUIRE);
}
inline bool is_sentinel(const double *p) { return load_entry(p) == kSentinel; }
double pairwise(std::vector<double> &w) {
  size_t n = w.size();
  while (n > 1) {
    const size_t half = (n + 1) / 2;
    for (size_t i = 0; i + half < n; ++i) w[i] += w[i + half];
    n = half;
  }
  return n ? w[0] : 0.0;
}
}  // namespace

double *partials_acquire() {
  PartialArena &A = arena();
  for (int t = 0; t < kSlots; ++t) {
    const int s = (A.next + t) % kSlots;
    if (!A.busy[s]) {
      // a slot is handed out clean: a leftover partial would be taken for a
      // new one (the first entries of the first two rows are where small
      // reductions write)
      const double *hs = A.h + (size_t)s * A.slot_len;
      for (size_t i = 0; i < 64; ++i)
        if (!is_sentinel(hs + i) || !is_sentinel(hs + partials_cap() + i))
          fail(ED_ERR_INTERNAL, "reduction slot " + std::to_string(s) + " holds a stale partial at entry " +
                                    std::to_string(is_sentinel(hs + i) ? partials_cap() + i : i));
      A.busy[s] = true;
      A.next = (s + 1) % kSlots;
      return A.d + (size_t)s * A.slot_len;
    }
  }
  // every slot taken: only possible after reductions abandoned by an error
  // path; once the stream is idle, restore all sentinels and start over
  ED_HIP(hipStreamSynchronize(stream()));
  uint64_t *u = reinterpret_cast<uint64_t *>(A.h);
  for (size_t i = 0; i < kSlots * A.slot_len; ++i) u[i] = kSentinel;
  for (int s = 0; s < kSlots; ++s) A.busy[s] = false;
  A.busy[0] = true;
  A.next = 1;
  return A.d;
}

double sum_partials_rows(const double *dpart, int nb, int m, size_t stride, std::vector<double> *out) {
  const int slot = arena_slot_of(dpart);
  if (out) out->assign(std::max(m, 0), 0.0);
  if (slot < 0) {
    // device partials: fold on the host after a copy
    if (m <= 0 || nb <= 0) return 0.0;
    std::vector<double> hbuf((size_t)(m - 1) * stride + nb);
    ED_HIP(hipMemcpyAsync(hbuf.data(), dpart, hbuf.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
    ED_HIP(hipStreamSynchronize(stream()));
    double first = 0.0;
    for (int r = 0; r < m; ++r) {
      std::vector<double> w(hbuf.begin() + (size_t)r * stride, hbuf.begin() + (size_t)r * stride + nb);
      const double s = pairwise(w);
      if (out) (*out)[r] = s;
      if (r == 0) first = s;
    }
    return first;
  }
  PartialArena &A = *t_arena.a;
  const double *hp = A.h + (dpart - A.d);
  if (m > 0 && nb > 0) {
    if ((size_t)(dpart - A.d) - (size_t)slot * A.slot_len + (size_t)(m - 1) * stride + nb > A.slot_len)
      fail(ED_ERR_INTERNAL, "reduction larger than a partials slot");
    bool synced = !params().getb("poll_reductions");
    if (synced) ED_HIP(hipStreamSynchronize(stream()));
    const auto t0 = std::chrono::steady_clock::now();
    unsigned spins = 0;
    for (int r = 0; r < m; ++r) {
      const double *row = hp + (size_t)r * stride;
      for (int i = 0; i < nb; ++i) {
        while (is_sentinel(row + i)) {
          if (synced) fail(ED_ERR_INTERNAL, "reduction partial missing after the stream completed");
          if ((++spins & 1023u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) {
            // long wait: let the runtime block (and report a failed stream)
            ED_HIP(hipStreamSynchronize(stream()));
            synced = true;
          }
        }
      }
    }
  }
  double first = 0.0;
  for (int r = 0; r < m; ++r) {
    double *row = const_cast<double *>(hp) + (size_t)r * stride;
    std::vector<double> w(std::max(nb, 0));
    for (int i = 0; i < nb; ++i) {
      const uint64_t u = load_entry(row + i);
      std::memcpy(&w[i], &u, 8);
    }
    const double s = pairwise(w);
    if (out) (*out)[r] = s;
    if (r == 0) first = s;
    uint64_t *u = reinterpret_cast<uint64_t *>(row);
    for (int i = 0; i < nb; ++i) u[i] = kSentinel;
  }
  A.busy[slot] = false;
  return first;
}

double sum_partials(const double *dpart, int nb, int m, std::vector<double> *out) {
  return sum_partials_rows(dpart, nb, m, (size_t)std::max(nb, 0), out);
}

Hamiltonian::~Hamiltonian() {
  if (pending) {
    try {
      sum_partials(pending, pending_nb);
    } catch (...) {
    }
  }
}

double *Hamiltonian::take_partials() {
  if (pending) sum_partials(pending, pending_nb);  // an unread result: wait for it and drop it
  pending = partials_acquire();
  pending_nb = 0;
  return pending;
}

double Hamiltonian::dot_result(int nb) {
  if (!pending) fail(ED_ERR_INTERNAL, "dot_result without a pending product");
  double *p = pending;
  pending = nullptr;
  return sum_partials(p, nb);
}

// ---------------------------------------------------------------------------
// per-spin merged symmetric one-body table + diagonal (host build)
// ---------------------------------------------------------------------------
struct HalfHost {
  uint32_t dim = 0;
  std::vector<uint32_t> ptr, ent;
  std::vector<double> val;
  std::vector<double> diag;
  int packed = 1, colbits = 1;
  uint32_t nval = 0;
};

static HalfHost build_half(const TermList &T, int spin, const HalfBasis &B) {
  const int L = T.L, off = spin ? L : 0;
  HalfHost H;
  H.dim = B.dim;
  H.diag.assign(B.dim, 0.0);
  std::vector<std::pair<uint64_t, double>> ent;  // key (row << 32 | col)
  for (const auto &tc : T.terms) {
    const Operator &op = *tc.first;
    const double coef = tc.second;
    if (op.kind == OP_ONE_BODY) {
      for (const auto &x : op.el) {
        if ((x.r >= L) != (spin == 1) || (x.c >= L) != (spin == 1)) continue;
        const int r = x.r - off, c = x.c - off;
        if (r == c) {
          const uint32_t m = 1u << r;
          for (uint32_t I = 0; I < B.dim; ++I)
            if (B.bin[I] & m) H.diag[I] += coef * x.v;
          continue;
        }
        // element v c+_r c_c: from state I (c occupied, r empty) to J = I with c -> r
        const uint32_t mc = 1u << c, mr = 1u << r;
        for (uint32_t I = 0; I < B.dim; ++I) {
          const uint32_t s = B.bin[I];
          if (!(s & mc) || (s & mr)) continue;
          const uint32_t J = B.index(s ^ mc ^ mr);
          if (J == B.dim) continue;
          const double X = coef * x.v * hop_sign(s, c, r);
          ent.push_back({((uint64_t)J << 32) | I, X});
        }
      }
    } else {
      for (const auto &x : op.el) {
        if ((x.r >= L) != (spin == 1) || (x.c >= L) != (spin == 1)) continue;
        const uint32_t m = (1u << (x.r - off)) | (1u << (x.c - off));
        for (uint32_t I = 0; I < B.dim; ++I)
          if ((B.bin[I] & m) == m) H.diag[I] += coef * x.v;
      }
    }
  }
  std::sort(ent.begin(), ent.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  std::vector<std::pair<uint64_t, double>> merged;
  for (const auto &e : ent) {
    if (!merged.empty() && merged.back().first == e.first) merged.back().second += e.second;
    else merged.push_back(e);
  }
  H.ptr.assign((size_t)B.dim + 1, 0);
  std::vector<uint32_t> cols;
  std::vector<double> vals;
  for (const auto &e : merged) {
    if (e.second == 0.0) continue;
    H.ptr[(e.first >> 32) + 1]++;
    cols.push_back((uint32_t)e.first);
    vals.push_back(e.second);
  }
  for (uint32_t r = 0; r < B.dim; ++r) H.ptr[r + 1] += H.ptr[r];
  // value table
  // columns 0..dim (dim itself is the sliced-ELL padding column) must fit
  H.colbits = 1;
  while (H.colbits < 32 && (1ull << H.colbits) <= (uint64_t)B.dim) ++H.colbits;
  std::vector<double> table;
  std::unordered_map<uint64_t, uint32_t> ids;
  bool ok = H.colbits < 32;
  const uint64_t cap = ok ? (1ull << (32 - H.colbits)) : 0;
  std::vector<uint32_t> vid(vals.size());
  for (size_t k = 0; ok && k < vals.size(); ++k) {
    uint64_t bits;
    std::memcpy(&bits, &vals[k], 8);
    auto it = ids.find(bits);
    if (it == ids.end()) {
      if (table.size() >= cap) { ok = false; break; }
      it = ids.emplace(bits, (uint32_t)table.size()).first;
      table.push_back(vals[k]);
    }
    vid[k] = it->second;
  }
  if (ok) {
    H.packed = 1;
    H.val = table;
    H.nval = (uint32_t)table.size();
    H.ent.resize(cols.size());
    for (size_t k = 0; k < cols.size(); ++k) H.ent[k] = cols[k] | (vid[k] << H.colbits);
  } else {
    H.packed = 0;
    H.ent = cols;
    H.val = vals;
    H.nval = 0;
  }
  return H;
}

// single magnitude test for the 16-bit sliced-ELL form: every value is
// +-t0 and columns 0..dim fit in 15 bits; returns t0 or 0
static double single_magnitude(const HalfHost &H) {
  if (!(H.dim < 32767 && params().getb("sell16"))) return 0.0;
  double t0 = 0.0;
  for (size_t e = 0; e < H.ent.size(); ++e) {
    const double v = std::fabs(H.packed ? H.val[H.ent[e] >> H.colbits] : H.val[e]);
    if (t0 == 0.0) t0 = v;
    if (v != t0) return 0.0;
  }
  return t0;
}

// Host image of one SellTable (layout described with SellTable): lane l of
// slice s holds state order[64 s + l]; the entries of a slice are padded to
// its longest row, rounded up to groups of G (G = 4 for 16-bit and packed
// entries, 1 for explicit values).
struct SellHost {
  std::vector<uint32_t> off, row, ent;
  std::vector<double> val;
  std::vector<uint16_t> e16;
};

static SellHost build_sell(const HalfHost &H, const std::vector<uint32_t> &order, double t0) {
  const uint32_t n = H.dim, ns = (n + 63) / 64;
  const bool single = t0 != 0.0;
  const uint32_t G = (single || H.packed) ? 4u : 1u;
  SellHost S;
  S.row.assign((size_t)ns * 64, n);
  for (uint32_t p = 0; p < n; ++p) S.row[p] = order[p];
  S.off.assign((size_t)ns + 1, 0);
  for (uint32_t s = 0; s < ns; ++s) {
    uint32_t w = 0;
    for (uint32_t p = s * 64; p < s * 64 + 64; ++p) {
      const uint32_t r = S.row[p];
      if (r < n) w = std::max(w, H.ptr[r + 1] - H.ptr[r]);
    }
    w = (w + G - 1) / G * G;
    S.off[s + 1] = S.off[s] + 64 * w;
  }
  const size_t tot = std::max<size_t>(S.off[ns], 1);
  S.ent.assign(single ? 1 : tot, 0u);
  S.val.assign((single || H.packed) ? 1 : tot, 0.0);
  S.e16.assign(single ? tot : 1, 0u);
  for (uint32_t s = 0; s < ns; ++s) {
    const uint32_t w = (S.off[s + 1] - S.off[s]) / 64;
    for (uint32_t lane = 0; lane < 64; ++lane) {
      const uint32_t r = S.row[(size_t)s * 64 + lane];
      const uint32_t len = r < n ? H.ptr[r + 1] - H.ptr[r] : 0;
      for (uint32_t j = 0; j < w; ++j) {
        const size_t dst = (size_t)S.off[s] + (size_t)(j / G) * 64 * G + (size_t)lane * G + j % G;
        const bool real = j < len;
        const size_t e = real ? H.ptr[r] + j : 0;
        if (single) {
          if (real) {
            const uint32_t col = H.packed ? (H.ent[e] & ((1u << H.colbits) - 1u)) : H.ent[e];
            const double v = H.packed ? H.val[H.ent[e] >> H.colbits] : H.val[e];
            S.e16[dst] = (uint16_t)(col | (v < 0.0 ? 0x8000u : 0u));
          } else {
            S.e16[dst] = (uint16_t)n;
          }
        } else if (real) {
          S.ent[dst] = H.ent[e];
          if (!H.packed) S.val[dst] = H.val[e];
        } else {
          S.ent[dst] = n;  // col = dim, value id 0 (any value times the zero slot)
          if (!H.packed) S.val[dst] = 0.0;
        }
      }
    }
  }
  return S;
}

struct DevSell {
  DevBuf<uint32_t> off, row, ent;
  DevBuf<double> val;
  DevBuf<uint16_t> e16;
  uint32_t nslice = 0;
  SellTable view(bool identity) const {
    SellTable T;
    T.nslice = nslice;
    T.off = off.p;
    T.row = identity ? nullptr : row.p;
    T.ent = ent.p;
    T.val = val.p;
    T.e16 = e16.p;
    return T;
  }
  void upload(const SellHost &S) {
    nslice = (uint32_t)S.off.size() - 1;
    off.upload(S.off);
    row.upload(S.row);
    ent.upload(S.ent);
    val.upload(S.val);
    e16.upload(S.e16);
  }
};

// breadth-first order of the states over the operator's hopping graph (each
// connected component from its lowest state): rows processed in this order
// meet their neighbour rows again sooner than in label order
static std::vector<uint32_t> bfs_order(const HalfHost &H) {
  std::vector<uint32_t> order;
  order.reserve(H.dim);
  std::vector<char> seen(H.dim, 0);
  const uint32_t mask = H.packed ? ((1u << H.colbits) - 1u) : 0xffffffffu;
  for (uint32_t s0 = 0; s0 < H.dim; ++s0) {
    if (seen[s0]) continue;
    seen[s0] = 1;
    size_t head = order.size();
    order.push_back(s0);
    while (head < order.size()) {
      const uint32_t u = order[head++];
      for (uint32_t e = H.ptr[u]; e < H.ptr[u + 1]; ++e) {
        const uint32_t c = H.ent[e] & mask;
        if (c < H.dim && !seen[c]) {
          seen[c] = 1;
          order.push_back(c);
        }
      }
    }
  }
  return order;
}

struct DevHalfStore {
  DevBuf<uint32_t> ptr, ent, states, order;
  DevBuf<double> val, diag;
  DevSell sell_nat, sell_sorted;
  DevHalf view;
  HalfHost host;
  void upload(const HalfHost &Hin, const HalfBasis &B) {
    HalfHost H = Hin;
    const double t0 = single_magnitude(H);
    std::vector<uint32_t> nat(H.dim), byrow(H.dim);
    for (uint32_t i = 0; i < H.dim; ++i) nat[i] = byrow[i] = i;
    // rows sorted by length inside windows of sell_sort_window states (0: no
    // sorting); a window bounds how far apart one wave's output stores land
    const long win = params().geti("sell_sort_window");
    const uint32_t w = win <= 0 ? 1u : (uint32_t)std::min<long>(win, (long)std::max<uint32_t>(H.dim, 1));
    for (uint32_t b0 = 0; b0 < H.dim; b0 += w)
      std::stable_sort(byrow.begin() + b0, byrow.begin() + std::min(H.dim, b0 + w), [&](uint32_t p, uint32_t q) {
        return H.ptr[p + 1] - H.ptr[p] > H.ptr[q + 1] - H.ptr[q];
      });
    sell_nat.upload(build_sell(H, nat, t0));
    sell_sorted.upload(build_sell(H, byrow, t0));
    host = H;
    ptr.upload(H.ptr);
    ent.upload(H.ent.empty() ? std::vector<uint32_t>{0u} : H.ent);
    val.upload(H.val.empty() ? std::vector<double>{0.0} : H.val);
    diag.upload(H.diag.empty() ? std::vector<double>{0.0} : H.diag);
    states.upload(B.bin.empty() ? std::vector<uint32_t>{0u} : B.bin);
    view.dim = H.dim;
    view.ptr = ptr.p;
    view.ent = ent.p;
    view.val = val.p;
    view.diag = diag.p;
    view.states = states.p;
    view.packed = H.packed;
    view.colbits = H.colbits;
    view.nval = H.nval;
    view.sell[0] = sell_nat.view(true);
    view.sell[1] = sell_sorted.view(w <= 1);
    view.mode16 = t0 != 0.0 ? 1 : 0;
    view.t0 = t0;
    order.upload(H.dim ? bfs_order(H) : std::vector<uint32_t>{0u});
    view.order = order.p;
  }
};

// cross-spin density-density coefficients W[i][j] (i up orbital, j down orbital)
static std::vector<double> cross_terms(const TermList &T) {
  const int L = T.L;
  std::vector<double> W((size_t)L * L, 0.0);
  for (const auto &tc : T.terms) {
    if (tc.first->kind != OP_INTERACTION) continue;
    for (const auto &x : tc.first->el) {
      const bool ru = x.r < L, cu = x.c < L;
      if (ru == cu) continue;
      const int i = ru ? x.r : x.c, j = ru ? x.c - L : x.r - L;
      W[(size_t)i * L + j] += tc.second * x.v;
    }
  }
  return W;
}

// Host full diagonal for general cross terms (diag_mode 0)
static std::vector<double> host_full_diag(const HalfHost &U, const HalfHost &D, const HalfBasis &BU,
                                          const HalfBasis &BD, const std::vector<double> &W, int L) {
  std::vector<double> dg((size_t)U.dim * D.dim);
#pragma omp parallel for schedule(static)
  for (long long d = 0; d < (long long)D.dim; ++d) {
    std::vector<double> wd(L, 0.0);
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j)
        if ((BD.bin[d] >> j) & 1u) wd[i] += W[(size_t)i * L + j];
    for (uint32_t u = 0; u < U.dim; ++u) {
      double s = U.diag[u] + D.diag[d];
      uint32_t m = BU.bin[u];
      while (m) {
        const int i = __builtin_ctz(m);
        s += wd[i];
        m &= m - 1;
      }
      dg[(size_t)d * U.dim + u] = s;
    }
  }
  return dg;
}

// ---------------------------------------------------------------------------
// factorized plan ('F' / 'G')
// ---------------------------------------------------------------------------
struct HamFactorized : Hamiltonian {
  DevHalfStore up, dn;
  DevBuf<double> dg;
  FactArgs args;
  std::shared_ptr<const HalfBasis> BU, BD;
  // autotuned schedule (variant 0): kernel variant + slab configuration
  bool tuned = false;
  int best_variant = HV_FUSED;
  SlabConfig best_sc;
  std::vector<double> tune_ms;

  HamFactorized(const TermList &T, const Sector &s) {
    sec = s;
    L = T.L;
    Nup = s.Nup();
    Ndw = s.Ndw();
    BU = half_basis(L, Nup);
    BD = half_basis(L, Ndw);
    dim = (size_t)BU->dim * BD->dim;
    format = 'F';
    // equal spin populations and no term that tells the spins apart: the
    // down tables are the up ones
    bool same = Nup == Ndw;
    for (const auto &tc : T.terms) same = same && !(tc.second != 0.0 && tc.first->spin_asym);
    HalfHost hu = build_half(T, 0, *BU);
    HalfHost hd = same ? hu : build_half(T, 1, *BD);
    up.upload(hu, *BU);
    dn.upload(hd, *BD);
    args.up = up.view;
    args.dn = dn.view;
    std::memset(args.Uc, 0, sizeof(args.Uc));
    std::vector<double> W = cross_terms(T);
    bool diagonal = true;
    for (int i = 0; i < L; ++i)
      for (int j = 0; j < L; ++j)
        if (i != j && W[(size_t)i * L + j] != 0.0) diagonal = false;
    if (diagonal) {
      bool uniform = true;
      for (int i = 0; i < L; ++i) {
        args.Uc[i] = W[(size_t)i * L + i];
        if (args.Uc[i] != args.Uc[0]) uniform = false;
      }
      args.diag_mode = uniform ? 2 : 1;
    } else {
      args.diag_mode = 0;
      dg.upload(host_full_diag(hu, hd, *BU, *BD, W, L));
      args.dg = dg.p;
    }
    variant = (int)params().geti("hv_variant");
    args.block = (int)params().geti("hv_block");
    args.row_order = (int)params().geti("hv_row_order");
  }
  // candidate schedules, in the order reported by ed_hamiltonian_autotune:
  // the fused kernel, the two-pass schedule with L2-resident slabs, and (when
  // a slab of dd rows fits) with LDS-resident slabs
  static std::vector<std::pair<int, SlabConfig>> candidates(uint32_t dd) {
    std::vector<std::pair<int, SlabConfig>> c;
    c.push_back({HV_FUSED, SlabConfig()});
    for (int sw : {8, 16, 32})
      for (int nt : {0, 1}) {
        SlabConfig s;
        s.sw = sw;
        s.nt = nt;
        c.push_back({HV_TWO_PASS, s});
      }
    for (int sw : {4, 8, 16})
      if ((size_t)dd * sw * sizeof(double) <= kDownLdsMax) {
        SlabConfig s;
        s.sw = sw;
        s.lds = 1;
        c.push_back({HV_TWO_PASS, s});
      }
    return c;
  }
  // force: time the candidates even when the shape's result is cached
  void autotune(bool force = false) {
    tuned = true;
    best_variant = HV_FUSED;
    const bool small = dim < ((size_t)1 << 20);
    // small sectors: the fused kernel or an LDS-slab two-pass schedule, timed
    // once per sector shape in the process (the tables' sizes key it)
    static std::map<std::array<uint64_t, 4>, std::pair<int, SlabConfig>> small_cache;
    static std::mutex small_mu;
    const std::array<uint64_t, 4> key{args.up.dim, args.dn.dim, (uint64_t)up.host.ptr.back(),
                                      (uint64_t)dn.host.ptr.back()};
    if (small && !force) {
      std::lock_guard<std::mutex> l(small_mu);
      auto it = small_cache.find(key);
      if (it != small_cache.end()) {
        best_variant = it->second.first;
        best_sc = it->second.second;
        return;
      }
    }
    DevBuf<double> x(dim), y(dim);
    fill(dim, x.p, 1.0 / std::sqrt((double)dim), stream());
    fill(dim, y.p, 0.0, stream());
    hipEvent_t e0, e1;
    ED_HIP(hipEventCreate(&e0));
    ED_HIP(hipEventCreate(&e1));
    double best = 1e300;
    tune_ms.clear();
    for (const auto &c : candidates(args.dn.dim)) {
      if (small && c.first == HV_TWO_PASS && !c.second.lds) {
        tune_ms.push_back(-1.0);  // L2-slab schedules: for sectors beyond the LDS slabs
        continue;
      }
      hv_factorized_cfg(args, x.p, y.p, 1.0, 0.0, nullptr, c.first, c.second, stream());
      ED_HIP(hipEventRecord(e0, stream()));
      for (int r = 0; r < 3; ++r) hv_factorized_cfg(args, x.p, y.p, 1.0, 0.0, nullptr, c.first, c.second, stream());
      ED_HIP(hipEventRecord(e1, stream()));
      ED_HIP(hipEventSynchronize(e1));
      float t = 0.f;
      ED_HIP(hipEventElapsedTime(&t, e0, e1));
      const double ms = t / 3.0;
      tune_ms.push_back(ms);
      if (ms < best) {
        best = ms;
        best_variant = c.first;
        best_sc = c.second;
      }
    }
    ED_HIP(hipEventDestroy(e0));
    ED_HIP(hipEventDestroy(e1));
    if (small) {
      std::lock_guard<std::mutex> l(small_mu);
      small_cache[key] = {best_variant, best_sc};
    }
  }
  int apply(const double *x, double *y, double ax, double beta, bool want_dot) override {
    int v = variant;
    SlabConfig sc;
    if (v == HV_AUTO) {
      if (!tuned) autotune();
      v = best_variant;
      sc = best_sc;
    }
    double *part = want_dot ? take_partials() : nullptr;
    const int nb = hv_factorized_cfg(args, x, y, ax, beta, part, v, sc, stream());
    if (part) pending_nb = nb;
    return nb;
  }
  void prepare() override {
    if (variant == HV_AUTO && !tuned) autotune();
  }
  int apply_fixed(const double *x, double *y, double ax, double beta, double *part, double *aux,
                  size_t aux_cap) override {
    int v = variant;
    SlabConfig sc;
    if (v == HV_AUTO && tuned) {
      v = best_variant;
      sc = best_sc;
    }
    return hv_factorized_cfg(args, x, y, ax, beta, part, v, sc, stream(), aux, aux_cap);
  }
  void diag(double *d) override { fact_full_diag(args, d, stream()); }
};

// ---------------------------------------------------------------------------
// CSR plan ('S' / 'E'): bit-exact strict-lower pattern of Hamiltonian_CSR,
// stored on device as the full symmetric matrix for a row-parallel SpMV.
// ---------------------------------------------------------------------------
void csr_lower(const TermList &T, int Nup, int Ndw, std::vector<int32_t> &Iptr, std::vector<int32_t> &J,
               std::vector<double> &v, std::vector<double> &diag) {
  const int L = T.L;
  auto BU = half_basis(L, Nup), BD = half_basis(L, Ndw);
  const size_t du = BU->dim, dd = BD->dim, dim = du * dd;
  if (dim >= 0x7fffffffull) fail(ED_ERR_UNSUPPORTED, "CSR format limited to dim < 2^31 (int32 indices)");
  diag.assign(dim, 0.0);
  std::vector<std::pair<uint64_t, double>> E;
  for (const auto &tc : T.terms) {
    const Operator &op = *tc.first;
    const double z = tc.second;
    for (size_t I = 0; I < dim; ++I) {
      const uint32_t iu = (uint32_t)(I % du), id = (uint32_t)(I / du);
      const uint32_t su = BU->bin[iu], sd = BD->bin[id];
      if (op.kind == OP_INTERACTION) {
        for (const auto &x : op.el) {
          const bool r_ok = (x.r < L) ? ((su >> x.r) & 1u) : ((sd >> (x.r - L)) & 1u);
          const bool c_ok = (x.c < L) ? ((su >> x.c) & 1u) : ((sd >> (x.c - L)) & 1u);
          if (r_ok && c_ok) diag[I] += z * x.v;
        }
        continue;
      }
      for (const auto &x : op.el) {
        if (x.r == x.c) {
          const bool occ = (x.r < L) ? ((su >> x.r) & 1u) : ((sd >> (x.r - L)) & 1u);
          if (occ && std::fabs(x.v) >= 1e-12) diag[I] += z * x.v;
          continue;
        }
        if ((x.r < L) != (x.c < L)) continue;  // spin flip leaves the (Nup, Ndw) sector
        const bool upspin = x.r < L;
        const int r = upspin ? x.r : x.r - L, c = upspin ? x.c : x.c - L;
        const uint32_t s = upspin ? su : sd;
        if (!((s >> c) & 1u) || ((s >> r) & 1u)) continue;
        const uint32_t s2 = s ^ (1u << c) ^ (1u << r);
        size_t Jo;
        if (upspin) {
          const uint32_t k = BU->index(s2);
          if (k == BU->dim) continue;
          Jo = (size_t)k + du * id;
        } else {
          const uint32_t k = BD->index(s2);
          if (k == BD->dim) continue;
          Jo = (size_t)iu + du * k;
        }
        const double X = x.v * hop_sign(s, c, r);
        if (std::fabs(X) < 1e-12) continue;
        if (!(I < Jo)) continue;  // HS_nondiagonal_operator::insert(J_out, I_in): keep I_in < J_out
        E.push_back({((uint64_t)Jo << 32) | I, X * z});
      }
    }
  }
  std::sort(E.begin(), E.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
  Iptr.assign(dim + 1, 0);
  J.clear();
  v.clear();
  for (size_t k = 0; k < E.size(); ++k) {
    if (!J.empty() && k > 0 && E[k - 1].first == E[k].first) {
      v.back() += E[k].second;
      continue;
    }
    Iptr[(E[k].first >> 32) + 1]++;
    J.push_back((int32_t)(E[k].first & 0xffffffffu));
    v.push_back(E[k].second);
  }
  for (size_t r = 0; r < dim; ++r) Iptr[r + 1] += Iptr[r];
}

struct HamCSR : Hamiltonian {
  DevBuf<int64_t> ptr;
  DevBuf<uint32_t> col;
  DevBuf<double> val, dgv;
  CsrArgs args;
  std::vector<int32_t> Iptr, J;
  std::vector<double> v, dgh;
  HamCSR(const TermList &T, const Sector &s, char fmt) {
    sec = s;
    L = T.L;
    Nup = s.Nup();
    Ndw = s.Ndw();
    format = fmt;
    csr_lower(T, Nup, Ndw, Iptr, J, v, dgh);
    dim = dgh.size();
    // full symmetric rows
    std::vector<int64_t> fp(dim + 1, 0);
    for (size_t i = 0; i < dim; ++i)
      for (int32_t e = Iptr[i]; e < Iptr[i + 1]; ++e) { fp[i + 1]++; fp[(size_t)J[e] + 1]++; }
    for (size_t i = 0; i < dim; ++i) fp[i + 1] += fp[i];
    std::vector<uint32_t> fc(fp[dim]);
    std::vector<double> fv(fp[dim]);
    std::vector<int64_t> pos(fp.begin(), fp.end() - 1);
    for (size_t i = 0; i < dim; ++i)
      for (int32_t e = Iptr[i]; e < Iptr[i + 1]; ++e) {
        fc[pos[i]] = (uint32_t)J[e];
        fv[pos[i]++] = v[e];
        fc[pos[J[e]]] = (uint32_t)i;
        fv[pos[J[e]]++] = v[e];
      }
    ptr.upload(fp);
    col.upload(fc.empty() ? std::vector<uint32_t>{0u} : fc);
    val.upload(fv.empty() ? std::vector<double>{0.0} : fv);
    dgv.upload(dgh.empty() ? std::vector<double>{0.0} : dgh);
    args.n = dim;
    args.ptr = ptr.p;
    args.col = col.p;
    args.val = val.p;
    args.diag = dgv.p;
  }
  int apply(const double *x, double *y, double ax, double beta, bool want_dot) override {
    double *part = want_dot ? take_partials() : nullptr;
    const int nb = hv_csr(args, x, y, ax, beta, part, stream());
    if (part) pending_nb = nb;
    return nb;
  }
  int apply_fixed(const double *x, double *y, double ax, double beta, double *part, double *, size_t) override {
    return hv_csr(args, x, y, ax, beta, part, stream());
  }
  void diag(double *d) override {
    if (dim) ED_HIP(hipMemcpyAsync(d, dgv.p, dim * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  }
};

// ---------------------------------------------------------------------------
// on-the-fly plan ('N')
// ---------------------------------------------------------------------------
struct HamOTF : Hamiltonian {
  DevHalfStore up, dn;   // used for the diagonal and the state tables
  DevBuf<double> dg;
  DevBuf<OtfHop> hops;
  DevBuf<uint32_t> binom;
  OtfArgs args;
  HamOTF(const TermList &T, const Sector &s) {
    sec = s;
    L = T.L;
    Nup = s.Nup();
    Ndw = s.Ndw();
    format = 'N';
    auto BU = half_basis(L, Nup), BD = half_basis(L, Ndw);
    dim = (size_t)BU->dim * BD->dim;
    HalfHost hu = build_half(T, 0, *BU), hd = build_half(T, 1, *BD);
    up.upload(hu, *BU);
    dn.upload(hd, *BD);
    std::vector<double> W = cross_terms(T);
    dg.upload(host_full_diag(hu, hd, *BU, *BD, W, L));
    std::map<std::pair<int, int>, double> hm;
    for (const auto &tc : T.terms) {
      if (tc.first->kind != OP_ONE_BODY) continue;
      for (const auto &x : tc.first->el) {
        if (x.r == x.c) continue;
        if ((x.r < L) != (x.c < L)) continue;
        // element v c+_r c_c: pull row I from J = I with r -> c
        hm[{x.r, x.c}] += tc.second * x.v;
      }
    }
    std::vector<OtfHop> hv;
    for (const auto &kv : hm)
      if (kv.second != 0.0) hv.push_back({kv.first.first, kv.first.second, kv.second});
    hops.upload(hv.empty() ? std::vector<OtfHop>{{0, 0, 0.0}} : hv);
    std::vector<uint32_t> bn((size_t)(L + 1) * (L + 1), 0);
    for (int n = 0; n <= L; ++n) {
      bn[(size_t)n * (L + 1)] = 1;
      for (int k = 1; k <= n; ++k)
        bn[(size_t)n * (L + 1) + k] = bn[(size_t)(n - 1) * (L + 1) + k - 1] + (k <= n - 1 ? bn[(size_t)(n - 1) * (L + 1) + k] : 0);
    }
    binom.upload(bn);
    args.L = L;
    args.du = BU->dim;
    args.dd = BD->dim;
    args.up_states = up.states.p;
    args.dn_states = dn.states.p;
    args.nhop = (int)hv.size();
    args.hops = hops.p;
    args.dg = dg.p;
    args.binom = binom.p;
  }
  int apply(const double *x, double *y, double ax, double beta, bool want_dot) override {
    double *part = want_dot ? take_partials() : nullptr;
    const int nb = hv_otf(args, x, y, ax, beta, part, stream());
    if (part) pending_nb = nb;
    return nb;
  }
  int apply_fixed(const double *x, double *y, double ax, double beta, double *part, double *, size_t) override {
    return hv_otf(args, x, y, ax, beta, part, stream());
  }
  void diag(double *d) override {
    if (dim) ED_HIP(hipMemcpyAsync(d, dg.p, dim * sizeof(double), hipMemcpyDeviceToDevice, stream()));
  }
};

std::vector<double> hamiltonian_autotune(Hamiltonian &H, int &best) {
  auto *F = dynamic_cast<HamFactorized *>(&H);
  if (!F) fail(ED_ERR_UNSUPPORTED, "autotuning applies to the factorized format only");
  F->autotune(true);
  const auto cands = HamFactorized::candidates(F->args.dn.dim);
  best = 0;
  for (size_t k = 0; k < cands.size(); ++k)
    if (cands[k].first == F->best_variant && cands[k].second.sw == F->best_sc.sw &&
        cands[k].second.nt == F->best_sc.nt && cands[k].second.lds == F->best_sc.lds)
      best = (int)k;
  return F->tune_ms;
}

static bool factorizable(const TermList &T) {
  for (const auto &tc : T.terms)
    if (tc.first->kind == OP_ONE_BODY && tc.first->spin_flip && tc.second != 0.0) return false;
  return true;
}

bool terms_factorizable(const TermList &T) { return factorizable(T); }

FactorizedHost factorized_tables(const TermList &T, const Sector &sec) {
  if (!factorizable(T)) fail(ED_ERR_UNSUPPORTED, "the instance's operators do not factorize by spin");
  if (terms_complex(T)) fail(ED_ERR_UNSUPPORTED, "the instance's operators have complex elements");
  if (sec.Nup() < 0 || sec.Ndw() < 0 || sec.Nup() > T.L || sec.Ndw() > T.L || sec.Nup() >= Sector::odd ||
      sec.Ndw() >= Sector::odd)
    fail(ED_ERR_SECTOR, "sector " + sec.name() + " has no factorized basis");
  const auto BU = half_basis(T.L, sec.Nup()), BD = half_basis(T.L, sec.Ndw());
  const HalfHost hu = build_half(T, 0, *BU), hd = build_half(T, 1, *BD);
  FactorizedHost F;
  F.du = BU->dim;
  F.dd = BD->dim;
  auto unpack = [](const HalfHost &H, std::vector<uint32_t> &ptr, std::vector<uint32_t> &col,
                   std::vector<double> &val) {
    ptr = H.ptr;
    col.resize(H.ent.size());
    val.resize(H.ent.size());
    for (size_t e = 0; e < H.ent.size(); ++e) {
      col[e] = H.packed ? (H.ent[e] & ((1u << H.colbits) - 1u)) : H.ent[e];
      val[e] = H.packed ? H.val[H.ent[e] >> H.colbits] : H.val[e];
    }
  };
  unpack(hu, F.up_ptr, F.up_col, F.up_val);
  unpack(hd, F.dn_ptr, F.dn_col, F.dn_val);
  F.diag = host_full_diag(hu, hd, *BU, *BD, cross_terms(T), T.L);
  return F;
}

bool terms_complex(const TermList &T) {
  for (const auto &tc : T.terms)
    if (tc.first->is_complex && tc.second != 0.0) return true;
  return false;
}

static std::unique_ptr<Hamiltonian> make_plan(const TermList &T, const Sector &sec, char format);

// Complex Hilbert space (model_instance<Complex>, src_ed/qcm_ED.cpp:127-158) on
// split vectors x = [x_r | x_i] of 2 n doubles: H = A + i B with A the real
// parts of every term (symmetric) and B the imaginary parts of the one-body
// terms (antisymmetric), so that
//   H x = [A x_r - B x_i | B x_r + A x_i],
// the real symmetric matrix [[A, -B], [B, A]]: its levels are those of H, each
// twice (psi and i psi).  A is the plan of the requested format; B, which has
// no diagonal, is factorized by spin (on the fly where a term flips spin).
// Every Krylov solver runs on it unchanged.
struct HamComplex : Hamiltonian {
  std::unique_ptr<Hamiltonian> A, B;  // B null when no term has imaginary parts
  size_t n = 0;
  HamComplex(const TermList &T, const Sector &s, char fmt) {
    A = make_plan(T, s, fmt);
    TermList Ti;
    Ti.L = T.L;
    for (const auto &tc : T.terms)
      if (tc.first->imag && tc.second != 0.0) Ti.terms.push_back({tc.first->imag.get(), tc.second});
    if (!Ti.terms.empty()) {
      if (fmt != 'N' && factorizable(Ti)) B.reset(new HamFactorized(Ti, s));
      else B.reset(new HamOTF(Ti, s));
      B->sec = s;
      B->terms = Ti;
    }
    sec = s;
    L = T.L;
    Nup = s.Nup();
    Ndw = s.Ndw();
    format = fmt;
    n = A->dim;
    dim = 2 * n;
  }
  void run(const double *x, double *y, double ax, double beta, double *aux, size_t aux_cap, bool fixed) {
    auto one = [&](Hamiltonian &P, const double *xx, double *yy, double a, double b) {
      if (fixed) P.apply_fixed(xx, yy, a, b, nullptr, aux, aux_cap);
      else P.apply(xx, yy, a, b, false);
    };
    one(*A, x, y, ax, beta);
    if (B) one(*B, x + n, y, -ax, 1.0);
    one(*A, x + n, y + n, ax, beta);
    if (B) one(*B, x, y + n, ax, 1.0);
  }
  int apply(const double *x, double *y, double ax, double beta, bool want_dot) override {
    run(x, y, ax, beta, nullptr, 0, false);
    if (!want_dot) return 0;
    double *part = take_partials();
    const double *b[1] = {x};
    const int nb = multi_dot(dim, y, 1, b, part, stream());
    pending_nb = nb;
    return nb;
  }
  int apply_fixed(const double *x, double *y, double ax, double beta, double *part, double *aux,
                  size_t aux_cap) override {
    run(x, y, ax, beta, aux, aux_cap, true);
    if (!part) return 0;
    const double *b[1] = {x};
    return multi_dot(dim, y, 1, b, part, stream());
  }
  void prepare() override {
    A->prepare();
    if (B) B->prepare();
  }
  void diag(double *d) override {
    A->diag(d);
    A->diag(d + n);
  }
};

std::unique_ptr<Hamiltonian> make_hamiltonian(const TermList &T, const Sector &sec, char format, bool force_complex) {
  if (!force_complex && !terms_complex(T)) return make_plan(T, sec, format);
  vlog("Hamiltonian %s: complex Hilbert space", sec.name().c_str());
  std::unique_ptr<Hamiltonian> H(new HamComplex(T, sec, format));
  H->sec = sec;
  H->terms = T;
  return H;
}

static std::unique_ptr<Hamiltonian> make_plan(const TermList &T, const Sector &sec, char format) {
  if (sec.Nup() >= Sector::odd || sec.Ndw() >= Sector::odd)
    fail(ED_ERR_UNSUPPORTED, "sector " + sec.name() + ": only sectors with conserved N and S are supported");
  if (sec.Nup() < 0 || sec.Ndw() < 0 || sec.Nup() > T.L || sec.Ndw() > T.L)
    fail(ED_ERR_SECTOR, "sector " + sec.name() + " is empty for " + std::to_string(T.L) + " orbitals");
  for (const auto &tc : T.terms)
    if (tc.first->spin_flip && tc.second != 0.0)
      fail(ED_ERR_UNSUPPORTED, "operator " + tc.first->name + " flips spin: spin-nonconserving sectors are not supported");
  vlog("Hamiltonian %s (format %c): build", sec.name().c_str(), format);
  std::unique_ptr<Hamiltonian> H;
  if (format == 'G' || format == 'F') {
    if (factorizable(T)) H.reset(new HamFactorized(T, sec));
    else H.reset(new HamCSR(T, sec, 'S'));
  } else if (format == 'S' || format == 'E' || format == 'O') {
    H.reset(new HamCSR(T, sec, format));
  } else if (format == 'N') {
    H.reset(new HamOTF(T, sec));
  } else {
    fail(ED_ERR_INVALID, std::string("Hamiltonian_format is not one of : G, N, S, O, F, E (got '") + format + "')");
  }
  H->sec = sec;
  H->terms = T;
  vlog("Hamiltonian %s: dim %zu ready", sec.name().c_str(), H->dim);
  return H;
}

// ---------------------------------------------------------------------------
// dense host helpers
// ---------------------------------------------------------------------------
// implicit QL on a symmetric tridiagonal matrix; rotations accumulated into the
// rows of Z given by zrows (n_z rows of length n, row-major) -- the full matrix
// (zrows = n, identity start) or only the last row (for the Ritz residual).
void ql_rows(int n, double *d, double *e, double *Z, int nz) {
  if (n <= 0) return;
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) fail(ED_ERR_NOT_CONVERGED, "tridiagonal eigensolver did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = fast_hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          r = fast_hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) { d[i + 1] -= p; e[m] = 0.0; break; }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < nz; ++k) {
            double *zr = Z + (size_t)k * n;
            const double fz = zr[i + 1];
            zr[i + 1] = s * zr[i] + c * fz;
            zr[i] = c * zr[i] - s * fz;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

// Lowest eigenpair of the symmetric tridiagonal matrix (d[0..n), e[0..n-1)).
// The eigenvalue comes from bisection on the Sturm count (inertia of the
// LDL^T factorisation of T - x), O(n) per step; the eigenvector from inverse
// iteration with T - (lambda - delta), which is positive definite, so the
// unpivoted LDL^T solve is stable; delta / gap is ~1e-12, and two sweeps reach
// rounding level.  O(n) memory and O(60 n) work, where the dense QL with
// eigenvectors costs O(n^2) to O(n^3) per call.
static int sturm_below(int n, const double *d, const double *e, double x) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < n; ++i) {
    if (q == 0.0) q = -1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

static double tridiag_min_eig(int n, const double *d, const double *e, double &norm) {
  double lo = d[0], hi = d[0];
  norm = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
    norm = std::max(norm, std::fabs(d[i]) + r);
  }
  // hi: at least one eigenvalue below; lo: none below
  hi += 1e-300 + 4.0 * std::numeric_limits<double>::epsilon() * norm;
  lo -= 1e-300 + 4.0 * std::numeric_limits<double>::epsilon() * norm;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int it = 0; it < 256; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (hi - lo <= 2.0 * eps * std::max(std::fabs(lo), std::fabs(hi)) + 1e-3 * eps * norm) break;
    if (sturm_below(n, d, e, mid) >= 1) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

static void tridiag_inverse_iteration(int n, const double *d, const double *e, double lam, double norm,
                                      std::vector<double> &z) {
  z.assign(n, 1.0);
  std::vector<double> D(n), l(n > 1 ? n - 1 : 1);
  double delta = (4.0 * std::max(n, 16)) * std::numeric_limits<double>::epsilon() * std::max(norm, 1e-300);
  for (int attempt = 0;; ++attempt) {
    const double sigma = lam - delta;
    bool ok = true;
    D[0] = d[0] - sigma;
    for (int i = 0; i < n - 1 && ok; ++i) {
      if (!(D[i] > 0.0)) { ok = false; break; }
      l[i] = e[i] / D[i];
      D[i + 1] = d[i + 1] - sigma - l[i] * e[i];
    }
    if (ok && !(D[n - 1] > 0.0)) ok = false;
    if (ok) break;
    if (attempt == 40) fail(ED_ERR_NOT_CONVERGED, "tridiagonal inverse iteration: no positive definite shift");
    delta *= 8.0;
  }
  for (int sweep = 0; sweep < 3; ++sweep) {
    for (int i = 1; i < n; ++i) z[i] -= l[i - 1] * z[i - 1];       // L y = z
    for (int i = 0; i < n; ++i) z[i] /= D[i];                       // D w = y
    for (int i = n - 2; i >= 0; --i) z[i] -= l[i] * z[i + 1];      // L^T x = w
    double s = 0.0, m = 0.0;
    for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(z[i]));
    for (int i = 0; i < n; ++i) { z[i] /= m; s += z[i] * z[i]; }
    s = 1.0 / std::sqrt(s);
    for (int i = 0; i < n; ++i) z[i] *= s;
  }
}

// Rayleigh-quotient iteration from the shift 'guess' (the previous Ritz value
// of a Lanczos run, close to the new lowest one): each sweep solves
// (T - sigma) x = z with the pivoted-free LDL^T of a shifted matrix and moves
// sigma to the Rayleigh quotient; cubic convergence, a few O(n) sweeps.  The
// result is accepted only when the Sturm count shows no eigenvalue below it
// (otherwise the caller falls back to bisection).
static bool tridiag_lowest_rqi(int n, const double *d, const double *e, double guess, double norm, double &lam,
                               std::vector<double> &z) {
  const double eps = std::numeric_limits<double>::epsilon();
  z.assign(n, 1.0 / std::sqrt((double)n));
  std::vector<double> D(n), l(n > 1 ? n - 1 : 1), x(n);
  double sigma = guess;
  for (int it = 0; it < 8; ++it) {
    // LDL^T of T - sigma (no pivoting; a zero pivot is nudged)
    D[0] = d[0] - sigma;
    for (int i = 0; i < n - 1; ++i) {
      if (D[i] == 0.0) D[i] = eps * std::max(norm, 1e-300);
      l[i] = e[i] / D[i];
      D[i + 1] = d[i + 1] - sigma - l[i] * e[i];
    }
    if (D[n - 1] == 0.0) D[n - 1] = eps * std::max(norm, 1e-300);
    x = z;
    for (int i = 1; i < n; ++i) x[i] -= l[i - 1] * x[i - 1];
    for (int i = 0; i < n; ++i) x[i] /= D[i];
    for (int i = n - 2; i >= 0; --i) x[i] -= l[i] * x[i + 1];
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i] * x[i];
    if (!(s > 0.0) || !std::isfinite(s)) return false;
    s = 1.0 / std::sqrt(s);
    for (int i = 0; i < n; ++i) z[i] = x[i] * s;
    // Rayleigh quotient and residual
    double rq = 0.0;
    for (int i = 0; i < n; ++i) {
      double tz = d[i] * z[i];
      if (i > 0) tz += e[i - 1] * z[i - 1];
      if (i < n - 1) tz += e[i] * z[i + 1];
      rq += z[i] * tz;
    }
    double r2 = 0.0;
    for (int i = 0; i < n; ++i) {
      double tz = (d[i] - rq) * z[i];
      if (i > 0) tz += e[i - 1] * z[i - 1];
      if (i < n - 1) tz += e[i] * z[i + 1];
      r2 += tz * tz;
    }
    sigma = rq;
    if (std::sqrt(r2) <= 8.0 * n * eps * std::max(norm, 1e-300)) break;
  }
  lam = sigma;
  const double tol = 16.0 * n * eps * std::max(norm, 1e-300);
  return sturm_below(n, d, e, lam - tol) == 0;
}

void tridiag_eigvec0(int n, const double *d, const double *e, double &ev0, std::vector<double> &z0, double guess) {
  if (n <= 0) fail(ED_ERR_INVALID, "empty tridiagonal matrix");
  if (n == 1) {
    // a one-state Krylov space (a sector of dimension 1, or an exact
    // eigenvector as start): the shifted factorisation below would divide by a
    // shift of the matrix's own (possibly zero) scale
    ev0 = d[0];
    z0.assign(1, 1.0);
    return;
  }
  double norm = 0.0;
  for (int i = 0; i < n; ++i)
    norm = std::max(norm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < n - 1 ? std::fabs(e[i]) : 0.0));
  if (std::isfinite(guess) && n > 2 && tridiag_lowest_rqi(n, d, e, guess, norm, ev0, z0)) {
    // refine the eigenvector with the shifted-definite inverse iteration
    tridiag_inverse_iteration(n, d, e, ev0, norm, z0);
    return;
  }
  ev0 = tridiag_min_eig(n, d, e, norm);
  tridiag_inverse_iteration(n, d, e, ev0, norm, z0);
}

void tridiag_lowest(int n, const double *d, const double *e, double &ev0, double &zlast0, double guess) {
  std::vector<double> z;
  tridiag_eigvec0(n, d, e, ev0, z, guess);
  zlast0 = z[n - 1];
}

// Householder tridiagonalisation (row-major working copy) + implicit QL
// Hermitian eigensystem by cyclic Jacobi sweeps.  For the pair (p, q) with
// a_pq = b e^{i phi}, U = diag(1, e^{-i phi}) R with the real plane rotation R
// of angle theta = atan2(2 b, a_qq - a_pp) / 2 makes (U^H A U)_pq zero:
//   columns  A_p <- c A_p - s e^{-i phi} A_q,  A_q <- s A_p + c e^{-i phi} A_q,
//   rows     A_p <- c A_p - s e^{i phi} A_q,   A_q <- s A_p + c e^{i phi} A_q.
void hermeig(int n, const cplx *Ain, double *w, cplx *V) {
  if (n <= 0) return;
  std::vector<cplx> a(Ain, Ain + (size_t)n * n), v((size_t)n * n, 0.0);
  auto A = [&](int i, int j) -> cplx & { return a[i + (size_t)j * n]; };
  for (int i = 0; i < n; ++i) v[i + (size_t)i * n] = 1.0;
  double fro = 0.0;
  for (const auto &x : a) fro += std::norm(x);
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += std::norm(A(p, q));
    if (!(off > 1e-32 * fro)) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double b = std::abs(A(p, q));
        if (b == 0.0) continue;
        const cplx ph = A(p, q) / b;
        const double th = 0.5 * std::atan2(2.0 * b, A(q, q).real() - A(p, p).real());
        const double c = std::cos(th), s = std::sin(th);
        const cplx em = std::conj(ph);  // e^{-i phi}
        for (int k = 0; k < n; ++k) {
          const cplx xp = A(k, p), xq = A(k, q);
          A(k, p) = c * xp - s * em * xq;
          A(k, q) = s * xp + c * em * xq;
          const cplx vp = v[k + (size_t)p * n], vq = v[k + (size_t)q * n];
          v[k + (size_t)p * n] = c * vp - s * em * vq;
          v[k + (size_t)q * n] = s * vp + c * em * vq;
        }
        for (int k = 0; k < n; ++k) {
          const cplx xp = A(p, k), xq = A(q, k);
          A(p, k) = c * xp - s * ph * xq;
          A(q, k) = s * xp + c * ph * xq;
        }
        A(p, q) = A(q, p) = 0.0;
        A(p, p) = A(p, p).real();
        A(q, q) = A(q, q).real();
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A(x, x).real() < A(y, y).real(); });
  for (int i = 0; i < n; ++i) {
    w[i] = A(idx[i], idx[i]).real();
    if (V)
      for (int k = 0; k < n; ++k) V[k + (size_t)i * n] = v[k + (size_t)idx[i] * n];
  }
}

void symeig(int n, const double *Ain, double *w, double *V) {
  if (n <= 0) return;
  std::vector<double> a((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = Ain[i + (size_t)j * n];
  std::vector<double> d(n), e(n);
  const bool vec = V != nullptr;
#define A_(i, j) a[(size_t)(i) * n + (j)]
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A_(i, k));
      if (scale == 0.0) {
        e[i] = A_(i, l);
      } else {
        for (int k = 0; k <= l; ++k) { A_(i, k) /= scale; h += A_(i, k) * A_(i, k); }
        double f = A_(i, l);
        double g = (f >= 0.0) ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A_(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          if (vec) A_(j, i) = A_(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A_(j, k) * A_(i, k);
          for (int k = j + 1; k <= l; ++k) g += A_(k, j) * A_(i, k);
          e[j] = g / h;
          f += e[j] * A_(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A_(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A_(j, k) -= (f * e[k] + g * A_(i, k));
        }
      }
    } else {
      e[i] = A_(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    if (vec) {
      const int l = i - 1;
      if (d[i] != 0.0) {
        for (int j = 0; j <= l; ++j) {
          double g = 0.0;
          for (int k = 0; k <= l; ++k) g += A_(i, k) * A_(k, j);
          for (int k = 0; k <= l; ++k) A_(k, j) -= g * A_(k, i);
        }
      }
      d[i] = A_(i, i);
      A_(i, i) = 1.0;
      for (int j = 0; j <= l; ++j) A_(j, i) = A_(i, j) = 0.0;
    } else {
      d[i] = A_(i, i);
    }
  }
#undef A_
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  ql_rows(n, d.data(), e.data(), vec ? a.data() : nullptr, vec ? n : 0);
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int k = 0; k < n; ++k) w[k] = d[idx[k]];
  if (vec)
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) V[i + (size_t)k * n] = a[(size_t)i * n + idx[k]];
}

}  // namespace ed
