// ed_batch.cpp -- host side of the batched solver: builds one work item per
// (instance, target sector) for the ground state and one per (state, spin,
// +-, orbital pair) for the continued fractions, runs each stage as a single
// launch, and hands the results to the instances exactly as their own solve
// would (finalize_states; GFRep 'F').
#include "ed_batch.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <tuple>

#include "ed_instance.hpp"
#include "ed_solvers.hpp"

namespace ed {
namespace {

// items whose sector has at most this many states run in a one-wave workgroup
constexpr uint32_t kOneWaveDim = 4096;

// packed device copy of many FactorizedHost tables
struct TablePack {
  std::vector<uint32_t> u32;
  std::vector<double> f64;
  struct Off { size_t up_ptr, up_col, dn_ptr, dn_col, up_val, dn_val, diag; uint32_t du, dd; };
  std::vector<Off> offs;
  DevBuf<uint32_t> du32;
  DevBuf<double> df64;
  int add(const FactorizedHost &F) {
    Off o;
    auto put32 = [&](const std::vector<uint32_t> &v) {
      const size_t at = u32.size();
      u32.insert(u32.end(), v.begin(), v.end());
      if (v.empty()) u32.push_back(0);
      return at;
    };
    auto put64 = [&](const std::vector<double> &v) {
      const size_t at = f64.size();
      f64.insert(f64.end(), v.begin(), v.end());
      if (v.empty()) f64.push_back(0.0);
      return at;
    };
    o.up_ptr = put32(F.up_ptr);
    o.up_col = put32(F.up_col);
    o.dn_ptr = put32(F.dn_ptr);
    o.dn_col = put32(F.dn_col);
    o.up_val = put64(F.up_val);
    o.dn_val = put64(F.dn_val);
    o.diag = put64(F.diag);
    o.du = F.du;
    o.dd = F.dd;
    offs.push_back(o);
    return (int)offs.size() - 1;
  }
  void upload() {
    du32.upload(u32);
    df64.upload(f64);
  }
  BatchSector view(int k) const {
    const Off &o = offs[k];
    BatchSector S;
    S.du = o.du;
    S.dd = o.dd;
    S.up_ptr = du32.p + o.up_ptr;
    S.up_col = du32.p + o.up_col;
    S.dn_ptr = du32.p + o.dn_ptr;
    S.dn_col = du32.p + o.dn_col;
    S.up_val = df64.p + o.up_val;
    S.dn_val = df64.p + o.dn_val;
    S.diag = df64.p + o.diag;
    return S;
  }
};

template <typename Item>
void launch_split(std::vector<Item> &items, const std::vector<uint32_t> &dims,
                  void (*launch)(const Item *, int, int, int, double, hipStream_t), int iters, double acc) {
  // one-wave and four-wave groups, each one launch
  std::vector<Item> small, large;
  for (size_t k = 0; k < items.size(); ++k) (dims[k] <= kOneWaveDim ? small : large).push_back(items[k]);
  DevBuf<Item> ds, dl;
  if (!small.empty()) {
    ds.alloc(small.size());
    ED_HIP(hipMemcpyAsync(ds.p, small.data(), small.size() * sizeof(Item), hipMemcpyHostToDevice, stream()));
    launch(ds.p, (int)small.size(), 64, iters, acc, stream());
  }
  if (!large.empty()) {
    dl.alloc(large.size());
    ED_HIP(hipMemcpyAsync(dl.p, large.data(), large.size() * sizeof(Item), hipMemcpyHostToDevice, stream()));
    launch(dl.p, (int)large.size(), 256, iters, acc, stream());
  }
  sync();
}

bool batchable(const Instance &I) {
  const Params &P = params();
  (void)P;
  if (I.gs_solved) return false;
  if (I.format != 'G' && I.format != 'F') return false;
  if (!terms_factorizable(I.terms())) return false;
  if (params().getc("Ground_state_method") != 'L') return false;
  if (params().geti("n_states") > 1) return false;
  if (params().getb("host_random_start")) return false;
  return true;
}

}  // namespace

void solve_batch(const std::vector<Instance *> &insts, bool want_gf) {
  Params &P = params();
  const int max_iter = (int)P.geti("max_iter_lanczos");
  const double accur = P.getd("accur_lanczos");
  const int max_cf = (int)P.geti("max_iter_CF");
  const double acc_cf = P.getd("accur_continued_fraction");
  const uint64_t seed = (uint64_t)P.geti("seed");

  std::vector<Instance *> todo;
  for (Instance *I : insts) {
    if (batchable(*I)) todo.push_back(I);
    else I->ground_state_solve();
  }
  vlog("batch: %zu instances batched, %zu solved one by one", todo.size(), insts.size() - todo.size());

  // ---- ground states: one item per (instance, target sector) ----
  if (!todo.empty()) {
    TablePack tabs;
    struct GSRef { Instance *I; Sector s; int tab; uint32_t dim; };
    std::vector<GSRef> refs;
    size_t ws_len = 0, gs_len = 0;
    for (Instance *I : todo) {
      const TermList T = I->terms();
      for (const auto &s : I->targets) {
        const size_t dim = sector_dim(I->model->L, s);
        if (dim == 0) continue;
        if (dim > 0xffffffffull) fail(ED_ERR_UNSUPPORTED, "batched sector too large");
        refs.push_back({I, s, tabs.add(factorized_tables(T, s)), (uint32_t)dim});
        ws_len += 3 * dim;
        gs_len += dim;
      }
    }
    tabs.upload();
    const size_t nab = 2 * (size_t)(max_iter + 2), ntri = 3 * (size_t)(max_iter + 2);
    DevBuf<double> ws(std::max<size_t>(ws_len, 1)), gsv(std::max<size_t>(gs_len, 1)),
        ab(std::max<size_t>(nab * refs.size(), 1)), tri(std::max<size_t>(ntri * refs.size(), 1)),
        out(std::max<size_t>(3 * refs.size(), 1));
    std::vector<BatchGSItem> items(refs.size());
    std::vector<uint32_t> dims(refs.size());
    size_t wo = 0, go = 0;
    for (size_t k = 0; k < refs.size(); ++k) {
      BatchGSItem &it = items[k];
      it.S = tabs.view(refs[k].tab);
      it.seed = seed;
      it.ws = ws.p + wo;
      it.gs = gsv.p + go;
      it.ab = ab.p + nab * k;
      it.tri = tri.p + ntri * k;
      it.out = out.p + 3 * k;
      dims[k] = refs[k].dim;
      wo += 3 * (size_t)refs[k].dim;
      go += refs[k].dim;
    }
    launch_split<BatchGSItem>(items, dims, batch_gs, max_iter, accur);
    const std::vector<double> res = out.to_host();
    vlog("batch: %zu ground-state items solved", refs.size());
    // states per instance, then the instance's own weighting
    std::map<Instance *, std::vector<std::unique_ptr<State>>> cand;
    for (size_t k = 0; k < refs.size(); ++k) {
      if (res[3 * k + 2] != 0.0)
        fail(ED_ERR_NOT_CONVERGED, "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations (instance " +
                                       std::to_string(refs[k].I->label) + ", sector " + refs[k].s.name() + ")");
      auto st = std::make_unique<State>();
      st->sec = refs[k].s;
      st->energy = res[3 * k];
      st->psi.alloc(refs[k].dim);
      ED_HIP(hipMemcpyAsync(st->psi.p, items[k].gs, sizeof(double) * refs[k].dim, hipMemcpyDeviceToDevice, stream()));
      cand[refs[k].I].push_back(std::move(st));
    }
    sync();
    for (Instance *I : todo) {
      I->states = std::move(cand[I]);
      I->finalize_states();
    }
  }
  if (!want_gf) return;

  // ---- Green functions: continued fractions batched, other methods one by one ----
  std::vector<Instance *> cf;
  for (Instance *I : insts) {
    if (I->gf_solved) continue;
    if (I->gf_method == 'F' && std::find(todo.begin(), todo.end(), I) != todo.end()) cf.push_back(I);
    else I->green_function_solve();
  }
  if (cf.empty()) return;
  TablePack tabs;
  std::map<std::pair<Instance *, Sector>, int> tab_of;
  struct MapSet { DevBuf<int32_t> map; DevBuf<int8_t> sgn, upar; };
  std::map<std::tuple<int, int, int, int, int>, std::unique_ptr<MapSet>> maps;  // (L, Nup, Ndw, pm, a)
  auto get_maps = [&](int L, const Sector &src, int pm, int a) -> MapSet & {
    auto key = std::make_tuple(L, src.Nup(), src.Ndw(), pm, a);
    auto it = maps.find(key);
    if (it != maps.end()) return *it->second;
    const CdagMaps C = cdag_maps(L, pm, a, src);
    auto M = std::make_unique<MapSet>();
    M->map.upload(C.map.empty() ? std::vector<int32_t>{-1} : C.map);
    M->sgn.upload(C.sgn.empty() ? std::vector<int8_t>{1} : C.sgn);
    M->upar.upload(C.upar.empty() ? std::vector<int8_t>{1} : C.upar);
    return *(maps[key] = std::move(M));
  };
  struct CFRef { Instance *I; State *st; int spin, pm, idx; uint32_t dim; int tab; };
  std::vector<CFRef> refs;
  std::vector<BatchCFItem> items;
  size_t ws_len = 0;
  for (Instance *I : cf) {
    const int L = I->model->L, ns = I->model->n_sites, np = ns * (ns + 1) / 2;
    const TermList T = I->terms();
    for (auto &stp : I->states) {
      State &st = *stp;
      for (int spin = 0; spin < ((I->mixing & 4) ? 2 : 1); ++spin) {
        GFRep &G = st.gf[spin];
        G = GFRep();
        G.kind = 'F';
        G.n = ns;
        G.ea.assign(np, {});
        G.eb.assign(np, {});
        G.ha.assign(np, {});
        G.hb.assign(np, {});
        for (int pm = -1; pm <= 1; pm += 2) {
          const Sector tgt = shift_sector(st.sec, pm, spin);
          if (!sector_valid(L, tgt)) continue;
          const size_t dt = sector_dim(L, tgt);
          if (dt == 0) continue;
          auto key = std::make_pair(I, tgt);
          auto ti = tab_of.find(key);
          const int tab = ti != tab_of.end() ? ti->second : (tab_of[key] = tabs.add(factorized_tables(T, tgt)));
          const uint32_t du_s = half_basis(L, st.sec.Nup())->dim;
          int idx = 0;
          for (int o1 = 0; o1 < ns; ++o1)
            for (int o2 = 0; o2 <= o1; ++o2, ++idx) {
              BatchCFItem it;
              it.gs = st.psi.p;
              it.du_s = du_s;
              it.spin = spin;
              it.norb = o2 == o1 ? 1 : 2;
              MapSet &m1 = get_maps(L, st.sec, pm, o1 + spin * L);
              it.map[0] = m1.map.p;
              it.sgn[0] = m1.sgn.p;
              it.upar = m1.upar.p;
              if (o2 != o1) {
                MapSet &m2 = get_maps(L, st.sec, pm, o2 + spin * L);
                it.map[1] = m2.map.p;
                it.sgn[1] = m2.sgn.p;
              }
              items.push_back(it);
              refs.push_back({I, &st, spin, pm, idx, (uint32_t)dt, tab});
              ws_len += 3 * dt;
            }
        }
      }
    }
  }
  tabs.upload();
  DevBuf<double> ws(std::max<size_t>(ws_len, 1)), av(std::max<size_t>((size_t)max_cf * refs.size(), 1)),
      bv(std::max<size_t>((size_t)(max_cf + 1) * refs.size(), 1)), out(std::max<size_t>(2 * refs.size(), 1));
  std::vector<uint32_t> dims(refs.size());
  size_t wo = 0;
  for (size_t k = 0; k < refs.size(); ++k) {
    BatchCFItem &it = items[k];
    it.S = tabs.view(refs[k].tab);
    it.ws = ws.p + wo;
    it.a = av.p + (size_t)max_cf * k;
    it.b = bv.p + (size_t)(max_cf + 1) * k;
    it.out = out.p + 2 * k;
    dims[k] = refs[k].dim;
    wo += 3 * (size_t)refs[k].dim;
  }
  launch_split<BatchCFItem>(items, dims, batch_cf, max_cf, acc_cf);
  const std::vector<double> res = out.to_host(), ah = av.to_host(), bh = bv.to_host();
  vlog("batch: %zu continued fractions built", refs.size());
  // continued_fraction transform (src_ed/continued_fraction.cpp:33-42), as build_cf
  for (size_t k = 0; k < refs.size(); ++k) {
    const int n = (int)res[2 * k];
    if (n < 0) continue;
    const CFRef &r = refs[k];
    std::vector<double> a(ah.begin() + (size_t)max_cf * k, ah.begin() + (size_t)max_cf * k + n);
    std::vector<double> b(bh.begin() + (size_t)(max_cf + 1) * k, bh.begin() + (size_t)(max_cf + 1) * k + n + 1);
    for (auto &x : b) x *= x;
    b[0] = res[2 * k + 1] * r.st->weight;
    if (r.pm == 1) {
      for (auto &x : a) x -= r.st->energy;
      r.st->gf[r.spin].ea[r.idx] = std::move(a);
      r.st->gf[r.spin].eb[r.idx] = std::move(b);
    } else {
      for (auto &x : a) x = r.st->energy - x;
      r.st->gf[r.spin].ha[r.idx] = std::move(a);
      r.st->gf[r.spin].hb[r.idx] = std::move(b);
    }
  }
  for (Instance *I : cf) I->gf_solved = true;
}

}  // namespace ed
