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

// packed device copy of many FactorizedHost tables
struct TablePack {
  std::vector<uint32_t> u32;
  std::vector<double> f64;
  struct Off { size_t up_ptr, up_col, dn_ptr, dn_col, up_val, dn_val, diag; uint32_t du, dd, nu, nd; };
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
    o.nu = (uint32_t)F.up_col.size();
    o.nd = (uint32_t)F.dn_col.size();
    offs.push_back(o);
    return (int)offs.size() - 1;
  }
  void upload() {
    du32.upload(u32);
    df64.upload(f64);
  }
  // LDS bytes of an item of this sector held in LDS
  size_t lds(int k, int nvec, size_t ntri) const {
    const Off &o = offs[k];
    return batch_lds_bytes(nvec, ntri, o.du, o.dd, o.nu, o.nd);
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

// execution class of an item: from LDS with one or four waves, global
// vectors with one or four waves
enum Cls { kLds = 0, kLdsWide = 1, kWave = 2, kBlock = 3 };
constexpr int kClasses = 4;

// items held in LDS with at least this many states get four waves (the
// steps are latency-bound; more waves per item hide more of it)
constexpr uint32_t kWideDim = 128;

// items whose sector has at most this many states and does not fit in LDS
// run in a one-wave workgroup with global vectors
constexpr uint32_t kOneWaveDim = 4096;

Cls classify(size_t lds, uint32_t dim) {
  if (lds <= kBatchLdsCap) return dim >= kWideDim ? kLdsWide : kLds;
  return dim <= kOneWaveDim ? kWave : kBlock;
}
bool in_lds(Cls c) { return c == kLds || c == kLdsWide; }

// one launch per non-empty class
template <typename Item, typename Launch>
void launch_classes(const std::vector<Item> &items, const std::vector<Cls> &cls, size_t lds_max, Launch launch) {
  std::vector<Item> part[kClasses];
  for (size_t k = 0; k < items.size(); ++k) part[cls[k]].push_back(items[k]);
  DevBuf<Item> d[kClasses];
  for (int c = 0; c < kClasses; ++c) {
    if (part[c].empty()) continue;
    d[c].upload(part[c]);
    launch(d[c].p, (int)part[c].size(), (c == kBlock || c == kLdsWide) ? 256 : 64, in_lds((Cls)c) ? lds_max : 0);
  }
  sync();
}

// The one-body tables and the diagonal are linear in the term values: the
// tables of every operator of a term list at unit value, merged on the union
// of their patterns, give each instance's tables as a weighted sum.
struct UnitTables {
  FactorizedHost pat;                             // union pattern (values unused)
  std::vector<std::vector<double>> up, dn, diag;  // per operator, on pat
};

UnitTables unit_tables(const TermList &T, const Sector &sec) {
  UnitTables U;
  std::vector<FactorizedHost> F;
  for (const auto &tc : T.terms) {
    TermList t1;
    t1.L = T.L;
    t1.terms = {{tc.first, 1.0}};
    F.push_back(factorized_tables(t1, sec));
  }
  U.pat.du = F[0].du;
  U.pat.dd = F[0].dd;
  // union of the rows of one spin, then every operator's values on it
  auto merge = [&](bool up, std::vector<uint32_t> &ptr, std::vector<uint32_t> &col,
                   std::vector<std::vector<double>> &val) {
    const uint32_t nrow = up ? U.pat.du : U.pat.dd;
    ptr.assign(1, 0);
    std::vector<uint32_t> row;
    for (uint32_t r = 0; r < nrow; ++r) {
      row.clear();
      for (const FactorizedHost &f : F) {
        const auto &fp = up ? f.up_ptr : f.dn_ptr;
        const auto &fc = up ? f.up_col : f.dn_col;
        row.insert(row.end(), fc.begin() + fp[r], fc.begin() + fp[r + 1]);
      }
      std::sort(row.begin(), row.end());
      row.erase(std::unique(row.begin(), row.end()), row.end());
      col.insert(col.end(), row.begin(), row.end());
      ptr.push_back((uint32_t)col.size());
    }
    val.assign(F.size(), std::vector<double>(col.size(), 0.0));
    for (size_t k = 0; k < F.size(); ++k) {
      const auto &fp = up ? F[k].up_ptr : F[k].dn_ptr;
      const auto &fc = up ? F[k].up_col : F[k].dn_col;
      const auto &fv = up ? F[k].up_val : F[k].dn_val;
      for (uint32_t r = 0; r < nrow; ++r)
        for (uint32_t e = fp[r]; e < fp[r + 1]; ++e) {
          const auto b = col.begin() + ptr[r], en = col.begin() + ptr[r + 1];
          val[k][std::lower_bound(b, en, fc[e]) - col.begin()] += fv[e];
        }
    }
  };
  merge(true, U.pat.up_ptr, U.pat.up_col, U.up);
  merge(false, U.pat.dn_ptr, U.pat.dn_col, U.dn);
  for (const FactorizedHost &f : F) U.diag.push_back(f.diag);
  return U;
}

// the tables of term list T in sector sec (same operators, in the same
// order, as the list the unit tables were built from)
FactorizedHost combine_tables(const UnitTables &U, const TermList &T) {
  FactorizedHost F = U.pat;
  auto sum = [&](const std::vector<std::vector<double>> &u, std::vector<double> &out) {
    out.assign(u[0].size(), 0.0);
    for (size_t k = 0; k < u.size(); ++k) {
      const double v = T.terms[k].second;
      if (v == 0.0) continue;
      for (size_t i = 0; i < out.size(); ++i) out[i] += v * u[k][i];
    }
  };
  sum(U.up, F.up_val);
  sum(U.dn, F.dn_val);
  sum(U.diag, F.diag);
  return F;
}

// unit tables per (operator list, sector), built once per batched solve
struct TableCache {
  std::map<std::pair<std::vector<const Operator *>, Sector>, UnitTables> m;
  FactorizedHost get(const TermList &T, const Sector &sec) {
    if (T.terms.empty()) return factorized_tables(T, sec);
    std::vector<const Operator *> ops;
    for (const auto &tc : T.terms) ops.push_back(tc.first);
    auto key = std::make_pair(std::move(ops), sec);
    auto it = m.find(key);
    if (it == m.end()) it = m.emplace(key, unit_tables(T, sec)).first;
    return combine_tables(it->second, T);
  }
};

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
  check_cancel();
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

  TableCache cache;
  // ---- ground states: one item per (instance, target sector) ----
  if (!todo.empty()) {
    TablePack tabs;
    struct GSRef { Instance *I; Sector s; int tab; uint32_t dim; };
    std::vector<GSRef> refs;
    size_t gs_len = 0;
    for (Instance *I : todo) {
      const TermList T = I->terms();
      for (const auto &s : I->targets) {
        const size_t dim = sector_dim(I->model->L, s);
        if (dim == 0) continue;
        if (dim > 0xffffffffull) fail(ED_ERR_UNSUPPORTED, "batched sector too large");
        refs.push_back({I, s, tabs.add(cache.get(T, s)), (uint32_t)dim});
        gs_len += dim;
      }
    }
    tabs.upload();
    vlog("batch: %zu ground-state tables built and uploaded", refs.size());
    const size_t ntri = (size_t)kTriArrays * (max_iter + 2);
    std::vector<Cls> cls(refs.size());
    size_t lds_max = 0, ws_len = 0, ntri_glob = 0;
    for (size_t k = 0; k < refs.size(); ++k) {
      const size_t b = tabs.lds(refs[k].tab, kGsVectors, ntri);
      cls[k] = classify(b, refs[k].dim);
      if (in_lds(cls[k])) {
        lds_max = std::max(lds_max, b);
      } else {
        ws_len += 2 * (size_t)refs[k].dim;
        ++ntri_glob;
      }
    }
    // the ground states stay here; each state's vector is a view kept alive by it
    auto gsv = std::make_shared<DevBuf<double>>(std::max<size_t>(gs_len, 1));
    DevBuf<double> ws(std::max<size_t>(ws_len, 1)), tri(std::max<size_t>(ntri * ntri_glob, 1)),
        out(std::max<size_t>(3 * refs.size(), 1));
    std::vector<BatchGSItem> items(refs.size());
    size_t wo = 0, go = 0, to = 0;
    for (size_t k = 0; k < refs.size(); ++k) {
      BatchGSItem &it = items[k];
      it.S = tabs.view(refs[k].tab);
      it.seed = seed;
      it.gs = gsv->p + go;
      it.out = out.p + 3 * k;
      go += refs[k].dim;
      if (!in_lds(cls[k])) {
        it.ws = ws.p + wo;
        it.tri = tri.p + ntri * to++;
        wo += 2 * (size_t)refs[k].dim;
      }
    }
    launch_classes(items, cls, lds_max, [&](const BatchGSItem *d, int n, int block, size_t lds) {
      batch_gs(d, n, block, lds, max_iter, accur, stream());
    });
    const std::vector<double> res = out.to_host();
    if (verbose()) {
      double sum = 0.0, mx = 0.0;
      for (size_t k = 0; k < refs.size(); ++k) {
        sum += res[3 * k + 1];
        mx = std::max(mx, res[3 * k + 1]);
      }
      vlog("batch: %zu ground-state items solved (Lanczos steps: mean %.1f, max %.0f)", refs.size(),
           refs.empty() ? 0.0 : sum / refs.size(), mx);
    }
    // states per instance, then the instance's own weighting
    std::map<Instance *, std::vector<std::unique_ptr<State>>> cand;
    for (size_t k = 0; k < refs.size(); ++k) {
      if (res[3 * k + 2] != 0.0)
        fail(ED_ERR_NOT_CONVERGED, "Lanczos procedure exceeded " + std::to_string(max_iter) + " iterations (instance " +
                                       std::to_string(refs[k].I->label) + ", sector " + refs[k].s.name() + ")");
      auto st = std::make_unique<State>();
      st->sec = refs[k].s;
      st->energy = res[3 * k];
      st->psi = DevBuf<double>::view(items[k].gs, refs[k].dim, gsv);
      cand[refs[k].I].push_back(std::move(st));
    }
    sync();
    vlog("batch: ground states copied out");
    for (Instance *I : todo) {
      I->states = std::move(cand[I]);
      I->finalize_states();
    }
    vlog("batch: states finalized");
  }
  if (!want_gf) return;
  check_cancel();

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
  struct CFRef { Instance *I; State *st; int spin, pm, idx; uint32_t dim; int tab; int fid; };
  std::vector<CFRef> refs;
  std::vector<BatchCFItem> items;
  auto store = std::make_shared<CfStore>();
  int nfrac = 0;
  for (Instance *I : cf) {
    const int L = I->model->L, ns = I->model->n_sites;
    const TermList T = I->terms();
    for (auto &stp : I->states) {
      State &st = *stp;
      for (int spin = 0; spin < ((I->mixing & 4) ? 2 : 1); ++spin) {
        GFRep &G = st.gf[spin];
        G = GFRep();
        G.kind = 'F';
        G.n = ns;
        G.store = store;
        G.frac0 = nfrac;
        nfrac += ns * (ns + 1);
        for (int pm = -1; pm <= 1; pm += 2) {
          const Sector tgt = shift_sector(st.sec, pm, spin);
          if (!sector_valid(L, tgt)) continue;
          const size_t dt = sector_dim(L, tgt);
          if (dt == 0) continue;
          auto key = std::make_pair(I, tgt);
          auto ti = tab_of.find(key);
          const int tab = ti != tab_of.end() ? ti->second : (tab_of[key] = tabs.add(cache.get(T, tgt)));
          const uint32_t du_s = half_basis(L, st.sec.Nup())->dim;
          std::vector<MapSet *> mo(ns);
          for (int o = 0; o < ns; ++o) mo[o] = &get_maps(L, st.sec, pm, o + spin * L);
          int idx = 0;
          for (int o1 = 0; o1 < ns; ++o1)
            for (int o2 = 0; o2 <= o1; ++o2, ++idx) {
              BatchCFItem it;
              it.gs = st.psi.p;
              it.du_s = du_s;
              it.spin = spin;
              it.norb = o2 == o1 ? 1 : 2;
              MapSet &m1 = *mo[o1];
              it.map[0] = m1.map.p;
              it.sgn[0] = m1.sgn.p;
              it.upar = m1.upar.p;
              if (o2 != o1) {
                MapSet &m2 = *mo[o2];
                it.map[1] = m2.map.p;
                it.sgn[1] = m2.sgn.p;
              }
              items.push_back(it);
              refs.push_back({I, &st, spin, pm, idx, (uint32_t)dt, tab, G.frac0 + 2 * idx + (pm == 1 ? 0 : 1)});
            }
        }
      }
    }
  }
  tabs.upload();
  vlog("batch: %zu continued-fraction items, tables and maps built", refs.size());
  std::vector<Cls> cls(refs.size());
  size_t lds_max = 0, ws_len = 0;
  for (size_t k = 0; k < refs.size(); ++k) {
    const size_t b = tabs.lds(refs[k].tab, kCfVectors, 0);
    cls[k] = classify(b, refs[k].dim);
    if (in_lds(cls[k])) lds_max = std::max(lds_max, b);
    else ws_len += 2 * (size_t)refs[k].dim;
  }
  DevBuf<double> ws(std::max<size_t>(ws_len, 1)), av(std::max<size_t>((size_t)max_cf * refs.size(), 1)),
      bv(std::max<size_t>((size_t)(max_cf + 1) * refs.size(), 1)), out(std::max<size_t>(2 * refs.size(), 1));
  size_t wo = 0;
  for (size_t k = 0; k < refs.size(); ++k) {
    BatchCFItem &it = items[k];
    it.S = tabs.view(refs[k].tab);
    it.a = av.p + (size_t)max_cf * k;
    it.b = bv.p + (size_t)(max_cf + 1) * k;
    it.out = out.p + 2 * k;
    if (!in_lds(cls[k])) {
      it.ws = ws.p + wo;
      wo += 2 * (size_t)refs[k].dim;
    }
  }
  launch_classes(items, cls, lds_max, [&](const BatchCFItem *d, int n, int block, size_t lds) {
    batch_cf(d, n, block, lds, max_cf, acc_cf, stream());
  });
  const std::vector<double> res = out.to_host();
  // the store: fraction f holds n_f coefficients (0 where psi vanished or no
  // target sector exists)
  std::vector<int> len(nfrac + 1, 0), fid(refs.size());
  std::vector<double> e0(refs.size()), wt(refs.size());
  std::vector<int8_t> pmv(refs.size());
  for (size_t k = 0; k < refs.size(); ++k) {
    const CFRef &r = refs[k];
    len[r.fid] = std::max(0, (int)res[2 * k]);
    fid[k] = r.fid;
    e0[k] = r.st->energy;
    wt[k] = r.st->weight;
    pmv[k] = (int8_t)r.pm;
  }
  std::vector<int> off(nfrac + 1, 0);
  size_t total = 0;
  for (int f = 0; f < nfrac; ++f) {
    off[f] = (int)total;
    total += (size_t)len[f];
    if (total > 0x7fffffffull) fail(ED_ERR_UNSUPPORTED, "batched continued fractions: too many coefficients");
  }
  off[nfrac] = (int)total;
  store->nfrac = nfrac;
  store->off.upload(off);
  store->A.alloc(std::max<size_t>(total, 1));
  store->B.alloc(std::max<size_t>(total, 1));
  DevBuf<int> dfid;
  DevBuf<double> de0, dwt;
  DevBuf<int8_t> dpm;
  dfid.upload(fid);
  de0.upload(e0);
  dwt.upload(wt);
  dpm.upload(pmv);
  batch_cf_store((int)refs.size(), max_cf, out.p, av.p, bv.p, dfid.p, store->off.p, de0.p, dwt.p, dpm.p,
                 store->A.p, store->B.p, stream());
  sync();
  vlog("batch: %zu continued fractions built and stored on the device (%zu coefficients)", refs.size(), total);
  for (Instance *I : cf) I->gf_solved = true;
  vlog("batch: continued fractions stored");
}

}  // namespace ed
