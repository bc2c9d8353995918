// ed_hdf5.cpp -- solved model instances to and from HDF5 files, in the layout
// of the reference's writers (ED::write_instance_hdf5 / read_instance_hdf5,
// src_ed/qcm_ED.cpp:373-402; model_instance<T>::write_hdf5 / read_hdf5,
// src_ed/model_instance.hpp:1351-1420; state<T>, src_ed/state.hpp:52-108;
// Q_matrix_set / Q_matrix, src_ed/Q_matrix_set.hpp:187-209,
// src_ed/Q_matrix.hpp:326-345; continued_fraction_set,
// src_ed/continued_fraction_set.cpp:135-176; mcf_set / h5_write_mcf,
// src_ed/mcf_set.hpp:166-195, src_ed/matrix_continued_fraction.hpp:512-547;
// attribute / vector / matrix encodings, src_util/hdf5_io.cpp).
//
// The HDF5 C library is opened at run time (dlopen), so the solver itself has
// no link-time dependency on it; without the library these entry points
// report ED_ERR_UNSUPPORTED.
#include <dlfcn.h>
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ed_instance.hpp"

namespace ed {

namespace {

// the few HDF5 C types and constants used here (HDF5 >= 1.10: 64-bit hid_t)
using hid_t = int64_t;
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;
constexpr hid_t kDefault = 0;         // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr int kScalar = 0;            // H5S_SCALAR
constexpr unsigned kAccRdonly = 0x0u, kAccRdwr = 0x1u, kAccTrunc = 0x2u;
constexpr size_t kVariable = (size_t)-1;  // H5T_VARIABLE

struct H5 {
  void *so = nullptr;
  herr_t (*open)();
  herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
  herr_t (*eset_auto)(hid_t, void *, void *);
  hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
  hid_t (*fopen)(const char *, unsigned, hid_t);
  herr_t (*fclose)(hid_t);
  htri_t (*lexists)(hid_t, const char *, hid_t);
  herr_t (*ldelete)(hid_t, const char *, hid_t);
  hid_t (*gcreate)(hid_t, const char *, hid_t, hid_t, hid_t);
  hid_t (*gopen)(hid_t, const char *, hid_t);
  herr_t (*gclose)(hid_t);
  hid_t (*screate)(int);
  hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
  herr_t (*sclose)(hid_t);
  int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
  hid_t (*acreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t);
  hid_t (*aopen)(hid_t, const char *, hid_t);
  htri_t (*aexists)(hid_t, const char *);
  herr_t (*awrite)(hid_t, hid_t, const void *);
  herr_t (*aread)(hid_t, hid_t, void *);
  hid_t (*aget_type)(hid_t);
  herr_t (*aclose)(hid_t);
  hid_t (*tcopy)(hid_t);
  herr_t (*tset_size)(hid_t, size_t);
  size_t (*tget_size)(hid_t);
  htri_t (*tis_vlstr)(hid_t);
  herr_t (*tclose)(hid_t);
  hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
  hid_t (*dopen)(hid_t, const char *, hid_t);
  herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
  hid_t (*dget_space)(hid_t);
  herr_t (*dclose)(hid_t);
  herr_t (*free_memory)(void *);
  hid_t *native_double = nullptr, *native_int = nullptr, *c_s1 = nullptr;
};

template <typename F>
void bind(void *so, const char *name, F &f) {
  f = reinterpret_cast<F>(dlsym(so, name));
  if (!f) fail(ED_ERR_UNSUPPORTED, std::string("HDF5 library lacks ") + name);
}

// the library: $PYQCM_HDF5_LIBRARY, else the usual sonames, else a conda tree
const H5 &h5() {
  static H5 L;
  static std::string err;
  static bool tried = false;
  if (tried) {
    if (!L.so) fail(ED_ERR_UNSUPPORTED, err);
    return L;
  }
  tried = true;
  std::vector<std::string> cand;
  if (const char *e = std::getenv("PYQCM_HDF5_LIBRARY")) cand.push_back(e);
  for (const char *s : {"libhdf5.so", "libhdf5.so.310", "libhdf5.so.200", "libhdf5.so.103",
                        "/opt/conda/lib/libhdf5.so"})
    cand.push_back(s);
  for (const auto &c : cand)
    if ((L.so = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!L.so) {
    err = "no HDF5 C library found (set PYQCM_HDF5_LIBRARY to its path)";
    fail(ED_ERR_UNSUPPORTED, err);
  }
  try {
    bind(L.so, "H5open", L.open);
    bind(L.so, "H5get_libversion", L.get_libversion);
    bind(L.so, "H5Eset_auto2", L.eset_auto);
    bind(L.so, "H5Fcreate", L.fcreate);
    bind(L.so, "H5Fopen", L.fopen);
    bind(L.so, "H5Fclose", L.fclose);
    bind(L.so, "H5Lexists", L.lexists);
    bind(L.so, "H5Ldelete", L.ldelete);
    bind(L.so, "H5Gcreate2", L.gcreate);
    bind(L.so, "H5Gopen2", L.gopen);
    bind(L.so, "H5Gclose", L.gclose);
    bind(L.so, "H5Screate", L.screate);
    bind(L.so, "H5Screate_simple", L.screate_simple);
    bind(L.so, "H5Sclose", L.sclose);
    bind(L.so, "H5Sget_simple_extent_dims", L.sget_dims);
    bind(L.so, "H5Acreate2", L.acreate);
    bind(L.so, "H5Aopen", L.aopen);
    bind(L.so, "H5Aexists", L.aexists);
    bind(L.so, "H5Awrite", L.awrite);
    bind(L.so, "H5Aread", L.aread);
    bind(L.so, "H5Aget_type", L.aget_type);
    bind(L.so, "H5Aclose", L.aclose);
    bind(L.so, "H5Tcopy", L.tcopy);
    bind(L.so, "H5Tset_size", L.tset_size);
    bind(L.so, "H5Tget_size", L.tget_size);
    bind(L.so, "H5Tis_variable_str", L.tis_vlstr);
    bind(L.so, "H5Tclose", L.tclose);
    bind(L.so, "H5Dcreate2", L.dcreate);
    bind(L.so, "H5Dopen2", L.dopen);
    bind(L.so, "H5Dwrite", L.dwrite);
    bind(L.so, "H5Dread", L.dread);
    bind(L.so, "H5Dget_space", L.dget_space);
    bind(L.so, "H5Dclose", L.dclose);
    bind(L.so, "H5free_memory", L.free_memory);
    L.native_double = reinterpret_cast<hid_t *>(dlsym(L.so, "H5T_NATIVE_DOUBLE_g"));
    L.native_int = reinterpret_cast<hid_t *>(dlsym(L.so, "H5T_NATIVE_INT_g"));
    L.c_s1 = reinterpret_cast<hid_t *>(dlsym(L.so, "H5T_C_S1_g"));
    if (!L.native_double || !L.native_int || !L.c_s1) fail(ED_ERR_UNSUPPORTED, "HDF5 library lacks its type ids");
    if (L.open() < 0) fail(ED_ERR_UNSUPPORTED, "H5open failed");
    L.eset_auto(kDefault, nullptr, nullptr);  // errors come back as status codes, not printed
  } catch (const Error &e) {
    err = e.what();
    dlclose(L.so);
    L.so = nullptr;
    throw;
  }
  return L;
}

hid_t ok(hid_t id, const std::string &what) {
  if (id < 0) fail(ED_ERR_INVALID, "HDF5 error: " + what);
  return id;
}

// an HDF5 object id closed on scope exit
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id &) = delete;
  Id(Id &&o) noexcept : id(o.id), close(o.close) { o.id = -1; }
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

Id group_create(hid_t loc, const std::string &name) {
  const H5 &L = h5();
  return Id(ok(L.gcreate(loc, name.c_str(), kDefault, kDefault, kDefault), "creating group " + name), L.gclose);
}
Id group_open(hid_t loc, const std::string &name) {
  const H5 &L = h5();
  return Id(ok(L.gopen(loc, name.c_str(), kDefault), "opening group " + name), L.gclose);
}
bool exists(hid_t loc, const std::string &name) { return h5().lexists(loc, name.c_str(), kDefault) > 0; }

// ---- attributes (h5_write_attr / h5_read_attr_*, src_util/hdf5_io.cpp:11-57) ----
void attr_num(hid_t g, const std::string &name, hid_t type, const void *v) {
  const H5 &L = h5();
  Id sp(ok(L.screate(kScalar), "dataspace"), L.sclose);
  Id a(ok(L.acreate(g, name.c_str(), type, sp, kDefault, kDefault), "creating attribute " + name), L.aclose);
  ok(L.awrite(a, type, v), "writing attribute " + name);
}
void attr(hid_t g, const std::string &name, double v) { attr_num(g, name, *h5().native_double, &v); }
void attr(hid_t g, const std::string &name, int v) { attr_num(g, name, *h5().native_int, &v); }
// variable-length C string, as H5::StrType(PredType::C_S1, H5T_VARIABLE)
void attr(hid_t g, const std::string &name, const std::string &v) {
  const H5 &L = h5();
  Id st(ok(L.tcopy(*L.c_s1), "string type"), L.tclose);
  ok(L.tset_size(st, kVariable), "string type");
  Id sp(ok(L.screate(kScalar), "dataspace"), L.sclose);
  Id a(ok(L.acreate(g, name.c_str(), st, sp, kDefault, kDefault), "creating attribute " + name), L.aclose);
  const char *cp = v.c_str();
  ok(L.awrite(a, st, &cp), "writing attribute " + name);
}
double attr_double(hid_t g, const std::string &name) {
  const H5 &L = h5();
  Id a(ok(L.aopen(g, name.c_str(), kDefault), "opening attribute " + name), L.aclose);
  double v = 0.0;
  ok(L.aread(a, *L.native_double, &v), "reading attribute " + name);
  return v;
}
int attr_int(hid_t g, const std::string &name) {
  const H5 &L = h5();
  Id a(ok(L.aopen(g, name.c_str(), kDefault), "opening attribute " + name), L.aclose);
  int v = 0;
  ok(L.aread(a, *L.native_int, &v), "reading attribute " + name);
  return v;
}
std::string attr_string(hid_t g, const std::string &name) {
  const H5 &L = h5();
  Id a(ok(L.aopen(g, name.c_str(), kDefault), "opening attribute " + name), L.aclose);
  Id ft(ok(L.aget_type(a), "attribute type"), L.tclose);
  Id mt(ok(L.tcopy(*L.c_s1), "string type"), L.tclose);
  if (L.tis_vlstr(ft) > 0) {
    ok(L.tset_size(mt, kVariable), "string type");
    char *p = nullptr;
    ok(L.aread(a, mt, &p), "reading attribute " + name);
    std::string s = p ? p : "";
    if (p) L.free_memory(p);
    return s;
  }
  const size_t n = L.tget_size(ft);
  ok(L.tset_size(mt, n + 1), "string type");
  std::vector<char> buf(n + 1, 0);
  ok(L.aread(a, mt, buf.data()), "reading attribute " + name);
  return std::string(buf.data());
}

// ---- 1-D double datasets and matrices (hdf5_io.cpp:63-145) ----
void vec(hid_t g, const std::string &name, const double *v, size_t n) {
  const H5 &L = h5();
  const hsize_t dims[1] = {n};
  Id sp(ok(L.screate_simple(1, dims, nullptr), "dataspace"), L.sclose);
  Id ds(ok(L.dcreate(g, name.c_str(), *L.native_double, sp, kDefault, kDefault, kDefault), "creating dataset " + name),
        L.dclose);
  if (n) ok(L.dwrite(ds, *L.native_double, kDefault, kDefault, kDefault, v), "writing dataset " + name);
}
void vec(hid_t g, const std::string &name, const std::vector<double> &v) { vec(g, name, v.data(), v.size()); }
std::vector<double> read_vec(hid_t g, const std::string &name) {
  const H5 &L = h5();
  Id ds(ok(L.dopen(g, name.c_str(), kDefault), "opening dataset " + name), L.dclose);
  Id sp(ok(L.dget_space(ds), "dataspace of " + name), L.sclose);
  hsize_t dims[1] = {0};
  if (L.sget_dims(sp, dims, nullptr) != 1) fail(ED_ERR_INVALID, "dataset " + name + " is not one-dimensional");
  std::vector<double> v((size_t)dims[0]);
  if (dims[0]) ok(L.dread(ds, *L.native_double, kDefault, kDefault, kDefault, v.data()), "reading dataset " + name);
  return v;
}
// real matrix r x c (column-major): group with attributes r, c and dataset "data"
void mat(hid_t g, const std::string &name, int r, int c, const double *v) {
  Id mg = group_create(g, name);
  attr(mg, "r", r);
  attr(mg, "c", c);
  vec(mg, "data", v, (size_t)r * c);
}
// complex matrix: datasets "re" and "im"
void cmat(hid_t g, const std::string &name, int r, int c, const double *re) {
  Id mg = group_create(g, name);
  attr(mg, "r", r);
  attr(mg, "c", c);
  vec(mg, "re", re, (size_t)r * c);
  std::vector<double> im((size_t)r * c, 0.0);
  vec(mg, "im", im);
}
std::vector<double> read_mat(hid_t g, const std::string &name, int &r, int &c, bool complex_ok) {
  Id mg = group_open(g, name);
  r = attr_int(mg, "r");
  c = attr_int(mg, "c");
  std::vector<double> v;
  if (exists(mg, "data")) {
    v = read_vec(mg, "data");
  } else if (complex_ok) {
    v = read_vec(mg, "re");
    const std::vector<double> im = read_vec(mg, "im");
    for (double x : im)
      if (x != 0.0) fail(ED_ERR_UNSUPPORTED, "complex weights in an HDF5 solution (real Hamiltonians only)");
  } else {
    fail(ED_ERR_INVALID, "matrix " + name + " has no data");
  }
  if (v.size() != (size_t)r * c) fail(ED_ERR_INVALID, "matrix " + name + " has the wrong size");
  return v;
}

// ---- Green-function representations ----
const char *format_name(char kind) { return kind == 'F' ? "cf" : kind == 'M' ? "mcf" : "bl"; }

// continued fractions (a, b) of pair q: host vectors, or read back from the
// device store (whose b holds b_0 .. b_{n-1}: the unused last b is not kept)
void cf_pair(const GFRep &R, int q, bool hole, std::vector<double> &a, std::vector<double> &b) {
  if (!R.store) {
    a = hole ? R.ha[q] : R.ea[q];
    b = hole ? R.hb[q] : R.eb[q];
    return;
  }
  const CfStore &S = *R.store;
  const int f = R.frac0 + 2 * q + (hole ? 1 : 0);
  if (f + 1 > S.nfrac) fail(ED_ERR_INTERNAL, "continued-fraction store shorter than its Green function");
  int o[2];
  d2h_pageable(o, S.off.p + f, sizeof o);
  a.assign((size_t)(o[1] - o[0]), 0.0);
  b.assign(a.size(), 0.0);
  if (!a.empty()) {
    d2h_pageable(a.data(), S.A.p + o[0], a.size() * sizeof(double));
    d2h_pageable(b.data(), S.B.p + o[0], b.size() * sizeof(double));
  }
}

void write_mcf(hid_t g, const GFRep::MCF &mc, int n) {
  attr(g, "p", mc.p);
  attr(g, "floors", mc.nfl);
  const int pp = mc.p * mc.p;
  for (int j = 0; j < mc.nfl; ++j) {
    Id ag = group_create(g, "A_" + std::to_string(j));
    mat(ag, "data", mc.p, mc.p, mc.A.data() + (size_t)j * pp);
  }
  const int nb = pp ? (int)(mc.B.size() / pp) : 0;
  for (int j = 0; j < nb; ++j) {
    Id bg = group_create(g, "B_" + std::to_string(j));
    mat(bg, "data", mc.p, mc.p, mc.B.data() + (size_t)j * pp);
  }
  Id wg = group_create(g, "W");
  cmat(wg, "data", mc.p, mc.p ? n : 0, mc.W.data());
}

GFRep::MCF read_mcf(hid_t g, int n) {
  GFRep::MCF mc;
  mc.p = attr_int(g, "p");
  mc.nfl = attr_int(g, "floors");
  const int pp = mc.p * mc.p;
  int r, c;
  for (int j = 0; j < mc.nfl; ++j) {
    Id ag = group_open(g, "A_" + std::to_string(j));
    const std::vector<double> A = read_mat(ag, "data", r, c, false);
    if (r != mc.p || c != mc.p) fail(ED_ERR_INVALID, "MCF floor of the wrong size");
    mc.A.insert(mc.A.end(), A.begin(), A.end());
    Id bg = group_open(g, "B_" + std::to_string(j));
    const std::vector<double> B = read_mat(bg, "data", r, c, false);
    if (r != mc.p || c != mc.p) fail(ED_ERR_INVALID, "MCF floor of the wrong size");
    mc.B.insert(mc.B.end(), B.begin(), B.end());
  }
  if (mc.nfl > 0) {
    Id wg = group_open(g, "W");
    mc.W = read_mat(wg, "data", r, c, true);
    if (r != mc.p || c != n) fail(ED_ERR_INVALID, "MCF weight of the wrong size");
  } else {
    mc.p = 0;
  }
  (void)pp;
  return mc;
}

void write_gf(hid_t g, const GFRep &R, int n) {
  if (R.kind == 'L') {
    // Q_matrix_set: one block (trivial point group), Q_matrix L, M, e, v
    attr(g, "nblocks", 1);
    Id bg = group_create(g, "block_0");
    const int M = (int)R.e.size();
    attr(bg, "L", n);
    attr(bg, "M", M);
    vec(bg, "e", R.e);
    mat(bg, "v", n, M, R.V.data());
  } else if (R.kind == 'F') {
    attr(g, "is_complex", 0);
    std::vector<double> a, b;
    for (int o1 = 0; o1 < n; ++o1)
      for (int o2 = 0; o2 <= o1; ++o2) {
        const int q = o1 * (o1 + 1) / 2 + o2;
        Id sg = group_create(g, "r0_a" + std::to_string(o1) + "_b" + std::to_string(o2));
        for (int hole = 0; hole < 2; ++hole) {
          cf_pair(R, q, hole == 1, a, b);
          Id fg = group_create(sg, hole ? "h" : "e");
          vec(fg, "a", a);
          vec(fg, "b", b);
        }
      }
  } else if (R.kind == 'M') {
    attr(g, "nblocks", 1);
    Id bg = group_create(g, "block_0");
    {
      Id eg = group_create(bg, "e");
      write_mcf(eg, R.me, n);
    }
    Id hg = group_create(bg, "h");
    write_mcf(hg, R.mh, n);
  }
}

GFRep read_gf(hid_t g, char kind, int n) {
  GFRep R;
  R.kind = kind;
  R.n = n;
  if (kind == 'L') {
    if (attr_int(g, "nblocks") != 1) fail(ED_ERR_UNSUPPORTED, "point-group blocks in an HDF5 solution");
    Id bg = group_open(g, "block_0");
    const int L = attr_int(bg, "L"), M = attr_int(bg, "M");
    if (L != n) fail(ED_ERR_INVALID, "Q matrix of the wrong dimension");
    R.e = read_vec(bg, "e");
    int r, c;
    R.V = read_mat(bg, "v", r, c, true);
    if ((int)R.e.size() != M || r != n || c != M) fail(ED_ERR_INVALID, "Q matrix of the wrong size");
  } else if (kind == 'F') {
    if (attr_int(g, "is_complex")) fail(ED_ERR_UNSUPPORTED, "complex continued fractions in an HDF5 solution");
    const int np = n * (n + 1) / 2;
    R.ea.assign(np, {});
    R.eb.assign(np, {});
    R.ha.assign(np, {});
    R.hb.assign(np, {});
    for (int o1 = 0; o1 < n; ++o1)
      for (int o2 = 0; o2 <= o1; ++o2) {
        const int q = o1 * (o1 + 1) / 2 + o2;
        Id sg = group_open(g, "r0_a" + std::to_string(o1) + "_b" + std::to_string(o2));
        for (int hole = 0; hole < 2; ++hole) {
          Id fg = group_open(sg, hole ? "h" : "e");
          std::vector<double> a = read_vec(fg, "a"), b = read_vec(fg, "b");
          if (b.size() < a.size()) fail(ED_ERR_INVALID, "continued fraction with fewer b than a");
          (hole ? R.ha : R.ea)[q] = std::move(a);
          (hole ? R.hb : R.eb)[q] = std::move(b);
        }
      }
  } else {
    if (attr_int(g, "nblocks") != 1) fail(ED_ERR_UNSUPPORTED, "point-group blocks in an HDF5 solution");
    Id bg = group_open(g, "block_0");
    {
      Id eg = group_open(bg, "e");
      R.me = read_mcf(eg, n);
    }
    Id hg = group_open(bg, "h");
    R.mh = read_mcf(hg, n);
  }
  return R;
}

// an HDF5 file id closed on scope exit
Id file_open(const std::string &filename, bool write) {
  const H5 &L = h5();
  if (!write) return Id(ok(L.fopen(filename.c_str(), kAccRdonly, kDefault), "opening " + filename), L.fclose);
  struct stat st {};
  if (stat(filename.c_str(), &st) == 0)
    return Id(ok(L.fopen(filename.c_str(), kAccRdwr, kDefault), "opening " + filename), L.fclose);
  return Id(ok(L.fcreate(filename.c_str(), kAccTrunc, kDefault, kDefault), "creating " + filename), L.fclose);
}

}  // namespace

std::string hdf5_version() {
  const H5 &L = h5();
  unsigned a = 0, b = 0, c = 0;
  L.get_libversion(&a, &b, &c);
  return std::to_string(a) + "." + std::to_string(b) + "." + std::to_string(c);
}

// ED::write_instance_hdf5 + model_instance<T>::write_hdf5 + state<T>::write_hdf5
void write_instance_hdf5(const std::string &filename, int label, const std::string &group) {
  Instance &I = get_instance(label);
  I.need_real("the HDF5 solution format");
  h5();
  if (!I.gf_solved) I.green_function_solve();
  Id file = file_open(filename, true);
  if (exists(file, group)) ok(h5().ldelete(file, group.c_str(), kDefault), "removing group " + group);
  Id grp = group_create(file, group);
  attr(grp, "GS_energy", I.GS_energy);
  attr(grp, "GS_sectors", I.gs_string());
  attr(grp, "GF_format", std::string(format_name(I.gf_method)));
  attr(grp, "mixing", I.mixing);
  attr(grp, "complex_HS", 0);
  {
    Id pg = group_create(grp, "params");
    for (const auto &kv : I.value) attr(pg, kv.first, kv.second);
  }
  attr(grp, "nstates", (int)I.states.size());
  const int n = I.gf_dim();
  int idx = 0;
  for (const auto &s : I.states) {
    Id sg = group_create(grp, "state_" + std::to_string(idx++));
    attr(sg, "sec", s->sec.name());
    attr(sg, "energy", s->energy);
    attr(sg, "weight", s->weight);
    attr(sg, "gf_format", std::string(format_name(s->gf[0].kind)));
    {
      Id gg = group_create(sg, "gf");
      write_gf(gg, s->gf[0], n);
    }
    if (I.mixing & 4) {
      Id gg = group_create(sg, "gf_down");
      write_gf(gg, s->gf[1], n);
    }
  }
}

// ED::read_instance_hdf5 + model_instance<T>::read_hdf5 + the state<T> reader:
// the instance takes the file's states and Green functions (no wavefunctions:
// the format holds none) and counts as solved
void read_instance_hdf5(const std::string &filename, int label, const std::string &group) {
  Instance &I = get_instance(label);
  I.need_real("the HDF5 solution format");
  h5();
  Id file = file_open(filename, false);
  Id grp = group_open(file, group);
  if (attr_int(grp, "complex_HS")) fail(ED_ERR_UNSUPPORTED, "complex Hilbert space in an HDF5 solution");
  const std::string fmt = attr_string(grp, "GF_format");
  const char kind = fmt == "bl" ? 'L' : fmt == "mcf" ? 'M' : 'F';
  const int mixing = attr_int(grp, "mixing");
  if (mixing != 0 && mixing != 4) fail(ED_ERR_UNSUPPORTED, "HDF5 solution with mixing " + std::to_string(mixing));
  {
    // parameter check (a warning in the reference)
    Id pg = group_open(grp, "params");
    for (const auto &kv : I.value)
      if (h5().aexists(pg, kv.first.c_str()) > 0) {
        const double stored = attr_double(pg, kv.first);
        if (std::fabs(stored - kv.second) > 1e-6)
          std::fprintf(stderr,
                       "WARNING : The value of %s from the HDF5 solution (%.10g) differs from the expected value "
                       "(%.10g).\n",
                       kv.first.c_str(), stored, kv.second);
      }
  }
  const int n = I.gf_dim();
  const int ns = attr_int(grp, "nstates");
  std::vector<std::unique_ptr<State>> states;
  for (int i = 0; i < ns; ++i) {
    Id sg = group_open(grp, "state_" + std::to_string(i));
    auto st = std::make_unique<State>();
    st->sec = Sector::parse(attr_string(sg, "sec"));
    st->energy = attr_double(sg, "energy");
    st->weight = attr_double(sg, "weight");
    const std::string f = attr_string(sg, "gf_format");
    const char k = f == "bl" ? 'L' : f == "mcf" ? 'M' : 'F';
    if (exists(sg, "gf")) {
      Id gg = group_open(sg, "gf");
      st->gf[0] = read_gf(gg, k, n);
    }
    if (exists(sg, "gf_down")) {
      Id gg = group_open(sg, "gf_down");
      st->gf[1] = read_gf(gg, k, n);
    }
    states.push_back(std::move(st));
  }
  I.states = std::move(states);
  I.mixing = mixing;
  I.gf_method = kind;
  I.GS_energy = attr_double(grp, "GS_energy");
  I.E0 = 0.0;
  for (const auto &s : I.states) I.E0 += s->weight * s->energy;
  I.averages.clear();
  I.averages_done = false;
  I.gs_solved = true;
  I.gf_solved = true;
}

}  // namespace ed
