// libgwhip_h5: host-side reader of green-mbpt's integral files and symmetry tables (HDF5), feeding a
// libgwhip handle.  Counterpart of df_integral_t (src/green/mbpt/df_integral_t.h:40-187) for a
// device that keeps every k pair resident: each stored chunk is read once, every pair k1 >= k2 is
// expanded with the pair-symmetry lists, and the packed array is uploaded in one call.
#include "gwhip_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct h5_error : std::runtime_error {
  int status;
  h5_error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// Owns one HDF5 identifier.
class Id {
public:
  Id(hid_t id, herr_t (*close)(hid_t)) : id_(id), close_(close) {}
  ~Id() {
    if (id_ >= 0) close_(id_);
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  hid_t get() const { return id_; }

private:
  hid_t id_;
  herr_t (*close_)(hid_t);
};

void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

hid_t open_file(const std::string& path) {
  quiet();
  const hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw h5_error(GWHIP_ERR_IO, "cannot open " + path);
  return f;
}

// Whether every component of `path` (relative to the file root) exists.
bool has_path(hid_t f, const std::string& path) {
  size_t pos = 0;
  while (true) {
    const size_t next = path.find('/', pos);
    const std::string part = path.substr(0, next);
    if (H5Lexists(f, part.c_str(), H5P_DEFAULT) <= 0) return false;
    if (next == std::string::npos) return true;
    pos = next + 1;
  }
}

// Dataset -> dims and a flat read in the given memory type.
std::vector<hsize_t> dims_of(hid_t f, const std::string& name) {
  Id ds(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (ds.get() < 0) throw h5_error(GWHIP_ERR_IO, "missing dataset " + name);
  Id sp(H5Dget_space(ds.get()), H5Sclose);
  const int nd = H5Sget_simple_extent_ndims(sp.get());
  std::vector<hsize_t> d(nd > 0 ? nd : 0);
  if (nd > 0) H5Sget_simple_extent_dims(sp.get(), d.data(), nullptr);
  return d;
}

size_t count_of(const std::vector<hsize_t>& d) {
  size_t n = 1;
  for (hsize_t v : d) n *= (size_t)v;
  return n;
}

void read_into(hid_t f, const std::string& name, hid_t memtype, void* buf, size_t capacity) {
  const size_t n = count_of(dims_of(f, name));
  if (n > capacity)
    throw h5_error(GWHIP_ERR_IO, name + " holds " + std::to_string(n) + " elements, expected at most " +
                                     std::to_string(capacity));
  Id ds(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (H5Dread(ds.get(), memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0)
    throw h5_error(GWHIP_ERR_IO, "cannot read " + name);
}

std::vector<int64_t> read_i64(hid_t f, const std::string& name) {
  std::vector<int64_t> v(count_of(dims_of(f, name)));
  read_into(f, name, H5T_NATIVE_INT64, v.data(), v.size());
  return v;
}

int64_t read_scalar_i64(hid_t f, const std::string& name) {
  const auto v = read_i64(f, name);
  if (v.size() != 1) throw h5_error(GWHIP_ERR_IO, name + " is not a scalar");
  return v[0];
}

// Complex128 dataset written as a two-field compound (h5py uses fields "r", "i").
void read_complex(hid_t f, const std::string& name, std::complex<double>* out, size_t count) {
  Id ds(H5Dopen2(f, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (ds.get() < 0) throw h5_error(GWHIP_ERR_IO, "missing dataset " + name);
  Id ft(H5Dget_type(ds.get()), H5Tclose);
  if (H5Tget_class(ft.get()) != H5T_COMPOUND || H5Tget_nmembers(ft.get()) != 2)
    throw h5_error(GWHIP_ERR_IO, name + " is not a complex (two-field compound) dataset");
  char* n0 = H5Tget_member_name(ft.get(), 0);
  char* n1 = H5Tget_member_name(ft.get(), 1);
  Id mt(H5Tcreate(H5T_COMPOUND, 2 * sizeof(double)), H5Tclose);
  H5Tinsert(mt.get(), n0, 0, H5T_NATIVE_DOUBLE);
  H5Tinsert(mt.get(), n1, sizeof(double), H5T_NATIVE_DOUBLE);
  H5free_memory(n0);
  H5free_memory(n1);
  const size_t n = count_of(dims_of(f, name));
  if (n != count) throw h5_error(GWHIP_ERR_IO, name + " has " + std::to_string(n) + " elements, expected " + std::to_string(count));
  if (H5Dread(ds.get(), mt.get(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    throw h5_error(GWHIP_ERR_IO, "cannot read " + name);
}

// Root attribute as a string (fixed or variable length); empty when absent.
std::string root_string_attr(hid_t f, const char* name) {
  if (H5Aexists(f, name) <= 0) return {};
  Id a(H5Aopen(f, name, H5P_DEFAULT), H5Aclose);
  Id ft(H5Aget_type(a.get()), H5Tclose);
  if (H5Tget_class(ft.get()) != H5T_STRING) throw h5_error(GWHIP_ERR_IO, std::string("attribute ") + name + " is not a string");
  Id mt(H5Tcopy(H5T_C_S1), H5Tclose);
  H5Tset_cset(mt.get(), H5Tget_cset(ft.get()));  // h5py writes UTF-8; HDF5 converts no character sets
  if (H5Tis_variable_str(ft.get()) > 0) {
    H5Tset_size(mt.get(), H5T_VARIABLE);
    char* s = nullptr;
    if (H5Aread(a.get(), mt.get(), &s) < 0) throw h5_error(GWHIP_ERR_IO, std::string("cannot read attribute ") + name);
    std::string out = s ? s : "";
    H5free_memory(s);
    return out;
  }
  const size_t len = H5Tget_size(ft.get());
  H5Tset_size(mt.get(), len + 1);
  std::vector<char> buf(len + 1, '\0');
  if (H5Aread(a.get(), mt.get(), buf.data()) < 0) throw h5_error(GWHIP_ERR_IO, std::string("cannot read attribute ") + name);
  return std::string(buf.data());
}

int version_cmp(const char* v) {
  int a = 0, b = 0, c = 0;
  if (!v || std::sscanf(v, "%d.%d.%d", &a, &b, &c) < 3) return -1;
  const int ref[3] = {1, 0, 0};
  const int in[3] = {a, b, c};
  for (int i = 0; i < 3; ++i)
    if (in[i] != ref[i]) return in[i] > ref[i] ? 1 : 0;
  return 1;
}

int64_t chunk_size_of(const std::string& df_path) {
  Id f(open_file(df_path + "/meta.h5"), H5Fclose);
  const std::string v = root_string_attr(f.get(), "__green_version__");
  const std::string fix = " (integral files at '" + df_path + "' are outdated; migrate them to format 1.0.0)";
  if (v.empty()) throw h5_error(GWHIP_ERR_OUTDATED_INPUT, "no __green_version__ attribute" + fix);
  const int c = version_cmp(v.c_str());
  if (c < 0) throw h5_error(GWHIP_ERR_IO, "malformed version string '" + v + "'");
  if (c == 0) throw h5_error(GWHIP_ERR_OUTDATED_INPUT, "format " + v + fix);
  const int64_t cs = read_scalar_i64(f.get(), "chunk_size");
  if (cs <= 0) throw h5_error(GWHIP_ERR_IO, "chunk_size must be positive");
  return cs;
}

struct Pairs {
  std::vector<int64_t> conj, trans, irre;  // irre: stored pair indices in storage order
};

Pairs read_pairs(const std::string& input_h5, int64_t nk) {
  const int64_t npair = nk * (nk + 1) / 2;
  Pairs p;
  Id f(open_file(input_h5), H5Fclose);
  if (has_path(f.get(), "symmetry/pairs")) {
    p.conj = read_i64(f.get(), "symmetry/pairs/conj_pairs_list");
    p.trans = read_i64(f.get(), "symmetry/pairs/trans_pairs_list");
    p.irre = read_i64(f.get(), "symmetry/pairs/kpair_irre_list");
    if ((int64_t)p.conj.size() != npair || (int64_t)p.trans.size() != npair)
      throw h5_error(GWHIP_ERR_IO, "pair lists do not have nk(nk+1)/2 = " + std::to_string(npair) + " entries");
  } else {  // no symmetry: every pair stored in order
    p.conj.resize(npair);
    p.trans.resize(npair);
    p.irre.resize(npair);
    for (int64_t i = 0; i < npair; ++i) p.conj[i] = p.trans[i] = p.irre[i] = i;
  }
  for (int64_t i = 0; i < npair; ++i)
    if (p.conj[i] < 0 || p.conj[i] >= npair || p.trans[i] < 0 || p.trans[i] >= npair)
      throw h5_error(GWHIP_ERR_IO, "pair list entry out of range");
  return p;
}

void read_packed(const std::string& df_path, const std::string& input_h5, int64_t nk, int64_t nao, int64_t NQ,
                 std::complex<double>* out) {
  if (nk <= 0 || nao <= 0 || NQ <= 0) throw h5_error(GWHIP_ERR_ARG, "nk, nao and NQ must be positive");
  const int64_t cs = chunk_size_of(df_path);
  const Pairs p = read_pairs(input_h5, nk);
  const int64_t npair = nk * (nk + 1) / 2, nirre = (int64_t)p.irre.size();
  const int64_t row = NQ * nao * nao;
  // position of each stored pair in the storage order (irre_pos_kpair)
  std::vector<int64_t> pos(npair, -1);
  for (int64_t i = 0; i < nirre; ++i) {
    if (p.irre[i] < 0 || p.irre[i] >= npair) throw h5_error(GWHIP_ERR_IO, "kpair_irre_list entry out of range");
    pos[p.irre[i]] = i;
  }
  // every stored row, chunk by chunk (read_a_chunk: the file's data fill a zeroed chunk buffer)
  std::vector<std::complex<double>> stored((size_t)nirre * row, std::complex<double>(0.0, 0.0));
  for (int64_t c0 = 0; c0 < nirre; c0 += cs) {
    const std::string name = std::to_string(c0);
    Id f(open_file(df_path + "/VQ_" + name + ".h5"), H5Fclose);
    const int64_t rows = std::min(cs, nirre - c0);
    const auto d = dims_of(f.get(), "/" + name);
    if (d.size() != 4 || (int64_t)d[2] != nao || (int64_t)d[3] != 2 * nao || (int64_t)d[0] > cs ||
        (int64_t)d[1] > NQ || ((int64_t)d[1] < NQ && d[0] > 1))
      throw h5_error(GWHIP_ERR_IO, "VQ_" + name + ".h5:/" + name + " is not [chunk][NQ][nao][2 nao] for nao = " +
                                       std::to_string(nao) + ", NQ = " + std::to_string(NQ));
    std::vector<std::complex<double>> buf((size_t)cs * row, std::complex<double>(0.0, 0.0));
    read_into(f.get(), "/" + name, H5T_NATIVE_DOUBLE, buf.data(), buf.size() * 2);
    std::memcpy(stored.data() + (size_t)c0 * row, buf.data(), (size_t)rows * row * sizeof(std::complex<double>));
  }
  // expand (read_integrals + symmetrize for k1 >= k2: sign +1)
  const int64_t n2 = nao * nao;
  for (int64_t k1 = 0; k1 < nk; ++k1)
    for (int64_t k2 = 0; k2 <= k1; ++k2) {
      const int64_t idx0 = k1 * (k1 + 1) / 2 + k2;
      int64_t idx = idx0;
      int op = 0;  // 0 direct, 1 conjugated, 2 transposed
      if (p.conj[idx] != idx) {
        idx = p.conj[idx];
        op = 1;
      } else if (p.trans[idx] != idx) {
        idx = p.trans[idx];
        op = 2;
      }
      if (pos[idx] < 0) throw h5_error(GWHIP_ERR_IO, "k pair " + std::to_string(idx) + " is not stored");
      const std::complex<double>* src = stored.data() + (size_t)pos[idx] * row;
      std::complex<double>* dst = out + (size_t)idx0 * row;
      for (int64_t Q = 0; Q < NQ; ++Q)
        for (int64_t i = 0; i < nao; ++i)
          for (int64_t j = 0; j < nao; ++j) {
            const std::complex<double> v = op == 2 ? src[Q * n2 + j * nao + i] : src[Q * n2 + i * nao + j];
            dst[Q * n2 + i * nao + j] = op == 1 ? std::conj(v) : v;
          }
    }
}

void kq_maps(const double* x, int64_t nk, int32_t* kp, int32_t* km) {
  auto find = [&](const double* y) -> int32_t {
    int32_t hit = -1;
    for (int64_t k = 0; k < nk; ++k) {
      bool same = true;
      for (int d = 0; d < 3 && same; ++d) {
        double diff = x[k * 3 + d] - y[d];
        diff -= std::nearbyint(diff);
        same = std::fabs(diff) < 1e-6;
      }
      if (same) {
        if (hit >= 0) throw h5_error(GWHIP_ERR_IO, "two mesh points coincide");
        hit = (int32_t)k;
      }
    }
    if (hit < 0) throw h5_error(GWHIP_ERR_IO, "k +- q is not a mesh point (the mesh is not closed under addition)");
    return hit;
  };
  for (int64_t q = 0; q < nk; ++q)
    for (int64_t k = 0; k < nk; ++k) {
      double yp[3], ym[3];
      for (int d = 0; d < 3; ++d) {
        yp[d] = x[k * 3 + d] + x[q * 3 + d];
        ym[d] = x[k * 3 + d] - x[q * 3 + d];
      }
      kp[q * nk + k] = find(yp);
      km[q * nk + k] = find(ym);
    }
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return GWHIP_OK;
  } catch (const h5_error& e) {
    g_err = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return GWHIP_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return GWHIP_ERR_IO;
  }
}

}  // namespace

extern "C" {

int gwhip_h5_check_version(const char* version) { return version_cmp(version); }

int gwhip_h5_read_meta(const char* df_path, int64_t* chunk_size) {
  if (!df_path || !chunk_size) return GWHIP_ERR_ARG;
  return guarded([&] { *chunk_size = chunk_size_of(df_path); });
}

int gwhip_h5_read_packed(const char* df_path, const char* input_h5, int64_t nk, int64_t nao, int64_t NQ,
                         double* V_packed) {
  if (!df_path || !input_h5 || !V_packed) return GWHIP_ERR_ARG;
  return guarded([&] { read_packed(df_path, input_h5, nk, nao, NQ, reinterpret_cast<std::complex<double>*>(V_packed)); });
}

int gwhip_h5_read_k_symmetry(const char* input_h5, int64_t nk, int64_t nso, int64_t* ink, int64_t* bz2ibz,
                             int64_t* ibz2bz, double* U, int64_t* tr_conj, double* mesh_scaled) {
  if (!input_h5 || !ink || !bz2ibz || !ibz2bz || !U || !tr_conj || !mesh_scaled) return GWHIP_ERR_ARG;
  return guarded([&] {
    Id f(open_file(input_h5), H5Fclose);
    const std::string k = "symmetry/k/";
    *ink = read_scalar_i64(f.get(), k + "ink");
    if (*ink < 1 || *ink > nk) throw h5_error(GWHIP_ERR_IO, "ink out of range");
    const auto b = read_i64(f.get(), k + "bz2ibz");
    const auto i = read_i64(f.get(), k + "ibz2bz");
    const auto t = read_i64(f.get(), k + "tr_conj");
    if ((int64_t)b.size() != nk || (int64_t)t.size() != nk || (int64_t)i.size() != *ink)
      throw h5_error(GWHIP_ERR_IO, "symmetry/k tables do not match nk = " + std::to_string(nk));
    std::copy(b.begin(), b.end(), bz2ibz);
    std::copy(i.begin(), i.end(), ibz2bz);
    std::copy(t.begin(), t.end(), tr_conj);
    read_complex(f.get(), k + "k_sym_transform_ao", reinterpret_cast<std::complex<double>*>(U), (size_t)(nk * nso * nso));
    read_into(f.get(), k + "mesh_scaled", H5T_NATIVE_DOUBLE, mesh_scaled, (size_t)nk * 3);
    if (count_of(dims_of(f.get(), k + "mesh_scaled")) != (size_t)nk * 3)
      throw h5_error(GWHIP_ERR_IO, "mesh_scaled is not [nk][3]");
  });
}

int gwhip_h5_read_q_symmetry(const char* input_h5, int64_t nq, int64_t NQ, int64_t* inq, int64_t* bz2ibz,
                             int64_t* ibz2bz, double* U, int64_t* tr_conj) {
  if (!input_h5 || !inq || !bz2ibz || !ibz2bz || !U || !tr_conj) return GWHIP_ERR_ARG;
  return guarded([&] {
    Id f(open_file(input_h5), H5Fclose);
    const std::string q = "symmetry/q/";
    *inq = read_scalar_i64(f.get(), q + "inq");
    if (*inq < 1 || *inq > nq) throw h5_error(GWHIP_ERR_IO, "inq out of range");
    const auto b = read_i64(f.get(), q + "bz2ibz");
    const auto i = read_i64(f.get(), q + "ibz2bz");
    const auto t = read_i64(f.get(), q + "tr_conj");
    if ((int64_t)b.size() != nq || (int64_t)t.size() != nq || (int64_t)i.size() != *inq)
      throw h5_error(GWHIP_ERR_IO, "symmetry/q tables do not match nq = " + std::to_string(nq));
    std::copy(b.begin(), b.end(), bz2ibz);
    std::copy(i.begin(), i.end(), ibz2bz);
    std::copy(t.begin(), t.end(), tr_conj);
    read_complex(f.get(), q + "k_sym_transform_p0", reinterpret_cast<std::complex<double>*>(U), (size_t)(nq * NQ * NQ));
  });
}

int gwhip_h5_kq_maps(const double* mesh_scaled, int64_t nk, int32_t* kplus, int32_t* kminus) {
  if (!mesh_scaled || !kplus || !kminus || nk <= 0) return GWHIP_ERR_ARG;
  return guarded([&] { kq_maps(mesh_scaled, nk, kplus, kminus); });
}

int gwhip_h5_load(gwhip_handle* h, const char* df_path, const char* input_h5, int with_symmetry) {
  if (!h || !df_path || !input_h5) return GWHIP_ERR_ARG;
  return guarded([&] {
    gwhip_dims d{};
    int st = gwhip_get_dims(h, &d);
    if (st != GWHIP_OK) throw h5_error(st, "gwhip_get_dims failed");
    auto call = [&](int s) {
      if (s != GWHIP_OK) throw h5_error(s, gwhip_last_error(h));
    };
    std::vector<int64_t> b(d.nk), i(d.nk), t(d.nk);
    std::vector<std::complex<double>> U((size_t)(d.nk * d.nso * d.nso));
    std::vector<double> mesh((size_t)d.nk * 3);
    int64_t ink = 0;
    Id f(open_file(input_h5), H5Fclose);
    const bool has_k = has_path(f.get(), "symmetry/k/mesh_scaled");
    if (has_k) {
      st = gwhip_h5_read_k_symmetry(input_h5, d.nk, d.nso, &ink, b.data(), i.data(), reinterpret_cast<double*>(U.data()),
                                    t.data(), mesh.data());
      if (st != GWHIP_OK) throw h5_error(st, g_err);
      std::vector<int32_t> kp((size_t)(d.nk * d.nk)), km((size_t)(d.nk * d.nk));
      kq_maps(mesh.data(), d.nk, kp.data(), km.data());
      call(gwhip_set_kq_maps(h, kp.data(), km.data()));
    } else if (with_symmetry) {
      throw h5_error(GWHIP_ERR_IO, "no /symmetry/k tables in " + std::string(input_h5));
    }
    std::vector<std::complex<double>> V((size_t)(d.nk * (d.nk + 1) / 2 * d.NQ * d.nao * d.nao));
    read_packed(df_path, input_h5, d.nk, d.nao, d.NQ, V.data());
    call(gwhip_upload_V(h, reinterpret_cast<const double*>(V.data())));
    if (with_symmetry) {
      call(gwhip_set_k_symmetry(h, ink, b.data(), i.data(), reinterpret_cast<const double*>(U.data()), t.data()));
      if (has_path(f.get(), "symmetry/q/k_sym_transform_p0")) {  // the star-of-q evaluation
        int64_t inq = 0;
        std::vector<int64_t> qb(d.nk), qi(d.nk), qt(d.nk);
        std::vector<std::complex<double>> Uq((size_t)(d.nk * d.NQ * d.NQ));
        st = gwhip_h5_read_q_symmetry(input_h5, d.nk, d.NQ, &inq, qb.data(), qi.data(),
                                      reinterpret_cast<double*>(Uq.data()), qt.data());
        if (st != GWHIP_OK) throw h5_error(st, g_err);
        call(gwhip_set_q_symmetry(h, inq, qb.data(), qi.data(), reinterpret_cast<const double*>(Uq.data()), qt.data()));
      } else {
        call(gwhip_set_q_symmetry(h, 0, nullptr, nullptr, nullptr, nullptr));
      }
    }
  });
}

int gwhip_h5_load_q0(gwhip_handle* h, const char* df_path, const char* input_h5, int64_t nk_S, const double* S,
                     int* enabled) {
  if (!h || !df_path || !input_h5 || !S || !enabled) return GWHIP_ERR_ARG;
  return guarded([&] {
    *enabled = 0;
    gwhip_dims d{};
    int st = gwhip_get_dims(h, &d);
    if (st != GWHIP_OK) throw h5_error(st, "gwhip_get_dims failed");
    const std::string aq = std::string(df_path) + "/AqQ.h5";
    quiet();
    if (H5Fis_hdf5(aq.c_str()) <= 0) return;  // no extrapolation data: the correction stays off
    Id f(open_file(aq), H5Fclose);
    const auto dims = dims_of(f.get(), "AqQ");
    if (dims.size() != 2 || (int64_t)dims[1] != d.NQ)
      throw h5_error(GWHIP_ERR_IO, "AqQ is not [inq][NQ] for NQ = " + std::to_string(d.NQ));
    const int64_t inq = (int64_t)dims[0];
    std::vector<std::complex<double>> A((size_t)(inq * d.NQ));
    read_complex(f.get(), "AqQ", A.data(), A.size());
    std::vector<double> qabs(inq);
    if (count_of(dims_of(f.get(), "q_abs")) != (size_t)inq) throw h5_error(GWHIP_ERR_IO, "q_abs is not [inq]");
    read_into(f.get(), "q_abs", H5T_NATIVE_DOUBLE, qabs.data(), qabs.size());
    double mad = 0.0;
    read_into(f.get(), "madelung", H5T_NATIVE_DOUBLE, &mad, 1);
    std::vector<int64_t> qir(inq);
    Id in(open_file(input_h5), H5Fclose);
    if (has_path(in.get(), "symmetry/q/ibz2bz")) {
      qir = read_i64(in.get(), "symmetry/q/ibz2bz");
      if ((int64_t)qir.size() != inq) throw h5_error(GWHIP_ERR_IO, "symmetry/q/ibz2bz does not have inq entries");
    } else {
      for (int64_t i = 0; i < inq; ++i) qir[i] = i;
    }
    st = gwhip_set_q0_extrapolation(h, inq, qir.data(), reinterpret_cast<const double*>(A.data()), qabs.data(), mad,
                                    nk_S, S);
    if (st != GWHIP_OK) throw h5_error(st, gwhip_last_error(h));
    *enabled = 1;
  });
}

const char* gwhip_h5_last_error(void) { return g_err.c_str(); }

}  // extern "C"
