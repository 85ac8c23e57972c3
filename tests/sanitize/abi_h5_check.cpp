// Host-side checks of the C ABI's argument handling and of the HDF5 integral reader, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (host code only; no GPU is touched: the library is
// compiled without device code and every call here fails, or succeeds, before a kernel launch).
// Usage: abi_h5_check <tests/golden/h5 directory>.  Exit status 0 when every expectation holds; a
// sanitizer finding aborts the run.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gwhip.h"
#include "gwhip_h5.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

gwhip_dims dims(int64_t nk, int64_t nao, int64_t nso, int64_t ns, int64_t NQ, int64_t nts) {
  gwhip_dims d;
  std::memset(&d, 0, sizeof(d));
  d.nk = nk;
  d.mesh[0] = nk;
  d.mesh[1] = d.mesh[2] = 1;
  d.nao = nao;
  d.nso = nso;
  d.ns = ns;
  d.NQ = NQ;
  d.nts = nts;
  d.nw_b = 3;
  d.beta = 10.0;
  return d;
}

int create(const gwhip_dims& d) {
  gwhip_handle* h = nullptr;
  const int st = gwhip_create(&d, 0, &h);
  if (st == GWHIP_OK) gwhip_destroy(h);
  else expect(h == nullptr, "failed create leaves the handle null");
  return st;
}

void abi_checks() {
  expect(gwhip_create(nullptr, 0, nullptr) == GWHIP_ERR_ARG, "create(null)");
  expect(create(dims(1, 2, 2, 1, 3, 5)) == GWHIP_ERR_WRONG_GRID, "odd nts -> wrong grid");
  expect(std::strstr(gwhip_last_error(nullptr), "even") != nullptr, "odd nts message");
  expect(create(dims(0, 2, 2, 1, 3, 4)) == GWHIP_ERR_ARG, "nk = 0");
  expect(create(dims(1, 2, 3, 1, 3, 4)) == GWHIP_ERR_ARG, "nso neither nao nor 2 nao");
  expect(create(dims(1, 2, 2, 3, 3, 4)) == GWHIP_ERR_ARG, "ns = 3");
  expect(create(dims(1, 2, 4, 2, 3, 4)) == GWHIP_ERR_ARG, "X2C with ns = 2");
  expect(create(dims(1, 2, 2, 1, -1, 4)) == GWHIP_ERR_ARG, "negative NQ");
  gwhip_dims m = dims(2, 2, 2, 1, 3, 4);
  m.mesh[0] = 1;
  expect(create(m) == GWHIP_ERR_ARG, "mesh product != nk");
  const int st = create(dims(1, 2, 2, 1, 3, 4));
  expect(st == GWHIP_OK || st == GWHIP_ERR_DEVICE, "valid dims reach the device check");

  expect(gwhip_set_grid(nullptr, nullptr, nullptr) == GWHIP_ERR_ARG, "set_grid(null)");
  expect(gwhip_upload_V(nullptr, nullptr) == GWHIP_ERR_ARG, "upload_V(null)");
  expect(gwhip_set_kq_maps(nullptr, nullptr, nullptr) == GWHIP_ERR_ARG, "set_kq_maps(null)");
  expect(gwhip_eval_sigma(nullptr, nullptr, nullptr, 0) == GWHIP_ERR_ARG, "eval_sigma(null)");
  expect(gwhip_eval_sigma_device(nullptr, nullptr, nullptr, 0) == GWHIP_ERR_ARG, "eval_sigma_device(null)");
  expect(gwhip_eval_sigma_device_noalloc(nullptr, nullptr, nullptr, 0) == GWHIP_ERR_ARG, "noalloc(null)");
  expect(gwhip_set_split_slices(nullptr, 5) == GWHIP_ERR_ARG, "set_split_slices(null)");
  expect(gwhip_vjp(nullptr, nullptr, nullptr, nullptr, 0) == GWHIP_ERR_ARG, "vjp(null)");
  expect(gwhip_hf(nullptr, nullptr, nullptr, 0.0, nullptr) == GWHIP_ERR_ARG, "hf(null)");
  expect(gwhip_device_bytes(nullptr) == 0, "device_bytes(null)");
  int tune[GWHIP_TUNING_COUNT];
  expect(gwhip_get_tuning(nullptr, tune, GWHIP_TUNING_COUNT) == GWHIP_ERR_ARG, "get_tuning(null)");
  gwhip_batch* b = nullptr;
  expect(gwhip_batch_create(0, nullptr, 0, &b) == GWHIP_ERR_ARG && b == nullptr, "batch_create(0)");
  gwhip_dyson* dy = nullptr;
  expect(gwhip_dyson_create(nullptr, 0, &dy) == GWHIP_ERR_ARG, "dyson_create(null)");
  expect(gwhip_debug_zgemm(0, 1, 1, nullptr, nullptr, nullptr, 0, 0, 0) == GWHIP_ERR_ARG, "debug_zgemm(0)");
}

void h5_checks(const std::string& dir) {
  expect(gwhip_h5_check_version("1.0.0") == 1, "version 1.0.0");
  expect(gwhip_h5_check_version("1.1.2b") == 1, "version 1.1.2b");
  expect(gwhip_h5_check_version("0.3.1b10") == 0, "version 0.3.1b10");
  expect(gwhip_h5_check_version("1.0") == -1, "version 1.0");
  expect(gwhip_h5_check_version("") == -1, "empty version");
  expect(gwhip_h5_check_version(nullptr) == -1, "null version");
  std::string longv(4096, '9');
  expect(gwhip_h5_check_version(longv.c_str()) == -1 || gwhip_h5_check_version(longv.c_str()) == 1,
         "overlong version is handled");

  int64_t chunk = 0;
  expect(gwhip_h5_read_meta((dir + "/versions/df_int").c_str(), &chunk) == GWHIP_ERR_OUTDATED_INPUT, "meta pre-1.0");
  expect(gwhip_h5_read_meta((dir + "/versions/df_int_1.1.0").c_str(), &chunk) == GWHIP_OK && chunk == 196,
         "meta 1.1.0");
  expect(gwhip_h5_read_meta((dir + "/no_such_dir").c_str(), &chunk) != GWHIP_OK, "meta missing");
  expect(gwhip_h5_read_meta(nullptr, &chunk) == GWHIP_ERR_ARG, "meta(null)");

  const std::string df = dir + "/gw_sym27_df", in = dir + "/sym27_input.h5";
  const int64_t nk = 27, nao = 2, NQ = 36;
  std::vector<double> V((size_t)nk * (nk + 1) / 2 * NQ * nao * nao * 2);
  expect(gwhip_h5_read_packed(df.c_str(), in.c_str(), nk, nao, NQ, V.data()) == GWHIP_OK, "read_packed");
  double norm = 0.0;
  for (double v : V) norm += v * v;
  expect(norm > 0.0, "read_packed fills V");
  std::vector<double> W((size_t)26 * 27 / 2 * NQ * 3 * 3 * 2);
  expect(gwhip_h5_read_packed(df.c_str(), in.c_str(), nk, 3, NQ, W.data()) == GWHIP_ERR_IO, "wrong nao");
  expect(gwhip_h5_read_packed(df.c_str(), in.c_str(), 26, nao, NQ, W.data()) == GWHIP_ERR_IO, "wrong mesh");
  expect(gwhip_h5_read_packed((dir + "/no_such_dir").c_str(), in.c_str(), nk, nao, NQ, V.data()) == GWHIP_ERR_IO,
         "missing integrals");

  int64_t ink = 0;
  std::vector<int64_t> bz2ibz(nk), ibz2bz(nk), tr(nk);
  std::vector<double> U((size_t)nk * nao * nao * 2), ms((size_t)nk * 3);
  expect(gwhip_h5_read_k_symmetry(in.c_str(), nk, nao, &ink, bz2ibz.data(), ibz2bz.data(), U.data(), tr.data(),
                                  ms.data()) == GWHIP_OK && ink == 4,
         "read_k_symmetry");
  std::vector<int32_t> kp((size_t)nk * nk), km((size_t)nk * nk);
  expect(gwhip_h5_kq_maps(ms.data(), nk, kp.data(), km.data()) == GWHIP_OK, "kq_maps");
  bool in_range = true;
  for (size_t i = 0; i < kp.size(); ++i) in_range = in_range && kp[i] >= 0 && kp[i] < nk && km[i] >= 0 && km[i] < nk;
  expect(in_range, "kq_maps in range");
  expect(gwhip_h5_read_k_symmetry((dir + "/nothing.h5").c_str(), nk, nao, &ink, bz2ibz.data(), ibz2bz.data(),
                                  U.data(), tr.data(), ms.data()) == GWHIP_ERR_IO,
         "missing input file");
  int64_t inq = 0;
  std::vector<double> Uq((size_t)nk * NQ * NQ * 2);
  const int stq = gwhip_h5_read_q_symmetry(in.c_str(), nk, NQ, &inq, bz2ibz.data(), ibz2bz.data(), Uq.data(),
                                           tr.data());
  expect(stq == GWHIP_OK && inq == 4, "read_q_symmetry");
  expect(gwhip_h5_load(nullptr, df.c_str(), in.c_str(), 1) == GWHIP_ERR_ARG, "h5_load(null handle)");
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <tests/golden/h5>\n", argv[0]);
    return 2;
  }
  abi_checks();
  h5_checks(argv[1]);
  if (failures) {
    std::fprintf(stderr, "%d expectation(s) failed\n", failures);
    return 1;
  }
  std::printf("abi_h5_check: all expectations hold\n");
  return 0;
}
