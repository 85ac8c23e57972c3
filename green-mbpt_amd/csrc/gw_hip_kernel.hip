// Host orchestration of the GW self-energy on one MI355X, and the C ABI (include/gwhip.h).
#include "gw_hip_kernel.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "../../include/gwhip.h"
#include "gw_kernels.hpp"
#include "gw_mfma_kernels.hpp"
#include "gw_vjp_kernels.hpp"
#include "zgemm_mfma.hpp"
#include "zgemm_ozaki.hpp"

namespace gwhip {

namespace {

int k_shift(int k, int q, const std::array<int64_t, 3>& mesh, int sign) {
  const int n1 = (int)mesh[1], n2 = (int)mesh[2];
  int a[3] = {k / (n1 * n2), (k / n2) % n1, k % n2};
  int b[3] = {q / (n1 * n2), (q / n2) % n1, q % n2};
  int c[3];
  for (int i = 0; i < 3; ++i) {
    const int n = (int)mesh[i];
    c[i] = ((a[i] + sign * b[i]) % n + n) % n;
  }
  return (c[0] * n1 + c[1]) * n2 + c[2];
}

int query_max_lds(int device) {
  int v = 65536;
  int a = 0;
  if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && a > v) v = a;
#if HIP_VERSION_MAJOR >= 6
  int b = 0;
  if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, device) == hipSuccess && b > v) v = b;
#endif
  (void)hipGetLastError();
  return v;
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) GWHIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

}  // namespace

template <typename T>
T* gw_hip_kernel::dalloc(size_t count) {
  void* p = nullptr;
  const size_t bytes = count * sizeof(T);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw status_error(GWHIP_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  bytes_ += bytes;
  return static_cast<T*>(p);
}

gw_hip_kernel::gw_hip_kernel(const Dims& d, int device, hipStream_t shared_stream) : d_(d), device_(device) {
  if (d_.nk <= 0 || d_.nao <= 0 || d_.nso <= 0 || d_.ns <= 0 || d_.NQ <= 0 || d_.nts <= 0 || d_.nw_b <= 0)
    throw status_error(GWHIP_ERR_ARG, "all dimensions must be positive");
  if (d_.mesh[0] * d_.mesh[1] * d_.mesh[2] != d_.nk) throw status_error(GWHIP_ERR_ARG, "mesh does not multiply to nk");
  if (d_.nso != d_.nao && d_.nso != 2 * d_.nao) throw status_error(GWHIP_ERR_ARG, "nso must be nao or 2*nao");
  if (d_.ns != 1 && d_.ns != 2) throw status_error(GWHIP_ERR_ARG, "ns must be 1 or 2");
  d_.X2C = d_.nso != d_.nao;
  if (d_.X2C && d_.ns != 1) throw status_error(GWHIP_ERR_ARG, "X2C requires ns == 1");
  if (d_.nts % 2 != 0) throw status_error(GWHIP_ERR_WRONG_GRID, "Number of tau points should be even!!!");
  d_.pns = d_.nso / d_.nao;
  d_.pseudo_ns = d_.X2C ? 4 : d_.ns;
  d_.prefactor = (d_.ns == 2 || d_.X2C) ? 1.0 : 2.0;

  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    (void)hipGetLastError();
    throw status_error(GWHIP_ERR_DEVICE, "no HIP device " + std::to_string(device));
  }
  GWHIP_CHECK(hipSetDevice(device_));
  if (shared_stream) {
    stream_ = shared_stream;
    own_stream_ = false;
    own_ws_ = false;
  } else {
    GWHIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  }
  max_lds_ = query_max_lds(device_);

  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao, NQ2 = NQ * NQ;
  Vq_ = dalloc<double2>((size_t)nk * nk * NQ * nao2);
  P_ = dalloc<double2>((size_t)nk * d_.nts * NQ2);
  Tbt_fold_ = dalloc<double2>((size_t)d_.nw_b * (d_.nts / 2));
  Ttb_ = dalloc<double2>((size_t)d_.nts * d_.nw_b);
  kplus_tab_ = dalloc<int>((size_t)nk * nk);
  kminus_tab_ = dalloc<int>((size_t)nk * nk);
  flag_ = dalloc<int>(1);
  dyson_flag_ = dalloc<int>(1);
  GWHIP_CHECK(hipMemset(dyson_flag_, 0, sizeof(int)));
  hcheck_ = dalloc<unsigned long long>(2);
  build_maps();
  if (!own_ws_) return;  // a batch binds a shared workspace later

  // chunk sizes for the transient buffers from the memory left on the device
  size_t free_b = 0, total_b = 0;
  GWHIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  const double budget = 0.7 * (double)free_b;
  const double per_q_dyson = 16.0 * (double)dyson_per_q();
  const double per_t_scr = (double)nk * NQ * nao2 * 16.0;
  const double dyson_share = std::min(0.3 * budget, per_q_dyson * nk);
  qchunk_ = std::max<int64_t>(1, std::min<int64_t>(nk, (int64_t)(dyson_share / per_q_dyson)));
  const double scr_budget = budget - qchunk_ * per_q_dyson;
  int64_t tmax = std::max<int64_t>(1, (int64_t)(scr_budget / per_t_scr));
  tchunk_sigma_ = std::min<int64_t>(d_.nts / 2, tmax);  // R is formed on the half tau grid
  tchunk_p0_ = std::min<int64_t>(d_.nts / 2, tmax);
  Pw_ = dalloc<double2>((size_t)qchunk_ * dyson_per_q());
  scratch_elems_ = std::max(tchunk_sigma_, tchunk_p0_) * nk * NQ * nao2;
  scratch_ = dalloc<double2>((size_t)scratch_elems_);
}

gw_hip_kernel::~gw_hip_kernel() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (void* p : {(void*)Vq_, (void*)P_, (void*)Tbt_fold_, (void*)Ttb_, (void*)G_dev_, (void*)S_dev_,
                  (void*)kplus_tab_, (void*)kminus_tab_, (void*)flag_, (void*)dyson_flag_, (void*)leak_P0_, (void*)hcheck_, (void*)hf_work_, (void*)hf_x_, (void*)ksym_idx_, (void*)ksym_U_, (void*)hf_sym_, (void*)q0_qiq_, (void*)q0_A_, (void*)q0_eps_dev_,
                  (void*)q0_S_, (void*)q0_tmp_, (void*)q0_epst_, (void*)qsym_U_, (void*)star_P_, (void*)star_V_,
                  (void*)star_Vs_, (void*)star_eVs_, (void*)star_km_, (void*)star_q0row_, (void*)orb_W_,
                  (void*)orb_G_, (void*)orb_part_, (void*)orb_O_, (void*)epi_part_})
    if (p) (void)hipFree(p);
  if (own_ws_)
    for (void* p : {(void*)Pw_, (void*)scratch_})
      if (p) (void)hipFree(p);
  free_slices();
  free_vjp_buffers();
  for (auto& pd : pending_) {
    (void)hipEventDestroy(pd.a);
    (void)hipEventDestroy(pd.b);
  }
  if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
}

void gw_hip_kernel::build_maps() {
  const int nk = (int)d_.nk;
  std::vector<int> kpq((size_t)nk * nk), kmq((size_t)nk * nk);
  for (int q = 0; q < nk; ++q)
    for (int k = 0; k < nk; ++k) {
      kpq[(size_t)q * nk + k] = k_shift(k, q, d_.mesh, +1);  // [q][k]: k + q
      kmq[(size_t)q * nk + k] = k_shift(k, q, d_.mesh, -1);  // [q][k]: k - q
    }
  upload_maps(kpq, kmq);
}

void gw_hip_kernel::upload_maps(const std::vector<int>& kpq, const std::vector<int>& kmq) {
  const int nk = (int)d_.nk;
  for (int v : kpq)
    if (v < 0 || v >= nk) throw status_error(GWHIP_ERR_ARG, "k + q map entry out of range");
  for (int v : kmq)
    if (v < 0 || v >= nk) throw status_error(GWHIP_ERR_ARG, "k - q map entry out of range");
  // V(k, k - q) lives in the store of the q' with k + q' = k - q for every k
  neg_q_.assign(nk, -1);
  for (int q = 0; q < nk; ++q)
    for (int qp = 0; qp < nk && neg_q_[q] < 0; ++qp)
      if (std::equal(kpq.begin() + (size_t)qp * nk, kpq.begin() + (size_t)(qp + 1) * nk, kmq.begin() + (size_t)q * nk))
        neg_q_[q] = qp;
  for (int q = 0; q < nk; ++q)
    if (neg_q_[q] < 0) throw status_error(GWHIP_ERR_ARG, "k - q map has no matching k + q' row");
  kplus_host_.assign((size_t)nk * nk, 0);  // [k][q] for the V expansion
  for (int q = 0; q < nk; ++q)
    for (int k = 0; k < nk; ++k) kplus_host_[(size_t)k * nk + q] = kpq[(size_t)q * nk + k];
  GWHIP_CHECK(hipMemcpy(kplus_tab_, kpq.data(), kpq.size() * sizeof(int), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(kminus_tab_, kmq.data(), kmq.size() * sizeof(int), hipMemcpyHostToDevice));
  kminus_host_ = kmq;
  star_dirty_ = true;
}

void gw_hip_kernel::set_kq_maps(const int32_t* kplus, const int32_t* kminus) {
  if (!kplus || !kminus) throw status_error(GWHIP_ERR_ARG, "null k-q map");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t n = (size_t)d_.nk * d_.nk;
  upload_maps(std::vector<int>(kplus, kplus + n), std::vector<int>(kminus, kminus + n));
  v_set_ = false;  // the V store is indexed by the maps
  stored_valid_ = false;
}

void gw_hip_kernel::set_grid(const double2* T_bt, const double2* T_tb) {
  if (!T_bt || !T_tb) throw status_error(GWHIP_ERR_ARG, "null transform matrix");
  const int64_t nts = d_.nts, half = nts / 2, nwb = d_.nw_b;
  // fold P(beta - tau) = P(tau): T'[n][t] = T[n][t] + T[n][nts-1-t], t < nts/2
  std::vector<double2> fold((size_t)nwb * half);
  for (int64_t n = 0; n < nwb; ++n)
    for (int64_t t = 0; t < half; ++t) {
      const double2 a = T_bt[n * nts + t], b = T_bt[n * nts + nts - 1 - t];
      fold[n * half + t] = make_double2(a.x + b.x, a.y + b.y);
    }
  GWHIP_CHECK(hipSetDevice(device_));
  GWHIP_CHECK(hipMemcpy(Tbt_fold_, fold.data(), fold.size() * sizeof(double2), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(Ttb_, T_tb, (size_t)nts * nwb * sizeof(double2), hipMemcpyHostToDevice));
  Ttb_host_.assign(T_tb, T_tb + (size_t)nts * nwb);
  grid_set_ = true;
  stored_valid_ = false;
}

void gw_hip_kernel::upload_V(const double2* V_packed) {
  if (!V_packed) throw status_error(GWHIP_ERR_ARG, "null V");
  GWHIP_CHECK(hipSetDevice(device_));
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  const size_t npk = (size_t)nk * (nk + 1) / 2 * NQ * nao2;
  double2* tmp = nullptr;
  int* tab = nullptr;
  if (hipMalloc(&tmp, npk * sizeof(double2)) != hipSuccess) {
    (void)hipGetLastError();
    throw status_error(GWHIP_ERR_OOM, "no room for the packed V staging buffer");
  }
  hipError_t e = hipMalloc(&tab, kplus_host_.size() * sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(tmp, V_packed, npk * sizeof(double2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(tab, kplus_host_.data(), kplus_host_.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const int64_t total = nk * nk * NQ * nao2;
    hipLaunchKernelGGL(expand_v_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream_, tmp, Vq_, tab,
                       (int)nk, (int)NQ, (int)d_.nao, total);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  }
  (void)hipFree(tmp);
  if (tab) (void)hipFree(tab);
  GWHIP_CHECK(e);
  v_changed();
  v_set_ = true;
  stored_valid_ = false;
}

void gw_hip_kernel::free_slices() {
  for (void* p : {(void*)Vs_p0_, (void*)eVs_p0_, (void*)Vs_s_, (void*)eVs_s_})
    if (p) (void)hipFree(p);
  if (own_ws_)
    for (void* p : {(void*)Zs_, (void*)eZs_})
      if (p) (void)hipFree(p);
  Vs_p0_ = Vs_s_ = nullptr;
  eVs_p0_ = eVs_s_ = nullptr;
  if (own_ws_) {
    Zs_ = Ps_ = nullptr;
    eZs_ = ePs_ = nullptr;
  }
  slices_ns_ = 0;
  ++generation_;
}

void gw_hip_kernel::set_split_slices(int ns) {
  if (ns < 3 || ns > 5) throw status_error(GWHIP_ERR_ARG, "split slices must be 3, 4 or 5");
  if (ns != split_ns_) {
    GWHIP_CHECK(hipSetDevice(device_));
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    free_slices();
  }
  split_ns_ = ns;
}

size_t gw_hip_kernel::scratch_per_t() const { return (size_t)d_.nk * d_.NQ * d_.nao * d_.nao; }
// three regions of nw_b matrices per q: P0(i Omega), (I - P0)^-1, P-tilde(i Omega)
size_t gw_hip_kernel::dyson_per_q() const { return (size_t)3 * d_.nw_b * d_.NQ * d_.NQ; }
size_t gw_hip_kernel::slice_bytes_per_t() const {
  const size_t kcp = (size_t)ceil_div(d_.nao * d_.nao, 64) * 64, kcs = (size_t)ceil_div(d_.NQ, 64) * 64;
  return std::max((size_t)d_.NQ * d_.nk * OZ_PLANES * split_ns_ * kcp, (size_t)d_.NQ * OZ_PLANES * split_ns_ * kcs);
}
size_t gw_hip_kernel::exps_per_t() const { return (size_t)d_.NQ * d_.nk; }

void gw_hip_kernel::bind_workspace(const Workspace& ws) {
  if (own_ws_) throw status_error(GWHIP_ERR_STATE, "this kernel owns its workspace");
  int64_t tmax = (int64_t)(ws.scratch_elems / scratch_per_t());
  const int64_t qmax = (int64_t)(ws.dyson_elems / dyson_per_q());
  if (ws.slices) {
    tmax = std::min<int64_t>(tmax, (int64_t)(ws.slice_bytes / slice_bytes_per_t()));
    tmax = std::min<int64_t>(tmax, (int64_t)(ws.exp_elems / exps_per_t()));
  }
  if (tmax < 1 || qmax < 1) throw status_error(GWHIP_ERR_OOM, "shared workspace too small for this system");
  tchunk_sigma_ = std::min<int64_t>(d_.nts / 2, tmax);  // R is formed on the half tau grid
  tchunk_p0_ = std::min<int64_t>(d_.nts / 2, tmax);
  qchunk_ = std::min<int64_t>(d_.nk, qmax);
  scratch_ = ws.scratch;
  scratch_elems_ = (int64_t)ws.scratch_elems;
  Pw_ = ws.dyson;
  Zs_ = Ps_ = ws.slices;
  eZs_ = ePs_ = ws.exps;
  ++generation_;
}

bool gw_hip_kernel::detect_hermitian_G(const double2* G) {
  GWHIP_CHECK(hipSetDevice(device_));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, reinterpret_cast<int*>(hcheck_), (int64_t)4);
  const int64_t nmat = d_.nts * d_.ns * d_.nk;
  hipLaunchKernelGGL(herm_check_kernel, dim3(1024), dim3(256), 0, stream_, G, nmat, (int)d_.nso, hcheck_);
  GWHIP_CHECK(hipGetLastError());
  unsigned long long v[2] = {0, 0};
  GWHIP_CHECK(hipMemcpyAsync(v, hcheck_, sizeof(v), hipMemcpyDeviceToHost, stream_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  double dmax, amax;
  std::memcpy(&dmax, &v[0], sizeof(double));
  std::memcpy(&amax, &v[1], sizeof(double));
  g_herm_ = std::isfinite(dmax) && dmax <= 1e-12 * amax;
  return g_herm_;
}

void gw_hip_kernel::check_ready(int precision) const {
  if (!grid_set_) throw status_error(GWHIP_ERR_STATE, "transform grid not set (gwhip_set_grid)");
  if (!v_set_) throw status_error(GWHIP_ERR_STATE, "integrals not set (gwhip_upload_V / gwhip_generate_V)");
  if (!precision_valid(precision)) throw status_error(GWHIP_ERR_ARG, "unknown precision");
  if (!scratch_) throw status_error(GWHIP_ERR_STATE, "no workspace bound");
  if (any_split(precision) && !own_ws_ && (!split_ready() || !Zs_))
    throw status_error(GWHIP_ERR_STATE, "split-precision stores or workspace not prepared");
}

void gw_hip_kernel::prepare_split() {
  GWHIP_CHECK(hipSetDevice(device_));
  ensure_slices();
}

// Slice stores of conj(V) (built once per V and slice count).
void gw_hip_kernel::ensure_slices() {
  if (slices_ns_ == split_ns_ && Vs_p0_) return;
  free_slices();
  const int ns = split_ns_;
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  kc_p0_ = ceil_div(nao2, 64) * 64;
  kc_s_ = ceil_div(NQ, 64) * 64;
  // int32 exactness inside a K chunk: one accumulator sums the products of up to `ns` slice pairs
  // and two planes, each |x y| <= 127^2 < 2^14, over kc elements
  const int64_t kc_max = split_kc_limit(ns);
  if (kc_p0_ > kc_max || kc_s_ > kc_max)
    throw status_error(GWHIP_ERR_UNSUPPORTED,
                       "split precision: K chunk of " + std::to_string(std::max(kc_p0_, kc_s_)) +
                           " exceeds the int32-exact limit " + std::to_string(kc_max) + " (nao^2 and naux must stay below it)");
  Vs_p0_ = dalloc<int8_t>((size_t)nk * NQ * nk * OZ_PLANES * ns * kc_p0_);
  eVs_p0_ = dalloc<int>((size_t)nk * NQ * nk);
  Vs_s_ = dalloc<int8_t>((size_t)nk * nk * nao2 * OZ_PLANES * ns * kc_s_);
  eVs_s_ = dalloc<int>((size_t)nk * nk * nao2);
  if (own_ws_) {  // Z slices (P0 stage) and P-tilde slices (Sigma stage) share one buffer
    Zs_ = Ps_ = dalloc<int8_t>(std::max((size_t)tchunk_p0_ * NQ * nk * OZ_PLANES * ns * kc_p0_,
                                        (size_t)tchunk_sigma_ * NQ * OZ_PLANES * ns * kc_s_));
    eZs_ = ePs_ = dalloc<int>(std::max((size_t)tchunk_p0_ * NQ * nk, (size_t)tchunk_sigma_ * NQ));
  }
  {  // P0 operand: rows (q, Q'), chunks k, elements (t, n) of V(k, k+q)
    SliceArgs a{};
    a.src = Vq_; a.nrows = nk * NQ; a.row_div = NQ; a.s_rhi = nk * NQ * nao2; a.s_rlo = nao2;
    a.nchunk = (int)nk; a.s_c = NQ * nao2; a.klen = (int)nao2; a.s_k = 1; a.kc = (int)kc_p0_;
    a.conj = 1; a.dst = Vs_p0_; a.exps = eVs_p0_;
    slice_rows(a, ns, stream_);
  }
  {  // Sigma operand: rows (q, k, (n, j)) = V^Q'_jn(k, k+q) (per-Q transposed in place), elements Q'
    SliceArgs a{};
    a.src = Vq_; a.nrows = nk * nk * nao2; a.row_div = nao2; a.s_rhi = NQ * nao2;
    a.row_div2 = d_.nao; a.s_rlo = 1; a.s_rlo2 = d_.nao;
    a.nchunk = 1; a.s_c = 0; a.klen = (int)NQ; a.s_k = nao2; a.kc = (int)kc_s_;
    a.conj = 1; a.dst = Vs_s_; a.exps = eVs_s_;
    slice_rows(a, ns, stream_);
  }
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  slices_ns_ = ns;
  ++generation_;
}

void gw_hip_kernel::v_changed() { free_slices(); }

void gw_hip_kernel::generate_V(uint64_t seed, double sigma) {
  GWHIP_CHECK(hipSetDevice(device_));
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  int* tab = nullptr;
  GWHIP_CHECK(hipMalloc(&tab, kplus_host_.size() * sizeof(int)));
  hipError_t e = hipMemcpy(tab, kplus_host_.data(), kplus_host_.size() * sizeof(int), hipMemcpyHostToDevice);
  // key = splitmix64(seed ^ splitmix64(stream)), stream id of V = 0x5611 (gwhip.inputs.STREAM_V)
  auto sm = [](uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const uint64_t key = sm(seed ^ sm(0x5611ull));
  if (e == hipSuccess) {
    const int64_t total = nk * nk * NQ * nao2;
    hipLaunchKernelGGL(generate_v_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, stream_, Vq_, tab,
                       (int)nk, (int)NQ, (int)d_.nao, total, key, sigma);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  }
  (void)hipFree(tab);
  GWHIP_CHECK(e);
  v_changed();
  v_set_ = true;
  stored_valid_ = false;
}

template <typename F>
void gw_hip_kernel::timed(int which, F&& launch) {
  if (!timing_) {  // capture-safe: no events
    launch();
    return;
  }
  Pending pd{which, nullptr, nullptr};
  GWHIP_CHECK(hipEventCreate(&pd.a));
  GWHIP_CHECK(hipEventCreate(&pd.b));
  GWHIP_CHECK(hipEventRecord(pd.a, stream_));
  launch();
  GWHIP_CHECK(hipEventRecord(pd.b, stream_));
  pending_.push_back(pd);
}

// Accumulate the per-launch timings recorded since the last call (the stream must have drained).
void gw_hip_kernel::resolve_timings() {
  for (auto& pd : pending_) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pd.a, pd.b) == hipSuccess) {
      kstat_[pd.which].launches += 1;
      kstat_[pd.which].ms += ms;
    }
    (void)hipEventDestroy(pd.a);
    (void)hipEventDestroy(pd.b);
  }
  (void)hipGetLastError();
  pending_.clear();
}

namespace {
template <int NB, typename T>
void launch_z_prep_t(dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G, const double2* Vt,
                     double2* Z, const int* kplus, int t0, int QPW) {
  allow_lds(z_prep_mfma_kernel<NB, T>, lds);
  const XcdGrid xg = make_xcd_grid(grid);
  hipLaunchKernelGGL((z_prep_mfma_kernel<NB, T>), xcd_launch_grid(xg), dim3(256), lds, st, xg, dd, G, Vt, Z, kplus, t0,
                     QPW);
  GWHIP_CHECK(hipGetLastError());
}
// The orbital-size products stay FP64 in both precision modes: FP32 matrix cores here put the
// Sigma error at ~2.4e-8 relative on the nk = 8, nao = 8, naux = 40 case, above the 1e-8 bar.
template <int NB>
void launch_z_prep(dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G, const double2* Vt,
                   double2* Z, const int* kplus, int t0, int QPW) {
  launch_z_prep_t<NB, double>(grid, lds, st, dd, G, Vt, Z, kplus, t0, QPW);
}
template <int NB, int NS, int SB>
void launch_z_prep_slice_t(dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G, const double2* Vt,
                           int8_t* dst, int* exps, int kc, const int* kplus, int t0, int QPW) {
  allow_lds(z_prep_slice_kernel<NB, NS, SB>, lds);
  const XcdGrid xg = make_xcd_grid(grid);
  hipLaunchKernelGGL((z_prep_slice_kernel<NB, NS, SB>), xcd_launch_grid(xg), dim3(256), lds, st, xg, dd, G, Vt, dst,
                     exps, kc, kplus, t0, QPW);
  GWHIP_CHECK(hipGetLastError());
}
template <int NB, int SB>
void launch_z_prep_slice_s(int ns, dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G,
                           const double2* Vt, int8_t* dst, int* exps, int kc, const int* kplus, int t0, int QPW) {
  if (ns == 5) launch_z_prep_slice_t<NB, 5, SB>(grid, lds, st, dd, G, Vt, dst, exps, kc, kplus, t0, QPW);
  else if (ns == 4) launch_z_prep_slice_t<NB, 4, SB>(grid, lds, st, dd, G, Vt, dst, exps, kc, kplus, t0, QPW);
  else launch_z_prep_slice_t<NB, 3, SB>(grid, lds, st, dd, G, Vt, dst, exps, kc, kplus, t0, QPW);
}
template <int NB>
void launch_z_prep_slice(int ns, dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G,
                         const double2* Vt, int8_t* dst, int* exps, int kc, const int* kplus, int t0, int QPW) {
  if (dd.pseudo_ns == 2) launch_z_prep_slice_s<NB, 2>(ns, grid, lds, st, dd, G, Vt, dst, exps, kc, kplus, t0, QPW);
  else launch_z_prep_slice_s<NB, 1>(ns, grid, lds, st, dd, G, Vt, dst, exps, kc, kplus, t0, QPW);
}
template <int NB, typename T>
void launch_sigma_epi_t(dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G, const double2* Vt,
                        const double2* R, double2* S, const int* kminus, int nks, int t0, double scale, int upper,
                        int mirror, double2* part) {
  allow_lds(sigma_epi_mfma_kernel<NB, T>, lds);
  const XcdGrid xg = make_xcd_grid(grid);
  hipLaunchKernelGGL((sigma_epi_mfma_kernel<NB, T>), xcd_launch_grid(xg), dim3(256), lds, st, xg, dd, G, Vt, R, S,
                     kminus, nks, t0, scale, upper, mirror, (int)grid.y, part);
  GWHIP_CHECK(hipGetLastError());
}
template <int NB>
void launch_sigma_epi(dim3 grid, size_t lds, hipStream_t st, const DevDims& dd, const double2* G,
                      const double2* Vt, const double2* R, double2* S, const int* kminus, int nks, int t0,
                      double scale, int upper, int mirror, double2* part) {
  launch_sigma_epi_t<NB, double>(grid, lds, st, dd, G, Vt, R, S, kminus, nks, t0, scale, upper, mirror, part);
}
}  // namespace

void gw_hip_kernel::zero(double2* X, int64_t n) {
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<int64_t>(4096, std::max<int64_t>(1, ceil_div(n, 256)))),
                     dim3(256), 0, stream_, X, n);
  GWHIP_CHECK(hipGetLastError());
}

// Z_Q = sum_s Gb_s V_Q Gq_s for q and tau in [t0, t0 + tc) into scratch_.
void gw_hip_kernel::p0_prep(int64_t q, const double2* G, int64_t t0, int64_t tc, int qpw) {
  if (orb_gemm()) {
    p0_prep_gemm(q, G, t0, tc);
    return;
  }
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  DevDims dd{(int)d_.nts, (int)d_.ns, (int)nk, (int)d_.nso, (int)d_.nao, (int)NQ, (int)d_.pns, (int)d_.pseudo_ns};
  const int NB = (int)ceil_div(d_.nao, 16);
  const size_t lds = (size_t)4 * (16 * NB) * (16 * NB + 1) * sizeof(double);
  const double2* Vtq = Vq_ + q * nk * NQ * nao2;
  const dim3 grid((unsigned)ceil_div(NQ, 4 * qpw), (unsigned)nk, (unsigned)tc);
  const int* kp = kplus_tab_ + q * nk;
  if (NB == 1) launch_z_prep<1>(grid, lds, stream_, dd, G, Vtq, scratch_, kp, (int)t0, qpw);
  else if (NB == 2) launch_z_prep<2>(grid, lds, stream_, dd, G, Vtq, scratch_, kp, (int)t0, qpw);
  else launch_z_prep<3>(grid, lds, stream_, dd, G, Vtq, scratch_, kp, (int)t0, qpw);
}

// Split precision, one or two collinear spin blocks: Z slices for q and tau in [t0, t0 + tc)
// straight into Zs_.
void gw_hip_kernel::p0_prep_slice(int64_t q, const double2* G, int64_t t0, int64_t tc, int qpw) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  DevDims dd{(int)d_.nts, (int)d_.ns, (int)nk, (int)d_.nso, (int)d_.nao, (int)NQ, (int)d_.pns, (int)d_.pseudo_ns};
  const int NB = (int)ceil_div(d_.nao, 16);
  const size_t lds = (size_t)d_.pseudo_ns * 4 * (16 * NB) * (16 * NB + 1) * sizeof(double);
  const double2* Vtq = Vq_ + q * nk * NQ * nao2;
  const dim3 grid((unsigned)ceil_div(NQ, 4 * qpw), (unsigned)nk, (unsigned)tc);
  const int* kp = kplus_tab_ + q * nk;
  const int ns = split_ns_;
  if (NB == 1) launch_z_prep_slice<1>(ns, grid, lds, stream_, dd, G, Vtq, Zs_, eZs_, (int)kc_p0_, kp, (int)t0, qpw);
  else if (NB == 2) launch_z_prep_slice<2>(ns, grid, lds, stream_, dd, G, Vtq, Zs_, eZs_, (int)kc_p0_, kp, (int)t0, qpw);
  else launch_z_prep_slice<3>(ns, grid, lds, stream_, dd, G, Vtq, Zs_, eZs_, (int)kc_p0_, kp, (int)t0, qpw);
}

// P0(t0 .. t0+tc) = -(pref/nk) Z conj(V)^T, FP64 (tile `shape`).
void gw_hip_kernel::p0_gemm(int64_t q, int64_t tc, double2* Pt0, int shape) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao, NQ2 = NQ * NQ;
  ZGemmArgs a{};
  a.M = (int)NQ; a.N = (int)NQ; a.K = (int)(nk * nao2); a.batch = (int)tc;
  a.kin = (int)nao2; a.nin = (int)NQ;
  a.A = scratch_; a.sAb = nk * NQ * nao2; a.sAm = nao2; a.sAko = NQ * nao2; a.sAki = 1;
  a.B = Vq_ + q * nk * NQ * nao2; a.sBb = 0; a.sBko = NQ * nao2; a.sBki = 1; a.sBno = 0; a.sBni = nao2;
  a.C = Pt0; a.sCb = NQ2; a.sCm = NQ; a.sCno = 0; a.sCni = 1;
  a.alpha = make_double2(-d_.prefactor / (double)nk, 0.0);
  a.accumulate = 0; a.conjA = 0; a.conjB = 1;
  a.upper = g_herm_ ? 1 : 0;
  zgemm_f64(a, true, true, stream_, shape);
}

// R = P-tilde(t0 .. t0+tc) conj(V) into scratch_, FP64 (tile `shape`): Pt0 = P-tilde at t0
// [tc][NQ^2], Vb = V(k, k - q) of nkb k points [nkb][NQ][nao][nao], read per-Q transposed
// (column (k, n, j) of the product is V^Q'_jn).
void gw_hip_kernel::sigma_gemm(const double2* Pt0, const double2* Vb, int64_t nkb, int64_t tc, int shape) {
  const int64_t NQ = d_.NQ, nao2 = d_.nao * d_.nao, NQ2 = NQ * NQ;
  ZGemmArgs a{};
  a.M = (int)NQ; a.N = (int)(nkb * nao2); a.K = (int)NQ; a.batch = (int)tc; a.kin = (int)NQ; a.nin = (int)nao2;
  a.A = Pt0; a.sAb = NQ2; a.sAm = NQ; a.sAko = 0; a.sAki = 1;
  a.B = Vb; a.sBb = 0; a.sBko = 0; a.sBki = nao2; a.sBno = NQ * nao2; a.sBni = 1;
  a.nin2 = (int)d_.nao; a.sBni2 = d_.nao;
  a.C = scratch_; a.sCb = nkb * NQ * nao2; a.sCm = nao2; a.sCno = NQ * nao2; a.sCni = 1;
  a.alpha = make_double2(1.0, 0.0);
  a.accumulate = 0; a.conjA = 0; a.conjB = 1;
  zgemm_f64(a, true, false, stream_, shape);
}

// Split precision, P0: int8 slices of Z (rows (tau, Q), chunks k) into Zs_.
void gw_hip_kernel::p0_slice(int64_t tc) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  SliceArgs sa{};
  sa.src = scratch_; sa.nrows = tc * NQ; sa.row_div = NQ; sa.s_rhi = nk * NQ * nao2; sa.s_rlo = nao2;
  sa.nchunk = (int)nk; sa.s_c = NQ * nao2; sa.klen = (int)nao2; sa.s_k = 1; sa.kc = (int)kc_p0_;
  sa.conj = 0; sa.dst = Zs_; sa.exps = eZs_;
  slice_rows(sa, split_ns_, stream_);
}

// Split precision, P0 = -(pref/nk) Z conj(V)^T from the slices (row-tile group `group`).
void gw_hip_kernel::p0_gemm_split(int64_t q, int64_t tc, double2* Pt0, int group, int kchunks) {
  const int64_t nk = d_.nk, NQ = d_.NQ, NQ2 = NQ * NQ;
  const int ns = split_ns_;
  const int64_t sc = OZ_PLANES * ns * kc_p0_;  // bytes of one k chunk of a row
  // K = nk chunks of kc: in launches over `kchunks` consecutive chunks each, the first writing C
  // and the others adding to it, so that one launch's A and B chunks are reused from L2 by all
  // its tiles instead of streaming the whole K past every tile (partial sums in FP64, as between
  // chunks inside a launch).  At nk = 64, naux = 300: 770 ms per evaluation in one launch per
  // (q, tau chunk), 604 ms with 4 chunks per launch.
  const int64_t nsl = kchunks <= 0 ? 1 : ceil_div(nk, std::min<int64_t>(kchunks, nk));
  for (int64_t sl = 0; sl < nsl; ++sl) {
    const int64_t c0 = nk * sl / nsl, c1 = nk * (sl + 1) / nsl;
    OzGemmArgs o{};
    o.M = (int)NQ; o.N = (int)NQ; o.batch = (int)tc; o.kc = (int)kc_p0_; o.nchunk = (int)(c1 - c0); o.nin = (int)NQ;
    o.A = Zs_ + c0 * sc; o.sAm = nk * sc; o.sAb = NQ * o.sAm; o.sAc = sc;
    o.eA = eZs_ + c0; o.seAb = NQ * nk; o.seAm = nk;
    o.B = Vs_p0_ + q * NQ * nk * sc + c0 * sc; o.sBb = 0; o.sBno = 0; o.sBni = nk * sc; o.sBc = sc;
    o.eB = eVs_p0_ + q * NQ * nk + c0; o.seBb = 0; o.seBno = 0; o.seBni = nk;
    o.C = Pt0; o.sCb = NQ2; o.sCm = NQ; o.sCno = 0; o.sCni = 1;
    o.alpha = make_double2(-d_.prefactor / (double)nk, 0.0);
    o.accumulate = sl > 0 ? 1 : 0;
    o.group_m = group;
    o.upper = g_herm_ ? 1 : 0;
    o.tile = tune_.p0_oz_tile;
    zgemm_ozaki(o, ns, stream_);
  }
}

// Split precision, Sigma: int8 slices of P-tilde(t0 .. t0+tc) (Pt0: [tc][NQ^2]) into Ps_.
void gw_hip_kernel::sigma_slice(const double2* Pt0, int64_t tc) {
  const int64_t NQ = d_.NQ;
  SliceArgs sa{};
  sa.src = Pt0; sa.nrows = tc * NQ; sa.row_div = tc * NQ; sa.s_rhi = 0; sa.s_rlo = NQ;
  sa.nchunk = 1; sa.s_c = 0; sa.klen = (int)NQ; sa.s_k = 1; sa.kc = (int)kc_s_;
  sa.conj = 0; sa.dst = Ps_; sa.exps = ePs_;
  slice_rows(sa, split_ns_, stream_);
}

// Split precision, R = P-tilde conj(V) into scratch_ from the slices: Vsb / eVsb = the slices and
// exponents of V(k, k - q) for nkb k points ([nkb][nao^2][2][ns][kc_s], [nkb][nao^2]).
void gw_hip_kernel::sigma_gemm_split(const int8_t* Vsb, const int* eVsb, int64_t nkb, int64_t tc, int group) {
  const int64_t NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  const int ns = split_ns_;
  OzGemmArgs o{};
  o.M = (int)NQ; o.N = (int)(nkb * nao2); o.batch = (int)tc; o.kc = (int)kc_s_; o.nchunk = 1; o.nin = (int)nao2;
  o.A = Ps_; o.sAm = OZ_PLANES * ns * kc_s_; o.sAb = NQ * o.sAm; o.sAc = 0;
  o.eA = ePs_; o.seAb = NQ; o.seAm = 1;
  o.B = Vsb; o.sBb = 0; o.sBno = nao2 * OZ_PLANES * ns * kc_s_;
  o.sBni = OZ_PLANES * ns * kc_s_; o.sBc = 0;
  o.eB = eVsb; o.seBb = 0; o.seBno = nao2; o.seBni = 1;
  o.C = scratch_; o.sCb = nkb * NQ * nao2; o.sCm = nao2; o.sCno = NQ * nao2; o.sCni = 1;
  o.alpha = make_double2(1.0, 0.0);
  o.accumulate = 0;
  o.group_m = group;
  o.tile = tune_.s_oz_tile;
  zgemm_ozaki(o, ns, stream_);
}

// P0 of every q in `qs` (all q of the mesh when null): the i-th into P_ row i.
void gw_hip_kernel::stage_p0(const double2* G, bool split, const std::vector<int>* qs) {
  const int64_t n = qs ? (int64_t)qs->size() : d_.nk;
  for (int64_t i = 0; i < n; ++i) p0_one_q(qs ? (*qs)[i] : i, G, split, P_ + i * d_.nts * d_.NQ * d_.NQ);
}

// P0(tau, q) for the half tau grid into Pq[t][Q][Q'] (t < nts/2), Hermitised.
void gw_hip_kernel::p0_one_q(int64_t q, const double2* G, bool split, double2* Pq) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao2 = d_.nao * d_.nao, NQ2 = NQ * NQ, half = d_.nts / 2;
  for (int64_t t0 = 0; t0 < half; t0 += tchunk_p0_) {
    const int64_t tc = std::min(tchunk_p0_, half - t0);
    if (split && fused_p0() && tune_.p0_fused) {
      timed(K_P0_PREP, [&] { p0_prep_slice(q, G, t0, tc, tune_.qpw); });
    } else {
      timed(K_P0_PREP, [&] { p0_prep(q, G, t0, tc, tune_.qpw); });
      if (split) p0_slice(tc);
    }
    if (split) {
      timed(K_P0_GEMM, [&] {
        p0_gemm_split(q, tc, Pq + t0 * NQ2, tune_.p0_group, tune_.p0_kchunks);
      });
    } else {
      timed(K_P0_GEMM, [&] { p0_gemm(q, tc, Pq + t0 * NQ2, tune_.p0_shape); });
    }
  }
  // Hermitise the half grid of this q (src/gw_cpu_kernel.cpp:217-225); for Hermitian G the
  // product is Hermitian and only its upper tiles were computed
  if (g_herm_)
    mirror_upper(Pq, (int)NQ, NQ2, half, stream_);
  else
    hermitize(Pq, (int)NQ, NQ2, half, stream_);
  GWHIP_CHECK(hipGetLastError());
}

bool gw_hip_kernel::dyson_failed() const {
  int f = 0;
  GWHIP_CHECK(hipMemcpy(&f, dyson_flag_, sizeof(int), hipMemcpyDeviceToHost));
  return f != 0;
}

void gw_hip_kernel::check_dyson_flag() {
  if (dyson_failed())
    throw status_error(GWHIP_ERR_NOT_POSITIVE_DEFINITE,
                       "I - P0(i Omega) is not positive definite for some (q, Omega): a pivot of its factorisation "
                       "is not positive (P0 is not the polarisation of a physical G)");
}

// Hermitian block Gauss-Jordan inverse (the sweep operator; hgj_panel_kernel states the invariant):
// only the upper block triangle of each matrix is kept current, and on return it holds the upper
// triangle of A^-1 (the diagonal tiles whole; the rest of the lower triangle is stale).  Per block
// column J (width nb = 64, the FP64 GEMM's tile, so pivot blocks are diagonal tiles):
//   P = the step's panel (pivot rows, and the swept rows' column J conjugated), D = A_JJ;
//   D <- D^-1 (gj_inverse_kernel, with the pivot guard);  Q = D P, columns < J0 negated (new pivot rows);
//   A -= P^H Q on the upper tiles;  A[i < J0][J] = -P^H D for the swept rows;  pivot rows <- Q, D.
// The update streams only the upper tiles: about 0.6 of the general sweep's GEMM work at NQ = 300.
void gw_hip_kernel::block_inverse(double2* X, int64_t nmat) {
  const int64_t n = d_.NQ, n2 = n * n;
  const int nb = (int)std::min<int64_t>(64, n);
  const int64_t per = 2 * n * nb + (int64_t)nb * nb;
  const int64_t mb_max = std::min<int64_t>(65535, scratch_elems_ / per);
  if (mb_max < 1) throw status_error(GWHIP_ERR_UNSUPPORTED, "scratch too small for the Dyson inverse panels");
  const size_t gj_lds = (size_t)(GJ_NB * GJ_NB + nb * GJ_NB) * sizeof(double2);
  for (int64_t m0 = 0; m0 < nmat; m0 += mb_max) {
    const int64_t mb = std::min(mb_max, nmat - m0);
    double2* A = X + m0 * n2;
    double2* P = scratch_;                 // [mb][jb][n]     (allocation stride nb n)
    double2* Q = P + mb * (int64_t)nb * n; // [mb][jb][n]
    double2* D = Q + mb * (int64_t)nb * n; // [mb][jb][jb]    (stride nb^2)
    const int64_t sP = (int64_t)nb * n, sD = (int64_t)nb * nb;
    for (int J0 = 0; J0 < n; J0 += nb) {
      const int jb = (int)std::min<int64_t>(nb, n - J0);
      hipLaunchKernelGGL(hgj_panel_kernel, dim3((unsigned)ceil_div(n * jb, 256), (unsigned)mb), dim3(256), 0, stream_,
                         A, (int)n, n2, J0, jb, P, sP, D, sD);
      GWHIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(gj_inverse_kernel, dim3((unsigned)mb), dim3(256), gj_lds, stream_, D, jb, sD, GJ_NB,
                         dyson_flag_);
      GWHIP_CHECK(hipGetLastError());
      {  // Q = D^-1 P
        ZGemmArgs a{};
        a.M = jb; a.N = (int)n; a.K = jb; a.batch = (int)mb; a.kin = jb; a.nin = (int)n;
        a.A = D; a.sAb = sD; a.sAm = jb; a.sAko = 0; a.sAki = 1;
        a.B = P; a.sBb = sP; a.sBko = 0; a.sBki = n; a.sBno = 0; a.sBni = 1;
        a.C = Q; a.sCb = sP; a.sCm = n; a.sCno = 0; a.sCni = 1;
        a.alpha = make_double2(1.0, 0.0);
        zgemm_f64(a, true, false, stream_);
      }
      if (J0 > 0) {
        hipLaunchKernelGGL(hgj_negate_kernel, dim3((unsigned)ceil_div((int64_t)jb * J0, 256), (unsigned)mb), dim3(256),
                           0, stream_, Q, (int)n, sP, J0, jb);
        GWHIP_CHECK(hipGetLastError());
      }
      {  // A -= P^H Q, upper tiles (the old column panel is P^H)
        ZGemmArgs a{};
        a.M = (int)n; a.N = (int)n; a.K = jb; a.batch = (int)mb; a.kin = jb; a.nin = (int)n;
        a.A = P; a.sAb = sP; a.sAm = 1; a.sAko = 0; a.sAki = n; a.conjA = 1;
        a.B = Q; a.sBb = sP; a.sBko = 0; a.sBki = n; a.sBno = 0; a.sBni = 1;
        a.C = A; a.sCb = n2; a.sCm = n; a.sCno = 0; a.sCni = 1;
        a.alpha = make_double2(-1.0, 0.0);
        a.accumulate = 1;
        a.upper = 1;
        zgemm_f64(a, false, false, stream_);
      }
      if (J0 > 0) {  // A[i][J] = -P^H D^-1 for the swept rows i < J0 (the rows below are not kept)
        ZGemmArgs a{};
        a.M = J0; a.N = jb; a.K = jb; a.batch = (int)mb; a.kin = jb; a.nin = jb;
        a.A = P; a.sAb = sP; a.sAm = 1; a.sAko = 0; a.sAki = n; a.conjA = 1;
        a.B = D; a.sBb = sD; a.sBko = 0; a.sBki = jb; a.sBno = 0; a.sBni = 1;
        a.C = A + J0; a.sCb = n2; a.sCm = n; a.sCno = 0; a.sCni = 1;
        a.alpha = make_double2(-1.0, 0.0);
        zgemm_f64(a, false, false, stream_);
      }
      hipLaunchKernelGGL(hgj_rows_kernel, dim3((unsigned)ceil_div((int64_t)jb * (n - J0), 256), (unsigned)mb), dim3(256),
                         0, stream_, A, (int)n, n2, J0, jb, Q, sP, D, sD);
      GWHIP_CHECK(hipGetLastError());
    }
  }
}

// P-tilde of the first nq rows of P_; q0_row[i] is the AqQ row of row i (-1: none), read when the
// q -> 0 data are being collected.  Per q chunk the Dyson work buffer holds three regions of
// qchunk_ x nw_b matrices: W0 = P0(i Omega) = T_bt' P0(tau); W1 = A = I - P0(i Omega), inverted in
// place and Hermitised; W2 = P-tilde(i Omega) = Herm(A^-1) P0(i Omega), the reference's product
// (src/gw_cpu_kernel.cpp:241-252; forming Herm(A^-1) - I instead would lose the relative accuracy of
// P-tilde where |P0| << 1); then P-tilde(tau) = T_tb P-tilde(i Omega).
void gw_hip_kernel::stage_dyson(int64_t nq, const int* q0_row) {
  const int64_t NQ = d_.NQ, NQ2 = NQ * NQ, half = d_.nts / 2, nwb = d_.nw_b, nts = d_.nts;
  const int64_t region = qchunk_ * nwb * NQ2;
  double2 *W0 = Pw_, *W1 = Pw_ + region, *W2 = Pw_ + 2 * region;

  Event e0(timing_), e1(timing_), e2(timing_), e3(timing_);
  auto mark = [&](hipEvent_t e) { if (timing_) GWHIP_CHECK(hipEventRecord(e, stream_)); };
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, dyson_flag_, (int64_t)1);
  float tw = 0.f, bse = 0.f, wt = 0.f;
  for (int64_t q0 = 0; q0 < nq; q0 += qchunk_) {
    const int64_t qc = std::min(qchunk_, nq - q0);
    mark(e0);
    {  // P0(tau_f) -> P0(i Omega_b) with the folded matrix
      ZGemmArgs a{};
      a.M = (int)nwb; a.N = (int)NQ2; a.K = (int)half; a.batch = (int)qc; a.kin = (int)half; a.nin = (int)NQ2;
      a.A = Tbt_fold_; a.sAb = 0; a.sAm = half; a.sAko = 0; a.sAki = 1;
      a.B = P_ + q0 * nts * NQ2; a.sBb = nts * NQ2; a.sBko = 0; a.sBki = NQ2; a.sBno = 0; a.sBni = 1;
      a.C = W0; a.sCb = nwb * NQ2; a.sCm = NQ2; a.sCno = 0; a.sCni = 1;
      a.alpha = make_double2(1.0, 0.0);
      zgemm_f64(a, true, false, stream_);
    }
    mark(e1);
    const int64_t nmat = qc * nwb;
    if (nmat > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "more than 65535 Dyson matrices in one q chunk");
    hipLaunchKernelGGL(identity_minus_kernel, dim3((unsigned)ceil_div(nmat * NQ2, 256)), dim3(256), 0, stream_, W1, W0,
                       (int)NQ, nmat);
    GWHIP_CHECK(hipGetLastError());
    timed(K_DYSON_INV, [&] { block_inverse(W1, nmat); });
    mirror_upper(W1, (int)NQ, NQ2, nmat, stream_);  // Herm(A^-1) from its upper triangle
    GWHIP_CHECK(hipGetLastError());
    {  // P-tilde(i Omega) = Herm(A^-1) P0(i Omega): A^-1 = (I - P0)^-1 commutes with P0, so the
       // product is Hermitian and only its upper tiles are formed, then mirrored
      ZGemmArgs a{};
      a.M = (int)NQ; a.N = (int)NQ; a.K = (int)NQ; a.batch = (int)nmat; a.kin = (int)NQ; a.nin = (int)NQ;
      a.A = W1; a.sAb = NQ2; a.sAm = NQ; a.sAko = 0; a.sAki = 1;
      a.B = W0; a.sBb = NQ2; a.sBko = 0; a.sBki = NQ; a.sBno = 0; a.sBni = 1;
      a.C = W2; a.sCb = NQ2; a.sCm = NQ; a.sCno = 0; a.sCni = 1;
      a.alpha = make_double2(1.0, 0.0);
      a.upper = 1;
      zgemm_f64(a, true, false, stream_);
    }
    mirror_upper(W2, (int)NQ, NQ2, nmat, stream_);
    GWHIP_CHECK(hipGetLastError());
    if (q0_eps_) {  // aux_to_PW_00 for the irreducible q of this chunk
      hipLaunchKernelGGL(q0_eps_kernel, dim3((unsigned)nwb, (unsigned)qc), dim3(256), 0, stream_, W2, q0_A_, q0_row,
                         (int)q0, (int)nwb, (int)NQ, (int)q0_inq_, q0_eps_dev_);
      GWHIP_CHECK(hipGetLastError());
    }
    mark(e2);
    {  // P-tilde(i Omega_b) -> P-tilde(tau_f), all nts
      ZGemmArgs a{};
      a.M = (int)nts; a.N = (int)NQ2; a.K = (int)nwb; a.batch = (int)qc; a.kin = (int)nwb; a.nin = (int)NQ2;
      a.A = Ttb_; a.sAb = 0; a.sAm = nwb; a.sAko = 0; a.sAki = 1;
      a.B = W2; a.sBb = nwb * NQ2; a.sBko = 0; a.sBki = NQ2; a.sBno = 0; a.sBni = 1;
      a.C = P_ + q0 * nts * NQ2; a.sCb = nts * NQ2; a.sCm = NQ2; a.sCno = 0; a.sCni = 1;
      a.alpha = make_double2(1.0, 0.0);
      zgemm_f64(a, true, false, stream_);
    }
    mark(e3);
    if (timing_) {
      GWHIP_CHECK(hipEventSynchronize(e3));
      float x;
      GWHIP_CHECK(hipEventElapsedTime(&x, e0, e1)); tw += x;
      GWHIP_CHECK(hipEventElapsedTime(&x, e1, e2)); bse += x;
      GWHIP_CHECK(hipEventElapsedTime(&x, e2, e3)); wt += x;
    }
  }
  if (!timing_) return;
  region_ms_[R_TW] += tw;
  region_ms_[R_BSE] += bse;
  region_ms_[R_WT] += wt;
}

void gw_hip_kernel::stage_sigma(const double2* G, double2* Sigma, bool split) {
  for (int64_t q = 0; q < d_.nk; ++q)
    sigma_q(G, Sigma, split, P_ + q * d_.nts * d_.NQ * d_.NQ, Vq_neg(q), Vs_neg(q), eVs_neg(q), kminus_tab_ + q * d_.nk,
            d_.nk);
}

// V(k, k - q) lives in the k + q' store of q' = neg_q_[q], and so do its slices.
const double2* gw_hip_kernel::Vq_neg(int64_t q) const {
  return Vq_ + (int64_t)neg_q_[q] * d_.nk * d_.NQ * d_.nao * d_.nao;
}
const int8_t* gw_hip_kernel::Vs_neg(int64_t q) const {
  return Vs_s_ ? Vs_s_ + (int64_t)neg_q_[q] * d_.nk * d_.nao * d_.nao * OZ_PLANES * split_ns_ * kc_s_ : nullptr;
}
const int* gw_hip_kernel::eVs_neg(int64_t q) const {
  return eVs_s_ ? eVs_s_ + (int64_t)neg_q_[q] * d_.nk * d_.nao * d_.nao : nullptr;
}

// Sigma += -1/nk sum over the nkb k points of one q: Pq = P-tilde of this q ([nts][NQ^2], the half
// grid read), Vb / Vsb / eVsb = V(k, k - q) of the nkb points ([nkb][NQ][nao][nao]; slices and
// exponents for split precision), km = the full-mesh index of k - q for each, Sigma
// [nts][ns][nkb][nso^2].
void gw_hip_kernel::sigma_q(const double2* G, double2* Sigma, bool split, const double2* Pq, const double2* Vb,
                            const int8_t* Vsb, const int* eVsb, const int* km, int64_t nkb) {
  const int64_t nk = d_.nk, NQ = d_.NQ, NQ2 = NQ * NQ, nts = d_.nts;
  if (nkb < 1 || nkb > nk) throw status_error(GWHIP_ERR_ARG, "Sigma stage k count out of range");

  // P0(beta - tau) = P0(tau) makes P-tilde(beta - tau) = P-tilde(tau) (to the rounding of the
  // tau <-> Omega transforms, ~1e-12 relative on the IR grids), so R = P-tilde conj(V) is formed on
  // the half grid only and serves both tau and beta - tau in the epilogue.
  const int64_t half = nts / 2;
  for (int64_t t0 = 0; t0 < half; t0 += tchunk_sigma_) {
    const int64_t tc = std::min(tchunk_sigma_, half - t0);
    if (split) {
      sigma_slice(Pq + t0 * NQ2, tc);
      timed(K_SIGMA_GEMM, [&] { sigma_gemm_split(Vsb, eVsb, nkb, tc, tune_.s_group); });
    } else {
      timed(K_SIGMA_GEMM, [&] { sigma_gemm(Pq + t0 * NQ2, Vb, nkb, tc, tune_.s_shape); });
    }
    timed(K_SIGMA_EPI, [&] { sigma_epi(G, Sigma, Vb, km, nkb, t0, tc, epi_form()); });
  }
}

// Sigma += -1/nk (V G) R for the 2 tc tau of the half-grid chunk at t0 (R in scratch_) with the
// register epilogue `form` (epi_form()), or the GEMM form of the orbital products.
void gw_hip_kernel::sigma_epi(const double2* G, double2* Sigma, const double2* Vb, const int* km, int64_t nkb,
                              int64_t t0, int64_t tc, int form) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nts = d_.nts;
  const double scale = -1.0 / (double)nk;
  if (orb_gemm()) {
    sigma_epi_gemm(G, Sigma, Vb, km, nkb, t0, tc, scale);
    return;
  }
  DevDims dd{(int)nts, (int)d_.ns, (int)nk, (int)d_.nso, (int)d_.nao, (int)NQ, (int)d_.pns, (int)d_.pseudo_ns};
  const int qs = epi_qsplit(nkb);
  double2* part = nullptr;
  if (qs > 1) {
    ensure_epi_part(nkb);
    part = epi_part_;
    zero(part, (int64_t)qs * nts * d_.ns * nkb * d_.nso * d_.nso);
  }
  const dim3 grid((unsigned)nkb, (unsigned)qs, (unsigned)(2 * tc));
  const int up = sigma_upper() ? 1 : 0;
  const int n = (int)nkb, t = (int)t0;
  const int nb4 = d_.nao <= 8 ? 2 : d_.nao <= 16 ? 4 : d_.nao <= 28 ? 7 : d_.nao <= 32 ? 8 : d_.nao <= 40 ? 10 : 12;
  const XcdGrid xg = make_xcd_grid(grid);
  auto go = [&](auto kern, int threads, size_t l4) {
    allow_lds(kern, l4);
    hipLaunchKernelGGL(kern, xcd_launch_grid(xg), dim3(threads), l4, stream_, xg, dd, G, Vb, scratch_, Sigma, km, n, t,
                       scale, up, 1, (int)grid.y, part);
  };
  if (form == 1 && d_.nao <= 32) {  // 4 x 4 x 4 MFMAs, every i group in each wave
    const size_t l4 = sigma_epi4_lds_bytes(nb4);
    if (nb4 == 2) go(sigma_epi4_kernel<2>, 256, l4);
    else if (nb4 == 4) go(sigma_epi4_kernel<4>, 256, l4);
    else if (nb4 == 7) go(sigma_epi4_kernel<7>, 256, l4);
    else go(sigma_epi4_kernel<8>, 256, l4);
  } else if (form == 2) {  // 4 x 4 x 4 MFMAs, one i group per wave
    if (nb4 == 2) go(sigma_epi4w_kernel<2, 4, 4>, 256, sigma_epi4w_lds_bytes<2, 4>());
    else if (nb4 == 4) go(sigma_epi4w_kernel<4, 4, 4>, 256, sigma_epi4w_lds_bytes<4, 4>());
    else if (nb4 == 7) go(sigma_epi4w_kernel<7, 2, 4>, 256, sigma_epi4w_lds_bytes<7, 2>());
    else if (nb4 == 8) go(sigma_epi4w_kernel<8, 2, 4>, 256, sigma_epi4w_lds_bytes<8, 2>());
    else if (nb4 == 10) go(sigma_epi4w_kernel<10, 1, 3>, 192, sigma_epi4w_lds_bytes<10, 1>());
    else go(sigma_epi4w_kernel<12, 1, 3>, 192, sigma_epi4w_lds_bytes<12, 1>());
  } else {  // 16 x 16 x 4 MFMAs
    const int NB = (int)ceil_div(d_.nao, 16);
    const size_t lds = ((size_t)2 * (16 * NB) * (16 * NB + 1) + (size_t)NB * NB * 8 * 64) * sizeof(double);
    if (NB == 1) launch_sigma_epi<1>(grid, lds, stream_, dd, G, Vb, scratch_, Sigma, km, n, t, scale, up, 1, part);
    else if (NB == 2) launch_sigma_epi<2>(grid, lds, stream_, dd, G, Vb, scratch_, Sigma, km, n, t, scale, up, 1, part);
    else launch_sigma_epi<3>(grid, lds, stream_, dd, G, Vb, scratch_, Sigma, km, n, t, scale, up, 1, part);
  }
  GWHIP_CHECK(hipGetLastError());
  if (qs > 1) {
    const int64_t per_tau = d_.ns * nkb * d_.nso * d_.nso;
    const int64_t nel = 2 * tc * per_tau;
    hipLaunchKernelGGL(epi_part_reduce_kernel, dim3((unsigned)std::min<int64_t>(4096, ceil_div(nel, 256))), dim3(256), 0,
                       stream_, part, Sigma, qs, nts * per_tau, per_tau, t, (int)tc, (int)nts);
    GWHIP_CHECK(hipGetLastError());
  }
}

// Sigma of Hermitian G is Hermitian; with one spin block per (tau, s, k) only its upper 16 x 16
// blocks were accumulated, so fill the rest from them.
bool gw_hip_kernel::sigma_upper() const { return g_herm_ && d_.pns == 1 && !orb_gemm(); }

// The fused P0 prep + slicing kernel takes one or two collinear spin blocks (their G planes share
// the work-group's LDS: 150 KB at nao = 48 with two).
bool gw_hip_kernel::fused_p0() const { return d_.pns == 1 && d_.pseudo_ns <= 2 && !orb_gemm(); }

// The register Sigma epilogue for this nao: the tuned form where it applies, else by measurement
// (one MI355X, split Sigma-evals): 4 x 4 x 4 MFMAs with all i groups per wave up to nao = 32
// (nao = 26: 12.2 ms per launch against 13.6 for 16 x 16 x 4), one i group per wave up to 40
// (nao = 40, ns = 2: 55.7 against 63.9 ms), and 16 x 16 x 4 above, where both pad to 48 (nao = 48:
// 29.0 against 31.3 ms).
int gw_hip_kernel::epi_form() const {
  if (tune_.epi_form == 1 && d_.nao <= 32) return 1;
  if (tune_.epi_form == 2 || tune_.epi_form == 3) return tune_.epi_form;
  return d_.nao <= 32 ? 1 : d_.nao <= 40 ? 2 : 3;
}

int gw_hip_kernel::epi_qsplit(int64_t nks) const {
  // the full-mesh launch has nks x nts work-groups; under ~2048 (eight per CU) the Q loop is split
  const int64_t blocks = nks * d_.nts;
  if (orb_gemm() || blocks >= 2048 || d_.NQ < 32) return 1;
  return (int)std::min<int64_t>(4, std::min<int64_t>(ceil_div(2048, blocks), d_.NQ / 16));
}

void gw_hip_kernel::ensure_epi_part(int64_t nks) {
  const int qs = epi_qsplit(nks);
  if (qs <= 1) return;
  const int64_t need = (int64_t)qs * d_.nts * d_.ns * nks * d_.nso * d_.nso;
  if (need <= epi_part_elems_) return;
  if (epi_part_) (void)hipFree(epi_part_), bytes_ -= (size_t)epi_part_elems_ * sizeof(double2), epi_part_ = nullptr;
  epi_part_elems_ = 0;
  epi_part_ = dalloc<double2>((size_t)need);
  epi_part_elems_ = need;
}

void gw_hip_kernel::ensure_orb_buffers() {
  ensure_epi_part(d_.nk);
  if (!orb_gemm()) return;
  if (!orb_W_) orb_W_ = dalloc<double2>((size_t)d_.nk * d_.NQ * d_.nao * d_.nao);
  if (!orb_G_) orb_G_ = dalloc<double2>((size_t)d_.nk * d_.nso * d_.nso);
  if (!orb_part_) orb_part_ = dalloc<double2>((size_t)kOrbKsplit * d_.nk * d_.nao * d_.nao);
  if (!orb_O_) orb_O_ = dalloc<double2>((size_t)d_.nk * d_.nao * d_.nao);
}

// orb_G_[k] = G(tau, is, map[k]) for k < n.
void gw_hip_kernel::gather_g(const double2* G, int64_t tau, int64_t is, const int* map, int64_t n) {
  gather_mats(G + (tau * d_.ns + is) * d_.nk * d_.nso * d_.nso, map, n);
}

// orb_G_[k] = base[map[k]] (nso x nso matrices) for k < n.
void gw_hip_kernel::gather_mats(const double2* base, const int* map, int64_t n) {
  const int64_t nso2 = d_.nso * d_.nso;
  if (n < 1 || n > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "gather: k count out of range");
  hipLaunchKernelGGL(gather_g_kernel, dim3((unsigned)ceil_div(nso2, 256), (unsigned)n), dim3(256), 0, stream_, base,
                     orb_G_, map, nso2);
  GWHIP_CHECK(hipGetLastError());
}

// The P0 orbital product as two batched FP64 GEMMs over k per (tau, spin block), for any nao:
//   W[k][Q][p][n] = sum_m V(k, k+q)_Q[p][m] Gq(k)[m][n],     Gq = G(tau, is, k + q) block (a, b)
//   Z[k][Q][t][n] (+)= sum_p Gb(k)[t][p] W[k][Q][p][n],      Gb = G(beta - tau, is, k) block (b, a)
// (the same Z_Q = Gb V_Q Gq as z_prep_mfma_kernel, src/gw_cpu_kernel.cpp:200-215).
void gw_hip_kernel::p0_prep_gemm(int64_t q, const double2* G, int64_t t0, int64_t tc) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao = d_.nao, nao2 = nao * nao, nso = d_.nso, nso2 = nso * nso;
  const int64_t vq = NQ * nao2;
  if (!orb_W_ || !orb_G_) throw status_error(GWHIP_ERR_STATE, "orbital GEMM buffers not allocated");
  for (int64_t t = 0; t < tc; ++t) {
    const int64_t tau = t0 + t, tt = d_.nts - 1 - tau;
    for (int64_t s = 0; s < d_.pseudo_ns; ++s) {
      const int64_t is = s / (d_.pns * d_.pns), a = (s / d_.pns) % d_.pns, b = s % d_.pns;
      gather_g(G, tau, is, kplus_tab_ + q * nk, nk);
      ZGemmArgs w{};
      w.M = (int)(NQ * nao); w.N = (int)nao; w.K = (int)nao; w.batch = (int)nk; w.kin = (int)nao; w.nin = (int)nao;
      w.A = Vq_ + q * nk * vq; w.sAb = vq; w.sAm = nao; w.sAko = 0; w.sAki = 1;
      w.B = orb_G_ + a * nao * nso + b * nao; w.sBb = nso2; w.sBko = 0; w.sBki = nso; w.sBno = 0; w.sBni = 1;
      w.C = orb_W_; w.sCb = vq; w.sCm = nao; w.sCno = 0; w.sCni = 1;
      w.alpha = make_double2(1.0, 0.0);
      zgemm_f64(w, true, false, stream_);
      ZGemmArgs z{};
      z.M = (int)nao; z.N = (int)(NQ * nao); z.K = (int)nao; z.batch = (int)nk; z.kin = (int)nao; z.nin = (int)nao;
      z.A = G + (tt * d_.ns + is) * nk * nso2 + b * nao * nso + a * nao; z.sAb = nso2; z.sAm = nso; z.sAko = 0;
      z.sAki = 1;
      z.B = orb_W_; z.sBb = vq; z.sBko = 0; z.sBki = nao; z.sBno = nao2; z.sBni = 1;
      z.C = scratch_ + t * nk * vq; z.sCb = vq; z.sCm = nao; z.sCno = nao2; z.sCni = 1;
      z.alpha = make_double2(1.0, 0.0);
      z.accumulate = s > 0 ? 1 : 0;
      zgemm_f64(z, true, false, stream_);
    }
  }
}

// The Sigma epilogue as two batched FP64 GEMMs over the nkb k points per (tau, spin block), for
// any nao; R(tz) in scratch_ [tz][nkb][Q][n][j] serves tau = t0 + tz and beta - tau:
//   XT[k][Q][n][i] = sum_m G(k)[m][n] V_Q[i][m],   G = G(tau, is, km[k]) block (a, b), i.e. X = V_Q G
//   Sigma(tau, is, k)[a,b](i, j) += scale sum_{Q, n} X[i][Q, n] R[k][Q][n][j]
// (sigma_epi_mfma_kernel's products, src/gw_cpu_kernel.cpp:359-373).
void gw_hip_kernel::sigma_epi_gemm(const double2* G, double2* Sigma, const double2* Vb, const int* km, int64_t nkb,
                                   int64_t t0, int64_t tc, double scale) {
  const int64_t NQ = d_.NQ, nao = d_.nao, nao2 = nao * nao, nso = d_.nso, nso2 = nso * nso, vq = NQ * nao2;
  if (!orb_W_ || !orb_G_) throw status_error(GWHIP_ERR_STATE, "orbital GEMM buffers not allocated");
  for (int64_t tz = 0; tz < tc; ++tz)
    for (int side = 0; side < 2; ++side) {
      const int64_t tau = side ? d_.nts - 1 - (t0 + tz) : t0 + tz;
      for (int64_t s = 0; s < d_.pseudo_ns; ++s) {
        const int64_t is = s / (d_.pns * d_.pns), a = (s / d_.pns) % d_.pns, b = s % d_.pns;
        gather_g(G, tau, is, km, nkb);
        ZGemmArgs x{};
        x.M = (int)nao; x.N = (int)(NQ * nao); x.K = (int)nao; x.batch = (int)nkb; x.kin = (int)nao;
        x.nin = (int)nao;
        x.A = orb_G_ + a * nao * nso + b * nao; x.sAb = nso2; x.sAm = 1; x.sAko = 0; x.sAki = nso;
        x.B = Vb; x.sBb = vq; x.sBko = 0; x.sBki = 1; x.sBno = nao2; x.sBni = nao;
        x.C = orb_W_; x.sCb = vq; x.sCm = nao; x.sCno = nao2; x.sCni = 1;
        x.alpha = make_double2(1.0, 0.0);
        zgemm_f64(x, false, false, stream_);
        ZGemmArgs y{};
        y.M = (int)nao; y.N = (int)nao; y.K = (int)(NQ * nao); y.batch = (int)nkb; y.kin = (int)nao;
        y.nin = (int)nao;
        y.A = orb_W_; y.sAb = vq; y.sAm = 1; y.sAko = nao2; y.sAki = nao;
        y.B = scratch_ + tz * nkb * vq; y.sBb = vq; y.sBko = nao2; y.sBki = nao; y.sBno = 0; y.sBni = 1;
        y.C = Sigma + (tau * d_.ns + is) * nkb * nso2 + a * nao * nso + b * nao; y.sCb = nso2; y.sCm = nso;
        y.sCno = 0; y.sCni = 1;
        y.alpha = make_double2(scale, 0.0);
        y.accumulate = 1;
        // only nkb x (nao / 64)^2 output tiles, each over K = NQ nao: cut K so that ~1024 blocks run
        const int64_t tiles = nkb * ceil_div(nao, 64) * ceil_div(nao, 64);
        const int ksplit = (int)std::min<int64_t>(kOrbKsplit, std::max<int64_t>(1, ceil_div(1024, tiles)));
        zgemm_f64_splitk(y, ksplit, orb_part_, false, false, stream_);
      }
    }
}

void gw_hip_kernel::finish_sigma(double2* Sigma, int64_t nkS) {
  if (!sigma_upper()) return;
  const int64_t n = d_.nso, nmat = d_.nts * d_.ns * nkS;
  for (int64_t m0 = 0; m0 < nmat; m0 += 65535) {  // grid.y limit
    const int64_t mc = std::min<int64_t>(65535, nmat - m0);
    mirror_upper(Sigma + m0 * n * n, (int)n, n * n, mc, stream_);
    GWHIP_CHECK(hipGetLastError());
  }
}

// One evaluation on the stream, kernel nodes only (Sigma zeroed by a fill kernel).
void gw_hip_kernel::enqueue_eval(const double2* G, double2* Sigma, int precision, int* nonfinite_flag) {
  const int64_t nsig = d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  zero(Sigma, nsig);
  stage_p0(G, p0_split(precision));
  stage_dyson(d_.nk, q0_qiq_);
  stage_sigma(G, Sigma, sigma_split(precision));
  finish_sigma(Sigma, d_.nk);
  if (nonfinite_flag) {
    hipLaunchKernelGGL(nonfinite_kernel, dim3(1024), dim3(256), 0, stream_, Sigma, nsig, nonfinite_flag);
    GWHIP_CHECK(hipGetLastError());
  }
}

void gw_hip_kernel::solve_device(const double2* G, double2* Sigma, int precision, bool q0_apply) {
  check_ready(precision);
  if (!G || !Sigma) throw status_error(GWHIP_ERR_ARG, "null G or Sigma");
  GWHIP_CHECK(hipSetDevice(device_));
  if (any_split(precision)) ensure_slices();
  timing_ = true;
  resolve_timings();
  detect_hermitian_G(G);
  for (double& r : region_ms_) r = 0.0;
  for (auto& k : kstat_) k = KernelStat{};
  const int64_t nsig = d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  ensure_orb_buffers();

  Event t0, t1, t2, t3;
  GWHIP_CHECK(hipEventRecord(t0, stream_));
  zero(Sigma, nsig);
  stored_valid_ = false;
  leak_valid_ = false;
  stage_p0(G, p0_split(precision));
  capture_leakage_p0();
  if (keep_p0_) {  // half-grid P0 of every q, for the stored-mode VJP
    const size_t row = (size_t)(d_.nts / 2) * d_.NQ * d_.NQ * sizeof(double2);
    GWHIP_CHECK(hipMemcpy2DAsync(P0keep_, row, P_, (size_t)d_.nts * d_.NQ * d_.NQ * sizeof(double2), row, (size_t)d_.nk,
                                 hipMemcpyDeviceToDevice, stream_));
  }
  GWHIP_CHECK(hipEventRecord(t1, stream_));
  q0_eps_ = q0_on_;
  stage_dyson(d_.nk, q0_qiq_);
  q0_eps_ = false;
  GWHIP_CHECK(hipEventRecord(t2, stream_));
  stage_sigma(G, Sigma, sigma_split(precision));
  finish_sigma(Sigma, d_.nk);
  if (q0_on_ && q0_apply) apply_q0(Sigma, G, d_.nk);
  GWHIP_CHECK(hipEventRecord(t3, stream_));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, flag_, (int64_t)1);
  hipLaunchKernelGGL(nonfinite_kernel, dim3(1024), dim3(256), 0, stream_, Sigma, nsig, flag_);
  GWHIP_CHECK(hipGetLastError());
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  resolve_timings();
  float x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t0, t1)); region_ms_[R_P0] = x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t2, t3)); region_ms_[R_SIGMA] = x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t0, t3)); region_ms_[R_TOTAL] = x;
  // non-finite first: a NaN or Inf in G also fails the Dyson step's pivot test
  int flag = 0;
  GWHIP_CHECK(hipMemcpy(&flag, flag_, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) throw status_error(GWHIP_ERR_NONFINITE, "Sigma contains non-finite values");
  check_dyson_flag();
  stored_valid_ = keep_p0_;
  leak_valid_ = leak_on_;
}

// Time the launch-parameter candidates on this system's own operands (q = 0, first tau chunk)
// and keep the fastest.  Overwrites P0/P-tilde of q = 0 and the scratch (no results kept).
void gw_hip_kernel::autotune(const double2* G, int precision) {
  check_ready(precision);
  if (!G) throw status_error(GWHIP_ERR_ARG, "null G");
  GWHIP_CHECK(hipSetDevice(device_));
  const bool split = any_split(precision);
  if (split) ensure_slices();
  ensure_orb_buffers();
  detect_hermitian_G(G);
  const int64_t half = d_.nts / 2, tcp = std::min(tchunk_p0_, half), tcs = std::min(tchunk_sigma_, half);
  Event a, b;
  auto best_of = [&](int ncand, const int* cand, auto&& run) {
    run(cand[0]);  // warm-up (module load, LDS attribute)
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    int best = cand[0];
    float best_ms = 1e30f;
    for (int i = 0; i < ncand; ++i) {
      float fastest = 1e30f;  // the fastest of three runs: one slow run (clock ramp, a neighbour) does not decide
      for (int r = 0; r < 3; ++r) {
        GWHIP_CHECK(hipEventRecord(a, stream_));
        run(cand[i]);
        GWHIP_CHECK(hipEventRecord(b, stream_));
        GWHIP_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        GWHIP_CHECK(hipEventElapsedTime(&ms, a, b));
        fastest = std::min(fastest, ms);
      }
      if (fastest < best_ms) {
        best_ms = fastest;
        best = cand[i];
      }
    }
    return best;
  };
  const int qpws[4] = {1, 2, 4, 8};
  tune_.qpw = best_of(4, qpws, [&](int c) { p0_prep(0, G, 0, tcp, c); });
  if (!split) {
    const int shapes[3] = {0, 1, 2};
    p0_prep(0, G, 0, tcp, tune_.qpw);
    tune_.p0_shape = best_of(3, shapes, [&](int c) { p0_gemm(0, tcp, P_, c); });
    tune_.s_shape = best_of(3, shapes, [&](int c) { sigma_gemm(P_, Vq_neg(0), d_.nk, tcs, c); });
  } else {
    if (fused_p0()) {  // fused prep + slicing vs the two kernels
      const int fz[2] = {0, 1};
      tune_.p0_fused = best_of(2, fz, [&](int c) {
        if (c) {
          p0_prep_slice(0, G, 0, tcp, tune_.qpw);
        } else {
          p0_prep(0, G, 0, tcp, tune_.qpw);
          p0_slice(tcp);
        }
      });
    }
    // block tile of each int8 GEMM, then the row-tile group of the L2-aware tile order
    const int groups[7] = {1, 2, 4, 8, 16, 32, 64};
    const int tiles[2] = {2, 4};
    p0_prep(0, G, 0, tcp, tune_.qpw);
    p0_slice(tcp);
    tune_.p0_oz_tile = best_of(2, tiles, [&](int c) {
      tune_.p0_oz_tile = c;
      p0_gemm_split(0, tcp, P_, tune_.p0_group, tune_.p0_kchunks);
    });
    tune_.p0_group = best_of(7, groups, [&](int c) { p0_gemm_split(0, tcp, P_, c, tune_.p0_kchunks); });
    const int kch[6] = {0, 16, 8, 4, 2, 1};
    tune_.p0_kchunks = best_of(6, kch, [&](int c) { p0_gemm_split(0, tcp, P_, tune_.p0_group, c); });
    sigma_slice(P_, tcs);
    tune_.s_oz_tile = best_of(2, tiles, [&](int c) {
      tune_.s_oz_tile = c;
      sigma_gemm_split(Vs_neg(0), eVs_neg(0), d_.nk, tcs, tune_.s_group);
    });
    tune_.s_group = best_of(7, groups, [&](int c) { sigma_gemm_split(Vs_neg(0), eVs_neg(0), d_.nk, tcs, c); });
  }
  // register Sigma epilogue form, on the R of q = 0 the Sigma GEMMs above left in scratch_ (Sigma
  // into a temporary buffer; where device memory cannot hold one, the form stays chosen by nao)
  double2* S = nullptr;
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  if (!orb_gemm()) {
    try {
      S = dalloc<double2>(nsig);
    } catch (const status_error& e) {
      if (e.status != GWHIP_ERR_OOM) throw;
      tune_.epi_form = 0;
    }
  }
  if (S) {
    const int forms[3] = {1, 2, 3};
    const int first = d_.nao <= 32 ? 0 : 1;  // the all-groups form covers nao <= 32
    try {
      tune_.epi_form = best_of(3 - first, forms + first,
                               [&](int c) { sigma_epi(G, S, Vq_neg(0), kminus_tab_, d_.nk, 0, tcs, c); });
    } catch (...) {
      (void)hipFree(S);
      bytes_ -= nsig * sizeof(double2);
      throw;
    }
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    (void)hipFree(S);
    bytes_ -= nsig * sizeof(double2);
  }
  stored_valid_ = false;
  ++generation_;
}

void gw_hip_kernel::solve_host(const double2* G, double2* Sigma, int precision) {
  if (!G || !Sigma) throw status_error(GWHIP_ERR_ARG, "null G or Sigma");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  if (!G_dev_) G_dev_ = dalloc<double2>(nsig);
  if (!S_dev_) S_dev_ = dalloc<double2>(nsig);
  GWHIP_CHECK(hipMemcpy(G_dev_, G, nsig * sizeof(double2), hipMemcpyHostToDevice));
  solve_device(G_dev_, S_dev_, precision);
  GWHIP_CHECK(hipMemcpy(Sigma, S_dev_, nsig * sizeof(double2), hipMemcpyDeviceToHost));
}

namespace {
constexpr size_t kWsAlign = 256;
inline size_t ws_round(size_t b) { return (b + kWsAlign - 1) / kWsAlign * kWsAlign; }
}  // namespace

// Byte layout of solve_device_in's workspace for tc tau points and qc q values per chunk:
// [flags 256 B | epilogue partials | Dyson work | orbital scratch | int8 slices | exponents].
struct gw_hip_kernel::WsPlan {
  size_t epi = 0, dyson = 0, scratch_elems = 0, scratch = 0, slices = 0, exps = 0, total = 0;
};

gw_hip_kernel::WsPlan gw_hip_kernel::ws_plan(int64_t tc, int64_t qc, bool split) const {
  WsPlan p;
  const int qs = epi_qsplit(d_.nk);
  if (qs > 1) p.epi = ws_round((size_t)qs * d_.nts * d_.ns * d_.nk * d_.nso * d_.nso * sizeof(double2));
  p.dyson = ws_round((size_t)qc * dyson_per_q() * sizeof(double2));
  const size_t nb = (size_t)std::min<int64_t>(64, d_.NQ);  // block_inverse's panels for one matrix
  p.scratch_elems = std::max((size_t)tc * scratch_per_t(), 2 * (size_t)d_.NQ * nb + nb * nb);
  p.scratch = ws_round(p.scratch_elems * sizeof(double2));
  if (split) {
    p.slices = ws_round((size_t)tc * slice_bytes_per_t());
    p.exps = ws_round((size_t)tc * exps_per_t() * sizeof(int));
  }
  p.total = kWsAlign + p.epi + p.dyson + p.scratch + p.slices + p.exps;
  return p;
}

size_t gw_hip_kernel::min_workspace_bytes(int precision) const { return ws_plan(1, 1, any_split(precision)).total; }

void gw_hip_kernel::solve_device_in(const double2* G, double2* Sigma, int precision, void* ws, size_t ws_bytes) {
  if (!grid_set_) throw status_error(GWHIP_ERR_STATE, "transform grid not set (gwhip_set_grid)");
  if (!v_set_) throw status_error(GWHIP_ERR_STATE, "integrals not set (gwhip_upload_V / gwhip_generate_V)");
  if (!precision_valid(precision)) throw status_error(GWHIP_ERR_ARG, "unknown precision");
  if (!G || !Sigma || !ws) throw status_error(GWHIP_ERR_ARG, "null G, Sigma or workspace");
  if ((uintptr_t)ws % kWsAlign != 0) throw status_error(GWHIP_ERR_ARG, "workspace not 256-byte aligned");
  if (orb_gemm())
    throw status_error(GWHIP_ERR_UNSUPPORTED, "the orbital GEMM form (nao > 48 or tuned on) needs buffers of its own");
  if (q0_on_) throw status_error(GWHIP_ERR_UNSUPPORTED, "the q -> 0 correction runs a host fit between stages");
  const bool split = any_split(precision);
  if (split && !split_ready())
    throw status_error(GWHIP_ERR_STATE, "int8 slice stores of V not built: evaluate once in a split precision first");
  const WsPlan p1 = ws_plan(1, 1, split);
  if (p1.total > ws_bytes)
    throw status_error(GWHIP_ERR_UNSUPPORTED, "this system needs " + std::to_string(p1.total) +
                                                  " bytes of workspace at one tau point and one q; " +
                                                  std::to_string(ws_bytes) + " are available");
  // as many tau points per chunk as fit (at most the half grid), then as many q
  const int64_t half = d_.nts / 2;
  int64_t tc = 1, qc = 1;
  while (tc < half && ws_plan(tc + 1, 1, split).total <= ws_bytes) ++tc;
  while (qc < d_.nk && ws_plan(tc, qc + 1, split).total <= ws_bytes) ++qc;
  const WsPlan p = ws_plan(tc, qc, split);
  GWHIP_CHECK(hipSetDevice(device_));

  // point the transient buffers into the workspace for this call; the handle's own come back after
  struct Swap {
    gw_hip_kernel& k;
    double2 *scratch, *Pw, *epi;
    int8_t *Zs, *Ps;
    int *eZs, *ePs, *flag, *dflag;
    unsigned long long* hcheck;
    int64_t scratch_elems, epi_elems, qc, tp0, tsig;
    bool timing;
    explicit Swap(gw_hip_kernel& x)
        : k(x), scratch(x.scratch_), Pw(x.Pw_), epi(x.epi_part_), Zs(x.Zs_), Ps(x.Ps_), eZs(x.eZs_), ePs(x.ePs_),
          flag(x.flag_), dflag(x.dyson_flag_), hcheck(x.hcheck_), scratch_elems(x.scratch_elems_),
          epi_elems(x.epi_part_elems_), qc(x.qchunk_), tp0(x.tchunk_p0_), tsig(x.tchunk_sigma_), timing(x.timing_) {}
    ~Swap() {
      k.scratch_ = scratch; k.Pw_ = Pw; k.epi_part_ = epi; k.Zs_ = Zs; k.Ps_ = Ps; k.eZs_ = eZs; k.ePs_ = ePs;
      k.flag_ = flag; k.dyson_flag_ = dflag; k.hcheck_ = hcheck; k.scratch_elems_ = scratch_elems;
      k.epi_part_elems_ = epi_elems; k.qchunk_ = qc; k.tchunk_p0_ = tp0; k.tchunk_sigma_ = tsig; k.timing_ = timing;
    }
  } restore(*this);
  char* base = static_cast<char*>(ws);
  flag_ = reinterpret_cast<int*>(base);
  dyson_flag_ = reinterpret_cast<int*>(base + 4);
  hcheck_ = reinterpret_cast<unsigned long long*>(base + 16);
  size_t off = kWsAlign;
  epi_part_ = p.epi ? reinterpret_cast<double2*>(base + off) : nullptr;
  epi_part_elems_ = (int64_t)(p.epi / sizeof(double2));
  off += p.epi;
  Pw_ = reinterpret_cast<double2*>(base + off);
  off += p.dyson;
  scratch_ = reinterpret_cast<double2*>(base + off);
  scratch_elems_ = (int64_t)p.scratch_elems;
  off += p.scratch;
  Zs_ = Ps_ = split ? reinterpret_cast<int8_t*>(base + off) : nullptr;
  off += p.slices;
  eZs_ = ePs_ = split ? reinterpret_cast<int*>(base + off) : nullptr;
  qchunk_ = qc;
  tchunk_p0_ = tchunk_sigma_ = tc;
  timing_ = false;
  stored_valid_ = false;
  leak_valid_ = false;

  detect_hermitian_G(G);
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, flag_, (int64_t)1);
  GWHIP_CHECK(hipGetLastError());
  enqueue_eval(G, Sigma, precision, flag_);
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  int flags[2] = {0, 0};
  GWHIP_CHECK(hipMemcpy(flags, base, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags[0]) throw status_error(GWHIP_ERR_NONFINITE, "Sigma contains non-finite values");
  if (flags[1])
    throw status_error(GWHIP_ERR_NOT_POSITIVE_DEFINITE,
                       "I - P0(i Omega) is not positive definite for some (q, Omega): a pivot of its factorisation "
                       "is not positive (P0 is not the polarisation of a physical G)");
}

void gw_hip_kernel::set_leakage_capture(bool on) {
  GWHIP_CHECK(hipSetDevice(device_));
  if (on && !leak_P0_) leak_P0_ = dalloc<double2>((size_t)(d_.nts / 2) * d_.NQ * d_.NQ);
  leak_on_ = on;
  leak_valid_ = false;
}

// Half-grid P0 of P_ row 0 (the first q evaluated) into leak_P0_, before the Dyson stage overwrites it.
void gw_hip_kernel::capture_leakage_p0() {
  if (!leak_on_) return;
  const size_t n = (size_t)(d_.nts / 2) * d_.NQ * d_.NQ;
  GWHIP_CHECK(hipMemcpyAsync(leak_P0_, P_, n * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
}

// P0 on the full tau grid from its half grid: P0(beta - tau) = P0(tau) (src/gw_cpu_kernel.cpp:217-225).
static void mirror_half_grid(std::vector<double2>& half, double2* full, int64_t nts, int64_t n2) {
  for (int64_t t = 0; t < nts / 2; ++t) {
    std::memcpy(full + t * n2, half.data() + t * n2, n2 * sizeof(double2));
    std::memcpy(full + (nts - 1 - t) * n2, half.data() + t * n2, n2 * sizeof(double2));
  }
}

void gw_hip_kernel::copy_leakage_inputs(double2* P0, double2* Pt) {
  if (!leak_on_ || !leak_valid_)
    throw status_error(GWHIP_ERR_STATE, "no leakage inputs: enable the capture, then evaluate");
  GWHIP_CHECK(hipSetDevice(device_));
  const int64_t n2 = d_.NQ * d_.NQ, half = d_.nts / 2;
  if (P0) {
    std::vector<double2> h((size_t)half * n2);
    GWHIP_CHECK(hipMemcpy(h.data(), leak_P0_, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    mirror_half_grid(h, P0, d_.nts, n2);
  }
  if (Pt) GWHIP_CHECK(hipMemcpy(Pt, P_, (size_t)d_.nts * n2 * sizeof(double2), hipMemcpyDeviceToHost));
}

void gw_hip_kernel::eval_one_q_host(const double2* G, int64_t q, int precision, double2* P0, double2* Pt,
                                    double2* Sq) {
  check_ready(precision);
  if (!G) throw status_error(GWHIP_ERR_ARG, "null G");
  if (q < 0 || q >= d_.nk) throw status_error(GWHIP_ERR_ARG, "q out of range");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  const int64_t n2 = d_.NQ * d_.NQ, half = d_.nts / 2;
  if (!G_dev_) G_dev_ = dalloc<double2>(nsig);
  if (!S_dev_) S_dev_ = dalloc<double2>(nsig);
  GWHIP_CHECK(hipMemcpy(G_dev_, G, nsig * sizeof(double2), hipMemcpyHostToDevice));
  if (any_split(precision)) ensure_slices();
  ensure_orb_buffers();
  stored_valid_ = false;
  leak_valid_ = false;
  detect_hermitian_G(G_dev_);
  p0_one_q(q, G_dev_, p0_split(precision), P_);
  if (P0) {
    std::vector<double2> h((size_t)half * n2);
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    GWHIP_CHECK(hipMemcpy(h.data(), P_, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    mirror_half_grid(h, P0, d_.nts, n2);
  }
  stage_dyson(1, nullptr);
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  check_dyson_flag();
  if (Pt) GWHIP_CHECK(hipMemcpy(Pt, P_, (size_t)d_.nts * n2 * sizeof(double2), hipMemcpyDeviceToHost));
  if (Sq) {
    zero(S_dev_, (int64_t)nsig);
    sigma_q(G_dev_, S_dev_, sigma_split(precision), P_, Vq_neg(q), Vs_neg(q), eVs_neg(q), kminus_tab_ + q * d_.nk,
            d_.nk);
    finish_sigma(S_dev_, d_.nk);
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    GWHIP_CHECK(hipMemcpy(Sq, S_dev_, nsig * sizeof(double2), hipMemcpyDeviceToHost));
  }
  resolve_timings();
}

void gw_hip_kernel::run_stages_host(const double2* G, int last_stage) {
  if (!grid_set_ || !v_set_) throw status_error(GWHIP_ERR_STATE, "grid or integrals not set");
  if (!G) throw status_error(GWHIP_ERR_ARG, "null G");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  if (!G_dev_) G_dev_ = dalloc<double2>(nsig);
  GWHIP_CHECK(hipMemcpy(G_dev_, G, nsig * sizeof(double2), hipMemcpyHostToDevice));
  stored_valid_ = false;
  ensure_orb_buffers();
  detect_hermitian_G(G_dev_);
  GWHIP_CHECK(hipMemsetAsync(P_, 0, (size_t)d_.nk * d_.nts * d_.NQ * d_.NQ * sizeof(double2), stream_));
  stage_p0(G_dev_, false);
  if (last_stage >= 2) stage_dyson(d_.nk, q0_qiq_);
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  resolve_timings();
}

}  // namespace gwhip

#include "gw_vjp_host.hpp"
#include "gw_hf_host.hpp"
#include "gw_symmetry.hpp"
#include "gw_q0.hpp"
#include "gw_dyson.hpp"
#include "gw_batch.hpp"

namespace gwhip {

void debug_zgemm(int M, int N, int K, const double2* A, const double2* B, double2* C, int conjA, int conjB,
                 int shape) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) throw status_error(GWHIP_ERR_ARG, "bad debug_zgemm arguments");
  GWHIP_CHECK(hipSetDevice(0));
  double2 *dA = nullptr, *dB = nullptr, *dC = nullptr;
  GWHIP_CHECK(hipMalloc(&dA, (size_t)M * K * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dB, (size_t)K * N * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dC, (size_t)M * N * sizeof(double2)));
  GWHIP_CHECK(hipMemcpy(dA, A, (size_t)M * K * sizeof(double2), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(dB, B, (size_t)K * N * sizeof(double2), hipMemcpyHostToDevice));
  ZGemmArgs a{};
  a.M = M; a.N = N; a.K = K; a.batch = 1; a.kin = K; a.nin = N;
  a.A = dA; a.sAm = K; a.sAki = 1;
  a.B = dB; a.sBki = N; a.sBni = 1;
  a.C = dC; a.sCm = N; a.sCni = 1;
  a.alpha = make_double2(1.0, 0.0);
  a.conjA = conjA; a.conjB = conjB;
  zgemm_f64(a, true, false, nullptr, shape);
  GWHIP_CHECK(hipDeviceSynchronize());
  GWHIP_CHECK(hipMemcpy(C, dC, (size_t)M * N * sizeof(double2), hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
}

void debug_zgemm_split(int M, int N, int K, const double2* A, const double2* B, double2* C, int ns, int tile) {
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C) throw status_error(GWHIP_ERR_ARG, "bad debug_zgemm_split arguments");
  GWHIP_CHECK(hipSetDevice(0));
  const int kc = (int)ceil_div(K, 64) * 64;
  if (ns < 3 || ns > 5) throw status_error(GWHIP_ERR_ARG, "split slices must be 3, 4 or 5");
  if (kc > split_kc_limit(ns)) throw status_error(GWHIP_ERR_UNSUPPORTED, "K exceeds the int32-exact chunk limit");
  double2 *dA = nullptr, *dB = nullptr, *dC = nullptr;
  int8_t *sA = nullptr, *sB = nullptr;
  int *eA = nullptr, *eB = nullptr;
  GWHIP_CHECK(hipMalloc(&dA, (size_t)M * K * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dB, (size_t)K * N * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&dC, (size_t)M * N * sizeof(double2)));
  GWHIP_CHECK(hipMalloc(&sA, (size_t)M * OZ_PLANES * ns * kc));
  GWHIP_CHECK(hipMalloc(&sB, (size_t)N * OZ_PLANES * ns * kc));
  GWHIP_CHECK(hipMalloc(&eA, (size_t)M * sizeof(int)));
  GWHIP_CHECK(hipMalloc(&eB, (size_t)N * sizeof(int)));
  GWHIP_CHECK(hipMemcpy(dA, A, (size_t)M * K * sizeof(double2), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(dB, B, (size_t)K * N * sizeof(double2), hipMemcpyHostToDevice));
  SliceArgs a{};
  a.src = dA; a.nrows = M; a.row_div = M; a.s_rhi = 0; a.s_rlo = K; a.nchunk = 1; a.s_c = 0; a.klen = K; a.s_k = 1;
  a.kc = kc; a.conj = 0; a.dst = sA; a.exps = eA;
  slice_rows(a, ns, nullptr);
  SliceArgs bsl{};
  bsl.src = dB; bsl.nrows = N; bsl.row_div = N; bsl.s_rhi = 0; bsl.s_rlo = 1; bsl.nchunk = 1; bsl.s_c = 0; bsl.klen = K;
  bsl.s_k = N; bsl.kc = kc; bsl.conj = 0; bsl.dst = sB; bsl.exps = eB;
  slice_rows(bsl, ns, nullptr);
  OzGemmArgs o{};
  o.M = M; o.N = N; o.batch = 1; o.kc = kc; o.nchunk = 1; o.nin = N;
  o.A = sA; o.sAm = OZ_PLANES * ns * kc; o.eA = eA; o.seAm = 1;
  o.B = sB; o.sBni = OZ_PLANES * ns * kc; o.eB = eB; o.seBni = 1;
  o.C = dC; o.sCm = N; o.sCni = 1;
  o.alpha = make_double2(1.0, 0.0);
  o.tile = tile;
  zgemm_ozaki(o, ns, nullptr);
  GWHIP_CHECK(hipDeviceSynchronize());
  GWHIP_CHECK(hipMemcpy(C, dC, (size_t)M * N * sizeof(double2), hipMemcpyDeviceToHost));
  for (void* p : {(void*)dA, (void*)dB, (void*)dC, (void*)sA, (void*)sB, (void*)eA, (void*)eB}) (void)hipFree(p);
}

void gw_hip_kernel::copy_P_tilde(double2* host) const {
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t n = (size_t)d_.nk * d_.nts * d_.NQ * d_.NQ;
  GWHIP_CHECK(hipMemcpy(host, P_, n * sizeof(double2), hipMemcpyDeviceToHost));
}

}  // namespace gwhip

// ------------------------------------------------------------------------------ C ABI
struct gwhip_handle {
  std::unique_ptr<gwhip::gw_hip_kernel> owned;
  gwhip::gw_hip_kernel* k = nullptr;
  bool borrowed = false;  // a system of a batch: the batch owns it
  std::string err;
};

struct gwhip_dyson {
  std::unique_ptr<gwhip::gw_dyson> d;
  std::string err;
};

struct gwhip_batch {
  std::unique_ptr<gwhip::gw_hip_batch> b;
  std::vector<std::unique_ptr<gwhip_handle>> systems;
  std::string err;
};

namespace {
thread_local std::string g_create_error;

gwhip::Dims to_dims(const gwhip_dims* dims) {
  gwhip::Dims d;
  d.nk = dims->nk; d.nao = dims->nao; d.nso = dims->nso; d.ns = dims->ns; d.NQ = dims->NQ;
  d.nts = dims->nts; d.nw_b = dims->nw_b; d.beta = dims->beta;
  d.mesh = {dims->mesh[0], dims->mesh[1], dims->mesh[2]};
  return d;
}


template <typename F>
int guarded(gwhip_handle* h, F&& f) {
  try {
    f();
    return GWHIP_OK;
  } catch (const gwhip::status_error& e) {
    if (h) h->err = e.what();
    return e.status;
  } catch (const gwhip::hip_error& e) {
    if (h) h->err = e.what();
    return e.code == hipErrorOutOfMemory ? GWHIP_ERR_OOM : GWHIP_ERR_DEVICE;
  } catch (const std::exception& e) {
    if (h) h->err = e.what();
    return GWHIP_ERR_DEVICE;
  }
}

template <typename F>
int guarded_d(gwhip_dyson* d, F&& f) {
  gwhip_handle tmp;
  const int st = guarded(&tmp, f);
  if (st != GWHIP_OK && d) d->err = tmp.err;
  return st;
}

template <typename F>
int guarded_b(gwhip_batch* b, F&& f) {
  gwhip_handle tmp;
  const int st = guarded(&tmp, f);
  if (st != GWHIP_OK && b) b->err = tmp.err;
  return st;
}
}  // namespace

extern "C" {

int gwhip_create(const gwhip_dims* dims, int device, gwhip_handle** out) {
  if (!dims || !out) {
    g_create_error = "null argument";
    return GWHIP_ERR_ARG;
  }
  *out = nullptr;
  gwhip::Dims d;
  d.nk = dims->nk; d.nao = dims->nao; d.nso = dims->nso; d.ns = dims->ns; d.NQ = dims->NQ;
  d.nts = dims->nts; d.nw_b = dims->nw_b; d.beta = dims->beta;
  d.mesh = {dims->mesh[0], dims->mesh[1], dims->mesh[2]};
  auto h = std::make_unique<gwhip_handle>();
  int st = guarded(h.get(), [&] {
    h->owned = std::make_unique<gwhip::gw_hip_kernel>(d, device);
    h->k = h->owned.get();
  });
  if (st != GWHIP_OK) {
    g_create_error = h->err;
    return st;
  }
  *out = h.release();
  return GWHIP_OK;
}

void gwhip_destroy(gwhip_handle* h) {
  if (h && !h->borrowed) delete h;
}

const char* gwhip_last_error(const gwhip_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int gwhip_set_grid(gwhip_handle* h, const double* T_bt, const double* T_tb) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_grid((const double2*)T_bt, (const double2*)T_tb); });
}

int gwhip_set_kq_maps(gwhip_handle* h, const int32_t* kplus, const int32_t* kminus) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_kq_maps(kplus, kminus); });
}

int gwhip_set_split_slices(gwhip_handle* h, int slices) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_split_slices(slices); });
}

int gwhip_upload_V(gwhip_handle* h, const double* V) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->upload_V((const double2*)V); });
}

int gwhip_generate_V(gwhip_handle* h, uint64_t seed, double sigma) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->generate_V(seed, sigma); });
}

int gwhip_eval_sigma(gwhip_handle* h, const double* G, double* Sigma, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->solve_host((const double2*)G, (double2*)Sigma, precision); });
}

int gwhip_eval_sigma_device(gwhip_handle* h, const void* G, void* Sigma, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->solve_device((const double2*)G, (double2*)Sigma, precision); });
}

int gwhip_autotune(gwhip_handle* h, const void* G_dev, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->autotune((const double2*)G_dev, precision); });
}

int gwhip_get_tuning(const gwhip_handle* h, int* values, int n) {
  if (!h || !values || n < gwhip::Tuning::count) return GWHIP_ERR_ARG;
  h->k->tuning().to_array(values);
  return GWHIP_OK;
}

int gwhip_set_tuning(gwhip_handle* h, const int* values, int n) {
  if (!h || !values || n < gwhip::Tuning::count) return GWHIP_ERR_ARG;
  gwhip::Tuning t;
  if (!t.from_array(values)) return GWHIP_ERR_ARG;
  h->k->set_tuning(t);
  return GWHIP_OK;
}

int gwhip_batch_create(int nsys, const gwhip_dims* dims, int device, gwhip_batch** out) {
  if (!dims || !out || nsys <= 0) {
    g_create_error = "null argument or empty batch";
    return GWHIP_ERR_ARG;
  }
  *out = nullptr;
  auto b = std::make_unique<gwhip_batch>();
  std::vector<gwhip::Dims> d;
  for (int i = 0; i < nsys; ++i) d.push_back(to_dims(&dims[i]));
  int st = guarded_b(b.get(), [&] { b->b = std::make_unique<gwhip::gw_hip_batch>(nsys, d.data(), device); });
  if (st != GWHIP_OK) {
    g_create_error = b->err;
    return st;
  }
  for (int i = 0; i < nsys; ++i) {
    auto h = std::make_unique<gwhip_handle>();
    h->k = &b->b->system(i);
    h->borrowed = true;
    b->systems.push_back(std::move(h));
  }
  *out = b.release();
  return GWHIP_OK;
}

void gwhip_batch_destroy(gwhip_batch* b) { delete b; }

const char* gwhip_batch_last_error(const gwhip_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

gwhip_handle* gwhip_batch_system(gwhip_batch* b, int i) {
  if (!b || i < 0 || i >= (int)b->systems.size()) return nullptr;
  return b->systems[i].get();
}

int gwhip_batch_autotune(gwhip_batch* b, const void* const* G_dev, int precision) {
  if (!b) return GWHIP_ERR_ARG;
  return guarded_b(b, [&] { b->b->autotune((const double2* const*)G_dev, precision); });
}

int gwhip_batch_eval_device(gwhip_batch* b, const void* const* G_dev, void* const* Sigma_dev, int precision) {
  if (!b) return GWHIP_ERR_ARG;
  return guarded_b(b, [&] { b->b->eval((const double2* const*)G_dev, (double2* const*)Sigma_dev, precision); });
}

int gwhip_batch_stats(const gwhip_batch* b, double* last_ms, int64_t* graph_nodes, int64_t* captures,
                      size_t* workspace_bytes) {
  if (!b) return GWHIP_ERR_ARG;
  if (last_ms) *last_ms = b->b->last_ms();
  if (graph_nodes) *graph_nodes = b->b->last_nodes();
  if (captures) *captures = b->b->captures();
  if (workspace_bytes) *workspace_bytes = b->b->workspace_bytes();
  return GWHIP_OK;
}

int gwhip_set_vjp_store(gwhip_handle* h, int keep) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_vjp_store(keep != 0); });
}

int gwhip_vjp(gwhip_handle* h, const double* G, const double* Sigma_bar, double* G_bar, int mode) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->vjp_host((const double2*)G, (const double2*)Sigma_bar, (double2*)G_bar, mode); });
}

int gwhip_vjp_device(gwhip_handle* h, const void* G, const void* Sigma_bar, void* G_bar, int mode) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->vjp_device((const double2*)G, (const double2*)Sigma_bar, (double2*)G_bar, mode); });
}

double gwhip_vjp_ms(const gwhip_handle* h) { return h ? h->k->vjp_ms() : 0.0; }

int gwhip_hf(gwhip_handle* h, const double* dm, const double* S, double madelung, double* F) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->hf_host((const double2*)dm, (const double2*)S, madelung, (double2*)F); });
}

int gwhip_hf_device(gwhip_handle* h, const void* dm, const void* S, double madelung, void* F) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->hf_device((const double2*)dm, (const double2*)S, madelung, (double2*)F); });
}

double gwhip_hf_ms(const gwhip_handle* h) { return h ? h->k->hf_ms() : 0.0; }

int gwhip_set_k_symmetry(gwhip_handle* h, int64_t ink, const int64_t* bz2ibz, const int64_t* ibz2bz, const double* U,
                         const int64_t* tr_conj) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_k_symmetry(ink, bz2ibz, ibz2bz, (const double2*)U, tr_conj); });
}

int gwhip_set_q_symmetry(gwhip_handle* h, int64_t inq, const int64_t* bz2ibz, const int64_t* ibz2bz, const double* U,
                         const int64_t* tr_conj) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_q_symmetry(inq, bz2ibz, ibz2bz, (const double2*)U, tr_conj); });
}

int gwhip_eval_sigma_ibz(gwhip_handle* h, const double* G_ibz, double* Sigma_ibz, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->solve_ibz_host((const double2*)G_ibz, (double2*)Sigma_ibz, precision); });
}

int gwhip_eval_sigma_ibz_device(gwhip_handle* h, const void* G_ibz, void* Sigma_ibz, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->solve_ibz_device((const double2*)G_ibz, (double2*)Sigma_ibz, precision); });
}

int gwhip_hf_ibz(gwhip_handle* h, const double* dm_ibz, const double* S_ibz, double madelung, double* F_ibz) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->hf_ibz_host((const double2*)dm_ibz, (const double2*)S_ibz, madelung, (double2*)F_ibz); });
}

int gwhip_hf_ibz_device(gwhip_handle* h, const void* dm_ibz, const void* S_ibz, double madelung, void* F_ibz) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h,
                 [&] { h->k->hf_ibz_device((const double2*)dm_ibz, (const double2*)S_ibz, madelung, (double2*)F_ibz); });
}

int gwhip_set_q0_extrapolation(gwhip_handle* h, int64_t inq, const int64_t* q_ir, const double* AqQ,
                               const double* q_abs, double madelung, int64_t nk_S, const double* S) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] {
    h->k->set_q0_extrapolation(inq, q_ir, (const double2*)AqQ, q_abs, madelung, nk_S, (const double2*)S);
  });
}

int gwhip_get_q0_eps(const gwhip_handle* h, double* eps_tau) {
  if (!h || !eps_tau) return GWHIP_ERR_ARG;
  const auto& e = h->k->q0_eps_tau();
  if (e.empty()) return GWHIP_ERR_STATE;
  std::memcpy(eps_tau, e.data(), e.size() * sizeof(double2));
  return GWHIP_OK;
}

int gwhip_dyson_create(const gwhip_dyson_dims* dims, int device, gwhip_dyson** out) {
  if (!dims || !out) {
    g_create_error = "null argument";
    return GWHIP_ERR_ARG;
  }
  *out = nullptr;
  auto d = std::make_unique<gwhip_dyson>();
  gwhip::DysonDims dd;
  dd.nts = dims->nts; dd.nw = dims->nw; dd.ns = dims->ns; dd.nk = dims->nk; dd.ink = dims->ink;
  dd.nao = dims->nao; dd.nso = dims->nso;
  const int st = guarded_d(d.get(), [&] { d->d = std::make_unique<gwhip::gw_dyson>(dd, device); });
  if (st != GWHIP_OK) {
    g_create_error = d->err;
    return st;
  }
  *out = d.release();
  return GWHIP_OK;
}

void gwhip_dyson_destroy(gwhip_dyson* d) { delete d; }

const char* gwhip_dyson_last_error(const gwhip_dyson* d) { return d ? d->err.c_str() : g_create_error.c_str(); }

int gwhip_dyson_set_grid(gwhip_dyson* d, const double* T_ft, const double* T_tf, const double* wn) {
  if (!d) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] { d->d->set_grid((const double2*)T_ft, (const double2*)T_tf, wn); });
}

int gwhip_dyson_set_system(gwhip_dyson* d, const double* H, const double* S, const double* weights) {
  if (!d) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] { d->d->set_system((const double2*)H, (const double2*)S, weights); });
}

int gwhip_dyson_compute_G(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu, double* G_tau) {
  if (!d || !Sigma1 || !Sigma_tau || !G_tau) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] { d->d->compute_G_host((double2*)Sigma1, (double2*)Sigma_tau, mu, (double2*)G_tau); });
}

int gwhip_dyson_compute_G_device(gwhip_dyson* d, void* Sigma1, void* Sigma_tau, double mu, void* G_tau) {
  if (!d) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] { d->d->compute_G_device((double2*)Sigma1, (double2*)Sigma_tau, mu, (double2*)G_tau); });
}

int gwhip_dyson_find_mu(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu0, double nel, double tol,
                        double* mu, double* nel_found) {
  if (!d || !mu || !nel_found) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] {
    const auto r = d->d->find_mu_host((double2*)Sigma1, (double2*)Sigma_tau, mu0, nel, tol);
    *nel_found = r.first;
    *mu = r.second;
  });
}

int gwhip_dyson_find_mu_device(gwhip_dyson* d, void* Sigma1, void* Sigma_tau, double mu0, double nel, double tol,
                               double* mu, double* nel_found) {
  if (!d || !mu || !nel_found) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] {
    const auto r = d->d->find_mu_device((double2*)Sigma1, (double2*)Sigma_tau, mu0, nel, tol);
    *nel_found = r.first;
    *mu = r.second;
  });
}

int gwhip_dyson_number_of_electrons(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu, double* nel) {
  if (!d || !nel) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] { *nel = d->d->number_of_electrons_host((double2*)Sigma1, (double2*)Sigma_tau, mu); });
}

int gwhip_dyson_energy(gwhip_dyson* d, const double* G_tau, const double* Sigma1, const double* Sigma_tau,
                       double* energies) {
  if (!d || !energies) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] {
    const auto e = d->d->energy_host((const double2*)G_tau, (const double2*)Sigma1, (const double2*)Sigma_tau);
    for (int i = 0; i < 3; ++i) energies[i] = e[i];
  });
}

int gwhip_dyson_energy_device(gwhip_dyson* d, const void* G_tau, const void* Sigma1, const void* Sigma_tau,
                              double* energies) {
  if (!d || !energies) return GWHIP_ERR_ARG;
  return guarded_d(d, [&] {
    const auto e = d->d->energy_device((const double2*)G_tau, (const double2*)Sigma1, (const double2*)Sigma_tau);
    for (int i = 0; i < 3; ++i) energies[i] = e[i];
  });
}

int gwhip_get_dims(const gwhip_handle* h, gwhip_dims* out) {
  if (!h || !out) return GWHIP_ERR_ARG;
  const gwhip::Dims& d = h->k->dims();
  out->nk = d.nk;
  for (int i = 0; i < 3; ++i) out->mesh[i] = d.mesh[i];
  out->nao = d.nao; out->nso = d.nso; out->ns = d.ns; out->NQ = d.NQ;
  out->nts = d.nts; out->nw_b = d.nw_b; out->beta = d.beta;
  return GWHIP_OK;
}

int gwhip_hf_exchange_stats(const gwhip_handle* h, double* ms, int64_t* launches) {
  if (!h) return GWHIP_ERR_ARG;
  if (ms) *ms = h->k->hf_exchange_ms();
  if (launches) *launches = h->k->hf_exchange_launches();
  return GWHIP_OK;
}

int gwhip_get_P_tilde(gwhip_handle* h, double* P) {
  if (!h || !P) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->copy_P_tilde((double2*)P); });
}

int gwhip_timings(const gwhip_handle* h, double* ms, int n) {
  if (!h || !ms) return 0;
  const int m = std::min(n, (int)gwhip::gw_hip_kernel::R_COUNT);
  for (int i = 0; i < m; ++i) ms[i] = h->k->region_ms()[i];
  return m;
}

int gwhip_kernel_stats(const gwhip_handle* h, int which, int64_t* launches, double* total_ms) {
  if (!h || !launches || !total_ms || which < 0 || which >= gwhip::gw_hip_kernel::K_COUNT) return GWHIP_ERR_ARG;
  *launches = h->k->kstat(which).launches;
  *total_ms = h->k->kstat(which).ms;
  return GWHIP_OK;
}

size_t gwhip_device_bytes(const gwhip_handle* h) { return h ? h->k->device_bytes() : 0; }

int gwhip_debug_zgemm(int M, int N, int K, const double* A, const double* B, double* C, int conjA, int conjB,
                      int shape) {
  gwhip_handle tmp;
  return guarded(&tmp, [&] {
    gwhip::debug_zgemm(M, N, K, (const double2*)A, (const double2*)B, (double2*)C, conjA, conjB, shape);
  });
}

int gwhip_debug_zgemm_split(int M, int N, int K, const double* A, const double* B, double* C, int slices, int tile) {
  gwhip_handle tmp;
  return guarded(&tmp, [&] {
    if (tile != 2 && tile != 4) throw gwhip::status_error(GWHIP_ERR_ARG, "int8 GEMM tile must be 2 or 4");
    gwhip::debug_zgemm_split(M, N, K, (const double2*)A, (const double2*)B, (double2*)C, slices, tile);
  });
}

int gwhip_debug_eval_q(gwhip_handle* h, const double* G, int64_t q, int precision, double* P0, double* P_tilde,
                       double* Sigma_q) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] {
    h->k->eval_one_q_host((const double2*)G, q, precision, (double2*)P0, (double2*)P_tilde, (double2*)Sigma_q);
  });
}

int gwhip_set_leakage_capture(gwhip_handle* h, int on) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->set_leakage_capture(on != 0); });
}

int gwhip_get_leakage_inputs(gwhip_handle* h, double* P0, double* P_tilde) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->copy_leakage_inputs((double2*)P0, (double2*)P_tilde); });
}

int gwhip_debug_stage(gwhip_handle* h, const double* G, int last_stage) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->run_stages_host((const double2*)G, last_stage); });
}

}  // extern "C"

namespace gwhip {
// Used by the no-allocation entry point (gwhip_noalloc.hip), which owns the device workspace.
int eval_sigma_in_workspace(gwhip_handle* h, const void* G, void* Sigma, int precision, void* ws, size_t ws_bytes) {
  if (!h) return GWHIP_ERR_ARG;
  return guarded(h, [&] { h->k->solve_device_in((const double2*)G, (double2*)Sigma, precision, ws, ws_bytes); });
}
int workspace_bytes_needed(const gwhip_handle* h, int precision, size_t* bytes) {
  if (!h || !bytes || !precision_valid(precision)) return GWHIP_ERR_ARG;
  *bytes = h->k->min_workspace_bytes(precision);
  return GWHIP_OK;
}
void set_handle_error(gwhip_handle* h, const char* msg) {
  if (h) h->err = msg;
}
int handle_device(const gwhip_handle* h) { return h ? h->k->device() : -1; }
}  // namespace gwhip
