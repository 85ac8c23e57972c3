// Irreducible-wedge (IBZ) inputs and outputs on the device (included by gw_hip_kernel.hip after
// gw_hf_host.hpp; one translation unit).
//
// The reference keeps G, Sigma, the density matrix and the Fock matrix on the ink irreducible k
// points and carries an irreducible-point object to any k with green-symmetry's value_AO
// (src/gw_cpu_kernel.cpp:177-178,336, src/hf_cpu_kernels.cpp:107):
//   X(k) = X(ik)                    at a representative (k = ibz2bz[ik])
//   X(k) = U_k op(X(ik)) U_k^H      elsewhere, op = complex conjugation when the symmetry
//                                   includes time reversal (tr_conj[k] = 1), else identity
// with ik the IBZ ordinal of bz2ibz[k] and U_k = k_sym_transform_ao[k] (nso x nso).
//
// Two evaluations of Sigma on the irreducible points:
// * with the q symmetry set (set_q_symmetry), the reference's own star-of-q scheme
//   (src/gw_cpu_kernel.cpp:44-47,74-134,281-354): P0 and P-tilde for the inq irreducible q only,
//   with G carried to the full mesh; Sigma for the ink irreducible k only, summed over every q of
//   each star with P-tilde carried to that q in the auxiliary basis,
//     P(q) = U_q P-tilde(q_ir) U_q^H, conjugated when q_tr_conj[q] = 1 (eval_p0_bz_from_ibz);
//   the work is inq/nq of the full-mesh P0 and Dyson stages and ink/nk of its Sigma stage;
// * without it, the irreducible input is unfolded to the full mesh, the full-mesh kernel runs,
//   and the representatives' rows are gathered back: Sigma(ik) = Sigma_full(ibz2bz[ik]).  This
//   equals the star-of-q result up to how exactly the system obeys its symmetry.
#pragma once

namespace gwhip {

// T[b][k] = op(X[b][ord[k]]) U_k^H (a representative: T = X); one thread per element.
__global__ void unfold_right_kernel(const double2* __restrict__ X, double2* __restrict__ T, const int* __restrict__ ord,
                                    const int* __restrict__ rep, const int* __restrict__ tr,
                                    const double2* __restrict__ U, int nk, int ink, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int k = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int i = e / n, j = e - (e / n) * n;
  const double2* x = X + ((b * ink + ord[k]) * n) * n;
  double2* t = T + ((b * nk + k) * n) * n;
  if (rep[k]) {
    t[e] = x[e];
    return;
  }
  const double s = tr[k] ? -1.0 : 1.0;
  const double2* u = U + ((int64_t)k * n + j) * n;  // row j of U_k
  double re = 0.0, im = 0.0;
  for (int m = 0; m < n; ++m) {
    const double2 xv = x[(int64_t)i * n + m], uv = u[m];
    const double xr = xv.x, xi = s * xv.y;
    // xv * conj(uv)
    re += xr * uv.x + xi * uv.y;
    im += xi * uv.x - xr * uv.y;
  }
  t[e] = make_double2(re, im);
}

// Y[b][k] = U_k T[b][k] (a representative: Y = T); one thread per element.
__global__ void unfold_left_kernel(const double2* __restrict__ T, double2* __restrict__ Y, const int* __restrict__ rep,
                                   const double2* __restrict__ U, int nk, int n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * n) return;
  const int k = blockIdx.y;
  const int64_t b = blockIdx.z;
  const int i = e / n, j = e - (e / n) * n;
  const double2* t = T + ((b * nk + k) * n) * n;
  double2* y = Y + ((b * nk + k) * n) * n;
  if (rep[k]) {
    y[e] = t[e];
    return;
  }
  const double2* u = U + ((int64_t)k * n + i) * n;  // row i of U_k
  double re = 0.0, im = 0.0;
  for (int m = 0; m < n; ++m) {
    const double2 uv = u[m], tv = t[(int64_t)m * n + j];
    re += uv.x * tv.x - uv.y * tv.y;
    im += uv.x * tv.y + uv.y * tv.x;
  }
  y[e] = make_double2(re, im);
}

// Dst[b][ik] = Src[b][ibz2bz[ik]]; one thread per element of one (b, ik) matrix.
__global__ void gather_ibz_kernel(const double2* __restrict__ Src, double2* __restrict__ Dst,
                                  const int* __restrict__ ibz2bz, int nk, int ink, int n2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n2) return;
  const int ik = blockIdx.y;
  const int64_t b = blockIdx.z;
  Dst[(b * ink + ik) * n2 + e] = Src[(b * nk + ibz2bz[ik]) * n2 + e];
}

// Dst[r][e] = Src[rows[r]][e] for rows of len elements; grid.y = row.
template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ src, T* __restrict__ dst, const int* __restrict__ rows,
                                   int64_t len) {
  const int64_t r = blockIdx.y;
  const T* s = src + (int64_t)rows[r] * len;
  T* d = dst + r * len;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x)
    d[e] = s[e];
}

// Y = conj(X), n elements.
__global__ void conj_copy_kernel(const double2* __restrict__ X, double2* __restrict__ Y, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const double2 x = X[e];
    Y[e] = make_double2(x.x, -x.y);
  }
}

template <typename T>
void launch_gather_rows(const T* src, T* dst, const int* rows, int64_t len, int64_t nrows, hipStream_t st) {
  if (nrows < 1 || nrows > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "gather: row count out of range");
  const dim3 grid((unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(len, 256))), (unsigned)nrows);
  hipLaunchKernelGGL(gather_rows_kernel<T>, grid, dim3(256), 0, st, src, dst, rows, len);
  GWHIP_CHECK(hipGetLastError());
}

void gw_hip_kernel::set_k_symmetry(int64_t ink, const int64_t* bz2ibz, const int64_t* ibz2bz, const double2* U,
                                   const int64_t* tr_conj) {
  const int64_t nk = d_.nk, n = d_.nso;
  if (!bz2ibz || !ibz2bz || !U || !tr_conj) throw status_error(GWHIP_ERR_ARG, "null symmetry array");
  if (ink < 1 || ink > nk) throw status_error(GWHIP_ERR_ARG, "ink must lie in [1, nk]");
  std::vector<int> ord(nk, -1), rep(nk, 0), tr(nk), i2b(ink);
  for (int64_t i = 0; i < ink; ++i) {
    const int64_t r = ibz2bz[i];
    if (r < 0 || r >= nk) throw status_error(GWHIP_ERR_ARG, "ibz2bz entry out of range");
    if (bz2ibz[r] != r) throw status_error(GWHIP_ERR_ARG, "a representative must map to itself in bz2ibz");
    if (rep[r]) throw status_error(GWHIP_ERR_ARG, "repeated representative in ibz2bz");
    rep[r] = 1;
    i2b[i] = (int)r;
  }
  for (int64_t k = 0; k < nk; ++k) {
    const int64_t r = bz2ibz[k];
    for (int64_t i = 0; i < ink && ord[k] < 0; ++i)
      if (ibz2bz[i] == r) ord[k] = (int)i;
    if (ord[k] < 0) throw status_error(GWHIP_ERR_ARG, "bz2ibz entry is not a representative");
    if (tr_conj[k] != 0 && tr_conj[k] != 1) throw status_error(GWHIP_ERR_ARG, "tr_conj entries must be 0 or 1");
    tr[k] = (int)tr_conj[k];
  }
  GWHIP_CHECK(hipSetDevice(device_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  if (!ksym_idx_) ksym_idx_ = dalloc<int>((size_t)4 * nk);
  if (!ksym_U_) ksym_U_ = dalloc<double2>((size_t)nk * n * n);
  std::vector<int> all;
  all.insert(all.end(), ord.begin(), ord.end());
  all.insert(all.end(), rep.begin(), rep.end());
  all.insert(all.end(), tr.begin(), tr.end());
  all.insert(all.end(), i2b.begin(), i2b.end());
  all.resize((size_t)4 * nk, 0);
  GWHIP_CHECK(hipMemcpy(ksym_idx_, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(ksym_U_, U, (size_t)nk * n * n * sizeof(double2), hipMemcpyHostToDevice));
  ink_ = ink;
  ksym_i2b_ = i2b;
  star_dirty_ = true;
}

void gw_hip_kernel::set_q_symmetry(int64_t inq, const int64_t* bz2ibz, const int64_t* ibz2bz, const double2* U,
                                   const int64_t* tr_conj) {
  const int64_t nq = d_.nk, NQ = d_.NQ, NQ2 = NQ * NQ;
  GWHIP_CHECK(hipSetDevice(device_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  if (inq == 0) {
    inq_ = 0;
    return;
  }
  if (!bz2ibz || !ibz2bz || !U || !tr_conj) throw status_error(GWHIP_ERR_ARG, "null q symmetry array");
  if (inq < 1 || inq > nq) throw status_error(GWHIP_ERR_ARG, "inq must lie in [0, nq]");
  std::vector<int> i2b(inq), tr(nq), ident(nq, 1), isrep(nq, 0);
  for (int64_t i = 0; i < inq; ++i) {
    const int64_t r = ibz2bz[i];
    if (r < 0 || r >= nq) throw status_error(GWHIP_ERR_ARG, "q ibz2bz entry out of range");
    if (bz2ibz[r] != r) throw status_error(GWHIP_ERR_ARG, "an irreducible q must map to itself in bz2ibz");
    if (isrep[r]) throw status_error(GWHIP_ERR_ARG, "repeated irreducible q in ibz2bz");
    isrep[r] = 1;
    i2b[i] = (int)r;
  }
  std::vector<std::vector<int>> stars(inq);
  for (int64_t q = 0; q < nq; ++q) {  // stars in ascending q, as input.h5:/symmetry/q/stars
    int64_t iq = -1;
    for (int64_t i = 0; i < inq && iq < 0; ++i)
      if (bz2ibz[q] == ibz2bz[i]) iq = i;
    if (iq < 0) throw status_error(GWHIP_ERR_ARG, "q bz2ibz entry is not an irreducible q");
    stars[iq].push_back((int)q);
    if (tr_conj[q] != 0 && tr_conj[q] != 1) throw status_error(GWHIP_ERR_ARG, "q tr_conj entries must be 0 or 1");
    tr[q] = (int)tr_conj[q];
    const double2* u = U + q * NQ2;
    for (int64_t a = 0; a < NQ && ident[q]; ++a)
      for (int64_t b = 0; b < NQ; ++b) {
        const double2 v = u[a * NQ + b];
        if (v.x != (a == b ? 1.0 : 0.0) || v.y != 0.0) {
          ident[q] = 0;
          break;
        }
      }
  }
  if (!qsym_U_) qsym_U_ = dalloc<double2>((size_t)nq * NQ2);
  GWHIP_CHECK(hipMemcpy(qsym_U_, U, (size_t)nq * NQ2 * sizeof(double2), hipMemcpyHostToDevice));
  qsym_i2b_ = i2b;
  qsym_tr_ = tr;
  qsym_ident_ = ident;
  qstar_ = stars;
  inq_ = inq;
  star_dirty_ = true;
}

// Device tables of the star-of-q evaluation (rebuilt after any change of the k or q symmetry, the
// k +- q maps or the q -> 0 data) and its buffers.
void gw_hip_kernel::star_tables() {
  const int64_t nq = d_.nk, nk = d_.nk, ink = ink_, NQ = d_.NQ, NQ2 = NQ * NQ, nao2 = d_.nao * d_.nao;
  if (star_dirty_ || star_ink_ != ink) {
    if (star_ink_ != ink) {
      for (void* p : {(void*)star_km_, (void*)star_V_})
        if (p) (void)hipFree(p);
      bytes_ -= (size_t)star_ink_ * (nq * sizeof(int) + NQ * nao2 * sizeof(double2));
      star_km_ = nullptr;
      star_V_ = nullptr;
    }
    if (!star_km_) star_km_ = dalloc<int>((size_t)nq * ink);
    if (!star_V_) star_V_ = dalloc<double2>((size_t)ink * NQ * nao2);
    std::vector<int> km((size_t)nq * ink);
    for (int64_t q = 0; q < nq; ++q)
      for (int64_t p = 0; p < ink; ++p) km[q * ink + p] = kminus_host_[q * nk + ksym_i2b_[p]];
    GWHIP_CHECK(hipMemcpy(star_km_, km.data(), km.size() * sizeof(int), hipMemcpyHostToDevice));
    std::vector<int> rows(inq_, -1);
    if (q0_on_) {  // AqQ row of each irreducible q; every row must be produced exactly once
      std::vector<int> seen(q0_inq_, 0);
      for (int64_t i = 0; i < inq_; ++i) {
        rows[i] = q0_qiq_host_[qsym_i2b_[i]];
        if (rows[i] >= 0) ++seen[rows[i]];
      }
      for (int64_t r = 0; r < q0_inq_; ++r)
        if (seen[r] != 1)
          throw status_error(GWHIP_ERR_ARG, "the q -> 0 data's irreducible q differ from the q symmetry's");
    }
    if (star_q0row_) (void)hipFree(star_q0row_), bytes_ -= (size_t)nq * sizeof(int), star_q0row_ = nullptr;
    star_q0row_ = dalloc<int>((size_t)nq);
    GWHIP_CHECK(hipMemcpy(star_q0row_, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice));
    star_ink_ = ink;
    star_dirty_ = false;
  }
  if (!star_P_) star_P_ = dalloc<double2>((size_t)2 * (d_.nts / 2) * NQ2);
}

// P-tilde of irreducible q row iq carried to q of its star, half tau grid: P_ row iq itself when
// U_q = I without time reversal, else star_P_ (eval_p0_bz_from_ibz, src/gw_cpu_kernel.cpp:281-293).
const double2* gw_hip_kernel::star_rotate(int64_t iq, int64_t q) {
  const int64_t NQ = d_.NQ, NQ2 = NQ * NQ, half = d_.nts / 2;
  const double2* Pi = P_ + iq * d_.nts * NQ2;
  if (qsym_ident_[q] && !qsym_tr_[q]) return Pi;
  double2* out = star_P_;
  if (qsym_ident_[q]) {  // conj(P) alone
    hipLaunchKernelGGL(conj_copy_kernel, dim3(1024), dim3(256), 0, stream_, Pi, out, half * NQ2);
    GWHIP_CHECK(hipGetLastError());
    return out;
  }
  // conj(U P U^H) = conj(U) conj(P) U^T: T = op(U) op(P), out = T op'(U)^T
  const int tr = qsym_tr_[q];
  double2* T = star_P_ + half * NQ2;
  const double2* Uq = qsym_U_ + q * NQ2;
  ZGemmArgs a{};
  a.M = (int)NQ; a.N = (int)NQ; a.K = (int)NQ; a.batch = (int)half; a.kin = (int)NQ; a.nin = (int)NQ;
  a.A = Uq; a.sAb = 0; a.sAm = NQ; a.sAko = 0; a.sAki = 1;
  a.B = Pi; a.sBb = NQ2; a.sBko = 0; a.sBki = NQ; a.sBno = 0; a.sBni = 1;
  a.C = T; a.sCb = NQ2; a.sCm = NQ; a.sCno = 0; a.sCni = 1;
  a.alpha = make_double2(1.0, 0.0);
  a.conjA = tr; a.conjB = tr;
  zgemm_f64(a, true, false, stream_);
  ZGemmArgs b{};
  b.M = (int)NQ; b.N = (int)NQ; b.K = (int)NQ; b.batch = (int)half; b.kin = (int)NQ; b.nin = (int)NQ;
  b.A = T; b.sAb = NQ2; b.sAm = NQ; b.sAko = 0; b.sAki = 1;
  b.B = Uq; b.sBb = 0; b.sBko = 0; b.sBki = 1; b.sBno = 0; b.sBni = NQ;  // B(k, n) = U[n][k]
  b.C = out; b.sCb = NQ2; b.sCm = NQ; b.sCno = 0; b.sCni = 1;
  b.alpha = make_double2(1.0, 0.0);
  b.conjB = tr ? 0 : 1;
  zgemm_f64(b, true, true, stream_);
  return out;
}

// gw_cpu_kernel::solve on the irreducible points (src/gw_cpu_kernel.cpp:31-134): G_ibz, Sigma_ibz
// [nts][ns][ink][nso^2] on the device.
void gw_hip_kernel::solve_star_device(const double2* G_ibz, double2* Sigma_ibz, int precision) {
  const int64_t nk = d_.nk, ink = ink_, NQ = d_.NQ, nao2 = d_.nao * d_.nao;
  const bool split = sigma_split(precision);
  if (any_split(precision)) ensure_slices();
  ensure_orb_buffers();
  star_tables();
  if (split) {
    const int64_t vs = ink * nao2 * OZ_PLANES * split_ns_ * kc_s_;
    if (vs != star_vs_bytes_) {
      if (star_Vs_) (void)hipFree(star_Vs_), bytes_ -= (size_t)star_vs_bytes_, star_Vs_ = nullptr;
      if (star_eVs_) (void)hipFree(star_eVs_), bytes_ -= (size_t)star_evs_elems_ * sizeof(int), star_eVs_ = nullptr;
      star_Vs_ = dalloc<int8_t>((size_t)vs);
      star_eVs_ = dalloc<int>((size_t)ink * nao2);
      star_vs_bytes_ = vs;
      star_evs_elems_ = ink * nao2;
    }
  }
  const size_t nfull = (size_t)d_.nts * d_.ns * nk * d_.nso * d_.nso;
  const int64_t nsig = d_.nts * d_.ns * ink * d_.nso * d_.nso;
  if (!G_dev_) G_dev_ = dalloc<double2>(nfull);
  if (!S_dev_) S_dev_ = dalloc<double2>(nfull);
  unfold(G_ibz, S_dev_, G_dev_, d_.nts * d_.ns);
  timing_ = true;
  resolve_timings();
  detect_hermitian_G(G_dev_);
  for (double& r : region_ms_) r = 0.0;
  for (auto& k : kstat_) k = KernelStat{};

  Event t0, t1, t2, t3;
  GWHIP_CHECK(hipEventRecord(t0, stream_));
  zero(Sigma_ibz, nsig);
  stored_valid_ = false;
  leak_valid_ = false;
  stage_p0(G_dev_, p0_split(precision), &qsym_i2b_);
  capture_leakage_p0();
  GWHIP_CHECK(hipEventRecord(t1, stream_));
  q0_eps_ = q0_on_;
  stage_dyson(inq_, star_q0row_);
  q0_eps_ = false;
  GWHIP_CHECK(hipEventRecord(t2, stream_));
  const int* rows = ksym_idx_ + 3 * nk;  // ibz2bz on the device
  for (int64_t iq = 0; iq < inq_; ++iq)
    for (int q : qstar_[iq]) {
      const double2* Pq = star_rotate(iq, q);
      launch_gather_rows(Vq_neg(q), star_V_, rows, NQ * nao2, ink, stream_);
      if (split) {
        const int64_t len = nao2 * OZ_PLANES * split_ns_ * kc_s_;  // a multiple of 64 bytes (kc_s is of 64)
        launch_gather_rows(reinterpret_cast<const uint4*>(Vs_neg(q)), reinterpret_cast<uint4*>(star_Vs_), rows,
                           len / 16, ink, stream_);
        launch_gather_rows(eVs_neg(q), star_eVs_, rows, nao2, ink, stream_);
      }
      sigma_q(G_dev_, Sigma_ibz, split, Pq, star_V_, star_Vs_, star_eVs_, star_km_ + (int64_t)q * ink, ink);
    }
  finish_sigma(Sigma_ibz, ink);
  if (q0_on_) apply_q0(Sigma_ibz, G_ibz, ink);  // on the irreducible points, as the reference does
  GWHIP_CHECK(hipEventRecord(t3, stream_));
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(64), 0, stream_, flag_, (int64_t)1);
  hipLaunchKernelGGL(nonfinite_kernel, dim3(1024), dim3(256), 0, stream_, Sigma_ibz, nsig, flag_);
  GWHIP_CHECK(hipGetLastError());
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  resolve_timings();
  float x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t0, t1)); region_ms_[R_P0] = x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t2, t3)); region_ms_[R_SIGMA] = x;
  GWHIP_CHECK(hipEventElapsedTime(&x, t0, t3)); region_ms_[R_TOTAL] = x;
  int flag = 0;  // non-finite first: a NaN or Inf in G also fails the Dyson step's pivot test
  GWHIP_CHECK(hipMemcpy(&flag, flag_, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) throw status_error(GWHIP_ERR_NONFINITE, "Sigma contains non-finite values");
  check_dyson_flag();
  leak_valid_ = leak_on_;
}

// X [nb][ink][n][n] -> Y [nb][nk][n][n] through the scratch T (same size as Y).
void gw_hip_kernel::unfold(const double2* X, double2* T, double2* Y, int64_t nb) {
  const int64_t nk = d_.nk, n = d_.nso;
  const int* ord = ksym_idx_;
  const int* rep = ksym_idx_ + nk;
  const int* tr = ksym_idx_ + 2 * nk;
  if (nb > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "more than 65535 (tau, spin) slices");
  const dim3 grid((unsigned)ceil_div(n * n, 256), (unsigned)nk, (unsigned)nb);
  hipLaunchKernelGGL(unfold_right_kernel, grid, dim3(256), 0, stream_, X, T, ord, rep, tr, ksym_U_, (int)nk, (int)ink_,
                     (int)n);
  GWHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(unfold_left_kernel, grid, dim3(256), 0, stream_, T, Y, rep, ksym_U_, (int)nk, (int)n);
  GWHIP_CHECK(hipGetLastError());
}

void gw_hip_kernel::gather_ibz(const double2* Full, double2* Ibz, int64_t nb) {
  const int64_t nk = d_.nk, n = d_.nso;
  if (nb > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "more than 65535 (tau, spin) slices");
  const dim3 grid((unsigned)ceil_div(n * n, 256), (unsigned)ink_, (unsigned)nb);
  hipLaunchKernelGGL(gather_ibz_kernel, grid, dim3(256), 0, stream_, Full, Ibz, ksym_idx_ + 3 * nk, (int)nk, (int)ink_,
                     (int)(n * n));
  GWHIP_CHECK(hipGetLastError());
}

void gw_hip_kernel::solve_ibz_device(const double2* G_ibz, double2* Sigma_ibz, int precision) {
  if (ink_ <= 0) throw status_error(GWHIP_ERR_STATE, "k symmetry not set (gwhip_set_k_symmetry)");
  if (!G_ibz || !Sigma_ibz) throw status_error(GWHIP_ERR_ARG, "null G or Sigma");
  check_ready(precision);
  GWHIP_CHECK(hipSetDevice(device_));
  if (inq_ > 0) {
    solve_star_device(G_ibz, Sigma_ibz, precision);
    return;
  }
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  if (!G_dev_) G_dev_ = dalloc<double2>(nsig);
  if (!S_dev_) S_dev_ = dalloc<double2>(nsig);
  unfold(G_ibz, S_dev_, G_dev_, d_.nts * d_.ns);  // Sigma's buffer is the scratch until the eval
  solve_device(G_dev_, S_dev_, precision, false);
  gather_ibz(S_dev_, Sigma_ibz, d_.nts * d_.ns);
  if (q0_on_) apply_q0(Sigma_ibz, G_ibz, ink_);  // on the irreducible points, as the reference does
  GWHIP_CHECK(hipStreamSynchronize(stream_));
}

void gw_hip_kernel::solve_ibz_host(const double2* G_ibz, double2* Sigma_ibz, int precision) {
  if (!G_ibz || !Sigma_ibz) throw status_error(GWHIP_ERR_ARG, "null G or Sigma");
  if (ink_ <= 0) throw status_error(GWHIP_ERR_STATE, "k symmetry not set (gwhip_set_k_symmetry)");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t n = (size_t)d_.nts * d_.ns * ink_ * d_.nso * d_.nso;
  double2* buf = dalloc<double2>(2 * n);
  try {
    GWHIP_CHECK(hipMemcpy(buf, G_ibz, n * sizeof(double2), hipMemcpyHostToDevice));
    solve_ibz_device(buf, buf + n, precision);
    GWHIP_CHECK(hipMemcpy(Sigma_ibz, buf + n, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(buf);
    bytes_ -= 2 * n * sizeof(double2);
    throw;
  }
  (void)hipFree(buf);
  bytes_ -= 2 * n * sizeof(double2);
}

void gw_hip_kernel::hf_ibz_device(const double2* dm_ibz, const double2* S_ibz, double madelung, double2* F_ibz) {
  if (ink_ <= 0) throw status_error(GWHIP_ERR_STATE, "k symmetry not set (gwhip_set_k_symmetry)");
  if (!dm_ibz || !S_ibz || !F_ibz) throw status_error(GWHIP_ERR_ARG, "null dm, S or F");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t nf = (size_t)d_.ns * d_.nk * d_.nso * d_.nso;
  if (!hf_sym_) hf_sym_ = dalloc<double2>(4 * nf);
  double2 *dm = hf_sym_, *S = hf_sym_ + nf, *F = hf_sym_ + 2 * nf, *T = hf_sym_ + 3 * nf;
  unfold(dm_ibz, T, dm, d_.ns);
  unfold(S_ibz, T, S, d_.ns);
  hf_device(dm, S, madelung, F);
  gather_ibz(F, F_ibz, d_.ns);
  GWHIP_CHECK(hipStreamSynchronize(stream_));
}

void gw_hip_kernel::hf_ibz_host(const double2* dm_ibz, const double2* S_ibz, double madelung, double2* F_ibz) {
  if (!dm_ibz || !S_ibz || !F_ibz) throw status_error(GWHIP_ERR_ARG, "null dm, S or F");
  if (ink_ <= 0) throw status_error(GWHIP_ERR_STATE, "k symmetry not set (gwhip_set_k_symmetry)");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t n = (size_t)d_.ns * ink_ * d_.nso * d_.nso;
  double2* buf = dalloc<double2>(3 * n);
  try {
    GWHIP_CHECK(hipMemcpy(buf, dm_ibz, n * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(buf + n, S_ibz, n * sizeof(double2), hipMemcpyHostToDevice));
    hf_ibz_device(buf, buf + n, madelung, buf + 2 * n);
    GWHIP_CHECK(hipMemcpy(F_ibz, buf + 2 * n, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(buf);
    bytes_ -= 3 * n * sizeof(double2);
    throw;
  }
  (void)hipFree(buf);
  bytes_ -= 3 * n * sizeof(double2);
}

}  // namespace gwhip
