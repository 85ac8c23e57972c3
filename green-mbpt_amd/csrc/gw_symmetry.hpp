// Irreducible-wedge (IBZ) inputs and outputs on the device (included by gw_hip_kernel.hip after
// gw_hf_host.hpp; one translation unit).
//
// The reference keeps G, Sigma, the density matrix and the Fock matrix on the ink irreducible k
// points and carries an irreducible-point object to any k with green-symmetry's value_AO
// (src/gw_cpu_kernel.cpp:177-178,336, src/hf_cpu_kernels.cpp:107):
//   X(k) = X(ik)                    at a representative (k = ibz2bz[ik])
//   X(k) = U_k op(X(ik)) U_k^H      elsewhere, op = complex conjugation when the symmetry
//                                   includes time reversal (tr_conj[k] = 1), else identity
// with ik the IBZ ordinal of bz2ibz[k] and U_k = k_sym_transform_ao[k] (nso x nso).  Here the
// irreducible input is unfolded to the full mesh on the device, the full-mesh kernel runs, and the
// representatives' rows are gathered back: Sigma(ik) = Sigma_full(ibz2bz[ik]).  Results equal the
// reference's star-of-q evaluation up to how exactly the system obeys its symmetry.
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
