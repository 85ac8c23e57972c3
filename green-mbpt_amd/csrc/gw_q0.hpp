// q -> 0 finite-size correction of Sigma (q0_treatment EXTRAPOLATE), included by gw_hip_kernel.hip
// after gw_symmetry.hpp; one translation unit.  Counterpart of mbpt_q0_utils_t
// (src/mbpt_q0_utils_t.cpp:11-178) and its call sites in gw_cpu_kernel (src/gw_cpu_kernel.cpp:59-64,
// 266-270):
//   during the Dyson stage, for every irreducible q (row iq of AqQ):
//     eps(n, iq) = conj(A_iq)^T P-tilde(i Omega_n, q) A_iq                      (q0_eps_kernel)
//   after Sigma:  per n, a degree-2 least-squares fit of Re eps(n, .) over the distinct |q| in
//     (0, 1] by the normal equations, eps_q0(n) = its constant term; eps_q0(tau) = T_tb eps_q0;
//     Sigma(t, s, k) -= madelung eps_q0(t) S(s, k) G(t, s, k) S(s, k)            (q0_gs/q0_apply)
//   X2C: the aa, bb and ab blocks with S_aa on both sides, ba -= (ab correction)^H.
// The fit is a 3 x 3 solve per frequency on the host; everything else stays on the device.  The
// eps kernel (q0_eps_kernel) sits with the Dyson-stage kernels in gw_kernels.hpp.
#pragma once

namespace gwhip {

// T[b] = G[b] S(s, k), b = (t, s, k) over [nts][ns][nkS]; X2C: every nao block of G times S_aa.
__global__ void q0_gs_kernel(const double2* __restrict__ G, const double2* __restrict__ S, double2* __restrict__ T,
                             int ns, int nkS, int nso, int nao, int x2c, int64_t nmat) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nso * nso) return;
  const int i = e / nso, j = e - (e / nso) * nso;
  for (int64_t b = blockIdx.y; b < nmat; b += gridDim.y) {
    const int k = (int)(b % nkS), s = (int)((b / nkS) % ns);
    const double2* g = G + b * nso * nso;
    const double2* sm = S + ((int64_t)s * nkS + k) * nso * nso;
    double re = 0.0, im = 0.0;
    if (!x2c) {
      for (int m = 0; m < nso; ++m) {
        const double2 x = g[(int64_t)i * nso + m], y = sm[(int64_t)m * nso + j];
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
      }
    } else {
      const int bj = j / nao, jj = j - bj * nao;
      for (int m = 0; m < nao; ++m) {
        const double2 x = g[(int64_t)i * nso + bj * nao + m], y = sm[(int64_t)m * nso + jj];
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
      }
    }
    T[b * nso * nso + e] = make_double2(re, im);
  }
}

// Sigma[b] -= madelung eps_t[t] S(s, k) T[b]  (X2C: S_aa, and ba from ab).
__global__ void q0_apply_kernel(double2* __restrict__ Sigma, const double2* __restrict__ T,
                                const double2* __restrict__ S, const double2* __restrict__ eps_t, double madelung,
                                int ns, int nkS, int nso, int nao, int x2c, int64_t nmat) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nso * nso) return;
  const int i = e / nso, j = e - (e / nso) * nso;
  for (int64_t b = blockIdx.y; b < nmat; b += gridDim.y) {
    const int k = (int)(b % nkS), s = (int)((b / nkS) % ns);
    const int t = (int)(b / ((int64_t)nkS * ns));
    const double2* tm = T + b * nso * nso;
    const double2* sm = S + ((int64_t)s * nkS + k) * nso * nso;
    const double2 ep = eps_t[t];
    const double cr = madelung * ep.x, ci = madelung * ep.y;
    double2* out = Sigma + b * nso * nso + e;
    double re = 0.0, im = 0.0;
    if (!x2c) {
      for (int m = 0; m < nso; ++m) {
        const double2 x = sm[(int64_t)i * nso + m], y = tm[(int64_t)m * nso + j];
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
      }
    } else {
      const int bi = i / nao, bj = j / nao, ii = i - bi * nao, jj = j - bj * nao;
      if (bi == 1 && bj == 0) {  // ba = (c S_aa G_ab S_aa)^H: element (ii, jj) = conj of ab's (jj, ii)
        double ar = 0.0, ai = 0.0;
        for (int m = 0; m < nao; ++m) {
          const double2 x = sm[(int64_t)jj * nso + m], y = tm[(int64_t)m * nso + nao + ii];
          ar += x.x * y.x - x.y * y.y;
          ai += x.x * y.y + x.y * y.x;
        }
        out->x -= cr * ar - ci * ai;
        out->y += cr * ai + ci * ar;  // minus the conjugate
        continue;
      }
      for (int m = 0; m < nao; ++m) {
        const double2 x = sm[(int64_t)ii * nso + m], y = tm[(int64_t)(bi * nao + m) * nso + j];
        re += x.x * y.x - x.y * y.y;
        im += x.x * y.y + x.y * y.x;
      }
    }
    out->x -= cr * re - ci * im;
    out->y -= cr * im + ci * re;
  }
}

void gw_hip_kernel::set_q0_extrapolation(int64_t inq, const int64_t* q_ir, const double2* AqQ, const double* q_abs,
                                         double madelung, int64_t nkS, const double2* S) {
  GWHIP_CHECK(hipSetDevice(device_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  star_dirty_ = true;
  if (inq == 0) {  // disable
    q0_on_ = false;
    return;
  }
  const int64_t nk = d_.nk, NQ = d_.NQ, nso = d_.nso;
  if (!q_ir || !AqQ || !q_abs || !S) throw status_error(GWHIP_ERR_ARG, "null q0 array");
  if (inq < 0 || inq > nk) throw status_error(GWHIP_ERR_ARG, "inq must lie in [0, nk]");
  if (nkS != nk && nkS != ink_) throw status_error(GWHIP_ERR_ARG, "S must cover nk or ink k points");
  std::vector<int> qiq(nk, -1);
  for (int64_t i = 0; i < inq; ++i) {
    if (q_ir[i] < 0 || q_ir[i] >= nk) throw status_error(GWHIP_ERR_ARG, "q_ir entry out of range");
    if (qiq[q_ir[i]] >= 0) throw status_error(GWHIP_ERR_ARG, "repeated q in q_ir");
    qiq[q_ir[i]] = (int)i;
  }
  q0_qabs_.assign(q_abs, q_abs + inq);
  // the fit needs three distinct |q| in (0, 1]
  std::vector<int> pick = q0_samples();
  if (pick.size() < 3) throw status_error(GWHIP_ERR_ARG, "fewer than 3 distinct |q| in (0, 1] for the q -> 0 fit");
  if (q0_qiq_) (void)hipFree(q0_qiq_), bytes_ -= (size_t)nk * sizeof(int), q0_qiq_ = nullptr;
  if (q0_A_) (void)hipFree(q0_A_), bytes_ -= (size_t)q0_inq_ * NQ * sizeof(double2), q0_A_ = nullptr;
  if (q0_eps_dev_) (void)hipFree(q0_eps_dev_), bytes_ -= (size_t)d_.nw_b * q0_inq_ * sizeof(double2), q0_eps_dev_ = nullptr;
  if (q0_S_) (void)hipFree(q0_S_), bytes_ -= (size_t)d_.ns * q0_nS_ * nso * nso * sizeof(double2), q0_S_ = nullptr;
  if (q0_tmp_) (void)hipFree(q0_tmp_), bytes_ -= (size_t)d_.nts * d_.ns * q0_nS_ * nso * nso * sizeof(double2), q0_tmp_ = nullptr;
  q0_inq_ = inq;
  q0_nS_ = nkS;
  q0_mad_ = madelung;
  q0_qiq_ = dalloc<int>((size_t)nk);
  q0_A_ = dalloc<double2>((size_t)inq * NQ);
  q0_eps_dev_ = dalloc<double2>((size_t)d_.nw_b * inq);
  q0_S_ = dalloc<double2>((size_t)d_.ns * nkS * nso * nso);
  q0_tmp_ = dalloc<double2>((size_t)d_.nts * d_.ns * nkS * nso * nso);
  if (!q0_epst_) q0_epst_ = dalloc<double2>((size_t)d_.nts);
  GWHIP_CHECK(hipMemcpy(q0_qiq_, qiq.data(), (size_t)nk * sizeof(int), hipMemcpyHostToDevice));
  q0_qiq_host_ = qiq;
  GWHIP_CHECK(hipMemcpy(q0_A_, AqQ, (size_t)inq * NQ * sizeof(double2), hipMemcpyHostToDevice));
  GWHIP_CHECK(hipMemcpy(q0_S_, S, (size_t)d_.ns * nkS * nso * nso * sizeof(double2), hipMemcpyHostToDevice));
  q0_on_ = true;
}

// filter_q_abs: first occurrence of each distinct |q| in (0, 1] (exact comparison).
std::vector<int> gw_hip_kernel::q0_samples() const {
  std::vector<int> idx;
  std::vector<double> seen;
  for (int i = 0; i < (int)q0_qabs_.size(); ++i) {
    const double v = q0_qabs_[i];
    if (v > 0.0 && v <= 1.0 && std::find(seen.begin(), seen.end(), v) == seen.end()) {
      seen.push_back(v);
      idx.push_back(i);
    }
  }
  return idx;
}

// Fit, transform to tau and apply to Sigma [nts][ns][nkS] with G of the same layout.
void gw_hip_kernel::apply_q0(double2* Sigma, const double2* G, int64_t nkS) {
  if (nkS != q0_nS_)
    throw status_error(GWHIP_ERR_STATE, "the q -> 0 overlap matrices were given for " + std::to_string(q0_nS_) +
                                            " k points, this entry point returns " + std::to_string(nkS));
  const int64_t nwb = d_.nw_b, nts = d_.nts, inq = q0_inq_, nso = d_.nso, nao = d_.nao;
  std::vector<double2> eps((size_t)nwb * inq);
  GWHIP_CHECK(hipMemcpyAsync(eps.data(), q0_eps_dev_, eps.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  const std::vector<int> pick = q0_samples();
  std::vector<double> eps_w(nwb);
  for (int64_t n = 0; n < nwb; ++n) {
    // normal equations (A^T A) c = A^T y, A = [1, q, q^2]; Gaussian elimination with full pivoting
    double M[3][4] = {{0}};
    for (int i : pick) {
      const double x = q0_qabs_[i], y = eps[(size_t)n * inq + i].x;
      const double p[3] = {1.0, x, x * x};
      for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) M[r][c] += p[r] * p[c];
        M[r][3] += p[r] * y;
      }
    }
    int col[3] = {0, 1, 2};
    for (int piv = 0; piv < 3; ++piv) {
      int br = piv, bc = piv;
      for (int r = piv; r < 3; ++r)
        for (int c = piv; c < 3; ++c)
          if (std::fabs(M[r][c]) > std::fabs(M[br][bc])) br = r, bc = c;
      for (int c = 0; c < 4; ++c) std::swap(M[piv][c], M[br][c]);
      for (int r = 0; r < 3; ++r) std::swap(M[r][piv], M[r][bc]);
      std::swap(col[piv], col[bc]);
      if (M[piv][piv] == 0.0) throw status_error(GWHIP_ERR_ARG, "singular q -> 0 fit");
      for (int r = piv + 1; r < 3; ++r) {
        const double f = M[r][piv] / M[piv][piv];
        for (int c = piv; c < 4; ++c) M[r][c] -= f * M[piv][c];
      }
    }
    double z[3];
    for (int r = 2; r >= 0; --r) {
      double v = M[r][3];
      for (int c = r + 1; c < 3; ++c) v -= M[r][c] * z[c];
      z[r] = v / M[r][r];
    }
    double coef[3];
    for (int r = 0; r < 3; ++r) coef[col[r]] = z[r];
    eps_w[n] = coef[0];
  }
  q0_eps_t_.assign(nts, make_double2(0.0, 0.0));
  for (int64_t t = 0; t < nts; ++t) {
    double re = 0.0, im = 0.0;
    for (int64_t n = 0; n < nwb; ++n) {
      const double2 m = Ttb_host_[(size_t)t * nwb + n];
      re += m.x * eps_w[n];
      im += m.y * eps_w[n];
    }
    q0_eps_t_[t] = make_double2(re, im);
  }
  GWHIP_CHECK(hipMemcpyAsync(q0_epst_, q0_eps_t_.data(), (size_t)nts * sizeof(double2), hipMemcpyHostToDevice, stream_));
  const int64_t nmat = nts * d_.ns * nkS;
  const dim3 grid((unsigned)ceil_div(nso * nso, 256), (unsigned)std::min<int64_t>(nmat, 65535));
  const int x2c = d_.X2C ? 1 : 0;
  hipLaunchKernelGGL(q0_gs_kernel, grid, dim3(256), 0, stream_, G, q0_S_, q0_tmp_, (int)d_.ns, (int)nkS, (int)nso,
                     (int)nao, x2c, nmat);
  GWHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(q0_apply_kernel, grid, dim3(256), 0, stream_, Sigma, q0_tmp_, q0_S_, q0_epst_, q0_mad_, (int)d_.ns,
                     (int)nkS, (int)nso, (int)nao, x2c, nmat);
  GWHIP_CHECK(hipGetLastError());
  GWHIP_CHECK(hipStreamSynchronize(stream_));
}

}  // namespace gwhip
