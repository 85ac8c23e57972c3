// Host orchestration of the vector-Jacobian product of Sigma[G] (included by gw_hip_kernel.hip,
// which owns the kernel definitions; one translation unit).
//
// For a cotangent Sbar the VJP is the Gbar with Re<Sbar, dSigma> = Re<Gbar, dG> for every dG.
// Per q, with c = -1/nk, alpha = -pref/nk (the forward pass is src/gw_cpu_kernel.cpp:31-373):
//   forward pieces:  P0h(t) (half grid, Hermitised), P0w = T_fold P0h, Binv = (I - P0w)^-1,
//                    Ainv = H(Binv), P-tilde(tau) = T_tb Ainv P0w          [recomputed or stored]
//   Sigma part:      W_Q' = Sbar_s V_Q'              Y = conj(P-tilde) W           (V = V(k, k-q))
//                    Gbar(tau, k-q)_s += c sum_Q V_Q^H Y_Q
//                    T_Q = V_Q G(tau, k-q)_s         Ptbar += c conj(T) W^T
//   Dyson adjoint:   Pwbar = T_tb^H Ptbar
//                    P0wbar = Ainv^H Pwbar + Binv^H H(Pwbar P0w^H) Binv^H
//                    Pbar = H(T_fold^H P0wbar)
//   P0 part:         Zbar = alpha Pbar V(k, k+q)
//                    Gbar(beta - tau, k)_(b,a) += sum_Q Zbar_Q Gq^H V_Q^H
//                    Gbar(tau, k + q)_(a,b)    += sum_Q V_Q^H Gb^H Zbar_Q
// H(X) = (X + X^H)/2 is its own adjoint; the inverse's adjoint is -Binv^H (.) Binv^H.
#pragma once

namespace gwhip {

namespace {

template <int NB, bool HL, bool HR>
void launch_sandwich_t(dim3 grid, hipStream_t st, const OrbArgs& a, const OrbOp& L, const OrbOp& V, const OrbOp& R,
                       const OrbOut& X) {
  const size_t lds = orb_sandwich_lds(NB);
  allow_lds(orb_sandwich_kernel<NB, HL, HR>, lds);
  hipLaunchKernelGGL((orb_sandwich_kernel<NB, HL, HR>), grid, dim3(256), lds, st, a, L, V, R, X);
  GWHIP_CHECK(hipGetLastError());
}

template <bool HL, bool HR>
void launch_sandwich(int NB, dim3 grid, hipStream_t st, const OrbArgs& a, const OrbOp& L, const OrbOp& V,
                     const OrbOp& R, const OrbOut& X) {
  if (NB == 1) launch_sandwich_t<1, HL, HR>(grid, st, a, L, V, R, X);
  else if (NB == 2) launch_sandwich_t<2, HL, HR>(grid, st, a, L, V, R, X);
  else launch_sandwich_t<3, HL, HR>(grid, st, a, L, V, R, X);
}

template <int NB, bool HM>
void launch_reduce_t(dim3 grid, hipStream_t st, const OrbArgs& a, const OrbOp& A, const OrbOp& M, const OrbOp& C,
                     const OrbOut& O) {
  const size_t lds = orb_reduce_lds(NB);
  allow_lds(orb_reduce_kernel<NB, HM>, lds);
  hipLaunchKernelGGL((orb_reduce_kernel<NB, HM>), grid, dim3(256), lds, st, a, A, M, C, O);
  GWHIP_CHECK(hipGetLastError());
}

template <bool HM>
void launch_reduce(int NB, dim3 grid, hipStream_t st, const OrbArgs& a, const OrbOp& A, const OrbOp& M,
                   const OrbOp& C, const OrbOut& O) {
  if (NB == 1) launch_reduce_t<1, HM>(grid, st, a, A, M, C, O);
  else if (NB == 2) launch_reduce_t<2, HM>(grid, st, a, A, M, C, O);
  else launch_reduce_t<3, HM>(grid, st, a, A, M, C, O);
}

// Plain (batched) FP64 complex GEMM over matrices with row stride ld: C = alpha op(A) op(B) (+ C).
// tA / tB: operand stored transposed (element (m, k) at k*ld + m); conj applies op = conj.
void mat_gemm(hipStream_t st, int n, int64_t batch, const double2* A, bool tA, bool cA, const double2* B, bool tB,
              bool cB, double2* C, double alpha, bool acc) {
  const int64_t n2 = (int64_t)n * n;
  ZGemmArgs g{};
  g.M = n; g.N = n; g.K = n; g.batch = (int)batch; g.kin = n; g.nin = n;
  g.A = A; g.sAb = n2; g.sAm = tA ? 1 : n; g.sAko = 0; g.sAki = tA ? n : 1;
  g.B = B; g.sBb = n2; g.sBko = 0; g.sBki = tB ? 1 : n; g.sBno = 0; g.sBni = tB ? n : 1;
  g.C = C; g.sCb = n2; g.sCm = n; g.sCno = 0; g.sCni = 1;
  g.alpha = make_double2(alpha, 0.0);
  g.accumulate = acc ? 1 : 0;
  g.conjA = cA ? 1 : 0;
  g.conjB = cB ? 1 : 0;
  zgemm_f64(g, !tA, tB, st);
}

// C[m] = sum_j op(T)[m][j] X[j] over stacks of NQ^2 matrices; T is rows x cols (row-major).
// trans: use T^H (element (m, j) = conj(T[j][m])).
void stack_transform(hipStream_t st, const double2* T, int rows, int cols, bool trans, const double2* X,
                     int64_t NQ2, double2* C) {
  ZGemmArgs g{};
  const int M = trans ? cols : rows, K = trans ? rows : cols;
  g.M = M; g.N = (int)NQ2; g.K = K; g.batch = 1; g.kin = K; g.nin = (int)NQ2;
  g.A = T; g.sAm = trans ? 1 : cols; g.sAko = 0; g.sAki = trans ? cols : 1;
  g.B = X; g.sBko = 0; g.sBki = NQ2; g.sBno = 0; g.sBni = 1;
  g.C = C; g.sCm = NQ2; g.sCno = 0; g.sCni = 1;
  g.alpha = make_double2(1.0, 0.0);
  g.conjA = trans ? 1 : 0;
  zgemm_f64(g, !trans, false, st);
}

}  // namespace

void gw_hip_kernel::orb_gather_view(const OrbArgs& a, const OrbOp& op, int tz, int64_t n) {
  const int64_t nao2 = d_.nao * d_.nao;
  if (n < 1 || n > 65535) throw status_error(GWHIP_ERR_UNSUPPORTED, "gather: k count out of range");
  hipLaunchKernelGGL(orb_gather_view_kernel, dim3((unsigned)ceil_div(nao2, 256), (unsigned)n), dim3(256), 0, stream_, a,
                     op, tz, orb_G_);
  GWHIP_CHECK(hipGetLastError());
}

// X_Q = L V_Q R for the grid's (k, tz): the kernel, or (orb_gemm()) one batched GEMM over k per
// tz with the other factor gathered per k.  The GEMM form takes the shapes the VJP uses: exactly
// one of L, R; V and X without a k map; rows (Q, r) of V and X equally spaced when R is present.
void gw_hip_kernel::vjp_sandwich(bool has_l, bool has_r, int NB, dim3 grid, const OrbArgs& a, const OrbOp& L,
                                 const OrbOp& V, const OrbOp& R, const OrbOut& X) {
  if (!orb_gemm()) {
    if (has_l && !has_r) launch_sandwich<true, false>(NB, grid, stream_, a, L, V, R, X);
    else if (!has_l && has_r) launch_sandwich<false, true>(NB, grid, stream_, a, L, V, R, X);
    else launch_sandwich<true, true>(NB, grid, stream_, a, L, V, R, X);
    return;
  }
  const int64_t nk = a.nk, nao = a.nao, NQ = a.NQ, nao2 = nao * nao;
  if (has_l == has_r || V.kmap != 0 || X.kmap != 0)
    throw status_error(GWHIP_ERR_UNSUPPORTED, "GEMM form of this orbital product not implemented");
  if (has_r && (V.s_q != nao * V.rs || X.s_q != nao * X.rs))
    throw status_error(GWHIP_ERR_UNSUPPORTED, "GEMM form needs equally spaced (Q, row) rows");
  for (int tz = 0; tz < (int)grid.z; ++tz) {
    orb_gather_view(a, has_l ? L : R, tz, nk);
    ZGemmArgs g{};
    g.batch = (int)nk; g.kin = (int)nao; g.nin = (int)nao; g.K = (int)nao;
    g.alpha = make_double2(1.0, 0.0);
    if (has_l) {  // X[k](r, (Q, c)) = sum_r' L[k](r, r') V_Q(r', c)
      g.M = (int)nao; g.N = (int)(NQ * nao);
      g.A = orb_G_; g.sAb = nao2; g.sAm = nao; g.sAko = 0; g.sAki = 1;
      g.B = V.p + tz * V.s_t; g.sBb = V.s_k; g.sBko = 0; g.sBki = V.rs; g.sBno = V.s_q; g.sBni = V.cs;
      g.conjB = V.conj;
      g.C = X.p + tz * X.s_t; g.sCb = X.s_k; g.sCm = X.rs; g.sCno = X.s_q; g.sCni = X.cs;
      zgemm_f64(g, true, false, stream_);
    } else {      // X[k]((Q, r), c) = sum_c' V_Q(r, c') R[k](c', c)
      g.M = (int)(NQ * nao); g.N = (int)nao;
      g.A = V.p + tz * V.s_t; g.sAb = V.s_k; g.sAm = V.rs; g.sAko = 0; g.sAki = V.cs;
      g.conjA = V.conj;
      g.B = orb_G_; g.sBb = nao2; g.sBko = 0; g.sBki = nao; g.sBno = 0; g.sBni = 1;
      g.C = X.p + tz * X.s_t; g.sCb = X.s_k; g.sCm = X.rs; g.sCno = 0; g.sCni = X.cs;
      zgemm_f64(g, V.cs == 1, false, stream_);
    }
  }
}

// O(tz, kidx(k)) += scale sum_Q A_Q M C_Q: the kernel, or (orb_gemm()) per tz Z'_Q = M C_Q as one
// batched GEMM (M gathered per k), then sum_{Q, m} A_Q(i, m) Z'_Q(m, j) as a split-K GEMM per k,
// then added to O through the k map.  The GEMM form needs A and C without a k map.
void gw_hip_kernel::vjp_reduce(bool has_m, int NB, dim3 grid, const OrbArgs& a, const OrbOp& A, const OrbOp& M,
                               const OrbOp& C, const OrbOut& O) {
  if (!orb_gemm()) {
    if (has_m) launch_reduce<true>(NB, grid, stream_, a, A, M, C, O);
    else launch_reduce<false>(NB, grid, stream_, a, A, M, C, O);
    return;
  }
  const int64_t nk = a.nk, nao = a.nao, NQ = a.NQ, nao2 = nao * nao, vq = NQ * nao2;
  if (A.kmap != 0 || C.kmap != 0)
    throw status_error(GWHIP_ERR_UNSUPPORTED, "GEMM form of this orbital reduction not implemented");
  const int64_t tiles = nk * ceil_div(nao, 64) * ceil_div(nao, 64);
  const int ksplit = (int)std::min<int64_t>(kOrbKsplit, std::max<int64_t>(1, ceil_div(1024, tiles)));
  for (int tz = 0; tz < (int)grid.y; ++tz) {
    const double2* cp = C.p + tz * C.s_t;
    int64_t c_sk = C.s_k, c_sq = C.s_q, c_rs = C.rs, c_cs = C.cs;
    int c_conj = C.conj;
    if (has_m) {  // Z'[k](i, (Q, j)) = sum_m M[k](i, m) C_Q(m, j), canonical [k][Q][nao^2]
      orb_gather_view(a, M, tz, nk);
      ZGemmArgs z{};
      z.M = (int)nao; z.N = (int)(NQ * nao); z.K = (int)nao; z.batch = (int)nk; z.kin = (int)nao; z.nin = (int)nao;
      z.A = orb_G_; z.sAb = nao2; z.sAm = nao; z.sAko = 0; z.sAki = 1;
      z.B = cp; z.sBb = C.s_k; z.sBko = 0; z.sBki = C.rs; z.sBno = C.s_q; z.sBni = C.cs;
      z.conjB = C.conj;
      z.C = orb_W_; z.sCb = vq; z.sCm = nao; z.sCno = nao2; z.sCni = 1;
      z.alpha = make_double2(1.0, 0.0);
      zgemm_f64(z, true, false, stream_);
      cp = orb_W_; c_sk = vq; c_sq = nao2; c_rs = nao; c_cs = 1; c_conj = 0;
    }
    ZGemmArgs y{};  // orb_O_[k](i, j) = sum_{Q, m} A_Q(i, m) Z'_Q(m, j)
    y.M = (int)nao; y.N = (int)nao; y.K = (int)(NQ * nao); y.batch = (int)nk; y.kin = (int)nao; y.nin = (int)nao;
    y.A = A.p + tz * A.s_t; y.sAb = A.s_k; y.sAm = A.rs; y.sAko = A.s_q; y.sAki = A.cs;
    y.conjA = A.conj;
    y.B = cp; y.sBb = c_sk; y.sBko = c_sq; y.sBki = c_rs; y.sBno = 0; y.sBni = c_cs;
    y.conjB = c_conj;
    y.C = orb_O_; y.sCb = nao2; y.sCm = nao; y.sCno = 0; y.sCni = 1;
    y.alpha = make_double2(1.0, 0.0);
    zgemm_f64_splitk(y, ksplit, orb_part_, A.cs == 1, c_rs == 1, stream_);
    hipLaunchKernelGGL(orb_scatter_add_kernel, dim3((unsigned)ceil_div(nao2, 256), (unsigned)nk), dim3(256), 0, stream_,
                       a, O, tz, orb_O_);
    GWHIP_CHECK(hipGetLastError());
  }
}

void gw_hip_kernel::set_vjp_store(bool keep) {
  GWHIP_CHECK(hipSetDevice(device_));
  keep_p0_ = keep;
  stored_valid_ = false;
  if (keep && !P0keep_) P0keep_ = dalloc<double2>((size_t)d_.nk * (d_.nts / 2) * d_.NQ * d_.NQ);
  if (!keep && P0keep_) {
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    (void)hipFree(P0keep_);
    bytes_ -= (size_t)d_.nk * (d_.nts / 2) * d_.NQ * d_.NQ * sizeof(double2);
    P0keep_ = nullptr;
  }
}

void gw_hip_kernel::ensure_vjp_buffers() {
  if (vW_) return;
  const int64_t NQ2 = d_.NQ * d_.NQ, nwb = d_.nw_b, nts = d_.nts, half = nts / 2;
  vP0h_ = dalloc<double2>((size_t)half * NQ2);
  vP0w_ = dalloc<double2>((size_t)nwb * NQ2);
  vBinv_ = dalloc<double2>((size_t)nwb * NQ2);
  vAinv_ = dalloc<double2>((size_t)nwb * NQ2);
  vM1_ = dalloc<double2>((size_t)nwb * NQ2);
  vM2_ = dalloc<double2>((size_t)nwb * NQ2);
  vM3_ = dalloc<double2>((size_t)nwb * NQ2);
  vPt_ = dalloc<double2>((size_t)nts * NQ2);
  vPtbar_ = dalloc<double2>((size_t)nts * NQ2);
  vPbar_ = dalloc<double2>((size_t)half * NQ2);
  // second orbital scratch of the same tau chunk as the forward scratch, halved until it fits
  const size_t per_t = (size_t)d_.nk * d_.NQ * d_.nao * d_.nao * sizeof(double2);
  int64_t tc = std::max(tchunk_sigma_, tchunk_p0_);
  while (tc >= 1) {
    void* p = nullptr;
    if (hipMalloc(&p, tc * per_t) == hipSuccess) {
      vW_ = static_cast<double2*>(p);
      bytes_ += tc * per_t;
      break;
    }
    (void)hipGetLastError();
    tc /= 2;
  }
  if (!vW_) throw status_error(GWHIP_ERR_OOM, "no room for the VJP scratch buffer");
  tchunk_vjp_ = tc;
}

void gw_hip_kernel::free_vjp_buffers() {
  for (double2** p : {&vP0h_, &vP0w_, &vBinv_, &vAinv_, &vM1_, &vM2_, &vM3_, &vPt_, &vPtbar_, &vPbar_, &vW_, &P0keep_}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
}

void gw_hip_kernel::vjp_device(const double2* G, const double2* Sbar, double2* Gbar, int mode) {
  if (!grid_set_) throw status_error(GWHIP_ERR_STATE, "transform grid not set (gwhip_set_grid)");
  if (!v_set_) throw status_error(GWHIP_ERR_STATE, "integrals not set (gwhip_upload_V / gwhip_generate_V)");
  if (!G || !Sbar || !Gbar) throw status_error(GWHIP_ERR_ARG, "null G, Sigma_bar or G_bar");
  if (mode != GWHIP_VJP_RECOMPUTE && mode != GWHIP_VJP_STORED) throw status_error(GWHIP_ERR_ARG, "unknown VJP mode");
  const bool stored = mode == GWHIP_VJP_STORED;
  if (stored && !stored_valid_)
    throw status_error(GWHIP_ERR_STATE, "no stored forward pass (gwhip_set_vjp_store(h, 1), then gwhip_eval_sigma)");
  GWHIP_CHECK(hipSetDevice(device_));
  ensure_vjp_buffers();
  ensure_orb_buffers();
  detect_hermitian_G(G);

  const int64_t nk = d_.nk, NQ = d_.NQ, nao = d_.nao, nso = d_.nso, nao2 = nao * nao, NQ2 = NQ * NQ;
  const int64_t nts = d_.nts, half = nts / 2, nwb = d_.nw_b, nso2 = nso * nso;
  const int64_t Gt = d_.ns * nk * nso2;                 // tau stride of G, Sigma
  const int64_t per_t = nk * NQ * nao2;                 // tau stride of the orbital scratch
  const double c = -1.0 / (double)nk, alpha = -d_.prefactor / (double)nk;
  const int NB = (int)ceil_div(nao, 16);
  const int QPW = 4;
  const unsigned gx = (unsigned)ceil_div(NQ, 4 * QPW);
  const int pns = (int)d_.pns;
  auto goff = [&](int64_t t, int64_t is) { return (t * d_.ns + is) * nk * nso2; };
  const size_t nsig = (size_t)nts * d_.ns * nk * nso2;

  Event e0, e1;
  GWHIP_CHECK(hipEventRecord(e0, stream_));
  GWHIP_CHECK(hipMemsetAsync(Gbar, 0, nsig * sizeof(double2), stream_));
  GWHIP_CHECK(hipMemsetAsync(dyson_flag_, 0, sizeof(int), stream_));

  for (int64_t q = 0; q < nk; ++q) {
    OrbArgs oa{(int)nao, (int)NQ, (int)nk, QPW, kplus_tab_ + q * nk, kminus_tab_ + q * nk, 1.0};
    // ---- forward pieces of this q
    const double2* P0h = P0keep_ ? P0keep_ + q * half * NQ2 : nullptr;
    if (!stored) {
      p0_one_q(q, G, false, vP0h_);
      P0h = vP0h_;
    }
    stack_transform(stream_, Tbt_fold_, (int)nwb, (int)half, false, P0h, NQ2, vP0w_);
    hipLaunchKernelGGL(identity_minus_kernel, dim3((unsigned)ceil_div(nwb * NQ2, 256)), dim3(256), 0, stream_, vBinv_,
                       vP0w_, (int)NQ, nwb);
    GWHIP_CHECK(hipGetLastError());
    // B = I - P0(i Omega) is Hermitian (P0(tau) is Hermitised and the folded transform is real on
    // it), so B^-1 comes from the Hermitian block sweep of the forward's Dyson step, its upper
    // triangle mirrored: the same B^-1 the forward used, Hermitian by construction (A^-1 = Herm(B^-1))
    block_inverse(vBinv_, nwb);
    mirror_upper(vBinv_, (int)NQ, NQ2, nwb, stream_);
    GWHIP_CHECK(hipGetLastError());
    GWHIP_CHECK(hipMemcpyAsync(vAinv_, vBinv_, nwb * NQ2 * sizeof(double2), hipMemcpyDeviceToDevice, stream_));
    const double2* Pt = P_ + q * nts * NQ2;
    if (!stored) {
      mat_gemm(stream_, (int)NQ, nwb, vAinv_, false, false, vP0w_, false, false, vM1_, 1.0, false);
      stack_transform(stream_, Ttb_, (int)nts, (int)nwb, false, vM1_, NQ2, vPt_);
      Pt = vPt_;
    }

    // ---- Sigma part: Gbar(tau, k - q) and Ptbar(tau)
    GWHIP_CHECK(hipMemsetAsync(vPtbar_, 0, nts * NQ2 * sizeof(double2), stream_));
    const double2* Vm = Vq_ + (int64_t)neg_q_[q] * nk * NQ * nao2;  // [k][Q] = V(k, k - q)
    const double2* Vp = Vq_ + q * nk * NQ * nao2;                   // [k][Q] = V(k, k + q)
    const OrbOp Vrow{Vm, 0, NQ * nao2, nao2, nao, 1, 0, 0};           // V(k, k-q)_Q (i, m)
    const OrbOp VmH{Vm, 0, NQ * nao2, nao2, 1, nao, 0, 1};            // V(k, k-q)_Q^H (m, i)
    const OrbOp none{nullptr, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t t0 = 0; t0 < nts; t0 += tchunk_vjp_) {
      const int64_t tc = std::min(tchunk_vjp_, nts - t0);
      for (int s = 0; s < d_.pseudo_ns; ++s) {
        const int64_t is = s / (pns * pns), a = (s / pns) % pns, b = s % pns;
        const int64_t blk = a * nao * nso + b * nao;
        // W_Q' = Sbar_s V_Q'
        const OrbOp Ls{Sbar + goff(t0, is) + blk, Gt, nso2, 0, nso, 1, 0, 0};
        const OrbOut W{vW_, per_t, NQ * nao2, nao2, nao, 1, 0};
        vjp_sandwich(true, false, NB, dim3(gx, (unsigned)nk, (unsigned)tc), oa, Ls, Vrow, none, W);
        {  // Y = conj(P-tilde) W
          ZGemmArgs g{};
          g.M = (int)NQ; g.N = (int)(nk * nao2); g.K = (int)NQ; g.batch = (int)tc; g.kin = (int)NQ; g.nin = (int)nao2;
          g.A = Pt + t0 * NQ2; g.sAb = NQ2; g.sAm = NQ; g.sAko = 0; g.sAki = 1;
          g.B = vW_; g.sBb = per_t; g.sBko = 0; g.sBki = nao2; g.sBno = NQ * nao2; g.sBni = 1;
          g.C = scratch_; g.sCb = per_t; g.sCm = nao2; g.sCno = NQ * nao2; g.sCni = 1;
          g.alpha = make_double2(1.0, 0.0);
          g.conjA = 1;
          zgemm_f64(g, true, false, stream_);
        }
        {  // Gbar(tau, k - q)_s += c sum_Q V_Q^H Y_Q
          OrbArgs oc = oa;
          oc.scale = c;
          const OrbOp Y{scratch_, per_t, NQ * nao2, nao2, nao, 1, 0, 0};
          const OrbOut O{Gbar + goff(t0, is) + blk, Gt, nso2, 0, nso, 1, 2};
          vjp_reduce(false, NB, dim3((unsigned)nk, (unsigned)tc), oc, VmH, none, Y, O);
        }
        {  // T_Q = V_Q G(tau, k - q)_s
          const OrbOp Gs{G + goff(t0, is) + blk, Gt, nso2, 0, nso, 1, 2, 0};
          const OrbOut T{scratch_, per_t, NQ * nao2, nao2, nao, 1, 0};
          vjp_sandwich(false, true, NB, dim3(gx, (unsigned)nk, (unsigned)tc), oa, none, Vrow, Gs, T);
        }
        {  // Ptbar += c conj(T) W^T
          ZGemmArgs g{};
          g.M = (int)NQ; g.N = (int)NQ; g.K = (int)(nk * nao2); g.batch = (int)tc; g.kin = (int)nao2; g.nin = (int)NQ;
          g.A = scratch_; g.sAb = per_t; g.sAm = nao2; g.sAko = NQ * nao2; g.sAki = 1;
          g.B = vW_; g.sBb = per_t; g.sBko = NQ * nao2; g.sBki = 1; g.sBno = 0; g.sBni = nao2;
          g.C = vPtbar_ + t0 * NQ2; g.sCb = NQ2; g.sCm = NQ; g.sCno = 0; g.sCni = 1;
          g.alpha = make_double2(c, 0.0);
          g.accumulate = 1;
          g.conjA = 1;
          zgemm_f64(g, true, true, stream_);
        }
      }
    }

    // ---- Dyson adjoint
    stack_transform(stream_, Ttb_, (int)nts, (int)nwb, true, vPtbar_, NQ2, vM1_);              // Pwbar
    mat_gemm(stream_, (int)NQ, nwb, vM1_, false, false, vP0w_, true, true, vM2_, 1.0, false);  // Pwbar P0w^H
    hermitize(vM2_, (int)NQ, NQ2, nwb, stream_);
    GWHIP_CHECK(hipGetLastError());
    mat_gemm(stream_, (int)NQ, nwb, vAinv_, true, true, vM1_, false, false, vM3_, 1.0, false);  // Ainv^H Pwbar
    mat_gemm(stream_, (int)NQ, nwb, vBinv_, true, true, vM2_, false, false, vM1_, 1.0, false);  // Binv^H H(.)
    mat_gemm(stream_, (int)NQ, nwb, vM1_, false, false, vBinv_, true, true, vM3_, 1.0, true);   // (.) Binv^H
    stack_transform(stream_, Tbt_fold_, (int)nwb, (int)half, true, vM3_, NQ2, vPbar_);          // T_fold^H P0wbar
    hermitize(vPbar_, (int)NQ, NQ2, half, stream_);
    GWHIP_CHECK(hipGetLastError());

    // ---- P0 part: Gbar(beta - tau, k) and Gbar(tau, k + q)
    const OrbOp VpH{Vp, 0, NQ * nao2, nao2, 1, nao, 0, 1};  // V(k, k+q)_Q^H
    for (int64_t t0 = 0; t0 < half; t0 += tchunk_vjp_) {
      const int64_t tc = std::min(tchunk_vjp_, half - t0);
      const int64_t tt0 = nts - 1 - t0;
      {  // Zbar = alpha Pbar V(k, k + q)
        ZGemmArgs g{};
        g.M = (int)NQ; g.N = (int)(nk * nao2); g.K = (int)NQ; g.batch = (int)tc; g.kin = (int)NQ; g.nin = (int)nao2;
        g.A = vPbar_ + t0 * NQ2; g.sAb = NQ2; g.sAm = NQ; g.sAko = 0; g.sAki = 1;
        g.B = Vp; g.sBb = 0; g.sBko = 0; g.sBki = nao2; g.sBno = NQ * nao2; g.sBni = 1;
        g.C = vW_; g.sCb = per_t; g.sCm = nao2; g.sCno = NQ * nao2; g.sCni = 1;
        g.alpha = make_double2(alpha, 0.0);
        zgemm_f64(g, true, false, stream_);
      }
      const OrbOp Zb{vW_, per_t, NQ * nao2, nao2, nao, 1, 0, 0};
      for (int s = 0; s < d_.pseudo_ns; ++s) {
        const int64_t is = s / (pns * pns), a = (s / pns) % pns, b = s % pns;
        const int64_t blk_ab = a * nao * nso + b * nao, blk_ba = b * nao * nso + a * nao;
        {  // Gbar(beta - tau, k)_(b,a) += sum_Q Zbar_Q Gq^H V_Q^H,  Gq = G(tau, k + q)_(a,b)
          const OrbOp GqH{G + goff(t0, is) + blk_ab, Gt, nso2, 0, 1, nso, 1, 1};
          const OrbOut O{Gbar + goff(tt0, is) + blk_ba, -Gt, nso2, 0, nso, 1, 0};
          vjp_reduce(true, NB, dim3((unsigned)nk, (unsigned)tc), oa, Zb, GqH, VpH, O);
        }
        {  // Gbar(tau, k + q)_(a,b) += sum_Q V_Q^H Gb^H Zbar_Q,  Gb = G(beta - tau, k)_(b,a)
          const OrbOp GbH{G + goff(tt0, is) + blk_ba, -Gt, nso2, 0, 1, nso, 0, 1};
          const OrbOut O{Gbar + goff(t0, is) + blk_ab, Gt, nso2, 0, nso, 1, 1};
          vjp_reduce(true, NB, dim3((unsigned)nk, (unsigned)tc), oa, VpH, GbH, Zb, O);
        }
      }
    }
  }
  GWHIP_CHECK(hipEventRecord(e1, stream_));
  GWHIP_CHECK(hipMemsetAsync(flag_, 0, sizeof(int), stream_));
  hipLaunchKernelGGL(nonfinite_kernel, dim3(1024), dim3(256), 0, stream_, Gbar, (int64_t)nsig, flag_);
  GWHIP_CHECK(hipGetLastError());
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  resolve_timings();
  float ms = 0.f;
  GWHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  vjp_ms_ = ms;
  int flag = 0;  // non-finite first: a NaN or Inf in G also fails the Dyson step's pivot test
  GWHIP_CHECK(hipMemcpy(&flag, flag_, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) throw status_error(GWHIP_ERR_NONFINITE, "G_bar contains non-finite values");
  check_dyson_flag();
}

void gw_hip_kernel::vjp_host(const double2* G, const double2* Sbar, double2* Gbar, int mode) {
  if (!G || !Sbar || !Gbar) throw status_error(GWHIP_ERR_ARG, "null G, Sigma_bar or G_bar");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t nsig = (size_t)d_.nts * d_.ns * d_.nk * d_.nso * d_.nso;
  if (!G_dev_) G_dev_ = dalloc<double2>(nsig);
  if (!S_dev_) S_dev_ = dalloc<double2>(nsig);
  double2* Gb_dev = dalloc<double2>(nsig);
  try {
    GWHIP_CHECK(hipMemcpy(G_dev_, G, nsig * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(S_dev_, Sbar, nsig * sizeof(double2), hipMemcpyHostToDevice));
    vjp_device(G_dev_, S_dev_, Gb_dev, mode);
    GWHIP_CHECK(hipMemcpy(Gbar, Gb_dev, nsig * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(Gb_dev);
    bytes_ -= nsig * sizeof(double2);
    throw;
  }
  (void)hipFree(Gb_dev);
  bytes_ -= nsig * sizeof(double2);
}

}  // namespace gwhip
