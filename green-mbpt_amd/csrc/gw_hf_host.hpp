// Hartree-Fock Fock contribution on the device from the resident integrals (included by
// gw_hip_kernel.hip after gw_vjp_host.hpp; one translation unit).  The counterpart of
// hf_scalar_cpu_kernel::solve / hf_x2c_cpu_kernel::solve (src/hf_cpu_kernels.cpp:24-279) for a
// full k mesh:
//   direct    J[Q] = (1/nk) sum_{k', s} sum_{ab} V(k', k')[Q, a, b] dm_s(k')[b, a]
//             F_s(k) += sum_Q J[Q] V(k, k)[Q]
//   exchange  F_s(k) -= pref / nk  sum_{k'} sum_Q V(k, k')_Q dm_s(k') V(k, k')_Q^H
//   Madelung  F_s(k) -= pref * madelung * S dm_s(k) S
// pref = 1/2 for ns = 1, 1 for ns = 2; X2C: spin blocks aa, bb, ab with pref 1, the Madelung
// term with the S_aa block on both sides, the direct term with dm_aa + dm_bb, F_ba = F_ab^H.
#pragma once

namespace gwhip {

// J[Q] over k' and spin; one block per Q.  dm: [nspin][nk][nso][nso].
__global__ __launch_bounds__(256) void hf_direct_upper_kernel(const double2* __restrict__ V0, const double2* __restrict__ dm,
                                                               int nk, int NQ, int nao, int nso, int nspin, int x2c,
                                                               double2* __restrict__ J) {
  const int Q = blockIdx.x;
  const int nao2 = nao * nao;
  double sr = 0.0, si = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)nk * nao2; e += blockDim.x) {
    const int k = (int)(e / nao2), ab = (int)(e - (int64_t)k * nao2), a = ab / nao, b = ab - (ab / nao) * nao;
    const double2 v = V0[((int64_t)k * NQ + Q) * nao2 + ab];
    double2 x = make_double2(0.0, 0.0);
    for (int s = 0; s < nspin; ++s) {
      const double2* d = dm + ((int64_t)s * nk + k) * nso * nso;
      const double2 y = d[(int64_t)b * nso + a];
      x.x += y.x;
      x.y += y.y;
      if (x2c) {
        const double2 z = d[(int64_t)(nao + b) * nso + nao + a];
        x.x += z.x;
        x.y += z.y;
      }
    }
    sr += v.x * x.x - v.y * x.y;
    si += v.x * x.y + v.y * x.x;
  }
  __shared__ double red[2][256];
  red[0][threadIdx.x] = sr;
  red[1][threadIdx.x] = si;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) J[Q] = make_double2(red[0][0] / nk, red[1][0] / nk);
}

// F_s(k)[i][j] += sum_Q J[Q] V(k, k)[Q][i][j] (and the bb block for X2C); one thread per (s, k, i, j).
__global__ void hf_direct_add_kernel(const double2* __restrict__ J, const double2* __restrict__ V0, double2* __restrict__ F,
                                     int nk, int NQ, int nao, int nso, int nspin, int x2c) {
  const int nao2 = nao * nao;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nspin * nk * nao2) return;
  const int s = (int)(e / ((int64_t)nk * nao2));
  const int64_t r = e - (int64_t)s * nk * nao2;
  const int k = (int)(r / nao2), ij = (int)(r - (int64_t)k * nao2), i = ij / nao, j = ij - (ij / nao) * nao;
  double sr = 0.0, si = 0.0;
  for (int Q = 0; Q < NQ; ++Q) {
    const double2 u = J[Q], v = V0[((int64_t)k * NQ + Q) * nao2 + ij];
    sr += u.x * v.x - u.y * v.y;
    si += u.x * v.y + u.y * v.x;
  }
  double2* f = F + ((int64_t)s * nk + k) * nso * nso;
  f[(int64_t)i * nso + j].x += sr;
  f[(int64_t)i * nso + j].y += si;
  if (x2c) {
    f[(int64_t)(nao + i) * nso + nao + j].x += sr;
    f[(int64_t)(nao + i) * nso + nao + j].y += si;
  }
}

// F(k)_blk[i][j] += scale * sum_q X[blk][q][k][i][j]; one thread per (blk, k, i, j).  Blocks: aa
// (offset 0), bb (nao, nao), ab (0, nao) of an nso x nso matrix.
__global__ void hf_exchange_sum_kernel(const double2* __restrict__ X, int nk, int nao, int nso, int nblk, double scale,
                                       double2* __restrict__ F) {
  const int nao2 = nao * nao;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nblk * nk * nao2) return;
  const int bl = (int)(e / ((int64_t)nk * nao2));
  const int64_t r = e - (int64_t)bl * nk * nao2;
  const int k = (int)(r / nao2), ij = (int)(r - (int64_t)k * nao2), i = ij / nao, j = ij - (ij / nao) * nao;
  const double2* x = X + (int64_t)bl * nk * nk * nao2 + (int64_t)k * nao2 + ij;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < nk; ++q) {
    const double2 v = x[(int64_t)q * nk * nao2];
    sr += v.x;
    si += v.y;
  }
  const int ro = (bl == 1) ? nao : 0, co = (bl == 0) ? 0 : nao;
  double2* f = F + (int64_t)k * nso * nso + (int64_t)(ro + i) * nso + co + j;
  f->x += scale * sr;
  f->y += scale * si;
}

// X2C: F_ba = F_ab^H per k.
__global__ void hf_x2c_adjoint_kernel(double2* F, int nk, int nao, int nso) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nk * nao * nao) return;
  const int k = (int)(e / (nao * nao)), ij = (int)(e - (int64_t)k * nao * nao), i = ij / nao, j = ij - (ij / nao) * nao;
  double2* f = F + (int64_t)k * nso * nso;
  const double2 v = f[(int64_t)j * nso + nao + i];  // (a, b) block element (j, i)
  f[(int64_t)(nao + i) * nso + j] = make_double2(v.x, -v.y);
}

namespace {
// C[b] (+)= alpha * A[b] B[b] for nb square n x n blocks with leading dimension ld and batch
// stride bs (blocks of larger matrices).
void block_gemm(hipStream_t st, int n, int64_t nb, const double2* A, const double2* B, double2* C, int64_t ld, int64_t bs,
                double alpha, bool acc) {
  ZGemmArgs g{};
  g.M = n; g.N = n; g.K = n; g.batch = (int)nb; g.kin = n; g.nin = n;
  g.A = A; g.sAb = bs; g.sAm = ld; g.sAko = 0; g.sAki = 1;
  g.B = B; g.sBb = bs; g.sBko = 0; g.sBki = ld; g.sBno = 0; g.sBni = 1;
  g.C = C; g.sCb = bs; g.sCm = ld; g.sCno = 0; g.sCni = 1;
  g.alpha = make_double2(alpha, 0.0);
  g.accumulate = acc ? 1 : 0;
  zgemm_f64(g, true, false, st);
}
}  // namespace

void gw_hip_kernel::hf_device(const double2* dm, const double2* S, double madelung, double2* F) {
  if (!v_set_) throw status_error(GWHIP_ERR_STATE, "integrals not set (gwhip_upload_V / gwhip_generate_V)");
  if (!dm || !S || !F) throw status_error(GWHIP_ERR_ARG, "null dm, S or F");
  GWHIP_CHECK(hipSetDevice(device_));
  ensure_orb_buffers();
  const int64_t nk = d_.nk, NQ = d_.NQ, nao = d_.nao, nso = d_.nso, nao2 = nao * nao, nso2 = nso * nso;
  const bool x2c = d_.X2C;
  const int64_t nspin = d_.ns;  // 1 for X2C
  const int64_t nmat = nspin * nk;
  const double pref = (x2c || d_.ns == 2) ? 1.0 : 0.5;
  // V(k, k) is the row q0 of the resident store whose k + q0 map is the identity
  int64_t q0 = -1;
  for (int64_t q = 0; q < nk && q0 < 0; ++q) {
    bool id = true;
    for (int64_t k = 0; k < nk && id; ++k) id = kplus_host_[(size_t)k * nk + q] == k;
    if (id) q0 = q;
  }
  if (q0 < 0) throw status_error(GWHIP_ERR_ARG, "k + q map has no identity row (q = 0)");
  const double2* Vd = Vq_ + q0 * nk * NQ * nao2;
  if (!hf_work_) hf_work_ = dalloc<double2>((size_t)std::max<int64_t>(NQ, nmat * nso2));
  Event e0, e1;
  GWHIP_CHECK(hipEventRecord(e0, stream_));
  zero(F, nmat * nso2);
  hipLaunchKernelGGL(hf_direct_upper_kernel, dim3((unsigned)NQ), dim3(256), 0, stream_, Vd, dm, (int)nk, (int)NQ,
                     (int)nao, (int)nso, (int)nspin, x2c ? 1 : 0, hf_work_);
  GWHIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(hf_direct_add_kernel, dim3((unsigned)ceil_div(nmat * nao2, 256)), dim3(256), 0, stream_, hf_work_, Vd,
                     F, (int)nk, (int)NQ, (int)nao, (int)nso, (int)nspin, x2c ? 1 : 0);
  GWHIP_CHECK(hipGetLastError());
  // exchange: X[blk][q][k] = sum_Q V(k, k+q)_Q dm(k+q)_blk V(k, k+q)_Q^H for all (k, q) in one
  // launch per spin block, then F(k)_blk -= pref/nk sum_q X[blk][q][k] (fixed order in q)
  const int NB = (int)ceil_div(nao, 16);
  const int nblk = x2c ? 3 : 1;
  if (!hf_x_) hf_x_ = dalloc<double2>((size_t)nblk * nk * nk * nao2);
  Event ex0, ex1;
  float ex_ms = 0.f;
  const OrbOp A{Vq_, nk * NQ * nao2, NQ * nao2, nao2, nao, 1, 0, 0};  // V(k, k+q)_Q (i, a)
  const OrbOp C{Vq_, nk * NQ * nao2, NQ * nao2, nao2, 1, nao, 0, 1};  // V(k, k+q)_Q^H (b, j)
  OrbArgs oa{(int)nao, (int)NQ, (int)nk, 4, kplus_tab_, kminus_tab_, 1.0};
  oa.kmap_st = nk;
  for (int64_t s = 0; s < nspin; ++s) {
    zero(hf_x_, nblk * nk * nk * nao2);
    if (orb_gemm()) {
      GWHIP_CHECK(hipEventRecord(ex0, stream_));
      hf_exchange_gemm(dm + s * nk * nso2, nblk);
      GWHIP_CHECK(hipEventRecord(ex1, stream_));
      GWHIP_CHECK(hipEventSynchronize(ex1));
      float m = 0.f;
      GWHIP_CHECK(hipEventElapsedTime(&m, ex0, ex1));
      ex_ms += m;
    }
    for (int bl = 0; bl < nblk && !orb_gemm(); ++bl) {
      const int64_t ro = (bl == 1) ? nao : 0, co = (bl == 0) ? 0 : nao;  // aa, bb, ab
      const OrbOp M{dm + s * nk * nso2 + ro * nso + co, 0, nso2, 0, nso, 1, 1, 0};
      const OrbOut O{hf_x_ + bl * nk * nk * nao2, nk * nao2, nao2, 0, nao, 1, 0};
      GWHIP_CHECK(hipEventRecord(ex0, stream_));
      launch_reduce<true>(NB, dim3((unsigned)nk, (unsigned)nk), stream_, oa, A, M, C, O);
      GWHIP_CHECK(hipEventRecord(ex1, stream_));
      GWHIP_CHECK(hipEventSynchronize(ex1));
      float m = 0.f;
      GWHIP_CHECK(hipEventElapsedTime(&m, ex0, ex1));
      ex_ms += m;
    }
    hipLaunchKernelGGL(hf_exchange_sum_kernel, dim3((unsigned)ceil_div(nblk * nk * nao2, 256)), dim3(256), 0, stream_,
                       hf_x_, (int)nk, (int)nao, (int)nso, nblk, -pref / (double)nk, F + s * nk * nso2);
    GWHIP_CHECK(hipGetLastError());
  }
  // Madelung: F -= c S dm S (X2C: S_aa on both sides of every block)
  if (madelung != 0.0) {
    double2* tmp = hf_work_;
    for (int bl = 0; bl < nblk; ++bl) {
      const int64_t ro = (bl == 1) ? nao : 0, co = (bl == 0) ? 0 : nao;
      const int n = (int)(x2c ? nao : nso);
      block_gemm(stream_, n, nmat, dm + ro * nso + co, S, tmp, nso, nso2, 1.0, false);             // dm S
      block_gemm(stream_, n, nmat, S, tmp, F + ro * nso + co, nso, nso2, -pref * madelung, true);  // -c S (dm S)
    }
  }
  if (x2c)
    hipLaunchKernelGGL(hf_x2c_adjoint_kernel, dim3((unsigned)ceil_div(nk * nao2, 256)), dim3(256), 0, stream_, F, (int)nk,
                       (int)nao, (int)nso);
  GWHIP_CHECK(hipGetLastError());
  GWHIP_CHECK(hipEventRecord(e1, stream_));
  GWHIP_CHECK(hipStreamSynchronize(stream_));
  float ms = 0.f;
  GWHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  hf_ms_ = ms;
  hf_exchange_ms_ = ex_ms;
  hf_exchange_launches_ = nspin * nblk;
}

// The exchange products X[blk][q][k] = sum_Q V(k, k+q)_Q dm(k+q)_blk V(k, k+q)_Q^H as batched FP64
// GEMMs over k per q (any nao): Y = V_Q dm with the rows (Q, i) stacked, then X = sum_{Q, b} Y V^H
// with K = (Q, b) cut into slices (zgemm_f64_splitk).  dm_s: one spin's [nk][nso^2].
void gw_hip_kernel::hf_exchange_gemm(const double2* dm_s, int nblk) {
  const int64_t nk = d_.nk, NQ = d_.NQ, nao = d_.nao, nao2 = nao * nao, nso = d_.nso, nso2 = nso * nso;
  const int64_t vq = NQ * nao2;
  const int64_t tiles = nk * ceil_div(nao, 64) * ceil_div(nao, 64);
  const int ksplit = (int)std::min<int64_t>(kOrbKsplit, std::max<int64_t>(1, ceil_div(1024, tiles)));
  for (int64_t q = 0; q < nk; ++q) {
    gather_mats(dm_s, kplus_tab_ + q * nk, nk);  // dm(k + q)
    for (int bl = 0; bl < nblk; ++bl) {
      const int64_t ro = (bl == 1) ? nao : 0, co = (bl == 0) ? 0 : nao;  // aa, bb, ab
      ZGemmArgs y{};
      y.M = (int)(NQ * nao); y.N = (int)nao; y.K = (int)nao; y.batch = (int)nk; y.kin = (int)nao; y.nin = (int)nao;
      y.A = Vq_ + q * nk * vq; y.sAb = vq; y.sAm = nao; y.sAko = 0; y.sAki = 1;
      y.B = orb_G_ + ro * nso + co; y.sBb = nso2; y.sBko = 0; y.sBki = nso; y.sBno = 0; y.sBni = 1;
      y.C = orb_W_; y.sCb = vq; y.sCm = nao; y.sCno = 0; y.sCni = 1;
      y.alpha = make_double2(1.0, 0.0);
      zgemm_f64(y, true, false, stream_);
      ZGemmArgs x{};
      x.M = (int)nao; x.N = (int)nao; x.K = (int)(NQ * nao); x.batch = (int)nk; x.kin = (int)nao; x.nin = (int)nao;
      x.A = orb_W_; x.sAb = vq; x.sAm = nao; x.sAko = nao2; x.sAki = 1;
      x.B = Vq_ + q * nk * vq; x.sBb = vq; x.sBko = nao2; x.sBki = 1; x.sBno = 0; x.sBni = nao;
      x.conjB = 1;
      x.C = hf_x_ + (bl * nk + q) * nk * nao2; x.sCb = nao2; x.sCm = nao; x.sCno = 0; x.sCni = 1;
      x.alpha = make_double2(1.0, 0.0);
      zgemm_f64_splitk(x, ksplit, orb_part_, true, true, stream_);
    }
  }
}

void gw_hip_kernel::hf_host(const double2* dm, const double2* S, double madelung, double2* F) {
  if (!dm || !S || !F) throw status_error(GWHIP_ERR_ARG, "null dm, S or F");
  GWHIP_CHECK(hipSetDevice(device_));
  const size_t n = (size_t)d_.ns * d_.nk * d_.nso * d_.nso;
  double2* buf = dalloc<double2>(3 * n);
  try {
    GWHIP_CHECK(hipMemcpy(buf, dm, n * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(buf + n, S, n * sizeof(double2), hipMemcpyHostToDevice));
    hf_device(buf, buf + n, madelung, buf + 2 * n);
    GWHIP_CHECK(hipMemcpy(F, buf + 2 * n, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(buf);
    bytes_ -= 3 * n * sizeof(double2);
    throw;
  }
  (void)hipFree(buf);
  bytes_ -= 3 * n * sizeof(double2);
}

}  // namespace gwhip
