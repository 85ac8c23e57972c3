// Dyson step and chemical-potential search on the device (included by gw_hip_kernel.hip after
// gw_q0.hpp; one translation unit).  Counterpart of dyson<G, S1, St>::compute_G, find_mu and
// compute_number_of_electrons (src/dyson.cpp:70-236):
//   Sigma1, Sigma(tau) made Hermitian in place (make_hermitian, :183-188)
//   Sigma(i w_n) = T_ft Sigma(tau)                                  one FP64 GEMM over all (s, k)
//   G(i w_n) = [(i w_n + mu) S - H - Sigma1 - Sigma(i w_n)]^-1       dyson_invert_kernel, one
//                                                                   workgroup per matrix, LDS
//                                                                   Gauss-Jordan with row pivoting
//   G(tau) = T_tf G(i w_n)                                          one FP64 GEMM
// The electron count N(mu) = -pref Re sum_n T_tf[nts-1][n] sum_{s,k} (w_k / nk) tr(G(i w_n; mu) S)
// is the reference's eigenvalue sum sum_i 1 / (i w_n + mu - eps_i) written as a trace, so every
// trial mu costs one batched inverse instead of an eigen-decomposition up front; the bisection
// is the reference's, step for step.
#pragma once

namespace gwhip {

struct DysonDims {
  int64_t nts = 0, nw = 0, ns = 0, nk = 0, ink = 0, nao = 0, nso = 0;
};

// Largest nso whose matrix the inverse keeps in LDS; larger ones are eliminated in global memory.
constexpr int DYSON_MAX_N = 96;

// M = z S - HS1 - Sw with z = i w_n + mu; TRACE = false: write M^-1; TRACE = true: write
// wk * tr(M^-1 S).  grid (nsk, nw).  GLOB = false: the matrix sits in LDS (dynamic LDS n^2 + 2n
// double2 + n int); GLOB = true (any n): it is eliminated in place in global memory, in `out`
// itself (TRACE = false) or in `work` (TRACE = true), both [nw][nsk][n^2] (dynamic LDS 2n double2
// + n int).  Same arithmetic in both.
template <bool TRACE, bool GLOB>
__global__ __launch_bounds__(256) void dyson_invert_kernel(const double2* __restrict__ HS1, const double2* __restrict__ S,
                                                           const double2* __restrict__ Sw, const double* __restrict__ wn,
                                                           double mu, int n, int nsk, int ink,
                                                           const double* __restrict__ wk, double2* __restrict__ out,
                                                           double2* __restrict__ work) {
  extern __shared__ double2 dl[];
  const int64_t mat = ((int64_t)blockIdx.y * nsk + blockIdx.x) * n * n;
  double2* A = GLOB ? (TRACE ? work : out) + mat : dl;
  double2* colp = GLOB ? dl : dl + n * n;
  double2* rowp = colp + n;
  int* perm = reinterpret_cast<int*>(rowp + n);
  __shared__ int piv;
  __shared__ double red[2][256];
  const int sk = blockIdx.x, w = blockIdx.y, tid = threadIdx.x;
  const int n2 = n * n;
  const double2* s = S + (int64_t)sk * n2;
  const double2* h = HS1 + (int64_t)sk * n2;
  const double2* x = Sw + ((int64_t)w * nsk + sk) * n2;
  const double zr = mu, zi = wn[w];
  for (int e = tid; e < n2; e += blockDim.x) {
    const double2 sv = s[e], hv = h[e], xv = x[e];
    A[e] = make_double2(zr * sv.x - zi * sv.y - hv.x - xv.x, zr * sv.y + zi * sv.x - hv.y - xv.y);
  }
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    if (tid < 64) {  // pivot: largest |A[i][p]|, i >= p
      double best = -1.0;
      int bi = p;
      for (int i = p + tid; i < n; i += 64) {
        const double2 a = A[i * n + p];
        const double m = a.x * a.x + a.y * a.y;
        if (m > best) best = m, bi = i;
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o);
        const int oi = __shfl_xor(bi, o);
        if (ob > best || (ob == best && oi < bi)) best = ob, bi = oi;
      }
      if (tid == 0) {
        piv = bi;
        perm[p] = bi;
      }
    }
    __syncthreads();
    const int r = piv;
    if (r != p)
      for (int j = tid; j < n; j += blockDim.x) {
        const double2 t = A[p * n + j];
        A[p * n + j] = A[r * n + j];
        A[r * n + j] = t;
      }
    __syncthreads();
    const double2 pv = A[p * n + p];
    const double den = pv.x * pv.x + pv.y * pv.y;
    const double2 inv = make_double2(pv.x / den, -pv.y / den);
    for (int j = tid; j < n; j += blockDim.x) rowp[j] = j == p ? inv : cmul(A[p * n + j], inv);
    for (int i = tid; i < n; i += blockDim.x) colp[i] = i == p ? make_double2(0.0, 0.0) : A[i * n + p];
    __syncthreads();
    for (int e = tid; e < n2; e += blockDim.x) {
      const int i = e / n, j = e - (e / n) * n;
      if (i == p) {
        A[e] = rowp[j];
      } else {
        const double2 base = j == p ? make_double2(0.0, 0.0) : A[e];
        const double2 c = colp[i], rv = rowp[j];
        A[e] = make_double2(base.x - (c.x * rv.x - c.y * rv.y), base.y - (c.x * rv.y + c.y * rv.x));
      }
    }
    __syncthreads();
  }
  for (int p = n - 1; p >= 0; --p) {  // undo the row interchanges on the columns
    const int r = perm[p];
    if (r != p)
      for (int i = tid; i < n; i += blockDim.x) {
        const double2 t = A[i * n + p];
        A[i * n + p] = A[i * n + r];
        A[i * n + r] = t;
      }
    __syncthreads();
  }
  if (!TRACE) {
    if (!GLOB) {
      double2* o = out + ((int64_t)w * nsk + sk) * n2;
      for (int e = tid; e < n2; e += blockDim.x) o[e] = A[e];
    }
    return;
  }
  double tr = 0.0, ti = 0.0;  // tr(A S) = sum_ij A_ij S_ji
  for (int e = tid; e < n2; e += blockDim.x) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 a = A[e], b = s[j * n + i];
    tr += a.x * b.x - a.y * b.y;
    ti += a.x * b.y + a.y * b.x;
  }
  red[0][tid] = tr;
  red[1][tid] = ti;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) {
      red[0][tid] += red[0][tid + o];
      red[1][tid] += red[1][tid + o];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double f = wk[sk % ink];
    out[(int64_t)w * nsk + sk] = make_double2(f * red[0][0], f * red[1][0]);
  }
}

// out[m] = w[m % ink] * tr(A_m B_m) = w sum_ij A_m[i][j] B_m[j][i] for nmat n x n matrices
// (A_m at A + m*n^2, B_m at B + (m % nb)*n^2); one work-group per matrix, fixed-order reduction.
__global__ __launch_bounds__(256) void trace_prod_kernel(const double2* __restrict__ A, const double2* __restrict__ B,
                                                         int n, int nb, int ink, const double* __restrict__ w,
                                                         double2* __restrict__ out) {
  __shared__ double red[2][256];
  const int64_t m = blockIdx.x;
  const int n2 = n * n, tid = threadIdx.x;
  const double2* a = A + m * n2;
  const double2* b = B + (m % nb) * n2;
  double tr = 0.0, ti = 0.0;
  for (int e = tid; e < n2; e += blockDim.x) {
    const int i = e / n, j = e - (e / n) * n;
    const double2 x = a[e], y = b[j * n + i];
    tr += x.x * y.x - x.y * y.y;
    ti += x.x * y.y + x.y * y.x;
  }
  red[0][tid] = tr;
  red[1][tid] = ti;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (tid < o) {
      red[0][tid] += red[0][tid + o];
      red[1][tid] += red[1][tid + o];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const double f = w[(m % nb) % ink];
    out[m] = make_double2(f * red[0][0], f * red[1][0]);
  }
}

// C = A + B elementwise (n elements).
__global__ void add2_kernel(const double2* __restrict__ A, const double2* __restrict__ B, double2* __restrict__ C, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    C[e] = make_double2(A[e].x + B[e].x, A[e].y + B[e].y);
}

class gw_dyson {
public:
  gw_dyson(const DysonDims& d, int device) : d_(d), device_(device) {
    if (d.nts < 2 || d.nw < 1 || (d.ns != 1 && d.ns != 2) || d.nk < 1 || d.ink < 1 || d.ink > d.nk || d.nao < 1 ||
        (d.nso != d.nao && d.nso != 2 * d.nao))
      throw status_error(GWHIP_ERR_ARG, "bad Dyson dimensions");
    if (d.nso == 2 * d.nao && d.ns != 1) throw status_error(GWHIP_ERR_ARG, "X2C requires ns == 1");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
      (void)hipGetLastError();
      throw status_error(GWHIP_ERR_DEVICE, "no usable HIP device");
    }
    GWHIP_CHECK(hipSetDevice(device_));
    GWHIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    try {
      allocate();
    } catch (...) {
      release();
      throw;
    }
  }
  ~gw_dyson() { release(); }
  gw_dyson(const gw_dyson&) = delete;
  gw_dyson& operator=(const gw_dyson&) = delete;

  void allocate() {
    const DysonDims& d = d_;
    const int64_t nsk = d.ns * d.ink, n2 = d.nso * d.nso;
    const size_t big = (size_t)std::max(d.nts, d.nw) * nsk * n2;
    Tft_ = alloc<double2>((size_t)d.nw * d.nts);
    Ttf_ = alloc<double2>((size_t)d.nts * d.nw);
    wn_ = alloc<double>((size_t)d.nw);
    wk_ = alloc<double>((size_t)d.ink);
    H_ = alloc<double2>((size_t)nsk * n2);
    S_ = alloc<double2>((size_t)nsk * n2);
    HS1_ = alloc<double2>((size_t)nsk * n2);
    Sw_ = alloc<double2>(big);
    Gw_ = alloc<double2>(big);
    tr_ = alloc<double2>((size_t)d.nw * nsk);
    in_ = alloc<double2>((size_t)(d.nts + 1) * nsk * n2);
    std::vector<double> ones(d.ink, 1.0 / (double)d.nk);
    GWHIP_CHECK(hipMemcpy(wk_, ones.data(), ones.size() * sizeof(double), hipMemcpyHostToDevice));
  }

  void release() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (void* p : {(void*)Tft_, (void*)Ttf_, (void*)wn_, (void*)wk_, (void*)H_, (void*)S_, (void*)HS1_, (void*)Sw_,
                    (void*)Gw_, (void*)tr_, (void*)in_, (void*)etr_, (void*)eG_})
      if (p) (void)hipFree(p);
    Tft_ = Ttf_ = H_ = S_ = HS1_ = Sw_ = Gw_ = tr_ = in_ = etr_ = eG_ = nullptr;
    wn_ = wk_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
  }

  const DysonDims& dims() const { return d_; }

  void set_grid(const double2* T_ft, const double2* T_tf, const double* wn) {
    if (!T_ft || !T_tf || !wn) throw status_error(GWHIP_ERR_ARG, "null transform");
    GWHIP_CHECK(hipSetDevice(device_));
    GWHIP_CHECK(hipMemcpy(Tft_, T_ft, (size_t)d_.nw * d_.nts * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(Ttf_, T_tf, (size_t)d_.nts * d_.nw * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(wn_, wn, (size_t)d_.nw * sizeof(double), hipMemcpyHostToDevice));
    TtB_.assign(T_tf + (size_t)(d_.nts - 1) * d_.nw, T_tf + (size_t)d_.nts * d_.nw);
    grid_set_ = true;
  }

  // H, S [ns][ink][nso][nso] (made Hermitian, as the reference's constructor does, :42-43);
  // weights [ink]: mesh points per irreducible point (null: one each).
  void set_system(const double2* H, const double2* S, const double* weights) {
    if (!H || !S) throw status_error(GWHIP_ERR_ARG, "null H or S");
    GWHIP_CHECK(hipSetDevice(device_));
    const int64_t nsk = d_.ns * d_.ink, n = d_.nso;
    GWHIP_CHECK(hipMemcpy(H_, H, (size_t)nsk * n * n * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(S_, S, (size_t)nsk * n * n * sizeof(double2), hipMemcpyHostToDevice));
    hermitize(H_, nsk);
    hermitize(S_, nsk);
    std::vector<double> w(d_.ink, 1.0 / (double)d_.nk);
    if (weights)
      for (int64_t k = 0; k < d_.ink; ++k) w[k] = weights[k] / (double)d_.nk;
    GWHIP_CHECK(hipMemcpy(wk_, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    sys_set_ = true;
  }

  // G(tau) from Sigma1, Sigma(tau) at mu (device pointers); Sigma1 and Sigma(tau) are made
  // Hermitian in place.
  void compute_G_device(double2* Sigma1, double2* Sigma_tau, double mu, double2* G) {
    if (!Sigma1 || !Sigma_tau || !G) throw status_error(GWHIP_ERR_ARG, "null Sigma1, Sigma or G");
    prepare(Sigma1, Sigma_tau);
    invert(mu, false);
    stack_transform(stream_, Ttf_, (int)d_.nts, (int)d_.nw, false, Gw_, nsk_n2(), G);
    GWHIP_CHECK(hipStreamSynchronize(stream_));
  }

  // Electron count at mu for the Sigma of the last prepare (device pointers; Hermitised in place).
  double number_of_electrons_device(double2* Sigma1, double2* Sigma_tau, double mu) {
    prepare(Sigma1, Sigma_tau);
    return count(mu);
  }

  // find_mu (src/dyson.cpp:100-172) from mu0; returns (N at the result, mu).
  std::pair<double, double> find_mu_device(double2* Sigma1, double2* Sigma_tau, double mu0, double nel_target,
                                           double tol) {
    if (!(nel_target > 0.0) || !(tol > 0.0)) throw status_error(GWHIP_ERR_ARG, "nel and tol must be positive");
    prepare(Sigma1, Sigma_tau);
    double mu = mu0, delta = 0.5;
    double nel = count(mu), nel1 = 0.0, nel2 = 0.0, mu1 = 0.0, mu2 = 0.0;
    auto off = [&](double x) { return std::fabs((x - nel_target) / nel_target) > tol; };
    if (off(nel)) {
      if (nel > nel_target) {
        mu1 = mu - delta;
        do {
          nel1 = count(mu1);
          mu1 -= delta;
        } while (nel1 > nel_target);
        mu2 = mu;
        nel2 = nel;
      } else {
        mu2 = mu + delta;
        do {
          nel2 = count(mu2);
          mu2 += delta;
        } while (nel2 < nel_target);
        mu1 = mu;
        nel1 = nel;
      }
      while (off(nel) && std::fabs(mu2 - mu1) > 0.01 * tol) {
        mu = 0.5 * (mu1 + mu2);
        const double nel_old = count(mu);
        if (off(nel_old) && std::fabs(nel - nel_old) < 0.001 * tol)
          throw status_error(GWHIP_ERR_MU_SEARCH, "Chemical potential search failed.");
        nel = nel_old;
        if (nel > nel_target) mu2 = mu;
        else mu1 = mu;
      }
    }
    (void)nel1;
    (void)nel2;
    return {nel, mu};
  }

  // compute_energy (src/common_utils.cpp:5-55) of G(tau), Sigma1, Sigma(tau) (device pointers, read
  // only) with the handle's H and k weights: {E_1b, E_HF, E_corr}.  G and Sigma go to the Matsubara
  // axis by one FP64 GEMM each; per (w, s, k) the trace of G(i w) Sigma(i w) and per (s, k) those of
  // dm H and dm Sigma1 (dm = -(2 or 1) G(beta)) are reduced on the device, the sums over (s, k) and
  // frequencies on the host in a fixed order.
  std::array<double, 3> energy_device(const double2* G, const double2* Sigma1, const double2* Sigma_tau) {
    if (!grid_set_ || !sys_set_) throw status_error(GWHIP_ERR_STATE, "grid or H/S not set");
    if (!G || !Sigma1 || !Sigma_tau) throw status_error(GWHIP_ERR_ARG, "null G, Sigma1 or Sigma");
    GWHIP_CHECK(hipSetDevice(device_));
    const int64_t nsk = d_.ns * d_.ink, N = nsk_n2(), n = d_.nso;
    stack_transform(stream_, Tft_, (int)d_.nw, (int)d_.nts, false, G, N, Gw_);
    stack_transform(stream_, Tft_, (int)d_.nw, (int)d_.nts, false, Sigma_tau, N, Sw_);
    if (d_.nw * nsk > 0x7fffffff) throw status_error(GWHIP_ERR_UNSUPPORTED, "too many (w, s, k) matrices");
    hipLaunchKernelGGL(trace_prod_kernel, dim3((unsigned)(d_.nw * nsk)), dim3(256), 0, stream_, Gw_, Sw_, (int)n,
                       (int)(d_.nw * nsk), (int)d_.ink, wk_, tr_);
    GWHIP_CHECK(hipGetLastError());
    // tr(G(beta) H) and tr(G(beta) Sigma1) per (s, k), after the nw * nsk traces of tr_
    const double2* Gb = G + (d_.nts - 1) * N;
    if (!etr_) etr_ = alloc<double2>((size_t)2 * nsk);
    hipLaunchKernelGGL(trace_prod_kernel, dim3((unsigned)nsk), dim3(256), 0, stream_, Gb, H_, (int)n, (int)nsk,
                       (int)d_.ink, wk_, etr_);
    hipLaunchKernelGGL(trace_prod_kernel, dim3((unsigned)nsk), dim3(256), 0, stream_, Gb, Sigma1, (int)n, (int)nsk,
                       (int)d_.ink, wk_, etr_ + nsk);
    GWHIP_CHECK(hipGetLastError());
    std::vector<double2> tr((size_t)d_.nw * nsk), et((size_t)2 * nsk);
    GWHIP_CHECK(hipMemcpyAsync(tr.data(), tr_, tr.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    GWHIP_CHECK(hipMemcpyAsync(et.data(), etr_, et.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    const bool x2c = d_.nso == 2 * d_.nao;
    const double cdm = (d_.ns == 2 || x2c) ? -1.0 : -2.0;     // dm = cdm G(beta)
    const double pref = (d_.ns == 1 && !x2c) ? 1.0 : 0.5;     // energy_prefactor
    double gs_r = 0.0;  // Re sum_w T_tf[beta][w] sum_sk w_k tr(G(i w) Sigma(i w))
    for (int64_t w = 0; w < d_.nw; ++w) {
      double sr = 0.0, si = 0.0;
      for (int64_t sk = 0; sk < nsk; ++sk) {
        sr += tr[(size_t)w * nsk + sk].x;
        si += tr[(size_t)w * nsk + sk].y;
      }
      gs_r += TtB_[w].x * sr - TtB_[w].y * si;
    }
    double e1 = 0.0, ehf = 0.0;
    for (int64_t sk = 0; sk < nsk; ++sk) {
      const double tH = cdm * et[sk].x, tS1 = cdm * et[nsk + sk].x;
      e1 += tH;
      ehf += 0.5 * (tS1 + 2.0 * tH);
    }
    return {e1, ehf, -gs_r * pref};
  }

  std::array<double, 3> energy_host(const double2* G, const double2* Sigma1, const double2* Sigma_tau) {
    upload(Sigma1, Sigma_tau);
    const int64_t N = nsk_n2();
    double2* g = nullptr;
    if (!eG_) eG_ = alloc<double2>((size_t)d_.nts * N);
    g = eG_;
    GWHIP_CHECK(hipMemcpy(g, G, (size_t)d_.nts * N * sizeof(double2), hipMemcpyHostToDevice));
    return energy_device(g, in_ + d_.nts * N, in_);
  }

  // Host-buffer variants: Sigma1 [ns][ink][nso^2], Sigma(tau), G [nts][ns][ink][nso^2]; the
  // Hermitised Sigma1 and Sigma(tau) are written back to the caller's arrays.
  void compute_G_host(double2* Sigma1, double2* Sigma_tau, double mu, double2* G) {
    upload(Sigma1, Sigma_tau);
    compute_G_device(in_ + d_.nts * nsk_n2(), in_, mu, Gw_scratch_out());
    download(Sigma1, Sigma_tau, G);
  }
  std::pair<double, double> find_mu_host(double2* Sigma1, double2* Sigma_tau, double mu0, double nel, double tol) {
    upload(Sigma1, Sigma_tau);
    auto r = find_mu_device(in_ + d_.nts * nsk_n2(), in_, mu0, nel, tol);
    download(Sigma1, Sigma_tau, nullptr);
    return r;
  }
  double number_of_electrons_host(double2* Sigma1, double2* Sigma_tau, double mu) {
    upload(Sigma1, Sigma_tau);
    const double r = number_of_electrons_device(in_ + d_.nts * nsk_n2(), in_, mu);
    download(Sigma1, Sigma_tau, nullptr);
    return r;
  }

private:
  DysonDims d_;
  int device_;
  hipStream_t stream_ = nullptr;
  bool grid_set_ = false, sys_set_ = false;
  double2 *Tft_ = nullptr, *Ttf_ = nullptr, *H_ = nullptr, *S_ = nullptr, *HS1_ = nullptr, *Sw_ = nullptr,
          *Gw_ = nullptr, *tr_ = nullptr, *in_ = nullptr, *etr_ = nullptr, *eG_ = nullptr;
  double *wn_ = nullptr, *wk_ = nullptr;
  std::vector<double2> TtB_;  // row nts-1 of T_tf (tau = beta)

  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      throw status_error(GWHIP_ERR_OOM, "hipMalloc failed in the Dyson solver");
    }
    return static_cast<T*>(p);
  }
  int64_t nsk_n2() const { return d_.ns * d_.ink * d_.nso * d_.nso; }
  // the host G is staged through Sw_ (free once G(i w_n) is formed)
  double2* Gw_scratch_out() { return Sw_; }

  void hermitize(double2* X, int64_t nmat) {
    const int64_t n = d_.nso;
    for (int64_t m0 = 0; m0 < nmat; m0 += 65535) {
      const int64_t mc = std::min<int64_t>(65535, nmat - m0);
      ::gwhip::hermitize(X + m0 * n * n, (int)n, n * n, mc, stream_);
      GWHIP_CHECK(hipGetLastError());
    }
  }

  void prepare(double2* Sigma1, double2* Sigma_tau) {
    if (!grid_set_ || !sys_set_) throw status_error(GWHIP_ERR_STATE, "grid or H/S not set");
    if (!Sigma1 || !Sigma_tau) throw status_error(GWHIP_ERR_ARG, "null Sigma1 or Sigma");
    GWHIP_CHECK(hipSetDevice(device_));
    const int64_t nsk = d_.ns * d_.ink;
    hermitize(Sigma1, nsk);
    hermitize(Sigma_tau, d_.nts * nsk);
    const int64_t N = nsk_n2();
    hipLaunchKernelGGL(add2_kernel, dim3((unsigned)std::min<int64_t>(4096, ceil_div(N, 256))), dim3(256), 0, stream_,
                       H_, Sigma1, HS1_, N);
    GWHIP_CHECK(hipGetLastError());
    stack_transform(stream_, Tft_, (int)d_.nw, (int)d_.nts, false, Sigma_tau, N, Sw_);
  }

  template <bool TRACE, bool GLOB>
  void launch_invert_t(double mu, double2* out) {
    const int n = (int)d_.nso;
    const size_t lds = (size_t)((GLOB ? 0 : n * n) + 2 * n) * sizeof(double2) + (size_t)n * sizeof(int);
    if (lds > 65536) GWHIP_CHECK(hipFuncSetAttribute((const void*)dyson_invert_kernel<TRACE, GLOB>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int nsk = (int)(d_.ns * d_.ink);
    if (d_.nw > 65535 || nsk > 0x7fffffff) throw status_error(GWHIP_ERR_UNSUPPORTED, "too many Dyson matrices");
    // the global-memory form works in Gw_ ([nw][nsk][n^2]), which TRACE mode does not otherwise use
    hipLaunchKernelGGL((dyson_invert_kernel<TRACE, GLOB>), dim3((unsigned)nsk, (unsigned)d_.nw), dim3(256), lds, stream_,
                       HS1_, S_, Sw_, wn_, mu, n, nsk, (int)d_.ink, wk_, out, Gw_);
    GWHIP_CHECK(hipGetLastError());
  }
  template <bool TRACE>
  void launch_invert(double mu, double2* out) {
    if (d_.nso <= DYSON_MAX_N) launch_invert_t<TRACE, false>(mu, out);
    else launch_invert_t<TRACE, true>(mu, out);
  }

  void invert(double mu, bool trace) {
    if (trace) launch_invert<true>(mu, tr_);
    else launch_invert<false>(mu, Gw_);
  }

  double count(double mu) {
    invert(mu, true);
    const int64_t nsk = d_.ns * d_.ink;
    std::vector<double2> tr((size_t)d_.nw * nsk);
    GWHIP_CHECK(hipMemcpyAsync(tr.data(), tr_, tr.size() * sizeof(double2), hipMemcpyDeviceToHost, stream_));
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    double re = 0.0;
    for (int64_t w = 0; w < d_.nw; ++w) {
      double sr = 0.0, si = 0.0;
      for (int64_t sk = 0; sk < nsk; ++sk) {
        sr += tr[(size_t)w * nsk + sk].x;
        si += tr[(size_t)w * nsk + sk].y;
      }
      re += TtB_[w].x * sr - TtB_[w].y * si;
    }
    const double pref = (d_.ns == 2 || d_.nso == 2 * d_.nao) ? 1.0 : 2.0;
    return pref * -re;
  }

  void upload(const double2* Sigma1, const double2* Sigma_tau) {
    if (!Sigma1 || !Sigma_tau) throw status_error(GWHIP_ERR_ARG, "null Sigma1 or Sigma");
    GWHIP_CHECK(hipSetDevice(device_));
    const int64_t N = nsk_n2();
    GWHIP_CHECK(hipMemcpy(in_, Sigma_tau, (size_t)d_.nts * N * sizeof(double2), hipMemcpyHostToDevice));
    GWHIP_CHECK(hipMemcpy(in_ + d_.nts * N, Sigma1, (size_t)N * sizeof(double2), hipMemcpyHostToDevice));
  }
  void download(double2* Sigma1, double2* Sigma_tau, double2* G) {
    const int64_t N = nsk_n2();
    GWHIP_CHECK(hipStreamSynchronize(stream_));
    GWHIP_CHECK(hipMemcpy(Sigma_tau, in_, (size_t)d_.nts * N * sizeof(double2), hipMemcpyDeviceToHost));
    GWHIP_CHECK(hipMemcpy(Sigma1, in_ + d_.nts * N, (size_t)N * sizeof(double2), hipMemcpyDeviceToHost));
    if (G) GWHIP_CHECK(hipMemcpy(G, Gw_scratch_out(), (size_t)d_.nts * N * sizeof(double2), hipMemcpyDeviceToHost));
  }
};

}  // namespace gwhip
