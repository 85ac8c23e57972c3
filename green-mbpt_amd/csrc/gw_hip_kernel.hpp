// gw_hip_kernel: the MI355X counterpart of green::mbpt::kernels::gw_cpu_kernel
// (src/green/mbpt/kernels.h:36-178, src/gw_cpu_kernel.cpp:31-373).
//
// Device-resident data: V for every ordered k pair, P0/P-tilde(tau, q) for every q, G and
// Sigma.  One evaluation runs three stages over all q on one HIP stream:
//   1. P0(tau, q) = -(pref/nk) sum_k Z_k(tau) V_k^H   for the half tau grid  (MFMA GEMM)
//      with Z_k = sum_s Gbar_s V_k G_s                                      (prep kernel)
//   2. Hermitise; P0(i Omega) = T_bt' P0(tau);  P-tilde = (I - P0)^-1 P0;  P-tilde(tau) = T_tb P-tilde(i Omega)
//   3. Sigma(tau, k) += -(1/nk) sum_{Q,n} (V G_s)[Q,i,n] R_k[Q,n,j],  R = P-tilde(tau, q) conj(V)
// V is held once, as V^Q_ab(k, k + q) for every (q, k); V(k, k - q) is the k + q' row of q' = -q.
// where T_bt' folds P0(beta - tau) = P0(tau) into the tau -> Omega matrix.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "common.hpp"

namespace gwhip {

struct Dims {
  int64_t nk = 0, nao = 0, nso = 0, ns = 0, NQ = 0, nts = 0, nw_b = 0;
  std::array<int64_t, 3> mesh{0, 0, 0};
  double beta = 0.0;
  // derived
  int64_t pns = 1, pseudo_ns = 1;
  bool X2C = false;
  double prefactor = 2.0;
};

struct KernelStat {
  int64_t launches = 0;
  double ms = 0.0;
};

// Launch parameters picked by the autotuner (defaults = the best measured at nk = 64, nao = 26,
// naux = 300): FP64 GEMM tile of the P0 and Sigma contractions (0: 64x64, 1: 128x64,
// 2: 64x128), the number of aux indices per wave of the P0 prep kernel, and for split precision
// the row-tile group of the int8 GEMM's L2-aware tile order and whether the P0 orbital kernel
// writes the int8 slices itself.
struct Tuning {
  int p0_shape = 0, s_shape = 0, qpw = 4;
  int p0_group = 1, s_group = 8;  // split precision: row tiles per L2 group of the int8 GEMMs
  int p0_oz_tile = 2, s_oz_tile = 2;  // int8 GEMM block tile: 2 (32 x 64, 16 x 16 x 64 MFMAs) or
                                      // 4 (64 x 64, 32 x 32 x 32 MFMAs, staged by LDS DMA)
  int p0_fused = 1;  // split precision, one or two collinear spin blocks: P0 orbital kernel writes int8
                     // slices directly
  int orb_gemm = 0;  // orbital contractions as batched FP64 GEMMs (always so for nao > 48)
  int p0_kchunks = 4;  // split precision: k chunks per launch of the int8 P0 GEMM (0: all in one), so
                       // that each launch's operand window stays in L2
  int epi_form = 0;  // register Sigma epilogue: 0 by nao, 1 4x4x4 all i groups per wave (nao <= 32),
                     // 2 4x4x4 one i group per wave, 3 16x16x4
  static constexpr int count = 11;
  void to_array(int* v) const {
    const int a[count] = {p0_shape, s_shape, qpw, p0_group, s_group, p0_oz_tile, s_oz_tile, p0_fused, orb_gemm,
                          p0_kchunks, epi_form};
    for (int i = 0; i < count; ++i) v[i] = a[i];
  }
  bool from_array(const int* v) {
    if (v[0] < 0 || v[0] > 2 || v[1] < 0 || v[1] > 2 || (v[2] != 1 && v[2] != 2 && v[2] != 4 && v[2] != 8) ||
        v[3] < 1 || v[4] < 1 || (v[5] != 2 && v[5] != 4) || (v[6] != 2 && v[6] != 4) || v[7] < 0 || v[7] > 1 || v[8] < 0 ||
        v[8] > 1 || v[9] < 0 || v[10] < 0 || v[10] > 3)
      return false;
    p0_shape = v[0]; s_shape = v[1]; qpw = v[2]; p0_group = v[3]; s_group = v[4]; p0_oz_tile = v[5]; s_oz_tile = v[6];
    p0_fused = v[7]; orb_gemm = v[8]; p0_kchunks = v[9]; epi_form = v[10];
    return true;
  }
};

// Transient device memory of one evaluation.  A stand-alone kernel owns its own; the kernels
// of a batch share one, sized for the largest system, and run one after another.
struct Workspace {
  double2* scratch = nullptr;  // Z or R, [tc][nk][NQ][nao^2]
  size_t scratch_elems = 0;
  double2* dyson = nullptr;    // Pw [qc][nw_b][NQ^2]
  size_t dyson_elems = 0;
  int8_t* slices = nullptr;    // int8 slices of Z (P0 stage) or P-tilde (Sigma stage)
  size_t slice_bytes = 0;
  int* exps = nullptr;         // their exponents
  size_t exp_elems = 0;
};

struct OrbOp;
struct OrbOut;
struct OrbArgs;

class gw_hip_kernel {
public:
  enum Region { R_P0 = 0, R_TW = 1, R_BSE = 2, R_WT = 3, R_SIGMA = 4, R_TOTAL = 5, R_COUNT = 6 };
  enum KStat { K_P0_GEMM = 0, K_SIGMA_GEMM = 1, K_P0_PREP = 2, K_SIGMA_EPI = 3, K_DYSON_INV = 4, K_COUNT = 5 };

  // shared_stream == nullptr: own stream and workspace.  Otherwise the kernel runs on the given
  // stream and waits for bind_workspace() before it can evaluate.
  gw_hip_kernel(const Dims& d, int device, hipStream_t shared_stream = nullptr);
  ~gw_hip_kernel();
  gw_hip_kernel(const gw_hip_kernel&) = delete;
  gw_hip_kernel& operator=(const gw_hip_kernel&) = delete;

  void set_grid(const double2* T_bt, const double2* T_tb);
  void set_kq_maps(const int32_t* kplus, const int32_t* kminus);
  void set_split_slices(int ns);
  void upload_V(const double2* V_packed);
  void generate_V(uint64_t seed, double sigma);
  // G_dev, Sigma_dev: device pointers [nts][ns][nk][nso][nso]
  // q0_apply: add the q -> 0 correction here (full-mesh entry); the irreducible entry applies it
  // after gathering instead
  void solve_device(const double2* G_dev, double2* Sigma_dev, int precision, bool q0_apply = true);
  void solve_host(const double2* G, double2* Sigma, int precision);
  // As solve_device, with every transient buffer (orbital scratch, Dyson work and panels, int8 slices
  // of Z and P-tilde, Sigma epilogue partials, status flags) carved from the device memory
  // [ws, ws + ws_bytes) instead of the handle's own workspace: no allocation, no timing events.
  // The tau and q chunks shrink to fit; GWHIP_ERR_UNSUPPORTED when one tau point and one q do not.
  // Needs the persistent stores ready (V; V's int8 slices for the split modes, built by an earlier
  // evaluation in a split mode); the q -> 0 correction and the orbital GEMM form are not run here.
  void solve_device_in(const double2* G_dev, double2* Sigma_dev, int precision, void* ws, size_t ws_bytes);
  // bytes solve_device_in needs at the smallest chunks (one tau point, one q) for `precision`
  size_t min_workspace_bytes(int precision) const;
  // test hook: stop after stage 1 (P0) or 2 (P-tilde); stage 3 runs Sigma as usual
  void run_stages_host(const double2* G, int last_stage);
  void copy_P_tilde(double2* host) const;
  // test hook: the stages of one q alone -- P0(tau, q) (full tau grid, after Hermitisation and the
  // tau mirror), P-tilde(tau, q) and q's contribution to Sigma (with the 1/nk of the full sum);
  // host buffers, any may be null
  void eval_one_q_host(const double2* G, int64_t q, int precision, double2* P0, double2* Pt, double2* Sq);
  // keep P0(tau) of the first q of each evaluation (the q_ir == 0 of the reference's leakage
  // print, src/gw_cpu_kernel.cpp:237-239,276-278); copy it and that q's P-tilde out (full tau grid)
  void set_leakage_capture(bool on);
  void copy_leakage_inputs(double2* P0, double2* Pt);
  // vector-Jacobian product of Sigma[G]: G_bar for the cotangent Sigma_bar (device pointers)
  void set_vjp_store(bool keep);
  void vjp_device(const double2* G_dev, const double2* Sbar_dev, double2* Gbar_dev, int mode);
  void vjp_host(const double2* G, const double2* Sbar, double2* Gbar, int mode);
  double vjp_ms() const { return vjp_ms_; }
  // Hartree-Fock Fock contribution from the resident integrals (device / host pointers):
  // dm, S, F [ns][nk][nso][nso]
  void hf_device(const double2* dm, const double2* S, double madelung, double2* F);
  void hf_host(const double2* dm, const double2* S, double madelung, double2* F);
  double hf_ms() const { return hf_ms_; }
  // Irreducible-wedge inputs/outputs: G_ibz, Sigma_ibz [nts][ns][ink][nso][nso]; dm, S, F
  // [ns][ink][nso][nso].  set_k_symmetry takes green-symmetry's tables (U: [nk][nso][nso]).
  void set_k_symmetry(int64_t ink, const int64_t* bz2ibz, const int64_t* ibz2bz, const double2* U,
                      const int64_t* tr_conj);
  int64_t ink() const { return ink_; }
  // q symmetry (green-symmetry's q_symmetry(), input.h5:/symmetry/q): inq irreducible q, bz2ibz
  // [nq], ibz2bz [inq], the auxiliary-basis matrices U [nq][NQ][NQ] (k_sym_transform_p0) and
  // tr_conj [nq].  With it (and the k symmetry) the irreducible entry runs the reference's
  // star-of-q evaluation.  inq = 0 switches it off.
  void set_q_symmetry(int64_t inq, const int64_t* bz2ibz, const int64_t* ibz2bz, const double2* U,
                      const int64_t* tr_conj);
  int64_t inq() const { return inq_; }
  // q -> 0 extrapolation (q0_treatment EXTRAPOLATE): inq irreducible q (full-mesh indices q_ir),
  // AqQ [inq][NQ], q_abs [inq], the Madelung constant of AqQ.h5 and the overlap S [ns][nkS][nso^2]
  // with nkS = nk (full-mesh entry) or ink (irreducible entry).  inq = 0 switches it off.
  void set_q0_extrapolation(int64_t inq, const int64_t* q_ir, const double2* AqQ, const double* q_abs,
                            double madelung, int64_t nkS, const double2* S);
  bool q0_enabled() const { return q0_on_; }
  const std::vector<double2>& q0_eps_tau() const { return q0_eps_t_; }
  void solve_ibz_device(const double2* G_ibz, double2* Sigma_ibz, int precision);
  void solve_ibz_host(const double2* G_ibz, double2* Sigma_ibz, int precision);
  void hf_ibz_device(const double2* dm_ibz, const double2* S_ibz, double madelung, double2* F_ibz);
  void hf_ibz_host(const double2* dm_ibz, const double2* S_ibz, double madelung, double2* F_ibz);
  double hf_exchange_ms() const { return hf_exchange_ms_; }
  int64_t hf_exchange_launches() const { return hf_exchange_launches_; }

  // batch support: workspace needs per tau point / per q, binding, and capture-safe enqueueing
  size_t scratch_per_t() const;
  size_t dyson_per_q() const;
  size_t slice_bytes_per_t() const;
  size_t exps_per_t() const;
  void bind_workspace(const Workspace& ws);
  void prepare_split();  // persistent slice stores for GWHIP_PREC_SPLIT
  bool split_ready() const { return slices_ns_ == split_ns_ && Vs_p0_ != nullptr; }
  // Enqueue one evaluation on the stream: no host synchronisation, no memset nodes, no
  // allocation (safe under stream capture once prepare_split / bind_workspace are done).
  void enqueue_eval(const double2* G_dev, double2* Sigma_dev, int precision, int* nonfinite_flag);
  void check_ready(int precision) const;
  // Whether G is Hermitian per (tau, spin, k) to 1e-12 relative; when it is, only the upper block
  // tiles of P0 are computed and mirrored.  Runs on the stream and synchronises.
  bool detect_hermitian_G(const double2* G_dev);
  void autotune(const double2* G_dev, int precision);
  const Tuning& tuning() const { return tune_; }
  void set_tuning(const Tuning& t) {
    tune_ = t;
    ++generation_;
  }
  void set_timing(bool on) { timing_ = on; }
  // buffers of the GEMM form of the orbital contractions and the Sigma epilogue's partial sums
  // over Q ranges (full mesh); allocates, so call before a capture
  void ensure_orb_buffers();
  // the register-chained orbital kernels handle nao up to 3 x 16
  static constexpr int64_t kOrbMaxNao = 48;
  bool has_inputs() const { return grid_set_ && v_set_; }
  // after the stream has drained: whether the last evaluation's Dyson step met a matrix I - P0
  // that is not positive definite
  bool dyson_failed() const;
  uint64_t generation() const { return generation_; }
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  const Dims& dims() const { return d_; }
  const double* region_ms() const { return region_ms_; }
  const KernelStat& kstat(int which) const { return kstat_[which]; }
  size_t device_bytes() const { return bytes_; }

private:
  Dims d_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true, own_ws_ = true, timing_ = true;
  struct Pending {
    int which;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending_;  // kernel timings read once the stream has drained
  Tuning tune_;
  uint64_t generation_ = 0;  // bumped whenever buffers or launch parameters change
  bool g_herm_ = false;      // G of the evaluation being enqueued is Hermitian
  unsigned long long* hcheck_ = nullptr;
  size_t bytes_ = 0;
  int max_lds_ = 65536;
  bool grid_set_ = false, v_set_ = false;

  // device buffers
  double2* Vq_ = nullptr;      // [nq][nk][NQ][nao^2]   V^Q_ab(k, k+q): the one resident copy of V
  double2* P_ = nullptr;       // [nq][nts][NQ^2]
  double2* Pw_ = nullptr;      // [qc][nw_b][NQ^2]: P0, then I - P0, its inverse and P-tilde on the bosonic grid
  double2* scratch_ = nullptr; // Z or R: [tc][nk][NQ][nao^2]; the Dyson stage's panels in between
  int64_t scratch_elems_ = 0;
  double2* Tbt_fold_ = nullptr;  // [nw_b][nts/2]
  double2* Ttb_ = nullptr;       // [nts][nw_b]
  double2* G_dev_ = nullptr;   // host-path staging
  double2* S_dev_ = nullptr;
  int* kplus_tab_ = nullptr;   // [nk][nq] : k + q
  int* kminus_tab_ = nullptr;  // [nq][nk] : k - q  (row q)
  int* flag_ = nullptr;
  int* dyson_flag_ = nullptr;  // set by the Dyson inverse when a pivot is not positive
  // split precision: int8 slice stores of conj(V) for the P0 and Sigma contractions, and the
  // per-evaluation slices of Z and P-tilde (zgemm_ozaki.hpp)
  int split_ns_ = 4, slices_ns_ = 0;
  int64_t kc_p0_ = 0, kc_s_ = 0;
  int8_t* Vs_p0_ = nullptr;   // [nq][NQ][nk][2][ns][kc_p0]
  int* eVs_p0_ = nullptr;     // [nq][NQ][nk]
  int8_t* Vs_s_ = nullptr;    // [nq][nk][nao^2][2][ns][kc_s]
  int* eVs_s_ = nullptr;      // [nq][nk][nao^2]
  int8_t* Zs_ = nullptr;      // [tc][NQ][nk][2][ns][kc_p0]  (P0 stage)
  int* eZs_ = nullptr;
  int8_t* Ps_ = nullptr;      // [tc][NQ][2][ns][kc_s]       (Sigma stage; same memory as Zs_)
  int* ePs_ = nullptr;
  int64_t qchunk_ = 1, tchunk_p0_ = 1, tchunk_sigma_ = 1;
  // VJP: per-q buffers, a second orbital scratch, and the kept half-grid P0 of the last
  // forward pass (stored mode)
  double2 *vP0h_ = nullptr, *vP0w_ = nullptr, *vBinv_ = nullptr, *vAinv_ = nullptr;
  double2 *vM1_ = nullptr, *vM2_ = nullptr, *vM3_ = nullptr, *vPt_ = nullptr, *vPtbar_ = nullptr, *vPbar_ = nullptr;
  double2* vW_ = nullptr;      // [tc][nk][NQ][nao^2]
  double2* P0keep_ = nullptr;  // [nq][nts/2][NQ^2]
  double2* leak_P0_ = nullptr; // [nts/2][NQ^2]: half-grid P0 of the first q of the last evaluation
  bool leak_on_ = false, leak_valid_ = false;
  void capture_leakage_p0();
  int64_t tchunk_vjp_ = 1;
  bool keep_p0_ = false, stored_valid_ = false;
  double vjp_ms_ = 0.0;
  double2* hf_work_ = nullptr;  // Hartree-Fock: J[Q], then S dm products [ns][nk][nso^2]
  double2* hf_x_ = nullptr;     // Hartree-Fock exchange per (q, k): [blocks][nq][nk][nao^2]
  double hf_ms_ = 0.0, hf_exchange_ms_ = 0.0;
  // k symmetry: [IBZ ordinal | representative flag | time reversal | ibz2bz] ints, U per k, and
  // the full-mesh dm, S, F and scratch of the irreducible HF entry
  int64_t ink_ = 0;
  // q -> 0 extrapolation
  bool q0_on_ = false, q0_eps_ = false;
  int64_t q0_inq_ = 0, q0_nS_ = 0;
  double q0_mad_ = 0.0;
  std::vector<double> q0_qabs_;
  std::vector<double2> q0_eps_t_, Ttb_host_;
  int* q0_qiq_ = nullptr;           // [nk] row of AqQ for an irreducible q, else -1
  double2* q0_A_ = nullptr;         // [inq][NQ]
  double2* q0_eps_dev_ = nullptr;   // [nw_b][inq]
  double2* q0_S_ = nullptr;         // [ns][nkS][nso^2]
  double2* q0_tmp_ = nullptr;       // G S, [nts][ns][nkS][nso^2]
  double2* q0_epst_ = nullptr;      // [nts]
  double2* orb_W_ = nullptr;  // GEMM form of the orbital contractions: [nk][NQ][nao^2] intermediate
  double2* orb_G_ = nullptr;  // and one (tau, spin) slice of G at mapped k points, [nk][nso^2]
  double2* orb_part_ = nullptr;  // split-K partial Sigma blocks, [kOrbKsplit][nk][nao^2]
  double2* orb_O_ = nullptr;     // one reduction result per k before it is added to its output, [nk][nao^2]
  double2* epi_part_ = nullptr;  // Sigma epilogue per Q range, [qsplit][nts][ns][nks][nso^2]
  int64_t epi_part_elems_ = 0;
  static constexpr int64_t kOrbKsplit = 64;
  int* ksym_idx_ = nullptr;
  double2* ksym_U_ = nullptr;
  std::vector<int> ksym_i2b_;       // host copy of ibz2bz
  // q symmetry and the star-of-q evaluation: irreducible q, stars, whether U_q is exactly the
  // identity, U_q on the device, and per-evaluation buffers (rotated P-tilde, the gathered
  // V(k, k - q) of the irreducible k and its slices, k - q of the irreducible k per q, AqQ rows)
  int64_t inq_ = 0;
  std::vector<int> qsym_i2b_, qsym_tr_, qsym_ident_, qs_q0row_host_;
  std::vector<std::vector<int>> qstar_;
  double2* qsym_U_ = nullptr;       // [nq][NQ^2]
  double2* star_P_ = nullptr;       // [2][nts/2][NQ^2]: rotated P-tilde and U P-tilde
  double2* star_V_ = nullptr;       // [ink][NQ][nao^2]
  int8_t* star_Vs_ = nullptr;       // [ink][nao^2][2][ns][kc_s]
  int* star_eVs_ = nullptr;         // [ink][nao^2]
  int* star_km_ = nullptr;          // [nq][ink]
  int* star_q0row_ = nullptr;       // [inq]
  int64_t star_ink_ = 0, star_vs_bytes_ = 0, star_evs_elems_ = 0;
  bool star_dirty_ = true;
  double2* hf_sym_ = nullptr;
  int64_t hf_exchange_launches_ = 0;
  std::vector<int> kplus_host_, kminus_host_, neg_q_, q0_qiq_host_;

  double region_ms_[R_COUNT] = {0};
  KernelStat kstat_[K_COUNT];

  template <typename T>
  T* dalloc(size_t count);
  struct WsPlan;
  WsPlan ws_plan(int64_t tc, int64_t qc, bool split) const;
  void build_maps();
  void upload_maps(const std::vector<int>& kpq, const std::vector<int>& kmq);
  void v_changed();  // V was (re)written: drop derived stores
  void ensure_slices();
  void free_slices();
  void ensure_vjp_buffers();
  void free_vjp_buffers();
  void zero(double2* X, int64_t n);
  void unfold(const double2* X, double2* T, double2* Y, int64_t nb);
  void gather_ibz(const double2* Full, double2* Ibz, int64_t nb);
  std::vector<int> q0_samples() const;
  void apply_q0(double2* Sigma, const double2* G, int64_t nkS);
  void p0_prep(int64_t q, const double2* G, int64_t t0, int64_t tc, int qpw);
  // orbital contractions as batched GEMMs (nao > 48, or tuning orb_gemm)
  bool orb_gemm() const { return d_.nao > kOrbMaxNao || tune_.orb_gemm != 0; }
  void p0_prep_gemm(int64_t q, const double2* G, int64_t t0, int64_t tc);
  void sigma_epi_gemm(const double2* G, double2* Sigma, const double2* Vb, const int* km, int64_t nkb, int64_t t0,
                      int64_t tc, double scale);
  void gather_g(const double2* G, int64_t tau, int64_t is, const int* map, int64_t n);
  void gather_mats(const double2* base, const int* map, int64_t n);
  void hf_exchange_gemm(const double2* dm_s, int nblk);
  // the VJP's orbital products: register-chained kernels, or batched GEMMs (orb_gemm())
  void vjp_sandwich(bool has_l, bool has_r, int NB, dim3 grid, const OrbArgs& a, const OrbOp& L, const OrbOp& V,
                    const OrbOp& R, const OrbOut& X);
  void vjp_reduce(bool has_m, int NB, dim3 grid, const OrbArgs& a, const OrbOp& A, const OrbOp& M, const OrbOp& C,
                  const OrbOut& O);
  void orb_gather_view(const OrbArgs& a, const OrbOp& op, int tz, int64_t n);
  void p0_prep_slice(int64_t q, const double2* G, int64_t t0, int64_t tc, int qpw);
  void p0_gemm(int64_t q, int64_t tc, double2* Pt0, int shape);
  void sigma_gemm(const double2* Pt0, const double2* Vb, int64_t nkb, int64_t tc, int shape);
  void p0_slice(int64_t tc);
  void p0_gemm_split(int64_t q, int64_t tc, double2* Pt0, int group, int kchunks);
  void sigma_slice(const double2* Pt0, int64_t tc);
  void sigma_gemm_split(const int8_t* Vsb, const int* eVsb, int64_t nkb, int64_t tc, int group);
  void stage_p0(const double2* G, bool split, const std::vector<int>* qs = nullptr);
  void p0_one_q(int64_t q, const double2* G, bool split, double2* Pq);
  // in-place inverse of nmat Hermitian positive-definite NQ x NQ matrices by block Gauss-Jordan steps
  // whose updates are batched FP64 MFMA GEMMs (panels in scratch_)
  void block_inverse(double2* X, int64_t nmat);
  void check_dyson_flag();  // reads dyson_flag_ (synchronous) and throws when it is set
  void stage_dyson(int64_t nq, const int* q0_row);
  void stage_sigma(const double2* G, double2* Sigma, bool split);
  void sigma_epi(const double2* G, double2* Sigma, const double2* Vb, const int* km, int64_t nkb, int64_t t0,
                 int64_t tc, int form);
  void sigma_q(const double2* G, double2* Sigma, bool split, const double2* Pq, const double2* Vb, const int8_t* Vsb,
               const int* eVsb, const int* km, int64_t nkb);
  const double2* Vq_neg(int64_t q) const;  // V(k, k - q) for every k: the k + q' rows of q' = -q
  const int8_t* Vs_neg(int64_t q) const;
  const int* eVs_neg(int64_t q) const;
  bool sigma_upper() const;
  int epi_form() const;
  bool fused_p0() const;
  // Q ranges of the register-chained Sigma epilogue for nks k points: 1 when nks x nts work-groups
  // fill the chip, up to 4 otherwise; ensure_epi_part sizes epi_part_ for it
  int epi_qsplit(int64_t nks) const;
  void ensure_epi_part(int64_t nks);
  void finish_sigma(double2* Sigma, int64_t nkS);
  void star_tables();
  const double2* star_rotate(int64_t iq, int64_t q);
  void solve_star_device(const double2* G_ibz, double2* Sigma_ibz, int precision);
  template <typename F>
  void timed(int kstat, F&& launch);
  void resolve_timings();
};

}  // namespace gwhip
