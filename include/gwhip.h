/*
 * gwhip -- MI355X (gfx950) GW self-energy kernel, C ABI.
 *
 * Drop-in for green-mbpt's GW hot path.  The reference's kernel-selection interface
 * (src/green/mbpt/kernel_factory.h:28-31,80-106) hands a custom kernel the solver's
 * dimensions, the imaginary-time transformer, the Brillouin-zone maps and the overlap,
 * and later calls  solve(G, Sigma)  (src/gw_cpu_kernel.cpp:31-72) with row-major
 * [nts, ns, ink, nso, nso] complex128 tensors.  This ABI exposes the same steps with
 * plain pointers and sizes; the C++ adapter in green-mbpt_amd/adapter/ maps the
 * reference's C++ types onto it.
 *
 * Complex numbers are interleaved (re, im) doubles.  All host buffers are owned by
 * the caller.  A handle owns one HIP stream and its device memory; calls on one handle
 * are not thread-safe.  Every function returns a status code (GWHIP_OK on success);
 * gwhip_last_error() gives the message of the last failure on a handle.
 */
#ifndef GWHIP_H
#define GWHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gwhip_handle gwhip_handle;

/* Status codes; the exception each mirrors is in src/green/mbpt/except.h:28-56. */
enum gwhip_status {
  GWHIP_OK = 0,
  GWHIP_ERR_ARG = 1,         /* bad dimensions or null pointer                            */
  GWHIP_ERR_WRONG_GRID = 2,  /* odd nts: mbpt_wrong_grid (src/gw_cpu_kernel.cpp:79-81)   */
  GWHIP_ERR_STATE = 3,       /* grid or integrals not set before eval                    */
  GWHIP_ERR_OOM = 4,         /* device allocation failed                                 */
  GWHIP_ERR_DEVICE = 5,      /* HIP runtime error or no usable device                    */
  GWHIP_ERR_NONFINITE = 6,   /* Sigma contains NaN/Inf                                   */
  GWHIP_ERR_UNSUPPORTED = 7, /* request outside what this build implements               */
  GWHIP_ERR_OUTDATED_INPUT = 8, /* integral files older than format 1.0.0: mbpt_outdated_input
                                   (src/green/mbpt/df_integral_t.h:42-51, except.h)          */
  GWHIP_ERR_IO = 9,          /* a file or dataset is missing or has the wrong shape        */
  GWHIP_ERR_MU_SEARCH = 10,  /* chemical-potential bisection stalled:
                                mbpt_chemical_potential_search_failure (src/dyson.cpp:148-150) */
  GWHIP_ERR_NOT_POSITIVE_DEFINITE = 11 /* some I - P0(i Omega_n, q) is not Hermitian positive definite
                                (a pivot of its factorisation is not positive): P0 is not the
                                negative semidefinite polarisation of a physical G.  The reference
                                factorises it with a pivoted LDLT (src/gw_cpu_kernel.cpp:245-250)
                                and carries on; this kernel reports it instead of returning a
                                degraded Sigma                                                  */
};

/* Arithmetic of the P0 and Sigma contractions (the Dyson step, transforms and
 * accumulation across k/q are FP64 in every mode). */
enum gwhip_precision {
  GWHIP_PREC_FP64 = 0,        /* complex128 throughout (reference default, P_sp = Sigma_sp = false) */
  GWHIP_PREC_SPLIT = 1,       /* split-precision P0 and Sigma contractions, FP64 accumulation
                                 (P_sp = Sigma_sp = true, src/green/mbpt/common_defs.h:162-163)   */
  GWHIP_PREC_SPLIT_P0 = 2,    /* split P0 contraction only (P_sp = true, Sigma_sp = false;
                                 the reference's flag at src/gw_cpu_kernel.cpp:96-100)            */
  GWHIP_PREC_SPLIT_SIGMA = 3  /* split Sigma contraction only (P_sp = false, Sigma_sp = true;
                                 src/gw_cpu_kernel.cpp:123-127)                                   */
};

/* Where the vector-Jacobian product takes P-tilde(tau, q) and P0(tau, q) from. */
enum gwhip_vjp_mode {
  GWHIP_VJP_RECOMPUTE = 0,   /* recompute them per q inside the VJP (no state kept)           */
  GWHIP_VJP_STORED = 1       /* use those of the last gwhip_eval_sigma*, which must have run
                                with gwhip_set_vjp_store(h, 1) on the same G                 */
};

/* Problem dimensions; names follow gw_cpu_kernel's members (src/green/mbpt/kernels.h:58-83).
 * The kernels run on the full k mesh (nq = nk); irreducible-wedge inputs (ink < nk) are unfolded
 * to it by the *_ibz entry points after gwhip_set_k_symmetry. */
typedef struct gwhip_dims {
  int64_t nk;        /* k points in the full Brillouin zone                          */
  int64_t mesh[3];   /* Monkhorst-Pack mesh, mesh[0]*mesh[1]*mesh[2] == nk;
                        k = (i0*mesh[1] + i1)*mesh[2] + i2, k+q componentwise mod mesh */
  int64_t nao;       /* atomic orbitals                                              */
  int64_t nso;       /* spin orbitals: nao, or 2*nao for X2C                         */
  int64_t ns;        /* spins (1 or 2; 1 for X2C)                                    */
  int64_t NQ;        /* auxiliary (density-fitting) basis size                       */
  int64_t nts;       /* fermionic tau nodes including 0 and beta (must be even)      */
  int64_t nw_b;      /* bosonic Matsubara frequencies                                */
  double beta;       /* inverse temperature                                          */
} gwhip_dims;

/* Create a kernel instance on HIP device `device`.
 * Replaces: gw_cpu_kernel::gw_cpu_kernel (src/green/mbpt/kernels.h:42-53) as returned by
 * gw_kernel_factory::get_kernel / CUSTOM_GW_KERNEL_CALL (src/green/mbpt/kernel_factory.h:28-31,85-91). */
int gwhip_create(const gwhip_dims* dims, int device, gwhip_handle** out);

/* Release device memory and the stream.  Replaces the kernel's shared_ptr<void> deleter
 * (src/green/mbpt/kernel_factory.h:89). */
void gwhip_destroy(gwhip_handle* h);

/* Message of the last failed call on h (static string for h == NULL). */
const char* gwhip_last_error(const gwhip_handle* h);

/* Replace the k +/- q maps derived from `mesh` by explicit tables [nq][nk]: kplus[q][k] =
 * k1_from_k2q(k, q) and kminus[q][k] = k2_from_k1q(k, q) (green-symmetry's k_q_map, used at
 * src/gw_cpu_kernel.cpp:90,117).  Every q must have a partner q' with kplus[q'] == kminus[q];
 * call before uploading or generating V. */
int gwhip_set_kq_maps(gwhip_handle* h, const int32_t* kplus, const int32_t* kminus);

/* Set the bosonic transforms: T_bt [nw_b][nts] maps P(tau_f) -> P(i Omega_b) and
 * T_tb [nts][nw_b] maps back; both complex.
 * Replaces: transformer_t::tau_f_to_w_b / w_b_to_tau_f as used at src/gw_cpu_kernel.cpp:234,264. */
int gwhip_set_grid(gwhip_handle* h, const double* T_bt, const double* T_tb);

/* Upload V^Q_ij(k1,k2) in the packed pair store of df_integral_t: [nk(nk+1)/2][NQ][nao][nao]
 * complex, pair index k1(k1+1)/2 + k2 for k1 >= k2; (k1 < k2) blocks are the per-Q adjoints.
 * V stays resident on the device across eval calls.
 * Replaces: df_integral_t::read_integrals + symmetrize (src/green/mbpt/df_integral_t.h:64-92,162-187). */
int gwhip_upload_V(gwhip_handle* h, const double* V_packed);

/* Fill the resident V on the device from the seeded counter-based generator
 * (SplitMix64 + Box-Muller, N(0, sigma^2) per real component, k1==k2 blocks Hermitised),
 * without a host copy.  Same counters as the Python generator gwhip.inputs.generate_V.
 * No reference counterpart: synthetic-input path for benchmarks. */
int gwhip_generate_V(gwhip_handle* h, uint64_t seed, double sigma);

/* Number of int8 slices per operand in the split-precision modes (3, 4 or 5; default 4): the P0
 * and Sigma contractions keep about 8 bits per slice of each operand relative to its row/column
 * maximum per K chunk (balanced radix-254 digits) and accumulate slice products exactly (int32),
 * then in FP64.
 * No reference counterpart (the reference's P_sp / Sigma_sp use complex<float>). */
int gwhip_set_split_slices(gwhip_handle* h, int slices);

/* Sigma(tau, s, k) from G(tau, s, k); host buffers [nts][ns][nk][nso][nso] complex.
 * Sigma is overwritten (the reference zeroes it first, src/gw_cpu_kernel.cpp:39-42).
 * Replaces: gw_cpu_kernel::solve (src/gw_cpu_kernel.cpp:31-72), reached via
 * gw_solver::solve (src/gw_solver.cpp:30-32). */
int gwhip_eval_sigma(gwhip_handle* h, const double* G, double* Sigma, int precision);

/* As gwhip_eval_sigma, with G and Sigma already in device memory (same layout).
 * Synchronous with respect to the host. */
int gwhip_eval_sigma_device(gwhip_handle* h, const void* G_dev, void* Sigma_dev, int precision);

/* Size of the device workspace of gwhip_eval_sigma_device_noalloc, per GPU: 256 MiB. */
#define GWHIP_NOALLOC_WORKSPACE_BYTES ((size_t)256 << 20)

/* As gwhip_eval_sigma_device, without allocating: no hipMalloc / hipFree or allocating library call
 * is made.  Every transient buffer of the evaluation (orbital scratch, Dyson work and panels, int8
 * slices of Z and P-tilde, Sigma epilogue partials, status flags) lives in one file-scope device
 * array of GWHIP_NOALLOC_WORKSPACE_BYTES per GPU, held from the first call on that GPU until the
 * process ends; the tau and q chunks shrink to fit it.  The handle's persistent state must exist:
 * grid and V (gwhip_set_grid, gwhip_upload_V / gwhip_generate_V) and, for the split modes, V's int8
 * slice stores, which the first gwhip_eval_sigma_device or gwhip_autotune in a split mode builds.
 * GWHIP_ERR_UNSUPPORTED when one tau point and one q do not fit the workspace (nk = 64, nao = 26,
 * naux = 300 needs about 0.64 GB in FP64 and does not), for nao > 48 (the orbital GEMM form) and
 * with the q -> 0 correction on.  Results are bit-identical to gwhip_eval_sigma_device.  All calls on one
 * GPU share its array: calls that overlap on one GPU (several host threads, several handles) race
 * and must be serialised by the caller.
 * No counterpart in the reference: for callers that cannot allocate device memory. */
int gwhip_eval_sigma_device_noalloc(gwhip_handle* h, const void* G_dev, void* Sigma_dev, int precision);

/* ---- launch-parameter autotuning ------------------------------------------------------
 * Time the candidate launch parameters of this system on its own operands (G in device memory)
 * and keep the fastest: the FP64 GEMM tiles of the P0 and Sigma contractions (0: 64x64,
 * 1: 128x64, 2: 64x128), the aux indices per wave of the P0 orbital kernel (1, 2, 4, 8) and, for
 * the split-precision modes, the block tile of each int8 GEMM, the number of row tiles that sweep
 * the column tiles together in the int8 GEMMs' L2-aware tile order (1 .. 64), the k chunks per
 * launch of the int8 P0 GEMM and the fused or separate slicing of the P0 operand; in both
 * precisions the form of the register Sigma epilogue.
 * Results change only by rounding (the epilogue forms and the k chunks add in another order).
 * No reference counterpart. */
int gwhip_autotune(gwhip_handle* h, const void* G_dev, int precision);
/* The tuning as an array of GWHIP_TUNING_COUNT ints: [0] FP64 P0 GEMM tile, [1] FP64 Sigma GEMM
 * tile, [2] aux indices per wave, [3] int8 P0 GEMM row-tile group, [4] int8 Sigma GEMM row-tile
 * group, [5] int8 P0 GEMM block tile and [6] int8 Sigma GEMM block tile (2: 32x64 of 16x16x64
 * MFMAs, A staged through registers; 4: 64x64 of 32x32x32 MFMAs; results bit-identical),
 * [7] split P0 orbital kernel writes the int8 slices itself (1) or via a separate pass (0),
 * [8] orbital-size products (Z = Gb V Gq of P0, V G R of Sigma) as batched FP64 GEMMs (1) or the
 * register-chained MFMA kernels (0); always 1 in effect for nao > 48, [9] int8 P0 GEMM: k chunks
 * per launch (0: all in one launch), so that each launch's operands stay in L2, [10] register
 * Sigma epilogue: 0 chosen by nao, 1 4x4x4 MFMAs with every 16-row group of i in each wave
 * (nao <= 32; otherwise as 0), 2 4x4x4 MFMAs with one 16-row group per wave, 3 16x16x4 MFMAs
 * (results agree to rounding). */
#define GWHIP_TUNING_COUNT 11
int gwhip_get_tuning(const gwhip_handle* h, int* values, int n);
int gwhip_set_tuning(gwhip_handle* h, const int* values, int n);

/* ---- several systems in one graph -------------------------------------------------------
 * A batch holds nsys independent systems (each its own dims, grid, V and G), one stream and one
 * transient workspace sized for the largest system.  gwhip_batch_eval_device evaluates Sigma of
 * every system; the first call for a given precision and set of G / Sigma pointers captures the
 * whole multi-system evaluation into a hipGraph, later calls replay it with one hipGraphLaunch.
 * Each system is configured through the handle gwhip_batch_system returns (gwhip_set_grid,
 * gwhip_upload_V / gwhip_generate_V, gwhip_set_kq_maps, gwhip_set_split_slices, gwhip_vjp...);
 * those handles belong to the batch (gwhip_destroy ignores them).  The reference evaluates one
 * system per gw_cpu_kernel (src/gw_cpu_kernel.cpp:31-72); this has no counterpart there. */
typedef struct gwhip_batch gwhip_batch;
int gwhip_batch_create(int nsys, const gwhip_dims* dims, int device, gwhip_batch** out);
void gwhip_batch_destroy(gwhip_batch* b);
const char* gwhip_batch_last_error(const gwhip_batch* b);
gwhip_handle* gwhip_batch_system(gwhip_batch* b, int i);
/* Autotune every system (G_dev[i] device pointers); drops captured graphs. */
int gwhip_batch_autotune(gwhip_batch* b, const void* const* G_dev, int precision);
/* Sigma_dev[i] = Sigma[G_dev[i]] for every system, device pointers, synchronous. */
int gwhip_batch_eval_device(gwhip_batch* b, const void* const* G_dev, void* const* Sigma_dev, int precision);
/* Milliseconds of the last graph launch, its node count, graphs captured so far, workspace bytes. */
int gwhip_batch_stats(const gwhip_batch* b, double* last_ms, int64_t* graph_nodes, int64_t* captures,
                      size_t* workspace_bytes);

/* Keep (keep = 1) or stop keeping (0) the half-grid P0(tau, q) of every q after each
 * evaluation, so that GWHIP_VJP_STORED can reuse P0 and P-tilde instead of recomputing them.
 * Costs nk * nts/2 * NQ^2 complex of device memory while enabled.
 * No reference counterpart (the reference has no derivative of Sigma). */
int gwhip_set_vjp_store(gwhip_handle* h, int keep);

/* Vector-Jacobian product of Sigma[G] (the map gw_cpu_kernel::solve computes,
 * src/gw_cpu_kernel.cpp:31-373) at G for the cotangent Sigma_bar: writes G_bar with
 * Re<Sigma_bar, dSigma> = Re<G_bar, dG> for every dG (<a,b> = sum conj(a) b; the convention of
 * torch.autograd for complex tensors).  Host buffers, all [nts][ns][nk][nso][nso] complex.
 * FP64 throughout.  `mode` is a gwhip_vjp_mode.  No reference counterpart. */
int gwhip_vjp(gwhip_handle* h, const double* G, const double* Sigma_bar, double* G_bar, int mode);

/* As gwhip_vjp with device buffers.  Synchronous with respect to the host. */
int gwhip_vjp_device(gwhip_handle* h, const void* G_dev, const void* Sigma_bar_dev, void* G_bar_dev, int mode);

/* Milliseconds of the last VJP (HIP events on the handle's stream). */
double gwhip_vjp_ms(const gwhip_handle* h);

/* Hartree-Fock Fock contribution from the handle's resident integrals (replaces
 * hf_scalar_cpu_kernel::solve and hf_x2c_cpu_kernel::solve, src/hf_cpu_kernels.cpp:24-146 and
 * :148-279, reached through the factory's HF hook, src/green/mbpt/kernel_factory.h:33-36):
 *   F_s(k) = sum_Q J[Q] V(k,k)_Q - pref/nk sum_{k'} sum_Q V(k,k')_Q dm_s(k') V(k,k')_Q^H
 *            - pref * madelung * S_s(k) dm_s(k) S_s(k),   J[Q] = 1/nk sum_{k',s} tr(V(k',k')_Q dm_s(k'))
 * pref = 1/2 for ns = 1 and 1 for ns = 2.  X2C (nso = 2 nao, ns = 1): spin blocks aa, bb, ab with
 * pref 1, the direct term from dm_aa + dm_bb into aa and bb, the Madelung term with S_aa on both
 * sides, F_ba = F_ab^H.  dm is the density matrix the reference forms (spin prefactor times
 * G(beta), src/hf_solver.cpp:9-16).  Host buffers dm, S, F: [ns][nk][nso][nso] complex. FP64. */
int gwhip_hf(gwhip_handle* h, const double* dm, const double* S, double madelung, double* F);

/* As gwhip_hf with device buffers.  Synchronous with respect to the host. */
int gwhip_hf_device(gwhip_handle* h, const void* dm_dev, const void* S_dev, double madelung, void* F_dev);

/* Milliseconds of the last gwhip_hf[_device] (HIP events on the handle's stream). */
double gwhip_hf_ms(const gwhip_handle* h);

/* Irreducible-wedge tables of green-symmetry's brillouin_zone_utils (input.h5:/symmetry/k), so
 * that G, Sigma, dm, S and F may be passed on the ink irreducible k points as the reference does
 * (G, Sigma: [nts][ns][ink][nso][nso]; dm, S, F: [ns][ink][nso][nso]):
 *   bz2ibz [nk]  full-mesh index of each k's representative;  ibz2bz [ink]  the representatives
 *   U [nk][nso][nso] complex  k_sym_transform_ao;  tr_conj [nk]  1 where time reversal applies.
 * An irreducible object is carried to k as X(k) = U_k op(X(ik)) U_k^H (op = conj when
 * tr_conj[k]), the identity at a representative: green-symmetry's value_AO as used at
 * src/gw_cpu_kernel.cpp:177-178,336 and src/hf_cpu_kernels.cpp:107. */
int gwhip_set_k_symmetry(gwhip_handle* h, int64_t ink, const int64_t* bz2ibz, const int64_t* ibz2bz, const double* U,
                         const int64_t* tr_conj);

/* The q symmetry of green-symmetry's brillouin_zone_utils (input.h5:/symmetry/q; the q mesh is
 * the k mesh): bz2ibz [nq], ibz2bz [inq] (the irreducible q), U [nq][NQ][NQ] complex
 * k_sym_transform_p0 (auxiliary-basis matrices) and tr_conj [nq].  A q of the star of an
 * irreducible q_ir is every q with bz2ibz[q] = q_ir, in ascending order.  P-tilde is carried from
 * q_ir to q as U_q P-tilde U_q^H, conjugated when tr_conj[q] = 1 (eval_p0_bz_from_ibz,
 * src/gw_cpu_kernel.cpp:281-293).  With both symmetries set, gwhip_eval_sigma_ibz runs the
 * reference's star-of-q evaluation.  inq = 0 switches it off. */
int gwhip_set_q_symmetry(gwhip_handle* h, int64_t inq, const int64_t* bz2ibz, const int64_t* ibz2bz, const double* U,
                         const int64_t* tr_conj);

/* Sigma on the irreducible k points from G on the irreducible k points: gw_cpu_kernel::solve with
 * symmetry (src/gw_cpu_kernel.cpp:31-134).  With the q symmetry set: the reference's scheme, P0
 * and P-tilde for the irreducible q over all k (G carried to the full mesh on the device), Sigma
 * for the irreducible k summed over the star of each irreducible q (:113-130).  Without it: the
 * full-mesh kernel runs on the unfolded G and the representatives' Sigma is returned (equal up to
 * how exactly the input obeys its symmetry).  Host buffers. */
int gwhip_eval_sigma_ibz(gwhip_handle* h, const double* G_ibz, double* Sigma_ibz, int precision);
int gwhip_eval_sigma_ibz_device(gwhip_handle* h, const void* G_ibz_dev, void* Sigma_ibz_dev, int precision);

/* Hartree-Fock on the irreducible k points (hf_scalar_cpu_kernel::solve with symmetry,
 * src/hf_cpu_kernels.cpp:24-146: dm is unfolded with value_AO, F returned at the
 * representatives).  dm, S, F: [ns][ink][nso][nso]. */
int gwhip_hf_ibz(gwhip_handle* h, const double* dm_ibz, const double* S_ibz, double madelung, double* F_ibz);
int gwhip_hf_ibz_device(gwhip_handle* h, const void* dm_ibz_dev, const void* S_ibz_dev, double madelung,
                        void* F_ibz_dev);

/* q -> 0 finite-size correction of Sigma (q0_treatment EXTRAPOLATE): mbpt_q0_utils_t
 * (src/mbpt_q0_utils_t.cpp:11-178) as gw_cpu_kernel::solve applies it (src/gw_cpu_kernel.cpp:59-64,
 * 266-270).  For each of the inq irreducible q (full-mesh index q_ir[iq]) the Dyson stage records
 * eps(n, iq) = conj(A_iq)^T P-tilde(i Omega_n, q) A_iq; after Sigma, per frequency, Re eps over the
 * distinct |q| in (0, 1] is fitted by a degree-2 polynomial (normal equations), the constant term
 * is carried to tau with the T_tb of gwhip_set_grid, and
 *   Sigma(t, s, k) -= madelung * eps_q0(t) * S(s, k) G(t, s, k) S(s, k)
 * (X2C: aa, bb, ab blocks with S_aa, ba = ab^H) on the k points the entry point returns.
 * AqQ [inq][NQ] complex, q_abs [inq] and madelung are AqQ.h5's; S [ns][nk_S][nso][nso] with
 * nk_S = nk for gwhip_eval_sigma or ink for gwhip_eval_sigma_ibz.  inq = 0 switches it off.
 * Needs at least three distinct |q| in (0, 1].  Not applied by the batch entry points. */
int gwhip_set_q0_extrapolation(gwhip_handle* h, int64_t inq, const int64_t* q_ir, const double* AqQ,
                               const double* q_abs, double madelung, int64_t nk_S, const double* S);

/* eps_q0 on the fermionic tau grid of the last evaluation with the correction on: [nts] complex. */
int gwhip_get_q0_eps(const gwhip_handle* h, double* eps_tau);

/* ---- Dyson step and chemical-potential search (dyson<G, S1, St>, src/dyson.cpp:19-236) ----
 * A separate handle: it needs the fermionic tau <-> i w_n transforms, not the bosonic ones.
 * Arrays on the irreducible k points as the reference keeps them: H, S, Sigma1 [ns][ink][nso][nso],
 * Sigma(tau), G(tau) [nts][ns][ink][nso][nso], complex. */
typedef struct gwhip_dyson gwhip_dyson;

typedef struct gwhip_dyson_dims {
  int64_t nts;   /* fermionic tau nodes                   */
  int64_t nw;    /* fermionic Matsubara frequencies       */
  int64_t ns;    /* spins (1 for X2C)                     */
  int64_t nk;    /* k points of the full mesh             */
  int64_t ink;   /* irreducible k points (nk without symmetry) */
  int64_t nao;   /* atomic orbitals                       */
  int64_t nso;   /* nao, or 2 nao for X2C                 */
} gwhip_dyson_dims;

/* Replaces: dyson<G, S1, St>::dyson (src/dyson.cpp:19-46) minus the file reading. */
int gwhip_dyson_create(const gwhip_dyson_dims* dims, int device, gwhip_dyson** out);
void gwhip_dyson_destroy(gwhip_dyson* d);
const char* gwhip_dyson_last_error(const gwhip_dyson* d);

/* green-grids' fermionic tau_to_omega / omega_to_tau as matrices: T_ft [nw][nts], T_tf [nts][nw]
 * complex, and the Matsubara frequencies w_n [nw] (real). */
int gwhip_dyson_set_grid(gwhip_dyson* d, const double* T_ft, const double* T_tf, const double* wn);

/* H and S (made Hermitian on the device, src/dyson.cpp:42-43) and the number of full-mesh points
 * each irreducible point stands for ([ink]; null: 1 each), used by the electron count. */
int gwhip_dyson_set_system(gwhip_dyson* d, const double* H, const double* S, const double* weights);

/* compute_G (src/dyson.cpp:174-236): G(tau) at mu.  Sigma1 and Sigma(tau) are made Hermitian in
 * place, as the reference does.  Host buffers; the _device variant takes device pointers. */
int gwhip_dyson_compute_G(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu, double* G_tau);
int gwhip_dyson_compute_G_device(gwhip_dyson* d, void* Sigma1_dev, void* Sigma_tau_dev, double mu, void* G_tau_dev);

/* find_mu (src/dyson.cpp:100-172) from mu0 to electron count nel within relative tolerance tol:
 * writes the chemical potential and the count found.  GWHIP_ERR_MU_SEARCH when the bisection stalls. */
int gwhip_dyson_find_mu(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu0, double nel, double tol,
                        double* mu, double* nel_found);
int gwhip_dyson_find_mu_device(gwhip_dyson* d, void* Sigma1_dev, void* Sigma_tau_dev, double mu0, double nel,
                               double tol, double* mu, double* nel_found);

/* compute_number_of_electrons (src/dyson.cpp:70-98) at mu, evaluated as
 * -pref Re sum_n T_tf[nts-1][n] sum_{s,k} (w_k / nk) tr(G(i w_n; mu) S). */
int gwhip_dyson_number_of_electrons(gwhip_dyson* d, double* Sigma1, double* Sigma_tau, double mu, double* nel);

/* compute_energy (src/common_utils.cpp:5-55, called by dyson::diff, src/dyson.cpp:290-309) on the
 * device: energies[0] = E_1b = nkpw sum_{s,k} w_k Re tr(dm H), energies[1] = E_HF = nkpw sum_{s,k}
 * w_k Re tr(dm (Sigma1 + 2 H)) / 2, energies[2] = E_corr (Galitskii-Migdal) = -pref nkpw Re
 * sum_n T_tf[nts-1][n] sum_{s,k} w_k tr(G(i w_n) Sigma(i w_n)), with dm = -(2 or 1) G(beta),
 * pref = 1 (ns = 1 without X2C) or 1/2, H and the weights of gwhip_dyson_set_system and the
 * fermionic transforms of gwhip_dyson_set_grid.  G, Sigma(tau) [nts][ns][ink][nso][nso],
 * Sigma1 [ns][ink][nso][nso]; read only (the reference does not Hermitise them here).  The
 * reference's Energy_HF record adds the nuclear repulsion to E_HF (src/dyson.cpp:326). */
int gwhip_dyson_energy(gwhip_dyson* d, const double* G_tau, const double* Sigma1, const double* Sigma_tau,
                       double* energies);
int gwhip_dyson_energy_device(gwhip_dyson* d, const void* G_tau, const void* Sigma1, const void* Sigma_tau,
                              double* energies);

/* The dimensions the handle was created with. */
int gwhip_get_dims(const gwhip_handle* h, gwhip_dims* out);

/* Of the last gwhip_hf[_device]: total milliseconds of its exchange-term launches (HIP events
 * around each) and their number (ns, or 3 spin blocks for X2C). */
int gwhip_hf_exchange_stats(const gwhip_handle* h, double* ms, int64_t* launches);

/* Copy P-tilde(tau, q) of the last eval to the host: [nk][nts][NQ][NQ] complex.
 * Exposes the intermediate the reference keeps in P0_tilde_s after eval_P_tilde
 * (src/gw_cpu_kernel.cpp:227-279); used by tests. */
int gwhip_get_P_tilde(gwhip_handle* h, double* P);

/* Keep (on = 1) P0(tau) of the first q of every evaluation, so that gwhip_get_leakage_inputs can
 * hand it out with that q's P-tilde: the two objects whose IR leakage the reference prints for
 * q_ir == 0 (ft.check_chebyshev + print_leakage, src/gw_cpu_kernel.cpp:237-239,276-278).
 * Costs nts/2 * NQ^2 complex of device memory while on. */
int gwhip_set_leakage_capture(gwhip_handle* h, int on);

/* P0(tau, q0) and P-tilde(tau, q0) of the last evaluation, q0 the first q evaluated (q = 0 on the
 * full mesh, the first irreducible q on a reduced one): host buffers [nts][NQ][NQ] complex, full
 * tau grid (P0 after Hermitisation, P0(beta - tau) = P0(tau)).  Either pointer may be null.
 * GWHIP_ERR_STATE unless the capture was on during that evaluation. */
int gwhip_get_leakage_inputs(gwhip_handle* h, double* P0, double* P_tilde);

/* Timings of the last eval, in milliseconds, for the region names of the reference's
 * statistics object (src/gw_cpu_kernel.cpp:43-68,205,233-274,363):
 *   0 "P0_zgemm" (P0 contraction), 1 "P0(t) -> P0(w)", 2 "GW-BSE" (Dyson),
 *   3 "P(w) -> P(t)", 4 "Selfenergy_zgemm" (Sigma contraction), 5 "total".
 * Measured with HIP events on the handle's stream.  Returns the number written. */
int gwhip_timings(const gwhip_handle* h, double* ms, int n);

/* Per-launch timing of one kernel family over the last eval:
 * which = 0 P0 big GEMM, 1 Sigma big GEMM, 2 P0 prep (G.V.G), 3 Sigma epilogue, 4 Dyson inverse.
 * Writes launch count and total milliseconds (HIP events around each launch). */
int gwhip_kernel_stats(const gwhip_handle* h, int which, int64_t* launches, double* total_ms);

/* Device bytes held by the handle. */
size_t gwhip_device_bytes(const gwhip_handle* h);

/* ---- test hooks (no reference counterpart; used by the parity tests) ---- */

/* C = opA(A) opB(B) with the library's FP64 MFMA GEMM on device 0; row-major host
 * arrays A [M][K], B [K][N], C [M][N] complex; op = conjugate when the flag is set;
 * shape selects the tile (0: 64x64, 1: 128x64, 2: 64x128). */
int gwhip_debug_zgemm(int M, int N, int K, const double* A, const double* B, double* C, int conjA, int conjB,
                      int shape);

/* C = A B with the split-precision (int8-slice) GEMM on device 0, `slices` slices per operand,
 * one K chunk, block tile `tile` (2: 32 x 64 of 16x16x64 MFMAs, 4: 64 x 64 of 32x32x32 MFMAs);
 * row-major host arrays as in gwhip_debug_zgemm. */
int gwhip_debug_zgemm_split(int M, int N, int K, const double* A, const double* B, double* C, int slices, int tile);

/* The stages of one q alone, for checking intermediates at full size: from G (host,
 * [nts][ns][nk][nso][nso]) compute P0(tau, q) (full tau grid after Hermitisation and the tau
 * mirror, [nts][NQ][NQ]), P-tilde(tau, q) ([nts][NQ][NQ]) and q's contribution to Sigma, including
 * the 1/nk of the full sum ([nts][ns][nk][nso][nso]) -- the intermediates of
 * selfenergy_innerloop(q) (src/gw_cpu_kernel.cpp:74-134).  Any output may be null. */
int gwhip_debug_eval_q(gwhip_handle* h, const double* G, int64_t q, int precision, double* P0, double* P_tilde,
                       double* Sigma_q);

/* Run an evaluation only up to stage `last_stage` (1: P0 on the half tau grid after
 * Hermitisation, 2: P-tilde on all tau) and leave the result where gwhip_get_P_tilde
 * reads it.  Sigma is not produced. */
int gwhip_debug_stage(gwhip_handle* h, const double* G, int last_stage);

#ifdef __cplusplus
}
#endif
#endif /* GWHIP_H */
