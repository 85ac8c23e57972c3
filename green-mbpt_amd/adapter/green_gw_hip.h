/*
 * green-mbpt custom-kernel adapter for the MI355X GW kernel (libgwhip).
 *
 * Registers through the reference's compile-time plug-in interface:
 *   cmake -DCUSTOM_KERNELS=<repo url> -DCUSTOM_KERNEL_HEADERS=green_gw_hip.h
 *         -DCUSTOM_KERNEL_ENUMS=HIP -DCUSTOM_KERNEL_NAMESPACES=green::mbpt::hip
 *         -DCUSTOM_KERNEL_LIBS=gwhip
 * (CMakeLists.txt:51-84, cmake/AddGreenKernel.cmake:1-24).  The factory macros then call the
 * three functions below (src/green/mbpt/kernel_factory.h:28-36, common_defs.h:169-180), and
 * `mbpt.exe --kernel HIP` selects it.
 *
 * Precision: the reference's own flags select the arithmetic of the two contractions --
 * --P_sp (polarisation P0) and --Sigma_sp (self-energy), src/green/mbpt/common_defs.h:162-163 --
 * each mapped onto the kernel's split-precision products for that contraction (GWHIP_PREC_SPLIT,
 * _SPLIT_P0, _SPLIT_SIGMA); the Dyson step, the transforms and all accumulation stay FP64.
 *
 * This header needs green-mbpt's own headers (params, grids, symmetry, utils, ndarray) and is
 * compiled inside a green-mbpt build (a CPU test compiles it against declaration stubs of those
 * types).  Everything device-side sits behind include/gwhip.h; the integral files and symmetry
 * tables are read by libgwhip_h5 (include/gwhip_h5.h), so link both gwhip and gwhip_h5.
 */
#ifndef GREEN_MBPT_GW_HIP_H
#define GREEN_MBPT_GW_HIP_H

#include <green/grids/transformer_t.h>
#include <green/mbpt/common_utils.h>
#include <green/mbpt/except.h>
#include <green/mbpt/kernels.h>
#include <green/params/params.h>
#include <green/symmetry/symmetry.h>
#include <green/utils/mpi_shared.h>

#include <functional>
#include <iostream>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "gwhip.h"
#include "gwhip_h5.h"

namespace green::mbpt::hip {

  using G_type     = utils::shared_object<ztensor<5>>;
  using bz_utils_t = symmetry::brillouin_zone_utils;

  /** Flags of this kernel (hook at src/green/mbpt/common_defs.h:169-180).  The precision is taken
   *  from the reference's --P_sp / --Sigma_sp, so the kernel defines no switch of its own for it. */
  inline void custom_kernel_parameters(params::params& p) {
    p.define<int>("hip_device", "HIP device used by the GW kernel", 0);
  }

  /** The kernel's precision mode for the reference's P_sp / Sigma_sp flags (common_defs.h:162-163:
   *  single-precision polarisation / self-energy contractions, gw_cpu_kernel.cpp:96-100,123-127). */
  inline int precision_mode(bool P_sp, bool Sigma_sp) {
    if (P_sp && Sigma_sp) return GWHIP_PREC_SPLIT;
    if (P_sp) return GWHIP_PREC_SPLIT_P0;
    if (Sigma_sp) return GWHIP_PREC_SPLIT_SIGMA;
    return GWHIP_PREC_FP64;
  }

  namespace detail {
    /** The GW kernel's handle and the integral directory it holds, so that the HF kernel reuses the
     *  resident integrals when it reads the same directory (node rank 0). */
    struct resident_t {
      std::weak_ptr<gwhip_handle> handle;
      std::string                 path;
    };
    inline resident_t& resident() {
      static resident_t r;
      return r;
    }

    inline void check(int st, const gwhip_handle* h) {
      if (st == GWHIP_OK) return;
      const std::string msg = h ? gwhip_last_error(h) : gwhip_h5_last_error();
      if (st == GWHIP_ERR_WRONG_GRID) throw mbpt_wrong_grid(msg);
      if (st == GWHIP_ERR_OUTDATED_INPUT) throw mbpt_outdated_input(msg);
      throw mbpt_kernel_error("HIP GW kernel: " + msg);
    }

    inline void require_one_node() {
      if (utils::context().internode_size != 1)
        throw mbpt_kernel_error("HIP GW kernel: one node (q is not split across nodes)");
    }

    // k + q and k - q exactly as the CPU kernels use them (gw_cpu_kernel.cpp:90,117)
    inline void set_maps(gwhip_handle* h, const bz_utils_t& bz) {
      const size_t nk = bz.nk();
      std::vector<int32_t> kp(nk * nk), km(nk * nk);
      for (size_t q = 0; q < nk; ++q)
        for (size_t k = 0; k < nk; ++k) {
          kp[q * nk + k] = bz.k_q_map().k1_from_k2q(k, q);
          km[q * nk + k] = bz.k_q_map().k2_from_k1q(k, q);
        }
      check(gwhip_set_kq_maps(h, kp.data(), km.data()), h);
    }

    // Integrals (every stored chunk, expanded to all k pairs, df_integral_t.h:40-187) and, on a
    // symmetry-reduced mesh, the k and q symmetry tables of the input file (so that
    // gwhip_eval_sigma_ibz runs the reference's star-of-q evaluation); then the reference's own
    // k +- q maps.
    inline void load_files(gwhip_handle* h, const bz_utils_t& bz, const std::string& df_path,
                           const std::string& input_file) {
      check(gwhip_h5_load(h, df_path.c_str(), input_file.c_str(), bz.ink() != bz.nk() ? 1 : 0), nullptr);
      set_maps(h, bz);
    }
  }  // namespace detail

  /** Owns one libgwhip handle with V resident on the device across SC iterations. */
  class gw_hip_kernel {
  public:
    gw_hip_kernel(bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ,
                  const grids::transformer_t& ft, const bz_utils_t& bz, const ztensor<4>& S_k) : _ft(ft), _bz(bz) {
      detail::require_one_node();
      // gw_solver sets X2C = (nao != nso) (src/green/mbpt/gw_solver.h:70); the kernel derives it the same way
      if (X2C != (nso != nao)) throw mbpt_kernel_error("HIP GW kernel: X2C flag disagrees with nso != nao");
      _ibz  = bz.ink() != bz.nk();
      _nts  = ft.sd().repn_fermi().nts();
      _nw_b = ft.sd().repn_bose().nw();
      gwhip_dims d{};
      d.nk   = bz.nk();
      d.mesh[0] = bz.nk(); d.mesh[1] = 1; d.mesh[2] = 1;   // replaced by the explicit maps below
      d.nao  = nao; d.nso = nso; d.ns = ns; d.NQ = NQ;
      d.nts  = _nts; d.nw_b = _nw_b;
      d.beta = p["BETA"];
      _NQ   = NQ;
      _prec = precision_mode(p["P_sp"], p["Sigma_sp"]);
      if (utils::context().node_rank == 0) {
        gwhip_handle* h = nullptr;
        check(gwhip_create(&d, p["hip_device"], &h), nullptr);
        _h.reset(h, gwhip_destroy);
        set_transforms();
        // P0 and P-tilde of the first q are kept for the reference's leakage print
        check(gwhip_set_leakage_capture(h, utils::context().global_rank == 0 ? 1 : 0), h);
        const std::string path = p["dfintegral_file"];
        detail::load_files(h, bz, path, p["input_file"]);
        detail::resident() = {_h, path};
        // q -> 0 treatment EXTRAPOLATE (kernels.h:48, mbpt_q0_utils_t.h:19-38): AqQ.h5 of the same
        // directory; S_k on the irreducible points, as the solver passes it
        const sigma_q0_treatment_e q0 = p["q0_treatment"];
        if (q0 == extrapolate) {
          int enabled = 0;
          const std::string input = p["input_file"];
          check(gwhip_h5_load_q0(h, path.c_str(), input.c_str(), bz.ink(), reinterpret_cast<const double*>(S_k.data()),
                                 &enabled),
                nullptr);
          if (!enabled)
            std::cout << "## Warning: " << path << "/AqQ.h5 is not found! Extrapolation treatment for GW self-energy will be disabled."
                      << std::endl;
        }
      }
    }

    /** Mirrors gw_cpu_kernel::solve (src/gw_cpu_kernel.cpp:31-72); collective over the node. */
    void solve(G_type& g, G_type& sigma_tau) {
      g.fence();
      sigma_tau.fence();
      if (utils::context().node_rank == 0) {
        const auto* G = reinterpret_cast<const double*>(g.object().data());
        auto*       S = reinterpret_cast<double*>(sigma_tau.object().data());
        check(_ibz ? gwhip_eval_sigma_ibz(_h.get(), G, S, _prec) : gwhip_eval_sigma(_h.get(), G, S, _prec), _h.get());
        if (utils::context().global_rank == 0) print_leakages();
      }
      sigma_tau.fence();
    }

  private:
    const grids::transformer_t& _ft;
    const bz_utils_t&           _bz;
    std::shared_ptr<gwhip_handle> _h;
    size_t _nts = 0, _nw_b = 0, _NQ = 0;
    int    _prec = GWHIP_PREC_FP64;
    bool   _ibz = false;

    static void check(int st, const gwhip_handle* h) { detail::check(st, h); }

    // The reference's leakage print for q_ir == 0 (gw_cpu_kernel.cpp:237-239,276-278): green-grids'
    // check_chebyshev on P0(tau) and on P-tilde(tau) of that q, through mbpt's print_leakage.
    void print_leakages() {
      ztensor<4> P0(_nts, 1, _NQ, _NQ), Pt(_nts, 1, _NQ, _NQ);
      check(gwhip_get_leakage_inputs(_h.get(), reinterpret_cast<double*>(P0.data()), reinterpret_cast<double*>(Pt.data())),
            _h.get());
      print_leakage(_ft.check_chebyshev(P0), "P0");
      print_leakage(_ft.check_chebyshev(Pt), "P");
    }

    // The transforms are linear maps per aux element: probe them with unit vectors so the
    // device applies exactly green-grids' tau_f_to_w_b / w_b_to_tau_f (gw_cpu_kernel.cpp:234,264).
    void set_transforms() {
      ztensor<4> ft_in(_nts, 1, 1, _nts), fw_out(_nw_b, 1, 1, _nts);
      for (size_t t = 0; t < _nts; ++t) ft_in(t, 0, 0, t) = 1.0;
      _ft.tau_f_to_w_b(ft_in, fw_out, 0, _nw_b, true);
      std::vector<std::complex<double>> T_bt(_nw_b * _nts), T_tb(_nts * _nw_b);
      for (size_t n = 0; n < _nw_b; ++n)
        for (size_t t = 0; t < _nts; ++t) T_bt[n * _nts + t] = fw_out(n, 0, 0, t);
      ztensor<4> fw_in(_nw_b, 1, 1, _nw_b), ft_out(_nts, 1, 1, _nw_b);
      for (size_t n = 0; n < _nw_b; ++n) fw_in(n, 0, 0, n) = 1.0;
      _ft.w_b_to_tau_f(fw_in, ft_out, 0, _nts, true);
      for (size_t t = 0; t < _nts; ++t)
        for (size_t n = 0; n < _nw_b; ++n) T_tb[t * _nw_b + n] = ft_out(t, 0, 0, n);
      check(gwhip_set_grid(_h.get(), reinterpret_cast<const double*>(T_bt.data()),
                           reinterpret_cast<const double*>(T_tb.data())),
            _h.get());
    }
  };

  /** GW factory hook (src/green/mbpt/kernel_factory.h:28-31). */
  inline std::tuple<std::shared_ptr<void>, std::function<void(G_type&, G_type&)>> custom_gw_kernel(
      bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, const grids::transformer_t& ft,
      const bz_utils_t& bz_utils, const ztensor<4>& S_k) {
    auto kernel = std::make_shared<gw_hip_kernel>(X2C, p, nao, nso, ns, NQ, ft, bz_utils, S_k);
    std::function<void(G_type&, G_type&)> callback = [kernel](G_type& g, G_type& s) { kernel->solve(g, s); };
    return std::tuple{std::static_pointer_cast<void>(kernel), callback};
  }

  /** Hartree-Fock on the device: the counterpart of hf_scalar_cpu_kernel / hf_x2c_cpu_kernel::solve
   *  (src/hf_cpu_kernels.cpp:24-279).  Uses the GW kernel's handle (V already resident) when one
   *  exists at the first solve and reads the same integral directory (dfintegral_hf_file ==
   *  dfintegral_file), else its own; the result is broadcast to every rank, as the CPU kernels'
   *  allreduce leaves it (hf_cpu_kernels.cpp:140). */
  class hf_hip_kernel {
  public:
    hf_hip_kernel(const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, double madelung,
                  const bz_utils_t& bz, const ztensor<4>& S_k) :
        _bz(bz), _S(S_k), _nao(nao), _nso(nso), _ns(ns), _NQ(NQ), _madelung(madelung),
        _device(p["hip_device"]), _path(p["dfintegral_hf_file"].as<std::string>()),
        _input(p["input_file"].as<std::string>()), _beta(p["BETA"]) {
      detail::require_one_node();
    }

    /** dm, S and the result on the irreducible k points: [ns][ink][nso][nso]. */
    ztensor<4> solve(const ztensor<4>& dm) {
      ztensor<4> F(_ns, _bz.ink(), _nso, _nso);
      if (utils::context().global_rank == 0) {
        gwhip_handle* h  = handle();
        const auto*   d  = reinterpret_cast<const double*>(dm.data());
        const auto*   S  = reinterpret_cast<const double*>(_S.data());
        auto*         Fp = reinterpret_cast<double*>(F.data());
        detail::check(_bz.ink() != _bz.nk() ? gwhip_hf_ibz(h, d, S, _madelung, Fp) : gwhip_hf(h, d, S, _madelung, Fp), h);
      }
      MPI_Bcast(F.data(), 2 * static_cast<int>(F.size()), MPI_DOUBLE, 0, utils::context().global);
      return F;
    }

  private:
    const bz_utils_t&             _bz;
    ztensor<4>                    _S;
    size_t                        _nao, _nso, _ns, _NQ;
    double                        _madelung;
    int                           _device;
    std::string                   _path, _input;
    double                        _beta;
    std::shared_ptr<gwhip_handle> _h;

    gwhip_handle* handle() {
      if (_h) return _h.get();
      if (detail::resident().path == _path) _h = detail::resident().handle.lock();
      if (_h) return _h.get();
      // no GW kernel on the same integrals: a handle of its own; HF reads no tau grid, so a minimal one
      gwhip_dims d{};
      d.nk = _bz.nk();
      d.mesh[0] = _bz.nk(); d.mesh[1] = 1; d.mesh[2] = 1;
      d.nao = _nao; d.nso = _nso; d.ns = _ns; d.NQ = _NQ;
      d.nts = 2; d.nw_b = 1; d.beta = _beta;
      gwhip_handle* h = nullptr;
      detail::check(gwhip_create(&d, _device, &h), nullptr);
      _h.reset(h, gwhip_destroy);
      detail::load_files(h, _bz, _path, _input);
      return h;
    }
  };

  /** HF factory hook (kernel_factory.h:33-36). */
  inline std::tuple<std::shared_ptr<void>, std::function<ztensor<4>(const ztensor<4>&)>> custom_hf_kernel(
      bool /*X2C*/, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, double madelung,
      const bz_utils_t& bz_utils, const ztensor<4>& S_k) {
    auto kernel = std::make_shared<hf_hip_kernel>(p, nao, nso, ns, NQ, madelung, bz_utils, S_k);
    std::function<ztensor<4>(const ztensor<4>&)> callback = [kernel](const ztensor<4>& dm) { return kernel->solve(dm); };
    return std::tuple{std::static_pointer_cast<void>(kernel), callback};
  }

}  // namespace green::mbpt::hip

#endif  // GREEN_MBPT_GW_HIP_H
