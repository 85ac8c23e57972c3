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
 * This header needs green-mbpt's own headers (params, grids, symmetry, utils, ndarray) and is
 * compiled only inside a green-mbpt build.  Everything device-side sits behind include/gwhip.h.
 */
#ifndef GREEN_MBPT_GW_HIP_H
#define GREEN_MBPT_GW_HIP_H

#include <green/grids/transformer_t.h>
#include <green/mbpt/except.h>
#include <green/mbpt/kernels.h>
#include <green/params/params.h>
#include <green/symmetry/symmetry.h>
#include <green/utils/mpi_shared.h>

#include <functional>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "gwhip.h"

namespace green::mbpt::hip {

  using G_type     = utils::shared_object<ztensor<5>>;
  using bz_utils_t = symmetry::brillouin_zone_utils;

  /** Flags of this kernel (hook at src/green/mbpt/common_defs.h:169-180). */
  inline void custom_kernel_parameters(params::params& p) {
    p.define<int>("hip_device", "HIP device used by the GW kernel", 0);
    p.define<bool>("hip_split_precision", "Split-precision P0/Sigma contractions in the HIP GW kernel", false);
  }

  /** Owns one libgwhip handle with V resident on the device across SC iterations. */
  class gw_hip_kernel {
  public:
    gw_hip_kernel(bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ,
                  const grids::transformer_t& ft, const bz_utils_t& bz) : _ft(ft), _bz(bz) {
      if (bz.ink() != bz.nk() || bz.inq() != bz.nq())
        throw mbpt_kernel_error("HIP GW kernel: symmetry-reduced k/q meshes are not supported");
      if (utils::context().internode_size != 1)
        throw mbpt_kernel_error("HIP GW kernel: one node (q is not split across nodes)");
      _nts  = ft.sd().repn_fermi().nts();
      _nw_b = ft.sd().repn_bose().nw();
      gwhip_dims d{};
      d.nk   = bz.nk();
      d.mesh[0] = bz.nk(); d.mesh[1] = 1; d.mesh[2] = 1;   // replaced by the explicit maps below
      d.nao  = nao; d.nso = nso; d.ns = ns; d.NQ = NQ;
      d.nts  = _nts; d.nw_b = _nw_b;
      d.beta = p["BETA"];
      _split = p["hip_split_precision"];
      if (utils::context().node_rank == 0) {
        gwhip_handle* h = nullptr;
        check(gwhip_create(&d, p["hip_device"], &h), nullptr);
        _h.reset(h, gwhip_destroy);
        set_maps();
        set_transforms();
        upload_integrals(p["dfintegral_file"], nao, NQ);
      }
    }

    /** Mirrors gw_cpu_kernel::solve (src/gw_cpu_kernel.cpp:31-72); collective over the node. */
    void solve(G_type& g, G_type& sigma_tau) {
      g.fence();
      sigma_tau.fence();
      if (utils::context().node_rank == 0) {
        const auto* G = reinterpret_cast<const double*>(g.object().data());
        auto*       S = reinterpret_cast<double*>(sigma_tau.object().data());
        check(gwhip_eval_sigma(_h.get(), G, S, _split ? GWHIP_PREC_SPLIT : GWHIP_PREC_FP64), _h.get());
      }
      sigma_tau.fence();
    }

  private:
    const grids::transformer_t& _ft;
    const bz_utils_t&           _bz;
    std::shared_ptr<gwhip_handle> _h;
    size_t _nts = 0, _nw_b = 0;
    bool   _split = false;

    static void check(int st, const gwhip_handle* h) {
      if (st == GWHIP_OK) return;
      const std::string msg = gwhip_last_error(h);
      if (st == GWHIP_ERR_WRONG_GRID) throw mbpt_wrong_grid(msg);
      throw mbpt_kernel_error("HIP GW kernel: " + msg);
    }

    // k + q and k - q exactly as the CPU kernel uses them (gw_cpu_kernel.cpp:90,117)
    void set_maps() {
      const size_t nk = _bz.nk();
      std::vector<int32_t> kp(nk * nk), km(nk * nk);
      for (size_t q = 0; q < nk; ++q)
        for (size_t k = 0; k < nk; ++k) {
          kp[q * nk + k] = _bz.k_q_map().k1_from_k2q(k, q);
          km[q * nk + k] = _bz.k_q_map().k2_from_k1q(k, q);
        }
      check(gwhip_set_kq_maps(_h.get(), kp.data(), km.data()), _h.get());
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

    // Read every stored k pair through the reference's own reader (df_integral_t.h:64-187)
    // into the packed layout the device expands.
    void upload_integrals(const std::string& path, size_t nao, size_t NQ) {
      const size_t nk = _bz.nk(), npair = nk * (nk + 1) / 2;
      std::vector<std::complex<double>> Vp(npair * NQ * nao * nao);
      df_integral_t reader(path, nao, NQ, _bz, utils::mpi_context(MPI_COMM_SELF));
      tensor<std::complex<double>, 3> v(NQ, nao, nao);
      for (size_t k1 = 0; k1 < nk; ++k1)
        for (size_t k2 = 0; k2 <= k1; ++k2) {
          reader.read_integrals(k1, k2);
          reader.symmetrize(v, k1, k2);
          std::copy(v.data(), v.data() + v.size(), Vp.data() + (k1 * (k1 + 1) / 2 + k2) * NQ * nao * nao);
        }
      check(gwhip_upload_V(_h.get(), reinterpret_cast<const double*>(Vp.data())), _h.get());
    }
  };

  /** GW factory hook (src/green/mbpt/kernel_factory.h:28-31). */
  inline std::tuple<std::shared_ptr<void>, std::function<void(G_type&, G_type&)>> custom_gw_kernel(
      bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, const grids::transformer_t& ft,
      const bz_utils_t& bz_utils, const ztensor<4>& /*S_k*/) {
    auto kernel = std::make_shared<gw_hip_kernel>(X2C, p, nao, nso, ns, NQ, ft, bz_utils);
    std::function<void(G_type&, G_type&)> callback = [kernel](G_type& g, G_type& s) { kernel->solve(g, s); };
    return std::tuple{std::static_pointer_cast<void>(kernel), callback};
  }

  /** HF factory hook (kernel_factory.h:33-36): delegates to the reference CPU kernels. */
  inline std::tuple<std::shared_ptr<void>, std::function<ztensor<4>(const ztensor<4>&)>> custom_hf_kernel(
      bool X2C, const params::params& p, size_t nao, size_t nso, size_t ns, size_t NQ, double madelung,
      const bz_utils_t& bz_utils, const ztensor<4>& S_k) {
    using x_type = ztensor<4>;
    if (X2C) {
      std::shared_ptr<void> kernel(new kernels::hf_x2c_cpu_kernel(p, nao, nso, ns, NQ, madelung, bz_utils, S_k));
      std::function callback = [kernel](const x_type& dm) -> x_type {
        return static_cast<kernels::hf_x2c_cpu_kernel*>(kernel.get())->solve(dm);
      };
      return std::tuple{kernel, callback};
    }
    std::shared_ptr<void> kernel(new kernels::hf_scalar_cpu_kernel(p, nao, nso, ns, NQ, madelung, bz_utils, S_k));
    std::function callback = [kernel](const x_type& dm) -> x_type {
      return static_cast<kernels::hf_scalar_cpu_kernel*>(kernel.get())->solve(dm);
    };
    return std::tuple{kernel, callback};
  }

}  // namespace green::mbpt::hip

#endif  // GREEN_MBPT_GW_HIP_H
