// Stub: the transformer_t calls the adapter makes (grid sizes, the two bosonic transforms applied
// to unit vectors, check_chebyshev).  Transforms here are identities padded with zeros.
#pragma once
#include <cstddef>

#include <green/mbpt/kernels.h>

namespace green::grids {
  struct repn_t {
    size_t n;
    size_t nts() const { return n; }
    size_t nw() const { return n; }
  };
  struct sparse_data_t {
    size_t nts_, nw_;
    repn_t repn_fermi() const { return {nts_}; }
    repn_t repn_bose() const { return {nw_}; }
  };
  class transformer_t {
  public:
    transformer_t(size_t nts, size_t nw_b) : sd_{nts, nw_b} {}
    const sparse_data_t& sd() const { return sd_; }
    void tau_f_to_w_b(const mbpt::ztensor<4>& in, mbpt::ztensor<4>& out, size_t, size_t, bool) const { (void)in, (void)out; }
    void w_b_to_tau_f(const mbpt::ztensor<4>& in, mbpt::ztensor<4>& out, size_t, size_t, bool) const { (void)in, (void)out; }
    double check_chebyshev(const mbpt::ztensor<4>&) const { return 0.0; }

  private:
    sparse_data_t sd_;
  };
}  // namespace green::grids
