// Stub: the tensor alias and q0 enum the adapter reads (green-mbpt's common_defs.h via kernels.h).
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <vector>

namespace green::mbpt {
  template <size_t N>
  class ztensor {
  public:
    template <typename... I>
    explicit ztensor(I... n) : shape_{static_cast<size_t>(n)...} {
      size_t s = 1;
      for (size_t v : shape_) s *= v;
      data_.assign(s, std::complex<double>(0.0, 0.0));
    }
    template <typename... I>
    std::complex<double>& operator()(I... i) {
      const size_t idx[N] = {static_cast<size_t>(i)...};
      size_t off = 0;
      for (size_t d = 0; d < N; ++d) off = off * shape_[d] + idx[d];
      return data_[off];
    }
    std::complex<double>*       data() { return data_.data(); }
    const std::complex<double>* data() const { return data_.data(); }
    size_t                      size() const { return data_.size(); }

  private:
    std::array<size_t, N>             shape_;
    std::vector<std::complex<double>> data_;
  };

  enum sigma_q0_treatment_e { ignore_G0, ewald_int, extrapolate };
}  // namespace green::mbpt
