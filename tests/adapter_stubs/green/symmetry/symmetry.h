// Stub: brillouin_zone_utils on a one-dimensional full mesh (k +/- q modulo nk).
#pragma once
#include <cstddef>

namespace green::symmetry {
  struct k_q_map_t {
    size_t nk;
    size_t k1_from_k2q(size_t k, size_t q) const { return (k + q) % nk; }
    size_t k2_from_k1q(size_t k, size_t q) const { return (k + nk - q) % nk; }
  };
  class brillouin_zone_utils {
  public:
    explicit brillouin_zone_utils(size_t nk) : map_{nk} {}
    size_t           nk() const { return map_.nk; }
    size_t           ink() const { return map_.nk; }
    const k_q_map_t& k_q_map() const { return map_; }

  private:
    k_q_map_t map_;
  };
}  // namespace green::symmetry
