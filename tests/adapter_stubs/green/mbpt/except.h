// Stub: the exception types the adapter throws (names as in green-mbpt's except.h).
#pragma once
#include <stdexcept>
#include <string>

namespace green::mbpt {
  struct mbpt_kernel_error : std::runtime_error {
    explicit mbpt_kernel_error(const std::string& w) : std::runtime_error(w) {}
  };
  struct mbpt_wrong_grid : std::runtime_error {
    explicit mbpt_wrong_grid(const std::string& w) : std::runtime_error(w) {}
  };
  struct mbpt_outdated_input : std::runtime_error {
    explicit mbpt_outdated_input(const std::string& w) : std::runtime_error(w) {}
  };
}  // namespace green::mbpt
