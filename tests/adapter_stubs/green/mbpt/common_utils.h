// Stub: mbpt's leakage print; counts its calls so the test can see the adapter reach it.
#pragma once
#include <iostream>
#include <string>

namespace green::mbpt {
  inline int& leakage_prints() {
    static int n = 0;
    return n;
  }
  inline void print_leakage(double leakage, const std::string& object) {
    ++leakage_prints();
    std::cout << "Leakage of " << object << ": " << leakage << std::endl;
  }
}  // namespace green::mbpt
