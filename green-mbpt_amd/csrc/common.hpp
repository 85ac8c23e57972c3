// Shared helpers for the gwhip HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace gwhip {

typedef double v4d __attribute__((ext_vector_type(4)));

struct hip_error : std::runtime_error {
  hipError_t code;
  hip_error(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

struct status_error : std::runtime_error {
  int status;
  status_error(int s, const std::string& what) : std::runtime_error(what), status(s) {}
};

#define GWHIP_CHECK(expr)                                                                        \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw ::gwhip::hip_error(_e, std::string(#expr " failed: ") + hipGetErrorString(_e) +      \
                                       " (" __FILE__ ":" + std::to_string(__LINE__) + ")");       \
  } while (0)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {  // c + a*b
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Precision modes of include/gwhip.h (gwhip_precision): which contractions use split products.
inline bool precision_valid(int p) { return p >= 0 && p <= 3; }
inline bool p0_split(int p) { return p == 1 || p == 2; }
inline bool sigma_split(int p) { return p == 1 || p == 3; }
inline bool any_split(int p) { return p >= 1 && p <= 3; }

// A HIP event owned by a scope: destroyed when the scope ends, also when an error unwinds it.
// Event(false) holds no event (for timing that is switched off).
struct Event {
  hipEvent_t e = nullptr;
  explicit Event(bool create = true) {
    if (create) GWHIP_CHECK(hipEventCreate(&e));
  }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
  operator hipEvent_t() const { return e; }
};

}  // namespace gwhip
