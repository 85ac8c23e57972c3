// The Sigma evaluation for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): gwhip_eval_sigma_device_noalloc
// runs gwhip_eval_sigma_device's kernels with every transient buffer inside one file-scope device
// array instead of the handle's workspace.
//
// The array lives in this translation unit's code object, so only a process that calls the entry
// point loads it; each device then holds its own copy until the process ends.  Host code never
// passes the array's name to a kernel: its device address on each GPU comes from
// hipGetSymbolAddress, looked up once per device (with that device current) on the first call there.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/gwhip.h"

namespace gwhip {
int eval_sigma_in_workspace(gwhip_handle* h, const void* G, void* Sigma, int precision, void* ws, size_t ws_bytes);
int workspace_bytes_needed(const gwhip_handle* h, int precision, size_t* bytes);
void set_handle_error(gwhip_handle* h, const char* msg);
int handle_device(const gwhip_handle* h);
}  // namespace gwhip

// The workspace: GWHIP_NOALLOC_WORKSPACE_BYTES (include/gwhip.h) per device.
static __device__ __align__(256) char scratch[GWHIP_NOALLOC_WORKSPACE_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (looked up once per device, with the device current).
int scratch_on(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return GWHIP_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_lock);
  if (!g_ptr[device]) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess || hipSetDevice(device) != hipSuccess) {
      (void)hipGetLastError();
      return GWHIP_ERR_DEVICE;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return GWHIP_ERR_DEVICE;
    }
    g_ptr[device] = p;
  }
  *out = g_ptr[device];
  return GWHIP_OK;
}

}  // namespace

extern "C" {

int gwhip_eval_sigma_device_noalloc(gwhip_handle* h, const void* G_dev, void* Sigma_dev, int precision) {
  if (!h) return GWHIP_ERR_ARG;
  size_t need = 0;
  int st = gwhip::workspace_bytes_needed(h, precision, &need);
  if (st != GWHIP_OK) {
    gwhip::set_handle_error(h, "unknown precision");
    return st;
  }
  if (need > (size_t)GWHIP_NOALLOC_WORKSPACE_BYTES) {
    gwhip::set_handle_error(h, "this system needs more workspace than the no-allocation entry point holds "
                               "(GWHIP_NOALLOC_WORKSPACE_BYTES at one tau point and one q)");
    return GWHIP_ERR_UNSUPPORTED;
  }
  void* ws = nullptr;
  st = scratch_on(gwhip::handle_device(h), &ws);
  if (st != GWHIP_OK) {
    gwhip::set_handle_error(h, "hipGetSymbolAddress of the no-allocation workspace failed");
    return st;
  }
  return gwhip::eval_sigma_in_workspace(h, G_dev, Sigma_dev, precision, ws, (size_t)GWHIP_NOALLOC_WORKSPACE_BYTES);
}

}  // extern "C"
