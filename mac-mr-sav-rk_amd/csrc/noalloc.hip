// Allocation-free variant of the FP64 inner product (mac_dot in model.hip): the same
// two-level fixed-order reduction, with its partial sums in a file-scope device array
// instead of memory the library allocates on first use.  For callers that must not
// allocate device memory (code being captured into a HIP graph, serving loops that
// forbid allocation).  The array lives once per device for the life of the process;
// calls that overlap on one device share it and race.
#include <mutex>
#include <string>

#include "host_common.h"
#include "internal.h"
#include "macstokes.h"

namespace {

// Partials of one reduction (at most 4096 blocks, see reduce_blocks) and its result.
constexpr int kScratchDoubles = 4096 + 64;
constexpr int kMaxDevices = 64;
constexpr int64_t kMaxN = MAC_NOALLOC_MAX_N;

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

std::mutex g_lock;
double* g_scratch_dev[kMaxDevices] = {};
thread_local std::string g_error;

// The device address of `scratch` on the current device, looked up once per device with
// that device current (each device holds its own copy of the array).
double* scratch_address(int device) {
  if (g_scratch_dev[device] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      throw macs::HipFailure(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double*>(p);
  }
  return g_scratch_dev[device];
}

}  // namespace

extern "C" {

const char* mac_noalloc_last_error(void) { return g_error.c_str(); }

int32_t mac_dot_noalloc(int64_t n, const double* a, const double* b, double* result) {
  try {
    g_error.clear();
    if (n < 1 || n > kMaxN || !a || !b || !result)
      throw macs::ArgFailure("mac_dot_noalloc: need 1 <= n <= 2^31 - 1 and non-null a, b, result");
    const int nb = macs::reduce_blocks(n);
    if (nb + 1 > kScratchDoubles) throw macs::ArgFailure("mac_dot_noalloc: reduction larger than its scratch");
    int device = 0;
    HIP_CHECK(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) throw macs::ArgFailure("mac_dot_noalloc: device ordinal too large");
    std::lock_guard<std::mutex> hold(g_lock);
    double* part = scratch_address(device);
    macs::reduce(a, b, n, n, 1, macs::RED_SUM, part, nullptr, 0);
    macs::finalize(part, nb, 1, macs::RED_SUM, part + nb, 0);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(result, part + nb, sizeof(double), hipMemcpyDeviceToHost));
    return MAC_OK;
  } catch (const macs::ArgFailure& e) {
    g_error = e.what();
    return MAC_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    g_error = e.what();
    return MAC_HIP_ERROR;
  }
}

}  // extern "C"
