// Linear solve that allocates nothing (pm_linear_solve_noalloc): for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The PCG work vectors, partial sums and scalars live in one file-scope device array of fixed size.  Its
// address on each device is taken once, at pm_initialize on that device (the array's host-side name is
// not a device address); the kernels receive that address as an argument.  Every call on a device shares
// that device's array and leaves its contents behind; every value a solve reads there is written by the
// same solve first.  Calls that overlap on one device race on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "pm_internal.hpp"

namespace pm {

namespace {

// 64 MiB: 4 vectors of up to kNoallocMaxDof doubles after 6,144 partial sums and 64 scalars
constexpr int64_t kScratchDoubles = (64ll << 20) / 8;
constexpr int64_t kScalars = 64, kPartials = 3 * 2048;
constexpr int64_t kNoallocMaxDof = 2000000;
static_assert(kScalars + kPartials + 4 * kNoallocMaxDof <= kScratchDoubles, "scratch holds the largest solve");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device (null until set up)
std::string g_setup_error[kMaxDevices];

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// Called by pm_initialize with the handle's device current.  A failure is kept and reported by
// pm_linear_solve_noalloc; the rest of the library does not depend on the array.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mutex);
  if (g_scratch[device]) return;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    g_setup_error[device] = std::string("cannot locate the no-allocation scratch array: ") + hipGetErrorString(e);
    return;
  }
  g_scratch[device] = static_cast<double*>(ptr);
}

}  // namespace pm

using namespace pm;

extern "C" int pm_linear_solve_noalloc(pm_fem* f, const double* d_b, double* d_x, int32_t* iterations) {
  if (!f) return 1;
  if (!f->initialized) { f->err = "pm_initialize has not been called"; return 1; }
  if (!d_b || !d_x || !iterations) { f->err = "null argument"; return 1; }
  if (f->ndof > kNoallocMaxDof) {
    f->err = "pm_linear_solve_noalloc: " + std::to_string(f->ndof) + " DOFs exceed the scratch maximum of " +
             std::to_string(kNoallocMaxDof);
    return 1;
  }
  double* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mutex);
    base = f->device >= 0 && f->device < kMaxDevices ? g_scratch[f->device] : nullptr;
    if (!base) {
      f->err = f->device >= 0 && f->device < kMaxDevices && !g_setup_error[f->device].empty()
                   ? g_setup_error[f->device] : "the no-allocation scratch array is not set up on this device";
      return 1;
    }
  }
  if (hipSetDevice(f->device) != hipSuccess) { f->err = "invalid HIP device"; return 1; }
  double* scal = base;
  double* red = base + kScalars;
  double* work = red + kPartials;
  build_block_jacobi(f);
  int it = 0;
  const int rc = pcg_solve_sym(f, d_b, d_x, work, red, scal, &it);
  *iterations = it;
  return rc;
}
