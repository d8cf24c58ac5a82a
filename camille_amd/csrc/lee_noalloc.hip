// Allocation-free boundary update: lee_update_boundary_noalloc.
//
// The overset interpolation is two-phase: every receiver value is first
// gathered from its donors into a staging buffer, and only then written to
// the receivers (receivers of one block may be donors of another, or of
// the same block across an O-grid seam).  lee_update_boundary stages
// through a buffer the handle allocated at create time and copies the
// state from and to host memory.  This variant works in place on device
// arrays the caller owns and stages through `scratch`, a static device
// array, so it makes no allocation and no host copy: it can run where the
// caller may not allocate (under HIP graph capture, inside a serving loop).
//
// Each device holds its own copy of `scratch` at its own address.  The
// address is looked up once per device at handle creation, with that
// device current (noalloc_prepare_device, called by lee_create); host code
// then only ever passes that looked-up address to kernels.  Every receiver
// slot of the staging area is written by the gather before the scatter
// reads it, so nothing needs clearing between calls.  All calls on one
// device share the array: calls that overlap in time on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "lee_internal.h"

using namespace lee;

// Receiver staging, LEE_NOALLOC_MAX_RECV * 5 doubles per device.
static __device__ __align__(256) double scratch[(size_t)LEE_NOALLOC_MAX_RECV * NV];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_scratch_mutex;
double *g_scratch_ptr[MAX_DEVICES] = {};         // address of `scratch` on each device
hipError_t g_scratch_err[MAX_DEVICES] = {};      // failed look-up, reported by the entry point

}  // namespace

namespace lee {

// Looks up this device's copy of `scratch` (the device must be current).
hipError_t noalloc_prepare_device(int device) {
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e == hipSuccess && !ptr) e = hipErrorInvalidSymbol;
  if (e != hipSuccess) {
    g_scratch_err[device] = e;
    (void)hipGetLastError();  // keep the failure out of later launch checks
    return e;
  }
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  return hipSuccess;
}

static double *noalloc_scratch(int device, hipError_t &err) {
  err = hipErrorInvalidDevice;
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  err = g_scratch_err[device];
  return g_scratch_ptr[device];
}

}  // namespace lee

extern "C" {

int lee_update_boundary_noalloc(lee_handle *h, double *const *ys_dev) {
  if (!h) return LEE_EINVAL;
  if (!ys_dev) {
    h->err = "bad argument";
    return LEE_EINVAL;
  }
  for (int ib = 0; ib < h->nblk; ib++)
    if (!ys_dev[ib]) {
      h->err = "null device array for block " + std::to_string(ib);
      return LEE_EINVAL;
    }
  // nrecv receivers stage nvars doubles each (5, plus the passive scalars)
  if ((size_t)h->ov.nrecv * nvars(h) > (size_t)LEE_NOALLOC_MAX_RECV * NV) {
    h->err = "lee_update_boundary_noalloc stages at most " + std::to_string((size_t)LEE_NOALLOC_MAX_RECV * NV) +
             " receiver values; this handle has " + std::to_string(h->ov.nrecv) + " receivers of " +
             std::to_string(nvars(h)) + " variables";
    return LEE_EINVAL;
  }
  hipError_t lookup = hipSuccess;
  double *buf = noalloc_scratch(h->device, lookup);
  if (!buf) {
    h->err = "the static staging area of device " + std::to_string(h->device) + " is not available" +
             (lookup != hipSuccess ? std::string(": ") + hipGetErrorString(lookup) : std::string());
    return LEE_EHIP;
  }
  hipStream_t s = h->stream;
  hipError_t e = hipSuccess;
  const double he = h->cfg.harmonic_wave == LEE_WAVE_ENTROPY ? harmonic_entropy(h->cfg, h->time_sol) : 0.0;
  for (int ib = 0; ib < h->nblk && e == hipSuccess; ib++)
    if (!h->b[ib].faces.empty()) e = apply_bcs(s, h, h->b[ib], ys_dev[ib], nullptr, false, nullptr, nullptr, he);
  if (e == hipSuccess) e = overset_interpolate_arrays(s, h, ys_dev, buf);
  if (e != hipSuccess) {
    h->err = std::string("HIP error: ") + hipGetErrorString(e);
    return LEE_EHIP;
  }
  return LEE_OK;
}

}  // extern "C"
