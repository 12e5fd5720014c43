// Allocation-free variant of the batched inner products hip_glsc3_many
// (math.hip:922) for callers that cannot call hipMalloc: code under HIP
// graph capture, tight serving loops, harnesses that forbid allocation.
//
// The per-workgroup partial sums live in a file-scope device array of the
// library (`scratch`, RED_BLOCKS doubles per vector, at most
// NM_NOALLOC_MAX_J vectors, 512 KiB), held by every GPU the library runs
// on until the process ends.  Its device address is looked up once per
// device (hipGetSymbolAddress, on the first call with that device current)
// and only that address is handed to kernels.  Each call writes every
// partial it reads, so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "neko_amd.h"
#include "nm_common.h"
#include "nm_kernels.h"

using namespace nm;

namespace {

constexpr int NM_NOALLOC_MAX_J = 64;
constexpr int NM_NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[RED_BLOCKS * NM_NOALLOC_MAX_J];

std::mutex g_scratch_mu;
double *g_scratch_ptr[NM_NOALLOC_MAX_DEVICES] = {};

// device address of this GPU's copy of `scratch` (never the host-side
// name of the array, which is no device address)
int scratch_on_current_device(double **out) {
  int dev = 0;
  NM_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NM_NOALLOC_MAX_DEVICES) {
    set_error("nm_glsc3_many_noalloc: device %d beyond the %d supported", dev,
              NM_NOALLOC_MAX_DEVICES);
    return NM_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    NM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return NM_OK;
}

}  // namespace

extern "C" int nm_glsc3_many_noalloc(double *h_dev, const double *w,
                                     const double *const *v, const double *mult,
                                     int j, int n, hipStream_t s) {
  if (j < 1 || j > NM_NOALLOC_MAX_J) {
    set_error("nm_glsc3_many_noalloc: j = %d vectors; the static scratch holds 1..%d", j,
              NM_NOALLOC_MAX_J);
    return NM_ERR_ARG;
  }
  if (n < 0 || !h_dev || !w || !v || !mult) {
    set_error("nm_glsc3_many_noalloc: n = %d, or a null pointer argument", n);
    return NM_ERR_ARG;
  }
  double *part = nullptr;
  int st = scratch_on_current_device(&part);
  if (st) return st;
  if ((st = launch_glsc3_many_partial(w, v, mult, j, n, part, s))) return st;
  return launch_combine(part, j, h_dev, s);
}
