// hr_misfit_noalloc: Compute_Misfit (Solver_PETSc_Inverse.F90:633-650) of the
// last forward with every scratch buffer in a fixed device array, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It has no counterpart in the reference.
//
// `scratch` holds the per-step partial sums and per-shot misfits.  Every GPU the
// library runs on holds its own copy until the process ends, and every call on
// a GPU shares it: calls that overlap on one GPU race.  In host code `scratch`
// names a host-side stand-in, never the device memory: kernels get the address
// hipGetSymbolAddress returns on the current device, kept per device below.

#include <hip/hip_runtime.h>
#include <mutex>
#include "hr_internal.h"
#include "../../include/hirensi.h"

namespace {

constexpr size_t HR_NOALLOC_SCRATCH = (size_t)1 << 20;   // doubles (8 MiB)
static __device__ __align__(256) double scratch[HR_NOALLOC_SCRATCH];

constexpr int HR_MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[HR_MAX_DEVICES] = {};

// device address of `scratch` on the current device (looked up once per device)
int scratch_address(hr_ctx* ctx, double** out) {
  int dev = -1;
  const hipError_t e0 = hipGetDevice(&dev);
  if (e0 != hipSuccess) return hr::ctx_fail(ctx, (int)e0, "hipGetDevice failed");
  if (dev < 0 || dev >= HR_MAX_DEVICES) return hr::ctx_fail(ctx, 1, "device index out of range for the fixed scratch");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hr::ctx_fail(ctx, e != hipSuccess ? (int)e : 3, "hipGetSymbolAddress of the misfit scratch failed");
    scratch_dev[dev] = (double*)p;
  }
  *out = scratch_dev[dev];
  return 0;
}

}  // namespace

extern "C" int hr_misfit_noalloc(hr_ctx* ctx, int nstep, double dt, int nrec, const double* dis_meas,
                                 double* misfit) {
  if (int rc = hr::misfit_check(ctx, nstep, nrec, dis_meas, misfit)) return rc;
  if (!hr::ctx_device_ptrs(ctx))
    return hr::ctx_fail(ctx, 1, "hr_misfit_noalloc needs a context created with HR_DEVICE_PTRS (host data would need staging)");
  if (hr::misfit_scratch_size(ctx) > HR_NOALLOC_SCRATCH)
    return hr::ctx_fail(ctx, 1, "hr_misfit_noalloc: nshot x (nstep + 1) + nshot partial sums exceed the fixed scratch of 1048576 doubles");
  double* scr = nullptr;
  if (int rc = scratch_address(ctx, &scr)) return rc;
  // every partial sum is written before it is read: nothing to clear
  if (int rc = hr::misfit_launch(ctx, dt, dis_meas, scr, misfit)) return rc;
  const hipError_t e = hipStreamSynchronize(hr::ctx_stream(ctx));
  if (e != hipSuccess) return hr::ctx_fail(ctx, (int)e, "hipStreamSynchronize failed");
  return 0;
}
