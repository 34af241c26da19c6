// Context operations shared between the translation units of the library
// (defined in hr_api.hip; hr_ctx stays opaque elsewhere).
#pragma once
#include <cstddef>
#include <hip/hip_runtime.h>

struct hr_ctx;

namespace hr {
// checks of hr_misfit's arguments against the last forward (0 = fine, else the error code, message set)
int misfit_check(hr_ctx* ctx, int nstep, int nrec, const void* dis_meas, const void* misfit);
// doubles of partial sums Compute_Misfit needs for the last forward: nshot x (nstep + 1) + nshot
size_t misfit_scratch_size(const hr_ctx* ctx);
// the misfit kernels on the context's stream, partial sums in part, the total in *out (device)
int misfit_launch(hr_ctx* ctx, double dt, const double* dobs, double* part, double* out);
bool ctx_device_ptrs(const hr_ctx* ctx);
hipStream_t ctx_stream(const hr_ctx* ctx);
int ctx_fail(hr_ctx* ctx, int code, const char* msg);
}  // namespace hr
