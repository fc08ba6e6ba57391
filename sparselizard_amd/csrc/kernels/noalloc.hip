// No-allocation variant of slx_plan_solve.  The solve's device workspace
// (FP64 refinement vectors, FP32 Krylov vectors, the FP32 copy of A's values,
// partial sums) lives in `scratch`, a fixed array of the library's code object,
// instead of a per-call allocation.  Each GPU holds its own copy of the array
// (at its own address) from the first use until the process ends; every call
// on a GPU shares it and finds the previous call's contents, which the solve
// never reads before writing (vectors are computed or cleared first).  Calls
// that overlap on one GPU race on it.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is taken once per device with hipGetSymbolAddress (at
// slx_create, with the device current) and only that address is passed on.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "../../../include/slx.h"

static __device__ __align__(256) unsigned char scratch[SLX_SOLVE_SCRATCH_BYTES];

namespace slx {

slx_status solve_on_workspace(slx_plan p, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner, double* x_out,
                              double* d_x_out, int32_t* outer_its, int32_t* inner_its, double* rel_residual, int32_t* method_used,
                              void* ws, size_t ws_bytes);
int plan_device(slx_plan p);
void set_last_error(const char* msg);

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device
}  // namespace

// Device address of `scratch` on `device` (which must be current), taken once.
hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

}  // namespace slx

extern "C" slx_status slx_plan_solve_noalloc(slx_plan plan, int32_t method, double rel_tol, int32_t max_outer, int32_t max_inner,
                                             double* x_out, double* d_x_out, int32_t* outer_its, int32_t* inner_its,
                                             double* rel_residual, int32_t* method_used) {
    const int dev = slx::plan_device(plan);
    if (dev < 0) {
        slx::set_last_error("solve: null plan or plan without a device context");
        return SLX_ERR_INVALID;
    }
    void* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(slx::g_scratch_mu);
        ptr = dev < slx::kMaxDevices ? slx::g_scratch_ptr[dev] : nullptr;
    }
    if (!ptr) {
        // the context was created on this device, so slx_create has taken the
        // address; this is the fallback for a device first seen here
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess) e = slx::scratch_init(dev);
        if (e != hipSuccess) {
            slx::set_last_error((std::string("solve: hipGetSymbolAddress of the scratch array: ") + hipGetErrorString(e)).c_str());
            return SLX_ERR_DEVICE;
        }
        std::lock_guard<std::mutex> lk(slx::g_scratch_mu);
        ptr = slx::g_scratch_ptr[dev];
    }
    return slx::solve_on_workspace(plan, method, rel_tol, max_outer, max_inner, x_out, d_x_out, outer_its, inner_its, rel_residual,
                                   method_used, ptr, SLX_SOLVE_SCRATCH_BYTES);
}
