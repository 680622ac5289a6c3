// noalloc.hip -- hz_adjoint_noalloc: the adjoint of hz_adjoint for callers
// that cannot allocate (code under HIP graph capture, serving loops that
// forbid hipMalloc).  No counterpart in the reference.
//
// hz_adjoint keeps the partial sums of its dL/dlambda reduction in device
// memory it allocates on first use, and the result in a pinned host word it
// waits for.  This entry point runs the same kernels with the partial sums in
// `scratch`, a file-scope device array of kDotPartials doubles (8 KiB, held
// by every GPU the library runs on until the process ends), and writes
// dL/dlambda to a device word the caller passes; it neither allocates nor
// synchronises.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress, looked up
// once per device on the first call with that device current, and only that
// address is handed to kernels.  Every call on one GPU shares that GPU's
// array; the reduction writes every partial sum it reads, so no clearing is
// needed, but calls that overlap on one GPU (two streams at once) race.
#include "hz_internal.h"

#include <mutex>

namespace hz {
namespace {

// partial sums of the dL/dlambda reduction: the first pass runs at most
// kDotPartials work-groups whatever the field size, so no input is too large
static __device__ __align__(256) double scratch[kDotPartials];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

// Device address of the current device's `scratch` (looked up once).
int scratch_address(int device, double **out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(HZ_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(HZ_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = (double *)ptr;
    }
    *out = g_scratch[device];
    return HZ_OK;
}

}  // namespace
}  // namespace hz

using namespace hz;

extern "C" int hz_adjoint_noalloc(hz_plan *plan, const double *ubar, const double *u, double *fbar, double *glbar,
                                  double *grbar, double *lambda_bar_dev, hz_stream_t stream)
{
    if (!plan) return set_error(HZ_ERR_ARG, "null argument");
    int dev = -1;
    HZ_HIP(hipGetDevice(&dev));
    if (dev != plan->device) return set_error(HZ_ERR_ARG, "the plan's device is not the current device");
    double *partial = nullptr;
    if (lambda_bar_dev) {
        if (int st = scratch_address(dev, &partial)) return st;
    }
    return adjoint_prebuilt(plan, ubar, u, fbar, glbar, grbar, lambda_bar_dev, partial, (hipStream_t)stream);
}
