// wb_noalloc.hip -- fixed device work space for the entry points that allocate nothing.
//
// Callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops) get one entry point, wb_constrained_vmult_noalloc, whose two n-vectors of work
// space (the source and the product in the device dof order) live in this file-scope
// array instead of the context's buffers.  Each GPU holds its own copy of the array
// until the process ends; its device address is looked up once per device, with that
// device current, when a context is created on it, and kept here.  Calls that overlap on
// one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "wb_internal.h"

namespace wb {

namespace {

static __device__ __align__(256) double scratch[2 * (size_t)NOALLOC_MAX_N];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[MAX_DEVICES] = {};

}  // namespace

hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) {
        e = hipSetDevice(device);
        if (e != hipSuccess) return e;
    }
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess) scratch_ptr[device] = (double *)p;
    if (cur != device) (void)hipSetDevice(cur);
    return e;
}

double *noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    return scratch_ptr[device];
}

}  // namespace wb
