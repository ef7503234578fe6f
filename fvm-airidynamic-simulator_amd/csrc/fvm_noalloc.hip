// fvm_noalloc.hip -- the entry point that allocates nothing.
//
// fvm_conservation_sums_static computes what fvm_conservation_sums computes
// (the FP64 domain totals of rho, rho_u, rho_v, E) for a state held in the
// caller's device memory, with the same two kernels and workgroup count, and
// writes the four totals into the caller's device memory on the caller's
// stream.  Its only scratch, the per-workgroup partial sums, is the file-scope
// device array `scratch` below: no hipMalloc, hipFree or allocating library
// call on any path, so it can be issued under HIP graph capture or from a
// loop that must not allocate.
//
// Every call on a GPU shares that GPU's copy of the array.  Each call
// overwrites every partial it reads before reading it, so nothing needs
// clearing between calls; calls that overlap on one GPU (on different streams)
// race on the array.
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>

#include "../../include/fvm/fvm.h"
#include "fvm_engine.hpp"
#include "fvm_kernels.hpp"

namespace {

// kConsBlocks partials of four doubles: 32 KiB per GPU, held until the
// process ends.
__device__ __align__(256) double scratch[fvm::kConsBlocks * 4];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch_ptr[kMaxDevices] = {};   // device address of `scratch` on each GPU

// Device address of this GPU's `scratch` (looked up once per device, with the
// device current).  The array's name is only a host-side handle; kernels get
// this address.
int scratch_for(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return fvm::set_error(FVM_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fvm::hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = (double*)p;
    }
    *out = g_scratch_ptr[device];
    return FVM_OK;
}

}  // namespace

extern "C" int fvm_conservation_sums_static(const void* dev_state, int elem_bytes, int nx, int ny, double* dev_out4,
                                            void* stream) {
    if (!dev_state || !dev_out4) return fvm::set_error(FVM_ERR_INVALID_ARGUMENT, "null device pointer");
    if (elem_bytes != 4 && elem_bytes != 8)
        return fvm::set_error(FVM_ERR_INVALID_ARGUMENT, "elem_bytes must be 4 (FP32 state) or 8 (FP64 state)");
    if (nx < 1 || ny < 1 || (long long)nx * ny > (long long)FVM_STATIC_MAX_CELLS)
        return fvm::set_error(FVM_ERR_INVALID_ARGUMENT,
                              "fvm_conservation_sums_static: nx and ny must be >= 1 and nx*ny at most FVM_STATIC_MAX_CELLS");
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return fvm::hip_fail(e, "hipGetDevice");
    double* partial = nullptr;
    int rc = scratch_for(device, &partial);
    if (rc) return rc;
    const size_t n = (size_t)nx * ny;
    const char* base = (const char*)dev_state;
    const void* planes[4];
    for (int f = 0; f < 4; f++) planes[f] = base + (size_t)f * n * elem_bytes;
    fvm::launch_cons(elem_bytes == 8, (hipStream_t)stream, planes, n, partial, fvm::kConsBlocks, dev_out4);
    e = hipGetLastError();
    if (e != hipSuccess) return fvm::hip_fail(e, "fvm_conservation_sums_static launch");
    return FVM_OK;
}
