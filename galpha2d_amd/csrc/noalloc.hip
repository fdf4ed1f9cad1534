// noalloc.hip -- the Delaunay entry point for callers that cannot allocate device memory.
//
// galphaComputeDeviceNoAlloc runs the same pipeline as galphaComputeDevice (no alpha stage) but
// carves every device buffer from `scratch`, a static device array of GALPHA_NOALLOC_SCRATCH_BYTES
// that each GPU holds from the library's load until the process ends.  The caller passes the
// points and the output array; the call makes no hipMalloc or hipFree.
//
// `scratch` in host code is a host-side stand-in for the symbol, not a device address: the
// device address of each GPU's copy comes from hipGetSymbolAddress, taken once per device at the
// first call on that device (with the device current), and only that pointer is handed on.
//
// Every call on a GPU shares that GPU's array and overwrites what the previous call left there;
// the pipeline clears every counter, flag and table it relies on at the start of a run.  Calls on
// one GPU are serialised by a per-device lock and return after their device work has finished,
// so two calls never overlap on one array through this entry point.
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>

#include "alpha_host.h"
#include "../../include/gpudt.h"

struct galpha_ctx;
galpha_ctx *galpha_noalloc_ctx_new(int device);
int galpha_noalloc_run(galpha_ctx *c, void *scratch, size_t scratch_bytes, const double *d_xy, int n, int fboSize,
                       int *d_tris9, int capacity, void *stream, size_t *high_water);
extern "C" void galpha_set_last_error(const char *msg);   // galpha.cpp (library-internal)

static __device__ __align__(256) unsigned char scratch[GALPHA_NOALLOC_SCRATCH_BYTES];

namespace {
const int kMaxDevices = 64;
struct DeviceScratch {
    std::mutex mu;
    void *ptr = nullptr;        // this device's copy of `scratch` (hipGetSymbolAddress)
    galpha_ctx *ctx = nullptr;  // host bookkeeping of the runs on this device
};
DeviceScratch g_dev[kMaxDevices];
}  // namespace

extern "C" int galphaComputeDeviceNoAlloc(int device, const double *d_xy, int n, int fboSize, int *d_tris9,
                                          int capacity, void *stream)
{
    try {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count || device >= kMaxDevices)
            throw GalphaError("no such device: " + std::to_string(device));
        DeviceScratch &d = g_dev[device];
        std::lock_guard<std::mutex> lk(d.mu);
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) throw GalphaError(std::string("hipSetDevice: ") + hipGetErrorString(e));
        if (!d.ptr) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                throw GalphaError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            d.ptr = p;
        }
        if (!d.ctx) d.ctx = galpha_noalloc_ctx_new(device);
        return galpha_noalloc_run(d.ctx, d.ptr, GALPHA_NOALLOC_SCRATCH_BYTES, d_xy, n, fboSize, d_tris9, capacity,
                                  stream, nullptr);
    } catch (const std::exception &ex) {
        galpha_set_last_error(ex.what());
        return -1;
    }
}
