// The allocation-free variant of the Doppler-centroid entry point. Its per-strip lag sums go to a fixed device
// array that lives in this file (one copy per GPU, held until the process ends) instead of the context's scratch
// buffer, and they come back to the host through a stack array, so no call here allocates device, pinned or heap
// memory. For callers that must not allocate (tight serving loops, harnesses that forbid hipMalloc).
#include <mutex>

#include "asarf.h"
#include "internal.h"

namespace {

// Lag sums of at most kDcMaxPartials (strip, block) pairs, re and im each.
constexpr int kScratchVals = 2 * asarf::kDcMaxPartials;
__device__ __align__(256) double scratch[kScratchVals];

// Device address of `scratch` per device, taken once per device with that device current. The array's name is
// a host-side stand-in, not a device address, so every host use goes through this pointer.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_ptr(int device, double** out) {
    if (device < 0 || device >= kMaxDevices)
        return asarf::set_error(ASARF_EINVAL, "asarf_doppler_centroid_noalloc: device %d", device);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch_ptr[device]) {
        hipError_t e = hipSetDevice(device);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return asarf::set_error(ASARF_EHIP, "asarf_doppler_centroid_noalloc: scratch address on device %d: %s",
                                    device, hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return ASARF_OK;
}

}  // namespace

extern "C" int asarf_doppler_centroid_noalloc(asarf_ctx* c, const asarf_image* img, double prf, int quirks,
                                              double* poly3, double* block_freq, double* block_x, int max_blocks,
                                              int* n_blocks) {
    using namespace asarf;
    if (!c || !poly3) return set_error(ASARF_EINVAL, "asarf_doppler_centroid_noalloc: null argument");
    DcLayout L;
    int st = dc_layout(img, prf, quirks, "asarf_doppler_centroid_noalloc", &L);
    if (st != ASARF_OK) return st;
    if (L.nvals > (size_t)kScratchVals)
        return set_error(ASARF_ESHAPE, "asarf_doppler_centroid_noalloc: %zu lag sums exceed the fixed scratch of %d",
                         L.nvals, kScratchVals);
    if ((st = ctx_init(c)) != ASARF_OK) return st;
    double* d = nullptr;
    if ((st = scratch_ptr(c->device, &d)) != ASARF_OK) return st;
    // every (strip, block) slot of the layout is written by the kernel, so nothing is cleared first
    double host[kScratchVals];
    st = dc_run(c, img, prf, L, d, host, poly3, block_freq, block_x, max_blocks, n_blocks);
    if (st != ASARF_OK) return st;
    return clear_error();
}
