// Entry point that allocates nothing: velocity extrapolation on caller-owned
// device buffers, with the BFS status grid in a fixed per-device array.
//
// It is for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops).  The array `scratch` below is a host-side
// stand-in in host code: its device address is looked up once per device
// with hipGetSymbolAddress when a context is created on that device
// (noalloc_scratch_init, called by flipmi_ctx_create with the device
// current), and only that address is ever handed to kernels.
//
// Every call on one GPU uses that GPU's one array: calls that overlap on a
// device (two streams, two host threads) race.  The array keeps the last
// call's contents; k_extrap_init overwrites every status byte a call reads,
// so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "flipmi_kernels.h"
#include "scene.h"

namespace flipmi {

constexpr size_t kNoallocScratchBytes = FLIPMI_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoallocScratchBytes];

namespace {
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Looks up this device's copy of `scratch`; the caller has made `device`
// current.  Later calls for the same device return at once.
hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

static void *noalloc_scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace flipmi

using namespace flipmi;

extern "C" int flipmi_extrapolate_noalloc(flipmi_ctx *ctx, int ni, int nj, int nk, float *grid_dev,
                                          const uint8_t *valid_dev, int layers, void *stream) {
    if (!ctx) return FLIPMI_ERR_INVALID_ARG;
    Ctx &c = ctx->c;
    if (!grid_dev || !valid_dev || ni < 1 || nj < 1 || nk < 1 || layers < 0 || layers > 250) {
        c.last_error = "flipmi_extrapolate_noalloc: invalid argument";
        return FLIPMI_ERR_INVALID_ARG;
    }
    const size_t n = (size_t)ni * (size_t)nj * (size_t)nk;
    if (n > kNoallocScratchBytes) {
        c.last_error = "flipmi_extrapolate_noalloc: grid of " + std::to_string(n) +
                       " faces exceeds the scratch limit of " + std::to_string(kNoallocScratchBytes) +
                       " (FLIPMI_NOALLOC_SCRATCH_BYTES)";
        return FLIPMI_ERR_INVALID_ARG;
    }
    uint8_t *status = (uint8_t *)noalloc_scratch_ptr(c.device);
    if (!status) {
        c.last_error = "flipmi_extrapolate_noalloc: scratch not set up for this device";
        return FLIPMI_ERR_HIP;
    }
    hipError_t e = hipSetDevice(c.device);
    if (e == hipSuccess) {
        hipStream_t s = stream ? (hipStream_t)stream : c.stream;
        e = launch_extrapolate(ni, nj, nk, grid_dev, valid_dev, status, layers, s);
    }
    if (e != hipSuccess) {
        c.last_error = std::string("flipmi_extrapolate_noalloc: ") + hipGetErrorString(e);
        return FLIPMI_ERR_HIP;
    }
    return FLIPMI_OK;
}
