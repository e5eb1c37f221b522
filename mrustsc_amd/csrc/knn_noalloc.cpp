// The allocation-free single call: scrust_hip_knn_device_noalloc runs the same pipeline as
// scrust_hip_knn_device with its workspace (the plan's arena and launch metadata) in a
// file-scope device array instead of the context's grow-only buffers, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// Each device the library runs on holds its own copy of `scratch` until the process ends.
// Its host-side name is only a stand-in for the symbol: the device address of a device's
// copy comes from hipGetSymbolAddress, taken once per device (at context creation, with
// that device current) and kept in `g_scratch`; every use below goes through that address.
// Every call on a device shares its copy, so calls that overlap on one device race. The
// pipeline clears what must start at zero (scalars, segment counts, column sums, the
// maximum norm) with kernels; everything else is written before it is read.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/scrust_hip.h"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

namespace {

constexpr size_t kScratchBytes = SCRUST_HIP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_error[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

}  // namespace

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch_done[device]) return;
    void* ptr = nullptr;
    const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_scratch_error[device] = err;
    g_scratch[device] = err == hipSuccess ? ptr : nullptr;
    g_scratch_done[device] = true;
}

static scrust_hip_status noalloc_workspace(int device, void** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (device < 0 || device >= kMaxDevices || !g_scratch_done[device])
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: no scratch address for device " + std::to_string(device));
    if (!g_scratch[device]) return hip_failure(g_scratch_error[device], "locate the scratch array");
    *out = g_scratch[device];
    return SCRUST_HIP_OK;
}

}  // namespace scrust_hip

using namespace scrust_hip;

extern "C" scrust_hip_status scrust_hip_knn_device_noalloc(scrust_hip_context* context,
                                                          const float* embedding, size_t n,
                                                          size_t d, size_t k,
                                                          const scrust_hip_knn_options* options,
                                                          uint32_t* out_indices,
                                                          float* out_distances,
                                                          scrust_hip_knn_stats* stats) {
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    void* workspace = nullptr;
    st = noalloc_workspace(context->device, &workspace);
    if (st != SCRUST_HIP_OK) return st;
    Workspace ws{workspace, kScratchBytes};
    return run_single(context, embedding, n, d, k, options, out_indices, out_distances, stats, nullptr, &ws);
}
