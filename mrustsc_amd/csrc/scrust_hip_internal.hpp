// Internal host types of scrust_hip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/scrust_hip.h"

namespace scrust_hip {

// A grow-only device allocation: a context reuses its buffers across calls, so a
// steady stream of same-sized calls allocates once.
struct Buffer {
    void* ptr = nullptr;
    size_t capacity = 0;
    Buffer() = default;
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer();
    hipError_t reserve(size_t bytes);
};

struct Workspace {
    Buffer input, out_indices, out_distances;  // host-input entry point only
    Buffer coords, norms, mean, acc, scalars, fail_list;
    Buffer pack_ref, pack_query;
    Buffer buf, cand, theta;
};

scrust_hip_status fail(scrust_hip_status status, const std::string& message);
scrust_hip_status invalid(const char* parameter, const char* constraint, const std::string& value);
scrust_hip_status hip_failure(hipError_t err, const char* step);
scrust_hip_status validate(const void* embedding, size_t n, size_t d, size_t k,
                           const void* out_indices, const void* out_distances,
                           const scrust_hip_knn_options* options);
const char* last_error();
void clear_error();
bool device_is_gfx950(int device, hipDeviceProp_t* props_out);

}  // namespace scrust_hip

struct scrust_hip_context {
    int device = 0;
    int compute_units = 0;
    hipStream_t stream = nullptr;
    hipEvent_t events[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    scrust_hip::Workspace ws;
    ~scrust_hip_context();
};

namespace scrust_hip {
scrust_hip_status run_knn(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                          const scrust_hip_knn_options* options, uint32_t* out_indices,
                          float* out_distances, scrust_hip_knn_stats* stats);
}
