// Host declarations of the fuzzy-simplicial-set (connectivities) pipeline.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>

namespace scrust_hip {

constexpr uint32_t kConnFlagBadId = 1u;        // a neighbour id >= n
constexpr uint32_t kConnFlagBadDistance = 2u;  // a negative or non-finite distance
constexpr uint32_t kConnFlagRepeatedId = 4u;   // a neighbour id listed twice in one row

// Byte offsets of the pipeline's device workspace.
struct ConnLayout {
    size_t keys[2], vals[2];  // sort input and output (2 n k entries each)
    size_t starts, slots;     // run starts and their exclusive scan
    size_t flags;             // validation flags (u32)
    size_t temp, temp_bytes;  // rocPRIM scratch
    size_t total;
};

// Bits of a row or column id in the sort keys ((row << bits) | column).
uint32_t conn_key_bits(size_t n);
ConnLayout conn_layout(size_t n, size_t k, size_t temp_bytes);
// rocPRIM scratch for an n x k graph (the layout's temp_bytes).
size_t conn_temp_bytes(size_t n, size_t k);
// Enqueues the pipeline on device buffers; out_indices / out_values hold `capacity`
// entries (2 n k always suffice) and out_indptr n + 1. The flags word in the workspace
// holds kConnFlag* bits once the stream reaches the end.
hipError_t conn_enqueue(const uint32_t* indices, const float* distances, size_t n, size_t k,
                        void* workspace, const ConnLayout& layout, uint32_t* out_indptr,
                        uint32_t* out_indices, float* out_values, size_t capacity,
                        hipStream_t stream);

}  // namespace scrust_hip
