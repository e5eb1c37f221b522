// Host declarations of the spatial row order of the candidate pass (knn_order.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace scrust_hip {

// Datasets with at least this many cells are ordered by default.
constexpr size_t kOrderMinCells = 65536;

// Byte offsets of the ordering workspace, relative to its base.
struct OrderLayout {
    size_t cells, counts, totals, total;
    uint32_t pivots, dims, chunks;
};

OrderLayout order_layout(size_t n, size_t d);

// Enqueues the cell assignment and a stable counting sort of the rows by cell, which writes
// perm[p] = input row at pass position p and inv_perm[row] = its position.
hipError_t order_enqueue(const float* x, size_t n, size_t d, void* workspace,
                         const OrderLayout& layout, uint32_t* perm, uint32_t* inv_perm,
                         hipStream_t stream);

}  // namespace scrust_hip
