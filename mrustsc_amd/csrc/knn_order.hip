// Spatial row order for the candidate pass (gfx950).
//
// The pass scans every reference for every query, so any order of the rows gives the same
// result; the order only changes how fast each row's append bound tightens. With rows
// grouped so that neighbours sit close together in the order, and each query block's scan
// starting at its own part of the order, a row's first tiles already hold its nearest
// neighbours: the bound is near its final value from the start and later tiles append
// almost nothing. The rerank's gathers of candidate rows also become local.
//
//   1. assign_kernel  each row's nearest of P pivot rows (rows i n / P, about 4096 rows
//                     per cell), squared distance over the first 16, 32, 48 or 64
//                     features in f32, the pivots in LDS
//   2. rocPRIM radix sort of (cell, row) pairs: stable, so rows of a cell keep input order
//   3. invert_kernel  inv_perm[perm[p]] = p
//
// The assignment is a heuristic: rounding or a different pivot choice only changes the
// order, never the graph.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <rocprim/device/device_radix_sort.hpp>

#include "knn_device.hpp"
#include "knn_order.hpp"

namespace scrust_hip {

namespace {

constexpr int kRowsPerCell = 4096;
constexpr int kMaxPivots = 256;
constexpr int kOrderDims = 64;  // features the assignment looks at

__device__ __forceinline__ uint32_t pivot_row(uint32_t p, uint32_t pivots, uint32_t n) {
    return static_cast<uint32_t>((static_cast<uint64_t>(p) * n) / pivots);
}

// DIMS: the features looked at, rounded up to a multiple of 16 (zero padded on both
// sides, which adds nothing to the distances).
template <int DIMS>
__global__ __launch_bounds__(256) void assign_kernel(const float* __restrict__ x, uint32_t n, uint32_t d,
                                                     uint32_t dims, uint32_t pivots,
                                                     uint32_t* __restrict__ cells,
                                                     uint32_t* __restrict__ rows) {
    __shared__ float4 piv[kMaxPivots * (DIMS / 4)];
    float* pf = reinterpret_cast<float*>(piv);
    for (uint32_t i = threadIdx.x; i < pivots * DIMS; i += blockDim.x) {
        const uint32_t p = i / DIMS, j = i - p * DIMS;
        pf[i] = j < dims ? x[static_cast<size_t>(pivot_row(p, pivots, n)) * d + j] : 0.0f;
    }
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    float xr[DIMS];
#pragma unroll
    for (int j = 0; j < DIMS; j++) xr[j] = j < static_cast<int>(dims) ? x[static_cast<size_t>(r) * d + j] : 0.0f;
    float best = __builtin_inff();
    uint32_t cell = 0;
    for (uint32_t p = 0; p < pivots; p++) {
        const float4* pv = piv + p * (DIMS / 4);
        float s = 0.0f;
#pragma unroll
        for (int j4 = 0; j4 < DIMS / 4; j4++) {
            const float4 v = pv[j4];
            const float t0 = xr[4 * j4] - v.x, t1 = xr[4 * j4 + 1] - v.y;
            const float t2 = xr[4 * j4 + 2] - v.z, t3 = xr[4 * j4 + 3] - v.w;
            s = __builtin_fmaf(t0, t0, s);
            s = __builtin_fmaf(t1, t1, s);
            s = __builtin_fmaf(t2, t2, s);
            s = __builtin_fmaf(t3, t3, s);
        }
        if (s < best) {
            best = s;
            cell = p;
        }
    }
    cells[r] = cell;
    rows[r] = r;
}

__global__ void invert_kernel(const uint32_t* __restrict__ perm, uint32_t n, uint32_t* __restrict__ inv) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) inv[perm[p]] = p;
}

}  // namespace

OrderLayout order_layout(size_t n, size_t d) {
    OrderLayout l{};
    l.pivots = static_cast<uint32_t>(std::clamp<size_t>(n / kRowsPerCell, 2, kMaxPivots));
    l.bits = 1;
    while ((1u << l.bits) < l.pivots) l.bits++;
    // Whole groups of 16 features when there are that many (the order is a heuristic).
    l.dims = static_cast<uint32_t>(d >= 16 ? std::min<size_t>(d, kOrderDims) / 16 * 16 : d);
    auto align = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t at = 0;
    l.cells = at;
    at = align(at + n * sizeof(uint32_t));
    l.cells_sorted = at;
    at = align(at + n * sizeof(uint32_t));
    l.rows = at;
    at = align(at + n * sizeof(uint32_t));
    l.temp = at;
    (void)rocprim::radix_sort_pairs(nullptr, l.temp_bytes, static_cast<uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), n, 0, l.bits);
    at = align(at + l.temp_bytes);
    l.total = at;
    return l;
}

hipError_t order_enqueue(const float* x, size_t n, size_t d, void* workspace, const OrderLayout& l,
                         uint32_t* perm, uint32_t* inv_perm, hipStream_t stream) {
    char* base = static_cast<char*>(workspace);
    uint32_t* cells = reinterpret_cast<uint32_t*>(base + l.cells);
    uint32_t* cells_sorted = reinterpret_cast<uint32_t*>(base + l.cells_sorted);
    uint32_t* rows = reinterpret_cast<uint32_t*>(base + l.rows);
    const uint32_t n32 = static_cast<uint32_t>(n);
    const unsigned blocks = (n32 + 255) / 256;
    const uint32_t d32 = static_cast<uint32_t>(d);
    if (l.dims <= 16)
        hipLaunchKernelGGL(assign_kernel<16>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells, rows);
    else if (l.dims <= 32)
        hipLaunchKernelGGL(assign_kernel<32>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells, rows);
    else if (l.dims <= 48)
        hipLaunchKernelGGL(assign_kernel<48>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells, rows);
    else
        hipLaunchKernelGGL(assign_kernel<64>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells, rows);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    size_t temp_bytes = l.temp_bytes;
    err = rocprim::radix_sort_pairs(base + l.temp, temp_bytes, cells, cells_sorted, rows, perm, n, 0, l.bits,
                                    stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(invert_kernel, dim3(blocks), dim3(256), 0, stream, perm, n32, inv_perm);
    return hipGetLastError();
}

}  // namespace scrust_hip
