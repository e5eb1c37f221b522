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
//                     features in f32 (as |p|^2 / 2 - x.p), the pivots in LDS
//   2. a stable counting sort of the rows by cell (P <= 256 cells): count_kernel counts
//      each wave chunk's cells, scan_kernel turns each cell's counts into the chunks'
//      first positions within the cell, scatter_kernel adds the cells' bases, places each
//      chunk's rows in input order and writes the inverse permutation
//
// The assignment is a heuristic: rounding or a different pivot choice only changes the
// order, never the graph.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "knn_device.hpp"
#include "knn_order.hpp"

namespace scrust_hip {

namespace {

constexpr int kRowsPerCell = 4096;
constexpr int kMaxPivots = 256;
constexpr int kOrderDims = 64;  // features the assignment looks at
constexpr uint32_t kSortChunk = 1024;  // rows of one wave's part of the counting sort

__device__ __forceinline__ uint32_t pivot_row(uint32_t p, uint32_t pivots, uint32_t n) {
    return static_cast<uint32_t>((static_cast<uint64_t>(p) * n) / pivots);
}

// DIMS: the features looked at, rounded up to a multiple of 16 (zero padded on both
// sides, which adds nothing to the distances).
template <int DIMS>
__global__ __launch_bounds__(256) void assign_kernel(const float* __restrict__ x, uint32_t n, uint32_t d,
                                                     uint32_t dims, uint32_t pivots,
                                                     uint32_t* __restrict__ cells) {
    __shared__ float4 piv[kMaxPivots * (DIMS / 4)];
    __shared__ float half_norm[kMaxPivots];
    float* pf = reinterpret_cast<float*>(piv);
    for (uint32_t i = threadIdx.x; i < pivots * DIMS; i += blockDim.x) {
        const uint32_t p = i / DIMS, j = i - p * DIMS;
        pf[i] = j < dims ? x[static_cast<size_t>(pivot_row(p, pivots, n)) * d + j] : 0.0f;
    }
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < pivots; p += blockDim.x) {
        float s = 0.0f;
        for (int j = 0; j < DIMS; j++) s = __builtin_fmaf(pf[p * DIMS + j], pf[p * DIMS + j], s);
        half_norm[p] = 0.5f * s;
    }
    __syncthreads();
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    // |x - p|^2 = |x|^2 + 2 (|p|^2 / 2 - x.p): the nearest pivot has the smallest bracket.
    float nx[DIMS];
#pragma unroll
    for (int j = 0; j < DIMS; j++) nx[j] = j < static_cast<int>(dims) ? -x[static_cast<size_t>(r) * d + j] : 0.0f;
    float best = __builtin_inff();
    uint32_t cell = 0;
    for (uint32_t p = 0; p < pivots; p++) {
        const float4* pv = piv + p * (DIMS / 4);
        float s = half_norm[p];
#pragma unroll
        for (int j4 = 0; j4 < DIMS / 4; j4++) {
            const float4 v = pv[j4];
            s = __builtin_fmaf(nx[4 * j4], v.x, s);
            s = __builtin_fmaf(nx[4 * j4 + 1], v.y, s);
            s = __builtin_fmaf(nx[4 * j4 + 2], v.z, s);
            s = __builtin_fmaf(nx[4 * j4 + 3], v.w, s);
        }
        if (s < best) {
            best = s;
            cell = p;
        }
    }
    cells[r] = cell;
}

// Counting sort, 1: wave c (of `chunks`) counts the cells of rows [c kSortChunk,
// (c + 1) kSortChunk) into counts[cell * chunks + c].
__global__ __launch_bounds__(256) void count_kernel(const uint32_t* __restrict__ cells, uint32_t n,
                                                    uint32_t pivots, uint32_t chunks,
                                                    uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[4][kMaxPivots];
    const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t c = blockIdx.x * 4 + w;
    for (uint32_t b = lane; b < static_cast<uint32_t>(kMaxPivots); b += kWave) hist[w][b] = 0;
    __syncthreads();
    if (c < chunks) {
        const uint32_t r0 = c * kSortChunk, r1 = n - r0 < kSortChunk ? n : r0 + kSortChunk;
        for (uint32_t r = r0 + lane; r < r1; r += kWave) atomicAdd(&hist[w][cells[r]], 1u);
    }
    __syncthreads();
    if (c < chunks)
        for (uint32_t b = lane; b < pivots; b += kWave) counts[static_cast<size_t>(b) * chunks + c] = hist[w][b];
}

// Counting sort, 2: workgroup b replaces cell b's counts (one per chunk) by their exclusive
// prefix sums, the chunks' first positions within the cell, and writes the cell's total.
__global__ __launch_bounds__(256) void scan_kernel(uint32_t* __restrict__ counts, uint32_t chunks,
                                                   uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[256];
    const uint32_t t = threadIdx.x;
    uint32_t* const cell = counts + static_cast<size_t>(blockIdx.x) * chunks;
    const uint32_t per = (chunks + 255) / 256;
    const uint32_t a = t * per < chunks ? t * per : chunks, e = a + per < chunks ? a + per : chunks;
    uint32_t sum = 0;
    for (uint32_t i = a; i < e; i++) sum += cell[i];
    part[t] = sum;
    __syncthreads();
    // Inclusive scan of the threads' sums (Hillis-Steele over LDS).
    for (uint32_t off = 1; off < 256u; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - sum;
    for (uint32_t i = a; i < e; i++) {
        const uint32_t v = cell[i];
        cell[i] = run;
        run += v;
    }
    if (t == 255) totals[blockIdx.x] = part[255];
}

// Counting sort, 3: wave c places its chunk's rows, 64 at a time in input order. Within a
// group of 64 the rows of one cell are ranked by lane (one ballot per distinct cell), so the
// rows of a cell keep their input order: the result of a stable sort by cell.
__global__ __launch_bounds__(256) void scatter_kernel(const uint32_t* __restrict__ cells, uint32_t n,
                                                      uint32_t pivots, uint32_t chunks,
                                                      const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ totals,
                                                      uint32_t* __restrict__ perm, uint32_t* __restrict__ inv_perm) {
    __shared__ uint32_t next[4][kMaxPivots];
    const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const uint32_t c = blockIdx.x * 4 + w;
    if (c >= chunks) return;
    const uint32_t r0 = c * kSortChunk, r1 = n - r0 < kSortChunk ? n : r0 + kSortChunk;
    // The chunk's first position in each cell (the cell's base, the exclusive prefix sum of
    // the totals, plus the chunk's first position within it), advanced as rows are placed:
    // lane l holds cells 4 l .. 4 l + 3.
    static_assert(kMaxPivots == 4 * kWave, "four cells a lane");
    uint32_t v[4], s4 = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t b = 4 * lane + i;
        v[i] = b < pivots ? totals[b] : 0u;
        s4 += v[i];
    }
    uint32_t incl = s4;
    for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t o = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), off, kWave));
        if (lane >= static_cast<uint32_t>(off)) incl += o;
    }
    uint32_t run = incl - s4;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t b = 4 * lane + i;
        next[w][b] = b < pivots ? run + first[static_cast<size_t>(b) * chunks + c] : 0u;
        run += v[i];
    }
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t base = r0; base < r1; base += kWave) {
        const uint32_t r = base + lane;
        const bool on = r < r1;
        const uint32_t cell = on ? cells[r] : ~0u;
        uint64_t pending = __ballot(on);
        while (pending) {
            const int leader = __builtin_ctzll(pending);
            const uint32_t b = static_cast<uint32_t>(__shfl(static_cast<int>(cell), leader, kWave));
            const uint64_t same = __ballot(cell == b);
            const uint32_t start = next[w][b];
            if (cell == b) {
                const uint32_t pos = start + static_cast<uint32_t>(__builtin_popcountll(same & below));
                perm[pos] = r;
                inv_perm[r] = pos;
            }
            if (lane == static_cast<uint32_t>(leader)) next[w][b] += static_cast<uint32_t>(__builtin_popcountll(same));
            pending &= ~same;
        }
    }
}

}  // namespace

OrderLayout order_layout(size_t n, size_t d) {
    OrderLayout l{};
    l.pivots = static_cast<uint32_t>(std::clamp<size_t>(n / kRowsPerCell, 2, kMaxPivots));
    l.chunks = static_cast<uint32_t>((n + kSortChunk - 1) / kSortChunk);
    // Whole groups of 16 features when there are that many (the order is a heuristic).
    l.dims = static_cast<uint32_t>(d >= 16 ? std::min<size_t>(d, kOrderDims) / 16 * 16 : d);
    auto align = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t at = 0;
    l.cells = at;
    at = align(at + n * sizeof(uint32_t));
    l.counts = at;
    at = align(at + static_cast<size_t>(l.pivots) * l.chunks * sizeof(uint32_t));
    l.totals = at;
    at = align(at + kMaxPivots * sizeof(uint32_t));
    l.total = at;
    return l;
}

hipError_t order_enqueue(const float* x, size_t n, size_t d, void* workspace, const OrderLayout& l,
                         uint32_t* perm, uint32_t* inv_perm, hipStream_t stream) {
    char* base = static_cast<char*>(workspace);
    uint32_t* cells = reinterpret_cast<uint32_t*>(base + l.cells);
    uint32_t* counts = reinterpret_cast<uint32_t*>(base + l.counts);
    uint32_t* totals = reinterpret_cast<uint32_t*>(base + l.totals);
    const uint32_t n32 = static_cast<uint32_t>(n);
    const unsigned blocks = (n32 + 255) / 256;
    const uint32_t d32 = static_cast<uint32_t>(d);
    if (l.dims <= 16)
        hipLaunchKernelGGL(assign_kernel<16>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells);
    else if (l.dims <= 32)
        hipLaunchKernelGGL(assign_kernel<32>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells);
    else if (l.dims <= 48)
        hipLaunchKernelGGL(assign_kernel<48>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells);
    else
        hipLaunchKernelGGL(assign_kernel<64>, dim3(blocks), dim3(256), 0, stream, x, n32, d32, l.dims, l.pivots, cells);
    const unsigned chunk_blocks = (l.chunks + 3) / 4;
    hipLaunchKernelGGL(count_kernel, dim3(chunk_blocks), dim3(256), 0, stream, cells, n32, l.pivots, l.chunks, counts);
    hipLaunchKernelGGL(scan_kernel, dim3(l.pivots), dim3(256), 0, stream, counts, l.chunks, totals);
    hipLaunchKernelGGL(scatter_kernel, dim3(chunk_blocks), dim3(256), 0, stream, cells, n32, l.pivots, l.chunks,
                       counts, totals, perm, inv_perm);
    return hipGetLastError();
}

}  // namespace scrust_hip
