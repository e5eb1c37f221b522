// UMAP's fuzzy simplicial set of a kNN graph on the device: the weighted graph scanpy
// stores in obsp["connectivities"] (crates/scrust-core/src/neighbors.rs:193-372).
//
//   1. edges_kernel, one thread per cell: rho (the first strictly positive distance), the
//      bandwidth sigma by 64 bisection steps on sum_j exp(-(d_j - rho)/sigma) = log2(k + 1)
//      (the running sum in f32, the comparison in f64, as neighbors.rs:249-293), the floor
//      1e-3 x the row mean over k + 1 entries (neighbors.rs:287-292), and the directed
//      weight of every neighbour (neighbors.rs:305-320). Each edge i -> j of weight w is
//      emitted twice: key (i, j) forward and key (j, i) backward, the direction in the
//      weight's sign bit. Weights are rounded from an f64 exponential (correctly rounded
//      f32 but for double-rounding cases).
//   2. A radix sort of all keys (rocPRIM) brings each output entry's forward and backward
//      halves together in (row, column) order.
//   3. runs_kernel marks the first key of every (row, column) run, an exclusive scan gives
//      each run its output slot, write_kernel forms w + r - w r (neighbors.rs:363) and
//      indptr_kernel finds each row's first slot by a binary search over the sorted keys.
//
// A row whose every neighbour sits at distance 0 has rho = 0 and every weight 1, whatever
// sigma is, so the global mean distance of neighbors.rs:244 (used only for that row's
// sigma floor) does not change the output and is not computed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "knn_connect.hpp"
#include "knn_device.hpp"

namespace scrust_hip {

namespace {

constexpr uint32_t kBackward = 0x80000000u;

// exp(x) for the weights: an f64 exponential rounded once to f32.
__device__ __forceinline__ float weight_exp(float x) {
    return static_cast<float>(exp(static_cast<double>(x)));
}

// ndarray 0.16.1's sum of a contiguous slice (numeric_util::unrolled_fold): eight running
// sums over elements 8i + j, folded pairwise, then the tail in order.
__device__ float ndarray_sum(const float* x, uint32_t len) {
    float p[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t i = 0;
    while (len - i >= 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) p[j] = p[j] + x[i + j];
        i += 8;
    }
    float acc = 0.0f;
    acc = acc + (p[0] + p[4]);
    acc = acc + (p[1] + p[5]);
    acc = acc + (p[2] + p[6]);
    acc = acc + (p[3] + p[7]);
    for (; i < len; i++) acc = acc + x[i];
    return acc;
}

__global__ __launch_bounds__(256) void edges_kernel(const uint32_t* __restrict__ indices,
                                                    const float* __restrict__ distances,
                                                    uint32_t n, uint32_t k, uint32_t bits,
                                                    uint64_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals,
                                                    uint32_t* __restrict__ flags) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    const float* d = distances + static_cast<size_t>(row) * k;
    const uint32_t* nb = indices + static_cast<size_t>(row) * k;
    uint32_t bad = 0;
    for (uint32_t j = 0; j < k; j++) {
        if (nb[j] >= n) bad |= kConnFlagBadId;
        if (!(d[j] >= 0.0f) || !__builtin_isfinite(d[j])) bad |= kConnFlagBadDistance;
        // A kNN graph lists each neighbour once per row; the key-run merge below folds at
        // most one forward and one backward entry per (row, column), so repeats are refused.
        for (uint32_t i = 0; i < j; i++)
            if (nb[i] == nb[j]) bad |= kConnFlagRepeatedId;
    }
    const uint64_t invalid = static_cast<uint64_t>(n) << bits;
    uint64_t* out_keys = keys + 2 * static_cast<size_t>(row) * k;
    uint32_t* out_vals = vals + 2 * static_cast<size_t>(row) * k;
    if (bad) {
        atomicOr(flags, bad);
        for (uint32_t j = 0; j < 2 * k; j++) {
            out_keys[j] = invalid;
            out_vals[j] = 0u;
        }
        return;
    }
    float rho = 0.0f;
    for (uint32_t j = 0; j < k; j++)
        if (d[j] > 0.0f) {
            rho = d[j];
            break;
        }
    const double target = log2(static_cast<double>(k + 1));
    float low = 0.0f, high = __FLT_MAX__, mid = 1.0f;
    for (int it = 0; it < 64; it++) {
        float cardinality = 0.0f;
        for (uint32_t j = 0; j < k; j++) {
            const float offset = d[j] - rho;
            cardinality = cardinality + (offset > 0.0f ? weight_exp(-offset / mid) : 1.0f);
        }
        const double c = static_cast<double>(cardinality);
        if (__builtin_fabs(c - target) < 1e-5) break;
        if (c > target) {
            high = mid;
            mid = (low + high) / 2.0f;
        } else {
            low = mid;
            mid = high >= __FLT_MAX__ ? mid * 2.0f : (low + high) / 2.0f;
        }
    }
    float sigma = mid;
    if (rho > 0.0f) {
        const float floor_value = 1e-3f * (ndarray_sum(d, k) / static_cast<float>(k + 1));
        sigma = mid > floor_value ? mid : floor_value;
    }
    for (uint32_t j = 0; j < k; j++) {
        const uint32_t col = nb[j];
        const float offset = d[j] - rho;
        const float w = (offset <= 0.0f || sigma == 0.0f) ? 1.0f : weight_exp(-offset / sigma);
        if (col == row || w == 0.0f) {
            out_keys[2 * j] = invalid;
            out_keys[2 * j + 1] = invalid;
            out_vals[2 * j] = 0u;
            out_vals[2 * j + 1] = 0u;
        } else {
            out_keys[2 * j] = (static_cast<uint64_t>(row) << bits) | col;
            out_keys[2 * j + 1] = (static_cast<uint64_t>(col) << bits) | row;
            out_vals[2 * j] = __float_as_uint(w);
            out_vals[2 * j + 1] = __float_as_uint(w) | kBackward;
        }
    }
}

__global__ void runs_kernel(const uint64_t* __restrict__ keys, size_t m, uint64_t invalid,
                            uint32_t* __restrict__ starts) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const uint64_t key = keys[p];
    starts[p] = (key < invalid && (p == 0 || keys[p - 1] != key)) ? 1u : 0u;
}

__global__ void write_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                             const uint32_t* __restrict__ starts, const uint32_t* __restrict__ slots,
                             size_t m, uint32_t bits, size_t capacity,
                             uint32_t* __restrict__ out_indices, float* __restrict__ out_values) {
    const size_t p = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= m || !starts[p]) return;
    const uint64_t key = keys[p];
    float weight = 0.0f, reverse = 0.0f;
    for (size_t q = p; q < m && q < p + 2 && keys[q] == key; q++) {
        const uint32_t v = vals[q];
        if (v & kBackward) reverse = __uint_as_float(v & ~kBackward);
        else weight = __uint_as_float(v);
    }
    const size_t slot = slots[p];
    if (slot >= capacity) return;
    out_indices[slot] = static_cast<uint32_t>(key & ((1ull << bits) - 1));
    out_values[slot] = weight + reverse - weight * reverse;
}

// indptr[r] = output slot of the first key of row r (nnz for r == n).
__global__ void indptr_kernel(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ slots,
                              const uint32_t* __restrict__ starts, size_t m, uint32_t n, uint32_t bits,
                              uint32_t* __restrict__ indptr) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    const uint64_t target = static_cast<uint64_t>(r) << bits;
    size_t lo = 0, hi = m;
    while (lo < hi) {
        const size_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < target) lo = mid + 1; else hi = mid;
    }
    indptr[r] = lo < m ? slots[lo] : (m ? slots[m - 1] + starts[m - 1] : 0u);
}

}  // namespace

uint32_t conn_key_bits(size_t n) {
    uint32_t bits = 1;
    while ((static_cast<size_t>(1) << bits) <= n) bits++;
    return bits;
}

ConnLayout conn_layout(size_t n, size_t k, size_t temp_bytes) {
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    const size_t m = 2 * n * k;
    ConnLayout l{};
    size_t at = 0;
    l.keys[0] = at; at += up(m * sizeof(uint64_t) + 1);
    l.keys[1] = at; at += up(m * sizeof(uint64_t) + 1);
    l.vals[0] = at; at += up(m * sizeof(uint32_t) + 1);
    l.vals[1] = at; at += up(m * sizeof(uint32_t) + 1);
    l.starts = at; at += up(m * sizeof(uint32_t) + 1);
    l.slots = at; at += up(m * sizeof(uint32_t) + 1);
    l.flags = at; at += up(sizeof(uint32_t) * 4);
    l.temp = at; at += up(temp_bytes + 1);
    l.temp_bytes = temp_bytes;
    l.total = at;
    return l;
}

size_t conn_temp_bytes(size_t n, size_t k) {
    const size_t m = 2 * n * k;
    const unsigned end_bit = 2 * conn_key_bits(n) + 1;
    size_t sort_bytes = 0, scan_bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr),
                                    static_cast<uint64_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), m, 0, end_bit);
    (void)rocprim::exclusive_scan(nullptr, scan_bytes, static_cast<uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), 0u, m, rocprim::plus<uint32_t>());
    return std::max(sort_bytes, scan_bytes);
}

hipError_t conn_enqueue(const uint32_t* indices, const float* distances, size_t n, size_t k,
                        void* workspace, const ConnLayout& l, uint32_t* out_indptr,
                        uint32_t* out_indices, float* out_values, size_t capacity,
                        hipStream_t stream) {
    char* base = static_cast<char*>(workspace);
    uint64_t* keys_a = reinterpret_cast<uint64_t*>(base + l.keys[0]);
    uint64_t* keys_b = reinterpret_cast<uint64_t*>(base + l.keys[1]);
    uint32_t* vals_a = reinterpret_cast<uint32_t*>(base + l.vals[0]);
    uint32_t* vals_b = reinterpret_cast<uint32_t*>(base + l.vals[1]);
    uint32_t* starts = reinterpret_cast<uint32_t*>(base + l.starts);
    uint32_t* slots = reinterpret_cast<uint32_t*>(base + l.slots);
    uint32_t* flags = reinterpret_cast<uint32_t*>(base + l.flags);
    void* temp = base + l.temp;
    const size_t m = 2 * n * k;
    const uint32_t bits = conn_key_bits(n);
    hipError_t err = hipMemsetAsync(flags, 0, sizeof(uint32_t) * 4, stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(edges_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream,
                       indices, distances, static_cast<uint32_t>(n), static_cast<uint32_t>(k), bits,
                       keys_a, vals_a, flags);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    size_t temp_bytes = l.temp_bytes;
    err = rocprim::radix_sort_pairs(temp, temp_bytes, keys_a, keys_b, vals_a, vals_b, m, 0,
                                    2 * bits + 1, stream);
    if (err != hipSuccess) return err;
    const unsigned blocks = static_cast<unsigned>((m + 255) / 256);
    hipLaunchKernelGGL(runs_kernel, dim3(blocks), dim3(256), 0, stream, keys_b, m,
                       static_cast<uint64_t>(n) << bits, starts);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    temp_bytes = l.temp_bytes;
    err = rocprim::exclusive_scan(temp, temp_bytes, starts, slots, 0u, m, rocprim::plus<uint32_t>(), stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(write_kernel, dim3(blocks), dim3(256), 0, stream, keys_b, vals_b, starts, slots, m,
                       bits, capacity, out_indices, out_values);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    hipLaunchKernelGGL(indptr_kernel, dim3(static_cast<unsigned>((n + 1 + 255) / 256)), dim3(256), 0, stream,
                       keys_b, slots, starts, m, static_cast<uint32_t>(n), bits, out_indptr);
    return hipGetLastError();
}

}  // namespace scrust_hip
