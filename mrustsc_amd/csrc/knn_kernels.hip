// kNN kernels for gfx950, except the candidate pass (knn_pass.hip).
//
// Pipeline for a set of embeddings (one HIP stream; a single call is a set of one):
//   1. colsum / colmean   exact f32 column sums -> f64 means (Euclidean only)
//   2. prepare            centred (or unit-scaled) f32 coordinates, f32 fma squared norms,
//                         packed MFMA operands, max norm, non-finite flag
//   3. candidate pass     (knn_pass.hip) MFMA contraction + per-row candidate selection
//   4. rerank             fp32 direct differences of the KP candidates, (d^2, id) order,
//                         fp64 boundary near-tie rule, completeness certificate
//   5. exact              fp32 direct differences over every reference for listed rows
// Every kernel reads per-dataset DatasetDesc records and maps flat row indices onto
// datasets. The reference's semantics these restate are cited at each kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knn_device.hpp"
#include "knn_launch.hpp"

namespace scrust_hip {

// ===========================================================================
// Small utilities.
// ===========================================================================

__global__ void fill_u32_kernel(uint32_t* p, size_t count, uint32_t value) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i < count; i += static_cast<size_t>(gridDim.x) * blockDim.x) p[i] = value;
}

void launch_fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t stream) {
    if (count == 0) return;
    unsigned blocks = static_cast<unsigned>((count + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, count, value);
}

// ===========================================================================
// 1. Column means (crates/scrust-core/src/neighbors.rs:144-154: f64 means).
//    The sum is exact and rounded once; see exact_sum_add.
// ===========================================================================

constexpr int kColsumWaves = 4;

// A block sums colsum_rows(n) rows of 64 columns; its four waves take every fourth row into
// their own LDS superaccumulators (a dependent chain of at most 512 adds per lane), which
// are combined and added to the column's global sum.
__global__ __launch_bounds__(256) void colsum_kernel(const DatasetDesc* __restrict__ sets) {
    __shared__ long long acc[kColsumWaves][kWave][kAccWords + 1];
    const DatasetDesc& s = sets[blockIdx.z];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t col = blockIdx.y * kWave + lane;
    const uint32_t rows = colsum_rows(s.n);
    const size_t r0 = static_cast<size_t>(blockIdx.x) * rows;
    if (r0 >= s.n || blockIdx.y * kWave >= s.d) return;
    for (int k = 0; k < kAccWords; k++) acc[w][lane][k] = 0;
    size_t r1 = r0 + rows;
    if (r1 > s.n) r1 = s.n;
    if (col < s.d)
        for (size_t r = r0 + w; r < r1; r += kColsumWaves) exact_sum_add(acc[w][lane], s.x[r * s.d + col]);
    __syncthreads();
    if (w == 0 && col < s.d) {
        for (int k = 0; k < kAccWords; k++) {
            long long v = 0;
            for (int ww = 0; ww < kColsumWaves; ww++) v += acc[ww][lane][k];
            if (v) atomicAdd(&s.acc[col * kAccWords + k], static_cast<unsigned long long>(v));
        }
    }
}

__global__ void colmean_kernel(const DatasetDesc* __restrict__ sets) {
    const DatasetDesc& s = sets[blockIdx.y];
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= s.d) return;
    long long words[kAccWords];
    for (int w = 0; w < kAccWords; w++) words[w] = static_cast<long long>(s.acc[col * kAccWords + w]);
    s.mean[col] = exact_sum_value(words) / static_cast<double>(s.n);
}

// Zero each dataset's column sums and max norm.
__global__ void clear_sums_kernel(const DatasetDesc* __restrict__ sets) {
    const DatasetDesc& s = sets[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < s.d * kAccWords; i += blockDim.x) s.acc[i] = 0ull;
    if (threadIdx.x == 0) *s.max_norm_bits = 0u;
}

// ===========================================================================
// 2. Coordinates, norms and MFMA operands.
//    Euclidean: (x as f64 - mean) rounded once (neighbors.rs:155-160); norms by f32
//    fma in dimension order (knn.rs:148-156). Cosine: x / |x| with the norm in f64.
//    Operand layouts are described in knn_pass.hip. Rows n.. of a dataset on the pass
//    (up to a whole pair of 64-row tiles) write sentinel reference rows (norm feature +inf).
//    prepare_kernel (one thread per row) serves the exact path and the fp32 pass;
//    prepare_rows_kernel (64 rows per block, coalesced) the bf16 pass.
// ===========================================================================

__global__ void prepare_kernel(const DatasetDesc* __restrict__ sets, uint32_t count,
                               uint32_t total_prep, int metric, int operand,
                               unsigned* __restrict__ flags) {
    const uint32_t flat = blockIdx.x * blockDim.x + threadIdx.x;
    float nrm = 0.0f;
    bool finite = true;
    uint32_t set = 0;
    bool has_norm = false;
    if (flat < total_prep) {
        set = find_set(sets, count, flat, true);
        const DatasetDesc& s = sets[set];
        const size_t row = flat - s.prep_begin;
        const int d = static_cast<int>(s.d);
        const int nch = static_cast<int>(s.nch), group = static_cast<int>(s.group);
        const int op = s.on_pass ? (s.local ? kOperandF16Local : operand) : kOperandNone;
        if (op == kOperandBf16 || op == kOperandF16Local) {
            // prepare_rows_kernel's (coalesced, 64 rows per block).
        } else if (row < s.n) {
            has_norm = true;
            // Prepare rows, coordinates and norms are in pass order; the input in input order.
            const size_t in_row = s.perm && op != kOperandNone ? s.perm[row] : row;
            const float* src = s.x + in_row * d;
            float* dst = s.coords + row * d;
            double length = 0.0;
            if (metric == kMetricCosine) {
                double acc = 0.0;
                for (int j = 0; j < d; j++) acc = __builtin_fma((double)src[j], (double)src[j], acc);
                length = __builtin_sqrt(acc);
            }
            for (int j = 0; j < d; j++) {
                const float v = src[j];
                finite = finite && __builtin_isfinite(v);
                float c;
                if (metric == kMetricCosine)
                    c = length > 0.0 ? static_cast<float>(static_cast<double>(v) / length) : 0.0f;
                else
                    c = static_cast<float>(static_cast<double>(v) - s.mean[j]);
                dst[j] = c;
                nrm = __builtin_fmaf(c, c, nrm);
            }
            s.norms[row] = nrm;
            if (op == kOperandF32) {
                float4* p4 = reinterpret_cast<float4*>(s.pack_f32 + row * (4 * group));
                for (int t4 = 0; t4 < group; t4++) {
                    float v[4];
                    for (int e = 0; e < 4; e++) {
                        const int t = t4 * 4 + e;
                        const int g = t / group, c = t % group;
                        const int f = g * nch + c;
                        v[e] = c >= nch ? 0.0f : (f < d ? dst[f] : (f == d ? nrm : 0.0f));
                    }
                    p4[t4] = make_float4(v[0], v[1], v[2], v[3]);
                }
            }
        } else if (op == kOperandF32) {
            float4* p4 = reinterpret_cast<float4*>(s.pack_f32 + row * (4 * group));
            for (int t4 = 0; t4 < group; t4++) {
                float v[4];
                for (int e = 0; e < 4; e++) {
                    const int t = t4 * 4 + e;
                    const int g = t / group, c = t % group;
                    v[e] = (c < nch && g * nch + c == d) ? kInf : 0.0f;
                }
                p4[t4] = make_float4(v[0], v[1], v[2], v[3]);
            }
        }
    }
    // Max squared norm per dataset (norms are >= 0: their bits order as unsigned ints).
    // Wave-reduce when the whole wave belongs to one dataset, else one atomic per lane.
    const uint32_t first_set = __shfl(set, 0, kWave);
    const bool uniform = __ballot(set != first_set) == 0ull;
    if (uniform) {
        unsigned bits = has_norm ? __float_as_uint(nrm) : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned other = __shfl_xor(bits, off, kWave);
            bits = other > bits ? other : bits;
        }
        if ((threadIdx.x & 63) == 0 && flat < total_prep) atomicMax(sets[set].max_norm_bits, bits);
    } else if (has_norm) {
        atomicMax(sets[set].max_norm_bits, __float_as_uint(nrm));
    }
    if (!finite) atomicOr(flags, kFlagNonFinite);
}

// Datasets on the bf16 pass: 64 rows per block. The rows' inputs are read into LDS with
// coalesced loads, one thread per row centres (or scales) its row there and sums its norm
// (the same arithmetic as prepare_kernel), and the coordinates and packed operands leave
// with coalesced stores: consecutive threads write consecutive 16-byte chunks. Rows
// n .. ref_rows - 1 get sentinel reference rows (norm feature +inf).
constexpr int kPrepRows = 64;

__global__ __launch_bounds__(256) void prepare_rows_kernel(const DatasetDesc* __restrict__ sets,
                                                           int metric, unsigned* __restrict__ flags) {
    __shared__ float xs[kPrepRows * kMaxPackedD];
    __shared__ double mean_s[kMaxPackedD];
    __shared__ float nrm_s[kPrepRows];
    constexpr uint16_t kBf16Inf = 0x7F80u;
    const DatasetDesc& s = sets[blockIdx.z];
    if (!s.on_pass) return;
    const uint32_t n = s.n, ref_rows = (n + 2 * kTileRefsHost - 1) / (2 * kTileRefsHost) * (2 * kTileRefsHost);
    const uint32_t r0 = blockIdx.x * kPrepRows;
    if (r0 >= ref_rows) return;
    const int tid = threadIdx.x;
    const int d = static_cast<int>(s.d), nch = static_cast<int>(s.nch);
    const uint32_t rows = r0 < n ? (n - r0 < kPrepRows ? n - r0 : kPrepRows) : 0u;
    // Rows in pass order: pass position r0 + t holds input row perm[r0 + t].
    const SCRUST_GLOBAL float* x = global_ptr(s.x);
    const SCRUST_GLOBAL uint32_t* perm = s.perm ? global_ptr(s.perm) : nullptr;
    if (perm) {
        for (uint32_t i = tid; i < rows * d; i += 256) {
            const uint32_t t = i / d;
            xs[i] = x[static_cast<size_t>(perm[r0 + t]) * d + (i - t * d)];
        }
    } else {
        for (uint32_t i = tid; i < rows * d; i += 256) xs[i] = x[static_cast<size_t>(r0) * d + i];
    }
    if (metric != kMetricCosine)
        for (int j = tid; j < d; j += 256) mean_s[j] = s.mean[j];
    __syncthreads();
    if (tid < kPrepRows) {
        float nrm = 0.0f;
        bool finite = true;
        if (static_cast<uint32_t>(tid) < rows) {
            float* row = xs + tid * d;
            double length = 0.0;
            if (metric == kMetricCosine) {
                double acc = 0.0;
                for (int j = 0; j < d; j++) acc = __builtin_fma((double)row[j], (double)row[j], acc);
                length = __builtin_sqrt(acc);
            }
            for (int j = 0; j < d; j++) {
                const float v = row[j];
                finite = finite && __builtin_isfinite(v);
                float c;
                if (metric == kMetricCosine)
                    c = length > 0.0 ? static_cast<float>(static_cast<double>(v) / length) : 0.0f;
                else
                    c = static_cast<float>(static_cast<double>(v) - mean_s[j]);
                row[j] = c;
                nrm = __builtin_fmaf(c, c, nrm);
            }
        }
        nrm_s[tid] = nrm;
        // Max squared norm (norms are >= 0: their bits order as unsigned ints).
        unsigned bits = __float_as_uint(nrm);
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned other = __shfl_xor(bits, off, kWave);
            bits = other > bits ? other : bits;
        }
        if (tid == 0 && rows > 0) atomicMax(s.max_norm_bits, bits);
        if (__ballot(!finite) != 0ull && tid == 0) atomicOr(flags, kFlagNonFinite);
    }
    __syncthreads();
    // Coordinates and norms are stored in pass order too (the rerank reads a row's
    // candidates, which sit near it in that order).
    SCRUST_GLOBAL float* coords = global_ptr(s.coords) + static_cast<size_t>(r0) * d;
    for (uint32_t i = tid; i < rows * d; i += 256) coords[i] = xs[i];
    if (static_cast<uint32_t>(tid) < rows) global_ptr(s.norms)[r0 + tid] = nrm_s[tid];
    const uint32_t prep_rows = ref_rows - r0 < kPrepRows ? ref_rows - r0 : kPrepRows;
    if (s.local) {
        // Local-origin pass: the coordinates again, padded with zeros to 32 nch floats per
        // row (sentinel rows n .. ref_rows - 1 all zero); the pass converts them tile by tile.
        const int width = 32 * nch;
        SCRUST_GLOBAL f32x4* src = reinterpret_cast<SCRUST_GLOBAL f32x4*>(global_ptr(s.pack_src) + static_cast<size_t>(r0) * width);
        for (uint32_t u = tid; u < prep_rows * (width / 4); u += 256) {
            const uint32_t row = u / (width / 4);
            const int f0 = static_cast<int>(u - row * (width / 4)) * 4;
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; e++) v[e] = (row < rows && f0 + e < d) ? xs[row * d + f0 + e] : 0.0f;
            src[u] = f32x4{v[0], v[1], v[2], v[3]};
        }
        return;
    }
    // Packed operands (see knn_pass.hip), in MFMA K order: reference [h | nh nl nl2 | l | h | 0],
    // query [h | -1/2 x3 | h | l | 0], 32 nch features per row. K index kappa sits in 16-byte
    // chunk (kappa % 32) / 8 * nch + kappa / 32 of the row (k_slot): MFMA step c reads chunk
    // g * nch + c in lane group g, so step c holds K indices 32 c .. 32 c + 31. Query rows are
    // stored row by row; reference rows tile by tile, coarse steps first (tile_chunk).
    const int chunks = 4 * nch;
    SCRUST_GLOBAL u32x4* rp = global_ptr(reinterpret_cast<u32x4*>(s.pack_ref));
    SCRUST_GLOBAL u32x4* qp = global_ptr(reinterpret_cast<u32x4*>(s.pack_query)) + static_cast<size_t>(r0) * chunks;
    for (uint32_t u = tid; u < prep_rows * chunks; u += 256) {
        const uint32_t row = u / chunks;
        const int chunk = static_cast<int>(u - row * chunks);
        const uint32_t ref_row = r0 + row;
        const size_t ref_at = static_cast<size_t>(ref_row / kTileRefsHost) * (kTileRefsHost * chunks) +
                              tile_chunk(static_cast<int>(ref_row % kTileRefsHost), chunk / nch, chunk % nch, nch);
        uint32_t rw[4];
        if (row < rows) {
            const float* c_row = xs + row * d;
            const float nrm = nrm_s[row];
            const uint16_t nh = bf16_rn(nrm);
            const float rest = nrm - bf16_value(nh);
            const uint16_t nl = bf16_rn(rest);
            const uint16_t nl2 = bf16_rn(rest - bf16_value(nl));
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                uint16_t rv[2];
#pragma unroll
                for (int h2 = 0; h2 < 2; h2++) {
                    const int f = k_index(chunk, e + h2, nch);
                    rv[h2] = 0;
                    if (f < d || (f >= d + 3 && f < 3 * d + 3)) {
                        // part 0: h.h, part 1: h_query.l_reference, part 2: l_query.h_reference
                        const int part = f < d ? 0 : (f < 2 * d + 3 ? 1 : 2);
                        const float c = c_row[part == 0 ? f : f - 3 - part * d];
                        const uint16_t h = bf16_rn(c);
                        rv[h2] = part == 1 ? bf16_rn(c - bf16_value(h)) : h;
                    } else if (f < d + 3) {
                        rv[h2] = f == d ? nh : (f == d + 1 ? nl : nl2);
                    }
                }
                rw[e >> 1] = static_cast<uint32_t>(rv[0]) | static_cast<uint32_t>(rv[1]) << 16;
            }
            rp[ref_at] = u32x4{rw[0], rw[1], rw[2], rw[3]};
            // The persistent pass forms its query fragments from the coordinates itself
            // (bf16_query_chunk) and has no query operand array.
            if (s.pack_query) qp[u] = bf16_query_chunk(c_row, d, chunk, nch);
        } else {
            // Sentinel reference row: +inf at the norm feature (K index d), zero elsewhere.
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                const int f = k_index(chunk, e, nch);
                rw[e >> 1] = (f == d ? kBf16Inf : 0u) | (f + 1 == d ? static_cast<uint32_t>(kBf16Inf) << 16 : 0u);
            }
            rp[ref_at] = u32x4{rw[0], rw[1], rw[2], rw[3]};
        }
    }
}

// Local-origin pass: one workgroup per kLocalRows query rows. The origin o is the mean of
// the superblock's rows (any f32 value serves: it only shifts both operands of every
// product); then each row's a'' = (c - o) sigma, computed as fma(c, sigma, -o sigma), which
// equals sigma fl(c - o) exactly (sigma is a power of two). Query operand, 32 nch fp16
// features: [a'' | 0 | -2^15 -2^15 q_e] with q_e >= local_c |a''| + local_c2 rounded up;
// the reference side puts |b''|^2 2^-16 (hi, lo) and an upper bound of |b''| in the same
// three places (knn_pass.hip). anorm holds a lower bound of |a''|^2.
__global__ __launch_bounds__(256) void local_queries_kernel(const DatasetDesc* __restrict__ sets) {
    __shared__ float o_s[kMaxPackedD + 32];
    const DatasetDesc& s = sets[blockIdx.y];
    if (!s.on_pass || !s.local) return;
    const uint32_t n = s.n, r0 = blockIdx.x * kLocalRows;
    if (r0 >= s.npad) return;
    const int d = static_cast<int>(s.d), width = 32 * static_cast<int>(s.nch);
    const uint32_t rows = r0 < n ? (n - r0 < kLocalRows ? n - r0 : kLocalRows) : 0u;
    const int tid = threadIdx.x;
    const SCRUST_GLOBAL float* coords = global_ptr(s.coords);
    if (tid < width) {
        float sum = 0.0f;
        if (tid < d)
            for (uint32_t r = 0; r < rows; r++) sum += coords[static_cast<size_t>(r0 + r) * d + tid];
        const float o = (tid < d && rows > 0) ? sum / static_cast<float>(rows) : 0.0f;
        o_s[tid] = o;
        global_ptr(s.origins)[static_cast<size_t>(blockIdx.x) * width + tid] = o;
    }
    __syncthreads();
    const float sigma = local_scale(__uint_as_float(*s.max_norm_bits));
    for (uint32_t r = tid; r < rows; r += 256) {
        const uint32_t row = r0 + r;
        const SCRUST_GLOBAL float* c = coords + static_cast<size_t>(row) * d;
        float nrm = 0.0f;
        for (int j = 0; j < d; j++) {
            const float a = __builtin_fmaf(c[j], sigma, -(o_s[j] * sigma));
            nrm = __builtin_fmaf(a, a, nrm);
        }
        // |a''|^2 within (1 +- 2^-16) of nrm (d fma roundings), so these bound it.
        const float lower = nrm * (1.0f - 0x1p-15f);
        const float len_up = __builtin_sqrtf(nrm * (1.0f + 0x1p-15f)) * (1.0f + 0x1p-20f);
        global_ptr(s.anorm)[row] = lower;
        const _Float16 qe = f16_up((s.local_c * len_up + s.local_c2) * (1.0f + 0x1p-20f));
        SCRUST_GLOBAL u32x4* dst = reinterpret_cast<SCRUST_GLOBAL u32x4*>(global_ptr(s.pack_query)) +
                                   static_cast<size_t>(row) * (width / 8);
        for (int f0 = 0; f0 < width; f0 += 8) {
            f16x8 h;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int f = f0 + e;
                float v = 0.0f;
                if (f < d) v = __builtin_fmaf(c[f], sigma, -(o_s[f] * sigma));
                h[e] = static_cast<_Float16>(v);
                if (f == width - 3 || f == width - 2) h[e] = static_cast<_Float16>(-32768.0f);
                if (f == width - 1) h[e] = qe;
            }
            dst[f0 / 8] = __builtin_bit_cast(u32x4, h);
        }
    }
}

void launch_prepare_all(const DatasetDesc* sets, const BatchShape& shape, unsigned* flags,
                        hipStream_t stream) {
    hipLaunchKernelGGL(clear_sums_kernel, dim3(shape.count), dim3(256), 0, stream, sets);
    if (shape.metric == kMetricEuclidean) {
        dim3 grid(shape.max_row_chunks, shape.max_col_groups, shape.count);
        hipLaunchKernelGGL(colsum_kernel, grid, dim3(kColsumWaves * kWave), 0, stream, sets);
        hipLaunchKernelGGL(colmean_kernel, dim3(shape.max_col_groups, shape.count), dim3(kWave), 0,
                           stream, sets);
    }
    // Datasets on a packed 16-bit pass are prepared by prepare_rows_kernel alone.
    const unsigned blocks = (shape.total_prep + 255) / 256;
    if (shape.direct_sets > 0)
        hipLaunchKernelGGL(prepare_kernel, dim3(blocks), dim3(256), 0, stream, sets, shape.count,
                           shape.total_prep, shape.metric, shape.operand, flags);
    if ((shape.operand == kOperandBf16 || shape.operand == kOperandF16Local) && shape.max_prep_blocks > 0)
        hipLaunchKernelGGL(prepare_rows_kernel, dim3(shape.max_prep_blocks, 1, shape.count), dim3(256), 0,
                           stream, sets, shape.metric, flags);
    if (shape.operand == kOperandF16Local && shape.max_local_blocks > 0)
        hipLaunchKernelGGL(local_queries_kernel, dim3(shape.max_local_blocks, shape.count), dim3(256), 0,
                           stream, sets);
}

// ===========================================================================
// 4. Rerank (crates/scrust-gpu/src/kernels/knn.rs:228-244 arithmetic; order by
//    (d^2, id) as knn.rs:179-182 / neighbors.rs:181-185), fp64 boundary rule and the
//    completeness certificate. KP lanes per query row.
// ===========================================================================

// Certificate of one pass row whose kept list was sorted: every reference outside the list
// has approximate score >= theta, so its fp32 squared distance is >= `lower`; if that
// clears the band's upper edge t + tau and every snapping threshold, the list holds every
// reference the oracle could place in the first k (or in the tie band).
// A list shorter than KP holds every reference appended over the whole scan; it is complete
// only if that is every other cell (scores that overflowed are never appended). A dataset
// past the pass's norm limit skipped the pass: every row is listed.
__device__ __forceinline__ bool row_certified(const DatasetDesc& s, uint32_t p, float nq, int d, double t,
                                              double tau, bool full, int n_valid) {
    const float max_norm = __uint_as_float(*s.max_norm_bits);
    bool certified = full || n_valid == static_cast<int>(s.n) - 1;
    if (!(max_norm <= kPassNormLimit)) certified = false;
    if (full && certified) {
        const double u = 0x1p-24;
        const double nqd = static_cast<double>(nq);
        const double nmax = static_cast<double>(max_norm);
        const double gamma = d * u / (1.0 - d * u);
        const double qn = __builtin_sqrt(nqd * (1.0 + gamma));
        const double rn = __builtin_sqrt(nmax * (1.0 + gamma));
        // |score + |a|^2 - exact d^2| <= err: f32 norms, the bf16 split of the
        // coordinates and of |b|^2, and the MFMA accumulation over every packed term.
        const double err = 1.25 * (gamma * (nqd + nmax) * 1.01 + 2.0 * s.split_rel * qn * rn +
                                   2.0 * s.acc_rel * (1.02 * qn * rn + 0.51 * nmax * (1.0 + gamma)) +
                                   s.norm_rel * nmax * (1.0 + gamma)) + 1e-30;
        const double cd = (d + 3) * u * 1.01;
        double lower = (static_cast<double>(s.theta[p]) + nqd - err) * (1.0 - cd);
        if (s.local) {
            // Local-origin pass, in units of sigma (local_scale): a reference not kept has
            // pass score >= theta, and the score is a lower bound of |b''|^2 - 2 a''.b'' up
            // to the per-row term 2 kq (fp16 rounding below the normal range), so
            // |a'' - b''|^2 >= theta + |a''|^2 - 2 kq. Then a' = fl(a - o) differs from a - o
            // by at most u |a - o| per coordinate: |a - b| >= |a' - b'| (1 - 2.02 u) - 2.02 u |a'|.
            const double sigma = static_cast<double>(local_scale(max_norm));
            const double an2 = static_cast<double>(s.anorm[p]);
            const double an_up = __builtin_sqrt(an2 / (1.0 - 0x1p-15) * (1.0 + 0x1p-16)) * (1.0 + 0x1p-40);
            const double kq = static_cast<double>(s.local_c2) * an_up + d * 0x1p-50;
            const double base = static_cast<double>(s.theta[p]) + an2 - 2.0 * kq;
            double len = base > 0.0 ? __builtin_sqrt(base) / sigma : 0.0;
            len = len * (1.0 - 2.02 * u) - 2.02 * u * (an_up / sigma);
            lower = len > 0.0 ? len * len * (1.0 - cd) : 0.0;
        }
        const double snap_max = (d + 2) * 0x1p-23 * (nqd + nmax) * (1.0 + 4.0 * u);
        const double need = (t + tau) > snap_max ? (t + tau) : snap_max;
        certified = lower > need;
    }
    return certified;
}

// f32 direct-difference squared distance of pass rows p and c (the rerank's arithmetic).
__device__ __forceinline__ float rerank_square(const DatasetDesc& s, const float* xq, float nq, uint32_t c, int d) {
    const float* xr = s.coords + static_cast<size_t>(c) * d;
    // Rows are 16- or 8-byte aligned when 4 or 2 divides d (the arena is 256-byte aligned).
    if ((d & 3) == 0) return direct_square_vec<4>(xq, xr, d, nq, s.norms[c]);
    if ((d & 1) == 0) return direct_square_vec<2>(xq, xr, d, nq, s.norms[c]);
    return direct_square_vec<1>(xq, xr, d, nq, s.norms[c]);
}

template <int KP>
__global__ __launch_bounds__(256) void rerank_kernel(const DatasetDesc* __restrict__ sets,
                                                     uint32_t count, uint32_t total_rows,
                                                     int metric, uint32_t* __restrict__ fail_list,
                                                     uint32_t* __restrict__ fail_count) {
    constexpr int QPW = kWave / KP;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int seg = lane & ~(KP - 1);
    const int j = lane & (KP - 1);
    const uint32_t flat0 = (blockIdx.x * 4 + w) * QPW + lane / KP;
    const uint32_t flat = flat0 < total_rows ? flat0 : total_rows - 1;
    const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
    const bool row_ok = flat0 < total_rows && s.on_pass && s.kp == KP;
    // Rows are visited in pass order: p indexes coordinates, norms, candidates and
    // thresholds; q = perm[p] is the input row the output belongs to. Candidate ids are
    // pass positions; keys carry input ids, so order and ties are the input's.
    const uint32_t p = flat - s.row_begin;
    const uint32_t n = s.n;
    const uint32_t q = row_ok && s.perm ? s.perm[p] : p;
    const int d = static_cast<int>(s.d);
    const float* xq = s.coords + static_cast<size_t>(p) * d;
    const float nq = s.norms[p];

    uint64_t raw = kNoKey;
    if (row_ok) raw = s.cand[static_cast<size_t>(p) * KP + j];
    const bool present = raw != kNoKey;
    const uint32_t c = key_id(raw);
    const bool valid = present && c != p && c < n;
    const float dsq = valid ? rerank_square(s, xq, nq, c, d) : kInf;
    uint64_t key = valid ? make_key(dsq, s.perm ? s.perm[c] : c) : kNoKey;
    key = segment_sort<KP>(key, lane);

    // Segment-local bookkeeping (every lane of a segment computes the same values).
    const uint64_t seg_mask = (KP == 64) ? ~0ull : (((1ull << (KP & 63)) - 1) << seg);
    const int n_valid = __popcll(__ballot(key != kNoKey) & seg_mask);
    const bool full = (__ballot(present) & seg_mask) == seg_mask;
    const float my_d = key_value(key);
    const uint32_t my_id = key_id(key);
    const int k = static_cast<int>(s.k);
    const int kk = row_ok ? k : 1;
    const float t_f = __shfl(my_d, seg + kk - 1, kWave);
    const double t = static_cast<double>(t_f);
    const double tau = t * static_cast<double>(d + 2) * 0x1p-23;
    const bool in_band = (j < n_valid) && __builtin_fabs(static_cast<double>(my_d) - t) <= tau;
    const uint64_t band_mask = __ballot(in_band) & seg_mask;
    const bool has_next = n_valid > kk;
    const double next = static_cast<double>(__shfl(my_d, seg + (kk < KP ? kk : KP - 1), kWave));
    const bool tie = row_ok && has_next && (next - t <= tau);

    int pos = j;
    if (__ballot(tie) != 0) {
        const int lo = band_mask ? __builtin_ctzll(band_mask) - seg : 0;
        const int hi = band_mask ? 63 - __builtin_clzll(band_mask) - seg : -1;
        double d64 = 0.0;
        if (tie && in_band && my_d != 0.0f) {
            const uint32_t mp = s.inv_perm ? s.inv_perm[my_id] : my_id;
            d64 = direct_square_f64(xq, s.coords + static_cast<size_t>(mp) * d, d);
        }
        int rank = 0;
        for (int i = 0; i < KP; i++) {
            const double od = __shfl(d64, seg + i, kWave);
            const uint32_t oid = __shfl(my_id, seg + i, kWave);
            if (i >= lo && i <= hi && (od < d64 || (od == d64 && oid < my_id))) rank++;
        }
        if (tie && in_band) pos = lo + rank;
    }
    if (!row_ok) return;

    const bool certified = row_certified(s, p, nq, d, t, tau, full, n_valid);
    if (certified) {
        if (pos < k && j < n_valid) {
            s.out_indices[static_cast<size_t>(q) * k + pos] = my_id;
            s.out_distances[static_cast<size_t>(q) * k + pos] =
                metric == kMetricCosine ? 0.5f * my_d : __builtin_sqrtf(my_d);
        }
    } else if (j == 0) {
        // Listed as the input row (the exact kernel works in input rows).
        const uint32_t slot = atomicAdd(fail_count, 1u);
        fail_list[slot] = s.row_begin + q;
    }
}

// Rerank of rows with KP = 64 E candidates (128 or 256): one wave per row, E keys per lane
// (sorted key e = i*64 + lane), otherwise as rerank_kernel. A tie band here can exceed
// kBandCap members; such a band keeps the (d^2, id) order, as the oracle's rule has it.
template <int E>
__global__ __launch_bounds__(256) void rerank_wide_kernel(const DatasetDesc* __restrict__ sets,
                                                          uint32_t count, uint32_t total_rows,
                                                          int metric, uint32_t* __restrict__ fail_list,
                                                          uint32_t* __restrict__ fail_count) {
    constexpr int KP = E * kWave;
    const int lane = threadIdx.x & 63;
    const uint32_t flat = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (flat >= total_rows) return;
    const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
    if (!s.on_pass || s.kp != KP) return;
    const uint32_t p = flat - s.row_begin;
    const uint32_t n = s.n;
    const uint32_t q = s.perm ? s.perm[p] : p;
    const int d = static_cast<int>(s.d);
    const float* xq = s.coords + static_cast<size_t>(p) * d;
    const float nq = s.norms[p];

    uint64_t key[E];
    bool full = true;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint64_t raw = s.cand[static_cast<size_t>(p) * KP + i * kWave + lane];
        const bool present = raw != kNoKey;
        full = full && __ballot(present) == ~0ull;
        const uint32_t c = key_id(raw);
        const bool valid = present && c != p && c < n;
        const float dsq = valid ? rerank_square(s, xq, nq, c, d) : kInf;
        key[i] = valid ? make_key(dsq, s.perm ? s.perm[c] : c) : kNoKey;
    }
    wave_sort<E>(key, lane);
    int n_valid = 0;
    float my_d[E];
    uint32_t my_id[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        n_valid += __popcll(__ballot(key[i] != kNoKey));
        my_d[i] = key_value(key[i]);
        my_id[i] = key_id(key[i]);
    }
    // Sorted key e's distance and id (e wave-uniform: word e / 64, lane e % 64).
    auto dist_at = [&](int e) {
        float v = 0.0f;
#pragma unroll
        for (int i = 0; i < E; i++)
            if (e / kWave == i) v = __shfl(my_d[i], e % kWave, kWave);
        return v;
    };
    auto id_at = [&](int e) {
        uint32_t v = 0u;
#pragma unroll
        for (int i = 0; i < E; i++)
            if (e / kWave == i) v = static_cast<uint32_t>(__shfl(static_cast<int>(my_id[i]), e % kWave, kWave));
        return v;
    };
    const int k = static_cast<int>(s.k);
    const double t = static_cast<double>(dist_at(k - 1));
    const double tau = t * static_cast<double>(d + 2) * 0x1p-23;
    const bool tie = n_valid > k && (static_cast<double>(dist_at(k)) - t <= tau);

    int pos[E];
#pragma unroll
    for (int i = 0; i < E; i++) pos[i] = i * kWave + lane;
    if (tie) {
        bool in_band[E];
        uint64_t band[E];
        int len = 0;
#pragma unroll
        for (int i = 0; i < E; i++) {
            in_band[i] = i * kWave + lane < n_valid && __builtin_fabs(static_cast<double>(my_d[i]) - t) <= tau;
            band[i] = __ballot(in_band[i]);
            len += __popcll(band[i]);
        }
        if (len <= kBandCap) {
            // First and last sorted positions of the band (it is contiguous in the sort).
            int lo = -1, hi = -1;
#pragma unroll
            for (int i = 0; i < E; i++) {
                if (band[i] && lo < 0) lo = i * kWave + __builtin_ctzll(band[i]);
                if (band[i]) hi = i * kWave + 63 - __builtin_clzll(band[i]);
            }
            double d64[E];
#pragma unroll
            for (int i = 0; i < E; i++) {
                d64[i] = 0.0;
                if (in_band[i] && my_d[i] != 0.0f) {
                    const uint32_t mp = s.inv_perm ? s.inv_perm[my_id[i]] : my_id[i];
                    d64[i] = direct_square_f64(xq, s.coords + static_cast<size_t>(mp) * d, d);
                }
            }
            int rank[E];
#pragma unroll
            for (int i = 0; i < E; i++) rank[i] = 0;
            for (int e = lo; e <= hi; e++) {
                double od = 0.0;
#pragma unroll
                for (int i = 0; i < E; i++)
                    if (e / kWave == i) od = __shfl(d64[i], e % kWave, kWave);
                const uint32_t oid = id_at(e);
#pragma unroll
                for (int i = 0; i < E; i++)
                    if (od < d64[i] || (od == d64[i] && oid < my_id[i])) rank[i]++;
            }
#pragma unroll
            for (int i = 0; i < E; i++)
                if (in_band[i]) pos[i] = lo + rank[i];
        }
    }
    if (row_certified(s, p, nq, d, t, tau, full, n_valid)) {
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (pos[i] < k && i * kWave + lane < n_valid) {
                s.out_indices[static_cast<size_t>(q) * k + pos[i]] = my_id[i];
                s.out_distances[static_cast<size_t>(q) * k + pos[i]] =
                    metric == kMetricCosine ? 0.5f * my_d[i] : __builtin_sqrtf(my_d[i]);
            }
        }
    } else if (lane == 0) {
        const uint32_t slot = atomicAdd(fail_count, 1u);
        fail_list[slot] = s.row_begin + q;
    }
}

void launch_rerank(const DatasetDesc* sets, const BatchShape& shape, unsigned kps,
                   uint32_t* fail_list, uint32_t* fail_count, hipStream_t stream) {
    if (shape.total_rows == 0) return;
    if (kps & 128u) {
        const unsigned blocks = (shape.total_rows + 3) / 4;
        hipLaunchKernelGGL(rerank_wide_kernel<2>, dim3(blocks), dim3(256), 0, stream, sets, shape.count,
                           shape.total_rows, shape.metric, fail_list, fail_count);
    }
    if (kps & 256u) {
        const unsigned blocks = (shape.total_rows + 3) / 4;
        hipLaunchKernelGGL(rerank_wide_kernel<4>, dim3(blocks), dim3(256), 0, stream, sets, shape.count,
                           shape.total_rows, shape.metric, fail_list, fail_count);
    }
    if (kps & 32u) {
        const unsigned blocks = (shape.total_rows + 7) / 8;
        hipLaunchKernelGGL(rerank_kernel<32>, dim3(blocks), dim3(256), 0, stream, sets, shape.count,
                           shape.total_rows, shape.metric, fail_list, fail_count);
    }
    if (kps & 64u) {
        const unsigned blocks = (shape.total_rows + 3) / 4;
        hipLaunchKernelGGL(rerank_kernel<64>, dim3(blocks), dim3(256), 0, stream, sets, shape.count,
                           shape.total_rows, shape.metric, fail_list, fail_count);
    }
}

// ===========================================================================
// 5. Exact rows: one workgroup per query row scans every reference with the f32
//    direct difference (the Metal kernel's per-query scan, knn.rs:222-244), keeping
//    the smallest keys in LDS; then the fp64 boundary rule.
// ===========================================================================

constexpr int kExactBuf = 1024;
constexpr int kExactKeep = 512;

// One thread: the first k of `m` keys sorted by (d^2, id) into out_id / out_d, then the fp64
// boundary rule (bd / bid / bf: kBandCap entries of scratch). `dropped`: keys past the m kept
// ones were discarded, so a band reaching the last kept key may be incomplete (flagged).
__device__ void band_order(const DatasetDesc& s, const uint64_t* keys, int m, int k, bool dropped,
                           const float* xq, uint32_t* out_id, float* out_d, double* bd, uint32_t* bid,
                           float* bf, unsigned* flags) {
    const int d = static_cast<int>(s.d);
    for (int j = 0; j < k; j++) {
        out_id[j] = key_id(keys[j]);
        out_d[j] = key_value(keys[j]);
    }
    const double t = static_cast<double>(key_value(keys[k - 1]));
    const double tau = t * static_cast<double>(d + 2) * 0x1p-23;
    // With t == 0 every band member was snapped to zero and keeps (0, id) order.
    if (t > 0.0 && m > k && static_cast<double>(key_value(keys[k])) - t <= tau) {
        int lo = k - 1;
        while (lo > 0 && t - static_cast<double>(key_value(keys[lo - 1])) <= tau) lo--;
        int hi = k;
        while (hi + 1 < m && hi - lo + 1 <= kBandCap && static_cast<double>(key_value(keys[hi + 1])) - t <= tau)
            hi++;
        // A band wider than kBandCap keeps the (d^2, id) order. With keys dropped, m is at
        // least k + kBandCap (kExactKeep, or kSplitKeep with k <= kMaxSplitK), so a band that reaches
        // the end of the kept keys is always wider than the cap.
        const int len = hi - lo + 1;
        if (len <= kBandCap && hi + 1 == m && dropped) {
            atomicOr(flags, kFlagBandOverflow);
        } else if (len <= kBandCap) {
            // Insertion sort of the band by (fp64 d^2, id); the band is small.
            for (int i = 0; i < len; i++) {
                const uint64_t key = keys[lo + i];
                const float v = key_value(key);
                const uint32_t id = key_id(key);
                const uint32_t at = s.inv_perm ? s.inv_perm[id] : id;
                const double v64 = v == 0.0f ? 0.0 : direct_square_f64(xq, s.coords + static_cast<size_t>(at) * d, d);
                int slot = i;
                while (slot > 0 && (bd[slot - 1] > v64 || (bd[slot - 1] == v64 && bid[slot - 1] > id))) {
                    bd[slot] = bd[slot - 1];
                    bid[slot] = bid[slot - 1];
                    bf[slot] = bf[slot - 1];
                    slot--;
                }
                bd[slot] = v64;
                bid[slot] = id;
                bf[slot] = v;
            }
            for (int p = lo; p < k; p++) {
                out_id[p] = bid[p - lo];
                out_d[p] = bf[p - lo];
            }
        }
    }
}

__device__ void lds_sort(uint64_t* keys, int count) {
    // Bitonic sort of kExactBuf keys by the workgroup (entries past count are kNoKey).
    for (int i = threadIdx.x; i < kExactBuf; i += blockDim.x)
        if (i >= count) keys[i] = kNoKey;
    __syncthreads();
    for (int size = 2; size <= kExactBuf; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int p = threadIdx.x; p < kExactBuf / 2; p += blockDim.x) {
                const int lo = (p / stride) * stride * 2 + (p % stride);
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                const uint64_t x = keys[lo], y = keys[hi];
                if ((x > y) == asc) {
                    keys[lo] = y;
                    keys[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void exact_kernel(const DatasetDesc* __restrict__ sets,
                                                    uint32_t count, uint32_t total_rows,
                                                    int metric, const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ list_count,
                                                    unsigned* __restrict__ flags, bool split_on) {
    extern __shared__ float xq[];
    __shared__ uint64_t keys[kExactBuf];
    __shared__ uint32_t kept;
    __shared__ uint64_t theta;
    __shared__ int dropped;
    __shared__ uint32_t out_id[kMaxK];
    __shared__ float out_d[kMaxK];
    __shared__ double bd[kBandCap];
    __shared__ uint32_t bid[kBandCap];
    __shared__ float bf[kBandCap];
    const int tid = threadIdx.x;
    const uint32_t total = list ? *list_count : total_rows;
    const bool split = split_on && total <= kSplitRowsMax;
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const uint32_t flat = list ? list[item] : item;
        const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
        if (split && s.on_pass && s.k <= kMaxSplitK) continue;  // exact_split_kernel's
        const uint32_t q = flat - s.row_begin;
        const int d = static_cast<int>(s.d), k = static_cast<int>(s.k);
        const uint32_t n = s.n;
        // Coordinates and norms of an ordered dataset are in pass order: pq is the row's
        // pass position, reference r (a pass position) has input id perm[r].
        const uint32_t pq = s.inv_perm ? s.inv_perm[q] : q;
        for (int j = tid; j < d; j += blockDim.x) xq[j] = s.coords[static_cast<size_t>(pq) * d + j];
        if (tid == 0) {
            kept = 0;
            theta = kNoKey;
            dropped = 0;
        }
        __syncthreads();
        const float nq = s.norms[pq];
        for (uint32_t r0 = 0; r0 < n; r0 += blockDim.x) {
            const uint32_t r = r0 + tid;
            if (r < n && r != pq) {
                const float v = direct_square(xq, s.coords + static_cast<size_t>(r) * d, d, nq, s.norms[r]);
                const uint64_t key = make_key(v, s.perm ? s.perm[r] : r);
                if (key < theta) {
                    const uint32_t slot = atomicAdd(&kept, 1u);
                    keys[slot] = key;
                }
            }
            __syncthreads();
            if (kept > static_cast<uint32_t>(kExactBuf) - blockDim.x) {
                lds_sort(keys, kept);
                if (tid == 0) {
                    kept = kExactKeep;
                    theta = keys[kExactKeep - 1];
                    dropped = 1;
                }
                __syncthreads();
            }
        }
        const int m = static_cast<int>(kept);
        lds_sort(keys, m);
        if (tid == 0) band_order(s, keys, m, k, dropped != 0, xq, out_id, out_d, bd, bid, bf, flags);
        __syncthreads();
        for (int j = tid; j < k; j += blockDim.x) {
            s.out_indices[static_cast<size_t>(q) * k + j] = out_id[j];
            s.out_distances[static_cast<size_t>(q) * k + j] =
                metric == kMetricCosine ? 0.5f * out_d[j] : __builtin_sqrtf(out_d[j]);
        }
        __syncthreads();
    }
}

// Split scan of a short exact list (pass datasets' rows): wave item = (list entry i, range j)
// over kSplitRange references, keeping its kSplitKeep smallest keys in `partial`. A wave
// holds its kept keys kSplitHeld per lane and takes kSplitE - kSplitHeld new keys per lane at
// a time; when more than kSplitKeep are held it keeps the kSplitKeep smallest (select_keys,
// through LDS) and drops later keys scoring above the kept maximum. Keys are (f32
// direct-difference d^2, input id), as in exact_kernel.
constexpr int kSplitE = 16;
constexpr int kSplitHeld = static_cast<int>(kSplitKeep) / kWave;
static_assert(kSplitHeld * kWave == static_cast<int>(kSplitKeep) && kSplitHeld < kSplitE, "split scan key layout");

__device__ void split_keep(uint64_t (&key)[kSplitE], float& cut, int lane, uint64_t* stage) {
    uint32_t total = 0;
#pragma unroll
    for (int e = 0; e < kSplitE; e++) total += static_cast<uint32_t>(__popcll(__ballot(key[e] != kNoKey)));
    if (total > kSplitKeep) {
        cut = select_keys<kSplitE>(key, static_cast<int>(kSplitKeep), lane, stage, 0u, 0u);
    } else {
        uint32_t base = 0;
#pragma unroll
        for (int e = 0; e < kSplitE; e++) {
            const uint64_t m = __ballot(key[e] != kNoKey);
            if (key[e] != kNoKey) stage[base + lanes_below(m)] = key[e];
            base += static_cast<uint32_t>(__popcll(m));
        }
        for (uint32_t p = base + static_cast<uint32_t>(lane); p < kSplitKeep; p += kWave) stage[p] = kNoKey;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int e = 0; e < kSplitE; e++) key[e] = e < kSplitHeld ? stage[e * kWave + lane] : kNoKey;
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(256) void exact_split_kernel(const DatasetDesc* __restrict__ sets, uint32_t count,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ list_count,
                                                          uint32_t ranges_max, uint64_t* __restrict__ partial) {
    __shared__ uint64_t stage_all[4][kSplitKeep];
    const uint32_t total = *list_count;
    if (total > kSplitRowsMax) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t* stage = stage_all[w];
    const uint32_t waves = gridDim.x * 4;
    for (uint32_t item = blockIdx.x * 4 + w; item < total * ranges_max; item += waves) {
        const uint32_t i = item / ranges_max, j = item - i * ranges_max;
        const uint32_t flat = list[i];
        const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
        if (!s.on_pass || s.k > kMaxSplitK) continue;
        const uint32_t n = s.n, q = flat - s.row_begin;
        const uint32_t pq = s.inv_perm ? s.inv_perm[q] : q;
        const int d = static_cast<int>(s.d);
        const float* xq = s.coords + static_cast<size_t>(pq) * d;
        const float nq = s.norms[pq];
        uint64_t key[kSplitE];
#pragma unroll
        for (int e = 0; e < kSplitE; e++) key[e] = kNoKey;
        float cut = kInf;
        const uint32_t r0 = j * kSplitRange;
        const uint32_t r1 = r0 < n ? (n - r0 < kSplitRange ? n : r0 + kSplitRange) : r0;
        for (uint32_t b = r0; b < r1; b += (kSplitE - kSplitHeld) * kWave) {
#pragma unroll
            for (int e = kSplitHeld; e < kSplitE; e++) {
                const uint32_t r = b + static_cast<uint32_t>(e - kSplitHeld) * kWave + static_cast<uint32_t>(lane);
                uint64_t kk = kNoKey;
                if (r < r1 && r != pq) {
                    const float v = direct_square(xq, s.coords + static_cast<size_t>(r) * d, d, nq, s.norms[r]);
                    if (!(v > cut)) kk = make_key(v, s.perm ? s.perm[r] : r);
                }
                key[e] = kk;
            }
            split_keep(key, cut, lane, stage);
        }
        uint64_t* out = partial + static_cast<size_t>(item) * kSplitKeep;
#pragma unroll
        for (int e = 0; e < kSplitHeld; e++) out[e * kWave + lane] = key[e];
    }
}

// One wave per listed row: the kSplitKeep smallest keys over the row's ranges, sorted, then
// the first k and the fp64 boundary rule (band_order) into the outputs.
__global__ __launch_bounds__(256) void exact_merge_kernel(const DatasetDesc* __restrict__ sets, uint32_t count,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ list_count,
                                                          uint32_t ranges_max, const uint64_t* __restrict__ partial,
                                                          int metric, unsigned* __restrict__ flags) {
    __shared__ uint64_t stage_all[4][kSplitKeep];
    __shared__ double bd_all[4][kBandCap];
    __shared__ uint32_t bid_all[4][kBandCap];
    __shared__ float bf_all[4][kBandCap];
    __shared__ uint32_t id_all[4][kMaxSplitK];
    __shared__ float dist_all[4][kMaxSplitK];
    const uint32_t total = *list_count;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * 4 + w;
    if (total > kSplitRowsMax || i >= total) return;
    const uint32_t flat = list[i];
    const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
    if (!s.on_pass || s.k > kMaxSplitK) return;
    uint64_t* stage = stage_all[w];
    const uint32_t n = s.n, q = flat - s.row_begin, k = s.k;
    const uint32_t pq = s.inv_perm ? s.inv_perm[q] : q;
    const uint32_t ranges = (n + kSplitRange - 1) / kSplitRange;
    const uint32_t keys_in = ranges * kSplitKeep;
    const uint64_t* src = partial + static_cast<size_t>(i) * ranges_max * kSplitKeep;
    uint64_t key[kSplitE];
#pragma unroll
    for (int e = 0; e < kSplitE; e++) key[e] = kNoKey;
    float cut = kInf;
    for (uint32_t b = 0; b < keys_in; b += (kSplitE - kSplitHeld) * kWave) {
#pragma unroll
        for (int e = kSplitHeld; e < kSplitE; e++) {
            const uint32_t p = b + static_cast<uint32_t>(e - kSplitHeld) * kWave + static_cast<uint32_t>(lane);
            const uint64_t kk = p < keys_in ? src[p] : kNoKey;
            key[e] = (kk != kNoKey && !(key_value(kk) > cut)) ? kk : kNoKey;
        }
        split_keep(key, cut, lane, stage);
    }
    // Sort the kept keys (kSplitHeld per lane) and finish on lane 0.
    uint64_t held[kSplitHeld];
#pragma unroll
    for (int e = 0; e < kSplitHeld; e++) held[e] = key[e];
    wave_sort<kSplitHeld>(held, lane);
#pragma unroll
    for (int e = 0; e < kSplitHeld; e++) stage[e * kWave + lane] = held[e];
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        int m = 0;
        while (m < static_cast<int>(kSplitKeep) && stage[m] != kNoKey) m++;
        const float* xq = s.coords + static_cast<size_t>(pq) * s.d;
        band_order(s, stage, m, static_cast<int>(k), n - 1 > kSplitKeep, xq, id_all[w], dist_all[w], bd_all[w],
                   bid_all[w], bf_all[w], flags);
        for (uint32_t j = 0; j < k; j++) {
            s.out_indices[static_cast<size_t>(q) * k + j] = id_all[w][j];
            s.out_distances[static_cast<size_t>(q) * k + j] =
                metric == kMetricCosine ? 0.5f * dist_all[w][j] : __builtin_sqrtf(dist_all[w][j]);
        }
    }
}

void launch_exact(const DatasetDesc* sets, const BatchShape& shape, const uint32_t* list,
                  const uint32_t* list_count, unsigned blocks, unsigned* flags,
                  uint64_t* partial, uint32_t ranges_max, hipStream_t stream) {
    if (blocks == 0) return;
    const bool split = partial != nullptr && ranges_max > 0;
    if (split) {
        hipLaunchKernelGGL(exact_split_kernel, dim3(1024), dim3(256), 0, stream, sets, shape.count, list, list_count,
                           ranges_max, partial);
        hipLaunchKernelGGL(exact_merge_kernel, dim3(kSplitRowsMax / 4), dim3(256), 0, stream, sets, shape.count, list,
                           list_count, ranges_max, static_cast<const uint64_t*>(partial), shape.metric, flags);
    }
    hipLaunchKernelGGL(exact_kernel, dim3(blocks), dim3(256), shape.max_d * sizeof(float), stream,
                       sets, shape.count, shape.total_rows, shape.metric, list, list_count, flags, split);
}

}  // namespace scrust_hip
