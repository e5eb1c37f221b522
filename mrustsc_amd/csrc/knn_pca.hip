// PCA by randomised SVD on the device: the producer of the embedding the kNN graph is built
// on (crates/scrust-core/src/pca.rs:74-215, bound as _scrust.pca in
// crates/scrust-py/src/preprocess.rs:154-185), with the embedding left in device memory.
//
// The steps are pca.rs's: per-gene f64 sums in row order (ColumnStats), a Gaussian sketch of
// width w = min(n_components + 24, min(n_cells, n_genes)), the range finder X_c Omega with
// orthonormalisation by two whitening passes (the w x w Gram's eigenvalues clamped at 1e-7 of
// the largest), 7 or 4 power iterations, the final Gram eigensolve of B^T = X_c^T Q, then the
// embedding Q V S, the loadings B^T V S^-1, the sign convention and the explained variance.
// Centring stays implicit: X_c Y = X Y - 1 (mu^T Y) and X_c^T Y = X^T Y - mu (1^T Y).
//
// On the device (f32): the CSR matrix is also sorted into CSC (rocPRIM radix sort of
// (gene, cell) keys) so that X^T Y is a per-gene gather in a fixed order; X Y runs one wave
// per cell, X^T Y one wave per gene, each lane holding two of the w <= 128 sketch columns;
// Grams are summed over row chunks whose partial results are added in chunk order; the
// whitening and basis transforms multiply by a w x w' matrix held in LDS. Only the w x w
// eigensolves run on the host (cyclic Jacobi in f64, as in the reference), one per
// whitening pass. Every reduction has a fixed order, so runs are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/scrust_hip.h"
#include "knn_device.hpp"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

namespace {

constexpr int kPcaMaxWidth = 128;     // sketch columns the kernels serve (2 per lane)
constexpr int kPcaOversampling = 24;  // pca.rs:21
constexpr uint32_t kChunkRows = 2048; // rows per Gram / column-sum partial
constexpr uint32_t kPcaFlagBadId = 1u;
constexpr uint32_t kPcaFlagNonFinite = 2u;

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------

__global__ void clear_word_kernel(uint32_t* p) { *p = 0u; }

// The cell ids of the sorted (gene, cell) keys.
__global__ void low_word_kernel(const uint64_t* keys, uint32_t nnz, uint32_t* out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < nnz) out[j] = static_cast<uint32_t>(keys[j]);
}

// CSR validity on the device (device input): indptr starts at 0, never decreases, ends at
// nnz = indptr[n]; every index names a gene; values are finite. A row reaching past nnz is
// flagged before any of its entries is read.
__global__ void csr_check_kernel(const uint32_t* indptr, const uint32_t* indices, const float* values,
                                 uint32_t n, uint32_t g, uint32_t nnz, uint32_t* flags) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t bad = 0;
    const uint32_t b = indptr[r], e = indptr[r + 1];
    if (e < b || e > nnz || (r == 0 && b != 0)) bad |= kPcaFlagBadId;
    for (uint32_t j = b; j < e && !(bad & kPcaFlagBadId); j++) {
        if (indices[j] >= g) bad |= kPcaFlagBadId;
        if (!__builtin_isfinite(values[j])) bad |= kPcaFlagNonFinite;
    }
    if (bad) atomicOr(flags, bad);
}

// (gene << 32 | cell) sort keys of every entry, values alongside.
__global__ void csc_keys_kernel(const uint32_t* indptr, const uint32_t* indices, const float* values, uint32_t n,
                                uint64_t* keys, float* vals) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (uint32_t j = indptr[r]; j < indptr[r + 1]; j++) {
        keys[j] = (static_cast<uint64_t>(indices[j]) << 32) | r;
        vals[j] = values[j];
    }
}

// colptr[c] = first sorted entry of gene c (a binary search), colptr[g] = nnz.
__global__ void colptr_kernel(const uint64_t* keys, uint32_t nnz, uint32_t g, uint32_t* colptr) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > g) return;
    const uint64_t key = static_cast<uint64_t>(c) << 32;
    uint32_t lo = 0, hi = nnz;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
    }
    colptr[c] = lo;
}

// Per gene: f64 sum and sum of squares in cell order (ColumnStats, pca.rs:226-255), and the
// f32 mean.
__global__ void colstats_kernel(const uint32_t* colptr, const float* cvals, uint32_t g, uint32_t n,
                                double* sums, double* squares, float* mean) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g) return;
    double s = 0.0, q = 0.0;
    for (uint32_t j = colptr[c]; j < colptr[c + 1]; j++) {
        const double v = static_cast<double>(cvals[j]);
        s += v;
        q += v * v;
    }
    sums[c] = s;
    squares[c] = q;
    mean[c] = static_cast<float>(s / static_cast<double>(n));
}

// shift[c] = sum_k mean[k] rhs[k][c]: the centring correction of X_c rhs.
__global__ void mean_times_kernel(const float* mean, const float* rhs, uint32_t g, uint32_t w, float* shift) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w) return;
    float s = 0.0f;
    for (uint32_t k = 0; k < g; k++) s = __builtin_fmaf(mean[k], rhs[static_cast<size_t>(k) * w + c], s);
    shift[c] = s;
}

// out = X rhs - 1 shift^T (sparse rows x dense rhs), one wave per output row; the same
// kernel serves X^T lhs - mean colsum^T on the CSC form (rows = genes, cols = cells,
// scale[r] = mean[r] multiplies shift).
__global__ __launch_bounds__(256) void spmm_kernel(const uint32_t* ptr, const uint32_t* idx, const float* val,
                                                   uint32_t rows, const float* rhs, uint32_t w, const float* shift,
                                                   const float* scale, float* out) {
    const uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    if (r >= rows) return;
    const uint32_t c0 = lane, c1 = lane + 64;
    const float f = scale ? scale[r] : 1.0f;
    float a0 = 0.0f, a1 = 0.0f;
    if (shift) {
        if (c0 < w) a0 = -(f * shift[c0]);
        if (c1 < w) a1 = -(f * shift[c1]);
    }
    for (uint32_t j = ptr[r]; j < ptr[r + 1]; j++) {
        const float v = val[j];
        const float* src = rhs + static_cast<size_t>(idx[j]) * w;
        if (c0 < w) a0 = __builtin_fmaf(v, src[c0], a0);
        if (c1 < w) a1 = __builtin_fmaf(v, src[c1], a1);
    }
    float* dst = out + static_cast<size_t>(r) * w;
    if (c0 < w) dst[c0] = a0;
    if (c1 < w) dst[c1] = a1;
}

// Column sums of y (rows x w), per chunk of kChunkRows rows (thread c sums column c).
__global__ void colsum_partial_kernel(const float* y, uint32_t rows, uint32_t w, double* partial) {
    const uint32_t chunk = blockIdx.x;
    const uint32_t c = threadIdx.x;
    if (c >= w) return;
    const uint32_t r0 = chunk * kChunkRows, r1 = min(rows, r0 + kChunkRows);
    double s = 0.0;
    for (uint32_t r = r0; r < r1; r++) s += static_cast<double>(y[static_cast<size_t>(r) * w + c]);
    partial[static_cast<size_t>(chunk) * w + c] = s;
}

// out[c] = sum over chunks in order (f32 result); with `center`, of (y - center)^2 partials.
__global__ void chunk_reduce_kernel(const double* partial, uint32_t chunks, uint32_t w, float* out, double* out64) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= w) return;
    double s = 0.0;
    for (uint32_t k = 0; k < chunks; k++) s += partial[static_cast<size_t>(k) * w + c];
    if (out) out[c] = static_cast<float>(s);
    if (out64) out64[c] = s;
}

// Sum over rows of (y[r][c] - center[c])^2 per chunk (the uncentred PCA's score variances).
__global__ void colvar_partial_kernel(const float* y, uint32_t rows, uint32_t w, const double* center,
                                      double* partial) {
    const uint32_t chunk = blockIdx.x;
    const uint32_t c = threadIdx.x;
    if (c >= w) return;
    const uint32_t r0 = chunk * kChunkRows, r1 = min(rows, r0 + kChunkRows);
    double s = 0.0;
    for (uint32_t r = r0; r < r1; r++) {
        const double d = static_cast<double>(y[static_cast<size_t>(r) * w + c]) - center[c];
        s += d * d;
    }
    partial[static_cast<size_t>(chunk) * w + c] = s;
}

// Partial Gram y^T y over one chunk of rows: the upper triangle, entries p = tid + 256 t
// in row-major triangle order; rows staged 32 at a time in LDS.
constexpr int kGramPer = (kPcaMaxWidth * (kPcaMaxWidth + 1) / 2 + 255) / 256;

__global__ __launch_bounds__(256) void gram_partial_kernel(const float* y, uint32_t rows, uint32_t w,
                                                           float* partial) {
    __shared__ float tile[32 * kPcaMaxWidth];
    const uint32_t chunk = blockIdx.x, tid = threadIdx.x;
    const uint32_t r0 = chunk * kChunkRows, r1 = min(rows, r0 + kChunkRows);
    const uint32_t pairs = w * (w + 1) / 2;
    int pi[kGramPer], pj[kGramPer];
    float acc[kGramPer];
#pragma unroll
    for (int t = 0; t < kGramPer; t++) {
        uint32_t p = tid + 256u * t;
        acc[t] = 0.0f;
        if (p >= pairs) {
            pi[t] = pj[t] = -1;
            continue;
        }
        // Row i of the triangle holds w - i entries.
        uint32_t i = 0;
        while (p >= w - i) {
            p -= w - i;
            i++;
        }
        pi[t] = static_cast<int>(i);
        pj[t] = static_cast<int>(i + p);
    }
    for (uint32_t b = r0; b < r1; b += 32) {
        const uint32_t count = min(32u, r1 - b);
        __syncthreads();
        for (uint32_t e = tid; e < count * w; e += 256) tile[e] = y[static_cast<size_t>(b) * w + e];
        __syncthreads();
        for (uint32_t r = 0; r < count; r++) {
#pragma unroll
            for (int t = 0; t < kGramPer; t++)
                if (pi[t] >= 0) acc[t] = __builtin_fmaf(tile[r * w + pi[t]], tile[r * w + pj[t]], acc[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < kGramPer; t++)
        if (pi[t] >= 0) partial[static_cast<size_t>(chunk) * pairs + tid + 256u * t] = acc[t];
}

// gram[i][j] (full w x w, f64) = sum over chunks in order of the partial triangles.
__global__ void gram_reduce_kernel(const float* partial, uint32_t chunks, uint32_t w, double* gram) {
    const uint32_t pairs = w * (w + 1) / 2;
    const uint32_t p0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (p0 >= pairs) return;
    uint32_t p = p0, i = 0;
    while (p >= w - i) {
        p -= w - i;
        i++;
    }
    const uint32_t j = i + p;
    double s = 0.0;
    for (uint32_t k = 0; k < chunks; k++) s += static_cast<double>(partial[static_cast<size_t>(k) * pairs + p0]);
    gram[static_cast<size_t>(i) * w + j] = s;
    gram[static_cast<size_t>(j) * w + i] = s;
}

// out (rows x wo) = y (rows x w) T (w x wo), T staged in LDS; one wave per row, y's row
// spread over the lanes (two values each) and broadcast by shuffles.
__global__ __launch_bounds__(256) void transform_kernel(const float* y, uint32_t rows, uint32_t w,
                                                        const float* t, uint32_t wo, float* out) {
    __shared__ float ts[kPcaMaxWidth * kPcaMaxWidth];
    for (uint32_t e = threadIdx.x; e < w * wo; e += 256) ts[e] = t[e];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t r = blockIdx.x * 4 + (threadIdx.x >> 6); r < rows; r += gridDim.x * 4) {
        const float* src = y + static_cast<size_t>(r) * w;
        const float y0 = lane < w ? src[lane] : 0.0f;
        const float y1 = lane + 64 < w ? src[lane + 64] : 0.0f;
        float a0 = 0.0f, a1 = 0.0f;
        for (uint32_t k = 0; k < w; k++) {
            const float yk = __shfl(k < 64 ? y0 : y1, static_cast<int>(k & 63), kWave);
            if (lane < wo) a0 = __builtin_fmaf(yk, ts[k * wo + lane], a0);
            if (lane + 64 < wo) a1 = __builtin_fmaf(yk, ts[k * wo + lane + 64], a1);
        }
        float* dst = out + static_cast<size_t>(r) * wo;
        if (lane < wo) dst[lane] = a0;
        if (lane + 64 < wo) dst[lane + 64] = a1;
    }
}

// Columns of y (rows x w) whose sign is -1 are negated.
__global__ void flip_kernel(float* y, uint32_t rows, uint32_t w, const float* sign) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(rows) * w) return;
    const uint32_t c = static_cast<uint32_t>(e % w);
    if (sign[c] < 0.0f) y[e] = -y[e];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// The sketch (and the oracle's): splitmix64 from `seed`, pairs of 53-bit uniforms through
// Box-Muller in f64, rounded to f32; row-major rows x cols.
void pca_sketch(size_t rows, size_t cols, uint64_t seed, float* out) {
    uint64_t state = seed;
    const size_t count = rows * cols;
    size_t i = 0;
    while (i < count) {
        double u[2];
        for (double& x : u) {
            state += 0x9E3779B97F4A7C15ull;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            x = static_cast<double>(z >> 11) * 0x1p-53;
        }
        const double uniform = std::max(u[0], DBL_MIN);
        const double radius = std::sqrt(-2.0 * std::log(uniform));
        const double angle = 6.283185307179586 * u[1];
        out[i++] = static_cast<float>(radius * std::cos(angle));
        if (i < count) out[i++] = static_cast<float>(radius * std::sin(angle));
    }
}

// Symmetric eigensolve by cyclic Jacobi rotations (the reference's pca.rs:390-445 method:
// the same sweeps, rotation and stopping rule, so the two agree to rounding). `a` (n x n,
// row-major) is destroyed; vectors are columns. False when 60 sweeps do not converge.
bool jacobi(std::vector<double>& a, size_t n, std::vector<double>& values, std::vector<double>& vectors) {
    vectors.assign(n * n, 0.0);
    for (size_t i = 0; i < n; i++) vectors[i * n + i] = 1.0;
    double total = 0.0;
    for (double v : a) total += v * v;
    const double tolerance = DBL_EPSILON * DBL_EPSILON * std::max(total, DBL_MIN);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (size_t r = 0; r < n; r++)
            for (size_t c = 0; c < n; c++)
                if (r != c) off += a[r * n + c] * a[r * n + c];
        if (off <= tolerance) {
            values.resize(n);
            for (size_t i = 0; i < n; i++) values[i] = a[i * n + i];
            return true;
        }
        for (size_t p = 0; p < n; p++) {
            for (size_t q = p + 1; q < n; q++) {
                const double pivot = a[p * n + q];
                if (pivot == 0.0) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * pivot);
                const double sgn = std::signbit(theta) ? -1.0 : 1.0;
                const double t = sgn / std::max(std::fabs(theta) + std::sqrt(theta * theta + 1.0), DBL_EPSILON);
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (size_t k = 0; k < n; k++) {
                    const double l = a[k * n + p], r = a[k * n + q];
                    a[k * n + p] = c * l - s * r;
                    a[k * n + q] = s * l + c * r;
                }
                for (size_t k = 0; k < n; k++) {
                    const double l = a[p * n + k], r = a[q * n + k];
                    a[p * n + k] = c * l - s * r;
                    a[q * n + k] = s * l + c * r;
                }
                a[p * n + q] = 0.0;
                a[q * n + p] = 0.0;
                for (size_t k = 0; k < n; k++) {
                    const double l = vectors[k * n + p], r = vectors[k * n + q];
                    vectors[k * n + p] = c * l - s * r;
                    vectors[k * n + q] = s * l + c * r;
                }
            }
        }
    }
    return false;
}

#define PCA_CHECK(expr, step)                                    \
    do {                                                         \
        hipError_t err_ = (expr);                                \
        if (err_ != hipSuccess) return hip_failure(err_, step);  \
    } while (0)

struct PcaRun {
    scrust_hip_context* ctx;
    hipStream_t stream;
    uint32_t n, g, nnz, w;
    // Device views into ctx->pca.
    uint32_t *colptr, *crow, *flags;
    float *cval, *mean, *omega, *ya, *yb, *ga, *gb, *shift, *colsum, *gram_part, *tmat, *emb_tmp;
    double *sums, *squares, *gram, *chunk_part, *col64;
    void* sort_temp;
    size_t sort_temp_bytes;
    uint64_t *keys_a, *keys_b;
    float *vals_a;

    scrust_hip_status gram_of(const float* y, uint32_t rows, std::vector<double>& out) {
        const uint32_t chunks = (rows + kChunkRows - 1) / kChunkRows;
        const uint32_t pairs = w * (w + 1) / 2;
        hipLaunchKernelGGL(gram_partial_kernel, dim3(chunks), dim3(256), 0, stream, y, rows, w, gram_part);
        hipLaunchKernelGGL(gram_reduce_kernel, dim3((pairs + 255) / 256), dim3(256), 0, stream, gram_part, chunks, w, gram);
        PCA_CHECK(hipGetLastError(), "launch Gram");
        out.resize(static_cast<size_t>(w) * w);
        PCA_CHECK(hipMemcpyAsync(out.data(), gram, out.size() * sizeof(double), hipMemcpyDeviceToHost, stream), "read Gram");
        PCA_CHECK(hipStreamSynchronize(stream), "read Gram");
        return SCRUST_HIP_OK;
    }

    // y (rows x w) <- y T with T = V diag(1 / sqrt(max(lambda, 1e-7 lambda_max))) (pca.rs:365-386);
    // the result lands in `to`.
    scrust_hip_status whiten(const float* y, uint32_t rows, float* to) {
        std::vector<double> gm, values, vectors;
        scrust_hip_status st = gram_of(y, rows, gm);
        if (st != SCRUST_HIP_OK) return st;
        if (!jacobi(gm, w, values, vectors))
            return fail(SCRUST_HIP_KERNEL, "HIP kernel pca: Jacobi eigenvalue iteration did not converge in 60 iterations");
        double largest = 0.0;
        for (double v : values) largest = std::max(largest, v);
        const double floor_value = largest * 1e-7;
        std::vector<float> t(static_cast<size_t>(w) * w);
        for (uint32_t c = 0; c < w; c++) {
            const double scale = floor_value > 0.0 ? 1.0 / std::sqrt(std::max(values[c], floor_value)) : 0.0;
            for (uint32_t r = 0; r < w; r++) t[static_cast<size_t>(r) * w + c] = static_cast<float>(vectors[static_cast<size_t>(r) * w + c] * scale);
        }
        return apply(y, rows, t, w, to);
    }

    scrust_hip_status apply(const float* y, uint32_t rows, const std::vector<float>& t, uint32_t wo, float* to) {
        PCA_CHECK(hipMemcpyAsync(tmat, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, stream), "upload transform");
        const unsigned blocks = std::min<unsigned>(4096, (rows + 3) / 4);
        hipLaunchKernelGGL(transform_kernel, dim3(blocks), dim3(256), 0, stream, y, rows, w, tmat, wo, to);
        PCA_CHECK(hipGetLastError(), "launch transform");
        // The host transform must stay alive until the copy has been consumed.
        PCA_CHECK(hipStreamSynchronize(stream), "transform");
        return SCRUST_HIP_OK;
    }

    // orthonormalize (pca.rs:352-363): whiten twice; y ends in `y` (via `spare`).
    scrust_hip_status orthonormalize(float* y, uint32_t rows, float* spare) {
        scrust_hip_status st = whiten(y, rows, spare);
        if (st != SCRUST_HIP_OK) return st;
        return whiten(spare, rows, y);
    }

    // out (n x w) = X_c rhs (rhs g x w).
    scrust_hip_status times(const uint32_t* indptr, const uint32_t* indices, const float* values, const float* rhs,
                            bool centre, float* out) {
        if (centre) hipLaunchKernelGGL(mean_times_kernel, dim3((w + 63) / 64), dim3(64), 0, stream, mean, rhs, g, w, shift);
        hipLaunchKernelGGL(spmm_kernel, dim3((n + 3) / 4), dim3(256), 0, stream, indptr, indices, values, n, rhs, w,
                           centre ? static_cast<const float*>(shift) : nullptr, static_cast<const float*>(nullptr), out);
        PCA_CHECK(hipGetLastError(), "launch X Y");
        return SCRUST_HIP_OK;
    }

    // out (g x w) = X_c^T lhs (lhs n x w), on the CSC form.
    scrust_hip_status transpose_times(const float* lhs, bool centre, float* out) {
        if (centre) {
            const uint32_t chunks = (n + kChunkRows - 1) / kChunkRows;
            hipLaunchKernelGGL(colsum_partial_kernel, dim3(chunks), dim3(kPcaMaxWidth), 0, stream, lhs, n, w, chunk_part);
            hipLaunchKernelGGL(chunk_reduce_kernel, dim3(1), dim3(kPcaMaxWidth), 0, stream, chunk_part, chunks, w, colsum,
                               static_cast<double*>(nullptr));
        }
        hipLaunchKernelGGL(spmm_kernel, dim3((g + 3) / 4), dim3(256), 0, stream, colptr, crow, cval, g, lhs, w,
                           centre ? static_cast<const float*>(colsum) : nullptr,
                           centre ? static_cast<const float*>(mean) : nullptr, out);
        PCA_CHECK(hipGetLastError(), "launch X^T Y");
        return SCRUST_HIP_OK;
    }
};

// Device workspace layout of one PCA.
struct PcaLayout {
    size_t colptr, crow, cval, flags, mean, sums, squares, omega, ya, yb, ga, gb, shift, colsum, gram_part, gram,
        tmat, chunk_part, col64, keys_a, keys_b, vals_a, sort_temp, sort_temp_bytes, total;
};

PcaLayout pca_layout(uint32_t n, uint32_t g, uint32_t nnz, uint32_t w, size_t sort_temp_bytes) {
    PcaLayout l{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t off = at;
        at += align256(bytes + 1);
        return off;
    };
    const uint32_t chunks = (std::max(n, g) + kChunkRows - 1) / kChunkRows;
    l.colptr = take((static_cast<size_t>(g) + 1) * 4);
    l.crow = take(static_cast<size_t>(nnz) * 4);
    l.cval = take(static_cast<size_t>(nnz) * 4);
    l.flags = take(4);
    l.mean = take(static_cast<size_t>(g) * 4);
    l.sums = take(static_cast<size_t>(g) * 8);
    l.squares = take(static_cast<size_t>(g) * 8);
    l.omega = take(static_cast<size_t>(g) * w * 4);
    l.ya = take(static_cast<size_t>(n) * w * 4);
    l.yb = take(static_cast<size_t>(n) * w * 4);
    l.ga = take(static_cast<size_t>(g) * w * 4);
    l.gb = take(static_cast<size_t>(g) * w * 4);
    l.shift = take(kPcaMaxWidth * 4);
    l.colsum = take(kPcaMaxWidth * 4);
    l.gram_part = take(static_cast<size_t>(chunks) * w * (w + 1) / 2 * 4);
    l.gram = take(static_cast<size_t>(w) * w * 8);
    l.tmat = take(static_cast<size_t>(kPcaMaxWidth) * kPcaMaxWidth * 4);
    l.chunk_part = take(static_cast<size_t>(chunks) * kPcaMaxWidth * 8);
    l.col64 = take(kPcaMaxWidth * 8);
    l.keys_a = take(static_cast<size_t>(nnz) * 8);
    l.keys_b = take(static_cast<size_t>(nnz) * 8);
    l.vals_a = take(static_cast<size_t>(nnz) * 4);
    l.sort_temp = take(sort_temp_bytes);
    l.sort_temp_bytes = sort_temp_bytes;
    l.total = at;
    return l;
}

uint32_t key_bits(uint32_t g) {
    uint32_t bits = 32;
    while ((1ull << (bits - 32)) < g) bits++;
    return bits;
}

// The whole PCA on device-resident CSR arrays; embedding (n x n_components) into device
// memory, loadings and variances into host memory.
scrust_hip_status run_pca(scrust_hip_context* ctx, const uint32_t* indptr, const uint32_t* indices,
                          const float* values, uint32_t n, uint32_t g, uint32_t nnz, uint32_t n_components,
                          bool zero_center, uint64_t seed, float* embedding, float* components,
                          float* explained_variance, float* explained_variance_ratio) {
    hipStream_t s = ctx->stream;
    const uint32_t max_components = std::min(n, g);
    const uint32_t w = std::min<uint32_t>(n_components + kPcaOversampling, max_components);
    size_t sort_bytes = 0;
    PCA_CHECK(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                                        static_cast<float*>(nullptr), static_cast<float*>(nullptr), nnz, 0, key_bits(g), s),
              "size the sort");
    const PcaLayout l = pca_layout(n, g, nnz, w, sort_bytes);
    PCA_CHECK(ctx->pca.reserve(l.total), "allocate PCA workspace");
    char* base = static_cast<char*>(ctx->pca.ptr);
    PcaRun r{};
    r.ctx = ctx;
    r.stream = s;
    r.n = n;
    r.g = g;
    r.nnz = nnz;
    r.w = w;
    auto at = [&](size_t off) { return static_cast<void*>(base + off); };
    r.colptr = static_cast<uint32_t*>(at(l.colptr));
    r.crow = static_cast<uint32_t*>(at(l.crow));
    r.cval = static_cast<float*>(at(l.cval));
    r.flags = static_cast<uint32_t*>(at(l.flags));
    r.mean = static_cast<float*>(at(l.mean));
    r.sums = static_cast<double*>(at(l.sums));
    r.squares = static_cast<double*>(at(l.squares));
    r.omega = static_cast<float*>(at(l.omega));
    r.ya = static_cast<float*>(at(l.ya));
    r.yb = static_cast<float*>(at(l.yb));
    r.ga = static_cast<float*>(at(l.ga));
    r.gb = static_cast<float*>(at(l.gb));
    r.shift = static_cast<float*>(at(l.shift));
    r.colsum = static_cast<float*>(at(l.colsum));
    r.gram_part = static_cast<float*>(at(l.gram_part));
    r.gram = static_cast<double*>(at(l.gram));
    r.tmat = static_cast<float*>(at(l.tmat));
    r.chunk_part = static_cast<double*>(at(l.chunk_part));
    r.col64 = static_cast<double*>(at(l.col64));
    r.keys_a = static_cast<uint64_t*>(at(l.keys_a));
    r.keys_b = static_cast<uint64_t*>(at(l.keys_b));
    r.vals_a = static_cast<float*>(at(l.vals_a));
    r.sort_temp = at(l.sort_temp);
    r.sort_temp_bytes = l.sort_temp_bytes;

    // Validity, then the CSC form by a stable radix sort of (gene, cell) keys.
    hipLaunchKernelGGL(clear_word_kernel, dim3(1), dim3(1), 0, s, r.flags);
    hipLaunchKernelGGL(csr_check_kernel, dim3((n + 255) / 256), dim3(256), 0, s, indptr, indices, values, n, g, nnz, r.flags);
    uint32_t flags = 0;
    PCA_CHECK(hipMemcpyAsync(&flags, r.flags, 4, hipMemcpyDeviceToHost, s), "read flags");
    PCA_CHECK(hipStreamSynchronize(s), "check the matrix");
    if (flags & kPcaFlagBadId)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected CSR row pointers from 0 to nnz, non-decreasing, and gene ids below " +
                                          std::to_string(g) + ", got a matrix that breaks them");
    if (flags & kPcaFlagNonFinite) return invalid("values", "finite", "NaN or infinity");
    hipLaunchKernelGGL(csc_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, s, indptr, indices, values, n, r.keys_a, r.vals_a);
    PCA_CHECK(hipGetLastError(), "launch CSC keys");
    size_t temp = r.sort_temp_bytes;
    PCA_CHECK(rocprim::radix_sort_pairs(r.sort_temp, temp, r.keys_a, r.keys_b, r.vals_a, r.cval, nnz, 0, key_bits(g), s),
              "sort into CSC");
    hipLaunchKernelGGL(colptr_kernel, dim3((g + 256) / 256), dim3(256), 0, s, r.keys_b, nnz, g, r.colptr);
    hipLaunchKernelGGL(low_word_kernel, dim3((nnz + 255) / 256 + 1), dim3(256), 0, s, r.keys_b, nnz, r.crow);
    hipLaunchKernelGGL(colstats_kernel, dim3((g + 255) / 256), dim3(256), 0, s, r.colptr, r.cval, g, n, r.sums, r.squares, r.mean);
    PCA_CHECK(hipGetLastError(), "launch column statistics");
    std::vector<double> sums(g), squares(g);
    PCA_CHECK(hipMemcpyAsync(sums.data(), r.sums, g * 8ull, hipMemcpyDeviceToHost, s), "read sums");
    PCA_CHECK(hipMemcpyAsync(squares.data(), r.squares, g * 8ull, hipMemcpyDeviceToHost, s), "read sums");
    // The sketch.
    std::vector<float> omega(static_cast<size_t>(g) * w);
    pca_sketch(g, w, seed, omega.data());
    PCA_CHECK(hipMemcpyAsync(r.omega, omega.data(), omega.size() * 4, hipMemcpyHostToDevice, s), "upload sketch");
    PCA_CHECK(hipStreamSynchronize(s), "upload sketch");
    const double ddof = zero_center ? 1.0 : 0.0;
    double total_variance = 0.0;
    for (uint32_t k = 0; k < g; k++)
        total_variance += std::max(squares[k] - sums[k] * sums[k] / n, 0.0) / (static_cast<double>(n) - ddof);

    // Range finder and power iterations (pca.rs:104-119).
    scrust_hip_status st = r.times(indptr, indices, values, r.omega, zero_center, r.ya);
    if (st == SCRUST_HIP_OK) st = r.orthonormalize(r.ya, n, r.yb);
    const int iterations = static_cast<double>(n_components) < 0.1 * max_components ? 7 : 4;
    for (int it = 0; it < iterations && st == SCRUST_HIP_OK; it++) {
        st = r.transpose_times(r.ya, zero_center, r.ga);
        if (st == SCRUST_HIP_OK) st = r.orthonormalize(r.ga, g, r.gb);
        if (st == SCRUST_HIP_OK) st = r.times(indptr, indices, values, r.ga, zero_center, r.ya);
        if (st == SCRUST_HIP_OK) st = r.orthonormalize(r.ya, n, r.yb);
    }
    if (st != SCRUST_HIP_OK) return st;
    // B^T = X_c^T Q (g x w) and its Gram (pca.rs:145-150).
    st = r.transpose_times(r.ya, zero_center, r.ga);
    std::vector<double> gm, values_h, vectors;
    if (st == SCRUST_HIP_OK) st = r.gram_of(r.ga, g, gm);
    if (st != SCRUST_HIP_OK) return st;
    if (!jacobi(gm, w, values_h, vectors))
        return fail(SCRUST_HIP_KERNEL, "HIP kernel pca: Jacobi eigenvalue iteration did not converge in 60 iterations");
    std::vector<uint32_t> order(w);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return values_h[a] > values_h[b]; });
    std::vector<double> singular(n_components);
    for (uint32_t c = 0; c < n_components; c++) singular[c] = std::sqrt(std::max(values_h[order[c]], 0.0));
    // Embedding Q (V S) and loadings B^T (V S^-1) (pca.rs:160-190).
    std::vector<float> vs(static_cast<size_t>(w) * n_components), vsi(static_cast<size_t>(w) * n_components);
    for (uint32_t c = 0; c < n_components; c++) {
        const double inverse = singular[c] > 0.0 ? 1.0 / singular[c] : 0.0;
        for (uint32_t k = 0; k < w; k++) {
            const double v = vectors[static_cast<size_t>(k) * w + order[c]];
            vs[static_cast<size_t>(k) * n_components + c] = static_cast<float>(v * singular[c]);
            vsi[static_cast<size_t>(k) * n_components + c] = static_cast<float>(v * inverse);
        }
    }
    st = r.apply(r.ya, n, vs, n_components, embedding);
    if (st == SCRUST_HIP_OK) st = r.apply(r.ga, g, vsi, n_components, r.gb);
    if (st != SCRUST_HIP_OK) return st;
    std::vector<float> loadings(static_cast<size_t>(g) * n_components);
    PCA_CHECK(hipMemcpyAsync(loadings.data(), r.gb, loadings.size() * 4, hipMemcpyDeviceToHost, s), "read loadings");
    PCA_CHECK(hipStreamSynchronize(s), "read loadings");
    // Signs: the largest-magnitude loading of every component positive (pca.rs:471-490).
    std::vector<float> sign(n_components, 1.0f);
    for (uint32_t c = 0; c < n_components; c++) {
        float extreme = 0.0f;
        for (uint32_t k = 0; k < g; k++) {
            const float v = loadings[static_cast<size_t>(k) * n_components + c];
            if (std::fabs(v) > std::fabs(extreme)) extreme = v;
        }
        if (extreme < 0.0f) sign[c] = -1.0f;
        for (uint32_t k = 0; k < g; k++)
            components[static_cast<size_t>(c) * g + k] = sign[c] * loadings[static_cast<size_t>(k) * n_components + c];
    }
    PCA_CHECK(hipMemcpyAsync(r.shift, sign.data(), n_components * 4ull, hipMemcpyHostToDevice, s), "upload signs");
    const size_t cells = static_cast<size_t>(n) * n_components;
    hipLaunchKernelGGL(flip_kernel, dim3(static_cast<unsigned>((cells + 255) / 256)), dim3(256), 0, s, embedding, n,
                       n_components, r.shift);
    PCA_CHECK(hipGetLastError(), "launch sign flip");
    // Explained variance (pca.rs:192-208).
    std::vector<double> var(n_components);
    if (zero_center) {
        for (uint32_t c = 0; c < n_components; c++) var[c] = singular[c] * singular[c] / (static_cast<double>(n) - 1.0);
        PCA_CHECK(hipStreamSynchronize(s), "sign flip");
    } else {
        const uint32_t chunks = (n + kChunkRows - 1) / kChunkRows;
        hipLaunchKernelGGL(colsum_partial_kernel, dim3(chunks), dim3(kPcaMaxWidth), 0, s, embedding, n, n_components, r.chunk_part);
        hipLaunchKernelGGL(chunk_reduce_kernel, dim3(1), dim3(kPcaMaxWidth), 0, s, r.chunk_part, chunks, n_components,
                           static_cast<float*>(nullptr), r.col64);
        PCA_CHECK(hipGetLastError(), "launch score means");
        std::vector<double> mean_h(n_components);
        PCA_CHECK(hipMemcpyAsync(mean_h.data(), r.col64, n_components * 8ull, hipMemcpyDeviceToHost, s), "read means");
        PCA_CHECK(hipStreamSynchronize(s), "read means");
        for (double& m : mean_h) m /= static_cast<double>(n);
        PCA_CHECK(hipMemcpyAsync(r.col64, mean_h.data(), n_components * 8ull, hipMemcpyHostToDevice, s), "upload means");
        hipLaunchKernelGGL(colvar_partial_kernel, dim3(chunks), dim3(kPcaMaxWidth), 0, s, embedding, n, n_components,
                           static_cast<const double*>(r.col64), r.chunk_part);
        hipLaunchKernelGGL(chunk_reduce_kernel, dim3(1), dim3(kPcaMaxWidth), 0, s, r.chunk_part, chunks, n_components,
                           static_cast<float*>(nullptr), r.col64);
        PCA_CHECK(hipGetLastError(), "launch score variances");
        PCA_CHECK(hipMemcpyAsync(var.data(), r.col64, n_components * 8ull, hipMemcpyDeviceToHost, s), "read variances");
        PCA_CHECK(hipStreamSynchronize(s), "read variances");
        for (double& v : var) v /= static_cast<double>(n);
    }
    for (uint32_t c = 0; c < n_components; c++) {
        explained_variance[c] = static_cast<float>(var[c]);
        explained_variance_ratio[c] =
            total_variance > 0.0 ? static_cast<float>(static_cast<double>(explained_variance[c]) / total_variance) : 0.0f;
    }
    return SCRUST_HIP_OK;
}

// Argument checks of pca.rs:84-100 (InvalidParameter), then this path's own limits.
scrust_hip_status check_pca(size_t n, size_t g, size_t n_components, const void* a, const void* b, const void* c,
                            const void* emb, const void* comp, const void* ev, const void* evr) {
    const size_t max_components = std::min(n, g);
    if (n_components == 0) return invalid("n_components", "at least 1", std::to_string(n_components));
    if (n_components > max_components)
        return invalid("n_components", "at most min(n_cells, n_genes)", std::to_string(n_components));
    if (n < 2) return invalid("n_cells", "at least 2", std::to_string(n));
    if (n_components + kPcaOversampling > kPcaMaxWidth && max_components > kPcaMaxWidth)
        return invalid("n_components", "at most 104 on the device (a sketch of at most 128 columns)",
                       std::to_string(n_components));
    if (n >= 0xFFFFFFFFull || g >= 0xFFFFFFFFull) return invalid("matrix", "fewer than 2^32 - 1 rows and columns", "more");
    if (!a || !b || !c || !emb || !comp || !ev || !evr) return invalid("buffers", "non-null", "a null pointer");
    return SCRUST_HIP_OK;
}

}  // namespace

}  // namespace scrust_hip

using namespace scrust_hip;

extern "C" {

scrust_hip_status scrust_hip_pca_device(scrust_hip_context* context, const uint32_t* indptr, const uint32_t* indices,
                                        const float* values, size_t n_rows, size_t n_cols, size_t n_components,
                                        int zero_center, uint64_t seed, float* out_embedding, float* out_components,
                                        float* out_explained_variance, float* out_explained_variance_ratio) {
    clear_error();
    scrust_hip_status st = check_pca(n_rows, n_cols, n_components, indptr, indices, values, out_embedding,
                                     out_components, out_explained_variance, out_explained_variance_ratio);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    uint32_t nnz = 0;
    hipError_t err = hipMemcpyAsync(&nnz, indptr + n_rows, 4, hipMemcpyDeviceToHost, context->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(context->stream);
    if (err != hipSuccess) return hip_failure(err, "read nnz");
    return run_pca(context, indptr, indices, values, static_cast<uint32_t>(n_rows), static_cast<uint32_t>(n_cols), nnz,
                   static_cast<uint32_t>(n_components), zero_center != 0, seed, out_embedding, out_components,
                   out_explained_variance, out_explained_variance_ratio);
}

scrust_hip_status scrust_hip_pca(scrust_hip_context* context, const uint32_t* indptr, const uint32_t* indices,
                                 const float* values, size_t n_rows, size_t n_cols, size_t n_components,
                                 int zero_center, uint64_t seed, float* out_embedding, float* out_components,
                                 float* out_explained_variance, float* out_explained_variance_ratio) {
    clear_error();
    scrust_hip_status st = check_pca(n_rows, n_cols, n_components, indptr, indices, values, out_embedding,
                                     out_components, out_explained_variance, out_explained_variance_ratio);
    if (st != SCRUST_HIP_OK) return st;
    // The CSR's shape on the host (csr_from_py's checks: crates/scrust-py/src/convert.rs).
    const size_t nnz = indptr[n_rows];
    if (indptr[0] != 0 || nnz >= 0xFFFFFFFFull)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected CSR row pointers from 0, got " + std::to_string(indptr[0]));
    for (size_t r = 0; r < n_rows; r++)
        if (indptr[r + 1] < indptr[r])
            return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected non-decreasing CSR row pointers, got a decrease at row " +
                                              std::to_string(r));
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipStream_t s = context->stream;
    const size_t ip = (n_rows + 1) * 4, ix = nnz * 4, vb = nnz * 4, eb = n_rows * n_components * 4;
    const size_t total = align256(ip) + align256(ix) + align256(vb) + align256(eb);
    hipError_t err = context->pca_io.reserve(total);
    if (err != hipSuccess) return hip_failure(err, "allocate PCA input");
    char* base = static_cast<char*>(context->pca_io.ptr);
    uint32_t* d_indptr = reinterpret_cast<uint32_t*>(base);
    uint32_t* d_indices = reinterpret_cast<uint32_t*>(base + align256(ip));
    float* d_values = reinterpret_cast<float*>(base + align256(ip) + align256(ix));
    float* d_emb = reinterpret_cast<float*>(base + align256(ip) + align256(ix) + align256(vb));
    err = hipMemcpyAsync(d_indptr, indptr, ip, hipMemcpyHostToDevice, s);
    if (err == hipSuccess && nnz) err = hipMemcpyAsync(d_indices, indices, ix, hipMemcpyHostToDevice, s);
    if (err == hipSuccess && nnz) err = hipMemcpyAsync(d_values, values, vb, hipMemcpyHostToDevice, s);
    if (err != hipSuccess) return hip_failure(err, "copy the matrix");
    st = run_pca(context, d_indptr, d_indices, d_values, static_cast<uint32_t>(n_rows), static_cast<uint32_t>(n_cols),
                 static_cast<uint32_t>(nnz), static_cast<uint32_t>(n_components), zero_center != 0, seed, d_emb,
                 out_components, out_explained_variance, out_explained_variance_ratio);
    if (st != SCRUST_HIP_OK) return st;
    err = hipMemcpyAsync(out_embedding, d_emb, eb, hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    if (err != hipSuccess) return hip_failure(err, "copy the embedding");
    return SCRUST_HIP_OK;
}

}  // extern "C"
