// The t-SNE input distances on gfx950: centring and the pairwise squared-distance matrix
// of crates/scrust-core/src/tsne.rs:173-188 (`squared_distances(&centred(&points))`,
// called at tsne.rs:86), the same centred expansion the kNN pass contracts on MFMA, with
// the n x n matrix written out instead of selected from.
//
//   1. tsne_colsum_kernel  exact f32 column sums (superaccumulators, any order gives the
//                          same bits), rounded once to f32
//   2. tsne_mean_kernel    mean = sum * f32(1/n)          (tsne.rs:174: sum, affine(1/n))
//   3. tsne_centre_kernel  c = x - mean in f32, |c|^2 as f32 squares summed in dimension
//                          order (tsne.rs:175, :182: broadcast_sub, sqr, sum_keepdim)
//   4. tsne_tile_kernel    128 x 128 output tiles of the upper triangle: Gram block on the
//                          fp32 MFMA (v_mfma_f32_16x16x4f32), then (|a|^2 + |b|^2) - 2 g
//                          clamped at 0 (tsne.rs:183-187), stored row-major and mirrored
//
// The matrix is n^2 * 4 bytes written once (1.6 GB at the reference's 20,000-cell cap),
// against 2 n^2 d flops: the tile kernel is bound by HBM writes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "knn_device.hpp"
#include "knn_launch.hpp"
#include "knn_tsne.hpp"

namespace scrust_hip {

namespace {

constexpr int kTsneTile = 128;  // output rows and columns per workgroup
constexpr int kTsneKc = 32;     // features per LDS stage
constexpr int kSumWaves = 4;
// Mirrored stores go through LDS, 32 columns of a wave's 64 x 64 block at a time, as rows
// of 64 values at a stride of 68 floats (16-byte aligned; lane groups fall in distinct banks).
constexpr int kMirrorCols = 32, kMirrorStride = 68;
constexpr int kTsneLds = 2 * kTsneTile * (kTsneKc + 1) > 4 * kMirrorCols * kMirrorStride
                             ? 2 * kTsneTile * (kTsneKc + 1)
                             : 4 * kMirrorCols * kMirrorStride;

// Column sums: a block takes colsum_rows(n) rows of 64 columns; each wave sums every
// fourth row into its own LDS superaccumulator, the four are combined and added to the
// column's global digits.
__global__ __launch_bounds__(256) void tsne_colsum_kernel(const float* __restrict__ x, uint32_t n,
                                                          uint32_t d, unsigned long long* __restrict__ acc) {
    __shared__ long long part[kSumWaves][kWave][kAccWords + 1];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const uint32_t col = blockIdx.y * kWave + lane;
    const uint32_t rows = colsum_rows(n);
    const size_t r0 = static_cast<size_t>(blockIdx.x) * rows;
    for (int k = 0; k < kAccWords; k++) part[w][lane][k] = 0;
    size_t r1 = r0 + rows;
    if (r1 > n) r1 = n;
    if (col < d)
        for (size_t r = r0 + w; r < r1; r += kSumWaves) exact_sum_add(part[w][lane], x[r * d + col]);
    __syncthreads();
    if (w == 0 && col < d) {
        for (int k = 0; k < kAccWords; k++) {
            long long v = 0;
            for (int ww = 0; ww < kSumWaves; ww++) v += part[ww][lane][k];
            if (v) atomicAdd(&acc[col * kAccWords + k], static_cast<unsigned long long>(v));
        }
    }
}

__global__ void tsne_mean_kernel(const unsigned long long* __restrict__ acc, uint32_t n, uint32_t d,
                                 float* __restrict__ mean) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= d) return;
    long long words[kAccWords];
    for (int k = 0; k < kAccWords; k++) words[k] = static_cast<long long>(acc[col * kAccWords + k]);
    const float sum = static_cast<float>(exact_sum_value(words));
    mean[col] = sum * static_cast<float>(1.0 / static_cast<double>(n));
}

// One thread per row: centred coordinates (zero padded to dpad) and the squared norm.
// Rows n .. npad - 1 get a zero norm so that the tile kernel reads norms without checks.
__global__ void tsne_centre_kernel(const float* __restrict__ x, uint32_t n, uint32_t npad, uint32_t d,
                                   uint32_t dpad, const float* __restrict__ mean, float* __restrict__ coords,
                                   float* __restrict__ norms, uint32_t* __restrict__ flags) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= npad) return;
    if (row >= n) {
        norms[row] = 0.0f;
        return;
    }
    const float* src = x + static_cast<size_t>(row) * d;
    float* dst = coords + static_cast<size_t>(row) * dpad;
    float nrm = 0.0f;
    bool finite = true;
    for (uint32_t j = 0; j < d; j++) {
        const float v = src[j];
        finite = finite && __builtin_isfinite(v);
        const float c = v - mean[j];
        dst[j] = c;
        const float sq = c * c;  // squared, rounded, then summed (no fused multiply-add)
        nrm = nrm + sq;
    }
    for (uint32_t j = d; j < dpad; j++) dst[j] = 0.0f;
    norms[row] = nrm;
    if (!finite) atomicOr(flags, kTsneFlagNonFinite);
}

// Workgroup (bx, by) with bx >= by computes rows [128 by, +128) x columns [128 bx, +128)
// and, off the diagonal, also writes them mirrored (the matrix is symmetric: row i's dot
// with row j is the same products in the same order as j's with i), so half the
// workgroups exit at once. Wave w takes the 64 x 64 quarter (w / 2, w % 2) as 4 x 4
// blocks of 16 x 16. MFMA operands: lane l gives A[m = l % 16][k = l / 16] from a column
// point and B[k = l / 16][n = l % 16] from a row point; it holds D[4 (l / 16) + r][l % 16]:
// one row (l % 16) and four consecutive columns (4 (l / 16) + r), stored as one float4.
__global__ __launch_bounds__(256) void tsne_tile_kernel(const float* __restrict__ coords,
                                                        const float* __restrict__ norms, uint32_t n,
                                                        uint32_t dpad, float* __restrict__ out) {
    if (blockIdx.x < blockIdx.y) return;
    // One LDS buffer: the two operand stages during the contraction, then the transposed
    // staging of the mirrored stores. Operand row stride 33: the 16 rows x 4 features one
    // read touches land in 64 distinct banks.
    __shared__ float lds[kTsneLds];
    typedef float Stage[kTsneKc + 1];
    Stage* sa = reinterpret_cast<Stage*>(lds);
    Stage* sb = reinterpret_cast<Stage*>(lds + kTsneTile * (kTsneKc + 1));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wr = w >> 1, wc = w & 1, lr = lane & 15, lk = lane >> 4;
    const uint32_t r0 = blockIdx.y * kTsneTile, c0 = blockIdx.x * kTsneTile;
    f32x4 acc[4][4];  // [row block][column block]
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

    for (uint32_t kc = 0; kc < dpad; kc += kTsneKc) {
        // 128 rows x 8 float4 per operand; features past dpad (and rows past n) are zero.
        for (int i = tid; i < kTsneTile * (kTsneKc / 4); i += 256) {
            const int row = i / (kTsneKc / 4), q = i % (kTsneKc / 4);
            const uint32_t k = kc + 4 * q;
            float4 va = make_float4(0.0f, 0.0f, 0.0f, 0.0f), vb = va;
            if (k < dpad) {
                if (r0 + row < n) va = *reinterpret_cast<const float4*>(coords + static_cast<size_t>(r0 + row) * dpad + k);
                if (c0 + row < n) vb = *reinterpret_cast<const float4*>(coords + static_cast<size_t>(c0 + row) * dpad + k);
            }
            sa[row][4 * q + 0] = va.x; sa[row][4 * q + 1] = va.y; sa[row][4 * q + 2] = va.z; sa[row][4 * q + 3] = va.w;
            sb[row][4 * q + 0] = vb.x; sb[row][4 * q + 1] = vb.y; sb[row][4 * q + 2] = vb.z; sb[row][4 * q + 3] = vb.w;
        }
        __syncthreads();
        // Features past dpad are zero in LDS; steps past them are skipped (uniformly).
        const uint32_t rest = dpad - kc;
        const int steps = rest >= static_cast<uint32_t>(kTsneKc) ? kTsneKc / 4 : static_cast<int>(rest / 4);
#pragma unroll
        for (int ks = 0; ks < kTsneKc / 4; ks++) {
            if (ks < steps) {
                float av[4], bv[4];
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    av[b] = sa[wr * 64 + b * 16 + lr][ks * 4 + lk];
                    bv[b] = sb[wc * 64 + b * 16 + lr][ks * 4 + lk];
                }
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int cb = 0; cb < 4; cb++)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[cb], av[rb], acc[rb][cb], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // (|a|^2 + |b|^2) - 2 g, clamped at zero (tsne.rs:183-187). Norms are padded to a
    // whole tile, so they load without checks.
    const uint32_t rbase = r0 + wr * 64 + lr, cbase = c0 + wc * 64 + 4 * lk;
    float nrow[4];
    float4 ncol[4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) nrow[rb] = norms[rbase + rb * 16];
#pragma unroll
    for (int cb = 0; cb < 4; cb++) ncol[cb] = *reinterpret_cast<const float4*>(norms + cbase + cb * 16);
    const bool mirror = blockIdx.x != blockIdx.y;
    // CHECKED: the tile crosses row or column n. VEC: rows start 16-byte aligned (n % 4 == 0).
    auto finish = [&](auto checked, auto vec) __attribute__((always_inline)) {
        constexpr bool kChecked = decltype(checked)::value, kVec = decltype(vec)::value;
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const uint32_t row = rbase + rb * 16;
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                const uint32_t col = cbase + cb * 16;
                const float nc[4] = {ncol[cb].x, ncol[cb].y, ncol[cb].z, ncol[cb].w};
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float sum = nrow[rb] + nc[r];
                    const float e = sum - 2.0f * acc[rb][cb][r];
                    v[r] = e > 0.0f ? e : 0.0f;
                }
                float* dst = out + static_cast<size_t>(row) * n + col;
                if (kVec && (!kChecked || (row < n && col + 3 < n))) {
                    *reinterpret_cast<float4*>(dst) = make_float4(v[0], v[1], v[2], v[3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (!kChecked || (row < n && col + r < n)) dst[r] = v[r];
                }
                if (mirror) {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (!kChecked || (row < n && col + r < n))
                            out[static_cast<size_t>(col + r) * n + row] = v[r];
                }
            }
        }
    };
    const bool interior = r0 + kTsneTile <= n && c0 + kTsneTile <= n;
    if (interior && n % 4 == 0 && mirror) {
        // Direct block: one float4 per (rb, cb). Mirrored block: each half of the wave's
        // columns is written to LDS transposed (column-major rows of 64), then read back
        // as float4 runs along the mirrored rows: 16 lanes store one 256-byte row segment.
        float* T = lds + w * (kMirrorCols * kMirrorStride);
#pragma unroll
        for (int rb = 0; rb < 4; rb++) {
            const uint32_t row = rbase + rb * 16;
#pragma unroll
            for (int cb = 0; cb < 4; cb++) {
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float sum = nrow[rb] + (r == 0 ? ncol[cb].x : r == 1 ? ncol[cb].y : r == 2 ? ncol[cb].z : ncol[cb].w);
                    const float e = sum - 2.0f * acc[rb][cb][r];
                    v[r] = e > 0.0f ? e : 0.0f;
                    acc[rb][cb][r] = v[r];
                }
                *reinterpret_cast<float4*>(out + static_cast<size_t>(row) * n + cbase + cb * 16) =
                    make_float4(v[0], v[1], v[2], v[3]);
            }
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            __syncthreads();
#pragma unroll
            for (int cbh = 0; cbh < 2; cbh++) {
                const int cb = 2 * h + cbh;
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        T[(cbh * 16 + 4 * lk + r) * kMirrorStride + rb * 16 + lr] = acc[rb][cb][r];
            }
            __syncthreads();
            // 32 columns x 16 float4: column c = i / 16, rows 4 (i % 16) .. + 3 of the wave's 64.
#pragma unroll
            for (int m = 0; m < 8; m++) {
                const int i = lane + 64 * m;
                const int c = i >> 4, q4 = i & 15;
                const float4 val = *reinterpret_cast<const float4*>(T + c * kMirrorStride + 4 * q4);
                const uint32_t col = c0 + wc * 64 + h * 32 + c;
                const uint32_t row = r0 + wr * 64 + 4 * q4;
                *reinterpret_cast<float4*>(out + static_cast<size_t>(col) * n + row) = val;
            }
        }
        return;
    }
    if (n % 4 == 0) {
        if (interior) finish(std::false_type{}, std::true_type{});
        else finish(std::true_type{}, std::true_type{});
    } else {
        if (interior) finish(std::false_type{}, std::false_type{});
        else finish(std::true_type{}, std::false_type{});
    }
}

}  // namespace

TsneLayout tsne_layout(size_t n, size_t d) {
    TsneLayout l{};
    auto align = [](size_t v) { return (v + 255) / 256 * 256; };
    l.dpad = static_cast<uint32_t>((d + 3) / 4 * 4);
    size_t at = 0;
    l.acc = at;
    at = align(at + d * kAccWords * sizeof(unsigned long long));
    l.mean = at;
    at = align(at + d * sizeof(float));
    l.coords = at;
    at = align(at + n * l.dpad * sizeof(float));
    l.norms = at;
    at = align(at + (n + kTsneTile - 1) / kTsneTile * kTsneTile * sizeof(float));
    l.flags = at;
    at = align(at + sizeof(uint32_t));
    l.total = at;
    return l;
}

hipError_t tsne_enqueue(const float* points, size_t n, size_t d, void* workspace,
                        const TsneLayout& l, float* out, hipStream_t stream) {
    char* base = static_cast<char*>(workspace);
    auto* acc = reinterpret_cast<unsigned long long*>(base + l.acc);
    auto* mean = reinterpret_cast<float*>(base + l.mean);
    auto* coords = reinterpret_cast<float*>(base + l.coords);
    auto* norms = reinterpret_cast<float*>(base + l.norms);
    auto* flags = reinterpret_cast<uint32_t*>(base + l.flags);
    const uint32_t n32 = static_cast<uint32_t>(n), d32 = static_cast<uint32_t>(d);
    launch_fill_u32(reinterpret_cast<uint32_t*>(acc), d * kAccWords * 2, 0u, stream);
    launch_fill_u32(flags, 1, 0u, stream);
    const uint32_t rows = colsum_rows(n32);
    const dim3 sum_grid((n32 + rows - 1) / rows, (d32 + kWave - 1) / kWave);
    hipLaunchKernelGGL(tsne_colsum_kernel, sum_grid, dim3(256), 0, stream, points, n32, d32, acc);
    hipLaunchKernelGGL(tsne_mean_kernel, dim3((d32 + 255) / 256), dim3(256), 0, stream, acc, n32, d32, mean);
    const uint32_t npad = (n32 + kTsneTile - 1) / kTsneTile * kTsneTile;
    hipLaunchKernelGGL(tsne_centre_kernel, dim3((npad + 255) / 256), dim3(256), 0, stream, points, n32, npad,
                       d32, l.dpad, mean, coords, norms, flags);
    const unsigned tiles = (n32 + kTsneTile - 1) / kTsneTile;
    hipLaunchKernelGGL(tsne_tile_kernel, dim3(tiles, tiles), dim3(256), 0, stream, coords, norms, n32, l.dpad,
                       out);
    return hipGetLastError();
}

}  // namespace scrust_hip
