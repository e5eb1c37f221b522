// kNN kernels for gfx950.
//
// Pipeline for one embedding of n cells x d dims (one HIP stream):
//   1. colsum / colmean   exact f32 column sums -> f64 means (Euclidean only)
//   2. prepare            centred (or unit-scaled) f32 coordinates, f32 fma squared norms,
//                         packed MFMA operands, max norm, non-finite flag
//   3. pass_bf16/pass_fp32  MFMA tile contraction of |b|^2 - 2 a.b fused with per-row
//                         candidate selection (KP = k + margin candidates kept per row)
//   4. rerank             fp32 direct differences of the KP candidates, (d^2, id) order,
//                         fp64 boundary near-tie rule, completeness certificate
//   5. exact              fp32 direct differences over every reference for rows listed by
//                         the rerank (or for every row on the exact path)
// The reference's semantics these restate are cited at each kernel.
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

constexpr int kColsumRows = 2048;

__global__ __launch_bounds__(64) void colsum_kernel(const float* __restrict__ x, size_t n,
                                                    int d, unsigned long long* acc_out) {
    __shared__ long long acc[kWave][kAccWords + 1];
    const int lane = threadIdx.x;
    const int col = blockIdx.y * kWave + lane;
    for (int w = 0; w < kAccWords; w++) acc[lane][w] = 0;
    const size_t r0 = static_cast<size_t>(blockIdx.x) * kColsumRows;
    size_t r1 = r0 + kColsumRows;
    if (r1 > n) r1 = n;
    if (col < d) {
        for (size_t r = r0; r < r1; r++) exact_sum_add(acc[lane], x[r * d + col]);
        for (int w = 0; w < kAccWords; w++) {
            long long v = acc[lane][w];
            if (v) atomicAdd(&acc_out[col * kAccWords + w], static_cast<unsigned long long>(v));
        }
    }
}

__global__ void colmean_kernel(const unsigned long long* acc, int d, size_t n, double* mean) {
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= d) return;
    long long words[kAccWords];
    for (int w = 0; w < kAccWords; w++) words[w] = static_cast<long long>(acc[col * kAccWords + w]);
    mean[col] = exact_sum_value(words) / static_cast<double>(n);
}

void launch_column_means(const float* x, size_t n, int d, unsigned long long* acc, double* mean,
                         hipStream_t stream) {
    launch_fill_u32(reinterpret_cast<uint32_t*>(acc), static_cast<size_t>(d) * kAccWords * 2, 0u,
                    stream);
    dim3 grid(static_cast<unsigned>((n + kColsumRows - 1) / kColsumRows),
              static_cast<unsigned>((d + kWave - 1) / kWave));
    hipLaunchKernelGGL(colsum_kernel, grid, dim3(kWave), 0, stream, x, n, d, acc);
    hipLaunchKernelGGL(colmean_kernel, dim3((d + 63) / 64), dim3(64), 0, stream, acc, d, n, mean);
}

// ===========================================================================
// 2. Coordinates, norms and MFMA operands.
//    Euclidean: (x as f64 - mean) rounded once (neighbors.rs:155-160); norms by f32
//    fma in dimension order (knn.rs:148-156). Cosine: x / |x| with the norm in f64.
//    Operand layouts are described in knn_pass.hip. Thread n writes the sentinel
//    reference row (norm feature +inf).
// ===========================================================================

__global__ void prepare_kernel(const float* __restrict__ x, size_t n, int d, int metric,
                               const double* __restrict__ mean, float* __restrict__ coords,
                               float* __restrict__ norms, int operand, uint16_t* __restrict__ pack_ref,
                               uint16_t* __restrict__ pack_query, float* __restrict__ pack_f32,
                               int nch, int group, unsigned* __restrict__ max_norm_bits,
                               unsigned* __restrict__ flags) {
    const size_t row = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    constexpr uint16_t kBf16Inf = 0x7F80u, kBf16MinusHalf = 0xBF00u;
    float nrm = 0.0f;
    bool finite = true;
    if (row < n) {
        const float* src = x + row * d;
        float* dst = coords + row * d;
        double length = 0.0;
        if (metric == kMetricCosine) {
            double s = 0.0;
            for (int j = 0; j < d; j++) s = __builtin_fma((double)src[j], (double)src[j], s);
            length = __builtin_sqrt(s);
        }
        for (int j = 0; j < d; j++) {
            float v = src[j];
            finite = finite && __builtin_isfinite(v);
            float c;
            if (metric == kMetricCosine)
                c = length > 0.0 ? static_cast<float>(static_cast<double>(v) / length) : 0.0f;
            else
                c = static_cast<float>(static_cast<double>(v) - mean[j]);
            dst[j] = c;
            nrm = __builtin_fmaf(c, c, nrm);
        }
        norms[row] = nrm;
        if (operand == kOperandBf16) {
            const int kc = 32 * nch;
            uint16_t* r = pack_ref + row * kc;
            uint16_t* q = pack_query + row * kc;
            for (int j = 0; j < d; j++) {
                float c = dst[j];
                uint16_t h = bf16_rn(c);
                uint16_t l = bf16_rn(c - bf16_value(h));
                r[j] = h;
                r[d + j] = l;
                r[2 * d + j] = h;
                q[j] = h;
                q[d + j] = h;
                q[2 * d + j] = l;
            }
            const uint16_t nh = bf16_rn(nrm);
            const float rest = nrm - bf16_value(nh);
            const uint16_t nl = bf16_rn(rest);
            const uint16_t nl2 = bf16_rn(rest - bf16_value(nl));
            r[3 * d] = nh;
            r[3 * d + 1] = nl;
            r[3 * d + 2] = nl2;
            q[3 * d] = kBf16MinusHalf;
            q[3 * d + 1] = kBf16MinusHalf;
            q[3 * d + 2] = kBf16MinusHalf;
            for (int p = 3 * d + 3; p < kc; p++) {
                r[p] = 0;
                q[p] = 0;
            }
        } else if (operand == kOperandF32) {
            float* p = pack_f32 + row * (4 * group);
            for (int s = 0; s < 4 * group; s++) {
                int g = s / group, c = s % group;
                int f = g * nch + c;
                p[s] = c >= nch ? 0.0f : (f < d ? dst[f] : (f == d ? nrm : 0.0f));
            }
        }
    } else if (row == n) {
        if (operand == kOperandBf16) {
            const int kc = 32 * nch;
            uint16_t* r = pack_ref + row * kc;
            for (int p = 0; p < kc; p++) r[p] = (p == 3 * d) ? kBf16Inf : 0;
        } else if (operand == kOperandF32) {
            float* p = pack_f32 + row * (4 * group);
            for (int s = 0; s < 4 * group; s++) {
                int g = s / group, c = s % group;
                p[s] = (c < nch && g * nch + c == d) ? kInf : 0.0f;
            }
        }
    }
    // Wave-reduced max norm (norms are >= 0, so their bits order as unsigned ints).
    unsigned bits = row < n ? __float_as_uint(nrm) : 0u;
    for (int off = 32; off > 0; off >>= 1) {
        unsigned other = __shfl_xor(bits, off, kWave);
        bits = other > bits ? other : bits;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(max_norm_bits, bits);
    if (!finite) atomicOr(flags, kFlagNonFinite);
}

void launch_prepare(const PrepareArgs& a, hipStream_t stream) {
    unsigned blocks = static_cast<unsigned>((a.n + 1 + 255) / 256);
    hipLaunchKernelGGL(prepare_kernel, dim3(blocks), dim3(256), 0, stream, a.x, a.n, a.d,
                       a.metric, a.mean, a.coords, a.norms, a.operand, a.pack_ref, a.pack_query,
                       a.pack_f32, a.nch, a.group, a.max_norm_bits, a.flags);
}

// ===========================================================================
// 4. Rerank (crates/scrust-gpu/src/kernels/knn.rs:228-244 arithmetic; order by
//    (d^2, id) as knn.rs:179-182 / neighbors.rs:181-185), fp64 boundary rule and the
//    completeness certificate. KP lanes per query row.
// ===========================================================================

template <int KP>
__global__ __launch_bounds__(256) void rerank_kernel(RerankArgs a) {
    constexpr int QPW = kWave / KP;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int seg = lane & ~(KP - 1);
    const int j = lane & (KP - 1);
    const uint32_t q = (blockIdx.x * 4 + w) * QPW + lane / KP;
    const bool row_ok = q < a.n;
    const uint32_t qq = row_ok ? q : a.n - 1;
    const int d = a.d;
    const float* xq = a.coords + static_cast<size_t>(qq) * d;
    const float nq = a.norms[qq];

    const uint64_t raw = a.cand[static_cast<size_t>(qq) * KP + j];
    const uint32_t id = key_id(raw);
    const bool present = raw != kNoKey;
    const bool valid = present && id != qq && id < a.n;
    float dsq = kInf;
    if (valid) dsq = direct_square(xq, a.coords + static_cast<size_t>(id) * d, d, nq, a.norms[id]);
    uint64_t key = valid ? make_key(dsq, id) : kNoKey;
    key = segment_sort<KP>(key, lane);

    // Segment-local bookkeeping (every lane of a segment computes the same values).
    const uint64_t seg_mask = (KP == 64) ? ~0ull : (((1ull << KP) - 1) << seg);
    const int n_valid = __popcll(__ballot(key != kNoKey) & seg_mask);
    const bool full = (__ballot(present) & seg_mask) == seg_mask;
    const float my_d = key_value(key);
    const uint32_t my_id = key_id(key);
    const int k = a.k;
    const float t_f = __shfl(my_d, seg + k - 1, kWave);
    const double t = static_cast<double>(t_f);
    const double tau = t * static_cast<double>(d + 2) * 0x1p-23;
    const bool in_band = (j < n_valid) && __builtin_fabs(static_cast<double>(my_d) - t) <= tau;
    const uint64_t band_mask = __ballot(in_band) & seg_mask;
    const bool has_next = n_valid > k;
    const double next = static_cast<double>(__shfl(my_d, seg + (k < KP ? k : KP - 1), kWave));
    const bool tie = has_next && (next - t <= tau);

    int pos = j;
    if (__ballot(tie && row_ok) != 0) {
        const int lo = band_mask ? __builtin_ctzll(band_mask) - seg : 0;
        const int hi = band_mask ? 63 - __builtin_clzll(band_mask) - seg : -1;
        double d64 = 0.0;
        if (in_band && my_d != 0.0f)
            d64 = direct_square_f64(xq, a.coords + static_cast<size_t>(my_id) * d, d);
        int rank = 0;
        for (int i = 0; i < KP; i++) {
            const double od = __shfl(d64, seg + i, kWave);
            const uint32_t oid = __shfl(my_id, seg + i, kWave);
            if (i >= lo && i <= hi && (od < d64 || (od == d64 && oid < my_id))) rank++;
        }
        if (tie && in_band) pos = lo + rank;
    }

    // Certificate: every reference outside the kept list has approximate score
    // >= theta, so its fp32 squared distance is >= L below; if L clears the band's
    // upper edge and every snapping threshold, the kept list holds every reference
    // the oracle could place in the first k (or in the tie band).
    bool certified = true;
    if (full) {
        const double u = 0x1p-24;
        const double nqd = static_cast<double>(nq);
        const double nmax = static_cast<double>(__uint_as_float(*a.max_norm_bits));
        const double gamma = d * u / (1.0 - d * u);
        const double qn = __builtin_sqrt(nqd * (1.0 + gamma));
        const double rn = __builtin_sqrt(nmax * (1.0 + gamma));
        // |score + |a|^2 - exact d^2| <= err: f32 norms, the bf16 split of the
        // coordinates and of |b|^2, and the MFMA accumulation over every packed term.
        const double err = 1.25 * (gamma * (nqd + nmax) * 1.01 + 2.0 * a.split_rel * qn * rn +
                                   2.0 * a.acc_rel * (1.02 * qn * rn + 0.51 * nmax * (1.0 + gamma)) +
                                   a.norm_rel * nmax * (1.0 + gamma)) + 1e-30;
        const double cd = (d + 3) * u * 1.01;
        const double lower = (static_cast<double>(a.theta[qq]) + nqd - err) * (1.0 - cd);
        const double snap_max = (d + 2) * 0x1p-23 * (nqd + nmax) * (1.0 + 4.0 * u);
        const double need = (t + tau) > snap_max ? (t + tau) : snap_max;
        certified = lower > need;
    }

    if (!row_ok) return;
    if (certified) {
        if (pos < k && j < n_valid) {
            a.out_indices[static_cast<size_t>(q) * k + pos] = my_id;
            a.out_distances[static_cast<size_t>(q) * k + pos] =
                a.metric == kMetricCosine ? 0.5f * my_d : __builtin_sqrtf(my_d);
        }
    } else if (j == 0) {
        const uint32_t slot = atomicAdd(a.fail_count, 1u);
        a.fail_list[slot] = q;
    }
}

void launch_rerank(const RerankArgs& a, hipStream_t stream) {
    const unsigned rows_per_block = 4 * (kWave / a.kp);
    const unsigned blocks = static_cast<unsigned>((a.n + rows_per_block - 1) / rows_per_block);
    if (a.kp == 32)
        hipLaunchKernelGGL(rerank_kernel<32>, dim3(blocks), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(rerank_kernel<64>, dim3(blocks), dim3(256), 0, stream, a);
}

// ===========================================================================
// 5. Exact rows: one workgroup per query row scans every reference with the f32
//    direct difference (the Metal kernel's per-query scan, knn.rs:222-244), keeping
//    the smallest keys in LDS; then the fp64 boundary rule.
// ===========================================================================

constexpr int kExactBuf = 1024;
constexpr int kExactKeep = 512;

__device__ void lds_sort(uint64_t* keys, int count) {
    // Bitonic sort of kExactBuf keys by 256 threads (entries past count are kNoKey).
    for (int i = threadIdx.x; i < kExactBuf; i += blockDim.x)
        if (i >= count) keys[i] = kNoKey;
    __syncthreads();
    for (int size = 2; size <= kExactBuf; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int p = threadIdx.x; p < kExactBuf / 2; p += blockDim.x) {
                const int lo = (p / stride) * stride * 2 + (p % stride);
                const int hi = lo + stride;
                const bool asc = (lo & size) == 0;
                uint64_t x = keys[lo], y = keys[hi];
                if ((x > y) == asc) {
                    keys[lo] = y;
                    keys[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(256) void exact_kernel(ExactArgs a) {
    extern __shared__ float xq[];
    __shared__ uint64_t keys[kExactBuf];
    __shared__ uint32_t count;
    __shared__ uint64_t theta;
    __shared__ int dropped;
    __shared__ uint32_t out_id[kMaxK];
    __shared__ float out_d[kMaxK];
    __shared__ double bd[kExactKeep];
    __shared__ uint32_t bid[kExactKeep];
    __shared__ float bf[kExactKeep];
    const int tid = threadIdx.x;
    const int d = a.d, k = a.k;
    const uint32_t total = a.list ? *a.list_count : a.n;
    for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
        const uint32_t q = a.list ? a.list[item] : item;
        for (int j = tid; j < d; j += blockDim.x) xq[j] = a.coords[static_cast<size_t>(q) * d + j];
        if (tid == 0) {
            count = 0;
            theta = kNoKey;
            dropped = 0;
        }
        __syncthreads();
        const float nq = a.norms[q];
        for (uint32_t r0 = 0; r0 < a.n; r0 += blockDim.x) {
            const uint32_t r = r0 + tid;
            if (r < a.n && r != q) {
                const float s = direct_square(xq, a.coords + static_cast<size_t>(r) * d, d, nq, a.norms[r]);
                const uint64_t key = make_key(s, r);
                if (key < theta) {
                    const uint32_t slot = atomicAdd(&count, 1u);
                    keys[slot] = key;
                }
            }
            __syncthreads();
            if (count > static_cast<uint32_t>(kExactBuf) - blockDim.x) {
                lds_sort(keys, count);
                if (tid == 0) {
                    count = kExactKeep;
                    theta = keys[kExactKeep - 1];
                    dropped = 1;
                }
                __syncthreads();
            }
        }
        const int m = static_cast<int>(count);
        lds_sort(keys, m);
        if (tid == 0) {
            // (d^2, id) order, then the fp64 boundary rule on the sorted keys.
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
                while (hi + 1 < m && static_cast<double>(key_value(keys[hi + 1])) - t <= tau) hi++;
                // A band wider than kBandCap keeps the (d^2, id) order. After a
                // compaction m >= kExactKeep and lo < k <= kMaxK, so a band that reaches
                // the end of the kept keys is always wider than the cap.
                if (hi - lo + 1 > kBandCap || (hi + 1 == m && dropped)) {
                    if (hi - lo + 1 <= kBandCap) atomicOr(a.flags, kFlagBandOverflow);
                } else {
                    // Insertion sort of the band by (fp64 d^2, id); the band is small.
                    const int len = hi - lo + 1;
                    for (int i = 0; i < len; i++) {
                        const uint64_t key = keys[lo + i];
                        const float v = key_value(key);
                        const uint32_t id = key_id(key);
                        const double v64 = v == 0.0f ? 0.0 : direct_square_f64(xq, a.coords + static_cast<size_t>(id) * d, d);
                        int s = i;
                        while (s > 0 && (bd[s - 1] > v64 || (bd[s - 1] == v64 && bid[s - 1] > id))) {
                            bd[s] = bd[s - 1];
                            bid[s] = bid[s - 1];
                            bf[s] = bf[s - 1];
                            s--;
                        }
                        bd[s] = v64;
                        bid[s] = id;
                        bf[s] = v;
                    }
                    for (int p = lo; p < k; p++) {
                        out_id[p] = bid[p - lo];
                        out_d[p] = bf[p - lo];
                    }
                }
            }
        }
        __syncthreads();
        for (int j = tid; j < k; j += blockDim.x) {
            a.out_indices[static_cast<size_t>(q) * k + j] = out_id[j];
            a.out_distances[static_cast<size_t>(q) * k + j] =
                a.metric == kMetricCosine ? 0.5f * out_d[j] : __builtin_sqrtf(out_d[j]);
        }
        __syncthreads();
    }
}

void launch_exact(const ExactArgs& a, unsigned blocks, hipStream_t stream) {
    if (blocks == 0) return;
    hipLaunchKernelGGL(exact_kernel, dim3(blocks), dim3(256), a.d * sizeof(float), stream, a);
}

}  // namespace scrust_hip
