// Device helpers shared by the kNN kernels (gfx950, wave64).
//
// Every distance the library reports is formed here, with the arithmetic the
// reference holds its Metal kernel to (crates/scrust-gpu/src/kernels/knn.rs:229-240):
// f32 differences, f32 fma accumulation in dimension order, and the snap of squared
// distances below (d+2)*eps*(|a|^2+|b|^2) to zero. The translation unit is compiled
// with -ffp-contract=off, so every fma is one written here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "knn_launch.hpp"

namespace scrust_hip {

constexpr int kWave = 64;
constexpr float kInf = __builtin_inff();
constexpr uint64_t kNoKey = ~0ull;
constexpr float kF32Epsilon = 1.1920928955078125e-07f;  // 2^-23, f32::EPSILON
constexpr int kAccWords = 10;                          // exact-sum digits, 32 bits each

// Global-memory address space. Pointers read from a descriptor in memory are generic to
// the compiler, which then emits FLAT accesses; FLAT loads also count against the LDS
// wait counter, so every LDS wait would wait for outstanding HBM loads as well.
#define SCRUST_GLOBAL __attribute__((address_space(1)))

template <class T>
__device__ __forceinline__ SCRUST_GLOBAL T* global_ptr(T* p) {
    return (SCRUST_GLOBAL T*)p;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
// A native 16-byte vector (HIP's uint4 is a class with unions, which keeps arrays of it
// out of registers).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Keys: (squared distance, id) packed so that u64 order is (d^2, id) order.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float unordered(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__device__ __forceinline__ uint64_t make_key(float v, uint32_t id) {
    return (static_cast<uint64_t>(ordered(v)) << 32) | id;
}

__device__ __forceinline__ float key_value(uint64_t key) {
    return unordered(static_cast<uint32_t>(key >> 32));
}

__device__ __forceinline__ uint32_t key_id(uint64_t key) { return static_cast<uint32_t>(key); }

// ---------------------------------------------------------------------------
// Distances.
// ---------------------------------------------------------------------------

// Snapping threshold, evaluated left to right in f32 exactly as knn.rs:237 does.
__device__ __forceinline__ float snap_threshold(int d, float na, float nb) {
    return (static_cast<float>(d) + 2.0f) * kF32Epsilon * (na + nb);
}

// f32 direct difference with fma in dimension order, then the snap (knn.rs:228-240).
__device__ __forceinline__ float direct_square(const float* __restrict__ a,
                                               const float* __restrict__ b, int d, float na,
                                               float nb) {
    float s = 0.0f;
    for (int j = 0; j < d; j++) {
        float delta = a[j] - b[j];
        s = __builtin_fmaf(delta, delta, s);
    }
    return s < snap_threshold(d, na, nb) ? 0.0f : s;
}

// direct_square over two rows in global memory, the loads widened to VEC floats (rows
// 4 * VEC-byte aligned: VEC divides d) and issued 16 floats at a time before their
// differences (one wait per batch instead of one per load); the fma chain is the same, in
// dimension order.
template <int VEC>
__device__ __forceinline__ float direct_square_vec(const float* __restrict__ a,
                                                   const float* __restrict__ b, int d, float na,
                                                   float nb) {
    typedef float vec_t __attribute__((ext_vector_type(VEC)));
    constexpr int U = 16 / VEC;
    const SCRUST_GLOBAL vec_t* av = reinterpret_cast<const SCRUST_GLOBAL vec_t*>(global_ptr(a));
    const SCRUST_GLOBAL vec_t* bv = reinterpret_cast<const SCRUST_GLOBAL vec_t*>(global_ptr(b));
    const int nv = d / VEC;
    float s = 0.0f;
    int j = 0;
    for (; j + U <= nv; j += U) {
        vec_t x[U], y[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            x[u] = av[j + u];
            y[u] = bv[j + u];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int e = 0; e < VEC; e++) {
                const float delta = x[u][e] - y[u][e];
                s = __builtin_fmaf(delta, delta, s);
            }
        }
    }
    for (; j < nv; j++) {
        const vec_t x = av[j], y = bv[j];
#pragma unroll
        for (int e = 0; e < VEC; e++) {
            const float delta = x[e] - y[e];
            s = __builtin_fmaf(delta, delta, s);
        }
    }
    return s < snap_threshold(d, na, nb) ? 0.0f : s;
}

// f64 recomputation for boundary near-ties: f64 differences of the same f32
// coordinates, f64 fma in dimension order.
__device__ __forceinline__ double direct_square_f64(const float* __restrict__ a,
                                                    const float* __restrict__ b, int d) {
    double s = 0.0;
    for (int j = 0; j < d; j++) {
        double delta = static_cast<double>(a[j]) - static_cast<double>(b[j]);
        s = __builtin_fma(delta, delta, s);
    }
    return s;
}

// ---------------------------------------------------------------------------
// bf16 rounding (round to nearest even; finite inputs).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint16_t bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ float bf16_value(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// Chunk m (16 bytes: K indices k_index(m, 0..7, nch)) of the bf16 query operand of a row with
// coordinates row[0 .. d): [h | -1/2 -1/2 -1/2 | h | l | 0], h = bf16(c), l = bf16(c - h). The
// parts against the reference's [h | norm terms | l | h | 0]: h.h, the -1/2 weights of the
// norm terms, h_query.l_reference and l_query.h_reference.
template <class Row>
__device__ __forceinline__ u32x4 bf16_query_chunk(Row row, int d, int m, int nch) {
    constexpr uint16_t kMinusHalf = 0xBF00u;
    uint32_t w[4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        uint16_t v[2];
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++) {
            const int f = k_index(m, e + h2, nch);
            v[h2] = 0;
            if (f < d || (f >= d + 3 && f < 3 * d + 3)) {
                const int part = f < d ? 0 : (f < 2 * d + 3 ? 1 : 2);
                const float c = row[part == 0 ? f : f - 3 - part * d];
                const uint16_t h = bf16_rn(c);
                v[h2] = part == 2 ? bf16_rn(c - bf16_value(h)) : h;
            } else if (f < d + 3) {
                v[h2] = kMinusHalf;
            }
        }
        w[e >> 1] = static_cast<uint32_t>(v[0]) | static_cast<uint32_t>(v[1]) << 16;
    }
    return u32x4{w[0], w[1], w[2], w[3]};
}

// ---------------------------------------------------------------------------
// Local-origin fp16 operands (the fp16 candidate pass): per-dataset scale.
// ---------------------------------------------------------------------------

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// sigma = 2^s with sqrt(max_norm) * sigma in [2^12, 2^13): every shifted, scaled coordinate
// |c - o| * sigma (o a mean of rows, so |o| <= sqrt(max_norm)) stays below 2^14, far inside
// fp16's range, and its squared norm below 2^28. Clamped to [2^-100, 2^100]; 1 when every
// norm is 0. A power of two, so scaling is exact.
__device__ __forceinline__ float local_scale(float max_norm) {
    if (!(max_norm > 0.0f)) return 1.0f;
    int e = 0;
    (void)__builtin_frexpf(__builtin_sqrtf(max_norm), &e);  // sqrt(max_norm) < 2^e
    int s = 13 - e;
    s = s < -100 ? -100 : (s > 100 ? 100 : s);
    return __builtin_ldexpf(1.0f, s);
}

// The smallest fp16 at or above v (v >= 0 finite, below 65504).
__device__ __forceinline__ _Float16 f16_up(float v) {
    _Float16 h = static_cast<_Float16>(v);
    if (static_cast<float>(h) < v) h = __builtin_bit_cast(_Float16, static_cast<uint16_t>(__builtin_bit_cast(uint16_t, h) + 1u));
    return h;
}

// ---------------------------------------------------------------------------
// Exact sums of f32 values: each f32 is an integer multiple of 2^-149; the sum is
// kept as a two's-complement integer in 32-bit digits held in int64 words and is
// rounded once to f64 at the end, so any summation order gives the same value.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void exact_sum_add(long long* acc, float v) {
    uint32_t bits = __float_as_uint(v);
    uint32_t biased = (bits >> 23) & 0xFFu;
    uint32_t mantissa = bits & 0x7FFFFFu;
    uint64_t m = biased ? (mantissa | 0x800000u) : mantissa;
    int position = biased ? static_cast<int>(biased) - 1 : 0;
    if (m == 0) return;
    uint64_t shifted = m << (position & 31);
    int word = position >> 5;
    long long lo = static_cast<long long>(shifted & 0xFFFFFFFFull);
    long long hi = static_cast<long long>(shifted >> 32);
    if (bits >> 31) {
        acc[word] -= lo;
        acc[word + 1] -= hi;
    } else {
        acc[word] += lo;
        acc[word + 1] += hi;
    }
}

__device__ inline double exact_sum_value(const long long* words) {
    uint32_t digit[kAccWords];
    long long carry = 0;
    for (int i = 0; i < kAccWords; i++) {
        long long v = words[i] + carry;
        digit[i] = static_cast<uint32_t>(static_cast<unsigned long long>(v) & 0xFFFFFFFFull);
        carry = v >> 32;
    }
    double sign = 1.0;
    if (carry < 0) {
        sign = -1.0;
        unsigned long long c = 1;
        for (int i = 0; i < kAccWords; i++) {
            unsigned long long v = static_cast<unsigned long long>(~digit[i]) + c;
            digit[i] = static_cast<uint32_t>(v);
            c = v >> 32;
        }
    }
    int top = -1;
    for (int i = kAccWords - 1; i >= 0; i--) {
        if (digit[i]) {
            top = i * 32 + 31 - __builtin_clz(digit[i]);
            break;
        }
    }
    if (top < 0) return 0.0;
    auto bit = [&](int b) -> uint32_t { return b < 0 ? 0u : (digit[b >> 5] >> (b & 31)) & 1u; };
    if (top <= 52) {
        unsigned long long v = 0;
        for (int b = top; b >= 0; b--) v = (v << 1) | bit(b);
        return sign * ldexp(static_cast<double>(v), -149);
    }
    unsigned long long m = 0;
    for (int b = top; b > top - 53; b--) m = (m << 1) | bit(b);
    uint32_t guard = bit(top - 53);
    uint32_t sticky = 0;
    for (int b = top - 54; b >= 0 && !sticky; b--) sticky |= bit(b);
    int exponent = top - 52 - 149;
    if (guard && (sticky || (m & 1ull))) {
        m += 1;
        if (m == (1ull << 53)) {
            m >>= 1;
            exponent += 1;
        }
    }
    return sign * ldexp(static_cast<double>(m), exponent);
}

// ---------------------------------------------------------------------------
// Wave-level bitonic sorts of u64 keys (ascending).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
    uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    lo = __shfl_xor(lo, mask, kWave);
    hi = __shfl_xor(hi, mask, kWave);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
    lo = __shfl(lo, src, kWave);
    hi = __shfl(hi, src, kWave);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Sort E*64 keys held E per lane; element index e = i*64 + lane.
template <int E>
__device__ __forceinline__ void wave_sort(uint64_t (&key)[E], int lane) {
#pragma unroll
    for (int size = 2; size <= E * kWave; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (stride >= kWave) {
                const int rs = stride / kWave;
#pragma unroll
                for (int i = 0; i < E; i++) {
                    if (i & rs) continue;
                    const int e = i * kWave + lane;
                    const bool asc = (e & size) == 0;
                    uint64_t x = key[i], y = key[i + rs];
                    bool swap = asc ? (x > y) : (x < y);
                    key[i] = swap ? y : x;
                    key[i + rs] = swap ? x : y;
                }
            } else {
#pragma unroll
                for (int i = 0; i < E; i++) {
                    const int e = i * kWave + lane;
                    const bool asc = (e & size) == 0;
                    const bool lower = (lane & stride) == 0;
                    uint64_t other = shfl_xor_u64(key[i], stride);
                    uint64_t lo = key[i] < other ? key[i] : other;
                    uint64_t hi = key[i] < other ? other : key[i];
                    key[i] = (lower == asc) ? lo : hi;
                }
            }
        }
    }
}

// Sort one key per lane within aligned segments of W lanes (W a power of two <= 64).
template <int W>
__device__ __forceinline__ uint64_t segment_sort(uint64_t key, int lane) {
    const int j = lane & (W - 1);
#pragma unroll
    for (int size = 2; size <= W; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const bool asc = (j & size) == 0;
            const bool lower = (j & stride) == 0;
            uint64_t other = shfl_xor_u64(key, stride);
            uint64_t lo = key < other ? key : other;
            uint64_t hi = key < other ? other : key;
            key = (lower == asc) ? lo : hi;
        }
    }
    return key;
}

// ---------------------------------------------------------------------------
// Wave-level selection of the smallest keys (ballot radix select).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Place of the j-th kept key: out[j], or with quarter > 0 back in the four segments
// (quarter keys at the start of each segment of `seg` slots).
__device__ __forceinline__ uint32_t kept_slot(uint32_t j, uint32_t quarter, uint32_t seg) {
    return quarter ? (j / quarter) * seg + j % quarter : j;
}

// Keeps the `keep` smallest keys (at least `keep` are valid), written unordered through
// kept_slot; returns the keep-th smallest score. Ballot radix select: a binary search over
// the ordered score bits, then over ids among keys that tie on the boundary score.
template <int E, class Out>
__device__ float select_keys(const uint64_t (&key)[E], int keep, int lane,
                             Out out, uint32_t quarter, uint32_t seg) {
    uint32_t lo = 0, hi = 0xFFFFFFFFu;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        int c = 0;
#pragma unroll
        for (int i = 0; i < E; i++) c += __popcll(__ballot(static_cast<uint32_t>(key[i] >> 32) <= mid));
        if (c >= keep) hi = mid; else lo = mid + 1;
    }
    const uint32_t v = lo;
    int below = 0, equal = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint32_t sc = static_cast<uint32_t>(key[i] >> 32);
        below += __popcll(__ballot(sc < v));
        equal += __popcll(__ballot(sc == v));
    }
    const int need = keep - below;
    uint32_t idcut = 0xFFFFFFFFu;
    if (equal > need) {
        uint32_t l2 = 0, h2 = 0xFFFFFFFFu;
        while (l2 < h2) {
            const uint32_t mid = l2 + (h2 - l2) / 2;
            int c = 0;
#pragma unroll
            for (int i = 0; i < E; i++)
                c += __popcll(__ballot(static_cast<uint32_t>(key[i] >> 32) == v &&
                                       static_cast<uint32_t>(key[i]) <= mid));
            if (c >= need) h2 = mid; else l2 = mid + 1;
        }
        idcut = l2;
    }
    uint32_t base = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint32_t sc = static_cast<uint32_t>(key[i] >> 32);
        const bool kept = sc < v || (sc == v && static_cast<uint32_t>(key[i]) <= idcut);
        const uint64_t m = __ballot(kept);
        if (kept) out[kept_slot(base + lanes_below(m), quarter, seg)] = key[i];
        base += static_cast<uint32_t>(__popcll(m));
    }
    return unordered(v);
}

}  // namespace scrust_hip
