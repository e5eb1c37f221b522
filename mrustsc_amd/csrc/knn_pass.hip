// Candidate pass: MFMA tile contraction of the squared-distance expansion fused with
// per-row candidate selection (gfx950, wave64).
//
// A workgroup owns 256 query rows (4 waves x 64) and streams every reference through LDS
// in tiles of 64, double buffered, read from HBM with 16-byte coalesced loads. A wave
// keeps its 64 rows' A fragments in registers for the whole scan; the reference tile
// supplies the B fragments. The reference's squared norm is one more packed feature, so
// the MFMA accumulator is acc = a.b - |b|^2 / 2 and the score s = |b|^2 - 2 a.b (the row's
// |a|^2 is a constant shift) is -2 acc: a lane tests each of its 16 accumulators against
// its rows' thresholds. References sit on the MFMA's M side and queries on its N side, so
// the four accumulators a lane holds per 16x16 block belong to one query row and are
// tested with one max and one compare; a block's tests form one wave-uniform branch.
//
// Selection: a pair with s < theta[row] is appended to the row's kCap-slot buffer in
// global memory (slot from an LDS counter). After each tile, a row whose buffer could
// overflow in the next tile is compacted by its wave (ballot radix select of the KP
// smallest keys); theta becomes the KP-th smallest score. theta only decreases, so every
// reference that is not kept has s >= theta_final: the bound the rerank certifies with.
//
// Operands (written by prepare_kernel):
//   bf16: coordinate c = h + l + r, h = bf16(c), l = bf16(c - h); reference rows
//         [h | l | h | nh nl nl2 | 0], query rows [h | h | l | -1/2 -1/2 -1/2 | 0]
//         (|b|^2 = nh + nl + nl2 + O(2^-24)); 32 * nch packed features per row.
//   fp32: feature f of 4 * group slots sits at (f / nch) * group + f % nch; features are
//         [c | |b|^2 | 0] for references; the kernel puts -1/2 at feature d for queries.
//   Row n of the reference operands is a sentinel whose norm feature is +inf, so the
//   padding columns of the last tile score -inf and are never candidates.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "knn_device.hpp"
#include "knn_launch.hpp"

namespace scrust_hip {

constexpr int kTileRefs = 64;       // references per LDS tile (4 column blocks of 16)
constexpr int kRowsPerBlock = 256;  // query rows per workgroup (4 waves x 64)

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Keep the `keep` smallest of the `count` keys at rowbuf[0..count) (count >= keep): they
// are written, unordered, to out[0..keep); returns the keep-th smallest score. Ballot
// radix select: a binary search over the ordered score bits, then over ids among keys
// that tie on the boundary score.
template <int E>
__device__ float select_smallest(const uint64_t* rowbuf, uint32_t count, int keep, int lane,
                                 uint64_t* out) {
    uint64_t key[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint32_t e = i * kWave + lane;
        key[i] = e < count ? rowbuf[e] : kNoKey;
    }
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
        const uint32_t s = static_cast<uint32_t>(key[i] >> 32);
        below += __popcll(__ballot(s < v));
        equal += __popcll(__ballot(s == v));
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
        const uint32_t s = static_cast<uint32_t>(key[i] >> 32);
        const bool kept = s < v || (s == v && static_cast<uint32_t>(key[i]) <= idcut);
        const uint64_t m = __ballot(kept);
        if (kept) out[base + lanes_below(m)] = key[i];
        base += static_cast<uint32_t>(__popcll(m));
    }
    return unordered(v);
}

template <int OP, int NCH>
struct Operand;

template <int NCH>
struct Operand<kOperandBf16, NCH> {
    static constexpr int kRowU4 = 4 * NCH;  // 16-byte chunks per packed row (32*NCH bf16)
    typedef bf16x8 Frag;
};

template <int NCH>
struct Operand<kOperandF32, NCH> {
    static constexpr int kGroup = (NCH + 3) / 4 * 4;  // floats per lane group, 16-B aligned
    static constexpr int kRowU4 = kGroup;              // 4 groups of kGroup floats
    typedef float Frag;
};

// At most two waves per SIMD: the allocator then keeps the tile prefetch in registers
// instead of spilling it to reach a higher occupancy the LDS budget would allow.
template <int OP, int NCH, int KP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) void pass_kernel(PassArgs p) {
    typedef Operand<OP, NCH> Op;
    constexpr int ROW_U4 = Op::kRowU4;
    constexpr int STRIDE = ROW_U4 + 1;  // LDS row stride in 16-byte chunks (bank spread)
    constexpr int LOADS = kTileRefs * ROW_U4 / 256;
    constexpr int E = kCap / kWave;
    static_assert(kTileRefs * ROW_U4 % 256 == 0, "tile load must split evenly");
    __shared__ uint4 tile[2][kTileRefs * STRIDE];
    __shared__ uint32_t cnt[kRowsPerBlock];
    __shared__ float phl[kRowsPerBlock];

    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, lc = lane & 15;
    const uint32_t n = p.n;
    const uint32_t qbase = blockIdx.x * kRowsPerBlock;
    const uint4* ref_rows = reinterpret_cast<const uint4*>(OP == kOperandBf16
                                                               ? static_cast<const void*>(p.pack_ref)
                                                               : static_cast<const void*>(p.pack_f32));

    // Query fragments: the B operand of every MFMA, held for the whole scan. Query block
    // qb's fragment in lane l carries row qb*16 + l%16.
    typename Op::Frag a[4][NCH];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        uint32_t row = qbase + w * kWave + rb * 16 + lc;
        if (row >= n) row = n - 1;
        if constexpr (OP == kOperandBf16) {
            const uint4* src = reinterpret_cast<const uint4*>(p.pack_query) + static_cast<size_t>(row) * ROW_U4 + g * NCH;
#pragma unroll
            for (int c = 0; c < NCH; c++) a[rb][c] = __builtin_bit_cast(bf16x8, src[c]);
        } else {
            const float* src = p.pack_f32 + static_cast<size_t>(row) * 4 * Op::kGroup + g * Op::kGroup;
#pragma unroll
            for (int c = 0; c < NCH; c++)
                a[rb][c] = (g * NCH + c == p.d) ? -0.5f : src[c];
        }
    }
    cnt[tid] = 0;
    phl[tid] = -kInf;
    // phi[qb] = -theta/2 of this lane's query row qb*16 + l%16.
    float phi[4];
#pragma unroll
    for (int qb = 0; qb < 4; qb++) phi[qb] = -kInf;

    const uint32_t n_tiles = (n + kTileRefs - 1) / kTileRefs;
    uint4 pre[LOADS];
    // Global -> registers for tile t (row n is the sentinel); registers -> LDS buffer b.
    // Indices are compile-time constants so `pre` stays in registers.
    auto load_tile = [&](uint32_t t) __attribute__((always_inline)) {
        static_for<0, LOADS>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int u = tid + 256 * i;
            const int rl = u / ROW_U4, part = u % ROW_U4;
            uint32_t ref = t * kTileRefs + rl;
            if (ref > n) ref = n;
            pre[i] = ref_rows[static_cast<size_t>(ref) * ROW_U4 + part];
        });
    };
    auto store_tile = [&](int b) __attribute__((always_inline)) {
        static_for<0, LOADS>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int u = tid + 256 * i;
            const int rl = u / ROW_U4, part = u % ROW_U4;
            tile[b][rl * STRIDE + part] = pre[i];
        });
    };
    auto mfma_block = [&](f32x4 (&acc)[4], int cur, int cb) __attribute__((always_inline)) {
#pragma unroll
        for (int rb = 0; rb < 4; rb++) acc[rb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const uint4* brow = &tile[cur][(cb * 16 + lc) * STRIDE];
        if constexpr (OP == kOperandBf16) {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const bf16x8 b = __builtin_bit_cast(bf16x8, brow[g * NCH + c]);
#pragma unroll
                for (int rb = 0; rb < 4; rb++)
                    acc[rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a[rb][c], acc[rb], 0, 0, 0);
            }
        } else {
            const float4* bq = reinterpret_cast<const float4*>(brow) + g * (Op::kGroup / 4);
#pragma unroll
            for (int c4 = 0; c4 < Op::kGroup / 4; c4++) {
                const float4 b4 = bq[c4];
                const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int cc = 0; cc < 4; cc++) {
                    if (c4 * 4 + cc < NCH) {
#pragma unroll
                        for (int rb = 0; rb < 4; rb++)
                            acc[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[cc], a[rb][c4 * 4 + cc], acc[rb], 0, 0, 0);
                    }
                }
            }
        }
    };
    // acc[qb][r] holds reference 4*(l/16) + r of the block against query qb*16 + l%16, so a
    // lane's four values share one threshold: one max and one compare per query block.
    auto filter = [&](const f32x4 (&acc)[4], uint32_t t, int cb) __attribute__((always_inline)) {
        uint64_t any = 0;
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            const float m = __builtin_fmaxf(__builtin_fmaxf(acc[qb][0], acc[qb][1]),
                                            __builtin_fmaxf(acc[qb][2], acc[qb][3]));
            any |= __ballot(m > phi[qb]);
        }
        if (any) {
#pragma unroll
            for (int qb = 0; qb < 4; qb++) {
                const int ql = w * kWave + qb * 16 + lc;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (acc[qb][r] > phi[qb]) {
                        const uint32_t id = t * kTileRefs + cb * 16 + g * 4 + r;
                        const uint32_t slot = atomicAdd(&cnt[ql], 1u);
                        p.buf[static_cast<size_t>(qbase + ql) * kCap + slot] = make_key(-2.0f * acc[qb][r], id);
                    }
                }
            }
        }
    };

    load_tile(0);
    store_tile(0);
    __syncthreads();
    for (uint32_t t = 0; t < n_tiles; t++) {
        const int cur = t & 1;
        // Unconditional prefetch (the last tile reloads itself into the idle buffer): a
        // conditional load keeps `pre` out of registers and serialises the prefetch.
        load_tile(t + 1 < n_tiles ? t + 1 : t);
        f32x4 acc0[4], acc1[4];
        // Software pipeline: the MFMAs of block cb are in flight while block cb-1 is filtered.
        mfma_block(acc0, cur, 0);
        mfma_block(acc1, cur, 1);
        filter(acc0, t, 0);
        mfma_block(acc0, cur, 2);
        filter(acc1, t, 1);
        mfma_block(acc1, cur, 3);
        filter(acc0, t, 2);
        filter(acc1, t, 3);
        store_tile(cur ^ 1);
        __syncthreads();
        // Compact every row of this wave whose buffer could overflow in the next tile.
        const uint32_t mine = cnt[w * kWave + lane];
        uint64_t full = __ballot(mine > static_cast<uint32_t>(kCap - kTileRefs));
        if (full) {
            while (full) {
                const int q = __builtin_ctzll(full);
                full &= full - 1;
                const int ql = w * kWave + q;
                uint64_t* rowbuf = p.buf + static_cast<size_t>(qbase + ql) * kCap;
                const uint32_t count = cnt[ql];
                const float theta = select_smallest<E>(rowbuf, count, KP, lane, rowbuf);
                if (lane == 0) {
                    cnt[ql] = KP;
                    phl[ql] = -0.5f * theta;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
#pragma unroll
            for (int qb = 0; qb < 4; qb++) phi[qb] = phl[w * kWave + qb * 16 + lc];
        }
    }
    // Final selection: the KP smallest keys of each row (unordered) and the KP-th score.
    for (int q = 0; q < kWave; q++) {
        const int ql = w * kWave + q;
        const uint32_t row = qbase + ql;
        if (row >= n) break;
        const uint64_t* rowbuf = p.buf + static_cast<size_t>(row) * kCap;
        uint64_t* out = p.cand + static_cast<size_t>(row) * KP;
        const uint32_t count = cnt[ql];
        if (count >= static_cast<uint32_t>(KP)) {
            const float theta = select_smallest<E>(rowbuf, count, KP, lane, out);
            if (lane == 0) p.theta[row] = theta;
        } else {
            if (lane < KP) out[lane] = static_cast<uint32_t>(lane) < count ? rowbuf[lane] : kNoKey;
            if (lane == 0) p.theta[row] = kInf;
        }
    }
}

// ---------------------------------------------------------------------------
// Dispatch over the compiled (operand, NCH, KP) variants.
// ---------------------------------------------------------------------------

template <int OP, int NCH, int KP>
static void run_pass(const PassArgs& a, hipStream_t stream) {
    const unsigned blocks = static_cast<unsigned>((a.n + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL((pass_kernel<OP, NCH, KP>), dim3(blocks), dim3(256), 0, stream, a);
}

template <int OP, int NCH>
static void run_kp(const PassArgs& a, hipStream_t stream) {
    if (a.kp == 32) run_pass<OP, NCH, 32>(a, stream);
    else run_pass<OP, NCH, 64>(a, stream);
}

static const int kBf16Chunks[] = {1, 2, 3, 4, 5, 6, 8, 10, 13};
static const int kF32Chunks[] = {1, 2, 3, 4, 6, 8, 10, 13, 16, 20, 26, 33};

int bf16_chunks(int d) {
    const int need = (3 * d + 3 + 31) / 32;
    for (int s : kBf16Chunks)
        if (s >= need) return s;
    return 0;
}

int fp32_chunks(int d) {
    const int need = (d + 1 + 3) / 4;
    for (int s : kF32Chunks)
        if (s >= need) return s;
    return 0;
}

int fp32_group(int nch) { return (nch + 3) / 4 * 4; }

#define SCRUST_PASS_CASE(OP, N) \
    case N:                     \
        run_kp<OP, N>(a, stream); \
        return true;

bool launch_pass(const PassArgs& a, hipStream_t stream) {
    if (a.operand == kOperandBf16) {
        switch (a.nch) {
            SCRUST_PASS_CASE(kOperandBf16, 1) SCRUST_PASS_CASE(kOperandBf16, 2)
            SCRUST_PASS_CASE(kOperandBf16, 3) SCRUST_PASS_CASE(kOperandBf16, 4)
            SCRUST_PASS_CASE(kOperandBf16, 5) SCRUST_PASS_CASE(kOperandBf16, 6)
            SCRUST_PASS_CASE(kOperandBf16, 8) SCRUST_PASS_CASE(kOperandBf16, 10)
            SCRUST_PASS_CASE(kOperandBf16, 13)
            default: return false;
        }
    }
    switch (a.nch) {
        SCRUST_PASS_CASE(kOperandF32, 1) SCRUST_PASS_CASE(kOperandF32, 2)
        SCRUST_PASS_CASE(kOperandF32, 3) SCRUST_PASS_CASE(kOperandF32, 4)
        SCRUST_PASS_CASE(kOperandF32, 6) SCRUST_PASS_CASE(kOperandF32, 8)
        SCRUST_PASS_CASE(kOperandF32, 10) SCRUST_PASS_CASE(kOperandF32, 13)
        SCRUST_PASS_CASE(kOperandF32, 16) SCRUST_PASS_CASE(kOperandF32, 20)
        SCRUST_PASS_CASE(kOperandF32, 26) SCRUST_PASS_CASE(kOperandF32, 33)
        default: return false;
    }
}

}  // namespace scrust_hip
