// Candidate pass: MFMA tile contraction of the squared-distance expansion fused with
// per-row candidate selection (gfx950, wave64).
//
// A work item is a block of 256 query rows (4 waves x 64) of one dataset. The workgroup
// streams every reference of that dataset through LDS in tiles of 64, double buffered,
// read from HBM with 16-byte coalesced loads. A wave keeps its 64 rows' query fragments
// in registers for the whole scan; the reference tile supplies the other MFMA operand.
// The reference's squared norm is one more packed feature, so the MFMA accumulator is
// acc = a.b - |b|^2 / 2 and the score s = |b|^2 - 2 a.b (the row's |a|^2 is a constant
// shift) is -2 acc. References sit on the MFMA's M side and queries on its N side, so the
// four accumulators a lane holds per 16x16 block belong to one query row and are tested
// with one max and one compare; a block's tests form one wave-uniform branch.
//
// Selection: a pair with s < theta[row] is appended to the row's `cap`-slot buffer in
// global memory. The buffer is four segments, one per lane group of the query row's four
// lanes, each with its own counter in a register: appends need no atomics and no LDS.
// After each tile, a row with a segment that could overflow in the next tile is compacted
// by its wave (ballot radix select of the KP smallest keys, KP/4 written back to each
// segment); theta becomes the KP-th smallest score. theta only decreases, so every
// reference that is not kept has s >= theta_final: the bound the rerank certifies with.
//
// Operands (written by prepare_kernel):
//   bf16: coordinate c = h + l + r, h = bf16(c), l = bf16(c - h); reference rows
//         [h | l | h | nh nl nl2 | 0], query rows [h | h | l | -1/2 -1/2 -1/2 | 0]
//         (|b|^2 = nh + nl + nl2 + O(2^-24)); 32 * nch packed features per row.
//   fp32: feature f of 4 * group slots sits at (f / nch) * group + f % nch; features are
//         [c | |b|^2 | 0] for references; the kernel puts -1/2 at feature d for queries.
//   The reference operands are padded to a whole pair of 64-row tiles with sentinel rows
//   whose norm feature is +inf: they score -inf and are never candidates.
//
// Two entry points share the block body: pass_kernel (one dataset, chunk count fixed at
// compile time) and pass_batch_kernel (a persistent grid that takes work items of many
// datasets from a queue; chunk count per dataset at run time, up to a compiled maximum).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "knn_device.hpp"
#include "knn_launch.hpp"


namespace scrust_hip {

constexpr int kTileRefs = 64;  // references per LDS tile (4 column blocks of 16)

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Row buffers: a row's `cap` slots are four segments of cap/4, segment g appended to by
// lane group g (lanes 16g..16g+15) with a per-lane counter; c[g] keys of segment g are
// valid.

// Loads a row's valid keys, E per lane (E * 64 == cap, key index e = i*64 + lane), and
// kNoKey in the other places.
template <int E>
__device__ __forceinline__ void gather_row(const SCRUST_GLOBAL uint64_t* rowbuf,
                                           const uint32_t (&c)[4], int lane, uint64_t (&key)[E]) {
    constexpr int kSeg = E * 16;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int e = i * kWave + lane;
        const int sg = e / kSeg, o = e % kSeg;
        const uint32_t cs = sg == 0 ? c[0] : sg == 1 ? c[1] : sg == 2 ? c[2] : c[3];
        key[i] = static_cast<uint32_t>(o) < cs ? rowbuf[e] : kNoKey;
    }
}

// Wave-wide min / max on the VALU (DPP): row shifts by 1, 2, 4 and 8 reduce each row of
// 16 lanes into its lane 15, row broadcasts 15 and 31 carry the rows into lane 63. Lanes
// without a source keep their own value.
template <bool MAX>
__device__ __forceinline__ uint32_t wave_extreme_u32(uint32_t v) {
    auto pick = [](uint32_t a, uint32_t b) { return MAX ? (a > b ? a : b) : (a < b ? a : b); };
    auto step = [&](uint32_t x, auto ctrl, auto row_mask) {
        const uint32_t t = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
            static_cast<int>(x), static_cast<int>(x), decltype(ctrl)::value, decltype(row_mask)::value, 0xF, false));
        return pick(x, t);
    };
    v = step(v, std::integral_constant<int, 0x111>{}, std::integral_constant<int, 0xF>{});  // row_shr:1
    v = step(v, std::integral_constant<int, 0x112>{}, std::integral_constant<int, 0xF>{});  // row_shr:2
    v = step(v, std::integral_constant<int, 0x114>{}, std::integral_constant<int, 0xF>{});  // row_shr:4
    v = step(v, std::integral_constant<int, 0x118>{}, std::integral_constant<int, 0xF>{});  // row_shr:8
    v = step(v, std::integral_constant<int, 0x142>{}, std::integral_constant<int, 0xA>{});  // row_bcast:15
    v = step(v, std::integral_constant<int, 0x143>{}, std::integral_constant<int, 0xC>{});  // row_bcast:31
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// In-scan compaction of one row (E * 64 == cap) whose `total` valid keys are in `key`
// (gather_row). With at least `keep` keys it keeps every
// key whose score is at most a cut-off v with keep <= count(score <= v) <= keep + keep/8,
// found by a binary search over the score range the keys span (a handful of steps; the
// exact KP-th score is only needed by the final selection); ties that defeat the window
// fall back to the exact selection. Keys are written back spread evenly over the four
// segments. Returns the number kept and sets theta (the new append bound, or +inf when
// the row holds fewer than `keep` keys and is only spread out).
template <int E>
__device__ uint32_t compact_row(SCRUST_GLOBAL uint64_t* rowbuf, const uint64_t (&key)[E],
                                uint32_t total, int keep, int lane, float& theta) {
    constexpr uint32_t kSeg = E * 16;
    uint32_t v = 0xFFFFFFFFu;
    uint32_t kept = total;
    theta = kInf;
    if (total >= static_cast<uint32_t>(keep)) {
        uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (key[i] != kNoKey) {
                const uint32_t sc = static_cast<uint32_t>(key[i] >> 32);
                lo = sc < lo ? sc : lo;
                hi = sc > hi ? sc : hi;
            }
        }
        lo = wave_extreme_u32<false>(lo);
        hi = wave_extreme_u32<true>(hi);
        const uint32_t most = static_cast<uint32_t>(keep + keep / 8);
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            uint32_t n_le = 0;
#pragma unroll
            for (int i = 0; i < E; i++)
                n_le += static_cast<uint32_t>(__popcll(__ballot(static_cast<uint32_t>(key[i] >> 32) <= mid)));
            if (n_le < static_cast<uint32_t>(keep)) {
                lo = mid + 1;
            } else if (n_le > most) {
                hi = mid;
            } else {
                lo = hi = mid;
            }
        }
        v = lo;
        kept = 0;
#pragma unroll
        for (int i = 0; i < E; i++)
            kept += static_cast<uint32_t>(__popcll(__ballot(static_cast<uint32_t>(key[i] >> 32) <= v)));
        if (kept > most || kept < static_cast<uint32_t>(keep)) {
            // Ties on the cut-off score (or a window the range bounds missed): exactly
            // `keep` keys by (score, id).
            theta = select_keys<E>(key, keep, lane, rowbuf, static_cast<uint32_t>(keep) / 4, kSeg);
            return static_cast<uint32_t>(keep);
        }
        theta = unordered(v);
    }
    // Keys with score <= v (all valid keys when spreading), quarter per segment.
    const uint32_t quarter = (kept + 3) / 4;
    uint64_t m[E];
#pragma unroll
    for (int i = 0; i < E; i++)
        m[i] = __ballot(key[i] != kNoKey && static_cast<uint32_t>(key[i] >> 32) <= v);
    uint32_t base = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
        if (key[i] != kNoKey && static_cast<uint32_t>(key[i] >> 32) <= v)
            rowbuf[kept_slot(base + lanes_below(m[i]), quarter, kSeg)] = key[i];
        base += static_cast<uint32_t>(__popcll(m[i]));
    }
    return kept;
}

template <int OP, int NCH>
struct Operand;

// LDS tile layout. A ds_read_b128 is serviced for lanes 0-31 and then lanes 32-63, each
// half over 128 four-byte banks, so a read is conflict-free when the 32 lanes of a half
// address 32 distinct 16-byte slots modulo 32. Lanes 16g..16g+15 read rows lc = 0..15 at
// lane group g's part, slot lc * stride + g * gstride + c: with stride = 2 (mod 4) the
// rows cover the 16 even residues and an odd gstride moves group g + 1 onto the odd ones.
__host__ __device__ constexpr int lds_gstride(int group_u4) { return group_u4 | 1; }
__host__ __device__ constexpr int lds_stride(int group_u4) {
    return (3 * lds_gstride(group_u4) + group_u4 + 1) / 4 * 4 + 2;
}

template <int NCH>
struct Operand<kOperandBf16, NCH> {
    typedef bf16x8 Frag;
    // 16-byte chunks per packed row, per lane group.
    static __device__ __forceinline__ int row_u4(int nch) { return 4 * nch; }
    static __device__ __forceinline__ int group_u4(int nch) { return nch; }
    static constexpr int kRowU4Max = 4 * NCH;
    static constexpr int kGroupU4Max = NCH;
    static constexpr int kLoadsMax = NCH;  // 64 rows x 4 nch chunks / 256 threads
};

// Local-origin fp16 operands: the same LDS layout as bf16 (4 nch 16-byte slots per row,
// lane group g's chunk c at g * gstride + c, holding features 8 (g nch + c) .. + 7), filled
// by converting the padded fp32 coordinates inside the pass.
template <int NCH>
struct Operand<kOperandF16Local, NCH> {
    typedef f16x8 Frag;
    static __device__ __forceinline__ int row_u4(int nch) { return 4 * nch; }
    static __device__ __forceinline__ int group_u4(int nch) { return nch; }
    static constexpr int kRowU4Max = 4 * NCH;
    static constexpr int kGroupU4Max = NCH;
    static constexpr int kLoadsMax = 1;
};

template <int NCH>
struct Operand<kOperandF32, NCH> {
    typedef float Frag;
    static __device__ __forceinline__ int row_u4(int nch) { return (nch + 3) / 4 * 4; }
    static __device__ __forceinline__ int group_u4(int nch) { return (nch + 3) / 4; }
    static constexpr int kRowU4Max = (NCH + 3) / 4 * 4;
    static constexpr int kGroupU4Max = (NCH + 3) / 4;
    static constexpr int kLoadsMax = kRowU4Max / 4;
};

// One 16x16x32 product of the 16-bit operands into acc.
__device__ __forceinline__ f32x4 mfma_k32(bf16x8 b, bf16x8 a, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, acc, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma_k32(f16x8 b, f16x8 a, f32x4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a, acc, 0, 0, 0);
}

template <int OP, int NCH>
struct PassSmem {
    u32x4 tile[2][kTileRefs * lds_stride(Operand<OP, NCH>::kGroupU4Max)];
    float ball[16 + kBallMaxDims];  // the ball bound's LDS (kBallLds floats)
};

// Three tile buffers for LDS-direct loads, each its own __shared__ array: the compiler tells
// loads into one buffer apart from reads of another only across distinct objects. Split
// tiles (below) have three more for the fine parts and two words of per-wave flags.
struct DmaTiles {
    u32x4* b0;
    u32x4* b1;
    u32x4* b2;
    // Persistent pass: the ball bound's LDS (kBallLds floats: the waves' thetas in two
    // barrier-parity halves, the block's radius, its centre), and each wave's staging of a
    // refined block's fine chunks (kFineStage 16-byte slots a wave).
    float* ball;
    u32x4* fine;
};

// Slots (16 bytes) of one LDS tile buffer.
template <int OP, int NCH>
constexpr int kTileSlots = kTileRefs * lds_stride(Operand<OP, NCH>::kGroupU4Max);

// The fixed-chunk bf16 pass loads tiles straight into LDS when three tile buffers fit in half
// of a CU's 160 KB (two workgroups per CU): up to 5 chunks. Measured at 1M x 50 on one GPU,
// A/B: the bf16 pass 1.5% faster; the fp32 pass 1.3% slower (its loads then have one tile of
// lead instead of two), so it keeps the register staging.
template <int OP, int NCH>
constexpr bool kPassDma = OP == kOperandBf16 && 3 * kTileSlots<OP, NCH> * 16 <= 80 * 1024;

// Split tiles of the two-stage bf16 pass (the persistent pass): the scan streams only a
// tile's coarse part (the first kc MFMA steps of every row, all a block needs unless its
// coarse scores come near a row's append bound) through LDS, 4 kc chunks per reference row; a
// block that needs more reads its 4 kf fine chunks per row from global memory. The coarse
// buffer has no padding: lane group g's step c of row r is at slot
// r 4 kc + ((g kc + c) ^ coarse_swizzle(r)). gfx950 serves a ds_read_b128 in four passes of
// 16 lanes, lanes 8j..8j+7 with 8j+16..8j+23, each over 64 four-byte banks (16 slots of 16
// bytes): a read is free of bank conflicts when each pass hits 16 distinct slots modulo 16.
// (This reading fits every layout measured: the padded whole tiles, 0 conflict cycles; a
// one-slot pad, 2-way; a swizzle by row bits 2 and 3 alone, 4 cycles per read.) The swizzle
// below takes row bits 1-3 so that it also keeps the 32 lanes of a half-wave on 32 distinct
// slots modulo 32.
template <int KC>
__host__ __device__ constexpr int coarse_swizzle(int r) {
    return KC == 2 ? (((r >> 2) & 1) | ((((r >> 1) ^ (r >> 3)) & 1) << 2)) : ((((r >> 2) ^ (r >> 3)) & 1) << 1);
}
template <int NCH>
constexpr int kCoarseSlots = kTileRefs * 4 * bf16_coarse_chunks(NCH);
// The persistent pass's ball-bound LDS (floats): the waves' thetas in two halves of 4 (by
// barrier parity) at 0, the block's radius at 8, its centre (at most kBallMaxD floats) at 16.
constexpr int kBallMaxD = static_cast<int>(kBallMaxDims);
constexpr int kBallLds = 16 + kBallMaxD;
static_assert(kBallLds <= 16 + static_cast<int>(kBallMaxDims), "the ball LDS fits PassSmem::ball");
constexpr int kFineStage = 3 * 64;  // fine chunks of one 16-reference block (NCH - coarse <= 3)
// Three ring slots of two tiles' coarse parts in half of a CU's LDS (two workgroups per CU).
template <int OP, int NCH>
constexpr bool kPassSplit = OP == kOperandBf16 && bf16_coarse_chunks(NCH) < NCH && bf16_coarse_chunks(NCH) <= 2 &&
                            3 * 2 * kCoarseSlots<NCH> * 16 + 8 <= 80 * 1024;
// The host's split_pass_nch names the chunk counts the split variants exist for.
static_assert(!kPassSplit<kOperandBf16, 1> && kPassSplit<kOperandBf16, 2> && kPassSplit<kOperandBf16, 3> &&
                  kPassSplit<kOperandBf16, 4> && kPassSplit<kOperandBf16, 5> && !kPassSplit<kOperandBf16, 6>,
              "split_pass_nch must match kPassSplit");

// Persistent pass: the end of a work item. Each wave selects, row after row, the kp smallest
// keys (unordered) of its 64 rows' buffers (row q's four segment counts in c(q)) exactly as
// select_kernel does: a whole block's rows into cand with their kp-th score in theta, a range
// item's rows (splits > 1, rows from tail_row0 on) into their range's place in tail_cand.
// A row with fewer than kp keys gets all of them, then kNoKey, and theta +inf. `empty`: no
// keys (a dataset that skipped the scan).
template <int E, class Counts>
__device__ __forceinline__ void finish_rows(const DatasetDesc& s, const PassItem& item,
                                            const SCRUST_GLOBAL uint64_t* rbuf, uint32_t row0,
                                            int lane, bool empty, Counts counts) {
    const uint32_t kp = s.kp, cap = s.cap;
    const bool tail = s.splits > 1 && row0 >= s.tail_row0;
    for (int q = 0; q < kWave; q++) {
        const uint32_t row = row0 + static_cast<uint32_t>(q);
        if (row >= s.n) break;
        uint32_t c[4] = {0u, 0u, 0u, 0u};
        if (!empty) counts(q, c);
        uint64_t key[E];
        gather_row<E>(rbuf + static_cast<size_t>(q) * cap, c, lane, key);
        SCRUST_GLOBAL uint64_t* out =
            tail ? global_ptr(s.tail_cand) + ((static_cast<size_t>(row - s.tail_row0) * s.splits + item.split) * kp)
                 : global_ptr(s.cand) + static_cast<size_t>(row) * kp;
        uint64_t m[E];
        uint32_t total = 0;
#pragma unroll
        for (int i = 0; i < E; i++) {
            m[i] = __ballot(key[i] != kNoKey);
            total += static_cast<uint32_t>(__popcll(m[i]));
        }
        float theta = kInf;
        if (total >= kp) {
            theta = select_keys<E>(key, static_cast<int>(kp), lane, out, 0u, 0u);
        } else {
            uint32_t base = 0;
#pragma unroll
            for (int i = 0; i < E; i++) {
                if (key[i] != kNoKey) out[base + lanes_below(m[i])] = key[i];
                base += static_cast<uint32_t>(__popcll(m[i]));
            }
            for (uint32_t slot = base + static_cast<uint32_t>(lane); slot < kp; slot += kWave) out[slot] = kNoKey;
        }
        if (!tail && lane == 0) global_ptr(s.theta)[row] = theta;
    }
}

// One work item: query rows [block * 256, block * 256 + 256) of dataset s against the
// references of tiles [tile_begin, tile_end), into buffer slot `split` (persistent pass:
// into the workgroup's own buffer region). FIXED: the dataset's chunk count equals NCH at
// compile time.
template <int OP, int NCH, bool FIXED, bool DMA = false, bool PAIRS = false, bool BIG = false>
__device__ __forceinline__ void pass_block(const DatasetDesc& s, const PassItem& item,
                                           PassSmem<OP, NCH>* sm, DmaTiles dma = {}) {
    typedef Operand<OP, NCH> Op;
    // Persistent pass: the workgroup's buffer region is its slot's (workgroup index).
    const bool persist = PAIRS && s.persist != 0u;
    // A dataset with a squared norm past kPassNormLimit is served by the exact kernel (the
    // rerank lists all its rows); its segment counts stay at the zeros they were cleared to
    // (persistent pass: its rows get empty candidate lists here).
    if (__uint_as_float(*global_ptr(s.max_norm_bits)) > kPassNormLimit) {
        if constexpr (PAIRS) {
            if (persist) {
                const int w0 = static_cast<int>(threadIdx.x) >> 6, l0 = static_cast<int>(threadIdx.x) & 63;
                finish_rows<2>(s, item, global_ptr(s.buf), item.block * kPassRows + w0 * kWave, l0, true,
                               [](int, uint32_t (&)[4]) {});
            }
        }
        return;
    }
    const int nch = FIXED ? NCH : static_cast<int>(s.nch);
    const int row_u4 = FIXED ? Op::kRowU4Max : Op::row_u4(nch);
    const int group_u4 = FIXED ? Op::kGroupU4Max : Op::group_u4(nch);
    const int gstride = lds_gstride(group_u4);
    const int stride = lds_stride(group_u4);
    const int loads = FIXED ? Op::kLoadsMax : row_u4 / 4;
    const int group = row_u4;  // fp32: floats per lane group (4 groups per row)
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, lc = lane & 15;
    const uint32_t n = s.n;
    const uint32_t cap = s.cap;
    const int kp = static_cast<int>(s.kp);
    const uint32_t qbase = item.block * kPassRows;
    const SCRUST_GLOBAL u32x4* ref_rows = global_ptr(reinterpret_cast<const u32x4*>(
        OP == kOperandBf16 ? static_cast<const void*>(s.pack_ref) : static_cast<const void*>(s.pack_f32)));
    // The item's buffer rows: this range's slot of the row buffers (indexed by the dataset's
    // row), or the workgroup's own region (persistent pass).
    SCRUST_GLOBAL uint64_t* const rbuf =
        global_ptr(s.buf) + (persist ? static_cast<size_t>(blockIdx.x) * kPassRows : range_row_base(s, item.split) + qbase) * s.cap;

    // Query fragments: the B operand of every MFMA, held for the whole scan. Query block
    // qb's fragment in lane l carries row qb*16 + l%16. The persistent pass has no query
    // operand array: the workgroup first writes its 256 rows' operand rows (as
    // prepare_rows_kernel writes them for the other paths, bf16_query_chunk) into the start
    // of its own buffer region, which the item's appends only reuse after the barrier that
    // follows the loads.
    const SCRUST_GLOBAL u32x4* qsrc = global_ptr(reinterpret_cast<const u32x4*>(s.pack_query));
    if constexpr (OP == kOperandBf16) {
        if (persist) {
            SCRUST_GLOBAL u32x4* stage = reinterpret_cast<SCRUST_GLOBAL u32x4*>(rbuf);
            const SCRUST_GLOBAL float* coords = global_ptr(s.coords);
#pragma unroll 1
            for (int u = tid; u < kPassRows * 4 * NCH; u += 256) {
                const int r = u / (4 * NCH), m = u - r * (4 * NCH);
                uint32_t row = qbase + static_cast<uint32_t>(r);
                if (row >= n) row = n - 1;
                stage[u] = bf16_query_chunk(coords + static_cast<size_t>(row) * s.d, static_cast<int>(s.d), m, NCH);
            }
            __syncthreads();
            qsrc = stage;
        }
    }
    typename Op::Frag a[4][NCH];
#pragma unroll
    for (int qb = 0; qb < 4; qb++) {
        uint32_t row = qbase + w * kWave + qb * 16 + lc;
        if (row >= n) row = n - 1;
        if constexpr (OP == kOperandBf16 || OP == kOperandF16Local) {
            typedef typename Op::Frag Frag;
            const size_t at = persist ? static_cast<size_t>(w * kWave + qb * 16 + lc) : static_cast<size_t>(row);
            const SCRUST_GLOBAL u32x4* src = qsrc + at * row_u4 + g * nch;
#pragma unroll
            for (int c = 0; c < NCH; c++)
                a[qb][c] = (FIXED || c < nch) ? __builtin_bit_cast(Frag, src[c]) : Frag{};
        } else {
            const SCRUST_GLOBAL float* src = global_ptr(s.pack_f32) + static_cast<size_t>(row) * 4 * group + g * group;
#pragma unroll
            for (int c = 0; c < NCH; c++)
                a[qb][c] = (FIXED || c < nch) ? ((g * nch + c == static_cast<int>(s.d)) ? -0.5f : src[c]) : 0.0f;
        }
    }
    if (persist) __syncthreads();
    // phi[qb] = -theta/2 of this lane's query row qb*16 + l%16; cnt[qb] = keys in this
    // lane's segment of that row's buffer.
    float phi[4];
    uint32_t cnt[4];
#pragma unroll
    for (int qb = 0; qb < 4; qb++) {
        phi[qb] = -kInf;
        cnt[qb] = 0;
    }
    const uint32_t seg = cap / 4;
    // A tile appends at most 16 keys to a segment (4 column blocks x 4 references).
    // Pair steps (split tiles with PAIRS) scan two tiles between compactions: 32 keys.
    // BIG (1024-slot buffers, KP = 256): a 256-slot segment could fill, and its count must
    // stay below 256 for the byte-wide segment counts, so it compacts one key earlier.
    const uint32_t trigger = seg - (PAIRS ? 32u : 16u) - (BIG ? 1u : 0u);
    SCRUST_GLOBAL uint64_t* const lane_seg =
        rbuf + static_cast<size_t>(w * kWave + lc) * cap + g * seg;

    // The reference operand holds whole pairs of tiles (rows n.. are sentinels), so a
    // tile is one contiguous run of tile_u4 16-byte chunks; ranges hold whole pairs.
    const uint32_t t_begin = item.tile_begin, t_end = item.tile_end;
    // With a rotation (item.pad > 0, spatial order) the scan visits tile (t + pad) mod the
    // dataset's tile count for t in [t_begin, t_end): a block's ranges cut the rotated
    // sequence, which starts at the block's own part of the order. item.pad and the tile
    // count are even, so pairs stay pairs. Without one, tile t itself.
    const uint32_t rot = item.pad;
    const uint32_t wrap = rot ? 2u * ((n + 2 * kTileRefs - 1) / (2 * kTileRefs)) : t_end;
    auto phys = [&](uint32_t t) __attribute__((always_inline)) {
        const uint32_t u = t + rot;
        return u >= wrap ? u - wrap : u;
    };
    const uint32_t tile_u4 = kTileRefs * row_u4;
    // LDS tile buffer `cur` (a compile-time constant at every use).
    auto tile_of = [&](int cur) __attribute__((always_inline)) -> u32x4* {
        if constexpr (DMA) return cur == 0 ? dma.b0 : cur == 1 ? dma.b1 : dma.b2;
        else return sm->tile[cur];
    };
    // Two register sets stage tiles t+1 and t+2 while tile t is consumed from LDS, so a
    // tile's loads are issued two tiles before they are stored (HBM/L3 latency exceeds
    // one tile's MFMA time).
    u32x4 pa[Op::kLoadsMax], pb[Op::kLoadsMax];
    // Thread chunk i of a tile is u = tid + 256 i: reference row u / row_u4, part
    // p = u % row_u4 of lane group p / group_u4, at LDS slot
    // row * stride + (p / group_u4) * gstride + p % group_u4.
    // bf16 tiles hold their rows' coarse chunks first, then the fine ones (tile_chunk).
    int lds_at[Op::kLoadsMax];
#pragma unroll
    for (int i = 0; i < Op::kLoadsMax; i++) {
        const int u = tid + 256 * i;
        if constexpr (OP == kOperandBf16) {
            int r = 0, gg = 0, c = 0;
            if (FIXED || i < loads) tile_chunk_of(u, nch, r, gg, c);
            lds_at[i] = r * stride + gg * gstride + c;
        } else {
            const int p = u % row_u4;
            lds_at[i] = (u / row_u4) * stride + (p / group_u4) * gstride + p % group_u4;
        }
    }
    // Global -> registers P for tile T; registers P -> LDS buffer B. Plain unrolled loops
    // with constant indices keep the sets in registers.
#define SCRUST_LOAD_TILE(P, T)                                                    \
    _Pragma("unroll") for (int i = 0; i < Op::kLoadsMax; i++) {                   \
        if (FIXED || i < loads)                                                   \
            P[i] = ref_rows[static_cast<size_t>(phys(T)) * tile_u4 + tid + 256 * i];   \
    }
#define SCRUST_STORE_TILE(P, B)                                                   \
    _Pragma("unroll") for (int i = 0; i < Op::kLoadsMax; i++) {                   \
        if (FIXED || i < loads) tile_of(B)[lds_at[i]] = P[i];                     \
    }
    auto mfma_block = [&](f32x4 (&acc)[4], int cur, int cb) __attribute__((always_inline)) {
#pragma unroll
        for (int qb = 0; qb < 4; qb++) acc[qb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        const u32x4* brow = &tile_of(cur)[(cb * 16 + lc) * stride];
        if constexpr (OP == kOperandBf16 || OP == kOperandF16Local) {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                if (FIXED || c < nch) {
                    const typename Op::Frag b = __builtin_bit_cast(typename Op::Frag, brow[g * gstride + c]);
#pragma unroll
                    for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(b, a[qb][c], acc[qb]);
                }
            }
        } else {
            const float4* bq = reinterpret_cast<const float4*>(brow) + g * gstride;
#pragma unroll
            for (int c4 = 0; c4 < Op::kRowU4Max / 4; c4++) {
                if (FIXED || c4 * 4 < nch) {
                    const float4 b4 = bq[c4];
                    const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                    for (int cc = 0; cc < 4; cc++) {
                        const int c = c4 * 4 + cc;
                        if (c < NCH && (FIXED || c < nch)) {
#pragma unroll
                            for (int qb = 0; qb < 4; qb++)
                                acc[qb] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[cc], a[qb][c], acc[qb], 0, 0, 0);
                        }
                    }
                }
            }
        }
    };
    // acc[qb][r] holds reference 4*(l/16) + r of the block against query qb*16 + l%16, so a
    // lane's four values share one threshold: one max and one compare per query block.
    auto filter = [&](const f32x4 (&acc)[4], uint32_t t, int cb) __attribute__((always_inline)) {
        uint64_t hit[4];
        uint64_t any = 0;
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            const float m = __builtin_fmaxf(__builtin_fmaxf(acc[qb][0], acc[qb][1]),
                                            __builtin_fmaxf(acc[qb][2], acc[qb][3]));
            hit[qb] = __ballot(m > phi[qb]);
            any |= hit[qb];
        }
        if (any) {
            const uint32_t id0 = t * kTileRefs + cb * 16 + g * 4;
#pragma unroll
            for (int qb = 0; qb < 4; qb++) {
                if (hit[qb]) {
                    SCRUST_GLOBAL uint64_t* const dst = lane_seg + static_cast<size_t>(qb * 16) * cap;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        if (acc[qb][r] > phi[qb]) {
                            dst[cnt[qb]] = make_key(-2.0f * acc[qb][r], id0 + r);
                            cnt[qb]++;
                        }
                    }
                }
            }
        }
    };
    // Segment counts of wave-local row q = qb*16 + lr, read from its four lanes.
    auto row_counts = [&](int q, uint32_t (&c)[4]) __attribute__((always_inline)) {
        const int qb = q >> 4, lr = q & 15;
        static_for<0, 4>([&](auto QB) {
            if (qb == QB) {
#pragma unroll
                for (int gg = 0; gg < 4; gg++)
                    c[gg] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cnt[QB]), gg * 16 + lr));
            }
        });
    };

    // Tile t: its MFMAs and filter on LDS buffer cur, then the next tile's registers to the
    // other buffer, a barrier, and the compaction of rows whose buffers could overflow.
    // bf16: a block's reference fragments, read from LDS one block ahead of their MFMAs.
    auto load_frags = [&](auto& b, int cur, int cb) __attribute__((always_inline)) {
        const u32x4* brow = &tile_of(cur)[(cb * 16 + lc) * stride];
#pragma unroll
        for (int c = 0; c < NCH; c++)
            if (FIXED || c < nch) b[c] = __builtin_bit_cast(std::remove_reference_t<decltype(b[0])>, brow[g * gstride + c]);
    };
    auto mfma_frags = [&](f32x4 (&acc)[4], const auto& b) __attribute__((always_inline)) {
#pragma unroll
        for (int qb = 0; qb < 4; qb++) acc[qb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            if (FIXED || c < nch) {
#pragma unroll
                for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(b[c], a[qb][c], acc[qb]);
            }
        }
    };
    // bf16, two stages (single datasets, NCH > kCoarse): the first kCoarse MFMA steps of a
    // block give every h.h and norm term (the packing puts them there); the other steps hold
    // only hi/lo cross terms, whose sum over any subset of features is at most
    // 2^-7 (1 + 2^-8)^2 |a||b| (|l| <= 2^-8 (1 + 2^-8) |c|, |h| <= (1 + 2^-8) |c|, Cauchy-
    // Schwarz). A block whose coarse accumulators all lie at least that bound (plus the
    // rounding of the remaining accumulation steps) below the row's append bound phi cannot
    // append a pair with its full accumulators either: its remaining steps and its filter are
    // skipped, and the buffers end exactly as with every step taken.
    constexpr int kCoarse = bf16_coarse_chunks(NCH);
    constexpr bool kTwoStage = OP == kOperandBf16 && FIXED && kCoarse < NCH;
    // Split tiles in pairs: the persistent pass (kPassSplit, PAIRS).
    constexpr bool kSplit = kTwoStage && DMA && kPassSplit<OP, NCH> && PAIRS;
    // gate[qb] = phi[qb] - skip_b[qb], the coarse bound of the block test; it changes only
    // when phi does (compaction).
    float skip_b[4], gate[4];
    if constexpr (kTwoStage) {
        const float mx = __uint_as_float(*global_ptr(s.max_norm_bits));
        const float rb = __builtin_sqrtf(mx * (1.0f + 0x1p-16f));
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            uint32_t row = qbase + w * kWave + qb * 16 + lc;
            if (row >= n) row = n - 1;
            const float qn = __builtin_sqrtf(global_ptr(s.norms)[row] * (1.0f + 0x1p-16f));
            skip_b[qb] = 0x1p-7f * 1.02f * qn * rb + 0x1p-13f * mx;
            gate[qb] = phi[qb] - skip_b[qb];
        }
    } else if constexpr (OP == kOperandF32 && FIXED) {
        // fp32 operands (the ball bound of the wide fp32 pass): the computed score lies within
        // acc_rel (|q||r| + |r|^2 / 2) of q.r - |r|^2 / 2, plus half the f32 rounding of the
        // |r|^2 feature ((d + 1) u |r|^2); skip_b_q bounds both, rounded up.
        const float mx = __uint_as_float(*global_ptr(s.max_norm_bits));
        const float rb = __builtin_sqrtf(mx * (1.0f + 0x1p-15f));
        const float ar = static_cast<float>(s.acc_rel) * 1.05f;
        const float nr = static_cast<float>(s.d + 2) * 0x1p-24f;
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            uint32_t row = qbase + w * kWave + qb * 16 + lc;
            if (row >= n) row = n - 1;
            const float qn = __builtin_sqrtf(global_ptr(s.norms)[row] * (1.0f + 0x1p-15f));
            skip_b[qb] = ar * (1.02f * qn * rb + 0.51f * mx) + nr * mx;
            gate[qb] = 0.0f;
        }
    }
    auto consume = [&](uint32_t t, int cur) __attribute__((always_inline)) {
        f32x4 acc0[4], acc1[4];
        if constexpr (kTwoStage) {
            typedef typename Op::Frag Frag;
            Frag c0[kCoarse], c1[kCoarse], f[NCH - kCoarse];
            auto load_part = [&](auto& b, int cb, int first, int count) __attribute__((always_inline)) {
                const u32x4* brow = &tile_of(cur)[(cb * 16 + lc) * stride + g * gstride + first];
#pragma unroll
                for (int c = 0; c < count; c++) b[c] = __builtin_bit_cast(Frag, brow[c]);
            };
            auto coarse = [&](f32x4 (&acc)[4], const Frag (&b)[kCoarse]) __attribute__((always_inline)) {
#pragma unroll
                for (int qb = 0; qb < 4; qb++) acc[qb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int c = 0; c < kCoarse; c++)
#pragma unroll
                    for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(b[c], a[qb][c], acc[qb]);
            };
            // The block's remaining steps and its filter, when some pair may pass.
            auto refine = [&](f32x4 (&acc)[4], int cb) __attribute__((always_inline)) {
                uint64_t need = 0;
#pragma unroll
                for (int qb = 0; qb < 4; qb++) {
                    const float m = __builtin_fmaxf(__builtin_fmaxf(acc[qb][0], acc[qb][1]),
                                                    __builtin_fmaxf(acc[qb][2], acc[qb][3]));
                    need |= __ballot(m > gate[qb]);
                }
                if (need) {
                    load_part(f, cb, kCoarse, NCH - kCoarse);
#pragma unroll
                    for (int c = kCoarse; c < NCH; c++)
#pragma unroll
                        for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(f[c - kCoarse], a[qb][c], acc[qb]);
                    filter(acc, t, cb);
                }
            };
            load_part(c0, 0, 0, kCoarse);
            load_part(c1, 1, 0, kCoarse);
            __builtin_amdgcn_sched_barrier(0);
            coarse(acc0, c0);
            __builtin_amdgcn_sched_barrier(0);
            load_part(c0, 2, 0, kCoarse);
            __builtin_amdgcn_sched_barrier(0);
            coarse(acc1, c1);
            __builtin_amdgcn_sched_barrier(0);
            load_part(c1, 3, 0, kCoarse);
            __builtin_amdgcn_sched_barrier(0);
            refine(acc0, 0);
            coarse(acc0, c0);
            refine(acc1, 1);
            coarse(acc1, c1);
            refine(acc0, 2);
            refine(acc1, 3);
        } else if constexpr ((OP == kOperandBf16 || OP == kOperandF16Local) && FIXED) {
            // Software pipeline: block cb+1's fragments are read while block cb's MFMAs
            // run, and block cb-1 is filtered while they run.
            // Scheduling barriers keep the compiler from sinking each read to its use.
            typename Op::Frag b0[NCH], b1[NCH];
            load_frags(b0, cur, 0);
            load_frags(b1, cur, 1);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags(acc0, b0);
            __builtin_amdgcn_sched_barrier(0);
            load_frags(b0, cur, 2);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags(acc1, b1);
            __builtin_amdgcn_sched_barrier(0);
            load_frags(b1, cur, 3);
            __builtin_amdgcn_sched_barrier(0);
            filter(acc0, t, 0);
            mfma_frags(acc0, b0);
            filter(acc1, t, 1);
            mfma_frags(acc1, b1);
            filter(acc0, t, 2);
            filter(acc1, t, 3);
        } else {
            // Software pipeline: the MFMAs of block cb are in flight while block cb-1 is filtered.
            mfma_block(acc0, cur, 0);
            mfma_block(acc1, cur, 1);
            filter(acc0, t, 0);
            mfma_block(acc0, cur, 2);
            filter(acc1, t, 1);
            mfma_block(acc1, cur, 3);
            filter(acc0, t, 2);
            filter(acc1, t, 3);
        }
    };
    // Compacts every row of this wave with a segment that could overflow in the next tile
    // (compact_row): the append bound tightens to the kept keys' cut-off score.
    // Rows are compacted one after another by the whole wave; the next row's keys are
    // gathered before the current row is processed, so their load latency overlaps it.
    auto compact_rows = [&](auto ec, uint64_t full) __attribute__((always_inline)) {
        constexpr int E = decltype(ec)::value;
        int q = __builtin_ctzll(full);
        full &= full - 1;
        uint32_t c[4];
        row_counts(q, c);
        uint64_t key[E];
        gather_row<E>(rbuf + static_cast<size_t>(w * kWave + q) * cap, c, lane, key);
        for (;;) {
            const bool more = full != 0;
            int qn = 0;
            uint32_t cn[4] = {0u, 0u, 0u, 0u};
            uint64_t keyn[E];
            if (more) {
                qn = __builtin_ctzll(full);
                full &= full - 1;
                row_counts(qn, cn);
                gather_row<E>(rbuf + static_cast<size_t>(w * kWave + qn) * cap, cn, lane, keyn);
            }
            float theta;
            const uint32_t kept = compact_row<E>(rbuf + static_cast<size_t>(w * kWave + q) * cap,
                                                 key, c[0] + c[1] + c[2] + c[3], kp, lane, theta);
            // Segment g now holds min(share, kept - g * share) keys.
            const uint32_t share = (kept + 3) / 4;
            const uint32_t start = static_cast<uint32_t>(g) * share;
            const uint32_t mine = kept > start ? (kept - start < share ? kept - start : share) : 0u;
            const int qb = q >> 4, lr = q & 15;
            static_for<0, 4>([&](auto QB) {
                if (qb == QB && lc == lr) {
                    cnt[QB] = mine;
                    if (theta < kInf) {
                        phi[QB] = -0.5f * theta;
                        if constexpr (kTwoStage) gate[QB] = phi[QB] - skip_b[QB];
                    }
                }
            });
            if (!more) break;
            q = qn;
#pragma unroll
            for (int i = 0; i < 4; i++) c[i] = cn[i];
#pragma unroll
            for (int i = 0; i < E; i++) key[i] = keyn[i];
        }
    };
    // Compacts every row of this wave with a segment that could overflow in the next tile
    // (compact_row): the append bound tightens to the kept keys' cut-off score.
    // Returns whether any row was compacted.
    auto compact = [&]() __attribute__((always_inline)) -> bool {
        uint64_t full = 0;
        static_for<0, 4>([&](auto QB) {
            const uint64_t f = __ballot(cnt[QB] > trigger);
            full |= ((f | f >> 16 | f >> 32 | f >> 48) & 0xFFFFull) << (16 * QB);
        });
        if (!full) return false;
        if constexpr (BIG) {
            compact_rows(std::integral_constant<int, 16>{}, full);
        } else {
            if (cap <= 128) compact_rows(std::integral_constant<int, 2>{}, full);
            else if (cap <= 256) compact_rows(std::integral_constant<int, 4>{}, full);
            else compact_rows(std::integral_constant<int, 8>{}, full);
        }
        return true;
    };

    // The persistent pass's split loop and the register-staged loop of wide bf16 rows and of
    // the fp32 pass (whose plans keep balls only past the packed operands' 128 features).
    constexpr bool kBall = (kTwoStage && ((DMA && PAIRS) || !DMA)) || (OP == kOperandF32 && FIXED && !DMA);
    // Ball bound (s.balls; the persistent pass): the steps of tile pair
    // P are skipped when no row q of the block can append a reference r of P. A row appends
    // only pairs whose computed score exceeds its bound phi_q, and the computed score lies
    // within skip_b_q of q.r - |r|^2 / 2 (the deviation the block test allows for the fine
    // steps alone), so a pair with |q - r|^2 > theta_q = |q|^2 - 2 phi_q + 4 skip_b_q cannot
    // append; and |q - r| >= |c_B - c_P| - rho_B - rho_P for the block's and the pair's
    // balls. A skipped step would have appended nothing, and its compaction test found
    // nothing, so the buffers end as with every step scanned. theta_q only falls during the
    // scan: a test made with an earlier block maximum stays valid. Each wave keeps its rows'
    // largest theta_q (wtheta, refreshed after its compactions) and writes it before each
    // barrier into the half of the ball LDS of that barrier's parity; after the barrier every
    // wave reads the same four values (theta_max), and so computes the same survivors and
    // scans the same steps.
    float* const blds = DMA ? dma.ball : (sm ? sm->ball : nullptr);
    const bool prune = kBall && blds != nullptr && s.balls != nullptr;
    const SCRUST_GLOBAL float* const balls = global_ptr(s.balls);
    const uint32_t bstride = s.ball_stride;
    const uint32_t ball_pairs = (n + 2 * kTileRefs - 1) / (2 * kTileRefs);
    const uint32_t ball_count = ball_pairs + s.npad / kPassRows;
    if (prune) {
        // The block's ball into LDS; the first read of it follows the first barrier.
        const SCRUST_GLOBAL float* cb = balls + static_cast<size_t>(ball_pairs + item.block) * bstride;
        if (tid < static_cast<int>(bstride)) blds[16 + tid] = cb[tid];
        if (tid == 0) blds[8] = balls[static_cast<size_t>(ball_count) * bstride + ball_pairs + item.block];
    }
    float wtheta = kInf, theta_max = kInf;
    uint32_t par = 0;
    auto refresh_theta = [&]() __attribute__((always_inline)) {
        float t = 0.0f;
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            uint32_t row = qbase + w * kWave + qb * 16 + lc;
            if (row >= n) row = n - 1;
            const float nq = global_ptr(s.norms)[row];
            const float th = nq - 2.0f * phi[qb] + 4.0f * skip_b[qb];
            // 2^-12 of the terms' magnitudes covers the rounding of th many times over.
            t = __builtin_fmaxf(t, th + 0x1p-12f * (nq + 2.0f * __builtin_fabsf(phi[qb]) + 4.0f * skip_b[qb]));
        }
#pragma unroll
        for (int m = 1; m < kWave; m <<= 1) t = __builtin_fmaxf(t, __shfl_xor(t, m, kWave));
        wtheta = t;
    };
    // Before a barrier: this wave's theta into the half of that barrier's parity. After
    // it: the four waves' values from that half.
    auto ball_publish = [&]() __attribute__((always_inline)) {
        if (prune && lane == 0) blds[par * 4 + w] = wtheta;
    };
    auto ball_absorb = [&]() __attribute__((always_inline)) {
        if (prune) {
            const float* h = blds + par * 4;
            theta_max = __builtin_fmaxf(__builtin_fmaxf(h[0], h[1]), __builtin_fmaxf(h[2], h[3]));
        }
        par ^= 1u;
    };
    uint32_t scanned = 0;  // tile-pair steps this item scans
    // Survivors among scan pairs [wb, wb + 64) (lane l tests pair wb + l): every pair of
    // the range while theta_max is infinite (some row has no bound yet).
    const uint32_t p_begin = t_begin / 2, pe = t_end / 2;
    auto window = [&](uint32_t wb) __attribute__((always_inline)) -> uint64_t {
        const uint32_t sp = wb + static_cast<uint32_t>(lane);
        bool keep = sp < pe;
        if (prune && theta_max < kInf && keep) {
            const uint32_t P = phys(2 * sp) >> 1;
            const float4* c = reinterpret_cast<const float4*>(s.balls + static_cast<size_t>(P) * bstride);
            const float4* o = reinterpret_cast<const float4*>(blds + 16);
            float d2 = 0.0f;
            // Four of the pair's centre loads in flight at a time (one at a time, each waiting
            // for its L2 round trip, the tests took 5.6% of a wave's time: pass 66.0-66.4 ->
            // 64.1-64.3 ms at 1M x 50, same-box A/B; eight gain nothing more).
#pragma unroll 4
            for (uint32_t j = 0; j < bstride / 4; j++) {
                const float4 a4 = c[j], b4 = o[j];
                const float e0 = a4.x - b4.x, e1 = a4.y - b4.y, e2 = a4.z - b4.z, e3 = a4.w - b4.w;
                d2 = __builtin_fmaf(e0, e0, d2);
                d2 = __builtin_fmaf(e1, e1, d2);
                d2 = __builtin_fmaf(e2, e2, d2);
                d2 = __builtin_fmaf(e3, e3, d2);
            }
            const float rho = balls[static_cast<size_t>(ball_count) * bstride + P] + blds[8];
            // Rounding of d2 (at most (d + 3) 2^-24 relative), of the square root and of
            // the differences: the 2^-12 factors leave the bound on the safe side.
            const float lb = __builtin_sqrtf(d2) * (1.0f - 0x1p-12f) - rho * (1.0f + 0x1p-12f);
            keep = !(lb > 0.0f && lb * lb * (1.0f - 0x1p-12f) > theta_max);
        }
        return __ballot(keep);
    };
    uint32_t wb = p_begin;
    uint64_t wm = window(p_begin);
    constexpr uint32_t kNone = ~0u;
    auto next = [&]() __attribute__((always_inline)) -> uint32_t {
        while (wm == 0) {
            wb += kWave;
            if (wb >= pe) return kNone;
            wm = window(wb);
        }
        const uint32_t b = static_cast<uint32_t>(__builtin_ctzll(wm));
        wm &= wm - 1;
        return wb + b;
    };
    if constexpr (OP == kOperandF16Local) {
        // 512 threads convert a 64-row tile: slot u = tid + 512 i is 16 bytes of LDS, row
        // u / (4 nch), part pp = u % (4 nch): features 8 pp .. 8 pp + 7 of the row, from 32
        // bytes of padded fp32 coordinates. 4 nch divides 512 (nch 1, 2 or 4), so a thread's
        // parts, and with them its origin values, are the same on every tile.
        static_assert(512 % (4 * NCH) == 0, "local-origin pass: nch must divide 128");
        constexpr int kRowSlots = 4 * NCH;
        constexpr int kSlots = kTileRefs * kRowSlots;
        constexpr int kPer = (kSlots + 511) / 512;
        const float sigma = local_scale(__uint_as_float(*global_ptr(s.max_norm_bits)));
        const SCRUST_GLOBAL float4* src4 = reinterpret_cast<const SCRUST_GLOBAL float4*>(global_ptr(s.pack_src));
        const SCRUST_GLOBAL float* origin = global_ptr(s.origins) + static_cast<size_t>(item.block / 2) * (32 * NCH);
        float os[kPer][8];  // -o sigma for this thread's parts
        int at[kPer];       // LDS slot of part i in a tile buffer
#pragma unroll
        for (int i = 0; i < kPer; i++) {
            const int u = tid + 512 * i;
            const int pp = u % kRowSlots;
#pragma unroll
            for (int e = 0; e < 8; e++) os[i][e] = -(origin[8 * pp + e] * sigma);
            at[i] = (u / kRowSlots) * stride + (pp / NCH) * gstride + pp % NCH;
        }
        const float limit_lo = 1.0f - 0x1p-15f, limit_up = 1.0f + 0x1p-15f;
        u32x4 pa[2 * kPer], pb[2 * kPer];
        auto load = [&](u32x4 (&P)[2 * kPer], uint32_t T) __attribute__((always_inline)) {
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                const int u = tid + 512 * i;
                if (u < kSlots) {
                    const size_t f4 = (static_cast<size_t>(phys(T)) * kTileRefs) * (8 * NCH) + 2 * static_cast<size_t>(u);
                    P[2 * i] = __builtin_bit_cast(u32x4, src4[f4]);
                    P[2 * i + 1] = __builtin_bit_cast(u32x4, src4[f4 + 1]);
                }
            }
        };
        // Registers of tile T -> fp16 operand rows in LDS buffer B. The row's last three
        // features: |b''|^2 2^-16 as fp16 hi and lo (scaled down by 2^-15 first, so their sum
        // is below |b''|^2 2^-16 despite the norm's own rounding and the MFMA's accumulation)
        // and |b''| rounded up; a sentinel row (past n) gets +inf, 0, 0 and scores -inf.
        auto store = [&](const u32x4 (&P)[2 * kPer], int B, uint32_t T) __attribute__((always_inline)) {
#pragma unroll
            for (int i = 0; i < kPer; i++) {
                const int u = tid + 512 * i;
                if (kSlots % 512 == 0 || u < kSlots) {
                    float x[8];
                    const float4 v0 = __builtin_bit_cast(float4, P[2 * i]), v1 = __builtin_bit_cast(float4, P[2 * i + 1]);
                    const float raw[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                    float part = 0.0f;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        x[e] = __builtin_fmaf(raw[e], sigma, os[i][e]);
                        part = __builtin_fmaf(x[e], x[e], part);
                    }
                    // Sum over the row's kRowSlots consecutive lanes.
#pragma unroll
                    for (int m = 1; m < kRowSlots; m <<= 1) part += __shfl_xor(part, m, kWave);
                    f16x8 h;
#pragma unroll
                    for (int e = 0; e < 8; e++) h[e] = static_cast<_Float16>(x[e]);
                    if (u % kRowSlots == kRowSlots - 1) {
                        const uint32_t ref = phys(T) * kTileRefs + static_cast<uint32_t>(u / kRowSlots);
                        const float nb = part * limit_lo * 0x1p-16f;
                        const _Float16 hi = static_cast<_Float16>(nb);
                        const _Float16 lo = static_cast<_Float16>(nb - static_cast<float>(hi));
                        const _Float16 up = f16_up(__builtin_sqrtf(part * limit_up) * (1.0f + 0x1p-20f));
                        const bool live = ref < n;
                        h[5] = live ? hi : static_cast<_Float16>(kInf);
                        h[6] = live ? lo : static_cast<_Float16>(0.0f);
                        h[7] = live ? up : static_cast<_Float16>(0.0f);
                    }
                    tile_of(B)[at[i]] = __builtin_bit_cast(u32x4, h);
                }
            }
        };
        const uint32_t last_load = t_end - 1;
        load(pa, t_begin);
        load(pb, t_begin + 1);
        store(pa, 0, t_begin);
        __syncthreads();
        for (uint32_t t = t_begin; t < t_end; t += 2) {
            load(pa, t + 2 < last_load ? t + 2 : last_load);
            consume(phys(t), 0);
            store(pb, 1, t + 1);
            __syncthreads();
            compact();
            load(pb, t + 3 < last_load ? t + 3 : last_load);
            consume(phys(t + 1), 1);
            store(pa, 0, t + 2 < last_load ? t + 2 : last_load);
            __syncthreads();
            compact();
        }
    } else if constexpr (kSplit) {
        // Split tiles in pairs, loaded straight into LDS like the whole tiles below: the coarse
        // parts of tiles t+4 and t+5 go into the ring slot tiles t-2 and t-1 left while tiles
        // t and t+1 are consumed (eight blocks between barriers); each wave waits for all its
        // loads before a plain s_barrier. A refined block reads its fine chunks from global
        // memory.
        typedef typename Op::Frag Frag;
        constexpr int KC = kCoarse, KF = NCH - kCoarse;
        static_assert(KF * kWave <= kFineStage, "a block's fine chunks fit the wave's staging");
        constexpr int kStep = 2;
        constexpr int kCRounds = KC * kStep;  // 4 KC wave-wide loads per tile, one per wave per round
        // Sources (chunk index from the step's first tile) of this lane's slots in each round.
        int csrc[kCRounds];
#pragma unroll
        for (int i = 0; i < kCRounds; i++) {
            const int u = (i * 4 + w) * kWave + lane;
            const int j = u / kCoarseSlots<NCH>, v = u % kCoarseSlots<NCH>;
            const int r = v / (4 * KC), p = v % (4 * KC);
            csrc[i] = j * kTileRefs * 4 * NCH + r * 4 * KC + (p ^ coarse_swizzle<KC>(r));
        }
        // This lane's coarse slots within a row (rows cb 16 + lc share lc's swizzle).
        int cofs[KC];
#pragma unroll
        for (int c = 0; c < KC; c++) cofs[c] = (g * KC + c) ^ coarse_swizzle<KC>(lc);
        auto coarse_of = [&](int cur) __attribute__((always_inline)) -> u32x4* {
            return cur == 0 ? dma.b0 : cur == 1 ? dma.b1 : dma.b2;
        };
        // The coarse parts of tiles T and T+1 into ring slot S.
        auto issue = [&](uint32_t T, auto slot) __attribute__((always_inline)) {
            constexpr int S = decltype(slot)::value;
            const SCRUST_GLOBAL u32x4* from = ref_rows + static_cast<size_t>(phys(T)) * tile_u4;
#pragma unroll
            for (int i = 0; i < kCRounds; i++)
                __builtin_amdgcn_global_load_lds(from + csrc[i],
                                                 (__attribute__((address_space(3))) void*)(coarse_of(S) + (i * 4 + w) * kWave),
                                                 16, 0, 0);
        };
        // Tiles t and t+1 (physical indices) from ring slot cur: the coarse steps of their
        // blocks, then the remaining steps and the filter of each block that may append.
        auto consume_pair = [&](uint32_t t, int cur) __attribute__((always_inline)) {
            f32x4 acc0[4], acc1[4];
            Frag c0[KC], c1[KC], f[KF];
            auto load_c = [&](Frag (&b)[KC], int cb) __attribute__((always_inline)) {
                const u32x4* row = coarse_of(cur) + (cb * 16 + lc) * (4 * KC);
#pragma unroll
                for (int c = 0; c < KC; c++) b[c] = __builtin_bit_cast(Frag, row[cofs[c]]);
            };
            auto coarse = [&](f32x4 (&acc)[4], const Frag (&b)[KC]) __attribute__((always_inline)) {
#pragma unroll
                for (int qb = 0; qb < 4; qb++) acc[qb] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int c = 0; c < KC; c++)
#pragma unroll
                    for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(b[c], a[qb][c], acc[qb]);
            };
            auto refine = [&](f32x4 (&acc)[4], int cb) __attribute__((always_inline)) {
                uint64_t need = 0;
#pragma unroll
                for (int qb = 0; qb < 4; qb++) {
                    const float m = __builtin_fmaxf(__builtin_fmaxf(acc[qb][0], acc[qb][1]),
                                                    __builtin_fmaxf(acc[qb][2], acc[qb][3]));
                    need |= __ballot(m > gate[qb]);
                }
                if (need) {
                    // The block's KF fine chunks go straight to this wave's LDS staging (lane l
                    // its own 16 bytes of each), all in flight at once, then one wait: loads
                    // into registers would be issued one at a time at this register budget.
                    // The previous block's reads of the staging are done (its MFMAs used them).
                    const SCRUST_GLOBAL u32x4* src = ref_rows + static_cast<size_t>(t + cb / 4) * tile_u4 +
                                                     kTileRefs * 4 * KC + ((cb % 4) * 16 + lc) * 4 * KF + g * KF;
                    u32x4* const stage = dma.fine + w * kFineStage;
#pragma unroll
                    for (int c = 0; c < KF; c++)
                        __builtin_amdgcn_global_load_lds(src + c, (__attribute__((address_space(3))) void*)(stage + c * kWave),
                                                         16, 0, 0);
                    __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
                    for (int c = KC; c < NCH; c++) {
                        f[c - KC] = __builtin_bit_cast(Frag, stage[(c - KC) * kWave + lane]);
#pragma unroll
                        for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_k32(f[c - KC], a[qb][c], acc[qb]);
                    }
                    filter(acc, t, cb);
                }
            };
            // Block cb's MFMAs run while block cb-1 is tested; block cb+2's fragments are read
            // while block cb's MFMAs run.
            load_c(c0, 0);
            load_c(c1, 1);
            __builtin_amdgcn_sched_barrier(0);
            coarse(acc0, c0);
            __builtin_amdgcn_sched_barrier(0);
            load_c(c0, 2);
            __builtin_amdgcn_sched_barrier(0);
            coarse(acc1, c1);
            __builtin_amdgcn_sched_barrier(0);
            load_c(c1, 3);
            __builtin_amdgcn_sched_barrier(0);
            static_for<0, 4 * kStep - 2>([&](auto CB) {
                constexpr int cb = decltype(CB)::value;
                if constexpr (cb % 2 == 0) {
                    refine(acc0, cb);
                    coarse(acc0, c0);
                    if constexpr (cb + 4 < 4 * kStep) load_c(c0, cb + 4);
                } else {
                    refine(acc1, cb);
                    coarse(acc1, c1);
                    if constexpr (cb + 4 < 4 * kStep) load_c(c1, cb + 4);
                }
            });
            refine(acc0, 4 * kStep - 2);
            refine(acc1, 4 * kStep - 1);
        };
        auto settle = [&]() __attribute__((always_inline)) {
            ball_publish();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            ball_absorb();
        };
        auto step = [&](uint32_t t, auto slot) __attribute__((always_inline)) {
            consume_pair(phys(t), decltype(slot)::value);
            scanned++;
            settle();
            if (compact() && prune) refresh_theta();
        };
        const std::integral_constant<int, 0> s0{};
        const std::integral_constant<int, 1> s1{};
        const std::integral_constant<int, 2> s2{};
        // Steps over the surviving pairs q0, q1, q2, ... in scan order (tile 2 q and 2 q + 1):
        // the coarse parts of the pair two steps ahead are loaded while a step is consumed.
        uint32_t q0 = next();
        if (q0 != kNone) {
            uint32_t q1 = next();
            issue(2 * q0, s0);
            if (q1 != kNone) issue(2 * q1, s1);
            settle();
            for (;;) {
                uint32_t q2 = q1 != kNone ? next() : kNone;
                if (q2 != kNone) issue(2 * q2, s2);
                step(2 * q0, s0);
                if (q1 == kNone) break;
                q0 = q1;
                q1 = q2;
                q2 = q1 != kNone ? next() : kNone;
                if (q2 != kNone) issue(2 * q2, s0);
                step(2 * q0, s1);
                if (q1 == kNone) break;
                q0 = q1;
                q1 = q2;
                q2 = q1 != kNone ? next() : kNone;
                if (q2 != kNone) issue(2 * q2, s1);
                step(2 * q0, s2);
                if (q1 == kNone) break;
                q0 = q1;
                q1 = q2;
            }
        }
        if (tid == 0) atomicAdd(&s.queue[kQueueScanned], scanned);
    } else if constexpr (!DMA) {
    if constexpr (kBall) {
        if (prune) {
            // Wide bf16 rows of a multi-round build on the spatial order: the same loop over
            // the surviving pairs (q, the next survivor's tiles loaded while q's are consumed),
            // the ball bound's theta exchange around each barrier.
            auto sync = [&]() __attribute__((always_inline)) {
                ball_publish();
                __syncthreads();
                ball_absorb();
            };
            auto after = [&]() __attribute__((always_inline)) {
                if (compact()) refresh_theta();
            };
            uint32_t q = next();
            if (q != kNone) {
                uint32_t qn = next();
                SCRUST_LOAD_TILE(pa, 2 * q)
                SCRUST_LOAD_TILE(pb, 2 * q + 1)
                SCRUST_STORE_TILE(pa, 0)
                sync();
                for (;;) {
                    scanned++;
                    // Past the last survivor the loads repeat its own tiles (never stored).
                    const uint32_t tn = qn != kNone ? 2 * qn : 2 * q;
                    SCRUST_LOAD_TILE(pa, tn)
                    consume(phys(2 * q), 0);
                    SCRUST_STORE_TILE(pb, 1)
                    sync();
                    after();
                    SCRUST_LOAD_TILE(pb, tn + 1)
                    consume(phys(2 * q + 1), 1);
                    if (qn == kNone) break;
                    SCRUST_STORE_TILE(pa, 0)
                    sync();
                    after();
                    q = qn;
                    qn = next();
                }
                sync();
                after();
            }
            if (tid == 0) atomicAdd(&s.queue[kQueueScanned], scanned);
            goto scan_done;
        }
    }
    {
    // The loop has one exit, at its bottom (the tile count is even): a mid-loop exit would
    // merge into the loop head, where the compiler must then assume the prefetch just
    // issued is outstanding and waits for it. Loads past the end reload the last tile.
    const uint32_t last_load = t_end - 1;
    SCRUST_LOAD_TILE(pa, t_begin)
    SCRUST_LOAD_TILE(pb, t_begin + 1)
    SCRUST_STORE_TILE(pa, 0)
    __syncthreads();
    for (uint32_t t = t_begin; t < t_end; t += 2) {
        SCRUST_LOAD_TILE(pa, t + 2 < last_load ? t + 2 : last_load)
        consume(phys(t), 0);
        SCRUST_STORE_TILE(pb, 1)
        __syncthreads();
        compact();
        SCRUST_LOAD_TILE(pb, t + 3 < last_load ? t + 3 : last_load)
        consume(phys(t + 1), 1);
        SCRUST_STORE_TILE(pa, 0)
        __syncthreads();
        compact();
    }
    }
    scan_done:;
    } else {
        // LDS-direct loads (global_load_lds_dwordx4): one wave-wide load fills 64 consecutive
        // 16-byte slots of a tile buffer, lane l the l-th. Slot u of a tile is reference row
        // u / stride, part q = u % stride: lane group q / gstride, chunk q % gstride, or
        // padding past the group's chunks (it loads the row's first chunk; never read).
        // Tile t+2 is loaded into the buffer that tile t-1 left while tile t is consumed;
        // before each barrier a wave waits for all its loads (so the next tile has landed
        // and its own reads of the tile just consumed are done), so a plain s_barrier
        // suffices and the compiler has no fence to drain the loads at.
        constexpr int kSlots = kTileSlots<OP, NCH>;
        constexpr int kRounds = (kSlots + 4 * kWave - 1) / (4 * kWave);
        int src_at[kRounds];
        bool src_on[kRounds];
#pragma unroll
        for (int i = 0; i < kRounds; i++) {
            const int u = (i * 4 + w) * kWave + lane;
            src_on[i] = u < kSlots;
            const int v = src_on[i] ? u : 0;
            const int q = v % stride, gg = q / gstride, c = q % gstride;
            src_at[i] = (gg < 4 && c < group_u4) ? tile_chunk(v / stride, gg, c, nch) : tile_chunk(v / stride, 0, 0, nch);
        }
        auto dma_tile = [&](uint32_t T, u32x4* to) __attribute__((always_inline)) {
            const SCRUST_GLOBAL u32x4* from = ref_rows + static_cast<size_t>(phys(T)) * tile_u4;
#pragma unroll
            for (int i = 0; i < kRounds; i++) {
                if (src_on[i])
                    __builtin_amdgcn_global_load_lds(from + src_at[i],
                                                     (__attribute__((address_space(3))) void*)(to + (i * 4 + w) * kWave),
                                                     16, 0, 0);
            }
        };
        auto settle = [&]() __attribute__((always_inline)) {
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_s_barrier();
        };
        dma_tile(t_begin, dma.b0);
        if (t_begin + 1 < t_end) dma_tile(t_begin + 1, dma.b1);
        settle();
        for (uint32_t t = t_begin;;) {
            if (t + 2 < t_end) dma_tile(t + 2, dma.b2);
            consume(phys(t), 0);
            settle();
            compact();
            if (++t >= t_end) break;
            if (t + 2 < t_end) dma_tile(t + 2, dma.b0);
            consume(phys(t), 1);
            settle();
            compact();
            if (++t >= t_end) break;
            if (t + 2 < t_end) dma_tile(t + 2, dma.b1);
            consume(phys(t), 2);
            settle();
            compact();
            if (++t >= t_end) break;
        }
    }
#undef SCRUST_LOAD_TILE
#undef SCRUST_STORE_TILE
    if constexpr (PAIRS) {
        if (persist) {
            // The item's selection from this wave's buffer rows (counts from the registers).
            auto counts = [&](int q, uint32_t (&c)[4]) __attribute__((always_inline)) { row_counts(q, c); };
            const SCRUST_GLOBAL uint64_t* wbuf = rbuf + static_cast<size_t>(w) * kWave * cap;
            if (cap <= 256) finish_rows<4>(s, item, wbuf, qbase + w * kWave, lane, false, counts);
            else finish_rows<8>(s, item, wbuf, qbase + w * kWave, lane, false, counts);
            return;
        }
    }
    // Segment counts for the final selection (select_kernel): byte g of the row's word.
    SCRUST_GLOBAL uint8_t* const counts = reinterpret_cast<SCRUST_GLOBAL uint8_t*>(
        global_ptr(s.seg_counts) + range_row_base(s, item.split));
#pragma unroll
    for (int qb = 0; qb < 4; qb++)
        counts[static_cast<size_t>(qbase + w * kWave + qb * 16 + lc) * 4 + g] =
            static_cast<uint8_t>(cnt[qb]);
}

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b to XCD b mod 8), and each
// XCD has its own L2. Maps workgroup b < count to a block so that each XCD takes a contiguous
// run of blocks in order: blocks of one group, which start their scans at the same tiles,
// then run on the same XCD and share its L2 while they are resident together.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t count) {
    constexpr uint32_t kXcds = 8;
    const uint32_t x = b % kXcds, j = b / kXcds;
    const uint32_t base = count / kXcds, extra = count % kXcds;
    return x * base + (x < extra ? x : extra) + j;
}

// At most two waves per SIMD: the allocator then keeps the tile prefetch in registers
// instead of spilling it to reach a higher occupancy the LDS budget would allow.
// Work item `v` of a single dataset's pass, with tail_block = tail_row0 / kPassRows (0 when
// every block scans in `splits` ranges): v < tail_block is a whole block's scan (block v),
// tail_block + g the g-th range item (range g / tail_blocks of block tail_block +
// g % tail_blocks). There are tail_block + (blocks - tail_block) splits items.
__device__ __forceinline__ PassItem single_item(const DatasetDesc& s, uint32_t v) {
    const uint32_t pairs = (s.n + 2 * kTileRefs - 1) / (2 * kTileRefs);
    const uint32_t blocks = s.npad / kPassRows;
    const uint32_t tail_block = s.tail_row0 / kPassRows;
    PassItem item;
    item.set = 0;
    if (v < tail_block) {
        item.split = 0;
        item.block = v;
        item.tile_begin = 0;
        item.tile_end = 2 * pairs;
    } else {
        const uint32_t g = v - tail_block, tail_blocks = blocks - tail_block;
        item.split = g / tail_blocks;
        item.block = tail_block + (g - item.split * tail_blocks);
        item.tile_begin = 2 * item.split * s.chunk_pairs;
        const uint32_t stop = (item.split + 1) * s.chunk_pairs;
        item.tile_end = 2 * (stop < pairs ? stop : pairs);
    }
    // In spatial order, a block's scan starts at the tiles of its own group of 16 blocks
    // (4096 rows, about one cell), so its rows meet their neighbours first; the 16 blocks
    // share the start, and its tiles stay in L2 for all of them.
    item.pad = s.rotate ? (4u * (item.block & ~15u)) % (2 * pairs) : 0u;
    return item;
}

// The persistent variant (PAIRS) is held to two waves per SIMD (256 registers a lane), which
// its two workgroups per CU need.
template <int OP, int NCH, bool PAIRS = false, bool BIG = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PAIRS ? 2 : 1, 2)))
void pass_kernel(const DatasetDesc* __restrict__ set) {
    const DatasetDesc& s = *set;
    const uint32_t blocks = s.npad / kPassRows;
    const uint32_t tail_block = s.tail_row0 / kPassRows;
    if constexpr (kPassSplit<OP, NCH> && PAIRS) {
        // The persistent pass (s.persist): workgroup b runs on XCD x = b mod 8 (round-robin
        // dispatch) and takes the blocks of x's contiguous run from queue x (the spatial order
        // keeps a run's blocks, which share their scans' first tiles, in one L2), then the
        // range items from queue 8. Every workgroup leaves once both are drained.
        __shared__ u32x4 b0[2 * kCoarseSlots<NCH>];
        __shared__ u32x4 b1[2 * kCoarseSlots<NCH>];
        __shared__ u32x4 b2[2 * kCoarseSlots<NCH>];
        __shared__ __attribute__((aligned(16))) float ball[kBallLds];
        __shared__ u32x4 fine[4 * kFineStage];
        __shared__ uint32_t next;
        const DmaTiles dma{b0, b1, b2, ball, fine};
        // Whole blocks are [0, whole): every block with one range, the blocks before the
        // last round's when that scans in ranges, none when every block does (items v < whole
        // are whole blocks under single_item's reading, the items from whole on ranges).
        const uint32_t whole = s.splits == 1 ? blocks : tail_block;
        const uint32_t x = blockIdx.x % 8u;
        const uint32_t base = whole / 8u, extra = whole % 8u;
        const uint32_t start = x * base + (x < extra ? x : extra), len = base + (x < extra ? 1u : 0u);
        const uint32_t tail_items = (blocks - whole) * s.splits;
        for (;;) {
            if (threadIdx.x == 0) {
                uint32_t v = ~0u;
                const uint32_t j = atomicAdd(&s.queue[x], 1u);
                if (j < len) {
                    v = start + j;
                } else {
                    const uint32_t gi = atomicAdd(&s.queue[8], 1u);
                    if (gi < tail_items) v = whole + gi;
                }
                next = v;
            }
            __syncthreads();
            const uint32_t v = next;
            __syncthreads();
            if (v == ~0u) break;
            pass_block<OP, NCH, true, true, true>(s, single_item(s, v), nullptr, dma);
            __syncthreads();
        }
        return;
    } else {
        // One work item per workgroup, whole tiles. Workgroups below the tail block scan
        // their block's whole range; from it on, workgroup tail_block + g scans range
        // j = g / tail_blocks of block tail_block + g % tail_blocks: the workgroups resident
        // together scan the same references (shared through L2).
        const PassItem item =
            single_item(s, blockIdx.x < tail_block && s.rotate ? xcd_block(blockIdx.x, tail_block) : blockIdx.x);
        if constexpr (kPassDma<OP, NCH>) {
            __shared__ u32x4 b0[kTileSlots<OP, NCH>];
            __shared__ u32x4 b1[kTileSlots<OP, NCH>];
            __shared__ u32x4 b2[kTileSlots<OP, NCH>];
            pass_block<OP, NCH, true, true, false, BIG>(s, item, nullptr, DmaTiles{b0, b1, b2});
        } else {
            __shared__ PassSmem<OP, NCH> sm;
            pass_block<OP, NCH, true, false, false, BIG>(s, item, &sm);
        }
    }
}

// Persistent pass, rows scanned in ranges (from tail_row0 on, splits > 1): the kp smallest
// keys (unordered) of the union of the rows' per-range top-kp lists (tail_cand), and the
// union's kp-th score, one wave per row:
// the same result as select_kernel over the ranges' whole buffers (the kp smallest of a
// union lie among each part's kp smallest; keys are totally ordered by (score, id)).
template <int E>
__global__ __launch_bounds__(256) void merge_kernel(const DatasetDesc* __restrict__ set) {
    const DatasetDesc& s = *set;
    const uint32_t row = s.tail_row0 + blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= s.n) return;
    const uint32_t kp = s.kp, total_keys = s.splits * kp;
    const SCRUST_GLOBAL uint64_t* src = global_ptr(s.tail_cand) + static_cast<size_t>(row - s.tail_row0) * total_keys;
    uint64_t key[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint32_t e = static_cast<uint32_t>(i * kWave + lane);
        key[i] = e < total_keys ? src[e] : kNoKey;
    }
    uint64_t m[E];
    uint32_t total = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
        m[i] = __ballot(key[i] != kNoKey);
        total += static_cast<uint32_t>(__popcll(m[i]));
    }
    SCRUST_GLOBAL uint64_t* out = global_ptr(s.cand) + static_cast<size_t>(row) * kp;
    float theta = kInf;
    if (total >= kp) {
        theta = select_keys<E>(key, static_cast<int>(kp), lane, out, 0u, 0u);
    } else {
        uint32_t base = 0;
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (key[i] != kNoKey) out[base + lanes_below(m[i])] = key[i];
            base += static_cast<uint32_t>(__popcll(m[i]));
        }
        for (uint32_t slot = base + static_cast<uint32_t>(lane); slot < kp; slot += kWave) out[slot] = kNoKey;
    }
    if (lane == 0) global_ptr(s.theta)[row] = theta;
}

void launch_merge(const DatasetDesc* set, const DatasetDesc& host, hipStream_t stream) {
    if (!host.persist || host.splits <= 1 || host.n <= host.tail_row0) return;
    const unsigned blocks = (host.n - host.tail_row0 + 3) / 4;
    const uint32_t keys = host.splits * host.kp;
    if (keys <= 128)
        hipLaunchKernelGGL(merge_kernel<2>, dim3(blocks), dim3(256), 0, stream, set);
    else if (keys <= 256)
        hipLaunchKernelGGL(merge_kernel<4>, dim3(blocks), dim3(256), 0, stream, set);
    else
        hipLaunchKernelGGL(merge_kernel<8>, dim3(blocks), dim3(256), 0, stream, set);
}

// Bounding balls of the persistent pass, one workgroup each: ball b < pairs holds reference
// rows [128 b, 128 b + 128), ball pairs + j query rows [256 j, 256 j + 256), both cut at n
// (rows past n are sentinels or clamped copies of row n - 1). The centre is the rows' mean
// (any point would do: the bound only needs the radius to cover the rows' distances from the
// stored centre); the radius is the largest f32 distance, scaled by 1 + 2^-12, which covers
// the rounding of the d-term sum ((d + 3) 2^-24 relative) and of the square root.
__global__ __launch_bounds__(256) void ball_kernel(const DatasetDesc* __restrict__ set) {
    const DatasetDesc& s = *set;
    const uint32_t n = s.n, d = s.d, stride = s.ball_stride;
    const uint32_t pairs = (n + 2 * kTileRefs - 1) / (2 * kTileRefs);
    const uint32_t count = pairs + s.npad / kPassRows;
    const uint32_t b = blockIdx.x, tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    const uint32_t r0 = b < pairs ? b * 2 * kTileRefs : (b - pairs) * kPassRows;
    const uint32_t span = b < pairs ? 2 * kTileRefs : kPassRows;
    const uint32_t r1 = r0 + span < n ? r0 + span : n;
    const SCRUST_GLOBAL float* coords = global_ptr(s.coords);
    SCRUST_GLOBAL float* balls = global_ptr(s.balls);
    __shared__ float part[4][kBallMaxD];
    __shared__ float centre[kBallMaxD];
    __shared__ float wfar[4];
    for (uint32_t j = lane; j < static_cast<uint32_t>(kBallMaxD); j += kWave) {
        float sum = 0.0f;
        if (j < d)
            for (uint32_t r = r0 + w; r < r1; r += 4) sum += coords[static_cast<size_t>(r) * d + j];
        part[w][j] = sum;
    }
    __syncthreads();
    if (tid < stride) {
        const float c = tid < d ? (part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]) / static_cast<float>(r1 - r0) : 0.0f;
        centre[tid] = c;
        balls[static_cast<size_t>(b) * stride + tid] = c;
    }
    __syncthreads();
    float far = 0.0f;
    if (r0 + tid < r1) {
        const SCRUST_GLOBAL float* x = coords + static_cast<size_t>(r0 + tid) * d;
        for (uint32_t j = 0; j < d; j++) {
            const float e = x[j] - centre[j];
            far = __builtin_fmaf(e, e, far);
        }
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) far = __builtin_fmaxf(far, __shfl_xor(far, m, kWave));
    if (lane == 0) wfar[w] = far;
    __syncthreads();
    if (tid == 0) {
        far = __builtin_fmaxf(__builtin_fmaxf(wfar[0], wfar[1]), __builtin_fmaxf(wfar[2], wfar[3]));
        balls[static_cast<size_t>(count) * stride + b] = __builtin_sqrtf(far) * (1.0f + 0x1p-12f);
    }
}

void launch_balls(const DatasetDesc* set, const DatasetDesc& host, hipStream_t stream) {
    const unsigned count = (host.n + 2 * kTileRefs - 1) / (2 * kTileRefs) + host.npad / kPassRows;
    hipLaunchKernelGGL(ball_kernel, dim3(count), dim3(256), 0, stream, set);
}

// Local-origin fp16 pass of one dataset: workgroup i scans every reference for the 512
// query rows of blocks 2i and 2i + 1 (npad is a multiple of 512), whose origin is
// origins[i]; in spatial order the scan starts at the tiles of the rows' own group.
template <int NCH>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2)))
void pass_local_kernel(const DatasetDesc* __restrict__ set) {
    const DatasetDesc& s = *set;
    const uint32_t pairs = (s.n + 2 * kTileRefs - 1) / (2 * kTileRefs);
    PassItem item;
    item.set = 0;
    item.split = 0;
    item.block = 2 * (s.rotate ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
    item.tile_begin = 0;
    item.tile_end = 2 * pairs;
    item.pad = s.rotate ? (4u * (item.block & ~15u)) % (2 * pairs) : 0u;
    __shared__ PassSmem<kOperandF16Local, NCH> sm;
    pass_block<kOperandF16Local, NCH, true>(s, item, &sm);
}

// Persistent grid over a queue of (dataset, query block, reference range) items, largest
// first. Every workgroup leaves once the queue is drained.
template <int OP, int NCH_MAX>
__device__ __forceinline__ void pass_batch_body(const DatasetDesc* __restrict__ sets,
                                                const PassItem* __restrict__ items,
                                                uint32_t n_items, uint32_t* __restrict__ counter,
                                                PassSmem<OP, NCH_MAX>& sm) {
    __shared__ uint32_t next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t item = next;
        __syncthreads();
        if (item >= n_items) break;
        const PassItem it = items[item];
        pass_block<OP, NCH_MAX, false>(sets[it.set], it, &sm);
        __syncthreads();
    }
}

template <int OP, int NCH_MAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2)))
void pass_batch_kernel(const DatasetDesc* __restrict__ sets, const PassItem* __restrict__ items,
                       uint32_t n_items, uint32_t* __restrict__ counter) {
    __shared__ PassSmem<OP, NCH_MAX> sm;
    pass_batch_body<OP, NCH_MAX>(sets, items, n_items, counter, sm);
}

// The narrowest class fits two waves per SIMD without spilling once asked to: the runtime
// chunk count costs it a few registers over the fixed-count kernel.
template <int OP, int NCH_MAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void pass_batch_narrow_kernel(const DatasetDesc* __restrict__ sets, const PassItem* __restrict__ items,
                              uint32_t n_items, uint32_t* __restrict__ counter) {
    __shared__ PassSmem<OP, NCH_MAX> sm;
    pass_batch_body<OP, NCH_MAX>(sets, items, n_items, counter, sm);
}

// Final candidate selection, one wave per query row: the KP smallest keys (unordered) of
// the row's buffers over every reference range, and the KP-th score. Non-candidates of a
// range scored at least that range's final theta, which is at least the KP-th smallest
// key of the union, so the union's KP-th score bounds every reference left out.
template <int E>
__global__ __launch_bounds__(256) void select_kernel(const DatasetDesc* __restrict__ sets,
                                                     uint32_t count, uint32_t total_rows) {
    const uint32_t flat = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (flat >= total_rows) return;
    const DatasetDesc& s = sets[find_set(sets, count, flat, false)];
    const uint32_t row = flat - s.row_begin;
    const bool whole = row < s.tail_row0;  // a row whose block scanned in one range
    if (!s.on_pass || (whole ? s.select_e1 : s.select_e) != E) return;
    const uint32_t cap = s.cap, seg = cap / 4, splits = whole ? 1u : s.splits;
    const int kp = static_cast<int>(s.kp);
    const SCRUST_GLOBAL uint64_t* buf = global_ptr(s.buf);
    const SCRUST_GLOBAL uint32_t* counts = global_ptr(s.seg_counts);
    // Key e = i*64 + lane is slot e % cap of range e / cap (uniform in i: cap >= 64).
    uint64_t key[E];
#pragma unroll
    for (int i = 0; i < E; i++) {
        const uint32_t e = static_cast<uint32_t>(i * kWave + lane);
        const uint32_t j = e / cap, p = e % cap;
        uint64_t k = kNoKey;
        if (j < splits) {
            const size_t at = range_row_base(s, j) + row;
            const uint32_t c = (counts[at] >> (8 * (p / seg))) & 0xFFu;
            if (p % seg < c) k = buf[at * cap + p];
        }
        key[i] = k;
    }
    uint64_t m[E];
    uint32_t total = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
        m[i] = __ballot(key[i] != kNoKey);
        total += static_cast<uint32_t>(__popcll(m[i]));
    }
    SCRUST_GLOBAL uint64_t* out = global_ptr(s.cand) + static_cast<size_t>(row) * kp;
    if (total >= static_cast<uint32_t>(kp)) {
        const float theta = select_keys<E>(key, kp, lane, out, 0u, 0u);
        if (lane == 0) global_ptr(s.theta)[row] = theta;
    } else {
        // Fewer keys than KP (n - 1 < KP): all of them, then kNoKey.
        uint32_t base = 0;
#pragma unroll
        for (int i = 0; i < E; i++) {
            if (key[i] != kNoKey) out[base + lanes_below(m[i])] = key[i];
            base += static_cast<uint32_t>(__popcll(m[i]));
        }
        for (uint32_t slot = base + static_cast<uint32_t>(lane); slot < static_cast<uint32_t>(kp); slot += kWave)
            out[slot] = kNoKey;
        if (lane == 0) global_ptr(s.theta)[row] = kInf;
    }
}

void launch_select(const DatasetDesc* sets, const BatchShape& shape, unsigned widths,
                   hipStream_t stream) {
    const unsigned blocks = (shape.total_rows + 3) / 4;
    if (blocks == 0) return;
    if (widths & 2u)
        hipLaunchKernelGGL(select_kernel<2>, dim3(blocks), dim3(256), 0, stream, sets, shape.count, shape.total_rows);
    if (widths & 4u)
        hipLaunchKernelGGL(select_kernel<4>, dim3(blocks), dim3(256), 0, stream, sets, shape.count, shape.total_rows);
    if (widths & 8u)
        hipLaunchKernelGGL(select_kernel<8>, dim3(blocks), dim3(256), 0, stream, sets, shape.count, shape.total_rows);
    if (widths & 16u)
        hipLaunchKernelGGL(select_kernel<16>, dim3(blocks), dim3(256), 0, stream, sets, shape.count, shape.total_rows);
}

// ---------------------------------------------------------------------------
// Dispatch over the compiled variants.
// ---------------------------------------------------------------------------

static const int kBf16Chunks[] = {1, 2, 3, 4, 5, 6, 8, 10, 13};
static const int kF32Chunks[] = {1, 2, 3, 4, 6, 8, 10, 13, 16, 20, 26, 33, 49, 57, 65};

int bf16_chunks(int d) {
    const int need = (3 * d + 3 + 31) / 32;
    for (int c : kBf16Chunks)
        if (c >= need) return c;
    return 0;
}

int fp32_chunks(int d) {
    const int need = (d + 1 + 3) / 4;
    for (int c : kF32Chunks)
        if (c >= need) return c;
    return 0;
}

int fp32_group(int nch) { return (nch + 3) / 4 * 4; }

int batch_chunks(int operand, int nch) {
    static const int bf16_max[] = {5, 10, 13};
    static const int f32_max[] = {13, 26, 33};
    const int* sizes = operand == kOperandBf16 ? bf16_max : f32_max;
    for (int i = 0; i < 3; i++)
        if (sizes[i] >= nch) return sizes[i];
    return 0;
}

template <int OP, int NCH>
static void run_pass(const DatasetDesc* set, const DatasetDesc& host, hipStream_t stream) {
    const unsigned tail_block = host.tail_row0 / kPassRows;
    const unsigned groups = tail_block + (host.npad / kPassRows - tail_block) * host.splits;
    // The persistent pass scans split tiles in pairs (its plans have buffers of 256 or 512
    // slots, room for the 32 appends per segment of a pair). Measured at 1M x 50 (A/B, same
    // box): pairs 6.5% faster than single split tiles with a ring of fine parts; pairs in a
    // ring of two slots 3.5% slower, and steps of four tiles (compactions inside the step) 5%
    // slower, than pairs in three. Every other single plan scans whole tiles in two stages:
    // at 2,638 and 100,000 cells and for server requests that is faster than split tiles
    // (A/B: 0.40 against 0.41 ms single-tile and 0.55 ms paired at 2,638; serving p99 58-65
    // against 82-109 ms), since those datasets refine most blocks.
    if constexpr (kPassSplit<OP, NCH>) {
        if (host.persist) {
            // The persistent pass launches one workgroup per slot.
            hipLaunchKernelGGL((pass_kernel<OP, NCH, true>), dim3(host.slots), dim3(256), 0, stream, set);
            return;
        }
    }
    // Plans keeping KP = 256 candidates (k > 111) have 1024-slot buffers: their own variant,
    // whose 16-keys-per-lane compaction would raise every other variant's register peak.
    if (host.cap > static_cast<uint32_t>(kCapMax)) {
        hipLaunchKernelGGL((pass_kernel<OP, NCH, false, true>), dim3(groups), dim3(256), 0, stream, set);
        return;
    }
    hipLaunchKernelGGL((pass_kernel<OP, NCH>), dim3(groups), dim3(256), 0, stream, set);
}

#define SCRUST_PASS_CASE(OP, N)          \
    case N:                              \
        run_pass<OP, N>(set, host, stream); \
        return true;

int local_chunks(int d) {
    const int need = (d + 3 + 31) / 32;
    return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : 0;
}

bool launch_pass(const DatasetDesc* set, const DatasetDesc& host, int operand,
                 hipStream_t stream) {
    if (operand == kOperandF16Local) {
        const dim3 grid(host.npad / kLocalRows), block(kLocalRows);
        switch (host.nch) {
            case 1: hipLaunchKernelGGL(pass_local_kernel<1>, grid, block, 0, stream, set); return true;
            case 2: hipLaunchKernelGGL(pass_local_kernel<2>, grid, block, 0, stream, set); return true;
            case 4: hipLaunchKernelGGL(pass_local_kernel<4>, grid, block, 0, stream, set); return true;
            default: return false;
        }
    }
    if (operand == kOperandBf16) {
        switch (host.nch) {
            SCRUST_PASS_CASE(kOperandBf16, 1) SCRUST_PASS_CASE(kOperandBf16, 2)
            SCRUST_PASS_CASE(kOperandBf16, 3) SCRUST_PASS_CASE(kOperandBf16, 4)
            SCRUST_PASS_CASE(kOperandBf16, 5) SCRUST_PASS_CASE(kOperandBf16, 6)
            SCRUST_PASS_CASE(kOperandBf16, 8) SCRUST_PASS_CASE(kOperandBf16, 10)
            SCRUST_PASS_CASE(kOperandBf16, 13)
            default: return false;
        }
    }
    switch (host.nch) {
        SCRUST_PASS_CASE(kOperandF32, 1) SCRUST_PASS_CASE(kOperandF32, 2)
        SCRUST_PASS_CASE(kOperandF32, 3) SCRUST_PASS_CASE(kOperandF32, 4)
        SCRUST_PASS_CASE(kOperandF32, 6) SCRUST_PASS_CASE(kOperandF32, 8)
        SCRUST_PASS_CASE(kOperandF32, 10) SCRUST_PASS_CASE(kOperandF32, 13)
        SCRUST_PASS_CASE(kOperandF32, 16) SCRUST_PASS_CASE(kOperandF32, 20)
        SCRUST_PASS_CASE(kOperandF32, 26) SCRUST_PASS_CASE(kOperandF32, 33)
        SCRUST_PASS_CASE(kOperandF32, 49) SCRUST_PASS_CASE(kOperandF32, 57)
        SCRUST_PASS_CASE(kOperandF32, 65)
        default: return false;
    }
}

template <int OP, int NCH_MAX>
static void run_batch(const DatasetDesc* sets, const PassItem* items, uint32_t n_items,
                      uint32_t* counter, unsigned blocks, hipStream_t stream) {
    constexpr bool kNarrow = (OP == kOperandBf16 && NCH_MAX == 5) || (OP == kOperandF32 && NCH_MAX == 13);
    if constexpr (kNarrow)
        hipLaunchKernelGGL((pass_batch_narrow_kernel<OP, NCH_MAX>), dim3(blocks), dim3(256), 0, stream,
                           sets, items, n_items, counter);
    else
        hipLaunchKernelGGL((pass_batch_kernel<OP, NCH_MAX>), dim3(blocks), dim3(256), 0, stream, sets,
                           items, n_items, counter);
}

bool launch_pass_batch(const DatasetDesc* sets, const PassItem* items, uint32_t n_items,
                       uint32_t* item_counter, int operand, int nch_max, unsigned blocks,
                       hipStream_t stream) {
    if (operand == kOperandBf16) {
        switch (nch_max) {
            case 5: run_batch<kOperandBf16, 5>(sets, items, n_items, item_counter, blocks, stream); return true;
            case 10: run_batch<kOperandBf16, 10>(sets, items, n_items, item_counter, blocks, stream); return true;
            case 13: run_batch<kOperandBf16, 13>(sets, items, n_items, item_counter, blocks, stream); return true;
            default: return false;
        }
    }
    switch (nch_max) {
        case 13: run_batch<kOperandF32, 13>(sets, items, n_items, item_counter, blocks, stream); return true;
        case 26: run_batch<kOperandF32, 26>(sets, items, n_items, item_counter, blocks, stream); return true;
        case 33: run_batch<kOperandF32, 33>(sets, items, n_items, item_counter, blocks, stream); return true;
        default: return false;
    }
}

}  // namespace scrust_hip
