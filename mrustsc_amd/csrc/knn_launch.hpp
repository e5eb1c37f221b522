// Host-side declarations of the kNN kernel launches and the per-dataset descriptor every
// kernel reads. One descriptor describes one embedding; a single call uses one, a
// batched launch many (the kernels map flat row indices onto datasets).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace scrust_hip {

constexpr int kMetricEuclidean = 0;
constexpr int kMetricCosine = 1;

constexpr int kOperandNone = 0;  // exact path: coordinates only
constexpr int kOperandBf16 = 1;
constexpr int kOperandF32 = 2;
// fp16 single products of local-origin coordinates (one origin per 512 query rows),
// converted from the padded fp32 coordinates tile by tile inside the pass.
constexpr int kOperandF16Local = 3;
constexpr int kLocalRows = 512;  // query rows per local-origin work item (8 waves x 64)
// Error model of the local-origin pass (see the certificate in rerank_kernel): fp16 rounding
// (unit 2^-11) of both operands, the MFMA's fp32 accumulation over 32 nch + 1 terms.
__host__ __device__ inline double local_product_coeff(int nch) {
    const double u = 0x1p-11, g = (32.0 * nch + 2.0) * 0x1p-24 * 1.01;
    return (2.0 * u + u * u + g * (1.0 + u) * (1.0 + u)) / (1.0 - g) * (1.0 + 0x1p-10);
}
// Absolute fp16 rounding error below the normal range (2^-25 per coordinate), over d terms.
__host__ __device__ inline double local_abs_coeff(int d) {
    double r = 1.0;
    while (r * r < d) r += 1.0;  // an integer >= sqrt(d)
    return 0x1p-25 * r * 1.01;
}

constexpr unsigned kFlagNonFinite = 1u;
constexpr unsigned kFlagBandOverflow = 2u;

constexpr int kCapMax = 512;  // largest candidate buffer (slots per query row) of the pass
// Widest boundary tie band resolved in fp64; wider bands keep the (d^2, id) order.
constexpr int kBandCap = 64;
constexpr int kMaxK = 256;      // largest k any path serves
// Candidates kept per row (KP) by the MFMA pass: 32 for k <= 15, 64 for k <= 47, 128 for
// k <= 111, 256 for k <= 239 (k + 17 <= KP: the list holds the first k and a margin for the
// boundary band). KP = 256 takes 1024-slot buffers and single builds (the batched kernel
// serves k <= kMaxBatchK).
constexpr int kMaxPassK = 239;   // largest k the MFMA candidate pass serves (KP = 256)
constexpr int kMaxBatchK = 111;  // largest k of the batched pass (KP = 128)
constexpr int kMaxLocalK = 47;   // largest k of the local-origin fp16 pass (KP = 64)
__host__ __device__ inline uint32_t pass_kp(uint32_t k) {
    return k + 17 <= 32 ? 32u : k + 17 <= 64 ? 64u : k + 17 <= 128 ? 128u : 256u;
}
constexpr int kMaxPackedD = 128;  // largest d of the packed 16-bit operands (bf16, fp16)
constexpr int kMaxPassD = 256;    // largest d with a compiled MFMA variant (fp32 past kMaxPackedD)
constexpr int kAccWordsHost = 10;
// Largest squared norm the MFMA candidate pass serves. Past it (coordinates beyond about
// 10^15, or a norm that overflowed f32) packed products, their sums or the bf16 rounding of
// the norm feature can overflow; such a dataset skips the pass and every row takes the
// exact kernel, whose f32 direct differences are the oracle's arithmetic at any magnitude.
constexpr float kPassNormLimit = 0x1p100f;
constexpr int kPassRows = 256;  // query rows per pass work item (4 waves x 64)

// bf16 operand rows hold 32 nch features in MFMA K order: element e of 16-byte chunk m is K
// index 32 (m % nch) + 8 (m / nch) + e, because MFMA step c reads chunk g * nch + c in lane
// group g and lane group g supplies K slots 8 g .. 8 g + 7 of the step.
__host__ __device__ inline int k_index(int chunk, int e, int nch) {
    return 32 * (chunk % nch) + 8 * (chunk / nch) + e;
}
// MFMA steps that hold every non-cross feature of the bf16 operand (h.h and the norm terms,
// K indices below d + 3) for every d a compiled chunk count serves; the later steps hold
// only hi/lo cross terms and zeros.
__host__ __device__ constexpr int bf16_coarse_chunks(int nch) {
    return nch <= 2 ? 1 : nch <= 5 ? 2 : nch <= 8 ? 3 : nch <= 10 ? 4 : 5;
}
constexpr int kTileRefsHost = 64;  // references per pass tile
// bf16 reference operands are stored tile by tile (64 rows, 4 nch 16-byte chunks each): a
// tile holds the coarse chunks of its rows first (steps c < kc = bf16_coarse_chunks(nch),
// 4 kc per row, lane group gg's step c at gg kc + c), then their fine chunks (steps c >= kc,
// 4 kf per row with kf = nch - kc, at gg kf + c - kc). The coarse part of a tile is then one
// contiguous run that the pass can stream without the fine part. Returns the chunk's index
// within its tile for row r of the tile.
__host__ __device__ inline int tile_chunk(int r, int gg, int c, int nch) {
    const int kc = bf16_coarse_chunks(nch), kf = nch - kc;
    return c < kc ? r * 4 * kc + gg * kc + c : kTileRefsHost * 4 * kc + r * 4 * kf + gg * kf + (c - kc);
}
// The inverse of tile_chunk: row r, lane group gg and step c of chunk u of a tile.
__host__ __device__ inline void tile_chunk_of(int u, int nch, int& r, int& gg, int& c) {
    const int kc = bf16_coarse_chunks(nch), kf = nch - kc;
    if (u < kTileRefsHost * 4 * kc) {
        r = u / (4 * kc);
        const int q = u - r * 4 * kc;
        gg = q / kc;
        c = q - gg * kc;
    } else {
        const int v = u - kTileRefsHost * 4 * kc;
        r = v / (4 * kf);
        const int q = v - r * 4 * kf;
        gg = q / kf;
        c = kc + (q - gg * kf);
    }
}
// Split scan of a short list of exact rows: at most kSplitRowsMax rows, each cut into ranges
// of kSplitRange references, each range keeping its kSplitKeep smallest keys. Rows with
// k <= kMaxSplitK take it (k + kBandCap <= kSplitKeep, so the merged keys hold the first k
// and any tie band the cap allows); rows with a larger k take the exact kernel.
constexpr uint32_t kSplitRowsMax = 256;
constexpr uint32_t kSplitRange = 8192;
constexpr uint32_t kSplitKeep = 256;
constexpr uint32_t kMaxSplitK = kSplitKeep - kBandCap;

// Device-visible description of one embedding and its workspace.
struct DatasetDesc {
    const float* x;               // input, n x d row-major
    float* coords;                // centred / unit-scaled coordinates, n x d
    float* norms;                 // f32 fma squared norms, n
    double* mean;                 // column means, d
    unsigned long long* acc;      // exact column sums, d x kAccWordsHost
    uint16_t* pack_ref;           // bf16 reference operands, ref_rows x 32 nch
    uint16_t* pack_query;         // bf16 query operands, n x 32 nch
    float* pack_f32;              // fp32 operands, ref_rows x 4 group
    uint64_t* buf;                // pass candidate buffers, range_rows() x cap
    uint32_t* seg_counts;         // keys per buffer segment (4 x u8), range_rows()
    uint64_t* cand;               // kept candidates, n x kp
    float* theta;                 // kp-th candidate score per row, n
    unsigned* max_norm_bits;      // max squared norm (as bits)
    uint32_t* out_indices;        // n x k
    float* out_distances;         // n x k
    // Pass order (null: input order): perm[p] = input row at pass position p, inv_perm the
    // inverse. Coordinates, norms, operands, buffers, candidates and thresholds are in pass
    // order, the input and the outputs in input order; ids are pass positions until the
    // rerank and the exact kernel key them by input id.
    const uint32_t* perm;
    const uint32_t* inv_perm;
    uint32_t rotate;              // 1: a block's scan starts at its own part of the order
    uint32_t n, d, k, kp, nch, group, cap;
    uint32_t row_begin;           // first flat row of this dataset
    uint32_t prep_begin;          // first flat prepare row (ref_rows per pass dataset)
    uint32_t on_pass;             // 1: candidate pass + rerank; 0: exact rows only
    uint32_t npad;                // n rounded up to whole query blocks
    uint32_t splits;              // reference ranges of a block's scan (buffer slots per row)
    uint32_t chunk_pairs;         // tile pairs per reference range (the last may be short)
    uint32_t select_e;            // keys per lane of the final selection (splits x cap / 64)
    // Single datasets: rows from tail_row0 on (whole query blocks at the end of the launch)
    // scan in `splits` ranges, the rows before it in one range, whose selection takes
    // select_e1 keys per lane. 0: every block scans in `splits` ranges.
    uint32_t tail_row0;
    uint32_t select_e1;
    // Error model of the pass (see the certificate in rerank_kernel).
    // Local-origin fp16 pass (operand kOperandF16Local): padded fp32 coordinates (ref_rows x
    // 32 nch, pass order, zero past d), one origin per kLocalRows query rows (32 nch floats
    // each), and per query row a lower bound of |a''|^2 (a'' = (a - o) sigma, local_scale).
    float* pack_src;
    float* origins;
    float* anorm;
    uint32_t local;    // 1: this dataset's pass is the local-origin fp16 pass
    float local_c;     // local_product_coeff(nch), rounded up to f32
    float local_c2;    // local_abs_coeff(d), rounded up to f32
    double split_rel;  // operand split error of a.b, relative to |a||b|
    double acc_rel;    // MFMA accumulation error, relative to the sum of |terms|
    double norm_rel;   // error of the |b|^2 feature, relative to |b|^2
    // Persistent pass (persist = 1: single datasets on split tiles whose blocks scan whole
    // ranges, the launch's last round excepted): `slots` workgroups, each with its own
    // kPassRows-row region of `buf`, take query blocks from per-XCD queues (queue[0..7]) and
    // then the last round's range items (queue[8]); the pass selects each whole block's
    // candidates itself (cand, theta), and each range item's top kp keys go to tail_cand
    // ((n - tail_row0) rows x splits ranges x kp keys) for merge_kernel.
    uint32_t persist;
    uint32_t slots;
    uint32_t* queue;
    uint64_t* tail_cand;
    // Bounding balls (pass order, coordinates) of the persistent pass and of the wide bf16
    // rows' register-staged pass of more than one round: ball b < pairs is tile
    // pair b (reference rows 128 b .. 128 b + 127), ball pairs + j query block j (rows 256 j ..
    // 256 j + 255); centre b at balls + b * ball_stride (d floats, zero to ball_stride), then
    // (pairs + blocks) radii, each an upper bound of the distance of the ball's rows from its
    // centre. Null: no step is skipped by the bound.
    float* balls;
    uint32_t ball_stride;
};

// Split-tile bf16 pass variants (the chunk counts knn_pass.hip compiles them for).
inline bool split_pass_nch(int operand, uint32_t nch) { return operand == kOperandBf16 && nch >= 2 && nch <= 5; }
// Persistent pass: 8 XCD queues, the range-item queue, and the count of tile-pair steps its
// items scanned (the others skipped by the ball bound).
constexpr int kQueueWords = 10;
constexpr int kQueueScanned = 9;
constexpr uint32_t kBallMaxDims = 256;  // widest dataset a pass keeps balls for

// Row offset of reference range j's buffers and segment counts: range 0 holds every row,
// a range j >= 1 only the rows from tail_row0 on (every row when tail_row0 is 0), so row r
// of range j sits at range_row_base(s, j) + r.
__host__ __device__ inline size_t range_row_base(const DatasetDesc& s, uint32_t j) {
    return j == 0 ? 0
                  : static_cast<size_t>(s.npad) + static_cast<size_t>(j - 1) * (s.npad - s.tail_row0) - s.tail_row0;
}

// Rows of buffers and counts over all ranges.
__host__ __device__ inline size_t range_rows(const DatasetDesc& s) {
    return static_cast<size_t>(s.npad) + static_cast<size_t>(s.splits - 1) * (s.npad - s.tail_row0);
}

// One pass work item: a query block against one reference range of its dataset.
struct PassItem {
    uint32_t set;         // dataset index
    uint32_t block;       // query block of kPassRows rows
    uint32_t split;       // buffer slot of the range (0 .. splits - 1)
    uint32_t tile_begin;  // first 64-row reference tile (even)
    uint32_t tile_end;    // one past the last tile (even)
    uint32_t pad;
};

struct BatchShape {
    uint32_t count;           // datasets
    uint32_t total_rows;      // sum of n
    uint32_t total_prep;      // sum of n + 1
    uint32_t max_row_chunks;  // max ceil(n / colsum_rows(n))
    uint32_t max_col_groups;  // max ceil(d / 64)
    uint32_t max_d;           // largest d
    uint32_t max_prep_blocks; // max ceil(ref_rows / 64) over datasets on the bf16 pass
    uint32_t max_local_blocks; // max npad / kLocalRows over datasets on the local-origin pass
    uint32_t direct_sets;      // datasets prepare_kernel serves (not on a packed 16-bit pass)
    int metric;
    int operand;
};

// Rows per column-sum block: about 512 blocks per dataset, 64 to 2048 rows each.
__host__ __device__ inline uint32_t colsum_rows(uint32_t n) {
    uint32_t rows = 64;
    while (rows < 2048 && rows * 512u < n) rows <<= 1;
    return rows;
}

// Dataset owning flat index `row`, given each dataset's first index (ascending).
__device__ __forceinline__ uint32_t find_set(const DatasetDesc* sets, uint32_t count,
                                             uint32_t row, bool prep) {
    uint32_t lo = 0, hi = count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        const uint32_t begin = prep ? sets[mid].prep_begin : sets[mid].row_begin;
        if (begin <= row) lo = mid; else hi = mid - 1;
    }
    return lo;
}

void launch_fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t stream);
// Column means, coordinates, norms and operands of every dataset.
void launch_prepare_all(const DatasetDesc* sets, const BatchShape& shape, unsigned* flags,
                        hipStream_t stream);
// Candidate pass of one dataset (compile-time chunk count): one workgroup per query block
// and reference range (splits ranges of chunk_pairs tile pairs).
bool launch_pass(const DatasetDesc* set, const DatasetDesc& host, int operand,
                 hipStream_t stream);
// Candidate pass of every item of a batch in one persistent launch.
bool launch_pass_batch(const DatasetDesc* sets, const PassItem* items, uint32_t n_items,
                       uint32_t* item_counter, int operand, int nch_max, unsigned blocks,
                       hipStream_t stream);
// Final candidate selection of every pass row over its reference ranges (one launch per
// selection width present; `widths` is a bit set of the select_e values 2, 4, 8, 16).
// Persistent pass (host.persist): the last round's rows from their per-range lists.
void launch_merge(const DatasetDesc* set, const DatasetDesc& host, hipStream_t stream);
// Bounding balls of the persistent pass's tile pairs and query blocks (DatasetDesc::balls).
void launch_balls(const DatasetDesc* set, const DatasetDesc& host, hipStream_t stream);
void launch_select(const DatasetDesc* sets, const BatchShape& shape, unsigned widths,
                   hipStream_t stream);
// Rerank of every pass row (one launch per candidate width present; `kps` is a bit set of
// the KP values 32, 64 and 128).
void launch_rerank(const DatasetDesc* sets, const BatchShape& shape, unsigned kps,
                   uint32_t* fail_list, uint32_t* fail_count, hipStream_t stream);
// Exact rows listed in `list` (flat rows, count on the device). With `partial` (the split
// scan's workspace, kSplitRowsMax x ranges_max x kSplitKeep keys), a list of at most
// kSplitRowsMax rows has its pass datasets' rows scanned by many waves per row instead.
void launch_exact(const DatasetDesc* sets, const BatchShape& shape, const uint32_t* list,
                  const uint32_t* list_count, unsigned blocks, unsigned* flags,
                  uint64_t* partial, uint32_t ranges_max, hipStream_t stream);

// MFMA chunk counts of the compiled pass variants for a dimension d (0: none).
int bf16_chunks(int d);
int fp32_chunks(int d);
int fp32_group(int nch);
// Chunks of the local-origin fp16 pass for d (d + 3 features, a compiled count 1, 2 or 4;
// 0: none, d > 125).
int local_chunks(int d);
// Chunk count of the persistent batched kernel able to serve nch (0: none).
int batch_chunks(int operand, int nch);

}  // namespace scrust_hip
