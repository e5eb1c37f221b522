// Internal host types of scrust_hip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/scrust_hip.h"
#include "knn_launch.hpp"
#include "knn_order.hpp"

namespace scrust_hip {

// A grow-only device allocation: a context reuses its buffers across calls, so a
// steady stream of same-sized calls allocates once.
struct Buffer {
    void* ptr = nullptr;
    size_t capacity = 0;
    Buffer() = default;
    Buffer(const Buffer&) = delete;
    Buffer& operator=(const Buffer&) = delete;
    ~Buffer();
    hipError_t reserve(size_t bytes);
};

// One embedding of a call or a batch, before device placement.
struct SetSpec {
    const float* x;
    size_t n, d, k;
    uint32_t* out_indices;
    float* out_distances;
};

// Items of one persistent batched pass launch (one compiled chunk class).
struct PassGroup {
    int nch_max;
    uint32_t begin, count;
};

// Work shape of one dataset's candidate pass. Every shape gives the same graph; the
// autotuner (knn_tune.cpp) times them and keeps the fastest per bucket.
struct TuneEntry {
    uint32_t cap;     // candidate buffer slots per row and reference range: 128, 256 or 512
    uint32_t splits;  // reference ranges per query block
    uint32_t order;   // 1 input order, 2 rows grouped by nearest pivot
    float ms;         // device time of the tuning call that chose it
};

// Bucket of the tuning table: path, metric, candidates kept, packed feature chunks (from d)
// and n in half-octaves.
struct TuneKey {
    uint32_t precision, metric, kp, nch, nbin;
    bool operator<(const TuneKey& o) const;
};

// The bucket of an (n, d) dataset keeping kp candidates per row on the given path.
TuneKey tune_key(int precision, int metric, uint32_t kp, size_t n, size_t d);
bool tune_lookup(const TuneKey& key, TuneEntry* out);

// Work shape of a batch as a whole (the batch tuner's search space): `cut` sets the longest
// reference range of one work item to 1/cut of a workgroup slot's share of the batch (0: 3),
// `cap` the candidate buffer of every dataset (0: the options' tile, else 256; at least
// 4 kp either way).
struct BatchKnobs {
    uint32_t cut;
    uint32_t cap;
};

// Device placement and launch parameters of a set of embeddings.
struct Plan {
    std::vector<DatasetDesc> sets;  // device pointers filled in
    std::vector<PassItem> items;    // persistent-pass work items, by class, largest first
    std::vector<PassGroup> groups;  // batched pass launches
    std::vector<uint32_t> exact_rows;  // flat rows of datasets on the exact path
    BatchShape shape{};
    int operand = kOperandNone;
    int precision = SCRUST_HIP_BF16;
    int nch_max = 0;
    unsigned kps = 0;  // bit set of the KP values present (32, 64, 128)
    unsigned select_widths = 0;  // select_e values present (bit set)
    size_t count_words = 0;      // segment counts of all datasets (u32 words)
    size_t counts_offset = 0;    // their arena offset
    void* arena_base = nullptr;  // set by plan_place
    bool any_pass = false;
    bool tuned = false;  // a dataset's work shape came from the tuning table
    // Spatial order of a single dataset's pass (knn_order.hip): its workspace's arena offset.
    bool ordered = false;
    OrderLayout order{};
    size_t order_offset = 0;
    // Split scan of a short list of uncertified rows (exact_split_kernel): ranges per row of
    // the largest pass dataset, and its key workspace's arena offset (0 ranges: none).
    uint32_t split_ranges = 0;
    size_t split_offset = 0;
    uint64_t pass_flops = 0;
    size_t arena_bytes = 0;
    size_t meta_bytes = 0;
    // Device views (valid after place()).
    DatasetDesc* d_sets = nullptr;
    PassItem* d_items = nullptr;
    uint32_t* d_list = nullptr;     // exact rows: prefilled exact-path rows, then uncertified
    uint32_t* d_scalars = nullptr;  // [0] flags, [1] list count, [2..] item counters
};

// Scalar slots of Plan::d_scalars.
constexpr int kScalarFlags = 0;
constexpr int kScalarListCount = 1;
constexpr int kScalarItemCounter = 2;  // first of one counter per pass group
constexpr int kScalarCount = 16;

scrust_hip_status fail(scrust_hip_status status, const std::string& message);
scrust_hip_status invalid(const char* parameter, const char* constraint, const std::string& value);
scrust_hip_status hip_failure(hipError_t err, const char* step);
scrust_hip_status validate(const void* embedding, size_t n, size_t d, size_t k,
                           const void* out_indices, const void* out_distances,
                           const scrust_hip_knn_options* options);
const char* last_error();
void clear_error();
bool device_is_gfx950(int device, hipDeviceProp_t* props_out);

// Lays out a plan for `specs` (no device work) for a device of `compute_units` CUs.
// `single`: the plan is enqueued as one call (plan_enqueue with single = true: exactly one
// dataset, the compile-time-specialised pass kernel); otherwise as a batch (the persistent
// pass over work items), whatever the dataset count. With `force`, a single dataset's pass
// takes that work shape (cap 0: the untuned defaults, table ignored); otherwise a single
// dataset with an automatic tile looks its bucket up in the tuning table. Batches take
// their shape from `knobs` (null: the defaults).
scrust_hip_status plan_layout(const std::vector<SetSpec>& specs,
                              const scrust_hip_knn_options* options, Plan& plan,
                              int compute_units, bool single, const TuneEntry* force = nullptr,
                              const BatchKnobs* knobs = nullptr, bool shared = false);
// Points the plan's descriptors into `arena` and `meta` (sized arena_bytes / meta_bytes)
// and uploads descriptors, items and the exact-row list on `stream`. With `pinned` (host
// memory of at least meta_bytes that stays untouched until the copy completes) the upload
// is one asynchronous copy; without it the copies read the plan's own vectors.
scrust_hip_status plan_place(Plan& plan, void* arena, void* meta, hipStream_t stream,
                             void* pinned = nullptr);
// Enqueues the whole pipeline. `single` uses the compile-time-specialised pass kernel
// (one dataset); otherwise the persistent batched pass. Events (6, optional) mark the
// stage boundaries.
scrust_hip_status plan_enqueue(const Plan& plan, bool single, int compute_units,
                               hipStream_t stream, hipEvent_t* events,
                               hipStream_t side = nullptr, hipEvent_t fork = nullptr,
                               hipEvent_t join = nullptr);

// Caller-provided device memory for a single call's arena and launch metadata.
struct Workspace {
    void* ptr;
    size_t bytes;
};

// One embedding already on the device, synchronously (the body of scrust_hip_knn_device).
// With `workspace`, the plan is placed there (no allocation; a plan that does not fit is an
// INVALID_PARAMETER error); otherwise in the context's grow-only buffers.
scrust_hip_status run_single(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                             const scrust_hip_knn_options* options, uint32_t* out_indices,
                             float* out_distances, scrust_hip_knn_stats* stats,
                             const TuneEntry* force = nullptr, const Workspace* workspace = nullptr);

// Takes the device address of the allocation-free entry point's scratch array on `device`
// (the current device), once per device; called at context creation.
void noalloc_setup(int device);

}  // namespace scrust_hip

struct scrust_hip_context {
    int device = 0;
    int compute_units = 0;
    hipStream_t stream = nullptr;
    hipEvent_t events[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    scrust_hip::Buffer arena, meta;                       // single-call plan workspace
    scrust_hip::Buffer input, out_indices, out_distances;  // host-input entry point
    scrust_hip::Buffer conn;                              // connectivities workspace
    scrust_hip::Buffer conn_in, conn_out;                 // its host-input staging
    scrust_hip::Buffer tsne, tsne_out;                    // t-SNE distances workspace, output staging
    scrust_hip::Buffer pca, pca_io;                       // PCA workspace, its host-input staging
    ~scrust_hip_context();
};

struct scrust_hip_batch {
    scrust_hip_context* context = nullptr;
    // What the batch was created from, kept so the batch tuner can lay it out again.
    std::vector<scrust_hip::SetSpec> specs;
    scrust_hip_knn_options options{};
    bool has_options = false;
    scrust_hip::BatchKnobs knobs{0, 0};
    scrust_hip::Plan plan;
    scrust_hip::Buffer arena, meta;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // A second stream and its fork / join events: the pass launches of different chunk
    // classes are captured as parallel branches of the graph.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    ~scrust_hip_batch();
};
