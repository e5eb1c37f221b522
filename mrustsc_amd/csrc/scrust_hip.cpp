// Host side of scrust_hip: context, argument validation, and the extern "C" boundary
// declared in include/scrust_hip.h.
//
// Mirrors the reference's Metal host code (crates/scrust-gpu/src/kernels/knn.rs:36-114
// and crates/scrust-gpu/src/context.rs:12-90): validate, prepare coordinates, dispatch,
// wait, check the status, read back. Centring and packing happen on the device here; the
// reference does them on the host (knn.rs:72, :127-158).
#include "../../include/scrust_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "knn_connect.hpp"
#include "knn_launch.hpp"
#include "knn_tsne.hpp"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

namespace {
thread_local std::string g_last_error;
}

scrust_hip_status fail(scrust_hip_status status, const std::string& message) {
    g_last_error = message;
    return status;
}

const char* last_error() { return g_last_error.c_str(); }

void clear_error() { g_last_error.clear(); }

// "{parameter} must be {constraint}, got {value}" -- the Display form of
// Error::InvalidParameter (crates/scrust-core/src/error.rs:12-17).
scrust_hip_status invalid(const char* parameter, const char* constraint, const std::string& value) {
    return fail(SCRUST_HIP_INVALID_PARAMETER,
                std::string(parameter) + " must be " + constraint + ", got " + value);
}

scrust_hip_status hip_failure(hipError_t err, const char* step) {
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation)
        return fail(SCRUST_HIP_OUT_OF_MEMORY,
                    std::string("HIP kernel knn: ") + step + ": " + hipGetErrorString(err));
    return fail(SCRUST_HIP_KERNEL, std::string("HIP kernel knn: ") + step + ": " + hipGetErrorString(err));
}

// Argument checks in the order of knn_metal (knn.rs:37-59). A cell is never its own
// neighbour, so k neighbours need k + 1 cells.
scrust_hip_status validate(const void* embedding, size_t n, size_t d, size_t k,
                           const void* out_indices, const void* out_distances,
                           const scrust_hip_knn_options* options) {
    if (k == 0) return invalid("k", "at least 1", std::to_string(k));
    if (d == 0) return invalid("embedding", "at least one dimension", std::to_string(d));
    if (n < k + 1) return invalid("k", "smaller than the cell count", std::to_string(k));
    if (k > static_cast<size_t>(kMaxK))
        return invalid("k", "within the supported maximum (256)", std::to_string(k));
    if (n >= 0xFFFFFFFFull) return invalid("embedding", "fewer than 2^32 - 1 cells", std::to_string(n));
    if (d > 8192) return invalid("embedding", "at most 8192 dimensions", std::to_string(d));
    if (!embedding) return invalid("embedding", "a non-null pointer", "null");
    if (!out_indices || !out_distances) return invalid("output", "a non-null pointer", "null");
    if (options) {
        if (options->metric > SCRUST_HIP_COSINE)
            return invalid("metric", "euclidean or cosine", std::to_string(options->metric));
        if (options->precision > SCRUST_HIP_FP16)
            return invalid("precision", "bf16, fp32, exact or fp16", std::to_string(options->precision));
        if (options->tile != 0 && options->tile != 128 && options->tile != 256 && options->tile != 512)
            return invalid("tile", "0, 128, 256 or 512", std::to_string(options->tile));
        if (options->order > 2)
            return invalid("order", "0 (automatic), 1 (input) or 2 (grouped)", std::to_string(options->order));
    }
    return SCRUST_HIP_OK;
}

Buffer::~Buffer() {
    if (ptr) (void)hipFree(ptr);
}

hipError_t Buffer::reserve(size_t bytes) {
    if (bytes <= capacity) return hipSuccess;
    if (ptr) {
        hipError_t err = hipFree(ptr);
        ptr = nullptr;
        capacity = 0;
        if (err != hipSuccess) return err;
    }
    size_t rounded = (bytes + 4095) / 4096 * 4096;
    hipError_t err = hipMalloc(&ptr, rounded);
    if (err != hipSuccess) {
        ptr = nullptr;
        return err;
    }
    capacity = rounded;
    return hipSuccess;
}

bool device_is_gfx950(int device, hipDeviceProp_t* props_out) {
    hipDeviceProp_t props;
    if (hipGetDeviceProperties(&props, device) != hipSuccess) return false;
    if (props_out) *props_out = props;
    return std::strncmp(props.gcnArchName, "gfx950", 6) == 0;
}

static double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return static_cast<double>(ms);
}

#define SCRUST_CHECK(expr, step)                                 \
    do {                                                         \
        hipError_t err_ = (expr);                                \
        if (err_ != hipSuccess) return hip_failure(err_, step);  \
    } while (0)

// Reads the plan's flags and exact-row count after completion; maps flags to errors.
static scrust_hip_status read_status(const Plan& plan, hipStream_t stream, uint32_t* list_count,
                                     uint32_t* scanned_steps = nullptr) {
    uint32_t host[kScalarCount] = {};
    SCRUST_CHECK(hipMemcpyAsync(host, plan.d_scalars, sizeof host, hipMemcpyDeviceToHost, stream),
                 "read status");
    SCRUST_CHECK(hipStreamSynchronize(stream), "wait for completion");
    if (host[kScalarFlags] & kFlagNonFinite) return invalid("embedding", "finite values", "NaN or infinity");
    if (host[kScalarFlags] & kFlagBandOverflow)
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: a boundary tie band ran past the kept keys");
    if (list_count) *list_count = host[kScalarListCount];
    if (scanned_steps) *scanned_steps = host[kScalarItemCounter + kQueueScanned];
    return SCRUST_HIP_OK;
}

scrust_hip_status run_single(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                             const scrust_hip_knn_options* options, uint32_t* out_indices,
                             float* out_distances, scrust_hip_knn_stats* stats,
                             const TuneEntry* force, const Workspace* workspace) {
    std::vector<SetSpec> specs{SetSpec{x, n, d, k, out_indices, out_distances}};
    Plan plan;
    scrust_hip_status st = plan_layout(specs, options, plan, ctx->compute_units, true, force);
    if (st != SCRUST_HIP_OK) return st;
    void* arena = nullptr;
    void* meta = nullptr;
    if (workspace) {
        // Arena first, metadata after it at a 256-byte boundary.
        auto need_of = [](const Plan& p) { return (p.arena_bytes + 255) / 256 * 256 + p.meta_bytes; };
        if (need_of(plan) > workspace->bytes && !force) {
            // The compact work shape (the smallest candidate buffer, one reference range,
            // automatic row order): the same graph in the least memory.
            const TuneEntry compact{128u, 1u, 0u, 0.0f};
            st = plan_layout(specs, options, plan, ctx->compute_units, true, &compact);
            if (st != SCRUST_HIP_OK) return st;
        }
        const size_t meta_at = (plan.arena_bytes + 255) / 256 * 256;
        const size_t need = meta_at + plan.meta_bytes;
        if (need > workspace->bytes)
            return invalid("embedding", ("a build whose workspace fits the " + std::to_string(workspace->bytes) +
                                         "-byte scratch array (this one needs " + std::to_string(need) + " bytes)").c_str(),
                           std::to_string(n) + " x " + std::to_string(d));
        arena = workspace->ptr;
        meta = static_cast<char*>(workspace->ptr) + meta_at;
    } else {
        SCRUST_CHECK(ctx->arena.reserve(plan.arena_bytes), "allocate workspace");
        SCRUST_CHECK(ctx->meta.reserve(plan.meta_bytes), "allocate metadata");
        arena = ctx->arena.ptr;
        meta = ctx->meta.ptr;
    }
    st = plan_place(plan, arena, meta, ctx->stream);
    if (st != SCRUST_HIP_OK) return st;
    st = plan_enqueue(plan, true, ctx->compute_units, ctx->stream, ctx->events);
    if (st != SCRUST_HIP_OK) return st;
    uint32_t listed = 0, scanned = 0;
    st = read_status(plan, ctx->stream, &listed, &scanned);
    if (st != SCRUST_HIP_OK) return st;
    if (stats) {
        hipEvent_t* ev = ctx->events;
        stats->prepare_ms = event_ms(ev[0], ev[1]);
        stats->pass_ms = event_ms(ev[1], ev[2]);
        stats->rerank_ms = event_ms(ev[2], ev[3]);
        stats->fallback_ms = event_ms(ev[3], ev[4]);
        stats->device_ms = event_ms(ev[0], ev[4]);
        stats->fallback_rows = listed;
        stats->candidates = plan.any_pass ? plan.sets[0].kp : 0;
        stats->pass_flops = plan.pass_flops;
        stats->path = plan.any_pass ? static_cast<uint32_t>(plan.precision) : SCRUST_HIP_EXACT;
        stats->tile = plan.sets[0].cap;
        const DatasetDesc& s0 = plan.sets[0];
        const bool tail = plan.any_pass && s0.tail_row0 != 0;
        stats->splits = plan.any_pass ? (tail ? 1u : s0.splits) : 0u;
        stats->tail_blocks = tail ? (s0.npad - s0.tail_row0) / kPassRows : 0u;
        stats->tail_splits = tail ? s0.splits : 0u;
        stats->order = plan.ordered ? 2u : 1u;
        stats->tuned = plan.tuned ? 1u : 0u;
        stats->workspace_bytes = plan.arena_bytes + plan.meta_bytes;
        stats->persistent = plan.any_pass && s0.persist ? 1u : 0u;
        stats->reserved = 0u;
        // A pass with the ball bound counts the tile-pair steps its items scanned (256 query
        // rows x 128 references each); every other pass scans each block against every pair.
        const uint64_t ref_rows = (static_cast<uint64_t>(s0.n) + 2 * kTileRefsHost - 1) / (2 * kTileRefsHost) * (2 * kTileRefsHost);
        stats->pairs_scored = !plan.any_pass ? 0u
                              : s0.balls ? static_cast<uint64_t>(scanned) * kPassRows * 2 * kTileRefsHost
                                         : static_cast<uint64_t>(s0.npad) * ref_rows;
    }
    return SCRUST_HIP_OK;
}

}  // namespace scrust_hip

namespace scrust_hip {

// Lays the batch out under `knobs`, uploads its metadata and captures the whole pipeline in
// a new graph (replacing the previous one).
static scrust_hip_status batch_build(scrust_hip_batch* batch, const BatchKnobs& knobs) {
    scrust_hip_context* context = batch->context;
    hipStream_t s = context->stream;
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for the batch");
    Plan plan;
    scrust_hip_status st = plan_layout(batch->specs, batch->has_options ? &batch->options : nullptr, plan,
                                       context->compute_units, false, nullptr, &knobs);
    if (st != SCRUST_HIP_OK) return st;
    if (batch->exec) {
        (void)hipGraphExecDestroy(batch->exec);
        batch->exec = nullptr;
    }
    if (batch->graph) {
        (void)hipGraphDestroy(batch->graph);
        batch->graph = nullptr;
    }
    SCRUST_CHECK(batch->arena.reserve(plan.arena_bytes), "allocate batch workspace");
    SCRUST_CHECK(batch->meta.reserve(plan.meta_bytes), "allocate batch workspace");
    st = plan_place(plan, batch->arena.ptr, batch->meta.ptr, s);
    if (st != SCRUST_HIP_OK) return st;
    SCRUST_CHECK(hipStreamSynchronize(s), "upload batch metadata");
    // Capture the whole pipeline once; runs replay it.
    SCRUST_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin graph capture");
    st = plan_enqueue(plan, false, context->compute_units, s, nullptr, batch->side, batch->fork, batch->join);
    hipGraph_t graph = nullptr;
    hipError_t err = hipStreamEndCapture(s, &graph);
    if (st != SCRUST_HIP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (err != hipSuccess) return hip_failure(err, "end graph capture");
    err = hipGraphInstantiate(&batch->exec, graph, nullptr, nullptr, 0);
    if (err != hipSuccess) {
        (void)hipGraphDestroy(graph);
        batch->exec = nullptr;
        return hip_failure(err, "instantiate graph");
    }
    batch->graph = graph;
    batch->plan = std::move(plan);
    batch->knobs = knobs;
    return SCRUST_HIP_OK;
}

// One replay of the batch's graph, waited for: device time and the exact-row count.
static scrust_hip_status batch_replay(scrust_hip_batch* batch, double* ms, uint32_t* listed) {
    scrust_hip_context* ctx = batch->context;
    if (!batch->exec) return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: the batch has no captured graph");
    hipStream_t s = ctx->stream;
    SCRUST_CHECK(hipEventRecord(ctx->events[0], s), "record event");
    SCRUST_CHECK(hipGraphLaunch(batch->exec, s), "launch graph");
    SCRUST_CHECK(hipEventRecord(ctx->events[1], s), "record event");
    scrust_hip_status st = read_status(batch->plan, s, listed);
    if (st != SCRUST_HIP_OK) return st;
    *ms = event_ms(ctx->events[0], ctx->events[1]);
    return SCRUST_HIP_OK;
}

// Copies every dataset's outputs to `ids` / `dist` (concatenated).
static scrust_hip_status batch_outputs(scrust_hip_batch* batch, std::vector<uint32_t>& ids, std::vector<float>& dist) {
    size_t total = 0;
    for (const SetSpec& sp : batch->specs) total += sp.n * sp.k;
    ids.resize(total);
    dist.resize(total);
    hipStream_t s = batch->context->stream;
    size_t at = 0;
    for (const SetSpec& sp : batch->specs) {
        SCRUST_CHECK(hipMemcpyAsync(ids.data() + at, sp.out_indices, sp.n * sp.k * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, s),
                     "copy batch output");
        SCRUST_CHECK(hipMemcpyAsync(dist.data() + at, sp.out_distances, sp.n * sp.k * sizeof(float),
                                    hipMemcpyDeviceToHost, s),
                     "copy batch output");
        at += sp.n * sp.k;
    }
    SCRUST_CHECK(hipStreamSynchronize(s), "copy batch output");
    return SCRUST_HIP_OK;
}

// The batch tuner: every (cut, buffer) shape, each checked against the first shape's
// outputs and timed over three replays; the fastest is kept (the creation shape unless
// another is at least 2% faster).
static scrust_hip_status batch_tune_shapes(scrust_hip_batch* batch, scrust_hip_batch_tune_result* result) {
    constexpr uint32_t kRepeats = 3;
    constexpr double kMinGain = 0.98;
    *result = scrust_hip_batch_tune_result{};
    const BatchKnobs created = batch->knobs;
    const uint32_t created_cap = batch->has_options && batch->options.tile ? batch->options.tile : 256u;
    result->default_cut = created.cut ? created.cut : 3u;
    result->default_tile = created.cap ? created.cap : created_cap;
    std::vector<BatchKnobs> shapes;
    shapes.push_back(BatchKnobs{result->default_cut, result->default_tile});
    for (uint32_t cut : {1u, 2u, 3u, 4u, 6u, 8u})
        for (uint32_t cap : {128u, 256u, 512u})
            if (cut != shapes[0].cut || cap != shapes[0].cap) shapes.push_back(BatchKnobs{cut, cap});
    std::vector<uint32_t> ref_ids, ids;
    std::vector<float> ref_dist, dist;
    BatchKnobs best = shapes[0];
    double best_ms = 0.0, default_ms = 0.0;
    for (size_t i = 0; i < shapes.size(); i++) {
        scrust_hip_status st = batch_build(batch, shapes[i]);
        if (st != SCRUST_HIP_OK) return st;
        double ms = 0.0;
        uint32_t listed = 0;
        st = batch_replay(batch, &ms, &listed);
        if (st != SCRUST_HIP_OK) return st;
        st = batch_outputs(batch, i == 0 ? ref_ids : ids, i == 0 ? ref_dist : dist);
        if (st != SCRUST_HIP_OK) return st;
        if (i > 0 && (ids != ref_ids || std::memcmp(dist.data(), ref_dist.data(), dist.size() * sizeof(float)) != 0))
            return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: batch shape cut " + std::to_string(shapes[i].cut) +
                                               ", tile " + std::to_string(shapes[i].cap) + " changed a graph");
        double fastest = 0.0;
        for (uint32_t r = 0; r < kRepeats; r++) {
            st = batch_replay(batch, &ms, &listed);
            if (st != SCRUST_HIP_OK) return st;
            if (r == 0 || ms < fastest) fastest = ms;
        }
        result->shapes++;
        result->worst_ms = std::max(result->worst_ms, fastest);
        if (i == 0) default_ms = best_ms = fastest;
        else if (fastest < best_ms) {
            best = shapes[i];
            best_ms = fastest;
        }
    }
    if (!(best_ms < default_ms * kMinGain)) {
        best = shapes[0];
        best_ms = default_ms;
    }
    scrust_hip_status st = batch_build(batch, best);
    if (st != SCRUST_HIP_OK) return st;
    result->cut = best.cut;
    result->tile = best.cap;
    result->best_ms = best_ms;
    result->default_ms = default_ms;
    return SCRUST_HIP_OK;
}

// A failed tuning run leaves the batch with its creation shape again (when that can still be
// built) and reports the first error.
static scrust_hip_status batch_tune(scrust_hip_batch* batch, scrust_hip_batch_tune_result* result) {
    const BatchKnobs created = batch->knobs;
    scrust_hip_status st = batch_tune_shapes(batch, result);
    if (st == SCRUST_HIP_OK) return st;
    const std::string message = last_error();
    (void)batch_build(batch, created);
    return fail(st, message);
}

}  // namespace scrust_hip

using namespace scrust_hip;

scrust_hip_context::~scrust_hip_context() {
    if (stream) {
        (void)hipSetDevice(device);
        (void)hipStreamSynchronize(stream);
    }
    for (hipEvent_t& e : events)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

scrust_hip_batch::~scrust_hip_batch() {
    if (context) (void)hipSetDevice(context->device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
}

// ---------------------------------------------------------------------------
// extern "C" boundary.
// ---------------------------------------------------------------------------

extern "C" {

const char* scrust_hip_last_error(void) { return last_error(); }

int scrust_hip_gpu_available(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
    for (int dev = 0; dev < count; dev++)
        if (device_is_gfx950(dev, nullptr)) return 1;
    return 0;
}

scrust_hip_status scrust_hip_context_create(int device, scrust_hip_context** out) {
    clear_error();
    if (!out) return invalid("out", "a non-null pointer", "null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(SCRUST_HIP_NO_GPU, "no HIP device is available on this machine");
    if (device < 0 || device >= count)
        return fail(SCRUST_HIP_NO_GPU, "no HIP device " + std::to_string(device));
    hipDeviceProp_t props;
    if (!device_is_gfx950(device, &props))
        return fail(SCRUST_HIP_NO_GPU, "HIP device " + std::to_string(device) + " is not gfx950");
    scrust_hip_context* ctx = new (std::nothrow) scrust_hip_context();
    if (!ctx) return fail(SCRUST_HIP_OUT_OF_MEMORY, "host allocation failed");
    ctx->device = device;
    ctx->compute_units = props.multiProcessorCount;
    hipError_t err = hipSetDevice(device);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    for (hipEvent_t& e : ctx->events)
        if (err == hipSuccess) err = hipEventCreate(&e);
    if (err != hipSuccess) {
        delete ctx;
        return hip_failure(err, "create context");
    }
    // The allocation-free entry point's scratch address on this device (once per device; a
    // failure is reported by that entry point, the others do not need it).
    noalloc_setup(device);
    *out = ctx;
    return SCRUST_HIP_OK;
}

void scrust_hip_context_destroy(scrust_hip_context* context) { delete context; }

size_t scrust_hip_max_supported_k(const scrust_hip_context* context) {
    (void)context;
    return static_cast<size_t>(kMaxK);
}

scrust_hip_status scrust_hip_knn_device(scrust_hip_context* context, const float* embedding,
                                        size_t n, size_t d, size_t k,
                                        const scrust_hip_knn_options* options,
                                        uint32_t* out_indices, float* out_distances,
                                        scrust_hip_knn_stats* stats) {
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    return run_single(context, embedding, n, d, k, options, out_indices, out_distances, stats);
}

scrust_hip_status scrust_hip_knn_ex(scrust_hip_context* context, const float* embedding,
                                    size_t n, size_t d, size_t k,
                                    const scrust_hip_knn_options* options,
                                    uint32_t* out_indices, float* out_distances,
                                    scrust_hip_knn_stats* stats) {
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipStream_t s = context->stream;
    SCRUST_CHECK(context->input.reserve(n * d * sizeof(float)), "allocate input");
    SCRUST_CHECK(context->out_indices.reserve(n * k * sizeof(uint32_t)), "allocate output");
    SCRUST_CHECK(context->out_distances.reserve(n * k * sizeof(float)), "allocate output");
    SCRUST_CHECK(hipMemcpyAsync(context->input.ptr, embedding, n * d * sizeof(float), hipMemcpyHostToDevice, s),
                 "copy input");
    st = run_single(context, static_cast<const float*>(context->input.ptr), n, d, k, options,
                    static_cast<uint32_t*>(context->out_indices.ptr),
                    static_cast<float*>(context->out_distances.ptr), stats);
    if (st != SCRUST_HIP_OK) return st;
    SCRUST_CHECK(hipMemcpyAsync(out_indices, context->out_indices.ptr, n * k * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s),
                 "copy indices");
    SCRUST_CHECK(hipMemcpyAsync(out_distances, context->out_distances.ptr, n * k * sizeof(float),
                                hipMemcpyDeviceToHost, s),
                 "copy distances");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for completion");
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_knn(scrust_hip_context* context, const float* embedding, size_t n,
                                 size_t d, size_t k, uint32_t* out_indices, float* out_distances) {
    return scrust_hip_knn_ex(context, embedding, n, d, k, nullptr, out_indices, out_distances, nullptr);
}

namespace {

using scrust_hip::fail;

// Argument checks of connectivities (neighbors.rs:209-230) on the sizes and pointers.
scrust_hip_status check_graph(const void* indices, const void* distances, size_t n, size_t k,
                              const void* indptr, const void* out_indices, const void* out_values,
                              const size_t* out_nnz) {
    if (n == 0 || k == 0)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected a neighbour graph with at least one cell and one neighbour, got " +
                                          std::to_string(n) + " x " + std::to_string(k));
    if (n >= 0xFFFFFFFFull)
        return scrust_hip::invalid("n", "below 2^32 - 1", std::to_string(n));
    if (!indices || !distances || !indptr || !out_indices || !out_values || !out_nnz)
        return scrust_hip::invalid("buffers", "non-null", "a null pointer");
    return SCRUST_HIP_OK;
}

// Runs the pipeline on device buffers and reads nnz and the validation flags.
scrust_hip_status run_connectivities(scrust_hip_context* ctx, const uint32_t* indices,
                                     const float* distances, size_t n, size_t k, uint32_t* indptr,
                                     uint32_t* out_indices, float* out_values, size_t capacity,
                                     size_t* out_nnz) {
    using namespace scrust_hip;
    const ConnLayout l = conn_layout(n, k, conn_temp_bytes(n, k));
    SCRUST_CHECK(ctx->conn.reserve(l.total), "allocate connectivities workspace");
    hipStream_t s = ctx->stream;
    SCRUST_CHECK(conn_enqueue(indices, distances, n, k, ctx->conn.ptr, l, indptr, out_indices, out_values,
                              capacity, s),
                 "launch connectivities");
    uint32_t flags = 0, nnz = 0;
    SCRUST_CHECK(hipMemcpyAsync(&flags, static_cast<char*>(ctx->conn.ptr) + l.flags, sizeof flags,
                                hipMemcpyDeviceToHost, s),
                 "read flags");
    SCRUST_CHECK(hipMemcpyAsync(&nnz, indptr + n, sizeof nnz, hipMemcpyDeviceToHost, s), "read nnz");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for connectivities");
    if (flags & kConnFlagBadId)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected neighbour ids below " + std::to_string(n) +
                                          ", got an out-of-range neighbour id");
    if (flags & kConnFlagBadDistance) return invalid("distances", "finite and >= 0", "a negative or non-finite value");
    if (flags & kConnFlagRepeatedId)
        return invalid("indices", "distinct neighbour ids within each row", "a repeated id");
    if (nnz > capacity)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected a capacity of " + std::to_string(nnz) +
                                          " entries, got " + std::to_string(capacity));
    *out_nnz = nnz;
    return SCRUST_HIP_OK;
}

}  // namespace

scrust_hip_status scrust_hip_connectivities_device(scrust_hip_context* context,
                                                   const uint32_t* indices,
                                                   const float* distances, size_t n, size_t k,
                                                   uint32_t* out_indptr, uint32_t* out_indices,
                                                   float* out_values, size_t capacity,
                                                   size_t* out_nnz) {
    scrust_hip::clear_error();
    scrust_hip_status st = check_graph(indices, distances, n, k, out_indptr, out_indices, out_values, out_nnz);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    return run_connectivities(context, indices, distances, n, k, out_indptr, out_indices, out_values,
                              capacity, out_nnz);
}

scrust_hip_status scrust_hip_connectivities(scrust_hip_context* context, const uint32_t* indices,
                                            const float* distances, size_t n, size_t k,
                                            uint32_t* out_indptr, uint32_t* out_indices,
                                            float* out_values, size_t capacity, size_t* out_nnz) {
    using namespace scrust_hip;
    clear_error();
    scrust_hip_status st = check_graph(indices, distances, n, k, out_indptr, out_indices, out_values, out_nnz);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipStream_t s = context->stream;
    const size_t nk = n * k, full = 2 * nk;
    SCRUST_CHECK(context->conn_in.reserve(nk * (sizeof(uint32_t) + sizeof(float))), "allocate graph");
    SCRUST_CHECK(context->conn_out.reserve((n + 1) * sizeof(uint32_t) + full * (sizeof(uint32_t) + sizeof(float)) + 512),
                 "allocate result");
    uint32_t* d_idx = static_cast<uint32_t*>(context->conn_in.ptr);
    float* d_dist = reinterpret_cast<float*>(d_idx + nk);
    uint32_t* d_indptr = static_cast<uint32_t*>(context->conn_out.ptr);
    uint32_t* d_cols = d_indptr + (n + 1);
    float* d_vals = reinterpret_cast<float*>(d_cols + full);
    SCRUST_CHECK(hipMemcpyAsync(d_idx, indices, nk * sizeof(uint32_t), hipMemcpyHostToDevice, s), "copy indices");
    SCRUST_CHECK(hipMemcpyAsync(d_dist, distances, nk * sizeof(float), hipMemcpyHostToDevice, s), "copy distances");
    size_t nnz = 0;
    st = run_connectivities(context, d_idx, d_dist, n, k, d_indptr, d_cols, d_vals, full, &nnz);
    if (st != SCRUST_HIP_OK) return st;
    if (nnz > capacity)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected a capacity of " + std::to_string(nnz) +
                                          " entries, got " + std::to_string(capacity));
    SCRUST_CHECK(hipMemcpyAsync(out_indptr, d_indptr, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                 "copy indptr");
    SCRUST_CHECK(hipMemcpyAsync(out_indices, d_cols, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                 "copy column ids");
    SCRUST_CHECK(hipMemcpyAsync(out_values, d_vals, nnz * sizeof(float), hipMemcpyDeviceToHost, s),
                 "copy weights");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for copies");
    *out_nnz = nnz;
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_neighbors(scrust_hip_context* context, const float* embedding, size_t n,
                                      size_t d, size_t k, const scrust_hip_knn_options* options,
                                      uint32_t* out_indices, float* out_distances,
                                      uint32_t* conn_indptr, uint32_t* conn_indices,
                                      float* conn_values, size_t capacity, size_t* conn_nnz) {
    using namespace scrust_hip;
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!conn_indptr || !conn_indices || !conn_values || !conn_nnz)
        return invalid("buffers", "non-null", "a null pointer");
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipStream_t s = context->stream;
    const size_t nk = n * k, full = 2 * nk;
    SCRUST_CHECK(context->input.reserve(n * d * sizeof(float)), "allocate input");
    SCRUST_CHECK(context->out_indices.reserve(nk * sizeof(uint32_t)), "allocate output");
    SCRUST_CHECK(context->out_distances.reserve(nk * sizeof(float)), "allocate output");
    SCRUST_CHECK(context->conn_out.reserve((n + 1) * sizeof(uint32_t) + full * (sizeof(uint32_t) + sizeof(float)) + 512),
                 "allocate result");
    SCRUST_CHECK(hipMemcpyAsync(context->input.ptr, embedding, n * d * sizeof(float), hipMemcpyHostToDevice, s),
                 "copy input");
    uint32_t* d_ids = static_cast<uint32_t*>(context->out_indices.ptr);
    float* d_dist = static_cast<float*>(context->out_distances.ptr);
    st = run_single(context, static_cast<const float*>(context->input.ptr), n, d, k, options, d_ids, d_dist,
                    nullptr);
    if (st != SCRUST_HIP_OK) return st;
    // The graph stays on the device for the fuzzy simplicial set.
    uint32_t* d_indptr = static_cast<uint32_t*>(context->conn_out.ptr);
    uint32_t* d_cols = d_indptr + (n + 1);
    float* d_vals = reinterpret_cast<float*>(d_cols + full);
    size_t nnz = 0;
    st = run_connectivities(context, d_ids, d_dist, n, k, d_indptr, d_cols, d_vals, full, &nnz);
    if (st != SCRUST_HIP_OK) return st;
    if (nnz > capacity)
        return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected a capacity of " + std::to_string(nnz) +
                                          " entries, got " + std::to_string(capacity));
    SCRUST_CHECK(hipMemcpyAsync(out_indices, d_ids, nk * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy indices");
    SCRUST_CHECK(hipMemcpyAsync(out_distances, d_dist, nk * sizeof(float), hipMemcpyDeviceToHost, s), "copy distances");
    SCRUST_CHECK(hipMemcpyAsync(conn_indptr, d_indptr, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                 "copy indptr");
    SCRUST_CHECK(hipMemcpyAsync(conn_indices, d_cols, nnz * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                 "copy column ids");
    SCRUST_CHECK(hipMemcpyAsync(conn_values, d_vals, nnz * sizeof(float), hipMemcpyDeviceToHost, s), "copy weights");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for copies");
    *conn_nnz = nnz;
    return SCRUST_HIP_OK;
}

}  // extern "C"

namespace {

using scrust_hip::fail;

// Argument checks of the t-SNE input distances: no features is a shape error
// (tsne.rs:120-122), more cells than the exact formulation allows an invalid parameter
// (tsne.rs:146-152).
scrust_hip_status check_tsne(const void* points, size_t n, size_t d, const void* out) {
    if (d == 0) return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected at least one feature, got zero features");
    if (n == 0) return fail(SCRUST_HIP_SHAPE, "shape mismatch: expected at least one cell, got zero cells");
    if (n > scrust_hip::kTsneMaxCells)
        return scrust_hip::invalid("n_cells", "at most 20000 for the exact O(n^2) formulation", std::to_string(n));
    if (d > 8192) return scrust_hip::invalid("points", "at most 8192 features", std::to_string(d));
    if (!points || !out) return scrust_hip::invalid("buffers", "non-null", "a null pointer");
    return SCRUST_HIP_OK;
}

// Centring and the distance tiles on device buffers; reads the non-finite flag.
scrust_hip_status run_tsne(scrust_hip_context* ctx, const float* points, size_t n, size_t d, float* out) {
    using namespace scrust_hip;
    const TsneLayout l = tsne_layout(n, d);
    SCRUST_CHECK(ctx->tsne.reserve(l.total), "allocate t-SNE workspace");
    hipStream_t s = ctx->stream;
    SCRUST_CHECK(tsne_enqueue(points, n, d, ctx->tsne.ptr, l, out, s), "launch t-SNE distances");
    uint32_t flags = 0;
    SCRUST_CHECK(hipMemcpyAsync(&flags, static_cast<char*>(ctx->tsne.ptr) + l.flags, sizeof flags,
                                hipMemcpyDeviceToHost, s),
                 "read flags");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for t-SNE distances");
    if (flags & kTsneFlagNonFinite) return invalid("points", "finite values", "NaN or infinity");
    return SCRUST_HIP_OK;
}

}  // namespace

extern "C" {

scrust_hip_status scrust_hip_tsne_squared_distances_device(scrust_hip_context* context, const float* points,
                                                           size_t n, size_t d, float* out) {
    using namespace scrust_hip;
    clear_error();
    scrust_hip_status st = check_tsne(points, n, d, out);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    return run_tsne(context, points, n, d, out);
}

scrust_hip_status scrust_hip_tsne_squared_distances(scrust_hip_context* context, const float* points,
                                                    size_t n, size_t d, float* out) {
    using namespace scrust_hip;
    clear_error();
    scrust_hip_status st = check_tsne(points, n, d, out);
    if (st != SCRUST_HIP_OK) return st;
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipStream_t s = context->stream;
    const size_t cells = n * n;
    SCRUST_CHECK(context->input.reserve(n * d * sizeof(float)), "allocate input");
    SCRUST_CHECK(context->tsne_out.reserve(cells * sizeof(float)), "allocate output");
    SCRUST_CHECK(hipMemcpyAsync(context->input.ptr, points, n * d * sizeof(float), hipMemcpyHostToDevice, s),
                 "copy input");
    float* d_out = static_cast<float*>(context->tsne_out.ptr);
    st = run_tsne(context, static_cast<const float*>(context->input.ptr), n, d, d_out);
    if (st != SCRUST_HIP_OK) return st;
    SCRUST_CHECK(hipMemcpyAsync(out, d_out, cells * sizeof(float), hipMemcpyDeviceToHost, s), "copy distances");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for copies");
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_neighbors_device(scrust_hip_context* context, const float* embedding,
                                             size_t n, size_t d, size_t k,
                                             const scrust_hip_knn_options* options,
                                             uint32_t* out_indices, float* out_distances,
                                             uint32_t* conn_indptr, uint32_t* conn_indices,
                                             float* conn_values, size_t capacity, size_t* conn_nnz) {
    using namespace scrust_hip;
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!conn_indptr || !conn_indices || !conn_values || !conn_nnz)
        return invalid("buffers", "non-null", "a null pointer");
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    st = run_single(context, embedding, n, d, k, options, out_indices, out_distances, nullptr);
    if (st != SCRUST_HIP_OK) return st;
    return run_connectivities(context, out_indices, out_distances, n, k, conn_indptr, conn_indices,
                              conn_values, capacity, conn_nnz);
}

scrust_hip_status scrust_hip_batch_create(scrust_hip_context* context,
                                          const scrust_hip_dataset* datasets, size_t count,
                                          const scrust_hip_knn_options* options,
                                          scrust_hip_batch** out) {
    clear_error();
    if (!out) return invalid("out", "a non-null pointer", "null");
    *out = nullptr;
    if (count == 0 || !datasets) return invalid("datasets", "at least one dataset", std::to_string(count));
    std::vector<SetSpec> specs;
    size_t total = 0;
    for (size_t i = 0; i < count; i++) {
        const scrust_hip_dataset& ds = datasets[i];
        scrust_hip_status st = validate(ds.embedding, ds.n, ds.d, ds.k, ds.out_indices, ds.out_distances, options);
        if (st != SCRUST_HIP_OK) return fail(st, "dataset " + std::to_string(i) + ": " + last_error());
        total += ds.n;
        specs.push_back(SetSpec{ds.embedding, ds.n, ds.d, ds.k, ds.out_indices, ds.out_distances});
    }
    if (total >= 0xFFFFFFFFull) return invalid("datasets", "fewer than 2^32 - 1 cells in total", std::to_string(total));
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    scrust_hip_batch* batch = new (std::nothrow) scrust_hip_batch();
    if (!batch) return fail(SCRUST_HIP_OUT_OF_MEMORY, "host allocation failed");
    batch->context = context;
    batch->specs = std::move(specs);
    batch->has_options = options != nullptr;
    if (options) batch->options = *options;
    // The side stream of the pass's parallel branches (default priority).
    hipError_t err = hipStreamCreateWithFlags(&batch->side, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&batch->fork, hipEventDisableTiming);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&batch->join, hipEventDisableTiming);
    if (err != hipSuccess) {
        delete batch;
        return hip_failure(err, "create batch stream");
    }
    scrust_hip_status st = batch_build(batch, BatchKnobs{0, 0});
    if (st != SCRUST_HIP_OK) {
        delete batch;
        return st;
    }
    *out = batch;
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_batch_run(scrust_hip_batch* batch, scrust_hip_batch_stats* stats) {
    clear_error();
    if (!batch) return invalid("batch", "a non-null pointer", "null");
    scrust_hip_context* ctx = batch->context;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(ctx->device));
    double ms = 0.0;
    uint32_t listed = 0;
    scrust_hip_status st = batch_replay(batch, &ms, &listed);
    if (st != SCRUST_HIP_OK) return st;
    if (stats) {
        const Plan& plan = batch->plan;
        stats->device_ms = ms;
        stats->datasets = plan.sets.size();
        stats->pass_items = plan.items.size();
        stats->exact_rows = listed;
        stats->pass_flops = plan.pass_flops;
        stats->cells = plan.shape.total_rows;
    }
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_batch_tune(scrust_hip_batch* batch, scrust_hip_batch_tune_result* result) {
    clear_error();
    if (!batch) return invalid("batch", "a non-null pointer", "null");
    if (!result) return invalid("result", "a non-null pointer", "null");
    scrust_hip_context* ctx = batch->context;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(ctx->device));
    return batch_tune(batch, result);
}

void scrust_hip_batch_destroy(scrust_hip_batch* batch) { delete batch; }

}  // extern "C"
