// Host side of scrust_hip: context, argument validation, the per-call kernel pipeline
// and the extern "C" boundary declared in include/scrust_hip.h.
//
// Mirrors the reference's Metal host code (crates/scrust-gpu/src/kernels/knn.rs:36-114
// and crates/scrust-gpu/src/context.rs:12-90): validate, prepare coordinates, dispatch,
// wait, check the command status, read back. Centring and packing happen on the device
// here; the reference does them on the host (knn.rs:72, :127-158).
#include "../../include/scrust_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "knn_launch.hpp"
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
        if (options->precision > SCRUST_HIP_EXACT)
            return invalid("precision", "bf16, fp32 or exact", std::to_string(options->precision));
    }
    return SCRUST_HIP_OK;
}

// ---------------------------------------------------------------------------
// Context.
// ---------------------------------------------------------------------------

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

// ---------------------------------------------------------------------------
// The pipeline on device-resident data.
// ---------------------------------------------------------------------------

namespace scrust_hip {

static double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return static_cast<double>(ms);
}

#define SCRUST_CHECK(expr, step)                       \
    do {                                               \
        hipError_t err_ = (expr);                      \
        if (err_ != hipSuccess) return hip_failure(err_, step); \
    } while (0)

scrust_hip_status run_knn(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                          const scrust_hip_knn_options* options, uint32_t* out_indices,
                          float* out_distances, scrust_hip_knn_stats* stats) {
    const int metric = options ? static_cast<int>(options->metric) : kMetricEuclidean;
    int precision = options ? static_cast<int>(options->precision) : SCRUST_HIP_BF16;
    if (k > static_cast<size_t>(kMaxPassK) || d > static_cast<size_t>(kMaxPassD)) precision = SCRUST_HIP_EXACT;
    const int di = static_cast<int>(d), ki = static_cast<int>(k);
    const uint32_t nu = static_cast<uint32_t>(n);

    int operand = kOperandNone, nch = 0, group = 0, kp = 0;
    double split_rel = 0.0, acc_rel = 0.0, norm_rel = 0.0;
    if (precision == SCRUST_HIP_BF16) {
        // c = h + l + r with |r| <= 2^-16 |c|: dropping l.l and the residuals costs
        // <= 3.02 * 2^-16 |a||b|; |b|^2 = nh + nl + nl2 to 2^-24; products of bf16 are
        // exact in f32, the accumulation is bounded as 2 * (terms + 1) * 2^-24.
        operand = kOperandBf16;
        nch = bf16_chunks(di);
        split_rel = 3.02 * 0x1p-16;
        acc_rel = 2.0 * (32.0 * nch + 1.0) * 0x1p-24 * 1.05;
        norm_rel = 0x1p-24 * 1.01;
    } else if (precision == SCRUST_HIP_FP32) {
        operand = kOperandF32;
        nch = fp32_chunks(di);
        group = fp32_group(nch);
        acc_rel = 2.0 * (4.0 * nch + 1.0) * 0x1p-24 * 1.05;
    }
    if (operand != kOperandNone) kp = (ki + 17 <= 32) ? 32 : 64;
    const size_t npad = (n + 255) / 256 * 256;

    Workspace& ws = ctx->ws;
    SCRUST_CHECK(ws.coords.reserve(n * d * sizeof(float)), "allocate coordinates");
    SCRUST_CHECK(ws.norms.reserve(n * sizeof(float)), "allocate norms");
    SCRUST_CHECK(ws.mean.reserve(d * sizeof(double)), "allocate means");
    SCRUST_CHECK(ws.acc.reserve(d * kAccWordsHost * sizeof(unsigned long long)), "allocate sums");
    SCRUST_CHECK(ws.scalars.reserve(16 * sizeof(uint32_t)), "allocate scalars");
    SCRUST_CHECK(ws.fail_list.reserve(n * sizeof(uint32_t)), "allocate fallback list");
    if (operand == kOperandBf16) {
        const size_t row = 32ull * nch * sizeof(uint16_t);
        SCRUST_CHECK(ws.pack_ref.reserve((n + 1) * row), "allocate operands");
        SCRUST_CHECK(ws.pack_query.reserve((n + 1) * row), "allocate operands");
    } else if (operand == kOperandF32) {
        SCRUST_CHECK(ws.pack_ref.reserve((n + 1) * 4ull * group * sizeof(float)), "allocate operands");
    }
    if (operand != kOperandNone) {
        SCRUST_CHECK(ws.buf.reserve(npad * kCap * sizeof(uint64_t)), "allocate candidate buffers");
        SCRUST_CHECK(ws.cand.reserve(n * kp * sizeof(uint64_t)), "allocate candidates");
        SCRUST_CHECK(ws.theta.reserve(n * sizeof(float)), "allocate thresholds");
    }
    uint32_t* scalars = static_cast<uint32_t*>(ws.scalars.ptr);
    unsigned* max_norm_bits = scalars + 0;
    unsigned* flags = scalars + 1;
    uint32_t* fail_count = scalars + 2;
    hipStream_t s = ctx->stream;
    hipEvent_t* ev = ctx->events;

    SCRUST_CHECK(hipEventRecord(ev[0], s), "record event");
    launch_fill_u32(scalars, 16, 0u, s);
    if (metric == kMetricEuclidean)
        launch_column_means(x, n, di, static_cast<unsigned long long*>(ws.acc.ptr),
                            static_cast<double*>(ws.mean.ptr), s);
    PrepareArgs pa{};
    pa.x = x;
    pa.n = n;
    pa.d = di;
    pa.metric = metric;
    pa.mean = static_cast<const double*>(ws.mean.ptr);
    pa.coords = static_cast<float*>(ws.coords.ptr);
    pa.norms = static_cast<float*>(ws.norms.ptr);
    pa.operand = operand;
    pa.pack_ref = static_cast<uint16_t*>(ws.pack_ref.ptr);
    pa.pack_query = static_cast<uint16_t*>(ws.pack_query.ptr);
    pa.pack_f32 = static_cast<float*>(ws.pack_ref.ptr);
    pa.nch = nch;
    pa.group = group;
    pa.max_norm_bits = max_norm_bits;
    pa.flags = flags;
    launch_prepare(pa, s);
    SCRUST_CHECK(hipGetLastError(), "launch prepare");
    SCRUST_CHECK(hipEventRecord(ev[1], s), "record event");

    ExactArgs ea{};
    ea.coords = pa.coords;
    ea.norms = pa.norms;
    ea.n = nu;
    ea.d = di;
    ea.k = ki;
    ea.metric = metric;
    ea.out_indices = out_indices;
    ea.out_distances = out_distances;
    ea.flags = flags;

    if (operand != kOperandNone) {
        PassArgs ps{};
        ps.operand = operand;
        ps.nch = nch;
        ps.kp = kp;
        ps.d = di;
        ps.pack_query = pa.pack_query;
        ps.pack_ref = pa.pack_ref;
        ps.pack_f32 = pa.pack_f32;
        ps.norms = pa.norms;
        ps.n = nu;
        ps.buf = static_cast<uint64_t*>(ws.buf.ptr);
        ps.cand = static_cast<uint64_t*>(ws.cand.ptr);
        ps.theta = static_cast<float*>(ws.theta.ptr);
        if (!launch_pass(ps, s))
            return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: no compiled pass for d = " + std::to_string(d));
        SCRUST_CHECK(hipGetLastError(), "launch candidate pass");
        SCRUST_CHECK(hipEventRecord(ev[2], s), "record event");

        RerankArgs ra{};
        ra.coords = pa.coords;
        ra.norms = pa.norms;
        ra.n = nu;
        ra.d = di;
        ra.k = ki;
        ra.kp = kp;
        ra.metric = metric;
        ra.cand = ps.cand;
        ra.theta = ps.theta;
        ra.max_norm_bits = max_norm_bits;
        ra.split_rel = split_rel;
        ra.acc_rel = acc_rel;
        ra.norm_rel = norm_rel;
        ra.out_indices = out_indices;
        ra.out_distances = out_distances;
        ra.fail_list = static_cast<uint32_t*>(ws.fail_list.ptr);
        ra.fail_count = fail_count;
        launch_rerank(ra, s);
        SCRUST_CHECK(hipGetLastError(), "launch rerank");
        SCRUST_CHECK(hipEventRecord(ev[3], s), "record event");

        ea.list = ra.fail_list;
        ea.list_count = fail_count;
        const unsigned blocks = static_cast<unsigned>(std::min<size_t>(n, 2048));
        launch_exact(ea, blocks, s);
        SCRUST_CHECK(hipGetLastError(), "launch exact rows");
        SCRUST_CHECK(hipEventRecord(ev[4], s), "record event");
    } else {
        SCRUST_CHECK(hipEventRecord(ev[2], s), "record event");
        SCRUST_CHECK(hipEventRecord(ev[3], s), "record event");
        ea.list = nullptr;
        ea.list_count = nullptr;
        const unsigned blocks = static_cast<unsigned>(std::min<size_t>(n, 8192));
        launch_exact(ea, blocks, s);
        SCRUST_CHECK(hipGetLastError(), "launch exact rows");
        SCRUST_CHECK(hipEventRecord(ev[4], s), "record event");
    }
    uint32_t host_scalars[4] = {0, 0, 0, 0};
    SCRUST_CHECK(hipMemcpyAsync(host_scalars, scalars, sizeof host_scalars, hipMemcpyDeviceToHost, s),
                 "read status");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for completion");
    if (host_scalars[1] & kFlagNonFinite)
        return invalid("embedding", "finite values", "NaN or infinity");
    if (host_scalars[1] & kFlagBandOverflow)
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: a boundary tie band exceeded 512 entries");
    if (stats) {
        stats->prepare_ms = event_ms(ev[0], ev[1]);
        stats->pass_ms = operand != kOperandNone ? event_ms(ev[1], ev[2]) : 0.0;
        stats->rerank_ms = operand != kOperandNone ? event_ms(ev[2], ev[3]) : 0.0;
        stats->fallback_ms = event_ms(ev[3], ev[4]);
        stats->device_ms = event_ms(ev[0], ev[4]);
        stats->fallback_rows = operand != kOperandNone ? host_scalars[2] : n;
        stats->candidates = static_cast<uint64_t>(kp);
        stats->pass_flops = operand != kOperandNone ? 2ull * n * n * d : 0ull;
        stats->path = static_cast<uint32_t>(precision);
        stats->tile = 0;
    }
    return SCRUST_HIP_OK;
}

}  // namespace scrust_hip

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
    return run_knn(context, embedding, n, d, k, options, out_indices, out_distances, stats);
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
    Workspace& ws = context->ws;
    hipStream_t s = context->stream;
    SCRUST_CHECK(ws.input.reserve(n * d * sizeof(float)), "allocate input");
    SCRUST_CHECK(ws.out_indices.reserve(n * k * sizeof(uint32_t)), "allocate output");
    SCRUST_CHECK(ws.out_distances.reserve(n * k * sizeof(float)), "allocate output");
    SCRUST_CHECK(hipMemcpyAsync(ws.input.ptr, embedding, n * d * sizeof(float), hipMemcpyHostToDevice, s),
                 "copy input");
    st = run_knn(context, static_cast<const float*>(ws.input.ptr), n, d, k, options,
                 static_cast<uint32_t*>(ws.out_indices.ptr), static_cast<float*>(ws.out_distances.ptr), stats);
    if (st != SCRUST_HIP_OK) return st;
    SCRUST_CHECK(hipMemcpyAsync(out_indices, ws.out_indices.ptr, n * k * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s),
                 "copy indices");
    SCRUST_CHECK(hipMemcpyAsync(out_distances, ws.out_distances.ptr, n * k * sizeof(float),
                                hipMemcpyDeviceToHost, s),
                 "copy distances");
    SCRUST_CHECK(hipStreamSynchronize(s), "wait for completion");
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_knn(scrust_hip_context* context, const float* embedding, size_t n,
                                 size_t d, size_t k, uint32_t* out_indices, float* out_distances) {
    return scrust_hip_knn_ex(context, embedding, n, d, k, nullptr, out_indices, out_distances, nullptr);
}

}  // extern "C"
