// Autotuner of the candidate pass's work shape, and the process-wide table it fills.
//
// A pass's shape is its candidate buffer (slots per row and reference range), the number of
// reference ranges each 256-row query block's scan is cut into, and the row order. The
// shapes trade different costs against each other: a small buffer is compacted more often
// but selected faster; more ranges fill the GPU on small inputs but rescan the first tiles
// of every range before a row's bound tightens; the grouped order tightens the bound early
// on clustered data at the price of L2 sharing between workgroups. The plan's defaults come
// from a cost model; the tuner measures instead. Every shape yields the same graph, which
// the tuner checks on every shape it times.
#include "../../include/scrust_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "knn_launch.hpp"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

bool TuneKey::operator<(const TuneKey& o) const {
    return std::tie(precision, metric, kp, nch, nbin) < std::tie(o.precision, o.metric, o.kp, o.nch, o.nbin);
}

namespace {

std::mutex g_table_mutex;

std::map<TuneKey, TuneEntry>& table() {
    static std::map<TuneKey, TuneEntry> t;
    return t;
}

constexpr uint32_t kRepeats = 3;       // timed runs per shape after the warm-up run
constexpr uint32_t kMaxRanges = 8;     // reference ranges per block tried (1024 / 128 slots)
constexpr uint32_t kMinRangePairs = 4; // shortest range tried, in pairs of 64-row tiles
constexpr uint32_t kMaxRounds = 4;     // more ranges only while items fit in this many rounds
// A measured shape replaces the default only when it is this much faster (the spread of
// repeated device times of one shape is about 1%).
constexpr double kMinGain = 0.98;
constexpr const char* kHeader =
    "# scrust_hip tuning table v1: precision metric kp nch n_bin tile splits order ms";

// The shapes worth timing for a dataset of `pairs` tile pairs in `blocks` query blocks on
// `slots` workgroup slots: every buffer size that holds 4 kp keys, every exact cut of the
// scan into 1..1024/cap ranges of at least kMinRangePairs pairs that keeps the work items
// within kMaxRounds rounds, and the grouped order for single-range scans.
std::vector<TuneEntry> shapes_for(uint32_t kp, uint32_t blocks, uint32_t pairs, uint64_t slots) {
    std::vector<TuneEntry> out;
    for (uint32_t cap : {128u, 256u, 512u}) {
        if (cap < 4 * kp) continue;
        for (uint32_t sp = 1; sp <= kMaxRanges && sp * cap <= 1024; sp++) {
            const uint32_t len = (pairs + sp - 1) / sp;
            if (sp > 1 && (len < kMinRangePairs || static_cast<uint64_t>(blocks) * sp > kMaxRounds * slots)) break;
            if ((pairs + len - 1) / len != sp) continue;
            out.push_back(TuneEntry{cap, sp, 1u, 0.0f});
            if (sp == 1) out.push_back(TuneEntry{cap, 1u, 2u, 0.0f});
        }
    }
    return out;
}

// Runs one shape: warm-up (its result compared with `ref_*`, or stored there when empty),
// then kRepeats timed runs; returns the best device time in `ms`.
scrust_hip_status time_shape(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                             const scrust_hip_knn_options* opt, const TuneEntry& shape,
                             std::vector<uint32_t>& ref_ids, std::vector<float>& ref_dist,
                             std::vector<uint32_t>& ids, std::vector<float>& dist, double* ms) {
    uint32_t* d_ids = static_cast<uint32_t*>(ctx->out_indices.ptr);
    float* d_dist = static_cast<float*>(ctx->out_distances.ptr);
    scrust_hip_knn_stats stats{};
    scrust_hip_status st = run_single(ctx, x, n, d, k, opt, d_ids, d_dist, &stats, &shape);
    if (st != SCRUST_HIP_OK) return st;
    if (stats.tile != shape.cap || stats.splits != shape.splits || stats.order != shape.order)
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: the plan did not take the tuned shape");
    hipError_t err = hipMemcpyAsync(ids.data(), d_ids, n * k * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(dist.data(), d_dist, n * k * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(ctx->stream);
    if (err != hipSuccess) return hip_failure(err, "copy tuning result");
    if (ref_ids.empty()) {
        ref_ids = ids;
        ref_dist = dist;
    } else if (std::memcmp(ids.data(), ref_ids.data(), ids.size() * sizeof(uint32_t)) != 0 ||
               std::memcmp(dist.data(), ref_dist.data(), dist.size() * sizeof(float)) != 0) {
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: tile " + std::to_string(shape.cap) + ", " +
                                           std::to_string(shape.splits) + " ranges, order " +
                                           std::to_string(shape.order) + " changed the graph");
    }
    double best = 0.0;
    for (uint32_t r = 0; r < kRepeats; r++) {
        st = run_single(ctx, x, n, d, k, opt, d_ids, d_dist, &stats, &shape);
        if (st != SCRUST_HIP_OK) return st;
        if (r == 0 || stats.device_ms < best) best = stats.device_ms;
    }
    *ms = best;
    return SCRUST_HIP_OK;
}

scrust_hip_status tune(scrust_hip_context* ctx, const float* x, size_t n, size_t d, size_t k,
                       const scrust_hip_knn_options* options, uint32_t flags, scrust_hip_tune_result* result) {
    *result = scrust_hip_tune_result{};
    scrust_hip_knn_options opt{};
    opt.metric = options ? options->metric : static_cast<uint32_t>(SCRUST_HIP_EUCLIDEAN);
    opt.precision = options ? options->precision : static_cast<uint32_t>(SCRUST_HIP_BF16);
    // The untuned plan: its shape is the baseline the tuned one has to beat.
    const TuneEntry defaults{0u, 0u, 0u, 0.0f};
    std::vector<SetSpec> specs{SetSpec{x, n, d, k, nullptr, nullptr}};
    Plan base;
    scrust_hip_status st = plan_layout(specs, &opt, base, ctx->compute_units, true, &defaults);
    if (st != SCRUST_HIP_OK) return st;
    if (!base.any_pass) return SCRUST_HIP_OK;  // exact path: no shape to choose
    // KP = 256 (k > 111) has one shape: 1024-slot buffers in one range.
    if (base.sets[0].kp > 128u) return SCRUST_HIP_OK;
    const DatasetDesc& s = base.sets[0];
    // (A launch's last-round blocks may scan in more ranges; the shape is the other blocks'.)
    const TuneEntry def{s.cap, s.tail_row0 ? 1u : s.splits, base.ordered ? 2u : 1u, 0.0f};
    result->default_tile = def.cap;
    result->default_splits = def.splits;
    result->default_order = def.order;
    const TuneKey key = tune_key(static_cast<int>(opt.precision), static_cast<int>(opt.metric), s.kp, n, d);
    if (flags & SCRUST_HIP_TUNE_SKIP_KNOWN) {
        TuneEntry known{};
        if (tune_lookup(key, &known)) {
            result->tile = known.cap;
            result->splits = known.splits;
            result->order = known.order;
            result->best_ms = known.ms;
            result->reused = 1;
            return SCRUST_HIP_OK;
        }
    }
    const uint32_t blocks = s.npad / kPassRows;
    const uint32_t pairs = static_cast<uint32_t>((n + 2 * kTileRefsHost - 1) / (2 * kTileRefsHost));
    std::vector<TuneEntry> shapes = shapes_for(s.kp, blocks, pairs, 2ull * ctx->compute_units);
    bool has_default = false;
    for (const TuneEntry& e : shapes)
        has_default |= e.cap == def.cap && e.splits == def.splits && e.order == def.order;
    if (!has_default) shapes.insert(shapes.begin(), def);

    hipError_t err = ctx->out_indices.reserve(n * k * sizeof(uint32_t));
    if (err == hipSuccess) err = ctx->out_distances.reserve(n * k * sizeof(float));
    if (err != hipSuccess) return hip_failure(err, "allocate tuning output");
    std::vector<uint32_t> ref_ids, ids(n * k);
    std::vector<float> ref_dist, dist(n * k);
    double default_ms = 0.0;
    TuneEntry best = def;
    double best_ms = 0.0;
    bool have_best = false;
    for (const TuneEntry& shape : shapes) {
        double ms = 0.0;
        st = time_shape(ctx, x, n, d, k, &opt, shape, ref_ids, ref_dist, ids, dist, &ms);
        if (st != SCRUST_HIP_OK) return st;
        result->shapes++;
        result->worst_ms = std::max(result->worst_ms, ms);
        if (shape.cap == def.cap && shape.splits == def.splits && shape.order == def.order) default_ms = ms;
        if (!have_best || ms < best_ms) {
            best = shape;
            best_ms = ms;
            have_best = true;
        }
    }
    if (!(best_ms < default_ms * kMinGain)) {
        best = def;
        best_ms = default_ms;
    }
    best.ms = static_cast<float>(best_ms);
    {
        std::lock_guard<std::mutex> lock(g_table_mutex);
        table()[key] = best;
    }
    result->tile = best.cap;
    result->splits = best.splits;
    result->order = best.order;
    result->best_ms = best_ms;
    result->default_ms = default_ms;
    return SCRUST_HIP_OK;
}

}  // namespace

TuneKey tune_key(int precision, int metric, uint32_t kp, size_t n, size_t d) {
    // Packed chunks of the pass operands: bf16 rows carry 3 d + 3 features in chunks of 32,
    // fp32 rows d + 1 in chunks of 4 (single and batched plans alike).
    const uint32_t nch = precision == SCRUST_HIP_BF16 ? static_cast<uint32_t>((3 * d + 3 + 31) / 32)
                                                      : static_cast<uint32_t>((d + 1 + 3) / 4);
    const uint32_t nbin = n > 1 ? static_cast<uint32_t>(std::lround(2.0 * std::log2(static_cast<double>(n)))) : 0u;
    return TuneKey{static_cast<uint32_t>(precision), static_cast<uint32_t>(metric), kp, nch, nbin};
}

bool tune_lookup(const TuneKey& key, TuneEntry* out) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    const auto& t = table();
    if (t.empty()) return false;
    auto it = t.find(key);
    if (it == t.end()) return false;
    *out = it->second;
    return true;
}

}  // namespace scrust_hip

using namespace scrust_hip;

extern "C" {

scrust_hip_status scrust_hip_tune_device(scrust_hip_context* context, const float* embedding,
                                         size_t n, size_t d, size_t k,
                                         const scrust_hip_knn_options* options, uint32_t flags,
                                         scrust_hip_tune_result* result) {
    clear_error();
    if (!result) return invalid("result", "a non-null pointer", "null");
    // The outputs live in the context; the embedding stands in for them in the null checks.
    scrust_hip_status st = validate(embedding, n, d, k, embedding, embedding, options);
    if (st != SCRUST_HIP_OK) return st;
    if (flags & ~SCRUST_HIP_TUNE_SKIP_KNOWN) return invalid("flags", "0 or SCRUST_HIP_TUNE_SKIP_KNOWN", std::to_string(flags));
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    return tune(context, embedding, n, d, k, options, flags, result);
}

scrust_hip_status scrust_hip_tune(scrust_hip_context* context, const float* embedding, size_t n,
                                  size_t d, size_t k, const scrust_hip_knn_options* options,
                                  uint32_t flags, scrust_hip_tune_result* result) {
    clear_error();
    if (!result) return invalid("result", "a non-null pointer", "null");
    scrust_hip_status st = validate(embedding, n, d, k, embedding, embedding, options);
    if (st != SCRUST_HIP_OK) return st;
    if (flags & ~SCRUST_HIP_TUNE_SKIP_KNOWN) return invalid("flags", "0 or SCRUST_HIP_TUNE_SKIP_KNOWN", std::to_string(flags));
    if (!context) return fail(SCRUST_HIP_NO_GPU, "no HIP device context");
    if (hipSetDevice(context->device) != hipSuccess)
        return fail(SCRUST_HIP_NO_GPU, "cannot select HIP device " + std::to_string(context->device));
    hipError_t err = context->input.reserve(n * d * sizeof(float));
    if (err == hipSuccess)
        err = hipMemcpyAsync(context->input.ptr, embedding, n * d * sizeof(float), hipMemcpyHostToDevice, context->stream);
    if (err != hipSuccess) return hip_failure(err, "copy input");
    return tune(context, static_cast<const float*>(context->input.ptr), n, d, k, options, flags, result);
}

scrust_hip_status scrust_hip_tune_save(const char* path) {
    clear_error();
    if (!path) return invalid("path", "a non-null pointer", "null");
    std::map<TuneKey, TuneEntry> snapshot;
    {
        std::lock_guard<std::mutex> lock(g_table_mutex);
        snapshot = table();
    }
    FILE* f = std::fopen(path, "w");
    if (!f) return invalid("path", "a writable file", path);
    bool ok = std::fprintf(f, "%s\n", kHeader) > 0;
    for (const auto& [key, e] : snapshot)
        ok &= std::fprintf(f, "%u %u %u %u %u %u %u %u %.4f\n", key.precision, key.metric, key.kp, key.nch,
                           key.nbin, e.cap, e.splits, e.order, static_cast<double>(e.ms)) > 0;
    ok &= std::fclose(f) == 0;
    if (!ok) return invalid("path", "a writable file", path);
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_tune_load(const char* path) {
    clear_error();
    if (!path) return invalid("path", "a non-null pointer", "null");
    FILE* f = std::fopen(path, "r");
    if (!f) return invalid("path", "a readable tuning table", path);
    std::map<TuneKey, TuneEntry> loaded;
    char line[512];
    unsigned lineno = 0;
    while (std::fgets(line, sizeof line, f)) {
        lineno++;
        const char* p = line;
        while (*p == ' ' || *p == '\t') p++;
        if (*p == '#' || *p == '\n' || *p == '\r' || *p == '\0') continue;
        TuneKey key{};
        TuneEntry e{};
        float ms = 0.0f;
        int used = 0;
        const int got = std::sscanf(p, "%u %u %u %u %u %u %u %u %f %n", &key.precision, &key.metric, &key.kp,
                                    &key.nch, &key.nbin, &e.cap, &e.splits, &e.order, &ms, &used);
        const bool valid = got == 9 && p[used] == '\0' && key.precision <= SCRUST_HIP_FP32 &&
                           key.metric <= SCRUST_HIP_COSINE && (key.kp == 32 || key.kp == 64 || key.kp == 128) &&
                           key.nch >= 1 && key.nch <= 1024 && key.nbin <= 64 &&
                           (e.cap == 128 || e.cap == 256 || e.cap == 512) && e.cap >= 4 * key.kp &&
                           e.splits >= 1 && e.splits * e.cap <= 1024 && (e.order == 1 || e.order == 2) &&
                           std::isfinite(ms) && ms >= 0.0f;
        if (!valid) {
            std::fclose(f);
            return invalid("tuning table", "lines of 9 fields as scrust_hip_tune_save writes them",
                           std::string(path) + " line " + std::to_string(lineno));
        }
        e.ms = ms;
        loaded[key] = e;
    }
    const bool read_error = std::ferror(f) != 0;
    std::fclose(f);
    if (read_error) return invalid("path", "a readable tuning table", path);
    std::lock_guard<std::mutex> lock(g_table_mutex);
    for (const auto& [key, e] : loaded) table()[key] = e;
    return SCRUST_HIP_OK;
}

void scrust_hip_tune_clear(void) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    table().clear();
}

size_t scrust_hip_tune_size(void) {
    std::lock_guard<std::mutex> lock(g_table_mutex);
    return table().size();
}

}  // extern "C"
