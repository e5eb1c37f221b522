// Serving: a request loop for kNN builds with admission by device memory and overlap of
// admitted builds on the device.
//
// The server allocates its device-memory budget once and hands out regions of it with a
// first-fit allocator. A scheduler thread admits queued requests in arrival order while
// a slot (one HIP stream, pinned staging buffers, a completion event) is idle and the
// request's region fits (past a head whose region does not, a later request that fits may go
// first, a bounded number of times); the build (upload, the single-embedding pipeline, download of
// the results into pinned memory) is enqueued on the slot's stream, so builds admitted
// together run concurrently on the device. Completion is observed by polling the slot's
// event; the results are copied into the caller's buffers and the region and slot are
// released, which lets the next queued request in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "knn_launch.hpp"
#include "scrust_hip_internal.hpp"

namespace scrust_hip {

// Admission: how far past a head that does not fit the scheduler looks, and how many
// requests may pass a head before it waits for memory itself.
constexpr size_t kBypassWindow = 16;
constexpr uint32_t kBypassMax = 32;

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr size_t kRegionAlign = 4096;

size_t round_region(size_t v) { return (v + kRegionAlign - 1) / kRegionAlign * kRegionAlign; }

// First-fit allocator over [0, size).
class RangeAllocator {
  public:
    void reset(size_t size) {
        free_.clear();
        free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t* offset) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t start = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[start + bytes] = rest;
            *offset = start;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t offset, size_t bytes) {
        used_ -= bytes;
        auto it = free_.emplace(offset, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t peak() const { return peak_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0, peak_ = 0;
};

struct Request {
    uint64_t ticket = 0;
    const float* x = nullptr;
    bool device_input = false;  // x is a device pointer (no staging, no upload)
    size_t n = 0, d = 0, k = 0;
    scrust_hip_knn_options options{};
    uint32_t* out_indices = nullptr;
    float* out_distances = nullptr;
    int64_t submit_ns = 0, admit_ns = 0, done_ns = 0;
    scrust_hip_status status = SCRUST_HIP_OK;
    std::string error;
    bool done = false;
    size_t offset = 0, bytes = 0;
    size_t arena_bytes = 0, meta_bytes = 0;
    uint32_t bypassed = 0;  // times a later request was admitted ahead of this one at the head
    uint32_t slot = 0;
    uint64_t fallback_rows = 0;
    Plan plan;  // laid out at submission, placed into the region at admission
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Request* req = nullptr;
    void* pinned = nullptr;  // [input | indices | distances | scalars | metadata]
    size_t pinned_bytes = 0;
};

// Byte layout of a request's device region and of its slot's pinned staging.
struct Layout {
    size_t input, indices, distances, arena, meta, total;
};

Layout region_layout(size_t n, size_t d, size_t k, size_t arena_bytes, size_t meta_bytes,
                     bool device_input) {
    Layout l{};
    l.input = 0;
    l.indices = l.input + (device_input ? 0 : round_region(n * d * sizeof(float)));
    l.distances = l.indices + round_region(n * k * sizeof(uint32_t));
    l.arena = l.distances + round_region(n * k * sizeof(float));
    l.meta = l.arena + round_region(arena_bytes);
    l.total = l.meta + round_region(meta_bytes);
    return l;
}

}  // namespace

}  // namespace scrust_hip

using namespace scrust_hip;

struct scrust_hip_server {
    int device = 0;
    int compute_units = 0;
    void* pool = nullptr;
    size_t pool_bytes = 0;
    RangeAllocator alloc;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    uint64_t submitted = 0, completed = 0, rejected = 0, in_flight = 0, peak_in_flight = 0;

    ~scrust_hip_server();
    void run();
    bool start(Slot& slot, Request& r);
    void finish(Slot& slot, Request& r);
};

scrust_hip_server::~scrust_hip_server() {
    if (worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv_work.notify_all();
        worker.join();
    }
    (void)hipSetDevice(device);
    for (Slot& s : slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        if (s.pinned) (void)hipHostFree(s.pinned);
    }
    if (pool) (void)hipFree(pool);
}

// Enqueues request r on the slot's stream. Returns false (and completes r with an error)
// when a HIP call fails.
bool scrust_hip_server::start(Slot& slot, Request& r) {
    auto failed = [&](scrust_hip_status st, const std::string& msg) {
        r.status = st;
        r.error = msg;
        return false;
    };
    const Layout l = region_layout(r.n, r.d, r.k, r.arena_bytes, r.meta_bytes, r.device_input);
    char* base = static_cast<char*>(pool) + r.offset;
    const size_t in_bytes = r.device_input ? 0 : r.n * r.d * sizeof(float);
    const size_t idx_bytes = r.n * r.k * sizeof(uint32_t);
    const size_t dist_bytes = r.n * r.k * sizeof(float);
    // Pinned staging: [input | indices | distances] at the region's offsets, then the
    // scalars, then the launch metadata.
    const size_t scalars_at = l.arena;
    const size_t meta_at = scalars_at + 256;
    const size_t need = meta_at + r.meta_bytes;
    if (slot.pinned_bytes < need) {
        if (slot.pinned) (void)hipHostFree(slot.pinned);
        slot.pinned = nullptr;
        slot.pinned_bytes = 0;
        const size_t grow = round_region(need * 5 / 4);
        if (hipHostMalloc(&slot.pinned, grow, hipHostMallocDefault) != hipSuccess)
            return failed(SCRUST_HIP_OUT_OF_MEMORY, "pinned staging allocation failed");
        slot.pinned_bytes = grow;
    }
    char* pin = static_cast<char*>(slot.pinned);
    if (in_bytes) std::memcpy(pin + l.input, r.x, in_bytes);
    // The plan was laid out at submission and sized this region; only the input and output
    // addresses inside the region are known now (plan_layout copies them, unread).
    DatasetDesc& desc = r.plan.sets[0];
    desc.x = r.device_input ? r.x : reinterpret_cast<const float*>(base + l.input);
    desc.out_indices = reinterpret_cast<uint32_t*>(base + l.indices);
    desc.out_distances = reinterpret_cast<float*>(base + l.distances);
    scrust_hip_status st = SCRUST_HIP_OK;
    hipError_t err = hipSuccess;
    if (in_bytes) err = hipMemcpyAsync(base + l.input, pin + l.input, in_bytes, hipMemcpyHostToDevice, slot.stream);
    if (err != hipSuccess) return failed(SCRUST_HIP_KERNEL, hipGetErrorString(err));
    st = plan_place(r.plan, base + l.arena, base + l.meta, slot.stream, pin + meta_at);
    if (st != SCRUST_HIP_OK) return failed(st, last_error());
    st = plan_enqueue(r.plan, true, compute_units, slot.stream, nullptr);
    if (st != SCRUST_HIP_OK) return failed(st, last_error());
    err = hipMemcpyAsync(pin + l.indices, base + l.indices, idx_bytes, hipMemcpyDeviceToHost, slot.stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(pin + l.distances, base + l.distances, dist_bytes, hipMemcpyDeviceToHost, slot.stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(pin + scalars_at, r.plan.d_scalars, kScalarCount * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, slot.stream);
    if (err == hipSuccess) err = hipEventRecord(slot.done, slot.stream);
    if (err != hipSuccess) return failed(SCRUST_HIP_KERNEL, hipGetErrorString(err));
    return true;
}

void scrust_hip_server::finish(Slot& slot, Request& r) {
    const Layout l = region_layout(r.n, r.d, r.k, r.arena_bytes, r.meta_bytes, r.device_input);
    const char* pin = static_cast<const char*>(slot.pinned);
    uint32_t scalars[kScalarCount];
    std::memcpy(scalars, pin + l.arena, sizeof scalars);
    if (scalars[kScalarFlags] & kFlagNonFinite) {
        r.status = SCRUST_HIP_INVALID_PARAMETER;
        r.error = "embedding must be finite values, got NaN or infinity";
    } else if (scalars[kScalarFlags] & kFlagBandOverflow) {
        r.status = SCRUST_HIP_KERNEL;
        r.error = "HIP kernel knn: a boundary tie band ran past the kept keys";
    } else {
        std::memcpy(r.out_indices, pin + l.indices, r.n * r.k * sizeof(uint32_t));
        std::memcpy(r.out_distances, pin + l.distances, r.n * r.k * sizeof(float));
    }
    r.fallback_rows = scalars[kScalarListCount];
}

void scrust_hip_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        bool busy = false;
        for (Slot& slot : slots) {
            if (!slot.req) continue;
            const hipError_t q = hipEventQuery(slot.done);
            if (q == hipErrorNotReady) {
                busy = true;
                continue;
            }
            Request& r = *slot.req;
            lk.unlock();
            if (q == hipSuccess) {
                finish(slot, r);
            } else {
                r.status = SCRUST_HIP_KERNEL;
                r.error = std::string("HIP kernel knn: ") + hipGetErrorString(q);
            }
            lk.lock();
            alloc.release(r.offset, r.bytes);
            r.done_ns = now_ns();
            r.done = true;
            slot.req = nullptr;
            completed++;
            in_flight--;
            progress = true;
            cv_done.notify_all();
        }
        // Admit while a slot is idle: the head request when its region fits; when it does
        // not, the first of the next kBypassWindow requests whose region does (a smaller build
        // need not wait behind a large one for memory). A head bypassed kBypassMax times
        // takes the next region that frees, so no request waits for ever.
        while (!queue.empty()) {
            Slot* idle = nullptr;
            for (Slot& slot : slots)
                if (!slot.req) {
                    idle = &slot;
                    break;
                }
            if (!idle) break;
            size_t pick = queue.size();
            size_t at = 0;
            if (alloc.alloc(queue.front()->bytes, &at)) {
                pick = 0;
            } else if (queue.front()->bypassed < kBypassMax) {
                const size_t window = std::min(queue.size(), kBypassWindow);
                for (size_t i = 1; i < window; i++)
                    if (alloc.alloc(queue[i]->bytes, &at)) {
                        pick = i;
                        queue.front()->bypassed++;
                        break;
                    }
            }
            if (pick == queue.size()) break;
            Request& r = *queue[pick];
            r.offset = at;
            queue.erase(queue.begin() + static_cast<std::ptrdiff_t>(pick));
            idle->req = &r;
            r.slot = static_cast<uint32_t>(idle - slots.data());
            r.admit_ns = now_ns();
            in_flight++;
            peak_in_flight = std::max(peak_in_flight, in_flight);
            lk.unlock();
            const bool ok = start(*idle, r);
            if (!ok) {
                // Copies and kernels start() queued before it failed may still read the
                // slot's staging and write the region: both are reused only after they end.
                const hipError_t drained = hipStreamSynchronize(idle->stream);
                if (drained != hipSuccess && r.status == SCRUST_HIP_OK) {
                    r.status = SCRUST_HIP_KERNEL;
                    r.error = std::string("HIP kernel knn: ") + hipGetErrorString(drained);
                }
            }
            lk.lock();
            if (!ok) {
                alloc.release(r.offset, r.bytes);
                r.done_ns = now_ns();
                r.done = true;
                idle->req = nullptr;
                completed++;
                in_flight--;
                cv_done.notify_all();
            } else {
                busy = true;
            }
            progress = true;
        }
        if (stop && queue.empty() && !busy && in_flight == 0) break;
        if (!progress) {
            if (busy) cv_work.wait_for(lk, std::chrono::microseconds(20));
            else cv_work.wait(lk);
        }
    }
}

extern "C" {

scrust_hip_status scrust_hip_server_create(int device, const scrust_hip_server_config* config,
                                           scrust_hip_server** out) {
    clear_error();
    if (!out) return invalid("out", "a non-null pointer", "null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return fail(SCRUST_HIP_NO_GPU, "no HIP device is available on this machine");
    hipDeviceProp_t props;
    if (device < 0 || device >= count || !device_is_gfx950(device, &props))
        return fail(SCRUST_HIP_NO_GPU, "HIP device " + std::to_string(device) + " is not a usable gfx950");
    const uint32_t n_slots = config && config->slots ? config->slots : 4u;
    if (n_slots > 64) return invalid("slots", "at most 64", std::to_string(n_slots));
    if (hipSetDevice(device) != hipSuccess) return fail(SCRUST_HIP_NO_GPU, "cannot select the HIP device");
    size_t free_b = 0, total_b = 0;
    hipError_t err = hipMemGetInfo(&free_b, &total_b);
    if (err != hipSuccess) return hip_failure(err, "query device memory");
    size_t budget = config && config->memory_bytes ? config->memory_bytes : free_b / 2;
    budget = round_region(budget);
    if (budget > free_b) return fail(SCRUST_HIP_OUT_OF_MEMORY, "the memory budget exceeds the free device memory");
    std::unique_ptr<scrust_hip_server> server(new (std::nothrow) scrust_hip_server());
    if (!server) return fail(SCRUST_HIP_OUT_OF_MEMORY, "host allocation failed");
    server->device = device;
    server->compute_units = props.multiProcessorCount;
    err = hipMalloc(&server->pool, budget);
    if (err != hipSuccess) return hip_failure(err, "allocate the server's device memory");
    server->pool_bytes = budget;
    server->alloc.reset(budget);
    server->slots.resize(n_slots);
    for (Slot& s : server->slots) {
        err = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (err != hipSuccess) return hip_failure(err, "create server streams");
    }
    scrust_hip_server* raw = server.get();
    server->worker = std::thread([raw] { raw->run(); });
    *out = server.release();
    return SCRUST_HIP_OK;
}

namespace {

scrust_hip_status submit_request(scrust_hip_server* server, const float* embedding, bool device_input,
                                 size_t n, size_t d, size_t k, const scrust_hip_knn_options* options,
                                 uint32_t* out_indices, float* out_distances, uint64_t* ticket) {
    clear_error();
    scrust_hip_status st = validate(embedding, n, d, k, out_indices, out_distances, options);
    if (st != SCRUST_HIP_OK) return st;
    if (!server) return fail(SCRUST_HIP_NO_GPU, "no server");
    if (!ticket) return invalid("ticket", "a non-null pointer", "null");
    auto r = std::make_unique<Request>();
    r->x = embedding;
    r->device_input = device_input;
    r->n = n;
    r->d = d;
    r->k = k;
    if (options) r->options = *options;
    r->out_indices = out_indices;
    r->out_distances = out_distances;
    // Lay the build out once, now: the work shape (a tuning-table lookup for an automatic
    // tile) and with it the region size are fixed at submission, so later changes to the
    // tuning table cannot outgrow the region admission hands out. Input and output
    // addresses are set at admission.
    std::vector<SetSpec> specs{SetSpec{embedding, n, d, k, nullptr, nullptr}};
    st = plan_layout(specs, options, r->plan, server->compute_units, true, nullptr, nullptr, true);
    if (st != SCRUST_HIP_OK) return st;
    r->arena_bytes = r->plan.arena_bytes;
    r->meta_bytes = r->plan.meta_bytes;
    r->bytes = region_layout(n, d, k, r->arena_bytes, r->meta_bytes, device_input).total;
    std::lock_guard<std::mutex> lk(server->mu);
    if (r->bytes > server->pool_bytes) {
        server->rejected++;
        return fail(SCRUST_HIP_OUT_OF_MEMORY,
                    "the build needs " + std::to_string(r->bytes) + " bytes of device memory, more than the server's budget of " +
                        std::to_string(server->pool_bytes));
    }
    r->ticket = server->next_ticket++;
    r->submit_ns = now_ns();
    *ticket = r->ticket;
    server->queue.push_back(r.get());
    server->requests.emplace(r->ticket, std::move(r));
    server->submitted++;
    server->cv_work.notify_all();
    return SCRUST_HIP_OK;
}

}  // namespace

scrust_hip_status scrust_hip_server_submit(scrust_hip_server* server, const float* embedding,
                                           size_t n, size_t d, size_t k,
                                           const scrust_hip_knn_options* options,
                                           uint32_t* out_indices, float* out_distances,
                                           uint64_t* ticket) {
    return submit_request(server, embedding, false, n, d, k, options, out_indices, out_distances, ticket);
}

scrust_hip_status scrust_hip_server_submit_device(scrust_hip_server* server, const float* embedding,
                                                  size_t n, size_t d, size_t k,
                                                  const scrust_hip_knn_options* options,
                                                  uint32_t* out_indices, float* out_distances,
                                                  uint64_t* ticket) {
    return submit_request(server, embedding, true, n, d, k, options, out_indices, out_distances, ticket);
}

scrust_hip_status scrust_hip_server_wait(scrust_hip_server* server, uint64_t ticket,
                                         int64_t timeout_ms, scrust_hip_request_info* info) {
    clear_error();
    if (!server) return fail(SCRUST_HIP_NO_GPU, "no server");
    std::unique_lock<std::mutex> lk(server->mu);
    auto it = server->requests.find(ticket);
    if (it == server->requests.end()) return invalid("ticket", "a pending ticket", std::to_string(ticket));
    Request& r = *it->second;
    auto ready = [&] { return r.done; };
    if (timeout_ms < 0) {
        server->cv_done.wait(lk, ready);
    } else if (!server->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return fail(SCRUST_HIP_KERNEL, "HIP kernel knn: request " + std::to_string(ticket) + " timed out");
    }
    if (info) {
        info->submit_ns = r.submit_ns;
        info->admit_ns = r.admit_ns;
        info->done_ns = r.done_ns;
        info->fallback_rows = r.fallback_rows;
        info->slot = r.slot;
        info->status = static_cast<uint32_t>(r.status);
    }
    const scrust_hip_status st = r.status;
    const std::string error = r.error;
    server->requests.erase(it);
    if (st != SCRUST_HIP_OK) return fail(st, error);
    return SCRUST_HIP_OK;
}

scrust_hip_status scrust_hip_server_stats_get(scrust_hip_server* server, scrust_hip_server_stats* stats) {
    clear_error();
    if (!server || !stats) return invalid("server", "a non-null pointer", "null");
    std::lock_guard<std::mutex> lk(server->mu);
    stats->submitted = server->submitted;
    stats->completed = server->completed;
    stats->rejected = server->rejected;
    stats->peak_in_flight = server->peak_in_flight;
    stats->memory_bytes = server->pool_bytes;
    stats->peak_memory_bytes = server->alloc.peak();
    return SCRUST_HIP_OK;
}

void scrust_hip_server_destroy(scrust_hip_server* server) { delete server; }

}  // extern "C"
