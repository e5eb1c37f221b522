/*
 * scrust_hip -- exact k-nearest-neighbour graph on AMD Instinct MI355X (gfx950).
 *
 * C ABI of the drop-in for the Metal kNN path of scrust. Every entry point returns a
 * scrust_hip_status; on failure scrust_hip_last_error() holds a message for the calling
 * thread. Nothing throws or aborts across this boundary. Outputs are caller-owned
 * buffers; the library owns all device memory it allocates.
 *
 * Output contract (crates/scrust-core/src/neighbors.rs:8-15): row i holds the k nearest
 * other cells of cell i as uint32 ids, nearest first, ties broken towards the smaller id,
 * and their float32 distances (Euclidean; cosine distance for SCRUST_HIP_COSINE).
 */
#ifndef SCRUST_HIP_H
#define SCRUST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors scrust_core::error::Error (crates/scrust-core/src/error.rs:8-33). The bindings
 * map SHAPE and INVALID_PARAMETER to ValueError and the rest to RuntimeError
 * (crates/scrust-py/src/lib.rs:24-35). */
typedef enum scrust_hip_status {
    SCRUST_HIP_OK = 0,
    SCRUST_HIP_SHAPE = 1,             /* Error::Shape */
    SCRUST_HIP_INVALID_PARAMETER = 2, /* Error::InvalidParameter */
    SCRUST_HIP_NO_GPU = 3,            /* Error::NoGpu */
    SCRUST_HIP_KERNEL = 4,            /* Error::Kernel: a launch or a device step failed */
    SCRUST_HIP_OUT_OF_MEMORY = 5      /* device allocation failed (folded into Kernel upstream) */
} scrust_hip_status;

typedef enum scrust_hip_metric {
    SCRUST_HIP_EUCLIDEAN = 0, /* the reference's only metric */
    SCRUST_HIP_COSINE = 1     /* 1 - cos, computed as |x/|x| - y/|y||^2 / 2; no centring */
} scrust_hip_metric;

typedef enum scrust_hip_precision {
    /* bf16 hi/lo split on MFMA (3 products, fp32 accumulate) selects k+margin candidates,
     * an fp32 direct-difference rerank orders them; rows whose candidate set cannot be
     * certified complete are recomputed exactly. */
    SCRUST_HIP_BF16 = 0,
    /* The same pipeline with an fp32 MFMA candidate pass: the plain FP32 kernel. */
    SCRUST_HIP_FP32 = 1,
    /* fp32 direct differences over every pair (no MFMA); the path for k > 47 or d > 128. */
    SCRUST_HIP_EXACT = 2
} scrust_hip_precision;

typedef struct scrust_hip_context scrust_hip_context;

typedef struct scrust_hip_knn_options {
    uint32_t metric;    /* scrust_hip_metric */
    uint32_t precision; /* scrust_hip_precision */
    uint32_t tile;      /* 0: autotuned or default tile shape; else a tile id (see DESIGN) */
    uint32_t reserved;
} scrust_hip_knn_options;

typedef struct scrust_hip_knn_stats {
    double prepare_ms;        /* centring / normalisation and operand packing */
    double pass_ms;           /* MFMA candidate pass (distance contraction + selection) */
    double rerank_ms;         /* fp32 direct-difference rerank + fp64 boundary rule */
    double fallback_ms;       /* exact recomputation of uncertified rows */
    double device_ms;         /* first to last device event of the call */
    uint64_t fallback_rows;   /* rows recomputed exactly */
    uint64_t candidates;      /* candidates kept per row by the pass (0 on the exact path) */
    uint64_t pass_flops;      /* 2 * rows * references * d: algorithmic flops of the pass */
    uint32_t path;            /* scrust_hip_precision actually used */
    uint32_t tile;            /* tile id used */
} scrust_hip_knn_stats;

/* Message for the last failure on the calling thread ("" if none). */
const char* scrust_hip_last_error(void);

/* 1 when a gfx950 device is usable. Replaces scrust_core::device::gpu_available
 * (crates/scrust-core/src/device.rs:39-41). */
int scrust_hip_gpu_available(void);

/* Replaces MetalContext::new (crates/scrust-gpu/src/context.rs:19-27): NO_GPU without a
 * usable device. A context owns one HIP stream and a grow-only device workspace; use it
 * from one thread at a time (the reference keeps one context per thread,
 * crates/scrust-py/src/embedding.rs:22-32). */
scrust_hip_status scrust_hip_context_create(int device, scrust_hip_context** out);
void scrust_hip_context_destroy(scrust_hip_context* context);

/* Replaces max_supported_k (crates/scrust-gpu/src/kernels/knn.rs:21-29). */
size_t scrust_hip_max_supported_k(const scrust_hip_context* context);

/* Replaces knn_metal (crates/scrust-gpu/src/kernels/knn.rs:31-114): embedding is n x d
 * float32 row-major host memory; out_indices / out_distances are n x k host buffers.
 * Euclidean, default precision. Synchronous. */
scrust_hip_status scrust_hip_knn(scrust_hip_context* context, const float* embedding, size_t n,
                                 size_t d, size_t k, uint32_t* out_indices,
                                 float* out_distances);

/* As scrust_hip_knn with options and optional stats (either may be NULL). */
scrust_hip_status scrust_hip_knn_ex(scrust_hip_context* context, const float* embedding,
                                    size_t n, size_t d, size_t k,
                                    const scrust_hip_knn_options* options,
                                    uint32_t* out_indices, float* out_distances,
                                    scrust_hip_knn_stats* stats);

/* Device-resident variant: embedding, out_indices and out_distances are device pointers
 * on the context's device. Work runs on the context's stream; the call returns after it
 * completes. No counterpart in the reference (its buffers are unified memory,
 * crates/scrust-gpu/src/context.rs:64-90). */
scrust_hip_status scrust_hip_knn_device(scrust_hip_context* context, const float* embedding,
                                        size_t n, size_t d, size_t k,
                                        const scrust_hip_knn_options* options,
                                        uint32_t* out_indices, float* out_distances,
                                        scrust_hip_knn_stats* stats);

#ifdef __cplusplus
}
#endif

#endif
