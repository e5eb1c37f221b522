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
    /* fp32 direct differences over every pair (no MFMA); the path for k > 239 or d > 256
     * (batches: k > 111 or d > 128). A bf16 or fp16 single call with 128 < d <= 256 takes
     * the fp32 pass. */
    SCRUST_HIP_EXACT = 2,
    /* fp16 single products of locally centred coordinates on MFMA (one origin per 512 query
     * rows of the spatial row order, fp32 accumulate, a rigorous per-pair error term folded
     * into the product) select the candidates; the same rerank and certificate follow. Single
     * builds whose rows are grouped (from 65,536 cells, or order = 2), k <= 47, d <= 125;
     * other shapes take the bf16 path. No counterpart in the reference. */
    SCRUST_HIP_FP16 = 3
} scrust_hip_precision;

typedef struct scrust_hip_context scrust_hip_context;

typedef struct scrust_hip_knn_options {
    uint32_t metric;    /* scrust_hip_metric */
    uint32_t precision; /* scrust_hip_precision */
    /* Candidate buffer of the pass, slots per query row: 0 = automatic (for a single build
     * the tuning table's work shape for its bucket when the table has one, else 256 slots
     * with the plan's own reference ranges), else 128, 256 or 512. Every setting gives the
     * same result. */
    uint32_t tile;
    /* Row order of the candidate pass: 0 = automatic (rows grouped by nearest pivot for
     * single embeddings of at least 65,536 cells whose query blocks each scan all references
     * in one range, so that each block meets its neighbours early), 1 = input order,
     * 2 = grouped. Every setting gives the same result. */
    uint32_t order;
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
    uint32_t tile;            /* candidate buffer slots per row used */
    uint32_t splits;          /* reference ranges per query block (0 on the exact path) */
    uint32_t order;           /* row order of the pass: 1 input, 2 grouped by nearest pivot */
    uint32_t tuned;           /* 1 when the work shape came from the tuning table */
    uint32_t tail_blocks;     /* query blocks of the launch's last round scanned in ranges */
    uint32_t tail_splits;     /* their reference ranges (0 when tail_blocks is 0) */
    uint64_t workspace_bytes; /* device workspace the call used (its arena and metadata) */
    uint32_t persistent;      /* 1: the pass ran one workgroup per slot with per-slot buffers */
    uint32_t reserved;
    uint64_t pairs_scored;    /* (query row, reference) pairs the pass scored, padding included:
                                 npad x ref_rows, less the tile-pair steps its ball bound skipped */
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

/* ---------------------------------------------------------------------------------------
 * PCA: the producer of the embedding the graph is built on. Replaces
 * scrust_core::pca::pca (crates/scrust-core/src/pca.rs:74-215; bound as _scrust.pca,
 * crates/scrust-py/src/preprocess.rs:154-185): truncated PCA of an n_rows x n_cols CSR matrix
 * (cells x genes; indptr has n_rows + 1 entries, indices and values indptr[n_rows]) by
 * randomised SVD with the reference's steps (sketch of n_components + 24 columns, 7 or 4 power
 * iterations, whitening orthonormalisation, final Gram eigensolve, sign convention, both
 * explained-variance statistics). Outputs: embedding n_rows x n_components (obsm["X_pca"]),
 * components n_components x n_cols, explained variance and its ratio (n_components each).
 * Errors: INVALID_PARAMETER for n_components 0 or above min(n_rows, n_cols), fewer than 2
 * rows (as the reference), n_components above 104 (this device path's sketch holds at most
 * 128 columns) and non-finite values; SHAPE for an inconsistent CSR matrix; KERNEL when a
 * Jacobi eigensolve does not converge. The sketch's normal draws are this library's own
 * (splitmix64 with Box-Muller), not the rand crate's: results agree with the reference to the
 * accuracy of a randomised SVD, not bit for bit.
 * ------------------------------------------------------------------------------------- */

/* Host CSR in, host outputs. */
scrust_hip_status scrust_hip_pca(scrust_hip_context* context, const uint32_t* indptr,
                                 const uint32_t* indices, const float* values, size_t n_rows,
                                 size_t n_cols, size_t n_components, int zero_center,
                                 uint64_t seed, float* out_embedding, float* out_components,
                                 float* out_explained_variance,
                                 float* out_explained_variance_ratio);

/* Device CSR in (device pointers on the context's device) and the embedding left in device
 * memory (out_embedding a device pointer: the input of scrust_hip_knn_device and
 * scrust_hip_neighbors_device, nothing crossing PCIe); components and variances in host
 * memory. No counterpart in the reference (its PCA returns host arrays). */
scrust_hip_status scrust_hip_pca_device(scrust_hip_context* context, const uint32_t* indptr,
                                        const uint32_t* indices, const float* values,
                                        size_t n_rows, size_t n_cols, size_t n_components,
                                        int zero_center, uint64_t seed, float* out_embedding,
                                        float* out_components, float* out_explained_variance,
                                        float* out_explained_variance_ratio);

/* Bytes of the device scratch array scrust_hip_knn_device_noalloc works in (256 MiB). */
#define SCRUST_HIP_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). As
 * scrust_hip_knn_device (same arguments, checks, results and errors), but the call's
 * workspace is a file-scope device array of SCRUST_HIP_NOALLOC_SCRATCH_BYTES that each
 * device holds until the process ends, so the call allocates nothing. When the default work
 * shape does not fit, the build takes the compact one (128-slot candidate buffer, one
 * reference range; the same graph); a build that fits in neither (past about 120,000 cells
 * at d = 50, k = 15) fails with INVALID_PARAMETER. Every call on a device shares that
 * device's array, which keeps the last call's contents: calls that overlap on one device
 * race. */
scrust_hip_status scrust_hip_knn_device_noalloc(scrust_hip_context* context,
                                                const float* embedding, size_t n, size_t d,
                                                size_t k, const scrust_hip_knn_options* options,
                                                uint32_t* out_indices, float* out_distances,
                                                scrust_hip_knn_stats* stats);

/* ---------------------------------------------------------------------------------------
 * Connectivities: UMAP's fuzzy simplicial set of a kNN graph, the weighted graph scanpy
 * stores in obsp["connectivities"]. Replaces scrust_core::neighbors::connectivities
 * (crates/scrust-core/src/neighbors.rs:208-233; bound as _scrust.connectivities,
 * crates/scrust-py/src/embedding.rs:79-100). indices / distances are the n x k kNN graph
 * (self excluded, as scrust_hip_knn returns it). The result is CSR: out_indptr (n + 1),
 * out_indices and out_values (capacity entries each; 2 n k always suffices), *out_nnz.
 * Errors: SHAPE for n == 0, k == 0, a neighbour id >= n or too small a capacity (as the
 * reference); INVALID_PARAMETER for a negative or non-finite distance (the reference does
 * not check). Weights agree with the reference's to its own audit bar (relative 1e-4, the
 * exponential being rounded from f64 here and taken from libm there).
 * ------------------------------------------------------------------------------------- */

/* Host buffers; synchronous. */
scrust_hip_status scrust_hip_connectivities(scrust_hip_context* context, const uint32_t* indices,
                                            const float* distances, size_t n, size_t k,
                                            uint32_t* out_indptr, uint32_t* out_indices,
                                            float* out_values, size_t capacity, size_t* out_nnz);

/* Device buffers on the context's device (e.g. the outputs of scrust_hip_knn_device);
 * returns after completion. */
scrust_hip_status scrust_hip_connectivities_device(scrust_hip_context* context,
                                                   const uint32_t* indices,
                                                   const float* distances, size_t n, size_t k,
                                                   uint32_t* out_indptr, uint32_t* out_indices,
                                                   float* out_values, size_t capacity,
                                                   size_t* out_nnz);

/* The device half of pp.neighbors (python/scrust/pp/_basics.py:186-215): the kNN graph and
 * its connectivities in one call, the graph kept on the device between the two steps. The
 * kNN outputs are scrust_hip_knn_ex's; the CSR is scrust_hip_connectivities'. The caller
 * forms obsp["distances"] from the kNN outputs (fixed-width rows: indptr = 0, k, 2k, ...). */
scrust_hip_status scrust_hip_neighbors(scrust_hip_context* context, const float* embedding,
                                      size_t n, size_t d, size_t k,
                                      const scrust_hip_knn_options* options,
                                      uint32_t* out_indices, float* out_distances,
                                      uint32_t* conn_indptr, uint32_t* conn_indices,
                                      float* conn_values, size_t capacity, size_t* conn_nnz);

/* scrust_hip_neighbors with every buffer on the context's device: an embedding already there
 * (the PCA output kept on the device; pp.neighbors reads obsm["X_pca"] at
 * python/scrust/pp/_basics.py:198) in, both graphs out, nothing crossing PCIe but the nnz.
 * Returns after completion. No counterpart in the reference (its arrays live in host memory). */
scrust_hip_status scrust_hip_neighbors_device(scrust_hip_context* context, const float* embedding,
                                             size_t n, size_t d, size_t k,
                                             const scrust_hip_knn_options* options,
                                             uint32_t* out_indices, float* out_distances,
                                             uint32_t* conn_indptr, uint32_t* conn_indices,
                                             float* conn_values, size_t capacity,
                                             size_t* conn_nnz);

/* ---------------------------------------------------------------------------------------
 * t-SNE input distances: the pairwise squared-distance matrix exact t-SNE starts from.
 * Replaces squared_distances(&centred(&points)) (crates/scrust-core/src/tsne.rs:86,
 * :173-188): column means subtracted in f32, then (|a|^2 + |b|^2) - 2 a.b clamped at 0, the
 * Gram block on the fp32 MFMA. out is n x n float32 row-major (the diagonal included).
 * The reference's expansion order is its gemm's, so entries agree with it to the
 * expansion's rounding bound, (2 d + 8) 2^-24 (|a|^2 + |b|^2), not bit for bit.
 * Errors: SHAPE for d == 0 (tsne.rs:120-122) or n == 0; INVALID_PARAMETER for
 * n > 20000, the exact formulation's cap (tsne.rs:40, :146-152), or non-finite input.
 * ------------------------------------------------------------------------------------- */

/* Host buffers; synchronous. */
scrust_hip_status scrust_hip_tsne_squared_distances(scrust_hip_context* context, const float* points,
                                                    size_t n, size_t d, float* out);

/* Device buffers on the context's device; returns after completion. */
scrust_hip_status scrust_hip_tsne_squared_distances_device(scrust_hip_context* context,
                                                           const float* points, size_t n, size_t d,
                                                           float* out);

/* ---------------------------------------------------------------------------------------
 * Autotuning of the candidate pass's work shape. No counterpart in the reference: its
 * Metal kernel's shape is fixed by threads_per_query (crates/scrust-gpu/src/kernels/knn.rs:
 * 116-125). The shape of a pass is its candidate buffer (128, 256 or 512 slots per row and
 * reference range), the reference ranges each 256-row query block's scan is cut into, and
 * the row order (input or grouped by nearest pivot). Every shape gives the same graph.
 * The tuner times the shapes on an embedding, checks that all of them return the same
 * indices and distances, and records the fastest in a process-wide table keyed by bucket:
 * path, metric, candidates kept (k <= 15: 32, k <= 47: 64, else 128), packed feature chunks (from d) and
 * n in half-octaves. Single builds and server requests with tile = 0 take the table's shape
 * for their bucket. The table is shared by every context and server of the process and can
 * be saved to and loaded from a text file, so tuning runs once per machine. Batches are
 * tuned as a whole instead (scrust_hip_batch_tune): their work items share one persistent
 * launch, so a shape that is fastest for a dataset alone is not the fastest inside one.
 * ------------------------------------------------------------------------------------- */

typedef struct scrust_hip_tune_result {
    uint32_t tile, splits, order;                         /* shape chosen */
    uint32_t default_tile, default_splits, default_order; /* shape of the untuned plan */
    uint32_t shapes;     /* shapes timed (0: the dataset is on the exact path, nothing to tune) */
    uint32_t reused;     /* 1: SCRUST_HIP_TUNE_SKIP_KNOWN and the bucket was already tuned */
    double best_ms;      /* best device time of the chosen shape (ms) */
    double default_ms;   /* best device time of the untuned shape (ms) */
    double worst_ms;     /* best device time of the slowest shape timed (ms) */
} scrust_hip_tune_result;

/* Leave a bucket that the table already holds as it is (no device work). */
#define SCRUST_HIP_TUNE_SKIP_KNOWN 1u

/* Tunes the bucket of an n x d embedding (device pointer on the context's device) for k
 * neighbours, with options->metric and options->precision (tile and order are what is
 * searched; NULL options: Euclidean bf16). Each shape runs once to warm up and check its
 * result, then three times; the best time counts. Synchronous. */
scrust_hip_status scrust_hip_tune_device(scrust_hip_context* context, const float* embedding,
                                         size_t n, size_t d, size_t k,
                                         const scrust_hip_knn_options* options, uint32_t flags,
                                         scrust_hip_tune_result* result);
/* As scrust_hip_tune_device with the embedding in host memory. */
scrust_hip_status scrust_hip_tune(scrust_hip_context* context, const float* embedding, size_t n,
                                  size_t d, size_t k, const scrust_hip_knn_options* options,
                                  uint32_t flags, scrust_hip_tune_result* result);
/* Writes the table as text (one bucket a line); INVALID_PARAMETER if the file cannot be
 * written. */
scrust_hip_status scrust_hip_tune_save(const char* path);
/* Merges a file written by scrust_hip_tune_save into the table (its buckets replace the
 * table's); INVALID_PARAMETER, with the line, for an unreadable file or a malformed line,
 * in which case the table is left unchanged. */
scrust_hip_status scrust_hip_tune_load(const char* path);
/* Empties the table: automatic shapes return to the plan's defaults. */
void scrust_hip_tune_clear(void);
/* Buckets in the table. */
size_t scrust_hip_tune_size(void);

/* ---------------------------------------------------------------------------------------
 * Batched builds: many embeddings of uneven n, d and k in one persistent MFMA launch,
 * the whole pipeline captured once in a HIP graph and replayed by scrust_hip_batch_run.
 * No counterpart in the reference (its kNN takes one embedding per call).
 * ------------------------------------------------------------------------------------- */

typedef struct scrust_hip_dataset {
    const float* embedding;  /* device pointer, n x d float32 row-major */
    size_t n, d, k;
    uint32_t* out_indices;   /* device pointer, n x k */
    float* out_distances;    /* device pointer, n x k */
} scrust_hip_dataset;

typedef struct scrust_hip_batch_stats {
    double device_ms;        /* graph launch to completion */
    uint64_t datasets;
    uint64_t pass_items;     /* (dataset, 256-row block) work items of the persistent pass */
    uint64_t exact_rows;     /* rows computed by the exact kernel (exact-path + uncertified) */
    uint64_t pass_flops;     /* sum over pass datasets of 2 * n * n * d */
    uint64_t cells;          /* sum of n */
} scrust_hip_batch_stats;

typedef struct scrust_hip_batch scrust_hip_batch;

/* Validates every dataset (as scrust_hip_knn does), allocates the batch's device workspace
 * and captures the pipeline in a HIP graph on the context's stream. The embedding and
 * output pointers are bound into the graph: their contents may change between runs, the
 * pointers may not. */
scrust_hip_status scrust_hip_batch_create(scrust_hip_context* context,
                                          const scrust_hip_dataset* datasets, size_t count,
                                          const scrust_hip_knn_options* options,
                                          scrust_hip_batch** out);
/* Replays the graph and waits for it. */
scrust_hip_status scrust_hip_batch_run(scrust_hip_batch* batch, scrust_hip_batch_stats* stats);

typedef struct scrust_hip_batch_tune_result {
    uint32_t cut, tile;                 /* shape chosen */
    uint32_t default_cut, default_tile; /* shape the batch was created with */
    uint32_t shapes;                    /* shapes timed */
    uint32_t reserved;
    double best_ms, default_ms;         /* best replay device time of each (ms) */
    double worst_ms;                    /* best replay device time of the slowest shape (ms) */
} scrust_hip_batch_tune_result;

/* Batch autotuner (no counterpart in the reference). Lays the batch out and captures it
 * again under each shape -- the longest reference range of one work item (1/cut of a
 * workgroup slot's share of the batch, cut in 1..8) and the candidate buffer of every
 * dataset (128, 256 or 512 slots, at least 4 x the candidates kept) -- replays each once
 * to warm up and check that every dataset's indices and distances equal the first shape's,
 * then three times, and keeps the fastest (the creation shape unless another is at least 2%
 * faster). Runs the batch's graph 4 times per shape, so the outputs are overwritten; the
 * inputs must hold valid embeddings. On an error the batch is built again with its
 * creation shape (if that fails too, scrust_hip_batch_run reports KERNEL). */
scrust_hip_status scrust_hip_batch_tune(scrust_hip_batch* batch, scrust_hip_batch_tune_result* result);
void scrust_hip_batch_destroy(scrust_hip_batch* batch);

/* ---------------------------------------------------------------------------------------
 * Serving: a request loop that admits kNN builds as device memory and slots free up and
 * overlaps admitted builds on the device (one HIP stream per slot). Requests are served in
 * arrival order; a request is admitted when a slot is idle and its workspace fits in the
 * server's device-memory budget, which the server allocates once at creation. No
 * counterpart in the reference.
 * ------------------------------------------------------------------------------------- */

typedef struct scrust_hip_server scrust_hip_server;

typedef struct scrust_hip_server_config {
    uint32_t slots;         /* concurrent builds on the device (0: 4) */
    uint32_t reserved;
    uint64_t memory_bytes;  /* device-memory budget (0: half of the free memory) */
} scrust_hip_server_config;

typedef struct scrust_hip_request_info {
    int64_t submit_ns;      /* steady-clock times of submission, admission, completion */
    int64_t admit_ns;
    int64_t done_ns;        /* results are in the caller's buffers */
    uint64_t fallback_rows; /* rows recomputed exactly */
    uint32_t slot;
    uint32_t status;        /* scrust_hip_status of the build */
} scrust_hip_request_info;

typedef struct scrust_hip_server_stats {
    uint64_t submitted, completed, rejected;
    uint64_t peak_in_flight;  /* most builds on the device at once */
    uint64_t memory_bytes;    /* the budget */
    uint64_t peak_memory_bytes;
} scrust_hip_server_stats;

scrust_hip_status scrust_hip_server_create(int device, const scrust_hip_server_config* config,
                                           scrust_hip_server** out);
/* Queues a build. `embedding` (n x d host memory) and the outputs (n x k host memory) must
 * stay valid until scrust_hip_server_wait returns for the ticket. Arguments are checked
 * here (as scrust_hip_knn does); a build larger than the whole budget is refused with
 * OUT_OF_MEMORY. */
scrust_hip_status scrust_hip_server_submit(scrust_hip_server* server, const float* embedding,
                                           size_t n, size_t d, size_t k,
                                           const scrust_hip_knn_options* options,
                                           uint32_t* out_indices, float* out_distances,
                                           uint64_t* ticket);
/* As scrust_hip_server_submit with a device-resident embedding (a pointer on the server's
 * device, e.g. a PCA result kept on the GPU): no staging and no upload; the results still
 * land in the host buffers. The embedding must stay valid until the wait returns. */
scrust_hip_status scrust_hip_server_submit_device(scrust_hip_server* server,
                                                  const float* embedding, size_t n, size_t d,
                                                  size_t k, const scrust_hip_knn_options* options,
                                                  uint32_t* out_indices, float* out_distances,
                                                  uint64_t* ticket);
/* Waits up to timeout_ms (< 0: forever) for a ticket; returns the build's status (KERNEL
 * with a "timed out" message when it is still running; the ticket stays valid). */
scrust_hip_status scrust_hip_server_wait(scrust_hip_server* server, uint64_t ticket,
                                         int64_t timeout_ms, scrust_hip_request_info* info);
scrust_hip_status scrust_hip_server_stats_get(scrust_hip_server* server,
                                              scrust_hip_server_stats* stats);
/* Finishes every queued and running build, then releases the server. */
void scrust_hip_server_destroy(scrust_hip_server* server);

#ifdef __cplusplus
}
#endif

#endif
