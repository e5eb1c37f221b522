/*
 * CPU oracle for the exact k-nearest-neighbour graph (test infrastructure only).
 *
 * This library is the checker for the HIP path, never part of it: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Semantics restated from the reference:
 *   - centring: crates/scrust-core/src/neighbors.rs:131-162 (f64 column means,
 *     one f32 rounding per coordinate) and its GPU twin
 *     crates/scrust-gpu/src/kernels/knn.rs:127-158 (f32 fma squared norms);
 *   - distances, snapping, order: the `cpu_knn` oracle the Metal kernel is held
 *     to, crates/scrust-gpu/src/kernels/knn.rs:277-313 (direct difference with
 *     fma in dimension order, snap below (d+2)*eps*(|a|^2+|b|^2), sort by
 *     (d^2, id)), with the selection of crates/scrust-core/src/neighbors.rs:171-191;
 *   - fp64 truth: tests/test_neighbors_audit.py:52-61;
 *   - CPU baseline (timed, not the oracle): crates/scrust-core/src/neighbors.rs:29-101.
 */
#ifndef SCRUST_KNN_ORACLE_H
#define SCRUST_KNN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_EUCLIDEAN = 0, ORACLE_COSINE = 1 };

/* Correctly rounded f64 value of the exact sum of n f32 values read with a stride. */
double oracle_exact_sum_f32(const float* values, size_t n, size_t stride);

/* Column means: exact sum of each column, rounded once to f64, divided by n in f64. */
void oracle_column_means(const float* x, size_t n, size_t d, double* means);

/* The coordinates every distance is computed from, plus their f32 fma squared norms.
 * Euclidean: centred columns. Cosine: rows scaled to unit f64 norm, not centred. */
void oracle_prepare(const float* x, size_t n, size_t d, int metric, float* coords,
                    float* norms);

/* Exact kNN under the oracle's semantics for the rows listed in `queries`
 * (all rows when queries == NULL). tie_rule != 0 applies the fp64 boundary
 * near-tie rule. Output rows follow the order of `queries`.
 * Returns 0, or -1 on invalid arguments. */
int oracle_knn(const float* x, size_t n, size_t d, size_t k, int metric, int tie_rule,
               const uint32_t* queries, size_t n_queries, int threads, uint32_t* out_indices,
               float* out_distances);

/* oracle_knn with the tie rule, vectorised across references (AVX-512; oracle_knn itself
 * on CPUs without it): the same arithmetic per pair and the same selection, so the same
 * output bit for bit, fast enough for whole graphs of 10^6 cells. */
int oracle_knn_fast(const float* x, size_t n, size_t d, size_t k, int metric,
                    const uint32_t* queries, size_t n_queries, int threads,
                    uint32_t* out_indices, float* out_distances);

/* Ground truth in f64 on the raw input (no centring): Euclidean distance, or
 * cosine distance 1 - x.y/(|x||y|). Ties by the smaller id. */
int oracle_knn_fp64(const float* x, size_t n, size_t d, size_t k, int metric,
                    const uint32_t* queries, size_t n_queries, int threads,
                    uint32_t* out_indices, double* out_distances);

/* The reference's CPU path, restated: f64-centred coordinates, the expansion
 * |a|^2+|b|^2-2a.b over tiles of tile_rows query rows, snapping at the
 * expansion's resolution, selection by (d^2, id). Used as the timed CPU baseline. */
int baseline_expansion_knn(const float* x, size_t n, size_t d, size_t k, size_t tile_rows,
                           const uint32_t* queries, size_t n_queries, int threads,
                           uint32_t* out_indices, float* out_distances);

/* Threads used when `threads` <= 0: every online host core. */
int oracle_default_threads(void);

/* ndarray 0.16.1's ArrayBase::sum over a contiguous f32 slice
 * (numeric_util::unrolled_fold: eight running sums, folded pairwise, then the tail). */
float oracle_ndarray_sum_f32(const float* x, size_t len);

/* UMAP's fuzzy simplicial set of a kNN graph, restating
 * crates/scrust-core/src/neighbors.rs:193-372 (connectivities, smooth_knn_distances,
 * directed_weights, symmetrise). indices/distances are n x k (self excluded). Writes the
 * CSR row pointers (n + 1), column ids and weights (capacity entries each; 2 n k always
 * suffices) and, when not NULL, sigmas and rhos (n each). Returns nnz, or -1 on invalid
 * arguments (an id >= n, n == 0, k == 0) or too small a capacity. */
long oracle_connectivities(const uint32_t* indices, const float* distances, size_t n, size_t k,
                           uint32_t* indptr, uint32_t* out_indices, float* out_values,
                           size_t capacity, float* sigmas, float* rhos);

/* The t-SNE input distances, squared_distances(&centred(&points))
 * (crates/scrust-core/src/tsne.rs:86, :173-188), for rows [row_begin, row_end): out holds
 * (row_end - row_begin) x n floats. Returns 0, or -1 on invalid arguments. */
int oracle_tsne_squared_distances(const float* x, size_t n, size_t d, size_t row_begin,
                                  size_t row_end, float* out, int threads);

/* PCA by randomised SVD restating crates/scrust-core/src/pca.rs:74-215 with f64 algebra:
 * CSR (indptr n + 1, indices, values) of n cells x g genes in; embedding n x n_components,
 * components n_components x g, explained variance and its ratio out. Returns 0, -1 for
 * invalid arguments (pca.rs:84-100), -2 when a Jacobi eigensolve does not converge. */
int oracle_pca(const uint32_t* indptr, const uint32_t* indices, const float* values, size_t n, size_t g,
               size_t n_components, int zero_center, uint64_t seed, float* embedding, float* components,
               float* explained_variance, float* explained_variance_ratio);

/* The sketch both the oracle and the device library draw: splitmix64 from `seed`, pairs of
 * 53-bit uniforms through Box-Muller in f64, rounded to f32; rows x cols row-major. */
void oracle_pca_sketch(size_t rows, size_t cols, uint64_t seed, float* out);

/* Cyclic Jacobi eigensolve of a symmetric n x n matrix (pca.rs:390-445); a is destroyed.
 * Returns 0, or -1 without convergence in 60 sweeps. */
int oracle_jacobi_eigen(double* a, size_t n, double* values, double* vectors);

#ifdef __cplusplus
}
#endif

#endif
