/*
 * CPU oracle for the exact k-nearest-neighbour graph -- TEST INFRASTRUCTURE.
 *
 * Nothing in the product links or calls this file: it is the checker the HIP
 * path is compared against (tests, smoke check) and the timed CPU baseline of
 * the benchmark. Every function cites the reference code it restates.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so that every fma below is
 * the explicit one and every other product is rounded on its own).
 */
#include "knn_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------ */
/* Exact summation of f32 values.                                           */
/*                                                                          */
/* The reference sums each column sequentially in f64                       */
/* (crates/scrust-core/src/neighbors.rs:146-154). A sequential f64 sum      */
/* cannot be reproduced by a parallel device reduction, so the oracle and   */
/* the device both take the exact sum and round it once to f64: every f32   */
/* is an integer multiple of 2^-149, so the sum is accumulated as a         */
/* two's-complement integer in 32-bit digits held in int64 words (carries   */
/* are resolved at the end). The result equals the sequential sum whenever  */
/* that sum is exact, and differs from it by at most its rounding error     */
/* otherwise.                                                               */
/* ------------------------------------------------------------------------ */

#define ACC_WORDS 10

typedef struct {
    int64_t w[ACC_WORDS];
} superacc;

static void acc_add(superacc* a, float v) {
    uint32_t bits;
    memcpy(&bits, &v, sizeof bits);
    uint32_t biased = (bits >> 23) & 0xFFu;
    uint32_t mantissa = bits & 0x7FFFFFu;
    uint64_t m;
    int position; /* bit position of the mantissa's lsb, in units of 2^-149 */
    if (biased == 0) {
        m = mantissa;
        position = 0;
    } else {
        m = mantissa | 0x800000u;
        position = (int)biased - 1;
    }
    if (m == 0) return;
    uint64_t shifted = m << (position & 31);
    int word = position >> 5;
    int64_t lo = (int64_t)(shifted & 0xFFFFFFFFu);
    int64_t hi = (int64_t)(shifted >> 32);
    if (bits >> 31) {
        a->w[word] -= lo;
        a->w[word + 1] -= hi;
    } else {
        a->w[word] += lo;
        a->w[word + 1] += hi;
    }
}

static double acc_to_double(const superacc* a) {
    uint32_t digit[ACC_WORDS];
    int64_t carry = 0;
    for (int i = 0; i < ACC_WORDS; i++) {
        int64_t v = a->w[i] + carry;
        digit[i] = (uint32_t)((uint64_t)v & 0xFFFFFFFFu);
        carry = v >> 32; /* arithmetic shift: floor division by 2^32 */
    }
    double sign = 1.0;
    if (carry < 0) { /* negative: take the two's complement magnitude */
        sign = -1.0;
        uint64_t c = 1;
        for (int i = 0; i < ACC_WORDS; i++) {
            uint64_t v = (uint64_t)(~digit[i]) + c;
            digit[i] = (uint32_t)v;
            c = v >> 32;
        }
    }
    int top = -1;
    for (int i = ACC_WORDS - 1; i >= 0; i--) {
        if (digit[i]) {
            top = i * 32 + 31 - __builtin_clz(digit[i]);
            break;
        }
    }
    if (top < 0) return 0.0;
#define BIT(b) ((b) < 0 ? 0u : (digit[(b) >> 5] >> ((b) & 31)) & 1u)
    if (top <= 52) {
        uint64_t v = 0;
        for (int b = top; b >= 0; b--) v = (v << 1) | BIT(b);
        return sign * ldexp((double)v, -149);
    }
    uint64_t m = 0;
    for (int b = top; b > top - 53; b--) m = (m << 1) | BIT(b);
    uint32_t guard = BIT(top - 53);
    uint32_t sticky = 0;
    for (int b = top - 54; b >= 0 && !sticky; b--) sticky |= BIT(b);
#undef BIT
    int exponent = top - 52 - 149;
    if (guard && (sticky || (m & 1u))) {
        m += 1;
        if (m == (1ull << 53)) {
            m >>= 1;
            exponent += 1;
        }
    }
    return sign * ldexp((double)m, exponent);
}

double oracle_exact_sum_f32(const float* values, size_t n, size_t stride) {
    superacc a;
    memset(&a, 0, sizeof a);
    for (size_t i = 0; i < n; i++) acc_add(&a, values[i * stride]);
    return acc_to_double(&a);
}

void oracle_column_means(const float* x, size_t n, size_t d, double* means) {
    for (size_t j = 0; j < d; j++) means[j] = oracle_exact_sum_f32(x + j, n, d) / (double)n;
}

/* Centring restates neighbors.rs:144-162 (value as f64 minus the f64 mean, rounded
 * once to f32); the squared norms restate knn.rs:148-156 (f32 fma in dimension
 * order). Cosine has no reference counterpart: rows are divided by their f64 norm
 * and rounded once; a zero row stays zero. */
void oracle_prepare(const float* x, size_t n, size_t d, int metric, float* coords,
                    float* norms) {
    if (metric == ORACLE_COSINE) {
        for (size_t r = 0; r < n; r++) {
            const float* row = x + r * d;
            double s = 0.0;
            for (size_t j = 0; j < d; j++) s = fma((double)row[j], (double)row[j], s);
            double length = sqrt(s);
            float nrm = 0.0f;
            for (size_t j = 0; j < d; j++) {
                float c = length > 0.0 ? (float)((double)row[j] / length) : 0.0f;
                coords[r * d + j] = c;
                nrm = fmaf(c, c, nrm);
            }
            norms[r] = nrm;
        }
        return;
    }
    double* means = (double*)malloc(d * sizeof(double));
    oracle_column_means(x, n, d, means);
    for (size_t r = 0; r < n; r++) {
        float nrm = 0.0f;
        for (size_t j = 0; j < d; j++) {
            float c = (float)((double)x[r * d + j] - means[j]);
            coords[r * d + j] = c;
            nrm = fmaf(c, c, nrm);
        }
        norms[r] = nrm;
    }
    free(means);
}

/* ------------------------------------------------------------------------ */
/* Parallel for over query rows.                                            */
/* ------------------------------------------------------------------------ */

int oracle_default_threads(void) {
    long c = sysconf(_SC_NPROCESSORS_ONLN);
    return c > 0 ? (int)c : 1;
}

typedef void (*row_fn)(void* ctx, size_t item, void* scratch);

typedef struct {
    row_fn fn;
    void* ctx;
    size_t n_items;
    size_t scratch_bytes;
    size_t next;
    pthread_mutex_t lock;
} pool_job;

static void* pool_worker(void* arg) {
    pool_job* job = (pool_job*)arg;
    void* scratch = job->scratch_bytes ? malloc(job->scratch_bytes) : NULL;
    for (;;) {
        pthread_mutex_lock(&job->lock);
        size_t item = job->next++;
        pthread_mutex_unlock(&job->lock);
        if (item >= job->n_items) break;
        job->fn(job->ctx, item, scratch);
    }
    free(scratch);
    return NULL;
}

static void parallel_for(row_fn fn, void* ctx, size_t n_items, size_t scratch_bytes,
                         int threads) {
    if (threads <= 0) threads = oracle_default_threads();
    if ((size_t)threads > n_items) threads = (int)(n_items ? n_items : 1);
    pool_job job = {fn, ctx, n_items, scratch_bytes, 0, PTHREAD_MUTEX_INITIALIZER};
    if (threads == 1) {
        pool_worker(&job);
        return;
    }
    pthread_t* ids = (pthread_t*)malloc((size_t)threads * sizeof(pthread_t));
    for (int t = 0; t < threads; t++) pthread_create(&ids[t], NULL, pool_worker, &job);
    for (int t = 0; t < threads; t++) pthread_join(ids[t], NULL);
    free(ids);
}

/* ------------------------------------------------------------------------ */
/* Keys: (squared distance, id) as one u64, ordered as (d^2, id) for d^2>=0. */
/* ------------------------------------------------------------------------ */

static inline uint32_t f32_bits(float f) {
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

static inline float bits_f32(uint32_t u) {
    float f;
    memcpy(&f, &u, sizeof f);
    return f;
}

/* Monotone map of f32 onto u32 (total order as f32::total_cmp). */
static inline uint32_t ordered(float f) {
    uint32_t u = f32_bits(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

static inline float unordered(uint32_t o) {
    return bits_f32((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

static inline uint64_t make_key(float d2, uint32_t id) {
    return ((uint64_t)ordered(d2) << 32) | id;
}

/* Keep the `keep` smallest keys of `keys[0..n)` in ascending order in `best`;
 * returns how many were kept. */
static size_t smallest_keys(const uint64_t* keys, size_t n, size_t keep, uint64_t* best) {
    size_t count = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t key = keys[i];
        if (count == keep && key >= best[keep - 1]) continue;
        size_t slot = count < keep ? count++ : keep - 1;
        while (slot > 0 && best[slot - 1] > key) {
            best[slot] = best[slot - 1];
            slot--;
        }
        best[slot] = key;
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* The oracle.                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float* coords;
    const float* norms;
    size_t n, d, k;
    int metric, tie_rule;
    const uint32_t* queries;
    uint32_t* out_indices;
    float* out_distances;
} oracle_ctx;

/* Squared distance as the Metal kernel and `cpu_knn` form it: f32 difference,
 * f32 fma accumulation in dimension order (knn.rs:229-232, :293-297), then the
 * snap below (d+2)*eps*(|a|^2+|b|^2) evaluated in f32 left to right
 * (knn.rs:237-240, :298-302). */
static inline float direct_square(const float* a, const float* b, size_t d, float na,
                                  float nb) {
    float s = 0.0f;
    for (size_t j = 0; j < d; j++) {
        float delta = a[j] - b[j];
        s = fmaf(delta, delta, s);
    }
    float threshold = ((float)d + 2.0f) * FLT_EPSILON * (na + nb);
    return s < threshold ? 0.0f : s;
}

/* The fp64 recomputation used inside a boundary tie band: f64 differences of the
 * same f32 coordinates, f64 fma in dimension order. A pair the f32 pass snapped to
 * zero stays at zero. */
static inline double direct_square_f64(const float* a, const float* b, size_t d) {
    double s = 0.0;
    for (size_t j = 0; j < d; j++) {
        double delta = (double)a[j] - (double)b[j];
        s = fma(delta, delta, s);
    }
    return s;
}

typedef struct {
    double d64;
    uint32_t id;
    float d2;
} band_entry;

static int band_cmp(const void* lhs, const void* rhs) {
    const band_entry* a = (const band_entry*)lhs;
    const band_entry* b = (const band_entry*)rhs;
    if (a->d64 < b->d64) return -1;
    if (a->d64 > b->d64) return 1;
    return a->id < b->id ? -1 : (a->id > b->id ? 1 : 0);
}

/* The fp64 boundary near-tie rule (no reference counterpart; shared verbatim in
 * meaning with the device rerank). With the sorted keys s_0 <= s_1 <= ... and
 * t = d2(s_{k-1}), tau = t*(d+2)*2^-23: if d2(s_k) - t <= tau, the band is every
 * position j with |d2(s_j) - t| <= tau; positions before the band stay, and the
 * band's slots up to k-1 take the band members in (fp64 d^2, id) order. A band of
 * more than ORACLE_BAND_CAP members (a near-equidistant crowd, e.g. every cell seen
 * from a zero row under cosine) keeps the (d^2, id) order.
 * Returns 1 when the kept keys may not hold the whole band. */
#define ORACLE_BAND_CAP 64
static int apply_tie_rule(const uint64_t* best, size_t kept, size_t k, size_t n_candidates,
                          const float* coords, size_t d, const float* q, uint32_t* ids,
                          float* d2s) {
    double t = (double)unordered((uint32_t)(best[k - 1] >> 32));
    double tau = t * (double)(d + 2) * 0x1p-23;
    if (kept <= k) return 0;
    double next = (double)unordered((uint32_t)(best[k] >> 32));
    /* Written as "not within tau": with t = +inf the gap inf - inf is NaN, and a tie of
     * infinities is left in (d^2, id) order, as on the device. */
    if (!(next - t <= tau)) return 0;
    size_t lo = k - 1;
    while (lo > 0 && t - (double)unordered((uint32_t)(best[lo - 1] >> 32)) <= tau) lo--;
    size_t hi = k;
    while (hi + 1 < kept && (double)unordered((uint32_t)(best[hi + 1] >> 32)) - t <= tau) hi++;
    if (hi - lo + 1 > ORACLE_BAND_CAP) return 0; /* too wide to resolve: (d^2, id) order */
    if (hi + 1 == kept && kept < n_candidates) return 1;
    size_t m = hi - lo + 1;
    band_entry* band = (band_entry*)malloc(m * sizeof(band_entry));
    for (size_t j = 0; j < m; j++) {
        uint32_t id = (uint32_t)best[lo + j];
        float d2 = unordered((uint32_t)(best[lo + j] >> 32));
        band[j].id = id;
        band[j].d2 = d2;
        band[j].d64 = d2 == 0.0f ? 0.0 : direct_square_f64(q, coords + (size_t)id * d, d);
    }
    qsort(band, m, sizeof(band_entry), band_cmp);
    for (size_t j = lo; j < k; j++) {
        ids[j] = band[j - lo].id;
        d2s[j] = band[j - lo].d2;
    }
    free(band);
    return 0;
}

static void oracle_row(void* vctx, size_t item, void* scratch) {
    oracle_ctx* c = (oracle_ctx*)vctx;
    size_t q = c->queries ? c->queries[item] : item;
    size_t n = c->n, d = c->d, k = c->k;
    uint64_t* keys = (uint64_t*)scratch;
    const float* qrow = c->coords + q * d;
    size_t m = 0;
    for (size_t r = 0; r < n; r++) {
        if (r == q) continue; /* a cell is not its own neighbour (knn.rs:224-226) */
        float s = direct_square(qrow, c->coords + r * d, d, c->norms[q], c->norms[r]);
        keys[m++] = make_key(s, (uint32_t)r);
    }
    uint32_t ids[4096];
    float d2s[4096];
    size_t keep = k + 32;
    for (;;) {
        if (keep > m) keep = m;
        uint64_t* best = (uint64_t*)malloc(keep * sizeof(uint64_t));
        size_t kept = smallest_keys(keys, m, keep, best);
        for (size_t j = 0; j < k; j++) {
            ids[j] = (uint32_t)best[j];
            d2s[j] = unordered((uint32_t)(best[j] >> 32));
        }
        int again = c->tie_rule
                        ? apply_tie_rule(best, kept, k, m, c->coords, d, qrow, ids, d2s)
                        : 0;
        free(best);
        if (!again) break;
        keep *= 2;
    }
    for (size_t j = 0; j < k; j++) {
        c->out_indices[item * k + j] = ids[j];
        c->out_distances[item * k + j] =
            c->metric == ORACLE_COSINE ? 0.5f * d2s[j] : sqrtf(d2s[j]);
    }
}

int oracle_knn(const float* x, size_t n, size_t d, size_t k, int metric, int tie_rule,
               const uint32_t* queries, size_t n_queries, int threads, uint32_t* out_indices,
               float* out_distances) {
    if (n == 0 || d == 0 || k == 0 || k >= n || k > 4096) return -1;
    if (!queries) n_queries = n;
    float* coords = (float*)malloc(n * d * sizeof(float));
    float* norms = (float*)malloc(n * sizeof(float));
    oracle_prepare(x, n, d, metric, coords, norms);
    oracle_ctx c = {coords, norms, n, d, k, metric, tie_rule, queries, out_indices,
                    out_distances};
    parallel_for(oracle_row, &c, n_queries, n * sizeof(uint64_t), threads);
    free(coords);
    free(norms);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The same oracle, vectorised across references (whole graphs at 10^6).    */
/*                                                                          */
/* oracle_knn_fast computes exactly what oracle_knn (tie rule on) computes: */
/* each pair's squared distance is the same f32 difference and f32 fma      */
/* chain in dimension order and the same snap (direct_square above, i.e.    */
/* knn.rs:229-240); only the pairs run side by side, 16 references to an    */
/* AVX-512 vector and 16 query rows per register block, over references      */
/* stored dimension-major in blocks of 16. Per row it keeps the k + 32      */
/* smallest keys exactly as smallest_keys does (references arrive in id     */
/* order, so a key whose d^2 is not below the kept maximum is never kept)   */
/* and finishes through the same tie rule; a row whose tie band needs more  */
/* keys than were kept is recomputed by oracle_row. Without AVX-512 it is   */
/* oracle_knn.                                                              */
/* ------------------------------------------------------------------------ */

#if defined(__x86_64__) && defined(__GNUC__)
#include <immintrin.h>
#define FAST_W 16
#define FAST_Q 16

typedef struct {
    const float* coords;
    const float* norms;
    const float* tblock; /* n_blocks x d x 16, references dimension-major per block */
    const float* tnorm;  /* n_blocks x 16, +inf past n */
    size_t n, d, k, n_queries, n_blocks;
    int metric;
    const uint32_t* queries;
    uint32_t* out_indices;
    float* out_distances;
} fast_ctx;

/* Insert `key` into the ascending list best[0..*count) of at most `keep` keys. */
static inline void keep_key(uint64_t* best, size_t* count, size_t keep, uint64_t key) {
    if (*count == keep && key >= best[keep - 1]) return;
    size_t slot = *count < keep ? (*count)++ : keep - 1;
    while (slot > 0 && best[slot - 1] > key) {
        best[slot] = best[slot - 1];
        slot--;
    }
    best[slot] = key;
}

__attribute__((target("avx512f"))) static void fast_rows(void* vctx, size_t item, void* scratch) {
    fast_ctx* c = (fast_ctx*)vctx;
    const size_t d = c->d, k = c->k, n = c->n;
    const size_t keep = k + 32 < n - 1 ? k + 32 : n - 1;
    const size_t first = item * FAST_Q;
    size_t rows = c->n_queries - first < FAST_Q ? c->n_queries - first : FAST_Q;
    size_t q[FAST_Q];
    float qn[FAST_Q];
    for (size_t i = 0; i < FAST_Q; i++) {
        size_t at = first + (i < rows ? i : 0);
        q[i] = c->queries ? c->queries[at] : at;
        qn[i] = c->norms[q[i]];
    }
    uint64_t* best = (uint64_t*)scratch; /* FAST_Q x keep */
    size_t count[FAST_Q] = {0};
    float worst[FAST_Q];
    for (size_t i = 0; i < FAST_Q; i++) worst[i] = INFINITY;
    const float scale = ((float)d + 2.0f) * FLT_EPSILON; /* exact: an integer times 2^-23 */
    const __m512 vscale = _mm512_set1_ps(scale);
    const __m512i lane = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    for (size_t b = 0; b < c->n_blocks; b++) {
        const float* t = c->tblock + b * d * FAST_W;
        __m512 acc[FAST_Q];
        for (size_t i = 0; i < FAST_Q; i++) acc[i] = _mm512_setzero_ps();
        for (size_t j = 0; j < d; j++) {
            const __m512 r = _mm512_loadu_ps(t + j * FAST_W);
            for (size_t i = 0; i < FAST_Q; i++) {
                const __m512 delta = _mm512_sub_ps(_mm512_set1_ps(c->coords[q[i] * d + j]), r);
                acc[i] = _mm512_fmadd_ps(delta, delta, acc[i]);
            }
        }
        const __m512 rn = _mm512_loadu_ps(c->tnorm + b * FAST_W);
        const __m512i ids = _mm512_add_epi32(_mm512_set1_epi32((int)(b * FAST_W)), lane);
        for (size_t i = 0; i < rows; i++) {
            /* snap: s < ((d + 2) eps) * (|a|^2 + |b|^2) -> 0, left to right in f32 */
            const __m512 th = _mm512_mul_ps(vscale, _mm512_add_ps(_mm512_set1_ps(qn[i]), rn));
            const __mmask16 snap = _mm512_cmp_ps_mask(acc[i], th, _CMP_LT_OQ);
            const __m512 v = _mm512_mask_blend_ps(snap, acc[i], _mm512_setzero_ps());
            __mmask16 m = count[i] < keep ? (__mmask16)0xFFFF
                                          : _mm512_cmp_ps_mask(v, _mm512_set1_ps(worst[i]), _CMP_LT_OQ);
            m &= (__mmask16)(b * FAST_W + FAST_W <= n ? 0xFFFF : ((1u << (n - b * FAST_W)) - 1u));
            m &= (__mmask16)~_mm512_cmpeq_epi32_mask(ids, _mm512_set1_epi32((int)q[i]));
            if (!m) continue;
            float vals[FAST_W];
            _mm512_storeu_ps(vals, v);
            while (m) {
                const int l = __builtin_ctz(m);
                m &= (__mmask16)(m - 1);
                keep_key(best + i * keep, &count[i], keep, make_key(vals[l], (uint32_t)(b * FAST_W + l)));
            }
            if (count[i] == keep) worst[i] = unordered((uint32_t)(best[i * keep + keep - 1] >> 32));
        }
    }
    for (size_t i = 0; i < rows; i++) {
        const size_t out = first + i;
        const uint64_t* mine = best + i * keep;
        uint32_t ids[4096];
        float d2s[4096];
        for (size_t j = 0; j < k; j++) {
            ids[j] = (uint32_t)mine[j];
            d2s[j] = unordered((uint32_t)(mine[j] >> 32));
        }
        if (apply_tie_rule(mine, count[i], k, n - 1, c->coords, d, c->coords + q[i] * d, ids, d2s)) {
            /* The band may run past the kept keys: the scalar oracle for this row. */
            oracle_ctx oc = {c->coords, c->norms, n, d, k, c->metric, 1, c->queries,
                             c->out_indices, c->out_distances};
            uint64_t* keys = (uint64_t*)malloc(n * sizeof(uint64_t));
            oracle_row(&oc, out, keys);
            free(keys);
            continue;
        }
        for (size_t j = 0; j < k; j++) {
            c->out_indices[out * k + j] = ids[j];
            c->out_distances[out * k + j] = c->metric == ORACLE_COSINE ? 0.5f * d2s[j] : sqrtf(d2s[j]);
        }
    }
}
#endif

int oracle_knn_fast(const float* x, size_t n, size_t d, size_t k, int metric,
                    const uint32_t* queries, size_t n_queries, int threads,
                    uint32_t* out_indices, float* out_distances) {
#if defined(__x86_64__) && defined(__GNUC__)
    if (n == 0 || d == 0 || k == 0 || k >= n || k > 4096) return -1;
    if (!__builtin_cpu_supports("avx512f"))
        return oracle_knn(x, n, d, k, metric, 1, queries, n_queries, threads, out_indices, out_distances);
    if (!queries) n_queries = n;
    float* coords = (float*)malloc(n * d * sizeof(float));
    float* norms = (float*)malloc(n * sizeof(float));
    oracle_prepare(x, n, d, metric, coords, norms);
    const size_t n_blocks = (n + FAST_W - 1) / FAST_W;
    float* tblock = (float*)calloc(n_blocks * d * FAST_W, sizeof(float));
    float* tnorm = (float*)malloc(n_blocks * FAST_W * sizeof(float));
    for (size_t r = 0; r < n_blocks * FAST_W; r++) {
        tnorm[r] = r < n ? norms[r] : INFINITY;
        if (r < n)
            for (size_t j = 0; j < d; j++) tblock[(r / FAST_W) * d * FAST_W + j * FAST_W + r % FAST_W] = coords[r * d + j];
    }
    fast_ctx c = {coords, norms, tblock, tnorm, n, d, k, n_queries, n_blocks, metric,
                  queries, out_indices, out_distances};
    const size_t keep = k + 32 < n - 1 ? k + 32 : n - 1;
    parallel_for(fast_rows, &c, (n_queries + FAST_Q - 1) / FAST_Q, FAST_Q * keep * sizeof(uint64_t), threads);
    free(tblock);
    free(tnorm);
    free(coords);
    free(norms);
    return 0;
#else
    return oracle_knn(x, n, d, k, metric, 1, queries, n_queries, threads, out_indices, out_distances);
#endif
}

/* ------------------------------------------------------------------------ */
/* fp64 ground truth (tests/test_neighbors_audit.py:52-61).                 */
/* ------------------------------------------------------------------------ */

typedef struct {
    const double* x;
    size_t n, d, k;
    int metric;
    const uint32_t* queries;
    uint32_t* out_indices;
    double* out_distances;
} fp64_ctx;

typedef struct {
    double v;
    uint32_t id;
} fp64_key;

static void fp64_row(void* vctx, size_t item, void* scratch) {
    fp64_ctx* c = (fp64_ctx*)vctx;
    size_t q = c->queries ? c->queries[item] : item;
    size_t n = c->n, d = c->d, k = c->k;
    fp64_key* best = (fp64_key*)scratch;
    size_t count = 0;
    const double* a = c->x + q * d;
    for (size_t r = 0; r < n; r++) {
        if (r == q) continue;
        const double* b = c->x + r * d;
        double v;
        if (c->metric == ORACLE_COSINE) {
            double dot = 0.0;
            for (size_t j = 0; j < d; j++) dot += a[j] * b[j];
            v = 1.0 - dot; /* rows were scaled to unit norm */
        } else {
            v = 0.0;
            for (size_t j = 0; j < d; j++) {
                double delta = a[j] - b[j];
                v += delta * delta;
            }
        }
        if (count == k && (v > best[k - 1].v || (v == best[k - 1].v && r > best[k - 1].id)))
            continue;
        size_t slot = count < k ? count++ : k - 1;
        while (slot > 0 && (best[slot - 1].v > v ||
                            (best[slot - 1].v == v && best[slot - 1].id > (uint32_t)r))) {
            best[slot] = best[slot - 1];
            slot--;
        }
        best[slot].v = v;
        best[slot].id = (uint32_t)r;
    }
    for (size_t j = 0; j < k; j++) {
        c->out_indices[item * k + j] = best[j].id;
        c->out_distances[item * k + j] =
            c->metric == ORACLE_COSINE ? best[j].v : sqrt(best[j].v > 0 ? best[j].v : 0.0);
    }
}

int oracle_knn_fp64(const float* x, size_t n, size_t d, size_t k, int metric,
                    const uint32_t* queries, size_t n_queries, int threads,
                    uint32_t* out_indices, double* out_distances) {
    if (n == 0 || d == 0 || k == 0 || k >= n) return -1;
    if (!queries) n_queries = n;
    double* x64 = (double*)malloc(n * d * sizeof(double));
    for (size_t r = 0; r < n; r++) {
        double s = 0.0;
        for (size_t j = 0; j < d; j++) {
            x64[r * d + j] = (double)x[r * d + j];
            s += x64[r * d + j] * x64[r * d + j];
        }
        if (metric == ORACLE_COSINE) {
            double length = sqrt(s);
            for (size_t j = 0; j < d; j++) x64[r * d + j] = length > 0 ? x64[r * d + j] / length : 0.0;
        }
    }
    fp64_ctx c = {x64, n, d, k, metric, queries, out_indices, out_distances};
    parallel_for(fp64_row, &c, n_queries, k * sizeof(fp64_key), threads);
    free(x64);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* The reference CPU path restated (timed baseline, not the oracle).        */
/* crates/scrust-core/src/neighbors.rs:29-101,127-129,144-191               */
/* ------------------------------------------------------------------------ */

#define BASE_QB 8    /* query rows that share one pass over the references */
#define BASE_CH 256  /* references per chunk */

typedef struct {
    const float* points;   /* centred, row-major n x d */
    const float* points_t; /* centred, d x n (the reference's points_t) */
    const float* norms;
    size_t n, d, k, tile_rows;
    const uint32_t* queries;
    size_t n_queries;
    uint32_t* out_indices;
    float* out_distances;
} base_ctx;

/* One tile of query rows (neighbors.rs:69-97): raw = (tile . P^T) * -2 + (|a|^2 + |b|^2);
 * floor = (|a|^2 + |b|^2) * (d+2)*eps; raw < floor -> 0; then per row the k smallest
 * by (max(d^2, 0), id) excluding the row itself, and sqrt. The tile is consumed in
 * chunks instead of being materialised whole; the values and the selection are the
 * same. */
static void base_tile(void* vctx, size_t tile, void* scratch) {
    base_ctx* c = (base_ctx*)vctx;
    size_t n = c->n, d = c->d, k = c->k;
    size_t begin = tile * c->tile_rows;
    size_t end = begin + c->tile_rows;
    if (end > c->n_queries) end = c->n_queries;
    uint64_t* best = (uint64_t*)scratch; /* BASE_QB x k */
    float resolution = ((float)d + 2.0f) * FLT_EPSILON; /* neighbors.rs:127-129 */
    for (size_t q0 = begin; q0 < end; q0 += BASE_QB) {
        size_t qb = end - q0 < BASE_QB ? end - q0 : BASE_QB;
        size_t rows[BASE_QB];
        float qv[BASE_QB][512];
        size_t count[BASE_QB];
        for (size_t i = 0; i < qb; i++) {
            rows[i] = c->queries ? c->queries[q0 + i] : q0 + i;
            for (size_t j = 0; j < d; j++) qv[i][j] = c->points[rows[i] * d + j];
            count[i] = 0;
        }
        for (size_t r0 = 0; r0 < n; r0 += BASE_CH) {
            size_t ch = n - r0 < BASE_CH ? n - r0 : BASE_CH;
            float acc[BASE_QB][BASE_CH];
            memset(acc, 0, sizeof acc);
            for (size_t j = 0; j < d; j++) {
                const float* prow = c->points_t + j * n + r0;
                for (size_t i = 0; i < qb; i++) {
                    float v = qv[i][j];
                    float* a = acc[i];
                    for (size_t r = 0; r < ch; r++) a[r] += v * prow[r];
                }
            }
            for (size_t i = 0; i < qb; i++) {
                float nq = c->norms[rows[i]];
                uint64_t* list = best + i * k;
                for (size_t r = 0; r < ch; r++) {
                    size_t col = r0 + r;
                    if (col == rows[i]) continue;
                    float sums = nq + c->norms[col];
                    float raw = acc[i][r] * -2.0f + 0.0f;
                    raw = raw + sums;
                    float floor_ = sums * resolution + 0.0f;
                    float sq = raw < floor_ ? 0.0f : raw;
                    if (!(sq > 0.0f)) sq = 0.0f; /* square.max(0.0) */
                    uint64_t key = make_key(sq, (uint32_t)col);
                    if (count[i] == k && key >= list[k - 1]) continue;
                    size_t slot = count[i] < k ? count[i]++ : k - 1;
                    while (slot > 0 && list[slot - 1] > key) {
                        list[slot] = list[slot - 1];
                        slot--;
                    }
                    list[slot] = key;
                }
            }
        }
        for (size_t i = 0; i < qb; i++) {
            for (size_t j = 0; j < k; j++) {
                uint64_t key = best[i * k + j];
                c->out_indices[(q0 + i) * k + j] = (uint32_t)key;
                c->out_distances[(q0 + i) * k + j] = sqrtf(unordered((uint32_t)(key >> 32)));
            }
        }
    }
}

int baseline_expansion_knn(const float* x, size_t n, size_t d, size_t k, size_t tile_rows,
                           const uint32_t* queries, size_t n_queries, int threads,
                           uint32_t* out_indices, float* out_distances) {
    if (n == 0 || d == 0 || k == 0 || k >= n || d > 512) return -1;
    if (!queries) n_queries = n;
    if (tile_rows == 0) {
        /* neighbors.rs:17-31: tiles of MAX_TILE_ELEMENTS / n rows */
        tile_rows = (64u * 1024u * 1024u) / n;
        if (tile_rows == 0) tile_rows = 1;
    }
    /* Parallel work is a slice of a tile so every thread has rows to select in, as
     * the reference's rayon iterator does within one tile. */
    size_t slice = tile_rows;
    int t = threads > 0 ? threads : oracle_default_threads();
    if (slice > (n_queries + (size_t)t - 1) / (size_t)t)
        slice = (n_queries + (size_t)t - 1) / (size_t)t;
    if (slice < BASE_QB) slice = BASE_QB;
    /* Centring: sequential f64 sums as in neighbors.rs:144-162. */
    double* means = (double*)calloc(d, sizeof(double));
    for (size_t r = 0; r < n; r++)
        for (size_t j = 0; j < d; j++) means[j] += (double)x[r * d + j];
    for (size_t j = 0; j < d; j++) means[j] /= (double)n;
    float* points = (float*)malloc(n * d * sizeof(float));
    float* points_t = (float*)malloc(n * d * sizeof(float));
    float* norms = (float*)malloc(n * sizeof(float));
    for (size_t r = 0; r < n; r++) {
        float s = 0.0f; /* points.sqr().sum_keepdim(1) */
        for (size_t j = 0; j < d; j++) {
            float c = (float)((double)x[r * d + j] - means[j]);
            points[r * d + j] = c;
            points_t[j * n + r] = c;
            s += c * c;
        }
        norms[r] = s;
    }
    free(means);
    base_ctx c = {points, points_t, norms, n, d, k, slice, queries, n_queries,
                  out_indices, out_distances};
    size_t tiles = (n_queries + slice - 1) / slice;
    parallel_for(base_tile, &c, tiles, BASE_QB * k * sizeof(uint64_t), t);
    free(points);
    free(points_t);
    free(norms);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Fuzzy simplicial set: crates/scrust-core/src/neighbors.rs:193-372.        */
/* ------------------------------------------------------------------------ */

/* ndarray 0.16.1 (Cargo.lock:736-737), numeric_util::unrolled_fold as used by
 * ArrayBase::sum on contiguous data: p0..p7 take elements 8i+j, the result is
 * ((((0 + (p0+p4)) + (p1+p5)) + (p2+p6)) + (p3+p7)) followed by the tail in order. */
float oracle_ndarray_sum_f32(const float* x, size_t len) {
    float p[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    size_t i = 0;
    while (len - i >= 8) {
        for (int j = 0; j < 8; j++) p[j] = p[j] + x[i + j];
        i += 8;
    }
    float acc = 0.0f;
    acc = acc + (p[0] + p[4]);
    acc = acc + (p[1] + p[5]);
    acc = acc + (p[2] + p[6]);
    acc = acc + (p[3] + p[7]);
    for (; i < len; i++) acc = acc + x[i];
    return acc;
}

typedef struct {
    uint32_t col;
    float w;
} conn_entry;

static int conn_by_col(const void* a, const void* b) {
    const uint32_t x = ((const conn_entry*)a)->col, y = ((const conn_entry*)b)->col;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* neighbors.rs:243-295: rho = first strictly positive distance; 64 bisection steps on
 * sigma with the cardinality summed in f32 and compared in f64 against log2(k + 1);
 * sigma floored at 1e-3 times the row mean over k + 1 entries (rho > 0) or the mean of
 * all distances (rho == 0). */
static void smooth_row(const float* row, size_t k, double target, float mean_distance,
                       float* sigma_out, float* rho_out) {
    float rho = 0.0f;
    for (size_t j = 0; j < k; j++)
        if (row[j] > 0.0f) {
            rho = row[j];
            break;
        }
    float low = 0.0f, high = FLT_MAX, mid = 1.0f;
    for (int it = 0; it < 64; it++) {
        float cardinality = 0.0f;
        for (size_t j = 0; j < k; j++) {
            const float offset = row[j] - rho;
            cardinality = cardinality + (offset > 0.0f ? expf(-offset / mid) : 1.0f);
        }
        const double c = (double)cardinality;
        if (fabs(c - target) < 1e-5) break;
        if (c > target) {
            high = mid;
            mid = (low + high) / 2.0f;
        } else {
            low = mid;
            mid = high >= FLT_MAX ? mid * 2.0f : (low + high) / 2.0f;
        }
    }
    const float reference = rho > 0.0f ? oracle_ndarray_sum_f32(row, k) / (float)(k + 1) : mean_distance;
    const float floor_value = 1e-3f * reference;
    *sigma_out = mid > floor_value ? mid : floor_value;
    *rho_out = rho;
}

long oracle_connectivities(const uint32_t* indices, const float* distances, size_t n, size_t k,
                           uint32_t* indptr, uint32_t* out_indices, float* out_values,
                           size_t capacity, float* sigmas, float* rhos) {
    if (n == 0 || k == 0) return -1;
    for (size_t i = 0; i < n * k; i++)
        if (indices[i] >= n) return -1;
    const double target = log2((double)(k + 1));
    const float mean_distance = oracle_ndarray_sum_f32(distances, n * k) / (float)(n * k);
    float* sg = (float*)malloc(n * sizeof(float));
    float* rh = (float*)malloc(n * sizeof(float));
    conn_entry* fwd = (conn_entry*)malloc(n * k * sizeof(conn_entry));
    size_t* flen = (size_t*)calloc(n, sizeof(size_t));
    size_t* indeg = (size_t*)calloc(n + 1, sizeof(size_t));
    for (size_t i = 0; i < n; i++) {
        smooth_row(distances + i * k, k, target, mean_distance, &sg[i], &rh[i]);
        /* neighbors.rs:298-327: weights of the directed graph, row sorted by column. */
        conn_entry* row = fwd + i * k;
        size_t m = 0;
        for (size_t j = 0; j < k; j++) {
            const uint32_t nb = indices[i * k + j];
            if (nb == i) continue;
            const float offset = distances[i * k + j] - rh[i];
            const float w = (offset <= 0.0f || sg[i] == 0.0f) ? 1.0f : expf(-offset / sg[i]);
            if (w != 0.0f) row[m++] = (conn_entry){nb, w};
        }
        qsort(row, m, sizeof(conn_entry), conn_by_col);
        flen[i] = m;
        for (size_t j = 0; j < m; j++) indeg[row[j].col + 1]++;
    }
    /* neighbors.rs:330-372: the transpose, then the sorted union a + a^T - a * a^T. */
    for (size_t i = 0; i < n; i++) indeg[i + 1] += indeg[i];
    conn_entry* back = (conn_entry*)malloc((indeg[n] + 1) * sizeof(conn_entry));
    size_t* fill = (size_t*)calloc(n, sizeof(size_t));
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < flen[i]; j++) {
            const uint32_t c = fwd[i * k + j].col;
            back[indeg[c] + fill[c]++] = (conn_entry){(uint32_t)i, fwd[i * k + j].w};
        }
    long nnz = 0;
    int ok = 1;
    indptr[0] = 0;
    for (size_t i = 0; i < n && ok; i++) {
        conn_entry* b = back + indeg[i];
        const size_t bl = indeg[i + 1] - indeg[i];
        qsort(b, bl, sizeof(conn_entry), conn_by_col);
        const conn_entry* f = fwd + i * k;
        size_t x = 0, y = 0;
        while (x < flen[i] || y < bl) {
            uint32_t col;
            if (x < flen[i] && y < bl) col = f[x].col < b[y].col ? f[x].col : b[y].col;
            else col = x < flen[i] ? f[x].col : b[y].col;
            float weight = 0.0f, reverse = 0.0f;
            if (x < flen[i] && f[x].col == col) weight = f[x++].w;
            if (y < bl && b[y].col == col) reverse = b[y++].w;
            if ((size_t)nnz >= capacity) {
                ok = 0;
                break;
            }
            out_indices[nnz] = col;
            out_values[nnz] = weight + reverse - weight * reverse;
            nnz++;
        }
        indptr[i + 1] = (uint32_t)nnz;
    }
    if (sigmas) memcpy(sigmas, sg, n * sizeof(float));
    if (rhos) memcpy(rhos, rh, n * sizeof(float));
    free(sg);
    free(rh);
    free(fwd);
    free(flen);
    free(indeg);
    free(back);
    free(fill);
    return ok ? nnz : -1;
}

/* ------------------------------------------------------------------------ */
/* t-SNE input distances: squared_distances(&centred(&points)),             */
/* crates/scrust-core/src/tsne.rs:86, :173-188.                              */
/*                                                                          */
/* centred (:173-177): column sums in f32, times f32(1/n) (candle's affine   */
/* casts the factor to f32), subtracted in f32. squared_distances           */
/* (:181-188): norms = row sums of squares, gram = points . points^T, then  */
/* (norm_i + norm_j) - 2 gram_ij, clamped at 0. candle's reductions and     */
/* gemm's blocking fix summation orders this restatement does not follow:   */
/* column sums are exact and rounded once, row sums and dot products run    */
/* in dimension order with every product rounded. Entries meet the          */
/* reference's to the expansion's rounding bound, not bit for bit.          */
/* ------------------------------------------------------------------------ */

typedef struct {
    const float* coords;
    const float* norms;
    size_t n, d, row_begin;
    float* out;
} tsne_ctx;

static void tsne_row(void* arg, size_t item, void* scratch) {
    (void)scratch;
    const tsne_ctx* c = (const tsne_ctx*)arg;
    const size_t i = c->row_begin + item;
    const float* a = c->coords + i * c->d;
    float* dst = c->out + item * c->n;
    for (size_t j = 0; j < c->n; j++) {
        const float* b = c->coords + j * c->d;
        float g = 0.0f;
        for (size_t t = 0; t < c->d; t++) {
            const float p = a[t] * b[t];
            g = g + p;
        }
        const float sum = c->norms[i] + c->norms[j];
        const float v = sum - 2.0f * g;
        dst[j] = v > 0.0f ? v : 0.0f;
    }
}

int oracle_tsne_squared_distances(const float* x, size_t n, size_t d, size_t row_begin,
                                  size_t row_end, float* out, int threads) {
    if (n == 0 || d == 0 || row_begin > row_end || row_end > n) return -1;
    float* coords = (float*)malloc(n * d * sizeof(float));
    float* norms = (float*)malloc(n * sizeof(float));
    float* mean = (float*)malloc(d * sizeof(float));
    if (!coords || !norms || !mean) {
        free(coords);
        free(norms);
        free(mean);
        return -1;
    }
    const float inv_n = (float)(1.0 / (double)n);
    for (size_t j = 0; j < d; j++) mean[j] = (float)oracle_exact_sum_f32(x + j, n, d) * inv_n;
    for (size_t i = 0; i < n; i++) {
        float s = 0.0f;
        for (size_t j = 0; j < d; j++) {
            const float c = x[i * d + j] - mean[j];
            coords[i * d + j] = c;
            const float sq = c * c;
            s = s + sq;
        }
        norms[i] = s;
    }
    tsne_ctx ctx = {coords, norms, n, d, row_begin, out};
    parallel_for(tsne_row, &ctx, row_end - row_begin, 0, threads);
    free(coords);
    free(norms);
    free(mean);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* PCA by randomised SVD (crates/scrust-core/src/pca.rs:74-215), in f64.   */
/*                                                                          */
/* The steps of pca.rs: per-gene f64 sums in row order (ColumnStats, :226-  */
/* 255), a Gaussian sketch of width min(n_components + 24, min(n, g))       */
/* (:21, :103), orthonormalisation by two whitening passes with the Gram    */
/* eigenvalues clamped at 1e-7 of the largest (:352-386), 7 or 4 power      */
/* iterations by the 0.1 rule (:111-119), the final Gram eigensolve of      */
/* B^T (:148-166), embedding and loadings (:168-190), the sign convention   */
/* (:471-490) and both explained-variance statistics (:192-208). Dense      */
/* algebra is f64 here (the reference runs it in f32 on candle); the        */
/* sketch's normal draws come from pca_sketch below, not from the rand      */
/* crate's StdRng (rand 0.8, not under the reference tree).                 */
/* ------------------------------------------------------------------------ */

/* splitmix64 stream from `seed`; pairs of 53-bit uniforms through Box-Muller in f64,
 * rounded to f32, row-major rows x cols. The device library draws the same values. */
void oracle_pca_sketch(size_t rows, size_t cols, uint64_t seed, float* out) {
    uint64_t state = seed;
    size_t count = rows * cols, i = 0;
    while (i < count) {
        double u[2];
        for (int t = 0; t < 2; t++) {
            state += 0x9E3779B97F4A7C15ull;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            u[t] = (double)(z >> 11) * 0x1p-53;
        }
        double uniform = u[0] > DBL_MIN ? u[0] : DBL_MIN;
        double radius = sqrt(-2.0 * log(uniform));
        double angle = 6.283185307179586 * u[1];
        out[i++] = (float)(radius * cos(angle));
        if (i < count) out[i++] = (float)(radius * sin(angle));
    }
}

/* Cyclic Jacobi rotations of a symmetric n x n row-major matrix (pca.rs:390-445): returns 0
 * and the eigenvalues (diagonal) and eigenvectors (columns), or -1 when 60 sweeps do not
 * bring the off-diagonal mass to eps^2 of the total. */
int oracle_jacobi_eigen(double* a, size_t n, double* values, double* vectors) {
    for (size_t i = 0; i < n * n; i++) vectors[i] = 0.0;
    for (size_t i = 0; i < n; i++) vectors[i * n + i] = 1.0;
    double scale = 0.0;
    for (size_t i = 0; i < n * n; i++) scale += a[i] * a[i];
    const double tolerance = DBL_EPSILON * DBL_EPSILON * (scale > DBL_MIN ? scale : DBL_MIN);
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (size_t r = 0; r < n; r++)
            for (size_t c = 0; c < n; c++)
                if (r != c) off += a[r * n + c] * a[r * n + c];
        if (off <= tolerance) {
            for (size_t i = 0; i < n; i++) values[i] = a[i * n + i];
            return 0;
        }
        for (size_t p = 0; p < n; p++) {
            for (size_t q = p + 1; q < n; q++) {
                double pivot = a[p * n + q];
                if (pivot == 0.0) continue;
                double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * pivot);
                double sgn = theta > 0.0 ? 1.0 : (theta < 0.0 ? -1.0 : (signbit(theta) ? -1.0 : 1.0));
                double den = fabs(theta) + sqrt(theta * theta + 1.0);
                double t = sgn / (den > DBL_EPSILON ? den : DBL_EPSILON);
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (size_t k = 0; k < n; k++) {
                    double l = a[k * n + p], r = a[k * n + q];
                    a[k * n + p] = c * l - s * r;
                    a[k * n + q] = s * l + c * r;
                }
                for (size_t k = 0; k < n; k++) {
                    double l = a[p * n + k], r = a[q * n + k];
                    a[p * n + k] = c * l - s * r;
                    a[q * n + k] = s * l + c * r;
                }
                a[p * n + q] = 0.0;
                a[q * n + p] = 0.0;
                for (size_t k = 0; k < n; k++) {
                    double l = vectors[k * n + p], r = vectors[k * n + q];
                    vectors[k * n + p] = c * l - s * r;
                    vectors[k * n + q] = s * l + c * r;
                }
            }
        }
    }
    return -1;
}

/* y (rows x w) <- y T with T = V diag(1 / sqrt(max(lambda, 1e-7 lambda_max))) of y^T y. */
static int pca_whiten(double* y, size_t rows, size_t w, double* scratch) {
    double* gram = scratch;
    double* values = gram + w * w;
    double* vectors = values + w;
    double* t = vectors + w * w;
    double* row = t + w * w;
    for (size_t i = 0; i < w * w; i++) gram[i] = 0.0;
    for (size_t r = 0; r < rows; r++)
        for (size_t i = 0; i < w; i++)
            for (size_t j = 0; j < w; j++) gram[i * w + j] += y[r * w + i] * y[r * w + j];
    if (oracle_jacobi_eigen(gram, w, values, vectors) != 0) return -1;
    double largest = 0.0;
    for (size_t i = 0; i < w; i++) largest = values[i] > largest ? values[i] : largest;
    double floor_value = largest * 1e-7;
    for (size_t c = 0; c < w; c++) {
        double v = values[c] > floor_value ? values[c] : floor_value;
        double sc = floor_value > 0.0 ? 1.0 / sqrt(v) : 0.0;
        for (size_t r = 0; r < w; r++) t[r * w + c] = vectors[r * w + c] * sc;
    }
    for (size_t r = 0; r < rows; r++) {
        for (size_t c = 0; c < w; c++) {
            double s = 0.0;
            for (size_t k = 0; k < w; k++) s += y[r * w + k] * t[k * w + c];
            row[c] = s;
        }
        memcpy(y + r * w, row, w * sizeof(double));
    }
    return 0;
}

/* X_c @ rhs (n x w) with rhs g x w; or X_c^T @ lhs (g x w) with lhs n x w. */
static void pca_times(const uint32_t* indptr, const uint32_t* indices, const float* values, size_t n, size_t g,
                      const double* mean, const double* rhs, size_t w, double* out) {
    double* shift = (double*)calloc(w, sizeof(double));
    if (mean)
        for (size_t k = 0; k < g; k++)
            for (size_t c = 0; c < w; c++) shift[c] += mean[k] * rhs[k * w + c];
    for (size_t r = 0; r < n; r++) {
        for (size_t c = 0; c < w; c++) out[r * w + c] = -shift[c];
        for (uint32_t j = indptr[r]; j < indptr[r + 1]; j++)
            for (size_t c = 0; c < w; c++) out[r * w + c] += (double)values[j] * rhs[(size_t)indices[j] * w + c];
    }
    free(shift);
}

static void pca_transpose_times(const uint32_t* indptr, const uint32_t* indices, const float* values, size_t n,
                                size_t g, const double* mean, const double* lhs, size_t w, double* out) {
    for (size_t i = 0; i < g * w; i++) out[i] = 0.0;
    for (size_t r = 0; r < n; r++)
        for (uint32_t j = indptr[r]; j < indptr[r + 1]; j++)
            for (size_t c = 0; c < w; c++) out[(size_t)indices[j] * w + c] += (double)values[j] * lhs[r * w + c];
    if (mean) {
        double* sums = (double*)calloc(w, sizeof(double));
        for (size_t r = 0; r < n; r++)
            for (size_t c = 0; c < w; c++) sums[c] += lhs[r * w + c];
        for (size_t k = 0; k < g; k++)
            for (size_t c = 0; c < w; c++) out[k * w + c] -= mean[k] * sums[c];
        free(sums);
    }
}

int oracle_pca(const uint32_t* indptr, const uint32_t* indices, const float* values, size_t n, size_t g,
               size_t n_components, int zero_center, uint64_t seed, float* embedding, float* components,
               float* explained_variance, float* explained_variance_ratio) {
    size_t max_components = n < g ? n : g;
    if (n_components == 0 || n_components > max_components || n < 2) return -1;
    /* ColumnStats: f64 sums and squares in row order. */
    double* sums = (double*)calloc(g, sizeof(double));
    double* squares = (double*)calloc(g, sizeof(double));
    for (size_t j = 0; j < indptr[n]; j++) {
        double v = (double)values[j];
        sums[indices[j]] += v;
        squares[indices[j]] += v * v;
    }
    double* mean = (double*)malloc(g * sizeof(double));
    double ddof = zero_center ? 1.0 : 0.0, total_variance = 0.0;
    for (size_t k = 0; k < g; k++) {
        mean[k] = (double)(float)(sums[k] / (double)n); /* the reference's mean is f32 */
        double var = squares[k] - sums[k] * sums[k] / (double)n;
        total_variance += (var > 0.0 ? var : 0.0) / ((double)n - ddof);
    }
    const double* centre = zero_center ? mean : NULL;
    size_t w = n_components + 24 < max_components ? n_components + 24 : max_components;
    float* omega32 = (float*)malloc(g * w * sizeof(float));
    oracle_pca_sketch(g, w, seed, omega32);
    double* omega = (double*)malloc(g * w * sizeof(double));
    for (size_t i = 0; i < g * w; i++) omega[i] = omega32[i];
    size_t big = n > g ? n : g;
    double* range = (double*)malloc(n * w * sizeof(double));
    double* genewise = (double*)malloc(g * w * sizeof(double));
    double* scratch = (double*)malloc((3 * w * w + 2 * w + big) * sizeof(double));
    int rc = 0;
    pca_times(indptr, indices, values, n, g, centre, omega, w, range);
    rc |= pca_whiten(range, n, w, scratch) | pca_whiten(range, n, w, scratch);
    int iterations = (double)n_components < 0.1 * (double)max_components ? 7 : 4;
    for (int it = 0; it < iterations && rc == 0; it++) {
        pca_transpose_times(indptr, indices, values, n, g, centre, range, w, genewise);
        rc |= pca_whiten(genewise, g, w, scratch) | pca_whiten(genewise, g, w, scratch);
        pca_times(indptr, indices, values, n, g, centre, genewise, w, range);
        rc |= pca_whiten(range, n, w, scratch) | pca_whiten(range, n, w, scratch);
    }
    double* bt = genewise; /* B^T = X_c^T range, g x w */
    double* gram = (double*)calloc(w * w, sizeof(double));
    double* evals = (double*)malloc(w * sizeof(double));
    double* evecs = (double*)malloc(w * w * sizeof(double));
    size_t* order = (size_t*)malloc(w * sizeof(size_t));
    if (rc == 0) {
        pca_transpose_times(indptr, indices, values, n, g, centre, range, w, bt);
        for (size_t k = 0; k < g; k++)
            for (size_t i = 0; i < w; i++)
                for (size_t j = 0; j < w; j++) gram[i * w + j] += bt[k * w + i] * bt[k * w + j];
        rc = oracle_jacobi_eigen(gram, w, evals, evecs);
    }
    if (rc == 0) {
        /* Descending eigenvalues, ties in index order (a stable sort). */
        for (size_t i = 0; i < w; i++) order[i] = i;
        for (size_t i = 1; i < w; i++) {
            size_t v = order[i], j = i;
            while (j > 0 && evals[order[j - 1]] < evals[v]) {
                order[j] = order[j - 1];
                j--;
            }
            order[j] = v;
        }
        double* singular = (double*)malloc(n_components * sizeof(double));
        for (size_t c = 0; c < n_components; c++) {
            double e = evals[order[c]];
            singular[c] = sqrt(e > 0.0 ? e : 0.0);
        }
        for (size_t r = 0; r < n; r++)
            for (size_t c = 0; c < n_components; c++) {
                double s = 0.0;
                for (size_t k = 0; k < w; k++) s += range[r * w + k] * evecs[k * w + order[c]];
                embedding[r * n_components + c] = (float)(s * singular[c]);
            }
        for (size_t c = 0; c < n_components; c++) {
            double inverse = singular[c] > 0.0 ? 1.0 / singular[c] : 0.0;
            for (size_t k = 0; k < g; k++) {
                double s = 0.0;
                for (size_t i = 0; i < w; i++) s += bt[k * w + i] * evecs[i * w + order[c]];
                components[c * g + k] = (float)(s * inverse);
            }
        }
        /* Signs: the largest-magnitude loading of every component is made positive. */
        for (size_t c = 0; c < n_components; c++) {
            float extreme = 0.0f;
            for (size_t k = 0; k < g; k++)
                if (fabsf(components[c * g + k]) > fabsf(extreme)) extreme = components[c * g + k];
            if (extreme < 0.0f) {
                for (size_t k = 0; k < g; k++) components[c * g + k] = -components[c * g + k];
                for (size_t r = 0; r < n; r++) embedding[r * n_components + c] = -embedding[r * n_components + c];
            }
        }
        for (size_t c = 0; c < n_components; c++) {
            if (zero_center) {
                explained_variance[c] = (float)(singular[c] * singular[c] / (double)(n - 1));
            } else {
                double m = 0.0, v = 0.0;
                for (size_t r = 0; r < n; r++) m += embedding[r * n_components + c];
                m /= (double)n;
                for (size_t r = 0; r < n; r++) {
                    double d = embedding[r * n_components + c] - m;
                    v += d * d;
                }
                explained_variance[c] = (float)(v / (double)n);
            }
            explained_variance_ratio[c] =
                total_variance > 0.0 ? (float)((double)explained_variance[c] / total_variance) : 0.0f;
        }
        free(singular);
    }
    free(sums); free(squares); free(mean); free(omega32); free(omega); free(range); free(genewise);
    free(scratch); free(gram); free(evals); free(evecs); free(order);
    return rc == 0 ? 0 : -2;
}
