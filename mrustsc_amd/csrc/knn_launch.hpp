// Host-side declarations of the kNN kernel launches (knn_kernels.hip).
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

constexpr unsigned kFlagNonFinite = 1u;
constexpr unsigned kFlagBandOverflow = 2u;

constexpr int kCap = 256;   // candidate buffer slots per query row in the pass
// Widest boundary tie band resolved in fp64; wider bands keep the (d^2, id) order.
// Every rerank band has at most KP - 1 <= 63 members, so the cap only binds on the
// exact path.
constexpr int kBandCap = 64;
constexpr int kMaxK = 256;  // largest k any path serves
constexpr int kMaxPassK = 47;  // largest k the MFMA candidate pass serves (KP = 64)
constexpr int kMaxPassD = 128; // largest d with a compiled MFMA variant

struct PrepareArgs {
    const float* x;
    size_t n;
    int d;
    int metric;
    const double* mean;
    float* coords;
    float* norms;
    int operand;
    uint16_t* pack_ref;
    uint16_t* pack_query;
    float* pack_f32;
    int nch;
    int group;
    unsigned* max_norm_bits;
    unsigned* flags;
};

struct PassArgs {
    int operand;
    int nch;
    int kp;
    int d;
    const uint16_t* pack_query;
    const uint16_t* pack_ref;
    const float* pack_f32;
    const float* norms;
    uint32_t n;
    uint64_t* buf;
    uint64_t* cand;
    float* theta;
};

struct RerankArgs {
    const float* coords;
    const float* norms;
    uint32_t n;
    int d;
    int k;
    int kp;
    int metric;
    const uint64_t* cand;
    const float* theta;
    const unsigned* max_norm_bits;
    // Error model of the pass (see the certificate in rerank_kernel):
    double split_rel;  // operand split error of a.b, relative to |a||b|
    double acc_rel;    // MFMA accumulation error, relative to the sum of |terms|
    double norm_rel;   // error of the |b|^2 feature, relative to |b|^2
    uint32_t* out_indices;
    float* out_distances;
    uint32_t* fail_list;
    uint32_t* fail_count;
};

struct ExactArgs {
    const float* coords;
    const float* norms;
    uint32_t n;
    int d;
    int k;
    int metric;
    const uint32_t* list;        // rows to compute; nullptr: every row
    const uint32_t* list_count;  // device count of `list`
    uint32_t* out_indices;
    float* out_distances;
    unsigned* flags;
};

void launch_fill_u32(uint32_t* p, size_t count, uint32_t value, hipStream_t stream);
void launch_column_means(const float* x, size_t n, int d, unsigned long long* acc, double* mean,
                         hipStream_t stream);
void launch_prepare(const PrepareArgs& a, hipStream_t stream);
bool launch_pass(const PassArgs& a, hipStream_t stream);
void launch_rerank(const RerankArgs& a, hipStream_t stream);
void launch_exact(const ExactArgs& a, unsigned blocks, hipStream_t stream);

// MFMA chunk counts of the compiled pass variants for a dimension d (0: none).
int bf16_chunks(int d);
int fp32_chunks(int d);
int fp32_group(int nch);

constexpr int kAccWordsHost = 10;

}  // namespace scrust_hip
