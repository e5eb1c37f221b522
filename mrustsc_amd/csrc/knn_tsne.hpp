// Host declarations of the t-SNE input distances: the centred pairwise squared-distance
// matrix exact t-SNE starts from (crates/scrust-core/src/tsne.rs:86, :173-188).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace scrust_hip {

// The reference's exact t-SNE refuses more cells than this before it forms the n x n
// matrix (tsne.rs:40, :146-152).
constexpr size_t kTsneMaxCells = 20000;
constexpr uint32_t kTsneFlagNonFinite = 1u;

// Byte offsets of the workspace.
struct TsneLayout {
    size_t acc;     // exact column sums (d x kAccWords int64)
    size_t mean;    // f32 column means (d)
    size_t coords;  // centred coordinates, n x dpad f32 (zero padded)
    size_t norms;   // f32 squared norms (n, zero padded to whole 128-row tiles)
    size_t flags;   // u32
    size_t total;
    uint32_t dpad;  // d rounded up to a multiple of 4 (the MFMA's k step)
};

TsneLayout tsne_layout(size_t n, size_t d);

// Enqueues centring and the distance tiles: out is n x n f32 row-major on the device.
// The flags word holds kTsneFlagNonFinite once the stream reaches the end.
hipError_t tsne_enqueue(const float* points, size_t n, size_t d, void* workspace,
                        const TsneLayout& layout, float* out, hipStream_t stream);

}  // namespace scrust_hip
