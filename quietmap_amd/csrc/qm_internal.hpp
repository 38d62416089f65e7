// The library's private handle types, shared by its translation units.
#pragma once

#include <mutex>
#include <vector>

#include "qm_path.hpp"

struct qm_gpu {
    int ordinal;
    char name[256];
    qm::BandTable bands;
    std::mutex lock;
    void* noalloc_scratch;  // this device's address of painter.hip's static scratch array
};

// Point sources bucketed on a square grid (device pointers): bucket (kx, ky) holds sources
// bucket_first[ky * buckets_x + kx] .. bucket_first[ky * buckets_x + kx + 1] of `sources`.
struct SourceGrid {
    const qm_source* sources;
    const uint32_t* bucket_first;
    double origin_x, origin_y, bucket_m;
    uint32_t buckets_x, buckets_y;
};

struct qm_square {
    qm_gpu* gpu;
    qm::Scene scene;  // device pointers
    const qm::Scene* d_scene;        // `scene` itself in device memory
    const qm::BandTable* d_bands;    // the GPU's band table in device memory
    SourceGrid grid;  // device pointers
    size_t n_sources;
    bool has_forest;  // a terrain node with forest cover (else the path kernels skip the foliage term)
    qm_source* d_sources_by_input;  // input order (for pairs)
    std::vector<qm_line_piece> lines;  // line candidates, numbered after the point sources
    std::vector<void*> allocations;
    size_t bytes;
};

struct qm_mlp;
namespace qm {
// The surrogate network's outputs (rows x 8, FP32) of BF16 input rows (rows x 40) on `stream`;
// with `d_rows_pairs`, the rows are 2 * *d_rows_pairs (at most `rows`), read on the device.
int mlp_infer_rows(qm_mlp* mlp, const uint16_t* d_rows, size_t rows, const uint32_t* d_rows_pairs, float* d_outputs,
                   hipStream_t stream, char* err, size_t errlen);
}  // namespace qm
