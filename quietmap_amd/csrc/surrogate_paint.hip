// The surrogate painter and its training samples (no counterpart in the reference; include/
// qm_surrogate.h states the contract).
//
//   samples_kernel   - per pair: the surrogate's inputs (calm and favourable rows) and the exact
//                      path's boundary terms per band as targets, with the envelope flag
//   tile_count_kernel, tile_list_kernel - per 16x16 receiver tile, the sources within the cutoff
//                      of its box: the (tile, source) items, 256 pairs each
//   features_kernel  - a workgroup per item (its lanes the tile's receivers as the exact grid
//                      painter maps them, the source uniform): BF16 input rows for the network, a
//                      flag (no path / network / exact) and the source altitude; pairs outside the
//                      envelope are appended to the exact list
//   exact_kernel     - the exact path of every listed pair: its energy per period
//   combine_kernel   - per receiver, over its tile's items: the network's A_ground + A_dif per band
//                      and state mixed into received energy as the exact path mixes it
//                      (physics/src/ray.rs:155-190, popup/src/evaluate.rs:119-203), or the exact
//                      pair's energy; FP64 sums per receiver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <vector>

#include "../../include/qm_surrogate.h"
#include "qm_features.hpp"
#include "qm_internal.hpp"

using namespace qm;

namespace {

__device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// The bucket range of sources that can lie within `cutoff` of (x, y).
struct BucketBox {
    int kx0, kx1, ky0, ky1;
};
__device__ BucketBox bucket_box(const SourceGrid& sg, double x, double y, double cutoff) {
    BucketBox b;
    b.kx0 = max((int)floor((x - cutoff - sg.origin_x) / sg.bucket_m), 0);
    b.kx1 = min((int)floor((x + cutoff - sg.origin_x) / sg.bucket_m), (int)sg.buckets_x - 1);
    b.ky0 = max((int)floor((y - cutoff - sg.origin_y) / sg.bucket_m), 0);
    b.ky1 = min((int)floor((y + cutoff - sg.origin_y) / sg.bucket_m), (int)sg.buckets_y - 1);
    return b;
}

}  // namespace

__global__ __launch_bounds__(QM_PATH_GROUP) __attribute__((amdgpu_waves_per_eu(2))) void samples_kernel(
    const Scene* __restrict__ scene, const BandTable* __restrict__ bands, const qm_receiver* receivers,
    const qm_source* sources, const uint32_t* pair_receiver, const uint32_t* pair_source, uint32_t n, float* inputs,
    float* targets, uint32_t* in_envelope) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PathScratch w;
    const qm_receiver& rcv = receivers[pair_receiver[i]];
    const qm_source src = sources[pair_source[i]];
    float f[FEATURES];
    const SurrogatePath sp = surrogate_features(sc, rcv, src, w.cross, f);
    PathResult r;
    evaluate_path(sc, B, rcv, src, w, r);
    float* x0 = inputs + (size_t)i * 2 * FEATURES;
    for (int k = 0; k < FEATURES; ++k) {
        x0[k] = f[k];
        x0[FEATURES + k] = k == 37 ? 1.0f : f[k];
    }
    float* t0 = targets + (size_t)i * 2 * BANDS;
    for (int b = 0; b < BANDS; ++b) {
        t0[b] = r.att[0][b];
        t0[BANDS + b] = r.att[1][b];
    }
    in_envelope[i] = (sp.in_envelope && r.failed == 0) ? 1u : 0u;
}

// The receiver grid in 16x16 tiles; a tile's 256 work-items map to its receivers as the exact grid
// painter maps them (four 8x8 waves), and a tile's sources are those of the bucket rows around it
// whose distance to the tile's box is within the cutoff.
struct TileGrid {
    const qm_receiver* receivers;
    uint32_t nx, ny, tiles_x;
    double cutoff;
};

__device__ void tile_box(const TileGrid& g, uint32_t tile, double& x0, double& y0, double& x1, double& y1) {
    const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
    const uint32_t ix0 = tx * 16, iy0 = ty * 16;
    const uint32_t ix1 = min(ix0 + 15, g.nx - 1), iy1 = min(iy0 + 15, g.ny - 1);
    const qm_receiver& a = g.receivers[(size_t)iy0 * g.nx + ix0];
    const qm_receiver& b = g.receivers[(size_t)iy1 * g.nx + ix1];
    x0 = fmin(a.x_m, b.x_m);
    x1 = fmax(a.x_m, b.x_m);
    y0 = fmin(a.y_m, b.y_m);
    y1 = fmax(a.y_m, b.y_m);
}

// The tile's sources, counted (`out` null) or listed from out[0].
__device__ uint32_t tile_sources(const TileGrid& g, const SourceGrid& sg, uint32_t tile, uint32_t* out) {
    double x0, y0, x1, y1;
    tile_box(g, tile, x0, y0, x1, y1);
    const int kx0 = max((int)floor((x0 - g.cutoff - sg.origin_x) / sg.bucket_m), 0);
    const int kx1 = min((int)floor((x1 + g.cutoff - sg.origin_x) / sg.bucket_m), (int)sg.buckets_x - 1);
    const int ky0 = max((int)floor((y0 - g.cutoff - sg.origin_y) / sg.bucket_m), 0);
    const int ky1 = min((int)floor((y1 + g.cutoff - sg.origin_y) / sg.bucket_m), (int)sg.buckets_y - 1);
    const double cut2 = g.cutoff * g.cutoff;
    uint32_t count = 0;
    for (int ky = ky0; ky <= ky1; ++ky) {
        const uint32_t s0 = sg.bucket_first[(size_t)ky * sg.buckets_x + kx0];
        const uint32_t s1 = sg.bucket_first[(size_t)ky * sg.buckets_x + kx1 + 1];
        for (uint32_t s = s0; s < s1; ++s) {
            const double sx = sg.sources[s].x_m, sy = sg.sources[s].y_m;
            const double dx = sx < x0 ? x0 - sx : (sx > x1 ? sx - x1 : 0.0);
            const double dy = sy < y0 ? y0 - sy : (sy > y1 ? sy - y1 : 0.0);
            if (dx * dx + dy * dy > cut2) continue;
            if (out) out[count] = s;
            ++count;
        }
    }
    return count;
}

__global__ __launch_bounds__(256) void tile_count_kernel(TileGrid g, SourceGrid sg, uint32_t n_tiles, uint32_t* counts) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_tiles) counts[t] = tile_sources(g, sg, t, nullptr);
}

// Items of a chunk of tiles first_tile..: item k is (item_tile[k], item_source[k]).
__global__ __launch_bounds__(256) void tile_list_kernel(TileGrid g, SourceGrid sg, uint32_t first_tile, uint32_t n,
                                                        const uint32_t* offsets, uint32_t* item_tile,
                                                        uint32_t* item_source) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = tile_sources(g, sg, first_tile + i, item_source + offsets[i]);
    for (uint32_t j = 0; j < k; ++j) item_tile[offsets[i] + j] = first_tile + i;
}

__device__ __forceinline__ bool tile_receiver(const TileGrid& g, uint32_t tile, uint32_t& index) {
    const int lane = threadIdx.x, wave = lane >> 6, l = lane & 63;
    const uint32_t ix = (tile % g.tiles_x) * 16 + (wave & 1) * 8 + (l & 7);
    const uint32_t iy = (tile / g.tiles_x) * 16 + (wave >> 1) * 8 + (l >> 3);
    index = (uint32_t)(min(iy, g.ny - 1) * g.nx + min(ix, g.nx - 1));
    return ix < g.nx && iy < g.ny;
}

// Pair flags: 0 no path (outside the grid or the cutoff), 1 the network's, 2 the exact path's.
__global__ __launch_bounds__(256) void features_kernel(Scene sc, TileGrid g, const qm_source* sources,
                                                       const uint32_t* item_tile, const uint32_t* item_source,
                                                       uint16_t* rows, uint32_t* row_count, uint32_t* row_of,
                                                       uint8_t* flags, float* source_alt, uint32_t* exact_list,
                                                       uint32_t* exact_count) {
    const uint32_t item = blockIdx.x;
    const size_t pair = (size_t)item * 256 + threadIdx.x;
    uint32_t index;
    const bool valid = tile_receiver(g, item_tile[item], index);
    const qm_receiver& rcv = g.receivers[index];
    const qm_source& src = sources[item_source[item]];
    const double dx = src.x_m - rcv.x_m, dy = src.y_m - rcv.y_m;
    if (!valid || dx * dx + dy * dy > g.cutoff * g.cutoff) {
        flags[pair] = 0;
        return;
    }
    Crossing cross[CROSS_CAP];
    float f[FEATURES];
    const SurrogatePath sp = surrogate_features(sc, rcv, src, cross, f);
    if (sp.in_envelope) {
        // The network's rows of this pair: the next free pair of rows (calm, favourable).
        const uint32_t slot = atomicAdd(row_count, 1u);
        row_of[pair] = slot;
        uint16_t* x0 = rows + (size_t)slot * 2 * FEATURES;
        for (int k = 0; k < FEATURES; ++k) {
            const uint16_t h = to_bf16(f[k]);
            x0[k] = h;
            x0[FEATURES + k] = k == 37 ? to_bf16(1.0f) : h;
        }
        flags[pair] = 1;
        // The source's altitude for the energy terms (its ground as the exact path samples it).
        const qm_terrain& T = sc.terrain;
        const double inv = 1.0 / T.spacing_m;
        GroundSample s0;
        ground_at(T, (src.x_m - T.origin_x_m) * inv, (src.y_m - T.origin_y_m) * inv, s0);
        source_alt[pair] = s0.z + src.height_m;
    } else {
        flags[pair] = 2;
        exact_list[atomicAdd(exact_count, 1u)] = (uint32_t)pair;
    }
}

__global__ __launch_bounds__(QM_PATH_GROUP) __attribute__((amdgpu_waves_per_eu(2))) void exact_kernel(
    const Scene* __restrict__ scene, const BandTable* __restrict__ bands, TileGrid g, const qm_source* sources,
    const uint32_t* item_tile, const uint32_t* item_source, const uint32_t* exact_list, const uint32_t* exact_count,
    float* exact_energy, uint8_t* exact_failed) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    const uint32_t count = *exact_count;
    PathScratch w;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const uint32_t pair = exact_list[k];
        const uint32_t item = pair / 256, lane = pair % 256;
        const uint32_t tile = item_tile[item];
        const uint32_t wave = lane >> 6, l = lane & 63;
        const uint32_t ix = (tile % g.tiles_x) * 16 + (wave & 1) * 8 + (l & 7);
        const uint32_t iy = (tile / g.tiles_x) * 16 + (wave >> 1) * 8 + (l >> 3);
        PathResult r;
        evaluate_path(sc, B, g.receivers[(size_t)iy * g.nx + ix], sources[item_source[item]], w, r);
        for (int q = 0; q < PERIODS; ++q) {
            const float s01 = r.bands[q][0] + r.bands[q][1], s23 = r.bands[q][2] + r.bands[q][3];
            const float s45 = r.bands[q][4] + r.bands[q][5], s67 = r.bands[q][6] + r.bands[q][7];
            exact_energy[(size_t)pair * PERIODS + q] = r.failed ? 0.0f : (s01 + s23) + (s45 + s67);
        }
        exact_failed[pair] = r.failed ? 1 : 0;
    }
}

// popup/src/evaluate.rs:119-203 and physics/src/ray.rs:155-190 as the exact path applies them (no
// foliage: the envelope has none), with the network's per-band A_ground + A_dif of both states.
// A workgroup per (tile, block of COMBINE_ITEMS of its items), a work-item per receiver: the
// block's partial energy and counts, reduced per tile in block order by reduce_kernel.
constexpr uint32_t COMBINE_ITEMS = 32;

struct Partial {
    double energy[PERIODS];
    uint32_t paths, fails, learned, pad;
};

__global__ __launch_bounds__(256) void combine_kernel(TileGrid g, uint32_t first_tile, const qm_source* sources,
                                                      const uint32_t* tile_items, uint32_t chunk_item0,
                                                      const uint32_t* block_tile, const uint32_t* block_item0,
                                                      const uint32_t* item_source, const uint8_t* flags,
                                                      const uint32_t* row_of, const float* source_alt, const float* att,
                                                      const float* exact_energy, const uint8_t* exact_failed,
                                                      Partial* partials) {
    const uint32_t blk = blockIdx.x;
    const uint32_t t = block_tile[blk];  // tile within the chunk
    uint32_t index;
    const bool valid = tile_receiver(g, first_tile + t, index);
    const qm_receiver& rcv = g.receivers[index];
    const uint32_t i0 = block_item0[blk], i1 = min(i0 + COMBINE_ITEMS, tile_items[t + 1]);
    Partial out{};
    for (uint32_t item = valid ? i0 : i1; item < i1; ++item) {
        const size_t p = (size_t)(item - chunk_item0) * 256 + threadIdx.x;
        const uint8_t flag = flags[p];
        if (flag == 0) continue;
        ++out.paths;
        float e[PERIODS];
        if (flag == 2) {
            if (exact_failed[p]) {
                ++out.fails;
                continue;
            }
            for (int q = 0; q < PERIODS; ++q) e[q] = exact_energy[p * PERIODS + q];
        } else {
            ++out.learned;
            const qm_source& src = sources[item_source[item]];
            const double dxd = rcv.x_m - src.x_m, dyd = rcv.y_m - src.y_m;
            const double length_d = hypot(dxd, dyd);
            const float H = (float)fmax(length_d, 1.0);
            const float dz = rcv.altitude_m - source_alt[p];
            const float slant = fmaxf(sqrtf(H * H + dz * dz), 1.0f);
            const float azimuth = atan2f((float)dyd, (float)dxd);
            float pf[PERIODS];
            for (int q = 0; q < PERIODS; ++q) pf[q] = favourable_at(rcv.favourable[q], azimuth);
            const float hd = fmaxf((float)length_d, src.exclusion_radius_m);
            const float d = fmaxf(sqrtf(hd * hd + dz * dz), 1.0f);
            const float weight = 1.0f / (d * d * 12.589254117941673f);
            const float* a0 = att + (size_t)row_of[p] * 2 * BANDS;
            float bands[PERIODS][BANDS];
            for (int b = 0; b < BANDS; ++b) {
                const float air = rcv.alpha_db_per_km[b] * slant * 1e-3f;
                const float eh = __expf(-(a0[b] + air) * (LN10 / 10.0f));
                const float ef = __expf(-(a0[BANDS + b] + air) * (LN10 / 10.0f));
                for (int q = 0; q < PERIODS; ++q)
                    bands[q][b] = weight * src.energy[q][b] * (pf[q] * ef + (1.0f - pf[q]) * eh);
            }
            for (int q = 0; q < PERIODS; ++q) {
                const float s01 = bands[q][0] + bands[q][1], s23 = bands[q][2] + bands[q][3];
                const float s45 = bands[q][4] + bands[q][5], s67 = bands[q][6] + bands[q][7];
                e[q] = (s01 + s23) + (s45 + s67);
            }
        }
        for (int q = 0; q < PERIODS; ++q) out.energy[q] += (double)e[q];
    }
    partials[(size_t)blk * 256 + threadIdx.x] = out;
}

// Per tile of the chunk: its blocks' partials summed in block order into the receivers.
__global__ __launch_bounds__(256) void reduce_kernel(TileGrid g, uint32_t first_tile, const uint32_t* tile_blocks,
                                                     const Partial* partials, double* period_energy, uint32_t* paths,
                                                     uint32_t* failures, uint32_t* surrogate_paths) {
    const uint32_t t = blockIdx.x;
    uint32_t index;
    if (!tile_receiver(g, first_tile + t, index)) return;
    double energy[PERIODS] = {0.0, 0.0, 0.0};
    uint32_t count = 0, fails = 0, learned = 0;
    for (uint32_t blk = tile_blocks[t]; blk < tile_blocks[t + 1]; ++blk) {
        const Partial& q = partials[(size_t)blk * 256 + threadIdx.x];
        for (int k = 0; k < PERIODS; ++k) energy[k] += q.energy[k];
        count += q.paths;
        fails += q.fails;
        learned += q.learned;
    }
    for (int k = 0; k < PERIODS; ++k) period_energy[(size_t)index * PERIODS + k] = energy[k];
    paths[index] = count;
    failures[index] = fails;
    if (surrogate_paths) surrogate_paths[index] = learned;
}

__global__ void zero_u32(uint32_t* p) { *p = 0u; }

namespace {

int fail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

#define SP_TRY(expr)                                                                   \
    do {                                                                               \
        const hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return fail(err, errlen, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

// Device allocations freed on the stream when the call returns (success or not).
struct StreamAllocations {
    hipStream_t stream;
    std::vector<void*> ptrs;
    ~StreamAllocations() {
        for (void* p : ptrs) (void)hipFreeAsync(p, stream);
    }
    hipError_t alloc(void** p, size_t bytes) {
        const hipError_t e = hipMallocAsync(p, std::max<size_t>(bytes, 256), stream);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
};

// (tile, source) items a chunk of the surrogate painter holds at most (256 pairs each).
constexpr size_t CHUNK_ITEMS = 1u << 15;

}  // namespace

extern "C" int qm_surrogate_samples_device(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers,
                                           const uint32_t* d_pair_receiver, const uint32_t* d_pair_source,
                                           size_t n_pairs, float* d_inputs, float* d_targets, uint32_t* d_in_envelope,
                                           void* stream, char* err, size_t errlen) {
    if (!gpu || !square || (n_pairs && (!d_receivers || !d_pair_receiver || !d_pair_source || !d_inputs ||
                                        !d_targets || !d_in_envelope)))
        return fail(err, errlen, "null argument");
    if (n_pairs == 0) return 0;
    if (n_pairs > 0xffffffffu) return fail(err, errlen, "too many pairs");
    if (square->scene.terrain.fine_detail || square->scene.terrain.row_y_m)
        return fail(err, errlen, "fine terrain or a row table: the network is trained on the coarse uniform walk");
    SP_TRY(hipSetDevice(gpu->ordinal));
    hipLaunchKernelGGL(samples_kernel, dim3(blocks_for(n_pairs)), dim3(QM_PATH_GROUP), 0, (hipStream_t)stream, square->d_scene,
                       square->d_bands, d_receivers, square->d_sources_by_input, d_pair_receiver, d_pair_source,
                       (uint32_t)n_pairs, d_inputs, d_targets, d_in_envelope);
    SP_TRY(hipGetLastError());
    return 0;
}

extern "C" int qm_surrogate_paint_receivers_device(qm_gpu* gpu, const qm_square* square, qm_mlp* mlp,
                                                   const qm_receiver* d_receivers, uint32_t nx, uint32_t ny,
                                                   double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                                                   uint32_t* d_failures, uint32_t* d_surrogate_paths, void* stream_ptr,
                                                   char* err, size_t errlen) {
    if (!gpu || !square || !mlp || !d_receivers || !d_period_energy || !d_paths || !d_failures)
        return fail(err, errlen, "null argument");
    if (!(cutoff_m >= 0.0)) return fail(err, errlen, "cutoff must be non-negative");
    if (square->scene.terrain.fine_detail || square->scene.terrain.row_y_m)
        return fail(err, errlen, "fine terrain or a row table: the network is trained on the coarse uniform walk");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    if (n > 0xffffffffu) return fail(err, errlen, "too many receivers");
    SP_TRY(hipSetDevice(gpu->ordinal));
    const hipStream_t s = (hipStream_t)stream_ptr;
    StreamAllocations mem{s, {}};
    TileGrid g{d_receivers, nx, ny, (nx + 15) / 16, cutoff_m};
    const uint32_t n_tiles = g.tiles_x * ((ny + 15) / 16);
    // Every tile's sources: counted, prefix-summed on the host, listed in one launch.
    uint32_t* d_counts = nullptr;
    SP_TRY(mem.alloc((void**)&d_counts, (size_t)n_tiles * 4));
    hipLaunchKernelGGL(tile_count_kernel, dim3(blocks_for(n_tiles)), dim3(256), 0, s, g, square->grid, n_tiles,
                       d_counts);
    SP_TRY(hipGetLastError());
    std::vector<uint32_t> counts(n_tiles);
    SP_TRY(hipMemcpyAsync(counts.data(), d_counts, (size_t)n_tiles * 4, hipMemcpyDeviceToHost, s));
    SP_TRY(hipStreamSynchronize(s));
    std::vector<uint32_t> offsets(n_tiles + 1, 0u);
    for (uint32_t t = 0; t < n_tiles; ++t) {
        if (counts[t] > CHUNK_ITEMS)
            return fail(err, errlen, "tile %u has %u sources in reach (at most %zu)", t, counts[t], CHUNK_ITEMS);
        if ((uint64_t)offsets[t] + counts[t] > 0xffffffffu) return fail(err, errlen, "too many (tile, source) items");
        offsets[t + 1] = offsets[t] + counts[t];
    }
    const size_t total_items = offsets[n_tiles];
    uint32_t *d_offsets, *d_item_tile, *d_item_source;
    SP_TRY(mem.alloc((void**)&d_offsets, ((size_t)n_tiles + 1) * 4));
    SP_TRY(mem.alloc((void**)&d_item_tile, std::max<size_t>(total_items, 1) * 4));
    SP_TRY(mem.alloc((void**)&d_item_source, std::max<size_t>(total_items, 1) * 4));
    SP_TRY(hipMemcpyAsync(d_offsets, offsets.data(), offsets.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(tile_list_kernel, dim3(blocks_for(n_tiles)), dim3(256), 0, s, g, square->grid, 0u, n_tiles,
                       d_offsets, d_item_tile, d_item_source);
    SP_TRY(hipGetLastError());
    // Chunks of whole tiles with at most CHUNK_ITEMS items.
    const size_t I = CHUNK_ITEMS, P = CHUNK_ITEMS * 256;
    const size_t max_blocks = I / COMBINE_ITEMS + (size_t)n_tiles;
    uint32_t *d_exact_list, *d_exact_count, *d_row_count, *d_row_of, *d_chunk, *d_block_tile, *d_block_item0,
        *d_tile_blocks;
    uint16_t* d_rows;
    uint8_t *d_flags, *d_exact_failed;
    float *d_source_alt, *d_att, *d_exact_energy;
    Partial* d_partials;
    SP_TRY(mem.alloc((void**)&d_exact_list, P * 4));
    SP_TRY(mem.alloc((void**)&d_exact_count, 4));
    SP_TRY(mem.alloc((void**)&d_row_count, 4));
    SP_TRY(mem.alloc((void**)&d_row_of, P * 4));
    SP_TRY(mem.alloc((void**)&d_rows, P * 2 * FEATURES * 2));
    SP_TRY(mem.alloc((void**)&d_flags, P));
    SP_TRY(mem.alloc((void**)&d_exact_failed, P));
    SP_TRY(mem.alloc((void**)&d_source_alt, P * 4));
    SP_TRY(mem.alloc((void**)&d_att, P * 2 * BANDS * 4));
    SP_TRY(mem.alloc((void**)&d_exact_energy, P * PERIODS * 4));
    SP_TRY(mem.alloc((void**)&d_partials, max_blocks * 256 * sizeof(Partial)));
    // Per chunk: the tiles' item offsets (relative to the chunk's tiles), the blocks' tile and first item.
    const size_t chunk_words = ((size_t)n_tiles + 1) + 2 * max_blocks + ((size_t)n_tiles + 1);
    SP_TRY(mem.alloc((void**)&d_chunk, chunk_words * 4));
    int cus = 0;
    SP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, gpu->ordinal));
    std::vector<std::vector<uint32_t>> staged;  // kept alive until the stream has read them
    uint32_t first = 0;
    while (first < n_tiles) {
        uint32_t last = first;
        while (last < n_tiles && offsets[last + 1] - offsets[first] <= I) ++last;
        const uint32_t m = last - first;
        const uint32_t item0 = offsets[first];
        const size_t items = offsets[last] - item0;
        // Host tables of the chunk: [tile_items m+1][block_tile B][block_item0 B][tile_blocks m+1].
        std::vector<uint32_t> tile_items(m + 1), block_tile, block_item0, tile_blocks(m + 1);
        for (uint32_t t = 0; t <= m; ++t) tile_items[t] = offsets[first + t];
        for (uint32_t t = 0; t < m; ++t) {
            tile_blocks[t] = (uint32_t)block_tile.size();
            for (uint32_t i = tile_items[t]; i < tile_items[t + 1]; i += COMBINE_ITEMS) {
                block_tile.push_back(t);
                block_item0.push_back(i);
            }
        }
        tile_blocks[m] = (uint32_t)block_tile.size();
        const uint32_t B = (uint32_t)block_tile.size();
        staged.emplace_back();
        std::vector<uint32_t>& h = staged.back();
        h.insert(h.end(), tile_items.begin(), tile_items.end());
        h.insert(h.end(), block_tile.begin(), block_tile.end());
        h.insert(h.end(), block_item0.begin(), block_item0.end());
        h.insert(h.end(), tile_blocks.begin(), tile_blocks.end());
        SP_TRY(hipMemcpyAsync(d_chunk, h.data(), h.size() * 4, hipMemcpyHostToDevice, s));
        const uint32_t* d_tile_items = d_chunk;
        d_block_tile = d_chunk + (m + 1);
        d_block_item0 = d_block_tile + B;
        d_tile_blocks = d_block_item0 + B;
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, s, d_exact_count);
        hipLaunchKernelGGL(zero_u32, dim3(1), dim3(1), 0, s, d_row_count);
        if (items) {
            hipLaunchKernelGGL(features_kernel, dim3((unsigned)items), dim3(256), 0, s, square->scene, g,
                               square->grid.sources, d_item_tile + item0, d_item_source + item0, d_rows, d_row_count,
                               d_row_of, d_flags, d_source_alt, d_exact_list, d_exact_count);
            SP_TRY(hipGetLastError());
            if (qm::mlp_infer_rows(mlp, d_rows, items * 256 * 2, d_row_count, d_att, s, err, errlen)) return 1;
            hipLaunchKernelGGL(exact_kernel, dim3((unsigned)std::max(cus, 1) * 4), dim3(QM_PATH_GROUP), 0, s, square->d_scene,
                               square->d_bands, g, square->grid.sources, d_item_tile + item0, d_item_source + item0,
                               d_exact_list, d_exact_count, d_exact_energy, d_exact_failed);
            SP_TRY(hipGetLastError());
        }
        if (B)
            hipLaunchKernelGGL(combine_kernel, dim3(B), dim3(256), 0, s, g, first, square->grid.sources, d_tile_items,
                               item0, d_block_tile, d_block_item0, d_item_source, d_flags, d_row_of, d_source_alt,
                               d_att, d_exact_energy, d_exact_failed, d_partials);
        hipLaunchKernelGGL(reduce_kernel, dim3(m), dim3(256), 0, s, g, first, d_tile_blocks, d_partials,
                           d_period_energy, d_paths, d_failures, d_surrogate_paths);
        SP_TRY(hipGetLastError());
        first = last;
    }
    SP_TRY(hipStreamSynchronize(s));
    return 0;
}
