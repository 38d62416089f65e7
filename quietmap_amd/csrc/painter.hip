// libqmpaint: kernels and the C ABI of include/qm_paint.h.
//
// Kernels (gfx950, wave64):
//   receivers_kernel  - paint-gpu kernels' point_receivers: ground + 4 m and the weather there
//   pairs_kernel      - Gpu::evaluate_pairs' ray transfers: a work-item per pair, the pairs taken
//                       in a coherent order (by source, then the receiver's Morton code on an 8 m
//                       grid, from pair_order_kernel and a device radix sort) so that a wave's
//                       rays leave one source for a compact block of receivers; each pair's
//                       results land at its own index
//   pair_order_kernel - the sort keys of that order
//   paint_kernel      - the grid painter: a persistent launch of 256-work-item groups, each group
//                       a 16x16 block of receivers (four 8x8 waves: neighbouring receivers in one
//                       wave follow near-identical rays), pulling tiles from per-XCD queues and
//                       stealing from the others when its own is empty; sources come from the
//                       bucket grid around the tile, uniform across the group (scalar loads).
//   crossings_kernel  - the exact crossing list of segments (parity hook)
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_eval.hpp"
#include "qm_internal.hpp"
#include "qm_queue.hpp"

using namespace qm;

namespace {

constexpr int TILE = 16;
constexpr int GROUP = QM_PATH_GROUP;

struct GridArgs {
    const qm_receiver* receivers;
    uint32_t nx, ny;
    uint32_t tiles_x, tiles_y, n_tiles;
    double cutoff;
    double* energy;
    uint32_t* paths;
    uint32_t* failures;
    uint32_t* counters;  // XCDS queue heads
};

// The allocation-free grid painter's memory (qm_paint_grid_device_noalloc): the XCDS queue heads
// in the first NOALLOC_HEAD_BYTES, then up to QM_NOALLOC_RECEIVERS receiver records. Each device
// holds its own copy until the process ends. Host code never passes this name to a kernel (on the
// host it is a stand-in, not a device address): it uses the address hipGetSymbolAddress returned
// on that device (qm_gpu::noalloc_scratch).
constexpr size_t NOALLOC_HEAD_BYTES = 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_HEAD_BYTES + sizeof(qm_receiver) * QM_NOALLOC_RECEIVERS];
static_assert(sizeof(uint32_t) * XCDS <= NOALLOC_HEAD_BYTES, "queue heads fit the scratch head");

}  // namespace

// paint/src/exact.rs:95-118 Point::with_receiver_properties (receiver_of in qm_eval.hpp).
__global__ __launch_bounds__(256) void receivers_kernel(Scene sc, const qm_point* points, uint32_t n,
                                                        qm_receiver* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = receiver_of(sc, points[i]);
}

#ifndef QM_PAINT_WAVES_PER_EU
#define QM_PAINT_WAVES_PER_EU 3
#endif
// Occupancy target of the path kernels (waves per SIMD): the register budget the compiler gets.
// At 3 the coarse paint kernel keeps 168 VGPRs and 7.8 KB of scratch a lane and runs about 1 %
// faster than at 2 (256 VGPRs, 7.6 KB) on the 1024 x 1024 receiver step.

// Gpu::evaluate_pairs' ray transfer (one ray per point pair): sums and point bands.
// FINE: the square holds fine terrain or the geographic frame's row table (the general
// instantiation, so the coarse kernel keeps its registers).
template <bool FINE, bool FOREST>
__global__ __launch_bounds__(QM_PATH_GROUP) __attribute__((amdgpu_waves_per_eu(QM_PAINT_WAVES_PER_EU))) void pairs_kernel(const Scene* __restrict__ scene, const BandTable* __restrict__ bands, const qm_receiver* receivers,
                                                    const qm_source* sources, const uint32_t* pair_point,
                                                    const uint32_t* pair_source, const float* pair_weight,
                                                    const uint32_t* order, uint32_t n, float* sums, uint32_t* failed,
                                                    float* point_bands, float* pair_bands) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t i = order ? order[k] : k;
    PathScratch w;
    PathResult r;
    const uint32_t p = pair_point[i];
    const qm_source src = sources[pair_source[i]];
    evaluate_path<FINE, FOREST>(sc, B, receivers[p], src, w, r);
    failed[i] = r.failed;
    float* out = sums + (size_t)i * QM_PAIR_SUMS;
    for (int q = 0; q < PERIODS; ++q) {
        float mixed = 0.0f;
        for (int b = 0; b < BANDS; ++b) mixed += r.bands[q][b];
        out[q] = r.failed ? 0.0f : mixed;
        out[3 + q] = r.failed ? 0.0f : r.states[0][q];
        out[6 + q] = r.failed ? 0.0f : r.states[1][q];
    }
    if (pair_bands) {
        float* dst = pair_bands + (size_t)i * PERIODS * BANDS;
        for (int q = 0; q < PERIODS; ++q)
            for (int b = 0; b < BANDS; ++b) dst[q * BANDS + b] = r.failed ? 0.0f : r.bands[q][b];
    }
    if (point_bands && !r.failed) {
        const float weight = pair_weight ? pair_weight[i] : 1.0f;
        float* dst = point_bands + ((size_t)p * QM_LAYERS + src.layer) * PERIODS * BANDS;
        for (int q = 0; q < PERIODS; ++q)
            for (int b = 0; b < BANDS; ++b) atomicAdd(dst + q * BANDS + b, weight * r.bands[q][b]);
    }
}

// The bits of x interleaved with zeros (bit b to bit 2b), for a 16-bit Morton code.
__device__ __forceinline__ uint32_t spread_bits16(uint32_t x) {
    x &= 0xffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// Sort keys of the pairs' coherent order: the source in the high word, then the Morton code of the
// receiver's position on an 8 m grid (modulo 524 km a side, which only a square far larger than any
// tile would wrap); the values are the pairs' own indices.
__global__ __launch_bounds__(256) void pair_order_kernel(const qm_receiver* receivers, const uint32_t* pair_point,
                                                         const uint32_t* pair_source, uint32_t n, uint64_t* keys,
                                                         uint32_t* indices) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const qm_receiver& r = receivers[pair_point[i]];
    const uint32_t qx = (uint32_t)(int64_t)floor(r.x_m * 0.125), qy = (uint32_t)(int64_t)floor(r.y_m * 0.125);
    keys[i] = ((uint64_t)pair_source[i] << 32) | (spread_bits16(qy) << 1) | spread_bits16(qx);
    indices[i] = i;
}

template <bool FINE, bool FOREST>
__global__ __launch_bounds__(GROUP) __attribute__((amdgpu_waves_per_eu(QM_PAINT_WAVES_PER_EU))) void paint_kernel(const Scene* __restrict__ scene, const BandTable* __restrict__ bands, SourceGrid sg, GridArgs g) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    __shared__ uint32_t s_tile;
    PathScratch w;
    const int lane = threadIdx.x;
    const int wave = lane >> 6, l = lane & 63;
    const int lx = (wave & 1) * 8 + (l & 7), ly = (wave >> 1) * 8 + (l >> 3);
    const int xcd = blockIdx.x % XCDS;
    phase_start();
    for (;;) {
        if (lane == 0) s_tile = next_tile(g.counters, g.n_tiles, xcd);
        __syncthreads();
        const uint32_t tile = s_tile;
        __syncthreads();
        if (tile >= g.n_tiles) break;
        const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
        const uint32_t ix = tx * TILE + lx, iy = ty * TILE + ly;
        const bool valid = ix < g.nx && iy < g.ny;
        const size_t index = valid ? (size_t)iy * g.nx + ix : 0;
        const qm_receiver& rcv = g.receivers[index];
        const double rx = rcv.x_m, ry = rcv.y_m;
        // The tile's box from its corner receivers (uniform across the group).
        const uint32_t ix0 = tx * TILE, iy0 = ty * TILE;
        const uint32_t ix1 = min(ix0 + TILE - 1, g.nx - 1), iy1 = min(iy0 + TILE - 1, g.ny - 1);
        const double bx0 = g.receivers[(size_t)iy0 * g.nx + ix0].x_m, by0 = g.receivers[(size_t)iy0 * g.nx + ix0].y_m;
        const double bx1 = g.receivers[(size_t)iy1 * g.nx + ix1].x_m, by1 = g.receivers[(size_t)iy1 * g.nx + ix1].y_m;
        const double lo_x = fmin(bx0, bx1) - g.cutoff, hi_x = fmax(bx0, bx1) + g.cutoff;
        const double lo_y = fmin(by0, by1) - g.cutoff, hi_y = fmax(by0, by1) + g.cutoff;
        const int kx0 = max((int)floor((lo_x - sg.origin_x) / sg.bucket_m), 0);
        const int kx1 = min((int)floor((hi_x - sg.origin_x) / sg.bucket_m), (int)sg.buckets_x - 1);
        const int ky0 = max((int)floor((lo_y - sg.origin_y) / sg.bucket_m), 0);
        const int ky1 = min((int)floor((hi_y - sg.origin_y) / sg.bucket_m), (int)sg.buckets_y - 1);
        double energy[PERIODS] = {0.0, 0.0, 0.0};
        uint32_t paths = 0, fails = 0;
        const double cut2 = g.cutoff * g.cutoff;
        for (int ky = ky0; ky <= ky1; ++ky) {
            const uint32_t s_begin = sg.bucket_first[(size_t)ky * sg.buckets_x + kx0];
            const uint32_t s_end = sg.bucket_first[(size_t)ky * sg.buckets_x + kx1 + 1];
            for (uint32_t s = s_begin; s < s_end; ++s) {
                const double sx = sg.sources[s].x_m, sy = sg.sources[s].y_m;
                // Whole-group reject: beyond the cutoff of the tile's box.
                const double ddx = sx < fmin(bx0, bx1) ? fmin(bx0, bx1) - sx : (sx > fmax(bx0, bx1) ? sx - fmax(bx0, bx1) : 0.0);
                const double ddy = sy < fmin(by0, by1) ? fmin(by0, by1) - sy : (sy > fmax(by0, by1) ? sy - fmax(by0, by1) : 0.0);
                if (ddx * ddx + ddy * ddy > cut2) continue;
                const double dx = sx - rx, dy = sy - ry;
                if (!valid || dx * dx + dy * dy > cut2) continue;
                PathResult r;
                evaluate_path<FINE, FOREST>(sc, B, rcv, sg.sources[s], w, r);
                ++paths;
                if (r.failed) {
                    ++fails;
                    continue;
                }
                for (int q = 0; q < PERIODS; ++q) {
                    // Pairwise sum of the eight bands, then FP64 across sources.
                    const float s01 = r.bands[q][0] + r.bands[q][1], s23 = r.bands[q][2] + r.bands[q][3];
                    const float s45 = r.bands[q][4] + r.bands[q][5], s67 = r.bands[q][6] + r.bands[q][7];
                    energy[q] += (double)((s01 + s23) + (s45 + s67));
                }
            }
        }
        if (valid) {
            for (int q = 0; q < PERIODS; ++q) g.energy[index * PERIODS + q] = energy[q];
            g.paths[index] = paths;
            g.failures[index] = fails;
        }
    }
    phase_flush();
}

__global__ __launch_bounds__(256) void crossings_kernel(Scene sc, const double* segments, uint32_t n, uint32_t capacity,
                                                        double* t, uint64_t* footprint, uint32_t* outline,
                                                        uint32_t* edge, uint32_t* flags, uint32_t* count,
                                                        uint32_t* failed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Crossing c[CROSS_CAP];
    int k = 0;
    failed[i] = find_crossings(sc, segments[4 * i], segments[4 * i + 1], segments[4 * i + 2],
                               segments[4 * i + 3], c, k);
    count[i] = (uint32_t)k;
    for (int j = 0; j < k && j < (int)capacity; ++j) {
        const size_t o = (size_t)i * capacity + j;
        t[o] = c[j].t;
        footprint[o] = c[j].footprint;
        outline[o] = c[j].outline;
        edge[o] = c[j].edge;
        flags[o] = c[j].flags & 3u;
    }
}

// =============================================================================================
// Host side.


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

#define QM_TRY(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return fail(err, errlen, "%s: %s", #call, hipGetErrorString(e_));  \
    } while (0)

BandTable make_band_table() {
    static const double F[8] = {63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0};
    static const double SHORT[8] = {0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 2.0, 3.0};
    static const double PER_M[8] = {0.02, 0.03, 0.04, 0.05, 0.06, 0.08, 0.09, 0.12};
    BandTable t;
    for (int b = 0; b < 8; ++b) {
        t.frequency[b] = (float)F[b];
        t.wavelength[b] = (float)(340.0 / F[b]);
        t.f_2_5[b] = (float)std::pow(F[b], 2.5);
        t.f_1_5[b] = (float)std::pow(F[b], 1.5);
        t.f_0_75[b] = (float)std::pow(F[b], 0.75);
        t.wavenumber[b] = (float)(2.0 * 3.141592653589793 * F[b] / 340.0);
        t.foliage_short_db[b] = (float)SHORT[b];
        t.foliage_db_per_m[b] = (float)PER_M[b];
        t.inv_wavelength[b] = (float)(F[b] / 340.0);
        t.inv_wavenumber[b] = (float)(340.0 / (2.0 * 3.141592653589793 * F[b]));
    }
    return t;
}

template <typename T>
int upload(qm_square* sq, const T* host, size_t count, T** out, char* err, size_t errlen) {
    *out = nullptr;
    if (count == 0) return 0;
    void* p = nullptr;
    QM_TRY(hipMalloc(&p, sizeof(T) * count));
    sq->allocations.push_back(p);
    sq->bytes += sizeof(T) * count;
    QM_TRY(hipMemcpy(p, host, sizeof(T) * count, hipMemcpyHostToDevice));
    *out = static_cast<T*>(p);
    return 0;
}

// A device buffer freed at scope end.
struct DeviceBuffer {
    void* p = nullptr;
    ~DeviceBuffer() {
        if (p) (void)hipFree(p);
    }
};

int device_alloc(DeviceBuffer& b, size_t bytes, char* err, size_t errlen) {
    QM_TRY(hipMalloc(&b.p, bytes ? bytes : 1));
    return 0;
}

int check_device(qm_gpu* gpu, char* err, size_t errlen) {
    QM_TRY(hipSetDevice(gpu->ordinal));
    return 0;
}

// The device address of `scratch` on each device, looked up once per device (with that device
// current) and kept until the process ends.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
void* g_scratch_address[MAX_DEVICES];

int scratch_address(int ordinal, void** out, char* err, size_t errlen) {
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (ordinal < 0 || ordinal >= MAX_DEVICES) return fail(err, errlen, "device ordinal %d beyond %d", ordinal, MAX_DEVICES);
    if (!g_scratch_address[ordinal]) {
        void* p = nullptr;
        QM_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(err, errlen, "hipGetSymbolAddress returned no address for the painter's scratch");
        g_scratch_address[ordinal] = p;
    }
    *out = g_scratch_address[ordinal];
    return 0;
}

// Device memory freed on the stream at scope end (hipMallocAsync's pool).
struct AsyncBuffer {
    void* p = nullptr;
    hipStream_t stream = nullptr;
    ~AsyncBuffer() {
        if (p) (void)hipFreeAsync(p, stream);
    }
};

// The grid painter over prepared receivers (nx*ny, row-major), its XCDS queue heads at
// `d_counters` cleared first by a kernel on the stream (no memset node where a caller captures).
int launch_paint(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers, uint32_t nx, uint32_t ny,
                 double cutoff_m, double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures,
                 uint32_t* d_counters, hipStream_t stream, char* err, size_t errlen) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, stream, d_counters, (uint32_t)XCDS);
    QM_TRY(hipGetLastError());
    GridArgs g;
    g.receivers = d_receivers;
    g.nx = nx;
    g.ny = ny;
    g.tiles_x = (nx + TILE - 1) / TILE;
    g.tiles_y = (ny + TILE - 1) / TILE;
    g.n_tiles = g.tiles_x * g.tiles_y;
    g.cutoff = cutoff_m;
    g.energy = d_period_energy;
    g.paths = d_paths;
    g.failures = d_failures;
    g.counters = d_counters;
    int cus = 0;
    QM_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, gpu->ordinal));
    int per_cu = 0;
    // The general instantiation: fine terrain or the geographic frame's row table.
    const bool fine = square->scene.terrain.fine_detail != nullptr || square->scene.terrain.row_y_m != nullptr;
    // Coarse squares without forest cover take the instantiation without the foliage term.
    const void* kernel = fine ? reinterpret_cast<const void*>(paint_kernel<true, true>)
                         : square->has_forest ? reinterpret_cast<const void*>(paint_kernel<false, true>)
                                              : reinterpret_cast<const void*>(paint_kernel<false, false>);
    QM_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, GROUP, 0));
    unsigned groups = (unsigned)std::max(1, cus * std::max(per_cu, 1));
    groups = std::min<unsigned>(groups, g.n_tiles);
    groups = std::max<unsigned>(groups, 1);
    if (fine)
        hipLaunchKernelGGL((paint_kernel<true, true>), dim3(groups), dim3(GROUP), 0, stream, square->d_scene,
                           square->d_bands, square->grid, g);
    else if (square->has_forest)
        hipLaunchKernelGGL((paint_kernel<false, true>), dim3(groups), dim3(GROUP), 0, stream, square->d_scene,
                           square->d_bands, square->grid, g);
    else
        hipLaunchKernelGGL((paint_kernel<false, false>), dim3(groups), dim3(GROUP), 0, stream, square->d_scene,
                           square->d_bands, square->grid, g);
    QM_TRY(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int qm_gpu_new(int ordinal, qm_gpu** out, char* err, size_t errlen) {
    *out = nullptr;
    int count = 0;
    QM_TRY(hipGetDeviceCount(&count));
    if (ordinal < 0 || ordinal >= count) return fail(err, errlen, "no GPU with ordinal %d (%d found)", ordinal, count);
    hipDeviceProp_t prop;
    QM_TRY(hipGetDeviceProperties(&prop, ordinal));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(err, errlen, "GPU %d is %s; libqmpaint is built for gfx950", ordinal, prop.gcnArchName);
    QM_TRY(hipSetDevice(ordinal));
    void* noalloc_scratch = nullptr;
    if (scratch_address(ordinal, &noalloc_scratch, err, errlen)) return 1;
    qm_gpu* gpu = new (std::nothrow) qm_gpu();
    if (!gpu) return fail(err, errlen, "out of host memory");
    gpu->ordinal = ordinal;
    gpu->noalloc_scratch = noalloc_scratch;
    std::snprintf(gpu->name, sizeof(gpu->name), "%s (%s)", prop.name, prop.gcnArchName);
    gpu->bands = make_band_table();
    *out = gpu;
    return 0;
}

int qm_gpu_name(const qm_gpu* gpu, char* out, size_t len) {
    if (!gpu || !out || !len) return 1;
    std::snprintf(out, len, "%s", gpu->name);
    return 0;
}

void qm_gpu_free(qm_gpu* gpu) { delete gpu; }

int qm_square_upload(qm_gpu* gpu, const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                     const qm_source* sources, size_t n_sources, double bucket_m, qm_square** out, char* err,
                     size_t errlen) {
    *out = nullptr;
    if (!gpu || !terrain || !obstacles || !weather) return fail(err, errlen, "null argument");
    if (terrain->rows < 2 || terrain->columns < 2 || !(terrain->spacing_m > 0.0))
        return fail(err, errlen, "terrain lattice must have 2x2 nodes and a positive spacing");
    if (weather->nodes_per_degree) {
        if (weather->rows != 180u * weather->nodes_per_degree + 1 || weather->columns != 360u * weather->nodes_per_degree)
            return fail(err, errlen, "world weather table must be (180 npd + 1) x 360 npd nodes");
        if (!(weather->east_m_per_unit > 0.0) || !(weather->north_m_per_unit > 0.0) ||
            !std::isfinite(weather->frame_x) || !std::isfinite(weather->frame_y))
            return fail(err, errlen, "weather frame must have positive metres per unit and a finite origin");
    } else if (weather->rows < 1 || weather->columns < 1 || !(weather->spacing_m > 0.0)) {
        return fail(err, errlen, "weather lattice must have a node and a positive spacing");
    }
    if (!(obstacles->metres_per_step > 0.0) || obstacles->cells_x == 0 || obstacles->cells_y == 0)
        return fail(err, errlen, "obstacle grid must have cells and a positive step");
    if (!(bucket_m > 0.0)) return fail(err, errlen, "bucket size must be positive");
    if (check_device(gpu, err, errlen)) return 1;
    qm_square* sq = new (std::nothrow) qm_square();
    if (!sq) return fail(err, errlen, "out of host memory");
    sq->gpu = gpu;
    sq->bytes = 0;
    auto cleanup = [&](int rc) {
        if (rc) qm_square_free(sq);
        return rc;
    };
    const size_t nodes = (size_t)terrain->rows * terrain->columns;
    uint32_t* d_nodes;
    uint8_t* d_canopy;
    if (upload(sq, terrain->nodes, nodes, &d_nodes, err, errlen)) return cleanup(1);
    std::vector<uint8_t> zero_canopy;
    const uint8_t* canopy = terrain->canopy_m;
    if (!canopy) {
        zero_canopy.assign(nodes, 0);
        canopy = zero_canopy.data();
    }
    if (upload(sq, canopy, nodes, &d_canopy, err, errlen)) return cleanup(1);
    sq->scene.terrain = *terrain;
    sq->has_forest = false;
    for (size_t k = 0; k < nodes && !sq->has_forest; ++k) sq->has_forest = (terrain->nodes[k] >> 24) != 0;
    sq->scene.terrain.nodes = d_nodes;
    sq->scene.terrain.canopy_m = d_canopy;
    sq->scene.terrain.fine_detail = nullptr;
    sq->scene.terrain.fine_blocks = nullptr;
    sq->scene.terrain.row_y_m = nullptr;
    if (!(terrain->column_m >= 0.0) || !std::isfinite(terrain->column_m))
        return cleanup(fail(err, errlen, "terrain column spacing must be finite and non-negative"));
    if (terrain->row_y_m) {
        for (uint32_t r = 0; r < terrain->rows; ++r)
            if (!std::isfinite(terrain->row_y_m[r]) || (r && !(terrain->row_y_m[r] > terrain->row_y_m[r - 1])))
                return cleanup(fail(err, errlen, "terrain row table not finite and ascending at row %u", r));
        double* d_rows;
        if (upload(sq, terrain->row_y_m, terrain->rows, &d_rows, err, errlen)) return cleanup(1);
        sq->scene.terrain.row_y_m = d_rows;
    }
    if (terrain->fine_detail) {
        if (!terrain->fine_blocks) return cleanup(fail(err, errlen, "fine terrain without its blocks"));
        const size_t fine = (size_t)(QM_FINE_PER_NODE * (terrain->rows - 1) + 1) *
                            (size_t)(QM_FINE_PER_NODE * (terrain->columns - 1) + 1);
        const size_t blocks = (size_t)((terrain->rows - 2) / QM_FINE_BLOCK_CELLS + 1) *
                              (size_t)((terrain->columns - 2) / QM_FINE_BLOCK_CELLS + 1);
        int16_t* d_detail;
        uint8_t* d_blocks;
        if (upload(sq, terrain->fine_detail, fine, &d_detail, err, errlen)) return cleanup(1);
        if (upload(sq, terrain->fine_blocks, blocks, &d_blocks, err, errlen)) return cleanup(1);
        sq->scene.terrain.fine_detail = d_detail;
        sq->scene.terrain.fine_blocks = d_blocks;
    }
    // Obstacles: validate the index before it reaches the device.
    const size_t cells = (size_t)obstacles->cells_x * obstacles->cells_y;
    if (obstacles->cell_first_run[cells] != obstacles->runs_count)
        return cleanup(fail(err, errlen, "obstacle cell index does not end at the run count"));
    for (size_t c = 0; c < cells; ++c)
        if (obstacles->cell_first_run[c] > obstacles->cell_first_run[c + 1])
            return cleanup(fail(err, errlen, "obstacle cell index not ascending at cell %zu", c));
    for (uint32_t r = 0; r < obstacles->runs_count; ++r) {
        const qm_run& run = obstacles->runs[r];
        if (run.outline >= obstacles->outlines_count) return cleanup(fail(err, errlen, "run %u: bad outline", r));
        const qm_outline& o = obstacles->outlines[run.outline];
        if ((size_t)run.first_edge + run.edge_count + 1 > o.vertex_count)
            return cleanup(fail(err, errlen, "run %u: edges beyond its outline", r));
    }
    for (uint32_t i = 0; i < obstacles->outlines_count; ++i) {
        const qm_outline& o = obstacles->outlines[i];
        if ((size_t)o.first_vertex + o.vertex_count > obstacles->vertices_count || o.vertex_count < 2)
            return cleanup(fail(err, errlen, "outline %u: vertices out of range", i));
    }
    uint32_t* d_first;
    qm_run* d_runs;
    qm_outline* d_outlines;
    int32_t* d_vertices;
    if (upload(sq, obstacles->cell_first_run, cells + 1, &d_first, err, errlen)) return cleanup(1);
    if (upload(sq, obstacles->runs, obstacles->runs_count, &d_runs, err, errlen)) return cleanup(1);
    if (upload(sq, obstacles->outlines, obstacles->outlines_count, &d_outlines, err, errlen)) return cleanup(1);
    if (upload(sq, obstacles->vertices, (size_t)obstacles->vertices_count * 2, &d_vertices, err, errlen))
        return cleanup(1);
    sq->scene.obstacles = *obstacles;
    sq->scene.obstacles.cell_first_run = d_first;
    sq->scene.obstacles.runs = d_runs;
    sq->scene.obstacles.outlines = d_outlines;
    sq->scene.obstacles.vertices = d_vertices;
    {
        // Occupancy words: a bit per cell holding runs, 8 x 8 cells a word.
        const uint32_t bx = (obstacles->cells_x + 7) / 8, by = (obstacles->cells_y + 7) / 8;
        std::vector<uint64_t> words((size_t)bx * by, 0);
        for (uint32_t y = 0; y < obstacles->cells_y; ++y)
            for (uint32_t x = 0; x < obstacles->cells_x; ++x) {
                const size_t c = (size_t)y * obstacles->cells_x + x;
                if (obstacles->cell_first_run[c + 1] > obstacles->cell_first_run[c])
                    words[(size_t)(y >> 3) * bx + (x >> 3)] |= 1ull << (8 * (y & 7) + (x & 7));
            }
        uint64_t* d_words = nullptr;
        if (!words.empty() && upload(sq, words.data(), words.size(), &d_words, err, errlen)) return cleanup(1);
        sq->scene.cell_blocks = d_words;
        sq->scene.blocks_x = bx;
        sq->scene.blocks_y = by;
        std::vector<uint32_t> run_vertex(obstacles->runs_count);
        for (uint32_t r = 0; r < obstacles->runs_count; ++r)
            run_vertex[r] = obstacles->outlines[obstacles->runs[r].outline].first_vertex + obstacles->runs[r].first_edge;
        uint32_t* d_run_vertex = nullptr;
        if (!run_vertex.empty() && upload(sq, run_vertex.data(), run_vertex.size(), &d_run_vertex, err, errlen))
            return cleanup(1);
        sq->scene.run_vertex = d_run_vertex;
    }
    uint8_t* d_weather;
    if (upload(sq, weather->nodes, (size_t)weather->rows * weather->columns * QM_WEATHER_NODE_BYTES, &d_weather, err,
               errlen))
        return cleanup(1);
    sq->scene.weather = *weather;
    sq->scene.weather.nodes = d_weather;
    // Sources: input order for pairs, and a counting sort into buckets for the painter.
    sq->n_sources = n_sources;
    if (upload(sq, sources, n_sources, &sq->d_sources_by_input, err, errlen)) return cleanup(1);
    double min_x = 0.0, min_y = 0.0, max_x = 0.0, max_y = 0.0;
    for (size_t i = 0; i < n_sources; ++i) {
        if (!std::isfinite(sources[i].x_m) || !std::isfinite(sources[i].y_m))
            return cleanup(fail(err, errlen, "source %zu: position not finite", i));
        if (sources[i].layer >= QM_LAYERS) return cleanup(fail(err, errlen, "source %zu: layer out of range", i));
        if (i == 0 || sources[i].x_m < min_x) min_x = sources[i].x_m;
        if (i == 0 || sources[i].y_m < min_y) min_y = sources[i].y_m;
        if (i == 0 || sources[i].x_m > max_x) max_x = sources[i].x_m;
        if (i == 0 || sources[i].y_m > max_y) max_y = sources[i].y_m;
    }
    const uint32_t bx = (uint32_t)std::floor((max_x - min_x) / bucket_m) + 1;
    const uint32_t by = (uint32_t)std::floor((max_y - min_y) / bucket_m) + 1;
    if ((uint64_t)bx * by > (1u << 26)) return cleanup(fail(err, errlen, "bucket grid too large"));
    std::vector<uint32_t> first((size_t)bx * by + 1, 0);
    std::vector<uint32_t> bucket(n_sources);
    for (size_t i = 0; i < n_sources; ++i) {
        const uint32_t kx = std::min((uint32_t)std::floor((sources[i].x_m - min_x) / bucket_m), bx - 1);
        const uint32_t ky = std::min((uint32_t)std::floor((sources[i].y_m - min_y) / bucket_m), by - 1);
        bucket[i] = ky * bx + kx;
        ++first[bucket[i] + 1];
    }
    for (size_t k = 0; k < (size_t)bx * by; ++k) first[k + 1] += first[k];
    std::vector<qm_source> sorted(n_sources);
    {
        std::vector<uint32_t> fill(first.begin(), first.end() - 1);
        for (size_t i = 0; i < n_sources; ++i) sorted[fill[bucket[i]]++] = sources[i];
    }
    qm_source* d_sorted;
    uint32_t* d_bucket_first;
    if (upload(sq, sorted.data(), n_sources, &d_sorted, err, errlen)) return cleanup(1);
    if (upload(sq, first.data(), first.size(), &d_bucket_first, err, errlen)) return cleanup(1);
    sq->grid.sources = d_sorted;
    sq->grid.bucket_first = d_bucket_first;
    sq->grid.origin_x = min_x;
    sq->grid.origin_y = min_y;
    sq->grid.bucket_m = bucket_m;
    sq->grid.buckets_x = bx;
    sq->grid.buckets_y = by;
    // The scene's and the band table's records in device memory: the path kernels read them
    // through pointers (scalar loads where used) rather than holding them in their arguments.
    qm::Scene* d_scene;
    qm::BandTable* d_bands;
    if (upload(sq, &sq->scene, 1, &d_scene, err, errlen)) return cleanup(1);
    if (upload(sq, &gpu->bands, 1, &d_bands, err, errlen)) return cleanup(1);
    sq->d_scene = d_scene;
    sq->d_bands = d_bands;
    *out = sq;
    return 0;
}

size_t qm_square_bytes(const qm_square* square) { return square ? square->bytes : 0; }

void qm_square_free(qm_square* square) {
    if (!square) return;
    if (square->gpu) (void)hipSetDevice(square->gpu->ordinal);
    for (void* p : square->allocations) (void)hipFree(p);
    delete square;
}

static int prepare_receivers(const qm_square* square, const qm_point* d_points, size_t n, qm_receiver* d_receivers,
                             hipStream_t stream, char* err, size_t errlen) {
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(receivers_kernel, dim3(blocks), dim3(256), 0, stream, square->scene, d_points, (uint32_t)n,
                       d_receivers);
    QM_TRY(hipGetLastError());
    return 0;
}

int qm_evaluate_pairs(qm_gpu* gpu, const qm_square* square, const qm_point* points, size_t n_points,
                      const uint32_t* pair_point, const uint32_t* pair_source, const float* pair_weight, size_t n_pairs,
                      float* sums, uint32_t* failed, float* point_bands, float* pair_bands, char* err,
                      size_t errlen) {
    if (!gpu || !square || (n_pairs && (!pair_point || !pair_source || !sums || !failed)))
        return fail(err, errlen, "null argument");
    if (n_pairs > 0x7fffffffu || n_points > 0xffffffffu) return fail(err, errlen, "too many pairs or points");
    for (size_t i = 0; i < n_points; ++i)
        if (!std::isfinite(points[i].x_m) || !std::isfinite(points[i].y_m))
            return fail(err, errlen, "point %zu: position not finite", i);
    for (size_t i = 0; i < n_pairs; ++i) {
        if (pair_point[i] >= n_points) return fail(err, errlen, "pair %zu: point %u out of range", i, pair_point[i]);
        if (pair_source[i] >= square->n_sources)
            return fail(err, errlen, "pair %zu: source %u out of range", i, pair_source[i]);
    }
    if (check_device(gpu, err, errlen)) return 1;
    // With pairs, the point bands come back whole from the card (cleared there); without, zeros.
    if (n_pairs == 0) {
        if (point_bands) std::memset(point_bands, 0, sizeof(float) * n_points * QM_LAYERS * PERIODS * BANDS);
        return 0;
    }
    hipStream_t stream;
    QM_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{stream};
    DeviceBuffer d_points, d_receivers, d_pp, d_ps, d_pw, d_sums, d_failed, d_bands, d_pair_bands;
    if (device_alloc(d_points, sizeof(qm_point) * n_points, err, errlen) ||
        device_alloc(d_receivers, sizeof(qm_receiver) * n_points, err, errlen) ||
        device_alloc(d_pp, sizeof(uint32_t) * n_pairs, err, errlen) ||
        device_alloc(d_ps, sizeof(uint32_t) * n_pairs, err, errlen) ||
        device_alloc(d_sums, sizeof(float) * n_pairs * QM_PAIR_SUMS, err, errlen) ||
        device_alloc(d_failed, sizeof(uint32_t) * n_pairs, err, errlen))
        return 1;
    if (pair_weight && device_alloc(d_pw, sizeof(float) * n_pairs, err, errlen)) return 1;
    const size_t band_floats = n_points * QM_LAYERS * PERIODS * BANDS;
    if (point_bands && device_alloc(d_bands, sizeof(float) * band_floats, err, errlen)) return 1;
    if (pair_bands && device_alloc(d_pair_bands, sizeof(float) * n_pairs * PERIODS * BANDS, err, errlen)) return 1;
    QM_TRY(hipMemcpyAsync(d_points.p, points, sizeof(qm_point) * n_points, hipMemcpyHostToDevice, stream));
    QM_TRY(hipMemcpyAsync(d_pp.p, pair_point, sizeof(uint32_t) * n_pairs, hipMemcpyHostToDevice, stream));
    QM_TRY(hipMemcpyAsync(d_ps.p, pair_source, sizeof(uint32_t) * n_pairs, hipMemcpyHostToDevice, stream));
    if (pair_weight)
        QM_TRY(hipMemcpyAsync(d_pw.p, pair_weight, sizeof(float) * n_pairs, hipMemcpyHostToDevice, stream));
    if (point_bands) QM_TRY(hipMemsetAsync(d_bands.p, 0, sizeof(float) * band_floats, stream));
    if (prepare_receivers(square, (const qm_point*)d_points.p, n_points, (qm_receiver*)d_receivers.p, stream, err,
                          errlen))
        return 1;
    const unsigned blocks = (unsigned)((n_pairs + 255) / 256);
    // The coherent order: keys, then a radix sort over the source's significant bits and the
    // 32-bit Morton code (the sort is stable, so pairs of one receiver and source keep their order).
    DeviceBuffer d_keys, d_keys_sorted, d_index, d_order, d_sort_temp;
    if (device_alloc(d_keys, sizeof(uint64_t) * n_pairs, err, errlen) ||
        device_alloc(d_keys_sorted, sizeof(uint64_t) * n_pairs, err, errlen) ||
        device_alloc(d_index, sizeof(uint32_t) * n_pairs, err, errlen) ||
        device_alloc(d_order, sizeof(uint32_t) * n_pairs, err, errlen))
        return 1;
    hipLaunchKernelGGL(pair_order_kernel, dim3(blocks), dim3(256), 0, stream, (const qm_receiver*)d_receivers.p,
                       (const uint32_t*)d_pp.p, (const uint32_t*)d_ps.p, (uint32_t)n_pairs, (uint64_t*)d_keys.p,
                       (uint32_t*)d_index.p);
    QM_TRY(hipGetLastError());
    int source_bits = 1;
    while (source_bits < 32 && (square->n_sources >> source_bits) != 0) ++source_bits;
    size_t sort_bytes = 0;
    QM_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const uint64_t*)d_keys.p, (uint64_t*)d_keys_sorted.p,
                                              (const uint32_t*)d_index.p, (uint32_t*)d_order.p, (int)n_pairs, 0,
                                              32 + source_bits, stream));
    if (device_alloc(d_sort_temp, sort_bytes, err, errlen)) return 1;
    QM_TRY(hipcub::DeviceRadixSort::SortPairs(d_sort_temp.p, sort_bytes, (const uint64_t*)d_keys.p,
                                              (uint64_t*)d_keys_sorted.p, (const uint32_t*)d_index.p,
                                              (uint32_t*)d_order.p, (int)n_pairs, 0, 32 + source_bits, stream));
    // The general instantiation for fine terrain or a row table; coarse squares without forest
    // cover without the foliage term.
    auto kernel = (square->scene.terrain.fine_detail || square->scene.terrain.row_y_m) ? pairs_kernel<true, true>
                  : square->has_forest                                                  ? pairs_kernel<false, true>
                                                                                        : pairs_kernel<false, false>;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), 0, stream, square->d_scene, square->d_bands,
                       (const qm_receiver*)d_receivers.p, square->d_sources_by_input, (const uint32_t*)d_pp.p,
                       (const uint32_t*)d_ps.p, (const float*)(pair_weight ? d_pw.p : nullptr),
                       (const uint32_t*)d_order.p, (uint32_t)n_pairs, (float*)d_sums.p, (uint32_t*)d_failed.p, (float*)(point_bands ? d_bands.p : nullptr),
                       (float*)(pair_bands ? d_pair_bands.p : nullptr));
    QM_TRY(hipGetLastError());
    QM_TRY(hipMemcpyAsync(sums, d_sums.p, sizeof(float) * n_pairs * QM_PAIR_SUMS, hipMemcpyDeviceToHost, stream));
    QM_TRY(hipMemcpyAsync(failed, d_failed.p, sizeof(uint32_t) * n_pairs, hipMemcpyDeviceToHost, stream));
    if (point_bands)
        QM_TRY(hipMemcpyAsync(point_bands, d_bands.p, sizeof(float) * band_floats, hipMemcpyDeviceToHost, stream));
    if (pair_bands)
        QM_TRY(hipMemcpyAsync(pair_bands, d_pair_bands.p, sizeof(float) * n_pairs * PERIODS * BANDS,
                              hipMemcpyDeviceToHost, stream));
    QM_TRY(hipStreamSynchronize(stream));
    return 0;
}

int qm_paint_grid_device(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, uint32_t nx, uint32_t ny,
                         double cutoff_m, double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures,
                         void* stream_ptr, char* err, size_t errlen) {
    if (!gpu || !square || !d_points || !d_period_energy || !d_paths || !d_failures)
        return fail(err, errlen, "null argument");
    if (!(cutoff_m >= 0.0)) return fail(err, errlen, "cutoff must be non-negative");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    if (check_device(gpu, err, errlen)) return 1;
    hipStream_t stream = (hipStream_t)stream_ptr;
    AsyncBuffer receivers, counters;
    receivers.stream = counters.stream = stream;
    QM_TRY(hipMallocAsync(&receivers.p, sizeof(qm_receiver) * n, stream));
    QM_TRY(hipMallocAsync(&counters.p, sizeof(uint32_t) * XCDS, stream));
    if (prepare_receivers(square, d_points, n, (qm_receiver*)receivers.p, stream, err, errlen)) return 1;
    return launch_paint(gpu, square, (const qm_receiver*)receivers.p, nx, ny, cutoff_m, d_period_energy, d_paths,
                        d_failures, (uint32_t*)counters.p, stream, err, errlen);
}

int qm_paint_grid_device_noalloc(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, uint32_t nx,
                                 uint32_t ny, double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                                 uint32_t* d_failures, void* stream_ptr, char* err, size_t errlen) {
    if (!gpu || !square || !d_points || !d_period_energy || !d_paths || !d_failures)
        return fail(err, errlen, "null argument");
    if (!(cutoff_m >= 0.0)) return fail(err, errlen, "cutoff must be non-negative");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    if (n > (size_t)QM_NOALLOC_RECEIVERS)
        return fail(err, errlen, "%zu receivers exceed the allocation-free painter's %d", n, QM_NOALLOC_RECEIVERS);
    if (!gpu->noalloc_scratch) return fail(err, errlen, "no scratch address for device %d", gpu->ordinal);
    if (check_device(gpu, err, errlen)) return 1;
    hipStream_t stream = (hipStream_t)stream_ptr;
    unsigned char* base = static_cast<unsigned char*>(gpu->noalloc_scratch);
    uint32_t* d_counters = reinterpret_cast<uint32_t*>(base);
    qm_receiver* d_receivers = reinterpret_cast<qm_receiver*>(base + NOALLOC_HEAD_BYTES);
    if (prepare_receivers(square, d_points, n, d_receivers, stream, err, errlen)) return 1;
    return launch_paint(gpu, square, d_receivers, nx, ny, cutoff_m, d_period_energy, d_paths, d_failures, d_counters,
                        stream, err, errlen);
}

int qm_receivers_device(qm_gpu* gpu, const qm_square* square, const qm_point* d_points, size_t n,
                        qm_receiver* d_receivers, void* stream, char* err, size_t errlen) {
    if (!gpu || !square || (n && (!d_points || !d_receivers))) return fail(err, errlen, "null argument");
    if (n > 0xffffffffu) return fail(err, errlen, "too many points");
    if (check_device(gpu, err, errlen)) return 1;
    return prepare_receivers(square, d_points, n, d_receivers, (hipStream_t)stream, err, errlen);
}

int qm_paint_receivers_device(qm_gpu* gpu, const qm_square* square, const qm_receiver* d_receivers, uint32_t nx,
                              uint32_t ny, double cutoff_m, double* d_period_energy, uint32_t* d_paths,
                              uint32_t* d_failures, void* stream_ptr, char* err, size_t errlen) {
    if (!gpu || !square || !d_receivers || !d_period_energy || !d_paths || !d_failures)
        return fail(err, errlen, "null argument");
    if (!(cutoff_m >= 0.0)) return fail(err, errlen, "cutoff must be non-negative");
    if ((size_t)nx * ny == 0) return 0;
    if (check_device(gpu, err, errlen)) return 1;
    hipStream_t stream = (hipStream_t)stream_ptr;
    AsyncBuffer counters;
    counters.stream = stream;
    QM_TRY(hipMallocAsync(&counters.p, sizeof(uint32_t) * XCDS, stream));
    return launch_paint(gpu, square, d_receivers, nx, ny, cutoff_m, d_period_energy, d_paths, d_failures,
                        (uint32_t*)counters.p, stream, err, errlen);
}

int qm_paint_grid(qm_gpu* gpu, const qm_square* square, const qm_point* points, uint32_t nx, uint32_t ny,
                  double cutoff_m, double* period_energy, uint32_t* paths, uint32_t* failures, char* err,
                  size_t errlen) {
    if (!gpu || !square || !points || !period_energy || !paths || !failures) return fail(err, errlen, "null argument");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(points[i].x_m) || !std::isfinite(points[i].y_m))
            return fail(err, errlen, "point %zu: position not finite", i);
    if (check_device(gpu, err, errlen)) return 1;
    hipStream_t stream;
    QM_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{stream};
    DeviceBuffer d_points, d_energy, d_paths, d_failures;
    if (device_alloc(d_points, sizeof(qm_point) * n, err, errlen) ||
        device_alloc(d_energy, sizeof(double) * n * PERIODS, err, errlen) ||
        device_alloc(d_paths, sizeof(uint32_t) * n, err, errlen) ||
        device_alloc(d_failures, sizeof(uint32_t) * n, err, errlen))
        return 1;
    QM_TRY(hipMemcpyAsync(d_points.p, points, sizeof(qm_point) * n, hipMemcpyHostToDevice, stream));
    if (qm_paint_grid_device(gpu, square, (const qm_point*)d_points.p, nx, ny, cutoff_m, (double*)d_energy.p,
                             (uint32_t*)d_paths.p, (uint32_t*)d_failures.p, stream, err, errlen))
        return 1;
    QM_TRY(hipMemcpyAsync(period_energy, d_energy.p, sizeof(double) * n * PERIODS, hipMemcpyDeviceToHost, stream));
    QM_TRY(hipMemcpyAsync(paths, d_paths.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
    QM_TRY(hipMemcpyAsync(failures, d_failures.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, stream));
    QM_TRY(hipStreamSynchronize(stream));
    return 0;
}

int qm_crossings(qm_gpu* gpu, const qm_square* square, const double* segments, size_t n_segments, uint32_t capacity,
                 double* t, uint64_t* footprint, uint32_t* outline, uint32_t* edge, uint32_t* flags, uint32_t* count,
                 uint32_t* failed, char* err, size_t errlen) {
    if (!gpu || !square || (n_segments && (!segments || !t || !footprint || !outline || !edge || !flags || !count || !failed)))
        return fail(err, errlen, "null argument");
    if (n_segments == 0) return 0;
    if (capacity == 0 || capacity > (uint32_t)CROSS_CAP) return fail(err, errlen, "capacity must be 1..%d", CROSS_CAP);
    for (size_t i = 0; i < 4 * n_segments; ++i)
        if (!std::isfinite(segments[i])) return fail(err, errlen, "segment %zu: not finite", i / 4);
    if (check_device(gpu, err, errlen)) return 1;
    DeviceBuffer d_seg, d_t, d_f, d_o, d_e, d_fl, d_c, d_failed;
    const size_t slots = n_segments * capacity;
    if (device_alloc(d_seg, sizeof(double) * 4 * n_segments, err, errlen) ||
        device_alloc(d_t, sizeof(double) * slots, err, errlen) ||
        device_alloc(d_f, sizeof(uint64_t) * slots, err, errlen) ||
        device_alloc(d_o, sizeof(uint32_t) * slots, err, errlen) ||
        device_alloc(d_e, sizeof(uint32_t) * slots, err, errlen) ||
        device_alloc(d_fl, sizeof(uint32_t) * slots, err, errlen) ||
        device_alloc(d_c, sizeof(uint32_t) * n_segments, err, errlen) ||
        device_alloc(d_failed, sizeof(uint32_t) * n_segments, err, errlen))
        return 1;
    QM_TRY(hipMemcpy(d_seg.p, segments, sizeof(double) * 4 * n_segments, hipMemcpyHostToDevice));
    const unsigned blocks = (unsigned)((n_segments + 255) / 256);
    hipLaunchKernelGGL(crossings_kernel, dim3(blocks), dim3(256), 0, 0, square->scene, (const double*)d_seg.p,
                       (uint32_t)n_segments, capacity, (double*)d_t.p, (uint64_t*)d_f.p, (uint32_t*)d_o.p,
                       (uint32_t*)d_e.p, (uint32_t*)d_fl.p, (uint32_t*)d_c.p, (uint32_t*)d_failed.p);
    QM_TRY(hipGetLastError());
    QM_TRY(hipMemcpy(t, d_t.p, sizeof(double) * slots, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(footprint, d_f.p, sizeof(uint64_t) * slots, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(outline, d_o.p, sizeof(uint32_t) * slots, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(edge, d_e.p, sizeof(uint32_t) * slots, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(flags, d_fl.p, sizeof(uint32_t) * slots, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(count, d_c.p, sizeof(uint32_t) * n_segments, hipMemcpyDeviceToHost));
    QM_TRY(hipMemcpy(failed, d_failed.p, sizeof(uint32_t) * n_segments, hipMemcpyDeviceToHost));
    return 0;
}

#ifdef QM_PHASE_CLOCKS
// The instrumented build only (not part of the ABI in include/): the phase clocks' totals of this
// translation unit's path kernels (qm::PathPhase order), cleared afterwards when `reset`.
int qm_phase_clocks(unsigned long long* out, int reset, char* err, size_t errlen) {
    if (!out) return fail(err, errlen, "out must not be null");
    QM_TRY(hipDeviceSynchronize());
    QM_TRY(hipMemcpyFromSymbol(out, HIP_SYMBOL(qm::qm_phase_cycles), sizeof(unsigned long long) * qm::PHASE_COUNT));
    if (reset) {
        const unsigned long long zero[qm::PHASE_COUNT] = {};
        QM_TRY(hipMemcpyToSymbol(HIP_SYMBOL(qm::qm_phase_cycles), zero, sizeof(zero)));
    }
    return 0;
}
#endif

}  // extern "C"
