// The batched painter of many uneven tiles (include/qm_tiles.h): one persistent, work-stealing
// launch over (tile, 16x16 receiver block, source block) items, captured with its clearing
// kernels in a hipGraph.
//
// tiles_kernel: a 256-work-item group takes an item from its XCD's stripe of the queue (stealing
// from the others when it is empty), maps its four waves to the four 8x8 quarters of the
// receiver block (neighbouring receivers in a wave follow near-identical rays), and runs the
// exact path of each receiver against every source of the block, the source uniform across the
// group (scalar loads). Per receiver the item's energy (FP64, pairwise band sums as the grid
// painter takes them) is added atomically: a receiver block's source blocks may run on different
// groups at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/qm_tiles.h"
#include "qm_eval.hpp"
#include "qm_internal.hpp"
#include "qm_queue.hpp"

using namespace qm;

namespace {

constexpr int GROUP = QM_PATH_GROUP;
constexpr int BLOCK = 16;  // receiver block side

struct TileItem {
    uint32_t tile, block, s0, s1;
};

struct TilesArgs {
    const qm_tile* tiles;
    const TileItem* items;
    uint32_t n_items;
    const qm_receiver* receivers;
    const qm_source* sources;
    double* energy;
    uint32_t* paths;
    uint32_t* failures;
    uint32_t* counters;  // XCDS queue heads
};

}  // namespace

template <bool FOREST>
__global__ __launch_bounds__(GROUP) __attribute__((amdgpu_waves_per_eu(3))) void tiles_kernel(
    const Scene* __restrict__ scene, const BandTable* __restrict__ bands, TilesArgs g) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    __shared__ uint32_t s_item;
    PathScratch w;
    const int lane = threadIdx.x;
    const int wave = lane >> 6, l = lane & 63;
    const int lx = (wave & 1) * 8 + (l & 7), ly = (wave >> 1) * 8 + (l >> 3);
    const int xcd = blockIdx.x % XCDS;
    for (;;) {
        if (lane == 0) s_item = next_tile(g.counters, g.n_items, xcd);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= g.n_items) break;
        const TileItem it = g.items[item];
        const qm_tile T = g.tiles[it.tile];
        const uint32_t blocks_x = (T.nx + BLOCK - 1) / BLOCK;
        const uint32_t ix = (it.block % blocks_x) * BLOCK + lx, iy = (it.block / blocks_x) * BLOCK + ly;
        const bool valid = ix < T.nx && iy < T.ny;
        const size_t index = (size_t)T.receiver_first + (valid ? (size_t)iy * T.nx + ix : 0);
        const qm_receiver& rcv = g.receivers[index];
        double energy[PERIODS] = {0.0, 0.0, 0.0};
        uint32_t fails = 0;
        for (uint32_t s = it.s0; s < it.s1; ++s) {
            if (!valid) continue;
            PathResult r;
            evaluate_path<false, FOREST>(sc, B, rcv, g.sources[s], w, r);
            if (r.failed) {
                ++fails;
                continue;
            }
            for (int q = 0; q < PERIODS; ++q) {
                const float s01 = r.bands[q][0] + r.bands[q][1], s23 = r.bands[q][2] + r.bands[q][3];
                const float s45 = r.bands[q][4] + r.bands[q][5], s67 = r.bands[q][6] + r.bands[q][7];
                energy[q] += (double)((s01 + s23) + (s45 + s67));
            }
        }
        if (valid) {
            for (int q = 0; q < PERIODS; ++q) atomicAdd(&g.energy[index * PERIODS + q], energy[q]);
            atomicAdd(&g.paths[index], it.s1 - it.s0);
            if (fails) atomicAdd(&g.failures[index], fails);
        }
    }
}

// Fills n words with zero (a kernel, not a memset node, inside the captured graph).
__global__ void clear_kernel(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

struct qm_tile_plan {
    qm_gpu* gpu = nullptr;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    qm_tile* d_tiles = nullptr;
    TileItem* d_items = nullptr;
    uint32_t* d_counters = nullptr;
    uint64_t paths = 0, items = 0;
};

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

#define T_TRY(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(err, errlen, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

void release(qm_tile_plan* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_tiles) (void)hipFree(p->d_tiles);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_counters) (void)hipFree(p->d_counters);
    if (p->stream) (void)hipStreamDestroy(p->stream);
}

int build(qm_tile_plan* p, const qm_square* square, const qm_tile* tiles, size_t n_tiles,
          const qm_receiver* d_receivers, size_t n_receivers, uint32_t block_sources, uint32_t groups_per_cu,
          double* d_energy, uint32_t* d_paths, uint32_t* d_failures, char* err, size_t errlen) {
    // Items, heaviest tiles first, then dealt round-robin into the XCDs' stripes.
    std::vector<uint32_t> order(n_tiles);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        return (uint64_t)tiles[a].source_count * tiles[a].nx * tiles[a].ny >
               (uint64_t)tiles[b].source_count * tiles[b].nx * tiles[b].ny;
    });
    std::vector<TileItem> list;
    for (uint32_t t : order) {
        const qm_tile& T = tiles[t];
        const uint32_t blocks = ((T.nx + BLOCK - 1) / BLOCK) * ((T.ny + BLOCK - 1) / BLOCK);
        for (uint32_t s0 = 0; s0 < T.source_count; s0 += block_sources)
            for (uint32_t b = 0; b < blocks; ++b)
                list.push_back(TileItem{t, b, T.source_first + s0,
                                        T.source_first + std::min(T.source_count, s0 + block_sources)});
        p->paths += (uint64_t)T.source_count * T.nx * T.ny;
    }
    if (list.size() > 0xffffffffu) return fail(err, errlen, "too many work items");
    // Stripes: a tile's source block (all its receiver blocks against the same sources) stays
    // together on one XCD (its L2 sees those sources' rays); source blocks are dealt round-robin
    // so that every stripe carries the same mix of heavy and light tiles.
    std::vector<std::vector<TileItem>> dealt(XCDS);
    {
        size_t k = 0;
        int q = 0;
        while (k < list.size()) {
            const TileItem first = list[k];
            while (k < list.size() && list[k].tile == first.tile && list[k].s0 == first.s0) dealt[q].push_back(list[k++]);
            q = (q + 1) % XCDS;
        }
    }
    // next_tile's stripe q is the index range n*q/8 .. n*(q+1)/8, which the concatenation lines up
    // with dealt[q] up to one source block (every item is taken once whatever the boundaries).
    std::vector<TileItem> striped;
    striped.reserve(list.size());
    for (int q = 0; q < XCDS; ++q) striped.insert(striped.end(), dealt[q].begin(), dealt[q].end());
    p->items = striped.size();
    T_TRY(hipMalloc((void**)&p->d_tiles, sizeof(qm_tile) * std::max<size_t>(n_tiles, 1)));
    T_TRY(hipMalloc((void**)&p->d_items, sizeof(TileItem) * std::max<size_t>(striped.size(), 1)));
    T_TRY(hipMalloc((void**)&p->d_counters, sizeof(uint32_t) * XCDS));
    T_TRY(hipMemcpy(p->d_tiles, tiles, sizeof(qm_tile) * n_tiles, hipMemcpyHostToDevice));
    if (!striped.empty())
        T_TRY(hipMemcpy(p->d_items, striped.data(), sizeof(TileItem) * striped.size(), hipMemcpyHostToDevice));
    int cus = 0;
    T_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->gpu->ordinal));
    unsigned groups = (unsigned)std::max(1, cus) * std::max(groups_per_cu, 1u);
    groups = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(groups, std::max<uint64_t>(p->items, 1)));
    TilesArgs g;
    g.tiles = p->d_tiles;
    g.items = p->d_items;
    g.n_items = (uint32_t)p->items;
    g.receivers = d_receivers;
    g.sources = square->d_sources_by_input;
    g.energy = d_energy;
    g.paths = d_paths;
    g.failures = d_failures;
    g.counters = p->d_counters;
    // The graph: clear the outputs and the queue heads, then paint.
    T_TRY(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    T_TRY(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(clear_kernel, dim3(1024), dim3(256), 0, p->stream, reinterpret_cast<uint32_t*>(d_energy),
                       n_receivers * PERIODS * 2);
    hipLaunchKernelGGL(clear_kernel, dim3(1024), dim3(256), 0, p->stream, d_paths, n_receivers);
    hipLaunchKernelGGL(clear_kernel, dim3(1024), dim3(256), 0, p->stream, d_failures, n_receivers);
    hipLaunchKernelGGL(clear_kernel, dim3(1), dim3(64), 0, p->stream, p->d_counters, (size_t)XCDS);
    if (square->has_forest)
        hipLaunchKernelGGL(tiles_kernel<true>, dim3(groups), dim3(GROUP), 0, p->stream, square->d_scene, square->d_bands, g);
    else
        hipLaunchKernelGGL(tiles_kernel<false>, dim3(groups), dim3(GROUP), 0, p->stream, square->d_scene, square->d_bands, g);
    const hipError_t launched = hipGetLastError();
    const hipError_t ended = hipStreamEndCapture(p->stream, &p->graph);
    T_TRY(launched);
    T_TRY(ended);
    T_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" int qm_tiles_plan(qm_gpu* gpu, const qm_square* square, const qm_tile* tiles, size_t n_tiles,
                             const qm_receiver* d_receivers, size_t n_receivers, uint32_t block_sources,
                             uint32_t groups_per_cu, double* d_period_energy, uint32_t* d_paths, uint32_t* d_failures,
                             qm_tile_plan** out, char* err, size_t errlen) {
    if (!gpu || !square || !out || (n_tiles && !tiles) || !d_receivers || !d_period_energy || !d_paths ||
        !d_failures)
        return fail(err, errlen, "null argument");
    *out = nullptr;
    if (block_sources == 0) return fail(err, errlen, "block_sources must be at least 1");
    if (square->scene.terrain.fine_detail || square->scene.terrain.row_y_m)
        return fail(err, errlen, "fine terrain or a row table: paint the square with qm_paint_grid or qm_evaluate_pairs");
    if (groups_per_cu == 0 || groups_per_cu > 8) return fail(err, errlen, "groups_per_cu must be 1..8");
    for (size_t t = 0; t < n_tiles; ++t) {
        const qm_tile& T = tiles[t];
        if ((uint64_t)T.receiver_first + (uint64_t)T.nx * T.ny > n_receivers)
            return fail(err, errlen, "tile %zu: receivers beyond the %zu given", t, n_receivers);
        if ((uint64_t)T.source_first + T.source_count > square->n_sources)
            return fail(err, errlen, "tile %zu: sources beyond the square's %zu", t, square->n_sources);
    }
    T_TRY(hipSetDevice(gpu->ordinal));
    qm_tile_plan* p = new (std::nothrow) qm_tile_plan;
    if (!p) return fail(err, errlen, "out of host memory");
    p->gpu = gpu;
    if (build(p, square, tiles, n_tiles, d_receivers, n_receivers, block_sources, groups_per_cu, d_period_energy,
              d_paths, d_failures, err, errlen)) {
        release(p);
        delete p;
        return 1;
    }
    *out = p;
    return 0;
}

extern "C" int qm_tiles_run(qm_tile_plan* p, void* stream, char* err, size_t errlen) {
    if (!p) return fail(err, errlen, "null argument");
    T_TRY(hipSetDevice(p->gpu->ordinal));
    T_TRY(hipGraphLaunch(p->exec, (hipStream_t)stream));
    return 0;
}

extern "C" uint64_t qm_tiles_paths(const qm_tile_plan* p) { return p ? p->paths : 0; }
extern "C" uint64_t qm_tiles_items(const qm_tile_plan* p) { return p ? p->items : 0; }

extern "C" void qm_tiles_plan_free(qm_tile_plan* p) {
    if (!p) return;
    (void)hipSetDevice(p->gpu->ordinal);
    (void)hipDeviceSynchronize();
    release(p);
    delete p;
}
