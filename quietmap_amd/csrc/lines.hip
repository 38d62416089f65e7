// Line-source pairs (include/qm_paint.h qm_evaluate_line_pairs): the reference's pair_rays /
// RayLister (paint-gpu/kernels/paint.cu:77-173) and ray_transfers (:229-290) for line candidates.
//
//   line_nodes_kernel - per pair: the piece's quadrature nodes seen from the receiver (qm_line.hpp:
//                       FP64 geometry, the skyline mask of wide buckets), at most LINE_NODE_CAP
//   scan_blocks_kernel, scan_offsets_kernel, list_rays_kernel
//                     - the exclusive scan of the node counts on the device (each block's 1,024
//                       counts by wave scans, then the blocks' offsets), and each pair's rays
//                       listed from its first: one ray per node, pair-major
//   line_rays_kernel  - per ray: the exact path from the node's point on the piece to the
//                       receiver, weighted by its dphi and the piece's divergence factor, walls
//                       searched unless the node stands in a clear gap; the pair's 9 sums and the
//                       point's bands accumulated atomically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_eval.hpp"
#include "qm_internal.hpp"
#include "qm_line.hpp"

using namespace qm;

__global__ __launch_bounds__(256) void line_nodes_kernel(Scene sc, const qm_receiver* receivers,
                                                         const qm_line_piece* pieces, const uint32_t* pair_point,
                                                         const uint32_t* pair_piece, uint32_t n, LineNode* nodes,
                                                         uint32_t* counts, float* divergence, uint32_t* failed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const qm_receiver& rcv = receivers[pair_point[i]];
    const qm_line_piece& piece = pieces[pair_piece[i]];
    counts[i] = 0;
    divergence[i] = 0.0f;
    failed[i] = rcv.failed;
    if (rcv.failed) return;
    // candidate.traced_ground_m: the ground at the piece's ends.
    const qm_terrain& T = sc.terrain;
    GroundSample ga, gb;
    uint32_t f = ground_at(T, column_position(T, piece.ax_m), row_position<true>(T, piece.ay_m), ga);
    f |= ground_at(T, column_position(T, piece.bx_m), row_position<true>(T, piece.by_m), gb);
    if (f) {
        failed[i] = f;
        return;
    }
    // On a bridge, the deck at the ends.
    const double ground_a = piece.on_deck ? (double)piece.deck_a_m : (double)ga.z;
    const double ground_b = piece.on_deck ? (double)piece.deck_b_m : (double)gb.z;
    const double start[3] = {piece.ax_m - rcv.x_m, piece.ay_m - rcv.y_m,
                             ground_a + piece.height_m - (double)rcv.altitude_m};
    const double end[3] = {piece.bx_m - rcv.x_m, piece.by_m - rcv.y_m,
                           ground_b + piece.height_m - (double)rcv.altitude_m};
    LineGeometry g;
    if (!line_geometry(start, end, g)) return;
    int k = 0;
    LineNode* out = nodes + (size_t)i * LINE_NODE_CAP;
    if (!line_nodes(sc.obstacles, rcv.x_m, rcv.y_m, fmax((double)piece.height_m, 0.0), g, out, k)) {
        failed[i] = QM_FAILED_CAPACITY;
        return;
    }
    // The node positions as fractions of the piece's length (its horizontal run is what a
    // node's point interpolates).
    for (int j = 0; j < k; ++j) out[j].along_m /= g.length_m;
    counts[i] = (uint32_t)k;
    divergence[i] = (float)line_divergence_factor(g);
}

template <bool FINE>
__global__ __launch_bounds__(QM_PATH_GROUP) __attribute__((amdgpu_waves_per_eu(2))) void line_rays_kernel(
    const Scene* __restrict__ scene, const BandTable* __restrict__ bands, const qm_receiver* receivers,
    const qm_line_piece* pieces, const uint32_t* pair_point,
    const uint32_t* pair_piece, const uint32_t* ray_pair, const uint32_t* ray_first, uint32_t n_rays,
    const LineNode* nodes, const float* divergence, float* sums, uint32_t* failed, float* pair_bands) {
    const Scene& sc = *scene;
    const BandTable& B = *bands;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_rays) return;
    PathScratch w;
    const uint32_t i = ray_pair[k];
    const LineNode node = nodes[(size_t)i * LINE_NODE_CAP + (k - ray_first[i])];
    const qm_line_piece& piece = pieces[pair_piece[i]];
    qm_source src;
    src.x_m = piece.ax_m + node.along_m * (piece.bx_m - piece.ax_m);
    src.y_m = piece.ay_m + node.along_m * (piece.by_m - piece.ay_m);
    src.height_m = piece.height_m;
    src.platform_half_width_m = 0.0f;
    src.exclusion_radius_m = 0.0f;
    src.ground_percent = QM_GROUND_FROM_TERRAIN;
    src.layer = piece.layer;
    src.on_deck = piece.on_deck;
    src.deck_m = piece.deck_a_m + node.along_m * (piece.deck_b_m - piece.deck_a_m);
    src.footprint_id = 0;
    for (int q = 0; q < PERIODS; ++q)
        for (int b = 0; b < BANDS; ++b) src.energy[q][b] = piece.energy_per_m[q][b];
    RayOptions opt;
    opt.obstacles = node.obstacles_on_ray != 0;
    opt.weight = node.weight_rad * divergence[i];
    PathResult r;
    evaluate_path<FINE>(sc, B, receivers[pair_point[i]], src, w, r, opt);
    if (r.failed) {
        atomicOr(&failed[i], r.failed);
        return;
    }
    float* out = sums + (size_t)i * QM_PAIR_SUMS;
    for (int q = 0; q < PERIODS; ++q) {
        float mixed = 0.0f;
        for (int b = 0; b < BANDS; ++b) mixed += r.bands[q][b];
        atomicAdd(out + q, mixed);
        atomicAdd(out + 3 + q, r.states[0][q]);
        atomicAdd(out + 6 + q, r.states[1][q]);
    }
    if (pair_bands) {
        float* dst = pair_bands + (size_t)i * PERIODS * BANDS;
        for (int q = 0; q < PERIODS; ++q)
            for (int b = 0; b < BANDS; ++b) atomicAdd(dst + q * BANDS + b, r.bands[q][b]);
    }
}

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

#define L_TRY(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return fail(err, errlen, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// Device buffers freed when the call returns.
struct Buffers {
    std::vector<void*> p;
    ~Buffers() {
        for (void* x : p) (void)hipFree(x);
    }
    template <class T>
    hipError_t alloc(T** out, size_t count) {
        const hipError_t e = hipMalloc((void**)out, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(*out);
        return e;
    }
};

}  // namespace

__global__ __launch_bounds__(256) void line_receivers_kernel(Scene sc, const qm_point* points, uint32_t n,
                                                             qm_receiver* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = receiver_of(sc, points[i]);
}

constexpr int SCAN_GROUP = 256;
constexpr int SCAN_PER_THREAD = 4;
constexpr int SCAN_BLOCK = SCAN_GROUP * SCAN_PER_THREAD;

__device__ __forceinline__ uint32_t wave_inclusive(uint32_t x, int lane) {
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

// Per block of SCAN_BLOCK counts: each count's exclusive prefix within the block (into `first`)
// and the block's total.
__global__ __launch_bounds__(SCAN_GROUP) void scan_blocks_kernel(const uint32_t* counts, uint32_t n, uint32_t* first,
                                                                 uint32_t* block_total) {
    __shared__ uint32_t wave_total[SCAN_GROUP / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * SCAN_BLOCK + (size_t)tid * SCAN_PER_THREAD;
    uint32_t v[SCAN_PER_THREAD], mine = 0;
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        v[k] = base + k < n ? counts[base + k] : 0u;
        mine += v[k];
    }
    const uint32_t inclusive = wave_inclusive(mine, lane);
    if (lane == 63) wave_total[wave] = inclusive;
    __syncthreads();
    uint32_t before = inclusive - mine;
    for (int w = 0; w < wave; ++w) before += wave_total[w];
    for (int k = 0; k < SCAN_PER_THREAD; ++k) {
        if (base + k < n) first[base + k] = before;
        before += v[k];
    }
    if (tid == SCAN_GROUP - 1) block_total[blockIdx.x] = before;
}

// One workgroup: the blocks' exclusive offsets (64-bit, in chunks of SCAN_GROUP blocks with a
// carry) and the grand total.
__global__ __launch_bounds__(SCAN_GROUP) void scan_offsets_kernel(const uint32_t* block_total, uint32_t blocks,
                                                                  uint64_t* block_first, uint64_t* total) {
    __shared__ uint64_t wave_total[SCAN_GROUP / 64];
    __shared__ uint64_t carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t at = 0; at < blocks; at += SCAN_GROUP) {
        const uint64_t mine = at + tid < blocks ? block_total[at + tid] : 0ull;
        uint64_t inclusive = mine;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(inclusive, d);
            if (lane >= d) inclusive += y;
        }
        if (lane == 63) wave_total[wave] = inclusive;
        __syncthreads();
        uint64_t before = carry + inclusive - mine;
        for (int w = 0; w < wave; ++w) before += wave_total[w];
        if (at + tid < blocks) block_first[at + tid] = before;
        __syncthreads();
        if (tid == SCAN_GROUP - 1) carry = before + mine;
        __syncthreads();
    }
    if (tid == 0) *total = carry;
}

// Each pair's first ray (its block's offset added) and the pair of each of its rays.
__global__ __launch_bounds__(256) void list_rays_kernel(const uint32_t* counts, uint32_t n,
                                                        const uint64_t* block_first, uint32_t* first,
                                                        uint32_t* ray_pair) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = (uint32_t)(block_first[i / SCAN_BLOCK] + first[i]);
    first[i] = f;
    for (uint32_t k = 0; k < counts[i]; ++k) ray_pair[f + k] = i;
}

// The point bands of every pair that did not fail (a failed pair goes to the caller's fallback
// whole, as its sums do).
__global__ __launch_bounds__(256) void line_bands_kernel(const uint32_t* pair_point, const qm_line_piece* pieces,
                                                         const uint32_t* pair_piece, uint32_t n, const uint32_t* failed,
                                                         const float* pair_bands, float* point_bands) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || failed[i]) return;
    float* dst = point_bands + ((size_t)pair_point[i] * QM_LAYERS + pieces[pair_piece[i]].layer) * PERIODS * BANDS;
    for (int k = 0; k < PERIODS * BANDS; ++k) atomicAdd(dst + k, pair_bands[(size_t)i * PERIODS * BANDS + k]);
}

extern "C" int qm_evaluate_line_pairs(qm_gpu* gpu, const qm_square* square, const qm_point* points, size_t n_points,
                                      const qm_line_piece* pieces, size_t n_pieces, const uint32_t* pair_point,
                                      const uint32_t* pair_piece, size_t n_pairs, float* sums, uint32_t* failed,
                                      uint32_t* rays, float* point_bands, float* pair_bands, char* err,
                                      size_t errlen) {
    if (!gpu || !square || (n_pairs && (!points || !pieces || !pair_point || !pair_piece || !sums || !failed)))
        return fail(err, errlen, "null argument");
    if (n_pairs > 0xffffffffu || n_points > 0xffffffffu) return fail(err, errlen, "too many pairs or points");
    for (size_t i = 0; i < n_points; ++i)
        if (!std::isfinite(points[i].x_m) || !std::isfinite(points[i].y_m))
            return fail(err, errlen, "point %zu: position not finite", i);
    for (size_t i = 0; i < n_pieces; ++i)
        if (!std::isfinite(pieces[i].ax_m) || !std::isfinite(pieces[i].ay_m) || !std::isfinite(pieces[i].bx_m) ||
            !std::isfinite(pieces[i].by_m))
            return fail(err, errlen, "piece %zu: ends not finite", i);
    for (size_t i = 0; i < n_pieces; ++i)
        if (pieces[i].layer >= QM_LAYERS) return fail(err, errlen, "piece %zu: layer out of range", i);
    for (size_t i = 0; i < n_pairs; ++i) {
        if (pair_point[i] >= n_points) return fail(err, errlen, "pair %zu: point %u out of range", i, pair_point[i]);
        if (pair_piece[i] >= n_pieces) return fail(err, errlen, "pair %zu: piece %u out of range", i, pair_piece[i]);
    }
    if (point_bands) std::memset(point_bands, 0, sizeof(float) * n_points * QM_LAYERS * PERIODS * BANDS);
    std::memset(sums, 0, sizeof(float) * n_pairs * QM_PAIR_SUMS);
    if (n_pairs == 0) return 0;
    L_TRY(hipSetDevice(gpu->ordinal));
    Buffers b;
    qm_point* d_points;
    qm_receiver* d_receivers;
    qm_line_piece* d_pieces;
    uint32_t *d_pp, *d_pl, *d_counts, *d_failed, *d_ray_pair, *d_ray_first;
    LineNode* d_nodes;
    float *d_div, *d_sums, *d_bands = nullptr, *d_pair_bands = nullptr;
    L_TRY(b.alloc(&d_points, n_points));
    L_TRY(b.alloc(&d_receivers, n_points));
    L_TRY(b.alloc(&d_pieces, n_pieces));
    L_TRY(b.alloc(&d_pp, n_pairs));
    L_TRY(b.alloc(&d_pl, n_pairs));
    L_TRY(b.alloc(&d_counts, n_pairs));
    L_TRY(b.alloc(&d_failed, n_pairs));
    L_TRY(b.alloc(&d_ray_first, n_pairs));
    L_TRY(b.alloc(&d_nodes, n_pairs * LINE_NODE_CAP));
    L_TRY(b.alloc(&d_div, n_pairs));
    L_TRY(b.alloc(&d_sums, n_pairs * QM_PAIR_SUMS));
    if (point_bands) L_TRY(b.alloc(&d_bands, n_points * QM_LAYERS * PERIODS * BANDS));
    if (point_bands || pair_bands) L_TRY(b.alloc(&d_pair_bands, n_pairs * PERIODS * BANDS));
    L_TRY(hipMemcpy(d_points, points, sizeof(qm_point) * n_points, hipMemcpyHostToDevice));
    L_TRY(hipMemcpy(d_pieces, pieces, sizeof(qm_line_piece) * n_pieces, hipMemcpyHostToDevice));
    L_TRY(hipMemcpy(d_pp, pair_point, 4 * n_pairs, hipMemcpyHostToDevice));
    L_TRY(hipMemcpy(d_pl, pair_piece, 4 * n_pairs, hipMemcpyHostToDevice));
    L_TRY(hipMemset(d_sums, 0, sizeof(float) * n_pairs * QM_PAIR_SUMS));
    if (d_bands) L_TRY(hipMemset(d_bands, 0, sizeof(float) * n_points * QM_LAYERS * PERIODS * BANDS));
    if (d_pair_bands) L_TRY(hipMemset(d_pair_bands, 0, sizeof(float) * n_pairs * PERIODS * BANDS));
    hipLaunchKernelGGL(line_receivers_kernel, dim3((unsigned)((n_points + 255) / 256)), dim3(256), 0, 0, square->scene,
                       d_points, (uint32_t)n_points, d_receivers);
    hipLaunchKernelGGL(line_nodes_kernel, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, 0, square->scene,
                       d_receivers, d_pieces, d_pp, d_pl, (uint32_t)n_pairs, d_nodes, d_counts, d_div, d_failed);
    L_TRY(hipGetLastError());
    // One ray per node, pair-major (the reference's count / scan / list passes, on the device; only
    // the total comes back, to size the ray list).
    const uint32_t blocks = (uint32_t)((n_pairs + SCAN_BLOCK - 1) / SCAN_BLOCK);
    uint32_t* d_block_total;
    uint64_t *d_block_first, *d_total;
    L_TRY(b.alloc(&d_block_total, blocks));
    L_TRY(b.alloc(&d_block_first, blocks));
    L_TRY(b.alloc(&d_total, 1));
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(blocks), dim3(SCAN_GROUP), 0, 0, (const uint32_t*)d_counts,
                       (uint32_t)n_pairs, d_ray_first, d_block_total);
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(SCAN_GROUP), 0, 0, (const uint32_t*)d_block_total, blocks,
                       d_block_first, d_total);
    L_TRY(hipGetLastError());
    uint64_t total = 0;
    L_TRY(hipMemcpy(&total, d_total, sizeof(uint64_t), hipMemcpyDeviceToHost));
    if (total > 0xffffffffu) return fail(err, errlen, "too many rays");
    L_TRY(b.alloc(&d_ray_pair, total));
    hipLaunchKernelGGL(list_rays_kernel, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, 0,
                       (const uint32_t*)d_counts, (uint32_t)n_pairs, (const uint64_t*)d_block_first, d_ray_first,
                       d_ray_pair);
    L_TRY(hipGetLastError());
    if (total) {
        if (square->scene.terrain.fine_detail || square->scene.terrain.row_y_m)
            hipLaunchKernelGGL(line_rays_kernel<true>, dim3((unsigned)((total + QM_PATH_GROUP - 1) / QM_PATH_GROUP)), dim3(QM_PATH_GROUP), 0, 0,
                               square->d_scene, square->d_bands, d_receivers, d_pieces, d_pp, d_pl, d_ray_pair, d_ray_first,
                               (uint32_t)total, d_nodes, d_div, d_sums, d_failed, d_pair_bands);
        else
            hipLaunchKernelGGL(line_rays_kernel<false>, dim3((unsigned)((total + QM_PATH_GROUP - 1) / QM_PATH_GROUP)), dim3(QM_PATH_GROUP), 0, 0,
                               square->d_scene, square->d_bands, d_receivers, d_pieces, d_pp, d_pl, d_ray_pair, d_ray_first,
                               (uint32_t)total, d_nodes, d_div, d_sums, d_failed, d_pair_bands);
        L_TRY(hipGetLastError());
    }
    if (d_bands) {
        hipLaunchKernelGGL(line_bands_kernel, dim3((unsigned)((n_pairs + 255) / 256)), dim3(256), 0, 0, d_pp, d_pieces,
                           d_pl, (uint32_t)n_pairs, d_failed, d_pair_bands, d_bands);
        L_TRY(hipGetLastError());
    }
    L_TRY(hipDeviceSynchronize());
    L_TRY(hipMemcpy(failed, d_failed, 4 * n_pairs, hipMemcpyDeviceToHost));
    L_TRY(hipMemcpy(sums, d_sums, sizeof(float) * n_pairs * QM_PAIR_SUMS, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n_pairs; ++i)
        if (failed[i])
            for (int j = 0; j < QM_PAIR_SUMS; ++j) sums[i * QM_PAIR_SUMS + j] = 0.0f;
    if (rays) {
        L_TRY(hipMemcpy(rays, d_counts, 4 * n_pairs, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n_pairs; ++i)
            if (failed[i]) rays[i] = 0u;
    }
    if (point_bands) {
        L_TRY(hipMemcpy(point_bands, d_bands, sizeof(float) * n_points * QM_LAYERS * PERIODS * BANDS,
                        hipMemcpyDeviceToHost));
    }
    if (pair_bands) {
        L_TRY(hipMemcpy(pair_bands, d_pair_bands, sizeof(float) * n_pairs * PERIODS * BANDS, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n_pairs; ++i)
            if (failed[i])
                for (int j = 0; j < PERIODS * BANDS; ++j) pair_bands[i * PERIODS * BANDS + j] = 0.0f;
    }
    return 0;
}
