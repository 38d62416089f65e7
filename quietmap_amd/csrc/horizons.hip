// A receiver's terrain and building horizons for aircraft screening, on the card: per direction
// sector and range band the steepest terrain and roof edge seen from the receiver, written as the
// floats the screening reads (popup/src/aircraft/horizons.rs:70-210 Horizons::build and write_f32;
// the reference builds them on the host for each of its card's points, paint-gpu/src/aircraft.rs:
// 252-258).
//
// A 64-lane workgroup per receiver. Terrain: lane s < 32 walks sector s's 48 samples growing
// geometrically from 30 m to 8 km (the ground with its fine detail, Ground::read_height_m);
// buildings: each lane takes four of the 256 sectors and searches the crossings of the ray to
// 512 m (in four 128 m pieces, so that a dense block's crossings stay within a path's capacity;
// the exact FP64 search of the painter), keeping building walls other than the receiver's own, a
// roof the ground under the wall plus its height. The ground read, ranges and tangents in FP64, written as FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/qm_paint.h"
#include "qm_eval.hpp"
#include "qm_internal.hpp"

namespace {

using namespace qm;

constexpr int TERRAIN_SAMPLES = 48;
constexpr double FIRST_SAMPLE_M = 30.0;
constexpr double BUILDING_REACH_M = 512.0;
constexpr int PIECES = 4;
constexpr double NEAREST_EDGE_M = 0.01;
constexpr int SECTOR_FLOATS = 1 + 2 * QM_HORIZON_BANDS;

__constant__ double TERRAIN_LIMITS[QM_HORIZON_BANDS] = {500.0, 1000.0, 2000.0, 3500.0, 5500.0, 8000.0};
__constant__ double BUILDING_LIMITS[QM_HORIZON_BANDS] = {16.0, 32.0, 64.0, 128.0, 256.0, 512.0};

__device__ int band_of(double range, const double* limits) {
    for (int k = 0; k < QM_HORIZON_BANDS; ++k)
        if (range <= limits[k] + 0.5) return k;
    return QM_HORIZON_BANDS - 1;
}

// One sector's edges (horizons.rs SectorEdges::offer): per band the steepest, and the steepest,
// compared in FP64 as offered.
struct Sector {
    double range[QM_HORIZON_BANDS], tangent[QM_HORIZON_BANDS];
    double steepest;
    __device__ void clear() {
        for (int b = 0; b < QM_HORIZON_BANDS; ++b) range[b] = -1.0;
        steepest = -INFINITY;
    }
    __device__ void offer(int band, double r, double t) {
        if (range[band] < 0.0 || t > tangent[band]) {
            range[band] = r;
            tangent[band] = t;
        }
        steepest = fmax(steepest, t);
    }
    __device__ void write(float* at) const {
        at[0] = (float)steepest;
        for (int b = 0; b < QM_HORIZON_BANDS; ++b) {
            at[1 + 2 * b] = range[b] < 0.0 ? NAN : (float)range[b];
            at[2 + 2 * b] = range[b] < 0.0 ? NAN : (float)tangent[b];
        }
    }
};

// Ground::read_height_m in FP64, as the host reads it (tiles/src/terrain.rs:342-385 sample_at, the
// fine detail where the point's block has it): the height at lattice position (cp, rp), or a
// failure where the cell has no terrain.
__device__ uint32_t height_at(const qm_terrain& T, double cp, double rp, double& height) {
    const double column = floor(cp), row = floor(rp);
    const double cf = cp - column, rf = rp - row;
    const int c = (int)column, r = (int)row;
    const int nc = c + (cf > 0.0 ? 1 : 0), nr = r + (rf > 0.0 ? 1 : 0);
    uint32_t n[4], k[4];
    if (column < -1.0 || row < -1.0 || column > (double)T.columns || row > (double)T.rows ||
        !node_at(T, r, c, n[0], k[0]) || !node_at(T, r, nc, n[1], k[1]) || !node_at(T, nr, c, n[2], k[2]) ||
        !node_at(T, nr, nc, n[3], k[3]))
        return QM_FAILED_NO_TERRAIN;
    const int nearest = (rf >= 0.5 ? 2 : 0) + (cf >= 0.5 ? 1 : 0);
    if ((n[nearest] >> 24) > 100u) return QM_FAILED_NO_TERRAIN;
    for (int i = 0; i < 4; ++i)
        if (!node_valid(n[i])) return QM_FAILED_NO_TERRAIN;
    const double w[4] = {(1.0 - rf) * (1.0 - cf), (1.0 - rf) * cf, rf * (1.0 - cf), rf * cf};
    double h = 0.0;
    for (int i = 0; i < 4; ++i) h += w[i] * (-500.0 + (double)(n[i] & 0xffffu) / 5.0);
    if (fine_block(T, cp, rp)) {
        const uint32_t rows = fine_rows(T), columns = fine_columns(T);
        const double fr_row = fmax(QM_FINE_PER_NODE * rp, 0.0), fr_column = fmax(QM_FINE_PER_NODE * cp, 0.0);
        const uint32_t fr0 = min((uint32_t)floor(fr_row), rows - 2), fc0 = min((uint32_t)floor(fr_column), columns - 2);
        const double fr = fr_row - fr0, fc = fr_column - fc0;
        const int16_t* d = T.fine_detail;
        const size_t at = (size_t)fr0 * columns + fc0;
        const double sum = (1.0 - fr) * ((1.0 - fc) * d[at] + fc * d[at + 1]) +
                           fr * ((1.0 - fc) * d[at + columns] + fc * d[at + columns + 1]);
        h += sum * QM_DETAIL_STEP_M;
    }
    height = h;
    return 0;
}

__device__ double wave_max(double x) {
    for (int offset = 32; offset > 0; offset >>= 1) x = fmax(x, __shfl_xor(x, offset, 64));
    return x;
}

__device__ uint32_t wave_or(uint32_t x) {
    for (int offset = 32; offset > 0; offset >>= 1) x |= __shfl_xor(x, offset, 64);
    return x;
}

__global__ __launch_bounds__(64) void horizons_kernel(Scene sc, const qm_receiver* receivers, const double* altitudes,
                                                      uint32_t n, double growth, float* out, uint32_t* failed_out) {
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const int lane = threadIdx.x;
    const qm_receiver& rcv = receivers[i];
    const double x = rcv.x_m, y = rcv.y_m;
    const double altitude = altitudes ? altitudes[i] : (double)rcv.altitude_m;
    const qm_terrain& T = sc.terrain;
    float* record = out + (size_t)i * QM_HORIZON_FLOATS;
    uint32_t failed = rcv.failed;
    // Terrain: a sector a lane.
    double steepest = -INFINITY;
    if (lane < QM_TERRAIN_SECTORS) {
        const double a = ((double)lane + 0.5) * (2.0 * M_PI) / QM_TERRAIN_SECTORS;
        const double east = cos(a), north = sin(a);
        Sector s;
        s.clear();
        double range = FIRST_SAMPLE_M;
        for (int k = 0; k < TERRAIN_SAMPLES; ++k) {
            double h;
            const double px = x + east * range, py = y + north * range;
            if (height_at(T, T.column_m > 0.0 ? column_position(T, px) : (px - T.origin_x_m) / T.spacing_m,
                          T.row_y_m ? row_position<true>(T, py) : (py - T.origin_y_m) / T.spacing_m, h) == 0)
                s.offer(band_of(range, TERRAIN_LIMITS), range, (h - altitude) / range);
            range *= growth;
        }
        s.write(record + 1 + lane * SECTOR_FLOATS);
        steepest = s.steepest;
    }
    steepest = wave_max(steepest);
    if (lane == 0) record[0] = (float)steepest;
    // Buildings: four sectors a lane.
    float* buildings = record + QM_TERRAIN_HORIZON_FLOATS;
    double steepest_b = -INFINITY;
    Crossing cross[CROSS_CAP];
    for (int sector = lane; sector < QM_BUILDING_SECTORS; sector += 64) {
        const double a = ((double)sector + 0.5) * (2.0 * M_PI) / QM_BUILDING_SECTORS;
        const double east = cos(a), north = sin(a);
        Sector s;
        s.clear();
        const double piece = BUILDING_REACH_M / PIECES;
        for (int p = 0; p < PIECES; ++p) {
            const double from = p * piece;
            int count = 0;
            failed |= find_crossings(sc, x + east * from, y + north * from, x + east * (from + piece),
                                     y + north * (from + piece), cross, count);
            for (int c = 0; c < count; ++c) {
                const Crossing& k = cross[c];
                if (!(k.flags & 1u) || k.footprint == rcv.own_footprint) continue;
                const double range = from + k.t * piece;
                if (range <= NEAREST_EDGE_M) continue;
                double h;
                const double px = x + east * range, py = y + north * range;
                if (height_at(T, T.column_m > 0.0 ? column_position(T, px) : (px - T.origin_x_m) / T.spacing_m,
                          T.row_y_m ? row_position<true>(T, py) : (py - T.origin_y_m) / T.spacing_m, h)) {
                    failed |= QM_FAILED_NO_TERRAIN;
                    continue;
                }
                const double roof = h + (double)sc.obstacles.outlines[k.outline].height_dm / 10.0;
                s.offer(band_of(range, BUILDING_LIMITS), range, (roof - altitude) / range);
            }
        }
        s.write(buildings + 1 + sector * SECTOR_FLOATS);
        steepest_b = fmax(steepest_b, s.steepest);
    }
    steepest_b = wave_max(steepest_b);
    if (lane == 0) buildings[0] = (float)steepest_b;
    failed = wave_or(failed);
    if (lane == 0) failed_out[i] = failed;
}

int hfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

}  // namespace

extern "C" int qm_horizons(qm_gpu* gpu, const qm_square* square, const qm_point* points, const double* altitudes_m,
                           size_t n_points, float* out, uint32_t* failed, char* err, size_t errlen) {
    if (!gpu || !square || (n_points && (!points || !out || !failed))) return hfail(err, errlen, "null argument");
    if (n_points == 0) return 0;
    if (n_points > 0xffffffffu) return hfail(err, errlen, "too many points");
    hipError_t e = hipSetDevice(gpu->ordinal);
    void *d_points = nullptr, *d_receivers = nullptr, *d_alt = nullptr, *d_out = nullptr, *d_failed = nullptr;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    alloc(&d_points, sizeof(qm_point) * n_points);
    alloc(&d_receivers, sizeof(qm_receiver) * n_points);
    if (altitudes_m) alloc(&d_alt, sizeof(double) * n_points);
    alloc(&d_out, sizeof(float) * QM_HORIZON_FLOATS * n_points);
    alloc(&d_failed, sizeof(uint32_t) * n_points);
    if (e == hipSuccess) e = hipMemcpy(d_points, points, sizeof(qm_point) * n_points, hipMemcpyHostToDevice);
    if (e == hipSuccess && altitudes_m)
        e = hipMemcpy(d_alt, altitudes_m, sizeof(double) * n_points, hipMemcpyHostToDevice);
    char inner[256] = {0};
    if (e == hipSuccess &&
        qm_receivers_device(gpu, square, (const qm_point*)d_points, n_points, (qm_receiver*)d_receivers, nullptr, inner,
                            sizeof inner)) {
        for (void* p : {d_points, d_receivers, d_alt, d_out, d_failed})
            if (p) (void)hipFree(p);
        return hfail(err, errlen, "receivers: %s", inner);
    }
    if (e == hipSuccess) {
        const double growth = std::pow(8000.0 / FIRST_SAMPLE_M, 1.0 / (double)(TERRAIN_SAMPLES - 1));
        hipLaunchKernelGGL(horizons_kernel, dim3((unsigned)n_points), dim3(64), 0, 0, square->scene,
                           (const qm_receiver*)d_receivers, (const double*)d_alt, (uint32_t)n_points, growth,
                           (float*)d_out, (uint32_t*)d_failed);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d_out, sizeof(float) * QM_HORIZON_FLOATS * n_points, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(failed, d_failed, sizeof(uint32_t) * n_points, hipMemcpyDeviceToHost);
    for (void* p : {d_points, d_receivers, d_alt, d_out, d_failed})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return hfail(err, errlen, "horizons: %s", hipGetErrorString(e));
    return 0;
}
