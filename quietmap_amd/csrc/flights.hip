// Aircraft boxes at the painter's points: per segment (one point's boxes of one tile, file and
// band) its boxes' SEL energy per period and their passes in the 48 pass classes, added to the
// segment's sum slot. Restates the reference card's flight_sums (paint-gpu/kernels/aircraft.cuh:
// 260-311) and through it popup/src/aircraft/boxes.rs:265-300 tile_flight_passes and
// physics/src/doc29/boxes.rs:159-284 box_sel_at_receiver / average_piece_sel.
//
// A 256-item group per segment, a box a work-item. The average pieces' closest points, the reach,
// the distance weight and the horizon sector are taken in FP64, the SEL terms in FP32 (as the reference's card); a
// box whose pass lies within 0.02 dB of a 3 dB class edge is re-read wholly in FP64 in the
// oracle's operation order, so every pass lands in the class the FP64 path gives it. A group
// gathers its passes in LDS (FP64 atomics) and its energies by wave reduction, then adds each
// non-zero class and the energies to the slot with one global FP64 atomic each: segments of one
// point share its slot.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "../../include/qm_paint.h"
#include "qm_doc29.hpp"
#include "qm_internal.hpp"

struct qm_flight_boxes {
    qm_gpu* gpu;
    qm_flight_box* d_boxes;
    size_t n;
};

namespace {

using namespace qm_doc29;

constexpr int GROUP = 256;
constexpr int CLASS_SLOTS = 3 * QM_PASS_CLASSES;
constexpr double EDGE_DB = 0.02;
const double NPD_FT[ND] = {200.0, 400.0, 630.0, 1000.0, 2000.0, 4000.0, 6300.0, 10000.0, 16000.0, 25000.0};
constexpr double SHARE_NEAR_M = 1000.0 * METRES_PER_FOOT;  // box_sums.rs INSTALLATION_SHARE_SLANTS_M
constexpr double SHARE_FAR_M = TAIL_ANCHOR_M;

struct FlightArgs {
    float log_npd[ND];
    double log_npd_d[ND];
    double seconds[3];
    float share_scale;     // 1 / lg(far / near)
    float tail_spherical;  // 20 lg(tail anchor / 25,000 ft)
};

// physics/src/doc29/segment.rs:200-268 closest_points of the average piece from `start` along
// `delta`, in FP64.
struct Closest {
    double along, length;
    double on_line[3], on_segment[3];
    double lateral, height, lambda_height;
};

__device__ Closest closest_of(const double s[3], const double d[3]) {
#pragma clang fp contract(off)
    Closest c;
    const double horizontal = d[0] * d[0] + d[1] * d[1];
    const double length_sq = horizontal + d[2] * d[2];
    c.length = fmax(sqrt(length_sq), 1.0);
    if (horizontal <= 1e-6) {
        c.along = 0.0;
        for (int a = 0; a < 3; ++a) c.on_line[a] = c.on_segment[a] = s[a];
        c.lateral = hypot(s[0], s[1]);
        c.height = c.lambda_height = s[2];
        return c;
    }
    c.along = -(s[0] * d[0] + s[1] * d[1] + s[2] * d[2]) / length_sq;
    const double clamped = fmin(fmax(c.along, 0.0), 1.0);
    for (int a = 0; a < 3; ++a) {
        c.on_line[a] = s[a] + c.along * d[a];
        c.on_segment[a] = s[a] + clamped * d[a];
    }
    c.lateral = fabs(s[0] * d[1] - s[1] * d[0]) / sqrt(horizontal);
    const double slant_sq = c.on_line[0] * c.on_line[0] + c.on_line[1] * c.on_line[1] + c.on_line[2] * c.on_line[2];
    c.height = copysign(sqrt(fmax(slant_sq - c.lateral * c.lateral, 0.0)), c.on_line[2]);
    const double cos_climb = sqrt(horizontal / length_sq);
    c.lambda_height = (c.along >= 0.0 && c.along <= 1.0) ? c.height : c.on_segment[2] / cos_climb;
    return c;
}

// The average piece centred on the box's centroid `c` (relative to the point) along its axis at
// `gradient`, `half` its half length (boxes.rs:114-125 average_piece_ends).
__device__ Closest piece_closest(const qm_flight_box& b, const double c[3], double gradient) {
#pragma clang fp contract(off)
    const double half = 0.5 * fmax(b.average_length_m, 1.0);
    const double step[3] = {b.cos_axis * half, b.sin_axis * half, gradient * half};
    const double s[3] = {c[0] - step[0], c[1] - step[1], c[2] - step[2]};
    const double d[3] = {(c[0] + step[0]) - s[0], (c[1] + step[1]) - s[1], (c[2] + step[2]) - s[2]};
    return closest_of(s, d);
}

// horizons.rs:84-111 SectorEdges::loss_db with the range and the sector decided in FP64 (the
// oracle's sector_of) and the edge losses in FP32.
__device__ float horizon_loss_mixed(const float* at, int sectors, const double p[3]) {
    const double range_d = hypot(p[0], p[1]);
    if (range_d <= 1.0) return 0.0f;
    const float range = (float)range_d, z = (float)p[2];
    if (p[2] >= (double)at[0] * range_d) return 0.0f;
    double angle = fmod(atan2(p[1], p[0]), 2.0 * PI_D);
    if (angle < 0.0) angle += 2.0 * PI_D;
    const int sector = min((int)(angle / (2.0 * PI_D) * (double)sectors), sectors - 1);
    const float* edges = at + 1 + sector * (1 + 2 * QM_HORIZON_BANDS);
    if (p[2] >= (double)edges[0] * range_d) return 0.0f;
    float loss = 0.0f;
    for (int band = 0; band < QM_HORIZON_BANDS; ++band) {
        const float r = edges[1 + 2 * band];
        if (isnan(r)) continue;
        loss = fmaxf(loss, edge_loss_f(r, edges[2 + 2 * band] * r, range, z));
    }
    return loss;
}

// One average piece's SEL energy per period (boxes.rs:205-284), FP32 terms.
__device__ void piece_sel_f(const FlightArgs& g, const qm_flight_box& b, const double c[3], double gradient,
                            const float* horizon, float sel[3]) {
    const Closest cp = piece_closest(b, c, gradient);
    const float slant = (float)sqrt(cp.on_line[0] * cp.on_line[0] + cp.on_line[1] * cp.on_line[1] +
                                    cp.on_line[2] * cp.on_line[2]);
    // NpdPosition::at (npd.rs:121-160).
    const float s = fmaxf(slant, (float)NPD_NEAREST_SLANT_M);
    const float log_d = log10_f(s * (float)(1.0 / METRES_PER_FOOT));
    // The last NPD distance below the position (0 when none), as a count of the rising distances
    // passed.
    int interval = 0;
#pragma unroll
    for (int k = 1; k <= ND - 2; ++k) interval += log_d > g.log_npd[k] ? 1 : 0;
    const float fraction = div_f(log_d - g.log_npd[interval], g.log_npd[interval + 1] - g.log_npd[interval]);
    const float lg_scaled = b.lg_scaled_distance[interval] +
                            fraction * (b.lg_scaled_distance[interval + 1] - b.lg_scaled_distance[interval]);
    const float scaled = exp10f(lg_scaled);
    const float length = (float)cp.length;
    const float along = (float)cp.along * length;
    const float stretch = sqrtf((float)gradient * (float)gradient + 1.0f);
    const float lateral = lateral_f((float)cp.lambda_height, (float)cp.lateral);
    const float height = (float)cp.height;
    const float corr[2] = {installation_f(QM_INSTALL_WING, height, slant), installation_f(QM_INSTALL_FUSELAGE, height, slant)};
    const float mix = fminf(fmaxf(log10f(slant / (float)SHARE_NEAR_M) * g.share_scale, 0.0f), 1.0f);
    float screening = 0.0f;
    if (b.screened && horizon) {
        const float terrain = horizon_loss_mixed(horizon, QM_TERRAIN_SECTORS, cp.on_line);
        const float buildings = horizon_loss_mixed(horizon + QM_TERRAIN_HORIZON_FLOATS, QM_BUILDING_SECTORS, cp.on_segment);
        screening = fmaxf(fmaxf(terrain, buildings) - lateral, 0.0f);
    }
    // screening.rs:47-54 screened_sel_db(0, ...) less the lateral attenuation, as a gain.
    const float gain = exp10f(-0.1f * (screening + lateral));
    const bool beyond = s >= (float)NPD_LAST_DISTANCE_M;
    for (int p = 0; p < 3; ++p) {
        if (b.silent[p]) {
            sel[p] = 0.0f;
            continue;
        }
        const float* lv = b.levels_db[p];
        float level;
        if (!beyond) {
            level = lv[interval] + fraction * (lv[interval + 1] - lv[interval]);
        } else {
            // boxes.rs:134-147 box_level_db beyond 25,000 ft: spherical spreading and the
            // absorption through the tail anchor.
            const float last = lv[ND - 1];
            const float alpha = fmaxf((last - g.tail_spherical - b.tail_levels_db[p]) /
                                          (float)(TAIL_ANCHOR_M - NPD_LAST_DISTANCE_M),
                                      0.0f);
            level = last - 20.0f * log10f(s / (float)NPD_LAST_DISTANCE_M) - alpha * (s - (float)NPD_LAST_DISTANCE_M);
        }
        const float piece = fmaxf(b.piece_length_m[p] * stretch, 1.0f);
        const float finite = finite_fraction_f(along + 0.5f * (piece - length), piece, scaled);
        const float* sh = &b.installation_shares[p][0][0];
        float installation = 0.0f;
        for (int k = 0; k < 3; ++k) {
            const float share = sh[k] + mix * (sh[3 + k] - sh[k]);
            installation += share * (k < 2 ? corr[k] : 1.0f);
        }
        if (!(installation > 0.0f)) installation = 1.0f;
        sel[p] = exp10f(0.1f * level) * finite * installation * gain;
    }
}

// The same in FP64, in the oracle's operation order (boxes.rs average_piece_sel, no contraction).
__device__ __noinline__ void piece_sel_d(const FlightArgs& g, const qm_flight_box& b, const double c[3],
                                         double gradient, const float* horizon, double sel[3]) {
#pragma clang fp contract(off)
    const Closest cp = piece_closest(b, c, gradient);
    const double slant = hypot(hypot(cp.on_line[0], cp.on_line[1]), cp.on_line[2]);
    const double s = fmax(slant, NPD_NEAREST_SLANT_M);
    const double log_d = log10(s / METRES_PER_FOOT);
    int interval = 0;
    for (int k = ND - 2; k >= 1; --k)
        if (log_d > g.log_npd_d[k]) {
            interval = k;
            break;
        }
    const double fraction = (log_d - g.log_npd_d[interval]) / (g.log_npd_d[interval + 1] - g.log_npd_d[interval]);
    const double lg0 = b.lg_scaled_distance[interval], lg1 = b.lg_scaled_distance[interval + 1];
    const double scaled = exp((lg0 + fraction * (lg1 - lg0)) * log(10.0));
    const double along = cp.along * cp.length;
    const double stretch = hypot(gradient, 1.0);
    const double lateral = lateral_d(cp.lambda_height, cp.lateral);
    const double corr[3] = {installation_d(QM_INSTALL_WING, cp.height, slant),
                            installation_d(QM_INSTALL_FUSELAGE, cp.height, slant), 1.0};
    const double mix = fmin(fmax(log10(slant / SHARE_NEAR_M) / log10(SHARE_FAR_M / SHARE_NEAR_M), 0.0), 1.0);
    double terrain = 0.0, buildings = 0.0;
    if (b.screened && horizon) {
        terrain = horizon_loss_d(horizon, QM_TERRAIN_SECTORS, cp.on_line);
        buildings = horizon_loss_d(horizon + QM_TERRAIN_HORIZON_FLOATS, QM_BUILDING_SECTORS, cp.on_segment);
    }
    const double gain = energy_d((0.0 - fmax(fmax(terrain, buildings) - lateral, 0.0)) - lateral);
    for (int p = 0; p < 3; ++p) {
        if (b.silent[p]) {
            sel[p] = 0.0;
            continue;
        }
        double total = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double s0 = b.installation_shares[p][0][k], s1 = b.installation_shares[p][1][k];
            total += (s0 + mix * (s1 - s0)) * corr[k];
        }
        const double installation = total > 0.0 ? total : 1.0;
        const float* lv = b.levels_db[p];
        double level;
        if (s < NPD_LAST_DISTANCE_M) {
            const double lower = lv[interval];
            level = lower + fraction * ((double)lv[interval + 1] - lower);
        } else {
            const double last = lv[ND - 1];
            const double alpha = fmax((last - 20.0 * log10(TAIL_ANCHOR_M / NPD_LAST_DISTANCE_M) - (double)b.tail_levels_db[p]) /
                                          (TAIL_ANCHOR_M - NPD_LAST_DISTANCE_M),
                                      0.0);
            level = last - 20.0 * log10(s / NPD_LAST_DISTANCE_M) - alpha * (s - NPD_LAST_DISTANCE_M);
        }
        const double piece = fmax((double)b.piece_length_m[p] * stretch, 1.0);
        const double finite = finite_fraction_d(along + 0.5 * (piece - cp.length), piece, scaled);
        sel[p] = energy_d(level) * finite * installation * gain;
    }
}

// paint/src/flight_boxes.rs:23-33 FlightWeight::at, in FP64 (its smoothstep's ends cancel in FP32).
__device__ double weight_d(uint32_t kind, double distance, double coarse) {
#pragma clang fp contract(off)
    if (kind == QM_FLIGHT_WEIGHT_ONE) return 1.0;
    const double f = fmin(fmax((distance - 1.5 * coarse) / (1.5 * coarse), 0.0), 1.0);
    const double far = f * f * (3.0 - 2.0 * f);
    return kind == QM_FLIGHT_WEIGHT_FAR ? far : 1.0 - far;
}

// popup/src/aircraft.rs:86-95 FlightPasses::add_pass's class, and whether the level lies within
// EDGE_DB of a class edge.
__device__ int pass_class_of(double energy, double passing, bool& near_edge) {
    const double level = 10.0 * log10(energy / passing);
    const double steps = (level - QM_PASS_LOWEST_DB) / QM_PASS_CLASS_DB;
    near_edge = fabs(steps - rint(steps)) * QM_PASS_CLASS_DB < EDGE_DB && steps > -0.5 &&
                steps < (double)QM_PASS_CLASSES - 0.5;
    const int c = (int)fmax(steps, 0.0);
    return min(c, QM_PASS_CLASSES - 1);
}

__device__ int pass_class_exact(double energy, double passing) {
#pragma clang fp contract(off)
    const double level = 10.0 * log10(energy / passing);
    const int c = (int)fmax((level - QM_PASS_LOWEST_DB) / QM_PASS_CLASS_DB, 0.0);
    return min(c, QM_PASS_CLASSES - 1);
}

// The arrays come as kernel arguments of their own (global pointers), the tables in `g`: pointers
// read out of an argument struct that is also passed by reference were generic, and every box
// field a flat load. At 4 waves per SIMD (128 VGPRs, some spilled around the FP64 re-read) the
// launch on 20,000 boxes x 4,096 points takes 21.2 ms against 26.2 ms at the 2 its 203 VGPRs
// would give (21.8 ms at 3): the box reads' latency, not the spills, bounds it.
__global__ __launch_bounds__(GROUP) __attribute__((amdgpu_waves_per_eu(4))) void flight_sums_kernel(const qm_flight_box* __restrict__ boxes,
                                                            const qm_flight_point* __restrict__ points,
                                                            const float* __restrict__ horizons,
                                                            const qm_flight_segment* __restrict__ segments,
                                                            double* __restrict__ sums, FlightArgs g) {
    __shared__ double classes[2 * CLASS_SLOTS];  // energies [period][class], then passing shares
    __shared__ double partial[GROUP / 64][3];
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * CLASS_SLOTS; i += GROUP) classes[i] = 0.0;
    __syncthreads();
    const qm_flight_segment sg = segments[blockIdx.x];
    const qm_flight_point& pt = points[sg.point];
    const double px = pt.x, py = pt.y, palt = pt.altitude_m;
    const float* horizon = pt.horizons == QM_NO_HORIZONS ? nullptr : horizons + pt.horizons;
    const double reach = fmin(sg.within_m, QM_AIRCRAFT_REACH_M);
    const double impedance_point = (double)pt.impedance_gain;
    double energy[3] = {0.0, 0.0, 0.0};
    for (uint32_t k = tid; k < sg.count; k += GROUP) {
        const qm_flight_box& b = boxes[sg.first + k];
        const double east = b.east_m - px, north = b.north_m - py;
        const double distance = hypot(east, north);
        if (distance > reach) continue;
        const double share = weight_d(sg.weight, distance, (double)sg.coarse_m);
        if (share <= 0.0) continue;
        const double c[3] = {east, north, b.centroid_altitude_m - palt};
        float sel[3];
        if (b.gradient_spread <= 0.0) {
            piece_sel_f(g, b, c, b.gradient, horizon, sel);
        } else {
            float upper[3];
            piece_sel_f(g, b, c, b.gradient - b.gradient_spread, horizon, sel);
            piece_sel_f(g, b, c, b.gradient + b.gradient_spread, horizon, upper);
            for (int p = 0; p < 3; ++p) sel[p] = 0.5f * (sel[p] + upper[p]);
        }
        const double impedance = b.helicopter ? 1.0 : impedance_point;
        double e[3], passing[3];
        int cls[3];
        bool near_edge = false;
        for (int p = 0; p < 3; ++p) {
            e[p] = share * (double)sel[p] * impedance / g.seconds[p];
            passing[p] = share * b.passing[p];
            cls[p] = -1;
            if (e[p] > 0.0 && passing[p] > 0.0) {
                bool edge;
                cls[p] = pass_class_of(e[p], passing[p], edge);
                near_edge = near_edge || edge;
            }
        }
        if (near_edge) {
            // The FP64 path: both pieces and the class decision as the oracle takes them.
#pragma clang fp contract(off)
            double sd[3];
            if (b.gradient_spread <= 0.0) {
                piece_sel_d(g, b, c, b.gradient, horizon, sd);
            } else {
                double lo[3], hi[3];
                piece_sel_d(g, b, c, b.gradient - b.gradient_spread, horizon, lo);
                piece_sel_d(g, b, c, b.gradient + b.gradient_spread, horizon, hi);
                for (int p = 0; p < 3; ++p) sd[p] = 0.5 * (lo[p] + hi[p]);
            }
            for (int p = 0; p < 3; ++p) {
                e[p] = share * sd[p] * impedance / g.seconds[p];
                passing[p] = share * b.passing[p];
                cls[p] = (e[p] > 0.0 && passing[p] > 0.0) ? pass_class_exact(e[p], passing[p]) : -1;
            }
        }
        for (int p = 0; p < 3; ++p) {
            energy[p] += e[p];
            if (cls[p] >= 0) {
                atomicAdd(&classes[p * QM_PASS_CLASSES + cls[p]], e[p]);
                atomicAdd(&classes[CLASS_SLOTS + p * QM_PASS_CLASSES + cls[p]], passing[p]);
            }
        }
    }
    for (int p = 0; p < 3; ++p) {
        double v = energy[p];
        for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
        if ((tid & 63) == 0) partial[tid >> 6][p] = v;
    }
    __syncthreads();
    double* at = sums + (size_t)sg.sum * QM_FLIGHT_SUMS;
    if (tid < 3) {
        double total = 0.0;
        for (int w = 0; w < GROUP / 64; ++w) total += partial[w][tid];
        if (total != 0.0) atomicAdd(&at[tid], total);
    }
    for (int i = tid; i < 2 * CLASS_SLOTS; i += GROUP)
        if (classes[i] != 0.0) atomicAdd(&at[3 + i], classes[i]);
}

int ffail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

FlightArgs args_of() {
    FlightArgs g;
    std::memset(&g, 0, sizeof g);
    for (int k = 0; k < ND; ++k) {
        g.log_npd_d[k] = std::log10(NPD_FT[k]);
        g.log_npd[k] = (float)g.log_npd_d[k];
    }
    for (int p = 0; p < 3; ++p) g.seconds[p] = QM_PERIOD_HOURS[p] * 3600.0;
    g.share_scale = (float)(1.0 / std::log10(SHARE_FAR_M / SHARE_NEAR_M));
    g.tail_spherical = (float)(20.0 * std::log10(TAIL_ANCHOR_M / NPD_LAST_DISTANCE_M));
    return g;
}

}  // namespace

extern "C" int qm_flight_boxes_upload(qm_gpu* gpu, const qm_flight_box* boxes, size_t n, qm_flight_boxes** out,
                                      char* err, size_t errlen) {
    if (!gpu || !out || (n && !boxes)) return ffail(err, errlen, "null argument");
    *out = nullptr;
    for (size_t i = 0; i < n; ++i) {
        const qm_flight_box& b = boxes[i];
        if (!std::isfinite(b.east_m) || !std::isfinite(b.north_m) || !std::isfinite(b.centroid_altitude_m) ||
            !std::isfinite(b.gradient) || !(b.gradient_spread >= 0.0) || !std::isfinite(b.average_length_m))
            return ffail(err, errlen, "box %zu: geometry not finite", i);
        for (int p = 0; p < 3; ++p)
            if (!(b.passing[p] >= 0.0) || !std::isfinite(b.passing[p]))
                return ffail(err, errlen, "box %zu: passing share of period %d", i, p);
    }
    hipError_t e = hipSetDevice(gpu->ordinal);
    qm_flight_box* d = nullptr;
    if (e == hipSuccess) e = hipMalloc((void**)&d, sizeof(qm_flight_box) * (n ? n : 1));
    if (e == hipSuccess && n) e = hipMemcpy(d, boxes, sizeof(qm_flight_box) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        return ffail(err, errlen, "flight boxes upload: %s", hipGetErrorString(e));
    }
    *out = new qm_flight_boxes{gpu, d, n};
    return 0;
}

extern "C" size_t qm_flight_boxes_count(const qm_flight_boxes* boxes) { return boxes ? boxes->n : 0; }

extern "C" void qm_flight_boxes_free(qm_flight_boxes* boxes) {
    if (!boxes) return;
    (void)hipSetDevice(boxes->gpu->ordinal);
    (void)hipFree(boxes->d_boxes);
    delete boxes;
}

extern "C" int qm_flight_sums_device(qm_gpu* gpu, const qm_flight_boxes* boxes, const qm_flight_point* d_points,
                                     const float* d_horizons, const qm_flight_segment* d_segments, size_t n_segments,
                                     double* d_sums, void* stream, char* err, size_t errlen) {
    if (!gpu || !boxes || (n_segments && (!d_points || !d_segments || !d_sums)))
        return ffail(err, errlen, "null argument");
    if (n_segments > 0x7fffffffu) return ffail(err, errlen, "too many segments");
    if (n_segments == 0) return 0;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) {
        const FlightArgs g = args_of();
        hipLaunchKernelGGL(flight_sums_kernel, dim3((unsigned)n_segments), dim3(GROUP), 0, (hipStream_t)stream,
                           (const qm_flight_box*)boxes->d_boxes, d_points, d_horizons, d_segments, d_sums, g);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return ffail(err, errlen, "flight sums: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_flight_sums(qm_gpu* gpu, const qm_flight_boxes* boxes, const qm_flight_point* points,
                              size_t n_points, const float* horizons, size_t n_horizon_floats,
                              const qm_flight_segment* segments, size_t n_segments, double* sums, size_t n_sums,
                              char* err, size_t errlen) {
    if (!gpu || !boxes || (n_segments && (!points || !segments)) || (n_sums && !sums))
        return ffail(err, errlen, "null argument");
    for (size_t i = 0; i < n_points; ++i) {
        const qm_flight_point& p = points[i];
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.altitude_m))
            return ffail(err, errlen, "point %zu: not finite", i);
        if (p.horizons != QM_NO_HORIZONS &&
            (!horizons || p.horizons > n_horizon_floats || n_horizon_floats - p.horizons < QM_HORIZON_FLOATS))
            return ffail(err, errlen, "point %zu: horizons at %llu of %zu floats", i, (unsigned long long)p.horizons,
                         n_horizon_floats);
    }
    for (size_t i = 0; i < n_segments; ++i) {
        const qm_flight_segment& s = segments[i];
        if (s.point >= n_points) return ffail(err, errlen, "segment %zu: point %u of %zu", i, s.point, n_points);
        if (s.sum >= n_sums) return ffail(err, errlen, "segment %zu: sum %u of %zu", i, s.sum, n_sums);
        if ((size_t)s.first + s.count > boxes->n)
            return ffail(err, errlen, "segment %zu: boxes %u+%u of %zu", i, s.first, s.count, boxes->n);
        if (s.weight > QM_FLIGHT_WEIGHT_NEAR) return ffail(err, errlen, "segment %zu: weight %u", i, s.weight);
        if (s.weight != QM_FLIGHT_WEIGHT_ONE && !(s.coarse_m > 0.0f))
            return ffail(err, errlen, "segment %zu: coarse spacing %g", i, (double)s.coarse_m);
    }
    hipError_t e = hipSetDevice(gpu->ordinal);
    void *d_points = nullptr, *d_horizons = nullptr, *d_segments = nullptr, *d_sums = nullptr;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 1);
    };
    alloc(&d_points, sizeof(qm_flight_point) * n_points);
    alloc(&d_horizons, sizeof(float) * n_horizon_floats);
    alloc(&d_segments, sizeof(qm_flight_segment) * n_segments);
    alloc(&d_sums, sizeof(double) * QM_FLIGHT_SUMS * n_sums);
    if (e == hipSuccess && n_points)
        e = hipMemcpy(d_points, points, sizeof(qm_flight_point) * n_points, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_horizon_floats)
        e = hipMemcpy(d_horizons, horizons, sizeof(float) * n_horizon_floats, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_segments)
        e = hipMemcpy(d_segments, segments, sizeof(qm_flight_segment) * n_segments, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_sums, 0, sizeof(double) * QM_FLIGHT_SUMS * n_sums);
    char inner[256] = {0};
    int rc = 0;
    if (e == hipSuccess)
        rc = qm_flight_sums_device(gpu, boxes, (const qm_flight_point*)d_points, (const float*)d_horizons,
                                   (const qm_flight_segment*)d_segments, n_segments, (double*)d_sums, nullptr, inner,
                                   sizeof inner);
    if (e == hipSuccess && rc == 0 && n_sums)
        e = hipMemcpy(sums, d_sums, sizeof(double) * QM_FLIGHT_SUMS * n_sums, hipMemcpyDeviceToHost);
    for (void* p : {d_points, d_horizons, d_segments, d_sums})
        if (p) (void)hipFree(p);
    if (rc) return ffail(err, errlen, "%s", inner);
    if (e != hipSuccess) return ffail(err, errlen, "flight sums: %s", hipGetErrorString(e));
    return 0;
}
