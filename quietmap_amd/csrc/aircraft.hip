// Doc 29 aircraft noise at a receiver grid from raw flight tracks: per receiver the SEL energy of
// every flight per period (Lden), the flights a day whose LAmax reaches 65 and 70 dB, an integer
// histogram of flights per 0.5 dB of LAmax and, optionally, the 48 pass classes per period.
//
// Reference behaviour restated: physics/src/doc29/segment.rs:200-381 (closest points, the
// segment's SEL and LAmax), npd.rs:121-330 (NPD positions, power rows, tails),
// corrections.rs:23-88 (Delta_V, Delta_F, lateral attenuation, Delta_I), screening.rs:56-63 (no
// horizons: unscreened); the reference card's flight_sums (paint-gpu/kernels/aircraft.cuh) runs
// the same physics per aircraft box.
//
// A work-item per receiver; a 256-item group is a 16x16 tile of receivers (four 8x8 waves) pulling
// tiles from per-XCD queues. Tracks and their segments are uniform across the group, read with
// scalar loads. SEL energies run in FP32 (per-track sums, FP64 across tracks) on the hardware
// reciprocal, log2 and exp2. An event's LAmax is the maximum over its segments: the FP32 pass
// keeps the three loudest segments; where that maximum lies within 0.02 dB of a 0.5 dB edge (the
// histogram's, 65 and 70 dB among them), the segments within 0.02 dB of the loudest are re-read in
// FP64 with the oracle's operation order (no contraction), so thresholds and histogram bins are
// decided on the same side of every edge as the FP64 value the CPU path computes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_doc29.hpp"
#include "qm_internal.hpp"
#include "qm_queue.hpp"

namespace {

using namespace qm_doc29;

constexpr float EPS_DB = 0.02f;
// N65 and N70 are decided at edges of the 0.5 dB histogram grid, so one edge test covers both.
static_assert((65.0 - QM_HIST_FLOOR_DB) / QM_HIST_STEP_DB == 50.0 && (70.0 - QM_HIST_FLOOR_DB) / QM_HIST_STEP_DB == 60.0,
              "65 and 70 dB lie on the histogram's edges");

// One power row as the device reads it (npd.rs:256-330), FP32 and FP64 copies.
struct DevRow {
    float sel[ND], lamax[ND], sml[ND];
    float tail_alpha, d0;
    uint32_t dipole;
    uint32_t pad;
    double lamax_d[ND];
    double tail_alpha_d;
    double sel_d[ND], sml_d[ND];
    double d0_d;
};

// A track with everything that is constant along it.
struct DevTrack {
    float x0, y0, x1, y1;       // bounding box of its vertices
    uint32_t first_vertex, vertex_count;
    uint32_t row_index;         // DevRow of its lower power row
    uint32_t installation;
    float weight, dv;           // power weight, Delta_V
    float sel_last, sel_alpha;  // the SEL tail past 25,000 ft (segment.rs:177-190)
    float lamax_last, lamax_alpha;
    double weight_d, lamax_last_d, lamax_alpha_d;
    double sel_last_d, sel_alpha_d, dv_d;
    double seconds_per_metre;  // of a pass (boxes.rs:245-257)
    uint32_t flights[3];
    uint32_t total_flights;
};

struct Tables {
    float log_npd[ND];
    double log_npd_d[ND];
};

struct AircraftArgs {
    const DevRow* rows;
    const DevTrack* tracks;
    uint32_t n_tracks;
    const float* vertices;
    const double* points;  // x, y
    uint32_t nx, ny, tiles_x, n_tiles;
    double receiver_altitude_m, reach_m;
    double* energy;
    uint32_t* n65;
    uint32_t* n70;
    uint32_t* hist;
    double* passes;  // QM_PASS_FLOATS a receiver, or null
    const float* horizons;          // QM_HORIZON_FLOATS a record, or null
    const uint32_t* horizon_of;     // a record per receiver (QM_NO_HORIZON: unscreened)
    const float* vertex_ground;     // terrain under each vertex, or null (0)
    double seconds[3];
    uint32_t* counters;
    Tables tables;
};

// ---------------------------------------------------------------------------------------------
// FP32 reads.
__device__ __forceinline__ void position_f(const Tables& T, float slant, int& interval, float& fraction, float& s) {
    s = fmaxf(slant, (float)NPD_NEAREST_SLANT_M);
    const float log_d = log10_f(s * (float)(1.0 / METRES_PER_FOOT));
    // The last NPD distance below the position (0 when none): the distances rise, so it is the
    // count of them passed, without a divergent search.
    interval = 0;
#pragma unroll
    for (int k = 1; k <= ND - 2; ++k) interval += log_d > T.log_npd[k] ? 1 : 0;
    fraction = div_f(log_d - T.log_npd[interval], T.log_npd[interval + 1] - T.log_npd[interval]);
}

__device__ __forceinline__ float linear_f(const float* c, int i, float f) { return c[i] + f * (c[i + 1] - c[i]); }

// The event LAmax of segment (a, b) in FP64 (segment.rs:357-381 over npd.rs and corrections.rs).
__device__ double segment_lmax_d(const AircraftArgs& g, const DevTrack& tr, const DevRow* rows, const float* a,
                                 const float* b, double rx, double ry) {
#pragma clang fp contract(off)
    const double s[3] = {(double)a[0] - rx, (double)a[1] - ry, (double)a[2] - g.receiver_altitude_m};
    const double e[3] = {(double)b[0] - rx, (double)b[1] - ry, (double)b[2] - g.receiver_altitude_m};
    const double d[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    const double hs = d[0] * d[0] + d[1] * d[1];
    const double ls = hs + d[2] * d[2];
    double along, on_line[3], on_seg[3], lateral, height;
    if (hs <= 1e-6) {
        along = 0.0;
        for (int k = 0; k < 3; ++k) on_line[k] = on_seg[k] = s[k];
        lateral = hypot(s[0], s[1]);
        height = s[2];
    } else {
        along = -(s[0] * d[0] + s[1] * d[1] + s[2] * d[2]) / ls;
        const double cl = fmin(fmax(along, 0.0), 1.0);
        for (int k = 0; k < 3; ++k) {
            on_line[k] = s[k] + along * d[k];
            on_seg[k] = s[k] + cl * d[k];
        }
        lateral = fabs(s[0] * d[1] - s[1] * d[0]) / sqrt(hs);
        const double ss = on_line[0] * on_line[0] + on_line[1] * on_line[1] + on_line[2] * on_line[2];
        height = copysign(sqrt(fmax(ss - lateral * lateral, 0.0)), on_line[2]);
    }
    const double distance = hypot(hypot(on_seg[0], on_seg[1]), on_seg[2]);
    // SegmentEmission::lamax_db (segment.rs:149-162)
    const double slant = fmax(distance, NPD_NEAREST_SLANT_M);
    double lamax;
    if (slant < NPD_LAST_DISTANCE_M) {
        const double log_d = log10(slant / METRES_PER_FOOT);
        int interval = 0;
        for (int k = ND - 2; k >= 1; --k)
            if (log_d > g.tables.log_npd_d[k]) {
                interval = k;
                break;
            }
        const double fraction = (log_d - g.tables.log_npd_d[interval]) /
                                (g.tables.log_npd_d[interval + 1] - g.tables.log_npd_d[interval]);
        const DevRow& lo = rows[tr.row_index];
        const double low = lo.lamax_d[interval] + fraction * (lo.lamax_d[interval + 1] - lo.lamax_d[interval]);
        if (tr.weight_d == 0.0) {
            lamax = low;
        } else {
            const DevRow& hi = rows[tr.row_index + 1];
            const double high = hi.lamax_d[interval] + fraction * (hi.lamax_d[interval + 1] - hi.lamax_d[interval]);
            lamax = low + tr.weight_d * (high - low);
        }
    } else {
        lamax = tr.lamax_last_d - 20.0 * log10(slant / NPD_LAST_DISTANCE_M) -
                tr.lamax_alpha_d * (slant - NPD_LAST_DISTANCE_M);
    }
    const double line_slant = hypot(hypot(on_line[0], on_line[1]), on_line[2]);
    const double lat = (along >= 0.0 && along <= 1.0) ? lateral_d(height, lateral)
                                                      : lateral_d(on_seg[2], hypot(on_seg[0], on_seg[1]));
    return lamax + 10.0 * log10(installation_d(tr.installation, height, line_slant)) - lat;
}

// One power row's SEL and scaled distance at an NPD position (npd.rs:290-330 PowerRow::read).
__device__ void row_read_d(const DevRow& r, int interval, double fraction, double slant, double& sel,
                           double& scaled) {
#pragma clang fp contract(off)
    if (r.dipole) {
        scaled = slant;
    } else {
        const double lower = r.sml_d[interval];
        scaled = r.d0_d * energy_d(lower + fraction * (r.sml_d[interval + 1] - lower));
    }
    if (slant < NPD_LAST_DISTANCE_M) {
        const double lower = r.sel_d[interval];
        sel = lower + fraction * (r.sel_d[interval + 1] - lower);
    } else {
        sel = r.sel_d[ND - 1] - 20.0 * log10(slant / NPD_LAST_DISTANCE_M) - r.tail_alpha_d * (slant - NPD_LAST_DISTANCE_M);
    }
}

// The SEL (dB) of segment (a, b) in FP64, in the oracle's operation order (segment.rs:318-350 over
// the closest points, SegmentEmission::read_npd and corrections.rs), unscreened.
// It takes the receiver altitude and the NPD distances' logarithms by value and the track by a
// pointer to the caller's copy: a reference into the kernel's argument block would make every
// pointer the kernel reads from that block a generic one.
__device__ __noinline__ double segment_sel_d(double receiver_altitude_m, const double* log_npd_d, const DevTrack& tr,
                                             const DevRow* rows, const float* a, const float* b, double rx, double ry,
                                             const float* horizon) {
#pragma clang fp contract(off)
    const double s[3] = {(double)a[0] - rx, (double)a[1] - ry, (double)a[2] - receiver_altitude_m};
    const double e[3] = {(double)b[0] - rx, (double)b[1] - ry, (double)b[2] - receiver_altitude_m};
    const double d[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    const double hs = d[0] * d[0] + d[1] * d[1];
    const double ls = hs + d[2] * d[2];
    double along, on_line[3], on_seg[3], lateral, height, lambda_height;
    if (hs <= 1e-6) {
        along = 0.0;
        for (int k = 0; k < 3; ++k) on_line[k] = on_seg[k] = s[k];
        lateral = hypot(s[0], s[1]);
        height = s[2];
        lambda_height = s[2];
    } else {
        along = -(s[0] * d[0] + s[1] * d[1] + s[2] * d[2]) / ls;
        const double cl = fmin(fmax(along, 0.0), 1.0);
        for (int k = 0; k < 3; ++k) {
            on_line[k] = s[k] + along * d[k];
            on_seg[k] = s[k] + cl * d[k];
        }
        lateral = fabs(s[0] * d[1] - s[1] * d[0]) / sqrt(hs);
        const double ss = on_line[0] * on_line[0] + on_line[1] * on_line[1] + on_line[2] * on_line[2];
        height = copysign(sqrt(fmax(ss - lateral * lateral, 0.0)), on_line[2]);
        const double cos_climb = sqrt(hs / ls);
        lambda_height = (along >= 0.0 && along <= 1.0) ? height : on_seg[2] / cos_climb;
    }
    const double length = fmax(sqrt(ls), 1.0);
    const double slant_in = hypot(hypot(on_line[0], on_line[1]), on_line[2]);
    // SegmentEmission::read_npd (segment.rs:126-146): the position, the power-interpolated table.
    const double slant = fmax(slant_in, NPD_NEAREST_SLANT_M);
    const double log_d = log10(slant / METRES_PER_FOOT);
    int interval = 0;
    for (int k = ND - 2; k >= 1; --k)
        if (log_d > log_npd_d[k]) {
            interval = k;
            break;
        }
    const double fraction = (log_d - log_npd_d[interval]) / (log_npd_d[interval + 1] - log_npd_d[interval]);
    double sel, scaled;
    row_read_d(rows[tr.row_index], interval, fraction, slant, sel, scaled);
    if (tr.weight_d != 0.0) {
        double sel_hi, scaled_hi;
        row_read_d(rows[tr.row_index + 1], interval, fraction, slant, sel_hi, scaled_hi);
        sel = sel + tr.weight_d * (sel_hi - sel);
        scaled = scaled + tr.weight_d * (scaled_hi - scaled);
    }
    if (!(slant < NPD_LAST_DISTANCE_M))
        sel = tr.sel_last_d - 20.0 * log10(slant / NPD_LAST_DISTANCE_M) - tr.sel_alpha_d * (slant - NPD_LAST_DISTANCE_M);
    const double finite = 10.0 * log10(finite_fraction_d(along * length, length, scaled));
    const double lat = lateral_d(lambda_height, lateral);
    const double installation = 10.0 * log10(installation_d(tr.installation, height, slant_in));
    const double free_sel = sel + tr.dv_d + installation - lat + finite;
    double terrain = 0.0, buildings = 0.0;
    if (horizon) {
        terrain = horizon_loss_d(horizon, QM_TERRAIN_SECTORS, on_line);
        buildings = horizon_loss_d(horizon + QM_TERRAIN_HORIZON_FLOATS, QM_BUILDING_SECTORS, on_seg);
    }
    return free_sel - fmax(fmax(terrain, buildings) - lat, 0.0);
}

// FlightPasses::add_pass's class in FP64 (popup/src/aircraft.rs:86-95).
__device__ int pass_class_d(double energy, double passing) {
#pragma clang fp contract(off)
    const double level_db = 10.0 * log10(energy / passing);
    const int c = (int)fmax((level_db - QM_PASS_LOWEST_DB) / QM_PASS_CLASS_DB, 0.0);
    return min(c, QM_PASS_CLASSES - 1);
}

// The reach rule in FP64 (the oracle's within_reach).
__device__ bool within_reach_d(const float* a, const float* b, double rx, double ry, double reach) {
#pragma clang fp contract(off)
    const double s0 = (double)a[0] - rx, s1 = (double)a[1] - ry;
    const double e0 = (double)b[0] - rx, e1 = (double)b[1] - ry;
    const double dx = e0 - s0, dy = e1 - s1;
    const double ll = dx * dx + dy * dy;
    double t = 0.0;
    if (ll > 0.0) t = fmin(fmax(-(s0 * dx + s1 * dy) / ll, 0.0), 1.0);
    const double px = s0 + t * dx, py = s1 + t * dy;
    return px * px + py * py <= reach * reach;
}

struct SegmentF {
    float sel_energy;  // 10^(SEL/10), 0 when not heard
    float lmax;        // -inf when not heard
};

// One segment's SEL energy and LAmax in FP32.
// `horizon`: the receiver's horizons when the segment is screened (its mean height above the
// ground under its ends below QM_SCREENING_CEILING_M), else null.
__device__ __forceinline__ SegmentF segment_f(const AircraftArgs& g, const DevTrack& tr, const DevRow* rows,
                                             const float* a, const float* b, float rxf, float ryf, double rx,
                                             double ry, const float* horizon) {
    SegmentF out{0.0f, -INFINITY};
    const float s[3] = {a[0] - rxf, a[1] - ryf, a[2] - (float)g.receiver_altitude_m};
    const float e[3] = {b[0] - rxf, b[1] - ryf, b[2] - (float)g.receiver_altitude_m};
    const float d[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
    const float reach = (float)g.reach_m;
    {
        const float ll = d[0] * d[0] + d[1] * d[1];
        const float t = ll > 0.0f ? fminf(fmaxf(div_f(-(s[0] * d[0] + s[1] * d[1]), ll), 0.0f), 1.0f) : 0.0f;
        const float px = s[0] + t * d[0], py = s[1] + t * d[1];
        const float q = px * px + py * py;
        if (q > (reach + 1.0f) * (reach + 1.0f)) return out;
        if (q > (reach - 1.0f) * (reach - 1.0f) && !within_reach_d(a, b, rx, ry, g.reach_m)) return out;
    }
    const float hs = d[0] * d[0] + d[1] * d[1];
    const float ls = hs + d[2] * d[2];
    float along, on_line[3], on_seg[3], lateral, height, lambda_height;
    if (hs <= 1e-6f) {
        along = 0.0f;
        for (int k = 0; k < 3; ++k) on_line[k] = on_seg[k] = s[k];
        lateral = sqrt_f(s[0] * s[0] + s[1] * s[1]);
        height = s[2];
        lambda_height = s[2];
    } else {
        along = div_f(-(s[0] * d[0] + s[1] * d[1] + s[2] * d[2]), ls);
        const float cl = fminf(fmaxf(along, 0.0f), 1.0f);
        for (int k = 0; k < 3; ++k) {
            on_line[k] = s[k] + along * d[k];
            on_seg[k] = s[k] + cl * d[k];
        }
        lateral = fabsf(s[0] * d[1] - s[1] * d[0]) * rsqrtf(hs);
        const float ss = on_line[0] * on_line[0] + on_line[1] * on_line[1] + on_line[2] * on_line[2];
        height = copysignf(sqrt_f(fmaxf(ss - lateral * lateral, 0.0f)), on_line[2]);
        lambda_height = (along >= 0.0f && along <= 1.0f) ? height : on_seg[2] * sqrt_f(div_f(ls, hs));
    }
    const float length = fmaxf(sqrt_f(ls), 1.0f);
    const float slant = sqrt_f(on_line[0] * on_line[0] + on_line[1] * on_line[1] + on_line[2] * on_line[2]);
    // SegmentEmission::read_npd (segment.rs:126-146).
    int interval;
    float fraction, ps;
    position_f(g.tables, slant, interval, fraction, ps);
    const DevRow& lo = rows[tr.row_index];
    float sel, scaled;
    {
        const float sml = linear_f(lo.sml, interval, fraction);
        float s_lo = ps < (float)NPD_LAST_DISTANCE_M ? linear_f(lo.sel, interval, fraction)
                                                    : lo.sel[ND - 1] - 20.0f * log10_f(ps * (float)(1.0 / NPD_LAST_DISTANCE_M)) -
                                                          lo.tail_alpha * (ps - (float)NPD_LAST_DISTANCE_M);
        float d_lo = lo.dipole ? ps : lo.d0 * exp10_f(sml * 0.1f);
        if (tr.weight != 0.0f) {
            const DevRow& hi = rows[tr.row_index + 1];
            const float smh = linear_f(hi.sml, interval, fraction);
            const float s_hi = ps < (float)NPD_LAST_DISTANCE_M
                                   ? linear_f(hi.sel, interval, fraction)
                                   : hi.sel[ND - 1] - 20.0f * log10_f(ps * (float)(1.0 / NPD_LAST_DISTANCE_M)) -
                                         hi.tail_alpha * (ps - (float)NPD_LAST_DISTANCE_M);
            const float d_hi = hi.dipole ? ps : hi.d0 * exp10_f(smh * 0.1f);
            s_lo = s_lo + tr.weight * (s_hi - s_lo);
            d_lo = d_lo + tr.weight * (d_hi - d_lo);
        }
        sel = ps < (float)NPD_LAST_DISTANCE_M
                  ? s_lo
                  : tr.sel_last - 20.0f * log10_f(ps * (float)(1.0 / NPD_LAST_DISTANCE_M)) -
                        tr.sel_alpha * (ps - (float)NPD_LAST_DISTANCE_M);
        scaled = d_lo;
    }
    const float finite = finite_fraction_f(along * length, length, scaled);
    const float lat = lateral_f(lambda_height, lateral);
    const float inst = installation_f(tr.installation, height, slant);
    // 10^(SEL/10) with SEL = NPD + Delta_V + Delta_I - Lambda + Delta_F (segment.rs:318-350).
    float screening = 0.0f;
    if (horizon) {
        // segment.rs:296-313: the terrain horizon at the line's closest point, the buildings' at
        // the segment's; their larger loss beyond the lateral attenuation (screening.rs:49-56).
        const float terrain = horizon_loss_f(horizon, QM_TERRAIN_SECTORS, on_line[0], on_line[1], on_line[2]);
        const float buildings = horizon_loss_f(horizon + QM_TERRAIN_HORIZON_FLOATS, QM_BUILDING_SECTORS, on_seg[0],
                                               on_seg[1], on_seg[2]);
        screening = fmaxf(fmaxf(terrain, buildings) - lat, 0.0f);
    }
    out.sel_energy = exp10_f((sel + tr.dv - lat - screening) * 0.1f) * inst * finite;
    // LAmax (segment.rs:357-381).
    const float distance = sqrt_f(on_seg[0] * on_seg[0] + on_seg[1] * on_seg[1] + on_seg[2] * on_seg[2]);
    float lamax;
    {
        int li;
        float lf, lps;
        position_f(g.tables, distance, li, lf, lps);
        if (lps < (float)NPD_LAST_DISTANCE_M) {
            lamax = linear_f(lo.lamax, li, lf);
            if (tr.weight != 0.0f) {
                const DevRow& hi = rows[tr.row_index + 1];
                lamax = lamax + tr.weight * (linear_f(hi.lamax, li, lf) - lamax);
            }
        } else {
            lamax = tr.lamax_last - 20.0f * log10_f(lps * (float)(1.0 / NPD_LAST_DISTANCE_M)) -
                    tr.lamax_alpha * (lps - (float)NPD_LAST_DISTANCE_M);
        }
    }
    // Within the segment's span the event reads the lateral attenuation at the closest point,
    // the SEL's own (lambda_height is the height there).
    const float lat2 = (along >= 0.0f && along <= 1.0f)
                           ? lat
                           : lateral_f(on_seg[2], sqrt_f(on_seg[0] * on_seg[0] + on_seg[1] * on_seg[1]));
    out.lmax = lamax + 10.0f * log10_f(installation_f(tr.installation, height, slant)) - lat2;
    return out;
}

}  // namespace

// PASSES: the pass classes are summed (g.passes is not null); SCREEN: the receivers' horizons
// screen the segments (g.horizons is not null). The kernel without either keeps the registers of
// the energies, counts and histogram alone.
template <bool PASSES, bool SCREEN>
__global__ __launch_bounds__(256) void aircraft_kernel(AircraftArgs g) {
    __shared__ uint32_t s_tile;
    const int lane = threadIdx.x;
    const int wave = lane >> 6, l = lane & 63;
    const int lx = (wave & 1) * 8 + (l & 7), ly = (wave >> 1) * 8 + (l >> 3);
    const int xcd = blockIdx.x % qm::XCDS;
    for (;;) {
        if (lane == 0) s_tile = qm::next_tile(g.counters, g.n_tiles, xcd);
        __syncthreads();
        const uint32_t tile = s_tile;
        __syncthreads();
        if (tile >= g.n_tiles) break;
        const uint32_t tx = tile % g.tiles_x, ty = tile / g.tiles_x;
        const uint32_t ix = tx * 16 + lx, iy = ty * 16 + ly;
        const bool valid = ix < g.nx && iy < g.ny;
        const size_t index = valid ? (size_t)iy * g.nx + ix : 0;
        const double rx = g.points[2 * index], ry = g.points[2 * index + 1];
        const float rxf = (float)rx, ryf = (float)ry;
        uint32_t* hist = g.hist + index * QM_HIST_BINS;
        if (valid)
            for (int k = 0; k < QM_HIST_BINS; ++k) hist[k] = 0u;
        const float* horizon = nullptr;
        if (SCREEN && valid && g.horizon_of[index] != QM_NO_HORIZON)
            horizon = g.horizons + (size_t)g.horizon_of[index] * QM_HORIZON_FLOATS;
        // Pass classes: a run of one class per period sums in registers and is added to the
        // receiver's row when the class changes.
        double* pass_row = (PASSES && valid) ? g.passes + index * QM_PASS_FLOATS : nullptr;
        if (pass_row)
            for (int k = 0; k < QM_PASS_FLOATS; ++k) pass_row[k] = 0.0;
        int run_class[3] = {-1, -1, -1};
        double run_energy[3] = {0.0, 0.0, 0.0}, run_passing[3] = {0.0, 0.0, 0.0};
        auto flush = [&](int p) {
            if (run_class[p] < 0) return;
            pass_row[p * QM_PASS_CLASSES + run_class[p]] += run_energy[p];
            pass_row[(3 + p) * QM_PASS_CLASSES + run_class[p]] += run_passing[p];
            run_energy[p] = run_passing[p] = 0.0;
        };
        // The tile's box (uniform).
        const uint32_t ix0 = tx * 16, iy0 = ty * 16;
        const uint32_t ix1 = min(ix0 + 15, g.nx - 1), iy1 = min(iy0 + 15, g.ny - 1);
        const float bx0 = (float)g.points[2 * ((size_t)iy0 * g.nx + ix0)];
        const float by0 = (float)g.points[2 * ((size_t)iy0 * g.nx + ix0) + 1];
        const float bx1 = (float)g.points[2 * ((size_t)iy1 * g.nx + ix1)];
        const float by1 = (float)g.points[2 * ((size_t)iy1 * g.nx + ix1) + 1];
        const float reach = (float)g.reach_m + 2.0f;
        double energy[3] = {0.0, 0.0, 0.0};
        uint32_t n65 = 0, n70 = 0;
        for (uint32_t t = 0; t < g.n_tracks; ++t) {
            const DevTrack tr = g.tracks[t];
            if (tr.x0 > bx1 + reach || tr.x1 < bx0 - reach || tr.y0 > by1 + reach || tr.y1 < by0 - reach) continue;
            float e_track = 0.0f;
            float m1 = -INFINITY, m2 = -INFINITY, m3 = -INFINITY;
            uint32_t i1 = 0, i2 = 0, i3 = 0;
            bool any = false;
            if (valid) {
                const float* v = g.vertices + 3 * (size_t)tr.first_vertex;
                for (uint32_t k = 0; k + 1 < tr.vertex_count; ++k) {
                    // Screened below the ceiling over the ground under its ends (segment.rs:296-313).
                    const float* screen = nullptr;
                    if (SCREEN && horizon) {
                        const size_t va = (size_t)tr.first_vertex + k;
                        const double ga = g.vertex_ground ? (double)g.vertex_ground[va] : 0.0;
                        const double gb = g.vertex_ground ? (double)g.vertex_ground[va + 1] : 0.0;
                        const double above = 0.5 * ((double)v[3 * k + 2] - ga + (double)v[3 * k + 5] - gb);
                        if (above < QM_SCREENING_CEILING_M) screen = horizon;
                    }
                    const SegmentF sf = segment_f(g, tr, g.rows, v + 3 * k, v + 3 * k + 3, rxf, ryf, rx, ry, screen);
                    if (sf.lmax == -INFINITY) continue;
                    any = true;
                    e_track += sf.sel_energy;
                    if (PASSES && pass_row) {
                        // The pass's class (aircraft.cuh flight_sums, pass_class): the FP32 level
                        // decides it unless within EPS_DB of a class edge, where the FP64 SEL does.
                        const float* pa = v + 3 * k;
                        const double d0 = (double)pa[3] - (double)pa[0], d1 = (double)pa[4] - (double)pa[1],
                                     d2 = (double)pa[5] - (double)pa[2];
                        const double length = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                        const float level = 10.0f * log10_f(div_f(sf.sel_energy, (float)(length * tr.seconds_per_metre)));
                        const float position = (level - (float)QM_PASS_LOWEST_DB) / (float)QM_PASS_CLASS_DB;
                        const int fast_class = min((int)fmaxf(position, 0.0f), QM_PASS_CLASSES - 1);
                        const float edge = rintf(position);
                        const bool near_edge =
                            edge >= 1.0f && edge <= (float)(QM_PASS_CLASSES - 1) &&
                            fabsf(level - ((float)QM_PASS_LOWEST_DB + (float)QM_PASS_CLASS_DB * edge)) < EPS_DB;
                        double exact = 0.0;
                        if (near_edge) {
                            double log_npd_d[ND];
                            for (int n = 0; n < ND; ++n) log_npd_d[n] = g.tables.log_npd_d[n];
                            const DevTrack track = tr;
                            exact = energy_d(segment_sel_d(g.receiver_altitude_m, log_npd_d, track, g.rows, pa, pa + 3,
                                                           rx, ry, screen));
                        }
                        for (int p = 0; p < 3; ++p) {
                            const double seconds = g.seconds[p];
                            const double e_p = (double)tr.flights[p] * (double)sf.sel_energy / seconds;
                            const double passing = (double)tr.flights[p] / seconds * length * tr.seconds_per_metre;
                            if (!(e_p > 0.0 && passing > 0.0)) continue;
                            const int c = near_edge ? pass_class_d((double)tr.flights[p] * exact / seconds, passing)
                                                    : fast_class;
                            if (c != run_class[p]) {
                                flush(p);
                                run_class[p] = c;
                            }
                            run_energy[p] += e_p;
                            run_passing[p] += passing;
                        }
                    }
                    if (sf.lmax > m1) {
                        m3 = m2, i3 = i2, m2 = m1, i2 = i1, m1 = sf.lmax, i1 = k;
                    } else if (sf.lmax > m2) {
                        m3 = m2, i3 = i2, m2 = sf.lmax, i2 = k;
                    } else if (sf.lmax > m3) {
                        m3 = sf.lmax, i3 = k;
                    }
                }
            }
            if (!any) continue;
            for (int p = 0; p < 3; ++p) energy[p] += (double)tr.flights[p] * (double)e_track;
            if (m1 < (float)QM_HIST_FLOOR_DB - EPS_DB) continue;
            // The FP32 maximum decides N65, N70 and the bin wherever it lies EPS_DB or more from
            // every 0.5 dB edge (65 and 70 dB are such edges): the FP64 maximum is then on the same
            // side of each. Within EPS_DB of one, the event's LAmax is re-read in FP64 over its
            // candidate segments.
            const float steps = rintf((m1 - (float)QM_HIST_FLOOR_DB) * (float)(1.0 / QM_HIST_STEP_DB));
            const bool near_edge = fabsf(m1 - ((float)QM_HIST_FLOOR_DB + (float)QM_HIST_STEP_DB * steps)) < EPS_DB;
            double lmax = (double)m1;
            const float* v = g.vertices + 3 * (size_t)tr.first_vertex;
            if (near_edge) lmax = segment_lmax_d(g, tr, g.rows, v + 3 * i1, v + 3 * i1 + 3, rx, ry);
            if (near_edge && m2 >= m1 - EPS_DB)
                lmax = fmax(lmax, segment_lmax_d(g, tr, g.rows, v + 3 * i2, v + 3 * i2 + 3, rx, ry));
            if (near_edge && m3 >= m1 - EPS_DB) {
                // Three or more within reach of the maximum: every such segment again.
                for (uint32_t k = 0; k + 1 < tr.vertex_count; ++k) {
                    const SegmentF sf = segment_f(g, tr, g.rows, v + 3 * k, v + 3 * k + 3, rxf, ryf, rx, ry, nullptr);
                    if (sf.lmax >= m1 - EPS_DB)
                        lmax = fmax(lmax, segment_lmax_d(g, tr, g.rows, v + 3 * k, v + 3 * k + 3, rx, ry));
                }
            }
            if (lmax >= 65.0) n65 += tr.total_flights;
            if (lmax >= 70.0) n70 += tr.total_flights;
            if (lmax >= QM_HIST_FLOOR_DB) {
                const double kb = floor((lmax - QM_HIST_FLOOR_DB) / QM_HIST_STEP_DB);
                const int bin = kb >= QM_HIST_BINS - 1 ? QM_HIST_BINS - 1 : (int)kb;
                hist[bin] += tr.total_flights;
            }
        }
        if (PASSES && pass_row)
            for (int p = 0; p < 3; ++p) flush(p);
        if (valid) {
            for (int p = 0; p < 3; ++p) g.energy[index * 3 + p] = energy[p];
            g.n65[index] = n65;
            g.n70[index] = n70;
        }
    }
}

// =============================================================================================
// Host side.

namespace {

int afail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

#define A_TRY(call)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return afail(err, errlen, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

const double NPD_FT[ND] = {200.0, 400.0, 630.0, 1000.0, 2000.0, 4000.0, 6300.0, 10000.0, 16000.0, 25000.0};

// The host-side power row (npd.rs:166-188, 290-310) in FP64.
struct HostRow {
    double sel[ND], lamax[ND], sml[ND], tail_alpha, d0;
    bool dipole;
};

HostRow host_row(const double* s, const double* l, double v_ref_kt) {
    HostRow r;
    bool placeholder = true;
    for (int k = 0; k < ND; ++k) {
        r.sel[k] = s[k];
        r.lamax[k] = l[k];
        r.sml[k] = s[k] - l[k];
        if (!(std::fabs(r.sml[k] - 12.0) < 1e-9)) placeholder = false;
    }
    r.dipole = placeholder;
    r.d0 = 2.0 / PI_D * (v_ref_kt * 1852.0 / 3600.0) * 1.0;
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (int i = 0; i < 3; ++i) {
        const int k = ND - 3 + i;
        const double d = NPD_FT[k] * METRES_PER_FOOT;
        const double y = s[k] + 20.0 * std::log10(d / NPD_LAST_DISTANCE_M);
        sx += d;
        sy += y;
        sxx += d * d;
        sxy += d * y;
    }
    r.tail_alpha = std::fmax(-(3.0 * sxy - sx * sy) / (3.0 * sxx - sx * sx), 0.0);
    return r;
}

// Level of a row at a slant (NpdPosition::level), FP64.
double host_level(const double* curve, double tail_alpha, double slant_m) {
    slant_m = std::fmax(slant_m, NPD_NEAREST_SLANT_M);
    if (slant_m < NPD_LAST_DISTANCE_M) {
        const double log_d = std::log10(slant_m / METRES_PER_FOOT);
        int interval = 0;
        for (int k = ND - 2; k >= 1; --k)
            if (log_d > std::log10(NPD_FT[k])) {
                interval = k;
                break;
            }
        const double f = (log_d - std::log10(NPD_FT[interval])) / (std::log10(NPD_FT[interval + 1]) - std::log10(NPD_FT[interval]));
        return curve[interval] + f * (curve[interval + 1] - curve[interval]);
    }
    return curve[ND - 1] - 20.0 * std::log10(slant_m / NPD_LAST_DISTANCE_M) - tail_alpha * (slant_m - NPD_LAST_DISTANCE_M);
}

double spherical(double s) { return 20.0 * std::log10(s / NPD_LAST_DISTANCE_M); }

}  // namespace

extern "C" int qm_aircraft_grid(qm_gpu* gpu, const qm_aircraft_class* classes, size_t n_classes, const qm_track* tracks,
                                size_t n_tracks, const float* vertices, size_t n_vertices, const double* points_xy,
                                uint32_t nx, uint32_t ny, double receiver_altitude_m, double reach_m,
                                double* period_energy, uint32_t* n65, uint32_t* n70, uint32_t* hist,
                                double* passes, const qm_aircraft_screening* screening, char* err,
                                size_t errlen) {
    if (!gpu || (n_tracks && (!classes || !tracks || !vertices)) || !points_xy || !period_energy || !n65 || !n70 || !hist)
        return afail(err, errlen, "null argument");
    const size_t n = (size_t)nx * ny;
    if (n == 0) return 0;
    if (!(reach_m > 1.0)) return afail(err, errlen, "reach must exceed 1 m");
    // Rows and tracks with their constants, in FP64 on the host.
    std::vector<DevRow> rows;
    std::vector<size_t> class_first(n_classes * 2);
    std::vector<HostRow> host_rows;
    for (size_t c = 0; c < n_classes; ++c) {
        if (classes[c].installation > QM_INSTALL_PROPELLER) return afail(err, errlen, "class %zu: installation", c);
        for (int dep = 0; dep < 2; ++dep) {
            if (classes[c].rows[dep] < 1 || classes[c].rows[dep] > QM_POWER_ROWS)
                return afail(err, errlen, "class %zu: %u power rows", c, classes[c].rows[dep]);
            class_first[2 * c + dep] = rows.size();
            for (uint32_t r = 0; r < classes[c].rows[dep]; ++r) {
                const HostRow h = host_row(classes[c].sel_db[dep][r], classes[c].lamax_db[dep][r], classes[c].v_ref_kt);
                host_rows.push_back(h);
                DevRow d;
                std::memset(&d, 0, sizeof d);
                for (int k = 0; k < ND; ++k) {
                    d.sel[k] = (float)h.sel[k];
                    d.lamax[k] = (float)h.lamax[k];
                    d.sml[k] = (float)h.sml[k];
                    d.lamax_d[k] = h.lamax[k];
                    d.sel_d[k] = h.sel[k];
                    d.sml_d[k] = h.sml[k];
                }
                d.d0_d = h.d0;
                d.tail_alpha = (float)h.tail_alpha;
                d.tail_alpha_d = h.tail_alpha;
                d.d0 = (float)h.d0;
                d.dipole = h.dipole ? 1u : 0u;
                rows.push_back(d);
            }
        }
    }
    std::vector<DevTrack> dev_tracks(n_tracks);
    for (size_t t = 0; t < n_tracks; ++t) {
        const qm_track& tr = tracks[t];
        if (tr.aircraft_class >= n_classes || tr.departure > 1 || tr.vertex_count < 2 ||
            (size_t)tr.first_vertex + tr.vertex_count > n_vertices || !(tr.power_weight >= 0.0f && tr.power_weight <= 1.0f))
            return afail(err, errlen, "track %zu: out of range", t);
        const qm_aircraft_class& c = classes[tr.aircraft_class];
        if (tr.power_row + (tr.power_weight > 0.0f ? 1u : 0u) >= c.rows[tr.departure])
            return afail(err, errlen, "track %zu: power bracket beyond its class's rows", t);
        DevTrack d;
        std::memset(&d, 0, sizeof d);
        float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
        for (uint32_t v = 0; v < tr.vertex_count; ++v) {
            const float* p = vertices + 3 * ((size_t)tr.first_vertex + v);
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]))
                return afail(err, errlen, "track %zu: vertex not finite", t);
            x0 = std::fmin(x0, p[0]);
            y0 = std::fmin(y0, p[1]);
            x1 = std::fmax(x1, p[0]);
            y1 = std::fmax(y1, p[1]);
        }
        d.x0 = x0;
        d.y0 = y0;
        d.x1 = x1;
        d.y1 = y1;
        d.first_vertex = tr.first_vertex;
        d.vertex_count = tr.vertex_count;
        d.row_index = (uint32_t)(class_first[2 * tr.aircraft_class + tr.departure] + tr.power_row);
        d.installation = c.installation;
        d.weight = tr.power_weight;
        d.weight_d = (double)tr.power_weight;
        d.dv_d = (double)tr.speed_kt > 10.0 ? 10.0 * std::log10(c.v_ref_kt / (double)tr.speed_kt) : 0.0;
        d.dv = (float)d.dv_d;
        d.seconds_per_metre = 3600.0 / (((double)tr.speed_kt > 10.0 ? (double)tr.speed_kt : c.v_ref_kt) * 1852.0);
        // table(25,000 ft) and table(16 km) of SEL and LAmax, power-interpolated (segment.rs:126-162).
        const HostRow& lo = host_rows[d.row_index];
        auto at = [&](const double* lo_curve, const double* hi_curve, double lo_alpha, double hi_alpha, double s) {
            const double low = host_level(lo_curve, lo_alpha, s);
            if (d.weight_d == 0.0) return low;
            return low + d.weight_d * (host_level(hi_curve, hi_alpha, s) - low);
        };
        const HostRow& hi = d.weight_d == 0.0 ? lo : host_rows[d.row_index + 1];
        const double sel_last = at(lo.sel, hi.sel, lo.tail_alpha, hi.tail_alpha, NPD_LAST_DISTANCE_M);
        const double sel_tail = at(lo.sel, hi.sel, lo.tail_alpha, hi.tail_alpha, TAIL_ANCHOR_M);
        const double lam_last = at(lo.lamax, hi.lamax, lo.tail_alpha, hi.tail_alpha, NPD_LAST_DISTANCE_M);
        const double lam_tail = at(lo.lamax, hi.lamax, lo.tail_alpha, hi.tail_alpha, TAIL_ANCHOR_M);
        const double sel_alpha = std::fmax((sel_last - spherical(TAIL_ANCHOR_M) - sel_tail) / (TAIL_ANCHOR_M - NPD_LAST_DISTANCE_M), 0.0);
        const double lam_alpha = std::fmax((lam_last - spherical(TAIL_ANCHOR_M) - lam_tail) / (TAIL_ANCHOR_M - NPD_LAST_DISTANCE_M), 0.0);
        d.sel_last = (float)sel_last;
        d.sel_alpha = (float)sel_alpha;
        d.lamax_last = (float)lam_last;
        d.lamax_alpha = (float)lam_alpha;
        d.lamax_last_d = lam_last;
        d.lamax_alpha_d = lam_alpha;
        d.sel_last_d = sel_last;
        d.sel_alpha_d = sel_alpha;
        for (int p = 0; p < 3; ++p) d.flights[p] = tr.flights[p];
        d.total_flights = tr.flights[0] + tr.flights[1] + tr.flights[2];
        dev_tracks[t] = d;
    }
    for (size_t i = 0; i < 2 * n; ++i)
        if (!std::isfinite(points_xy[i])) return afail(err, errlen, "point %zu: not finite", i / 2);
    const bool screen = screening && screening->horizons;
    if (screen) {
        if (!screening->horizon_of) return afail(err, errlen, "screening: horizons without horizon_of");
        for (size_t i = 0; i < n; ++i)
            if (screening->horizon_of[i] != QM_NO_HORIZON && screening->horizon_of[i] >= screening->n_horizons)
                return afail(err, errlen, "receiver %zu: horizon %u of %zu", i, screening->horizon_of[i],
                             screening->n_horizons);
    }
    A_TRY(hipSetDevice(gpu->ordinal));
    AircraftArgs g;
    std::memset(&g, 0, sizeof g);
    for (int k = 0; k < ND; ++k) {
        g.tables.log_npd_d[k] = std::log10(NPD_FT[k]);
        g.tables.log_npd[k] = (float)g.tables.log_npd_d[k];
    }
    void *d_rows = nullptr, *d_tracks = nullptr, *d_vertices = nullptr, *d_points = nullptr, *d_energy = nullptr,
         *d_n65 = nullptr, *d_n70 = nullptr, *d_hist = nullptr, *d_counters = nullptr, *d_passes = nullptr,
         *d_horizons = nullptr, *d_horizon_of = nullptr, *d_ground = nullptr;
    auto cleanup = [&]() {
        for (void* p : {d_rows, d_tracks, d_vertices, d_points, d_energy, d_n65, d_n70, d_hist, d_counters, d_passes,
                        d_horizons, d_horizon_of, d_ground})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipSuccess;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes ? bytes : 1);
    };
    alloc(&d_rows, sizeof(DevRow) * rows.size());
    alloc(&d_tracks, sizeof(DevTrack) * n_tracks);
    alloc(&d_vertices, sizeof(float) * 3 * n_vertices);
    alloc(&d_points, sizeof(double) * 2 * n);
    alloc(&d_energy, sizeof(double) * 3 * n);
    alloc(&d_n65, sizeof(uint32_t) * n);
    alloc(&d_n70, sizeof(uint32_t) * n);
    alloc(&d_hist, sizeof(uint32_t) * QM_HIST_BINS * n);
    alloc(&d_counters, sizeof(uint32_t) * qm::XCDS);
    if (passes) alloc(&d_passes, sizeof(double) * QM_PASS_FLOATS * n);
    if (screen) {
        alloc(&d_horizons, sizeof(float) * QM_HORIZON_FLOATS * screening->n_horizons);
        alloc(&d_horizon_of, sizeof(uint32_t) * n);
        if (screening->vertex_ground_m) alloc(&d_ground, sizeof(float) * n_vertices);
        if (e == hipSuccess && screening->n_horizons)
            e = hipMemcpy(d_horizons, screening->horizons, sizeof(float) * QM_HORIZON_FLOATS * screening->n_horizons,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(d_horizon_of, screening->horizon_of, sizeof(uint32_t) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess && d_ground && n_vertices)
            e = hipMemcpy(d_ground, screening->vertex_ground_m, sizeof(float) * n_vertices, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && !rows.empty()) e = hipMemcpy(d_rows, rows.data(), sizeof(DevRow) * rows.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && n_tracks) e = hipMemcpy(d_tracks, dev_tracks.data(), sizeof(DevTrack) * n_tracks, hipMemcpyHostToDevice);
    if (e == hipSuccess && n_vertices) e = hipMemcpy(d_vertices, vertices, sizeof(float) * 3 * n_vertices, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_points, points_xy, sizeof(double) * 2 * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        return afail(err, errlen, "aircraft upload: %s", hipGetErrorString(e));
    }
    g.rows = (const DevRow*)d_rows;
    g.tracks = (const DevTrack*)d_tracks;
    g.n_tracks = (uint32_t)n_tracks;
    g.vertices = (const float*)d_vertices;
    g.points = (const double*)d_points;
    g.nx = nx;
    g.ny = ny;
    g.tiles_x = (nx + 15) / 16;
    g.n_tiles = g.tiles_x * ((ny + 15) / 16);
    g.receiver_altitude_m = receiver_altitude_m;
    g.reach_m = reach_m;
    g.energy = (double*)d_energy;
    g.n65 = (uint32_t*)d_n65;
    g.n70 = (uint32_t*)d_n70;
    g.hist = (uint32_t*)d_hist;
    g.passes = (double*)d_passes;
    g.horizons = (const float*)d_horizons;
    g.horizon_of = (const uint32_t*)d_horizon_of;
    g.vertex_ground = (const float*)d_ground;
    for (int p = 0; p < 3; ++p) g.seconds[p] = QM_PERIOD_HOURS[p] * 3600.0;
    g.counters = (uint32_t*)d_counters;
    hipLaunchKernelGGL(qm::zero_u32_kernel, dim3(1), dim3(64), 0, 0, g.counters, (uint32_t)qm::XCDS);
    int cus = 0, per_cu = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, gpu->ordinal);
    const void* kernels[2][2] = {{reinterpret_cast<const void*>(aircraft_kernel<false, false>),
                                  reinterpret_cast<const void*>(aircraft_kernel<false, true>)},
                                 {reinterpret_cast<const void*>(aircraft_kernel<true, false>),
                                  reinterpret_cast<const void*>(aircraft_kernel<true, true>)}};
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernels[passes != nullptr][screen], 256, 0);
    unsigned groups = (unsigned)std::max(1, cus * std::max(per_cu, 1));
    groups = std::min<unsigned>(groups, g.n_tiles);
    if (passes && screen)
        hipLaunchKernelGGL((aircraft_kernel<true, true>), dim3(groups), dim3(256), 0, 0, g);
    else if (passes)
        hipLaunchKernelGGL((aircraft_kernel<true, false>), dim3(groups), dim3(256), 0, 0, g);
    else if (screen)
        hipLaunchKernelGGL((aircraft_kernel<false, true>), dim3(groups), dim3(256), 0, 0, g);
    else
        hipLaunchKernelGGL((aircraft_kernel<false, false>), dim3(groups), dim3(256), 0, 0, g);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(period_energy, d_energy, sizeof(double) * 3 * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(n65, d_n65, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(n70, d_n70, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hist, d_hist, sizeof(uint32_t) * QM_HIST_BINS * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess && passes)
        e = hipMemcpy(passes, d_passes, sizeof(double) * QM_PASS_FLOATS * n, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return afail(err, errlen, "aircraft_kernel: %s", hipGetErrorString(e));
    return 0;
}
