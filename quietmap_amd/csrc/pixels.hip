// libqmpaint: pixel loudness on the card (include/qm_paint.h qm_pixel_loudness), the reference's
// Gpu::pixel_loudness (paint-gpu/src/loudness.rs). What it computes is the reference's CPU path:
// layered_distributions (popup/src/percentiles.rs:88, layer_states, moment), the level
// distributions of popup/src/distribution.rs:95-233 (runs, with_line, with_distribution, over),
// the line shapes of popup/src/distribution.rs:17-31,283-299 and popup/src/percentiles/lines.rs
// over Kurze's quantiles (physics/src/percentile.rs:122-202), and Nden per curve
// (popup/src/loudness.rs:122 Curve::at, :253 loudness, day).
//
// Design (gfx950, wave64).
// - A workgroup of four waves per pixel walks its three periods in turn. Within a period the
//   weather states that differ are dealt to the waves, at most three a wave (a state whose every
//   line keeps the previous state's mean is that state again, and counts twice where it is
//   mixed). A state's distributions live in registers: lane l holds bins l and l + 64 of the 90
//   two-decibel bins, each as (time share, share x intensity); silence is a wave-uniform share.
// - A line's shape (its relative intensities ascending, with FP64 prefix sums of share and of
//   share x value) is built once per (period, hour group) into LDS by wave 0 and applied by every
//   wave to each of its states that the line moves; two shape buffers alternate, so one barrier
//   a line keeps the waves in step.
// - Adding a line to a distribution is a gather. For each occupied source bin (taken in turn from
//   a ballot of the occupied bins, its values read from its lane), the lane of target bin b counts
//   the line's values that put the sum below b's upper energy edge (one binary search in LDS);
//   the left neighbour's count (a lane shuffle) is the count below b's lower edge, and the prefix
//   sums between the two counts are the share and energy that land in b. No atomics, and each
//   bin's sums come in source order.
// - A layer's state distribution joins the place's through the same gather over the layer's
//   occupied bins, compacted into an ascending value list by ballot prefix counts.
// - Mixed over states and hour groups, the four waves' period mixtures are summed in LDS in wave
//   order. Nden reads each curve at the whole decibels its bins fall between; every step a bin
//   needs is marked, the marked steps compacted, and each evaluated once by ISO 532-1
//   (qm_zwicker.hpp) across the lanes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_internal.hpp"
#include "qm_zwicker.hpp"

namespace {

constexpr int PERIODS = 3;
constexpr int STATES = 10;
constexpr int LAYERS = 6;
constexpr int CURVES = QM_LOUD_CURVES;
constexpr int THIRDS = 28;
constexpr int BINS = QM_LOUD_BINS;
constexpr double BIN_DB = 2.0;
constexpr double LEVEL_MIN_DB = -30.0;
constexpr int ROWS = 41;
constexpr int QUANTILES = 100;
constexpr double LAMBDA_MIN = 1e-4;
constexpr double STEPS_PER_DECADE = 8.0;
constexpr int NODES = 124;           // node_probabilities: 99 even, 24 in the tail, one last
constexpr int PROFILES = 3;          // traffic.rs HOURLY_SHARES
constexpr int MAX_GROUPS = 4;        // the map's hour groups per period: 4, 1, 3
constexpr double HOLDS_MEAN_LAMBDA = 10.0;
constexpr double PI_D = 3.14159265358979323846;
constexpr double PASSES_HOLD_MEAN = 2.0 * PI_D * HOLDS_MEAN_LAMBDA;
constexpr double HELD_EXCEEDED = 1e-3;
constexpr double HELD_SHARE = 1e-2;
constexpr int CURVE_STEPS = 191;     // whole decibels from -30 to 160 dB
constexpr int LINES_MAX = 65535;     // lines a layer may carry

// A packed line (paint/src/sone.rs pack_line): energy per period, the mean in each weather state
// per period, lambda per period, the hourly profile (or -1), duty per period (or NaN), passes in
// progress per period (or NaN).
constexpr int LINE_FLOATS = 43;
constexpr int AT_MEANS = 3, AT_LAMBDA = 33, AT_PROFILE = 36, AT_DUTY = 37, AT_PASSING = 40;

constexpr int WAVES = STATES + 1;    // a wave per weather state and the shape builder
constexpr int BUILDER = STATES;
constexpr int GROUP = 64 * WAVES;
constexpr int SHAPE_CAP = 128;       // the longest shape: 128 pass values below the hold mean
constexpr int LIST_CAP = 96;         // a distribution as values: silence + 90 bins
constexpr int STEP_SLOTS = 192;

const float PERIOD_HOURS_F[PERIODS] = {12.0f, 4.0f, 8.0f};
const float PERIOD_PENALTY_F[PERIODS] = {0.0f, 5.0f, 10.0f};

// Everything the kernel reads besides the pixels, built on the host per call.
struct Tables {
    double node_cum[NODES + 1];             // shares of the nodes before each
    float node_share[NODES];
    float node_ln_p[NODES];                 // ln p of each node
    float top_ln_p;                         // ln of the table's last quantile probability
    float exceed_ln_p;                      // ln(1 - HELD_EXCEEDED)
    int first_tail;                         // nodes from here lie above the last quantile
    int pad;
    float log_rows[NODES + 1][ROWS];        // ln intensity at each node's quantile, per row; [NODES]: the top
    int groups[PERIODS];                    // hour groups of a pixel with a profiled line
    float group_weight[PERIODS][MAX_GROUPS];       // a group's hours / (period hours x STATES)
    double factor[PROFILES][PERIODS][MAX_GROUPS];  // a profile's mean flow factor over a group
    double ln_factorial[SHAPE_CAP + 1];     // ln c! for the passes' Poisson shares
    double fold_top[ROWS];                  // the map fold's FP64 ln intensity at the last quantile, per row
    double fold_ln_top, fold_ln_exceed;     // ln of the last quantile's and the held test's probability
};

// A list of values ascending with the prefix sums before each (LineValues).
template <int CAP>
struct Values {
    float v[CAP];
    double q[CAP + 1];  // shares of the values before each
    double w[CAP + 1];  // share x value of the values before each
    int n;
    int zeros;          // values <= 0 (silent over silence)
};

struct Shared {
    Values<SHAPE_CAP> shape[2];      // line shapes, alternating
    Values<LIST_CAP> list[STATES];   // a state wave's layer distribution as values
    float acc[STATES][96][2];        // a state wave's distribution being summed
    float2 sources[STATES][96];      // a state wave's distribution as (intensity, share) sources
    float mix[CURVES][BINS][2];      // the period's mixtures
    struct {
        double floor, compare;
    } pre[LAYERS][STATES];           // a layer's floor and the held test's energy, per state
    unsigned char moves[LAYERS][STATES];  // whether any of a layer's lines moves, per state
    uint32_t layer_at[LAYERS];       // where each layer of the pixel starts
    uint32_t layer_lines[LAYERS];    // its lines
    float nden[CURVES + 1];
    float sone[CURVES][STEP_SLOTS];
    unsigned short steps[CURVES][STEP_SLOTS];
    unsigned char need[CURVES][STEP_SLOTS];
};

// A distribution held across a wave: bins lane and lane + 64 as (share, share x intensity).
struct Dist {
    float s0, w0, s1, w1;
    float silent;
};

enum Kind { PASS_MOVES, PASS_HOLDS, DUTY_MOVES, DUTY_SILENT, KURZE_SHAPED, KURZE_HOLDS };

__device__ __forceinline__ float read_lane(float x, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}
__device__ __forceinline__ int read_lane_i(int x, int k) { return __builtin_amdgcn_readlane(x, k); }
__device__ __forceinline__ double read_lane_d(double x, int k) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ unsigned long long lanes_below(int lane) { return (1ull << lane) - 1ull; }
// Lane i gets lane i - 1's x; lane 0 gets `first` (a wave shift right by one).
__device__ __forceinline__ int from_lane_below(int x, int first) {
    return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false);
}

// LDS written by some lanes of a wave and read by others of the same wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane i gets lane i - d's x within its row of 16 lanes, 0 below the row's start (DPP row_shr).
template <int D>
__device__ __forceinline__ double row_shift(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), 0x110 + D, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + D, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Inclusive prefix sum across the wave: within each row of 16 lanes by four DPP shifts, then each
// row raised by the rows below it (their last lanes read as scalars).
__device__ __forceinline__ double wave_scan(double x, int lane) {
    x += row_shift<1>(x);
    x += row_shift<2>(x);
    x += row_shift<4>(x);
    x += row_shift<8>(x);
    const double r0 = read_lane_d(x, 15), r1 = read_lane_d(x, 31), r2 = read_lane_d(x, 47);
    const int row = lane >> 4;
    return x + (row == 0 ? 0.0 : (row == 1 ? r0 : (row == 2 ? r0 + r1 : r0 + r1 + r2)));
}

__device__ __forceinline__ float wave_sum(float x) {
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d);
    return x;
}

// The kind of a line in a period at its hour group's flow factor (lines.rs Line::shape and
// layer_states' sorting of a layer's lines).
__device__ Kind kind_of(const float* line, int p, double factor) {
    if (!isnan(line[AT_PASSING])) {
        const double mean = (double)line[AT_PASSING + p];
        return mean >= 2.2250738585072014e-308 && mean < PASSES_HOLD_MEAN ? PASS_MOVES : PASS_HOLDS;
    }
    if (!isnan(line[AT_DUTY])) return line[AT_DUTY + p] > 0.0f ? DUTY_MOVES : DUTY_SILENT;
    const double lambda = (double)line[AT_LAMBDA + p] * factor;
    return lambda >= HOLDS_MEAN_LAMBDA ? KURZE_HOLDS : KURZE_SHAPED;
}

__device__ __forceinline__ double factor_of(const Tables& T, const float* line, int p, int g, bool profiled) {
    const int profile = (int)line[AT_PROFILE];
    return profiled && profile >= 0 ? T.factor[profile][p][g] : 1.0;
}

// Kurze's relative intensity at a node (physics/src/percentile.rs relative_intensity): within the
// table, the log intensity between the rows about lambda; above its last quantile, the top's
// intensity plus the nearest emitter's passing; sparser than the table, scaled down with lambda.
__device__ float table_intensity(const Tables& T, int node, float lambda) {
    const float position = fminf(fmaxf(log10f(lambda / (float)LAMBDA_MIN) * (float)STEPS_PER_DECADE, 0.0f),
                                 (float)(ROWS - 1));
    const int row = (int)position;
    const float fraction = position - (float)row;
    const float lo = T.log_rows[node][row], hi = T.log_rows[node][min(row + 1, ROWS - 1)];
    const float v = expf(lo + fraction * (hi - lo));
    return lambda < (float)LAMBDA_MIN ? lambda / (float)LAMBDA_MIN * v : v;
}
__device__ __forceinline__ float passing_intensity(float lambda, float ln_p) {
    const float s = -ln_p / (2.0f * lambda);
    return 1.0f / (lambda * (float)PI_D * (1.0f + s * s));
}
__device__ float relative_intensity(const Tables& T, float lambda, int node) {
    if (isnan(lambda)) return 1.0f;
    if (!(lambda > 0.0f)) return 0.0f;
    if (node < T.first_tail) return table_intensity(T, node, lambda);
    return table_intensity(T, NODES, lambda) + passing_intensity(lambda, T.node_ln_p[node]) -
           passing_intensity(lambda, T.top_ln_p);
}
__device__ float exceeded_intensity(const Tables& T, float lambda) {
    if (isnan(lambda)) return 1.0f;
    if (!(lambda > 0.0f)) return 0.0f;
    return table_intensity(T, NODES, lambda) + passing_intensity(lambda, T.exceed_ln_p) -
           passing_intensity(lambda, T.top_ln_p);
}

// A line's shape into `out` (wave 0 of the group): the Kurze values at the nodes (a lane per
// node), an event's silence and loud share, or the passes in progress (distribution.rs
// pass_values, in FP64 on one lane as the Rust loop runs).
__device__ void build_shape(const Tables& T, const float* line, Kind kind, int p, double factor,
                            Values<SHAPE_CAP>& out, int lane) {
    if (kind == KURZE_SHAPED) {
        const float lambda = (float)((double)line[AT_LAMBDA + p] * factor);
        const int n0 = lane, n1 = lane + 64;
        const float v0 = relative_intensity(T, lambda, n0);
        const float v1 = n1 < NODES ? relative_intensity(T, lambda, n1) : 0.0f;
        out.v[n0] = v0;
        if (n1 < NODES) out.v[n1] = v1;
        const double x0 = (double)T.node_share[n0] * (double)v0;
        const double x1 = n1 < NODES ? (double)T.node_share[n1] * (double)v1 : 0.0;
        const double i0 = wave_scan(x0, lane), i1 = wave_scan(x1, lane);
        const double half = read_lane_d(i0, 63);
        out.w[n0 + 1] = i0;
        if (n1 < NODES) out.w[n1 + 1] = half + i1;
        for (int k = lane; k <= NODES; k += 64) out.q[k] = T.node_cum[k];
        const int zeros = __popcll(__ballot(v0 <= 0.0f)) + __popcll(__ballot(n1 < NODES && v1 <= 0.0f));
        if (lane == 0) {
            out.w[0] = 0.0;
            out.n = NODES;
            out.zeros = zeros;
        }
        return;
    }
    if (kind == DUTY_MOVES) {
        if (lane == 0) {
            const double duty = (double)line[AT_DUTY + p];
            out.v[0] = 0.0f;
            out.v[1] = (float)(1.0 / duty);
            out.q[0] = out.w[0] = 0.0;
            out.q[1] = 1.0 - duty;
            out.w[1] = 0.0;
            out.q[2] = out.q[1] + duty;
            out.w[2] = duty * (1.0 / duty);
            out.n = 2;
            out.zeros = 1;
        }
        return;
    }
    // PASS_MOVES: the passes in progress (distribution.rs pass_values), the counts' Poisson
    // shares in parallel (count c = lane + 1 and lane + 65) and their running sums by scans; the
    // list stops at the first count after which no more than 1e-12 of the time is left, that rest
    // lumped at its mean.
    const double mean = (double)line[AT_PASSING + p];
    const double none = exp(-mean), lead = -expm1(-mean), ln_mean = log(mean);
    const double ca = (double)(lane + 1), cb = (double)(lane + 65);
    const double share_a = exp(-mean + ca * ln_mean - T.ln_factorial[lane + 1]);
    const double share_b = exp(-mean + cb * ln_mean - T.ln_factorial[lane + 65]);
    const double sa = wave_scan(share_a, lane), ma = wave_scan(share_a * ca / mean, lane);
    const double sb = read_lane_d(sa, 63) + wave_scan(share_b, lane);
    const double mb = read_lane_d(ma, 63) + wave_scan(share_b * cb / mean, lane);
    const unsigned long long stop_a = __ballot(lead - sa <= 1e-12), stop_b = __ballot(lead - sb <= 1e-12);
    int last = stop_a ? __ffsll((long long)stop_a) : (stop_b ? 64 + __ffsll((long long)stop_b) : SHAPE_CAP - 1);
    last = min(last, SHAPE_CAP - 1);  // the lumped rest's index; counts 1 .. last - 1 before it
    if (lane + 1 < last) {
        out.v[lane + 1] = (float)(ca / mean);
        out.q[lane + 2] = none + sa;
        out.w[lane + 2] = ma;
    }
    if (lane + 65 < last) {
        out.v[lane + 65] = (float)(cb / mean);
        out.q[lane + 66] = none + sb;
        out.w[lane + 66] = mb;
    }
    const int before = last - 1;  // counts summed before the rest
    const double s_before = before == 0 ? 0.0 : (before <= 64 ? read_lane_d(sa, before - 1) : read_lane_d(sb, before - 65));
    const double m_before = before == 0 ? 0.0 : (before <= 64 ? read_lane_d(ma, before - 1) : read_lane_d(mb, before - 65));
    if (lane == 0) {
        const double time_left = lead - s_before, mean_left = 1.0 - m_before;
        out.v[0] = 0.0f;
        out.q[0] = out.w[0] = 0.0;
        out.q[1] = none;
        out.w[1] = 0.0;
        out.v[last] = (float)(mean_left / time_left);
        out.q[last + 1] = none + s_before + time_left;
        out.w[last + 1] = m_before + mean_left;
        out.n = last + 1;
        out.zeros = 1;
    }
}

// The bin of a sum x: its level rounded to the grid from -30 dB, clamped to the grid (runs' bin):
// (10 lg x + 30) / 2 from the hardware log2.
__device__ __forceinline__ int bin_of(float x) {
    const float half_steps = __log2f(x) * 1.50514997831990598f + 15.0f;
    return (int)fminf(fmaxf(floorf(half_steps + 0.5f), 0.0f), (float)(BINS - 1));
}

// Distribution::with_line: every occupied bin (and silence) of `d` with every value x scale of the
// list. For one source bin the lanes take the list's values (value j on lane j % 64), each finds
// the bin of its sum; the sums rise with j, so equal bins form runs, found by ballots of where the
// bin changes. The first lane of a run adds the run's share and energy, from the prefix sums at
// its two ends, into the run's bin of the wave's accumulator in LDS: one writer a bin per source,
// sources in bin order. Silence over silence stays silent.
template <int CAP>
__device__ void add_values(Dist& d, const Values<CAP>& list, float scale, float (*acc)[2], float2* sources,
                           int lane) {
    static_assert(CAP <= 128, "two values a lane");
    const int n = list.n, zeros = list.zeros;
    const bool ok_a = lane < n, ok_b = lane + 64 < n;
    const float va = ok_a ? list.v[lane] : 0.0f, vb = ok_b ? list.v[lane + 64] : 0.0f;
    const double qa = list.q[lane], wa = list.w[lane];
    const double qb = ok_b ? list.q[lane + 64] : 0.0, wb = ok_b ? list.w[lane + 64] : 0.0;
    // The sources (intensity, share): silence first, then the occupied bins in order.
    const int lead = d.silent > 0.0f ? 1 : 0;
    const unsigned long long m0 = __ballot(d.s0 > 0.0f), m1 = __ballot(d.s1 > 0.0f);
    if (d.s0 > 0.0f) sources[lead + __popcll(m0 & lanes_below(lane))] = make_float2(d.w0 / d.s0, d.s0);
    if (d.s1 > 0.0f) sources[lead + __popcll(m0) + __popcll(m1 & lanes_below(lane))] = make_float2(d.w1 / d.s1, d.s1);
    if (lead && lane == 0) sources[0] = make_float2(0.0f, d.silent);
    const int count = lead + __popcll(m0) + __popcll(m1);
    acc[lane][0] = acc[lane][1] = 0.0f;
    if (lane + 64 < BINS) acc[lane + 64][0] = acc[lane + 64][1] = 0.0f;
    wave_sync();
    double silent = 0.0;
    // One source: the bins of its sums with this lane's values, the runs of equal bins, and each
    // run's share and energy added by its first lane.
    auto source = [&](float intensity, float share) {
        const bool quiet = !(intensity > 0.0f);
        const bool live_a = ok_a && !(quiet && lane < zeros);
        const bool live_b = ok_b && !(quiet && lane + 64 < zeros);
        const int ta = live_a ? bin_of(intensity + scale * va) : -1;
        const int tb = live_b ? bin_of(intensity + scale * vb) : -1;
        const int pa = from_lane_below(ta, -1), pb = from_lane_below(tb, read_lane_i(ta, 63));
        const bool start_a = live_a && ta != pa, start_b = live_b && tb != pb;
        const unsigned long long runs_a = __ballot(start_a), runs_b = __ballot(start_b);
        const unsigned long long after = lane == 63 ? 0ull : ~0ull << (lane + 1);
        if (start_a) {
            const unsigned long long next = runs_a & after;
            const int end = next ? __ffsll((long long)next) - 1 : (runs_b ? 64 + __ffsll((long long)runs_b) - 1 : n);
            const float dq = (float)(list.q[end] - qa), dw = (float)(list.w[end] - wa);
            atomicAdd(&acc[ta][0], share * dq);
            atomicAdd(&acc[ta][1], share * (intensity * dq + scale * dw));
        }
        if (start_b) {
            const unsigned long long next = runs_b & after;
            const int end = next ? 64 + __ffsll((long long)next) - 1 : n;
            const float dq = (float)(list.q[end] - qb), dw = (float)(list.w[end] - wb);
            atomicAdd(&acc[tb][0], share * dq);
            atomicAdd(&acc[tb][1], share * (intensity * dq + scale * dw));
        }
        if (quiet) silent += (double)share * list.q[zeros];
    };
    if (n > 64) {
        for (int k = 0; k < count; ++k) {
            const float2 src = sources[k];
            source(src.x, src.y);
        }
    } else if (n > 0) {
        // A short list: 64 / n sources a pass, source k + lane / n on the lanes of its segment,
        // value lane % n on each. A run ends at the next run's first lane or at its segment's end.
        const int per = 64 / n, seg = lane / n, at = lane - seg * n;
        const bool mine = seg < per;
        const float v = mine ? list.v[at] : 0.0f;
        const double q0 = list.q[at], w0 = list.w[at];
        const unsigned long long edges = __ballot(mine && at == 0 && seg > 0) | (per * n < 64 ? 1ull << (per * n) : 0ull);
        const unsigned long long after = lane == 63 ? 0ull : ~0ull << (lane + 1);
        if (lead) silent += (double)d.silent * list.q[zeros];
        for (int k = 0; k < count; k += per) {
            const bool has = mine && k + seg < count;
            const float2 src = has ? sources[k + seg] : make_float2(0.0f, 0.0f);
            const bool live = has && !(!(src.x > 0.0f) && at < zeros);
            const int t = live ? bin_of(src.x + scale * v) : -1;
            const int before = from_lane_below(t, -1);
            const bool start = live && (t != before || at == 0);
            const unsigned long long next = (__ballot(start) | edges) & after;
            if (start) {
                const int end = (next ? __ffsll((long long)next) - 1 : 64) - seg * n;
                const float dq = (float)(list.q[end] - q0), dw = (float)(list.w[end] - w0);
                atomicAdd(&acc[t][0], src.y * dq);
                atomicAdd(&acc[t][1], src.y * (src.x * dq + scale * dw));
            }
        }
    }
    wave_sync();
    d.s0 = acc[lane][0];
    d.w0 = acc[lane][1];
    d.s1 = lane + 64 < BINS ? acc[lane + 64][0] : 0.0f;
    d.w1 = lane + 64 < BINS ? acc[lane + 64][1] : 0.0f;
    d.silent = (float)silent;
    wave_sync();
}

// Distribution::with_line for a list of few values (an event's silence and loud moments): the
// lanes take the source bins, each adding its sums, value by value, into the accumulator's bins
// (LDS atomics).
template <int CAP>
__device__ void add_few(Dist& d, const Values<CAP>& list, float scale, float (*acc)[2], int lane) {
    const int n = list.n, zeros = list.zeros;
    acc[lane][0] = acc[lane][1] = 0.0f;
    if (lane + 64 < BINS) acc[lane + 64][0] = acc[lane + 64][1] = 0.0f;
    wave_sync();
    const float i0 = d.s0 > 0.0f ? d.w0 / d.s0 : 0.0f, i1 = d.s1 > 0.0f ? d.w1 / d.s1 : 0.0f;
    double silent = 0.0;
    const double sc = (double)scale;
    for (int j = 0; j < n; ++j) {
        const float v = list.v[j];
        const double share = list.q[j + 1] - list.q[j], weighted = list.w[j + 1] - list.w[j];
        if (d.s0 > 0.0f) {
            const int t = bin_of(i0 + scale * v);
            atomicAdd(&acc[t][0], (float)((double)d.s0 * share));
            atomicAdd(&acc[t][1], (float)((double)d.s0 * ((double)i0 * share + sc * weighted)));
        }
        if (d.s1 > 0.0f) {
            const int t = bin_of(i1 + scale * v);
            atomicAdd(&acc[t][0], (float)((double)d.s1 * share));
            atomicAdd(&acc[t][1], (float)((double)d.s1 * ((double)i1 * share + sc * weighted)));
        }
        if (d.silent > 0.0f) {
            if (j < zeros) {
                silent += (double)d.silent * share;
            } else if (lane == 0) {
                const int t = bin_of(scale * v);
                atomicAdd(&acc[t][0], (float)((double)d.silent * share));
                atomicAdd(&acc[t][1], (float)((double)d.silent * sc * weighted));
            }
        }
    }
    wave_sync();
    d.s0 = acc[lane][0];
    d.w0 = acc[lane][1];
    d.s1 = lane + 64 < BINS ? acc[lane + 64][0] : 0.0f;
    d.w1 = lane + 64 < BINS ? acc[lane + 64][1] : 0.0f;
    d.silent = (float)silent;
    wave_sync();
}

// Distribution::over on silence: one bin at `x` (by its rounded level, as runs starts), or
// silence when there is nothing.
__device__ void set_over(Dist& d, double x, int lane) {
    d.s0 = d.w0 = d.s1 = d.w1 = 0.0f;
    if (!(x > 0.0)) {
        d.silent = 1.0f;
        return;
    }
    d.silent = 0.0f;
    const double r = round((10.0 * log10(x) - LEVEL_MIN_DB) / BIN_DB);
    const int b = r > 0.0 ? (r >= (double)(BINS - 1) ? BINS - 1 : (int)r) : 0;
    if (b == lane) {
        d.s0 = 1.0f;
        d.w0 = (float)x;
    }
    if (b == lane + 64) {
        d.s1 = 1.0f;
        d.w1 = (float)x;
    }
}

// A distribution's occupied bins (and silence, first) as an ascending list of values with their
// prefix sums (Distribution::with_distribution's LineValues), compacted by ballot prefix counts.
__device__ void list_of(const Dist& d, Values<LIST_CAP>& out, int lane) {
    const int base = d.silent > 0.0f ? 1 : 0;
    const unsigned long long m0 = __ballot(d.s0 > 0.0f), m1 = __ballot(d.s1 > 0.0f);
    const int e0 = base + __popcll(m0 & lanes_below(lane));
    const int e1 = base + __popcll(m0) + __popcll(m1 & lanes_below(lane));
    const double i0 = wave_scan(d.s0 > 0.0f ? (double)d.s0 : 0.0, lane);
    const double j0 = wave_scan(d.s0 > 0.0f ? (double)d.w0 : 0.0, lane);
    const double i1 = wave_scan(d.s1 > 0.0f ? (double)d.s1 : 0.0, lane);
    const double j1 = wave_scan(d.s1 > 0.0f ? (double)d.w1 : 0.0, lane);
    const double lead = base ? (double)d.silent : 0.0;
    const double half_q = lead + read_lane_d(i0, 63), half_w = read_lane_d(j0, 63);
    if (d.s0 > 0.0f) {
        out.v[e0] = d.w0 / d.s0;
        out.q[e0 + 1] = lead + i0;
        out.w[e0 + 1] = j0;
    }
    if (d.s1 > 0.0f) {
        out.v[e1] = d.w1 / d.s1;
        out.q[e1 + 1] = half_q + i1;
        out.w[e1 + 1] = half_w + j1;
    }
    if (lane == 0) {
        out.q[0] = out.w[0] = 0.0;
        if (base) {
            out.v[0] = 0.0f;
            out.q[1] = lead;
            out.w[1] = 0.0;
        }
        out.n = base + __popcll(m0) + __popcll(m1);
        out.zeros = base;
    }
}

// A state's distribution into a curve's mixture, `weight` of the time (LDS atomics: the states'
// waves add into one mixture).
__device__ __forceinline__ void accumulate(float (*mix)[2], const Dist& d, float weight, int lane) {
    if (d.s0 > 0.0f) {
        atomicAdd(&mix[lane][0], weight * d.s0);
        atomicAdd(&mix[lane][1], weight * d.w0);
    }
    if (d.s1 > 0.0f) {
        atomicAdd(&mix[lane + 64][0], weight * d.s1);
        atomicAdd(&mix[lane + 64][1], weight * d.w1);
    }
}

// ISO 532-1 loudness of a curve's spectrum raised by `offset_db` (a call of its own, so that the
// painter's state and the method's working set are not live together).
__device__ __noinline__ float curve_step_sone(const float* spectrum, float offset_db) {
    float levels[THIRDS];
    for (int t = 0; t < THIRDS; ++t) levels[t] = spectrum[t] + offset_db;
    return zwicker_sone(levels);
}

// The lines of a period and hour group that move some weather state, in layer order and, within a
// layer, moment()'s order: events and streams of passes, then the Kurze lines whose loud moments
// pass the held test (percentiles.rs layer_states). Every wave walks them alike.
struct Step {
    int layer;
    const float* line;
    Kind kind;
    double factor;
    unsigned movers;  // the distinct states the line moves
};

struct StepWalk {
    int layer = 0, pass = 0;
    uint32_t j = 0;
};

__device__ bool next_step(StepWalk& w, Step& out, const Shared& sh, const Tables& T, const float* in, int p, int g,
                          bool profiled, unsigned distinct, int lane) {
    const int state = min(lane, STATES - 1);
    for (; w.layer < LAYERS; ++w.layer, w.pass = 0, w.j = 0) {
        const float* lines = in + sh.layer_at[w.layer] + PERIODS + 1;
        const uint32_t n = sh.layer_lines[w.layer];
        for (; w.pass < 2; ++w.pass, w.j = 0)
            for (; w.j < n; ++w.j) {
                const float* line = lines + w.j * LINE_FLOATS;
                const double factor = factor_of(T, line, p, g, profiled);
                const Kind kind = kind_of(line, p, factor);
                if (w.pass == 0 ? (kind != PASS_MOVES && kind != DUTY_MOVES) : kind != KURZE_SHAPED) continue;
                if (line[p] <= 0.0f) continue;
                const double mean = (double)line[AT_MEANS + p * STATES + state] * factor;
                bool moves = !(mean <= 0.0);
                if (w.pass == 1 && moves) {
                    const double exceeded =
                        (double)exceeded_intensity(T, (float)((double)line[AT_LAMBDA + p] * factor));
                    moves = !(mean * exceeded < HELD_SHARE * sh.pre[w.layer][state].compare);
                }
                const unsigned movers = (unsigned)__ballot(lane < STATES && moves) & distinct;
                if (!movers) continue;
                out.layer = w.layer;
                out.line = line;
                out.kind = kind;
                out.factor = factor;
                out.movers = movers;
                ++w.j;
                return true;
            }
    }
    return false;
}

#ifndef QM_PIXELS_WAVES_PER_EU
#define QM_PIXELS_WAVES_PER_EU 3
#endif
// One workgroup per pixel: a wave per weather state (wave s keeps state s's distributions) and a
// builder wave that makes each moving line's shape one line ahead of the state waves, the two
// shape buffers alternating, one barrier a line.
__global__ __launch_bounds__(GROUP) __attribute__((amdgpu_waves_per_eu(QM_PIXELS_WAVES_PER_EU))) void pixel_loudness_kernel(
    const float* __restrict__ input, const uint64_t* __restrict__ offsets, const Tables* __restrict__ tables,
    float* mixtures, float* ndens) {
    __shared__ Shared sh;
    const Tables& T = *tables;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool builder = wave == BUILDER;
    const uint32_t pixel = blockIdx.x;
    const float* in = input + offsets[pixel];

    // The pixel's layers: steady energy per period, then its lines; the curves after the last.
    bool profiled = false;
    uint32_t curves_at = 0;
    {
        uint32_t at = 0;
        for (int l = 0; l < LAYERS; ++l) {
            const uint32_t n = (uint32_t)in[at + PERIODS];
            if (tid == 0) {
                sh.layer_at[l] = at;
                sh.layer_lines[l] = n;
            }
            for (uint32_t j = 0; j < n; ++j)
                profiled = profiled || in[at + PERIODS + 1 + j * LINE_FLOATS + AT_PROFILE] >= 0.0f;
            at += PERIODS + 1 + n * LINE_FLOATS;
        }
        curves_at = at;
    }
    if (tid <= CURVES) sh.nden[tid] = 0.0f;
    if (wave < CURVES)
        for (int k = lane; k < STEP_SLOTS; k += 64) sh.need[wave][k] = 0;
    const int state = min(lane, STATES - 1);  // the state of a lane in the per-state decisions
    __syncthreads();

    for (int p = 0; p < PERIODS; ++p) {
        // States equal to the one before (every line's mean the same) are that state again: the
        // wave of the first of a run of equal states counts for the run.
        bool differs = false;
        for (int l = 0; l < LAYERS; ++l)
            for (uint32_t j = 0; j < sh.layer_lines[l]; ++j) {
                const float* means = in + sh.layer_at[l] + PERIODS + 1 + j * LINE_FLOATS + AT_MEANS + p * STATES;
                differs = differs || (state > 0 && means[state] != means[state - 1]);
            }
        const unsigned distinct = ((unsigned)__ballot(lane < STATES && differs) & ((1u << STATES) - 1u)) | 1u;
        const bool active = wave < STATES && (distinct >> wave & 1u);
        int copies = 1;
        while (wave + copies < STATES && !(distinct >> (wave + copies) & 1u)) ++copies;
        for (int k = tid; k < CURVES * BINS * 2; k += GROUP) (&sh.mix[0][0][0])[k] = 0.0f;

        const int groups = profiled ? T.groups[p] : 1;
        for (int g = 0; g < groups; ++g) {
            const float weight = (profiled ? T.group_weight[p][g] : 1.0f / STATES) * (float)copies;
            // Wave l for layer l, a lane a state: the floor, the energy the held test compares
            // against, and whether any line moves (percentiles.rs moment / layer_states).
            if (wave < LAYERS) {
                const int l = wave;
                double floor = (double)in[sh.layer_at[l] + p], held = 0.0;
                int moving = 0;
                const float* lines = in + sh.layer_at[l] + PERIODS + 1;
                const uint32_t n = sh.layer_lines[l];
                for (uint32_t j = 0; j < n; ++j) {
                    const float* line = lines + j * LINE_FLOATS;
                    const double factor = factor_of(T, line, p, g, profiled);
                    const double mean = (double)line[AT_MEANS + p * STATES + state] * factor;
                    if (line[p] <= 0.0f || mean <= 0.0) continue;
                    switch (kind_of(line, p, factor)) {
                    case PASS_MOVES:
                    case DUTY_MOVES: ++moving; break;
                    case PASS_HOLDS:
                    case KURZE_HOLDS: floor += mean; break;
                    case KURZE_SHAPED: held += mean; break;
                    case DUTY_SILENT: break;
                    }
                }
                const double compare = floor + held;
                for (uint32_t j = 0; j < n; ++j) {
                    const float* line = lines + j * LINE_FLOATS;
                    const double factor = factor_of(T, line, p, g, profiled);
                    const double mean = (double)line[AT_MEANS + p * STATES + state] * factor;
                    if (line[p] <= 0.0f || mean <= 0.0 || kind_of(line, p, factor) != KURZE_SHAPED) continue;
                    const double exceeded =
                        (double)exceeded_intensity(T, (float)((double)line[AT_LAMBDA + p] * factor));
                    if (mean * exceeded < HELD_SHARE * compare) floor += mean;
                    else ++moving;
                }
                if (lane < STATES) {
                    sh.pre[l][lane].floor = floor;
                    sh.pre[l][lane].compare = compare;
                    sh.moves[l][lane] = moving != 0;
                }
            }
            __syncthreads();
            StepWalk walk;
            Step step;
            if (builder) {
                // Shape k into buffer k % 2 before barrier k; the state waves apply it after.
                int k = 0;
                while (next_step(walk, step, sh, T, in, p, g, profiled, distinct, lane)) {
                    build_shape(T, step.line, step.kind, p, step.factor, sh.shape[k & 1], lane);
                    __syncthreads();
                    ++k;
                }
            } else {
                unsigned steady = 0;
                if (wave < STATES)
                    for (int l = 0; l < LAYERS; ++l)
                        if (!sh.moves[l][wave]) steady |= 1u << l;
                // The place starts from the steady layers' floors.
                Dist place, layer;
                if (active) {
                    double base = 0.0;
                    for (int l = 0; l < LAYERS; ++l)
                        if (steady >> l & 1u) base += sh.pre[l][wave].floor;
                    set_over(place, base, lane);
                }
                bool pending = next_step(walk, step, sh, T, in, p, g, profiled, distinct, lane);
                int k = 0;
                for (int l = 0; l < LAYERS; ++l) {
                    if (active) set_over(layer, sh.pre[l][wave].floor, lane);
                    for (; pending && step.layer == l; pending = next_step(walk, step, sh, T, in, p, g, profiled,
                                                                           distinct, lane)) {
                        __syncthreads();
                        const Values<SHAPE_CAP>& shape = sh.shape[k & 1];
                        ++k;
                        if (wave < STATES && (step.movers >> wave & 1u)) {
                            const float scale = (float)((double)step.line[AT_MEANS + p * STATES + wave] * step.factor);
                            if (shape.n <= 4) add_few(layer, shape, scale, sh.acc[wave], lane);
                            else add_values(layer, shape, scale, sh.acc[wave], sh.sources[wave], lane);
                        }
                    }
                    if (active) {
                        accumulate(sh.mix[l], layer, weight, lane);
                        if (!(steady >> l & 1u)) {
                            list_of(layer, sh.list[wave], lane);
                            wave_sync();
                            add_values(place, sh.list[wave], 1.0f, sh.acc[wave], sh.sources[wave], lane);
                        }
                    }
                }
                if (active) accumulate(sh.mix[LAYERS], place, weight, lane);
            }
            __syncthreads();  // sh.pre and sh.moves are the next group's
        }

        if (mixtures)
            for (int k = tid; k < CURVES * BINS * 2; k += GROUP)
                mixtures[((size_t)pixel * PERIODS + p) * (CURVES * BINS * 2) + k] = (&sh.mix[0][0][0])[k];

        // Nden's period term per curve (Curve::at over the mixture's bins), a wave a curve.
        if (wave < CURVES) {
            const int c = wave;
            const float* spectrum = in + curves_at + (c * PERIODS + p) * THIRDS;
            if (!isnan(spectrum[0])) {
                float share[2], position[2];
                int below[2];
                for (int h = 0; h < 2; ++h) {
                    const int b = lane + 64 * h;
                    share[h] = b < BINS ? sh.mix[c][b][0] : 0.0f;
                    position[h] = 0.0f;
                    below[h] = 0;
                    if (share[h] > 0.0f) {
                        const float level = 10.0f * log10f(sh.mix[c][b][1] / share[h]) + PERIOD_PENALTY_F[p];
                        position[h] = fminf(fmaxf(level - (float)LEVEL_MIN_DB, 0.0f), (float)(CURVE_STEPS - 1));
                        below[h] = min((int)floorf(position[h]), CURVE_STEPS - 2);
                        sh.need[c][below[h]] = 1;
                        sh.need[c][below[h] + 1] = 1;
                    }
                }
                wave_sync();
                int count = 0;
                for (int r = 0; r < STEP_SLOTS / 64; ++r) {
                    const int step = r * 64 + lane;
                    const bool needed = sh.need[c][step] != 0;
                    const unsigned long long m = __ballot(needed);
                    if (needed) sh.steps[c][count + __popcll(m & lanes_below(lane))] = (unsigned short)step;
                    count += __popcll(m);
                }
                wave_sync();
                for (int k = lane; k < count; k += 64) {
                    const int step = sh.steps[c][k];
                    sh.sone[c][step] = curve_step_sone(spectrum, (float)LEVEL_MIN_DB + (float)step);
                }
                wave_sync();
                float sum = 0.0f;
                for (int h = 0; h < 2; ++h)
                    if (share[h] > 0.0f) {
                        const float lo = sh.sone[c][below[h]], hi = sh.sone[c][below[h] + 1];
                        sum += share[h] * (lo + (hi - lo) * (position[h] - (float)below[h]));
                    }
                sum = wave_sum(sum);
                if (lane == 0) sh.nden[c] += PERIOD_HOURS_F[p] * sum;
                for (int k = lane; k < STEP_SLOTS; k += 64) sh.need[c][k] = 0;
                wave_sync();
            }
        }
        __syncthreads();
    }
    if (tid < CURVES) ndens[(size_t)pixel * CURVES + tid] = sh.nden[tid] / 24.0f;
}

// =============================================================================================
// The map's folding (paint/src/sone_fold.rs fold_placed_for_map, popup/src/percentiles/weather.rs
// merge_close_states, paint/src/sone.rs write_packed): unfolded pixels (the packed layout, every
// value f64) into the card's packed f32 pixels at the same offsets. A work-item per pixel: the
// folding is a sequential scan of its lines, cheap beside the distributions; its per-layer sums
// live in the work-item's private memory, its lines' keep flags and duty ratios in global scratch
// (one entry per input line), the folded lines in f64 in `work` (the pixel's own range).

constexpr double DUTY_FOLD_PEAK_SHARE = 0.12;
constexpr double DUTY_FOLD_STD_SHARE = 0.08;
constexpr double KURZE_FOLD_HELD_SHARE = 0.05;
constexpr double FOLD_SPREAD = 1.4125375446227544;  // 10^(1.5 / 10)
constexpr int AT_ENERGY = 0;

__device__ __forceinline__ bool has_duty_of(const double* line) { return !isnan(line[AT_DUTY]); }
__device__ __forceinline__ bool has_passing_of(const double* line) { return !isnan(line[AT_PASSING]); }
__device__ __forceinline__ int profile_of(const double* line) { return (int)line[AT_PROFILE]; }

// Kurze's relative intensity at the held test's probability (1 - 1e-3, above the table's last
// quantile) in FP64, for lambda below the simulated maximum: the top's intensity between the rows
// about lambda (scaled down with lambda below the table), plus the nearest emitter's passing there
// less its passing at the top (physics/src/percentile.rs relative_intensity).
__device__ double exceeded_d(const Tables& T, double lambda) {
    if (!isfinite(lambda)) return 1.0;
    if (lambda <= 0.0) return 0.0;
    const double at = fmax(lambda, LAMBDA_MIN);
    const double position = fmin(fmax(log10(at / LAMBDA_MIN) * STEPS_PER_DECADE, 0.0), (double)(ROWS - 1));
    const int row = (int)position;
    const double fraction = position - trunc(position);
    const double lo = T.fold_top[row], hi = T.fold_top[min(row + 1, ROWS - 1)];
    double top = exp(lo + fraction * (hi - lo));
    if (lambda < LAMBDA_MIN) top = lambda / LAMBDA_MIN * top;
    auto passing = [&](double ln_p) {
        const double s = -ln_p / (2.0 * lambda);
        return 1.0 / (lambda * PI_D * (1.0 + s * s));
    };
    return top + passing(T.fold_ln_exceed) - passing(T.fold_ln_top);
}

// Line::steady into `out`: lambda infinite, no profile unless given, no duty or passes.
__device__ void write_steady_line(double* out, const double (*means)[STATES], int profile) {
    for (int p = 0; p < PERIODS; ++p) {
        bool any = false;
        for (int k = 0; k < STATES; ++k) {
            out[AT_MEANS + p * STATES + k] = means[p][k];
            any = any || means[p][k] > 0.0;
        }
        out[AT_ENERGY + p] = any ? 1.0 : 0.0;
        out[AT_LAMBDA + p] = INFINITY;
        out[AT_DUTY + p] = NAN;
        out[AT_PASSING + p] = NAN;
    }
    out[AT_PROFILE] = (double)profile;
}

__global__ __launch_bounds__(64) void fold_kernel(const double* __restrict__ input, const uint64_t* __restrict__ offsets,
                                                  const uint64_t* __restrict__ line_first, uint32_t pixels,
                                                  const Tables* __restrict__ tables, double* work, uint8_t* keep,
                                                  double* ratio, float* out) {
    const uint32_t pixel = blockIdx.x * blockDim.x + threadIdx.x;
    if (pixel >= pixels) return;
    const Tables& T = *tables;
    const double* in = input + offsets[pixel];
    double* w = work + offsets[pixel];
    uint8_t* kept = keep + line_first[pixel];
    double* rat = ratio + line_first[pixel];
    bool any_profiled = false;
    {
        size_t at = 0;
        for (int l = 0; l < LAYERS; ++l) {
            const uint32_t n = (uint32_t)in[at + PERIODS];
            for (uint32_t j = 0; j < n; ++j)
                any_profiled = any_profiled || profile_of(in + at + PERIODS + 1 + (size_t)j * LINE_FLOATS) >= 0;
            at += PERIODS + 1 + (size_t)n * LINE_FLOATS;
        }
    }
    size_t at = 0, wat = 0, line_base = 0;
    size_t layer_w[LAYERS];
    for (int l = 0; l < LAYERS; ++l) {
        const double* steady = in + at;
        const uint32_t n = (uint32_t)in[at + PERIODS];
        const double* lines = in + at + PERIODS + 1;
        layer_w[l] = wat;
        for (int p = 0; p < PERIODS; ++p) w[wat + p] = steady[p];
        double* count_at = w + wat + PERIODS;
        wat += PERIODS + 1;
        if (n == 0) {
            *count_at = 0.0;
            at += PERIODS + 1;
            continue;
        }
        bool has_kurze = false, has_duty = false;
        for (uint32_t j = 0; j < n; ++j) {
            const double* line = lines + (size_t)j * LINE_FLOATS;
            if (has_duty_of(line)) has_duty = true;
            else if (!has_passing_of(line))
                for (int p = 0; p < PERIODS; ++p) has_kurze = has_kurze || isfinite(line[AT_LAMBDA + p]);
        }
        double held[PERIODS][MAX_GROUPS][STATES];
        if (has_kurze) {
            for (int p = 0; p < PERIODS; ++p)
                for (int g = 0; g < MAX_GROUPS; ++g)
                    for (int k = 0; k < STATES; ++k) held[p][g][k] = steady[p];
            for (uint32_t j = 0; j < n; ++j) {
                const double* line = lines + (size_t)j * LINE_FLOATS;
                if (has_duty_of(line) || has_passing_of(line)) continue;
                for (int p = 0; p < PERIODS; ++p) {
                    if (line[AT_ENERGY + p] <= 0.0) continue;
                    const int ng = any_profiled ? T.groups[p] : 1;
                    for (int g = 0; g < ng; ++g) {
                        const int profile = profile_of(line);
                        const double factor = any_profiled && profile >= 0 ? T.factor[profile][p][g] : 1.0;
                        for (int k = 0; k < STATES; ++k) {
                            const double m = line[AT_MEANS + p * STATES + k] * factor;
                            if (m > 0.0) held[p][g][k] += m;
                        }
                    }
                }
            }
        }
        double total[PERIODS][STATES];
        if (has_duty) {
            for (int p = 0; p < PERIODS; ++p)
                for (int k = 0; k < STATES; ++k) total[p][k] = steady[p];
            for (uint32_t j = 0; j < n; ++j) {
                const double* line = lines + (size_t)j * LINE_FLOATS;
                for (int p = 0; p < PERIODS; ++p) {
                    if (line[AT_ENERGY + p] <= 0.0) continue;
                    for (int k = 0; k < STATES; ++k)
                        if (line[AT_MEANS + p * STATES + k] > 0.0) total[p][k] += line[AT_MEANS + p * STATES + k];
                }
            }
        }
        double unprof[PERIODS][STATES], prof[PROFILES][PERIODS][STATES];
        for (int p = 0; p < PERIODS; ++p)
            for (int k = 0; k < STATES; ++k) {
                unprof[p][k] = 0.0;
                for (int q = 0; q < PROFILES; ++q) prof[q][p][k] = 0.0;
            }
        bool has_unprof = false, has_prof[PROFILES] = {false, false, false};
        uint32_t candidates = 0;
        for (uint32_t j = 0; j < n; ++j) {
            const double* line = lines + (size_t)j * LINE_FLOATS;
            kept[line_base + j] = 0;
            rat[line_base + j] = NAN;
            if (has_duty_of(line)) {
                bool moves = false, eligible = true;
                double max_ratio = 0.0;
                for (int p = 0; p < PERIODS; ++p) {
                    const double d = line[AT_DUTY + p];
                    if (line[AT_ENERGY + p] <= 0.0 || d <= 0.0) continue;
                    const double var_scale = fmax(1.0 - d, 0.0) / d;
                    for (int k = 0; k < STATES; ++k) {
                        const double mean = line[AT_MEANS + p * STATES + k];
                        if (mean <= 0.0) continue;
                        moves = true;
                        const double t = total[p][k];
                        if (t <= 0.0 || mean / d > DUTY_FOLD_PEAK_SHARE * t) {
                            eligible = false;
                        } else {
                            const double r = mean / t;
                            max_ratio = fmax(max_ratio, r * r * var_scale);
                        }
                    }
                }
                if (moves) {
                    if (eligible) {
                        rat[line_base + j] = max_ratio;
                        ++candidates;
                    } else {
                        kept[line_base + j] = 1;
                    }
                }
                continue;
            }
            if (has_passing_of(line)) {
                bool moves = false;
                for (int p = 0; p < PERIODS; ++p) {
                    bool any = false;
                    for (int k = 0; k < STATES; ++k) any = any || line[AT_MEANS + p * STATES + k] > 0.0;
                    const double passing = line[AT_PASSING + p];
                    moves = moves || (line[AT_ENERGY + p] > 0.0 && passing > 0.0 && passing < PASSES_HOLD_MEAN && any);
                }
                if (moves) {
                    kept[line_base + j] = 1;
                } else {
                    for (int p = 0; p < PERIODS; ++p) {
                        if (line[AT_ENERGY + p] <= 0.0) continue;
                        for (int k = 0; k < STATES; ++k)
                            if (line[AT_MEANS + p * STATES + k] > 0.0) {
                                unprof[p][k] += line[AT_MEANS + p * STATES + k];
                                has_unprof = true;
                            }
                    }
                }
                continue;
            }
            const int profile = profile_of(line);
            bool always_held = true;
            for (int p = 0; p < PERIODS && always_held; ++p) {
                if (line[AT_ENERGY + p] <= 0.0) continue;
                const int ng = any_profiled ? T.groups[p] : 1;
                for (int g = 0; g < ng && always_held; ++g) {
                    const double factor = any_profiled && profile >= 0 ? T.factor[profile][p][g] : 1.0;
                    const double lambda = line[AT_LAMBDA + p] * factor;
                    if (lambda >= HOLDS_MEAN_LAMBDA) continue;
                    const double exc = exceeded_d(T, lambda);
                    for (int k = 0; k < STATES; ++k) {
                        const double m = line[AT_MEANS + p * STATES + k] * factor;
                        if (m > 0.0 && (m * exc >= KURZE_FOLD_HELD_SHARE * held[p][g][k] || !isfinite(m * exc))) {
                            always_held = false;
                            break;
                        }
                    }
                }
            }
            if (always_held) {
                for (int p = 0; p < PERIODS; ++p) {
                    if (line[AT_ENERGY + p] <= 0.0) continue;
                    double* target;
                    if (profile >= 0) {
                        has_prof[profile] = true;
                        target = prof[profile][p];
                    } else {
                        has_unprof = true;
                        target = unprof[p];
                    }
                    for (int k = 0; k < STATES; ++k)
                        if (line[AT_MEANS + p * STATES + k] > 0.0) target[k] += line[AT_MEANS + p * STATES + k];
                }
            } else {
                kept[line_base + j] = 1;
            }
        }
        // The events eligible to fold, in order of their variance ratio (then their index), while
        // the folded variance stays within the budget.
        if (candidates > 0) {
            double folded_var[PERIODS][STATES];
            for (int p = 0; p < PERIODS; ++p)
                for (int k = 0; k < STATES; ++k) folded_var[p][k] = 0.0;
            const double budget_sq = DUTY_FOLD_STD_SHARE * DUTY_FOLD_STD_SHARE;
            double last_ratio = -1.0;
            int64_t last_index = -1;
            for (uint32_t c = 0; c < candidates; ++c) {
                int64_t pick = -1;
                double pick_ratio = 0.0;
                for (uint32_t j = 0; j < n; ++j) {
                    const double r = rat[line_base + j];
                    if (isnan(r)) continue;
                    const bool after = r > last_ratio || (r == last_ratio && (int64_t)j > last_index);
                    const bool before = pick < 0 || r < pick_ratio || (r == pick_ratio && (int64_t)j < pick);
                    if (after && before) {
                        pick = j;
                        pick_ratio = r;
                    }
                }
                last_ratio = pick_ratio;
                last_index = pick;
                const double* line = lines + (size_t)pick * LINE_FLOATS;
                bool fits = true;
                for (int p = 0; p < PERIODS && fits; ++p) {
                    const double d = line[AT_DUTY + p];
                    if (line[AT_ENERGY + p] <= 0.0 || d <= 0.0) continue;
                    const double var_scale = fmax(1.0 - d, 0.0) / d;
                    for (int k = 0; k < STATES; ++k) {
                        const double m = line[AT_MEANS + p * STATES + k];
                        if (m <= 0.0) continue;
                        if (!(folded_var[p][k] + m * m * var_scale <= budget_sq * total[p][k] * total[p][k])) {
                            fits = false;
                            break;
                        }
                    }
                }
                if (fits) {
                    for (int p = 0; p < PERIODS; ++p) {
                        const double d = line[AT_DUTY + p];
                        if (line[AT_ENERGY + p] <= 0.0 || d <= 0.0) continue;
                        const double var_scale = fmax(1.0 - d, 0.0) / d;
                        for (int k = 0; k < STATES; ++k) {
                            const double m = line[AT_MEANS + p * STATES + k];
                            if (m > 0.0) {
                                folded_var[p][k] += m * m * var_scale;
                                unprof[p][k] += m;
                                has_unprof = true;
                            }
                        }
                    }
                } else {
                    kept[line_base + pick] = 1;
                }
            }
        }
        // The folded steady lines first (unprofiled, then by profile), then the kept lines in order.
        uint32_t count = 0;
        if (has_unprof) {
            write_steady_line(w + wat, unprof, -1);
            wat += LINE_FLOATS;
            ++count;
        }
        for (int q = 0; q < PROFILES; ++q)
            if (has_prof[q]) {
                write_steady_line(w + wat, prof[q], q);
                wat += LINE_FLOATS;
                ++count;
            }
        for (uint32_t j = 0; j < n; ++j)
            if (kept[line_base + j]) {
                const double* line = lines + (size_t)j * LINE_FLOATS;
                for (int k = 0; k < LINE_FLOATS; ++k) w[wat + k] = line[k];
                wat += LINE_FLOATS;
                ++count;
            }
        *count_at = (double)count;
        line_base += n;
        at += PERIODS + 1 + (size_t)n * LINE_FLOATS;
    }
    const size_t curves_in = at;
    // merge_close_states over every line of the pixel, period by period.
    for (int p = 0; p < PERIODS; ++p) {
        int start = 0;
        while (start < STATES) {
            int end = start + 1;
            for (; end < STATES; ++end) {
                bool all = true;
                for (int l = 0; l < LAYERS && all; ++l) {
                    const size_t base = layer_w[l];
                    const uint32_t n = (uint32_t)w[base + PERIODS];
                    for (uint32_t j = 0; j < n && all; ++j) {
                        const double* m = w + base + PERIODS + 1 + (size_t)j * LINE_FLOATS + AT_MEANS + p * STATES;
                        double low = INFINITY, high = 0.0;
                        for (int k = start; k <= end; ++k) {
                            if (!isfinite(m[k])) {
                                all = false;
                                break;
                            }
                            low = fmin(low, m[k]);
                            high = fmax(high, m[k]);
                        }
                        all = all && high <= FOLD_SPREAD * low;
                    }
                }
                if (!all) break;
            }
            if (end - start > 1)
                for (int l = 0; l < LAYERS; ++l) {
                    const size_t base = layer_w[l];
                    const uint32_t n = (uint32_t)w[base + PERIODS];
                    for (uint32_t j = 0; j < n; ++j) {
                        double* m = w + base + PERIODS + 1 + (size_t)j * LINE_FLOATS + AT_MEANS + p * STATES;
                        bool differs = false;
                        for (int k = start + 1; k < end; ++k) differs = differs || m[k] != m[start];
                        if (!differs) continue;
                        double sum = 0.0;
                        for (int k = start; k < end; ++k) sum += m[k];
                        const double mean = sum / (double)(end - start);
                        for (int k = start; k < end; ++k) m[k] = mean;
                    }
                }
            start = end;
        }
    }
    // Each layer's first line, if an unprofiled steady line whose states all read alike in FP32,
    // goes back into the layer's steady floor; then the packed f32 pixel.
    float* o = out + offsets[pixel];
    size_t oat = 0;
    for (int l = 0; l < LAYERS; ++l) {
        const size_t base = layer_w[l];
        double steady[PERIODS] = {w[base], w[base + 1], w[base + 2]};
        uint32_t n = (uint32_t)w[base + PERIODS];
        const double* first = w + base + PERIODS + 1;
        uint32_t skip = 0;
        if (n > 0 && profile_of(first) < 0 && !has_duty_of(first) && !has_passing_of(first)) {
            bool collapse = true;
            for (int p = 0; p < PERIODS; ++p) {
                collapse = collapse && !isfinite(first[AT_LAMBDA + p]);
                for (int k = 0; k < STATES; ++k)
                    collapse = collapse && (float)first[AT_MEANS + p * STATES + k] == (float)first[AT_MEANS + p * STATES];
            }
            if (collapse) {
                for (int p = 0; p < PERIODS; ++p) {
                    double sum = 0.0;
                    for (int k = 0; k < STATES; ++k) sum += first[AT_MEANS + p * STATES + k];
                    steady[p] += sum / (double)STATES;
                }
                skip = 1;
            }
        }
        for (int p = 0; p < PERIODS; ++p) o[oat + p] = (float)steady[p];
        o[oat + PERIODS] = (float)(n - skip);
        oat += PERIODS + 1;
        for (uint32_t j = skip; j < n; ++j) {
            const double* line = w + base + PERIODS + 1 + (size_t)j * LINE_FLOATS;
            for (int k = 0; k < LINE_FLOATS; ++k) o[oat + k] = (float)line[k];
            // An event's duty and a passing line's passes go whole or not at all (the first tells).
            for (int p = 0; p < PERIODS; ++p) {
                if (!has_duty_of(line)) o[oat + AT_DUTY + p] = NAN;
                if (!has_passing_of(line)) o[oat + AT_PASSING + p] = NAN;
            }
            oat += LINE_FLOATS;
        }
    }
    for (int k = 0; k < CURVES * PERIODS * THIRDS; ++k) o[oat + k] = (float)in[curves_in + k];
}

// =============================================================================================
// Host side.

int pfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

// Kurze's quantile table by the reference's seeded simulation (physics/src/percentile_tests.rs
// writes_the_table): SplitMix64 from 0x5eed_1971, a row per lambda from 1e-4 at eight a decade,
// each row's samples the intensity of a Poisson line over at least 200 distances either side
// (the rest of the line as its mean), sorted, read at the quantiles (k + 0.5) / 100, kept as FP32.
class Random {
  public:
    explicit Random(uint64_t seed) : state_(seed) {}
    double uniform() {
        state_ += 0x9e3779b97f4a7c15ull;
        uint64_t z = state_;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return (double)((z ^ (z >> 31)) >> 11) / 9007199254740992.0;
    }
    size_t poisson(double mean) {
        if (mean > 50.0) {
            const double radius = std::sqrt(-2.0 * std::log(1.0 - uniform()));
            const double normal = radius * std::cos(2.0 * PI_D * uniform());
            return (size_t)std::max(std::round(mean + std::sqrt(mean) * normal), 0.0);
        }
        const double stop = std::exp(-mean);
        size_t n = 0;
        double product = uniform();
        while (product > stop) {
            ++n;
            product *= uniform();
        }
        return n;
    }

  private:
    uint64_t state_;
};

const std::vector<float>& kurze_quantiles() {
    static std::vector<float> table;
    static std::once_flag once;
    std::call_once(once, [] {
        Random random(0x5eed1971ull);
        table.resize((size_t)ROWS * QUANTILES);
        std::vector<double> draws;
        for (int row = 0; row < ROWS; ++row) {
            const double lambda = LAMBDA_MIN * std::pow(10.0, (double)row / STEPS_PER_DECADE);
            const double half = std::max(5.0 / lambda, 200.0);
            const double inside_mean = 2.0 * lambda * half;
            const double beyond = 2.0 * lambda * (PI_D / 2.0 - std::atan(half));
            const double mean = lambda * PI_D;
            const size_t samples = (size_t)std::min(20000.0 * std::max(0.01 / lambda, 1.0), 1000000.0);
            draws.assign(samples, 0.0);
            for (double& sample : draws) {
                const size_t emitters = random.poisson(inside_mean);
                double near = 0.0;
                for (size_t e = 0; e < emitters; ++e) {
                    const double x = (2.0 * random.uniform() - 1.0) * half;
                    near += 1.0 / (1.0 + x * x);
                }
                sample = (near + beyond) / mean;
            }
            std::sort(draws.begin(), draws.end());
            for (int k = 0; k < QUANTILES; ++k)
                table[(size_t)row * QUANTILES + k] =
                    (float)draws[(size_t)(((double)k + 0.5) / QUANTILES * (double)samples)];
        }
    });
    return table;
}

// traffic.rs HOURLY_SHARES (the three road profiles' share of the day's flow in each hour) and
// the periods' first hours.
const double HOURLY_SHARES[PROFILES][24] = {
    {0.01309, 0.00918, 0.00785, 0.00838, 0.01246, 0.02353, 0.03973, 0.05108, 0.05364, 0.05517, 0.05799, 0.05978,
     0.06152, 0.06281, 0.06338, 0.06522, 0.06718, 0.06558, 0.05905, 0.04848, 0.03909, 0.03131, 0.02568, 0.01881},
    {0.00684, 0.00423, 0.00330, 0.00331, 0.00691, 0.02047, 0.04139, 0.05610, 0.05161, 0.05350, 0.05888, 0.06151,
     0.06501, 0.06797, 0.07131, 0.07483, 0.08065, 0.07760, 0.06261, 0.04433, 0.03251, 0.02441, 0.01922, 0.01150},
    {0.00545, 0.00336, 0.00242, 0.00236, 0.00555, 0.02012, 0.03996, 0.05863, 0.05247, 0.05348, 0.05843, 0.06219,
     0.06493, 0.06839, 0.07266, 0.07679, 0.08515, 0.08071, 0.06255, 0.04356, 0.03107, 0.02300, 0.01719, 0.00958},
};
const int PERIOD_FIRST_HOUR[PERIODS] = {7, 19, 23};
const int PERIOD_HOURS[PERIODS] = {12, 4, 8};
const int MAP_HOUR_GROUPS[PERIODS] = {4, 1, 3};

double hour_factor(int profile, int period, int slot) {
    const int hours = PERIOD_HOURS[period];
    double mean = 0.0;
    for (int s = 0; s < hours; ++s) mean += HOURLY_SHARES[profile][(PERIOD_FIRST_HOUR[period] + s) % 24];
    mean /= hours;
    return HOURLY_SHARES[profile][(PERIOD_FIRST_HOUR[period] + slot) % 24] / mean;
}

double quantile_probability(int k) { return ((double)k + 0.5) / QUANTILES; }

// The tables: node probabilities and shares (distribution.rs node_probabilities), each node's log
// intensity on every row of the quantile table at its quantile position (percentile.rs
// quantile_position / log_quantile, as AtProbabilities reads them), and the map's
// hour groups (traffic.rs hour_groups: a period's hours by their summed flow, cut into equal runs)
// with each profile's mean flow factor over each group.
void build_tables(const float* quantiles, Tables& T) {
    std::memset(&T, 0, sizeof(T));
    double p[NODES], share[NODES];
    for (int k = 0; k < 99; ++k) {
        p[k] = ((double)k + 0.5) / 100.0;
        share[k] = 0.01;
    }
    for (int k = 0; k < 24; ++k) {
        const double a = 2.0 + 0.25 * k, b = a + 0.25;
        p[99 + k] = 1.0 - std::pow(10.0, -(a + b) / 2.0);
        share[99 + k] = std::pow(10.0, -a) - std::pow(10.0, -b);
    }
    p[NODES - 1] = 1.0 - 0.5e-8;
    share[NODES - 1] = 1e-8;
    const double top = quantile_probability(QUANTILES - 1);
    T.first_tail = NODES;
    double cum = 0.0;
    for (int n = 0; n < NODES; ++n) {
        T.node_cum[n] = cum;
        cum += share[n];
        T.node_share[n] = (float)share[n];
        T.node_ln_p[n] = (float)std::log(p[n]);
        if (p[n] > top && n < T.first_tail) T.first_tail = n;
    }
    T.node_cum[NODES] = cum;
    T.top_ln_p = (float)std::log(top);
    for (int c = 0; c <= SHAPE_CAP; ++c) T.ln_factorial[c] = std::lgamma((double)c + 1.0);
    T.exceed_ln_p = (float)std::log(1.0 - HELD_EXCEEDED);
    T.fold_ln_exceed = std::log(1.0 - HELD_EXCEEDED);
    T.fold_ln_top = std::log(top);
    auto log_row = [&](double prob, float* out) {
        prob = std::max(prob, quantile_probability(0));
        const int k = std::min((int)std::floor(prob * QUANTILES - 0.5), QUANTILES - 2);
        const double t = (std::log(1.0 - prob) - std::log(1.0 - quantile_probability(k))) /
                         (std::log(1.0 - quantile_probability(k + 1)) - std::log(1.0 - quantile_probability(k)));
        for (int row = 0; row < ROWS; ++row) {
            const float* q = quantiles + (size_t)row * QUANTILES;
            const double here = std::log(std::max((double)q[k], 1e-30));
            const double next = std::log(std::max((double)q[k + 1], 1e-30));
            out[row] = (float)(here + t * (next - here));
            if (prob == top) T.fold_top[row] = here + t * (next - here);
        }
    };
    for (int n = 0; n < T.first_tail; ++n) log_row(p[n], T.log_rows[n]);
    log_row(top, T.log_rows[NODES]);
    for (int period = 0; period < PERIODS; ++period) {
        const int hours = PERIOD_HOURS[period];
        std::vector<int> slots(hours);
        std::vector<double> flow(hours, 0.0);
        for (int s = 0; s < hours; ++s) {
            slots[s] = s;
            for (int profile = 0; profile < PROFILES; ++profile) flow[s] += hour_factor(profile, period, s);
        }
        std::sort(slots.begin(), slots.end(),
                  [&](int a, int b) { return flow[a] < flow[b] || (flow[a] == flow[b] && a < b); });
        const int runs = MAP_HOUR_GROUPS[period];
        T.groups[period] = runs;
        for (int g = 0; g < runs; ++g) {
            const int first = g * hours / runs, last = (g + 1) * hours / runs;
            T.group_weight[period][g] = (float)((double)(last - first) / ((double)hours * STATES));
            std::vector<int> group(slots.begin() + first, slots.begin() + last);
            std::sort(group.begin(), group.end());
            for (int profile = 0; profile < PROFILES; ++profile) {
                double f = 0.0;
                for (int s : group) f += hour_factor(profile, period, s);
                T.factor[profile][period][g] = f / (double)(last - first);
            }
        }
    }
}

// The packed pixels' shape, checked on the host: each layer's line count a whole number the
// input holds, each line's profile none or a known one, the curves within the input.
template <typename V>
int validate(const V* input, size_t input_floats, const uint64_t* offsets, size_t pixels, char* err,
             size_t errlen) {
    for (size_t i = 0; i < pixels; ++i) {
        size_t at = offsets[i];
        for (int l = 0; l < LAYERS; ++l) {
            if (at > input_floats || input_floats - at < (size_t)PERIODS + 1)
                return pfail(err, errlen, "pixel %zu: layer %d beyond the input", i, l);
            const double count = (double)input[at + PERIODS];
            if (!(count >= 0.0 && count <= (double)LINES_MAX) || count != std::floor(count))
                return pfail(err, errlen, "pixel %zu: layer %d line count %g", i, l, (double)count);
            const size_t lines = (size_t)count;
            at += PERIODS + 1;
            if (input_floats - at < lines * LINE_FLOATS)
                return pfail(err, errlen, "pixel %zu: layer %d lines beyond the input", i, l);
            for (size_t j = 0; j < lines; ++j) {
                const double profile = (double)input[at + j * LINE_FLOATS + AT_PROFILE];
                if (!(profile == -1.0 || profile == 0.0 || profile == 1.0 || profile == 2.0))
                    return pfail(err, errlen, "pixel %zu: layer %d line %zu profile %g", i, l, j, (double)profile);
            }
            at += lines * LINE_FLOATS;
        }
        if (at > input_floats || input_floats - at < (size_t)CURVES * PERIODS * THIRDS)
            return pfail(err, errlen, "pixel %zu: curves beyond the input", i);
    }
    return 0;
}

struct Device {
    void* p = nullptr;
    ~Device() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" void qm_kurze_table(float* out) {
    const std::vector<float>& t = kurze_quantiles();
    std::copy(t.begin(), t.end(), out);
}

extern "C" int qm_pixel_loudness(qm_gpu* gpu, const float* input, size_t input_floats, const uint64_t* offsets,
                                 size_t pixels, const float* kurze, float* mixtures, float* ndens, char* err,
                                 size_t errlen) {
    if (!gpu || (pixels && (!input || !offsets || !ndens))) return pfail(err, errlen, "null argument");
    if (pixels == 0) return 0;
    if (pixels > (size_t)INT32_MAX) return pfail(err, errlen, "%zu pixels exceed one launch", pixels);
    if (validate(input, input_floats, offsets, pixels, err, errlen)) return 1;
    static_assert(sizeof(Shared) <= 64 * 1024, "the workgroup's LDS");
    Tables* tables = new (std::nothrow) Tables;
    if (!tables) return pfail(err, errlen, "out of host memory");
    build_tables(kurze ? kurze : kurze_quantiles().data(), *tables);
    const size_t mix_floats = pixels * PERIODS * CURVES * BINS * 2;
    Device d_input, d_offsets, d_tables, d_mix, d_ndens;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) e = hipMalloc(&d_input.p, sizeof(float) * std::max<size_t>(input_floats, 1));
    if (e == hipSuccess) e = hipMalloc(&d_offsets.p, sizeof(uint64_t) * pixels);
    if (e == hipSuccess) e = hipMalloc(&d_tables.p, sizeof(Tables));
    if (e == hipSuccess && mixtures) e = hipMalloc(&d_mix.p, sizeof(float) * mix_floats);
    if (e == hipSuccess) e = hipMalloc(&d_ndens.p, sizeof(float) * pixels * CURVES);
    if (e == hipSuccess) e = hipMemcpy(d_input.p, input, sizeof(float) * input_floats, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_offsets.p, offsets, sizeof(uint64_t) * pixels, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tables.p, tables, sizeof(Tables), hipMemcpyHostToDevice);
    delete tables;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pixel_loudness_kernel, dim3((unsigned)pixels), dim3(GROUP), 0, 0, (const float*)d_input.p,
                           (const uint64_t*)d_offsets.p, (const Tables*)d_tables.p, (float*)d_mix.p,
                           (float*)d_ndens.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(ndens, d_ndens.p, sizeof(float) * pixels * CURVES, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mixtures)
        e = hipMemcpy(mixtures, d_mix.p, sizeof(float) * mix_floats, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return pfail(err, errlen, "pixel loudness: %s", hipGetErrorString(e));
    return 0;
}

extern "C" int qm_pixel_loudness_unfolded(qm_gpu* gpu, const double* unfolded, size_t input_floats,
                                          const uint64_t* offsets, size_t pixels, const float* kurze,
                                          float* mixtures, float* ndens, float* folded, char* err, size_t errlen) {
    if (!gpu || (pixels && (!unfolded || !offsets || !ndens))) return pfail(err, errlen, "null argument");
    if (pixels == 0) return 0;
    if (pixels > (size_t)INT32_MAX) return pfail(err, errlen, "%zu pixels exceed one launch", pixels);
    if (validate(unfolded, input_floats, offsets, pixels, err, errlen)) return 1;
    // Each pixel's first line among all the input's lines: the fold's per-line flags live there.
    std::vector<uint64_t> line_first(pixels);
    uint64_t lines = 0;
    for (size_t i = 0; i < pixels; ++i) {
        line_first[i] = lines;
        size_t at = offsets[i];
        for (int l = 0; l < LAYERS; ++l) {
            const uint64_t n = (uint64_t)unfolded[at + PERIODS];
            lines += n;
            at += PERIODS + 1 + n * LINE_FLOATS;
        }
    }
    Tables* tables = new (std::nothrow) Tables;
    if (!tables) return pfail(err, errlen, "out of host memory");
    build_tables(kurze ? kurze : kurze_quantiles().data(), *tables);
    const size_t mix_floats = pixels * PERIODS * CURVES * BINS * 2;
    Device d_input, d_offsets, d_first, d_tables, d_work, d_keep, d_ratio, d_out, d_mix, d_ndens;
    hipError_t e = hipSetDevice(gpu->ordinal);
    if (e == hipSuccess) e = hipMalloc(&d_input.p, sizeof(double) * std::max<size_t>(input_floats, 1));
    if (e == hipSuccess) e = hipMalloc(&d_work.p, sizeof(double) * std::max<size_t>(input_floats, 1));
    if (e == hipSuccess) e = hipMalloc(&d_out.p, sizeof(float) * std::max<size_t>(input_floats, 1));
    if (e == hipSuccess) e = hipMalloc(&d_offsets.p, sizeof(uint64_t) * pixels);
    if (e == hipSuccess) e = hipMalloc(&d_first.p, sizeof(uint64_t) * pixels);
    if (e == hipSuccess) e = hipMalloc(&d_keep.p, std::max<uint64_t>(lines, 1));
    if (e == hipSuccess) e = hipMalloc(&d_ratio.p, sizeof(double) * std::max<uint64_t>(lines, 1));
    if (e == hipSuccess) e = hipMalloc(&d_tables.p, sizeof(Tables));
    if (e == hipSuccess && mixtures) e = hipMalloc(&d_mix.p, sizeof(float) * mix_floats);
    if (e == hipSuccess) e = hipMalloc(&d_ndens.p, sizeof(float) * pixels * CURVES);
    if (e == hipSuccess) e = hipMemcpy(d_input.p, unfolded, sizeof(double) * input_floats, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_offsets.p, offsets, sizeof(uint64_t) * pixels, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d_first.p, line_first.data(), sizeof(uint64_t) * pixels, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tables.p, tables, sizeof(Tables), hipMemcpyHostToDevice);
    delete tables;
    // Floats past a folded pixel's end read NaN (all bits set).
    if (e == hipSuccess) e = hipMemset(d_out.p, 0xFF, sizeof(float) * input_floats);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(fold_kernel, dim3((unsigned)((pixels + 63) / 64)), dim3(64), 0, 0,
                           (const double*)d_input.p, (const uint64_t*)d_offsets.p, (const uint64_t*)d_first.p,
                           (uint32_t)pixels, (const Tables*)d_tables.p, (double*)d_work.p, (uint8_t*)d_keep.p,
                           (double*)d_ratio.p, (float*)d_out.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(pixel_loudness_kernel, dim3((unsigned)pixels), dim3(GROUP), 0, 0, (const float*)d_out.p,
                           (const uint64_t*)d_offsets.p, (const Tables*)d_tables.p, (float*)d_mix.p,
                           (float*)d_ndens.p);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(ndens, d_ndens.p, sizeof(float) * pixels * CURVES, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mixtures)
        e = hipMemcpy(mixtures, d_mix.p, sizeof(float) * mix_floats, hipMemcpyDeviceToHost);
    if (e == hipSuccess && folded)
        e = hipMemcpy(folded, d_out.p, sizeof(float) * input_floats, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return pfail(err, errlen, "pixel loudness of unfolded pixels: %s", hipGetErrorString(e));
    return 0;
}
