// A painted pixel's loudness (Nden per layer and of the place) from its packed sound, on the card:
// the level distributions of each period, hour group and weather state (popup/src/distribution.rs,
// popup/src/percentiles.rs layered_distributions, layer_states, moment), Kurze's quantile line
// shapes, events and streams of passes (physics/src/percentile.rs, distribution.rs line_values,
// pass_values), mixed into each period's distributions; then the loudness of each mixture on its
// curve at the whole decibels its levels need (popup/src/loudness.rs Curve::at, loudness) with
// ISO 532-1 per step. The reference card runs the same two passes (paint-gpu/kernels/
// distribution.cuh pixel_mixtures, loudness.cuh pixel_ndens) in FP32 at the map's resolution.
//
// Work decomposition. `mixtures_kernel`: a 64-lane workgroup per (pixel, period, hour group,
// weather state). Its distributions live in LDS (a layer's in two buffers, the place's in two, a
// line's values); every lane works a share of each step (the line's values, the source bins of a
// sum, whose runs it adds with LDS atomics), and the results are added with the group's share of
// the period into the pixel's mixtures with global atomics. A pixel with no line on an hourly
// profile has one hour group; its other groups' workgroups return at once. `ndens_kernel`: a
// workgroup per pixel, a period at a time: the whole-decibel steps each curve's bins need, ISO
// 532-1 at those steps (qm_zwicker.hpp), then each curve's mean over its bins.
//
// Host side: the quantile table (the reference's seeded simulation, restated from
// physics/src/percentile_tests.rs:1-100, or the caller's), the nodes' parts (AtProbabilities), the
// map's hour groups and profile factors (popup/src/percentiles/traffic.rs hour_groups,
// hour_factor), validation of the packed pixels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/qm_paint.h"
#include "qm_internal.hpp"
#include "qm_zwicker.hpp"

namespace {

constexpr int PERIODS = 3;
constexpr int STATES = 10;
constexpr int LAYERS = 6;
constexpr int CURVES = LAYERS + 1;
constexpr int THIRDS = 28;
constexpr int LINE_FLOATS = PERIODS + PERIODS * STATES + PERIODS + 1 + 2 * PERIODS;
constexpr int NODES = 124;             // distribution.rs node_probabilities
constexpr int TOP_NODE = NODES;        // the table's last quantile
constexpr int HELD_NODE = NODES + 1;   // 1 - HELD_EXCEEDED
constexpr int ROW_NODES = NODES + 2;
constexpr int ROWS = 41;
constexpr int QUANTILES = 100;
constexpr int BINS = QM_LOUD_BINS;     // the map's 2 dB grid over 180 dB
constexpr int LAST = BINS - 1;
constexpr int VALUES_MAX = NODES;      // >= BINS + 1
constexpr int GROUPS_MAX = 4;
constexpr int PROFILES = 3;
constexpr int CURVE_STEPS = 190;       // whole decibels from -30 to 160
constexpr int LINES_MAX = 1 << 16;
constexpr float BIN_DB = 2.0f;
constexpr float LEVEL_MIN_DB = -30.0f;
constexpr float CURVE_LOW_DB = -30.0f;
constexpr float LAMBDA_MIN = 1e-4f;
constexpr float STEPS_PER_DECADE = 8.0f;
constexpr float SIMULATED_LAMBDA_MAX = 10.0f;
constexpr float HOLDS_MEAN_LAMBDA = 10.0f;
constexpr float HELD_SHARE = 1e-2f;
constexpr float PASS_TAIL = 1e-9f;
constexpr float PI_F = 3.14159265358979323846f;
constexpr double PI_D = 3.14159265358979323846;
constexpr double PASSES_HOLD_MEAN_D = 2.0 * PI_D * 10.0;
constexpr float LN10_F = 2.302585092994045684f;
static_assert(LINE_FLOATS == 43, "packed line");
static_assert(VALUES_MAX >= BINS + 1, "a distribution's bins fit the values");

// The tables the kernels read (global memory: the lanes read different entries).
struct LoudTables {
    float normal[ROW_NODES];
    float ln_p[ROW_NODES];
    uint32_t tail[ROW_NODES];
    float rows[ROWS * ROW_NODES];  // row-major: a row's nodes together
    float shares[NODES];
    uint32_t group_count[PERIODS];
    uint32_t group_slots[PERIODS * GROUPS_MAX];
    float factor[PROFILES * PERIODS * GROUPS_MAX];
};

__constant__ float PERIOD_HOURS_F[PERIODS] = {12.0f, 4.0f, 8.0f};
__constant__ float PENALTY_DB_F[PERIODS] = {0.0f, 5.0f, 10.0f};

__device__ __forceinline__ float energy_f(float db) { return expf(db * (LN10_F / 10.0f)); }

// distribution.rs Distribution, FP32: the share of the time in silence and per bin its share and
// share x mean intensity.
struct Dist {
    float silent;
    float bins[BINS][2];
};

// A line's values ascending with the sums of shares and of shares x values from each value up
// (from the top, so that a rare loud pass keeps its share in FP32).
struct Values {
    int count;
    float value[VALUES_MAX];
    float shares_from[VALUES_MAX + 1];
    float weighted_from[VALUES_MAX + 1];
};

struct Work {
    Dist layer[2];
    Dist whole[2];
    Values values;
    int start[LAYERS], count[LAYERS];
    float floor[LAYERS], held[LAYERS];
    int moves[LAYERS];
    int profiled;
};

__device__ __forceinline__ int bin_of(float level_db) {
    const float position = rintf((level_db - LEVEL_MIN_DB) / BIN_DB);
    return position <= 0.0f ? 0 : (position >= (float)LAST ? LAST : (int)position);
}

__device__ __forceinline__ float ceiling_of(int bin) { return energy_f(LEVEL_MIN_DB + ((float)bin + 0.5f) * BIN_DB); }

// Sums from the top of the values' shares (held in shares_from on entry): each lane sums two
// neighbouring entries, then a suffix scan over the lanes.
__device__ void sums_from_top(Values& v) {
    const int lane = threadIdx.x;
    const int count = v.count;
    const int k0 = 2 * lane, k1 = 2 * lane + 1;
    const float s0 = k0 < count ? v.shares_from[k0] : 0.0f, s1 = k1 < count ? v.shares_from[k1] : 0.0f;
    const float w0 = k0 < count ? s0 * v.value[k0] : 0.0f, w1 = k1 < count ? s1 * v.value[k1] : 0.0f;
    float s = s0 + s1, w = w0 + w1;
    // Inclusive suffix sums over lanes: lane l gets the sum of lanes >= l.
#pragma unroll
    for (int offset = 1; offset < 64; offset <<= 1) {
        const float so = __shfl_down(s, offset, 64), wo = __shfl_down(w, offset, 64);
        if (lane + offset < 64) {
            s += so;
            w += wo;
        }
    }
    __syncthreads();
    if (k0 < count) {
        v.shares_from[k0] = s;
        v.weighted_from[k0] = w;
    }
    if (k1 < count) {
        v.shares_from[k1] = s - s0;
        v.weighted_from[k1] = w - w0;
    }
    if (lane == 0) v.shares_from[count] = v.weighted_from[count] = 0.0f;
    __syncthreads();
}

// distribution.rs first_at_or_above: the first value from `from` reaching `limit`, galloping.
__device__ __forceinline__ int first_at_or_above(const Values& v, int from, float limit) {
    int low = from, step = 1;
    while (low + step - 1 < v.count && v.value[low + step - 1] < limit) {
        low += step;
        step *= 2;
    }
    int high = min(low + step - 1, v.count);
    while (low < high) {
        const int middle = (low + high) / 2;
        if (v.value[middle] < limit) low = middle + 1;
        else high = middle;
    }
    return low;
}

// Distribution::runs, one lane: a moment at `intensity` (0 silent) holding `share` of the time with
// each of the values x `scale` added, each bin's run of values added at once (LDS atomics).
__device__ void add_runs(Dist& out, float intensity, float share, const Values& v, float scale) {
    int from = 0;
    if (intensity <= 0.0f) {
        while (from < v.count && v.value[from] <= 0.0f) ++from;
        if (from > 0) atomicAdd(&out.silent, share * (v.shares_from[0] - v.shares_from[from]));
    }
    if (from == v.count) return;
    const float step = energy_f(BIN_DB);
    int bin = bin_of(10.0f * log10f(intensity + scale * v.value[from]));
    float ceiling = ceiling_of(bin);
    for (;;) {
        const int to = bin == LAST ? v.count : first_at_or_above(v, from + 1, (ceiling - intensity) / scale);
        const float shares = v.shares_from[from] - v.shares_from[to];
        const float weighted = scale * (v.weighted_from[from] - v.weighted_from[to]);
        atomicAdd(&out.bins[bin][0], share * shares);
        atomicAdd(&out.bins[bin][1], share * (intensity * shares + weighted));
        from = to;
        if (from >= v.count) break;
        const float value = v.value[from];
        do {
            ++bin;
            ceiling *= step;
        } while (!(bin == LAST || intensity + scale * value < ceiling));
    }
}

__device__ void clear(Dist& d) {
    for (int k = threadIdx.x; k < 2 * BINS; k += 64) (&d.bins[0][0])[k] = 0.0f;
    if (threadIdx.x == 0) d.silent = 0.0f;
    __syncthreads();
}

// Distribution::with_line: `out` = `d` with the values x `scale` added at every moment.
__device__ void with_line(const Dist& d, const Values& v, float scale, Dist& out) {
    clear(out);
    if (threadIdx.x == 0 && d.silent > 0.0f) add_runs(out, 0.0f, d.silent, v, scale);
    for (int b = threadIdx.x; b < BINS; b += 64)
        if (d.bins[b][0] > 0.0f) add_runs(out, d.bins[b][1] / d.bins[b][0], d.bins[b][0], v, scale);
    __syncthreads();
}

// Distribution::over on silence: every moment at `intensity` (silence at none).
__device__ void steady(float intensity, Dist& out) {
    clear(out);
    if (threadIdx.x == 0) {
        if (intensity > 0.0f) {
            const int b = bin_of(10.0f * log10f(intensity));
            out.bins[b][0] = 1.0f;
            out.bins[b][1] = intensity;
        } else {
            out.silent = 1.0f;
        }
    }
    __syncthreads();
}

// Distribution::with_distribution: `d` with an independent `other` added; other's levels (and its
// silence, a value of nothing) become a line's values in order.
__device__ void with_distribution(const Dist& d, const Dist& other, Values& v, Dist& out) {
    if (threadIdx.x == 0) {
        int count = 0;
        if (other.silent > 0.0f) {
            v.value[0] = 0.0f;
            v.shares_from[0] = other.silent;
            count = 1;
        }
        for (int b = 0; b < BINS; ++b)
            if (other.bins[b][0] > 0.0f) {
                v.value[count] = other.bins[b][1] / other.bins[b][0];
                v.shares_from[count] = other.bins[b][0];
                ++count;
            }
        v.count = count;
    }
    __syncthreads();
    sums_from_top(v);
    with_line(d, v, 1.0f, out);
}

// Where lambda falls in the table: its row and the fraction toward the next.
struct Row {
    int row;
    float fraction;
};

__device__ __forceinline__ Row row_of(float lambda) {
    const float position =
        fminf(fmaxf(log10f(lambda / LAMBDA_MIN) * STEPS_PER_DECADE, 0.0f), (float)(ROWS - 1));
    const int row = (int)floorf(position);
    return Row{row, position - (float)row};
}

// percentile.rs AtProbabilities::relative_intensities at node k, the parts that hang on lambda
// alone given (`at_top`, `passing_top`, the row).
__device__ float relative_intensity(const LoudTables& T, float lambda, int k, float at_top, float passing_top, Row at) {
    if (isinf(lambda)) return 1.0f;
    if (lambda <= 0.0f) return 0.0f;
    if (lambda >= SIMULATED_LAMBDA_MAX) {
        const float deviation = 1.0f / sqrtf(2.0f * PI_F * lambda);
        return fmaxf(1.0f + deviation * T.normal[k], 0.0f);
    }
    if (T.tail[k]) {
        const float s = -T.ln_p[k] / (2.0f * lambda);
        return at_top + 1.0f / (lambda * PI_F * (1.0f + s * s)) - passing_top;
    }
    if (lambda < LAMBDA_MIN) return lambda / LAMBDA_MIN * expf(T.rows[k]);
    const float low = T.rows[at.row * ROW_NODES + k];
    const float high = T.rows[min(at.row + 1, ROWS - 1) * ROW_NODES + k];
    return expf(low + at.fraction * (high - low));
}

__device__ void top_parts(const LoudTables& T, float lambda, Row at, float& at_top, float& passing_top) {
    at_top = passing_top = 0.0f;
    if (isinf(lambda) || lambda <= 0.0f || lambda >= SIMULATED_LAMBDA_MAX) return;
    at_top = relative_intensity(T, lambda, TOP_NODE, 0.0f, 0.0f, at);
    const float s = -T.ln_p[TOP_NODE] / (2.0f * lambda);
    passing_top = 1.0f / (lambda * PI_F * (1.0f + s * s));
}

// The intensity a line of `lambda` exceeds a thousandth of the time, relative to its mean.
__device__ float exceeded_of(const LoudTables& T, float lambda) {
    const Row at = row_of(lambda);
    float at_top, passing_top;
    top_parts(T, lambda, at, at_top, passing_top);
    return relative_intensity(T, lambda, HELD_NODE, at_top, passing_top, at);
}

// Kurze's quantiles at lambda relative to the line's mean (lines.rs Line::shape).
__device__ void kurze_values(const LoudTables& T, float lambda, Values& v) {
    const Row at = row_of(lambda);
    float at_top, passing_top;
    top_parts(T, lambda, at, at_top, passing_top);
    for (int k = threadIdx.x; k < NODES; k += 64) {
        v.value[k] = relative_intensity(T, lambda, k, at_top, passing_top, at);
        v.shares_from[k] = T.shares[k];
    }
    if (threadIdx.x == 0) v.count = NODES;
    __syncthreads();
    sums_from_top(v);
}

// An event's values relative to its mean: silent off its duty, its mean over the duty on it.
__device__ void event_values(float duty, Values& v) {
    if (threadIdx.x == 0) {
        v.count = 2;
        v.value[0] = 0.0f;
        v.value[1] = 1.0f / duty;
        v.shares_from[0] = 1.0f - duty;
        v.shares_from[1] = duty;
    }
    __syncthreads();
    sums_from_top(v);
}

// distribution.rs pass_values in FP32: k passes in progress (Poisson with `mean`) at k / mean for
// k below the first count past the mean whose next share falls under PASS_TAIL (or the values'
// room), the last value the tail's mean over its share (64 further counts summed).
__device__ void pass_values(float mean, Values& v) {
    const float ln_mean = logf(mean);
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        int last = max(1, (int)ceilf(mean));
        while (last < VALUES_MAX - 1 &&
               -mean + (float)(last + 1) * ln_mean - lgammaf((float)(last + 2)) > logf(PASS_TAIL))
            ++last;
        s_last = last;
    }
    __syncthreads();
    const int last = s_last;
    for (int k = threadIdx.x; k < last; k += 64) {
        v.value[k] = (float)k / mean;
        v.shares_from[k] = expf(-mean + (float)k * ln_mean - lgammaf((float)(k + 1)));
    }
    if (threadIdx.x == 0) {
        float share = expf(-mean + (float)last * ln_mean - lgammaf((float)(last + 1)));
        float left = 0.0f, energy = 0.0f;
        for (int k = last; k < last + 64 && share > 0.0f; ++k) {
            left += share;
            energy += share * (float)k / mean;
            share *= mean / (float)(k + 1);
        }
        v.count = last + 1;
        v.value[last] = left > 0.0f ? energy / left : (float)last / mean;
        v.shares_from[last] = left;
    }
    __syncthreads();
    sums_from_top(v);
}

// One packed line (paint/src/sone.rs pack_line).
struct Line {
    const float* f;
    __device__ float energy(int p) const { return f[p]; }
    __device__ float mean(int p, int state) const { return f[PERIODS + p * STATES + state]; }
    __device__ float lambda(int p) const { return f[PERIODS + PERIODS * STATES + p]; }
    __device__ int profile() const { return (int)f[2 * PERIODS + PERIODS * STATES]; }
    __device__ float duty(int p) const { return f[2 * PERIODS + PERIODS * STATES + 1 + p]; }
    __device__ float passing(int p) const { return f[3 * PERIODS + PERIODS * STATES + 1 + p]; }
    // Line::shape of a stream of passes: it moves the level unless so many are in progress that
    // it holds its mean.
    __device__ bool passes_move(int p) const {
        const float m = passing(p);
        return m > 0.0f && m < (float)PASSES_HOLD_MEAN_D;
    }
    __device__ float factor(const LoudTables& T, int p, int group) const {
        const int pr = profile();
        return pr >= 0 ? T.factor[(pr * PERIODS + p) * GROUPS_MAX + group] : 1.0f;
    }
};

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
    return x;
}

__device__ __forceinline__ bool wave_any(bool b) { return __ballot(b) != 0; }

// percentiles.rs moment(): a layer's energy that holds steady (its steady energy, the lines that
// hold their mean, the Kurze lines whose loud moments stay under a hundredth of the steady
// energy), the energy the held test compares against, and whether a line moves the level.
__device__ float layer_floor(const LoudTables& T, const float* lines, int count, float steady_energy, int p, int group,
                             int state, float& held, bool& moves) {
    float floor_part = threadIdx.x == 0 ? steady_energy : 0.0f, candidates = 0.0f;
    bool moved = false;
    for (int i = threadIdx.x; i < count; i += 64) {
        const Line line{lines + (size_t)i * LINE_FLOATS};
        const float factor = line.factor(T, p, group);
        const float mean = line.mean(p, state) * factor;
        if (line.energy(p) <= 0.0f || mean <= 0.0f) continue;
        const float duty = line.duty(p);
        if (!isnan(duty)) {
            if (duty > 0.0f) moved = true;
            continue;
        }
        if (!isnan(line.passing(p))) {
            if (line.passes_move(p)) moved = true;
            else floor_part += mean;
            continue;
        }
        if (line.lambda(p) * factor >= HOLDS_MEAN_LAMBDA) floor_part += mean;
        else candidates += mean;
    }
    float floor = wave_sum(floor_part);
    const float candidate_sum = wave_sum(candidates);
    held = floor + candidate_sum;
    if (candidate_sum > 0.0f) {
        float extra = 0.0f;
        for (int i = threadIdx.x; i < count; i += 64) {
            const Line line{lines + (size_t)i * LINE_FLOATS};
            const float factor = line.factor(T, p, group);
            const float mean = line.mean(p, state) * factor;
            const float lambda = line.lambda(p) * factor;
            if (line.energy(p) <= 0.0f || mean <= 0.0f || !isnan(line.duty(p)) || !isnan(line.passing(p)) ||
                lambda >= HOLDS_MEAN_LAMBDA)
                continue;
            if (mean * exceeded_of(T, lambda) < HELD_SHARE * held) extra += mean;
            else moved = true;
        }
        floor += wave_sum(extra);
    }
    moves = wave_any(moved);
    return floor;
}

// percentiles.rs layer_states for one state: the layer's floor, then its moving lines added in
// moment()'s order (events and streams of passes, then the Kurze lines that move). Returns which
// buffer of w.layer holds the result.
__device__ int layer_distribution(const LoudTables& T, Work& w, const float* lines, int count, float floor, float held,
                                  bool moves, int p, int group, int state) {
    steady(floor, w.layer[0]);
    int now = 0;
    if (!moves) return now;
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < count; ++i) {
            const Line line{lines + (size_t)i * LINE_FLOATS};
            const float factor = line.factor(T, p, group);
            const float mean = line.mean(p, state) * factor;
            if (line.energy(p) <= 0.0f || mean <= 0.0f) continue;
            const float duty = line.duty(p), lambda = line.lambda(p) * factor;
            const bool passes = !isnan(line.passing(p));
            if (pass == 0) {
                if (passes ? !line.passes_move(p) : (isnan(duty) || duty <= 0.0f)) continue;
                if (passes) pass_values(line.passing(p), w.values);
                else event_values(duty, w.values);
            } else {
                if (!isnan(duty) || passes || lambda >= HOLDS_MEAN_LAMBDA) continue;
                if (mean * exceeded_of(T, lambda) < HELD_SHARE * held) continue;
                kurze_values(T, lambda, w.values);
            }
            with_line(w.layer[now], w.values, mean, w.layer[1 - now]);
            now = 1 - now;
        }
    return now;
}

__device__ void accumulate(float* mixture, const Dist& d, float weight) {
    for (int b = threadIdx.x; b < BINS; b += 64)
        if (d.bins[b][0] > 0.0f) {
            atomicAdd(&mixture[2 * b], weight * d.bins[b][0]);
            atomicAdd(&mixture[2 * b + 1], weight * d.bins[b][1]);
        }
}

// One workgroup per (pixel, period, group, state): blockIdx.x = ((pixel * PERIODS + p) *
// GROUPS_MAX + group) * STATES + state.
__global__ __launch_bounds__(64) void mixtures_kernel(const float* __restrict__ input, const uint64_t* __restrict__ offsets,
                                                      const LoudTables* __restrict__ tables, uint32_t pixels,
                                                      float* __restrict__ mixtures) {
    __shared__ Work w;
    const LoudTables& T = *tables;
    const uint32_t task = blockIdx.x;
    const int state = task % STATES;
    const int group = (task / STATES) % GROUPS_MAX;
    const int p = (task / (STATES * GROUPS_MAX)) % PERIODS;
    const uint32_t pixel = task / (STATES * GROUPS_MAX * PERIODS);
    if (pixel >= pixels || group >= (int)T.group_count[p]) return;
    const float* in = input + offsets[pixel];
    if (threadIdx.x == 0) {
        int at = 0;
        for (int l = 0; l < LAYERS; ++l) {
            w.count[l] = (int)in[at + PERIODS];
            w.start[l] = at + PERIODS + 1;
            at = w.start[l] + w.count[l] * LINE_FLOATS;
        }
    }
    __syncthreads();
    bool profiled = false;
    for (int l = 0; l < LAYERS; ++l)
        for (int i = threadIdx.x; i < w.count[l]; i += 64)
            profiled = profiled || Line{in + w.start[l] + (size_t)i * LINE_FLOATS}.profile() >= 0;
    profiled = wave_any(profiled);
    if (!profiled && group > 0) return;
    const float weight = profiled ? (float)T.group_slots[p * GROUPS_MAX + group] / (PERIOD_HOURS_F[p] * STATES)
                                  : 1.0f / STATES;
    float* mixture = mixtures + ((size_t)pixel * PERIODS + p) * CURVES * BINS * 2;
    float floor = 0.0f;
    for (int l = 0; l < LAYERS; ++l) {
        float held;
        bool moves;
        const float f = layer_floor(T, in + w.start[l], w.count[l], in[w.start[l] - PERIODS - 1 + p], p, group, state,
                                    held, moves);
        if (threadIdx.x == 0) {
            w.floor[l] = f;
            w.held[l] = held;
            w.moves[l] = moves;
        }
        if (!moves) floor += f;
    }
    __syncthreads();
    steady(floor, w.whole[0]);
    int whole = 0;
    for (int l = 0; l < LAYERS; ++l) {
        const int own = layer_distribution(T, w, in + w.start[l], w.count[l], w.floor[l], w.held[l], w.moves[l] != 0, p,
                                           group, state);
        accumulate(mixture + (size_t)l * BINS * 2, w.layer[own], weight);
        if (w.moves[l]) {
            with_distribution(w.whole[whole], w.layer[own], w.values, w.whole[1 - whole]);
            whole = 1 - whole;
        }
        __syncthreads();
    }
    accumulate(mixture + (size_t)LAYERS * BINS * 2, w.whole[whole], weight);
}

// A mixture bin's place on its curve: the whole decibel below its level and the fraction above.
__device__ __forceinline__ int curve_step(const float* bin, int p, float& fraction) {
    const float level = 10.0f * log10f(bin[1] / bin[0]) + PENALTY_DB_F[p];
    const float position = fminf(fmaxf(level - CURVE_LOW_DB, 0.0f), (float)CURVE_STEPS);
    const int below = min((int)floorf(position), CURVE_STEPS - 1);
    fraction = position - (float)below;
    return below;
}

// Nden per curve of each pixel from its mixtures: a 64-lane workgroup per pixel, a period at a
// time (loudness.rs loudness and day; Curve::at at the needed whole decibels).
__global__ __launch_bounds__(64) void ndens_kernel(const float* __restrict__ input, const uint64_t* __restrict__ offsets,
                                                   const float* __restrict__ mixtures, uint32_t pixels,
                                                   float* __restrict__ ndens) {
    const uint32_t pixel = blockIdx.x;
    if (pixel >= pixels) return;
    const float* in = input + offsets[pixel];
    __shared__ float curve[CURVES][CURVE_STEPS + 1];
    __shared__ unsigned char needed[CURVES][CURVE_STEPS + 1];
    __shared__ float nden[CURVES];
    __shared__ int curves_at;
    const int t = threadIdx.x;
    if (t == 0) {
        int at = 0;
        for (int l = 0; l < LAYERS; ++l) at += PERIODS + 1 + (int)in[at + PERIODS] * LINE_FLOATS;
        curves_at = at;
    }
    if (t < CURVES) nden[t] = 0.0f;
    __syncthreads();
    for (int p = 0; p < PERIODS; ++p) {
        const float* mixture = mixtures + ((size_t)pixel * PERIODS + p) * CURVES * BINS * 2;
        for (int k = t; k < CURVES * (CURVE_STEPS + 1); k += 64) (&needed[0][0])[k] = 0;
        __syncthreads();
        for (int k = t; k < CURVES * BINS; k += 64) {
            const int c = k / BINS;
            const float* bin = mixture + 2 * k;
            if (bin[0] <= 0.0f || isnan(in[curves_at + (c * PERIODS + p) * THIRDS])) continue;
            float fraction;
            const int below = curve_step(bin, p, fraction);
            needed[c][below] = 1;
            needed[c][below + 1] = 1;
        }
        __syncthreads();
        for (int k = t; k < CURVES * (CURVE_STEPS + 1); k += 64) {
            const int c = k / (CURVE_STEPS + 1), step = k % (CURVE_STEPS + 1);
            if (!needed[c][step]) continue;
            const float* shape = in + curves_at + (c * PERIODS + p) * THIRDS;
            float levels[THIRDS];
            for (int n = 0; n < THIRDS; ++n) levels[n] = shape[n] + CURVE_LOW_DB + (float)step;
            curve[c][step] = zwicker_sone(levels);
        }
        __syncthreads();
        if (t < CURVES && !isnan(in[curves_at + (t * PERIODS + p) * THIRDS])) {
            float mean = 0.0f;
            for (int b = 0; b < BINS; ++b) {
                const float* bin = mixture + 2 * (t * BINS + b);
                if (bin[0] <= 0.0f) continue;
                float fraction;
                const int below = curve_step(bin, p, fraction);
                mean += bin[0] * (curve[t][below] + (curve[t][below + 1] - curve[t][below]) * fraction);
            }
            nden[t] += PERIOD_HOURS_F[p] * mean / 24.0f;
        }
        __syncthreads();
    }
    if (t < CURVES) ndens[(size_t)pixel * CURVES + t] = nden[t];
}

// =============================================================================================
// Host side.

int lfail(char* err, size_t errlen, const char* fmt, ...) {
    if (err && errlen) {
        va_list args;
        va_start(args, fmt);
        vsnprintf(err, errlen, fmt, args);
        va_end(args);
    }
    return 1;
}

// The quantile table's seeded simulation (physics/src/percentile_tests.rs: SplitMix64 from
// 0x5eed_1971; per row 20,000 samples, up to a million for sparse lines, of a Poisson line over at
// least 200 distances each side; the quantiles at (k + 0.5) / 100, as FP32).
struct SplitMix {
    uint64_t state;
    double uniform() {
        state += 0x9e3779b97f4a7c15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return (double)((z ^ (z >> 31)) >> 11) / (double)(1ull << 53);
    }
    size_t poisson(double mean) {
        if (mean > 50.0) {
            const double radius = std::sqrt(-2.0 * std::log(1.0 - uniform()));
            const double normal = radius * std::cos(2.0 * PI_D * uniform());
            return (size_t)std::max(std::round(mean + std::sqrt(mean) * normal), 0.0);
        }
        const double limit = std::exp(-mean);
        size_t count = 0;
        for (double product = uniform(); product > limit; product *= uniform()) ++count;
        return count;
    }
};

const std::vector<float>& seeded_table() {
    static std::vector<float> table;
    static std::once_flag once;
    std::call_once(once, [] {
        SplitMix r{0x5eed1971ull};
        table.assign(ROWS * QUANTILES, 0.0f);
        std::vector<double> samples;
        for (int row = 0; row < ROWS; ++row) {
            const double lambda = 1e-4 * std::pow(10.0, row / 8.0);
            const double half = std::max(5.0 / lambda, 200.0);
            const double inside_mean = lambda * 2.0 * half;
            const double outside = lambda * 2.0 * (PI_D / 2.0 - std::atan(half));
            const double mean = lambda * PI_D;
            const size_t count = (size_t)std::min(20000.0 * std::max(0.01 / lambda, 1.0), 1000000.0);
            samples.resize(count);
            for (size_t i = 0; i < count; ++i) {
                const size_t n = r.poisson(inside_mean);
                double inside = 0.0;
                for (size_t k = 0; k < n; ++k) {
                    const double x = (2.0 * r.uniform() - 1.0) * half;
                    inside += 1.0 / (1.0 + x * x);
                }
                samples[i] = (inside + outside) / mean;
            }
            std::sort(samples.begin(), samples.end());
            for (int k = 0; k < QUANTILES; ++k)
                table[row * QUANTILES + k] = (float)samples[(size_t)(((double)k + 0.5) / QUANTILES * (double)count)];
        }
    });
    return table;
}

double probability(int k) { return ((double)k + 0.5) / QUANTILES; }

// Acklam's inverse of the standard normal distribution (P. J. Acklam, 2003), percentile.rs
// normal_quantile.
double normal_quantile(double p) {
    const double a[6] = {-3.969683028665376e1, 2.209460984245205e2, -2.759285104469687e2,
                         1.38357751867269e2,   -3.066479806614716e1, 2.506628277459239};
    const double b[5] = {-5.447609879822406e1, 1.615858368580409e2, -1.556989798598866e2, 6.680131188771972e1,
                         -1.328068155288572e1};
    const double c[6] = {-7.784894002430293e-3, -3.223964580411365e-1, -2.400758277161838,
                         -2.549671010366478,    4.374664141464968,     2.938163982698783};
    const double d[4] = {7.784695709041462e-3, 3.224671290700398e-1, 2.445134137142996, 3.754408661907416};
    auto tail = [&](double q) {
        return (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
               ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1.0);
    };
    if (p < 0.02425) return tail(std::sqrt(-2.0 * std::log(p)));
    if (p > 1.0 - 0.02425) return -tail(std::sqrt(-2.0 * std::log(1.0 - p)));
    const double q = p - 0.5, r = q * q;
    return (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
           (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1.0);
}

// traffic.rs HOURLY_SHARES and PERIOD_FIRST_HOUR: the three road profiles' hourly flow shares.
const double HOURLY_SHARES[PROFILES][24] = {
    {0.01309, 0.00918, 0.00785, 0.00838, 0.01246, 0.02353, 0.03973, 0.05108, 0.05364, 0.05517, 0.05799, 0.05978,
     0.06152, 0.06281, 0.06338, 0.06522, 0.06718, 0.06558, 0.05905, 0.04848, 0.03909, 0.03131, 0.02568, 0.01881},
    {0.00684, 0.00423, 0.00330, 0.00331, 0.00691, 0.02047, 0.04139, 0.05610, 0.05161, 0.05350, 0.05888, 0.06151,
     0.06501, 0.06797, 0.07131, 0.07483, 0.08065, 0.07760, 0.06261, 0.04433, 0.03251, 0.02441, 0.01922, 0.01150},
    {0.00545, 0.00336, 0.00242, 0.00236, 0.00555, 0.02012, 0.03996, 0.05863, 0.05247, 0.05348, 0.05843, 0.06219,
     0.06493, 0.06839, 0.07266, 0.07679, 0.08515, 0.08071, 0.06255, 0.04356, 0.03107, 0.02300, 0.01719, 0.00958},
};
const int FIRST_HOUR[PERIODS] = {7, 19, 23};
const int HOURS[PERIODS] = {12, 4, 8};
const int MAP_GROUPS[PERIODS] = {4, 1, 3};

double hour_factor(int profile, int period, int slot) {
    double mean = 0.0;
    for (int s = 0; s < HOURS[period]; ++s) mean += HOURLY_SHARES[profile][(FIRST_HOUR[period] + s) % 24];
    mean /= HOURS[period];
    return HOURLY_SHARES[profile][(FIRST_HOUR[period] + slot) % 24] / mean;
}

// paint-gpu/src/loudness.rs loudness_constants: the nodes' parts (AtProbabilities::new over the
// nodes, the table's last quantile and the held test's probability), the shares, the map's hour
// groups (traffic.rs hour_groups: the period's hours ordered by summed flow, cut into equal runs)
// and the profiles' mean factors per group.
LoudTables build_tables(const float* table) {
    LoudTables T;
    std::memset(&T, 0, sizeof T);
    std::vector<double> p, shares;
    for (int k = 0; k < 99; ++k) {
        p.push_back(((double)k + 0.5) / 100.0);
        shares.push_back(0.01);
    }
    auto at = [](double t) { return 1.0 - std::pow(10.0, -t); };
    for (int k = 0; k < 24; ++k) {
        const double a = 2.0 + k * 0.25, b = 2.25 + k * 0.25;
        p.push_back(at((a + b) / 2.0));
        shares.push_back(at(b) - at(a));
    }
    p.push_back(1.0 - 0.5e-8);
    shares.push_back(1e-8);
    p.push_back(probability(QUANTILES - 1));
    p.push_back(1.0 - 1e-3);
    const double bottom = probability(0), top = probability(QUANTILES - 1);
    for (int n = 0; n < ROW_NODES; ++n) {
        T.normal[n] = (float)normal_quantile(std::min(std::max(p[n], bottom), top));
        T.ln_p[n] = (float)std::log(p[n]);
        T.tail[n] = p[n] > top ? 1u : 0u;
        if (!T.tail[n]) {
            const double q = std::max(p[n], bottom);
            const int k = std::min((int)std::floor(q * QUANTILES - 0.5), QUANTILES - 2);
            const double t = (std::log(1.0 - q) - std::log(1.0 - probability(k))) /
                             (std::log(1.0 - probability(k + 1)) - std::log(1.0 - probability(k)));
            for (int row = 0; row < ROWS; ++row) {
                const double here = std::log(std::max((double)table[row * QUANTILES + k], 1e-30));
                const double next = std::log(std::max((double)table[row * QUANTILES + k + 1], 1e-30));
                T.rows[row * ROW_NODES + n] = (float)(here + t * (next - here));
            }
        }
    }
    for (int n = 0; n < NODES; ++n) T.shares[n] = (float)shares[n];
    for (int period = 0; period < PERIODS; ++period) {
        std::vector<int> slots(HOURS[period]);
        for (int s = 0; s < HOURS[period]; ++s) slots[s] = s;
        auto flow = [&](int s) {
            double f = 0.0;
            for (int pr = 0; pr < PROFILES; ++pr) f += hour_factor(pr, period, s);
            return f;
        };
        std::sort(slots.begin(), slots.end(), [&](int a, int b) {
            const double fa = flow(a), fb = flow(b);
            return fa < fb || (fa == fb && a < b);
        });
        const int runs = MAP_GROUPS[period], hours = HOURS[period];
        T.group_count[period] = (uint32_t)runs;
        for (int g = 0; g < runs; ++g) {
            const int first = g * hours / runs, end = (g + 1) * hours / runs;
            T.group_slots[period * GROUPS_MAX + g] = (uint32_t)(end - first);
            for (int pr = 0; pr < PROFILES; ++pr) {
                double f = 0.0;
                for (int k = first; k < end; ++k) f += hour_factor(pr, period, slots[k]);
                T.factor[(pr * PERIODS + period) * GROUPS_MAX + g] = (float)(f / (end - first));
            }
        }
        for (int g = runs; g < GROUPS_MAX; ++g)
            for (int pr = 0; pr < PROFILES; ++pr) T.factor[(pr * PERIODS + period) * GROUPS_MAX + g] = 1.0f;
    }
    return T;
}

// The packed pixels' shape, checked on the host: every layer's count a whole number the input
// holds, the curves within the input.
int validate(const float* input, size_t input_floats, const uint64_t* offsets, size_t pixels, char* err,
             size_t errlen) {
    for (size_t i = 0; i < pixels; ++i) {
        size_t at = offsets[i];
        for (int l = 0; l < LAYERS; ++l) {
            if (at + PERIODS + 1 > input_floats) return lfail(err, errlen, "pixel %zu: layer %d beyond the input", i, l);
            const float c = input[at + PERIODS];
            if (!(c >= 0.0f && c <= (float)LINES_MAX) || c != std::floor(c))
                return lfail(err, errlen, "pixel %zu: layer %d line count %g", i, l, (double)c);
            at += PERIODS + 1 + (size_t)c * LINE_FLOATS;
        }
        if (at + (size_t)CURVES * PERIODS * THIRDS > input_floats)
            return lfail(err, errlen, "pixel %zu: curves beyond the input", i);
    }
    return 0;
}

}  // namespace

extern "C" void qm_kurze_table(float* out) {
    const std::vector<float>& t = seeded_table();
    std::copy(t.begin(), t.end(), out);
}

extern "C" int qm_pixel_loudness(qm_gpu* gpu, const float* input, size_t input_floats, const uint64_t* offsets,
                                 size_t pixels, const float* kurze, float* mixtures, float* ndens, char* err,
                                 size_t errlen) {
    if (!gpu || (pixels && (!input || !offsets || !ndens))) return lfail(err, errlen, "null argument");
    if (pixels == 0) return 0;
    if (pixels > (size_t)UINT32_MAX / (PERIODS * GROUPS_MAX * STATES))
        return lfail(err, errlen, "%zu pixels exceed one launch", pixels);
    if (validate(input, input_floats, offsets, pixels, err, errlen)) return 1;
    const LoudTables tables = build_tables(kurze ? kurze : seeded_table().data());
    hipError_t e = hipSetDevice(gpu->ordinal);
    void *d_input = nullptr, *d_offsets = nullptr, *d_tables = nullptr, *d_mix = nullptr, *d_ndens = nullptr;
    const size_t mix_floats = pixels * PERIODS * CURVES * BINS * 2;
    auto alloc = [&](void** p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, bytes);
    };
    alloc(&d_input, sizeof(float) * input_floats);
    alloc(&d_offsets, sizeof(uint64_t) * pixels);
    alloc(&d_tables, sizeof(LoudTables));
    alloc(&d_mix, sizeof(float) * mix_floats);
    alloc(&d_ndens, sizeof(float) * pixels * CURVES);
    if (e == hipSuccess) e = hipMemcpy(d_input, input, sizeof(float) * input_floats, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_offsets, offsets, sizeof(uint64_t) * pixels, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_tables, &tables, sizeof(LoudTables), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_mix, 0, sizeof(float) * mix_floats);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mixtures_kernel, dim3((unsigned)(pixels * PERIODS * GROUPS_MAX * STATES)), dim3(64), 0, 0,
                           (const float*)d_input, (const uint64_t*)d_offsets, (const LoudTables*)d_tables,
                           (uint32_t)pixels, (float*)d_mix);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(ndens_kernel, dim3((unsigned)pixels), dim3(64), 0, 0, (const float*)d_input,
                           (const uint64_t*)d_offsets, (const float*)d_mix, (uint32_t)pixels, (float*)d_ndens);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(ndens, d_ndens, sizeof(float) * pixels * CURVES, hipMemcpyDeviceToHost);
    if (e == hipSuccess && mixtures) e = hipMemcpy(mixtures, d_mix, sizeof(float) * mix_floats, hipMemcpyDeviceToHost);
    for (void* p : {d_input, d_offsets, d_tables, d_mix, d_ndens})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return lfail(err, errlen, "pixel loudness: %s", hipGetErrorString(e));
    return 0;
}
