// TEST INFRASTRUCTURE — FP64 oracle: a painted pixel's level distributions and loudness (Nden),
// restated from the reference's CPU path. Not part of the product path.
//
//   physics/src/percentile.rs            Kurze's quantiles: relative_intensity, normal_quantile,
//                                        AtProbabilities (the same to the bit, by construction)
//   physics/src/percentile_tests.rs:1-100  the quantile table's seeded simulation (SplitMix64)
//   popup/src/distribution.rs            node_probabilities, Distribution (runs, with_line,
//                                        with_distribution, over, add_scaled, exceeded_db),
//                                        LineValues, line_values, pass_values, first_at_or_above
//   popup/src/percentiles.rs             layered_distributions, layer_states, moment
//   popup/src/percentiles/lines.rs       Line::shape and the held / holding constants
//   popup/src/percentiles/traffic.rs     hour_factor, hour_groups (HOURLY_SHARES)
//   popup/src/loudness.rs                Curve::at, loudness, day (Nden)
//   paint/src/sone.rs:28-52              the packed pixel the card reads (pack_line)
//
// A pixel arrives packed as the reference's card reads it (paint-gpu/kernels/distribution.cuh:
// per layer its steady energy per period and its line count, then each line's 43 floats; then per
// curve (the six layers', then the place's) and period the spectrum's 28 third-octave levels at
// 0 dB(A), the first NaN for none). The map's resolution: 2 dB bins, hour groups [4, 1, 3].
#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <mutex>
#include <vector>

#include "oracle.hpp"
#include "scene.hpp"

namespace qmo {
double zwicker_loudness_sone(const double in[28]);
}  // namespace qmo

namespace qmo {
namespace loud {

constexpr int PERIODS = 3;
constexpr int STATES = 10;
constexpr int LAYERS = 6;
constexpr int CURVES = LAYERS + 1;
constexpr int THIRDS = 28;
constexpr int LINE_FLOATS = PERIODS + PERIODS * STATES + PERIODS + 1 + 2 * PERIODS;
static_assert(LINE_FLOATS == 43, "packed line");
constexpr double LEVEL_MIN_DB = -30.0;
constexpr double SPAN_DB = 180.0;
constexpr double LAMBDA_MIN = 1e-4;
constexpr double STEPS_PER_DECADE = 8.0;
constexpr double SIMULATED_LAMBDA_MAX = 10.0;
constexpr int QUANTILES = 100;
constexpr int ROWS = 41;
constexpr double HOLDS_MEAN_LAMBDA = 10.0;
constexpr double PASSES_HOLD_MEAN = 2.0 * M_PI * HOLDS_MEAN_LAMBDA;
constexpr double HELD_EXCEEDED = 1e-3;
constexpr double HELD_SHARE = 1e-2;
constexpr double CURVE_LOW_DB = -30.0;
constexpr double CURVE_HIGH_DB = 160.0;
const double PERIOD_HOURS[PERIODS] = {12.0, 4.0, 8.0};
const double PERIOD_PENALTY_DB[PERIODS] = {0.0, 5.0, 10.0};
const int MAP_HOUR_GROUPS[PERIODS] = {4, 1, 3};

// ---------------------------------------------------------------------------------------------
// The quantile table (percentile_tests.rs: SplitMix64 from 0x5eed_1971, a row per lambda).

struct SplitMix {
    uint64_t state;
    double uniform() {
        state += 0x9e3779b97f4a7c15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return (double)((z ^ (z >> 31)) >> 11) / (double)(1ull << 53);
    }
};

size_t poisson(SplitMix& r, double mean) {
    if (mean > 50.0) {
        const double a = std::sqrt(-2.0 * std::log(1.0 - r.uniform()));
        const double normal = a * std::cos(2.0 * M_PI * r.uniform());
        return (size_t)std::max(std::round(mean + std::sqrt(mean) * normal), 0.0);
    }
    const double limit = std::exp(-mean);
    size_t count = 0;
    double product = r.uniform();
    while (product > limit) {
        ++count;
        product *= r.uniform();
    }
    return count;
}

std::vector<double> simulate_row(double lambda, SplitMix& r) {
    const double half = std::max(5.0 / lambda, 200.0);
    const double mean_inside = lambda * 2.0 * half;
    const double outside = lambda * 2.0 * (M_PI / 2.0 - std::atan(half));
    const double mean = lambda * M_PI;
    const size_t count = (size_t)std::min(20000.0 * std::max(0.01 / lambda, 1.0), 1000000.0);
    std::vector<double> samples(count);
    for (size_t i = 0; i < count; ++i) {
        const size_t n = poisson(r, mean_inside);
        double inside = 0.0;
        for (size_t k = 0; k < n; ++k) {
            const double x = (2.0 * r.uniform() - 1.0) * half;
            inside += 1.0 / (1.0 + x * x);
        }
        samples[i] = (inside + outside) / mean;
    }
    std::sort(samples.begin(), samples.end());
    std::vector<double> q(QUANTILES);
    for (int k = 0; k < QUANTILES; ++k) q[k] = samples[(size_t)(((double)k + 0.5) / QUANTILES * (double)count)];
    return q;
}

const std::vector<float>& kurze_table() {
    static std::vector<float> table;
    static std::once_flag once;
    std::call_once(once, [] {
        SplitMix r{0x5eed1971ull};
        table.resize(ROWS * QUANTILES);
        for (int row = 0; row < ROWS; ++row) {
            const double lambda = LAMBDA_MIN * std::pow(10.0, row / STEPS_PER_DECADE);
            const std::vector<double> v = simulate_row(lambda, r);
            for (int k = 0; k < QUANTILES; ++k) table[row * QUANTILES + k] = (float)v[k];
        }
    });
    return table;
}

// ---------------------------------------------------------------------------------------------
// percentile.rs

double probability(int k) { return ((double)k + 0.5) / QUANTILES; }

// Acklam's inverse normal (P. J. Acklam, "An algorithm for computing the inverse normal cumulative
// distribution function", 2003), as percentile.rs normal_quantile.
double normal_quantile(double p) {
    static const double A[6] = {-3.969683028665376e1, 2.209460984245205e2, -2.759285104469687e2,
                                1.38357751867269e2,   -3.066479806614716e1, 2.506628277459239};
    static const double B[5] = {-5.447609879822406e1, 1.615858368580409e2, -1.556989798598866e2,
                                6.680131188771972e1,  -1.328068155288572e1};
    static const double C[6] = {-7.784894002430293e-3, -3.223964580411365e-1, -2.400758277161838,
                                -2.549671010366478,    4.374664141464968,     2.938163982698783};
    static const double D[4] = {7.784695709041462e-3, 3.224671290700398e-1, 2.445134137142996, 3.754408661907416};
    auto tail = [&](double q) {
        return (((((C[0] * q + C[1]) * q + C[2]) * q + C[3]) * q + C[4]) * q + C[5]) /
               ((((D[0] * q + D[1]) * q + D[2]) * q + D[3]) * q + 1.0);
    };
    if (p < 0.02425) return tail(std::sqrt(-2.0 * std::log(p)));
    if (p > 1.0 - 0.02425) return -tail(std::sqrt(-2.0 * std::log(1.0 - p)));
    const double q = p - 0.5, r = q * q;
    return (((((A[0] * r + A[1]) * r + A[2]) * r + A[3]) * r + A[4]) * r + A[5]) * q /
           (((((B[0] * r + B[1]) * r + B[2]) * r + B[3]) * r + B[4]) * r + 1.0);
}

double passing_intensity(double lambda, double ln_p) {
    const double s = -ln_p / (2.0 * lambda);
    return 1.0 / (lambda * M_PI * (1.0 + s * s));
}

double normal_intensity(double lambda, double normal) {
    const double deviation = 1.0 / std::sqrt(2.0 * M_PI * lambda);
    return std::max(1.0 + deviation * normal, 0.0);
}

void table_position(double lambda, int& row, double& fraction) {
    const double position = std::min(std::max(std::log10(lambda / LAMBDA_MIN) * STEPS_PER_DECADE, 0.0), (double)(ROWS - 1));
    row = (int)std::floor(position);
    fraction = position - std::trunc(position);
}

void quantile_position(double p, int& k, double& t) {
    p = std::max(p, probability(0));
    k = std::min((int)std::floor(p * QUANTILES - 0.5), QUANTILES - 2);
    t = (std::log(1.0 - p) - std::log(1.0 - probability(k))) /
        (std::log(1.0 - probability(k + 1)) - std::log(1.0 - probability(k)));
}

double log_quantile(const float* table, int row, int k, double t) {
    const float* values = table + row * QUANTILES;
    const double here = std::log(std::max((double)values[k], 1e-30));
    const double next = std::log(std::max((double)values[k + 1], 1e-30));
    return here + t * (next - here);
}

double between_rows(double low, double high, double fraction) { return std::exp(low + fraction * (high - low)); }

double relative_intensity(const float* table, double lambda, double p) {
    if (!std::isfinite(lambda)) return 1.0;
    if (lambda <= 0.0) return 0.0;
    const double bottom = probability(0), top = probability(QUANTILES - 1);
    if (lambda >= SIMULATED_LAMBDA_MAX) return normal_intensity(lambda, normal_quantile(std::min(std::max(p, bottom), top)));
    if (p > top)
        return relative_intensity(table, lambda, top) + passing_intensity(lambda, std::log(p)) -
               passing_intensity(lambda, std::log(top));
    if (lambda < LAMBDA_MIN) return lambda / LAMBDA_MIN * relative_intensity(table, LAMBDA_MIN, p);
    int row, k;
    double fraction, t;
    table_position(lambda, row, fraction);
    quantile_position(p, k, t);
    return between_rows(log_quantile(table, row, k, t), log_quantile(table, std::min(row + 1, ROWS - 1), k, t), fraction);
}

// ---------------------------------------------------------------------------------------------
// distribution.rs

std::vector<std::pair<double, double>> node_probabilities() {
    std::vector<std::pair<double, double>> nodes;
    for (int k = 0; k < 99; ++k) nodes.push_back({((double)k + 0.5) / 100.0, 0.01});
    auto at = [](double t) { return 1.0 - std::pow(10.0, -t); };
    for (int k = 0; k < 24; ++k) {
        const double a = 2.0 + k * 0.25, b = 2.25 + k * 0.25;
        nodes.push_back({at((a + b) / 2.0), at(b) - at(a)});
    }
    nodes.push_back({1.0 - 0.5e-8, 1e-8});
    return nodes;
}

struct LineValues {
    std::vector<std::pair<double, double>> pairs;  // (value, share), ascending
    std::vector<std::array<double, 2>> sums;       // sums[k]: of the first k, shares and shares x values

    static LineValues of(std::vector<std::pair<double, double>> p) {
        std::stable_sort(p.begin(), p.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        LineValues v;
        v.pairs = std::move(p);
        v.sums.push_back({0.0, 0.0});
        for (const auto& [value, share] : v.pairs) {
            const auto last = v.sums.back();
            v.sums.push_back({last[0] + share, last[1] + share * value});
        }
        return v;
    }
};

LineValues line_values(const float* table, double lambda, const double* duty) {
    if (duty) return LineValues::of({{0.0, 1.0 - *duty}, {1.0 / *duty, *duty}});
    const auto nodes = node_probabilities();
    std::vector<std::pair<double, double>> p;
    for (const auto& [prob, share] : nodes) p.push_back({relative_intensity(table, lambda, prob), share});
    return LineValues::of(p);
}

LineValues pass_values(double mean) {
    double share = std::exp(-mean);
    std::vector<std::pair<double, double>> pairs{{0.0, share}};
    double time_left = -std::expm1(-mean), mean_left = 1.0, count = 0.0;
    for (;;) {
        count += 1.0;
        share *= mean / count;
        if (time_left - share <= 1e-12) {
            pairs.push_back({mean_left / time_left, time_left});
            return LineValues::of(pairs);
        }
        time_left -= share;
        mean_left -= share * count / mean;
        pairs.push_back({count / mean, share});
    }
}

size_t first_at_or_above(const std::vector<std::pair<double, double>>& pairs, size_t from, double limit) {
    size_t low = from, step = 1;
    while (low + step - 1 < pairs.size() && pairs[low + step - 1].first < limit) {
        low += step;
        step *= 2;
    }
    const size_t high = std::min(low + step - 1, pairs.size());
    size_t k = low;
    while (k < high && pairs[k].first < limit) ++k;
    return k;
}

struct Distribution {
    double silent = 0.0;
    double bin_db = 2.0;
    std::vector<std::array<double, 2>> bins;

    static Distribution empty(double bin_db) {
        Distribution d;
        d.bin_db = bin_db;
        d.bins.assign((size_t)std::round(SPAN_DB / bin_db), {0.0, 0.0});
        return d;
    }
    size_t bin(double level_db) const {
        const double r = std::round((level_db - LEVEL_MIN_DB) / bin_db);
        size_t b = 0;
        if (r > 0.0) b = r >= (double)bins.size() ? bins.size() - 1 : (size_t)r;  // a saturating cast
        return std::min(b, bins.size() - 1);
    }
    template <class F>
    void intensities(F&& f) const {
        for (const auto& b : bins)
            if (b[0] > 0.0) f(b[1] / b[0], b[0]);
    }
    // visit(bin or -1 for silence, shares, weighted)
    template <class V>
    void runs(double intensity, const LineValues& line, double scale, V&& visit) const {
        const auto& pairs = line.pairs;
        const auto& sums = line.sums;
        auto run = [&](size_t from, size_t to, double out[2]) {
            out[0] = sums[to][0] - sums[from][0];
            out[1] = scale * (sums[to][1] - sums[from][1]);
        };
        size_t from = 0;
        double r[2];
        if (intensity <= 0.0) {
            while (from < pairs.size() && pairs[from].first <= 0.0) ++from;
            if (from > 0) {
                run(0, from, r);
                visit(-1, r[0], r[1]);
            }
        }
        if (from == pairs.size()) return;
        const size_t last = bins.size() - 1;
        const double step = energy(bin_db);
        size_t b = bin(10.0 * std::log10(intensity + scale * pairs[from].first));
        double ceiling = energy(LEVEL_MIN_DB + ((double)b + 0.5) * bin_db);
        for (;;) {
            const size_t to = b == last ? pairs.size() : first_at_or_above(pairs, from + 1, (ceiling - intensity) / scale);
            run(from, to, r);
            visit((long)b, r[0], r[1]);
            from = to;
            if (from >= pairs.size()) break;
            const double value = pairs[from].first;
            for (;;) {
                ++b;
                ceiling *= step;
                if (b == last || intensity + scale * value < ceiling) break;
            }
        }
    }
    Distribution with_line(const LineValues& line, double scale) const {
        Distribution sum = empty(bin_db);
        auto add = [&](double intensity, double share) {
            runs(intensity, line, scale, [&](long b, double shares, double weighted) {
                if (b >= 0) {
                    sum.bins[b][0] += share * shares;
                    sum.bins[b][1] += share * (intensity * shares + weighted);
                } else {
                    sum.silent += share * shares;
                }
            });
        };
        if (silent > 0.0) add(0.0, silent);
        intensities([&](double i, double s) { add(i, s); });
        return sum;
    }
    Distribution with_distribution(const Distribution& other) const {
        std::vector<std::pair<double, double>> v;
        other.intensities([&](double i, double s) { v.push_back({i, s}); });
        if (other.silent > 0.0) v.push_back({0.0, other.silent});
        return with_line(LineValues::of(v), 1.0);
    }
    Distribution over(double intensity) const { return with_line(LineValues::of({{intensity, 1.0}}), 1.0); }
    void add_scaled(const Distribution& o, double w) {
        silent += o.silent * w;
        for (size_t k = 0; k < bins.size(); ++k) {
            bins[k][0] += o.bins[k][0] * w;
            bins[k][1] += o.bins[k][1] * w;
        }
    }
    double exceeded_db(double exceeded) const {
        double above = 0.0;
        for (size_t k = bins.size(); k-- > 0;) {
            above += bins[k][0];
            if (bins[k][0] > 0.0 && above > exceeded * (1.0 - 1e-12)) return 10.0 * std::log10(bins[k][1] / bins[k][0]);
        }
        return -std::numeric_limits<double>::infinity();
    }
};

// ---------------------------------------------------------------------------------------------
// traffic.rs: hourly flow shares of the three road profiles and the periods' first hours.

const double HOURLY_SHARES[3][24] = {
    {0.01309, 0.00918, 0.00785, 0.00838, 0.01246, 0.02353, 0.03973, 0.05108, 0.05364, 0.05517, 0.05799, 0.05978,
     0.06152, 0.06281, 0.06338, 0.06522, 0.06718, 0.06558, 0.05905, 0.04848, 0.03909, 0.03131, 0.02568, 0.01881},
    {0.00684, 0.00423, 0.00330, 0.00331, 0.00691, 0.02047, 0.04139, 0.05610, 0.05161, 0.05350, 0.05888, 0.06151,
     0.06501, 0.06797, 0.07131, 0.07483, 0.08065, 0.07760, 0.06261, 0.04433, 0.03251, 0.02441, 0.01922, 0.01150},
    {0.00545, 0.00336, 0.00242, 0.00236, 0.00555, 0.02012, 0.03996, 0.05863, 0.05247, 0.05348, 0.05843, 0.06219,
     0.06493, 0.06839, 0.07266, 0.07679, 0.08515, 0.08071, 0.06255, 0.04356, 0.03107, 0.02300, 0.01719, 0.00958},
};
const int PERIOD_FIRST_HOUR[PERIODS] = {7, 19, 23};
constexpr int PROFILES = 3;

double hour_factor(int profile, int period, int slot) {
    const int hours = (int)PERIOD_HOURS[period];
    auto share = [&](int s) { return HOURLY_SHARES[profile][(PERIOD_FIRST_HOUR[period] + s) % 24]; };
    double mean = 0.0;
    for (int s = 0; s < hours; ++s) mean += share(s);
    mean /= hours;
    return share(slot) / mean;
}

// hour_groups(profiled, period, Resolution::MAP.hour_groups): the slots of each group and the hours.
std::vector<std::vector<int>> hour_groups(bool profiled, int period, int& hours) {
    if (!profiled) {
        hours = 1;
        return {{0}};
    }
    hours = (int)PERIOD_HOURS[period];
    auto flow = [&](int slot) {
        double f = 0.0;
        for (int p = 0; p < PROFILES; ++p) f += hour_factor(p, period, slot);
        return f;
    };
    std::vector<int> slots(hours);
    for (int s = 0; s < hours; ++s) slots[s] = s;
    std::sort(slots.begin(), slots.end(), [&](int a, int b) {
        const double fa = flow(a), fb = flow(b);
        return fa < fb || (fa == fb && a < b);
    });
    const int runs = MAP_HOUR_GROUPS[period];
    std::vector<std::vector<int>> cut;
    for (int run = 0; run < runs; ++run) {
        std::vector<int> g(slots.begin() + run * hours / runs, slots.begin() + (run + 1) * hours / runs);
        std::sort(g.begin(), g.end());
        cut.push_back(g);
    }
    return cut;
}

// ---------------------------------------------------------------------------------------------
// A packed line (sone.rs pack_line) and percentiles.rs.

struct Line {
    double energy[PERIODS];
    double means[PERIODS][STATES];
    double lambda[PERIODS];
    int profile;  // -1 none
    bool has_duty, has_passing;
    double duty[PERIODS], passing[PERIODS];

    static Line unpack(const float* f) {
        Line l;
        for (int p = 0; p < PERIODS; ++p) l.energy[p] = f[p];
        for (int p = 0; p < PERIODS; ++p)
            for (int s = 0; s < STATES; ++s) l.means[p][s] = f[PERIODS + p * STATES + s];
        for (int p = 0; p < PERIODS; ++p) l.lambda[p] = f[PERIODS + PERIODS * STATES + p];
        l.profile = (int)f[2 * PERIODS + PERIODS * STATES];
        l.has_duty = !std::isnan(f[2 * PERIODS + PERIODS * STATES + 1]);
        l.has_passing = !std::isnan(f[3 * PERIODS + PERIODS * STATES + 1]);
        for (int p = 0; p < PERIODS; ++p) {
            l.duty[p] = f[2 * PERIODS + PERIODS * STATES + 1 + p];
            l.passing[p] = f[3 * PERIODS + PERIODS * STATES + 1 + p];
        }
        return l;
    }
};

struct Shape {
    bool some = false;
    LineValues values;
    bool has_exceeded = false;
    double exceeded = 0.0;
};

// lines.rs Line::shape.
Shape shape_of(const float* table, const Line& line, int period, double factor) {
    Shape s;
    if (line.has_passing) {
        const double mean = line.passing[period];
        if (mean >= std::numeric_limits<double>::min() && mean < PASSES_HOLD_MEAN) {
            s.some = true;
            s.values = pass_values(mean);
        }
        return s;
    }
    const double lambda = line.lambda[period] * factor;
    if (line.has_duty) {
        const double duty = line.duty[period];
        if (duty > 0.0) {
            s.some = true;
            s.values = line_values(table, 0.0, &duty);
        }
        return s;
    }
    if (lambda >= HOLDS_MEAN_LAMBDA) return s;
    s.some = true;
    s.values = line_values(table, lambda, nullptr);
    s.has_exceeded = true;
    s.exceeded = relative_intensity(table, lambda, 1.0 - HELD_EXCEEDED);
    return s;
}

struct StateResult {
    Distribution d;
    bool steady;
    double floor;
};

// percentiles.rs layer_states over moment().
std::vector<StateResult> layer_states(const float* table, const std::vector<Line>& lines, double steady, int period,
                                      const std::vector<int>& slots, double bin_db) {
    std::vector<double> factors(lines.size());
    for (size_t i = 0; i < lines.size(); ++i) {
        if (lines[i].profile < 0) {
            factors[i] = 1.0;
            continue;
        }
        double f = 0.0;
        for (int s : slots) f += hour_factor(lines[i].profile, period, s);
        factors[i] = f / (double)slots.size();
    }
    std::vector<Shape> shapes(lines.size());
    for (size_t i = 0; i < lines.size(); ++i) shapes[i] = shape_of(table, lines[i], period, factors[i]);
    std::vector<StateResult> states;
    for (int state = 0; state < STATES; ++state) {
        if (state > 0) {
            bool same = true;
            for (const Line& l : lines) same = same && l.means[period][state] == l.means[period][state - 1];
            if (same) {
                states.push_back(states.back());
                continue;
            }
        }
        double floor = steady;
        std::vector<std::pair<size_t, double>> moving;  // (line, mean)
        struct Held {
            size_t i;
            double mean, exceeded;
        };
        std::vector<Held> held;
        for (size_t i = 0; i < lines.size(); ++i) {
            const double mean = lines[i].means[period][state] * factors[i];
            if (lines[i].energy[period] <= 0.0 || mean <= 0.0) continue;
            if (shapes[i].some && !shapes[i].has_exceeded) {
                moving.push_back({i, mean});
            } else if (shapes[i].some) {
                held.push_back({i, mean, shapes[i].exceeded});
            } else if (lines[i].has_duty) {
            } else {
                floor += mean;
            }
        }
        double steady_energy = floor;
        for (const Held& h : held) steady_energy += h.mean;
        for (const Held& h : held) {
            if (h.mean * h.exceeded < HELD_SHARE * steady_energy) floor += h.mean;
            else moving.push_back({h.i, h.mean});
        }
        Distribution sum = Distribution::empty(bin_db);
        sum.silent = 1.0;
        sum = sum.over(floor);
        for (const auto& [i, mean] : moving) sum = sum.with_line(shapes[i].values, mean);
        states.push_back({sum, moving.empty(), floor});
    }
    return states;
}

struct PixelResult {
    Distribution each[LAYERS][PERIODS];
    Distribution whole[PERIODS];
};

// percentiles.rs layered_distributions at the map's resolution.
PixelResult layered_distributions(const float* table, const std::vector<std::vector<Line>>& layers,
                                  const double steady[LAYERS][PERIODS], double bin_db) {
    bool profiled = false;
    std::vector<bool> follows(LAYERS, false);
    for (int l = 0; l < LAYERS; ++l) {
        for (const Line& line : layers[l]) follows[l] = follows[l] || line.profile >= 0;
        profiled = profiled || follows[l];
    }
    PixelResult out;
    for (int p = 0; p < PERIODS; ++p) {
        int hours = 1;
        const auto groups = hour_groups(profiled, p, hours);
        std::vector<std::vector<StateResult>> alike(LAYERS);
        for (int l = 0; l < LAYERS; ++l)
            if (!follows[l]) alike[l] = layer_states(table, layers[l], steady[l][p], p, groups[0], bin_db);
        for (int l = 0; l < LAYERS; ++l) {
            out.each[l][p] = Distribution::empty(bin_db);
            if (!follows[l])
                for (const StateResult& s : alike[l]) out.each[l][p].add_scaled(s.d, 1.0 / STATES);
        }
        out.whole[p] = Distribution::empty(bin_db);
        for (const auto& slots : groups) {
            std::vector<std::vector<StateResult>> own(LAYERS);
            for (int l = 0; l < LAYERS; ++l)
                if (follows[l]) own[l] = layer_states(table, layers[l], steady[l][p], p, slots, bin_db);
            const double weight = (double)slots.size() / (double)(hours * STATES);
            for (int l = 0; l < LAYERS; ++l)
                if (follows[l])
                    for (const StateResult& s : own[l]) out.each[l][p].add_scaled(s.d, weight);
            for (int state = 0; state < STATES; ++state) {
                double floor = 0.0;
                for (int l = 0; l < LAYERS; ++l) {
                    const StateResult& s = follows[l] ? own[l][state] : alike[l][state];
                    if (s.steady) floor += s.floor;
                }
                Distribution whole = Distribution::empty(bin_db);
                whole.silent = 1.0;
                whole = whole.over(floor);
                for (int l = 0; l < LAYERS; ++l) {
                    const StateResult& s = follows[l] ? own[l][state] : alike[l][state];
                    if (!s.steady) whole = whole.with_distribution(s.d);
                }
                out.whole[p].add_scaled(whole, weight);
            }
        }
    }
    return out;
}

// loudness.rs Curve::at over the whole decibels' loudness.
struct Curve {
    bool some = false;
    double shape[THIRDS];
    std::vector<double> values;
    std::vector<bool> known;
    double value(int step) {
        if (!known[step]) {
            double levels[THIRDS];
            for (int k = 0; k < THIRDS; ++k) levels[k] = shape[k] + (CURVE_LOW_DB + step);
            values[step] = qmo::zwicker_loudness_sone(levels);
            known[step] = true;
        }
        return values[step];
    }
    double at(double level_db) {
        if (!some) return 0.0;
        const int last = (int)(CURVE_HIGH_DB - CURVE_LOW_DB);
        const double position = std::min(std::max(level_db - CURVE_LOW_DB, 0.0), (double)last);
        const int below = std::min((int)std::floor(position), last - 1);
        const double low = value(below), high = value(below + 1);
        return low + (high - low) * (position - below);
    }
    static Curve of(const float* f) {
        Curve c;
        c.some = !std::isnan(f[0]);
        for (int k = 0; k < THIRDS; ++k) c.shape[k] = f[k];
        const int steps = (int)(CURVE_HIGH_DB - CURVE_LOW_DB) + 1;
        c.values.assign(steps, 0.0);
        c.known.assign(steps, false);
        return c;
    }
};

double nden(const Distribution* per_period, Curve* curves) {
    double total = 0.0;
    for (int p = 0; p < PERIODS; ++p) {
        double mean = 0.0;
        per_period[p].intensities([&](double intensity, double share) {
            mean += share * curves[p].at(10.0 * std::log10(intensity) + PERIOD_PENALTY_DB[p]);
        });
        total += PERIOD_HOURS[p] * mean;
    }
    return total / 24.0;
}

// ---------------------------------------------------------------------------------------------
// paint/src/sone_fold.rs fold_placed_for_map and popup/src/percentiles/weather.rs
// merge_close_states, over lines in f64 (popup/src/percentiles/lines.rs Line, Line::steady), then
// paint/src/sone.rs write_packed / pack_line into the card's f32 layout.

constexpr double DUTY_FOLD_PEAK_SHARE = 0.12;
constexpr double DUTY_FOLD_STD_SHARE = 0.08;
constexpr double KURZE_FOLD_HELD_SHARE = 0.05;
constexpr int MAX_GROUPS = 4;

struct FLine {
    double energy[PERIODS];
    double means[PERIODS][STATES];
    double lambda[PERIODS];
    int profile;  // -1 none
    bool has_duty, has_passing;
    double duty[PERIODS], passing[PERIODS];

    static FLine unpack(const double* f) {
        FLine l;
        for (int p = 0; p < PERIODS; ++p) l.energy[p] = f[p];
        for (int p = 0; p < PERIODS; ++p)
            for (int s = 0; s < STATES; ++s) l.means[p][s] = f[PERIODS + p * STATES + s];
        for (int p = 0; p < PERIODS; ++p) l.lambda[p] = f[PERIODS + PERIODS * STATES + p];
        l.profile = (int)f[2 * PERIODS + PERIODS * STATES];
        l.has_duty = !std::isnan(f[2 * PERIODS + PERIODS * STATES + 1]);
        l.has_passing = !std::isnan(f[3 * PERIODS + PERIODS * STATES + 1]);
        for (int p = 0; p < PERIODS; ++p) {
            l.duty[p] = f[2 * PERIODS + PERIODS * STATES + 1 + p];
            l.passing[p] = f[3 * PERIODS + PERIODS * STATES + 1 + p];
        }
        return l;
    }
    // Line::steady
    static FLine steady(const double means[PERIODS][STATES], int profile) {
        FLine l;
        for (int p = 0; p < PERIODS; ++p) {
            bool any = false;
            for (int s = 0; s < STATES; ++s) {
                l.means[p][s] = means[p][s];
                any = any || means[p][s] > 0.0;
            }
            l.energy[p] = any ? 1.0 : 0.0;
            l.lambda[p] = std::numeric_limits<double>::infinity();
            l.duty[p] = l.passing[p] = std::nan("");
        }
        l.profile = profile;
        l.has_duty = l.has_passing = false;
        return l;
    }
    template <typename V>
    void pack(V* out) const {
        for (int p = 0; p < PERIODS; ++p) out[p] = (V)energy[p];
        for (int p = 0; p < PERIODS; ++p)
            for (int s = 0; s < STATES; ++s) out[PERIODS + p * STATES + s] = (V)means[p][s];
        for (int p = 0; p < PERIODS; ++p) out[PERIODS + PERIODS * STATES + p] = (V)lambda[p];
        out[2 * PERIODS + PERIODS * STATES] = (V)profile;
        for (int p = 0; p < PERIODS; ++p) {
            out[2 * PERIODS + PERIODS * STATES + 1 + p] = has_duty ? (V)duty[p] : (V)std::nan("");
            out[3 * PERIODS + PERIODS * STATES + 1 + p] = has_passing ? (V)passing[p] : (V)std::nan("");
        }
    }
};

struct FLayer {
    std::vector<FLine> lines;
    double steady[PERIODS];
};

// The map's hour groups (count per period) and each profile's mean factor per group.
struct MapHours {
    int groups[PERIODS];
    double factor[PROFILES][PERIODS][MAX_GROUPS];
};
const MapHours& map_hours() {
    static MapHours h;
    static std::once_flag once;
    std::call_once(once, [] {
        for (int p = 0; p < PERIODS; ++p) {
            int hours = 0;
            const auto g = hour_groups(true, p, hours);
            h.groups[p] = (int)g.size();
            for (int profile = 0; profile < PROFILES; ++profile)
                for (int k = 0; k < MAX_GROUPS; ++k) {
                    double f = 1.0;
                    if (k < (int)g.size()) {
                        f = 0.0;
                        for (int slot : g[k]) f += hour_factor(profile, p, slot);
                        f /= (double)g[k].size();
                    }
                    h.factor[profile][p][k] = f;
                }
        }
    });
    return h;
}

void merge_close_states(std::vector<FLayer>& layers, double spread) {
    auto within = [&](const double* means, int n) {
        double low = std::numeric_limits<double>::infinity(), high = 0.0;
        for (int k = 0; k < n; ++k) {
            if (!std::isfinite(means[k])) return false;
            low = std::min(low, means[k]);
            high = std::max(high, means[k]);
        }
        return high <= spread * low;
    };
    for (int p = 0; p < PERIODS; ++p) {
        int start = 0;
        while (start < STATES) {
            int end = start + 1;
            for (;;) {
                if (end >= STATES) break;
                bool all = true;
                for (const FLayer& layer : layers)
                    for (const FLine& line : layer.lines) all = all && within(&line.means[p][start], end - start + 1);
                if (!all) break;
                ++end;
            }
            if (end - start > 1)
                for (FLayer& layer : layers)
                    for (FLine& line : layer.lines) {
                        double* run = &line.means[p][start];
                        bool differs = false;
                        for (int k = 1; k < end - start; ++k) differs = differs || run[k] != run[0];
                        if (differs) {
                            double sum = 0.0;
                            for (int k = 0; k < end - start; ++k) sum += run[k];
                            const double mean = sum / (double)(end - start);
                            for (int k = 0; k < end - start; ++k) run[k] = mean;
                        }
                    }
            start = end;
        }
    }
}

void fold_placed_for_map(std::vector<FLayer>& placed, double spread) {
    const MapHours& hours = map_hours();
    const float* table = kurze_table().data();
    bool any_profiled = false;
    for (const FLayer& layer : placed)
        for (const FLine& line : layer.lines) any_profiled = any_profiled || line.profile >= 0;
    for (FLayer& layer : placed) {
        std::vector<FLine>& lines = layer.lines;
        if (lines.empty()) continue;
        bool has_kurze = false, has_duty = false;
        for (const FLine& line : lines) {
            if (line.has_duty) {
                has_duty = true;
            } else if (!line.has_passing) {
                for (int p = 0; p < PERIODS; ++p) has_kurze = has_kurze || std::isfinite(line.lambda[p]);
            }
        }
        double held[PERIODS][MAX_GROUPS][STATES] = {};
        if (has_kurze) {
            for (int p = 0; p < PERIODS; ++p) {
                const int ng = any_profiled ? hours.groups[p] : 1;
                for (int g = 0; g < ng; ++g)
                    for (int s = 0; s < STATES; ++s) held[p][g][s] = layer.steady[p];
            }
            for (const FLine& line : lines) {
                if (line.has_duty || line.has_passing) continue;
                for (int p = 0; p < PERIODS; ++p) {
                    if (line.energy[p] <= 0.0) continue;
                    const int ng = any_profiled ? hours.groups[p] : 1;
                    for (int g = 0; g < ng; ++g) {
                        const double factor = any_profiled && line.profile >= 0 ? hours.factor[line.profile][p][g] : 1.0;
                        for (int s = 0; s < STATES; ++s) {
                            const double m = line.means[p][s] * factor;
                            if (m > 0.0) held[p][g][s] += m;
                        }
                    }
                }
            }
        }
        double layer_total[PERIODS][STATES] = {};
        if (has_duty) {
            for (int p = 0; p < PERIODS; ++p)
                for (int s = 0; s < STATES; ++s) layer_total[p][s] = layer.steady[p];
            for (const FLine& line : lines)
                for (int p = 0; p < PERIODS; ++p) {
                    if (line.energy[p] <= 0.0) continue;
                    for (int s = 0; s < STATES; ++s)
                        if (line.means[p][s] > 0.0) layer_total[p][s] += line.means[p][s];
                }
        }
        double unprof[PERIODS][STATES] = {};
        bool has_unprof = false;
        double prof[PROFILES][PERIODS][STATES] = {};
        bool has_prof[PROFILES] = {false, false, false};
        std::vector<bool> keep(lines.size(), false);
        std::vector<std::pair<double, size_t>> duty_candidates;
        for (size_t idx = 0; idx < lines.size(); ++idx) {
            const FLine& line = lines[idx];
            if (line.has_duty) {
                bool moves = false, eligible = true;
                double max_var_ratio = 0.0;
                for (int p = 0; p < PERIODS; ++p) {
                    if (line.energy[p] <= 0.0 || line.duty[p] <= 0.0) continue;
                    const double d = line.duty[p];
                    const double var_scale = std::max(1.0 - d, 0.0) / d;
                    for (int s = 0; s < STATES; ++s) {
                        const double mean = line.means[p][s];
                        if (mean <= 0.0) continue;
                        moves = true;
                        const double total = layer_total[p][s];
                        if (total <= 0.0 || mean / d > DUTY_FOLD_PEAK_SHARE * total) {
                            eligible = false;
                        } else {
                            const double r = mean / total;
                            max_var_ratio = std::max(max_var_ratio, r * r * var_scale);
                        }
                    }
                }
                if (moves) {
                    if (eligible) duty_candidates.push_back({max_var_ratio, idx});
                    else keep[idx] = true;
                }
                continue;
            }
            if (line.has_passing) {
                bool moves = false;
                for (int p = 0; p < PERIODS; ++p) {
                    bool any = false;
                    for (int s = 0; s < STATES; ++s) any = any || line.means[p][s] > 0.0;
                    moves = moves || (line.energy[p] > 0.0 && line.passing[p] > 0.0 &&
                                      line.passing[p] < PASSES_HOLD_MEAN && any);
                }
                if (moves) {
                    keep[idx] = true;
                } else {
                    for (int p = 0; p < PERIODS; ++p) {
                        if (line.energy[p] <= 0.0) continue;
                        for (int s = 0; s < STATES; ++s)
                            if (line.means[p][s] > 0.0) {
                                unprof[p][s] += line.means[p][s];
                                has_unprof = true;
                            }
                    }
                }
                continue;
            }
            bool always_held = true;
            for (int p = 0; p < PERIODS && always_held; ++p) {
                if (line.energy[p] <= 0.0) continue;
                const int ng = any_profiled ? hours.groups[p] : 1;
                for (int g = 0; g < ng && always_held; ++g) {
                    const double factor = any_profiled && line.profile >= 0 ? hours.factor[line.profile][p][g] : 1.0;
                    const double lam = line.lambda[p] * factor;
                    if (lam >= HOLDS_MEAN_LAMBDA) continue;
                    const double exc = relative_intensity(table, lam, 1.0 - HELD_EXCEEDED);
                    for (int s = 0; s < STATES; ++s) {
                        const double m = line.means[p][s] * factor;
                        if (m > 0.0 && (m * exc >= KURZE_FOLD_HELD_SHARE * held[p][g][s] || !std::isfinite(m * exc))) {
                            always_held = false;
                            break;
                        }
                    }
                }
            }
            if (always_held) {
                for (int p = 0; p < PERIODS; ++p) {
                    if (line.energy[p] <= 0.0) continue;
                    double* target;
                    if (line.profile >= 0) {
                        has_prof[line.profile] = true;
                        target = prof[line.profile][p];
                    } else {
                        has_unprof = true;
                        target = unprof[p];
                    }
                    for (int s = 0; s < STATES; ++s)
                        if (line.means[p][s] > 0.0) target[s] += line.means[p][s];
                }
            } else {
                keep[idx] = true;
            }
        }
        if (!duty_candidates.empty()) {
            std::sort(duty_candidates.begin(), duty_candidates.end(), [](const auto& a, const auto& b) {
                return a.first < b.first || (a.first == b.first && a.second < b.second);
            });
            double folded_var[PERIODS][STATES] = {};
            const double budget_sq = DUTY_FOLD_STD_SHARE * DUTY_FOLD_STD_SHARE;
            for (const auto& candidate : duty_candidates) {
                const FLine& line = lines[candidate.second];
                bool fits = true;
                for (int p = 0; p < PERIODS && fits; ++p) {
                    if (line.energy[p] <= 0.0 || line.duty[p] <= 0.0) continue;
                    const double var_scale = std::max(1.0 - line.duty[p], 0.0) / line.duty[p];
                    for (int s = 0; s < STATES; ++s) {
                        const double m = line.means[p][s];
                        if (m <= 0.0) continue;
                        const double total = layer_total[p][s];
                        if (!(folded_var[p][s] + m * m * var_scale <= budget_sq * total * total)) {
                            fits = false;
                            break;
                        }
                    }
                }
                if (fits) {
                    for (int p = 0; p < PERIODS; ++p) {
                        if (line.energy[p] <= 0.0 || line.duty[p] <= 0.0) continue;
                        const double var_scale = std::max(1.0 - line.duty[p], 0.0) / line.duty[p];
                        for (int s = 0; s < STATES; ++s) {
                            const double m = line.means[p][s];
                            if (m > 0.0) {
                                folded_var[p][s] += m * m * var_scale;
                                unprof[p][s] += m;
                                has_unprof = true;
                            }
                        }
                    }
                } else {
                    keep[candidate.second] = true;
                }
            }
        }
        std::vector<FLine> folded;
        if (has_unprof) folded.push_back(FLine::steady(unprof, -1));
        for (int k = 0; k < PROFILES; ++k)
            if (has_prof[k]) folded.push_back(FLine::steady(prof[k], k));
        for (size_t idx = 0; idx < lines.size(); ++idx)
            if (keep[idx]) folded.push_back(lines[idx]);
        lines = folded;
    }
    merge_close_states(placed, spread);
    for (FLayer& layer : placed) {
        if (layer.lines.empty()) continue;
        const FLine& first = layer.lines.front();
        bool collapse = first.profile < 0 && !first.has_duty && !first.has_passing;
        for (int p = 0; p < PERIODS; ++p) collapse = collapse && !std::isfinite(first.lambda[p]);
        for (int p = 0; p < PERIODS; ++p)
            for (int s = 0; s < STATES; ++s) collapse = collapse && (float)first.means[p][s] == (float)first.means[p][0];
        if (collapse) {
            for (int p = 0; p < PERIODS; ++p) {
                double sum = 0.0;
                for (int s = 0; s < STATES; ++s) sum += first.means[p][s];
                layer.steady[p] += sum / (double)STATES;
            }
            layer.lines.erase(layer.lines.begin());
        }
    }
}

}  // namespace loud
}  // namespace qmo

using namespace qmo::loud;

extern "C" {

// The map's folding of unfolded pixels (the packed layout, every value f64) into the card's packed
// f32 pixels at the same offsets, and when `out64` is given the same in f64; `spread` the weather
// merge's (10^0.15 for the map). Returns the floats each folded pixel takes in `used` (at most its
// unfolded count).
void qmo_fold_pixels(const double* input, const uint64_t* offsets, size_t pixels, double spread, float* out,
                     double* out64, uint64_t* used) {
    for (size_t i = 0; i < pixels; ++i) {
        const double* in = input + offsets[i];
        std::vector<FLayer> placed(LAYERS);
        size_t at = 0;
        for (int l = 0; l < LAYERS; ++l) {
            for (int p = 0; p < PERIODS; ++p) placed[l].steady[p] = in[at + p];
            const int count = (int)in[at + PERIODS];
            at += PERIODS + 1;
            for (int k = 0; k < count; ++k) placed[l].lines.push_back(FLine::unpack(in + at + (size_t)k * LINE_FLOATS));
            at += (size_t)count * LINE_FLOATS;
        }
        fold_placed_for_map(placed, spread);
        float* o = out + offsets[i];
        double* o64 = out64 ? out64 + offsets[i] : nullptr;
        size_t w = 0;
        for (int l = 0; l < LAYERS; ++l) {
            for (int p = 0; p < PERIODS; ++p) o[w + p] = (float)placed[l].steady[p];
            o[w + PERIODS] = (float)placed[l].lines.size();
            if (o64) {
                for (int p = 0; p < PERIODS; ++p) o64[w + p] = placed[l].steady[p];
                o64[w + PERIODS] = (double)placed[l].lines.size();
            }
            w += PERIODS + 1;
            for (const FLine& line : placed[l].lines) {
                line.pack(o + w);
                if (o64) line.pack(o64 + w);
                w += LINE_FLOATS;
            }
        }
        for (size_t k = 0; k < (size_t)CURVES * PERIODS * THIRDS; ++k) {
            o[w + k] = (float)in[at + k];
            if (o64) o64[w + k] = in[at + k];
        }
        used[i] = w + (size_t)CURVES * PERIODS * THIRDS;
    }
}

// merge_close_states alone over `layers` layers of packed f64 lines (43 values each, layer l
// holding counts[l] of them one after another), in place.
void qmo_merge_close_states(double* lines, const uint32_t* counts, size_t layers, double spread) {
    std::vector<FLayer> placed(layers);
    size_t at = 0;
    for (size_t l = 0; l < layers; ++l) {
        for (int p = 0; p < PERIODS; ++p) placed[l].steady[p] = 0.0;
        for (uint32_t k = 0; k < counts[l]; ++k) placed[l].lines.push_back(FLine::unpack(lines + (at + k) * LINE_FLOATS));
        at += counts[l];
    }
    merge_close_states(placed, spread);
    at = 0;
    for (size_t l = 0; l < layers; ++l)
        for (const FLine& line : placed[l].lines) line.pack(lines + (at++) * LINE_FLOATS);
}

}  // extern "C"

extern "C" {

// The seeded quantile table, ROWS x QUANTILES floats.
void qmo_kurze_table(float* out) {
    const auto& t = kurze_table();
    std::copy(t.begin(), t.end(), out);
}

double qmo_relative_intensity(double lambda, double p) { return relative_intensity(kurze_table().data(), lambda, p); }

// node_probabilities: `n` (p, share) pairs written when `out` has room for them; returns the count.
size_t qmo_node_probabilities(double* out, size_t room) {
    const auto nodes = node_probabilities();
    for (size_t k = 0; k < nodes.size() && k < room; ++k) {
        out[2 * k] = nodes[k].first;
        out[2 * k + 1] = nodes[k].second;
    }
    return nodes.size();
}

// pass_values(mean) (kind 0), line_values(lambda, none) (kind 1) or an event of duty `x` (kind 2):
// the (value, share) pairs ascending; returns the count.
size_t qmo_line_values(int kind, double x, double* out, size_t room) {
    LineValues v = kind == 0 ? pass_values(x) : (kind == 1 ? line_values(kurze_table().data(), x, nullptr)
                                                           : line_values(kurze_table().data(), 0.0, &x));
    for (size_t k = 0; k < v.pairs.size() && k < room; ++k) {
        out[2 * k] = v.pairs[k].first;
        out[2 * k + 1] = v.pairs[k].second;
    }
    return v.pairs.size();
}

// Distribution::runs of a moment at `intensity` over (values x scale) on the grid of `bin_db`:
// per bin its (shares, weighted) into `out` (bins + 1 pairs, the last silence), and whether any
// bin was visited twice (returned 1).
int qmo_runs(const double* values, const double* shares, size_t n, double scale, double intensity, double bin_db,
             double* out) {
    std::vector<std::pair<double, double>> p;
    for (size_t k = 0; k < n; ++k) p.push_back({values[k], shares[k]});
    const LineValues line = LineValues::of(p);
    const Distribution grid = Distribution::empty(bin_db);
    const size_t bins = grid.bins.size();
    std::vector<bool> seen(bins + 1, false);
    int twice = 0;
    for (size_t k = 0; k < 2 * (bins + 1); ++k) out[k] = 0.0;
    grid.runs(intensity, line, scale, [&](long b, double s, double w) {
        const size_t at = b < 0 ? bins : (size_t)b;
        if (seen[at]) twice = 1;
        seen[at] = true;
        out[2 * at] = s;
        out[2 * at + 1] = w;
    });
    return twice;
}

// `count` lines of lambda `lambda` added over silence one at a time (with_line); returns the
// summed intensity over the lines' summed mean.
double qmo_many_lines_energy(double lambda, int count, double bin_db) {
    const LineValues v = line_values(kurze_table().data(), lambda, nullptr);
    double mean = 0.0;
    for (const auto& [value, share] : v.pairs) mean += value * share;
    Distribution sum = Distribution::empty(bin_db);
    sum.silent = 1.0;
    for (int k = 0; k < count; ++k) sum = sum.with_line(v, 1.0);
    double total = 0.0;
    sum.intensities([&](double i, double s) { total += i * s; });
    return total / (count * mean);
}

// The level exceeded `exceeded` of the time by a distribution of one line of `lambda` at unit mean.
double qmo_line_exceeded_db(double lambda, double exceeded, double bin_db) {
    Distribution d = Distribution::empty(bin_db);
    d.silent = 1.0;
    d = d.with_line(line_values(kurze_table().data(), lambda, nullptr), 1.0);
    return d.exceeded_db(exceeded);
}

// The map's hour groups: per period the slot count of each group (up to 4, zero-padded) and the
// mean flow factor per (profile, period, group), as with any line following a profile.
void qmo_hour_groups(int* slots /* 3 x 4 */, double* factors /* 3 profiles x 3 x 4 */) {
    for (int p = 0; p < PERIODS; ++p) {
        int hours = 0;
        const auto groups = hour_groups(true, p, hours);
        for (int g = 0; g < 4; ++g) slots[p * 4 + g] = g < (int)groups.size() ? (int)groups[g].size() : 0;
        for (int profile = 0; profile < PROFILES; ++profile)
            for (int g = 0; g < 4; ++g) {
                double f = 1.0;
                if (g < (int)groups.size()) {
                    f = 0.0;
                    for (int s : groups[g]) f += hour_factor(profile, p, s);
                    f /= (double)groups[g].size();
                }
                factors[(profile * PERIODS + p) * 4 + g] = f;
            }
    }
}

// Per pixel (packed, `offsets[i]` floats into `input`): the mixtures (PERIODS x CURVES x bins x 2:
// each layer's distribution, then the place's, per bin share and share x intensity; may be null)
// and Nden per curve (CURVES floats: each layer alone on its own curves, then the place).
int qmo_pixel_loudness(const float* input, const uint64_t* offsets, size_t pixels, double bin_db, double* mixtures,
                       double* ndens, int threads) {
    const float* table = kurze_table().data();
    const size_t bins = (size_t)std::round(SPAN_DB / bin_db);
    qmo::parallel_for(pixels, std::max(threads, 1), [&](size_t i, int) {
        const float* in = input + offsets[i];
        std::vector<std::vector<Line>> layers(LAYERS);
        double steady[LAYERS][PERIODS];
        size_t at = 0;
        for (int l = 0; l < LAYERS; ++l) {
            for (int p = 0; p < PERIODS; ++p) steady[l][p] = in[at + p];
            const int count = (int)in[at + PERIODS];
            at += PERIODS + 1;
            for (int k = 0; k < count; ++k) layers[l].push_back(Line::unpack(in + at + (size_t)k * LINE_FLOATS));
            at += (size_t)count * LINE_FLOATS;
        }
        const PixelResult r = layered_distributions(table, layers, steady, bin_db);
        std::vector<Curve> curves;
        for (int c = 0; c < CURVES; ++c)
            for (int p = 0; p < PERIODS; ++p) curves.push_back(Curve::of(in + at + (size_t)(c * PERIODS + p) * THIRDS));
        for (int c = 0; c < CURVES; ++c) {
            const Distribution* d = c < LAYERS ? r.each[c] : r.whole;
            ndens[i * CURVES + c] = nden(d, &curves[c * PERIODS]);
            if (mixtures)
                for (int p = 0; p < PERIODS; ++p)
                    for (size_t b = 0; b < bins; ++b) {
                        double* m = mixtures + (((i * PERIODS + p) * CURVES + c) * bins + b) * 2;
                        m[0] = d[p].bins[b][0];
                        m[1] = d[p].bins[b][1];
                    }
        }
    });
    return 0;
}

}  // extern "C"
