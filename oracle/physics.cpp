// TEST INFRASTRUCTURE — FP64 oracle: CNOSSOS-EU boundary terms, the ray transfer, foliage, bands
// and air absorption, restated from the reference's Rust (physics crate). Each function cites the
// file:line it follows. Not part of the product path.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace qmo {

const double BAND_FREQUENCY_HZ[BANDS] = {63.0, 125.0, 250.0, 500.0, 1000.0, 2000.0, 4000.0, 8000.0};
const double A_WEIGHTING_DB[BANDS] = {-26.2, -16.1, -8.6, -3.2, 0.0, 1.2, 1.0, -1.1};
const double PERIOD_HOURS[PERIODS] = {12.0, 4.0, 8.0};
const double PERIOD_PENALTY_DB[PERIODS] = {0.0, 5.0, 10.0};

static const double LN_10 = 2.302585092994045684017991454684364208;

// physics/src/bands.rs:36-38
double energy(double level_db) { return std::exp(level_db * (LN_10 / 10.0)); }
// bands.rs:41-43
double amplitude(double level_db) { return std::exp(level_db * (LN_10 / 20.0)); }
// bands.rs:47-56
double level_db(double e) { return e > 0.0 ? 10.0 * std::log10(e) : -std::numeric_limits<double>::infinity(); }
// bands.rs:59-66
double lden_energy(const double period_energy[PERIODS]) {
    double total = 0.0;
    for (int p = 0; p < PERIODS; ++p) total += PERIOD_HOURS[p] * period_energy[p] * energy(PERIOD_PENALTY_DB[p]);
    return total / 24.0;
}

// physics/src/atmosphere.rs:17-19
double exact_mid_band_frequency_hz(int band) { return 1000.0 * std::pow(10.0, 0.3 * (band - 4.0)); }

// atmosphere.rs:22-45 (ISO 9613-1:1993 eqs 3-5, Annex B)
double alpha_db_per_km(double frequency_hz, double temperature_c, double relative_humidity_pct, double pressure_kpa) {
    const double temperature = temperature_c + 273.15;
    const double reference_temperature = 293.15, triple_point = 273.16;
    const double pressure_ratio = pressure_kpa / 101.325;
    const double saturation_exponent = -6.8346 * std::pow(triple_point / temperature, 1.261) + 4.6151;
    const double humidity = relative_humidity_pct * std::pow(10.0, saturation_exponent) / pressure_ratio;
    const double oxygen = pressure_ratio * (24.0 + 4.04e4 * humidity * (0.02 + humidity) / (0.391 + humidity));
    const double t_ratio = temperature / reference_temperature;
    const double nitrogen = pressure_ratio * std::pow(t_ratio, -0.5) *
                            (9.0 + 280.0 * humidity * std::exp(-4.170 * (std::pow(t_ratio, -1.0 / 3.0) - 1.0)));
    const double f2 = frequency_hz * frequency_hz;
    return 8.686 * f2 *
           (1.84e-11 / pressure_ratio * std::sqrt(t_ratio) +
            std::pow(t_ratio, -2.5) * (0.01275 * std::exp(-2239.1 / temperature) / (oxygen + f2 / oxygen) +
                                       0.1068 * std::exp(-3352.0 / temperature) / (nitrogen + f2 / nitrogen))) *
           1000.0;
}

// atmosphere.rs:48-56
void alpha_bands(double temperature_c, double relative_humidity_pct, double out[BANDS]) {
    for (int b = 0; b < BANDS; ++b)
        out[b] = alpha_db_per_km(exact_mid_band_frequency_hz(b), temperature_c, relative_humidity_pct, 101.325);
}

// ---------------------------------------------------------------------------------------------
// physics/src/cnossos/mod.rs:62-88 interpolate_with_steps
static double interpolate_with_steps(const double* xs, const double* ys, size_t n, double x, bool before) {
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    size_t upper = before ? (size_t)(std::lower_bound(xs, xs + n, x) - xs)
                          : (size_t)(std::upper_bound(xs, xs + n, x) - xs);
    if (before && xs[upper] == x) return ys[upper];
    const double x0 = xs[upper - 1], x1 = xs[upper];
    if (x1 == x0) return ys[upper];
    return ys[upper - 1] + (x - x0) / (x1 - x0) * (ys[upper] - ys[upper - 1]);
}

double VerticalProfile::ground_altitude_at(double x) const { return interpolate_with_steps(d, z, n, x, false); }
double VerticalProfile::ground_altitude_before(double x) const { return interpolate_with_steps(d, z, n, x, true); }

// mod.rs:40-58
double VerticalProfile::mean_ground_factor(double start_m, double end_m) const {
    if (end_m - start_m < 1e-3) return interpolate_with_steps(d, g, n, start_m, false);
    const size_t first = (size_t)(std::upper_bound(d, d + n, start_m) - d);
    const size_t last = (size_t)(std::lower_bound(d, d + n, end_m) - d);
    double px = start_m, pg = interpolate_with_steps(d, g, n, start_m, false);
    double total = 0.0;
    for (size_t i = first; i < last; ++i) {
        total += 0.5 * (pg + g[i]) * (d[i] - px);
        px = d[i];
        pg = g[i];
    }
    total += 0.5 * (pg + interpolate_with_steps(d, g, n, end_m, true)) * (end_m - px);
    return total / (end_m - start_m);
}

// physics/src/cnossos/mean_plane.rs:19-64
MeanPlane fit_mean_plane(const VerticalProfile& profile, double start_m, double end_m) {
    const double reference_z = profile.ground_altitude_at(start_m);
    const double length = end_m - start_m;
    if (length < 1e-3) return MeanPlane{0.0, reference_z};
    double integral_z = 0.0, integral_xz = 0.0;
    auto accumulate = [&](double x0, double z0, double x1, double z1) {
        const double u0 = x0 - start_m, u1 = x1 - start_m;
        const double w0 = z0 - reference_z, w1 = z1 - reference_z;
        const double width = u1 - u0;
        if (width <= 0.0) return;
        integral_z += 0.5 * (w0 + w1) * width;
        integral_xz += width * (u0 * (2.0 * w0 + w1) + u1 * (w0 + 2.0 * w1)) / 6.0;
    };
    const double* x = profile.d;
    const double* z = profile.z;
    const size_t first = (size_t)(std::upper_bound(x, x + profile.n, start_m) - x);
    const size_t last = (size_t)(std::lower_bound(x, x + profile.n, end_m) - x);
    double px = start_m, pz = profile.ground_altitude_at(start_m);
    for (size_t i = first; i < last; ++i) {
        accumulate(px, pz, x[i], z[i]);
        px = x[i];
        pz = z[i];
    }
    accumulate(px, pz, end_m, profile.ground_altitude_before(end_m));
    const double slope = 12.0 * (integral_xz - 0.5 * length * integral_z) / (length * length * length);
    const double intercept_relative = integral_z / length - 0.5 * slope * length;
    return MeanPlane{slope, reference_z + intercept_relative - slope * start_m};
}

// mean_plane.rs:74-96
double MeanPlane::signed_height(double x, double z) const { return (z - altitude_at(x)) / std::sqrt(1.0 + slope * slope); }
double MeanPlane::projected_distance(P2 from, P2 to) const {
    return std::fabs((to.x - from.x) + slope * (to.z - from.z)) / std::sqrt(1.0 + slope * slope);
}
P2 MeanPlane::mirror(P2 p) const {
    const double norm_sq = 1.0 + slope * slope;
    const double offset = (p.z - altitude_at(p.x)) / norm_sq;
    return P2{p.x + 2.0 * offset * slope, p.z - 2.0 * offset};
}

// mean_plane.rs:109-120
EquivalentGeometry EquivalentGeometry::over(MeanPlane plane, P2 from, P2 to) {
    const double sh = plane.signed_height(from.x, from.z);
    const double rh = plane.signed_height(to.x, to.z);
    EquivalentGeometry g;
    g.plane = plane;
    g.source_side_height_m = std::max(sh, 0.0);
    g.receiver_side_height_m = std::max(rh, 0.0);
    g.source_side_below_plane = sh < 0.0;
    g.receiver_side_below_plane = rh < 0.0;
    g.projected_distance_m = plane.projected_distance(from, to);
    return g;
}

// ---------------------------------------------------------------------------------------------
// physics/src/cnossos/ground.rs
static const double SHORT_PATH_HEIGHT_FACTOR = 30.0;
static const double FAVOURABLE_CURVATURE_A0_PER_M = 2e-4;
static const double FAVOURABLE_TERRAIN_HEIGHT_COEFFICIENT = 6e-3;
static const double MINIMUM_HEIGHT_SUM_M = 1e-3;
static const double HARD_GROUND_BELOW = 1e-3;

// ground.rs:39-51
double ground_factor_near_source(const EquivalentGeometry& g, double path_ground_factor, double source_ground_factor) {
    const double height_sum = std::max(g.source_side_height_m + g.receiver_side_height_m, MINIMUM_HEIGHT_SUM_M);
    const double ratio = g.projected_distance_m / (SHORT_PATH_HEIGHT_FACTOR * height_sum);
    if (ratio <= 1.0) return path_ground_factor * ratio + source_ground_factor * (1.0 - ratio);
    return path_ground_factor;
}

struct BandConstants {
    double f_2_5, f_1_5, f_0_75, wavenumber;
};
// ground.rs:61-74
static const BandConstants& band_constants(int band) {
    static BandConstants table[BANDS];
    static bool ready = [] {
        for (int b = 0; b < BANDS; ++b) {
            const double f = BAND_FREQUENCY_HZ[b];
            table[b] = BandConstants{std::pow(f, 2.5), std::pow(f, 1.5), std::pow(f, 0.75),
                                     2.0 * M_PI * f / SPEED_OF_SOUND_M_PER_S};
        }
        return true;
    }();
    (void)ready;
    return table[band];
}

// ground.rs:78-89
static double analytic_ground_db(int band, double dp, double zs, double zr, double gw13) {
    const BandConstants& c = band_constants(band);
    const double k = c.wavenumber;
    const double gw26 = gw13 * gw13;
    const double w = 0.0185 * c.f_2_5 * gw26 / (c.f_1_5 * gw26 + 1.3e3 * c.f_0_75 * gw13 + 1.16e6);
    const double wd = w * dp;
    const double cf = dp * (1.0 + 3.0 * wd * std::exp(-std::sqrt(wd))) / (1.0 + wd);
    const double root = std::sqrt(2.0 * cf / k);
    const double product = (zs * zs - root * zs + cf / k) * (zr * zr - root * zr + cf / k);
    return -10.0 * std::log10(4.0 * k * k / (dp * dp) * product);
}

// ground.rs:92-102
static double favourable_floor_db(const EquivalentGeometry& g, double floor_factor) {
    const double height_sum = std::max(g.source_side_height_m + g.receiver_side_height_m, MINIMUM_HEIGHT_SUM_M);
    const double base = -3.0 * (1.0 - floor_factor);
    const double limit = SHORT_PATH_HEIGHT_FACTOR * height_sum;
    if (g.projected_distance_m <= limit) return base;
    return base * (1.0 + 2.0 * (1.0 - limit / g.projected_distance_m));
}

// ground.rs:105-137
void ground_attenuation_bands(const EquivalentGeometry& g, GroundFactors factors, State state, double out[BANDS]) {
    const double dp = std::max(g.projected_distance_m, 1e-6);
    const double zs = g.source_side_height_m, zr = g.receiver_side_height_m;
    const double gw13 = std::pow(std::max(factors.impedance, 0.0), 1.3);
    if (state == HOMOGENEOUS) {
        if (factors.path < HARD_GROUND_BELOW) {
            for (int b = 0; b < BANDS; ++b) out[b] = -3.0;
            return;
        }
        const double floor = -3.0 * (1.0 - factors.floor);
        for (int b = 0; b < BANDS; ++b) out[b] = std::max(analytic_ground_db(b, dp, zs, zr, gw13), floor);
        return;
    }
    const double floor = favourable_floor_db(g, factors.floor);
    if (factors.path < HARD_GROUND_BELOW) {
        for (int b = 0; b < BANDS; ++b) out[b] = floor;
        return;
    }
    const double height_sum = std::max(zs + zr, MINIMUM_HEIGHT_SUM_M);
    const double curvature = FAVOURABLE_CURVATURE_A0_PER_M * dp * dp / 2.0;
    const double terrain = FAVOURABLE_TERRAIN_HEIGHT_COEFFICIENT * dp / height_sum;
    const double rs = zs / height_sum, rr = zr / height_sum;
    const double zs_f = zs + curvature * (rs * rs) + terrain;
    const double zr_f = zr + curvature * (rr * rr) + terrain;
    for (int b = 0; b < BANDS; ++b) out[b] = std::max(analytic_ground_db(b, dp, zs_f, zr_f, gw13), floor);
}

// ---------------------------------------------------------------------------------------------
// physics/src/cnossos/rubber_band.rs
double distance(P2 a, P2 b) {
    const double dx = b.x - a.x, dz = b.z - a.z;
    return std::sqrt(dx * dx + dz * dz);
}

// rubber_band.rs:182-193
double triangle_excess(P2 s, P2 o, P2 r) {
    const double ux = o.x - s.x, uz = o.z - s.z, vx = r.x - o.x, vz = r.z - o.z;
    const double lu = std::sqrt(ux * ux + uz * uz), lv = std::sqrt(vx * vx + vz * vz);
    const double wx = ux + vx, wz = uz + vz;
    const double lw = std::sqrt(wx * wx + wz * wz);
    const double dot = ux * vx + uz * vz;
    if (dot <= 0.0) return lu + lv - lw;
    const double cross = ux * vz - uz * vx;
    return 2.0 * cross * cross / ((lu * lv + dot) * (lu + lv + lw));
}

static double chord_altitude(P2 from, P2 to, double x) {
    return from.z + (to.z - from.z) * (x - from.x) / (to.x - from.x);
}

// rubber_band.rs:18-21, 48-55
StateRay StateRay::between(State state, P2 source, P2 receiver) {
    return StateRay{state, std::max(1000.0, 8.0 * distance(source, receiver))};
}

// rubber_band.rs:65-80
double StateRay::arc_excess(double chord) const {
    if (state == HOMOGENEOUS) return 0.0;
    const double x = chord / (2.0 * radius_m);
    if (x > 1.0 / 16.0) return 2.0 * radius_m * std::asin(std::min(x, 1.0)) - chord;
    const double x2 = x * x;
    return 2.0 * radius_m * x * x2 *
           (1.0 / 6.0 + x2 * (3.0 / 40.0 + x2 * (5.0 / 112.0 + x2 * (35.0 / 1152.0))));
}

// rubber_band.rs:57-61
double StateRay::length(P2 from, P2 to) const {
    const double chord = distance(from, to);
    return chord + arc_excess(chord);
}

// rubber_band.rs:86-121
double StateRay::path_difference(P2 from, const P2* points, size_t n, P2 to) const {
    if (n == 0) return 0.0;
    double chords = 0.0, arcs = 0.0;
    P2 start = from;
    for (size_t i = 0; i < n; ++i) {
        chords += triangle_excess(start, points[i], to);
        arcs += arc_excess(distance(start, points[i]));
        start = points[i];
    }
    arcs += arc_excess(distance(start, to)) - arc_excess(distance(from, to));
    const double excess = chords + arcs;
    if (n > 1) return excess;
    const P2 first = points[0];
    const double chord = chord_altitude(from, to, first.x);
    if (state == HOMOGENEOUS) return first.z >= chord ? excess : -excess;
    if (first.z >= chord) return excess;
    const P2 on_chord{first.x, chord};
    return 2.0 * triangle_excess(from, on_chord, to) - triangle_excess(from, first, to) +
           2.0 * arc_excess(distance(from, on_chord)) + 2.0 * arc_excess(distance(on_chord, to)) -
           arc_excess(distance(from, first)) - arc_excess(distance(first, to)) - arc_excess(distance(from, to));
}

// rubber_band.rs:125-138 (Sag constants) and 349-359 (Sag::at), taken per call.
double StateRay::sag_at(P2 source, P2 receiver, double x_from_source) const {
    if (state != FAVOURABLE) return 0.0;
    const double horizontal = receiver.x - source.x;
    const double chord = distance(source, receiver);
    const double gamma2 = radius_m * radius_m;
    const double along_per_x = chord / horizontal;
    const double centre_term = std::sqrt(gamma2 - 0.25 * chord * chord);
    const double along = x_from_source * along_per_x;
    const double offset = along - 0.5 * chord;
    const double perpendicular = along * (chord - along) / (std::sqrt(std::max(gamma2 - offset * offset, 0.0)) + centre_term);
    return perpendicular * along_per_x;
}

// rubber_band.rs:202-264
static void diffraction_path(const StateRay& ray, P2 source, P2 receiver, const std::vector<P2>& candidates,
                             VerticalScratch& s, DiffractionPath& path) {
    path.points.clear();
    path.blocked = false;
    s.blocking.clear();
    s.blocking_index.clear();
    for (size_t index = 0; index < candidates.size(); ++index) {
        const P2 c = candidates[index];
        if (c.z > chord_altitude(source, receiver, c.x) + ray.sag_at(source, receiver, c.x - source.x)) {
            if (!s.blocking.empty() && s.blocking.back().x == c.x) {
                if (c.z > s.blocking.back().z) {
                    s.blocking.back() = c;
                    s.blocking_index.back() = index;
                }
            } else {
                s.blocking.push_back(c);
                s.blocking_index.push_back(index);
            }
        }
    }
    if (!s.blocking.empty()) {
        path.blocked = true;
        s.hull.clear();
        s.hull_index.clear();
        const size_t NONE = (size_t)-1;
        auto visit = [&](P2 point, size_t index) {
            while (s.hull.size() >= 2) {
                const P2 a = s.hull[s.hull.size() - 2], b = s.hull[s.hull.size() - 1];
                const double ray_at_b = chord_altitude(a, point, b.x) + ray.sag_at(a, point, b.x - a.x);
                if (b.z <= ray_at_b) {
                    s.hull.pop_back();
                    s.hull_index.pop_back();
                } else {
                    break;
                }
            }
            s.hull.push_back(point);
            s.hull_index.push_back(index);
        };
        visit(source, NONE);
        for (size_t i = 0; i < s.blocking.size(); ++i) visit(s.blocking[i], s.blocking_index[i]);
        visit(receiver, NONE);
        for (size_t i = 1; i + 1 < s.hull.size(); ++i) path.points.push_back(candidates[s.hull_index[i]]);
        return;
    }
    bool have = false;
    double best_value = 0.0;
    P2 best{};
    for (const P2& c : candidates) {
        const double delta = ray.path_difference(source, &c, 1, receiver);
        if (!have || delta > best_value) {
            have = true;
            best_value = delta;
            best = c;
        }
    }
    if (have) path.points.push_back(best);
}

// ---------------------------------------------------------------------------------------------
// physics/src/cnossos/diffraction.rs
static const double DIFFRACTION_CAP_DB = 25.0;
static const double MULTIPLE_DIFFRACTION_MINIMUM_SPAN_M = 0.3;

// diffraction.rs:19-30
static void diffraction_db(double path_difference_m, const double c_second[BANDS], double out[BANDS]) {
    for (int b = 0; b < BANDS; ++b) {
        const double lambda = SPEED_OF_SOUND_M_PER_S / BAND_FREQUENCY_HZ[b];
        const double x = 40.0 * c_second[b] * path_difference_m / lambda;
        out[b] = x >= -2.0 ? 10.0 * std::log10(3.0 + x) : 0.0;
    }
}

// diffraction.rs:33-35
static double ground_split_db(double ground_db, double mirrored_db, double direct_db) {
    return -20.0 * std::log10(1.0 + (amplitude(-ground_db) - 1.0) * amplitude(-(mirrored_db - direct_db)));
}

struct DirectPath {
    EquivalentGeometry geometry;
    double path_ground;
    double near_source;
};

// diffraction.rs:46-63
static DirectPath direct_path_of(const VerticalPath& path) {
    const double length = path.receiver.x;
    DirectPath d;
    d.geometry = EquivalentGeometry::over(fit_mean_plane(path.profile, 0.0, length), path.source, path.receiver);
    d.path_ground = path.profile.mean_ground_factor(0.0, length);
    d.near_source = ground_factor_near_source(d.geometry, d.path_ground, path.source_ground_factor);
    return d;
}

// diffraction.rs:199-213
static GroundFactors source_side_factors(State state, double path_ground, double near_source) {
    return GroundFactors{path_ground, state == HOMOGENEOUS ? near_source : path_ground, near_source};
}

// diffraction.rs:66-196
static void boundary_for_candidates(const VerticalPath& path, State state, const DirectPath& shared,
                                    const std::vector<P2>& candidates, VerticalScratch& s,
                                    DiffractionPath& diffraction, StateBoundary& boundary) {
    const P2 source = path.source, receiver = path.receiver;
    const double length = receiver.x;
    const VerticalProfile& profile = path.profile;
    ground_attenuation_bands(shared.geometry, source_side_factors(state, shared.path_ground, shared.near_source), state,
                             boundary.attenuation_db);
    for (int b = 0; b < BANDS; ++b) boundary.without_ground_db[b] = 0.0;
    const StateRay ray = StateRay::between(state, source, receiver);
    diffraction_path(ray, source, receiver, candidates, s, diffraction);
    const std::vector<P2>& points = diffraction.points;
    if (points.empty()) return;
    const bool blocked = diffraction.blocked;
    const P2 first = points.front(), last = points.back();
    const size_t n = points.size();
    const double delta = ray.path_difference(source, points.data(), n, receiver);
    double span = 0.0;
    if (n > 1)
        for (size_t i = 0; i + 1 < n; ++i) span += ray.length(points[i], points[i + 1]);
    double c_second[BANDS];
    for (int b = 0; b < BANDS; ++b) {
        if (n > 1 && span > MULTIPLE_DIFFRACTION_MINIMUM_SPAN_M) {
            const double x = 5.0 * SPEED_OF_SOUND_M_PER_S / BAND_FREQUENCY_HZ[b] / span;
            c_second[b] = (1.0 + x * x) / (1.0 / 3.0 + x * x);
        } else {
            c_second[b] = 1.0;
        }
    }
    const EquivalentGeometry source_side = EquivalentGeometry::over(fit_mean_plane(profile, 0.0, first.x), source, first);
    const EquivalentGeometry receiver_side =
        EquivalentGeometry::over(fit_mean_plane(profile, last.x, length), last, receiver);
    const P2 source_image = source_side.plane.mirror(source);
    const P2 receiver_image = receiver_side.plane.mirror(receiver);
    double direct_db[BANDS], source_image_db[BANDS], receiver_image_db[BANDS];
    diffraction_db(delta, c_second, direct_db);
    diffraction_db(ray.path_difference(source_image, points.data(), n, receiver), c_second, source_image_db);
    diffraction_db(ray.path_difference(source, points.data(), n, receiver_image), c_second, receiver_image_db);
    const double source_side_ground = profile.mean_ground_factor(0.0, first.x);
    const double source_side_near = ground_factor_near_source(source_side, source_side_ground, path.source_ground_factor);
    double ground_so[BANDS], ground_or[BANDS];
    ground_attenuation_bands(source_side, source_side_factors(state, source_side_ground, source_side_near), state,
                             ground_so);
    const double receiver_side_ground = profile.mean_ground_factor(last.x, length);
    ground_attenuation_bands(receiver_side,
                             GroundFactors{receiver_side_ground, receiver_side_ground, receiver_side_ground}, state,
                             ground_or);
    const double rayleigh_star =
        blocked ? 0.0 : ray.path_difference(source_image, points.data(), n, receiver_image);
    for (int b = 0; b < BANDS; ++b) {
        const double lambda = SPEED_OF_SOUND_M_PER_S / BAND_FREQUENCY_HZ[b];
        const bool admitted = blocked || (delta > -lambda / 20.0 && delta > lambda / 4.0 - rayleigh_star);
        if (!admitted) continue;
        double sr = direct_db[b];
        const double ground_free = std::min(std::max(sr, 0.0), DIFFRACTION_CAP_DB);
        double split_so;
        if (source_side.source_side_below_plane) {
            sr = source_image_db[b];
            split_so = ground_so[b];
        } else {
            split_so = ground_split_db(ground_so[b], source_image_db[b], direct_db[b]);
        }
        if (!std::isfinite(split_so)) {
            split_so = ground_so[b];
            sr = source_image_db[b];
        }
        double split_or;
        if (receiver_side.receiver_side_below_plane) {
            sr = receiver_image_db[b];
            split_or = ground_or[b];
        } else {
            split_or = ground_split_db(ground_or[b], receiver_image_db[b], sr);
        }
        if (!std::isfinite(split_or)) {
            split_or = ground_or[b];
            sr = receiver_image_db[b];
        }
        boundary.attenuation_db[b] = std::min(std::max(sr, 0.0), DIFFRACTION_CAP_DB) + split_so + split_or;
        boundary.without_ground_db[b] = ground_free;
    }
}

// physics/src/cnossos/mod.rs:160-184 collect_candidates
static void collect_candidates(const VerticalPath& path, std::vector<P2>& out) {
    out.clear();
    const double length = path.receiver.x;
    auto inside = [&](const P2& p) { return p.x > 0.0 && p.x < length; };
    size_t a = 0, b = 0;
    auto skip = [&](const P2* list, size_t n, size_t& i) {
        while (i < n && !inside(list[i])) ++i;
    };
    skip(path.terrain_candidates, path.terrain_count, a);
    skip(path.obstacle_tops, path.tops_count, b);
    while (a < path.terrain_count || b < path.tops_count) {
        if (a < path.terrain_count && b < path.tops_count) {
            if (path.obstacle_tops[b].x < path.terrain_candidates[a].x) {
                out.push_back(path.obstacle_tops[b++]);
            } else {
                out.push_back(path.terrain_candidates[a++]);
            }
        } else if (a < path.terrain_count) {
            out.push_back(path.terrain_candidates[a++]);
        } else {
            out.push_back(path.obstacle_tops[b++]);
        }
        skip(path.terrain_candidates, path.terrain_count, a);
        skip(path.obstacle_tops, path.tops_count, b);
    }
}

// mod.rs:123-135 state_boundaries
void state_boundaries(const VerticalPath& path, VerticalScratch& scratch, StateBoundary out[2], WorkCount* work) {
    collect_candidates(path, scratch.candidates);
    if (work) work->candidates += scratch.candidates.size();
    const DirectPath direct = direct_path_of(path);
    boundary_for_candidates(path, HOMOGENEOUS, direct, scratch.candidates, scratch, scratch.path, out[0]);
    std::swap(scratch.calm_path, scratch.path);
    boundary_for_candidates(path, FAVOURABLE, direct, scratch.candidates, scratch, scratch.path, out[1]);
}

// ---------------------------------------------------------------------------------------------
// physics/src/foliage.rs:13-30
void foliage_attenuation_db(double depth_m, double out[BANDS]) {
    static const double SHORT_DB[BANDS] = {0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 2.0, 3.0};
    static const double DB_PER_M[BANDS] = {0.02, 0.03, 0.04, 0.05, 0.06, 0.08, 0.09, 0.12};
    for (int b = 0; b < BANDS; ++b) {
        if (depth_m < 10.0) out[b] = 0.0;
        else if (depth_m < 20.0) out[b] = SHORT_DB[b];
        else out[b] = DB_PER_M[b] * std::min(depth_m, 200.0);
    }
}

// foliage.rs:36-64
double foliage_depth_m(const Profile& profile, const StateRay& ray, P2 source, P2 receiver) {
    const double length = receiver.x - source.x;
    auto point = [&](size_t sample) {
        const double t = profile.t[sample];
        return P2{t * length, source.z + t * (receiver.z - source.z) + ray.sag_at(source, receiver, t * length)};
    };
    auto weight = [&](size_t sample, P2 p) {
        const double above = p.z - profile.ground_m[sample];
        return (above > 0.0 && above <= profile.canopy_m[sample]) ? profile.forest_cover[sample] : 0.0;
    };
    const std::vector<double>& cover = profile.forest_cover;
    double depth = 0.0;
    for (size_t s = 1; s < profile.t.size(); ++s) {
        if (cover[s - 1] <= 0.0 && cover[s] <= 0.0) continue;
        const P2 a = point(s - 1), b = point(s);
        depth += 0.5 * (weight(s - 1, a) + weight(s, b)) * distance(a, b);
    }
    return depth;
}

// ---------------------------------------------------------------------------------------------
// physics/src/ray.rs
static const double CREST_ENDS_M = 300.0;
static const double CREST_SPACING_M = 150.0;

// ray.rs:412-421
static double interpolate_points(const std::vector<P2>& points, double at) {
    size_t upper = 0;
    while (upper < points.size() && points[upper].x <= at) ++upper;  // partition_point
    upper = std::min(std::max(upper, (size_t)1), points.size() - 1);
    const P2 p0 = points[upper - 1], p1 = points[upper];
    double f = 0.0;
    if (p1.x > p0.x) f = std::min(std::max((at - p0.x) / (p1.x - p0.x), 0.0), 1.0);
    return p0.z + f * (p1.z - p0.z);
}

// ray.rs:201-226
void containing_footprints(const std::vector<Crossing>& crossings, uint64_t own_footprint, std::vector<uint64_t>& out) {
    out.clear();
    if (own_footprint == 0) return;
    for (const Crossing& c : crossings)
        if (c.bounds_roof) out.push_back(c.footprint_id);
    std::sort(out.begin(), out.end());
    size_t odd = 0, start = 0;
    while (start < out.size()) {
        const uint64_t id = out[start];
        size_t end = start;
        while (end < out.size() && out[end] == id) ++end;
        if ((end - start) % 2 == 1) out[odd++] = id;
        start = end;
    }
    out.resize(odd);
    out.push_back(own_footprint);
    std::sort(out.begin(), out.end());
}

// ray.rs:360-385
static void crests_of(const std::vector<P2>& terrain, double length, std::vector<P2>& out) {
    out.clear();
    auto middle = [&](double x) { return x > CREST_ENDS_M && x < length - CREST_ENDS_M; };
    bool have = false;
    P2 best{};
    int64_t at = 0;
    for (const P2& point : terrain) {
        const int64_t group = (int64_t)std::floor(point.x / CREST_SPACING_M);
        if (have && (!middle(point.x) || at != group)) {
            out.push_back(best);
            have = false;
        }
        if (!middle(point.x)) {
            out.push_back(point);
            continue;
        }
        if (have && point.z <= best.z) {
            // keep
        } else {
            have = true;
            best = point;
            at = group;
        }
    }
    if (have) out.push_back(best);
}

// ray.rs:389-404
static void clip_roofs_in_closing_order(std::vector<RayScratch::Roof>& roofs) {
    std::stable_sort(roofs.begin(), roofs.end(), [](const RayScratch::Roof& a, const RayScratch::Roof& b) {
        if (a.x1 != b.x1) return a.x1 < b.x1;
        return a.x0 < b.x0;
    });
    double covered_to = -std::numeric_limits<double>::infinity();
    size_t kept = 0;
    for (size_t i = 0; i < roofs.size(); ++i) {
        RayScratch::Roof roof = roofs[i];
        if (roof.x1 <= covered_to) continue;
        if (roof.x0 < covered_to) {
            const double span = roof.x1 - roof.x0;
            roof.top0 += (covered_to - roof.x0) / span * (roof.top1 - roof.top0);
            roof.x0 = covered_to;
        }
        covered_to = roof.x1;
        roofs[kept++] = roof;
    }
    roofs.resize(kept);
}

// ray.rs:255-354 PathBuffers::fill
void fill_path(RayScratch& s, const Profile& profile, const std::vector<Crossing>& crossings, const RayEnds& ends) {
    const double length = std::max(profile.horizontal_m, 1.0);
    s.terrain.clear();
    const double source_ground = ends.has_source_ground ? ends.source_ground_m : profile.ground_m[0];
    for (size_t i = 0; i < profile.t.size(); ++i) {
        const double x = profile.t[i] * length;
        const double z = profile.ground_m[i];
        s.terrain.push_back(P2{x, x < ends.platform_half_width_m ? std::min(z, source_ground) : z});
    }
    crests_of(s.terrain, length, s.crests);
    const std::vector<P2>& terrain = s.terrain;
    auto terrain_at = [&](double x) { return interpolate_points(terrain, x); };
    auto ground_factor_at = [&](double x) {
        size_t upper = 0;
        while (upper < terrain.size() && terrain[upper].x <= x) ++upper;
        upper = std::min(std::max(upper, (size_t)1), terrain.size() - 1);
        const double x0 = terrain[upper - 1].x, x1 = terrain[upper].x;
        double f = 0.0;
        if (x1 > x0) f = std::min(std::max((x - x0) / (x1 - x0), 0.0), 1.0);
        const double a = profile.ground_factor[upper - 1], b = profile.ground_factor[upper];
        return a + f * (b - a);
    };
    s.tops.clear();
    s.roofs.clear();
    s.buildings.clear();
    containing_footprints(crossings, ends.own_footprint, s.containing);
    for (const Crossing& c : crossings) {
        if (std::binary_search(s.containing.begin(), s.containing.end(), c.footprint_id)) continue;
        const double x = c.t * length;
        const double top = terrain_at(x) + c.height_m;
        s.tops.push_back(P2{x, top});
        if (c.bounds_roof) s.buildings.push_back(RayScratch::Building{c.footprint_id, x, top});
    }
    std::stable_sort(s.buildings.begin(), s.buildings.end(),
                     [](const RayScratch::Building& a, const RayScratch::Building& b) {
                         if (a.id != b.id) return a.id < b.id;
                         return a.x < b.x;
                     });
    for (size_t start = 0; start < s.buildings.size();) {
        size_t end = start;
        while (end < s.buildings.size() && s.buildings[end].id == s.buildings[start].id) ++end;
        for (size_t k = start; k + 1 < end; k += 2)
            s.roofs.push_back(RayScratch::Roof{s.buildings[k].x, s.buildings[k + 1].x, s.buildings[k].top,
                                               s.buildings[k + 1].top});
        start = end;
    }
    clip_roofs_in_closing_order(s.roofs);
    s.distance_m.clear();
    s.altitude_m.clear();
    s.ground.clear();
    auto push = [&](double x, double z, double g) {
        s.distance_m.push_back(x);
        s.altitude_m.push_back(z);
        s.ground.push_back(g);
    };
    push(0.0, terrain[0].z, profile.ground_factor[0]);
    size_t r = 0;
    double covered_to = 0.0;
    for (size_t index = 1; index < terrain.size(); ++index) {
        const double x = terrain[index].x, altitude = terrain[index].z;
        while (r < s.roofs.size()) {
            const RayScratch::Roof& roof = s.roofs[r];
            if (roof.x0 >= x) break;
            push(roof.x0, terrain_at(roof.x0), ground_factor_at(roof.x0));
            push(roof.x0, roof.top0, 0.0);
            push(roof.x1, roof.top1, 0.0);
            push(roof.x1, terrain_at(roof.x1), ground_factor_at(roof.x1));
            covered_to = roof.x1;
            ++r;
        }
        if (x > covered_to) push(x, altitude, profile.ground_factor[index]);
    }
}

// ray.rs:125-191 ray_terms (calm_edge is the popup's display; not on the painter's path)
void ray_terms(const Profile& profile, const std::vector<Crossing>& crossings, const RayEnds& ends,
               const double favourable_probability[PERIODS], const double alpha_db_per_km[BANDS], RayScratch& s,
               RayTerms& out, WorkCount* work) {
    const double horizontal = std::max(profile.horizontal_m, 1.0);
    const double source_altitude = (ends.has_source_ground ? ends.source_ground_m : profile.ground_m[0]) + ends.source_height_m;
    const double slant = std::max(std::hypot(horizontal, ends.receiver_altitude_m - source_altitude), 1.0);
    fill_path(s, profile, crossings, ends);
    const double length = s.distance_m.back();
    VerticalPath path;
    path.profile = VerticalProfile{s.distance_m.data(), s.altitude_m.data(), s.ground.data(), s.distance_m.size()};
    path.source = P2{0.0, source_altitude};
    path.receiver = P2{length, ends.receiver_altitude_m};
    path.source_ground_factor = ends.has_source_ground_factor ? ends.source_ground_factor : profile.ground_factor[0];
    path.terrain_candidates = s.crests.data();
    path.terrain_count = s.crests.size();
    path.obstacle_tops = s.tops.data();
    path.tops_count = s.tops.size();
    state_boundaries(path, s.vertical, out.boundaries, work);
    const P2 source{0.0, source_altitude}, receiver{horizontal, ends.receiver_altitude_m};
    for (int st = 0; st < 2; ++st) {
        const StateRay ray = StateRay::between((State)st, source, receiver);
        foliage_attenuation_db(foliage_depth_m(profile, ray, source, receiver), out.foliage_db[st]);
    }
    double air[BANDS];
    for (int b = 0; b < BANDS; ++b) {
        out.air_db[b] = alpha_db_per_km[b] * slant / 1000.0;
        air[b] = std::exp(-out.air_db[b] * (LN_10 / 10.0));
    }
    double state_energy[2][BANDS];
    for (int st = 0; st < 2; ++st)
        for (int b = 0; b < BANDS; ++b)
            state_energy[st][b] =
                std::exp(-(out.boundaries[st].attenuation_db[b] + out.foliage_db[st][b]) * (LN_10 / 10.0));
    out.transfer.slant_m = slant;
    for (int p = 0; p < PERIODS; ++p) {
        const double pf = favourable_probability[p];
        for (int b = 0; b < BANDS; ++b)
            out.transfer.periods[p][b] = air[b] * (pf * state_energy[1][b] + (1.0 - pf) * state_energy[0][b]);
    }
    for (int st = 0; st < 2; ++st)
        for (int b = 0; b < BANDS; ++b) out.transfer.states[st][b] = air[b] * state_energy[st][b];
}

}  // namespace qmo
