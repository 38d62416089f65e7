// TEST INFRASTRUCTURE — FP64 oracle: the line-source quadrature (CNOSSOS-EU point sum of a
// straight piece, Directive 2015/996 2.5.3) and the receiver skyline that places its wide
// buckets' nodes, restated from physics/src/line.rs:1-328 and popup/src/obstacles/skyline.rs:1-108,
// and a line pair's received energy (popup/src/evaluate.rs:171-267, one ray per node). Not part
// of the product path.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "scene.hpp"

namespace qmo {

// physics/src/line.rs:14-37
constexpr int LINE_BUCKET_COUNT = 5;
const double ONE_BUCKET_SPAN_RAD = 3.0 * M_PI / 180.0;
const double WIDE_BUCKET_MIN_AZIMUTH_SPAN_RAD = 3.0 * M_PI / 180.0;
constexpr int WIDE_BUCKET_MASK_BINS = 128;
constexpr double WIDE_BUCKET_PART_MAX_SPAN_RAD = 0.26;
constexpr int WIDE_BUCKET_MAX_PARTS_PER_RUN = 9;
constexpr double LINE_PERPENDICULAR_FLOOR_M = 0.5;
constexpr double POINT_DIVERGENCE_LINEAR = 12.589254117941673;

// line.rs:40-50
double wrap_to_pi(double angle) {
    const double a = std::fmod(angle, 2.0 * M_PI);
    if (a > M_PI) return a - 2.0 * M_PI;
    if (a <= -M_PI) return a + 2.0 * M_PI;
    return a;
}

// line.rs:52-188 LinePieceGeometry (receiver-relative, x east, y north, z up).
bool line_geometry(const double start[3], const double end[3], LineGeometry& g) {
    const double along[3] = {end[0] - start[0], end[1] - start[1], end[2] - start[2]};
    g.length_m = std::sqrt(along[0] * along[0] + along[1] * along[1] + along[2] * along[2]);
    if (g.length_m < 1e-3) return false;
    for (int i = 0; i < 3; ++i) {
        g.start[i] = start[i];
        g.unit[i] = along[i] / g.length_m;
    }
    g.foot_along_m = -(start[0] * g.unit[0] + start[1] * g.unit[1] + start[2] * g.unit[2]);
    double foot[3];
    for (int i = 0; i < 3; ++i) foot[i] = start[i] + g.foot_along_m * g.unit[i];
    g.perpendicular_m =
        std::max(std::sqrt(foot[0] * foot[0] + foot[1] * foot[1] + foot[2] * foot[2]), LINE_PERPENDICULAR_FLOOR_M);
    g.start_angle_rad = std::atan((0.0 - g.foot_along_m) / g.perpendicular_m);
    g.end_angle_rad = std::atan((g.length_m - g.foot_along_m) / g.perpendicular_m);
    return true;
}
double line_subtended(const LineGeometry& g) { return g.end_angle_rad - g.start_angle_rad; }
double line_divergence_factor(const LineGeometry& g) {
    return std::hypot(g.unit[0], g.unit[1]) / (POINT_DIVERGENCE_LINEAR * g.perpendicular_m);
}
static double in_plane_angle_at(const LineGeometry& g, double along) {
    return std::atan((along - g.foot_along_m) / g.perpendicular_m);
}
static double along_at_in_plane_angle(const LineGeometry& g, double angle) {
    return std::clamp(g.foot_along_m + g.perpendicular_m * std::tan(angle), 0.0, g.length_m);
}
static void point_at(const LineGeometry& g, double along, double p[3]) {
    for (int i = 0; i < 3; ++i) p[i] = g.start[i] + along * g.unit[i];
}
static bool along_at_azimuth(const LineGeometry& g, double azimuth, double& out) {
    const double dx = std::cos(azimuth), dy = std::sin(azimuth);
    const double ux = g.unit[0] * g.length_m, uy = g.unit[1] * g.length_m;
    const double denominator = dx * uy - dy * ux;
    if (std::fabs(denominator) < 1e-12) return false;
    const double fraction = (dy * g.start[0] - dx * g.start[1]) / denominator;
    out = std::clamp(fraction, 0.0, 1.0) * g.length_m;
    return true;
}
static double azimuth_at(const LineGeometry& g, double along) {
    double p[3];
    point_at(g, along, p);
    return std::atan2(p[1], p[0]);
}
static double horizontal_range_at(const LineGeometry& g, double along) {
    double p[3];
    point_at(g, along, p);
    return std::hypot(p[0], p[1]);
}

// line.rs:294-325 mark_blocked_bins
static void mark_blocked_bins(const LineGeometry& g, const SkylineArc& arc, double need_radius, double span_lo,
                              double span_hi, double bin_width, bool* blocked) {
    if (arc.nearest_m > need_radius) return;
    for (double shift : {0.0, 2.0 * M_PI, -2.0 * M_PI}) {
        const double piece_lo = std::max(arc.lo_rad + shift, span_lo);
        const double piece_hi = std::min(arc.hi_rad + shift, span_hi);
        if (piece_hi <= piece_lo) continue;
        double along;
        if (!along_at_azimuth(g, 0.5 * (piece_lo + piece_hi), along)) continue;
        if (horizontal_range_at(g, along) - arc.nearest_m <= 1.0) continue;
        const int first = (int)std::max(std::floor((piece_lo - span_lo) / bin_width), 0.0);
        long last_raw = (long)std::ceil((piece_hi - span_lo) / bin_width);
        long last = last_raw >= 1 ? last_raw - 1 : 0;  // saturating_sub(1)
        last = std::min<long>(std::max<long>(last, first), WIDE_BUCKET_MASK_BINS - 1);
        for (long k = first; k <= last; ++k) blocked[k] = true;
    }
}

// line.rs:205-290 push_wide_bucket_nodes
static void push_wide_bucket_nodes(const LineGeometry& g, double angle_lo, double angle_hi, const SkylineQuery& skyline,
                                   std::vector<LineNode>& nodes) {
    const double along_lo = along_at_in_plane_angle(g, angle_lo), along_hi = along_at_in_plane_angle(g, angle_hi);
    const double azimuth_a = azimuth_at(g, along_lo);
    const double turn = wrap_to_pi(azimuth_at(g, along_hi) - azimuth_a);
    const double span = std::fabs(turn);
    if (span < WIDE_BUCKET_MIN_AZIMUTH_SPAN_RAD) return;
    const double span_lo = turn < 0.0 ? azimuth_a + turn : azimuth_a;
    const double span_hi = turn < 0.0 ? azimuth_a : azimuth_a + turn;
    const double range_lo = horizontal_range_at(g, along_lo), range_hi = horizontal_range_at(g, along_hi);
    double a[3], b[3];
    point_at(g, along_lo, a);
    point_at(g, along_hi, b);
    const double chord = std::hypot(a[0] - b[0], a[1] - b[1]);
    const double centre_range = horizontal_range_at(g, along_at_in_plane_angle(g, 0.5 * (angle_lo + angle_hi)));
    const double need_radius = std::min(std::min(range_lo, range_hi), centre_range) + chord;
    const double bin_width = span / WIDE_BUCKET_MASK_BINS;
    bool blocked[WIDE_BUCKET_MASK_BINS] = {};
    skyline(span_lo, span_hi, need_radius, [&](const SkylineArc& arc) {
        mark_blocked_bins(g, arc, need_radius, span_lo, span_hi, bin_width, blocked);
    });
    bool any = false;
    for (bool x : blocked) any = any || x;
    if (!any) return;
    const double lo_clamp = std::min(along_lo, along_hi), hi_clamp = std::max(along_lo, along_hi);
    auto angle_at_azimuth = [&](double azimuth, double fallback) {
        double along;
        if (!along_at_azimuth(g, azimuth, along)) return fallback;
        return in_plane_angle_at(g, std::clamp(along, lo_clamp, hi_clamp));
    };
    const double angle_at_span_lo = turn < 0.0 ? angle_hi : angle_lo;
    const double angle_at_span_hi = turn < 0.0 ? angle_lo : angle_hi;
    int bin = 0;
    while (bin < WIDE_BUCKET_MASK_BINS) {
        const bool run_blocked = blocked[bin];
        int run_end = bin;
        while (run_end < WIDE_BUCKET_MASK_BINS && blocked[run_end] == run_blocked) ++run_end;
        const double run_lo = span_lo + bin * bin_width;
        const double run_hi = run_end == WIDE_BUCKET_MASK_BINS ? span_hi : span_lo + run_end * bin_width;
        const int parts = (int)std::clamp(std::ceil((run_hi - run_lo) / WIDE_BUCKET_PART_MAX_SPAN_RAD), 1.0,
                                          (double)WIDE_BUCKET_MAX_PARTS_PER_RUN);
        const double step = (run_hi - run_lo) / parts;
        for (int part = 0; part < parts; ++part) {
            const double part_lo = run_lo + part * step;
            const double part_hi = (part + 1 == parts && run_end == WIDE_BUCKET_MASK_BINS) ? span_hi : part_lo + step;
            const double edge_lo = (bin == 0 && part == 0) ? angle_at_span_lo : angle_at_azimuth(part_lo, angle_at_span_lo);
            const double edge_hi = (run_end == WIDE_BUCKET_MASK_BINS && part + 1 == parts)
                                       ? angle_at_span_hi
                                       : angle_at_azimuth(part_hi, angle_at_span_hi);
            double along;
            if (!along_at_azimuth(g, 0.5 * (part_lo + part_hi), along))
                along = along_at_in_plane_angle(g, 0.5 * (edge_lo + edge_hi));
            nodes.push_back(LineNode{along, std::fabs(edge_hi - edge_lo), run_blocked});
        }
        bin = run_end;
    }
}

// line.rs:181-203 line_quadrature_nodes
void line_quadrature_nodes(const LineGeometry& g, const SkylineQuery& skyline, std::vector<LineNode>& nodes) {
    nodes.clear();
    const double span = line_subtended(g);
    const int buckets = span < ONE_BUCKET_SPAN_RAD ? 1 : LINE_BUCKET_COUNT;
    const double bucket_angle = span / buckets;
    for (int bucket = 0; bucket < buckets; ++bucket) {
        const double angle_lo = g.start_angle_rad + bucket * bucket_angle;
        const double angle_hi = angle_lo + bucket_angle;
        const size_t before = nodes.size();
        push_wide_bucket_nodes(g, angle_lo, angle_hi, skyline, nodes);
        if (nodes.size() == before)
            nodes.push_back(LineNode{along_at_in_plane_angle(g, angle_lo + 0.5 * bucket_angle), angle_hi - angle_lo, true});
    }
}

// popup/src/obstacles/skyline.rs:8-20
static double origin_to_segment_distance(const double a[2], const double b[2]) {
    const double ex = b[0] - a[0], ey = b[1] - a[1];
    const double length2 = ex * ex + ey * ey;
    const double t = length2 > 0.0 ? std::clamp(-(a[0] * ex + a[1] * ey) / length2, 0.0, 1.0) : 0.0;
    const double x = a[0] + t * ex, y = a[1] + t * ey;
    return std::sqrt(x * x + y * y);
}

// skyline.rs:23-108 Scene::skyline_arcs over the planar cell index (cells of QM_CELL_STEPS steps,
// rows growing south); a cell's tallest outline is found from its runs.
void skyline_arcs(const qm_obstacles& O, const double origin[2], double lo_rad, double hi_rad, double radius_m,
                  double los_floor_m, const std::function<void(const SkylineArc&)>& visit) {
    double steps[2];
    steps_of(O, origin[0], origin[1], steps);
    const double ox = steps[0], oy = steps[1];
    const double mps = O.metres_per_step;
    const double mpy = O.metres_per_step_y > 0.0 ? O.metres_per_step_y : mps;
    const double cell = QM_CELL_STEPS;
    const double low[2] = {std::cos(lo_rad), std::sin(lo_rad)}, high[2] = {std::cos(hi_rad), std::sin(hi_rad)};
    auto relative = [&](int32_t vx, int32_t vy, double out[2]) {
        out[0] = ((double)vx - ox) * mps;
        out[1] = (oy - (double)vy) * mpy;
    };
    const long row0 = (long)std::floor((oy - radius_m / mpy) / cell), row1 = (long)std::floor((oy + radius_m / mpy) / cell);
    const long col0 = (long)std::floor((ox - radius_m / mps) / cell), col1 = (long)std::floor((ox + radius_m / mps) / cell);
    for (long row = row0; row <= row1; ++row) {
        const double south = (oy - (row + 1) * cell) * mpy;
        const double north = (oy - row * cell) * mpy;
        const double dy = std::max(std::max(south, -north), 0.0);
        for (long column = col0; column <= col1; ++column) {
            const double west = (column * cell - ox) * mps;
            const double east = ((column + 1) * cell - ox) * mps;
            const double dx = std::max(std::max(west, -east), 0.0);
            if (dx * dx + dy * dy > radius_m * radius_m) continue;
            const double corners[4][2] = {{west, south}, {east, south}, {west, north}, {east, north}};
            bool below = true, above = true;
            for (const auto& c : corners) {
                below = below && (low[0] * c[1] - low[1] * c[0] < 0.0);
                above = above && (c[0] * high[1] - c[1] * high[0] < 0.0);
            }
            if (below || above) continue;
            const long cx = column - O.cell_x0, cy = row - O.cell_y0;
            if (cx < 0 || cy < 0 || cx >= (long)O.cells_x || cy >= (long)O.cells_y) continue;  // not read
            const size_t index = (size_t)cy * O.cells_x + (size_t)cx;
            const uint32_t r0 = O.cell_first_run[index], r1 = O.cell_first_run[index + 1];
            double tallest = 0.0;
            for (uint32_t r = r0; r < r1; ++r) tallest = std::max(tallest, O.outlines[O.runs[r].outline].height_dm * 0.1);
            if (tallest <= los_floor_m) continue;
            for (uint32_t r = r0; r < r1; ++r) {
                const qm_run& run = O.runs[r];
                const qm_outline& record = O.outlines[run.outline];
                if (record.kind != QM_KIND_BUILDING && record.height_dm * 0.1 <= los_floor_m) continue;
                const uint32_t first = record.first_vertex + run.first_edge;
                double a[2];
                relative(O.vertices[2 * first], O.vertices[2 * first + 1], a);
                double azimuth_a = std::atan2(a[1], a[0]);
                for (uint32_t next = first + 1; next <= first + run.edge_count; ++next) {
                    double b[2];
                    relative(O.vertices[2 * next], O.vertices[2 * next + 1], b);
                    const double azimuth_b = std::atan2(b[1], b[0]);
                    const double nearest_m = origin_to_segment_distance(a, b);
                    if (nearest_m >= 1e-6 && nearest_m <= radius_m) {
                        const double unwrapped_b = azimuth_a + wrap_to_pi(azimuth_b - azimuth_a);
                        visit(SkylineArc{std::min(azimuth_a, unwrapped_b), std::max(azimuth_a, unwrapped_b), nearest_m});
                    }
                    a[0] = b[0];
                    a[1] = b[1];
                    azimuth_a = azimuth_b;
                }
            }
        }
    }
}

// popup/src/evaluate.rs:171-267 source_rays for a line piece: every quadrature node a ray
// from its point to the receiver, weighted by its dphi times the piece's divergence factor, the
// wall search skipped inside the skyline's clear gaps; received as add_ray (evaluate.rs:119-152).
int evaluate_line_pair(const Scene& S, const Receiver& receiver, const qm_line_piece& piece, PairScratch& s,
                       PairResult& out, std::vector<LineNode>& nodes, WorkCount* work) {
    for (int p = 0; p < PERIODS; ++p) {
        out.mixed[p] = 0.0;
        out.states[0][p] = out.states[1][p] = 0.0;
        for (int b = 0; b < BANDS; ++b) out.bands[p][b] = 0.0;
    }
    for (int st = 0; st < 2; ++st)
        for (int b = 0; b < BANDS; ++b) out.boundary[st][b] = 0.0;
    // candidate.traced_ground_m (candidates.rs:151-158): the deck at the piece's ends, else the
    // terrain there.
    GroundSample ga, gb;
    int failed = ground_at(S.terrain, piece.ax_m, piece.ay_m, ga);
    if (!failed) failed = ground_at(S.terrain, piece.bx_m, piece.by_m, gb);
    if (failed) return failed;
    const double ground_a = piece.on_deck ? (double)piece.deck_a_m : ga.height_m;
    const double ground_b = piece.on_deck ? (double)piece.deck_b_m : gb.height_m;
    const double start[3] = {piece.ax_m - receiver.x, piece.ay_m - receiver.y,
                             ground_a + piece.height_m - receiver.altitude_m};
    const double end[3] = {piece.bx_m - receiver.x, piece.by_m - receiver.y,
                           ground_b + piece.height_m - receiver.altitude_m};
    LineGeometry g;
    if (!line_geometry(start, end, g)) return 0;
    const double origin[2] = {receiver.x, receiver.y};
    const double floor_m = std::max((double)piece.height_m, 0.0);
    line_quadrature_nodes(
        g,
        [&](double lo, double hi, double radius, const std::function<void(const SkylineArc&)>& visit) {
            skyline_arcs(S.obstacles, origin, lo, hi, radius, floor_m, visit);
        },
        nodes);
    const double divergence = line_divergence_factor(g);
    for (const LineNode& node : nodes) {
        const double fraction = node.along_m / g.length_m;
        const double point[2] = {piece.ax_m + fraction * (piece.bx_m - piece.ax_m),
                                 piece.ay_m + fraction * (piece.by_m - piece.ay_m)};
        const double rcv[2] = {receiver.x, receiver.y};
        failed = fill_profile(S.terrain, point, rcv, s.profile);
        if (failed) return failed;
        if (work) work->vertices += s.profile.t.size();
        s.crossings.clear();
        if (node.obstacles_on_ray) {
            failed = crossings_raw(S.obstacles, point, rcv, s.raw, work);
            if (failed) return failed;
            for (const RawCrossing& r : s.raw)
                if (receiver.own_footprint == 0 || r.c.footprint_id != receiver.own_footprint) s.crossings.push_back(r.c);
        }
        const double azimuth = std::atan2(receiver.y - point[1], receiver.x - point[0]);
        double p[PERIODS];
        for (int period = 0; period < PERIODS; ++period) p[period] = favourable_at(receiver.favourable, period, azimuth);
        RayEnds ends;
        ends.source_height_m = piece.height_m;
        // evaluate.rs:182-186 deck_at(fraction): a node on a deck stands on it.
        ends.has_source_ground = piece.on_deck != 0;
        ends.source_ground_m =
            piece.on_deck ? (double)piece.deck_a_m + fraction * ((double)piece.deck_b_m - (double)piece.deck_a_m) : 0.0;
        ends.receiver_altitude_m = receiver.altitude_m;
        ends.has_source_ground_factor = false;
        ends.source_ground_factor = 0.0;
        ends.platform_half_width_m = 0.0;
        ends.own_footprint = 0;
        ray_terms(s.profile, s.crossings, ends, p, receiver.alpha, s.ray, s.terms, work);
        const double weight = node.weight_rad * divergence;
        const Transfer& tr = s.terms.transfer;
        for (int period = 0; period < PERIODS; ++period)
            for (int b = 0; b < BANDS; ++b) {
                const double power = weight * (double)piece.energy_per_m[period][b];
                out.bands[period][b] += power * tr.periods[period][b];
                for (int st = 0; st < 2; ++st) out.states[st][period] += power * tr.states[st][b];
            }
    }
    for (int period = 0; period < PERIODS; ++period) {
        double sum = 0.0;
        for (int b = 0; b < BANDS; ++b) sum += out.bands[period][b];
        out.mixed[period] = sum;
    }
    return 0;
}

}  // namespace qmo
