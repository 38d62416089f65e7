// Line sources on the device: the quadrature nodes of one straight piece seen from a receiver
// (CNOSSOS-EU 2.5.3 point sum, uniform in the in-plane angle) with the skyline mask that places
// a wide bucket's nodes, in FP64 (node positions and weights; each node's ray is then an exact
// path). Restates physics/src/line.rs:40-325 (LinePieceGeometry, line_quadrature_nodes,
// push_wide_bucket_nodes, mark_blocked_bins) and popup/src/obstacles/skyline.rs:23-108
// (skyline_arcs) over the planar cell index.
#pragma once

#include "qm_path.hpp"

namespace qm {

constexpr int LINE_NODE_CAP = 64;  // nodes of one piece (beyond: QM_FAILED_CAPACITY)
constexpr int LINE_BUCKET_COUNT = 5;
constexpr double ONE_BUCKET_SPAN_RAD = 3.0 * PI_D / 180.0;
constexpr double WIDE_BUCKET_MIN_AZIMUTH_SPAN_RAD = 3.0 * PI_D / 180.0;
constexpr int WIDE_BUCKET_MASK_BINS = 128;
constexpr double WIDE_BUCKET_PART_MAX_SPAN_RAD = 0.26;
constexpr int WIDE_BUCKET_MAX_PARTS_PER_RUN = 9;
constexpr double LINE_PERPENDICULAR_FLOOR_M = 0.5;
constexpr double POINT_DIVERGENCE_LINEAR = 12.589254117941673;

struct LineNode {
    double along_m;
    float weight_rad;
    uint32_t obstacles_on_ray;
};

struct LineGeometry {
    double start[3], unit[3];
    double length_m, foot_along_m, perpendicular_m, start_angle_rad, end_angle_rad;
};

QM_DEV double wrap_to_pi(double angle) {
    const double a = fmod(angle, 2.0 * PI_D);
    if (a > PI_D) return a - 2.0 * PI_D;
    if (a <= -PI_D) return a + 2.0 * PI_D;
    return a;
}

QM_DEV bool line_geometry(const double start[3], const double end[3], LineGeometry& g) {
    const double along[3] = {end[0] - start[0], end[1] - start[1], end[2] - start[2]};
    g.length_m = sqrt(along[0] * along[0] + along[1] * along[1] + along[2] * along[2]);
    if (g.length_m < 1e-3) return false;
    for (int i = 0; i < 3; ++i) {
        g.start[i] = start[i];
        g.unit[i] = along[i] / g.length_m;
    }
    g.foot_along_m = -(start[0] * g.unit[0] + start[1] * g.unit[1] + start[2] * g.unit[2]);
    double foot[3];
    for (int i = 0; i < 3; ++i) foot[i] = start[i] + g.foot_along_m * g.unit[i];
    g.perpendicular_m = fmax(sqrt(foot[0] * foot[0] + foot[1] * foot[1] + foot[2] * foot[2]), LINE_PERPENDICULAR_FLOOR_M);
    g.start_angle_rad = atan((0.0 - g.foot_along_m) / g.perpendicular_m);
    g.end_angle_rad = atan((g.length_m - g.foot_along_m) / g.perpendicular_m);
    return true;
}
QM_DEV double line_divergence_factor(const LineGeometry& g) {
    return hypot(g.unit[0], g.unit[1]) / (POINT_DIVERGENCE_LINEAR * g.perpendicular_m);
}
QM_DEV double in_plane_angle_at(const LineGeometry& g, double along) {
    return atan((along - g.foot_along_m) / g.perpendicular_m);
}
QM_DEV double along_at_in_plane_angle(const LineGeometry& g, double angle) {
    return clampd(g.foot_along_m + g.perpendicular_m * tan(angle), 0.0, g.length_m);
}
QM_DEV bool along_at_azimuth(const LineGeometry& g, double azimuth, double& out) {
    const double dx = cos(azimuth), dy = sin(azimuth);
    const double ux = g.unit[0] * g.length_m, uy = g.unit[1] * g.length_m;
    const double denominator = dx * uy - dy * ux;
    if (fabs(denominator) < 1e-12) return false;
    out = clampd((dy * g.start[0] - dx * g.start[1]) / denominator, 0.0, 1.0) * g.length_m;
    return true;
}
QM_DEV double line_azimuth_at(const LineGeometry& g, double along) {
    return atan2(g.start[1] + along * g.unit[1], g.start[0] + along * g.unit[0]);
}
QM_DEV double horizontal_range_at(const LineGeometry& g, double along) {
    return hypot(g.start[0] + along * g.unit[0], g.start[1] + along * g.unit[1]);
}

// The wedge's blocked mask, 128 bits.
struct Mask {
    uint32_t w[WIDE_BUCKET_MASK_BINS / 32];
};
QM_DEV bool mask_bit(const Mask& m, int k) { return (m.w[k >> 5] >> (k & 31)) & 1u; }

// mark_blocked_bins (line.rs:294-325) for one arc.
QM_DEV void mark_blocked(const LineGeometry& g, double lo, double hi, double nearest, double need_radius,
                         double span_lo, double span_hi, double bin_width, Mask& m) {
    if (nearest > need_radius) return;
    for (int k = 0; k < 3; ++k) {
        const double shift = k == 0 ? 0.0 : (k == 1 ? 2.0 * PI_D : -2.0 * PI_D);
        const double piece_lo = fmax(lo + shift, span_lo), piece_hi = fmin(hi + shift, span_hi);
        if (piece_hi <= piece_lo) continue;
        double along;
        if (!along_at_azimuth(g, 0.5 * (piece_lo + piece_hi), along)) continue;
        if (horizontal_range_at(g, along) - nearest <= 1.0) continue;
        const int first = (int)fmax(floor((piece_lo - span_lo) / bin_width), 0.0);
        const long long raw = (long long)ceil((piece_hi - span_lo) / bin_width);
        long long last = raw >= 1 ? raw - 1 : 0;
        last = min(max(last, (long long)first), (long long)WIDE_BUCKET_MASK_BINS - 1);
        for (int b = first; b <= (int)last; ++b) m.w[b >> 5] |= 1u << (b & 31);
    }
}

// skyline_arcs (skyline.rs:23-108) of the receiver at (x, y) over [lo, hi] within `radius`, each
// arc marked straight into the mask.
QM_DEV void skyline_mark(const qm_obstacles& O, double x, double y, double lo, double hi, double radius,
                         double los_floor, const LineGeometry& g, double need_radius, double span_lo, double span_hi,
                         double bin_width, Mask& m) {
    const double mps = O.metres_per_step;
    const double mpy = O.metres_per_step_y > 0.0 ? O.metres_per_step_y : mps;
    const double ox = O.origin_steps_x + x / mps, oy = O.origin_steps_y - y / mpy;
    const double cell = (double)QM_CELL_STEPS;
    const double lox = cos(lo), loy = sin(lo), hix = cos(hi), hiy = sin(hi);
    const long long row0 = (long long)floor((oy - radius / mpy) / cell), row1 = (long long)floor((oy + radius / mpy) / cell);
    const long long col0 = (long long)floor((ox - radius / mps) / cell), col1 = (long long)floor((ox + radius / mps) / cell);
    for (long long row = row0; row <= row1; ++row) {
        const double south = (oy - (row + 1) * cell) * mpy, north = (oy - row * cell) * mpy;
        const double dy = fmax(fmax(south, -north), 0.0);
        for (long long column = col0; column <= col1; ++column) {
            const double west = (column * cell - ox) * mps, east = ((column + 1) * cell - ox) * mps;
            const double dx = fmax(fmax(west, -east), 0.0);
            if (dx * dx + dy * dy > radius * radius) continue;
            const double cx[4] = {west, east, west, east}, cy[4] = {south, south, north, north};
            bool below = true, above = true;
            for (int k = 0; k < 4; ++k) {
                below = below && (lox * cy[k] - loy * cx[k] < 0.0);
                above = above && (cx[k] * hiy - cy[k] * hix < 0.0);
            }
            if (below || above) continue;
            const long long gx = column - O.cell_x0, gy = row - O.cell_y0;
            if (gx < 0 || gy < 0 || gx >= (long long)O.cells_x || gy >= (long long)O.cells_y) continue;
            const size_t index = (size_t)gy * O.cells_x + (size_t)gx;
            const uint32_t r0 = O.cell_first_run[index], r1 = O.cell_first_run[index + 1];
            double tallest = 0.0;
            for (uint32_t r = r0; r < r1; ++r) tallest = fmax(tallest, O.outlines[O.runs[r].outline].height_dm * 0.1);
            if (tallest <= los_floor) continue;
            for (uint32_t r = r0; r < r1; ++r) {
                const qm_run run = O.runs[r];
                const qm_outline& record = O.outlines[run.outline];
                if (record.kind != QM_KIND_BUILDING && record.height_dm * 0.1 <= los_floor) continue;
                const uint32_t first = record.first_vertex + run.first_edge;
                double ax = ((double)O.vertices[2 * first] - ox) * mps, ay = (oy - (double)O.vertices[2 * first + 1]) * mpy;
                double azimuth_a = atan2(ay, ax);
                for (uint32_t next = first + 1; next <= first + run.edge_count; ++next) {
                    const double bx = ((double)O.vertices[2 * next] - ox) * mps;
                    const double by = (oy - (double)O.vertices[2 * next + 1]) * mpy;
                    const double azimuth_b = atan2(by, bx);
                    // origin_to_segment_distance (skyline.rs:8-20)
                    const double ex = bx - ax, ey = by - ay, length2 = ex * ex + ey * ey;
                    const double t = length2 > 0.0 ? clampd(-(ax * ex + ay * ey) / length2, 0.0, 1.0) : 0.0;
                    const double px = ax + t * ex, py = ay + t * ey;
                    const double nearest = sqrt(px * px + py * py);
                    if (nearest >= 1e-6 && nearest <= radius) {
                        const double unwrapped_b = azimuth_a + wrap_to_pi(azimuth_b - azimuth_a);
                        mark_blocked(g, fmin(azimuth_a, unwrapped_b), fmax(azimuth_a, unwrapped_b), nearest, need_radius,
                                     span_lo, span_hi, bin_width, m);
                    }
                    ax = bx;
                    ay = by;
                    azimuth_a = azimuth_b;
                }
            }
        }
    }
}

QM_DEV bool push_node(LineNode* nodes, int& n, double along, double weight, bool obstacles) {
    if (n >= LINE_NODE_CAP) return false;
    nodes[n++] = LineNode{along, (float)weight, obstacles ? 1u : 0u};
    return true;
}

// push_wide_bucket_nodes (line.rs:205-290); false when the node buffer is full.
QM_DEV bool wide_bucket_nodes(const qm_obstacles& O, double rx, double ry, double los_floor, const LineGeometry& g,
                              double angle_lo, double angle_hi, LineNode* nodes, int& n, bool& placed) {
    placed = false;
    const double along_lo = along_at_in_plane_angle(g, angle_lo), along_hi = along_at_in_plane_angle(g, angle_hi);
    const double azimuth_a = line_azimuth_at(g, along_lo);
    const double turn = wrap_to_pi(line_azimuth_at(g, along_hi) - azimuth_a);
    const double span = fabs(turn);
    if (span < WIDE_BUCKET_MIN_AZIMUTH_SPAN_RAD) return true;
    const double span_lo = turn < 0.0 ? azimuth_a + turn : azimuth_a;
    const double span_hi = turn < 0.0 ? azimuth_a : azimuth_a + turn;
    const double range_lo = horizontal_range_at(g, along_lo), range_hi = horizontal_range_at(g, along_hi);
    const double chord = hypot((g.start[0] + along_lo * g.unit[0]) - (g.start[0] + along_hi * g.unit[0]),
                               (g.start[1] + along_lo * g.unit[1]) - (g.start[1] + along_hi * g.unit[1]));
    const double centre_range = horizontal_range_at(g, along_at_in_plane_angle(g, 0.5 * (angle_lo + angle_hi)));
    const double need_radius = fmin(fmin(range_lo, range_hi), centre_range) + chord;
    const double bin_width = span / WIDE_BUCKET_MASK_BINS;
    Mask m = {{0u, 0u, 0u, 0u}};
    skyline_mark(O, rx, ry, span_lo, span_hi, need_radius, los_floor, g, need_radius, span_lo, span_hi, bin_width, m);
    if (!(m.w[0] | m.w[1] | m.w[2] | m.w[3])) return true;
    placed = true;
    const double lo_clamp = fmin(along_lo, along_hi), hi_clamp = fmax(along_lo, along_hi);
    const double angle_at_span_lo = turn < 0.0 ? angle_hi : angle_lo;
    const double angle_at_span_hi = turn < 0.0 ? angle_lo : angle_hi;
    int bin = 0;
    while (bin < WIDE_BUCKET_MASK_BINS) {
        const bool run_blocked = mask_bit(m, bin);
        int run_end = bin;
        while (run_end < WIDE_BUCKET_MASK_BINS && mask_bit(m, run_end) == run_blocked) ++run_end;
        const double run_lo = span_lo + bin * bin_width;
        const double run_hi = run_end == WIDE_BUCKET_MASK_BINS ? span_hi : span_lo + run_end * bin_width;
        const int parts = (int)clampd(ceil((run_hi - run_lo) / WIDE_BUCKET_PART_MAX_SPAN_RAD), 1.0,
                                      (double)WIDE_BUCKET_MAX_PARTS_PER_RUN);
        const double step = (run_hi - run_lo) / parts;
        for (int part = 0; part < parts; ++part) {
            const double part_lo = run_lo + part * step;
            const double part_hi = (part + 1 == parts && run_end == WIDE_BUCKET_MASK_BINS) ? span_hi : part_lo + step;
            double edge_lo = angle_at_span_lo, edge_hi = angle_at_span_hi, along;
            if (!(bin == 0 && part == 0) && along_at_azimuth(g, part_lo, along))
                edge_lo = in_plane_angle_at(g, clampd(along, lo_clamp, hi_clamp));
            if (!(run_end == WIDE_BUCKET_MASK_BINS && part + 1 == parts) && along_at_azimuth(g, part_hi, along))
                edge_hi = in_plane_angle_at(g, clampd(along, lo_clamp, hi_clamp));
            if (!along_at_azimuth(g, 0.5 * (part_lo + part_hi), along))
                along = along_at_in_plane_angle(g, 0.5 * (edge_lo + edge_hi));
            if (!push_node(nodes, n, along, fabs(edge_hi - edge_lo), run_blocked)) return false;
        }
        bin = run_end;
    }
    return true;
}

// line_quadrature_nodes (line.rs:181-203) with the scene's skyline; false when more than
// LINE_NODE_CAP nodes were needed.
QM_DEV bool line_nodes(const qm_obstacles& O, double rx, double ry, double los_floor, const LineGeometry& g,
                       LineNode* nodes, int& n) {
    n = 0;
    const double span = g.end_angle_rad - g.start_angle_rad;
    const int buckets = span < ONE_BUCKET_SPAN_RAD ? 1 : LINE_BUCKET_COUNT;
    const double bucket_angle = span / buckets;
    for (int bucket = 0; bucket < buckets; ++bucket) {
        const double angle_lo = g.start_angle_rad + bucket * bucket_angle;
        const double angle_hi = angle_lo + bucket_angle;
        bool placed;
        if (!wide_bucket_nodes(O, rx, ry, los_floor, g, angle_lo, angle_hi, nodes, n, placed)) return false;
        if (!placed && !push_node(nodes, n, along_at_in_plane_angle(g, angle_lo + 0.5 * bucket_angle),
                                  angle_hi - angle_lo, true))
            return false;
    }
    return true;
}

}  // namespace qm
