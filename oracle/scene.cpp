// TEST INFRASTRUCTURE — FP64 oracle: the ground under a ray, its obstacle crossings, the weather of
// a receiver and one receiver-source pair end to end, restated from the reference's Rust
// (popup/src/scene.rs, popup/src/obstacles/crossings.rs, physics/src/weather.rs,
// popup/src/evaluate.rs, paint/src/exact.rs) on the planar scene of qm_scene.h. Not part of the
// product path.
#include "scene.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>

namespace qmo {

// tiles/src/terrain.rs:36-38
static double height_m_of_code(uint32_t code) { return -500.0 + (double)code / 5.0; }

struct Node {
    uint32_t height_code, impervious, forest, canopy;
};

static bool node_at(const qm_terrain& T, int64_t row, int64_t column, Node& out) {
    if (row < 0 || column < 0 || row >= (int64_t)T.rows || column >= (int64_t)T.columns) return false;
    const size_t index = (size_t)row * T.columns + (size_t)column;
    const uint32_t v = T.nodes[index];
    out = Node{v & 0xffffu, (v >> 16) & 0xffu, (v >> 24) & 0xffu, T.canopy_m ? (uint32_t)T.canopy_m[index] : 0u};
    return true;
}

static bool node_valid(const Node& n) { return n.height_code != 0xffffu && n.impervious <= 100u; }

// tiles/src/terrain.rs:342-385 Terrain::sample_at at lattice position (cp, rp), the fractions
// measured north and east from the cell's south-west node.
static int cell_sample(const qm_terrain& T, double cp, double rp, GroundSample& out) {
    const double column = std::floor(cp), row = std::floor(rp);
    const double cf = cp - column, rf = rp - row;
    const int64_t c = (int64_t)column, r = (int64_t)row;
    const int64_t nc = c + (cf > 0.0 ? 1 : 0), nr = r + (rf > 0.0 ? 1 : 0);
    Node corners[4];
    if (!node_at(T, r, c, corners[0]) || !node_at(T, r, nc, corners[1]) || !node_at(T, nr, c, corners[2]) ||
        !node_at(T, nr, nc, corners[3]))
        return QM_FAILED_NO_TERRAIN;
    const Node& nearest = corners[(rf >= 0.5 ? 2 : 0) + (cf >= 0.5 ? 1 : 0)];
    if (nearest.forest > 100u) return QM_FAILED_NO_TERRAIN;
    for (const Node& n : corners)
        if (!node_valid(n)) return QM_FAILED_NO_TERRAIN;
    const double w[4] = {(1.0 - rf) * (1.0 - cf), (1.0 - rf) * cf, rf * (1.0 - cf), rf * cf};
    double height = 0.0, impervious = 0.0;
    for (int k = 0; k < 4; ++k) {
        height += w[k] * height_m_of_code(corners[k].height_code);
        impervious += w[k] * (double)corners[k].impervious;
    }
    out.height_m = height;
    out.ground_factor = 1.0 - impervious / 100.0;
    out.forest_cover = (double)nearest.forest / 100.0;
    out.canopy_m = (double)nearest.canopy;
    return 0;
}

// Fine terrain (tiles/src/terrain_fine.rs): detail in steps of QM_DETAIL_STEP_M on the lattice
// QM_FINE_PER_NODE times as fine, read where the block of QM_FINE_BLOCK_CELLS square cells holding
// the point has fine terrain (the reference's tiles with a fine_offset).
static uint32_t fine_rows(const qm_terrain& T) { return QM_FINE_PER_NODE * (T.rows - 1) + 1; }
static uint32_t fine_columns(const qm_terrain& T) { return QM_FINE_PER_NODE * (T.columns - 1) + 1; }

static bool fine_block(const qm_terrain& T, double cp, double rp) {
    if (!T.fine_detail || !T.fine_blocks) return false;
    const int64_t bx_n = ((int64_t)T.columns - 2) / QM_FINE_BLOCK_CELLS + 1;
    const int64_t by_n = ((int64_t)T.rows - 2) / QM_FINE_BLOCK_CELLS + 1;
    const int64_t bx = std::min(std::max((int64_t)std::floor(cp / QM_FINE_BLOCK_CELLS), (int64_t)0), bx_n - 1);
    const int64_t by = std::min(std::max((int64_t)std::floor(rp / QM_FINE_BLOCK_CELLS), (int64_t)0), by_n - 1);
    return T.fine_blocks[by * bx_n + bx] != 0;
}

// paint-gpu/kernels/terrain.cuh:91-103 fine_detail_at (tiles/src/terrain_fine.rs): the detail's
// bilinear value at fine position (row, column).
static double detail_at(const qm_terrain& T, double row, double column) {
    const uint32_t rows = fine_rows(T), columns = fine_columns(T);
    row = std::max(row, 0.0);
    column = std::max(column, 0.0);
    const uint32_t r = std::min((uint32_t)std::floor(row), rows - 2), c = std::min((uint32_t)std::floor(column), columns - 2);
    const double fr = row - r, fc = column - c;
    const int16_t* d = T.fine_detail;
    const size_t at = (size_t)r * columns + c;
    const double sum = (1.0 - fr) * ((1.0 - fc) * d[at] + fc * d[at + 1]) +
                       fr * ((1.0 - fc) * d[at + columns] + fc * d[at + columns + 1]);
    return sum * QM_DETAIL_STEP_M;
}

// The geographic frame (qm_terrain.row_y_m, column_m): the one arc-second rows at the table's
// north metres, fine rows evenly between them (popup/src/scene.rs:303-330 row_line_y,
// rows_north_of), the columns evenly. Without a table, row r at origin_y + r * spacing.
static double column_metres(const qm_terrain& T) { return T.column_m > 0.0 ? T.column_m : T.spacing_m; }
static double table_row_y(const qm_terrain& T, int64_t k) {
    const int64_t last = (int64_t)T.rows - 1;
    if (k < 0) return T.row_y_m[0] + (double)k * (T.row_y_m[1] - T.row_y_m[0]);
    if (k > last) return T.row_y_m[last] + (double)(k - last) * (T.row_y_m[last] - T.row_y_m[last - 1]);
    return T.row_y_m[k];
}
// North metres of row line k of the lattice `scale` times as fine (a table's rows only).
static double row_line_y(const qm_terrain& T, int scale, int64_t k) {
    const int64_t coarse = k >= 0 ? k / scale : -((-k + scale - 1) / scale);
    const int64_t sub = k - coarse * scale;
    const double y0 = table_row_y(T, coarse);
    if (sub == 0) return y0;
    return y0 + (table_row_y(T, coarse + 1) - y0) * (double)sub / (double)scale;
}
static double column_position(const qm_terrain& T, double x) {
    return T.column_m > 0.0 ? (x - T.origin_x_m) / T.column_m : (x - T.origin_x_m) / T.spacing_m;
}
// The fractional row (node index) of north metres y: the table's rows about it.
static double row_position(const qm_terrain& T, double y) {
    if (!T.row_y_m) return (y - T.origin_y_m) / T.spacing_m;
    const int64_t last = (int64_t)T.rows - 1;
    int64_t lo = 0;
    if (y < T.row_y_m[0]) {
        lo = 0;
    } else if (y >= T.row_y_m[last]) {
        lo = last - 1;
    } else {
        int64_t hi = last;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (T.row_y_m[mid] <= y) lo = mid;
            else hi = mid;
        }
    }
    return (double)lo + (y - T.row_y_m[lo]) / (T.row_y_m[lo + 1] - T.row_y_m[lo]);
}

// popup/src/scene.rs Ground::at: the 1/6" height where the point's block has fine terrain.
int ground_at(const qm_terrain& T, double x, double y, GroundSample& out) {
    const double cp = column_position(T, x);
    const double rp = row_position(T, y);
    const int failed = cell_sample(T, cp, rp, out);
    if (failed) return failed;
    if (fine_block(T, cp, rp)) out.height_m += detail_at(T, QM_FINE_PER_NODE * rp, QM_FINE_PER_NODE * cp);
    return 0;
}

// popup/src/scene.rs Ground::fine_sample (paint-gpu/kernels/terrain_walk.cuh:105-160): a vertex of
// the fine walk on fine row line `line` (on_row) or fine column line `line` between fine rows
// `south` and south + 1, `f` along the line: the coarse cell's sample at the vertex's place (its
// fine indices over QM_FINE_PER_NODE) plus, in a block with fine terrain, the detail between the
// line's two fine nodes.
static int fine_sample(const qm_terrain& T, bool on_row, int64_t line, int64_t south, double position, double f,
                       GroundSample& out) {
    const int64_t rows = fine_rows(T), columns = fine_columns(T);
    int64_t r0, c0;
    if (on_row) {
        r0 = line;
        c0 = (int64_t)std::floor(position);
    } else {
        r0 = south;
        c0 = line;
    }
    const int64_t r1 = on_row ? r0 : r0 + 1, c1 = on_row ? c0 + 1 : c0;
    if (r0 < 0 || c0 < 0 || r0 >= rows || c0 >= columns || (f > 0.0 && (r1 >= rows || c1 >= columns)))
        return QM_FAILED_NO_TERRAIN;
    const double rp = (on_row ? (double)r0 : (double)r0 + f) / QM_FINE_PER_NODE;
    const double cp = (on_row ? (double)c0 + f : (double)c0) / QM_FINE_PER_NODE;
    const int failed = cell_sample(T, cp, rp, out);
    if (failed) return failed;
    if (fine_block(T, cp, rp)) {
        const int16_t* d = T.fine_detail;
        double d0 = f < 1.0 ? (double)d[r0 * columns + c0] : 0.0;
        double d1 = f > 0.0 ? (double)d[r1 * columns + c1] : 0.0;
        if (f <= 0.0) d1 = d0;
        if (f >= 1.0) d0 = d1;
        out.height_m += (d0 + f * (d1 - d0)) * QM_DETAIL_STEP_M;
    }
    return 0;
}

// popup/src/scene.rs:532-561 edge_sample
static int edge_sample(bool has_a, const Node& a_in, bool has_b, const Node& b_in, double f, GroundSample& out) {
    Node a, b;
    if (has_a && has_b) {
        a = a_in;
        b = b_in;
    } else if (has_a && f <= 0.0) {
        a = b = a_in;
    } else if (has_b && f >= 1.0) {
        a = b = b_in;
    } else {
        return QM_FAILED_NO_TERRAIN;
    }
    const Node& nearer = f < 0.5 ? a : b;
    if (nearer.forest > 100u || !node_valid(a) || !node_valid(b)) return QM_FAILED_NO_TERRAIN;
    const double ha = height_m_of_code(a.height_code), hb = height_m_of_code(b.height_code);
    out.height_m = ha + f * (hb - ha);
    const double ia = (double)a.impervious, ib = (double)b.impervious;
    out.ground_factor = 1.0 - (ia + f * (ib - ia)) / 100.0;
    out.forest_cover = (double)nearer.forest / 100.0;
    out.canopy_m = (double)nearer.canopy;
    return 0;
}

static void push_sample(Profile& profile, double t, const GroundSample& s) {
    profile.push(t, s.height_m, s.ground_factor, s.forest_cover, s.canopy_m);
}

// popup/src/scene.rs:333-420 Ground::walk on the planar lattice `scale` times as fine as the
// coarse one: a vertex wherever the ray a -> b crosses a row or column line strictly between
// fractions `from` and `to`, read between that line's two nodes (the fine sample on the fine
// lattice); at a node (both lines at once) the row line's reading, both advanced.
static int walk(const qm_terrain& T, int scale, const double a[2], const double b[2], double from, double to,
                Profile& profile) {
    const double unit = T.spacing_m / scale;
    const double cunit = column_metres(T) / scale;
    const double ca = (a[0] - T.origin_x_m) / cunit, cb = (b[0] - T.origin_x_m) / cunit;
    // With a row table the row lines lie at their north metres (rise in metres, ra 0) and the row
    // index of the start comes from the table (scene.rs walk: rows_north_of).
    const bool geo = T.row_y_m != nullptr;
    const double ra = geo ? 0.0 : (a[1] - T.origin_y_m) / unit, rb = geo ? b[1] - a[1] : (b[1] - T.origin_y_m) / unit;
    const double columns = cb - ca, rise = rb - ra;
    const double column_from = ca + from * columns;
    const double row_from = geo ? scale * row_position(T, a[1] + from * rise) : ra + from * rise;
    double next_column, column_step;
    if (columns > 0.0) {
        next_column = std::floor(column_from) + 1.0;
        column_step = 1.0;
    } else {
        next_column = std::ceil(column_from) - 1.0;
        column_step = -1.0;
    }
    // `south`: the south row of the cell the ray is in; a start on a row line belongs to the
    // cell the ray heads into.
    double next_row, south;
    if (rise > 0.0) {
        next_row = std::floor(row_from) + 1.0;
        south = std::floor(row_from);
    } else if (rise < 0.0) {
        next_row = std::ceil(row_from) - 1.0;
        south = std::ceil(row_from) - 1.0;
    } else {
        next_row = 0.0;
        south = std::floor(row_from);
    }
    const double inf = std::numeric_limits<double>::infinity();
    for (;;) {
        const double t_column = columns != 0.0 ? (next_column - ca) / columns : inf;
        const double t_row = rise == 0.0 ? inf
                             : geo      ? (row_line_y(T, scale, (int64_t)next_row) - a[1]) / rise
                                        : (next_row - ra) / rise;
        // A column vertex's fraction between the row lines about it.
        auto row_fraction = [&](double t) {
            if (!geo) return std::min(std::max(ra + t * rise - south, 0.0), 1.0);
            const double y = a[1] + t * rise, y0 = row_line_y(T, scale, (int64_t)south);
            const double y1 = row_line_y(T, scale, (int64_t)south + 1);
            return std::min(std::max((y - y0) / (y1 - y0), 0.0), 1.0);
        };
        const double t = std::min(t_column, t_row);
        if (t >= to) break;
        GroundSample sample;
        int failed;
        if (scale == 1) {
            Node n0{}, n1{};
            bool h0, h1;
            double f;
            if (t_row <= t_column) {
                const double c = ca + t * columns;
                const double west = std::floor(c);
                f = c - west;
                h0 = node_at(T, (int64_t)next_row, (int64_t)west, n0);
                h1 = node_at(T, (int64_t)next_row, (int64_t)west + 1, n1);
            } else {
                f = row_fraction(t);
                h0 = node_at(T, (int64_t)south, (int64_t)next_column, n0);
                h1 = node_at(T, (int64_t)south + 1, (int64_t)next_column, n1);
            }
            failed = edge_sample(h0, n0, h1, n1, f, sample);
        } else if (t_row <= t_column) {
            const double c = ca + t * columns;
            failed = fine_sample(T, true, (int64_t)next_row, 0, c, c - std::floor(c), sample);
        } else {
            failed = fine_sample(T, false, (int64_t)next_column, (int64_t)south, 0.0, row_fraction(t), sample);
        }
        if (failed) return failed;
        push_sample(profile, t, sample);
        if (t_row <= t_column) {
            if (rise > 0.0) {
                south = next_row;
                next_row += 1.0;
            } else {
                south = next_row - 1.0;
                next_row -= 1.0;
            }
        }
        if (t_column <= t_row) next_column += column_step;
    }
    return 0;
}

// popup/src/scene.rs Ground::fine_near: whether a block under the stretch a -> b (its box's
// corners) has fine terrain.
static bool fine_near(const qm_terrain& T, const double a[2], const double b[2]) {
    if (!T.fine_detail) return false;
    const double xs[2] = {std::min(a[0], b[0]), std::max(a[0], b[0])};
    const double ys[2] = {std::min(a[1], b[1]), std::max(a[1], b[1])};
    for (double x : xs)
        for (double y : ys)
            if (fine_block(T, column_position(T, x), row_position(T, y))) return true;
    return false;
}

// popup/src/scene.rs:242-309 Ground::fill_profile: the 1/6" lattice within QM_FINE_ENDS_M of either
// end where a block under that stretch has fine terrain, the coarse lattice elsewhere, neighbours
// on one lattice joined, a vertex (Ground::at) where the lattice changes.
int fill_profile(const qm_terrain& T, const double source[2], const double receiver[2], Profile& profile) {
    const double offset[2] = {source[0] - receiver[0], source[1] - receiver[1]};
    const double length = std::hypot(offset[0], offset[1]);
    profile.clear(std::max(length, 1.0));
    GroundSample start, end;
    int failed = ground_at(T, source[0], source[1], start);
    if (failed) return failed;
    push_sample(profile, 0.0, start);
    auto along = [&](double t, double out[2]) {
        out[0] = source[0] + t * (receiver[0] - source[0]);
        out[1] = source[1] + t * (receiver[1] - source[1]);
    };
    struct Stretch {
        double from, to;
        int scale;
    };
    Stretch stretches[3] = {{0.0, 1.0, 1}, {0.0, 1.0, 1}, {0.0, 1.0, 1}};
    int count = 1;
    if (!T.fine_detail) {
        count = 1;
    } else if (length <= 2.0 * QM_FINE_ENDS_M) {
        stretches[0].scale = fine_near(T, source, receiver) ? QM_FINE_PER_NODE : 1;
    } else {
        const double e = QM_FINE_ENDS_M / length;
        double pe[2], pf[2];
        along(e, pe);
        along(1.0 - e, pf);
        const Stretch candidates[3] = {{0.0, e, fine_near(T, source, pe) ? QM_FINE_PER_NODE : 1},
                                       {e, 1.0 - e, 1},
                                       {1.0 - e, 1.0, fine_near(T, pf, receiver) ? QM_FINE_PER_NODE : 1}};
        count = 0;
        for (const Stretch& c : candidates) {
            if (count > 0 && stretches[count - 1].scale == c.scale) stretches[count - 1].to = c.to;
            else stretches[count++] = c;
        }
    }
    for (int k = 0; k < count; ++k) {
        failed = walk(T, stretches[k].scale, source, receiver, stretches[k].from, stretches[k].to, profile);
        if (failed) return failed;
        if (stretches[k].to < 1.0) {
            const double to = stretches[k].to;
            GroundSample sw;
            failed = ground_at(T, source[0] - to * offset[0], source[1] - to * offset[1], sw);
            if (failed) return failed;
            push_sample(profile, to, sw);
        }
    }
    failed = ground_at(T, receiver[0], receiver[1], end);
    if (failed) return failed;
    push_sample(profile, 1.0, end);
    return 0;
}

// ---------------------------------------------------------------------------------------------
// popup/src/obstacles/mod.rs:61-76 Lattice::steps (snapped to a lattice point within 1e-6)
static double snap(double steps) {
    const double nearest = std::round(steps);
    return std::fabs(steps - nearest) < 1e-6 ? nearest : steps;
}
void steps_of(const qm_obstacles& O, double x, double y, double out[2]) {
    out[0] = snap(O.origin_steps_x + x / O.metres_per_step);
    out[1] = snap(O.origin_steps_y - y / (O.metres_per_step_y > 0.0 ? O.metres_per_step_y : O.metres_per_step));
}

// crossings.rs:13-28
static bool segment_intersection_t(const double start[2], const double delta[2], const double a[2], const double b[2],
                                   double& t_out) {
    const double ex = b[0] - a[0], ey = b[1] - a[1];
    const double denominator = delta[0] * ey - delta[1] * ex;
    if (denominator == 0.0) return false;
    const double wx = a[0] - start[0], wy = a[1] - start[1];
    const double t = (wx * ey - wy * ex) / denominator;
    const double u = (wx * delta[1] - wy * delta[0]) / denominator;
    if (t > 0.0 && t < 1.0 && u >= 0.0 && u <= 1.0) {
        t_out = t;
        return true;
    }
    return false;
}

static void vertex_steps(const qm_obstacles& O, size_t index, double out[2]) {
    out[0] = (double)O.vertices[2 * index];
    out[1] = (double)O.vertices[2 * index + 1];
}

// crossings.rs:80-111 SceneTile::edge_bounds_roof
static bool edge_bounds_roof(const qm_obstacles& O, const double start[2], const double delta[2], size_t edge,
                             size_t first, size_t count) {
    auto side = [&](size_t index) {
        double v[2];
        vertex_steps(O, index, v);
        return delta[0] * (v[1] - start[1]) - delta[1] * (v[0] - start[0]);
    };
    const double a = side(edge), b = side(edge + 1);
    if (a != 0.0 && b != 0.0) return true;
    const size_t distinct = count - 1;
    const size_t vertex = ((a == 0.0 ? edge : edge + 1) - first) % distinct;
    double n0 = 0.0, n1 = 0.0;
    for (size_t offset = 1; offset < distinct; ++offset) {
        if (n0 == 0.0) n0 = side(first + (vertex + distinct - offset) % distinct);
        if (n1 == 0.0) n1 = side(first + (vertex + offset) % distinct);
        if (n0 != 0.0 && n1 != 0.0) break;
    }
    return n0 * n1 < 0.0;
}

// crossings.rs:34-56 sort_and_deduplicate; ties of (t, footprint) ordered by outline and edge so
// that the kept crossing is one fixed choice.
void sort_and_deduplicate(std::vector<RawCrossing>& list) {
    std::sort(list.begin(), list.end(), [](const RawCrossing& a, const RawCrossing& b) {
        if (a.c.t != b.c.t) return a.c.t < b.c.t;
        if (a.c.footprint_id != b.c.footprint_id) return a.c.footprint_id < b.c.footprint_id;
        if (a.outline != b.outline) return a.outline < b.outline;
        return a.edge < b.edge;
    });
    size_t kept = 0;
    for (size_t i = 0; i < list.size(); ++i) {
        const RawCrossing candidate = list[i];
        bool merged = false;
        for (size_t k = kept; k-- > 0;) {
            if (!(candidate.c.t - list[k].c.t < 1e-9)) break;
            if (list[k].c.footprint_id == candidate.c.footprint_id) {
                list[k].c.bounds_roof = list[k].c.bounds_roof || candidate.c.bounds_roof;
                merged = true;
                break;
            }
        }
        if (!merged) list[kept++] = candidate;
    }
    list.resize(kept);
}

static RawCrossing make_crossing(const qm_obstacles& O, const double start[2], const double delta[2], uint32_t outline,
                                 size_t edge_vertex, double t) {
    const qm_outline& record = O.outlines[outline];
    RawCrossing r;
    r.c.t = t;
    r.c.height_m = (double)record.height_dm / 10.0;
    r.c.building = record.kind == QM_KIND_BUILDING;
    r.c.bounds_roof = r.c.building &&
                      edge_bounds_roof(O, start, delta, edge_vertex, record.first_vertex, record.vertex_count);
    r.c.footprint_id = record.footprint_id;
    r.outline = outline;
    r.edge = (uint32_t)edge_vertex;
    return r;
}

// crossings.rs:163-230 Scene::crossings: an Amanatides-Woo walk over the cells, the padded piece
// box filter (crossings.rs:60-75) and the exact test on each listed edge.
int crossings_raw(const qm_obstacles& O, const double from[2], const double to[2], std::vector<RawCrossing>& out,
                  WorkCount* work) {
    out.clear();
    double start[2], end[2];
    steps_of(O, from[0], from[1], start);
    steps_of(O, to[0], to[1], end);
    const double delta[2] = {end[0] - start[0], end[1] - start[1]};
    const double cs = (double)QM_CELL_STEPS;
    int64_t cell[2] = {(int64_t)std::floor(start[0] / cs), (int64_t)std::floor(start[1] / cs)};
    const int64_t last[2] = {(int64_t)std::floor(end[0] / cs), (int64_t)std::floor(end[1] / cs)};
    const int64_t step[2] = {delta[0] >= 0.0 ? 1 : -1, delta[1] >= 0.0 ? 1 : -1};
    const double inf = std::numeric_limits<double>::infinity();
    double t_delta[2], t_max[2];
    for (int axis = 0; axis < 2; ++axis) {
        t_delta[axis] = delta[axis] != 0.0 ? std::fabs(cs / delta[axis]) : inf;
        if (delta[axis] == 0.0) {
            t_max[axis] = inf;
        } else {
            const double boundary = (double)(cell[axis] + (delta[axis] >= 0.0 ? 1 : 0)) * cs;
            t_max[axis] = std::fabs((boundary - start[axis]) / delta[axis]);
        }
    }
    const double pad = 1e-9 * (1.0 + std::fabs(delta[0]) + std::fabs(delta[1]));
    double t_enter = 0.0;
    for (;;) {
        const int64_t gx = cell[0] - O.cell_x0, gy = cell[1] - O.cell_y0;
        if (gx < 0 || gy < 0 || gx >= (int64_t)O.cells_x || gy >= (int64_t)O.cells_y) return QM_FAILED_NOT_READ;
        const size_t index = (size_t)gy * O.cells_x + (size_t)gx;
        const uint32_t r0 = O.cell_first_run[index], r1 = O.cell_first_run[index + 1];
        if (work) work->cells += 1;
        if (r1 > r0) {
            if (work) work->cells_occupied += 1;
            const double t_exit = std::min(std::min(t_max[0], t_max[1]), 1.0);
            const double te = std::min(std::max(t_enter, 0.0), 1.0), tx = std::min(std::max(t_exit, 0.0), 1.0);
            const double p[2] = {start[0] + delta[0] * te, start[1] + delta[1] * te};
            const double q[2] = {start[0] + delta[0] * tx, start[1] + delta[1] * tx};
            int32_t low[2], high[2];
            for (int axis = 0; axis < 2; ++axis) {
                low[axis] = (int32_t)std::floor(std::min(p[axis], q[axis]) - pad);
                high[axis] = (int32_t)std::ceil(std::max(p[axis], q[axis]) + pad);
            }
            for (uint32_t ri = r0; ri < r1; ++ri) {
                const qm_run& run = O.runs[ri];
                const size_t first = (size_t)O.outlines[run.outline].first_vertex + run.first_edge;
                int32_t ax = O.vertices[2 * first], ay = O.vertices[2 * first + 1];
                for (size_t next = first + 1; next <= first + run.edge_count; ++next) {
                    const int32_t bx = O.vertices[2 * next], by = O.vertices[2 * next + 1];
                    if (work) work->edges += 1;
                    const bool meets = high[0] >= std::min(ax, bx) && std::max(ax, bx) >= low[0] &&
                                       high[1] >= std::min(ay, by) && std::max(ay, by) >= low[1];
                    double t;
                    const double av[2] = {(double)ax, (double)ay}, bv[2] = {(double)bx, (double)by};
                    if (meets && segment_intersection_t(start, delta, av, bv, t))
                        out.push_back(make_crossing(O, start, delta, run.outline, next - 1, t));
                    ax = bx;
                    ay = by;
                }
            }
        }
        if (cell[0] == last[0] && cell[1] == last[1]) break;
        t_enter = std::min(t_max[0], t_max[1]);
        int axis;
        if (cell[0] == last[0]) axis = 1;
        else if (cell[1] == last[1] || t_max[0] < t_max[1]) axis = 0;
        else axis = 1;
        t_max[axis] += t_delta[axis];
        cell[axis] += step[axis];
    }
    sort_and_deduplicate(out);
    if (work) work->crossings += out.size();
    return 0;
}

// Brute force: every edge of every outline, no cells and no box filter; the crossing set a correct
// cell walk must reproduce exactly (test support for the cell walk).
void crossings_brute_force(const qm_obstacles& O, const double from[2], const double to[2],
                           std::vector<RawCrossing>& out) {
    out.clear();
    double start[2], end[2];
    steps_of(O, from[0], from[1], start);
    steps_of(O, to[0], to[1], end);
    const double delta[2] = {end[0] - start[0], end[1] - start[1]};
    for (uint32_t o = 0; o < O.outlines_count; ++o) {
        const qm_outline& record = O.outlines[o];
        for (size_t v = record.first_vertex; v + 1 < (size_t)record.first_vertex + record.vertex_count; ++v) {
            double a[2], b[2], t;
            vertex_steps(O, v, a);
            vertex_steps(O, v + 1, b);
            if (segment_intersection_t(start, delta, a, b, t)) out.push_back(make_crossing(O, start, delta, o, v, t));
        }
    }
    sort_and_deduplicate(out);
}

// ---------------------------------------------------------------------------------------------
// physics/src/weather.rs:175-214 corners + place_of: on the world table by the place's latitude
// and longitude (the geographic frame), or on a planar weather lattice (clamped at its edges).
void place_weather(const qm_weather& W, double x, double y, double favourable[PERIODS][SECTORS], double alpha[BANDS]) {
    uint32_t row, column, r1, c1;
    double fy, fx;
    if (W.nodes_per_degree) {
        // geo::Mercator::to_degrees of the frame point (z12 units, 4096 a side), then the world
        // table's corners (weather.rs:186-200), longitude wrapping.
        const double n = 4096.0;
        const double mx = W.frame_x + x / W.east_m_per_unit, my = W.frame_y - y / W.north_m_per_unit;
        const double lat = std::atan(std::sinh(M_PI * (1.0 - 2.0 * my / n))) * (180.0 / M_PI);
        const double lon = mx / n * 360.0 - 180.0;
        const double nodes = (double)W.nodes_per_degree;
        const double yy = std::min(std::max((90.0 - lat) * nodes, 0.0), (double)(W.rows - 1));
        double wrapped = std::fmod(lon, 360.0);
        if (wrapped < 0.0) wrapped += 360.0;
        const double xx = wrapped * nodes;
        row = std::min((uint32_t)std::floor(yy), W.rows - 2);
        const uint32_t at = (uint32_t)std::floor(xx);
        fy = yy - row;
        fx = xx - at;
        column = at % W.columns;
        c1 = (at + 1) % W.columns;
        r1 = row + 1;
    } else {
        const double yy = std::min(std::max((y - W.origin_y_m) / W.spacing_m, 0.0), (double)(W.rows - 1));
        const double xx = std::min(std::max((x - W.origin_x_m) / W.spacing_m, 0.0), (double)(W.columns - 1));
        row = std::min((uint32_t)std::floor(yy), W.rows >= 2 ? W.rows - 2 : 0u);
        column = std::min((uint32_t)std::floor(xx), W.columns >= 2 ? W.columns - 2 : 0u);
        fy = yy - row;
        fx = xx - column;
        r1 = W.rows >= 2 ? row + 1 : row;
        c1 = W.columns >= 2 ? column + 1 : column;
    }
    const uint8_t* n[4] = {W.nodes + ((size_t)row * W.columns + column) * QM_WEATHER_NODE_BYTES,
                           W.nodes + ((size_t)row * W.columns + c1) * QM_WEATHER_NODE_BYTES,
                           W.nodes + ((size_t)r1 * W.columns + column) * QM_WEATHER_NODE_BYTES,
                           W.nodes + ((size_t)r1 * W.columns + c1) * QM_WEATHER_NODE_BYTES};
    const double w[4] = {(1.0 - fy) * (1.0 - fx), (1.0 - fy) * fx, fy * (1.0 - fx), fy * fx};
    for (int p = 0; p < PERIODS; ++p)
        for (int s = 0; s < SECTORS; ++s) {
            double total = 0.0;
            for (int k = 0; k < 4; ++k) total += w[k] * (double)n[k][p * SECTORS + s];
            favourable[p][s] = total / 100.0;
        }
    for (int b = 0; b < BANDS; ++b) {
        double total = 0.0;
        const int at = PERIODS * SECTORS + 2 * b;
        for (int k = 0; k < 4; ++k) total += w[k] * (double)((uint32_t)n[k][at] | ((uint32_t)n[k][at + 1] << 8));
        alpha[b] = total / 100.0;
    }
}

// weather.rs:61-70 FavourableProbability::at
double favourable_at(const double favourable[PERIODS][SECTORS], int period, double azimuth_rad) {
    double bearing = std::fmod(90.0 - azimuth_rad * (180.0 / M_PI), 360.0);
    if (bearing < 0.0) bearing += 360.0;
    const double position = bearing / (360.0 / SECTORS);
    const size_t lower = (size_t)std::floor(position) % SECTORS;
    const size_t upper = (lower + 1) % SECTORS;
    const double fraction = position - std::floor(position);
    const double* row = favourable[period];
    return row[lower] + fraction * (row[upper] - row[lower]);
}

// paint/src/exact.rs:95-118 Point::with_receiver_properties: ground + 4 m and the weather there.
int receiver_of(const Scene& S, const qm_point& point, Receiver& out) {
    GroundSample g;
    const int failed = ground_at(S.terrain, point.x_m, point.y_m, g);
    out.x = point.x_m;
    out.y = point.y_m;
    out.own_footprint = point.own_footprint;
    out.altitude_m = failed ? 0.0 : g.height_m + QM_RECEIVER_HEIGHT_M;
    place_weather(S.weather, point.x_m, point.y_m, out.favourable, out.alpha);
    return failed;
}

// popup/src/evaluate.rs:171-221 source_rays (a point: one ray), 300-330 ray, 119-152 add_ray;
// paint/src/exact.rs:186-205 whole_energy.
int evaluate_pair(const Scene& S, const Receiver& receiver, const qm_source& source, PairScratch& s, PairResult& out,
                  WorkCount* work) {
    for (int p = 0; p < PERIODS; ++p) {
        out.mixed[p] = 0.0;
        out.states[0][p] = out.states[1][p] = 0.0;
        for (int b = 0; b < BANDS; ++b) out.bands[p][b] = 0.0;
    }
    const double src[2] = {source.x_m, source.y_m};
    const double rcv[2] = {receiver.x, receiver.y};
    int failed = fill_profile(S.terrain, src, rcv, s.profile);
    if (failed) return failed;
    if (work) work->vertices += s.profile.t.size();
    failed = crossings_raw(S.obstacles, src, rcv, s.raw, work);
    if (failed) return failed;
    s.crossings.clear();
    for (const RawCrossing& r : s.raw)  // evaluate.rs:293-297 without_own_walls
        if (receiver.own_footprint == 0 || r.c.footprint_id != receiver.own_footprint) s.crossings.push_back(r.c);
    const double azimuth = std::atan2(receiver.y - source.y_m, receiver.x - source.x_m);
    double p[PERIODS];
    for (int period = 0; period < PERIODS; ++period) p[period] = favourable_at(receiver.favourable, period, azimuth);
    RayEnds ends;
    ends.source_height_m = source.height_m;
    ends.has_source_ground = source.on_deck != 0;  // evaluate.rs:193-196 a deck's ground (deck_at(0))
    ends.source_ground_m = source.on_deck ? (double)source.deck_m : 0.0;
    ends.receiver_altitude_m = receiver.altitude_m;
    ends.has_source_ground_factor = source.ground_percent != QM_GROUND_FROM_TERRAIN;
    ends.source_ground_factor = (double)source.ground_percent / 100.0;
    ends.platform_half_width_m = source.platform_half_width_m;
    ends.own_footprint = source.footprint_id;
    ray_terms(s.profile, s.crossings, ends, p, receiver.alpha, s.ray, s.terms, work);
    // evaluate.rs:195-203: the point divergence at the candidate's distance.
    const double horizontal = std::hypot(source.x_m - receiver.x, source.y_m - receiver.y);
    // candidate.traced_ground_m (candidates.rs:151-158): the deck, else the terrain at the source.
    const double ground0 = source.on_deck ? (double)source.deck_m : s.profile.ground_m[0];
    const double d = std::max(std::hypot(std::max(horizontal, (double)source.exclusion_radius_m),
                                         receiver.altitude_m - (ground0 + (double)source.height_m)),
                              1.0);
    const double weight = 1.0 / (d * d * energy(11.0));
    const Transfer& tr = s.terms.transfer;
    for (int st = 0; st < 2; ++st)
        for (int b = 0; b < BANDS; ++b) out.boundary[st][b] = s.terms.boundaries[st].attenuation_db[b];
    for (int period = 0; period < PERIODS; ++period)
        for (int b = 0; b < BANDS; ++b) {
            const double power = weight * (double)source.energy[period][b];
            out.bands[period][b] += power * tr.periods[period][b];
            for (int st = 0; st < 2; ++st) out.states[st][period] += power * tr.states[st][b];
        }
    for (int period = 0; period < PERIODS; ++period) {
        double sum = 0.0;
        for (int b = 0; b < BANDS; ++b) sum += out.bands[period][b];
        out.mixed[period] = sum;
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
void parallel_for(size_t n, int threads, const std::function<void(size_t, int)>& body) {
    if (threads <= 1 || n < 2) {
        for (size_t i = 0; i < n; ++i) body(i, 0);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int k = 0; k < threads; ++k)
        pool.emplace_back([&, k] {
            for (;;) {
                const size_t i = next.fetch_add(1);
                if (i >= n) break;
                body(i, k);
            }
        });
    for (std::thread& t : pool) t.join();
}

}  // namespace qmo
