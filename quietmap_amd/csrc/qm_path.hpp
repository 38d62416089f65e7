// One receiver-source path on the device: obstacle crossings (exact FP64), the ground under the
// ray (lattice walk), the vertical path with roofs, both meteorological states' rubber bands and
// CNOSSOS-EU boundary terms, foliage, air absorption and the period mix.
//
// Reference behaviour restated (all file:line into the reference):
//   crossings        popup/src/obstacles/crossings.rs:13-230
//   ground walk      popup/src/scene.rs:242-420, 532-561; tiles/src/terrain.rs:342-385
//   vertical path    physics/src/ray.rs:125-191 (ray_terms), 201-226 (containing footprints),
//                    255-354 (PathBuffers::fill), 360-385 (crests_of), 389-404 (roof clipping)
//   boundary terms   physics/src/cnossos/mod.rs:123-184, mean_plane.rs:19-120, ground.rs:39-137,
//                    rubber_band.rs:48-264, diffraction.rs:19-213
//   foliage          physics/src/foliage.rs:22-64
//   pair energy      popup/src/evaluate.rs:119-221, 300-330; paint/src/exact.rs:186-205
//
// Design (this file's own): one pass over the lattice walk per path. The wall tops and roof
// grounds are read before the walk by bracketing each crossing between its two lattice-line
// vertices directly, so the walk streams the merged profile (terrain plus roofs) with no lag. Prefix
// moments of that profile (integral of z, x*z and G from the source) are kept in FP64 at every
// vertex; each diffraction candidate carries its prefix, so a side plane over any range is a
// difference of two prefixes without FP32 cancellation. Candidates (thinned terrain crests merged
// with wall tops) feed both states' rubber bands as monotone chains and the unblocked-best search
// in the same pass. Path differences (delta) are formed in FP64; band terms in FP32.
#pragma once

#include "qm_device.hpp"

namespace qm {

// Capacities of the per-path buffers; a path beyond any of them reports QM_FAILED_CAPACITY.
constexpr int CROSS_CAP = 48;
constexpr int ROOF_CAP = 24;
constexpr int HULL_CAP = 24;
constexpr int LOG_CAP = 64;  // candidate log entries; compacted when full (live entries <= 2*HULL_CAP + 4)

struct Scene {
    qm_terrain terrain;
    qm_obstacles obstacles;
    qm_weather weather;
    // Built at upload from the obstacles' cell index: one 64-bit word per block of 8 x 8 cells
    // (block row-major, blocks_x a row; bit 8 * (y & 7) + (x & 7) set where the cell holds runs),
    // so the cell walk tests an empty cell from the word it holds instead of reading the index.
    const uint64_t* cell_blocks;
    uint32_t blocks_x, blocks_y;
    // Built at upload: each run's first vertex (its outline's first vertex plus the run's first
    // edge), so a cell's edges load right after its runs, the outline record only at a hit.
    const uint32_t* run_vertex;
};

struct GroundSample {
    float z, g, forest, canopy;
};

// ---------------------------------------------------------------------------------------------
// Terrain: tiles/src/terrain.rs:36-38 height code, popup/src/scene.rs:532-561 edge_sample,
// tiles/src/terrain.rs:342-385 sample_at (fractions north and east of the south-west node).
QM_DEV bool node_at(const qm_terrain& T, int row, int column, uint32_t& node, uint32_t& index) {
    if (row < 0 || column < 0 || row >= (int)T.rows || column >= (int)T.columns) return false;
    index = (uint32_t)row * T.columns + (uint32_t)column;
    node = T.nodes[index];
    return true;
}
// The canopy byte is read only under forest (the only reader is the foliage term).
QM_DEV float canopy_of(const qm_terrain& T, uint32_t node, uint32_t index) {
    return (node >> 24) ? (float)T.canopy_m[index] : 0.0f;
}
QM_DEV bool node_valid(uint32_t n) { return (n & 0xffffu) != 0xffffu && ((n >> 16) & 0xffu) <= 100u; }
QM_DEV float height_of(uint32_t n) { return -500.0f + (float)(n & 0xffffu) * 0.2f; }
QM_DEV float impervious_of(uint32_t n) { return (float)((n >> 16) & 0xffu); }
QM_DEV float forest_of(uint32_t n) { return (float)(n >> 24); }

QM_DEV uint32_t edge_sample(const qm_terrain& T, int r0, int c0, int r1, int c1, float f, GroundSample& out) {
    uint32_t a, b, ca, cb;
    const bool ha = node_at(T, r0, c0, a, ca), hb = node_at(T, r1, c1, b, cb);
    if (!(ha && hb)) {
        if (ha && f <= 0.0f) {
            b = a;
            cb = ca;
        } else if (hb && f >= 1.0f) {
            a = b;
            ca = cb;
        } else {
            return QM_FAILED_NO_TERRAIN;
        }
    }
    const uint32_t nearer = f < 0.5f ? a : b;
    const uint32_t nearer_index = f < 0.5f ? ca : cb;
    if ((nearer >> 24) > 100u || !node_valid(a) || !node_valid(b)) return QM_FAILED_NO_TERRAIN;
    const float za = height_of(a), zb = height_of(b);
    const float ia = impervious_of(a), ib = impervious_of(b);
    out.z = za + f * (zb - za);
    out.g = 1.0f - (ia + f * (ib - ia)) * 0.01f;
    out.forest = forest_of(nearer) * 0.01f;
    out.canopy = canopy_of(T, nearer, nearer_index);
    return 0;
}

// Bilinear ground at a point given in lattice units (column, row).
QM_DEV uint32_t ground_at(const qm_terrain& T, double cp, double rp, GroundSample& out) {
    const double column = floor(cp), row = floor(rp);
    const float cf = (float)(cp - column), rf = (float)(rp - row);
    const int c = (int)column, r = (int)row;
    const int nc = c + (cf > 0.0f ? 1 : 0), nr = r + (rf > 0.0f ? 1 : 0);
    uint32_t n[4], k[4];
    if (!node_at(T, r, c, n[0], k[0]) || !node_at(T, r, nc, n[1], k[1]) || !node_at(T, nr, c, n[2], k[2]) ||
        !node_at(T, nr, nc, n[3], k[3]))
        return QM_FAILED_NO_TERRAIN;
    const int nearest = (rf >= 0.5f ? 2 : 0) + (cf >= 0.5f ? 1 : 0);
    if ((n[nearest] >> 24) > 100u) return QM_FAILED_NO_TERRAIN;
    for (int i = 0; i < 4; ++i)
        if (!node_valid(n[i])) return QM_FAILED_NO_TERRAIN;
    const float w[4] = {(1.0f - rf) * (1.0f - cf), (1.0f - rf) * cf, rf * (1.0f - cf), rf * cf};
    float h = 0.0f, imp = 0.0f;
    for (int i = 0; i < 4; ++i) {
        h += w[i] * height_of(n[i]);
        imp += w[i] * impervious_of(n[i]);
    }
    out.z = h;
    out.g = 1.0f - imp * 0.01f;
    out.forest = forest_of(n[nearest]) * 0.01f;
    out.canopy = canopy_of(T, n[nearest], k[nearest]);
    return 0;
}

// The geographic frame (qm_terrain.row_y_m, column_m): the one arc-second rows lie at the table's
// north metres (uneven: latitude through Mercator, popup/src/scene.rs:303-330 row_line_y and
// rows_north_of), the columns evenly at column_m. GEO: the caller may meet such a table (the
// general instantiations); without one, or with GEO false, row r lies at origin_y + r * spacing.
QM_DEV double column_metres(const qm_terrain& T) { return T.column_m > 0.0 ? T.column_m : T.spacing_m; }

// North metres of row k of the table, the end spacings extended beyond it.
QM_DEV double table_row_y(const qm_terrain& T, int k) {
    const int last = (int)T.rows - 1;
    if (k < 0) return T.row_y_m[0] + (double)k * (T.row_y_m[1] - T.row_y_m[0]);
    if (k > last) return T.row_y_m[last] + (double)(k - last) * (T.row_y_m[last] - T.row_y_m[last - 1]);
    return T.row_y_m[k];
}

// The fractional column and row (node indices) of frame metres x and y.
QM_DEV double column_position(const qm_terrain& T, double x) { return (x - T.origin_x_m) * (1.0 / column_metres(T)); }
template <bool GEO = false>
QM_DEV double row_position(const qm_terrain& T, double y) {
    if (!GEO || !T.row_y_m) return (y - T.origin_y_m) * (1.0 / T.spacing_m);
    const int last = (int)T.rows - 1;
    int lo = 0;
    if (y < T.row_y_m[0]) {
        lo = 0;
    } else if (y >= T.row_y_m[last]) {
        lo = last - 1;
    } else {
        int hi = last;  // row_y_m[lo] <= y < row_y_m[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (T.row_y_m[mid] <= y) lo = mid;
            else hi = mid;
        }
    }
    return (double)lo + (y - T.row_y_m[lo]) / (T.row_y_m[lo + 1] - T.row_y_m[lo]);
}

// The ray's lattice geometry: position of the source (ca, ra) in lattice units, the offsets to
// the receiver, and the first lines strictly ahead of the source.
// `scale`: the lattice scale times as fine as the coarse one (1, or QM_FINE_PER_NODE on a fine
// stretch); the first lines are those strictly after fraction `from` of the ray.
// With a row table (`geo`), the row units are north metres from the source (ra 0, rise the ray's
// north extent) and row line k of the lattice lies on the table, a fine line evenly between the
// table's rows.
struct Lattice {
    float ca, ra, columns, rise, inv_columns, inv_rise;
    int first_column, column_step, first_row, row_step, start_south;
    int scale;
    float span;  // lines the ray crosses, both families (the walk's step budget)
    bool geo;
    const qm_terrain* terrain;
    double y0, dy;
};

// North metres of row line k of a geo lattice.
QM_DEV double row_line_y(const Lattice& L, int k) {
    const int s = L.scale;
    const int coarse = k >= 0 ? k / s : -((-k + s - 1) / s);
    const int sub = k - coarse * s;
    const double y0 = table_row_y(*L.terrain, coarse);
    if (sub == 0) return y0;
    return y0 + (table_row_y(*L.terrain, coarse + 1) - y0) * (double)sub / (double)s;
}

QM_DEV Lattice lattice_of(const qm_terrain& T, double sx, double sy, double rx, double ry, int scale = 1,
                          float from = 0.0f, bool geo_ok = false) {
    Lattice L;
    L.scale = scale;
    L.geo = geo_ok && T.row_y_m != nullptr;
    L.terrain = &T;
    const double inv = (double)scale / column_metres(T);
    const double ca = (sx - T.origin_x_m) * inv, cb = (rx - T.origin_x_m) * inv;
    L.ca = (float)ca;
    L.columns = (float)(cb - ca);
    L.inv_columns = L.columns != 0.0f ? 1.0f / L.columns : 0.0f;
    const float c0 = L.ca + from * L.columns;
    if (L.columns > 0.0f) {
        L.first_column = (int)floorf(c0) + 1;
        L.column_step = 1;
    } else {
        L.first_column = (int)ceilf(c0) - 1;
        L.column_step = -1;
    }
    float r0;
    if (L.geo) {
        L.y0 = sy;
        L.dy = ry - sy;
        L.ra = 0.0f;
        L.rise = (float)L.dy;
        L.inv_rise = L.rise != 0.0f ? 1.0f / L.rise : 0.0f;
        const double pa = (double)scale * row_position<true>(T, sy), pb = (double)scale * row_position<true>(T, ry);
        r0 = (float)((double)scale * row_position<true>(T, sy + (double)from * L.dy));
        L.span = fabsf(L.columns) + (float)fabs(pb - pa);
    } else {
        const double rinv = (double)scale / T.spacing_m;
        const double ra = (sy - T.origin_y_m) * rinv, rb = (ry - T.origin_y_m) * rinv;
        L.ra = (float)ra;
        L.rise = (float)(rb - ra);
        L.inv_rise = L.rise != 0.0f ? 1.0f / L.rise : 0.0f;
        r0 = L.ra + from * L.rise;
        L.y0 = L.dy = 0.0;
        L.span = fabsf(L.columns) + fabsf(L.rise);
    }
    if (L.rise > 0.0f) {
        L.first_row = (int)floorf(r0) + 1;
        L.row_step = 1;
        L.start_south = (int)floorf(r0);
    } else if (L.rise < 0.0f) {
        L.first_row = (int)ceilf(r0) - 1;
        L.row_step = -1;
        L.start_south = (int)ceilf(r0) - 1;
    } else {
        L.first_row = 0;
        L.row_step = 0;
        L.start_south = (int)floorf(r0);
    }
    return L;
}

// The fractional row line (this lattice's units) of the ray at fraction t.
QM_DEV float row_units_at(const Lattice& L, float t) {
    if (L.geo) return (float)((double)L.scale * row_position<true>(*L.terrain, L.y0 + (double)t * L.dy));
    return L.ra + t * L.rise;
}

// ---------------------------------------------------------------------------------------------
// Fine terrain (tiles/src/terrain_fine.rs; paint-gpu/kernels/terrain.cuh:91-103, 155-175 and
// terrain_walk.cuh:59-160): detail in QM_DETAIL_STEP_M steps on the lattice QM_FINE_PER_NODE times
// as fine, read in the blocks of QM_FINE_BLOCK_CELLS square cells flagged as having it.
QM_DEV bool fine_block(const qm_terrain& T, double cp, double rp) {
    if (!T.fine_detail) return false;
    const int bx_n = ((int)T.columns - 2) / QM_FINE_BLOCK_CELLS + 1;
    const int by_n = ((int)T.rows - 2) / QM_FINE_BLOCK_CELLS + 1;
    const int bx = min(max((int)floor(cp / QM_FINE_BLOCK_CELLS), 0), bx_n - 1);
    const int by = min(max((int)floor(rp / QM_FINE_BLOCK_CELLS), 0), by_n - 1);
    return T.fine_blocks[by * bx_n + bx] != 0;
}

QM_DEV uint32_t fine_rows(const qm_terrain& T) { return QM_FINE_PER_NODE * (T.rows - 1) + 1; }
QM_DEV uint32_t fine_columns(const qm_terrain& T) { return QM_FINE_PER_NODE * (T.columns - 1) + 1; }

// The detail's bilinear value (m) at fine position (row, column).
QM_DEV float detail_at(const qm_terrain& T, float row, float column) {
    const uint32_t rows = fine_rows(T), columns = fine_columns(T);
    row = fmaxf(row, 0.0f);
    column = fmaxf(column, 0.0f);
    const uint32_t r = min((uint32_t)floorf(row), rows - 2), c = min((uint32_t)floorf(column), columns - 2);
    const float fr = row - (float)r, fc = column - (float)c;
    const int16_t* d = T.fine_detail;
    const size_t at = (size_t)r * columns + c;
    const float sum = (1.0f - fr) * ((1.0f - fc) * (float)d[at] + fc * (float)d[at + 1]) +
                      fr * ((1.0f - fc) * (float)d[at + columns] + fc * (float)d[at + columns + 1]);
    return sum * (float)QM_DETAIL_STEP_M;
}

// popup/src/scene.rs Ground::at: the coarse ground plus, in a block with fine terrain, its detail.
QM_DEV uint32_t ground_at_fine(const qm_terrain& T, double cp, double rp, GroundSample& out) {
    const uint32_t failed = ground_at(T, cp, rp, out);
    if (!failed && fine_block(T, cp, rp))
        out.z += detail_at(T, (float)(QM_FINE_PER_NODE * rp), (float)(QM_FINE_PER_NODE * cp));
    return failed;
}

// A vertex of the fine walk (popup/src/scene.rs Ground::fine_sample): on fine row line `line` (or
// fine column line `line` between fine rows `south` and south + 1), `position` along the fine
// lattice and `f` the fraction between the line's two fine nodes: the coarse cell's sample at the
// vertex's place plus, in a block with fine terrain, the detail between those nodes.
QM_DEV uint32_t fine_vertex(const qm_terrain& T, bool on_row, int line, int south, float position, float f,
                            GroundSample& out) {
    const int rows = (int)fine_rows(T), columns = (int)fine_columns(T);
    const int r0 = on_row ? line : south;
    const int c0 = on_row ? (int)floorf(position) : line;
    const int r1 = on_row ? r0 : r0 + 1, c1 = on_row ? c0 + 1 : c0;
    if (r0 < 0 || c0 < 0 || r0 >= rows || c0 >= columns || (f > 0.0f && (r1 >= rows || c1 >= columns)))
        return QM_FAILED_NO_TERRAIN;
    const double rp = (on_row ? (double)r0 : (double)r0 + (double)f) / QM_FINE_PER_NODE;
    const double cp = (on_row ? (double)c0 + (double)f : (double)c0) / QM_FINE_PER_NODE;
    const uint32_t failed = ground_at(T, cp, rp, out);
    if (failed) return failed;
    if (fine_block(T, cp, rp)) {
        const int16_t* d = T.fine_detail;
        float d0 = f < 1.0f ? (float)d[(size_t)r0 * columns + c0] : 0.0f;
        float d1 = f > 0.0f ? (float)d[(size_t)r1 * columns + c1] : 0.0f;
        if (f <= 0.0f) d1 = d0;
        if (f >= 1.0f) d0 = d1;
        out.z += (d0 + f * (d1 - d0)) * (float)QM_DETAIL_STEP_M;
    }
    return 0;
}

QM_DEV float t_of_column(const Lattice& L, int c) { return ((float)c - L.ca) * L.inv_columns; }
QM_DEV float t_of_row(const Lattice& L, int k) {
    if (L.geo) return (float)((row_line_y(L, k) - L.y0) / L.dy);
    return ((float)k - L.ra) * L.inv_rise;
}

// Sample on row line k at fraction t (the fine sample on a fine lattice when FINE).
template <bool FINE = false>
QM_DEV uint32_t row_vertex(const qm_terrain& T, const Lattice& L, int k, float t, GroundSample& out) {
    const float c = L.ca + t * L.columns;
    const float west = floorf(c);
    if (FINE && L.scale != 1) return fine_vertex(T, true, k, 0, c, c - west, out);
    return edge_sample(T, k, (int)west, k, (int)west + 1, c - west, out);
}
// Sample on column line c at fraction t, the cell's south row given.
template <bool FINE = false>
QM_DEV uint32_t column_vertex(const qm_terrain& T, const Lattice& L, int c, int south, float t, GroundSample& out) {
    const float r = L.ra + t * L.rise;
    float f = clampf(r - (float)south, 0.0f, 1.0f);
    if (FINE && L.geo) {
        // Between the row lines about the vertex, by north metres (scene.rs walk's fraction).
        const double y = L.y0 + (double)t * L.dy, y0 = row_line_y(L, south), y1 = row_line_y(L, south + 1);
        f = (float)clampd((y - y0) / (y1 - y0), 0.0, 1.0);
    }
    if (FINE && L.scale != 1) return fine_vertex(T, false, c, south, r, f, out);
    return edge_sample(T, south, c, south + 1, c, f, out);
}

// The walk's terrain vertex nearest before (or after) fraction t, directly: popup/src/scene.rs's
// lattice walk visits every line strictly between the ends; this finds the last line at or before
// t (`after` false) or the first strictly after t (`after` true) of each family and reads it.
// Returns the vertex fraction (0 or 1 for the ends, which the caller samples).
// With FINE the lattice is the stretch's that holds t, [from, to) of the ray, and a vertex at a
// stretch's bound (`from` before, `to` after) that is not an end is returned as is_end with t_out
// at that bound, for the caller to read (Ground::at, the walk's switch vertex).
template <bool FINE = false>
QM_DEV uint32_t bracket_vertex(const qm_terrain& T, const Lattice& L, float t, float length, bool after,
                               float& t_out, GroundSample& out, bool& is_end, float from = 0.0f, float to = 1.0f) {
    const float x = t * length;
    float best_t = after ? to : from;
    int kind = 0;  // 0 end, 1 column, 2 row
    int line = 0;
    if (L.columns != 0.0f) {
        const float pos = L.ca + t * L.columns;
        int c = L.column_step > 0 ? (int)floorf(pos) : (int)ceilf(pos);
        // Adjust to the last line at or before x (by the products the walk's vertices use).
        for (int k = 0; k < 2; ++k)
            if (t_of_column(L, c) * length > x) c -= L.column_step;
        for (int k = 0; k < 2; ++k)
            if (t_of_column(L, c + L.column_step) * length <= x) c += L.column_step;
        if (after) c += L.column_step;
        const bool started = L.column_step > 0 ? c >= L.first_column : c <= L.first_column;
        if (started) {
            const float tc = t_of_column(L, c);
            if (tc < to && tc > from) {
                if (after ? tc < best_t : tc > best_t) {
                    best_t = tc;
                    kind = 1;
                    line = c;
                }
            }
        }
    }
    if (L.rise != 0.0f) {
        const float pos = FINE ? row_units_at(L, t) : L.ra + t * L.rise;
        int k = L.row_step > 0 ? (int)floorf(pos) : (int)ceilf(pos);
        for (int j = 0; j < 2; ++j)
            if (t_of_row(L, k) * length > x) k -= L.row_step;
        for (int j = 0; j < 2; ++j)
            if (t_of_row(L, k + L.row_step) * length <= x) k += L.row_step;
        if (after) k += L.row_step;
        const bool started = L.row_step > 0 ? k >= L.first_row : k <= L.first_row;
        if (started) {
            const float tr = t_of_row(L, k);
            if (tr < to && tr > from) {
                // At a tie the row line's reading (the walk's rule).
                if (after ? tr <= best_t : tr >= best_t) {
                    best_t = tr;
                    kind = 2;
                    line = k;
                }
            }
        }
    }
    t_out = best_t;
    is_end = kind == 0;
    if (kind == 2) return row_vertex<FINE>(T, L, line, best_t, out);
    if (kind == 1) {
        const float r = FINE ? row_units_at(L, best_t) : L.ra + best_t * L.rise;
        const int south = L.rise >= 0.0f ? (int)floorf(r) : (int)ceilf(r) - 1;
        return column_vertex<FINE>(T, L, line, L.rise == 0.0f ? L.start_south : south, best_t, out);
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Obstacles: exact crossings, FP64, with the reference's operation order and no contraction so
// that every t rounds as the reference's f64 does.
struct Crossing {
    double t;
    uint64_t footprint;
    uint32_t outline;
    uint32_t edge;
    float height_m;
    uint32_t flags;  // bit 0 building, bit 1 bounds roof, bit 2 dropped (own or containing)
    float x, top;    // filled after the bracket reads
    float z_at, g_at;
    double pz, pxz, pg;  // prefix moments at x, filled by the walk
};

QM_DEV bool crossing_before(const Crossing& a, double t, uint64_t footprint, uint32_t outline, uint32_t edge) {
    if (a.t != t) return a.t < t;
    if (a.footprint != footprint) return a.footprint < footprint;
    if (a.outline != outline) return a.outline < outline;
    return a.edge < edge;
}

QM_DEV double side_of(const qm_obstacles& O, const double start[2], const double delta[2], size_t index) {
#pragma clang fp contract(off)
    const double vx = (double)O.vertices[2 * index], vy = (double)O.vertices[2 * index + 1];
    return delta[0] * (vy - start[1]) - delta[1] * (vx - start[0]);
}

// popup/src/obstacles/crossings.rs:80-111
QM_DEV bool edge_bounds_roof(const qm_obstacles& O, const double start[2], const double delta[2], size_t edge,
                             size_t first, size_t count) {
    const double a = side_of(O, start, delta, edge), b = side_of(O, start, delta, edge + 1);
    if (a != 0.0 && b != 0.0) return true;
    const size_t distinct = count - 1;
    const size_t vertex = ((a == 0.0 ? edge : edge + 1) - first) % distinct;
    double n0 = 0.0, n1 = 0.0;
    for (size_t offset = 1; offset < distinct; ++offset) {
        if (n0 == 0.0) n0 = side_of(O, start, delta, first + (vertex + distinct - offset) % distinct);
        if (n1 == 0.0) n1 = side_of(O, start, delta, first + (vertex + offset) % distinct);
        if (n0 != 0.0 && n1 != 0.0) break;
    }
    return n0 * n1 < 0.0;
}

QM_DEV double snap_steps(double steps) {
    const double nearest = rint(steps);
    return fabs(steps - nearest) < 1e-6 ? nearest : steps;
}

// popup/src/obstacles/crossings.rs:163-230: the cell walk, the padded piece box and the exact
// test; every hit inserted in (t, footprint, outline, edge) order. Returns failure bits.
QM_DEV uint32_t find_crossings(const Scene& sc, double fx, double fy, double tx, double ty, Crossing* out,
                               int& n) {
#pragma clang fp contract(off)
    const qm_obstacles& O = sc.obstacles;
    n = 0;
    const double mpy = O.metres_per_step_y > 0.0 ? O.metres_per_step_y : O.metres_per_step;
    const double start[2] = {snap_steps(O.origin_steps_x + fx / O.metres_per_step),
                             snap_steps(O.origin_steps_y - fy / mpy)};
    const double end[2] = {snap_steps(O.origin_steps_x + tx / O.metres_per_step),
                           snap_steps(O.origin_steps_y - ty / mpy)};
    const double delta[2] = {end[0] - start[0], end[1] - start[1]};
    const double cs = (double)QM_CELL_STEPS;
    int64_t cell[2] = {(int64_t)floor(start[0] / cs), (int64_t)floor(start[1] / cs)};
    const int64_t last[2] = {(int64_t)floor(end[0] / cs), (int64_t)floor(end[1] / cs)};
    const int step[2] = {delta[0] >= 0.0 ? 1 : -1, delta[1] >= 0.0 ? 1 : -1};
    const double inf = __builtin_inf();
    double t_delta[2], t_max[2];
    for (int a = 0; a < 2; ++a) {
        t_delta[a] = delta[a] != 0.0 ? fabs(cs / delta[a]) : inf;
        if (delta[a] == 0.0) {
            t_max[a] = inf;
        } else {
            const double boundary = (double)(cell[a] + (delta[a] >= 0.0 ? 1 : 0)) * cs;
            t_max[a] = fabs((boundary - start[a]) / delta[a]);
        }
    }
    const double pad = 1e-9 * (1.0 + fabs(delta[0]) + fabs(delta[1]));
    double t_enter = 0.0;
    uint32_t failed = 0;
    // The occupancy word of the block the walk is in (re-read when the walk leaves the block).
    uint32_t held_block = 0xffffffffu;
    uint64_t held_word = 0;
    for (;;) {
        const int64_t gx = cell[0] - O.cell_x0, gy = cell[1] - O.cell_y0;
        if (gx < 0 || gy < 0 || gx >= (int64_t)O.cells_x || gy >= (int64_t)O.cells_y) return QM_FAILED_NOT_READ;
        bool occupied = true;
        if (sc.cell_blocks) {
            const uint32_t block = (uint32_t)(gy >> 3) * sc.blocks_x + (uint32_t)(gx >> 3);
            if (block != held_block) {
                held_word = sc.cell_blocks[block];
                held_block = block;
            }
            occupied = (held_word >> (8 * (gy & 7) + (gx & 7))) & 1ull;
        }
        uint32_t r0 = 0, r1 = 0;
        if (occupied) {
            const size_t index = (size_t)gy * O.cells_x + (size_t)gx;
            r0 = O.cell_first_run[index];
            r1 = O.cell_first_run[index + 1];
        }
        if (r1 > r0) {
            const double t_exit = fmin(fmin(t_max[0], t_max[1]), 1.0);
            const double te = clampd(t_enter, 0.0, 1.0), tx2 = clampd(t_exit, 0.0, 1.0);
            const double p0 = start[0] + delta[0] * te, p1 = start[1] + delta[1] * te;
            const double q0 = start[0] + delta[0] * tx2, q1 = start[1] + delta[1] * tx2;
            const int32_t low0 = (int32_t)floor(fmin(p0, q0) - pad), high0 = (int32_t)ceil(fmax(p0, q0) + pad);
            const int32_t low1 = (int32_t)floor(fmin(p1, q1) - pad), high1 = (int32_t)ceil(fmax(p1, q1) + pad);
            for (uint32_t ri = r0; ri < r1; ++ri) {
                const qm_run run = O.runs[ri];
                const size_t first = sc.run_vertex ? (size_t)sc.run_vertex[ri]
                                                   : (size_t)O.outlines[run.outline].first_vertex + run.first_edge;
                int32_t ax = O.vertices[2 * first], ay = O.vertices[2 * first + 1];
                for (size_t next = first + 1; next <= first + run.edge_count; ++next) {
                    const int32_t bx = O.vertices[2 * next], by = O.vertices[2 * next + 1];
                    const bool meets = high0 >= min(ax, bx) && max(ax, bx) >= low0 && high1 >= min(ay, by) &&
                                       max(ay, by) >= low1;
                    if (meets) {
                        const double ex = (double)bx - (double)ax, ey = (double)by - (double)ay;
                        const double den = delta[0] * ey - delta[1] * ex;
                        if (den != 0.0) {
                            const double wx = (double)ax - start[0], wy = (double)ay - start[1];
                            const double t = (wx * ey - wy * ex) / den;
                            const double u = (wx * delta[1] - wy * delta[0]) / den;
                            if (t > 0.0 && t < 1.0 && u >= 0.0 && u <= 1.0) {
                                if (n >= CROSS_CAP) {
                                    failed |= QM_FAILED_CAPACITY | QM_DETAIL_CROSSINGS;
                                } else {
                                    const qm_outline& record = O.outlines[run.outline];
                                    const uint32_t edge = (uint32_t)(next - 1);
                                    int k = n;
                                    while (k > 0 && !crossing_before(out[k - 1], t, record.footprint_id, run.outline, edge)) {
                                        out[k] = out[k - 1];
                                        --k;
                                    }
                                    Crossing c;
                                    c.t = t;
                                    c.footprint = record.footprint_id;
                                    c.outline = run.outline;
                                    c.edge = edge;
                                    c.height_m = (float)((double)record.height_dm / 10.0);
                                    const bool building = record.kind == QM_KIND_BUILDING;
                                    const bool roof = building && edge_bounds_roof(O, start, delta, edge,
                                                                                   record.first_vertex,
                                                                                   record.vertex_count);
                                    c.flags = (building ? 1u : 0u) | (roof ? 2u : 0u);
                                    out[k] = c;
                                    ++n;
                                }
                            }
                        }
                    }
                    ax = bx;
                    ay = by;
                }
            }
        }
        if (cell[0] == last[0] && cell[1] == last[1]) break;
        t_enter = fmin(t_max[0], t_max[1]);
        int axis;
        if (cell[0] == last[0]) axis = 1;
        else if (cell[1] == last[1] || t_max[0] < t_max[1]) axis = 0;
        else axis = 1;
        t_max[axis] += t_delta[axis];
        cell[axis] += step[axis];
    }
    // crossings.rs:34-56: one crossing per footprint within 1e-9 of the one kept before it.
    int kept = 0;
    for (int i = 0; i < n; ++i) {
        const Crossing c = out[i];
        bool merged = false;
        for (int k = kept - 1; k >= 0; --k) {
            if (!(c.t - out[k].t < 1e-9)) break;
            if (out[k].footprint == c.footprint) {
                out[k].flags |= (c.flags & 2u);
                merged = true;
                break;
            }
        }
        if (!merged) out[kept++] = c;
    }
    n = kept;
    return failed;
}

// ---------------------------------------------------------------------------------------------
// The vertical plane. P2 = (x, z).
struct P2f {
    float x, z;
};
struct P2d {
    double x, z;
};

// A diffraction candidate: its point, the profile's altitude and ground factor just after it, and
// the prefix moments of the profile from the source up to it (about (0, z0)). `id` is its entry
// in the path's candidate log once a state has kept it, -1 before.
struct Cand {
    float x, z, z_at, g_at;
    double pz, pxz, pg;
    int id;
};

// What the boundary terms read of a kept candidate beyond its point (one log entry).
struct CandData {
    float z_at, g_at;
    double pz, pxz, pg;
};

// A kept candidate as the states hold it: its point and its log entry.
struct CandRef {
    float x, z;
    int id;
};

// physics/src/cnossos/rubber_band.rs:182-193 (FP64; the FP32 form with its legs is below).
QM_DEV double triangle_excess_d(P2d s, P2d o, P2d r) {
    const double ux = o.x - s.x, uz = o.z - s.z, vx = r.x - o.x, vz = r.z - o.z;
    const double lu = sqrt(ux * ux + uz * uz), lv = sqrt(vx * vx + vz * vz);
    const double wx = ux + vx, wz = uz + vz;
    const double lw = sqrt(wx * wx + wz * wz);
    const double dot = ux * vx + uz * vz;
    if (dot <= 0.0) return lu + lv - lw;
    const double cross = ux * vz - uz * vx;
    return 2.0 * cross * cross / ((lu * lv + dot) * (lu + lv + lw));
}
QM_DEV float dist_f(P2f a, P2f b) {
    const float dx = b.x - a.x, dz = b.z - a.z;
    return fsqrt(dx * dx + dz * dz);
}
QM_DEV double dist_d(P2d a, P2d b) {
    const double dx = b.x - a.x, dz = b.z - a.z;
    return sqrt(dx * dx + dz * dz);
}

// rubber_band.rs:65-80 arc_excess (favourable state only; radius > 0 marks favourable).
QM_DEV float arc_excess_f(float radius, float chord) {
    if (radius <= 0.0f) return 0.0f;
    const float x = chord * frcp(2.0f * radius);
    if (x > 1.0f / 16.0f) return 2.0f * radius * asinf(fminf(x, 1.0f)) - chord;
    const float x2 = x * x;
    return 2.0f * radius * x * x2 *
           (1.0f / 6.0f + x2 * (3.0f / 40.0f + x2 * (5.0f / 112.0f + x2 * (35.0f / 1152.0f))));
}
QM_DEV double arc_excess_d(double radius, double chord) {
    if (radius <= 0.0) return 0.0;
    const double x = chord / (2.0 * radius);
    if (x > 1.0 / 16.0) return 2.0 * radius * asin(fmin(x, 1.0)) - chord;
    const double x2 = x * x;
    return 2.0 * radius * x * x2 * (1.0 / 6.0 + x2 * (3.0 / 40.0 + x2 * (5.0 / 112.0 + x2 * (35.0 / 1152.0))));
}

// triangle_excess_f with the two legs' lengths |so| and |or| it forms.
QM_DEV float triangle_excess_legs_f(P2f s, P2f o, P2f r, float& lu, float& lv) {
    const float ux = o.x - s.x, uz = o.z - s.z, vx = r.x - o.x, vz = r.z - o.z;
    lu = fsqrt(ux * ux + uz * uz);
    lv = fsqrt(vx * vx + vz * vz);
    const float wx = ux + vx, wz = uz + vz;
    const float lw = fsqrt(wx * wx + wz * wz);
    const float dot = ux * vx + uz * vz;
    if (dot <= 0.0f) return lu + lv - lw;
    const float cross = ux * vz - uz * vx;
    return fdiv(2.0f * cross * cross, (lu * lv + dot) * (lu + lv + lw));
}

// rubber_band.rs:86-121 path_difference over n points, FP64 (the delta of the boundary terms).
QM_DEV double path_difference_d(double radius, P2d from, const P2d* points, int n, P2d to) {
    if (n == 0) return 0.0;
    double chords = 0.0, arcs = 0.0;
    P2d start = from;
    for (int i = 0; i < n; ++i) {
        chords += triangle_excess_d(start, points[i], to);
        arcs += arc_excess_d(radius, dist_d(start, points[i]));
        start = points[i];
    }
    arcs += arc_excess_d(radius, dist_d(start, to)) - arc_excess_d(radius, dist_d(from, to));
    const double excess = chords + arcs;
    if (n > 1) return excess;
    const P2d first = points[0];
    const double chord = from.z + (to.z - from.z) * (first.x - from.x) / (to.x - from.x);
    if (first.z >= chord) return excess;
    if (radius <= 0.0) return -excess;
    const P2d a{first.x, chord};
    return 2.0 * triangle_excess_d(from, a, to) - triangle_excess_d(from, first, to) +
           2.0 * arc_excess_d(radius, dist_d(from, a)) + 2.0 * arc_excess_d(radius, dist_d(a, to)) -
           arc_excess_d(radius, dist_d(from, first)) - arc_excess_d(radius, dist_d(first, to)) -
           arc_excess_d(radius, dist_d(from, to));
}

// rubber_band.rs:125-138, 349-359: the favourable ray's height above the chord from `from` to
// `to` at horizontal distance `dx` from `from` (0 in calm air, radius 0).
QM_DEV float sag_f(float radius, P2f from, P2f to, float dx) {
    if (radius <= 0.0f) return 0.0f;
    const float horizontal = to.x - from.x;
    const float chord = dist_f(from, to);
    const float gamma2 = radius * radius;
    const float along_per_x = fdiv(chord, horizontal);
    const float centre_term = fsqrt(gamma2 - 0.25f * chord * chord);
    const float along = dx * along_per_x;
    const float offset = along - 0.5f * chord;
    const float perpendicular = fdiv(along * (chord - along), fsqrt(fmaxf(gamma2 - offset * offset, 0.0f)) + centre_term);
    return perpendicular * along_per_x;
}

// A state's ray from S to R with its per-path constants taken once (rubber_band.rs:125-138): the
// altitude of the straight chord, and the favourable arc's height above it.
struct StateLine {
    float sx, sz, slope;        // chord: sz + slope * (x - sx)
    float along_per_x, chord, gamma2, centre_term;
    bool favourable;
};

QM_DEV StateLine state_line(float radius, P2f from, P2f to) {
    StateLine L;
    L.sx = from.x;
    L.sz = from.z;
    L.slope = fdiv(to.z - from.z, to.x - from.x);
    L.favourable = radius > 0.0f;
    L.chord = dist_f(from, to);
    L.gamma2 = radius * radius;
    L.along_per_x = fdiv(L.chord, to.x - from.x);
    L.centre_term = L.favourable ? fsqrt(L.gamma2 - 0.25f * L.chord * L.chord) : 0.0f;
    return L;
}

// The state's ray altitude at horizontal distance x (Sag::at, rubber_band.rs:349-359).
QM_DEV float line_altitude(const StateLine& L, float x) {
    const float dx = x - L.sx;
    float z = L.sz + L.slope * dx;
    if (L.favourable) {
        const float along = dx * L.along_per_x;
        const float offset = along - 0.5f * L.chord;
        z += fdiv(along * (L.chord - along), fsqrt(fmaxf(L.gamma2 - offset * offset, 0.0f)) + L.centre_term) *
             L.along_per_x;
    }
    return z;
}

// rubber_band.rs:86-121 path_difference with one point o (FP32, the candidates' screening) for
// a state's line S -> R. The excess over the chord is the triangle excess plus the arcs' excesses
// over the legs; below the straight chord in the favourable state it is mirrored about the chord
// point A = (o.x, chord(o.x)), whose legs lie on the chord itself (T(S, A, R) = 0, |SA| = along,
// |AR| = |SR| - along), so the arcs come from lengths the line already holds.
QM_DEV float single_path_difference_f(const StateLine& L, float radius, float arc_sr, P2f from, P2f o, P2f to) {
    float lu, lv;
    const float chords = triangle_excess_legs_f(from, o, to, lu, lv);
    const float excess = chords + arc_excess_f(radius, lu) + arc_excess_f(radius, lv) - arc_sr;
    const float dx = o.x - L.sx;
    const float chord = L.sz + L.slope * dx;
    if (o.z >= chord) return excess;
    if (radius <= 0.0f) return -excess;
    const float along = dx * L.along_per_x;
    return 2.0f * (arc_excess_f(radius, along) + arc_excess_f(radius, L.chord - along) - arc_sr) - excess;
}

// ---------------------------------------------------------------------------------------------
// Mean planes (mean_plane.rs:19-120) from FP64 prefix moments.
struct Plane {
    double slope, intercept;
};

// The plane over [a, b] given the prefixes at both ends and the profile's altitude just after a.
QM_DEV Plane plane_over(double a, double b, double z0, double z_at_a, double pz_a, double pxz_a, double pz_b,
                        double pxz_b) {
    const double length = b - a;
    if (length < 1e-3) return Plane{0.0, z_at_a};
    const double inv = 1.0 / length;
    const double dz = pz_b - pz_a;
    const double iz = dz - (z_at_a - z0) * length;
    const double ixz = (pxz_b - pxz_a) - a * dz - (z_at_a - z0) * 0.5 * length * length;
    const double slope = 12.0 * (ixz - 0.5 * length * iz) * (inv * inv * inv);
    const double intercept_relative = iz * inv - 0.5 * slope * length;
    return Plane{slope, z_at_a + intercept_relative - slope * a};
}

struct Geometry {
    Plane plane;
    float hs, hr;        // end heights over the plane, negative taken as 0
    bool s_below, r_below;
    float dp;            // projected distance
};

// mean_plane.rs:109-120
QM_DEV Geometry geometry_over(Plane p, P2d from, P2d to) {
    const double inv_norm = 1.0 / sqrt(1.0 + p.slope * p.slope);
    const double hs = (from.z - (p.slope * from.x + p.intercept)) * inv_norm;
    const double hr = (to.z - (p.slope * to.x + p.intercept)) * inv_norm;
    Geometry g;
    g.plane = p;
    g.hs = (float)fmax(hs, 0.0);
    g.hr = (float)fmax(hr, 0.0);
    g.s_below = hs < 0.0;
    g.r_below = hr < 0.0;
    g.dp = (float)(fabs((to.x - from.x) + p.slope * (to.z - from.z)) * inv_norm);
    return g;
}

// mean_plane.rs:91-96
QM_DEV P2d mirror(Plane p, P2d q) {
    const double norm_sq = 1.0 + p.slope * p.slope;
    const double offset = (q.z - (p.slope * q.x + p.intercept)) / norm_sq;
    return P2d{q.x + 2.0 * offset * p.slope, q.z - 2.0 * offset};
}

// ground.rs:39-51
QM_DEV float near_source_factor(const Geometry& g, float path_g, float source_g) {
    const float height_sum = fmaxf(g.hs + g.hr, MINIMUM_HEIGHT_SUM_M);
    const float ratio = fdiv(g.dp, SHORT_PATH_HEIGHT_FACTOR * height_sum);
    return ratio <= 1.0f ? path_g * ratio + source_g * (1.0f - ratio) : path_g;
}

// ground.rs:78-89, with the hardware reciprocal and square root (1 ulp) and the band's 1/k from
// the table: 4k^2/dp^2 taken as 4(k/dp)^2.
QM_DEV float analytic_ground_db(const BandTable& B, int band, float dp, float inv_dp, float zs, float zr, float gw13,
                                float gw26) {
    const float w = fdiv(0.0185f * B.f_2_5[band] * gw26, B.f_1_5[band] * gw26 + 1.3e3f * B.f_0_75[band] * gw13 + 1.16e6f);
    const float wd = w * dp;
    const float cf = fdiv(dp * (1.0f + 3.0f * wd * __expf(-fsqrt(wd))), 1.0f + wd);
    const float ck = cf * B.inv_wavenumber[band];
    const float root = fsqrt(2.0f * ck);
    const float product = (zs * zs - root * zs + ck) * (zr * zr - root * zr + ck);
    const float kd = B.wavenumber[band] * inv_dp;
    return -10.0f * __log10f(4.0f * kd * kd * product);
}

// ground.rs:105-137 A_ground of one (sub-)path in one state, per band.
QM_DEV void ground_bands(const BandTable& B, const Geometry& g, float path, float impedance, float floor_factor,
                         bool favourable, float out[BANDS]) {
    const float dp = fmaxf(g.dp, 1e-6f);
    const float inv_dp = frcp(dp);
    const float zs = g.hs, zr = g.hr;
    const float gw13 = __powf(fmaxf(impedance, 0.0f), 1.3f);
    const float gw26 = gw13 * gw13;
    if (!favourable) {
        if (path < HARD_GROUND_BELOW) {
            for (int b = 0; b < BANDS; ++b) out[b] = -3.0f;
            return;
        }
        const float floor = -3.0f * (1.0f - floor_factor);
        for (int b = 0; b < BANDS; ++b) out[b] = fmaxf(analytic_ground_db(B, b, dp, inv_dp, zs, zr, gw13, gw26), floor);
        return;
    }
    const float height_sum0 = fmaxf(g.hs + g.hr, MINIMUM_HEIGHT_SUM_M);
    const float base = -3.0f * (1.0f - floor_factor);
    const float limit = SHORT_PATH_HEIGHT_FACTOR * height_sum0;
    const float floor = g.dp <= limit ? base : base * (1.0f + 2.0f * (1.0f - fdiv(limit, g.dp)));
    if (path < HARD_GROUND_BELOW) {
        for (int b = 0; b < BANDS; ++b) out[b] = floor;
        return;
    }
    const float height_sum = fmaxf(zs + zr, MINIMUM_HEIGHT_SUM_M);
    const float curvature = FAVOURABLE_CURVATURE_A0_PER_M * dp * dp * 0.5f;
    const float inv_height_sum = frcp(height_sum);
    const float terrain = FAVOURABLE_TERRAIN_HEIGHT_COEFFICIENT * dp * inv_height_sum;
    const float rs = zs * inv_height_sum, rr = zr * inv_height_sum;
    const float zsf = zs + curvature * (rs * rs) + terrain;
    const float zrf = zr + curvature * (rr * rr) + terrain;
    for (int b = 0; b < BANDS; ++b) out[b] = fmaxf(analytic_ground_db(B, b, dp, inv_dp, zsf, zrf, gw13, gw26), floor);
}

// diffraction.rs:19-30
QM_DEV float diffraction_db(const BandTable& B, int band, float delta, float c_second) {
    const float x = 40.0f * c_second * delta * B.inv_wavelength[band];
    return x >= -2.0f ? 10.0f * __log10f(3.0f + x) : 0.0f;
}
// diffraction.rs:33-35
QM_DEV float ground_split_db(float ground_db, float mirrored_db, float direct_db) {
    return -20.0f * __log10f(1.0f + (amplitude_of(-ground_db) - 1.0f) * amplitude_of(-(mirrored_db - direct_db)));
}

// The per-state streaming rubber band and unblocked-best search (rubber_band.rs:202-264).
struct StateStream {
    StateLine line;  // the state's ray from S to R
    float radius;  // 0 calm; the favourable arc's radius
    float arc_sr;  // the arc's excess over the whole chord S -> R
    bool blocking;
    bool has_pending;
    bool have_best;
    int hull_n;    // entries of the state's hull (kept apart, HULL_CAP CandRefs) after the source
    float best_delta;
    CandRef pending;
    CandRef best;
    CandRef top[2];  // the hull's last two entries (top[1] the last), mirrored from its buffer
};

// Work-items of a group of every kernel that inlines the path walk: 1-D groups of at most this
// many (each such kernel's __launch_bounds__ names it, and its launches are dim3(n) x dim3(k <=
// QM_PATH_GROUP)). The LDS below costs sizeof(StateStream) x 2 x QM_PATH_GROUP bytes a group.
#define QM_PATH_GROUP 256

// The two states of every work-item of a group, in LDS: the scalars touched at every candidate,
// read and written with LDS instructions wherever the walk is inlined (a pointer to them carried
// in the per-path buffers made them generic accesses that waited on the scratch traffic).
static __shared__ StateStream qm_lane_states[QM_PATH_GROUP][2];
QM_DEV StateStream* lane_states() { return qm_lane_states[threadIdx.x]; }

// A state's hull: the diffraction points of its rubber band so far, source excluded.
struct Hull {
    CandRef e[HULL_CAP];
};

// The candidate log of one path: the data of every candidate a state has kept. Its count is the
// walk's own register (`log_n`), not a field, so that a kept candidate waits on no memory read.
struct CandLog {
    CandData entry[LOG_CAP];
};

// Drops the entries no state refers to any more and renumbers the rest in order (the log is full;
// at most 2*HULL_CAP + 4 entries are live, so this frees at least 12).
// Returns the new count. Called where a candidate enters (at most one new entry follows).
QM_DEV int log_compact(CandLog& log, StateStream* states, Hull* hulls) {
    uint64_t live = 0;
    for (int s = 0; s < 2; ++s) {
        const StateStream& st = states[s];
        for (int i = 0; i < st.hull_n; ++i) live |= 1ull << hulls[s].e[i].id;
        if (st.has_pending) live |= 1ull << st.pending.id;
        if (st.have_best && !st.blocking) live |= 1ull << st.best.id;
    }
    int next = 0;
    for (int i = 0; i < LOG_CAP; ++i)
        if ((live >> i) & 1ull) log.entry[next++] = log.entry[i];
    for (int s = 0; s < 2; ++s) {
        StateStream& st = states[s];
        for (int i = 0; i < st.hull_n; ++i)
            hulls[s].e[i].id = __popcll(live & ((1ull << hulls[s].e[i].id) - 1ull));
        if (st.hull_n >= 1) st.top[1] = hulls[s].e[st.hull_n - 1];
        if (st.hull_n >= 2) st.top[0] = hulls[s].e[st.hull_n - 2];
        if (st.has_pending) st.pending.id = __popcll(live & ((1ull << st.pending.id) - 1ull));
        if (st.have_best && !st.blocking) st.best.id = __popcll(live & ((1ull << st.best.id) - 1ull));
    }
    return next;
}

// The candidate as a state keeps it, logging its data on first use (shared by both states; the
// caller leaves a free entry).
QM_DEV CandRef keep(CandLog& log, int& log_n, Cand& c) {
    if (c.id < 0) {
        c.id = log_n++;
        CandData& d = log.entry[c.id];
        d.z_at = c.z_at;
        d.g_at = c.g_at;
        d.pz = c.pz;
        d.pxz = c.pxz;
        d.pg = c.pg;
    }
    return CandRef{c.x, c.z, c.id};
}

QM_DEV void hull_push(StateStream& st, Hull& hull, P2f S, const CandRef& c, uint32_t& failed) {
    while (st.hull_n >= 1) {
        const P2f a = st.hull_n >= 2 ? P2f{st.top[0].x, st.top[0].z} : S;
        const P2f b{st.top[1].x, st.top[1].z};
        const P2f p{c.x, c.z};
        const float ray_at_b = a.z + fdiv((p.z - a.z) * (b.x - a.x), p.x - a.x) + sag_f(st.radius, a, p, b.x - a.x);
        if (b.z <= ray_at_b) {
            --st.hull_n;
            st.top[1] = st.top[0];
            if (st.hull_n >= 2) st.top[0] = hull.e[st.hull_n - 2];
        } else {
            break;
        }
    }
    if (st.hull_n >= HULL_CAP) {
        failed |= QM_FAILED_CAPACITY | QM_DETAIL_HULL;
        return;
    }
    hull.e[st.hull_n++] = c;
    st.top[0] = st.top[1];
    st.top[1] = c;
}

// One candidate (in order of distance) into state `s`'s stream.
QM_DEV void state_candidate(StateStream* states, Hull* hulls, int s, CandLog& log, int& log_n, P2f S, P2f R, Cand& c,
                            uint32_t& failed) {
    StateStream& st = states[s];
    const P2f p{c.x, c.z};
    if (p.z > line_altitude(st.line, p.x)) {
        if (st.has_pending && st.pending.x == p.x) {
            if (p.z > st.pending.z) st.pending = keep(log, log_n, c);
        } else {
            if (st.has_pending) hull_push(st, hulls[s], S, st.pending, failed);
            st.pending = keep(log, log_n, c);
            st.has_pending = true;
        }
        st.blocking = true;
        return;
    }
    if (st.blocking) return;  // a blocked state's unblocked-best is never read
    const float delta = single_path_difference_f(st.line, st.radius, st.arc_sr, S, p, R);
    if (!st.have_best || delta > st.best_delta) {
        st.best_delta = delta;
        st.best = keep(log, log_n, c);
        st.have_best = true;
    }
}

// ---------------------------------------------------------------------------------------------
// The merged profile as the walk streams it: the last vertex emitted with its prefix moments
// (integrals from the source of z - z0, x*(z - z0) and G), and the top segment of the last roof
// emitted, whose near end keeps its prefix, for the walk vertices that roof covers.
struct Vertex {
    float x, z, g;
    double pz, pxz, pg;
};

struct Profile {
    Vertex head;       // the last vertex emitted
    Vertex roof_near;  // the last roof's near-wall top
    float roof_x1, roof_top1;
    double z0;         // altitude of the first vertex (the moments' origin)
};

// The vertex (x, z, g) after `a`, its prefix moments extended over the segment between them.
QM_DEV Vertex extend(const Vertex& a, float x, float z, float g, double z0) {
    Vertex v{x, z, g, a.pz, a.pxz, a.pg};
    const double u0 = a.x, u1 = x;
    const double width = u1 - u0;
    if (width > 0.0) {
        const double w0 = (double)a.z - z0, w1 = (double)z - z0;
        v.pz += 0.5 * (w0 + w1) * width;
        v.pxz += width * (u0 * (2.0 * w0 + w1) + u1 * (w0 + 2.0 * w1)) * (1.0 / 6.0);
        v.pg += 0.5 * ((double)a.g + (double)g) * width;
    }
    return v;
}

// The profile at x on the segment from `a` toward (bx, bz, bg), a.x <= x < bx (the upper side of
// a step: the last vertex at or before x, interpolated toward the next), with the prefix to x.
// The prefix takes the point on the segment in FP64: a receiver-side plane over the few
// centimetres from a wall to the receiver is fitted from the difference of the end's prefix and
// this one, both spanning the whole segment, and an FP32-rounded point there (1e-5 m at 200 m)
// turned that difference, and the fitted slope, into noise.
QM_DEV void segment_at(const Vertex& a, float bx, float bz, float bg, float x, double z0, Cand& c) {
    if (a.x == x) {
        c.z_at = a.z;
        c.g_at = a.g;
        c.pz = a.pz;
        c.pxz = a.pxz;
        c.pg = a.pg;
        return;
    }
    const float f = fdiv(x - a.x, bx - a.x);
    c.z_at = a.z + f * (bz - a.z);
    c.g_at = a.g + f * (bg - a.g);
    const double u0 = a.x, u1 = x;
    const double fd = (u1 - u0) / ((double)bx - u0);
    const double w0 = (double)a.z - z0, w1 = (double)a.z + fd * ((double)bz - (double)a.z) - z0;
    const double gd = (double)a.g + fd * ((double)bg - (double)a.g);
    const double width = u1 - u0;
    c.pz = a.pz + 0.5 * (w0 + w1) * width;
    c.pxz = a.pxz + width * (u0 * (2.0 * w0 + w1) + u1 * (w0 + 2.0 * w1)) * (1.0 / 6.0);
    c.pg = a.pg + 0.5 * ((double)a.g + gd) * width;
}

// The profile at a walk vertex's x just after the walk has handled it: the vertex itself when it
// was emitted, else the roof covering it (x in (near wall, far wall]).
QM_DEV void profile_at_walk(const Profile& P, float x, Cand& c) {
    if (P.head.x == x) segment_at(P.head, x, P.head.z, P.head.g, x, P.z0, c);
    else segment_at(P.roof_near, P.roof_x1, P.roof_top1, 0.0f, x, P.z0, c);
}

struct Roof {
    float x0, x1, top0, top1;
    float ground0, g0, ground1, g1;  // terrain and G under each wall
};

// The result of one path.
struct PathResult {
    float bands[PERIODS][BANDS];  // received A-weighted energy per period and band (weighted)
    float states[2][PERIODS];     // calm and favourable alone, per period
    float att[2][BANDS];          // A_ground + A_dif per state and band (the boundary terms), dB
    uint32_t failed;
};

// ---------------------------------------------------------------------------------------------
// One state's boundary term (diffraction.rs:66-196), FP64 deltas, FP32 bands.
QM_DEV void state_boundary(const BandTable& B, StateStream& st, Hull& hull, const CandLog& log, bool favourable, P2f Sf, P2f Rf, double z0,
                           double L, double pz_L, double pxz_L, double pg_L, const Geometry& direct,
                           float path_ground, float near_source, float source_g, float g_start,
                           uint32_t& failed, float att[BANDS]) {
    // Whole path (Table 2.5.b: calm G'/G', favourable G/G').
    ground_bands(B, direct, path_ground, favourable ? path_ground : near_source, near_source, favourable, att);
    // The diffraction points.
    if (st.blocking) {
        if (st.has_pending) hull_push(st, hull, Sf, st.pending, failed);
        // The receiver closes the band.
        hull_push(st, hull, Sf, CandRef{Rf.x, Rf.z, -1}, failed);
        --st.hull_n;  // drop the receiver itself
    } else if (st.have_best) {
        hull.e[0] = st.best;
        st.hull_n = 1;
    } else {
        return;  // no candidate at all
    }
    const int n = st.hull_n;
    if (n <= 0) return;
    P2d pts[HULL_CAP];
    for (int i = 0; i < n; ++i) pts[i] = P2d{(double)hull.e[i].x, (double)hull.e[i].z};
    const P2d S{(double)Sf.x, (double)Sf.z}, R{L, (double)Rf.z};
    const double radius = favourable ? (double)st.radius : 0.0;
    const double delta = path_difference_d(radius, S, pts, n, R);
    double span = 0.0;
    for (int i = 0; i + 1 < n; ++i) span += dist_d(pts[i], pts[i + 1]) + arc_excess_d(radius, dist_d(pts[i], pts[i + 1]));
    const CandRef& first = hull.e[0];
    const CandRef& last = hull.e[n - 1];
    const CandData& fd = log.entry[first.id];
    const CandData& ld = log.entry[last.id];
    const Plane ps = plane_over(0.0, first.x, z0, z0, 0.0, 0.0, fd.pz, fd.pxz);
    const Plane pr = plane_over(last.x, L, z0, ld.z_at, ld.pz, ld.pxz, pz_L, pxz_L);
    const Geometry gs = geometry_over(ps, S, pts[0]);
    const Geometry gr = geometry_over(pr, pts[n - 1], R);
    const P2d s_image = mirror(ps, S), r_image = mirror(pr, R);
    const float d_direct = (float)delta;
    const float d_s = (float)path_difference_d(radius, s_image, pts, n, R);
    const float d_r = (float)path_difference_d(radius, S, pts, n, r_image);
    const bool blocked = st.blocking;
    const float rayleigh_star = blocked ? 0.0f : (float)path_difference_d(radius, s_image, pts, n, r_image);
    // Which bands diffraction takes (Rayleigh, 2021/1226 point (9)(c)).
    bool any = false;
    bool admitted[BANDS];
    for (int b = 0; b < BANDS; ++b) {
        const float lambda = B.wavelength[b];
        admitted[b] = blocked || (d_direct > -lambda / 20.0f && d_direct > lambda / 4.0f - rayleigh_star);
        any = any || admitted[b];
    }
    if (!any) return;
    const float span_f = (float)span;
    const double ls = (double)first.x;
    const float g_so = ls < 1e-3 ? g_start : (float)(fd.pg / ls);
    const double lr = L - (double)last.x;
    const float g_or = lr < 1e-3 ? ld.g_at : (float)((pg_L - ld.pg) / lr);
    const float near_so = near_source_factor(gs, g_so, source_g);
    float ground_so[BANDS], ground_or[BANDS];
    ground_bands(B, gs, g_so, favourable ? g_so : near_so, near_so, favourable, ground_so);
    ground_bands(B, gr, g_or, g_or, g_or, favourable, ground_or);
    for (int b = 0; b < BANDS; ++b) {
        if (!admitted[b]) continue;
        float c_second = 1.0f;
        if (n > 1 && span_f > MULTIPLE_DIFFRACTION_MINIMUM_SPAN_M) {
            const float x = 5.0f * B.wavelength[b] * frcp(span_f);
            c_second = fdiv(1.0f + x * x, 1.0f / 3.0f + x * x);
        }
        const float direct_db = diffraction_db(B, b, d_direct, c_second);
        const float s_db = diffraction_db(B, b, d_s, c_second);
        const float r_db = diffraction_db(B, b, d_r, c_second);
        float sr = direct_db;
        float split_so;
        if (gs.s_below) {
            sr = s_db;
            split_so = ground_so[b];
        } else {
            split_so = ground_split_db(ground_so[b], s_db, direct_db);
        }
        if (!isfinite(split_so)) {
            split_so = ground_so[b];
            sr = s_db;
        }
        float split_or;
        if (gr.r_below) {
            sr = r_db;
            split_or = ground_or[b];
        } else {
            split_or = ground_split_db(ground_or[b], r_db, sr);
        }
        if (!isfinite(split_or)) {
            split_or = ground_or[b];
            sr = r_db;
        }
        att[b] = clampf(sr, 0.0f, DIFFRACTION_CAP_DB) + split_so + split_or;
    }
}

}  // namespace qm
