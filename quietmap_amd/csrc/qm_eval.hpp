// One receiver-source pair end to end on the device: popup/src/evaluate.rs:171-221 (a point's
// one ray), 300-330 (the ray), 119-152 (add_ray) and paint/src/exact.rs:186-205 (whole_energy),
// over physics/src/ray.rs:125-354. See qm_path.hpp for the pieces and the design.
#pragma once

#include "qm_path.hpp"

namespace qm {

// Phase clocks, compiled only into the instrumented build (QM_PHASE_CLOCKS, libqmpaint_phases.so)
// that the profiling workload uses to split a path's time: each mark adds the shader clock cycles
// since the wave's previous mark to the phase that mark ends (the wave's first active lane keeps
// the wave's totals in LDS), so the phases partition every wave's time; a kernel that calls
// phase_start and phase_flush adds its waves' totals to the device's. Each translation unit keeps
// its own.
enum PathPhase {
    PHASE_OUTSIDE,     // the caller's loop (tile queue, source loop, sums)
    PHASE_CROSSINGS,   // obstacle crossings, own and containing footprints
    PHASE_SETUP,       // the ends' ground, wall tops' terrain, roofs, the states' lines
    PHASE_VERTEX,      // the lattice walk's next vertex and its terrain nodes
    PHASE_PROFILE,     // roofs and the vertex into the merged profile, wall tops' reads, foliage
    PHASE_CREST,       // crest thinning, the tops' merge ahead of each candidate
    PHASE_CANDIDATES,  // both states' streams over the candidates
    PHASE_TAIL,        // the last tops, the direct path's plane
    PHASE_BOUNDARY,    // both states' boundary terms
    PHASE_MIX,         // foliage, air absorption, the period mix
    PHASE_COUNT
};
#ifdef QM_PHASE_CLOCKS
static __device__ unsigned long long qm_phase_cycles[PHASE_COUNT];
static __shared__ unsigned long long qm_phase_last[QM_PATH_GROUP / 64];
static __shared__ unsigned long long qm_phase_acc[QM_PATH_GROUP / 64][PHASE_COUNT];
QM_DEV bool phase_lane() { return (int)(threadIdx.x & 63) == __builtin_ctzll(__builtin_amdgcn_read_exec()); }
// At a path kernel's start: the wave's totals cleared and its clock taken.
QM_DEV void phase_start() {
    if (phase_lane()) {
        for (int k = 0; k < PHASE_COUNT; ++k) qm_phase_acc[threadIdx.x >> 6][k] = 0;
        qm_phase_last[threadIdx.x >> 6] = __builtin_readcyclecounter();
    }
}
QM_DEV void phase_mark(PathPhase k) {
    const unsigned long long now = __builtin_readcyclecounter();
    if (phase_lane()) {
        qm_phase_acc[threadIdx.x >> 6][k] += now - qm_phase_last[threadIdx.x >> 6];
        qm_phase_last[threadIdx.x >> 6] = now;
    }
}
// At the kernel's end: the wave's totals into the device's.
QM_DEV void phase_flush() {
    phase_mark(PHASE_OUTSIDE);
    if (phase_lane())
        for (int k = 0; k < PHASE_COUNT; ++k) atomicAdd(&qm_phase_cycles[k], qm_phase_acc[threadIdx.x >> 6][k]);
}
#else
QM_DEV void phase_start() {}
QM_DEV void phase_mark(PathPhase) {}
QM_DEV void phase_flush() {}
#endif

// The per-path buffers (private to the work-item).
struct PathScratch {
    Crossing cross[CROSS_CAP];
    Roof roofs[ROOF_CAP];
    CandLog log;
    Hull hull[2];
};

// The crest thinning of physics/src/ray.rs:360-385: the open group's best sample.
struct CrestGroup {
    bool open;
    int group;
    Cand best;
};

// Wall tops in distance order: the next to read the profile at (`resolved`) and the next to
// emit as a candidate (`emitted`), with their x (+inf when none is left).
struct Tops {
    int resolved, emitted;
    float resolve_x, emit_x;
};

// Favourable probability of `period` for sound travelling along `azimuth` (weather.rs:61-70).
QM_DEV float favourable_at(const float* row, float azimuth) {
    float bearing = fmodf(90.0f - azimuth * (180.0f / 3.14159265358979f), 360.0f);
    if (bearing < 0.0f) bearing += 360.0f;
    const float position = bearing * (SECTORS / 360.0f);
    const float fl = floorf(position);
    const int lower = ((int)fl) % SECTORS;
    const int upper = (lower + 1) % SECTORS;
    const float fraction = position - fl;
    return row[lower] + fraction * (row[upper] - row[lower]);
}

QM_DEV void candidate_out(PathScratch& w, int& log_n, P2f S, P2f R, float length, Cand& c, uint32_t& failed) {
    if (!(c.x > 0.0f && c.x < length)) return;
    phase_mark(PHASE_CREST);
    // A candidate adds at most one log entry (both states share it): room for one.
    if (log_n >= LOG_CAP) log_n = log_compact(w.log, lane_states(), w.hull);
    c.id = -1;
    state_candidate(lane_states(), w.hull, 0, w.log, log_n, S, R, c, failed);
    state_candidate(lane_states(), w.hull, 1, w.log, log_n, S, R, c, failed);
    phase_mark(PHASE_CANDIDATES);
}

// The x of the first top at or after index i not dropped (+inf when none).
QM_DEV float top_x_from(const PathScratch& w, int& i, int n) {
    while (i < n && (w.cross[i].flags & 4u)) ++i;
    return i < n ? w.cross[i].x : __builtin_inff();
}

// Emits every resolved top before x (the candidate merge of physics/src/cnossos/mod.rs:160-184:
// a top first only when strictly nearer). Inlined: out of line, its accesses to the per-path
// buffers went through generic pointers (flat instructions waiting on all memory traffic), which
// cost the bench step 30 %.
QM_DEV void tops_before_slow(PathScratch& w, int& log_n, Tops& tp, int n, P2f S, P2f R, float length, float x,
                             uint32_t& failed) {
    while (tp.emit_x < x && tp.emitted < tp.resolved) {
        const Crossing& c = w.cross[tp.emitted];
        Cand k;
        k.x = c.x;
        k.z = c.top;
        k.z_at = c.z_at;
        k.g_at = c.g_at;
        k.pz = c.pz;
        k.pxz = c.pxz;
        k.pg = c.pg;
        candidate_out(w, log_n, S, R, length, k, failed);
        ++tp.emitted;
        tp.emit_x = top_x_from(w, tp.emitted, n);
    }
}
QM_DEV void tops_before(PathScratch& w, int& log_n, Tops& tp, int n, P2f S, P2f R, float length, float x,
                        uint32_t& failed) {
    if (tp.emit_x < x && tp.emitted < tp.resolved) tops_before_slow(w, log_n, tp, n, S, R, length, x, failed);
}

// Reads the profile for every top before x on the segment from the profile's head to the vertex
// (x, z, g) about to be emitted.
__device__ void resolve_tops_slow(PathScratch& w, Tops& tp, int n, const Profile& P, float x, float z, float g) {
    while (tp.resolve_x < x) {
        Crossing& c = w.cross[tp.resolved];
        Cand k;
        segment_at(P.head, x, z, g, c.x, P.z0, k);
        c.z_at = k.z_at;
        c.g_at = k.g_at;
        c.pz = k.pz;
        c.pxz = k.pxz;
        c.pg = k.pg;
        ++tp.resolved;
        tp.resolve_x = top_x_from(w, tp.resolved, n);
    }
}

// Appends a vertex to the merged profile (tops before it read their profile first).
QM_DEV void profile_emit(PathScratch& w, Tops& tp, int n, Profile& P, float x, float z, float g) {
    if (tp.resolve_x < x) resolve_tops_slow(w, tp, n, P, x, z, g);
    P.head = extend(P.head, x, z, g, P.z0);
}

QM_DEV void crest_vertex(PathScratch& w, int& log_n, const Profile& P, CrestGroup& cg, Tops& tp, int n, P2f S, P2f R,
                         float length, float x, float z, uint32_t& failed) {
    const bool middle = x > CREST_ENDS_M && x < length - CREST_ENDS_M;
    const int group = (int)floorf(x / CREST_SPACING_M);
    if (cg.open && (!middle || cg.group != group)) {
        tops_before(w, log_n, tp, n, S, R, length, cg.best.x, failed);
        candidate_out(w, log_n, S, R, length, cg.best, failed);
        cg.open = false;
    }
    if (!middle) {
        Cand k;
        profile_at_walk(P, x, k);
        k.x = x;
        k.z = z;
        tops_before(w, log_n, tp, n, S, R, length, x, failed);
        candidate_out(w, log_n, S, R, length, k, failed);
        return;
    }
    if (!cg.open || z > cg.best.z) {
        profile_at_walk(P, x, cg.best);
        cg.best.x = x;
        cg.best.z = z;
        cg.group = group;
        cg.open = true;
    }
}

// The path's crossings (exact, FP64) with flag bit 2 on those that never screen it: the receiver's
// own walls (evaluate.rs:293-297) and, for a source in a building, the walls of every footprint
// containing it (ray.rs:201-226).
QM_DEV uint32_t path_crossings(const Scene& sc, const qm_receiver& rcv, const qm_source& src, Crossing* cross, int& n) {
    n = 0;
    const uint32_t failed = find_crossings(sc, src.x_m, src.y_m, rcv.x_m, rcv.y_m, cross, n);
    if (failed & QM_FAILED_NOT_READ) return failed;
    for (int i = 0; i < n; ++i) {
        cross[i].flags &= 3u;
        if (rcv.own_footprint != 0 && cross[i].footprint == rcv.own_footprint) cross[i].flags |= 4u;
    }
    if (src.footprint_id != 0) {
        for (int i = 0; i < n; ++i) {
            if (cross[i].flags & 4u) continue;
            const uint64_t id = cross[i].footprint;
            int count = 0;
            for (int j = 0; j < n; ++j)
                if (!(cross[j].flags & 4u) && (cross[j].flags & 2u) && cross[j].footprint == id) ++count;
            if ((count & 1) || id == src.footprint_id) cross[i].flags |= 8u;
        }
        for (int i = 0; i < n; ++i)
            if (cross[i].flags & 8u) cross[i].flags = (cross[i].flags & ~8u) | 4u;
    }
    return failed;
}

// A ray's foliage depth in one state (foliage.rs:36-64), streamed over the terrain samples.
struct Foliage {
    float depth[2];
    float prev_x, prev_cover, prev_ground, prev_canopy;
};

// Points are taken lazily: a stretch counts only where either end has cover, so a ray over open
// ground computes nothing here.
QM_DEV void foliage_sample(Foliage& f, const PathScratch& w, float x, float ground, float cover, float canopy,
                           bool first) {
    if (first || !(f.prev_cover > 0.0f || cover > 0.0f)) {
        f.prev_x = x;
        f.prev_cover = cover;
        f.prev_ground = ground;
        f.prev_canopy = canopy;
        return;
    }
    for (int s = 0; s < 2; ++s) {
        const float za = line_altitude(lane_states()[s].line, f.prev_x), zb = line_altitude(lane_states()[s].line, x);
        const float above_a = za - f.prev_ground, above_b = zb - ground;
        const float wa = (above_a > 0.0f && above_a <= f.prev_canopy) ? f.prev_cover : 0.0f;
        const float wb = (above_b > 0.0f && above_b <= canopy) ? cover : 0.0f;
        const float dx = x - f.prev_x, dz = zb - za;
        f.depth[s] += 0.5f * (wa + wb) * fsqrt(dx * dx + dz * dz);
    }
    f.prev_x = x;
    f.prev_cover = cover;
    f.prev_ground = ground;
    f.prev_canopy = canopy;
}

// The ray's stretches by lattice (popup/src/scene.rs:242-309 Ground::fill_profile): within
// QM_FINE_ENDS_M of either end the fine lattice where a block under that stretch has fine terrain,
// the coarse one elsewhere, neighbours on one lattice joined. `to` ends each stretch (the last 1).
struct StretchPlan {
    int count;
    int scale[3];
    float to[3];
    double to_d[3];
};

// popup/src/scene.rs Ground::fine_near: a block at a corner of the stretch's box has fine terrain.
QM_DEV bool fine_near(const qm_terrain& T, double ax, double ay, double bx, double by) {
    const double xs[2] = {fmin(ax, bx), fmax(ax, bx)}, ys[2] = {fmin(ay, by), fmax(ay, by)};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            if (fine_block(T, column_position(T, xs[i]), row_position<true>(T, ys[j]))) return true;
    return false;
}

QM_DEV StretchPlan stretch_plan(const qm_terrain& T, double sx, double sy, double rx, double ry, double length) {
    StretchPlan P;
    P.count = 1;
    P.scale[0] = 1;
    P.to[0] = 1.0f;
    P.to_d[0] = 1.0;
    if (!T.fine_detail) return P;
    if (length <= 2.0 * QM_FINE_ENDS_M) {
        P.scale[0] = fine_near(T, sx, sy, rx, ry) ? QM_FINE_PER_NODE : 1;
        return P;
    }
    const double e = QM_FINE_ENDS_M / length;
    const double ax = sx + e * (rx - sx), ay = sy + e * (ry - sy);
    const double bx = sx + (1.0 - e) * (rx - sx), by = sy + (1.0 - e) * (ry - sy);
    const int scales[3] = {fine_near(T, sx, sy, ax, ay) ? QM_FINE_PER_NODE : 1, 1,
                           fine_near(T, bx, by, rx, ry) ? QM_FINE_PER_NODE : 1};
    const double ends[3] = {e, 1.0 - e, 1.0};
    P.count = 0;
    for (int k = 0; k < 3; ++k) {
        if (P.count > 0 && P.scale[P.count - 1] == scales[k]) {
            P.to_d[P.count - 1] = ends[k];
        } else {
            P.scale[P.count] = scales[k];
            P.to_d[P.count] = ends[k];
            ++P.count;
        }
    }
    for (int k = 0; k < P.count; ++k) P.to[k] = (float)P.to_d[k];
    return P;
}

// The switch vertex at fraction `to` of the ray (Ground::at where the lattice changes).
QM_DEV uint32_t switch_vertex(const qm_terrain& T, double sx, double sy, double rx, double ry, double to,
                              GroundSample& out) {
    const double x = sx - to * (sx - rx), y = sy - to * (sy - ry);
    return ground_at_fine(T, column_position(T, x), row_position<true>(T, y), out);
}

// The ray's ends and stretches, for terrain_bracket on a fine scene.
struct RayEnds {
    double sx, sy, rx, ry;
    StretchPlan plan;
};

// terrain_at and ground_factor_at of physics/src/ray.rs:265-283 at x = t * length, from the two
// walk vertices bracketing it (platform-clamped as the terrain polyline is). With FINE, within the
// stretch that holds t, a stretch's switch vertex standing in where no line lies between.
template <bool FINE = false>
QM_DEV uint32_t terrain_bracket(const qm_terrain& T, const Lattice& L, float t, float length, float platform,
                                float source_ground, const GroundSample& start, const GroundSample& end, float& z_out,
                                float& g_out, const RayEnds* ray = nullptr) {
    float tp, tn;
    GroundSample a, b;
    bool ea, eb;
    uint32_t failed;
    if (FINE && ray && !(ray->plan.count == 1 && ray->plan.scale[0] == 1)) {
        const StretchPlan& P = ray->plan;
        int k = 0;
        while (k + 1 < P.count && !(t < P.to[k])) ++k;
        const float from = k ? P.to[k - 1] : 0.0f, to = P.to[k];
        const Lattice Lk = lattice_of(T, ray->sx, ray->sy, ray->rx, ray->ry, P.scale[k], from, true);
        failed = bracket_vertex<true>(T, Lk, t, length, false, tp, a, ea, from, to);
        failed |= bracket_vertex<true>(T, Lk, t, length, true, tn, b, eb, from, to);
        if (ea) {
            if (k == 0) a = start;
            else failed |= switch_vertex(T, ray->sx, ray->sy, ray->rx, ray->ry, P.to_d[k - 1], a);
        }
        if (eb) {
            if (k + 1 == P.count) b = end;
            else failed |= switch_vertex(T, ray->sx, ray->sy, ray->rx, ray->ry, P.to_d[k], b);
        }
    } else {
        failed = bracket_vertex<FINE>(T, L, t, length, false, tp, a, ea);
        failed |= bracket_vertex<FINE>(T, L, t, length, true, tn, b, eb);
        if (ea) a = start;
        if (eb) b = end;
    }
    if (failed) return failed;
    const float x = t * length;
    const float xa = tp * length, xb = tn * length;
    const float za = xa < platform ? fminf(a.z, source_ground) : a.z;
    const float zb = xb < platform ? fminf(b.z, source_ground) : b.z;
    const float f = xb > xa ? clampf(fdiv(x - xa, xb - xa), 0.0f, 1.0f) : 0.0f;
    z_out = za + f * (zb - za);
    g_out = a.g + f * (b.g - a.g);
    return 0;
}

// paint/src/exact.rs:95-118 Point::with_receiver_properties and physics/src/weather.rs:163-214.
QM_DEV qm_receiver receiver_of(const Scene& sc, const qm_point& p) {
    qm_receiver r;
    r.x_m = p.x_m;
    r.y_m = p.y_m;
    r.own_footprint = p.own_footprint;
    const qm_terrain& T = sc.terrain;
    GroundSample g;
    r.failed = ground_at_fine(T, column_position(T, p.x_m), row_position<true>(T, p.y_m), g);
    r.altitude_m = r.failed ? 0.0f : g.z + (float)QM_RECEIVER_HEIGHT_M;
    const qm_weather& W = sc.weather;
    uint32_t row, col, r1, c1;
    float fy, fx;
    if (W.nodes_per_degree) {
        // The place's latitude and longitude (geo::Mercator::to_degrees), then the world table's
        // four nodes about it, longitude wrapping (weather.rs corners).
        const double n = 4096.0;
        const double mx = W.frame_x + p.x_m / W.east_m_per_unit, my = W.frame_y - p.y_m / W.north_m_per_unit;
        const double lat = atan(sinh(3.14159265358979323846 * (1.0 - 2.0 * my / n))) * (180.0 / 3.14159265358979323846);
        const double lon = mx / n * 360.0 - 180.0;
        const double npd = (double)W.nodes_per_degree;
        const double yy = clampd((90.0 - lat) * npd, 0.0, (double)(W.rows - 1));
        double lw = fmod(lon, 360.0);
        if (lw < 0.0) lw += 360.0;
        const double xx = lw * npd;
        row = min((uint32_t)floor(yy), W.rows - 2);
        const uint32_t column = (uint32_t)floor(xx);
        fy = (float)(yy - row);
        fx = (float)(xx - column);
        col = column % W.columns;
        c1 = (column + 1) % W.columns;
        r1 = row + 1;
    } else {
        const double yy = clampd((p.y_m - W.origin_y_m) / W.spacing_m, 0.0, (double)(W.rows - 1));
        const double xx = clampd((p.x_m - W.origin_x_m) / W.spacing_m, 0.0, (double)(W.columns - 1));
        row = min((uint32_t)floor(yy), W.rows >= 2 ? W.rows - 2 : 0u);
        col = min((uint32_t)floor(xx), W.columns >= 2 ? W.columns - 2 : 0u);
        fy = (float)(yy - row);
        fx = (float)(xx - col);
        r1 = W.rows >= 2 ? row + 1 : row;
        c1 = W.columns >= 2 ? col + 1 : col;
    }
    const uint8_t* nd[4] = {W.nodes + ((size_t)row * W.columns + col) * QM_WEATHER_NODE_BYTES,
                            W.nodes + ((size_t)row * W.columns + c1) * QM_WEATHER_NODE_BYTES,
                            W.nodes + ((size_t)r1 * W.columns + col) * QM_WEATHER_NODE_BYTES,
                            W.nodes + ((size_t)r1 * W.columns + c1) * QM_WEATHER_NODE_BYTES};
    const float w[4] = {(1.0f - fy) * (1.0f - fx), (1.0f - fy) * fx, fy * (1.0f - fx), fy * fx};
    for (int q = 0; q < PERIODS; ++q)
        for (int s = 0; s < SECTORS; ++s) {
            float total = 0.0f;
            for (int k = 0; k < 4; ++k) total += w[k] * (float)nd[k][q * SECTORS + s];
            r.favourable[q][s] = total * 0.01f;
        }
    for (int b = 0; b < BANDS; ++b) {
        float total = 0.0f;
        const int at = PERIODS * SECTORS + 2 * b;
        for (int k = 0; k < 4; ++k) total += w[k] * (float)((uint32_t)nd[k][at] | ((uint32_t)nd[k][at + 1] << 8));
        r.alpha_db_per_km[b] = total * 0.01f;
    }
    return r;
}

// How a ray is taken: whether walls are searched on it (a line node inside a clear gap of its
// skyline mask has none: popup/src/evaluate.rs:300-314) and its weight (negative: the point
// source's divergence, evaluate.rs:195-203; a line node's dphi times the piece's divergence
// factor otherwise, evaluate.rs:245-265).
struct RayOptions {
    bool obstacles = true;
    float weight = -1.0f;
};

// FINE: the scene may hold fine terrain (qm_terrain.fine_detail): the walk's stretches near the
// ends go over the fine lattice (popup/src/scene.rs:242-309); without it the coarse walk alone.
// FOREST: the terrain may have forest cover; a square with none takes the instantiation without
// the foliage term's per-vertex state (its depths stay 0, as they would over open ground).
template <bool FINE = false, bool FOREST = true>
QM_DEV void evaluate_path(const Scene& sc, const BandTable& B, const qm_receiver& rcv, const qm_source& src,
                          PathScratch& w, PathResult& out, const RayOptions& opt = RayOptions{}) {
    phase_mark(PHASE_OUTSIDE);
    uint32_t failed = rcv.failed;
    for (int p = 0; p < PERIODS; ++p) {
        for (int b = 0; b < BANDS; ++b) out.bands[p][b] = 0.0f;
        out.states[0][p] = out.states[1][p] = 0.0f;
    }
    for (int b = 0; b < BANDS; ++b) out.att[0][b] = out.att[1][b] = 0.0f;
    out.failed = failed;
    if (failed) return;
    const qm_terrain& T = sc.terrain;
    const double sx = src.x_m, sy = src.y_m, rx = rcv.x_m, ry = rcv.y_m;
    if (!isfinite(sx) || !isfinite(sy) || !isfinite(rx) || !isfinite(ry)) {
        out.failed = QM_FAILED_NOT_READ;
        return;
    }
    const double length_d = hypot(sx - rx, sy - ry);
    const float H = (float)fmax(length_d, 1.0);

    int n = 0;
    if (opt.obstacles) failed |= path_crossings(sc, rcv, src, w.cross, n);
    phase_mark(PHASE_CROSSINGS);
    if (failed & QM_FAILED_NOT_READ) {
        out.failed = failed;
        return;
    }

    // The ends (tiles/src/terrain.rs:342-385 at both ends of popup/src/scene.rs:242-309).
    const Lattice L = lattice_of(T, sx, sy, rx, ry, 1, 0.0f, FINE);
    GroundSample s0, r1;
    RayEnds ray;
    if (FINE) {
        ray = RayEnds{sx, sy, rx, ry, stretch_plan(T, sx, sy, rx, ry, length_d)};
        failed |= ground_at_fine(T, column_position(T, sx), row_position<true>(T, sy), s0);
        failed |= ground_at_fine(T, column_position(T, rx), row_position<true>(T, ry), r1);
    } else {
        failed |= ground_at(T, column_position(T, sx), row_position(T, sy), s0);
        failed |= ground_at(T, column_position(T, rx), row_position(T, ry), r1);
    }
    if (failed & QM_FAILED_NO_TERRAIN) {
        out.failed = failed;
        return;
    }
    // The source stands on a bridge's deck where it has one, else on the terrain (ray.rs:133-134,
    // 259-268 source_ground_m).
    const float source_ground = src.on_deck ? src.deck_m : s0.z;
    const float source_alt = source_ground + src.height_m;
    const float receiver_alt = rcv.altitude_m;
    const float platform = src.platform_half_width_m;
    const float source_g = src.ground_percent != QM_GROUND_FROM_TERRAIN ? (float)src.ground_percent * 0.01f : s0.g;

    // Tops on the terrain polyline (ray.rs:304-316) and the roofs they bound (ray.rs:317-335).
    int nroofs = 0;
    for (int i = 0; i < n; ++i) {
        Crossing& c = w.cross[i];
        if (c.flags & 4u) continue;
        const float t = (float)c.t;
        c.x = (float)(c.t * (double)H);
        float z, g;
        failed |= terrain_bracket<FINE>(T, L, t, H, platform, source_ground, s0, r1, z, g, FINE ? &ray : nullptr);
        c.top = z + c.height_m;
        c.z_at = z;  // terrain under the wall, until the walk resolves the profile there
        c.g_at = g;
    }
    // Pair each footprint's bounding crossings in order of distance.
    for (int i = 0; i < n; ++i) w.cross[i].flags &= ~16u;
    for (int i = 0; i < n; ++i) {
        Crossing& a = w.cross[i];
        if ((a.flags & 4u) || !(a.flags & 2u) || (a.flags & 16u)) continue;
        a.flags |= 16u;
        int open = i;  // an entry waiting for its exit
        for (int j = i + 1; j < n; ++j) {
            Crossing& b = w.cross[j];
            if ((b.flags & 4u) || !(b.flags & 2u) || (b.flags & 16u) || b.footprint != a.footprint) continue;
            b.flags |= 16u;
            if (open < 0) {
                open = j;
                continue;
            }
            if (nroofs >= ROOF_CAP) {
                failed |= QM_FAILED_CAPACITY | QM_DETAIL_ROOFS;
            } else {
                const Crossing& e = w.cross[open];
                Roof r;
                r.x0 = e.x;
                r.x1 = b.x;
                r.top0 = e.top;
                r.top1 = b.top;
                r.ground0 = e.z_at;
                r.g0 = e.g_at;
                r.ground1 = b.z_at;
                r.g1 = b.g_at;
                // Insert in closing order (x1, then x0), stable.
                int k = nroofs;
                while (k > 0 && (w.roofs[k - 1].x1 > r.x1 || (w.roofs[k - 1].x1 == r.x1 && w.roofs[k - 1].x0 > r.x0))) {
                    w.roofs[k] = w.roofs[k - 1];
                    --k;
                }
                w.roofs[k] = r;
                ++nroofs;
            }
            open = -1;
        }
    }
    // Clip in closing order (ray.rs:389-404).
    {
        float covered = -__builtin_inff();
        float covered_ground = 0.0f, covered_g = 0.0f;
        int kept = 0;
        for (int i = 0; i < nroofs; ++i) {
            Roof r = w.roofs[i];
            if (r.x1 <= covered) continue;
            if (r.x0 < covered) {
                const float span = r.x1 - r.x0;
                r.top0 += fdiv(covered - r.x0, span) * (r.top1 - r.top0);
                r.x0 = covered;
                r.ground0 = covered_ground;
                r.g0 = covered_g;
            }
            covered = r.x1;
            covered_ground = r.ground1;
            covered_g = r.g1;
            w.roofs[kept++] = r;
        }
        nroofs = kept;
    }

    // The walk: one pass streaming the merged profile, the candidates and both states.
    const P2f S{0.0f, source_alt};
    const P2f R{H, receiver_alt};
    {
        const float direct = sqrtf(H * H + (receiver_alt - source_alt) * (receiver_alt - source_alt));
        lane_states()[0].radius = 0.0f;
        lane_states()[1].radius = fmaxf(1000.0f, 8.0f * direct);
        for (int s = 0; s < 2; ++s) {
            lane_states()[s].line = state_line(lane_states()[s].radius, S, R);
            lane_states()[s].arc_sr = arc_excess_f(lane_states()[s].radius, lane_states()[s].line.chord);
        }
        for (int s = 0; s < 2; ++s) {
            lane_states()[s].blocking = false;
            lane_states()[s].has_pending = false;
            lane_states()[s].have_best = false;
            lane_states()[s].hull_n = 0;
            lane_states()[s].best_delta = 0.0f;
        }
    }
    int log_n = 0;
    Tops tp;
    tp.resolved = tp.emitted = 0;
    tp.resolve_x = top_x_from(w, tp.resolved, n);
    tp.emitted = tp.resolved;
    tp.emit_x = tp.resolve_x;
    CrestGroup cg;
    cg.open = false;
    cg.group = 0;
    Foliage fol;
    fol.depth[0] = fol.depth[1] = 0.0f;
    int r = 0;
    float next_roof = nroofs > 0 ? w.roofs[0].x0 : __builtin_inff();
    float covered_to = 0.0f;

    // Vertex 0: the source's ground (cut to a deck under it within the platform).
    Profile P;
    const float ground0 = 0.0f < platform ? fminf(s0.z, source_ground) : s0.z;
    P.z0 = ground0;
    P.head = Vertex{0.0f, ground0, s0.g, 0.0, 0.0, 0.0};
    P.roof_near = P.head;
    P.roof_x1 = P.roof_top1 = 0.0f;
    if (FOREST) foliage_sample(fol, w, 0.0f, s0.z, s0.forest, s0.canopy, true);
    crest_vertex(w, log_n, P, cg, tp, n, S, R, H, 0.0f, ground0, failed);

    // The walk's lattice: the coarse one, or with FINE each stretch's in turn.
    Lattice WL = L;
    int stretch = 0;
    float stretch_to = 1.0f;
    if (FINE && !(ray.plan.count == 1 && ray.plan.scale[0] == 1)) {
        WL = lattice_of(T, sx, sy, rx, ry, ray.plan.scale[0], 0.0f, true);
        stretch_to = ray.plan.to[0];
    }
    int next_column = WL.first_column, next_row = WL.first_row, south = WL.start_south;
    phase_mark(PHASE_SETUP);
    const float inf = __builtin_inff();
    // Every line crossed is one vertex: a stretch ends within |columns| + |rise| + 2 steps of its
    // lattice over its share of the ray.
    int budget = (int)(WL.span * stretch_to) + 4;
    bool last = false;
    while (!last) {
        if (--budget < 0) {
            failed |= QM_FAILED_CAPACITY | QM_DETAIL_WALK;
            break;
        }
        float t;
        GroundSample v;
        const float t_column = WL.columns != 0.0f ? t_of_column(WL, next_column) : inf;
        const float t_row = WL.rise != 0.0f ? t_of_row(WL, next_row) : inf;
        t = fminf(t_column, t_row);
        if (t >= stretch_to) {
            if (!FINE || stretch_to >= 1.0f) {
                t = 1.0f;
                v = r1;
                last = true;
            } else {
                // The switch vertex where the lattice changes, then the next stretch's lattice.
                t = stretch_to;
                failed |= switch_vertex(T, sx, sy, rx, ry, ray.plan.to_d[stretch], v);
                const float from = stretch_to;
                ++stretch;
                stretch_to = ray.plan.to[stretch];
                WL = lattice_of(T, sx, sy, rx, ry, ray.plan.scale[stretch], from, true);
                next_column = WL.first_column;
                next_row = WL.first_row;
                south = WL.start_south;
                budget += (int)(WL.span * (stretch_to - from)) + 4;
                if (failed & QM_FAILED_NO_TERRAIN) break;
            }
        } else {
            if (FINE) {
                if (t_row <= t_column) failed |= row_vertex<FINE>(T, WL, next_row, t, v);
                else failed |= column_vertex<FINE>(T, WL, next_column, south, t, v);
            } else {
                // row_vertex or column_vertex with the line's two nodes chosen by selects, so the
                // lanes of a wave that meet rows and columns in different order read in one pass.
                const bool on_row = t_row <= t_column;
                const float c = WL.ca + t * WL.columns;
                const float west = floorf(c);
                const float r = WL.ra + t * WL.rise;
                const int r0 = on_row ? next_row : south, r1 = on_row ? next_row : south + 1;
                const int c0 = on_row ? (int)west : next_column, c1 = on_row ? (int)west + 1 : next_column;
                const float f = on_row ? c - west : clampf(r - (float)south, 0.0f, 1.0f);
                failed |= edge_sample(T, r0, c0, r1, c1, f, v);
            }
            if (t_row <= t_column) {
                if (WL.rise > 0.0f) {
                    south = next_row;
                    next_row += 1;
                } else {
                    south = next_row - 1;
                    next_row -= 1;
                }
            }
            if (t_column <= t_row) next_column += WL.column_step;
            if (failed & QM_FAILED_NO_TERRAIN) break;
        }
        phase_mark(PHASE_VERTEX);
        const float x = t * H;
        const float z = x < platform ? fminf(v.z, source_ground) : v.z;
        if (FOREST) foliage_sample(fol, w, x, v.z, v.forest, v.canopy, false);
        // Roofs starting before this vertex, then the vertex unless a roof covers it.
        while (next_roof < x) {
            const Roof f = w.roofs[r];
            profile_emit(w, tp, n, P, f.x0, f.ground0, f.g0);
            profile_emit(w, tp, n, P, f.x0, f.top0, 0.0f);
            P.roof_near = P.head;
            profile_emit(w, tp, n, P, f.x1, f.top1, 0.0f);
            profile_emit(w, tp, n, P, f.x1, f.ground1, f.g1);
            P.roof_x1 = f.x1;
            P.roof_top1 = f.top1;
            covered_to = f.x1;
            ++r;
            next_roof = r < nroofs ? w.roofs[r].x0 : inf;
        }
        if (x > covered_to) profile_emit(w, tp, n, P, x, z, v.g);
        phase_mark(PHASE_PROFILE);
        crest_vertex(w, log_n, P, cg, tp, n, S, R, H, x, z, failed);
        phase_mark(PHASE_CREST);
    }
    phase_mark(PHASE_CREST);
    if (failed & (QM_FAILED_NO_TERRAIN | QM_FAILED_CAPACITY)) {
        out.failed = failed;
        return;
    }
    if (cg.open) {
        tops_before(w, log_n, tp, n, S, R, H, cg.best.x, failed);
        candidate_out(w, log_n, S, R, H, cg.best, failed);
    }
    // Tops at the last vertex's x take it; none can lie beyond it.
    while (tp.resolved < n) {
        Crossing& c = w.cross[tp.resolved];
        if (!(c.flags & 4u)) {
            if (c.x != P.head.x) {
                failed |= QM_FAILED_CAPACITY | QM_DETAIL_PROFILE;
                break;
            }
            c.z_at = P.head.z;
            c.g_at = P.head.g;
            c.pz = P.head.pz;
            c.pxz = P.head.pxz;
            c.pg = P.head.pg;
        }
        ++tp.resolved;
    }
    tops_before(w, log_n, tp, n, S, R, H, inf, failed);

    // The direct path (diffraction.rs:46-63) from the whole profile's prefix.
    const Vertex& end = P.head;
    const double Ld = (double)H;
    const double z0 = P.z0;
    const Plane direct_plane = plane_over(0.0, Ld, z0, z0, 0.0, 0.0, end.pz, end.pxz);
    const Geometry direct = geometry_over(direct_plane, P2d{0.0, (double)source_alt}, P2d{Ld, (double)receiver_alt});
    const float path_ground = (float)(end.pg / Ld);
    const float near_source = near_source_factor(direct, path_ground, source_g);
    float att[2][BANDS];
    phase_mark(PHASE_TAIL);
    for (int s = 0; s < 2; ++s)
        state_boundary(B, lane_states()[s], w.hull[s], w.log, s == 1, S, R, z0, Ld, end.pz, end.pxz, end.pg, direct, path_ground,
                       near_source, source_g, s0.g, failed, att[s]);
    for (int b = 0; b < BANDS; ++b) {
        out.att[0][b] = att[0][b];
        out.att[1][b] = att[1][b];
    }

    phase_mark(PHASE_BOUNDARY);
    // ray.rs:166-190: foliage, air at the slant distance, the period mix; evaluate.rs:119-152.
    const float slant = fmaxf(sqrtf(H * H + (receiver_alt - source_alt) * (receiver_alt - source_alt)), 1.0f);
    const float azimuth = atan2f((float)(ry - sy), (float)(rx - sx));
    float p[PERIODS];
    for (int q = 0; q < PERIODS; ++q) p[q] = favourable_at(rcv.favourable[q], azimuth);
    // evaluate.rs:195-203: the point divergence.
    const float hd = fmaxf((float)length_d, src.exclusion_radius_m);
    const float dz = receiver_alt - source_alt;
    const float d = fmaxf(sqrtf(hd * hd + dz * dz), 1.0f);
    const float weight = opt.weight >= 0.0f ? opt.weight : frcp(d * d * 12.589254117941673f);
    float fol_db[2][BANDS];
    for (int s = 0; s < 2; ++s) {
        const float depth = fol.depth[s];
        for (int b = 0; b < BANDS; ++b)
            fol_db[s][b] = depth < 10.0f ? 0.0f
                           : depth < 20.0f ? B.foliage_short_db[b]
                                           : B.foliage_db_per_m[b] * fminf(depth, 200.0f);
    }
    for (int b = 0; b < BANDS; ++b) {
        const float air = __expf(-rcv.alpha_db_per_km[b] * slant * 1e-3f * (LN10 / 10.0f));
        const float eh = air * __expf(-(att[0][b] + fol_db[0][b]) * (LN10 / 10.0f));
        const float ef = air * __expf(-(att[1][b] + fol_db[1][b]) * (LN10 / 10.0f));
        for (int q = 0; q < PERIODS; ++q) {
            const float power = weight * src.energy[q][b];
            out.bands[q][b] = power * (p[q] * ef + (1.0f - p[q]) * eh);
            out.states[0][q] += power * eh;
            out.states[1][q] += power * ef;
        }
    }
    phase_mark(PHASE_MIX);
    out.failed = failed;
}

}  // namespace qm
