// TEST INFRASTRUCTURE — the FP64 oracle's C entry points, for the tests (ctypes) and the bench's
// CPU-baseline leg. Not part of the product path.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "scene.hpp"

using namespace qmo;

extern "C" {

void qmo_alpha_bands(double temperature_c, double relative_humidity_pct, double* out) {
    alpha_bands(temperature_c, relative_humidity_pct, out);
}

double qmo_lden_energy(const double* period_energy) { return lden_energy(period_energy); }

// physics/src/cnossos/mod.rs:123 state_boundaries on an explicit vertical path.
int qmo_state_boundaries(size_t n, const double* d, const double* z, const double* g, double sx, double sz, double rx,
                         double rz, double source_ground_factor, size_t n_terrain, const double* terrain_x,
                         const double* terrain_z, size_t n_tops, const double* tops_x, const double* tops_z,
                         double* attenuation, double* without_ground) {
    std::vector<P2> terrain(n_terrain), tops(n_tops);
    for (size_t i = 0; i < n_terrain; ++i) terrain[i] = P2{terrain_x[i], terrain_z[i]};
    for (size_t i = 0; i < n_tops; ++i) tops[i] = P2{tops_x[i], tops_z[i]};
    VerticalPath path;
    path.profile = VerticalProfile{d, z, g, n};
    path.source = P2{sx, sz};
    path.receiver = P2{rx, rz};
    path.source_ground_factor = source_ground_factor;
    path.terrain_candidates = terrain.data();
    path.terrain_count = terrain.size();
    path.obstacle_tops = tops.data();
    path.tops_count = tops.size();
    VerticalScratch scratch;
    StateBoundary out[2];
    state_boundaries(path, scratch, out, nullptr);
    for (int s = 0; s < 2; ++s)
        for (int b = 0; b < BANDS; ++b) {
            attenuation[s * BANDS + b] = out[s].attenuation_db[b];
            without_ground[s * BANDS + b] = out[s].without_ground_db[b];
        }
    return 0;
}

// physics/src/ray.rs:125 ray_terms on an explicit profile and crossing list.
int qmo_ray_terms(size_t n, double horizontal_m, const double* t, const double* ground_m, const double* ground_factor,
                  const double* forest_cover, const double* canopy_m, size_t n_crossings, const double* crossing_t,
                  const double* crossing_height, const uint8_t* crossing_building, const uint8_t* crossing_roof,
                  const uint64_t* crossing_footprint, double source_height_m, int has_source_ground,
                  double source_ground_m, double receiver_altitude_m, int has_source_ground_factor,
                  double source_ground_factor, double platform_half_width_m, uint64_t own_footprint,
                  const double* favourable, const double* alpha, double* periods, double* states, double* slant,
                  double* boundary_db) {
    Profile profile;
    profile.clear(horizontal_m);
    for (size_t i = 0; i < n; ++i) profile.push(t[i], ground_m[i], ground_factor[i], forest_cover[i], canopy_m[i]);
    std::vector<Crossing> crossings(n_crossings);
    for (size_t i = 0; i < n_crossings; ++i)
        crossings[i] = Crossing{crossing_t[i], crossing_height[i], crossing_building[i] != 0, crossing_roof[i] != 0,
                                crossing_footprint[i]};
    RayEnds ends{source_height_m, has_source_ground != 0, source_ground_m, receiver_altitude_m,
                 has_source_ground_factor != 0, source_ground_factor, platform_half_width_m, own_footprint};
    RayScratch scratch;
    RayTerms terms;
    ray_terms(profile, crossings, ends, favourable, alpha, scratch, terms, nullptr);
    for (int p = 0; p < PERIODS; ++p)
        for (int b = 0; b < BANDS; ++b) periods[p * BANDS + b] = terms.transfer.periods[p][b];
    for (int s = 0; s < 2; ++s)
        for (int b = 0; b < BANDS; ++b) {
            states[s * BANDS + b] = terms.transfer.states[s][b];
            boundary_db[s * BANDS + b] = terms.boundaries[s].attenuation_db[b];
        }
    *slant = terms.transfer.slant_m;
    return 0;
}

// physics/src/ray.rs:255-354 PathBuffers::fill: the merged vertical profile and the tops.
int qmo_vertical_path(size_t n, double horizontal_m, const double* t, const double* ground_m, const double* ground_factor,
                      size_t n_crossings, const double* crossing_t, const double* crossing_height,
                      const uint8_t* crossing_building, const uint8_t* crossing_roof, const uint64_t* crossing_footprint,
                      double source_height_m, double platform_half_width_m, uint64_t own_footprint, int has_source_ground,
                      double source_ground_m, size_t capacity, double* d, double* z, double* g, size_t* n_out,
                      double* tops, size_t* n_tops, double* crests, size_t* n_crests) {
    Profile profile;
    profile.clear(horizontal_m);
    for (size_t i = 0; i < n; ++i) profile.push(t[i], ground_m[i], ground_factor[i], 0.0, 0.0);
    std::vector<Crossing> crossings(n_crossings);
    for (size_t i = 0; i < n_crossings; ++i)
        crossings[i] = Crossing{crossing_t[i], crossing_height[i], crossing_building[i] != 0, crossing_roof[i] != 0,
                                crossing_footprint[i]};
    RayEnds ends{source_height_m, has_source_ground != 0, source_ground_m, 0.0, false, 0.0, platform_half_width_m,
                 own_footprint};
    RayScratch s;
    fill_path(s, profile, crossings, ends);
    *n_out = s.distance_m.size();
    *n_tops = s.tops.size();
    *n_crests = s.crests.size();
    if (s.distance_m.size() > capacity || s.tops.size() > capacity || s.crests.size() > capacity) return -1;
    for (size_t i = 0; i < s.distance_m.size(); ++i) {
        d[i] = s.distance_m[i];
        z[i] = s.altitude_m[i];
        g[i] = s.ground[i];
    }
    for (size_t i = 0; i < s.tops.size(); ++i) {
        tops[2 * i] = s.tops[i].x;
        tops[2 * i + 1] = s.tops[i].z;
    }
    for (size_t i = 0; i < s.crests.size(); ++i) {
        crests[2 * i] = s.crests[i].x;
        crests[2 * i + 1] = s.crests[i].z;
    }
    return 0;
}

int qmo_ground_at(const qm_terrain* terrain, double x, double y, double* out) {
    GroundSample g{};
    const int failed = ground_at(*terrain, x, y, g);
    out[0] = g.height_m;
    out[1] = g.ground_factor;
    out[2] = g.forest_cover;
    out[3] = g.canopy_m;
    return failed;
}

int qmo_fill_profile(const qm_terrain* terrain, double sx, double sy, double rx, double ry, size_t capacity,
                     double* t, double* z, double* g, double* forest, double* canopy, size_t* n, double* horizontal) {
    Profile profile;
    const double s[2] = {sx, sy}, r[2] = {rx, ry};
    const int failed = fill_profile(*terrain, s, r, profile);
    *n = profile.t.size();
    *horizontal = profile.horizontal_m;
    if (failed) return failed;
    if (profile.t.size() > capacity) return -1;
    for (size_t i = 0; i < profile.t.size(); ++i) {
        t[i] = profile.t[i];
        z[i] = profile.ground_m[i];
        g[i] = profile.ground_factor[i];
        forest[i] = profile.forest_cover[i];
        canopy[i] = profile.canopy_m[i];
    }
    return 0;
}

int qmo_crossings(const qm_obstacles* obstacles, double fx, double fy, double tx, double ty, size_t capacity,
                  double* t, double* height, uint8_t* building, uint8_t* roof, uint64_t* footprint, uint32_t* outline,
                  uint32_t* edge, size_t* n, int brute_force) {
    std::vector<RawCrossing> list;
    const double from[2] = {fx, fy}, to[2] = {tx, ty};
    int failed = 0;
    if (brute_force) crossings_brute_force(*obstacles, from, to, list);
    else failed = crossings_raw(*obstacles, from, to, list, nullptr);
    *n = list.size();
    if (failed) return failed;
    if (list.size() > capacity) return -1;
    for (size_t i = 0; i < list.size(); ++i) {
        t[i] = list[i].c.t;
        height[i] = list[i].c.height_m;
        building[i] = list[i].c.building;
        roof[i] = list[i].c.bounds_roof;
        footprint[i] = list[i].c.footprint_id;
        outline[i] = list[i].outline;
        edge[i] = list[i].edge;
    }
    return 0;
}

// Receiver properties of every point: altitude (ground + 4 m), air absorption and favourable
// percentages (paint/src/exact.rs:95-118, physics/src/weather.rs:163-214).
int qmo_receivers(const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                  const qm_point* points, size_t n, double* altitude, double* alpha, double* favourable,
                  uint32_t* failed) {
    Scene S{*terrain, *obstacles, *weather};
    for (size_t i = 0; i < n; ++i) {
        Receiver r;
        failed[i] = (uint32_t)receiver_of(S, points[i], r);
        altitude[i] = r.altitude_m;
        for (int b = 0; b < BANDS; ++b) alpha[i * BANDS + b] = r.alpha[b];
        for (int p = 0; p < PERIODS; ++p)
            for (int s = 0; s < SECTORS; ++s) favourable[(i * PERIODS + p) * SECTORS + s] = r.favourable[p][s];
    }
    return 0;
}

// paint/src/exact.rs:218-258 evaluate_all on the CPU: every pair's energy (mixed[3], then
// states[2][3]), its failure bits, every point's bands [layer][period][band] (weights optional).
int qmo_evaluate_pairs(const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                       const qm_point* points, size_t n_points, const qm_source* sources, size_t n_sources,
                       const uint32_t* pair_point, const uint32_t* pair_source, const float* pair_weight,
                       size_t n_pairs, double* sums, uint32_t* failed, double* point_bands, int threads,
                       uint64_t* work_out) {
    Scene S{*terrain, *obstacles, *weather};
    for (size_t i = 0; i < n_pairs; ++i)
        if (pair_point[i] >= n_points || pair_source[i] >= n_sources) return -1;
    std::vector<Receiver> receivers(n_points);
    std::vector<uint32_t> receiver_failed(n_points);
    for (size_t i = 0; i < n_points; ++i) receiver_failed[i] = (uint32_t)receiver_of(S, points[i], receivers[i]);
    const int workers = std::max(threads, 1);
    std::vector<PairScratch> scratch(workers);
    std::vector<WorkCount> work(workers);
    std::vector<PairResult> results(n_pairs);
    parallel_for(n_pairs, workers, [&](size_t i, int k) {
        const uint32_t p = pair_point[i];
        failed[i] = receiver_failed[p];
        if (!failed[i])
            failed[i] = (uint32_t)evaluate_pair(S, receivers[p], sources[pair_source[i]], scratch[k], results[i], &work[k]);
        double* out = sums + i * QM_PAIR_SUMS;
        if (failed[i]) {
            for (int j = 0; j < QM_PAIR_SUMS; ++j) out[j] = 0.0;
            return;
        }
        for (int q = 0; q < PERIODS; ++q) {
            out[q] = results[i].mixed[q];
            out[3 + q] = results[i].states[0][q];
            out[6 + q] = results[i].states[1][q];
        }
    });
    if (point_bands) {
        std::memset(point_bands, 0, sizeof(double) * n_points * QM_LAYERS * PERIODS * BANDS);
        for (size_t i = 0; i < n_pairs; ++i) {
            if (failed[i]) continue;
            const qm_source& src = sources[pair_source[i]];
            const double w = pair_weight ? (double)pair_weight[i] : 1.0;
            double* dst = point_bands + ((size_t)pair_point[i] * QM_LAYERS + src.layer) * PERIODS * BANDS;
            for (int q = 0; q < PERIODS; ++q)
                for (int b = 0; b < BANDS; ++b) dst[q * BANDS + b] += w * results[i].bands[q][b];
        }
    }
    if (work_out) {
        WorkCount total;
        for (const WorkCount& w : work) {
            total.vertices += w.vertices;
            total.cells += w.cells;
            total.cells_occupied += w.cells_occupied;
            total.edges += w.edges;
            total.crossings += w.crossings;
            total.candidates += w.candidates;
        }
        work_out[0] = total.vertices;
        work_out[1] = total.cells;
        work_out[2] = total.cells_occupied;
        work_out[3] = total.edges;
        work_out[4] = total.crossings;
        work_out[5] = total.candidates;
    }
    return 0;
}

// The boundary terms (A_ground + A_dif per state and band, physics/src/cnossos/diffraction.rs:
// 64-196) of every pair, `boundary` n_pairs x 2 x 8, with each pair's failure bits.
int qmo_pair_boundaries(const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                        const qm_point* points, size_t n_points, const qm_source* sources, size_t n_sources,
                        const uint32_t* pair_point, const uint32_t* pair_source, size_t n_pairs, double* boundary,
                        uint32_t* failed, int threads) {
    Scene S{*terrain, *obstacles, *weather};
    for (size_t i = 0; i < n_pairs; ++i)
        if (pair_point[i] >= n_points || pair_source[i] >= n_sources) return -1;
    std::vector<Receiver> receivers(n_points);
    std::vector<uint32_t> receiver_failed(n_points);
    for (size_t i = 0; i < n_points; ++i) receiver_failed[i] = (uint32_t)receiver_of(S, points[i], receivers[i]);
    const int workers = std::max(threads, 1);
    std::vector<PairScratch> scratch(workers);
    parallel_for(n_pairs, workers, [&](size_t i, int k) {
        const uint32_t p = pair_point[i];
        PairResult r{};
        failed[i] = receiver_failed[p];
        if (!failed[i]) failed[i] = (uint32_t)evaluate_pair(S, receivers[p], sources[pair_source[i]], scratch[k], r, nullptr);
        for (int st = 0; st < 2; ++st)
            for (int b = 0; b < BANDS; ++b) boundary[(i * 2 + st) * BANDS + b] = failed[i] ? 0.0 : r.boundary[st][b];
    });
    return 0;
}

// physics/src/line.rs line_quadrature_nodes for a piece from `start` to `end` (receiver-relative,
// metres), the skyline handing every given arc (lo, hi, nearest; n_arcs x 3) to every wide bucket.
// info: length, subtended in-plane angle, divergence factor. Returns the node count (nodes beyond
// `capacity` not written) or -1 for a degenerate piece.
long qmo_line_nodes(const double* start, const double* end, const double* arcs, size_t n_arcs, size_t capacity,
                    double* along, double* weight, uint8_t* obstacles, double* info) {
    LineGeometry g;
    if (!line_geometry(start, end, g)) return -1;
    std::vector<LineNode> nodes;
    line_quadrature_nodes(
        g,
        [&](double, double, double, const std::function<void(const SkylineArc&)>& visit) {
            for (size_t i = 0; i < n_arcs; ++i) visit(SkylineArc{arcs[3 * i], arcs[3 * i + 1], arcs[3 * i + 2]});
        },
        nodes);
    for (size_t i = 0; i < nodes.size() && i < capacity; ++i) {
        along[i] = nodes[i].along_m;
        weight[i] = nodes[i].weight_rad;
        obstacles[i] = nodes[i].obstacles_on_ray ? 1 : 0;
    }
    info[0] = g.length_m;
    info[1] = line_subtended(g);
    info[2] = line_divergence_factor(g);
    return (long)nodes.size();
}

// The skyline arcs (popup/src/obstacles/skyline.rs) of a receiver at (x, y) over a wedge.
long qmo_skyline_arcs(const qm_obstacles* obstacles, double x, double y, double lo, double hi, double radius,
                      double floor_m, size_t capacity, double* out) {
    const double origin[2] = {x, y};
    size_t n = 0;
    skyline_arcs(*obstacles, origin, lo, hi, radius, floor_m, [&](const SkylineArc& a) {
        if (n < capacity) {
            out[3 * n] = a.lo_rad;
            out[3 * n + 1] = a.hi_rad;
            out[3 * n + 2] = a.nearest_m;
        }
        ++n;
    });
    return (long)n;
}

// Line pairs as qmo_evaluate_pairs: per pair the 9 sums, the failure bits and the rays (nodes).
int qmo_evaluate_line_pairs(const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                            const qm_point* points, size_t n_points, const qm_line_piece* pieces, size_t n_pieces,
                            const uint32_t* pair_point, const uint32_t* pair_piece, size_t n_pairs, double* sums,
                            uint32_t* failed, uint32_t* rays, int threads) {
    Scene S{*terrain, *obstacles, *weather};
    for (size_t i = 0; i < n_pairs; ++i)
        if (pair_point[i] >= n_points || pair_piece[i] >= n_pieces) return -1;
    std::vector<Receiver> receivers(n_points);
    std::vector<uint32_t> receiver_failed(n_points);
    for (size_t i = 0; i < n_points; ++i) receiver_failed[i] = (uint32_t)receiver_of(S, points[i], receivers[i]);
    const int workers = std::max(threads, 1);
    std::vector<PairScratch> scratch(workers);
    std::vector<std::vector<LineNode>> nodes(workers);
    parallel_for(n_pairs, workers, [&](size_t i, int k) {
        const uint32_t p = pair_point[i];
        PairResult r{};
        failed[i] = receiver_failed[p];
        if (!failed[i])
            failed[i] = (uint32_t)evaluate_line_pair(S, receivers[p], pieces[pair_piece[i]], scratch[k], r, nodes[k],
                                                     nullptr);
        rays[i] = failed[i] ? 0u : (uint32_t)nodes[k].size();
        double* out = sums + i * QM_PAIR_SUMS;
        for (int q = 0; q < PERIODS; ++q) {
            out[q] = failed[i] ? 0.0 : r.mixed[q];
            out[3 + q] = failed[i] ? 0.0 : r.states[0][q];
            out[6 + q] = failed[i] ? 0.0 : r.states[1][q];
        }
    });
    return 0;
}

// The painted grid: per receiver the summed mixed energy per period of every source within
// `cutoff_m` (horizontal), the paths counted and failures; Lden follows from lden_energy.
int qmo_paint_points(const qm_terrain* terrain, const qm_obstacles* obstacles, const qm_weather* weather,
                     const qm_point* points, size_t n_points, const qm_source* sources, size_t n_sources,
                     double cutoff_m, double* period_energy, uint64_t* paths, uint32_t* failures, int threads) {
    Scene S{*terrain, *obstacles, *weather};
    const int workers = std::max(threads, 1);
    std::vector<PairScratch> scratch(workers);
    // Sources bucketed by the cutoff (every source when the cutoff is unbounded).
    const bool bucketed = std::isfinite(cutoff_m) && cutoff_m > 0.0 && n_sources > 0;
    double min_x = 0.0, min_y = 0.0;
    int64_t bx = 1, by = 1;
    std::vector<uint32_t> first(2, 0), order(n_sources);
    if (bucketed) {
        double max_x = sources[0].x_m, max_y = sources[0].y_m;
        min_x = sources[0].x_m;
        min_y = sources[0].y_m;
        for (size_t j = 0; j < n_sources; ++j) {
            min_x = std::min(min_x, sources[j].x_m);
            min_y = std::min(min_y, sources[j].y_m);
            max_x = std::max(max_x, sources[j].x_m);
            max_y = std::max(max_y, sources[j].y_m);
        }
        bx = (int64_t)std::floor((max_x - min_x) / cutoff_m) + 1;
        by = (int64_t)std::floor((max_y - min_y) / cutoff_m) + 1;
        first.assign((size_t)(bx * by + 1), 0);
        std::vector<uint32_t> bucket(n_sources);
        for (size_t j = 0; j < n_sources; ++j) {
            const int64_t kx = std::min((int64_t)std::floor((sources[j].x_m - min_x) / cutoff_m), bx - 1);
            const int64_t ky = std::min((int64_t)std::floor((sources[j].y_m - min_y) / cutoff_m), by - 1);
            bucket[j] = (uint32_t)(ky * bx + kx);
            ++first[bucket[j] + 1];
        }
        for (size_t k = 0; k + 1 < first.size(); ++k) first[k + 1] += first[k];
        std::vector<uint32_t> fill(first.begin(), first.end() - 1);
        for (size_t j = 0; j < n_sources; ++j) order[fill[bucket[j]]++] = (uint32_t)j;
    } else {
        for (size_t j = 0; j < n_sources; ++j) order[j] = (uint32_t)j;
        first[1] = (uint32_t)n_sources;
    }
    parallel_for(n_points, workers, [&](size_t i, int k) {
        Receiver r;
        double e[PERIODS] = {0.0, 0.0, 0.0};
        uint64_t count = 0;
        uint32_t failed_count = 0;
        if (receiver_of(S, points[i], r) == 0) {
            PairResult result;
            int64_t kx0 = 0, kx1 = 0, ky0 = 0, ky1 = 0;
            if (bucketed) {
                const int64_t kx = (int64_t)std::floor((r.x - min_x) / cutoff_m);
                const int64_t ky = (int64_t)std::floor((r.y - min_y) / cutoff_m);
                kx0 = std::max<int64_t>(kx - 1, 0);
                kx1 = std::min<int64_t>(kx + 1, bx - 1);
                ky0 = std::max<int64_t>(ky - 1, 0);
                ky1 = std::min<int64_t>(ky + 1, by - 1);
            }
            // Sources in ascending index order, as the unbucketed loop takes them.
            std::vector<uint32_t> near;
            for (int64_t y = ky0; y <= ky1; ++y)
                for (int64_t x = kx0; x <= kx1; ++x)
                    if (x >= 0 && y >= 0 && x < bx && y < by)
                        for (uint32_t s = first[y * bx + x]; s < first[y * bx + x + 1]; ++s) near.push_back(order[s]);
            std::sort(near.begin(), near.end());
            for (uint32_t j : near) {
                const double h = std::hypot(sources[j].x_m - r.x, sources[j].y_m - r.y);
                if (!(h <= cutoff_m)) continue;
                ++count;
                if (evaluate_pair(S, r, sources[j], scratch[k], result, nullptr)) {
                    ++failed_count;
                    continue;
                }
                for (int q = 0; q < PERIODS; ++q) e[q] += result.mixed[q];
            }
        } else {
            failed_count = 1;
        }
        for (int q = 0; q < PERIODS; ++q) period_energy[i * PERIODS + q] = e[q];
        paths[i] = count;
        failures[i] = failed_count;
    });
    return 0;
}

}  // extern "C"
