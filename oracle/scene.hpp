// TEST INFRASTRUCTURE — FP64 oracle: scene-side declarations (the ground under a ray, obstacle
// crossings, weather, one pair end to end). Not part of the product path.
#pragma once

#include <functional>

#include "oracle.hpp"
#include "../quietmap_amd/csrc/qm_scene.h"

namespace qmo {

struct GroundSample {
    double height_m, ground_factor, forest_cover, canopy_m;
};

struct RawCrossing {
    Crossing c;
    uint32_t outline;
    uint32_t edge;  // index of the edge's first vertex
};

struct Scene {
    qm_terrain terrain;
    qm_obstacles obstacles;
    qm_weather weather;
};

struct Receiver {
    double x, y;
    uint64_t own_footprint;
    double altitude_m;
    double favourable[PERIODS][SECTORS];
    double alpha[BANDS];
};

struct PairResult {
    double bands[PERIODS][BANDS];
    double mixed[PERIODS];
    double states[2][PERIODS];
    double boundary[2][BANDS];  // A_ground + A_dif per state and band
};

struct PairScratch {
    Profile profile;
    std::vector<RawCrossing> raw;
    std::vector<Crossing> crossings;
    RayScratch ray;
    RayTerms terms;
};

// The line-source quadrature (oracle/line.cpp; physics/src/line.rs).
struct LineGeometry {
    double start[3], unit[3];
    double length_m, foot_along_m, perpendicular_m, start_angle_rad, end_angle_rad;
};
struct LineNode {
    double along_m, weight_rad;
    bool obstacles_on_ray;
};
struct SkylineArc {
    double lo_rad, hi_rad, nearest_m;
};
using SkylineQuery =
    std::function<void(double lo, double hi, double radius, const std::function<void(const SkylineArc&)>& visit)>;
double wrap_to_pi(double angle);
bool line_geometry(const double start[3], const double end[3], LineGeometry& g);
double line_subtended(const LineGeometry& g);
double line_divergence_factor(const LineGeometry& g);
void line_quadrature_nodes(const LineGeometry& g, const SkylineQuery& skyline, std::vector<LineNode>& nodes);
void skyline_arcs(const qm_obstacles& O, const double origin[2], double lo_rad, double hi_rad, double radius_m,
                  double los_floor_m, const std::function<void(const SkylineArc&)>& visit);

int ground_at(const qm_terrain& T, double x, double y, GroundSample& out);
int fill_profile(const qm_terrain& T, const double source[2], const double receiver[2], Profile& profile);
void steps_of(const qm_obstacles& O, double x, double y, double out[2]);
void sort_and_deduplicate(std::vector<RawCrossing>& list);
int crossings_raw(const qm_obstacles& O, const double from[2], const double to[2], std::vector<RawCrossing>& out,
                  WorkCount* work);
void crossings_brute_force(const qm_obstacles& O, const double from[2], const double to[2],
                           std::vector<RawCrossing>& out);
void place_weather(const qm_weather& W, double x, double y, double favourable[PERIODS][SECTORS], double alpha[BANDS]);
double favourable_at(const double favourable[PERIODS][SECTORS], int period, double azimuth_rad);
int receiver_of(const Scene& S, const qm_point& point, Receiver& out);
int evaluate_pair(const Scene& S, const Receiver& receiver, const qm_source& source, PairScratch& s, PairResult& out,
                  WorkCount* work);
void parallel_for(size_t n, int threads, const std::function<void(size_t, int)>& body);
int evaluate_line_pair(const Scene& S, const Receiver& receiver, const qm_line_piece& piece, PairScratch& s,
                       PairResult& out, std::vector<LineNode>& nodes, WorkCount* work);

}  // namespace qmo
