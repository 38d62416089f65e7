// TEST INFRASTRUCTURE — the FP64 CPU oracle of the noise painter's hot path.
//
// A C++ restatement, in double precision, of the reference's Rust CPU path for one
// receiver-source pair: physics/src/{bands,atmosphere,foliage,ray}.rs,
// physics/src/cnossos/{mod,mean_plane,ground,rubber_band,diffraction}.rs,
// popup/src/evaluate.rs, popup/src/obstacles/crossings.rs and paint/src/exact.rs, each function
// citing the file:line it follows. Only the tests, the smoke check and the bench's CPU-baseline
// leg may load it, and only as the checker: the product path (quietmap_amd) never calls it.
//
// The scene the oracle walks is the planar synthetic scene of quietmap_amd/csrc/qm_scene.h
// (a regular terrain lattice in metres and an obstacle cell grid in integer steps), with the
// reference's semantics: a profile vertex wherever the ray crosses a lattice line
// (popup/src/scene.rs:242-420) and crossings by a cell walk with the exact segment test
// (popup/src/obstacles/crossings.rs:163-230).
//
// Parity pins (tests/test_oracle_*.py): ISO/TR 17534-4 TC01-TC28 (state_boundaries, ground,
// diffraction, rubber band, mean planes), the reference's ray_tests constants (ray_terms), the
// mean-plane closed forms, Lden(60,55,50) = 60, ISO 9613-1 alpha tables, crossing semantics
// against brute force over all edges, and MoSQITo ISO 532-1 loudness values.
#pragma once

#include <cstdint>
#include <cstddef>
#include <vector>

namespace qmo {

// physics/src/bands.rs:3-21
constexpr int BANDS = 8;
constexpr int PERIODS = 3;
constexpr int SECTORS = 16;
extern const double BAND_FREQUENCY_HZ[BANDS];
constexpr double SPEED_OF_SOUND_M_PER_S = 340.0;
extern const double A_WEIGHTING_DB[BANDS];
extern const double PERIOD_HOURS[PERIODS];
extern const double PERIOD_PENALTY_DB[PERIODS];

double energy(double level_db);          // bands.rs:36-38
double amplitude(double level_db);       // bands.rs:41-43
double lden_energy(const double period_energy[PERIODS]);  // bands.rs:59-66
double level_db(double energy);          // bands.rs:47-56 (returns -inf for 0)

// physics/src/atmosphere.rs:17-56
double exact_mid_band_frequency_hz(int band);
double alpha_db_per_km(double frequency_hz, double temperature_c, double relative_humidity_pct, double pressure_kpa);
void alpha_bands(double temperature_c, double relative_humidity_pct, double out[BANDS]);

struct P2 {
    double x;  // horizontal distance from the source (m)
    double z;  // altitude (m)
};
inline bool operator==(const P2& a, const P2& b) { return a.x == b.x && a.z == b.z; }

// physics/src/cnossos/mod.rs:23-27
struct VerticalProfile {
    const double* d;
    const double* z;
    const double* g;
    size_t n;
    double ground_altitude_at(double x) const;      // mod.rs:30-33
    double ground_altitude_before(double x) const;  // mod.rs:35-38
    double mean_ground_factor(double start_m, double end_m) const;  // mod.rs:40-58
};

// physics/src/cnossos/mean_plane.rs:7-12
struct MeanPlane {
    double slope;
    double intercept_m;
    double altitude_at(double x) const { return slope * x + intercept_m; }
    double signed_height(double x, double z) const;
    double projected_distance(P2 from, P2 to) const;
    P2 mirror(P2 p) const;
};
MeanPlane fit_mean_plane(const VerticalProfile& profile, double start_m, double end_m);

// mean_plane.rs:99-121
struct EquivalentGeometry {
    MeanPlane plane;
    double source_side_height_m;
    double receiver_side_height_m;
    bool source_side_below_plane;
    bool receiver_side_below_plane;
    double projected_distance_m;
    static EquivalentGeometry over(MeanPlane plane, P2 from, P2 to);
};

enum State { HOMOGENEOUS = 0, FAVOURABLE = 1 };

// ground.rs:31-36
struct GroundFactors {
    double path;
    double impedance;
    double floor;
};
double ground_factor_near_source(const EquivalentGeometry& g, double path_ground_factor, double source_ground_factor);
void ground_attenuation_bands(const EquivalentGeometry& g, GroundFactors factors, State state, double out[BANDS]);

// rubber_band.rs:40-46
struct StateRay {
    State state;
    double radius_m;
    static StateRay between(State state, P2 source, P2 receiver);
    double length(P2 from, P2 to) const;
    double arc_excess(double chord) const;
    double path_difference(P2 from, const P2* points, size_t n, P2 to) const;
    double sag_at(P2 source, P2 receiver, double x_from_source) const;
};
double distance(P2 a, P2 b);
double triangle_excess(P2 s, P2 o, P2 r);

struct DiffractionPath {
    std::vector<P2> points;
    bool blocked = false;
};

struct StateBoundary {
    double attenuation_db[BANDS];
    double without_ground_db[BANDS];
};

// Work counters of one ray (for the roofline's algorithmic FLOP count).
struct WorkCount {
    uint64_t vertices = 0;     // V: lattice vertices of the profile (ends included)
    uint64_t cells = 0;        // C: obstacle cells visited
    uint64_t cells_occupied = 0;
    uint64_t edges = 0;        // E: edges box-tested
    uint64_t crossings = 0;    // X: crossings kept
    uint64_t candidates = 0;   // K: candidates strictly between S and R
};

// mod.rs:60-76
struct VerticalPath {
    VerticalProfile profile;
    P2 source;
    P2 receiver;
    double source_ground_factor;
    const P2* terrain_candidates;
    size_t terrain_count;
    const P2* obstacle_tops;
    size_t tops_count;
};

struct VerticalScratch {
    std::vector<P2> candidates;
    std::vector<P2> blocking;
    std::vector<size_t> blocking_index;
    std::vector<P2> hull;
    std::vector<size_t> hull_index;
    DiffractionPath path;
    DiffractionPath calm_path;
};

void state_boundaries(const VerticalPath& path, VerticalScratch& scratch, StateBoundary out[2], WorkCount* work);

// physics/src/foliage.rs
struct Profile {
    double horizontal_m = 0.0;
    std::vector<double> t, ground_m, ground_factor, forest_cover, canopy_m;
    void clear(double horizontal) {
        horizontal_m = horizontal;
        t.clear(); ground_m.clear(); ground_factor.clear(); forest_cover.clear(); canopy_m.clear();
    }
    void push(double tt, double z, double g, double forest, double canopy) {
        t.push_back(tt); ground_m.push_back(z); ground_factor.push_back(g);
        forest_cover.push_back(forest); canopy_m.push_back(canopy);
    }
};
void foliage_attenuation_db(double depth_m, double out[BANDS]);
double foliage_depth_m(const Profile& profile, const StateRay& ray, P2 source, P2 receiver);

// physics/src/ray.rs:29-43
struct Crossing {
    double t;
    double height_m;
    bool building;
    bool bounds_roof;
    uint64_t footprint_id;
};

// ray.rs:47-61; `has_*` stand in for Option
struct RayEnds {
    double source_height_m;
    bool has_source_ground;
    double source_ground_m;
    double receiver_altitude_m;
    bool has_source_ground_factor;
    double source_ground_factor;
    double platform_half_width_m;
    uint64_t own_footprint;
};

struct Transfer {
    double slant_m;
    double periods[PERIODS][BANDS];
    double states[2][BANDS];
};

struct RayTerms {
    Transfer transfer;
    StateBoundary boundaries[2];
    double foliage_db[2][BANDS];
    double air_db[BANDS];
};

struct RayScratch {
    std::vector<double> distance_m, altitude_m, ground;
    std::vector<P2> terrain, crests, tops;
    struct Roof { double x0, x1, top0, top1; };
    std::vector<Roof> roofs;
    struct Building { uint64_t id; double x; double top; };
    std::vector<Building> buildings;
    std::vector<uint64_t> containing;
    VerticalScratch vertical;
};

void fill_path(RayScratch& s, const Profile& profile, const std::vector<Crossing>& crossings, const RayEnds& ends);
void containing_footprints(const std::vector<Crossing>& crossings, uint64_t own_footprint, std::vector<uint64_t>& out);
void ray_terms(const Profile& profile, const std::vector<Crossing>& crossings, const RayEnds& ends,
               const double favourable_probability[PERIODS], const double alpha_db_per_km[BANDS],
               RayScratch& scratch, RayTerms& out, WorkCount* work);

}  // namespace qmo
