// TEST INFRASTRUCTURE — FP64 oracle: a receiver's terrain and building horizons for aircraft
// screening, restated from popup/src/aircraft/horizons.rs:70-210 (Horizons::build, SectorEdges,
// write_f32) over the oracle's ground (Ground::read_height_m / Ground::at, fine detail included)
// and crossings (Scene::crossings). Not part of the product path.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <vector>

#include "../quietmap_amd/csrc/qm_aircraft.h"
#include "oracle.hpp"
#include "scene.hpp"

namespace qmo {
namespace {

constexpr int TERRAIN_SAMPLES = 48;
constexpr double TERRAIN_FIRST_SAMPLE_M = 30.0;
const double TERRAIN_BAND_LIMITS_M[QM_HORIZON_BANDS] = {500.0, 1000.0, 2000.0, 3500.0, 5500.0, 8000.0};
const double BUILDING_BAND_LIMITS_M[QM_HORIZON_BANDS] = {16.0, 32.0, 64.0, 128.0, 256.0, 512.0};
constexpr double BUILDING_NEAREST_EDGE_M = 0.01;

struct Edge {
    bool some = false;
    double range_m = 0.0, tangent = 0.0;
};

// horizons.rs:43-80 SectorEdges.
struct SectorEdges {
    std::vector<std::array<Edge, QM_HORIZON_BANDS>> sectors;
    std::vector<double> steepest;
    double steepest_of_all = -std::numeric_limits<double>::infinity();
    explicit SectorEdges(int n) : sectors(n), steepest(n, -std::numeric_limits<double>::infinity()) {}
    void offer(int sector, int band, double range_m, double tangent) {
        Edge& slot = sectors[sector][band];
        if (!slot.some || tangent > slot.tangent) slot = Edge{true, range_m, tangent};
        steepest[sector] = std::max(steepest[sector], tangent);
        steepest_of_all = std::max(steepest_of_all, tangent);
    }
    void write_f32(float* out) const {
        *out++ = (float)steepest_of_all;
        for (size_t s = 0; s < sectors.size(); ++s) {
            *out++ = (float)steepest[s];
            for (const Edge& e : sectors[s]) {
                *out++ = e.some ? (float)e.range_m : std::numeric_limits<float>::quiet_NaN();
                *out++ = e.some ? (float)e.tangent : std::numeric_limits<float>::quiet_NaN();
            }
        }
    }
};

void sector_direction(int sector, int sectors, double& east, double& north) {
    const double a = ((double)sector + 0.5) * (2.0 * M_PI) / (double)sectors;
    east = std::cos(a);
    north = std::sin(a);
}

int band_of(double range_m, const double* limits) {
    for (int k = 0; k < QM_HORIZON_BANDS; ++k)
        if (range_m <= limits[k] + 0.5) return k;
    return QM_HORIZON_BANDS - 1;
}

}  // namespace
}  // namespace qmo

using namespace qmo;

// Horizons::build at (x, y) and `altitude_m`, the walls of `own_footprint` no horizon, into
// QM_HORIZON_FLOATS floats (write_f32); returns a failure bit where a crossing search fails.
extern "C" int qmo_horizons(const qm_terrain* terrain, const qm_obstacles* obstacles, double x, double y,
                            double altitude_m, uint64_t own_footprint, float* out) {
    SectorEdges ter(QM_TERRAIN_SECTORS);
    const double growth = std::pow(8000.0 / TERRAIN_FIRST_SAMPLE_M, 1.0 / (double)(TERRAIN_SAMPLES - 1));
    for (int s = 0; s < QM_TERRAIN_SECTORS; ++s) {
        double east, north;
        sector_direction(s, QM_TERRAIN_SECTORS, east, north);
        double range_m = TERRAIN_FIRST_SAMPLE_M;
        for (int k = 0; k < TERRAIN_SAMPLES; ++k) {
            GroundSample g;
            if (ground_at(*terrain, x + east * range_m, y + north * range_m, g) == 0)
                ter.offer(s, band_of(range_m, TERRAIN_BAND_LIMITS_M), range_m, (g.height_m - altitude_m) / range_m);
            range_m *= growth;
        }
    }
    SectorEdges bld(QM_BUILDING_SECTORS);
    const double reach_m = BUILDING_BAND_LIMITS_M[QM_HORIZON_BANDS - 1];
    int failed = 0;
    std::vector<RawCrossing> list;
    for (int s = 0; s < QM_BUILDING_SECTORS; ++s) {
        double east, north;
        sector_direction(s, QM_BUILDING_SECTORS, east, north);
        const double from[2] = {x, y}, to[2] = {x + east * reach_m, y + north * reach_m};
        list.clear();
        failed |= crossings_raw(*obstacles, from, to, list, nullptr);
        for (const RawCrossing& rc : list) {
            const Crossing& c = rc.c;
            if (!c.building || c.footprint_id == own_footprint) continue;
            const double range_m = c.t * reach_m;
            if (range_m <= BUILDING_NEAREST_EDGE_M) continue;
            GroundSample g;
            if (ground_at(*terrain, x + east * range_m, y + north * range_m, g)) {
                failed |= QM_FAILED_NO_TERRAIN;
                continue;
            }
            const double roof_m = g.height_m + c.height_m;
            bld.offer(s, band_of(range_m, BUILDING_BAND_LIMITS_M), range_m, (roof_m - altitude_m) / range_m);
        }
    }
    ter.write_f32(out);
    bld.write_f32(out + QM_TERRAIN_HORIZON_FLOATS);
    return failed;
}
