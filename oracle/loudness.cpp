// TEST INFRASTRUCTURE — FP64 oracle: ISO 532-1:2017 stationary loudness (Zwicker, frontal free
// field) of 28 third-octave levels, restated from physics/src/loudness.rs:1-182. Not part of the
// product path.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <limits>

#include "oracle.hpp"

namespace qmo {
namespace {

// The method's tables (ISO 532-1:2017 / DIN 45631; physics/src/loudness.rs:26-96).
const double RAP[8] = {45.0, 55.0, 65.0, 71.0, 80.0, 90.0, 100.0, 120.0};
const double DLL[8][11] = {
    {-32.0, -24.0, -16.0, -10.0, -5.0, 0.0, -7.0, -3.0, 0.0, -2.0, 0.0},
    {-29.0, -22.0, -15.0, -10.0, -4.0, 0.0, -7.0, -2.0, 0.0, -2.0, 0.0},
    {-27.0, -19.0, -14.0, -9.0, -4.0, 0.0, -6.0, -2.0, 0.0, -2.0, 0.0},
    {-25.0, -17.0, -12.0, -9.0, -3.0, 0.0, -5.0, -2.0, 0.0, -2.0, 0.0},
    {-23.0, -16.0, -11.0, -7.0, -3.0, 0.0, -4.0, -1.0, 0.0, -1.0, 0.0},
    {-20.0, -14.0, -10.0, -6.0, -3.0, 0.0, -4.0, -1.0, 0.0, -1.0, 0.0},
    {-18.0, -12.0, -9.0, -6.0, -2.0, 0.0, -3.0, -1.0, 0.0, -1.0, 0.0},
    {-15.0, -10.0, -8.0, -4.0, -2.0, 0.0, -3.0, -1.0, 0.0, -1.0, 0.0},
};
const double LTQ[20] = {30.0, 18.0, 12.0, 8.0, 7.0, 6.0, 5.0, 4.0, 3.0, 3.0,
                        3.0,  3.0,  3.0,  3.0, 3.0, 3.0, 3.0, 3.0, 3.0, 3.0};
const double A0[20] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -0.5, -1.6, -3.2, -5.4, -5.6, -4.0, -1.5, 2.0, 5.0, 12.0};
const double DCB[20] = {-0.25, -0.6, -0.8, -0.8, -0.5, 0.0, 0.5, 1.1, 1.5, 1.7,
                        1.8,   1.8,  1.7,  1.6,  1.4,  1.2, 0.8, 0.5, 0.0, -0.5};
const double ZUP[21] = {0.9,  1.8,  2.8,  3.5,  4.4,  5.4,  6.6,  7.9,  9.2,  10.6, 12.3,
                        13.8, 15.2, 16.7, 18.1, 19.3, 20.6, 21.8, 22.7, 23.6, 24.0};
const double RNS[18] = {21.5, 18.0, 15.1, 11.5, 9.0, 6.1, 4.4, 3.1, 2.13, 1.36, 0.82, 0.42, 0.30, 0.22, 0.15, 0.10, 0.035, 0.0};
const double USL[18][8] = {
    {13.0, 8.2, 6.3, 5.5, 5.5, 5.5, 5.5, 5.5},      {9.0, 7.5, 6.0, 5.1, 4.5, 4.5, 4.5, 4.5},
    {7.8, 6.7, 5.6, 4.9, 4.4, 3.9, 3.9, 3.9},       {6.2, 5.4, 4.6, 4.0, 3.5, 3.2, 3.2, 3.2},
    {4.5, 3.8, 3.6, 3.2, 2.9, 2.7, 2.7, 2.7},       {3.7, 3.0, 2.8, 2.35, 2.2, 2.2, 2.2, 2.2},
    {2.9, 2.3, 2.1, 1.9, 1.8, 1.7, 1.7, 1.7},       {2.4, 1.7, 1.5, 1.35, 1.3, 1.3, 1.3, 1.3},
    {1.95, 1.45, 1.3, 1.15, 1.1, 1.1, 1.1, 1.1},    {1.5, 1.2, 0.94, 0.86, 0.82, 0.82, 0.82, 0.82},
    {0.72, 0.67, 0.64, 0.63, 0.62, 0.62, 0.62, 0.62}, {0.59, 0.53, 0.51, 0.50, 0.42, 0.42, 0.42, 0.42},
    {0.40, 0.33, 0.26, 0.24, 0.24, 0.22, 0.22, 0.22}, {0.27, 0.21, 0.20, 0.18, 0.17, 0.17, 0.17, 0.17},
    {0.16, 0.15, 0.14, 0.12, 0.11, 0.11, 0.11, 0.11}, {0.12, 0.11, 0.10, 0.08, 0.08, 0.08, 0.08, 0.08},
    {0.09, 0.08, 0.07, 0.06, 0.06, 0.06, 0.06, 0.05}, {0.06, 0.05, 0.03, 0.02, 0.02, 0.02, 0.02, 0.02},
};

// physics/src/loudness.rs:99-144 core_loudness
void core_loudness(const double levels[28], double core[21]) {
    double corrected[11];
    for (int band = 0; band < 11; ++band) {
        const double level = levels[band];
        int range = 7;
        for (int j = 0; j < 8; ++j)
            if (level <= RAP[j] - DLL[j][band]) {
                range = j;
                break;
            }
        corrected[band] = energy(level + DLL[range][band]);
    }
    double lowest[3];
    const int spans[3][2] = {{0, 6}, {6, 9}, {9, 11}};
    for (int k = 0; k < 3; ++k) {
        double sum = 0.0;
        for (int b = spans[k][0]; b < spans[k][1]; ++b) sum += corrected[b];
        lowest[k] = sum > 0.0 ? 10.0 * std::log10(sum) : -std::numeric_limits<double>::infinity();
    }
    for (int band = 0; band < 21; ++band) core[band] = 0.0;
    for (int band = 0; band < 20; ++band) {
        double level = (band < 3 ? lowest[band] : levels[band + 8]) - A0[band];
        if (level > LTQ[band]) {
            level -= DCB[band];
            const double s = 0.25;
            const double threshold = energy(LTQ[band] / 4.0);
            core[band] =
                std::max(0.0635 * threshold * (std::sqrt(std::sqrt(1.0 - s + s * energy(level - LTQ[band]))) - 1.0), 0.0);
        }
    }
    const double correction = 0.4 + 0.32 * std::pow(core[0], 0.2);
    if (correction <= 1.0) core[0] *= correction;
}

}  // namespace

// physics/src/loudness.rs:148-182 zwicker_loudness_sone
double zwicker_loudness_sone(const double in[28]) {
    double levels[28];
    for (int i = 0; i < 28; ++i) levels[i] = std::max(in[i], -60.0);
    double core[21];
    core_loudness(levels, core);
    double total = 0.0, z1 = 0.0, n1 = 0.0;
    for (int band = 0; band < 21; ++band) {
        const double upper = ZUP[band] + 0.0001;
        const int group = std::min(band > 0 ? band - 1 : 0, 7);
        const double c = core[band];
        while (z1 < upper) {
            if (n1 <= c) {
                total += c * (upper - z1);
                z1 = upper;
                n1 = c;
            } else {
                int range = 17;
                for (int j = 0; j < 18; ++j)
                    if (n1 > RNS[j]) {
                        range = j;
                        break;
                    }
                const double steepness = USL[range][group];
                double n2 = std::max(RNS[range], c);
                double z2 = z1 + (n1 - n2) / steepness;
                if (z2 > upper) {
                    z2 = upper;
                    n2 = n1 - (z2 - z1) * steepness;
                }
                total += (z2 - z1) * (n1 + n2) / 2.0;
                z1 = z2;
                n1 = n2;
            }
        }
    }
    return std::max(total, 0.0);
}

}  // namespace qmo

extern "C" void qmo_zwicker(const double* levels, size_t n, double* sone) {
    for (size_t i = 0; i < n; ++i) sone[i] = qmo::zwicker_loudness_sone(levels + 28 * i);
}
