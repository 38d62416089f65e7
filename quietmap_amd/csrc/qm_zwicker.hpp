// ISO 532-1:2017 stationary loudness (Zwicker, frontal free field) of 28 third-octave levels,
// 25 Hz - 12.5 kHz, in FP32 on the device: physics/src/loudness.rs:99-182 (core loudness, the
// specific loudness pattern and its area), the reference card's loud_zwicker
// (paint-gpu/kernels/loudness.cuh:85-153). The method's tables are the standard's.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace {

constexpr float LTQ[20] = {30.0f, 18.0f, 12.0f, 8.0f, 7.0f, 6.0f, 5.0f, 4.0f, 3.0f, 3.0f,
                           3.0f,  3.0f,  3.0f,  3.0f, 3.0f, 3.0f, 3.0f, 3.0f, 3.0f, 3.0f};
constexpr float A0[20] = {0.0f,  0.0f,  0.0f,  0.0f,  0.0f,  0.0f,  0.0f, 0.0f, 0.0f, 0.0f,
                          -0.5f, -1.6f, -3.2f, -5.4f, -5.6f, -4.0f, -1.5f, 2.0f, 5.0f, 12.0f};
constexpr float DCB[20] = {-0.25f, -0.6f, -0.8f, -0.8f, -0.5f, 0.0f, 0.5f, 1.1f, 1.5f, 1.7f,
                           1.8f,   1.8f,  1.7f,  1.6f,  1.4f,  1.2f, 0.8f, 0.5f, 0.0f, -0.5f};
constexpr float ZUP[21] = {0.9f,  1.8f,  2.8f,  3.5f,  4.4f,  5.4f,  6.6f,  7.9f,  9.2f,  10.6f, 12.3f,
                           13.8f, 15.2f, 16.7f, 18.1f, 19.3f, 20.6f, 21.8f, 22.7f, 23.6f, 24.0f};
__constant__ float RNS[18] = {21.5f, 18.0f, 15.1f, 11.5f, 9.0f, 6.1f, 4.4f, 3.1f, 2.13f,
                              1.36f, 0.82f, 0.42f, 0.30f, 0.22f, 0.15f, 0.10f, 0.035f, 0.0f};
__constant__ float USL[18][8] = {
    {13.0f, 8.2f, 6.3f, 5.5f, 5.5f, 5.5f, 5.5f, 5.5f},        {9.0f, 7.5f, 6.0f, 5.1f, 4.5f, 4.5f, 4.5f, 4.5f},
    {7.8f, 6.7f, 5.6f, 4.9f, 4.4f, 3.9f, 3.9f, 3.9f},         {6.2f, 5.4f, 4.6f, 4.0f, 3.5f, 3.2f, 3.2f, 3.2f},
    {4.5f, 3.8f, 3.6f, 3.2f, 2.9f, 2.7f, 2.7f, 2.7f},         {3.7f, 3.0f, 2.8f, 2.35f, 2.2f, 2.2f, 2.2f, 2.2f},
    {2.9f, 2.3f, 2.1f, 1.9f, 1.8f, 1.7f, 1.7f, 1.7f},         {2.4f, 1.7f, 1.5f, 1.35f, 1.3f, 1.3f, 1.3f, 1.3f},
    {1.95f, 1.45f, 1.3f, 1.15f, 1.1f, 1.1f, 1.1f, 1.1f},      {1.5f, 1.2f, 0.94f, 0.86f, 0.82f, 0.82f, 0.82f, 0.82f},
    {0.72f, 0.67f, 0.64f, 0.63f, 0.62f, 0.62f, 0.62f, 0.62f}, {0.59f, 0.53f, 0.51f, 0.50f, 0.42f, 0.42f, 0.42f, 0.42f},
    {0.40f, 0.33f, 0.26f, 0.24f, 0.24f, 0.22f, 0.22f, 0.22f}, {0.27f, 0.21f, 0.20f, 0.18f, 0.17f, 0.17f, 0.17f, 0.17f},
    {0.16f, 0.15f, 0.14f, 0.12f, 0.11f, 0.11f, 0.11f, 0.11f}, {0.12f, 0.11f, 0.10f, 0.08f, 0.08f, 0.08f, 0.08f, 0.08f},
    {0.09f, 0.08f, 0.07f, 0.06f, 0.06f, 0.06f, 0.06f, 0.05f}, {0.06f, 0.05f, 0.03f, 0.02f, 0.02f, 0.02f, 0.02f, 0.02f},
};

// The lowest bands' range thresholds RAP[j] - DLL[j][band] (ISO 532-1 table A.2), increasing in j
// for every band, so a level's range is the count of thresholds it exceeds.
constexpr float range_threshold(int j, int band) {
    constexpr float rap[8] = {45.0f, 55.0f, 65.0f, 71.0f, 80.0f, 90.0f, 100.0f, 120.0f};
    constexpr float dll[8][11] = {
        {-32.0f, -24.0f, -16.0f, -10.0f, -5.0f, 0.0f, -7.0f, -3.0f, 0.0f, -2.0f, 0.0f},
        {-29.0f, -22.0f, -15.0f, -10.0f, -4.0f, 0.0f, -7.0f, -2.0f, 0.0f, -2.0f, 0.0f},
        {-27.0f, -19.0f, -14.0f, -9.0f, -4.0f, 0.0f, -6.0f, -2.0f, 0.0f, -2.0f, 0.0f},
        {-25.0f, -17.0f, -12.0f, -9.0f, -3.0f, 0.0f, -5.0f, -2.0f, 0.0f, -2.0f, 0.0f},
        {-23.0f, -16.0f, -11.0f, -7.0f, -3.0f, 0.0f, -4.0f, -1.0f, 0.0f, -1.0f, 0.0f},
        {-20.0f, -14.0f, -10.0f, -6.0f, -3.0f, 0.0f, -4.0f, -1.0f, 0.0f, -1.0f, 0.0f},
        {-18.0f, -12.0f, -9.0f, -6.0f, -2.0f, 0.0f, -3.0f, -1.0f, 0.0f, -1.0f, 0.0f},
        {-15.0f, -10.0f, -8.0f, -4.0f, -2.0f, 0.0f, -3.0f, -1.0f, 0.0f, -1.0f, 0.0f},
    };
    return rap[j] - dll[j][band];
}
constexpr float range_correction(int j, int band) {
    constexpr float rap[8] = {45.0f, 55.0f, 65.0f, 71.0f, 80.0f, 90.0f, 100.0f, 120.0f};
    return rap[j] - range_threshold(j, band);  // DLL[j][band]
}
// RNS, the specific loudness at the slopes' range boundaries (table A.6), as literals for the
// branch-free range count.
constexpr float RNS_C[18] = {21.5f, 18.0f, 15.1f, 11.5f, 9.0f, 6.1f, 4.4f, 3.1f, 2.13f,
                             1.36f, 0.82f, 0.42f, 0.30f, 0.22f, 0.15f, 0.10f, 0.035f, 0.0f};

constexpr float LN10 = 2.302585092994045684f;
constexpr float LOG2_10 = 3.321928094887362f;

// 10^(db/10) on the hardware exp2 (about 1 ulp; the energies feed sums compared at 0.01 sone).
__device__ __forceinline__ float energy_of(float db) { return __builtin_amdgcn_exp2f(db * (LOG2_10 / 10.0f)); }

// physics/src/loudness.rs:148-182 over the core loudness of 99-144. The hardware exp2, log2,
// square root and reciprocal throughout (ISO 532-1's tables hold two or three digits); the range
// searches as counts of thresholds passed, with the tables' entries as literals where the band
// is fixed.
__device__ float zwicker_sone(const float in[28]) {
    float levels[28];
#pragma unroll
    for (int i = 0; i < 28; ++i) levels[i] = fmaxf(in[i], -60.0f);
    float lowest[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int band = 0; band < 11; ++band) {
        const float level = levels[band];
        // The range is the smallest j with level <= RAP[j] - DLL[j][band] (7 when none): the
        // correction of the last threshold passed.
        float correction = range_correction(0, band);
#pragma unroll
        for (int j = 0; j < 7; ++j)
            if (level > range_threshold(j, band)) correction = range_correction(j + 1, band);
        lowest[band < 6 ? 0 : (band < 9 ? 1 : 2)] += energy_of(level + correction);
    }
    float core[21];
#pragma unroll
    for (int band = 0; band < 20; ++band) {
        const float base = band < 3 ? (lowest[band] > 0.0f ? (10.0f / LOG2_10) * __builtin_amdgcn_logf(lowest[band])
                                                          : -INFINITY)
                                    : levels[band + 8];
        float level = base - A0[band];
        core[band] = 0.0f;
        if (level > LTQ[band]) {
            level -= DCB[band];
            const float threshold = energy_of(LTQ[band] * 0.25f);
            const float root = __builtin_amdgcn_sqrtf(__builtin_amdgcn_sqrtf(0.75f + 0.25f * energy_of(level - LTQ[band])));
            core[band] = fmaxf(0.0635f * threshold * (root - 1.0f), 0.0f);
        }
    }
    core[20] = 0.0f;
    const float fifth = core[0] > 0.0f ? __builtin_amdgcn_exp2f(0.2f * __builtin_amdgcn_logf(core[0])) : 0.0f;
    const float correction = 0.4f + 0.32f * fifth;
    if (correction <= 1.0f) core[0] *= correction;
    float total = 0.0f, z1 = 0.0f, n1 = 0.0f;
#pragma unroll
    for (int band = 0; band < 21; ++band) {
        const float upper = ZUP[band] + 0.0001f;
        const int group = band > 0 ? (band - 1 < 7 ? band - 1 : 7) : 0;
        const float c = core[band];
        while (z1 < upper) {
            if (n1 <= c) {
                total += c * (upper - z1);
                z1 = upper;
                n1 = c;
            } else {
                // The smallest j with n1 > RNS[j] (17 when none): RNS falls, so it counts the
                // entries n1 does not exceed.
                int range = 0;
#pragma unroll
                for (int j = 0; j < 17; ++j) range += n1 <= RNS_C[j] ? 1 : 0;
                const float steepness = USL[range][group];
                float n2 = fmaxf(RNS[range], c);
                float z2 = z1 + (n1 - n2) * __builtin_amdgcn_rcpf(steepness);
                if (z2 > upper) {
                    z2 = upper;
                    n2 = n1 - (z2 - z1) * steepness;
                }
                total += (z2 - z1) * (n1 + n2) * 0.5f;
                z1 = z2;
                n1 = n2;
            }
        }
    }
    return fmaxf(total, 0.0f);
}

}  // namespace
