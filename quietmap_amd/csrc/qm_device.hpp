// Device-side constants and small helpers of the painter (gfx950, wave64).
//
// Constants restate physics/src/bands.rs:3-21, physics/src/cnossos/ground.rs:8-21,
// physics/src/cnossos/diffraction.rs:13-16, physics/src/cnossos/rubber_band.rs:13-21,
// physics/src/foliage.rs:13-19 and physics/src/ray.rs:17-26.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qm_scene.h"

#define QM_DEV __device__ __forceinline__

namespace qm {

constexpr int BANDS = QM_BANDS;
constexpr int PERIODS = QM_PERIODS;
constexpr int SECTORS = QM_SECTORS;

constexpr float SPEED_OF_SOUND = 340.0f;
constexpr float LN10 = 2.302585092994045684f;
constexpr double PI_D = 3.141592653589793238462643383279502884;

constexpr float SHORT_PATH_HEIGHT_FACTOR = 30.0f;
constexpr float FAVOURABLE_CURVATURE_A0_PER_M = 2e-4f;
constexpr float FAVOURABLE_TERRAIN_HEIGHT_COEFFICIENT = 6e-3f;
constexpr float MINIMUM_HEIGHT_SUM_M = 1e-3f;
constexpr float HARD_GROUND_BELOW = 1e-3f;
constexpr float DIFFRACTION_CAP_DB = 25.0f;
constexpr float MULTIPLE_DIFFRACTION_MINIMUM_SPAN_M = 0.3f;
constexpr float CREST_ENDS_M = 300.0f;
constexpr float CREST_SPACING_M = 150.0f;

// Per-band constants: frequency, wavelength and the powers of (2.5.17), computed on the host in
// FP64 (physics/src/cnossos/ground.rs:61-74) and rounded once to FP32.
struct BandTable {
    float frequency[BANDS];
    float wavelength[BANDS];
    float f_2_5[BANDS];
    float f_1_5[BANDS];
    float f_0_75[BANDS];
    float wavenumber[BANDS];
    float foliage_short_db[BANDS];
    float foliage_db_per_m[BANDS];
    float inv_wavelength[BANDS];  // 1 / wavelength and 1 / wavenumber, so the band terms multiply
    float inv_wavenumber[BANDS];
};

QM_DEV float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// The hardware's 1-ulp square root and reciprocal (v_sqrt_f32, v_rcp_f32) for the FP32 geometry of
// the walk and the candidates, where the IEEE-rounded forms cost a correction sequence each.
QM_DEV float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
QM_DEV float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
QM_DEV float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
QM_DEV double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// physics/src/bands.rs:36-43 energy, amplitude
QM_DEV float energy_of(float level_db) { return __expf(level_db * (LN10 / 10.0f)); }
QM_DEV float amplitude_of(float level_db) { return __expf(level_db * (LN10 / 20.0f)); }

}  // namespace qm
