// Doc 29 corrections and horizon screening as the card reads them, shared by the raw-track kernel
// (aircraft.hip) and the aircraft-box kernel (flights.hip): FP32 forms for the fast pass and FP64
// forms in the oracle's operation order (no contraction) for re-reads near a decision edge.
// physics/src/doc29/corrections.rs:23-88 (Delta_F, lateral attenuation, Delta_I),
// screening.rs:19-44 (path difference and edge losses), popup/src/aircraft/horizons.rs:84-111
// (SectorEdges::loss_db, sector_of); npd.rs:13-40 for the NPD distances' constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "qm_aircraft.h"

namespace qm_doc29 {

constexpr int ND = QM_NPD_DISTANCES;
constexpr double METRES_PER_FOOT = 0.3048;
constexpr double NPD_NEAREST_SLANT_M = 100.0 * METRES_PER_FOOT;
constexpr double NPD_LAST_DISTANCE_M = 25000.0 * METRES_PER_FOOT;
constexpr double TAIL_ANCHOR_M = 16000.0;
constexpr double PI_D = 3.141592653589793238462643383279502884;

// The fast pass's FP32 arithmetic on the hardware reciprocal, log2 and exp2 (about 1 ulp each, no
// IEEE division sequences): its error stays orders of magnitude under the 0.02 dB margin within
// which every decision (pass class, N65 / N70, histogram bin) is re-read in FP64.
__device__ __forceinline__ float div_f(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float log10_f(float x) { return __builtin_amdgcn_logf(x) * 0.301029995663981195f; }
__device__ __forceinline__ float exp10_f(float x) { return __builtin_amdgcn_exp2f(x * 3.32192809488736235f); }
__device__ __forceinline__ float sqrt_f(float x) { return __builtin_amdgcn_sqrtf(x); }

// corrections.rs:36-45: (e(a1) - e(a0)) / pi with e(a) = a / (1 + a^2) + atan(a), a1 - a0 =
// length / scaled. Where both ends lie two or more scaled distances to one side, e is within 0.03
// of pi / 2 at both and the FP32 difference loses its digits (a segment far behind the receiver
// read the 1e-15 floor for a fraction of 1e-8, which moved its pass to the lowest class); there
// the integral of e' = 2 / (1 + a^2)^2 is taken in u = 1 / |a|, 2 u^2 / (1 + u^2)^2, whose
// antiderivative sum_k (-1)^(k+1) 2k / (2k+1) u^(2k+1) gives the difference term by term as
// (u0 - u1) sum_k c_k P_2k, P_n = (u0^(n+1) - u1^(n+1)) / (u0 - u1), u <= 1/2.
__device__ __forceinline__ float finite_fraction_f(float along, float length, float scaled) {
    if (length < 1.0f || scaled < 1.0f) return 1.0f;
    const float inv_scaled = __builtin_amdgcn_rcpf(scaled);
    const float a0 = -along * inv_scaled, a1 = -(along - length) * inv_scaled;
    if (a0 * a1 > 0.0f && fminf(fabsf(a0), fabsf(a1)) >= 2.0f) {
        const float u0 = __builtin_amdgcn_rcpf(fminf(fabsf(a0), fabsf(a1))), u1 = __builtin_amdgcn_rcpf(fmaxf(fabsf(a0), fabsf(a1)));
        const float du = div_f(length * inv_scaled, a0 * a1);
        float p = 1.0f, power = 1.0f, sum = 0.0f, sign = 1.0f;
#pragma unroll
        for (int n = 1; n <= 28; ++n) {
            power *= u0;
            p = u1 * p + power;
            if ((n & 1) == 0) {
                const float k = (float)(n / 2);
                sum += sign * (2.0f * k / (2.0f * k + 1.0f)) * p;
                sign = -sign;
            }
        }
        return fmaxf(du * sum * (float)(1.0 / PI_D), 1e-15f);
    }
    const float e0 = div_f(a0, 1.0f + a0 * a0) + atanf(a0);
    const float e1 = div_f(a1, 1.0f + a1 * a1) + atanf(a1);
    return fmaxf((e1 - e0) * (float)(1.0 / PI_D), 1e-15f);
}

__device__ __forceinline__ float lateral_f(float height, float lateral) {
    const float beta = atanf(div_f(height, fmaxf(lateral, 0.01f))) * (float)(180.0 / PI_D);
    if (beta > 50.0f) return 0.0f;
    const float gamma = lateral <= 914.0f ? 1.089f * (1.0f - __expf(-0.00274f * lateral)) : 1.0f;
    const float lb = beta < 0.0f ? 10.857f : 1.137f - 0.0229f * beta + 9.72f * __expf(-0.142f * beta);
    return gamma * lb;
}

__device__ __forceinline__ float installation_f(uint32_t inst, float height, float slant) {
    if (inst == QM_INSTALL_PROPELLER) return 1.0f;
    const float a = inst == QM_INSTALL_WING ? 0.0039f : 0.1225f;
    const float b = inst == QM_INSTALL_WING ? 0.062f : 0.329f;
    const float c = inst == QM_INSTALL_WING ? 0.8786f : 1.0f;
    height = fmaxf(height, 0.0f);
    const float sin2 = div_f(height * height, fmaxf(slant * slant, 1e-12f));
    const float cos2 = 1.0f - sin2;
    const float num = a * cos2 + sin2;
    const float d = cos2 - sin2;
    const float den = c * 4.0f * sin2 * cos2 + d * d;
    return (den > 0.0f && num > 0.0f) ? div_f(__powf(num, b), den) : 1.0f;
}

// ---------------------------------------------------------------------------------------------
// FP64 reads, in the oracle's operation order.
inline __device__ double lateral_d(double height, double lateral) {
#pragma clang fp contract(off)
    const double beta = atan(height / fmax(lateral, 0.01)) * (180.0 / PI_D);
    if (beta > 50.0) return 0.0;
    const double gamma = lateral <= 914.0 ? 1.089 * (1.0 - exp(-0.00274 * lateral)) : 1.0;
    const double lb = beta < 0.0 ? 10.857 : 1.137 - 0.0229 * beta + 9.72 * exp(-0.142 * beta);
    return gamma * lb;
}

inline __device__ double installation_d(uint32_t inst, double height, double slant) {
#pragma clang fp contract(off)
    if (inst == QM_INSTALL_PROPELLER) return 1.0;
    const double a = inst == QM_INSTALL_WING ? 0.0039 : 0.1225;
    const double b = inst == QM_INSTALL_WING ? 0.062 : 0.329;
    const double c = inst == QM_INSTALL_WING ? 0.8786 : 1.0;
    height = fmax(height, 0.0);
    const double sin2 = height * height / fmax(slant * slant, 1e-12);
    const double cos2 = 1.0 - sin2;
    const double num = a * cos2 + sin2;
    const double d = cos2 - sin2;
    const double den = c * 4.0 * sin2 * cos2 + d * d;
    return (den > 0.0 && num > 0.0) ? pow(num, b) / den : 1.0;
}

// 10^(level / 10) as the oracle takes it (bands.rs:36-38).
inline __device__ double energy_d(double level_db) {
#pragma clang fp contract(off)
    return exp(level_db * (2.302585092994045684017991454684364208 / 10.0));
}

inline __device__ double finite_fraction_d(double along, double length, double scaled) {
#pragma clang fp contract(off)
    if (length < 1.0 || scaled < 1.0) return 1.0;
    const double a0 = -along / scaled, a1 = -(along - length) / scaled;
    const double e0 = a0 / (1.0 + a0 * a0) + atan(a0);
    const double e1 = a1 / (1.0 + a1 * a1) + atan(a1);
    return fmax((e1 - e0) * (1.0 / PI_D), 1e-15);
}

// screening.rs:19-44 and horizons.rs:84-111 in FP64, in the oracle's operation order.
inline __device__ double path_difference_loss_d(double delta) {
#pragma clang fp contract(off)
    const double raw = 10.0 * log10(3.0 + QM_PATH_DIFFERENCE_SLOPE_PER_M * fmax(delta, 0.0));
    return fmin(fmax(raw - QM_GRAZING_LOSS_DB, 0.0), QM_SCREENING_CAP_DB);
}

inline __device__ double edge_loss_d(double er, double eh, double sr, double sh) {
#pragma clang fp contract(off)
    if (!(er < sr && eh * sr > sh * er)) return 0.0;
    return path_difference_loss_d(hypot(er, eh) + hypot(sr - er, sh - eh) - hypot(sr, sh));
}

inline __device__ double horizon_loss_d(const float* at, int sectors, const double* p) {
#pragma clang fp contract(off)
    const double range = hypot(p[0], p[1]);
    if (range <= 1.0 || p[2] >= (double)at[0] * range) return 0.0;
    double angle = fmod(atan2(p[1], p[0]), 2.0 * PI_D);
    if (angle < 0.0) angle += 2.0 * PI_D;
    const int sector = min((int)(angle / (2.0 * PI_D) * (double)sectors), sectors - 1);
    const float* edges = at + 1 + sector * (1 + 2 * QM_HORIZON_BANDS);
    if (p[2] >= (double)edges[0] * range) return 0.0;
    double loss = 0.0;
    for (int band = 0; band < QM_HORIZON_BANDS; ++band) {
        const double r = edges[1 + 2 * band], t = edges[2 + 2 * band];
        if (isnan(r)) continue;
        loss = fmax(loss, edge_loss_d(r, t * r, range, p[2]));
    }
    return loss;
}

// screening.rs:19-44 in FP32: the single-edge path difference |RE| + |ES| - |RS| as
// 2 (u x v)^2 / ((|u||v| + u.v)(|u| + |v| + |u + v|)), u = RE, v = ES, which keeps its digits
// where the direct sum of three kilometre lengths loses the metre-scale difference.
__device__ __forceinline__ float path_difference_loss_f(float delta) {
    const float raw = 10.0f * log10_f(3.0f + (float)QM_PATH_DIFFERENCE_SLOPE_PER_M * fmaxf(delta, 0.0f));
    return fminf(fmaxf(raw - (float)QM_GRAZING_LOSS_DB, 0.0f), (float)QM_SCREENING_CAP_DB);
}

__device__ __forceinline__ float edge_loss_f(float er, float eh, float sr, float sh) {
    if (!(er < sr && eh * sr > sh * er)) return 0.0f;
    const float vx = sr - er, vy = sh - eh;
    const float a = sqrt_f(er * er + eh * eh), b = sqrt_f(vx * vx + vy * vy), c = sqrt_f(sr * sr + sh * sh);
    const float cross = er * vy - eh * vx, dot = er * vx + eh * vy;
    return path_difference_loss_f(div_f(2.0f * cross * cross, (a * b + dot) * (a + b + c)));
}

// horizons.rs:84-104 SectorEdges::loss_db in FP32 over the card's floats.
inline __device__ float horizon_loss_f(const float* at, int sectors, float x, float y, float z) {
    const float range = sqrt_f(x * x + y * y);
    if (range <= 1.0f || z >= at[0] * range) return 0.0f;
    float angle = atan2f(y, x);
    if (angle < 0.0f) angle += (float)(2.0 * PI_D);
    const int sector = min((int)(angle * (float)(sectors / (2.0 * PI_D))), sectors - 1);
    const float* edges = at + 1 + sector * (1 + 2 * QM_HORIZON_BANDS);
    if (z >= edges[0] * range) return 0.0f;
    float loss = 0.0f;
    for (int band = 0; band < QM_HORIZON_BANDS; ++band) {
        const float r = edges[1 + 2 * band];
        if (isnan(r)) continue;
        loss = fmaxf(loss, edge_loss_f(r, edges[2 + 2 * band] * r, range, z));
    }
    return loss;
}

}  // namespace qm_doc29
