// The inputs of the excess-attenuation surrogate for one receiver-source path (no counterpart in
// the reference). Everything is read from the same scene as the exact path: the exact crossings
// (with the receiver's own walls and a building source's containing footprints dropped, as the
// exact path drops them), the ground at both ends, and 48 bilinear ground samples along the path
// (about the density of a 30 m lattice's lines on a 1.3 km path).
//
// Inputs (40, FP32), heights relative to the straight line from the source to the receiver:
//    0      horizontal length / 1000 m
//    1, 2   source and receiver heights above their ground / 4 m
//    3      where the highest ground sample over the line lies, x / length
//    4..15  per twelfth of the path, the highest of its four ground samples over the line, / 10 m
//   16..19  per quarter of the path, the mean G of its twelve samples
//   20..35  up to 8 wall tops in distance order (the 8 highest over the line): x / length and the
//           top's height over the line / 10 m; empty slots (0, -3)
//   36      wall tops / 8
//   37      0 calm, 1 favourable (the caller's; this function writes 0)
//   38      G under the source (as the exact path takes it)
//   39      the roofed fraction of the path
// The envelope the network is trained and used in: no failure, length >= 500 m, at most 8 wall
// tops, no forest at the samples or the ends. Outside it the exact path is the answer: the paths
// within 500 m carry much of a receiver's energy, so the exact path keeps them and the network's
// per-band error moves the receiver's level less.
#pragma once

#include "qm_eval.hpp"

namespace qm {

constexpr int FEATURES = 40;
constexpr int FEATURE_SAMPLES = 48;
constexpr float ENVELOPE_MIN_LENGTH_M = 500.0f;
constexpr int ENVELOPE_TOPS = 8;

struct SurrogatePath {
    uint32_t failed;   // QM_FAILED_* bits of the path's preparation
    bool in_envelope;
};

QM_DEV SurrogatePath surrogate_features(const Scene& sc, const qm_receiver& rcv, const qm_source& src, Crossing* cross,
                                        float f[FEATURES]) {
    SurrogatePath out{rcv.failed, false};
    for (int i = 0; i < FEATURES; ++i) f[i] = 0.0f;
    // A source on a bridge's deck takes the exact path: the network knows sources on the terrain.
    if (rcv.failed || src.on_deck) return out;
    const qm_terrain& T = sc.terrain;
    const double sx = src.x_m, sy = src.y_m, rx = rcv.x_m, ry = rcv.y_m;
    if (!isfinite(sx) || !isfinite(sy) || !isfinite(rx) || !isfinite(ry)) {
        out.failed = QM_FAILED_NOT_READ;
        return out;
    }
    const float H = (float)fmax(hypot(sx - rx, sy - ry), 1.0);
    int n = 0;
    uint32_t failed = path_crossings(sc, rcv, src, cross, n);
    const double inv = 1.0 / T.spacing_m;
    GroundSample s0, r1;
    failed |= ground_at(T, (sx - T.origin_x_m) * inv, (sy - T.origin_y_m) * inv, s0);
    failed |= ground_at(T, (rx - T.origin_x_m) * inv, (ry - T.origin_y_m) * inv, r1);
    if (failed & (QM_FAILED_NOT_READ | QM_FAILED_NO_TERRAIN)) {
        out.failed = failed;
        return out;
    }
    const float source_alt = s0.z + src.height_m, receiver_alt = rcv.altitude_m;
    const float slope = (receiver_alt - source_alt) / H;
    f[0] = H * 1e-3f;
    f[1] = src.height_m * 0.25f;
    f[2] = (receiver_alt - r1.z) * 0.25f;
    bool forest = s0.forest > 0.0f || r1.forest > 0.0f;
    float g_sum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float highest = -__builtin_inff();
    for (int j = 0; j < FEATURE_SAMPLES; ++j) {
        const double t = (j + 0.5) / FEATURE_SAMPLES;
        GroundSample g;
        failed |= ground_at(T, (sx + (rx - sx) * t - T.origin_x_m) * inv, (sy + (ry - sy) * t - T.origin_y_m) * inv, g);
        const float x = (float)t * H;
        const float above = (g.z - (source_alt + slope * x)) * 0.1f;
        const int k = 4 + j / 4;
        f[k] = (j & 3) ? fmaxf(f[k], above) : above;
        if (above > highest) {
            highest = above;
            f[3] = (float)t;
        }
        g_sum[j / 12] += g.g;
        forest = forest || g.forest > 0.0f;
    }
    for (int q = 0; q < 4; ++q) f[16 + q] = g_sum[q] * (1.0f / 12.0f);
    // Wall tops: the 8 highest over the line, then in distance order.
    float top_x[ENVELOPE_TOPS], top_h[ENVELOPE_TOPS];
    int tops = 0, kept = 0;
    float roofed = 0.0f, entry_x = 0.0f;
    uint64_t open_footprint = 0;
    bool open = false;
    for (int i = 0; i < n; ++i) {
        const Crossing& c = cross[i];
        if (c.flags & 4u) continue;
        const float x = (float)(c.t * (double)H);
        GroundSample g;
        failed |= ground_at(T, (sx + (rx - sx) * c.t - T.origin_x_m) * inv, (sy + (ry - sy) * c.t - T.origin_y_m) * inv, g);
        const float h = (g.z + c.height_m - (source_alt + slope * x)) * 0.1f;
        ++tops;
        if (kept < ENVELOPE_TOPS) {
            top_x[kept] = x;
            top_h[kept] = h;
            ++kept;
        } else {
            int low = 0;
            for (int k = 1; k < ENVELOPE_TOPS; ++k)
                if (top_h[k] < top_h[low]) low = k;
            if (h > top_h[low]) {
                for (int k = low; k + 1 < ENVELOPE_TOPS; ++k) {
                    top_x[k] = top_x[k + 1];
                    top_h[k] = top_h[k + 1];
                }
                top_x[ENVELOPE_TOPS - 1] = x;
                top_h[ENVELOPE_TOPS - 1] = h;
            }
        }
        // Roofs: a footprint's bounding crossings pair up in distance order.
        if (c.flags & 2u) {
            if (open && c.footprint == open_footprint) {
                roofed += x - entry_x;
                open = false;
            } else if (!open) {
                open = true;
                open_footprint = c.footprint;
                entry_x = x;
            }
        }
    }
    for (int k = 0; k < ENVELOPE_TOPS; ++k) {
        f[20 + 2 * k] = k < kept ? top_x[k] / H : 0.0f;
        f[21 + 2 * k] = k < kept ? top_h[k] : -3.0f;
    }
    f[36] = (float)tops * (1.0f / ENVELOPE_TOPS);
    f[38] = src.ground_percent != QM_GROUND_FROM_TERRAIN ? (float)src.ground_percent * 0.01f : s0.g;
    f[39] = fminf(roofed / H, 1.0f);
    out.failed = failed;
    out.in_envelope = !failed && H >= ENVELOPE_MIN_LENGTH_M && tops <= ENVELOPE_TOPS && !forest;
    return out;
}

}  // namespace qm
